// HIP kernels of the RealQM grid solver for gfx950 (MI355X).
//
// All fields are FP32 on the (S, S, S) node grid, flat id = i*S*S + j*S + k.
// Stencil kernels are 2.5D: a (bx, by) workgroup tile covers (k, j) and each
// thread marches `ich` consecutive i-planes, carrying the centre line
// (i-1, i, i+1) in registers so each plane is fetched once per tile; the
// in-plane neighbours come through the L1/L2 caches.  Only interior nodes
// (1..NN-1 on every axis) are written; faces are fixed Dirichlet data.
//
// Floating-point order follows the reference kernels term by term and the
// library is built with -ffp-contract=off, so FP32 results match an IEEE
// single-precision evaluation of the same expressions (the CPU oracle's
// FP32-faithful mode).  Reductions accumulate FP32 per-voxel terms in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <mutex>

#include "rq_internal.h"
#include "rq_stencil.h"

namespace rq {

namespace {

inline dim3 stencil_grid(int n, const Tile& t) {
  return dim3((n + t.bx - 1) / t.bx, (n + t.by - 1) / t.by, (n + t.ich - 1) / t.ich);
}

// Interior coordinate of this thread in a 2.5D launch; false if outside.
struct March {
  int k, j, i0, i1;
};

// Tile of this workgroup.  Workgroups are dispatched round-robin over the 8
// XCDs (linear id mod 8), each with its own L2; the linear id is remapped so
// that every XCD sweeps one contiguous range of tiles (neighbouring j-rows
// and i-chunks, whose halos overlap, then share an L2).
__device__ __forceinline__ bool march_coords(int n, int ich, March& m) {
  const int gx = gridDim.x, gy = gridDim.y;
  const int nblk = gx * gy * gridDim.z;
  const int b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const int per = nblk >> 3, rem = nblk & 7;
  const int xcd = b & 7, local = b >> 3;
  const int t = (xcd < rem) ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
  const int tx = t % gx, ty = (t / gx) % gy, tz = t / (gx * gy);
  m.k = 1 + tx * blockDim.x + threadIdx.x;
  m.j = 1 + ty * blockDim.y + threadIdx.y;
  m.i0 = 1 + tz * ich;
  m.i1 = min(m.i0 + ich, n + 1);
  return m.k <= n && m.j <= n;
}

// Angular sector test of a split atom (molecule.js:194-225 inSplitAtom):
// d = node - nucleus (bohr), r = |d|.  HEMI owns the half-space axis.d < 0
// (split_idx 0) or >= 0 (1); THIRD owns one 120-degree sector about the
// axis, measured from the in-plane frame (e1, axis x e1) less split_rot.
__device__ __forceinline__ bool in_split(float dx, float dy, float dz, float r, const AtomS2& e) {
  if (e.st <= 1 || r < 1e-6f) return true;
  const float axL = fmaxf(sqrtf(e.ax * e.ax + e.ay * e.ay + e.az * e.az), 1e-12f);
  const float a0 = e.ax / axL, a1 = e.ay / axL, a2 = e.az / axL;
  const float dotp = dx * a0 + dy * a1 + dz * a2;
  if (e.st == 2) return e.sidx == 0 ? dotp < 0.0f : dotp >= 0.0f;
  if (e.st == 3) {
    const float px = dx - dotp * a0, py = dy - dotp * a1, pz = dz - dotp * a2;
    float e1x, e1y, e1z;
    if (fabsf(a2) < 0.9f) {
      e1x = a1;
      e1y = -a0;
      e1z = 0.0f;
    } else {
      e1x = 0.0f;
      e1y = a2;
      e1z = -a1;
    }
    const float e1L = fmaxf(sqrtf(e1x * e1x + e1y * e1y + e1z * e1z), 1e-12f);
    e1x = e1x / e1L;
    e1y = e1y / e1L;
    e1z = e1z / e1L;
    const float e2x = a1 * e1z - a2 * e1y, e2y = a2 * e1x - a0 * e1z, e2z = a0 * e1y - a1 * e1x;
    const float pe1 = px * e1x + py * e1y + pz * e1z;
    const float pe2 = px * e2x + py * e2y + pz * e2z;
    float th = atan2f(pe2, pe1) - e.rot;
    if (th < -3.14159265f) th = th + 6.28318530f;
    if (th > 3.14159265f) th = th - 6.28318530f;
    // a negative quotient converts to sector 0, as WGSL's saturating u32()
    const unsigned sector = (unsigned)fmaxf(floorf((th + 3.14159265f) / 2.09439510f), 0.0f) % 3u;
    return sector == (unsigned)e.sidx;
  }
  return true;
}

// --------------------------------------------------------------------------
// Initialisation: trial functions, labels, K (cloud/solver.py:16-59 and
// molecule.js:1377-1449).
// --------------------------------------------------------------------------
__global__ void k_init_trial(int mode, int S, float h, float h2, float r_sing, const Atom* __restrict__ atoms,
                             const AtomS2* __restrict__ ext, int n_atoms, float* __restrict__ U,
                             float* __restrict__ U2, uint16_t* __restrict__ lab, float* __restrict__ K,
                             float* __restrict__ bestU, float* __restrict__ W) {
  const int S2 = S * S;
  const int S3 = S2 * S;
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= S3) return;
  const int i = id / S2;
  const int j = (id / S) % S;
  const int k = id % S;
  const int NN = S - 1;
  const bool face = (i < 1 || i >= NN || j < 1 || j >= NN || k < 1 || k >= NN);
  W[id] = 1.0f;
  if (mode == RQ_INIT_S1 && face) {
    U[id] = 0.0f;
    U2[id] = 0.0f;
    lab[id] = 0;
    K[id] = 0.0f;
    bestU[id] = 0.0f;
    return;
  }
  const float xi = (float)i * h, yj = (float)j * h, zk = (float)k * h;
  const float inv_sqrt_pi = 0.5641895835f;
  float bU = 0.0f;
  int bestN = 0;
  float Kv = 0.0f;
  if (mode == RQ_INIT_S2) {
    // molecule.js gpuInitAccum (:1401-1431): u32 positions, K from the
    // nuclear charge Z_nuc (bare nuclei too), the trial of decay init_zeff
    // only within init_rcut, on atoms with electrons and inside the atom's
    // split sector
    for (int n = 0; n < n_atoms; ++n) {
      const float Za = atoms[n].Z;
      const AtomS2 e = ext[n];
      const float Zn = e.Zn;
      if (Za <= 0.0f && Zn <= 0.0f) continue;
      const float rc = atoms[n].rc;
      const float dx = xi - floorf(atoms[n].i) * h;
      const float dy = yj - floorf(atoms[n].j) * h;
      const float dz = zk - floorf(atoms[n].k) * h;
      const float r2 = dx * dx + dy * dy + dz * dz;
      const float r = (rc > 0.0f) ? fmaxf(sqrtf(r2 + 0.04f * h2), rc) : fmaxf(sqrtf(r2), r_sing);
      Kv += Zn / r;
      const float Ze = e.zeff;
      const float rReal = sqrtf(r2);
      const bool ok = !(rReal > e.rcut || Za <= 0.0f) && in_split(dx, dy, dz, rReal, e);
      const float ut = ok ? Ze * Ze * inv_sqrt_pi * expf(-Ze * r) : 0.0f;
      if (ut > bU) {
        bU = ut;
        bestN = n;
      }
    }
  } else {
    for (int n = 0; n < n_atoms; ++n) {
      const float Za = atoms[n].Z;
      if (!(Za > 0.0f)) continue;
      const float rc = atoms[n].rc;
      const float dx = xi - atoms[n].i * h;
      const float dy = yj - atoms[n].j * h;
      const float dz = zk - atoms[n].k * h;
      const float r2 = dx * dx + dy * dy + dz * dz;
      const float r = (rc > 0.0f) ? fmaxf(sqrtf(r2 + 0.04f * h2), rc) : sqrtf(r2 + h2);
      Kv += Za / r;
      const float ut = Za * Za * inv_sqrt_pi * expf(-Za * r);
      if (ut > bU) {
        bU = ut;
        bestN = n;
      }
    }
  }
  K[id] = Kv;
  bestU[id] = bU;
  lab[id] = (uint16_t)bestN;
  U[id] = bU;
  U2[id] = (mode == RQ_INIT_S2) ? bU : 0.0f;
}

__global__ void k_recompute_k(int mode, int S, float h, float h2, float r_sing, const Atom* __restrict__ atoms,
                              const AtomS2* __restrict__ ext, int n_atoms, float* __restrict__ K) {
  const int S2 = S * S;
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= S2 * S) return;
  const int i = id / S2;
  const int j = (id / S) % S;
  const int k = id % S;
  const int NN = S - 1;
  if (mode == RQ_INIT_S1 && (i < 1 || i >= NN || j < 1 || j >= NN || k < 1 || k >= NN)) {
    K[id] = 0.0f;
    return;
  }
  float Kv = 0.0f;
  if (mode == RQ_INIT_S2) {
    // molecule.js recomputeK (:1337-1366): offsets formed in grid units, the
    // nuclear charge Z_nuc, r~ = max(r, rc) or max(r, 2h) (no 0.04 h^2 term)
    for (int n = 0; n < n_atoms; ++n) {
      const float Za = atoms[n].Z;
      const float Zn = ext[n].Zn;
      if (Za <= 0.0f && Zn <= 0.0f) continue;
      const float rc = atoms[n].rc;
      const float di = ((float)i - floorf(atoms[n].i)) * h;
      const float dj = ((float)j - floorf(atoms[n].j)) * h;
      const float dk = ((float)k - floorf(atoms[n].k)) * h;
      const float r2 = di * di + dj * dj + dk * dk;
      const float r = (rc > 0.0f) ? fmaxf(sqrtf(r2), rc) : fmaxf(sqrtf(r2), r_sing);
      Kv += Zn / r;
    }
  } else {
    const float xi = (float)i * h, yj = (float)j * h, zk = (float)k * h;
    for (int n = 0; n < n_atoms; ++n) {
      const float Za = atoms[n].Z;
      if (!(Za > 0.0f)) continue;
      const float rc = atoms[n].rc;
      const float dx = xi - atoms[n].i * h;
      const float dy = yj - atoms[n].j * h;
      const float dz = zk - atoms[n].k * h;
      const float r2 = dx * dx + dy * dy + dz * dz;
      const float r = (rc > 0.0f) ? fmaxf(sqrtf(r2 + 0.04f * h2), rc) : sqrtf(r2 + h2);
      Kv += Za / r;
    }
  }
  K[id] = Kv;
}

// rc flag (bit 15 of the label): node lies strictly inside the rc sphere of
// the nucleus of its own domain (molecule.js:242-246 isInsideRc, integer
// nucleus positions).
__global__ void k_rc_flags(int S, float h, const Atom* __restrict__ atoms, int n_atoms, uint16_t* __restrict__ lab) {
  const int S2 = S * S;
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= S2 * S) return;
  const int i = id / S2;
  const int j = (id / S) % S;
  const int k = id % S;
  const uint16_t raw = lab[id];
  const int L = raw & kLabelMask;
  uint16_t out = (uint16_t)L;
  if (L < n_atoms) {
    const float rc = atoms[L].rc;
    if (rc > 0.0f) {
      const float dx = ((float)i - floorf(atoms[L].i)) * h;
      const float dy = ((float)j - floorf(atoms[L].j)) * h;
      const float dz = ((float)k - floorf(atoms[L].k)) * h;
      if (sqrtf(dx * dx + dy * dy + dz * dz) < rc) out |= kRcFlag;
    }
  }
  lab[id] = out;
}

// --------------------------------------------------------------------------
// ITP update (cloud/solver.py:61-90; molecule.js:228-343 rc rule via flag).
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_itp(int S, int ich, const float* __restrict__ U, float* __restrict__ Un,
                                             const float* __restrict__ K, const float* __restrict__ P,
                                             const uint16_t* __restrict__ lab, float half_dv, float dtv) {
  March m;
  if (!march_coords(S - 2, ich, m)) return;
  itp_col(S, m.k, m.j, m.i0, m.i1, U, Un, K, P, lab, half_dv, dtv);
}

// --------------------------------------------------------------------------
// Poisson: damped-Jacobi smoother P' = a P + (sum6 P + rhs) / c.
// FINE: rhs = (alpha * U^2) * beta (cloud/solver.py:206 ``coeff*rho*h2``).
// COARSE: rhs = g.
// --------------------------------------------------------------------------
template <bool FINE>
__global__ __launch_bounds__(256) void k_smooth(int S, int ich, const float* __restrict__ P, float* __restrict__ Pn,
                                                const float* __restrict__ src, float a, float c, float alpha,
                                                float beta) {
  March m;
  if (!march_coords(S - 2, ich, m)) return;
  smooth_col<FINE>(S, m.k, m.j, m.i0, m.i1, P, Pn, src, a, c, alpha, beta);
}

// r = rhs - (6 x - sum6 x) on the interior.
template <bool FINE>
__global__ __launch_bounds__(256) void k_residual(int S, int ich, const float* __restrict__ x,
                                                  const float* __restrict__ src, float alpha, float beta,
                                                  float* __restrict__ r) {
  March m;
  if (!march_coords(S - 2, ich, m)) return;
  residual_col<FINE>(S, m.k, m.j, m.i0, m.i1, x, src, alpha, beta, r);
}

// Coarse rhs = sign * 4 * FW(r) at coarse interior nodes; FW = 27-point full
// weighting (8,4,2,1)/64 summed centre, faces, edges, corners in fixed order.
__global__ __launch_bounds__(256) void k_restrict4(int Sf, const float* __restrict__ r, float* __restrict__ gc,
                                                   float sign) {
  const int nc = (Sf - 1) / 2 - 1;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nc * nc * nc) return;
  restrict4_pt(Sf, tid / (nc * nc) + 1, (tid / nc) % nc + 1, tid % nc + 1, r, gc, sign);
}

// Trilinear prolongation (k axis first, then j, then i) and correction
// x += scale * P e on the fine interior.


// ---- fused multigrid passes (bit-identical to the unfused sequences) ----

// Two smoother sweeps from a zero initial guess: e1 = g / c, then
// e2 = a e1 + (sum6 e1 + g) / c, with e1 formed on the fly (g is zero on faces).
__global__ __launch_bounds__(256) void k_smooth2_from_zero(int S, int ich, const float* __restrict__ g,
                                                           float* __restrict__ e_out, float a, float c) {
  March m;
  if (!march_coords(S - 2, ich, m)) return;
  smooth2z_col(S, m.k, m.j, m.i0, m.i1, g, e_out, a, c);
}


// Linear workgroup index remapped so each XCD (linear id mod 8) takes one
// contiguous range of tiles (the remap of march_coords).
__device__ __forceinline__ int xcd_block() {
  const int nblk = gridDim.x * gridDim.y * gridDim.z;
  const int b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const int per = nblk >> 3, rem = nblk & 7;
  const int xcd = b & 7, local = b >> 3;
  return (xcd < rem) ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// Residual + 4 FW restriction through LDS.  A workgroup owns a CK x CJ x CI
// tile of coarse interior nodes: it first evaluates the residual once at each
// of the (2CK+1)(2CJ+1)(2CI+1) fine nodes the tile's restriction stencils
// cover (residual_at, the expression of k_residual), stores them in LDS, then
// every thread forms one coarse value with restrict27 (the sums of
// k_restrict4).  Same values as k_residual + k_restrict4, without the fine
// residual array in HBM and without recomputing each residual 27 times.
template <bool FINE, int CK, int CJ, int CI>
__global__ __launch_bounds__(CK * CJ * CI) void k_resres_tiled(int Sf, const float* __restrict__ x,
                                                                const float* __restrict__ src, float alpha,
                                                                float beta, float* __restrict__ gc, float sign) {
  __shared__ float r[(2 * CK + 1) * (2 * CJ + 1) * (2 * CI + 1)];
  const int nc = (Sf - 1) / 2 - 1;
  const int gx = (nc + CK - 1) / CK, gy = (nc + CJ - 1) / CJ;
  const int t = xcd_block();
  resres_tile<FINE, CK, CJ, CI>(Sf, 1 + (t % gx) * CK, 1 + ((t / gx) % gy) * CJ, 1 + (t / (gx * gy)) * CI, x, src,
                                alpha, beta, gc, sign, r);
}


// Residual + 4 FW restriction as a march along i for the large levels: one
// workgroup per 32 x CJ coarse-column tile and CI coarse planes
// (resres_march_tile; with CP = 2, two of every four waves restrict and a
// barrier pair serves two coarse planes).
template <bool FINE, int CJ, int CP>
__global__ __launch_bounds__(128 * CJ) void k_resres_march(int Sf, int CI, const float* __restrict__ x,
                                                            const float* __restrict__ src, float alpha, float beta,
                                                            float* __restrict__ gc, float sign) {
  constexpr int PL = 65 * (2 * CJ + 1), RING = 2 * CP + 2;
  __shared__ float ring[RING * PL];
  const int nc = (Sf - 1) / 2 - 1;
  const int gx = (nc + 31) / 32, gy = (nc + CJ - 1) / CJ;
  const int t = xcd_block();
  resres_march_tile<FINE, CJ, CP>(Sf, CI, 1 + (t % gx) * 32, 1 + ((t / gx) % gy) * CJ, 1 + (t / (gx * gy)) * CI, x,
                                  src, alpha, beta, gc, sign, ring);
}

// Two damped-Jacobi sweeps fused as a march along i (temporal blocking).  A
// workgroup owns TK (k) x TJ (j) interior columns and planes [I0, I1).  The
// first sweep is formed on the (TK+2) x (TJ+2) region the second sweep's
// stencils read (each thread carries the i-1 / i values of its region nodes
// in registers) into an LDS ring of CP+2 planes; face nodes keep x there (the
// two ping-pong buffers hold the same faces).  Each round forms CP planes of
// the first sweep and then CP planes of the second from the ring.  Same
// smooth_col expressions in the same order, so the output equals two k_smooth
// launches bit for bit, with one read of x and src and one write.
template <bool FINE, int TK, int TJ, int CP>
__global__ __launch_bounds__(256) void k_smooth2_march(int S, int CI, const float* __restrict__ x,
                                                       float* __restrict__ xn, const float* __restrict__ src,
                                                       float a, float c, float alpha, float beta) {
  constexpr int NT = 256, RK = TK + 2, RJ = TJ + 2, PL = RK * RJ, PTS = (PL + NT - 1) / NT, RING = CP + 2;
  constexpr int OUT = TK * TJ * CP, OPT = (OUT + NT - 1) / NT;
  __shared__ float ring[RING * PL];
  const int n = S - 2, S2 = S * S;
  const int gx = (n + TK - 1) / TK, gy = (n + TJ - 1) / TJ;
  const int t = xcd_block();
  const int K0 = 1 + (t % gx) * TK, J0 = 1 + ((t / gx) % gy) * TJ, I0 = 1 + (t / (gx * gy)) * CI;
  const int I1 = min(I0 + CI, n + 1);
  const int tid = threadIdx.x;
  // region nodes of this thread: (j, k) clamped into the grid for the loads;
  // interior nodes read their four in-plane neighbours at offsets oj / ok,
  // faces (and unused slots beyond S - 1) read themselves and keep x, so every
  // load of a plane is issued without a branch
  int col[PTS], oj[PTS], ok[PTS];
  bool upd[PTS];
  float xm[PTS], xc[PTS];
#pragma unroll
  for (int p = 0; p < PTS; ++p) {
    const int q = min(tid + p * NT, PL - 1);
    const int j = J0 - 1 + q / RK, k = K0 - 1 + q % RK;
    upd[p] = tid + p * NT < PL && j >= 1 && j <= n && k >= 1 && k <= n;
    oj[p] = upd[p] ? S : 0;
    ok[p] = upd[p] ? 1 : 0;
    col[p] = min(j, S - 1) * S + min(k, S - 1);
    xc[p] = x[(I0 - 1) * S2 + col[p]];
    xm[p] = (I0 >= 2) ? x[(I0 - 2) * S2 + col[p]] : 0.0f;
  }
  // first sweep at plane f (I0 - 1 <= f <= I1) into ring slot f % RING
  auto sweep1 = [&](int f) {
    float* dst = ring + (f % RING) * PL + tid;
    if (f >= 1 && f <= n) {
      float xp[PTS], nb[PTS][4], u[PTS];
#pragma unroll
      for (int p = 0; p < PTS; ++p) {
        const int id = f * S2 + col[p];
        xp[p] = x[id + S2];
        nb[p][0] = x[id + oj[p]];
        nb[p][1] = x[id - oj[p]];
        nb[p][2] = x[id + ok[p]];
        nb[p][3] = x[id - ok[p]];
        u[p] = src[id];
      }
#pragma unroll
      for (int p = 0; p < PTS; ++p) {
        const float s6 = xp[p] + xm[p] + nb[p][0] + nb[p][1] + nb[p][2] + nb[p][3];
        const float rhs = FINE ? (alpha * (u[p] * u[p])) * beta : u[p];
        float v = (s6 + rhs) / c;
        if (a != 0.0f) v = a * xc[p] + v;
        if (tid + p * NT < PL) dst[p * NT] = upd[p] ? v : xc[p];
        xm[p] = xc[p];
        xc[p] = xp[p];
      }
    } else {
      // face plane (f == 0 or f == S - 1): the first sweep keeps x
#pragma unroll
      for (int p = 0; p < PTS; ++p) {
        if (tid + p * NT < PL) dst[p * NT] = xc[p];
        const float xp = (f + 1 <= S - 1) ? x[(f + 1) * S2 + col[p]] : 0.0f;
        xm[p] = xc[p];
        xc[p] = xp;
      }
    }
  };
  // second-sweep outputs of this thread in a round: plane offset, LDS centre
  int oh[OPT], occ[OPT], oid[OPT];
#pragma unroll
  for (int q = 0; q < OPT; ++q) {
    const int o = min(tid + q * NT, OUT - 1);
    const int tj = (o / TK) % TJ, tk = o % TK;
    oh[q] = o / (TK * TJ);
    occ[q] = (tj + 1) * RK + (tk + 1);
    const bool inside = tid + q * NT < OUT && J0 + tj <= n && K0 + tk <= n;
    oid[q] = inside ? (J0 + tj) * S + K0 + tk : -1;
  }
  sweep1(I0 - 1);
  sweep1(I0);
  for (int I = I0; I < I1; I += CP) {
#pragma unroll
    for (int h = 1; h <= CP; ++h)
      if (I + h <= I1) sweep1(I + h);
    float u[OPT];
#pragma unroll
    for (int q = 0; q < OPT; ++q) {
      const int Ih = min(I + oh[q], I1 - 1);
      u[q] = src[Ih * S2 + max(oid[q], 0)];
    }
    lds_barrier();
#pragma unroll
    for (int q = 0; q < OPT; ++q) {
      const int Ih = I + oh[q];
      const int cc = occ[q];
      const int s0 = (Ih % RING) * PL + cc, sm = ((Ih + RING - 1) % RING) * PL + cc, sp = ((Ih + 1) % RING) * PL + cc;
      const float s6 = ring[sp] + ring[sm] + ring[s0 + RK] + ring[s0 - RK] + ring[s0 + 1] + ring[s0 - 1];
      const float rhs = FINE ? (alpha * (u[q] * u[q])) * beta : u[q];
      float v = (s6 + rhs) / c;
      if (a != 0.0f) v = a * ring[s0] + v;
      if (oid[q] >= 0 && Ih < I1) xn[Ih * S2 + oid[q]] = v;
    }
    lds_barrier();
  }
}

// The fine level's pre-smoothing pair, residual and 4 FW restriction as one
// march along i (temporal blocking of the whole down step).  A workgroup
// owns CK x CJ coarse interior columns (K0.., J0..) and coarse planes
// [I0, I1).  Fine regions around the tile, per plane: R0 (2CK+1) x (2CJ+1)
// nodes whose residuals the tile's restriction reads, R1 = R0 + 1 (second
// sweep) and R2 = R0 + 2 (first sweep, from x in global memory as
// k_smooth2_march forms it).  Fine plane p of the march: first sweep at p
// into an LDS ring (with U, which the later stages read from there), second
// sweep at p - 1 from it into a second ring (and to xn where the tile owns
// the node), residual at p - 2 into a third ring, and the coarse plane whose
// three fine planes are complete is restricted from the third ring.  The
// global loads of plane p + 1 are issued right after plane p's first sweep,
// so they are in flight across the later stages and barriers (the barriers
// wait for LDS traffic only).  Rings of three planes with a barrier after
// each of the first three stages: a slot is rewritten only after the
// barriers that end every read of it.  The smooth_col, residual and
// restrict27 expressions in the same order, so xn and gc equal two k_smooth
// launches followed by k_resres_march bit for bit.
template <int CK, int CJ, int NT>
__global__ __launch_bounds__(NT) void k_smooth2_resres_march(int Sf, int CI, const float* __restrict__ x,
                                                              float* __restrict__ xn, const float* __restrict__ src,
                                                              float a, float c, float alpha, float beta,
                                                              float* __restrict__ gc) {
  constexpr int RING = 3;
  constexpr int W0 = 2 * CK + 1, H0 = 2 * CJ + 1, W1 = W0 + 2, H1 = H0 + 2, W2 = W0 + 4, H2 = H0 + 4;
  // LDS rows: the first-sweep, U and second-sweep rings share the row pitch
  // W2, and every stage maps its threads onto rows of W2 slots (the last
  // ones idle on the narrower regions), so a wave's reads and writes are
  // consecutive words even across row ends.  The residual ring stores each
  // row's even-k entries, then from R0O (= 16 mod 32) its odd-k ones: the
  // restriction's stride-2 reads of a row become three unit-stride reads
  // and the residual's stores fall on disjoint banks.
  constexpr int E0 = CK + 1, O0 = CK;  // even / odd entries of an R0 row
  constexpr int R0O = 16 + 32 * ((E0 > 16 ? E0 - 16 + 31 : 0) / 32), R0P = R0O + O0;
  constexpr int PL0 = R0P * H0, PL1 = W2 * H1, PL2 = W2 * H2, Q0 = W2 * H0;
  constexpr int P0 = (Q0 + NT - 1) / NT, P1 = (PL1 + NT - 1) / NT, P2 = (PL2 + NT - 1) / NT;
  static_assert(CK * CJ <= NT, "one thread per coarse column");
  __shared__ float ring2[RING * PL2];  // first sweep
  __shared__ float ringu[RING * PL2];  // U on the first-sweep region
  __shared__ float ring1[RING * PL1];  // second sweep
  __shared__ float ring0[RING * PL0];  // residual (rows split even / odd k)
  const int n = Sf - 2, nc = (Sf - 1) / 2 - 1, Sc = nc + 2, S2 = Sf * Sf;
  const int gx = (nc + CK - 1) / CK, gy = (nc + CJ - 1) / CJ;
  const int t = xcd_block();
  const int K0 = 1 + (t % gx) * CK, J0 = 1 + ((t / gx) % gy) * CJ, I0 = 1 + (t / (gx * gy)) * CI;
  const int I1 = min(I0 + CI, nc + 1);
  const int tid = threadIdx.x;
  // fine nodes the tile writes: [2K0 - 1, 2K0 + 2CK - 2] per axis, to n on the last tile
  const int k_lo = 2 * K0 - 1, k_hi = (K0 + CK - 1 >= nc) ? n : 2 * K0 + 2 * CK - 2;
  const int j_lo = 2 * J0 - 1, j_hi = (J0 + CJ - 1 >= nc) ? n : 2 * J0 + 2 * CJ - 2;
  const int i_lo = 2 * I0 - 1, i_hi = (I1 == nc + 1) ? n : 2 * I1 - 2;
  // first-sweep region nodes of this thread: byte offset of the column
  // clamped into the grid.  On a plane p in [1, n] every load of the sweep
  // (x at p +- 1 and the four in-plane neighbours, U at p) stays inside the
  // grid for a clamped column; the values loaded for a non-interior node are
  // discarded by the select that keeps its x.
  uint32_t col4[P2];
  bool upd[P2];
  float xm[P2], xc[P2], xp[P2], nb[P2][4], u[P2];
  const int p_first = 2 * I0 - 3, p_last = 2 * I1 + 1;
#pragma unroll
  for (int q = 0; q < P2; ++q) {
    const int o = min(tid + q * NT, PL2 - 1);
    const int j = 2 * J0 - 3 + o / W2, k = 2 * K0 - 3 + o % W2;
    upd[q] = tid + q * NT < PL2 && j >= 1 && j <= n && k >= 1 && k <= n;
    col4[q] = 4u * static_cast<uint32_t>(min(max(j, 0), Sf - 1) * Sf + min(max(k, 0), Sf - 1));
    xm[q] = (p_first - 1 >= 0) ? ld_at(x + (p_first - 1) * S2, col4[q]) : 0.0f;
    xc[q] = (p_first >= 0) ? ld_at(x + p_first * S2, col4[q]) : 0.0f;
  }
  // second-sweep nodes: LDS index of the node in a first-sweep slot, and the
  // byte offset of the store for nodes the tile owns (~0u: no store)
  int c1[P1];
  uint32_t st1[P1];
  bool in1[P1], ok1[P1];
#pragma unroll
  for (int q = 0; q < P1; ++q) {
    const int o = min(tid + q * NT, PL1 - 1);
    const int jj = o / W2, kk = o % W2;
    const int j = 2 * J0 - 2 + jj, k = 2 * K0 - 2 + kk;
    c1[q] = o + W2 + 1;  // (jj + 1) W2 + kk + 1
    ok1[q] = tid + q * NT < PL1 && kk < W1;
    in1[q] = j >= 1 && j <= n && k >= 1 && k <= n;
    st1[q] = (in1[q] && j >= j_lo && j <= j_hi && k >= k_lo && k <= k_hi) ? 4u * static_cast<uint32_t>(j * Sf + k)
                                                                          : ~0u;
  }
  // inputs of the first sweep at plane p: x at p + 1, the in-plane
  // neighbours and U at p (only x at p + 1 outside the interior planes)
  {
    const int PP = p_first;
    if (PP >= 1 && PP <= n) {
      const float* xq = x + PP * S2;
      const float* uq = src + PP * S2;
#pragma unroll
      for (int q = 0; q < P2; ++q) {
        xp[q] = ld_at(xq + S2, col4[q]);
        nb[q][0] = ld_at(xq + Sf, col4[q]);
        nb[q][1] = ld_at(xq - Sf, col4[q]);
        nb[q][2] = ld_at(xq + 1, col4[q]);
        nb[q][3] = ld_at(xq - 1, col4[q]);
        u[q] = ld_at(uq, col4[q]);
      }
    } else {
      const bool in = PP + 1 >= 0 && PP + 1 <= Sf - 1;
#pragma unroll
      for (int q = 0; q < P2; ++q) xp[q] = in ? ld_at(x + (PP + 1) * S2, col4[q]) : 0.0f;
    }
  }

  // ring slots of planes p, p - 1, p - 2 (p_first >= -1)
  int s_p = (p_first + RING) % RING;
  for (int p = p_first; p <= p_last; ++p) {
    const int s_m1 = (s_p + RING - 1) % RING, s_m2 = (s_p + RING - 2) % RING;
    // ---- first sweep at fine plane p ----
    {
      float* dst = ring2 + s_p * PL2 + tid;
      float* dstu = ringu + s_p * PL2 + tid;
      if (p >= 1 && p <= n) {
#pragma unroll
        for (int q = 0; q < P2; ++q) {
          const float s6 = xp[q] + xm[q] + nb[q][0] + nb[q][1] + nb[q][2] + nb[q][3];
          const float rhs = (alpha * (u[q] * u[q])) * beta;
          float v = (s6 + rhs) / c;
          if (a != 0.0f) v = a * xc[q] + v;
          if (tid + q * NT < PL2) {
            dst[q * NT] = upd[q] ? v : xc[q];
            dstu[q * NT] = u[q];
          }
        }
      } else if (p == 0 || p == Sf - 1) {
        // face planes keep x
#pragma unroll
        for (int q = 0; q < P2; ++q)
          if (tid + q * NT < PL2) dst[q * NT] = xc[q];
      }
#pragma unroll
      for (int q = 0; q < P2; ++q) {
        xm[q] = xc[q];
        xc[q] = xp[q];
      }
      if (p < p_last) {
        const int PP = p + 1;
        if (PP >= 1 && PP <= n) {
          const float* xq = x + PP * S2;
          const float* uq = src + PP * S2;
#pragma unroll
          for (int q = 0; q < P2; ++q) {
            xp[q] = ld_at(xq + S2, col4[q]);
            nb[q][0] = ld_at(xq + Sf, col4[q]);
            nb[q][1] = ld_at(xq - Sf, col4[q]);
            nb[q][2] = ld_at(xq + 1, col4[q]);
            nb[q][3] = ld_at(xq - 1, col4[q]);
            u[q] = ld_at(uq, col4[q]);
          }
        } else {
          const bool in = PP + 1 >= 0 && PP + 1 <= Sf - 1;
#pragma unroll
          for (int q = 0; q < P2; ++q) xp[q] = in ? ld_at(x + (PP + 1) * S2, col4[q]) : 0.0f;
        }
      }
    }
    lds_barrier();
    // ---- second sweep at fine plane p - 1 ----
    const int f1 = p - 1;
    if (f1 >= 2 * I0 - 2 && f1 >= 0 && f1 <= Sf - 1) {
      const bool pin = f1 >= 1 && f1 <= n;
      const bool own_plane = f1 >= i_lo && f1 <= i_hi;
      float* xf = xn + f1 * S2;
      // planes f1 + 1 = p, f1, f1 - 1 in the slots of p, p - 1, p - 2
      const float* rp = ring2 + s_p * PL2;
      const float* r0 = ring2 + s_m1 * PL2;
      const float* rm = ring2 + s_m2 * PL2;
      const float* u0 = ringu + s_m1 * PL2;
#pragma unroll
      for (int q = 0; q < P1; ++q) {
        if (ok1[q]) {
          const int cc = c1[q];
          float v = r0[cc];
          if (pin && in1[q]) {
            const float s6 = rp[cc] + rm[cc] + r0[cc + W2] + r0[cc - W2] + r0[cc + 1] + r0[cc - 1];
            const float uu = u0[cc];
            const float rhs = (alpha * (uu * uu)) * beta;
            float w = (s6 + rhs) / c;
            if (a != 0.0f) w = a * v + w;
            v = w;
            if (own_plane && st1[q] != ~0u) st_at(xf, st1[q], v);
          }
          ring1[s_m1 * PL1 + tid + q * NT] = v;
        }
      }
    }
    lds_barrier();
    // ---- residual at fine plane p - 2 ----
    const int f0 = p - 2;
    if (f0 >= 2 * I0 - 1 && f0 <= 2 * I1 - 1) {
      const float* q0 = ring1 + s_m2 * PL1;
      const float* qa = ring1 + s_m1 * PL1;
      const float* qb = ring1 + s_p * PL1;
      const float* uq = ringu + s_m2 * PL2;
#pragma unroll
      for (int q = 0; q < P0; ++q) {
        const int o = tid + q * NT;
        const int jj = o / W2, kk = o % W2;
        if (o < Q0 && kk < W0) {
          const int j = 2 * J0 - 1 + jj, k = 2 * K0 - 1 + kk;
          const int cc = o + W2 + 1;  // (jj + 1) W2 + kk + 1
          const float s6 = qa[cc] + qb[cc] + q0[cc + W2] + q0[cc - W2] + q0[cc + 1] + q0[cc - 1];
          const float uu = uq[cc + W2 + 1];  // (jj + 2) W2 + kk + 2
          const float rhs = (alpha * (uu * uu)) * beta;
          const float r = rhs - (6.0f * q0[cc] - s6);
          ring0[s_m2 * PL0 + jj * R0P + ((kk & 1) ? R0O + (kk >> 1) : (kk >> 1))] = (j <= n && k <= n) ? r : 0.0f;
        }
      }
    }
    lds_barrier();
    // ---- restriction of coarse plane (f0 - 1) / 2 once its three fine planes are in ----
    if ((f0 & 1) && f0 >= 2 * I0 + 1 && f0 <= 2 * I1 - 1 && tid < CK * CJ) {
      const int I = (f0 - 1) / 2, tk = tid % CK, tj = tid / CK;
      const int K = K0 + tk, J = J0 + tj;
      if (K <= nc && J <= nc) {
        // the centre (row 2 tj + 1, k entry 2 tk + 1: odd entry tk); its k-1 /
        // k+1 neighbours are the even entries tk and tk + 1
        const int cc = (2 * tj + 1) * R0P + R0O + tk;
        // planes f0 - 2 = p - 4, f0 - 1 = p - 3 and f0 = p - 2 sit in the
        // slots of p - 1, p and p - 2 (three-slot rings)
        gc[I * Sc * Sc + J * Sc + K] =
            restrict27_planes_k(ring0, s_m1 * PL0 + cc, s_p * PL0 + cc, s_m2 * PL0 + cc, R0P, 1 - R0O, -R0O);
      }
    }
    s_p = (s_p + 1) % RING;
  }
}

// The fine level's up step and the ITP update as one march along i: the
// coarse correction (x + P e), two damped-Jacobi post-sweeps and the ITP
// update on the new potential (temporal blocking of prolongation, sweep
// pair and update).  A workgroup owns TK x TJ fine interior columns and
// planes [I0, I1).  Per plane: stage 0 loads, on the region R3 = tile + 2,
// x with the trilinear coarse correction (each coarse-plane interpolant
// formed once, prolong_correct_col's operations), U and the label into LDS
// rings; stage 1 forms the first sweep
// on R2 = tile + 1 from them; stage 2 forms the second sweep on the tile
// (written to xn) and the ITP update of U with it (written to Un).  Stage 0
// runs at plane s, stage 1 at s - 1, stage 2 at s - 2; rings of three (x,
// first sweep) and four (U, label) planes with a barrier after each stage.
// The expressions of prolong_correct_col (scale 1), smooth_col and itp_col
// in the same order: xn and Un equal k_prolong_correct, two k_smooth and
// k_itp bit for bit.
template <int TK, int TJ, int RU>
__global__ __launch_bounds__(TK * TJ) void k_post_march(int Sf, int CI, const float* __restrict__ x,
                                                        const float* __restrict__ e, float* __restrict__ xn,
                                                        const float* __restrict__ U, float* __restrict__ Un,
                                                        const float* __restrict__ K, const uint16_t* __restrict__ lab,
                                                        float a, float c, float alpha, float beta, float half_dv,
                                                        float dtv) {
  constexpr int NT = TK * TJ;
  constexpr int W3 = TK + 4, H3 = TJ + 4, W2 = TK + 2, H2 = TJ + 2;
  // every ring has the row pitch W3 and stage 1 maps its threads onto rows
  // of W3 slots (two idle per row), so a wave's LDS accesses are consecutive
  // words across row ends
  constexpr int PL3 = W3 * H3, PL2 = W3 * H2;
  constexpr int P3 = (PL3 + NT - 1) / NT, P2 = (PL2 + NT - 1) / NT;
  __shared__ float rx[3 * PL3];       // corrected x
  static_assert(RU == 4 || RU == 5, "U / label ring of four (barrier after stage 2) or five planes");
  __shared__ float ru[RU * PL3];      // U
  __shared__ uint16_t rl[RU * PL3];   // label
  __shared__ float r1[3 * PL2];       // first sweep
  const int n = Sf - 2, Sc = (Sf - 1) / 2 + 1, S2 = Sf * Sf, Sc2 = Sc * Sc;
  const int gx = (n + TK - 1) / TK, gy = (n + TJ - 1) / TJ;
  const int t = xcd_block();
  const int K0 = 1 + (t % gx) * TK, J0 = 1 + ((t / gx) % gy) * TJ, I0 = 1 + (t / (gx * gy)) * CI;
  const int I1 = min(I0 + CI, n + 1);
  const int tid = threadIdx.x;
  // stage-0 nodes of this thread on R3: byte offsets of the column clamped
  // into the grid (x, U; the label at half of it), and, for interior
  // columns, of the four coarse nodes of the bilinear in-plane interpolant
  // (lerp_k along k on rows j >> 1 and (j >> 1) + (j & 1)); other columns
  // read coarse node 0, which is in the grid, and discard the value
  uint32_t col4[P3], col2[P3], e00[P3], e01[P3], e10[P3], e11[P3];
  bool cin[P3];
#pragma unroll
  for (int q = 0; q < P3; ++q) {
    const int o = min(tid + q * NT, PL3 - 1);
    const int jr = J0 - 2 + o / W3, kr = K0 - 2 + o % W3;
    cin[q] = tid + q * NT < PL3 && jr >= 1 && jr <= n && kr >= 1 && kr <= n;
    const uint32_t col = static_cast<uint32_t>(min(max(jr, 0), Sf - 1) * Sf + min(max(kr, 0), Sf - 1));
    col4[q] = 4u * col;
    col2[q] = 2u * col;
    const uint32_t b = cin[q] ? static_cast<uint32_t>((jr >> 1) * Sc + (kr >> 1)) : 0u;
    const uint32_t dk = cin[q] ? static_cast<uint32_t>(kr & 1) : 0u;
    const uint32_t dj = cin[q] ? static_cast<uint32_t>((jr & 1) * Sc) : 0u;
    e00[q] = 4u * b;
    e01[q] = 4u * (b + dk);
    e10[q] = 4u * (b + dj);
    e11[q] = 4u * (b + dj + dk);
  }
  // first-sweep nodes on R2: LDS index in an R3 slot, slot in use, interior flag
  int c2[P2];
  bool ok2[P2], in2[P2];
#pragma unroll
  for (int q = 0; q < P2; ++q) {
    const int o = min(tid + q * NT, PL2 - 1);
    const int jj = o / W3, kk = o % W3;
    const int j = J0 - 1 + jj, k = K0 - 1 + kk;
    c2[q] = o + W3 + 1;  // (jj + 1) W3 + kk + 1
    ok2[q] = tid + q * NT < PL2 && kk < W2;
    in2[q] = j >= 1 && j <= n && k >= 1 && k <= n;
  }
  // stage-2 node of this thread (one per thread)
  const int tk = tid % TK, tj = tid / TK;
  const int jo = J0 + tj, ko = K0 + tk;
  const bool own = jo <= n && ko <= n;
  const uint32_t own4 = own ? 4u * static_cast<uint32_t>(jo * Sf + ko) : 0u;
  // Coarse-plane interpolants of the prolongation: fine plane s uses
  // T(s >> 1) and T((s + 1) >> 1) (tA, tB; the same plane for even s), so
  // each odd plane after the first brings in one new coarse plane.
  auto tplane_raw = [&](int ic, int q, float* r) {
    const float* ep = e + ic * Sc2;
    r[0] = ld_at(ep, e00[q]);
    r[1] = ld_at(ep, e01[q]);
    r[2] = ld_at(ep, e10[q]);
    r[3] = ld_at(ep, e11[q]);
  };
  auto tcombine = [](const float* r) -> float {
    const float l0 = 0.5f * (r[0] + r[1]);
    const float l1 = 0.5f * (r[2] + r[3]);
    return 0.5f * (l0 + l1);
  };
  // Every global input is fetched one iteration ahead, so its latency is
  // spent behind the current plane's stages and barriers: x, U and the label
  // of plane s + 1, the raw coarse values of the coarse plane that plane s + 1
  // brings in, and K at the owned column of the plane stage 2 handles next.
  float xq[P3] = {}, uq[P3] = {}, tA[P3] = {}, tB[P3] = {}, pend[P3][4] = {};
  uint16_t lq[P3] = {};
  float kq = 0.0f;
  const int s0 = max(I0 - 2, 1);  // first plane of the march that takes the correction
  if (s0 <= n) {
#pragma unroll
    for (int q = 0; q < P3; ++q) {
      float r[4];
      tplane_raw(s0 >> 1, q, r);
      tA[q] = tcombine(r);
      tplane_raw((s0 + 1) >> 1, q, r);
      tB[q] = tcombine(r);
    }
  }
  auto fetch = [&](int s) {
    if (s >= 0 && s <= Sf - 1) {
      const float* xs = x + s * S2;
      const float* us = U + s * S2;
      const uint16_t* ls = lab + s * S2;
#pragma unroll
      for (int q = 0; q < P3; ++q) {
        xq[q] = ld_at(xs, col4[q]);
        uq[q] = ld_at(us, col4[q]);
        lq[q] = ld_at16(ls, col2[q]);
      }
      if ((s & 1) && s > s0 && s <= n) {
#pragma unroll
        for (int q = 0; q < P3; ++q) tplane_raw((s + 1) >> 1, q, pend[q]);
      }
    }
  };
  fetch(I0 - 2);
  for (int s = I0 - 2; s <= I1 + 1; ++s) {
    float xv[P3], uv[P3];
    uint16_t lv[P3];
#pragma unroll
    for (int q = 0; q < P3; ++q) {
      xv[q] = xq[q];
      uv[q] = uq[q];
      lv[q] = lq[q];
    }
    const float kv = kq;
    // ---- coarse interpolants of plane s (before pend is refilled) ----
    if (s > s0 && s <= n) {
#pragma unroll
      for (int q = 0; q < P3; ++q) {
        tA[q] = tB[q];
        if (s & 1) tB[q] = tcombine(pend[q]);
      }
    }
    if (s + 1 <= I1 + 1) fetch(s + 1);
    if (own && s - 1 >= I0 && s - 1 < I1) kq = ld_at(K + (s - 1) * S2, own4);
    // ---- stage 0: corrected x, U and label at plane s ----
    if (s >= 0 && s <= Sf - 1) {
      const int sx = s % 3, su = s % RU;
      if (s >= 1 && s <= n) {
#pragma unroll
        for (int q = 0; q < P3; ++q)
          if (cin[q]) xv[q] = xv[q] + 0.5f * (tA[q] + tB[q]);
      }
#pragma unroll
      for (int q = 0; q < P3; ++q) {
        if (tid + q * NT < PL3) {
          rx[sx * PL3 + tid + q * NT] = xv[q];
          ru[su * PL3 + tid + q * NT] = uv[q];
          rl[su * PL3 + tid + q * NT] = lv[q];
        }
      }
    }
    lds_barrier();
    // ---- stage 1: first sweep at plane s - 1 on R2 ----
    const int f1 = s - 1;
    if (f1 >= I0 - 1 && f1 >= 0 && f1 <= Sf - 1) {
      const bool pin = f1 >= 1 && f1 <= n;
      const float* xp = rx + ((f1 + 1) % 3) * PL3;
      const float* x0 = rx + (f1 % 3) * PL3;
      const float* xm = rx + ((f1 + 2) % 3) * PL3;
      const float* u0 = ru + (f1 % RU) * PL3;
#pragma unroll
      for (int q = 0; q < P2; ++q) {
        if (ok2[q]) {
          const int cc = c2[q];
          float v = x0[cc];
          if (pin && in2[q]) {
            const float s6 = xp[cc] + xm[cc] + x0[cc + W3] + x0[cc - W3] + x0[cc + 1] + x0[cc - 1];
            const float u = u0[cc];
            const float rhs = (alpha * (u * u)) * beta;
            float w = (s6 + rhs) / c;
            if (a != 0.0f) w = a * v + w;
            v = w;
          }
          r1[(f1 % 3) * PL2 + tid + q * NT] = v;
        }
      }
    }
    lds_barrier();
    // ---- stage 2: second sweep and ITP update at plane s - 2 on the tile ----
    const int f2 = s - 2;
    if (f2 >= I0 && f2 < I1 && own) {
      const int bp = ((f2 + 1) % 3) * PL2, b0 = (f2 % 3) * PL2, bm = ((f2 + 2) % 3) * PL2;
      const int cc = (tj + 1) * W3 + (tk + 1);
      const float s6 = r1[bp + cc] + r1[bm + cc] + r1[b0 + cc + W3] + r1[b0 + cc - W3] + r1[b0 + cc + 1] +
                       r1[b0 + cc - 1];
      const int cu = (tj + 2) * W3 + (tk + 2);
      const int u0 = (f2 % RU) * PL3 + cu, up_ = ((f2 + 1) % RU) * PL3 + cu, um_ = ((f2 + RU - 1) % RU) * PL3 + cu;
      const float uc = ru[u0];
      const float rhs = (alpha * (uc * uc)) * beta;
      float v = (s6 + rhs) / c;
      if (a != 0.0f) v = a * r1[b0 + cc] + v;
      st_at(xn + f2 * S2, own4, v);
      const uint16_t lc = rl[u0];
      float out;
      if (lc & kRcFlag) {
        out = 0.0f;
      } else {
        // all six neighbours read unconditionally, then selected (a read
        // under the label test would be a branch per neighbour)
        const float n0 = ru[up_], n1 = ru[um_], n2 = ru[u0 + W3], n3 = ru[u0 - W3], n4 = ru[u0 + 1], n5 = ru[u0 - 1];
        const float t0 = (rl[up_] == lc) ? n0 : uc;
        const float t1 = (rl[um_] == lc) ? n1 : uc;
        const float t2 = (rl[u0 + W3] == lc) ? n2 : uc;
        const float t3 = (rl[u0 - W3] == lc) ? n3 : uc;
        const float t4 = (rl[u0 + 1] == lc) ? n4 : uc;
        const float t5 = (rl[u0 - 1] == lc) ? n5 : uc;
        const float lap = t0 + t1 + t2 + t3 + t4 + t5 - 6.0f * uc;
        out = uc + half_dv * lap + dtv * (kv - 2.0f * v) * uc;
      }
      st_at(Un + f2 * S2, own4, out);
    }
    // with four U / label slots the next stage 0 overwrites plane s - 3,
    // which this stage read; with five it writes a slot no stage reads now
    if (RU == 4) lds_barrier();
  }
}

// ---- one launch per coarse level and direction (temporal blocking in LDS) ----
//
// The global loads that stage a tile are unrolled with clamped (always valid)
// indices, so a thread has all of them in flight at once instead of paying
// one memory latency per loop trip.
//
// Down leg of coarse level l (S nodes, zero initial guess, two pre-sweeps):
// e = two sweeps from zero (smooth2z expression), r = g - A e, coarse rhs =
// 4 FW(r).  A workgroup owns CK x CJ x CI coarse interior nodes; the fine
// region R0 their restriction reads, R1 = R0 + 1 (where e is needed) and
// R2 = R0 + 2 (where g is needed) live in LDS.  e is also written to e_out
// on R0 (neighbouring tiles write the same values on their shared plane).
template <int CK, int CJ, int CI>
__global__ __launch_bounds__(256) void k_down2(int S, const float* __restrict__ g, float* __restrict__ e_out,
                                               float* __restrict__ gc, float a, float c) {
  constexpr int NT = 256;
  static_assert(CK * CJ * CI <= NT, "one thread per coarse node of the tile");
  constexpr int K0n = 2 * CK + 1, J0n = 2 * CJ + 1, I0n = 2 * CI + 1;  // R0
  constexpr int K1n = K0n + 2, J1n = J0n + 2, I1n = I0n + 2;           // R1
  constexpr int K2n = K0n + 4, J2n = J0n + 4, I2n = I0n + 4;           // R2
  constexpr int N0 = K0n * J0n * I0n, N1 = K1n * J1n * I1n, N2 = K2n * J2n * I2n;
  constexpr int IT2 = (N2 + NT - 1) / NT;
  // G (and g / c in Gc, one division per node) padded to whole loop trips so
  // the staging stores need no branch
  __shared__ float G[IT2 * NT], Gc[IT2 * NT], E[N1], R[N0];
  const int n = S - 2, nc = (S - 1) / 2 - 1, Sc = nc + 2, S2 = S * S;
  const int gx = (nc + CK - 1) / CK, gy = (nc + CJ - 1) / CJ;
  const int t = xcd_block();
  const int K0 = 1 + (t % gx) * CK, J0 = 1 + ((t / gx) % gy) * CJ, I0 = 1 + (t / (gx * gy)) * CI;
  // fine coordinates of R0's first node; R1 / R2 start 1 / 2 lower
  const int fk = 2 * K0 - 1, fj = 2 * J0 - 1, fi = 2 * I0 - 1;
  const int tid = threadIdx.x;
  {
    float v[IT2];
    unsigned long long inb = 0;
#pragma unroll
    for (int it = 0; it < IT2; ++it) {
      const int q = tid + it * NT;
      const int qq = q < N2 ? q : N2 - 1;
      const int i = fi - 2 + qq / (J2n * K2n), j = fj - 2 + (qq / K2n) % J2n, k = fk - 2 + qq % K2n;
      if (i >= 0 && i < S && j >= 0 && j < S && k >= 0 && k < S) inb |= 1ull << it;
      v[it] = g[min(max(i, 0), S - 1) * S2 + min(max(j, 0), S - 1) * S + min(max(k, 0), S - 1)];
    }
    // keep every load above the first store
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int it = 0; it < IT2; ++it) {
      const float gv = (inb >> it) & 1 ? v[it] : 0.0f;
      G[tid + it * NT] = gv;
      Gc[tid + it * NT] = gv / c;
    }
  }
  __syncthreads();
  constexpr int dI2 = J2n * K2n, dJ2 = K2n;
  for (int q = tid; q < N1; q += NT) {
    const int li = q / (J1n * K1n), lj = (q / K1n) % J1n, lk = q % K1n;
    const int i = fi - 1 + li, j = fj - 1 + lj, k = fk - 1 + lk;
    float v = 0.0f;
    if (i >= 1 && i <= n && j >= 1 && j <= n && k >= 1 && k <= n) {
      const int q2 = (li + 1) * dI2 + (lj + 1) * dJ2 + (lk + 1);
      const float s6 = Gc[q2 + dI2] + Gc[q2 - dI2] + Gc[q2 + dJ2] + Gc[q2 - dJ2] + Gc[q2 + 1] + Gc[q2 - 1];
      v = (s6 + G[q2]) / c;
      if (a != 0.0f) v = a * Gc[q2] + v;
      if (li >= 1 && li <= I0n && lj >= 1 && lj <= J0n && lk >= 1 && lk <= K0n) e_out[i * S2 + j * S + k] = v;
    }
    E[q] = v;
  }
  __syncthreads();
  constexpr int dI1 = J1n * K1n, dJ1 = K1n;
  for (int q = tid; q < N0; q += NT) {
    const int li = q / (J0n * K0n), lj = (q / K0n) % J0n, lk = q % K0n;
    const int i = fi + li, j = fj + lj, k = fk + lk;
    float r = 0.0f;
    if (i <= n && j <= n && k <= n) {
      const int q1 = (li + 1) * dI1 + (lj + 1) * dJ1 + (lk + 1);
      const int q2 = (li + 2) * dI2 + (lj + 2) * dJ2 + (lk + 2);
      const float s6 = E[q1 + dI1] + E[q1 - dI1] + E[q1 + dJ1] + E[q1 - dJ1] + E[q1 + 1] + E[q1 - 1];
      r = G[q2] - (6.0f * E[q1] - s6);
    }
    R[q] = r;
  }
  __syncthreads();
  const int tk = tid % CK, tj = (tid / CK) % CJ, ti = tid / (CK * CJ);
  const int K = K0 + tk, J = J0 + tj, I = I0 + ti;
  if (tid >= CK * CJ * CI || K > nc || J > nc || I > nc) return;
  gc[I * Sc * Sc + J * Sc + K] =
      restrict27(R, (2 * ti + 1) * (J0n * K0n) + (2 * tj + 1) * K0n + (2 * tk + 1), J0n * K0n, K0n, 1);
}

// Up leg of coarse level l: xc = x + P e (faces keep x), then NU sweeps
// (the k_prolong_correct + k_smooth expressions).  A workgroup owns a
// TK x TJ x TI output tile; xc and g are staged on the tile plus a halo of
// NU and each sweep shrinks the valid region by one node.
template <int TK, int TJ, int TI, int NU>
__global__ __launch_bounds__(256) void k_up(int S, const float* __restrict__ x, const float* __restrict__ e,
                                            float* __restrict__ x_out, const float* __restrict__ g, float a,
                                            float c) {
  constexpr int RK = TK + 2 * NU, RJ = TJ + 2 * NU, RI = TI + 2 * NU, NR = RK * RJ * RI;
  constexpr int dI = RJ * RK, dJ = RK, ITR = (NR + 255) / 256;
  // padded to whole loop trips so the staging stores need no branch
  __shared__ float buf[2][ITR * 256], gl[ITR * 256];
  const int n = S - 2, Sc = (S - 1) / 2 + 1, S2 = S * S;
  const int gx = (n + TK - 1) / TK, gy = (n + TJ - 1) / TJ;
  const int t = xcd_block();
  const int k0 = 1 + (t % gx) * TK - NU, j0 = 1 + ((t / gx) % gy) * TJ - NU, i0 = 1 + (t / (gx * gy)) * TI - NU;
  const int tid = threadIdx.x;
  // staged in groups of 4 trips: the loads of a group are issued before any
  // of its stores
#pragma unroll
  for (int it0 = 0; it0 < ITR; it0 += 4) {
    float xv[4], gv[4], pe[4];
    unsigned flags = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = tid + (it0 + u) * 256;
      const int qq = q < NR ? q : NR - 1;
      const int i = i0 + qq / dI, j = j0 + (qq / dJ) % RJ, k = k0 + qq % RK;
      if (i >= 0 && i < S && j >= 0 && j < S && k >= 0 && k < S) flags |= 1u << (2 * u);
      if (i >= 1 && i <= n && j >= 1 && j <= n && k >= 1 && k <= n) flags |= 2u << (2 * u);
      const int id = min(max(i, 0), S - 1) * S2 + min(max(j, 0), S - 1) * S + min(max(k, 0), S - 1);
      xv[u] = x[id];
      gv[u] = g[id];
      pe[u] = prolong_at(e, Sc, min(max(i, 1), n), min(max(j, 1), n), min(max(k, 1), n));
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (it0 + u >= ITR) break;
      const int q = tid + (it0 + u) * 256;
      const bool grid = (flags >> (2 * u)) & 1, inner = (flags >> (2 * u)) & 2;
      buf[0][q] = grid ? (inner ? xv[u] + pe[u] : xv[u]) : 0.0f;
      gl[q] = gv[u];
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 1; s <= NU; ++s) {
    const float* A = buf[(s - 1) & 1];
    float* B = buf[s & 1];
    for (int q = tid; q < NR; q += 256) {
      const int li = q / dI, lj = (q / dJ) % RJ, lk = q % RK;
      if (li < s || li >= RI - s || lj < s || lj >= RJ - s || lk < s || lk >= RK - s) continue;
      const int i = i0 + li, j = j0 + lj, k = k0 + lk;
      if (i < 1 || i > n || j < 1 || j > n || k < 1 || k > n) {
        B[q] = A[q];
        continue;
      }
      const float s6 = A[q + dI] + A[q - dI] + A[q + dJ] + A[q - dJ] + A[q + 1] + A[q - 1];
      float v = (s6 + gl[q]) / c;
      if (a != 0.0f) v = a * A[q] + v;
      if (s == NU)
        x_out[i * S2 + j * S + k] = v;
      else
        B[q] = v;
    }
    if (s < NU) __syncthreads();
  }
}

// Prolongation + correction followed by one smoother sweep:
// x_out = a xc + (sum6 xc + rhs) / c with xc = x + P e on interior nodes
// (faces keep x).  Same values as k_prolong_correct then k_smooth.
template <bool FINE>
__global__ __launch_bounds__(256) void k_prolong_smooth(int Sf, int ich, const float* __restrict__ x,
                                                        const float* __restrict__ e, float* __restrict__ x_out,
                                                        const float* __restrict__ src, float a, float c, float alpha,
                                                        float beta) {
  March m;
  if (!march_coords(Sf - 2, ich, m)) return;
  prolong_smooth_col<FINE>(Sf, m.k, m.j, m.i0, m.i1, x, e, x_out, src, a, c, alpha, beta);
}

// Last fine smoother sweep fused with the ITP update that consumes it:
// P' = a P + (sum6 P + (alpha U^2) beta) / c, then U' = ITP(U, K, P', label)
// with P' taken at the centre (the only place the update reads P).
__global__ __launch_bounds__(256) void k_smooth_itp(int S, int ich, const float* __restrict__ P,
                                                    float* __restrict__ Pn, const float* __restrict__ U,
                                                    float* __restrict__ Un, const float* __restrict__ K,
                                                    const uint16_t* __restrict__ lab, float a, float c, float alpha,
                                                    float beta, float half_dv, float dtv) {
  March m;
  if (!march_coords(S - 2, ich, m)) return;
  smooth_itp_col(S, m.k, m.j, m.i0, m.i1, P, Pn, U, Un, K, lab, a, c, alpha, beta, half_dv, dtv);
}

template <bool SET>
__global__ __launch_bounds__(256) void k_prolong_correct(int Sf, int ich, float* __restrict__ x,
                                                         const float* __restrict__ e, float scale) {
  March m;
  if (!march_coords(Sf - 2, ich, m)) return;
  prolong_correct_col<SET>(Sf, m.k, m.j, m.i0, m.i1, x, e, scale);
}

// Exact coarsest-level solve of 6e - sum6 e = g (zero Dirichlet faces) in
// FP64: e = S D^-1 S g with S the orthonormal type-I DST along each axis.
// One workgroup; the m^3 block lives in LDS; each 1-D transform reads a line
// into registers and writes it back in place.
constexpr int kMaxM = kMaxCoarseInterior;

__device__ void dst_lines(double* A, const double* Sm, int m, int axis) {
  const int nlines = m * m;
  const int stride = (axis == 2) ? 1 : (axis == 1 ? m : m * m);
  for (int line = threadIdx.x; line < nlines; line += blockDim.x) {
    const int a = line / m, b = line % m;
    int base;
    if (axis == 2) base = a * m * m + b * m;
    else if (axis == 1) base = a * m * m + b;
    else base = a * m + b;
    double v[kMaxM];
#pragma unroll
    for (int q = 0; q < kMaxM; ++q) v[q] = (q < m) ? A[base + q * stride] : 0.0;
    for (int p = 0; p < m; ++p) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < kMaxM; ++q)
        if (q < m) s += Sm[p * m + q] * v[q];
      A[base + p * stride] = s;
    }
  }
}

__global__ __launch_bounds__(256) void k_coarse_exact(int Sc, const float* __restrict__ g, float* __restrict__ e) {
  extern __shared__ double lds[];
  const int m = Sc - 2;
  double* Sm = lds;             // m*m
  double* lam = Sm + m * m;     // m
  double* A = lam + m;          // m^3
  const double pi = 3.14159265358979323846;
  const double nrm = sqrt(2.0 / (double)(m + 1));
  for (int t = threadIdx.x; t < m * m; t += blockDim.x) {
    const int p = t / m, q = t % m;
    Sm[t] = nrm * sin(pi * (double)((p + 1) * (q + 1)) / (double)(m + 1));
  }
  for (int t = threadIdx.x; t < m; t += blockDim.x) lam[t] = 2.0 - 2.0 * cos(pi * (double)(t + 1) / (double)(m + 1));
  const int Sc2 = Sc * Sc;
  const int m3 = m * m * m;
  for (int t = threadIdx.x; t < m3; t += blockDim.x) {
    const int a = t / (m * m), b = (t / m) % m, c = t % m;
    A[t] = (double)g[(a + 1) * Sc2 + (b + 1) * Sc + (c + 1)];
  }
  __syncthreads();
  dst_lines(A, Sm, m, 2);
  __syncthreads();
  dst_lines(A, Sm, m, 1);
  __syncthreads();
  dst_lines(A, Sm, m, 0);
  __syncthreads();
  for (int t = threadIdx.x; t < m3; t += blockDim.x) {
    const int a = t / (m * m), b = (t / m) % m, c = t % m;
    A[t] = A[t] / (lam[a] + lam[b] + lam[c]);
  }
  __syncthreads();
  dst_lines(A, Sm, m, 2);
  __syncthreads();
  dst_lines(A, Sm, m, 1);
  __syncthreads();
  dst_lines(A, Sm, m, 0);
  __syncthreads();
  for (int t = threadIdx.x; t < m3; t += blockDim.x) {
    const int a = t / (m * m), b = (t / m) % m, c = t % m;
    e[(a + 1) * Sc2 + (b + 1) * Sc + (c + 1)] = (float)A[t];
  }
}

// One orthonormal type-I DST along `axis` of the m^3 block `in` (LDS), all
// threads of the workgroup producing outputs directly: out[t] = sum_q
// Sm[p][q] in[line(t) + q stride] with even / odd q in two partial sums.
// Used by the single-workgroup bottom; k_coarse_exact_rows evaluates the same
// sums with the matrix row in registers.
__device__ __forceinline__ void dst_axis(const double* __restrict__ in, double* __restrict__ out,
                                         const double* __restrict__ Sm, int m, int axis) {
  const int m2 = m * m, m3 = m2 * m;
  for (int t = threadIdx.x; t < m3; t += blockDim.x) {
    const int a = t / m2, b = (t / m) % m, c = t % m;
    int p, base, stride;
    if (axis == 2) {
      p = c;
      base = a * m2 + b * m;
      stride = 1;
    } else if (axis == 1) {
      p = b;
      base = a * m2 + c;
      stride = m;
    } else {
      p = a;
      base = b * m + c;
      stride = m2;
    }
    const double* row = Sm + p * m;
    double s0 = 0.0, s1 = 0.0;
    int q = 0;
    for (; q + 1 < m; q += 2) {
      s0 += row[q] * in[base + q * stride];
      s1 += row[q + 1] * in[base + (q + 1) * stride];
    }
    if (q < m) s0 += row[q] * in[base + q * stride];
    out[t] = s0 + s1;
  }
}

// The last two V-cycle levels in one workgroup, everything in LDS: on level L
// (S <= kBottomMaxS nodes per axis) pre-smoothing, residual, 4 FW restriction,
// the exact FP64 DST solve of level L+1, prolongation + correction and
// post-smoothing; only the level-L correction is written back.  Each step is
// the expression of the multi-launch sequence (k_smooth2_from_zero / k_smooth
// / k_residual / k_restrict4 / k_coarse_exact_rows / k_prolong_correct) in the
// same order, so the result is bit-identical to it; Jacobi sweeps run in
// place (new values held in registers across a barrier).
// LDS: A, B (m^3 f64), Sm (m^2 f64), E, R (S^3 f32), G (Sc^3 f32).
constexpr int kBotThreads = 1024;
constexpr int kBotPer = ((kBottomMaxS - 2) * (kBottomMaxS - 2) * (kBottomMaxS - 2) + kBotThreads - 1) / kBotThreads;
constexpr int kBotGroup = 4;  // nodes whose loads the scheduler may interleave

// Every thread owns kBotPer interior nodes (packed i | j << 8 | k << 16); a
// thread with fewer real nodes repeats node (1, 1, 1) for the rest and masks
// their stores, so the per-node loops carry no branches.
__device__ __forceinline__ int bot_id(int p, int S) {
  return (p & 0xff) * S * S + ((p >> 8) & 0xff) * S + (p >> 16);
}

__global__ __launch_bounds__(kBotThreads) void k_bottom(int S, int nu1, int nu2, float a, float c,
                                                        const float* __restrict__ g, float* __restrict__ e_out,
                                                        const double* __restrict__ Smat,
                                                        const double* __restrict__ rden) {
  extern __shared__ double lds[];
  const int Sc = (S - 1) / 2 + 1;
  const int m = Sc - 2, m2 = m * m, m3 = m2 * m;
  const int S2 = S * S, S3 = S2 * S, Sc2 = Sc * Sc;
  const int n = S - 2, n2 = n * n, n3 = n2 * n;
  double* A = lds;
  double* B = A + m3;
  double* Sm = B + m3;
  float* cur = reinterpret_cast<float*>(Sm + m2);  // current iterate (faces 0)
  float* oth = cur + S3;                            // sweep target / residual (faces 0)
  float* G = oth + S3;                              // level L+1: rhs, then solution (faces 0)
  const int tid = threadIdx.x;
  int pk[kBotPer];
  float gr[kBotPer];
  unsigned live = 0;
#pragma unroll
  for (int q = 0; q < kBotPer; ++q) {
    const int p = tid + q * kBotThreads;
    if (p < n3) {
      pk[q] = (p / n2 + 1) | (((p / n) % n + 1) << 8) | ((p % n + 1) << 16);
      live |= 1u << q;
    } else {
      pk[q] = 1 | (1 << 8) | (1 << 16);
    }
  }
#pragma unroll
  for (int q = 0; q < kBotPer; ++q) gr[q] = g[bot_id(pk[q], S)];
  // reciprocal eigenvalue sums of this thread's transform outputs, loaded early
  const double rd0 = tid < m3 ? rden[tid] : 0.0;
  const double rd1 = tid + kBotThreads < m3 ? rden[tid + kBotThreads] : 0.0;
  for (int t = tid; t < 2 * S3; t += kBotThreads) cur[t] = 0.0f;
  for (int t = tid; t < Sc * Sc2; t += kBotThreads) G[t] = 0.0f;
  for (int t = tid; t < m2; t += kBotThreads) Sm[t] = Smat[t];
  __syncthreads();
  // pre-smoothing from a zero guess (k_smooth2_from_zero, then k_smooth)
  int pre = nu1;
  if (pre >= 2) {
#pragma unroll
    for (int q = 0; q < kBotPer; ++q) {
      const int id = bot_id(pk[q], S);
      const float e_ip = g[id + S2] / c, e_im = g[id - S2] / c, e_jp = g[id + S] / c, e_jm = g[id - S] / c;
      const float e_kp = g[id + 1] / c, e_km = g[id - 1] / c;
      const float gc = gr[q];
      const float s6 = e_ip + e_im + e_jp + e_jm + e_kp + e_km;
      float w = (s6 + gc) / c;
      if (a != 0.0f) w = a * (gc / c) + w;
      if (live & (1u << q)) cur[id] = w;
      if (q % kBotGroup == kBotGroup - 1) __builtin_amdgcn_sched_barrier(0);
    }
    pre -= 2;
    __syncthreads();
  }
  auto sweep = [&]() {
#pragma unroll
    for (int q = 0; q < kBotPer; ++q) {
      const int id = bot_id(pk[q], S);
      const float s6 = cur[id + S2] + cur[id - S2] + cur[id + S] + cur[id - S] + cur[id + 1] + cur[id - 1];
      float w = (s6 + gr[q]) / c;
      if (a != 0.0f) w = a * cur[id] + w;
      if (live & (1u << q)) oth[id] = w;
      if (q % kBotGroup == kBotGroup - 1) __builtin_amdgcn_sched_barrier(0);
    }
    float* t = cur;
    cur = oth;
    oth = t;
    __syncthreads();
  };
  for (int it = 0; it < pre; ++it) sweep();
  // residual (k_residual), then 4 FW onto the coarse interior (k_restrict4)
#pragma unroll
  for (int q = 0; q < kBotPer; ++q) {
    const int id = bot_id(pk[q], S);
    const float s6 = cur[id + S2] + cur[id - S2] + cur[id + S] + cur[id - S] + cur[id + 1] + cur[id - 1];
    const float r = gr[q] - (6.0f * cur[id] - s6);
    if (live & (1u << q)) oth[id] = r;
    if (q % kBotGroup == kBotGroup - 1) __builtin_amdgcn_sched_barrier(0);
  }
  __syncthreads();
  for (int t = tid; t < m3; t += kBotThreads)
    restrict4_pt(S, t / m2 + 1, (t / m) % m + 1, t % m + 1, oth, G, 1.0f);
  __syncthreads();
  // exact solve on level L+1 (the transforms of k_coarse_exact_rows)
  for (int t = tid; t < m3; t += kBotThreads) {
    const int a3 = t / m2, b3 = (t / m) % m, c3 = t % m;
    A[t] = (double)G[(a3 + 1) * Sc2 + (b3 + 1) * Sc + (c3 + 1)];
  }
  __syncthreads();
  dst_axis(A, B, Sm, m, 2);
  __syncthreads();
  dst_axis(B, A, Sm, m, 1);
  __syncthreads();
  dst_axis(A, B, Sm, m, 0);
  __syncthreads();
  if (tid < m3) B[tid] *= rd0;
  if (tid + kBotThreads < m3) B[tid + kBotThreads] *= rd1;
  __syncthreads();
  dst_axis(B, A, Sm, m, 2);
  __syncthreads();
  dst_axis(A, B, Sm, m, 1);
  __syncthreads();
  dst_axis(B, A, Sm, m, 0);
  __syncthreads();
  for (int t = tid; t < m3; t += kBotThreads) {
    const int a3 = t / m2, b3 = (t / m) % m, c3 = t % m;
    G[(a3 + 1) * Sc2 + (b3 + 1) * Sc + (c3 + 1)] = (float)A[t];
  }
  __syncthreads();
  // prolongation + correction in place (k_prolong_correct, scale 1; faces keep 0)
#pragma unroll
  for (int q = 0; q < kBotPer; ++q) {
    const int i = pk[q] & 0xff, j = (pk[q] >> 8) & 0xff, k = pk[q] >> 16;
    const int id = bot_id(pk[q], S);
    const float x = cur[id] + prolong_at(G, Sc, i, j, k);
    if (live & (1u << q)) cur[id] = x;
    if (q % kBotGroup == kBotGroup - 1) __builtin_amdgcn_sched_barrier(0);
  }
  __syncthreads();
  for (int it = 0; it < nu2; ++it) sweep();
#pragma unroll
  for (int q = 0; q < kBotPer; ++q) {
    const int id = bot_id(pk[q], S);
    if (live & (1u << q)) e_out[id] = cur[id];
  }
}

// Exact coarsest solve with the orthonormal type-I DST matrix Sm (m x m,
// symmetric, m <= kMaxCoarsePP) and the reciprocal eigenvalue sums rden (m^3)
// precomputed on the host: e = S D^-1 S g along the three axes.  The
// workgroup is (m, L) threads and thread (p, l) produces output p of lines l,
// l + L, ... on every axis, so its DST matrix row Sm[p][*] is loaded into
// registers once for all six transforms and only the input lines are read
// from LDS.  The 1 / eigenvalue-sum scaling (prefetched at the start) is
// applied as the third transform stores.  M > 0 fixes m at compile time so
// each line's loads are issued together; M = 0 takes m at run time.  Same
// products and sums, in the same order, as dst_axis.
constexpr int kExactLines = 8;  // lines per thread: ceil(m^2 / floor(1024 / m)) for m <= 19

template <int M>
__global__ __launch_bounds__(1024) void k_coarse_exact_rows(int Sc, const float* __restrict__ g,
                                                            float* __restrict__ e, const double* __restrict__ Smat,
                                                            const double* __restrict__ rden) {
  extern __shared__ double lds[];
  constexpr int QM = M > 0 ? M : kMaxCoarsePP;
  const int m = M > 0 ? M : Sc - 2;
  const int m2 = m * m, m3 = m2 * m;
  double* A = lds;
  double* B = A + m3;
  const int p = threadIdx.x, L = blockDim.y;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x, nt = blockDim.x * blockDim.y;
  double row[QM];
#pragma unroll
  for (int q = 0; q < QM; ++q) row[q] = (M > 0 || q < m) ? Smat[p * m + q] : 0.0;
  double rd[kExactLines];
#pragma unroll
  for (int it = 0; it < kExactLines; ++it) {
    const int l = threadIdx.y + it * L;
    rd[it] = l < m2 ? rden[(l / m) * m + (l % m) + p * m2] : 0.0;
  }
  const int Sc2 = Sc * Sc;
  for (int t = tid; t < m3; t += nt) {
    const int a = t / m2, b = (t / m) % m, c = t % m;
    A[t] = (double)g[(a + 1) * Sc2 + (b + 1) * Sc + (c + 1)];
  }
  __syncthreads();
  auto pass = [&](const double* __restrict__ in, double* __restrict__ out, int axis, bool scale) {
#pragma unroll
    for (int it = 0; it < kExactLines; ++it) {
      const int l = threadIdx.y + it * L;
      if (l >= m2) break;
      const int u = l / m, w = l % m;
      int base, stride, o;
      if (axis == 2) {
        base = u * m2 + w * m;
        stride = 1;
        o = base + p;
      } else if (axis == 1) {
        base = u * m2 + w;
        stride = m;
        o = base + p * m;
      } else {
        base = u * m + w;
        stride = m2;
        o = base + p * m2;
      }
      double s0 = 0.0, s1 = 0.0;
      if (M > 0) {
#pragma unroll
        for (int q = 0; q + 1 < QM; q += 2) {
          s0 += row[q] * in[base + q * stride];
          s1 += row[q + 1] * in[base + (q + 1) * stride];
        }
        if (QM & 1) s0 += row[QM - 1] * in[base + (QM - 1) * stride];
      } else {
        int q = 0;
        for (; q + 1 < m; q += 2) {
          s0 += row[q] * in[base + q * stride];
          s1 += row[q + 1] * in[base + (q + 1) * stride];
        }
        if (q < m) s0 += row[q] * in[base + q * stride];
      }
      double v = s0 + s1;
      if (scale) v *= rd[it];
      out[o] = v;
    }
  };
  pass(A, B, 2, false);
  __syncthreads();
  pass(B, A, 1, false);
  __syncthreads();
  pass(A, B, 0, true);
  __syncthreads();
  pass(B, A, 2, false);
  __syncthreads();
  pass(A, B, 1, false);
  __syncthreads();
  pass(B, A, 0, false);
  __syncthreads();
  for (int t = tid; t < m3; t += nt) {
    const int a = t / m2, b = (t / m) % m, c = t % m;
    e[(a + 1) * Sc2 + (b + 1) * Sc + (c + 1)] = (float)A[t];
  }
}

// ---- direct per-domain potential (molecule.js "direct Pother") ----
// rhs_m = (alpha U^2) beta on interior nodes labelled != m, 0 on nodes of m
// (molecule.js:597-621 computeRhoOther, in the V-cycle's 2 pi h^2 scaling)
__global__ __launch_bounds__(256) void k_rho_other(int S, int ich, const float* __restrict__ U,
                                                   const uint16_t* __restrict__ lab, int m, float alpha, float beta,
                                                   float* __restrict__ rhs) {
  March q;
  if (!march_coords(S - 2, ich, q)) return;
  const int S2 = S * S;
  for (int i = q.i0; i < q.i1; ++i) {
    const int id = i * S2 + q.j * S + q.k;
    const float u = U[id];
    rhs[id] = ((lab[id] & kLabelMask) != m) ? (alpha * (u * u)) * beta : 0.0f;
  }
}

// molecule.js initPdirect (:624-662): the start of domain m's persistent
// direct potential, P_m = sum over the other atoms with Z > 0 of
// 0.5 / sqrt(r^2 + h^2) on the interior (positions as u32 with floor_pos),
// also added into the frame potential.
__global__ __launch_bounds__(256) void k_pdirect_init(int S, int ich, float h, float h2, const Atom* __restrict__ atoms,
                                                      int n_atoms, int m, int floor_pos, float* __restrict__ Pm,
                                                      float* __restrict__ Pother) {
  March q;
  if (!march_coords(S - 2, ich, q)) return;
  const int S2 = S * S;
  const float yj = (float)q.j * h, zk = (float)q.k * h;
  for (int i = q.i0; i < q.i1; ++i) {
    const int id = i * S2 + q.j * S + q.k;
    const float xi = (float)i * h;
    float val = 0.0f;
    for (int n = 0; n < n_atoms; ++n) {
      if (atoms[n].Z <= 0.0f || n == m) continue;
      const float ai = floor_pos ? floorf(atoms[n].i) : atoms[n].i;
      const float aj = floor_pos ? floorf(atoms[n].j) : atoms[n].j;
      const float ak = floor_pos ? floorf(atoms[n].k) : atoms[n].k;
      const float dx = xi - ai * h, dy = yj - aj * h, dz = zk - ak * h;
      const float r = sqrtf(dx * dx + dy * dy + dz * dz + h2);
      val += 0.5f / r;
    }
    Pm[id] = val;
    Pother[id] += val;
  }
}

// molecule.js copyPotherForLabel (:664-688): Pother = P_m on the nodes labelled m
__global__ __launch_bounds__(256) void k_copy_for_label(int S, int ich, const float* __restrict__ Pm,
                                                        const uint16_t* __restrict__ lab, int m,
                                                        float* __restrict__ Pother) {
  March q;
  if (!march_coords(S - 2, ich, q)) return;
  const int S2 = S * S;
  for (int i = q.i0; i < q.i1; ++i) {
    const int id = i * S2 + q.j * S + q.k;
    if ((lab[id] & kLabelMask) == m) Pother[id] = Pm[id];
  }
}

// rhs of domain m's self-potential: (alpha u^2) beta where label == m, else 0
// (molecule.js:572-593 computeRhoSelf, in the build's rhs form)
__global__ __launch_bounds__(256) void k_rho_self(int S, int ich, const float* __restrict__ U,
                                                  const uint16_t* __restrict__ lab, int m, float alpha, float beta,
                                                  float* __restrict__ rhs) {
  March q;
  if (!march_coords(S - 2, ich, q)) return;
  const int S2 = S * S;
  for (int i = q.i0; i < q.i1; ++i) {
    const int id = i * S2 + q.j * S + q.k;
    const float u = U[id];
    rhs[id] = ((lab[id] & kLabelMask) == m) ? (alpha * (u * u)) * beta : 0.0f;
  }
}

// Pother -= frac * P_self on domain m's nodes (molecule.js:694-721 subtractPself)
__global__ __launch_bounds__(256) void k_sub_pself(int S, int ich, const float* __restrict__ Pself,
                                                   const uint16_t* __restrict__ lab, int m, float frac,
                                                   float* __restrict__ Pother) {
  March q;
  if (!march_coords(S - 2, ich, q)) return;
  const int S2 = S * S;
  for (int i = q.i0; i < q.i1; ++i) {
    const int id = i * S2 + q.j * S + q.k;
    if ((lab[id] & kLabelMask) == m) Pother[id] = Pother[id] - Pself[id] * frac;
  }
}

// P(x) = P_label(x)(x) on the interior (molecule.js:664-688 copyPotherForLabel)
// The ITP update (itp_col) with the direct potential gathered on the fly:
// the centre P is the node's own domain potential Pdom[label] (0 past the
// domains), also stored into P -- the values of k_pother_gather then k_itp
// in one pass, without re-reading the labels and P.
__global__ __launch_bounds__(256) void k_itp_gather(int S, int ich, const float* __restrict__ U,
                                                    float* __restrict__ Un, const float* __restrict__ K,
                                                    const uint16_t* __restrict__ lab,
                                                    const float* const* __restrict__ Pdom, int n_dom,
                                                    float* __restrict__ P, float half_dv, float dtv) {
  March q;
  if (!march_coords(S - 2, ich, q)) return;
  const int S2 = S * S;
  int id = q.i0 * S2 + q.j * S + q.k;
  float um = U[id - S2], uc = U[id];
  uint16_t lm = lab[id - S2], lc = lab[id];
  for (int i = q.i0; i < q.i1; ++i, id += S2) {
    const float up = U[id + S2];
    const uint16_t lp = lab[id + S2];
    const int L = lc & kLabelMask;
    const float pv = (L < n_dom) ? Pdom[L][id] : 0.0f;
    P[id] = pv;
    float out;
    if (lc & kRcFlag) {
      out = 0.0f;
    } else {
      const float ujp = U[id + S], ujm = U[id - S], ukp = U[id + 1], ukm = U[id - 1];
      const uint16_t ljp = lab[id + S], ljm = lab[id - S], lkp = lab[id + 1], lkm = lab[id - 1];
      const float t0 = (lp == lc) ? up : uc;
      const float t1 = (lm == lc) ? um : uc;
      const float t2 = (ljp == lc) ? ujp : uc;
      const float t3 = (ljm == lc) ? ujm : uc;
      const float t4 = (lkp == lc) ? ukp : uc;
      const float t5 = (lkm == lc) ? ukm : uc;
      const float lap = t0 + t1 + t2 + t3 + t4 + t5 - 6.0f * uc;
      out = uc + half_dv * lap + dtv * (K[id] - 2.0f * pv) * uc;
    }
    Un[id] = out;
    um = uc;
    uc = up;
    lm = lc;
    lc = lp;
  }
}

__global__ __launch_bounds__(256) void k_pother_gather(int S, int ich, const uint16_t* __restrict__ lab,
                                                       const float* const* __restrict__ Pdom, int n_dom,
                                                       float* __restrict__ P) {
  March q;
  if (!march_coords(S - 2, ich, q)) return;
  const int S2 = S * S;
  for (int i = q.i0; i < q.i1; ++i) {
    const int id = i * S2 + q.j * S + q.k;
    const int L = lab[id] & kLabelMask;
    P[id] = (L < n_dom) ? Pdom[L][id] : 0.0f;
  }
}

// Display slice at k = sk (molecule.js:1108-1192 extract, W-cutoff 0):
// out[0 .. S^2) density U^2, out[S^2 .. 2S^2) label + 1000 Z of the label's
// atom, out[2S^2 .. 3S^2) boundary marks (1 = labels differ along +i or +j
// with density on either side, 2 = they differ with none), then the K line
// K(i, sk, sk) and the cutoff-weight line (1 everywhere).
__global__ void k_display_slice(int S, int sk, const float* __restrict__ U, const uint16_t* __restrict__ lab,
                                const float* __restrict__ K, const Atom* __restrict__ atoms, int n_atoms,
                                float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
  const int NN = S - 1, S2 = S * S, SS = S * S;
  if (i > NN || j > NN) return;
  if (j == 0) {
    out[3 * SS + i] = K[i * S2 + sk * S + sk];
    out[3 * SS + S + i] = 1.0f;
  }
  if (i < 1 || i >= NN || j < 1 || j >= NN) {
    out[i * S + j] = 0.0f;
    out[SS + i * S + j] = 0.0f;
    out[2 * SS + i * S + j] = 0.0f;
    return;
  }
  const int idx = i * S2 + j * S + sk;
  const float u = U[idx];
  const int lbl = lab[idx] & kLabelMask;
  const float Zl = lbl < n_atoms ? atoms[lbl].Z : 0.0f;
  out[i * S + j] = u * u;
  out[SS + i * S + j] = (float)lbl + Zl * 1000.0f;
  const float dens = u * u;
  float bnd = 0.0f;
  if (i > 1 && i < NN - 1) {
    const int nb = idx + S2;
    if (lbl != (lab[nb] & kLabelMask)) {
      const float nd = U[nb] * U[nb];
      bnd = (dens > 1e-6f || nd > 1e-6f) ? 1.0f : 2.0f;
    }
  }
  if (bnd < 1.5f && j > 1 && j < NN - 1) {
    const int nb = idx + S;
    if (lbl != (lab[nb] & kLabelMask)) {
      const float nd = U[nb] * U[nb];
      bnd = (dens > 1e-6f || nd > 1e-6f) ? fmaxf(bnd, 1.0f) : 2.0f;
    }
  }
  out[2 * SS + i * S + j] = bnd;
}

__global__ void k_fill(float* __restrict__ p, size_t n, float v) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

__global__ void k_square(int S3, const float* __restrict__ U, float* __restrict__ rho) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < S3) {
    const float u = U[t];
    rho[t] = u * u;
  }
}

__global__ void k_lab16to32(size_t n, const uint16_t* __restrict__ in, int32_t* __restrict__ out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = (int32_t)(in[t] & kLabelMask);
}

__global__ void k_lab32to16(size_t n, const int32_t* __restrict__ in, uint16_t* __restrict__ out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    int v = in[t];
    v = v < 0 ? 0 : (v > (int)kLabelMask ? (int)kLabelMask : v);
    out[t] = (uint16_t)v;
  }
}

// --------------------------------------------------------------------------
// Per-domain reductions.  Each thread accumulates a run of equal labels in
// FP64 registers, flushing to the workgroup's LDS bins when the label
// changes; after a barrier each non-zero bin is added to the global result
// once per workgroup.  Bins and results are two-word fixed-point sums
// (fx_add, fx_value), so the totals do not depend on the order in which
// waves and workgroups add to them.  Labels are clipped to [0, n_dom-1]
// (cloud/solver.py:427 ``label.clip``).
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_norm_sums(int S, int ich, float h3, const float* __restrict__ U,
                                                   const uint16_t* __restrict__ lab, int n_dom,
                                                   Fx* __restrict__ norms, int* __restrict__ bad) {
  extern __shared__ Fx fbins[];
  const int nthr = blockDim.x * blockDim.y;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  for (int t = tid; t < kFxWords * n_dom; t += nthr) fbins[t] = 0;
  __syncthreads();
  March m;
  if (march_coords(S - 2, ich, m)) norm_sums_col(S, m.k, m.j, m.i0, m.i1, h3, U, lab, n_dom, fbins, bad);
  __syncthreads();
  for (int t = tid; t < n_dom; t += nthr) fx_add_words(norms + kFxWords * t, fbins + kFxWords * t);
}

// per-domain scale sqrt(Z / n) (1 for empty or uncharged domains), once per
// normalisation rather than in every workgroup of the scaling pass
__global__ __launch_bounds__(256) void k_norm_scales(int n_dom, const Fx* __restrict__ norms,
                                                     const float* __restrict__ Z, float* __restrict__ scale) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_dom) return;
  const double z = (double)Z[t];
  const double nv = fx_value(norms + kFxWords * t);
  scale[t] = (z > 0.0 && nv > 1e-20) ? (float)sqrt(z / nv) : 1.0f;
}

__global__ __launch_bounds__(256) void k_normalize(int S, int ich, float* __restrict__ U,
                                                   const uint16_t* __restrict__ lab, const float* __restrict__ scale,
                                                   int n_dom) {
  March m;
  if (!march_coords(S - 2, ich, m)) return;
  normalize_col(S, m.k, m.j, m.i0, m.i1, U, lab, scale, n_dom);
}

// --------------------------------------------------------------------------
// Boundary rules.
// --------------------------------------------------------------------------
// The score rule's decision at a node with enough foreign neighbours, from
// the node's best atom bN (the first strict maximum of Z_a / sqrt(r_a^2 +
// h^2) over atoms with Z > 0, or kBestNone): flip to bN if it is not the
// node's own domain and its score beats the own atom's score by `ratio`
// (cloud/solver.py:128-145).  The best atom depends only on the geometry, so
// the kernels keep it per node (best_atom, kBestUnknown until first needed
// after the atoms were last set) and each later application of the rule
// evaluates two scores instead of searching the atoms again.
constexpr int kBestUnknown = 0xFFFF, kBestNone = 0xFFFE;
template <class Score>
__device__ __forceinline__ uint16_t flip_label(uint16_t raw, int cur, int bN, int n_atoms,
                                               const Atom* __restrict__ atoms, float ratio, Score score) {
  if (bN == kBestNone || bN == cur) return raw;
  const float best = score(bN);
  const float cur_score = (cur < n_atoms && atoms[cur].Z > 0.0f) ? score(cur) : 0.0f;
  return best > cur_score * ratio ? (uint16_t)bN : raw;
}

// Best-atom map of the score rule for every interior node, one pass per
// geometry (the rule's own search then never runs).  A workgroup owns a
// block of kMapB^3 nodes.  It keeps the atoms that can be the best anywhere
// in the block -- those whose score bound at the block's nearest point
// reaches the largest bound of any atom at its farthest point (with 1e-4
// slack for FP32 rounding) -- in index order; each node then takes the first
// strict maximum of Z / sqrt(r^2 + h^2) over that list, with the expression
// of the rule's scan.  An atom left out cannot reach the block maximum, so
// the result is the scan's.  More than kMapMaxCand candidates: the block is
// left unknown and the rule searches those nodes itself.
constexpr int kMapB = 8, kMapMaxCand = 512;
__global__ __launch_bounds__(256) void k_score_map(int S, float h, float h2, const Atom* __restrict__ atoms,
                                                   int n_atoms, uint16_t* __restrict__ best_atom) {
  __shared__ float wred[4];
  __shared__ int wcnt[4];
  __shared__ int cand[kMapMaxCand];
  __shared__ int ncand;
  const int n = S - 2, nb = (n + kMapB - 1) / kMapB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int b = blockIdx.x;
  const int i0 = 1 + (b / (nb * nb)) * kMapB, j0 = 1 + ((b / nb) % nb) * kMapB, k0 = 1 + (b % nb) * kMapB;
  const int i1 = min(i0 + kMapB, n + 1), j1 = min(j0 + kMapB, n + 1), k1 = min(k0 + kMapB, n + 1);
  const float xl = (float)i0 * h, xh = (float)(i1 - 1) * h, yl = (float)j0 * h, yh = (float)(j1 - 1) * h;
  const float zl = (float)k0 * h, zh = (float)(k1 - 1) * h;
  // largest lower bound (farthest block point) over the atoms
  float mL = 0.0f;
  for (int a = tid; a < n_atoms; a += 256) {
    const Atom at = atoms[a];
    if (!(at.Z > 0.0f)) continue;
    const float ax = at.i * h, ay = at.j * h, az = at.k * h;
    const float dx = fmaxf(fabsf(xl - ax), fabsf(xh - ax)), dy = fmaxf(fabsf(yl - ay), fabsf(yh - ay));
    const float dz = fmaxf(fabsf(zl - az), fabsf(zh - az));
    mL = fmaxf(mL, at.Z / sqrtf(dx * dx + dy * dy + dz * dz + h2));
  }
  for (int off = 32; off > 0; off >>= 1) mL = fmaxf(mL, __shfl_xor(mL, off, 64));
  if (lane == 0) wred[wave] = mL;
  if (tid == 0) ncand = 0;
  __syncthreads();
  const float thr = fmaxf(fmaxf(wred[0], wred[1]), fmaxf(wred[2], wred[3])) * 0.9999f;
  // candidates in index order: 256 atoms per round, appended in order
  for (int a0 = 0; a0 < n_atoms; a0 += 256) {
    const int a = a0 + tid;
    bool keep = false;
    if (a < n_atoms) {
      const Atom at = atoms[a];
      if (at.Z > 0.0f) {
        const float ax = at.i * h, ay = at.j * h, az = at.k * h;
        const float dx = fmaxf(0.0f, fmaxf(xl - ax, ax - xh)), dy = fmaxf(0.0f, fmaxf(yl - ay, ay - yh));
        const float dz = fmaxf(0.0f, fmaxf(zl - az, az - zh));
        keep = at.Z / sqrtf(dx * dx + dy * dy + dz * dz + h2) >= thr;
      }
    }
    const unsigned long long m = __ballot(keep);
    if (lane == 0) wcnt[wave] = __popcll(m);
    __syncthreads();
    const int base = ncand;
    int before = 0;
    for (int w = 0; w < wave; ++w) before += wcnt[w];
    if (keep) {
      const int pos = base + before + __popcll(m & ((1ull << lane) - 1ull));
      if (pos < kMapMaxCand) cand[pos] = a;
    }
    __syncthreads();
    if (tid == 0) ncand = base + wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
    __syncthreads();
  }
  const int nc = ncand;
  const int S2 = S * S;
  for (int q = tid; q < kMapB * kMapB * kMapB; q += 256) {
    const int i = i0 + q / (kMapB * kMapB), j = j0 + (q / kMapB) % kMapB, k = k0 + q % kMapB;
    if (i >= i1 || j >= j1 || k >= k1) continue;
    int bN = kBestUnknown;
    if (nc <= kMapMaxCand) {
      const float xi = (float)i * h, yj = (float)j * h, zk = (float)k * h;
      float best = 0.0f;
      bN = kBestNone;
      for (int c = 0; c < nc; ++c) {
        const int a = cand[c];
        const float dx = xi - atoms[a].i * h;
        const float dy = yj - atoms[a].j * h;
        const float dz = zk - atoms[a].k * h;
        const float r2 = dx * dx + dy * dy + dz * dz + h2;
        const float sc = atoms[a].Z / sqrtf(r2);
        if (sc > best) {
          best = sc;
          bN = a;
        }
      }
    }
    best_atom[i * S2 + j * S + k] = (uint16_t)bN;
  }
}

// Nuclear-score flip (cloud/solver.py:92-147), reading labels from `lin`
// and writing `lout` (double-buffered).
// Score rule by a scan over all atoms (cloud/solver.py:92-147; the form for
// small molecules, where a scan of a few dozen atoms beats any search).
__global__ __launch_bounds__(256) void k_boundary_scan(int S, int ich, float h, float h2,
                                                       const Atom* __restrict__ atoms, int n_atoms,
                                                       const uint16_t* __restrict__ lin, uint16_t* __restrict__ lout,
                                                       float ratio, int min_foreign, uint16_t* __restrict__ best_atom) {
  const int n = S - 2;
  March m;
  if (!march_coords(n, ich, m)) return;
  const int S2 = S * S;
  for (int i = m.i0; i < m.i1; ++i) {
    const int id = i * S2 + m.j * S + m.k;
    const uint16_t raw = lin[id];
    const int cur = raw & kLabelMask;
    int other = 0;
    other += ((lin[id + S2] & kLabelMask) != cur);
    other += ((lin[id - S2] & kLabelMask) != cur);
    other += ((lin[id + S] & kLabelMask) != cur);
    other += ((lin[id - S] & kLabelMask) != cur);
    other += ((lin[id + 1] & kLabelMask) != cur);
    other += ((lin[id - 1] & kLabelMask) != cur);
    uint16_t out = raw;
    // a flip needs other >= min_foreign (cloud/solver.py:144), so nodes with
    // fewer foreign neighbours keep their label without the score search
    if (other > 0 && other >= min_foreign) {
      const float xi = (float)i * h, yj = (float)m.j * h, zk = (float)m.k * h;
      auto score = [&](int a) -> float {
        const float dx = xi - atoms[a].i * h;
        const float dy = yj - atoms[a].j * h;
        const float dz = zk - atoms[a].k * h;
        const float r2 = dx * dx + dy * dy + dz * dz + h2;
        return atoms[a].Z / sqrtf(r2);
      };
      int bN = best_atom[id];
      if (bN == kBestUnknown) {
        // the reference's scan: the first strict maximum among atoms with Z > 0
        float best = 0.0f;
        bN = kBestNone;
        for (int a = 0; a < n_atoms; ++a) {
          if (!(atoms[a].Z > 0.0f)) continue;
          const float sc = score(a);
          if (sc > best) {
            best = sc;
            bN = a;
          }
        }
        best_atom[id] = (uint16_t)bN;
      }
      out = flip_label(raw, cur, bN, n_atoms, atoms, ratio, score);
    }
    lout[id] = out;
  }
}

__global__ __launch_bounds__(256) void k_boundary_score(int S, int ich, float h, float h2,
                                                        const Atom* __restrict__ atoms, int n_atoms,
                                                        const uint16_t* __restrict__ lin,
                                                        uint16_t* __restrict__ lout, float ratio, int min_foreign,
                                                        int cs, int ncell, float zmax,
                                                        const int32_t* __restrict__ cell_start,
                                                        const int32_t* __restrict__ cell_atoms,
                                                        uint16_t* __restrict__ best_atom) {
  const int n = S - 2;
  March m;
  if (!march_coords(n, ich, m)) return;
  const int S2 = S * S;
  for (int i = m.i0; i < m.i1; ++i) {
    const int id = i * S2 + m.j * S + m.k;
    const uint16_t raw = lin[id];
    const int cur = raw & kLabelMask;
    int other = 0;
    other += ((lin[id + S2] & kLabelMask) != cur);
    other += ((lin[id - S2] & kLabelMask) != cur);
    other += ((lin[id + S] & kLabelMask) != cur);
    other += ((lin[id - S] & kLabelMask) != cur);
    other += ((lin[id + 1] & kLabelMask) != cur);
    other += ((lin[id - 1] & kLabelMask) != cur);
    uint16_t out = raw;
    // a flip needs other >= min_foreign (cloud/solver.py:144), so nodes with
    // fewer foreign neighbours keep their label without the score search
    if (other > 0 && other >= min_foreign) {
      const float xi = (float)i * h, yj = (float)m.j * h, zk = (float)m.k * h;
      // score of atom a: Z / sqrt(|x - R_a|^2 + h^2) (cloud/solver.py:110-128)
      auto score = [&](int a) -> float {
        const float dx = xi - atoms[a].i * h;
        const float dy = yj - atoms[a].j * h;
        const float dz = zk - atoms[a].k * h;
        const float r2 = dx * dx + dy * dy + dz * dz + h2;
        return atoms[a].Z / sqrtf(r2);
      };
      // the reference scans all atoms in index order keeping the first
      // strict maximum; here the candidates come from shells of cells, so a
      // tie keeps the lower index -- the same winner.
      int bN = best_atom[id];
      if (bN == kBestUnknown) {
        float best = 0.0f;
        int bestN = kBestNone;
        bool found = false;
        const int hi = min(i / cs, ncell - 1), hj = min(m.j / cs, ncell - 1), hk = min(m.k / cs, ncell - 1);
        const int rmax = max(max(max(hi, ncell - 1 - hi), max(hj, ncell - 1 - hj)), max(hk, ncell - 1 - hk));
        for (int r = 0; r <= rmax; ++r) {
          for (int ci = max(hi - r, 0); ci <= min(hi + r, ncell - 1); ++ci) {
            for (int cj = max(hj - r, 0); cj <= min(hj + r, ncell - 1); ++cj) {
              const bool edge_ij = abs(ci - hi) == r || abs(cj - hj) == r;
              for (int ck = max(hk - r, 0); ck <= min(hk + r, ncell - 1); ++ck) {
                if (!edge_ij && abs(ck - hk) != r) continue;
                const int cell = (ci * ncell + cj) * ncell + ck;
                for (int q = cell_start[cell]; q < cell_start[cell + 1]; ++q) {
                  const int a = cell_atoms[q];
                  const float sc = score(a);
                  if (sc > best || (found && sc == best && a < bestN)) {
                    best = sc;
                    bestN = a;
                    found = true;
                  }
                }
              }
            }
          }
          // every atom in shells > r lies at least r cells (r * cs nodes) away;
          // 1e-4 relative slack covers the FP32 rounding of the scores
          const float dmin = (float)(r * cs) * h * 0.9999f;
          if (best > 0.0f && zmax / sqrtf(dmin * dmin + h2) < best) break;
        }
        bN = bestN;
        best_atom[id] = (uint16_t)bN;
      }
      out = flip_label(raw, cur, bN, n_atoms, atoms, ratio, score);
    }
    lout[id] = out;
  }
}

// Level-set rule (molecule.js:426-540; angular-sector constraints omitted).
// sector test of node (i, j, k) for domain n (molecule.js:438-443 cellSectorOK)
__device__ __forceinline__ bool cell_sector_ok(int i, int j, int k, int n, float h, const Atom* __restrict__ atoms,
                                               const AtomS2* __restrict__ ext, int n_atoms) {
  if (!ext || n >= n_atoms) return true;
  const AtomS2 e = ext[n];
  if (e.st <= 1) return true;
  const float dx = ((float)i - floorf(atoms[n].i)) * h;
  const float dy = ((float)j - floorf(atoms[n].j)) * h;
  const float dz = ((float)k - floorf(atoms[n].k)) * h;
  return in_split(dx, dy, dz, sqrtf(dx * dx + dy * dy + dz * dz), e);
}

__global__ __launch_bounds__(256) void k_levelset(int S, int ich, float h, const Atom* __restrict__ atoms,
                                                  const AtomS2* __restrict__ ext,
                                                  int n_atoms, const int32_t* __restrict__ lock_group,
                                                  const uint16_t* __restrict__ lin, uint16_t* __restrict__ lout,
                                                  const float* __restrict__ U, float* __restrict__ W, float dt_w,
                                                  float curv_reg, int rc_flags) {
  const int n = S - 2;
  March m;
  if (!march_coords(n, ich, m)) return;
  const int S2 = S * S;
  for (int i = m.i0; i < m.i1; ++i) {
    const int id = i * S2 + m.j * S + m.k;
    const uint16_t raw = lin[id];
    const int myL = raw & kLabelMask;
    const float uc = U[id];
    const float myRho = uc * uc;
    int bestL = myL;
    float bestRho = 0.0f;
    float cnt = 0.0f;
    const int nb[6] = {id + S2, id - S2, id + S, id - S, id + 1, id - 1};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int l = lin[nb[q]] & kLabelMask;
      if (l == myL) {
        cnt += 1.0f;
      } else {
        const float un = U[nb[q]];
        const float r = un * un;
        if (r > bestRho) {
          bestRho = r;
          bestL = l;
        }
      }
    }
    float w = W[id];
    if (cnt >= 6.0f) {
      W[id] = fmaxf(w, 1.0f);
      lout[id] = raw;
      continue;
    }
    if (lock_group) {
      const int gm = myL < n_atoms ? lock_group[myL] : 0;
      const int gb = bestL < n_atoms ? lock_group[bestL] : 0;
      if (gm > 0 && gm == gb) {
        W[id] = fmaxf(w, 0.5f);
        lout[id] = raw;
        continue;
      }
    }
    const float denom = myRho + bestRho;
    const float vel = (denom < 1e-20f) ? 0.0f : (myRho - bestRho) / denom;
    w += dt_w * vel;
    const float curv = (cnt - 3.0f) / 3.0f;
    w += curv_reg * curv;
    int newL = myL;
    if (w < 0.0f) {
      newL = bestL;
      w = 0.1f;
    }
    if (ext) {
      // split siblings own only the nodes of their sector (molecule.js:521-536)
      const bool myOK = cell_sector_ok(i, m.j, m.k, myL, h, atoms, ext, n_atoms);
      if (!cell_sector_ok(i, m.j, m.k, newL, h, atoms, ext, n_atoms)) {
        newL = myL;
        w = fmaxf(w, 0.1f);
        if (!myOK && bestL != myL && cell_sector_ok(i, m.j, m.k, bestL, h, atoms, ext, n_atoms)) {
          newL = bestL;
          w = 0.1f;
        }
      } else if (!myOK && newL == myL) {
        if (bestL != myL && cell_sector_ok(i, m.j, m.k, bestL, h, atoms, ext, n_atoms)) {
          newL = bestL;
          w = 0.1f;
        }
      }
    }
    W[id] = w;

    uint16_t out = (uint16_t)newL;
    if (newL == myL) {
      out = raw;
    } else if (rc_flags && newL < n_atoms) {
      const float rc = atoms[newL].rc;
      if (rc > 0.0f) {
        const float dx = ((float)i - floorf(atoms[newL].i)) * h;
        const float dy = ((float)m.j - floorf(atoms[newL].j)) * h;
        const float dz = ((float)m.k - floorf(atoms[newL].k)) * h;
        if (sqrtf(dx * dx + dy * dy + dz * dz) < rc) out |= kRcFlag;
      }
    }
    lout[id] = out;
  }
}

// --------------------------------------------------------------------------
// Energy (cloud/solver.py:227-280; molecule.js:950-1075 for mode S2).
// Per-domain T, V_eK, V_ee bins plus the three dipole sums.
// --------------------------------------------------------------------------
__device__ __forceinline__ bool excl_s2(int i, int j, int k, const Atom& a, float h, float r_ex) {
  const float dx = ((float)i - floorf(a.i)) * h;
  const float dy = ((float)j - floorf(a.j)) * h;
  const float dz = ((float)k - floorf(a.k)) * h;
  return sqrtf(dx * dx + dy * dy + dz * dz) < r_ex;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Energy terms per domain.  The LDS bins cover the domain window [d0, d0 +
// nw) (3 sums per domain, then the 3 dipole sums, which only the window
// d0 = 0 adds to the result); the host runs one launch per window when the
// bins of all domains would not fit the LDS.  A run or dipole sum that is
// not finite sets *bad (the fixed-point sums cannot carry a NaN or Inf).
__global__ __launch_bounds__(256) void k_energy(int mode, int S, int ich, float h, float h3, float r_sing,
                                                const float* __restrict__ U, const float* __restrict__ K,
                                                const float* __restrict__ P, const uint16_t* __restrict__ lab,
                                                const Atom* __restrict__ atoms, int n_dom, int d0, int nw,
                                                Fx* __restrict__ per_dom, Fx* __restrict__ dip,
                                                int* __restrict__ bad) {
  extern __shared__ Fx fbins[];  // nw * 3 sums, then 3 dipole sums
  const int nthr = blockDim.x * blockDim.y;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  for (int t = tid; t < kFxWords * (nw * 3 + 3); t += nthr) fbins[t] = 0;
  __syncthreads();
  // flush of one run of label L into the window's bins
  auto flush = [&](int L, double aT, double aK, double aE) {
    if (!isfinite(aT + aK + aE)) *bad = 1;
    const int w = L - d0;
    if (w >= 0 && w < nw) {
      fx_add(fbins + kFxWords * (w * 3 + 0), aT);
      fx_add(fbins + kFxWords * (w * 3 + 1), aK);
      fx_add(fbins + kFxWords * (w * 3 + 2), aE);
    }
  };
  const int n = S - 2;
  March m;
  double dx_acc = 0.0, dy_acc = 0.0, dz_acc = 0.0;
  if (march_coords(n, ich, m)) {
    const int S2 = S * S;
    int cur = -1;
    double aT = 0.0, aK = 0.0, aE = 0.0;
    const float yj = (float)m.j * h, zk = (float)m.k * h;
    for (int i = m.i0; i < m.i1; ++i) {
      const int id = i * S2 + m.j * S + m.k;
      int L = lab[id] & kLabelMask;
      L = L < n_dom ? L : n_dom - 1;
      const float u = U[id];
      const float rho = u * u;
      float T, VeK, Vee;
      if (mode == RQ_ENERGY_S2) {
        const Atom a = atoms[L];
        const float r_ex = fmaxf(a.rc, r_sing);
        if (excl_s2(i, m.j, m.k, a, h, r_ex)) continue;
        const bool sip = ((lab[id + S2] & kLabelMask) == L) && !excl_s2(i + 1, m.j, m.k, a, h, r_ex);
        const bool sjp = ((lab[id + S] & kLabelMask) == L) && !excl_s2(i, m.j + 1, m.k, a, h, r_ex);
        const bool skp = ((lab[id + 1] & kLabelMask) == L) && !excl_s2(i, m.j, m.k + 1, a, h, r_ex);
        const float da = sip ? U[id + S2] - u : 0.0f;
        const float db = sjp ? U[id + S] - u : 0.0f;
        const float dc = skp ? U[id + 1] - u : 0.0f;
        T = 0.5f * (da * da + db * db + dc * dc) * h;
      } else {
        const float da = ((lab[id + S2] & kLabelMask) == L) ? U[id + S2] - u : 0.0f;
        const float db = ((lab[id + S] & kLabelMask) == L) ? U[id + S] - u : 0.0f;
        const float dc = ((lab[id + 1] & kLabelMask) == L) ? U[id + 1] - u : 0.0f;
        T = 0.5f * (da * da + db * db + dc * dc) * h;
      }
      VeK = -K[id] * rho * h3;
      Vee = P[id] * rho * h3;
      const float xi = (float)i * h;
      dx_acc += (double)(rho * xi * h3);
      dy_acc += (double)(rho * yj * h3);
      dz_acc += (double)(rho * zk * h3);
      if (L != cur) {
        if (cur >= 0) flush(cur, aT, aK, aE);
        cur = L;
        aT = aK = aE = 0.0;
      }
      aT += (double)T;
      aK += (double)VeK;
      aE += (double)Vee;
    }
    if (cur >= 0) flush(cur, aT, aK, aE);
  }
  dx_acc = wave_sum(dx_acc);
  dy_acc = wave_sum(dy_acc);
  dz_acc = wave_sum(dz_acc);
  if ((tid & 63) == 0) {
    if (!isfinite(dx_acc + dy_acc + dz_acc)) *bad = 1;
    fx_add(fbins + kFxWords * (nw * 3 + 0), dx_acc);
    fx_add(fbins + kFxWords * (nw * 3 + 1), dy_acc);
    fx_add(fbins + kFxWords * (nw * 3 + 2), dz_acc);
  }
  __syncthreads();
  for (int t = tid; t < nw * 3; t += nthr) fx_add_words(per_dom + kFxWords * (3 * d0 + t), fbins + kFxWords * t);
  if (d0 == 0 && tid < 3) fx_add_words(dip + kFxWords * tid, fbins + kFxWords * (nw * 3 + tid));
}

// --------------------------------------------------------------------------
// Hellmann-Feynman force (cloud/solver.py:282-315).  Each thread holds a
// column of `kCol` densities in registers and loops the atoms in chunks of
// four; per-atom sums are FP64, reduced across the wave with shuffles and
// across the workgroup through LDS in a fixed order, then added to the
// result (a fixed-point sum, fx_add) once per atom component and workgroup.
// --------------------------------------------------------------------------
constexpr int kCol = 16;
constexpr int kChunk = 4;

__global__ __launch_bounds__(256) void k_forces_hf(int S, float h, float h2, float h3, const float* __restrict__ U,
                                                   const Atom* __restrict__ atoms, int first, int count,
                                                   Fx* __restrict__ out) {
  __shared__ double red[4][kChunk * 3];
  const int n = S - 2;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  const int k = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int i0 = 1 + blockIdx.z * kCol;
  const bool active = (k <= n && j <= n);
  const int S2 = S * S;
  float rho[kCol];
#pragma unroll
  for (int c = 0; c < kCol; ++c) {
    const int i = i0 + c;
    float r = 0.0f;
    if (active && i <= n) {
      const float u = U[i * S2 + j * S + k];
      r = u * u;
    }
    rho[c] = (r < 1e-20f) ? 0.0f : r;  // skipped voxels contribute nothing
  }
  const float yj = (float)j * h, zk = (float)k * h;
  for (int a0 = 0; a0 < count; a0 += kChunk) {
    double acc[kChunk * 3];
#pragma unroll
    for (int q = 0; q < kChunk * 3; ++q) acc[q] = 0.0;
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      const int a = a0 + q;
      if (a >= count) break;
      const Atom at = atoms[first + a];
      if (!(at.Z > 0.0f)) continue;
      const float dy = yj - at.j * h;
      const float dz = zk - at.k * h;
#pragma unroll
      for (int c = 0; c < kCol; ++c) {
        if (rho[c] == 0.0f) continue;
        const float xi = (float)(i0 + c) * h;
        const float dx = xi - at.i * h;
        const float r2 = dx * dx + dy * dy + dz * dz + h2;
        const float inv_r3 = 1.0f / (r2 * sqrtf(r2));
        const float w = at.Z * rho[c] * inv_r3 * h3;
        acc[q * 3 + 0] += (double)(w * dx);
        acc[q * 3 + 1] += (double)(w * dy);
        acc[q * 3 + 2] += (double)(w * dz);
      }
    }
#pragma unroll
    for (int q = 0; q < kChunk * 3; ++q) acc[q] = wave_sum(acc[q]);
    const int wave = tid >> 6;
    if ((tid & 63) == 0) {
#pragma unroll
      for (int q = 0; q < kChunk * 3; ++q) red[wave][q] = acc[q];
    }
    __syncthreads();
    if (tid < kChunk * 3) {
      const int a = a0 + tid / 3;
      if (a < count) {
        const double s = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
        fx_add(out + kFxWords * (a * 3 + tid % 3), s);
      }
    }
    __syncthreads();
  }
}

// molecule.js:1199-1270: F = 2 Z <grad P> averaged over the ball of `radius`
// cells around the (integer) nucleus position; central differences.  One
// workgroup per atom; FP64 sums.
__global__ __launch_bounds__(256) void k_forces_gradp(int S, float h, const float* __restrict__ P,
                                                      const Atom* __restrict__ atoms,
                                                      const AtomS2* __restrict__ ext, int n_atoms, int R,
                                                      double* __restrict__ out) {
  __shared__ double red[4][4];
  const int a = blockIdx.x;
  if (a >= n_atoms) return;
  const Atom at = atoms[a];
  const int tid = threadIdx.x;
  // the force charge: Z, or Z_nuc for a bare nucleus (molecule.js:1214)
  const float ZA = (ext && at.Z <= 0.0f) ? ext[a].Zn : at.Z;
  if (!(ZA > 0.0f)) {
    if (tid < 3) out[a * 3 + tid] = 0.0;
    return;
  }
  const int ci = (int)floorf(at.i), cj = (int)floorf(at.j), ck = (int)floorf(at.k);
  const float inv2h = 0.5f * (1.0f / h);
  const int D = 2 * R + 1;
  const int total = D * D * D;
  const int S2 = S * S;
  double sx = 0.0, sy = 0.0, sz = 0.0, cnt = 0.0;
  for (int t = tid; t < total; t += blockDim.x) {
    const int di = t / (D * D) - R, dj = (t / D) % D - R, dk = t % D - R;
    const int ii = ci + di, jj = cj + dj, kk = ck + dk;
    if (ii < 1 || ii >= S - 1 || jj < 1 || jj >= S - 1 || kk < 1 || kk >= S - 1) continue;
    if (di * di + dj * dj + dk * dk > R * R) continue;
    const int id = ii * S2 + jj * S + kk;
    sx += (double)((P[id + S2] - P[id - S2]) * inv2h);
    sy += (double)((P[id + S] - P[id - S]) * inv2h);
    sz += (double)((P[id + 1] - P[id - 1]) * inv2h);
    cnt += 1.0;
  }
  sx = wave_sum(sx);
  sy = wave_sum(sy);
  sz = wave_sum(sz);
  cnt = wave_sum(cnt);
  if ((tid & 63) == 0) {
    red[tid >> 6][0] = sx;
    red[tid >> 6][1] = sy;
    red[tid >> 6][2] = sz;
    red[tid >> 6][3] = cnt;
  }
  __syncthreads();
  if (tid < 3) {
    const int nw = (blockDim.x + 63) / 64;
    double s = 0.0, c = 0.0;
    for (int w = 0; w < nw; ++w) {
      s += red[w][tid];
      c += red[w][3];
    }
    if (c > 0.0) s /= c;
    out[a * 3 + tid] = 2.0 * (double)ZA * s;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline unsigned nblk(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_init_trial(hipStream_t st, int mode, int S, float h, float h2, const Atom* atoms, const AtomS2* ext,
                       int n_atoms, float* U, float* U2, uint16_t* lab, float* K, float* bestU, float* W) {
  const size_t S3 = (size_t)S * S * S;
  const float r_sing = 2.0f * h;
  hipLaunchKernelGGL(k_init_trial, dim3(nblk(S3, 256)), dim3(256), 0, st, mode, S, h, h2, r_sing, atoms, ext,
                     n_atoms, U, U2, lab, K, bestU, W);
}

void launch_recompute_k(hipStream_t st, int mode, int S, float h, float h2, const Atom* atoms, const AtomS2* ext,
                        int n_atoms, float* K) {
  const size_t S3 = (size_t)S * S * S;
  hipLaunchKernelGGL(k_recompute_k, dim3(nblk(S3, 256)), dim3(256), 0, st, mode, S, h, h2, 2.0f * h, atoms, ext,
                     n_atoms, K);
}

void launch_rc_flags(hipStream_t st, int S, float h, const Atom* atoms, int n_atoms, uint16_t* lab) {
  const size_t S3 = (size_t)S * S * S;
  hipLaunchKernelGGL(k_rc_flags, dim3(nblk(S3, 256)), dim3(256), 0, st, S, h, atoms, n_atoms, lab);
}

void launch_itp(hipStream_t st, const Tile& t, int S, const float* U, float* Un, const float* K, const float* P,
                const uint16_t* lab, float half_dv, float dtv) {
  hipLaunchKernelGGL(k_itp, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, U, Un, K, P, lab, half_dv,
                     dtv);
}

void launch_smooth_fine(hipStream_t st, const Tile& t, int S, const float* P, float* Pn, const float* U, float a,
                        float c, float alpha, float beta) {
  hipLaunchKernelGGL(k_smooth<true>, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, P, Pn, U, a, c,
                     alpha, beta);
}

void launch_smooth_coarse(hipStream_t st, const Tile& t, int S, const float* E, float* En, const float* g, float a,
                          float c) {
  hipLaunchKernelGGL(k_smooth<false>, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, E, En, g, a, c,
                     1.0f, 1.0f);
}

void launch_residual_fine(hipStream_t st, const Tile& t, int S, const float* P, const float* U, float alpha,
                          float beta, float* r) {
  hipLaunchKernelGGL(k_residual<true>, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, P, U, alpha, beta,
                     r);
}

void launch_residual_coarse(hipStream_t st, const Tile& t, int S, const float* E, const float* g, float* r) {
  hipLaunchKernelGGL(k_residual<false>, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, E, g, 1.0f, 1.0f,
                     r);
}

void launch_restrict4(hipStream_t st, int Sf, const float* r, float* gc, float sign) {
  const int nc = (Sf - 1) / 2 - 1;
  const size_t tot = (size_t)nc * nc * nc;
  hipLaunchKernelGGL(k_restrict4, dim3(nblk(tot, 256)), dim3(256), 0, st, Sf, r, gc, sign);
}

void launch_prolong_correct(hipStream_t st, const Tile& t, int Sf, float* x, const float* ec, float scale) {
  hipLaunchKernelGGL(k_prolong_correct<false>, stencil_grid(Sf - 2, t), dim3(t.bx, t.by), 0, st, Sf, t.ich, x, ec,
                     scale);
}

void launch_prolong_set(hipStream_t st, const Tile& t, int Sf, float* x, const float* ec) {
  hipLaunchKernelGGL(k_prolong_correct<true>, stencil_grid(Sf - 2, t), dim3(t.bx, t.by), 0, st, Sf, t.ich, x, ec,
                     1.0f);
}

template <int M>
static void exact_rows(hipStream_t st, int Sc, const float* g, float* e, const double* Smat, const double* rden,
                       size_t lds) {
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)k_coarse_exact_rows<M>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    attr = true;
  }
  const int m = Sc - 2;
  hipLaunchKernelGGL((k_coarse_exact_rows<M>), dim3(1), dim3(m, 1024 / m), lds, st, Sc, g, e, Smat, rden);
}

void launch_coarse_exact(hipStream_t st, int Sc, const float* g, float* e, const double* Smat, const double* rden) {
  const int m = Sc - 2;
  if (Smat && rden && m <= kMaxCoarsePP) {
    const size_t lds = sizeof(double) * 2 * (size_t)m * m * m;
    // rows variants with m fixed at compile time for the odd m that 2^k-cell
    // grids produce; m^2 * 8 * 2 bytes of LDS never exceeds 110 KB (m <= 19)
    switch (m) {
      case 3: exact_rows<3>(st, Sc, g, e, Smat, rden, lds); return;
      case 5: exact_rows<5>(st, Sc, g, e, Smat, rden, lds); return;
      case 7: exact_rows<7>(st, Sc, g, e, Smat, rden, lds); return;
      case 9: exact_rows<9>(st, Sc, g, e, Smat, rden, lds); return;
      case 11: exact_rows<11>(st, Sc, g, e, Smat, rden, lds); return;
      case 13: exact_rows<13>(st, Sc, g, e, Smat, rden, lds); return;
      case 15: exact_rows<15>(st, Sc, g, e, Smat, rden, lds); return;
      case 19: exact_rows<19>(st, Sc, g, e, Smat, rden, lds); return;
      default: exact_rows<0>(st, Sc, g, e, Smat, rden, lds); return;
    }
  }
  const size_t lds = sizeof(double) * ((size_t)m * m + m + (size_t)m * m * m);
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void*)k_coarse_exact, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr_set = true;
  }
  hipLaunchKernelGGL(k_coarse_exact, dim3(1), dim3(256), lds, st, Sc, g, e);
}

size_t bottom_lds_bytes(int S) {
  const size_t Sc = (S - 1) / 2 + 1, m = Sc - 2;
  return sizeof(double) * (2 * m * m * m + m * m) + sizeof(float) * (2 * (size_t)S * S * S + Sc * Sc * Sc);
}

bool bottom_fits(int S) {
  return S >= 7 && S <= kBottomMaxS && (S - 1) % 2 == 0 && bottom_lds_bytes(S) <= 160 * 1024;
}

void launch_bottom(hipStream_t st, int S, int nu1, int nu2, float a, float c, const float* g, float* e,
                   const double* Smat, const double* rden) {
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)k_bottom, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr = true;
  }
  hipLaunchKernelGGL(k_bottom, dim3(1), dim3(kBotThreads), bottom_lds_bytes(S), st, S, nu1, nu2, a, c, g, e, Smat,
                     rden);
}

void launch_smooth2_from_zero(hipStream_t st, const Tile& t, int S, const float* g, float* e_out, float a, float c) {
  hipLaunchKernelGGL(k_smooth2_from_zero, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, g, e_out, a, c);
}

template <int CK, int CJ, int CI>
void resres_tiled(hipStream_t st, bool fine, int Sf, const float* x, const float* src, float alpha, float beta,
                  float* gc, float sign) {
  const int nc = (Sf - 1) / 2 - 1;
  const dim3 grid((nc + CK - 1) / CK, (nc + CJ - 1) / CJ, (nc + CI - 1) / CI);
  if (fine)
    hipLaunchKernelGGL((k_resres_tiled<true, CK, CJ, CI>), grid, dim3(CK * CJ * CI), 0, st, Sf, x, src, alpha, beta,
                       gc, sign);
  else
    hipLaunchKernelGGL((k_resres_tiled<false, CK, CJ, CI>), grid, dim3(CK * CJ * CI), 0, st, Sf, x, src, 1.0f, 1.0f,
                       gc, sign);
}

namespace {
template <int TK, int TJ, int CP>
void smooth2_march(hipStream_t st, bool fine, int S, const float* x, float* xn, const float* src, float a, float c,
                   float alpha, float beta) {
  const int n = S - 2;
  const int gx = (n + TK - 1) / TK, gy = (n + TJ - 1) / TJ;
  const int gz = std::max(1, std::min(n, kSmooth2Blocks / (gx * gy)));
  const int CI = std::max(2 * CP, (n + gz - 1) / gz);
  const dim3 grid(gx, gy, (n + CI - 1) / CI);
  if (fine)
    hipLaunchKernelGGL((k_smooth2_march<true, TK, TJ, CP>), grid, dim3(256), 0, st, S, CI, x, xn, src, a, c, alpha,
                       beta);
  else
    hipLaunchKernelGGL((k_smooth2_march<false, TK, TJ, CP>), grid, dim3(256), 0, st, S, CI, x, xn, src, a, c, 1.0f,
                       1.0f);
}
}  // namespace

void launch_smooth2(hipStream_t st, bool fine, int S, const float* x, float* xn, const float* src, float a, float c,
                    float alpha, float beta) {
  // 64 x 8 columns, one plane per round: the fastest of the 32/64 x 4/8,
  // 1/2-plane tiles on the 513^3 grid
  smooth2_march<64, 8, 1>(st, fine, S, x, xn, src, a, c, alpha, beta);
}

namespace {
template <int CK, int CJ, int NT>
void smooth2_resres(hipStream_t st, int Sf, const float* x, float* xn, const float* src, float a, float c,
                    float alpha, float beta, float* gc, int blocks) {
  const int nc = (Sf - 1) / 2 - 1;
  const int gx = (nc + CK - 1) / CK, gy = (nc + CJ - 1) / CJ;
  const int gz = std::max(1, std::min(nc, blocks / (gx * gy)));
  const int CI = (nc + gz - 1) / gz;
  const dim3 grid(gx, gy, (nc + CI - 1) / CI);
  hipLaunchKernelGGL((k_smooth2_resres_march<CK, CJ, NT>), grid, dim3(NT), 0, st, Sf, CI, x, xn, src, a, c, alpha,
                     beta, gc);
}
}  // namespace

void launch_smooth2_resres(hipStream_t st, int Sf, const float* x, float* xn, const float* src, float a, float c,
                           float alpha, float beta, float* gc, int blocks) {
  // 32 x 8 coarse columns (64 x 16 fine), 768 threads, ~2048 workgroups:
  // each stage's region fills two rounds of the workgroup (1.5 thread slots
  // per owned node and stage, against 2 for 32 x 4 with 512 threads).  On
  // the 513^3 grid: 656 us, against 683 with 4096 workgroups, 777 for
  // 32 x 4 / 512, 712 for 32 x 8 / 512 and 778 for 16 x 16 / 768
  // (RQ_SMOOTH2_RESRES_SHAPE 1-6, RQ_SMOOTH2_RESRES_BLOCKS)
  static const int shape = [] {
    const char* e = std::getenv("RQ_SMOOTH2_RESRES_SHAPE");
    return e ? std::atoi(e) : 0;
  }();
  const int b = blocks > 0 ? blocks : kSmooth2ResresBlocks;
  switch (shape) {
    case 1: smooth2_resres<16, 8, 512>(st, Sf, x, xn, src, a, c, alpha, beta, gc, b); break;
    case 2: smooth2_resres<16, 4, 256>(st, Sf, x, xn, src, a, c, alpha, beta, gc, b); break;
    case 3: smooth2_resres<32, 8, 1024>(st, Sf, x, xn, src, a, c, alpha, beta, gc, b); break;
    case 4: smooth2_resres<32, 4, 512>(st, Sf, x, xn, src, a, c, alpha, beta, gc, b); break;
    case 5: smooth2_resres<32, 8, 512>(st, Sf, x, xn, src, a, c, alpha, beta, gc, b); break;
    case 6: smooth2_resres<16, 16, 768>(st, Sf, x, xn, src, a, c, alpha, beta, gc, b); break;
    default: smooth2_resres<32, 8, 768>(st, Sf, x, xn, src, a, c, alpha, beta, gc, b); break;
  }
}

namespace {
template <int TK, int TJ, int RU = 4>
void post_march(hipStream_t st, int Sf, const float* x, const float* e, float* xn, const float* U, float* Un,
                const float* K, const uint16_t* lab, float a, float c, float alpha, float beta, float half_dv,
                float dtv, int blocks) {
  const int n = Sf - 2;
  const int gx = (n + TK - 1) / TK, gy = (n + TJ - 1) / TJ;
  const int gz = std::max(1, std::min(n, blocks / (gx * gy)));
  const int CI = (n + gz - 1) / gz;
  const dim3 grid(gx, gy, (n + CI - 1) / CI);
  hipLaunchKernelGGL((k_post_march<TK, TJ, RU>), grid, dim3(TK * TJ), 0, st, Sf, CI, x, e, xn, U, Un, K, lab, a, c,
                     alpha, beta, half_dv, dtv);
}
}  // namespace

void launch_post_itp(hipStream_t st, int Sf, const float* x, const float* e, float* xn, const float* U, float* Un,
                     const float* K, const uint16_t* lab, float a, float c, float alpha, float beta, float half_dv,
                     float dtv, int blocks) {
  // 32 x 8 columns with five U / label planes (no barrier after the update
  // stage), ~8192 workgroups: 1042 us on the 513^3 grid, against 1086 with
  // four planes, 1150 for 64 x 8 and 1125 for 16 x 16 (RQ_POST_SHAPE 1-3)
  static const int shape = [] {
    const char* v = std::getenv("RQ_POST_SHAPE");
    return v ? std::atoi(v) : 0;
  }();
  const int b = blocks > 0 ? blocks : kPostBlocks;
  switch (shape) {
    case 1: post_march<32, 8, 4>(st, Sf, x, e, xn, U, Un, K, lab, a, c, alpha, beta, half_dv, dtv, b); break;
    case 2: post_march<64, 8, 5>(st, Sf, x, e, xn, U, Un, K, lab, a, c, alpha, beta, half_dv, dtv, b); break;
    case 3: post_march<16, 16, 4>(st, Sf, x, e, xn, U, Un, K, lab, a, c, alpha, beta, half_dv, dtv, b); break;
    default: post_march<32, 8, 5>(st, Sf, x, e, xn, U, Un, K, lab, a, c, alpha, beta, half_dv, dtv, b); break;
  }
}

void launch_residual_restrict(hipStream_t st, bool fine, int Sf, const float* x, const float* src, float alpha,
                              float beta, float* gc, float sign) {
  const int nc = (Sf - 1) / 2 - 1;
  if (nc > kResResTiledMaxNc) {
    // march: 32 x 2 coarse columns per workgroup, CI planes for ~2048 workgroups
    constexpr int CJ = kResResMarchCJ, CP = kResResMarchCP;
    const int gx = (nc + 31) / 32, gy = (nc + CJ - 1) / CJ;
    const int gz = std::max(1, std::min(nc, 2048 / (gx * gy)));
    const int CI = std::max(4, (nc + gz - 1) / gz);
    const dim3 grid(gx, gy, (nc + CI - 1) / CI);
    if (fine)
      hipLaunchKernelGGL((k_resres_march<true, CJ, CP>), grid, dim3(128 * CJ), 0, st, Sf, CI, x, src, alpha, beta, gc,
                         sign);
    else
      hipLaunchKernelGGL((k_resres_march<false, CJ, CP>), grid, dim3(128 * CJ), 0, st, Sf, CI, x, src, 1.0f, 1.0f, gc,
                         sign);
    return;
  }
  // 8x8x4 coarse tiles (1.27 residuals per fine node) where there are enough
  // of them to fill the chip, 4x4x4 (1.42) on the small levels
  if (nc >= 40)
    resres_tiled<8, 8, 4>(st, fine, Sf, x, src, alpha, beta, gc, sign);
  else
    resres_tiled<4, 4, 4>(st, fine, Sf, x, src, alpha, beta, gc, sign);
}

template <int CK, int CJ, int CI>
static void down2_tiled(hipStream_t st, int S, const float* g, float* e_out, float* gc, float a, float c) {
  const int nc = (S - 1) / 2 - 1;
  const dim3 grid((nc + CK - 1) / CK, (nc + CJ - 1) / CJ, (nc + CI - 1) / CI);
  hipLaunchKernelGGL((k_down2<CK, CJ, CI>), grid, dim3(256), 0, st, S, g, e_out, gc, a, c);
}

// 256 threads per workgroup on every level; smaller coarse tiles on smaller
// levels keep a few nodes per thread (these launches are latency-bound)
void launch_pdirect_init(hipStream_t st, const Tile& t, int S, float h, float h2, const Atom* atoms, int n_atoms,
                         int m, int floor_pos, float* Pm, float* Pother) {
  hipLaunchKernelGGL(k_pdirect_init, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, h, h2, atoms, n_atoms,
                     m, floor_pos, Pm, Pother);
}

void launch_copy_for_label(hipStream_t st, const Tile& t, int S, const float* Pm, const uint16_t* lab, int m,
                           float* Pother) {
  hipLaunchKernelGGL(k_copy_for_label, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, Pm, lab, m, Pother);
}

void launch_rho_other(hipStream_t st, const Tile& t, int S, const float* U, const uint16_t* lab, int m, float alpha,
                      float beta, float* rhs) {
  hipLaunchKernelGGL(k_rho_other, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, U, lab, m, alpha, beta,
                     rhs);
}

void launch_rho_self(hipStream_t st, const Tile& t, int S, const float* U, const uint16_t* lab, int m, float alpha,
                     float beta, float* rhs) {
  hipLaunchKernelGGL(k_rho_self, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, U, lab, m, alpha, beta,
                     rhs);
}

void launch_sub_pself(hipStream_t st, const Tile& t, int S, const float* Pself, const uint16_t* lab, int m, float frac,
                      float* Pother) {
  hipLaunchKernelGGL(k_sub_pself, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, Pself, lab, m, frac,
                     Pother);
}

void launch_pother_gather(hipStream_t st, const Tile& t, int S, const uint16_t* lab, const float* const* Pdom,
                          int n_dom, float* P) {
  hipLaunchKernelGGL(k_pother_gather, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, lab, Pdom, n_dom, P);
}

void launch_itp_gather(hipStream_t st, const Tile& t, int S, const float* U, float* Un, const float* K,
                       const uint16_t* lab, const float* const* Pdom, int n_dom, float* P, float half_dv, float dtv) {
  hipLaunchKernelGGL(k_itp_gather, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, U, Un, K, lab, Pdom,
                     n_dom, P, half_dv, dtv);
}

void launch_display_slice(hipStream_t st, int S, int sk, const float* U, const uint16_t* lab, const float* K,
                          const Atom* atoms, int n_atoms, float* out) {
  hipLaunchKernelGGL(k_display_slice, dim3((S + 15) / 16, (S + 15) / 16), dim3(16, 16), 0, st, S, sk, U, lab, K, atoms,
                     n_atoms, out);
}

void launch_down2(hipStream_t st, int S, const float* g, float* e_out, float* gc, float a, float c) {
  const int nc = (S - 1) / 2 - 1;
  if (nc >= 32)
    down2_tiled<8, 8, 4>(st, S, g, e_out, gc, a, c);
  else if (nc >= 16)
    down2_tiled<4, 4, 4>(st, S, g, e_out, gc, a, c);
  else
    down2_tiled<4, 4, 2>(st, S, g, e_out, gc, a, c);
}

template <int TK, int TJ, int TI>
static bool up_tiled(hipStream_t st, int S, int nu, const float* x, const float* e, float* x_out, const float* g,
                     float a, float c) {
  const int n = S - 2;
  const dim3 grid((n + TK - 1) / TK, (n + TJ - 1) / TJ, (n + TI - 1) / TI);
  switch (nu) {
    case 1:
      hipLaunchKernelGGL((k_up<TK, TJ, TI, 1>), grid, dim3(256), 0, st, S, x, e, x_out, g, a, c);
      return true;
    case 2:
      hipLaunchKernelGGL((k_up<TK, TJ, TI, 2>), grid, dim3(256), 0, st, S, x, e, x_out, g, a, c);
      return true;
    case 3:
      hipLaunchKernelGGL((k_up<TK, TJ, TI, 3>), grid, dim3(256), 0, st, S, x, e, x_out, g, a, c);
      return true;
    default:
      return false;
  }
}

bool launch_up(hipStream_t st, int S, int nu, const float* x, const float* e, float* x_out, const float* g, float a,
               float c) {
  if (S - 2 >= 64) return up_tiled<16, 8, 4>(st, S, nu, x, e, x_out, g, a, c);
  return up_tiled<8, 8, 4>(st, S, nu, x, e, x_out, g, a, c);
}

void launch_prolong_smooth(hipStream_t st, const Tile& t, bool fine, int Sf, const float* x, const float* e,
                           float* x_out, const float* src, float a, float c, float alpha, float beta) {
  if (fine)
    hipLaunchKernelGGL(k_prolong_smooth<true>, stencil_grid(Sf - 2, t), dim3(t.bx, t.by), 0, st, Sf, t.ich, x, e,
                       x_out, src, a, c, alpha, beta);
  else
    hipLaunchKernelGGL(k_prolong_smooth<false>, stencil_grid(Sf - 2, t), dim3(t.bx, t.by), 0, st, Sf, t.ich, x, e,
                       x_out, src, a, c, 1.0f, 1.0f);
}

void launch_smooth_itp(hipStream_t st, const Tile& t, int S, const float* P, float* Pn, const float* U, float* Un,
                       const float* K, const uint16_t* lab, float a, float c, float alpha, float beta, float half_dv,
                       float dtv) {
  hipLaunchKernelGGL(k_smooth_itp, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, P, Pn, U, Un, K, lab, a,
                     c, alpha, beta, half_dv, dtv);
}

void launch_zero(hipStream_t st, float* p, size_t n) { launch_fill(st, p, n, 0.0f); }

void launch_fill(hipStream_t st, float* p, size_t n, float v) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill, dim3(nblk(n, 256)), dim3(256), 0, st, p, n, v);
}

void launch_norm_sums(hipStream_t st, const Tile& t, int S, float h3, const float* U, const uint16_t* lab, int n_dom,
                      Fx* norms, int* bad) {
  // longer columns than the sweeps' tile: a few thousand workgroups, so a
  // few thousand atomics into the per-domain sums instead of ~10^4
  Tile r = t;
  r.ich = std::max(t.ich, kReduceIch);
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void*)k_norm_sums, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr_set = true;
  }
  hipLaunchKernelGGL(k_norm_sums, stencil_grid(S - 2, r), dim3(r.bx, r.by), sizeof(Fx) * kFxWords * n_dom, st, S,
                     r.ich, h3, U, lab, n_dom, norms, bad);
}

void launch_normalize(hipStream_t st, const Tile& t, int S, float* U, const uint16_t* lab, const Fx* norms,
                      const float* Z, int n_dom, float* scale) {
  hipLaunchKernelGGL(k_norm_scales, dim3((n_dom + 255) / 256), dim3(256), 0, st, n_dom, norms, Z, scale);
  hipLaunchKernelGGL(k_normalize, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, U, lab, scale, n_dom);
}

void launch_score_map(hipStream_t st, int S, float h, float h2, const Atom* atoms, int n_atoms, uint16_t* best_atom) {
  const int nb = (S - 2 + kMapB - 1) / kMapB;
  hipLaunchKernelGGL(k_score_map, dim3((unsigned)nb * nb * nb), dim3(256), 0, st, S, h, h2, atoms, n_atoms, best_atom);
}

void launch_boundary_score(hipStream_t st, const Tile& t, int S, float h, float h2, const Atom* atoms, int n_atoms,
                           const uint16_t* lin, uint16_t* lout, float ratio, int min_foreign, const AtomCells& cells,
                           const int32_t* cell_start, const int32_t* cell_atoms, uint16_t* best_atom) {
  if (n_atoms <= kScoreScanMaxAtoms) {
    hipLaunchKernelGGL(k_boundary_scan, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, h, h2, atoms,
                       n_atoms, lin, lout, ratio, min_foreign, best_atom);
    return;
  }
  hipLaunchKernelGGL(k_boundary_score, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, h, h2, atoms,
                     n_atoms, lin, lout, ratio, min_foreign, cells.cs, cells.n, cells.zmax, cell_start, cell_atoms,
                     best_atom);
}

void launch_levelset(hipStream_t st, const Tile& t, int S, float h, const Atom* atoms, const AtomS2* ext,
                     int n_atoms, const int32_t* lock_group, const uint16_t* lin, uint16_t* lout, const float* U,
                     float* W, float dt_w, float curv_reg, int rc_flags) {
  hipLaunchKernelGGL(k_levelset, stencil_grid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, h, atoms, ext, n_atoms,
                     lock_group, lin, lout, U, W, dt_w, curv_reg, rc_flags);
}

void launch_energy(hipStream_t st, const Tile& t, int mode, int S, float h, float h3, const float* U, const float* K,
                   const float* P, const uint16_t* lab, const Atom* atoms, int n_dom, Fx* per_dom, Fx* dip,
                   int* bad) {
  // domain windows whose bins (3 sums each + 3 dipole sums) fit 150 KiB of LDS
  constexpr int kWindow = 3200;
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void*)k_energy, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr_set = true;
  }
  for (int d0 = 0; d0 < n_dom; d0 += kWindow) {
    const int nw = std::min(kWindow, n_dom - d0);
    const size_t lds = sizeof(Fx) * kFxWords * ((size_t)nw * 3 + 3);
    hipLaunchKernelGGL(k_energy, stencil_grid(S - 2, t), dim3(t.bx, t.by), lds, st, mode, S, t.ich, h, h3, 2.0f * h,
                       U, K, P, lab, atoms, n_dom, d0, nw, per_dom, dip, bad);
  }
}

void launch_forces_hf(hipStream_t st, int S, float h, float h2, float h3, const float* U, const Atom* atoms,
                      int first, int count, Fx* out3) {
  const int n = S - 2;
  dim3 grid((n + 63) / 64, (n + 3) / 4, (n + kCol - 1) / kCol);
  hipLaunchKernelGGL(k_forces_hf, grid, dim3(64, 4), 0, st, S, h, h2, h3, U, atoms, first, count, out3);
}

void launch_forces_gradp(hipStream_t st, int S, float h, const float* P, const Atom* atoms, const AtomS2* ext,
                         int n_atoms, int radius, double* out3n) {
  hipLaunchKernelGGL(k_forces_gradp, dim3(n_atoms), dim3(256), 0, st, S, h, P, atoms, ext, n_atoms, radius, out3n);
}

void launch_square(hipStream_t st, int S, const float* U, float* rho) {
  const int S3 = S * S * S;
  hipLaunchKernelGGL(k_square, dim3(nblk(S3, 256)), dim3(256), 0, st, S3, U, rho);
}

void launch_label_u16_to_i32(hipStream_t st, size_t n, const uint16_t* in, int32_t* out) {
  hipLaunchKernelGGL(k_lab16to32, dim3(nblk(n, 256)), dim3(256), 0, st, n, in, out);
}

void launch_label_i32_to_u16(hipStream_t st, size_t n, const int32_t* in, uint16_t* out) {
  hipLaunchKernelGGL(k_lab32to16, dim3(nblk(n, 256)), dim3(256), 0, st, n, in, out);
}

// ---- scratch of the allocation-free HF force entry point ----
// One array per GPU for the whole process: the fixed-point sums of up to
// kNoAllocMaxAtoms atoms (3 components each), then their values as doubles.
// In host code `scratch` names a host-side stand-in, not device memory: the
// device address of each GPU's copy comes from hipGetSymbolAddress, taken
// once per device (noalloc_setup) and kept in g_scratch_dev.
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevicesTracked = 64;
void* g_scratch_dev[kMaxDevicesTracked] = {};
std::mutex g_scratch_mu;

__global__ void k_fx_values(int n, const Fx* __restrict__ w, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = fx_value(w + kFxWords * t);
}
}  // namespace

hipError_t noalloc_setup() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevicesTracked) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[dev]) return hipSuccess;
  void* p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_dev[dev] = p;
  return hipSuccess;
}

void* noalloc_scratch() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevicesTracked) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_dev[dev];
}

void launch_fx_values(hipStream_t st, int n, const Fx* w, double* out) {
  hipLaunchKernelGGL(k_fx_values, dim3(nblk(n, 256)), dim3(256), 0, st, n, w, out);
}

}  // namespace rq
