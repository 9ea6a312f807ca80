// Reverse-mode (adjoint) kernels for dE/dR of the fixed-step S1 trajectory
// with frozen labels.  Forward map per step: Poisson (linear in P and in
// rhs = 2 pi h^2 U^2), ITP U' = U + (dv/2) L_N U + dv h^2 (K - 2P) U with the
// symmetric label-Neumann Laplacian L_N, per-domain normalise.  The adjoint
// of the ITP is the same stencil applied to the adjoint field; the Jacobi
// smoother is symmetric; the V-cycle with nu1 = nu2 and R = Pt/2 is
// symmetric, so its coarse part is reused as is.
//
// Adjoint fields are FP32 (same grid/layout as the forward fields, zero
// faces); per-domain and per-atom sums are FP64.
#include <hip/hip_runtime.h>

#include <cmath>

#include "rq_internal.h"

namespace rq {
namespace {

inline dim3 sgrid(int n, const Tile& t) {
  return dim3((n + t.bx - 1) / t.bx, (n + t.by - 1) / t.by, (n + t.ich - 1) / t.ich);
}

struct M3 {
  int k, j, i0, i1;
};

// Thread coordinates of a 2.5D launch, with the linear workgroup id remapped
// so each XCD (workgroups are dealt to the 8 XCDs round-robin) takes one
// contiguous range of tiles: neighbouring tiles, whose halos overlap, then
// share that XCD's L2.
__device__ __forceinline__ bool coords(int n, int ich, M3& m) {
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

template <typename T>
__device__ __forceinline__ T wsum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// K(R) and the frozen-label trial U0 = Z_L^2/sqrt(pi) exp(-Z_L r~_L) on the
// interior (cloud/solver.py:36-57 with the argmax replaced by the given label).
__global__ void k_init_frozen(int S, float h, float h2, const Atom* __restrict__ atoms, int n_atoms,
                              const uint16_t* __restrict__ lab, float* __restrict__ U, float* __restrict__ K) {
  const int S2 = S * S;
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= S2 * S) return;
  const int i = id / S2, j = (id / S) % S, k = id % S;
  const int NN = S - 1;
  if (i < 1 || i >= NN || j < 1 || j >= NN || k < 1 || k >= NN) return;
  const float xi = (float)i * h, yj = (float)j * h, zk = (float)k * h;
  const int L = lab[id] & kLabelMask;
  float Kv = 0.0f, u0 = 0.0f;
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
    if (n == L) u0 = Za * Za * 0.5641895835f * expf(-Za * r);
  }
  K[id] = Kv;
  U[id] = u0;
}

// Seed of the backward pass: dE/dU, dE/dK, dE/dP of E = T + V_eK + V_ee
// (cloud/solver.py:227-257) at the final state.
__global__ __launch_bounds__(256) void k_energy_grad(int S, int ich, float h, float h3, const float* __restrict__ U,
                                                     const float* __restrict__ K, const float* __restrict__ P,
                                                     const uint16_t* __restrict__ lab, float* __restrict__ lamU,
                                                     float* __restrict__ gK, float* __restrict__ lamP) {
  const int n = S - 2;
  M3 m;
  if (!coords(n, ich, m)) return;
  const int S2 = S * S;
  const int d[3] = {S2, S, 1};
  const int c3[3] = {0, 0, 0};
  (void)c3;
  for (int i = m.i0; i < m.i1; ++i) {
    const int id = i * S2 + m.j * S + m.k;
    const int L = lab[id] & kLabelMask;
    const float u = U[id];
    const int coord[3] = {i, m.j, m.k};
    float gT = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if ((lab[id + d[a]] & kLabelMask) == L) gT += u - U[id + d[a]];
      // backward edge from v - e counts only when v - e is interior
      if (coord[a] > 1 && (lab[id - d[a]] & kLabelMask) == L) gT += u - U[id - d[a]];
    }
    lamU[id] = h * gT - 2.0f * K[id] * u * h3 + 2.0f * P[id] * u * h3;
    gK[id] = -(u * u) * h3;
    lamP[id] = (u * u) * h3;
  }
}

// ITP adjoint: lam_out = lam + hd L_N lam + dtv (K - 2P) lam;
// gK += dtv U lam; lamP += -2 dtv U lam.
__global__ __launch_bounds__(256) void k_itp_adj(int S, int ich, const float* __restrict__ lam,
                                                 float* __restrict__ lam_out, const float* __restrict__ U,
                                                 const float* __restrict__ K, const float* __restrict__ P,
                                                 const uint16_t* __restrict__ lab, float half_dv, float dtv,
                                                 float* __restrict__ gK, float* __restrict__ lamP) {
  const int n = S - 2;
  M3 m;
  if (!coords(n, ich, m)) return;
  const int S2 = S * S;
  // an i-march: labels and lam of planes i - 1 and i stay in registers
  int id = m.i0 * S2 + m.j * S + m.k;
  uint16_t lbm = lab[id - S2], lc = lab[id];
  float lvm = lam[id - S2], lc_v = lam[id];
  for (int i = m.i0; i < m.i1; ++i, id += S2) {
    const uint16_t lbp = lab[id + S2];
    const float lvp = lam[id + S2];
    const float t0 = (lbp == lc) ? lvp : lc_v;
    const float t1 = (lbm == lc) ? lvm : lc_v;
    const float t2 = (lab[id + S] == lc) ? lam[id + S] : lc_v;
    const float t3 = (lab[id - S] == lc) ? lam[id - S] : lc_v;
    const float t4 = (lab[id + 1] == lc) ? lam[id + 1] : lc_v;
    const float t5 = (lab[id - 1] == lc) ? lam[id - 1] : lc_v;
    const float lap = t0 + t1 + t2 + t3 + t4 + t5 - 6.0f * lc_v;
    const float kk = K[id], pp = P[id];
    lam_out[id] = lc_v + half_dv * lap + dtv * (kk - 2.0f * pp) * lc_v;
    const float ul = U[id] * lc_v;
    gK[id] += dtv * ul;
    lamP[id] += -2.0f * dtv * ul;
    lbm = lc;
    lc = lbp;
    lvm = lc_v;
    lc_v = lvp;
  }
}

// Smoother adjoint: lam_out = a lam + sum6 lam / c, and by MODE
//   kAdjAcc:   acc += coef * lam                   (V-cycle rhs adjoint, coef = 1 / c)
//   kAdjSrc:   acc += coef * src * lam             (Jacobi mode: src = U, acc = lamU, coef = 2 alpha beta / c)
//   kAdjFirst: acc  = 0 + coef * lam               (first sweep of a V-cycle adjoint: no zero fill before it)
//   kAdjLast:  lamU += coef2 * src * (acc + coef * lam)
//              (last sweep: the accumulated rhs adjoint goes straight into lamU, src = U)
// The same FP32 operations as accumulating into a zeroed acc and adding
// coef2 U acc to lamU in a pass of its own.  An i-march: lam of planes i - 1
// and i stay in registers.
template <int MODE>
__global__ __launch_bounds__(256) void k_smooth_adj(int S, int ich, const float* __restrict__ lam,
                                                    float* __restrict__ lam_out, const float* __restrict__ src,
                                                    float* __restrict__ acc, float* __restrict__ lamU, float a,
                                                    float c, float coef, float coef2) {
  const int n = S - 2;
  M3 m;
  if (!coords(n, ich, m)) return;
  const int S2 = S * S;
  int id = m.i0 * S2 + m.j * S + m.k;
  float lm = lam[id - S2], lc = lam[id];
  for (int i = m.i0; i < m.i1; ++i, id += S2) {
    const float lp = lam[id + S2];
    const float s6 = lp + lm + lam[id + S] + lam[id - S] + lam[id + 1] + lam[id - 1];
    float v = s6 / c;
    if (a != 0.0f) v = a * lc + v;
    lam_out[id] = v;
    if (MODE == kAdjAcc) {
      acc[id] += coef * lc;
    } else if (MODE == kAdjSrc) {
      acc[id] += coef * src[id] * lc;
    } else if (MODE == kAdjFirst) {
      acc[id] = 0.0f + coef * lc;
    } else {
      const float af = acc[id] + coef * lc;
      lamU[id] += coef2 * src[id] * af;
    }
    lm = lc;
    lc = lp;
  }
}

// V-cycle correction adjoint: lam_out = lam - (6 z - sum6 z); acc += z.
__global__ __launch_bounds__(256) void k_correct_adj(int S, int ich, const float* __restrict__ lam,
                                                     const float* __restrict__ z, float* __restrict__ lam_out,
                                                     float* __restrict__ acc) {
  const int n = S - 2;
  M3 m;
  if (!coords(n, ich, m)) return;
  const int S2 = S * S;
  // an i-march: z of planes i - 1 and i stay in registers
  int id = m.i0 * S2 + m.j * S + m.k;
  float zm = z[id - S2], zc = z[id];
  for (int i = m.i0; i < m.i1; ++i, id += S2) {
    const float zp = z[id + S2];
    const float s6 = zp + zm + z[id + S] + z[id - S] + z[id + 1] + z[id - 1];
    lam_out[id] = lam[id] - (6.0f * zc - s6);
    acc[id] += zc;
    zm = zc;
    zc = zp;
  }
}

// lamU += coef * U * acc
__global__ __launch_bounds__(256) void k_axpy_src(int S, int ich, const float* __restrict__ U,
                                                  const float* __restrict__ acc, float* __restrict__ lamU,
                                                  float coef) {
  const int n = S - 2;
  M3 m;
  if (!coords(n, ich, m)) return;
  const int S2 = S * S;
  for (int i = m.i0; i < m.i1; ++i) {
    const int id = i * S2 + m.j * S + m.k;
    lamU[id] += coef * U[id] * acc[id];
  }
}

// Per-domain sum of (A * B) * w, FP64 runs into fixed-point sums (labels
// clipped to n_dom - 1).
__global__ __launch_bounds__(256) void k_domain_dot(int S, int ich, float w, const float* __restrict__ A,
                                                    const float* __restrict__ B, const uint16_t* __restrict__ lab,
                                                    int n_dom, Fx* __restrict__ out) {
  extern __shared__ Fx fbins[];
  const int nthr = blockDim.x * blockDim.y;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  for (int t = tid; t < kFxWords * n_dom; t += nthr) fbins[t] = 0;
  __syncthreads();
  const int n = S - 2;
  M3 m;
  if (coords(n, ich, m)) {
    const int S2 = S * S;
    int cur = -1;
    double accv = 0.0;
    for (int i = m.i0; i < m.i1; ++i) {
      const int id = i * S2 + m.j * S + m.k;
      int L = lab[id] & kLabelMask;
      L = L < n_dom ? L : n_dom - 1;
      const float term = (A[id] * B[id]) * w;
      if (L != cur) {
        if (cur >= 0) fx_add(fbins + kFxWords * cur, accv);
        cur = L;
        accv = 0.0;
      }
      accv += (double)term;
    }
    if (cur >= 0) fx_add(fbins + kFxWords * cur, accv);
  }
  __syncthreads();
  for (int t = tid; t < n_dom; t += nthr) fx_add_words(out + kFxWords * t, fbins + kFxWords * t);
}

// Normalise adjoint: for domains with Z > 0 and n > 1e-20,
//   lam <- s (lam - (h^3 / n) U_pre m),  s = sqrt(Z / n),  m = sum_d U_pre lam.
__global__ __launch_bounds__(256) void k_normalize_adj(int S, int ich, float h3, const float* __restrict__ Upre,
                                                       float* __restrict__ lam, const uint16_t* __restrict__ lab,
                                                       const Fx* __restrict__ norms,
                                                       const Fx* __restrict__ dots, const float* __restrict__ Z,
                                                       int n_dom) {
  __shared__ float sc[kMaxDomains];
  __shared__ float be[kMaxDomains];
  const int nthr = blockDim.x * blockDim.y;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  for (int t = tid; t < n_dom; t += nthr) {
    const double z = (double)Z[t], nv = fx_value(norms + kFxWords * t);
    if (z > 0.0 && nv > 1e-20) {
      const double s = sqrt(z / nv);
      sc[t] = (float)s;
      be[t] = (float)(s * (double)h3 * fx_value(dots + kFxWords * t) / nv);
    } else {
      sc[t] = 1.0f;
      be[t] = 0.0f;
    }
  }
  __syncthreads();
  const int n = S - 2;
  M3 m;
  if (!coords(n, ich, m)) return;
  const int S2 = S * S;
  for (int i = m.i0; i < m.i1; ++i) {
    const int id = i * S2 + m.j * S + m.k;
    const int L = lab[id] & kLabelMask;
    if (L < n_dom) lam[id] = sc[L] * lam[id] - be[L] * Upre[id];
  }
}

// dE/dR (per bohr) from gK = dE/dK and lamU0 = dE/dU0 (pre-normalisation
// trial):  sum_v gK_v Z d/s^3 + [L_v = n] lamU0_v Z U0_v d / s,  d = x_v - R_n,
// s = sqrt(d^2 + eps h^2), terms vanish where r~ is clamped at rc.
constexpr int kCol = 16;
constexpr int kChunk = 4;

__global__ __launch_bounds__(256) void k_pos_grad(int S, float h, float h2, const float* __restrict__ gK,
                                                  const float* __restrict__ lamU0, const uint16_t* __restrict__ lab,
                                                  const Atom* __restrict__ atoms, int n_atoms,
                                                  Fx* __restrict__ out) {
  __shared__ double red[4][kChunk * 3];
  const int n = S - 2;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  const int k = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int i0 = 1 + blockIdx.z * kCol;
  const bool active = (k <= n && j <= n);
  const int S2 = S * S;
  float gk[kCol], lu[kCol];
  int lb[kCol];
#pragma unroll
  for (int c = 0; c < kCol; ++c) {
    const int i = i0 + c;
    gk[c] = 0.0f;
    lu[c] = 0.0f;
    lb[c] = -1;
    if (active && i <= n) {
      const int id = i * S2 + j * S + k;
      gk[c] = gK[id];
      lu[c] = lamU0[id];
      lb[c] = lab[id] & kLabelMask;
    }
  }
  const float yj = (float)j * h, zk = (float)k * h;
  for (int a0 = 0; a0 < n_atoms; a0 += kChunk) {
    double acc[kChunk * 3];
#pragma unroll
    for (int q = 0; q < kChunk * 3; ++q) acc[q] = 0.0;
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      const int a = a0 + q;
      if (a >= n_atoms) break;
      const Atom at = atoms[a];
      if (!(at.Z > 0.0f)) continue;
      const float eps = (at.rc > 0.0f) ? 0.04f * h2 : h2;
      const float dy = yj - at.j * h;
      const float dz = zk - at.k * h;
#pragma unroll
      for (int c = 0; c < kCol; ++c) {
        if (lb[c] < 0) continue;
        const float xi = (float)(i0 + c) * h;
        const float dx = xi - at.i * h;
        const float s2 = dx * dx + dy * dy + dz * dz + eps;
        const float s = sqrtf(s2);
        if (at.rc > 0.0f && s < at.rc) continue;  // clamped: no dependence on R
        float w = gk[c] * at.Z / (s2 * s);
        if (lb[c] == a) {
          const float u0 = at.Z * at.Z * 0.5641895835f * expf(-at.Z * s);
          w += lu[c] * at.Z * u0 / s;
        }
        acc[q * 3 + 0] += (double)(w * dx);
        acc[q * 3 + 1] += (double)(w * dy);
        acc[q * 3 + 2] += (double)(w * dz);
      }
    }
#pragma unroll
    for (int q = 0; q < kChunk * 3; ++q) acc[q] = wsum(acc[q]);
    const int wave = tid >> 6;
    if ((tid & 63) == 0) {
#pragma unroll
      for (int q = 0; q < kChunk * 3; ++q) red[wave][q] = acc[q];
    }
    __syncthreads();
    if (tid < kChunk * 3) {
      const int a = a0 + tid / 3;
      if (a < n_atoms) fx_add(out + kFxWords * (a * 3 + tid % 3), red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid]);
    }
    __syncthreads();
  }
}

}  // namespace

static inline unsigned nb256(size_t n) { return (unsigned)((n + 255) / 256); }

void launch_init_frozen(hipStream_t st, int S, float h, float h2, const Atom* atoms, int n_atoms, const uint16_t* lab,
                        float* U, float* K) {
  const size_t S3 = (size_t)S * S * S;
  hipLaunchKernelGGL(k_init_frozen, dim3(nb256(S3)), dim3(256), 0, st, S, h, h2, atoms, n_atoms, lab, U, K);
}

void launch_energy_grad(hipStream_t st, const Tile& t, int S, float h, float h3, const float* U, const float* K,
                        const float* P, const uint16_t* lab, float* lamU, float* gK, float* lamP) {
  hipLaunchKernelGGL(k_energy_grad, sgrid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, h, h3, U, K, P, lab, lamU,
                     gK, lamP);
}

void launch_itp_adj(hipStream_t st, const Tile& t, int S, const float* lam, float* lam_out, const float* U,
                    const float* K, const float* P, const uint16_t* lab, float half_dv, float dtv, float* gK,
                    float* lamP) {
  hipLaunchKernelGGL(k_itp_adj, sgrid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, lam, lam_out, U, K, P, lab,
                     half_dv, dtv, gK, lamP);
}

void launch_smooth_adj(hipStream_t st, const Tile& t, int S, int mode, const float* lam, float* lam_out,
                       const float* src, float* acc, float* lamU, float a, float c, float coef, float coef2) {
  const dim3 g = sgrid(S - 2, t), b(t.bx, t.by);
  switch (mode) {
    case kAdjAcc:
      hipLaunchKernelGGL(k_smooth_adj<kAdjAcc>, g, b, 0, st, S, t.ich, lam, lam_out, src, acc, lamU, a, c, coef, coef2);
      break;
    case kAdjSrc:
      hipLaunchKernelGGL(k_smooth_adj<kAdjSrc>, g, b, 0, st, S, t.ich, lam, lam_out, src, acc, lamU, a, c, coef, coef2);
      break;
    case kAdjFirst:
      hipLaunchKernelGGL(k_smooth_adj<kAdjFirst>, g, b, 0, st, S, t.ich, lam, lam_out, src, acc, lamU, a, c, coef,
                         coef2);
      break;
    default:
      hipLaunchKernelGGL(k_smooth_adj<kAdjLast>, g, b, 0, st, S, t.ich, lam, lam_out, src, acc, lamU, a, c, coef,
                         coef2);
      break;
  }
}

void launch_correct_adj(hipStream_t st, const Tile& t, int S, const float* lam, const float* z, float* lam_out,
                        float* acc) {
  hipLaunchKernelGGL(k_correct_adj, sgrid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, lam, z, lam_out, acc);
}

void launch_axpy_src(hipStream_t st, const Tile& t, int S, const float* U, const float* acc, float* lamU,
                     float coef) {
  hipLaunchKernelGGL(k_axpy_src, sgrid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, U, acc, lamU, coef);
}

void launch_domain_dot(hipStream_t st, const Tile& t, int S, float w, const float* A, const float* B,
                       const uint16_t* lab, int n_dom, Fx* out) {
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void*)k_domain_dot, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr_set = true;
  }
  hipLaunchKernelGGL(k_domain_dot, sgrid(S - 2, t), dim3(t.bx, t.by), sizeof(Fx) * kFxWords * n_dom, st, S, t.ich, w,
                     A, B, lab, n_dom, out);
}

void launch_normalize_adj(hipStream_t st, const Tile& t, int S, float h3, const float* Upre, float* lam,
                          const uint16_t* lab, const Fx* norms, const Fx* dots, const float* Z, int n_dom) {
  hipLaunchKernelGGL(k_normalize_adj, sgrid(S - 2, t), dim3(t.bx, t.by), 0, st, S, t.ich, h3, Upre, lam, lab, norms,
                     dots, Z, n_dom);
}

void launch_pos_grad(hipStream_t st, int S, float h, float h2, const float* gK, const float* lamU0,
                     const uint16_t* lab, const Atom* atoms, int n_atoms, Fx* out) {
  const int n = S - 2;
  dim3 grid((n + 63) / 64, (n + 3) / 4, (n + kCol - 1) / kCol);
  hipLaunchKernelGGL(k_pos_grad, grid, dim3(64, 4), 0, st, S, h, h2, gK, lamU0, lab, atoms, n_atoms, out);
}

}  // namespace rq
