// Batched solver: many molecules with uneven grids advanced together.
//
// Every pass of the step (V-cycle levels, ITP, normalise, energy, init) is
// one launch over the concatenated tile space of all molecules: one
// workgroup per work item (the grid rounded up to a multiple of 8; the
// grid-stride loop only matters beyond 2^20 items), each item decodes its
// molecule by binary search over per-level descriptors and runs the same
// per-column body as the single-grid kernels (rq_stencil.h).  The per-step
// launch count is that of one molecule, independent of the batch size, and
// step chunks replay from HIP graphs.  Labels are frozen (no boundary rule),
// the schedule is the cloud/solver.py loop body with a multigrid Poisson step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rq_internal.h"
#include "rq_adjoint_cols.h"
#include "rq_march.h"
#include "rq_stencil.h"

namespace rq {
namespace {

constexpr int kMaxBatchDom = 256;  // domains per molecule (LDS bins per work item)

struct BLev {
  int S, ich, gx, gy, gz;
  int tile0;     // first tile of this molecule in the level's tile space
  int cnode0;    // first coarse interior node (restriction items from this level)
  int ctile0;    // first coarse tile (kBatchResCK x CJ x CI coarse nodes) of the tiled restriction
  int dtile0;    // first tile of the down march (kDownCK x kDownCJ coarse columns, down_ci planes)
  int utile0;    // first tile of the up march (kUpTK x kUpTJ columns, up_ci planes)
  long long off; // element offset of this molecule in the level's arrays
};

// coarse tile of the batch's residual + restriction (the single-grid tile)
constexpr int kBatchResCK = 8, kBatchResCJ = 8, kBatchResCI = 4;
// fine-level marches (the single grid's shapes): down = two pre-sweeps +
// residual + restriction over kDownCK x kDownCJ coarse columns, up =
// correction + two post-sweeps + ITP over kUpTK x kUpTJ fine columns
constexpr int kDownCK = 32, kDownCJ = 8, kDownNT = 768, kUpTK = 32, kUpTJ = 8, kUpRU = 5;
constexpr int kDownCI = 16, kUpCI = 32;  // default coarse / fine planes per march item

struct BMol {
  int atom0, n_atoms, dom0, node0;
  float half_dv, dtv, h, h2, h3, two_pi;
  int dst0, rden0;  // offsets into the coarsest-level DST tables
};

__device__ __forceinline__ int find_mol(const BLev* __restrict__ L, int nm, int t) {
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L[mid].tile0 <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int find_cnode(const BLev* __restrict__ L, int nm, int t) {
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L[mid].cnode0 <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int find_ctile(const BLev* __restrict__ L, int nm, int t) {
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L[mid].ctile0 <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int find_dtile(const BLev* __restrict__ L, int nm, int t) {
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L[mid].dtile0 <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int find_utile(const BLev* __restrict__ L, int nm, int t) {
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L[mid].utile0 <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int find_node(const BMol* __restrict__ M, int nm, int t) {
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (M[mid].node0 <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Work item t -> XCD-contiguous item (the grid is a multiple of 8, so item
// t runs on XCD t mod 8).
__device__ __forceinline__ int xcd_item(int t, int n) {
  const int per = n >> 3, rem = n & 7;
  const int xcd = t & 7, local = t >> 3;
  return (xcd < rem) ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

struct Col {
  int m, k, j, i0, i1;
};

__device__ __forceinline__ bool item_col(const BLev* __restrict__ L, int nm, int ntiles, int t, Col& c) {
  const int u = xcd_item(t, ntiles);
  const int m = find_mol(L, nm, u);
  const BLev d = L[m];
  const int local = u - d.tile0;
  const int tx = local % d.gx, ty = (local / d.gx) % d.gy, tz = local / (d.gx * d.gy);
  const int n = d.S - 2;
  c.m = m;
  c.k = 1 + tx * blockDim.x + threadIdx.x;
  c.j = 1 + ty * blockDim.y + threadIdx.y;
  c.i0 = 1 + tz * d.ich;
  c.i1 = min(c.i0 + d.ich, n + 1);
  return c.k <= n && c.j <= n;
}

template <bool FINE>
__global__ __launch_bounds__(256) void kb_smooth(const BLev* __restrict__ L, const BMol* __restrict__ M, int nm,
                                                 int ntiles, const float* __restrict__ P, float* __restrict__ Pn,
                                                 const float* __restrict__ src, float a, float c) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    Col q;
    if (!item_col(L, nm, ntiles, t, q)) continue;
    const long long o = L[q.m].off;
    const float al = FINE ? M[q.m].two_pi : 1.0f, be = FINE ? M[q.m].h2 : 1.0f;
    smooth_col<FINE>(L[q.m].S, q.k, q.j, q.i0, q.i1, P + o, Pn + o, src + o, a, c, al, be);
  }
}

__global__ __launch_bounds__(256) void kb_smooth2z(const BLev* __restrict__ L, int nm, int ntiles,
                                                   const float* __restrict__ g, float* __restrict__ e, float a,
                                                   float c) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    Col q;
    if (!item_col(L, nm, ntiles, t, q)) continue;
    const long long o = L[q.m].off;
    smooth2z_col(L[q.m].S, q.k, q.j, q.i0, q.i1, g + o, e + o, a, c);
  }
}

template <bool FINE>
__global__ __launch_bounds__(256) void kb_prolong_smooth(const BLev* __restrict__ Lf, const BLev* __restrict__ Lc,
                                                         const BMol* __restrict__ M, int nm, int ntiles,
                                                         const float* __restrict__ x, const float* __restrict__ e,
                                                         float* __restrict__ x_out, const float* __restrict__ src,
                                                         float a, float c) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    Col q;
    if (!item_col(Lf, nm, ntiles, t, q)) continue;
    const long long o = Lf[q.m].off, oc = Lc[q.m].off;
    const float al = FINE ? M[q.m].two_pi : 1.0f, be = FINE ? M[q.m].h2 : 1.0f;
    prolong_smooth_col<FINE>(Lf[q.m].S, q.k, q.j, q.i0, q.i1, x + o, e + oc, x_out + o, src + o, a, c, al, be);
  }
}

__global__ __launch_bounds__(256) void kb_smooth_itp(const BLev* __restrict__ L, const BMol* __restrict__ M, int nm,
                                                     int ntiles, const float* __restrict__ P, float* __restrict__ Pn,
                                                     const float* __restrict__ U, float* __restrict__ Un,
                                                     const float* __restrict__ K, const uint16_t* __restrict__ lab,
                                                     float a, float c) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    Col q;
    if (!item_col(L, nm, ntiles, t, q)) continue;
    const long long o = L[q.m].off;
    const BMol mm = M[q.m];
    smooth_itp_col(L[q.m].S, q.k, q.j, q.i0, q.i1, P + o, Pn + o, U + o, Un + o, K + o, lab + o, a, c, mm.two_pi,
                   mm.h2, mm.half_dv, mm.dtv);
  }
}

__global__ __launch_bounds__(256) void kb_itp(const BLev* __restrict__ L, const BMol* __restrict__ M, int nm,
                                              int ntiles, const float* __restrict__ U, float* __restrict__ Un,
                                              const float* __restrict__ K, const float* __restrict__ P,
                                              const uint16_t* __restrict__ lab) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    Col q;
    if (!item_col(L, nm, ntiles, t, q)) continue;
    const long long o = L[q.m].off;
    itp_col(L[q.m].S, q.k, q.j, q.i0, q.i1, U + o, Un + o, K + o, P + o, lab + o, M[q.m].half_dv, M[q.m].dtv);
  }
}

// items = coarse interior nodes of level l+1 over all molecules
// Residual + 4 FW restriction over the concatenated coarse tiles of every
// molecule (resres_tile: residuals of a tile's fine region once into LDS,
// then 27-point sums) -- the values of kb_residual then kb_restrict4 without
// the fine residual array.
template <bool FINE>
__global__ __launch_bounds__(kBatchResCK * kBatchResCJ * kBatchResCI) void kb_resres_tiled(
    const BLev* __restrict__ Lf, const BLev* __restrict__ Lc, const BMol* __restrict__ M, int nm, int nct,
    const float* __restrict__ x, const float* __restrict__ src, float* __restrict__ gc) {
  constexpr int CK = kBatchResCK, CJ = kBatchResCJ, CI = kBatchResCI;
  __shared__ float r[(2 * CK + 1) * (2 * CJ + 1) * (2 * CI + 1)];
  for (int t = blockIdx.x; t < nct; t += gridDim.x) {
    const int u = xcd_item(t, nct);
    const int m = find_ctile(Lf, nm, u);
    const BLev d = Lf[m];
    const int nc = (d.S - 1) / 2 - 1;
    const int gx = (nc + CK - 1) / CK, gy = (nc + CJ - 1) / CJ;
    const int local = u - d.ctile0;
    const float al = FINE ? M[m].two_pi : 1.0f, be = FINE ? M[m].h2 : 1.0f;
    resres_tile<FINE, CK, CJ, CI>(d.S, 1 + (local % gx) * CK, 1 + ((local / gx) % gy) * CJ,
                                  1 + (local / (gx * gy)) * CI, x + d.off, src + d.off, al, be, gc + Lc[m].off,
                                  1.0f, r);
    __syncthreads();
  }
}

// Fine-level down march on every molecule: each work item is one tile of
// smooth2_resres_march_tile (the single grid's kernel body), x -> xn after
// two sweeps and the coarse right-hand side into gc.
__global__ __launch_bounds__(kDownNT) void kb_down_march(const BLev* __restrict__ Lf, const BLev* __restrict__ Lc,
                                                         const BMol* __restrict__ M, int nm, int nit, int CI,
                                                         const float* __restrict__ x, float* __restrict__ xn,
                                                         const float* __restrict__ src, float a, float c,
                                                         float* __restrict__ gc) {
  for (int t = blockIdx.x; t < nit; t += gridDim.x) {
    const int u = xcd_item(t, nit);
    const int m = find_dtile(Lf, nm, u);
    const BLev d = Lf[m];
    smooth2_resres_march_tile<kDownCK, kDownCJ, kDownNT>(d.S, CI, u - d.dtile0, x + d.off, xn + d.off, src + d.off,
                                                         a, c, M[m].two_pi, M[m].h2, gc + Lc[m].off);
    __syncthreads();
  }
}

// Fine-level up march on every molecule: post_march_tile per work item
// (correction from e, two sweeps x -> xn, ITP update U -> Un).
__global__ __launch_bounds__(kUpTK * kUpTJ) void kb_up_march(const BLev* __restrict__ Lf, const BLev* __restrict__ Lc,
                                                             const BMol* __restrict__ M, int nm, int nit, int CI,
                                                             const float* __restrict__ x, const float* __restrict__ e,
                                                             float* __restrict__ xn, const float* __restrict__ U,
                                                             float* __restrict__ Un, const float* __restrict__ K,
                                                             const uint16_t* __restrict__ lab, float a, float c) {
  for (int t = blockIdx.x; t < nit; t += gridDim.x) {
    const int u = xcd_item(t, nit);
    const int m = find_utile(Lf, nm, u);
    const BLev d = Lf[m];
    const BMol mm = M[m];
    post_march_tile<kUpTK, kUpTJ, kUpRU>(d.S, CI, u - d.utile0, x + d.off, e + Lc[m].off, xn + d.off, U + d.off,
                                         Un + d.off, K + d.off, lab + d.off, a, c, mm.two_pi, mm.h2, mm.half_dv,
                                         mm.dtv);
    __syncthreads();
  }
}

template <bool FINE>
__global__ __launch_bounds__(256) void kb_residual_restrict(const BLev* __restrict__ Lf, const BLev* __restrict__ Lc,
                                                            const BMol* __restrict__ M, int nm, int nitems,
                                                            const float* __restrict__ x,
                                                            const float* __restrict__ src, float* __restrict__ gc) {
  const int stride = gridDim.x * blockDim.x * blockDim.y;
  for (int t = blockIdx.x * blockDim.x * blockDim.y + threadIdx.y * blockDim.x + threadIdx.x; t < nitems;
       t += stride) {
    const int m = find_cnode(Lf, nm, t);
    const int Sf = Lf[m].S;
    const int nc = (Sf - 1) / 2 - 1;
    const int local = t - Lf[m].cnode0;
    const float al = FINE ? M[m].two_pi : 1.0f, be = FINE ? M[m].h2 : 1.0f;
    residual_restrict_pt<FINE>(Sf, local / (nc * nc) + 1, (local / nc) % nc + 1, local % nc + 1, x + Lf[m].off,
                               src + Lf[m].off, al, be, gc + Lc[m].off, 1.0f);
  }
}

__global__ __launch_bounds__(256) void kb_prolong_correct(const BLev* __restrict__ Lf, const BLev* __restrict__ Lc,
                                                          int nm, int ntiles, float* __restrict__ x,
                                                          const float* __restrict__ e) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    Col q;
    if (!item_col(Lf, nm, ntiles, t, q)) continue;
    prolong_correct_col(Lf[q.m].S, q.k, q.j, q.i0, q.i1, x + Lf[q.m].off, e + Lc[q.m].off, 1.0f);
  }
}

// exact coarsest solve, one workgroup per molecule (ping-pong DST as in the single-grid kernel)
__device__ __forceinline__ void dst_axis_b(const double* __restrict__ in, double* __restrict__ out,
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

__global__ __launch_bounds__(1024) void kb_exact(const BLev* __restrict__ L, const BMol* __restrict__ M,
                                                 const float* __restrict__ g, float* __restrict__ e,
                                                 const double* __restrict__ Stab, const double* __restrict__ Rtab) {
  extern __shared__ double lds[];
  const int mol = blockIdx.x;
  const int Sc = L[mol].S;
  const int m = Sc - 2, m3 = m * m * m;
  const long long o = L[mol].off;
  const double* Smat = Stab + M[mol].dst0;
  const double* rden = Rtab + M[mol].rden0;
  double* Sm = lds;
  double* A = Sm + m * m;
  double* B = A + m3;
  const int Sc2 = Sc * Sc;
  for (int t = threadIdx.x; t < m * m; t += blockDim.x) Sm[t] = Smat[t];
  for (int t = threadIdx.x; t < m3; t += blockDim.x) {
    const int a = t / (m * m), b = (t / m) % m, c = t % m;
    A[t] = (double)g[o + (a + 1) * Sc2 + (b + 1) * Sc + (c + 1)];
  }
  __syncthreads();
  dst_axis_b(A, B, Sm, m, 2);
  __syncthreads();
  dst_axis_b(B, A, Sm, m, 1);
  __syncthreads();
  dst_axis_b(A, B, Sm, m, 0);
  __syncthreads();
  for (int t = threadIdx.x; t < m3; t += blockDim.x) B[t] *= rden[t];
  __syncthreads();
  dst_axis_b(B, A, Sm, m, 2);
  __syncthreads();
  dst_axis_b(A, B, Sm, m, 1);
  __syncthreads();
  dst_axis_b(B, A, Sm, m, 0);
  __syncthreads();
  for (int t = threadIdx.x; t < m3; t += blockDim.x) {
    const int a = t / (m * m), b = (t / m) % m, c = t % m;
    e[o + (a + 1) * Sc2 + (b + 1) * Sc + (c + 1)] = (float)A[t];
  }
}

__global__ __launch_bounds__(256) void kb_norm_sums(const BLev* __restrict__ L, const BMol* __restrict__ M, int nm,
                                                    int ntiles, const float* __restrict__ U,
                                                    const uint16_t* __restrict__ lab, Fx* __restrict__ norms) {
  __shared__ Fx bins[kFxWords * kMaxBatchDom];
  const int nthr = blockDim.x * blockDim.y;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int u = xcd_item(t, ntiles);
    const int m = find_mol(L, nm, u);  // uniform per workgroup
    const int nd = M[m].n_atoms;
    for (int d = tid; d < kFxWords * nd; d += nthr) bins[d] = 0;
    __syncthreads();
    Col q;
    if (item_col(L, nm, ntiles, t, q))
      norm_sums_col(L[m].S, q.k, q.j, q.i0, q.i1, M[m].h3, U + L[m].off, lab + L[m].off, nd, bins);
    __syncthreads();
    for (int d = tid; d < nd; d += nthr) fx_add_words(norms + kFxWords * (M[m].dom0 + d), bins + kFxWords * d);
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void kb_normalize(const BLev* __restrict__ L, const BMol* __restrict__ M, int nm,
                                                    int ntiles, float* __restrict__ U,
                                                    const uint16_t* __restrict__ lab,
                                                    const Fx* __restrict__ norms, const float* __restrict__ Z) {
  __shared__ float scale[kMaxBatchDom];
  const int nthr = blockDim.x * blockDim.y;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int u = xcd_item(t, ntiles);
    const int m = find_mol(L, nm, u);
    const int nd = M[m].n_atoms, d0 = M[m].dom0;
    __syncthreads();
    for (int d = tid; d < nd; d += nthr) {
      const double z = (double)Z[d0 + d], nv = fx_value(norms + kFxWords * (d0 + d));
      scale[d] = (z > 0.0 && nv > 1e-20) ? (float)sqrt(z / nv) : 1.0f;
    }
    __syncthreads();
    Col q;
    if (item_col(L, nm, ntiles, t, q)) normalize_col(L[m].S, q.k, q.j, q.i0, q.i1, U + L[m].off, lab + L[m].off, scale, nd);
  }
}

__global__ __launch_bounds__(256) void kb_energy(const BLev* __restrict__ L, const BMol* __restrict__ M, int nm,
                                                 int ntiles, const float* __restrict__ U, const float* __restrict__ K,
                                                 const float* __restrict__ P, const uint16_t* __restrict__ lab,
                                                 Fx* __restrict__ out) {
  __shared__ Fx bins[kFxWords * 3 * kMaxBatchDom];
  const int nthr = blockDim.x * blockDim.y;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int u = xcd_item(t, ntiles);
    const int m = find_mol(L, nm, u);
    const int nd = M[m].n_atoms;
    for (int d = tid; d < kFxWords * 3 * nd; d += nthr) bins[d] = 0;
    __syncthreads();
    Col q;
    if (item_col(L, nm, ntiles, t, q)) {
      const long long o = L[m].off;
      energy_s1_col(L[m].S, q.k, q.j, q.i0, q.i1, M[m].h, M[m].h3, U + o, K + o, P + o, lab + o, nd, bins);
    }
    __syncthreads();
    for (int d = tid; d < 3 * nd; d += nthr) fx_add_words(out + kFxWords * (3 * M[m].dom0 + d), bins + kFxWords * d);
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void kb_init(const BMol* __restrict__ M, const BLev* __restrict__ L, int nm,
                                               int nnodes, const Atom* __restrict__ atoms, float* __restrict__ U,
                                               uint16_t* __restrict__ lab, float* __restrict__ K) {
  const int stride = gridDim.x * blockDim.x;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nnodes; t += stride) {
    const int m = find_node(M, nm, t);
    const long long o = L[m].off;
    init_trial_s1_node(L[m].S, t - M[m].node0, M[m].h, M[m].h2, atoms + M[m].atom0, M[m].n_atoms, U + o, lab + o,
                       K + o);
  }
}

// ---- reverse sweep of the frozen-label map over all molecules (the
// single-grid adjoint's column bodies, rq_adjoint_cols.h) ----

__global__ __launch_bounds__(256) void kb_init_frozen(const BMol* __restrict__ M, const BLev* __restrict__ L, int nm,
                                                      int nnodes, const Atom* __restrict__ atoms,
                                                      const uint16_t* __restrict__ lab, float* __restrict__ U,
                                                      float* __restrict__ K) {
  const int stride = gridDim.x * blockDim.x;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nnodes; t += stride) {
    const int m = find_node(M, nm, t);
    const long long o = L[m].off;
    init_frozen_node(L[m].S, t - M[m].node0, M[m].h, M[m].h2, atoms + M[m].atom0, M[m].n_atoms, lab + o, U + o,
                     K + o);
  }
}

__global__ __launch_bounds__(256) void kb_energy_grad(const BLev* __restrict__ L, const BMol* __restrict__ M, int nm,
                                                      int ntiles, const float* __restrict__ U,
                                                      const float* __restrict__ K, const float* __restrict__ P,
                                                      const uint16_t* __restrict__ lab, float* __restrict__ lamU,
                                                      float* __restrict__ gK, float* __restrict__ lamP) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    Col q;
    if (!item_col(L, nm, ntiles, t, q)) continue;
    const long long o = L[q.m].off;
    energy_grad_col(L[q.m].S, q.k, q.j, q.i0, q.i1, M[q.m].h, M[q.m].h3, U + o, K + o, P + o, lab + o, lamU + o,
                    gK + o, lamP + o);
  }
}

__global__ __launch_bounds__(256) void kb_itp_adj(const BLev* __restrict__ L, const BMol* __restrict__ M, int nm,
                                                  int ntiles, const float* __restrict__ lam,
                                                  float* __restrict__ lam_out, const float* __restrict__ U,
                                                  const float* __restrict__ K, const float* __restrict__ P,
                                                  const uint16_t* __restrict__ lab, float* __restrict__ gK,
                                                  float* __restrict__ lamP) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    Col q;
    if (!item_col(L, nm, ntiles, t, q)) continue;
    const long long o = L[q.m].off;
    itp_adj_col(L[q.m].S, q.k, q.j, q.i0, q.i1, lam + o, lam_out + o, U + o, K + o, P + o, lab + o, M[q.m].half_dv,
                M[q.m].dtv, gK + o, lamP + o);
  }
}

// coef2 of kAdjLast is each molecule's 2 alpha beta = 2 (2 pi) h^2
template <int MODE>
__global__ __launch_bounds__(256) void kb_smooth_adj(const BLev* __restrict__ L, const BMol* __restrict__ M, int nm,
                                                     int ntiles, const float* __restrict__ lam,
                                                     float* __restrict__ lam_out, const float* __restrict__ src,
                                                     float* __restrict__ acc, float* __restrict__ lamU, float a,
                                                     float c, float coef) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    Col q;
    if (!item_col(L, nm, ntiles, t, q)) continue;
    const long long o = L[q.m].off;
    const float ab2 = 2.0f * M[q.m].two_pi * M[q.m].h2;
    smooth_adj_col<MODE>(L[q.m].S, q.k, q.j, q.i0, q.i1, lam + o, lam_out + o, src ? src + o : nullptr, acc + o,
                         lamU ? lamU + o : nullptr, a, c, coef, ab2);
  }
}

__global__ __launch_bounds__(256) void kb_correct_adj(const BLev* __restrict__ L, int nm, int ntiles,
                                                      const float* __restrict__ lam, const float* __restrict__ z,
                                                      float* __restrict__ lam_out, float* __restrict__ acc) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    Col q;
    if (!item_col(L, nm, ntiles, t, q)) continue;
    const long long o = L[q.m].off;
    correct_adj_col(L[q.m].S, q.k, q.j, q.i0, q.i1, lam + o, z + o, lam_out + o, acc + o);
  }
}

__global__ __launch_bounds__(256) void kb_axpy_src(const BLev* __restrict__ L, const BMol* __restrict__ M, int nm,
                                                   int ntiles, const float* __restrict__ U,
                                                   const float* __restrict__ acc, float* __restrict__ lamU) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    Col q;
    if (!item_col(L, nm, ntiles, t, q)) continue;
    const long long o = L[q.m].off;
    axpy_src_col(L[q.m].S, q.k, q.j, q.i0, q.i1, U + o, acc + o, lamU + o, 2.0f * M[q.m].two_pi * M[q.m].h2);
  }
}

__global__ __launch_bounds__(256) void kb_domain_dot(const BLev* __restrict__ L, const BMol* __restrict__ M, int nm,
                                                     int ntiles, const float* __restrict__ A,
                                                     const float* __restrict__ B, const uint16_t* __restrict__ lab,
                                                     Fx* __restrict__ out) {
  __shared__ Fx bins[kFxWords * kMaxBatchDom];
  const int nthr = blockDim.x * blockDim.y;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int u = xcd_item(t, ntiles);
    const int m = find_mol(L, nm, u);  // uniform per workgroup
    const int nd = M[m].n_atoms;
    for (int d = tid; d < kFxWords * nd; d += nthr) bins[d] = 0;
    __syncthreads();
    Col q;
    if (item_col(L, nm, ntiles, t, q)) {
      const long long o = L[m].off;
      domain_dot_col(L[m].S, q.k, q.j, q.i0, q.i1, 1.0f, A + o, B + o, lab + o, nd, bins);
    }
    __syncthreads();
    for (int d = tid; d < nd; d += nthr) fx_add_words(out + kFxWords * (M[m].dom0 + d), bins + kFxWords * d);
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void kb_normalize_adj(const BLev* __restrict__ L, const BMol* __restrict__ M, int nm,
                                                        int ntiles, const float* __restrict__ Upre,
                                                        float* __restrict__ lam, const uint16_t* __restrict__ lab,
                                                        const Fx* __restrict__ norms, const Fx* __restrict__ dots,
                                                        const float* __restrict__ Z) {
  __shared__ float sc[kMaxBatchDom], be[kMaxBatchDom];
  const int nthr = blockDim.x * blockDim.y;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int u = xcd_item(t, ntiles);
    const int m = find_mol(L, nm, u);
    const int nd = M[m].n_atoms, d0 = M[m].dom0;
    __syncthreads();
    for (int d = tid; d < nd; d += nthr)
      normalize_adj_coef(Z[d0 + d], norms + kFxWords * (d0 + d), dots + kFxWords * (d0 + d), M[m].h3, sc[d], be[d]);
    __syncthreads();
    Col q;
    if (item_col(L, nm, ntiles, t, q)) {
      const long long o = L[m].off;
      normalize_adj_col(L[m].S, q.k, q.j, q.i0, q.i1, Upre + o, lam + o, lab + o, sc, be, nd);
    }
  }
}

// position gradient: per molecule the single-grid block layout (64 x 4
// columns, kAdjCol planes), blocks numbered from ptile0[m]
__global__ __launch_bounds__(256) void kb_pos_grad(const int* __restrict__ ptile0, const BLev* __restrict__ L,
                                                   const BMol* __restrict__ M, int nm, int nblocks,
                                                   const float* __restrict__ gK, const float* __restrict__ lamU0,
                                                   const uint16_t* __restrict__ lab, const Atom* __restrict__ atoms,
                                                   Fx* __restrict__ out) {
  __shared__ double red[4][kAdjChunk * 3];
  for (int t = blockIdx.x; t < nblocks; t += gridDim.x) {
    int lo = 0, hi = nm - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (ptile0[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    const int m = lo;
    const int S = L[m].S, n = S - 2;
    const int gx = (n + 63) / 64, gy = (n + 3) / 4;
    const int local = t - ptile0[m];
    const int k = 1 + (local % gx) * 64 + threadIdx.x;
    const int j = 1 + ((local / gx) % gy) * 4 + threadIdx.y;
    const int i0 = 1 + (local / (gx * gy)) * kAdjCol;
    const long long o = L[m].off;
    pos_grad_block(S, M[m].h, M[m].h2, k, j, i0, gK + o, lamU0 + o, lab + o, atoms + M[m].atom0, M[m].n_atoms,
                   out + kFxWords * 3 * (size_t)M[m].atom0, red);
  }
}

}  // namespace
}  // namespace rq

using namespace rq;

namespace {
thread_local std::string g_batch_error;

struct BatchPtrs {
  float *U, *U2, *P, *P2;
  std::vector<float*> e0, e1;
  bool operator==(const BatchPtrs& o) const {
    return U == o.U && U2 == o.U2 && P == o.P && P2 == o.P2 && e0 == o.e0 && e1 == o.e1;
  }
};

struct BAdj {
  float *lamU = nullptr, *lamU2 = nullptr, *lamP = nullptr, *lamP2 = nullptr, *gK = nullptr, *acc = nullptr;
  float *t1 = nullptr, *t3 = nullptr, *zero = nullptr;
  std::vector<float*> ckU, ckP, segU, segP;
  Fx* red = nullptr;
  int* ptile0 = nullptr;  // first position-gradient block of each molecule
  int npg = 0;
};

struct BatchGraph {
  BatchPtrs before, after;
  int pi, ni;
  hipGraph_t graph;
  hipGraphExec_t exec;
};
}  // namespace

struct rq_batch {
  rq_params p{};
  int nm = 0, L = 0;
  int bx = 64, by = 4;
  std::vector<int> NN;
  std::vector<double> hh, dv;
  std::vector<std::vector<BLev>> lev;  // [level][mol]
  std::vector<BMol> mol;
  std::vector<int> ntiles, ncnodes, nctiles;  // per level
  int ndtiles = 0, nutiles = 0;                // fine-level march items
  int down_ci = kDownCI, up_ci = kUpCI;        // planes per march item (RQ_BATCH_DOWN_CI / RQ_BATCH_UP_CI)
  bool march = true;                           // fine-level marches when nu1 = nu2 = 2 (RQ_BATCH_MARCH=0: off)
  std::vector<size_t> len;             // elements per level array
  std::vector<int> maxS;               // largest node count per axis on each level
  int nnodes = 0, ndom = 0;
  int grid_stencil = 2048;
  bool fuse_small = false;  // one-pass residual+restriction / prolongation+sweep on levels <= 65^3 (RQ_BATCH_FUSE_SMALL=1)
  bool fuse_itp = true;     // last post-sweep carries the ITP update (RQ_BATCH_FUSE_ITP=0: two passes)
  BLev* lev_d = nullptr;  // L * nm
  BMol* mol_d = nullptr;
  Atom* atoms_d = nullptr;
  float* Z_d = nullptr;
  double *Stab = nullptr, *Rtab = nullptr;
  size_t lds_exact = 0;
  float *U = nullptr, *U2 = nullptr, *K = nullptr, *P = nullptr, *P2 = nullptr;
  uint16_t* lab = nullptr;
  std::vector<float*> e0, e1, g;  // per level (index 0 unused)
  Fx* red = nullptr;              // fixed-point sums: norms (ndom) + energies (3 ndom)
  hipStream_t st = nullptr;
  bool initialized = false, use_graph = false, timing = false;
  std::vector<BatchGraph> graphs;
  std::vector<Atom> atoms_h;  // all molecules' atoms, in order
  BAdj adj;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
  double fam_ms[F_COUNT] = {0};
  int64_t fam_n[F_COUNT] = {0};
  std::string err;
};

namespace {

int bfail(rq_batch* b, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (b) b->err = buf;
  else g_batch_error = buf;
  return code;
}

#define BCHK(b, call)                                                                         \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) return bfail((b), RQ_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define BLAUNCH(b)                                                                                  \
  do {                                                                                              \
    hipError_t e_ = hipGetLastError();                                                              \
    if (e_ != hipSuccess) return bfail((b), RQ_EHIP, "kernel launch: %s", hipGetErrorString(e_)); \
  } while (0)

struct BTimed {
  rq_batch* b;
  int fam;
  hipEvent_t a = nullptr;
  BTimed(rq_batch* b_, int f) : b(b_), fam(f) {
    if (b->timing) {
      (void)hipEventCreateWithFlags(&a, hipEventDisableSystemFence);
      (void)hipEventRecord(a, b->st);
    }
  }
  ~BTimed() {
    if (b->timing && a) {
      hipEvent_t e = nullptr;
      (void)hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
      (void)hipEventRecord(e, b->st);
      b->ev.push_back({fam, {a, e}});
    }
  }
};

void resolve(rq_batch* b) {
  if (b->ev.empty()) return;
  (void)hipStreamSynchronize(b->st);
  for (auto& x : b->ev) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, x.second.first, x.second.second) == hipSuccess) {
      b->fam_ms[x.first] += ms;
      b->fam_n[x.first] += 1;
    }
    (void)hipEventDestroy(x.second.first);
    (void)hipEventDestroy(x.second.second);
  }
  b->ev.clear();
}

const BLev* LD(rq_batch* b, int l) { return b->lev_d + (size_t)l * b->nm; }

// One workgroup per item up to 2^20 items (the grid-stride loops then run
// once): the hardware keeps as many workgroups resident as occupancy allows,
// which a capped persistent grid (2048, 8 per CU) did not -- it left the
// fine passes at ~3 TB/s.  A multiple of 8 keeps item t on XCD t mod 8.
int grid_for(int items) { return std::max(8, std::min(((items + 7) / 8) * 8, 1 << 20)); }

int b_vcycle_coarse(rq_batch* b, int l) {
  const float a = b->p.mg_a, c = b->p.mg_c;
  if (l == b->L - 1) {
    BTimed t(b, F_COARSE);
    hipLaunchKernelGGL(kb_exact, dim3(b->nm), dim3(1024), b->lds_exact, b->st, LD(b, l), b->mol_d, b->g[l], b->e0[l],
                       b->Stab, b->Rtab);
    BLAUNCH(b);
    return RQ_OK;
  }
  const dim3 blk(b->bx, b->by);
  const int nt = b->ntiles[l];
  int pre = b->p.mg_nu1;
  if (pre >= 2) {
    BTimed t(b, F_SMOOTH_C);
    hipLaunchKernelGGL(kb_smooth2z, dim3(grid_for(nt)), blk, 0, b->st, LD(b, l), b->nm, nt, b->g[l], b->e0[l], a, c);
    pre -= 2;
  } else {
    launch_fill(b->st, b->e0[l], b->len[l], 0.0f);
  }
  for (int it = 0; it < pre; ++it) {
    {
      BTimed t(b, F_SMOOTH_C);
      hipLaunchKernelGGL(kb_smooth<false>, dim3(grid_for(nt)), blk, 0, b->st, LD(b, l), b->mol_d, b->nm, nt, b->e0[l],
                         b->e1[l], b->g[l], a, c);
    }
    std::swap(b->e0[l], b->e1[l]);
  }
  // separate residual and restriction passes and separate correction and
  // sweeps: the one-pass variants recompute (27 residuals per coarse node, 7
  // prolongations per sweep node) and measured slower on gfx950
  const bool small = b->fuse_small && b->maxS[l] <= kFuseMaxS;
  {
    BTimed t(b, F_RESTRICT_C);
    const int ni = b->ncnodes[l];
    if (small) {
      hipLaunchKernelGGL(kb_residual_restrict<false>, dim3(std::min((ni + 255) / 256, 4096)), dim3(256), 0, b->st,
                         LD(b, l), LD(b, l + 1), b->mol_d, b->nm, ni, b->e0[l], b->g[l], b->g[l + 1]);
    } else {
      const int nct = b->nctiles[l];
      hipLaunchKernelGGL(kb_resres_tiled<false>, dim3(grid_for(nct)), dim3(kBatchResCK * kBatchResCJ * kBatchResCI),
                         0, b->st, LD(b, l), LD(b, l + 1), b->mol_d, b->nm, nct, b->e0[l], b->g[l], b->g[l + 1]);
    }
  }
  BLAUNCH(b);
  if (int rc = b_vcycle_coarse(b, l + 1)) return rc;
  int post = b->p.mg_nu2;
  if (small && post >= 1) {
    {
      BTimed t(b, F_PROLONG_C);
      hipLaunchKernelGGL(kb_prolong_smooth<false>, dim3(grid_for(nt)), blk, 0, b->st, LD(b, l), LD(b, l + 1),
                         b->mol_d, b->nm, nt, b->e0[l], b->e0[l + 1], b->e1[l], b->g[l], a, c);
    }
    std::swap(b->e0[l], b->e1[l]);
    --post;
  } else {
    BTimed t(b, F_PROLONG_C);
    hipLaunchKernelGGL(kb_prolong_correct, dim3(grid_for(nt)), blk, 0, b->st, LD(b, l), LD(b, l + 1), b->nm, nt,
                       b->e0[l], b->e0[l + 1]);
  }
  for (int it = 0; it < post; ++it) {
    {
      BTimed t(b, F_SMOOTH_C);
      hipLaunchKernelGGL(kb_smooth<false>, dim3(grid_for(nt)), blk, 0, b->st, LD(b, l), b->mol_d, b->nm, nt, b->e0[l],
                         b->e1[l], b->g[l], a, c);
    }
    std::swap(b->e0[l], b->e1[l]);
  }
  BLAUNCH(b);
  return RQ_OK;
}

// one V-cycle on all molecules; the last post-sweep carries the ITP update when fuse_itp
int b_vcycle(rq_batch* b, bool fuse_itp, bool* did_itp) {
  *did_itp = false;
  const float a = b->p.mg_a, c = b->p.mg_c;
  const dim3 blk(b->bx, b->by);
  const int nt = b->ntiles[0];
  const dim3 grd(grid_for(nt));
  // the fine level as two marches (the single grid's fused V-cycle legs):
  // same values as the separate sweeps, restriction, correction and ITP
  if (b->march && fuse_itp && b->p.mg_nu1 == 2 && b->p.mg_nu2 == 2 && !(b->fuse_small && b->maxS[0] <= kFuseMaxS)) {
    {
      BTimed t(b, F_RESTRICT);
      hipLaunchKernelGGL(kb_down_march, dim3(grid_for(b->ndtiles)), dim3(kDownNT), 0, b->st, LD(b, 0), LD(b, 1),
                         b->mol_d, b->nm, b->ndtiles, b->down_ci, b->P, b->P2, b->U, a, c, b->g[1]);
    }
    std::swap(b->P, b->P2);
    BLAUNCH(b);
    if (int rc = b_vcycle_coarse(b, 1)) return rc;
    {
      BTimed t(b, F_SMOOTH_ITP);
      hipLaunchKernelGGL(kb_up_march, dim3(grid_for(b->nutiles)), dim3(kUpTK * kUpTJ), 0, b->st, LD(b, 0), LD(b, 1),
                         b->mol_d, b->nm, b->nutiles, b->up_ci, b->P, b->e0[1], b->P2, b->U, b->U2, b->K, b->lab, a,
                         c);
    }
    std::swap(b->P, b->P2);
    std::swap(b->U, b->U2);
    *did_itp = true;
    BLAUNCH(b);
    return RQ_OK;
  }
  for (int it = 0; it < b->p.mg_nu1; ++it) {
    {
      BTimed t(b, F_SMOOTH);
      hipLaunchKernelGGL(kb_smooth<true>, grd, blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, b->P, b->P2, b->U, a, c);
    }
    std::swap(b->P, b->P2);
  }
  const bool small = b->fuse_small && b->maxS[0] <= kFuseMaxS;
  {
    BTimed t(b, F_RESTRICT);
    const int ni = b->ncnodes[0];
    if (small) {
      hipLaunchKernelGGL(kb_residual_restrict<true>, dim3(std::min((ni + 255) / 256, 4096)), dim3(256), 0, b->st,
                         LD(b, 0), LD(b, 1), b->mol_d, b->nm, ni, b->P, b->U, b->g[1]);
    } else {
      const int nct = b->nctiles[0];
      hipLaunchKernelGGL(kb_resres_tiled<true>, dim3(grid_for(nct)), dim3(kBatchResCK * kBatchResCJ * kBatchResCI),
                         0, b->st, LD(b, 0), LD(b, 1), b->mol_d, b->nm, nct, b->P, b->U, b->g[1]);
    }
  }
  BLAUNCH(b);
  if (int rc = b_vcycle_coarse(b, 1)) return rc;
  int post = b->p.mg_nu2;
  if (small) {
    {
      BTimed t(b, F_PROLONG);
      hipLaunchKernelGGL(kb_prolong_smooth<true>, grd, blk, 0, b->st, LD(b, 0), LD(b, 1), b->mol_d, b->nm, nt, b->P,
                         b->e0[1], b->P2, b->U, a, c);
    }
    std::swap(b->P, b->P2);
    --post;
  } else {
    BTimed t(b, F_PROLONG);
    hipLaunchKernelGGL(kb_prolong_correct, grd, blk, 0, b->st, LD(b, 0), LD(b, 1), b->nm, nt, b->P, b->e0[1]);
  }
  for (int it = 0; it < post; ++it) {
    if (fuse_itp && it == post - 1) {
      {
        BTimed t(b, F_SMOOTH_ITP);
        hipLaunchKernelGGL(kb_smooth_itp, grd, blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, b->P, b->P2, b->U, b->U2,
                           b->K, b->lab, a, c);
      }
      std::swap(b->P, b->P2);
      std::swap(b->U, b->U2);
      *did_itp = true;
      break;
    }
    {
      BTimed t(b, F_SMOOTH);
      hipLaunchKernelGGL(kb_smooth<true>, grd, blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, b->P, b->P2, b->U, a, c);
    }
    std::swap(b->P, b->P2);
  }
  BLAUNCH(b);
  return RQ_OK;
}

int b_normalize(rq_batch* b) {
  const dim3 blk(b->bx, b->by);
  const int nt = b->ntiles[0];
  launch_fill(b->st, reinterpret_cast<float*>(b->red), 2 * kFxWords * (size_t)b->ndom, 0.0f);
  BTimed t(b, F_NORM);
  hipLaunchKernelGGL(kb_norm_sums, dim3(grid_for(nt)), blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, b->U, b->lab,
                     b->red);
  hipLaunchKernelGGL(kb_normalize, dim3(grid_for(nt)), blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, b->U, b->lab,
                     b->red, b->Z_d);
  BLAUNCH(b);
  return RQ_OK;
}

int b_step(rq_batch* b, int64_t step, const rq_schedule* sc) {
  bool did = false;
  if (step % sc->poisson_interval == 0)
    if (int rc = b_vcycle(b, b->fuse_itp, &did)) return rc;
  if (!did) {
    BTimed t(b, F_ITP);
    const int nt = b->ntiles[0];
    hipLaunchKernelGGL(kb_itp, dim3(grid_for(nt)), dim3(b->bx, b->by), 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, b->U,
                       b->U2, b->K, b->P, b->lab);
    BLAUNCH(b);
    std::swap(b->U, b->U2);
  }
  if (step % sc->norm_interval == 0)
    if (int rc = b_normalize(b)) return rc;
  return RQ_OK;
}

BatchPtrs bptrs(const rq_batch* b) { return BatchPtrs{b->U, b->U2, b->P, b->P2, b->e0, b->e1}; }

void set_bptrs(rq_batch* b, const BatchPtrs& p) {
  b->U = p.U;
  b->U2 = p.U2;
  b->P = p.P;
  b->P2 = p.P2;
  b->e0 = p.e0;
  b->e1 = p.e1;
}

// Drains the batch stream before destroying its cached step graphs: a
// launched executable may still be running (HIP does not document that
// hipGraphExecDestroy waits for it).
void drop_bgraphs(rq_batch* b) {
  if (b->graphs.empty()) return;
  if (b->st) (void)hipStreamSynchronize(b->st);
  for (auto& g : b->graphs) {
    (void)hipGraphExecDestroy(g.exec);
    (void)hipGraphDestroy(g.graph);
  }
  b->graphs.clear();
}

int b_graph_chunk(rq_batch* b, int64_t c0, const rq_schedule* sc) {
  const BatchPtrs cur = bptrs(b);
  for (auto& ge : b->graphs)
    if (ge.pi == sc->poisson_interval && ge.ni == sc->norm_interval && ge.before == cur) {
      BCHK(b, hipGraphLaunch(ge.exec, b->st));
      set_bptrs(b, ge.after);
      return RQ_OK;
    }
  if (b->graphs.size() >= 8) drop_bgraphs(b);
  BatchGraph ge;
  ge.before = cur;
  ge.pi = sc->poisson_interval;
  ge.ni = sc->norm_interval;
  BCHK(b, hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
  int rc = RQ_OK;
  for (int64_t s = c0 + 1; s <= c0 + sc->norm_interval && rc == RQ_OK; ++s) rc = b_step(b, s, sc);
  hipGraph_t g = nullptr;
  hipError_t ee = hipStreamEndCapture(b->st, &g);
  if (rc != RQ_OK || ee != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    set_bptrs(b, cur);
    return rc != RQ_OK ? rc : bfail(b, RQ_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(ee));
  }
  ge.graph = g;
  BCHK(b, hipGraphInstantiate(&ge.exec, g, nullptr, nullptr, 0));
  ge.after = bptrs(b);
  b->graphs.push_back(ge);
  BCHK(b, hipGraphLaunch(ge.exec, b->st));
  return RQ_OK;
}

void bfree(rq_batch* b) {
  auto fr = [](void* p) {
    if (p) (void)hipFree(p);
  };
  drop_bgraphs(b);
  for (auto& x : b->ev) {
    (void)hipEventDestroy(x.second.first);
    (void)hipEventDestroy(x.second.second);
  }
  fr(b->lev_d);
  fr(b->mol_d);
  fr(b->atoms_d);
  fr(b->Z_d);
  fr(b->Stab);
  fr(b->Rtab);
  fr(b->U);
  fr(b->U2);
  fr(b->K);
  fr(b->P);
  fr(b->P2);
  fr(b->lab);
  fr(b->red);
  {
    BAdj& A = b->adj;
    for (float* q : {A.lamU, A.lamU2, A.lamP, A.lamP2, A.gK, A.acc, A.t1, A.t3, A.zero}) fr(q);
    for (auto* v : {&A.ckU, &A.ckP, &A.segU, &A.segP})
      for (float* q : *v) fr(q);
    fr(A.red);
    fr(A.ptile0);
  }
  for (size_t l = 0; l < b->e0.size(); ++l) {
    fr(b->e0[l]);
    fr(b->e1[l]);
    fr(b->g[l]);
  }
  if (b->st) (void)hipStreamDestroy(b->st);
}

// ---- frozen-label map and its reverse sweep over all molecules ----
// The rq_adjoint_gradient map of every molecule at once (trial start,
// multigrid Poisson): K(R) and the trial at the batch's labels, normalise,
// bootstrap V-cycles, n_steps loop-body steps, E = T + V_eK + V_ee per
// molecule; the reverse sweep mirrors frozen_run's, one launch per pass
// over all molecules' tiles.

int b_adj_alloc(rq_batch* b, float** q) {
  if (*q) return RQ_OK;
  if (hipMalloc((void**)q, b->len[0] * 4) != hipSuccess) {
    (void)hipGetLastError();
    return bfail(b, RQ_ENOMEM, "batch adjoint: allocation of a %zu-node field failed", b->len[0]);
  }
  BCHK(b, hipMemsetAsync(*q, 0, b->len[0] * 4, b->st));
  return RQ_OK;
}

int b_adj_alloc_vec(rq_batch* b, std::vector<float*>& v, size_t n) {
  while (v.size() < n) v.push_back(nullptr);
  for (size_t i = 0; i < n; ++i)
    if (int rc = b_adj_alloc(b, &v[i])) return rc;
  return RQ_OK;
}

int b_copy(rq_batch* b, float* dst, const float* src) {
  BCHK(b, hipMemcpyAsync(dst, src, b->len[0] * 4, hipMemcpyDeviceToDevice, b->st));
  return RQ_OK;
}

// per-domain normalisation of the level-0 field U (the batch's b_normalize on any field)
int b_normalize_on(rq_batch* b, float* U) {
  const dim3 blk(b->bx, b->by);
  const int nt = b->ntiles[0];
  launch_fill(b->st, reinterpret_cast<float*>(b->red), 2 * kFxWords * (size_t)b->ndom, 0.0f);
  hipLaunchKernelGGL(kb_norm_sums, dim3(grid_for(nt)), blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, U, b->lab,
                     b->red);
  hipLaunchKernelGGL(kb_normalize, dim3(grid_for(nt)), blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, U, b->lab,
                     b->red, b->Z_d);
  BLAUNCH(b);
  return RQ_OK;
}

// one Poisson step (V-cycle) of b->P on b->U; with fuse the last post-sweep
// carries the ITP update (then *did and the new U is b->U)
int b_poisson(rq_batch* b, bool fuse, bool* did) { return b_vcycle(b, fuse, did); }

int b_itp(rq_batch* b, const float* U, float* Un, const float* P) {
  const int nt = b->ntiles[0];
  hipLaunchKernelGGL(kb_itp, dim3(grid_for(nt)), dim3(b->bx, b->by), 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, U, Un,
                     b->K, P, b->lab);
  BLAUNCH(b);
  return RQ_OK;
}

// adjoint of one V-cycle Poisson step on b->adj lamP, accumulating into lamU (U: its input)
int b_poisson_adj(rq_batch* b, const float* U) {
  BAdj& A = b->adj;
  const float a = b->p.mg_a, c = b->p.mg_c;
  const dim3 blk(b->bx, b->by);
  const int nt = b->ntiles[0];
  const dim3 grd(grid_for(nt));
  if (b->p.mg_nu2 == 0) launch_fill(b->st, A.acc, b->len[0], 0.0f);
  for (int it = 0; it < b->p.mg_nu2; ++it) {
    if (it == 0)
      hipLaunchKernelGGL(kb_smooth_adj<kAdjFirst>, grd, blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, A.lamP, A.lamP2,
                         nullptr, A.acc, nullptr, a, c, 1.0f / c);
    else
      hipLaunchKernelGGL(kb_smooth_adj<kAdjAcc>, grd, blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, A.lamP, A.lamP2,
                         nullptr, A.acc, nullptr, a, c, 1.0f / c);
    std::swap(A.lamP, A.lamP2);
  }
  // plain 4 FW restriction of lamP: the residual + restriction pass with x = 0
  const int nct = b->nctiles[0];
  hipLaunchKernelGGL(kb_resres_tiled<false>, dim3(grid_for(nct)), dim3(kBatchResCK * kBatchResCJ * kBatchResCI), 0,
                     b->st, LD(b, 0), LD(b, 1), b->mol_d, b->nm, nct, A.zero, A.lamP, b->g[1]);
  BLAUNCH(b);
  if (int rc = b_vcycle_coarse(b, 1)) return rc;
  launch_fill(b->st, A.t3, b->len[0], 0.0f);
  hipLaunchKernelGGL(kb_prolong_correct, grd, blk, 0, b->st, LD(b, 0), LD(b, 1), b->nm, nt, A.t3, b->e0[1]);
  hipLaunchKernelGGL(kb_correct_adj, grd, blk, 0, b->st, LD(b, 0), b->nm, nt, A.lamP, A.t3, A.lamP2, A.acc);
  std::swap(A.lamP, A.lamP2);
  for (int it = 0; it < b->p.mg_nu1; ++it) {
    const bool last = it == b->p.mg_nu1 - 1;
    if (last)
      hipLaunchKernelGGL(kb_smooth_adj<kAdjLast>, grd, blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, A.lamP, A.lamP2,
                         U, A.acc, A.lamU, a, c, 1.0f / c);
    else
      hipLaunchKernelGGL(kb_smooth_adj<kAdjAcc>, grd, blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, A.lamP, A.lamP2,
                         nullptr, A.acc, nullptr, a, c, 1.0f / c);
    std::swap(A.lamP, A.lamP2);
  }
  if (b->p.mg_nu1 == 0)
    hipLaunchKernelGGL(kb_axpy_src, grd, blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, U, A.acc, A.lamU);
  BLAUNCH(b);
  return RQ_OK;
}

// adjoint of normalise on lamU in place, given the pre-normalise field
int b_normalize_adj(rq_batch* b, const float* Upre) {
  BAdj& A = b->adj;
  const dim3 blk(b->bx, b->by);
  const int nt = b->ntiles[0];
  Fx* norms = A.red;
  Fx* dots = A.red + kFxWords * (size_t)b->ndom;
  launch_fill(b->st, reinterpret_cast<float*>(A.red), 4 * kFxWords * (size_t)b->ndom, 0.0f);
  hipLaunchKernelGGL(kb_norm_sums, dim3(grid_for(nt)), blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, Upre, b->lab,
                     norms);
  hipLaunchKernelGGL(kb_domain_dot, dim3(grid_for(nt)), blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, Upre, A.lamU,
                     b->lab, dots);
  hipLaunchKernelGGL(kb_normalize_adj, dim3(grid_for(nt)), blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, Upre, A.lamU,
                     b->lab, norms, dots, b->Z_d);
  BLAUNCH(b);
  return RQ_OK;
}

int b_frozen_run(rq_batch* b, const rq_adjoint_spec* sp, bool backward, double* E3, double* dEdR) {
  if (!sp) return bfail(b, RQ_EINVAL, "null spec");
  for (int r : sp->reserved)
    if (r) return bfail(b, RQ_EINVAL, "rq_adjoint_spec.reserved must be zero");
  if (sp->start != RQ_ADJ_START_TRIAL) return bfail(b, RQ_EINVAL, "batches differentiate the trial-start map");
  const rq_schedule& sc = sp->schedule;
  if (sp->n_steps < 0 || sp->bootstrap < 0 || sc.poisson_interval < 1 || sc.norm_interval < 1)
    return bfail(b, RQ_EINVAL, "bad adjoint spec");
  if (backward && b->p.mg_nu1 != b->p.mg_nu2) return bfail(b, RQ_EINVAL, "adjoint of the V-cycle needs mg_nu1 == mg_nu2");
  if (!b->initialized) return bfail(b, RQ_ESTATE, "the frozen map needs the labels of rq_batch_initialize");
  BAdj& A = b->adj;
  const int64_t Tn = sp->n_steps;
  const int C = sp->checkpoint_interval > 0 ? sp->checkpoint_interval : 64;
  const int64_t nck = backward ? (Tn + C - 1) / C + 1 : 0;
  int rc;
  // fixed-point sums: norms, dots (ndom each), dE/dR (3 per atom), energies (3 per domain)
  if (!A.red) BCHK(b, hipMalloc((void**)&A.red, sizeof(Fx) * kFxWords * 8 * (size_t)b->ndom));
  if (backward) {
    for (float** q : {&A.lamU, &A.lamU2, &A.lamP, &A.lamP2, &A.gK, &A.acc, &A.t1, &A.t3, &A.zero})
      if ((rc = b_adj_alloc(b, q))) return rc;
    if ((rc = b_adj_alloc_vec(b, A.ckU, nck)) || (rc = b_adj_alloc_vec(b, A.ckP, nck))) return rc;
    const size_t segn = (size_t)std::min<int64_t>(C, std::max<int64_t>(Tn, 1));
    if ((rc = b_adj_alloc_vec(b, A.segU, segn + 1)) || (rc = b_adj_alloc_vec(b, A.segP, segn))) return rc;
    if (!A.ptile0) {
      std::vector<int> p0(b->nm);
      int tot = 0;
      for (int m = 0; m < b->nm; ++m) {
        const int n = b->lev[0][m].S - 2;
        p0[m] = tot;
        tot += ((n + 63) / 64) * ((n + 3) / 4) * ((n + kAdjCol - 1) / kAdjCol);
      }
      A.npg = tot;
      BCHK(b, hipMalloc((void**)&A.ptile0, sizeof(int) * b->nm));
      BCHK(b, hipMemcpyAsync(A.ptile0, p0.data(), sizeof(int) * b->nm, hipMemcpyHostToDevice, b->st));
    }
  }
  const int nt = b->ntiles[0];
  const dim3 blk(b->bx, b->by), grd(grid_for(nt));
  // ---- forward ----
  hipLaunchKernelGGL(kb_init_frozen, dim3(std::min((b->nnodes + 255) / 256, 65535)), dim3(256), 0, b->st, b->mol_d,
                     LD(b, 0), b->nm, b->nnodes, b->atoms_d, b->lab, b->U, b->K);
  BLAUNCH(b);
  if ((rc = b_normalize_on(b, b->U))) return rc;
  launch_fill(b->st, b->P, b->len[0], 0.0f);
  for (int i = 0; i < sp->bootstrap; ++i) {
    bool did = false;
    if ((rc = b_poisson(b, false, &did))) return rc;
  }
  if (backward && ((rc = b_copy(b, A.ckU[0], b->U)) || (rc = b_copy(b, A.ckP[0], b->P)))) return rc;
  const int64_t nseg = backward ? (Tn + C - 1) / C : 0;
  const int64_t t0_last = nseg > 0 ? (nseg - 1) * C : Tn;
  for (int64_t t = 1; t <= Tn; ++t) {
    const bool rec = backward && t > t0_last;
    const int idx = rec ? (int)(t - t0_last - 1) : 0;
    if (rec && idx == 0 && (rc = b_copy(b, A.segU[0], b->U))) return rc;
    bool did = false;
    if (t % sc.poisson_interval == 0) {
      if ((rc = b_poisson(b, b->fuse_itp, &did))) return rc;
      if (rec && (rc = b_copy(b, A.segP[idx], b->P))) return rc;
    }
    if (!did) {
      if ((rc = b_itp(b, b->U, b->U2, b->P))) return rc;
      std::swap(b->U, b->U2);
    }
    if (t % sc.norm_interval == 0 && (rc = b_normalize_on(b, b->U))) return rc;
    if (rec && t < Tn && (rc = b_copy(b, A.segU[idx + 1], b->U))) return rc;
    if (backward && t % C == 0 && t < Tn) {
      if ((rc = b_copy(b, A.ckU[t / C], b->U)) || (rc = b_copy(b, A.ckP[t / C], b->P))) return rc;
    }
  }
  // energies of the final state, per molecule
  {
    Fx* per = A.red + 5 * kFxWords * (size_t)b->ndom;
    launch_fill(b->st, reinterpret_cast<float*>(per), 6 * kFxWords * (size_t)b->ndom, 0.0f);
    hipLaunchKernelGGL(kb_energy, grd, blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, b->U, b->K, b->P, b->lab, per);
    BLAUNCH(b);
    std::vector<Fx> hv(kFxWords * 3 * (size_t)b->ndom);
    BCHK(b, hipMemcpyAsync(hv.data(), per, sizeof(Fx) * hv.size(), hipMemcpyDeviceToHost, b->st));
    BCHK(b, hipStreamSynchronize(b->st));
    for (int m = 0; m < b->nm; ++m) {
      double e[3] = {0, 0, 0};
      for (int d = 0; d < b->mol[m].n_atoms; ++d)
        for (int q = 0; q < 3; ++q) e[q] += fx_value(&hv[kFxWords * (3 * (size_t)(b->mol[m].dom0 + d) + q)]);
      if (!std::isfinite(e[0] + e[1] + e[2])) return bfail(b, RQ_EUNSTABLE, "molecule %d: non-finite energy", m);
      if (E3)
        for (int q = 0; q < 3; ++q) E3[3 * m + q] = e[q];
    }
  }
  if (!backward) return RQ_OK;
  // ---- reverse sweep ----
  hipLaunchKernelGGL(kb_energy_grad, grd, blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, b->U, b->K, b->P, b->lab,
                     A.lamU, A.gK, A.lamP);
  BLAUNCH(b);
  std::vector<const float*> Pused(C);
  for (int64_t sg = nseg - 1; sg >= 0; --sg) {
    const int64_t t0 = sg * C, t1 = std::min<int64_t>(Tn, t0 + C);
    const bool recompute = sg != nseg - 1;
    if (recompute) {
      // the segment's inputs, recomputed from its checkpoint with the batch state
      if ((rc = b_copy(b, b->U, A.ckU[sg])) || (rc = b_copy(b, b->P, A.ckP[sg]))) return rc;
      for (int64_t t = t0 + 1; t <= t1; ++t) {
        const int idx = (int)(t - t0 - 1);
        if ((rc = b_copy(b, A.segU[idx], b->U))) return rc;
        bool did = false;
        if (t % sc.poisson_interval == 0) {
          if ((rc = b_poisson(b, b->fuse_itp && t < t1, &did))) return rc;
          if ((rc = b_copy(b, A.segP[idx], b->P))) return rc;
        }
        if (t < t1) {
          if (!did) {
            if ((rc = b_itp(b, b->U, b->U2, b->P))) return rc;
            std::swap(b->U, b->U2);
          }
          if (t % sc.norm_interval == 0 && (rc = b_normalize_on(b, b->U))) return rc;
        }
      }
    }
    // P in use at each step of the segment (the last Poisson step's output)
    const float* Pprev = A.ckP[sg];
    for (int64_t t = t0 + 1; t <= t1; ++t) {
      const int idx = (int)(t - t0 - 1);
      Pused[idx] = (t % sc.poisson_interval == 0) ? A.segP[idx] : Pprev;
      Pprev = Pused[idx];
    }
    for (int64_t t = t1; t > t0; --t) {
      const int idx = (int)(t - t0 - 1);
      const float* Uin = A.segU[idx];
      if (t % sc.norm_interval == 0) {
        if ((rc = b_itp(b, Uin, A.t1, Pused[idx]))) return rc;
        if ((rc = b_normalize_adj(b, A.t1))) return rc;
      }
      hipLaunchKernelGGL(kb_itp_adj, grd, blk, 0, b->st, LD(b, 0), b->mol_d, b->nm, nt, A.lamU, A.lamU2, Uin, b->K,
                         Pused[idx], b->lab, A.gK, A.lamP);
      BLAUNCH(b);
      std::swap(A.lamU, A.lamU2);
      if (t % sc.poisson_interval == 0 && (rc = b_poisson_adj(b, Uin))) return rc;
    }
  }
  for (int i = 0; i < sp->bootstrap; ++i)
    if ((rc = b_poisson_adj(b, A.ckU[0]))) return rc;
  // initial normalise: recompute the trial field into t1 (K is unchanged)
  hipLaunchKernelGGL(kb_init_frozen, dim3(std::min((b->nnodes + 255) / 256, 65535)), dim3(256), 0, b->st, b->mol_d,
                     LD(b, 0), b->nm, b->nnodes, b->atoms_d, b->lab, A.t1, b->K);
  BLAUNCH(b);
  if ((rc = b_normalize_adj(b, A.t1))) return rc;
  Fx* gout = A.red + 2 * kFxWords * (size_t)b->ndom;  // 3 per atom (one domain per atom)
  launch_fill(b->st, reinterpret_cast<float*>(gout), 2 * 3 * kFxWords * (size_t)b->ndom, 0.0f);
  hipLaunchKernelGGL(kb_pos_grad, dim3(std::min(A.npg, 1 << 20)), dim3(64, 4), 0, b->st, A.ptile0, LD(b, 0), b->mol_d,
                     b->nm, A.npg, A.gK, A.lamU, b->lab, b->atoms_d, gout);
  BLAUNCH(b);
  std::vector<Fx> hg(kFxWords * 3 * (size_t)b->ndom);
  BCHK(b, hipMemcpyAsync(hg.data(), gout, sizeof(Fx) * hg.size(), hipMemcpyDeviceToHost, b->st));
  BCHK(b, hipStreamSynchronize(b->st));
  if (dEdR)
    for (size_t q = 0; q < 3 * (size_t)b->ndom; ++q) dEdR[q] = fx_value(&hg[kFxWords * q]);
  return RQ_OK;
}

int upload_descriptors(rq_batch* b) {
  std::vector<BLev> flat;
  for (int l = 0; l < b->L; ++l) flat.insert(flat.end(), b->lev[l].begin(), b->lev[l].end());
  BCHK(b, hipMemcpy(b->lev_d, flat.data(), sizeof(BLev) * flat.size(), hipMemcpyHostToDevice));
  return RQ_OK;
}

// tile counts and offsets of every level for the current (bx, by, ich)
void layout(rq_batch* b, const std::vector<int>* ich_fine) {
  b->ntiles.assign(b->L, 0);
  b->ncnodes.assign(b->L, 0);
  b->nctiles.assign(b->L, 0);
  for (int l = 0; l < b->L; ++l) {
    int tile = 0, cnode = 0, ctile = 0;
    for (int m = 0; m < b->nm; ++m) {
      BLev& d = b->lev[l][m];
      const int n = d.S - 2;
      if (l == 0 && ich_fine && (*ich_fine)[m] > 0) d.ich = (*ich_fine)[m];
      else if (d.ich <= 0) d.ich = std::max(1, std::min(8, n / 8));
      d.gx = (n + b->bx - 1) / b->bx;
      d.gy = (n + b->by - 1) / b->by;
      d.gz = (n + d.ich - 1) / d.ich;
      d.tile0 = tile;
      tile += d.gx * d.gy * d.gz;
      d.cnode0 = cnode;
      d.ctile0 = ctile;
      const int nc = (d.S - 1) / 2 - 1;
      if (l + 1 < b->L) {
        cnode += nc * nc * nc;
        ctile += ((nc + kBatchResCK - 1) / kBatchResCK) * ((nc + kBatchResCJ - 1) / kBatchResCJ) *
                 ((nc + kBatchResCI - 1) / kBatchResCI);
      }
    }
    b->ntiles[l] = tile;
    b->ncnodes[l] = cnode;
    b->nctiles[l] = ctile;
  }
  int dt = 0, ut = 0;
  for (int m = 0; m < b->nm; ++m) {
    BLev& d = b->lev[0][m];
    const int n = d.S - 2, nc = (d.S - 1) / 2 - 1;
    d.dtile0 = dt;
    d.utile0 = ut;
    dt += ((nc + kDownCK - 1) / kDownCK) * ((nc + kDownCJ - 1) / kDownCJ) * ((nc + b->down_ci - 1) / b->down_ci);
    ut += ((n + kUpTK - 1) / kUpTK) * ((n + kUpTJ - 1) / kUpTJ) * ((n + b->up_ci - 1) / b->up_ci);
  }
  b->ndtiles = dt;
  b->nutiles = ut;
}

}  // namespace

extern "C" {

const char* rq_batch_last_error(const rq_batch* b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

int rq_batch_create(const rq_params* pp, const rq_molecule* mols, int32_t n_mols, rq_batch** out) {
  if (!pp || !mols || !out || n_mols < 1) return bfail(nullptr, RQ_EINVAL, "null argument or empty batch");
  *out = nullptr;
  const rq_params& p = *pp;
  if (p.poisson_mode != RQ_POISSON_MULTIGRID) return bfail(nullptr, RQ_EINVAL, "batches run the multigrid Poisson mode");
  if (p.mg_nu2 < 1) return bfail(nullptr, RQ_EINVAL, "batches need mg_nu2 >= 1");
  const int L = p.mg_levels;
  if (L < 2 || L > 10) return bfail(nullptr, RQ_EINVAL, "mg_levels %d outside [2, 10]", L);
  rq_batch* b = new rq_batch();
  b->p = p;
  if (const char* fs = std::getenv("RQ_BATCH_FUSE_SMALL")) b->fuse_small = std::atoi(fs) != 0;
  if (const char* fi = std::getenv("RQ_BATCH_FUSE_ITP")) b->fuse_itp = std::atoi(fi) != 0;
  if (const char* mv = std::getenv("RQ_BATCH_MARCH")) b->march = std::atoi(mv) != 0;
  if (const char* dc = std::getenv("RQ_BATCH_DOWN_CI")) b->down_ci = std::max(1, std::atoi(dc));
  if (const char* uc = std::getenv("RQ_BATCH_UP_CI")) b->up_ci = std::max(1, std::atoi(uc));
  b->nm = n_mols;
  b->L = L;
  std::vector<Atom> atoms;
  std::vector<float> Z;
  b->lev.assign(L, std::vector<BLev>(n_mols));
  b->len.assign(L, 0);
  int node0 = 0;
  for (int m = 0; m < n_mols; ++m) {
    const rq_molecule& mo = mols[m];
    const int div = 1 << (L - 1);
    if (mo.NN < 8 || mo.NN > 1024 || mo.NN % div || !(mo.screen > 0) || mo.n_atoms < 1 ||
        mo.n_atoms > kMaxBatchDom || !mo.atoms) {
      delete b;
      return bfail(nullptr, RQ_EINVAL, "molecule %d: NN=%d (divisible by %d), n_atoms=%d (1..%d)", m, mo.NN, div,
                   mo.n_atoms, kMaxBatchDom);
    }
    const int mc = mo.NN / div - 1;
    if (mc < 1 || mc > kMaxCoarsePP) {
      delete b;
      return bfail(nullptr, RQ_ERANGE, "molecule %d: coarsest interior %d outside [1, %d]", m, mc, kMaxCoarsePP);
    }
    const double h = mo.screen / mo.NN;
    int n_elec = 0;
    for (int a = 0; a < mo.n_atoms; ++a) {
      atoms.push_back({mo.atoms[a].i, mo.atoms[a].j, mo.atoms[a].k, mo.atoms[a].Z, mo.atoms[a].rc});
      Z.push_back(mo.atoms[a].Z);
      if (mo.atoms[a].Z > 0.f) ++n_elec;
    }
    const double dv = p.dv > 0.0 ? p.dv : rq_dv_for_electrons(n_elec);
    BMol bm;
    bm.atom0 = (int)atoms.size() - mo.n_atoms;
    bm.n_atoms = mo.n_atoms;
    bm.dom0 = bm.atom0;
    bm.node0 = node0;
    node0 += (mo.NN + 1) * (mo.NN + 1) * (mo.NN + 1);
    bm.half_dv = (float)(0.5 * dv);
    bm.dtv = (float)(dv * (h * h));
    bm.h = (float)h;
    bm.h2 = (float)(h * h);
    bm.h3 = (float)std::pow(h, 3.0);
    bm.two_pi = (float)(2.0 * M_PI);
    bm.dst0 = bm.rden0 = 0;
    b->mol.push_back(bm);
    for (int l = 0; l < L; ++l) {
      BLev& d = b->lev[l][m];
      d.S = (mo.NN >> l) + 1;
      d.ich = 0;
      d.off = (long long)b->len[l];
      const size_t n3 = (size_t)d.S * d.S * d.S;
      b->len[l] += (n3 + 63) / 64 * 64;
    }
  }
  b->nnodes = node0;
  b->ndom = (int)atoms.size();
  b->atoms_h = atoms;
  b->maxS.assign(L, 0);
  for (int l = 0; l < L; ++l)
    for (int m = 0; m < n_mols; ++m) b->maxS[l] = std::max(b->maxS[l], b->lev[l][m].S);
  layout(b, nullptr);
  // DST tables of every molecule's coarsest level
  std::vector<double> Stab, Rtab;
  int mmax = 0;
  for (int m = 0; m < n_mols; ++m) {
    const int mc = b->lev[L - 1][m].S - 2;
    mmax = std::max(mmax, mc);
    b->mol[m].dst0 = (int)Stab.size();
    b->mol[m].rden0 = (int)Rtab.size();
    std::vector<double> lam(mc);
    const double nrm = std::sqrt(2.0 / (mc + 1));
    for (int pi = 0; pi < mc; ++pi) {
      lam[pi] = 2.0 - 2.0 * std::cos(M_PI * (pi + 1) / (mc + 1));
      for (int q = 0; q < mc; ++q) Stab.push_back(nrm * std::sin(M_PI * (double)((pi + 1) * (q + 1)) / (mc + 1)));
    }
    for (int a = 0; a < mc; ++a)
      for (int c = 0; c < mc; ++c)
        for (int e = 0; e < mc; ++e) Rtab.push_back(1.0 / (lam[a] + lam[c] + lam[e]));
  }
  b->lds_exact = sizeof(double) * ((size_t)mmax * mmax + 2 * (size_t)mmax * mmax * mmax);
  auto alloc = [&](void** q, size_t bytes) { return hipMalloc(q, bytes) == hipSuccess; };
  bool ok = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) == hipSuccess;
  const size_t n0 = b->len[0];
  ok = ok && alloc((void**)&b->lev_d, sizeof(BLev) * L * n_mols) && alloc((void**)&b->mol_d, sizeof(BMol) * n_mols) &&
       alloc((void**)&b->atoms_d, sizeof(Atom) * atoms.size()) && alloc((void**)&b->Z_d, sizeof(float) * Z.size()) &&
       alloc((void**)&b->Stab, sizeof(double) * Stab.size()) && alloc((void**)&b->Rtab, sizeof(double) * Rtab.size()) &&
       alloc((void**)&b->U, n0 * 4) && alloc((void**)&b->U2, n0 * 4) && alloc((void**)&b->K, n0 * 4) &&
       alloc((void**)&b->P, n0 * 4) && alloc((void**)&b->P2, n0 * 4) && alloc((void**)&b->lab, n0 * 2) &&
       alloc((void**)&b->red, sizeof(Fx) * kFxWords * 4 * (size_t)b->ndom);
  b->e0.assign(L, nullptr);
  b->e1.assign(L, nullptr);
  b->g.assign(L, nullptr);
  for (int l = 1; l < L && ok; ++l)
    ok = alloc((void**)&b->e0[l], b->len[l] * 4) && alloc((void**)&b->e1[l], b->len[l] * 4) &&
         alloc((void**)&b->g[l], b->len[l] * 4);
  if (ok) {
    ok = hipMemcpy(b->mol_d, b->mol.data(), sizeof(BMol) * n_mols, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->atoms_d, atoms.data(), sizeof(Atom) * atoms.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->Z_d, Z.data(), sizeof(float) * Z.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->Stab, Stab.data(), sizeof(double) * Stab.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->Rtab, Rtab.data(), sizeof(double) * Rtab.size(), hipMemcpyHostToDevice) == hipSuccess;
    for (float* q : {b->U, b->U2, b->K, b->P, b->P2}) ok = ok && hipMemset(q, 0, n0 * 4) == hipSuccess;
    ok = ok && hipMemset(b->lab, 0, n0 * 2) == hipSuccess;
    for (int l = 1; l < L && ok; ++l)
      for (float* q : {b->e0[l], b->e1[l], b->g[l]}) ok = ok && hipMemset(q, 0, b->len[l] * 4) == hipSuccess;
  }
  if (ok) ok = upload_descriptors(b) == RQ_OK;
  if (ok)
    ok = hipFuncSetAttribute((const void*)kb_exact, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) ==
         hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    bfree(b);
    delete b;
    return bfail(nullptr, RQ_ENOMEM, "batch allocation/upload failed (%d molecules, %zu fine nodes)", n_mols, n0);
  }
  *out = b;
  return RQ_OK;
}

int rq_batch_destroy(rq_batch* b) {
  if (!b) return RQ_OK;
  (void)hipStreamSynchronize(b->st);
  bfree(b);
  delete b;
  return RQ_OK;
}

int rq_batch_info(const rq_batch* b, int32_t* n_mols, int64_t* interior_voxels) {
  if (!b) return RQ_EINVAL;
  if (n_mols) *n_mols = b->nm;
  if (interior_voxels) {
    int64_t v = 0;
    for (int m = 0; m < b->nm; ++m) {
      const int64_t n = b->lev[0][m].S - 2;
      v += n * n * n;
    }
    *interior_voxels = v;
  }
  return RQ_OK;
}

int rq_batch_set_tiles(rq_batch* b, int32_t bx, int32_t by, const int32_t* ich) {
  if (!b) return bfail(nullptr, RQ_EINVAL, "null batch");
  if (bx < 1 || by < 1 || bx * by > 256 || (bx * by) % 64) return bfail(b, RQ_EINVAL, "tile %dx%d invalid", bx, by);
  b->bx = bx;
  b->by = by;
  std::vector<int> ichv;
  if (ich) ichv.assign(ich, ich + b->nm);
  for (int l = 0; l < b->L; ++l)
    for (auto& d : b->lev[l]) d.ich = 0;
  layout(b, ich ? &ichv : nullptr);
  drop_bgraphs(b);
  return upload_descriptors(b);
}

int rq_batch_set_graph(rq_batch* b, int32_t on) {
  if (!b) return bfail(nullptr, RQ_EINVAL, "null batch");
  b->use_graph = on != 0;
  return RQ_OK;
}

int rq_batch_initialize(rq_batch* b, int32_t bootstrap) {
  if (!b) return bfail(nullptr, RQ_EINVAL, "null batch");
  hipLaunchKernelGGL(kb_init, dim3(std::min((b->nnodes + 255) / 256, 65535)), dim3(256), 0, b->st, b->mol_d,
                     LD(b, 0), b->nm, b->nnodes, b->atoms_d, b->U, b->lab, b->K);
  BLAUNCH(b);
  launch_fill(b->st, b->P, b->len[0], 0.0f);
  if (int rc = b_normalize(b)) return rc;
  for (int i = 0; i < bootstrap; ++i) {
    bool did = false;
    if (int rc = b_vcycle(b, false, &did)) return rc;
  }
  b->initialized = true;
  return RQ_OK;
}

int rq_batch_step(rq_batch* b, int64_t first, int32_t n, const rq_schedule* sc) {
  if (!b) return bfail(nullptr, RQ_EINVAL, "null batch");
  if (!b->initialized) return bfail(b, RQ_ESTATE, "rq_batch_step before rq_batch_initialize");
  rq_schedule d;
  rq_default_schedule(&d);
  if (!sc) sc = &d;
  if (sc->poisson_interval < 1 || sc->norm_interval < 1) return bfail(b, RQ_EINVAL, "bad schedule");
  const bool graphable = b->use_graph && !b->timing && sc->norm_interval % sc->poisson_interval == 0;
  int64_t s = first + 1;
  const int64_t last = first + n;
  while (s <= last) {
    const int64_t c0 = s - 1;
    if (graphable && c0 % sc->norm_interval == 0 && c0 + sc->norm_interval <= last) {
      if (int rc = b_graph_chunk(b, c0, sc)) return rc;
      s = c0 + sc->norm_interval + 1;
      continue;
    }
    if (int rc = b_step(b, s, sc)) return rc;
    ++s;
  }
  return RQ_OK;
}

int rq_batch_set_atoms(rq_batch* b, int32_t m, const rq_atom* atoms, int32_t n) {
  if (!b || !atoms) return bfail(b, RQ_EINVAL, "null argument");
  if (m < 0 || m >= b->nm) return bfail(b, RQ_ERANGE, "molecule %d outside [0, %d)", m, b->nm);
  const BMol& mo = b->mol[m];
  if (n != mo.n_atoms) return bfail(b, RQ_EINVAL, "molecule %d has %d atoms, got %d", m, mo.n_atoms, n);
  for (int a = 0; a < n; ++a) {
    Atom& at = b->atoms_h[mo.atom0 + a];
    if (atoms[a].Z != at.Z) return bfail(b, RQ_EINVAL, "molecule %d atom %d: Z changes from %g", m, a, at.Z);
    at = {atoms[a].i, atoms[a].j, atoms[a].k, atoms[a].Z, atoms[a].rc};
  }
  BCHK(b, hipMemcpyAsync(b->atoms_d + mo.atom0, b->atoms_h.data() + mo.atom0, sizeof(Atom) * n, hipMemcpyHostToDevice,
                         b->st));
  BCHK(b, hipStreamSynchronize(b->st));
  drop_bgraphs(b);
  return RQ_OK;
}

int rq_batch_solve_frozen(rq_batch* b, const rq_adjoint_spec* spec, double* energy3) {
  if (!b) return bfail(nullptr, RQ_EINVAL, "null batch");
  return b_frozen_run(b, spec, false, energy3, nullptr);
}

int rq_batch_adjoint_gradient(rq_batch* b, const rq_adjoint_spec* spec, double* energy3, double* dEdR) {
  if (!b) return bfail(nullptr, RQ_EINVAL, "null batch");
  return b_frozen_run(b, spec, true, energy3, dEdR);
}

int rq_batch_sync(rq_batch* b) {
  if (!b) return bfail(nullptr, RQ_EINVAL, "null batch");
  BCHK(b, hipStreamSynchronize(b->st));
  return RQ_OK;
}

int rq_batch_energy(rq_batch* b, double* out3) {
  if (!b || !out3) return bfail(b, RQ_EINVAL, "null argument");
  Fx* per = b->red + kFxWords * (size_t)b->ndom;
  launch_fill(b->st, reinterpret_cast<float*>(per), 6 * kFxWords * (size_t)b->ndom, 0.0f);
  {
    BTimed t(b, F_ENERGY);
    const int nt = b->ntiles[0];
    hipLaunchKernelGGL(kb_energy, dim3(grid_for(nt)), dim3(b->bx, b->by), 0, b->st, LD(b, 0), b->mol_d, b->nm, nt,
                       b->U, b->K, b->P, b->lab, per);
  }
  BLAUNCH(b);
  std::vector<Fx> hv(kFxWords * 3 * (size_t)b->ndom);
  BCHK(b, hipMemcpyAsync(hv.data(), per, sizeof(Fx) * hv.size(), hipMemcpyDeviceToHost, b->st));
  BCHK(b, hipStreamSynchronize(b->st));
  for (int m = 0; m < b->nm; ++m) {
    double e[3] = {0, 0, 0};
    for (int d = 0; d < b->mol[m].n_atoms; ++d)
      for (int q = 0; q < 3; ++q) e[q] += fx_value(&hv[kFxWords * (3 * (size_t)(b->mol[m].dom0 + d) + q)]);
    for (int q = 0; q < 3; ++q) out3[3 * m + q] = e[q];
  }
  return RQ_OK;
}

int rq_batch_read_field(rq_batch* b, int32_t m, int32_t field, void* host, size_t bytes) {
  if (!b || !host) return bfail(b, RQ_EINVAL, "null argument");
  if (m < 0 || m >= b->nm) return bfail(b, RQ_ERANGE, "molecule %d outside [0, %d)", m, b->nm);
  const BLev& d = b->lev[0][m];
  const size_t n3 = (size_t)d.S * d.S * d.S;
  if (bytes != n3 * 4) return bfail(b, RQ_EINVAL, "molecule %d needs %zu bytes", m, n3 * 4);
  const float* src = field == RQ_FIELD_U ? b->U : field == RQ_FIELD_K ? b->K : field == RQ_FIELD_P ? b->P : nullptr;
  if (field == RQ_FIELD_LABEL) {
    std::vector<uint16_t> tmp(n3);
    BCHK(b, hipMemcpyAsync(tmp.data(), b->lab + d.off, n3 * 2, hipMemcpyDeviceToHost, b->st));
    BCHK(b, hipStreamSynchronize(b->st));
    int32_t* o = (int32_t*)host;
    for (size_t i = 0; i < n3; ++i) o[i] = tmp[i] & kLabelMask;
    return RQ_OK;
  }
  if (!src) return bfail(b, RQ_EINVAL, "field %d not available in batches (U, K, P, label)", field);
  BCHK(b, hipMemcpyAsync(host, src + d.off, n3 * 4, hipMemcpyDeviceToHost, b->st));
  BCHK(b, hipStreamSynchronize(b->st));
  return RQ_OK;
}

int rq_batch_timing_enable(rq_batch* b, int32_t on) {
  if (!b) return bfail(nullptr, RQ_EINVAL, "null batch");
  resolve(b);
  b->timing = on != 0;
  for (int f = 0; f < F_COUNT; ++f) {
    b->fam_ms[f] = 0;
    b->fam_n[f] = 0;
  }
  return RQ_OK;
}

int rq_batch_timing_read(rq_batch* b, double* ms, int64_t* launches, int32_t n) {
  if (!b) return bfail(nullptr, RQ_EINVAL, "null batch");
  resolve(b);
  for (int f = 0; f < n && f < F_COUNT; ++f) {
    if (ms) ms[f] = b->fam_ms[f];
    if (launches) launches[f] = b->fam_n[f];
  }
  return RQ_OK;
}

}  // extern "C"
