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
#include "rq_adjoint_cols.h"

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

__global__ void k_init_frozen(int S, float h, float h2, const Atom* __restrict__ atoms, int n_atoms,
                              const uint16_t* __restrict__ lab, float* __restrict__ U, float* __restrict__ K) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= S * S * S) return;
  init_frozen_node(S, id, h, h2, atoms, n_atoms, lab, U, K);
}

__global__ __launch_bounds__(256) void k_energy_grad(int S, int ich, float h, float h3, const float* __restrict__ U,
                                                     const float* __restrict__ K, const float* __restrict__ P,
                                                     const uint16_t* __restrict__ lab, float* __restrict__ lamU,
                                                     float* __restrict__ gK, float* __restrict__ lamP) {
  M3 m;
  if (!coords(S - 2, ich, m)) return;
  energy_grad_col(S, m.k, m.j, m.i0, m.i1, h, h3, U, K, P, lab, lamU, gK, lamP);
}

__global__ __launch_bounds__(256) void k_itp_adj(int S, int ich, const float* __restrict__ lam,
                                                 float* __restrict__ lam_out, const float* __restrict__ U,
                                                 const float* __restrict__ K, const float* __restrict__ P,
                                                 const uint16_t* __restrict__ lab, float half_dv, float dtv,
                                                 float* __restrict__ gK, float* __restrict__ lamP) {
  M3 m;
  if (!coords(S - 2, ich, m)) return;
  itp_adj_col(S, m.k, m.j, m.i0, m.i1, lam, lam_out, U, K, P, lab, half_dv, dtv, gK, lamP);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_smooth_adj(int S, int ich, const float* __restrict__ lam,
                                                    float* __restrict__ lam_out, const float* __restrict__ src,
                                                    float* __restrict__ acc, float* __restrict__ lamU, float a,
                                                    float c, float coef, float coef2) {
  M3 m;
  if (!coords(S - 2, ich, m)) return;
  smooth_adj_col<MODE>(S, m.k, m.j, m.i0, m.i1, lam, lam_out, src, acc, lamU, a, c, coef, coef2);
}

__global__ __launch_bounds__(256) void k_correct_adj(int S, int ich, const float* __restrict__ lam,
                                                     const float* __restrict__ z, float* __restrict__ lam_out,
                                                     float* __restrict__ acc) {
  M3 m;
  if (!coords(S - 2, ich, m)) return;
  correct_adj_col(S, m.k, m.j, m.i0, m.i1, lam, z, lam_out, acc);
}

__global__ __launch_bounds__(256) void k_axpy_src(int S, int ich, const float* __restrict__ U,
                                                  const float* __restrict__ acc, float* __restrict__ lamU,
                                                  float coef) {
  M3 m;
  if (!coords(S - 2, ich, m)) return;
  axpy_src_col(S, m.k, m.j, m.i0, m.i1, U, acc, lamU, coef);
}

__global__ __launch_bounds__(256) void k_domain_dot(int S, int ich, float w, const float* __restrict__ A,
                                                    const float* __restrict__ B, const uint16_t* __restrict__ lab,
                                                    int n_dom, Fx* __restrict__ out) {
  extern __shared__ Fx fbins[];
  const int nthr = blockDim.x * blockDim.y;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  for (int t = tid; t < kFxWords * n_dom; t += nthr) fbins[t] = 0;
  __syncthreads();
  M3 m;
  if (coords(S - 2, ich, m)) domain_dot_col(S, m.k, m.j, m.i0, m.i1, w, A, B, lab, n_dom, fbins);
  __syncthreads();
  for (int t = tid; t < n_dom; t += nthr) fx_add_words(out + kFxWords * t, fbins + kFxWords * t);
}

__global__ __launch_bounds__(256) void k_normalize_adj(int S, int ich, float h3, const float* __restrict__ Upre,
                                                       float* __restrict__ lam, const uint16_t* __restrict__ lab,
                                                       const Fx* __restrict__ norms,
                                                       const Fx* __restrict__ dots, const float* __restrict__ Z,
                                                       int n_dom) {
  __shared__ float sc[kMaxDomains];
  __shared__ float be[kMaxDomains];
  const int nthr = blockDim.x * blockDim.y;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  for (int t = tid; t < n_dom; t += nthr)
    normalize_adj_coef(Z[t], norms + kFxWords * t, dots + kFxWords * t, h3, sc[t], be[t]);
  __syncthreads();
  M3 m;
  if (!coords(S - 2, ich, m)) return;
  normalize_adj_col(S, m.k, m.j, m.i0, m.i1, Upre, lam, lab, sc, be, n_dom);
}

__global__ __launch_bounds__(256) void k_pos_grad(int S, float h, float h2, const float* __restrict__ gK,
                                                  const float* __restrict__ lamU0, const uint16_t* __restrict__ lab,
                                                  const Atom* __restrict__ atoms, int n_atoms,
                                                  Fx* __restrict__ out) {
  __shared__ double red[4][kAdjChunk * 3];
  const int k = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int i0 = 1 + blockIdx.z * kAdjCol;
  pos_grad_block(S, h, h2, k, j, i0, gK, lamU0, lab, atoms, n_atoms, out, red);
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
  dim3 grid((n + 63) / 64, (n + 3) / 4, (n + kAdjCol - 1) / kAdjCol);
  hipLaunchKernelGGL(k_pos_grad, grid, dim3(64, 4), 0, st, S, h, h2, gK, lamU0, lab, atoms, n_atoms, out);
}

}  // namespace rq
