// Per-node and per-column bodies of the reverse-mode (adjoint) passes,
// shared by the single-grid kernels (rq_adjoint.hip) and the batch engine's
// kernels (rq_batch.hip), so a batched molecule's reverse sweep runs the
// same operations as its own handle's.  Forward map per step: Poisson
// (linear in P and in rhs = 2 pi h^2 U^2), ITP U' = U + (dv/2) L_N U +
// dv h^2 (K - 2P) U with the symmetric label-Neumann Laplacian L_N,
// per-domain normalise.
#pragma once

#include <hip/hip_runtime.h>

#include "rq_internal.h"

namespace rq {

template <typename T>
__device__ __forceinline__ T adj_wsum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// K(R) and the frozen-label trial U0 = Z_L^2/sqrt(pi) exp(-Z_L r~_L) at node
// id (cloud/solver.py:36-57 with the argmax replaced by the given label).
// The map that the adjoint differentiates forms 1/r~ with the hardware
// reciprocal square root and reciprocal (1 ulp each; Z / max(r, rc) =
// Z min(1 / r, 1 / rc)): K
// then carries ~1e-7 relative rounding, the same order as the IEEE form's,
// at a third of its instructions per node and atom.
__device__ __forceinline__ void init_frozen_node(int S, int id, float h, float h2, const Atom* __restrict__ atoms,
                                                 int n_atoms, const uint16_t* __restrict__ lab, float* __restrict__ U,
                                                 float* __restrict__ K) {
  const int S2 = S * S;
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
    const float s2 = r2 + ((rc > 0.0f) ? 0.04f * h2 : h2);
    const float rs = __builtin_amdgcn_rsqf(s2);
    const float inv_r = (rc > 0.0f) ? fminf(rs, __builtin_amdgcn_rcpf(rc)) : rs;
    Kv += Za * inv_r;
    if (n == L) {
      const float r = (rc > 0.0f) ? fmaxf(s2 * rs, rc) : s2 * rs;
      u0 = Za * Za * 0.5641895835f * expf(-Za * r);
    }
  }
  K[id] = Kv;
  U[id] = u0;
}

// Seed of the backward pass: dE/dU, dE/dK, dE/dP of E = T + V_eK + V_ee
// (cloud/solver.py:227-257) at the final state, along one column.
__device__ __forceinline__ void energy_grad_col(int S, int k, int j, int i0, int i1, float h, float h3,
                                                const float* __restrict__ U, const float* __restrict__ K,
                                                const float* __restrict__ P, const uint16_t* __restrict__ lab,
                                                float* __restrict__ lamU, float* __restrict__ gK,
                                                float* __restrict__ lamP) {
  const int S2 = S * S;
  const int d[3] = {S2, S, 1};
  for (int i = i0; i < i1; ++i) {
    const int id = i * S2 + j * S + k;
    const int L = lab[id] & kLabelMask;
    const float u = U[id];
    const int coord[3] = {i, j, k};
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
// gK += dtv U lam; lamP += -2 dtv U lam.  An i-march: labels and lam of
// planes i - 1 and i stay in registers.
__device__ __forceinline__ void itp_adj_col(int S, int k, int j, int i0, int i1, const float* __restrict__ lam,
                                            float* __restrict__ lam_out, const float* __restrict__ U,
                                            const float* __restrict__ K, const float* __restrict__ P,
                                            const uint16_t* __restrict__ lab, float half_dv, float dtv,
                                            float* __restrict__ gK, float* __restrict__ lamP) {
  const int S2 = S * S;
  int id = i0 * S2 + j * S + k;
  uint16_t lbm = lab[id - S2], lc = lab[id];
  float lvm = lam[id - S2], lc_v = lam[id];
  for (int i = i0; i < i1; ++i, id += S2) {
    const uint16_t lbp = lab[id + S2];
    const float lvp = lam[id + S2];
    // the in-plane neighbours loaded unconditionally, then selected (a load
    // under the label test would be a branch and a dependent wait each)
    const uint16_t ljp = lab[id + S], ljm = lab[id - S], lkp = lab[id + 1], lkm = lab[id - 1];
    const float vjp = lam[id + S], vjm = lam[id - S], vkp = lam[id + 1], vkm = lam[id - 1];
    const float t0 = (lbp == lc) ? lvp : lc_v;
    const float t1 = (lbm == lc) ? lvm : lc_v;
    const float t2 = (ljp == lc) ? vjp : lc_v;
    const float t3 = (ljm == lc) ? vjm : lc_v;
    const float t4 = (lkp == lc) ? vkp : lc_v;
    const float t5 = (lkm == lc) ? vkm : lc_v;
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
__device__ __forceinline__ void smooth_adj_col(int S, int k, int j, int i0, int i1, const float* __restrict__ lam,
                                               float* __restrict__ lam_out, const float* __restrict__ src,
                                               float* __restrict__ acc, float* __restrict__ lamU, float a, float c,
                                               float coef, float coef2) {
  const int S2 = S * S;
  int id = i0 * S2 + j * S + k;
  float lm = lam[id - S2], lc = lam[id];
  for (int i = i0; i < i1; ++i, id += S2) {
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
__device__ __forceinline__ void correct_adj_col(int S, int k, int j, int i0, int i1, const float* __restrict__ lam,
                                                const float* __restrict__ z, float* __restrict__ lam_out,
                                                float* __restrict__ acc) {
  const int S2 = S * S;
  int id = i0 * S2 + j * S + k;
  float zm = z[id - S2], zc = z[id];
  for (int i = i0; i < i1; ++i, id += S2) {
    const float zp = z[id + S2];
    const float s6 = zp + zm + z[id + S] + z[id - S] + z[id + 1] + z[id - 1];
    lam_out[id] = lam[id] - (6.0f * zc - s6);
    acc[id] += zc;
    zm = zc;
    zc = zp;
  }
}

// lamU += coef * U * acc
__device__ __forceinline__ void axpy_src_col(int S, int k, int j, int i0, int i1, const float* __restrict__ U,
                                             const float* __restrict__ acc, float* __restrict__ lamU, float coef) {
  const int S2 = S * S;
  for (int i = i0; i < i1; ++i) {
    const int id = i * S2 + j * S + k;
    lamU[id] += coef * U[id] * acc[id];
  }
}

// Per-domain sum of (A * B) * w along a column: FP64 runs into LDS
// fixed-point bins (labels clipped to n_dom - 1).
__device__ __forceinline__ void domain_dot_col(int S, int k, int j, int i0, int i1, float w,
                                               const float* __restrict__ A, const float* __restrict__ B,
                                               const uint16_t* __restrict__ lab, int n_dom, Fx* bins) {
  const int S2 = S * S;
  int cur = -1;
  double accv = 0.0;
  for (int i = i0; i < i1; ++i) {
    const int id = i * S2 + j * S + k;
    int L = lab[id] & kLabelMask;
    L = L < n_dom ? L : n_dom - 1;
    const float term = (A[id] * B[id]) * w;
    if (L != cur) {
      if (cur >= 0) fx_add(bins + kFxWords * cur, accv);
      cur = L;
      accv = 0.0;
    }
    accv += (double)term;
  }
  if (cur >= 0) fx_add(bins + kFxWords * cur, accv);
}

// Normalise-adjoint coefficients of one domain: for Z > 0 and n > 1e-20,
// lam <- s (lam - (h^3 / n) U_pre m), s = sqrt(Z / n), m = sum_d U_pre lam;
// sc = s, be = s h^3 m / n (else 1, 0).
__device__ __forceinline__ void normalize_adj_coef(float Zd, const Fx* __restrict__ norm, const Fx* __restrict__ dot,
                                                   float h3, float& sc, float& be) {
  const double z = (double)Zd, nv = fx_value(norm);
  if (z > 0.0 && nv > 1e-20) {
    const double s = sqrt(z / nv);
    sc = (float)s;
    be = (float)(s * (double)h3 * fx_value(dot) / nv);
  } else {
    sc = 1.0f;
    be = 0.0f;
  }
}

__device__ __forceinline__ void normalize_adj_col(int S, int k, int j, int i0, int i1, const float* __restrict__ Upre,
                                                  float* __restrict__ lam, const uint16_t* __restrict__ lab,
                                                  const float* sc, const float* be, int n_dom) {
  const int S2 = S * S;
  for (int i = i0; i < i1; ++i) {
    const int id = i * S2 + j * S + k;
    const int L = lab[id] & kLabelMask;
    if (L < n_dom) lam[id] = sc[L] * lam[id] - be[L] * Upre[id];
  }
}

// dE/dR (per bohr) of one 64 x 4-column, kAdjCol-plane block from gK = dE/dK
// and lamU0 = dE/dU0 (pre-normalisation trial):  sum_v gK_v Z d/s^3 +
// [L_v = n] lamU0_v Z U0_v d / s,  d = x_v - R_n, s = sqrt(d^2 + eps h^2),
// terms vanish where r~ is clamped at rc.  1/s comes from the hardware
// reciprocal square root (1 ulp), and each thread sums its kAdjCol planes'
// terms per atom in FP32 before one conversion: a gradient needs no IEEE
// division per node and atom, and the 4e11 node-atom terms of a 3000-atom
// 512^3 grid are the whole cost.  Per-atom FP64 sums are reduced across the
// wave and the four waves (red: 4 x kAdjChunk x 3 doubles of LDS) and added
// to the fixed-point result once per workgroup and component.
constexpr int kAdjCol = 16;
constexpr int kAdjChunk = 4;

__device__ __forceinline__ void pos_grad_block(int S, float h, float h2, int k, int j, int i0,
                                               const float* __restrict__ gK, const float* __restrict__ lamU0,
                                               const uint16_t* __restrict__ lab, const Atom* __restrict__ atoms,
                                               int n_atoms, Fx* __restrict__ out, double (*red)[kAdjChunk * 3]) {
  const int n = S - 2;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  const bool active = (k <= n && j <= n);
  const int S2 = S * S;
  float gk[kAdjCol], lu[kAdjCol];
  int lb[kAdjCol];
#pragma unroll
  for (int c = 0; c < kAdjCol; ++c) {
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
  for (int a0 = 0; a0 < n_atoms; a0 += kAdjChunk) {
    double acc[kAdjChunk * 3];
#pragma unroll
    for (int q = 0; q < kAdjChunk * 3; ++q) acc[q] = 0.0;
#pragma unroll
    for (int q = 0; q < kAdjChunk; ++q) {
      const int a = a0 + q;
      if (a >= n_atoms) break;
      const Atom at = atoms[a];
      if (!(at.Z > 0.0f)) continue;
      const float eps = (at.rc > 0.0f) ? 0.04f * h2 : h2;
      const float dy = yj - at.j * h;
      const float dz = zk - at.k * h;
      float fx = 0.0f, fy = 0.0f, fz = 0.0f;
#pragma unroll
      for (int c = 0; c < kAdjCol; ++c) {
        if (lb[c] < 0) continue;
        const float xi = (float)(i0 + c) * h;
        const float dx = xi - at.i * h;
        const float s2 = dx * dx + dy * dy + dz * dz + eps;
        const float rs = __builtin_amdgcn_rsqf(s2);
        const float s = s2 * rs;
        if (at.rc > 0.0f && s < at.rc) continue;  // clamped: no dependence on R
        float w = gk[c] * at.Z * (rs * rs * rs);
        if (lb[c] == a) {
          const float u0 = at.Z * at.Z * 0.5641895835f * expf(-at.Z * s);
          w += lu[c] * at.Z * u0 * rs;
        }
        fx += w * dx;
        fy += w * dy;
        fz += w * dz;
      }
      acc[q * 3 + 0] = (double)fx;
      acc[q * 3 + 1] = (double)fy;
      acc[q * 3 + 2] = (double)fz;
    }
#pragma unroll
    for (int q = 0; q < kAdjChunk * 3; ++q) acc[q] = adj_wsum(acc[q]);
    const int wave = tid >> 6;
    if ((tid & 63) == 0) {
#pragma unroll
      for (int q = 0; q < kAdjChunk * 3; ++q) red[wave][q] = acc[q];
    }
    __syncthreads();
    if (tid < kAdjChunk * 3) {
      const int a = a0 + tid / 3;
      if (a < n_atoms) fx_add(out + kFxWords * (a * 3 + tid % 3), red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid]);
    }
    __syncthreads();
  }
}

}  // namespace rq
