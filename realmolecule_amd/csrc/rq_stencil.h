// Per-column stencil bodies shared by the single-grid kernels
// (rq_kernels.hip) and the batched persistent kernels (rq_batch.hip), so both
// evaluate identical FP32 expressions in identical order.
//
// A "column" is one (k, j) interior position and the i-planes [i0, i1) a
// thread marches through; S is the node count per axis of that grid.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rq_internal.h"

namespace rq {

// Access through a wave-uniform base pointer plus a per-lane 32-bit byte
// offset: the base stays in scalar registers (plane and neighbour shifts are
// scalar additions) and the access takes the offset as is, with no 64-bit
// address arithmetic per lane.  The empty asm keeps the offset in the
// access's own block: hoisted out of a loop as a 64-bit zero extension, it
// would be added per lane again.
__device__ __forceinline__ float ld_at(const float* __restrict__ base, uint32_t off) {
  asm volatile("" : "+v"(off));
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + off);
}
__device__ __forceinline__ uint16_t ld_at16(const uint16_t* __restrict__ base, uint32_t off) {
  asm volatile("" : "+v"(off));
  return *reinterpret_cast<const uint16_t*>(reinterpret_cast<const char*>(base) + off);
}
__device__ __forceinline__ void st_at(float* __restrict__ base, uint32_t off, float v) {
  asm volatile("" : "+v"(off));
  *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + off) = v;
}

// label-Neumann ITP update (cloud/solver.py:61-90; rc rule via label bit 15)
__device__ __forceinline__ void itp_col(int S, int k, int j, int i0, int i1, const float* __restrict__ U,
                                        float* __restrict__ Un, const float* __restrict__ K,
                                        const float* __restrict__ P, const uint16_t* __restrict__ lab, float half_dv,
                                        float dtv) {
  const int S2 = S * S;
  int id = i0 * S2 + j * S + k;
  float um = U[id - S2], uc = U[id];
  uint16_t lm = lab[id - S2], lc = lab[id];
  for (int i = i0; i < i1; ++i, id += S2) {
    const float up = U[id + S2];
    const uint16_t lp = lab[id + S2];
    const float ujp = U[id + S], ujm = U[id - S], ukp = U[id + 1], ukm = U[id - 1];
    const uint16_t ljp = lab[id + S], ljm = lab[id - S], lkp = lab[id + 1], lkm = lab[id - 1];
    float out;
    if (lc & kRcFlag) {
      out = 0.0f;
    } else {
      const float t0 = (lp == lc) ? up : uc;
      const float t1 = (lm == lc) ? um : uc;
      const float t2 = (ljp == lc) ? ujp : uc;
      const float t3 = (ljm == lc) ? ujm : uc;
      const float t4 = (lkp == lc) ? ukp : uc;
      const float t5 = (lkm == lc) ? ukm : uc;
      const float lap = t0 + t1 + t2 + t3 + t4 + t5 - 6.0f * uc;
      out = uc + half_dv * lap + dtv * (K[id] - 2.0f * P[id]) * uc;
    }
    Un[id] = out;
    um = uc;
    uc = up;
    lm = lc;
    lc = lp;
  }
}

// damped Jacobi P' = a P + (sum6 P + rhs) / c; FINE: rhs = (alpha U^2) beta, else rhs = g
template <bool FINE>
__device__ __forceinline__ void smooth_col(int S, int k, int j, int i0, int i1, const float* __restrict__ P,
                                           float* __restrict__ Pn, const float* __restrict__ src, float a, float c,
                                           float alpha, float beta) {
  const int S2 = S * S;
  int id = i0 * S2 + j * S + k;
  float pm = P[id - S2], pc = P[id];
#pragma unroll 2
  for (int i = i0; i < i1; ++i, id += S2) {
    const float pp = P[id + S2];
    const float s6 = pp + pm + P[id + S] + P[id - S] + P[id + 1] + P[id - 1];
    float rhs;
    if (FINE) {
      const float u = src[id];
      rhs = (alpha * (u * u)) * beta;
    } else {
      rhs = src[id];
    }
    float v = (s6 + rhs) / c;
    if (a != 0.0f) v = a * pc + v;
    Pn[id] = v;
    pm = pc;
    pc = pp;
  }
}

// two sweeps from a zero guess: e1 = g / c, e2 = a e1 + (sum6 e1 + g) / c
__device__ __forceinline__ void smooth2z_col(int S, int k, int j, int i0, int i1, const float* __restrict__ g,
                                             float* __restrict__ e_out, float a, float c) {
  const int S2 = S * S;
  for (int i = i0; i < i1; ++i) {
    const int id = i * S2 + j * S + k;
    const float e_ip = g[id + S2] / c, e_im = g[id - S2] / c, e_jp = g[id + S] / c, e_jm = g[id - S] / c;
    const float e_kp = g[id + 1] / c, e_km = g[id - 1] / c;
    const float gc = g[id];
    const float s6 = e_ip + e_im + e_jp + e_jm + e_kp + e_km;
    float v = (s6 + gc) / c;
    if (a != 0.0f) v = a * (gc / c) + v;
    e_out[id] = v;
  }
}

// Linear interpolation along one axis written without a branch: for an even
// fine index the second sample is the first one again and 0.5f * (a + a) == a
// exactly in IEEE arithmetic, so the result equals the two-case rule
// (copy at even, average at odd) bit for bit while every lane of a wave runs
// the same instructions.
__device__ __forceinline__ float lerp_k(const float* __restrict__ e, int base, int k) {
  const int kc = k >> 1;
  return 0.5f * (e[base + kc] + e[base + kc + (k & 1)]);
}

// trilinear prolongation of coarse e at fine interior node (i, j, k): k axis, then j, then i
__device__ __forceinline__ float prolong_at(const float* __restrict__ e, int Sc, int i, int j, int k) {
  const int Sc2 = Sc * Sc;
  const int ic = i >> 1, jc = j >> 1;
  const int jo = (j & 1) * Sc, io = (i & 1) * Sc2;
  const int b0 = ic * Sc2 + jc * Sc;
  const float ejk0 = 0.5f * (lerp_k(e, b0, k) + lerp_k(e, b0 + jo, k));
  const float ejk1 = 0.5f * (lerp_k(e, b0 + io, k) + lerp_k(e, b0 + io + jo, k));
  return 0.5f * (ejk0 + ejk1);
}

// bilinear (k, then j) interpolation of coarse plane ic at fine column (j, k):
// the per-plane term of prolong_at
__device__ __forceinline__ float prolong_plane(const float* __restrict__ e, int Sc, int ic, int j, int k) {
  const int b0 = ic * Sc * Sc + (j >> 1) * Sc;
  return 0.5f * (lerp_k(e, b0, k) + lerp_k(e, b0 + (j & 1) * Sc, k));
}

// x_out = smoother applied to xc = x + P e (interior; faces keep x)
template <bool FINE>
__device__ __forceinline__ void prolong_smooth_col(int Sf, int k, int j, int i0, int i1, const float* __restrict__ x,
                                                   const float* __restrict__ e, float* __restrict__ x_out,
                                                   const float* __restrict__ src, float a, float c, float alpha,
                                                   float beta) {
  const int n = Sf - 2;
  const int Sc = (Sf - 1) / 2 + 1;
  const int S2 = Sf * Sf;
  auto xc = [&](int ii, int jj, int kk) -> float {
    const int id = ii * S2 + jj * Sf + kk;
    if (ii < 1 || ii > n || jj < 1 || jj > n || kk < 1 || kk > n) return x[id];
    return x[id] + prolong_at(e, Sc, ii, jj, kk);
  };
  for (int i = i0; i < i1; ++i) {
    const int id = i * S2 + j * Sf + k;
    const float s6 = xc(i + 1, j, k) + xc(i - 1, j, k) + xc(i, j + 1, k) + xc(i, j - 1, k) + xc(i, j, k + 1) +
                     xc(i, j, k - 1);
    float rhs;
    if (FINE) {
      const float u = src[id];
      rhs = (alpha * (u * u)) * beta;
    } else {
      rhs = src[id];
    }
    float v = (s6 + rhs) / c;
    if (a != 0.0f) v = a * xc(i, j, k) + v;
    x_out[id] = v;
  }
}

// Coarse-plane interpolant cache of one fine column for a march along i:
// prolong_at(i) = 0.5 (T(i>>1) + T((i>>1) + (i&1))) with T = prolong_plane,
// each T formed once.  i advances uniformly across a workgroup, so the reuse
// branches do not diverge.
struct PlaneCache {
  int ca = -1, cb = -1;
  float ta = 0.0f, tb = 0.0f;
  __device__ __forceinline__ float corr(const float* __restrict__ e, int Sc, int i, int j, int k) {
    const int ic = i >> 1;
    if (ic != ca) {
      ta = (ic == cb) ? tb : prolong_plane(e, Sc, ic, j, k);
      ca = ic;
    }
    float t1 = ta;
    if (i & 1) {
      if (cb != ic + 1) {
        tb = prolong_plane(e, Sc, ic + 1, j, k);
        cb = ic + 1;
      }
      t1 = tb;
    }
    return 0.5f * (ta + t1);
  }
};

// last fine sweep fused with the ITP update reading the new P at the centre
__device__ __forceinline__ void smooth_itp_col(int S, int k, int j, int i0, int i1, const float* __restrict__ P,
                                               float* __restrict__ Pn, const float* __restrict__ U,
                                               float* __restrict__ Un, const float* __restrict__ K,
                                               const uint16_t* __restrict__ lab, float a, float c, float alpha,
                                               float beta, float half_dv, float dtv) {
  const int S2 = S * S;
  int id = i0 * S2 + j * S + k;
  float pm = P[id - S2], pc = P[id];
  float um = U[id - S2], uc = U[id];
  uint16_t lm = lab[id - S2], lc = lab[id];
#pragma unroll 2
  for (int i = i0; i < i1; ++i, id += S2) {
    const float pp = P[id + S2];
    const float s6 = pp + pm + P[id + S] + P[id - S] + P[id + 1] + P[id - 1];
    const float rhs = (alpha * (uc * uc)) * beta;
    float v = (s6 + rhs) / c;
    if (a != 0.0f) v = a * pc + v;
    Pn[id] = v;
    const float up = U[id + S2];
    const uint16_t lp = lab[id + S2];
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
      out = uc + half_dv * lap + dtv * (K[id] - 2.0f * v) * uc;
    }
    Un[id] = out;
    pm = pc;
    pc = pp;
    um = uc;
    uc = up;
    lm = lc;
    lc = lp;
  }
}

// residual r = rhs - (6 x - sum6 x) at one fine node
template <bool FINE>
__device__ __forceinline__ float residual_at(const float* __restrict__ x, const float* __restrict__ src, int id, int S2,
                                             int S, float alpha, float beta) {
  const float s6 = x[id + S2] + x[id - S2] + x[id + S] + x[id - S] + x[id + 1] + x[id - 1];
  float rhs;
  if (FINE) {
    const float u = src[id];
    rhs = (alpha * (u * u)) * beta;
  } else {
    rhs = src[id];
  }
  return rhs - (6.0f * x[id] - s6);
}

// coarse rhs at coarse interior node (I, J, K2) = sign * 4 FW(r), 27 residuals on the fly
template <bool FINE>
__device__ __forceinline__ void residual_restrict_pt(int Sf, int I, int J, int K2, const float* __restrict__ x,
                                                     const float* __restrict__ src, float alpha, float beta,
                                                     float* __restrict__ gc, float sign) {
  const int Sc = (Sf - 1) / 2 + 1;
  const int S2 = Sf * Sf;
  const int c = (2 * I) * S2 + (2 * J) * Sf + 2 * K2;
  const int di = S2, dj = Sf, dk = 1;
#define RQ_R(off) residual_at<FINE>(x, src, c + (off), S2, Sf, alpha, beta)
  float fs = RQ_R(di);
  fs = fs + RQ_R(-di);
  fs = fs + RQ_R(dj);
  fs = fs + RQ_R(-dj);
  fs = fs + RQ_R(dk);
  fs = fs + RQ_R(-dk);
  float es = RQ_R(di + dj);
  es = es + RQ_R(di - dj);
  es = es + RQ_R(-di + dj);
  es = es + RQ_R(-di - dj);
  es = es + RQ_R(di + dk);
  es = es + RQ_R(di - dk);
  es = es + RQ_R(-di + dk);
  es = es + RQ_R(-di - dk);
  es = es + RQ_R(dj + dk);
  es = es + RQ_R(dj - dk);
  es = es + RQ_R(-dj + dk);
  es = es + RQ_R(-dj - dk);
  float cs = RQ_R(di + dj + dk);
  cs = cs + RQ_R(di + dj - dk);
  cs = cs + RQ_R(di - dj + dk);
  cs = cs + RQ_R(di - dj - dk);
  cs = cs + RQ_R(-di + dj + dk);
  cs = cs + RQ_R(-di + dj - dk);
  cs = cs + RQ_R(-di - dj + dk);
  cs = cs + RQ_R(-di - dj - dk);
  const float val = 8.0f * RQ_R(0) + 4.0f * fs + 2.0f * es + cs;
#undef RQ_R
  float out = val * 0.0625f;
  if (sign < 0.0f) out = -out;
  gc[I * Sc * Sc + J * Sc + K2] = out;
}

// residual r = rhs - (6 x - sum6 x) along one column (written to r)
template <bool FINE>
__device__ __forceinline__ void residual_col(int S, int k, int j, int i0, int i1, const float* __restrict__ x,
                                             const float* __restrict__ src, float alpha, float beta,
                                             float* __restrict__ r) {
  const int S2 = S * S;
  int id = i0 * S2 + j * S + k;
  float xm = x[id - S2], xc = x[id];
  for (int i = i0; i < i1; ++i, id += S2) {
    const float xp = x[id + S2];
    const float s6 = xp + xm + x[id + S] + x[id - S] + x[id + 1] + x[id - 1];
    float rhs;
    if (FINE) {
      const float u = src[id];
      rhs = (alpha * (u * u)) * beta;
    } else {
      rhs = src[id];
    }
    r[id] = rhs - (6.0f * xc - s6);
    xm = xc;
    xc = xp;
  }
}

// 4 FW(r) / 4 = (8 centre + 4 faces + 2 edges + corners) / 16 about r[c0],
// where r[cm] / r[cp] are its i-1 / i+1 neighbours, dj the in-plane j
// stride, and kp / km the offsets of the k+1 / k-1 neighbours (dk and -dk
// in a plain row; other values for a row stored as its even then its odd
// entries); faces, edges and corners each summed in a fixed order
__device__ __forceinline__ float restrict27_planes_k(const float* r, int cm, int c0, int cp, int dj, int kp,
                                                     int km) {
  float fs = r[cp];
  fs = fs + r[cm];
  fs = fs + r[c0 + dj];
  fs = fs + r[c0 - dj];
  fs = fs + r[c0 + kp];
  fs = fs + r[c0 + km];
  float es = r[cp + dj];
  es = es + r[cp - dj];
  es = es + r[cm + dj];
  es = es + r[cm - dj];
  es = es + r[cp + kp];
  es = es + r[cp + km];
  es = es + r[cm + kp];
  es = es + r[cm + km];
  es = es + r[c0 + dj + kp];
  es = es + r[c0 + dj + km];
  es = es + r[c0 - dj + kp];
  es = es + r[c0 - dj + km];
  float cs = r[cp + dj + kp];
  cs = cs + r[cp + dj + km];
  cs = cs + r[cp - dj + kp];
  cs = cs + r[cp - dj + km];
  cs = cs + r[cm + dj + kp];
  cs = cs + r[cm + dj + km];
  cs = cs + r[cm - dj + kp];
  cs = cs + r[cm - dj + km];
  const float val = 8.0f * r[c0] + 4.0f * fs + 2.0f * es + cs;
  return val * 0.0625f;
}

// the same in a plain row (k stride dk)
__device__ __forceinline__ float restrict27_planes(const float* r, int cm, int c0, int cp, int dj, int dk) {
  return restrict27_planes_k(r, cm, c0, cp, dj, dk, -dk);
}

// the same about r[c] with i-stride di
__device__ __forceinline__ float restrict27(const float* r, int c, int di, int dj, int dk) {
  return restrict27_planes(r, c - di, c, c + di, dj, dk);
}

// coarse rhs = sign * 4 FW(r) at coarse interior node (I, J, K2) from a stored residual
__device__ __forceinline__ void restrict4_pt(int Sf, int I, int J, int K2, const float* __restrict__ r,
                                             float* __restrict__ gc, float sign) {
  const int Sc = (Sf - 1) / 2 + 1;
  const int S2 = Sf * Sf;
  float out = restrict27(r, (2 * I) * S2 + (2 * J) * Sf + 2 * K2, S2, Sf, 1);
  if (sign < 0.0f) out = -out;
  gc[I * Sc * Sc + J * Sc + K2] = out;
}

// x += scale * P e along one fine column, each coarse-plane interpolant
// formed once (PlaneCache; same values and operations as prolong_at).
// x += scale * P e along the column, two fine planes (2m, 2m+1) per trip over
// coarse plane m: P e = 0.5 (T(m) + T(m)) at 2m and 0.5 (T(m) + T(m+1)) at
// 2m+1 (T = prolong_plane, the prolong_at terms), T(m+1) carried to the next
// trip, so every load of a trip is issued without a branch.
// SET: x = 0 + correction, without reading x (the sum into a zero-filled x)
template <bool SET = false>
__device__ __forceinline__ void prolong_correct_col(int Sf, int k, int j, int i0, int i1, float* __restrict__ x,
                                                    const float* __restrict__ e, float scale) {
  const int Sc = (Sf - 1) / 2 + 1;
  const int S2 = Sf * Sf;
  const int col = j * Sf + k;
  auto xat = [&](int id) -> float { return SET ? 0.0f : x[id]; };
  auto apply = [&](int id, float xv, float corr) { x[id] = (scale == 1.0f) ? xv + corr : xv + scale * corr; };
  int i = i0, m = i0 >> 1;
  float ta = prolong_plane(e, Sc, m, j, k);
  if ((i & 1) && i < i1) {
    const float tb = prolong_plane(e, Sc, m + 1, j, k);
    const int id = i * S2 + col;
    apply(id, xat(id), 0.5f * (ta + tb));
    ta = tb;
    ++i;
    ++m;
  }
#pragma unroll 2
  for (; i + 1 < i1; i += 2, ++m) {
    const float tb = prolong_plane(e, Sc, m + 1, j, k);
    const int id = i * S2 + col;
    const float x0 = xat(id), x1 = xat(id + S2);
    apply(id, x0, 0.5f * (ta + ta));
    apply(id + S2, x1, 0.5f * (ta + tb));
    ta = tb;
  }
  if (i < i1) {
    const int id = i * S2 + col;
    apply(id, xat(id), 0.5f * (ta + ta));
  }
}

// run-length FP64 accumulation of (U*U)*h3 per clipped label into LDS
// fixed-point bins (fx_add); a run whose sum is not finite (a NaN or Inf
// amplitude) sets *bad when bad is given
__device__ __forceinline__ void norm_sums_col(int S, int k, int j, int i0, int i1, float h3,
                                              const float* __restrict__ U, const uint16_t* __restrict__ lab,
                                              int n_dom, Fx* bins, int* __restrict__ bad = nullptr) {
  // planes in batches of B: the batch's loads are issued together (indices
  // clamped to the column's last plane, whose repeats are skipped), then the
  // run-length accumulation walks them in order, so the sums are those of a
  // plane-by-plane walk
  constexpr int B = 8;
  const int S2 = S * S, col = j * S + k;
  int cur = -1;
  double acc = 0.0;
  for (int ib = i0; ib < i1; ib += B) {
    int Lb[B];
    float vb[B];
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const int id = min(ib + q, i1 - 1) * S2 + col;
      Lb[q] = lab[id] & kLabelMask;
      vb[q] = U[id];
    }
#pragma unroll
    for (int q = 0; q < B; ++q) {
      if (ib + q >= i1) break;
      const int L = Lb[q] < n_dom ? Lb[q] : n_dom - 1;
      const float term = (vb[q] * vb[q]) * h3;
      if (L != cur) {
        if (cur >= 0) {
          if (bad && !isfinite(acc)) *bad = 1;
          fx_add(bins + kFxWords * cur, acc);
        }
        cur = L;
        acc = 0.0;
      }
      acc += (double)term;
    }
  }
  if (cur >= 0) {
    if (bad && !isfinite(acc)) *bad = 1;
    fx_add(bins + kFxWords * cur, acc);
  }
}

// U *= scale[label] for labels inside [0, n_dom)
__device__ __forceinline__ void normalize_col(int S, int k, int j, int i0, int i1, float* __restrict__ U,
                                              const uint16_t* __restrict__ lab, const float* scale, int n_dom) {
  // planes in batches of B: labels and amplitudes loaded together (indices
  // clamped to the column's last plane, whose repeats are not stored), then
  // the batch's scale gathers, then the products
  constexpr int B = 8;
  const int S2 = S * S, col = j * S + k;
  for (int ib = i0; ib < i1; ib += B) {
    int Lb[B];
    float vb[B], sb[B];
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const int id = min(ib + q, i1 - 1) * S2 + col;
      Lb[q] = lab[id] & kLabelMask;
      vb[q] = U[id];
    }
#pragma unroll
    for (int q = 0; q < B; ++q) sb[q] = Lb[q] < n_dom ? scale[Lb[q]] : 1.0f;
#pragma unroll
    for (int q = 0; q < B; ++q)
      if (ib + q < i1 && Lb[q] < n_dom) U[(ib + q) * S2 + col] = vb[q] * sb[q];
  }
}

// S1 energy terms of one column into per-domain LDS bins (3 per domain)
__device__ __forceinline__ void energy_s1_col(int S, int k, int j, int i0, int i1, float h, float h3,
                                              const float* __restrict__ U, const float* __restrict__ K,
                                              const float* __restrict__ P, const uint16_t* __restrict__ lab,
                                              int n_dom, Fx* bins) {
  const int S2 = S * S;
  int cur = -1;
  double aT = 0.0, aK = 0.0, aE = 0.0;
  for (int i = i0; i < i1; ++i) {
    const int id = i * S2 + j * S + k;
    int L = lab[id] & kLabelMask;
    L = L < n_dom ? L : n_dom - 1;
    const float u = U[id];
    const float rho = u * u;
    const float da = ((lab[id + S2] & kLabelMask) == L) ? U[id + S2] - u : 0.0f;
    const float db = ((lab[id + S] & kLabelMask) == L) ? U[id + S] - u : 0.0f;
    const float dc = ((lab[id + 1] & kLabelMask) == L) ? U[id + 1] - u : 0.0f;
    const float T = 0.5f * (da * da + db * db + dc * dc) * h;
    const float VeK = -K[id] * rho * h3;
    const float Vee = P[id] * rho * h3;
    if (L != cur) {
      if (cur >= 0) {
        fx_add(bins + kFxWords * (cur * 3 + 0), aT);
        fx_add(bins + kFxWords * (cur * 3 + 1), aK);
        fx_add(bins + kFxWords * (cur * 3 + 2), aE);
      }
      cur = L;
      aT = aK = aE = 0.0;
    }
    aT += (double)T;
    aK += (double)VeK;
    aE += (double)Vee;
  }
  if (cur >= 0) {
    fx_add(bins + kFxWords * (cur * 3 + 0), aT);
    fx_add(bins + kFxWords * (cur * 3 + 1), aK);
    fx_add(bins + kFxWords * (cur * 3 + 2), aE);
  }
}

// trial init (S1) at one node: K, trial argmax label, best trial
__device__ __forceinline__ void init_trial_s1_node(int S, int id, float h, float h2, const Atom* __restrict__ atoms,
                                                   int n_atoms, float* __restrict__ U, uint16_t* __restrict__ lab,
                                                   float* __restrict__ K) {
  const int S2 = S * S;
  const int i = id / S2, j = (id / S) % S, k = id % S;
  const int NN = S - 1;
  if (i < 1 || i >= NN || j < 1 || j >= NN || k < 1 || k >= NN) {
    U[id] = 0.0f;
    lab[id] = 0;
    K[id] = 0.0f;
    return;
  }
  const float xi = (float)i * h, yj = (float)j * h, zk = (float)k * h;
  const float inv_sqrt_pi = 0.5641895835f;
  float bU = 0.0f, Kv = 0.0f;
  int bestN = 0;
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
  K[id] = Kv;
  lab[id] = (uint16_t)bestN;
  U[id] = bU;
}

// Workgroup barrier that orders LDS only: global loads issued before it stay
// in flight (a full __syncthreads would also wait for this wave's global
// stores, which on gfx9 share one counter with its loads).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Residual + 4 FW restriction of one coarse tile as a march along i: the
// workgroup (128 CJ threads) owns 32 (k) x CJ (j) coarse columns and CI
// coarse planes from I0; per fine plane it evaluates the residual on the
// (65 x 2CJ+1) fine region those columns' stencils cover (each thread keeps the
// i-1 / i values of its nodes in registers, so a fine plane is fetched once)
// into a (2CP+2)-plane LDS ring ``ring``.  Each round forms 2CP fine planes
// and then restricts CP coarse planes.  Same residual_at / restrict27
// expressions as the separate residual and restriction passes.  Used by the
// single-grid and the batched kernels; every thread of the workgroup calls it
// with the same tile.
template <bool FINE, int CJ, int CP>
__device__ __forceinline__ void resres_march_tile(int Sf, int CI, int K0, int J0, int I0, const float* __restrict__ x,
                                                  const float* __restrict__ src, float alpha, float beta,
                                                  float* __restrict__ gc, float sign, float* ring) {
  constexpr int CK = 32, RK = 2 * CK + 1, RJ = 2 * CJ + 1, PL = RK * RJ;
  constexpr int NT = 128 * CJ, PTS = (PL + NT - 1) / NT, RING = 2 * CP + 2;
  static_assert(CP * CK * CJ <= NT, "one thread per coarse column and plane");
  const int nc = (Sf - 1) / 2 - 1, n = Sf - 2;
  const int Sc = nc + 2, S2 = Sf * Sf;
  const int I1 = min(I0 + CI, nc + 1);
  const int tid = threadIdx.x;
  // this thread's region nodes (clamped into the interior for the loads)
  int col[PTS];
  bool in[PTS];
  float xm[PTS], xc[PTS];
  const int f0 = 2 * I0 - 1;
#pragma unroll
  for (int p = 0; p < PTS; ++p) {
    const int q = min(tid + p * NT, PL - 1);
    const int j = 2 * J0 - 1 + q / RK, k = 2 * K0 - 1 + q % RK;
    in[p] = tid + p * NT < PL && j <= n && k <= n;
    col[p] = min(j, n) * Sf + min(k, n);
    xm[p] = x[(f0 - 1) * S2 + col[p]];
    xc[p] = x[f0 * S2 + col[p]];
  }
  // residual of fine plane f into ring slot f % RING (carried xm / xc advance)
  auto plane = [&](int f) {
#pragma unroll
    for (int p = 0; p < PTS; ++p) {
      const int id = f * S2 + col[p];
      const float xp = x[id + S2];
      const float s6 = xp + xm[p] + x[id + Sf] + x[id - Sf] + x[id + 1] + x[id - 1];
      float rhs;
      if (FINE) {
        const float u = src[id];
        rhs = (alpha * (u * u)) * beta;
      } else {
        rhs = src[id];
      }
      const float r = rhs - (6.0f * xc[p] - s6);
      if (tid + p * NT < PL) ring[(f % RING) * PL + tid + p * NT] = in[p] ? r : 0.0f;
      xm[p] = xc[p];
      xc[p] = xp;
    }
  };
  plane(f0);
  // thread -> (coarse plane of the round, coarse column)
  const int half = tid / (CK * CJ), tk = tid % CK, tj = (tid / CK) % CJ;
  for (int I = I0; I < I1; I += CP) {
#pragma unroll
    for (int h = 0; h < CP; ++h) {
      if (I + h < I1) {
        plane(2 * (I + h));
        plane(2 * (I + h) + 1);
      }
    }
    lds_barrier();
    const int Ih = I + half;
    if (half < CP && Ih < I1) {
      const int K = K0 + tk, J = J0 + tj;
      if (K <= nc && J <= nc) {
        const int c = (2 * tj + 1) * RK + (2 * tk + 1);
        float out = restrict27_planes(ring, ((2 * Ih - 1) % RING) * PL + c, ((2 * Ih) % RING) * PL + c,
                                      ((2 * Ih + 1) % RING) * PL + c, RK, 1);
        if (sign < 0.0f) out = -out;
        gc[Ih * Sc * Sc + J * Sc + K] = out;
      }
    }
    lds_barrier();
  }
}

// Residual + 4 FW restriction of one CK x CJ x CI tile of coarse interior
// nodes starting at (I0, J0, K0): the residual once at each of the
// (2CK+1)(2CJ+1)(2CI+1) fine nodes the tile's restriction stencils cover
// (residual_at) into the LDS array r, then one coarse value per thread with
// restrict27 -- the values of the residual pass then the restriction pass,
// without the fine residual array.  The caller synchronises before reusing r.
template <bool FINE, int CK, int CJ, int CI>
__device__ __forceinline__ void resres_tile(int Sf, int K0, int J0, int I0, const float* __restrict__ x,
                                            const float* __restrict__ src, float alpha, float beta,
                                            float* __restrict__ gc, float sign, float* r) {
  constexpr int FK = 2 * CK + 1, FJ = 2 * CJ + 1, FI = 2 * CI + 1;
  constexpr int NT = CK * CJ * CI, NF = FK * FJ * FI;
  const int nc = (Sf - 1) / 2 - 1, n = Sf - 2;
  const int Sc = nc + 2, S2 = Sf * Sf;
  const int fk0 = 2 * K0 - 1, fj0 = 2 * J0 - 1, fi0 = 2 * I0 - 1;
  const int tid = threadIdx.x;
  // all residual loads of a thread are independent and (indices clamped to
  // the interior) unconditional, so the unrolled loop has them in flight
  // together; nodes past the interior store 0 (no valid coarse node reads them)
#pragma unroll
  for (int it = 0; it < (NF + NT - 1) / NT; ++it) {
    const int q = tid + it * NT;
    const int qq = q < NF ? q : NF - 1;
    const int li = qq / (FJ * FK), lj = (qq / FK) % FJ, lk = qq % FK;
    const int fi = fi0 + li, fj = fj0 + lj, fk = fk0 + lk;
    const bool in = fi <= n && fj <= n && fk <= n;
    const float v = residual_at<FINE>(x, src, min(fi, n) * S2 + min(fj, n) * Sf + min(fk, n), S2, Sf, alpha, beta);
    if (q < NF) r[q] = in ? v : 0.0f;
  }
  __syncthreads();
  const int tk = tid % CK, tj = (tid / CK) % CJ, ti = tid / (CK * CJ);
  const int K = K0 + tk, J = J0 + tj, I = I0 + ti;
  // no early return: a caller looping over tiles synchronises after this
  if (K <= nc && J <= nc && I <= nc) {
    float out = restrict27(r, (2 * ti + 1) * (FJ * FK) + (2 * tj + 1) * FK + (2 * tk + 1), FJ * FK, FK, 1);
    if (sign < 0.0f) out = -out;
    gc[I * Sc * Sc + J * Sc + K] = out;
  }
}

}  // namespace rq
