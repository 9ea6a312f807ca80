// Fine-level marches of the multigrid V-cycle as per-tile device functions
// for the batched solver (rq_batch.hip): the pre-smoothing pair + residual +
// 4 FW restriction (down) and the correction + post-smoothing pair + ITP
// update (up), each for tile `tile` of one grid's column tiles x plane
// chunks of CI planes (k fastest, then j, then the chunk).  The bodies are
// those of the single-grid kernels k_smooth2_resres_march and k_post_march
// (rq_kernels.hip), which keep their own copy: routed through these
// functions, the single-grid kernels compile to a different instruction
// schedule that measured 1.7 % slower per 513^3 step.  Both copies are held
// to the separate passes bit for bit by the GPU tests.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rq_internal.h"
#include "rq_stencil.h"

namespace rq {

template <int CK, int CJ, int NT>
__device__ __forceinline__ void smooth2_resres_march_tile(int Sf, int CI, int tile, const float* __restrict__ x,
                                                          float* __restrict__ xn,
                                                          const float* __restrict__ src, float a, float c,
                                                          float alpha, float beta, float* __restrict__ gc) {
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
  const int t = tile;
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

template <int TK, int TJ, int RU>
__device__ __forceinline__ void post_march_tile(int Sf, int CI, int tile, const float* __restrict__ x,
                                                const float* __restrict__ e, float* __restrict__ xn,
                                                const float* __restrict__ U, float* __restrict__ Un,
                                                const float* __restrict__ K, const uint16_t* __restrict__ lab,
                                                float a, float c, float alpha, float beta, float half_dv, float dtv) {
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
  const int t = tile;
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

}  // namespace rq
