// Fused per-electron RealQM solver of mol_fast.js on gfx950, behind the C ABI
// of include/rq_molfast.h.
//
// Fields (HBM, FP32): U, W, P as NELEC contiguous S^3 planes-of-planes per
// ping-pong half (electron m at offset m*S^3, the reference's layout), K one
// S^3 field, plus the 2-grid scratch (fine residual, coarse rhs, two coarse
// error buffers) and FP64 reduction partials.
//
// The fused update runs one thread per interior node for ALL electrons: the
// cross-electron terms (max of the other amplitudes, the Pauli density, P
// total and the split-group P) read each electron's u and P once per node
// instead of once per (node, electron) pair as the reference's (WG, NELEC)
// dispatch does.  Kernels are templated on NELEC so the per-electron values
// stay in registers.  Every FP32 expression keeps the WGSL operand order
// (built with -ffp-contract=off); reductions sum the per-node FP32 terms in
// FP64 in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rq_molfast.h"

namespace rqmf {
namespace {

constexpr int kMaxE = RQ_MF_MAX_ELEC;
constexpr int kFinalThreads = 256;
// LDS plane ring of the update: 2 x 3 NELEC planes of 66 x 6 floats (28.5 KB
// at 3 electrons, 57 KB at 6); above 6 electrons the per-node kernel runs
constexpr int kMfLdsMaxE = 6;
// three-plane ring of the update with K: 3 x (3 NELEC + 1) planes of
// kMfRingPlane + 1 floats (46.5 KiB at 3 electrons; the last slot of a plane
// is the write sink of the lanes past the tile); at 4 electrons it would hold
// 2 waves per SIMD (208 VGPRs, 61.8 KB) against the two-plane march's 3
constexpr int kMfRingMaxE = 3;
constexpr int kMfRingPlane = 396;  // (TK + 2)(TJ + 2) bound of a ring tile (66 x 6)
constexpr int kMfRingPlanes = 8;   // target planes per ring workgroup
#ifndef MF_LDS_WAVES
#define MF_LDS_WAVES 3  // waves per SIMD the LDS march is compiled for (4 spills to scratch)
#endif
constexpr int kMfLdsBlocks = 4096;  // target workgroups of the LDS march

// FP32 constants of the WGSL passes (writeParams and the templated arrays,
// mol_fast.js:96-131, 838-876)
struct Orb {
  float rc, target, tilt_x, tilt_y, r_out;
  float ax, ay, az, rot;               // split axis (normalised, toFixed(6))
  float e1x, e1y, e1z, e2x, e2y, e2z;  // in-plane basis of the sector split
  int split, sidx, fix, group;
  // per-electron constants of the update, formed on the host in FP32 with the
  // shader's expressions: the self-interaction fraction f_m and hd = half_d k_mult
  float fm, hd;
  // squared grid distance beyond which sqrt(d2) h >= rc surely (skips the cusp test)
  float rc_skip2;
};

struct Kc {
  int NN, S, S2, S3, N2, shell;
  float h, h2, inv_h, inv_h2, dt, half_d, h3, two_pi, R_out, field_x, alpha_tf, full_self, kc, kc_multi, beta;
  float px[kMaxE], py[kMaxE], pz[kMaxE], Z[kMaxE];
  Orb orb[kMaxE];
};

// double-precision set-up of uploadInitialData (mol_fast.js:941-1046)
struct Kd {
  int S;
  double hv, h2v, soft, mid_i0;
  double px[kMaxE], py[kMaxE], pz[kMaxE], Z[kMaxE], zi[kMaxE], tgw[kMaxE], rc[kMaxE], ro[kMaxE], hs[kMaxE];
  double a0[kMaxE], a1[kMaxE], a2[kMaxE];
  double e1x[kMaxE], e1y[kMaxE], e1z[kMaxE], e2x[kMaxE], e2y[kMaxE], e2z[kMaxE];
  int split[kMaxE], sidx[kMaxE];
};

// ---------------------------------------------------------------------------
// set-up (double)
// ---------------------------------------------------------------------------
__device__ int sector_d(const Kd& d, int a, double dx, double dy, double dz, double dot) {
  const double px = dx - dot * d.a0[a], py = dy - dot * d.a1[a], pz = dz - dot * d.a2[a];
  const double pe1 = px * d.e1x[a] + py * d.e1y[a] + pz * d.e1z[a];
  const double pe2 = px * d.e2x[a] + py * d.e2y[a] + pz * d.e2z[a];
  double theta = atan2(pe2, pe1) - 0.0;
  if (theta < -M_PI) theta += 2 * M_PI;
  if (theta > M_PI) theta -= 2 * M_PI;
  return (int)floor((theta + M_PI) / (2 * M_PI / 3)) % 3;
}

// inSplit (mol_fast.js:878-939)
__device__ bool in_split_d(const Kd& d, int a, double dx, double dy, double dz, double r) {
  const int split = d.split[a], idx = d.sidx[a];
  if (split == RQ_MF_SPLIT_SPHERE || r < 1e-6) return true;
  const double dot = dx * d.a0[a] + dy * d.a1[a] + dz * d.a2[a];
  switch (split) {
    case RQ_MF_SPLIT_HEMI:
      return idx == 0 ? dot < 0 : dot >= 0;
    case RQ_MF_SPLIT_THIRD:
      return sector_d(d, a, dx, dy, dz, dot) == idx;
    case RQ_MF_SPLIT_TETRA: {
      const double u0 = dx / r, u1 = dy / r, u2 = dz / r;
      const double vs[4][3] = {{1, 1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
      int best = 0;
      double bd = -INFINITY;
      for (int v = 0; v < 4; ++v) {
        const double dd = vs[v][0] * u0 + vs[v][1] * u1 + vs[v][2] * u2;
        if (dd > bd) {
          bd = dd;
          best = v;
        }
      }
      return best == idx;
    }
    case RQ_MF_SPLIT_HEMI_THIRD:
      if (idx == 0) return dot >= 0;
      if (dot >= 0) return false;
      return sector_d(d, a, dx, dy, dz, dot) == idx - 1;
    default:
      return true;
  }
}

__device__ double smoothcut_d(double r, double rc, double hv) {
  const double edge = rc - 3 * hv;
  const double t = fmax(0.0, fmin(1.0, (r - edge) / (rc - edge)));
  return t * t * (3 - 2 * t);
}

template <int NE>
__global__ __launch_bounds__(256) void k_mf_init(Kd d, float* __restrict__ K, float* __restrict__ U0,
                                                  float* __restrict__ W0, float* __restrict__ P0) {
  const int S = d.S;
  const size_t S3 = (size_t)S * S * S;
  const size_t id = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (id >= S3) return;
  const int k = (int)(id % S), j = (int)((id / S) % S), i = (int)(id / ((size_t)S * S));
  double K_acc = 0.0, best_u = -1.0;
  int best_m = -1;
  double r_l[NE], u_l[NE];
#pragma unroll
  for (int a = 0; a < NE; ++a) {
    const double dx = (i - d.px[a]) * d.hv, dy = (j - d.py[a]) * d.hv, dz = (k - d.pz[a]) * d.hv;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r_soft = sqrt(r2 + d.soft);
    const double r = sqrt(r2);
    K_acc += d.Z[a] / r_soft;
    const double uh = exp(-d.zi[a] * r);
    r_l[a] = r;
    u_l[a] = uh;
    const double uw = sqrt(d.tgw[a] * d.zi[a] * d.zi[a] * d.zi[a]) * uh;
    const bool er = !(d.ro[a] > 0 && r > d.ro[a] + d.hv);
    const bool ehs = d.hs[a] == 0 || ((d.hs[a] > 0) == (i > d.mid_i0));
    if (er && ehs && in_split_d(d, a, dx, dy, dz, r) && uw > best_u) {
      best_u = uw;
      best_m = a;
    }
  }
  K[id] = (float)K_acc;
  double r_best = 0.0;
#pragma unroll
  for (int m = 0; m < NE; ++m) {
    double u = u_l[m];
    if (d.rc[m] > 0 && r_l[m] < d.rc[m]) u *= smoothcut_d(r_l[m], d.rc[m], d.hv);
    U0[m * S3 + id] = (float)u;
    P0[m * S3 + id] = (float)(0.5 * d.Z[m] / (r_l[m] + d.h2v));
    W0[m * S3 + id] = 0.0f;
    if (m == best_m) r_best = r_l[m];
  }
  if (best_m >= 0 && r_best < 3.0) {
    double w = 1.0;
    double rcb = 0.0, rob = 0.0;
#pragma unroll
    for (int m = 0; m < NE; ++m)
      if (m == best_m) {
        rcb = d.rc[m];
        rob = d.ro[m];
      }
    if (rcb > 0 && r_best < rcb) w = smoothcut_d(r_best, rcb, d.hv);
    if (rob > 0 && r_best > rob + d.hv) w = 0.0;
    W0[best_m * S3 + id] = (float)w;
  }
}

// per-electron sum of U^2 over all nodes (the pre-normalisation), FP64 partials
template <int NE>
__global__ __launch_bounds__(256) void k_mf_sumsq(int S, const float* __restrict__ U, double* __restrict__ part) {
  const size_t S3 = (size_t)S * S * S;
  const size_t per = (S3 + gridDim.x - 1) / gridDim.x;
  const size_t b0 = blockIdx.x * per, b1 = std::min(S3, b0 + per);
  double acc[NE];
#pragma unroll
  for (int m = 0; m < NE; ++m) acc[m] = 0.0;
  for (size_t id = b0 + threadIdx.x; id < b1; id += 256) {
#pragma unroll
    for (int m = 0; m < NE; ++m) {
      const double v = U[m * S3 + id];
      acc[m] += v * v;
    }
  }
  __shared__ double red[256];
#pragma unroll
  for (int m = 0; m < NE; ++m) {
    red[threadIdx.x] = acc[m];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x * NE + m] = red[0];
    __syncthreads();
  }
}

// Ud[i] *= sqrt(target / (norm h^3)) in double, stored as float
// (mol_fast.js:1029-1041); sums[m] = sum of U^2 over all nodes
template <int NE>
__global__ __launch_bounds__(256) void k_mf_prenorm(Kd d, const double* __restrict__ sums,
                                                     const double* __restrict__ tgt, float* __restrict__ U) {
  const int S = d.S;
  const size_t S3 = (size_t)S * S * S;
  const size_t id = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (id >= S3) return;
#pragma unroll
  for (int m = 0; m < NE; ++m) {
    double norm = sums[m];
    norm *= d.hv * d.hv * d.hv;
    if (norm > 0 && tgt[m] > 0) U[m * S3 + id] = (float)((double)U[m * S3 + id] * sqrt(tgt[m] / norm));
  }
}

// ---------------------------------------------------------------------------
// FP32 passes
// ---------------------------------------------------------------------------
__device__ __forceinline__ float smooth_rc(float r, float rc, float h) {
  const float edge = rc - 3.0f * h;
  const float t = fminf(fmaxf((r - edge) / (rc - edge), 0.0f), 1.0f);
  return (t * t) * (3.0f - 2.0f * t);
}

__device__ __forceinline__ unsigned sector_f(const Orb& o, float rx, float ry, float rz, float dotA) {
  const float ppx = rx - dotA * o.ax, ppy = ry - dotA * o.ay, ppz = rz - dotA * o.az;
  const float d2 = (ppx * o.e2x + ppy * o.e2y) + ppz * o.e2z;
  const float d1 = (ppx * o.e1x + ppy * o.e1y) + ppz * o.e1z;
  const float pi = 3.141592653589793f, tpi = 6.283185307179586f;
  float theta = atan2f(d2, d1) - o.rot;
  if (theta < -pi) theta = theta + tpi;
  if (theta > pi) theta = theta - tpi;
  return ((unsigned)floorf((theta + pi) / 2.094395102393195f)) % 3u;
}

// angular split of the update shader (mol_fast.js:194-263)
__device__ bool in_sector_f(const Orb& o, float rx, float ry, float rz, float r_m) {
  const float dotA = (rx * o.ax + ry * o.ay) + rz * o.az;
  switch (o.split) {
    case RQ_MF_SPLIT_HEMI:
      return o.sidx == 0 ? dotA < 0.0f : dotA >= 0.0f;
    case RQ_MF_SPLIT_THIRD:
      return sector_f(o, rx, ry, rz, dotA) == (unsigned)o.sidx;
    case RQ_MF_SPLIT_TETRA: {
      if (!(r_m > 1e-6f)) return true;
      const float ux = rx / r_m, uy = ry / r_m, uz = rz / r_m;
      int best = 0;
      float bd = (ux + uy) + uz;
      const float d1 = (ux + -uy) + -uz;
      if (d1 > bd) {
        bd = d1;
        best = 1;
      }
      const float d2 = (-ux + uy) + -uz;
      if (d2 > bd) {
        bd = d2;
        best = 2;
      }
      const float d3 = (-ux + -uy) + uz;
      if (d3 > bd) {
        bd = d3;
        best = 3;
      }
      return best == o.sidx;
    }
    case RQ_MF_SPLIT_HEMI_THIRD:
      if (o.sidx == 0) return dotA >= 0.0f;
      if (dotA >= 0.0f) return false;
      return sector_f(o, rx, ry, rz, dotA) == (unsigned)(o.sidx - 1);
    default:
      return true;
  }
}

// Per-node quantities shared by every electron's update at that node.
struct NodeIn {
  float fi, fj, fk, x_au, y_au, Kv, P_total, tmin;
};

// frozen-core cusps: w_m = min(w_m, t_a) over every nucleus a with r_a < rc_a;
// the t_a do not depend on m and min is exact, so their minimum is formed once
template <int NE>
__device__ __forceinline__ float cusp_min(const Kc& c, float fi, float fj, float fk) {
  float tmin = INFINITY;
  if (!c.shell) {
#pragma unroll
    for (int a = 0; a < NE; ++a) {
      const float rca = c.orb[a].rc;
      if (rca > 0.0f) {
        const float dx = fi - c.px[a], dy = fj - c.py[a], dz = fk - c.pz[a];
        const float d2 = (dx * dx + dy * dy) + dz * dz;
        if (d2 <= c.orb[a].rc_skip2) {
          const float r = sqrtf(d2) * c.h;
          if (r < rca) tmin = fminf(tmin, smooth_rc(r, rca, c.h));
        }
      }
    }
  }
  return tmin;
}

// Electron m's fused update (mol_fast.js:155-350) at one interior node from
// its 7-point values of w, u and P (centre, +i, -i, +j, -j, +k, -k) and the
// centre values un / pn of every electron.
template <int NE>
__device__ __forceinline__ void mf_electron(const Kc& c, int m, const NodeIn& nd, const float* un, const float* pn,
                                            const float* w7, const float* u7, const float* p7, float& w_out,
                                            float& u_out, float& p_out) {
  const Orb& ob = c.orb[m];
  const float uc = un[m];
  float umax = 0.0f;
#pragma unroll
  for (int e = 0; e < NE; ++e)
    if (e != m && un[e] > umax) umax = un[e];
  const float cm = uc - umax;
  // w: free-boundary advection + diffusion, then cusps / confinement / split
  const float wc = w7[0], wip = w7[1], wim = w7[2], wjp = w7[3], wjm = w7[4], wkp = w7[5], wkm = w7[6];
  const float lw = (((((wip + wim) + wjp) + wjm) + wkp) + wkm - 6.0f * wc) * c.inv_h2;
  const float gx = (wip - wim) * c.inv_h, gy = (wjp - wjm) * c.inv_h, gz = (wkp - wkm) * c.inv_h;
  float nw = (wc + ((0.05f * c.dt) * fabsf(cm)) * lw) + ((1.0f * c.dt) * cm) * sqrtf((gx * gx + gy * gy) + gz * gz);
  nw = fminf(fmaxf(nw, 0.0f), 1.0f);
  if (c.shell) {
    const float rx = nd.fi - c.px[m], ry = nd.fj - c.py[m], rz = nd.fk - c.pz[m];
    const float r_m = sqrtf((rx * rx + ry * ry) + rz * rz) * c.h;
    const float ro_m = ob.r_out * c.R_out;
    if (ob.rc > 0.0f && r_m < ob.rc) nw = fminf(nw, smooth_rc(r_m, ob.rc, c.h));
    if (ro_m > 0.0f && r_m > ro_m) nw = (r_m <= ro_m + c.h) ? nw : 0.0f;
    if (ob.split != 0 && ob.fix == 1 && !in_sector_f(ob, rx, ry, rz, r_m)) nw = 0.0f;
  } else {
    nw = fminf(nw, nd.tmin);
  }
  w_out = nw;
  // u: w-weighted Laplacian + potential terms
  float P_group = 0.0f;
#pragma unroll
  for (int e = 0; e < NE; ++e)
    if (c.orb[e].group == ob.group) P_group = P_group + pn[e];
  const float selfN = ob.target;
  const float Vother = 2.0f * (nd.P_total - ob.fm * P_group);
  const float rho_m = uc * uc;
  const float V_TF =
      (selfN > 1.0f && c.alpha_tf > 0.0f) ? c.alpha_tf * powf(fmaxf(rho_m, 1e-20f), 0.6666667f) : 0.0f;
  const float V_tilt = (-ob.tilt_x) * nd.x_au - ob.tilt_y * nd.y_au;
  float rho_other = 0.0f;
#pragma unroll
  for (int e = 0; e < NE; ++e)
    if (e != m) rho_other = rho_other + un[e] * un[e];
  const float V_pauli = c.beta * rho_other;
  const float uip = u7[1], uim = u7[2], ujp = u7[3], ujm = u7[4], ukp = u7[5], ukm = u7[6];
  const float hd = ob.hd;
  const float A = hd * (((uip - uc) * (wip + nw)) * 0.5f - ((uc - uim) * (nw + wim)) * 0.5f);
  const float B = hd * (((ujp - uc) * (wjp + nw)) * 0.5f - ((uc - ujm) * (nw + wjm)) * 0.5f);
  const float C = hd * (((ukp - uc) * (wkp + nw)) * 0.5f - ((uc - ukm) * (nw + wkm)) * 0.5f);
  const float D = ((c.dt * (((((nd.Kv - Vother) - V_TF) - V_tilt) - V_pauli) - c.field_x * nd.x_au)) * uc) * wc;
  u_out = (((uc + A) + B) + C) + D;
  // P_m: pseudo-time Jacobi step sourced by electron m's own density
  const float Pc = pn[m];
  const float s6 = ((((p7[1] + p7[2]) + p7[3]) + p7[4]) + p7[5]) + p7[6];
  p_out = (Pc + (c.dt * (s6 - 6.0f * Pc)) * c.inv_h2) + ((c.two_pi * c.dt) * uc) * uc;
}

// The fused update at one interior node per thread, every electron; the
// 7-point values come straight from global memory (used above kMfLdsMaxE
// electrons, where the LDS ring below would not fit).
template <int NE>
__global__ __launch_bounds__(256) void k_mf_update(Kc c, const float* __restrict__ K, const float* __restrict__ Ui,
                                                    const float* __restrict__ Wi, const float* __restrict__ Pi,
                                                    float* __restrict__ Uo, float* __restrict__ Wo,
                                                    float* __restrict__ Po) {
  const int n = c.NN - 1;
  const int k = 1 + blockIdx.x * 64 + threadIdx.x, j = 1 + blockIdx.y * 4 + threadIdx.y, i = 1 + blockIdx.z;
  if (k > n || j > n) return;
  const int S = c.S, S2 = c.S2;
  const size_t S3 = (size_t)c.S3;
  const int id = i * S2 + j * S + k;  // per-lane offsets 32-bit, electron bases uniform
  float un[NE], pn[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    un[e] = (Ui + e * S3)[id];
    pn[e] = (Pi + e * S3)[id];
  }
  NodeIn nd;
  nd.fi = (float)i;
  nd.fj = (float)j;
  nd.fk = (float)k;
  nd.x_au = (nd.fi - (float)c.N2) * c.h;
  nd.y_au = (nd.fj - (float)c.N2) * c.h;
  nd.Kv = K[id];
  nd.P_total = 0.0f;
#pragma unroll
  for (int e = 0; e < NE; ++e) nd.P_total = nd.P_total + pn[e];
  nd.tmin = cusp_min<NE>(c, nd.fi, nd.fj, nd.fk);
#pragma unroll
  for (int m = 0; m < NE; ++m) {
    const float* __restrict__ Um = Ui + m * S3;
    const float* __restrict__ Wm = Wi + m * S3;
    const float* __restrict__ Pm = Pi + m * S3;
    const float w7[7] = {Wm[id], Wm[id + S2], Wm[id - S2], Wm[id + S], Wm[id - S], Wm[id + 1], Wm[id - 1]};
    const float u7[7] = {un[m], Um[id + S2], Um[id - S2], Um[id + S], Um[id - S], Um[id + 1], Um[id - 1]};
    const float p7[7] = {pn[m], Pm[id + S2], Pm[id - S2], Pm[id + S], Pm[id - S], Pm[id + 1], Pm[id - 1]};
    float wn, unw, pnw;
    mf_electron<NE>(c, m, nd, un, pn, w7, u7, p7, wn, unw, pnw);
    (Wo + m * S3)[id] = wn;
    (Uo + m * S3)[id] = unw;
    (Po + m * S3)[id] = pnw;
  }
}

// Linear block id remapped so that each of the 8 XCDs (workgroups are dealt
// round-robin over them) sweeps one contiguous range of tiles, whose halo
// rows then stay in that XCD's L2.
__device__ __forceinline__ int xcd_tile() {
  const int gx = gridDim.x, gy = gridDim.y;
  const int nblk = gx * gy * gridDim.z;
  const int b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const int per = nblk >> 3, rem = nblk & 7;
  const int xcd = b & 7, local = b >> 3;
  return (xcd < rem) ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// The fused update as a march along i with an LDS plane ring: a workgroup
// owns 64 (k) x 4 (j) columns and ci planes.  Per plane every field of every
// electron (u, w, P) is staged once into LDS with a one-node halo, the i-1 /
// i+1 values stay in the owning thread's registers, and each node's 7-point
// values are read from LDS and registers: about 3 NELEC + 1 global loads per
// node instead of 21 NELEC + 1.  Double-buffered LDS, one barrier per plane.
template <int NE>
__global__ __launch_bounds__(256, MF_LDS_WAVES) void k_mf_update_lds(Kc c, int ci, const float* __restrict__ K,
                                                        const float* __restrict__ Ui, const float* __restrict__ Wi,
                                                        const float* __restrict__ Pi, float* __restrict__ Uo,
                                                        float* __restrict__ Wo, float* __restrict__ Po) {
  constexpr int TK = 64, TJ = 4, RK = TK + 2, RJ = TJ + 2, PL = RK * RJ, NF = 3 * NE, NH = 2 * RK + 2 * TJ;
  __shared__ float lds[2][NF][PL];
  const int n = c.NN - 1, S = c.S, S2 = c.S2;
  const size_t S3 = (size_t)c.S3;
  const int t = xcd_tile();
  const int gx = gridDim.x, gy = gridDim.y;
  const int K0 = 1 + (t % gx) * TK, J0 = 1 + ((t / gx) % gy) * TJ, I0 = 1 + (t / (gx * gy)) * ci;
  const int I1 = min(I0 + ci, n + 1);
  const int tk = threadIdx.x, tj = threadIdx.y, tid = tj * TK + tk;
  const int k = K0 + tk, j = J0 + tj;
  const bool active = k <= n && j <= n;
  const int col = min(j, S - 1) * S + min(k, S - 1);
  const int oslot = (tj + 1) * RK + (tk + 1);
  // this thread's halo node of the (RJ x RK) region, if any
  const bool has_halo = tid < NH;
  int hcol = 0, hslot = 0;
  if (has_halo) {
    int hj, hk;
    if (tid < RK) {
      hj = 0;
      hk = tid;
    } else if (tid < 2 * RK) {
      hj = RJ - 1;
      hk = tid - RK;
    } else {
      const int q = tid - 2 * RK;
      hj = 1 + (q >> 1);
      hk = (q & 1) ? RK - 1 : 0;
    }
    hcol = min(J0 - 1 + hj, S - 1) * S + min(K0 - 1 + hk, S - 1);
    hslot = hj * RK + hk;
  }
  // field f: u of electron f (f < NE), w of f - NE, P of f - 2 NE
  auto fld = [&](int f) -> const float* {
    return f < NE ? Ui + f * S3 : (f < 2 * NE ? Wi + (f - NE) * S3 : Pi + (f - 2 * NE) * S3);
  };
  float prv[NF], cur[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    prv[f] = fld(f)[(I0 - 1) * S2 + col];
    cur[f] = fld(f)[I0 * S2 + col];
  }
  const float fj = (float)j, fk = (float)k;
  int b = 0;
  for (int i = I0; i < I1; ++i) {
    const int id = i * S2 + col;
    float nxt[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) nxt[f] = fld(f)[id + S2];
    float halo[NF];
    if (has_halo) {
#pragma unroll
      for (int f = 0; f < NF; ++f) halo[f] = fld(f)[i * S2 + hcol];
    }
    const float Kv = K[id];
#pragma unroll
    for (int f = 0; f < NF; ++f) lds[b][f][oslot] = cur[f];
    if (has_halo) {
#pragma unroll
      for (int f = 0; f < NF; ++f) lds[b][f][hslot] = halo[f];
    }
    __syncthreads();
    if (active) {
      float un[NE], pn[NE];
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        un[e] = cur[e];
        pn[e] = cur[2 * NE + e];
      }
      NodeIn nd;
      nd.fi = (float)i;
      nd.fj = fj;
      nd.fk = fk;
      nd.x_au = (nd.fi - (float)c.N2) * c.h;
      nd.y_au = (fj - (float)c.N2) * c.h;
      nd.Kv = Kv;
      nd.P_total = 0.0f;
#pragma unroll
      for (int e = 0; e < NE; ++e) nd.P_total = nd.P_total + pn[e];
      nd.tmin = cusp_min<NE>(c, nd.fi, fj, fk);
#pragma unroll
      for (int m = 0; m < NE; ++m) {
        const int fu = m, fw = NE + m, fp = 2 * NE + m;
        const float w7[7] = {cur[fw], nxt[fw], prv[fw], lds[b][fw][oslot + RK], lds[b][fw][oslot - RK],
                             lds[b][fw][oslot + 1], lds[b][fw][oslot - 1]};
        const float u7[7] = {cur[fu], nxt[fu], prv[fu], lds[b][fu][oslot + RK], lds[b][fu][oslot - RK],
                             lds[b][fu][oslot + 1], lds[b][fu][oslot - 1]};
        const float p7[7] = {cur[fp], nxt[fp], prv[fp], lds[b][fp][oslot + RK], lds[b][fp][oslot - RK],
                             lds[b][fp][oslot + 1], lds[b][fp][oslot - 1]};
        float wn, unw, pnw;
        mf_electron<NE>(c, m, nd, un, pn, w7, u7, p7, wn, unw, pnw);
        (Wo + m * S3)[id] = wn;
        (Uo + m * S3)[id] = unw;
        (Po + m * S3)[id] = pnw;
      }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      prv[f] = cur[f];
      cur[f] = nxt[f];
    }
    b ^= 1;
  }
}

// Store sink of the lanes of k_mf_update_ring that lie past the last
// interior node (tile overhang): their stores go here, so every store
// instruction is unconditional.
__device__ float g_mf_sink[256];

// The fused update as an i-march with a three-plane LDS ring (NELEC <= 3).
// While plane i is computed the ring holds planes i - 1 (own columns), i
// (with its one-node halo) and i + 1 of every field of every electron and K,
// and the registers hold plane i + 2, loaded during the previous plane.
// After plane i's stores, plane i + 2 goes into the slot of plane i - 1 and
// the loads of plane i + 3 are issued, so a load has a whole plane's compute
// to arrive and no plane waits on a load issued in its own iteration.  Every
// register set is written to LDS before it is reloaded (no register rotation
// that would force a wait), and the overhang lanes store to a sink instead of
// branching around their stores: every path reaching the LDS write then has
// the same stores after the loads it needs, so the wait for those loads does
// not wait for the stores.  Every node reads its 7-point values from LDS.
//
// Grid (k-tiles, j-tiles, gz): each tile's planes are split into gz ranges
// of balanced length (plane counts differ by at most one), so the host can
// pick gz to make the workgroup count a whole number of resident rounds.
// Workgroups go to the XCDs in slab order (all tiles of one plane range
// before the next), so neighbouring tiles march the same planes at the same
// time and their halo reads meet in L2.
template <int NE>
__global__ __launch_bounds__(256, MF_LDS_WAVES) void k_mf_update_ring(Kc c, int TK, int TJ,
                                                         const float* __restrict__ K, const float* __restrict__ Ui,
                                                         const float* __restrict__ Wi, const float* __restrict__ Pi,
                                                         float* __restrict__ Uo, float* __restrict__ Wo,
                                                         float* __restrict__ Po) {
  // a (TK x TJ)-node tile per 256-thread workgroup, TK * TJ <= 256 and
  // (TK + 2)(TJ + 2) <= PL; the LDS field planes keep the fixed stride PL
  constexpr int PL = kMfRingPlane + 1, NF = 3 * NE, NL = NF + 1;
  __shared__ float ring[3][NL][PL];
  const int RK = TK + 2, RJ = TJ + 2, NH = 2 * RK + 2 * TJ;
  const int n = c.NN - 1, S = c.S, S2 = c.S2;
  const size_t S3 = (size_t)c.S3;
  const int gxt = gridDim.x, gyt = gridDim.y, gz = gridDim.z;
  const int t = xcd_tile();
  const int tile = t % (gxt * gyt), z = t / (gxt * gyt);
  const int I0 = 1 + (z * n) / gz, I1 = 1 + ((z + 1) * n) / gz;
  const int tid = threadIdx.x;
  // threads past TK * TJ repeat a node of the last row: they read its LDS
  // slots, but write their own-column values to the sink slot kMfRingPlane
  // (never read) and their results to the global sink.  Writing the owner's
  // slot instead would race: a spare in another wave could store plane i + 2
  // into the slot of plane i - 1 while the owner still reads plane i - 1.
  const int tk = tid % TK, tj = min(tid / TK, TJ - 1);
  const bool spare = tid >= TK * TJ;
  const int oslot = (tj + 1) * RK + (tk + 1);
  const int wslot = spare ? kMfRingPlane : oslot;
  // this thread's halo node of the (RJ x RK) region, if any
  const bool has_halo = tid < NH;
  int hj = 0, hk = 0;
  if (has_halo) {
    if (tid < RK) {
      hk = tid;
    } else if (tid < 2 * RK) {
      hj = RJ - 1;
      hk = tid - RK;
    } else {
      const int q = tid - 2 * RK;
      hj = 1 + (q >> 1);
      hk = (q & 1) ? RK - 1 : 0;
    }
  }
  const int hslot = hj * RK + hk;
  // field f: u of electron f (f < NE), w of f - NE, P of f - 2 NE, K (f = NF)
  auto fld = [&](int f) -> const float* {
    return f < NE ? Ui + f * S3 : (f < 2 * NE ? Wi + (f - NE) * S3 : (f < NF ? Pi + (f - 2 * NE) * S3 : K));
  };
  {
    const int K0 = 1 + (tile % gxt) * TK, J0 = 1 + (tile / gxt) * TJ;
    const int k = K0 + tk, j = J0 + tj;
    const bool active = k <= n && j <= n && !spare;
    const int col = min(j, S - 1) * S + min(k, S - 1);
    // threads without a halo node load their own column again (a cache
    // hit) so the halo load is unconditional
    const int hcol = has_halo ? min(J0 - 1 + hj, S - 1) * S + min(K0 - 1 + hk, S - 1) : col;
    // prologue: planes I0 - 1 (own columns, slot 2), I0 (slot 0) and I0 + 1
    // (slot 1) with their halos, loads issued together; plane I0 + 2 into
    // registers.  Planes past I1 (the i + 1 neighbour of the range's last
    // plane) are clamped to I1 <= S - 1: their values are never read, and
    // the repeated loads of plane I1 hit in cache instead of fetching planes
    // no plane of the range needs.
    {
      const int i1 = min(I0 + 1, S - 1);
      float A[NL], B[NL], HB[NF], C[NL], HC[NF];
#pragma unroll
      for (int f = 0; f < NL; ++f) {
        A[f] = fld(f)[(I0 - 1) * S2 + col];
        B[f] = fld(f)[I0 * S2 + col];
        C[f] = fld(f)[i1 * S2 + col];
      }
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        HB[f] = fld(f)[I0 * S2 + hcol];
        HC[f] = fld(f)[i1 * S2 + hcol];
      }
#pragma unroll
      for (int f = 0; f < NL; ++f) {
        ring[2][f][wslot] = A[f];
        ring[0][f][wslot] = B[f];
        ring[1][f][wslot] = C[f];
      }
      if (has_halo) {
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          ring[0][f][hslot] = HB[f];
          ring[1][f][hslot] = HC[f];
        }
      }
    }
    float R[NL], H[NF];
    {
      const int i2 = min(I0 + 2, I1);
#pragma unroll
      for (int f = 0; f < NL; ++f) R[f] = fld(f)[i2 * S2 + col];
#pragma unroll
      for (int f = 0; f < NF; ++f) H[f] = fld(f)[i2 * S2 + hcol];
    }
    const float fj = (float)j, fk = (float)k;
    int q = 0;
    for (int i = I0; i < I1; ++i) {
      const int qn = q == 2 ? 0 : q + 1, qp = q == 0 ? 2 : q - 1;
      __syncthreads();
      const int id = i * S2 + col;
      float un[NE], pn[NE];
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        un[e] = ring[q][e][oslot];
        pn[e] = ring[q][2 * NE + e][oslot];
      }
      NodeIn nd;
      nd.fi = (float)i;
      nd.fj = fj;
      nd.fk = fk;
      nd.x_au = (nd.fi - (float)c.N2) * c.h;
      nd.y_au = (fj - (float)c.N2) * c.h;
      nd.Kv = ring[q][NF][oslot];
      nd.P_total = 0.0f;
#pragma unroll
      for (int e = 0; e < NE; ++e) nd.P_total = nd.P_total + pn[e];
      nd.tmin = cusp_min<NE>(c, nd.fi, fj, fk);
      float wo[NE], uo[NE], po[NE];
#pragma unroll
      for (int m = 0; m < NE; ++m) {
        const int fu = m, fw = NE + m, fp = 2 * NE + m;
        const float w7[7] = {ring[q][fw][oslot],      ring[qn][fw][oslot],     ring[qp][fw][oslot],
                             ring[q][fw][oslot + RK], ring[q][fw][oslot - RK], ring[q][fw][oslot + 1],
                             ring[q][fw][oslot - 1]};
        const float u7[7] = {un[m],                   ring[qn][fu][oslot],     ring[qp][fu][oslot],
                             ring[q][fu][oslot + RK], ring[q][fu][oslot - RK], ring[q][fu][oslot + 1],
                             ring[q][fu][oslot - 1]};
        const float p7[7] = {pn[m],                   ring[qn][fp][oslot],     ring[qp][fp][oslot],
                             ring[q][fp][oslot + RK], ring[q][fp][oslot - RK], ring[q][fp][oslot + 1],
                             ring[q][fp][oslot - 1]};
        mf_electron<NE>(c, m, nd, un, pn, w7, u7, p7, wo[m], uo[m], po[m]);
      }
      // all stores after the last branch of the compute
#pragma unroll
      for (int m = 0; m < NE; ++m) {
        *(active ? Wo + m * S3 + id : g_mf_sink + tid) = wo[m];
        *(active ? Uo + m * S3 + id : g_mf_sink + tid) = uo[m];
        *(active ? Po + m * S3 + id : g_mf_sink + tid) = po[m];
      }
      // plane i + 2 into the slot of plane i - 1: its halo and neighbour
      // reads ended before this plane's barrier, its own-column reads are
      // this thread's and done; then the loads of plane i + 3 (clamped)
#pragma unroll
      for (int f = 0; f < NL; ++f) ring[qp][f][wslot] = R[f];
      if (has_halo) {
#pragma unroll
        for (int f = 0; f < NF; ++f) ring[qp][f][hslot] = H[f];
      }
      const int i3 = min(i + 3, I1);
#pragma unroll
      for (int f = 0; f < NL; ++f) R[f] = fld(f)[i3 * S2 + col];
#pragma unroll
      for (int f = 0; f < NF; ++f) H[f] = fld(f)[i3 * S2 + hcol];
      q = qn;
    }
  }
}

// reduce (mol_fast.js:363-436): per-node FP32 terms, FP64 per-block partials
template <int NE>
__global__ __launch_bounds__(256) void k_mf_reduce(Kc c, int ich, const float* __restrict__ K,
                                                    const float* __restrict__ U, const float* __restrict__ W,
                                                    const float* __restrict__ P, double* __restrict__ part) {
  constexpr int NR = NE + 6;
  const int n = c.NN - 1;
  const int k = 1 + blockIdx.x * 64 + threadIdx.x, j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int i0 = 1 + blockIdx.z * ich, i1 = min(i0 + ich, n + 1);
  const int S = c.S, S2 = c.S2;
  const size_t S3 = (size_t)c.S3;
  double acc[NR];
#pragma unroll
  for (int x = 0; x < NR; ++x) acc[x] = 0.0;
  if (k <= n && j <= n) {
    const float yc = ((float)j - (float)c.N2) * c.h, zc = ((float)k - (float)c.N2) * c.h;
    for (int i = i0; i < i1; ++i) {
      const size_t id = (size_t)i * S2 + (size_t)j * S + k;
      const float xc = ((float)i - (float)c.N2) * c.h;
      const float Kv = K[id];
      float P_tot = 0.0f;
#pragma unroll
      for (int e = 0; e < NE; ++e) P_tot = P_tot + P[e * S3 + id];
      float T = 0.0f, VeK = 0.0f, Vee = 0.0f, dex = 0.0f, dey = 0.0f, dez = 0.0f;
#pragma unroll
      for (int m = 0; m < NE; ++m) {
        const size_t o = m * S3 + id;
        const Orb& ob = c.orb[m];
        const float v = U[o];
        acc[m] += (double)((v * v) * c.h3);
        const float selfN = ob.target;
        if (W[o] > 0.1f) {
          const float a = U[o + S2] - v, b = U[o + S] - v, cc = U[o + 1] - v;
          const float km = (selfN > 1.0f) ? c.kc_multi : 1.0f;
          T = T + ((km * c.kc) * ((a * a + b * b) + cc * cc)) * c.h;
        }
        VeK = VeK - ((Kv * v) * v) * c.h3;
        const float rho_cell = (v * v) * c.h3;
        dex = dex - rho_cell * xc;
        dey = dey - rho_cell * yc;
        dez = dez - rho_cell * zc;
        const float fm = (selfN > 1.0f) ? 1.0f / selfN : 1.0f;
        Vee = Vee + (((P_tot - fm * P[o]) * v) * v) * c.h3;
        if (selfN > 1.0f && c.alpha_tf > 0.0f)
          T = T + ((0.6f * c.alpha_tf) * powf(fmaxf(v * v, 1e-20f), 1.6666667f)) * c.h3;
        if (ob.tilt_y != 0.0f) T = T + ((((-ob.tilt_y) * yc) * v) * v) * c.h3;
      }
      acc[NE] += (double)T;
      acc[NE + 1] += (double)VeK;
      acc[NE + 2] += (double)Vee;
      acc[NE + 3] += (double)dex;
      acc[NE + 4] += (double)dey;
      acc[NE + 5] += (double)dez;
    }
  }
  __shared__ double red[4][NR];
  const int tid = threadIdx.y * 64 + threadIdx.x, lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int x = 0; x < NR; ++x) {
    double v = acc[x];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wv][x] = v;
  }
  __syncthreads();
  if (tid < NR) {
    const size_t b = blockIdx.x + (size_t)gridDim.x * (blockIdx.y + (size_t)gridDim.y * blockIdx.z);
    part[b * NR + tid] = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
  }
}

// fixed-order sum of the partials into sums[NR]
__global__ __launch_bounds__(kFinalThreads) void k_mf_finalize(int nparts, int NR, const double* __restrict__ part,
                                                                double* __restrict__ sums) {
  __shared__ double red[kFinalThreads];
  for (int x = 0; x < NR; ++x) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nparts; b += kFinalThreads) v += part[(size_t)b * NR + x];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kFinalThreads / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) sums[x] = red[0];
    __syncthreads();
  }
}

// normalize (mol_fast.js:468-489) with the FP32 norm of the last reduction
template <int NE>
__global__ __launch_bounds__(256) void k_mf_normalize(Kc c, const double* __restrict__ sums, float* __restrict__ U) {
  const int n = c.NN - 1;
  const int k = 1 + blockIdx.x * 64 + threadIdx.x, j = 1 + blockIdx.y * 4 + threadIdx.y, i = 1 + blockIdx.z;
  if (k > n || j > n) return;
  const size_t id = (size_t)i * c.S2 + (size_t)j * c.S + k;
#pragma unroll
  for (int m = 0; m < NE; ++m) {
    const float nm = (float)sums[m], tgt = c.orb[m].target;
    if (nm > 0.0f && tgt > 0.0f) {
      const size_t o = m * (size_t)c.S3 + id;
      U[o] = U[o] * sqrtf(tgt / nm);
    }
  }
}

// ---- per-electron 2-grid correction (mol_fast.js:645-789) ----
__global__ __launch_bounds__(256) void k_mf_residual(Kc c, const float* __restrict__ u, const float* __restrict__ p,
                                                     float* __restrict__ res) {
  const int n = c.NN - 1;
  const int k = 1 + blockIdx.x * 64 + threadIdx.x, j = 1 + blockIdx.y * 4 + threadIdx.y, i = 1 + blockIdx.z;
  if (k > n || j > n) return;
  const int S = c.S, S2 = c.S2;
  const size_t id = (size_t)i * S2 + (size_t)j * S + k;
  const float uc = u[id];
  const float rho = uc * uc;
  const float Pc = p[id];
  const float lap = (((((((p[id + S2] + p[id - S2]) + p[id + S]) + p[id - S]) + p[id + 1]) + p[id - 1]) - 6.0f * Pc)) *
                    c.inv_h2;
  res[id] = (-c.two_pi) * rho - lap;
}

__global__ __launch_bounds__(256) void k_mf_restrict(Kc c, const float* __restrict__ f, float* __restrict__ co) {
  const int NC = c.NN / 2, SC = NC + 1, ncm = NC - 1;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= ncm * ncm * ncm) return;
  const int K2 = t % ncm + 1, J = (t / ncm) % ncm + 1, I = t / (ncm * ncm) + 1;
  const int S = c.S, S2 = c.S2;
  const size_t b = (size_t)(2 * I) * S2 + (size_t)(2 * J) * S + 2 * K2;
  auto at = [&](int di, int dj, int dk) { return f[b + (long)di * S2 + (long)dj * S + dk]; };
  float val = 8.0f * at(0, 0, 0);
  val = val + 4.0f * (((((at(1, 0, 0) + at(-1, 0, 0)) + at(0, 1, 0)) + at(0, -1, 0)) + at(0, 0, 1)) + at(0, 0, -1));
  val = val + 2.0f * (((((((((((at(1, 1, 0) + at(1, -1, 0)) + at(-1, 1, 0)) + at(-1, -1, 0)) + at(1, 0, 1)) +
                              at(1, 0, -1)) +
                             at(-1, 0, 1)) +
                            at(-1, 0, -1)) +
                           at(0, 1, 1)) +
                          at(0, 1, -1)) +
                         at(0, -1, 1)) +
                        at(0, -1, -1));
  val = val + (((((((at(1, 1, 1) + at(1, 1, -1)) + at(1, -1, 1)) + at(1, -1, -1)) + at(-1, 1, 1)) + at(-1, 1, -1)) +
                at(-1, -1, 1)) +
               at(-1, -1, -1));
  co[(size_t)I * SC * SC + (size_t)J * SC + K2] = val * 0.015625f;
}

__global__ __launch_bounds__(256) void k_mf_coarse_smooth(Kc c, const float* __restrict__ Ein, float* __restrict__ Eout,
                                                          const float* __restrict__ rhs) {
  const int NC = c.NN / 2, SC = NC + 1, SC2 = SC * SC, ncm = NC - 1;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= ncm * ncm * ncm) return;
  const int K2 = t % ncm + 1, J = (t / ncm) % ncm + 1, I = t / (ncm * ncm) + 1;
  const int cid = I * SC2 + J * SC + K2;
  const float ec = Ein[cid];
  const float sn = ((((Ein[cid + SC2] + Ein[cid - SC2]) + Ein[cid + SC]) + Ein[cid - SC]) + Ein[cid + 1]) + Ein[cid - 1];
  const float hc2 = 4.0f * c.h2;
  Eout[cid] = 0.3333f * ec + (sn + hc2 * rhs[cid]) / 9.0f;
}

__global__ __launch_bounds__(256) void k_mf_prolong(Kc c, const float* __restrict__ Ec, float* __restrict__ p) {
  const int n = c.NN - 1, NC = c.NN / 2, SC = NC + 1, SC2 = SC * SC;
  const int k = 1 + blockIdx.x * 64 + threadIdx.x, j = 1 + blockIdx.y * 4 + threadIdx.y, i = 1 + blockIdx.z;
  if (k > n || j > n) return;
  const int ci = i / 2, cj = j / 2, ck = k / 2;
  const int ci1 = min(ci + (i & 1), NC), cj1 = min(cj + (j & 1), NC), ck1 = min(ck + (k & 1), NC);
  const float wi = (i & 1) ? 0.5f : 1.0f, wj = (j & 1) ? 0.5f : 1.0f, wk = (k & 1) ? 0.5f : 1.0f;
  const float wi1 = 1.0f - wi, wj1 = 1.0f - wj, wk1 = 1.0f - wk;
  float corr = 0.0f;
  corr = corr + ((wi * wj) * wk) * Ec[ci * SC2 + cj * SC + ck];
  corr = corr + ((wi * wj) * wk1) * Ec[ci * SC2 + cj * SC + ck1];
  corr = corr + ((wi * wj1) * wk) * Ec[ci * SC2 + cj1 * SC + ck];
  corr = corr + ((wi * wj1) * wk1) * Ec[ci * SC2 + cj1 * SC + ck1];
  corr = corr + ((wi1 * wj) * wk) * Ec[ci1 * SC2 + cj * SC + ck];
  corr = corr + ((wi1 * wj) * wk1) * Ec[ci1 * SC2 + cj * SC + ck1];
  corr = corr + ((wi1 * wj1) * wk) * Ec[ci1 * SC2 + cj1 * SC + ck];
  corr = corr + ((wi1 * wj1) * wk1) * Ec[ci1 * SC2 + cj1 * SC + ck1];
  const size_t fid = (size_t)i * c.S2 + (size_t)j * c.S + k;
  p[fid] = p[fid] + 0.5f * corr;
}

__global__ void k_mf_fill(float* __restrict__ x, size_t n) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) x[t] = 0.0f;
}

// electronic force per atom (mol_fast.js:503-560): sequential in the shader's order
__global__ void k_mf_forces(Kc c, int NE, const float* __restrict__ P, float* __restrict__ out) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= NE) return;
  const float Za = c.Z[a];
  if (Za <= 0.0f) {
    out[3 * a] = out[3 * a + 1] = out[3 * a + 2] = 0.0f;
    return;
  }
  const int ci = (int)c.px[a], cj = (int)c.py[a], ck = (int)c.pz[a];
  const int R = 4, S = c.S, S2 = c.S2;
  const size_t S3 = (size_t)c.S3;
  const float R2f = (float)(R * R), inv2h = 0.5f * c.inv_h;
  float sFi = 0.0f, sFj = 0.0f, sFk = 0.0f, count = 0.0f;
  for (int di = -R; di <= R; ++di) {
    const int ii = ci + di;
    if (ii < 1 || ii >= S - 1) continue;
    for (int dj = -R; dj <= R; ++dj) {
      const int jj = cj + dj;
      if (jj < 1 || jj >= S - 1) continue;
      for (int dk = -R; dk <= R; ++dk) {
        const int kk = ck + dk;
        if (kk < 1 || kk >= S - 1) continue;
        if ((float)(di * di + dj * dj + dk * dk) > R2f) continue;
        const size_t id = (size_t)ii * S2 + (size_t)jj * S + kk;
        float xp = 0.0f, xm = 0.0f, yp = 0.0f, ym = 0.0f, zp = 0.0f, zm = 0.0f;
        for (int m = 0; m < NE; ++m) {
          const size_t o = m * S3 + id;
          xp = xp + P[o + S2];
          xm = xm + P[o - S2];
          yp = yp + P[o + S];
          ym = ym + P[o - S];
          zp = zp + P[o + 1];
          zm = zm + P[o - 1];
        }
        sFi = sFi + (xp - xm) * inv2h;
        sFj = sFj + (yp - ym) * inv2h;
        sFk = sFk + (zp - zm) * inv2h;
        count = count + 1.0f;
      }
    }
  }
  const float avg = (count > 0.0f) ? 1.0f / count : 0.0f;
  out[3 * a] = ((2.0f * Za) * sFi) * avg;
  out[3 * a + 1] = ((2.0f * Za) * sFj) * avg;
  out[3 * a + 2] = ((2.0f * Za) * sFk) * avg;
}

// recomputeK (mol_fast.js:569-585), every node
__global__ __launch_bounds__(256) void k_mf_recompute_k(Kc c, int NE, float* __restrict__ K) {
  const size_t S3 = (size_t)c.S3;
  const size_t id = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (id >= S3) return;
  const int S = c.S;
  const int k = (int)(id % S), j = (int)((id / S) % S), i = (int)(id / ((size_t)S * S));
  const float soft = 4.0f * c.h2;
  float Kv = 0.0f;
  for (int a = 0; a < NE; ++a) {
    const float Za = c.Z[a];
    if (Za <= 0.0f) continue;
    const float di = ((float)i - c.px[a]) * c.h, dj = ((float)j - c.py[a]) * c.h, dk = ((float)k - c.pz[a]) * c.h;
    Kv = Kv + Za / sqrtf(((di * di + dj * dj) + dk * dk) + soft);
  }
  K[id] = Kv;
}

// extract (mol_fast.js:599-640)
__global__ void k_mf_extract(Kc c, int NE, const float* __restrict__ U, const float* __restrict__ W,
                             const float* __restrict__ P, float* __restrict__ out) {
  const int i = blockIdx.x * 16 + threadIdx.x, j = blockIdx.y * 16 + threadIdx.y;
  const int SS = c.NN + 1, N2 = c.N2;
  if (i > c.NN || j > c.NN) return;
  const size_t S3 = (size_t)c.S3, A = (size_t)SS * SS;
  for (int m = 0; m < NE; ++m) {
    const size_t idx2 = m * S3 + (size_t)i * c.S2 + (size_t)j * c.S + N2;
    out[m * A + (size_t)i * SS + j] = (W[idx2] > 0.1f) ? U[idx2] : 0.0f;
    if (j == 0) {
      const size_t line = (size_t)i * c.S2 + (size_t)N2 * c.S + N2, ml = m * S3 + line;
      out[NE * A + (size_t)m * SS + i] = U[ml];
      out[2 * NE * A + (size_t)m * SS + i] = W[ml];
      float ps = 0.0f;
      for (int e = 0; e < NE; ++e) ps = ps + P[e * S3 + line];
      if (m == 0) out[NE * A + (size_t)NE * SS + i] = 2.0f * ps;
      float vp = 0.0f;
      if (c.orb[m].target > 1.0f && c.alpha_tf > 0.0f) {
        const float rho = U[ml] * U[ml];
        vp = c.alpha_tf * powf(fmaxf(rho, 1e-20f), 0.6666667f);
      }
      out[NE * A + (size_t)(NE + 1 + m) * SS + i] = vp;
    }
  }
}

// ---- template dispatch on NELEC ----
#define MF_DISPATCH(NE_, CALL)          \
  switch (NE_) {                        \
    case 1: { constexpr int E = 1; CALL; } break;   \
    case 2: { constexpr int E = 2; CALL; } break;   \
    case 3: { constexpr int E = 3; CALL; } break;   \
    case 4: { constexpr int E = 4; CALL; } break;   \
    case 5: { constexpr int E = 5; CALL; } break;   \
    case 6: { constexpr int E = 6; CALL; } break;   \
    case 7: { constexpr int E = 7; CALL; } break;   \
    case 8: { constexpr int E = 8; CALL; } break;   \
    case 9: { constexpr int E = 9; CALL; } break;   \
    case 10: { constexpr int E = 10; CALL; } break; \
    case 11: { constexpr int E = 11; CALL; } break; \
    case 12: { constexpr int E = 12; CALL; } break; \
    case 13: { constexpr int E = 13; CALL; } break; \
    case 14: { constexpr int E = 14; CALL; } break; \
    case 15: { constexpr int E = 15; CALL; } break; \
    default: { constexpr int E = 16; CALL; } break; \
  }

#define MF_DISPATCH_RING(NE_, CALL)                \
  switch (NE_) {                                  \
    case 1: { constexpr int E = 1; CALL; } break; \
    case 2: { constexpr int E = 2; CALL; } break; \
    default: { constexpr int E = 3; CALL; } break; \
  }

#define MF_DISPATCH_LDS(NE_, CALL)                 \
  switch (NE_) {                                  \
    case 1: { constexpr int E = 1; CALL; } break; \
    case 2: { constexpr int E = 2; CALL; } break; \
    case 3: { constexpr int E = 3; CALL; } break; \
    case 4: { constexpr int E = 4; CALL; } break; \
    case 5: { constexpr int E = 5; CALL; } break; \
    default: { constexpr int E = 6; CALL; } break; \
  }

float fixed6(double x) {
  char b[64];
  snprintf(b, sizeof b, "%.6f", x);
  return (float)strtod(b, nullptr);
}

}  // namespace
}  // namespace rqmf

using namespace rqmf;

struct rq_mf {
  rq_mf_params p{};
  std::vector<rq_mf_nucleus> nuc;
  int NE = 0, NN = 0, S = 0, NC = 0, SC = 0;
  size_t S3 = 0, SC3 = 0;
  bool r_out_enforce = true;
  bool initialized = false;
  Kc kc{};
  Kd kd{};
  hipStream_t st = nullptr;
  float* K = nullptr;
  float* U[2] = {nullptr, nullptr};
  float* W[2] = {nullptr, nullptr};
  float* P[2] = {nullptr, nullptr};
  float* res = nullptr;
  float* rhsC = nullptr;
  float* errC[2] = {nullptr, nullptr};
  float* slice = nullptr;
  float* force_d = nullptr;
  double* part = nullptr;
  double* sums = nullptr;
  double* tgt_d = nullptr;
  int cur = 0;
  int red_ich = 8;
  int nparts = 0;
  // update kernel: 2 = three-plane LDS ring (NELEC <= kMfRingMaxE, else the
  // two-plane march), 1 = two-plane LDS march, 0 = per-node (RQ_MF_LDS)
  int lds_march = 2;
  int lds_ci = 1;
  dim3 lds_grid;
  int ring_tk = 64, ring_tj = 4;  // tile and grid of the three-plane ring
  dim3 ring_grid;
  bool timing = false;
  std::vector<hipEvent_t> ev;
  int ev_used = 0;
  std::string err;
};

namespace {

std::string g_err;

int fail(rq_mf* h, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (h)
    h->err = buf;
  else
    g_err = buf;
  return code;
}

#define MFCHK(h, call)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) return fail((h), RQ_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define MFLAUNCH(h)                                                                              \
  do {                                                                                           \
    hipError_t e_ = hipGetLastError();                                                           \
    if (e_ != hipSuccess) return fail((h), RQ_EHIP, "kernel launch: %s", hipGetErrorString(e_)); \
  } while (0)

dim3 node_grid(int NN) {
  const int n = NN - 1;
  return dim3((n + 63) / 64, (n + 3) / 4, n);
}

// FP32 constants from the parameters, positions and the rOutEnforce flag
void build_kc(rq_mf* h) {
  const rq_mf_params& p = h->p;
  Kc& c = h->kc;
  const double hv = p.screen / p.NN, h2v = hv * hv, h3v = hv * hv * hv, dv = 0.12;
  c.NN = p.NN;
  c.S = p.NN + 1;
  c.S2 = c.S * c.S;
  c.S3 = c.S2 * c.S;
  c.N2 = (int)std::lround(p.NN / 2.0);
  c.shell = p.shell_mode ? 1 : 0;
  c.h = (float)hv;
  c.h2 = (float)h2v;
  c.inv_h = (float)(1.0 / hv);
  c.inv_h2 = (float)(1.0 / h2v);
  c.dt = (float)(dv * h2v);
  c.half_d = (float)(p.kinetic_coeff * dv);
  c.h3 = (float)h3v;
  c.two_pi = (float)(2.0 * M_PI);
  c.R_out = h->r_out_enforce ? 1.0f : 0.0f;
  c.field_x = (float)p.field_x;
  c.alpha_tf = (float)p.alpha_tf;
  c.full_self = p.full_self ? 1.0f : 0.0f;
  c.kc = fixed6(p.kinetic_coeff);
  c.kc_multi = fixed6(p.kc_multi);
  c.beta = fixed6(p.beta_pauli);
  for (int a = 0; a < h->NE; ++a) {
    const rq_mf_nucleus& n = h->nuc[a];
    c.px[a] = (float)n.i;
    c.py[a] = (float)n.j;
    c.pz[a] = (float)n.k;
    c.Z[a] = (float)n.Z;
    Orb& o = c.orb[a];
    o.rc = (float)n.rc;
    o.target = (float)n.target;
    o.tilt_x = (float)n.tilt_x;
    o.tilt_y = (float)n.tilt_y;
    o.r_out = fixed6(n.r_out);
    double L = std::hypot(n.axis[0], n.axis[1], n.axis[2]);
    if (!(L != 0.0)) L = 1.0;
    o.ax = fixed6(n.axis[0] / L);
    o.ay = fixed6(n.axis[1] / L);
    o.az = fixed6(n.axis[2] / L);
    o.rot = 0.0f;
    // e1 = normalize(...), e2 = cross(ax, e1) in FP32 (mol_fast.js:204-208)
    float vx, vy, vz;
    if (std::fabs(o.az) < 0.9f) {
      vx = o.ay;
      vy = -o.ax;
      vz = 0.0f;
    } else {
      vx = 0.0f;
      vy = o.az;
      vz = -o.ay;
    }
    const float ln = std::sqrt((vx * vx + vy * vy) + vz * vz);
    o.e1x = vx / ln;
    o.e1y = vy / ln;
    o.e1z = vz / ln;
    o.e2x = o.ay * o.e1z - o.az * o.e1y;
    o.e2y = o.az * o.e1x - o.ax * o.e1z;
    o.e2z = o.ax * o.e1y - o.ay * o.e1x;
    o.split = n.split;
    o.sidx = n.split_idx;
    o.fix = n.split_fix ? 1 : 0;
    int g = a;
    if (n.split != RQ_MF_SPLIT_SPHERE) {
      for (int b = 0; b < a; ++b) {
        const rq_mf_nucleus& q = h->nuc[b];
        if (q.split == n.split && q.i == n.i && q.j == n.j && q.k == n.k) {
          g = b;
          break;
        }
      }
    }
    o.group = g;
    const float selfN = o.target;
    const float fm_multi = (c.full_self > 0.5f) ? 0.0f : 1.0f / selfN;
    o.fm = (selfN > 1.0f) ? fm_multi : 1.0f;
    o.hd = c.half_d * ((selfN > 1.0f) ? c.kc_multi : 1.0f);
    const double rg = (double)o.rc / (double)c.h * (1.0 + 1e-4);
    o.rc_skip2 = (float)(rg * rg);
  }
}

void build_kd(rq_mf* h) {
  Kd& d = h->kd;
  const double hv = h->p.screen / h->p.NN;
  d.S = h->S;
  d.hv = hv;
  d.h2v = hv * hv;
  d.soft = 4 * d.h2v;
  d.mid_i0 = (h->nuc[0].i + h->nuc[h->NE - 1].i) / 2;
  for (int a = 0; a < h->NE; ++a) {
    const rq_mf_nucleus& n = h->nuc[a];
    d.px[a] = n.i;
    d.py[a] = n.j;
    d.pz[a] = n.k;
    d.Z[a] = n.Z;
    d.zi[a] = std::isnan(n.z_init) ? (n.Z > 0 ? n.Z : 1.0) : n.z_init;
    d.tgw[a] = (n.target != 0.0 && !std::isnan(n.target)) ? n.target : 1.0;
    d.rc[a] = n.rc;
    d.ro[a] = n.r_out;
    d.hs[a] = n.hs;
    double L = std::hypot(n.axis[0], n.axis[1], n.axis[2]);
    if (!(L != 0.0)) L = 1.0;
    const double a0 = n.axis[0] / L, a1 = n.axis[1] / L, a2 = n.axis[2] / L;
    d.a0[a] = a0;
    d.a1[a] = a1;
    d.a2[a] = a2;
    double e1x, e1y, e1z;
    if (std::fabs(a2) < 0.9) {
      e1x = a1;
      e1y = -a0;
      e1z = 0;
    } else {
      e1x = 0;
      e1y = a2;
      e1z = -a1;
    }
    double e1l = std::hypot(e1x, e1y, e1z);
    if (!(e1l != 0.0)) e1l = 1.0;
    e1x /= e1l;
    e1y /= e1l;
    e1z /= e1l;
    d.e1x[a] = e1x;
    d.e1y[a] = e1y;
    d.e1z[a] = e1z;
    d.e2x[a] = a1 * e1z - a2 * e1y;
    d.e2y[a] = a2 * e1x - a0 * e1z;
    d.e2z[a] = a0 * e1y - a1 * e1x;
    d.split[a] = n.split;
    d.sidx[a] = n.split_idx;
  }
}

int launch_reduce(rq_mf* h, int half) {
  const int n = h->NN - 1;
  const dim3 grid((n + 63) / 64, (n + 3) / 4, (n + h->red_ich - 1) / h->red_ich);
  const Kc& c = h->kc;
  const int ich = h->red_ich;
  MF_DISPATCH(h->NE, hipLaunchKernelGGL((k_mf_reduce<E>), grid, dim3(64, 4), 0, h->st, c, ich, h->K, h->U[half],
                                        h->W[half], h->P[half], h->part));
  MFLAUNCH(h);
  hipLaunchKernelGGL(k_mf_finalize, dim3(1), dim3(kFinalThreads), 0, h->st, h->nparts, h->NE + 6, h->part, h->sums);
  MFLAUNCH(h);
  return RQ_OK;
}

int twogrid(rq_mf* h, int half) {
  const Kc& c = h->kc;
  const int ncm = h->NC - 1;
  const int cblocks = (ncm * ncm * ncm + 255) / 256;
  for (int m = 0; m < h->NE; ++m) {
    hipLaunchKernelGGL(k_mf_residual, node_grid(h->NN), dim3(64, 4), 0, h->st, c, h->U[half] + m * h->S3,
                       h->P[half] + m * h->S3, h->res);
    hipLaunchKernelGGL(k_mf_restrict, dim3(cblocks), dim3(256), 0, h->st, c, h->res, h->rhsC);
    for (int b = 0; b < 2; ++b)
      hipLaunchKernelGGL(k_mf_fill, dim3((unsigned)((h->SC3 + 255) / 256)), dim3(256), 0, h->st, h->errC[b], h->SC3);
    for (int cs = 0; cs < h->p.coarse_iters; ++cs) {
      const int src = cs % 2;
      hipLaunchKernelGGL(k_mf_coarse_smooth, dim3(cblocks), dim3(256), 0, h->st, c, h->errC[src], h->errC[1 - src],
                         h->rhsC);
    }
    // the reference binds error buffer 0 to the prolongation whatever the iteration count
    hipLaunchKernelGGL(k_mf_prolong, node_grid(h->NN), dim3(64, 4), 0, h->st, c, h->errC[0], h->P[half] + m * h->S3);
    MFLAUNCH(h);
  }
  return RQ_OK;
}

void free_all(rq_mf* h) {
  float* fl[] = {h->K, h->U[0], h->U[1], h->W[0], h->W[1], h->P[0], h->P[1], h->res, h->rhsC, h->errC[0], h->errC[1],
                 h->slice, h->force_d};
  for (float* f : fl)
    if (f) (void)hipFree(f);
  if (h->part) (void)hipFree(h->part);
  if (h->sums) (void)hipFree(h->sums);
  if (h->tgt_d) (void)hipFree(h->tgt_d);
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  if (h->st) (void)hipStreamDestroy(h->st);
}

}  // namespace

extern "C" {

void rq_mf_default_params(rq_mf_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->NN = 200;
  p->screen = 15.0;
  p->kinetic_coeff = 0.5;
  p->kc_multi = 1.0;
  p->vcycle = 1;
  p->vcycle_interval = 50;
  p->coarse_iters = 10;
  p->norm_interval = 20;
}

const char* rq_mf_last_error(const rq_mf* h) { return h ? h->err.c_str() : g_err.c_str(); }

void rq_mf_ring_tile(int32_t NN, int32_t* tk, int32_t* tj, int32_t* sink_slot) {
  // the fewest k-tiles of at most 64 nodes, split evenly, so little of a
  // wave falls past the last interior node (199 nodes: 4 x 50 instead of
  // 4 x 64), then as many j-rows as 256 threads and the LDS plane allow
  const int nn = std::max(1, NN - 1);
  const int ntk = (nn + 63) / 64;
  const int tkr = (nn + ntk - 1) / ntk;
  const int tjr = std::max(1, std::min({256 / tkr, kMfRingPlane / (tkr + 2) - 2, 16}));
  if (tk) *tk = tkr;
  if (tj) *tj = tjr;
  if (sink_slot) *sink_slot = kMfRingPlane;
}

int rq_mf_create(const rq_mf_params* p, const rq_mf_nucleus* nuclei, int32_t n, rq_mf** out) {
  if (!p || !nuclei || !out) return fail(nullptr, RQ_EINVAL, "null argument");
  *out = nullptr;
  if (n < 1 || n > RQ_MF_MAX_ELEC) return fail(nullptr, RQ_ERANGE, "mol_fast: 1..%d nuclei, got %d", RQ_MF_MAX_ELEC, n);
  if (p->NN < 4 || p->NN % 2) return fail(nullptr, RQ_EINVAL, "mol_fast: NN must be even (and >= 4), got %d", p->NN);
  if (!(p->screen > 0)) return fail(nullptr, RQ_EINVAL, "screen must be > 0");
  if (p->vcycle_interval < 1 || p->coarse_iters < 0 || p->norm_interval < 1)
    return fail(nullptr, RQ_EINVAL, "bad vcycle_interval / coarse_iters / norm_interval");
  for (int a = 0; a < n; ++a) {
    const rq_mf_nucleus& q = nuclei[a];
    if (!std::isfinite(q.i) || !std::isfinite(q.j) || !std::isfinite(q.k) || !std::isfinite(q.Z))
      return fail(nullptr, RQ_EINVAL, "nucleus %d: non-finite position or charge", a);
    const int s = q.split;
    if (!(s == RQ_MF_SPLIT_SPHERE || s == RQ_MF_SPLIT_HEMI || s == RQ_MF_SPLIT_THIRD || s == RQ_MF_SPLIT_TETRA ||
          s == RQ_MF_SPLIT_HEMI_THIRD))
      return fail(nullptr, RQ_EINVAL, "nucleus %d: unknown split code %d", a, s);
  }
  const size_t S = (size_t)p->NN + 1, S3 = S * S * S;
  const size_t bytes = (size_t)n * S3 * 4 * 6 + S3 * 8;
  if (bytes > ((size_t)200 << 30)) return fail(nullptr, RQ_ERANGE, "mol_fast: %zu bytes of fields exceed 200 GiB", bytes);
  rq_mf* h = new rq_mf();
  h->p = *p;
  h->nuc.assign(nuclei, nuclei + n);
  h->NE = n;
  h->NN = p->NN;
  h->S = (int)S;
  h->S3 = S3;
  h->NC = p->NN / 2;
  h->SC = h->NC + 1;
  h->SC3 = (size_t)h->SC * h->SC * h->SC;
  build_kc(h);
  build_kd(h);
  auto bad = [&](hipError_t e, const char* what) {
    const int rc = fail(nullptr, RQ_ENOMEM, "%s: %s", what, hipGetErrorString(e));
    free_all(h);
    delete h;
    return rc;
  };
  hipError_t e = hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking);
  if (e != hipSuccess) return bad(e, "hipStreamCreate");
  const size_t fb = (size_t)n * S3 * sizeof(float);
  float** big[] = {&h->U[0], &h->U[1], &h->W[0], &h->W[1], &h->P[0], &h->P[1]};
  for (float** f : big) {
    if ((e = hipMalloc(f, fb)) != hipSuccess) return bad(e, "hipMalloc field");
    if ((e = hipMemsetAsync(*f, 0, fb, h->st)) != hipSuccess) return bad(e, "hipMemset");
  }
  if ((e = hipMalloc(&h->K, S3 * 4)) != hipSuccess) return bad(e, "hipMalloc K");
  if ((e = hipMalloc(&h->res, S3 * 4)) != hipSuccess) return bad(e, "hipMalloc residual");
  if ((e = hipMemsetAsync(h->K, 0, S3 * 4, h->st)) != hipSuccess) return bad(e, "hipMemset");
  if ((e = hipMemsetAsync(h->res, 0, S3 * 4, h->st)) != hipSuccess) return bad(e, "hipMemset");
  for (float** f : {&h->rhsC, &h->errC[0], &h->errC[1]}) {
    if ((e = hipMalloc(f, h->SC3 * 4)) != hipSuccess) return bad(e, "hipMalloc coarse");
    if ((e = hipMemsetAsync(*f, 0, h->SC3 * 4, h->st)) != hipSuccess) return bad(e, "hipMemset");
  }
  const size_t slice_n = (size_t)3 * n * S * S;
  if ((e = hipMalloc(&h->slice, slice_n * 4)) != hipSuccess) return bad(e, "hipMalloc slice");
  if ((e = hipMemsetAsync(h->slice, 0, slice_n * 4, h->st)) != hipSuccess) return bad(e, "hipMemset");
  if ((e = hipMalloc(&h->force_d, (size_t)3 * n * 4)) != hipSuccess) return bad(e, "hipMalloc forces");
  const int nn = h->NN - 1;
  h->nparts = ((nn + 63) / 64) * ((nn + 3) / 4) * ((nn + h->red_ich - 1) / h->red_ich);
  const size_t pn = std::max<size_t>((size_t)h->nparts * (n + 6), 1024 * (size_t)n);
  if ((e = hipMalloc(&h->part, pn * 8)) != hipSuccess) return bad(e, "hipMalloc partials");
  if ((e = hipMalloc(&h->sums, (size_t)(n + 6) * 8)) != hipSuccess) return bad(e, "hipMalloc sums");
  if ((e = hipMemsetAsync(h->sums, 0, (size_t)(n + 6) * 8, h->st)) != hipSuccess) return bad(e, "hipMemset");
  if ((e = hipMalloc(&h->tgt_d, (size_t)n * 8)) != hipSuccess) return bad(e, "hipMalloc targets");
  std::vector<double> tg(n);
  for (int a = 0; a < n; ++a) tg[a] = nuclei[a].target;
  if ((e = hipMemcpyAsync(h->tgt_d, tg.data(), n * 8, hipMemcpyHostToDevice, h->st)) != hipSuccess)
    return bad(e, "hipMemcpy");
  if ((e = hipStreamSynchronize(h->st)) != hipSuccess) return bad(e, "hipStreamSynchronize");
  {
    const int gx = (nn + 63) / 64, gy = (nn + 3) / 4;
    int blocks = kMfLdsBlocks;  // RQ_MF_BLOCKS: target workgroups (launch-shape sweeps)
    if (const char* bv = std::getenv("RQ_MF_BLOCKS")) blocks = std::max(1, std::atoi(bv));
    const int gz = std::max(1, std::min(nn, blocks / (gx * gy)));
    h->lds_ci = (nn + gz - 1) / gz;
    h->lds_grid = dim3(gx, gy, (nn + h->lds_ci - 1) / h->lds_ci);
    // ring tiles (rq_mf_ring_tile)
    int tkr = 0, tjr = 0;
    rq_mf_ring_tile(h->NN, &tkr, &tjr, nullptr);
    h->ring_tk = tkr;
    h->ring_tj = tjr;
    // plane ranges per tile: about kMfRingPlanes planes each, rounded so
    // the workgroups fill whole resident rounds (workgroups per CU the
    // kernel holds x CUs); RQ_MF_RING_GZ sets the count directly
    const int tiles = ((nn + tkr - 1) / tkr) * ((nn + tjr - 1) / tjr);
    int per_cu = 0, cus = 0, dev = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return bad(e, "hipGetDevice");
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
      return bad(e, "hipDeviceGetAttribute");
    MF_DISPATCH_RING(h->NE, e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mf_update_ring<E>, 256, 0));
    if (e != hipSuccess) return bad(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    const long slots = (long)std::max(1, per_cu) * std::max(1, cus);
    const long rounds = std::max(1L, std::lround((double)tiles * nn / kMfRingPlanes / slots));
    int rgz = (int)std::max(1L, std::lround((double)rounds * slots / tiles));
    if (const char* gv = std::getenv("RQ_MF_RING_GZ")) rgz = std::atoi(gv);
    rgz = std::max(1, std::min(nn, rgz));
    h->ring_grid = dim3((nn + tkr - 1) / tkr, (nn + tjr - 1) / tjr, rgz);
    if (const char* lv = std::getenv("RQ_MF_LDS")) h->lds_march = std::min(2, std::max(0, std::atoi(lv)));
  }
  *out = h;
  return RQ_OK;
}

int rq_mf_destroy(rq_mf* h) {
  if (!h) return RQ_OK;
  if (h->st) (void)hipStreamSynchronize(h->st);
  free_all(h);
  delete h;
  return RQ_OK;
}

int rq_mf_sync(rq_mf* h) {
  if (!h) return fail(nullptr, RQ_EINVAL, "null handle");
  MFCHK(h, hipStreamSynchronize(h->st));
  return RQ_OK;
}

int rq_mf_init(rq_mf* h) {
  if (!h) return fail(nullptr, RQ_EINVAL, "null handle");
  const unsigned blocks = (unsigned)((h->S3 + 255) / 256);
  const Kd& d = h->kd;
  MF_DISPATCH(h->NE, hipLaunchKernelGGL((k_mf_init<E>), dim3(blocks), dim3(256), 0, h->st, d, h->K, h->U[0], h->W[0],
                                        h->P[0]));
  MFLAUNCH(h);
  const int nparts = 1024;
  MF_DISPATCH(h->NE, hipLaunchKernelGGL((k_mf_sumsq<E>), dim3(nparts), dim3(256), 0, h->st, h->S, h->U[0], h->part));
  hipLaunchKernelGGL(k_mf_finalize, dim3(1), dim3(kFinalThreads), 0, h->st, nparts, h->NE, h->part, h->sums);
  MF_DISPATCH(h->NE, hipLaunchKernelGGL((k_mf_prenorm<E>), dim3(blocks), dim3(256), 0, h->st, d, h->sums, h->tgt_d,
                                        h->U[0]));
  MFLAUNCH(h);
  const size_t fb = (size_t)h->NE * h->S3 * sizeof(float);
  MFCHK(h, hipMemcpyAsync(h->U[1], h->U[0], fb, hipMemcpyDeviceToDevice, h->st));
  MFCHK(h, hipMemcpyAsync(h->W[1], h->W[0], fb, hipMemcpyDeviceToDevice, h->st));
  MFCHK(h, hipMemcpyAsync(h->P[1], h->P[0], fb, hipMemcpyDeviceToDevice, h->st));
  MFCHK(h, hipStreamSynchronize(h->st));
  h->cur = 0;
  h->initialized = true;
  return RQ_OK;
}

int rq_mf_set_timing(rq_mf* h, int32_t on) {
  if (!h) return fail(nullptr, RQ_EINVAL, "null handle");
  h->timing = on != 0;
  return RQ_OK;
}

int rq_mf_do_steps(rq_mf* h, int32_t n, double* sums) {
  if (!h) return fail(nullptr, RQ_EINVAL, "null handle");
  if (!h->initialized) return fail(h, RQ_ESTATE, "rq_mf_do_steps before rq_mf_init");
  if (n < 0) return fail(h, RQ_EINVAL, "negative step count");
  if (h->timing && (int)h->ev.size() < 2 * n) {
    const size_t old = h->ev.size();
    h->ev.resize(2 * n);
    for (size_t q = old; q < h->ev.size(); ++q) MFCHK(h, hipEventCreateWithFlags(&h->ev[q], hipEventDisableSystemFence));
  }
  h->ev_used = 0;
  const dim3 grid = node_grid(h->NN), blk(64, 4);
  const Kc& c = h->kc;
  for (int s = 0; s < n; ++s) {
    const int cur = h->cur, nxt = 1 - cur;
    if (h->timing) MFCHK(h, hipEventRecord(h->ev[2 * s], h->st));
    if (h->NE <= kMfRingMaxE && h->lds_march == 2) {
      MF_DISPATCH_RING(h->NE, hipLaunchKernelGGL((k_mf_update_ring<E>), h->ring_grid, dim3(256), 0, h->st,
                                                 c, h->ring_tk, h->ring_tj,
                                                 h->K, h->U[cur], h->W[cur], h->P[cur], h->U[nxt], h->W[nxt],
                                                 h->P[nxt]));
    } else if (h->NE <= kMfLdsMaxE && h->lds_march) {
      const int ci = h->lds_ci;
      MF_DISPATCH_LDS(h->NE, hipLaunchKernelGGL((k_mf_update_lds<E>), h->lds_grid, dim3(64, 4), 0, h->st, c, ci,
                                                h->K, h->U[cur], h->W[cur], h->P[cur], h->U[nxt], h->W[nxt],
                                                h->P[nxt]));
    } else {
      MF_DISPATCH(h->NE, hipLaunchKernelGGL((k_mf_update<E>), grid, blk, 0, h->st, c, h->K, h->U[cur], h->W[cur],
                                            h->P[cur], h->U[nxt], h->W[nxt], h->P[nxt]));
    }
    MFLAUNCH(h);
    if (h->timing) {
      MFCHK(h, hipEventRecord(h->ev[2 * s + 1], h->st));
      h->ev_used = s + 1;
    }
    if ((s + 1) % h->p.norm_interval == 0 || s == n - 1) {
      if (int rc = launch_reduce(h, nxt)) return rc;
      MF_DISPATCH(h->NE, hipLaunchKernelGGL((k_mf_normalize<E>), grid, blk, 0, h->st, c, h->sums, h->U[nxt]));
      MFLAUNCH(h);
    }
    if (h->p.vcycle && (s + 1) % h->p.vcycle_interval == 0)
      if (int rc = twogrid(h, nxt)) return rc;
    h->cur = nxt;
  }
  if (sums) {
    MFCHK(h, hipMemcpyAsync(sums, h->sums, (size_t)(h->NE + 6) * 8, hipMemcpyDeviceToHost, h->st));
    MFCHK(h, hipStreamSynchronize(h->st));
    for (int x = 0; x < h->NE + 6; ++x)
      if (!std::isfinite(sums[x])) return fail(h, RQ_EUNSTABLE, "numerical instability: non-finite reduction %d", x);
  }
  return RQ_OK;
}

int rq_mf_update_time(rq_mf* h, double* ms, int32_t* launches) {
  if (!h || !ms) return fail(h, RQ_EINVAL, "null argument");
  MFCHK(h, hipStreamSynchronize(h->st));
  double t = 0.0;
  for (int s = 0; s < h->ev_used; ++s) {
    float e = 0.0f;
    MFCHK(h, hipEventElapsedTime(&e, h->ev[2 * s], h->ev[2 * s + 1]));
    t += e;
  }
  *ms = t;
  if (launches) *launches = h->ev_used;
  return RQ_OK;
}

int rq_mf_set_r_out(rq_mf* h, int32_t enforce) {
  if (!h) return fail(nullptr, RQ_EINVAL, "null handle");
  h->r_out_enforce = enforce != 0;
  build_kc(h);
  return RQ_OK;
}

int rq_mf_forces(rq_mf* h, float* out3n) {
  if (!h || !out3n) return fail(h, RQ_EINVAL, "null argument");
  if (!h->initialized) return fail(h, RQ_ESTATE, "rq_mf_forces before rq_mf_init");
  hipLaunchKernelGGL(k_mf_forces, dim3(1), dim3(64), 0, h->st, h->kc, h->NE, h->P[h->cur], h->force_d);
  MFLAUNCH(h);
  MFCHK(h, hipMemcpyAsync(out3n, h->force_d, (size_t)3 * h->NE * 4, hipMemcpyDeviceToHost, h->st));
  MFCHK(h, hipStreamSynchronize(h->st));
  return RQ_OK;
}

int rq_mf_set_positions(rq_mf* h, const double* ijk3n) {
  if (!h || !ijk3n) return fail(h, RQ_EINVAL, "null argument");
  for (int a = 0; a < h->NE; ++a) {
    if (!std::isfinite(ijk3n[3 * a]) || !std::isfinite(ijk3n[3 * a + 1]) || !std::isfinite(ijk3n[3 * a + 2]))
      return fail(h, RQ_EINVAL, "nucleus %d: non-finite position", a);
    h->nuc[a].i = ijk3n[3 * a];
    h->nuc[a].j = ijk3n[3 * a + 1];
    h->nuc[a].k = ijk3n[3 * a + 2];
  }
  build_kc(h);
  return RQ_OK;
}

int rq_mf_recompute_k(rq_mf* h) {
  if (!h) return fail(nullptr, RQ_EINVAL, "null handle");
  hipLaunchKernelGGL(k_mf_recompute_k, dim3((unsigned)((h->S3 + 255) / 256)), dim3(256), 0, h->st, h->kc, h->NE,
                     h->K);
  MFLAUNCH(h);
  MFCHK(h, hipStreamSynchronize(h->st));
  return RQ_OK;
}

int rq_mf_extract(rq_mf* h, float* out, size_t n_floats) {
  if (!h || !out) return fail(h, RQ_EINVAL, "null argument");
  const size_t need = (size_t)3 * h->NE * h->S * h->S;
  if (n_floats < need) return fail(h, RQ_ERANGE, "extract needs %zu floats, got %zu", need, n_floats);
  const dim3 g((h->S + 15) / 16, (h->S + 15) / 16);
  hipLaunchKernelGGL(k_mf_extract, g, dim3(16, 16), 0, h->st, h->kc, h->NE, h->U[h->cur], h->W[h->cur], h->P[h->cur],
                     h->slice);
  MFLAUNCH(h);
  MFCHK(h, hipMemcpyAsync(out, h->slice, need * 4, hipMemcpyDeviceToHost, h->st));
  MFCHK(h, hipStreamSynchronize(h->st));
  return RQ_OK;
}

static float* mf_field(rq_mf* h, int32_t field, int32_t m, int half) {
  switch (field) {
    case RQ_MF_FIELD_U: return h->U[half] + (size_t)m * h->S3;
    case RQ_MF_FIELD_W: return h->W[half] + (size_t)m * h->S3;
    case RQ_MF_FIELD_P: return h->P[half] + (size_t)m * h->S3;
    case RQ_MF_FIELD_K: return h->K;
    default: return nullptr;
  }
}

int rq_mf_read_field(rq_mf* h, int32_t field, int32_t m, float* host, size_t n_floats) {
  if (!h || !host) return fail(h, RQ_EINVAL, "null argument");
  if (m < 0 || m >= h->NE) return fail(h, RQ_EINVAL, "electron %d out of range", m);
  float* f = mf_field(h, field, m, h->cur);
  if (!f) return fail(h, RQ_EINVAL, "unknown field %d", field);
  if (n_floats < h->S3) return fail(h, RQ_ERANGE, "field needs %zu floats", h->S3);
  MFCHK(h, hipMemcpyAsync(host, f, h->S3 * 4, hipMemcpyDeviceToHost, h->st));
  MFCHK(h, hipStreamSynchronize(h->st));
  return RQ_OK;
}

int rq_mf_write_field(rq_mf* h, int32_t field, int32_t m, const float* host, size_t n_floats) {
  if (!h || !host) return fail(h, RQ_EINVAL, "null argument");
  if (m < 0 || m >= h->NE) return fail(h, RQ_EINVAL, "electron %d out of range", m);
  if (n_floats < h->S3) return fail(h, RQ_ERANGE, "field needs %zu floats", h->S3);
  // both halves take the data, so the faces agree (kernels never write faces)
  for (int half = 0; half < 2; ++half) {
    float* f = mf_field(h, field, m, half);
    if (!f) return fail(h, RQ_EINVAL, "unknown field %d", field);
    MFCHK(h, hipMemcpyAsync(f, host, h->S3 * 4, hipMemcpyHostToDevice, h->st));
    if (field == RQ_MF_FIELD_K) break;
  }
  MFCHK(h, hipStreamSynchronize(h->st));
  h->initialized = true;
  return RQ_OK;
}

}  // extern "C"
