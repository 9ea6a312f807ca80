// C ABI of the RealQM grid solver (declared in include/rq.h).
//
// One handle = one molecular grid + one HIP stream.  The handle owns all
// device memory; host buffers are caller-owned.  Pass order of the run() loop
// body follows cloud/solver.py:561-586.
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

using namespace rq;

namespace {
thread_local std::string g_create_error;

struct EventPair {
  int fam;
  hipEvent_t a, b;
};

// Pointer state of the ping-pong buffers; a captured step chunk is valid for
// the state it was captured in and leaves the solver in `after`.
struct PtrState {
  float *U, *U2, *P, *P2;
  uint16_t* lab;
  std::vector<float*> e0, e1;
  bool operator==(const PtrState& o) const {
    return U == o.U && U2 == o.U2 && P == o.P && P2 == o.P2 && lab == o.lab && e0 == o.e0 && e1 == o.e1;
  }
};

struct GraphEntry {
  PtrState before, after;
  int pi, ni;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

// Device buffers of the reverse sweep (allocated on first use, grown as needed).
struct AdjState {
  float *lamU = nullptr, *lamU2 = nullptr, *lamP = nullptr, *lamP2 = nullptr, *gK = nullptr, *acc = nullptr;
  float *t1 = nullptr, *t2 = nullptr, *t3 = nullptr;
  float *wU = nullptr, *wP = nullptr;  // saved start of RQ_ADJ_START_SAVED maps
  std::vector<float*> ckU, ckP;    // checkpoints (U after the step, P in use)
  std::vector<float*> segU, segP;  // one segment of recomputed inputs
  Fx* red = nullptr;  // fixed-point sums: 2 n_dom (norms, dots) + 3 n_atoms (dE/dR) + 3 n_dom + 3 (energy)
};
}  // namespace

struct rq_solver {
  rq_params p{};
  int NN = 0, S = 0;
  size_t S3 = 0;
  double h64 = 0, dv = 0;
  float h = 0, h2 = 0, h3 = 0, half_dv = 0, dtv = 0, two_pi = 0;
  int n_atoms = 0;
  std::vector<Atom> atoms_h;
  // molecule.js per-atom fields (rq_set_atoms_s2) as given, and resolved
  // against the current Z on the device; any_split: some atom has a sector
  std::vector<rq_atom_s2> ext_h;
  AtomS2* ext_d = nullptr;
  bool any_split = false;
  // uniform cell list of the atoms (Z > 0) for the score rule's argmax search
  AtomCells cells;
  int32_t *cell_start_d = nullptr, *cell_atoms_d = nullptr;
  // score rule: each node's best atom, filled on first need after the atoms
  // were last set (0xFFFF = not yet searched); allocated at the first use
  uint16_t* best_atom_d = nullptr;
  bool best_atom_stale = true;
  std::vector<int32_t> lock_h;
  Atom* atoms_d = nullptr;
  float* Z_d = nullptr;
  int32_t* lock_d = nullptr;
  float *U = nullptr, *U2 = nullptr, *K = nullptr, *P = nullptr, *P2 = nullptr, *bestU = nullptr, *W = nullptr;
  void* arena = nullptr;  // one allocation holding the fine fields and labels
  // direct per-domain potential (pother_mode DIRECT): persistent P_m per
  // atom, its device pointer table and the masked right-hand side
  std::vector<float*> Pdom;
  float** Pdom_d = nullptr;
  float* rhs_dom = nullptr;
  // concurrent per-domain V-cycles (pother_mode DIRECT): domain m runs on
  // context m % dom.size(), each with its own stream, scratch and coarse levels
  struct DomCtx {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    float* x2 = nullptr;
    float* rhs = nullptr;
    std::vector<float*> e0, e1, g;
  };
  std::vector<DomCtx> dom;
  hipEvent_t dom_fork = nullptr;
  // frame potential (pother_mode FRAME): Pother, and the self-potential solve's
  // iterate pair and rhs (sic_interval > 0)
  float* Pother = nullptr;
  float* sic_x = nullptr;
  float* sic_x2 = nullptr;
  float* sic_g = nullptr;
  uint16_t *lab = nullptr, *lab2 = nullptr;
  // multigrid levels 1..L-1: S, two iterate buffers and the rhs
  int levels = 0;
  std::vector<int> Sl;
  std::vector<float*> e0, e1, g;
  Fx* red = nullptr;  // red_len fixed-point sums (per-domain energies + dipole, or norms)
  double *dst_S = nullptr, *dst_rden = nullptr;  // coarsest-level DST tables
  size_t red_len = 0;
  hipStream_t st = nullptr;
  Tile tile;                 // fine-level stencil launch shape
  Tile ptile;                // fine-level prolongation launch shape (two fine planes per trip want longer chunks)
  std::vector<Tile> ctile;   // coarse levels (index = level)
  bool initialized = false;
  bool timing = false;
  std::vector<EventPair> ev_pending;
  std::vector<hipEvent_t> ev_free;
  double fam_ms[F_COUNT] = {0};
  int64_t fam_n[F_COUNT] = {0};
  AdjState adj;
  bool use_graph = false;
  bool fused = true;  // fused multigrid passes (bit-identical to the unfused sequence)
  bool bottom = false;  // last two levels in one LDS-resident workgroup (with fused)
  bool legs = true;     // one launch per coarse level and direction (with fused)
  bool smooth2 = true;  // sweep pairs as one two-sweep march (with fused); RQ_SMOOTH2=0 turns it off
  // smallest level (nodes per axis) whose sweep pairs run as one march: below
  // it a level's three fields stay in the 256 MB MALL and two plain sweeps
  // measured faster (RQ_SMOOTH2_MIN_S overrides)
  int smooth2_min_s = kSmooth2MinS;
  // with smooth2: the fine level's pre-sweep pair, residual and restriction
  // as one march (RQ_SMOOTH2_RESRES=0 runs the pair and the restriction
  // march as two launches); s2rr_blocks: its target workgroup count
  bool s2rr = true;
  int s2rr_blocks = 0;
  // with smooth2 and an ITP update due: the fine correction, post-sweep pair
  // and ITP update as one march (RQ_POST_FUSE=0: separate launches)
  bool post_fuse = true;
  int post_blocks = 0;
  // largest level (nodes per axis) that runs the one-pass prolongation+sweep;
  // 0: none -- on gfx950 the separate correction and sweep launches measured
  // faster on every level of the 192^3 workload.  RQ_FUSE_MAX_S overrides it.
  int fuse_max_s = 0;
  int itp_fuse_max_s = kItpFuseMaxS;  // largest grid whose last sweep carries the ITP update (RQ_ITP_FUSE_MAX_S)
  std::vector<GraphEntry> graphs;
  // set by the normalisation's sums when an amplitude is not finite (4 B,
  // cleared by rq_init_trial and rq_write_field)
  int* health = nullptr;
  // after a HIP runtime error the handle refuses further device work
  // (molecule.js device.lost, :2258-2261); broken_msg keeps the first error
  bool broken = false;
  std::string broken_msg;
  int32_t fail_after = 0;  // rq_debug_fail_after: the n-th HIP check fails
  std::string err;
};

namespace {

int fail(rq_solver* h, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (h) h->err = buf;
  else g_create_error = buf;
  if (h && code == RQ_EHIP && !h->broken) {
    h->broken = true;
    h->broken_msg = buf;
  }
  return code;
}

// the armed rq_debug_fail_after countdown: true on the call it reaches
bool inject_due(rq_solver* h) { return h && h->fail_after > 0 && --h->fail_after == 0; }

// null handle, or a handle whose device work failed earlier
int gate(rq_solver* h) {
  if (!h) return fail(nullptr, RQ_EINVAL, "null handle");
  if (h->broken) return fail(h, RQ_EHIP, "handle unusable after an earlier HIP error (%s)", h->broken_msg.c_str());
  return RQ_OK;
}

#define HIPCHK(h, call)                                                                          \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ == hipSuccess && inject_due(h)) e_ = hipErrorLaunchFailure;                           \
    if (e_ != hipSuccess) return fail((h), RQ_EHIP, "%s: %s", #call, hipGetErrorString(e_));     \
  } while (0)

#define LAUNCHCHK(h)                                                                             \
  do {                                                                                           \
    hipError_t e_ = hipGetLastError();                                                           \
    if (e_ == hipSuccess && inject_due(h)) e_ = hipErrorLaunchFailure;                           \
    if (e_ != hipSuccess) return fail((h), RQ_EHIP, "kernel launch: %s", hipGetErrorString(e_)); \
  } while (0)

hipEvent_t take_event(rq_solver* s) {
  if (!s->ev_free.empty()) {
    hipEvent_t e = s->ev_free.back();
    s->ev_free.pop_back();
    return e;
  }
  // timing-only events: the host synchronises the stream before reading them,
  // so the system-scope fence (a cache write-back per event) is not needed and
  // would be charged to the bracketed kernel
  hipEvent_t e = nullptr;
  (void)hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
  return e;
}

void resolve_events(rq_solver* s) {
  if (s->ev_pending.empty()) return;
  (void)hipStreamSynchronize(s->st);
  for (auto& ep : s->ev_pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ep.a, ep.b) == hipSuccess) {
      s->fam_ms[ep.fam] += ms;
      s->fam_n[ep.fam] += 1;
    }
    s->ev_free.push_back(ep.a);
    s->ev_free.push_back(ep.b);
  }
  s->ev_pending.clear();
}

// Brackets a kernel launch with events when timing is on.
struct Timed {
  rq_solver* s;
  int fam;
  hipEvent_t a = nullptr;
  Timed(rq_solver* s_, int f) : s(s_), fam(f) {
    if (s->timing) {
      a = take_event(s);
      (void)hipEventRecord(a, s->st);
    }
  }
  ~Timed() {
    if (s->timing && a) {
      hipEvent_t b = take_event(s);
      (void)hipEventRecord(b, s->st);
      s->ev_pending.push_back({fam, a, b});
      if (s->ev_pending.size() > 4096) resolve_events(s);
    }
  }
};

// Built-in launch shape per level, from sweeps on gfx950 (192^3 workload):
// large levels 64x4 tiles marching 2..8 planes for about ``want`` workgroups
// (8192: 4 planes at 193^3 for the sweeps; the prolongation, which advances
// two fine planes per trip, is launched with 4096: 7 planes); mid levels
// 32x8; levels of <= 64 interior nodes per axis one plane per thread, since
// there a launch is latency-bound and a march only lengthens each thread's
// dependent chain.
Tile tile_for(int S, int want = 8192) {
  Tile t;
  const int n = S - 2;
  t.bx = n >= 128 ? 64 : (n >= 24 ? 32 : 16);
  t.by = 256 / t.bx;
  if (n >= 255) {
    // fields beyond the MALL stream from HBM: 128 x 2 tiles; the sweeps'
    // target (8192) becomes 16384 workgroups, 4 planes at 257^3 (sweep 33.3
    // us against 36.0 for 64 x 4 x 8, the ITP pass 77 against 80), 8 at
    // 513^3; the prolongation's (4096) stays, 8 planes
    t.bx = 128;
    t.by = 2;
    if (want >= 8192) want = 16384;
  }
  if (n <= 64) {
    t.ich = 1;
    return t;
  }
  const int gx = (n + t.bx - 1) / t.bx, gy = (n + t.by - 1) / t.by;
  int nz = std::max(1, std::min(n, want / std::max(1, gx * gy)));
  t.ich = std::min(8, std::max(2, (n + nz - 1) / nz));
  return t;
}

// Zeroes n fixed-point sums with a fill kernel (not hipMemsetAsync) so the
// pass is graph-capturable.
void zero_fx(rq_solver* s, Fx* p, size_t n) {
  launch_fill(s->st, reinterpret_cast<float*>(p), n * kFxWords * (sizeof(Fx) / sizeof(float)), 0.0f);
}

// Values of n fixed-point sums copied from the device (synchronises the stream).
// the normalisation's per-domain factors: n_atoms floats after the n_atoms
// norm sums at the start of red (red holds 3 n_atoms + 3 sums)
float* norm_scales(rq_solver* s) { return reinterpret_cast<float*>(s->red + kFxWords * (size_t)s->n_atoms); }

// health word idx (0: normalisation sums, 1: energy sums) after the stream drains
int read_flag(rq_solver* s, int idx, int* out) {
  HIPCHK(s, hipMemcpyAsync(out, s->health + idx, sizeof(int), hipMemcpyDeviceToHost, s->st));
  HIPCHK(s, hipStreamSynchronize(s->st));
  return RQ_OK;
}

int read_fx(rq_solver* s, const Fx* d, size_t n, double* out) {
  std::vector<Fx> w(n * kFxWords);
  HIPCHK(s, hipMemcpyAsync(w.data(), d, sizeof(Fx) * w.size(), hipMemcpyDeviceToHost, s->st));
  HIPCHK(s, hipStreamSynchronize(s->st));
  for (size_t i = 0; i < n; ++i) out[i] = fx_value(&w[i * kFxWords]);
  return RQ_OK;
}

int upload_atoms(rq_solver* s) {
  HIPCHK(s, hipMemcpyAsync(s->atoms_d, s->atoms_h.data(), sizeof(Atom) * s->n_atoms, hipMemcpyHostToDevice, s->st));
  // the S2 fields with their defaults resolved against the current Z
  // (molecule.js:22-25, 1887-1892: Z_nuc || Z, initZeff || Z, initRcut || 1e6)
  std::vector<AtomS2> ext(s->n_atoms);
  s->any_split = false;
  for (int a = 0; a < s->n_atoms; ++a) {
    const rq_atom_s2& x = s->ext_h[a];
    const float Z = s->atoms_h[a].Z;
    ext[a] = AtomS2{x.Z_nuc > 0.0f ? x.Z_nuc : Z, x.init_zeff > 0.0f ? x.init_zeff : Z,
                    x.init_rcut > 0.0f ? x.init_rcut : 1e6f, x.split_type, x.split_idx,
                    x.split_axis[0], x.split_axis[1], x.split_axis[2], x.split_rot};
    s->any_split |= x.split_type > 1;
  }
  HIPCHK(s, hipMemcpyAsync(s->ext_d, ext.data(), sizeof(AtomS2) * s->n_atoms, hipMemcpyHostToDevice, s->st));
  HIPCHK(s, hipStreamSynchronize(s->st));
  s->best_atom_stale = true;
  std::vector<float> Z(s->n_atoms);
  for (int a = 0; a < s->n_atoms; ++a) Z[a] = s->atoms_h[a].Z;
  HIPCHK(s, hipMemcpyAsync(s->Z_d, Z.data(), sizeof(float) * s->n_atoms, hipMemcpyHostToDevice, s->st));
  // cell list: atoms with Z > 0 binned by floor(position / cs), clamped
  AtomCells& c = s->cells;
  const int nc3 = c.n * c.n * c.n;
  std::vector<int32_t> start(nc3 + 1, 0), order;
  std::vector<int> cell_of(s->n_atoms, -1);
  c.zmax = 0.0f;
  auto bin = [&](float x) { return std::min(c.n - 1, std::max(0, (int)std::floor(x / (float)c.cs))); };
  for (int a = 0; a < s->n_atoms; ++a) {
    const Atom& at = s->atoms_h[a];
    if (!(at.Z > 0.0f)) continue;
    c.zmax = std::max(c.zmax, at.Z);
    cell_of[a] = (bin(at.i) * c.n + bin(at.j)) * c.n + bin(at.k);
    ++start[cell_of[a] + 1];
  }
  for (int q = 0; q < nc3; ++q) start[q + 1] += start[q];
  order.assign(std::max(1, start[nc3]), 0);
  std::vector<int32_t> fill(start.begin(), start.end() - 1);
  for (int a = 0; a < s->n_atoms; ++a)
    if (cell_of[a] >= 0) order[fill[cell_of[a]]++] = a;
  HIPCHK(s, hipMemcpyAsync(s->cell_start_d, start.data(), sizeof(int32_t) * (nc3 + 1), hipMemcpyHostToDevice, s->st));
  HIPCHK(s, hipMemcpyAsync(s->cell_atoms_d, order.data(), sizeof(int32_t) * order.size(), hipMemcpyHostToDevice,
                           s->st));
  HIPCHK(s, hipStreamSynchronize(s->st));
  return RQ_OK;
}

// ---- Poisson ----
int smooth_fine(rq_solver* s, float a, float c) {
  {
    Timed t(s, F_SMOOTH);
    launch_smooth_fine(s->st, s->tile, s->S, s->P, s->P2, s->U, a, c, s->two_pi, s->h2);
  }
  LAUNCHCHK(s);
  std::swap(s->P, s->P2);
  return RQ_OK;
}

// count fine sweeps; with fused, pairs on large grids run as one two-sweep march (bit-identical)
int smooth_fine_n(rq_solver* s, float a, float c, int count) {
  for (; s->fused && s->smooth2 && s->S >= s->smooth2_min_s && count >= 2; count -= 2) {
    {
      Timed t(s, F_SMOOTH);
      launch_smooth2(s->st, true, s->S, s->P, s->P2, s->U, a, c, s->two_pi, s->h2);
    }
    LAUNCHCHK(s);
    std::swap(s->P, s->P2);
  }
  for (; count > 0; --count)
    if (int rc = smooth_fine(s, a, c)) return rc;
  return RQ_OK;
}

// V-cycle on coarse level l (x = e0[l], rhs g[l]); zero initial guess.
int vcycle_coarse(rq_solver* s, int l) {
  const int S = s->Sl[l];
  const size_t n3 = (size_t)S * S * S;
  if (l == s->levels - 1) {
    Timed t(s, F_COARSE);
    launch_coarse_exact(s->st, S, s->g[l], s->e0[l], s->dst_S, s->dst_rden);
    LAUNCHCHK(s);
    return RQ_OK;
  }
  if (s->fused && s->bottom && l == s->levels - 2 && s->dst_S && bottom_fits(S)) {
    {
      Timed t(s, F_BOTTOM);
      launch_bottom(s->st, S, s->p.mg_nu1, s->p.mg_nu2, s->p.mg_a, s->p.mg_c, s->g[l], s->e0[l], s->dst_S,
                    s->dst_rden);
    }
    LAUNCHCHK(s);
    return RQ_OK;
  }
  const Tile tl = s->ctile[l];
  const float a = s->p.mg_a, c = s->p.mg_c;
  // per-level legs pay off where launches are latency-bound (<= 65^3 nodes);
  // above, their halo recomputation costs more than the launches they save
  if (s->fused && s->legs && S <= kLegsMaxS && s->p.mg_nu1 == 2 && s->p.mg_nu2 >= 1 && s->p.mg_nu2 <= 3) {
    {
      Timed t(s, F_RESTRICT_C);
      launch_down2(s->st, S, s->g[l], s->e0[l], s->g[l + 1], a, c);
    }
    LAUNCHCHK(s);
    int rc = vcycle_coarse(s, l + 1);
    if (rc) return rc;
    {
      Timed t(s, F_PROLONG_C);
      launch_up(s->st, S, s->p.mg_nu2, s->e0[l], s->e0[l + 1], s->e1[l], s->g[l], a, c);
    }
    LAUNCHCHK(s);
    std::swap(s->e0[l], s->e1[l]);
    return RQ_OK;
  }
  int pre = s->p.mg_nu1;
  if (s->fused && pre >= 2) {
    {
      Timed t(s, F_SMOOTH_C);
      launch_smooth2_from_zero(s->st, tl, S, s->g[l], s->e0[l], a, c);
    }
    pre -= 2;
  } else {
    launch_zero(s->st, s->e0[l], n3);
  }
  for (int it = 0; it < pre; ++it) {
    {
      Timed t(s, F_SMOOTH_C);
      launch_smooth_coarse(s->st, tl, S, s->e0[l], s->e1[l], s->g[l], a, c);
    }
    std::swap(s->e0[l], s->e1[l]);
  }
  // the one-pass prolongation+sweep trades extra cache traffic for a launch:
  // it pays off only on small, latency-bound levels
  const bool small = s->fused && S <= s->fuse_max_s;
  {
    Timed t(s, F_RESTRICT_C);
    if (s->fused) {
      launch_residual_restrict(s->st, false, S, s->e0[l], s->g[l], 1.0f, 1.0f, s->g[l + 1], 1.0f);
    } else {
      launch_residual_coarse(s->st, tl, S, s->e0[l], s->g[l], s->e1[l]);
      launch_restrict4(s->st, S, s->e1[l], s->g[l + 1], 1.0f);
    }
  }
  LAUNCHCHK(s);
  int rc = vcycle_coarse(s, l + 1);
  if (rc) return rc;
  int post = s->p.mg_nu2;
  if (small && post >= 1) {
    {
      Timed t(s, F_PROLONG_C);
      launch_prolong_smooth(s->st, tl, false, S, s->e0[l], s->e0[l + 1], s->e1[l], s->g[l], a, c, 1.0f, 1.0f);
    }
    std::swap(s->e0[l], s->e1[l]);
    --post;
  } else {
    Timed t(s, F_PROLONG_C);
    launch_prolong_correct(s->st, tl, S, s->e0[l], s->e0[l + 1], 1.0f);
  }
  for (; s->fused && s->smooth2 && S >= s->smooth2_min_s && post >= 2; post -= 2) {
    {
      Timed t(s, F_SMOOTH_C);
      launch_smooth2(s->st, false, S, s->e0[l], s->e1[l], s->g[l], a, c, 1.0f, 1.0f);
    }
    std::swap(s->e0[l], s->e1[l]);
  }
  for (int it = 0; it < post; ++it) {
    {
      Timed t(s, F_SMOOTH_C);
      launch_smooth_coarse(s->st, tl, S, s->e0[l], s->e1[l], s->g[l], a, c);
    }
    std::swap(s->e0[l], s->e1[l]);
  }
  LAUNCHCHK(s);
  return RQ_OK;
}

// One V-cycle on the fine level.  With fuse_itp the last post-smoothing sweep
// and the ITP update run as one kernel and *did_itp is set.
int vcycle_fine(rq_solver* s, bool fuse_itp, bool* did_itp) {
  if (did_itp) *did_itp = false;
  const bool itp_due = fuse_itp && did_itp;
  fuse_itp = fuse_itp && s->S <= s->itp_fuse_max_s;
  const float a = s->p.mg_a, c = s->p.mg_c;
  const bool small = s->fused && s->S <= s->fuse_max_s;
  if (s->fused && s->smooth2 && s->s2rr && s->S >= s->smooth2_min_s && s->p.mg_nu1 == 2) {
    {
      Timed t(s, F_SMOOTH_RESTRICT);
      launch_smooth2_resres(s->st, s->S, s->P, s->P2, s->U, a, c, s->two_pi, s->h2, s->g[1], s->s2rr_blocks);
    }
    LAUNCHCHK(s);
    std::swap(s->P, s->P2);
  } else {
    if (int rc = smooth_fine_n(s, a, c, s->p.mg_nu1)) return rc;
    {
      Timed t(s, F_RESTRICT);
      if (s->fused) {
        launch_residual_restrict(s->st, true, s->S, s->P, s->U, s->two_pi, s->h2, s->g[1], 1.0f);
      } else {
        launch_residual_fine(s->st, s->tile, s->S, s->P, s->U, s->two_pi, s->h2, s->P2);
        launch_restrict4(s->st, s->S, s->P2, s->g[1], 1.0f);
      }
    }
    LAUNCHCHK(s);
  }
  int rc = vcycle_coarse(s, 1);
  if (rc) return rc;
  int post = s->p.mg_nu2;
  if (itp_due && s->fused && s->smooth2 && s->post_fuse && s->S >= s->smooth2_min_s && post == 2 && !small) {
    {
      Timed t(s, F_POST_ITP);
      launch_post_itp(s->st, s->S, s->P, s->e0[1], s->P2, s->U, s->U2, s->K, s->lab, a, c, s->two_pi, s->h2,
                      s->half_dv, s->dtv, s->post_blocks);
    }
    LAUNCHCHK(s);
    std::swap(s->P, s->P2);
    std::swap(s->U, s->U2);
    *did_itp = true;
    return RQ_OK;
  }
  if (small && post >= 1) {
    {
      Timed t(s, F_PROLONG);
      launch_prolong_smooth(s->st, s->tile, true, s->S, s->P, s->e0[1], s->P2, s->U, a, c, s->two_pi, s->h2);
    }
    std::swap(s->P, s->P2);
    --post;
  } else {
    Timed t(s, F_PROLONG);
    launch_prolong_correct(s->st, s->ptile, s->S, s->P, s->e0[1], 1.0f);
  }
  LAUNCHCHK(s);
  // plain sweeps before the last one (which may carry the ITP update)
  if (s->fused && fuse_itp && post >= 3) {
    if (int rc2 = smooth_fine_n(s, a, c, post - 1)) return rc2;
    post = 1;
  } else if (!(s->fused && fuse_itp)) {
    return smooth_fine_n(s, a, c, post);
  }
  for (int it = 0; it < post; ++it) {
    if (s->fused && fuse_itp && it == post - 1) {
      {
        Timed t(s, F_SMOOTH_ITP);
        launch_smooth_itp(s->st, s->tile, s->S, s->P, s->P2, s->U, s->U2, s->K, s->lab, a, c, s->two_pi, s->h2,
                          s->half_dv, s->dtv);
      }
      LAUNCHCHK(s);
      std::swap(s->P, s->P2);
      std::swap(s->U, s->U2);
      if (did_itp) *did_itp = true;
      break;
    }
    rc = smooth_fine(s, a, c);
    if (rc) return rc;
  }
  return RQ_OK;
}

// molecule.js two-grid correction (sign-inverted, 0.5 damping), molecule.js:2879-2907.
int twogrid_correction(rq_solver* s) {
  const int S1 = s->Sl[1];
  const size_t n3 = (size_t)S1 * S1 * S1;
  const Tile tl = s->ctile[1];
  {
    Timed t(s, F_RESTRICT);
    launch_residual_fine(s->st, s->tile, s->S, s->P, s->U, s->two_pi, s->h2, s->P2);
    launch_restrict4(s->st, s->S, s->P2, s->g[1], -1.0f);
  }
  launch_zero(s->st, s->e0[1], n3);
  for (int it = 0; it < s->p.twogrid_coarse_sweeps; ++it) {
    {
      Timed t(s, F_SMOOTH_C);
      launch_smooth_coarse(s->st, tl, S1, s->e0[1], s->e1[1], s->g[1], s->p.smooth_a, s->p.smooth_c);
    }
    std::swap(s->e0[1], s->e1[1]);
  }
  {
    Timed t(s, F_PROLONG);
    launch_prolong_correct(s->st, s->ptile, s->S, s->P, s->e0[1], 0.5f);
  }
  LAUNCHCHK(s);
  return RQ_OK;
}

int copy_field(rq_solver* s, float* dst, const float* src);

// One V-cycle on the fine grid for x (in place, x2 scratch) with the
// precomputed right-hand side rhs: the vcycle_fine sequence with rhs = src[id]
// in place of (2 pi U^2) h^2.
int vcycle_rhs(rq_solver* s, float* x, float* x2, const float* rhs) {
  const float a = s->p.mg_a, c = s->p.mg_c;
  float* cur = x;
  float* oth = x2;
  for (int it = 0; it < s->p.mg_nu1; ++it) {
    {
      Timed t(s, F_SMOOTH);
      launch_smooth_coarse(s->st, s->tile, s->S, cur, oth, rhs, a, c);
    }
    std::swap(cur, oth);
  }
  {
    Timed t(s, F_RESTRICT);
    if (s->fused) {
      launch_residual_restrict(s->st, false, s->S, cur, rhs, 1.0f, 1.0f, s->g[1], 1.0f);
    } else {
      launch_residual_coarse(s->st, s->tile, s->S, cur, rhs, oth);
      launch_restrict4(s->st, s->S, oth, s->g[1], 1.0f);
    }
  }
  LAUNCHCHK(s);
  if (int rc = vcycle_coarse(s, 1)) return rc;
  {
    Timed t(s, F_PROLONG);
    launch_prolong_correct(s->st, s->ptile, s->S, cur, s->e0[1], 1.0f);
  }
  for (int it = 0; it < s->p.mg_nu2; ++it) {
    {
      Timed t(s, F_SMOOTH);
      launch_smooth_coarse(s->st, s->tile, s->S, cur, oth, rhs, a, c);
    }
    std::swap(cur, oth);
  }
  LAUNCHCHK(s);
  if (cur != x) return copy_field(s, x, cur);
  return RQ_OK;
}

// molecule.js direct Pother (molecule.js:2834-2858): per domain with Z > 0,
// rho_other -> one V-cycle on its persistent potential; then the gather,
// or with fuse_itp the gather fused into the step's ITP update (*did_itp).
int pother_direct_step(rq_solver* s, bool fuse_itp, bool* did_itp) {
  if (s->dom.empty()) {
    for (int m = 0; m < s->n_atoms; ++m) {
      if (!(s->atoms_h[m].Z > 0.0f)) continue;
      launch_rho_other(s->st, s->tile, s->S, s->U, s->lab, m, s->two_pi, s->h2, s->rhs_dom);
      if (int rc = vcycle_rhs(s, s->Pdom[m], s->P2, s->rhs_dom)) return rc;
    }
  } else {
    // fork: every context stream waits for the main stream (U current), runs
    // its domains' V-cycles with its own scratch, and the main stream joins
    // them before the gather.  The launch code reads the stream and the
    // coarse buffers from the handle, so each context's are swapped in while
    // its launches are enqueued (host-side; the streams then run concurrently).
    HIPCHK(s, hipEventRecord(s->dom_fork, s->st));
    for (auto& d : s->dom) HIPCHK(s, hipStreamWaitEvent(d.st, s->dom_fork, 0));
    hipStream_t main_st = s->st;
    float* main_x2 = s->P2;
    int rc = RQ_OK;
    for (int m = 0; m < s->n_atoms && rc == RQ_OK; ++m) {
      if (!(s->atoms_h[m].Z > 0.0f)) continue;
      auto& d = s->dom[m % s->dom.size()];
      s->st = d.st;
      std::swap(s->e0, d.e0);
      std::swap(s->e1, d.e1);
      std::swap(s->g, d.g);
      launch_rho_other(s->st, s->tile, s->S, s->U, s->lab, m, s->two_pi, s->h2, d.rhs);
      rc = vcycle_rhs(s, s->Pdom[m], d.x2, d.rhs);
      std::swap(s->e0, d.e0);
      std::swap(s->e1, d.e1);
      std::swap(s->g, d.g);
      s->st = main_st;
    }
    s->P2 = main_x2;
    if (rc) return rc;
    for (auto& d : s->dom) {
      HIPCHK(s, hipEventRecord(d.done, d.st));
      HIPCHK(s, hipStreamWaitEvent(s->st, d.done, 0));
    }
  }
  if (fuse_itp) {
    {
      Timed t(s, F_ITP);
      launch_itp_gather(s->st, s->tile, s->S, s->U, s->U2, s->K, s->lab, s->Pdom_d, s->n_atoms, s->P, s->half_dv,
                        s->dtv);
    }
    LAUNCHCHK(s);
    std::swap(s->U, s->U2);
    if (did_itp) *did_itp = true;
    return RQ_OK;
  }
  launch_pother_gather(s->st, s->tile, s->S, s->lab, s->Pdom_d, s->n_atoms, s->P);
  LAUNCHCHK(s);
  return RQ_OK;
}

// molecule.js direct Pother, its own scheme (molecule.js:2833-2858): for
// each domain m with Z != 0, rho_other_m, kDirectSweeps damped sweeps of the
// persistent P_m (ping-pong with P2, whose faces are zero like P_m's), and
// Pother = P_m on the nodes labelled m.  The rhs is (h^2 2pi) rho, the
// page's product order.
int pother_sweeps_step(rq_solver* s) {
  const float alpha = s->h2 * s->two_pi;
  for (int m = 0; m < s->n_atoms; ++m) {
    if (s->atoms_h[m].Z == 0.0f) continue;
    launch_rho_other(s->st, s->tile, s->S, s->U, s->lab, m, alpha, 1.0f, s->rhs_dom);
    float* x = s->Pdom[m];
    float* y = s->P2;
    for (int it = 0; it < kDirectSweeps; ++it) {
      Timed t(s, F_SMOOTH);
      launch_smooth_coarse(s->st, s->tile, s->S, x, y, s->rhs_dom, s->p.smooth_a, s->p.smooth_c);
      std::swap(x, y);
    }
    if (x != s->Pdom[m]) HIPCHK(s, hipMemcpyAsync(s->Pdom[m], x, s->S3 * sizeof(float), hipMemcpyDeviceToDevice, s->st));
    launch_copy_for_label(s->st, s->tile, s->S, s->Pdom[m], s->lab, m, s->Pother);
  }
  LAUNCHCHK(s);
  return RQ_OK;
}

// the potential the ITP update, the energy and the grad-P forces read
const float* itp_potential(const rq_solver* s) {
  const int pm = s->p.pother_mode;
  return (pm == RQ_POTHER_FRAME || pm == RQ_POTHER_DIRECT_SWEEPS) ? s->Pother : s->P;
}
const float* gradp_potential(const rq_solver* s) {
  return s->p.pother_mode == RQ_POTHER_DIRECT_SWEEPS ? s->Pother : s->P;
}

int poisson_step(rq_solver* s, bool twogrid_due, bool fuse_itp = false, bool* did_itp = nullptr) {
  if (did_itp) *did_itp = false;
  if (s->p.pother_mode == RQ_POTHER_DIRECT_SWEEPS) return pother_sweeps_step(s);
  if (s->p.pother_mode == RQ_POTHER_DIRECT) return pother_direct_step(s, fuse_itp && s->fused, did_itp);
  if (s->p.poisson_mode == RQ_POISSON_MULTIGRID) return vcycle_fine(s, fuse_itp, did_itp);
  if (int rc = smooth_fine_n(s, s->p.smooth_a, s->p.smooth_c, s->p.jacobi_sweeps)) return rc;
  if (s->p.poisson_mode == RQ_POISSON_TWOGRID && twogrid_due) return twogrid_correction(s);
  return RQ_OK;
}

int update_step(rq_solver* s) {
  {
    Timed t(s, F_ITP);
    const float* Pv = itp_potential(s);
    launch_itp(s->st, s->tile, s->S, s->U, s->U2, s->K, Pv, s->lab, s->half_dv, s->dtv);
  }
  LAUNCHCHK(s);
  std::swap(s->U, s->U2);
  return RQ_OK;
}

int normalize_step(rq_solver* s) {
  Fx* norms = s->red;
  zero_fx(s, norms, s->n_atoms);
  {
    Timed t(s, F_NORM);
    launch_norm_sums(s->st, s->tile, s->S, s->h3, s->U, s->lab, s->n_atoms, norms, s->health);
    launch_normalize(s->st, s->tile, s->S, s->U, s->lab, norms, s->Z_d, s->n_atoms, norm_scales(s));
  }
  LAUNCHCHK(s);
  return RQ_OK;
}

int boundary_step(rq_solver* s) {
  if (s->p.boundary_mode == RQ_BOUNDARY_OFF) return RQ_OK;
  {
    Timed t(s, F_BOUNDARY);
    if (s->p.boundary_mode == RQ_BOUNDARY_SCORE) {
      // S^3 uint16 entries, rounded up to whole floats for the fill kernel
      const size_t nf = ((size_t)s->S3 + 1) / 2;
      if (!s->best_atom_d) HIPCHK(s, hipMalloc(&s->best_atom_d, nf * sizeof(float)));
      if (s->best_atom_stale) {
        // the best atom of every interior node, once per geometry: the rule
        // then evaluates two scores per candidate node and never searches
        launch_score_map(s->st, s->S, s->h, s->h2, s->atoms_d, s->n_atoms, s->best_atom_d);
        s->best_atom_stale = false;
      }
      launch_boundary_score(s->st, s->tile, s->S, s->h, s->h2, s->atoms_d, s->n_atoms, s->lab, s->lab2,
                            s->p.flip_ratio, s->p.flip_min_foreign, s->cells, s->cell_start_d, s->cell_atoms_d,
                            s->best_atom_d);
    } else {
      bool any_lock = false;
      for (int32_t g : s->lock_h) any_lock |= (g > 0);
      launch_levelset(s->st, s->tile, s->S, s->h, s->atoms_d, s->any_split ? s->ext_d : nullptr, s->n_atoms,
                      any_lock ? s->lock_d : nullptr, s->lab,
                      s->lab2, s->U, s->W, s->p.ls_dt_w, s->p.ls_curv_reg, s->p.rc_dirichlet);
    }
  }
  LAUNCHCHK(s);
  // faces are never written by the boundary kernels: keep them in the new buffer
  std::swap(s->lab, s->lab2);
  // copy the whole buffer back so both halves agree on the faces and interior
  HIPCHK(s, hipMemcpyAsync(s->lab2, s->lab, s->S3 * sizeof(uint16_t), hipMemcpyDeviceToDevice, s->st));
  return RQ_OK;
}

float* field_ptr(rq_solver* s, int f) {
  switch (f) {
    case RQ_FIELD_U: return s->U;
    case RQ_FIELD_U2: return s->U2;
    case RQ_FIELD_K: return s->K;
    case RQ_FIELD_P: return s->P;
    case RQ_FIELD_P2: return s->P2;
    case RQ_FIELD_BESTU: return s->bestU;
    case RQ_FIELD_W: return s->W;
    case RQ_FIELD_POTHER: return s->Pother;
    default: return nullptr;
  }
}

void free_all(rq_solver* s) {
  // drain the stream, then destroy the step graphs before the buffers their
  // kernel nodes reference are freed
  if (s->st) (void)hipStreamSynchronize(s->st);
  for (auto& ge : s->graphs) {
    (void)hipGraphExecDestroy(ge.exec);
    (void)hipGraphDestroy(ge.graph);
  }
  s->graphs.clear();
  auto fr = [](void* p) {
    if (p) (void)hipFree(p);
  };
  fr(s->atoms_d);
  fr(s->ext_d);
  fr(s->best_atom_d);
  fr(s->Z_d);
  fr(s->lock_d);
  fr(s->arena);
  for (float* q : s->Pdom) fr(q);
  fr(s->Pdom_d);
  fr(s->rhs_dom);
  for (auto& d : s->dom) {
    fr(d.x2);
    fr(d.rhs);
    for (float* q : d.e0) fr(q);
    for (float* q : d.e1) fr(q);
    for (float* q : d.g) fr(q);
    if (d.done) (void)hipEventDestroy(d.done);
    if (d.st) (void)hipStreamDestroy(d.st);
  }
  if (s->dom_fork) (void)hipEventDestroy(s->dom_fork);
  fr(s->Pother);
  fr(s->sic_x);
  fr(s->sic_x2);
  fr(s->sic_g);
  fr(s->cell_start_d);
  fr(s->cell_atoms_d);
  fr(s->red);
  fr(s->health);
  fr(s->dst_S);
  fr(s->dst_rden);
  for (size_t l = 0; l < s->e0.size(); ++l) {
    fr(s->e0[l]);
    fr(s->e1[l]);
    fr(s->g[l]);
  }
  for (auto& ep : s->ev_pending) {
    (void)hipEventDestroy(ep.a);
    (void)hipEventDestroy(ep.b);
  }
  for (auto e : s->ev_free) (void)hipEventDestroy(e);
  AdjState& A = s->adj;
  for (float* p : {A.lamU, A.lamU2, A.lamP, A.lamP2, A.gK, A.acc, A.t1, A.t2, A.t3, A.wU, A.wP}) fr(p);
  for (float* p : A.ckU) fr(p);
  for (float* p : A.ckP) fr(p);
  for (float* p : A.segU) fr(p);
  for (float* p : A.segP) fr(p);
  fr(A.red);
  if (s->st) (void)hipStreamDestroy(s->st);
}

// ---------------------------------------------------------------------------
// Frozen-label trajectory and its adjoint (rq_solve_frozen / rq_adjoint_gradient)
// ---------------------------------------------------------------------------
int adj_alloc(rq_solver* s, float** p) {
  if (*p) return RQ_OK;
  if (hipMalloc((void**)p, s->S3 * sizeof(float)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(s, RQ_ENOMEM, "adjoint: device allocation of a %zu-node field failed", s->S3);
  }
  HIPCHK(s, hipMemsetAsync(*p, 0, s->S3 * sizeof(float), s->st));
  return RQ_OK;
}

int adj_alloc_vec(rq_solver* s, std::vector<float*>& v, size_t n) {
  while (v.size() < n) {
    float* p = nullptr;
    int rc = adj_alloc(s, &p);
    if (rc) return rc;
    v.push_back(p);
  }
  return RQ_OK;
}

int copy_field(rq_solver* s, float* dst, const float* src) {
  HIPCHK(s, hipMemcpyAsync(dst, src, s->S3 * sizeof(float), hipMemcpyDeviceToDevice, s->st));
  return RQ_OK;
}

// the buffer of {t1, t2} that is not `busy`
float* other_of(float* busy, float* t1, float* t2) { return busy == t1 ? t2 : t1; }

// One Poisson step P_out = Poisson(P_in, U) with explicit buffers; t1, t2 are
// scratch distinct from P_in, P_out, U and U_out.  With U_out given, a fused
// multigrid step runs its last post-smoothing sweep together with the ITP
// update U -> U_out on the new P (k_smooth_itp, the values of the sweep then
// k_itp) and sets *did_itp; otherwise the caller applies the ITP itself.
int poisson_into(rq_solver* s, const float* P_in, float* P_out, const float* U, float* t1, float* t2,
                 float* U_out = nullptr, bool* did_itp = nullptr) {
  const Tile& T = s->tile;
  if (did_itp) *did_itp = false;
  if (s->p.poisson_mode == RQ_POISSON_MULTIGRID) {
    float* x;
    const float a = s->p.mg_a, c = s->p.mg_c;
    // on the grids where the solver's own step marches (>= smooth2_min_s
    // nodes per axis, V(2,2)), the same marches: bit-identical values
    const bool march = s->fused && s->smooth2 && s->s2rr && s->S >= s->smooth2_min_s && s->p.mg_nu1 == 2;
    if (march) {
      launch_smooth2_resres(s->st, s->S, P_in, t1, U, a, c, s->two_pi, s->h2, s->g[1], s->s2rr_blocks);
      LAUNCHCHK(s);
      if (int rc = vcycle_coarse(s, 1)) return rc;
      if (U_out && s->post_fuse && s->p.mg_nu2 == 2 && s->S > s->fuse_max_s) {
        launch_post_itp(s->st, s->S, t1, s->e0[1], P_out, U, U_out, s->K, s->lab, a, c, s->two_pi, s->h2,
                        s->half_dv, s->dtv, s->post_blocks);
        LAUNCHCHK(s);
        if (did_itp) *did_itp = true;
        return RQ_OK;
      }
      x = t1;
    } else if (s->p.mg_nu1 == 0) {
      if (int rc = copy_field(s, t1, P_in)) return rc;
      x = t1;
    } else {
      const float* src = P_in;
      float* dst = t1;
      for (int it = 0; it < s->p.mg_nu1; ++it) {
        launch_smooth_fine(s->st, T, s->S, src, dst, U, a, c, s->two_pi, s->h2);
        src = dst;
        dst = other_of(dst, t1, t2);
      }
      x = const_cast<float*>(src);
    }
    if (!march) {
      // residual + restriction in one pass (the values of k_residual then
      // k_restrict4, without the fine residual array)
      if (s->fused) {
        launch_residual_restrict(s->st, true, s->S, x, U, s->two_pi, s->h2, s->g[1], 1.0f);
      } else {
        launch_residual_fine(s->st, T, s->S, x, U, s->two_pi, s->h2, P_out);
        launch_restrict4(s->st, s->S, P_out, s->g[1], 1.0f);
      }
      LAUNCHCHK(s);
      if (int rc = vcycle_coarse(s, 1)) return rc;
    }
    launch_prolong_correct(s->st, s->ptile, s->S, x, s->e0[1], 1.0f);
    if (s->p.mg_nu2 == 0) return copy_field(s, P_out, x);
    float* src = x;
    for (int it = 0; it < s->p.mg_nu2; ++it) {
      const bool last = it == s->p.mg_nu2 - 1;
      float* dst = last ? P_out : other_of(src, t1, t2);
      if (last && U_out && s->fused && s->S <= s->itp_fuse_max_s) {
        launch_smooth_itp(s->st, T, s->S, src, dst, U, U_out, s->K, s->lab, a, c, s->two_pi, s->h2, s->half_dv,
                          s->dtv);
        if (did_itp) *did_itp = true;
      } else {
        launch_smooth_fine(s->st, T, s->S, src, dst, U, a, c, s->two_pi, s->h2);
      }
      src = dst;
    }
    LAUNCHCHK(s);
    return RQ_OK;
  }
  const int n = s->p.jacobi_sweeps;
  if (n == 0) return copy_field(s, P_out, P_in);
  const float* src = P_in;
  for (int it = 0; it < n; ++it) {
    float* dst = (it == n - 1) ? P_out : (src == t1 ? t2 : t1);
    launch_smooth_fine(s->st, T, s->S, src, dst, U, s->p.smooth_a, s->p.smooth_c, s->two_pi, s->h2);
    src = dst;
  }
  LAUNCHCHK(s);
  return RQ_OK;
}

int normalize_on(rq_solver* s, float* U) {
  zero_fx(s, s->red, s->n_atoms);
  launch_norm_sums(s->st, s->tile, s->S, s->h3, U, s->lab, s->n_atoms, s->red, s->health);
  launch_normalize(s->st, s->tile, s->S, U, s->lab, s->red, s->Z_d, s->n_atoms, norm_scales(s));
  LAUNCHCHK(s);
  return RQ_OK;
}

// Adjoint of one Poisson step: lamP (w.r.t. P_out) -> w.r.t. P_in, in place
// via lamP2; lamU += (rhs adjoint) * 2 alpha beta U.
int poisson_adj(rq_solver* s, const float* U) {
  AdjState& A = s->adj;
  const Tile& T = s->tile;
  const float ab2 = 2.0f * s->two_pi * s->h2;
  if (s->p.poisson_mode == RQ_POISSON_MULTIGRID) {
    const float a = s->p.mg_a, c = s->p.mg_c;
    // the first sweep's adjoint writes acc (no zero fill before it); the
    // last one adds the rhs adjoint into lamU (no separate pass)
    if (s->p.mg_nu2 == 0) launch_fill(s->st, A.acc, s->S3, 0.0f);
    for (int it = 0; it < s->p.mg_nu2; ++it) {
      launch_smooth_adj(s->st, T, s->S, it == 0 ? kAdjFirst : kAdjAcc, A.lamP, A.lamP2, nullptr, A.acc, nullptr, a,
                        c, 1.0f / c, 0.0f);
      std::swap(A.lamP, A.lamP2);
    }
    launch_restrict4(s->st, s->S, A.lamP, s->g[1], 1.0f);
    LAUNCHCHK(s);
    if (int rc = vcycle_coarse(s, 1)) return rc;
    // t3's faces are zero since its allocation and never written
    launch_prolong_set(s->st, T, s->S, A.t3, s->e0[1]);
    launch_correct_adj(s->st, T, s->S, A.lamP, A.t3, A.lamP2, A.acc);
    std::swap(A.lamP, A.lamP2);
    for (int it = 0; it < s->p.mg_nu1; ++it) {
      const bool last = it == s->p.mg_nu1 - 1;
      launch_smooth_adj(s->st, T, s->S, last ? kAdjLast : kAdjAcc, A.lamP, A.lamP2, last ? U : nullptr, A.acc,
                        last ? A.lamU : nullptr, a, c, 1.0f / c, ab2);
      std::swap(A.lamP, A.lamP2);
    }
    if (s->p.mg_nu1 == 0) launch_axpy_src(s->st, T, s->S, U, A.acc, A.lamU, ab2);
    LAUNCHCHK(s);
    return RQ_OK;
  }
  const float a = s->p.smooth_a, c = s->p.smooth_c;
  for (int it = 0; it < s->p.jacobi_sweeps; ++it) {
    launch_smooth_adj(s->st, T, s->S, kAdjSrc, A.lamP, A.lamP2, U, A.lamU, nullptr, a, c, ab2 / c, 0.0f);
    std::swap(A.lamP, A.lamP2);
  }
  LAUNCHCHK(s);
  return RQ_OK;
}

// Adjoint of normalise (on lamU in place) given the pre-normalise field Upre.
int normalize_adj(rq_solver* s, const float* Upre) {
  AdjState& A = s->adj;
  Fx* norms = A.red;
  Fx* dots = A.red + kFxWords * (size_t)s->n_atoms;
  zero_fx(s, A.red, 2 * (size_t)s->n_atoms);
  launch_norm_sums(s->st, s->tile, s->S, s->h3, Upre, s->lab, s->n_atoms, norms, s->health);
  launch_domain_dot(s->st, s->tile, s->S, 1.0f, Upre, A.lamU, s->lab, s->n_atoms, dots);
  launch_normalize_adj(s->st, s->tile, s->S, s->h3, Upre, A.lamU, s->lab, norms, dots, s->Z_d, s->n_atoms);
  LAUNCHCHK(s);
  return RQ_OK;
}

int frozen_run(rq_solver* s, const rq_adjoint_spec* sp, bool backward, double* E3, double* dEdR) {
  if (s->p.pother_mode != RQ_POTHER_TOTAL)
    return fail(s, RQ_EINVAL, "the adjoint differentiates the total-potential map (pother_mode TOTAL)");
  if (!sp) return fail(s, RQ_EINVAL, "null spec");
  for (int r : sp->reserved)
    if (r) return fail(s, RQ_EINVAL, "rq_adjoint_spec.reserved must be zero");
  if (sp->start != RQ_ADJ_START_TRIAL && sp->start != RQ_ADJ_START_SAVED)
    return fail(s, RQ_EINVAL, "rq_adjoint_spec.start %d", sp->start);
  const bool saved = sp->start == RQ_ADJ_START_SAVED;
  if (saved && !s->adj.wU) return fail(s, RQ_ESTATE, "RQ_ADJ_START_SAVED before rq_adjoint_save_start");
  const rq_schedule& sc = sp->schedule;
  if (sp->n_steps < 0 || sp->bootstrap < 0 || sc.poisson_interval < 1 || sc.norm_interval < 1)
    return fail(s, RQ_EINVAL, "bad adjoint spec");
  if (s->p.poisson_mode == RQ_POISSON_TWOGRID)
    return fail(s, RQ_EINVAL, "adjoint supports the Jacobi and multigrid Poisson modes");
  if (backward && s->p.poisson_mode == RQ_POISSON_MULTIGRID && s->p.mg_nu1 != s->p.mg_nu2)
    return fail(s, RQ_EINVAL, "adjoint of the V-cycle needs mg_nu1 == mg_nu2");
  if (s->p.rc_dirichlet || s->p.init_mode != RQ_INIT_S1 || s->p.energy_mode != RQ_ENERGY_S1)
    return fail(s, RQ_EINVAL, "adjoint implements the cloud/solver.py semantics (init/energy S1, no rc rule)");
  AdjState& A = s->adj;
  const int64_t Tn = sp->n_steps;
  // checkpoint interval; <= 0: the whole trajectory as one segment when its
  // 2 T + 1 fields (beyond those already held) fit in a quarter of the free
  // device memory -- the forward pass then records every step's inputs and
  // the reverse sweep recomputes nothing -- else 64
  int C = sp->checkpoint_interval;
  if (C <= 0) {
    C = 64;
    const int64_t T1 = std::max<int64_t>(Tn, 1);
    size_t free_b = 0, total_b = 0;
    if (backward && T1 <= (int64_t(1) << 20) && hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
      const size_t want = 2 * (size_t)T1 + 1, have = A.segU.size() + A.segP.size();
      const size_t extra = want > have ? want - have : 0;
      if (extra * s->S3 * sizeof(float) <= free_b / 4) C = (int)T1;
    }
  }
  const int64_t nck = backward ? (Tn + C - 1) / C + 1 : 0;
  int rc;
  for (float** p : {&A.t1, &A.t2, &A.t3})
    if ((rc = adj_alloc(s, p))) return rc;
  if (!A.red) {
    const size_t nr = 2 * (size_t)s->n_atoms + 3 * (size_t)s->n_atoms + 3 * (size_t)s->n_atoms + 3;
    HIPCHK(s, hipMalloc((void**)&A.red, sizeof(Fx) * kFxWords * nr));
  }
  if (backward) {
    for (float** p : {&A.lamU, &A.lamU2, &A.lamP, &A.lamP2, &A.gK, &A.acc})
      if ((rc = adj_alloc(s, p))) return rc;
    if ((rc = adj_alloc_vec(s, A.ckU, nck)) || (rc = adj_alloc_vec(s, A.ckP, nck))) return rc;
    const size_t segn = (size_t)std::min<int64_t>(C, std::max<int64_t>(Tn, 1));
    if ((rc = adj_alloc_vec(s, A.segU, segn + 1)) || (rc = adj_alloc_vec(s, A.segP, segn))) return rc;
  }
  const Tile& T = s->tile;
  // ---- forward ----
  if (saved) {
    // K(R) as in the trial pass (its trial amplitudes go to scratch); U and P
    // are the saved constants
    launch_init_frozen(s->st, s->S, s->h, s->h2, s->atoms_d, s->n_atoms, s->lab, A.t1, s->K);
    LAUNCHCHK(s);
    if ((rc = copy_field(s, s->U, A.wU)) || (rc = copy_field(s, s->P, A.wP))) return rc;
  } else {
    launch_init_frozen(s->st, s->S, s->h, s->h2, s->atoms_d, s->n_atoms, s->lab, s->U, s->K);
    LAUNCHCHK(s);
    if ((rc = normalize_on(s, s->U))) return rc;
    launch_fill(s->st, s->P, s->S3, 0.0f);
    for (int b = 0; b < sp->bootstrap; ++b) {
      if ((rc = poisson_into(s, s->P, s->P2, s->U, A.t1, A.t2))) return rc;
      std::swap(s->P, s->P2);
    }
  }
  if (backward) {
    if ((rc = copy_field(s, A.ckU[0], s->U)) || (rc = copy_field(s, A.ckP[0], s->P))) return rc;
  }
  // the last segment's inputs (U before each step, P of each Poisson step)
  // are recorded here, so the reverse sweep does not recompute that segment
  const int64_t nseg = backward ? (Tn + C - 1) / C : 0;
  const int64_t t0_last = nseg > 0 ? (nseg - 1) * C : Tn;
  for (int64_t t = 1; t <= Tn; ++t) {
    const bool rec = backward && t > t0_last;
    const int idx = rec ? (int)(t - t0_last - 1) : 0;
    if (rec && idx == 0 && (rc = copy_field(s, A.segU[0], s->U))) return rc;
    bool did_itp = false;
    if (t % sc.poisson_interval == 0) {
      if ((rc = poisson_into(s, s->P, s->P2, s->U, A.t1, A.t2, s->U2, &did_itp))) return rc;
      std::swap(s->P, s->P2);
      if (rec && (rc = copy_field(s, A.segP[idx], s->P))) return rc;
    }
    if (!did_itp) launch_itp(s->st, T, s->S, s->U, s->U2, s->K, s->P, s->lab, s->half_dv, s->dtv);
    LAUNCHCHK(s);
    std::swap(s->U, s->U2);
    if (t % sc.norm_interval == 0 && (rc = normalize_on(s, s->U))) return rc;
    if (rec && t < Tn && (rc = copy_field(s, A.segU[idx + 1], s->U))) return rc;
    if (backward && t % C == 0 && t < Tn) {
      if ((rc = copy_field(s, A.ckU[t / C], s->U)) || (rc = copy_field(s, A.ckP[t / C], s->P))) return rc;
    }
  }
  // energy of the final state
  {
    Fx* per = A.red + kFxWords * 5 * (size_t)s->n_atoms;
    Fx* dip = per + kFxWords * 3 * (size_t)s->n_atoms;
    zero_fx(s, per, 3 * (size_t)s->n_atoms + 3);
    launch_fill(s->st, reinterpret_cast<float*>(s->health + 1), 1, 0.0f);
    launch_energy(s->st, T, RQ_ENERGY_S1, s->S, s->h, s->h3, s->U, s->K, s->P, s->lab, s->atoms_d, s->n_atoms, per,
                  dip, s->health + 1);
    LAUNCHCHK(s);
    std::vector<double> hv(3 * (size_t)s->n_atoms);
    if ((rc = read_fx(s, per, hv.size(), hv.data()))) return rc;
    int bad = 0;
    if ((rc = read_flag(s, 1, &bad))) return rc;
    if (bad) return fail(s, RQ_EUNSTABLE, "non-finite energy in the frozen solve");
    double e[3] = {0, 0, 0};
    for (int d = 0; d < s->n_atoms; ++d)
      for (int q = 0; q < 3; ++q) e[q] += hv[d * 3 + q];
    if (E3)
      for (int q = 0; q < 3; ++q) E3[q] = e[q];
    if (!std::isfinite(e[0] + e[1] + e[2])) return fail(s, RQ_EUNSTABLE, "non-finite energy in the frozen solve");
  }
  if (!backward) return RQ_OK;
  // ---- reverse sweep ----
  launch_energy_grad(s->st, T, s->S, s->h, s->h3, s->U, s->K, s->P, s->lab, A.lamU, A.gK, A.lamP);
  LAUNCHCHK(s);
  std::vector<const float*> Pused(C);
  for (int64_t sg = nseg - 1; sg >= 0; --sg) {
    const int64_t t0 = sg * C, t1 = std::min<int64_t>(Tn, t0 + C);
    // the last segment was recorded by the forward pass
    const bool recompute = sg != nseg - 1;
    if (recompute && (rc = copy_field(s, A.segU[0], A.ckU[sg]))) return rc;
    const float* Pprev = A.ckP[sg];
    for (int64_t t = t0 + 1; t <= t1; ++t) {
      const int idx = (int)(t - t0 - 1);
      bool did_itp = false;
      if (t % sc.poisson_interval == 0) {
        if (recompute && (rc = poisson_into(s, Pprev, A.segP[idx], A.segU[idx], A.t1, A.t2,
                                            t < t1 ? A.segU[idx + 1] : nullptr, &did_itp)))
          return rc;
        Pused[idx] = A.segP[idx];
      } else {
        Pused[idx] = Pprev;
      }
      Pprev = Pused[idx];
      if (recompute && t < t1) {
        if (!did_itp)
          launch_itp(s->st, T, s->S, A.segU[idx], A.segU[idx + 1], s->K, Pused[idx], s->lab, s->half_dv, s->dtv);
        LAUNCHCHK(s);
        if (t % sc.norm_interval == 0 && (rc = normalize_on(s, A.segU[idx + 1]))) return rc;
      }
    }
    for (int64_t t = t1; t > t0; --t) {
      const int idx = (int)(t - t0 - 1);
      const float* Uin = A.segU[idx];
      if (t % sc.norm_interval == 0) {
        launch_itp(s->st, T, s->S, Uin, A.t1, s->K, Pused[idx], s->lab, s->half_dv, s->dtv);
        LAUNCHCHK(s);
        if ((rc = normalize_adj(s, A.t1))) return rc;
      }
      launch_itp_adj(s->st, T, s->S, A.lamU, A.lamU2, Uin, s->K, Pused[idx], s->lab, s->half_dv, s->dtv, A.gK,
                     A.lamP);
      LAUNCHCHK(s);
      std::swap(A.lamU, A.lamU2);
      if (t % sc.poisson_interval == 0 && (rc = poisson_adj(s, Uin))) return rc;
    }
  }
  if (saved) {
    // the start does not depend on R: only K(R) carries the gradient
    launch_fill(s->st, A.lamU, s->S3, 0.0f);
    LAUNCHCHK(s);
  } else {
    // bootstrap Poisson steps (P_in = 0 at the first one; its adjoint is dropped)
    for (int b = 0; b < sp->bootstrap; ++b)
      if ((rc = poisson_adj(s, A.ckU[0]))) return rc;
    // initial normalise: recompute the trial field into t1
    launch_init_frozen(s->st, s->S, s->h, s->h2, s->atoms_d, s->n_atoms, s->lab, A.t1, A.t2);
    LAUNCHCHK(s);
    if ((rc = normalize_adj(s, A.t1))) return rc;
  }
  Fx* gout = A.red + kFxWords * 2 * (size_t)s->n_atoms;
  zero_fx(s, gout, 3 * (size_t)s->n_atoms);
  launch_pos_grad(s->st, s->S, s->h, s->h2, A.gK, A.lamU, s->lab, s->atoms_d, s->n_atoms, gout);
  LAUNCHCHK(s);
  std::vector<double> g(3 * (size_t)s->n_atoms);
  if ((rc = read_fx(s, gout, g.size(), g.data()))) return rc;
  if (dEdR) std::memcpy(dEdR, g.data(), sizeof(double) * g.size());
  return RQ_OK;
}

}  // namespace

extern "C" {

int rq_abi_version(void) { return RQ_ABI_VERSION; }

const char* rq_last_error(const rq_solver* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

double rq_dv_for_electrons(int32_t n) {
  if (n > 500) return 0.005;
  if (n > 200) return 0.01;
  if (n > 100) return 0.02;
  if (n > 30) return 0.03;
  return 0.12;
}

void rq_default_params(rq_params* p) {
  std::memset(p, 0, sizeof *p);
  p->NN = 96;
  p->screen = 9.6;
  p->dv = 0.0;
  p->init_mode = RQ_INIT_S1;
  p->poisson_mode = RQ_POISSON_JACOBI;
  p->jacobi_sweeps = 2;
  p->smooth_a = 0.0f;
  p->smooth_c = 6.0f;
  p->mg_levels = 4;
  p->mg_nu1 = 2;
  p->mg_nu2 = 2;
  p->mg_a = 1.0f / 7.0f;
  p->mg_c = 7.0f;
  p->twogrid_interval = 50;
  p->twogrid_coarse_sweeps = 10;
  p->boundary_mode = RQ_BOUNDARY_SCORE;
  p->flip_ratio = 1.1f;
  p->flip_min_foreign = 2;
  p->ls_dt_w = 0.5f;
  p->ls_curv_reg = 0.15f;
  p->energy_mode = RQ_ENERGY_S1;
  p->rc_dirichlet = 0;
}

void rq_default_schedule(rq_schedule* s) {
  s->poisson_interval = 2;
  s->norm_interval = 20;
  s->boundary_interval = 100;
}

int rq_create(const rq_params* pp, const rq_atom* atoms, int32_t n_atoms, rq_solver** out) {
  if (!pp || !out || (!atoms && n_atoms > 0)) return fail(nullptr, RQ_EINVAL, "null argument");
  *out = nullptr;
  const rq_params& p = *pp;
  for (int r : p.reserved)
    if (r != 0) return fail(nullptr, RQ_EINVAL, "rq_params.reserved must be zero");
  if (p.NN < 8 || p.NN > 1024 || (p.NN & 1)) return fail(nullptr, RQ_EINVAL, "NN=%d must be even in [8, 1024]", p.NN);
  if (!(p.screen > 0.0)) return fail(nullptr, RQ_EINVAL, "screen must be > 0");
  if (n_atoms < 1 || n_atoms > kMaxDomains)
    return fail(nullptr, RQ_ERANGE, "n_atoms=%d outside [1, %d]", n_atoms, kMaxDomains);
  if (p.poisson_mode < 0 || p.poisson_mode > 2) return fail(nullptr, RQ_EINVAL, "bad poisson_mode");
  if (p.boundary_mode < 0 || p.boundary_mode > 2) return fail(nullptr, RQ_EINVAL, "bad boundary_mode");
  if (p.init_mode < 0 || p.init_mode > 1) return fail(nullptr, RQ_EINVAL, "bad init_mode");
  if (p.energy_mode < 0 || p.energy_mode > 1) return fail(nullptr, RQ_EINVAL, "bad energy_mode");
  if (p.pother_mode < 0 || p.pother_mode > 3) return fail(nullptr, RQ_EINVAL, "bad pother_mode");
  if (p.sic_interval < 0) return fail(nullptr, RQ_EINVAL, "negative sic_interval");
  if (p.sic_interval > 0 && p.pother_mode != RQ_POTHER_FRAME)
    return fail(nullptr, RQ_EINVAL, "sic_interval needs pother_mode FRAME");
  if (p.pother_mode == RQ_POTHER_FRAME && p.sic_interval > 0 && p.poisson_mode == RQ_POISSON_JACOBI)
    return fail(nullptr, RQ_EINVAL, "the self-interaction solve needs a coarse level (TWOGRID or MULTIGRID)");
  if (p.pother_mode == RQ_POTHER_DIRECT && p.poisson_mode != RQ_POISSON_MULTIGRID)
    return fail(nullptr, RQ_EINVAL, "pother_mode DIRECT needs RQ_POISSON_MULTIGRID");
  if (!(p.smooth_c != 0.f) || !(p.mg_c != 0.f)) return fail(nullptr, RQ_EINVAL, "smoother c must be non-zero");
  int levels = 0;
  if (p.poisson_mode == RQ_POISSON_MULTIGRID) {
    levels = p.mg_levels;
    if (levels < 2 || levels > 10) return fail(nullptr, RQ_EINVAL, "mg_levels=%d outside [2, 10]", levels);
    const int div = 1 << (levels - 1);
    if (p.NN % div) return fail(nullptr, RQ_EINVAL, "NN=%d not divisible by 2^(levels-1)=%d", p.NN, div);
    const int mc = p.NN / div - 1;
    if (mc < 1 || mc > kMaxCoarseInterior)
      return fail(nullptr, RQ_ERANGE, "coarsest interior %d outside [1, %d]; change mg_levels", mc,
                  kMaxCoarseInterior);
    if (p.mg_nu1 < 0 || p.mg_nu2 < 0) return fail(nullptr, RQ_EINVAL, "negative smoothing count");
  } else if (p.poisson_mode == RQ_POISSON_TWOGRID) {
    levels = 2;
    if (p.NN % 2 || p.NN / 2 < 3) return fail(nullptr, RQ_EINVAL, "two-grid needs NN/2 >= 3");
    if (p.twogrid_interval < 1 || p.twogrid_coarse_sweeps < 0) return fail(nullptr, RQ_EINVAL, "bad two-grid knobs");
  }
  if (p.jacobi_sweeps < 0) return fail(nullptr, RQ_EINVAL, "negative jacobi_sweeps");

  rq_solver* s = new rq_solver();
  s->p = p;
  if (const char* fm = std::getenv("RQ_FUSE_MAX_S")) s->fuse_max_s = std::atoi(fm);
  if (const char* fi = std::getenv("RQ_ITP_FUSE_MAX_S")) s->itp_fuse_max_s = std::atoi(fi);
  if (const char* s2 = std::getenv("RQ_SMOOTH2")) s->smooth2 = std::atoi(s2) != 0;
  if (const char* s2m = std::getenv("RQ_SMOOTH2_MIN_S")) s->smooth2_min_s = std::atoi(s2m);
  if (const char* e = std::getenv("RQ_SMOOTH2_RESRES")) s->s2rr = std::atoi(e) != 0;
  if (const char* e = std::getenv("RQ_SMOOTH2_RESRES_BLOCKS")) s->s2rr_blocks = std::atoi(e);
  if (const char* e = std::getenv("RQ_POST_FUSE")) s->post_fuse = std::atoi(e) != 0;
  if (const char* e = std::getenv("RQ_POST_BLOCKS")) s->post_blocks = std::atoi(e);
  s->NN = p.NN;
  s->S = p.NN + 1;
  s->S3 = (size_t)s->S * s->S * s->S;
  s->h64 = p.screen / p.NN;
  int n_elec = 0;
  s->atoms_h.resize(n_atoms);
  for (int a = 0; a < n_atoms; ++a) {
    s->atoms_h[a] = {atoms[a].i, atoms[a].j, atoms[a].k, atoms[a].Z, atoms[a].rc};
    if (atoms[a].Z > 0.f) ++n_elec;
  }
  s->lock_h.assign(n_atoms, 0);
  s->n_atoms = n_atoms;
  s->dv = p.dv > 0.0 ? p.dv : rq_dv_for_electrons(n_elec);
  s->h = (float)s->h64;
  s->h2 = (float)(s->h64 * s->h64);
  s->h3 = (float)std::pow(s->h64, 3.0);  // Python h ** 3 (cloud/solver.py:328)
  s->half_dv = (float)(0.5 * s->dv);
  s->dtv = (float)(s->dv * (s->h64 * s->h64));
  s->two_pi = (float)(2.0 * M_PI);
  s->tile = tile_for(s->S);
  s->ptile = tile_for(s->S, 4096);
  s->levels = levels;

  // device address of the allocation-free entry point's scratch array, once
  // per device, with the handle's device current
  if (hipError_t se = noalloc_setup(); se != hipSuccess) {
    delete s;
    return fail(nullptr, RQ_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(se));
  }
  auto alloc = [&](void** ptr, size_t bytes) -> bool { return hipMalloc(ptr, bytes) == hipSuccess; };
  bool ok = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) == hipSuccess;
  const size_t fb = s->S3 * sizeof(float), lb = s->S3 * sizeof(uint16_t);
  ok = ok && alloc((void**)&s->atoms_d, sizeof(Atom) * n_atoms) && alloc((void**)&s->Z_d, sizeof(float) * n_atoms) &&
       alloc((void**)&s->lock_d, sizeof(int32_t) * n_atoms) && alloc((void**)&s->ext_d, sizeof(AtomS2) * n_atoms);
  {
    rq_atom_s2 dflt;
    std::memset(&dflt, 0, sizeof dflt);
    dflt.split_axis[2] = 1.0f;  // molecule.js SPLIT_AXIS default [0, 0, 1]
    s->ext_h.assign(n_atoms, dflt);
  }
  {
    // cells of about 2 bohr (at least 4 nodes) for the score rule's search
    AtomCells& c = s->cells;
    c.cs = std::max(4, (int)std::lround(2.0 / (p.screen / p.NN)));
    c.n = (s->S + c.cs - 1) / c.cs;
    const size_t nc3 = (size_t)c.n * c.n * c.n;
    ok = ok && alloc((void**)&s->cell_start_d, sizeof(int32_t) * (nc3 + 1)) &&
         alloc((void**)&s->cell_atoms_d, sizeof(int32_t) * std::max(1, n_atoms));
  }
  {
    // The fine fields and both label maps live in one arena.  Consecutive
    // fields start `stagger` bytes apart modulo 2 MiB, so the same node of the
    // arrays a kernel streams together does not fall on the same HBM channel
    // (RQ_FIELD_STAGGER overrides the offset for tuning).
    size_t stagger = kFieldStagger;
    if (const char* fs = std::getenv("RQ_FIELD_STAGGER")) stagger = (size_t)std::atoll(fs);
    const size_t two_mb = size_t(2) << 20;
    const size_t sizes[9] = {fb, fb, fb, fb, fb, fb, fb, lb, lb};
    size_t offs[9], off = 0;
    for (int f = 0; f < 9; ++f) {
      offs[f] = off + (size_t)f * stagger % two_mb;
      off += (sizes[f] + two_mb - 1) / two_mb * two_mb + two_mb;
    }
    ok = ok && alloc(&s->arena, off);
    if (ok) {
      char* base = static_cast<char*>(s->arena);
      float** ff[7] = {&s->U, &s->U2, &s->K, &s->P, &s->P2, &s->bestU, &s->W};
      for (int f = 0; f < 7; ++f) *ff[f] = reinterpret_cast<float*>(base + offs[f]);
      s->lab = reinterpret_cast<uint16_t*>(base + offs[7]);
      s->lab2 = reinterpret_cast<uint16_t*>(base + offs[8]);
    }
  }
  s->red_len = (size_t)n_atoms * 3 + 3;
  ok = ok && alloc((void**)&s->red, sizeof(Fx) * kFxWords * s->red_len);
  ok = ok && alloc((void**)&s->health, 256);
  if (ok) launch_fill(s->st, reinterpret_cast<float*>(s->health), 1, 0.0f);
  if (ok && levels > 0) {
    s->Sl.assign(levels, 0);
    s->e0.assign(levels, nullptr);
    s->e1.assign(levels, nullptr);
    s->g.assign(levels, nullptr);
    s->Sl[0] = s->S;
    s->ctile.assign(levels, Tile());
    for (int l = 1; l < levels && ok; ++l) {
      s->Sl[l] = (s->NN >> l) + 1;
      s->ctile[l] = tile_for(s->Sl[l]);
      const size_t b = (size_t)s->Sl[l] * s->Sl[l] * s->Sl[l] * sizeof(float);
      ok = alloc((void**)&s->e0[l], b) && alloc((void**)&s->e1[l], b) && alloc((void**)&s->g[l], b);
      if (ok) {
        ok = hipMemsetAsync(s->e0[l], 0, b, s->st) == hipSuccess && hipMemsetAsync(s->e1[l], 0, b, s->st) == hipSuccess &&
             hipMemsetAsync(s->g[l], 0, b, s->st) == hipSuccess;
      }
    }
  }
  if (ok && levels > 0) {
    // orthonormal type-I DST matrix and 1/(l_a + l_b + l_c) of the coarsest level
    const int m = s->Sl[levels - 1] - 2;
    std::vector<double> Sm((size_t)m * m), rd((size_t)m * m * m), lam(m);
    const double nrm = std::sqrt(2.0 / (m + 1));
    for (int p = 0; p < m; ++p) {
      lam[p] = 2.0 - 2.0 * std::cos(M_PI * (p + 1) / (m + 1));
      for (int q = 0; q < m; ++q) Sm[(size_t)p * m + q] = nrm * std::sin(M_PI * (double)((p + 1) * (q + 1)) / (m + 1));
    }
    for (int a = 0; a < m; ++a)
      for (int b = 0; b < m; ++b)
        for (int c = 0; c < m; ++c) rd[((size_t)a * m + b) * m + c] = 1.0 / (lam[a] + lam[b] + lam[c]);
    ok = alloc((void**)&s->dst_S, sizeof(double) * Sm.size()) && alloc((void**)&s->dst_rden, sizeof(double) * rd.size());
    ok = ok && hipMemcpy(s->dst_S, Sm.data(), sizeof(double) * Sm.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(s->dst_rden, rd.data(), sizeof(double) * rd.size(), hipMemcpyHostToDevice) == hipSuccess;
  }
  if (ok && (p.pother_mode == RQ_POTHER_DIRECT || p.pother_mode == RQ_POTHER_DIRECT_SWEEPS)) {
    const double bytes = (double)(n_atoms + 1) * (double)s->S3 * sizeof(float);
    if (bytes > 64.0 * (1ull << 30)) {
      free_all(s);
      delete s;
      return fail(nullptr, RQ_ENOMEM, "pother_mode DIRECT needs %.1f GiB for %d per-domain potentials", bytes / (1ull << 30),
                  n_atoms);
    }
    s->Pdom.assign(n_atoms, nullptr);
    for (int a = 0; a < n_atoms && ok; ++a) {
      ok = alloc((void**)&s->Pdom[a], fb) && hipMemsetAsync(s->Pdom[a], 0, fb, s->st) == hipSuccess;
    }
    ok = ok && alloc((void**)&s->rhs_dom, fb) && hipMemsetAsync(s->rhs_dom, 0, fb, s->st) == hipSuccess &&
         alloc((void**)&s->Pdom_d, sizeof(float*) * n_atoms) &&
         hipMemcpy(s->Pdom_d, s->Pdom.data(), sizeof(float*) * n_atoms, hipMemcpyHostToDevice) == hipSuccess;
    // concurrent contexts: one per charged domain up to kDomStreams (RQ_DOM_STREAMS overrides; <= 1: sequential)
    int n_ctx = 0;
    for (int a = 0; a < n_atoms; ++a) n_ctx += atoms[a].Z > 0.0f;
    if (p.pother_mode == RQ_POTHER_DIRECT_SWEEPS) n_ctx = 0;  // sweeps run on the handle's stream
    n_ctx = std::min(n_ctx, kDomStreams);
    if (const char* ds = std::getenv("RQ_DOM_STREAMS")) n_ctx = std::min(n_ctx, std::max(0, std::atoi(ds)));
    if (ok && n_ctx > 1) {
      ok = hipEventCreateWithFlags(&s->dom_fork, hipEventDisableTiming) == hipSuccess;
      s->dom.resize(n_ctx);
      for (auto& d : s->dom) {
        ok = ok && hipStreamCreateWithFlags(&d.st, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&d.done, hipEventDisableTiming) == hipSuccess &&
             alloc((void**)&d.x2, fb) && hipMemsetAsync(d.x2, 0, fb, s->st) == hipSuccess &&
             alloc((void**)&d.rhs, fb) && hipMemsetAsync(d.rhs, 0, fb, s->st) == hipSuccess;
        d.e0.assign(levels, nullptr);
        d.e1.assign(levels, nullptr);
        d.g.assign(levels, nullptr);
        for (int l = 1; l < levels && ok; ++l) {
          const size_t b = (size_t)s->Sl[l] * s->Sl[l] * s->Sl[l] * sizeof(float);
          for (float** q : {&d.e0[l], &d.e1[l], &d.g[l]})
            ok = ok && alloc((void**)q, b) && hipMemsetAsync(*q, 0, b, s->st) == hipSuccess;
        }
      }
    }
  }
  if (ok && (p.pother_mode == RQ_POTHER_FRAME || p.pother_mode == RQ_POTHER_DIRECT_SWEEPS)) {
    ok = alloc((void**)&s->Pother, fb) && hipMemsetAsync(s->Pother, 0, fb, s->st) == hipSuccess;
    if (p.sic_interval > 0)
      for (float** q : {&s->sic_x, &s->sic_x2, &s->sic_g})
        ok = ok && alloc((void**)q, fb) && hipMemsetAsync(*q, 0, fb, s->st) == hipSuccess;
  }
  if (!ok) {
    (void)hipGetLastError();
    free_all(s);
    delete s;
    return fail(nullptr, RQ_ENOMEM, "device allocation failed for NN=%d (%zu nodes)", p.NN, (size_t)(p.NN + 1) * (p.NN + 1) * (p.NN + 1));
  }
  // zero every field so faces start as Dirichlet zeros
  const void* fields[] = {s->U, s->U2, s->K, s->P, s->P2, s->bestU, s->W};
  for (const void* f : fields) ok = ok && hipMemsetAsync((void*)f, 0, fb, s->st) == hipSuccess;
  ok = ok && hipMemsetAsync(s->lab, 0, lb, s->st) == hipSuccess && hipMemsetAsync(s->lab2, 0, lb, s->st) == hipSuccess;
  ok = ok && hipMemsetAsync(s->lock_d, 0, sizeof(int32_t) * n_atoms, s->st) == hipSuccess;
  if (!ok || upload_atoms(s) != RQ_OK) {
    std::string m = s->err;
    free_all(s);
    delete s;
    return fail(nullptr, RQ_EHIP, "device initialisation failed %s", m.c_str());
  }
  *out = s;
  return RQ_OK;
}

int rq_destroy(rq_solver* h) {
  if (!h) return RQ_OK;
  (void)hipStreamSynchronize(h->st);
  free_all(h);
  delete h;
  return RQ_OK;
}

int rq_sync(rq_solver* h) {
  if (int g_ = gate(h)) return g_;
  HIPCHK(h, hipStreamSynchronize(h->st));
  return RQ_OK;
}

int rq_get_info(const rq_solver* h, int32_t* NN, int32_t* S, int32_t* n_atoms, double* hg, double* dv) {
  if (!h) return RQ_EINVAL;
  if (NN) *NN = h->NN;
  if (S) *S = h->S;
  if (n_atoms) *n_atoms = h->n_atoms;
  if (hg) *hg = h->h64;
  if (dv) *dv = h->dv;
  return RQ_OK;
}

int rq_set_atoms(rq_solver* h, const rq_atom* atoms, int32_t n) {
  if (int g_ = gate(h)) return g_;
  if (!h || !atoms) return fail(h, RQ_EINVAL, "null argument");
  if (n != h->n_atoms) return fail(h, RQ_EINVAL, "n_atoms %d != %d", n, h->n_atoms);
  for (int a = 0; a < n; ++a) h->atoms_h[a] = {atoms[a].i, atoms[a].j, atoms[a].k, atoms[a].Z, atoms[a].rc};
  return upload_atoms(h);
}

int rq_get_atoms(const rq_solver* h, rq_atom* atoms, int32_t n) {
  if (!h || !atoms || n != h->n_atoms) return RQ_EINVAL;
  for (int a = 0; a < n; ++a) atoms[a] = {h->atoms_h[a].i, h->atoms_h[a].j, h->atoms_h[a].k, h->atoms_h[a].Z, h->atoms_h[a].rc};
  return RQ_OK;
}

int rq_set_atoms_s2(rq_solver* h, const rq_atom_s2* ext, int32_t n) {
  if (int g_ = gate(h)) return g_;
  if (!h || !ext) return fail(h, RQ_EINVAL, "null argument");
  if (n != h->n_atoms) return fail(h, RQ_EINVAL, "n_atoms %d != %d", n, h->n_atoms);
  for (int a = 0; a < n; ++a) {
    const rq_atom_s2& x = ext[a];
    if (x.split_type < 0 || x.split_type > 4)
      return fail(h, RQ_EINVAL, "atom %d: split_type %d (0/1 sphere, 2 hemi, 3 third, 4 none)", a, x.split_type);
    if (x.split_idx < 0 || x.split_idx > 2) return fail(h, RQ_EINVAL, "atom %d: split_idx %d", a, x.split_idx);
    if (!std::isfinite(x.Z_nuc) || !std::isfinite(x.init_zeff) || !std::isfinite(x.init_rcut))
      return fail(h, RQ_EINVAL, "atom %d: non-finite S2 field", a);
  }
  h->ext_h.assign(ext, ext + n);
  return upload_atoms(h);
}

int rq_set_lock_groups(rq_solver* h, const int32_t* group, int32_t n) {
  if (int g_ = gate(h)) return g_;
  if (!h || !group || n != h->n_atoms) return fail(h, RQ_EINVAL, "lock groups need n_atoms entries");
  h->lock_h.assign(group, group + n);
  HIPCHK(h, hipMemcpyAsync(h->lock_d, group, sizeof(int32_t) * n, hipMemcpyHostToDevice, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return RQ_OK;
}

int rq_init_trial(rq_solver* h) {
  if (int g_ = gate(h)) return g_;
  launch_init_trial(h->st, h->p.init_mode, h->S, h->h, h->h2, h->atoms_d, h->ext_d, h->n_atoms, h->U, h->U2, h->lab, h->K,
                    h->bestU, h->W);
  LAUNCHCHK(h);
  if (h->p.rc_dirichlet) {
    launch_rc_flags(h->st, h->S, h->h, h->atoms_d, h->n_atoms, h->lab);
    LAUNCHCHK(h);
  }
  HIPCHK(h, hipMemcpyAsync(h->lab2, h->lab, h->S3 * sizeof(uint16_t), hipMemcpyDeviceToDevice, h->st));
  launch_fill(h->st, reinterpret_cast<float*>(h->health), 1, 0.0f);
  HIPCHK(h, hipMemsetAsync(h->P, 0, h->S3 * sizeof(float), h->st));
  HIPCHK(h, hipMemsetAsync(h->P2, 0, h->S3 * sizeof(float), h->st));
  if (h->p.pother_mode == RQ_POTHER_DIRECT_SWEEPS) {
    // molecule.js:2678-2693: Pother cleared, then every domain with Z != 0
    // starts from the others' 0.5/r and adds it into Pother
    HIPCHK(h, hipMemsetAsync(h->Pother, 0, h->S3 * sizeof(float), h->st));
    for (int m = 0; m < h->n_atoms; ++m) {
      if (h->atoms_h[m].Z == 0.0f) continue;
      launch_pdirect_init(h->st, h->tile, h->S, h->h, h->h2, h->atoms_d, h->n_atoms, m,
                          h->p.init_mode == RQ_INIT_S2, h->Pdom[m], h->Pother);
    }
    LAUNCHCHK(h);
  }
  h->initialized = true;
  return RQ_OK;
}

int rq_initialize(rq_solver* h, int32_t bootstrap) {
  int rc = rq_init_trial(h);
  if (rc) return rc;
  rc = normalize_step(h);
  if (rc) return rc;
  for (int i = 0; i < bootstrap; ++i) {
    rc = poisson_step(h, false);
    if (rc) return rc;
  }
  // the frame potential starts as P (molecule.js:2091)
  if (h->p.pother_mode == RQ_POTHER_FRAME)
    HIPCHK(h, hipMemcpyAsync(h->Pother, h->P, h->S3 * sizeof(float), hipMemcpyDeviceToDevice, h->st));
  return RQ_OK;
}

int rq_poisson(rq_solver* h, int32_t n) {
  if (int g_ = gate(h)) return g_;
  for (int i = 0; i < n; ++i) {
    int rc = poisson_step(h, false);
    if (rc) return rc;
  }
  return RQ_OK;
}

int rq_update(rq_solver* h) {
  if (int g_ = gate(h)) return g_;
  return update_step(h);
}

int rq_normalize(rq_solver* h) {
  if (int g_ = gate(h)) return g_;
  return normalize_step(h);
}

int rq_boundary(rq_solver* h) {
  if (int g_ = gate(h)) return g_;
  return boundary_step(h);
}

int rq_recompute_k(rq_solver* h) {
  if (int g_ = gate(h)) return g_;
  launch_recompute_k(h->st, h->p.init_mode, h->S, h->h, h->h2, h->atoms_d, h->ext_d, h->n_atoms, h->K);
  LAUNCHCHK(h);
  if (h->p.rc_dirichlet) {
    launch_rc_flags(h->st, h->S, h->h, h->atoms_d, h->n_atoms, h->lab);
    LAUNCHCHK(h);
    HIPCHK(h, hipMemcpyAsync(h->lab2, h->lab, h->S3 * sizeof(uint16_t), hipMemcpyDeviceToDevice, h->st));
  }
  return RQ_OK;
}

int rq_twogrid_correct(rq_solver* h) {
  if (int g_ = gate(h)) return g_;
  if (h->p.poisson_mode != RQ_POISSON_TWOGRID) return fail(h, RQ_EINVAL, "two-grid correction needs RQ_POISSON_TWOGRID");
  return twogrid_correction(h);
}

static int eager_step(rq_solver* h, int64_t step, const rq_schedule* sc, bool with_boundary) {
  int rc;
  bool did_itp = false;
  if (step % sc->poisson_interval == 0) {
    const bool tg = h->p.poisson_mode == RQ_POISSON_TWOGRID && step % h->p.twogrid_interval == 0;
    // the fused last sweep + ITP reads the fresh P; the frame mode's ITP reads Pother
    if ((rc = poisson_step(h, tg, h->p.pother_mode != RQ_POTHER_FRAME, &did_itp))) return rc;
  }
  if (!did_itp && (rc = update_step(h))) return rc;
  if (step % sc->norm_interval == 0 && (rc = normalize_step(h))) return rc;
  if (with_boundary && h->p.boundary_mode != RQ_BOUNDARY_OFF && step % sc->boundary_interval == 0 &&
      (rc = boundary_step(h)))
    return rc;
  return RQ_OK;
}

static PtrState ptr_state(const rq_solver* h) { return PtrState{h->U, h->U2, h->P, h->P2, h->lab, h->e0, h->e1}; }

static void set_ptr_state(rq_solver* h, const PtrState& p) {
  h->U = p.U;
  h->U2 = p.U2;
  h->P = p.P;
  h->P2 = p.P2;
  h->lab = p.lab;
  h->e0 = p.e0;
  h->e1 = p.e1;
}

// Destroys every cached step graph.  A launched executable may still be
// running on the handle's stream (or be read by a tracing tool's completion
// callback), and HIP does not document that hipGraphExecDestroy waits for
// it, so the stream is drained first.
static void drop_graphs(rq_solver* h) {
  if (h->graphs.empty()) return;
  (void)hipStreamSynchronize(h->st);
  for (auto& ge : h->graphs) {
    (void)hipGraphExecDestroy(ge.exec);
    (void)hipGraphDestroy(ge.graph);
  }
  h->graphs.clear();
}

// One aligned chunk of norm_interval steps (no boundary rule inside) from a
// cached HIP graph, capturing it on first use for the current buffer state.
static int graph_chunk(rq_solver* h, int64_t c0, const rq_schedule* sc) {
  const PtrState cur = ptr_state(h);
  for (auto& ge : h->graphs) {
    if (ge.pi == sc->poisson_interval && ge.ni == sc->norm_interval && ge.before == cur) {
      HIPCHK(h, hipGraphLaunch(ge.exec, h->st));
      set_ptr_state(h, ge.after);
      return RQ_OK;
    }
  }
  if (h->graphs.size() >= 8) drop_graphs(h);
  GraphEntry ge;
  ge.before = cur;
  ge.pi = sc->poisson_interval;
  ge.ni = sc->norm_interval;
  HIPCHK(h, hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
  int rc = RQ_OK;
  for (int64_t step = c0 + 1; step <= c0 + sc->norm_interval && rc == RQ_OK; ++step) rc = eager_step(h, step, sc, false);
  hipGraph_t g = nullptr;
  hipError_t ee = hipStreamEndCapture(h->st, &g);
  if (rc != RQ_OK) {
    if (g) (void)hipGraphDestroy(g);
    set_ptr_state(h, cur);
    return rc;
  }
  if (ee != hipSuccess) {
    set_ptr_state(h, cur);
    return fail(h, RQ_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(ee));
  }
  ge.graph = g;
  HIPCHK(h, hipGraphInstantiate(&ge.exec, g, nullptr, nullptr, 0));
  ge.after = ptr_state(h);
  h->graphs.push_back(ge);
  HIPCHK(h, hipGraphLaunch(ge.exec, h->st));
  return RQ_OK;
}

int rq_step(rq_solver* h, int64_t first, int32_t n, const rq_schedule* sc) {
  if (int g_ = gate(h)) return g_;
  if (!h->initialized) return fail(h, RQ_ESTATE, "rq_step before rq_init_trial/rq_initialize");
  rq_schedule d;
  rq_default_schedule(&d);
  if (!sc) sc = &d;
  if (sc->poisson_interval < 1 || sc->norm_interval < 1 || sc->boundary_interval < 1)
    return fail(h, RQ_EINVAL, "schedule intervals must be >= 1");
  const bool graphable = h->use_graph && !h->timing && h->p.poisson_mode != RQ_POISSON_TWOGRID &&
                         sc->norm_interval % sc->poisson_interval == 0 &&
                         sc->boundary_interval % sc->norm_interval == 0;
  int64_t step = first + 1;
  const int64_t last = first + n;
  while (step <= last) {
    int rc;
    const int64_t c0 = step - 1;
    if (graphable && c0 % sc->norm_interval == 0 && c0 + sc->norm_interval <= last) {
      if ((rc = graph_chunk(h, c0, sc))) return rc;
      step = c0 + sc->norm_interval;
      if (h->p.boundary_mode != RQ_BOUNDARY_OFF && step % sc->boundary_interval == 0 && (rc = boundary_step(h)))
        return rc;
      ++step;
      continue;
    }
    if ((rc = eager_step(h, step, sc, true))) return rc;
    ++step;
  }
  // a normalisation ran: its sums flag amplitudes that are not finite
  // (molecule.js:3198-3201 stops on a non-finite energy the same way)
  if (last / sc->norm_interval > first / sc->norm_interval) {
    int bad = 0;
    if (int rc = read_flag(h, 0, &bad)) return rc;
    if (bad)
      return fail(h, RQ_EUNSTABLE, "Numerical instability: non-finite amplitude at a normalisation in steps %lld..%lld",
                  (long long)first + 1, (long long)last);
  }
  return RQ_OK;
}

int rq_adjoint_save_start(rq_solver* h) {
  if (int g_ = gate(h)) return g_;
  if (!h->initialized) return fail(h, RQ_ESTATE, "rq_adjoint_save_start before the fields are set");
  if (int rc = adj_alloc(h, &h->adj.wU)) return rc;
  if (int rc = adj_alloc(h, &h->adj.wP)) return rc;
  if (int rc = copy_field(h, h->adj.wU, h->U)) return rc;
  if (int rc = copy_field(h, h->adj.wP, h->P)) return rc;
  HIPCHK(h, hipStreamSynchronize(h->st));
  return RQ_OK;
}

int rq_debug_fail_after(rq_solver* h, int32_t n_checks) {
  if (int g_ = gate(h)) return g_;
  if (n_checks < 0) return fail(h, RQ_EINVAL, "n_checks %d < 0", n_checks);
  h->fail_after = n_checks;
  return RQ_OK;
}

int rq_set_fused(rq_solver* h, int32_t on) {
  if (int g_ = gate(h)) return g_;
  h->fused = on != 0;
  h->bottom = on == 3;
  h->legs = on != 2;
  drop_graphs(h);
  return RQ_OK;
}

int rq_set_graph(rq_solver* h, int32_t on) {
  if (int g_ = gate(h)) return g_;
  h->use_graph = on != 0;
  return RQ_OK;
}

int rq_set_tile(rq_solver* h, int32_t level, int32_t bx, int32_t by, int32_t ich) {
  if (int g_ = gate(h)) return g_;
  const int nl = std::max(1, h->levels);
  if (level < -1 || level >= nl) return fail(h, RQ_ERANGE, "level %d outside [-1, %d)", level, nl);
  const bool dflt = bx <= 0;
  if (!dflt && (by < 1 || ich < 1 || bx * by > 256 || (bx * by) % 64 != 0))
    return fail(h, RQ_EINVAL, "tile %dx%d (ich %d): need bx*by a multiple of 64 and <= 256", bx, by, ich);
  for (int l = 0; l < nl; ++l) {
    if (level != -1 && level != l) continue;
    const int S = (l == 0) ? h->S : h->Sl[l];
    Tile t = tile_for(S);
    Tile pt = tile_for(S, 4096);
    if (!dflt) {
      t.bx = bx;
      t.by = by;
      t.ich = ich;
      pt = t;
    }
    if (l == 0) {
      h->tile = t;
      h->ptile = pt;
    } else {
      h->ctile[l] = t;
    }
  }
  drop_graphs(h);
  return RQ_OK;
}

int rq_get_tile(const rq_solver* h, int32_t level, int32_t* bx, int32_t* by, int32_t* ich) {
  if (!h) return RQ_EINVAL;
  const int nl = std::max(1, h->levels);
  if (level < 0 || level >= nl) return RQ_ERANGE;
  const Tile& t = (level == 0) ? h->tile : h->ctile[level];
  if (bx) *bx = t.bx;
  if (by) *by = t.by;
  if (ich) *ich = t.ich;
  return RQ_OK;
}

int rq_energy(rq_solver* h, double* out6, double* per_domain) {
  if (int g_ = gate(h)) return g_;
  if (!h || !out6) return fail(h, RQ_EINVAL, "null argument");
  const int nd = h->n_atoms;
  Fx* per = h->red;
  Fx* dip = h->red + kFxWords * (size_t)nd * 3;
  zero_fx(h, h->red, h->red_len);
  launch_fill(h->st, reinterpret_cast<float*>(h->health + 1), 1, 0.0f);
  {
    Timed t(h, F_ENERGY);
    const float* Pv = itp_potential(h);
    launch_energy(h->st, h->tile, h->p.energy_mode, h->S, h->h, h->h3, h->U, h->K, Pv, h->lab, h->atoms_d, nd, per,
                  dip, h->health + 1);
  }
  LAUNCHCHK(h);
  std::vector<double> hostv(h->red_len);
  if (int rc = read_fx(h, h->red, h->red_len, hostv.data())) return rc;
  int bad = 0;
  if (int rc = read_flag(h, 1, &bad)) return rc;
  if (bad) return fail(h, RQ_EUNSTABLE, "non-finite energy terms (a NaN or Inf amplitude or potential)");
  double T = 0, K = 0, E = 0;
  for (int d = 0; d < nd; ++d) {
    T += hostv[d * 3 + 0];
    K += hostv[d * 3 + 1];
    E += hostv[d * 3 + 2];
  }
  out6[0] = T;
  out6[1] = K;
  out6[2] = E;
  out6[3] = hostv[nd * 3 + 0];
  out6[4] = hostv[nd * 3 + 1];
  out6[5] = hostv[nd * 3 + 2];
  if (per_domain) std::memcpy(per_domain, hostv.data(), sizeof(double) * nd * 3);
  if (!std::isfinite(T + K + E)) return fail(h, RQ_EUNSTABLE, "non-finite energy (T=%g VeK=%g Vee=%g)", T, K, E);
  return RQ_OK;
}

int rq_domain_norms(rq_solver* h, double* norms) {
  if (int g_ = gate(h)) return g_;
  if (!h || !norms) return fail(h, RQ_EINVAL, "null argument");
  zero_fx(h, h->red, h->n_atoms);
  launch_norm_sums(h->st, h->tile, h->S, h->h3, h->U, h->lab, h->n_atoms, h->red, h->health);
  LAUNCHCHK(h);
  return read_fx(h, h->red, h->n_atoms, norms);
}

int rq_forces_hf(rq_solver* h, int32_t first, int32_t count, double* out3) {
  if (int g_ = gate(h)) return g_;
  if (!h || !out3) return fail(h, RQ_EINVAL, "null argument");
  if (first < 0 || count < 0 || first + count > h->n_atoms) return fail(h, RQ_ERANGE, "atom range out of bounds");
  if (count == 0) return RQ_OK;
  Fx* buf = nullptr;
  HIPCHK(h, hipMallocAsync((void**)&buf, sizeof(Fx) * kFxWords * count * 3, h->st));
  zero_fx(h, buf, (size_t)count * 3);
  {
    Timed t(h, F_FORCE);
    launch_forces_hf(h->st, h->S, h->h, h->h2, h->h3, h->U, h->atoms_d, first, count, buf);
  }
  LAUNCHCHK(h);
  const int rc = read_fx(h, buf, (size_t)count * 3, out3);
  HIPCHK(h, hipFreeAsync(buf, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return rc;
}

int rq_forces_hf_noalloc(rq_solver* h, int32_t first, int32_t count, double* out3) {
  if (int g_ = gate(h)) return g_;
  if (!h || !out3) return fail(h, RQ_EINVAL, "null argument");
  if (first < 0 || count < 0 || first + count > h->n_atoms) return fail(h, RQ_ERANGE, "atom range out of bounds");
  if (count > kNoAllocMaxAtoms)
    return fail(h, RQ_ERANGE, "rq_forces_hf_noalloc: count %d above its scratch limit of %d atoms per call", count,
                kNoAllocMaxAtoms);
  if (count == 0) return RQ_OK;
  unsigned char* base = static_cast<unsigned char*>(noalloc_scratch());
  if (!base) return fail(h, RQ_ESTATE, "rq_forces_hf_noalloc: no scratch address for the current device");
  Fx* sums = reinterpret_cast<Fx*>(base);
  double* vals = reinterpret_cast<double*>(base + kNoAllocFxBytes);
  // the array keeps the previous call's contents: clear the sums first
  zero_fx(h, sums, (size_t)count * 3);
  {
    Timed t(h, F_FORCE);
    launch_forces_hf(h->st, h->S, h->h, h->h2, h->h3, h->U, h->atoms_d, first, count, sums);
  }
  launch_fx_values(h->st, count * 3, sums, vals);
  LAUNCHCHK(h);
  HIPCHK(h, hipMemcpyAsync(out3, vals, sizeof(double) * 3 * count, hipMemcpyDeviceToHost, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return RQ_OK;
}

int rq_forces_gradp(rq_solver* h, int32_t radius, double* out3n) {
  if (int g_ = gate(h)) return g_;
  if (!h || !out3n) return fail(h, RQ_EINVAL, "null argument");
  int R = radius > 0 ? radius : std::max(5, (int)std::lround(3.0 / h->h64));
  if (R > h->NN) return fail(h, RQ_ERANGE, "radius %d beyond grid", R);
  double* buf = nullptr;
  HIPCHK(h, hipMallocAsync((void**)&buf, sizeof(double) * h->n_atoms * 3, h->st));
  {
    Timed t(h, F_FORCE);
    launch_forces_gradp(h->st, h->S, h->h, gradp_potential(h), h->atoms_d, h->ext_d, h->n_atoms, R, buf);
  }
  LAUNCHCHK(h);
  HIPCHK(h, hipMemcpyAsync(out3n, buf, sizeof(double) * h->n_atoms * 3, hipMemcpyDeviceToHost, h->st));
  HIPCHK(h, hipFreeAsync(buf, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return RQ_OK;
}

int rq_read_field(rq_solver* h, int32_t f, void* host, size_t bytes) {
  if (int g_ = gate(h)) return g_;
  if (!h || !host) return fail(h, RQ_EINVAL, "null argument");
  const size_t need = h->S3 * 4;
  if (bytes != need) return fail(h, RQ_EINVAL, "field %d needs %zu bytes, got %zu", f, need, bytes);
  if (f == RQ_FIELD_LABEL || f == RQ_FIELD_RHO) {
    void* tmp = nullptr;
    HIPCHK(h, hipMallocAsync(&tmp, need, h->st));
    if (f == RQ_FIELD_LABEL) launch_label_u16_to_i32(h->st, h->S3, h->lab, (int32_t*)tmp);
    else launch_square(h->st, h->S, h->U, (float*)tmp);
    LAUNCHCHK(h);
    HIPCHK(h, hipMemcpyAsync(host, tmp, need, hipMemcpyDeviceToHost, h->st));
    HIPCHK(h, hipFreeAsync(tmp, h->st));
  } else {
    float* p = field_ptr(h, f);
    if (!p) return fail(h, RQ_EINVAL, "unknown field %d", f);
    HIPCHK(h, hipMemcpyAsync(host, p, need, hipMemcpyDeviceToHost, h->st));
  }
  HIPCHK(h, hipStreamSynchronize(h->st));
  return RQ_OK;
}

// Self-potential of domain m (molecule.js:3018-3069): from zero, 4 sweeps of
// the S2 smoother, one two-grid correction (restricted residual with the
// inverted sign, 10 coarse sweeps from zero, half the prolongation added),
// 4 sweeps; the result is in s->sic_x.
static int sic_solve(rq_solver* s, int m) {
  const float a = s->p.smooth_a, c = s->p.smooth_c;
  launch_rho_self(s->st, s->tile, s->S, s->U, s->lab, m, s->two_pi, s->h2, s->sic_g);
  launch_zero(s->st, s->sic_x, s->S3);
  for (int it = 0; it < 4; ++it) {
    launch_smooth_coarse(s->st, s->tile, s->S, s->sic_x, s->sic_x2, s->sic_g, a, c);
    std::swap(s->sic_x, s->sic_x2);
  }
  const int S1 = s->Sl[1];
  launch_residual_coarse(s->st, s->tile, s->S, s->sic_x, s->sic_g, s->sic_x2);
  launch_restrict4(s->st, s->S, s->sic_x2, s->g[1], -1.0f);
  launch_zero(s->st, s->e0[1], (size_t)S1 * S1 * S1);
  for (int it = 0; it < 10; ++it) {
    launch_smooth_coarse(s->st, s->ctile[1], S1, s->e0[1], s->e1[1], s->g[1], a, c);
    std::swap(s->e0[1], s->e1[1]);
  }
  launch_prolong_correct(s->st, s->ptile, s->S, s->sic_x, s->e0[1], 0.5f);
  for (int it = 0; it < 4; ++it) {
    launch_smooth_coarse(s->st, s->tile, s->S, s->sic_x, s->sic_x2, s->sic_g, a, c);
    std::swap(s->sic_x, s->sic_x2);
  }
  LAUNCHCHK(s);
  return RQ_OK;
}

int rq_pother_frame(rq_solver* h, int64_t frame) {
  if (int g_ = gate(h)) return g_;
  if (h->p.pother_mode != RQ_POTHER_FRAME) return fail(h, RQ_ESTATE, "rq_pother_frame needs pother_mode FRAME");
  if (!h->initialized) return fail(h, RQ_ESTATE, "rq_pother_frame before initialisation");
  HIPCHK(h, hipMemcpyAsync(h->Pother, h->P, h->S3 * sizeof(float), hipMemcpyDeviceToDevice, h->st));
  if (h->p.sic_interval > 0 && frame > 0 && frame % h->p.sic_interval == 0) {
    for (int m = 0; m < h->n_atoms; ++m) {
      const float Z = h->atoms_h[m].Z;
      if (Z == 0.0f) continue;
      if (int rc = sic_solve(h, m)) return rc;
      const float frac = (Z > 1.0f) ? 1.0f / Z : 1.0f;
      launch_sub_pself(h->st, h->tile, h->S, h->sic_x, h->lab, m, frac, h->Pother);
      LAUNCHCHK(h);
    }
  }
  HIPCHK(h, hipStreamSynchronize(h->st));
  return RQ_OK;
}

int rq_display_slice(rq_solver* h, int32_t sk, float* host, size_t n) {
  if (int g_ = gate(h)) return g_;
  if (!h || !host) return fail(h, RQ_EINVAL, "null argument");
  const size_t S = h->S, need = 3 * S * S + 2 * S;
  if (n != need) return fail(h, RQ_EINVAL, "display slice needs %zu floats, got %zu", need, n);
  if (sk < 0 || sk > h->NN) return fail(h, RQ_ERANGE, "slice_k %d outside [0, %d]", sk, h->NN);
  float* tmp = nullptr;
  HIPCHK(h, hipMallocAsync((void**)&tmp, need * sizeof(float), h->st));
  launch_display_slice(h->st, h->S, sk, h->U, h->lab, h->K, h->atoms_d, h->n_atoms, tmp);
  LAUNCHCHK(h);
  HIPCHK(h, hipMemcpyAsync(host, tmp, need * sizeof(float), hipMemcpyDeviceToHost, h->st));
  HIPCHK(h, hipFreeAsync(tmp, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return RQ_OK;
}

int rq_write_field(rq_solver* h, int32_t f, const void* host, size_t bytes) {
  if (int g_ = gate(h)) return g_;
  if (!h || !host) return fail(h, RQ_EINVAL, "null argument");
  const size_t need = h->S3 * 4;
  if (bytes != need) return fail(h, RQ_EINVAL, "field %d needs %zu bytes, got %zu", f, need, bytes);
  if (f == RQ_FIELD_RHO) return fail(h, RQ_EINVAL, "rho is read-only (computed from U)");
  if (f == RQ_FIELD_LABEL) {
    void* tmp = nullptr;
    HIPCHK(h, hipMallocAsync(&tmp, need, h->st));
    HIPCHK(h, hipMemcpyAsync(tmp, host, need, hipMemcpyHostToDevice, h->st));
    launch_label_i32_to_u16(h->st, h->S3, (const int32_t*)tmp, h->lab);
    LAUNCHCHK(h);
    if (h->p.rc_dirichlet) {
      launch_rc_flags(h->st, h->S, h->h, h->atoms_d, h->n_atoms, h->lab);
      LAUNCHCHK(h);
    }
    HIPCHK(h, hipMemcpyAsync(h->lab2, h->lab, h->S3 * sizeof(uint16_t), hipMemcpyDeviceToDevice, h->st));
    HIPCHK(h, hipFreeAsync(tmp, h->st));
  } else {
    float* p = field_ptr(h, f);
    if (!p) return fail(h, RQ_EINVAL, "unknown field %d", f);
    HIPCHK(h, hipMemcpyAsync(p, host, need, hipMemcpyHostToDevice, h->st));
    // the ping-pong partner takes the same face values (kernels never write faces)
    float* partner = (f == RQ_FIELD_U) ? h->U2 : (f == RQ_FIELD_P) ? h->P2 : nullptr;
    if (partner) HIPCHK(h, hipMemcpyAsync(partner, p, need, hipMemcpyDeviceToDevice, h->st));
  }
  launch_fill(h->st, reinterpret_cast<float*>(h->health), 1, 0.0f);
  LAUNCHCHK(h);
  HIPCHK(h, hipStreamSynchronize(h->st));
  h->initialized = true;
  return RQ_OK;
}

int rq_swap_u(rq_solver* h) {
  if (int g_ = gate(h)) return g_;
  std::swap(h->U, h->U2);
  return RQ_OK;
}

int rq_field_device_ptr(rq_solver* h, int32_t f, void** dptr) {
  if (int g_ = gate(h)) return g_;
  if (!h || !dptr) return fail(h, RQ_EINVAL, "null argument");
  if (f == RQ_FIELD_LABEL) {
    *dptr = h->lab;
    return RQ_OK;
  }
  float* p = field_ptr(h, f);
  if (!p) return fail(h, RQ_EINVAL, "field %d has no device buffer", f);
  *dptr = p;
  return RQ_OK;
}

int rq_stream(rq_solver* h, void** s) {
  if (!h || !s) return fail(h, RQ_EINVAL, "null argument");
  *s = (void*)h->st;
  return RQ_OK;
}

int rq_solve_frozen(rq_solver* h, const rq_adjoint_spec* spec, double* energy3) {
  if (int g_ = gate(h)) return g_;
  return frozen_run(h, spec, false, energy3, nullptr);
}

int rq_adjoint_gradient(rq_solver* h, const rq_adjoint_spec* spec, double* energy3, double* dEdR) {
  if (int g_ = gate(h)) return g_;
  return frozen_run(h, spec, true, energy3, dEdR);
}

int rq_timing_enable(rq_solver* h, int32_t on) {
  if (int g_ = gate(h)) return g_;
  resolve_events(h);
  h->timing = on != 0;
  for (int f = 0; f < F_COUNT; ++f) {
    h->fam_ms[f] = 0;
    h->fam_n[f] = 0;
  }
  return RQ_OK;
}

int rq_timing_read(rq_solver* h, double* ms, int64_t* launches, int32_t n) {
  if (int g_ = gate(h)) return g_;
  resolve_events(h);
  for (int f = 0; f < n && f < F_COUNT; ++f) {
    if (ms) ms[f] = h->fam_ms[f];
    if (launches) launches[f] = h->fam_n[f];
  }
  return RQ_OK;
}

}  // extern "C"
