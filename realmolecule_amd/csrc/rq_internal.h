// Internal declarations shared by the kernel and API translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rq.h"

namespace rq {

// Label storage: uint16 per node; bit 15 marks "inside the pseudopotential
// radius rc of the own domain's nucleus" (molecule.js isInsideRc rule), the
// low 15 bits are the domain index.
constexpr uint16_t kRcFlag = 0x8000u;
constexpr uint16_t kLabelMask = 0x7fffu;
constexpr int kMaxDomains = 4096;       // LDS bins of the per-domain reductions
constexpr int kMaxCoarseInterior = 25;  // exact FP64 solve held in LDS (m^3 doubles)
constexpr int kMaxCoarsePP = 19;        // table-driven exact solve: 2 m^3 doubles of LDS, matrix row in registers
constexpr int kFuseMaxS = 65;           // one-pass prolongation+sweep up to 65^3 nodes
constexpr size_t kFieldStagger = 4096;   // byte offset between consecutive fine fields (mod 2 MiB)
constexpr int kScoreScanMaxAtoms = 128;  // score rule: scan all atoms up to this count, cell search above
constexpr int kLegsMaxS = 65;           // one down-leg + one up-leg launch per coarse level up to 65^3 nodes
constexpr int kResResMarchCJ = 2;       // coarse j-columns per residual+restriction march workgroup (128 CJ threads)
constexpr int kResResMarchCP = 2;       // coarse planes restricted per barrier round of that march
constexpr int kSmooth2Blocks = 8192;    // target workgroup count of the two-sweep march
constexpr int kSmooth2ResresBlocks = 2048;  // target workgroup count of the pre-smoothing + restriction march
constexpr int kPostBlocks = 8192;  // target workgroup count of the correction + post-sweeps + ITP march
constexpr int kSmooth2MinS = 320;       // smallest level (nodes per axis) that runs sweep pairs as that march
constexpr int kDirectSweeps = 10;       // damped sweeps per step of each direct potential (molecule.js:2835 JACOBI_DIRECT)
constexpr int kDomStreams = 4;          // concurrent per-domain V-cycles of the direct potential (one stream each)
constexpr int kReduceIch = 24;          // planes per thread of the per-domain reductions (fewer global atomics)
// largest fine grid (nodes per axis) whose last post-sweep carries the ITP
// update: fused while the fields fit the 256 MB MALL (193^3: 28.5 us against
// 14.4 + 22.8 separate); from HBM the two streaming passes win (257^3:
// 0.379 against 0.385 ms per step, 513^3: 2.96 against 3.08)
constexpr int kItpFuseMaxS = 200;
constexpr int kResResTiledMaxNc = 63;   // 3D-tiled residual+restriction up to 63 coarse interior nodes, march above
constexpr int kBottomMaxS = 25;         // last two V-cycle levels in one LDS-resident workgroup up to 25^3 nodes

// Order-independent FP64 sums.  A sum is held as two signed 64-bit words in
// fixed point: w[0] counts units of 2^-30 and w[1] units of 2^-70.  A
// contribution v (|v| < 2^32) enters as w[0] += rint(v 2^30) and w[1] +=
// rint((v 2^30 - rint(v 2^30)) 2^40), both parts exact in FP64, so each
// contribution is represented to within 2^-71.  Integer additions are exact
// and associative: however the workgroups' and waves' atomic additions
// interleave, the words -- and the value read from them -- are the same, so
// every per-domain reduction is bit-reproducible run to run.  The low word
// takes up to 2^24 contributions per sum before it could overflow.
typedef unsigned long long Fx;
constexpr int kFxWords = 2;
__host__ __device__ inline double fx_value(const Fx* w) {
  return (double)(long long)w[0] * 0x1p-30 + (double)(long long)w[1] * 0x1p-70;
}
#ifdef __HIP__
// Adds v to the fixed-point sum w; w may live in LDS or in global memory.
__device__ __forceinline__ void fx_add(Fx* w, double v) {
  const double s = v * 0x1p30;
  const double q = rint(s);
  atomicAdd(w, (Fx)(long long)q);
  atomicAdd(w + 1, (Fx)(long long)rint((s - q) * 0x1p40));
}
// Adds the words of one fixed-point sum to another (exact).
__device__ __forceinline__ void fx_add_words(Fx* w, const Fx* src) {
  const Fx a = src[0], b = src[1];
  if (a | b) {
    atomicAdd(w, a);
    atomicAdd(w + 1, b);
  }
}
#endif

// Launch shape of the 2.5D stencil kernels: (bx, by) tile in (k, j), marching
// ich planes along i.
struct Tile {
  int bx = 64, by = 4, ich = 8;
};

struct Atom {
  float i, j, k, Z, rc;
};

// molecule.js per-atom fields of the S2 semantics (rq_atom_s2), defaults
// resolved on the host: Zn = Z_nuc > 0 ? Z_nuc : Z, zeff = init_zeff > 0 ?
// init_zeff : Z, rcut = init_rcut > 0 ? init_rcut : 1e6.
struct AtomS2 {
  float Zn, zeff, rcut;
  int st, sidx;
  float ax, ay, az, rot;
};

// Uniform cell list over the grid for nearest-atom searches.
struct AtomCells {
  int cs = 4;          // cell edge in nodes
  int n = 1;           // cells per axis
  float zmax = 0.0f;   // largest Z of the binned atoms
};

// Kernel families timed separately by rq_timing_*: fine-grid (level 0)
// kernels are split from the coarse-level ones so each can be priced against
// its own algorithmic bytes.
enum Family {
  F_ITP = 0, F_SMOOTH, F_SMOOTH_C, F_RESTRICT, F_RESTRICT_C, F_PROLONG, F_PROLONG_C, F_COARSE, F_NORM,
  F_BOUNDARY, F_ENERGY, F_FORCE, F_SMOOTH_ITP, F_BOTTOM, F_SMOOTH_RESTRICT, F_POST_ITP, F_COUNT
};

// ---- launchers (rq_kernels.hip) ----
void launch_init_trial(hipStream_t st, int mode, int S, float h, float h2, const Atom* atoms, const AtomS2* ext,
                       int n_atoms, float* U, float* U2, uint16_t* lab, float* K, float* bestU, float* W);
void launch_recompute_k(hipStream_t st, int mode, int S, float h, float h2, const Atom* atoms, const AtomS2* ext,
                        int n_atoms, float* K);
void launch_rc_flags(hipStream_t st, int S, float h, const Atom* atoms, int n_atoms, uint16_t* lab);
void launch_itp(hipStream_t st, const Tile& t, int S, const float* U, float* Un, const float* K, const float* P,
                const uint16_t* lab, float half_dv, float dtv);
void launch_smooth_fine(hipStream_t st, const Tile& t, int S, const float* P, float* Pn, const float* U, float a,
                        float c, float alpha, float beta);
void launch_smooth_coarse(hipStream_t st, const Tile& t, int S, const float* E, float* En, const float* g, float a,
                          float c);
void launch_residual_fine(hipStream_t st, const Tile& t, int S, const float* P, const float* U, float alpha,
                          float beta, float* r);
void launch_residual_coarse(hipStream_t st, const Tile& t, int S, const float* E, const float* g, float* r);
void launch_restrict4(hipStream_t st, int Sf, const float* r, float* gc, float sign);
void launch_prolong_correct(hipStream_t st, const Tile& t, int Sf, float* x, const float* ec, float scale);
// x = 0 + prolongation of ec on the interior (x's faces untouched): the
// correction into a zero-filled x without the fill or the read of x
void launch_prolong_set(hipStream_t st, const Tile& t, int Sf, float* x, const float* ec);
// Smat (m x m DST matrix) and rden (m^3 reciprocal eigenvalue sums) may be
// null, which selects the line-transform variant that computes them itself.
void launch_coarse_exact(hipStream_t st, int Sc, const float* g, float* e, const double* Smat, const double* rden);
// Levels L (S nodes per axis) and L+1 of a V-cycle in one workgroup: nu1 / nu2
// sweeps around the exact DST solve (tables of level L+1); e gets the level-L
// correction.  bottom_fits(S): S is odd, <= kBottomMaxS and the LDS fits.
bool bottom_fits(int S);
// One launch per coarse level and direction: down leg (two pre-sweeps from a
// zero guess, residual, 4 FW restriction into gc; e_out gets the iterate) and
// up leg (prolongation + correction and nu post-sweeps, nu in 1..3: returns
// false for other nu without launching).
void launch_down2(hipStream_t st, int S, const float* g, float* e_out, float* gc, float a, float c);
// display slice (3 S^2 + 2 S floats) at k = sk
void launch_display_slice(hipStream_t st, int S, int sk, const float* U, const uint16_t* lab, const float* K,
                          const Atom* atoms, int n_atoms, float* out);
// direct per-domain potential: masked right-hand side and the label gather
// molecule.js direct Pother (RQ_POTHER_DIRECT_SWEEPS): the start of P_m
// (also added into Pother) and the copy of P_m onto the nodes labelled m
void launch_pdirect_init(hipStream_t st, const Tile& t, int S, float h, float h2, const Atom* atoms, int n_atoms,
                         int m, int floor_pos, float* Pm, float* Pother);
void launch_copy_for_label(hipStream_t st, const Tile& t, int S, const float* Pm, const uint16_t* lab, int m,
                           float* Pother);
void launch_rho_other(hipStream_t st, const Tile& t, int S, const float* U, const uint16_t* lab, int m, float alpha,
                      float beta, float* rhs);
void launch_rho_self(hipStream_t st, const Tile& t, int S, const float* U, const uint16_t* lab, int m, float alpha,
                     float beta, float* rhs);
void launch_sub_pself(hipStream_t st, const Tile& t, int S, const float* Pself, const uint16_t* lab, int m, float frac,
                      float* Pother);
void launch_pother_gather(hipStream_t st, const Tile& t, int S, const uint16_t* lab, const float* const* Pdom,
                          int n_dom, float* P);
// k_pother_gather then the ITP update on the gathered P, in one pass
void launch_itp_gather(hipStream_t st, const Tile& t, int S, const float* U, float* Un, const float* K,
                       const uint16_t* lab, const float* const* Pdom, int n_dom, float* P, float half_dv, float dtv);
bool launch_up(hipStream_t st, int S, int nu, const float* x, const float* e, float* x_out, const float* g, float a,
               float c);
void launch_bottom(hipStream_t st, int S, int nu1, int nu2, float a, float c, const float* g, float* e,
                   const double* Smat, const double* rden);
// fused multigrid passes (bit-identical to the unfused sequences)
void launch_smooth2_from_zero(hipStream_t st, const Tile& t, int S, const float* g, float* e_out, float a, float c);
// two damped-Jacobi sweeps x -> xn in one march (equal to two launch_smooth_* calls)
void launch_smooth2(hipStream_t st, bool fine, int S, const float* x, float* xn, const float* src, float a, float c,
                    float alpha, float beta);
// fine level (NN even): two sweeps x -> xn and the coarse rhs 4 FW(residual
// of xn) into gc, one march (k_smooth2_resres_march); blocks <= 0: default
void launch_smooth2_resres(hipStream_t st, int Sf, const float* x, float* xn, const float* src, float a, float c,
                           float alpha, float beta, float* gc, int blocks);
// fine level: x + P e, two sweeps -> xn, and the ITP update of U with xn ->
// Un, one march (k_post_march); blocks <= 0: default
void launch_post_itp(hipStream_t st, int Sf, const float* x, const float* e, float* xn, const float* U, float* Un,
                     const float* K, const uint16_t* lab, float a, float c, float alpha, float beta, float half_dv,
                     float dtv, int blocks);
void launch_residual_restrict(hipStream_t st, bool fine, int Sf, const float* x, const float* src, float alpha,
                              float beta, float* gc, float sign);
void launch_prolong_smooth(hipStream_t st, const Tile& t, bool fine, int Sf, const float* x, const float* e,
                           float* x_out, const float* src, float a, float c, float alpha, float beta);
void launch_smooth_itp(hipStream_t st, const Tile& t, int S, const float* P, float* Pn, const float* U, float* Un,
                       const float* K, const uint16_t* lab, float a, float c, float alpha, float beta, float half_dv,
                       float dtv);
void launch_zero(hipStream_t st, float* p, size_t n);
void launch_norm_sums(hipStream_t st, const Tile& t, int S, float h3, const float* U, const uint16_t* lab,
                      int n_dom, Fx* norms /* n_dom sums */, int* bad /* set to 1 on a non-finite sum */);
// scale: n_dom floats of scratch (the per-domain factors)
void launch_normalize(hipStream_t st, const Tile& t, int S, float* U, const uint16_t* lab, const Fx* norms,
                      const float* Z, int n_dom, float* scale);
// Score rule.  Up to kScoreScanMaxAtoms atoms: a scan over all atoms.  Above:
// the argmax over atoms is searched through a uniform cell list (cells of cs
// nodes, n per axis; cell_start[n^3 + 1], cell_atoms sorted by cell; zmax =
// largest Z), visiting shells of cells until no farther atom can reach the
// best score found.  Both give the labels of the scan.
void launch_boundary_score(hipStream_t st, const Tile& t, int S, float h, float h2, const Atom* atoms, int n_atoms,
                           const uint16_t* lin, uint16_t* lout, float ratio, int min_foreign, const AtomCells& cells,
                           const int32_t* cell_start, const int32_t* cell_atoms,
                           uint16_t* best_atom);
// The score rule's best atom of every interior node (Z > 0 atoms, the first
// strict maximum of Z / sqrt(r^2 + h^2) in index order), the map the rule
// reads; nodes of blocks with too many candidate atoms are left unknown.
void launch_score_map(hipStream_t st, int S, float h, float h2, const Atom* atoms, int n_atoms, uint16_t* best_atom);
void launch_levelset(hipStream_t st, const Tile& t, int S, float h, const Atom* atoms, const AtomS2* ext,
                     int n_atoms, const int32_t* lock_group, const uint16_t* lin, uint16_t* lout, const float* U,
                     float* W, float dt_w, float curv_reg, int rc_flags);
void launch_energy(hipStream_t st, const Tile& t, int mode, int S, float h, float h3, const float* U,
                   const float* K, const float* P, const uint16_t* lab, const Atom* atoms, int n_dom,
                   Fx* per_dom /* n_dom*3 sums */, Fx* dip /* 3 sums */, int* bad /* 1 on a non-finite sum */);
void launch_forces_hf(hipStream_t st, int S, float h, float h2, float h3, const float* U, const Atom* atoms,
                      int first, int count, Fx* out3 /* 3 count sums */);
void launch_forces_gradp(hipStream_t st, int S, float h, const float* P, const Atom* atoms, const AtomS2* ext,
                         int n_atoms, int radius, double* out3n);
void launch_square(hipStream_t st, int S, const float* U, float* rho);
void launch_label_u16_to_i32(hipStream_t st, size_t n, const uint16_t* in, int32_t* out);
void launch_label_i32_to_u16(hipStream_t st, size_t n, const int32_t* in, uint16_t* out);
void launch_fill(hipStream_t st, float* p, size_t n, float v);

// Allocation-free HF forces (rq_forces_hf_noalloc): a file-scope device
// array in rq_kernels.hip holds kNoAllocMaxAtoms atoms' fixed-point sums and
// their double values.  noalloc_setup() takes the current device's address
// of it once (hipGetSymbolAddress); noalloc_scratch() returns that address
// for the current device (nullptr before setup).
constexpr int kNoAllocMaxAtoms = 1024;
constexpr size_t kNoAllocFxBytes = sizeof(Fx) * kFxWords * 3 * (size_t)kNoAllocMaxAtoms;
constexpr size_t kNoAllocScratchBytes = kNoAllocFxBytes + sizeof(double) * 3 * (size_t)kNoAllocMaxAtoms;
hipError_t noalloc_setup();
void* noalloc_scratch();
void launch_fx_values(hipStream_t st, int n, const Fx* w, double* out);

// ---- adjoint launchers (rq_adjoint.hip) ----
void launch_init_frozen(hipStream_t st, int S, float h, float h2, const Atom* atoms, int n_atoms, const uint16_t* lab,
                        float* U, float* K);
void launch_energy_grad(hipStream_t st, const Tile& t, int S, float h, float h3, const float* U, const float* K,
                        const float* P, const uint16_t* lab, float* lamU, float* gK, float* lamP);
void launch_itp_adj(hipStream_t st, const Tile& t, int S, const float* lam, float* lam_out, const float* U,
                    const float* K, const float* P, const uint16_t* lab, float half_dv, float dtv, float* gK,
                    float* lamP);
// smoother-adjoint accumulation modes (k_smooth_adj)
enum { kAdjAcc = 0, kAdjSrc = 1, kAdjFirst = 2, kAdjLast = 3 };
void launch_smooth_adj(hipStream_t st, const Tile& t, int S, int mode, const float* lam, float* lam_out,
                       const float* src, float* acc, float* lamU, float a, float c, float coef, float coef2);
void launch_correct_adj(hipStream_t st, const Tile& t, int S, const float* lam, const float* z, float* lam_out,
                        float* acc);
void launch_axpy_src(hipStream_t st, const Tile& t, int S, const float* U, const float* acc, float* lamU, float coef);
void launch_domain_dot(hipStream_t st, const Tile& t, int S, float w, const float* A, const float* B,
                       const uint16_t* lab, int n_dom, Fx* out);
void launch_normalize_adj(hipStream_t st, const Tile& t, int S, float h3, const float* Upre, float* lam,
                          const uint16_t* lab, const Fx* norms, const Fx* dots, const float* Z, int n_dom);
void launch_pos_grad(hipStream_t st, int S, float h, float h2, const float* gK, const float* lamU0,
                     const uint16_t* lab, const Atom* atoms, int n_atoms, Fx* out /* 3 n_atoms sums */);

}  // namespace rq
