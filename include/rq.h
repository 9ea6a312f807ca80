/*
 * rq.h -- C ABI of the RealQM grid solver for AMD Instinct MI355X (gfx950).
 *
 * One opaque handle owns the device fields of one molecular grid and one HIP
 * stream.  All calls on a handle are issued in order on that stream; calls
 * that return data to the host synchronise the stream.  A handle is not
 * thread-safe; distinct handles may be used from distinct threads.
 *
 * Every function returns 0 (RQ_OK) on success or a negative RQ_E* code; the
 * message of the last failure on a handle is returned by rq_last_error(h)
 * (rq_last_error(NULL) gives the last failure of rq_create / batch create).
 *
 * Grid convention (cloud/solver.py:322-329): NN cells and S = NN + 1 nodes
 * per axis, h = screen / NN, flat node index id = i*S*S + j*S + k (k fastest).
 * Face nodes (any index 0 or NN) are Dirichlet data: kernels write interior
 * nodes only.  Atom positions are fractional grid units (cloud/solver.py:335).
 *
 * Reference interfaces replaced (file:line in Claes542/RealMolecule):
 *   rq_create            RealQMSolver.__init__          cloud/solver.py:321-400
 *   rq_initialize        RealQMSolver.initialize        cloud/solver.py:402-419
 *   rq_init_trial        initTrial kernel launch        cloud/solver.py:404-410, 16-59
 *   rq_poisson           RealQMSolver._poisson_step     cloud/solver.py:434-444
 *   rq_update            updateU launch + U/U2 swap     cloud/solver.py:567-572, 61-90
 *   rq_normalize         RealQMSolver._normalize        cloud/solver.py:421-432
 *   rq_boundary          boundaryEvolve launch          cloud/solver.py:579-586, 92-147
 *                        evolveBoundary pass + copy     molecule.js:3093-3104, 426-540
 *   rq_step              body of RealQMSolver.run       cloud/solver.py:561-586
 *                        per-step part of doSteps       molecule.js:2829-2996
 *   rq_energy            RealQMSolver._compute_energy   cloud/solver.py:446-460
 *                        reduceEnergy+finalizeEnergy    molecule.js:950-1075, 3078-3091
 *   rq_forces_hf         computeForceAll batches        cloud/solver.py:478-502, 282-315
 *   rq_forces_gradp      gradPtotal pass                molecule.js:1199-1270, 2997-3005
 *   rq_set_atoms         atom_data upload after moves   cloud/solver.py:538-543
 *   rq_recompute_k       recomputeK launch              cloud/solver.py:544-548, 149-179
 *   rq_read_field /      CuPy array .get() / assignment of U, U2, P, P2, K_buf,
 *   rq_write_field       rho, label, bestU              cloud/solver.py:365-372;
 *                        cloud/run_h2o_debug.py:20-57; cloud/benchmark.py:104
 */
#ifndef RQ_H
#define RQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RQ_ABI_VERSION 1

/* status codes */
#define RQ_OK         0
#define RQ_EINVAL    -1  /* bad argument or unsupported parameter combination */
#define RQ_ENOMEM    -2  /* device allocation failed */
#define RQ_EHIP      -3  /* HIP runtime error (message has the HIP error string); the handle
                            then refuses further calls with RQ_EHIP, except rq_destroy and
                            rq_last_error (molecule.js device.lost, :2258-2261) */
#define RQ_ESTATE    -4  /* call out of order (e.g. step before initialise) */
#define RQ_ERANGE    -5  /* size beyond a documented limit */
#define RQ_EUNSTABLE -6  /* non-finite energy/field detected */

/* Poisson solver for the Hartree potential, lap P = -2 pi rho */
#define RQ_POISSON_JACOBI    0  /* cloud/solver.py: sweeps of P'=(a P + sum6 P + 2pi h^2 rho)/c */
#define RQ_POISSON_MULTIGRID 1  /* V-cycle per Poisson step, exact FP64 coarsest solve */
#define RQ_POISSON_TWOGRID   2  /* molecule.js: smoother sweeps + two-grid correction every
                                   twogrid_interval steps (sign-inverted, x0.5; molecule.js:2879-2907) */

/* domain-boundary rule */
#define RQ_BOUNDARY_SCORE    0  /* cloud/solver.py:92-147 nuclear-score flip */
#define RQ_BOUNDARY_LEVELSET 1  /* molecule.js:426-540 level set W */
#define RQ_BOUNDARY_OFF      2

/* trial initialisation / softening semantics */
#define RQ_INIT_S1 0  /* cloud/solver.py:16-59: interior only, r~ = sqrt(r^2+h^2) or max(sqrt(r^2+.04h^2), rc) */
#define RQ_INIT_S2 1  /* molecule.js:1377-1449: all nodes, integer positions, r~ = max(r, 2h) or
                         max(sqrt(r^2+.04h^2), rc); initZeff = Z; faces keep trial values */

/* energy semantics */
#define RQ_ENERGY_S1 0  /* cloud/solver.py:227-280 */
#define RQ_ENERGY_S2 1  /* molecule.js:950-1075: r < max(rc, 2h) exclusion of the own nucleus */

/* field ids for rq_read_field / rq_write_field / rq_field_device_ptr */
#define RQ_FIELD_U      0  /* float32 S^3, current amplitude */
#define RQ_FIELD_U2     1  /* float32 S^3, ping-pong partner of U */
#define RQ_FIELD_K      2  /* float32 S^3, nuclear potential sum Z/r~ */
#define RQ_FIELD_P      3  /* float32 S^3, current Hartree potential */
#define RQ_FIELD_P2     4  /* float32 S^3, ping-pong partner of P */
#define RQ_FIELD_RHO    5  /* float32 S^3, read only: U*U (computed on demand) */
#define RQ_FIELD_LABEL  6  /* int32 S^3 on the host (stored as uint16 on the device) */
#define RQ_FIELD_BESTU  7  /* float32 S^3, trial maximum from initialisation */
#define RQ_FIELD_W      8  /* float32 S^3, level-set value (RQ_BOUNDARY_LEVELSET) */
#define RQ_FIELD_POTHER 9  /* float32 S^3, frame potential (RQ_POTHER_FRAME) */

typedef struct rq_solver rq_solver;

/* One atom row, as cloud/solver.py:378-382 atom_data: fractional grid
 * position (i, j, k), electron charge Z (domain Z_eff; Z <= 0 = inactive) and
 * pseudopotential radius rc (bohr, 0 = bare). */
typedef struct rq_atom {
  float i, j, k, Z, rc;
} rq_atom;

typedef struct rq_params {
  /* screen and dv are double: the reference forms h, h^2, h^3 and dv*h^2 in
   * double precision on the host and rounds each to FP32 once
   * (cloud/solver.py:326-328, 352-353, 408). */
  double  screen;         /* box edge in bohr */
  double  dv;             /* ITP step in h^2 units; <= 0 selects cloud/solver.py:341-351 */
  int32_t NN;             /* cells per axis (S = NN+1 nodes); even; 8 <= NN <= 1024 */
  int32_t init_mode;      /* RQ_INIT_* */
  int32_t poisson_mode;   /* RQ_POISSON_* */
  int32_t jacobi_sweeps;  /* sweeps per Poisson step (JACOBI/TWOGRID), default 2 */
  float   smooth_a;       /* smoother P' = a*P + (sum6 P + rhs)/c; S1 (0, 6), molecule.js (0.3333, 9) */
  float   smooth_c;
  int32_t mg_levels;      /* levels incl. the finest (>= 2); NN divisible by 2^(levels-1) */
  int32_t mg_nu1;         /* pre-smoothing sweeps */
  int32_t mg_nu2;         /* post-smoothing sweeps */
  float   mg_a;           /* V-cycle smoother (a, c); default omega = 6/7: (1/7, 7) */
  float   mg_c;
  int32_t twogrid_interval;      /* RQ_POISSON_TWOGRID: correction every n steps (molecule.js:110) */
  int32_t twogrid_coarse_sweeps; /* molecule.js: 10 */
  int32_t boundary_mode;  /* RQ_BOUNDARY_* */
  float   flip_ratio;     /* score rule: flip if best > ratio * current (1.1) */
  int32_t flip_min_foreign; /* score rule: foreign neighbours needed (2) */
  float   ls_dt_w;        /* level set: boundarySpeed dt_w (0.5, molecule.js:125) */
  float   ls_curv_reg;    /* level set: curvature weight (0.15, molecule.js:34) */
  int32_t energy_mode;    /* RQ_ENERGY_* */
  int32_t rc_dirichlet;   /* 1: molecule.js rc rule in the ITP update (U'=0 inside own rc) */
  int32_t pother_mode;    /* RQ_POTHER_*: potential the ITP, energy and forces see */
  int32_t sic_interval;   /* RQ_POTHER_FRAME: self-interaction correction every n frames (0: never) */
  int32_t reserved[6];    /* must be zero */
} rq_params;

/* pother_mode.  TOTAL: P solves the Poisson equation of all electrons (the
 * cloud/solver.py potential).  DIRECT: molecule.js "direct Pother"
 * (molecule.js:597-621, 664-688, 2834-2858): for every domain m with Z > 0 a
 * persistent P_m solves nabla^2 P_m = -2 pi rho_other_m (the density of all
 * OTHER domains) with one multigrid V-cycle per Poisson step, and P holds
 * P_label(x)(x) -- each electron sees the others' repulsion, not its own.
 * Needs RQ_POISSON_MULTIGRID. */
#define RQ_POTHER_TOTAL 0
#define RQ_POTHER_DIRECT 1
/* FRAME: molecule.js's frame-stale PotherBuf (molecule.js:2436-2470,
 * 3009-3076).  The ITP update and the energy read Pother, which
 * rq_pother_frame rebuilds at each frame's end as a copy of P, minus, on
 * frames that are a multiple of sic_interval, each domain's self-potential
 * on its own nodes (scaled by 1/Z for Z > 1).  Needs a coarse level
 * (RQ_POISSON_TWOGRID or RQ_POISSON_MULTIGRID) when sic_interval > 0. */
#define RQ_POTHER_FRAME 2
/* DIRECT_SWEEPS: molecule.js's own direct Pother, its default for pages of
 * at most 5 electrons (molecule.js:111, 624-688, 2833-2858).  For every
 * domain m with Z != 0 a persistent P_m, set by rq_initialize to
 * sum_{n != m, Z_n > 0} 0.5 / sqrt(r_n^2 + h^2), takes per Poisson step 10
 * damped sweeps P' = a P + (sum6 P + h^2 2pi rho_other_m) / c with the
 * handle's smoother (smooth_a, smooth_c; the page's 0.3333 and 9), and the
 * frame potential Pother takes P_m on the nodes labelled m.  The ITP update,
 * the energy and the grad-P forces read Pother.  Any poisson_mode. */
#define RQ_POTHER_DIRECT_SWEEPS 3

/* Scheduling intervals of the run() loop (cloud/solver.py:550-586). */
typedef struct rq_schedule {
  int32_t poisson_interval;  /* Poisson step when step % n == 0 (2) */
  int32_t norm_interval;     /* normalise when step % n == 0 (20) */
  int32_t boundary_interval; /* boundary rule when step % n == 0 (5*norm_interval = 100) */
} rq_schedule;

/* --- library --- */
int         rq_abi_version(void);
const char* rq_last_error(const rq_solver* h);
void        rq_default_params(rq_params* p);       /* cloud/solver.py semantics, Jacobi Poisson */
void        rq_default_schedule(rq_schedule* s);   /* (2, 20, 100) */
double      rq_dv_for_electrons(int32_t n_elec);   /* cloud/solver.py:341-351 */

/* --- lifetime --- */
int rq_create(const rq_params* p, const rq_atom* atoms, int32_t n_atoms, rq_solver** out);
int rq_destroy(rq_solver* h);
int rq_sync(rq_solver* h);
/* Test hook, no counterpart in the reference: the n_checks-th HIP runtime
 * check after this call reports hipErrorLaunchFailure, as a failing device
 * would, so callers' lost-device handling can be exercised; 0 disarms. */
int rq_debug_fail_after(rq_solver* h, int32_t n_checks);
int rq_get_info(const rq_solver* h, int32_t* NN, int32_t* S, int32_t* n_atoms, double* h_grid, double* dv);

/* --- atoms --- */
int rq_set_atoms(rq_solver* h, const rq_atom* atoms, int32_t n_atoms); /* same n as create */
int rq_get_atoms(const rq_solver* h, rq_atom* atoms, int32_t n_atoms);
/* Lock groups of the level-set rule (molecule.js:487-500): group[d] > 0 marks
 * domain d; a boundary between two domains of one group does not evolve. */
int rq_set_lock_groups(rq_solver* h, const int32_t* group, int32_t n_atoms);

/* The molecule.js per-atom fields beyond rq_atom (struct Atom,
 * molecule.js:183-191, filled by fillAtomBuf, :1875-1903), used with
 * init_mode RQ_INIT_S2: the trial init and the K refresh, the grad-P forces
 * and the level set's sector rule.  Until set, every atom has the defaults
 * (Z_nuc = Z, init_zeff = Z, init_rcut = 1e6 bohr, no split). */
#define RQ_SPLIT_SPHERE 0  /* no angular restriction */
#define RQ_SPLIT_HEMI   2  /* split_idx 0: the half-space axis.d < 0, 1: axis.d >= 0 */
#define RQ_SPLIT_THIRD  3  /* split_idx 0..2: 120-degree sector about the axis, from split_rot */
typedef struct rq_atom_s2 {
  float   Z_nuc;      /* nuclear charge of K, E_KK and forces (molecule.js:22-25); <= 0: Z.
                         Z = 0 with Z_nuc > 0 is a bare nucleus: in K, no electron domain */
  float   init_zeff;  /* decay rate of the trial Zeff^2/sqrt(pi) exp(-Zeff r~) (:1889-1891); <= 0: Z */
  float   init_rcut;  /* trial zero beyond this distance from the nucleus, bohr (:1892); <= 0: 1e6 */
  int32_t split_type; /* RQ_SPLIT_*: the domain owns only nodes in its sector (molecule.js:194-225) */
  int32_t split_idx;
  float   split_axis[3];
  float   split_rot;  /* angular offset of the THIRD sectors (rad) */
} rq_atom_s2;
/* molecule.js USER_Z_NUC, USER_INIT_ZEFF / INIT_ZEFF, USER_INIT_RCUT /
 * INIT_RCUT and USER_SPLIT_* per atom; takes effect at the next rq_init_trial
 * / rq_initialize / rq_recompute_k (K, labels, trial) and immediately in the
 * forces and the level set. */
int rq_set_atoms_s2(rq_solver* h, const rq_atom_s2* ext, int32_t n_atoms);

/* --- passes --- */
int rq_init_trial(rq_solver* h);                 /* K, label, U, bestU (and W = 1) */
int rq_initialize(rq_solver* h, int32_t bootstrap_poisson_steps); /* init_trial + normalise + n Poisson steps */
int rq_poisson(rq_solver* h, int32_t n_steps);   /* n Poisson steps in the configured mode */
int rq_update(rq_solver* h);                     /* one ITP step + swap */
int rq_normalize(rq_solver* h);                  /* per-domain renormalisation */
int rq_boundary(rq_solver* h);                   /* one application of the boundary rule */
int rq_recompute_k(rq_solver* h);                /* K after rq_set_atoms */
/* One molecule.js two-grid correction of P (molecule.js:2879-2907): residual,
 * full-weighting restriction, twogrid_coarse_sweeps damped sweeps from zero on
 * the NN/2 grid with the sign-inverted right-hand side, P += 0.5 * prolongation.
 * Needs poisson_mode RQ_POISSON_TWOGRID. */
int rq_twogrid_correct(rq_solver* h);
/* Steps first_step+1 .. first_step+n_steps of the run() loop body
 * (cloud/solver.py:561-592).  When a normalisation falls in the range, the
 * call waits for the stream and returns RQ_EUNSTABLE if the normalisation
 * met an amplitude that is not finite (molecule.js:3198-3201 stops the same
 * way on a non-finite energy). */
int rq_step(rq_solver* h, int64_t first_step, int32_t n_steps, const rq_schedule* s);
/* on != 0: rq_step replays aligned chunks of norm_interval steps from cached
 * HIP graphs (captured on first use; Jacobi/multigrid modes, timing off,
 * norm_interval a multiple of poisson_interval, boundary_interval a multiple
 * of norm_interval); other steps run eagerly.  Results are identical. */
int rq_set_graph(rq_solver* h, int32_t on);
/* on != 0 (default): multigrid passes run fused -- two pre-sweeps from the zero
 * coarse guess in one pass, residual+restriction without the residual array,
 * prolongation+correction with the first post-sweep, and the last fine
 * post-sweep with the ITP update; with two pre-sweeps and 1..3 post-sweeps
 * every coarse level runs as one down-leg launch (sweeps, residual,
 * restriction) and one up-leg launch (prolongation, sweeps).  on == 2: fused
 * passes without the per-level legs.  on == 3: as 1, and the last two levels
 * of the V-cycle (up to 25^3 nodes) run in one LDS-resident workgroup.
 * Results are bit-identical to on == 0. */
int rq_set_fused(rq_solver* h, int32_t on);
/* Launch shape of the 2.5D stencil kernels on multigrid level `level` (0 =
 * finest, -1 = all): a (bx, by) workgroup tile in (k, j), bx*by a multiple of
 * 64 and <= 256, each thread marching `ich` planes along i.  bx <= 0 restores
 * the built-in choice.  Used by the autotuner (realmolecule_amd/autotune.py). */
int rq_set_tile(rq_solver* h, int32_t level, int32_t bx, int32_t by, int32_t ich);
int rq_get_tile(const rq_solver* h, int32_t level, int32_t* bx, int32_t* by, int32_t* ich);

/* --- reductions (FP64) --- */
/* out[0..5] = T, V_eK, V_ee, sum rho x h^3, sum rho y h^3, sum rho z h^3 (bohr).
 * per_domain (nullable) = n_atoms x 3 (T, V_eK, V_ee) keyed by voxel label. */
int rq_energy(rq_solver* h, double* out6, double* per_domain);
/* per-domain sum U^2 h^3 (n_atoms values) */
int rq_domain_norms(rq_solver* h, double* norms);
/* Electronic HF force on atoms [first, first+count): Z_n sum rho h^3 d/(d^2+h^2)^1.5 */
int rq_forces_hf(rq_solver* h, int32_t first, int32_t count, double* out3);
/* rq_forces_hf without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as rq_forces_hf; its partial sums
 * live in one library-owned array per GPU (72 KiB, kept until the process
 * ends) instead of a buffer allocated per call, so at most 1024 atoms per
 * call (count > 1024: RQ_ERANGE; call it in batches of atoms).  The array is
 * shared by every handle on that GPU: calls that overlap on one GPU race
 * (issue them from one thread, or serialise them). */
int rq_forces_hf_noalloc(rq_solver* h, int32_t first, int32_t count, double* out3);
/* molecule.js sphere-averaged 2 Z <grad P> over radius `radius` cells (<=0: max(5, round(3/h))) */
int rq_forces_gradp(rq_solver* h, int32_t radius, double* out3n);

/* --- fields --- */
int rq_read_field(rq_solver* h, int32_t field, void* host, size_t bytes);
int rq_write_field(rq_solver* h, int32_t field, const void* host, size_t bytes);
int rq_swap_u(rq_solver* h);                     /* exchange U and U2 (cloud/benchmark.py:104) */
int rq_field_device_ptr(rq_solver* h, int32_t field, void** dptr);
int rq_stream(rq_solver* h, void** hip_stream);
/* Display slice at k = slice_k into host (3 S^2 + 2 S floats, S = NN+1):
 * density U^2, label + 1000 Z, boundary marks (1 active, 2 broken) as S x S
 * (i, j) planes, then the K line K(i, k, k) and the cutoff-weight line (all
 * 1: no W cutoff).  Replaces molecule.js's extract pass + slice readback
 * (molecule.js:1108-1192, 1818-1835). */
int rq_display_slice(rq_solver* h, int32_t slice_k, float* host, size_t n_floats);
/* End-of-frame potential of RQ_POTHER_FRAME: Pother = P, then for frame > 0
 * and frame % sic_interval == 0, for every domain m with Z != 0, the
 * self-potential of its density (4 sweeps from zero, one two-grid
 * correction, 4 sweeps, in the solver's S2 smoother) is subtracted on the
 * domain's nodes.  Replaces the Pother block of the molecule.js frame
 * (molecule.js:3009-3076). */
int rq_pother_frame(rq_solver* h, int64_t frame);

/* --- differentiation: frozen-label fixed-step trajectory --- */
/* The differentiated map: K(R) and the trial U0 = Z_L^2/sqrt(pi) exp(-Z_L r~)
 * of each voxel's CURRENT label L (labels are frozen: call rq_init_trial at
 * the reference geometry first), per-domain normalise, `bootstrap` Poisson
 * steps from P = 0, then steps 1..n_steps of the run() loop body without the
 * boundary rule; E = T + V_eK + V_ee of the final state.  Positions are the
 * handle's atoms.  Replaces the per-molecule finite-difference force path
 * (6N+1 solves) built on RealQMSolver.run (cloud/solver.py:550-634). */
typedef struct rq_adjoint_spec {
  int64_t     n_steps;
  int32_t     bootstrap;           /* Poisson steps after the initial normalise */
  int32_t     checkpoint_interval; /* steps per stored (U, P) checkpoint; <= 0: the whole
                                      trajectory when its 2 n_steps + 1 fields fit in a quarter
                                      of the free device memory (no recomputation), else 64 */
  rq_schedule schedule;            /* boundary_interval is ignored (frozen labels) */
  int32_t     start;               /* RQ_ADJ_START_TRIAL or RQ_ADJ_START_SAVED */
  int32_t     reserved[3];         /* must be zero */
} rq_adjoint_spec;
/* The map starts from the trial U0(R), normalised, and `bootstrap` Poisson
 * steps from P = 0 (the reference's initialize, cloud/solver.py:402-419). */
#define RQ_ADJ_START_TRIAL 0
/* The map starts from the U and P that rq_adjoint_save_start stored (e.g. a
 * converged state at a reference geometry); they are constants of the map,
 * so dE/dR is the derivative of the n_steps that follow.  bootstrap is
 * ignored.  No counterpart in the reference: it re-converges each FD
 * geometry from the trial. */
#define RQ_ADJ_START_SAVED 1
/* Stores the handle's current U and P as the start of RQ_ADJ_START_SAVED
 * maps (two fields of device memory, kept until rq_destroy). */
int rq_adjoint_save_start(rq_solver* h);

/* Forward solve only; energy3 = (T, V_eK, V_ee). */
int rq_solve_frozen(rq_solver* h, const rq_adjoint_spec* spec, double* energy3);
/* Forward solve + reverse sweep: energy3 as above and dEdR (n_atoms x 3,
 * Ha/bohr) of the electronic energy; the nucleus-nucleus term is the caller's.
 * Multigrid mode requires mg_nu1 == mg_nu2 (symmetric V-cycle). */
int rq_adjoint_gradient(rq_solver* h, const rq_adjoint_spec* spec, double* energy3, double* dEdR);

/* --- batches: many molecules with uneven grids in one launch per pass --- */
/* A batch advances n molecules together: every pass (each V-cycle level,
 * the ITP update, normalise, energy, init) is one launch over the
 * concatenated tiles of all molecules (one workgroup per tile of every
 * molecule, not a fixed resident grid), and step chunks replay from HIP
 * graphs.  Semantics per molecule: cloud/solver.py trial init, normalise,
 * multigrid Poisson (rq_params: mg_levels, mg_nu1, mg_nu2, mg_a, mg_c; every
 * NN divisible by 2^(mg_levels-1), coarsest interior <= 19), ITP, normalise
 * every norm_interval; labels frozen (no boundary rule).  Replaces running the
 * reference's per-molecule solver once per conformer in a geometry scan. */
typedef struct rq_batch rq_batch;
typedef struct rq_molecule {
  int32_t        NN;       /* cells per axis of this molecule's grid */
  double         screen;   /* box edge (bohr) */
  int32_t        n_atoms;  /* 1..256 */
  const rq_atom* atoms;
} rq_molecule;

const char* rq_batch_last_error(const rq_batch* b);
int rq_batch_create(const rq_params* p, const rq_molecule* mols, int32_t n_mols, rq_batch** out);
int rq_batch_destroy(rq_batch* b);
int rq_batch_info(const rq_batch* b, int32_t* n_mols, int64_t* interior_voxels);
/* (bx, by) workgroup tile for all stencil passes; ich[m] planes marched per
 * thread on molecule m's fine grid (nullable: built-in choice). */
int rq_batch_set_tiles(rq_batch* b, int32_t bx, int32_t by, const int32_t* ich);
int rq_batch_set_graph(rq_batch* b, int32_t on);
int rq_batch_initialize(rq_batch* b, int32_t bootstrap_vcycles);
int rq_batch_step(rq_batch* b, int64_t first_step, int32_t n_steps, const rq_schedule* s);
int rq_batch_sync(rq_batch* b);
/* out3[3m..3m+2] = (T, V_eK, V_ee) of molecule m, FP64 */
int rq_batch_energy(rq_batch* b, double* out3);
/* fields U, K, P (float32) or label (int32) of molecule m, (NN_m+1)^3 values */
int rq_batch_read_field(rq_batch* b, int32_t mol, int32_t field, void* host, size_t bytes);
/* Positions (and rc) of molecule mol's atoms for the frozen-label map below;
 * Z must not change.  Replaces re-creating a per-molecule solver for each
 * geometry of a scan (cloud/solver.py:321-400). */
int rq_batch_set_atoms(rq_batch* b, int32_t mol, const rq_atom* atoms, int32_t n_atoms);
/* rq_solve_frozen / rq_adjoint_gradient of every molecule at once, one
 * launch per pass over all molecules' tiles (forward map and reverse
 * sweep): labels frozen at those of rq_batch_initialize, start
 * RQ_ADJ_START_TRIAL, multigrid Poisson.  energy3: (T, V_eK, V_ee) per
 * molecule (3 n_mols); dEdR: the electronic dE/dR of every atom, molecules
 * in order (3 per atom, Ha/bohr).  Overwrites the batch's U, P and K (call
 * rq_batch_initialize before stepping again).  Replaces running the
 * per-molecule adjoint, or central differences, once per geometry. */
int rq_batch_solve_frozen(rq_batch* b, const rq_adjoint_spec* spec, double* energy3);
int rq_batch_adjoint_gradient(rq_batch* b, const rq_adjoint_spec* spec, double* energy3, double* dEdR);
int rq_batch_timing_enable(rq_batch* b, int32_t on);
int rq_batch_timing_read(rq_batch* b, double* ms, int64_t* launches, int32_t n_families);

/* --- timing (HIP events on the handle's stream) --- */
/* Per-kernel-family accumulated device time in ms since the last enable, and
 * launch counts.  Families: 0 ITP update, 1 fine smoother sweep, 2 coarse
 * smoother sweep, 3 fine residual+restriction, 4 coarse residual+restriction,
 * 5 fine prolongation+correction, 6 coarse prolongation+correction,
 * 7 exact coarsest solve, 8 normalise, 9 boundary rule, 10 energy, 11 forces,
 * 12 last fine smoother sweep fused with the ITP update, 13 the last two
 * levels of the V-cycle in one single-workgroup launch, 14 the fine
 * pre-smoothing pair with the residual and restriction in one launch, 15
 * the fine correction, post-smoothing pair and ITP update in one launch.  With
 * fused passes (the default) family 3 is the fused residual+restriction and
 * family 5 the fused prolongation+first post-smoothing sweep.  Events bracket
 * each launch on the handle's stream. */
#define RQ_TIMING_FAMILIES 16
int rq_timing_enable(rq_solver* h, int32_t on);
int rq_timing_read(rq_solver* h, double* ms, int64_t* launches, int32_t n_families);

#ifdef __cplusplus
}
#endif
#endif /* RQ_H */
