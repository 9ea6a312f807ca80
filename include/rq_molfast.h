/*
 * rq_molfast.h -- C ABI of the fused per-electron RealQM solver (mol_fast.js)
 * for AMD Instinct MI355X (gfx950).
 *
 * mol_fast.js keeps, per electron m (one per nucleus, NELEC <= 16), an
 * amplitude u_m, a partition weight w_m and a Hartree potential P_m on one
 * S^3 grid (S = NN + 1, h = screen / NN, node id = i*S*S + j*S + k) and
 * advances all three in one fused update.  One handle owns those fields on
 * the device and one HIP stream; calls are issued in order on it and the
 * ones that return data synchronise it.  Not thread-safe per handle.
 *
 * Every function returns 0 (RQ_OK) or a negative RQ_E* code (rq.h); the
 * last failure's message is rq_mf_last_error(h) (NULL: the last failed
 * rq_mf_create).
 *
 * Reference interfaces replaced (file:line in Claes542/RealMolecule):
 *   rq_mf_create        USER_* globals + nuclei table       mol_fast.js:13-93
 *                       writeParams / templated constants   mol_fast.js:96-131, 838-876
 *   rq_mf_init          uploadInitialData                    mol_fast.js:941-1046
 *   rq_mf_do_steps      doSteps(n), GPU part                 mol_fast.js:1237-1304
 *                       (fused update 134-351, reduce +
 *                       finalize 352-466, normalize 468-489,
 *                       per-electron 2-grid correction 642-789)
 *   rq_mf_set_r_out     rOutEnforce toggle + writeParams     mol_fast.js:1239-1244
 *   rq_mf_forces        forcePL dispatch + readback          mol_fast.js:496-560, 1339-1354
 *   rq_mf_set_positions writeParams after stepNuclei         mol_fast.js:1398-1401
 *   rq_mf_recompute_k   recomputeK dispatch                  mol_fast.js:563-586, 1402-1408
 *   rq_mf_extract       extract dispatch + slice readback    mol_fast.js:592-641, 1289-1291
 *   rq_mf_read_field /  U_buf / W_buf / P_buf / K_buf contents
 *   rq_mf_write_field
 */
#ifndef RQ_MOLFAST_H
#define RQ_MOLFAST_H

#include <stddef.h>
#include <stdint.h>

#include "rq.h"

#ifdef __cplusplus
extern "C" {
#endif

#define RQ_MF_MAX_ELEC 16

/* angular split of an orbital (mol_fast.js:118, 878-939) */
#define RQ_MF_SPLIT_SPHERE     0
#define RQ_MF_SPLIT_HEMI       2
#define RQ_MF_SPLIT_THIRD      3
#define RQ_MF_SPLIT_TETRA      4
#define RQ_MF_SPLIT_HEMI_THIRD 5

/* fields for rq_mf_read_field / rq_mf_write_field (current ping-pong half) */
#define RQ_MF_FIELD_U 0
#define RQ_MF_FIELD_W 1
#define RQ_MF_FIELD_P 2
#define RQ_MF_FIELD_K 3

/* one USER_NUCLEI entry (mol_fast.js:29-46) with its USER_NORM_TARGETS entry */
typedef struct rq_mf_nucleus {
  double i, j, k;   /* grid position */
  double Z;         /* nuclear charge; Z <= 0 exerts and feels no force */
  double rc;        /* cusp radius (au), 0: none */
  double r_out;     /* outer confinement radius (au), 0: none */
  double tilt_x, tilt_y; /* V = -tilt_x x - tilt_y y */
  double hs;        /* half-space of the initial partition: -1, 0, +1 */
  double z_init;    /* initial decay rate; NaN: Z > 0 ? Z : 1 */
  double target;    /* occupation (norm target) */
  double axis[3];   /* split axis (normalised by the library) */
  int32_t split;    /* RQ_MF_SPLIT_* */
  int32_t split_idx;
  int32_t split_fix; /* 1: enforce the split every step, 0: at set-up only */
  int32_t reserved;
} rq_mf_nucleus;

typedef struct rq_mf_params {
  int32_t NN;              /* cells per axis, even */
  int32_t shell_mode;      /* USER_SHELL_MODE: rc / r_out / split of the own orbital only */
  int32_t full_self;       /* USER_FULL_SELF */
  int32_t vcycle;          /* USER_VCYCLE !== false */
  int32_t vcycle_interval; /* USER_VCYCLE_INTERVAL (50) */
  int32_t coarse_iters;    /* USER_COARSE_ITERS (10) */
  int32_t norm_interval;   /* 20 */
  int32_t reserved0;
  double screen;           /* USER_SCREEN, au */
  double kinetic_coeff;    /* USER_KINETIC_COEFF (0.5) */
  double kc_multi;         /* USER_KC_MULTI (1) */
  double beta_pauli;       /* USER_BETA_PAULI (0) */
  double alpha_tf;         /* USER_ALPHA_TF (0) */
  double field_x;          /* USER_FIELD_X (0) */
  int32_t reserved[8];
} rq_mf_params;

typedef struct rq_mf rq_mf;

void        rq_mf_default_params(rq_mf_params* p);
const char* rq_mf_last_error(const rq_mf* h);
int rq_mf_create(const rq_mf_params* p, const rq_mf_nucleus* nuclei, int32_t n, rq_mf** out);
int rq_mf_destroy(rq_mf* h);
int rq_mf_sync(rq_mf* h);

/* K, U (pre-normalised), W, P on every node, both ping-pong halves */
int rq_mf_init(rq_mf* h);
/* n fused steps; sums (NELEC + 6 doubles, may be NULL) receives the last
 * reduction: norm_m * h^3 sums, then T, V_eK, V_ee, dipole x, y, z */
int rq_mf_do_steps(rq_mf* h, int32_t n, double* sums);
int rq_mf_set_r_out(rq_mf* h, int32_t enforce);
/* electronic force per nucleus, 3 floats each (2 Z <grad sum_m P_m>) */
int rq_mf_forces(rq_mf* h, float* out3n);
int rq_mf_set_positions(rq_mf* h, const double* ijk3n);
int rq_mf_recompute_k(rq_mf* h);
/* 3 NELEC S^2 floats */
int rq_mf_extract(rq_mf* h, float* out, size_t n_floats);
int rq_mf_read_field(rq_mf* h, int32_t field, int32_t m, float* host, size_t n_floats);
int rq_mf_write_field(rq_mf* h, int32_t field, int32_t m, const float* host, size_t n_floats);
/* with timing on, rq_mf_do_steps brackets every fused update with HIP
 * events on the handle's stream; rq_mf_update_time returns their summed
 * device time (ms) and count for the last rq_mf_do_steps call */
int rq_mf_set_timing(rq_mf* h, int32_t on);
int rq_mf_update_time(rq_mf* h, double* ms, int32_t* launches);
/* Host-only query (no device call): the (tk, tj) node tile of the
 * three-plane LDS-ring update for a grid of NN cells, and the LDS slot its
 * lanes past tk * tj write instead of a node's slot.  A tile's LDS region
 * is (tk + 2) x (tj + 2) slots, all below sink_slot.  No counterpart in the
 * reference (the WGSL update has no LDS tiles); exported for tests. */
void rq_mf_ring_tile(int32_t NN, int32_t* tk, int32_t* tj, int32_t* sink_slot);

#ifdef __cplusplus
}
#endif
#endif
