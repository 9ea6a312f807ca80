/*
 * cstep.c -- C + OpenMP restatement of the cloud/solver.py step with a
 * multigrid Poisson step (TEST INFRASTRUCTURE ONLY).
 *
 * The all-core CPU baseline that bench.py times beside the GPU, on the same
 * fields.  Each pass performs the FP32 operations of oracle/s1.py and
 * oracle/mg.py in the same order (built with -ffp-contract=off):
 *   itp_update   U' = (U + half_dv*lap) + (dtv*(K - 2P))*U, lap = the six
 *                label-Neumann terms (+i,-i,+j,-j,+k,-k) - 6U
 *                (cloud/solver.py:61-90; oracle/s1.py itp_update)
 *   smooth       x' = a x + (sum6 x + g)/c           (oracle/s1.py smooth)
 *   residual     r = g - (6 x - sum6 x)              (oracle/mg.py residual)
 *   restrict4    (8 c + 4 faces + 2 edges + corners) * 0.0625  (restrict4)
 *   prolong      linear along k, then j, then i      (prolong)
 *   exact solve  FP64 type-I DST diagonalisation     (exact_solve)
 *   normalise    per-domain FP64 sums of (U*U)*h3, U *= (float)sqrt(Z/n)
 *                (cloud/solver.py:421-432; oracle/s1.py normalize)
 * Fields are (S,S,S) with k fastest and zero faces.  Checked against the
 * NumPy oracle in tests/test_oracle_cstep.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(S, i, j, k) ((((size_t)(i)) * (S) + (size_t)(j)) * (S) + (size_t)(k))

int cs_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static float sum6(const float* x, int S, int i, int j, int k) {
  const size_t id = IDX(S, i, j, k), S2 = (size_t)S * S;
  return x[id + S2] + x[id - S2] + x[id + S] + x[id - S] + x[id + 1] + x[id - 1];
}

/* x' = a x + (sum6 x + g) / c on the interior; faces of xn = faces of x */
static void smooth(int S, const float* x, float* xn, const float* g, float a, float c) {
  const int n = S - 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 1; i < n; ++i)
    for (int j = 1; j < n; ++j)
      for (int k = 1; k < n; ++k) {
        const size_t id = IDX(S, i, j, k);
        float v = (sum6(x, S, i, j, k) + g[id]) / c;
        if (a != 0.0f) v = a * x[id] + v;
        xn[id] = v;
      }
}

static void residual(int S, const float* x, const float* g, float* r) {
  const int n = S - 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 1; i < n; ++i)
    for (int j = 1; j < n; ++j)
      for (int k = 1; k < n; ++k) {
        const size_t id = IDX(S, i, j, k);
        r[id] = g[id] - (6.0f * x[id] - sum6(x, S, i, j, k));
      }
}

static const int FACES[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
static const int EDGES[12][3] = {{1, 1, 0},  {1, -1, 0}, {-1, 1, 0}, {-1, -1, 0}, {1, 0, 1},  {1, 0, -1},
                                 {-1, 0, 1}, {-1, 0, -1}, {0, 1, 1}, {0, 1, -1},  {0, -1, 1}, {0, -1, -1}};
static const int CORNERS[8][3] = {{1, 1, 1},  {1, 1, -1},  {1, -1, 1},  {1, -1, -1},
                                  {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};

/* coarse rhs gc = 4 FW(r) at coarse interior nodes (r has zero faces) */
static void restrict4(int Sf, const float* r, float* gc) {
  const int Sc = (Sf - 1) / 2 + 1, nc = Sc - 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (int I = 1; I < nc; ++I)
    for (int J = 1; J < nc; ++J)
      for (int K = 1; K < nc; ++K) {
        const int i = 2 * I, j = 2 * J, k = 2 * K;
#define AT(o) r[IDX(Sf, i + (o)[0], j + (o)[1], k + (o)[2])]
        float fs = AT(FACES[0]);
        for (int q = 1; q < 6; ++q) fs = fs + AT(FACES[q]);
        float es = AT(EDGES[0]);
        for (int q = 1; q < 12; ++q) es = es + AT(EDGES[q]);
        float cs = AT(CORNERS[0]);
        for (int q = 1; q < 8; ++q) cs = cs + AT(CORNERS[q]);
#undef AT
        const float val = 8.0f * r[IDX(Sf, i, j, k)] + 4.0f * fs + 2.0f * es + cs;
        gc[IDX(Sc, I, J, K)] = val * 0.0625f;
      }
}

/* x += P e on the fine interior: linear along k, then j, then i */
static void prolong_correct(int Sf, float* x, const float* e) {
  const int Sc = (Sf - 1) / 2 + 1, n = Sf - 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 1; i < n; ++i)
    for (int j = 1; j < n; ++j)
      for (int k = 1; k < n; ++k) {
        const int ci = i / 2, cj = j / 2, ck = k / 2;
        float fi[2];
        for (int a = 0; a < 2; ++a) {
          const int ii = ci + a;
          if (ii >= Sc) {
            fi[a] = 0.0f;
            continue;
          }
          float fj[2];
          for (int b = 0; b < 2; ++b) {
            const int jj = cj + b;
            if (jj >= Sc) {
              fj[b] = 0.0f;
              continue;
            }
            const float* row = e + IDX(Sc, ii, jj, 0);
            fj[b] = (k & 1) ? 0.5f * (row[ck] + row[ck + 1]) : row[ck];
          }
          fi[a] = (j & 1) ? 0.5f * (fj[0] + fj[1]) : fj[0];
        }
        const float corr = (i & 1) ? 0.5f * (fi[0] + fi[1]) : fi[0];
        const size_t id = IDX(Sf, i, j, k);
        x[id] = x[id] + corr;
      }
}

/* exact FP64 solve of A e = g on the coarsest grid (m interior per axis) */
static void exact_solve(int Sc, const float* g, float* e) {
  const int m = Sc - 2;
  const size_t m3 = (size_t)m * m * m;
  double* Smat = (double*)malloc(sizeof(double) * m * m);
  double* lam = (double*)malloc(sizeof(double) * m);
  double* A = (double*)malloc(sizeof(double) * m3);
  double* B = (double*)malloc(sizeof(double) * m3);
  const double nrm = sqrt(2.0 / (m + 1));
  for (int p = 0; p < m; ++p) {
    lam[p] = 2.0 - 2.0 * cos((p + 1) * M_PI / (m + 1));
    for (int q = 0; q < m; ++q) Smat[p * m + q] = nrm * sin(M_PI * (p + 1) * (q + 1) / (m + 1));
  }
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < m; ++b)
      for (int c = 0; c < m; ++c) A[((size_t)a * m + b) * m + c] = g[IDX(Sc, a + 1, b + 1, c + 1)];
  /* S along each axis (S is symmetric and orthonormal), scale, S again */
  for (int pass = 0; pass < 2; ++pass) {
    for (int axis = 0; axis < 3; ++axis) {
      for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b)
          for (int c = 0; c < m; ++c) {
            double s = 0.0;
            for (int q = 0; q < m; ++q) {
              const size_t src = axis == 0 ? ((size_t)q * m + b) * m + c
                                 : axis == 1 ? ((size_t)a * m + q) * m + c
                                             : ((size_t)a * m + b) * m + q;
              const int p = axis == 0 ? a : axis == 1 ? b : c;
              s += Smat[p * m + q] * A[src];
            }
            B[((size_t)a * m + b) * m + c] = s;
          }
      double* t = A;
      A = B;
      B = t;
    }
    if (pass == 0)
      for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b)
          for (int c = 0; c < m; ++c) A[((size_t)a * m + b) * m + c] /= lam[a] + lam[b] + lam[c];
  }
  memset(e, 0, sizeof(float) * (size_t)Sc * Sc * Sc);
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < m; ++b)
      for (int c = 0; c < m; ++c) e[IDX(Sc, a + 1, b + 1, c + 1)] = (float)A[((size_t)a * m + b) * m + c];
  free(Smat);
  free(lam);
  free(A);
  free(B);
}

typedef struct {
  int levels, nu1, nu2;
  float a, c;
  int* S;       /* nodes per axis of each level */
  float** x;    /* iterate (levels >= 1) */
  float** x2;   /* sweep partner */
  float** g;    /* right-hand side */
  float** r;    /* residual */
} MG;

static float* zalloc(size_t n) { return (float*)calloc(n, sizeof(float)); }

/* one V-cycle on level l for iterate xin -> result in *xout (ping-pong) */
static void vcycle(MG* mg, int l, float** xcur, float** xalt) {
  const int S = mg->S[l];
  const size_t S3 = (size_t)S * S * S;
  if (l == mg->levels - 1) {
    exact_solve(S, mg->g[l], *xcur);
    return;
  }
  for (int s = 0; s < mg->nu1; ++s) {
    smooth(S, *xcur, *xalt, mg->g[l], mg->a, mg->c);
    float* t = *xcur;
    *xcur = *xalt;
    *xalt = t;
  }
  memset(mg->r[l], 0, sizeof(float) * S3);
  residual(S, *xcur, mg->g[l], mg->r[l]);
  const int Sc = mg->S[l + 1];
  memset(mg->g[l + 1], 0, sizeof(float) * (size_t)Sc * Sc * Sc);
  restrict4(S, mg->r[l], mg->g[l + 1]);
  memset(mg->x[l + 1], 0, sizeof(float) * (size_t)Sc * Sc * Sc);
  memset(mg->x2[l + 1], 0, sizeof(float) * (size_t)Sc * Sc * Sc);
  vcycle(mg, l + 1, &mg->x[l + 1], &mg->x2[l + 1]);
  prolong_correct(S, *xcur, mg->x[l + 1]);
  for (int s = 0; s < mg->nu2; ++s) {
    smooth(S, *xcur, *xalt, mg->g[l], mg->a, mg->c);
    float* t = *xcur;
    *xcur = *xalt;
    *xalt = t;
  }
}

static void itp_update(int S, const float* U, float* Un, const float* K, const float* P, const int32_t* L,
                       float half_dv, float dtv) {
  const int n = S - 1;
  const size_t S2 = (size_t)S * S;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 1; i < n; ++i)
    for (int j = 1; j < n; ++j)
      for (int k = 1; k < n; ++k) {
        const size_t id = IDX(S, i, j, k);
        const float c = U[id];
        const int32_t l = L[id];
        const size_t nb[6] = {id + S2, id - S2, id + S, id - S, id + 1, id - 1};
        float t[6];
        for (int q = 0; q < 6; ++q) t[q] = (L[nb[q]] == l) ? U[nb[q]] : c;
        const float lap = t[0] + t[1] + t[2] + t[3] + t[4] + t[5] - 6.0f * c;
        Un[id] = c + half_dv * lap + dtv * (K[id] - 2.0f * P[id]) * c;
      }
}

static void normalize(int S, float* U, const int32_t* L, const float* Z, int n_dom, float h3) {
  const int n = S - 1;
  double* norms = (double*)calloc((size_t)n_dom, sizeof(double));
#pragma omp parallel
  {
    double* loc = (double*)calloc((size_t)n_dom, sizeof(double));
#pragma omp for collapse(2) schedule(static)
    for (int i = 1; i < n; ++i)
      for (int j = 1; j < n; ++j)
        for (int k = 1; k < n; ++k) {
          const size_t id = IDX(S, i, j, k);
          int l = L[id];
          l = l < 0 ? 0 : (l > n_dom - 1 ? n_dom - 1 : l);
          loc[l] += (double)((U[id] * U[id]) * h3);
        }
#pragma omp critical
    for (int d = 0; d < n_dom; ++d) norms[d] += loc[d];
    free(loc);
  }
  float* sc = (float*)malloc(sizeof(float) * (size_t)n_dom);
  for (int d = 0; d < n_dom; ++d)
    sc[d] = (Z[d] > 0.0f && norms[d] > 1e-20) ? (float)sqrt((double)Z[d] / norms[d]) : 1.0f;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 1; i < n; ++i)
    for (int j = 1; j < n; ++j)
      for (int k = 1; k < n; ++k) {
        const size_t id = IDX(S, i, j, k);
        const int l = L[id];
        if (l >= 0 && l < n_dom) U[id] = U[id] * sc[l];
      }
  free(sc);
  free(norms);
}

/*
 * Steps first+1 .. first+n of the loop body: V-cycle every poisson_interval
 * steps (rhs (two_pi*(U*U))*h2), ITP update, normalise every norm_interval
 * steps.  U and P are updated in place (S^3 floats, zero faces).  Returns 0,
 * or -1 on a bad argument / allocation failure.
 */
int cs_steps(int NN, int levels, int nu1, int nu2, float a, float c, float h2, float h3, float two_pi,
             float half_dv, float dtv, float* U, float* P, const float* K, const int32_t* L, const float* Z,
             int n_dom, long first, int n, int poisson_interval, int norm_interval) {
  if (levels < 2 || NN % (1 << (levels - 1)) || n < 0 || poisson_interval < 1 || norm_interval < 1) return -1;
  const int S = NN + 1;
  const size_t S3 = (size_t)S * S * S;
  MG mg;
  mg.levels = levels;
  mg.nu1 = nu1;
  mg.nu2 = nu2;
  mg.a = a;
  mg.c = c;
  mg.S = (int*)malloc(sizeof(int) * levels);
  mg.x = (float**)calloc(levels, sizeof(float*));
  mg.x2 = (float**)calloc(levels, sizeof(float*));
  mg.g = (float**)calloc(levels, sizeof(float*));
  mg.r = (float**)calloc(levels, sizeof(float*));
  int ok = mg.S && mg.x && mg.x2 && mg.g && mg.r;
  for (int l = 0; ok && l < levels; ++l) {
    mg.S[l] = (NN >> l) + 1;
    const size_t n3 = (size_t)mg.S[l] * mg.S[l] * mg.S[l];
    mg.g[l] = zalloc(n3);
    mg.r[l] = zalloc(n3);
    mg.x2[l] = zalloc(n3);
    mg.x[l] = l ? zalloc(n3) : NULL;
    ok = mg.g[l] && mg.r[l] && mg.x2[l] && (l == 0 || mg.x[l]);
  }
  float* Un = ok ? zalloc(S3) : NULL;
  if (!Un) ok = 0;
  if (ok) {
    memcpy(mg.x2[0], P, sizeof(float) * S3);
    float* pc = P;
    float* pa = mg.x2[0];
    float* uc = U;
    float* ua = Un;
    for (long step = first + 1; step <= first + n; ++step) {
      if (step % poisson_interval == 0) {
        const int nn = S - 1;
        float* g0 = mg.g[0];
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 1; i < nn; ++i)
          for (int j = 1; j < nn; ++j)
            for (int k = 1; k < nn; ++k) {
              const size_t id = IDX(S, i, j, k);
              g0[id] = (two_pi * (uc[id] * uc[id])) * h2;
            }
        vcycle(&mg, 0, &pc, &pa);
      }
      memcpy(ua, uc, sizeof(float) * S3); /* faces */
      itp_update(S, uc, ua, K, pc, L, half_dv, dtv);
      float* t = uc;
      uc = ua;
      ua = t;
      if (step % norm_interval == 0) normalize(S, uc, L, Z, n_dom, h3);
    }
    if (uc != U) memcpy(U, uc, sizeof(float) * S3);
    if (pc != P) memcpy(P, pc, sizeof(float) * S3);
  }
  free(Un);
  for (int l = 0; l < levels; ++l) {
    if (mg.g) free(mg.g[l]);
    if (mg.r) free(mg.r[l]);
    if (mg.x2) free(mg.x2[l]);
    if (mg.x) free(mg.x[l]);
  }
  free(mg.S);
  free(mg.x);
  free(mg.x2);
  free(mg.g);
  free(mg.r);
  return ok ? 0 : -1;
}
