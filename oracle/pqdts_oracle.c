/*
 * CPU oracle for the pqdts solver -- TEST INFRASTRUCTURE ONLY (see
 * pqdts_oracle.h). Sequential restatement of the reference; each function
 * cites the pqdts.f90 / condat_simplexproj.c lines it follows.
 *
 * Single-precision literals of the reference are reproduced as widened floats
 * ((double)1e-5f etc.), because the reference writes them as default-real
 * constants assigned to real64 variables.
 */
#include "pqdts_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* pqdts.f90:354/404/450  dx = 1e-5 (default real) */
#define LS_DX ((double)1e-5f)
/* pqdts.f90:424/430  xmin .ge. -1e-12 (default real) */
#define LS_FEAS ((double)-1e-12f)
/* pqdts.f90:1207  tactivetol = 1e-8 (default real) */
#define ACTIVE_TOL ((double)1e-8f)
/* pqdts.f90:1955  abs(dlalpha) .lt. 1e-4 (default real) */
#define STAGE1_EXIT ((double)1e-4f)

struct orc {
  int64_t M, N, D;
  /* F rows as dense spans, as the reference's Frows (pqdts.f90:1345-1363):
   * row d covers columns lo[d] .. lo[d]+len[d]-1, values at val[off[d]..] */
  int64_t *lo, *len, *off;
  double *val;
  int *has;          /* row has >= 1 COO entry (block_red(d)%c > 0, pqdts.f90:946) */
  double *P;         /* D x N */
  double gamma;
  double *buf[ORC_NSLOT];
  double *O;         /* D x N */
  double *diag;      /* M     */
  int32_t *maxr;     /* M     */
  int8_t *ik;        /* M x N */
  int full;          /* tfullspace */
  int mode;
  orc_event_cb cb;
  void *user;
};

static void emit(orc_t *c, int kind, const double *v, int nv) {
  if (c->cb) c->cb(c->user, kind, v, nv);
}

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  return p;
}

/* F ingest, pqdts.f90:1318-1365: per-row span [min col, max col], zero fill,
 * duplicate (row,col) entries overwritten in file order (1359-1363). Done in
 * O(nnz) instead of the reference's O(D*nnz) scan; the result is identical. */
orc_t *orc_create(int64_t M, int64_t N, int64_t D,
                  int64_t nF, const int32_t *Frow, const int32_t *Fcol, const double *Fval,
                  int64_t nP, const int32_t *Prow, const int32_t *Pcol, const double *Pval,
                  double gamma) {
  orc_t *c = (orc_t *)xcalloc(1, sizeof(orc_t));
  c->M = M;
  c->N = N;
  c->D = D;
  c->gamma = gamma;
  c->lo = (int64_t *)xcalloc(D, sizeof(int64_t));
  c->len = (int64_t *)xcalloc(D, sizeof(int64_t));
  c->off = (int64_t *)xcalloc(D + 1, sizeof(int64_t));
  c->has = (int *)xcalloc(D, sizeof(int));
  int64_t *hi = (int64_t *)xcalloc(D, sizeof(int64_t));
  for (int64_t d = 0; d < D; d++) {
    c->lo[d] = INT64_MAX;
    hi[d] = -1;
  }
  for (int64_t e = 0; e < nF; e++) {
    int64_t d = Frow[e], k = Fcol[e];
    if (k < c->lo[d]) c->lo[d] = k;
    if (k > hi[d]) hi[d] = k;
    c->has[d] = 1;
  }
  for (int64_t d = 0; d < D; d++) {
    c->len[d] = c->has[d] ? hi[d] - c->lo[d] + 1 : 0;
    if (!c->has[d]) c->lo[d] = 0;
    c->off[d + 1] = c->off[d] + c->len[d];
  }
  c->val = (double *)xcalloc(c->off[D], sizeof(double));
  for (int64_t e = 0; e < nF; e++) {
    int64_t d = Frow[e];
    c->val[c->off[d] + (Fcol[e] - c->lo[d])] = Fval[e];
  }
  free(hi);
  /* P densify, pqdts.f90:1527-1535: P(Pj,Pi) = Pd, last duplicate wins */
  c->P = (double *)xcalloc(D * N, sizeof(double));
  for (int64_t e = 0; e < nP; e++) c->P[(int64_t)Prow[e] * N + Pcol[e]] = Pval[e];
  for (int s = 0; s < ORC_NSLOT; s++) c->buf[s] = (double *)xcalloc(M * N, sizeof(double));
  c->O = (double *)xcalloc(D * N, sizeof(double));
  c->diag = (double *)xcalloc(M, sizeof(double));
  c->maxr = (int32_t *)xcalloc(M, sizeof(int32_t));
  for (int64_t i = 0; i < M; i++) c->maxr[i] = -1;
  c->ik = (int8_t *)xcalloc(M * N, sizeof(int8_t));
  c->full = 1;
  c->mode = 2;
  return c;
}

void orc_destroy(orc_t *c) {
  if (!c) return;
  free(c->lo);
  free(c->len);
  free(c->off);
  free(c->has);
  free(c->val);
  free(c->P);
  for (int s = 0; s < ORC_NSLOT; s++) free(c->buf[s]);
  free(c->O);
  free(c->diag);
  free(c->maxr);
  free(c->ik);
  free(c);
}

double *orc_buf(orc_t *c, int slot) { return c->buf[slot]; }
double *orc_O(orc_t *c) { return c->O; }
double *orc_diag(orc_t *c) { return c->diag; }
int32_t *orc_maxr(orc_t *c) { return c->maxr; }
int8_t *orc_ik(orc_t *c) { return c->ik; }
void orc_set_fullspace(orc_t *c, int full) { c->full = full; }
int orc_get_fullspace(const orc_t *c) { return c->full; }
void orc_set_mode(orc_t *c, int mode) { c->mode = mode; }

static double rowsum(const double *x, int64_t N) {
  double s = 0.0;
  for (int64_t n = 0; n < N; n++) s += x[n];
  return s;
}

/* X(j,maxr) = 1 - (sum(X(j,:)) - X(j,maxr)): pqdts.f90:931-933, 986-988, 1047-1049 */
static void reconstruct(orc_t *c, double *X) {
  const int64_t M = c->M, N = c->N;
  for (int64_t j = 0; j < M; j++) {
    double *row = X + j * N;
    int32_t r = c->maxr[j];
    row[r] = 1.0 - (rowsum(row, N) - row[r]);
  }
}

/* redO local part, pqdts.f90:724-736 */
static void redO(orc_t *c, const double *X) {
  const int64_t N = c->N, D = c->D;
  for (int64_t d = 0; d < D; d++) {
    double *o = c->O + d * N;
    if (!c->has[d]) {
      for (int64_t n = 0; n < N; n++) o[n] = 0.0;
      continue;
    }
    const double *f = c->val + c->off[d];
    const double *xs = X + c->lo[d] * N;
    for (int64_t n = 0; n < N; n++) {
      double s = 0.0;
      for (int64_t k = 0; k < c->len[d]; k++) s += f[k] * xs[k * N + n];
      o[n] = s;
    }
  }
}

void orc_redO(orc_t *c, int x) { redO(c, c->buf[x]); }

/* L, pqdts.f90:917-973 */
static double Lfun(orc_t *c, double *X) {
  const int64_t M = c->M, N = c->N, D = c->D;
  if (!c->full) reconstruct(c, X);
  redO(c, X);
  double val = 0.0;
  for (int64_t d = 0; d < D; d++) {
    if (!c->has[d]) continue;
    const double *o = c->O + d * N, *p = c->P + d * N;
    double s = 0.0;
    for (int64_t n = 0; n < N; n++) {
      double r = o[n] - p[n];
      s += r * r;
    }
    val += s;
  }
  if (c->gamma != 0.0) {
    for (int64_t n = 0; n < N; n++)
      for (int64_t j = 0; j + 1 < M; j++) {
        double t = X[j * N + n] - X[(j + 1) * N + n];
        val += c->gamma * (t * t);
      }
  }
  return val;
}

double orc_L(orc_t *c, int x) { return Lfun(c, c->buf[x]); }

/* dL, pqdts.f90:975-1032 */
static void dLfun(orc_t *c, double *X, double *G) {
  const int64_t M = c->M, N = c->N, D = c->D;
  if (!c->full) reconstruct(c, X);
  memset(G, 0, sizeof(double) * M * N);
  redO(c, X);
  for (int64_t d = 0; d < D; d++) {
    if (!c->has[d]) continue;
    const double *f = c->val + c->off[d];
    const double *o = c->O + d * N, *p = c->P + d * N;
    for (int64_t k = 0; k < c->len[d]; k++) {
      double *g = G + (c->lo[d] + k) * N;
      for (int64_t n = 0; n < N; n++) g[n] += 2.0 * (o[n] - p[n]) * f[k];
    }
  }
  if (c->gamma != 0.0) {
    const double g2 = 2.0 * c->gamma;
    for (int64_t n = 0; n < N; n++)
      for (int64_t j = 0; j < M; j++) {
        if (j + 1 < M) G[j * N + n] = G[j * N + n] + g2 * (X[j * N + n] - X[(j + 1) * N + n]);
        if (j > 0) G[j * N + n] = G[j * N + n] - g2 * (X[(j - 1) * N + n] - X[j * N + n]);
      }
  }
  if (!c->full) {
    for (int64_t j = 0; j < M; j++) {
      double *g = G + j * N;
      double gm = g[c->maxr[j]];
      for (int64_t n = 0; n < N; n++) g[n] -= gm;
    }
  }
}

void orc_dL(orc_t *c, int x, int g) { dLfun(c, c->buf[x], c->buf[g]); }

/* Bdiag, pqdts.f90:208-244. The reference fills an M x N array whose columns are
 * identical; the oracle keeps one column. */
void orc_Bdiag(orc_t *c) {
  const int64_t M = c->M, D = c->D;
  double *bd = c->diag;
  for (int64_t k = 0; k < M; k++) bd[k] = 0.0;
  for (int64_t d = 0; d < D; d++) {
    if (!c->has[d]) continue;
    const double *f = c->val + c->off[d];
    for (int64_t k = 0; k < c->len[d]; k++) bd[c->lo[d] + k] += 2.0 * (f[k] * f[k]);
  }
  if (c->gamma != 0.0) {
    for (int64_t k = 0; k + 1 < M; k++) bd[k] += 2.0 * c->gamma;
    for (int64_t k = 1; k < M; k++) bd[k] += 2.0 * c->gamma;
  }
  if (!c->full)
    for (int64_t k = 0; k < M; k++) bd[k] *= 2.0;
}

/* BF (Hessian times vector), pqdts.f90:246-342 */
static void BFfun(orc_t *c, const double *dd, double *BX) {
  const int64_t M = c->M, N = c->N, D = c->D;
  const int32_t *maxr = c->maxr;
  if (!c->full) {
    memcpy(BX, dd, sizeof(double) * M * N);
    for (int64_t j = 0; j < M; j++) BX[j * N + maxr[j]] = -(rowsum(dd + j * N, N) - dd[j * N + maxr[j]]);
    redO(c, BX);
  } else {
    redO(c, dd);
  }
  memset(BX, 0, sizeof(double) * M * N);
  for (int64_t d = 0; d < D; d++) {
    if (!c->has[d]) continue;
    const double *f = c->val + c->off[d];
    const double *o = c->O + d * N;
    for (int64_t k = 0; k < c->len[d]; k++) {
      double *bx = BX + (c->lo[d] + k) * N;
      for (int64_t n = 0; n < N; n++) bx[n] += 2.0 * f[k] * o[n];
    }
  }
  if (c->gamma != 0.0) {
    const double g2 = 2.0 * c->gamma;
    if (!c->full) {
      /* maxr-aware stencil on the expanded vector, pqdts.f90:294-317 */
      double *others = (double *)xcalloc(M, sizeof(double));
      for (int64_t j = 0; j < M; j++) others[j] = rowsum(dd + j * N, N) - dd[j * N + maxr[j]];
      for (int64_t i = 0; i < N; i++)
        for (int64_t j = 0; j < M; j++) {
          double *b = BX + j * N + i;
          if (j + 1 < M) {
            *b = *b + g2 * (dd[j * N + i] - dd[(j + 1) * N + i]);
            if (i == maxr[j]) *b = *b + g2 * (-dd[j * N + i] - others[j]);
            if (i == maxr[j + 1]) *b = *b + g2 * (+dd[(j + 1) * N + i] + others[j + 1]);
          }
          if (j > 0) {
            *b = *b - g2 * (dd[(j - 1) * N + i] - dd[j * N + i]);
            if (i == maxr[j - 1]) *b = *b - g2 * (-dd[(j - 1) * N + i] - others[j - 1]);
            if (i == maxr[j]) *b = *b - g2 * (+dd[j * N + i] + others[j]);
          }
        }
      free(others);
    } else {
      /* pqdts.f90:319-330 */
      for (int64_t i = 0; i < N; i++)
        for (int64_t j = 0; j < M; j++) {
          double *b = BX + j * N + i;
          if (j + 1 < M) *b = *b + g2 * (dd[j * N + i] - dd[(j + 1) * N + i]);
          if (j > 0) *b = *b - g2 * (dd[(j - 1) * N + i] - dd[j * N + i]);
        }
    }
  }
  if (!c->full) {
    /* fold, pqdts.f90:334-340 */
    for (int64_t j = 0; j < M; j++) {
      double *b = BX + j * N;
      double bm = b[maxr[j]];
      for (int64_t n = 0; n < N; n++) b[n] -= bm;
    }
  }
}

void orc_BF(orc_t *c, int dd, int bx) { BFfun(c, c->buf[dd], c->buf[bx]); }

/* Condat's projection onto the simplex {x >= 0, sum x = 1}, restated from the
 * published algorithm (L. Condat, "Fast projection onto the simplex and the l1
 * ball", Math. Program. 158, 2016, Fig. 2) in the operation order of
 * condat_simplexproj.c:60-95, so that tau is bit-identical:
 *   pass 1: online filter; a candidate list `act` with running pivot tau,
 *           restarted (older entries parked in `park`) when tau drops to y-1;
 *   pass 2: parked entries re-examined newest-first and put in front of `act`;
 *   pass 3: repeated sweeps drop entries <= tau until nothing changes. */
void orc_simplexproj_condat(const double *y, double *x, int n) {
  if (n <= 0) return;
  double *park = (double *)xcalloc((size_t)n, sizeof(double));
  double *act = (double *)xcalloc((size_t)n, sizeof(double));
  int npark = 0, nact = 1;
  double tau = y[0] - 1.0;
  act[0] = y[0];
  for (int i = 1; i < n; i++) {
    if (!(y[i] > tau)) continue;
    double t = tau + (y[i] - tau) / (double)(nact + 1);
    if (t <= y[i] - 1.0) {
      /* restart: everything collected so far is parked */
      for (int k = 0; k < nact; k++) park[npark++] = act[k];
      tau = y[i] - 1.0;
      act[0] = y[i];
      nact = 1;
    } else {
      tau = t;
      act[nact++] = y[i];
    }
  }
  if (npark > 0) {
    /* newest-first over the parked values; accepted ones are prepended, so the
     * sweep order below is: accepted parked values in original order, then act */
    double *front = (double *)xcalloc((size_t)n, sizeof(double));
    int nfront = 0;
    for (int k = npark - 1; k >= 0; k--) {
      if (park[k] > tau) {
        front[nfront++] = park[k];
        nact++;
        tau += (park[k] - tau) / (double)nact;
      }
    }
    double *merged = (double *)xcalloc((size_t)n, sizeof(double));
    int m = 0;
    for (int k = nfront - 1; k >= 0; k--) merged[m++] = front[k];
    for (int k = 0; k < nact - nfront; k++) merged[m++] = act[k];
    memcpy(act, merged, sizeof(double) * (size_t)m);
    free(front);
    free(merged);
  }
  for (;;) {
    int before = nact, kept = 0;
    for (int i = 0; i < before; i++) {
      if (act[i] > tau) {
        act[kept++] = act[i];
      } else {
        int remaining = (before - 1 - i) + kept;
        tau += (tau - act[i]) / (double)remaining;
      }
    }
    nact = kept;
    if (kept == before) break;
  }
  for (int i = 0; i < n; i++) x[i] = y[i] > tau ? y[i] - tau : 0.0;
  free(park);
  free(act);
}

static int cmp_desc(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return (x < y) - (x > y);
}

/* Sort-based projection: tau = (sum_{k<=K} u_k - 1)/K for the largest K with
 * u_K > (sum_{k<=K} u_k - 1)/K, u sorted descending (Held, Wolfe & Crowder 1974;
 * Duchi et al. 2008). Independent of Condat; used to cross-check it. */
void orc_simplexproj_sort(const double *y, double *x, int n) {
  double *u = (double *)xcalloc((size_t)n, sizeof(double));
  memcpy(u, y, sizeof(double) * (size_t)n);
  qsort(u, (size_t)n, sizeof(double), cmp_desc);
  double s = 0.0, tau = 0.0;
  for (int k = 0; k < n; k++) {
    s += u[k];
    double t = (s - 1.0) / (double)(k + 1);
    if (u[k] > t) tau = t;
  }
  for (int i = 0; i < n; i++) x[i] = y[i] > tau ? y[i] - tau : 0.0;
  free(u);
}

/* doproj, pqdts.f90:1120-1140 */
static void doproj(orc_t *c, double *X) {
  const int64_t M = c->M, N = c->N;
  double *tmp = (double *)xcalloc(N, sizeof(double));
  for (int64_t i = 0; i < M; i++) {
    orc_simplexproj_condat(X + i * N, tmp, (int)N);
    memcpy(X + i * N, tmp, sizeof(double) * N);
  }
  free(tmp);
}

void orc_doproj(orc_t *c, int x) { doproj(c, c->buf[x]); }

/* doproj2, pqdts.f90:1142-1164 */
static double doproj2(orc_t *c, double *X) {
  const int64_t M = c->M, N = c->N;
  double xmin = 0.0;
  for (int64_t i = 0; i < M; i++) {
    double *row = X + i * N;
    for (int64_t j = 0; j < N; j++)
      if (row[j] < 0.0) row[j] = 0.0;
    int32_t r = c->maxr[i];
    row[r] = 1.0 - (rowsum(row, N) - row[r]);
    if (row[r] < xmin) xmin = row[r];
  }
  return xmin;
}

double orc_doproj2(orc_t *c, int x) { return doproj2(c, c->buf[x]); }

static double pdot(const orc_t *c, const double *x, const double *y) {
  /* pqdts.f90:601-605: column sums accumulated column by column */
  const int64_t M = c->M, N = c->N;
  double val = 0.0;
  for (int64_t n = 0; n < N; n++) {
    double s = 0.0;
    for (int64_t j = 0; j < M; j++) s += x[j * N + n] * y[j * N + n];
    val += s;
  }
  return val;
}

double orc_pdot(orc_t *c, int x, int y) { return pdot(c, c->buf[x], c->buf[y]); }

/* conv_test, pqdts.f90:1034-1067 (forces tfullspace = .true., 1052) */
static double conv_test(orc_t *c, double *X, double *G, int print) {
  const int64_t M = c->M, N = c->N;
  if (!c->full) reconstruct(c, X);
  c->full = 1;
  dLfun(c, X, G);
  double conv = 0.0;
  for (int64_t i = 0; i < M; i++) {
    const double *g = G + i * N, *x = X + i * N;
    double gmin = g[0];
    for (int64_t n = 1; n < N; n++)
      if (g[n] < gmin) gmin = g[n];
    double lambda = fmax(0.0, -gmin);
    double s = 0.0;
    for (int64_t n = 0; n < N; n++) {
      double t = (g[n] + lambda) * x[n];
      s += t * t;
    }
    conv += s;
    if (print) {
      double mn = g[0] + lambda;
      for (int64_t n = 1; n < N; n++)
        if (g[n] + lambda < mn) mn = g[n] + lambda;
      double v[3] = {(double)(i + 1), lambda, mn};
      emit(c, ORC_EV_LAMBDA, v, 3);
    }
  }
  return sqrt(conv / ((double)N * (double)M));
}

double orc_conv_test(orc_t *c, int x, int g, int print) { return conv_test(c, c->buf[x], c->buf[g], print); }

/* pgaxpy2(m,n,1,X,a,S,xtmp): xtmp = a*S + 1*X, pqdts.f90:515-528 */
static void trial_point(orc_t *c, double a) {
  const int64_t MN = c->M * c->N;
  const double *X = c->buf[ORC_X], *S = c->buf[ORC_SN];
  double *T = c->buf[ORC_XTMP];
  for (int64_t e = 0; e < MN; e++) T[e] = a * S[e] + 1.0 * X[e];
}

/* Lalpha_min2, pqdts.f90:394-438 (tproj = .true. as called from the driver) */
void orc_lalpha_min2(orc_t *c, double *alpha_out, double *dlalpha) {
  double *T = c->buf[ORC_XTMP];
  double xmin;
  trial_point(c, 0.0);
  doproj2(c, T);
  double L0 = Lfun(c, T);
  trial_point(c, 0.0 + LS_DX);
  doproj2(c, T);
  double Lp = Lfun(c, T);
  double df = (Lp - L0) / LS_DX;
  *dlalpha = df;
  double alpha = 1.0, lm = DBL_MAX, alphamin = 0.0;
  for (int i = 1; i <= 100; i++) {
    trial_point(c, alpha);
    xmin = doproj2(c, T);
    Lp = Lfun(c, T);
    if (Lp < lm && xmin >= LS_FEAS) {
      lm = Lp;
      alphamin = alpha;
    }
    double v[5] = {(double)i, Lp, alpha, df, xmin};
    emit(c, ORC_EV_LS, v, 5);
    if (Lp > L0 + 0.1 * alpha * df || xmin < LS_FEAS)
      alpha = alpha * 0.75;
    else
      break;
  }
  *alpha_out = alphamin;
}

/* Lalpha_min_auglag, pqdts.f90:440-483 (tproj = .true.) */
void orc_lalpha_min_auglag(orc_t *c, double *alpha_out, double *dlalpha) {
  double *T = c->buf[ORC_XTMP];
  trial_point(c, 0.0);
  doproj(c, T);
  double L0 = Lfun(c, T);
  trial_point(c, 0.0 + LS_DX);
  doproj(c, T);
  double Lp = Lfun(c, T);
  double df = (Lp - L0) / LS_DX;
  *dlalpha = df;
  double alpha = 1.0, lm = DBL_MAX, alphamin = 0.0;
  for (int i = 1; i <= 100; i++) {
    trial_point(c, alpha);
    doproj(c, T);
    Lp = Lfun(c, T);
    if (Lp < lm) {
      lm = Lp;
      alphamin = alpha;
    }
    double v[5] = {(double)i, Lp, alpha, df, 0.0};
    emit(c, ORC_EV_LS, v, 5);
    if (Lp > L0 + 0.1 * alpha * df)
      alpha = alpha * 0.75;
    else
      break;
  }
  *alpha_out = alphamin;
}

/* reduced-space transform, pqdts.f90:1970-1980 */
void orc_transform(orc_t *c) {
  const int64_t M = c->M, N = c->N;
  double *X = c->buf[ORC_X];
  for (int64_t j = 0; j < M; j++) {
    double *row = X + j * N;
    if (c->maxr[j] != -1) row[c->maxr[j]] = 1.0 - (rowsum(row, N) - row[c->maxr[j]]);
    int32_t best = 0;
    for (int64_t n = 1; n < N; n++)
      if (row[n] > row[best]) best = (int32_t)n;
    c->maxr[j] = best;
    row[best] = 1.0;
  }
}

/* Minimisation driver, pqdts.f90:1908-2167 (mode 1 and mode 2). */
double orc_solve(orc_t *c, int mode, int maxiter, double conv_thres, int start_stage,
                 int output, orc_event_cb cb, void *user) {
  const int64_t M = c->M, N = c->N, MN = M * N;
  double *X = c->buf[ORC_X], *G = c->buf[ORC_G], *HP = c->buf[ORC_HP];
  double *XT = c->buf[ORC_XTMP], *SN = c->buf[ORC_SN], *Z = c->buf[ORC_Z];
  double *bd = c->diag;
  c->cb = cb;
  c->user = user;
  c->mode = mode;
  double conv = 0.0, alpha = 0.0, dlalpha, norm_der, L0, xmin;

  /* pqdts.f90:1908-1915 */
  L0 = Lfun(c, X);
  dLfun(c, X, G);
  norm_der = pdot(c, G, G);
  {
    double v[2] = {L0, sqrt(norm_der)};
    emit(c, ORC_EV_ITER0, v, 2);
  }
  for (int oiter = start_stage; oiter <= 2; oiter++) {
    dlalpha = -DBL_MAX;
    for (int iiter = 1; iiter <= maxiter; iiter++) {
      if (mode == 1) {
        /* projected gradient, pqdts.f90:1934-1951 */
        c->full = 0;
        for (int64_t j = 0; j < M; j++) {
          double *row = X + j * N;
          int32_t best = 0;
          for (int64_t n = 1; n < N; n++)
            if (row[n] > row[best]) best = (int32_t)n;
          c->maxr[j] = best;
          row[best] = 1.0;
        }
        dLfun(c, X, G);
        for (int64_t e = 0; e < MN; e++) SN[e] = G[e] * -1.0;
        orc_lalpha_min2(c, &alpha, &dlalpha);
      } else {
        if ((fabs(dlalpha) < STAGE1_EXIT || dlalpha > 0) && oiter == 1) {
          emit(c, ORC_EV_STOP1, &dlalpha, 1);
          break;
        }
        c->full = (oiter == 1);
        if (!c->full) {
          orc_transform(c);
        } else {
          for (int64_t j = 0; j < M; j++) c->maxr[j] = -1;
        }
        dLfun(c, X, G);
        /* active set, pqdts.f90:1986-2005 */
        int64_t nact = 0;
        for (int64_t e = 0; e < MN; e++) {
          if (X[e] < ACTIVE_TOL && G[e] > 0) {
            G[e] = 0.0;
            c->ik[e] = 1;
            nact++;
          } else {
            c->ik[e] = 0;
          }
        }
        {
          double v[2] = {(double)nact, (double)MN};
          emit(c, ORC_EV_ACTIVE, v, 2);
        }
        norm_der = pdot(c, G, G);
        memset(SN, 0, sizeof(double) * MN);
        memset(Z, 0, sizeof(double) * MN);
        for (int64_t e = 0; e < MN; e++) G[e] = G[e] * -1.0; /* r = -g */
        orc_Bdiag(c);
        for (int64_t j = 0; j < M; j++)
          for (int64_t n = 0; n < N; n++) Z[j * N + n] = G[j * N + n] / sqrt(bd[j]);
        double eps = fmin(0.5, sqrt(sqrt(norm_der))) * sqrt(norm_der);
        memcpy(XT, Z, sizeof(double) * MN);
        /* PCG, pqdts.f90:2036-2090 */
        for (int j = 1; j <= 50; j++) {
          BFfun(c, XT, HP);
          for (int64_t e = 0; e < MN; e++)
            if (c->ik[e] == 1) HP[e] = 0.0;
          double b1 = pdot(c, G, Z);
          double v2 = pdot(c, XT, HP);
          double a = b1 / v2;
          for (int64_t e = 0; e < MN; e++) SN[e] = SN[e] + a * XT[e];
          for (int64_t e = 0; e < MN; e++) G[e] = G[e] + (-a) * HP[e];
          orc_Bdiag(c);
          for (int64_t jj = 0; jj < M; jj++)
            for (int64_t n = 0; n < N; n++) Z[jj * N + n] = G[jj * N + n] / sqrt(bd[jj]);
          v2 = pdot(c, G, Z);
          double beta = v2 / b1;
          for (int64_t e = 0; e < MN; e++) XT[e] = beta * XT[e] + 1.0 * Z[e];
          double b2 = pdot(c, G, G);
          double v[5] = {(double)j, sqrt(b2), eps, b1, v2};
          emit(c, ORC_EV_PCG, v, 5);
          if (sqrt(b2) <= eps) break;
        }
        /* two-metric diagonal step, pqdts.f90:2092-2104 */
        dLfun(c, X, G);
        orc_Bdiag(c);
        for (int64_t k = 0; k < M; k++)
          for (int64_t i = 0; i < N; i++)
            if (c->ik[k * N + i] == 1 && c->maxr[k] != i) SN[k * N + i] = -G[k * N + i] / bd[k];
        if (c->full)
          orc_lalpha_min_auglag(c, &alpha, &dlalpha);
        else
          orc_lalpha_min2(c, &alpha, &dlalpha);
      }
      /* step, projection, re-evaluation, pqdts.f90:2111-2128 */
      if (alpha != 0.0)
        for (int64_t e = 0; e < MN; e++) X[e] = X[e] + alpha * SN[e];
      xmin = 0.0;
      if (c->full)
        doproj(c, X);
      else
        xmin = doproj2(c, X);
      L0 = Lfun(c, X);
      dLfun(c, X, G);
      norm_der = pdot(c, G, G);
      conv = conv_test(c, X, G, 0);
      {
        double v[9] = {(double)oiter, (double)iiter, L0, xmin, conv, sqrt(norm_der),
                       (double)mode, alpha, dlalpha};
        emit(c, ORC_EV_ITER, v, 9);
      }
      if (conv < conv_thres) break;
    }
    conv = conv_test(c, X, G, output >= 3);
  }
  c->cb = NULL;
  c->user = NULL;
  return conv;
}
