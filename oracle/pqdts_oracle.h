/*
 * CPU oracle for the pqdts two-metric projected truncated-Newton solver.
 *
 * TEST INFRASTRUCTURE ONLY. This is a sequential, plain-C restatement of the
 * reference solver (pqdts.f90 module `basics` and the minimisation loop of
 * `program pqdts`, plus condat_simplexproj.c). It exists to check the HIP
 * product path and to serve as a timed CPU baseline. Only tests/, the smoke
 * entry point and the CPU-baseline leg of the benchmark may load it; the
 * product library never links or calls it.
 *
 * Layout: row-major. X[i*N + n] is Pi(i+1, n+1) of the reference (which stores
 * X(M,N) column-major); O and P are D x N row-major, so O[d*N + n] is the
 * reference's O(n+1, d+1). All arithmetic is IEEE double, summed sequentially
 * in the loop order of the reference.
 */
#ifndef PQDTS_ORACLE_H
#define PQDTS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

/* buffers owned by the oracle context */
enum orc_slot {
  ORC_X = 0,     /* iterate X (M x N)                               */
  ORC_G = 1,     /* gradient / CG residual, the reference's dLdX     */
  ORC_HP = 2,    /* Hessian-vector product output (bd in the CG loop) */
  ORC_XTMP = 3,  /* CG direction / line-search trial point          */
  ORC_SN = 4,    /* Newton step                                     */
  ORC_Z = 5,     /* preconditioned residual                         */
  ORC_NSLOT = 6
};

/* events reported during a solve; values as printed by the reference */
enum orc_event {
  ORC_EV_ITER = 0,     /* v: oiter iiter L xmin conv normder mode alpha dlalpha (pqdts.f90:2131) */
  ORC_EV_PCG = 1,      /* v: j sqrt(b2) eps b1 v2                               (pqdts.f90:2084) */
  ORC_EV_LS = 2,       /* v: i L alpha df xmin                                  (pqdts.f90:429/474) */
  ORC_EV_ACTIVE = 3,   /* v: nact M*N                                           (pqdts.f90:2002) */
  ORC_EV_STOP1 = 4,    /* v: dlalpha                                            (pqdts.f90:1956) */
  ORC_EV_ITER0 = 5,    /* v: L0 sqrt(normder)                                   (pqdts.f90:1914) */
  ORC_EV_LAMBDA = 6    /* v: i lambda min(G+lambda)                             (pqdts.f90:1059) */
};
typedef void (*orc_event_cb)(void *user, int kind, const double *v, int nv);

/* COO inputs are 0-based int32 as in the data/ row and col files; duplicates: last wins. */
orc_t *orc_create(int64_t M, int64_t N, int64_t D,
                  int64_t nF, const int32_t *Frow, const int32_t *Fcol, const double *Fval,
                  int64_t nP, const int32_t *Prow, const int32_t *Pcol, const double *Pval,
                  double gamma);
void orc_destroy(orc_t *c);

double *orc_buf(orc_t *c, int slot);
double *orc_O(orc_t *c);
double *orc_diag(orc_t *c);         /* Bdiag output, length M (same for every column) */
int32_t *orc_maxr(orc_t *c);         /* 0-based column, or -1                          */
int8_t *orc_ik(orc_t *c);
void orc_set_fullspace(orc_t *c, int full);
int orc_get_fullspace(const orc_t *c);
void orc_set_mode(orc_t *c, int mode);

/* operators; `x`/`y` are slot ids */
void orc_redO(orc_t *c, int x);                         /* O = F X          */
double orc_L(orc_t *c, int x);                          /* objective        */
void orc_dL(orc_t *c, int x, int g);                    /* gradient         */
void orc_BF(orc_t *c, int dd, int bx);                  /* Hessian * dd     */
void orc_Bdiag(orc_t *c);                               /* Hessian diagonal */
void orc_doproj(orc_t *c, int x);                       /* Condat per row   */
double orc_doproj2(orc_t *c, int x);                    /* clip+reconstruct */
double orc_conv_test(orc_t *c, int x, int g, int print);
double orc_pdot(orc_t *c, int x, int y);
void orc_lalpha_min2(orc_t *c, double *alpha, double *dlalpha);
void orc_lalpha_min_auglag(orc_t *c, double *alpha, double *dlalpha);
void orc_transform(orc_t *c);                           /* reduced-space maxr transform */

/* the reference's Condat FFI, restated: y in, x out (x must not alias y) */
void orc_simplexproj_condat(const double *y, double *x, int n);
/* textbook sort-based projection (Held-Wolfe-Crowder / Duchi et al.) for cross-checks */
void orc_simplexproj_sort(const double *y, double *x, int n);

/* full solve from the current X (set X first); returns the final conv */
double orc_solve(orc_t *c, int mode, int maxiter, double conv_thres, int start_stage,
                 int output, orc_event_cb cb, void *user);

#ifdef __cplusplus
}
#endif
#endif
