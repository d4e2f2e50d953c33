/*
 * pqdts_amd -- C ABI of the MI355X (gfx950) two-metric projected truncated-Newton
 * solver for quantum detector tomography:
 *
 *   min_Pi ||P - F Pi||_F^2 + gamma * sum_n sum_i (Pi[i,n] - Pi[i+1,n])^2,
 *   every row of Pi (M x N) on the probability simplex.
 *
 * The reference (pc2/pqdts) is one Fortran program whose solver lives in
 * `module basics` (pqdts.f90:24-1189) and in the minimisation loop of
 * `program pqdts` (pqdts.f90:1908-2168). Its only in-process foreign call is
 * simplexproj_condat_ (condat_simplexproj.c:60). This header exposes the same
 * operators, one entry point per reference subroutine, on a device-resident
 * problem context, plus the whole minimisation loop. The Fortran drop-in driver
 * (pqdts_omp.x) binds these through iso_c_binding.
 *
 * Conventions
 *   - every function returns 0 on success, or a nonzero code; pqdts_last_error()
 *     then describes the failure (a HIP error, a bad argument, a missing GPU).
 *   - host arrays are plain pointers owned by the caller; the library copies.
 *   - M x N host arrays are row-major (C order, element (i,n) at i*N+n) when
 *     layout == PQDTS_ROW_MAJOR, or column-major like the reference's X(M,N)
 *     when layout == PQDTS_COL_MAJOR.
 *   - COO index arrays are 0-based int32, exactly the data/[FP]_{row,col}.bin
 *     files (pqdts.f90:1080-1107); duplicate (row,col) entries: the last wins
 *     (pqdts.f90:1359-1363, 1533-1535).
 *   - calls on one context are stream-ordered and not thread-safe; different
 *     contexts are independent.
 */
#ifndef PQDTS_AMD_H
#define PQDTS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pqdts_ctx pqdts_ctx;

/* device-resident M x N buffers of a context (the reference's arrays) */
enum pqdts_slot {
  PQDTS_X = 0,     /* iterate X                    (pqdts.f90:1540)        */
  PQDTS_G = 1,     /* gradient / CG residual, dLdX (pqdts.f90:1542)        */
  PQDTS_HP = 2,    /* Hessian * direction, bd in the CG loop (2037)        */
  PQDTS_XTMP = 3,  /* CG direction / line-search point, Xtmp (1548)        */
  PQDTS_SN = 4,    /* Newton step sn (1550)                                */
  PQDTS_Z = 5      /* preconditioned residual z (1552)                     */
};

enum pqdts_layout { PQDTS_ROW_MAJOR = 0, PQDTS_COL_MAJOR = 1 };

/* Progress events of pqdts_solve, with the values the reference prints. */
enum pqdts_event {
  PQDTS_EV_ITER = 0,    /* oiter iiter L xmin conv ||der||^2 mode alpha dlalpha  (pqdts.f90:2131) */
  PQDTS_EV_PCG = 1,     /* j sqrt(b2) eps b1 v2                                (pqdts.f90:2084) */
  PQDTS_EV_LS = 2,      /* i L alpha df xmin is_min2 (1: Lalpha_min2, 0: auglag) (pqdts.f90:429/474) */
  PQDTS_EV_ACTIVE = 3,  /* nact M*N                                            (pqdts.f90:2002) */
  PQDTS_EV_STOP1 = 4,   /* dlalpha                                             (pqdts.f90:1956) */
  PQDTS_EV_ITER0 = 5,   /* L0 sqrt(||der||^2)                                  (pqdts.f90:1914) */
  PQDTS_EV_LAMBDA = 6,  /* i lambda min(G+lambda)                              (pqdts.f90:1059) */
  PQDTS_EV_STAGE = 7,   /* oiter conv: end of a stage, before its output file  (pqdts.f90:2139) */
  /* polish (pqdts_set_polish; no counterpart in the reference) */
  PQDTS_EV_POLISH = 8,  /* iteration active_entries changed_entries cg_rounds step_max          */
                        /* (method 1: then largest column block, band of S, factorisation failed) */
  PQDTS_EV_POLISHED = 9 /* status(1 converged, 2 cap) iterations cg_rounds conv L               */
};
typedef void (*pqdts_event_cb)(void *user, int kind, const double *values, int nvalues);

/* Message describing the last failure on this thread ("" if none). */
const char *pqdts_last_error(void);

/* Number of visible GPUs (0 when none); never fails. */
int pqdts_device_count(void);
/* Compute units of the current device (256 on MI355X). No counterpart in the
 * reference. */
int pqdts_device_cu_count(void);

/* Builds a problem on the current GPU from COO F (D x M) and P (D x N):
 * the reference's F ingest (pqdts.f90:1318-1365), P densify (1527-1538) and
 * work-array allocation (1540-1563). X starts at 0, maxr at -1, full space. */
int pqdts_create(int64_t M, int64_t N, int64_t D,
                 int64_t nF, const int32_t *F_row, const int32_t *F_col, const double *F_val,
                 int64_t nP, const int32_t *P_row, const int32_t *P_col, const double *P_val,
                 double gamma, pqdts_ctx **out);
int pqdts_destroy(pqdts_ctx *ctx);

/* host <-> device copies of a slot */
int pqdts_put(pqdts_ctx *ctx, int slot, const double *host, int layout);
int pqdts_get(pqdts_ctx *ctx, int slot, double *host, int layout);
/* O = F*X as left by the last objective/gradient/HVP evaluation, D x N row-major */
int pqdts_get_O(pqdts_ctx *ctx, double *host);
/* maxr (pqdts.f90:93), 0-based column index per row or -1 */
int pqdts_set_maxr(pqdts_ctx *ctx, const int32_t *maxr);
int pqdts_get_maxr(pqdts_ctx *ctx, int32_t *maxr);
/* active-set mask ik (pqdts.f90:92), M x N row-major */
int pqdts_get_ik(pqdts_ctx *ctx, int8_t *ik);
/* the reference's global tfullspace flag (pqdts.f90:97) */
int pqdts_set_fullspace(pqdts_ctx *ctx, int full);
int pqdts_get_fullspace(pqdts_ctx *ctx, int *full);

/* ---- operators: one per reference subroutine --------------------------- */
/* L (pqdts.f90:917-973): objective at slot x; reconstructs x's implicit column in
 * place when not in full space. */
int pqdts_L(pqdts_ctx *ctx, int x, double *val);
/* dL (pqdts.f90:975-1032): gradient of slot x into slot g. */
int pqdts_dL(pqdts_ctx *ctx, int x, int g);
/* BF (pqdts.f90:246-342): Hessian times slot dd into slot bx. */
int pqdts_BF(pqdts_ctx *ctx, int dd, int bx);
/* Bdiag (pqdts.f90:208-244): Hessian diagonal; every column of the reference's
 * M x N result is the same, so this returns the M-vector (diag may be NULL). */
int pqdts_Bdiag(pqdts_ctx *ctx, double *diag);
/* doproj (pqdts.f90:1120-1140): Euclidean projection of every row of slot x
 * onto the simplex, the reference's simplexproj_condat_ per row. */
int pqdts_doproj(pqdts_ctx *ctx, int x);
/* doproj2 (pqdts.f90:1142-1164): clip negatives, rebuild the implicit column,
 * xmin = min(0, min_i x[i,maxr(i)]). */
int pqdts_doproj2(pqdts_ctx *ctx, int x, double *xmin);
/* conv_test (pqdts.f90:1034-1067): KKT measure; sets full space as the reference does. */
int pqdts_conv_test(pqdts_ctx *ctx, int x, int g, double *conv);
/* pdot (pqdts.f90:589-614) */
int pqdts_pdot(pqdts_ctx *ctx, int x, int y, double *val);
/* BLAS-1 of the reference (pqdts.f90:485-587), on slots */
int pqdts_paxpy(pqdts_ctx *ctx, double alpha, int x, int y);                /* y += alpha x        */
int pqdts_pgaxpy(pqdts_ctx *ctx, double alpha, int x, double beta, int y);  /* y = beta y + alpha x */
int pqdts_pgaxpy2(pqdts_ctx *ctx, double alpha, int x, double beta, int y, int z); /* z = beta y + alpha x */
int pqdts_pset(pqdts_ctx *ctx, double alpha, int x);
int pqdts_pcopy(pqdts_ctx *ctx, int x, int y);
int pqdts_pscale(pqdts_ctx *ctx, double alpha, int x);
int pqdts_pcopy_scale(pqdts_ctx *ctx, double alpha, int x, int y);
/* reduced-space transform of X (pqdts.f90:1970-1980) */
int pqdts_transform(pqdts_ctx *ctx);
/* Line searches along slot SN from slot X, trial points in XTMP:
 * Lalpha_min2 (pqdts.f90:394-438) and Lalpha_min_auglag (pqdts.f90:440-483).
 * On return XTMP holds the last trial point evaluated. */
int pqdts_lalpha_min2(pqdts_ctx *ctx, double *alpha, double *dlalpha);
int pqdts_lalpha_min_auglag(pqdts_ctx *ctx, double *alpha, double *dlalpha);

/* The minimisation (pqdts.f90:1908-2167) from the current X: mode 1 (projected
 * gradient) or 2 (two-metric truncated Newton), stages start_stage..2, at most
 * maxiter Newton iterations per stage, stop when conv < eps. `cb` (may be NULL)
 * receives every progress event in the reference's print order; output >= 3
 * also reports the per-row lambda lines. A PQDTS_EV_STAGE event closes each
 * stage after its final conv_test; the callback may then call pqdts_get on the
 * same context to write the per-stage file. *conv receives the last KKT value. */
int pqdts_solve(pqdts_ctx *ctx, int mode, int maxiter, double eps, int start_stage, int output,
                pqdts_event_cb cb, void *user, double *conv);

/* Polish (no counterpart in the reference). The reference stops at conv < eps
 * (pqdts.f90:1034-1067, 2134); conv weights each KKT residual by X, so on
 * problems with weak curvature (small gamma) points far apart pass the same eps.
 * With max_outer > 0, every later pqdts_solve and batch admission of this
 * context continues from the minimisation's result with a primal-dual
 * active-set Newton iteration on the reduced problem (the implicit column of
 * each row is its largest entry): an entry is held at 0 when G - d X > 0 (G the
 * reduced gradient, d the reduced Hessian diagonal of its row); the free entries
 * take a full Newton step, solved by CG with the Jacobi preconditioner to the
 * relative residual tol / (the last step's largest entry), kept within
 * [cg_rtol, 1e-3] (cg_rtol for the first step). It stops when the active set
 * repeats, the last step is <= 100 tol and that solve's error estimate (its
 * relative residual times its step) is <= tol: a KKT point of the strictly
 * convex problem, i.e. its unique minimiser to about tol (status 1). It also stops after max_outer iterations
 * or, with budget > 0, once its CG rounds reach budget times the PCG rounds of
 * the minimisation itself (status 2); X is then the minimisation's result,
 * unchanged. Gradient, active set, steps and the CG scalars are FP64; the
 * Hessian products are FP64 unless fp32 = 1 and the solve is mixed precision.
 * max_outer = 0 turns it off (the default). */
int pqdts_set_polish(pqdts_ctx *ctx, int max_outer, double cg_rtol, double tol, double budget, int fp32);
/* The polish's method (no counterpart in the reference): 0 (the default) as
 * above; 1 direct face solves. Method 1 iterates in the full space with the row
 * sums as constraints and multipliers lambda: an entry is held at 0 when
 * G + lambda - d X > -1e-13 d (d the diagonal of K = 2 F^T F + 2 gamma Lap; a row keeps
 * its largest entry free). On each new active set it forms the inverse of K on
 * the free rows of every column and the Schur complement of the row sums (M x M,
 * banded), factorises it (band Cholesky) and applies 2 corrections from the
 * current point, each with a fresh gradient (iterative refinement), all in FP64;
 * a column whose free rows did not change keeps its inverse.
 * It stops when the active set repeats and the last correction is <= 100 tol
 * (status 1: the unique minimiser), or with status 2 (X unchanged: the
 * minimisation's result) after max_outer iterations, when M > 3072 or a column
 * has more than 1024 free rows, or when the factorisation fails. cg_rtol, budget
 * and fp32 do not apply. Its buffers, about 8 M N + 8 M^2 + 16 sum_n m_n^2 bytes
 * (m_n the free rows of column n), are held while the polish runs and are not
 * part of pqdts_footprint. */
int pqdts_set_polish_method(pqdts_ctx *ctx, int method);
/* outcome of the last polish on ctx: stats[0..2] = status (0 not run, 1 converged,
 * 2 stopped at a cap), iterations, CG rounds; *nact active entries, *step
 * the last step's largest entry (each pointer may be NULL) */
int pqdts_polish_result(pqdts_ctx *ctx, int *stats, int64_t *nact, double *step);

/* The named sections timed on this context so far, under the reference's timer
 * names (pqdts.f90:103-186): min, L, L_O, dL, dL_O, dL_dLdX, BF, BF_O, BF_BX, CG,
 * active_dLdX, Lalpha_min, doproj, doproj2, pdot and the BLAS-1 names. Sections
 * the reference times but this library fuses into a neighbouring kernel
 * (precond1/2 into the CG kernels, active into the product's epilogue, L_gam,
 * dL_gam, BF_gam into the products, Bdiag computed once per context) are not
 * listed. Seconds are host wall-clock time around the asynchronous launches
 * unless pqdts_device_timers is on. Returns the number of timers; fills up to
 * `cap` entries. */
int pqdts_timers(pqdts_ctx *ctx, int cap, char (*names)[24], int *counts, double *seconds);
/* on != 0: time every section on the device, between HIP events recorded on the
 * context's stream at its start and end (the device time of the section's
 * launches); 0: host wall clock. Costs two event records per section. */
int pqdts_device_timers(pqdts_ctx *ctx, int on);

/* Per-kernel device timing with HIP events on the context's stream: `on`
 * enables it (1: every launch; k > 1: every k-th operator launch, a sample of
 * the durations whose flops and bytes are those of the sampled launches; 0: off)
 * and clears the totals; pqdts_kernel_stats returns the number of
 * kernel names and fills launch counts and total milliseconds. Band products:
 * "spmm_F" = O = F*V, "spmmT_F" = 2 F^T W + stencil, "hvp_G" = 2 G p with the
 * Gram matrix (names ending in 32: the FP32 mixed-precision variants).
 * Streaming kernels: "cg_init", "cg_update", "cg_dir", "foldval_gram",
 * "project", "conv_rows", "axpby". No counterpart in the reference (its timers
 * are the wall-clock ones above, pqdts.f90:103-186). */
int pqdts_kernel_timing(pqdts_ctx *ctx, int on);
/* flops, bytes (each may be NULL): algorithmic work of the timed launches.
 * flops = 2 x stored elements of the banded operand x N per problem and band
 * product; bytes = each operand read once and each result written once (the
 * banded operand at its stored precision, M x N and D x N arrays in FP64). */
int pqdts_kernel_stats(pqdts_ctx *ctx, int cap, char (*names)[24], int *counts, double *ms, double *flops,
                       double *bytes);

/* Sizes of a context: info[0..13] = M, N, D, stored F elements (sum of row span
 * lengths, the reference's Frows storage), packed elements of F's band tiles,
 * #F tiles, #F^T tiles, Gram path active (0/1), packed elements of G's band
 * tiles, stored elements of G's row bands, packed elements of F^T's band
 * tiles, device bytes held by the problem, split of the Gram product over band
 * chunks (1 = none; see pqdts_set_hvp_split), most band chunks of one G tile. */
int pqdts_info(pqdts_ctx *ctx, int64_t *info);

/* How BF (and every CG step) applies the Hessian: 1 = the reference's two
 * products 2 F^T (F p) (pqdts.f90:272-288); 2 = one product with the banded Gram
 * matrix G = F^T F, formed once on the GPU; 0 = automatic (G when its band tiles
 * cost less than 0.8 of the two products). Same operator, different rounding.
 * No counterpart in the reference. */
int pqdts_set_hvp_path(pqdts_ctx *ctx, int path);

/* Split of the Gram product over the band chunks of each output tile: split k
 * has k workgroups multiply disjoint chunk ranges of one tile into k partial
 * M x N arrays (k x M x N doubles of device memory), which a second pass sums
 * in a fixed order before the stencil, fold, mask and dot. Shortens the longest
 * dependency chain of products too small to fill the GPU. 0 = by the problem's
 * shape (the default; FP64 products only), 1 = no split, 2..16 fixed (FP64 and
 * FP32 products). Same operator, different
 * rounding; a problem's split never depends on what it is batched with. Fails
 * while the context is in a batch. No counterpart in the reference. */
int pqdts_set_hvp_split(pqdts_ctx *ctx, int split);

/* ---- batched / served form (no counterpart in the reference) -------------
 * A batch advances many independent problems together: every operator of the
 * minimisation is one launch over all admitted problems (uneven M, N, D, gamma
 * allowed), and one host synchronisation per step serves all of them. Each
 * problem follows exactly the reference's algorithm on its own data (its own
 * stages, CG exits and line searches); a problem's result does not depend on
 * what else is in the batch. Contexts stay owned by the caller and must not be
 * used through the single-context calls while admitted. A context is in at most
 * one slot at a time (admitting it again fails until pqdts_batch_release);
 * pqdts_set_hvp_path and pqdts_set_hvp_split fail while it is admitted;
 * pqdts_destroy of an admitted context first releases its slot. */
typedef struct pqdts_batch pqdts_batch;
int pqdts_batch_create(int capacity, pqdts_batch **out);
int pqdts_batch_destroy(pqdts_batch *b);
/* Admits ctx (its X is the starting point) into a free slot. precision: 0 = FP64
 * throughout; 1 = mixed: the Hessian products inside CG in FP32 MFMA with
 * per-chunk FP64 accumulation, everything else (objective, gradient, line
 * search, stopping tests, projections) in FP64. */
int pqdts_batch_admit(pqdts_batch *b, pqdts_ctx *ctx, int mode, int maxiter, double eps, int start_stage,
                      int output, int precision, pqdts_event_cb cb, void *user, int *slot);
/* One Newton iteration (plus stage switches) of every unfinished problem.
 * Returns the number of finished problems still admitted (their slots in
 * finished[0..cap)), or -1. */
int pqdts_batch_step(pqdts_batch *b, int *finished, int cap);
/* stats[0..3] = finished, Newton iterations, PCG iterations, line-search trials */
int pqdts_batch_result(pqdts_batch *b, int slot, double *conv, int *stats);
int pqdts_batch_release(pqdts_batch *b, int slot);
int pqdts_batch_active(pqdts_batch *b);
/* Restricts an empty batch's launches to a subset of the device's compute units:
 * bit i of mask[i / 32] selects CU i (nwords 32-bit words). Lets a server keep
 * small, latency-bound problems in their own batch on CUs that large problems'
 * batches do not use. No counterpart in the reference. */
int pqdts_batch_set_cu_mask(pqdts_batch *b, const uint32_t *mask, int nwords);
int pqdts_batch_kernel_timing(pqdts_batch *b, int on);
int pqdts_batch_kernel_stats(pqdts_batch *b, int cap, char (*names)[24], int *counts, double *ms,
                             double *flops, double *bytes);

/* ---- request server (served form; no counterpart in the reference) --------
 * Queues requests (a problem as COO F and P, as pqdts_create takes them, with its
 * solve parameters) and solves them on the GPU with continuous batching: each of
 * `nlanes` size classes (by M*N: lane g holds lane_bounds[g-1] <= M*N <
 * lane_bounds[g]) is one batch on its own stream, advanced by its own host thread.
 * A lane admits the head of its queue (FIFO) when a slot is free and the
 * request's device footprint fits the budget: the bytes pqdts_footprint reports
 * (persistent + transient) are held from admission until the context exists,
 * then its actual bytes until it finishes; a finished problem's Pi is copied to
 * the host and its device memory freed. The device bytes of all admitted
 * problems together never exceed budget_bytes (0 = 90 % of the free device
 * memory at creation; each lane's own engine arrays, about 2.4 KB per slot, are
 * not counted). lane_cus (may be NULL): per lane a consecutive CU range
 * lo, hi, or -1, -1 for all CUs. A request larger than the budget is refused at
 * submission. Every problem computes the same bits as solved alone. A lane's
 * Newton iterations replay captured hipGraphs (the launches before the line
 * search and after the step; PQDTS_GRAPHS in the environment overrides: 0 eager,
 * 1 CG rounds in blocks of 8, 2 this). */
typedef struct pqdts_server pqdts_server;
int pqdts_server_create(int nlanes, const int64_t *lane_bounds, const int32_t *lane_cus, int capacity,
                        int64_t budget_bytes, pqdts_server **out);
/* stops the lanes (unfinished requests are dropped) and frees everything */
int pqdts_server_destroy(pqdts_server *s);
/* copies the request; X0 (row-major M x N) may be NULL for Pi = 1/N; mode,
 * maxiter, eps as pqdts_solve (from stage 1); precision as pqdts_batch_admit */
int pqdts_server_submit(pqdts_server *s, int64_t M, int64_t N, int64_t D, int64_t nF, const int32_t *F_row,
                        const int32_t *F_col, const double *F_val, int64_t nP, const int32_t *P_row,
                        const int32_t *P_col, const double *P_val, double gamma, int mode, int maxiter,
                        double eps, int precision, const double *X0, int64_t *id);
/* ids of requests finished since the last call (at most cap); waits up to
 * timeout_ms (-1: until one finishes or none is left) when none is ready.
 * Returns the count, or -1. */
int pqdts_server_wait(pqdts_server *s, int64_t *ids, int cap, int timeout_ms);
/* result of a finished request, which is then forgotten: X (row-major M x N,
 * may be NULL), conv, stats[7] (finished, Newton iterations, PCG iterations,
 * line-search trials, polish status / iterations / CG rounds as
 * pqdts_polish_result), times[3] seconds (queued, admitted to done, total).
 * Returns the request's own status (nonzero if its solve failed). */
int pqdts_server_result(pqdts_server *s, int64_t id, double *X, double *conv, int *stats, double *times);
/* polish settings (as pqdts_set_polish) for the requests submitted from now on */
int pqdts_server_set_polish(pqdts_server *s, int max_outer, double cg_rtol, double tol, double budget, int fp32);
/* polish method (as pqdts_set_polish_method) for the requests submitted from now on */
int pqdts_server_set_polish_method(pqdts_server *s, int method);
/* info[0..7] = queued, running, finished not yet waited for, device bytes held
 * now, peak device bytes held, budget, submitted, completed */
int pqdts_server_info(pqdts_server *s, int64_t *info);
/* Device bytes a problem will hold (persistent: its arrays, band packs, Gram
 * matrix, mixed-precision copies, Gram-split partials) and, on top, while its
 * context is created (transient: the dense matrices of the Gram formation), from
 * F's pattern alone; what the server's admission counts. */
int pqdts_footprint(int64_t M, int64_t N, int64_t D, int64_t nF, const int32_t *F_row, const int32_t *F_col,
                    int precision, int64_t *persistent, int64_t *transient);

/* simplexproj_condat_ (condat_simplexproj.c:60) over many rows at once: rows of
 * length n from y (host, row-major) projected onto the simplex into x (host).
 * x must not alias y, as in the reference. */
int pqdts_simplexproj_rows(const double *y, double *x, int64_t rows, int64_t n);

/* pqdts_simplexproj_rows without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc). y and x are
 * DEVICE pointers to rows x n doubles (row-major); x == y projects in place,
 * otherwise y is copied to x first and must not overlap it. The work is
 * enqueued on `stream` (a hipStream_t; NULL = the default stream) and the call
 * returns without waiting. The projection's per-block partials live in a static
 * device array of the library (2 MiB per GPU, held until the process ends), so
 * at most 1048576 rows per call; larger calls fail with an error. Calls that
 * overlap on one GPU race on that array (its contents, never the projected
 * rows). Same kernel arithmetic as pqdts_simplexproj_rows: bit-identical rows. */
int pqdts_simplexproj_rows_noalloc(const double *y, double *x, int64_t rows, int64_t n, void *stream);

/* The reference's only foreign symbol, same name, signature and contract
 * (condat_simplexproj.c:60, called from pqdts.f90:1133 through an implicit
 * Fortran interface, hence the trailing underscore and by-reference length):
 * projects y[0..*plength) onto the simplex into x; y is not modified, x is the
 * output and also the scratch, so x must not overlap y. Computed on the GPU
 * with Condat's filter in the reference's operation order: x is bit-identical
 * to the reference's. No status return (the reference has none): a failure is
 * printed to stderr and aborts. */
void simplexproj_condat_(double *y, double *x, int *plength);

#ifdef __cplusplus
}
#endif
#endif
