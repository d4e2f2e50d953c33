// Direct face solves of the polish (no counterpart in the reference).
//
// The polish method 1 is a primal-dual active-set Newton method in the full space
// with the row-sum constraints kept as constraints. The objective's Hessian acts on
// every column n of Pi with the same M x M matrix K = 2 F^T F + 2 gamma Lap. On a
// face (the active entries held at 0) the minimiser satisfies, per column n with
// free rows I_n,
//
//     K[I_n, I_n] x_n + lambda[I_n] = b_n,     sum_n x[i, n] = 1 for every row i,
//
// so x_n = K_n^-1 (b_n - lambda[I_n]) and the row multipliers solve
// S lambda = r with the Schur complement S = sum_n E_n K_n^-1 E_n^T (M x M).
// Every free set of a column is small (tens of rows) and S is banded (two rows are
// coupled only through a column where both are free), so one face solve costs
// the inverses of the N column blocks, S, and a band Cholesky factorisation of S.
//
// Each solve is a correction from the current point: with the gradient g of the
// objective there and the current multipliers lambda,
//
//     y_n = K_n^-1 (g + lambda)[I_n],  S delta = -(1 - X 1) - sum_n E_n y_n,
//     dx_n = -(y_n + K_n^-1 delta[I_n]),  X += dx,  lambda += delta,
//
// repeated (iterative refinement): the gradient is the solver's own band product,
// so the fixed point is the face minimiser of the objective the minimisation
// used, and the explicit inverses only need to be good enough for the refinement
// to contract. Solving for the increment delta keeps the rounding of the S solve
// relative to delta, not to lambda.
//
// The active set follows the rule mu - d X > -theta d with mu = g + lambda, d = diag K
// and theta = 1e-13 (a margin above the gradient's rounding, so that an entry at 0
// whose mu is 0 to rounding, as at a degenerate minimiser, stays put instead of
// flipping with the rounding); a row whose entries would all be active keeps its
// largest entry free. When the rule returns the face just solved, the point
// satisfies the KKT conditions of the strictly convex problem to that margin: it
// is the minimiser.
//
// Every sum runs in a fixed order (per column over its free rows in increasing
// row order, per row over its free columns in increasing column order), so a
// problem computes the same bits alone and in any batch.
#pragma once

namespace pq {

// Per-problem buffers of the direct polish (device pointers).
struct Face {
  int *pos;        // M x N: index of entry (i, n) in column n's list of free rows, -1 if active
  int *crow;       // N x M: crow[n * M + k] = k-th free row of column n
  int *cnt;        // N: free rows per column
  int64_t *koff;   // N + 1: offset of column n's block in kb
  double *kb;      // column blocks, cnt[n]^2 each (row-major): K on the free rows, then its inverse
  int64_t *koff_old;  // the previous iteration's offsets and blocks
  double *kb_old;
  uint8_t *cdirty;    // N: the column's free set changed in this iteration's active set
  uint8_t *cfresh;    // N: the column's block is formed anew in this iteration
  double *S;       // M x M: the Schur complement, then its Cholesky factor U (S = U^T U, upper)
  double *lam;     // M row multipliers
  double *rhs;     // M: right-hand side of the S solve, then the increment delta
  int *slast;      // M: last nonzero column of each row of S
  double *cstep;   // N: largest |dx| of each column in the last correction
};

constexpr double FACE_THETA = 1e-13;  // margin of the active-set rule, relative to d
constexpr int FACE_MAX_M = 1024;     // largest column block
constexpr int FACE_MAX_ROWS = 3072;  // largest M (rows of S held in LDS)

#define FACE_PROLOGUE() \
  BATCH_PROLOGUE();     \
  const Face &FC = faces[wl.pid[wi_]]

// band chunk of the band packs (columns of a tile's band come in chunks of this many)
constexpr int FACE_KC = 16;

// K[i, j] = 2 G[i, j] + 2 gamma Lap[i, j] with G = F^T F: read from G's band pack
// where the problem holds one (row tile i / 64, columns klo .. klo + 16 nch - 1),
// else summed over the probes d in increasing order from F's row spans (rows
// kdlo[k] .. kdhi[k] of F are those whose span can hold column k)
__device__ __forceinline__ double kentry(const Prob &P, int i, int j) {
  double s = 0.0;
  if (P.bg.pack) {
    const int t = i / GT;
    const int k0 = gp(P.bg.klo)[t], kw = gp(P.bg.nch)[t] * FACE_KC;
    if (j >= k0 && j < k0 + kw) s = gp(P.bg.pack)[gp(P.bg.off)[t] + (int64_t)(j - k0) * GT + (i - t * GT)];
  } else {
    const int d0 = max(gp(P.kdlo)[i], gp(P.kdlo)[j]);
    const int d1 = min(gp(P.kdhi)[i], gp(P.kdhi)[j]);
    for (int d = d0; d <= d1; d++) {
      const int lo = gp(P.flo)[d], len = gp(P.flen)[d];
      const int ri = i - lo, rj = j - lo;
      if (ri >= 0 && ri < len && rj >= 0 && rj < len) {
        const GLOBAL double *fv = gp(P.fval) + gp(P.foff)[d];
        s += fv[ri] * fv[rj];
      }
    }
  }
  double k = 2.0 * s;
  if (P.gamma != 0.0) {
    if (i == j) {
      if (i + 1 < P.M) k += 2.0 * P.gamma;
      if (i > 0) k += 2.0 * P.gamma;
    } else if (i - j == 1 || j - i == 1) {
      k -= 2.0 * P.gamma;
    }
  }
  return k;
}

// One wave per row: lambda[j] = -g[j, n*] at the row's largest entry n* (the
// first one), which makes mu vanish there.
__global__ __launch_bounds__(BLOCK) void kf_lam0(const Prob *__restrict__ probs, const Face *__restrict__ faces,
                                                 WorkList wl) {
  FACE_PROLOGUE();
  const int j = lb * WAVES_PER_BLOCK + wave_id();
  if (j >= P.M) return;
  const int N = P.N;
  const GLOBAL double *xr = gp(P.op[0]) + (int64_t)j * N;
  const GLOBAL double *gr = gp(P.op[1]) + (int64_t)j * N;
  double v = -INFINITY;
  int idx = N;
  for (int n = lane_id(); n < N; n += WAVE) {
    const double x = xr[n];
    if (x > v) {
      v = x;
      idx = n;
    }
  }
  wave_argmax(v, idx);
  if (lane_id() == 0) gp(FC.lam)[j] = idx < N ? -gr[idx] : 0.0;
}

// One wave per row: the active set (ik) of the next face, X = 0 on it; partials of
// the entries whose activity changed (region 0) and of the active entries
// (region 1), as kb_pdas_active.
__global__ __launch_bounds__(BLOCK) void kf_active(const Prob *__restrict__ probs, const Face *__restrict__ faces,
                                                   WorkList wl) {
  __shared__ double sh[WAVES_PER_BLOCK];
  FACE_PROLOGUE();
  const int j = lb * WAVES_PER_BLOCK + wave_id();
  const int N = P.N;
  double nchg = 0.0, nact = 0.0;
  if (j < P.M) {
    const double d = gp(P.diag_full)[j], lj = gp(FC.lam)[j], th = -FACE_THETA * d;
    const int64_t base = (int64_t)j * N;
    GLOBAL double *xr = gp(P.op[0]) + base;
    const GLOBAL double *gr = gp(P.op[1]) + base;
    GLOBAL uint8_t *ikr = gp(P.ik) + base;
    // the row's largest entry, and whether any entry stays free
    double v = -INFINITY;
    int idx = N;
    double nfree = 0.0;
    for (int n = lane_id(); n < N; n += WAVE) {
      const double x = xr[n];
      if (x > v) {
        v = x;
        idx = n;
      }
      if (!((gr[n] + lj) - d * x > th)) nfree += 1.0;
    }
    wave_argmax(v, idx);
    const int keep = wave_sum(nfree) > 0.0 ? -1 : idx;
    for (int n = lane_id(); n < N; n += WAVE) {
      const double x = xr[n];
      const uint8_t a = (n != keep && (gr[n] + lj) - d * x > th) ? 1 : 0;
      if (a != ikr[n]) {
        nchg += 1.0;
        gp(FC.cdirty)[n] = 1;
      }
      ikr[n] = a;
      if (a) {
        nact += 1.0;
        xr[n] = 0.0;
      }
    }
  }
  double t = block_sum(nchg, sh);
  if (threadIdx.x == 0) gp(P.part)[lb] = t;
  t = block_sum(nact, sh);
  if (threadIdx.x == 0) gp(P.part)[(int64_t)P.pcap + lb] = t;
  if (threadIdx.x == 0 && lb == 0) gp(P.pn)[0] = gp(P.pn)[1] = nblk;
}

// One thread per column: the column's free rows in increasing order (crow), their
// positions (pos) and count (cnt); whether its block must be formed anew (cfresh,
// from the changes kf_active marked in cdirty, which starts at 1 everywhere).
__global__ __launch_bounds__(BLOCK) void kf_cols(const Prob *__restrict__ probs, const Face *__restrict__ faces,
                                                 WorkList wl) {
  FACE_PROLOGUE();
  const int n = lb * BLOCK + threadIdx.x;
  const int M = P.M, N = P.N;
  if (n >= N) return;
  const GLOBAL uint8_t *ik = gp(P.ik);
  GLOBAL int *pos = gp(FC.pos);
  GLOBAL int *cr = gp(FC.crow) + (int64_t)n * M;
  int k = 0;
  for (int i = 0; i < M; i++) {
    const int64_t e = (int64_t)i * N + n;
    if (ik[e]) {
      pos[e] = -1;
    } else {
      pos[e] = k;
      cr[k] = i;
      k++;
    }
  }
  gp(FC.cnt)[n] = k;
  gp(FC.cfresh)[n] = gp(FC.cdirty)[n];
  gp(FC.cdirty)[n] = 0;
}

// One block per problem: block offsets koff (exclusive prefix sums of cnt^2),
// the largest block (scal[i_max]) and the total size (scal[i_tot]).
__global__ __launch_bounds__(BLOCK) void kf_offsets(const Prob *__restrict__ probs, const Face *__restrict__ faces,
                                                    WorkList wl, int i_max, int i_tot) {
  __shared__ int64_t tsum[BLOCK];
  __shared__ int tmax[BLOCK];
  FACE_PROLOGUE();
  const int N = P.N, t = threadIdx.x;
  const int chunk = (N + BLOCK - 1) / BLOCK;
  const int n0 = min(N, t * chunk), n1 = min(N, n0 + chunk);
  const GLOBAL int *cnt = gp(FC.cnt);
  int64_t s = 0;
  int mx = 0;
  for (int n = n0; n < n1; n++) {
    const int c = cnt[n];
    s += (int64_t)c * c;
    mx = max(mx, c);
  }
  tsum[t] = s;
  tmax[t] = mx;
  __syncthreads();
  if (t == 0) {
    int64_t acc = 0;
    int m = 0;
    for (int u = 0; u < BLOCK; u++) {
      const int64_t v = tsum[u];
      tsum[u] = acc;
      acc += v;
      m = max(m, tmax[u]);
    }
    gp(FC.koff)[N] = acc;
    gp(P.scal)[i_max] = (double)m;
    gp(P.scal)[i_tot] = (double)acc;
  }
  __syncthreads();
  int64_t o = tsum[t];
  for (int n = n0; n < n1; n++) {
    gp(FC.koff)[n] = o;
    o += (int64_t)cnt[n] * cnt[n];
  }
}

// Size classes of the column blocks: up to FACE_LDS_S rows in a small LDS array
// (several workgroups per CU), up to FACE_LDS_L rows in one of nearly all the CU's
// 160 KB, larger ones (up to FACE_MAX_M) in global memory.
constexpr int FACE_LDS_S = 48;
constexpr int FACE_LDS_L = 140;

// One workgroup per column, for the blocks of one size class held in LDS (LM =
// FACE_LDS_S or FACE_LDS_L; kf_factor_big takes the larger ones). A column whose free
// set did not change since the last factorisation (cfresh[n] = 0) copies its
// inverse from the previous block storage (kb_old at koff_old[n]). Otherwise: the
// block K[I_n, I_n], then its inverse in place by symmetric sweeps (Gauss-Jordan
// on a symmetric positive definite matrix: pivot k is the k-th Schur complement's
// leading entry, positive); sweeping every pivot leaves -K^-1.
template <int LM>
__global__ __launch_bounds__(BLOCK) void kf_factor(const Prob *__restrict__ probs, const Face *__restrict__ faces,
                                                   WorkList wl) {
  constexpr int CAP = LM > 0 ? LM : FACE_MAX_M;
  __shared__ double lA[LM > 0 ? LM * LM : 1];
  __shared__ double cv[CAP];
  __shared__ int rows[CAP];
  FACE_PROLOGUE();
  const int n = lb, M = P.M;
  const int m = gp(FC.cnt)[n];
  constexpr int LO = LM == FACE_LDS_S ? 0 : (LM == FACE_LDS_L ? FACE_LDS_S : FACE_LDS_L);
  if (m <= LO || m > CAP) return;
  GLOBAL double *A = gp(FC.kb) + gp(FC.koff)[n];
  if (!gp(FC.cfresh)[n]) {
    const GLOBAL double *Ao = gp(FC.kb_old) + gp(FC.koff_old)[n];
    for (int e = threadIdx.x; e < m * m; e += BLOCK) A[e] = Ao[e];
    return;
  }
  const GLOBAL int *cr = gp(FC.crow) + (int64_t)n * M;
  for (int a = threadIdx.x; a < m; a += BLOCK) rows[a] = cr[a];
  __syncthreads();
  double *W = LM > 0 ? lA : (double *)A;
  const int w = wave_id(), l = lane_id();
  for (int a = w; a < m; a += WAVES_PER_BLOCK)
    for (int b = a + l; b < m; b += WAVE) {
      const double v = kentry(P, rows[a], rows[b]);
      W[a * m + b] = v;
      W[b * m + a] = v;
    }
  __syncthreads();
  for (int k = 0; k < m; k++) {
    for (int i = threadIdx.x; i < m; i += BLOCK) cv[i] = W[i * m + k];
    __syncthreads();
    const double r = 1.0 / cv[k];
    for (int i = w; i < m; i += WAVES_PER_BLOCK) {
      const double ci = cv[i];
      for (int jj = l; jj < m; jj += WAVE) {
        double v;
        if (i == k)
          v = jj == k ? -r : cv[jj] * r;
        else if (jj == k)
          v = ci * r;
        else
          v = W[i * m + jj] - (ci * cv[jj]) * r;
        W[i * m + jj] = v;
      }
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < m * m; e += BLOCK) A[e] = -W[e];
}

// Pivots per pass of the sweep of blocks above FACE_LDS_L rows (in global memory).
constexpr int FACE_SB = 16;

// One workgroup per column whose block has more than FACE_LDS_L rows (up to
// FACE_MAX_M), in global memory: the same single-pivot sweeps as kf_factor, in the
// same arithmetic, FACE_SB pivots per pass over the matrix. A pass first forms in
// LDS the pivot columns as each will be when its own sweep comes (column k0 + q
// swept by the pivots k0 .. k0 + q - 1 of the pass), then applies the FACE_SB
// sweeps to every element in registers: one read and one write per element per
// pass instead of per pivot. (Sweeping the pivots of a pass as one block instead,
// W -= W[:, B] W[B, B]^-1 W[B, :], is algebraically the same but loses accuracy on
// ill-conditioned blocks.)
__global__ __launch_bounds__(BLOCK) void kf_factor_big(const Prob *__restrict__ probs,
                                                       const Face *__restrict__ faces, WorkList wl) {
  __shared__ double C[FACE_SB * FACE_MAX_M];
  __shared__ double rq[FACE_SB];
  __shared__ int rows[FACE_MAX_M];
  FACE_PROLOGUE();
  const int n = lb, M = P.M;
  const int m = gp(FC.cnt)[n];
  if (m <= FACE_LDS_L || m > FACE_MAX_M) return;
  GLOBAL double *A = gp(FC.kb) + gp(FC.koff)[n];
  const int t = threadIdx.x, w = wave_id(), l = lane_id();
  if (!gp(FC.cfresh)[n]) {
    const GLOBAL double *Ao = gp(FC.kb_old) + gp(FC.koff_old)[n];
    for (int e = t; e < m * m; e += BLOCK) A[e] = Ao[e];
    return;
  }
  const GLOBAL int *cr = gp(FC.crow) + (int64_t)n * M;
  for (int a = t; a < m; a += BLOCK) rows[a] = cr[a];
  __syncthreads();
  for (int a = w; a < m; a += WAVES_PER_BLOCK)
    for (int b = a + l; b < m; b += WAVE) {
      const double v = kentry(P, rows[a], rows[b]);
      A[(int64_t)a * m + b] = v;
      A[(int64_t)b * m + a] = v;
    }
  __syncthreads();
  for (int k0 = 0; k0 < m; k0 += FACE_SB) {
    const int nb = min(FACE_SB, m - k0);
    // the pass's pivot columns (rows, by the exact symmetry single sweeps keep)
    for (int e = t; e < nb * m; e += BLOCK) {
      const int q = e / m, i = e - q * m;
      C[q * FACE_MAX_M + i] = A[(int64_t)(k0 + q) * m + i];
    }
    __syncthreads();
    // pivot k0 + q applied to the later pivot columns of the pass
    for (int q = 0; q + 1 < nb; q++) {
      const int k = k0 + q;
      const double r = 1.0 / C[q * FACE_MAX_M + k];
      const int cnt2 = (nb - 1 - q) * m;
      for (int e = t; e < cnt2; e += BLOCK) {
        const int q2 = q + 1 + e / m, i = e - (e / m) * m;
        const double pj = C[q * FACE_MAX_M + k0 + q2];
        const double v = i == k ? pj * r : C[q2 * FACE_MAX_M + i] - (C[q * FACE_MAX_M + i] * pj) * r;
        C[q2 * FACE_MAX_M + i] = v;
      }
      __syncthreads();
    }
    if (t < nb) rq[t] = 1.0 / C[t * FACE_MAX_M + k0 + t];
    __syncthreads();
    // every element through the pass's sweeps, in pivot order
    for (int i = w; i < m; i += WAVES_PER_BLOCK) {
      double pi[FACE_SB];
#pragma unroll
      for (int q = 0; q < FACE_SB; q++) pi[q] = q < nb ? C[q * FACE_MAX_M + i] : 0.0;
      GLOBAL double *Ar = A + (int64_t)i * m;
      for (int j = l; j < m; j += WAVE) {
        double v = Ar[j];
#pragma unroll
        for (int q = 0; q < FACE_SB; q++) {
          if (q < nb) {
            const int k = k0 + q;
            const double r = rq[q];
            if (i == k)
              v = j == k ? -r : C[q * FACE_MAX_M + j] * r;
            else if (j == k)
              v = pi[q] * r;
            else
              v = v - (pi[q] * C[q * FACE_MAX_M + j]) * r;
          }
        }
        Ar[j] = v;
      }
    }
    __syncthreads();
  }
  for (int e = t; e < m * m; e += BLOCK) A[e] = -A[e];
}

// One workgroup per row i of S: S[i, :] = sum over the free columns n of row i, in
// increasing n, of row pos(i, n) of K_n^-1 scattered to the columns' free rows;
// slast[i] = its last nonzero column. The row's free columns are taken BLOCK at a
// time (compacted in order through wave ballots).
__global__ __launch_bounds__(BLOCK) void kf_schur(const Prob *__restrict__ probs, const Face *__restrict__ faces,
                                                  WorkList wl) {
  __shared__ double srow[FACE_MAX_ROWS];
  __shared__ int ln[BLOCK], la[BLOCK];
  __shared__ int wcnt[WAVES_PER_BLOCK];
  __shared__ int lastw[WAVES_PER_BLOCK];
  FACE_PROLOGUE();
  const int i = lb, M = P.M, N = P.N;
  if (M > FACE_MAX_ROWS) return;
  for (int c = threadIdx.x; c < M; c += BLOCK) srow[c] = 0.0;
  __syncthreads();
  const GLOBAL int *posr = gp(FC.pos) + (int64_t)i * N;
  const int w = wave_id(), l = lane_id();
  for (int c0 = 0; c0 < N; c0 += BLOCK) {
    const int n = c0 + threadIdx.x;
    const int a = n < N ? posr[n] : -1;
    const uint64_t mask = __ballot(a >= 0);
    if (l == 0) wcnt[w] = __popcll(mask);
    __syncthreads();
    int off = 0, tot = 0;
    for (int u = 0; u < WAVES_PER_BLOCK; u++) {
      if (u < w) off += wcnt[u];
      tot += wcnt[u];
    }
    if (a >= 0) {
      const int q = off + __popcll(mask & ((1ull << l) - 1ull));
      ln[q] = n;
      la[q] = a;
    }
    __syncthreads();
    for (int q = 0; q < tot; q++) {
      const int nn = ln[q];
      const int m = gp(FC.cnt)[nn];
      const GLOBAL double *Ar = gp(FC.kb) + gp(FC.koff)[nn] + (int64_t)la[q] * m;
      const GLOBAL int *cr = gp(FC.crow) + (int64_t)nn * M;
      for (int b = threadIdx.x; b < m; b += BLOCK) srow[cr[b]] += Ar[b];
      __syncthreads();
    }
  }
  int last = -1;
  GLOBAL double *Sr = gp(FC.S) + (int64_t)i * M;
  for (int c = threadIdx.x; c < M; c += BLOCK) {
    const double v = srow[c];
    Sr[c] = v;
    if (v != 0.0) last = c;
  }
  for (int o = WAVE / 2; o > 0; o >>= 1) last = max(last, __shfl_xor(last, o, WAVE));
  if (l == 0) lastw[w] = last;
  __syncthreads();
  if (threadIdx.x == 0) {
    int r = lastw[0];
    for (int u = 1; u < WAVES_PER_BLOCK; u++) r = max(r, lastw[u]);
    gp(FC.slast)[i] = r;
  }
}

// Blocked band Cholesky of S: FACE_NB rows of U per step; the rows' panel (FACE_NB x
// at most FACE_PANEL columns) is held in LDS, so each step reads and writes the
// trailing band of S once for FACE_NB rank-1 updates.
constexpr int FACE_NB = 16;
constexpr int FACE_PANEL = 1216;  // widest panel: the band of S (bw) must be below it

// One workgroup per problem: S = U^T U in place (U upper, row-major) within the
// band |i - j| <= bw, bw = max_i (slast[i] - i) -> scal[i_bw]; scal[i_fail] = 1
// where a pivot is not positive or the band is too wide (the host then abandons
// this polish). Step k0 factors the diagonal block D = S[k0:k0+nb, k0:k0+nb] in LDS,
// solves the panel U[k0:k0+nb, c0:e] = D^-T S[k0:k0+nb, c0:e] (c0 = k0 + nb, e the
// band's last column for these rows), and subtracts panel^T panel from the trailing
// window S[c0:e, c0:e] (upper triangle; outside the band both factors vanish).
__global__ __launch_bounds__(BLOCK) void kf_chol(const Prob *__restrict__ probs, const Face *__restrict__ faces,
                                                 WorkList wl, int i_bw, int i_fail) {
  __shared__ double D[FACE_NB * FACE_NB];
  __shared__ double Pn[FACE_NB * FACE_PANEL];
  __shared__ int bw_w[WAVES_PER_BLOCK];
  __shared__ int fail_s;
  FACE_PROLOGUE();
  const int M = P.M;
  if (M > FACE_MAX_ROWS) return;
  const int w = wave_id(), l = lane_id(), t = threadIdx.x;
  int b = 0;
  for (int i = t; i < M; i += BLOCK) b = max(b, gp(FC.slast)[i] - i);
  for (int o = WAVE / 2; o > 0; o >>= 1) b = max(b, __shfl_xor(b, o, WAVE));
  if (l == 0) bw_w[w] = b;
  if (t == 0) fail_s = 0;
  __syncthreads();
  int bw = 0;
  for (int u = 0; u < WAVES_PER_BLOCK; u++) bw = max(bw, bw_w[u]);
  GLOBAL double *S = gp(FC.S);
  if (bw >= FACE_PANEL) {
    if (t == 0) {
      gp(P.scal)[i_bw] = (double)bw;
      gp(P.scal)[i_fail] = 1.0;
    }
    return;
  }
  for (int k0 = 0; k0 < M; k0 += FACE_NB) {
    const int nb = min(FACE_NB, M - k0);
    const int c0 = k0 + nb;
    const int e = min(M - 1, k0 + nb - 1 + bw);
    const int W = e >= c0 ? e - c0 + 1 : 0;  // panel columns c0 .. e
    // diagonal block (upper triangle), then its Cholesky factor in place
    {
      const int r = t / FACE_NB, c = t % FACE_NB;
      if (r < nb && c < nb) D[r * FACE_NB + c] = c >= r ? S[(int64_t)(k0 + r) * M + (k0 + c)] : 0.0;
    }
    __syncthreads();
    for (int k = 0; k < nb; k++) {
      const double piv = D[k * FACE_NB + k];
      const double u = sqrt(piv);
      __syncthreads();
      if (t == 0 && !(piv > 0.0)) fail_s = 1;
      if (t >= k && t < nb) D[k * FACE_NB + t] = t == k ? u : D[k * FACE_NB + t] / u;
      __syncthreads();
      {
        const int r = t / FACE_NB, c = t % FACE_NB;
        if (r > k && r < nb && c >= r && c < nb) D[r * FACE_NB + c] -= D[k * FACE_NB + r] * D[k * FACE_NB + c];
      }
      __syncthreads();
    }
    // panel: column j solves D^T x = S[k0:k0+nb, j] by forward substitution
    for (int jj = t; jj < W; jj += BLOCK) {
      const int j = c0 + jj;
      double x[FACE_NB];
#pragma unroll
      for (int r = 0; r < FACE_NB; r++) {
        if (r < nb) {
          double v = S[(int64_t)(k0 + r) * M + j];
          for (int q = 0; q < r; q++) v -= D[q * FACE_NB + r] * x[q];
          x[r] = v / D[r * FACE_NB + r];
          Pn[r * FACE_PANEL + jj] = x[r];
          S[(int64_t)(k0 + r) * M + j] = x[r];
        } else {
          Pn[r * FACE_PANEL + jj] = 0.0;  // the trailing update multiplies all FACE_NB rows
        }
      }
    }
    // the factor's diagonal block back into S
    {
      const int r = t / FACE_NB, c = t % FACE_NB;
      if (r < nb && c >= r && c < nb) S[(int64_t)(k0 + r) * M + (k0 + c)] = D[r * FACE_NB + c];
    }
    __syncthreads();
    // trailing window: S[i][j] -= sum_r Pn[r][i] Pn[r][j], c0 <= i <= j <= e
    for (int ii = w; ii < W; ii += WAVES_PER_BLOCK) {
      double pi[FACE_NB];
#pragma unroll
      for (int r = 0; r < FACE_NB; r++) pi[r] = r < nb ? Pn[r * FACE_PANEL + ii] : 0.0;
      GLOBAL double *Srow = S + (int64_t)(c0 + ii) * M + c0;
      for (int jj = ii + l; jj < W; jj += WAVE) {
        double acc = 0.0;
#pragma unroll
        for (int r = 0; r < FACE_NB; r++) acc += pi[r] * Pn[r * FACE_PANEL + jj];
        Srow[jj] = Srow[jj] - acc;
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    gp(P.scal)[i_bw] = (double)bw;
    gp(P.scal)[i_fail] = fail_s ? 1.0 : 0.0;
  }
}

// One workgroup per column: y_n = K_n^-1 (g + lambda)[I_n] into Y (operand slot
// iy) at the column's free rows; one wave per row of the block.
__global__ __launch_bounds__(BLOCK) void kf_y(const Prob *__restrict__ probs, const Face *__restrict__ faces,
                                              WorkList wl, int iy) {
  __shared__ double v[FACE_MAX_M];
  FACE_PROLOGUE();
  const int n = lb, M = P.M, N = P.N;
  const int m = gp(FC.cnt)[n];
  if (m == 0 || m > FACE_MAX_M) return;
  const GLOBAL int *cr = gp(FC.crow) + (int64_t)n * M;
  const GLOBAL double *A = gp(FC.kb) + gp(FC.koff)[n];
  const GLOBAL double *G = gp(P.op[1]);
  for (int b = threadIdx.x; b < m; b += BLOCK) {
    const int i = cr[b];
    v[b] = G[(int64_t)i * N + n] + gp(FC.lam)[i];
  }
  __syncthreads();
  GLOBAL double *Y = gp(P.op[iy]);
  for (int a = wave_id(); a < m; a += WAVES_PER_BLOCK) {
    double s = 0.0;
    for (int b = lane_id(); b < m; b += WAVE) s += A[(int64_t)a * m + b] * v[b];
    s = wave_sum(s);
    if (lane_id() == 0) Y[(int64_t)cr[a] * N + n] = s;
  }
}

// One wave per row: rhs[j] = -(1 - sum_n X[j, n]) - sum over free n of Y[j, n].
__global__ __launch_bounds__(BLOCK) void kf_rhs(const Prob *__restrict__ probs, const Face *__restrict__ faces,
                                                WorkList wl, int iy) {
  FACE_PROLOGUE();
  const int j = lb * WAVES_PER_BLOCK + wave_id();
  if (j >= P.M) return;
  const int N = P.N;
  const int64_t base = (int64_t)j * N;
  const GLOBAL double *xr = gp(P.op[0]) + base;
  const GLOBAL double *yr = gp(P.op[iy]) + base;
  const GLOBAL uint8_t *ikr = gp(P.ik) + base;
  double sx = 0.0, sy = 0.0;
  for (int n = lane_id(); n < N; n += WAVE) {
    sx += xr[n];
    if (!ikr[n]) sy += yr[n];
  }
  sx = wave_sum(sx);
  sy = wave_sum(sy);
  if (lane_id() == 0) gp(FC.rhs)[j] = -(1.0 - sx) - sy;
}

// One workgroup per problem: U^T U delta = rhs with the band factor of kf_chol
// (bandwidth scal[i_bw]); delta replaces rhs and lambda += delta.
__global__ __launch_bounds__(BLOCK) void kf_solve(const Prob *__restrict__ probs, const Face *__restrict__ faces,
                                                  WorkList wl, int i_bw) {
  __shared__ double r[FACE_MAX_ROWS];
  __shared__ double z[FACE_MAX_ROWS];
  FACE_PROLOGUE();
  const int M = P.M;
  if (M > FACE_MAX_ROWS) return;
  const int bw = (int)gp(P.scal)[i_bw];
  const GLOBAL double *U = gp(FC.S);
  GLOBAL double *out = gp(FC.rhs);
  for (int i = threadIdx.x; i < M; i += BLOCK) r[i] = out[i];
  __syncthreads();
  // U^T z = r: z[k] = r[k] / U[k][k], then r[j] -= U[k][j] z[k] for j in (k, k + bw]
  for (int k = 0; k < M; k++) {
    const double zk = r[k] / U[(int64_t)k * M + k];
    const int e = min(M - 1, k + bw);
    for (int j = k + 1 + threadIdx.x; j <= e; j += BLOCK) r[j] -= U[(int64_t)k * M + j] * zk;
    if (threadIdx.x == 0) z[k] = zk;
    __syncthreads();
  }
  // U delta = z: delta[k] = z[k] / U[k][k], then z[i] -= U[i][k] delta[k] for i in [k - bw, k)
  for (int k = M - 1; k >= 0; k--) {
    const double dk = z[k] / U[(int64_t)k * M + k];
    const int s0 = max(0, k - bw);
    for (int i = s0 + threadIdx.x; i < k; i += BLOCK) z[i] -= U[(int64_t)i * M + k] * dk;
    if (threadIdx.x == 0) r[k] = dk;
    __syncthreads();
  }
  for (int i = threadIdx.x; i < M; i += BLOCK) {
    const double d = r[i];
    out[i] = d;
    gp(FC.lam)[i] = gp(FC.lam)[i] + d;
  }
}

// One workgroup per column: dx = -(y_n + K_n^-1 delta[I_n]), X += dx at the free
// rows; cstep[n] = the column's largest |dx| (0 without free rows).
__global__ __launch_bounds__(BLOCK) void kf_update(const Prob *__restrict__ probs, const Face *__restrict__ faces,
                                                   WorkList wl, int iy) {
  __shared__ double v[FACE_MAX_M];
  __shared__ double mw[WAVES_PER_BLOCK];
  FACE_PROLOGUE();
  const int n = lb, M = P.M, N = P.N;
  const int m = gp(FC.cnt)[n];
  if (m == 0 || m > FACE_MAX_M) {
    if (threadIdx.x == 0) gp(FC.cstep)[n] = 0.0;
    return;
  }
  const GLOBAL int *cr = gp(FC.crow) + (int64_t)n * M;
  const GLOBAL double *A = gp(FC.kb) + gp(FC.koff)[n];
  const GLOBAL double *delta = gp(FC.rhs);
  for (int b = threadIdx.x; b < m; b += BLOCK) v[b] = delta[cr[b]];
  __syncthreads();
  GLOBAL double *X = gp(P.op[0]);
  const GLOBAL double *Y = gp(P.op[iy]);
  double mx = 0.0;
  for (int a = wave_id(); a < m; a += WAVES_PER_BLOCK) {
    double s = 0.0;
    for (int b = lane_id(); b < m; b += WAVE) s += A[(int64_t)a * m + b] * v[b];
    s = wave_sum(s);
    if (lane_id() == 0) {
      const int64_t e = (int64_t)cr[a] * N + n;
      const double dx = -(Y[e] + s);
      X[e] = X[e] + dx;
      mx = fmax(mx, fabs(dx));
    }
  }
  if (lane_id() == 0) mw[wave_id()] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = mw[0];
    for (int u = 1; u < WAVES_PER_BLOCK; u++) r = fmax(r, mw[u]);
    gp(FC.cstep)[n] = r;
  }
}

// One block per problem: scal[dst] = -(largest cstep), the sign convention of the
// polish step scalar.
__global__ __launch_bounds__(BLOCK) void kf_stepmax(const Prob *__restrict__ probs, const Face *__restrict__ faces,
                                                    WorkList wl, int dst) {
  __shared__ double sh[WAVES_PER_BLOCK];
  FACE_PROLOGUE();
  double m = 0.0;
  for (int n = threadIdx.x; n < P.N; n += BLOCK) m = fmax(m, gp(FC.cstep)[n]);
  const double r = block_min(-m, sh);
  if (threadIdx.x == 0) gp(P.scal)[dst] = r;
}

// One wave per row: X = max(X, 0). The certified face solution's free entries can
// sit a rounding below 0; its row sums are 1 to rounding.
__global__ __launch_bounds__(BLOCK) void kf_clip(const Prob *__restrict__ probs, WorkList wl) {
  BATCH_PROLOGUE();
  const int j = lb * WAVES_PER_BLOCK + wave_id();
  if (j >= P.M) return;
  GLOBAL double *xr = gp(P.op[0]) + (int64_t)j * P.N;
  for (int n = lane_id(); n < P.N; n += WAVE)
    if (!(xr[n] > 0.0)) xr[n] = 0.0;
}

}  // namespace pq
