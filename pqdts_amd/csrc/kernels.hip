// Device kernels of the pqdts_amd solver (gfx950, wave64).
//
// Data layout in HBM (one problem):
//   X, G, HP, XTMP, SN, Z : M x N float64, row-major (photon number i, outcome n),
//                           so every row-wise operation of the reference (row
//                           sums, maxloc, simplex projection) reads one contiguous
//                           row, and every column-parallel loop of the reference
//                           becomes a coalesced sweep along n.
//   O, P                  : D x N float64, row-major (probe d, outcome n).
//   F                     : D rows as dense spans (the reference's Frows,
//                           pqdts.f90:1345-1363): row d covers columns
//                           flo[d] .. flo[d]+flen[d]-1, values at fval[foff[d]..].
//   maxr                  : M int32 (0-based, -1 = none); ik: M x N uint8.
//
// Reductions are deterministic: every block writes one partial, and a single
// block later sums the partials in index order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pq {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;
constexpr int WAVES_PER_BLOCK = BLOCK / WAVE;

// single-precision literals of the reference, widened (pqdts.f90:354, 424, 1207)
constexpr double kActiveTol = (double)1e-8f;

__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }
__device__ __forceinline__ int wave_id() { return threadIdx.x / WAVE; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, WAVE));
  return v;
}

// (value, index) max with the smallest index among equal values (Fortran maxloc)
__device__ __forceinline__ void wave_argmax(double &v, int &i) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    double ov = __shfl_xor(v, o, WAVE);
    int oi = __shfl_xor(i, o, WAVE);
    if (ov > v || (ov == v && oi < i)) {
      v = ov;
      i = oi;
    }
  }
}

// Block-wide sum, result valid in thread 0. `sh` holds WAVES_PER_BLOCK doubles.
__device__ __forceinline__ double block_sum(double v, double *sh) {
  v = wave_sum(v);
  if (lane_id() == 0) sh[wave_id()] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < WAVES_PER_BLOCK; w++) r += sh[w];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double block_min(double v, double *sh) {
  v = wave_min(v);
  if (lane_id() == 0) sh[wave_id()] = v;
  __syncthreads();
  double r = sh[0];
  if (threadIdx.x == 0)
    for (int w = 1; w < WAVES_PER_BLOCK; w++) r = fmin(r, sh[w]);
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// Final reduction of partials into scalars: out[dst] = sum (or min) of
// part[0..n). One block; fixed order, so repeated runs give identical bits.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_reduce(const double *__restrict__ part, int n, double *__restrict__ out,
                                                  int dst, int is_min) {
  __shared__ double sh[WAVES_PER_BLOCK];
  double v = 0.0;
  if (is_min) {
    v = INFINITY;
    for (int i = threadIdx.x; i < n; i += BLOCK) v = fmin(v, part[i]);
    v = block_min(v, sh);
  } else {
    for (int i = threadIdx.x; i < n; i += BLOCK) v += part[i];
    v = block_sum(v, sh);
  }
  if (threadIdx.x == 0) out[dst] = v;
}

// ---------------------------------------------------------------------------
// Epilogue modes of the transposed product G = 2 F^T W + stencil.
// ---------------------------------------------------------------------------
enum TMode {
  T_GRAD = 0,  // dL: W = O - P, stencil on X          (pqdts.f90:998-1021)
  T_HVP = 1    // BF: W = O,     stencil on direction  (pqdts.f90:280-330)
};

// E(j,n): the stencil source. In the reduced HVP, the direction's implicit
// column is replaced by ev[j] (the expanded vector, pqdts.f90:293-317).
template <bool EXP>
__device__ __forceinline__ double stencil_src(const double *__restrict__ E, const int *__restrict__ maxr,
                                              const double *__restrict__ ev, int j, int n, int N) {
  if (EXP && n == maxr[j]) return ev[j];
  return E[(int64_t)j * N + n];
}

// ---------------------------------------------------------------------------
// Band GEMM on FP64 MFMA: C[r, n] = sum_k A[r, k] B[k, n] for a banded A.
//
// Both products of the solver are this kernel:
//   O = F V        A = F   (D x M), B = V (M x N)             (redO, pqdts.f90:724-736)
//   G = 2 F^T W    A = F^T (M x D), B = W = O - P or O (D x N) (dL 998-1006, BF 280-288)
// A is pre-packed per tile of TR rows as a dense, zero-filled, transposed band
// [kchunks*KC][TR] covering the union of the tile's row spans, so staging a
// chunk of A is one contiguous copy. B chunks (KC rows x TC columns, rows
// contiguous along n) are staged through double-buffered LDS while the
// previous chunk feeds v_mfma_f64_16x16x4f64. Four waves, each a 32 x 32
// output tile (2 x 2 MFMA tiles, accumulators in registers).
// ---------------------------------------------------------------------------
typedef double d4_t __attribute__((ext_vector_type(4)));

enum BMode { B_PLAIN = 0, B_EXPAND = 1, B_DIFF = 2 };
enum EMode { E_STORE = 0, E_LOSS = 1, E_STORE_LOSS = 2, E_GRAD = 3, E_HVP = 4 };

struct BandArgs {
  const double *apack;    // packed A tiles
  const int64_t *aoff;    // per tile: offset into apack
  const int *klo;         // per tile: first k of the band
  const int *kchunks;     // per tile: number of KC chunks
  int nrows;              // rows of C
  int nk;                 // rows of B
  int N;                  // columns
  const double *B;        // B operand (V or O)
  const double *B2;       // P for B_DIFF
  const int *maxr;        // implicit columns (B_EXPAND on rows of B; reduced epilogues on rows of C)
  const double *ev;       // expansion values, indexed like maxr
  double *C;              // output
  const double *P;        // E_LOSS: data
  const uint8_t *has;     // E_LOSS: rows of F that have entries
  double gamma;           // E_GRAD / E_HVP: smoothing weight
  const double *E;        // stencil source (X or the direction)
  const double *gm;       // fold values G[k, maxr[k]]
  const uint8_t *ik;      // active set
  const double *dd;       // direction for the v2 partial
  double *partial;        // per-block partials (loss or dot)
};

template <int TR, int TC, int KC, int BM, int EM, bool RED, bool MASK, bool DOT>
__global__ __launch_bounds__(BLOCK) void k_band_gemm(BandArgs a) {
  static_assert(TR == 64 && TC == 64, "wave layout assumes a 64 x 64 block tile");
  constexpr int WC = 2;                  // waves along columns
  constexpr int MR = 2, MC = 2;          // 16 x 16 MFMA tiles per wave along rows / columns
  constexpr int AE = KC * TR / BLOCK;    // A elements staged per thread per chunk
  constexpr int BE = KC * TC / BLOCK;    // B elements staged per thread per chunk
  __shared__ double As[2][KC * TR];      // [kk][r]
  __shared__ double Bs[2][KC * TC];      // [kk][c]
  __shared__ double sh[WAVES_PER_BLOCK];
  const int tile = blockIdx.y;
  const int r0 = tile * TR, n0 = blockIdx.x * TC;
  const int tid = threadIdx.x, lane = lane_id(), w = wave_id();
  const int wr0 = (w / WC) * 32, wc0 = (w % WC) * 32;
  const double *ap = a.apack + a.aoff[tile];
  const int k0 = a.klo[tile], nch = a.kchunks[tile];
  const int N = a.N;

  double ra[AE], rb[BE];
  auto gload = [&](int c) {
    const double *src = ap + (int64_t)c * (KC * TR);
#pragma unroll
    for (int i = 0; i < AE; i++) ra[i] = src[tid + BLOCK * i];
#pragma unroll
    for (int i = 0; i < BE; i++) {
      const int e = tid + BLOCK * i;
      const int kk = e / TC, cc = e - kk * TC;
      const int k = k0 + c * KC + kk, n = n0 + cc;
      double v = 0.0;
      if (k < a.nk && n < N) {
        const int64_t idx = (int64_t)k * N + n;
        v = a.B[idx];
        if (BM == B_DIFF) v -= a.B2[idx];
        if (BM == B_EXPAND && n == a.maxr[k]) v = a.ev[k];
      }
      rb[i] = v;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < AE; i++) As[buf][tid + BLOCK * i] = ra[i];
#pragma unroll
    for (int i = 0; i < BE; i++) Bs[buf][tid + BLOCK * i] = rb[i];
  };

  d4_t acc[MR][MC];
#pragma unroll
  for (int i = 0; i < MR; i++)
#pragma unroll
    for (int j = 0; j < MC; j++) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};

  if (nch > 0) {
    gload(0);
    sstore(0);
  }
  __syncthreads();
  const int fk = lane >> 4, fi = lane & 15;  // fragment coordinates: k within the step, row/column
  for (int c = 0; c < nch; c++) {
    const int buf = c & 1;
    if (c + 1 < nch) gload(c + 1);
#pragma unroll
    for (int ks = 0; ks < KC; ks += 4) {
      double af[MR], bf[MC];
#pragma unroll
      for (int i = 0; i < MR; i++) af[i] = As[buf][(ks + fk) * TR + wr0 + i * 16 + fi];
#pragma unroll
      for (int j = 0; j < MC; j++) bf[j] = Bs[buf][(ks + fk) * TC + wc0 + j * 16 + fi];
#pragma unroll
      for (int i = 0; i < MR; i++)
#pragma unroll
        for (int j = 0; j < MC; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    if (c + 1 < nch) sstore(buf ^ 1);
    __syncthreads();
  }

  // epilogue: register q of lane l holds row 4*q + l/16, column l%16 of each
  // 16 x 16 tile (the FP64 MFMA output layout: four groups of one row)
  double red = 0.0;
  const double g2 = 2.0 * a.gamma;
  constexpr bool EXP = (EM == E_HVP) && RED;
#pragma unroll
  for (int i = 0; i < MR; i++)
#pragma unroll
    for (int j = 0; j < MC; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = r0 + wr0 + i * 16 + 4 * q + fk;
        const int n = n0 + wc0 + j * 16 + fi;
        if (row >= a.nrows || n >= N) continue;
        const int64_t idx = (int64_t)row * N + n;
        const double v = acc[i][j][q];
        if (EM == E_STORE || EM == E_STORE_LOSS) a.C[idx] = v;
        if ((EM == E_LOSS || EM == E_STORE_LOSS) && a.has[row]) {
          const double t = v - a.P[idx];
          red = fma(t, t, red);
        }
        if (EM == E_GRAD || EM == E_HVP) {
          double g = 2.0 * v;
          if (a.gamma != 0.0) {
            const double e0 = stencil_src<EXP>(a.E, a.maxr, a.ev, row, n, N);
            if (row + 1 < a.nrows) g = g + g2 * (e0 - stencil_src<EXP>(a.E, a.maxr, a.ev, row + 1, n, N));
            if (row > 0) g = g - g2 * (stencil_src<EXP>(a.E, a.maxr, a.ev, row - 1, n, N) - e0);
          }
          if (RED) g = (n == a.maxr[row]) ? 0.0 : g - a.gm[row];
          if (MASK && a.ik[idx]) g = 0.0;
          a.C[idx] = g;
          if (DOT) red = fma(a.dd[idx], g, red);
        }
      }
  if (EM == E_LOSS || EM == E_STORE_LOSS || DOT) {
    const double b = block_sum(red, sh);
    if (threadIdx.x == 0) a.partial[blockIdx.y * gridDim.x + blockIdx.x] = b;
  }
}

// Pre-fold value G[k, maxr[k]] of the transposed product (one wave per k), used
// by the fold in the transposed band GEMM; same formula as its epilogue at n = maxr[k].
template <int MODE>
__global__ __launch_bounds__(BLOCK) void k_foldval(const double *__restrict__ O, const double *__restrict__ P,
                                                   int N, int M, const int *__restrict__ flo,
                                                   const int *__restrict__ flen, const int64_t *__restrict__ foff,
                                                   const double *__restrict__ fval, const int *__restrict__ kdlo,
                                                   const int *__restrict__ kdhi, double gamma,
                                                   const double *__restrict__ E, const int *__restrict__ maxr,
                                                   const double *__restrict__ ev, double *__restrict__ gm) {
  const int k = blockIdx.x * WAVES_PER_BLOCK + wave_id();
  if (k >= M) return;
  const int c = maxr[k];
  double acc = 0.0;
  for (int d = kdlo[k] + lane_id(); d <= kdhi[k]; d += WAVE) {
    const int rel = k - flo[d];
    if (rel >= 0 && rel < flen[d]) {
      double w = O[(int64_t)d * N + c];
      if (MODE == T_GRAD) w -= P[(int64_t)d * N + c];
      acc = fma(fval[foff[d] + rel], w, acc);
    }
  }
  acc = wave_sum(acc);
  if (lane_id() == 0) {
    constexpr bool EXP = (MODE == T_HVP);
    double g = 2.0 * acc;
    if (gamma != 0.0) {
      const double g2 = 2.0 * gamma;
      const double e0 = stencil_src<EXP>(E, maxr, ev, k, c, N);
      if (k + 1 < M) g = g + g2 * (e0 - stencil_src<EXP>(E, maxr, ev, k + 1, c, N));
      if (k > 0) g = g - g2 * (stencil_src<EXP>(E, maxr, ev, k - 1, c, N) - e0);
    }
    gm[k] = g;
  }
}

// Gram path of BF: pre-fold value (2 G p~ + stencil)[k, maxr[k]] from the packed
// band of G = F^T F (layout of k_band_gemm's A operand), one wave per k. p~ is
// the reduced-space expansion of dd (implicit column replaced by ev).
__global__ __launch_bounds__(BLOCK) void k_foldval_gram(const double *__restrict__ gpack,
                                                        const int64_t *__restrict__ goff,
                                                        const int *__restrict__ glo, const int *__restrict__ gnch,
                                                        int kc, int tile, int N, int M, double gamma,
                                                        const double *__restrict__ dd, const int *__restrict__ maxr,
                                                        const double *__restrict__ ev, double *__restrict__ gm) {
  const int k = blockIdx.x * WAVES_PER_BLOCK + wave_id();
  if (k >= M) return;
  const int c = maxr[k];
  const int t = k / tile, r = k - t * tile;
  const int l0 = glo[t], nl = gnch[t] * kc;
  const double *gp = gpack + goff[t] + r;
  double acc = 0.0;
  for (int j = lane_id(); j < nl; j += WAVE) {
    const int l = l0 + j;
    if (l >= M) break;
    const double p = (c == maxr[l]) ? ev[l] : dd[(int64_t)l * N + c];
    acc = fma(gp[(int64_t)j * tile], p, acc);
  }
  acc = wave_sum(acc);
  if (lane_id() == 0) {
    double g = 2.0 * acc;
    if (gamma != 0.0) {
      const double g2 = 2.0 * gamma;
      const double e0 = stencil_src<true>(dd, maxr, ev, k, c, N);
      if (k + 1 < M) g = g + g2 * (e0 - stencil_src<true>(dd, maxr, ev, k + 1, c, N));
      if (k > 0) g = g - g2 * (stencil_src<true>(dd, maxr, ev, k - 1, c, N) - e0);
    }
    gm[k] = g;
  }
}

// Dense copy of F (D x M, row-major) from its row spans, one wave per row; used
// once to form the Gram matrix G = F^T F with the band GEMM.
__global__ __launch_bounds__(BLOCK) void k_dense_rows(const int *__restrict__ flo, const int *__restrict__ flen,
                                                      const int64_t *__restrict__ foff,
                                                      const double *__restrict__ fval, int D, int M,
                                                      double *__restrict__ Fd) {
  const int d = blockIdx.x * WAVES_PER_BLOCK + wave_id();
  if (d >= D) return;
  double *row = Fd + (int64_t)d * M;
  for (int k = lane_id(); k < M; k += WAVE) {
    const int rel = k - flo[d];
    row[k] = (rel >= 0 && rel < flen[d]) ? fval[foff[d] + rel] : 0.0;
  }
}

// Bdiag (pqdts.f90:208-244), one wave per k: 2 sum_d F[d,k]^2 + 2 gamma [k<M-1]
// + 2 gamma [k>0]; the reduced-space copy is twice that (pqdts.f90:235-236).
__global__ __launch_bounds__(BLOCK) void k_bdiag(int M, const int *__restrict__ flo, const int *__restrict__ flen,
                                                 const int64_t *__restrict__ foff, const double *__restrict__ fval,
                                                 const int *__restrict__ kdlo, const int *__restrict__ kdhi,
                                                 double gamma, double *__restrict__ diag_full,
                                                 double *__restrict__ diag_red) {
  const int k = blockIdx.x * WAVES_PER_BLOCK + wave_id();
  if (k >= M) return;
  double acc = 0.0;
  for (int d = kdlo[k] + lane_id(); d <= kdhi[k]; d += WAVE) {
    const int rel = k - flo[d];
    if (rel >= 0 && rel < flen[d]) {
      const double f = fval[foff[d] + rel];
      acc += 2.0 * (f * f);
    }
  }
  acc = wave_sum(acc);
  if (lane_id() == 0) {
    if (gamma != 0.0) {
      if (k + 1 < M) acc += 2.0 * gamma;
      if (k > 0) acc += 2.0 * gamma;
    }
    diag_full[k] = acc;
    diag_red[k] = acc * 2.0;
  }
}

// ---------------------------------------------------------------------------
// Row kernels: one wave per row j of an M x N array.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double row_sum(const double *__restrict__ row, int N) {
  double s = 0.0;
  for (int n = lane_id(); n < N; n += WAVE) s += row[n];
  return wave_sum(s);
}

// X[j,maxr] = 1 - (sum X[j,:] - X[j,maxr])   (pqdts.f90:931-933)
__global__ __launch_bounds__(BLOCK) void k_reconstruct(double *__restrict__ X, int M, int N,
                                                       const int *__restrict__ maxr) {
  const int j = blockIdx.x * WAVES_PER_BLOCK + wave_id();
  if (j >= M) return;
  double *row = X + (int64_t)j * N;
  const double s = row_sum(row, N);
  if (lane_id() == 0) {
    const int r = maxr[j];
    row[r] = 1.0 - (s - row[r]);
  }
}

// ev[j] = -(sum V[j,:] - V[j,maxr])   (pqdts.f90:264)
__global__ __launch_bounds__(BLOCK) void k_others(const double *__restrict__ V, int M, int N,
                                                  const int *__restrict__ maxr, double *__restrict__ ev) {
  const int j = blockIdx.x * WAVES_PER_BLOCK + wave_id();
  if (j >= M) return;
  const double *row = V + (int64_t)j * N;
  const double s = row_sum(row, N);
  if (lane_id() == 0) ev[j] = -(s - row[maxr[j]]);
}

// Reduced-space transform (pqdts.f90:1970-1980; mode 1 without the first step,
// pqdts.f90:1940-1944): rebuild the old implicit entry, pick the first maximum
// of the row as the new implicit column, store the placeholder 1.0 there.
__global__ __launch_bounds__(BLOCK) void k_transform(double *__restrict__ X, int M, int N, int *__restrict__ maxr,
                                                     int rebuild) {
  const int j = blockIdx.x * WAVES_PER_BLOCK + wave_id();
  if (j >= M) return;
  double *row = X + (int64_t)j * N;
  const int old = maxr[j];
  double newv = 0.0;
  if (rebuild && old != -1) {
    const double s = row_sum(row, N);
    newv = 1.0 - (s - row[old]);
  }
  double best = -INFINITY;
  int bi = 0x7fffffff;
  for (int n = lane_id(); n < N; n += WAVE) {
    double v = (rebuild && old != -1 && n == old) ? newv : row[n];
    if (v > best || bi == 0x7fffffff) {
      best = v;
      bi = n;
    }
  }
  wave_argmax(best, bi);
  if (lane_id() == 0) {
    if (rebuild && old != -1) row[old] = newv;
    row[bi] = 1.0;
    maxr[j] = bi;
  }
}

// doproj2 (pqdts.f90:1149-1158) on a row, optionally first forming the trial
// point T = a*S + 1*X (pgaxpy2, pqdts.f90:524), and optionally applying the
// objective's reconstruction again (L's prep, pqdts.f90:929-935). Per-block
// partial of xmin = min(0, min_j T[j,maxr]).
__global__ __launch_bounds__(BLOCK) void k_clip_rows(double *__restrict__ T, const double *__restrict__ X,
                                                     const double *__restrict__ S, double a, int M, int N,
                                                     const int *__restrict__ maxr, int form_trial, int recon_again,
                                                     double *__restrict__ partial) {
  // Each lane only re-reads the elements it wrote itself; values of other
  // lanes travel by wave reductions.
  __shared__ double sh[WAVES_PER_BLOCK];
  const int j = blockIdx.x * WAVES_PER_BLOCK + wave_id();
  double xm = 0.0;
  if (j < M) {
    double *row = T + (int64_t)j * N;
    const int r = maxr[j];
    double s = 0.0, own = 0.0;
    for (int n = lane_id(); n < N; n += WAVE) {
      double v;
      if (form_trial) {
        const int64_t idx = (int64_t)j * N + n;
        v = a * S[idx] + 1.0 * X[idx];
      } else {
        v = row[n];
      }
      if (v < 0.0) v = 0.0;
      row[n] = v;
      s += v;
      if (n == r) own = v;
    }
    s = wave_sum(s);
    own = wave_sum(own);
    double v = 1.0 - (s - own);
    xm = fmin(0.0, v);
    if (recon_again) {
      double s2 = 0.0;
      for (int n = lane_id(); n < N; n += WAVE) s2 += (n == r) ? v : row[n];
      s2 = wave_sum(s2);
      v = 1.0 - (s2 - v);
    }
    if (lane_id() == (r & (WAVE - 1))) row[r] = v;
  }
  xm = wave_min(xm);
  if (lane_id() == 0) sh[wave_id()] = xm;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = sh[0];
    for (int w = 1; w < WAVES_PER_BLOCK; w++) v = fmin(v, sh[w]);
    partial[blockIdx.x] = v;
  }
}

// Euclidean projection of a row onto the simplex (doproj, pqdts.f90:1120-1140,
// whose per-row kernel is Condat's simplexproj_condat_). On a wave: Michelot's
// fixed-point iteration tau <- (sum_{y>tau} y - 1) / #{y > tau}, starting from
// the mean, which increases tau monotonically to the same threshold as
// Condat's algorithm; each step is one sweep of the row plus two wave
// reductions. Optionally forms the trial point T = a*S + 1*X first.
__global__ __launch_bounds__(BLOCK) void k_simplex_rows(double *__restrict__ T, const double *__restrict__ X,
                                                        const double *__restrict__ S, double a, int M, int N,
                                                        int form_trial) {
  const int j = blockIdx.x * WAVES_PER_BLOCK + wave_id();
  if (j >= M) return;
  double *row = T + (int64_t)j * N;
  double s = 0.0;
  for (int n = lane_id(); n < N; n += WAVE) {
    double v;
    if (form_trial) {
      const int64_t idx = (int64_t)j * N + n;
      v = a * S[idx] + 1.0 * X[idx];
      row[n] = v;
    } else {
      v = row[n];
    }
    s += v;
  }
  s = wave_sum(s);
  // each lane re-reads only the elements it wrote itself
  double tau = (s - 1.0) / (double)N;
  int cnt = N;
  for (int it = 0; it < N + 1; it++) {
    double ss = 0.0;
    int c = 0;
    for (int n = lane_id(); n < N; n += WAVE) {
      const double v = row[n];
      if (v > tau) {
        ss += v;
        c++;
      }
    }
    ss = wave_sum(ss);
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, WAVE);
    if (c == cnt) break;
    cnt = c;
    tau = (ss - 1.0) / (double)c;
  }
  for (int n = lane_id(); n < N; n += WAVE) {
    const double v = row[n];
    row[n] = v > tau ? v - tau : 0.0;
  }
}

// conv_test rows (pqdts.f90:1055-1060): lambda_i = max(0, -min_n G[i,n]),
// partial of sum ((G + lambda) X)^2; optionally stores lambda and min(G+lambda).
__global__ __launch_bounds__(BLOCK) void k_conv_rows(const double *__restrict__ X, const double *__restrict__ G,
                                                     int M, int N, double *__restrict__ lam,
                                                     double *__restrict__ lmin, double *__restrict__ partial) {
  __shared__ double sh[WAVES_PER_BLOCK];
  const int j = blockIdx.x * WAVES_PER_BLOCK + wave_id();
  double s = 0.0;
  if (j < M) {
    const double *g = G + (int64_t)j * N, *x = X + (int64_t)j * N;
    double mn = INFINITY;
    for (int n = lane_id(); n < N; n += WAVE) mn = fmin(mn, g[n]);
    mn = wave_min(mn);
    const double lambda = fmax(0.0, -mn);
    for (int n = lane_id(); n < N; n += WAVE) {
      const double t = (g[n] + lambda) * x[n];
      s = fma(t, t, s);
    }
    if (lam != nullptr) {
      double m2 = INFINITY;
      for (int n = lane_id(); n < N; n += WAVE) m2 = fmin(m2, g[n] + lambda);
      m2 = wave_min(m2);
      if (lane_id() == 0) {
        lam[j] = lambda;
        lmin[j] = m2;
      }
    }
  }
  double b = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

// Smoothing part of the objective (pqdts.f90:956-962): partial of
// sum gamma (X[j,n] - X[j+1,n])^2 over j < M-1.
__global__ __launch_bounds__(BLOCK) void k_smooth(const double *__restrict__ X, int M, int N, double gamma,
                                                  double *__restrict__ partial) {
  __shared__ double sh[WAVES_PER_BLOCK];
  const int64_t total = (int64_t)(M - 1) * N;
  double s = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < total; e += (int64_t)gridDim.x * BLOCK) {
    const double t = X[e] - X[e + N];
    s += gamma * (t * t);
  }
  double b = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

// pdot (pqdts.f90:589-614)
__global__ __launch_bounds__(BLOCK) void k_dot(const double *__restrict__ x, const double *__restrict__ y,
                                               int64_t total, double *__restrict__ partial) {
  __shared__ double sh[WAVES_PER_BLOCK];
  double s = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < total; e += (int64_t)gridDim.x * BLOCK)
    s = fma(x[e], y[e], s);
  double b = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

// z = beta*y + alpha*x (pgaxpy / pgaxpy2 / paxpy / pscale / pcopy / pset)
__global__ __launch_bounds__(BLOCK) void k_axpby(double *__restrict__ z, const double *__restrict__ x,
                                                 const double *__restrict__ y, double alpha, double beta,
                                                 int64_t total, int kind) {
  for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < total; e += (int64_t)gridDim.x * BLOCK) {
    double v;
    switch (kind) {
      case 0: v = beta * y[e] + alpha * x[e]; break;  // pgaxpy2 / pgaxpy
      case 1: v = y[e] + alpha * x[e]; break;         // paxpy
      case 2: v = x[e] * alpha; break;                // pscale / pcopy_scale
      case 3: v = x[e]; break;                        // pcopy
      default: v = alpha; break;                      // pset
    }
    z[e] = v;
  }
}

// ---------------------------------------------------------------------------
// Newton-CG vector kernels (pqdts.f90:1986-2104), one wave per row.
// ---------------------------------------------------------------------------

// Active set + CG start (pqdts.f90:1986-2031): ik = X < tol and G > 0, G[ik] = 0;
// partials of nact and |G|^2; r = -G (in G); z = r / sqrt(diag); p = z (in P);
// sn = 0; partial of r.z; ev[j] = -(sum p[j,:] - p[j,maxr]) when reduced.
__global__ __launch_bounds__(BLOCK) void k_cg_init(const double *__restrict__ X, double *__restrict__ G,
                                                   uint8_t *__restrict__ ik, double *__restrict__ Z,
                                                   double *__restrict__ Pd, double *__restrict__ SN,
                                                   const double *__restrict__ diag, int M, int N,
                                                   const int *__restrict__ maxr, double *__restrict__ ev,
                                                   int reduced, int tactive, double *__restrict__ part_nact,
                                                   double *__restrict__ part_nd, double *__restrict__ part_b1) {
  __shared__ double sh[WAVES_PER_BLOCK];
  const int j = blockIdx.x * WAVES_PER_BLOCK + wave_id();
  double nact = 0.0, nd = 0.0, b1 = 0.0;
  if (j < M) {
    const double sq = sqrt(diag[j]);
    const int mr = reduced ? maxr[j] : -1;
    double psum = 0.0, own = 0.0;
    for (int n = lane_id(); n < N; n += WAVE) {
      const int64_t idx = (int64_t)j * N + n;
      double g = G[idx];
      uint8_t a = 0;
      if (tactive && X[idx] < kActiveTol && g > 0.0) {
        a = 1;
        g = 0.0;
        nact += 1.0;
      }
      ik[idx] = a;
      nd = fma(g, g, nd);
      const double r = g * -1.0;
      const double z = r / sq;
      G[idx] = r;
      Z[idx] = z;
      Pd[idx] = z;
      SN[idx] = 0.0;
      b1 = fma(r, z, b1);
      psum += z;
      if (n == mr) own = z;
    }
    if (reduced) {
      psum = wave_sum(psum);
      own = wave_sum(own);
      if (lane_id() == 0) ev[j] = -(psum - own);
    }
  }
  double t;
  t = block_sum(nact, sh);
  if (threadIdx.x == 0) part_nact[blockIdx.x] = t;
  t = block_sum(nd, sh);
  if (threadIdx.x == 0) part_nd[blockIdx.x] = t;
  t = block_sum(b1, sh);
  if (threadIdx.x == 0) part_b1[blockIdx.x] = t;
}

// CG update (pqdts.f90:2060-2070): a = b1/v2; sn += a p; r += (-a) Hp;
// z = r / sqrt(diag); partials of r.z and r.r.
__global__ __launch_bounds__(BLOCK) void k_cg_update(double *__restrict__ SN, const double *__restrict__ Pd,
                                                     double *__restrict__ R, const double *__restrict__ HP,
                                                     double *__restrict__ Z, const double *__restrict__ diag,
                                                     int M, int N, const double *__restrict__ scal, int i_b1,
                                                     int i_v2, double *__restrict__ part_rz,
                                                     double *__restrict__ part_rr) {
  __shared__ double sh[WAVES_PER_BLOCK];
  const int j = blockIdx.x * WAVES_PER_BLOCK + wave_id();
  const double a = scal[i_b1] / scal[i_v2];
  double rz = 0.0, rr = 0.0;
  if (j < M) {
    const double sq = sqrt(diag[j]);
    for (int n = lane_id(); n < N; n += WAVE) {
      const int64_t idx = (int64_t)j * N + n;
      SN[idx] = SN[idx] + a * Pd[idx];
      const double r = R[idx] + (-a) * HP[idx];
      R[idx] = r;
      const double z = r / sq;
      Z[idx] = z;
      rz = fma(r, z, rz);
      rr = fma(r, r, rr);
    }
  }
  double t = block_sum(rz, sh);
  if (threadIdx.x == 0) part_rz[blockIdx.x] = t;
  t = block_sum(rr, sh);
  if (threadIdx.x == 0) part_rr[blockIdx.x] = t;
}

// New direction (pqdts.f90:2075): p = (v2/b1) p + 1 z; ev for the next BF.
__global__ __launch_bounds__(BLOCK) void k_cg_dir(double *__restrict__ Pd, const double *__restrict__ Z, int M,
                                                  int N, const double *__restrict__ scal, int i_b1, int i_v2new,
                                                  const int *__restrict__ maxr, double *__restrict__ ev,
                                                  int reduced) {
  const int j = blockIdx.x * WAVES_PER_BLOCK + wave_id();
  if (j >= M) return;
  const double beta = scal[i_v2new] / scal[i_b1];
  const int mr = reduced ? maxr[j] : -1;
  double psum = 0.0, own = 0.0;
  for (int n = lane_id(); n < N; n += WAVE) {
    const int64_t idx = (int64_t)j * N + n;
    const double p = beta * Pd[idx] + 1.0 * Z[idx];
    Pd[idx] = p;
    psum += p;
    if (n == mr) own = p;
  }
  if (reduced) {
    psum = wave_sum(psum);
    own = wave_sum(own);
    if (lane_id() == 0) ev[j] = -(psum - own);
  }
}

// Two-metric diagonal step (pqdts.f90:2095-2102): sn = -G/diag on active
// entries outside the implicit column.
__global__ __launch_bounds__(BLOCK) void k_twometric(double *__restrict__ SN, const double *__restrict__ G,
                                                     const uint8_t *__restrict__ ik,
                                                     const double *__restrict__ diag, int M, int N,
                                                     const int *__restrict__ maxr) {
  const int j = blockIdx.x * WAVES_PER_BLOCK + wave_id();
  if (j >= M) return;
  const double dj = diag[j];
  const int mr = maxr[j];
  for (int n = lane_id(); n < N; n += WAVE) {
    const int64_t idx = (int64_t)j * N + n;
    if (ik[idx] == 1 && mr != n) SN[idx] = -G[idx] / dj;
  }
}

}  // namespace pq
