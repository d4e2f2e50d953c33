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
// O = F V  (the reference's redO local part, pqdts.f90:724-736)
//
// Block = RD consecutive F rows x 256 consecutive columns n. The union k-range
// of the RD row spans is swept in chunks of KC; each chunk of F is staged in
// LDS (zero-filled outside a row's span) and every V element fetched from HBM/L2
// is reused by RD fused multiply-adds from registers.
//   EXPAND: the reduced-space BF input (pqdts.f90:260-267): V[k, maxr[k]] is
//           replaced by ev[k] = -(sum_n V[k,n] - V[k,maxr[k]]).
//   LOSS:   the objective's data term (pqdts.f90:944-952): per-block partial of
//           sum (O - P)^2 over rows that have entries.
//   STORE:  write O.
// ---------------------------------------------------------------------------
template <int RD, int KC, bool EXPAND, bool LOSS, bool STORE>
__global__ __launch_bounds__(BLOCK) void k_spmm_rows(
    const double *__restrict__ V, int N, int D, const int *__restrict__ flo, const int *__restrict__ flen,
    const int64_t *__restrict__ foff, const double *__restrict__ fval, const uint8_t *__restrict__ fhas,
    const int *__restrict__ tlo, const int *__restrict__ thi, const int *__restrict__ maxr,
    const double *__restrict__ ev, const double *__restrict__ P, double *__restrict__ O,
    double *__restrict__ partial) {
  __shared__ double Fs[RD][KC];
  __shared__ int s_lo[RD], s_len[RD];
  __shared__ int64_t s_off[RD];
  __shared__ double sh[WAVES_PER_BLOCK];
  const int tile = blockIdx.y;
  const int d0 = tile * RD;
  const int n = blockIdx.x * BLOCK + threadIdx.x;
  const bool active = n < N;
  if (threadIdx.x < RD) {
    int d = d0 + threadIdx.x;
    bool ok = d < D && fhas[d];
    s_lo[threadIdx.x] = ok ? flo[d] : 0;
    s_len[threadIdx.x] = ok ? flen[d] : 0;
    s_off[threadIdx.x] = ok ? foff[d] : 0;
  }
  double acc[RD];
#pragma unroll
  for (int r = 0; r < RD; r++) acc[r] = 0.0;
  const int ulo = tlo[tile], uhi = thi[tile];
  __syncthreads();
  for (int k0 = ulo; k0 <= uhi; k0 += KC) {
    const int kc = min(KC, uhi - k0 + 1);
    for (int e = threadIdx.x; e < RD * KC; e += BLOCK) {
      const int r = e / KC, kk = e - r * KC;
      const int k = k0 + kk;
      double v = 0.0;
      const int rel = k - s_lo[r];
      if (kk < kc && rel >= 0 && rel < s_len[r]) v = fval[s_off[r] + rel];
      Fs[r][kk] = v;
    }
    __syncthreads();
    if (active) {
      const double *vp = V + (int64_t)k0 * N + n;
      int kk = 0;
      for (; kk + 1 < kc; kk += 2) {
        double v0 = vp[(int64_t)kk * N];
        double v1 = vp[(int64_t)(kk + 1) * N];
        if (EXPAND) {
          if (n == maxr[k0 + kk]) v0 = ev[k0 + kk];
          if (n == maxr[k0 + kk + 1]) v1 = ev[k0 + kk + 1];
        }
#pragma unroll
        for (int r = 0; r < RD; r++) {
          acc[r] = fma(Fs[r][kk], v0, acc[r]);
          acc[r] = fma(Fs[r][kk + 1], v1, acc[r]);
        }
      }
      if (kk < kc) {
        double v0 = vp[(int64_t)kk * N];
        if (EXPAND && n == maxr[k0 + kk]) v0 = ev[k0 + kk];
#pragma unroll
        for (int r = 0; r < RD; r++) acc[r] = fma(Fs[r][kk], v0, acc[r]);
      }
    }
    __syncthreads();
  }
  double loss = 0.0;
  if (active) {
#pragma unroll
    for (int r = 0; r < RD; r++) {
      const int d = d0 + r;
      if (d < D) {
        if (STORE) O[(int64_t)d * N + n] = acc[r];
        if (LOSS && s_len[r] > 0) {
          double t = acc[r] - P[(int64_t)d * N + n];
          loss = fma(t, t, loss);
        }
      }
    }
  }
  if (LOSS) {
    double b = block_sum(loss, sh);
    if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = b;
  }
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
// G[k,n] = 2 sum_d F[d,k] W[d,n] (+ 2 gamma second difference of E along k),
// then in reduced space the fold G[k,:] -= G[k,maxr[k]] (pqdts.f90:334-340,
// 1024-1030), then the active-set mask (pqdts.f90:2040-2052), plus a per-block
// partial of sum dd*G (the CG's v2 = p.Hp, pqdts.f90:2058).
//
// Block = RK consecutive k x 256 consecutive n; the union d-range of the RK
// columns of F is swept in chunks of KC staged in LDS as F^T.
// ---------------------------------------------------------------------------
template <int RK, int KC, int MODE, bool REDUCED, bool MASK, bool DOT>
__global__ __launch_bounds__(BLOCK) void k_spmmT_cols(
    const double *__restrict__ O, const double *__restrict__ P, int N, int M, int D,
    const int *__restrict__ flo, const int *__restrict__ flen, const int64_t *__restrict__ foff,
    const double *__restrict__ fval, const int *__restrict__ tlo, const int *__restrict__ thi, double gamma,
    const double *__restrict__ E, const int *__restrict__ maxr, const double *__restrict__ ev,
    const double *__restrict__ gm, const uint8_t *__restrict__ ik, const double *__restrict__ dd,
    double *__restrict__ G, double *__restrict__ partial) {
  __shared__ double Ft[RK][KC];
  __shared__ double sh[WAVES_PER_BLOCK];
  const int tile = blockIdx.y;
  const int k0 = tile * RK;
  const int n = blockIdx.x * BLOCK + threadIdx.x;
  const bool active = n < N;
  double acc[RK];
#pragma unroll
  for (int r = 0; r < RK; r++) acc[r] = 0.0;
  const int ulo = tlo[tile], uhi = thi[tile];
  for (int c0 = ulo; c0 <= uhi; c0 += KC) {
    const int kc = min(KC, uhi - c0 + 1);
    for (int e = threadIdx.x; e < RK * KC; e += BLOCK) {
      const int r = e / KC, dd_ = e - r * KC;
      const int d = c0 + dd_;
      double v = 0.0;
      if (dd_ < kc) {
        const int rel = k0 + r - flo[d];
        if (rel >= 0 && rel < flen[d]) v = fval[foff[d] + rel];
      }
      Ft[r][dd_] = v;
    }
    __syncthreads();
    if (active) {
      const double *op = O + (int64_t)c0 * N + n;
      const double *pp = P + (int64_t)c0 * N + n;
      for (int j = 0; j < kc; j++) {
        double w = op[(int64_t)j * N];
        if (MODE == T_GRAD) w -= pp[(int64_t)j * N];
#pragma unroll
        for (int r = 0; r < RK; r++) acc[r] = fma(Ft[r][j], w, acc[r]);
      }
    }
    __syncthreads();
  }
  double dot = 0.0;
  if (active) {
    const double g2 = 2.0 * gamma;
    constexpr bool EXP = (MODE == T_HVP) && REDUCED;
#pragma unroll
    for (int r = 0; r < RK; r++) {
      const int k = k0 + r;
      if (k >= M) break;
      double g = 2.0 * acc[r];
      if (gamma != 0.0) {
        const double e0 = stencil_src<EXP>(E, maxr, ev, k, n, N);
        if (k + 1 < M) g = g + g2 * (e0 - stencil_src<EXP>(E, maxr, ev, k + 1, n, N));
        if (k > 0) g = g - g2 * (stencil_src<EXP>(E, maxr, ev, k - 1, n, N) - e0);
      }
      const int64_t idx = (int64_t)k * N + n;
      if (REDUCED) g = (n == maxr[k]) ? 0.0 : g - gm[k];
      if (MASK && ik[idx]) g = 0.0;
      G[idx] = g;
      if (DOT) dot = fma(dd[idx], g, dot);
    }
  }
  if (DOT) {
    double b = block_sum(dot, sh);
    if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = b;
  }
}

// Pre-fold value G[k, maxr[k]] of the transposed product (one wave per k), used
// by the fold of k_spmmT_cols. Same formula as k_spmmT_cols at n = maxr[k].
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
