// Device kernels of the pqdts_amd solver (gfx950, wave64).
//
// Every kernel is batched: it receives the device array of problem
// descriptors (Prob) and, by value in its kernel arguments, a WorkList that
// maps its flat grid onto up to MAXB problems (a contiguous range of blocks per
// problem) together with per-problem flags and one per-problem scalar. A single
// problem is a batch of one, so the plain solve and the batched solve run the
// same code.
//
// Data layout in HBM (per problem):
//   X, G, HP, XTMP, SN, Z : M x N float64, row-major (photon number i, outcome n):
//                           every row-wise operation of the reference (row sums,
//                           maxloc, simplex projection) reads one contiguous row,
//                           every column-parallel loop becomes a coalesced sweep.
//   O, P                  : D x N float64, row-major (probe d, outcome n).
//   F                     : D rows as dense spans (the reference's Frows,
//                           pqdts.f90:1345-1363): flo/flen/foff/fval.
//   band packs            : F, F^T and optionally G = F^T F as per-tile dense
//                           transposed bands for the MFMA band GEMM (FP64, and an
//                           FP32 copy for the mixed-precision Hessian products).
//   maxr                  : M int32 (0-based, -1 = none); ik: M x N uint8.
//
// Reductions are deterministic: every block writes one partial into its
// problem's partial region, and one block per problem later sums them in index
// order (kb_reduce).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pq {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;
constexpr int WAVES_PER_BLOCK = BLOCK / WAVE;
constexpr int MAXB = 128;      // problems per launch
// element kernels over the M x N arrays: blocks of BLOCK x ELEMS elements
constexpr int ELEMS = 8;
constexpr int EBLOCK = BLOCK * ELEMS;
constexpr int NSLOT = 8;       // X G HP XTMP SN Z, then O and R = O - P as operands 6, 7
constexpr int OP_O = 6;
constexpr int OP_R = 7;

// single-precision literals of the reference, widened (pqdts.f90:1207)
constexpr double kActiveTol = (double)1e-8f;

struct Band {
  const double *pack;    // FP64 tiles
  const float *pack32;   // FP32 copy (mixed precision), may be null
  const int64_t *off;    // per tile: offset into pack
  const int *klo;        // per tile: first column of the band
  const int *nch;        // per tile: number of KC chunks
  int ntiles;
};

// Immutable per-problem descriptor (device copy built once per problem).
struct Prob {
  int M, N, D, pcap;
  double gamma;
  double *op[NSLOT];     // X G HP XTMP SN Z O R
  double *P;
  const double *diag_full, *diag_red;
  int *maxr;
  uint8_t *ik;
  double *ev, *gm, *lam, *lmin;
  const int *flo, *flen;
  const int64_t *foff;
  const double *fval;
  const uint8_t *fhas;
  const int *kdlo, *kdhi;
  Band bf, bt, bg;
  double *part;          // NREG partial regions of pcap doubles
  double *scal;          // NSCAL device scalars
  double *cglog;         // CG_LOG_ROUNDS x 5 values logged by the device-side CG loop
  int *pn;               // NREG: number of partials the last producer of each region stored
  int hs;                // K splits of the Gram Hessian product (1: epilogue in the product)
  double *hw;            // hs x M x N partial products when hs > 1
  float *p32;            // FP32 copy of the CG direction, implicit column expanded (mixed precision), or null
  double *pb;            // F_LATE: per row, sum of the direction over the free entries outside maxr
};

constexpr int CG_LOG_ROUNDS = 50;

constexpr int NREG = 4;
constexpr int LS_K = 2;                // line-search trials evaluated per host round trip, at most
constexpr int NSCAL = 21 + 3 * LS_K;   // solver scalars, then per-trial objective terms and xmin

// flags of a problem within one launch
constexpr unsigned F_FULL = 1u;  // full space (tfullspace); otherwise reduced
// Mixed-precision CG of a reduced problem: the Hessian product stores its result
// unfolded with the fold value of each row (gm), and kb_cg_update applies the fold
constexpr unsigned F_LATE = 2u;
// Polish iteration of a reduced problem (primal-dual active-set Newton after the
// minimisation): the CG keeps the active set it is given (ik), uses the Jacobi
// preconditioner z = r / diag and stops at a relative residual (the launch's
// per-problem argument of kb_cg_start)
constexpr unsigned F_POLISH = 4u;

struct WorkList {
  int np;
  int guard;  // scalar slot that must be nonzero for a problem to run (-1: none)
  int pid[MAXB];
  int boff[MAXB + 1];
  unsigned char flags[MAXB];
  double a[MAXB];  // per-problem scalar argument (step length, ...)
};

// Every array a descriptor points to is in global memory, but pointers loaded from
// memory are generic to the compiler, and generic (flat) accesses also count
// against the LDS wait counter: a wait for an LDS operand would then wait for the
// global loads in flight too. Accesses go through GLOBAL-qualified pointers.
#define GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ GLOBAL T *gp(T *p) {
  return (GLOBAL T *)p;
}

__device__ __forceinline__ int64_t lmin64(int64_t a, int64_t b) { return a < b ? a : b; }
__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }
__device__ __forceinline__ int wave_id() { return threadIdx.x / WAVE; }

// Work-list entry of flat block bid: the last i with boff[i] <= bid (boff is
// nondecreasing from boff[0] = 0), i.e. the number of such entries minus one.
// Every lane tests MAXB / WAVE entries and a ballot counts them: one round of
// independent loads instead of a binary search's chain of dependent ones.
// Wave-uniform result.
__device__ __forceinline__ int wl_find(const WorkList &wl, int bid) {
  const int l = lane_id();
  int cnt = 0;
#pragma unroll
  for (int b0 = 0; b0 < MAXB; b0 += WAVE) {
    const int i = b0 + l;
    cnt += __popcll(__ballot(i < wl.np && wl.boff[i] <= bid));
  }
  return cnt - 1;
}

#define BATCH_PROLOGUE()                                   \
  const int wi_ = wl_find(wl, blockIdx.x);                 \
  const Prob &P = probs[wl.pid[wi_]];                      \
  if (wl.guard >= 0 && gp(P.scal)[wl.guard] == 0.0) return;    \
  const int lb = blockIdx.x - wl.boff[wi_];                \
  const int nblk = wl.boff[wi_ + 1] - wl.boff[wi_];        \
  const bool full = (wl.flags[wi_] & F_FULL) != 0;         \
  const double arg = wl.a[wi_];                            \
  const bool late = (wl.flags[wi_] & F_LATE) != 0;         \
  const bool polish = (wl.flags[wi_] & F_POLISH) != 0;     \
  (void)late;                                              \
  (void)polish;                                            \
  (void)full;                                              \
  (void)nblk;                                              \
  (void)arg

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
    const double ov = __shfl_xor(v, o, WAVE);
    const int oi = __shfl_xor(i, o, WAVE);
    if (ov > v || (ov == v && oi < i)) {
      v = ov;
      i = oi;
    }
  }
}

// Block-wide sum / min, result valid in thread 0. `sh` holds WAVES_PER_BLOCK doubles.
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
// Per-problem reductions. A producing launch stores one partial per block of a
// problem (entry i of the work list has boff[i+1]-boff[i] blocks); the partials
// are summed (or min-ed) in index order, the same order wherever that happens:
//  * kb_reduce, one block per work-list entry, right after the producer;
//  * kb_reduce_pn, several regions at once, counts taken from pn[region];
//  * inside the next kernel, each of whose blocks sums the partials itself
//    (reduce_parts_all): no extra launch and no inter-block synchronisation.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double reduce_parts(const GLOBAL double *part, int n, bool is_min, double *sh) {
  double v;
  if (is_min) {
    v = INFINITY;
    for (int i = threadIdx.x; i < n; i += BLOCK) v = fmin(v, part[i]);
    v = block_min(v, sh);
  } else {
    v = 0.0;
    for (int i = threadIdx.x; i < n; i += BLOCK) v += part[i];
    v = block_sum(v, sh);
  }
  return v;
}

// reduce_parts with the result in every thread (`bc` is one shared double)
__device__ __forceinline__ double reduce_parts_all(const GLOBAL double *part, int n, double *sh, double *bc) {
  const double v = reduce_parts(part, n, false, sh);
  if (threadIdx.x == 0) *bc = v;
  __syncthreads();
  const double r = *bc;
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(BLOCK) void kb_reduce(const Prob *__restrict__ probs, WorkList wl, int region,
                                                   int dst, int is_min) {
  __shared__ double sh[WAVES_PER_BLOCK];
  const Prob &P = probs[wl.pid[blockIdx.x]];
  if (wl.guard >= 0 && gp(P.scal)[wl.guard] == 0.0) return;
  const int n = wl.boff[blockIdx.x + 1] - wl.boff[blockIdx.x];
  const double v = reduce_parts(gp(P.part) + (int64_t)region * P.pcap, n, is_min != 0, sh);
  if (threadIdx.x == 0) gp(P.scal)[dst] = v;
}

// Sums of up to three regions (scal[d_i] = sum of pn[r_i] partials of region r_i;
// the minimum instead for the regions whose bit is set in minmask), one block per
// work-list entry.
__global__ __launch_bounds__(BLOCK) void kb_reduce_pn(const Prob *__restrict__ probs, WorkList wl, int nr, int r0,
                                                      int d0, int r1, int d1, int r2, int d2, int minmask) {
  __shared__ double sh[WAVES_PER_BLOCK];
  const Prob &P = probs[wl.pid[blockIdx.x]];
  if (wl.guard >= 0 && gp(P.scal)[wl.guard] == 0.0) return;
  const int rr[3] = {r0, r1, r2}, dd[3] = {d0, d1, d2};
  for (int k = 0; k < nr; k++) {
    const double v = reduce_parts(gp(P.part) + (int64_t)rr[k] * P.pcap, gp(P.pn)[rr[k]], (minmask >> k) & 1, sh);
    if (threadIdx.x == 0) gp(P.scal)[dd[k]] = v;
  }
}

// scal[dst] = value for every problem of the list (one thread per entry)
__global__ void kb_set_scalar(const Prob *__restrict__ probs, WorkList wl, int dst, double value) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < wl.np) probs[wl.pid[i]].scal[dst] = value;
}

// ---------------------------------------------------------------------------
// Epilogues and operands of the band GEMM
// ---------------------------------------------------------------------------
enum TMode {
  T_GRAD = 0,  // dL: W = O - P, stencil on X          (pqdts.f90:998-1021)
  T_HVP = 1    // BF: W = O,     stencil on direction  (pqdts.f90:280-330)
};

// E(j,n): the stencil source. In the reduced HVP the direction's implicit column
// is replaced by ev[j] (the expanded vector, pqdts.f90:293-317).
__device__ __forceinline__ double stencil_src(bool exp, const GLOBAL double *__restrict__ E,
                                              const GLOBAL int *__restrict__ maxr, const GLOBAL double *__restrict__ ev,
                                              int j, int n, int N) {
  if (exp && n == maxr[j]) return ev[j];
  return E[(int64_t)j * N + n];
}

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

// B operand: as stored, with the implicit column expanded (reduced problems), or
// the FP32 expanded copy of the CG direction (P.p32, written where the direction
// is formed)
// B_SPAN: B is F itself (rows d, columns k) read from its span storage, zero
// outside each row's span: the Gram matrix F^T F formed without a dense copy of F.
enum BMode { B_PLAIN = 0, B_EXPAND = 1, B_P32 = 3, B_SPAN = 4 };
// E_GBAND: the product is the Gram matrix, stored straight into G's band pack
// (P.bg); output tiles outside the band of their row tile are not computed.
// E_STORE_R: store O = F X and R = O - P (the gradient's operand, formed once here
// instead of in every row tile of the transposed product that reads it).
enum EMode { E_STORE = 0, E_LOSS = 1, E_GRAD = 3, E_HVP = 4, E_PART = 5, E_GBAND = 6, E_STORE_R = 7 };
enum BandSel { BAND_F = 0, BAND_FT = 1, BAND_G = 2 };

struct GemmOps {
  int band;    // BandSel
  int b;       // operand index of B (rows of B = columns of the band)
  int c;       // operand index of C
  int e;       // stencil source operand (E_GRAD / E_HVP)
  int dd;      // direction operand for the v2 = dd.C partial
  int region;  // partial region
  int sm = -1;  // E_LOSS: region of the smoothing partials of B computed by extra blocks (-1: none)
};

template <typename T>
struct MfmaT;
template <>
struct MfmaT<double> {
  typedef d4_t acc_t;
  __device__ static acc_t mfma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // output register q of lane l sits at row 4*q + l/16 (four groups of one row)
  __device__ static int row(int q, int fk) { return 4 * q + fk; }
  __device__ static const double *pack(const Band &b) { return b.pack; }
};
template <>
struct MfmaT<float> {
  typedef f4_t acc_t;
  __device__ static acc_t mfma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // output register q of lane l sits at row 4*(l/16) + q (one group of four rows)
  __device__ static int row(int q, int fk) { return 4 * fk + q; }
  __device__ static const float *pack(const Band &b) { return b.pack32; }
};

// ---------------------------------------------------------------------------
// Band GEMM on MFMA: C[r, n] = sum_k A[r, k] B[k, n] for a banded A.
//
// All products of the solver are this kernel:
//   O = F V        A = F   (D x M), B = V (M x N)              (redO, pqdts.f90:724-736)
//   G = 2 F^T W    A = F^T (M x D), B = W = R = O - P or O (D x N)  (dL 998-1006, BF 280-288)
//   H p = 2 G p    A = G = F^T F (M x M), B = p (M x N)         (BF through the Gram matrix)
// A is pre-packed per tile of TR rows as a dense, zero-filled, transposed band
// [nch*KC][TR], so staging a chunk of A is one contiguous copy. B chunks (KC rows
// x TC columns) go through double-buffered LDS while the previous chunk feeds
// the MFMAs. Four waves, each a 32 x 32 output tile (2 x 2 MFMA tiles).
// T = double: v_mfma_f64_16x16x4f64. T = float (mixed precision): B is rounded
// to FP32 when staged, v_mfma_f32_16x16x4f32 accumulates in FP32 within a chunk
// and every chunk's sum is added into FP64 accumulators (compensating the FP32
// accumulation length to one chunk).
// ---------------------------------------------------------------------------
//
// band_gemm_tile computes output tile lb (64 x 64) of one problem with the whole
// workgroup, LDS buffers supplied by the caller, and adds the tile's loss or
// v2 contributions to redv; kb_band_gemm runs one tile per workgroup over a work
// list.
// ---------------------------------------------------------------------------
constexpr int GT = 64;  // rows and columns of an output tile
// Staging of band chunks in LDS. Two forms:
//  * registers (every operand but the FP32 direction): each thread loads its share
//    of chunk c+1 into registers while chunk c feeds the MFMAs, then stores it into
//    the other of two LDS buffers. Row stride lds_stride(): the MFMA operand reads of
//    a wave take 16 consecutive elements from each of 4 consecutive rows, and FP32
//    rows padded to 80 elements put those 4 groups on disjoint LDS banks (FP64 reads,
//    8 bytes a lane, are conflict-free unpadded).
//  * direct (FP32 products whose B is the FP32 CG direction, B_P32): global loads
//    write LDS without passing through registers, into a ring of DMA_STAGES buffers,
//    so DMA_STAGES - 1 chunks are in flight while one is multiplied. Rows are 64
//    elements, row r rotated by 16 (r mod 4) elements for the same bank spread.
constexpr int DMA_STAGES = 4;
template <typename T>
constexpr int lds_stride() {
  return sizeof(T) == 4 ? GT + 16 : GT;
}
template <typename T, int BM>
constexpr bool gemm_dma() {
  return sizeof(T) == 4 && BM == B_P32;
}
// buffers per operand, and elements per buffer
template <typename T, int BM>
constexpr int gemm_nbuf() {
  return gemm_dma<T, BM>() ? DMA_STAGES : 2;
}
template <typename T, int BM, int KC>
constexpr int gemm_lsz() {
  return gemm_dma<T, BM>() ? KC * GT : KC * lds_stride<T>();
}
// s_waitcnt immediate (gfx9 encoding) that waits until at most n vector memory
// operations of the wave are outstanding and does not wait on the other counters
constexpr int vmcnt_le(int n) {
  return (n & 15) | (7 << 4) | (15 << 8) | ((n >> 4) << 14);
}
#define LDS_AS __attribute__((address_space(3)))

// Direct form: wave w copies rows 4w .. 4w+3 of a chunk, A as one 16-byte piece a
// lane (asrc: this lane's piece), B one element a lane and one row an operation (5
// vector memory operations a wave in all). B is read at indices clamped into the
// direction (k < nk, n < N): where the clamp applies, the row's A is zero (the band
// pack is zero-filled past the band) or the column is outside the output, so the
// products are those of a zero-filled B. Row r of a buffer holds element x at
// (x + 16 (r mod 4)) mod 64.
constexpr int DMA_OPS = 5;
template <int KC>
__device__ __forceinline__ void dma_issue(float *wa, float *wb, const GLOBAL float *asrc, const GLOBAL float *p32,
                                          int kb, int nk, int N, int nb, int w, int lane) {
  __builtin_amdgcn_global_load_lds((const GLOBAL void *)asrc, (LDS_AS void *)(wa + w * 4 * GT), 16, 0, 0);
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int kk = 4 * w + u;
    const int k = kb + kk, n = nb + ((lane - 16 * u) & (GT - 1));
    const int kc = k < nk ? k : nk - 1, nc = n < N ? n : N - 1;
    __builtin_amdgcn_global_load_lds((const GLOBAL void *)(p32 + (int64_t)kc * N + nc),
                                     (LDS_AS void *)(wb + kk * GT), 4, 0, 0);
  }
}
// One chunk of the direct form: copies a later chunk (when `issue`) into wa/wb and
// multiplies the chunk in ra/rb into acc. Every pointer the copies and reads use is a
// restrict-qualified parameter, so that the compiler knows the copies do not write
// what is read here (it would otherwise wait for every copy in flight before the
// first read).
template <int KC>
__device__ __forceinline__ void dma_chunk(const float *__restrict__ ra, const float *__restrict__ rb,
                                          float *__restrict__ wa, float *__restrict__ wb, bool issue,
                                          const GLOBAL float *__restrict__ asrc, const GLOBAL float *__restrict__ p32,
                                          int kb, int nk, int N, int nb, int w, int lane, int wr0, int wc0,
                                          f4_t (&acc)[2][2]) {
  if (issue) dma_issue<KC>(wa, wb, asrc, p32, kb, nk, N, nb, w, lane);
  const int fk = lane >> 4, fi = lane & 15;
#pragma unroll
  for (int ks = 0; ks < KC; ks += 4) {
    const int r = ks + fk;  // r mod 4 = fk
    float af[2], bf[2];
#pragma unroll
    for (int i = 0; i < 2; i++) af[i] = ra[r * GT + ((wr0 + i * 16 + fi + 16 * fk) & (GT - 1))];
#pragma unroll
    for (int j = 0; j < 2; j++) bf[j] = rb[r * GT + ((wc0 + j * 16 + fi + 16 * fk) & (GT - 1))];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) acc[i][j] = MfmaT<float>::mfma(af[i], bf[j], acc[i][j]);
  }
}

// lds: A buffers then B buffers, gemm_nbuf() of gemm_lsz() elements each
template <typename T, int KC, int BM, int EM, bool MASK, bool DOT, bool LAT>
__device__ __forceinline__ void band_gemm_tile(const Prob &P, const GemmOps &ops, int lb, bool full, bool late, T *lds,
                                               double &redv) {
  constexpr int TR = GT, TC = GT, WC = 2, MR = 2, MC = 2;
  constexpr int AE = KC * TR / BLOCK;
  constexpr int BE = KC * TC / BLOCK;
  constexpr int LS = lds_stride<T>();
  constexpr bool FP32 = sizeof(T) == 4;
  // FP32: each thread stages 4 consecutive elements of A (one 16-byte load)
  static_assert(!FP32 || AE == 4, "FP32 A staging takes one float4 per thread");
  typedef MfmaT<T> Mf;
  const bool red = !full;
  const Band &band = ops.band == BAND_F ? P.bf : (ops.band == BAND_FT ? P.bt : P.bg);
  const int N = P.N;
  const int gx = (N + TC - 1) / TC;
  // E_PART: block lb is split `sp` of output tile lt; it multiplies band chunks
  // [c0, c1) of the tile and stores the raw sums into partial array sp
  const int sp = EM == E_PART ? lb / (gx * band.ntiles) : 0;
  const int lt = lb - sp * gx * band.ntiles;
  const int tile = lt / gx, n0 = (lt - tile * gx) * TC;
  const int r0 = tile * TR;
  const int nrows = ops.band == BAND_F ? P.D : P.M;
  const int nk = ops.band == BAND_FT ? P.D : P.M;
  const GLOBAL double *Bp = gp(P.op[ops.b]);
  GLOBAL double *Cp = EM == E_PART ? gp(P.hw) + (int64_t)sp * P.M * N : gp(P.op[ops.c]);
  const int tid = threadIdx.x, lane = lane_id(), w = wave_id();
  const int wr0 = (w / WC) * 32, wc0 = (w % WC) * 32;
  const int nch_t = gp(band.nch)[tile];
  const int c0 = EM == E_PART ? (sp * nch_t) / P.hs : 0;
  const int nch = EM == E_PART ? ((sp + 1) * nch_t) / P.hs - c0 : nch_t;
  const GLOBAL T *ap = gp(Mf::pack(band)) + gp(band.off)[tile] + (int64_t)c0 * (KC * TR);
  const int k0 = gp(band.klo)[tile] + c0 * KC;
  // E_GBAND: columns [gk0, gk0 + gkw) of row tile `tile` form its band of G
  const int gk0 = EM == E_GBAND ? gp(P.bg.klo)[tile] : 0;
  const int gkw = EM == E_GBAND ? gp(P.bg.nch)[tile] * KC : 0;
  if (EM == E_GBAND && (n0 + TC <= gk0 || n0 >= gk0 + gkw)) return;  // uniform over the block

  constexpr bool DMA = gemm_dma<T, BM>();
  constexpr int NBUF = gemm_nbuf<T, BM>(), LSZ = gemm_lsz<T, BM, KC>();
  static_assert(!DMA || (KC == 16 && BLOCK == 256 && DMA_STAGES == 4),
                "the direct staging maps 4 waves onto 16-row chunks and a 4-stage ring");
  T *const As = lds;               // As + b * LSZ: A buffer b
  T *const Bs = lds + NBUF * LSZ;  // Bs + b * LSZ: B buffer b

  static_assert(BM != B_P32 || FP32, "the FP32 direction copy feeds FP32 products only");
  T ra[AE], rb[BE];  // register form only
  auto gload = [&](int c) {
    const GLOBAL T *src = ap + (int64_t)c * (KC * TR);
    if constexpr (FP32) {
      const f4_t a4 = *(const GLOBAL f4_t *)(src + 4 * tid);
#pragma unroll
      for (int i = 0; i < AE; i++) ra[i] = a4[i];
    } else {
#pragma unroll
      for (int i = 0; i < AE; i++) ra[i] = src[tid + BLOCK * i];
    }
#pragma unroll
    for (int i = 0; i < BE; i++) {
      const int e = tid + BLOCK * i;
      const int kk = e / TC, cc = e - kk * TC;
      const int k = k0 + c * KC + kk, n = n0 + cc;
      double v = 0.0;
      if (BM == B_SPAN && k < nk && n < N) {
        const int rel = n - gp(P.flo)[k];
        if (rel >= 0 && rel < gp(P.flen)[k]) v = gp(P.fval)[gp(P.foff)[k] + rel];
      } else if (k < nk && n < N) {
        const int64_t idx = (int64_t)k * N + n;
        v = Bp[idx];
        if (BM == B_EXPAND && red) {
          // independent loads, selected afterwards
          const int mr = gp(P.maxr)[k];
          const double ek = gp(P.ev)[k];
          if (n == mr) v = ek;
        }
      }
      rb[i] = (T)v;
    }
  };
  auto sstore = [&](int buf) {
    T *a = As + buf * LSZ, *b = Bs + buf * LSZ;
    if constexpr (FP32) {
      // thread t holds elements 4t .. 4t+3 of the chunk: row t/16, columns 4 (t mod 16) ..
#pragma unroll
      for (int i = 0; i < AE; i++) a[(tid >> 4) * LS + 4 * (tid & 15) + i] = ra[i];
    } else {
#pragma unroll
      for (int i = 0; i < AE; i++) a[tid + BLOCK * i] = ra[i];
    }
#pragma unroll
    for (int i = 0; i < BE; i++) {
      const int e = tid + BLOCK * i;
      b[(e / TC) * LS + (e % TC)] = rb[i];
    }
  };
  const GLOBAL float *p32 = gp(P.p32);
  // this lane's 16-byte piece of A within a chunk (row 4w + lane/16, rotated)
  const int a_piece = (4 * w + (lane >> 4)) * TR + ((4 * (lane & 15) - 16 * (lane >> 4)) & (GT - 1));

  typename Mf::acc_t acc[MR][MC];
  d4_t acc64[MR][MC];  // FP32 path: chunk sums accumulated in FP64
#pragma unroll
  for (int i = 0; i < MR; i++)
#pragma unroll
    for (int j = 0; j < MC; j++) {
      acc[i][j] = typename Mf::acc_t{0, 0, 0, 0};
      acc64[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
    }

  if constexpr (DMA) {
#pragma unroll
    for (int c = 0; c < DMA_STAGES - 1; c++)
      if (c < nch)
        dma_issue<KC>(As + c * LSZ, Bs + c * LSZ, (const GLOBAL float *)ap + (int64_t)c * (KC * TR) + a_piece, p32,
                      k0 + c * KC, nk, N, n0, w, lane);
  } else {
    if (nch > 0) {
      gload(0);
      sstore(0);
    }
    __syncthreads();
  }
  const int fk = lane >> 4, fi = lane & 15;
  for (int c = 0; c < nch; c++) {
    if constexpr (DMA) {
      // this wave's copies of chunk c have landed once at most the copies of the
      // chunks issued after it are outstanding; the barrier then makes the whole
      // chunk visible and tells every wave that buffer (c - 1) mod 4 has been read,
      // so chunk c + 3 may overwrite it
      const int later = nch - 1 - c;
      if (later >= DMA_STAGES - 2)
        __builtin_amdgcn_s_waitcnt(vmcnt_le((DMA_STAGES - 2) * DMA_OPS));
      else if (later == 1)
        __builtin_amdgcn_s_waitcnt(vmcnt_le(DMA_OPS));
      else
        __builtin_amdgcn_s_waitcnt(vmcnt_le(0));
      __builtin_amdgcn_s_barrier();
      const int st = c & (DMA_STAGES - 1), wt = (c + DMA_STAGES - 1) & (DMA_STAGES - 1);
      const int cn = c + DMA_STAGES - 1;
      dma_chunk<KC>(As + st * LSZ, Bs + st * LSZ, As + wt * LSZ, Bs + wt * LSZ, cn < nch,
                      (const GLOBAL float *)ap + (int64_t)cn * (KC * TR) + a_piece, p32, k0 + cn * KC, nk, N, n0, w,
                      lane, wr0, wc0, acc);
    } else {
      const int buf = c & 1;
      if (c + 1 < nch) gload(c + 1);
      const T *a = As + buf * LSZ, *b = Bs + buf * LSZ;
#pragma unroll
      for (int ks = 0; ks < KC; ks += 4) {
        T af[MR], bf[MC];
#pragma unroll
        for (int i = 0; i < MR; i++) af[i] = a[(ks + fk) * LS + wr0 + i * 16 + fi];
#pragma unroll
        for (int j = 0; j < MC; j++) bf[j] = b[(ks + fk) * LS + wc0 + j * 16 + fi];
#pragma unroll
        for (int i = 0; i < MR; i++)
#pragma unroll
          for (int j = 0; j < MC; j++) acc[i][j] = Mf::mfma(af[i], bf[j], acc[i][j]);
      }
    }
    if (FP32) {
#pragma unroll
      for (int i = 0; i < MR; i++)
#pragma unroll
        for (int j = 0; j < MC; j++)
#pragma unroll
          for (int q = 0; q < 4; q++) {
            acc64[i][j][q] += (double)acc[i][j][q];
            acc[i][j][q] = 0;
          }
    }
    if constexpr (!DMA) {
      if (c + 1 < nch) sstore((c & 1) ^ 1);
      __syncthreads();
    }
  }

  // Epilogue. LAT (latency form, for launches too small to fill the chip): in
  // batches of QB outputs (q) of one MFMA tile (i, j) per lane, every operand of a
  // batch is loaded first (indices clamped into range, values of out-of-range
  // outputs discarded) so the loads overlap instead of each waiting on the
  // previous output's store; then the outputs are formed in (i, j, q) order.
  // Otherwise one output at a time (fewer registers: more waves per SIMD, which
  // hides the same latency when the launch fills the chip). Same arithmetic, in
  // the same order, so the two forms give the same bits.
  constexpr int QB = 2;
  const double g2 = 2.0 * P.gamma;
  constexpr bool GH = (EM == E_GRAD || EM == E_HVP);
  constexpr bool LOSS = (EM == E_LOSS);
  constexpr bool WITH_P = (EM == E_LOSS || EM == E_STORE_R);
  const bool exp = (EM == E_HVP) && red;
  const bool sten = GH && P.gamma != 0.0;
  const GLOBAL double *__restrict__ E = GH ? gp(P.op[ops.e]) : nullptr;
  const GLOBAL double *__restrict__ dd = DOT ? gp(P.op[ops.dd]) : nullptr;
  // the late fold exists for the masked FP32 Hessian product with its dot (CG)
  constexpr bool LATE_OK = FP32 && EM == E_HVP && MASK && DOT;
  // the late fold writes gm (the row's fold value) instead of reading it
  const bool lfold = LATE_OK && late && red;
  static_assert(!(LAT && EM == E_GBAND), "the Gram band is formed in the occupancy form only");
  if constexpr (LAT) {
#pragma unroll
  for (int i = 0; i < MR; i++)
#pragma unroll
    for (int j = 0; j < MC; j++)
#pragma unroll
      for (int q0 = 0; q0 < 4; q0 += QB) {
        // each batch's loads come after the previous batch's stores, which bounds
        // the registers held by loads in flight to one batch
        __asm__ volatile("" ::: "memory");
        const int n = n0 + wc0 + j * 16 + fi;
        const int nc = n < N ? n : N - 1;
        int row[QB], mr0[QB];
        double s0[QB], sp[QB], sm[QB], gmv[QB], pv[QB], dv[QB], pbv[QB];
        uint8_t ikv[QB], fh[QB];
#pragma unroll
        for (int u = 0; u < QB; u++) {
          row[u] = r0 + wr0 + i * 16 + Mf::row(q0 + u, fk);
          const int rc = row[u] < nrows ? row[u] : nrows - 1;
          const int rp = rc + 1 < nrows ? rc + 1 : rc;
          const int rm = rc > 0 ? rc - 1 : rc;
          const int64_t idx = (int64_t)rc * N + nc;
          if (LOSS) fh[u] = gp(P.fhas)[rc];
          if (WITH_P) pv[u] = gp(P.P)[idx];
          if (GH) {
            mr0[u] = red ? gp(P.maxr)[rc] : -1;
            gmv[u] = red && !lfold ? gp(P.gm)[rc] : 0.0;
            if (LATE_OK) pbv[u] = lfold ? gp(P.pb)[rc] : 0.0;
            if (sten) {
              s0[u] = E[idx];
              sp[u] = E[(int64_t)rp * N + nc];
              sm[u] = E[(int64_t)rm * N + nc];
              if (exp) {
                const int mrp = gp(P.maxr)[rp], mrm = gp(P.maxr)[rm];
                const double e0 = gp(P.ev)[rc], ep = gp(P.ev)[rp], em = gp(P.ev)[rm];
                if (nc == mr0[u]) s0[u] = e0;
                if (nc == mrp) sp[u] = ep;
                if (nc == mrm) sm[u] = em;
              }
            }
          }
          if (MASK) ikv[u] = gp(P.ik)[idx];
          if (DOT) dv[u] = dd[idx];
        }
#pragma unroll
        for (int u = 0; u < QB; u++) {
          const int q = q0 + u;
          if (row[u] >= nrows || n >= N) continue;
          const int64_t idx = (int64_t)row[u] * N + n;
          const double v = FP32 ? acc64[i][j][q] : (double)acc[i][j][q];
          if (EM == E_STORE || EM == E_STORE_R || EM == E_PART) Cp[idx] = v;
          if (EM == E_STORE_R) gp(P.op[OP_R])[idx] = v - pv[u];
          if (LOSS && fh[u]) {
            const double t = v - pv[u];
            redv = fma(t, t, redv);
          }
          if (GH) {
            double g = 2.0 * v;
            if (sten) {
              if (row[u] + 1 < nrows) g = g + g2 * (s0[u] - sp[u]);
              if (row[u] > 0) g = g - g2 * (sm[u] - s0[u]);
            }
            if (LATE_OK && lfold) {
              // the late fold, as in the occupancy form below
              const bool m = ikv[u] != 0;
              if (n == mr0[u]) {
                gp(P.gm)[row[u]] = g;
                redv = fma(-g, pbv[u], redv);
              } else if (!m) {
                redv = fma(dv[u], g, redv);
              }
              Cp[idx] = m ? 0.0 : g;
            } else {
              if (red) g = (n == mr0[u]) ? 0.0 : g - gmv[u];
              if (MASK && ikv[u]) g = 0.0;
              Cp[idx] = g;
              if (DOT) redv = fma(dv[u], g, redv);
            }
          }
        }
      }
  } else {
#pragma unroll
    for (int i = 0; i < MR; i++)
#pragma unroll
      for (int j = 0; j < MC; j++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int row = r0 + wr0 + i * 16 + Mf::row(q, fk);
          const int n = n0 + wc0 + j * 16 + fi;
          if (row >= nrows || n >= N) continue;
          const int64_t idx = (int64_t)row * N + n;
          const double v = FP32 ? acc64[i][j][q] : (double)acc[i][j][q];
          if (EM == E_STORE || EM == E_STORE_R || EM == E_PART) Cp[idx] = v;
          if (EM == E_STORE_R) gp(P.op[OP_R])[idx] = v - gp(P.P)[idx];
          if (EM == E_GBAND && n >= gk0 && n < gk0 + gkw)
            gp(const_cast<double *>(P.bg.pack))[gp(P.bg.off)[tile] + (int64_t)(n - gk0) * GT + (row - r0)] = v;
          if (LOSS && gp(P.fhas)[row]) {
            const double t = v - gp(P.P)[idx];
            redv = fma(t, t, redv);
          }
          if (GH) {
            double g = 2.0 * v;
            if (sten) {
              const double e0 = stencil_src(exp, E, gp(P.maxr), gp(P.ev), row, n, N);
              if (row + 1 < nrows) g = g + g2 * (e0 - stencil_src(exp, E, gp(P.maxr), gp(P.ev), row + 1, n, N));
              if (row > 0) g = g - g2 * (stencil_src(exp, E, gp(P.maxr), gp(P.ev), row - 1, n, N) - e0);
            }
            if (LATE_OK && late && red) {
              // fold left to kb_cg_update: store g (0 where masked) and, at the
              // implicit column, the row's fold value; the dot uses
              // sum_{n != maxr, free} p (g - gm) = sum p g - gm sum p
              const int mr = gp(P.maxr)[row];
              const bool m = gp(P.ik)[idx] != 0;
              if (n == mr) {
                gp(P.gm)[row] = g;
                redv = fma(-g, gp(P.pb)[row], redv);
              } else if (!m) {
                redv = fma(dd[idx], g, redv);
              }
              Cp[idx] = m ? 0.0 : g;
            } else {
              if (red) g = (n == gp(P.maxr)[row]) ? 0.0 : g - gp(P.gm)[row];
              if (MASK && gp(P.ik)[idx]) g = 0.0;
              Cp[idx] = g;
              if (DOT) redv = fma(dd[idx], g, redv);
            }
          }
        }
  }
}

// Smoothing part of the objective (pqdts.f90:956-962), block lb of nblk: partial
// of sum gamma (X[j,n] - X[j+1,n])^2 over j < M-1, elements [lb, lb+1) x EBLOCK
// of the (M-1) x N differences, into part[region][lb]; pn[region] = nblk.
__device__ __forceinline__ void smooth_part(const Prob &P, int x, int lb, int nblk, int region, double *sh) {
  const int64_t total = (int64_t)(P.M - 1) * P.N;
  const GLOBAL double *X = gp(P.op[x]);
  double s = 0.0;
  for (int64_t e = (int64_t)lb * EBLOCK + threadIdx.x; e < lmin64(total, (int64_t)(lb + 1) * EBLOCK);
       e += BLOCK) {
    const double t = X[e] - X[e + P.N];
    s += P.gamma * (t * t);
  }
  const double b = block_sum(s, sh);
  if (threadIdx.x == 0) gp(P.part)[(int64_t)region * P.pcap + lb] = b;
  if (threadIdx.x == 0 && lb == 0) gp(P.pn)[region] = nblk;
}

// With ops.sm >= 0 (E_LOSS, B = X) the blocks past the product's own compute
// the smoothing partials of X: one launch for both terms of the objective.
template <typename T, int KC, int BM, int EM, bool MASK, bool DOT, bool LAT>
__global__ __launch_bounds__(BLOCK) void kb_band_gemm(const Prob *__restrict__ probs, WorkList wl, GemmOps ops) {
  // One LDS array for the chunk buffers and the block-sum scratch, which overlaps the
  // first buffer and is used once the product is done (after a barrier). With a second
  // LDS variable in the kernel the compiler tags every LDS access with the variable it
  // belongs to and loses the distinction between the ring buffers of the direct form,
  // whose reads it would then make wait for every copy in flight; and 32 KB exactly
  // leaves room for 5 workgroups a CU.
  __shared__ __align__(16) T lds[2 * gemm_nbuf<T, BM>() * gemm_lsz<T, BM, KC>()];
  double *const sh = reinterpret_cast<double *>(lds);
  BATCH_PROLOGUE();
  int ng = nblk;  // blocks of the product itself
  if (EM == E_LOSS && ops.sm >= 0) {
    ng = ((P.N + GT - 1) / GT) * P.bf.ntiles;
    if (lb >= ng) {
      smooth_part(P, ops.b, lb - ng, nblk - ng, ops.sm, sh);
      return;
    }
  }
  double redv = 0.0;
  band_gemm_tile<T, KC, BM, EM, MASK, DOT, LAT>(P, ops, lb, full, late, lds, redv);
  if (EM == E_LOSS || DOT) {
    __syncthreads();  // every wave is done with the chunk buffers that sh overlaps
    const double b = block_sum(redv, sh);
    if (threadIdx.x == 0) gp(P.part)[(int64_t)ops.region * P.pcap + lb] = b;
    if (threadIdx.x == 0 && lb == 0) gp(P.pn)[ops.region] = ng;
  }
}

// Pre-fold value G[k, maxr[k]] of the transposed product (one wave per k), used
// by the fold in the transposed band GEMM; same formula as its epilogue at
// n = maxr[k]. Reduced-space problems only.
template <int MODE>
__global__ __launch_bounds__(BLOCK) void kb_foldval(const Prob *__restrict__ probs, WorkList wl, int esrc) {
  BATCH_PROLOGUE();
  const int k = lb * WAVES_PER_BLOCK + wave_id();
  if (k >= P.M) return;
  const int N = P.N;
  const GLOBAL double *E = gp(P.op[esrc]);
  const GLOBAL double *O = gp(P.op[OP_O]);
  const int c = gp(P.maxr)[k];
  double acc = 0.0;
  for (int d = gp(P.kdlo)[k] + lane_id(); d <= gp(P.kdhi)[k]; d += WAVE) {
    const int rel = k - gp(P.flo)[d];
    if (rel >= 0 && rel < gp(P.flen)[d]) {
      double wv = O[(int64_t)d * N + c];
      if (MODE == T_GRAD) wv -= gp(P.P)[(int64_t)d * N + c];
      acc = fma(gp(P.fval)[gp(P.foff)[d] + rel], wv, acc);
    }
  }
  acc = wave_sum(acc);
  if (lane_id() == 0) {
    const bool exp = (MODE == T_HVP);
    double g = 2.0 * acc;
    if (P.gamma != 0.0) {
      const double g2 = 2.0 * P.gamma;
      const double e0 = stencil_src(exp, E, gp(P.maxr), gp(P.ev), k, c, N);
      if (k + 1 < P.M) g = g + g2 * (e0 - stencil_src(exp, E, gp(P.maxr), gp(P.ev), k + 1, c, N));
      if (k > 0) g = g - g2 * (stencil_src(exp, E, gp(P.maxr), gp(P.ev), k - 1, c, N) - e0);
    }
    gp(P.gm)[k] = g;
  }
}

// Gram path of BF: pre-fold value (2 G p~ + stencil)[k, maxr[k]] from the packed
// band of G = F^T F, one wave per k; p~ is dd with its implicit column replaced
// by ev. Reduced-space problems only.
__device__ __forceinline__ void foldval_gram_row(const Prob &P, const GLOBAL double *__restrict__ dd, int k, int kc,
                                                 int tile) {
  const int N = P.N, M = P.M;
  const int c = gp(P.maxr)[k];
  const int t = k / tile, r = k - t * tile;
  const int l0 = gp(P.bg.klo)[t], nl = gp(P.bg.nch)[t] * kc;
  const GLOBAL double *gpk = gp(P.bg.pack) + gp(P.bg.off)[t] + r;
  double acc = 0.0;
  for (int j = lane_id(); j < nl; j += WAVE) {
    const int l = l0 + j;
    if (l >= M) break;
    const double p = (c == gp(P.maxr)[l]) ? gp(P.ev)[l] : dd[(int64_t)l * N + c];
    acc = fma(gpk[(int64_t)j * tile], p, acc);
  }
  acc = wave_sum(acc);
  if (lane_id() == 0) {
    double g = 2.0 * acc;
    if (P.gamma != 0.0) {
      const double g2 = 2.0 * P.gamma;
      const double e0 = stencil_src(true, dd, gp(P.maxr), gp(P.ev), k, c, N);
      if (k + 1 < M) g = g + g2 * (e0 - stencil_src(true, dd, gp(P.maxr), gp(P.ev), k + 1, c, N));
      if (k > 0) g = g - g2 * (stencil_src(true, dd, gp(P.maxr), gp(P.ev), k - 1, c, N) - e0);
    }
    gp(P.gm)[k] = g;
  }
}

__global__ __launch_bounds__(BLOCK) void kb_foldval_gram(const Prob *__restrict__ probs, WorkList wl, int ddsrc,
                                                         int kc, int tile) {
  BATCH_PROLOGUE();
  const int k = lb * WAVES_PER_BLOCK + wave_id();
  if (k >= P.M) return;
  foldval_gram_row(P, gp(P.op[ddsrc]), k, kc, tile);
}

// Bdiag (pqdts.f90:208-244), one wave per k: 2 sum_d F[d,k]^2 + 2 gamma [k<M-1]
// + 2 gamma [k>0]; the reduced-space copy is twice that (pqdts.f90:235-236).
__global__ __launch_bounds__(BLOCK) void kb_bdiag(const Prob *__restrict__ probs, WorkList wl) {
  BATCH_PROLOGUE();
  const int k = lb * WAVES_PER_BLOCK + wave_id();
  if (k >= P.M) return;
  double acc = 0.0;
  for (int d = gp(P.kdlo)[k] + lane_id(); d <= gp(P.kdhi)[k]; d += WAVE) {
    const int rel = k - gp(P.flo)[d];
    if (rel >= 0 && rel < gp(P.flen)[d]) {
      const double f = gp(P.fval)[gp(P.foff)[d] + rel];
      acc += 2.0 * (f * f);
    }
  }
  acc = wave_sum(acc);
  if (lane_id() == 0) {
    if (P.gamma != 0.0) {
      if (k + 1 < P.M) acc += 2.0 * P.gamma;
      if (k > 0) acc += 2.0 * P.gamma;
    }
    gp(const_cast<double *>(P.diag_full))[k] = acc;
    gp(const_cast<double *>(P.diag_red))[k] = acc * 2.0;
  }
}

// ---------------------------------------------------------------------------
// Row kernels: one wave per row j of an M x N array. Each lane only re-reads
// the elements it wrote itself; other lanes' values travel by wave reductions.
// ---------------------------------------------------------------------------
#define ROW_PROLOGUE()                                \
  BATCH_PROLOGUE();                                   \
  const int j = lb * WAVES_PER_BLOCK + wave_id();     \
  const int N = P.N;                                  \
  (void)N

__device__ __forceinline__ double row_sum(const GLOBAL double *__restrict__ row, int N) {
  double s = 0.0;
  for (int n = lane_id(); n < N; n += WAVE) s += row[n];
  return wave_sum(s);
}

// X[j,maxr] = 1 - (sum X[j,:] - X[j,maxr])   (pqdts.f90:931-933); reduced problems only
__global__ __launch_bounds__(BLOCK) void kb_reconstruct(const Prob *__restrict__ probs, WorkList wl, int x) {
  ROW_PROLOGUE();
  if (j >= P.M || full) return;
  GLOBAL double *row = gp(P.op[x]) + (int64_t)j * N;
  const double s = row_sum(row, N);
  if (lane_id() == 0) {
    const int r = gp(P.maxr)[j];
    row[r] = 1.0 - (s - row[r]);
  }
}

// ev[j] = -(sum V[j,:] - V[j,maxr])   (pqdts.f90:264); reduced problems only
__global__ __launch_bounds__(BLOCK) void kb_others(const Prob *__restrict__ probs, WorkList wl, int v) {
  ROW_PROLOGUE();
  if (j >= P.M || full) return;
  const GLOBAL double *row = gp(P.op[v]) + (int64_t)j * N;
  const double s = row_sum(row, N);
  if (lane_id() == 0) gp(P.ev)[j] = -(s - row[gp(P.maxr)[j]]);
}

// Reduced-space transform of X (pqdts.f90:1970-1980; mode 1 without the first
// step, pqdts.f90:1940-1944): rebuild the old implicit entry, pick the first
// maximum of the row, store the placeholder 1.0 there. Full-space problems get
// maxr = -1 instead (pqdts.f90:1982).
__global__ __launch_bounds__(BLOCK) void kb_transform(const Prob *__restrict__ probs, WorkList wl, int rebuild) {
  ROW_PROLOGUE();
  if (j >= P.M) return;
  if (full) {
    if (lane_id() == 0) gp(P.maxr)[j] = -1;
    return;
  }
  GLOBAL double *row = gp(P.op[0]) + (int64_t)j * N;
  const int old = gp(P.maxr)[j];
  const bool rb = rebuild && old >= 0 && old < N;
  double newv = 0.0;
  if (rb) {
    const double s = row_sum(row, N);
    newv = 1.0 - (s - row[old]);
  }
  double best = -INFINITY;
  int bi = 0x7fffffff;
  for (int n = lane_id(); n < N; n += WAVE) {
    const double v = (rb && n == old) ? newv : row[n];
    if (v > best || bi == 0x7fffffff) {
      best = v;
      bi = n;
    }
  }
  wave_argmax(best, bi);
  if (lane_id() == 0) {
    if (rb) row[old] = newv;
    row[bi] = 1.0;
    gp(P.maxr)[j] = bi;
  }
}

// Projection of the rows of `t` (optionally first forming t = a*SN + 1*X with a
// per-problem `a`, pgaxpy2 at pqdts.f90:524):
//   full space:    Euclidean projection onto the simplex (doproj, pqdts.f90:1120-1140;
//                  the reference calls Condat's simplexproj_condat_ per row). Here:
//                  Michelot's fixed point tau <- (sum_{y>tau} y - 1)/#{y>tau} from
//                  the row mean, which increases monotonically to the same threshold.
//   reduced space: doproj2 (pqdts.f90:1149-1158): clip negatives, rebuild the
//                  implicit entry; per-block partial of xmin = min(0, min_j t[j,maxr]);
//                  with `recon_again` the objective's own rebuild (pqdts.f90:929-935)
//                  follows.
__device__ __forceinline__ void project_block(const Prob &P, int lb, int nblk, bool full, double arg, int t,
                                              int form_trial, int recon_again, int region, double *sh) {
  const int j = lb * WAVES_PER_BLOCK + wave_id();
  const int N = P.N;
  double xm = 0.0;
  if (j < P.M) {
    GLOBAL double *row = gp(P.op[t]) + (int64_t)j * N;
    const GLOBAL double *xr = gp(P.op[0]) + (int64_t)j * N;
    const GLOBAL double *sr = gp(P.op[4]) + (int64_t)j * N;
    if (full) {
      double s = 0.0;
      for (int n = lane_id(); n < N; n += WAVE) {
        double v;
        if (form_trial) {
          v = arg * sr[n] + 1.0 * xr[n];
          row[n] = v;
        } else {
          v = row[n];
        }
        s += v;
      }
      s = wave_sum(s);
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
    } else {
      const int r = gp(P.maxr)[j];
      double s = 0.0, own = 0.0;
      for (int n = lane_id(); n < N; n += WAVE) {
        double v = form_trial ? arg * sr[n] + 1.0 * xr[n] : row[n];
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
  }
  xm = wave_min(xm);
  if (lane_id() == 0) sh[wave_id()] = xm;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = sh[0];
    for (int w = 1; w < WAVES_PER_BLOCK; w++) v = fmin(v, sh[w]);
    gp(P.part)[(int64_t)region * P.pcap + lb] = v;
    if (lb == 0) gp(P.pn)[region] = nblk;
  }
}

__global__ __launch_bounds__(BLOCK) void kb_project(const Prob *__restrict__ probs, WorkList wl, int t,
                                                    int form_trial, int recon_again, int region) {
  __shared__ double sh[WAVES_PER_BLOCK];
  BATCH_PROLOGUE();
  project_block(P, lb, nblk, full, arg, t, form_trial, recon_again, region, sh);
}

// The full-space projection of the rows of operand 0 with the descriptor passed
// by value (kernel arguments), for callers that keep no descriptor in device
// memory: one block per WAVES_PER_BLOCK rows, the same arithmetic as kb_project.
__global__ __launch_bounds__(BLOCK) void kb_project_byval(const Prob P, int region) {
  __shared__ double sh[WAVES_PER_BLOCK];
  project_block(P, blockIdx.x, gridDim.x, true, 0.0, 0, 0, 0, region, sh);
}

// Condat's simplex projection (condat_simplexproj.c:60-95) of one row, in the
// reference's summation order so that tau is bit-identical, one lane per row: the
// filter is sequential. `x` (the output row) doubles as the candidate list, as
// in the reference: candidates parked by a restart of the online filter sit in
// x[0, base), the live candidates in x[base, base + nlive).
__device__ void condat_row(const GLOBAL double *__restrict__ y, GLOBAL double *x, int n) {
  // pass 1 (lines 66-75): online filter with restarts
  int base = 0, nlive = 1;
  x[0] = y[0];
  double tau = y[0] - 1.0;
  for (int i = 1; i < n; i++) {
    const double v = y[i];
    if (!(v > tau)) continue;
    x[base + nlive] = v;
    tau += (v - tau) / (double)(nlive + 1);
    if (tau <= v - 1.0) {
      // restart: everything before v is parked, v alone is live
      tau = v - 1.0;
      base += nlive;
      nlive = 1;
    } else {
      nlive++;
    }
  }
  // pass 2 (lines 76-82): parked values, newest first, re-enter in front of the
  // live list when above tau
  int head = base;
  for (int k = base - 1; k >= 0; k--) {
    const double v = x[k];
    if (v > tau) {
      x[--head] = v;
      nlive++;
      tau += (v - tau) / (double)nlive;
    }
  }
  GLOBAL double *live = x + head;
  // pass 3 (lines 83-89): sweeps dropping values <= tau until none is dropped
  for (;;) {
    const int last = nlive - 1;
    int kept = 0;
    for (int i = 0; i <= last; i++) {
      const double v = live[i];
      if (v > tau)
        live[kept++] = v;
      else
        tau += (tau - v) / (double)(last - i + kept);
    }
    const bool dropped = kept <= last;
    nlive = kept;
    if (!dropped) break;
  }
  // lines 90-91
  for (int i = 0; i < n; i++) x[i] = y[i] > tau ? y[i] - tau : 0.0;
}

// rows x n values from y into x, one lane per row (x must not alias y)
__global__ __launch_bounds__(BLOCK) void kb_condat_rows(const double *__restrict__ y, double *x, int64_t rows,
                                                        int n) {
  const int64_t r = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (r >= rows) return;
  condat_row(gp(y) + r * n, gp(x) + r * n, n);
}

// conv_test rows (pqdts.f90:1055-1060): lambda_i = max(0, -min_n G[i,n]), partial
// of sum ((G + lambda) X)^2; with `store`, lambda and min(G+lambda) per row.
__global__ __launch_bounds__(BLOCK) void kb_conv_rows(const Prob *__restrict__ probs, WorkList wl, int x, int g,
                                                      int store, int region) {
  __shared__ double sh[WAVES_PER_BLOCK];
  ROW_PROLOGUE();
  double s = 0.0;
  if (j < P.M) {
    const GLOBAL double *gr = gp(P.op[g]) + (int64_t)j * N, *xr = gp(P.op[x]) + (int64_t)j * N;
    double mn = INFINITY;
    for (int n = lane_id(); n < N; n += WAVE) mn = fmin(mn, gr[n]);
    mn = wave_min(mn);
    const double lambda = fmax(0.0, -mn);
    for (int n = lane_id(); n < N; n += WAVE) {
      const double t = (gr[n] + lambda) * xr[n];
      s = fma(t, t, s);
    }
    if (store) {
      double m2 = INFINITY;
      for (int n = lane_id(); n < N; n += WAVE) m2 = fmin(m2, gr[n] + lambda);
      m2 = wave_min(m2);
      if (lane_id() == 0) {
        gp(P.lam)[j] = lambda;
        gp(P.lmin)[j] = m2;
      }
    }
  }
  const double b = block_sum(s, sh);
  if (threadIdx.x == 0) gp(P.part)[(int64_t)region * P.pcap + lb] = b;
}

// ---------------------------------------------------------------------------
// Element kernels over the M x N arrays: blocks of BLOCK x ELEMS elements.
// ---------------------------------------------------------------------------

// pdot (pqdts.f90:589-614)
__global__ __launch_bounds__(BLOCK) void kb_dot(const Prob *__restrict__ probs, WorkList wl, int x, int y,
                                                int region) {
  __shared__ double sh[WAVES_PER_BLOCK];
  BATCH_PROLOGUE();
  const int64_t total = (int64_t)P.M * P.N;
  const GLOBAL double *xa = gp(P.op[x]), *ya = gp(P.op[y]);
  double s = 0.0;
  for (int64_t e = (int64_t)lb * EBLOCK + threadIdx.x; e < lmin64(total, (int64_t)(lb + 1) * EBLOCK);
       e += BLOCK)
    s = fma(xa[e], ya[e], s);
  const double b = block_sum(s, sh);
  if (threadIdx.x == 0) gp(P.part)[(int64_t)region * P.pcap + lb] = b;
}

// Epilogue of a K-split Gram Hessian product (pqdts.f90:280-330 with G = F^T F),
// one thread per element: the hs partial products are summed in split order, then
// bx = 2 v + the gamma stencil of dd~ (implicit column expanded), the reduced-space
// fold, the active-set mask, and the dd.bx partial (region) - the arithmetic of the
// product's own E_HVP epilogue. The fold value (the same expression at column
// maxr[j], which kb_foldval_gram forms for the unsplit product) is formed here
// from the partials of that column.
__device__ __forceinline__ double hvp_epi_val(const Prob &P, bool red, const GLOBAL double *__restrict__ E,
                                              const GLOBAL double *__restrict__ W, int64_t MN, int j, int n) {
  const int M = P.M, N = P.N;
  const int64_t e = (int64_t)j * N + n;
  double v = W[e];
  for (int s2 = 1; s2 < P.hs; s2++) v += W[s2 * MN + e];
  double g = 2.0 * v;
  if (P.gamma != 0.0) {
    const double g2 = 2.0 * P.gamma;
    const double e0 = stencil_src(red, E, gp(P.maxr), gp(P.ev), j, n, N);
    if (j + 1 < M) g = g + g2 * (e0 - stencil_src(red, E, gp(P.maxr), gp(P.ev), j + 1, n, N));
    if (j > 0) g = g - g2 * (stencil_src(red, E, gp(P.maxr), gp(P.ev), j - 1, n, N) - e0);
  }
  return g;
}

template <bool MASK, bool DOT>
__global__ __launch_bounds__(BLOCK) void kb_hvp_epi(const Prob *__restrict__ probs, WorkList wl, int dd, int bx,
                                                    int region) {
  __shared__ double sh[WAVES_PER_BLOCK];
  BATCH_PROLOGUE();
  const bool red = !full;
  const int N = P.N;
  const int64_t MN = (int64_t)P.M * N;
  const GLOBAL double *E = gp(P.op[dd]);
  GLOBAL double *C = gp(P.op[bx]);
  const GLOBAL double *W = gp(P.hw);
  double redv = 0.0;
  int64_t e = (int64_t)lb * EBLOCK + threadIdx.x;
  const int64_t eend = lmin64(MN, (int64_t)(lb + 1) * EBLOCK);
  // (j, n) of e, advanced without a 64-bit division per element
  int j = (int)(e / N), n = (int)(e - (int64_t)j * N);
  for (; e < eend; e += BLOCK) {
    double g = hvp_epi_val(P, red, E, W, MN, j, n);
    if (red) {
      const int mr = gp(P.maxr)[j];
      g = (n == mr) ? 0.0 : g - hvp_epi_val(P, red, E, W, MN, j, mr);
    }
    if (MASK && gp(P.ik)[e]) g = 0.0;
    C[e] = g;
    if (DOT) redv = fma(E[e], g, redv);
    n += BLOCK;
    if (n >= N) {
      const int q = n / N;
      j += q;
      n -= q * N;
    }
  }
  if (DOT) {
    const double b = block_sum(redv, sh);
    if (threadIdx.x == 0) gp(P.part)[(int64_t)region * P.pcap + lb] = b;
    if (threadIdx.x == 0 && lb == 0) gp(P.pn)[region] = nblk;
  }
}

// BLAS-1 of the reference (pqdts.f90:485-587): z = beta*y + alpha*x and its
// special cases; alpha = the per-problem argument when `per_problem_alpha`.
enum AxKind { AX_GAXPY = 0, AX_AXPY = 1, AX_SCALE = 2, AX_COPY = 3, AX_SET = 4 };
__global__ __launch_bounds__(BLOCK) void kb_axpby(const Prob *__restrict__ probs, WorkList wl, int z, int x, int y,
                                                  double alpha, double beta, int kind, int per_problem_alpha) {
  BATCH_PROLOGUE();
  const int64_t total = (int64_t)P.M * P.N;
  const double al = per_problem_alpha ? arg : alpha;
  GLOBAL double *za = gp(P.op[z]);
  const GLOBAL double *xa = x >= 0 ? gp(P.op[x]) : nullptr, *ya = y >= 0 ? gp(P.op[y]) : nullptr;
  for (int64_t e = (int64_t)lb * EBLOCK + threadIdx.x; e < lmin64(total, (int64_t)(lb + 1) * EBLOCK);
       e += BLOCK) {
    double v;
    switch (kind) {
      case AX_GAXPY: v = beta * ya[e] + al * xa[e]; break;
      case AX_AXPY: v = ya[e] + al * xa[e]; break;
      case AX_SCALE: v = xa[e] * al; break;
      case AX_COPY: v = xa[e]; break;
      default: v = al; break;
    }
    za[e] = v;
  }
}

// ---------------------------------------------------------------------------
// Newton-CG vector kernels (pqdts.f90:1986-2104), one wave per row.
// Operands: X=0, G=1 (r), HP=2, XTMP=3 (p), SN=4, Z=5.
// ---------------------------------------------------------------------------

// FP32 copy of the CG direction (operand 3) with the implicit column (mr, -1 in
// full space) replaced by ev[j]: the B operand of the mixed-precision Hessian
// product, rounded once where the direction is formed instead of in every tile
// that reads it. p32_elem runs for each entry v = p[j,n] in the loop that forms
// the row; with `late` (F_LATE, reduced problems) it also sums b = p over the
// free entries (ik = 0) outside the implicit column, which p32_row_end stores as
// pb[j] for the Hessian product's dot, with the implicit column's entry.
__device__ __forceinline__ void p32_elem(const Prob &P, int64_t idx, int n, int mr, bool late, bool active, double v,
                                         double &b) {
  if (n != mr) {
    gp(P.p32)[idx] = (float)v;
    if (late && !active) b += v;
  }
}
__device__ __forceinline__ void p32_row_end(const Prob &P, int j, int mr, bool late, double evj, double b) {
  if (mr >= 0 && lane_id() == 0) gp(P.p32)[(int64_t)j * P.N + mr] = (float)evj;
  if (late) {
    b = wave_sum(b);
    if (lane_id() == 0) gp(P.pb)[j] = b;
  }
}

// Preconditioner divisor of row j: sqrt(diag) in the reference's PCG
// (pqdts.f90:2028, 2067), diag itself (Jacobi) in a polish iteration.
__device__ __forceinline__ double prec_div(const Prob &P, int j, bool full, bool polish) {
  const double d = full ? gp(P.diag_full)[j] : gp(P.diag_red)[j];
  return polish ? d : sqrt(d);
}

// Active set + CG start (pqdts.f90:1986-2031): ik = X < tol and G > 0, G[ik] = 0;
// partials of nact (region 0), |G|^2 (region 1) and r.z (region 2); r = -G;
// p = z = r / sqrt(diag) (z itself is not stored); sn = 0;
// ev[j] = -(sum p[j,:] - p[j,maxr]) when reduced; the FP32 copy of p when P.p32.
__global__ __launch_bounds__(BLOCK) void kb_cg_init(const Prob *__restrict__ probs, WorkList wl, int tactive) {
  __shared__ double sh[WAVES_PER_BLOCK];
  ROW_PROLOGUE();
  double nact = 0.0, nd = 0.0, b1 = 0.0;
  if (j < P.M) {
    const double sq = prec_div(P, j, full, polish);
    const int mr = full ? -1 : gp(P.maxr)[j];
    double psum = 0.0, own = 0.0, pb = 0.0;
    const bool lt = late && !full;
    const int64_t base = (int64_t)j * N;
    for (int n = lane_id(); n < N; n += WAVE) {
      const int64_t idx = base + n;
      double g = gp(P.op[1])[idx];
      uint8_t a = 0;
      if (polish) {
        // the active set of the polish iteration, chosen by kb_pdas_active
        a = gp(P.ik)[idx];
        if (a) {
          g = 0.0;
          nact += 1.0;
        }
      } else if (tactive && gp(P.op[0])[idx] < kActiveTol && g > 0.0) {
        a = 1;
        g = 0.0;
        nact += 1.0;
      }
      gp(P.ik)[idx] = a;
      nd = fma(g, g, nd);
      const double r = g * -1.0;
      const double zz = r / sq;
      gp(P.op[1])[idx] = r;
      gp(P.op[3])[idx] = zz;
      gp(P.op[4])[idx] = 0.0;
      b1 = fma(r, zz, b1);
      psum += zz;
      if (n == mr) own = zz;
      if (P.p32) p32_elem(P, idx, n, mr, lt, a != 0, zz, pb);
    }
    double evj = 0.0;
    if (!full) {
      psum = wave_sum(psum);
      own = wave_sum(own);
      evj = -(psum - own);
      if (lane_id() == 0) gp(P.ev)[j] = evj;
    }
    if (P.p32) p32_row_end(P, j, mr, lt, evj, pb);
  }
  double t = block_sum(nact, sh);
  if (threadIdx.x == 0) gp(P.part)[lb] = t;
  t = block_sum(nd, sh);
  if (threadIdx.x == 0) gp(P.part)[(int64_t)P.pcap + lb] = t;
  t = block_sum(b1, sh);
  if (threadIdx.x == 0) gp(P.part)[2 * (int64_t)P.pcap + lb] = t;
  if (threadIdx.x == 0 && lb == 0) gp(P.pn)[0] = gp(P.pn)[1] = gp(P.pn)[2] = nblk;
}

// One row j of the CG update, one wave: sn += a p; r += (-a) Hp; z = r / sqrt(diag);
// r.z and r.r accumulate into rz, rr (pqdts.f90:2060-2070). z is not stored: the
// direction update forms the same quotient from r.
// With `lf` (F_LATE in reduced space) Hp is folded here: 0 at the implicit column
// and on the active set, Hp - gm[j] elsewhere.
// The row's arrays are restrict-qualified parameters (they are distinct arrays), and
// each lane loads the operands of ROW_U elements before it stores any, so that the
// loads overlap instead of each waiting on the previous element's stores.
constexpr int ROW_U = 4;  // elements a lane loads before it computes and stores them
template <bool LF>
__device__ __forceinline__ void cg_update_span(GLOBAL double *__restrict__ sn, const GLOBAL double *__restrict__ p,
                                               const GLOBAL double *__restrict__ hp, GLOBAL double *__restrict__ r,
                                               const GLOBAL uint8_t *__restrict__ ik, int N, int mr, double gmj,
                                               double sq, double a, double &rz, double &rr) {
  for (int n0 = lane_id(); n0 < N; n0 += ROW_U * WAVE) {
    // the operands of ROW_U elements first (indices clamped into the row), so that
    // their loads overlap; then the elements in increasing n
    double sv[ROW_U], pv[ROW_U], hv[ROW_U], rv[ROW_U];
    uint8_t kv[ROW_U];
#pragma unroll
    for (int u = 0; u < ROW_U; u++) {
      const int n = n0 + u * WAVE, nc = n < N ? n : N - 1;
      sv[u] = sn[nc];
      pv[u] = p[nc];
      hv[u] = hp[nc];
      rv[u] = r[nc];
      kv[u] = LF ? ik[nc] : 0;
    }
#pragma unroll
    for (int u = 0; u < ROW_U; u++) {
      const int n = n0 + u * WAVE;
      if (n >= N) break;
      sn[n] = sv[u] + a * pv[u];
      double h = hv[u];
      if (LF) h = (n == mr || kv[u]) ? 0.0 : h - gmj;
      const double x = rv[u] + (-a) * h;
      r[n] = x;
      const double zz = x / sq;
      rz = fma(x, zz, rz);
      rr = fma(x, x, rr);
    }
  }
}
__device__ __forceinline__ void cg_update_row(const Prob &P, int j, bool full, bool lf, bool pol, double a,
                                              double &rz, double &rr) {
  const int N = P.N;
  const double sq = prec_div(P, j, full, pol);
  const int mr = lf ? gp(P.maxr)[j] : -1;
  const double gmj = lf ? gp(P.gm)[j] : 0.0;
  const int64_t base = (int64_t)j * N;
  if (lf)
    cg_update_span<true>(gp(P.op[4]) + base, gp(P.op[3]) + base, gp(P.op[2]) + base, gp(P.op[1]) + base,
                         gp(P.ik) + base, N, mr, gmj, sq, a, rz, rr);
  else
    cg_update_span<false>(gp(P.op[4]) + base, gp(P.op[3]) + base, gp(P.op[2]) + base, gp(P.op[1]) + base,
                          gp(P.ik) + base, N, mr, gmj, sq, a, rz, rr);
}

// One row j of the new direction, one wave: p = beta p + 1 z with z = r / sqrt(diag)
// (pqdts.f90:2075) and, in reduced space, ev[j] = -(sum p[j,:] - p[j,maxr]) for
// the next product; with C32 the FP32 copy of p outside the implicit column and,
// with LT (the late fold), the sum pb of p over the free entries, as p32_elem forms
// them. Restrict-qualified arrays and batched loads, as in cg_update_span.
template <bool C32, bool LT>
__device__ __forceinline__ void cg_dir_span(GLOBAL double *__restrict__ p, const GLOBAL double *__restrict__ r,
                                            const GLOBAL uint8_t *__restrict__ ik, GLOBAL float *__restrict__ p32,
                                            int N, int mr, double sq, double beta, double &psum, double &own,
                                            double &pb) {
  for (int n0 = lane_id(); n0 < N; n0 += ROW_U * WAVE) {
    double pv[ROW_U], rv[ROW_U];
    uint8_t kv[ROW_U];
#pragma unroll
    for (int u = 0; u < ROW_U; u++) {
      const int n = n0 + u * WAVE, nc = n < N ? n : N - 1;
      pv[u] = p[nc];
      rv[u] = r[nc];
      kv[u] = (C32 && LT) ? ik[nc] : 0;
    }
#pragma unroll
    for (int u = 0; u < ROW_U; u++) {
      const int n = n0 + u * WAVE;
      if (n >= N) break;
      const double v = beta * pv[u] + 1.0 * (rv[u] / sq);
      p[n] = v;
      psum += v;
      if (n == mr) own = v;
      if (C32 && n != mr) {
        p32[n] = (float)v;
        if (LT && !kv[u]) pb += v;
      }
    }
  }
}
__device__ __forceinline__ void cg_dir_row(const Prob &P, int j, bool full, bool late, bool pol, double beta) {
  const int N = P.N;
  const int mr = full ? -1 : gp(P.maxr)[j];
  const double sq = prec_div(P, j, full, pol);
  const bool lt = late && !full;
  double psum = 0.0, own = 0.0, pb = 0.0;
  const int64_t base = (int64_t)j * N;
  GLOBAL double *pr = gp(P.op[3]) + base;
  const GLOBAL double *rr = gp(P.op[1]) + base;
  const GLOBAL uint8_t *ikr = gp(P.ik) + base;
  if (!P.p32)
    cg_dir_span<false, false>(pr, rr, ikr, nullptr, N, mr, sq, beta, psum, own, pb);
  else if (lt)
    cg_dir_span<true, true>(pr, rr, ikr, gp(P.p32) + base, N, mr, sq, beta, psum, own, pb);
  else
    cg_dir_span<true, false>(pr, rr, ikr, gp(P.p32) + base, N, mr, sq, beta, psum, own, pb);
  double evj = 0.0;
  if (!full) {
    psum = wave_sum(psum);
    own = wave_sum(own);
    evj = -(psum - own);
    if (lane_id() == 0) gp(P.ev)[j] = evj;
  }
  if (P.p32) p32_row_end(P, j, mr, lt, evj, pb);
}

// CG update (pqdts.f90:2060-2070). Every block first sums the Hessian product's
// v2 = p.Hp partials (region i_reg_v2, stored by the band GEMM), then a = b1/v2;
// sn += a p; r += (-a) Hp; z = r / sqrt(diag); partials of r.z (region 0) and
// r.r (region 2) for kb_cg_dir.
__global__ __launch_bounds__(BLOCK) void kb_cg_update(const Prob *__restrict__ probs, WorkList wl, int i_b1,
                                                      int i_reg_v2, int i_v2) {
  __shared__ double sh[WAVES_PER_BLOCK];
  __shared__ double bc;
  ROW_PROLOGUE();
  const double v2 = reduce_parts_all(gp(P.part) + (int64_t)i_reg_v2 * P.pcap, gp(P.pn)[i_reg_v2], sh, &bc);
  if (threadIdx.x == 0 && lb == 0) gp(P.scal)[i_v2] = v2;
  // a polish face without free entries (or with a zero gradient on them) has
  // r = p = 0 and v2 = 0: no step instead of 0/0
  const double a = (polish && !(v2 > 0.0)) ? 0.0 : gp(P.scal)[i_b1] / v2;
  double rz = 0.0, rr = 0.0;
  if (j < P.M) cg_update_row(P, j, full, late && !full, polish, a, rz, rr);
  double t = block_sum(rz, sh);
  if (threadIdx.x == 0) gp(P.part)[lb] = t;
  t = block_sum(rr, sh);
  if (threadIdx.x == 0) gp(P.part)[2 * (int64_t)P.pcap + lb] = t;
  if (threadIdx.x == 0 && lb == 0) gp(P.pn)[0] = gp(P.pn)[2] = nblk;
}

// End of CG round j (pqdts.f90:2072-2089). Every block sums the r.z and r.r
// partials of kb_cg_update; block 0 stores them, logs (j, sqrt(r.r), eps, b1,
// r.z) and, once sqrt(r.r) <= eps, ends this problem's CG (scal[i_act] = 0,
// scal[i_n] = rounds done). Otherwise p = (r.z/b1) p + 1 z, and ev for the next
// BF. A problem whose CG ends keeps its p: nothing reads it before the next
// kb_cg_init (every block takes the same decision, whenever it starts).
__global__ __launch_bounds__(BLOCK) void kb_cg_dir(const Prob *__restrict__ probs, WorkList wl, int i_b1,
                                                   int i_rz, int i_rr, int j_round, int i_eps, int i_act,
                                                   int i_n) {
  __shared__ double sh[WAVES_PER_BLOCK];
  __shared__ double bc;
  ROW_PROLOGUE();
  const int np_ = gp(P.pn)[0];
  const double vrz = reduce_parts_all(gp(P.part), np_, sh, &bc);
  const double vrr = reduce_parts_all(gp(P.part) + 2 * (int64_t)P.pcap, np_, sh, &bc);
  GLOBAL double *sc = gp(P.scal);
  const double b2 = sqrt(vrr);
  const bool stop = b2 <= sc[i_eps];
  if (threadIdx.x == 0 && lb == 0) {
    if (j_round >= 1 && j_round <= CG_LOG_ROUNDS) {
      GLOBAL double *row = gp(P.cglog) + (int64_t)(j_round - 1) * 5;
      row[0] = (double)j_round;
      row[1] = b2;
      row[2] = sc[i_eps];
      row[3] = sc[i_b1];
      row[4] = vrz;
    }
    sc[i_rz] = vrz;
    sc[i_rr] = vrr;
    sc[i_n] = (double)j_round;
    if (stop) sc[i_act] = 0.0;
  }
  if (stop || j >= P.M) return;
  cg_dir_row(P, j, full, late, polish, vrz / sc[i_b1]);
}

// Start of PCG (pqdts.f90:2024-2031), one thread block per problem: sums of
// kb_cg_init's partials (nact, |G|^2, r.z) into scal[i_nact], scal[i_nd],
// scal[i_b1]; CG tolerance eps = min(0.5, sqrt(|G|)) |G| (a polish iteration:
// its relative residual, the launch argument, times |G|); CG marked running
// with no rounds done.
__global__ __launch_bounds__(BLOCK) void kb_cg_start(const Prob *__restrict__ probs, WorkList wl, int i_nact,
                                                     int i_nd, int i_b1, int i_eps, int i_act, int i_n) {
  __shared__ double sh[WAVES_PER_BLOCK];
  const Prob &P = probs[wl.pid[blockIdx.x]];
  const bool polish = (wl.flags[blockIdx.x] & F_POLISH) != 0;
  const double rtol = wl.a[blockIdx.x];
  const int n = gp(P.pn)[0];
  const double nact = reduce_parts(gp(P.part), n, false, sh);
  const double nd = reduce_parts(gp(P.part) + (int64_t)P.pcap, n, false, sh);
  const double b1 = reduce_parts(gp(P.part) + 2 * (int64_t)P.pcap, n, false, sh);
  if (threadIdx.x == 0) {
    GLOBAL double *sc = gp(P.scal);
    sc[i_nact] = nact;
    sc[i_nd] = nd;
    sc[i_b1] = b1;
    sc[i_eps] = polish ? rtol * sqrt(nd) : fmin(0.5, sqrt(sqrt(nd))) * sqrt(nd);
    sc[i_act] = 1.0;
    sc[i_n] = 0.0;
  }
}

// A CG that continues into the next batch step (a polish iteration's CG runs at
// most one step's rounds at a time): the round count restarts at 1, so the r.z of
// the last round done moves to the slot round 1 reads (i_b1). Thread 0 per problem.
__global__ void kb_cg_carry(const Prob *__restrict__ probs, WorkList wl, int i_n, int i_rz0, int i_rz1, int i_b1) {
  if (threadIdx.x != 0) return;
  GLOBAL double *sc = gp(probs[wl.pid[blockIdx.x]].scal);
  const int n = (int)sc[i_n];
  if (n >= 1) sc[i_b1] = sc[(n & 1) ? i_rz1 : i_rz0];
  sc[i_n] = 0.0;
}

// ---------------------------------------------------------------------------
// Polish: primal-dual active-set Newton on the reduced problem (no counterpart in
// the reference). With G the reduced gradient (the multiplier estimate of each
// entry: G = g - g[maxr], zero on the free entries at the optimum) and d the
// reduced Hessian diagonal of the row, an entry outside the implicit column is
// active (held at 0) when G - d X > 0. Such a fixed point with X >= 0 on the free
// entries satisfies the KKT conditions of the (strictly convex) problem.
// ---------------------------------------------------------------------------

// One wave per row: the new active set into ik, X = 0 on it; partials of the
// entries whose activity changed (region 0) and of the active entries (region 1).
__global__ __launch_bounds__(BLOCK) void kb_pdas_active(const Prob *__restrict__ probs, WorkList wl) {
  __shared__ double sh[WAVES_PER_BLOCK];
  ROW_PROLOGUE();
  double nchg = 0.0, nact = 0.0;
  if (j < P.M) {
    const double d = gp(P.diag_red)[j];
    const int mr = gp(P.maxr)[j];
    const int64_t base = (int64_t)j * N;
    GLOBAL double *xr = gp(P.op[0]) + base;
    const GLOBAL double *gr = gp(P.op[1]) + base;
    GLOBAL uint8_t *ikr = gp(P.ik) + base;
    for (int n = lane_id(); n < N; n += WAVE) {
      const double x = xr[n];
      const uint8_t a = (n != mr && gr[n] - d * x > 0.0) ? 1 : 0;
      if (a != ikr[n]) nchg += 1.0;
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

// One wave per row: the full Newton step X += SN outside the implicit column (SN
// is 0 on the active set); per block the partial -max |SN| (region), so that the
// minimum over the blocks is minus the step's largest entry.
__global__ __launch_bounds__(BLOCK) void kb_pdas_step(const Prob *__restrict__ probs, WorkList wl, int region) {
  __shared__ double sh[WAVES_PER_BLOCK];
  ROW_PROLOGUE();
  double m = 0.0;
  if (j < P.M) {
    const int mr = gp(P.maxr)[j];
    const int64_t base = (int64_t)j * N;
    GLOBAL double *xr = gp(P.op[0]) + base;
    const GLOBAL double *sr = gp(P.op[4]) + base;
    for (int n = lane_id(); n < N; n += WAVE) {
      if (n == mr) continue;
      const double v = sr[n];
      xr[n] = xr[n] + v;
      m = fmax(m, fabs(v));
    }
  }
  const double b = block_min(-m, sh);
  if (threadIdx.x == 0) gp(P.part)[(int64_t)region * P.pcap + lb] = b;
  if (threadIdx.x == 0 && lb == 0) gp(P.pn)[region] = nblk;
}

// Two-metric diagonal step (pqdts.f90:2095-2102): sn = -G/diag on active entries
// outside the implicit column.
__global__ __launch_bounds__(BLOCK) void kb_twometric(const Prob *__restrict__ probs, WorkList wl) {
  ROW_PROLOGUE();
  if (j >= P.M) return;
  const double dj = full ? gp(P.diag_full)[j] : gp(P.diag_red)[j];
  const int mr = full ? -1 : gp(P.maxr)[j];
  const int64_t base = (int64_t)j * N;
  for (int n = lane_id(); n < N; n += WAVE) {
    const int64_t idx = base + n;
    if (gp(P.ik)[idx] == 1 && mr != n) gp(P.op[4])[idx] = -gp(P.op[1])[idx] / dj;
  }
}

// FP32 copy of a band pack (mixed precision)
__global__ void k_round_f32(const double *__restrict__ src, float *__restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

}  // namespace pq
