// Host side of pqdts_amd: problem context, the C ABI declared in
// include/pqdts_amd.h, and the minimisation loop (pqdts.f90:1908-2167) driving
// the kernels of kernels.hip on one HIP stream.
#include "../../include/pqdts_amd.h"
#include "kernels.hip"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

using namespace pq;

namespace {

thread_local std::string g_err;

struct Fail : std::runtime_error {
  int code;
  Fail(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw Fail((int)e, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPC(x) hip_check((x), #x)

void require(bool ok, const std::string &msg) {
  if (!ok) throw Fail(-1, msg);
}

template <class F>
int guarded(F &&f) {
  try {
    g_err.clear();
    f();
    return 0;
  } catch (const Fail &e) {
    g_err = e.what();
    return e.code ? e.code : -1;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// tile heights of the two products
constexpr int TILE = 64;  // rows and columns of a band-GEMM block tile
constexpr int KC = 16;    // band chunk staged per LDS buffer (4 MFMA k-steps)

// device scalar slots
enum {
  S_LA = 0, S_LB, S_XMIN, S_NACT, S_ND, S_B1, S_V2, S_RZ0, S_RZ1, S_RR, S_CONV, S_DOT, S_NSCAL = 16
};

// partial-sum regions (each PCAP doubles)
enum { R_A = 0, R_B, R_C, R_D, R_NREG };

inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

struct Timer {
  int count = 0;
  double sec = 0.0;
  std::chrono::steady_clock::time_point t0;
  bool running = false;
};

}  // namespace

struct pqdts_ctx {
  int M = 0, N = 0, D = 0;
  int64_t MN = 0, DN = 0;
  int64_t nspan = 0;       // stored F elements: sum of row span lengths
  int64_t tile_elems = 0;  // packed elements of F's band tiles (the MFMA work of O = F V)
  double gamma = 0.0;
  int device = 0;
  hipStream_t s = nullptr;
  // F as dense row spans
  int *flo = nullptr, *flen = nullptr;
  int64_t *foff = nullptr;
  double *fval = nullptr;
  uint8_t *fhas = nullptr;
  // F (tiles of D rows) and F^T (tiles of M rows) packed as dense transposed bands
  struct Band {
    int ntiles = 0;
    double *pack = nullptr;
    int64_t *off = nullptr;
    int *klo = nullptr, *nch = nullptr;
  } bf, bt, bg;  // bg: the Gram matrix G = F^T F (M x M), built on demand
  int nrt = 0, nct = 0;
  // HVP path: the two products through O = F p (F path) or one product with G
  int hvp_pref = 0;           // 0 auto (cost model), 1 F path, 2 Gram path
  bool gram_auto = false;     // what the cost model chose at create
  int64_t gram_elems = 0;     // packed elements of G's band tiles (0 until known)
  std::vector<int> glo, gw;   // G's row bands on the host (start, width)
  bool use_gram() const { return hvp_pref == 2 || (hvp_pref == 0 && gram_auto); }
  int *kd_lo = nullptr, *kd_hi = nullptr;
  double *P = nullptr, *O = nullptr;
  double *slot[6] = {};
  double *diag_full = nullptr, *diag_red = nullptr;
  int *maxr = nullptr;
  uint8_t *ik = nullptr;
  double *ev = nullptr, *gm = nullptr, *lam = nullptr, *lmin = nullptr;
  double *part = nullptr;
  int pcap = 0;
  double *scal = nullptr, *hscal = nullptr;
  bool full = true;
  bool maxr_valid = false;  // every maxr[i] in [0, N)
  int mode = 2;
  bool sync_timers = false;
  std::map<std::string, Timer> timers;
  std::vector<std::string> timer_order;
  pqdts_event_cb cb = nullptr;
  void *user = nullptr;
  // per-kernel HIP-event timing (pqdts_kernel_timing): pending start/stop pairs
  // on the context's stream, folded into totals at every host synchronisation
  bool ktiming = false;
  struct KPair {
    hipEvent_t a, b;
    int kid;
  };
  std::vector<KPair> kpending;
  std::vector<hipEvent_t> kpool;
  std::vector<std::string> knames;
  std::vector<int> kcount;
  std::vector<double> kms;

  int kernel_id(const char *name) {
    for (size_t i = 0; i < knames.size(); i++)
      if (knames[i] == name) return (int)i;
    knames.push_back(name);
    kcount.push_back(0);
    kms.push_back(0.0);
    return (int)knames.size() - 1;
  }
  hipEvent_t kevent() {
    if (!kpool.empty()) {
      hipEvent_t e = kpool.back();
      kpool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPC(hipEventCreate(&e));
    return e;
  }
  // returns the start event recorded before a launch (or nullptr)
  hipEvent_t kbegin() {
    if (!ktiming) return nullptr;
    hipEvent_t a = kevent();
    HIPC(hipEventRecord(a, s));
    return a;
  }
  void kend(hipEvent_t a, const char *name) {
    if (!a) return;
    hipEvent_t b = kevent();
    HIPC(hipEventRecord(b, s));
    kpending.push_back({a, b, kernel_id(name)});
  }
  // called after the stream has been synchronised
  void kfold() {
    for (auto &p : kpending) {
      float ms = 0.f;
      HIPC(hipEventElapsedTime(&ms, p.a, p.b));
      kcount[p.kid]++;
      kms[p.kid] += ms;
      kpool.push_back(p.a);
      kpool.push_back(p.b);
    }
    kpending.clear();
  }

  double *reg(int r) { return part + (int64_t)r * pcap; }
  const double *diag() const { return full ? diag_full : diag_red; }
  int row_blocks() const { return cdiv(M, WAVES_PER_BLOCK); }

  void tstart(const char *name) {
    auto it = timers.find(name);
    if (it == timers.end()) {
      it = timers.emplace(name, Timer{}).first;
      timer_order.push_back(name);
    }
    it->second.count++;
    it->second.running = true;
    it->second.t0 = std::chrono::steady_clock::now();
  }
  void tstop(const char *name) {
    if (sync_timers) (void)hipStreamSynchronize(s);
    auto &t = timers[name];
    t.sec += std::chrono::duration<double>(std::chrono::steady_clock::now() - t.t0).count();
    t.running = false;
  }
  void emit(int kind, std::initializer_list<double> v) {
    if (cb) {
      std::vector<double> tmp(v);
      cb(user, kind, tmp.data(), (int)tmp.size());
    }
  }
  // copy all device scalars to the host and wait for the stream
  void fetch() {
    HIPC(hipMemcpyAsync(hscal, scal, sizeof(double) * S_NSCAL, hipMemcpyDeviceToHost, s));
    HIPC(hipStreamSynchronize(s));
    if (!kpending.empty()) kfold();
  }
  void reduce(int region, int n, int dst, bool is_min = false) {
    k_reduce<<<1, BLOCK, 0, s>>>(reg(region), n, scal, dst, is_min ? 1 : 0);
  }
};

namespace {

struct Scope {
  pqdts_ctx *c;
  const char *n;
  Scope(pqdts_ctx *c_, const char *n_) : c(c_), n(n_) { c->tstart(n); }
  ~Scope() { c->tstop(n); }
};

double *slot_ptr(pqdts_ctx *c, int slot) {
  require(slot >= 0 && slot < 6, "invalid slot id " + std::to_string(slot));
  return c->slot[slot];
}

template <class T>
T *dalloc(int64_t n) {
  void *p = nullptr;
  HIPC(hipMalloc(&p, sizeof(T) * (size_t)std::max<int64_t>(n, 1)));
  return (T *)p;
}

// ---------------------------------------------------------------------------
// operators
// ---------------------------------------------------------------------------

// Reduced-space kernels index rows by maxr; refuse to launch them otherwise.
void need_maxr(pqdts_ctx *c) {
  require(c->maxr_valid, "reduced space needs a valid maxr (call pqdts_transform or pqdts_set_maxr)");
}

void op_reconstruct(pqdts_ctx *c, double *X) {
  need_maxr(c);
  k_reconstruct<<<c->row_blocks(), BLOCK, 0, c->s>>>(X, c->M, c->N, c->maxr);
}

BandArgs band_args(const pqdts_ctx::Band &b) {
  BandArgs a{};
  a.apack = b.pack;
  a.aoff = b.off;
  a.klo = b.klo;
  a.kchunks = b.nch;
  return a;
}

// O = F V (redO, pqdts.f90:724-736); returns the number of LOSS partials in R_A
int op_spmm(pqdts_ctx *c, const double *V, bool expand, bool loss, bool store) {
  BandArgs a = band_args(c->bf);
  a.nrows = c->D;
  a.nk = c->M;
  a.N = c->N;
  a.B = V;
  a.maxr = c->maxr;
  a.ev = c->ev;
  a.C = c->O;
  a.P = c->P;
  a.has = c->fhas;
  a.partial = c->reg(R_A);
  dim3 grid(cdiv(c->N, TILE), c->bf.ntiles);
  hipEvent_t ev0 = c->kbegin();
  if (expand)
    k_band_gemm<TILE, TILE, KC, B_EXPAND, E_STORE, false, false, false><<<grid, BLOCK, 0, c->s>>>(a);
  else if (loss && store)
    k_band_gemm<TILE, TILE, KC, B_PLAIN, E_STORE_LOSS, false, false, false><<<grid, BLOCK, 0, c->s>>>(a);
  else if (loss)
    k_band_gemm<TILE, TILE, KC, B_PLAIN, E_LOSS, false, false, false><<<grid, BLOCK, 0, c->s>>>(a);
  else
    k_band_gemm<TILE, TILE, KC, B_PLAIN, E_STORE, false, false, false><<<grid, BLOCK, 0, c->s>>>(a);
  HIPC(hipGetLastError());
  c->kend(ev0, "spmm_F");
  return grid.x * grid.y;
}

// G = 2 F^T W + stencil (+ fold / mask / dot) (dL 998-1030, BF 276-340);
// returns the number of DOT partials in R_B
int op_spmmT(pqdts_ctx *c, int mode, const double *E, const double *dd, double *G, bool mask, bool dot) {
  const bool red = !c->full;
  if (red) need_maxr(c);
  if (red) {
    if (mode == T_GRAD)
      k_foldval<T_GRAD><<<c->row_blocks(), BLOCK, 0, c->s>>>(c->O, c->P, c->N, c->M, c->flo, c->flen, c->foff,
                                                            c->fval, c->kd_lo, c->kd_hi, c->gamma, E, c->maxr,
                                                            c->ev, c->gm);
    else
      k_foldval<T_HVP><<<c->row_blocks(), BLOCK, 0, c->s>>>(c->O, c->P, c->N, c->M, c->flo, c->flen, c->foff,
                                                           c->fval, c->kd_lo, c->kd_hi, c->gamma, E, c->maxr,
                                                           c->ev, c->gm);
  }
  BandArgs a = band_args(c->bt);
  a.nrows = c->M;
  a.nk = c->D;
  a.N = c->N;
  a.B = c->O;
  a.B2 = c->P;
  a.maxr = c->maxr;
  a.ev = c->ev;
  a.C = G;
  a.gamma = c->gamma;
  a.E = E;
  a.gm = c->gm;
  a.ik = c->ik;
  a.dd = dd;
  a.partial = c->reg(R_B);
  dim3 grid(cdiv(c->N, TILE), c->bt.ntiles);
  hipEvent_t ev0 = c->kbegin();
#define BG(BM, EM, R, MK, DT) k_band_gemm<TILE, TILE, KC, BM, EM, R, MK, DT><<<grid, BLOCK, 0, c->s>>>(a)
  if (mode == T_GRAD) {
    if (red)
      BG(B_DIFF, E_GRAD, true, false, false);
    else
      BG(B_DIFF, E_GRAD, false, false, false);
  } else {
    if (red) {
      if (mask && dot)
        BG(B_PLAIN, E_HVP, true, true, true);
      else
        BG(B_PLAIN, E_HVP, true, false, false);
    } else {
      if (mask && dot)
        BG(B_PLAIN, E_HVP, false, true, true);
      else
        BG(B_PLAIN, E_HVP, false, false, false);
    }
  }
#undef BG
  HIPC(hipGetLastError());
  c->kend(ev0, "spmmT_F");
  return grid.x * grid.y;
}

// objective (pqdts.f90:917-973); `recon` = apply the reduced-space prep
double op_L(pqdts_ctx *c, double *X, bool recon = true, bool fetch = true) {
  Scope t(c, "L");
  if (!c->full && recon) op_reconstruct(c, X);
  const int na = op_spmm(c, X, false, true, false);
  c->reduce(R_A, na, S_LA);
  if (c->gamma != 0.0) {
    const int nb = std::min(cdiv((int64_t)(c->M - 1) * c->N, BLOCK), c->pcap);
    if (nb > 0) {
      k_smooth<<<nb, BLOCK, 0, c->s>>>(X, c->M, c->N, c->gamma, c->reg(R_B));
      c->reduce(R_B, nb, S_LB);
    } else {
      HIPC(hipMemsetAsync(c->scal + S_LB, 0, sizeof(double), c->s));
    }
  } else {
    HIPC(hipMemsetAsync(c->scal + S_LB, 0, sizeof(double), c->s));
  }
  if (!fetch) return 0.0;
  c->fetch();
  return c->hscal[S_LA] + c->hscal[S_LB];
}

// gradient (pqdts.f90:975-1032)
void op_dL(pqdts_ctx *c, double *X, double *G) {
  Scope t(c, "dL");
  if (!c->full) op_reconstruct(c, X);
  op_spmm(c, X, false, false, true);
  op_spmmT(c, T_GRAD, X, nullptr, G, false, false);
}

// Hessian times dd (pqdts.f90:246-342). `have_ev`: ev already holds the
// expansion of dd (computed by the CG kernels).
void ensure_gram(pqdts_ctx *c);

// Gram path of BF: BX = 2 G p~ + stencil (+ fold / mask / dot) with G = F^T F,
// mathematically the same operator as 2 F^T (F p~) + stencil; no O is formed.
int op_BF_gram(pqdts_ctx *c, const double *dd, double *BX, bool mask, bool dot) {
  const bool red = !c->full;
  if (red)
    k_foldval_gram<<<c->row_blocks(), BLOCK, 0, c->s>>>(c->bg.pack, c->bg.off, c->bg.klo, c->bg.nch, KC, TILE,
                                                         c->N, c->M, c->gamma, dd, c->maxr, c->ev, c->gm);
  BandArgs a = band_args(c->bg);
  a.nrows = c->M;
  a.nk = c->M;
  a.N = c->N;
  a.B = dd;
  a.maxr = c->maxr;
  a.ev = c->ev;
  a.C = BX;
  a.gamma = c->gamma;
  a.E = dd;
  a.gm = c->gm;
  a.ik = c->ik;
  a.dd = dd;
  a.partial = c->reg(R_B);
  dim3 grid(cdiv(c->N, TILE), c->bg.ntiles);
  hipEvent_t ev0 = c->kbegin();
#define BG(BM, R, MK, DT) k_band_gemm<TILE, TILE, KC, BM, E_HVP, R, MK, DT><<<grid, BLOCK, 0, c->s>>>(a)
  if (red) {
    if (mask && dot)
      BG(B_EXPAND, true, true, true);
    else
      BG(B_EXPAND, true, false, false);
  } else {
    if (mask && dot)
      BG(B_PLAIN, false, true, true);
    else
      BG(B_PLAIN, false, false, false);
  }
#undef BG
  HIPC(hipGetLastError());
  c->kend(ev0, "hvp_G");
  return grid.x * grid.y;
}

int op_BF(pqdts_ctx *c, const double *dd, double *BX, bool have_ev, bool mask, bool dot) {
  Scope t(c, "BF");
  if (!c->full) need_maxr(c);
  if (!c->full && !have_ev)
    k_others<<<c->row_blocks(), BLOCK, 0, c->s>>>(dd, c->M, c->N, c->maxr, c->ev);
  if (c->use_gram()) {
    ensure_gram(c);
    return op_BF_gram(c, dd, BX, mask, dot);
  }
  op_spmm(c, dd, !c->full, false, true);
  return op_spmmT(c, T_HVP, dd, dd, BX, mask, dot);
}

double op_pdot(pqdts_ctx *c, const double *x, const double *y) {
  Scope t(c, "pdot");
  const int nb = std::min(cdiv(c->MN, BLOCK), std::min(c->pcap, 1024));
  k_dot<<<nb, BLOCK, 0, c->s>>>(x, y, c->MN, c->reg(R_A));
  c->reduce(R_A, nb, S_DOT);
  c->fetch();
  return c->hscal[S_DOT];
}

void op_doproj(pqdts_ctx *c, double *X) {
  Scope t(c, "doproj");
  k_simplex_rows<<<c->row_blocks(), BLOCK, 0, c->s>>>(X, nullptr, nullptr, 0.0, c->M, c->N, 0);
  HIPC(hipGetLastError());
}

double op_doproj2(pqdts_ctx *c, double *X, bool fetch = true) {
  Scope t(c, "doproj2");
  require(!c->full, "doproj2 needs the reduced space (maxr set)");
  need_maxr(c);
  const int nb = c->row_blocks();
  k_clip_rows<<<nb, BLOCK, 0, c->s>>>(X, nullptr, nullptr, 0.0, c->M, c->N, c->maxr, 0, 0, c->reg(R_C));
  c->reduce(R_C, nb, S_XMIN, true);
  if (!fetch) return 0.0;
  c->fetch();
  return c->hscal[S_XMIN];
}

// conv_test (pqdts.f90:1034-1067)
double op_conv(pqdts_ctx *c, double *X, double *G, bool print) {
  if (!c->full) op_reconstruct(c, X);
  c->full = true;
  op_dL(c, X, G);
  const int nb = c->row_blocks();
  k_conv_rows<<<nb, BLOCK, 0, c->s>>>(X, G, c->M, c->N, print ? c->lam : nullptr, print ? c->lmin : nullptr,
                                      c->reg(R_C));
  c->reduce(R_C, nb, S_CONV);
  c->fetch();
  if (print && c->cb) {
    std::vector<double> lam(c->M), lmin(c->M);
    HIPC(hipMemcpy(lam.data(), c->lam, sizeof(double) * c->M, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(lmin.data(), c->lmin, sizeof(double) * c->M, hipMemcpyDeviceToHost));
    for (int i = 0; i < c->M; i++) c->emit(PQDTS_EV_LAMBDA, {(double)(i + 1), lam[i], lmin[i]});
  }
  return std::sqrt(c->hscal[S_CONV] / ((double)c->N * (double)c->M));
}

// Objective at the trial point X + a*SN after projection (both line searches).
// Returns L; *xmin receives the doproj2 minimum (0 for Condat).
double trial_L(pqdts_ctx *c, double a, bool condat, double *xmin) {
  double *T = c->slot[PQDTS_XTMP];
  const int nb = c->row_blocks();
  if (condat) {
    Scope t(c, "doproj");
    k_simplex_rows<<<nb, BLOCK, 0, c->s>>>(T, c->slot[PQDTS_X], c->slot[PQDTS_SN], a, c->M, c->N, 1);
  } else {
    Scope t(c, "doproj2");
    k_clip_rows<<<nb, BLOCK, 0, c->s>>>(T, c->slot[PQDTS_X], c->slot[PQDTS_SN], a, c->M, c->N, c->maxr, 1, 1,
                                        c->reg(R_C));
    c->reduce(R_C, nb, S_XMIN, true);
  }
  HIPC(hipGetLastError());
  // the reduced-space prep of L was applied inside k_clip_rows
  double L = op_L(c, T, false, true);
  *xmin = condat ? 0.0 : c->hscal[S_XMIN];
  return L;
}

// Lalpha_min2 (pqdts.f90:394-438) and Lalpha_min_auglag (pqdts.f90:440-483)
void op_linesearch(pqdts_ctx *c, bool condat, double *alpha_out, double *dlalpha) {
  Scope t(c, "Lalpha_min");
  if (!condat) {
    require(!c->full, "Lalpha_min2 needs the reduced space (maxr set)");
    need_maxr(c);
  }
  const double dx = (double)1e-5f;        // pqdts.f90:404
  const double feas = (double)-1e-12f;    // pqdts.f90:424
  double xmin;
  const double L0 = trial_L(c, 0.0, condat, &xmin);
  double Lp = trial_L(c, 0.0 + dx, condat, &xmin);
  const double df = (Lp - L0) / dx;
  *dlalpha = df;
  double alpha = 1.0, lm = std::numeric_limits<double>::max(), alphamin = 0.0;
  for (int i = 1; i <= 100; i++) {
    Lp = trial_L(c, alpha, condat, &xmin);
    const bool feasible = condat || xmin >= feas;
    if (Lp < lm && feasible) {
      lm = Lp;
      alphamin = alpha;
    }
    c->emit(PQDTS_EV_LS, {(double)i, Lp, alpha, df, xmin, condat ? 0.0 : 1.0});
    if (Lp > L0 + 0.1 * alpha * df || !feasible)
      alpha = alpha * 0.75;
    else
      break;
  }
  *alpha_out = alphamin;
}

void op_axpby(pqdts_ctx *c, double *z, const double *x, const double *y, double alpha, double beta, int kind) {
  const int nb = std::max(1, std::min(cdiv(c->MN, BLOCK), 4096));
  k_axpby<<<nb, BLOCK, 0, c->s>>>(z, x, y, alpha, beta, c->MN, kind);
  HIPC(hipGetLastError());
}

void op_transform(pqdts_ctx *c, bool rebuild) {
  k_transform<<<c->row_blocks(), BLOCK, 0, c->s>>>(c->slot[PQDTS_X], c->M, c->N, c->maxr, rebuild ? 1 : 0);
  HIPC(hipGetLastError());
  c->maxr_valid = true;
}

void fill_maxr(pqdts_ctx *c, int v) {
  std::vector<int> h(c->M, v);
  HIPC(hipMemcpyAsync(c->maxr, h.data(), sizeof(int) * c->M, hipMemcpyHostToDevice, c->s));
  HIPC(hipStreamSynchronize(c->s));
  c->maxr_valid = v >= 0 && v < c->N;
}

// ---------------------------------------------------------------------------
// The minimisation loop (pqdts.f90:1908-2167)
// ---------------------------------------------------------------------------
double solve(pqdts_ctx *c, int mode, int maxiter, double conv_thres, int start_stage, int output) {
  Scope tmin(c, "min");
  double *X = c->slot[PQDTS_X], *G = c->slot[PQDTS_G], *HP = c->slot[PQDTS_HP];
  double *XT = c->slot[PQDTS_XTMP], *SN = c->slot[PQDTS_SN], *Z = c->slot[PQDTS_Z];
  const int M = c->M, N = c->N;
  const int64_t MN = c->MN;
  c->mode = mode;
  double conv = 0.0, alpha = 0.0, dlalpha, norm_der, L0, xmin;

  L0 = op_L(c, X);
  op_dL(c, X, G);
  norm_der = op_pdot(c, G, G);
  c->emit(PQDTS_EV_ITER0, {L0, std::sqrt(norm_der)});

  for (int oiter = start_stage; oiter <= 2; oiter++) {
    dlalpha = -std::numeric_limits<double>::max();
    for (int iiter = 1; iiter <= maxiter; iiter++) {
      if (mode == 1) {
        // projected gradient, pqdts.f90:1934-1951
        c->full = false;
        op_transform(c, false);
        op_dL(c, X, G);
        op_axpby(c, SN, G, nullptr, -1.0, 0.0, 2);
        op_linesearch(c, false, &alpha, &dlalpha);
      } else {
        if ((std::fabs(dlalpha) < (double)1e-4f || dlalpha > 0) && oiter == 1) {
          c->emit(PQDTS_EV_STOP1, {dlalpha});
          break;
        }
        c->full = (oiter == 1);
        if (!c->full)
          op_transform(c, true);
        else
          fill_maxr(c, -1);
        op_dL(c, X, G);
        const bool red = !c->full;
        if (red) need_maxr(c);
        const int nb = c->row_blocks();
        {
          Scope t(c, "active_dLdX");
          k_cg_init<<<nb, BLOCK, 0, c->s>>>(X, G, c->ik, Z, XT, SN, c->diag(), M, N, c->maxr, c->ev, red ? 1 : 0,
                                            1, c->reg(R_A), c->reg(R_B), c->reg(R_C));
          HIPC(hipGetLastError());
          c->reduce(R_A, nb, S_NACT);
          c->reduce(R_B, nb, S_ND);
          c->reduce(R_C, nb, S_B1);
          c->fetch();
        }
        c->emit(PQDTS_EV_ACTIVE, {c->hscal[S_NACT], (double)MN});
        norm_der = c->hscal[S_ND];
        const double eps = std::fmin(0.5, std::sqrt(std::sqrt(norm_der))) * std::sqrt(norm_der);
        int i_b1 = S_B1;
        for (int j = 1; j <= 50; j++) {
          const int ndot = op_BF(c, XT, HP, true, true, true);
          Scope t(c, "CG");
          c->reduce(R_B, ndot, S_V2);
          const int i_rz = (j & 1) ? S_RZ1 : S_RZ0;
          k_cg_update<<<nb, BLOCK, 0, c->s>>>(SN, XT, G, HP, Z, c->diag(), M, N, c->scal, i_b1, S_V2, c->reg(R_A),
                                              c->reg(R_C));
          HIPC(hipGetLastError());
          c->reduce(R_A, nb, i_rz);
          c->reduce(R_C, nb, S_RR);
          k_cg_dir<<<nb, BLOCK, 0, c->s>>>(XT, Z, M, N, c->scal, i_b1, i_rz, c->maxr, c->ev, red ? 1 : 0);
          HIPC(hipGetLastError());
          c->fetch();
          const double b1 = c->hscal[i_b1], v2 = c->hscal[i_rz], b2 = c->hscal[S_RR];
          c->emit(PQDTS_EV_PCG, {(double)j, std::sqrt(b2), eps, b1, v2});
          if (std::sqrt(b2) <= eps) break;
          i_b1 = i_rz;
        }
        // two-metric diagonal step, pqdts.f90:2092-2104
        op_dL(c, X, G);
        k_twometric<<<nb, BLOCK, 0, c->s>>>(SN, G, c->ik, c->diag(), M, N, c->maxr);
        HIPC(hipGetLastError());
        op_linesearch(c, c->full, &alpha, &dlalpha);
      }
      // step, projection and re-evaluation, pqdts.f90:2111-2128
      if (alpha != 0.0) op_axpby(c, X, SN, X, alpha, 0.0, 1);
      xmin = 0.0;
      if (c->full)
        op_doproj(c, X);
      else
        xmin = op_doproj2(c, X);
      L0 = op_L(c, X);
      op_dL(c, X, G);
      norm_der = op_pdot(c, G, G);
      conv = op_conv(c, X, G, false);
      c->emit(PQDTS_EV_ITER,
              {(double)oiter, (double)iiter, L0, xmin, conv, std::sqrt(norm_der), (double)mode, alpha, dlalpha});
      if (conv < conv_thres) break;
    }
    conv = op_conv(c, X, G, output >= 3);
    c->emit(PQDTS_EV_STAGE, {(double)oiter, conv});
  }
  HIPC(hipStreamSynchronize(c->s));
  c->kfold();
  return conv;
}

void destroy_ctx(pqdts_ctx *c) {
  if (!c) return;
  if (c->s) (void)hipStreamSynchronize(c->s);
  for (auto &p : c->kpending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (hipEvent_t e : c->kpool) (void)hipEventDestroy(e);
  void *ptrs[] = {c->flo,      c->flen,      c->foff,   c->fval,   c->fhas,   c->bf.pack, c->bf.off,
                  c->bf.klo,   c->bf.nch,    c->bt.pack, c->bt.off, c->bt.klo, c->bt.nch,  c->bg.pack,
                  c->bg.off,   c->bg.klo,    c->bg.nch,  c->kd_lo,
                  c->kd_hi,    c->P,         c->O,      c->diag_full, c->diag_red, c->maxr, c->ik,
                  c->ev,       c->gm,        c->lam,    c->lmin,   c->part,   c->scal};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (double *p : c->slot)
    if (p) (void)hipFree(p);
  if (c->hscal) (void)hipHostFree(c->hscal);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
}

// Packs the rows of a banded matrix into per-tile dense transposed bands for
// k_band_gemm: tile t covers rows [t*TILE, (t+1)*TILE); its band starts at
// column klo[t] and holds nch[t]*KC zero-filled columns; element (row r, col k)
// of the tile sits at off[t] + (k - klo[t])*TILE + (r - t*TILE).
struct HostBand {
  std::vector<double> pack;
  std::vector<int64_t> off;
  std::vector<int> klo, nch;
};

template <class Span, class Val>
HostBand pack_band(int R, Span span, Val val) {
  const int nt = cdiv(R, TILE);
  HostBand b;
  b.off.assign(nt, 0);
  b.klo.assign(nt, 0);
  b.nch.assign(nt, 0);
  int64_t total = 0;
  for (int t = 0; t < nt; t++) {
    int lo = INT32_MAX, hi = -1;
    for (int r = t * TILE; r < std::min(R, (t + 1) * TILE); r++) {
      const auto sp = span(r);
      if (sp.second > 0) {
        lo = std::min(lo, sp.first);
        hi = std::max(hi, sp.first + sp.second - 1);
      }
    }
    b.off[t] = total;
    if (hi >= 0) {
      b.klo[t] = lo;
      b.nch[t] = cdiv(hi - lo + 1, KC);
      total += (int64_t)b.nch[t] * KC * TILE;
    }
  }
  b.pack.assign((size_t)std::max<int64_t>(total, 1), 0.0);
  for (int t = 0; t < nt; t++)
    for (int r = t * TILE; r < std::min(R, (t + 1) * TILE); r++) {
      const auto sp = span(r);
      for (int k = sp.first; k < sp.first + sp.second; k++)
        b.pack[b.off[t] + (int64_t)(k - b.klo[t]) * TILE + (r - t * TILE)] = val(r, k);
    }
  return b;
}

void upload_band(pqdts_ctx *c, pqdts_ctx::Band &dst, const HostBand &h) {
  dst.ntiles = (int)h.off.size();
  dst.pack = dalloc<double>((int64_t)h.pack.size());
  dst.off = dalloc<int64_t>(dst.ntiles);
  dst.klo = dalloc<int>(dst.ntiles);
  dst.nch = dalloc<int>(dst.ntiles);
  HIPC(hipMemcpyAsync(dst.pack, h.pack.data(), sizeof(double) * h.pack.size(), hipMemcpyHostToDevice, c->s));
  HIPC(hipMemcpyAsync(dst.off, h.off.data(), sizeof(int64_t) * dst.ntiles, hipMemcpyHostToDevice, c->s));
  HIPC(hipMemcpyAsync(dst.klo, h.klo.data(), sizeof(int) * dst.ntiles, hipMemcpyHostToDevice, c->s));
  HIPC(hipMemcpyAsync(dst.nch, h.nch.data(), sizeof(int) * dst.ntiles, hipMemcpyHostToDevice, c->s));
}

// Packed elements of the band tiles of a banded matrix (the MFMA work per output column).
template <class Span>
int64_t band_elems(int R, Span span) {
  int64_t total = 0;
  for (int t = 0; t < cdiv(R, TILE); t++) {
    int lo = INT32_MAX, hi = -1;
    for (int r = t * TILE; r < std::min(R, (t + 1) * TILE); r++) {
      const auto sp = span(r);
      if (sp.second > 0) {
        lo = std::min(lo, sp.first);
        hi = std::max(hi, sp.first + sp.second - 1);
      }
    }
    if (hi >= 0) total += (int64_t)cdiv(hi - lo + 1, KC) * KC * TILE;
  }
  return total;
}

// Forms G = F^T F once, on the device: F is expanded to a dense D x M matrix and
// multiplied by the packed F^T band with the band GEMM itself; the rows of G are
// then packed into band tiles like F.
void ensure_gram(pqdts_ctx *c) {
  if (c->bg.pack) return;
  const int M = c->M, D = c->D;
  double *Fd = nullptr, *Gd = nullptr;
  try {
    Fd = dalloc<double>((int64_t)D * M);
    Gd = dalloc<double>((int64_t)M * M);
    k_dense_rows<<<cdiv(D, WAVES_PER_BLOCK), BLOCK, 0, c->s>>>(c->flo, c->flen, c->foff, c->fval, D, M, Fd);
    BandArgs a = band_args(c->bt);
    a.nrows = M;
    a.nk = D;
    a.N = M;
    a.B = Fd;
    a.C = Gd;
    dim3 grid(cdiv(M, TILE), c->bt.ntiles);
    k_band_gemm<TILE, TILE, KC, B_PLAIN, E_STORE, false, false, false><<<grid, BLOCK, 0, c->s>>>(a);
    HIPC(hipGetLastError());
    std::vector<double> G((size_t)M * M);
    HIPC(hipMemcpyAsync(G.data(), Gd, sizeof(double) * G.size(), hipMemcpyDeviceToHost, c->s));
    HIPC(hipStreamSynchronize(c->s));
    HostBand h = pack_band(
        M, [&](int k) { return std::make_pair(c->glo[k], c->gw[k]); },
        [&](int k, int l) { return G[(size_t)k * M + l]; });
    upload_band(c, c->bg, h);
    HIPC(hipStreamSynchronize(c->s));
  } catch (...) {
    if (Fd) (void)hipFree(Fd);
    if (Gd) (void)hipFree(Gd);
    throw;
  }
  (void)hipFree(Fd);
  (void)hipFree(Gd);
}

pqdts_ctx *create_ctx(int64_t M, int64_t N, int64_t D, int64_t nF, const int32_t *Fr, const int32_t *Fc,
                      const double *Fv, int64_t nP, const int32_t *Pr, const int32_t *Pc, const double *Pv,
                      double gamma) {
  require(M >= 1 && N >= 1 && D >= 1, "need M >= 1, N >= 1, D >= 1");
  require(M * N < (int64_t)INT32_MAX * 8 && M < INT32_MAX && N < INT32_MAX && D < INT32_MAX, "problem too large");
  require(nF >= 0 && nP >= 0, "negative nnz");
  require(nF == 0 || (Fr && Fc && Fv), "null F arrays");
  require(nP == 0 || (Pr && Pc && Pv), "null P arrays");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Fail(-2, "no GPU visible to pqdts_amd");

  // F ingest: O(nnz) span build with zero fill and last-wins duplicates
  // (same result as the O(D*nnz) scan of pqdts.f90:1328-1364)
  std::vector<int> lo(D, INT32_MAX), hi(D, -1), len(D, 0);
  std::vector<uint8_t> has(D, 0);
  for (int64_t e = 0; e < nF; e++) {
    const int d = Fr[e], k = Fc[e];
    require(d >= 0 && d < D && k >= 0 && k < M, "F index out of range at entry " + std::to_string(e));
    lo[d] = std::min(lo[d], k);
    hi[d] = std::max(hi[d], k);
    has[d] = 1;
  }
  std::vector<int64_t> off(D + 1, 0);
  for (int d = 0; d < D; d++) {
    if (!has[d]) lo[d] = 0;
    len[d] = has[d] ? hi[d] - lo[d] + 1 : 0;
    off[d + 1] = off[d] + len[d];
  }
  std::vector<double> val(off[D], 0.0);
  for (int64_t e = 0; e < nF; e++) val[off[Fr[e]] + (Fc[e] - lo[Fr[e]])] = Fv[e];
  // per-k d-range (column spans of F)
  std::vector<int> kdlo(M, INT32_MAX), kdhi(M, -1);
  for (int d = 0; d < D; d++)
    for (int k = lo[d]; k < lo[d] + len[d]; k++) {
      kdlo[k] = std::min(kdlo[k], d);
      kdhi[k] = std::max(kdhi[k], d);
    }
  for (int k = 0; k < M; k++)
    if (kdhi[k] < 0) kdlo[k] = 0;
  // F and F^T packed for the band GEMM
  HostBand hbf = pack_band(
      (int)D, [&](int d) { return std::make_pair(lo[d], len[d]); },
      [&](int d, int k) { return val[off[d] + (k - lo[d])]; });
  HostBand hbt = pack_band(
      (int)M, [&](int k) { return std::make_pair(kdlo[k], kdhi[k] >= 0 ? kdhi[k] - kdlo[k] + 1 : 0); },
      [&](int k, int d) {
        const int rel = k - lo[d];
        return (rel >= 0 && rel < len[d]) ? val[off[d] + rel] : 0.0;
      });
  // P densify (pqdts.f90:1527-1535), last duplicate wins
  std::vector<double> Ph((size_t)(D * N), 0.0);
  for (int64_t e = 0; e < nP; e++) {
    require(Pr[e] >= 0 && Pr[e] < D && Pc[e] >= 0 && Pc[e] < N, "P index out of range at entry " + std::to_string(e));
    Ph[(size_t)Pr[e] * N + Pc[e]] = Pv[e];
  }

  pqdts_ctx *c = new pqdts_ctx();
  try {
    c->M = (int)M;
    c->N = (int)N;
    c->D = (int)D;
    c->MN = M * N;
    c->DN = D * N;
    c->gamma = gamma;
    c->sync_timers = getenv("PQDTS_SYNC_TIMERS") && atoi(getenv("PQDTS_SYNC_TIMERS")) != 0;
    HIPC(hipGetDevice(&c->device));
    HIPC(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    c->nspan = off[D];
    c->tile_elems = (int64_t)hbf.pack.size();
    c->nrt = (int)hbf.off.size();
    c->nct = (int)hbt.off.size();
    // G = F^T F is banded: row k reaches from the first to the last column of
    // every F row that covers k. The Gram HVP is chosen when its band tiles are
    // clearly cheaper than the two F products (and G fits in HBM many times over).
    c->glo.assign(M, INT32_MAX);
    c->gw.assign(M, 0);
    {
      std::vector<int> ghi(M, -1);
      for (int d = 0; d < D; d++)
        for (int k = lo[d]; k < lo[d] + len[d]; k++) {
          c->glo[k] = std::min(c->glo[k], lo[d]);
          ghi[k] = std::max(ghi[k], lo[d] + len[d] - 1);
        }
      for (int k = 0; k < M; k++) {
        if (ghi[k] < 0) c->glo[k] = 0;
        c->gw[k] = ghi[k] >= 0 ? ghi[k] - c->glo[k] + 1 : 0;
      }
    }
    c->gram_elems = band_elems((int)M, [&](int k) { return std::make_pair(c->glo[k], c->gw[k]); });
    const int64_t fpath = (int64_t)hbf.pack.size() + (int64_t)hbt.pack.size();
    c->gram_auto = (double)c->gram_elems < 0.8 * (double)fpath && M * M * 8 < ((int64_t)64 << 30);
    if (const char *e = getenv("PQDTS_HVP")) c->hvp_pref = (e[0] == 'G') ? 2 : (e[0] == 'F') ? 1 : 0;
    c->flo = dalloc<int>(D);
    c->flen = dalloc<int>(D);
    c->foff = dalloc<int64_t>(D);
    c->fval = dalloc<double>(off[D]);
    c->fhas = dalloc<uint8_t>(D);
    c->kd_lo = dalloc<int>(M);
    c->kd_hi = dalloc<int>(M);
    c->P = dalloc<double>(c->DN);
    c->O = dalloc<double>(c->DN);
    for (int s = 0; s < 6; s++) c->slot[s] = dalloc<double>(c->MN);
    c->diag_full = dalloc<double>(M);
    c->diag_red = dalloc<double>(M);
    c->maxr = dalloc<int>(M);
    c->ik = dalloc<uint8_t>(c->MN);
    c->ev = dalloc<double>(M);
    c->gm = dalloc<double>(M);
    c->lam = dalloc<double>(M);
    c->lmin = dalloc<double>(M);
    const int gx = cdiv(N, TILE);
    c->pcap = std::max({gx * c->nrt, gx * c->nct, cdiv(M, WAVES_PER_BLOCK), 4096});
    c->part = dalloc<double>((int64_t)c->pcap * R_NREG);
    c->scal = dalloc<double>(S_NSCAL);
    HIPC(hipHostMalloc((void **)&c->hscal, sizeof(double) * S_NSCAL, hipHostMallocDefault));
    auto up = [&](void *dst, const void *src, size_t bytes) {
      if (bytes) HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->s));
    };
    up(c->flo, lo.data(), sizeof(int) * D);
    up(c->flen, len.data(), sizeof(int) * D);
    up(c->foff, off.data(), sizeof(int64_t) * D);
    up(c->fval, val.data(), sizeof(double) * off[D]);
    up(c->fhas, has.data(), D);
    upload_band(c, c->bf, hbf);
    upload_band(c, c->bt, hbt);
    up(c->kd_lo, kdlo.data(), sizeof(int) * M);
    up(c->kd_hi, kdhi.data(), sizeof(int) * M);
    up(c->P, Ph.data(), sizeof(double) * c->DN);
    for (int s = 0; s < 6; s++) HIPC(hipMemsetAsync(c->slot[s], 0, sizeof(double) * c->MN, c->s));
    HIPC(hipMemsetAsync(c->O, 0, sizeof(double) * c->DN, c->s));
    HIPC(hipMemsetAsync(c->ik, 0, c->MN, c->s));
    HIPC(hipMemsetAsync(c->scal, 0, sizeof(double) * S_NSCAL, c->s));
    k_bdiag<<<c->row_blocks(), BLOCK, 0, c->s>>>(c->M, c->flo, c->flen, c->foff, c->fval, c->kd_lo, c->kd_hi,
                                                 gamma, c->diag_full, c->diag_red);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(c->s));  // host vectors die at return
    fill_maxr(c, -1);
  } catch (...) {
    destroy_ctx(c);
    throw;
  }
  return c;
}

void copy_slot(pqdts_ctx *c, int slot, double *host, const double *src, int layout, bool to_device) {
  double *dev = slot_ptr(c, slot);
  const int M = c->M, N = c->N;
  require(layout == PQDTS_ROW_MAJOR || layout == PQDTS_COL_MAJOR, "invalid layout");
  require(host != nullptr || src != nullptr, "null host pointer");
  std::vector<double> tmp;
  if (to_device) {
    const double *h = src;
    if (layout == PQDTS_COL_MAJOR) {
      tmp.resize(c->MN);
      for (int n = 0; n < N; n++)
        for (int i = 0; i < M; i++) tmp[(size_t)i * N + n] = src[(size_t)n * M + i];
      h = tmp.data();
    }
    HIPC(hipMemcpyAsync(dev, h, sizeof(double) * c->MN, hipMemcpyHostToDevice, c->s));
    HIPC(hipStreamSynchronize(c->s));
  } else {
    double *h = host;
    if (layout == PQDTS_COL_MAJOR) {
      tmp.resize(c->MN);
      h = tmp.data();
    }
    HIPC(hipMemcpyAsync(h, dev, sizeof(double) * c->MN, hipMemcpyDeviceToHost, c->s));
    HIPC(hipStreamSynchronize(c->s));
    if (layout == PQDTS_COL_MAJOR)
      for (int n = 0; n < N; n++)
        for (int i = 0; i < M; i++) host[(size_t)n * M + i] = tmp[(size_t)i * N + n];
  }
}

pqdts_ctx *checked(pqdts_ctx *c) {
  require(c != nullptr, "null context");
  return c;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *pqdts_last_error(void) { return g_err.c_str(); }

int pqdts_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int pqdts_create(int64_t M, int64_t N, int64_t D, int64_t nF, const int32_t *F_row, const int32_t *F_col,
                 const double *F_val, int64_t nP, const int32_t *P_row, const int32_t *P_col, const double *P_val,
                 double gamma, pqdts_ctx **out) {
  return guarded([&] {
    require(out != nullptr, "null output pointer");
    *out = nullptr;
    *out = create_ctx(M, N, D, nF, F_row, F_col, F_val, nP, P_row, P_col, P_val, gamma);
  });
}

int pqdts_destroy(pqdts_ctx *ctx) {
  return guarded([&] { destroy_ctx(ctx); });
}

int pqdts_put(pqdts_ctx *ctx, int slot, const double *host, int layout) {
  return guarded([&] { copy_slot(checked(ctx), slot, nullptr, host, layout, true); });
}

int pqdts_get(pqdts_ctx *ctx, int slot, double *host, int layout) {
  return guarded([&] { copy_slot(checked(ctx), slot, host, nullptr, layout, false); });
}

int pqdts_get_O(pqdts_ctx *ctx, double *host) {
  return guarded([&] {
    checked(ctx);
    require(host != nullptr, "null host pointer");
    HIPC(hipMemcpyAsync(host, ctx->O, sizeof(double) * ctx->DN, hipMemcpyDeviceToHost, ctx->s));
    HIPC(hipStreamSynchronize(ctx->s));
  });
}

int pqdts_set_maxr(pqdts_ctx *ctx, const int32_t *maxr) {
  return guarded([&] {
    checked(ctx);
    require(maxr != nullptr, "null maxr");
    bool valid = true;
    for (int i = 0; i < ctx->M; i++) {
      require(maxr[i] >= -1 && maxr[i] < ctx->N, "maxr out of range at row " + std::to_string(i));
      valid = valid && maxr[i] >= 0;
    }
    HIPC(hipMemcpyAsync(ctx->maxr, maxr, sizeof(int) * ctx->M, hipMemcpyHostToDevice, ctx->s));
    HIPC(hipStreamSynchronize(ctx->s));
    ctx->maxr_valid = valid;
  });
}

int pqdts_get_maxr(pqdts_ctx *ctx, int32_t *maxr) {
  return guarded([&] {
    checked(ctx);
    HIPC(hipMemcpyAsync(maxr, ctx->maxr, sizeof(int) * ctx->M, hipMemcpyDeviceToHost, ctx->s));
    HIPC(hipStreamSynchronize(ctx->s));
  });
}

int pqdts_get_ik(pqdts_ctx *ctx, int8_t *ik) {
  return guarded([&] {
    checked(ctx);
    HIPC(hipMemcpyAsync(ik, ctx->ik, ctx->MN, hipMemcpyDeviceToHost, ctx->s));
    HIPC(hipStreamSynchronize(ctx->s));
  });
}

int pqdts_set_fullspace(pqdts_ctx *ctx, int full) {
  return guarded([&] { checked(ctx)->full = full != 0; });
}

int pqdts_get_fullspace(pqdts_ctx *ctx, int *full) {
  return guarded([&] { *full = checked(ctx)->full ? 1 : 0; });
}

int pqdts_L(pqdts_ctx *ctx, int x, double *val) {
  return guarded([&] { *val = op_L(checked(ctx), slot_ptr(ctx, x)); });
}

int pqdts_dL(pqdts_ctx *ctx, int x, int g) {
  return guarded([&] {
    op_dL(checked(ctx), slot_ptr(ctx, x), slot_ptr(ctx, g));
    HIPC(hipStreamSynchronize(ctx->s));
  });
}

int pqdts_BF(pqdts_ctx *ctx, int dd, int bx) {
  return guarded([&] {
    require(dd != bx, "BF input and output must differ");
    op_BF(checked(ctx), slot_ptr(ctx, dd), slot_ptr(ctx, bx), false, false, false);
    HIPC(hipStreamSynchronize(ctx->s));
  });
}

int pqdts_Bdiag(pqdts_ctx *ctx, double *diag) {
  return guarded([&] {
    checked(ctx);
    if (diag) {
      HIPC(hipMemcpyAsync(diag, ctx->diag(), sizeof(double) * ctx->M, hipMemcpyDeviceToHost, ctx->s));
      HIPC(hipStreamSynchronize(ctx->s));
    }
  });
}

int pqdts_doproj(pqdts_ctx *ctx, int x) {
  return guarded([&] {
    op_doproj(checked(ctx), slot_ptr(ctx, x));
    HIPC(hipStreamSynchronize(ctx->s));
  });
}

int pqdts_doproj2(pqdts_ctx *ctx, int x, double *xmin) {
  return guarded([&] { *xmin = op_doproj2(checked(ctx), slot_ptr(ctx, x)); });
}

int pqdts_conv_test(pqdts_ctx *ctx, int x, int g, double *conv) {
  return guarded([&] { *conv = op_conv(checked(ctx), slot_ptr(ctx, x), slot_ptr(ctx, g), false); });
}

int pqdts_pdot(pqdts_ctx *ctx, int x, int y, double *val) {
  return guarded([&] { *val = op_pdot(checked(ctx), slot_ptr(ctx, x), slot_ptr(ctx, y)); });
}

#define BLAS1_BODY(z, x, y, a, b, kind)                  \
  return guarded([&] {                                   \
    op_axpby(checked(ctx), z, x, y, a, b, kind);         \
    HIPC(hipStreamSynchronize(ctx->s));                  \
  })

int pqdts_paxpy(pqdts_ctx *ctx, double alpha, int x, int y) {
  BLAS1_BODY(slot_ptr(ctx, y), slot_ptr(ctx, x), slot_ptr(ctx, y), alpha, 0.0, 1);
}
int pqdts_pgaxpy(pqdts_ctx *ctx, double alpha, int x, double beta, int y) {
  BLAS1_BODY(slot_ptr(ctx, y), slot_ptr(ctx, x), slot_ptr(ctx, y), alpha, beta, 0);
}
int pqdts_pgaxpy2(pqdts_ctx *ctx, double alpha, int x, double beta, int y, int z) {
  BLAS1_BODY(slot_ptr(ctx, z), slot_ptr(ctx, x), slot_ptr(ctx, y), alpha, beta, 0);
}
int pqdts_pset(pqdts_ctx *ctx, double alpha, int x) {
  BLAS1_BODY(slot_ptr(ctx, x), nullptr, nullptr, alpha, 0.0, 4);
}
int pqdts_pcopy(pqdts_ctx *ctx, int x, int y) {
  BLAS1_BODY(slot_ptr(ctx, y), slot_ptr(ctx, x), nullptr, 1.0, 0.0, 3);
}
int pqdts_pscale(pqdts_ctx *ctx, double alpha, int x) {
  BLAS1_BODY(slot_ptr(ctx, x), slot_ptr(ctx, x), nullptr, alpha, 0.0, 2);
}
int pqdts_pcopy_scale(pqdts_ctx *ctx, double alpha, int x, int y) {
  BLAS1_BODY(slot_ptr(ctx, y), slot_ptr(ctx, x), nullptr, alpha, 0.0, 2);
}

int pqdts_transform(pqdts_ctx *ctx) {
  return guarded([&] {
    op_transform(checked(ctx), true);
    HIPC(hipStreamSynchronize(ctx->s));
  });
}

int pqdts_lalpha_min2(pqdts_ctx *ctx, double *alpha, double *dlalpha) {
  return guarded([&] { op_linesearch(checked(ctx), false, alpha, dlalpha); });
}

int pqdts_lalpha_min_auglag(pqdts_ctx *ctx, double *alpha, double *dlalpha) {
  return guarded([&] { op_linesearch(checked(ctx), true, alpha, dlalpha); });
}

int pqdts_solve(pqdts_ctx *ctx, int mode, int maxiter, double eps, int start_stage, int output, pqdts_event_cb cb,
                void *user, double *conv) {
  return guarded([&] {
    checked(ctx);
    require(mode == 1 || mode == 2, "mode must be 1 or 2");
    require(start_stage == 1 || start_stage == 2, "start_stage must be 1 or 2");
    ctx->cb = cb;
    ctx->user = user;
    try {
      double cv = solve(ctx, mode, maxiter, eps, start_stage, output);
      if (conv) *conv = cv;
    } catch (...) {
      ctx->cb = nullptr;
      ctx->user = nullptr;
      throw;
    }
    ctx->cb = nullptr;
    ctx->user = nullptr;
  });
}

int pqdts_timers(pqdts_ctx *ctx, int cap, char (*names)[24], int *counts, double *seconds) {
  int n = 0;
  int rc = guarded([&] {
    checked(ctx);
    for (const auto &name : ctx->timer_order) {
      if (n < cap) {
        const Timer &t = ctx->timers[name];
        if (names) {
          std::strncpy(names[n], name.c_str(), 23);
          names[n][23] = 0;
        }
        if (counts) counts[n] = t.count;
        if (seconds) seconds[n] = t.sec;
      }
      n++;
    }
  });
  return rc == 0 ? n : -1;
}

int pqdts_kernel_timing(pqdts_ctx *ctx, int on) {
  return guarded([&] {
    checked(ctx);
    HIPC(hipStreamSynchronize(ctx->s));
    ctx->kfold();
    ctx->ktiming = on != 0;
    for (size_t i = 0; i < ctx->kcount.size(); i++) {
      ctx->kcount[i] = 0;
      ctx->kms[i] = 0.0;
    }
  });
}

int pqdts_kernel_stats(pqdts_ctx *ctx, int cap, char (*names)[24], int *counts, double *ms) {
  int n = 0;
  int rc = guarded([&] {
    checked(ctx);
    HIPC(hipStreamSynchronize(ctx->s));
    ctx->kfold();
    n = (int)ctx->knames.size();
    for (int i = 0; i < std::min(n, cap); i++) {
      if (names) {
        std::strncpy(names[i], ctx->knames[i].c_str(), 23);
        names[i][23] = 0;
      }
      if (counts) counts[i] = ctx->kcount[i];
      if (ms) ms[i] = ctx->kms[i];
    }
  });
  return rc == 0 ? n : -1;
}

int pqdts_info(pqdts_ctx *ctx, int64_t *info) {
  return guarded([&] {
    checked(ctx);
    require(info != nullptr, "null info");
    info[0] = ctx->M;
    info[1] = ctx->N;
    info[2] = ctx->D;
    info[3] = ctx->nspan;
    info[4] = ctx->tile_elems;
    info[5] = ctx->nrt;
    info[6] = ctx->nct;
    info[7] = ctx->use_gram() ? 1 : 0;
    info[8] = ctx->gram_elems;
    int64_t gnnz = 0;
    for (int w : ctx->gw) gnnz += w;
    info[9] = gnnz;
    info[10] = 0;
    info[11] = 0;
  });
}

int pqdts_set_hvp_path(pqdts_ctx *ctx, int path) {
  return guarded([&] {
    checked(ctx);
    require(path >= 0 && path <= 2, "hvp path must be 0 (auto), 1 (F products) or 2 (Gram)");
    ctx->hvp_pref = path;
    if (ctx->use_gram()) ensure_gram(ctx);
  });
}

int pqdts_simplexproj_rows(const double *y, double *x, int64_t rows, int64_t n) {
  return guarded([&] {
    require(rows >= 0 && n >= 1 && rows * n < (int64_t)INT32_MAX * 8, "bad shape");
    require(y != nullptr && x != nullptr && x != y, "x must be a distinct non-null buffer");
    if (rows == 0) return;
    double *d = dalloc<double>(rows * n);
    hipError_t e1 = hipMemcpy(d, y, sizeof(double) * rows * n, hipMemcpyHostToDevice);
    if (e1 == hipSuccess) {
      k_simplex_rows<<<cdiv(rows, WAVES_PER_BLOCK), BLOCK>>>(d, nullptr, nullptr, 0.0, (int)rows, (int)n, 0);
      e1 = hipGetLastError();
    }
    if (e1 == hipSuccess) e1 = hipMemcpy(x, d, sizeof(double) * rows * n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    HIPC(e1);
  });
}

}  // extern "C"
