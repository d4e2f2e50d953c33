// Host side of pqdts_amd: problem contexts, the batched launch engine, the
// lockstep minimisation loop (pqdts.f90:1908-2167) and the C ABI declared in
// include/pqdts_amd.h.
//
// A context (pqdts_ctx) owns one problem's device arrays and an immutable device
// descriptor. An Engine drives any number of contexts on one HIP stream: every
// operator is one launch (per precision / Hessian path) over all the problems
// that need it, and every host synchronisation fetches the scalars of all of
// them at once. The plain solve is an engine with one problem.
#include "../../include/pqdts_amd.h"
#include "kernels.hip"
#include "face.hip"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
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

constexpr int TILE = 64;  // rows and columns of a band-GEMM block tile
// loosest relative CG residual of a polish iteration
constexpr double kPolishLooseRtol = 1e-3;
constexpr int KC = 16;    // band chunk staged per LDS buffer (4 MFMA k-steps)
static_assert(KC == FACE_KC, "the direct polish reads G's band pack");

// device scalar slots per problem
enum {
  S_LA = 0, S_LB, S_XMIN, S_NACT, S_ND, S_B1, S_V2, S_RZ0, S_RZ1, S_RR, S_CONV, S_DOT,
  S_EPSCG, S_CGACT, S_CGN,  // device-side CG loop: eps, still-iterating flag, rounds done
  S_NCHG, S_STEP,            // polish: active entries that changed, minus the step's largest entry
  S_FMAX, S_FTOT, S_FBW, S_FFAIL,  // direct polish: largest column block, total block size, band of
                                   // the Schur complement, its factorisation failed
  S_LT = 21,                 // line-search trial k: data term S_LT + k, smoothing S_LTB + k, xmin S_LTX + k
  S_LTB = S_LT + LS_K,
  S_LTX = S_LTB + LS_K
};
static_assert(S_FFAIL < S_LT && S_LTX + LS_K <= NSCAL, "scalar slots");
// partial-sum regions per problem
enum { R_A = 0, R_B = 1, R_C = 2 };

inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Band GEMM launch. Launches with fewer workgroups than fill the chip (4 per CU)
// take the latency form of the epilogue, larger ones the occupancy form (for FP64
// the two measured 4-13 % apart per CG round, in opposite directions, in the two
// regimes). Both forms compute the same bits.
constexpr int LAT_BLOCKS = 1024;
template <typename T, int BM, int EM, bool MASK, bool DOT>
void launch_gemm(int tot, hipStream_t s, const Prob *probs, const WorkList &wl, const GemmOps &ops) {
  if (tot < LAT_BLOCKS) {
    kb_band_gemm<T, KC, BM, EM, MASK, DOT, true><<<tot, BLOCK, 0, s>>>(probs, wl, ops);
    return;
  }
  kb_band_gemm<T, KC, BM, EM, MASK, DOT, false><<<tot, BLOCK, 0, s>>>(probs, wl, ops);
}

template <class T>
T *dalloc(int64_t n) {
  void *p = nullptr;
  HIPC(hipMalloc(&p, sizeof(T) * (size_t)std::max<int64_t>(n, 1)));
  return (T *)p;
}

struct Timer {
  int count = 0;
  double sec = 0.0;
  std::chrono::steady_clock::time_point t0;
  std::vector<hipEvent_t> open;  // device timing: start events of the running sections
};

struct Engine;

}  // namespace

// ===========================================================================
// problem context
// ===========================================================================
struct pqdts_ctx {
  int M = 0, N = 0, D = 0;
  int64_t MN = 0, DN = 0;
  double gamma = 0.0;
  hipStream_t s = nullptr;
  // device storage
  int *flo = nullptr, *flen = nullptr;
  int64_t *foff = nullptr;
  double *fval = nullptr;
  uint8_t *fhas = nullptr;
  int *kd_lo = nullptr, *kd_hi = nullptr;
  double *P = nullptr, *O = nullptr, *R = nullptr;  // R = O - P, the gradient's operand
  double *slot[6] = {};
  double *diag_full = nullptr, *diag_red = nullptr;
  int *maxr = nullptr;
  uint8_t *ik = nullptr;
  double *ev = nullptr, *gm = nullptr, *lam = nullptr, *lmin = nullptr;
  double *part = nullptr, *scal = nullptr;
  int *pn = nullptr;  // partial counts per region (see Prob::pn)
  int pcap = 0;
  struct DBand {
    int ntiles = 0;
    double *pack = nullptr;
    float *pack32 = nullptr;
    int64_t *off = nullptr;
    int *klo = nullptr, *nch = nullptr;
    int64_t elems = 0;
  } bf, bt, bg;
  Prob desc{};               // host copy of the device descriptor
  Prob *ddesc = nullptr;     // device descriptor (standalone use)
  int64_t nspan = 0;
  // state shared with the reference's module variables
  bool full = true;          // tfullspace (pqdts.f90:97)
  bool maxr_valid = false;   // every maxr[i] in [0, N)
  // Hessian path and precision
  int hvp_pref = 0;          // 0 auto, 1 F products, 2 Gram
  bool gram_auto = false;
  int64_t gram_elems = 0, gram_nnz = 0;
  std::vector<int> glo, gw;
  int hs = 1;                // split of the Gram product over band chunks (Prob::hs)
  int hs_pref = 0;           // 0: by shape (apply_gram_split), else fixed
  int gram_nch = 0;          // most band chunks of one G tile
  double *hw = nullptr;      // hs x M x N partial products when hs > 1
  float *p32 = nullptr;      // FP32 expanded CG direction (mixed precision; Prob::p32)
  double *pb = nullptr;      // Prob::pb (mixed precision)
  bool use_gram() const { return hvp_pref == 2 || (hvp_pref == 0 && gram_auto); }
  // polish after the minimisation (pqdts_set_polish) and the outcome of the last one
  struct Polish {
    int max_outer = 0;        // 0: no polish
    double cg_rtol = 1e-6;    // the tightest relative CG residual
    double tol = 1e-10;       // target distance of the result from the minimiser
    double budget = 0.0;      // CG rounds of the whole polish, at most this multiple of the
                              // minimisation's PCG rounds (<= 0: no limit)
    int fp32 = 0;             // 1: FP32 Hessian products where the solve is mixed
    int method = 0;           // 0: Newton-CG in the reduced space; 1: direct face solves
  } pol;
  struct PolishOut {
    int status = 0;           // 0 not run, 1 converged, 2 stopped at a cap (result: the
                              // minimisation's own)
    int iters = 0, rounds = 0;
    int64_t nact = 0;
    double step = 0.0;
  } pol_out;
  std::unique_ptr<Engine> eng;  // standalone engine (this problem alone, on stream s)
  Engine *batch = nullptr;      // the batch engine whose slot holds this context, if any
  Engine &engine();
};

namespace {

Band to_band(const pqdts_ctx::DBand &b) {
  Band r{};
  r.pack = b.pack;
  r.pack32 = b.pack32;
  r.off = b.off;
  r.klo = b.klo;
  r.nch = b.nch;
  r.ntiles = b.ntiles;
  return r;
}

void refresh_desc(pqdts_ctx *c) {
  Prob &p = c->desc;
  p.M = c->M;
  p.N = c->N;
  p.D = c->D;
  p.pcap = c->pcap;
  p.gamma = c->gamma;
  for (int i = 0; i < 6; i++) p.op[i] = c->slot[i];
  p.op[OP_O] = c->O;
  p.op[OP_R] = c->R;
  p.P = c->P;
  p.diag_full = c->diag_full;
  p.diag_red = c->diag_red;
  p.maxr = c->maxr;
  p.ik = c->ik;
  p.ev = c->ev;
  p.gm = c->gm;
  p.lam = c->lam;
  p.lmin = c->lmin;
  p.flo = c->flo;
  p.flen = c->flen;
  p.foff = c->foff;
  p.fval = c->fval;
  p.fhas = c->fhas;
  p.kdlo = c->kd_lo;
  p.kdhi = c->kd_hi;
  p.bf = to_band(c->bf);
  p.bt = to_band(c->bt);
  p.bg = to_band(c->bg);
  p.part = c->part;
  p.scal = c->scal;
  p.pn = c->pn;
  p.hs = c->hs;
  p.hw = c->hw;
  HIPC(hipMemcpyAsync(c->ddesc, &c->desc, sizeof(Prob), hipMemcpyHostToDevice, c->s));
  HIPC(hipStreamSynchronize(c->s));
}

// ===========================================================================
// Engine: batched launches over problem slots
// ===========================================================================
struct PState;

struct Engine {
  hipStream_t s = nullptr;
  bool own_stream = false;
  int cap = 0;
  std::vector<pqdts_ctx *> ctx;   // slot -> problem (nullptr: free)
  Prob *dprobs = nullptr;         // device descriptors [cap]
  double *dscal = nullptr;        // device scalars [cap * NSCAL] (batched engines)
  double *hscal = nullptr;        // pinned host copy
  bool private_scal = false;      // standalone: scalars live in the context
  // Captured CG blocks (hipGraph): the launches of CG rounds j0..j1 for one set of
  // problems, replayed while that set, its spaces and precisions stay the same
  // (the tail of a batch repeats one set for thousands of Newton iterations).
  // Batched engines only, without kernel timing or synchronising timers.
  std::map<std::vector<int64_t>, hipGraphExec_t> cg_graphs;
  // 0: eager launches; 1: CG rounds in captured blocks of 8 between host checks;
  // 2: a Newton iteration as two captured graphs around the host-driven line search
  int graphs_on = 0;
  std::vector<int64_t> gen;  // per slot: attach count, so a graph never outlives its problem
  int64_t gen_next = 0;
  std::vector<double> arg;        // per-slot scalar argument of the next launch
  int guard = -1;                 // launches skip problems whose scal[guard] is 0
  double *dlog = nullptr;         // device-side CG logs [cap][CG_LOG_ROUNDS][5]
  double *hlog = nullptr;         // pinned host copy
  std::vector<int> fp32;          // per-slot: mixed precision Hessian products
  // direct polish: per-slot buffers, their device copies and the capacity (doubles)
  // of each slot's block storage
  Face *dfaces = nullptr;
  Face *hface = nullptr;  // pinned host copy of dfaces (cap entries)
  std::vector<int64_t> fkcap, fkcap_old;
  // Stream-ordered allocation (hipMallocAsync) for the polish's buffers where the
  // device supports memory pools. Elsewhere (MI355X on ROCm 7.2 reports no pool
  // support) the buffers come from hipMalloc at admission, and those the polish
  // gives back are kept on a list and freed when the engine is idle or destroyed,
  // or once kDeferMax have gathered: hipFree waits for the whole device, so freeing
  // them while the engine runs would stall every other lane's stream.
  bool pools = false;
  std::vector<void *> deferred;
  static constexpr size_t kDeferMax = 4096;
  void release_deferred() {
    if (deferred.empty()) return;
    (void)hipStreamSynchronize(s);
    for (void *q : deferred) (void)hipFree(q);
    deferred.clear();
  }
  template <class T>
  T *salloc(int64_t n) {
    void *p = nullptr;
    const size_t b = sizeof(T) * (size_t)std::max<int64_t>(n, 1);
    if (pools)
      HIPC(hipMallocAsync(&p, b, s));
    else
      HIPC(hipMalloc(&p, b));
    return (T *)p;
  }
  void sfree(void *p) {
    if (!p) return;
    if (pools)
      (void)hipFreeAsync(p, s);
    else
      deferred.push_back(p);
  }
  // frees the deferred buffers once the engine is idle or too many have gathered
  void settle_deferred() {
    if (deferred.size() > kDeferMax || !any_active()) release_deferred();
  }
  // timers: host wall clock around the (asynchronous) launches, or with
  // dev_timers the device time between HIP events recorded on the stream at the
  // start and the end of each section
  std::map<std::string, Timer> timers;
  std::vector<std::string> timer_order;
  bool dev_timers = false;
  struct TPending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<TPending> tpending;
  // per-kernel HIP-event timing
  bool ktiming = false;
  int ksample = 1;    // time every ksample-th operator launch (a sample of the durations)
  uint64_t kseq = 0;
  struct Work {
    double flops = 0.0, bytes = 0.0;
  };
  struct KPair {
    hipEvent_t a, b;
    int kid;
    Work work;
    int round;                                // CG round of a guarded launch, else 0
    bool idle_at_stop;                        // no work in the round where a problem's CG stops
    std::vector<std::pair<int, Work>> parts;  // guarded launch: (slot, work) per problem
  };
  std::vector<KPair> kpending, kdeferred;
  int kround = 0;  // current CG round while CG launches are guarded
  std::vector<hipEvent_t> kpool;
  std::vector<std::string> knames;
  std::vector<int> kcount;
  std::vector<double> kms, kwork, kbytes;
  std::vector<std::unique_ptr<PState>> st;  // solver state per slot

  ~Engine();

  // Replaces the engine's own stream by one restricted to the CUs of `mask`.
  void set_cu_mask(const uint32_t *mask, int nwords) {
    require(own_stream, "only an engine with its own stream can be restricted");
    HIPC(hipStreamSynchronize(s));
    hipStream_t ns = nullptr;
    HIPC(hipExtStreamCreateWithCUMask(&ns, (uint32_t)nwords, mask));
    (void)hipStreamDestroy(s);
    s = ns;
  }
  void init(int capacity, hipStream_t stream, bool make_stream) {
    cap = capacity;
    ctx.assign(cap, nullptr);
    arg.assign(cap, 0.0);
    fp32.assign(cap, 0);
    st.resize(cap);
    if (make_stream) {
      HIPC(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      own_stream = true;
    } else {
      s = stream;
    }
    dprobs = dalloc<Prob>(cap);
    dfaces = dalloc<Face>(cap);
    {
      int dev = 0, v = 0;
      HIPC(hipGetDevice(&dev));
      if (hipDeviceGetAttribute(&v, hipDeviceAttributeMemoryPoolsSupported, dev) == hipSuccess && v) {
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
          // keep freed blocks in the pool for the next problems' polish
          uint64_t keep = UINT64_MAX;
          (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
          pools = true;
        }
      }
    }
    HIPC(hipHostMalloc((void **)&hface, sizeof(Face) * cap, hipHostMallocDefault));
    for (int p = 0; p < cap; p++) hface[p] = Face{};
    fkcap.assign(cap, 0);
    fkcap_old.assign(cap, 0);
    HIPC(hipMemcpyAsync(dfaces, hface, sizeof(Face) * cap, hipMemcpyHostToDevice, s));
    HIPC(hipStreamSynchronize(s));
    HIPC(hipHostMalloc((void **)&hscal, sizeof(double) * NSCAL * cap, hipHostMallocDefault));
    dlog = dalloc<double>((int64_t)cap * CG_LOG_ROUNDS * 5);
    HIPC(hipHostMalloc((void **)&hlog, sizeof(double) * cap * CG_LOG_ROUNDS * 5, hipHostMallocDefault));
    graphs_on = getenv("PQDTS_GRAPHS") ? std::max(0, std::min(2, atoi(getenv("PQDTS_GRAPHS")))) : 0;
  }
  void drop_graphs() {
    for (auto &kv : cg_graphs) (void)hipGraphExecDestroy(kv.second);
    cg_graphs.clear();
  }
  // Runs CG rounds j0..j1 of the problems S through a captured graph: captured
  // on first use for this key (rounds, slots and their attach generations,
  // spaces, precisions, Hessian paths and splits), replayed afterwards. The
  // launches are the eager ones, so the results are the same bits.
  template <class R>
  void graph_rounds(const std::vector<int> &S, int j0, int j1, R &&rounds) {
    graph_run(S, j0, j1, [&] { rounds(j0, j1); });
  }
  // Runs the launches fn records for the problems S through a graph captured on
  // first use for this key (tag0, tag1, the problems, their attach generations,
  // spaces, precisions, Hessian paths and splits) and replayed afterwards. fn may
  // change host state only in ways that the same key determines.
  template <class Fn>
  void graph_run(const std::vector<int> &S, int tag0, int tag1, Fn &&fn) {
    std::vector<int64_t> key{tag0, tag1, (int64_t)S.size()};
    for (int p : S) {
      const pqdts_ctx *c = ctx[p];
      key.push_back(p);
      key.push_back(gen[p]);
      key.push_back((c->full ? 1 : 0) | (late(p) ? 2 : 0) | (prod32(p) ? 4 : 0) | (c->use_gram() ? 8 : 0) |
                    ((int64_t)c->hs << 4) | ((int64_t)c->hs_pref << 12) | ((int64_t)polishing(p) << 20));
    }
    auto it = cg_graphs.find(key);
    if (it == cg_graphs.end()) {
      if (cg_graphs.size() >= 256) drop_graphs();
      hipGraph_t g = nullptr;
      HIPC(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        fn();
      } catch (...) {
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      HIPC(hipStreamEndCapture(s, &g));
      hipGraphExec_t ex = nullptr;
      const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPC(e);
      it = cg_graphs.emplace(std::move(key), ex).first;
    }
    HIPC(hipGraphLaunch(it->second, s));
  }

  // Places a context in a slot; a batched engine redirects its scalars into the
  // engine's contiguous array so that one copy fetches every problem's scalars.
  void attach(int slot, pqdts_ctx *c) {
    ctx[slot] = c;
    if ((int)gen.size() < cap) gen.assign(cap, -1);
    gen[slot] = ++gen_next;
    Prob d = c->desc;
    d.cglog = dlog + (int64_t)slot * CG_LOG_ROUNDS * 5;
    // the CG kernels write the FP32 direction copy only for mixed-precision slots
    d.p32 = fp32[slot] ? c->p32 : nullptr;
    d.pb = c->pb;
    if (!private_scal) {
      if (!dscal) dscal = dalloc<double>((int64_t)NSCAL * cap);
      d.scal = dscal + (int64_t)NSCAL * slot;
    }
    HIPC(hipMemcpyAsync(dprobs + slot, &d, sizeof(Prob), hipMemcpyHostToDevice, s));
    HIPC(hipStreamSynchronize(s));
  }
  void detach(int slot) {
    face_free(slot);
    if (ctx[slot] && ctx[slot]->batch == this) ctx[slot]->batch = nullptr;
    ctx[slot] = nullptr;
    st[slot].reset();
  }

  // ---- direct polish buffers (Face) of a slot
  void face_upload_all() {
    HIPC(hipMemcpyAsync(dfaces, hface, sizeof(Face) * cap, hipMemcpyHostToDevice, s));
    HIPC(hipStreamSynchronize(s));
  }
  void face_free(int p) {
    if (!hface || p < 0 || p >= cap) return;
    Face &f = hface[p];
    if (!f.pos) return;
    void *bufs[] = {f.pos, f.crow, f.cnt, f.koff, f.kb, f.koff_old, f.kb_old, f.cdirty, f.cfresh,
                    f.S, f.lam, f.rhs, f.slast, f.cstep};
    for (void *q : bufs) sfree(q);
    f = Face{};
    fkcap[p] = fkcap_old[p] = 0;
  }
  // every buffer but the block storage; every column's block is formed in the first
  // iteration (cdirty = 1)
  void face_alloc(int p) {
    face_free(p);
    const int64_t M = ctx[p]->M, N = ctx[p]->N;
    Face &f = hface[p];
    try {
      f.pos = salloc<int>(M * N);
      f.crow = salloc<int>(M * N);
      f.cnt = salloc<int>(N);
      f.koff = salloc<int64_t>(N + 1);
      f.koff_old = salloc<int64_t>(N + 1);
      f.cdirty = salloc<uint8_t>(N);
      f.cfresh = salloc<uint8_t>(N);
      f.S = salloc<double>(M * M);
      f.lam = salloc<double>(M);
      f.rhs = salloc<double>(M);
      f.slast = salloc<int>(M);
      f.cstep = salloc<double>(N);
      HIPC(hipMemsetAsync(f.cdirty, 1, (size_t)N, s));
    } catch (...) {
      face_free(p);
      throw;
    }
  }
  // the last iteration's blocks become the old ones
  void face_swap(int p) {
    Face &f = hface[p];
    std::swap(f.kb, f.kb_old);
    std::swap(f.koff, f.koff_old);
    std::swap(fkcap[p], fkcap_old[p]);
  }
  // block storage for `need` doubles
  void face_grow(int p, int64_t need) {
    if (need <= fkcap[p]) return;
    Face &f = hface[p];
    sfree(f.kb);
    f.kb = nullptr;
    fkcap[p] = 0;
    const int64_t c2 = need + need / 4 + 1024;
    f.kb = salloc<double>(c2);
    fkcap[p] = c2;
  }
  // A problem admitted with the direct polish gets its polish buffers now, with block
  // storage for 4 M N doubles in each of the two generations, so that the polish
  // itself allocates only when a problem needs more.
  void face_prepare(int p) {
    const pqdts_ctx *c = ctx[p];
    if (c->pol.method != 1 || c->pol.max_outer <= 0 || c->M > FACE_MAX_ROWS) return;
    face_alloc(p);
    const int64_t r = 4 * c->MN;
    face_grow(p, r);
    face_swap(p);
    face_grow(p, r);
    face_swap(p);
    face_upload_all();
  }
  void face_step(const std::vector<int> &QD, std::vector<int> &Fz);

  // ---- timing (never throws: tstop runs in destructors)
  hipEvent_t tevent() {
    if (!kpool.empty()) {
      hipEvent_t e = kpool.back();
      kpool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    return hipEventCreate(&e) == hipSuccess ? e : nullptr;
  }
  void tstart(const char *name) {
    auto it = timers.find(name);
    if (it == timers.end()) {
      it = timers.emplace(name, Timer{}).first;
      timer_order.push_back(name);
    }
    it->second.count++;
    if (dev_timers) {
      hipEvent_t a = tevent();
      if (a && hipEventRecord(a, s) != hipSuccess) a = nullptr;
      it->second.open.push_back(a);
    } else {
      it->second.t0 = std::chrono::steady_clock::now();
    }
  }
  void tstop(const char *name) {
    auto &t = timers[name];
    if (!t.open.empty()) {
      hipEvent_t a = t.open.back();
      t.open.pop_back();
      hipEvent_t b = a ? tevent() : nullptr;
      if (b && hipEventRecord(b, s) == hipSuccess) {
        tpending.push_back({name, a, b});
      } else {
        if (a) kpool.push_back(a);
        if (b) kpool.push_back(b);
      }
    } else if (!dev_timers) {
      t.sec += std::chrono::duration<double>(std::chrono::steady_clock::now() - t.t0).count();
    }
  }
  // adds the device time of the finished sections (the stream has been synchronised)
  void tresolve() {
    for (auto &p : tpending) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) timers[p.name].sec += ms / 1e3;
      kpool.push_back(p.a);
      kpool.push_back(p.b);
    }
    tpending.clear();
  }
  int kernel_id(const std::string &name) {
    for (size_t i = 0; i < knames.size(); i++)
      if (knames[i] == name) return (int)i;
    knames.push_back(name);
    kcount.push_back(0);
    kms.push_back(0.0);
    kwork.push_back(0.0);
    kbytes.push_back(0.0);
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
  hipEvent_t kbegin() {
    if (!ktiming) return nullptr;
    if (ksample > 1 && (kseq++ % (uint64_t)ksample) != 0) return nullptr;
    hipEvent_t a = kevent();
    HIPC(hipEventRecord(a, s));
    return a;
  }
  // Algorithmic work of one band product of problem p: flops = 2 x stored
  // elements of the banded operand (F for F V and F^T W, G for G p) x N; bytes =
  // that operand once (FP32 copy in mixed precision) + the dense operand read once
  // + the result written once (+ P for the loss and the gradient, the stencil
  // source for the gradient, the active-set mask for the Hessian product).
  // em: the epilogue (E_STORE_R also reads P and writes R; E_GRAD reads R and the
  // stencil source X; E_HVP reads the direction for the stencil and the mask)
  Work gemm_work(int p, int band, bool f32 = false, int em = E_HVP) const {
    const pqdts_ctx *c = ctx[p];
    const double MN = (double)c->MN, DN = (double)c->D * c->N, ab = f32 ? 4.0 : 8.0;
    Work w;
    if (band == BAND_G) {
      w.flops = 2.0 * (double)c->gram_nnz * c->N;
      w.bytes = ab * (double)c->gram_nnz + 16.0 * MN + MN;
    } else if (band == BAND_F) {
      w.flops = 2.0 * (double)c->nspan * c->N;
      w.bytes = ab * (double)c->nspan + 8.0 * MN + (em == E_STORE_R ? 24.0 : 8.0) * DN;
    } else {
      w.flops = 2.0 * (double)c->nspan * c->N;
      w.bytes = ab * (double)c->nspan + 8.0 * DN + 16.0 * MN + (em == E_GRAD ? 0.0 : MN);
    }
    return w;
  }
  // Streaming kernels: `per_elem` bytes per element of the M x N arrays.
  Work stream_work(int p, double per_elem) const {
    Work w;
    w.bytes = per_elem * (double)ctx[p]->MN;
    return w;
  }
  // Closes a timed launch over S (work(p) per problem). Inside the guarded CG
  // loop a problem whose CG already stopped does no work, so its share is
  // settled by kresolve() once the round counts are known.
  template <class WF>
  void kend(hipEvent_t a, const std::string &name, const std::vector<int> &S, WF work,
            bool idle_at_stop = false) {
    if (!a) return;
    hipEvent_t b = kevent();
    HIPC(hipEventRecord(b, s));
    KPair k{a, b, kernel_id(name), Work{}, 0, idle_at_stop, {}};
    if (guard >= 0 && kround > 0) {
      k.round = kround;
      for (int p : S) k.parts.push_back({p, work(p)});
    } else {
      for (int p : S) {
        const Work w = work(p);
        k.work.flops += w.flops;
        k.work.bytes += w.bytes;
      }
    }
    kpending.push_back(std::move(k));
  }
  void kfold() {
    for (auto &p : kpending) {
      float ms = 0.f;
      HIPC(hipEventElapsedTime(&ms, p.a, p.b));
      kcount[p.kid]++;
      kms[p.kid] += ms;
      kwork[p.kid] += p.work.flops;
      kbytes[p.kid] += p.work.bytes;
      kpool.push_back(p.a);
      kpool.push_back(p.b);
      if (p.round > 0) kdeferred.push_back(std::move(p));
    }
    kpending.clear();
  }
  // Work of guarded CG launches: problem p ran rounds 1..scal[S_CGN]; a launch
  // marked idle_at_stop did nothing in the last of them if the CG stopped there.
  void kresolve() {
    for (auto &d : kdeferred)
      for (auto &pw : d.parts) {
        const double *h = host_scal(pw.first);
        const int n = (int)h[S_CGN];
        if (d.round < n || (d.round == n && !(d.idle_at_stop && h[S_CGACT] == 0.0))) {
          kwork[d.kid] += pw.second.flops;
          kbytes[d.kid] += pw.second.bytes;
        }
      }
    kdeferred.clear();
  }

  // ---- scalars
  double *host_scal(int slot) { return hscal + (int64_t)NSCAL * slot; }
  void fetch() {
    if (private_scal) {
      HIPC(hipMemcpyAsync(hscal, ctx[0]->scal, sizeof(double) * NSCAL, hipMemcpyDeviceToHost, s));
    } else {
      HIPC(hipMemcpyAsync(hscal, dscal, sizeof(double) * NSCAL * cap, hipMemcpyDeviceToHost, s));
    }
    HIPC(hipStreamSynchronize(s));
    if (!kpending.empty()) kfold();
    if (!tpending.empty()) tresolve();
  }

  // ---- work lists: groups of at most MAXB slots; nb(slot) blocks each
  template <class NB, class L>
  void launch(const std::vector<int> &S, NB nb, L fn) {
    for (size_t g = 0; g < S.size(); g += MAXB) {
      WorkList wl;
      wl.guard = guard;
      wl.np = (int)std::min<size_t>(MAXB, S.size() - g);
      int tot = 0;
      for (int i = 0; i < wl.np; i++) {
        const int p = S[g + i];
        wl.pid[i] = p;
        wl.boff[i] = tot;
        tot += nb(p);
        wl.flags[i] = (ctx[p]->full ? F_FULL : 0) | (late(p) ? F_LATE : 0) | (polishing(p) ? F_POLISH : 0);
        wl.a[i] = arg[p];
      }
      wl.boff[wl.np] = tot;
      if (tot > 0) fn(wl, tot);
      HIPC(hipGetLastError());
    }
  }
  // a producer launch followed by the per-problem reduction of its partials
  template <class NB, class L>
  void launch_reduce(const std::vector<int> &S, NB nb, L fn, int region, int dst, bool is_min = false) {
    launch(S, nb, [&](const WorkList &wl, int tot) {
      fn(wl, tot);
      kb_reduce<<<wl.np, BLOCK, 0, s>>>(dprobs, wl, region, dst, is_min ? 1 : 0);
    });
  }

  int nb_rows(int p) const { return cdiv(ctx[p]->M, WAVES_PER_BLOCK); }
  int nb_elem(int p) const { return std::max(1, cdiv(ctx[p]->MN, EBLOCK)); }
  int nb_gemm(int p, int band) const {
    const pqdts_ctx *c = ctx[p];
    const int nt = band == BAND_F ? c->bf.ntiles : band == BAND_FT ? c->bt.ntiles : c->bg.ntiles;
    return cdiv(c->N, TILE) * nt;
  }

  // FP32 Hessian products for problem p: a mixed-precision slot, except while it
  // polishes with FP64 products; polishing(p): in its polish phase
  bool prod32(int p) const;
  bool polishing(int p) const;

  // Inside the CG of a mixed-precision solve (cg_late), the fold of the Hessian
  // product moves from the product's epilogue to kb_cg_update (F_LATE), except
  // for an explicitly split Gram product, whose summing pass folds.
  bool cg_late = false;
  bool late(int p) const {
    return cg_late && prod32(p) && !(ctx[p]->use_gram() && ctx[p]->hs > 1 && ctx[p]->hs_pref > 0);
  }
  std::vector<int> early_of(const std::vector<int> &S) const {
    std::vector<int> r;
    for (int p : S)
      if (!late(p)) r.push_back(p);
    return r;
  }

  std::vector<int> reduced_of(const std::vector<int> &S) const {
    std::vector<int> r;
    for (int p : S)
      if (!ctx[p]->full) r.push_back(p);
    return r;
  }
  void need_maxr(const std::vector<int> &S) const {
    for (int p : S)
      require(ctx[p]->full || ctx[p]->maxr_valid,
              "reduced space needs a valid maxr (call pqdts_transform or pqdts_set_maxr)");
  }

  // ---- band GEMM dispatch
  template <typename T, int BM, int EM, bool MASK, bool DOT>
  void gemm_t(const std::vector<int> &S, GemmOps ops, const char *kname) {
    hipEvent_t e0 = kbegin();
    launch(S, [&](int p) { return nb_gemm(p, ops.band); }, [&](const WorkList &wl, int tot) {
      launch_gemm<T, BM, EM, MASK, DOT>(tot, s, dprobs, wl, ops);
    });
    kend(e0, kname, S, [&](int p) { return gemm_work(p, ops.band, sizeof(T) == 4, EM); });
  }

  // ================= operators (reference subroutine in brackets) =================

  // X[j,maxr] = 1 - (sum - X[j,maxr]) for reduced problems (L/dL/conv prep)
  void reconstruct(const std::vector<int> &S, int x) {
    need_maxr(S);
    launch(S, [&](int p) { return nb_rows(p); },
           [&](const WorkList &wl, int tot) { kb_reconstruct<<<tot, BLOCK, 0, s>>>(dprobs, wl, x); });
  }

  // L (pqdts.f90:917-973): data term -> scal[ia], smoothing -> scal[ib]
  // ixmin >= 0: the xmin partials kb_project left in region R_C are reduced in the
  // same launch (a line-search trial)
  void L(const std::vector<int> &S, int x, bool recon, int ia = S_LA, int ib = S_LB, int ixmin = -1) {
    struct Scope sc(this, "L");
    if (recon) reconstruct(S, x);
    // the loss product, and in the same launch the smoothing partials (region
    // R_B): at least one smoothing block per problem, so that M = 1 still stores
    // a zero partial
    GemmOps ops{BAND_F, x, OP_O, 0, 0, R_A};
    ops.sm = R_B;
    auto nsm = [&](int p) { return std::max(1, cdiv((int64_t)(ctx[p]->M - 1) * ctx[p]->N, EBLOCK)); };
    hipEvent_t e0 = kbegin();
    tstart("L_O");
    launch(S, [&](int p) { return nb_gemm(p, BAND_F) + nsm(p); }, [&](const WorkList &wl, int tot) {
      launch_gemm<double, B_PLAIN, E_LOSS, false, false>(tot, s, dprobs, wl, ops);
    });
    tstop("L_O");
    kend(e0, "spmm_F", S, [&](int p) { return gemm_work(p, BAND_F, false, E_LOSS); });
    launch(S, [](int) { return 1; }, [&](const WorkList &wl, int) {
      if (ixmin >= 0)
        kb_reduce_pn<<<wl.np, BLOCK, 0, s>>>(dprobs, wl, 3, R_A, ia, R_B, ib, R_C, ixmin, 4);
      else
        kb_reduce_pn<<<wl.np, BLOCK, 0, s>>>(dprobs, wl, 2, R_A, ia, R_B, ib, 0, 0, 0);
    });
  }
  double Lval(int p) { return host_scal(p)[S_LA] + host_scal(p)[S_LB]; }

  // dL (pqdts.f90:975-1032)
  void dL(const std::vector<int> &S, int x, int g) {
    struct Scope sc(this, "dL");
    reconstruct(S, x);
    tstart("dL_O");
    gemm_t<double, B_PLAIN, E_STORE_R, false, false>(S, GemmOps{BAND_F, x, OP_O, 0, 0, 0}, "spmm_F");
    tstop("dL_O");
    tstart("dL_dLdX");
    std::vector<int> R = reduced_of(S);
    if (!R.empty())
      launch(R, [&](int p) { return nb_rows(p); },
             [&](const WorkList &wl, int tot) { kb_foldval<T_GRAD><<<tot, BLOCK, 0, s>>>(dprobs, wl, x); });
    gemm_t<double, B_PLAIN, E_GRAD, false, false>(S, GemmOps{BAND_FT, OP_R, g, x, 0, 0}, "spmmT_F");
    tstop("dL_dLdX");
  }

  // BF (pqdts.f90:246-342): bx = H dd; with `dot` the partial dd.bx goes to S_V2.
  // p32: dd is the CG direction, whose FP32 expanded copy the CG kernels keep
  // (Prob::p32) for the mixed-precision problems of S.
  void BF(const std::vector<int> &S, int dd, int bx, bool have_ev, bool mask, bool dot, bool p32 = false) {
    struct Scope sc(this, "BF");
    need_maxr(S);
    std::vector<int> R = reduced_of(S);
    if (!have_ev && !R.empty())
      launch(R, [&](int p) { return nb_rows(p); },
             [&](const WorkList &wl, int tot) { kb_others<<<tot, BLOCK, 0, s>>>(dprobs, wl, dd); });
    std::vector<int> g64, g32, f64, f32;
    for (int p : S) {
      const bool gram = ctx[p]->use_gram();
      (gram ? (prod32(p) ? g32 : g64) : (prod32(p) ? f32 : f64)).push_back(p);
    }
    // Gram path: fold values, then one product 2 G p~ with the epilogue.
    // Problems with a split product write partials instead; their summing pass
    // forms the fold values itself. Problems whose fold is left to kb_cg_update
    // (late) need no fold values either.
    std::vector<int> g64s, g32s;
    auto split_off = [&](std::vector<int> &v, std::vector<int> &sv) {
      std::vector<int> keep;
      for (int p : v) (ctx[p]->hs > 1 && (!prod32(p) || ctx[p]->hs_pref > 0) ? sv : keep).push_back(p);
      v.swap(keep);
    };
    split_off(g64, g64s);
    split_off(g32, g32s);
    std::vector<int> G = g64;
    G.insert(G.end(), g32.begin(), g32.end());
    std::vector<int> GR = reduced_of(early_of(G));
    if (!GR.empty()) {
      hipEvent_t e0 = kbegin();
      launch(GR, [&](int p) { return nb_rows(p); }, [&](const WorkList &wl, int tot) {
        kb_foldval_gram<<<tot, BLOCK, 0, s>>>(dprobs, wl, dd, KC, TILE);
      });
      // a G row band and the gathered column of p~ per row
      kend(e0, "foldval_gram", GR, [&](int p) { Work w; w.bytes = 16.0 * (double)ctx[p]->gram_nnz; return w; });
    }
    const GemmOps hg{BAND_G, dd, bx, dd, dd, R_B};
    // B = p with its implicit columns expanded (reduced problems)
    std::vector<int> GS;
    const bool any_gram = !g64.empty() || !g32.empty() || !g64s.empty() || !g32s.empty();
    if (any_gram) tstart("BF_O");
    hvp_gemm<double, B_EXPAND>(g64, hg, mask, dot, "hvp_G");
    if (p32)
      hvp_gemm<float, B_P32>(g32, hg, mask, dot, "hvp_G32");
    else
      hvp_gemm<float, B_EXPAND>(g32, hg, mask, dot, "hvp_G32");
    part_gemm<double, B_EXPAND>(g64s, hg, "hvp_Gs");
    if (p32)
      part_gemm<float, B_P32>(g32s, hg, "hvp_G32s");
    else
      part_gemm<float, B_EXPAND>(g32s, hg, "hvp_G32s");
    GS = g64s;
    GS.insert(GS.end(), g32s.begin(), g32s.end());
    if (!GS.empty()) {
      hipEvent_t e0 = kbegin();
      launch(GS, [&](int p) { return nb_elem(p); }, [&](const WorkList &wl, int tot) {
        if (mask && dot)
          kb_hvp_epi<true, true><<<tot, BLOCK, 0, s>>>(dprobs, wl, dd, bx, R_B);
        else
          kb_hvp_epi<false, false><<<tot, BLOCK, 0, s>>>(dprobs, wl, dd, bx, R_B);
      });
      kend(e0, "hvp_epi", GS, [&](int p) { return stream_work(p, 8.0 * (ctx[p]->hs + 3) + 1.0); });
    }
    if (any_gram) tstop("BF_O");
    // F path: O = F p~, fold values, then 2 F^T O with the epilogue
    for (int pass = 0; pass < 2; pass++) {
      const std::vector<int> &FS = pass ? f32 : f64;
      if (FS.empty()) continue;
      tstart("BF_O");
      if (pass && p32)
        gemm_t<float, B_P32, E_STORE, false, false>(FS, GemmOps{BAND_F, dd, OP_O, 0, 0, 0}, "spmm_F32");
      else if (pass)
        gemm_t<float, B_EXPAND, E_STORE, false, false>(FS, GemmOps{BAND_F, dd, OP_O, 0, 0, 0}, "spmm_F32");
      else
        gemm_t<double, B_EXPAND, E_STORE, false, false>(FS, GemmOps{BAND_F, dd, OP_O, 0, 0, 0}, "spmm_F");
      tstop("BF_O");
      tstart("BF_BX");
      std::vector<int> FR = reduced_of(early_of(FS));
      if (!FR.empty())
        launch(FR, [&](int p) { return nb_rows(p); },
               [&](const WorkList &wl, int tot) { kb_foldval<T_HVP><<<tot, BLOCK, 0, s>>>(dprobs, wl, dd); });
      const GemmOps ht{BAND_FT, OP_O, bx, dd, dd, R_B};
      if (pass)
        hvp_gemm<float, B_PLAIN>(FS, ht, mask, dot, "spmmT_F32");
      else
        hvp_gemm<double, B_PLAIN>(FS, ht, mask, dot, "spmmT_F");
      tstop("BF_BX");
    }
  }
  // split Gram product: hs partial products per problem (nothing fused)
  template <typename T, int BM>
  void part_gemm(const std::vector<int> &S, GemmOps ops, const char *kname) {
    if (S.empty()) return;
    hipEvent_t e0 = kbegin();
    launch(S, [&](int p) { return nb_gemm(p, ops.band) * ctx[p]->hs; }, [&](const WorkList &wl, int tot) {
      launch_gemm<T, BM, E_PART, false, false>(tot, s, dprobs, wl, ops);
    });
    kend(e0, kname, S, [&](int p) { return gemm_work(p, ops.band, sizeof(T) == 4); });
  }
  template <typename T, int BM>
  void hvp_gemm(const std::vector<int> &S, GemmOps ops, bool mask, bool dot, const char *kname) {
    if (S.empty()) return;
    hipEvent_t e0 = kbegin();
    auto nb = [&](int p) { return nb_gemm(p, ops.band); };
    if (mask && dot) {
      // the v2 partials (region R_B) are summed by kb_cg_update
      launch(S, nb, [&](const WorkList &wl, int tot) {
        launch_gemm<T, BM, E_HVP, true, true>(tot, s, dprobs, wl, ops);
      });
    } else {
      launch(S, nb, [&](const WorkList &wl, int tot) {
        launch_gemm<T, BM, E_HVP, false, false>(tot, s, dprobs, wl, ops);
      });
    }
    kend(e0, kname, S, [&](int p) { return gemm_work(p, ops.band, sizeof(T) == 4); });
  }

  // pdot (pqdts.f90:589-614) -> S_DOT
  void pdot(const std::vector<int> &S, int x, int y) {
    struct Scope sc(this, "pdot");
    launch_reduce(S, [&](int p) { return nb_elem(p); },
                  [&](const WorkList &wl, int tot) { kb_dot<<<tot, BLOCK, 0, s>>>(dprobs, wl, x, y, R_A); }, R_A,
                  S_DOT);
  }

  // doproj / doproj2 (pqdts.f90:1120-1164) on operand t; with form_trial the
  // trial point t = arg*SN + X is formed first. xmin -> S_XMIN (ixmin < 0: the
  // xmin partials stay in region R_C with their count in pn, for L to reduce).
  void project(const std::vector<int> &S, int t, bool form_trial, bool recon_again, int ixmin = S_XMIN) {
    need_maxr(S);
    bool any_full = false, any_red = false;
    for (int p : S) (ctx[p]->full ? any_full : any_red) = true;
    struct Scope sc(this, any_red && !any_full ? "doproj2" : "doproj");
    hipEvent_t e0 = kbegin();
    auto proj = [&](const WorkList &wl, int tot) {
      kb_project<<<tot, BLOCK, 0, s>>>(dprobs, wl, t, form_trial ? 1 : 0, recon_again ? 1 : 0, R_C);
    };
    if (ixmin >= 0)
      launch_reduce(S, [&](int p) { return nb_rows(p); }, proj, R_C, ixmin, true);
    else
      launch(S, [&](int p) { return nb_rows(p); }, proj);
    kend(e0, "project", S, [&](int p) { return stream_work(p, form_trial ? 24.0 : 16.0); });
  }

  // conv_test (pqdts.f90:1034-1067): sets full space; -> S_CONV
  void conv(const std::vector<int> &S, int x, int g, bool store) {
    reconstruct(S, x);
    for (int p : S) ctx[p]->full = true;
    dL(S, x, g);
    hipEvent_t e0 = kbegin();
    launch_reduce(S, [&](int p) { return nb_rows(p); }, [&](const WorkList &wl, int tot) {
      kb_conv_rows<<<tot, BLOCK, 0, s>>>(dprobs, wl, x, g, store ? 1 : 0, R_C);
    }, R_C, S_CONV);
    kend(e0, "conv_rows", S, [&](int p) { return stream_work(p, store ? 24.0 : 16.0); });
  }
  double convval(int p) {
    return std::sqrt(host_scal(p)[S_CONV] / ((double)ctx[p]->N * (double)ctx[p]->M));
  }

  void axpby(const std::vector<int> &S, int z, int x, int y, double alpha, double beta, int kind,
             bool per_problem_alpha) {
    hipEvent_t e0 = kbegin();
    launch(S, [&](int p) { return nb_elem(p); }, [&](const WorkList &wl, int tot) {
      kb_axpby<<<tot, BLOCK, 0, s>>>(dprobs, wl, z, x, y, alpha, beta, kind, per_problem_alpha ? 1 : 0);
    });
    kend(e0, "axpby", S, [&](int p) { return stream_work(p, 24.0); });
  }

  void transform(const std::vector<int> &S, bool rebuild) {
    launch(S, [&](int p) { return nb_rows(p); },
           [&](const WorkList &wl, int tot) { kb_transform<<<tot, BLOCK, 0, s>>>(dprobs, wl, rebuild ? 1 : 0); });
    for (int p : S) ctx[p]->maxr_valid = !ctx[p]->full;
  }

  // ---- Newton-CG pieces (pqdts.f90:1986-2104)
  void cg_init(const std::vector<int> &S) {
    struct Scope sc(this, "active_dLdX");
    need_maxr(S);
    hipEvent_t e0 = kbegin();
    launch(S, [&](int p) { return nb_rows(p); }, [&](const WorkList &wl, int tot) {
      kb_cg_init<<<tot, BLOCK, 0, s>>>(dprobs, wl, 1);
      kb_cg_start<<<wl.np, BLOCK, 0, s>>>(dprobs, wl, S_NACT, S_ND, S_B1, S_EPSCG, S_CGACT, S_CGN);
    });
    // reads X, G; writes the mask, r, p, sn
    kend(e0, "cg_init", S, [&](int p) { return stream_work(p, 41.0); });
  }
  // CG round j: update (after H p with its v2 partials), then the direction,
  // the round's log and the exit test
  void cg_update(const std::vector<int> &S, int i_b1) {
    hipEvent_t e0 = kbegin();
    launch(S, [&](int p) { return nb_rows(p); }, [&](const WorkList &wl, int tot) {
      kb_cg_update<<<tot, BLOCK, 0, s>>>(dprobs, wl, i_b1, R_B, S_V2);
    });
    // reads sn, p, r, Hp; writes sn, r
    kend(e0, "cg_update", S, [&](int p) { return stream_work(p, 48.0); });
  }
  void cg_dir(const std::vector<int> &S, int i_b1, int i_rz, int j) {
    hipEvent_t e0 = kbegin();
    launch(S, [&](int p) { return nb_rows(p); }, [&](const WorkList &wl, int tot) {
      kb_cg_dir<<<tot, BLOCK, 0, s>>>(dprobs, wl, i_b1, i_rz, S_RR, j, S_EPSCG, S_CGACT, S_CGN);
    });
    // reads p, r; writes p
    kend(e0, "cg_dir", S, [&](int p) { return stream_work(p, 24.0); }, true);
  }
  void twometric(const std::vector<int> &S) {
    launch(S, [&](int p) { return nb_rows(p); },
           [&](const WorkList &wl, int tot) { kb_twometric<<<tot, BLOCK, 0, s>>>(dprobs, wl); });
  }

  struct Scope {
    Engine *e;
    const char *n;
    Scope(Engine *e_, const char *n_) : e(e_), n(n_) { e->tstart(n); }
    ~Scope() { e->tstop(n); }
  };

  // ---- solver (defined below)
  void line_search(const std::vector<int> &A);
  void tick();
  bool any_active() const;
};

// Per-problem progress of the minimisation (the reference's loop variables).
struct PState {
  int mode = 2, maxiter = 0, start_stage = 1, output = 0;
  double eps = 0.0;
  pqdts_event_cb cb = nullptr;
  void *user = nullptr;
  int phase = 0;  // 0 initial evaluation pending, 1 iterating, 2 done
  int stage = 1, iiter = 0;
  bool stage_exit = false;
  double dlalpha = 0.0, alpha = 0.0, conv = 0.0;
  // line search
  double ls_L0 = 0.0, ls_df = 0.0, ls_alpha = 0.0, ls_lm = 0.0, ls_amin = 0.0;
  // counters
  int newton = 0, pcg = 0, trials = 0;
  // polish (phase 3): settings copied from the context, then the iteration state
  pqdts_ctx::Polish pol;
  int pol_it = 0;         // outer iterations started
  int pol_cgr = 0;        // CG rounds of the current outer iteration
  int64_t pol_budget = 0; // CG rounds the polish may use (0: no limit)
  int pol_rounds = 0;     // CG rounds of all outer iterations
  bool pol_incg = false;  // the current outer iteration's CG continues next step
  bool pol_entered = false;
  double pol_rtol = 1.0;  // relative CG residual of the last solve
  double pol_step = 0.0;  // largest entry of the last step
  int pol_nchg = 0;
  int64_t pol_nact = 0;
  int pol_status = 0;
  void emit(int kind, std::initializer_list<double> v) {
    if (cb) {
      std::vector<double> t(v);
      cb(user, kind, t.data(), (int)t.size());
    }
  }
};

Engine::~Engine() {
  if (s) (void)hipStreamSynchronize(s);
  drop_graphs();
  for (auto &p : kpending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (hipEvent_t e : kpool) (void)hipEventDestroy(e);
  for (int p = 0; p < cap; p++) face_free(p);
  release_deferred();
  if (s) (void)hipStreamSynchronize(s);
  if (hface) (void)hipHostFree(hface);
  if (dfaces) (void)hipFree(dfaces);
  if (dprobs) (void)hipFree(dprobs);
  if (dscal) (void)hipFree(dscal);
  if (hscal) (void)hipHostFree(hscal);
  if (dlog) (void)hipFree(dlog);
  if (hlog) (void)hipHostFree(hlog);
  if (own_stream && s) (void)hipStreamDestroy(s);
}

bool Engine::polishing(int p) const { return st[p] && st[p]->phase == 3; }
bool Engine::prod32(int p) const { return fp32[p] && !(polishing(p) && !st[p]->pol.fp32); }

bool Engine::any_active() const {
  for (int p = 0; p < cap; p++)
    if (ctx[p] && st[p] && st[p]->phase != 2) return true;
  return false;
}

// Lalpha_min_auglag (stage 1, pqdts.f90:440-483) for full-space problems and
// Lalpha_min2 (pqdts.f90:394-438) for reduced ones, in lockstep: every round
// evaluates one trial point for each problem still searching.
// Lalpha_min2 / Lalpha_min_auglag (pqdts.f90:394-483) for every problem of A. The
// trials of a round run back to back on the device, each into its own scalar
// slots, and are read with one host round trip; the host walks them in the
// reference's order and stops at the first accepted one, so trial values, counts
// and events do not depend on the round size. The two evaluations at alpha = 0
// and dx share a round; the Armijo trials go one per round (LS_ROUND): evaluating
// more ahead cost the batched mix more GPU time than the round trips it saved.
void Engine::line_search(const std::vector<int> &A) {
  Scope sc(this, "Lalpha_min");
  const double dx = (double)1e-5f;      // pqdts.f90:354/404/450
  const double feas = (double)-1e-12f;  // pqdts.f90:424/430
  // trials k = 0..K-1 of every problem of S at step lengths alpha(p, k)
  auto trials = [&](const std::vector<int> &S, int K, const std::function<double(int, int)> &alpha) {
    for (int k = 0; k < K; k++) {
      for (int p : S) arg[p] = alpha(p, k);
      project(S, PQDTS_XTMP, true, true, -1);
      // the reduced-space prep was applied by kb_project; xmin is reduced with L
      L(S, PQDTS_XTMP, false, S_LT + k, S_LTB + k, S_LTX + k);
    }
    fetch();
  };
  auto Lk = [&](int p, int k) { return host_scal(p)[S_LT + k] + host_scal(p)[S_LTB + k]; };
  // L at alpha = 0 and at alpha = dx (the slope probe) as one round
  trials(A, 2, [&](int, int k) { return k == 0 ? 0.0 : 0.0 + dx; });
  for (int p : A) {
    PState &q = *st[p];
    q.ls_L0 = Lk(p, 0);
    q.ls_df = (Lk(p, 1) - q.ls_L0) / dx;
    q.dlalpha = q.ls_df;
    q.ls_alpha = 1.0;
    q.ls_lm = std::numeric_limits<double>::max();
    q.ls_amin = 0.0;
  }
  std::vector<int> S = A;
  constexpr int LS_ROUND = 1;
  for (int i = 1; i <= 100 && !S.empty(); i += LS_ROUND) {
    const int Kr = std::min(LS_ROUND, 101 - i);
    // alpha of trial k: the current one times 0.75^k, multiplied step by step as the
    // reference does
    trials(S, Kr, [&](int p, int k) {
      double a = st[p]->ls_alpha;
      for (int t = 0; t < k; t++) a = a * 0.75;
      return a;
    });
    std::vector<int> next;
    for (int p : S) {
      PState &q = *st[p];
      const bool min2 = !ctx[p]->full;
      bool accepted = false;
      for (int k = 0; k < Kr && !accepted; k++) {
        const double Lp = Lk(p, k), xmin = min2 ? host_scal(p)[S_LTX + k] : 0.0;
        const bool feasible = !min2 || xmin >= feas;
        q.trials++;
        if (Lp < q.ls_lm && feasible) {
          q.ls_lm = Lp;
          q.ls_amin = q.ls_alpha;
        }
        q.emit(PQDTS_EV_LS, {(double)(i + k), Lp, q.ls_alpha, q.ls_df, xmin, min2 ? 1.0 : 0.0});
        if (Lp > q.ls_L0 + 0.1 * q.ls_alpha * q.ls_df || !feasible)
          q.ls_alpha = q.ls_alpha * 0.75;
        else
          accepted = true;
      }
      if (!accepted) next.push_back(p);
    }
    S.swap(next);
  }
  for (int p : A) st[p]->alpha = st[p]->ls_amin;
}

// One Newton iteration (or the initial evaluation / stage switches) for every
// problem of the engine, pqdts.f90:1908-2167.
// corrections (face solves from the current point) per iteration of the direct polish
constexpr int kFacePasses = 2;

// Direct polish (polish method 1; no counterpart in the reference): one
// primal-dual active-set iteration per step for the problems QD, in the full space
// with the row sums as constraints. On entry the minimisation's result is kept in
// Z and the multipliers start at -G at each row's largest entry. Each iteration
// takes the active set from mu - d X > -theta d (mu = G + lambda) and, unless it repeats
// the face just solved, forms the inverses of K on every column's free rows, the
// Schur complement of the row sums and its band Cholesky factor, then applies
// kFacePasses corrections from the current point, each with a fresh gradient.
// Finished problems go to Fz with their status: 1 when the active set repeats
// after a last correction of at most 100 tol, 2 at the iteration cap, for a
// problem larger than the method handles (M > FACE_MAX_ROWS or a column with more
// than FACE_MAX_M free rows), or when the factorisation fails.
void Engine::face_step(const std::vector<int> &QD, std::vector<int> &Fz) {
  Scope sc(this, "polish");
  const int X = PQDTS_X, G = PQDTS_G, Y = PQDTS_HP;
  // launches timed like the operators' (HIP events, when kernel timing is on)
  auto timed = [&](const char *name, const std::vector<int> &S, auto nb, auto fn) {
    hipEvent_t e0 = kbegin();
    launch(S, nb, fn);
    kend(e0, name, S, [](int) { return Work{}; });
  };
  std::vector<int> ent, init, Qa;
  for (int p : QD)
    if (!st[p]->pol_entered) {
      st[p]->pol_entered = true;
      ent.push_back(p);
    }
  // the minimisation's result (full space): the result if the polish stops at a cap
  if (!ent.empty()) axpby(ent, PQDTS_Z, X, -1, 1.0, 0.0, AX_COPY, false);
  for (int p : ent) {
    if (ctx[p]->M > FACE_MAX_ROWS || !ctx[p]->full) {
      st[p]->pol_status = 2;
      Fz.push_back(p);
    } else {
      if (hface[p].pos)
        HIPC(hipMemsetAsync(hface[p].cdirty, 1, (size_t)ctx[p]->N, s));
      else
        face_alloc(p);
      init.push_back(p);
    }
  }
  if (!init.empty()) face_upload_all();
  if (!init.empty()) {
    dL(init, X, G);
    timed("kf_lam0", init, [&](int p) { return nb_rows(p); }, [&](const WorkList &wl, int tot) {
      kf_lam0<<<tot, BLOCK, 0, s>>>(dprobs, dfaces, wl);
    });
  }
  for (int p : QD)
    if (st[p]->pol_status == 0) Qa.push_back(p);
  if (Qa.empty()) return;
  for (int p : Qa) face_swap(p);
  face_upload_all();
  timed("kf_active", Qa, [&](int p) { return nb_rows(p); }, [&](const WorkList &wl, int tot) {
    kf_active<<<tot, BLOCK, 0, s>>>(dprobs, dfaces, wl);
    kb_reduce_pn<<<wl.np, BLOCK, 0, s>>>(dprobs, wl, 2, R_A, S_NCHG, R_B, S_NACT, 0, 0, 0);
  });
  timed("kf_cols", Qa, [&](int p) { return cdiv(ctx[p]->N, BLOCK); }, [&](const WorkList &wl, int tot) {
    kf_cols<<<tot, BLOCK, 0, s>>>(dprobs, dfaces, wl);
  });
  timed("kf_offsets", Qa, [](int) { return 1; }, [&](const WorkList &wl, int tot) {
    kf_offsets<<<tot, BLOCK, 0, s>>>(dprobs, dfaces, wl, S_FMAX, S_FTOT);
  });
  fetch();
  std::vector<int> go;
  for (int p : Qa) {
    PState &q = *st[p];
    const double *h = host_scal(p);
    q.pol_nchg = (int)h[S_NCHG];
    q.pol_nact = (int64_t)h[S_NACT];
    if (q.pol_it > 0 && q.pol_nchg == 0 && q.pol_step <= 100.0 * q.pol.tol) {
      q.pol_status = 1;
      Fz.push_back(p);
    } else if (q.pol_it >= q.pol.max_outer || h[S_FMAX] > FACE_MAX_M) {
      q.pol_status = 2;
      Fz.push_back(p);
    } else {
      q.pol_it++;
      face_grow(p, (int64_t)h[S_FTOT]);
      go.push_back(p);
    }
  }
  if (go.empty()) return;
  face_upload_all();
  timed("kf_factor", go, [&](int p) { return ctx[p]->N; }, [&](const WorkList &wl, int tot) {
    kf_factor<FACE_LDS_S><<<tot, BLOCK, 0, s>>>(dprobs, dfaces, wl);
    kf_factor<FACE_LDS_L><<<tot, BLOCK, 0, s>>>(dprobs, dfaces, wl);
    kf_factor_big<<<tot, BLOCK, 0, s>>>(dprobs, dfaces, wl);
  });
  timed("kf_schur", go, [&](int p) { return ctx[p]->M; }, [&](const WorkList &wl, int tot) {
    kf_schur<<<tot, BLOCK, 0, s>>>(dprobs, dfaces, wl);
  });
  timed("kf_chol", go, [](int) { return 1; }, [&](const WorkList &wl, int tot) {
    kf_chol<<<tot, BLOCK, 0, s>>>(dprobs, dfaces, wl, S_FBW, S_FFAIL);
  });
  for (int pass = 0; pass < kFacePasses; pass++) {
    dL(go, X, G);
    timed("kf_y", go, [&](int p) { return ctx[p]->N; }, [&](const WorkList &wl, int tot) {
      kf_y<<<tot, BLOCK, 0, s>>>(dprobs, dfaces, wl, Y);
    });
    timed("kf_rhs", go, [&](int p) { return nb_rows(p); }, [&](const WorkList &wl, int tot) {
      kf_rhs<<<tot, BLOCK, 0, s>>>(dprobs, dfaces, wl, Y);
    });
    timed("kf_solve", go, [](int) { return 1; }, [&](const WorkList &wl, int tot) {
      kf_solve<<<tot, BLOCK, 0, s>>>(dprobs, dfaces, wl, S_FBW);
    });
    timed("kf_update", go, [&](int p) { return ctx[p]->N; }, [&](const WorkList &wl, int tot) {
      kf_update<<<tot, BLOCK, 0, s>>>(dprobs, dfaces, wl, Y);
    });
    timed("kf_stepmax", go, [](int) { return 1; }, [&](const WorkList &wl, int tot) {
      kf_stepmax<<<tot, BLOCK, 0, s>>>(dprobs, dfaces, wl, S_STEP);
    });
  }
  dL(go, X, G);  // the gradient at the new point, for the next active set
  fetch();
  for (int p : go) {
    PState &q = *st[p];
    const double *h = host_scal(p);
    q.pol_step = -h[S_STEP];
    q.emit(PQDTS_EV_POLISH, {(double)q.pol_it, (double)q.pol_nact, (double)q.pol_nchg, 0.0, q.pol_step,
                             h[S_FMAX], h[S_FBW], h[S_FFAIL]});
    if (h[S_FFAIL] != 0.0 || !std::isfinite(q.pol_step)) {
      q.pol_status = 2;
      Fz.push_back(p);
    }
  }
}

void Engine::tick() {
  const int X = PQDTS_X, G = PQDTS_G, HP = PQDTS_HP, XT = PQDTS_XTMP, SN = PQDTS_SN;
  std::vector<int> init;
  for (int p = 0; p < cap; p++)
    if (ctx[p] && st[p] && st[p]->phase == 0) init.push_back(p);
  if (!init.empty()) {
    // pqdts.f90:1908-1915 (full space, as set at program start)
    L(init, X, true);
    dL(init, X, G);
    pdot(init, G, G);
    fetch();
    for (int p : init) {
      PState &q = *st[p];
      q.emit(PQDTS_EV_ITER0, {Lval(p), std::sqrt(host_scal(p)[S_DOT])});
      q.phase = 1;
      q.stage = q.start_stage;
      q.iiter = 0;
      q.dlalpha = -std::numeric_limits<double>::max();
      q.stage_exit = false;
    }
  }
  // stage control: exits at the top of an iteration and at maxiter
  for (;;) {
    std::vector<int> E, Elam;
    for (int p = 0; p < cap; p++) {
      if (!ctx[p] || !st[p] || st[p]->phase != 1) continue;
      PState &q = *st[p];
      bool end = q.stage_exit || q.iiter >= q.maxiter;
      if (!end && q.mode == 2 && q.stage == 1 && (std::fabs(q.dlalpha) < (double)1e-4f || q.dlalpha > 0)) {
        q.emit(PQDTS_EV_STOP1, {q.dlalpha});  // pqdts.f90:1955-1957
        end = true;
      }
      if (end) (q.output >= 3 ? Elam : E).push_back(p);
    }
    if (E.empty() && Elam.empty()) break;
    // end of stage: conv_test (pqdts.f90:2139) and the per-stage output event
    if (!E.empty()) conv(E, X, G, false);
    if (!Elam.empty()) conv(Elam, X, G, true);
    fetch();
    std::vector<int> both = E;
    both.insert(both.end(), Elam.begin(), Elam.end());
    for (int p : both) {
      PState &q = *st[p];
      q.conv = convval(p);
      if (q.output >= 3 && q.cb) {
        std::vector<double> lam(ctx[p]->M), lmin(ctx[p]->M);
        HIPC(hipMemcpy(lam.data(), ctx[p]->lam, sizeof(double) * ctx[p]->M, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(lmin.data(), ctx[p]->lmin, sizeof(double) * ctx[p]->M, hipMemcpyDeviceToHost));
        for (int i = 0; i < ctx[p]->M; i++) q.emit(PQDTS_EV_LAMBDA, {(double)(i + 1), lam[i], lmin[i]});
      }
      q.emit(PQDTS_EV_STAGE, {(double)q.stage, q.conv});
      q.stage++;
      if (q.stage > 2) {
        q.phase = q.pol.max_outer > 0 ? 3 : 2;
      } else {
        q.iiter = 0;
        q.dlalpha = -std::numeric_limits<double>::max();
        q.stage_exit = false;
      }
    }
  }
  std::vector<int> A, A1, A2;
  std::vector<int> Q;  // polishing
  for (int p = 0; p < cap; p++) {
    if (!ctx[p] || !st[p]) continue;
    if (st[p]->phase == 3) {
      Q.push_back(p);
      continue;
    }
    if (st[p]->phase != 1) continue;
    A.push_back(p);
    PState &q = *st[p];
    q.iiter++;
    q.newton++;
    if (q.mode == 1) {
      ctx[p]->full = false;
      A1.push_back(p);
    } else {
      ctx[p]->full = (q.stage == 1);
      A2.push_back(p);
    }
  }
  if (A.empty() && Q.empty()) return;
  // PQDTS_GRAPHS=2: with every problem in the Newton-CG of mode 2 (none polishing),
  // the launches before the line search and those after the step each replay a
  // captured graph; the line search stays host-driven. The launches are the eager
  // ones, except that all 50 CG rounds are launched (problems whose CG has stopped
  // skip them at the guard) instead of leaving at a host check every 8 rounds.
  const bool whole = graphs_on == 2 && !private_scal && !ktiming && !dev_timers && Q.empty() && A1.empty() &&
                     !A2.empty();
  if (!A2.empty() && !whole) transform(A2, true);   // pqdts.f90:1970-1983
  if (!A1.empty()) transform(A1, false);  // pqdts.f90:1940-1944
  std::vector<int> Fz;  // polish: finished in this step
  {
    // the direct polish runs one active-set iteration per step on its own
    std::vector<int> QD, QK;
    for (int p : Q) (st[p]->pol.method == 1 ? QD : QK).push_back(p);
    if (!QD.empty()) face_step(QD, Fz);
    Q.swap(QK);
  }
  // Polish (no counterpart in the reference): problems starting an outer iteration
  // (Qs; on entry the reduced-space transform of the minimisation's result),
  // problems whose CG continues from the last step (Qc)
  std::vector<int> Qs, Qc, Qent;
  for (int p : Q) {
    PState &q = *st[p];
    (q.pol_incg ? Qc : Qs).push_back(p);
    if (!q.pol_entered) {
      q.pol_entered = true;
      q.pol_budget = q.pol.budget > 0.0 ? std::max<int64_t>(1, (int64_t)(q.pol.budget * q.pcg)) : 0;
      Qent.push_back(p);
    }
  }
  if (!Qent.empty()) {
    // the minimisation's result (full space) is kept in Z: the result if the polish
    // stops at a cap
    axpby(Qent, PQDTS_Z, X, -1, 1.0, 0.0, AX_COPY, false);
    for (int p : Qent) ctx[p]->full = false;
    transform(Qent, true);
  }
  std::vector<int> AQ = whole ? std::vector<int>() : A;
  AQ.insert(AQ.end(), Qs.begin(), Qs.end());
  if (!AQ.empty()) dL(AQ, X, G);
  std::vector<int> Qn;  // polish: new outer iteration
  if (!Qs.empty()) {
    Scope sc(this, "polish");
    launch(Qs, [&](int p) { return nb_rows(p); }, [&](const WorkList &wl, int tot) {
      kb_pdas_active<<<tot, BLOCK, 0, s>>>(dprobs, wl);
      kb_reduce_pn<<<wl.np, BLOCK, 0, s>>>(dprobs, wl, 2, R_A, S_NCHG, R_B, S_NACT, 0, 0, 0);
    });
    fetch();
    for (int p : Qs) {
      PState &q = *st[p];
      const double *h = host_scal(p);
      q.pol_nchg = (int)h[S_NCHG];
      q.pol_nact = (int64_t)h[S_NACT];
      // converged: the active set repeats (the last solve was on this same face),
      // that solve's step was a correction of at most 100 tol, and its own error,
      // about its relative residual times that step, is within the target (the
      // estimate alone can be off by the conditioning: a larger last step needs
      // one more, loose, solve)
      if (q.pol_it > 0 && q.pol_nchg == 0 && q.pol_step <= 100.0 * q.pol.tol &&
          q.pol_rtol * q.pol_step <= q.pol.tol) {
        q.pol_status = 1;
        Fz.push_back(p);
      } else if (q.pol_it >= q.pol.max_outer || (q.pol_budget > 0 && q.pol_rounds >= q.pol_budget)) {
        q.pol_status = 2;
        Fz.push_back(p);
      } else {
        // each step is solved to the relative residual that makes its error about
        // the target, given the last step's size: tight while the steps are large
        // (an inexact step there can make the active set diverge), looser once
        // only a small correction is left
        q.pol_it++;
        q.pol_cgr = 0;
        q.pol_incg = true;
        const double want = q.pol_it == 1 ? q.pol.cg_rtol : q.pol.tol / std::max(q.pol_step, 1e-300);
        q.pol_rtol = std::min(kPolishLooseRtol, std::max(q.pol.cg_rtol, want));
        arg[p] = q.pol_rtol;
        Qn.push_back(p);
      }
    }
    // X changed on the new active set
    if (!Qn.empty()) dL(Qn, X, G);
  }
  std::vector<int> CG0 = A2;  // CG from a fresh start
  CG0.insert(CG0.end(), Qn.begin(), Qn.end());
  std::vector<int> Qstep;     // polish problems whose CG ended in this step
  // round j of the PCG: r.z of the previous round is in S_B1 (j = 1), else S_RZ0 / S_RZ1
  auto ib1_of = [](int j) { return j == 1 ? (int)S_B1 : ((j - 1) & 1) ? (int)S_RZ1 : (int)S_RZ0; };
  if (whole) {
    // transform, gradient, active set, PCG (pqdts.f90:1970-2090), gradient and the
    // two-metric step (2093-2102) as one graph
    Scope sc(this, "CG");
    graph_run(A2, -1, 0, [&] {
      transform(A2, true);
      dL(A2, X, G);
      cg_late = true;
      cg_init(A2);
      guard = S_CGACT;
      for (int j = 1; j <= 50; j++) {
        const int ib1 = ib1_of(j), i_rz = (j & 1) ? S_RZ1 : S_RZ0;
        BF(A2, XT, HP, true, true, true, true);
        cg_update(A2, ib1);
        cg_dir(A2, ib1, i_rz, j);
      }
      guard = -1;
      cg_late = false;
      dL(A2, X, G);
      twometric(A2);
    });
    for (int p : A2) ctx[p]->maxr_valid = !ctx[p]->full;  // what transform records on the host
  }
  if (!CG0.empty() || !Qc.empty()) {
    if (!whole) cg_late = true;
    if (!CG0.empty() && !whole) cg_init(CG0);  // pqdts.f90:1986-2031
    if (!Qc.empty())
      launch(Qc, [](int) { return 1; }, [&](const WorkList &wl, int) {
        kb_cg_carry<<<wl.np, 64, 0, s>>>(dprobs, wl, S_CGN, S_RZ0, S_RZ1, S_B1);
      });
    // PCG, pqdts.f90:2036-2090. The exit test runs on the device: every CG launch
    // skips problems whose flag dropped, and the host only checks every few rounds.
    // A polish CG runs at most 50 rounds per step too and continues in the next.
    if (!whole) {
      Scope sc(this, "CG");
      guard = S_CGACT;
      // problems still in CG; those whose CG stopped leave the launches at the
      // next host check (their blocks would only exit at the guard)
      std::vector<int> C = CG0;
      C.insert(C.end(), Qc.begin(), Qc.end());
      auto rounds = [&](int j0, int j1) {
        for (int j = j0; j <= j1; j++) {
          kround = j;
          const int ib1 = ib1_of(j), i_rz = (j & 1) ? S_RZ1 : S_RZ0;
          BF(C, XT, HP, true, true, true, true);
          cg_update(C, ib1);
          cg_dir(C, ib1, i_rz, j);
        }
      };
      const bool graph = graphs_on == 1 && !private_scal && !ktiming && !dev_timers;
      for (int j0 = 1; j0 <= 50; j0 += 8) {
        const int j1 = std::min(j0 + 7, 50);
        if (graph)
          graph_rounds(C, j0, j1, rounds);
        else
          rounds(j0, j1);
        fetch();
        std::vector<int> next;
        for (int p : C)
          if (host_scal(p)[S_CGACT] != 0.0) next.push_back(p);
        if (next.empty()) break;
        C.swap(next);
      }
      guard = -1;
      kround = 0;
    }
    cg_late = false;
    fetch();
    kresolve();
    if (!A2.empty()) {
      HIPC(hipMemcpyAsync(hlog, dlog, sizeof(double) * cap * CG_LOG_ROUNDS * 5, hipMemcpyDeviceToHost, s));
      HIPC(hipStreamSynchronize(s));
    }
    for (int p : A2) {
      PState &q = *st[p];
      const double *h = host_scal(p);
      q.emit(PQDTS_EV_ACTIVE, {h[S_NACT], (double)ctx[p]->MN});
      const int nr = (int)h[S_CGN];
      const double *lg = hlog + (int64_t)p * CG_LOG_ROUNDS * 5;
      for (int r = 0; r < nr; r++) q.emit(PQDTS_EV_PCG, {lg[5 * r], lg[5 * r + 1], lg[5 * r + 2], lg[5 * r + 3],
                                                         lg[5 * r + 4]});
      q.pcg += nr;
    }
    std::vector<int> QC = Qn;
    QC.insert(QC.end(), Qc.begin(), Qc.end());
    for (int p : QC) {
      PState &q = *st[p];
      const double *h = host_scal(p);
      const int nr = (int)h[S_CGN];
      q.pol_cgr += nr;
      q.pol_rounds += nr;
      if (h[S_CGACT] == 0.0 || (q.pol_budget > 0 && q.pol_rounds >= q.pol_budget)) {
        q.pol_incg = false;
        Qstep.push_back(p);
      }
    }
    if (!A2.empty() && !whole) {
      dL(A2, X, G);    // pqdts.f90:2093
      twometric(A2);   // pqdts.f90:2095-2102
    }
  }
  // polish: the full Newton step and its largest entry
  if (!Qstep.empty()) {
    Scope sc(this, "polish");
    launch_reduce(Qstep, [&](int p) { return nb_rows(p); }, [&](const WorkList &wl, int tot) {
      kb_pdas_step<<<tot, BLOCK, 0, s>>>(dprobs, wl, R_C);
    }, R_C, S_STEP, true);
  }
  if (!A1.empty()) axpby(A1, SN, G, -1, -1.0, 0.0, AX_SCALE, false);  // pqdts.f90:1950
  if (!A.empty()) line_search(A);
  // step, projection and re-evaluation, pqdts.f90:2111-2128; a finished polish
  // projects its result (clip + implicit column) and is evaluated the same way
  std::vector<int> stepS;
  for (int p : A)
    if (st[p]->alpha != 0.0) {
      arg[p] = st[p]->alpha;
      stepS.push_back(p);
    }
  if (!stepS.empty()) axpby(stepS, X, SN, X, 0.0, 0.0, AX_AXPY, true);
  // a polish stopped at a cap returns the minimisation's result unchanged
  // AD: certified direct polishes, whose result is feasible to rounding: clipped at
  // 0 instead of projected (a projection would spread the rows' rounding-level sum
  // error over every entry, zeros included)
  std::vector<int> Fr, AF = A, AD;
  for (int p : Fz) (st[p]->pol_status == 1 ? (st[p]->pol.method == 1 ? AD : AF) : Fr).push_back(p);
  if (whole) {
    graph_run(A, -2, 0, [&] {
      project(A, X, false, false);
      L(A, X, true);
      dL(A, X, G);
      pdot(A, G, G);
      conv(A, X, G, false);
    });
    for (int p : A) ctx[p]->full = true;  // what conv records on the host
    AF.clear();
  }
  if (!AF.empty()) project(AF, X, false, false);
  if (!AD.empty()) {
    launch(AD, [&](int p) { return nb_rows(p); }, [&](const WorkList &wl, int tot) {
      kf_clip<<<tot, BLOCK, 0, s>>>(dprobs, wl);
    });
    AF.insert(AF.end(), AD.begin(), AD.end());
  }
  if (!Fr.empty()) {
    axpby(Fr, X, PQDTS_Z, -1, 1.0, 0.0, AX_COPY, false);
    for (int p : Fr) ctx[p]->full = true;
    AF.insert(AF.end(), Fr.begin(), Fr.end());
  }
  if (!AF.empty()) {
    L(AF, X, true);
    dL(AF, X, G);
    pdot(AF, G, G);
    conv(AF, X, G, false);
  }
  fetch();
  for (int p : A) {
    PState &q = *st[p];
    const double *h = host_scal(p);
    q.conv = convval(p);
    // xmin of doproj2 (0 where Condat projected: its rows write 0 partials)
    q.emit(PQDTS_EV_ITER, {(double)q.stage, (double)q.iiter, Lval(p), std::fmin(0.0, h[S_XMIN]), q.conv,
                           std::sqrt(h[S_DOT]), (double)q.mode, q.alpha, q.dlalpha});
    if (q.conv < q.eps) q.stage_exit = true;
  }
  for (int p : Qstep) {
    PState &q = *st[p];
    q.pol_step = -host_scal(p)[S_STEP];
    q.emit(PQDTS_EV_POLISH, {(double)q.pol_it, (double)q.pol_nact, (double)q.pol_nchg, (double)q.pol_cgr,
                             q.pol_step});
  }
  for (int p : Fz) {
    PState &q = *st[p];
    q.conv = convval(p);
    q.phase = 2;
    face_free(p);
    pqdts_ctx::PolishOut &o = ctx[p]->pol_out;
    o.status = q.pol_status;
    o.iters = q.pol_it;
    o.rounds = q.pol_rounds;
    o.nact = q.pol_nact;
    o.step = q.pol_step;
    q.emit(PQDTS_EV_POLISHED, {(double)q.pol_status, (double)q.pol_it, (double)q.pol_rounds, q.conv, Lval(p)});
  }
}

// ===========================================================================
// context creation
// ===========================================================================

// Packs the rows of a banded matrix into per-tile dense transposed bands for
// kb_band_gemm: tile t covers rows [t*TILE, (t+1)*TILE); its band starts at
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

void upload_band(pqdts_ctx *c, pqdts_ctx::DBand &dst, const HostBand &h) {
  dst.ntiles = (int)h.off.size();
  dst.elems = (int64_t)h.pack.size();
  dst.pack = dalloc<double>(dst.elems);
  dst.off = dalloc<int64_t>(dst.ntiles);
  dst.klo = dalloc<int>(dst.ntiles);
  dst.nch = dalloc<int>(dst.ntiles);
  HIPC(hipMemcpyAsync(dst.pack, h.pack.data(), sizeof(double) * h.pack.size(), hipMemcpyHostToDevice, c->s));
  HIPC(hipMemcpyAsync(dst.off, h.off.data(), sizeof(int64_t) * dst.ntiles, hipMemcpyHostToDevice, c->s));
  HIPC(hipMemcpyAsync(dst.klo, h.klo.data(), sizeof(int) * dst.ntiles, hipMemcpyHostToDevice, c->s));
  HIPC(hipMemcpyAsync(dst.nch, h.nch.data(), sizeof(int) * dst.ntiles, hipMemcpyHostToDevice, c->s));
  HIPC(hipStreamSynchronize(c->s));  // host vectors die with the caller
}

void free_band(pqdts_ctx::DBand &b) {
  void *p[] = {b.pack, b.pack32, b.off, b.klo, b.nch};
  for (void *q : p)
    if (q) (void)hipFree(q);
  b = pqdts_ctx::DBand{};
}

// FP32 copies of the band packs used by the mixed-precision Hessian products.
void ensure_fp32(pqdts_ctx *c) {
  for (pqdts_ctx::DBand *b : {&c->bf, &c->bt, &c->bg}) {
    if (!b->pack || b->pack32) continue;
    b->pack32 = dalloc<float>(b->elems);
    k_round_f32<<<std::min(4096, cdiv(b->elems, 256)), 256, 0, c->s>>>(b->pack, b->pack32, b->elems);
    HIPC(hipGetLastError());
  }
  if (!c->p32) c->p32 = dalloc<float>(c->MN);
  if (!c->pb) c->pb = dalloc<double>(c->M);
  HIPC(hipStreamSynchronize(c->s));
  refresh_desc(c);
  if (c->eng) c->eng->attach(0, c);
}

void apply_gram_split(pqdts_ctx *c);

// Forms G = F^T F once, on the device, straight into its band tiles: the band
// GEMM multiplies F^T's band by F read from its span storage (B_SPAN), computes
// only the output tiles that meet the band of their row tile, and stores them in
// the band pack (E_GBAND). No dense copy of F or G exists at any point.
void ensure_gram(pqdts_ctx *c) {
  if (c->bg.pack) return;
  const int M = c->M;
  // band layout of G's rows (columns glo[k] .. glo[k] + gw[k] - 1), values 0
  pqdts_ctx::DBand &g = c->bg;
  const int nt = cdiv(M, TILE);
  std::vector<int64_t> off(nt, 0);
  std::vector<int> klo(nt, 0), nch(nt, 0);
  int64_t total = 0;
  int mx = 0;
  for (int t = 0; t < nt; t++) {
    int lo = INT32_MAX, hi = -1;
    for (int k = t * TILE; k < std::min(M, (t + 1) * TILE); k++)
      if (c->gw[k] > 0) {
        lo = std::min(lo, c->glo[k]);
        hi = std::max(hi, c->glo[k] + c->gw[k] - 1);
      }
    off[t] = total;
    if (hi >= 0) {
      klo[t] = lo;
      nch[t] = cdiv(hi - lo + 1, KC);
      total += (int64_t)nch[t] * KC * TILE;
      mx = std::max(mx, nch[t]);
    }
  }
  Prob *dd = nullptr;
  try {
    g.ntiles = nt;
    g.elems = std::max<int64_t>(total, 1);
    g.pack = dalloc<double>(g.elems);
    g.off = dalloc<int64_t>(nt);
    g.klo = dalloc<int>(nt);
    g.nch = dalloc<int>(nt);
    HIPC(hipMemsetAsync(g.pack, 0, sizeof(double) * g.elems, c->s));
    HIPC(hipMemcpyAsync(g.off, off.data(), sizeof(int64_t) * nt, hipMemcpyHostToDevice, c->s));
    HIPC(hipMemcpyAsync(g.klo, klo.data(), sizeof(int) * nt, hipMemcpyHostToDevice, c->s));
    HIPC(hipMemcpyAsync(g.nch, nch.data(), sizeof(int) * nt, hipMemcpyHostToDevice, c->s));
    // a temporary descriptor: N := M (the columns of G), the band of G as output
    Prob d = c->desc;
    d.N = M;
    d.bg = to_band(g);
    dd = dalloc<Prob>(1);
    HIPC(hipMemcpyAsync(dd, &d, sizeof(Prob), hipMemcpyHostToDevice, c->s));
    WorkList wl;
    wl.guard = -1;
    wl.np = 1;
    wl.pid[0] = 0;
    wl.flags[0] = F_FULL;
    wl.a[0] = 0.0;
    wl.boff[0] = 0;
    wl.boff[1] = cdiv(M, TILE) * c->bt.ntiles;
    // G = F^T F: band F^T (rows k, columns d), B = F (rows d, columns l)
    GemmOps ops{BAND_FT, 0, 1, 0, 0, 0};
    kb_band_gemm<double, KC, B_SPAN, E_GBAND, false, false, false><<<wl.boff[1], BLOCK, 0, c->s>>>(dd, wl, ops);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(c->s));
  } catch (...) {
    if (dd) (void)hipFree(dd);
    free_band(c->bg);
    throw;
  }
  (void)hipFree(dd);
  c->gram_nch = mx;
  refresh_desc(c);
  if (c->bf.pack32) ensure_fp32(c);
  apply_gram_split(c);
}

// Split of the Gram product over the band chunks of each tile (Prob::hs). A
// product of fewer tiles than the chip holds at once (about 4 workgroups per CU)
// runs at the latency of its longest tile, the one whose band spans the most
// chunks; splitting the chunks over hs workgroups shortens that chain hs-fold at
// the price of hs partial results in HBM and one summing pass (which also forms
// the fold values, so the separate fold-value pass goes away). By shape: the
// largest hs in {2, 4} with hs x tiles <= 1024 and at least 6 chunks per slice
// of the longest tile; FP64 products only: the FP32 products of mixed-precision
// solves run batched, where the chip is already full and the summing pass costs
// more than the shorter chain saves. An explicit split applies to both. The
// split depends on the problem and its precision alone, so a problem computes
// the same bits alone and in a batch.
// The table TUNED_SPLIT (written by the offline autotuner, pqdts_amd/autotune.py,
// and compiled in) gives the fastest measured split per bucket of
// (log2 output tiles, log2 most band chunks of a tile); 0 marks a bucket no tuned
// shape reached, which falls back to the rule: the largest k in {2, 4} with
// k x tiles <= 1024 and at least 6 chunks per slice of the longest tile.
#include "tuned_split.inc"

int ilog2_floor(int64_t v) {
  int b = 0;
  while (v > 1) {
    v >>= 1;
    b++;
  }
  return b;
}

int split_by_shape(int64_t N, int gram_tiles, int gram_nch) {
  const int64_t ntile = (int64_t)cdiv(N, TILE) * gram_tiles;
  const int b1 = std::min(TUNED_TILE_BUCKETS - 1, ilog2_floor(std::max<int64_t>(ntile, 1)));
  const int b2 = std::min(TUNED_CHUNK_BUCKETS - 1, ilog2_floor(std::max(gram_nch, 1)));
  const int tuned = TUNED_SPLIT[b1][b2];
  if (tuned > 0) return std::min(tuned, std::max(gram_nch, 1));
  int hs = 1;
  for (int k = 2; k <= 4; k *= 2)
    if (k * ntile <= 1024 && gram_nch >= 6 * k) hs = k;
  return hs;
}

void apply_gram_split(pqdts_ctx *c) {
  int hs = 1;
  if (c->hs_pref > 0) {
    hs = c->hs_pref;
  } else if (c->bg.pack) {
    hs = split_by_shape(c->N, c->bg.ntiles, c->gram_nch);
  }
  // the partials are only needed once G exists
  const bool need = hs > 1 && c->bg.pack;
  if (hs != c->hs || need != (c->hw != nullptr)) {
    if (c->hw) {
      HIPC(hipStreamSynchronize(c->s));
      (void)hipFree(c->hw);
    }
    c->hw = nullptr;
    c->hs = 1;
    if (need) c->hw = dalloc<double>((int64_t)hs * c->MN);
    c->hs = hs;
  }
  refresh_desc(c);
  if (c->eng) c->eng->attach(0, c);
}

void destroy_ctx(pqdts_ctx *c) {
  if (!c) return;
  if (c->batch) {
    // still in a batch slot: leave it first, so the batch never sees a freed context
    Engine *e = c->batch;
    if (e->s) (void)hipStreamSynchronize(e->s);
    for (int p = 0; p < e->cap; p++)
      if (e->ctx[p] == c) e->detach(p);
    c->batch = nullptr;
  }
  if (c->s) (void)hipStreamSynchronize(c->s);
  c->eng.reset();
  free_band(c->bf);
  free_band(c->bt);
  free_band(c->bg);
  void *ptrs[] = {c->flo,       c->flen,     c->foff, c->fval, c->fhas, c->kd_lo, c->kd_hi, c->P,
                  c->O,         c->R,         c->diag_full, c->diag_red, c->maxr, c->ik,  c->ev,    c->gm,    c->lam,
                  c->lmin,      c->part,     c->scal, c->ddesc, c->pn, c->hw, c->p32, c->pb};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (double *p : c->slot)
    if (p) (void)hipFree(p);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
}

// The shape of a problem as its context will hold it, from the COO pattern of F
// alone: the row spans of F (the reference's Frows, pqdts.f90:1328-1364, built in
// O(nnz) with zero fill), the column spans (rows of F^T), the row bands of
// G = F^T F, the band tiles of each, the Hessian-product path and the Gram split.
struct Plan {
  int64_t M = 0, N = 0, D = 0;
  std::vector<int> lo, len, kdlo, kdhi, glo, gw;
  std::vector<uint8_t> has;
  std::vector<int64_t> off;
  int64_t nspan = 0, gram_nnz = 0;
  int64_t bf_elems = 0, bt_elems = 0, bg_elems = 0;
  int bf_tiles = 0, bt_tiles = 0, bg_tiles = 0, bg_nch = 0;
  bool gram_auto = false;
};

void check_shape(int64_t M, int64_t N, int64_t D, int64_t nF, int64_t nP) {
  require(M >= 1 && N >= 1 && D >= 1, "need M >= 1, N >= 1, D >= 1");
  require(M * N < (int64_t)INT32_MAX * 8 && D * N < (int64_t)INT32_MAX * 8 && M < INT32_MAX && N < INT32_MAX &&
              D < INT32_MAX,
          "problem too large");
  require(nF >= 0 && nP >= 0, "negative nnz");
}

// elements and most chunks of one tile of a band stored per tile of TILE rows
template <class Span>
std::pair<int64_t, int> band_shape(int R, Span span) {
  int64_t total = 0;
  int mx = 0;
  for (int t = 0; t < cdiv(R, TILE); t++) {
    int lo = INT32_MAX, hi = -1;
    for (int r = t * TILE; r < std::min(R, (t + 1) * TILE); r++) {
      const auto sp = span(r);
      if (sp.second > 0) {
        lo = std::min(lo, sp.first);
        hi = std::max(hi, sp.first + sp.second - 1);
      }
    }
    if (hi >= 0) {
      const int nch = cdiv(hi - lo + 1, KC);
      total += (int64_t)nch * KC * TILE;
      mx = std::max(mx, nch);
    }
  }
  return {total, mx};
}

Plan make_plan(int64_t M, int64_t N, int64_t D, int64_t nF, const int32_t *Fr, const int32_t *Fc) {
  check_shape(M, N, D, nF, 0);
  require(nF == 0 || (Fr && Fc), "null F arrays");
  Plan p;
  p.M = M;
  p.N = N;
  p.D = D;
  p.lo.assign(D, INT32_MAX);
  p.len.assign(D, 0);
  p.has.assign(D, 0);
  std::vector<int> hi(D, -1);
  for (int64_t e = 0; e < nF; e++) {
    const int d = Fr[e], k = Fc[e];
    require(d >= 0 && d < D && k >= 0 && k < M, "F index out of range at entry " + std::to_string(e));
    p.lo[d] = std::min(p.lo[d], k);
    hi[d] = std::max(hi[d], k);
    p.has[d] = 1;
  }
  p.off.assign(D + 1, 0);
  for (int d = 0; d < D; d++) {
    if (!p.has[d]) p.lo[d] = 0;
    p.len[d] = p.has[d] ? hi[d] - p.lo[d] + 1 : 0;
    p.off[d + 1] = p.off[d] + p.len[d];
  }
  p.nspan = p.off[D];
  p.kdlo.assign(M, INT32_MAX);
  p.kdhi.assign(M, -1);
  p.glo.assign(M, INT32_MAX);
  p.gw.assign(M, 0);
  std::vector<int> ghi(M, -1);
  for (int d = 0; d < D; d++)
    for (int k = p.lo[d]; k < p.lo[d] + p.len[d]; k++) {
      p.kdlo[k] = std::min(p.kdlo[k], d);
      p.kdhi[k] = std::max(p.kdhi[k], d);
      p.glo[k] = std::min(p.glo[k], p.lo[d]);
      ghi[k] = std::max(ghi[k], p.lo[d] + p.len[d] - 1);
    }
  for (int k = 0; k < M; k++) {
    if (p.kdhi[k] < 0) p.kdlo[k] = 0;
    if (ghi[k] < 0) p.glo[k] = 0;
    p.gw[k] = ghi[k] >= 0 ? ghi[k] - p.glo[k] + 1 : 0;
    p.gram_nnz += p.gw[k];
  }
  p.bf_tiles = cdiv(D, TILE);
  p.bt_tiles = p.bg_tiles = cdiv(M, TILE);
  p.bf_elems = band_shape((int)D, [&](int d) { return std::make_pair(p.lo[d], p.len[d]); }).first;
  p.bt_elems = band_shape((int)M, [&](int k) {
                 return std::make_pair(p.kdlo[k], p.kdhi[k] >= 0 ? p.kdhi[k] - p.kdlo[k] + 1 : 0);
               }).first;
  const auto g = band_shape((int)M, [&](int k) { return std::make_pair(p.glo[k], p.gw[k]); });
  p.bg_elems = g.first;
  p.bg_nch = g.second;
  // the Gram HVP when its band tiles clearly cost less than the two F products
  p.gram_auto = (double)p.bg_elems < 0.8 * (double)(p.bf_elems + p.bt_elems) && M * M * 8 < ((int64_t)64 << 30);
  return p;
}

// Device bytes a context holds: every array its creation, the Gram matrix, mixed
// precision and the Gram split allocate (the stream and host copies aside).
int64_t footprint_bytes(int64_t M, int64_t N, int64_t D, int64_t nspan, int64_t bf_elems, int bf_tiles,
                        int64_t bt_elems, int bt_tiles, bool gram, int64_t bg_elems, int bg_tiles, bool fp32,
                        int hs) {
  const int64_t MN = M * N, DN = D * N;
  const int64_t pcap = std::max<int64_t>({(int64_t)cdiv(N, TILE) * std::max(bf_tiles, bt_tiles),
                                          cdiv(M, WAVES_PER_BLOCK), cdiv(MN, EBLOCK), 1});
  int64_t b = D * (4 + 4 + 8 + 1) + 8 * nspan + M * (4 + 4) + 24 * DN + 48 * MN + M * (8 + 8 + 4) + MN +
              M * 4 * 8 + 8 * pcap * NREG + 8 * NSCAL + 4 * NREG + (int64_t)sizeof(Prob);
  // a band pack holds at least one element
  auto band = [&](int64_t elems, int tiles) { return 8 * std::max<int64_t>(elems, 1) + 16 * (int64_t)tiles; };
  b += band(bf_elems, bf_tiles) + band(bt_elems, bt_tiles) + (gram ? band(bg_elems, bg_tiles) : 0);
  if (fp32)
    b += 4 * (std::max<int64_t>(bf_elems, 1) + std::max<int64_t>(bt_elems, 1) +
              (gram ? std::max<int64_t>(bg_elems, 1) : 0)) +
         4 * MN + 8 * M;
  if (hs > 1 && gram) b += 8 * (int64_t)hs * MN;
  return b;
}

// bytes held only while the Gram matrix is formed: a temporary descriptor
int64_t gram_transient_bytes() { return (int64_t)sizeof(Prob); }

// what a context of this plan will hold (persistent) and, on top, while it is created
void plan_footprint(const Plan &p, int hvp_pref, int hs_pref, bool fp32, int64_t *persistent, int64_t *transient) {
  const bool gram = hvp_pref == 2 || (hvp_pref == 0 && p.gram_auto);
  const int hs = !gram ? 1 : hs_pref > 0 ? hs_pref : split_by_shape(p.N, p.bg_tiles, p.bg_nch);
  *persistent = footprint_bytes(p.M, p.N, p.D, p.nspan, p.bf_elems, p.bf_tiles, p.bt_elems, p.bt_tiles, gram,
                                p.bg_elems, p.bg_tiles, fp32, hs);
  *transient = gram ? gram_transient_bytes() : 0;
}

int env_hvp_pref() {
  const char *e = getenv("PQDTS_HVP");
  return e ? ((e[0] == 'G') ? 2 : (e[0] == 'F') ? 1 : 0) : 0;
}

pqdts_ctx *create_ctx(int64_t M, int64_t N, int64_t D, int64_t nF, const int32_t *Fr, const int32_t *Fc,
                      const double *Fv, int64_t nP, const int32_t *Pr, const int32_t *Pc, const double *Pv,
                      double gamma) {
  check_shape(M, N, D, nF, nP);
  require(nF == 0 || (Fr && Fc && Fv), "null F arrays");
  require(nP == 0 || (Pr && Pc && Pv), "null P arrays");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Fail(-2, "no GPU visible to pqdts_amd");

  // F ingest: O(nnz) span build with zero fill and last-wins duplicates (the same
  // spans as the O(D*nnz) scan of pqdts.f90:1328-1364)
  const Plan pl = make_plan(M, N, D, nF, Fr, Fc);
  const std::vector<int> &lo = pl.lo, &len = pl.len, &kdlo = pl.kdlo, &kdhi = pl.kdhi, &glo = pl.glo, &gw = pl.gw;
  const std::vector<uint8_t> &has = pl.has;
  const std::vector<int64_t> &off = pl.off;
  std::vector<double> val(off[D], 0.0);
  for (int64_t e = 0; e < nF; e++) val[off[Fr[e]] + (Fc[e] - lo[Fr[e]])] = Fv[e];
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
    c->nspan = off[D];
    c->glo = glo;
    c->gw = gw;
    c->gram_nnz = pl.gram_nnz;
    c->gram_elems = pl.bg_elems;
    c->gram_auto = pl.gram_auto;
    c->hvp_pref = env_hvp_pref();
    HIPC(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    c->flo = dalloc<int>(D);
    c->flen = dalloc<int>(D);
    c->foff = dalloc<int64_t>(D);
    c->fval = dalloc<double>(off[D]);
    c->fhas = dalloc<uint8_t>(D);
    c->kd_lo = dalloc<int>(M);
    c->kd_hi = dalloc<int>(M);
    c->P = dalloc<double>(c->DN);
    c->O = dalloc<double>(c->DN);
    c->R = dalloc<double>(c->DN);
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
    const int nbt = std::max((int)hbf.off.size(), (int)hbt.off.size());
    c->pcap = std::max({gx * nbt, cdiv(M, WAVES_PER_BLOCK), cdiv(c->MN, EBLOCK), 1});
    c->part = dalloc<double>((int64_t)c->pcap * NREG);
    c->scal = dalloc<double>(NSCAL);
    c->pn = dalloc<int>(NREG);
    HIPC(hipMemsetAsync(c->pn, 0, sizeof(int) * NREG, c->s));
    c->ddesc = dalloc<Prob>(1);
    auto up = [&](void *dst, const void *src, size_t bytes) {
      if (bytes) HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->s));
    };
    up(c->flo, lo.data(), sizeof(int) * D);
    up(c->flen, len.data(), sizeof(int) * D);
    up(c->foff, off.data(), sizeof(int64_t) * D);
    up(c->fval, val.data(), sizeof(double) * off[D]);
    up(c->fhas, has.data(), D);
    up(c->kd_lo, kdlo.data(), sizeof(int) * M);
    up(c->kd_hi, kdhi.data(), sizeof(int) * M);
    up(c->P, Ph.data(), sizeof(double) * c->DN);
    for (int s = 0; s < 6; s++) HIPC(hipMemsetAsync(c->slot[s], 0, sizeof(double) * c->MN, c->s));
    HIPC(hipMemsetAsync(c->O, 0, sizeof(double) * c->DN, c->s));
    HIPC(hipMemsetAsync(c->ik, 0, c->MN, c->s));
    HIPC(hipMemsetAsync(c->scal, 0, sizeof(double) * NSCAL, c->s));
    std::vector<int> m1(M, -1);
    up(c->maxr, m1.data(), sizeof(int) * M);
    HIPC(hipStreamSynchronize(c->s));
    upload_band(c, c->bf, hbf);
    upload_band(c, c->bt, hbt);
    refresh_desc(c);
    // Bdiag once (stage-constant; the reference recomputes it at every CG step)
    WorkList wl;
    wl.guard = -1;
    wl.np = 1;
    wl.pid[0] = 0;
    wl.boff[0] = 0;
    wl.boff[1] = cdiv(M, WAVES_PER_BLOCK);
    wl.flags[0] = F_FULL;
    wl.a[0] = 0.0;
    kb_bdiag<<<wl.boff[1], BLOCK, 0, c->s>>>(c->ddesc, wl);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(c->s));
    if (c->use_gram()) ensure_gram(c);
  } catch (...) {
    destroy_ctx(c);
    throw;
  }
  return c;
}

// device bytes the context holds now (the footprint formula on its actual arrays)
int64_t ctx_device_bytes(const pqdts_ctx *c) {
  return footprint_bytes(c->M, c->N, c->D, c->nspan, c->bf.elems, c->bf.ntiles, c->bt.elems, c->bt.ntiles,
                         c->bg.pack != nullptr, c->bg.elems, c->bg.ntiles, c->p32 != nullptr, c->hw ? c->hs : 1);
}

pqdts_ctx *checked(pqdts_ctx *c) {
  require(c != nullptr, "null context");
  return c;
}

}  // namespace

Engine &pqdts_ctx::engine() {
  if (!eng) {
    eng.reset(new Engine());
    eng->private_scal = true;
    eng->init(1, s, false);
    eng->attach(0, this);
  }
  return *eng;
}

namespace {

double *slot_ptr(pqdts_ctx *c, int slot) {
  require(slot >= 0 && slot < 6, "invalid slot id " + std::to_string(slot));
  return c->slot[slot];
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

// standalone operator on one context
const std::vector<int> ONE{0};

Engine &E1(pqdts_ctx *c) { return checked(c)->engine(); }

}  // namespace

// ===========================================================================
// batches (served / batched form)
// ===========================================================================
struct pqdts_batch {
  Engine eng;
  std::vector<int> finished;  // slots finished since the last query
};

namespace {
// Places ctx (its X is the starting point) in a free slot of e with a fresh solver
// state; returns the slot.
int engine_admit(Engine &e, pqdts_ctx *ctx, int mode, int maxiter, double eps, int start_stage, int output,
                 int precision, pqdts_event_cb cb, void *user) {
  require(mode == 1 || mode == 2, "mode must be 1 or 2");
  require(start_stage == 1 || start_stage == 2, "start_stage must be 1 or 2");
  require(precision == 0 || precision == 1, "precision must be 0 (FP64) or 1 (mixed)");
  require(ctx->batch == nullptr, "the context is already in a batch (release it first)");
  int p = -1;
  for (int i = 0; i < e.cap; i++)
    if (!e.ctx[i]) {
      p = i;
      break;
    }
  require(p >= 0, "batch is full");
  HIPC(hipStreamSynchronize(ctx->s));
  if (precision) ensure_fp32(ctx);
  e.fp32[p] = precision;
  e.attach(p, ctx);
  ctx->batch = &e;
  PState *q = new PState();
  q->mode = mode;
  q->maxiter = maxiter;
  q->eps = eps;
  q->start_stage = start_stage;
  q->output = output;
  q->cb = cb;
  q->user = user;
  q->pol = ctx->pol;
  ctx->pol_out = pqdts_ctx::PolishOut{};
  e.st[p].reset(q);
  ctx->full = true;
  e.face_prepare(p);
  return p;
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

int pqdts_device_cu_count(void) {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
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
  return guarded([&] {
    Engine &e = E1(ctx);
    slot_ptr(ctx, x);
    e.L(ONE, x, true);
    e.fetch();
    *val = e.Lval(0);
  });
}

int pqdts_dL(pqdts_ctx *ctx, int x, int g) {
  return guarded([&] {
    Engine &e = E1(ctx);
    slot_ptr(ctx, x);
    slot_ptr(ctx, g);
    e.dL(ONE, x, g);
    HIPC(hipStreamSynchronize(e.s));
  });
}

int pqdts_BF(pqdts_ctx *ctx, int dd, int bx) {
  return guarded([&] {
    Engine &e = E1(ctx);
    slot_ptr(ctx, dd);
    slot_ptr(ctx, bx);
    require(dd != bx, "BF input and output must differ");
    e.BF(ONE, dd, bx, false, false, false);
    HIPC(hipStreamSynchronize(e.s));
  });
}

int pqdts_Bdiag(pqdts_ctx *ctx, double *diag) {
  return guarded([&] {
    checked(ctx);
    if (diag) {
      HIPC(hipMemcpyAsync(diag, ctx->full ? ctx->diag_full : ctx->diag_red, sizeof(double) * ctx->M,
                          hipMemcpyDeviceToHost, ctx->s));
      HIPC(hipStreamSynchronize(ctx->s));
    }
  });
}

int pqdts_doproj(pqdts_ctx *ctx, int x) {
  return guarded([&] {
    Engine &e = E1(ctx);
    slot_ptr(ctx, x);
    // doproj is the Condat projection whatever the space (pqdts.f90:1120-1140)
    const bool f = ctx->full;
    ctx->full = true;
    try {
      e.project(ONE, x, false, false);
      HIPC(hipStreamSynchronize(e.s));
    } catch (...) {
      ctx->full = f;
      throw;
    }
    ctx->full = f;
  });
}

int pqdts_doproj2(pqdts_ctx *ctx, int x, double *xmin) {
  return guarded([&] {
    Engine &e = E1(ctx);
    slot_ptr(ctx, x);
    require(!ctx->full, "doproj2 needs the reduced space (maxr set)");
    e.project(ONE, x, false, false);
    e.fetch();
    *xmin = std::fmin(0.0, e.host_scal(0)[S_XMIN]);
  });
}

int pqdts_conv_test(pqdts_ctx *ctx, int x, int g, double *conv) {
  return guarded([&] {
    Engine &e = E1(ctx);
    slot_ptr(ctx, x);
    slot_ptr(ctx, g);
    e.conv(ONE, x, g, false);
    e.fetch();
    *conv = e.convval(0);
  });
}

int pqdts_pdot(pqdts_ctx *ctx, int x, int y, double *val) {
  return guarded([&] {
    Engine &e = E1(ctx);
    slot_ptr(ctx, x);
    slot_ptr(ctx, y);
    e.pdot(ONE, x, y);
    e.fetch();
    *val = e.host_scal(0)[S_DOT];
  });
}

// BLAS-1 on slots; an operand of -1 is not used by that kind
static int blas1(pqdts_ctx *ctx, int z, int x, int y, double a, double b, int kind, const char *timer) {
  return guarded([&] {
    Engine &e = E1(ctx);
    Engine::Scope sc(&e, timer);
    slot_ptr(ctx, z);
    if (x >= 0) slot_ptr(ctx, x);
    if (y >= 0) slot_ptr(ctx, y);
    e.axpby(ONE, z, x, y, a, b, kind, false);
    HIPC(hipStreamSynchronize(e.s));
  });
}

int pqdts_paxpy(pqdts_ctx *ctx, double alpha, int x, int y) {
  return blas1(ctx, y, x, y, alpha, 0.0, AX_AXPY, "paxpy");
}
int pqdts_pgaxpy(pqdts_ctx *ctx, double alpha, int x, double beta, int y) {
  return blas1(ctx, y, x, y, alpha, beta, AX_GAXPY, "gpaxpy");
}
int pqdts_pgaxpy2(pqdts_ctx *ctx, double alpha, int x, double beta, int y, int z) {
  return blas1(ctx, z, x, y, alpha, beta, AX_GAXPY, "pgpaxpy2");
}
int pqdts_pset(pqdts_ctx *ctx, double alpha, int x) { return blas1(ctx, x, -1, -1, alpha, 0.0, AX_SET, "pset"); }
int pqdts_pcopy(pqdts_ctx *ctx, int x, int y) { return blas1(ctx, y, x, -1, 1.0, 0.0, AX_COPY, "pcopy"); }
int pqdts_pscale(pqdts_ctx *ctx, double alpha, int x) {
  return blas1(ctx, x, x, -1, alpha, 0.0, AX_SCALE, "pscale");
}
int pqdts_pcopy_scale(pqdts_ctx *ctx, double alpha, int x, int y) {
  return blas1(ctx, y, x, -1, alpha, 0.0, AX_SCALE, "pcopy_scale");
}

int pqdts_transform(pqdts_ctx *ctx) {
  return guarded([&] {
    Engine &e = E1(ctx);
    require(!ctx->full, "the reduced-space transform needs tfullspace = false");
    e.transform(ONE, true);
    HIPC(hipStreamSynchronize(e.s));
  });
}

static int linesearch_one(pqdts_ctx *ctx, bool condat, double *alpha, double *dlalpha) {
  return guarded([&] {
    Engine &e = E1(ctx);
    if (!condat) require(!ctx->full, "Lalpha_min2 needs the reduced space (maxr set)");
    const bool f = ctx->full;
    ctx->full = condat;
    e.st[0].reset(new PState());
    try {
      e.line_search(ONE);
    } catch (...) {
      ctx->full = f;
      throw;
    }
    ctx->full = f;
    *alpha = e.st[0]->alpha;
    *dlalpha = e.st[0]->dlalpha;
    e.st[0].reset();
  });
}

int pqdts_lalpha_min2(pqdts_ctx *ctx, double *alpha, double *dlalpha) {
  return linesearch_one(ctx, false, alpha, dlalpha);
}

int pqdts_lalpha_min_auglag(pqdts_ctx *ctx, double *alpha, double *dlalpha) {
  return linesearch_one(ctx, true, alpha, dlalpha);
}

int pqdts_solve(pqdts_ctx *ctx, int mode, int maxiter, double eps, int start_stage, int output, pqdts_event_cb cb,
                void *user, double *conv) {
  return guarded([&] {
    Engine &e = E1(ctx);
    require(mode == 1 || mode == 2, "mode must be 1 or 2");
    require(start_stage == 1 || start_stage == 2, "start_stage must be 1 or 2");
    PState *q = new PState();
    e.st[0].reset(q);
    q->mode = mode;
    q->maxiter = maxiter;
    q->eps = eps;
    q->start_stage = start_stage;
    q->output = output;
    q->cb = cb;
    q->user = user;
    q->pol = ctx->pol;
    ctx->pol_out = pqdts_ctx::PolishOut{};
    e.fp32[0] = 0;
    e.face_prepare(0);
    Engine::Scope sc(&e, "min");
    // the reference starts in full space (pqdts.f90:1225)
    ctx->full = true;
    try {
      while (e.any_active()) e.tick();
    } catch (...) {
      e.st[0].reset();
      throw;
    }
    HIPC(hipStreamSynchronize(e.s));
    if (conv) *conv = q->conv;
    e.st[0].reset();
    e.release_deferred();
  });
}

int pqdts_set_polish(pqdts_ctx *ctx, int max_outer, double cg_rtol, double tol, double budget, int fp32) {
  return guarded([&] {
    checked(ctx);
    require(max_outer >= 0 && cg_rtol > 0.0 && cg_rtol < 1.0 && tol > 0.0 && (fp32 == 0 || fp32 == 1),
            "bad polish settings");
    require(ctx->batch == nullptr, "the context is in a batch (release it first)");
    ctx->pol.max_outer = max_outer;
    ctx->pol.cg_rtol = cg_rtol;
    ctx->pol.tol = tol;
    ctx->pol.budget = budget;
    ctx->pol.fp32 = fp32;
  });
}

int pqdts_set_polish_method(pqdts_ctx *ctx, int method) {
  return guarded([&] {
    checked(ctx);
    require(method == 0 || method == 1, "bad polish method (0 Newton-CG, 1 direct face solves)");
    require(ctx->batch == nullptr, "the context is in a batch (release it first)");
    ctx->pol.method = method;
  });
}

int pqdts_polish_result(pqdts_ctx *ctx, int *stats, int64_t *nact, double *step) {
  return guarded([&] {
    checked(ctx);
    const pqdts_ctx::PolishOut &o = ctx->pol_out;
    if (stats) {
      stats[0] = o.status;
      stats[1] = o.iters;
      stats[2] = o.rounds;
    }
    if (nact) *nact = o.nact;
    if (step) *step = o.step;
  });
}

int pqdts_timers(pqdts_ctx *ctx, int cap, char (*names)[24], int *counts, double *seconds) {
  int n = 0;
  int rc = guarded([&] {
    Engine &e = E1(ctx);
    HIPC(hipStreamSynchronize(e.s));
    e.tresolve();
    for (const auto &name : e.timer_order) {
      if (n < cap) {
        const Timer &t = e.timers[name];
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

static int kernel_stats_of(Engine &e, int cap, char (*names)[24], int *counts, double *ms, double *flops,
                           double *bytes) {
  HIPC(hipStreamSynchronize(e.s));
  e.kfold();
  const int n = (int)e.knames.size();
  for (int i = 0; i < std::min(n, cap); i++) {
    if (names) {
      std::strncpy(names[i], e.knames[i].c_str(), 23);
      names[i][23] = 0;
    }
    if (counts) counts[i] = e.kcount[i];
    if (ms) ms[i] = e.kms[i];
    if (flops) flops[i] = e.kwork[i];
    if (bytes) bytes[i] = e.kbytes[i];
  }
  return n;
}

static void kernel_timing_of(Engine &e, int on) {
  HIPC(hipStreamSynchronize(e.s));
  e.kfold();
  e.ktiming = on != 0;
  e.ksample = std::max(1, on);
  e.kseq = 0;
  for (size_t i = 0; i < e.kcount.size(); i++) {
    e.kcount[i] = 0;
    e.kms[i] = 0.0;
    e.kwork[i] = 0.0;
    e.kbytes[i] = 0.0;
  }
}

int pqdts_device_timers(pqdts_ctx *ctx, int on) {
  return guarded([&] {
    Engine &e = E1(ctx);
    HIPC(hipStreamSynchronize(e.s));
    e.tresolve();
    e.dev_timers = on != 0;
  });
}

int pqdts_kernel_timing(pqdts_ctx *ctx, int on) {
  return guarded([&] { kernel_timing_of(E1(ctx), on); });
}

int pqdts_kernel_stats(pqdts_ctx *ctx, int cap, char (*names)[24], int *counts, double *ms, double *flops,
                       double *bytes) {
  int n = 0;
  int rc = guarded([&] { n = kernel_stats_of(E1(ctx), cap, names, counts, ms, flops, bytes); });
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
    info[4] = ctx->bf.elems;
    info[5] = ctx->bf.ntiles;
    info[6] = ctx->bt.ntiles;
    info[7] = ctx->use_gram() ? 1 : 0;
    info[8] = ctx->gram_elems;
    info[9] = ctx->gram_nnz;
    info[10] = ctx->bt.elems;
    // device bytes held by this problem
    info[11] = ctx_device_bytes(ctx);
    info[12] = ctx->hs;
    info[13] = ctx->gram_nch;
  });
}

int pqdts_set_hvp_path(pqdts_ctx *ctx, int path) {
  return guarded([&] {
    checked(ctx);
    require(path >= 0 && path <= 2, "hvp path must be 0 (auto), 1 (F products) or 2 (Gram)");
    require(ctx->batch == nullptr, "the context is in a batch (release it first)");
    ctx->hvp_pref = path;
    if (ctx->use_gram()) ensure_gram(ctx);
  });
}

int pqdts_set_hvp_split(pqdts_ctx *ctx, int split) {
  return guarded([&] {
    checked(ctx);
    require(split >= 0 && split <= 16, "split must be 0 (by shape) or 1..16");
    require(ctx->batch == nullptr, "the context is in a batch (release it first)");
    ctx->hs_pref = split;
    apply_gram_split(ctx);
  });
}

int pqdts_simplexproj_rows(const double *y, double *x, int64_t rows, int64_t n) {
  return guarded([&] {
    require(rows >= 0 && n >= 1 && rows * n < (int64_t)INT32_MAX * 8, "bad shape");
    require(y != nullptr && x != nullptr && x != y, "x must be a distinct non-null buffer");
    if (rows == 0) return;
    // a context-free descriptor with the rows as its X operand
    double *d = dalloc<double>(rows * n);
    Prob *pd = nullptr;
    hipError_t e1 = hipMemcpy(d, y, sizeof(double) * rows * n, hipMemcpyHostToDevice);
    double *part = nullptr;
    int *pn = nullptr;
    if (e1 == hipSuccess) {
      Prob p{};
      p.M = (int)rows;
      p.N = (int)n;
      p.pcap = cdiv(rows, WAVES_PER_BLOCK);
      p.op[0] = d;
      part = dalloc<double>((int64_t)p.pcap * NREG);
      p.part = part;
      // kb_project records its partial count
      pn = dalloc<int>(NREG);
      p.pn = pn;
      pd = dalloc<Prob>(1);
      e1 = hipMemcpy(pd, &p, sizeof(Prob), hipMemcpyHostToDevice);
      if (e1 == hipSuccess) {
        WorkList wl;
        wl.guard = -1;
        wl.np = 1;
        wl.pid[0] = 0;
        wl.boff[0] = 0;
        wl.boff[1] = p.pcap;
        wl.flags[0] = F_FULL;
        wl.a[0] = 0.0;
        kb_project<<<wl.boff[1], BLOCK>>>(pd, wl, 0, 0, 0, R_C);
        e1 = hipGetLastError();
      }
    }
    if (e1 == hipSuccess) e1 = hipMemcpy(x, d, sizeof(double) * rows * n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (pd) (void)hipFree(pd);
    if (part) (void)hipFree(part);
    if (pn) (void)hipFree(pn);
    HIPC(e1);
  });
}

}  // extern "C"

// ---- the projection without allocation ------------------------------------
// pqdts_simplexproj_rows allocates a device copy of the rows, a descriptor and the
// projection's per-block partials on every call. The variant below takes the rows
// in device memory and a stream, passes the descriptor as kernel arguments and
// keeps the partials (one double per block of WAVES_PER_BLOCK rows) and their
// count in this file-scope device array. Each GPU holds its own copy until the
// process ends; its address is looked up once per device.
namespace {
constexpr int64_t NOALLOC_MAX_ROWS = (int64_t)1 << 20;
constexpr int64_t NOALLOC_PARTS = NOALLOC_MAX_ROWS / WAVES_PER_BLOCK;
constexpr int NOALLOC_MAX_DEV = 64;
static __device__ __align__(256) double scratch[NOALLOC_PARTS + NREG];
std::mutex g_scratch_mu;
double *g_scratch[NOALLOC_MAX_DEV] = {};

// this device's address of `scratch` (the name itself is a host-side stand-in)
double *scratch_on_current_device() {
  int dev = 0;
  HIPC(hipGetDevice(&dev));
  require(dev >= 0 && dev < NOALLOC_MAX_DEV, "device index beyond the no-allocation table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr, "no device address for the scratch array");
    g_scratch[dev] = (double *)p;
  }
  return g_scratch[dev];
}
}  // namespace

extern "C" {

int pqdts_simplexproj_rows_noalloc(const double *y, double *x, int64_t rows, int64_t n, void *stream) {
  return guarded([&] {
    require(rows >= 0 && n >= 1 && n < INT32_MAX && rows * n < (int64_t)INT32_MAX * 8, "bad shape");
    require(rows <= NOALLOC_MAX_ROWS, "pqdts_simplexproj_rows_noalloc: at most 1048576 rows per call");
    require(y != nullptr && x != nullptr, "null device pointer");
    require(x == y || x + rows * n <= y || y + rows * n <= x, "x and y overlap without being the same rows");
    if (rows == 0) return;
    double *sp = scratch_on_current_device();
    hipStream_t s = (hipStream_t)stream;
    if (x != y) HIPC(hipMemcpyAsync(x, y, sizeof(double) * rows * n, hipMemcpyDeviceToDevice, s));
    Prob p{};
    p.M = (int)rows;
    p.N = (int)n;
    p.pcap = cdiv(rows, WAVES_PER_BLOCK);
    p.op[0] = x;
    p.part = sp;
    p.pn = (int *)(sp + NOALLOC_PARTS);
    kb_project_byval<<<p.pcap, BLOCK, 0, s>>>(p, 0);
    HIPC(hipGetLastError());
  });
}

// The reference's own foreign function (condat_simplexproj.c:60), called per row
// by its doproj (pqdts.f90:1133) through an implicit interface: all arguments by
// reference, y read only, x written (and used as scratch), no aliasing, no
// status. The row is projected on the GPU by Condat's sequential filter in the
// reference's order, so x is bit-identical. Having no error channel, a failure
// (no GPU, HIP error, aliasing) is reported on stderr and aborts the process.
void simplexproj_condat_(double *y, double *x, int *plength) {
  const int rc = guarded([&] {
    require(plength != nullptr, "null length");
    const int64_t n = *plength;
    if (n <= 0) return;
    require(y != nullptr && x != nullptr, "null row pointer");
    require(x + n <= y || y + n <= x, "x and y overlap (x is the output and the scratch)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Fail(-2, "no GPU visible to pqdts_amd");
    double *d = dalloc<double>(2 * n);
    hipError_t e = hipMemcpy(d, y, sizeof(double) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      kb_condat_rows<<<1, BLOCK>>>(d, d + n, 1, (int)n);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(x, d + n, sizeof(double) * n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    HIPC(e);
  });
  if (rc != 0) {
    std::fprintf(stderr, "simplexproj_condat_: %s\n", g_err.c_str());
    std::abort();
  }
}

// ---- batches --------------------------------------------------------------
int pqdts_batch_create(int capacity, pqdts_batch **out) {
  return guarded([&] {
    require(out != nullptr && capacity >= 1, "bad batch arguments");
    *out = nullptr;
    pqdts_batch *b = new pqdts_batch();
    try {
      b->eng.init(capacity, nullptr, true);
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}

int pqdts_batch_destroy(pqdts_batch *b) {
  return guarded([&] {
    if (b)
      for (int p = 0; p < b->eng.cap; p++)
        if (b->eng.ctx[p] && b->eng.ctx[p]->batch == &b->eng) b->eng.ctx[p]->batch = nullptr;
    delete b;
  });
}

int pqdts_batch_admit(pqdts_batch *b, pqdts_ctx *ctx, int mode, int maxiter, double eps, int start_stage,
                      int output, int precision, pqdts_event_cb cb, void *user, int *slot) {
  return guarded([&] {
    require(b != nullptr, "null batch");
    const int p = engine_admit(b->eng, checked(ctx), mode, maxiter, eps, start_stage, output, precision, cb, user);
    if (slot) *slot = p;
  });
}

int pqdts_batch_step(pqdts_batch *b, int *finished, int cap) {
  int n = 0;
  int rc = guarded([&] {
    require(b != nullptr, "null batch");
    Engine &e = b->eng;
    if (e.any_active()) e.tick();
    HIPC(hipStreamSynchronize(e.s));
    e.settle_deferred();
    for (int p = 0; p < e.cap; p++)
      if (e.ctx[p] && e.st[p] && e.st[p]->phase == 2) {
        if (finished && n < cap) finished[n] = p;
        n++;
      }
  });
  return rc == 0 ? n : -1;
}

int pqdts_batch_result(pqdts_batch *b, int slot, double *conv, int *stats) {
  return guarded([&] {
    require(b != nullptr && slot >= 0 && slot < b->eng.cap && b->eng.ctx[slot] && b->eng.st[slot],
            "no problem in that slot");
    PState &q = *b->eng.st[slot];
    if (conv) *conv = q.conv;
    if (stats) {
      stats[0] = q.phase == 2 ? 1 : 0;
      stats[1] = q.newton;
      stats[2] = q.pcg;
      stats[3] = q.trials;
    }
  });
}

int pqdts_batch_release(pqdts_batch *b, int slot) {
  return guarded([&] {
    require(b != nullptr && slot >= 0 && slot < b->eng.cap, "bad slot");
    HIPC(hipStreamSynchronize(b->eng.s));
    b->eng.detach(slot);
  });
}

int pqdts_batch_set_cu_mask(pqdts_batch *b, const uint32_t *mask, int nwords) {
  return guarded([&] {
    // The mask is checked before anything else (and before any HIP call). Only
    // one consecutive range of CUs is accepted: a host crash was observed while
    // a stream with a non-consecutive mask (all CUs but every 8th) was set up,
    // with no cause found in this library, so such masks are refused.
    require(mask != nullptr && nwords >= 1 && nwords <= 64, "bad CU mask arguments");
    int lo = -1, hi = -1, count = 0;
    for (int i = 0; i < 32 * nwords; i++)
      if ((mask[i / 32] >> (i % 32)) & 1u) {
        if (lo < 0) lo = i;
        hi = i;
        count++;
      }
    require(count > 0, "the CU mask selects no CU");
    require(count == hi - lo + 1, "the CU mask must select one consecutive range of CUs");
    require(b != nullptr, "null batch");
    for (int p = 0; p < b->eng.cap; p++) require(b->eng.ctx[p] == nullptr, "the batch must be empty");
    b->eng.set_cu_mask(mask, nwords);
  });
}

int pqdts_batch_active(pqdts_batch *b) {
  int n = 0;
  int rc = guarded([&] {
    require(b != nullptr, "null batch");
    for (int p = 0; p < b->eng.cap; p++)
      if (b->eng.ctx[p]) n++;
  });
  return rc == 0 ? n : -1;
}

int pqdts_batch_kernel_timing(pqdts_batch *b, int on) {
  return guarded([&] {
    require(b != nullptr, "null batch");
    kernel_timing_of(b->eng, on);
  });
}

int pqdts_batch_kernel_stats(pqdts_batch *b, int cap, char (*names)[24], int *counts, double *ms,
                             double *flops, double *bytes) {
  int n = 0;
  int rc = guarded([&] {
    require(b != nullptr, "null batch");
    n = kernel_stats_of(b->eng, cap, names, counts, ms, flops, bytes);
  });
  return rc == 0 ? n : -1;
}

}  // extern "C"

#include "server.hip"
