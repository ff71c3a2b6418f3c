// abc_capi.hip -- host side of libabc.so: the C ABI declared in include/abc.h.
//
// Orchestrates, per call, on one HIP stream:
//   prior_kernel -> rates_kernel -> bucket order of particles by expected cost
//   (heaviest first) -> item offsets -> burn_kernel (burn-in of every cell up
//   to its labelling window) -> ssa_kernel (the window, persistent, lane-level
//   work stealing) -> stats_kernel -> error / acceptance kernels.
// Device buffers are owned by the context and grow on demand.  The *_dev
// sequences use kernels only (no memset, no host wait), so they can be
// captured into HIP graphs (abc_graph_begin / abc_graph_end).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/abc.h"
#include "abc_kernels.h"

// Device memory of abc_simulate_noalloc_dev: one array per GPU, alive until
// the process ends.  Host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (scratch_address).
#define ABC_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
static __device__ __align__(256) unsigned char scratch[ABC_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The address of this device's scratch array (the device must be current);
// looked up once per device.
hipError_t scratch_address(int device, void** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  const bool cached = device >= 0 && device < kMaxDevices;
  if (cached && g_scratch_ptr[device]) {
    *out = g_scratch_ptr[device];
    return hipSuccess;
  }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (cached) g_scratch_ptr[device] = p;
  *out = p;
  return hipSuccess;
}

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

// State of an acceptance session (abc_accept_begin .. abc_accept_end).
struct AcceptState {
  bool active = false;
  ErrSpec E{};
  int Kp = 0;               // K padded to the MFMA k-chunk (32)
  int64_t G = 0, Gpad = 0;  // genes, padded to the 64-gene tile
  int mode = 0;
  double eps = 0.0, margin_rel = 0.0;
  int64_t k = 0;
  int64_t n_candidates = 0, n_screened = 0;
  unsigned long long cand_cap = 0;
  // top-k (mode 1): per-gene pools of capg = k + n_sub entries in device
  // memory; batches are screened n_sub rows at a time
  bool pooled = false;
  int64_t n_sub = 0, capg = 0;
  std::vector<double> thr_host;                                // per-gene threshold
  std::vector<std::vector<std::pair<double, int64_t>>> kept;   // per-gene (err, index)
};

}  // namespace

struct abc_ctx {
  int device = 0;
  void* scratch = nullptr;   // this device's scratch array (abc_simulate_noalloc_dev)
  hipStream_t stream = nullptr;
  bool own_stream = true;
  int num_cu = 0;
  int ssa_blocks = 0;        // 0 = occupancy default
  int ssa_claim = 2;         // SSA work items a lane claims per queue atomic
  int ssa_blocks_default = 0;
  int burn_blocks = 0;       // occupancy default of burn_kernel
  DesignDev hd{};
  DesignDev* dd = nullptr;
  double* d_betas = nullptr;
  int64_t* d_goff = nullptr;
  DevBuf theta, rates, cost, hist, order, bsum, off, cellrec, cu, cl,
      cb, steps, react, queue, stats, moments, superset, edata, ese, err, cand, ccount, ah, al,
      pflag, bh, bl, cg, gflag, thr, wf, sy, pool, pooln, poolb, presc, apidx, thrlow, spass, tally, mss, mmap, mom, mth, mst, selc, selo, coff, pitems, mpidx, mk1, mcost, mord, dmeta, dw, dboot, dflag, du, dl, dout, dsort, dpt, dxs, selb;
  AcceptState acc;
  hipEvent_t ev[5]{};
  bool timed[4]{};
  // ring of (start, after burn_kernel, after ssa_kernel) events around every
  // simulation; tally[0] counts burn_kernel's draws since the last reset
  static constexpr int kRing = 256;
  hipEvent_t ring[kRing][3]{};
  int ring_used = 0;
  double burn_ms_acc = 0.0, ssa_ms_acc = 0.0;
  int64_t ssa_launches = 0;
  // bumped whenever ensure() replaces a buffer: graphs captured before hold
  // the old address and must not replay
  uint64_t gen = 0;
  // stream capture (abc_graph_begin .. abc_graph_end)
  bool capturing = false;
  bool capture_failed = false;
  std::string msg;
};

struct abc_graph {
  hipGraphExec_t exec = nullptr;
  const abc_ctx* ctx = nullptr;  // the context whose stream was captured
  uint64_t gen = 0;              // its buffer generation at capture
};

namespace {

int fail(abc_ctx* c, int code, const std::string& m) {
  if (c) {
    c->msg = m;
    if (c->capturing) c->capture_failed = true;  // abc_graph_end reports and discards
  }
  return code;
}

// Calls that wait for the device or copy host memory cannot be captured.
#define NO_CAPTURE(ctx)                                                             \
  do {                                                                             \
    if ((ctx)->capturing)                                                          \
      return fail(ctx, ABC_E_STATE, "only *_dev entry points can be captured");    \
  } while (0)

// Timing events are left out of captured graphs.
hipError_t record(abc_ctx* c, hipEvent_t e) {
  return c->capturing ? hipSuccess : hipEventRecord(e, c->stream);
}

#define HIPCHK(ctx, expr)                                                          \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(ctx, ABC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

int ensure(abc_ctx* c, DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (b.bytes >= bytes) return ABC_OK;
  if (c->capturing)
    return fail(c, ABC_E_STATE,
                "a device buffer would grow during graph capture: run the same call once "
                "before capturing it");
  if (b.p) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, ABC_E_HIP, hipGetErrorString(e));
    hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    c->gen++;
  }
  size_t want = bytes + bytes / 4;
  if (hipMalloc(&b.p, want) != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    return fail(c, ABC_E_OOM, "device allocation of " + std::to_string(want) + " bytes failed");
  }
  b.bytes = want;
  return ABC_OK;
}

template <typename T>
T* P(DevBuf& b) { return reinterpret_cast<T*>(b.p); }

// Philox key (seed_lo, seed_hi), the same for every model and particle; the
// model goes to the counter's cell word (cell | model << 24; cells < 2^24)
uint32_t key0(uint64_t seed) { return (uint32_t)seed; }
uint32_t key1(uint64_t seed) { return (uint32_t)(seed >> 32); }
uint32_t mtag(int model) { return (uint32_t)model << 24; }

int check_range(abc_ctx* c, int model, int64_t first, int64_t n) {
  if (model < 1 || model > 5) return fail(c, ABC_E_MODEL, "model must be 1..5");
  if (n < 0 || first < 0 || first + n > 0xFFFFFFFFll)
    return fail(c, ABC_E_ARG, "particle range must lie in [0, 2^32)");
  return ABC_OK;
}

int grid1(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

// grid of the kernels that loop over particles one workgroup at a time: one
// workgroup per particle up to 2^20 (a grid stays far below 2^32 lanes)
unsigned pgrid(int64_t n) { return (unsigned)std::min<int64_t>(n, (int64_t)1 << 20); }

// Device memory one simulation of n particles works in: per-particle rates,
// queue order and item offsets, each cell's burn-in record, sub-lane offsets
// and observed counts, and the counters.  Filled from the context's buffers
// (ws_from_ctx) or carved from one block of memory (ws_carve).
struct SimWs {
  PRates* rates = nullptr;
  float* cost = nullptr;
  unsigned* hist = nullptr;             // 2 * ABC_NBUCKET: histogram, cursors
  int32_t* order = nullptr;
  int64_t* bsum = nullptr;
  int64_t* off = nullptr;               // n + 1
  CellStart* cellrec = nullptr;         // n * C
  int32_t* coff = nullptr;              // n * (C + 1)
  int64_t* pitems = nullptr;
  uint32_t* cu = nullptr;               // n * C, rounded up to whole u64 words
  uint32_t* cl = nullptr;
  uint8_t* cb = nullptr;                // n * C
  unsigned long long* react = nullptr;
  unsigned long long* queue = nullptr;  // 8 words: burn-in and SSA work queues
  unsigned long long* tally = nullptr;  // 2 words: burn-in draws
  unsigned long long* steps = nullptr;  // n: draws per particle when the caller keeps none
  bool timed = true;                    // burn-in / SSA events into the timing ring
};

int64_t cell_words(int64_t n, int C) { return (n * (int64_t)C + 1) / 2; }

// Bytes of a carved workspace for n particles of C cells (each array aligned
// to 256 bytes); with base != NULL the pointers of w are set into it.
size_t ws_carve(char* base, int64_t n, int C, SimWs* w) {
  const size_t nn = (size_t)n, cells = (size_t)n * (size_t)C;
  const int64_t nblk = (n + ABC_SCAN_BLOCK - 1) / ABC_SCAN_BLOCK;
  size_t at = 0;
  auto take = [&](size_t bytes) -> char* {
    char* q = base ? base + at : nullptr;
    at += (std::max<size_t>(bytes, 16) + 255) / 256 * 256;
    return q;
  };
  SimWs v;
  v.rates = reinterpret_cast<PRates*>(take(sizeof(PRates) * nn));
  v.cost = reinterpret_cast<float*>(take(sizeof(float) * nn));
  v.hist = reinterpret_cast<unsigned*>(take(sizeof(unsigned) * 2 * ABC_NBUCKET));
  v.order = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * nn));
  v.bsum = reinterpret_cast<int64_t*>(take(sizeof(int64_t) * (size_t)nblk));
  v.off = reinterpret_cast<int64_t*>(take(sizeof(int64_t) * (nn + 1)));
  v.cellrec = reinterpret_cast<CellStart*>(take(sizeof(CellStart) * cells));
  v.coff = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * nn * (size_t)(C + 1)));
  v.pitems = reinterpret_cast<int64_t*>(take(sizeof(int64_t) * nn));
  v.cu = reinterpret_cast<uint32_t*>(take(sizeof(uint64_t) * (size_t)cell_words(n, C)));
  v.cl = reinterpret_cast<uint32_t*>(take(sizeof(uint64_t) * (size_t)cell_words(n, C)));
  v.cb = reinterpret_cast<uint8_t*>(take(cells));
  v.react = reinterpret_cast<unsigned long long*>(take(sizeof(uint64_t) * nn));
  v.queue = reinterpret_cast<unsigned long long*>(take(64));
  v.tally = reinterpret_cast<unsigned long long*>(take(16));
  v.steps = reinterpret_cast<unsigned long long*>(take(sizeof(uint64_t) * nn));
  if (base && w) *w = v;
  return at;
}

// The context's buffers, grown to n particles.  The burn-in tally is created
// cleared once (abc_ssa_time reads and clears it).
int ws_from_ctx(abc_ctx* c, int64_t n, SimWs* w) {
  int rc;
  const int C = c->hd.C;
  const size_t nn = (size_t)n, cells = (size_t)n * (size_t)C;
  const int64_t nblk = (n + ABC_SCAN_BLOCK - 1) / ABC_SCAN_BLOCK;
  if ((rc = ensure(c, c->rates, sizeof(PRates) * nn))) return rc;
  if ((rc = ensure(c, c->cost, sizeof(float) * nn))) return rc;
  if ((rc = ensure(c, c->hist, sizeof(unsigned) * 2 * ABC_NBUCKET))) return rc;
  if ((rc = ensure(c, c->order, sizeof(int32_t) * nn))) return rc;
  if ((rc = ensure(c, c->bsum, sizeof(int64_t) * (size_t)nblk))) return rc;
  if ((rc = ensure(c, c->off, sizeof(int64_t) * (nn + 1)))) return rc;
  if ((rc = ensure(c, c->cellrec, sizeof(CellStart) * cells))) return rc;
  if ((rc = ensure(c, c->coff, sizeof(int32_t) * nn * (size_t)(C + 1)))) return rc;
  if ((rc = ensure(c, c->pitems, sizeof(int64_t) * nn))) return rc;
  if ((rc = ensure(c, c->cu, sizeof(uint64_t) * (size_t)cell_words(n, C)))) return rc;
  if ((rc = ensure(c, c->cl, sizeof(uint64_t) * (size_t)cell_words(n, C)))) return rc;
  if ((rc = ensure(c, c->cb, cells))) return rc;
  if ((rc = ensure(c, c->react, sizeof(uint64_t) * nn))) return rc;
  if ((rc = ensure(c, c->queue, 64))) return rc;
  if ((rc = ensure(c, c->steps, sizeof(uint64_t) * nn))) return rc;
  if (!c->tally.p) {
    if (c->capturing) return fail(c, ABC_E_STATE, "simulate once before capturing it");
    if ((rc = ensure(c, c->tally, 16))) return rc;
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, c->stream,
                       P<unsigned long long>(c->tally), (int64_t)2);
    HIPCHK(c, hipGetLastError());
  }
  w->rates = P<PRates>(c->rates);
  w->cost = P<float>(c->cost);
  w->hist = P<unsigned>(c->hist);
  w->order = P<int32_t>(c->order);
  w->bsum = P<int64_t>(c->bsum);
  w->off = P<int64_t>(c->off);
  w->cellrec = P<CellStart>(c->cellrec);
  w->coff = P<int32_t>(c->coff);
  w->pitems = P<int64_t>(c->pitems);
  w->cu = P<uint32_t>(c->cu);
  w->cl = P<uint32_t>(c->cl);
  w->cb = P<uint8_t>(c->cb);
  w->react = P<unsigned long long>(c->react);
  w->queue = P<unsigned long long>(c->queue);
  w->tally = P<unsigned long long>(c->tally);
  w->steps = P<unsigned long long>(c->steps);
  w->timed = true;
  return ABC_OK;
}

// rates of n particles of one model into rates/cost from index a on
void launch_rates(abc_ctx* c, const SimWs& w, int model, int64_t n, const double* theta_dev,
                  int64_t a) {
  hipLaunchKernelGGL(rates_kernel, dim3(grid1(n, 256)), dim3(256), 0, c->stream, c->dd, model, n,
                     theta_dev, w.rates + a, w.cost + a);
}

// heaviest-first queue order of all n particles (bucket sort of the cost)
int launch_order(abc_ctx* c, const SimWs& w, int64_t n) {
  unsigned* cursor = w.hist + ABC_NBUCKET;
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(256), 0, c->stream,
                     reinterpret_cast<unsigned long long*>(w.hist), (int64_t)(ABC_NBUCKET / 2));
  hipLaunchKernelGGL(bucket_hist_kernel, dim3(std::min(grid1(n, 256), 1024)), dim3(256), 0,
                     c->stream, w.cost, n, w.hist);
  hipLaunchKernelGGL(bucket_scan_kernel, dim3(1), dim3(ABC_NBUCKET), 0, c->stream, w.hist, cursor);
  hipLaunchKernelGGL(bucket_scatter_kernel, dim3(grid1(n, 256)), dim3(256), 0, c->stream,
                     w.cost, n, cursor, w.order);
  HIPCHK(c, hipGetLastError());
  return ABC_OK;
}

// after the burn-in: each cell's sub-lanes -> per-particle cell offsets and
// item counts -> the item offsets of the particles in queue order
int stage_items(abc_ctx* c, const SimWs& w, int64_t n) {
  const int64_t nblk = (n + ABC_SCAN_BLOCK - 1) / ABC_SCAN_BLOCK;
  const int C = c->hd.C;
  hipLaunchKernelGGL(cell_items_kernel, dim3(pgrid(n)), dim3(256), 0, c->stream, n, C, w.cellrec,
                     w.coff, w.pitems);
  hipLaunchKernelGGL(items_scan_kernel, dim3((unsigned)nblk), dim3(256), 0, c->stream, w.pitems,
                     w.order, n, w.off, w.bsum);
  hipLaunchKernelGGL(bsum_scan_kernel, dim3(1), dim3(256), 0, c->stream, w.bsum, nblk);
  hipLaunchKernelGGL(add_bsum_kernel, dim3(grid1(n, 256)), dim3(256), 0, c->stream, w.off, n,
                     w.bsum);
  HIPCHK(c, hipGetLastError());
  return ABC_OK;
}

// fold the pending burn_kernel / ssa_kernel event triples into the running totals
int drain_ring(abc_ctx* c) {
  for (int q = 0; q < c->ring_used; ++q) {
    float vb = 0.f, vs = 0.f;
    HIPCHK(c, hipEventSynchronize(c->ring[q][2]));
    HIPCHK(c, hipEventElapsedTime(&vb, c->ring[q][0], c->ring[q][1]));
    HIPCHK(c, hipEventElapsedTime(&vs, c->ring[q][1], c->ring[q][2]));
    c->burn_ms_acc += vb;
    c->ssa_ms_acc += vs;
    c->ssa_launches++;
  }
  c->ring_used = 0;
  return ABC_OK;
}

// cells of n particles (rates and queue order staged in w): burn-in, sub-lane
// offsets, the window SSA.  Particle p is global particle pidx[p] (first + p
// without pidx) with model tag mtags[p] (mtag_all without mtags).  Cell
// counts go to w.cu / w.cl / w.cb, Philox draws per particle to steps.
int simulate_core(abc_ctx* c, const SimWs& w, uint64_t seed, int64_t n, int64_t first,
                  uint32_t mtag_all, const uint32_t* mtags, const int64_t* pidx_dev,
                  unsigned long long* steps) {
  int rc;
  const int64_t words = cell_words(n, c->hd.C);  // u32 counts cleared as u64 words
  // counters cleared by a kernel (no memset nodes, so the sequence can be captured)
  hipLaunchKernelGGL(zero_u64_kernel, dim3(grid1(words, 256)), dim3(256), 0, c->stream,
                     reinterpret_cast<unsigned long long*>(w.cu), words);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(grid1(words, 256)), dim3(256), 0, c->stream,
                     reinterpret_cast<unsigned long long*>(w.cl), words);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(grid1(n, 256)), dim3(256), 0, c->stream, steps, n);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(grid1(n, 256)), dim3(256), 0, c->stream, w.react, n);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, c->stream, w.queue, (int64_t)8);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, record(c, c->ev[1]));
  hipEvent_t* trio = nullptr;
  if (!c->capturing && w.timed) {
    if (c->ring_used == abc_ctx::kRing) {
      if ((rc = drain_ring(c))) return rc;
    }
    trio = c->ring[c->ring_used++];
    HIPCHK(c, hipEventRecord(trio[0], c->stream));
  }
  const int blocks = c->ssa_blocks > 0 ? c->ssa_blocks : c->ssa_blocks_default;
  hipLaunchKernelGGL(burn_kernel, dim3(c->burn_blocks), dim3(SSA_WG), 0, c->stream, c->dd,
                     w.rates, w.order, n, (uint32_t)first, key0(seed), key1(seed), mtag_all,
                     mtags, c->d_betas, c->d_goff, w.queue, w.cellrec, steps, w.react, pidx_dev,
                     w.tally);
  HIPCHK(c, hipGetLastError());
  if ((rc = stage_items(c, w, n))) return rc;
  if (trio) HIPCHK(c, hipEventRecord(trio[1], c->stream));
  hipLaunchKernelGGL(ssa_kernel, dim3(blocks), dim3(SSA_WG), 0, c->stream, c->dd, w.rates,
                     w.order, w.off, n, (uint32_t)first, key0(seed), key1(seed), mtag_all, mtags,
                     w.cellrec, w.coff, w.queue + 1, w.cu, w.cl, w.cb, steps, pidx_dev,
                     (uint32_t)c->ssa_claim);
  HIPCHK(c, hipGetLastError());
  if (trio) HIPCHK(c, hipEventRecord(trio[2], c->stream));
  HIPCHK(c, record(c, c->ev[2]));
  return ABC_OK;
}

// One model's particles through rates, queue order, burn-in, SSA and the
// statistics, in workspace w.
int simulate_ws(abc_ctx* c, const SimWs& w, int model, uint64_t seed, int64_t first, int64_t n,
                const double* theta_dev, double* stats_dev, double* moments_dev,
                double* superset_dev, int64_t* steps_dev, const int64_t* pidx_dev) {
  int rc;
  HIPCHK(c, record(c, c->ev[0]));
  launch_rates(c, w, model, n, theta_dev, 0);
  if ((rc = launch_order(c, w, n))) return rc;
  unsigned long long* steps =
      steps_dev ? reinterpret_cast<unsigned long long*>(steps_dev) : w.steps;
  if ((rc = simulate_core(c, w, seed, n, first, mtag(model), nullptr, pidx_dev, steps)))
    return rc;
  hipLaunchKernelGGL(stats_kernel, dim3(pgrid(n)), dim3(256), 0, c->stream, c->dd, w.cu, w.cl,
                     w.cb, n, stats_dev, moments_dev, superset_dev);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, record(c, c->ev[3]));
  c->timed[0] = c->timed[1] = c->timed[2] = !c->capturing;
  return ABC_OK;
}

int simulate_dev(abc_ctx* c, int model, uint64_t seed, int64_t first, int64_t n,
                 const double* theta_dev, double* stats_dev, double* moments_dev,
                 double* superset_dev, int64_t* steps_dev, const int64_t* pidx_dev = nullptr) {
  int rc;
  if ((rc = check_range(c, model, first, n))) return rc;
  if (n == 0) return ABC_OK;
  if (n > 0x7FFFFFFF) return fail(c, ABC_E_ARG, "at most 2^31 - 1 particles per call");
  SimWs w;
  if ((rc = ws_from_ctx(c, n, &w))) return rc;
  return simulate_ws(c, w, model, seed, first, n, theta_dev, stats_dev, moments_dev,
                     superset_dev, steps_dev, pidx_dev);
}

int make_spec(abc_ctx* c, int K, const int32_t* parts, int n_parts, double clamp, ErrSpec* E) {
  if (K <= 0 || K > ABC_MAX_STATS) return fail(c, ABC_E_ARG, "K must be 1..256");
  if (n_parts <= 0 || n_parts > ABC_MAX_PARTS || !parts)
    return fail(c, ABC_E_ARG, "n_parts must be 1..16");
  int s = 0;
  for (int q = 0; q < n_parts; ++q) {
    if (parts[q] < 0) return fail(c, ABC_E_ARG, "negative part size");
    E->parts[q] = parts[q];
    s += parts[q];
  }
  if (s != K) return fail(c, ABC_E_ARG, "part sizes must sum to K");
  E->K = K;
  E->n_parts = n_parts;
  E->clamp = clamp;
  return ABC_OK;
}

}  // namespace

extern "C" {

int32_t abc_create(abc_ctx** out, int32_t device, const abc_design* d) {
  if (!out || !d) return ABC_E_ARG;
  *out = nullptr;
  abc_ctx* c = new abc_ctx();
  auto bad = [&](const char* m) {
    c->msg = m;
    return ABC_E_ARG;
  };
  // validation keeps every kernel index in bounds
  if (!(d->cycle_h > 0.0)) { int r = bad("cycle_h must be > 0"); delete c; return r; }
  if (d->n_bins < 1 || d->n_bins > ABC_MAX_BINS) { delete c; return ABC_E_ARG; }
  if (d->n_conds < 1 || d->n_conds > ABC_MAX_CONDS) { delete c; return ABC_E_ARG; }
  if (d->cells_per_particle < 1 || d->cells_per_particle > (1 << 24)) { delete c; return ABC_E_ARG; }
  if (!d->pulse_h || !d->chase_h || !d->is_chase || !d->age_weights) { delete c; return ABC_E_ARG; }
  if (d->age_mode != 0 && d->age_mode != 1) { delete c; return ABC_E_ARG; }
  if (d->capture_mode < 0 || d->capture_mode > 2) { delete c; return ABC_E_ARG; }
  if (d->capture_mode == 1 &&
      (d->beta_a < 1 || d->beta_a > 4 || d->beta_b < 1 || d->beta_a + d->beta_b - 1 > 32)) {
    delete c; return ABC_E_ARG;
  }
  if (d->prior_mode != 0 && d->prior_mode != 1) { delete c; return ABC_E_ARG; }
  if (!(d->burn_eps > 0.0) || d->burn_max < 1 || d->burn_max > 64) { delete c; return ABC_E_ARG; }
  DesignDev& h = c->hd;
  h.T = d->cycle_h;
  h.nb = d->n_bins;
  h.J = d->n_conds;
  h.C = d->cells_per_particle;
  h.age_mode = d->age_mode;
  h.cap_mode = d->capture_mode;
  h.beta_a = d->beta_a;
  h.beta_b = d->beta_b;
  h.prior_mode = d->prior_mode;
  h.burn_eps = d->burn_eps;
  h.splice_log10 = d->splice_log10;
  if (d->split_target < 0.0 || d->split_max < 1 || d->split_max > 4096) { delete c; return ABC_E_ARG; }
  // a cell's sub-lanes reach at most 2 * split_max - 1 (powers of two up to
  // the cap); a particle's item offsets are 32-bit
  if ((int64_t)d->cells_per_particle * (2 * (int64_t)d->split_max) > 0x7FFFFFFFll) {
    delete c;
    return ABC_E_ARG;
  }
  h.split_target = d->split_target;
  h.split_max = d->split_max;
  h.split_scale = d->capture_mode == 1 ? (double)d->beta_a / (double)(d->beta_a + d->beta_b)
                                       : (d->capture_mode == 2 ? 0.25 : 1.0);
  double wmax = 0.0;
  for (int j = 0; j < h.J; ++j) {
    if (!(d->pulse_h[j] > 0.0) || d->chase_h[j] < 0.0 || (d->is_chase[j] != 0 && d->is_chase[j] != 1)) {
      delete c; return ABC_E_ARG;
    }
    h.pulse[j] = d->pulse_h[j];
    h.chase[j] = d->chase_h[j];
    h.is_chase[j] = d->is_chase[j];
    wmax = std::max(wmax, d->pulse_h[j] + d->chase_h[j]);
  }
  h.wmax = wmax;
  h.bmin = std::max(1, (int)std::ceil(wmax / h.T));
  h.bmax = std::max(h.bmin, (int)d->burn_max);
  for (int j = 0; j < h.nb; ++j) h.bound[j] = (h.T * (double)(h.nb + j + 1)) / (double)h.nb;
  for (int q = 0; q < h.nb * h.J; ++q) h.weights[q] = d->age_weights[q];
  // capture: grouped empirical list (group = bin*2 + is_chase)
  std::vector<double> grouped;
  std::vector<int64_t> goff(2 * h.nb + 1, 0);
  if (h.cap_mode == 2) {
    if (!d->betas || !d->beta_bin || !d->beta_is_chase || d->n_betas <= 0) { delete c; return ABC_E_ARG; }
    std::vector<std::vector<double>> g(2 * h.nb);
    for (int64_t q = 0; q < d->n_betas; ++q) {
      int b = d->beta_bin[q], ch = d->beta_is_chase[q];
      if (b < 0 || b >= h.nb || (ch != 0 && ch != 1)) { delete c; return ABC_E_ARG; }
      g[b * 2 + ch].push_back(d->betas[q]);
    }
    double msum = 0.0;
    for (int q = 0; q < 2 * h.nb; ++q) {
      // every (bin, group) used by a condition needs at least one efficiency
      bool used = false;
      for (int j = 0; j < h.J; ++j) if (h.is_chase[j] == (q & 1)) used = true;
      if (used && g[q].empty()) {
        c->msg = "capture list has no efficiency for an age bin / group";
        delete c; return ABC_E_ARG;
      }
      goff[q + 1] = goff[q] + (int64_t)g[q].size();
      grouped.insert(grouped.end(), g[q].begin(), g[q].end());
      for (double x : g[q]) msum += x;
    }
    h.capture_mean = msum / (double)grouped.size();
  } else if (h.cap_mode == 1) {
    h.capture_mean = (double)h.beta_a / (double)(h.beta_a + h.beta_b);
  } else {
    h.capture_mean = 1.0;
  }
  if (grouped.empty()) grouped.push_back(0.0);
  c->device = device;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) { delete c; return ABC_E_HIP; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete c; return ABC_E_HIP; }
  c->num_cu = prop.multiProcessorCount;
  if (scratch_address(device, &c->scratch) != hipSuccess || !c->scratch) {
    (void)hipGetLastError();
    delete c;
    return ABC_E_HIP;
  }
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ssa_kernel, SSA_WG, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  c->ssa_blocks_default = c->num_cu * per_cu;
  int per_cu_burn = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_burn, burn_kernel, SSA_WG, 0) !=
          hipSuccess ||
      per_cu_burn < 1)
    per_cu_burn = 1;
  c->burn_blocks = c->num_cu * per_cu_burn;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return ABC_E_HIP; }
  for (auto& ev : c->ev)
    if (hipEventCreate(&ev) != hipSuccess) { abc_destroy(c); return ABC_E_HIP; }
  for (auto& pr : c->ring)
    for (auto& ev : pr)
      if (hipEventCreate(&ev) != hipSuccess) { abc_destroy(c); return ABC_E_HIP; }
  if (hipMalloc(&c->dd, sizeof(DesignDev)) != hipSuccess ||
      hipMalloc(&c->d_betas, sizeof(double) * grouped.size()) != hipSuccess ||
      hipMalloc(&c->d_goff, sizeof(int64_t) * goff.size()) != hipSuccess) {
    abc_destroy(c);
    return ABC_E_OOM;
  }
  if (hipMemcpy(c->dd, &h, sizeof(DesignDev), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_betas, grouped.data(), sizeof(double) * grouped.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_goff, goff.data(), sizeof(int64_t) * goff.size(), hipMemcpyHostToDevice) != hipSuccess) {
    abc_destroy(c);
    return ABC_E_HIP;
  }
  *out = c;
  return ABC_OK;
}

int32_t abc_destroy(abc_ctx* c) {
  if (!c) return ABC_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  DevBuf* bufs[] = {&c->theta, &c->rates, &c->cost, &c->hist, &c->order, &c->bsum, &c->off,
                    &c->cellrec,
                    &c->cu, &c->cl, &c->cb, &c->steps, &c->react, &c->queue,
                    &c->stats, &c->moments, &c->superset, &c->edata, &c->ese, &c->err,
                    &c->cand, &c->ccount, &c->ah, &c->al, &c->pflag, &c->bh, &c->bl,
                    &c->cg, &c->gflag, &c->thr, &c->wf, &c->sy, &c->tally, &c->mss, &c->mmap, &c->coff, &c->pitems, &c->mpidx, &c->mk1, &c->mcost, &c->mord, &c->dmeta, &c->dw,
                    &c->dboot, &c->dflag, &c->du, &c->dl, &c->dout, &c->dsort, &c->dpt, &c->dxs, &c->selb,
                    &c->pool, &c->pooln, &c->poolb, &c->presc, &c->apidx, &c->thrlow, &c->spass,
                    &c->mom, &c->mth, &c->mst, &c->selc, &c->selo};
  for (DevBuf* b : bufs)
    if (b->p) hipFree(b->p);
  if (c->dd) hipFree(c->dd);
  if (c->d_betas) hipFree(c->d_betas);
  if (c->d_goff) hipFree(c->d_goff);
  for (auto& ev : c->ev)
    if (ev) hipEventDestroy(ev);
  for (auto& pr : c->ring)
    for (auto& ev : pr)
      if (ev) hipEventDestroy(ev);
  if (c->stream && c->own_stream) hipStreamDestroy(c->stream);
  delete c;
  return ABC_OK;
}

const char* abc_last_error(const abc_ctx* c) { return c ? c->msg.c_str() : "null context"; }

int32_t abc_n_params(const abc_ctx* c, int32_t model) {
  if (!c || model < 1 || model > 5) return ABC_E_ARG;
  return abc_layout_n_params(model, c->hd.nb);
}

int32_t abc_n_stats(const abc_ctx* c) { return c ? 4 * c->hd.nb + 3 * c->hd.J : ABC_E_ARG; }

int32_t abc_set_stream(abc_ctx* c, void* s) {
  if (!c) return ABC_E_ARG;
  NO_CAPTURE(c);
  hipStreamSynchronize(c->stream);
  if (c->own_stream) hipStreamDestroy(c->stream);
  if (s) {
    c->stream = (hipStream_t)s;
    c->own_stream = false;
  } else {
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return ABC_E_HIP;
    c->own_stream = true;
  }
  return ABC_OK;
}

int32_t abc_synchronize(abc_ctx* c) {
  if (!c) return ABC_E_ARG;
  NO_CAPTURE(c);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ABC_OK;
}

int32_t abc_set_ssa_blocks(abc_ctx* c, int32_t blocks) {
  if (!c || blocks < 0) return ABC_E_ARG;
  c->ssa_blocks = blocks;
  return ABC_OK;
}

int32_t abc_set_ssa_claim(abc_ctx* c, int32_t claim) {
  if (!c) return ABC_E_ARG;
  if (claim != 1 && claim != 2 && claim != 4 && claim != 8)
    return fail(c, ABC_E_ARG, "claim must be 1, 2, 4 or 8");
  c->ssa_claim = claim;
  return ABC_OK;
}

int32_t abc_get_ssa_claim(const abc_ctx* c) { return c ? c->ssa_claim : ABC_E_ARG; }

int32_t abc_get_ssa_blocks(const abc_ctx* c) {
  if (!c) return ABC_E_ARG;
  return c->ssa_blocks > 0 ? c->ssa_blocks : c->ssa_blocks_default;
}

int32_t abc_sample_prior_dev(abc_ctx* c, int32_t model, uint64_t seed, int64_t first, int64_t n,
                             double* theta_dev) {
  if (!c) return ABC_E_ARG;
  int rc = check_range(c, model, first, n);
  if (rc) return rc;
  if (n == 0) return ABC_OK;
  hipLaunchKernelGGL(prior_kernel, dim3(grid1(n, 256)), dim3(256), 0, c->stream, c->dd, model,
                     key0(seed), key1(seed), (uint32_t)first, n, theta_dev);
  HIPCHK(c, hipGetLastError());
  return ABC_OK;
}

int32_t abc_sample_prior(abc_ctx* c, int32_t model, uint64_t seed, int64_t first, int64_t n,
                         double* theta) {
  if (!c || (!theta && n > 0)) return ABC_E_ARG;
  NO_CAPTURE(c);
  int rc = check_range(c, model, first, n);
  if (rc) return rc;
  if (n == 0) return ABC_OK;
  const int np = abc_layout_n_params(model, c->hd.nb);
  if ((rc = ensure(c, c->theta, sizeof(double) * n * np))) return rc;
  if ((rc = abc_sample_prior_dev(c, model, seed, first, n, P<double>(c->theta)))) return rc;
  HIPCHK(c, hipMemcpyAsync(theta, c->theta.p, sizeof(double) * n * np, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ABC_OK;
}

int32_t abc_math_probe_dev(abc_ctx* c, int32_t kind, int64_t n, const double* x_dev,
                           double* y_dev) {
  if (!c || kind < 0 || kind > 2 || n < 0 || (n > 0 && (!x_dev || !y_dev))) return ABC_E_ARG;
  if (n == 0) return ABC_OK;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(math_probe_kernel, dim3((unsigned)blocks), dim3(256), 0, c->stream, kind, n,
                     x_dev, y_dev);
  HIPCHK(c, hipGetLastError());
  return ABC_OK;
}

int32_t abc_simulate_dev(abc_ctx* c, int32_t model, uint64_t seed, int64_t first, int64_t n,
                         const double* theta_dev, double* stats_dev, double* moments_dev,
                         double* superset_dev, int64_t* steps_dev) {
  if (!c || (n > 0 && (!theta_dev || !stats_dev))) return ABC_E_ARG;
  return simulate_dev(c, model, seed, first, n, theta_dev, stats_dev, moments_dev, superset_dev,
                      steps_dev);
}

}  // extern "C"

namespace {

// The particles of several models through one burn-in and one SSA launch:
// model m's particles are first[m] + i, or pidx_dev[m][i] when pidx_dev is
// given (then first is not read).
int simulate_multi(abc_ctx* c, int32_t n_models, const int32_t* models, uint64_t seed,
                   const int64_t* first, const int64_t* const* pidx_dev, const int64_t* n,
                   const double* const* theta_dev, double* const* stats_dev,
                   double* const* moments_dev, double* const* superset_dev,
                   int64_t* const* steps_dev) {
  if (n_models < 1 || n_models > 16 || !models || !(first || pidx_dev) || !n || !theta_dev ||
      !stats_dev)
    return fail(c, ABC_E_ARG, "need 1..16 models with first (or pidx), n, theta and stats");
  int rc;
  int64_t total = 0, a[17];
  for (int m = 0; m < n_models; ++m) {
    if ((rc = check_range(c, models[m], pidx_dev ? 0 : first[m], n[m]))) return rc;
    if (n[m] > 0 && (!theta_dev[m] || !stats_dev[m] || (pidx_dev && !pidx_dev[m])))
      return fail(c, ABC_E_ARG,
                  "theta and stats (and pidx) are required for every model with particles");
    a[m] = total;
    total += n[m];
  }
  a[n_models] = total;
  if (total == 0) return ABC_OK;
  if (total > 0x7FFFFFFF) return fail(c, ABC_E_ARG, "at most 2^31 - 1 particles per call");
  HIPCHK(c, record(c, c->ev[0]));
  SimWs w;
  if ((rc = ws_from_ctx(c, total, &w))) return rc;
  if ((rc = ensure(c, c->mpidx, sizeof(int64_t) * total))) return rc;
  if ((rc = ensure(c, c->mk1, sizeof(uint32_t) * total))) return rc;
  for (int m = 0; m < n_models; ++m) {
    if (n[m] == 0) continue;
    launch_rates(c, w, models[m], n[m], theta_dev[m], a[m]);
    hipLaunchKernelGGL(fill_ids_kernel, dim3(grid1(n[m], 256)), dim3(256), 0, c->stream, n[m],
                       pidx_dev ? (int64_t)0 : first[m], pidx_dev ? pidx_dev[m] : nullptr,
                       mtag(models[m]), P<int64_t>(c->mpidx) + a[m], P<uint32_t>(c->mk1) + a[m]);
  }
  HIPCHK(c, hipGetLastError());
  if ((rc = launch_order(c, w, total))) return rc;
  unsigned long long* steps = w.steps;
  if ((rc = simulate_core(c, w, seed, total, 0, 0u, P<uint32_t>(c->mk1), P<int64_t>(c->mpidx),
                          steps)))
    return rc;
  const size_t C = (size_t)c->hd.C;
  for (int m = 0; m < n_models; ++m) {
    if (n[m] == 0) continue;
    const size_t o = (size_t)a[m] * C;
    hipLaunchKernelGGL(stats_kernel, dim3(pgrid(n[m])), dim3(256), 0, c->stream, c->dd,
                       w.cu + o, w.cl + o, w.cb + o, n[m],
                       stats_dev[m], moments_dev ? moments_dev[m] : nullptr,
                       superset_dev ? superset_dev[m] : nullptr);
    if (steps_dev && steps_dev[m])
      hipLaunchKernelGGL(copy_u64_kernel, dim3(grid1(n[m], 256)), dim3(256), 0, c->stream,
                         steps + a[m], reinterpret_cast<unsigned long long*>(steps_dev[m]), n[m]);
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, record(c, c->ev[3]));
  c->timed[0] = c->timed[1] = c->timed[2] = !c->capturing;
  return ABC_OK;
}

}  // namespace

extern "C" {

int32_t abc_simulate_multi_dev(abc_ctx* c, int32_t n_models, const int32_t* models, uint64_t seed,
                               const int64_t* first, const int64_t* n,
                               const double* const* theta_dev, double* const* stats_dev,
                               double* const* moments_dev, double* const* superset_dev,
                               int64_t* const* steps_dev) {
  if (!c) return ABC_E_ARG;
  if (!first) return fail(c, ABC_E_ARG, "first is required");
  return simulate_multi(c, n_models, models, seed, first, nullptr, n, theta_dev, stats_dev,
                        moments_dev, superset_dev, steps_dev);
}

int32_t abc_simulate_multi_idx_dev(abc_ctx* c, int32_t n_models, const int32_t* models,
                                   uint64_t seed, const int64_t* n,
                                   const int64_t* const* pidx_dev,
                                   const double* const* theta_dev, double* const* stats_dev,
                                   double* const* moments_dev, double* const* superset_dev,
                                   int64_t* const* steps_dev) {
  if (!c) return ABC_E_ARG;
  if (!pidx_dev) return fail(c, ABC_E_ARG, "pidx is required");
  return simulate_multi(c, n_models, models, seed, nullptr, pidx_dev, n, theta_dev, stats_dev,
                        moments_dev, superset_dev, steps_dev);
}

int32_t abc_simulate_noalloc_dev(abc_ctx* c, int32_t model, uint64_t seed, int64_t first,
                                 int64_t n, const double* theta_dev, double* stats_dev,
                                 double* moments_dev, double* superset_dev, int64_t* steps_dev) {
  if (!c || (n > 0 && (!theta_dev || !stats_dev))) return ABC_E_ARG;
  int rc;
  if ((rc = check_range(c, model, first, n))) return rc;
  if (n == 0) return ABC_OK;
  const int C = c->hd.C;
  if (ws_carve(nullptr, n, C, nullptr) > ABC_NOALLOC_SCRATCH_BYTES) {
    int64_t lo = 0, hi = (int64_t)1 << 31;  // the largest n whose workspace fits
    while (hi - lo > 1) {
      const int64_t mid = lo + (hi - lo) / 2;
      (ws_carve(nullptr, mid, C, nullptr) <= ABC_NOALLOC_SCRATCH_BYTES ? lo : hi) = mid;
    }
    return fail(c, ABC_E_ARG,
                "abc_simulate_noalloc_dev takes at most " + std::to_string(lo) +
                    " particles of " + std::to_string(C) + " cells per call (" +
                    std::to_string(ABC_NOALLOC_SCRATCH_BYTES >> 20) + " MiB of scratch)");
  }
  SimWs w;
  ws_carve(static_cast<char*>(c->scratch), n, C, &w);
  w.timed = false;  // its burn-in draws are not in the context's tally
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, c->stream, w.tally, (int64_t)2);
  HIPCHK(c, hipGetLastError());
  return simulate_ws(c, w, model, seed, first, n, theta_dev, stats_dev, moments_dev,
                     superset_dev, steps_dev, nullptr);
}

int32_t abc_simulate_idx_dev(abc_ctx* c, int32_t model, uint64_t seed, int64_t n,
                             const int64_t* pidx_dev, const double* theta_dev, double* stats_dev,
                             double* moments_dev, double* superset_dev, int64_t* steps_dev) {
  if (!c || (n > 0 && (!theta_dev || !stats_dev || !pidx_dev))) return ABC_E_ARG;
  return simulate_dev(c, model, seed, 0, n, theta_dev, stats_dev, moments_dev, superset_dev,
                      steps_dev, pidx_dev);
}

int32_t abc_simulate(abc_ctx* c, int32_t model, uint64_t seed, int64_t first, int64_t n,
                     const double* theta, double* stats, double* moments, double* superset,
                     int64_t* steps, uint32_t* cell_u, uint32_t* cell_l, uint8_t* cell_bin) {
  if (!c || (n > 0 && (!theta || !stats))) return ABC_E_ARG;
  NO_CAPTURE(c);
  int rc = check_range(c, model, first, n);
  if (rc) return rc;
  if (n == 0) return ABC_OK;
  const int np = abc_layout_n_params(model, c->hd.nb);
  const int K = 4 * c->hd.nb + 3 * c->hd.J, NBJ = c->hd.nb * c->hd.J;
  if ((rc = ensure(c, c->theta, sizeof(double) * n * np))) return rc;
  if ((rc = ensure(c, c->stats, sizeof(double) * n * K))) return rc;
  if (moments && (rc = ensure(c, c->moments, sizeof(double) * n * NBJ * 6))) return rc;
  if (superset && (rc = ensure(c, c->superset, sizeof(double) * n * NBJ * 2))) return rc;
  if ((rc = ensure(c, c->steps, sizeof(uint64_t) * n))) return rc;
  HIPCHK(c, hipMemcpyAsync(c->theta.p, theta, sizeof(double) * n * np, hipMemcpyHostToDevice, c->stream));
  rc = simulate_dev(c, model, seed, first, n, P<double>(c->theta), P<double>(c->stats),
                    moments ? P<double>(c->moments) : nullptr,
                    superset ? P<double>(c->superset) : nullptr, P<int64_t>(c->steps));
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(stats, c->stats.p, sizeof(double) * n * K, hipMemcpyDeviceToHost, c->stream));
  if (moments)
    HIPCHK(c, hipMemcpyAsync(moments, c->moments.p, sizeof(double) * n * NBJ * 6, hipMemcpyDeviceToHost, c->stream));
  if (superset)
    HIPCHK(c, hipMemcpyAsync(superset, c->superset.p, sizeof(double) * n * NBJ * 2, hipMemcpyDeviceToHost, c->stream));
  if (steps)
    HIPCHK(c, hipMemcpyAsync(steps, c->steps.p, sizeof(int64_t) * n, hipMemcpyDeviceToHost, c->stream));
  const size_t cells = (size_t)n * (size_t)c->hd.C;
  if (cell_u) HIPCHK(c, hipMemcpyAsync(cell_u, c->cu.p, 4 * cells, hipMemcpyDeviceToHost, c->stream));
  if (cell_l) HIPCHK(c, hipMemcpyAsync(cell_l, c->cl.p, 4 * cells, hipMemcpyDeviceToHost, c->stream));
  if (cell_bin) HIPCHK(c, hipMemcpyAsync(cell_bin, c->cb.p, cells, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ABC_OK;
}

int32_t abc_errors_dev(abc_ctx* c, int64_t n, int32_t K, const double* stats_dev, int64_t G,
                       const double* data_dev, const double* se_dev, const int32_t* parts,
                       int32_t n_parts, double clamp, double* err_dev) {
  if (!c || n < 0 || G < 0) return ABC_E_ARG;
  ErrSpec E;
  int rc = make_spec(c, K, parts, n_parts, clamp, &E);
  if (rc) return rc;
  if (n == 0 || G == 0) return ABC_OK;
  if (G > 65535 * 64) return fail(c, ABC_E_ARG, "too many genes for one call");
  HIPCHK(c, record(c, c->ev[3]));
  // grid.y is limited to 65535: split genes into slices
  for (int64_t g0 = 0; g0 < G; g0 += 65535) {
    const int64_t gn = std::min<int64_t>(65535, G - g0);
    hipLaunchKernelGGL(err_fp64_kernel, dim3(grid1(n, 256), (unsigned)gn), dim3(256), 0, c->stream,
                       stats_dev, n, data_dev, se_dev, G, g0, E, err_dev + n * g0);
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, record(c, c->ev[4]));
  c->timed[3] = !c->capturing;
  return ABC_OK;
}

int32_t abc_errors(abc_ctx* c, int64_t n, int32_t K, const double* stats, int64_t G,
                   const double* data, const double* se, const int32_t* parts, int32_t n_parts,
                   double clamp, double* err) {
  if (!c || n < 0 || G < 0 || (n > 0 && G > 0 && (!stats || !data || !se || !err))) return ABC_E_ARG;
  NO_CAPTURE(c);
  ErrSpec E;
  int rc = make_spec(c, K, parts, n_parts, clamp, &E);
  if (rc) return rc;
  if (n == 0 || G == 0) return ABC_OK;
  if ((rc = ensure(c, c->stats, sizeof(double) * n * K))) return rc;
  if ((rc = ensure(c, c->edata, sizeof(double) * G * K))) return rc;
  if ((rc = ensure(c, c->ese, sizeof(double) * G * K))) return rc;
  if ((rc = ensure(c, c->err, sizeof(double) * n * G))) return rc;
  HIPCHK(c, hipMemcpyAsync(c->stats.p, stats, sizeof(double) * n * K, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->edata.p, data, sizeof(double) * G * K, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->ese.p, se, sizeof(double) * G * K, hipMemcpyHostToDevice, c->stream));
  for (int64_t g0 = 0; g0 < G; g0 += 65535) {
    const int64_t gn = std::min<int64_t>(65535, G - g0);
    hipLaunchKernelGGL(err_fp64_kernel, dim3(grid1(n, 256), (unsigned)gn), dim3(256), 0, c->stream,
                       P<double>(c->stats), n, P<double>(c->edata), P<double>(c->ese), G, g0, E,
                       P<double>(c->err) + n * g0);
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(err, c->err.p, sizeof(double) * n * G, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ABC_OK;
}

}  // extern "C"

// ------------------------------------------------------------ accept session
namespace {

struct GeneHit {
  double err;
  int64_t idx;  // global 0-based particle index
  bool operator<(const GeneHit& o) const { return err < o.err || (err == o.err && idx < o.idx); }
};

int64_t pad_to(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Top-k batches (mode 1): screened, rescored and selected per gene entirely
// on the device, n_sub rows at a time, with no host wait.  pidx_host (n
// global indices) is staged through device memory; without it row i is
// particle first_index + i.
int screen_batch_pooled(abc_ctx* c, int64_t n, const double* stats_dev, int64_t first_index,
                        const int64_t* pidx_host) {
  AcceptState& A = c->acc;
  const int grid = (int)std::min<int64_t>(A.G, (int64_t)c->num_cu * 8);
  for (int64_t off = 0; off < n; off += A.n_sub) {
    const int64_t nb = std::min(A.n_sub, n - off);
    const int64_t npad = pad_to(nb, 64);
    const int64_t* pidx_dev = nullptr;
    if (pidx_host) {
      HIPCHK(c, hipMemcpyAsync(c->apidx.p, pidx_host + off, sizeof(int64_t) * nb,
                               hipMemcpyHostToDevice, c->stream));
      pidx_dev = P<int64_t>(c->apidx);
    }
    hipLaunchKernelGGL(particle_prep_kernel, dim3(grid1(npad, 256)), dim3(256), 0, c->stream,
                       stats_dev + off, nb, n, A.E.K, A.Kp, npad, P<__bf16>(c->ah),
                       P<__bf16>(c->al), P<uint8_t>(c->pflag));
    hipLaunchKernelGGL(gemm_screen_kernel, dim3((unsigned)(npad / 64), (unsigned)(A.Gpad / 64)),
                       dim3(256), 0, c->stream, P<__bf16>(c->ah), P<__bf16>(c->al), nb, A.Kp,
                       P<__bf16>(c->bh), P<__bf16>(c->bl), A.G, P<double>(c->cg),
                       P<uint8_t>(c->gflag), P<uint8_t>(c->pflag), P<double>(c->thr),
                       A.margin_rel, A.E.clamp, (Candidate*)nullptr,
                       (unsigned long long*)nullptr, 0ull, (float*)nullptr, (float*)nullptr,
                       P<PoolEntry>(c->pool), A.capg, P<uint32_t>(c->pooln));
    // four workgroups share each gene's new entries
    hipLaunchKernelGGL(pool_rescore_kernel, dim3(grid, 4), dim3(256), 0, c->stream,
                       stats_dev + off, n, P<double>(c->edata), P<double>(c->ese), A.G, A.E,
                       P<PoolEntry>(c->pool), A.capg, P<uint32_t>(c->pooln),
                       P<uint32_t>(c->poolb), first_index + off, pidx_dev);
    hipLaunchKernelGGL(pool_select_kernel, dim3(grid), dim3(256), 0, c->stream,
                       P<PoolEntry>(c->pool), A.capg, P<uint32_t>(c->pooln),
                       P<uint32_t>(c->poolb), A.G, A.k, P<double>(c->thr),
                       P<unsigned long long>(c->presc));
    HIPCHK(c, hipGetLastError());
    if (pidx_host) HIPCHK(c, hipStreamSynchronize(c->stream));  // apidx is reused
  }
  A.n_screened += n;
  return ABC_OK;
}

int screen_batch(abc_ctx* c, int64_t n, const double* stats_dev, int64_t first_index,
                 const int64_t* pidx_host = nullptr) {
  AcceptState& A = c->acc;
  if (A.pooled) return screen_batch_pooled(c, n, stats_dev, first_index, pidx_host);
  int rc;
  const int64_t npad = pad_to(n, 64);
  if ((rc = ensure(c, c->ah, sizeof(uint16_t) * npad * 2 * A.Kp))) return rc;
  if ((rc = ensure(c, c->al, sizeof(uint16_t) * npad * 2 * A.Kp))) return rc;
  if ((rc = ensure(c, c->pflag, npad))) return rc;
  if ((rc = ensure(c, c->ccount, 64))) return rc;
  hipLaunchKernelGGL(particle_prep_kernel, dim3(grid1(npad, 256)), dim3(256), 0, c->stream,
                     stats_dev, n, n, A.E.K, A.Kp, npad, P<__bf16>(c->ah), P<__bf16>(c->al),
                     P<uint8_t>(c->pflag));
  HIPCHK(c, hipGetLastError());
  unsigned long long cap = std::max<unsigned long long>(A.cand_cap, 1ull << 16);
  for (;;) {
    if ((rc = ensure(c, c->cand, sizeof(Candidate) * cap))) return rc;
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, c->stream,
                       P<unsigned long long>(c->ccount), (int64_t)1);
    hipLaunchKernelGGL(gemm_screen_kernel, dim3((unsigned)(npad / 64), (unsigned)(A.Gpad / 64)),
                       dim3(256), 0, c->stream, P<__bf16>(c->ah), P<__bf16>(c->al), n, A.Kp,
                       P<__bf16>(c->bh), P<__bf16>(c->bl), A.G, P<double>(c->cg),
                       P<uint8_t>(c->gflag), P<uint8_t>(c->pflag), P<double>(c->thr),
                       A.margin_rel, A.E.clamp, P<Candidate>(c->cand),
                       P<unsigned long long>(c->ccount), cap, (float*)nullptr, (float*)nullptr,
                       (PoolEntry*)nullptr, (int64_t)0, (uint32_t*)nullptr);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(rescore_kernel, dim3(grid1((int64_t)cap, 256)), dim3(256), 0, c->stream,
                       stats_dev, n, P<double>(c->edata), P<double>(c->ese), A.G, A.E,
                       P<Candidate>(c->cand), P<unsigned long long>(c->ccount), cap);
    HIPCHK(c, hipGetLastError());
    unsigned long long cnt = 0;
    HIPCHK(c, hipMemcpyAsync(&cnt, c->ccount.p, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (cnt > cap) {  // grow and screen the batch again (the GEMM is cheap)
      cap = cnt + cnt / 8;
      A.cand_cap = cap;
      continue;
    }
    A.n_candidates += (int64_t)cnt;
    std::vector<Candidate> h(cnt);
    if (cnt)
      HIPCHK(c, hipMemcpy(h.data(), c->cand.p, sizeof(Candidate) * cnt, hipMemcpyDeviceToHost));
    std::vector<char> touched(A.mode == 1 ? A.G : 0, 0);
    for (const Candidate& x : h) {
      if (!(x.err == x.err)) continue;  // NaN errors are never accepted
      if (A.mode == 0) {
        if (x.err <= A.eps) A.kept[x.gene].push_back({x.err, pidx_host ? pidx_host[x.particle] : first_index + x.particle});
      } else {
        A.kept[x.gene].push_back({x.err, pidx_host ? pidx_host[x.particle] : first_index + x.particle});
        touched[x.gene] = 1;
      }
    }
    if (A.mode == 1) {
      bool changed = false;
      for (int64_t g = 0; g < A.G; ++g) {
        if (!touched[g]) continue;
        auto& v = A.kept[g];
        if ((int64_t)v.size() > A.k) {
          std::nth_element(v.begin(), v.begin() + A.k, v.end(),
                           [](const auto& a, const auto& b) { return GeneHit{a.first, a.second} < GeneHit{b.first, b.second}; });
          v.resize(A.k);
        }
        if ((int64_t)v.size() == A.k && A.k > 0) {
          double t = v[0].first;
          for (auto& e : v) t = std::max(t, e.first);
          if (t != A.thr_host[g]) { A.thr_host[g] = t; changed = true; }
        }
      }
      if (changed)
        HIPCHK(c, hipMemcpyAsync(c->thr.p, A.thr_host.data(), sizeof(double) * A.G,
                                 hipMemcpyHostToDevice, c->stream));
    }
    break;
  }
  A.n_screened += n;
  return ABC_OK;
}

}  // namespace

extern "C" {

int32_t abc_accept_begin(abc_ctx* c, int32_t K, int64_t G, const double* data, const double* se,
                         const int32_t* parts, int32_t n_parts, double clamp, int32_t mode,
                         double eps, int64_t k, double margin_rel) {
  if (!c || G < 0 || (G > 0 && (!data || !se))) return ABC_E_ARG;
  NO_CAPTURE(c);
  if (mode != 0 && mode != 1) return fail(c, ABC_E_ARG, "mode must be 0 (fixed eps) or 1 (top-k)");
  if (mode == 1 && k < 0) return fail(c, ABC_E_ARG, "k must be >= 0");
  if (!(margin_rel >= 0.0)) return fail(c, ABC_E_ARG, "margin_rel must be >= 0");
  AcceptState& A = c->acc;
  int rc = make_spec(c, K, parts, n_parts, clamp, &A.E);
  if (rc) return rc;
  if (K > 128) return fail(c, ABC_E_ARG, "screening supports K <= 128");
  A.G = G;
  A.Kp = (int)pad_to(K, 32);
  A.Gpad = pad_to(std::max<int64_t>(G, 1), 64);
  A.mode = mode;
  A.eps = eps;
  A.k = k;
  A.margin_rel = margin_rel;
  A.n_candidates = 0;
  A.n_screened = 0;
  A.kept.assign(mode == 1 ? 0 : G, {});
  A.thr_host.assign(std::max<int64_t>(G, 1), mode == 0 ? eps : HUGE_VAL);
  A.pooled = mode == 1;
  A.active = true;
  if (G == 0) return ABC_OK;
  if (A.pooled) {
    // pools of k + n_sub entries per gene: n_sub rows per screened block, as
    // many as about 1 GiB of pools allows (256..65536)
    if (k > (int64_t)1 << 30) return fail(c, ABC_E_ARG, "k must be <= 2^30 in top-k mode");
    const int64_t budget = (int64_t)1 << 30;
    int64_t ns = budget / (16 * G) - k;
    ns = std::max<int64_t>(256, std::min<int64_t>(65536, ns)) / 64 * 64;
    A.n_sub = ns;
    A.capg = k + ns;
    const int64_t npad = pad_to(ns, 64);
    if ((rc = ensure(c, c->pool, sizeof(PoolEntry) * (size_t)G * (size_t)A.capg))) return rc;
    if ((rc = ensure(c, c->pooln, 8 * (size_t)((G + 1) / 2)))) return rc;
    if ((rc = ensure(c, c->poolb, 8 * (size_t)((G + 1) / 2)))) return rc;
    if ((rc = ensure(c, c->presc, 8))) return rc;
    if ((rc = ensure(c, c->apidx, sizeof(int64_t) * (size_t)ns))) return rc;
    if ((rc = ensure(c, c->ah, sizeof(uint16_t) * npad * 2 * A.Kp))) return rc;
    if ((rc = ensure(c, c->al, sizeof(uint16_t) * npad * 2 * A.Kp))) return rc;
    if ((rc = ensure(c, c->pflag, npad))) return rc;
    hipLaunchKernelGGL(zero_u64_kernel, dim3(grid1((G + 1) / 2, 256)), dim3(256), 0, c->stream,
                       P<unsigned long long>(c->pooln), (G + 1) / 2);
    hipLaunchKernelGGL(zero_u64_kernel, dim3(grid1((G + 1) / 2, 256)), dim3(256), 0, c->stream,
                       P<unsigned long long>(c->poolb), (G + 1) / 2);
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, c->stream,
                       P<unsigned long long>(c->presc), (int64_t)1);
    HIPCHK(c, hipGetLastError());
  }
  if ((rc = ensure(c, c->edata, sizeof(double) * G * K))) return rc;
  if ((rc = ensure(c, c->ese, sizeof(double) * G * K))) return rc;
  if ((rc = ensure(c, c->bh, sizeof(uint16_t) * A.Gpad * 2 * A.Kp))) return rc;
  if ((rc = ensure(c, c->bl, sizeof(uint16_t) * A.Gpad * 2 * A.Kp))) return rc;
  if ((rc = ensure(c, c->cg, sizeof(double) * G))) return rc;
  if ((rc = ensure(c, c->gflag, G))) return rc;
  if ((rc = ensure(c, c->thr, sizeof(double) * G))) return rc;
  if ((rc = ensure(c, c->wf, sizeof(float) * G * K))) return rc;
  HIPCHK(c, hipMemcpyAsync(c->edata.p, data, sizeof(double) * G * K, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->ese.p, se, sizeof(double) * G * K, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->thr.p, A.thr_host.data(), sizeof(double) * G, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(gene_prep_kernel, dim3(grid1(A.Gpad, 256)), dim3(256), 0, c->stream,
                     P<double>(c->edata), P<double>(c->ese), G, K, A.Kp, A.Gpad,
                     P<__bf16>(c->bh), P<__bf16>(c->bl), P<double>(c->cg), P<uint8_t>(c->gflag),
                     P<float>(c->wf));
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ABC_OK;
}

namespace {

int screen_run(abc_ctx* c, abc_mlp* mlp, int64_t n, const float* x_dev, const float* mean_dev,
               const float* sd_dev, double z, double margin, uint8_t* pass_dev, float* score_dev,
               int64_t* n_pass);

}  // namespace

int32_t abc_screen_dev(abc_ctx* c, abc_mlp* mlp, int64_t n, const float* x_dev,
                       const float* mean_dev, const float* sd_dev, double z, double margin,
                       uint8_t* pass_dev, int64_t* n_pass) {
  if (!c || !mlp || n < 0 || (n > 0 && (!x_dev || !mean_dev || !sd_dev || !pass_dev)))
    return ABC_E_ARG;
  return screen_run(c, mlp, n, x_dev, mean_dev, sd_dev, z, margin, pass_dev, nullptr, n_pass);
}

int32_t abc_screen_score_dev(abc_ctx* c, abc_mlp* mlp, int64_t n, const float* x_dev,
                             const float* mean_dev, const float* sd_dev, double margin,
                             float* score_dev) {
  if (!c || !mlp || n < 0 || (n > 0 && (!x_dev || !mean_dev || !sd_dev || !score_dev)))
    return ABC_E_ARG;
  if (n == 0) return ABC_OK;
  int rc;
  if ((rc = ensure(c, c->spass, (size_t)n))) return rc;
  return screen_run(c, mlp, n, x_dev, mean_dev, sd_dev, 0.0, margin, P<uint8_t>(c->spass),
                    score_dev, nullptr);
}

}  // extern "C"

namespace {

int screen_run(abc_ctx* c, abc_mlp* mlp, int64_t n, const float* x_dev, const float* mean_dev,
               const float* sd_dev, double z, double margin, uint8_t* pass_dev, float* score_dev,
               int64_t* n_pass) {
  NO_CAPTURE(c);
  AcceptState& A = c->acc;
  if (!A.active) return fail(c, ABC_E_ARG, "abc_accept_begin was not called");
  if (abc_mlp_n_stats(mlp) != A.E.K) return fail(c, ABC_E_ARG, "surrogate and genes disagree on K");
  if (n == 0) { if (n_pass) *n_pass = 0; return ABC_OK; }
  int rc;
  if ((rc = ensure(c, c->sy, sizeof(float) * n * 2 * A.E.K))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  rc = abc_mlp_forward_dev(mlp, n, x_dev, P<float>(c->sy));
  if (rc) return fail(c, rc, "surrogate forward failed");
  if ((rc = abc_mlp_synchronize(mlp))) return fail(c, rc, "surrogate forward failed");
  hipLaunchKernelGGL(screen_kernel, dim3((unsigned)((n + 15) / 16)), dim3(256), 0, c->stream,
                     P<float>(c->sy), n, A.E.K, mean_dev, sd_dev, (float)z, P<float>(c->wf),
                     P<double>(c->edata), A.G, P<double>(c->thr), (float)margin, A.E.clamp,
                     pass_dev, score_dev);
  HIPCHK(c, hipGetLastError());
  if (n_pass) {
    std::vector<uint8_t> h(n);
    HIPCHK(c, hipMemcpyAsync(h.data(), pass_dev, n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int64_t s = 0;
    for (uint8_t b : h) s += b;
    *n_pass = s;
  }
  return ABC_OK;
}

}  // namespace

extern "C" {

int32_t abc_accept_add_dev(abc_ctx* c, int64_t n, const double* stats_dev, int64_t first_index) {
  if (!c || n < 0 || (n > 0 && !stats_dev) || first_index < 0) return ABC_E_ARG;
  if (c->capturing && !c->acc.pooled)
    return fail(c, ABC_E_STATE, "only top-k acceptance batches can be captured");
  if (!c->acc.active) return fail(c, ABC_E_ARG, "abc_accept_begin was not called");
  if (n == 0 || c->acc.G == 0) return ABC_OK;
  return screen_batch(c, n, stats_dev, first_index);
}

int32_t abc_accept_add_idx_dev(abc_ctx* c, int64_t n, const double* stats_dev,
                               const int64_t* pidx_host) {
  if (!c || n < 0 || (n > 0 && (!stats_dev || !pidx_host))) return ABC_E_ARG;
  NO_CAPTURE(c);
  if (!c->acc.active) return fail(c, ABC_E_ARG, "abc_accept_begin was not called");
  if (n == 0 || c->acc.G == 0) return ABC_OK;
  return screen_batch(c, n, stats_dev, 0, pidx_host);
}

int32_t abc_accept_set_thresholds(abc_ctx* c, const double* thr) {
  if (!c || !thr) return ABC_E_ARG;
  NO_CAPTURE(c);
  AcceptState& A = c->acc;
  if (!A.active) return fail(c, ABC_E_ARG, "abc_accept_begin was not called");
  for (int64_t g = 0; g < A.G; ++g) A.thr_host[g] = std::min(A.thr_host[g], thr[g]);
  if (A.G > 0 && A.pooled) {
    // the device thresholds may already be lower (k-th errors found so far)
    int rc;
    if ((rc = ensure(c, c->thrlow, sizeof(double) * A.G))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->thrlow.p, thr, sizeof(double) * A.G, hipMemcpyHostToDevice,
                             c->stream));
    hipLaunchKernelGGL(thr_min_kernel, dim3(grid1(A.G, 256)), dim3(256), 0, c->stream,
                       P<double>(c->thr), P<double>(c->thrlow), A.G);
    HIPCHK(c, hipGetLastError());
  } else if (A.G > 0) {
    HIPCHK(c, hipMemcpyAsync(c->thr.p, A.thr_host.data(), sizeof(double) * A.G,
                             hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ABC_OK;
}

int32_t abc_accept_add(abc_ctx* c, int64_t n, const double* stats, int64_t first_index) {
  if (!c || n < 0 || (n > 0 && !stats) || first_index < 0) return ABC_E_ARG;
  NO_CAPTURE(c);
  if (!c->acc.active) return fail(c, ABC_E_ARG, "abc_accept_begin was not called");
  if (n == 0 || c->acc.G == 0) return ABC_OK;
  int rc;
  if ((rc = ensure(c, c->stats, sizeof(double) * n * c->acc.E.K))) return rc;
  HIPCHK(c, hipMemcpyAsync(c->stats.p, stats, sizeof(double) * n * c->acc.E.K,
                           hipMemcpyHostToDevice, c->stream));
  return screen_batch(c, n, P<double>(c->stats), first_index);
}

int32_t abc_accept_end(abc_ctx* c, int64_t* counts, int64_t* idx, double* err_out, int64_t cap,
                       int64_t* n_candidates) {
  if (!c || cap < 0 || (cap > 0 && !idx)) return ABC_E_ARG;
  NO_CAPTURE(c);
  AcceptState& A = c->acc;
  if (!A.active) return fail(c, ABC_E_ARG, "abc_accept_begin was not called");
  if (A.G > 0 && !counts) return ABC_E_ARG;
  if (A.pooled && A.G > 0) {
    // each gene's kept entries: sorted by (error, index) on the device, then copied
    if (A.k <= ABC_POOL_SORT_MAX) {
      const int grid = (int)std::min<int64_t>(A.G, (int64_t)c->num_cu * 8);
      hipLaunchKernelGGL(pool_sort_kernel, dim3(grid), dim3(256), 0, c->stream,
                         P<PoolEntry>(c->pool), A.capg, P<uint32_t>(c->pooln), A.G);
      HIPCHK(c, hipGetLastError());
    }
    std::vector<uint32_t> cnt(A.G);
    unsigned long long nres = 0;
    HIPCHK(c, hipMemcpyAsync(cnt.data(), c->pooln.p, sizeof(uint32_t) * A.G,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&nres, c->presc.p, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    uint32_t mx = 0;
    for (uint32_t v : cnt) mx = std::max(mx, v);
    std::vector<PoolEntry> h((size_t)A.G * mx);
    if (mx)
      HIPCHK(c, hipMemcpy2D(h.data(), sizeof(PoolEntry) * mx, c->pool.p,
                            sizeof(PoolEntry) * (size_t)A.capg, sizeof(PoolEntry) * mx, A.G,
                            hipMemcpyDeviceToHost));
    A.kept.assign(A.G, {});
    for (int64_t g = 0; g < A.G; ++g) {
      auto& v = A.kept[g];
      v.reserve(cnt[g]);
      for (uint32_t q = 0; q < cnt[g]; ++q)
        v.push_back({h[(size_t)g * mx + q].err, h[(size_t)g * mx + q].idx});
    }
    A.n_candidates = (int64_t)nres;
  }
  int64_t total = 0;
  for (int64_t g = 0; g < A.G; ++g) {
    auto& v = A.kept[g];
    std::sort(v.begin(), v.end());
    counts[g] = (int64_t)v.size();
    for (auto& pr : v) {
      if (total < cap) {
        idx[total] = pr.second + 1;
        if (err_out) err_out[total] = pr.first;
      }
      total++;
    }
  }
  if (n_candidates) *n_candidates = A.n_candidates;
  if (total > cap) return fail(c, ABC_E_CAPACITY, "accepted entries exceed cap");
  A.active = false;
  A.kept.clear();
  return ABC_OK;
}

int32_t abc_accept(abc_ctx* c, int64_t n, int32_t K, const double* stats, int64_t G,
                   const double* data, const double* se, const int32_t* parts, int32_t n_parts,
                   double clamp, int32_t mode, double eps, int64_t k, int64_t first_index,
                   int64_t* counts, int64_t* idx, double* err_out, int64_t cap) {
  if (!c || n < 0 || G < 0 || (G > 0 && !counts) || cap < 0 || (cap > 0 && !idx)) return ABC_E_ARG;
  NO_CAPTURE(c);
  int rc = abc_accept_begin(c, K, G, data, se, parts, n_parts, clamp, mode, eps, k, 0.02);
  if (rc) return rc;
  if ((rc = abc_accept_add(c, n, stats, first_index))) { c->acc.active = false; return rc; }
  rc = abc_accept_end(c, counts, idx, err_out, cap, nullptr);
  c->acc.active = false;
  return rc;
}

int32_t abc_screen_errors(abc_ctx* c, int64_t n, int32_t K, const double* stats, int64_t G,
                          const double* data, const double* se, const int32_t* parts,
                          int32_t n_parts, double clamp, float* approx, float* bound) {
  if (!c || n <= 0 || G <= 0 || !stats || !approx || !bound) return ABC_E_ARG;
  NO_CAPTURE(c);
  int rc = abc_accept_begin(c, K, G, data, se, parts, n_parts, clamp, 0, 0.0, 0, 0.0);
  if (rc) return rc;
  AcceptState& A = c->acc;
  const int64_t npad = pad_to(n, 64);
  if ((rc = ensure(c, c->stats, sizeof(double) * n * K))) return rc;
  if ((rc = ensure(c, c->ah, sizeof(uint16_t) * npad * 2 * A.Kp))) return rc;
  if ((rc = ensure(c, c->al, sizeof(uint16_t) * npad * 2 * A.Kp))) return rc;
  if ((rc = ensure(c, c->pflag, npad))) return rc;
  if ((rc = ensure(c, c->err, sizeof(float) * 2 * n * G))) return rc;
  HIPCHK(c, hipMemcpyAsync(c->stats.p, stats, sizeof(double) * n * K, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(particle_prep_kernel, dim3(grid1(npad, 256)), dim3(256), 0, c->stream,
                     P<double>(c->stats), n, n, K, A.Kp, npad, P<__bf16>(c->ah),
                     P<__bf16>(c->al), P<uint8_t>(c->pflag));
  float* ap = P<float>(c->err);
  float* bp = ap + n * G;
  hipLaunchKernelGGL(gemm_screen_kernel, dim3((unsigned)(npad / 64), (unsigned)(A.Gpad / 64)),
                     dim3(256), 0, c->stream, P<__bf16>(c->ah), P<__bf16>(c->al), n, A.Kp,
                     P<__bf16>(c->bh), P<__bf16>(c->bl), G, P<double>(c->cg), P<uint8_t>(c->gflag),
                     P<uint8_t>(c->pflag), P<double>(c->thr), 0.0, clamp, (Candidate*)nullptr,
                     (unsigned long long*)nullptr, 0ull, ap, bp, (PoolEntry*)nullptr, (int64_t)0,
                     (uint32_t*)nullptr);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(approx, ap, sizeof(float) * n * G, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(bound, bp, sizeof(float) * n * G, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  A.active = false;
  return ABC_OK;
}

// -------------------------------------------------------------- moment model

namespace {

int moment_cfg(abc_ctx* c, int model, const abc_moment_cfg* u, MomCfgDev* m) {
  if (model < 1 || model > 5) return fail(c, ABC_E_MODEL, "model must be 1..5");
  if (!u) return fail(c, ABC_E_ARG, "moment configuration missing");
  if (!(u->cycle > 0.0) || u->n_steps < 1 || u->n_steps > ABC_MAX_BINS)
    return fail(c, ABC_E_ARG, "cycle must be > 0 and n_steps 1..8");
  if (u->n_ages < 1 || u->n_ages > ABC_MOM_MAX_AGES || !u->ages)
    return fail(c, ABC_E_ARG, "n_ages must be 1..8 with ages given");
  if (u->n_conds < 1 || u->n_conds > ABC_MOM_MAX_CONDS || !u->pulse || !u->chase)
    return fail(c, ABC_E_ARG, "n_conds must be 1..16 with pulse and chase given");
  if (!(u->rtol > 0.0) || !(u->atol > 0.0)) return fail(c, ABC_E_ARG, "rtol and atol must be > 0");
  if (u->transient_mode != 0 && u->transient_mode != 1)
    return fail(c, ABC_E_ARG, "transient_mode must be 0 or 1");
  if (u->method != 0 && u->method != 1) return fail(c, ABC_E_ARG, "method must be 0 or 1");
  if (!std::isfinite(u->t0) || u->t0 > 0.0) return fail(c, ABC_E_ARG, "t0 must be finite and <= 0");
  std::memset(m, 0, sizeof(*m));
  m->cycle = u->cycle;
  m->n_steps = u->n_steps;
  m->n_ages = u->n_ages;
  m->n_conds = u->n_conds;
  m->transient_mode = u->transient_mode;
  m->t0 = u->t0;
  for (int i = 0; i < 9; ++i) m->iv[i] = u->iv[i];
  m->rtol = u->rtol;
  m->atol = u->atol;
  m->method = u->method;
  for (int a = 0; a < u->n_ages; ++a) {
    if (!std::isfinite(u->ages[a]) || u->ages[a] < 0.0)
      return fail(c, ABC_E_ARG, "ages must be finite and >= 0");
    m->ages[a] = u->ages[a];
  }
  for (int j = 0; j < u->n_conds; ++j) {
    if (!(u->pulse[j] >= 0.0) || !(u->chase[j] >= 0.0))
      return fail(c, ABC_E_ARG, "pulse and chase must be >= 0");
    m->pulse[j] = u->pulse[j];
    m->chase[j] = u->chase[j];
    m->is_chase[j] = u->is_chase ? (u->is_chase[j] != 0) : 0;
  }
  m->pulse_cond = m->chase_cond = -1;
  return ABC_OK;
}

// transient -> trajectories (-> downsampling and statistics when stats_dev)
int moment_run(abc_ctx* c, int model, int64_t n, const double* theta_dev, int64_t ld,
               const MomCfgDev& m, double* out_dev, int64_t* steps_dev, double* stats_dev) {
  int rc;
  if ((rc = ensure(c, c->mss, sizeof(double) * 9 * (size_t)n))) return rc;
  if ((rc = ensure(c, c->mmap, sizeof(double) * 90 * (size_t)n))) return rc;
  if ((rc = ensure(c, c->mcost, sizeof(unsigned long long) * 2 * (size_t)n))) return rc;
  if ((rc = ensure(c, c->mord, sizeof(int32_t) * 2 * (size_t)n))) return rc;
  unsigned long long* steps = reinterpret_cast<unsigned long long*>(steps_dev);
  unsigned long long* key = P<unsigned long long>(c->mcost);
  unsigned long long* cost = key + n;
  int32_t* ord_map = P<int32_t>(c->mord);
  int32_t* ord = ord_map + n;
  // Lanes of a wave run in lockstep, so each kernel takes the particles in an
  // order of descending cost (the results do not depend on it): the cycle
  // maps by the cost predicted from the rates, the trajectories by the steps
  // the cycle maps took.
  hipLaunchKernelGGL(moment_key_kernel, dim3(grid1(n, 256)), dim3(256), 0, c->stream, m, model, n,
                     theta_dev, ld, key);
  hipLaunchKernelGGL(moment_order_kernel, dim3(1), dim3(ABC_MOM_ORDER_WG), 0, c->stream, n, key,
                     ord_map);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(grid1(n, 256)), dim3(256), 0, c->stream, cost, n);
  if (m.method == 1)
    hipLaunchKernelGGL(moment_cyclemap_kernel<1>, dim3((unsigned)((n + 63) / 64 * 10)), dim3(64),
                       0, c->stream, m, model, n, theta_dev, ld, ord_map, P<double>(c->mmap), cost,
                       steps);
  else
    hipLaunchKernelGGL(moment_cyclemap_kernel<0>, dim3((unsigned)((n + 63) / 64 * 10)), dim3(64),
                       0, c->stream, m, model, n, theta_dev, ld, ord_map, P<double>(c->mmap), cost,
                       steps);
  HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(moment_order_kernel, dim3(1), dim3(ABC_MOM_ORDER_WG), 0, c->stream, n, cost,
                     ord);
  HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(moment_transient_kernel, dim3(grid1(n, 64)), dim3(64), 0, c->stream, m, n,
                     P<double>(c->mmap), P<double>(c->mss));
  HIPCHK(c, hipGetLastError());
  // one wave per (group of 64 ordered particles, condition x age item)
  const int64_t waves = (n + 63) / 64 * (int64_t)m.n_conds * m.n_ages;
  if (m.method == 1)
    hipLaunchKernelGGL(moment_traj_kernel<1>, dim3((unsigned)waves), dim3(64), 0, c->stream, m,
                       model, n, theta_dev, ld, P<double>(c->mss), P<double>(c->mmap), ord,
                       out_dev, steps);
  else
    hipLaunchKernelGGL(moment_traj_kernel<0>, dim3((unsigned)waves), dim3(64), 0, c->stream, m,
                       model, n, theta_dev, ld, P<double>(c->mss), P<double>(c->mmap), ord,
                       out_dev, steps);
  HIPCHK(c, hipGetLastError());
  if (stats_dev) {
    hipLaunchKernelGGL(moment_stats_kernel, dim3(grid1(n, 64)), dim3(64), 0, c->stream, m, n,
                       out_dev, stats_dev);
    HIPCHK(c, hipGetLastError());
  }
  return ABC_OK;
}

int moment_reference_cols(int model, int nb) {
  return (model >= 3 ? nb : 1) + 4;  // kon/koff/alpha/gamma/lambda, one of the first four per bin
}

}  // namespace

int32_t abc_recover_dev(abc_ctx* c, int32_t model, int64_t n, const double* theta_dev,
                        int64_t ld, const abc_moment_cfg* cfg, double* out_dev,
                        int64_t* steps_dev) {
  if (!c) return ABC_E_ARG;
  MomCfgDev m;
  int rc = moment_cfg(c, model, cfg, &m);
  if (rc) return rc;
  if (n < 0 || ld < n) return fail(c, ABC_E_ARG, "need n >= 0 and ld >= n");
  if (n == 0) return ABC_OK;
  if (!theta_dev || !out_dev) return fail(c, ABC_E_ARG, "theta and out are required");
  if (n > 0x7FFFFFFFll / (m.n_conds * m.n_ages)) return fail(c, ABC_E_ARG, "too many particles per call");
  if (steps_dev)
    hipLaunchKernelGGL(zero_u64_kernel, dim3(grid1(n, 256)), dim3(256), 0, c->stream,
                       reinterpret_cast<unsigned long long*>(steps_dev), n);
  return moment_run(c, model, n, theta_dev, ld, m, out_dev, steps_dev, nullptr);
}

int32_t abc_recover(abc_ctx* c, int32_t model, int64_t n, const double* theta, int64_t ld,
                    const abc_moment_cfg* cfg, double* out, int64_t* steps) {
  if (!c) return ABC_E_ARG;
  NO_CAPTURE(c);
  MomCfgDev m;
  int rc = moment_cfg(c, model, cfg, &m);
  if (rc) return rc;
  if (n < 0 || ld < n) return fail(c, ABC_E_ARG, "need n >= 0 and ld >= n");
  if (n == 0) return ABC_OK;
  if (!theta || !out) return fail(c, ABC_E_ARG, "theta and out are required");
  const int cols = moment_reference_cols(model, m.n_steps);
  const size_t outw = (size_t)m.n_conds * m.n_ages * 5;
  if ((rc = ensure(c, c->mth, sizeof(double) * (size_t)n * cols))) return rc;
  if ((rc = ensure(c, c->mom, sizeof(double) * (size_t)n * outw))) return rc;
  if ((rc = ensure(c, c->mst, sizeof(int64_t) * (size_t)n))) return rc;
  HIPCHK(c, hipMemcpy2DAsync(c->mth.p, sizeof(double) * n, theta, sizeof(double) * ld,
                             sizeof(double) * n, cols, hipMemcpyHostToDevice, c->stream));
  rc = abc_recover_dev(c, model, n, P<double>(c->mth), n, cfg, P<double>(c->mom),
                       steps ? P<int64_t>(c->mst) : nullptr);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(out, c->mom.p, sizeof(double) * n * outw, hipMemcpyDeviceToHost, c->stream));
  if (steps)
    HIPCHK(c, hipMemcpyAsync(steps, c->mst.p, sizeof(int64_t) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ABC_OK;
}

int32_t abc_moment_sim_dev(abc_ctx* c, int32_t model, int64_t n, const double* theta_dev,
                           int64_t ld, const abc_moment_cfg* cfg, const double* capture,
                           const double* weights, int32_t pulse_cond, int32_t chase_cond,
                           double* stats_dev, double* moments_dev, int64_t* steps_dev) {
  if (!c) return ABC_E_ARG;
  MomCfgDev m;
  int rc = moment_cfg(c, model, cfg, &m);
  if (rc) return rc;
  if (n < 0 || ld < n) return fail(c, ABC_E_ARG, "need n >= 0 and ld >= n");
  if (!weights) return fail(c, ABC_E_ARG, "age weights are required");
  if (pulse_cond < -1 || pulse_cond >= m.n_conds || chase_cond < -1 || chase_cond >= m.n_conds)
    return fail(c, ABC_E_ARG, "pulse_cond / chase_cond must be -1 or a condition index");
  if (n == 0) return ABC_OK;
  if (!theta_dev || !stats_dev) return fail(c, ABC_E_ARG, "theta and stats are required");
  if (n > 0x7FFFFFFFll / (m.n_conds * m.n_ages)) return fail(c, ABC_E_ARG, "too many particles per call");
  for (int i = 0; i < m.n_ages * m.n_conds; ++i) m.weights[i] = weights[i];
  m.downsample = capture != nullptr;
  if (capture)
    for (int i = 0; i < m.n_ages * 2 * 3; ++i) m.cap[i] = capture[i];
  m.pulse_cond = pulse_cond;
  m.chase_cond = chase_cond;
  double* mom = moments_dev;
  if (!mom) {
    if ((rc = ensure(c, c->mom, sizeof(double) * (size_t)n * m.n_conds * m.n_ages * 5))) return rc;
    mom = P<double>(c->mom);
  }
  if (steps_dev)
    hipLaunchKernelGGL(zero_u64_kernel, dim3(grid1(n, 256)), dim3(256), 0, c->stream,
                       reinterpret_cast<unsigned long long*>(steps_dev), n);
  return moment_run(c, model, n, theta_dev, ld, m, mom, steps_dev, stats_dev);
}

int32_t abc_moment_sim(abc_ctx* c, int32_t model, int64_t n, const double* theta, int64_t ld,
                       const abc_moment_cfg* cfg, const double* capture, const double* weights,
                       int32_t pulse_cond, int32_t chase_cond, double* stats, double* moments,
                       int64_t* steps) {
  if (!c) return ABC_E_ARG;
  NO_CAPTURE(c);
  MomCfgDev m;
  int rc = moment_cfg(c, model, cfg, &m);
  if (rc) return rc;
  if (n < 0 || ld < n) return fail(c, ABC_E_ARG, "need n >= 0 and ld >= n");
  if (n == 0) return ABC_OK;
  if (!theta || !stats) return fail(c, ABC_E_ARG, "theta and stats are required");
  const int cols = moment_reference_cols(model, m.n_steps);
  const size_t outw = (size_t)m.n_conds * m.n_ages * 5;
  const int K = 4 * m.n_ages + 3 * m.n_conds;
  if ((rc = ensure(c, c->mth, sizeof(double) * (size_t)n * cols))) return rc;
  if ((rc = ensure(c, c->mom, sizeof(double) * (size_t)n * outw))) return rc;
  if ((rc = ensure(c, c->stats, sizeof(double) * (size_t)n * K))) return rc;
  if ((rc = ensure(c, c->mst, sizeof(int64_t) * (size_t)n))) return rc;
  HIPCHK(c, hipMemcpy2DAsync(c->mth.p, sizeof(double) * n, theta, sizeof(double) * ld,
                             sizeof(double) * n, cols, hipMemcpyHostToDevice, c->stream));
  rc = abc_moment_sim_dev(c, model, n, P<double>(c->mth), n, cfg, capture, weights, pulse_cond,
                          chase_cond, P<double>(c->stats), P<double>(c->mom),
                          steps ? P<int64_t>(c->mst) : nullptr);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(stats, c->stats.p, sizeof(double) * n * K, hipMemcpyDeviceToHost, c->stream));
  if (moments)
    HIPCHK(c, hipMemcpyAsync(moments, c->mom.p, sizeof(double) * n * outw, hipMemcpyDeviceToHost, c->stream));
  if (steps)
    HIPCHK(c, hipMemcpyAsync(steps, c->mst.p, sizeof(int64_t) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ABC_OK;
}

// ------------------------------------------------------ model probabilities

int32_t abc_model_probs(abc_ctx* c, int64_t G, int32_t M, const int64_t* counts, uint64_t seed,
                        int32_t n_boot, double alpha, double* prob, double* lb, double* ub) {
  if (!c) return ABC_E_ARG;
  NO_CAPTURE(c);
  if (G < 0 || M < 1 || M > ABC_SEL_MAX_GROUPS)
    return fail(c, ABC_E_ARG, "need G >= 0 and 1..8 model groups");
  if (n_boot < 1 || n_boot > ABC_SEL_MAX_BOOT) return fail(c, ABC_E_ARG, "n_boot must be 1..256");
  if (!(alpha >= 0.0 && alpha <= 1.0)) return fail(c, ABC_E_ARG, "alpha must lie in [0, 1]");
  if (G == 0) return ABC_OK;
  if (G > 0x7FFFFFFF) return fail(c, ABC_E_ARG, "at most 2^31 - 1 genes per call");
  if (!counts || !prob || !lb || !ub) return fail(c, ABC_E_ARG, "counts and outputs are required");
  for (int64_t i = 0; i < G * M; ++i)
    if (counts[i] < 0 || counts[i] > 0xFFFFFFFFll)
      return fail(c, ABC_E_ARG, "counts must lie in [0, 2^32)");
  for (int64_t g = 0; g < G; ++g) {
    int64_t S = 0;
    for (int m = 0; m < M; ++m) S += counts[g + G * m];
    if (S > 0xFFFFFFFFll) return fail(c, ABC_E_ARG, "at most 2^32 - 1 accepted particles per gene");
  }
  int rc;
  const size_t gm = (size_t)G * M;
  if ((rc = ensure(c, c->selc, sizeof(int64_t) * gm))) return rc;
  if ((rc = ensure(c, c->selo, sizeof(double) * 3 * gm))) return rc;
  HIPCHK(c, hipMemcpyAsync(c->selc.p, counts, sizeof(int64_t) * gm, hipMemcpyHostToDevice, c->stream));
  double* o = P<double>(c->selo);
  // the resampled shares of a launch's genes stay within 256 MiB
  const int64_t per_gene = (int64_t)M * n_boot * (int64_t)sizeof(double);
  int64_t chunk = (256ll << 20) / per_gene;
  if (chunk > G) chunk = G;
  if ((rc = ensure(c, c->selb, (size_t)(chunk * per_gene)))) return rc;
  const int rpb = ABC_SEL_BOOT_PER_WG;
  const unsigned parts = (unsigned)((n_boot + rpb - 1) / rpb);
  for (int64_t g0 = 0; g0 < G; g0 += chunk) {
    const int64_t cnt = G - g0 < chunk ? G - g0 : chunk;
    hipLaunchKernelGGL(model_prob_kernel, dim3((unsigned)cnt, parts), dim3(ABC_SEL_WG), 0,
                       c->stream, G, g0, M, P<int64_t>(c->selc), (uint32_t)seed,
                       (uint32_t)(seed >> 32) ^ 0x5851F42Du, n_boot, rpb, P<double>(c->selb), o,
                       o + gm, o + 2 * gm);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(model_prob_bounds_kernel, dim3((unsigned)cnt), dim3(ABC_SEL_WG), 0,
                       c->stream, G, g0, M, P<int64_t>(c->selc), n_boot, alpha,
                       P<double>(c->selb), o, o + gm, o + 2 * gm);
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipMemcpyAsync(prob, o, sizeof(double) * gm, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(lb, o + gm, sizeof(double) * gm, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(ub, o + 2 * gm, sizeof(double) * gm, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ABC_OK;
}

// --------------------------------------------- data-side summary statistics

namespace {

constexpr int64_t kDataMaxCells = 1ll << 21;  // with counts < 2^20 every sum fits 64 bits

// validate the cells' metadata and copy it to the device
int data_meta(abc_ctx* c, int64_t G, int64_t n, int32_t nc, const int32_t* age, int32_t J,
              const int32_t* cond, const int32_t* pulse, int64_t np, const int32_t* chase,
              int64_t nch, const double* w, int32_t n_boot) {
  if (G < 0 || G > 0x7FFFFFFFll) return fail(c, ABC_E_ARG, "need 0 <= genes < 2^31");
  if (n < 0 || n > kDataMaxCells) return fail(c, ABC_E_ARG, "need 0 <= cells <= 2^21");
  if (nc < 1 || nc > ABC_DATA_MAX_CLUSTERS) return fail(c, ABC_E_ARG, "n_clusters must be 1..8");
  if (J < 1 || J > ABC_DATA_MAX_CONDS) return fail(c, ABC_E_ARG, "n_conds must be 1..16");
  if (n_boot < 2 || n_boot > (1 << 20)) return fail(c, ABC_E_ARG, "n_boot must be 2..2^20");
  if (np < 0 || nch < 0 || np > kDataMaxCells || nch > kDataMaxCells)
    return fail(c, ABC_E_ARG, "pulse and chase lists: 0..2^21 cells");
  if (n > 0 && (!age || !cond)) return fail(c, ABC_E_ARG, "age and cond are required");
  if ((np > 0 && !pulse) || (nch > 0 && !chase) || !w)
    return fail(c, ABC_E_ARG, "pulse, chase and age_id_dist are required");
  for (int64_t i = 0; i < n; ++i) {
    if (age[i] < 0 || age[i] >= nc) return fail(c, ABC_E_ARG, "age clusters must lie in 0..n_clusters-1");
    if (cond[i] < -1 || cond[i] >= J) return fail(c, ABC_E_ARG, "conditions must lie in -1..n_conds-1");
  }
  for (int64_t i = 0; i < np; ++i)
    if (pulse[i] < 0 || pulse[i] >= n) return fail(c, ABC_E_ARG, "pulse cells must lie in 0..n-1");
  for (int64_t i = 0; i < nch; ++i)
    if (chase[i] < 0 || chase[i] >= n) return fail(c, ABC_E_ARG, "chase cells must lie in 0..n-1");
  for (int64_t i = 0; i < (int64_t)nc * J; ++i)
    if (!std::isfinite(w[i])) return fail(c, ABC_E_ARG, "age_id_dist must be finite");
  int rc;
  const size_t ni = (size_t)(2 * n + np + nch);
  if ((rc = ensure(c, c->dmeta, sizeof(int32_t) * (ni > 0 ? ni : 1)))) return rc;
  if ((rc = ensure(c, c->dw, sizeof(double) * (size_t)nc * J))) return rc;
  int32_t* m = P<int32_t>(c->dmeta);
  if (n > 0) {
    HIPCHK(c, hipMemcpyAsync(m, age, sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(m + n, cond, sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
  }
  if (np > 0)
    HIPCHK(c, hipMemcpyAsync(m + 2 * n, pulse, sizeof(int32_t) * np, hipMemcpyHostToDevice, c->stream));
  if (nch > 0)
    HIPCHK(c, hipMemcpyAsync(m + 2 * n + np, chase, sizeof(int32_t) * nch, hipMemcpyHostToDevice,
                             c->stream));
  HIPCHK(c, hipMemcpyAsync(c->dw.p, w, sizeof(double) * nc * J, hipMemcpyHostToDevice, c->stream));
  // each family's positions in group order (a stable counting sort) and the
  // group of each sorted position, for data_stats_sorted_kernel: pulse and
  // chase cells by cluster, all cells by condition, all cells by
  // (condition, cluster); -1 for a cell in no condition
  const int64_t total = np + nch + 2 * n;
  std::vector<int32_t> perm((size_t)(total > 0 ? total : 1)), grp((size_t)(total > 0 ? total : 1));
  auto order = [&](int64_t base, int64_t m, int ngrp, auto key) {
    // counts per key (key -1 counted in slot ngrp, placed last), then offsets
    std::vector<int64_t> at((size_t)ngrp + 1, 0);
    for (int64_t i = 0; i < m; ++i) {
      const int k = key(i);
      at[(size_t)(k < 0 ? ngrp : k)]++;
    }
    int64_t acc = 0;
    for (int k = 0; k <= ngrp; ++k) {
      const int64_t t = at[(size_t)k];
      at[(size_t)k] = acc;
      acc += t;
    }
    for (int64_t i = 0; i < m; ++i) {
      const int k = key(i);
      const int64_t to = at[(size_t)(k < 0 ? ngrp : k)]++;
      perm[(size_t)(base + to)] = (int32_t)i;
      grp[(size_t)(base + to)] = k;
    }
  };
  order(0, np, nc, [&](int64_t i) { return (int)age[pulse[i]]; });
  order(np, nch, nc, [&](int64_t i) { return (int)age[chase[i]]; });
  order(np + nch, n, J, [&](int64_t i) { return (int)cond[i]; });
  order(np + nch + n, n, J * nc, [&](int64_t i) { return cond[i] < 0 ? -1 : (int)(cond[i] * nc + age[i]); });
  if ((rc = ensure(c, c->dsort, sizeof(int32_t) * 2 * (size_t)(total > 0 ? total : 1)))) return rc;
  if (total > 0) {
    HIPCHK(c, hipMemcpyAsync(c->dsort.p, perm.data(), sizeof(int32_t) * total, hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(P<int32_t>(c->dsort) + total, grp.data(), sizeof(int32_t) * total,
                             hipMemcpyHostToDevice, c->stream));
    // the host vectors must outlive the copies
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return ABC_OK;
}

// the kernel over every gene, in launches whose resample scratch stays near 1 GiB
int data_run(abc_ctx* c, int64_t G, int64_t n, const uint32_t* u, const uint32_t* l, int32_t nc,
             int32_t J, int64_t np, int64_t nch, int32_t n_boot, uint64_t seed, double* out) {
  int rc;
  const int64_t K1 = 4 * nc + 3 * J;
  // families of at most ABC_DATA_SORTED_MAX cells: each gene's cells packed in
  // group order, per-cell draw counts in LDS, then sums over the cells in
  // group order; larger ones: sums per draw
  int64_t mmax = n;
  if (np > mmax) mmax = np;
  if (nch > mmax) mmax = nch;
  const bool sorted = mmax <= ABC_DATA_SORTED_MAX;
  const int64_t total = np + nch + 2 * n;
  const int64_t slots = (int64_t)ABC_DATA_WG *
                        ((np + ABC_DATA_WG - 1) / ABC_DATA_WG + (nch + ABC_DATA_WG - 1) / ABC_DATA_WG +
                         2 * ((n + ABC_DATA_WG - 1) / ABC_DATA_WG));
  const int64_t per_gene = (int64_t)n_boot * K1 * (int64_t)sizeof(double) +
                           (sorted ? slots * (int64_t)sizeof(unsigned long long) : 0);
  int64_t chunk = (1ll << 30) / per_gene;
  if (chunk < 1) chunk = 1;
  if (chunk > G) chunk = G;
  if ((rc = ensure(c, c->dboot, (size_t)(chunk * n_boot * K1) * sizeof(double)))) return rc;
  if ((rc = ensure(c, c->dpt, sizeof(double) * (size_t)(chunk * K1)))) return rc;
  if ((rc = ensure(c, c->dflag, sizeof(unsigned long long)))) return rc;
  if (sorted && slots > 0 && (rc = ensure(c, c->dxs, sizeof(unsigned long long) * (size_t)(chunk * slots))))
    return rc;
  unsigned long long* flag = P<unsigned long long>(c->dflag);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, c->stream, flag, (int64_t)1);
  const int32_t* m = P<int32_t>(c->dmeta);
  const int32_t* perm = P<int32_t>(c->dsort);
  for (int64_t g0 = 0; g0 < G; g0 += chunk) {
    const int64_t cnt = G - g0 < chunk ? G - g0 : chunk;
    if (sorted) {
      if (slots > 0)
        hipLaunchKernelGGL(data_pack_kernel,
                           dim3((unsigned)((cnt * (slots / ABC_DATA_WG) + 7) / 8 * 8)),
                           dim3(ABC_DATA_WG), 0, c->stream, n, u, l, g0, m + 2 * n, np,
                           m + 2 * n + np, nch, perm, perm + total, slots, cnt,
                           P<unsigned long long>(c->dxs), reinterpret_cast<unsigned int*>(flag));
      HIPCHK(c, hipGetLastError());
      const int spb = ABC_DATA_SETS_PER_WG;
      const int nsb = (n_boot + 1 + spb - 1) / spb;
      const int64_t items = cnt * nsb;
      hipLaunchKernelGGL(data_stats_sorted_kernel, dim3((unsigned)((items + 7) / 8 * 8)),
                         dim3(ABC_DATA_WG),
                         (size_t)((mmax + 1) / 2 > 0 ? (mmax + 1) / 2 : 1) * sizeof(uint32_t),
                         c->stream, cnt, nsb, g0, P<unsigned long long>(c->dxs), slots, n, np, nch,
                         nc, J, P<double>(c->dw), n_boot, spb, (uint32_t)seed,
                         (uint32_t)(seed >> 32) ^ 0x2545F491u, P<double>(c->dpt),
                         P<double>(c->dboot));
      HIPCHK(c, hipGetLastError());
      hipLaunchKernelGGL(data_se_kernel, dim3(grid1(cnt * K1, 256)), dim3(256), 0, c->stream, G,
                         g0, cnt, nc, J, n_boot, P<double>(c->dpt), P<double>(c->dboot), out);
    } else
      hipLaunchKernelGGL(data_stats_kernel, dim3((unsigned)cnt), dim3(ABC_DATA_WG), 0, c->stream,
                         G, g0, n, u, l, m, m + n, nc, J, m + 2 * n, np, m + 2 * n + np, nch,
                         P<double>(c->dw), n_boot, (uint32_t)seed,
                         (uint32_t)(seed >> 32) ^ 0x2545F491u, P<double>(c->dboot), out,
                         reinterpret_cast<unsigned int*>(flag));
    HIPCHK(c, hipGetLastError());
  }
  unsigned long long bad = 0;
  HIPCHK(c, hipMemcpyAsync(&bad, flag, sizeof(bad), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (bad) return fail(c, ABC_E_ARG, "counts must be below 2^20");
  return ABC_OK;
}

}  // namespace

int32_t abc_data_stats_dev(abc_ctx* c, int64_t n_genes, int64_t n_cells, const uint32_t* u_dev,
                           const uint32_t* l_dev, int32_t n_clusters, const int32_t* age,
                           int32_t n_conds, const int32_t* cond, const int32_t* pulse,
                           int64_t n_pulse, const int32_t* chase, int64_t n_chase,
                           const double* age_id_dist, int32_t n_boot, uint64_t seed,
                           double* out_dev) {
  if (!c) return ABC_E_ARG;
  NO_CAPTURE(c);
  int rc = data_meta(c, n_genes, n_cells, n_clusters, age, n_conds, cond, pulse, n_pulse, chase,
                     n_chase, age_id_dist, n_boot);
  if (rc) return rc;
  if (n_genes == 0) return ABC_OK;
  if ((n_cells > 0 && (!u_dev || !l_dev)) || !out_dev)
    return fail(c, ABC_E_ARG, "u, l and out are required");
  return data_run(c, n_genes, n_cells, u_dev, l_dev, n_clusters, n_conds, n_pulse, n_chase, n_boot,
                  seed, out_dev);
}

int32_t abc_data_stats(abc_ctx* c, int64_t n_genes, int64_t n_cells, const uint32_t* u,
                       const uint32_t* l, int32_t n_clusters, const int32_t* age, int32_t n_conds,
                       const int32_t* cond, const int32_t* pulse, int64_t n_pulse,
                       const int32_t* chase, int64_t n_chase, const double* age_id_dist,
                       int32_t n_boot, uint64_t seed, double* out) {
  if (!c) return ABC_E_ARG;
  NO_CAPTURE(c);
  int rc = data_meta(c, n_genes, n_cells, n_clusters, age, n_conds, cond, pulse, n_pulse, chase,
                     n_chase, age_id_dist, n_boot);
  if (rc) return rc;
  if (n_genes == 0) return ABC_OK;
  if ((n_cells > 0 && (!u || !l)) || !out) return fail(c, ABC_E_ARG, "u, l and out are required");
  const size_t nu = (size_t)n_genes * (size_t)n_cells;
  for (size_t i = 0; i < nu; ++i)
    if ((u[i] | l[i]) >= (1u << 20)) return fail(c, ABC_E_ARG, "counts must be below 2^20");
  const size_t K = (size_t)(8 * n_clusters + 6 * n_conds);
  if ((rc = ensure(c, c->du, sizeof(uint32_t) * (nu > 0 ? nu : 1)))) return rc;
  if ((rc = ensure(c, c->dl, sizeof(uint32_t) * (nu > 0 ? nu : 1)))) return rc;
  if ((rc = ensure(c, c->dout, sizeof(double) * (size_t)n_genes * K))) return rc;
  if (nu > 0) {
    HIPCHK(c, hipMemcpyAsync(c->du.p, u, sizeof(uint32_t) * nu, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->dl.p, l, sizeof(uint32_t) * nu, hipMemcpyHostToDevice, c->stream));
  }
  rc = data_run(c, n_genes, n_cells, P<uint32_t>(c->du), P<uint32_t>(c->dl), n_clusters, n_conds,
                n_pulse, n_chase, n_boot, seed, P<double>(c->dout));
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(out, c->dout.p, sizeof(double) * (size_t)n_genes * K,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ABC_OK;
}

int32_t abc_ssa_time(abc_ctx* c, int32_t reset, double* ms2, int64_t* burn_steps,
                     int64_t* launches) {
  if (!c || !ms2 || !burn_steps || !launches) return ABC_E_ARG;
  NO_CAPTURE(c);
  int rc = drain_ring(c);
  if (rc) return rc;
  ms2[0] = c->burn_ms_acc;
  ms2[1] = c->ssa_ms_acc;
  *launches = c->ssa_launches;
  uint64_t tally = 0;
  if (c->tally.p) {
    HIPCHK(c, hipMemcpyAsync(&tally, c->tally.p, sizeof(tally), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  *burn_steps = (int64_t)tally;
  if (reset) {
    c->burn_ms_acc = c->ssa_ms_acc = 0.0;
    c->ssa_launches = 0;
    if (c->tally.p) {
      hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, c->stream,
                         P<unsigned long long>(c->tally), (int64_t)2);
      HIPCHK(c, hipGetLastError());
    }
  }
  return ABC_OK;
}

int32_t abc_last_timing(abc_ctx* c, double* ms) {
  if (!c || !ms) return ABC_E_ARG;
  NO_CAPTURE(c);
  for (int q = 0; q < 4; ++q) {
    ms[q] = 0.0;
    if (!c->timed[q]) continue;
    float v = 0.f;
    HIPCHK(c, hipEventSynchronize(c->ev[q + 1]));
    HIPCHK(c, hipEventElapsedTime(&v, c->ev[q], c->ev[q + 1]));
    ms[q] = v;
  }
  return ABC_OK;
}

// ------------------------------------------------------------ HIP graphs

int32_t abc_graph_begin(abc_ctx* c) {
  if (!c) return ABC_E_ARG;
  if (c->capturing) return fail(c, ABC_E_STATE, "a capture is already open");
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  c->capture_failed = false;
  return ABC_OK;
}

int32_t abc_graph_end(abc_ctx* c, abc_graph** g) {
  if (!c || !g) return ABC_E_ARG;
  if (!c->capturing) return fail(c, ABC_E_STATE, "no capture is open");
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &graph);
  c->capturing = false;
  if (c->capture_failed || e != hipSuccess || !graph) {
    if (graph) hipGraphDestroy(graph);
    (void)hipGetLastError();
    const bool own = c->capture_failed;
    c->capture_failed = false;
    return own ? ABC_E_STATE
               : fail(c, ABC_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  if (*g && (*g)->exec) {
    (*g)->ctx = c;
    (*g)->gen = c->gen;
    // same topology with new arguments (other particles, seeds or buffers):
    // update the executable graph in place
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult res = hipGraphExecUpdateError;
    if (hipGraphExecUpdate((*g)->exec, graph, &bad, &res) == hipSuccess &&
        res == hipGraphExecUpdateSuccess) {
      hipGraphDestroy(graph);
      return ABC_OK;
    }
    (void)hipGetLastError();
    hipGraphExecDestroy((*g)->exec);
    (*g)->exec = nullptr;
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (ei != hipSuccess) return fail(c, ABC_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  if (!*g) *g = new abc_graph;
  (*g)->exec = exec;
  (*g)->ctx = c;
  (*g)->gen = c->gen;
  return ABC_OK;
}

int32_t abc_graph_launch(abc_ctx* c, abc_graph* g) {
  if (!c || !g || !g->exec) return ABC_E_ARG;
  NO_CAPTURE(c);
  if (g->ctx != c)
    return fail(c, ABC_E_STATE, "the graph was captured on another context");
  if (g->gen != c->gen)
    return fail(c, ABC_E_STATE,
                "a device buffer the graph uses was reallocated by a larger call after the "
                "capture: capture the graph again");
  HIPCHK(c, hipGraphLaunch(g->exec, c->stream));
  return ABC_OK;
}

int32_t abc_graph_destroy(abc_graph* g) {
  if (!g) return ABC_OK;
  if (g->exec) hipGraphExecDestroy(g->exec);
  delete g;
  return ABC_OK;
}

// -------------------------------------------------------------- autotuner

int32_t abc_autotune_ssa(abc_ctx* c, int32_t n_models, const int32_t* models, uint64_t seed,
                         int64_t first, int64_t n, int32_t* best_blocks, double* best_ms) {
  if (!c || n_models <= 0 || n_models > 16 || !models || n <= 0) return ABC_E_ARG;
  NO_CAPTURE(c);
  int rc;
  for (int q = 0; q < n_models; ++q)
    if ((rc = check_range(c, models[q], first, n))) return rc;
  const int nb = c->hd.nb, K = 4 * nb + 3 * c->hd.J;
  int np_max = 0;
  for (int q = 0; q < n_models; ++q) np_max = std::max(np_max, abc_layout_n_params(models[q], nb));
  if ((rc = ensure(c, c->theta, sizeof(double) * (size_t)n * np_max * n_models))) return rc;
  if ((rc = ensure(c, c->stats, sizeof(double) * (size_t)n * K * n_models))) return rc;
  double* theta = P<double>(c->theta);
  for (int q = 0; q < n_models; ++q)
    if ((rc = abc_sample_prior_dev(c, models[q], seed, first, n, theta + (size_t)q * n * np_max)))
      return rc;
  // candidates: 1..16 resident workgroups of SSA_WG lanes per CU, up to occupancy
  const int per_cu = std::max(1, c->ssa_blocks_default / std::max(1, c->num_cu));
  std::vector<int> cand;
  for (int w : {1, 2, 3, 4, 6, 8, 12, 16})
    if (w <= per_cu) cand.push_back(w * c->num_cu);
  if (std::find(cand.begin(), cand.end(), c->ssa_blocks_default) == cand.end())
    cand.push_back(c->ssa_blocks_default);
  // keep the caller's running SSA time out of the trials
  if ((rc = drain_ring(c))) return rc;
  const double acc0 = c->ssa_ms_acc, bacc0 = c->burn_ms_acc;
  const int64_t l0 = c->ssa_launches;
  uint64_t tally0 = 0;
  if (c->tally.p) {
    HIPCHK(c, hipMemcpyAsync(&tally0, c->tally.p, sizeof(tally0), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  const int saved = c->ssa_blocks;
  // every trial is one multi-model call (all models' particles in one burn-in
  // and one SSA launch), the shape of a batch of abc_simulate_multi_dev
  std::vector<int64_t> firsts(n_models, first), ns(n_models, n);
  std::vector<const double*> th_m(n_models);
  std::vector<double*> st_m(n_models);
  for (int q = 0; q < n_models; ++q) {
    th_m[q] = theta + (size_t)q * n * np_max;
    st_m[q] = P<double>(c->stats) + (size_t)q * n * K;
  }
  auto run = [&](int blocks, double* ms) -> int {
    c->ssa_blocks = blocks;
    const double a = c->ssa_ms_acc;
    {
      const int r = simulate_multi(c, n_models, models, seed, firsts.data(), nullptr, ns.data(),
                                   th_m.data(), st_m.data(), nullptr, nullptr, nullptr);
      if (r) return r;
    }
    const int r = drain_ring(c);
    *ms = c->ssa_ms_acc - a;
    return r;
  };
  double ms = 0.0, best = 1e300;
  int bestb = c->ssa_blocks_default;
  const int saved_claim = c->ssa_claim;
  c->ssa_claim = 1;
  rc = run(c->ssa_blocks_default, &ms);  // warm-up (code load, buffer growth)
  for (size_t q = 0; rc == ABC_OK && q < cand.size(); ++q) {
    rc = run(cand[q], &ms);
    if (rc == ABC_OK && ms < best) {
      best = ms;
      bestb = cand[q];
    }
  }
  // then the work items a lane claims per queue atomic, on that grid: the
  // better of two trials each (claims 2 and 4 differ by about 1 %, about the
  // spread of one trial)
  int bestc = 1;
  for (int cl : {2, 4, 8}) {
    c->ssa_claim = cl;
    for (int trial = 0; rc == ABC_OK && trial < 2; ++trial) {
      rc = run(bestb, &ms);
      if (rc == ABC_OK && ms < best) {
        best = ms;
        bestc = cl;
      }
    }
  }
  c->ssa_claim = rc == ABC_OK ? bestc : saved_claim;
  c->ssa_ms_acc = acc0;
  c->burn_ms_acc = bacc0;
  c->ssa_launches = l0;
  if (rc == ABC_OK && c->tally.p) {
    HIPCHK(c, hipMemcpyAsync(c->tally.p, &tally0, sizeof(tally0), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  if (rc) {
    c->ssa_blocks = saved;
    return rc;
  }
  c->ssa_blocks = bestb;
  if (best_blocks) *best_blocks = bestb;
  if (best_ms) *best_ms = best;
  return ABC_OK;
}

}  // extern "C"
