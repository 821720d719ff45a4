// Batched contractions (tensor-network sweeps): many independent
// contractions with one label triple and per-instance extents, replayed from
// one captured HIP graph.  No counterpart in the reference, whose contract
// handles one pair of tensors per call (include/t1m/t1m.h:560-577).
//
// Creation builds every instance's index maps (one plan per distinct shape),
// writes the instances' kernel arguments to device arrays, and lists every
// (instance, tile) pair, sorted by decreasing contracted extent (the work of
// a tile is proportional to K: longest-processing-time-first order).  The
// graph holds a persistent launch of a few workgroups per CU that walk that
// list with a stride of the grid size -- one launch per tile family when
// some instances take larger tiles than the rest.  Which instances take
// which family is a tiling mode; with `autotune` the candidate modes are
// timed on a scratch copy of C and the fastest is kept (and remembered in
// the tuning file).  When the instances cannot share one kernel (mixed
// gather directions or offset widths, or the simple kernel), the graph
// holds one launch per instance instead.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "internal.h"
#include "runtime.h"

using namespace t1m_hip;

// One persistent launch of a batch: a tile family, the kernel arguments of
// the instances it covers and its LPT-ordered work list.
struct BatchPhase {
  int kernel = 0;
  void* d_args = nullptr;
  BatchItem* d_items = nullptr;
  unsigned* d_tickets = nullptr;   // work-queue counters: one per domain + finished workgroups
  int nitems = 0;    // work-list entries, padding included
  int ntiles = 0;    // tiles (entries that are not padding)
  int ninst = 0;
  int grid = 0;
  int xcds = 1;      // L2 domains the work list is ordered for
  bool pairs = false;  // every instance allows two-element gathers (complex<float>)
};

struct t1m_batch {
  std::vector<std::shared_ptr<Plan>> plans;
  std::vector<GemmProblem> problems;
  std::vector<BatchPhase> phases;   // launched one after another
  int mode = 0;                     // tiling mode of the phases
  bool persistent = false;
  hipStream_t capture = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int synchronize = 1;
  std::string description;
  void release_device() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (capture) (void)hipStreamDestroy(capture);
    for (BatchPhase& ph : phases) {
      if (ph.d_args) (void)hipFree(ph.d_args);
      if (ph.d_items) (void)hipFree(ph.d_items);
      if (ph.d_tickets) (void)hipFree(ph.d_tickets);
    }
    exec = nullptr;
    graph = nullptr;
    capture = nullptr;
    phases.clear();
    persistent = false;
  }
  ~t1m_batch() { release_device(); }
};

namespace {

int cu_count_batch() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 256;
  return cus > 0 ? cus : 256;
}

// Tile family of a batch's instances: the small tiles, which keep the many
// small instances of a sweep busy (one instance rarely fills the 144-wide
// tiles).
int batch_kernel(int dtype, int precision) {
  if (const char* f = env_override("T1M_FORCE_KERNEL")) {
    const int k = atoi(f);
    // batches run the small square tiles only
    if (kernel_fits(k, dtype, precision) && k != K_C64_F32_144 && k != K_C64_BF16X3_144 &&
        k != K_C64_F32_THIN && k != K_C128_F64_THIN)
      return k;
  }
  if (precision == T1M_PRECISION_SIMPLE) return K_SIMPLE;
  switch (dtype) {
    case T1M_COMPLEX64:
      return precision == T1M_PRECISION_BF16X3 ? K_C64_BF16X3_64 : K_C64_F32_64;
    case T1M_COMPLEX128:
      return K_C128_F64_32;
    case T1M_FLOAT32:
      return K_F32_64;
    case T1M_FLOAT64:
      return K_F64_32;
    default:
      return K_SIMPLE;
  }
}

// Tiling modes (which instances run on which tile family):
//   0  every instance on the batch family (batch_kernel);
//   1  complex<float>: instances whose 144 x 144 tiling wastes < 10 % of its
//      area and whose K >= 128 on the 144 x 144 tiles, the rest on 64 x 64;
//   2  complex<float>: every instance at least 144 x 144 on the 144 x 144
//      tiles, the rest on 64 x 64;
//   3  complex<double>: every instance on the 64 x 64 tiles.
// Measured on the 512-instance MPS sweep (bond dimensions 16..256, extents
// multiples of 64): mode 0 84.7 TFLOP/s, mode 1 80.4, mode 2 67 -- ragged
// 144-wide tiles and one 144 x 144 workgroup per CU lose there; batches of
// large, evenly tiled instances gain from them.
std::vector<int> tiling_modes(int dtype, int precision) {
  if (env_override("T1M_FORCE_KERNEL")) return {0};
  if (dtype == T1M_COMPLEX64 && precision == T1M_PRECISION_NATIVE) return {0, 1, 2};
  if (dtype == T1M_COMPLEX128 && precision == T1M_PRECISION_NATIVE) return {0, 3};
  return {0};
}

int kernel_for(const GemmProblem& g, int family, int mode) {
  auto util = [](int64_t x) { return double(x) / double((x + 143) / 144 * 144); };
  switch (mode) {
    case 1:
      return util(g.M) * util(g.N) >= 0.9 && g.K >= 128 ? K_C64_F32_144 : family;
    case 2:
      return g.M >= 144 && g.N >= 144 ? K_C64_F32_144 : family;
    case 3:
      return K_C128_F64_64;
    default:
      return family;
  }
}

// L2 domains (XCDs) of the device: gfx950 and gfx942 parts dispatch the
// workgroups of a launch round-robin over their XCDs (workgroup b on XCD
// b mod n), each with its own L2.  T1M_BATCH_XCDS overrides (1: no XCD
// ordering).
int xcd_count() {
  if (const char* e = env_override("T1M_BATCH_XCDS")) {
    const int n = atoi(e);
    return n >= 1 && n <= 64 ? n : 1;
  }
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    (void)hipGetLastError();
    return 1;
  }
  const int cus = cu_count_batch();
  if (strncmp(prop.gcnArchName, "gfx950", 6) == 0 && cus % 8 == 0) return 8;
  if (strncmp(prop.gcnArchName, "gfx942", 6) == 0) return cus % 8 == 0 ? 8 : cus % 6 == 0 ? 6 : 1;
  return 1;
}

// Work list ordered for n L2 domains.  Each instance goes whole to one
// domain (instances by decreasing work, each to the least loaded domain),
// so its A and B panels are fetched into one L2 instead of every domain's;
// each domain's items run longest first (LPT) and the domains' lists are
// interleaved, entry p going to domain p mod n, which is where workgroup
// p mod grid runs when the grid is a multiple of n.  Shorter lists are
// padded with entries of instance -1, which the kernels skip.
std::vector<BatchItem> xcd_order(const std::vector<BatchItem>& items, size_t ninst,
                                 const std::vector<int64_t>& inst_work, int n) {
  std::vector<size_t> order(ninst);
  for (size_t i = 0; i < ninst; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t x, size_t y) { return inst_work[x] > inst_work[y]; });
  std::vector<int> dom(ninst, 0);
  std::vector<int64_t> load(static_cast<size_t>(n), 0);
  std::vector<size_t> rank(ninst, 0);
  for (size_t r = 0; r < ninst; ++r) {
    const size_t i = order[r];
    int best = 0;
    for (int d = 1; d < n; ++d)
      if (load[size_t(d)] < load[size_t(best)]) best = d;
    dom[i] = best;
    load[size_t(best)] += inst_work[i];
    rank[i] = r;
  }
  std::vector<std::vector<BatchItem>> lists(static_cast<size_t>(n));
  std::vector<BatchItem> sorted = items;   // instance-major, then by rank
  std::stable_sort(sorted.begin(), sorted.end(), [&](const BatchItem& x, const BatchItem& y) {
    return rank[size_t(x.inst)] < rank[size_t(y.inst)];
  });
  for (const BatchItem& it : sorted) lists[size_t(dom[size_t(it.inst)])].push_back(it);
  size_t longest = 0;
  for (auto& l : lists) {
    std::stable_sort(l.begin(), l.end(),
                     [](const BatchItem& x, const BatchItem& y) { return x.work > y.work; });
    longest = std::max(longest, l.size());
  }
  std::vector<BatchItem> out(longest * size_t(n), BatchItem{-1, 0, 0, 0});
  for (int d = 0; d < n; ++d)
    for (size_t k = 0; k < lists[size_t(d)].size(); ++k)
      out[k * size_t(n) + size_t(d)] = lists[size_t(d)][k];
  return out;
}

int build_phase(t1m_batch& bt, int kern, const std::vector<size_t>& which) {
  if (which.empty()) return T1M_SUCCESS;
  const GemmProblem& g0 = bt.problems[which[0]];
  const size_t asz = batch_args_size(g0.idx32 != 0);
  std::vector<char> args(asz * which.size());
  std::vector<BatchItem> items;
  std::vector<int64_t> inst_work(which.size(), 0);
  for (size_t w = 0; w < which.size(); ++w) {
    int tm = 0, tn = 0;
    batch_make_args(bt.problems[which[w]], kern, args.data() + asz * w, &tm, &tn);
    const int32_t work = int32_t(bt.problems[which[w]].K);
    inst_work[w] = int64_t(tm) * tn * work;
    for (int a = 0; a < tm; ++a)
      for (int b = 0; b < tn; ++b) items.push_back({int32_t(w), a, b, work});
  }
  const size_t ntiles = items.size();
  bool pairs = true;
  for (size_t w : which) pairs = pairs && bt.problems[w].a_pairs && bt.problems[w].b_pairs;
  const int per_cu =
      batch_blocks_per_cu(kern, g0.idx32 != 0, g0.a_mfast != 0, g0.b_kfast != 0, pairs);
  const int64_t slots = int64_t(cu_count_batch()) * per_cu;
  const int nx = xcd_count();
  if (nx > 1 && slots % nx == 0) {
    items = xcd_order(items, which.size(), inst_work, nx);
  } else {
    std::stable_sort(items.begin(), items.end(),
                     [](const BatchItem& x, const BatchItem& y) { return x.work > y.work; });
  }
  if (items.size() > size_t(INT32_MAX))
    return set_error(T1M_ERROR_UNSUPPORTED, "batch has too many tiles");
  bt.phases.emplace_back();
  BatchPhase& ph = bt.phases.back();   // freed with the batch, also on failure
  ph.kernel = kern;
  ph.ninst = int(which.size());
  const int ndom = (nx > 1 && slots % nx == 0) ? nx : 1;
  if (hipMalloc(&ph.d_args, args.size()) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&ph.d_items), items.size() * sizeof(BatchItem)) !=
          hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&ph.d_tickets), sizeof(unsigned) * (ndom + 1)) !=
          hipSuccess) {
    (void)hipGetLastError();
    return set_error(T1M_ERROR_OUT_OF_MEMORY, "hipMalloc for the batch work list failed");
  }
  HIP_TRY(hipMemcpy(ph.d_args, args.data(), args.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ph.d_items, items.data(), items.size() * sizeof(BatchItem),
                    hipMemcpyHostToDevice));
  // zeroed once here (not in the captured graph); the kernels leave them zero
  HIP_TRY(hipMemset(ph.d_tickets, 0, sizeof(unsigned) * (ndom + 1)));
  ph.grid = int(std::min<int64_t>(slots, int64_t(items.size())));
  ph.nitems = int(items.size());
  ph.ntiles = int(ntiles);
  ph.xcds = ndom;
  ph.pairs = pairs;
  return T1M_SUCCESS;
}

// The persistent phases for tiling mode `mode`, largest tiles first; none
// when the instances cannot share one kernel.
int build_persistent(t1m_batch& bt, int precision, int mode) {
  const size_t n = bt.problems.size();
  bt.mode = mode;
  if (n == 0) return T1M_SUCCESS;
  const GemmProblem& g0 = bt.problems[0];
  const int family = batch_kernel(g0.dtype, precision);
  if (family == K_SIMPLE) return T1M_SUCCESS;
  for (const GemmProblem& g : bt.problems)
    if (g.a_mfast != g0.a_mfast || g.b_kfast != g0.b_kfast || g.idx32 != g0.idx32 ||
        g.K >= (int64_t(1) << 31))
      return T1M_SUCCESS;
  std::vector<int> kerns;
  std::vector<std::vector<size_t>> groups;
  for (size_t i = 0; i < n; ++i) {
    const int k = kernel_for(bt.problems[i], family, mode);
    size_t gi = 0;
    while (gi < kerns.size() && kerns[gi] != k) ++gi;
    if (gi == kerns.size()) {
      kerns.push_back(k);
      groups.emplace_back();
    }
    groups[gi].push_back(i);
  }
  // Larger tiles first (their items are longer).
  std::vector<size_t> order(kerns.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
    return kernel_info(kerns[x]).bm > kernel_info(kerns[y]).bm;
  });
  int rc;
  for (size_t gi : order)
    if ((rc = build_phase(bt, kerns[gi], groups[gi]))) return rc;
  bt.persistent = true;
  return T1M_SUCCESS;
}

// Captures the batch's launches (persistent phases, or one launch per
// instance) into a graph and instantiates it.
int capture_batch(t1m_batch& bt) {
  HIP_TRY(hipStreamCreateWithFlags(&bt.capture, hipStreamNonBlocking));
  HIP_TRY(hipStreamBeginCapture(bt.capture, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  if (bt.persistent) {
    const GemmProblem& g0 = bt.problems[0];
    for (size_t i = 0; i < bt.phases.size() && e == hipSuccess; ++i) {
      const BatchPhase& ph = bt.phases[i];
      e = launch_contract_batched(ph.kernel, g0.idx32 != 0, g0.a_mfast != 0, g0.b_kfast != 0,
                                  ph.pairs, ph.d_args, ph.d_items, ph.nitems, ph.grid,
                                  ph.d_tickets, ph.xcds, bt.capture);
    }
  } else {
    for (size_t i = 0; i < bt.problems.size() && e == hipSuccess; ++i)
      e = launch_contract(bt.problems[i], bt.plans[i]->kernel.load(), bt.capture);
  }
  if (e != hipSuccess) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(bt.capture, &g);
    if (g) (void)hipGraphDestroy(g);
    return set_error(T1M_ERROR_DEVICE, "batch capture failed: %s", hipGetErrorString(e));
  }
  HIP_TRY(hipStreamEndCapture(bt.capture, &bt.graph));
  HIP_TRY(hipGraphInstantiate(&bt.exec, bt.graph, nullptr, nullptr, 0));
  return T1M_SUCCESS;
}

// Key of a batch in the tuning file: device, dtype, precision and a hash of
// every problem's shape and gather directions.
std::string batch_key(const t1m_batch& bt, int precision) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&h](int64_t v) {
    for (int i = 0; i < 8; ++i) h = (h ^ uint8_t(uint64_t(v) >> (8 * i))) * 1099511628211ull;
  };
  for (const GemmProblem& g : bt.problems) {
    mix(g.M); mix(g.N); mix(g.K);
    mix(g.a_mfast); mix(g.b_kfast); mix(g.idx32); mix(g.c_overlap);
  }
  char buf[160];
  snprintf(buf, sizeof(buf), "batch|%s|dt%d|p%d|n%zu|h%016llx", device_tag().c_str(),
           bt.problems.empty() ? 0 : bt.problems[0].dtype, precision, bt.problems.size(),
           (unsigned long long)h);
  return buf;
}

// Times tiling mode `mode` on a copy of the batch whose outputs are moved
// into `scratch` (C's values there are arbitrary: only time is measured).
int time_mode(const t1m_batch& bt, int precision, int mode,
              const std::vector<std::ptrdiff_t>& shift, float* ms_out) {
  t1m_batch trial;
  trial.plans = bt.plans;
  trial.problems = bt.problems;
  for (size_t i = 0; i < trial.problems.size(); ++i)
    trial.problems[i].C = static_cast<char*>(trial.problems[i].C) + shift[i];
  int rc;
  if ((rc = build_persistent(trial, precision, mode))) return rc;
  if ((rc = capture_batch(trial))) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return set_error(T1M_ERROR_DEVICE, "hipEventCreate for batch tuning failed");
  }
  hipError_t e = hipGraphLaunch(trial.exec, trial.capture);   // warm-up
  if (e == hipSuccess) e = hipGraphLaunch(trial.exec, trial.capture);
  float best = 1e30f;
  for (int rep = 0; rep < 5 && e == hipSuccess; ++rep) {
    e = hipEventRecord(e0, trial.capture);
    if (e == hipSuccess) e = hipGraphLaunch(trial.exec, trial.capture);
    if (e == hipSuccess) e = hipEventRecord(e1, trial.capture);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && ms < best) best = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess)
    return set_error(T1M_ERROR_DEVICE, "batch tuning launch failed: %s", hipGetErrorString(e));
  *ms_out = best;
  return T1M_SUCCESS;
}

void describe(t1m_batch& bt) {
  char buf[128];
  bt.description.clear();
  if (!bt.persistent) {
    snprintf(buf, sizeof(buf), "per-instance launches: %zu", bt.problems.size());
    bt.description = buf;
    return;
  }
  snprintf(buf, sizeof(buf), "mode %d", bt.mode);
  bt.description = buf;
  for (const BatchPhase& ph : bt.phases) {
    snprintf(buf, sizeof(buf), "; %s: %d instances, %d tiles, grid %d, %d L2 domains",
             kernel_info(ph.kernel).name, ph.ninst, ph.ntiles, ph.grid, ph.xcds);
    bt.description += buf;
  }
}

}  // namespace

extern "C" {

int t1m_batch_create(int64_t count, const t1m_tensor* a, const char* la,
                     const t1m_tensor* b, const char* lb, const t1m_tensor* c,
                     const char* lc, const t1m_options* opt, t1m_batch** out) {
  clear_error();
  if (!out || count < 0 || (count > 0 && (!a || !b || !c)))
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "bad batch arguments");
  *out = nullptr;
  t1m_options o;
  if (opt) o = *opt; else t1m_options_init(&o);
  if (o.precision == T1M_PRECISION_BF16X3 && count > 0 && a[0].dtype != T1M_COMPLEX64)
    return set_error(T1M_ERROR_UNSUPPORTED, "BF16x3 needs complex<float> data");
  std::unique_ptr<t1m_batch> bt(new t1m_batch());
  bt->synchronize = o.synchronize;
  const hipStream_t stream = static_cast<hipStream_t>(o.stream);
  int rc;
  struct Span { uintptr_t lo, hi; };
  std::vector<Span> spans;              // C byte range of each instance
  std::vector<size_t> owner;            // instance (index into spans) of each problem
  for (int64_t i = 0; i < count; ++i) {
    Problem P;
    if ((rc = parse_problem(a + i, la, b + i, lb, c + i, lc, P))) return rc;
    if (i > 0 && P.dtype != a[0].dtype)
      return set_error(T1M_ERROR_INVALID_ARGUMENT, "batch instances must share one dtype");
    if (P.M == 0 || P.N == 0 || P.K == 0 || P.L == 0) continue;
    if (!is_device_pointer(a[i].data) || !is_device_pointer(b[i].data) ||
        !is_device_pointer(c[i].data))
      return set_error(T1M_ERROR_INVALID_ARGUMENT,
                       "batch instance %lld: tensors must be device memory", (long long)i);
    std::shared_ptr<Plan> plan;
    if ((rc = get_plan(P, o.precision, stream, plan))) return rc;
    GemmProblem g;
    g.A = a[i].data; g.B = b[i].data; g.C = c[i].data;
    g.rsA = plan->rsA; g.csA = plan->csA; g.rsB = plan->rsB; g.csB = plan->csB;
    g.rsC = plan->rsC; g.csC = plan->csC;
    g.M = plan->M; g.N = plan->N; g.K = plan->K;
    const bool cplx = P.dtype == T1M_COMPLEX64 || P.dtype == T1M_COMPLEX128;
    g.alpha_re = o.alpha_re; g.alpha_im = cplx ? o.alpha_im : 0.0;
    g.conj_a = cplx && o.conj_a; g.conj_b = cplx && o.conj_b;
    g.a_mfast = plan->a_mfast; g.b_kfast = plan->b_kfast;
    g.a_pairs = plan->a_pairs; g.b_pairs = plan->b_pairs;
    g.idx32 = plan->idx32; g.dtype = P.dtype;
    g.split_k = 1;   // whole tiles per workgroup (no split K)
    g.cus = 0;
    g.c_overlap = plan->c_overlap ? 1 : 0;
    const uintptr_t c0 = reinterpret_cast<uintptr_t>(c[i].data);
    const int64_t es = dtype_size(P.dtype);
    spans.push_back({c0 + plan->c_lo * es, c0 + (plan->c_hi + 1) * es});
    // An instance with batch labels contributes one problem per batch index.
    for (int64_t z = 0; z < plan->L; ++z) {
      GemmProblem gz = g;
      if (plan->L > 1) {
        gz.A = static_cast<const char*>(g.A) + plan->hboff[size_t(z)];
        gz.B = static_cast<const char*>(g.B) + plan->hboff[size_t(plan->L + z)];
        gz.C = static_cast<char*>(g.C) + plan->hboff[size_t(2 * plan->L + z)];
      }
      if (plan->swap_ab) transpose_problem(gz);   // as the plan's kernel choice assumes
      bt->plans.push_back(plan);
      bt->problems.push_back(gz);
      owner.push_back(spans.size() - 1);
    }
  }
  // Instances run side by side: when two write into one range of memory,
  // every instance adds atomically.
  {
    std::vector<Span> sorted = spans;
    std::sort(sorted.begin(), sorted.end(),
              [](const Span& x, const Span& y) { return x.lo < y.lo; });
    for (size_t i = 1; i < sorted.size(); ++i)
      if (sorted[i].lo < sorted[i - 1].hi) {
        for (GemmProblem& g : bt->problems) g.c_overlap = 1;
        break;
      }
  }

  // Tiling mode: from the tuning file, else timed (autotune), else mode 0.
  const std::vector<int> modes = tiling_modes(bt->problems.empty() ? 0 : bt->problems[0].dtype,
                                              o.precision);
  int mode = 0;
  const char* forced = env_override("T1M_BATCH_MODE");   // tests / tuning
  if (forced && std::find(modes.begin(), modes.end(), atoi(forced)) != modes.end()) {
    mode = atoi(forced);
  } else if (modes.size() > 1 && !bt->problems.empty()) {
    const std::string key = batch_key(*bt, o.precision);
    const int known = tuned_batch_mode(key);
    if (known >= 0 && std::find(modes.begin(), modes.end(), known) != modes.end()) {
      mode = known;
    } else if (o.autotune) {
      // Scratch outputs: each instance's C span, packed one after another.
      std::vector<size_t> offs(spans.size());
      size_t total = 0;
      for (size_t i = 0; i < spans.size(); ++i) {
        offs[i] = total;
        total += (spans[i].hi - spans[i].lo + 255) / 256 * 256;
      }
      char* scratch = nullptr;
      if (hipMalloc(&scratch, total) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(T1M_ERROR_OUT_OF_MEMORY,
                         "hipMalloc of %zu bytes for batch tuning failed", total);
      }
      std::vector<std::ptrdiff_t> shift(bt->problems.size());
      for (size_t p = 0; p < shift.size(); ++p) {
        const Span& sp = spans[owner[p]];
        shift[p] = std::ptrdiff_t(reinterpret_cast<uintptr_t>(scratch) + offs[owner[p]]) -
                   std::ptrdiff_t(sp.lo);
      }
      // The scratch outputs sit packed in one allocation, the real ones
      // wherever the caller put them: another mode replaces the default
      // only when it is faster by 3 % or more.
      float best = 1e30f;
      for (int m : modes) {
        float ms = 0;
        rc = time_mode(*bt, o.precision, m, shift, &ms);
        if (rc) break;
        if (m == modes[0] || ms < 0.97f * best) { best = ms; mode = m; }
      }
      (void)hipFree(scratch);
      if (rc) return rc;
      store_batch_mode(key, mode);
    }
  }
  if ((rc = build_persistent(*bt, o.precision, mode))) return rc;
  if (!bt->persistent && o.autotune) {
    for (size_t i = 0; i < bt->plans.size(); ++i)
      if (!bt->plans[i]->tuned.load() &&
          (rc = autotune_plan(*bt->plans[i], bt->problems[i], stream)))
        return rc;
  }
  if ((rc = capture_batch(*bt))) return rc;
  describe(*bt);
  *out = bt.release();
  return T1M_SUCCESS;
}

int t1m_batch_launch(t1m_batch* batch, void* stream) {
  clear_error();
  if (!batch) return set_error(T1M_ERROR_INVALID_ARGUMENT, "null batch");
  const hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(hipGraphLaunch(batch->exec, s));
  if (batch->synchronize) HIP_TRY(hipStreamSynchronize(s));
  return T1M_SUCCESS;
}

int t1m_batch_destroy(t1m_batch* batch) {
  clear_error();
  delete batch;
  return T1M_SUCCESS;
}

const char* t1m_batch_describe(const t1m_batch* batch) {
  return batch ? batch->description.c_str() : "";
}

}  // extern "C"
