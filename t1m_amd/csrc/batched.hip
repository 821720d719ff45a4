// Batched contractions (tensor-network sweeps): many independent
// contractions with one label triple and per-instance extents, replayed from
// one captured HIP graph.  No counterpart in the reference, whose contract
// handles one pair of tensors per call (include/t1m/t1m.h:560-577).
//
// Creation builds every instance's index maps (one plan per distinct shape),
// writes the instances' kernel arguments to one device array, and lists
// every (instance, tile) pair, sorted by decreasing contracted extent (the
// work of a tile is proportional to K: longest-processing-time-first order).
// The graph holds a single persistent launch of a few workgroups per CU that
// walk that list with a stride of the grid size.  When the instances cannot
// share one kernel (mixed gather directions or offset widths, or the simple
// kernel), the graph holds one launch per instance instead.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "internal.h"
#include "runtime.h"

using namespace t1m_hip;

struct t1m_batch {
  std::vector<std::shared_ptr<Plan>> plans;
  std::vector<GemmProblem> problems;
  void* d_args = nullptr;
  BatchItem* d_items = nullptr;
  int kernel = 0;
  int nitems = 0;
  int grid = 0;
  bool persistent = false;
  hipStream_t capture = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int synchronize = 1;
  ~t1m_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (capture) (void)hipStreamDestroy(capture);
    if (d_args) (void)hipFree(d_args);
    if (d_items) (void)hipFree(d_items);
  }
};

namespace {

int cu_count_batch() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 256;
  return cus > 0 ? cus : 256;
}

// Tile family for a whole batch: the small tiles, which keep the many small
// instances of a sweep busy (one instance rarely fills the 144-wide tiles).
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

int build_persistent(t1m_batch& bt, int precision) {
  const size_t n = bt.problems.size();
  if (n == 0) return T1M_SUCCESS;
  const GemmProblem& g0 = bt.problems[0];
  const int kern = batch_kernel(g0.dtype, precision);
  if (kern == K_SIMPLE) return T1M_SUCCESS;
  for (const GemmProblem& g : bt.problems)
    if (g.a_mfast != g0.a_mfast || g.b_kfast != g0.b_kfast || g.idx32 != g0.idx32 ||
        g.K >= (int64_t(1) << 31))
      return T1M_SUCCESS;
  const size_t asz = batch_args_size(g0.idx32 != 0);
  std::vector<char> args(asz * n);
  std::vector<BatchItem> items;
  for (size_t i = 0; i < n; ++i) {
    int tm = 0, tn = 0;
    batch_make_args(bt.problems[i], kern, args.data() + asz * i, &tm, &tn);
    const int32_t work = int32_t(bt.problems[i].K);
    for (int a = 0; a < tm; ++a)
      for (int b = 0; b < tn; ++b) items.push_back({int32_t(i), a, b, work});
  }
  std::stable_sort(items.begin(), items.end(),
                   [](const BatchItem& x, const BatchItem& y) { return x.work > y.work; });
  if (items.size() > size_t(INT32_MAX))
    return set_error(T1M_ERROR_UNSUPPORTED, "batch has too many tiles");
  if (hipMalloc(&bt.d_args, args.size()) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&bt.d_items), items.size() * sizeof(BatchItem)) !=
          hipSuccess) {
    (void)hipGetLastError();
    return set_error(T1M_ERROR_OUT_OF_MEMORY, "hipMalloc for the batch work list failed");
  }
  HIP_TRY(hipMemcpy(bt.d_args, args.data(), args.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(bt.d_items, items.data(), items.size() * sizeof(BatchItem),
                    hipMemcpyHostToDevice));
  const int per_cu = batch_blocks_per_cu(kern, g0.idx32 != 0, g0.a_mfast != 0, g0.b_kfast != 0);
  const int64_t slots = int64_t(cu_count_batch()) * per_cu;
  bt.grid = int(std::min<int64_t>(slots, int64_t(items.size())));
  bt.nitems = int(items.size());
  bt.kernel = kern;
  bt.persistent = true;
  return T1M_SUCCESS;
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
  std::vector<std::pair<uintptr_t, uintptr_t>> spans;   // C byte range of each instance
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
    g.idx32 = plan->idx32; g.dtype = P.dtype;
    g.split_k = 1;   // whole tiles per workgroup (no split K)
    g.cus = 0;
    g.c_overlap = plan->c_overlap ? 1 : 0;
    const uintptr_t c0 = reinterpret_cast<uintptr_t>(c[i].data);
    const int64_t es = dtype_size(P.dtype);
    spans.emplace_back(c0 + plan->c_lo * es, c0 + (plan->c_hi + 1) * es);
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
    }
  }
  // Instances run side by side: when two write into one range of memory,
  // every instance adds atomically.
  std::sort(spans.begin(), spans.end());
  for (size_t i = 1; i < spans.size(); ++i)
    if (spans[i].first < spans[i - 1].second) {
      for (GemmProblem& g : bt->problems) g.c_overlap = 1;
      break;
    }
  if ((rc = build_persistent(*bt, o.precision))) return rc;
  if (!bt->persistent && o.autotune) {
    for (size_t i = 0; i < bt->plans.size(); ++i)
      if (!bt->plans[i]->tuned.load() &&
          (rc = autotune_plan(*bt->plans[i], bt->problems[i], stream)))
        return rc;
  }
  HIP_TRY(hipStreamCreateWithFlags(&bt->capture, hipStreamNonBlocking));
  HIP_TRY(hipStreamBeginCapture(bt->capture, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  if (bt->persistent) {
    const GemmProblem& g0 = bt->problems[0];
    e = launch_contract_batched(bt->kernel, g0.idx32 != 0, g0.a_mfast != 0, g0.b_kfast != 0,
                                bt->d_args, bt->d_items, bt->nitems, bt->grid, bt->capture);
  } else {
    for (size_t i = 0; i < bt->problems.size() && e == hipSuccess; ++i)
      e = launch_contract(bt->problems[i], bt->plans[i]->kernel.load(), bt->capture);
  }
  if (e != hipSuccess) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(bt->capture, &g);
    if (g) (void)hipGraphDestroy(g);
    return set_error(T1M_ERROR_DEVICE, "batch capture failed: %s", hipGetErrorString(e));
  }
  HIP_TRY(hipStreamEndCapture(bt->capture, &bt->graph));
  HIP_TRY(hipGraphInstantiate(&bt->exec, bt->graph, nullptr, nullptr, 0));
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

}  // extern "C"
