// Batched contractions (tensor-network sweeps): many independent
// contractions with one label triple and per-instance extents, replayed as
// one captured HIP graph.  No counterpart in the reference, whose contract
// handles one pair of tensors per call (include/t1m/t1m.h:560-577).
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "internal.h"
#include "runtime.h"

using namespace t1m_hip;

struct t1m_batch {
  std::vector<std::shared_ptr<Plan>> plans;
  std::vector<GemmProblem> problems;
  hipStream_t capture = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int synchronize = 1;
  ~t1m_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (capture) (void)hipStreamDestroy(capture);
  }
};

extern "C" {

int t1m_batch_create(int64_t count, const t1m_tensor* a, const char* la,
                     const t1m_tensor* b, const char* lb, const t1m_tensor* c,
                     const char* lc, const t1m_options* opt, t1m_batch** out) {
  clear_error();
  if (!out || count < 0 || (count > 0 && (!a || !b || !c)))
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "bad batch arguments");
  t1m_options o;
  if (opt) o = *opt; else t1m_options_init(&o);
  if (o.precision == T1M_PRECISION_BF16X3 && count > 0 && a[0].dtype != T1M_COMPLEX64)
    return set_error(T1M_ERROR_UNSUPPORTED, "BF16x3 needs complex<float> data");
  std::unique_ptr<t1m_batch> bt(new t1m_batch());
  bt->synchronize = o.synchronize;
  const hipStream_t stream = static_cast<hipStream_t>(o.stream);
  int rc;
  for (int64_t i = 0; i < count; ++i) {
    Problem P;
    if ((rc = parse_problem(a + i, la, b + i, lb, c + i, lc, P))) return rc;
    if (P.M == 0 || P.N == 0 || P.K == 0) continue;
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
    if (o.autotune && !plan->tuned.load() && (rc = autotune_plan(*plan, g, stream))) return rc;
    bt->plans.push_back(plan);
    bt->problems.push_back(g);
  }
  HIP_TRY(hipStreamCreateWithFlags(&bt->capture, hipStreamNonBlocking));
  HIP_TRY(hipStreamBeginCapture(bt->capture, hipStreamCaptureModeThreadLocal));
  for (size_t i = 0; i < bt->problems.size(); ++i) {
    const hipError_t e = launch_contract(bt->problems[i], bt->plans[i]->kernel, bt->capture);
    if (e != hipSuccess) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(bt->capture, &g);
      if (g) (void)hipGraphDestroy(g);
      return set_error(T1M_ERROR_DEVICE, "capture of instance %zu failed: %s", i,
                       hipGetErrorString(e));
    }
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
