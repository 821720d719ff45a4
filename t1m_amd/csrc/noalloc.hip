// t1m_contract_noalloc: t1m_contract for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  No counterpart in the reference.
//
// The only device memory t1m_contract needs beyond A, B and C is a plan's
// index maps (six scatter vectors and, with batch labels, the batch byte
// offsets), which the plan cache allocates once per shape.  Here they live
// in `scratch`, a file-scope device array of kScratchBytes, and are rebuilt
// on every call (one k_scatter_maps launch, two with batch labels) ahead of
// the contraction kernel on the same stream.  Nothing waits in between.
//
// Every call on a GPU shares that GPU's array, which keeps the last call's
// maps: calls that overlap on one GPU (other streams, other threads, or an
// earlier call's kernels still running on another stream) race on it.
// Callers serialise them (one stream, or a wait between streams).  The
// kernels write every map entry they read, so nothing needs clearing.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "internal.h"
#include "runtime.h"

using namespace t1m_hip;

// T1M_NOALLOC_SCRATCH_BYTES (8 MiB): the maps of M + N + K up to ~1M entries
// (int32 offsets) or ~0.5M (int64), plus batch offsets.  Each GPU the
// library runs on holds one copy until the process ends.
constexpr size_t kScratchBytes = T1M_NOALLOC_SCRATCH_BYTES;

static __device__ __align__(256) char scratch[kScratchBytes];

namespace {

// The array's device address on each GPU.  In host code `scratch` names a
// host-side stand-in, never device memory: kernels get these pointers.
std::mutex g_scratch_mu;
std::vector<char*>* g_scratch_ptr = new std::vector<char*>();

int scratch_on_current_device(char** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (size_t(dev) >= g_scratch_ptr->size()) g_scratch_ptr->resize(size_t(dev) + 1, nullptr);
  char*& p = (*g_scratch_ptr)[size_t(dev)];
  if (!p) {
    void* addr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !addr)
      return set_error(T1M_ERROR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                       hipGetErrorString(e));
    p = static_cast<char*>(addr);
  }
  *out = p;
  return T1M_SUCCESS;
}

}  // namespace

extern "C" {

int t1m_contract_noalloc(const t1m_tensor* a, const char* la, const t1m_tensor* b,
                         const char* lb, const t1m_tensor* c, const char* lc,
                         const t1m_options* opt) {
  clear_error();
  t1m_options o;
  if (opt) o = *opt; else t1m_options_init(&o);
  if (o.precision < 0 || o.precision > T1M_PRECISION_SIMPLE)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "unknown precision %d", o.precision);
  if (o.autotune)
    return set_error(T1M_ERROR_INVALID_ARGUMENT,
                     "autotune times candidates on an allocated scratch C; "
                     "not available without allocation");
  Problem P0;
  int rc;
  if ((rc = parse_problem(a, la, b, lb, c, lc, P0))) return rc;
  if (o.precision == T1M_PRECISION_BF16X3 && P0.dtype != T1M_COMPLEX64)
    return set_error(T1M_ERROR_UNSUPPORTED, "BF16x3 needs complex<float> data");
  if (P0.M == 0 || P0.N == 0 || P0.K == 0 || P0.L == 0) return T1M_SUCCESS;   // C += 0
  if (!a->data || !b->data || !c->data)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "null data pointer");
  if (!is_device_pointer(a->data) || !is_device_pointer(b->data) || !is_device_pointer(c->data))
    return set_error(T1M_ERROR_INVALID_ARGUMENT,
                     "t1m_contract_noalloc needs device memory (host tensors are staged "
                     "through allocated buffers)");
  Plan pl;   // host-side; its maps point into the scratch array (pl.maps stays null)
  Problem P;
  plan_layout(P0, o.precision, pl, P);
  apply_tuning_file(pl);
  const size_t need = plan_device_bytes(pl);
  if (need > kScratchBytes)
    return set_error(T1M_ERROR_UNSUPPORTED,
                     "index maps need %zu bytes; the no-allocation scratch holds %zu "
                     "(M + N + K too large: use t1m_contract)",
                     need, kScratchBytes);
  char* base = nullptr;
  if ((rc = scratch_on_current_device(&base))) return rc;
  const hipStream_t stream = static_cast<hipStream_t>(o.stream);
  hipError_t e = plan_build_maps(pl, P, base, stream);
  if (e != hipSuccess)
    return set_error(T1M_ERROR_DEVICE, "index map construction failed: %s", hipGetErrorString(e));
  const GemmProblem g = gemm_for(pl, a->data, b->data, c->data, o);
  e = launch_contract(g, pl.kernel, stream);
  if (e != hipSuccess)
    return set_error(T1M_ERROR_DEVICE, "kernel %s launch failed: %s",
                     kernel_info(pl.kernel.load()).name, hipGetErrorString(e));
  if (o.synchronize) HIP_TRY(hipStreamSynchronize(stream));
  return T1M_SUCCESS;
}

}  // extern "C"
