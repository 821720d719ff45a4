// Host-side runtime declarations (plans, problems, staging, errors).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "internal.h"

namespace t1m_hip {

int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();
const char* last_error();

#define HIP_TRY(expr)                                                         \
  do {                                                                        \
    const hipError_t e_ = (expr);                                             \
    if (e_ != hipSuccess)                                                     \
      return ::t1m_hip::set_error(T1M_ERROR_DEVICE, "%s: %s", #expr,          \
                                  hipGetErrorString(e_));                     \
  } while (0)

// One mode (label) of the contraction with its extent and its stride in
// each operand (0 where the operand lacks the mode).
struct Mode {
  char label;
  int64_t dim;
  int64_t sa, sb, sc;
};

// A validated contraction: free modes of A (m), of B (n), contracted (k),
// each in the reference's alphabetical bundle order (t1m.h:99-133).
struct Problem {
  int dtype = 0;
  std::string la, lb, lc;
  std::vector<Mode> m, n, k;
  std::vector<Mode> batch;          // labels in A, B and C (einsum batch modes)
  int64_t M = 1, N = 1, K = 1, L = 1;
  std::vector<int64_t> da, sa, db, sb, dc, sc;
};

// Device-resident index maps for one (dtype, labels, extents, strides).
struct Plan {
  int dtype = 0;
  int precision = 0;
  std::atomic<int> kernel{0};
  std::atomic<bool> tuned{false};
  std::mutex tune_mu;
  int64_t c_lo = 0, c_hi = 0;                    // C's element span
  bool c_overlap = false;                        // two index tuples of C may share an address
  bool swap_ab = false;                          // run as C^T = B^T A^T (thin m, see get_plan)
  int64_t M = 0, N = 0, K = 0, L = 1;
  int a_mfast = 1, b_kfast = 1, idx32 = 1;
  int a_pairs = 0, b_pairs = 0;                  // two-element gathers allowed (GemmProblem)
  int64_t min_rsA = 0, min_csA = 0, min_rsB = 0, min_csB = 0, min_rsC = 0, min_csC = 0;
  int span32 = 0;                                // A, B spans < 2^32 bytes (GemmProblem)
  void* maps = nullptr;                          // owned device memory (null: borrowed)
  // Batch problems (L > 1): byte offsets into A, B, C, device [3][L] and host.
  int64_t* boff = nullptr;
  std::vector<int64_t> hboff;
  const void *rsA = nullptr, *rsC = nullptr, *csA = nullptr, *rsB = nullptr,
             *csB = nullptr, *csC = nullptr;
  std::vector<Mode> modes_m, modes_n, modes_k;   // kernel traversal order
  ~Plan();
};

// A host tensor copied through device memory for one call.  The device
// buffer comes from a per-device pool of staging buffers and goes back to
// it when the call ends; a call that did not reach its final stream
// synchronisation (an error path) synchronises before giving it back, so no
// queued copy or kernel still uses a pooled buffer.
struct Staged {
  void* buf = nullptr;
  size_t cap = 0;                // pooled buffer's size (>= bytes)
  int dev = 0;
  hipStream_t stream = nullptr;
  bool synced = false;           // the stream was synchronised after the last use
  void* device = nullptr;        // element-0 address inside buf
  const char* host_base = nullptr;
  size_t bytes = 0;
  int stage(const void* host, const std::vector<int64_t>& dims,
            const std::vector<int64_t>& strides, int esize, bool copy_in,
            hipStream_t stream);
  int copy_back(void* host, hipStream_t stream);
  ~Staged();
};
// Frees every pooled staging buffer (t1m_plan_cache_clear).
void staging_pool_clear();

int dtype_size(int dtype);
int check_tensor(const t1m_tensor* t, const char* labels, const char* what,
                 std::vector<int64_t>& strides);
int parse_problem(const t1m_tensor* a, const char* la, const t1m_tensor* b,
                  const char* lb, const t1m_tensor* c, const char* lc,
                  Problem& P);
int choose_kernel(const Problem& P, int precision);
bool may_alias(const std::vector<int64_t>& dims, const std::vector<int64_t>& strides);
// A test / tuning override from the environment (T1M_FORCE_KERNEL,
// T1M_FORCE_IDX64); unset and empty read alike (nullptr).
const char* env_override(const char* name);
void transpose_problem(GemmProblem& g);
int get_plan(const Problem& P, int precision, hipStream_t stream,
             std::shared_ptr<Plan>& out);
// The parts of get_plan: host-side layout (P receives the bundles in
// traversal order), the device bytes the maps need, and their construction
// into caller-provided device memory (queued on `stream`, not waited for).
void plan_layout(const Problem& P0, int precision, Plan& pl, Problem& P);
size_t plan_device_bytes(const Plan& pl);
hipError_t plan_build_maps(Plan& pl, const Problem& P, char* base, hipStream_t stream);
// A tuning-file entry for the plan's key, if it fits the plan (host only).
void apply_tuning_file(Plan& pl);
// "<arch>/<CUs>" of the current device (tuning-file keys).
std::string device_tag();
// A batch's tiling mode from the tuning file (-1: none), and storing one.
int tuned_batch_mode(const std::string& key);
void store_batch_mode(const std::string& key, int mode);
// The kernel arguments of a call of plan `pl` on device data A, B, C.
GemmProblem gemm_for(const Plan& pl, const void* A, const void* B, void* C,
                     const t1m_options& opt);
int cu_count();
int64_t plan_cache_size();
void plan_cache_clear();
// While `stream` is capturing, ties the plan's lifetime to the graph.
int retain_plan_for_capture(const std::shared_ptr<Plan>& plan, hipStream_t stream);
bool is_device_pointer(const void* ptr);
// Kernel ids worth timing for a plan's dtype and precision.
std::vector<int> kernel_candidates(int dtype, int precision);
// Whether kernel id `kernel` computes this dtype at this precision.
bool kernel_fits(int kernel, int dtype, int precision);
// Times the candidates on a scratch copy of C's extent and stores the
// fastest in the plan (and in the tuning cache file).
int autotune_plan(Plan& plan, const GemmProblem& g, hipStream_t stream);
int contract_impl(const t1m_tensor* a, const char* la, const t1m_tensor* b,
                  const char* lb, const t1m_tensor* c, const char* lc,
                  const t1m_options& opt);

}  // namespace t1m_hip
