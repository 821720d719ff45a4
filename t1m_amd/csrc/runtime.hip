// Host runtime of libt1m_hip.so: the extern "C" entry points of t1m_hip.h.
//
// t1m_contract follows t1m::internal::contract (include/t1m/t1m.h:432-557):
//   1. index bundles from the labels (t1m.h:440-441 -> 90-136), validated;
//   2. scatter vectors for the rows / columns of A, B and C (t1m.h:442-444 ->
//      246-268), built on the device by one kernel launch and cached in a
//      plan keyed by dtype, labels, extents and strides;
//   3. one fused contraction kernel (gather -> MFMA -> scatter-accumulate),
//      in place of the reference's pack / micro-kernel / unpack loops.
// Where the reference keeps the bundle order of t1m.h:99-133 (alphabetical)
// for the traversal, the kernels here may traverse a bundle's modes in
// another order (a sum over K or an enumeration of M / N is order-free):
// each bundle is ordered so that the operand staged along it has its
// smallest stride first, which makes the gathers coalesce.  The exported
// index maps (t1m_index_maps) keep the reference order bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sys/stat.h>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "internal.h"
#include "runtime.h"

namespace t1m_hip {

namespace {
thread_local std::string g_last_error;
}

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

void clear_error() { g_last_error.clear(); }
const char* last_error() { return g_last_error.c_str(); }

// ---------------------------------------------------------------------------
// Labels -> modes
// ---------------------------------------------------------------------------

int dtype_size(int dtype) {
  switch (dtype) {
    case T1M_FLOAT32: return 4;
    case T1M_FLOAT64: return 8;
    case T1M_COMPLEX64: return 8;
    case T1M_COMPLEX128: return 16;
    default: return 0;
  }
}

int check_tensor(const t1m_tensor* t, const char* labels, const char* what,
                 std::vector<int64_t>& strides) {
  if (!t) return set_error(T1M_ERROR_INVALID_ARGUMENT, "%s: null tensor", what);
  if (!labels) return set_error(T1M_ERROR_INVALID_ARGUMENT, "%s: null labels", what);
  if (dtype_size(t->dtype) == 0)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "%s: unknown dtype %d", what, t->dtype);
  if (t->rank < 0 || t->rank > T1M_MAX_RANK)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "%s: rank %d out of range [0, %d]",
                     what, t->rank, T1M_MAX_RANK);
  if (int(strlen(labels)) != t->rank)
    return set_error(T1M_ERROR_INVALID_ARGUMENT,
                     "%s: %d labels \"%s\" for a rank-%d tensor", what,
                     int(strlen(labels)), labels, t->rank);
  if (t->rank > 0 && !t->dims)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "%s: null dims", what);
  for (int i = 0; i < t->rank; ++i) {
    if (t->dims[i] < 0)
      return set_error(T1M_ERROR_INVALID_ARGUMENT, "%s: negative extent", what);
    for (int j = 0; j < i; ++j)
      if (labels[i] == labels[j])
        return set_error(T1M_ERROR_INVALID_ARGUMENT,
                         "%s: label '%c' repeated in \"%s\"", what, labels[i], labels);
  }
  strides.assign(t->rank, 0);
  if (t->strides) {
    for (int i = 0; i < t->rank; ++i) strides[i] = t->strides[i];
  } else {
    int64_t acc = 1;   // column-major, t1m.h:42-50
    for (int i = 0; i < t->rank; ++i) { strides[i] = acc; acc *= t->dims[i]; }
  }
  return T1M_SUCCESS;
}

int parse_problem(const t1m_tensor* a, const char* la, const t1m_tensor* b,
                  const char* lb, const t1m_tensor* c, const char* lc,
                  Problem& P) {
  std::vector<int64_t> sa, sb, sc;
  int rc;
  if ((rc = check_tensor(a, la, "A", sa))) return rc;
  if ((rc = check_tensor(b, lb, "B", sb))) return rc;
  if ((rc = check_tensor(c, lc, "C", sc))) return rc;
  if (a->dtype != b->dtype || a->dtype != c->dtype)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "A, B and C must share one dtype");
  P.dtype = a->dtype;
  P.la = la; P.lb = lb; P.lc = lc;
  P.m.clear(); P.n.clear(); P.k.clear(); P.batch.clear();
  auto pos = [](const char* s, char ch) -> int {
    const char* q = strchr(s, ch);
    return (q && ch) ? int(q - s) : -1;
  };
  // Contracted and free modes, alphabetical (t1m.h:99-111).
  std::string sorted_a(la), sorted_b(lb);
  std::sort(sorted_a.begin(), sorted_a.end());
  std::sort(sorted_b.begin(), sorted_b.end());
  for (char ch : sorted_a) {
    const int pa = pos(la, ch), pb = pos(lb, ch), pc = pos(lc, ch);
    if (pb >= 0 && pc >= 0) {
      // In A, B and C: a batch mode, C[..z..] = sum_k A[..z..] B[..z..]
      // (einsum semantics; the reference would sum over it, t1m.h:106-119).
      if (a->dims[pa] != b->dims[pb] || a->dims[pa] != c->dims[pc])
        return set_error(T1M_ERROR_INVALID_ARGUMENT,
                         "extent mismatch for batch label '%c'", ch);
      P.batch.push_back({ch, a->dims[pa], sa[pa], sb[pb], sc[pc]});
    } else if (pb >= 0) {
      if (a->dims[pa] != b->dims[pb])
        return set_error(T1M_ERROR_INVALID_ARGUMENT,
                         "extent mismatch for contracted label '%c': %lld vs %lld", ch,
                         (long long)a->dims[pa], (long long)b->dims[pb]);
      P.k.push_back({ch, a->dims[pa], sa[pa], sb[pb], 0});
    } else {
      if (pc < 0)
        return set_error(T1M_ERROR_INVALID_ARGUMENT,
                         "label '%c' of A is neither in B nor in C", ch);
      if (a->dims[pa] != c->dims[pc])
        return set_error(T1M_ERROR_INVALID_ARGUMENT,
                         "extent mismatch for label '%c' of A and C", ch);
      P.m.push_back({ch, a->dims[pa], sa[pa], 0, sc[pc]});
    }
  }
  for (char ch : sorted_b) {
    const int pb = pos(lb, ch), pc = pos(lc, ch);
    if (pos(la, ch) >= 0) continue;
    if (pc < 0)
      return set_error(T1M_ERROR_INVALID_ARGUMENT,
                       "label '%c' of B is neither in A nor in C", ch);
    if (b->dims[pb] != c->dims[pc])
      return set_error(T1M_ERROR_INVALID_ARGUMENT,
                       "extent mismatch for label '%c' of B and C", ch);
    P.n.push_back({ch, b->dims[pb], 0, sb[pb], sc[pc]});
  }
  for (const char* q = lc; *q; ++q)
    if (pos(la, *q) < 0 && pos(lb, *q) < 0)
      return set_error(T1M_ERROR_INVALID_ARGUMENT,
                       "label '%c' of C is in neither A nor B", *q);
  P.M = P.N = P.K = P.L = 1;
  for (auto& x : P.m) P.M *= x.dim;
  for (auto& x : P.n) P.N *= x.dim;
  for (auto& x : P.k) P.K *= x.dim;
  for (auto& x : P.batch) P.L *= x.dim;
  if (P.L > INT32_MAX)
    return set_error(T1M_ERROR_UNSUPPORTED, "more than 2^31 - 1 batch problems");
  P.sa = sa; P.sb = sb; P.sc = sc;
  P.da.assign(a->dims, a->dims + a->rank);
  P.db.assign(b->dims, b->dims + b->rank);
  P.dc.assign(c->dims, c->dims + c->rank);
  return T1M_SUCCESS;
}

// ---------------------------------------------------------------------------
// Kernel choice
// ---------------------------------------------------------------------------

int cu_count() {
  static int n = [] {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 256;
    return cus > 0 ? cus : 256;
  }();
  return n;
}

std::vector<int> kernel_candidates(int dtype, int precision) {
  if (precision == T1M_PRECISION_SIMPLE) return {K_SIMPLE};
  switch (dtype) {
    case T1M_COMPLEX64:
      if (precision == T1M_PRECISION_BF16X3) return {K_C64_BF16X3_144, K_C64_BF16X3_64};
      return {K_C64_F32_144, K_C64_F32_64, K_C64_F32_THIN};
    case T1M_COMPLEX128:
      return {K_C128_F64_64, K_C128_F64_64W, K_C128_F64_32, K_C128_F64_THIN};
    case T1M_FLOAT32:
      return {K_F32_128, K_F32_64};
    case T1M_FLOAT64:
      return {K_F64_64, K_F64_32};
    default:
      return {K_SIMPLE};
  }
}

bool kernel_fits(int kernel, int dtype, int precision) {
  if (kernel == K_SIMPLE) return true;
  for (int k : kernel_candidates(dtype, precision))
    if (k == kernel) return true;
  return false;
}

// Static rule, used when no tuning result is known.  complex<float>: the
// thin 128 x 16 tiles for C at most 16 wide; else the 144 x 144 tiles (one
// workgroup of 2 waves per SIMD per CU) once they nearly fill the CUs and a
// relative-time model favours them: 1 / (rate x tile utilisation x wave
// efficiency), rates 0.72 (144) and 0.57 (64) of peak for FP32 and 1.0 /
// 0.6 for BF16x3 as measured on 48^4, utilisation = the useful fraction of
// the padded tile grid, wave efficiency = tiles / (rounds x CUs) for the
// one-per-CU large tiles (the small ones run 3 per CU and fill in).
// Calibrated on the 48^4, 256x4x256 and reference-suite contractions.
int choose_kernel(const Problem& P, int precision) {
  if (P.K >= (int64_t(1) << 31)) return K_SIMPLE;     // MFMA kernels use int k
  if (const char* f = env_override("T1M_FORCE_KERNEL")) {   // tests / tuning; ignored
    const int k = atoi(f);                            // when it does not fit
    if (kernel_fits(k, P.dtype, precision)) return k;
  }
  if (precision == T1M_PRECISION_SIMPLE) return K_SIMPLE;
  auto tiles = [&](int bm, int bn) {
    return double((P.M + bm - 1) / bm) * double((P.N + bn - 1) / bn) * double(P.L);
  };
  switch (P.dtype) {
    case T1M_COMPLEX64: {
      const bool bf = precision == T1M_PRECISION_BF16X3;
      if (!bf && P.N <= 16) return K_C64_F32_THIN;   // one 16-column block
      auto util = [&](int bm, int bn) {
        return double(P.M) / double((P.M + bm - 1) / bm * bm) *
               (double(P.N) / double((P.N + bn - 1) / bn * bn));
      };
      const double cus = cu_count(), t144 = tiles(144, 144);
      const double wave = t144 / (std::ceil(t144 / cus) * cus);
      const double r144 = bf ? 1.0 : 0.72, r64 = bf ? 0.6 : 0.57;
      const bool big = t144 >= 0.75 * cus &&
                       r144 * util(144, 144) * wave >= r64 * util(64, 64);
      if (bf) return big ? K_C64_BF16X3_144 : K_C64_BF16X3_64;
      return big ? K_C64_F32_144 : K_C64_F32_64;
    }
    case T1M_COMPLEX128:
      if (P.N <= 16) return K_C128_F64_THIN;
      return tiles(64, 64) >= cu_count() ? K_C128_F64_64 : K_C128_F64_32;
    case T1M_FLOAT32:   // narrow C (N <= 32): the small tiles waste less
      if (P.N <= 32) return K_F32_64;
      return tiles(128, 128) >= cu_count() ? K_F32_128 : K_F32_64;
    case T1M_FLOAT64:
      if (P.N <= 32) return K_F64_32;
      return tiles(64, 64) >= cu_count() ? K_F64_64 : K_F64_32;
    default:
      return K_SIMPLE;
  }
}

// ---------------------------------------------------------------------------
// Autotuning
// ---------------------------------------------------------------------------

namespace {

struct TuneCache {
  std::mutex mu;
  bool loaded = false;
  std::unordered_map<std::string, int> best;          // key -> kernel id
  std::unordered_map<std::string, int> batch_modes;   // batch key -> tiling mode
  std::string path;
};

TuneCache& tune_cache() {
  static TuneCache* t = new TuneCache();
  return *t;
}

std::string tune_path() {
  if (const char* p = getenv("T1M_TUNE_CACHE"); p && *p) return p;
  const char* home = getenv("HOME");
  return std::string(home ? home : "/tmp") + "/.cache/t1m_mi355x/tuning.txt";
}

int kernel_by_name(const std::string& name) {
  for (int k = 0; k < K_COUNT; ++k)
    if (name == kernel_info(k).name) return k;
  return -1;
}

void load_tune_cache(TuneCache& T) {
  if (T.loaded) return;
  T.loaded = true;
  T.path = tune_path();
  FILE* f = fopen(T.path.c_str(), "r");
  if (!f) return;
  char key[512], name[128];
  while (fscanf(f, "%511s %127s", key, name) == 2) {
    if (strncmp(name, "batch_mode_", 11) == 0) {
      const int m = atoi(name + 11);
      if (m >= 0 && m < 16) T.batch_modes[key] = m;
      continue;
    }
    const int k = kernel_by_name(name);
    if (k >= 0) T.best[key] = k;
  }
  fclose(f);
}

void append_tune_line(TuneCache& T, const std::string& key, const char* value) {
  // Create the file's directory without a shell: mkdir each prefix (a bare
  // file name lives in the working directory).
  const size_t slash = T.path.find_last_of('/');
  const std::string dir = slash == std::string::npos ? std::string() : T.path.substr(0, slash);
  for (size_t i = 1; i <= dir.size(); ++i)
    if (i == dir.size() || dir[i] == '/') (void)mkdir(dir.substr(0, i).c_str(), 0755);
  FILE* f = fopen(T.path.c_str(), "a");
  if (!f) return;
  fprintf(f, "%s %s\n", key.c_str(), value);
  fclose(f);
}

void append_tune_cache(TuneCache& T, const std::string& key, int kernel) {
  append_tune_line(T, key, kernel_info(kernel).name);
}

}  // namespace

std::string device_tag() {
  int dev = 0;
  (void)hipGetDevice(&dev);
  hipDeviceProp_t prop;
  std::string arch = "gpu";
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) arch = prop.gcnArchName;
  for (char& ch : arch) if (ch == ':' || ch == ' ') ch = '_';
  return arch + "/" + std::to_string(cu_count());
}

int tuned_batch_mode(const std::string& key) {
  TuneCache& T = tune_cache();
  std::lock_guard<std::mutex> lock(T.mu);
  load_tune_cache(T);
  auto it = T.batch_modes.find(key);
  return it == T.batch_modes.end() ? -1 : it->second;
}

void store_batch_mode(const std::string& key, int mode) {
  TuneCache& T = tune_cache();
  std::lock_guard<std::mutex> lock(T.mu);
  load_tune_cache(T);
  T.batch_modes[key] = mode;
  append_tune_line(T, key, ("batch_mode_" + std::to_string(mode)).c_str());
}

namespace {

// A kernel id read from the tuning file is used only if it computes the
// plan's dtype at its precision and the plan's K fits the MFMA kernels' int
// k (a stale or edited file must not bind, say, an FP64 kernel to complex
// float data).
bool tuned_kernel_ok(const Plan& pl, int kernel) {
  if (kernel < 0 || kernel >= K_COUNT || !kernel_fits(kernel, pl.dtype, pl.precision))
    return false;
  return kernel == K_SIMPLE || pl.K < (int64_t(1) << 31);
}

std::string tune_key(const Plan& pl) {
  char buf[256];
  snprintf(buf, sizeof(buf), "%s|dt%d|p%d|M%lld|N%lld|K%lld|a%d|b%d|i%d", device_tag().c_str(),
           pl.dtype, pl.precision, (long long)pl.M, (long long)pl.N,
           (long long)pl.K, pl.a_mfast, pl.b_kfast, pl.idx32);
  if (pl.L > 1) {
    const size_t n = strlen(buf);
    snprintf(buf + n, sizeof(buf) - n, "|L%lld", (long long)pl.L);
  }
  if (pl.swap_ab) {
    const size_t n = strlen(buf);
    snprintf(buf + n, sizeof(buf) - n, "|T");
  }
  return buf;
}

}  // namespace

int autotune_plan(Plan& pl, const GemmProblem& g0, hipStream_t stream) {
  std::lock_guard<std::mutex> guard(pl.tune_mu);
  if (pl.tuned.load()) return T1M_SUCCESS;
  const std::vector<int> cands = kernel_candidates(pl.dtype, pl.precision);
  if (cands.size() < 2 || env_override("T1M_FORCE_KERNEL")) {
    pl.tuned = true;
    return T1M_SUCCESS;
  }
  TuneCache& T = tune_cache();
  const std::string key = tune_key(pl);
  {
    std::lock_guard<std::mutex> lock(T.mu);
    load_tune_cache(T);
    auto it = T.best.find(key);
    if (it != T.best.end() && tuned_kernel_ok(pl, it->second)) {
      pl.kernel = it->second;
      pl.tuned = true;
      return T1M_SUCCESS;
    }
  }
  const int es = dtype_size(pl.dtype);
  const size_t bytes = size_t(pl.c_hi - pl.c_lo + 1) * es;
  char* scratch = nullptr;
  if (hipMalloc(&scratch, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(T1M_ERROR_OUT_OF_MEMORY, "hipMalloc of %zu bytes for autotuning failed", bytes);
  }
  GemmProblem g = g0;
  g.C = scratch - pl.c_lo * es;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    const hipError_t err = hipGetLastError();
    if (e0) (void)hipEventDestroy(e0);
    (void)hipFree(scratch);
    return set_error(T1M_ERROR_DEVICE, "hipEventCreate for autotuning failed: %s",
                     hipGetErrorString(err));
  }
  // Each candidate runs as a burst of back-to-back launches (sized to about 6 ms from one
  // single launch, at most 256): a single launch after an idle GPU pays a start-up cost of
  // the order of the 4 % that separates some candidates (44 against 46 us at 1024 x 1024 x
  // 256 complex<double>), and bursts of 16 such launches (0.7 ms) still picked the slower
  // one on some devices.  Three rounds, the middle one in the opposite candidate order, so a
  // clock that is still settling favours none of them; a candidate's time is its fastest
  // round.
  std::vector<int> todo;
  for (int k : cands) {
    // thin tiles only where C is narrow (elsewhere they cannot win)
    if ((k == K_C64_F32_THIN || k == K_C128_F64_THIN) && (pl.swap_ab ? pl.M : pl.N) > 64)
      continue;
    todo.push_back(k);
  }
  std::vector<float> per(todo.size(), 1e30f);
  std::vector<int> burst(todo.size(), 1);
  hipError_t e = hipSuccess;
  int failed = -1;
  auto time_launches = [&](int k, int n, float& ms) {
    e = hipEventRecord(e0, stream);
    for (int i = 0; i < n && e == hipSuccess; ++i) e = launch_contract(g, k, stream);
    if (e == hipSuccess) e = hipEventRecord(e1, stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  };
  for (int round = 0; round < 3 && e == hipSuccess; ++round) {
    for (size_t i = 0; i < todo.size() && e == hipSuccess; ++i) {
      const size_t c = round != 1 ? i : todo.size() - 1 - i;
      const int k = todo[c];
      if (round == 0) {
        e = launch_contract(g, k, stream);                 // warm-up
        float one = 0;
        if (e == hipSuccess) time_launches(k, 1, one);
        if (e == hipSuccess)
          burst[c] = one > 0 ? std::max(1, std::min(256, int(6.0f / one))) : 256;
      }
      float ms = 0;
      if (e == hipSuccess) time_launches(k, burst[c], ms);
      if (e == hipSuccess) per[c] = std::min(per[c], ms / float(burst[c]));
      if (e != hipSuccess) failed = k;
    }
  }
  if (e != hipSuccess) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(scratch);
    return set_error(T1M_ERROR_DEVICE, "autotuning %s failed: %s",
                     kernel_info(failed >= 0 ? failed : int(pl.kernel)).name, hipGetErrorString(e));
  }
  int best = int(pl.kernel);
  float best_ms = 1e30f;
  for (size_t c = 0; c < todo.size(); ++c)
    if (per[c] < best_ms) { best_ms = per[c]; best = todo[c]; }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(scratch);
  pl.kernel = best;
  pl.tuned = true;
  std::lock_guard<std::mutex> lock(T.mu);
  T.best[key] = best;
  append_tune_cache(T, key, best);
  return T1M_SUCCESS;
}

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------

namespace {

struct PlanCache {
  std::mutex mu;
  std::list<std::string> lru;   // front = most recent
  std::unordered_map<std::string, std::pair<std::shared_ptr<Plan>, std::list<std::string>::iterator>> map;
  static constexpr size_t kMax = 256;
};

PlanCache& cache() {
  static PlanCache* c = new PlanCache();   // never destroyed (no HIP at exit)
  return *c;
}

// Plans that a captured graph held until it was destroyed.  The graph's
// user-object destructor may run on a runtime thread where HIP calls are not
// allowed, so it only parks its reference here; the next plan creation or
// cache clear drops the parked references (freeing the plans' device memory
// when nothing else holds them).
std::mutex g_parked_mu;
std::vector<std::shared_ptr<Plan>>* g_parked = new std::vector<std::shared_ptr<Plan>>();

void park_plan_ref(void* p) {
  auto* ref = static_cast<std::shared_ptr<Plan>*>(p);
  {
    std::lock_guard<std::mutex> lock(g_parked_mu);
    g_parked->push_back(std::move(*ref));
  }
  delete ref;
}

void drop_parked_plans() {
  std::vector<std::shared_ptr<Plan>> dead;
  {
    std::lock_guard<std::mutex> lock(g_parked_mu);
    dead.swap(*g_parked);
  }
}   // `dead` releases its plans here, outside the lock

std::string plan_key(const Problem& P, int precision, int device) {
  std::string k;
  char buf[64];
  // Test / tuning overrides change the plan, so they are part of its key.
  const char* fk = env_override("T1M_FORCE_KERNEL");
  const char* f64 = env_override("T1M_FORCE_IDX64");
  const char* fp = env_override("T1M_PAIRS");
  snprintf(buf, sizeof(buf), "%d|%d|%d|%.8s|%.8s|%.8s|", device, P.dtype, precision,
           fk ? fk : "", f64 ? f64 : "", fp ? fp : "");
  k += buf;
  k += P.la; k += ','; k += P.lb; k += ','; k += P.lc; k += '|';
  auto add = [&](const std::vector<int64_t>& v) {
    for (int64_t x : v) { snprintf(buf, sizeof(buf), "%lld,", (long long)x); k += buf; }
    k += '|';
  };
  add(P.da); add(P.sa); add(P.db); add(P.sb); add(P.dc); add(P.sc);
  return k;
}

int64_t max_abs_offset(const std::vector<Mode>& modes, int which) {
  int64_t s = 0;
  for (const Mode& m : modes) {
    const int64_t st = which == 0 ? m.sa : which == 1 ? m.sb : m.sc;
    s += (m.dim > 0 ? m.dim - 1 : 0) * (st < 0 ? -st : st);
  }
  return s;
}

// Smallest element offset of a bundle's scatter vector (<= 0).
int64_t min_offset(const std::vector<Mode>& modes, int which) {
  int64_t s = 0;
  for (const Mode& m : modes) {
    const int64_t st = which == 0 ? m.sa : which == 1 ? m.sb : m.sc;
    if (st < 0) s += (m.dim > 0 ? m.dim - 1 : 0) * st;
  }
  return s;
}

int64_t min_stride(const std::vector<Mode>& modes, int which) {
  int64_t best = INT64_MAX;
  for (const Mode& m : modes) {
    if (m.dim <= 1) continue;
    int64_t st = which == 0 ? m.sa : which == 1 ? m.sb : m.sc;
    st = st < 0 ? -st : st;
    best = std::min(best, st);
  }
  return best;
}

void order_by(std::vector<Mode>& modes, int which) {
  std::stable_sort(modes.begin(), modes.end(), [which](const Mode& x, const Mode& y) {
    auto key = [which](const Mode& m) {
      if (m.dim <= 1) return INT64_MAX;
      int64_t st = which == 0 ? m.sa : which == 1 ? m.sb : m.sc;
      return st < 0 ? -st : st;
    };
    return key(x) < key(y);
  });
}

// Two tensors reading one bundle with different unit-stride modes (A and B
// over k when both are gathered along k; A and C over m, B and C over n
// when the operand is gathered along that bundle): a traversal led by
// either one leaves the other strided (every lane of a stage, or every row
// of the epilogue, in its own cache line).  Splitting each one's fastest
// mode into 4 x rest and leading with both 4-blocks gives runs of four
// consecutive elements in both.  The traversal order is free: all maps of
// a bundle are built from the same mode list.  `first` leads.
void block_order(std::vector<Mode>& v, int first, int second) {
  auto stride = [](const Mode& m, int which) {
    return which == 0 ? m.sa : which == 1 ? m.sb : m.sc;
  };
  auto fastest = [&](int which) {
    int best = -1;
    int64_t bs = INT64_MAX;
    for (size_t i = 0; i < v.size(); ++i) {
      if (v[i].dim <= 1) continue;
      int64_t st = stride(v[i], which);
      st = st < 0 ? -st : st;
      if (st < bs) { bs = st; best = int(i); }
    }
    return best;
  };
  std::vector<Mode>& k = v;
  const int ia = fastest(first), ib = fastest(second);
  if (ia < 0 || ib < 0 || ia == ib) return;
  auto small = [](int64_t st) { return st >= -2 && st <= 2; };   // 4 elements within a line
  if (!small(stride(k[ia], first)) || !small(stride(k[ib], second))) return;
  if (k[ia].dim % 4 != 0 || k[ib].dim % 4 != 0) return;
  if (k.size() + 2 > size_t(T1M_MAX_RANK)) return;
  auto split = [](const Mode& m, Mode& lo, Mode& hi) {
    lo = m;
    lo.dim = 4;
    hi = m;
    hi.dim = m.dim / 4;
    hi.sa = 4 * m.sa; hi.sb = 4 * m.sb; hi.sc = 4 * m.sc;
  };
  Mode alo, ahi, blo, bhi;
  split(k[ia], alo, ahi);
  split(k[ib], blo, bhi);
  std::vector<Mode> out = {alo, blo, ahi, bhi};
  for (size_t i = 0; i < k.size(); ++i)
    if (int(i) != ia && int(i) != ib) out.push_back(k[i]);
  k.swap(out);
}

void fill_map(MapDesc& d, const std::vector<Mode>& modes, int which, bool out32,
              const void* out) {
  d.nmodes = int32_t(modes.size());
  d.out32 = out32 ? 1 : 0;
  d.len = 1;
  for (size_t i = 0; i < modes.size(); ++i) {
    d.dims[i] = modes[i].dim;
    d.strides[i] = which == 0 ? modes[i].sa : which == 1 ? modes[i].sb : modes[i].sc;
    d.len *= modes[i].dim;
  }
  d.out = const_cast<void*>(out);
}

}  // namespace

// C += A.B as C^T += B^T.A^T: operands (data, maps, conjugation, batch
// offsets, gather directions) exchanged, C's row and column maps swapped.
void transpose_problem(GemmProblem& g) {
  std::swap(g.A, g.B);
  const void* rsA = g.rsA;
  const void* csA = g.csA;
  g.rsA = g.csB; g.csA = g.rsB;          // rows of B^T are B's columns
  g.rsB = csA; g.csB = rsA;              // columns of A^T are A's rows
  std::swap(g.rsC, g.csC);
  std::swap(g.M, g.N);
  std::swap(g.conj_a, g.conj_b);
  const int32_t amf = g.a_mfast;
  g.a_mfast = g.b_kfast ? 0 : 1;         // B gathered along n <=> B^T along its rows
  g.b_kfast = amf ? 0 : 1;               // A gathered along k <=> A^T along k
  std::swap(g.boffA, g.boffB);
  std::swap(g.a_pairs, g.b_pairs);       // a property of each tensor's gather map
  const int64_t mra = g.min_rsA, mca = g.min_csA;
  g.min_rsA = g.min_csB; g.min_csA = g.min_rsB;
  g.min_rsB = mca; g.min_csB = mra;
  std::swap(g.min_rsC, g.min_csC);
}

// True unless the layout provably gives every index tuple its own element:
// with the modes of extent > 1 sorted by |stride|, each stride must pass the
// reach of all smaller ones (a zero stride, or strides that interleave,
// fail).  Outputs that may alias are then added atomically, so every
// contribution lands as in a serial loop.
const char* env_override(const char* name) {
  const char* v = getenv(name);
  return (v && *v) ? v : nullptr;
}

bool may_alias(const std::vector<int64_t>& dims, const std::vector<int64_t>& strides) {
  std::vector<std::pair<int64_t, int64_t>> m;   // (|stride|, extent)
  for (size_t i = 0; i < dims.size(); ++i)
    if (dims[i] > 1) m.emplace_back(strides[i] < 0 ? -strides[i] : strides[i], dims[i]);
  std::sort(m.begin(), m.end());
  int64_t reach = 0;                              // largest offset of the modes so far
  for (const auto& [s, e] : m) {
    if (s <= reach) return true;
    reach += (e - 1) * s;
  }
  return false;
}

// Host half of a plan: orientation, kernel choice, gather directions, the
// traversal order of each bundle and the offset width.  P receives the
// problem with its bundles in traversal order.
void plan_layout(const Problem& P0, int precision, Plan& pl, Problem& P) {
  P = P0;
  pl.M = P.M; pl.N = P.N; pl.K = P.K; pl.L = P.L;
  pl.dtype = P.dtype;
  pl.precision = precision;
  // Transposed plans, C^T = B^T . A^T (products commute): every kernel sees
  // M and N swapped, A and B (maps, data, conjugation, gather directions)
  // exchanged.  (1) A complex contraction with at most 16 free rows of A
  // and a wider C, so the thin 128 x 16 tiles cover it; (2) C whose unit
  // stride lies on a free mode of B (n-major), so the epilogue's 16-lane
  // runs (consecutive rows) become contiguous -- unless C is thin in n.
  const bool cplx_native = (P.dtype == T1M_COMPLEX64 || P.dtype == T1M_COMPLEX128) &&
                           precision == T1M_PRECISION_NATIVE;
  const bool thin_m = cplx_native && P.M <= 16 && P.N > 16;
  const bool n_major = precision != T1M_PRECISION_SIMPLE && min_stride(P.n, 2) < min_stride(P.m, 2) &&
                       !(cplx_native && P.N <= 16) && P.M > 16;
  pl.swap_ab = thin_m || n_major;
  {
    Problem Q = P;
    if (pl.swap_ab) std::swap(Q.M, Q.N);
    pl.kernel = choose_kernel(Q, precision);
  }
  pl.c_lo = pl.c_hi = 0;
  for (size_t i = 0; i < P.dc.size(); ++i) {
    const int64_t span = (P.dc[i] > 0 ? P.dc[i] - 1 : 0) * P.sc[i];
    if (span < 0) pl.c_lo += span; else pl.c_hi += span;
  }
  pl.c_overlap = may_alias(P.dc, P.sc);

  // Staging directions and traversal order of each bundle.
  pl.a_mfast = min_stride(P.m, 0) <= min_stride(P.k, 0) ? 1 : 0;
  pl.b_kfast = min_stride(P.k, 1) <= min_stride(P.n, 1) ? 1 : 0;
  order_by(P.m, pl.a_mfast ? 0 : 2);
  order_by(P.k, pl.b_kfast ? 1 : (pl.a_mfast ? 1 : 0));
  order_by(P.n, pl.b_kfast ? 2 : 1);
  if (!pl.a_mfast && pl.b_kfast) block_order(P.k, 0, 1);
  // C leads on m and n: its rows are read and written (measured against
  // the operand leading: C first was faster on every contraction where the
  // two differ, up to 1.9x on the reference suite's thin ones).
  if (pl.a_mfast) block_order(P.m, 2, 0);
  if (!pl.b_kfast) block_order(P.n, 2, 1);

  const int64_t lim = (int64_t(1) << 31) - 1;
  const int64_t offA = max_abs_offset(P.m, 0) + max_abs_offset(P.k, 0);
  const int64_t offB = max_abs_offset(P.k, 1) + max_abs_offset(P.n, 1);
  const int64_t offC = max_abs_offset(P.m, 2) + max_abs_offset(P.n, 2);
  pl.idx32 = (offA <= lim && offB <= lim && offC <= lim) ? 1 : 0;
  if (const char* f = env_override("T1M_FORCE_IDX64"))
    if (atoi(f) != 0) pl.idx32 = 0;
  pl.min_rsA = min_offset(P.m, 0);
  pl.min_csA = min_offset(P.k, 0);
  pl.min_rsB = min_offset(P.k, 1);
  pl.min_csB = min_offset(P.n, 1);
  pl.min_rsC = min_offset(P.m, 2);
  pl.min_csC = min_offset(P.n, 2);
  // A map entry minus its map's minimum is at most the map's span, so a
  // row + column byte offset is at most (offX + 1) elements: below 2^32 bytes.
  const int64_t esz = dtype_size(P.dtype);
  const int64_t lim32 = int64_t(1) << 32;
  pl.span32 = ((offA + 1) * esz < lim32 && (offB + 1) * esz < lim32 &&
               (offC + 1) * esz < lim32) ? 1 : 0;
  pl.modes_m = P.m; pl.modes_n = P.n; pl.modes_k = P.k;
  // Two-element gathers (complex<float>): the leading mode (extent > 1) of
  // the operand's gather direction has unit stride and even extent, so
  // map[i + 1] = map[i] + 1 for even i and the direction's extent is even.
  auto unit_pairs = [](const std::vector<Mode>& modes, int which) {
    for (const Mode& md : modes) {
      if (md.dim <= 1) continue;
      const int64_t st = which == 0 ? md.sa : md.sb;
      return st == 1 && md.dim % 2 == 0;
    }
    return false;
  };
  const char* pe = env_override("T1M_PAIRS");   // tests: 0 keeps one-element gathers
  const bool c64 = P.dtype == T1M_COMPLEX64 && !(pe && atoi(pe) == 0);
  pl.a_pairs = c64 && unit_pairs(pl.a_mfast ? P.m : P.k, 0) ? 1 : 0;
  pl.b_pairs = c64 && unit_pairs(pl.b_kfast ? P.k : P.n, 1) ? 1 : 0;
}

// Device bytes of a plan's maps: rsA, rsC (M entries), csA, rsB (K), csB,
// csC (N), then the [3][L] batch byte offsets when L > 1, each of the seven
// parts 256-B aligned.
size_t plan_device_bytes(const Plan& pl) {
  auto round = [](size_t x) { return (x + 255) / 256 * 256; };
  const size_t isz = pl.idx32 ? 4 : 8;
  size_t total = 2 * (round(size_t(pl.M) * isz) + round(size_t(pl.K) * isz) +
                      round(size_t(pl.N) * isz));
  if (pl.L > 1) total += round(3 * size_t(pl.L) * 8);
  return total;
}

// Points the plan's maps into `base` (plan_device_bytes bytes of device
// memory) and queues their construction on `stream`: one launch for the six
// scatter vectors, one for the batch offsets (scatter vectors of the batch
// modes with byte strides).  Nothing waits.
hipError_t plan_build_maps(Plan& pl, const Problem& P, char* base, hipStream_t stream) {
  auto round = [](size_t x) { return (x + 255) / 256 * 256; };
  const size_t isz = pl.idx32 ? 4 : 8;
  const size_t szM = round(size_t(P.M) * isz), szK = round(size_t(P.K) * isz),
               szN = round(size_t(P.N) * isz);
  pl.rsA = base;
  pl.rsC = base + szM;
  pl.csA = base + 2 * szM;
  pl.rsB = base + 2 * szM + szK;
  pl.csB = base + 2 * szM + 2 * szK;
  pl.csC = base + 2 * szM + 2 * szK + szN;
  MapBatch mb;
  mb.count = 6;
  fill_map(mb.maps[0], P.m, 0, pl.idx32, pl.rsA);
  fill_map(mb.maps[1], P.m, 2, pl.idx32, pl.rsC);
  fill_map(mb.maps[2], P.k, 0, pl.idx32, pl.csA);
  fill_map(mb.maps[3], P.k, 1, pl.idx32, pl.rsB);
  fill_map(mb.maps[4], P.n, 1, pl.idx32, pl.csB);
  fill_map(mb.maps[5], P.n, 2, pl.idx32, pl.csC);
  hipError_t e = launch_scatter_maps(mb, stream);
  pl.boff = nullptr;
  if (e != hipSuccess || P.L <= 1) return e;
  // Batch problem z reads A, B, C at sum_j digit_j(z) * stride_j * esize
  // bytes (mixed radix over the batch modes, first mode fastest).
  char* boff = base + 2 * (szM + szK + szN);
  pl.boff = reinterpret_cast<int64_t*>(boff);
  const int64_t es = dtype_size(P.dtype);
  std::vector<Mode> bytes = P.batch;
  for (Mode& md : bytes) { md.sa *= es; md.sb *= es; md.sc *= es; }
  MapBatch ob;
  ob.count = 3;
  for (int w = 0; w < 3; ++w) fill_map(ob.maps[w], bytes, w, false, boff + w * size_t(P.L) * 8);
  return launch_scatter_maps(ob, stream);
}

int get_plan(const Problem& P0, int precision, hipStream_t stream,
             std::shared_ptr<Plan>& out) {
  int device = 0;
  HIP_TRY(hipGetDevice(&device));
  const std::string key = plan_key(P0, precision, device);
  PlanCache& C = cache();
  {
    std::lock_guard<std::mutex> lock(C.mu);
    auto it = C.map.find(key);
    if (it != C.map.end()) {
      C.lru.splice(C.lru.begin(), C.lru, it->second.second);
      out = it->second.first;
      return T1M_SUCCESS;
    }
  }
  drop_parked_plans();
  auto plan = std::make_shared<Plan>();
  Plan& pl = *plan;
  Problem P;
  plan_layout(P0, precision, pl, P);

  // One allocation for the maps and the batch offsets.
  const size_t total = plan_device_bytes(pl);
  char* base = nullptr;
  if (hipMalloc(&base, total) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(T1M_ERROR_OUT_OF_MEMORY, "hipMalloc of %zu bytes for index maps failed", total);
  }
  pl.maps = base;
  hipError_t e = plan_build_maps(pl, P, base, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess)
    return set_error(T1M_ERROR_DEVICE, "index map construction failed: %s", hipGetErrorString(e));
  if (P.L > 1) {
    // The same offsets on the host (the persistent batch flattens instances
    // into one problem per batch index with them).
    const int64_t es = dtype_size(P.dtype);
    pl.hboff.assign(size_t(3 * P.L), 0);
    for (int64_t z = 0; z < P.L; ++z) {
      int64_t r = z, oa = 0, ob = 0, oc = 0;
      for (const Mode& md : P.batch) {
        const int64_t d = r % md.dim;
        r /= md.dim;
        oa += d * md.sa; ob += d * md.sb; oc += d * md.sc;
      }
      pl.hboff[size_t(z)] = oa * es;
      pl.hboff[size_t(P.L + z)] = ob * es;
      pl.hboff[size_t(2 * P.L + z)] = oc * es;
    }
  }
  apply_tuning_file(pl);          // a tuning result from an earlier run

  std::lock_guard<std::mutex> lock(C.mu);
  auto it = C.map.find(key);
  if (it != C.map.end()) {          // another thread won the race
    out = it->second.first;
    return T1M_SUCCESS;
  }
  C.lru.push_front(key);
  C.map.emplace(key, std::make_pair(plan, C.lru.begin()));
  while (C.map.size() > PlanCache::kMax) {
    C.map.erase(C.lru.back());
    C.lru.pop_back();
  }
  out = plan;
  return T1M_SUCCESS;
}

Plan::~Plan() {
  if (maps) (void)hipFree(maps);   // holds the batch offsets too
}

int64_t plan_cache_size() {
  PlanCache& C = cache();
  std::lock_guard<std::mutex> lock(C.mu);
  return int64_t(C.map.size());
}

namespace {
std::atomic<uint64_t> g_plan_generation{0};
}  // namespace

void plan_cache_clear() {
  {
    PlanCache& C = cache();
    std::lock_guard<std::mutex> lock(C.mu);
    C.map.clear();
    C.lru.clear();
    g_plan_generation.fetch_add(1);
  }
  {
    TuneCache& T = tune_cache();   // re-read the tuning file on next use
    std::lock_guard<std::mutex> lock(T.mu);
    T.loaded = false;
    T.best.clear();
    T.batch_modes.clear();
  }
  drop_parked_plans();
  staging_pool_clear();
}

// A call captured into a graph bakes the plan's map pointers (and batch
// offsets) into the graph's kernel nodes.  The graph then holds a reference
// to the plan (a user object released when the graph and its executable
// graphs are destroyed), so evicting or clearing the plan cache cannot free
// memory a later replay reads.
int retain_plan_for_capture(const std::shared_ptr<Plan>& plan, hipStream_t stream) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &st) != hipSuccess) {
    (void)hipGetLastError();   // e.g. the null stream beside a global capture:
    return T1M_SUCCESS;        // the launch itself reports that case
  }
  if (st != hipStreamCaptureStatusActive) return T1M_SUCCESS;
  hipGraph_t graph = nullptr;
  unsigned long long id = 0;
  HIP_TRY(hipStreamGetCaptureInfo_v2(stream, &st, &id, &graph, nullptr, nullptr));
  if (!graph) return set_error(T1M_ERROR_DEVICE, "capturing stream without a graph");
  auto* ref = new std::shared_ptr<Plan>(plan);
  hipUserObject_t obj = nullptr;
  hipError_t e = hipUserObjectCreate(&obj, ref, park_plan_ref, 1,
                                     hipUserObjectNoDestructorSync);
  if (e != hipSuccess) {
    delete ref;
    return set_error(T1M_ERROR_DEVICE, "hipUserObjectCreate: %s", hipGetErrorString(e));
  }
  e = hipGraphRetainUserObject(graph, obj, 1, hipGraphUserObjectMove);
  if (e != hipSuccess) {
    (void)hipUserObjectRelease(obj, 1);   // runs park_plan_ref
    return set_error(T1M_ERROR_DEVICE, "hipGraphRetainUserObject: %s", hipGetErrorString(e));
  }
  return T1M_SUCCESS;
}

// ---------------------------------------------------------------------------
// Host pointers: stage the strided extent through device memory.
// ---------------------------------------------------------------------------

bool is_device_pointer(const void* ptr) {
  if (!ptr) return false;
  hipPointerAttribute_t attr;
  const hipError_t e = hipPointerGetAttributes(&attr, ptr);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

namespace {

// Staging buffers kept between calls, per device: a call with host tensors
// takes the smallest free buffer that fits (else allocates one) and returns
// it afterwards, so repeated host-pointer calls pay no hipMalloc / hipFree
// (hipFree also waits for the whole device).  At most kPoolBytes stay
// cached per device; larger returns are freed.
constexpr size_t kPoolBytes = size_t(4) << 30;
struct StagingPool {
  std::mutex mu;
  std::vector<std::vector<std::pair<void*, size_t>>> free;   // [device] -> (buffer, size)
  std::vector<size_t> held;                                  // [device] cached bytes
};
StagingPool& staging_pool() {
  static StagingPool* p = new StagingPool();
  return *p;
}

void* pool_take(int dev, size_t bytes, size_t& cap) {
  StagingPool& P = staging_pool();
  {
    std::lock_guard<std::mutex> lock(P.mu);
    if (size_t(dev) < P.free.size()) {
      auto& v = P.free[size_t(dev)];
      size_t best = v.size();
      for (size_t i = 0; i < v.size(); ++i)
        if (v[i].second >= bytes && (best == v.size() || v[i].second < v[best].second)) best = i;
      if (best < v.size()) {
        void* b = v[best].first;
        cap = v[best].second;
        P.held[size_t(dev)] -= cap;
        v.erase(v.begin() + std::ptrdiff_t(best));
        return b;
      }
    }
  }
  void* b = nullptr;
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipGetLastError();
    staging_pool_clear();                  // cached buffers may be what is missing
    if (hipMalloc(&b, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  cap = bytes;
  return b;
}

void pool_give(int dev, void* b, size_t cap) {
  StagingPool& P = staging_pool();
  {
    std::lock_guard<std::mutex> lock(P.mu);
    if (size_t(dev) >= P.free.size()) {
      P.free.resize(size_t(dev) + 1);
      P.held.resize(size_t(dev) + 1, 0);
    }
    if (P.held[size_t(dev)] + cap <= kPoolBytes) {
      P.free[size_t(dev)].emplace_back(b, cap);
      P.held[size_t(dev)] += cap;
      return;
    }
  }
  (void)hipFree(b);
}

}  // namespace

void staging_pool_clear() {
  std::vector<std::vector<std::pair<void*, size_t>>> drop;
  {
    StagingPool& P = staging_pool();
    std::lock_guard<std::mutex> lock(P.mu);
    drop.swap(P.free);
    P.held.assign(P.held.size(), 0);
  }
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (size_t d = 0; d < drop.size(); ++d) {
    if (drop[d].empty()) continue;
    (void)hipSetDevice(int(d));
    for (auto& [b, n] : drop[d]) (void)hipFree(b);
  }
  (void)hipSetDevice(cur);
}

int Staged::stage(const void* host, const std::vector<int64_t>& dims,
                  const std::vector<int64_t>& strides, int esize, bool copy_in,
                  hipStream_t stream_) {
  int64_t lo = 0, hi = 0;
  bool empty = false;
  for (size_t i = 0; i < dims.size(); ++i) {
    if (dims[i] == 0) empty = true;
    const int64_t span = (dims[i] - 1) * strides[i];
    if (span < 0) lo += span; else hi += span;
  }
  host_base = static_cast<const char*>(host) + lo * esize;
  bytes = empty ? 0 : size_t(hi - lo + 1) * esize;
  if (bytes == 0) { device = const_cast<void*>(host); return T1M_SUCCESS; }
  HIP_TRY(hipGetDevice(&dev));
  stream = stream_;
  buf = pool_take(dev, bytes, cap);
  if (!buf)
    return set_error(T1M_ERROR_OUT_OF_MEMORY, "hipMalloc of %zu bytes for staging failed", bytes);
  device = static_cast<char*>(buf) - lo * esize;
  if (copy_in) HIP_TRY(hipMemcpyAsync(buf, host_base, bytes, hipMemcpyHostToDevice, stream));
  return T1M_SUCCESS;
}

int Staged::copy_back(void* host, hipStream_t stream) {
  (void)host;
  if (!buf || bytes == 0) return T1M_SUCCESS;
  HIP_TRY(hipMemcpyAsync(const_cast<char*>(host_base), buf, bytes, hipMemcpyDeviceToHost, stream));
  return T1M_SUCCESS;
}

Staged::~Staged() {
  if (!buf) return;
  if (!synced) (void)hipStreamSynchronize(stream);   // error paths: nothing may still use it
  pool_give(dev, buf, cap);
}

// ---------------------------------------------------------------------------
// Contraction driver
// ---------------------------------------------------------------------------

namespace {

// Per-thread memo of recent calls: the raw call signature (dtypes, ranks,
// extents, strides, labels, precision, device, test overrides) -> validated
// problem and plan, so a repeated shape skips label parsing and the plan
// cache's key building.  Entries die with a plan cache clear (generation).
struct RecentCall {
  std::vector<int64_t> sig;
  Problem P;
  std::shared_ptr<Plan> plan;
};
constexpr size_t kRecentCalls = 8;
thread_local std::vector<RecentCall> tl_recent;
thread_local size_t tl_recent_next = 0;

void call_signature(const t1m_tensor* const* t, const char* const* l, int precision,
                    std::vector<int64_t>& sig) {
  sig.clear();
  int dev = 0;
  (void)hipGetDevice(&dev);
  sig.push_back(int64_t(g_plan_generation.load()));
  sig.push_back(dev);
  sig.push_back(precision);
  for (const char* env : {"T1M_FORCE_KERNEL", "T1M_FORCE_IDX64", "T1M_PAIRS"}) {
    const char* v = env_override(env);
    int64_t h = v ? 1469598103934665603LL : 0;
    if (v)
      for (const char* q = v; *q; ++q) h = (h ^ uint8_t(*q)) * 1099511628211LL;
    sig.push_back(h);
  }
  for (int i = 0; i < 3; ++i) {
    sig.push_back(t[i]->dtype);
    sig.push_back(t[i]->rank);
    for (int d = 0; d < t[i]->rank; ++d) sig.push_back(t[i]->dims[d]);
    sig.push_back(t[i]->strides ? 1 : 0);
    if (t[i]->strides)
      for (int d = 0; d < t[i]->rank; ++d) sig.push_back(t[i]->strides[d]);
    for (const char* q = l[i]; *q; ++q) sig.push_back(uint8_t(*q));
    sig.push_back(-1);
  }
}

}  // namespace

GemmProblem gemm_for(const Plan& pl, const void* A, const void* B, void* C,
                     const t1m_options& opt) {
  GemmProblem g;
  g.A = A; g.B = B; g.C = C;
  g.rsA = pl.rsA; g.csA = pl.csA; g.rsB = pl.rsB; g.csB = pl.csB;
  g.rsC = pl.rsC; g.csC = pl.csC;
  g.M = pl.M; g.N = pl.N; g.K = pl.K;
  g.alpha_re = opt.alpha_re; g.alpha_im = opt.alpha_im;
  const bool cplx = pl.dtype == T1M_COMPLEX64 || pl.dtype == T1M_COMPLEX128;
  g.conj_a = cplx && opt.conj_a; g.conj_b = cplx && opt.conj_b;
  if (!cplx) g.alpha_im = 0.0;
  g.a_mfast = pl.a_mfast; g.b_kfast = pl.b_kfast;
  g.a_pairs = pl.a_pairs; g.b_pairs = pl.b_pairs;
  g.idx32 = pl.idx32;
  g.min_rsA = pl.min_rsA; g.min_csA = pl.min_csA;
  g.min_rsB = pl.min_rsB; g.min_csB = pl.min_csB;
  g.min_rsC = pl.min_rsC; g.min_csC = pl.min_csC;
  g.span32 = pl.idx32 && pl.span32 ? 1 : 0;
  g.dtype = pl.dtype;
  g.split_k = opt.deterministic ? 1 : 0;
  g.cus = cu_count();
  g.c_overlap = pl.c_overlap ? 1 : 0;
  g.L = pl.L;
  if (pl.L > 1) {
    g.boffA = pl.boff;
    g.boffB = pl.boff + pl.L;
    g.boffC = pl.boff + 2 * pl.L;
  }
  if (pl.swap_ab) transpose_problem(g);
  return g;
}

void apply_tuning_file(Plan& pl) {
  if (env_override("T1M_FORCE_KERNEL")) return;
  TuneCache& T = tune_cache();
  std::lock_guard<std::mutex> lock(T.mu);
  load_tune_cache(T);
  auto it = T.best.find(tune_key(pl));
  if (it != T.best.end() && tuned_kernel_ok(pl, it->second)) {
    pl.kernel = it->second;
    pl.tuned = true;
  }
}

int contract_impl(const t1m_tensor* a, const char* la, const t1m_tensor* b,
                  const char* lb, const t1m_tensor* c, const char* lc,
                  const t1m_options& opt) {
  if (opt.precision < 0 || opt.precision > T1M_PRECISION_SIMPLE)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "unknown precision %d", opt.precision);
  hipStream_t stream = static_cast<hipStream_t>(opt.stream);
  int rc;
  // Fast path: the same call shape as a recent call on this thread.
  const bool well_formed = a && b && c && la && lb && lc && a->rank >= 0 && b->rank >= 0 &&
                           c->rank >= 0 && a->rank <= T1M_MAX_RANK && b->rank <= T1M_MAX_RANK &&
                           c->rank <= T1M_MAX_RANK && (a->rank == 0 || a->dims) &&
                           (b->rank == 0 || b->dims) && (c->rank == 0 || c->dims);
  thread_local std::vector<int64_t> sig;
  const Problem* Pp = nullptr;
  std::shared_ptr<Plan> plan;
  Problem Pnew;
  if (well_formed) {
    const t1m_tensor* ts[3] = {a, b, c};
    const char* ls[3] = {la, lb, lc};
    call_signature(ts, ls, opt.precision, sig);
    for (RecentCall& r : tl_recent)
      if (r.sig == sig) {
        Pp = &r.P;
        plan = r.plan;
        break;
      }
  }
  if (!Pp) {
    if ((rc = parse_problem(a, la, b, lb, c, lc, Pnew))) return rc;
    if (opt.precision == T1M_PRECISION_BF16X3 && Pnew.dtype != T1M_COMPLEX64)
      return set_error(T1M_ERROR_UNSUPPORTED, "BF16x3 needs complex<float> data");
    if (Pnew.M == 0 || Pnew.N == 0 || Pnew.K == 0 || Pnew.L == 0) return T1M_SUCCESS;   // C += 0
    if ((rc = get_plan(Pnew, opt.precision, stream, plan))) return rc;
    Pp = &Pnew;
    if (well_formed) {
      RecentCall r{sig, Pnew, plan};
      if (tl_recent.size() < kRecentCalls) {
        tl_recent.push_back(std::move(r));
      } else {
        tl_recent[tl_recent_next] = std::move(r);
        tl_recent_next = (tl_recent_next + 1) % kRecentCalls;
      }
    }
  }
  const Problem& P = *Pp;
  if (!a->data || !b->data || !c->data)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "null data pointer");

  const int es = dtype_size(P.dtype);
  Staged sa, sb, sc;
  const void* dA = a->data;
  const void* dB = b->data;
  void* dC = c->data;
  const bool hostA = !is_device_pointer(a->data);
  const bool hostB = !is_device_pointer(b->data);
  const bool hostC = !is_device_pointer(c->data);
  if (hostA) { if ((rc = sa.stage(a->data, P.da, P.sa, es, true, stream))) return rc; dA = sa.device; }
  if (hostB) { if ((rc = sb.stage(b->data, P.db, P.sb, es, true, stream))) return rc; dB = sb.device; }
  if (hostC) { if ((rc = sc.stage(c->data, P.dc, P.sc, es, true, stream))) return rc; dC = sc.device; }

  GemmProblem g = gemm_for(*plan, dA, dB, dC, opt);
  if (opt.autotune && !plan->tuned.load() && (rc = autotune_plan(*plan, g, stream))) return rc;
  if ((rc = retain_plan_for_capture(plan, stream))) return rc;
  hipError_t e = launch_contract(g, plan->kernel, stream);
  if (e != hipSuccess)
    return set_error(T1M_ERROR_DEVICE, "kernel %s launch failed: %s",
                     kernel_info(plan->kernel.load()).name, hipGetErrorString(e));
  if (hostC && (rc = sc.copy_back(c->data, stream))) return rc;
  if (opt.synchronize || hostA || hostB || hostC) {
    HIP_TRY(hipStreamSynchronize(stream));
    sa.synced = sb.synced = sc.synced = true;
  }
  return T1M_SUCCESS;
}

}  // namespace t1m_hip

// ===========================================================================
// extern "C"
// ===========================================================================

using namespace t1m_hip;

extern "C" {

void t1m_options_init(t1m_options* opt) {
  if (!opt) return;
  opt->alpha_re = 1.0;
  opt->alpha_im = 0.0;
  opt->precision = T1M_PRECISION_NATIVE;
  opt->conj_a = 0;
  opt->conj_b = 0;
  opt->synchronize = 1;
  opt->stream = nullptr;
  opt->autotune = 0;
  opt->deterministic = 0;
}

const char* t1m_last_error(void) { return last_error(); }

const char* t1m_version(void) { return "t1m-mi355x 0.1.0 (gfx950)"; }

int64_t t1m_plan_cache_size(void) { return plan_cache_size(); }

const char* t1m_autotune_cache_path(void) {
  static thread_local std::string p;
  p = tune_path();
  return p.c_str();
}

const char* t1m_plan_kernel(const t1m_tensor* a, const char* la, const t1m_tensor* b,
                            const char* lb, const t1m_tensor* c, const char* lc,
                            const t1m_options* opt) {
  clear_error();
  t1m_options o;
  if (opt) o = *opt; else t1m_options_init(&o);
  Problem P;
  if (parse_problem(a, la, b, lb, c, lc, P)) return nullptr;
  if (P.M == 0 || P.N == 0 || P.K == 0 || P.L == 0) return "none";
  std::shared_ptr<Plan> plan;
  if (get_plan(P, o.precision, static_cast<hipStream_t>(o.stream), plan)) return nullptr;
  return kernel_info(plan->kernel.load()).name;
}

void t1m_plan_cache_clear(void) { plan_cache_clear(); }

int t1m_contract(const t1m_tensor* a, const char* la, const t1m_tensor* b,
                 const char* lb, const t1m_tensor* c, const char* lc,
                 const t1m_options* opt) {
  clear_error();
  t1m_options o;
  if (opt) o = *opt; else t1m_options_init(&o);
  return contract_impl(a, la, b, lb, c, lc, o);
}

int t1m_contract_vjp(const t1m_tensor* a, const char* la, const t1m_tensor* b,
                     const char* lb, const t1m_tensor* grad_c, const char* lc,
                     const t1m_tensor* grad_a, const t1m_tensor* grad_b,
                     const t1m_options* opt) {
  clear_error();
  t1m_options o;
  if (opt) o = *opt; else t1m_options_init(&o);
  // Validate the forward problem once so errors name A, B and C.
  Problem P;
  int rc = parse_problem(a, la, b, lb, grad_c, lc, P);
  if (rc) return rc;
  // The forward is C += alpha * A' . B' with A' = conj(A) when conj_a is set
  // (B' likewise); gradients are taken with respect to A' and B':
  //   grad_a += conj(alpha) * grad_c . conj(B'),  conj(B') = B if conj_b;
  //   grad_b += conj(alpha) * conj(A') . grad_c.
  const bool cplx = P.dtype == T1M_COMPLEX64 || P.dtype == T1M_COMPLEX128;
  t1m_options g = o;
  g.alpha_im = -o.alpha_im;
  if (grad_a) {
    g.conj_a = 0;
    g.conj_b = cplx && !o.conj_b;
    if ((rc = contract_impl(grad_c, lc, b, lb, grad_a, la, g))) return rc;
  }
  if (grad_b) {
    g.conj_a = cplx && !o.conj_a;
    g.conj_b = 0;
    if ((rc = contract_impl(a, la, grad_c, lc, grad_b, lb, g))) return rc;
  }
  return T1M_SUCCESS;
}

}  // extern "C"
