// extern "C" exports of the index maps and the reference panel formats, so
// they can be compared bit for bit with the reference semantics:
//   t1m_index_bundles  get_index_bundles   include/t1m/t1m.h:90-136 (host)
//   t1m_index_maps     get_scatter + get_block_scatter, t1m.h:141-205 (device)
//   t1m_pack_block     pack_block / pack_cell, t1m.h:279-404 (device)
//   t1m_unpack         unpack, t1m.h:406-430 (device)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <iterator>
#include <string>

#include "internal.h"
#include "runtime.h"

using namespace t1m_hip;

namespace {

// A view's counts and arrays: non-negative counts, and every array the
// kernels index present when the view is not empty.  pack_block reads the
// block strides of every block its rows / columns start (rbs[r0 / br] for
// r0 < nrs), so with `whole_blocks` the block-scatter vectors must cover
// ceil(nrs / br) and ceil(ncs / bc) entries; unpack reads entry 0 only.
bool view_ok(const t1m_view& v, bool whole_blocks) {
  if (v.nrs < 0 || v.ncs < 0 || v.nrbs < 0 || v.ncbs < 0) return false;
  if (v.nrs == 0 || v.ncs == 0) return true;
  if (!(v.rs && v.cs && v.rbs && v.cbs && v.nrbs > 0 && v.ncbs > 0)) return false;
  if (!whole_blocks) return true;
  if (v.br <= 0 || v.bc <= 0) return false;
  return v.nrbs >= (v.nrs + v.br - 1) / v.br && v.ncbs >= (v.ncs + v.bc - 1) / v.bc;
}

}  // namespace

extern "C" {

int t1m_index_bundles(const char* la, const char* lb, const char* lc,
                      int64_t* bundles, int64_t* counts) {
  clear_error();
  if (!la || !lb || !lc || !bundles || !counts)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "null argument");
  const std::string A(la), B(lb), Cs(lc);
  if (A.size() > T1M_MAX_RANK || B.size() > T1M_MAX_RANK || Cs.size() > T1M_MAX_RANK)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "more than %d labels", T1M_MAX_RANK);
  std::string a = A, b = B;
  std::sort(a.begin(), a.end());
  std::sort(b.begin(), b.end());
  std::string contracted, free_a, free_b;
  std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(contracted));
  std::set_difference(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(free_a));
  std::set_difference(b.begin(), b.end(), a.begin(), a.end(), std::back_inserter(free_b));
  auto find = [](const std::string& s, char ch) -> int64_t {
    const size_t p = s.find(ch);
    return p == std::string::npos ? int64_t(-1) : int64_t(p);
  };
  int64_t* AX = bundles + 0 * T1M_MAX_RANK;
  int64_t* AY = bundles + 1 * T1M_MAX_RANK;
  int64_t* BX = bundles + 2 * T1M_MAX_RANK;
  int64_t* BY = bundles + 3 * T1M_MAX_RANK;
  int64_t* CX = bundles + 4 * T1M_MAX_RANK;
  int64_t* CY = bundles + 5 * T1M_MAX_RANK;
  for (size_t i = 0; i < contracted.size(); ++i) {
    AY[i] = find(A, contracted[i]);
    BX[i] = find(B, contracted[i]);
  }
  for (size_t i = 0; i < free_a.size(); ++i) {
    AX[i] = find(A, free_a[i]);
    CX[i] = find(Cs, free_a[i]);
  }
  for (size_t i = 0; i < free_b.size(); ++i) {
    BY[i] = find(B, free_b[i]);
    CY[i] = find(Cs, free_b[i]);
  }
  counts[0] = int64_t(free_a.size());
  counts[1] = int64_t(contracted.size());
  counts[2] = int64_t(contracted.size());
  counts[3] = int64_t(free_b.size());
  counts[4] = int64_t(free_a.size());
  counts[5] = int64_t(free_b.size());
  return T1M_SUCCESS;
}

int t1m_index_maps(const t1m_tensor* t, const int64_t* modes, int32_t nmodes,
                   int64_t block, uint64_t* scat, uint64_t* block_scat,
                   void* stream) {
  clear_error();
  if (!t || (nmodes > 0 && !modes) || !scat)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "null argument");
  if (nmodes < 0 || nmodes > T1M_MAX_RANK)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "nmodes %d out of range", nmodes);
  if (block_scat && block <= 0)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "block size must be positive");
  if (t->rank < 0 || t->rank > T1M_MAX_RANK || (t->rank > 0 && !t->dims))
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "bad tensor descriptor");
  MapBatch mb;
  mb.count = 1;
  MapDesc& d = mb.maps[0];
  d.nmodes = nmodes;
  d.out32 = 0;
  d.len = 1;
  d.out = scat;
  for (int i = 0; i < nmodes; ++i) {
    const int64_t x = modes[i];
    if (x < 0 || x >= t->rank)
      return set_error(T1M_ERROR_INVALID_ARGUMENT, "mode %lld out of range", (long long)x);
    if (t->dims[x] < 0)
      return set_error(T1M_ERROR_INVALID_ARGUMENT, "negative extent %lld", (long long)t->dims[x]);
    d.dims[i] = t->dims[x];
    if (t->strides) {
      d.strides[i] = t->strides[x];
    } else {
      int64_t s = 1;   // column-major, t1m.h:42-50
      for (int64_t j = 0; j < x; ++j) s *= t->dims[j];
      d.strides[i] = s;
    }
    d.len *= d.dims[i];
  }
  const hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(launch_scatter_maps(mb, s));
  if (block_scat) HIP_TRY(launch_block_scatter(scat, d.len, block, block_scat, s));
  return T1M_SUCCESS;
}

int t1m_pack_block(int32_t dtype, int32_t label, const t1m_view* view,
                   int64_t length, const void* src, void* dest, void* stream) {
  clear_error();
  if (!view || !src || !dest)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "null argument");
  if (dtype_size(dtype) == 0)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
  if (label != 0 && label != 1)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "label must be 0 (A) or 1 (B)");
  const bool cplx = dtype == T1M_COMPLEX64 || dtype == T1M_COMPLEX128;
  if (view->br <= 0 || view->bc <= 0 || (cplx && (view->br < 2 || (label == 0 && view->bc < 2))))
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "block sizes too small for the panel format");
  if (!view_ok(*view, true) || length < 0)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "bad view counts or arrays");
  HIP_TRY(launch_pack_block(dtype, label, *view, length, src, dest,
                            static_cast<hipStream_t>(stream)));
  return T1M_SUCCESS;
}

int t1m_unpack(int32_t dtype, const t1m_view* view, const void* src,
               void* dest, void* stream) {
  clear_error();
  if (!view || !src || !dest)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "null argument");
  if (dtype_size(dtype) == 0)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
  if (!view_ok(*view, false) || view->br <= 0)
    return set_error(T1M_ERROR_INVALID_ARGUMENT, "bad view counts, arrays or block size");
  HIP_TRY(launch_unpack(dtype, *view, src, dest, static_cast<hipStream_t>(stream)));
  return T1M_SUCCESS;
}

}  // extern "C"
