// Device construction of t1m's index maps and the reference panel formats.
//
//  * k_scatter_maps: scatter vectors (get_scatter, include/t1m/t1m.h:141-174)
//    for up to six mode lists in one launch.  Entry e of a bundle's vector is
//    sum_j digit_j(e) * stride_j with e's mixed-radix digits over the bundle's
//    extents, first mode fastest -- the closed form of the reference's fold.
//    Arithmetic is modulo 2^64 like std::size_t.
//  * k_block_scatter: get_block_scatter (t1m.h:179-205), one lane per block.
//  * k_pack_block / k_unpack: pack_block / pack_cell (t1m.h:279-404) and
//    unpack (t1m.h:406-430) producing / consuming the reference's
//    micro-panel format, one lane per matrix element.
//
// The contraction kernels consume the scatter vectors directly (they gather
// from strided HBM into LDS); the reference panel format is exported so the
// packing index maps can be compared bit for bit.
#include "internal.h"

namespace t1m_hip {

namespace {

template <typename U>
__device__ inline uint64_t scatter_entry(const MapDesc& d, int64_t e) {
  U r = U(e);
  uint64_t off = 0;
  for (int j = 0; j < d.nmodes; ++j) {
    const U ext = U(d.dims[j]);
    const U q = r / ext;
    off += uint64_t(r - q * ext) * uint64_t(d.strides[j]);
    r = q;
  }
  return off;
}

__global__ void k_scatter_maps(MapBatch batch) {
  const MapDesc& d = batch.maps[blockIdx.y];
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  // A bundle of fewer than 2^32 entries has every extent below 2^32 too:
  // its digits come from 32-bit divisions (64-bit ones cost ~4x the
  // instructions and made the kernel issue-bound at a quarter of HBM speed).
  const bool narrow = uint64_t(d.len) <= 0xffffffffull;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < d.len;
       e += stride) {
    const uint64_t off = narrow ? scatter_entry<uint32_t>(d, e) : scatter_entry<uint64_t>(d, e);
    if (d.out32)
      static_cast<int32_t*>(d.out)[e] = int32_t(int64_t(off));
    else
      static_cast<uint64_t*>(d.out)[e] = off;
  }
}

// t1m.h:179-205: a size-1 block gives 0; otherwise the block's adjacent
// differences must all be equal, and that difference is the block stride,
// else 0.  Differences wrap modulo 2^64.
__global__ void k_block_scatter(const uint64_t* __restrict__ scat, int64_t len,
                                int64_t b, uint64_t* __restrict__ out) {
  const int64_t nblk = len / b + (len % b > 0);
  const int64_t blk = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (blk >= nblk) return;
  const int64_t i0 = blk * b;
  const int64_t bs = (b < len - i0) ? b : len - i0;
  if (bs == 1) { out[blk] = 0; return; }
  const uint64_t s = scat[i0 + 1] - scat[i0];
  bool uniform = true;
  for (int64_t t = 2; t < bs; ++t)
    if (scat[i0 + t] - scat[i0 + t - 1] != s) { uniform = false; break; }
  out[blk] = uniform ? s : 0;
}

// Source index of element (k, l) of the cell whose first row / column are
// view rows r0 / cols c0: the dense path of pack_cell (t1m.h:285-289) uses
// the cell's first block strides; otherwise rs[k] + cs[l].
__device__ inline uint64_t cell_src(const t1m_view& v, int64_t r0, int64_t c0,
                                    int64_t k, int64_t l) {
  const uint64_t rsc = v.rbs[r0 / v.br];
  const uint64_t csc = v.cbs[c0 / v.bc];
  if (rsc > 0 && csc > 0)
    return uint64_t(k) * rsc + uint64_t(l) * csc + (v.rs[r0] + v.cs[c0]);
  return v.rs[r0 + k] + v.cs[c0 + l];
}

template <typename R, bool kComplex, bool kLabelA>
__global__ void k_pack_block(t1m_view v, int64_t length,
                             const R* __restrict__ src, R* __restrict__ dest) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= v.nrs * v.ncs) return;
  const int64_t r = e % v.nrs, c = e / v.nrs;
  if constexpr (!kComplex) {
    // t1m.h:346-370: br x bc cells; A column-major (ld br), B row-major (ld bc)
    const int64_t r0 = r / v.br * v.br, c0 = c / v.bc * v.bc;
    const int64_t k = r - r0, l = c - c0;
    const R val = src[cell_src(v, r0, c0, k, l)];
    if constexpr (kLabelA)
      dest[c0 * v.br + r0 * length + k + l * v.br] = val;
    else
      dest[r0 * v.bc + c0 * length + l + k * v.bc] = val;
  } else if constexpr (kLabelA) {
    // t1m.h:375-389 (cells of br/2 x bc/2 complex), t1m.h:310-325
    const int64_t hbr = v.br / 2, hbc = v.bc / 2;
    const int64_t r0 = r / hbr * hbr, c0 = c / hbc * hbc;
    const int64_t k = r - r0, l = c - c0;
    const uint64_t s = cell_src(v, r0, c0, k, l);
    const R re = src[2 * s], im = src[2 * s + 1];
    R* d = dest + 2 * (c0 * v.br + r0 * length) + 2 * k + 2 * l * v.br;
    d[0] = re;
    d[1] = im;
    d[v.br] = -im;
    d[v.br + 1] = re;
  } else {
    // t1m.h:390-402 (cells of br/2 x bc complex), t1m.h:326-337
    const int64_t hbr = v.br / 2;
    const int64_t r0 = r / hbr * hbr, c0 = c / v.bc * v.bc;
    const int64_t k = r - r0, l = c - c0;
    const uint64_t s = cell_src(v, r0, c0, k, l);
    R* d = dest + 2 * r0 * v.bc + c0 * length + l + 2 * k * v.bc;
    d[0] = src[2 * s];
    d[v.bc] = src[2 * s + 1];
  }
}

template <typename R, bool kComplex>
__global__ void k_unpack(t1m_view v, const R* __restrict__ src,
                         R* __restrict__ dest) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= v.nrs * v.ncs) return;
  const int64_t k = e % v.nrs, l = e / v.nrs;
  // t1m.h:411-420: dense detection on the block's first strides
  const uint64_t rsc = v.rbs[0], csc = v.cbs[0];
  const uint64_t di = (rsc > 0 && csc > 0)
                          ? uint64_t(k) * rsc + uint64_t(l) * csc + (v.rs[0] + v.cs[0])
                          : v.rs[k] + v.cs[l];
  if constexpr (!kComplex) {
    dest[di] += src[k + l * v.br];
  } else {
    const int64_t si = 2 * k + l * v.br;   // t1m.h:424-427
    dest[2 * di] += src[si];
    dest[2 * di + 1] += src[si + 1];
  }
}

inline unsigned blocks_for(int64_t n, int threads) {
  const int64_t b = (n + threads - 1) / threads;
  return unsigned(b < 1 ? 1 : b);
}

}  // namespace

hipError_t launch_scatter_maps(const MapBatch& batch, hipStream_t stream) {
  int64_t maxlen = 1;
  for (int i = 0; i < batch.count; ++i)
    maxlen = batch.maps[i].len > maxlen ? batch.maps[i].len : maxlen;
  unsigned gx = blocks_for(maxlen, 256);
  if (gx > 4096) gx = 4096;
  hipLaunchKernelGGL(k_scatter_maps, dim3(gx, batch.count), dim3(256), 0,
                     stream, batch);
  return hipGetLastError();
}

hipError_t launch_block_scatter(const uint64_t* scat, int64_t len, int64_t b,
                                uint64_t* out, hipStream_t stream) {
  const int64_t nblk = len / b + (len % b > 0);
  if (nblk == 0) return hipSuccess;
  hipLaunchKernelGGL(k_block_scatter, dim3(blocks_for(nblk, 256)), dim3(256),
                     0, stream, scat, len, b, out);
  return hipGetLastError();
}

hipError_t launch_pack_block(int dtype, int label, const t1m_view& v,
                             int64_t length, const void* src, void* dest,
                             hipStream_t stream) {
  const int64_t n = v.nrs * v.ncs;
  if (n == 0) return hipSuccess;
  const dim3 g(blocks_for(n, 256)), b(256);
  const bool A = label == 0;
  switch (dtype) {
    case T1M_FLOAT32:
      if (A) hipLaunchKernelGGL((k_pack_block<float, false, true>), g, b, 0, stream, v, length, (const float*)src, (float*)dest);
      else hipLaunchKernelGGL((k_pack_block<float, false, false>), g, b, 0, stream, v, length, (const float*)src, (float*)dest);
      break;
    case T1M_FLOAT64:
      if (A) hipLaunchKernelGGL((k_pack_block<double, false, true>), g, b, 0, stream, v, length, (const double*)src, (double*)dest);
      else hipLaunchKernelGGL((k_pack_block<double, false, false>), g, b, 0, stream, v, length, (const double*)src, (double*)dest);
      break;
    case T1M_COMPLEX64:
      if (A) hipLaunchKernelGGL((k_pack_block<float, true, true>), g, b, 0, stream, v, length, (const float*)src, (float*)dest);
      else hipLaunchKernelGGL((k_pack_block<float, true, false>), g, b, 0, stream, v, length, (const float*)src, (float*)dest);
      break;
    default:
      if (A) hipLaunchKernelGGL((k_pack_block<double, true, true>), g, b, 0, stream, v, length, (const double*)src, (double*)dest);
      else hipLaunchKernelGGL((k_pack_block<double, true, false>), g, b, 0, stream, v, length, (const double*)src, (double*)dest);
      break;
  }
  return hipGetLastError();
}

hipError_t launch_unpack(int dtype, const t1m_view& v, const void* src,
                         void* dest, hipStream_t stream) {
  const int64_t n = v.nrs * v.ncs;
  if (n == 0) return hipSuccess;
  const dim3 g(blocks_for(n, 256)), b(256);
  switch (dtype) {
    case T1M_FLOAT32:
      hipLaunchKernelGGL((k_unpack<float, false>), g, b, 0, stream, v, (const float*)src, (float*)dest);
      break;
    case T1M_FLOAT64:
      hipLaunchKernelGGL((k_unpack<double, false>), g, b, 0, stream, v, (const double*)src, (double*)dest);
      break;
    case T1M_COMPLEX64:
      hipLaunchKernelGGL((k_unpack<float, true>), g, b, 0, stream, v, (const float*)src, (float*)dest);
      break;
    default:
      hipLaunchKernelGGL((k_unpack<double, true>), g, b, 0, stream, v, (const double*)src, (double*)dest);
      break;
  }
  return hipGetLastError();
}

}  // namespace t1m_hip
