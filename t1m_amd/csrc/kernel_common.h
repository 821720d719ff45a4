// Device-side helpers shared by the contraction kernels.
#pragma once

#include <type_traits>

#include "internal.h"

namespace t1m_hip {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T> struct Vec;
template <> struct Vec<float> { using c2 = float2; using acc = f32x4; };
template <> struct Vec<double> { using c2 = double2; using acc = f64x4; };

__device__ inline f32x4 mfma16(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ inline f64x4 mfma16(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <typename Idx>
struct KArgs {
  const void* A;
  const void* B;
  void* C;
  const Idx* rsA;
  const Idx* csA;
  const Idx* rsB;
  const Idx* csB;
  const Idx* rsC;
  const Idx* csC;
  int64_t M, N, K;
  double alpha_re, alpha_im;
  int32_t conj_a, conj_b;
  int32_t a_mfast, b_kfast;
  int32_t tiles_m, tiles_n;
  int32_t ksplit;    // workgroups per tile along K (1: no split)
  int32_t kchunk;    // contracted extent per split, a multiple of the stage depth
  int32_t atomic_c;  // add into C atomically (split K, or outputs that alias)
  int32_t nbatch;    // batch problems in this launch (1: none)
  const int64_t* boffA;   // byte offsets of the batch problems (nbatch > 1)
  const int64_t* boffB;
  const int64_t* boffC;
  int64_t min_rsA, min_csA, min_rsB, min_csB, min_rsC, min_csC;   // smallest map entries
  int32_t span32;   // A, B and C span < 2^32 bytes from their rebased bases
};

// Moves the operand pointers of q to batch problem z (of this launch).
template <typename Idx>
__device__ inline void batch_shift(KArgs<Idx>& q, int z) {
  if (q.boffA) {
    q.A = static_cast<const char*>(q.A) + q.boffA[z];
    q.B = static_cast<const char*>(q.B) + q.boffB[z];
    q.C = static_cast<char*>(q.C) + q.boffC[z];
  }
}

// Every buffer a kernel here reads or writes through a pointer is global
// memory, but pointers read from device memory (the persistent kernels'
// per-instance records) are generic to the compiler, which then emits flat
// loads, stores and atomics.  A flat operation counts in lgkmcnt as well as
// vmcnt, so each wait for an LDS operand would also wait for the global loads
// in flight.  The gathers and the C updates access memory through gmem(),
// a view of the pointer in the global address space, so the compiler emits
// global operations whatever the pointer's origin.
template <typename T>
using global_t = __attribute__((address_space(1))) T;
template <typename T>
__device__ inline global_t<T>* gmem(T* q) {
  return (global_t<T>*)q;
}
// *q read through the global view.  HIP's vector types (float2, double2)
// are classes whose assignment cannot bind an object of another address
// space, so their bits are read as a plain vector of words.
template <typename T>
__device__ inline T gload(const T* q) {
  if constexpr (std::is_class_v<T>) {
    static_assert(sizeof(T) % 4 == 0 && sizeof(T) <= 16, "vector type");
    typedef unsigned W __attribute__((ext_vector_type(sizeof(T) / 4)));
    typedef const __attribute__((address_space(1))) W GW;
    return __builtin_bit_cast(T, *(GW*)q);
  } else {
    return *gmem(q);
  }
}
// Unsafe (no-return, denormal-agnostic) atomic add into global memory.
template <typename T>
__device__ inline void atomic_add_gmem(T* q, T v) {
  [[clang::atomic(ignore_denormal_mode)]] {
    __hip_atomic_fetch_add(gmem(q), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The value of the neighbouring lane (lanes 2t and 2t+1 swap), by a DPP
// quad permutation [1, 0, 3, 2]: no LDS round trip.
__device__ inline float swap_pair(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1,
                                                            0xF, 0xF, false));
}
__device__ inline double swap_pair(double v) {
  const int2 w = __builtin_bit_cast(int2, v);
  int2 r;
  r.x = __builtin_amdgcn_mov_dpp(w.x, 0xB1, 0xF, 0xF, false);
  r.y = __builtin_amdgcn_mov_dpp(w.y, 0xB1, 0xF, 0xF, false);
  return __builtin_bit_cast(double, r);
}

// Epilogue of a tile: C[addr(r, c, q)] += val(r, c, q) for this lane's
// outputs, NR row groups x NC column groups x 4, where has(r, c) says which
// (row group, column group) pairs the lane holds (compile-time) and ok(r, c,
// q) which outputs exist (addr of the others must still be a valid address
// of C).  With `atomic` (K split over workgroups) the adds are atomic.
// Otherwise C is read for GR row groups at a time, then updated and written:
// a compiler cannot move a C read above an earlier C write (they might
// alias), so an element-by-element read-modify-write would wait out a
// memory latency per element.
template <typename T, int NR, int NC, int GR, typename Has, typename Ok, typename Val,
          typename Addr>
__device__ __forceinline__ void store_tile(bool atomic, Has has, Ok ok, Val val, Addr addr) {
  if (atomic) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if (has(r, c))
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const T v = val(r, c, q);
            if (ok(r, c, q)) atomic_add_gmem(addr(r, c, q), v);
          }
    return;
  }
#pragma unroll
  for (int r0 = 0; r0 < NR; r0 += GR) {
    T cv[GR][NC][4];
#pragma unroll
    for (int r = r0; r < r0 + GR && r < NR; ++r)
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if (has(r, c))
#pragma unroll
          for (int q = 0; q < 4; ++q) cv[r - r0][c][q] = *gmem(addr(r, c, q));
#pragma unroll
    for (int r = r0; r < r0 + GR && r < NR; ++r)
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if (has(r, c))
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const T v = val(r, c, q);
            if (ok(r, c, q)) *gmem(addr(r, c, q)) = cv[r - r0][c][q] + v;
          }
  }
}

// The two halves of store_tile's non-atomic path for a caller that loads
// the C values early: all of them into cv, then cv + value stored.
template <typename T, int NR, int NC, typename Has, typename Addr>
__device__ __forceinline__ void load_c_tile(Has has, Addr addr, T (&cv)[NR][NC][4]) {
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int c = 0; c < NC; ++c)
      if (has(r, c))
#pragma unroll
        for (int q = 0; q < 4; ++q) cv[r][c][q] = *gmem(addr(r, c, q));
}

template <typename T, int NR, int NC, typename Has, typename Ok, typename Val, typename Addr>
__device__ __forceinline__ void add_store_c_tile(Has has, Ok ok, Val val, Addr addr,
                                                 const T (&cv)[NR][NC][4]) {
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int c = 0; c < NC; ++c)
      if (has(r, c))
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const T v = val(r, c, q);
          if (ok(r, c, q)) *gmem(addr(r, c, q)) = cv[r][c][q] + v;
        }
}

// C[idx] += v; atomically when several workgroups share the tile (split K).
template <typename T>
__device__ inline void accumulate(T* dst, T v, bool atomic) {
  if (atomic)
    atomic_add_gmem(dst, v);
  else
    *gmem(dst) += v;
}

// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2).
// Renumber so that XCD x receives a contiguous range of logical tiles, and
// order logical tiles in bands of kGroupM tile-rows (column-major inside a
// band) so a range covers a compact 2-D block of C that shares A row panels
// and B column panels.
constexpr int kXcds = 8;
constexpr int kGroupM = 4;

__device__ inline void tile_coords_of(int w, int tiles_m, int tiles_n, int& tm, int& tn) {
  const int nwg = tiles_m * tiles_n;
  if (nwg % kXcds == 0) w = (w % kXcds) * (nwg / kXcds) + w / kXcds;
  const int band = kGroupM * tiles_n;
  const int b0 = w / band * kGroupM;
  const int rows = (tiles_m - b0) < kGroupM ? (tiles_m - b0) : kGroupM;
  const int r = w % band;
  tm = b0 + r % rows;
  tn = r / rows;
}


// Gather of one operand tile (BX free-index values x BK contracted values)
// from strided global memory by NT lanes.  Element e = tid + NT u of the
// tile sits at (x, k) = (e % BX, e / BX) when lanes run along the free index
// (XF) and at (e / BK, e % BK) otherwise; XF follows the operand's unit-stride
// direction, so consecutive lanes read consecutive addresses.  Loads are
// issued without branches: out-of-range rows / k read a valid in-range
// element and are zeroed when the tile is written to LDS, after the main
// loop's multiply has overlapped their latency.  The k offsets of the next
// stage are fetched one stage ahead (load_k) so no data load waits on an
// index load.  When BX*BK is not a multiple of NT the last slot is only
// used by the first lanes (valid()).
template <typename T2, typename Idx, int BX, int BK, int NT, bool XF>
struct TileGather {
  static constexpr int TOTAL = BX * BK;
  static constexpr int E = (TOTAL + NT - 1) / NT;
  static constexpr bool RAGGED = (TOTAL % NT) != 0;
  Idx xoff[E];
  Idx koff[E];
  T2 val[E];
  uint32_t xok;

  __device__ static constexpr int xc(unsigned e) {
    return int(XF ? e % unsigned(BX) : e / unsigned(BK));
  }
  __device__ static constexpr int kc(unsigned e) {
    return int(XF ? e / unsigned(BX) : e % unsigned(BK));
  }
  __device__ static bool valid(unsigned tid, int u) {
    return !RAGGED || u + 1 < E || tid + unsigned(NT) * u < unsigned(TOTAL);
  }
  __device__ static unsigned elem(unsigned tid, int u) {
    // Clamp the unused tail slots onto element 0 so every address is valid.
    return valid(tid, u) ? tid + unsigned(NT) * u : 0u;
  }

  __device__ void init(unsigned tid, int64_t x0, int64_t X, const Idx* xmap) {
    xok = 0;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int64_t xg = x0 + xc(elem(tid, u));
      xoff[u] = gmem(xmap)[xg < X ? xg : X - 1];
      xok |= ((xg < X && valid(tid, u)) ? 1u : 0u) << u;
    }
  }
  __device__ void load_k(unsigned tid, int p0, int K, const Idx* kmap) {
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int kg = p0 + kc(elem(tid, u));
      koff[u] = gmem(kmap)[kg < K ? kg : K - 1];
    }
  }
  __device__ void load_data(const T2* __restrict__ base) { load_data_to(base, val); }
  __device__ void load_data_to(const T2* __restrict__ base, T2 (&out)[E]) const {
#pragma unroll
    for (int u = 0; u < E; ++u) out[u] = gload(base + (xoff[u] + koff[u]));
  }
  // Entry u of a stage starting at p0: zero when out of range, conjugated
  // if asked (sign = -1).
  template <typename T>
  __device__ T2 finished(unsigned tid, int p0, int K, T sign, int u, T2 v) const {
    const int kg = p0 + kc(elem(tid, u));
    const bool ok = ((xok >> u) & 1u) && kg < K;
    v.x = ok ? v.x : T(0);
    v.y = ok ? sign * v.y : T(0);
    return v;
  }
  // Per-element view shared with PairGather: slot u holds PER = 1 element.
  static constexpr int PER = 1;
  using V = T2;
  __device__ static int xcj(unsigned tid, int u, int) { return xc(elem(tid, u)); }
  __device__ static int kcj(unsigned tid, int u, int) { return kc(elem(tid, u)); }
  template <typename T>
  __device__ T2 finished(unsigned tid, int p0, int K, T sign, int u, int, T2 v) const {
    return finished(tid, p0, K, sign, u, v);
  }
  // Zero out-of-range entries (real data).
  __device__ void finish_real(unsigned tid, int p0, int K) {
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int kg = p0 + kc(elem(tid, u));
      const bool ok = ((xok >> u) & 1u) && kg < K;
      val[u] = ok ? val[u] : T2(0);
    }
  }
  // Zero out-of-range entries and conjugate if asked (sign = -1).
  template <typename T>
  __device__ void finish_one(unsigned tid, int p0, int K, T sign, int u) {
    const int kg = p0 + kc(elem(tid, u));
    const bool ok = ((xok >> u) & 1u) && kg < K;
    val[u].x = ok ? val[u].x : T(0);
    val[u].y = ok ? sign * val[u].y : T(0);
  }
  template <typename T>
  __device__ void finish(unsigned tid, int p0, int K, T sign) {
#pragma unroll
    for (int u = 0; u < E; ++u) finish_one(tid, p0, K, sign, u);
  }
};

// complex<float> tile gather with two neighbours along the operand's unit-
// stride direction per lane: slot u holds the pair (x, x+1) at one k (XF)
// or (k, k+1) at one x, read with one 16-byte load.  Valid when the gather
// direction's map has map[i + 1] = map[i] + 1 for every even i (leading mode
// of unit stride and even extent; the extent of the direction is then even,
// so a clamped pair stays inside the tensor) -- half the load instructions
// and address arithmetic of TileGather for the same bytes.
typedef float f4a8 __attribute__((ext_vector_type(4), aligned(8)));

template <typename Idx, int BX, int BK, int NT, bool XF>
struct PairGather {
  static_assert(BX % 2 == 0 && BK % 2 == 0, "pairs");
  static constexpr int PX = XF ? BX / 2 : BX;   // pair grid
  static constexpr int PK = XF ? BK : BK / 2;
  static constexpr int TOTAL = PX * PK;
  static constexpr int E = (TOTAL + NT - 1) / NT;
  static constexpr bool RAGGED = (TOTAL % NT) != 0;
  static constexpr int PER = 2;
  static_assert(2 * E <= 32, "validity bits");
  using V = f4a8;   // (re, im) of the pair's first element, then of the second
  Idx xoff[E];
  Idx koff[E];
  V val[E];
  uint32_t xok;     // bit 2u + j: element j of slot u has its x inside

  __device__ static bool valid(unsigned tid, int u) {
    return !RAGGED || u + 1 < E || tid + unsigned(NT) * u < unsigned(TOTAL);
  }
  __device__ static unsigned elem(unsigned tid, int u) {
    return valid(tid, u) ? tid + unsigned(NT) * u : 0u;
  }
  __device__ static int xcj(unsigned tid, int u, int j) {
    const unsigned e = elem(tid, u);
    return XF ? 2 * int(e % unsigned(PX)) + j : int(e / unsigned(PK));
  }
  __device__ static int kcj(unsigned tid, int u, int j) {
    const unsigned e = elem(tid, u);
    return XF ? int(e / unsigned(PX)) : 2 * int(e % unsigned(PK)) + j;
  }
  __device__ void init(unsigned tid, int64_t x0, int64_t X, const Idx* xmap) {
    xok = 0;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int64_t xg = x0 + xcj(tid, u, 0);
      // an out-of-range pair reads the last pair (XF: X is even) or row
      xoff[u] = gmem(xmap)[xg < X ? xg : (XF ? X - 2 : X - 1)];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int64_t xj = x0 + xcj(tid, u, j);
        xok |= ((xj < X && valid(tid, u)) ? 1u : 0u) << (2 * u + j);
      }
    }
  }
  __device__ void load_k(unsigned tid, int p0, int K, const Idx* kmap) {
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int kg = p0 + kcj(tid, u, 0);
      koff[u] = gmem(kmap)[kg < K ? kg : (XF ? K - 1 : K - 2)];
    }
  }
  __device__ void load_data(const float2* __restrict__ base) { load_data_to(base, val); }
  __device__ void load_data_to(const float2* __restrict__ base, V (&out)[E]) const {
    // (the pair is 8-byte aligned: a typed global view keeps that alignment)
    typedef const __attribute__((address_space(1))) V GV;
#pragma unroll
    for (int u = 0; u < E; ++u) out[u] = *(GV*)(base + (xoff[u] + koff[u]));
  }
  // Element j of slot u of a stage starting at p0: zero when out of range,
  // conjugated if asked (sign = -1).
  template <typename T>
  __device__ float2 finished(unsigned tid, int p0, int K, T sign, int u, int j, V v) const {
    const int kg = p0 + kcj(tid, u, j);
    const bool ok = ((xok >> (2 * u + j)) & 1u) && kg < K;
    const float re = j == 0 ? v.x : v.z, im = j == 0 ? v.y : v.w;
    return float2{ok ? re : 0.f, ok ? float(sign) * im : 0.f};
  }
};

__device__ inline void tile_coords(int tiles_m, int tiles_n, int& tm, int& tn) {
  tile_coords_of(int(blockIdx.x), tiles_m, tiles_n, tm, tn);
}

// Tile, K range and batch problem of the workgroup: the grid is
// tiles x ksplit x nbatch, tile-major; q's operands are moved to the batch
// problem.
template <typename Idx>
__device__ inline void tile_and_krange(KArgs<Idx>& q, int& tm, int& tn, int& kbeg,
                                       int& kend) {
  const KArgs<Idx>& p = q;
  const int ntiles = p.tiles_m * p.tiles_n;
  const int per = ntiles * p.ksplit;
  const int z = int(blockIdx.x) / per;
  const int r = int(blockIdx.x) - z * per;
  batch_shift(q, z);
  const int w = r % ntiles;
  const int split = r / ntiles;
  tile_coords_of(w, p.tiles_m, p.tiles_n, tm, tn);
  kbeg = split * p.kchunk;
  const int64_t e = int64_t(kbeg) + p.kchunk;
  kend = e < p.K ? int(e) : int(p.K);
}


// Kernel arguments of problem g for bm x bn tiles with stage depth bk.  K is
// split over several workgroups per tile when the tiles cannot fill the CUs
// (fewer than half as many tiles as CUs) and each split keeps >= 8 stages.
template <typename Idx>
KArgs<Idx> make_args(const GemmProblem& g, int bm, int bn, int bk = 16) {
  KArgs<Idx> a;
  a.A = g.A; a.B = g.B; a.C = g.C;
  a.rsA = static_cast<const Idx*>(g.rsA);
  a.csA = static_cast<const Idx*>(g.csA);
  a.rsB = static_cast<const Idx*>(g.rsB);
  a.csB = static_cast<const Idx*>(g.csB);
  a.rsC = static_cast<const Idx*>(g.rsC);
  a.csC = static_cast<const Idx*>(g.csC);
  a.M = g.M; a.N = g.N; a.K = g.K;
  a.alpha_re = g.alpha_re; a.alpha_im = g.alpha_im;
  a.conj_a = g.conj_a; a.conj_b = g.conj_b;
  a.a_mfast = g.a_mfast; a.b_kfast = g.b_kfast;
  a.tiles_m = bm > 0 ? int((g.M + bm - 1) / bm) : 0;
  a.tiles_n = bn > 0 ? int((g.N + bn - 1) / bn) : 0;
  a.ksplit = 1;
  a.kchunk = int(g.K);
  a.nbatch = int32_t(g.L > 0 ? g.L : 1);
  a.boffA = g.L > 1 ? g.boffA : nullptr;
  a.boffB = g.L > 1 ? g.boffB : nullptr;
  a.boffC = g.L > 1 ? g.boffC : nullptr;
  a.min_rsA = g.min_rsA; a.min_csA = g.min_csA;
  a.min_rsB = g.min_rsB; a.min_csB = g.min_csB;
  a.min_rsC = g.min_rsC; a.min_csC = g.min_csC;
  a.span32 = g.span32;
  const int64_t tiles = int64_t(a.tiles_m) * a.tiles_n * (g.L > 0 ? g.L : 1);
  const int cus = g.cus > 0 ? g.cus : 256;
  if (g.split_k == 0 && bm > 0 && tiles > 0 && 2 * tiles < cus) {
    const int64_t want = (2 * int64_t(cus) + tiles - 1) / tiles;     // ~2 waves of workgroups
    const int64_t most = g.K / (8 * int64_t(bk));                     // >= 8 stages per split
    const int64_t split = want < most ? want : most;
    if (split > 1) {
      const int64_t chunk = ((g.K + split - 1) / split + bk - 1) / bk * bk;
      a.kchunk = int(chunk);
      a.ksplit = int((g.K + chunk - 1) / chunk);
    }
  }
  a.atomic_c = (a.ksplit > 1 || g.c_overlap) ? 1 : 0;
  return a;
}

// Runs launch(q, grid) over blocks_per_problem x nbatch workgroups of
// `threads` lanes, in several launches when one grid would pass 2^31 lanes
// (each launch sees its batch problems as 0.. with the offset tables moved).
template <typename Idx, typename F>
hipError_t launch_batches(const KArgs<Idx>& a, int64_t blocks_per_problem, int threads, F&& launch) {
  const int64_t limit = ((int64_t(1) << 31) - 1) / threads;
  if (blocks_per_problem <= 0) return hipSuccess;
  if (blocks_per_problem > limit) return hipErrorInvalidConfiguration;
  const int64_t L = a.nbatch > 0 ? a.nbatch : 1;
  const int64_t zmax = limit / blocks_per_problem;
  for (int64_t z0 = 0; z0 < L; z0 += zmax) {
    KArgs<Idx> q = a;
    const int64_t nz = L - z0 < zmax ? L - z0 : zmax;
    q.nbatch = int32_t(nz);
    if (q.boffA) { q.boffA += z0; q.boffB += z0; q.boffC += z0; }
    launch(q, unsigned(blocks_per_problem * nz));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace t1m_hip
