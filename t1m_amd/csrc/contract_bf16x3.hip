// complex<float> contraction in split-BF16 arithmetic ("BF16x3") on gfx950.
//
// Every real operand value x of the 1M GEMM (include/t1m/t1m.h:310-337) is
// split once, while it is staged into LDS, into x_hi = bf16_rne(x) and
// x_lo = bf16_rne(x - x_hi) (x - x_hi is exact in FP32).  Each product is
// then a_hi*b_hi + a_hi*b_lo + a_lo*b_hi (the a_lo*b_lo term, below 2^-16
// relative, is dropped), each term on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation (Henry, Tang & Heinecke, "Leveraging the bfloat16 artificial
// intelligence datatype for higher-precision computations", ARITH 2019).
//
// Layout is the transposed 1M form of contract_mfma.hip with 32 real k per
// MFMA: lane (j or real row c, g = lane/16) supplies 8 bf16 = 4 complex k
// values 4g..4g+3 as (re, im) pairs.  For B (MFMA rows = columns j of C) the
// pairs are B's own (re, im), i.e. the packed words straight from LDS.  For
// A (MFMA columns = real rows 2i+s of C_r) the 1M expansion of A_r row 2i+s
// is (re, -im) for s = 0 and (im, re) for s = 1 per complex k: one v_perm
// (swap the bf16 halves) and one xor (flip the sign bit of the high half)
// per packed word, with lane-constant selectors.  Splitting commutes with
// negation under round-to-nearest-even, so the expanded hi/lo words are
// exactly the splits of the expanded values.
//
// LDS holds one 32-bit word per complex element and plane (hi, lo), rows of
// m (for A) / n (for B) with the k index contiguous, so a lane's fragment is
// one 16-byte read per plane.  Operands gathered along m / n are staged in
// quads of four consecutive k (one 16-byte LDS store per plane and row).
// That is the 64 x 64 family below; the 144 x 144 tiles use the row-pair
// form further down.
#include "internal.h"
#include "kernel_common.h"

#include <type_traits>

// Layout choices for the 64 x 64 tiles, measured on complex<float> 48^4
// contractions in the four gather-direction combinations (timing variants
// built from switches that no longer exist): A unexpanded (expanded rows
// cost occupancy), quad staging of A (+5..15 %) and of B.

namespace t1m_hip {

namespace {

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// (re, im) -> packed bf16 hi word and lo word (round to nearest even).  The
// residuals x - hi are exact in FP32; each is one v_dot2c_f32_bf16 of the hi
// word with (-1, 0) or (0, -1) accumulated onto x, in place of unpacking hi
// and subtracting (4 VALU operations per complex value instead of 6: the
// staging VALU work, not the MFMAs, bounds these kernels' issue).
// The (-1, 0) and (0, -1) bf16 pairs are moved into scalar registers by
// hand: as an inline constant the compiler encodes (-1, 0) as "-1.0", which
// this instruction reads as the f32 bit pattern, i.e. (0, -1).
__device__ inline bf16x2 sreg_bf16x2(uint32_t bits) {
  uint32_t r;
  asm("s_mov_b32 %0, %1" : "=s"(r) : "n"(bits));
  return __builtin_bit_cast(bf16x2, r);
}
__device__ inline void split_pair(float re, float im, uint32_t& hi,
                                  uint32_t& lo) {
  const f32x2 v = {re, im};
  const bf16x2 h = __builtin_convertvector(v, bf16x2);
  const bf16x2 mre = sreg_bf16x2(0x0000BF80u);   // (-1, 0)
  const bf16x2 mim = sreg_bf16x2(0xBF800000u);   // (0, -1)
  const f32x2 r = {__builtin_amdgcn_fdot2_f32_bf16(h, mre, re, false),
                   __builtin_amdgcn_fdot2_f32_bf16(h, mim, im, false)};
  const bf16x2 l = __builtin_convertvector(r, bf16x2);
  hi = __builtin_bit_cast(uint32_t, h);
  lo = __builtin_bit_cast(uint32_t, l);
}

__device__ inline f32x4 mfma_bf16(const u32x4& a, const u32x4& b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(
      __builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// Gather of a BX x BK tile along x (the operand's unit-stride direction) in
// quads: lane work unit (x, kq) holds the four consecutive k = 4kq..4kq+3 of
// row x, so it can store them as one 16-byte LDS write per plane into the
// k-contiguous rows the MFMA fragments read.  Lanes run along x, so each of
// the four global loads of a unit is coalesced across lanes.
template <typename Idx, int BX, int BK, int NT>
struct QuadGather {
  static constexpr int NQ = BX * (BK / 4);
  static constexpr int E = (NQ + NT - 1) / NT;
  static constexpr bool RAGGED = (NQ % NT) != 0;
  Idx xoff[E];
  Idx koff[E][4];
  float2 val[E][4];
  uint32_t xok;

  __device__ static constexpr int xc(unsigned e) { return int(e % unsigned(BX)); }
  __device__ static constexpr int kq(unsigned e) { return int(e / unsigned(BX)); }
  __device__ static bool valid(unsigned tid, int u) {
    return !RAGGED || u + 1 < E || tid + unsigned(NT) * u < unsigned(NQ);
  }
  __device__ static unsigned elem(unsigned tid, int u) {
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
    for (int u = 0; u < E; ++u)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int kg = p0 + 4 * kq(elem(tid, u)) + t;
        koff[u][t] = gmem(kmap)[kg < K ? kg : K - 1];
      }
  }
  __device__ void load_data(const float2* __restrict__ base) { load_data_to(base, val); }
  __device__ void load_data_to(const float2* __restrict__ base, float2 (&out)[E][4]) const {
#pragma unroll
    for (int u = 0; u < E; ++u)
#pragma unroll
      for (int t = 0; t < 4; ++t) out[u][t] = gload(base + (xoff[u] + koff[u][t]));
  }
  // Whether entry (u, t) of a stage starting at p0 lies inside the problem.
  __device__ bool inside(unsigned tid, int p0, int K, int u, int t) const {
    return ((xok >> u) & 1u) && p0 + 4 * kq(elem(tid, u)) + t < K;
  }
  __device__ void finish_one(unsigned tid, int p0, int K, float sign, int u) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int kg = p0 + 4 * kq(elem(tid, u)) + t;
      const bool ok = ((xok >> u) & 1u) && kg < K;
      val[u][t].x = ok ? val[u][t].x : 0.f;
      val[u][t].y = ok ? sign * val[u][t].y : 0.f;
    }
  }
  __device__ void finish(unsigned tid, int p0, int K, float sign) {
#pragma unroll
    for (int u = 0; u < E; ++u) finish_one(tid, p0, K, sign, u);
  }
};

template <typename G> struct IsQuad : std::false_type {};
template <typename Idx, int BX, int BK, int NT>
struct IsQuad<QuadGather<Idx, BX, BK, NT>> : std::true_type {};

template <typename Idx, int BM, int BN, int BK, int WGM, int WGN, bool AF, bool BF>
__device__ __forceinline__ void contract_tile_bf16x3(const KArgs<Idx>& p, int tm, int tn, int kbeg, int kend) {
  constexpr int NT = 64 * WGM * WGN;
  constexpr int WTM = BM / WGM;
  constexpr int MB = WTM / 8;
  constexpr int NBT = BN / 16;
  constexpr int NBW = NBT / WGN;
  constexpr bool XT = NBT % WGN != 0;
  constexpr int XB = XT ? (MB + WGN - 1) / WGN : 0;
  constexpr int XB_ALLOC = XT ? XB : 1;
  constexpr int PITCH = BK + 4;   // words per LDS row (padded: conflict-light reads)
  static_assert(BM < 144, "144 x 144 tiles use the row-pair kernel");
  __shared__ __attribute__((aligned(16))) uint32_t sAh[2][BM][PITCH];
  __shared__ __attribute__((aligned(16))) uint32_t sAl[2][BM][PITCH];
  __shared__ __attribute__((aligned(16))) uint32_t sBh[2][BN][PITCH];
  __shared__ __attribute__((aligned(16))) uint32_t sBl[2][BN][PITCH];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  const int wm = wid % WGM, wn = wid / WGM;
  const int64_t i0 = int64_t(tm) * BM, j0 = int64_t(tn) * BN;
  const int64_t M = p.M, N = p.N;
  const int K = kend;     // contracted range [kbeg, kend) of this workgroup
  const float2* __restrict__ A = static_cast<const float2*>(p.A);
  const float2* __restrict__ B = static_cast<const float2*>(p.B);

  const unsigned utid = unsigned(tid);
  // Gathers along x store quads (16-byte writes); gathers along k store
  // single words, consecutive lanes filling consecutive words of a row.
  constexpr bool QA = AF, QB = BF;
  using GA = std::conditional_t<QA, QuadGather<Idx, BM, BK, NT>,
                                TileGather<float2, Idx, BM, BK, NT, AF>>;
  using GB = std::conditional_t<QB, QuadGather<Idx, BN, BK, NT>,
                                TileGather<float2, Idx, BN, BK, NT, BF>>;
  GA ga;
  GB gb;
  ga.init(utid, i0, M, p.rsA);
  gb.init(utid, j0, N, p.csB);
  const float sa = p.conj_a ? -1.f : 1.f;
  const float sb = p.conj_b ? -1.f : 1.f;

  // Stage one gathered unit (an element, or a quad of four k) of A / B into
  // LDS buffer buf: zero / conjugate, split into hi and lo words, store.
  auto write_a = [&](int buf, int p0, int u) {
    ga.finish_one(utid, p0, K, sa, u);
    const unsigned e = GA::elem(utid, u);
    if constexpr (QA) {
      u32x4 hi, lo;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        uint32_t h, l;
        split_pair(ga.val[u][t].x, ga.val[u][t].y, h, l);
        hi[t] = h;
        lo[t] = l;
      }
      if (GA::valid(utid, u)) {
        const int x = GA::xc(e), kw = 4 * GA::kq(e);
        *reinterpret_cast<u32x4*>(&sAh[buf][x][kw]) = hi;
        *reinterpret_cast<u32x4*>(&sAl[buf][x][kw]) = lo;
      }
    } else {
      uint32_t hi, lo;
      split_pair(ga.val[u].x, ga.val[u].y, hi, lo);
      if (GA::valid(utid, u)) {
        const int x = GA::xc(e), k = GA::kc(e);
        sAh[buf][x][k] = hi;
        sAl[buf][x][k] = lo;
      }
    }
  };
  auto write_b = [&](int buf, int p0, int u) {
    gb.finish_one(utid, p0, K, sb, u);
    const unsigned e = GB::elem(utid, u);
    if constexpr (QB) {
      u32x4 hi, lo;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        uint32_t h, l;
        split_pair(gb.val[u][t].x, gb.val[u][t].y, h, l);
        hi[t] = h;
        lo[t] = l;
      }
      if (GB::valid(utid, u)) {
        const int x = GB::xc(e), kw = 4 * GB::kq(e);
        *reinterpret_cast<u32x4*>(&sBh[buf][x][kw]) = hi;
        *reinterpret_cast<u32x4*>(&sBl[buf][x][kw]) = lo;
      }
    } else {
      uint32_t hi, lo;
      split_pair(gb.val[u].x, gb.val[u].y, hi, lo);
      if (GB::valid(utid, u)) {
        const int x = GB::xc(e), k = GB::kc(e);
        sBh[buf][x][k] = hi;
        sBl[buf][x][k] = lo;
      }
    }
  };
  constexpr int NU = GA::E + GB::E;
  auto write_unit = [&](int buf, int p0, int i) {
    if (i < GA::E) write_a(buf, p0, i);
    else write_b(buf, p0, i - GA::E);
  };
  auto swrite = [&](int buf, int p0) {
#pragma unroll
    for (int i = 0; i < NU; ++i) write_unit(buf, p0, i);
  };

  f32x4 acc[MB][NBW];
  f32x4 accx[XB_ALLOC];
#pragma unroll
  for (int a = 0; a < MB; ++a)
#pragma unroll
    for (int b = 0; b < NBW; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int x = 0; x < XB_ALLOC; ++x) accx[x] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int g = lane >> 4;
  const int l16 = lane & 15;
  const bool odd = (lane & 1) != 0;
  const int am0 = wm * WTM + (l16 >> 1);
  const int bn0 = wn * NBW * 16 + l16;
  const int bnx = NBW * WGN * 16 + l16;
  int xmb[XB_ALLOC];
#pragma unroll
  for (int x = 0; x < XB_ALLOC; ++x) {
    const int mb = wn * XB + x;
    xmb[x] = mb < MB ? mb : MB - 1;
  }
  // 1M expansion of a packed (re, im) word for real row parity s:
  // s = 0 -> (re, -im), s = 1 -> (im, re).
  const uint32_t psel = odd ? 0x01000302u : 0x03020100u;
  const uint32_t pxor = odd ? 0u : 0x80000000u;
  auto expand = [&](u32x4 w) {
    u32x4 r;
#pragma unroll
    for (int t = 0; t < 4; ++t) r[t] = __builtin_amdgcn_perm(w[t], w[t], psel) ^ pxor;
    return r;
  };
  // Word column of this lane's 4 complex k in the rows it reads.
  const int colA = 4 * g, colB = 4 * g;
  auto ldsA = [&](const uint32_t (*plane)[PITCH], int row) {
    return expand(*reinterpret_cast<const u32x4*>(&plane[row][colA]));
  };
  auto ldsB = [&](const uint32_t (*plane)[PITCH], int row) {
    return *reinterpret_cast<const u32x4*>(&plane[row][colB]);
  };

  // B fragments stay in registers for a k-chunk; A fragments stream per row
  // block.  Three products per tile: lo*hi and hi*lo first, then hi*hi.
  auto compute = [&](int buf) {
    static_assert(BK == 16, "one MFMA k-chunk per stage");
    // Raised wave priority over the MFMA stream: +4..7 % on 64 x 64 tiles
    // (measured on four 48^4 layouts).
    __builtin_amdgcn_s_setprio(1);
    u32x4 bh[NBW], bl[NBW];
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb) {
      bh[nb] = ldsB(sBh[buf], bn0 + nb * 16);
      bl[nb] = ldsB(sBl[buf], bn0 + nb * 16);
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      const u32x4 ah = ldsA(sAh[buf], am0 + mb * 8);
      const u32x4 al = ldsA(sAl[buf], am0 + mb * 8);
#pragma unroll
      for (int nb = 0; nb < NBW; ++nb) {
        acc[mb][nb] = mfma_bf16(bl[nb], ah, acc[mb][nb]);
        acc[mb][nb] = mfma_bf16(bh[nb], al, acc[mb][nb]);
        acc[mb][nb] = mfma_bf16(bh[nb], ah, acc[mb][nb]);
      }
    }
    if constexpr (XT) {
      const u32x4 xbh = ldsB(sBh[buf], bnx);
      const u32x4 xbl = ldsB(sBl[buf], bnx);
#pragma unroll
      for (int x = 0; x < XB; ++x) {
        const u32x4 ah = ldsA(sAh[buf], am0 + xmb[x] * 8);
        const u32x4 al = ldsA(sAl[buf], am0 + xmb[x] * 8);
        accx[x] = mfma_bf16(xbl, ah, accx[x]);
        accx[x] = mfma_bf16(xbh, al, accx[x]);
        accx[x] = mfma_bf16(xbh, ah, accx[x]);
      }
    }
    __builtin_amdgcn_s_setprio(0);
  };

  const int nst = (K - kbeg + BK - 1) / BK;
  ga.load_k(utid, kbeg, K, p.csA);
  gb.load_k(utid, kbeg, K, p.rsB);
  ga.load_data(A);
  gb.load_data(B);
  ga.load_k(utid, kbeg + BK, K, p.csA);
  gb.load_k(utid, kbeg + BK, K, p.rsB);
  swrite(0, kbeg);
  __syncthreads();
  // The loop body is branch-free (a conditional prefetch made the compiler
  // wait for every load in flight at the top of each iteration); the last
  // stage is peeled off so no data is fetched past the end.  The k-offset
  // prefetch of the stage after the last reads clamped, valid indices.
  for (int s = 0; s + 1 < nst; ++s) {
    ga.load_data(A);
    gb.load_data(B);
    ga.load_k(utid, kbeg + (s + 2) * BK, K, p.csA);
    gb.load_k(utid, kbeg + (s + 2) * BK, K, p.rsB);
    compute(s & 1);
    swrite((s + 1) & 1, kbeg + (s + 1) * BK);
    __syncthreads();
  }
  compute((nst - 1) & 1);

  // Epilogue (as in the FP32 1M kernel): complex rows i0 + am0 + 8 r' (re
  // or im by lane parity) times 4 columns per 16-column block.
  float* __restrict__ Cr = static_cast<float*>(p.C);
  const bool cplx_alpha = p.alpha_im != 0.0;
  const float ar = float(p.alpha_re), ai = float(p.alpha_im);
  constexpr int NR = MB + (XT ? XB : 0);
  constexpr int NC = NBW + (XT ? 1 : 0);
  Idx rofs[NR];
  bool rok[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int blk = r < MB ? r : wn * XB + (r - MB);
    const int64_t i = i0 + am0 + (blk < MB ? blk : MB - 1) * 8;
    rok[r] = i < M && blk < MB;
    rofs[r] = rok[r] ? gmem(p.rsC)[i] : Idx(0);
  }
  Idx cofs[NC][4];
  bool cok[NC][4];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t jb = c < NBW ? j0 + wn * NBW * 16 + c * 16 : j0 + NBW * WGN * 16;
      const int64_t j = jb + 4 * g + q;
      cok[c][q] = j < N;
      cofs[c][q] = cok[c][q] ? gmem(p.csC)[j] : Idx(0);
    }
  auto scaled = [&](float v) {
    if (cplx_alpha) {
      const float other = swap_pair(v);
      return odd ? ar * v + ai * other : ar * v - ai * other;
    }
    return ar * v;
  };
  auto value = [&](int r, int c, int q) -> float {
    if constexpr (XT) {
      if (r >= MB) return scaled(accx[r - MB][q]);
    }
    return scaled(acc[r][c][q]);
  };
  auto has = [](int r, int c) { return r < MB ? c < NBW : c == NC - 1; };
  auto ok = [&](int r, int c, int q) { return rok[r] && cok[c][q]; };
  auto addr = [&](int r, int c, int q) {
    return Cr + 2 * int64_t(rofs[r] + cofs[c][q]) + (odd ? 1 : 0);
  };
  store_tile<float, NR, NC, 3>(p.atomic_c != 0, has, ok, value, addr);
}

template <typename Idx, int BM, int BN, int BK, int WGM, int WGN, bool AF, bool BF>
__global__ void __launch_bounds__(64 * WGM * WGN)
k_contract_bf16x3(KArgs<Idx> p) {
  int tm, tn, kbeg, kend;
  tile_and_krange(p, tm, tn, kbeg, kend);
  contract_tile_bf16x3<Idx, BM, BN, BK, WGM, WGN, AF, BF>(p, tm, tn, kbeg, kend);
}

template <typename Idx, int BM, int BN, int BK, int WGM, int WGN, bool AF, bool BF>
__global__ void __launch_bounds__(64 * WGM * WGN)
k_contract_bf16x3_batched(const KArgs<Idx>* __restrict__ inst,
                          const BatchItem* __restrict__ items, int nitems) {
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const BatchItem w = items[it];
    if (w.inst < 0) continue;   // padding of an XCD-ordered work list
    const KArgs<Idx> p = inst[w.inst];
    contract_tile_bf16x3<Idx, BM, BN, BK, WGM, WGN, AF, BF>(p, w.tm, w.tn, 0, int(p.K));
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// 144 x 144 tiles in the row-pair form.
//
// MFMA columns are 16 complex rows i of C (not 16 real rows of C_r), and
// the 1M expansion sits on B: LDS holds A's packed (re, im) words as they
// are and, for every B element, the two words that pair with them,
// (re, -im) for the real part of C and (im, re) for the imaginary part
// (sum_p a_re b_re - a_im b_im and a_re b_im + a_im b_re).  Conjugating A
// or B only flips sign bits of those two words (two xor masks).  LDS holds
// B's first form only: the second is its half-swapped copy with another sign
// mask, formed in registers when a fragment is read (two VALU operations a
// word, beside the MFMA stream), which takes a third of the LDS writes and a
// quarter of the fragment reads off the LDS.  A lane's outputs are whole complex elements, C[i][j]
// for i = l16 of a 16-row unit and j = 4 g + t of a 16-column unit.
//
// The tile is 9 x 9 units of 16 complex rows x 16 columns.  Each wave owns
// a rectangle of units (its A row fragments and B column fragments per
// k-chunk, reused across the rectangle), laid out so that the two waves of
// each SIMD (waves w and w + 4) hold 20 or 21 units between them.  LDS rows
// are 16 words (16 complex k) without padding; the 16-byte chunk c of row x
// sits at chunk c ^ ((x >> 2) & 3): fragment reads (16 rows, one chunk) and
// the quad and word writes of the gathers all run at the contiguous rate
// (measured with an LDS pattern probe on gfx950).
//
// Out-of-range rows, columns and k (edge tiles, the last stage) are zeroed
// when staged; tiles and stages that are wholly inside skip those selects.

struct WaveRect { int r0, nr, c0, nc; };
constexpr WaveRect kRect144[8] = {
    {0, 5, 0, 2}, {0, 5, 2, 2}, {5, 4, 0, 3}, {5, 4, 3, 3},
    {0, 5, 4, 2}, {0, 5, 6, 2}, {0, 9, 8, 1}, {5, 4, 6, 2}};

constexpr int kP144 = 144;             // tile rows / columns
constexpr int kPW = 16;                // words per LDS row (16 complex k)
constexpr int kPPlane = kP144 * kPW;   // words per plane
constexpr int kPPlanes = 4;            // A hi, A lo, B re-hi, B re-lo

// Word offset of word w of LDS row x (chunk swizzle): rows of 16 words, the
// 16-byte chunk c of row x at c ^ ((x >> 2) & 3).
__device__ inline int pword(int x, int w) {
  return x * kPW + ((((w >> 2) ^ (x >> 2)) & 3) << 2) + (w & 3);
}

// The same plane in units of 16 rows: a unit holds its rows' four 16-byte
// chunks (words 4c .. 4c + 3) as four blocks of 16 rows, chunk-major.  A
// fragment read, lane (g, r) taking chunk g of row r0 + r, then puts every
// 16-lane group on one block in the same lane order, which gfx950 serves
// without bank conflicts, and so are 16-byte writes of 16 rows of one chunk.
// Measured with this kernel's accesses in isolation: the row layout above
// logged 4 conflict cycles per 16-byte read (1.8x the time of these units)
// and 8 per 16-byte write; its word writes (one k per lane along a row) are
// conflict-free, the units' are not.  So an operand staged in 16-byte quads
// (gathered along its rows) uses the units, one staged in words (gathered
// along k) the rows.
__device__ inline int pword_units(int x, int w) {
  return (x >> 4) * 256 + (w >> 2) * 64 + ((x & 15) << 2) + (w & 3);
}
template <bool UNITS>
__device__ inline int plane_word(int x, int w) {
  return UNITS ? pword_units(x, w) : pword(x, w);
}

// MFMA side of a 144 x 144 row-pair tile, shared by the staging paths:
// wave (r0, c0)'s NR x NC units, their accumulators, the products of one
// LDS stage, and the epilogue.  Per k-chunk every fragment is read from LDS
// once (and every B column's second form made once); the fragments of one
// dimension stay in registers while the other's stream -- B's (two forms of
// two planes, 16 registers a column) when they take no more registers than
// A's (two planes, 8 a row).
template <typename Idx, int NR, int NC>
struct PairUnits {
  static constexpr bool HOLD_B = 2 * NC <= NR;
  f32x4 acc[NR][NC][2];
  int r0, c0, g, l16;
  uint32_t mim_r;

  __device__ __forceinline__ void init(const KArgs<Idx>& p, int r0_, int c0_) {
    r0 = r0_;
    c0 = c0_;
    const int lane = int(threadIdx.x) & 63;
    g = lane >> 4;
    l16 = lane & 15;
    // B is stored in its re-form w ^ mre; the im-form is swap(w) ^ mim =
    // swap(r) ^ swap(mre) ^ mim.
    const uint32_t mre = (p.conj_a != 0) == (p.conj_b != 0) ? 0x80000000u : 0u;
    const uint32_t mim = (p.conj_b ? 0x00008000u : 0u) | (p.conj_a ? 0x80000000u : 0u);
    mim_r = mim ^ (mre ? 0x00008000u : 0u);
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        acc[r][c][0] = f32x4{0.f, 0.f, 0.f, 0.f};
        acc[r][c][1] = f32x4{0.f, 0.f, 0.f, 0.f};
      }
  }
  __device__ __forceinline__ u32x4 im_form(const u32x4& r) const {
    u32x4 x;
#pragma unroll
    for (int t = 0; t < 4; ++t) x[t] = __builtin_amdgcn_alignbit(r[t], r[t], 16) ^ mim_r;
    return x;
  }
  template <bool UNITS>
  __device__ __forceinline__ u32x4 frag(const uint32_t* pl, int row) const {
    return *reinterpret_cast<const u32x4*>(pl + plane_word<UNITS>(row, 4 * g));
  }
  // Three products per unit and part: lo*hi, hi*lo, hi*hi.
  __device__ __forceinline__ void unit(int r, int c, const u32x4& ah, const u32x4& al,
                                       const u32x4& rh, const u32x4& rl, const u32x4& ih,
                                       const u32x4& il) {
    acc[r][c][0] = mfma_bf16(rl, ah, acc[r][c][0]);
    acc[r][c][1] = mfma_bf16(il, ah, acc[r][c][1]);
    acc[r][c][0] = mfma_bf16(rh, al, acc[r][c][0]);
    acc[r][c][1] = mfma_bf16(ih, al, acc[r][c][1]);
    acc[r][c][0] = mfma_bf16(rh, ah, acc[r][c][0]);
    acc[r][c][1] = mfma_bf16(ih, ah, acc[r][c][1]);
  }
  // One k-chunk from the planes of an LDS stage (A hi, A lo, B re-hi, B re-lo);
  // UA / UB: A's / B's planes in 16-row units (pword_units), else in rows.
  template <bool UA = false, bool UB = false>
  __device__ __forceinline__ void compute(const uint32_t* stage) {
    const uint32_t* pl0 = stage;
    const uint32_t* pl1 = stage + kPPlane;
    const uint32_t* pl2 = stage + 2 * kPPlane;
    const uint32_t* pl3 = stage + 3 * kPPlane;
    if constexpr (HOLD_B) {
      u32x4 h0[NC], h1[NC], h2[NC], h3[NC];
#pragma unroll
      for (int x = 0; x < NC; ++x) {
        const int row = (c0 + x) * 16 + l16;
        h0[x] = frag<UB>(pl2, row);
        h1[x] = frag<UB>(pl3, row);
        h2[x] = im_form(h0[x]);
        h3[x] = im_form(h1[x]);
      }
#pragma unroll
      for (int y = 0; y < NR; ++y) {
        const int row = (r0 + y) * 16 + l16;
        const u32x4 ah = frag<UA>(pl0, row), al = frag<UA>(pl1, row);
#pragma unroll
        for (int x = 0; x < NC; ++x) unit(y, x, ah, al, h0[x], h1[x], h2[x], h3[x]);
      }
    } else {
      u32x4 h0[NR], h1[NR];
#pragma unroll
      for (int y = 0; y < NR; ++y) {
        const int row = (r0 + y) * 16 + l16;
        h0[y] = frag<UA>(pl0, row);
        h1[y] = frag<UA>(pl1, row);
      }
#pragma unroll
      for (int x = 0; x < NC; ++x) {
        const int row = (c0 + x) * 16 + l16;
        const u32x4 rh = frag<UB>(pl2, row), rl = frag<UB>(pl3, row);
        const u32x4 ih = im_form(rh), il = im_form(rl);
#pragma unroll
        for (int y = 0; y < NR; ++y) unit(y, x, h0[y], h1[y], rh, rl, ih, il);
      }
    }
  }
  // Epilogue: lane holds C[i][j] (re, im) for i = unit row + l16 and
  // j = unit column + 4 g + t.  C is read a row unit at a time (all of its
  // columns) before any of it is written back.  OFF32 (span32 plans):
  // 32-bit byte offsets from C's rebased base, one register an address.
  template <bool OFF32 = false>
  __device__ __forceinline__ void store(const KArgs<Idx>& p, int tm, int tn) const {
    const int64_t i0 = int64_t(tm) * kP144, j0 = int64_t(tn) * kP144;
    using Off = std::conditional_t<OFF32, uint32_t, Idx>;
    char* __restrict__ Cb = static_cast<char*>(p.C);
    if constexpr (OFF32) Cb += 8 * (p.min_rsC + p.min_csC);
    auto off = [&](const Idx* map, int64_t x, int64_t mn) -> Off {
      if constexpr (OFF32) return uint32_t(int64_t(map[x]) - mn) * 8u;
      else return map[x];
    };
    auto at = [&](Off ro, Off co) -> float2* {
      if constexpr (OFF32) return reinterpret_cast<float2*>(Cb + uint64_t(ro + co));
      else return reinterpret_cast<float2*>(Cb) + (int64_t(ro) + int64_t(co));
    };
    const float ar = float(p.alpha_re), ai = float(p.alpha_im);
    Off cofs[NC][4];
    bool cok[NC][4];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int64_t j = j0 + (c0 + c) * 16 + 4 * g + t;
        cok[c][t] = j < p.N;
        cofs[c][t] = off(p.csC, cok[c][t] ? j : p.N - 1, p.min_csC);
      }
    const bool atomic = p.atomic_c != 0;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int64_t i = i0 + (r0 + r) * 16 + l16;
      const bool rok = i < p.M;
      const Off ro = off(p.rsC, rok ? i : p.M - 1, p.min_rsC);
      auto val = [&](int c, int t) {
        const float re = acc[r][c][0][t], im = acc[r][c][1][t];
        return float2{ar * re - ai * im, ar * im + ai * re};
      };
      if (atomic) {
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int t = 0; t < 4; ++t)
            if (rok && cok[c][t]) {
              const float2 v = val(c, t);
              float* d = reinterpret_cast<float*>(at(ro, cofs[c][t]));
              atomic_add_gmem(d, v.x);
              atomic_add_gmem(d + 1, v.y);
            }
      } else {
        float2 cv[NC][4];
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int t = 0; t < 4; ++t) cv[c][t] = *at(ro, cofs[c][t]);
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int t = 0; t < 4; ++t)
            if (rok && cok[c][t]) {
              const float2 v = val(c, t);
              *at(ro, cofs[c][t]) = float2{cv[c][t].x + v.x, cv[c][t].y + v.y};
            }
      }
    }
  }
};

// Staging with the k offsets read from the global maps (any offsets).
template <typename Idx, bool AF, bool BF, int NR, int NC>
__device__ __forceinline__ void bf16x3_pair_tile(const KArgs<Idx>& p, uint32_t* __restrict__ smem,
                                                 int tm, int tn, int kbeg, int kend, int r0,
                                                 int c0) {
  constexpr int BM = kP144, BN = kP144, BK = 16, NT = 512;
  // smem: [buf][plane][row][16 words]
  auto plane = [&](int buf, int pl) { return smem + (buf * kPPlanes + pl) * kPPlane; };

  const int tid = threadIdx.x;
  const int64_t i0 = int64_t(tm) * BM, j0 = int64_t(tn) * BN;
  const int64_t M = p.M, N = p.N;
  const int K = kend;
  const float2* __restrict__ A = static_cast<const float2*>(p.A);
  const float2* __restrict__ B = static_cast<const float2*>(p.B);
  const unsigned utid = unsigned(tid);

  using GA = std::conditional_t<AF, QuadGather<Idx, BM, BK, NT>,
                                TileGather<float2, Idx, BM, BK, NT, false>>;
  using GB = std::conditional_t<BF, QuadGather<Idx, BN, BK, NT>,
                                TileGather<float2, Idx, BN, BK, NT, false>>;
  GA ga;
  GB gb;
  ga.init(utid, i0, M, p.rsA);
  gb.init(utid, j0, N, p.csB);
  const bool a_inside = i0 + BM <= M, b_inside = j0 + BN <= N;
  const uint32_t mre = (p.conj_a != 0) == (p.conj_b != 0) ? 0x80000000u : 0u;

  // Stage slot u of a gather (values v) into LDS: zero out-of-range entries
  // (edge tiles and the last stage only), split into hi / lo, store (A raw,
  // B in its re-form).
  auto put_quad = [&](const auto& gt, bool is_b, int buf, int p0, int u, bool edge,
                      const float2 (&v)[4]) {
    using G = std::remove_cv_t<std::remove_reference_t<decltype(gt)>>;
    const unsigned e = G::elem(utid, u);
    u32x4 hi, lo;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      float re = v[t].x, im = v[t].y;
      if (edge && !gt.inside(utid, p0, K, u, t)) re = im = 0.f;
      uint32_t h, l;
      split_pair(re, im, h, l);
      hi[t] = is_b ? h ^ mre : h;
      lo[t] = is_b ? l ^ mre : l;
    }
    if (G::valid(utid, u)) {
      const int w = pword(G::xc(e), 4 * G::kq(e));
      *reinterpret_cast<u32x4*>(plane(buf, is_b ? 2 : 0) + w) = hi;
      *reinterpret_cast<u32x4*>(plane(buf, is_b ? 3 : 1) + w) = lo;
    }
  };
  auto put_word = [&](const auto& gt, bool is_b, int buf, int p0, int u, bool edge, float2 v) {
    using G = std::remove_cv_t<std::remove_reference_t<decltype(gt)>>;
    const unsigned e = G::elem(utid, u);
    if (edge) v = gt.finished(utid, p0, K, 1.f, u, v);
    uint32_t hi, lo;
    split_pair(v.x, v.y, hi, lo);
    if (G::valid(utid, u)) {
      const int w = pword(G::xc(e), G::kc(e));
      plane(buf, is_b ? 2 : 0)[w] = is_b ? hi ^ mre : hi;
      plane(buf, is_b ? 3 : 1)[w] = is_b ? lo ^ mre : lo;
    }
  };
  // Gathered values of one stage, between their request and their staging.
  using VA = std::conditional_t<AF, float2[GA::E][4], float2[GA::E]>;
  using VB = std::conditional_t<BF, float2[GB::E][4], float2[GB::E]>;
  VA va;
  VB vb;
  auto load_stage = [&] {
    ga.load_data_to(A, va);
    gb.load_data_to(B, vb);
  };
  auto swrite = [&](int buf, int p0) {
    const bool kin = p0 + BK <= K;
#pragma unroll
    for (int u = 0; u < GA::E; ++u) {
      if constexpr (AF) put_quad(ga, false, buf, p0, u, !(a_inside && kin), va[u]);
      else put_word(ga, false, buf, p0, u, !(a_inside && kin), va[u]);
    }
#pragma unroll
    for (int u = 0; u < GB::E; ++u) {
      if constexpr (BF) put_quad(gb, true, buf, p0, u, !(b_inside && kin), vb[u]);
      else put_word(gb, true, buf, p0, u, !(b_inside && kin), vb[u]);
    }
  };

  PairUnits<Idx, NR, NC> mm;
  mm.init(p, r0, c0);

  // Pipeline (LDS buffers alternate by stage): stage s + 1 is requested at
  // the top of stage s and staged after its products.  Requests past the
  // last stage read clamped, valid offsets.
  const int nst = (K - kbeg + BK - 1) / BK;
  ga.load_k(utid, kbeg, K, p.csA);
  gb.load_k(utid, kbeg, K, p.rsB);
  load_stage();
  ga.load_k(utid, kbeg + BK, K, p.csA);
  gb.load_k(utid, kbeg + BK, K, p.rsB);
  swrite(0, kbeg);
  __syncthreads();
  for (int s = 0; s + 1 < nst; ++s) {
    load_stage();
    ga.load_k(utid, kbeg + (s + 2) * BK, K, p.csA);
    gb.load_k(utid, kbeg + (s + 2) * BK, K, p.rsB);
    mm.compute(plane(s & 1, 0));
    swrite((s + 1) & 1, kbeg + (s + 1) * BK);
    __syncthreads();
  }
  mm.compute(plane((nst - 1) & 1, 0));
  mm.store(p, tm, tn);
}

// Gather slots of one operand for the LDS-map path: offsets are 32-bit
// byte offsets from the operand's rebased base.  QUAD: lanes along the free
// index x, a slot holds the four consecutive k = kc .. kc + 3 of one row
// (one 16-byte LDS write per plane).  Otherwise lanes along k: a slot holds
// one element of row e / 16 at k = tid % 16 (the same k for every slot).
template <bool QUAD, int NT = 512>
struct LdsGather {
  static constexpr int TOTAL = QUAD ? kP144 * 4 : kP144 * 16;
  static constexpr int E = (TOTAL + NT - 1) / NT;
  static constexpr bool RAGGED = (TOTAL % NT) != 0;
  uint32_t xb[E];   // byte offset of the slot's row
  uint32_t xok;     // bit u: slot u's row lies inside the problem

  __device__ static bool valid(unsigned tid, int u) {
    return !RAGGED || u + 1 < E || tid + unsigned(NT) * u < unsigned(TOTAL);
  }
  __device__ static unsigned elem(unsigned tid, int u) {
    return valid(tid, u) ? tid + unsigned(NT) * u : 0u;
  }
  __device__ static int xc(unsigned e) { return QUAD ? int(e % unsigned(kP144)) : int(e / 16u); }
  // First k of slot u within a stage.
  __device__ static int kc(unsigned tid, int u) {
    return QUAD ? 4 * int(elem(tid, u) / unsigned(kP144)) : int(tid & 15u);
  }
  __device__ void init(unsigned tid, int64_t x0, int64_t X, const int32_t* xmap, int64_t xmin) {
    xok = 0;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int64_t xg = x0 + xc(elem(tid, u));
      xb[u] = uint32_t(int64_t(xmap[xg < X ? xg : X - 1]) - xmin) * 8u;
      xok |= ((xg < X && valid(tid, u)) ? 1u : 0u) << u;
    }
  }
};

// k-map entries per operand staged in LDS per launch (the launcher splits
// longer contracted ranges into sequential launches).
constexpr int kKMapMax = 8192;

// Dedicated staging waves for the LDS-map path: kPairProducers waves after
// the kPairConsumers MFMA waves gather, split and stage every stage of A and
// B, two stages ahead of the products (stage s + 1 is staged while the MFMA
// waves multiply stage s; the loads of stage s + 2 are in flight
// meanwhile).  With the staging's VALU work cut down to its core (32-bit
// addresses, LDS k offsets, the dot2 split), the two roles overlap on each
// SIMD instead of alternating phases at every barrier.
constexpr int kPairConsumers = 8;
constexpr int kPairProducers = 4;
constexpr int kPairWsThreads = 64 * (kPairConsumers + kPairProducers);

template <bool AF, bool BF>
__device__ __forceinline__ void pair_ws_producer(const KArgs<int32_t>& p,
                                                 uint32_t* __restrict__ smem,
                                                 const uint32_t* __restrict__ kmA,
                                                 const uint32_t* __restrict__ kmB, int tm, int tn,
                                                 int kbeg, int kend, unsigned ptid) {
  constexpr int BM = kP144, BN = kP144, BK = 16, NT = 64 * kPairProducers;
  using GA = LdsGather<AF, NT>;
  using GB = LdsGather<BF, NT>;
  auto plane = [&](int buf, int pl) { return smem + (buf * kPPlanes + pl) * kPPlane; };
  const int64_t i0 = int64_t(tm) * BM, j0 = int64_t(tn) * BN;
  const int K = kend;
  const int nst = (K - kbeg + BK - 1) / BK;
  const int nfull = (K - kbeg) / BK;   // stages 0 .. nfull - 1 lie wholly inside K
  const char* __restrict__ Ab = static_cast<const char*>(p.A) + 8 * (p.min_rsA + p.min_csA);
  const char* __restrict__ Bb = static_cast<const char*>(p.B) + 8 * (p.min_rsB + p.min_csB);
  GA ga;
  GB gb;
  ga.init(ptid, i0, p.M, p.rsA, p.min_rsA);
  gb.init(ptid, j0, p.N, p.csB, p.min_csB);
  const bool inside = i0 + BM <= p.M && j0 + BN <= p.N;
  const uint32_t mre = (p.conj_a != 0) == (p.conj_b != 0) ? 0x80000000u : 0u;

  using VA = std::conditional_t<AF, float2[GA::E][4], float2[GA::E]>;
  using VB = std::conditional_t<BF, float2[GB::E][4], float2[GB::E]>;
  // Requests stage t (clamped to the last stage: the requests past the end
  // read valid, unused data).
  auto fetch = [&](const char* base, const uint32_t* km, const auto& gt, auto& v, int t) {
    using G = std::remove_cv_t<std::remove_reference_t<decltype(gt)>>;
    const int tt = t < nst ? t : nst - 1;
    if constexpr (std::is_same_v<std::remove_cv_t<std::remove_reference_t<decltype(v[0])>>,
                                 float2>) {
      const uint32_t k1 = km[tt * BK + G::kc(ptid, 0)];
#pragma unroll
      for (int u = 0; u < G::E; ++u)
        v[u] = *reinterpret_cast<const float2*>(base + uint64_t(gt.xb[u] + k1));
    } else {
#pragma unroll
      for (int u = 0; u < G::E; ++u) {
        const u32x4 k4 = *reinterpret_cast<const u32x4*>(km + tt * BK + G::kc(ptid, u));
#pragma unroll
        for (int j = 0; j < 4; ++j)
          v[u][j] = *reinterpret_cast<const float2*>(base + uint64_t(gt.xb[u] + k4[j]));
      }
    }
  };
  auto request = [&](int t, VA& va, VB& vb) {
    fetch(Ab, kmA, ga, va, t);
    fetch(Bb, kmB, gb, vb, t);
  };
  // Stage t's values into LDS buffer buf; EDGE zeroes entries outside the
  // problem (rows of edge tiles, k past the end).
  auto put = [&](const auto& gt, bool is_b, auto edge_tag, int buf, int t, const auto& v) {
    using G = std::remove_cv_t<std::remove_reference_t<decltype(gt)>>;
    constexpr bool EDGE = decltype(edge_tag)::value;
    const int kst = kbeg + t * BK;
    uint32_t* ph = plane(buf, is_b ? 2 : 0);
    uint32_t* pl = plane(buf, is_b ? 3 : 1);
    const uint32_t m = is_b ? mre : 0u;
#pragma unroll
    for (int u = 0; u < G::E; ++u) {
      const unsigned e = G::elem(ptid, u);
      const int x = G::xc(e), k0 = G::kc(ptid, u);
      const bool rok = (gt.xok >> u) & 1u;
      if constexpr (std::is_same_v<std::remove_cv_t<std::remove_reference_t<decltype(v[0])>>,
                                   float2>) {
        float2 w = v[u];
        if (EDGE && !(rok && kst + k0 < K)) w = float2{0.f, 0.f};
        uint32_t h, l;
        split_pair(w.x, w.y, h, l);
        if (G::valid(ptid, u)) {
          const int a = pword(x, k0);
          ph[a] = h ^ m;
          pl[a] = l ^ m;
        }
      } else {
        u32x4 hi, lo;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          float2 w = v[u][j];
          if (EDGE && !(rok && kst + k0 + j < K)) w = float2{0.f, 0.f};
          uint32_t h, l;
          split_pair(w.x, w.y, h, l);
          hi[j] = h ^ m;
          lo[j] = l ^ m;
        }
        if (G::valid(ptid, u)) {   // quads: the 16-row unit layout
          const int a = pword_units(x, k0);
          *reinterpret_cast<u32x4*>(ph + a) = hi;
          *reinterpret_cast<u32x4*>(pl + a) = lo;
        }
      }
    }
  };
  auto swrite = [&](int buf, int t, const VA& va, const VB& vb) {
    if (inside && t < nfull) {
      put(ga, false, std::false_type{}, buf, t, va);
      put(gb, true, std::false_type{}, buf, t, vb);
    } else {
      put(ga, false, std::true_type{}, buf, t, va);
      put(gb, true, std::true_type{}, buf, t, vb);
    }
  };

  // Three register sets where they fit (B gathered along k): the loads of
  // the next two stages are in flight while one stage is staged; else two.
  VA va0, va1, va2;
  VB vb0, vb1, vb2;
  if constexpr (BF) {
    request(0, va0, vb0);
    request(1, va1, vb1);
    swrite(0, 0, va0, vb0);
    __syncthreads();
    // Iteration s stages stage s + 1 and requests stage s + 2 (one barrier
    // each, matching the MFMA waves' loop), unrolled by two so the register
    // sets alternate statically.
    for (int s = 0; s + 1 < nst; s += 2) {
      request(s + 2, va0, vb0);
      swrite(1, s + 1, va1, vb1);
      __syncthreads();
      if (s + 2 >= nst) break;
      request(s + 3, va1, vb1);
      swrite(0, s + 2, va0, vb0);
      __syncthreads();
    }
    return;
  }
  request(0, va0, vb0);
  request(1, va1, vb1);
  request(2, va2, vb2);
  swrite(0, 0, va0, vb0);
  __syncthreads();
  // Iteration s stages stage s + 1 and requests stage s + 3 (one barrier
  // each), unrolled by three so the register sets rotate statically.
  for (int s = 0; s + 1 < nst; s += 3) {
    request(s + 3, va0, vb0);
    swrite((s + 1) & 1, s + 1, va1, vb1);
    __syncthreads();
    if (s + 2 >= nst) break;
    request(s + 4, va1, vb1);
    swrite((s + 2) & 1, s + 2, va2, vb2);
    __syncthreads();
    if (s + 3 >= nst) break;
    request(s + 5, va2, vb2);
    swrite((s + 3) & 1, s + 3, va0, vb0);
    __syncthreads();
  }
}

template <bool AF, bool BF, int NR, int NC>
__device__ __forceinline__ void pair_ws_consumer(const KArgs<int32_t>& p,
                                                 const uint32_t* __restrict__ smem, int tm,
                                                 int tn, int kbeg, int kend, int r0, int c0) {
  PairUnits<int32_t, NR, NC> mm;
  mm.init(p, r0, c0);
  const int nst = (kend - kbeg + 15) / 16;
  __syncthreads();
  for (int s = 0; s + 1 < nst; ++s) {
    mm.template compute<AF, BF>(smem + (s & 1) * kPPlanes * kPPlane);
    __syncthreads();
  }
  mm.template compute<AF, BF>(smem + ((nst - 1) & 1) * kPPlanes * kPPlane);
  mm.template store<true>(p, tm, tn);
}

template <bool AF, bool BF>
__global__ void __launch_bounds__(kPairWsThreads) k_contract_bf16x3_pair_ws(KArgs<int32_t> p) {
  __shared__ __attribute__((aligned(16))) uint32_t smem[2 * kPPlanes * kPPlane];
  __shared__ __attribute__((aligned(16))) uint32_t kmap[2 * kKMapMax];
  int tm, tn, kbeg, kend;
  tile_and_krange(p, tm, tn, kbeg, kend);
  // The k maps of this workgroup's range, as byte offsets from the rebased
  // bases (every wave helps; the loop's first barrier orders them).
  {
    const int K = kend;
    const int n = (K - kbeg + 15) / 16 * 16;
    for (int i = int(threadIdx.x); i < n; i += kPairWsThreads) {
      const int k = kbeg + i < K ? kbeg + i : K - 1;
      kmap[i] = uint32_t(int64_t(p.csA[k]) - p.min_csA) * 8u;
      kmap[kKMapMax + i] = uint32_t(int64_t(p.rsB[k]) - p.min_rsB) * 8u;
    }
  }
  __syncthreads();
  const int wid = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
  if (wid >= kPairConsumers) {
    pair_ws_producer<AF, BF>(p, smem, kmap, kmap + kKMapMax, tm, tn, kbeg, kend,
                             unsigned(threadIdx.x) - 64u * kPairConsumers);
    return;
  }
  const int r0 = kRect144[wid].r0, c0 = kRect144[wid].c0;
  switch (wid) {
    case 2:
    case 3:
      pair_ws_consumer<AF, BF, 4, 3>(p, smem, tm, tn, kbeg, kend, r0, c0);
      break;
    case 6:
      pair_ws_consumer<AF, BF, 9, 1>(p, smem, tm, tn, kbeg, kend, r0, c0);
      break;
    case 7:
      pair_ws_consumer<AF, BF, 4, 2>(p, smem, tm, tn, kbeg, kend, r0, c0);
      break;
    default:
      pair_ws_consumer<AF, BF, 5, 2>(p, smem, tm, tn, kbeg, kend, r0, c0);
      break;
  }
}

template <typename Idx, bool AF, bool BF>
__global__ void __launch_bounds__(512) k_contract_bf16x3_pair(KArgs<Idx> p) {
  __shared__ __attribute__((aligned(16))) uint32_t smem[2 * kPPlanes * kPPlane];
  int tm, tn, kbeg, kend;
  tile_and_krange(p, tm, tn, kbeg, kend);
  const int wid = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
  const int r0 = kRect144[wid & 7].r0, c0 = kRect144[wid & 7].c0;
  switch (wid) {
    case 2:
    case 3:
      bf16x3_pair_tile<Idx, AF, BF, 4, 3>(p, smem, tm, tn, kbeg, kend, r0, c0);
      break;
    case 6:
      bf16x3_pair_tile<Idx, AF, BF, 9, 1>(p, smem, tm, tn, kbeg, kend, r0, c0);
      break;
    case 7:
      bf16x3_pair_tile<Idx, AF, BF, 4, 2>(p, smem, tm, tn, kbeg, kend, r0, c0);
      break;
    default:
      bf16x3_pair_tile<Idx, AF, BF, 5, 2>(p, smem, tm, tn, kbeg, kend, r0, c0);
      break;
  }
}

// 32-bit plans whose operands span less than 2^32 bytes take the LDS-map
// kernel with staging waves, in sequential launches of at most kKMapMax
// contracted indices (each adds its partial products into C; launches on
// one stream run in order); others the global-map kernel.
template <typename Idx>
hipError_t launch_pair144(const GemmProblem& g, hipStream_t s) {
  const bool af = g.a_mfast != 0, bf = g.b_kfast == 0;
  if constexpr (std::is_same_v<Idx, int32_t>) {
    if (g.span32) {
      const dim3 block(kPairWsThreads);
      for (int64_t k0 = 0; k0 < g.K; k0 += kKMapMax) {
        GemmProblem h = g;
        h.K = g.K - k0 < kKMapMax ? g.K - k0 : kKMapMax;
        h.csA = static_cast<const int32_t*>(g.csA) + k0;
        h.rsB = static_cast<const int32_t*>(g.rsB) + k0;
        const KArgs<int32_t> a = make_args<int32_t>(h, kP144, kP144, 16);
        const int64_t per = int64_t(a.tiles_m) * a.tiles_n * a.ksplit;
        const hipError_t e = launch_batches(a, per, kPairWsThreads,
                                            [&](const KArgs<int32_t>& q, unsigned grid) {
          if (af && bf)
            hipLaunchKernelGGL((k_contract_bf16x3_pair_ws<true, true>), grid, block, 0, s, q);
          else if (af)
            hipLaunchKernelGGL((k_contract_bf16x3_pair_ws<true, false>), grid, block, 0, s, q);
          else if (bf)
            hipLaunchKernelGGL((k_contract_bf16x3_pair_ws<false, true>), grid, block, 0, s, q);
          else
            hipLaunchKernelGGL((k_contract_bf16x3_pair_ws<false, false>), grid, block, 0, s, q);
        });
        if (e != hipSuccess) return e;
      }
      return hipSuccess;
    }
  }
  const KArgs<Idx> a = make_args<Idx>(g, kP144, kP144, 16);
  const int64_t per = int64_t(a.tiles_m) * a.tiles_n * a.ksplit;
  const dim3 block(512);
  return launch_batches(a, per, 512, [&](const KArgs<Idx>& q, unsigned grid) {
    if (af && bf)
      hipLaunchKernelGGL((k_contract_bf16x3_pair<Idx, true, true>), grid, block, 0, s, q);
    else if (af)
      hipLaunchKernelGGL((k_contract_bf16x3_pair<Idx, true, false>), grid, block, 0, s, q);
    else if (bf)
      hipLaunchKernelGGL((k_contract_bf16x3_pair<Idx, false, true>), grid, block, 0, s, q);
    else
      hipLaunchKernelGGL((k_contract_bf16x3_pair<Idx, false, false>), grid, block, 0, s, q);
  });
}

template <typename Idx, int BM, int BN, int BK, int WGM, int WGN>
hipError_t launch_t(const GemmProblem& g, hipStream_t s) {
  const KArgs<Idx> a = make_args<Idx>(g, BM, BN, BK);
  const int64_t per = int64_t(a.tiles_m) * a.tiles_n * a.ksplit;
  const dim3 block(64 * WGM * WGN);
  const bool af = g.a_mfast != 0, bf = g.b_kfast == 0;
  return launch_batches(a, per, 64 * WGM * WGN, [&](const KArgs<Idx>& q, unsigned grid) {
    if (af && bf)
      hipLaunchKernelGGL((k_contract_bf16x3<Idx, BM, BN, BK, WGM, WGN, true, true>), grid, block, 0, s, q);
    else if (af)
      hipLaunchKernelGGL((k_contract_bf16x3<Idx, BM, BN, BK, WGM, WGN, true, false>), grid, block, 0, s, q);
    else if (bf)
      hipLaunchKernelGGL((k_contract_bf16x3<Idx, BM, BN, BK, WGM, WGN, false, true>), grid, block, 0, s, q);
    else
      hipLaunchKernelGGL((k_contract_bf16x3<Idx, BM, BN, BK, WGM, WGN, false, false>), grid, block, 0, s, q);
  });
}

template <typename Idx>
hipError_t dispatch(const GemmProblem& g, int32_t kernel, hipStream_t s) {
  if (kernel == K_C64_BF16X3_144) return launch_pair144<Idx>(g, s);
  return launch_t<Idx, 64, 64, 16, 2, 2>(g, s);
}

}  // namespace

hipError_t launch_bf16x3(const GemmProblem& g, int32_t kernel, hipStream_t s) {
  return g.idx32 ? dispatch<int32_t>(g, kernel, s) : dispatch<int64_t>(g, kernel, s);
}

namespace {

template <typename Idx, int BM, int BN, int BK, int WGM, int WGN>
struct BatchedBf {
  using K = void (*)(const KArgs<Idx>*, const BatchItem*, int);
  static K pick(bool af, bool bf) {
    if (af && bf) return k_contract_bf16x3_batched<Idx, BM, BN, BK, WGM, WGN, true, true>;
    if (af) return k_contract_bf16x3_batched<Idx, BM, BN, BK, WGM, WGN, true, false>;
    if (bf) return k_contract_bf16x3_batched<Idx, BM, BN, BK, WGM, WGN, false, true>;
    return k_contract_bf16x3_batched<Idx, BM, BN, BK, WGM, WGN, false, false>;
  }
  static int per_cu(bool af, bool bf) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pick(af, bf), 64 * WGM * WGN, 0) !=
        hipSuccess) {
      (void)hipGetLastError();
      return 1;
    }
    return n > 0 ? n : 1;
  }
  static hipError_t launch(bool af, bool bf, const void* inst, const BatchItem* items,
                           int nitems, int grid, hipStream_t s) {
    hipLaunchKernelGGL(pick(af, bf), dim3(grid), dim3(64 * WGM * WGN), 0, s,
                       static_cast<const KArgs<Idx>*>(inst), items, nitems);
    return hipGetLastError();
  }
};

template <typename Idx, typename F>
auto with_bf_config(int32_t kernel, F&& f) {
  (void)kernel;   // batches run the small tiles only
  return f(BatchedBf<Idx, 64, 64, 16, 2, 2>());
}

}  // namespace

int bf16x3_batched_blocks_per_cu(int32_t kernel, bool idx32, bool af, bool bf) {
  if (idx32) return with_bf_config<int32_t>(kernel, [&](auto c) { return c.per_cu(af, bf); });
  return with_bf_config<int64_t>(kernel, [&](auto c) { return c.per_cu(af, bf); });
}

hipError_t launch_bf16x3_batched(int32_t kernel, bool idx32, bool af, bool bf,
                                 const void* inst, const BatchItem* items, int nitems,
                                 int grid, hipStream_t s) {
  if (idx32)
    return with_bf_config<int32_t>(kernel, [&](auto c) {
      return c.launch(af, bf, inst, items, nitems, grid, s);
    });
  return with_bf_config<int64_t>(kernel, [&](auto c) {
    return c.launch(af, bf, inst, items, nitems, grid, s);
  });
}

}  // namespace t1m_hip
