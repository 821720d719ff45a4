// Transposition-free complex contraction on gfx950 matrix cores (1M method).
//
// The reference executes a complex contraction as a real GEMM on 1M panels:
// A becomes 2x2 real blocks [re -im; im re] and B becomes [re; im]
// (include/t1m/t1m.h:310-337), so C_r (2m x n) = A_r (2m x 2k) . B_r (2k x n)
// costs exactly 8mnk real flops (t1m.h:508-554, BLIS ukernel at 546-547).
//
// Here one workgroup owns a BM x BN tile of complex C and walks the whole K
// range (no packing pass, no workspace):
//   * gather: each lane loads complex A / B elements straight from strided
//     HBM through the scatter vectors (rs/cs as in t1m.h:251-252), lanes
//     along the operand's unit-stride direction, and stores them to LDS as
//     (re, im) pairs -- the real-valued panels of 1M.  Two LDS buffers: the
//     global loads of stage s+1 are in flight while stage s is multiplied.
//   * multiply: v_mfma_f32_16x16x4f32 (FP32) or v_mfma_f64_16x16x4f64 (FP64)
//     in the transposed 1M form: MFMA rows are 16 columns j of C, MFMA columns
//     are 16 real rows of C_r (8 complex rows i, re/im interleaved).  Lane
//     (g = lane/16) reads one complex B[q_g][j] and one complex A[i][q_g]; the
//     first MFMA uses B's real parts with A_r's even real columns
//     (s ? im : re), the second B's imaginary parts with A_r's odd real
//     columns (s ? re : -im), s = row parity.  Two MFMAs therefore cover four
//     complex k.  For FP32 the LDS tile of A holds both parities' operand
//     pairs, (re, -im) and (im, re), written once per element, so the MFMAs
//     read their operands straight from LDS; for FP64 the pair is selected
//     in registers.
//   * scatter epilogue: lane l holds re (even l) or im (odd l) of
//     C[i][j]; C is read-modified-written through rsC/csC (C += alpha*A.B,
//     t1m.h:543-548), so arbitrarily strided C works and rows with unit stride
//     give contiguous 64-byte runs per 16 lanes.  The offsets are loaded
//     first and C is read and written in groups of row blocks (store_tile),
//     atomically when K is split or C's elements may share addresses.
//
// Wave layout: WGM x WGN waves.  A wave owns BM/WGM complex rows (MB blocks
// of 8) and NBW = floor(BN/16 / WGN) blocks of 16 columns.  When BN/16 is
// not a multiple of WGN, the one leftover 16-column block is shared: each
// n-wave takes MB/WGN or one more of its row blocks, reusing the A operands
// it reads for its full blocks (a wave-uniform branch per row block; +0.5
// to +3.5 % over reading them again; the same change in the BF16x3 kernel,
// as a branch or with selects, lost 8-13 %).  For 144 x 144 tiles on 2 x 4
// waves (2 waves per SIMD) a SIMD computes 41 MFMA tiles per k chunk
// against an even share of 40.5.
#include "internal.h"
#include "kernel_common.h"

#include <cstring>
#include <type_traits>

// Design choices below were measured on the 48^4 complex<float> and the
// 1024 x 1024 x 256 / 2304^3 complex<double> contractions (timing variants
// built from switches that no longer exist; results kept with each choice).

namespace t1m_hip {

namespace {

// One BM x BN tile (tm, tn) of one contraction, whole K range.  PREC: load
// C's offsets and values before the main loop (see the epilogue set-up).
// PAIRS (complex<float>): both operands gathered two neighbours per lane
// (PairGather).
template <typename T, typename Idx, int BM, int BN, int BK, int WGM, int WGN,
          bool AF, bool BF, bool PREC, bool PAIRS = false>
__device__ __forceinline__ void contract_tile_1m(const KArgs<Idx>& p, int tm, int tn, int kbeg, int kend) {
  using C2 = typename Vec<T>::c2;
  using Acc = typename Vec<T>::acc;
  constexpr int NT = 64 * WGM * WGN;
  constexpr int WTM = BM / WGM;          // complex rows per wave
  constexpr int MB = WTM / 8;            // 8-row blocks per wave
  constexpr int NBT = BN / 16;           // 16-column blocks per workgroup
  constexpr int NBW = NBT / WGN;         // full column blocks per wave
  constexpr bool XT = NBT % WGN != 0;    // one shared column block
  constexpr int XB = XT ? (MB + WGN - 1) / WGN : 0;
  constexpr int XB_ALLOC = XT ? XB : 1;
  // A pre-expanded in LDS for FP32 (+7.5 %: no per-chunk selects and their
  // MFMA hazard waits); FP64 MFMAs take twice as long and the doubled A tile
  // cost more than the selects (-9 %).
  constexpr bool AEXP = sizeof(T) == 4;
  // Global loads two stages ahead (a second set of data registers) on the
  // 144 x 144 FP32 tiles with 32-bit offsets: on 8 waves +1.2..3.3 % at
  // 2304^3 over the four gather-direction pairs, on 12 waves +1.4..3.2 %
  // although the pair variants then spill 2-12 registers; and on the
  // persistent FP32 batch's 64 x 64 tiles with two-element gathers, whose
  // kernel is bounded to 3 workgroups per CU (154-167 registers, no spills;
  // one-element gathers would spill): MPS sweep forward +2.4 %, VJP +1.5 %.
  // On the single-launch 64 x 64 tiles the extra registers cost a workgroup
  // per CU (-6 %; with the same bound -1.8..+2 %), on FP64 they measured
  // -1..+1 %.
  constexpr bool DEEP = AEXP && (BM == 144 || (PREC && PAIRS)) && sizeof(Idx) == 4;
  // Next stage's LDS writes in the middle of the stage; 64-bit offsets on
  // the 144 x 144 tiles leave no registers for them.  The persistent FP32
  // batch (PREC, with the two-stage prefetch above) writes at the end of the
  // stage: equal on short K, +1.2 % on the long-K VJP launches of a 512-item
  // MPS sweep.
  constexpr bool MIDW = (sizeof(Idx) == 4 || BM * BN < 144 * 144) && !(PREC && DEEP);
  // LDS layouts.  A holds each complex element twice, pre-expanded for the
  // two row parities of 1M: (re, -im) for even lanes, (im, re) for odd ones,
  // so the MFMA operands come straight from the LDS reads.  Both tiles are
  // XOR-swizzled by k: element (k, x) sits in slot x ^ sw(k), sw(k) =
  // (k/2) mod (half a bank sweep), and the row pitch is an odd multiple of
  // 128 bytes.  A 16-lane read group (8 A units or 16 B units, aligned) stays
  // one contiguous bank sweep, and a gather along k writes 256 consecutive
  // bytes of banks per 256 bytes stored.
  struct alignas(2 * sizeof(C2)) AExp {
    C2 e, o;
  };
  using AU = std::conditional_t<AEXP, AExp, C2>;
  constexpr int SA = 256 / int(sizeof(AU));   // units per 256-byte bank sweep
  constexpr int SB = 256 / int(sizeof(C2));
  // FP64 (16-byte units): a 16-byte read by 64 lanes is conflict-free on
  // gfx950 when lane j of each 16-lane group uses the same 16-byte bank slot
  // ((address / 16) mod 16), i.e. when the groups' rows start at the same
  // bank slot and share their swizzle.  So A, gathered along its rows (AF),
  // gets a row pitch of a multiple of 256 bytes (no swizzle: its writes are
  // whole rows); B, when gathered along its columns, the same pitch with a
  // swizzle that is one value for the four k rows a read covers,
  // sw(k) = (k/4) mod 8.  A B gathered along k writes columns, which that
  // layout serves worse than the pitch-and-swizzle above; F64_BK picks its
  // layout.
  constexpr bool F64_BK = false;
  // (64 x 64 tiles only: the larger pitches would cost the 32 x 32 and thin
  // tiles a workgroup per CU of LDS)
  constexpr bool F64L = !AEXP && BM == 64 && BN == 64;
  constexpr bool F64A = F64L && AF;
  constexpr bool F64B = F64L && (BF || F64_BK);
  constexpr int PADA = AEXP ? ((SA / 2 - BM % SA) % SA + SA) % SA : F64A ? 16 : 4;
  constexpr int PADB = F64B ? 16 : ((SB / 2 - BN % SB) + SB) % SB;
  constexpr int LDA = BM + PADA, LDB = BN + PADB;
  static_assert(!AEXP || (LDA * int(sizeof(AU))) % 256 == 128, "A pitch");
  static_assert(!F64A || (LDA * int(sizeof(AU))) % 256 == 0, "A pitch (FP64)");
  static_assert(F64B ? (LDB * int(sizeof(C2))) % 256 == 0
                     : (LDB * int(sizeof(C2))) % 256 == 128, "B pitch");
  static_assert(WTM % 8 == 0 && BN % 16 == 0, "tile shape");
  static_assert(NBT - NBW * WGN <= 1, "at most one shared column block");
  static_assert(BK % 4 == 0, "k chunk");
  // FP32 64 x 64 tile, 8-k stages, B gathered along k in pairs: lane e
  // writes (k, k+1) of column e / 4 for k = 2 (e mod 4), so the 32 lanes of
  // a half-wave write 8 columns of four even (then odd) rows.  A swizzle by
  // whole 64-byte segments per k pair, and odd k in the other 32-byte half
  // of its segment, spreads them over the bank sweep: 3.96 conflict cycles
  // per write instead of 23.5 with the (k/2) mod 16 swizzle (reads of both
  // layouts: none).  The swizzle then crosses 16-column blocks, so the
  // fragment reads swizzle the full column within its 32-column pair.
  constexpr bool BKP = PAIRS && !BF && AEXP && BN == 64 && BK == 8;
  // A read group spans 8 aligned units, a B group 16.
  auto swa = [](int k) { return AEXP ? (k >> 1) & (SA / 2 - 1 < 7 ? SA / 2 - 1 : 7) : 0; };
  auto swb = [](int k) {
    if constexpr (BKP) return ((k >> 1) & 3) * 8 + (k & 1) * 4;
    return (k >> (F64B ? 2 : 1)) & (SB / 2 - 1 < 15 ? SB / 2 - 1 : 15);
  };
  static_assert(!BKP || (SB == 32 && LDB >= BN), "B k-pair swizzle: 32-unit bank sweep");

  __shared__ AU sA[2][BK][LDA];
  __shared__ C2 sB[2][BK][LDB];

  const int tid = threadIdx.x;
  // With a shared column block the wave index is made scalar, so the
  // per-row-block test below is a scalar branch rather than an exec-mask one
  // (+0.6..0.8 % at 144 x 144; on the 64 x 64 tiles, which have no such
  // test, the scalar index measured -1 %).
  const int lane = tid & 63;
  const int wid = XT ? __builtin_amdgcn_readfirstlane(tid >> 6) : tid >> 6;
  const int wm = wid % WGM, wn = wid / WGM;
  // Shared column block: n-wave wn takes row blocks [xlo, xhi), MB / WGN
  // each and one more for the first MB % WGN waves.  Waves wn and wn + 2
  // share a SIMD (wave id = wm + WGM wn, WGM = 2), so the extra blocks land
  // on different SIMDs: 41 + 41 + 40 + 40 MFMA tiles per SIMD pair at
  // 144 x 144, where ceil(MB / WGN) per wave gave 42 + 42 + 39 + 39.  A
  // block mb sits in accx[mb % XB] (a range of at most XB blocks has
  // distinct residues).
  const int xlo = wn * (MB / WGN) + (wn < MB % WGN ? wn : MB % WGN);
  const int xhi = xlo + MB / WGN + (wn < MB % WGN ? 1 : 0);
  const int64_t i0 = int64_t(tm) * BM, j0 = int64_t(tn) * BN;
  const int64_t M = p.M, N = p.N;
  const int K = kend;     // contracted range [kbeg, kend) of this workgroup

  const C2* __restrict__ A = static_cast<const C2*>(p.A);
  const C2* __restrict__ B = static_cast<const C2*>(p.B);

  // ---- gathers: A tile (BM x BK), B tile (BN x BK) -----------------------
  const unsigned utid = unsigned(tid);
  static_assert(!PAIRS || sizeof(T) == 4, "pairs of complex<float>");
  using GA = std::conditional_t<PAIRS, PairGather<Idx, BM, BK, NT, AF>,
                                TileGather<C2, Idx, BM, BK, NT, AF>>;
  using GB = std::conditional_t<PAIRS, PairGather<Idx, BN, BK, NT, BF>,
                                TileGather<C2, Idx, BN, BK, NT, BF>>;
  GA ga;
  GB gb;
  ga.init(utid, i0, M, p.rsA);
  gb.init(utid, j0, N, p.csB);
  const T sa = p.conj_a ? T(-1) : T(1);
  const T sb = p.conj_b ? T(-1) : T(1);

  // Data registers of a stage: one set, or two when the loads run two
  // stages ahead (DEEP).
  using VA = typename GA::V[GA::E];
  using VB = typename GB::V[GB::E];
  auto swrite = [&](int buf, int p0, const VA& va, const VB& vb) {
#pragma unroll
    for (int u = 0; u < GA::E; ++u) {
      if (GA::valid(utid, u)) {
#pragma unroll
        for (int j = 0; j < GA::PER; ++j) {
          const int k = GA::kcj(utid, u, j), x = GA::xcj(utid, u, j);
          const C2 v = ga.finished(utid, p0, K, sa, u, j, va[u]);
          if constexpr (AEXP) {
            AExp w;
            w.e.x = v.x; w.e.y = -v.y;
            w.o.x = v.y; w.o.y = v.x;
            sA[buf][k][x ^ swa(k)] = w;
          } else {
            sA[buf][k][x] = v;
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < GB::E; ++u) {
      if (GB::valid(utid, u)) {
#pragma unroll
        for (int j = 0; j < GB::PER; ++j) {
          const int k = GB::kcj(utid, u, j), x = GB::xcj(utid, u, j);
          sB[buf][k][x ^ swb(k)] = gb.finished(utid, p0, K, sb, u, j, vb[u]);
        }
      }
    }
  };

  Acc acc[MB][NBW];
  Acc accx[XB_ALLOC];
#pragma unroll
  for (int a = 0; a < MB; ++a)
#pragma unroll
    for (int b = 0; b < NBW; ++b) acc[a][b] = Acc{0, 0, 0, 0};
#pragma unroll
  for (int x = 0; x < XB_ALLOC; ++x) accx[x] = Acc{0, 0, 0, 0};

  const int g = lane >> 4;           // k sub-index inside a 4-k chunk
  const int l16 = lane & 15;
  const bool odd = (lane & 1) != 0;  // real row parity (re / im of C)
  const int am0 = wm * WTM + (l16 >> 1);
  const int bn0 = wn * NBW * 16 + l16;
  const int bnx = NBW * WGN * 16 + l16;

  // A operands: the two 1M operands of this lane's row parity (AEXP) or
  // the raw (re, im) pair.
  auto aread = [&](int buf, int q, int row0) -> C2 {
    const int x = row0 + (((l16 >> 1) ^ swa(q)) - (l16 >> 1));
    if constexpr (AEXP) {
      const C2* u = reinterpret_cast<const C2*>(&sA[buf][q][x]);
      return u[odd ? 1 : 0];
    } else {
      return sA[buf][q][x];
    }
  };
  // Two MFMAs per tile: B's real parts against A_r's even real columns,
  // B's imaginary parts against the odd.
  auto operands = [&](const C2& a, T& re, T& im) {
    if constexpr (AEXP) {
      re = a.x;
      im = a.y;
    } else {
      re = odd ? a.y : a.x;    // A_r[2i+s][2q]
      im = odd ? a.x : -a.y;   // A_r[2i+s][2q+1]
    }
  };
  // Operand pipeline: B fragments one 4-k chunk ahead, A operands LA row
  // groups ahead of the MFMAs that use them (a rolling window over the
  // KCH x NG (chunk, row group) steps of a stage; the scheduling barriers
  // keep the compiler from sinking the LDS reads).  Loading whole chunks of
  // fragments ahead needed far more registers for the same speed; LA from 1
  // to 6 measured alike.  `mid` (the next stage's LDS writes) runs after
  // half the stage's MFMAs.  The wave's priority is raised over the MFMA
  // stream and lowered for the writes (+1-2 %).
  constexpr int KCH = BK / 4;
  constexpr int NG = MB;                     // row groups per chunk
  constexpr int TOT = KCH * NG;
  constexpr int LA = 2;
  struct BFr {
    C2 b[NBW];
    C2 bx;
  };
  auto load_b = [&](int buf, int kc, BFr& f) {
    const int q = kc * 4 + g;
    if constexpr (BKP) {
      static_assert(NBW % 2 == 0 && !XT, "32-column pairs per wave");
      const int s = swb(q), b0 = bn0 - l16;   // b0: a multiple of 32
#pragma unroll
      for (int nb = 0; nb < NBW; ++nb)
        f.b[nb] = sB[buf][q][b0 + (nb & ~1) * 16 + (((nb & 1) * 16 + l16) ^ s)];
      return;
    }
    const int bsw = (l16 ^ swb(q)) - l16;
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb) f.b[nb] = sB[buf][q][bn0 + bsw + nb * 16];
    if constexpr (XT) f.bx = sB[buf][q][bnx + bsw];
  };
  auto load_a = [&](int buf, int t) -> C2 {
    const int kc = t / NG, grp = t % NG;
    return aread(buf, kc * 4 + g, am0 + grp * 8);
  };
  // The shared column block reuses the A operands of this wave's share of
  // row blocks (xlo <= grp < xhi, a wave-uniform branch) instead of reading
  // them again.
  auto mma_sel = [&](int grp, T re, T im, const BFr& f) {
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb) {
      acc[grp][nb] = mfma16(f.b[nb].x, re, acc[grp][nb]);
      acc[grp][nb] = mfma16(f.b[nb].y, im, acc[grp][nb]);
    }
    if constexpr (XT) {
      if (grp >= xlo && grp < xhi) {
        Acc& ax = accx[grp % XB];
        ax = mfma16(f.bx.x, re, ax);
        ax = mfma16(f.bx.y, im, ax);
      }
    }
  };
  // FP64 (A not pre-expanded): the row-parity selects of step t+1 run
  // before the MFMAs of step t, so no MFMA waits on a just-written operand.
  auto compute_fine = [&](int buf, auto&& mid) {
    BFr fb[2];
    C2 aw[LA + 1];
    T sre[2], sim[2];
    __builtin_amdgcn_s_setprio(1);
    load_b(buf, 0, fb[0]);
#pragma unroll
    for (int t = 0; t < LA; ++t) aw[t] = load_a(buf, t);
    if constexpr (!AEXP) operands(aw[0], sre[0], sim[0]);
#pragma unroll
    for (int t = 0; t < TOT; ++t) {
      const int kc = t / NG, grp = t % NG;
      if (grp == 0 && kc + 1 < KCH) load_b(buf, kc + 1, fb[(kc + 1) & 1]);
      if (t + LA < TOT) aw[(t + LA) % (LA + 1)] = load_a(buf, t + LA);
      if constexpr (!AEXP) {
        if (t + 1 < TOT) operands(aw[(t + 1) % (LA + 1)], sre[(t + 1) & 1], sim[(t + 1) & 1]);
      }
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (AEXP) {
        T re, im;
        operands(aw[t % (LA + 1)], re, im);
        mma_sel(grp, re, im, fb[kc & 1]);
      } else {
        mma_sel(grp, sre[t & 1], sim[t & 1], fb[kc & 1]);
      }
      __builtin_amdgcn_sched_barrier(0);
      if (MIDW && t == TOT / 2 - 1) {
        __builtin_amdgcn_s_setprio(0);
        mid();
        __builtin_amdgcn_s_setprio(1);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    __builtin_amdgcn_s_setprio(0);
    if (!MIDW) mid();
  };
  auto compute = [&](int buf, auto&& mid) { compute_fine(buf, mid); };

  // ---- epilogue set-up: C[rsC[i] + csC[j]] += alpha * acc ----------------
  // This lane's outputs: complex rows i (one per row block; re or im by
  // lane parity) times columns j (4 per 16-column block).  All row and
  // column offsets are loaded first (those of missing outputs are 0, a
  // valid element), then C is updated in groups (store_tile).  With PREC
  // the offsets and the C values are loaded before the main loop and stay
  // in flight behind it (C's elements belong to this workgroup alone unless
  // the adds are atomic, and then nothing is prefetched).  Measured: +6 % on
  // the persistent FP32 batch (short K ranges, one item after another), no
  // gain for the single-launch FP32 64 x 64 and FP64 tiles, and the 144 x 144
  // tiles have no registers for it.
  T* __restrict__ Cr = static_cast<T*>(p.C);
  const bool cplx_alpha = p.alpha_im != 0.0;
  const T ar = T(p.alpha_re), ai = T(p.alpha_im);
  constexpr int NR = MB + (XT ? XB : 0);        // row blocks (shared ones last)
  constexpr int NC = NBW + (XT ? 1 : 0);        // column blocks (shared one last)
  Idx rofs[NR];
  bool rok[NR];
  Idx cofs[NC][4];
  bool cok[NC][4];
  auto offsets = [&] {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      // shared slot x = r - MB holds block xlo + ((x - xlo) mod XB), if < xhi
      const int blk = r < MB ? r : xlo + ((r - MB) - xlo % XB_ALLOC + XB_ALLOC) % XB_ALLOC;
      const int64_t i = i0 + am0 + (blk < MB ? blk : MB - 1) * 8;
      rok[r] = i < M && (r < MB || blk < xhi);
      rofs[r] = rok[r] ? gmem(p.rsC)[i] : Idx(0);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        // MFMA output row of (lane, q): 4(lane/16) + q for the f32 16x16x4
        // form, lane/16 + 4q for the f64 one.
        const int64_t jb = c < NBW ? j0 + wn * NBW * 16 + c * 16 : j0 + NBW * WGN * 16;
        const int64_t j = jb + (sizeof(T) == 4 ? 4 * g + q : g + 4 * q);
        cok[c][q] = j < N;
        cofs[c][q] = cok[c][q] ? gmem(p.csC)[j] : Idx(0);
      }
  };
  auto scaled = [&](T v) {
    if (cplx_alpha) {
      const T other = swap_pair(v);     // the neighbour lane holds the other part
      return odd ? ar * v + ai * other : ar * v - ai * other;
    }
    return ar * v;
  };
  auto value = [&](int r, int c, int q) -> T {
    if constexpr (XT) {
      if (r >= MB) return scaled(accx[r - MB][q]);
    }
    return scaled(acc[r][c][q]);
  };
  // (row block, column block) pairs this lane holds: full blocks x full
  // columns, and the shared row blocks x the shared column.
  auto has = [](int r, int c) { return r < MB ? c < NBW : c == NC - 1; };
  auto ok = [&](int r, int c, int q) { return rok[r] && cok[c][q]; };
  auto addr = [&](int r, int c, int q) {
    return Cr + 2 * int64_t(rofs[r] + cofs[c][q]) + (odd ? 1 : 0);
  };

  // ---- main loop: double-buffered LDS, one barrier per stage ------------
  // Iteration s multiplies stage s from LDS while the data loads of stage
  // s+1 and the k-offset loads of stage s+2 are in flight.
  const bool atomic = p.atomic_c != 0;
  T cv[NR][NC][4];
  if constexpr (PREC) offsets();
  const int nst = (K - kbeg + BK - 1) / BK;
  ga.load_k(utid, kbeg, K, p.csA);
  gb.load_k(utid, kbeg, K, p.rsB);
  ga.load_data(A);
  gb.load_data(B);
  ga.load_k(utid, kbeg + BK, K, p.csA);
  gb.load_k(utid, kbeg + BK, K, p.rsB);
  if constexpr (DEEP) {
    // Two stages of data in flight: iteration s loads stage s + 2 into the
    // set that stage s left, while the other set (stage s + 1) is written
    // to LDS in the middle of the stage.  Loads past the last stage read
    // clamped, valid k offsets and are never written.
    VA va1;
    VB vb1;
    ga.load_data_to(A, va1);
    gb.load_data_to(B, vb1);
    ga.load_k(utid, kbeg + 2 * BK, K, p.csA);
    gb.load_k(utid, kbeg + 2 * BK, K, p.rsB);
    if constexpr (PREC) {
      if (!atomic) load_c_tile<T, NR, NC>(has, addr, cv);
    }
    swrite(0, kbeg, ga.val, gb.val);
    __syncthreads();
    auto body = [&](int s, VA& la, VB& lb, const VA& wa, const VB& wb) {
      ga.load_data_to(A, la);
      gb.load_data_to(B, lb);
      ga.load_k(utid, kbeg + (s + 3) * BK, K, p.csA);
      gb.load_k(utid, kbeg + (s + 3) * BK, K, p.rsB);
      compute(s & 1, [&] { swrite((s + 1) & 1, kbeg + (s + 1) * BK, wa, wb); });
      __syncthreads();
    };
    int s = 0;
    for (; s + 2 < nst; s += 2) {
      body(s, ga.val, gb.val, va1, vb1);
      body(s + 1, va1, vb1, ga.val, gb.val);
    }
    if (s + 1 < nst) {
      body(s, ga.val, gb.val, va1, vb1);
      ++s;
    }
    compute(s & 1, [] {});
  } else {
    if constexpr (PREC) {
      if (!atomic) load_c_tile<T, NR, NC>(has, addr, cv);
    }
    swrite(0, kbeg, ga.val, gb.val);
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
      compute(s & 1, [&] { swrite((s + 1) & 1, kbeg + (s + 1) * BK, ga.val, gb.val); });
      __syncthreads();
    }
    compute((nst - 1) & 1, [] {});
  }

  if constexpr (PREC) {
    if (atomic)
      store_tile<T, NR, NC, NR>(true, has, ok, value, addr);
    else
      add_store_c_tile<T, NR, NC>(has, ok, value, addr, cv);
  } else {
    offsets();
    store_tile<T, NR, NC, 3>(atomic, has, ok, value, addr);
  }
}

template <typename T, typename Idx, int BM, int BN, int BK, int WGM, int WGN,
          bool AF, bool BF, bool PAIRS>
__global__ void __launch_bounds__(64 * WGM * WGN)
k_contract_1m(KArgs<Idx> p) {
  int tm, tn, kbeg, kend;
  tile_and_krange(p, tm, tn, kbeg, kend);
  contract_tile_1m<T, Idx, BM, BN, BK, WGM, WGN, AF, BF, false, PAIRS>(p, tm, tn, kbeg, kend);
}

// Persistent batched form over a work queue.  The work list holds ndom
// interleaved per-domain lists (entry k * ndom + d of domain d, each sorted by
// decreasing work, padded with instance -1); workgroup b serves domain
// b mod ndom, the L2 domain (XCD) it runs on, and takes that list's entries
// in order through the counter tickets[d], so a workgroup whose tiles ran
// short takes more of them and the grid (a few workgroups per CU) drains
// the list in LPT order in one launch.  Lane 0 takes the next ticket while
// the current tile runs.  tickets[ndom] counts finished workgroups; the last
// one clears every counter, so the next launch (a graph replay) starts from
// zero without a memset in the graph.  Launches of one list must not
// overlap.
template <typename T, typename Idx, int BM, int BN, int BK, int WGM, int WGN,
          bool AF, bool BF, bool PAIRS>
// (3 workgroups per CU on the FP32 64 x 64 tiles with 32-bit offsets: the
// two-stage prefetch must not cost a workgroup)
__global__ void __launch_bounds__(64 * WGM * WGN,
                                  (sizeof(T) == 4 && BM < 144 && sizeof(Idx) == 4) ? 3 : 1)
k_contract_1m_batched(const KArgs<Idx>* __restrict__ inst, const BatchItem* __restrict__ items,
                      int nitems, unsigned* __restrict__ tickets, int ndom) {
  __shared__ int s_ticket[2];
  const int d = int(blockIdx.x) % ndom;
  const int per = nitems / ndom;             // entries of each domain's list
  if (threadIdx.x == 0) s_ticket[0] = int(atomicAdd(&tickets[d], 1u));
  __syncthreads();
  int k = s_ticket[0];
  for (int it = 1; k < per; ++it) {
    unsigned next = 0;
    if (threadIdx.x == 0) next = atomicAdd(&tickets[d], 1u);
    const BatchItem w = items[int64_t(k) * ndom + d];
    if (w.inst >= 0) {   // not padding
      const KArgs<Idx> p = inst[w.inst];
      // C is prefetched before the main loop on the small FP32 tiles; the
      // 144 x 144 tiles have no registers for it.
      contract_tile_1m<T, Idx, BM, BN, BK, WGM, WGN, AF, BF, sizeof(T) == 4 && BM < 144,
                       PAIRS>(p, w.tm, w.tn, 0, int(p.K));
    }
    // Tickets alternate between two slots, so one barrier both publishes the
    // next ticket and ends this tile's LDS use: every wave read slot it & 1
    // (two items ago) before the previous item's closing barrier.
    if (threadIdx.x == 0) s_ticket[it & 1] = int(next);
    __syncthreads();
    k = s_ticket[it & 1];
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&tickets[ndom], 1u) == gridDim.x - 1)
      for (int i = 0; i <= ndom; ++i) atomicExch(&tickets[i], 0u);
  }
}

// One lane per output element: a plain loop over the contracted index.
// Works for every dtype; used as the cross-check path and for problems the
// MFMA tiles do not cover (real dtypes).
template <typename T, bool kComplex, typename Idx>
__global__ void __launch_bounds__(256) k_contract_simple(KArgs<Idx> p) {
  const int64_t per = (p.M * p.N + 255) / 256;       // blocks per batch problem
  const int z = int(int64_t(blockIdx.x) / per);
  batch_shift(p, z);
  const int64_t e = (int64_t(blockIdx.x) - z * per) * blockDim.x + threadIdx.x;
  if (e >= p.M * p.N) return;
  const int64_t i = e % p.M, j = e / p.M;
  const Idx ra = p.rsA[i], cb = p.csB[j];
  const T* A = static_cast<const T*>(p.A);
  const T* B = static_cast<const T*>(p.B);
  T* C = static_cast<T*>(p.C);
  if constexpr (kComplex) {
    const T sa = p.conj_a ? T(-1) : T(1), sb = p.conj_b ? T(-1) : T(1);
    T re = 0, im = 0;
    for (int64_t q = 0; q < p.K; ++q) {
      const int64_t oa = 2 * int64_t(ra + p.csA[q]);
      const int64_t ob = 2 * int64_t(p.rsB[q] + cb);
      const T xr = A[oa], xi = sa * A[oa + 1];
      const T yr = B[ob], yi = sb * B[ob + 1];
      re = fma(xr, yr, re);
      re = fma(-xi, yi, re);
      im = fma(xr, yi, im);
      im = fma(xi, yr, im);
    }
    const T ar = T(p.alpha_re), ai = T(p.alpha_im);
    const int64_t oc = 2 * int64_t(p.rsC[i] + p.csC[j]);
    accumulate(C + oc, ar * re - ai * im, p.atomic_c != 0);
    accumulate(C + oc + 1, ar * im + ai * re, p.atomic_c != 0);
  } else {
    T acc = 0;
    for (int64_t q = 0; q < p.K; ++q)
      acc = fma(A[ra + p.csA[q]], B[p.rsB[q] + cb], acc);
    accumulate(C + int64_t(p.rsC[i] + p.csC[j]), T(p.alpha_re) * acc, p.atomic_c != 0);
  }
}

template <typename Idx>
hipError_t launch_simple(const GemmProblem& g, hipStream_t s) {
  const KArgs<Idx> a = make_args<Idx>(g, 0, 0);
  const int64_t n = g.M * g.N;
  return launch_batches(a, (n + 255) / 256, 256, [&](const KArgs<Idx>& q, unsigned grid) {
    switch (g.dtype) {
      case T1M_FLOAT32: hipLaunchKernelGGL((k_contract_simple<float, false, Idx>), grid, 256, 0, s, q); break;
      case T1M_FLOAT64: hipLaunchKernelGGL((k_contract_simple<double, false, Idx>), grid, 256, 0, s, q); break;
      case T1M_COMPLEX64: hipLaunchKernelGGL((k_contract_simple<float, true, Idx>), grid, 256, 0, s, q); break;
      default: hipLaunchKernelGGL((k_contract_simple<double, true, Idx>), grid, 256, 0, s, q); break;
    }
  });
}

// The kernel for the gather directions (and, with ALLOW_PAIRS, two-element
// gathers when both operands allow them).
template <typename T, typename Idx, int BM, int BN, int BK, int WGM, int WGN, bool ALLOW_PAIRS>
using Kernel1M = void (*)(KArgs<Idx>);
template <typename T, typename Idx, int BM, int BN, int BK, int WGM, int WGN, bool ALLOW_PAIRS>
Kernel1M<T, Idx, BM, BN, BK, WGM, WGN, ALLOW_PAIRS> pick_1m(bool af, bool bf, bool pairs) {
  if constexpr (ALLOW_PAIRS) {
    if (pairs) {
      if (af && bf) return k_contract_1m<T, Idx, BM, BN, BK, WGM, WGN, true, true, true>;
      if (af) return k_contract_1m<T, Idx, BM, BN, BK, WGM, WGN, true, false, true>;
      if (bf) return k_contract_1m<T, Idx, BM, BN, BK, WGM, WGN, false, true, true>;
      return k_contract_1m<T, Idx, BM, BN, BK, WGM, WGN, false, false, true>;
    }
  }
  if (af && bf) return k_contract_1m<T, Idx, BM, BN, BK, WGM, WGN, true, true, false>;
  if (af) return k_contract_1m<T, Idx, BM, BN, BK, WGM, WGN, true, false, false>;
  if (bf) return k_contract_1m<T, Idx, BM, BN, BK, WGM, WGN, false, true, false>;
  return k_contract_1m<T, Idx, BM, BN, BK, WGM, WGN, false, false, false>;
}

template <typename T, typename Idx, int BM, int BN, int BK, int WGM, int WGN,
          bool ALLOW_PAIRS = false>
hipError_t launch_1m(const GemmProblem& g, hipStream_t s) {
  const KArgs<Idx> a = make_args<Idx>(g, BM, BN, BK);
  const int64_t per = int64_t(a.tiles_m) * a.tiles_n * a.ksplit;
  const dim3 block(64 * WGM * WGN);
  const auto kern = pick_1m<T, Idx, BM, BN, BK, WGM, WGN, ALLOW_PAIRS>(
      g.a_mfast != 0, g.b_kfast == 0, g.a_pairs && g.b_pairs);
  return launch_batches(a, per, 64 * WGM * WGN, [&](const KArgs<Idx>& q, unsigned grid) {
    hipLaunchKernelGGL(kern, grid, block, 0, s, q);
  });
}

template <typename Idx>
hipError_t dispatch(const GemmProblem& g, int32_t kernel, hipStream_t s) {
  switch (kernel) {
    case K_C64_F32_144:
      // 32-bit offsets: 12 waves (3 x 4, 48 x 36 per wave, 3 per SIMD) with
      // the two-stage prefetch, although its pair variants spill 2-12
      // registers: at 48^4 abcd,dbef / bdac,dbef / abcd,ebdf / bdac,ebdf
      // 794 / 850 / 793 / 803 us on 8 waves (2 x 4) -> 779 / 775 / 770 / 779
      // (without the prefetch 12 waves were 795 / 786 / 795 / 803).  64-bit
      // offsets keep 8 waves.
      if constexpr (sizeof(Idx) == 4)
        return launch_1m<float, Idx, 144, 144, 16, 3, 4, true>(g, s);
      else
        return launch_1m<float, Idx, 144, 144, 16, 2, 4, true>(g, s);
    // 8-k stages: 3 workgroups per CU (+5 % and +13 % on two 48^4 / 24^4
    // layouts, -1 % on a third).
    case K_C64_F32_64: return launch_1m<float, Idx, 64, 64, 8, 2, 2, true>(g, s);
    // 4 x 2 waves (16 x 32 complex per wave, 2 waves per SIMD): +23 % over
    // 2 x 2 at 1024 x 1024 x 256, +3 % at 2304^3.
    case K_C128_F64_64: return launch_1m<double, Idx, 64, 64, 16, 4, 2>(g, s);
    // 8 x 2 waves (4 per SIMD, 8 x 32 complex per wave): +2..3 % at
    // 1024 x 1024 x 256, -3..-10 % on the reference suite's other
    // complex<double> contractions -- an autotuning candidate only.
    case K_C128_F64_64W: return launch_1m<double, Idx, 64, 64, 16, 8, 2>(g, s);
    case K_C128_F64_32: return launch_1m<double, Idx, 32, 32, 16, 2, 2>(g, s);
    // Thin C (one 16-column block): 4 waves stacked along m.
    case K_C64_F32_THIN: return launch_1m<float, Idx, 128, 16, 16, 4, 1>(g, s);
    case K_C128_F64_THIN: return launch_1m<double, Idx, 128, 16, 16, 4, 1>(g, s);
    default: return launch_simple<Idx>(g, s);
  }
}

const KernelInfo kInfo[K_COUNT] = {
    {"simple", 0, 0},
    {"c64_f32mfma_144x144", 144, 144},
    {"c64_f32mfma_64x64", 64, 64},
    {"c128_f64mfma_64x64", 64, 64},
    {"c128_f64mfma_32x32", 32, 32},
    {"c64_bf16x3_144x144", 144, 144},
    {"c64_bf16x3_64x64", 64, 64},
    {"f32mfma_128x128", 128, 128},
    {"f32mfma_64x64", 64, 64},
    {"f64mfma_64x64", 64, 64},
    {"f64mfma_32x32", 32, 32},
    {"c64_f32mfma_128x16", 128, 16},
    {"c128_f64mfma_128x16", 128, 16},
    {"c128_f64mfma_64x64w16", 64, 64},
};

}  // namespace

hipError_t launch_bf16x3(const GemmProblem& g, int32_t kernel, hipStream_t s);
bool is_real_kernel(int32_t kernel);
hipError_t launch_real(const GemmProblem& g, int32_t kernel, hipStream_t s);
int real_batched_blocks_per_cu(int32_t kernel, bool idx32, bool af, bool bf);
hipError_t launch_real_batched(int32_t kernel, bool idx32, bool af, bool bf, const void* inst,
                               const BatchItem* items, int nitems, int grid, hipStream_t s);
int bf16x3_batched_blocks_per_cu(int32_t kernel, bool idx32, bool af, bool bf);
hipError_t launch_bf16x3_batched(int32_t kernel, bool idx32, bool af, bool bf,
                                 const void* inst, const BatchItem* items, int nitems,
                                 int grid, hipStream_t s);

namespace {

template <typename T, typename Idx, int BM, int BN, int BK, int WGM, int WGN,
          bool ALLOW_PAIRS = false>
struct Batched1M {
  using K = void (*)(const KArgs<Idx>*, const BatchItem*, int, unsigned*, int);
  static K pick(bool af, bool bf, bool pairs) {
    if constexpr (ALLOW_PAIRS) {
      if (pairs) {
        if (af && bf) return k_contract_1m_batched<T, Idx, BM, BN, BK, WGM, WGN, true, true, true>;
        if (af) return k_contract_1m_batched<T, Idx, BM, BN, BK, WGM, WGN, true, false, true>;
        if (bf) return k_contract_1m_batched<T, Idx, BM, BN, BK, WGM, WGN, false, true, true>;
        return k_contract_1m_batched<T, Idx, BM, BN, BK, WGM, WGN, false, false, true>;
      }
    }
    if (af && bf) return k_contract_1m_batched<T, Idx, BM, BN, BK, WGM, WGN, true, true, false>;
    if (af) return k_contract_1m_batched<T, Idx, BM, BN, BK, WGM, WGN, true, false, false>;
    if (bf) return k_contract_1m_batched<T, Idx, BM, BN, BK, WGM, WGN, false, true, false>;
    return k_contract_1m_batched<T, Idx, BM, BN, BK, WGM, WGN, false, false, false>;
  }
  static int per_cu(bool af, bool bf, bool pairs) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pick(af, bf, pairs), 64 * WGM * WGN,
                                                     0) != hipSuccess) {
      (void)hipGetLastError();
      return 1;
    }
    return n > 0 ? n : 1;
  }
  static hipError_t launch(bool af, bool bf, bool pairs, const void* inst,
                           const BatchItem* items, int nitems, int grid, unsigned* tickets,
                           int ndom, hipStream_t s) {
    hipLaunchKernelGGL(pick(af, bf, pairs), dim3(grid), dim3(64 * WGM * WGN), 0, s,
                       static_cast<const KArgs<Idx>*>(inst), items, nitems, tickets, ndom);
    return hipGetLastError();
  }
};

template <typename Idx, typename F>
auto with_1m_config(int32_t kernel, F&& f) {
  switch (kernel) {
    // The large tiles (a batch's large instances in tiling modes 1 and 2).
    case K_C64_F32_144: return f(Batched1M<float, Idx, 144, 144, 16, 2, 4, true>());
    case K_C64_F32_THIN:
    // Stages of 8 k: half the LDS and 160 registers, 3 workgroups per CU
    // instead of 2 (+2 % on 512 MPS contractions; items are short in k).
    case K_C64_F32_64: return f(Batched1M<float, Idx, 64, 64, 8, 2, 2, true>());
    case K_C128_F64_64:
    case K_C128_F64_64W: return f(Batched1M<double, Idx, 64, 64, 16, 2, 2>());
    default: return f(Batched1M<double, Idx, 32, 32, 16, 2, 2>());
  }
}

}  // namespace

size_t batch_args_size(bool idx32) {
  return idx32 ? sizeof(KArgs<int32_t>) : sizeof(KArgs<int64_t>);
}

void batch_make_args(const GemmProblem& g0, int32_t kernel, void* out, int* tiles_m,
                     int* tiles_n) {
  const KernelInfo& ki = kInfo[(kernel >= 0 && kernel < K_COUNT) ? kernel : 0];
  GemmProblem g = g0;
  g.split_k = 1;        // a batched launch covers whole tiles
  if (g.idx32) {
    const KArgs<int32_t> a = make_args<int32_t>(g, ki.bm, ki.bn);
    memcpy(out, &a, sizeof(a));
    *tiles_m = a.tiles_m; *tiles_n = a.tiles_n;
  } else {
    const KArgs<int64_t> a = make_args<int64_t>(g, ki.bm, ki.bn);
    memcpy(out, &a, sizeof(a));
    *tiles_m = a.tiles_m; *tiles_n = a.tiles_n;
  }
}

int batch_blocks_per_cu(int32_t kernel, bool idx32, bool af, bool bk, bool pairs) {
  const bool bf = !bk;
  if (kernel == K_C64_BF16X3_144 || kernel == K_C64_BF16X3_64)
    return bf16x3_batched_blocks_per_cu(kernel, idx32, af, bf);
  if (is_real_kernel(kernel)) return real_batched_blocks_per_cu(kernel, idx32, af, bf);
  if (idx32)
    return with_1m_config<int32_t>(kernel, [&](auto c) { return c.per_cu(af, bf, pairs); });
  return with_1m_config<int64_t>(kernel, [&](auto c) { return c.per_cu(af, bf, pairs); });
}

hipError_t launch_contract_batched(int32_t kernel, bool idx32, bool af, bool bk, bool pairs,
                                   const void* inst, const BatchItem* items, int nitems,
                                   int grid, unsigned* tickets, int ndom, hipStream_t s) {
  const bool bf = !bk;
  if (nitems == 0) return hipSuccess;
  if (kernel == K_C64_BF16X3_144 || kernel == K_C64_BF16X3_64)
    return launch_bf16x3_batched(kernel, idx32, af, bf, inst, items, nitems, grid, s);
  if (is_real_kernel(kernel))
    return launch_real_batched(kernel, idx32, af, bf, inst, items, nitems, grid, s);
  // every domain needs a workgroup and the list its equal share of entries
  if (!tickets || ndom < 1 || grid < ndom || nitems % ndom != 0) return hipErrorInvalidValue;
  if (idx32)
    return with_1m_config<int32_t>(kernel, [&](auto c) {
      return c.launch(af, bf, pairs, inst, items, nitems, grid, tickets, ndom, s);
    });
  return with_1m_config<int64_t>(kernel, [&](auto c) {
    return c.launch(af, bf, pairs, inst, items, nitems, grid, tickets, ndom, s);
  });
}

const KernelInfo& kernel_info(int32_t id) {
  return kInfo[(id >= 0 && id < K_COUNT) ? id : 0];
}

hipError_t launch_contract(const GemmProblem& g, int32_t kernel,
                           hipStream_t s) {
  if (g.M == 0 || g.N == 0) return hipSuccess;
  if (kernel == K_C64_BF16X3_144 || kernel == K_C64_BF16X3_64)
    return launch_bf16x3(g, kernel, s);
  if (is_real_kernel(kernel)) return launch_real(g, kernel, s);
  return g.idx32 ? dispatch<int32_t>(g, kernel, s) : dispatch<int64_t>(g, kernel, s);
}

}  // namespace t1m_hip
