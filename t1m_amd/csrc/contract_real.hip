// Real contraction C += alpha * A . B (t1m::contract for float / double,
// include/t1m/t1m.h:560-568; real branch of the driver t1m.h:466-507, whose
// micro-kernel call is bli_sgemm_ukernel / bli_dgemm_ukernel,
// bli/mappings.h:91-99) on gfx950 matrix cores.
//
// Same structure as the complex kernels (contract_mfma.hip): one workgroup
// per BM x BN tile of C, strided gathers into double-buffered, XOR-swizzled
// LDS, B fragments one 4-k chunk and A operands two row blocks ahead of
// their MFMAs, the next stage's LDS writes in the middle of the stage,
// grouped scatter-accumulate epilogue through the scatter vectors (the
// pipeline and swizzle: +4..14 % over chunk-wise fragments and plain rows).  Without 1M the MFMA is the plain real product: in
// the transposed form the MFMA rows are 16 columns j of C and its columns
// are 16 rows i, so lane l holds C[i = l % 16][.] and 16 lanes write 16
// consecutive rows.  v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64, one MFMA
// per 4 k per 16 x 16 tile.
#include "internal.h"
#include "kernel_common.h"

#include <cstring>

namespace t1m_hip {

namespace {

template <typename T, typename Idx, int BM, int BN, int BK, int WGM, int WGN,
          bool AF, bool BF>
__device__ __forceinline__ void contract_tile_real(const KArgs<Idx>& p, int tm, int tn, int kbeg, int kend) {
  using Acc = typename Vec<T>::acc;
  constexpr int NT = 64 * WGM * WGN;
  constexpr int WTM = BM / WGM;          // rows per wave
  constexpr int MB = WTM / 16;           // 16-row blocks per wave
  constexpr int NBT = BN / 16;
  constexpr int NBW = NBT / WGN;
  constexpr bool XT = NBT % WGN != 0;
  constexpr int XB = XT ? (MB + WGN - 1) / WGN : 0;
  constexpr int XB_ALLOC = XT ? XB : 1;
  // LDS rows of BX + PAD elements, PAD making the row pitch 16 words past a
  // multiple of 64 words (bank = 16 k + x for 4-byte elements), and element
  // (k, x) stored at column x ^ sw(k), sw(k) = 4 (k/4 mod 4) (4-byte) or
  // 2 (k/4 mod 4) (8-byte): a gather along k (4 or 2 rows x 16 k per write)
  // then hits every bank once, and the MFMA reads (16 consecutive x of one
  // aligned group, 4 k) stay a permutation of one bank sweep.
  constexpr int EW = int(sizeof(T)) / 4;             // words per element
  constexpr int PADA = ((16 / EW - BM % (64 / EW)) % (64 / EW) + 64 / EW) % (64 / EW);
  constexpr int PADB = ((16 / EW - BN % (64 / EW)) % (64 / EW) + 64 / EW) % (64 / EW);
  constexpr int LDA = BM + PADA, LDB = BN + PADB;
  static_assert((LDA * EW) % 64 == 16 && (LDB * EW) % 64 == 16, "row pitch");
  static_assert(WTM % 16 == 0 && BN % 16 == 0, "tile shape");
  static_assert(NBT - NBW * WGN <= 1, "at most one shared column block");
  static_assert(BK % 4 == 0, "k chunk");
  auto sw = [](int k) { return ((k >> 2) & 3) * (4 / EW); };

  __shared__ T sA[2][BK][LDA];
  __shared__ T sB[2][BK][LDB];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  const int wm = wid % WGM, wn = wid / WGM;
  // Shared column block: n-wave wn takes row blocks [xlo, xhi) (MB / WGN
  // each, one more for the first MB % WGN waves); block mb sits in
  // accx[mb % XB].
  const int xlo = wn * (MB / WGN) + (wn < MB % WGN ? wn : MB % WGN);
  const int xhi = xlo + MB / WGN + (wn < MB % WGN ? 1 : 0);
  const int64_t i0 = int64_t(tm) * BM, j0 = int64_t(tn) * BN;
  const int64_t M = p.M, N = p.N;
  const int K = kend;     // contracted range [kbeg, kend) of this workgroup
  const T* __restrict__ A = static_cast<const T*>(p.A);
  const T* __restrict__ B = static_cast<const T*>(p.B);

  const unsigned utid = unsigned(tid);
  using GA = TileGather<T, Idx, BM, BK, NT, AF>;
  using GB = TileGather<T, Idx, BN, BK, NT, BF>;
  GA ga;
  GB gb;
  ga.init(utid, i0, M, p.rsA);
  gb.init(utid, j0, N, p.csB);

  auto swrite = [&](int buf, int p0) {
    ga.finish_real(utid, p0, K);
    gb.finish_real(utid, p0, K);
#pragma unroll
    for (int u = 0; u < GA::E; ++u) {
      const unsigned e = GA::elem(utid, u);
      const int k = GA::kc(e);
      if (GA::valid(utid, u)) sA[buf][k][GA::xc(e) ^ sw(k)] = ga.val[u];
    }
#pragma unroll
    for (int u = 0; u < GB::E; ++u) {
      const unsigned e = GB::elem(utid, u);
      const int k = GB::kc(e);
      if (GB::valid(utid, u)) sB[buf][k][GB::xc(e) ^ sw(k)] = gb.val[u];
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

  const int g = lane >> 4;
  const int l16 = lane & 15;
  const int am0 = wm * WTM + l16;
  const int bn0 = wn * NBW * 16 + l16;
  const int bnx = NBW * WGN * 16 + l16;

  // Operand pipeline (as in the complex kernels): B fragments one 4-k
  // chunk ahead, A operands LA row blocks ahead of their MFMAs, scheduling
  // barriers keeping the order, the next stage's LDS writes after half the
  // stage, wave priority raised over the MFMA stream.  The shared column
  // block reuses the wave's A operands under a wave-uniform branch.
  constexpr int KCH = BK / 4;
  constexpr int TOT = KCH * MB;
  constexpr int LA = 2;
  struct BFr {
    T b[NBW];
    T bx;
  };
  auto load_b = [&](int buf, int kc, BFr& f) {
    const int q = kc * 4 + g;
    const int s = sw(q);
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb) f.b[nb] = sB[buf][q][(bn0 + nb * 16) ^ s];
    if constexpr (XT) f.bx = sB[buf][q][bnx ^ s];
  };
  auto load_a = [&](int buf, int t) -> T {
    const int kc = t / MB, grp = t % MB;
    const int q = kc * 4 + g;
    return sA[buf][q][(am0 + grp * 16) ^ sw(q)];
  };
  auto mma_grp = [&](int grp, T a, const BFr& f) {
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb) acc[grp][nb] = mfma16(f.b[nb], a, acc[grp][nb]);
    if constexpr (XT) {
      if (grp >= xlo && grp < xhi) {
        Acc& ax = accx[grp % XB];
        ax = mfma16(f.bx, a, ax);
      }
    }
  };
  auto compute = [&](int buf, auto&& mid) {
    BFr fb[2];
    T aw[LA + 1];
    __builtin_amdgcn_s_setprio(1);
    load_b(buf, 0, fb[0]);
#pragma unroll
    for (int t = 0; t < LA; ++t) aw[t] = load_a(buf, t);
#pragma unroll
    for (int t = 0; t < TOT; ++t) {
      const int kc = t / MB, grp = t % MB;
      if (grp == 0 && kc + 1 < KCH) load_b(buf, kc + 1, fb[(kc + 1) & 1]);
      if (t + LA < TOT) aw[(t + LA) % (LA + 1)] = load_a(buf, t + LA);
      __builtin_amdgcn_sched_barrier(0);
      mma_grp(grp, aw[t % (LA + 1)], fb[kc & 1]);
      __builtin_amdgcn_sched_barrier(0);
      if (t == TOT / 2 - 1) {
        __builtin_amdgcn_s_setprio(0);
        mid();
        __builtin_amdgcn_s_setprio(1);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    __builtin_amdgcn_s_setprio(0);
    if constexpr (TOT < 2) mid();
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
  for (int s = 0; s + 1 < nst; ++s) {
    ga.load_data(A);
    gb.load_data(B);
    ga.load_k(utid, kbeg + (s + 2) * BK, K, p.csA);
    gb.load_k(utid, kbeg + (s + 2) * BK, K, p.rsB);
    compute(s & 1, [&] { swrite((s + 1) & 1, kbeg + (s + 1) * BK); });
    __syncthreads();
  }
  compute((nst - 1) & 1, [] {});

  // Epilogue: this lane holds rows i0 + am0 + 16 r' (one per row block)
  // times 4 columns per 16-column block; offsets first, then store_tile.
  T* __restrict__ C = static_cast<T*>(p.C);
  const T alpha = T(p.alpha_re);
  constexpr int NR = MB + (XT ? XB : 0);        // row blocks (shared ones last)
  constexpr int NC = NBW + (XT ? 1 : 0);        // column blocks (shared one last)
  Idx rofs[NR];
  bool rok[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    // shared slot x = r - MB holds block xlo + ((x - xlo) mod XB), if < xhi
    const int blk = r < MB ? r : xlo + ((r - MB) - xlo % XB_ALLOC + XB_ALLOC) % XB_ALLOC;
    const int64_t i = i0 + am0 + (blk < MB ? blk : MB - 1) * 16;
    rok[r] = i < M && (r < MB || blk < xhi);
    rofs[r] = rok[r] ? gmem(p.rsC)[i] : Idx(0);
  }
  Idx cofs[NC][4];
  bool cok[NC][4];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      // MFMA output row of (lane, q): 4(lane/16) + q (f32), lane/16 + 4q (f64).
      const int64_t jb = c < NBW ? j0 + wn * NBW * 16 + c * 16 : j0 + NBW * WGN * 16;
      const int64_t j = jb + (sizeof(T) == 4 ? 4 * g + q : g + 4 * q);
      cok[c][q] = j < N;
      cofs[c][q] = cok[c][q] ? gmem(p.csC)[j] : Idx(0);
    }
  auto value = [&](int r, int c, int q) -> T {
    if constexpr (XT) {
      if (r >= MB) return alpha * accx[r - MB][q];
    }
    return alpha * acc[r][c][q];
  };
  auto has = [](int r, int c) { return r < MB ? c < NBW : c == NC - 1; };
  auto ok = [&](int r, int c, int q) { return rok[r] && cok[c][q]; };
  auto addr = [&](int r, int c, int q) { return C + int64_t(rofs[r] + cofs[c][q]); };
  store_tile<T, NR, NC, 3>(p.atomic_c != 0, has, ok, value, addr);
}

template <typename T, typename Idx, int BM, int BN, int BK, int WGM, int WGN, bool AF, bool BF>
__global__ void __launch_bounds__(64 * WGM * WGN) k_contract_real(KArgs<Idx> p) {
  int tm, tn, kbeg, kend;
  tile_and_krange(p, tm, tn, kbeg, kend);
  contract_tile_real<T, Idx, BM, BN, BK, WGM, WGN, AF, BF>(p, tm, tn, kbeg, kend);
}

template <typename T, typename Idx, int BM, int BN, int BK, int WGM, int WGN, bool AF, bool BF>
__global__ void __launch_bounds__(64 * WGM * WGN)
k_contract_real_batched(const KArgs<Idx>* __restrict__ inst, const BatchItem* __restrict__ items,
                        int nitems) {
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const BatchItem w = items[it];
    if (w.inst < 0) continue;   // padding of an XCD-ordered work list
    const KArgs<Idx> p = inst[w.inst];
    contract_tile_real<T, Idx, BM, BN, BK, WGM, WGN, AF, BF>(p, w.tm, w.tn, 0, int(p.K));
    __syncthreads();
  }
}

template <typename T, typename Idx, int BM, int BN, int BK, int WGM, int WGN>
struct RealCfg {
  static constexpr int bm = BM, bn = BN;
  using KB = void (*)(const KArgs<Idx>*, const BatchItem*, int);
  template <bool AF, bool BF>
  static hipError_t launch_one(const KArgs<Idx>& a, hipStream_t s) {
    const int64_t per = int64_t(a.tiles_m) * a.tiles_n * a.ksplit;
    return launch_batches(a, per, 64 * WGM * WGN, [&](const KArgs<Idx>& q, unsigned grid) {
      hipLaunchKernelGGL((k_contract_real<T, Idx, BM, BN, BK, WGM, WGN, AF, BF>), grid,
                         dim3(64 * WGM * WGN), 0, s, q);
    });
  }
  static hipError_t launch(const GemmProblem& g, hipStream_t s) {
    const KArgs<Idx> a = make_args<Idx>(g, BM, BN, BK);
    const bool af = g.a_mfast != 0, bf = g.b_kfast == 0;
    if (af && bf) return launch_one<true, true>(a, s);
    if (af) return launch_one<true, false>(a, s);
    if (bf) return launch_one<false, true>(a, s);
    return launch_one<false, false>(a, s);
  }
  static KB pick(bool af, bool bf) {
    if (af && bf) return k_contract_real_batched<T, Idx, BM, BN, BK, WGM, WGN, true, true>;
    if (af) return k_contract_real_batched<T, Idx, BM, BN, BK, WGM, WGN, true, false>;
    if (bf) return k_contract_real_batched<T, Idx, BM, BN, BK, WGM, WGN, false, true>;
    return k_contract_real_batched<T, Idx, BM, BN, BK, WGM, WGN, false, false>;
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
  static hipError_t launch_batched(bool af, bool bf, const void* inst, const BatchItem* items,
                                   int nitems, int grid, hipStream_t s) {
    hipLaunchKernelGGL(pick(af, bf), dim3(grid), dim3(64 * WGM * WGN), 0, s,
                       static_cast<const KArgs<Idx>*>(inst), items, nitems);
    return hipGetLastError();
  }
};

template <typename Idx, typename F>
auto with_real_config(int32_t kernel, F&& f) {
  switch (kernel) {
    // 8 waves (64 x 32 each): 2 waves per SIMD (+30 % over 4 waves of 64 x 64).
    // (256 x 128 on 4 x 2 waves: -6 %.)
    case K_F32_128: return f(RealCfg<float, Idx, 128, 128, 16, 2, 4>());
    case K_F32_64: return f(RealCfg<float, Idx, 64, 64, 16, 2, 2>());
    // (4 x 2 waves: -2..+7 %; 128 x 128 on 2 x 4: -10..-25 %.)
    case K_F64_64: return f(RealCfg<double, Idx, 64, 64, 16, 2, 2>());
    default: return f(RealCfg<double, Idx, 32, 32, 16, 2, 2>());
  }
}

}  // namespace

bool is_real_kernel(int32_t kernel) {
  return kernel == K_F32_128 || kernel == K_F32_64 || kernel == K_F64_64 || kernel == K_F64_32;
}

hipError_t launch_real(const GemmProblem& g, int32_t kernel, hipStream_t s) {
  if (g.idx32) return with_real_config<int32_t>(kernel, [&](auto c) { return c.launch(g, s); });
  return with_real_config<int64_t>(kernel, [&](auto c) { return c.launch(g, s); });
}

int real_batched_blocks_per_cu(int32_t kernel, bool idx32, bool af, bool bf) {
  if (idx32) return with_real_config<int32_t>(kernel, [&](auto c) { return c.per_cu(af, bf); });
  return with_real_config<int64_t>(kernel, [&](auto c) { return c.per_cu(af, bf); });
}

hipError_t launch_real_batched(int32_t kernel, bool idx32, bool af, bool bf, const void* inst,
                               const BatchItem* items, int nitems, int grid, hipStream_t s) {
  if (idx32)
    return with_real_config<int32_t>(kernel, [&](auto c) {
      return c.launch_batched(af, bf, inst, items, nitems, grid, s);
    });
  return with_real_config<int64_t>(kernel, [&](auto c) {
    return c.launch_batched(af, bf, inst, items, nitems, grid, s);
  });
}

}  // namespace t1m_hip
