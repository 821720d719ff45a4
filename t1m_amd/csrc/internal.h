// Internal declarations shared by the translation units of libt1m_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "t1m_hip.h"

namespace t1m_hip {

// ---- index maps ---------------------------------------------------------

// One scatter vector to build: entry e = sum_j (e_j) * strides[j] with the
// mixed-radix digits e_j of e over dims[] (first mode fastest).
struct MapDesc {
  int32_t nmodes;
  int32_t out32;          // 1: write int32, 0: write 64-bit words
  int64_t len;            // product of dims (1 for no modes)
  int64_t dims[T1M_MAX_RANK];
  int64_t strides[T1M_MAX_RANK];
  void* out;
};

constexpr int kMaxMapsPerLaunch = 6;
struct MapBatch {
  int32_t count;
  MapDesc maps[kMaxMapsPerLaunch];
};

// Builds up to six scatter vectors in one launch (grid.y = map).
hipError_t launch_scatter_maps(const MapBatch& batch, hipStream_t stream);
// get_block_scatter on the device, one lane per block.
hipError_t launch_block_scatter(const uint64_t* scat, int64_t len, int64_t b,
                                uint64_t* out, hipStream_t stream);
hipError_t launch_pack_block(int dtype, int label, const t1m_view& v,
                             int64_t length, const void* src, void* dest,
                             hipStream_t stream);
hipError_t launch_unpack(int dtype, const t1m_view& v, const void* src,
                         void* dest, hipStream_t stream);

// ---- contraction kernels ------------------------------------------------

// A matricised contraction: C[rsC[i] + csC[j]] += alpha * sum_p
// A[rsA[i] + csA[p]] * B[rsB[p] + csB[j]], offsets in elements.
struct GemmProblem {
  const void* A;
  const void* B;
  void* C;
  const void* rsA;   // M entries (int32 or int64, see idx32)
  const void* csA;   // K
  const void* rsB;   // K
  const void* csB;   // N
  const void* rsC;   // M
  const void* csC;   // N
  int64_t M, N, K;
  double alpha_re, alpha_im;
  int32_t conj_a, conj_b;
  int32_t a_mfast;   // stage A with lanes along m (else along k)
  int32_t b_kfast;   // stage B with lanes along k (else along n)
  int32_t idx32;     // offsets are int32 (else int64)
  int32_t dtype;     // t1m_dtype
  int32_t split_k;   // 0: split K when the tiles cannot fill the CUs; 1: never
  int32_t cus;       // compute units of the device (split-K sizing)
  int32_t c_overlap = 0;   // distinct outputs may share an address: add atomically
  // complex<float>: the operand's gather-direction map has map[i + 1] =
  // map[i] + 1 for every even i, so lanes may load neighbour pairs.
  int32_t a_pairs = 0, b_pairs = 0;
  // Batch modes (labels in A, B and C): L independent problems, problem z
  // reading A, B, C at byte offsets boffA[z], boffB[z], boffC[z] (device).
  int64_t L = 1;
  const int64_t* boffA = nullptr;
  const int64_t* boffB = nullptr;
  const int64_t* boffC = nullptr;
  // Smallest entry of each map (element offsets; <= 0), and whether the
  // byte spans of A, B and C from those minima stay below 2^32 (kernels may
  // then address them by 32-bit byte offsets from rebased bases).
  int64_t min_rsA = 0, min_csA = 0, min_rsB = 0, min_csB = 0, min_rsC = 0, min_csC = 0;
  int32_t span32 = 0;
};

// Kernel families.
enum KernelId : int32_t {
  K_SIMPLE = 0,        // one lane per output element (any dtype)
  K_C64_F32_144 = 1,   // complex<float>, 1M on FP32 MFMA, 144x144 tiles
  K_C64_F32_64 = 2,    // complex<float>, 1M on FP32 MFMA, 64x64 tiles
  K_C128_F64_64 = 3,   // complex<double>, 1M on FP64 MFMA, 64x64 tiles
  K_C128_F64_32 = 4,   // complex<double>, 1M on FP64 MFMA, 32x32 tiles
  K_C64_BF16X3_144 = 5,// complex<float>, 1M BF16x3 MFMA, 144x144 tiles
  K_C64_BF16X3_64 = 6, // complex<float>, 1M BF16x3 MFMA, 64x64 tiles
  K_F32_128 = 7,       // float, FP32 MFMA, 128x128 tiles
  K_F32_64 = 8,        // float, FP32 MFMA, 64x64 tiles
  K_F64_64 = 9,        // double, FP64 MFMA, 64x64 tiles
  K_F64_32 = 10,       // double, FP64 MFMA, 32x32 tiles
  K_C64_F32_THIN = 11, // complex<float>, 1M on FP32 MFMA, 128x16 tiles (N <= 16)
  K_C128_F64_THIN = 12,// complex<double>, 1M on FP64 MFMA, 128x16 tiles (N <= 16)
  K_C128_F64_64W = 13, // complex<double>, 1M on FP64 MFMA, 64x64 tiles on 16 waves
  K_COUNT
};

struct KernelInfo {
  const char* name;
  int32_t bm, bn;      // workgroup tile, in complex (or real) outputs
};
const KernelInfo& kernel_info(int32_t id);

hipError_t launch_contract(const GemmProblem& p, int32_t kernel,
                           hipStream_t stream);

// ---- persistent batched launch --------------------------------------------

// One unit of a batched launch: tile (tm, tn) of instance `inst`.
struct BatchItem {
  int32_t inst;
  int32_t tm;
  int32_t tn;
  int32_t work;   // sort key (contracted extent)
};

// Size in bytes of one instance's kernel arguments for a batched launch.
size_t batch_args_size(bool idx32);
// Writes the kernel arguments of problem g for kernel `kernel` into out
// (batch_args_size bytes) and returns its tile grid (tiles_m x tiles_n).
void batch_make_args(const GemmProblem& g, int32_t kernel, void* out,
                     int* tiles_m, int* tiles_n);
// Resident workgroups per CU of the batched form of `kernel`.
// pairs: every instance allows two-element gathers of both operands.
int batch_blocks_per_cu(int32_t kernel, bool idx32, bool a_mfast, bool b_kfast, bool pairs);
// Launches the persistent batched kernel over `nitems` work items: `ndom`
// interleaved per-domain lists (entry k * ndom + d of domain d).  tickets:
// ndom + 1 zeroed counters (the 1M kernels take the entries of their
// domain's list through them and leave them zero at exit; the other
// kernels deal the list statically and ignore them).
hipError_t launch_contract_batched(int32_t kernel, bool idx32, bool a_mfast,
                                   bool b_kfast, bool pairs, const void* inst,
                                   const BatchItem* items, int nitems, int grid,
                                   unsigned* tickets, int ndom, hipStream_t stream);

}  // namespace t1m_hip
