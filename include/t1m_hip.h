/*
 * t1m_hip.h -- C ABI of libt1m_hip.so, the MI355X (gfx950) implementation of
 * t1m's transposition-free tensor contraction.
 *
 * The reference library is header-only C++ with no FFI; its boundary is the
 * pair of templates t1m::contract (include/t1m/t1m.h:560-568 real,
 * 570-577 complex) over t1m::tensor (t1m.h:31-66).  This header is the plain-C
 * layer those templates call in this project (include/t1m/t1m.h here), and
 * what a ctypes / cgo / JNI binding would bind.  No exceptions, no STL and no
 * torch types cross it: plain pointers, sizes and int status codes.
 *
 * Pointers: tensor data may be device memory (the fast path; nothing is
 * copied) or host memory (staged through device memory and copied back, so
 * the reference's host-pointer examples run unchanged).
 *
 * Semantics follow the reference: C += alpha * A . B, with beta ignored
 * (t1m.h:497-502, 543-548: the micro-kernel writes a zeroed tile that unpack
 * adds into C).
 */
#ifndef T1M_HIP_H_
#define T1M_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define T1M_MAX_RANK 16

/* Element types: t1m.h:15-23 (TensorScalarArithmetic / TensorScalarCompound). */
typedef enum t1m_dtype {
  T1M_FLOAT32 = 0,
  T1M_FLOAT64 = 1,
  T1M_COMPLEX64 = 2,   /* std::complex<float>, interleaved (re, im) */
  T1M_COMPLEX128 = 3   /* std::complex<double> */
} t1m_dtype;

/* Arithmetic used for the contraction.  The reference has one path (BLIS
 * sgemm/dgemm micro-kernel on 1M panels, mappings.h:91-99). */
typedef enum t1m_precision {
  T1M_PRECISION_NATIVE = 0,    /* FP32 MFMA for float data, FP64 MFMA for double */
  T1M_PRECISION_BF16X3 = 1,    /* float data only: bf16 hi/lo operand splits,
                                  hi*hi + hi*lo + lo*hi on BF16 MFMA, FP32 sums */
  T1M_PRECISION_SIMPLE = 2     /* one lane per output element, plain FMA loop;
                                  no MFMA (a cross-check path) */
} t1m_precision;

typedef enum t1m_status {
  T1M_SUCCESS = 0,
  T1M_ERROR_INVALID_ARGUMENT = 1,  /* labels, extents, ranks, dtypes */
  T1M_ERROR_OUT_OF_MEMORY = 2,     /* the reference throws std::bad_alloc (t1m.h:455-458) */
  T1M_ERROR_DEVICE = 3,            /* a HIP runtime error */
  T1M_ERROR_UNSUPPORTED = 4        /* e.g. BF16x3 on double data */
} t1m_status;

/* A strided tensor view: t1m::tensor<T, ndim> (t1m.h:31-66) plus explicit
 * strides (the reference derives col-major strides on demand, t1m.h:42-50). */
typedef struct t1m_tensor {
  int32_t dtype;            /* t1m_dtype */
  int32_t rank;             /* 0 .. T1M_MAX_RANK */
  const int64_t* dims;      /* rank extents */
  const int64_t* strides;   /* rank strides in elements; NULL = column-major */
  void* data;               /* device or host pointer */
} t1m_tensor;

typedef struct t1m_options {
  double alpha_re;          /* C += alpha * A.B; the complex reference entry */
  double alpha_im;          /* point fixes alpha = 1 (t1m.h:576) */
  int32_t precision;        /* t1m_precision */
  int32_t conj_a;           /* use conj(A) (complex only) */
  int32_t conj_b;           /* use conj(B) (complex only) */
  int32_t synchronize;      /* nonzero: the result is visible on return */
  void* stream;             /* hipStream_t; NULL = the null stream */
  int32_t autotune;         /* nonzero: the first call of a new shape times
                               the candidate tile shapes (on a scratch C) and
                               keeps the fastest; see t1m_autotune_cache_path */
  int32_t deterministic;    /* nonzero: never split K across workgroups, so
                               results are bitwise reproducible; by default a
                               contraction whose C tiles cannot fill the GPU
                               splits K and adds the partial sums into C with
                               atomics (order of those additions varies) */
} t1m_options;

/* alpha = 1, NATIVE, no conjugation, synchronize = 1, null stream,
 * autotune = 0 (tile shape from the tuning cache, else a static rule),
 * deterministic = 0. */
void t1m_options_init(t1m_options* opt);

/*
 * C += alpha * A . B over einsum-style labels (one character per mode).
 * Replaces t1m::internal::contract (t1m.h:432-557) as called by both public
 * overloads (t1m.h:560-577).  Labels shared by A and B only are contracted;
 * every other label of A and of B must appear in C, and C has no other
 * labels.  A label in A, B and C is a batch mode, C[..z..] += alpha *
 * sum_k A[..z..] B[..z..] (einsum semantics; the reference sums over such a
 * label and leaves C's mode unindexed, t1m.h:106-119).  Unlike the
 * reference, malformed labels and mismatched extents are reported
 * (T1M_ERROR_INVALID_ARGUMENT) instead of being undefined behaviour.
 */
int t1m_contract(const t1m_tensor* a, const char* la, const t1m_tensor* b,
                 const char* lb, const t1m_tensor* c, const char* lc,
                 const t1m_options* opt);

/*
 * Vector-Jacobian product of C = alpha * A' . B' (no counterpart in the
 * reference), where A' = conj(A) if opt->conj_a is set, else A (B' likewise),
 * exactly the operands t1m_contract uses with the same options:
 *   grad_a += conj(alpha) * grad_c . conj(B')   (over B's free modes)
 *   grad_b += conj(alpha) * conj(A') . grad_c   (over A's free modes)
 * The gradients are with respect to A' and B' (what a caller holding a lazy
 * conjugate view differentiates).  grad_a / grad_b may be NULL to skip one.
 * For real dtypes conj is identity.  With torch's convention, grad_c of a
 * real loss is dL/dRe + i dL/dIm.
 */
int t1m_contract_vjp(const t1m_tensor* a, const char* la, const t1m_tensor* b,
                     const char* lb, const t1m_tensor* grad_c, const char* lc,
                     const t1m_tensor* grad_a, const t1m_tensor* grad_b,
                     const t1m_options* opt);

/*
 * t1m_contract without device allocation (no counterpart in the reference).
 * For callers that cannot allocate device memory: code under HIP graph
 * capture (including the first call of a shape), tight serving loops,
 * harnesses that forbid hipMalloc.  Same semantics and options as
 * t1m_contract, except: A, B and C must be device memory; autotune is
 * rejected; no plan is cached -- every call rebuilds the contraction's index
 * maps (one extra small launch) in a static device array of
 * T1M_NOALLOC_SCRATCH_BYTES per GPU, and a contraction whose maps do not fit
 * (2 x (M + N + K) offsets of 4 or 8 bytes, plus 24 bytes per batch problem)
 * returns T1M_ERROR_UNSUPPORTED.  All calls on one GPU share that array:
 * calls that overlap on one GPU race (serialise them on one stream, or wait
 * between streams).
 */
#define T1M_NOALLOC_SCRATCH_BYTES (8u << 20)
int t1m_contract_noalloc(const t1m_tensor* a, const char* la, const t1m_tensor* b,
                         const char* lb, const t1m_tensor* c, const char* lc,
                         const t1m_options* opt);

/*
 * Batched contraction (no counterpart in the reference): count independent
 * contractions sharing one label triple and dtype, each with its own extents,
 * strides and device pointers (tensor-network sweeps).  create() builds the
 * index maps and an LPT-ordered work list on the device and captures the
 * launch in a HIP graph; launch() replays it.  All data must be device memory,
 * and stays in use by the graph until destroy() (the caller keeps it alive).
 * Launches of one batch must not overlap (serialise them on one stream): they
 * add into the same C and share the work list's device counters.
 */
typedef struct t1m_batch t1m_batch;
int t1m_batch_create(int64_t count, const t1m_tensor* a, const char* la,
                     const t1m_tensor* b, const char* lb, const t1m_tensor* c,
                     const char* lc, const t1m_options* opt, t1m_batch** out);
int t1m_batch_launch(t1m_batch* batch, void* stream);
int t1m_batch_destroy(t1m_batch* batch);
/* What the batch's graph runs: its tiling mode and, per persistent launch,
 * the tile family, instances, tiles and grid (or the per-instance launch
 * count).  With opt->autotune, create() times the candidate tiling modes
 * (which instances take the large tiles) on a scratch copy of C and keeps
 * the fastest; the choice is kept in the tuning file for later batches of
 * the same shapes. */
const char* t1m_batch_describe(const t1m_batch* batch);

/* ---- index maps (exported so they can be checked bit for bit) ---------- */

/*
 * Index bundles AX, AY, BX, BY, CX, CY of get_index_bundles (t1m.h:90-136):
 * bundles is [6][T1M_MAX_RANK], counts is [6].  Host only.
 */
int t1m_index_bundles(const char* la, const char* lb, const char* lc,
                      int64_t* bundles, int64_t* counts);

/*
 * Scatter vector of the modes `modes` of tensor t (get_scatter, t1m.h:141-174:
 * first listed mode fastest) and its block-scatter vector at block size
 * `block` (get_block_scatter, t1m.h:179-205), built on the device into the
 * device arrays scat (prod of the extents) and block_scat (ceil(len/block)).
 * Values are modulo 2^64, as std::size_t in the reference.  block_scat may be
 * NULL.  Runs on `stream` and returns after it has been queued.
 */
int t1m_index_maps(const t1m_tensor* t, const int64_t* modes, int32_t nmodes,
                   int64_t block, uint64_t* scat, uint64_t* block_scat,
                   void* stream);

/* matrix_view (t1m.h:210-241); the four vectors are device arrays. */
typedef struct t1m_view {
  int64_t br, bc;
  const uint64_t* rs;  int64_t nrs;
  const uint64_t* cs;  int64_t ncs;
  const uint64_t* rbs; int64_t nrbs;
  const uint64_t* cbs; int64_t ncbs;
} t1m_view;

/*
 * Device pack_block<T, A|B> (t1m.h:341-404, cells t1m.h:279-339) into the
 * reference's micro-panel format (real panels; 1M for complex dtypes).
 * label 0 = A (column-major panels), 1 = B (row-major panels).  As in the
 * reference, only panel entries backed by the view are written; the caller
 * zero-fills dest first (t1m.h:518, 527).
 */
int t1m_pack_block(int32_t dtype, int32_t label, const t1m_view* view,
                   int64_t length, const void* src, void* dest, void* stream);

/* Device unpack (t1m.h:406-430): dest[idx] += tile, tile column-major with
 * leading dimension view->br (real rows; complex pairs for complex dtypes). */
int t1m_unpack(int32_t dtype, const t1m_view* view, const void* src,
               void* dest, void* stream);

/* ---- diagnostics -------------------------------------------------------- */

/* Message of the last failing call on this thread ("" if none). */
const char* t1m_last_error(void);
/* Library version string. */
const char* t1m_version(void);
/* File that keeps autotuning results across processes: $T1M_TUNE_CACHE, else
 * $HOME/.cache/t1m_mi355x/tuning.txt.  Lines are "<key> <kernel name>". */
const char* t1m_autotune_cache_path(void);
/* Name of the kernel the plan for this contraction uses (tuning / tests);
 * builds the plan if needed.  Returns NULL on error. */
const char* t1m_plan_kernel(const t1m_tensor* a, const char* la,
                            const t1m_tensor* b, const char* lb,
                            const t1m_tensor* c, const char* lc,
                            const t1m_options* opt);
/* Number of plans in the plan cache (index maps kept on the device). */
int64_t t1m_plan_cache_size(void);
/* Drop every cached plan and free its device memory (and the pooled host-staging buffers). */
void t1m_plan_cache_clear(void);

#ifdef __cplusplus
}
#endif

#endif  /* T1M_HIP_H_ */
