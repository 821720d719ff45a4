/*
 * t1m_oracle.c -- CPU restatement of the reference t1m contraction path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (t1m_amd/, include/)
 * links, loads or calls this file.  Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py use it, and only as the checker / the CPU
 * timing baseline.
 *
 * Every function restates one piece of include/t1m/t1m.h (the reference
 * header) and cites the lines it follows.  Index arithmetic is done in
 * uint64_t with wrap-around, exactly like the reference's std::size_t.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   orc_bundles          reference KATs test_scatter.cpp:7-29 + brute force
 *                        over random label sets against Python set algebra.
 *   orc_scatter          KAT test_scatter.cpp:31-40 + closed-form mixed radix
 *                        (numpy) + brute force over tiny tensors.
 *   orc_block_scatter    KATs test_scatter.cpp:42-71 + closed-form rule.
 *   orc_pack_block       KATs test_packing.cpp:23-169 (real), 210-385 (1M),
 *                        + closed form [re -im; im re] / [re; im] on random
 *                        views against a numpy gather.
 *   orc_unpack           KATs test_packing.cpp:171-204, 387-428 + closed form.
 *   orc_contract         GEMM-ordered labels reduce to a plain GEMM
 *                        (numpy.matmul, the reference's own oracle at
 *                        test_contraction.cpp:93-105) and permuted labels
 *                        against numpy.einsum in complex128.
 *   orc_einsum_*         numpy.einsum (a library routine) on random shapes,
 *                        plus brute-force hand loops on tiny inputs.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAX_RANK 32

/* ---------------------------------------------------------------------- */
/* Index bundles: t1m.h:90-136.                                            */
/* ---------------------------------------------------------------------- */

static int cmp_char(const void* x, const void* y) {
  return (int)(*(const unsigned char*)x) - (int)(*(const unsigned char*)y);
}

/* std::string::find for one character; npos is reported as UINT64_MAX. */
static uint64_t find_char(const char* s, char ch) {
  const char* p = strchr(s, ch);
  if (!p || ch == '\0') return UINT64_MAX;
  return (uint64_t)(p - s);
}

/*
 * Sorted-multiset intersection / difference with the merge semantics of
 * std::set_intersection and std::set_difference (t1m.h:99-111).
 */
static size_t ms_intersection(const char* a, size_t na, const char* b,
                              size_t nb, char* out) {
  size_t i = 0, j = 0, n = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) ++i;
    else if (b[j] < a[i]) ++j;
    else { out[n++] = a[i]; ++i; ++j; }
  }
  return n;
}

static size_t ms_difference(const char* a, size_t na, const char* b,
                            size_t nb, char* out) {
  size_t i = 0, j = 0, n = 0;
  while (i < na) {
    if (j == nb) { out[n++] = a[i++]; continue; }
    if (a[i] < b[j]) out[n++] = a[i++];
    else { if (!(b[j] < a[i])) ++i; ++j; }
  }
  return n;
}

/*
 * out[0..5] receive AX, AY, BX, BY, CX, CY (each up to ORC_MAX_RANK
 * entries, row-major [6][ORC_MAX_RANK]); counts[0..5] their lengths.
 * t1m.h:90-136.  Returns -1 if any label string exceeds ORC_MAX_RANK.
 */
int orc_bundles(const char* la, const char* lb, const char* lc, uint64_t* out,
                int64_t* counts) {
  size_t na = strlen(la), nb = strlen(lb), nc = strlen(lc);
  if (na > ORC_MAX_RANK || nb > ORC_MAX_RANK || nc > ORC_MAX_RANK) return -1;
  char a[ORC_MAX_RANK + 1], b[ORC_MAX_RANK + 1];
  memcpy(a, la, na); memcpy(b, lb, nb);
  qsort(a, na, 1, cmp_char);           /* t1m.h:99-101 */
  qsort(b, nb, 1, cmp_char);
  char contracted[ORC_MAX_RANK], free_a[ORC_MAX_RANK], free_b[ORC_MAX_RANK];
  size_t nk = ms_intersection(a, na, b, nb, contracted);  /* t1m.h:105-106 */
  size_t nfa = ms_difference(a, na, b, nb, free_a);       /* t1m.h:107-108 */
  size_t nfb = ms_difference(b, nb, a, na, free_b);       /* t1m.h:109-110 */
  uint64_t* AX = out + 0 * ORC_MAX_RANK;
  uint64_t* AY = out + 1 * ORC_MAX_RANK;
  uint64_t* BX = out + 2 * ORC_MAX_RANK;
  uint64_t* BY = out + 3 * ORC_MAX_RANK;
  uint64_t* CX = out + 4 * ORC_MAX_RANK;
  uint64_t* CY = out + 5 * ORC_MAX_RANK;
  for (size_t i = 0; i < nk; ++i) {       /* t1m.h:114-119 */
    AY[i] = find_char(la, contracted[i]);
    BX[i] = find_char(lb, contracted[i]);
  }
  for (size_t i = 0; i < nfa; ++i) {      /* t1m.h:121-126 */
    AX[i] = find_char(la, free_a[i]);
    CX[i] = find_char(lc, free_a[i]);
  }
  for (size_t i = 0; i < nfb; ++i) {      /* t1m.h:128-133 */
    BY[i] = find_char(lb, free_b[i]);
    CY[i] = find_char(lc, free_b[i]);
  }
  counts[0] = (int64_t)nfa; counts[1] = (int64_t)nk;
  counts[2] = (int64_t)nk;  counts[3] = (int64_t)nfb;
  counts[4] = (int64_t)nfa; counts[5] = (int64_t)nfb;
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Scatter vectors: t1m.h:141-174.                                         */
/* ---------------------------------------------------------------------- */

/* Length of the scatter vector of a bundle: product of its extents. */
uint64_t orc_scatter_len(const uint64_t* idx, int64_t nidx,
                         const uint64_t* dims) {
  uint64_t n = 1;
  for (int64_t i = 0; i < nidx; ++i) n *= dims[idx[i]];
  return n;
}

/*
 * The reference builds per-mode vectors v_j[x] = x * stride (t1m.h:148-159)
 * and folds them with f(av, bv)[b + a*|bv|] = av[a] + bv[b] starting from the
 * last mode (t1m.h:162-173).  The second argument varies fastest, so the
 * first mode of the bundle ends up fastest.  This restates the fold literally
 * (sequential left fold over the reversed parts; the combiner is associative
 * and order-preserving, so libstdc++'s unrolled std::reduce yields the same
 * vector).  An empty bundle is undefined in the reference (parts.back() on an
 * empty vector); here it yields the one-element vector {0}.
 */
void orc_scatter(const uint64_t* idx, int64_t nidx, const uint64_t* dims,
                 const uint64_t* strides, uint64_t* out) {
  if (nidx == 0) { out[0] = 0; return; }
  uint64_t total = orc_scatter_len(idx, nidx, dims);
  uint64_t* acc = (uint64_t*)malloc(sizeof(uint64_t) * (total ? total : 1));
  uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * (total ? total : 1));
  /* init = parts.back() */
  uint64_t len = dims[idx[nidx - 1]];
  for (uint64_t j = 0; j < len; ++j) acc[j] = j * strides[idx[nidx - 1]];
  /* fold with parts[nidx-2], ..., parts[0] as the second argument */
  for (int64_t p = nidx - 2; p >= 0; --p) {
    uint64_t d = dims[idx[p]], s = strides[idx[p]];
    for (uint64_t a = 0; a < len; ++a)
      for (uint64_t b = 0; b < d; ++b) tmp[b + a * d] = acc[a] + b * s;
    len *= d;
    uint64_t* t = acc; acc = tmp; tmp = t;
  }
  memcpy(out, acc, sizeof(uint64_t) * len);
  free(acc); free(tmp);
}

/* ---------------------------------------------------------------------- */
/* Block scatter: t1m.h:179-205.                                           */
/* ---------------------------------------------------------------------- */

uint64_t orc_block_scatter_len(uint64_t n, uint64_t b) {
  return n / b + (n % b > 0);   /* div_ceil, t1m.h:73-75 */
}

void orc_block_scatter(const uint64_t* scat, uint64_t n, uint64_t b,
                       uint64_t* out) {
  uint64_t o = 0;
  for (uint64_t i = 0; i < n; i += b) {
    uint64_t bs = (b < n - i) ? b : n - i;   /* t1m.h:185 */
    const uint64_t* blk = scat + i;
    if (bs == 1) { out[o++] = 0; continue; } /* t1m.h:188-191 */
    /* adjacent_difference: dist[0] = blk[0], dist[t] = blk[t] - blk[t-1]. */
    uint64_t last = blk[bs - 1] - blk[bs - 2];
    int uniform = 1;
    /* adjacent_find over dist[1..] with not_equal_to (t1m.h:195-199). */
    for (uint64_t t = 2; t < bs; ++t)
      if (blk[t] - blk[t - 1] != blk[t - 1] - blk[t - 2]) { uniform = 0; break; }
    out[o++] = uniform ? last : 0;
  }
}

/* ---------------------------------------------------------------------- */
/* Matrix views: t1m.h:210-241.                                            */
/* ---------------------------------------------------------------------- */

typedef struct {
  uint64_t br; const uint64_t* rbs; uint64_t nrbs; const uint64_t* rs; uint64_t nrs;
  uint64_t bc; const uint64_t* cbs; uint64_t ncbs; const uint64_t* cs; uint64_t ncs;
} orc_view;

static uint64_t div_ceil(uint64_t x, uint64_t y) { return x / y + (x % y > 0); }

/* t1m.h:218-237: rows slice (offset, n), block strides (offset/br, ceil(n/br)). */
static orc_view subview(orc_view v, uint64_t ri, uint64_t ci, uint64_t nr,
                        uint64_t nc) {
  orc_view w = v;
  w.rs = v.rs + ri; w.nrs = nr;
  w.rbs = v.rbs + ri / v.br; w.nrbs = div_ceil(nr, v.br);
  w.cs = v.cs + ci; w.ncs = nc;
  w.cbs = v.cbs + ci / v.bc; w.ncbs = div_ceil(nc, v.bc);
  return w;
}

/* ---------------------------------------------------------------------- */
/* 1M / real packing: t1m.h:279-404, unpack t1m.h:406-430.                 */
/* dtype: 0 float, 1 double, 2 complex<float>, 3 complex<double>.          */
/* ---------------------------------------------------------------------- */

/* Source index of cell element (k, l), with the dense fast path of
 * t1m.h:285-289 (wrap-around arithmetic, identical to rs[k] + cs[l] when the
 * cell's block strides are uniform). */
static inline uint64_t cell_src(const orc_view* c, uint64_t k, uint64_t l) {
  uint64_t rsc = c->rbs[0], csc = c->cbs[0];
  if (rsc > 0 && csc > 0) return k * rsc + l * csc + (c->rs[0] + c->cs[0]);
  return c->rs[k] + c->cs[l];
}

#define DEF_PACK_CELL_REAL(NAME, T)                                          \
  static void NAME(const orc_view* c, int labelA, const T* src, T* dest) {   \
    uint64_t nr = c->nrs, nc = c->ncs;                                       \
    if (labelA) { /* t1m.h:292-299 column major, ld br */                    \
      for (uint64_t l = 0; l < nc; ++l)                                      \
        for (uint64_t k = 0; k < nr; ++k)                                    \
          dest[k + l * c->br] = src[cell_src(c, k, l)];                      \
    } else { /* t1m.h:300-307 row major, ld bc */                            \
      for (uint64_t k = 0; k < nr; ++k)                                      \
        for (uint64_t l = 0; l < nc; ++l)                                    \
          dest[l + k * c->bc] = src[cell_src(c, k, l)];                      \
    }                                                                        \
  }
DEF_PACK_CELL_REAL(pack_cell_f, float)
DEF_PACK_CELL_REAL(pack_cell_d, double)

/* 1M packing of a complex cell; src is interleaved (re, im) of type T. */
#define DEF_PACK_CELL_1M(NAME, T)                                            \
  static void NAME(const orc_view* c, int labelA, const T* src, T* dest) {   \
    uint64_t nr = c->nrs, nc = c->ncs;                                       \
    if (labelA) { /* t1m.h:310-325 */                                        \
      for (uint64_t l = 0; l < nc; ++l)                                      \
        for (uint64_t k = 0; k < nr; ++k) {                                  \
          uint64_t s = cell_src(c, k, l);                                    \
          T re = src[2 * s], im = src[2 * s + 1];                            \
          uint64_t base = 2 * k + 2 * l * c->br;                             \
          dest[base] = re;                                                   \
          dest[base + 1] = im;                                               \
          dest[base + c->br] = -im;                                          \
          dest[base + c->br + 1] = re;                                       \
        }                                                                    \
    } else { /* t1m.h:326-337 */                                             \
      for (uint64_t k = 0; k < nr; ++k)                                      \
        for (uint64_t l = 0; l < nc; ++l) {                                  \
          uint64_t s = cell_src(c, k, l);                                    \
          uint64_t base = l + 2 * k * c->bc;                                 \
          dest[base] = src[2 * s];                                           \
          dest[base + c->bc] = src[2 * s + 1];                               \
        }                                                                    \
    }                                                                        \
  }
DEF_PACK_CELL_1M(pack_cell_c, float)
DEF_PACK_CELL_1M(pack_cell_z, double)

static void pack_cell_any(int dtype, const orc_view* c, int labelA,
                          const void* src, void* dest) {
  switch (dtype) {
    case 0: pack_cell_f(c, labelA, (const float*)src, (float*)dest); break;
    case 1: pack_cell_d(c, labelA, (const double*)src, (double*)dest); break;
    case 2: pack_cell_c(c, labelA, (const float*)src, (float*)dest); break;
    default: pack_cell_z(c, labelA, (const double*)src, (double*)dest); break;
  }
}

static size_t real_size(int dtype) { return (dtype == 0 || dtype == 2) ? 4 : 8; }

/* t1m.h:341-404.  dest offsets are in real (underlying) elements. */
static void pack_block_v(int dtype, int labelA, const orc_view* blk,
                         uint64_t length, const void* src, void* dest) {
  uint64_t nr = blk->nrs, nc = blk->ncs;
  char* d = (char*)dest;
  size_t rsz = real_size(dtype);
  if (dtype < 2) {
    if (labelA) {                                   /* t1m.h:346-357 */
      for (uint64_t r = 0; r < nr; r += blk->br)
        for (uint64_t c = 0; c < nc; c += blk->bc) {
          uint64_t off = c * blk->br + r * length;
          uint64_t cr = blk->br < nr - r ? blk->br : nr - r;
          uint64_t cc = blk->bc < nc - c ? blk->bc : nc - c;
          orc_view cell = subview(*blk, r, c, cr, cc);
          pack_cell_any(dtype, &cell, 1, src, d + off * rsz);
        }
    } else {                                        /* t1m.h:358-370 */
      for (uint64_t c = 0; c < nc; c += blk->bc)
        for (uint64_t r = 0; r < nr; r += blk->br) {
          uint64_t off = r * blk->bc + c * length;
          uint64_t cr = blk->br < nr - r ? blk->br : nr - r;
          uint64_t cc = blk->bc < nc - c ? blk->bc : nc - c;
          orc_view cell = subview(*blk, r, c, cr, cc);
          pack_cell_any(dtype, &cell, 0, src, d + off * rsz);
        }
    }
  } else {
    uint64_t hbr = blk->br / 2, hbc = blk->bc / 2;  /* t1m.h:372-373 */
    if (labelA) {                                   /* t1m.h:375-389 */
      for (uint64_t r = 0; r < nr; r += hbr)
        for (uint64_t c = 0; c < nc; c += hbc) {
          uint64_t off = 2 * (c * blk->br + r * length);
          uint64_t cr = hbr < nr - r ? hbr : nr - r;
          uint64_t cc = hbc < nc - c ? hbc : nc - c;
          orc_view cell = subview(*blk, r, c, cr, cc);
          pack_cell_any(dtype, &cell, 1, src, d + off * rsz);
        }
    } else {                                        /* t1m.h:390-402 */
      for (uint64_t c = 0; c < nc; c += blk->bc)
        for (uint64_t r = 0; r < nr; r += hbr) {
          uint64_t off = 2 * r * blk->bc + c * length;
          uint64_t cr = hbr < nr - r ? hbr : nr - r;
          uint64_t cc = blk->bc < nc - c ? blk->bc : nc - c;
          orc_view cell = subview(*blk, r, c, cr, cc);
          pack_cell_any(dtype, &cell, 0, src, d + off * rsz);
        }
    }
  }
}

/* ctypes entry: a view given by its four vectors and two block sizes. */
void orc_pack_block(int dtype, int labelA, uint64_t br, const uint64_t* rbs,
                    uint64_t nrbs, const uint64_t* rs, uint64_t nrs, uint64_t bc,
                    const uint64_t* cbs, uint64_t ncbs, const uint64_t* cs,
                    uint64_t ncs, uint64_t length, const void* src, void* dest) {
  orc_view v = {br, rbs, nrbs, rs, nrs, bc, cbs, ncbs, cs, ncs};
  pack_block_v(dtype, labelA, &v, length, src, dest);
}

/* t1m.h:406-430: dest[idx] += tile element (complex from 1M pairs). */
static void unpack_v(int dtype, const orc_view* blk, const void* src,
                     void* dest) {
  uint64_t nr = blk->nrs, nc = blk->ncs;
  for (uint64_t l = 0; l < nc; ++l)
    for (uint64_t k = 0; k < nr; ++k) {
      uint64_t di = cell_src(blk, k, l);
      switch (dtype) {
        case 0: ((float*)dest)[di] += ((const float*)src)[k + l * blk->br]; break;
        case 1: ((double*)dest)[di] += ((const double*)src)[k + l * blk->br]; break;
        case 2: {
          uint64_t si = 2 * k + l * blk->br;
          ((float*)dest)[2 * di] += ((const float*)src)[si];
          ((float*)dest)[2 * di + 1] += ((const float*)src)[si + 1];
        } break;
        default: {
          uint64_t si = 2 * k + l * blk->br;
          ((double*)dest)[2 * di] += ((const double*)src)[si];
          ((double*)dest)[2 * di + 1] += ((const double*)src)[si + 1];
        } break;
      }
    }
}

void orc_unpack(int dtype, uint64_t br, const uint64_t* rbs, uint64_t nrbs,
                const uint64_t* rs, uint64_t nrs, uint64_t bc,
                const uint64_t* cbs, uint64_t ncbs, const uint64_t* cs,
                uint64_t ncs, const void* src, void* dest) {
  orc_view v = {br, rbs, nrbs, rs, nrs, bc, cbs, ncbs, cs, ncs};
  unpack_v(dtype, &v, src, dest);
}

/* ---------------------------------------------------------------------- */
/* Five-loop contraction: t1m.h:432-557 with a plain C micro-kernel in the */
/* place of bli_?gemm_ukernel (C_tile := alpha * A_panel * B_panel, the    */
/* tile pre-zeroed, so beta never matters: t1m.h:497-502, 543-548).        */
/* ---------------------------------------------------------------------- */

typedef struct {
  int32_t rank;
  const uint64_t* dims;
  const uint64_t* strides;   /* elements */
} orc_tensor;

typedef struct {
  uint64_t *rs, *cs, *rbs, *cbs;
  uint64_t nrs, ncs, nrbs, ncbs, br, bc;
} orc_layout;

/* t1m.h:246-268 */
static void layout_make(orc_layout* L, const orc_tensor* t, const uint64_t* X,
                        int64_t nX, const uint64_t* Y, int64_t nY, uint64_t br,
                        uint64_t bc) {
  L->nrs = orc_scatter_len(X, nX, t->dims);
  L->ncs = orc_scatter_len(Y, nY, t->dims);
  L->rs = (uint64_t*)malloc(8 * L->nrs);
  L->cs = (uint64_t*)malloc(8 * L->ncs);
  orc_scatter(X, nX, t->dims, t->strides, L->rs);
  orc_scatter(Y, nY, t->dims, t->strides, L->cs);
  L->nrbs = orc_block_scatter_len(L->nrs, br);
  L->ncbs = orc_block_scatter_len(L->ncs, bc);
  L->rbs = (uint64_t*)malloc(8 * L->nrbs);
  L->cbs = (uint64_t*)malloc(8 * L->ncbs);
  orc_block_scatter(L->rs, L->nrs, br, L->rbs);
  orc_block_scatter(L->cs, L->ncs, bc, L->cbs);
  L->br = br; L->bc = bc;
}

static void layout_free(orc_layout* L) {
  free(L->rs); free(L->cs); free(L->rbs); free(L->cbs);
}

static orc_view layout_view(const orc_layout* L) {
  orc_view v = {L->br, L->rbs, L->nrbs, L->rs, L->nrs,
                L->bc, L->cbs, L->ncbs, L->cs, L->ncs};
  return v;
}

/* Micro-kernel stand-in for bli_sgemm_ukernel / bli_dgemm_ukernel
 * (mappings.h:91-99): C[i + j*ldc] = alpha * sum_p A[i + p*MR] * B[p*NR + j]
 * for i < m, j < n; the caller has zeroed C.  The p-sum runs in order. */
#define DEF_UKERNEL(NAME, T)                                                  \
  static void NAME(uint64_t m, uint64_t n, uint64_t k, T alpha, const T* A,  \
                   const T* B, T* C, uint64_t MR, uint64_t NR) {             \
    for (uint64_t i = 0; i < m; ++i)                                          \
      for (uint64_t j = 0; j < n; ++j) {                                      \
        T acc = 0;                                                            \
        for (uint64_t p = 0; p < k; ++p) acc += A[i + p * MR] * B[p * NR + j]; \
        C[i + j * MR] = alpha * acc;                                          \
      }                                                                       \
  }
DEF_UKERNEL(ukernel_s, float)
DEF_UKERNEL(ukernel_d, double)

/* Register-blocked variant for the common full-tile case, used when the
 * tile is MRxNR = 6x16 (float) / 6x8 (double); same arithmetic order per
 * element (p ascending), vectorizable over j. */
#define DEF_UKERNEL_FULL(NAME, T, MRC, NRC)                                   \
  static void NAME(uint64_t k, T alpha, const T* restrict A,                  \
                   const T* restrict B, T* restrict C) {                      \
    T acc[MRC][NRC];                                                          \
    for (int i = 0; i < MRC; ++i)                                             \
      for (int j = 0; j < NRC; ++j) acc[i][j] = 0;                            \
    for (uint64_t p = 0; p < k; ++p) {                                        \
      const T* a = A + p * MRC;                                               \
      const T* b = B + p * NRC;                                               \
      for (int i = 0; i < MRC; ++i)                                           \
        for (int j = 0; j < NRC; ++j) acc[i][j] += a[i] * b[j];               \
    }                                                                         \
    for (int i = 0; i < MRC; ++i)                                             \
      for (int j = 0; j < NRC; ++j) C[i + j * MRC] = alpha * acc[i][j];       \
  }
DEF_UKERNEL_FULL(ukernel_s_6x16, float, 6, 16)
DEF_UKERNEL_FULL(ukernel_d_6x8, double, 6, 8)

static void ukernel(int dtype, uint64_t m, uint64_t n, uint64_t k, double alpha,
                    const void* A, const void* B, void* C, uint64_t MR,
                    uint64_t NR) {
  if (dtype == 0 || dtype == 2) {
    if (m == 6 && n == 16 && MR == 6 && NR == 16)
      ukernel_s_6x16(k, (float)alpha, (const float*)A, (const float*)B, (float*)C);
    else
      ukernel_s(m, n, k, (float)alpha, (const float*)A, (const float*)B,
                (float*)C, MR, NR);
  } else {
    if (m == 6 && n == 8 && MR == 6 && NR == 8)
      ukernel_d_6x8(k, alpha, (const double*)A, (const double*)B, (double*)C);
    else
      ukernel_d(m, n, k, alpha, (const double*)A, (const double*)B,
                (double*)C, MR, NR);
  }
}

/*
 * Whole contraction, t1m.h:432-557.  bs = {MR, NR, KP, MC, KC, NC}
 * (mappings.h:105-142; KP is 4 in the reference).  alpha is real as in the
 * reference (the complex entry point passes 1, t1m.h:576).  C += alpha*A*B.
 * Returns 0, or -1 on allocation failure (the reference throws bad_alloc,
 * t1m.h:455-458), or -2 on a bundle error.
 */
/*
 * nthreads > 1 runs loop 3 (the MC blocks of A) on that many OpenMP threads,
 * each with its own A and C buffers; the B panel is packed once per (jc, pc)
 * as in the serial loop.  Different MC blocks write disjoint rows of C.
 * The reference is serial; this is the all-cores CPU baseline.
 */
int orc_contract_threads(int dtype, double alpha, const void* a, int32_t ra,
                         const uint64_t* da, const uint64_t* sa, const char* la,
                         const void* b, int32_t rb, const uint64_t* db,
                         const uint64_t* sb, const char* lb, void* c, int32_t rc,
                         const uint64_t* dc, const uint64_t* sc, const char* lc,
                         const uint64_t* bs, int nthreads);

int orc_contract(int dtype, double alpha, const void* a, int32_t ra,
                 const uint64_t* da, const uint64_t* sa, const char* la,
                 const void* b, int32_t rb, const uint64_t* db,
                 const uint64_t* sb, const char* lb, void* c, int32_t rc,
                 const uint64_t* dc, const uint64_t* sc, const char* lc,
                 const uint64_t* bs) {
  return orc_contract_threads(dtype, alpha, a, ra, da, sa, la, b, rb, db, sb, lb, c, rc, dc,
                              sc, lc, bs, 1);
}

int orc_contract_threads(int dtype, double alpha, const void* a, int32_t ra,
                         const uint64_t* da, const uint64_t* sa, const char* la,
                         const void* b, int32_t rb, const uint64_t* db,
                         const uint64_t* sb, const char* lb, void* c, int32_t rc,
                         const uint64_t* dc, const uint64_t* sc, const char* lc,
                         const uint64_t* bs, int nthreads) {
  uint64_t MR = bs[0], NR = bs[1], KP = bs[2], MC = bs[3], KC = bs[4], NC = bs[5];
  uint64_t bund[6 * ORC_MAX_RANK];
  int64_t cnt[6];
  if (orc_bundles(la, lb, lc, bund, cnt)) return -2;
  orc_tensor ta = {ra, da, sa}, tb = {rb, db, sb}, tc = {rc, dc, sc};
  orc_layout LA, LB, LC;
  layout_make(&LA, &ta, bund + 0 * ORC_MAX_RANK, cnt[0], bund + 1 * ORC_MAX_RANK, cnt[1], MR, KP);
  layout_make(&LB, &tb, bund + 2 * ORC_MAX_RANK, cnt[2], bund + 3 * ORC_MAX_RANK, cnt[3], KP, NR);
  layout_make(&LC, &tc, bund + 4 * ORC_MAX_RANK, cnt[4], bund + 5 * ORC_MAX_RANK, cnt[5], MR, NR);
  orc_view VA = layout_view(&LA), VB = layout_view(&LB), VC = layout_view(&LC);

  size_t rsz = real_size(dtype);
  uint64_t szA = MC * KC, szB = KC * NC, szC = MR * NR;   /* t1m.h:450-453 */
  if (nthreads < 1) nthreads = 1;
  /* B panel, then one (A block, C tile) pair per thread, 64-byte aligned. */
  const size_t per = ((szA + szC) * rsz + 63) / 64 * 64, bbytes = (szB * rsz + 63) / 64 * 64;
  char* space_b = (char*)aligned_alloc(64, bbytes + per * (size_t)nthreads);
  if (!space_b) { layout_free(&LA); layout_free(&LB); layout_free(&LC); return -1; }

  uint64_t M = VA.nrs, K = VA.ncs, N = VB.ncs;              /* t1m.h:462-464 */
  int cplx = dtype >= 2;
  uint64_t kstep = cplx ? KC / 2 : KC, mstep = cplx ? MC / 2 : MC;

  for (uint64_t jc = 0; jc < N; jc += NC) {                 /* loop 5 */
    uint64_t ncn = NC < N - jc ? NC : N - jc;
    for (uint64_t pc = 0; pc < K; pc += kstep) {            /* loop 4 */
      uint64_t k = kstep < K - pc ? kstep : K - pc;
      uint64_t kreal = cplx ? 2 * k : k;
      orc_view vb = subview(VB, pc, jc, k, ncn);
      memset(space_b, 0, szB * rsz);
      pack_block_v(dtype, 0, &vb, KC, b, space_b);
      int64_t nblk = (int64_t)((M + mstep - 1) / mstep);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) if (nthreads > 1)
      for (int64_t blk = 0; blk < nblk; ++blk) {            /* loop 3 */
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        char* space_a = space_b + bbytes + per * (size_t)tid;
        char* space_c = space_a + szA * rsz;
        uint64_t ic = (uint64_t)blk * mstep;
        uint64_t mcm = mstep < M - ic ? mstep : M - ic;
        uint64_t mreal = cplx ? 2 * mcm : mcm;
        orc_view va = subview(VA, ic, pc, mcm, k);
        memset(space_a, 0, szA * rsz);
        pack_block_v(dtype, 1, &va, KC, a, space_a);
        for (uint64_t jr = 0; jr < ncn; jr += NR) {         /* loop 2 */
          const char* sliver_b = space_b + KC * jr * rsz;
          uint64_t n = NR < ncn - jr ? NR : ncn - jr;
          for (uint64_t ir = 0; ir < mreal; ir += MR) {     /* loop 1 */
            const char* sliver_a = space_a + ir * KC * rsz;
            uint64_t m = MR < mreal - ir ? MR : mreal - ir;
            uint64_t cri = cplx ? ic + ir / 2 : ic + ir;
            orc_view vc = subview(VC, cri, jc + jr, cplx ? m / 2 : m, n);
            memset(space_c, 0, szC * rsz);
            ukernel(dtype, m, n, kreal, alpha, sliver_a, sliver_b, space_c, MR, NR);
            unpack_v(dtype, &vc, space_c, c);
          }
        }
      }
    }
  }
  free(space_b);
  layout_free(&LA); layout_free(&LB); layout_free(&LC);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Naive label loops (numerics oracle, independent of bundles/packing).    */
/* C[c] += alpha * sum over contracted labels A[a] * op(B[b]).             */
/* Accumulates each output element in long double (complex).              */
/* conj_a / conj_b conjugate the operand (used for the VJP checks).        */
/* ---------------------------------------------------------------------- */

typedef long double _Complex lcplx;

static int label_pos(const char* s, char ch) {
  const char* p = strchr(s, ch);
  return p ? (int)(p - s) : -1;
}

/* dtype 0..3 as above; alpha complex (re, im). */
int orc_einsum(int dtype, double alpha_re, double alpha_im, int conj_a,
               int conj_b, const void* a, const int64_t* da, const int64_t* sa,
               const char* la, const void* b, const int64_t* db,
               const int64_t* sb, const char* lb, void* c, const int64_t* dc,
               const int64_t* sc, const char* lc) {
  /* Collect the distinct labels: C's labels first, then contracted ones. */
  char labels[ORC_MAX_RANK * 3];
  int64_t ext[ORC_MAX_RANK * 3];
  int nl = 0, nfree;
  size_t na = strlen(la), nb = strlen(lb), nc = strlen(lc);
  for (size_t i = 0; i < nc; ++i) { labels[nl] = lc[i]; ext[nl] = dc[i]; ++nl; }
  nfree = nl;
  for (size_t i = 0; i < na; ++i)
    if (label_pos(lc, la[i]) < 0) {
      if (label_pos(lb, la[i]) < 0) return -2;   /* summed single-operand label */
      int seen = 0;
      for (int t = nfree; t < nl; ++t) seen |= labels[t] == la[i];
      if (!seen) { labels[nl] = la[i]; ext[nl] = da[i]; ++nl; }
    }
  int64_t strA[ORC_MAX_RANK * 3], strB[ORC_MAX_RANK * 3], strC[ORC_MAX_RANK * 3];
  for (int t = 0; t < nl; ++t) {
    int pa = label_pos(la, labels[t]), pb = label_pos(lb, labels[t]);
    int pc = label_pos(lc, labels[t]);
    strA[t] = pa >= 0 ? sa[pa] : 0;
    strB[t] = pb >= 0 ? sb[pb] : 0;
    strC[t] = pc >= 0 ? sc[pc] : 0;
  }
  (void)nb; (void)db;
  int64_t idx[ORC_MAX_RANK * 3];
  uint64_t nout = 1, nsum = 1;
  for (int t = 0; t < nfree; ++t) nout *= (uint64_t)ext[t];
  for (int t = nfree; t < nl; ++t) nsum *= (uint64_t)ext[t];
  lcplx alpha = (long double)alpha_re + (long double)alpha_im * I;
  for (uint64_t o = 0; o < nout; ++o) {
    uint64_t r = o;
    int64_t offA = 0, offB = 0, offC = 0;
    for (int t = 0; t < nfree; ++t) {
      idx[t] = (int64_t)(r % (uint64_t)ext[t]); r /= (uint64_t)ext[t];
      offA += idx[t] * strA[t]; offB += idx[t] * strB[t]; offC += idx[t] * strC[t];
    }
    lcplx acc = 0;
    for (uint64_t s = 0; s < nsum; ++s) {
      uint64_t q = s;
      int64_t oa = offA, ob = offB;
      for (int t = nfree; t < nl; ++t) {
        int64_t x = (int64_t)(q % (uint64_t)ext[t]); q /= (uint64_t)ext[t];
        oa += x * strA[t]; ob += x * strB[t];
      }
      lcplx va, vb;
      switch (dtype) {
        case 0: va = ((const float*)a)[oa]; vb = ((const float*)b)[ob]; break;
        case 1: va = ((const double*)a)[oa]; vb = ((const double*)b)[ob]; break;
        case 2: va = (long double)((const float*)a)[2 * oa] + (long double)((const float*)a)[2 * oa + 1] * I;
                vb = (long double)((const float*)b)[2 * ob] + (long double)((const float*)b)[2 * ob + 1] * I; break;
        default: va = (long double)((const double*)a)[2 * oa] + (long double)((const double*)a)[2 * oa + 1] * I;
                 vb = (long double)((const double*)b)[2 * ob] + (long double)((const double*)b)[2 * ob + 1] * I; break;
      }
      if (conj_a) va = conjl(va);
      if (conj_b) vb = conjl(vb);
      acc += va * vb;
    }
    acc *= alpha;
    switch (dtype) {
      case 0: ((float*)c)[offC] += (float)creall(acc); break;
      case 1: ((double*)c)[offC] += (double)creall(acc); break;
      case 2: ((float*)c)[2 * offC] += (float)creall(acc);
              ((float*)c)[2 * offC + 1] += (float)cimagl(acc); break;
      default: ((double*)c)[2 * offC] += (double)creall(acc);
               ((double*)c)[2 * offC + 1] += (double)cimagl(acc); break;
    }
  }
  return 0;
}
