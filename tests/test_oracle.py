"""Pins the CPU oracle (oracle/t1m_oracle.c) before anything is checked
against it: the reference's own known-answer vectors, closed forms, brute force
and numpy library routines.  CPU only."""
import itertools
import random
import string

import numpy as np
import pytest

from oracle import oracle as orc


def _fill(kind, n, dtype):
    if kind == "iota1":
        return np.arange(1, n + 1, dtype=dtype)
    # test_packing.cpp:11-16: element i = (2i+1) + (2i+2)i
    i = np.arange(n)
    return ((2 * i + 1) + 1j * (2 * i + 2)).astype(dtype)


# ---------------------------------------------------------------- KATs ----

def test_kat_bundles(kats):
    for case in kats["bundles"]:
        got = orc.bundles(case["la"], case["lb"], case["lc"])
        for key in ("AX", "AY", "BX", "BY", "CX", "CY"):
            assert got[key] == case[key], (case["source"], key)


def test_kat_scatter(kats):
    for case in kats["scatter"]:
        dims = case["dims"]
        got = orc.scatter(case["indices"], dims, orc.col_major_strides(dims))
        assert got.tolist() == case["expected"], case["source"]


def test_kat_block_scatter(kats):
    for case in kats["block_scatter"]:
        got = orc.block_scatter(case["scat"], case["b"])
        assert got.tolist() == case["expected"], case["source"]


def test_kat_pack(kats):
    for case in kats["pack"]:
        dt = np.dtype(case["dtype"])
        src = _fill(case["fill"], case["numel"], dt)
        dims = case["dims"]
        lay = orc.Layout(dims, orc.col_major_strides(dims), case["X"], case["Y"],
                         case["br"], case["bc"])
        real = np.float32 if dt.kind == "c" else dt
        dest = np.zeros(case["dest_len"], dtype=real)
        orc.pack_block(lay, case["label"], case["length"], src, dest)
        assert dest.tolist() == [float(x) for x in case["expected"]], case["name"]


def test_kat_unpack(kats):
    for case in kats["unpack"]:
        dt = np.dtype(case["dtype"])
        src = np.arange(1, case["src_len"] + 1, dtype=np.float32)
        dims = case["dims"]
        lay = orc.Layout(dims, orc.col_major_strides(dims), case["X"], case["Y"],
                         case["br"], case["bc"])
        dest = np.zeros(case["dest_numel"], dtype=dt)
        orc.unpack(lay, src, dest)
        flat = dest.view(np.float32) if dt.kind == "c" else dest
        assert flat.tolist() == [float(x) for x in case["expected"]], case["name"]


# ------------------------------------------------------- closed forms ----

def _py_bundles(la, lb, lc):
    """Set algebra, alphabetical order, positions by first occurrence."""
    sa, sb = set(la), set(lb)
    contracted = sorted(sa & sb)
    fa, fb = sorted(sa - sb), sorted(sb - sa)
    return {"AX": [la.index(x) for x in fa], "AY": [la.index(x) for x in contracted],
            "BX": [lb.index(x) for x in contracted], "BY": [lb.index(x) for x in fb],
            "CX": [lc.index(x) for x in fa], "CY": [lc.index(x) for x in fb]}


def test_bundles_brute_force():
    rng = random.Random(11)
    for _ in range(300):
        pool = rng.sample(string.ascii_letters, rng.randint(2, 12))
        nk = rng.randint(0, len(pool) - 1)
        contracted, rest = pool[:nk], pool[nk:]
        cut = rng.randint(0, len(rest))
        fa, fb = rest[:cut], rest[cut:]
        la = "".join(rng.sample(contracted + fa, len(contracted + fa)))
        lb = "".join(rng.sample(contracted + fb, len(contracted + fb)))
        lc = "".join(rng.sample(fa + fb, len(fa + fb)))
        assert orc.bundles(la, lb, lc) == _py_bundles(la, lb, lc), (la, lb, lc)


def _mixed_radix(indices, dims, strides):
    """scat[i] = sum_j i_j * stride[X_j], first bundle mode fastest."""
    ext = [dims[x] for x in indices]
    out = []
    for flat in range(int(np.prod(ext, dtype=np.int64)) if ext else 1):
        off, r = 0, flat
        for x, e in zip(indices, ext):
            off += (r % e) * strides[x]
            r //= e
        out.append(off % (1 << 64))
    return out


def test_scatter_closed_form_and_bruteforce():
    rng = np.random.default_rng(5)
    for _ in range(200):
        nd = int(rng.integers(1, 6))
        dims = [int(x) for x in rng.integers(1, 5, nd)]
        strides = [int(x) for x in rng.integers(0, 50, nd)]
        k = int(rng.integers(1, nd + 1))
        idx = [int(x) for x in rng.permutation(nd)[:k]]
        got = orc.scatter(idx, dims, strides).tolist()
        assert got == _mixed_radix(idx, dims, strides)
    # brute force: the scatter of a col-major tensor over all modes in order
    # enumerates every element offset exactly once, in order.
    dims = [3, 4, 2]
    got = orc.scatter([0, 1, 2], dims, orc.col_major_strides(dims))
    assert got.tolist() == list(range(24))
    # a last-mode-fastest fold (a plausible mistake) fails the KAT.
    wrong = _mixed_radix([1, 0], [3, 2, 2, 3], orc.col_major_strides([3, 2, 2, 3]))
    assert wrong != [0, 1, 2, 3, 4, 5]


def _py_block_scatter(scat, b):
    out = []
    for i in range(0, len(scat), b):
        blk = [int(x) for x in scat[i:i + b]]
        if len(blk) == 1:
            out.append(0)
            continue
        d = [(blk[t] - blk[t - 1]) % (1 << 64) for t in range(1, len(blk))]
        out.append(d[-1] if all(x == d[0] for x in d) else 0)
    return out


def test_block_scatter_closed_form():
    rng = np.random.default_rng(6)
    for _ in range(300):
        n = int(rng.integers(1, 40))
        if rng.random() < 0.5:
            base, s = int(rng.integers(0, 100)), int(rng.integers(0, 9))
            scat = [base + s * t for t in range(n)]
            for t in rng.integers(0, n, int(rng.integers(0, 3))):
                scat[int(t)] += int(rng.integers(0, 5))
        else:
            scat = [int(x) for x in rng.integers(0, 60, n)]
        b = int(rng.integers(1, 9))
        assert orc.block_scatter(scat, b).tolist() == _py_block_scatter(scat, b)


def _np_pack_1m_a(A, MR, KC, length):
    """Closed form of the 1M A panel (t1m.h:310-325, 378-389) for a matrix A
    (m x k complex): real rows 2i+s, real cols 2p+t hold [[re,-im],[im,re]],
    MR-row micro-panels of KC real columns, col-major inside (ld MR)."""
    m, k = A.shape
    Ar = np.zeros((2 * m, 2 * k), dtype=A.real.dtype)
    Ar[0::2, 0::2] = A.real
    Ar[1::2, 0::2] = A.imag
    Ar[0::2, 1::2] = -A.imag
    Ar[1::2, 1::2] = A.real
    npan = -(-2 * m // MR)
    out = np.zeros(npan * MR * length, dtype=A.real.dtype)
    for pn in range(npan):
        blk = Ar[pn * MR:(pn + 1) * MR]
        for col in range(blk.shape[1]):
            for row in range(blk.shape[0]):
                out[pn * MR * length + col * MR + row] = blk[row, col]
    return out


def _np_pack_1m_b(B, NR, length):
    """Closed form of the 1M B panel (t1m.h:326-337, 390-402): real rows
    2p+t hold [re; im], NR-column micro-panels, row-major inside (ld NR)."""
    k, n = B.shape
    Br = np.zeros((2 * k, n), dtype=B.real.dtype)
    Br[0::2] = B.real
    Br[1::2] = B.imag
    npan = -(-n // NR)
    out = np.zeros(npan * NR * length, dtype=B.real.dtype)
    for pn in range(npan):
        blk = Br[:, pn * NR:(pn + 1) * NR]
        for row in range(blk.shape[0]):
            for col in range(blk.shape[1]):
                out[pn * NR * length + row * NR + col] = blk[row, col]
    return out


@pytest.mark.parametrize("seed", range(6))
def test_pack_1m_closed_form(seed):
    rng = np.random.default_rng(seed)
    # a random col-major 4-mode tensor matricised with a random bundle split
    dims = [int(x) for x in rng.integers(1, 5, 4)]
    strides = orc.col_major_strides(dims)
    perm = [int(x) for x in rng.permutation(4)]
    cut = int(rng.integers(1, 4))
    X, Y = perm[:cut], perm[cut:]
    T = (rng.standard_normal(int(np.prod(dims))) +
         1j * rng.standard_normal(int(np.prod(dims)))).astype(np.complex64)
    MR, KP, NR = 2 * int(rng.integers(1, 4)), 4, int(rng.integers(1, 5))
    lay_a = orc.Layout(dims, strides, X, Y, MR, KP)
    m, k = len(lay_a.rs), len(lay_a.cs)
    A = T[(lay_a.rs[:, None] + lay_a.cs[None, :]).astype(np.int64)]
    length = 2 * k + 2 * int(rng.integers(0, 3))     # KC >= 2k (one k block)
    dest = np.zeros(-(-2 * m // MR) * MR * length, dtype=np.float32)
    orc.pack_block(lay_a, "A", length, T, dest)
    np.testing.assert_array_equal(dest, _np_pack_1m_a(A, MR, length // 2 * 2, length))
    # B side: the same tensor, rows = X (the k index), cols = Y
    lay_b = orc.Layout(dims, strides, X, Y, KP, NR)
    B = T[(lay_b.rs[:, None] + lay_b.cs[None, :]).astype(np.int64)]
    kk, n = B.shape
    length = 2 * kk
    dest = np.zeros(-(-n // NR) * NR * length, dtype=np.float32)
    orc.pack_block(lay_b, "B", length, T, dest)
    np.testing.assert_array_equal(dest, _np_pack_1m_b(B, NR, length))


def test_unpack_closed_form():
    rng = np.random.default_rng(3)
    dims = [3, 5, 2]
    strides = orc.col_major_strides(dims)
    lay = orc.Layout(dims, strides, [0, 2], [1], 12, 5)   # 6 complex rows x 5
    tile = rng.standard_normal(12 * 5).astype(np.float32)  # col-major, ld 12
    C = (rng.standard_normal(30) + 1j * rng.standard_normal(30)).astype(np.complex64)
    want = C.copy()
    for l in range(5):
        for k in range(6):
            want[lay.rs[k] + lay.cs[l]] += tile[2 * k + 12 * l] + 1j * tile[2 * k + 1 + 12 * l]
    orc.unpack(lay, tile, C)
    np.testing.assert_array_equal(C, want)


# ----------------------------------------------------------- numerics ----

def _rand(rng, shape, dtype):
    x = rng.uniform(-1, 1, shape)
    if np.dtype(dtype).kind == "c":
        x = x + 1j * rng.uniform(-1, 1, shape)
    return np.asfortranarray(x.astype(dtype))


GEMM_LIKE = [("ab", "bc", "ac"), ("abc", "bce", "ae"), ("abc", "cd", "abd"),
             ("abc", "cde", "abde"), ("abcd", "cde", "abe"), ("abcd", "cdef", "abef")]


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64, np.complex128])
@pytest.mark.parametrize("labels", GEMM_LIKE)
def test_contract_gemm_like(dtype, labels):
    """test_contraction.cpp:33-112 pattern: GEMM-ordered labels make the
    contraction a plain GEMM of the tensor memory; checked with numpy."""
    la, lb, lc = labels
    rng = np.random.default_rng(len(la) * 10 + len(lc))
    for d in (2, 3, 7):
        A = _rand(rng, (d,) * len(la), dtype)
        B = _rand(rng, (d,) * len(lb), dtype)
        C = np.zeros((d,) * len(lc), dtype=dtype, order="F")
        orc.contract(A, la, B, lb, C, lc)
        m = d ** (len(la) - len(set(la) & set(lb)))
        k = d ** len(set(la) & set(lb))
        n = d ** (len(lb) - len(set(la) & set(lb)))
        ref = (A.reshape(m, k, order="F").astype(np.complex128) @
               B.reshape(k, n, order="F").astype(np.complex128))
        got = C.reshape(m, n, order="F")
        eps = np.finfo(np.dtype(dtype).type(0).real.dtype).eps
        err = np.linalg.norm(got - ref) / (eps * d * np.linalg.norm(A) * np.linalg.norm(B))
        assert err <= 1.0


PERMUTED = [("abcd", "dbef", "acef"), ("asb", "btc", "astc"), ("abc", "bd", "acd"),
            ("efbad", "cf", "abcde"), ("dega", "gfbc", "abcdef"), ("aebf", "fdec", "abcd")]


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128, np.float32, np.float64])
@pytest.mark.parametrize("labels", PERMUTED)
def test_contract_permuted_vs_einsum(dtype, labels):
    la, lb, lc = labels
    rng = np.random.default_rng(sum(map(ord, la + lb)))
    ext = {ch: int(rng.integers(2, 6)) for ch in set(la + lb)}
    A = _rand(rng, [ext[x] for x in la], dtype)
    B = _rand(rng, [ext[x] for x in lb], dtype)
    C0 = _rand(rng, [ext[x] for x in lc], dtype)
    C = C0.copy(order="F")
    orc.contract(A, la, B, lb, C, lc)
    ref = C0.astype(np.complex128) + np.einsum(f"{la},{lb}->{lc}", A.astype(np.complex128),
                                               B.astype(np.complex128))
    tol = 1e-5 if np.dtype(dtype) in (np.float32, np.complex64) else 1e-12
    assert np.linalg.norm(C - ref) / np.linalg.norm(ref) < tol
    # the naive label loops agree too
    C2 = C0.copy(order="F")
    orc.einsum(A, la, B, lb, C2, lc)
    assert np.linalg.norm(C2 - ref) / np.linalg.norm(ref) < tol


def test_contract_small_blocks_hit_every_edge():
    """Tiny block sizes force partial micro-panels and many KC/MC chunks.
    As in BLIS, NC is a multiple of NR and MC of MR (the reference relies on
    it: a last B panel is KC x NR wide, t1m.h:390-402)."""
    rng = np.random.default_rng(9)
    A = _rand(rng, (5, 3, 7), np.complex128)
    B = _rand(rng, (7, 3, 4), np.complex128)
    C = np.zeros((5, 4), dtype=np.complex128, order="F")
    orc.contract(A, "abc", B, "cbd", C, "ad", block_sizes=(4, 3, 4, 8, 8, 6))
    ref = np.einsum("abc,cbd->ad", A, B)
    np.testing.assert_allclose(C, ref, rtol=1e-12, atol=1e-12)


def test_einsum_conj_and_alpha():
    rng = np.random.default_rng(4)
    A = _rand(rng, (3, 4), np.complex128)
    B = _rand(rng, (4, 5), np.complex128)
    for ca, cb in itertools.product([False, True], repeat=2):
        C = np.zeros((3, 5), dtype=np.complex128, order="F")
        orc.einsum(A, "ab", B, "bc", C, "ac", alpha=0.5 - 2j, conj_a=ca, conj_b=cb)
        ref = (0.5 - 2j) * ((A.conj() if ca else A) @ (B.conj() if cb else B))
        np.testing.assert_allclose(C, ref, rtol=1e-13)
    # tiny brute force by hand loops
    C = np.zeros((3, 5), dtype=np.complex128, order="F")
    orc.einsum(A, "ab", B, "bc", C, "ac")
    for i in range(3):
        for j in range(5):
            s = sum(A[i, p] * B[p, j] for p in range(4))
            assert abs(C[i, j] - s) < 1e-13


@pytest.mark.parametrize("labels", [("zab", "zbc", "zac"), ("azbw", "wbcz", "cazw"),
                                    ("ab", "ab", "ab"), ("abk", "bk", "ab")])
def test_einsum_batch_labels(labels):
    """Labels in A, B and C are batch modes (einsum semantics): the oracle's
    label loops against numpy.einsum, and a brute-force loop for the first."""
    la, lb, lc = labels
    rng = np.random.default_rng(11)
    ext = {ch: int(rng.integers(2, 5)) for ch in set(la + lb + lc)}
    A = _rand(rng, [ext[x] for x in la], np.complex128)
    B = _rand(rng, [ext[x] for x in lb], np.complex128)
    C = np.zeros([ext[x] for x in lc], dtype=np.complex128, order="F")
    orc.einsum(A, la, B, lb, C, lc)
    np.testing.assert_allclose(C, np.einsum(f"{la},{lb}->{lc}", A, B), rtol=1e-13, atol=1e-13)
    if labels[0] == "zab":
        for z in range(ext["z"]):
            np.testing.assert_allclose(C[z], A[z] @ B[z], rtol=1e-13)


def test_contract_threads_match_serial():
    """The all-cores variant of the five-loop driver (loop 3 over threads)
    gives the serial result bit for bit: each C element is accumulated by
    one thread, in the same order."""
    rng = np.random.default_rng(21)
    for dt in (np.complex64, np.float64):
        A = _rand(rng, (40, 7, 9), dt)
        B = _rand(rng, (9, 7, 30), dt)
        C1 = np.zeros((40, 30), dtype=dt, order="F")
        C4 = np.zeros((40, 30), dtype=dt, order="F")
        bs = (4, 8, 4, 12, 8, 16)
        orc.contract(A, "abc", B, "cbd", C1, "ad", block_sizes=bs)
        orc.contract(A, "abc", B, "cbd", C4, "ad", block_sizes=bs, threads=4)
        np.testing.assert_array_equal(C1, C4)
        np.testing.assert_allclose(C1, np.einsum("abc,cbd->ad", A, B), rtol=1e-4)
