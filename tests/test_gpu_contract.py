"""GPU parity tests: the HIP path through the C ABI against the CPU oracle
(small sizes), torch complex128 / Freivalds properties (full sizes) and the
reference's golden vectors.  Bar: index maps and packing bit-exact; values
within the relative Frobenius tolerance written in each test."""
import contextlib
import ctypes
import os
import subprocess

import numpy as np
import pytest
import torch

import t1m_amd
from oracle import oracle as orc
from util import extents, rand, rel_err, to_dev, to_host

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# Relative Frobenius tolerances against a long-double / complex128 reference.
TOL = {"fp32": 2e-6, "bf16x3": 1e-5, "fp64": 1e-13, "simple32": 2e-6}


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.nodeid}", flush=True)
    yield
    torch.cuda.synchronize()


@contextlib.contextmanager
def forced(kernel=None, idx64=False):
    old_k, old_i = os.environ.get("T1M_FORCE_KERNEL"), os.environ.get("T1M_FORCE_IDX64")
    try:
        if kernel is not None:
            os.environ["T1M_FORCE_KERNEL"] = str(t1m_amd.KERNELS[kernel])
        if idx64:
            os.environ["T1M_FORCE_IDX64"] = "1"
        yield
    finally:
        for k, v in (("T1M_FORCE_KERNEL", old_k), ("T1M_FORCE_IDX64", old_i)):
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def u64(t):
    return t.cpu().numpy().view(np.uint64)


# ------------------------------------------------------------ index maps --

MAP_CASES = [
    # (dims, strides or None = col-major, modes, block)
    ([3, 2, 2, 3], None, [0, 1], 3),                 # test_scatter.cpp:31-53
    ([3, 2, 2, 3], None, [2, 3], 3),
    ([48, 48, 48, 48], None, [0, 2], 6),             # A rows at 48^4 (a, c)
    ([48, 48, 48, 48], None, [1, 3], 4),             # A cols (b, d), KP = 4
    ([48, 48, 48, 48], None, [1, 0], 4),             # B rows of 'dbef' (b, d)
    ([48, 48, 48, 48], None, [2, 3], 16),            # B cols (e, f), NR = 16
    ([256, 4, 4, 256], None, [3, 2], 8),             # C cols of 'astc' (c, t)
    ([7, 5, 3], [15, 1, 5], [0, 2, 1], 4),           # row-major-ish strides
    ([5, 4, 3], [-1, 5, 40], [0, 1], 3),             # negative stride: wraps
    ([6, 3], [0, 7], [0, 1], 5),                     # zero stride
    ([4, 5], None, [], 2),                           # empty bundle -> {0}
]


@pytest.mark.parametrize("case", range(len(MAP_CASES)))
def test_index_maps_bit_exact(case):
    dims, strides, modes, block = MAP_CASES[case]
    scat, bscat = t1m_amd.index_maps(dims, strides, modes, block)
    st = strides if strides is not None else orc.col_major_strides(dims)
    want = orc.scatter(modes, dims, [s % (1 << 64) for s in st])
    np.testing.assert_array_equal(u64(scat), want)
    np.testing.assert_array_equal(u64(bscat), orc.block_scatter(want, block))


def test_index_maps_random_bit_exact():
    rng = np.random.default_rng(17)
    for _ in range(60):
        nd = int(rng.integers(1, 6))
        dims = [int(x) for x in rng.integers(1, 7, nd)]
        strides = [int(x) for x in rng.integers(-30, 200, nd)]
        modes = [int(x) for x in rng.permutation(nd)[:int(rng.integers(1, nd + 1))]]
        block = int(rng.integers(1, 9))
        scat, bscat = t1m_amd.index_maps(dims, strides, modes, block)
        want = orc.scatter(modes, dims, [s % (1 << 64) for s in strides])
        np.testing.assert_array_equal(u64(scat), want)
        np.testing.assert_array_equal(u64(bscat), orc.block_scatter(want, block))


# --------------------------------------------------- reference packing ----

def _fill(kind, n, dtype):
    if kind == "iota1":
        return np.arange(1, n + 1, dtype=dtype)
    i = np.arange(n)
    return ((2 * i + 1) + 1j * (2 * i + 2)).astype(dtype)


def _dev_layout(dims, X, Y, br, bc):
    rs, rbs = t1m_amd.index_maps(dims, None, X, br)
    cs, cbs = t1m_amd.index_maps(dims, None, Y, bc)
    return rs, cs, rbs, cbs


DT = {"float32": t1m_amd.FLOAT32, "complex64": t1m_amd.COMPLEX64}


def test_pack_kats_on_device(kats):
    for case in kats["pack"]:
        dt = np.dtype(case["dtype"])
        src = torch.from_numpy(_fill(case["fill"], case["numel"], dt)).cuda()
        rs, cs, rbs, cbs = _dev_layout(case["dims"], case["X"], case["Y"], case["br"], case["bc"])
        dest = torch.zeros(case["dest_len"], dtype=torch.float32, device="cuda")
        t1m_amd.pack_block(DT[case["dtype"]], case["label"], case["br"], case["bc"],
                           rs, cs, rbs, cbs, case["length"], src, dest)
        assert dest.cpu().tolist() == [float(x) for x in case["expected"]], case["name"]


def test_unpack_kats_on_device(kats):
    for case in kats["unpack"]:
        src = torch.arange(1, case["src_len"] + 1, dtype=torch.float32, device="cuda")
        rs, cs, rbs, cbs = _dev_layout(case["dims"], case["X"], case["Y"], case["br"], case["bc"])
        tdt = torch.complex64 if case["dtype"] == "complex64" else torch.float32
        dest = torch.zeros(case["dest_numel"], dtype=tdt, device="cuda")
        t1m_amd.unpack(DT[case["dtype"]], case["br"], case["bc"], rs, cs, rbs, cbs, src, dest)
        flat = torch.view_as_real(dest).flatten() if dest.is_complex() else dest
        assert flat.cpu().tolist() == [float(x) for x in case["expected"]], case["name"]


@pytest.mark.parametrize("seed", range(8))
def test_pack_1m_random_views_bit_exact(seed):
    """Device 1M / real panels against the oracle's pack_block on random
    tensors, bundle splits and block sizes (odd sizes hit partial cells)."""
    rng = np.random.default_rng(100 + seed)
    dims = [int(x) for x in rng.integers(1, 6, 4)]
    perm = [int(x) for x in rng.permutation(4)]
    cut = int(rng.integers(1, 4))
    X, Y = perm[:cut], perm[cut:]
    n = int(np.prod(dims))
    cplx = seed % 2 == 0
    host = (rng.standard_normal(n) + (1j * rng.standard_normal(n) if cplx else 0)).astype(
        np.complex64 if cplx else np.float32)
    for label in ("A", "B"):
        br = 2 * int(rng.integers(1, 4)) if (cplx or label == "A") else int(rng.integers(1, 6))
        bc = 4 if label == "A" else int(rng.integers(1, 6))
        if label == "B":
            br, bc = 4, bc
        lay = orc.Layout(dims, orc.col_major_strides(dims), X, Y, br, bc)
        m, k = len(lay.rs), len(lay.cs)
        length = (2 if cplx else 1) * (k if label == "A" else m) + int(rng.integers(0, 3))
        npan = -(-(2 * m if cplx else m) // br) if label == "A" else -(-k // bc)
        size = npan * (br if label == "A" else bc) * length + 64
        want = np.zeros(size, dtype=np.float32)
        orc.pack_block(lay, label, length, host, want)
        rs, cs, rbs, cbs = _dev_layout(dims, X, Y, br, bc)
        got = torch.zeros(size, dtype=torch.float32, device="cuda")
        t1m_amd.pack_block(t1m_amd.COMPLEX64 if cplx else t1m_amd.FLOAT32, label, br, bc,
                           rs, cs, rbs, cbs, length, torch.from_numpy(host).cuda(), got)
        np.testing.assert_array_equal(got.cpu().numpy(), want)


# ----------------------------------------------------------- contraction --

def run(a_np, la, b_np, lb, c_np, lc, **kw):
    a, b, c = to_dev(a_np), to_dev(b_np), to_dev(c_np)
    t1m_amd.contract(a, la, b, lb, c, lc, **kw)
    return to_host(c)


def oracle_einsum(a, la, b, lb, c0, lc, alpha=1.0, conj_a=False, conj_b=False):
    c = c0.copy(order="F")
    orc.einsum(a, la, b, lb, c, lc, alpha=alpha, conj_a=conj_a, conj_b=conj_b)
    return c


def test_readme_example_c1_all_paths():
    """A[10,10,10] 'abc' x B[10,10] 'bd' -> C[10,10,10] 'acd' (example.cpp)."""
    rng = np.random.default_rng(1)
    A, B = rand(rng, (10, 10, 10), np.complex64), rand(rng, (10, 10), np.complex64)
    C0 = np.zeros((10, 10, 10), np.complex64, order="F")
    want = C0.copy(order="F")
    orc.contract(A, "abc", B, "bd", want, "acd")      # the reference's 5 loops
    exact = oracle_einsum(A, "abc", B, "bd", C0, "acd")
    assert rel_err(want, exact) < 1e-6
    for kern, prec, tol in [("c64_f32mfma_144x144", "native", TOL["fp32"]),
                            ("c64_f32mfma_64x64", "native", TOL["fp32"]),
                            ("c64_bf16x3_144x144", "bf16x3", TOL["bf16x3"]),
                            ("c64_bf16x3_64x64", "bf16x3", TOL["bf16x3"]),
                            ("simple", "native", TOL["simple32"])]:
        with forced(kern):
            got = run(A, "abc", B, "bd", C0, "acd", precision=prec)
        assert rel_err(got, exact) < tol, kern
        assert rel_err(got, want) < 2 * tol, kern


RAGGED = [
    # labels, extents (M, N, K all ragged w.r.t. 144 / 64 / 32 tiles)
    (("abcd", "dbef", "acef"), {"a": 7, "b": 5, "c": 23, "d": 13, "e": 9, "f": 17}),
    (("asb", "btc", "astc"), {"a": 37, "s": 4, "b": 19, "t": 4, "c": 29}),
    (("efbad", "cf", "abcde"), {"a": 6, "b": 5, "c": 11, "d": 4, "e": 7, "f": 9}),
    (("ab", "bc", "ac"), {"a": 300, "b": 3, "c": 290}),
    (("ba", "cb", "ca"), {"a": 150, "b": 41, "c": 70}),
]

KERNELS_C64 = [("c64_f32mfma_144x144", "native", "fp32"), ("c64_f32mfma_64x64", "native", "fp32"),
               ("c64_f32mfma_128x16", "native", "fp32"),
               ("c64_bf16x3_144x144", "bf16x3", "bf16x3"), ("c64_bf16x3_64x64", "bf16x3", "bf16x3"),
               ("simple", "native", "simple32")]
KERNELS_C128 = [("c128_f64mfma_64x64", "native", "fp64"), ("c128_f64mfma_32x32", "native", "fp64"),
                ("c128_f64mfma_128x16", "native", "fp64"),
                ("c128_f64mfma_64x64w16", "native", "fp64"),
                ("simple", "native", "fp64")]


@pytest.mark.parametrize("idx64", [False, True])
@pytest.mark.parametrize("case", range(len(RAGGED)))
def test_ragged_all_kernels(case, idx64):
    (la, lb, lc), ext = RAGGED[case]
    rng = np.random.default_rng(case)
    for dtype, kernels in ((np.complex64, KERNELS_C64), (np.complex128, KERNELS_C128)):
        A = rand(rng, [ext[x] for x in la], dtype)
        B = rand(rng, [ext[x] for x in lb], dtype)
        C0 = rand(rng, [ext[x] for x in lc], dtype)
        want = oracle_einsum(A, la, B, lb, C0, lc)
        for kern, prec, tolkey in kernels:
            with forced(kern, idx64):
                got = run(A, la, B, lb, C0, lc, precision=prec)
            err = rel_err(got - C0, want - C0)
            print(f"  {la},{lb}->{lc} {np.dtype(dtype).name} {kern} idx64={idx64} rel={err:.2e}")
            assert err < TOL[tolkey], (kern, err)


@pytest.mark.parametrize("layout", [("ik", "kj"), ("ki", "kj"), ("ik", "jk"), ("ki", "jk")])
def test_144_tiles_every_stage_count(layout):
    """The 144 x 144 FP32 tiles keep two stages of global loads in flight
    (pairs of stages per loop trip, one stage left over or not): K of 1 to 6
    stages of 16, ragged or whole, for each pair of gather directions."""
    la, lb = layout
    rng = np.random.default_rng(11)
    for k in (1, 15, 16, 17, 32, 47, 48, 64, 80, 96):
        ext = {"i": 150, "j": 161, "k": k}
        A = rand(rng, [ext[x] for x in la], np.complex64)
        B = rand(rng, [ext[x] for x in lb], np.complex64)
        C0 = rand(rng, [ext[x] for x in "ij"], np.complex64)
        want = oracle_einsum(A, la, B, lb, C0, "ij")
        with forced("c64_f32mfma_144x144"):
            got = run(A, la, B, lb, C0, "ij")
        err = rel_err(got - C0, want - C0)
        print(f"  {la},{lb}->ij k={k} rel={err:.2e}")
        assert err < TOL["fp32"], (k, err)


@pytest.mark.parametrize("layout", [("ik", "kj"), ("ki", "kj"), ("ik", "jk"), ("ki", "jk")])
def test_bf16x3_144_every_gather_direction(layout):
    """The 144 x 144 BF16x3 tiles stage an operand gathered along its rows in
    16-byte quads into a plane of 16-row chunk-major units and one gathered
    along k in words into a plane of rows; the MFMA waves read each plane in
    its own layout.  Every pair of gather directions (ik / ki: A along i / k;
    kj / jk: B along k / j), edge tiles in i and j, K of one to seven stages,
    ragged or whole, against the oracle."""
    la, lb = layout
    rng = np.random.default_rng(13)
    for k in (1, 16, 37, 100):
        ext = {"i": 150, "j": 161, "k": k}
        A = rand(rng, [ext[x] for x in la], np.complex64)
        B = rand(rng, [ext[x] for x in lb], np.complex64)
        C0 = rand(rng, [ext[x] for x in "ij"], np.complex64)
        want = oracle_einsum(A, la, B, lb, C0, "ij")
        with forced("c64_bf16x3_144x144"):
            got = run(A, la, B, lb, C0, "ij", precision="bf16x3")
        err = rel_err(got - C0, want - C0)
        print(f"  {la},{lb}->ij k={k} rel={err:.2e}")
        assert err < TOL["bf16x3"], (k, err)


REAL_KERNELS = {np.float32: ["f32mfma_128x128", "f32mfma_64x64", "simple"],
                np.float64: ["f64mfma_64x64", "f64mfma_32x32", "simple"]}


@pytest.mark.parametrize("idx64", [False, True])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_real_contractions(dtype, idx64):
    """Real contract(alpha, ...) (t1m.h:560-568): C += alpha * A . B, every
    real kernel on ragged shapes."""
    rng = np.random.default_rng(5)
    for (la, lb, lc), ext in RAGGED:
        A = rand(rng, [ext[x] for x in la], dtype)
        B = rand(rng, [ext[x] for x in lb], dtype)
        C0 = rand(rng, [ext[x] for x in lc], dtype)
        want = oracle_einsum(A, la, B, lb, C0, lc, alpha=0.75)
        for kern in REAL_KERNELS[dtype]:
            with forced(kern, idx64):
                got = run(A, la, B, lb, C0, lc, alpha=0.75)
            err = rel_err(got - C0, want - C0)
            assert err < (2e-6 if dtype == np.float32 else 1e-13), (la, lb, lc, kern, err)


SPLIT_KERNELS = [(np.complex64, "c64_f32mfma_144x144", "native", "fp32"),
                 (np.complex64, "c64_f32mfma_64x64", "native", "fp32"),
                 (np.complex64, "c64_bf16x3_64x64", "bf16x3", "bf16x3"),
                 (np.complex64, "c64_bf16x3_144x144", "bf16x3", "bf16x3"),
                 (np.complex128, "c128_f64mfma_32x32", "native", "fp64"),
                 (np.complex128, "c128_f64mfma_64x64", "native", "fp64"),
                 (np.float32, "f32mfma_128x128", "native", "fp32"),
                 (np.float64, "f64mfma_32x32", "native", "fp64")]


@pytest.mark.parametrize("dtype,kern,prec,tolkey", SPLIT_KERNELS)
def test_split_k_small_output_long_k(dtype, kern, prec, tolkey):
    """cad,dcb->ab-like (benchmark.cpp:163-166): a C of one tile and K in the
    thousands splits K over workgroups (atomic accumulation into C); the
    deterministic option keeps one workgroup per tile and is bitwise
    reproducible."""
    rng = np.random.default_rng(30)
    la, lb, lc = "cad", "dcb", "ab"
    ext = {"a": 40, "b": 37, "c": 61, "d": 70}
    A = rand(rng, [ext[x] for x in la], dtype)
    B = rand(rng, [ext[x] for x in lb], dtype)
    C0 = rand(rng, [ext[x] for x in lc], dtype)
    want = oracle_einsum(A, la, B, lb, C0, lc, alpha=0.5 if np.dtype(dtype).kind == "f" else 0.5 + 1j)
    alpha = 0.5 if np.dtype(dtype).kind == "f" else 0.5 + 1j
    with forced(kern):
        got = run(A, la, B, lb, C0, lc, precision=prec, alpha=alpha)
        assert rel_err(got - C0, want - C0) < TOL[tolkey], "split"
        d1 = run(A, la, B, lb, C0, lc, precision=prec, alpha=alpha, deterministic=True)
        d2 = run(A, la, B, lb, C0, lc, precision=prec, alpha=alpha, deterministic=True)
    assert rel_err(d1 - C0, want - C0) < TOL[tolkey], "deterministic"
    np.testing.assert_array_equal(d1, d2)


# Batch modes: labels in A, B and C, C[..z..] = sum_k A[..z..] B[..z..]
# (einsum semantics; the reference leaves such labels undefined).
BATCH_CASES = [
    (("zab", "zbc", "zac"), {"z": 5, "a": 37, "b": 19, "c": 29}),            # batched GEMM
    (("abz", "bcz", "acz"), {"z": 3, "a": 70, "b": 33, "c": 41}),            # batch mode slowest
    (("azbw", "wbcz", "cazw"), {"z": 2, "w": 3, "a": 23, "b": 9, "c": 17}),  # two, interleaved
    (("ab", "ab", "ab"), {"a": 31, "b": 7}),                                 # Hadamard product
    (("abk", "bk", "ab"), {"a": 20, "b": 6, "k": 5}),                        # batch, no free B mode
    (("zab", "zbc", "zac"), {"z": 2, "a": 40, "b": 2000, "c": 50}),          # batch + split K
]
REAL_BATCH = {np.float32: [(k, "native", "fp32") for k in REAL_KERNELS[np.float32]],
              np.float64: [(k, "native", "fp64") for k in REAL_KERNELS[np.float64]]}


@pytest.mark.parametrize("case", range(len(BATCH_CASES)))
def test_batch_labels_all_kernels(case):
    (la, lb, lc), ext = BATCH_CASES[case]
    rng = np.random.default_rng(40 + case)
    for dtype, kernels in ((np.complex64, KERNELS_C64), (np.complex128, KERNELS_C128),
                           (np.float32, REAL_BATCH[np.float32]),
                           (np.float64, REAL_BATCH[np.float64])):
        A = rand(rng, [ext[x] for x in la], dtype)
        B = rand(rng, [ext[x] for x in lb], dtype)
        C0 = rand(rng, [ext[x] for x in lc], dtype)
        alpha = 0.75 if np.dtype(dtype).kind == "f" else 0.5 - 0.25j
        want = oracle_einsum(A, la, B, lb, C0, lc, alpha=alpha)
        exact = np.einsum(f"{la},{lb}->{lc}", A.astype(np.complex128), B.astype(np.complex128))
        assert rel_err(want - C0, alpha * exact) < 1e-6          # oracle vs numpy
        for kern, prec, tolkey in kernels:
            with forced(kern):
                got = run(A, la, B, lb, C0, lc, precision=prec, alpha=alpha)
            err = rel_err(got - C0, want - C0)
            print(f"  {la},{lb}->{lc} {np.dtype(dtype).name} {kern} rel={err:.2e}")
            assert err < TOL[tolkey], (kern, err)


def test_batch_labels_empty_batch_and_vjp():
    rng = np.random.default_rng(47)
    A = rand(rng, (0, 5, 6), np.complex64)          # z = 0: nothing to add
    B = rand(rng, (0, 6, 7), np.complex64)
    C0 = rand(rng, (0, 5, 7), np.complex64)
    assert run(A, "zab", B, "zbc", C0, "zac").shape == (0, 5, 7)
    la, lb, lc = "azb", "bcz", "acz"
    ext = {"z": 3, "a": 6, "b": 5, "c": 7}
    A = rand(rng, [ext[x] for x in la], np.complex128)
    B = rand(rng, [ext[x] for x in lb], np.complex128)
    G = rand(rng, [ext[x] for x in lc], np.complex128)
    ga_want = oracle_einsum(G, lc, B, lb, np.zeros_like(A, order="F"), la, conj_b=True)
    gb_want = oracle_einsum(A, la, G, lc, np.zeros_like(B, order="F"), lb, conj_a=True)
    ga = to_dev(np.zeros_like(A, order="F"))
    gb = to_dev(np.zeros_like(B, order="F"))
    t1m_amd.contract_vjp(to_dev(A), la, to_dev(B), lb, to_dev(G), lc, ga, gb)
    assert rel_err(to_host(ga), ga_want) < 1e-13
    assert rel_err(to_host(gb), gb_want) < 1e-13


def test_batch_labels_in_persistent_batch():
    """Instances with batch labels are flattened into one problem per batch
    index of the persistent launch."""
    rng = np.random.default_rng(48)
    la, lb, lc = "zasb", "zbtc", "zastc"
    As, Bs, Cs, wants = [], [], [], []
    for _ in range(5):
        z, dl, db, dr = (int(x) for x in rng.choice([2, 3, 16, 24, 40], 4))
        A = rand(rng, (z, dl, 2, db), np.complex64)
        B = rand(rng, (z, db, 2, dr), np.complex64)
        C0 = rand(rng, (z, dl, 2, 2, dr), np.complex64)
        wants.append(oracle_einsum(A, la, B, lb, C0, lc) - C0)
        As.append(to_dev(A)); Bs.append(to_dev(B)); Cs.append(to_dev(C0))
        Cs[-1]._c0 = C0
    batch = t1m_amd.Batch(As, la, Bs, lb, Cs, lc)
    batch.launch()
    torch.cuda.synchronize()
    for c, w in zip(Cs, wants):
        assert rel_err(to_host(c) - c._c0, w) < TOL["fp32"]
    batch.close()


def test_conj_alpha_and_accumulate():
    rng = np.random.default_rng(8)
    la, lb, lc = "abcd", "dbef", "acef"
    ext = {"a": 9, "b": 6, "c": 17, "d": 11, "e": 8, "f": 19}
    for dtype, kernels in ((np.complex64, KERNELS_C64), (np.complex128, KERNELS_C128)):
        A = rand(rng, [ext[x] for x in la], dtype)
        B = rand(rng, [ext[x] for x in lb], dtype)
        C0 = rand(rng, [ext[x] for x in lc], dtype)
        for kern, prec, tolkey in kernels:
            for ca, cb, alpha in [(True, False, 1.0), (False, True, 0.5 - 2j), (True, True, -1j)]:
                want = oracle_einsum(A, la, B, lb, C0, lc, alpha=alpha, conj_a=ca, conj_b=cb)
                with forced(kern):
                    got = run(A, la, B, lb, C0, lc, alpha=alpha, conj_a=ca, conj_b=cb,
                              precision=prec)
                assert rel_err(got - C0, want - C0) < TOL[tolkey], (kern, ca, cb, alpha)
        # C += semantics: a second call adds the product again
        a, b, c = to_dev(A), to_dev(B), to_dev(C0)
        t1m_amd.contract(a, la, b, lb, c, lc)
        t1m_amd.contract(a, la, b, lb, c, lc)
        twice = oracle_einsum(A, la, B, lb, oracle_einsum(A, la, B, lb, C0, lc), lc)
        tol = TOL["fp32"] if dtype == np.complex64 else TOL["fp64"]
        assert rel_err(to_host(c) - C0, twice - C0) < tol


def test_edge_shapes():
    rng = np.random.default_rng(9)
    cases = [
        ("a", "b", "ab", {"a": 70, "b": 33}),                 # outer product (K = 1)
        ("ab", "ab", "", {"a": 20, "b": 31}),                 # full contraction -> scalar
        ("abc", "c", "ab", {"a": 5, "b": 9, "c": 300}),       # N = 1 (matrix-vector)
        ("c", "cb", "b", {"b": 150, "c": 77}),                # M = 1
        ("ab", "bc", "ac", {"a": 1, "b": 1, "c": 1}),         # 1 x 1 x 1
        ("abc", "cbd", "da", {"a": 17, "b": 3, "c": 2, "d": 145}),
    ]
    for la, lb, lc, ext in cases:
        for dtype in (np.complex64, np.complex128):
            A = rand(rng, [ext[x] for x in la], dtype)
            B = rand(rng, [ext[x] for x in lb], dtype)
            C0 = rand(rng, [ext[x] for x in lc], dtype)
            want = oracle_einsum(A, la, B, lb, C0, lc)
            got = run(A, la, B, lb, C0, lc)
            tol = TOL["fp32"] if dtype == np.complex64 else TOL["fp64"]
            assert rel_err(got - C0, want - C0) < tol, (la, lb, lc, dtype)


def test_row_major_and_negative_strides():
    rng = np.random.default_rng(10)
    la, lb, lc = "abc", "cbd", "ad"
    ext = {"a": 33, "b": 5, "c": 21, "d": 40}
    A = np.ascontiguousarray(rand(rng, [ext[x] for x in la], np.complex64))   # row-major
    B = rand(rng, [ext[x] for x in lb], np.complex64)
    C0 = np.ascontiguousarray(rand(rng, [ext[x] for x in lc], np.complex64))
    want = oracle_einsum(A, la, B, lb, C0, lc)
    a, b = torch.from_numpy(A).cuda(), to_dev(B)
    c = torch.from_numpy(C0).cuda()
    t1m_amd.contract(a, la, b, lb, c, lc)
    assert rel_err(c.cpu().numpy() - C0, want - C0) < TOL["fp32"]
    # A viewed backwards along 'a' through a descriptor with a negative stride
    Af = A[::-1].copy()                       # Af[i] = A[n-1-i]
    af = torch.from_numpy(np.ascontiguousarray(Af)).cuda()
    last = af.data_ptr() + (ext["a"] - 1) * af.stride(0) * af.element_size()
    d = t1m_amd._Desc(af, strides=[-af.stride(0), af.stride(1), af.stride(2)], ptr=last)
    c2 = torch.from_numpy(C0).cuda()
    db, dc = t1m_amd._Desc(b), t1m_amd._Desc(c2)
    o = t1m_amd._options()
    rc = t1m_amd.load().t1m_contract(ctypes.byref(d.desc), la.encode(), ctypes.byref(db.desc),
                                      lb.encode(), ctypes.byref(dc.desc), lc.encode(),
                                      ctypes.byref(o))
    assert rc == 0
    assert rel_err(c2.cpu().numpy() - C0, want - C0) < TOL["fp32"]


def test_host_pointers_are_staged():
    """Host memory, as the reference's callers pass it, through the C ABI."""
    rng = np.random.default_rng(11)
    A = rand(rng, (12, 7, 9), np.complex64)
    B = rand(rng, (7, 11), np.complex64)
    C = rand(rng, (12, 9, 11), np.complex64)
    C0 = C.copy(order="F")
    want = oracle_einsum(A, "abc", B, "bd", C0, "acd")

    def desc(x):
        dims = (ctypes.c_int64 * x.ndim)(*x.shape)
        st = (ctypes.c_int64 * x.ndim)(*[s // x.itemsize for s in x.strides])
        return t1m_amd.TensorDesc(t1m_amd.COMPLEX64, x.ndim, dims, st, x.ctypes.data), dims, st

    da, db, dc = desc(A), desc(B), desc(C)
    o = t1m_amd._options(stream=0)
    rc = t1m_amd.load().t1m_contract(ctypes.byref(da[0]), b"abc", ctypes.byref(db[0]), b"bd",
                                      ctypes.byref(dc[0]), b"acd", ctypes.byref(o))
    assert rc == 0, t1m_amd.load().t1m_last_error()
    assert rel_err(C - C0, want - C0) < TOL["fp32"]


def test_host_staging_buffers_are_reused():
    """Host-pointer calls take their device buffers from a per-device pool:
    a larger call after smaller ones, a smaller one reusing a larger buffer
    (stale bytes beyond its extent must not matter), a pool emptied by a
    plan cache clear, and host operands mixed with device ones."""
    rng = np.random.default_rng(12)
    for step, (a_, b_, c_, d_) in enumerate([(12, 7, 9, 11), (40, 33, 21, 17), (5, 3, 4, 2),
                                             (40, 33, 21, 17), (9, 8, 7, 6)]):
        if step == 3:
            t1m_amd.plan_cache_clear()
        A = rand(rng, (a_, b_, c_), np.complex64)
        B = rand(rng, (b_, d_), np.complex64)
        C = rand(rng, (a_, c_, d_), np.complex64)
        C0 = C.copy()
        want = oracle_einsum(A, "abc", B, "bd", C0, "acd")
        if step == 4:   # device A, host B and C
            a = to_dev(A)
            c = torch.from_numpy(C)
            t1m_amd.contract(a, "abc", torch.from_numpy(B), "bd", c, "acd")
            got = c.numpy()
        else:
            c = torch.from_numpy(C)
            t1m_amd.contract(torch.from_numpy(A), "abc", torch.from_numpy(B), "bd", c, "acd")
            got = c.numpy()
        assert rel_err(got - C0, want - C0) < TOL["fp32"], step


def test_autotune_picks_a_candidate_and_persists():
    rng = np.random.default_rng(21)
    la, lb, lc = "abcd", "dbef", "acef"
    ext = {"a": 24, "b": 10, "c": 24, "d": 12, "e": 24, "f": 24}
    A = rand(rng, [ext[x] for x in la], np.complex64)
    B = rand(rng, [ext[x] for x in lb], np.complex64)
    C0 = rand(rng, [ext[x] for x in lc], np.complex64)
    want = oracle_einsum(A, la, B, lb, C0, lc)
    for prec, cands in (("native", ("c64_f32mfma_144x144", "c64_f32mfma_64x64")),
                        ("bf16x3", ("c64_bf16x3_144x144", "c64_bf16x3_64x64"))):
        a, b, c = to_dev(A), to_dev(B), to_dev(C0)
        t1m_amd.contract(a, la, b, lb, c, lc, precision=prec, autotune=True)
        assert rel_err(to_host(c) - C0, want - C0) < TOL["fp32" if prec == "native" else "bf16x3"]
        assert t1m_amd.plan_kernel(a, la, b, lb, c, lc, precision=prec) in cands
    path = t1m_amd.autotune_cache_path()
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    assert any("M576|N576|K120" in ln[0] for ln in lines), lines


def test_plan_cache_reuse():
    rng = np.random.default_rng(12)
    A, B = rand(rng, (8, 9), np.complex64), rand(rng, (9, 10), np.complex64)
    a, b = to_dev(A), to_dev(B)
    c = torch.zeros(8, 10, dtype=torch.complex64, device="cuda")
    t1m_amd.contract(a, "ab", b, "bc", c, "ac")
    n = t1m_amd.plan_cache_size()
    for _ in range(5):
        t1m_amd.contract(a, "ab", b, "bc", c, "ac")
    assert t1m_amd.plan_cache_size() == n


# ----------------------------------------------------- full-size configs --

def _freivalds(A, la, B, lb, C_delta, lc, rng):
    """Check C_delta = A.B with a random vector over C's B-free labels:
    (A.B).x == A.(B.x), all in complex128 numpy (size independent)."""
    fb = [ch for ch in lb if ch not in la]
    x = rng.standard_normal([B.shape[lb.index(ch)] for ch in fb]) + 0j
    fbl = "".join(fb)
    lhs = np.einsum(f"{lc},{fbl}->{''.join(ch for ch in lc if ch not in fbl)}", C_delta, x,
                    optimize=True)
    bx = np.einsum(f"{lb},{fbl}->{''.join(ch for ch in lb if ch not in fbl)}", B, x,
                   optimize=True)
    rhs = np.einsum(f"{la},{''.join(ch for ch in lb if ch not in fbl)}->"
                    f"{''.join(ch for ch in lc if ch not in fbl)}", A, bx, optimize=True)
    return rel_err(lhs, rhs)


def _torch_ref(a, la, b, lb, lc):
    return torch.einsum(f"{la},{lb}->{lc}", a.to(torch.complex128), b.to(torch.complex128))


def test_config2_48x4_fp32_full_size():
    """c2: complex<float> A[48^4] 'abcd' x B[48^4] 'dbef' -> C[48^4] 'acef'."""
    rng = np.random.default_rng(2)
    sh = (48, 48, 48, 48)
    A, B = rand(rng, sh, np.complex64), rand(rng, sh, np.complex64)
    a, b = to_dev(A), to_dev(B)
    c = to_dev(np.zeros(sh, np.complex64, order="F"))
    t1m_amd.contract(a, "abcd", b, "dbef", c, "acef")
    ref = _torch_ref(a, "abcd", b, "dbef", "acef")
    err = torch.linalg.vector_norm(c.to(torch.complex128) - ref) / torch.linalg.vector_norm(ref)
    print(f"  c2 fp32 rel_err vs complex128 = {err.item():.3e}")
    assert err.item() < TOL["fp32"]
    C = to_host(c)
    f = _freivalds(A.astype(np.complex128), "abcd", B.astype(np.complex128), "dbef", C, "acef", rng)
    assert f < TOL["fp32"]


def test_config3_tn_step_fp64_full_size():
    """c3: complex<double> A[256,4,256] 'asb' x B[256,4,256] 'btc' -> C 'astc'."""
    rng = np.random.default_rng(3)
    A = rand(rng, (256, 4, 256), np.complex128, "normal") / 16.0
    B = rand(rng, (256, 4, 256), np.complex128, "normal") / 16.0
    a, b = to_dev(np.asfortranarray(A)), to_dev(np.asfortranarray(B))
    c = to_dev(np.zeros((256, 4, 4, 256), np.complex128, order="F"))
    t1m_amd.contract(a, "asb", b, "btc", c, "astc")
    ref = _torch_ref(a, "asb", b, "btc", "astc")
    err = (torch.linalg.vector_norm(c - ref) / torch.linalg.vector_norm(ref)).item()
    print(f"  c3 fp64 rel_err vs complex128 = {err:.3e}")
    assert err < TOL["fp64"]


def test_config4_precision_bf16x3_vs_fp32():
    """c4: c2 shapes, |z| log-uniform in [1e-3, 1e3], random phase."""
    rng = np.random.default_rng(4)
    sh = (48, 48, 48, 48)
    A, B = rand(rng, sh, np.complex64, "logmag"), rand(rng, sh, np.complex64, "logmag")
    a, b = to_dev(A), to_dev(B)
    ref = _torch_ref(a, "abcd", b, "dbef", "acef")
    errs = {}
    for prec in ("native", "bf16x3"):
        c = to_dev(np.zeros(sh, np.complex64, order="F"))
        t1m_amd.contract(a, "abcd", b, "dbef", c, "acef", precision=prec)
        errs[prec] = (torch.linalg.vector_norm(c.to(torch.complex128) - ref) /
                      torch.linalg.vector_norm(ref)).item()
    print(f"  c4 rel_err fp32={errs['native']:.3e} bf16x3={errs['bf16x3']:.3e}")
    assert errs["native"] < TOL["fp32"]
    assert errs["bf16x3"] < TOL["bf16x3"]


# ------------------------------------------------------------------ VJP --

def test_vjp_against_oracle_and_finite_differences():
    """VJP of L = sum |C|^2 (grad_c = 2C): grad_a = grad_c . conj(B),
    grad_b = conj(A) . grad_c, checked against the oracle's conjugated label
    loops and complex128 central differences (h = 1e-6)."""
    rng = np.random.default_rng(13)
    la, lb, lc = "asb", "btc", "astc"
    ext = {"a": 6, "s": 4, "b": 5, "t": 4, "c": 7}
    A = rand(rng, [ext[x] for x in la], np.complex128)
    B = rand(rng, [ext[x] for x in lb], np.complex128)
    Cf = oracle_einsum(A, la, B, lb, np.zeros([ext[x] for x in lc], np.complex128, order="F"), lc)
    G = 2 * Cf
    ga_want = oracle_einsum(G, lc, B, lb, np.zeros_like(A, order="F"), la, conj_b=True)
    gb_want = oracle_einsum(A, la, G, lc, np.zeros_like(B, order="F"), lb, conj_a=True)
    a, b, g = to_dev(A), to_dev(B), to_dev(np.asfortranarray(G))
    ga = to_dev(np.zeros_like(A, order="F"))
    gb = to_dev(np.zeros_like(B, order="F"))
    t1m_amd.contract_vjp(a, la, b, lb, g, lc, ga, gb)
    assert rel_err(to_host(ga), ga_want) < 1e-13
    assert rel_err(to_host(gb), gb_want) < 1e-13

    def loss(Ax, Bx):
        return float(np.sum(np.abs(np.einsum(f"{la},{lb}->{lc}", Ax, Bx)) ** 2))

    h = 1e-6
    GA = to_host(ga)
    for _ in range(6):
        idx = tuple(int(rng.integers(0, s)) for s in A.shape)
        for unit, part in ((1.0, "re"), (1j, "im")):
            Ap, Am = A.copy(), A.copy()
            Ap[idx] += unit * h
            Am[idx] -= unit * h
            fd = (loss(Ap, B) - loss(Am, B)) / (2 * h)
            an = GA[idx].real if part == "re" else GA[idx].imag
            assert abs(fd - an) <= 1e-5 * max(1.0, abs(fd)), (idx, part, fd, an)


def test_autograd_harness_complex64():
    rng = np.random.default_rng(14)
    A = to_dev(rand(rng, (9, 4, 7), np.complex64)).requires_grad_(True)
    B = to_dev(rand(rng, (7, 4, 8), np.complex64)).requires_grad_(True)
    C = t1m_amd.contract_autograd(A, "asb", B, "btc", "astc")
    (C.abs() ** 2).sum().backward()
    A2 = A.detach().to(torch.complex128).requires_grad_(True)
    B2 = B.detach().to(torch.complex128).requires_grad_(True)
    (torch.einsum("asb,btc->astc", A2, B2).abs() ** 2).sum().backward()
    assert rel_err(A.grad.cpu().numpy(), A2.grad.cpu().numpy()) < 1e-5
    assert rel_err(B.grad.cpu().numpy(), B2.grad.cpu().numpy()) < 1e-5


# ---------------------------------------------------------------- batch --

@pytest.mark.parametrize("dtype,prec,tol", [(np.complex64, "native", TOL["fp32"]),
                                            (np.complex64, "bf16x3", TOL["bf16x3"]),
                                            (np.complex128, "native", TOL["fp64"]),
                                            (np.float32, "native", TOL["fp32"]),
                                            (np.float64, "native", TOL["fp64"])])
def test_batch_families(dtype, prec, tol):
    """Persistent batched launch (complex) / per-instance graph (real), with
    ragged bond dimensions, conj and alpha, against the oracle."""
    rng = np.random.default_rng(16)
    la, lb, lc = "asb", "btc", "astc"
    As, Bs, Cs, C0s, wants = [], [], [], [], []
    for _ in range(9):
        dl, db, dr = (int(x) for x in rng.integers(3, 90, 3))
        A = rand(rng, (dl, 4, db), dtype)
        B = rand(rng, (db, 4, dr), dtype)
        C0 = rand(rng, (dl, 4, 4, dr), dtype)
        alpha = 0.5 - 1j if np.dtype(dtype).kind == "c" else 0.5
        wants.append(oracle_einsum(A, la, B, lb, C0, lc, alpha=alpha, conj_b=True) - C0)
        As.append(to_dev(A)); Bs.append(to_dev(B)); Cs.append(to_dev(C0)); C0s.append(C0)
    batch = t1m_amd.Batch(As, la, Bs, lb, Cs, lc, alpha=alpha, conj_b=True, precision=prec)
    batch.launch()
    torch.cuda.synchronize()
    for c, c0, w in zip(Cs, C0s, wants):
        assert rel_err(to_host(c) - c0, w) < tol
    batch.close()


def test_batched_vjp_mps_sweep():
    """The MPS sweep's gradient as the benchmark computes it: forward batch,
    then grad_A = grad_C . conj(B) and grad_B = conj(A) . grad_C as two more
    persistent batches (complex<float>, L = sum |C|^2, grad_C = 2 C), checked
    against the oracle's conjugated label loops and complex128 central
    differences (h = 1e-6) of the loss."""
    rng = np.random.default_rng(23)
    la, lb, lc = "asb", "btc", "astc"
    dims = [tuple(int(x) for x in rng.choice(np.arange(16, 81, 16), 3)) for _ in range(6)]
    As = [rand(rng, (dl, 4, db), np.complex64) for dl, db, dr in dims]
    Bs = [rand(rng, (db, 4, dr), np.complex64) for dl, db, dr in dims]
    Cs = [to_dev(np.zeros((dl, 4, 4, dr), np.complex64, order="F")) for dl, db, dr in dims]
    a_dev, b_dev = [to_dev(A) for A in As], [to_dev(B) for B in Bs]
    fwd = t1m_amd.Batch(a_dev, la, b_dev, lb, Cs, lc)
    fwd.launch()
    torch.cuda.synchronize()
    Gs = [2 * c for c in Cs]
    GA = [torch.zeros_like(a) for a in a_dev]
    GB = [torch.zeros_like(b) for b in b_dev]
    vja = t1m_amd.Batch(Gs, lc, b_dev, lb, GA, la, conj_b=True)
    vjb = t1m_amd.Batch(a_dev, la, Gs, lc, GB, lb, conj_a=True)
    vja.launch()
    vjb.launch()
    torch.cuda.synchronize()
    for bt in (fwd, vja, vjb):
        bt.close()
    for i, (A, B) in enumerate(zip(As, Bs)):
        A2, B2 = A.astype(np.complex128), B.astype(np.complex128)
        C = oracle_einsum(A2, la, B2, lb, np.zeros(Cs[i].shape, np.complex128, order="F"), lc)
        ga_want = oracle_einsum(2 * C, lc, B2, lb, np.zeros_like(A2, order="F"), la, conj_b=True)
        gb_want = oracle_einsum(A2, la, 2 * C, lc, np.zeros_like(B2, order="F"), lb, conj_a=True)
        assert rel_err(to_host(GA[i]), ga_want) < 1e-5, i
        assert rel_err(to_host(GB[i]), gb_want) < 1e-5, i

        def loss(Ax, Bx):
            return float(np.sum(np.abs(np.einsum(f"{la},{lb}->{lc}", Ax, Bx)) ** 2))

        h = 1e-6
        ga, gb = to_host(GA[i]), to_host(GB[i])
        for _ in range(2):
            ia = tuple(int(rng.integers(0, s)) for s in A2.shape)
            ib = tuple(int(rng.integers(0, s)) for s in B2.shape)
            for unit, part in ((1.0, "re"), (1j, "im")):
                Ap, Am = A2.copy(), A2.copy()
                Ap[ia] += unit * h
                Am[ia] -= unit * h
                fd = (loss(Ap, B2) - loss(Am, B2)) / (2 * h)
                an = ga[ia].real if part == "re" else ga[ia].imag
                assert abs(fd - an) <= 1e-4 * max(1.0, abs(fd)), (i, ia, part, fd, an)
                Bp, Bm = B2.copy(), B2.copy()
                Bp[ib] += unit * h
                Bm[ib] -= unit * h
                fd = (loss(A2, Bp) - loss(A2, Bm)) / (2 * h)
                an = gb[ib].real if part == "re" else gb[ib].imag
                assert abs(fd - an) <= 1e-4 * max(1.0, abs(fd)), (i, ib, part, fd, an)


def test_batch_matches_single_calls():
    rng = np.random.default_rng(15)
    la, lb, lc = "asb", "btc", "astc"
    As, Bs, Cs, wants = [], [], [], []
    for _ in range(12):
        dl, db, dr = (int(x) for x in rng.choice([16, 32, 48, 64, 96, 128], 3))
        A = rand(rng, (dl, 4, db), np.complex64)
        B = rand(rng, (db, 4, dr), np.complex64)
        C0 = rand(rng, (dl, 4, 4, dr), np.complex64)
        wants.append(oracle_einsum(A, la, B, lb, C0, lc) - C0)
        As.append(to_dev(A)); Bs.append(to_dev(B)); Cs.append(to_dev(C0))
        Cs[-1]._c0 = C0
    batch = t1m_amd.Batch(As, la, Bs, lb, Cs, lc)
    for rep in range(3):
        batch.launch()
        torch.cuda.synchronize()
        for c, w in zip(Cs, wants):
            assert rel_err(to_host(c) - c._c0, (rep + 1) * w) < TOL["fp32"]
    batch.close()


@pytest.mark.parametrize("pairs", ["1", "0"])
def test_batch_every_stage_count(pairs):
    """The persistent complex<float> batch over bond dimensions that give one,
    an odd and an even number of 8-k stages, whole or with a ragged last
    stage, and ragged tile edges: the forward, grad_A and grad_B launches of
    an MPS step against the oracle.  With even extents the tiles gather
    two-element pairs and keep two stages of loads in flight (whose loop
    peels odd and even stage counts differently); T1M_PAIRS=0 forces the
    one-element gathers, one stage ahead."""
    rng = np.random.default_rng(29)
    la, lb, lc = "asb", "btc", "astc"
    bonds = [2, 8, 10, 24, 26, 40, 56, 64]     # forward: 1, 1, 2, 3, 4, 5, 7, 8 stages
    dims = [(int(rng.choice(np.arange(2, 131, 2))), db, int(rng.choice(np.arange(2, 131, 2))))
            for db in bonds]
    As = [rand(rng, (dl, 4, db), np.complex64) for dl, db, dr in dims]
    Bs = [rand(rng, (db, 4, dr), np.complex64) for dl, db, dr in dims]
    C0s = [rand(rng, (dl, 4, 4, dr), np.complex64) for dl, db, dr in dims]
    Gs = [rand(rng, (dl, 4, 4, dr), np.complex64) for dl, db, dr in dims]
    a_dev, b_dev = [to_dev(A) for A in As], [to_dev(B) for B in Bs]
    c_dev, g_dev = [to_dev(C) for C in C0s], [to_dev(G) for G in Gs]
    ga_dev = [torch.zeros_like(a) for a in a_dev]
    gb_dev = [torch.zeros_like(b) for b in b_dev]
    old = os.environ.get("T1M_PAIRS")
    os.environ["T1M_PAIRS"] = pairs
    try:
        batches = [t1m_amd.Batch(a_dev, la, b_dev, lb, c_dev, lc),
                   t1m_amd.Batch(g_dev, lc, b_dev, lb, ga_dev, la, conj_b=True),
                   t1m_amd.Batch(a_dev, la, g_dev, lc, gb_dev, lb, conj_a=True)]
    finally:
        if old is None:
            os.environ.pop("T1M_PAIRS", None)
        else:
            os.environ["T1M_PAIRS"] = old
    for bt in batches:
        bt.launch()
    torch.cuda.synchronize()
    for bt in batches:
        bt.close()
    for i, (A, B, C0, G) in enumerate(zip(As, Bs, C0s, Gs)):
        want_c = oracle_einsum(A, la, B, lb, C0, lc) - C0
        want_ga = oracle_einsum(G, lc, B, lb, np.zeros_like(A, order="F"), la, conj_b=True)
        want_gb = oracle_einsum(A, la, G, lc, np.zeros_like(B, order="F"), lb, conj_a=True)
        print(f"  instance {i} dims {dims[i]}", flush=True)
        assert rel_err(to_host(c_dev[i]) - C0, want_c) < TOL["fp32"], ("forward", i, dims[i])
        assert rel_err(to_host(ga_dev[i]), want_ga) < TOL["fp32"], ("grad_A", i, dims[i])
        assert rel_err(to_host(gb_dev[i]), want_gb) < TOL["fp32"], ("grad_B", i, dims[i])


BLOCKED_LAYOUTS = [
    # k: A and B gathered along k, unit strides on different contracted modes
    (("bdac", "dbef", "acef"), dict(a=5, b=8, c=7, d=12, e=6, f=5)),
    (("bdac", "dbef", "acef"), dict(a=9, b=4, c=3, d=20, e=3, f=7)),
    (("bdac", "dbef", "acef"), dict(a=40, b=16, c=2, d=24, e=33, f=2)),
    # m: A gathered along m with unit stride on e, C's on a
    (("efbad", "cf", "abcde"), dict(a=8, b=4, c=3, d=4, e=12, f=4)),
    # n: B gathered along n with unit stride on d, C's on c
    (("ka", "dck", "cad"), dict(a=8, k=12, c=8, d=16)),
]


@pytest.mark.parametrize("case", range(len(BLOCKED_LAYOUTS)))
def test_blocked_bundle_orders(case):
    """Layouts where two tensors of one bundle have their unit stride on
    different modes, so the planner splits both into 4-blocks and
    interleaves them in the traversal (k: A and B; m: C and A; n: C and
    B); every kernel against the oracle."""
    (la, lb, lc), ext = BLOCKED_LAYOUTS[case]
    rng = np.random.default_rng(23 + case)
    for dtype in (np.complex64, np.complex128, np.float32, np.float64):
        A = rand(rng, [ext[x] for x in la], dtype)
        B = rand(rng, [ext[x] for x in lb], dtype)
        C0 = rand(rng, [ext[x] for x in lc], dtype)
        want = oracle_einsum(A, la, B, lb, C0, lc)
        for kern, prec, tolkey in ALL_KERNELS[dtype]:
            with forced(kern):
                got = run(A, la, B, lb, C0, lc, precision=prec)
            err = rel_err(got - C0, want - C0)
            print(f"  {la},{lb}->{lc} {ext} {np.dtype(dtype).name} {kern} rel={err:.2e}")
            assert err < TOL[tolkey], (la, lb, lc, ext, dtype, kern, err)


def test_aliased_outputs_add_every_contribution():
    """C views whose index tuples share addresses (a zero stride, interleaved
    strides) get every contribution, as a serial loop over the outputs would
    give; expected values from numpy's unbuffered np.add.at."""
    rng = np.random.default_rng(21)
    M, K, N = 150, 37, 70
    for dtype in (np.complex64, np.complex128, np.float32, np.float64):
        A = rand(rng, (M, K), dtype)
        B = rand(rng, (K, N), dtype)
        P = A.astype(np.complex128 if np.dtype(dtype).kind == "c" else np.float64) @ B
        # (stride_m, stride_n, buffer length): zero stride on m, then m and n interleaved
        for sm, sn in ((0, 1), (1, 2), (3, 1)):
            length = (M - 1) * sm + (N - 1) * sn + 1
            idx = (np.arange(M)[:, None] * sm + np.arange(N)[None, :] * sn).ravel()
            buf0 = rand(rng, (length,), dtype)
            want = buf0.astype(P.dtype)
            np.add.at(want, idx, (0.5 * P).ravel())
            for kern, prec, tolkey in ALL_KERNELS[dtype]:
                buf = to_dev(buf0)
                c = buf.as_strided((M, N), (sm, sn))
                with forced(kern):
                    t1m_amd.contract(to_dev(A), "ak", to_dev(B), "kn", c, "an",
                                     precision=prec, alpha=0.5)
                err = rel_err(to_host(buf) - buf0, want - buf0)
                print(f"  strides ({sm},{sn}) {np.dtype(dtype).name} {kern} rel={err:.2e}")
                assert err < TOL[tolkey], (sm, sn, dtype, kern, err)


def test_batch_instances_sharing_one_output():
    """Batch instances that write into one C add both contributions."""
    rng = np.random.default_rng(22)
    la, lb, lc = "asb", "btc", "astc"
    A1, A2 = (rand(rng, (48, 4, 32), np.complex64) for _ in range(2))
    B1, B2 = (rand(rng, (32, 4, 40), np.complex64) for _ in range(2))
    C0 = rand(rng, (48, 4, 4, 40), np.complex64)
    want = oracle_einsum(A2, la, B2, lb, oracle_einsum(A1, la, B1, lb, C0, lc), lc) - C0
    c = to_dev(C0)
    batch = t1m_amd.Batch([to_dev(A1), to_dev(A2)], la, [to_dev(B1), to_dev(B2)], lb, [c, c], lc)
    batch.launch()
    torch.cuda.synchronize()
    assert rel_err(to_host(c) - C0, want) < TOL["fp32"]
    batch.close()


# ------------------------------------------------------- C++ drop-in ----

def test_cpp_dropin_program():
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "dropin_test")
    assert os.path.exists(exe), "build tests/cpp first (__graft_entry__.build())"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]


# ------------------------------------------------------------ fuzzing ----

ALL_KERNELS = {np.complex64: KERNELS_C64, np.complex128: KERNELS_C128,
               np.float32: REAL_BATCH[np.float32], np.float64: REAL_BATCH[np.float64]}


def _scaled_err(got, want, A, la, B, lb, C0, lc, ext, alpha=1.0):
    """Error of C0 + alpha A.B against the oracle, relative to the scale of
    what was summed: |alpha| |A|.|B| (the contraction of the magnitudes)
    plus |C0|.  Small sums with cancellation (dot products, contractions to
    a scalar) are so held to the precision of their terms, not of their
    possibly tiny result."""
    absa = np.asarray(np.abs(A), dtype=np.float64)     # stays an array for 0-d tensors
    absb = np.asarray(np.abs(B), dtype=np.float64)
    mag = oracle_einsum(absa, la, absb, lb, np.zeros([ext[x] for x in lc], np.float64), lc)
    d = np.asarray(got, np.complex128).ravel() - np.asarray(want, np.complex128).ravel()
    scale = abs(alpha) * np.linalg.norm(mag.ravel()) + np.linalg.norm(np.abs(C0).ravel())
    return np.linalg.norm(d) / max(scale, 1e-300)


def _random_problem(rng):
    """Random labels (contracted, free A, free B, sometimes a batch label)
    with extents 1..9."""
    letters = list("abcdefghijklmnop")
    rng.shuffle(letters)
    nk, nm, nn, nb = (int(rng.integers(0, 3)), int(rng.integers(0, 3)),
                      int(rng.integers(0, 3)), int(rng.integers(0, 2)) if rng.random() < 0.3 else 0)
    k, m, n, z = (letters[:nk], letters[nk:nk + nm], letters[nk + nm:nk + nm + nn],
                  letters[nk + nm + nn:nk + nm + nn + nb])
    la = "".join(rng.permutation(k + m + z)) if k + m + z else ""
    lb = "".join(rng.permutation(k + n + z)) if k + n + z else ""
    lc = "".join(rng.permutation(m + n + z)) if m + n + z else ""
    ext = {ch: int(rng.integers(1, 10)) for ch in la + lb + lc}
    return la, lb, lc, ext


def _random_layout(rng, X):
    """Device tensor holding X's values in a random memory layout: permuted
    storage order of the modes and padding on one of them."""
    if X.ndim == 0:
        return to_dev(X)
    perm = [int(p) for p in rng.permutation(X.ndim)]
    shape_p = [X.shape[p] for p in perm]
    pad = [0] * X.ndim
    pad[int(rng.integers(0, X.ndim))] = int(rng.integers(0, 3))
    big = torch.zeros([s_ + q for s_, q in zip(shape_p, pad)],
                      dtype=torch.from_numpy(np.zeros(1, X.dtype)).dtype, device="cuda")
    view = big[tuple(slice(0, s_) for s_ in shape_p)]
    view.copy_(torch.from_numpy(np.ascontiguousarray(X.transpose(perm))))
    return view.permute(*[int(q) for q in np.argsort(perm)])


@pytest.mark.parametrize("seed", range(12))
def test_fuzz_random_contractions(seed):
    """Random label sets (incl. batch labels, empty bundles, extents of 1),
    random strides and every kernel of each dtype, against the oracle."""
    rng = np.random.default_rng(1000 + seed)
    for case in range(6):
        la, lb, lc, ext = _random_problem(rng)
        dtype = [np.complex64, np.complex128, np.float32, np.float64][int(rng.integers(0, 4))]
        A = rand(rng, [ext[x] for x in la], dtype)
        B = rand(rng, [ext[x] for x in lb], dtype)
        C0 = rand(rng, [ext[x] for x in lc], dtype)
        cplx = np.dtype(dtype).kind == "c"
        alpha = complex(rng.uniform(-1, 1), rng.uniform(-1, 1)) if cplx else float(rng.uniform(-1, 1))
        conj_a = cplx and rng.random() < 0.3
        conj_b = cplx and rng.random() < 0.3
        want = oracle_einsum(A, la, B, lb, C0, lc, alpha=alpha, conj_a=conj_a, conj_b=conj_b)
        for kern, prec, tolkey in ALL_KERNELS[dtype]:
            a, b, c = _random_layout(rng, A), _random_layout(rng, B), _random_layout(rng, C0)
            with forced(kern):
                t1m_amd.contract(a, la, b, lb, c, lc, precision=prec, alpha=alpha,
                                 conj_a=conj_a, conj_b=conj_b)
            got = c.cpu().numpy()
            err = _scaled_err(got, want, A, la, B, lb, C0, lc, ext, alpha)
            print(f"  {la},{lb}->{lc} {ext} {np.dtype(dtype).name} {kern} err={err:.2e}")
            assert err < TOL[tolkey], (la, lb, lc, ext, kern, err)


@pytest.mark.parametrize("seed", range(8))
def test_fuzz_blocked_orders(seed):
    """Random label sets with extents in {4, 8, 12, 16} and randomly
    permuted storage orders, so that the planner's 4-block interleaving of
    k, m and n traversals (unit strides of two tensors on different modes)
    is exercised often; every kernel of each dtype against the oracle."""
    rng = np.random.default_rng(2000 + seed)
    for case in range(4):
        la, lb, lc, _ = _random_problem(rng)
        ext = {ch: int(rng.choice([4, 8, 12, 16])) for ch in la + lb + lc}
        dtype = [np.complex64, np.complex128, np.float32, np.float64][int(rng.integers(0, 4))]
        A = rand(rng, [ext[x] for x in la], dtype)
        B = rand(rng, [ext[x] for x in lb], dtype)
        C0 = rand(rng, [ext[x] for x in lc], dtype)
        want = oracle_einsum(A, la, B, lb, C0, lc)
        for kern, prec, tolkey in ALL_KERNELS[dtype]:
            a, b, c = _random_layout(rng, A), _random_layout(rng, B), _random_layout(rng, C0)
            with forced(kern):
                t1m_amd.contract(a, la, b, lb, c, lc, precision=prec)
            err = _scaled_err(c.cpu().numpy(), want, A, la, B, lb, C0, lc, ext)
            print(f"  {la},{lb}->{lc} {ext} {np.dtype(dtype).name} {kern} err={err:.2e}")
            assert err < TOL[tolkey], (la, lb, lc, ext, kern, err)


def test_static_tile_rule():
    """Without tuning: thin C (at most 16 wide) takes the 128 x 16 tiles, a
    48^4 contraction the 144 x 144 tiles, a 324-tile GEMM (1.27 rounds of
    144-tiles) and a 36-row C the 64 x 64 tiles; FP64 thin C its thin
    tiles."""
    def kern(la, lb, lc, ext, dtype=torch.complex64, precision="native"):
        mk = lambda lab: torch.zeros([ext[x] for x in lab], dtype=dtype, device="cuda")
        return t1m_amd.plan_kernel(mk(la), la, mk(lb), lb, mk(lc), lc, precision=precision)
    with forced(None):
        assert kern("efbad", "cf", "abcde", dict(a=24, b=16, c=12, d=16, e=24, f=16)) == "c64_f32mfma_128x16"
        assert kern("efbad", "cf", "abcde", dict(a=24, b=16, c=12, d=16, e=24, f=16),
                    torch.complex128) == "c128_f64mfma_128x16"
        assert kern("abcd", "dbef", "acef", {ch: 48 for ch in "abcdef"}) == "c64_f32mfma_144x144"
        assert kern("abcd", "dbef", "acef", {ch: 48 for ch in "abcdef"},
                    precision="bf16x3") == "c64_bf16x3_144x144"
        assert kern("ac", "cb", "ab", dict(a=2568, b=2560, c=64)) == "c64_f32mfma_64x64"
        assert kern("ea", "ebcd", "abcd", dict(a=36, b=36, c=36, d=36, e=36)) == "c64_f32mfma_64x64"
    # an empty override reads as unset
    old = os.environ.get("T1M_FORCE_KERNEL")
    os.environ["T1M_FORCE_KERNEL"] = ""
    try:
        assert kern("abcd", "dbef", "acef", {ch: 48 for ch in "abcdef"}) == "c64_f32mfma_144x144"
    finally:
        if old is None:
            del os.environ["T1M_FORCE_KERNEL"]
        else:
            os.environ["T1M_FORCE_KERNEL"] = old


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_thin_m_runs_transposed(dtype):
    """A with at most 16 free rows (and a wider C) runs as C^T = B^T A^T on
    every kernel, with conjugation, complex alpha, a batch label and host
    pointers; the static rule picks the thin tiles for it."""
    rng = np.random.default_rng(31)
    kernels = ALL_KERNELS[dtype][:-1] if dtype == np.complex64 else ALL_KERNELS[dtype]
    for la, lb, lc, ext in (("ea", "ebcd", "abcd", dict(a=12, b=9, c=10, d=7, e=20)),
                            ("zea", "zeb", "zab", dict(z=3, a=5, b=70, e=33)),
                            ("ae", "ebc", "cba", dict(a=16, b=40, c=3, e=8))):
        A = rand(rng, [ext[x] for x in la], dtype)
        B = rand(rng, [ext[x] for x in lb], dtype)
        C0 = rand(rng, [ext[x] for x in lc], dtype)
        alpha = complex(0.5, -0.75)
        want = oracle_einsum(A, la, B, lb, C0, lc, alpha=alpha, conj_a=True)
        for kern, prec, tolkey in kernels:
            with forced(kern):
                got = run(A, la, B, lb, C0, lc, precision=prec, alpha=alpha, conj_a=True)
            err = rel_err(got - C0, want - C0)
            print(f"  {la},{lb}->{lc} {ext} {kern} rel={err:.2e}")
            assert err < TOL[tolkey], (la, lb, lc, kern, err)
    thin = "c64_f32mfma_128x16" if dtype == np.complex64 else "c128_f64mfma_128x16"
    tdt = torch.complex64 if dtype == np.complex64 else torch.complex128
    mk = lambda lab, ext: torch.zeros([ext[x] for x in lab], dtype=tdt, device="cuda")
    ext = dict(a=12, b=36, c=36, d=36, e=36)
    with forced(None):
        assert t1m_amd.plan_kernel(mk("ea", ext), "ea", mk("ebcd", ext), "ebcd",
                                   mk("abcd", ext), "abcd") == thin


def test_concurrent_calls_from_threads():
    """Reentrancy: four host threads, each on a stream of its own, issue
    contractions of shared and distinct shapes (new plans, cached plans,
    the autotuner's first call) while a fifth thread keeps clearing the plan
    cache; every result matches a single-threaded call."""
    import threading
    import time
    rng = np.random.default_rng(41)
    problems = []
    for la, lb, lc, ext in [("abcd", "dbef", "acef", dict(a=7, b=5, c=9, d=6, e=8, f=4)),
                            ("asb", "btc", "astc", dict(a=33, s=4, b=21, t=4, c=17)),
                            ("ab", "bc", "ac", dict(a=130, b=70, c=150)),
                            ("efbad", "cf", "abcde", dict(a=6, b=5, c=11, d=4, e=7, f=9))]:
        for dtype in (np.complex64, np.complex128):
            A = rand(rng, [ext[x] for x in la], dtype)
            B = rand(rng, [ext[x] for x in lb], dtype)
            C0 = rand(rng, [ext[x] for x in lc], dtype)
            problems.append((la, lb, lc, to_dev(A), to_dev(B), C0))
    want = []
    for la, lb, lc, a, b, C0 in problems:
        c = to_dev(C0)
        t1m_amd.contract(a, la, b, lb, c, lc)
        want.append(to_host(c))
    torch.cuda.synchronize()
    t1m_amd.plan_cache_clear()
    errors = []
    stop = threading.Event()

    def worker(w):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for it in range(24):
                    i = (3 * w + it) % len(problems)
                    la, lb, lc, a, b, C0 = problems[i]
                    c = to_dev(C0)
                    t1m_amd.contract(a, la, b, lb, c, lc, autotune=(it % 5 == 0))
                    err = rel_err(to_host(c) - C0, want[i] - C0)
                    tol = TOL["fp32"] if c.dtype == torch.complex64 else TOL["fp64"]
                    if not err < tol:
                        errors.append((w, it, i, err))
        except Exception as e:  # reported below, on the main thread
            errors.append((w, repr(e)))

    def clearer():
        while not stop.is_set():
            t1m_amd.plan_cache_clear()
            time.sleep(0.002)

    threads = [threading.Thread(target=worker, args=(w,)) for w in range(4)]
    cl = threading.Thread(target=clearer)
    cl.start()
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=90)
    stop.set()
    cl.join(timeout=10)
    torch.cuda.synchronize()
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors[:5]


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_contract_captured_in_a_caller_graph(dtype):
    """A caller captures t1m_contract (cached plan, asynchronous) into its
    own HIP graph: every replay adds A.B to C again, with eager work and
    synchronizations between the replays; a K small enough for split K
    (atomic adds into C) is included."""
    rng = np.random.default_rng(52)
    cases = [("abcd", "dbef", "acef", dict(a=12, b=8, c=12, d=8, e=16, f=8)),
             ("asb", "btc", "astc", dict(a=40, s=4, b=2048, t=4, c=24))]
    for la, lb, lc, ext in cases:
        A = rand(rng, [ext[x] for x in la], dtype)
        B = rand(rng, [ext[x] for x in lb], dtype)
        C0 = rand(rng, [ext[x] for x in lc], dtype)
        prod = oracle_einsum(A, la, B, lb, np.zeros_like(C0, order="F"), lc)
        a, b, c = to_dev(A), to_dev(B), to_dev(C0)
        t1m_amd.contract(a, la, b, lb, c, lc)          # builds the plan eagerly
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            t1m_amd.contract(a, la, b, lb, c, lc, synchronize=False)
        tol = TOL["fp32"] if dtype == np.complex64 else TOL["fp64"]
        base = to_host(c)                               # after the eager call only
        for r in range(3):
            g.replay()
            torch.cuda.synchronize()
            got = to_host(c)
            assert rel_err(got - base, prod) < tol, (la, r)
            base = got
            scratch = to_dev(C0)                        # eager work between replays
            t1m_amd.contract(a, la, b, lb, scratch, lc)
            assert rel_err(to_host(scratch) - C0, prod) < tol, (la, r)
        del g


def test_offsets_beyond_int32_take_64bit_maps():
    """A strided view whose element offsets pass 2^31 (a 25 GB buffer): the
    plan switches to 64-bit maps by itself (the int32 maps would wrap);
    checked against a complex128 product of a dense copy, both directions
    of the operand roles."""
    step = 1 << 26                                   # 47 * 2^26 > 2^31 - 1
    buf = torch.zeros(47 * step + 64, dtype=torch.complex64, device="cuda")
    try:
        big = torch.as_strided(buf, (64, 48), (1, step))
        big.copy_(torch.randn(64, 48, dtype=torch.complex64, device="cuda"))
        B = torch.randn(48, 40, dtype=torch.complex64, device="cuda")
        C = torch.zeros(64, 40, dtype=torch.complex64, device="cuda")
        t1m_amd.contract(big, "ik", B, "kj", C, "ij")
        ref = big.contiguous().to(torch.complex128) @ B.to(torch.complex128)
        assert rel_err(C.cpu().numpy(), ref.cpu().numpy()) < TOL["fp32"]
        A = torch.randn(40, 64, dtype=torch.complex64, device="cuda")
        C2 = torch.zeros(40, 48, dtype=torch.complex64, device="cuda")
        t1m_amd.contract(A, "ik", big, "kj", C2, "ij")
        ref2 = A.to(torch.complex128) @ big.contiguous().to(torch.complex128)
        assert rel_err(C2.cpu().numpy(), ref2.cpu().numpy()) < TOL["fp32"]
    finally:
        del buf
        torch.cuda.empty_cache()


def _reversed_desc(t, dim):
    """A descriptor of torch tensor t viewed backwards along `dim` (negative
    stride, data pointer at the last element of that dimension)."""
    st = list(t.stride())
    last = t.data_ptr() + (t.shape[dim] - 1) * st[dim] * t.element_size()
    st[dim] = -st[dim]
    return t1m_amd._Desc(t, strides=st, ptr=last)


def test_bf16x3_144_staging_paths():
    """Both staging paths of the 144 x 144 BF16x3 tiles.  Plans with 32-bit
    offsets whose operands span less than 2^32 bytes run the kernel that
    copies the contracted-index maps into LDS and addresses A, B and C by
    32-bit byte offsets from bases rebased by the maps' minima: checked with
    negative strides on all three operands, interior and edge tiles, K not a
    multiple of 16, and K past one LDS map (8192 indices: sequential
    launches).  A 32-bit plan whose operand spans more than 2^32 bytes takes
    the global-map kernel (both operand roles).  Against complex128 products."""
    tol = TOL["bf16x3"]
    g = torch.Generator(device="cuda").manual_seed(77)

    def crandn(*shape):
        return torch.randn(*shape, dtype=torch.complex64, device="cuda", generator=g)

    lib = t1m_amd.load()
    for M, N, K in ((300, 290, 37), (160, 150, 8492)):
        a = crandn(K, M).t()            # col-major M x K ("ik")
        b = crandn(N, K).t()            # col-major K x N ("kj")
        c = crandn(N, M).t()            # col-major M x N ("ij")
        c0 = c.cpu().numpy().astype(np.complex128)
        # A backwards along i, B backwards along k, C backwards along j
        da, db, dc = _reversed_desc(a, 0), _reversed_desc(b, 0), _reversed_desc(c, 1)
        o = t1m_amd._options(precision="bf16x3")
        with forced("c64_bf16x3_144x144"):
            rc = lib.t1m_contract(ctypes.byref(da.desc), b"ik", ctypes.byref(db.desc), b"kj",
                                  ctypes.byref(dc.desc), b"ij", ctypes.byref(o))
        assert rc == 0, lib.t1m_last_error()
        An = a.cpu().numpy().astype(np.complex128)[::-1, :]
        Bn = b.cpu().numpy().astype(np.complex128)[::-1, :]
        want = c0.copy()
        want[:, ::-1] += An @ Bn
        err = rel_err(c.cpu().numpy() - c0, want - c0)
        print(f"  M={M} N={N} K={K} reversed operands rel={err:.2e}")
        assert err < tol, (M, N, K, err)

    # offsets below 2^31 elements, byte span above 2^32: 36 * 2^24 + 149
    step = 1 << 24
    buf = torch.zeros(36 * step + 160, dtype=torch.complex64, device="cuda")
    try:
        big = torch.as_strided(buf, (150, 37), (1, step))
        big.copy_(crandn(150, 37))
        bigd = big.contiguous().to(torch.complex128)
        B = crandn(37, 160)
        C = torch.zeros(150, 160, dtype=torch.complex64, device="cuda")
        with forced("c64_bf16x3_144x144"):
            t1m_amd.contract(big, "ik", B, "kj", C, "ij", precision="bf16x3")
        ref = bigd @ B.to(torch.complex128)
        err = rel_err(C.cpu().numpy(), ref.cpu().numpy())
        print(f"  span > 2^32 bytes as A rel={err:.2e}")
        assert err < tol, err
        A = crandn(160, 150)
        C2 = torch.zeros(160, 37, dtype=torch.complex64, device="cuda")
        with forced("c64_bf16x3_144x144"):
            t1m_amd.contract(A, "ik", big, "kj", C2, "ij", precision="bf16x3")
        ref2 = A.to(torch.complex128) @ bigd
        err = rel_err(C2.cpu().numpy(), ref2.cpu().numpy())
        print(f"  span > 2^32 bytes as B rel={err:.2e}")
        assert err < tol, err
    finally:
        del buf
        torch.cuda.empty_cache()


PAIR_CASES = [
    # (labels, extents): both operands pairable (unit stride, even leading
    # extent) in all four gather-direction combinations, ragged tiles and K
    (("ik", "kj", "ij"), {"i": 150, "j": 161, "k": 48}),
    (("ki", "kj", "ij"), {"i": 146, "j": 70, "k": 34}),
    (("ik", "jk", "ij"), {"i": 98, "j": 66, "k": 50}),
    (("ki", "jk", "ij"), {"i": 40, "j": 290, "k": 18}),
    (("abcd", "dbef", "acef"), {"a": 12, "b": 6, "c": 11, "d": 8, "e": 10, "f": 7}),
    (("asb", "btc", "astc"), {"a": 38, "s": 4, "b": 30, "t": 4, "c": 22}),
    # a batch label and an odd leading extent (pairs on one operand only)
    (("zik", "zkj", "zij"), {"z": 3, "i": 70, "k": 16, "j": 33}),
    (("ik", "kj", "ij"), {"i": 71, "j": 64, "k": 40}),
]


@pytest.mark.parametrize("case", range(len(PAIR_CASES)))
def test_two_element_gathers(case):
    """complex<float> gathers of two neighbours per lane (16-byte loads),
    taken when an operand's gather direction leads with a unit-stride mode of
    even extent, against the oracle and against the one-element gathers
    (T1M_PAIRS=0), on both FP32 tile families, the thin tiles and BF16x3."""
    (la, lb, lc), ext = PAIR_CASES[case]
    rng = np.random.default_rng(300 + case)
    A = rand(rng, [ext[x] for x in la], np.complex64)
    B = rand(rng, [ext[x] for x in lb], np.complex64)
    C0 = rand(rng, [ext[x] for x in lc], np.complex64)
    want = oracle_einsum(A, la, B, lb, C0, lc)
    for kern, prec, tolkey in (("c64_f32mfma_144x144", "native", "fp32"),
                               ("c64_f32mfma_64x64", "native", "fp32"),
                               ("c64_bf16x3_144x144", "bf16x3", "bf16x3")):
        got = {}
        for pairs in ("1", "0"):
            old = os.environ.get("T1M_PAIRS")
            os.environ["T1M_PAIRS"] = pairs
            try:
                with forced(kern):
                    got[pairs] = run(A, la, B, lb, C0, lc, precision=prec)
            finally:
                if old is None:
                    os.environ.pop("T1M_PAIRS", None)
                else:
                    os.environ["T1M_PAIRS"] = old
            err = rel_err(got[pairs] - C0, want - C0)
            print(f"  {la},{lb}->{lc} {kern} pairs={pairs} rel={err:.2e}")
            assert err < TOL[tolkey], (kern, pairs, err)
        # the same products in the same order: identical sums
        assert np.array_equal(got["1"], got["0"]), kern
