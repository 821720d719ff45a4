"""GPU tests of the host interface's safety rules: torch's lazy conj / neg
views, plans held by captured graphs, tuning-file validation and the
lifetime of a batch's tensors.  Values against the oracle or complex128
references computed on the CPU; tolerances (relative Frobenius) as written."""
import gc
import os

import numpy as np
import pytest
import torch

import t1m_amd
from oracle import oracle as orc
from util import rand, rel_err, to_dev, to_host

pytestmark = pytest.mark.gpu
TOL32, TOL64 = 2e-6, 1e-13


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.nodeid}", flush=True)
    yield
    torch.cuda.synchronize()


def _einsum(A, la, B, lb, C0, lc, **kw):
    c = C0.copy(order="F")
    orc.einsum(A, la, B, lb, c, lc, **kw)
    return c


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_conj_views_are_read_as_conjugates(dtype):
    """contract(a.conj(), b) must use conj(a): the view's memory is
    unconjugated, the library reads it with the conj flag.  Both operands,
    combined with explicit flags (which then cancel), and a neg view."""
    rng = np.random.default_rng(101)
    la, lb, lc = "asb", "btc", "astc"
    A = rand(rng, (7, 4, 9), dtype)
    B = rand(rng, (9, 4, 6), dtype)
    C0 = rand(rng, (7, 4, 4, 6), dtype)
    tol = TOL32 if dtype == np.complex64 else TOL64
    a, b = to_dev(A), to_dev(B)
    cases = [
        (a.conj(), b, {}, dict(conj_a=True)),
        (a, b.conj(), {}, dict(conj_b=True)),
        (a.conj(), b.conj(), {}, dict(conj_a=True, conj_b=True)),
        (a.conj(), b, dict(conj_a=True), {}),                 # flag and view cancel
        (torch._neg_view(a), b, {}, dict(alpha=-1.0)),
    ]
    for i, (x, y, flags, want_kw) in enumerate(cases):
        c = to_dev(C0)
        t1m_amd.contract(x, la, y, lb, c, lc, **flags)
        want = _einsum(A, la, B, lb, C0, lc, **want_kw)
        assert rel_err(to_host(c) - C0, want - C0) < tol, i


def test_conj_or_neg_outputs_are_refused():
    c = torch.zeros(4, 5, dtype=torch.complex64, device="cuda")
    a = torch.ones(4, 3, dtype=torch.complex64, device="cuda")
    b = torch.ones(3, 5, dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError):
        t1m_amd.contract(a, "ik", b, "kj", c.conj(), "ij")
    with pytest.raises(ValueError):
        t1m_amd.contract(a, "ik", b, "kj", torch._neg_view(c), "ij")
    ga = torch.zeros_like(a)
    with pytest.raises(ValueError):
        t1m_amd.contract_vjp(a, "ik", b, "kj", c, "ij", ga.conj(), None)


def test_vjp_conj_flags_differentiate_the_conjugated_operands():
    """t1m_contract_vjp with conj_a / conj_b: the forward is C = A'.B' with
    A' = conj(A); gradients are with respect to A' and B' (grad_a = grad_c .
    conj(B'), grad_b = conj(A') . grad_c), checked against the oracle."""
    rng = np.random.default_rng(102)
    la, lb, lc = "asb", "btc", "astc"
    A = rand(rng, (6, 4, 5), np.complex128)
    B = rand(rng, (5, 4, 7), np.complex128)
    G = rand(rng, (6, 4, 4, 7), np.complex128)
    a, b, g = to_dev(A), to_dev(B), to_dev(G)
    for cj_a in (False, True):
        for cj_b in (False, True):
            Ap = np.conj(A) if cj_a else A
            Bp = np.conj(B) if cj_b else B
            ga_want = _einsum(G, lc, Bp, lb, np.zeros_like(A, order="F"), la, conj_b=True)
            gb_want = _einsum(Ap, la, G, lc, np.zeros_like(B, order="F"), lb, conj_a=True)
            ga = to_dev(np.zeros_like(A, order="F"))
            gb = to_dev(np.zeros_like(B, order="F"))
            x = a.conj() if cj_a else a        # the Python layer maps the bits
            y = b.conj() if cj_b else b        # onto the ABI's conj flags
            t1m_amd.contract_vjp(x, la, y, lb, g, lc, ga, gb)
            assert rel_err(to_host(ga), ga_want) < TOL64, (cj_a, cj_b)
            assert rel_err(to_host(gb), gb_want) < TOL64, (cj_a, cj_b)


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_autograd_through_conjugated_output_and_inputs(dtype):
    """A loss through C.conj() makes torch hand backward() a lazy conj view of
    grad_C; inputs given as conj views too.  Gradients against torch's own
    einsum autograd in complex128 on the CPU."""
    torch.manual_seed(103)
    A = torch.randn(6, 4, 5, dtype=dtype)
    B = torch.randn(5, 4, 7, dtype=dtype)
    W = torch.randn(6, 4, 4, 7, dtype=torch.complex128)
    tol = 1e-5 if dtype == torch.complex64 else 1e-12

    def loss_t1m(a, b):
        C = t1m_amd.contract_autograd(a.conj(), "asb", b, "btc", "astc")
        return (C.conj() * W.to(dtype).cuda()).real.sum()

    def loss_ref(a, b):
        C = torch.einsum("asb,btc->astc", a.conj(), b)
        return (C.conj() * W).real.sum()

    a = A.cuda().requires_grad_(True)
    b = B.cuda().requires_grad_(True)
    loss_t1m(a, b).backward()
    a2 = A.to(torch.complex128).requires_grad_(True)
    b2 = B.to(torch.complex128).requires_grad_(True)
    loss_ref(a2, b2).backward()
    # (torch may hand back the gradient of a conj input as a conj view)
    assert rel_err(a.grad.resolve_conj().cpu().numpy(), a2.grad.resolve_conj().numpy()) < tol
    assert rel_err(b.grad.resolve_conj().cpu().numpy(), b2.grad.resolve_conj().numpy()) < tol


def test_captured_graph_survives_plan_cache_clear():
    """A graph captured around a cached plan keeps the plan (index maps)
    alive: clearing the plan cache and building other plans in between must
    not change what replays compute."""
    rng = np.random.default_rng(104)
    la, lb, lc = "abcd", "dbef", "acef"
    ext = dict(a=12, b=8, c=12, d=8, e=16, f=8)
    A = rand(rng, [ext[x] for x in la], np.complex64)
    B = rand(rng, [ext[x] for x in lb], np.complex64)
    C0 = rand(rng, [ext[x] for x in lc], np.complex64)
    prod = _einsum(A, la, B, lb, np.zeros_like(C0, order="F"), lc)
    a, b, c = to_dev(A), to_dev(B), to_dev(C0)
    t1m_amd.contract(a, la, b, lb, c, lc)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        t1m_amd.contract(a, la, b, lb, c, lc, synchronize=False)
    base = to_host(c)
    for r in range(3):
        t1m_amd.plan_cache_clear()
        assert t1m_amd.plan_cache_size() == 0
        # new plans reuse freed device memory if the old maps were freed
        for n in range(4):
            x = torch.randn(16 + n, 8, dtype=torch.complex64, device="cuda")
            y = torch.randn(8, 12 + n, dtype=torch.complex64, device="cuda")
            z = torch.zeros(16 + n, 12 + n, dtype=torch.complex64, device="cuda")
            t1m_amd.contract(x, "ik", y, "kj", z, "ij")
        g.replay()
        torch.cuda.synchronize()
        got = to_host(c)
        assert rel_err(got - base, prod) < TOL32, r
        base = got
    del g
    gc.collect()
    t1m_amd.plan_cache_clear()          # releases the plan the graph parked


def test_corrupt_tuning_entries_are_ignored(tmp_path):
    """A tuning-file line naming a kernel of another dtype for a plan's key
    (stale or edited file) is ignored: the plan keeps a kernel that fits."""
    path = tmp_path / "tuning.txt"
    old = os.environ.get("T1M_TUNE_CACHE")
    os.environ["T1M_TUNE_CACHE"] = str(path)
    try:
        t1m_amd.plan_cache_clear()
        rng = np.random.default_rng(105)
        A = rand(rng, (40, 30), np.complex64)
        B = rand(rng, (30, 50), np.complex64)
        C0 = np.zeros((40, 50), np.complex64, order="F")
        want = _einsum(A, "ik", B, "kj", C0, "ij")
        a, b, c = to_dev(A), to_dev(B), to_dev(C0)
        t1m_amd.contract(a, "ik", b, "kj", c, "ij", autotune=True)
        lines = path.read_text().split("\n")
        key = next(ln.split()[0] for ln in lines if ln.strip())
        for bad in ("c128_f64mfma_64x64", "f32mfma_128x128", "no_such_kernel"):
            path.write_text(f"{key} {bad}\n")
            t1m_amd.plan_cache_clear()            # re-reads the tuning file
            k = t1m_amd.plan_kernel(a, "ik", b, "kj", c, "ij")
            assert k.startswith("c64_f32mfma"), (bad, k)
            c.zero_()
            t1m_amd.contract(a, "ik", b, "kj", c, "ij")
            assert rel_err(to_host(c), want) < TOL32, bad
        path.write_text(f"{key} c64_f32mfma_64x64\n")   # a fitting entry is used
        t1m_amd.plan_cache_clear()
        assert t1m_amd.plan_kernel(a, "ik", b, "kj", c, "ij") == "c64_f32mfma_64x64"
    finally:
        if old is None:
            os.environ.pop("T1M_TUNE_CACHE", None)
        else:
            os.environ["T1M_TUNE_CACHE"] = old
        t1m_amd.plan_cache_clear()


def test_batch_keeps_temporary_inputs_alive():
    """Batch built from temporaries (the only references are inside the
    batch): after the caller's lists are gone and torch's allocator has
    handed out fresh blocks of the same sizes, replays still read the
    original operands."""
    rng = np.random.default_rng(106)
    la, lb, lc = "asb", "btc", "astc"
    dims = [(16, 32, 24), (40, 16, 16), (24, 48, 32)]
    As = [rand(rng, (dl, 4, db), np.complex64) for dl, db, dr in dims]
    Bs = [rand(rng, (db, 4, dr), np.complex64) for dl, db, dr in dims]
    Cs = [to_dev(np.zeros((dl, 4, 4, dr), np.complex64, order="F")) for dl, db, dr in dims]
    batch = t1m_amd.Batch([to_dev(A).contiguous() for A in As], la,
                          [to_dev(B).contiguous() for B in Bs], lb, Cs, lc)
    gc.collect()
    junk = [torch.full((dl, 4, db), 7.0 + 0j, dtype=torch.complex64, device="cuda")
            for dl, db, dr in dims] + \
           [torch.full((db, 4, dr), 7.0 + 0j, dtype=torch.complex64, device="cuda")
            for dl, db, dr in dims]
    batch.launch()
    torch.cuda.synchronize()
    for A, B, c in zip(As, Bs, Cs):
        want = _einsum(A, la, B, lb, np.zeros(c.shape, np.complex64, order="F"), lc)
        assert rel_err(to_host(c), want) < TOL32
    batch.close()
    with pytest.raises(RuntimeError):
        batch.launch()
    del junk


def _batch_case(rng, shapes):
    As, Bs, Cs, C0s = [], [], [], []
    for m, k, n in shapes:
        A = rand(rng, (m, k), np.complex64)
        B = rand(rng, (k, n), np.complex64)
        C0 = rand(rng, (m, n), np.complex64)
        As.append(to_dev(A)); Bs.append(to_dev(B)); Cs.append(to_dev(C0)); C0s.append(C0)
    return As, Bs, Cs, C0s


def _check_batch(As, Bs, Cs, C0s, tol=TOL32):
    for a, b, c, c0 in zip(As, Bs, Cs, C0s):
        want = (a.to(torch.complex128) @ b.to(torch.complex128)).cpu().numpy()
        assert rel_err(to_host(c) - c0, want) < tol


# Large instances tiled evenly by 144 x 144 (576 = 4 x 144, K >= 128) beside
# small ones: every tiling mode puts different instances on the large tiles.
MIXED_SHAPES = [(576, 256, 576), (288, 160, 432), (576, 128, 288),
                (64, 32, 64), (96, 48, 80), (128, 16, 64), (200, 64, 150)]


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_batch_tiling_modes_compute_the_same(mode):
    """Each tiling mode of a complex<float> batch (all 64 x 64; evenly tiled
    large instances on 144 x 144; every instance >= 144 x 144 on 144 x 144)
    gives the products, on instances of both kinds, against complex128."""
    rng = np.random.default_rng(107 + mode)
    As, Bs, Cs, C0s = _batch_case(rng, MIXED_SHAPES)
    old = os.environ.get("T1M_BATCH_MODE")
    os.environ["T1M_BATCH_MODE"] = str(mode)
    try:
        bt = t1m_amd.Batch(As, "ik", Bs, "kj", Cs, "ij")
    finally:
        if old is None:
            os.environ.pop("T1M_BATCH_MODE", None)
        else:
            os.environ["T1M_BATCH_MODE"] = old
    desc = bt.describe()
    print(" ", desc)
    assert desc.startswith(f"mode {mode}")
    if mode > 0:
        assert "c64_f32mfma_144x144" in desc
    bt.launch()
    torch.cuda.synchronize()
    _check_batch(As, Bs, Cs, C0s)
    bt.close()


@pytest.mark.parametrize("xcds", ["1", "8", "3"])
@pytest.mark.parametrize("mode", [0, 1])
def test_batch_xcd_ordered_work_lists(xcds, mode):
    """Work lists ordered for several L2 domains (whole instances per
    domain, interleaved, padded with skipped entries) give the products of
    the plain LPT list.  3 domains divide a grid of 3 workgroups per CU
    (the 64 x 64 phase) but not the 144 x 144 phase's one per CU, which then
    keeps the plain list."""
    rng = np.random.default_rng(109)
    As, Bs, Cs, C0s = _batch_case(rng, MIXED_SHAPES)
    saved = {k: os.environ.get(k) for k in ("T1M_BATCH_MODE", "T1M_BATCH_XCDS")}
    os.environ["T1M_BATCH_MODE"] = str(mode)
    os.environ["T1M_BATCH_XCDS"] = xcds
    try:
        bt = t1m_amd.Batch(As, "ik", Bs, "kj", Cs, "ij")
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    desc = bt.describe()
    print(" ", desc)
    if xcds == "3":
        assert "3 L2 domains" in desc or "1 L2 domains" in desc
    else:
        assert desc.count(f"{xcds} L2 domains") == desc.count("grid")
    bt.launch()
    torch.cuda.synchronize()
    _check_batch(As, Bs, Cs, C0s)
    bt.launch()                 # a second launch adds the products again
    torch.cuda.synchronize()
    for a, b, c, c0 in zip(As, Bs, Cs, C0s):
        want = 2 * (a.to(torch.complex128) @ b.to(torch.complex128)).cpu().numpy()
        assert rel_err(to_host(c) - c0, want) < 2e-6
    bt.close()


def test_batch_autotune_times_modes_and_remembers(tmp_path):
    """autotune=True times the tiling modes on scratch outputs (C is touched
    once, by the kept graph), writes the choice to the tuning file, and a
    later batch of the same shapes reads it back."""
    path = tmp_path / "tuning.txt"
    old = os.environ.get("T1M_TUNE_CACHE")
    os.environ["T1M_TUNE_CACHE"] = str(path)
    try:
        t1m_amd.plan_cache_clear()
        rng = np.random.default_rng(108)
        As, Bs, Cs, C0s = _batch_case(rng, MIXED_SHAPES)
        bt = t1m_amd.Batch(As, "ik", Bs, "kj", Cs, "ij", autotune=True)
        torch.cuda.synchronize()
        for c, c0 in zip(Cs, C0s):                     # tuning wrote to scratch only
            assert np.array_equal(to_host(c), c0)
        desc = bt.describe()
        print(" ", desc)
        lines = [ln.split() for ln in path.read_text().splitlines() if "batch|" in ln]
        assert len(lines) == 1 and lines[0][1].startswith("batch_mode_"), lines
        mode = int(lines[0][1][len("batch_mode_"):])
        assert desc.startswith(f"mode {mode}")
        bt.launch()
        torch.cuda.synchronize()
        _check_batch(As, Bs, Cs, C0s)
        bt.close()
        As2, Bs2, Cs2, C0s2 = _batch_case(rng, MIXED_SHAPES)
        bt2 = t1m_amd.Batch(As2, "ik", Bs2, "kj", Cs2, "ij")
        assert bt2.describe().startswith(f"mode {mode}")
        bt2.launch()
        torch.cuda.synchronize()
        _check_batch(As2, Bs2, Cs2, C0s2)
        bt2.close()
    finally:
        if old is None:
            os.environ.pop("T1M_TUNE_CACHE", None)
        else:
            os.environ["T1M_TUNE_CACHE"] = old
        t1m_amd.plan_cache_clear()
