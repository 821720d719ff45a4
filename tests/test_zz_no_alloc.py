"""t1m_contract_noalloc against t1m_contract, through the C ABI.

With deterministic = 1 both run the same kernels on the same index maps, so
their results must agree bit for bit; each is also checked against the
oracle's label loops (relative Frobenius 2e-6 float, 1e-13 double).  Every
case runs two calls in a row, the second on a smaller input, so the second
call's maps overwrite the first's in the shared scratch array.  Also: a
first call of a new shape captured into a graph and replayed, and the
refusals (host memory, autotune, maps larger than the scratch array)."""
import ctypes

import numpy as np
import pytest
import torch

import t1m_amd
from oracle import oracle as orc
from util import rand, rel_err, to_dev, to_host

pytestmark = pytest.mark.gpu

# (labels, extents of the larger call, extents of the smaller call)
CASES = [
    (("abcd", "dbef", "acef"), dict(a=24, b=12, c=20, d=16, e=24, f=12),
     dict(a=9, b=5, c=7, d=6, e=11, f=3)),
    (("asb", "btc", "astc"), dict(a=64, s=4, b=96, t=4, c=48),
     dict(a=16, s=4, b=24, t=4, c=20)),
    (("zab", "zbc", "zac"), dict(z=6, a=40, b=30, c=50),          # batch label
     dict(z=3, a=10, b=7, c=9)),
    (("ik", "kj", "ij"), dict(i=8, k=4096, j=8),                  # split-K sized
     dict(i=5, k=300, j=6)),
]
DTYPES = [(np.complex64, "native", 2e-6), (np.complex64, "bf16x3", 1e-5),
          (np.complex128, "native", 1e-13), (np.float32, "native", 2e-6),
          (np.float64, "native", 1e-13)]


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.nodeid}", flush=True)
    yield
    torch.cuda.synchronize()


class _Call:
    """Descriptors for one call through the C ABI (kept alive here)."""

    def __init__(self, a, b, c):
        self.d = [t1m_amd._Desc(t) for t in (a, b, c)]

    def run(self, fn, la, lb, lc, precision, deterministic=1, stream=None, synchronize=True):
        o = t1m_amd._options(precision=precision, stream=stream, deterministic=deterministic,
                             synchronize=synchronize)
        return fn(ctypes.byref(self.d[0].desc), la.encode(), ctypes.byref(self.d[1].desc),
                  lb.encode(), ctypes.byref(self.d[2].desc), lc.encode(), ctypes.byref(o))


def _one(lib, labels, ext, dtype, prec, tol, rng):
    la, lb, lc = labels
    A = rand(rng, [ext[x] for x in la], dtype)
    B = rand(rng, [ext[x] for x in lb], dtype)
    C0 = rand(rng, [ext[x] for x in lc], dtype)
    a, b = to_dev(A), to_dev(B)
    c_ref, c_new = to_dev(C0), to_dev(C0)
    call = _Call(a, b, c_ref)
    assert call.run(lib.t1m_contract, la, lb, lc, prec) == 0, lib.t1m_last_error()
    call = _Call(a, b, c_new)
    assert call.run(lib.t1m_contract_noalloc, la, lb, lc, prec) == 0, lib.t1m_last_error()
    torch.cuda.synchronize()
    got, ref = to_host(c_new), to_host(c_ref)
    bits = lambda x: np.ascontiguousarray(x).view(np.uint8)
    assert np.array_equal(bits(got), bits(ref)), (labels, dtype, prec)
    want = C0.copy(order="F")
    orc.einsum(A, la, B, lb, want, lc)
    assert rel_err(got - C0, want - C0) < tol, (labels, dtype, prec)


@pytest.mark.parametrize("dtype,prec,tol", DTYPES)
def test_noalloc_matches_contract_two_calls(dtype, prec, tol):
    lib = t1m_amd.load()
    rng = np.random.default_rng(301)
    for labels, big, small in CASES:
        if prec == "bf16x3" and labels[0] == "ik":
            continue
        _one(lib, labels, big, dtype, prec, tol, rng)
        _one(lib, labels, small, dtype, prec, tol, rng)


def test_noalloc_first_call_captured_in_a_graph():
    """A shape never seen before, first called under graph capture (which
    t1m_contract cannot do: it allocates and waits); every replay adds A.B."""
    lib = t1m_amd.load()
    rng = np.random.default_rng(302)
    la, lb, lc = "abc", "cbd", "ad"
    A = rand(rng, (37, 5, 11), np.complex64)
    B = rand(rng, (11, 5, 29), np.complex64)
    C0 = rand(rng, (37, 29), np.complex64)
    prod = C0.copy(order="F") * 0
    orc.einsum(A, la, B, lb, prod, lc)
    a, b, c = to_dev(A), to_dev(B), to_dev(C0)
    # one eager call of another shape first: the library looks up the scratch
    # array's device address on a device's first call, outside any capture
    w = torch.zeros(3, 2, dtype=torch.complex64, device="cuda")
    t1m_amd.contract_noalloc(w, "ik", w.T.contiguous(), "kj",
                             torch.zeros(3, 3, dtype=torch.complex64, device="cuda"), "ij")
    call = _Call(a, b, c)
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        rc = call.run(lib.t1m_contract_noalloc, la, lb, lc, "native",
                      stream=torch.cuda.current_stream().cuda_stream, synchronize=False)
    assert rc == 0, lib.t1m_last_error()
    base = to_host(c)
    assert np.array_equal(base, C0)                  # capture ran nothing
    for r in range(3):
        g.replay()
        torch.cuda.synchronize()
        got = to_host(c)
        assert rel_err(got - base, prod) < 2e-6, r
        base = got
        # another no-alloc call in between overwrites the scratch maps;
        # the next replay rebuilds its own first
        x = torch.randn(12, 7, dtype=torch.complex64, device="cuda")
        y = torch.randn(7, 9, dtype=torch.complex64, device="cuda")
        z = torch.zeros(12, 9, dtype=torch.complex64, device="cuda")
        t1m_amd.contract_noalloc(x, "ik", y, "kj", z, "ij")
        assert rel_err(z.cpu().numpy(), (x @ y).cpu().numpy()) < 2e-6
    del g


def test_noalloc_refusals():
    lib = t1m_amd.load()
    x = torch.ones(4, 3, dtype=torch.complex64)
    y = torch.ones(3, 5, dtype=torch.complex64)
    z = torch.zeros(4, 5, dtype=torch.complex64)
    call = _Call(x, y, z)                               # host memory
    assert call.run(lib.t1m_contract_noalloc, "ik", "kj", "ij", "native") == 1
    assert b"device memory" in lib.t1m_last_error()
    xd, yd, zd = x.cuda(), y.cuda(), z.cuda()
    call = _Call(xd, yd, zd)
    o = t1m_amd._options(autotune=True)
    d = call.d
    assert lib.t1m_contract_noalloc(ctypes.byref(d[0].desc), b"ik", ctypes.byref(d[1].desc),
                                    b"kj", ctypes.byref(d[2].desc), b"ij",
                                    ctypes.byref(o)) == 1
    assert b"autotune" in lib.t1m_last_error()
    # M = 1.1e6 rows: 2 x (M + N + K) int32 offsets exceed the 8 MiB array
    big_m = 1_100_000
    xa = torch.ones(big_m, 1, dtype=torch.complex64, device="cuda")
    yb = torch.ones(1, 1, dtype=torch.complex64, device="cuda")
    zc = torch.zeros(big_m, 1, dtype=torch.complex64, device="cuda")
    call = _Call(xa, yb, zc)
    assert call.run(lib.t1m_contract_noalloc, "ik", "kj", "ij", "native") == 4
    assert b"scratch" in lib.t1m_last_error()
    assert float(zc.abs().sum()) == 0.0               # nothing ran
    t1m_amd.contract(xa, "ik", yb, "kj", zc, "ij")    # the allocating path takes it
    assert float(zc.real.sum()) == big_m
