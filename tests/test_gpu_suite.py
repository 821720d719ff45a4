"""The reference's benchmark suite as a parity test: all 24 contractions of
benchmark/benchmark.cpp:87-187 at their real extents (odd 148 / 156, the
2568 x 2560 x 2568 GEMM, 5- and 6-mode tensors, every permutation class),
for float, double, complex<float> and complex<double> (benchmark.cpp:189-194).

Each result is checked against a float64 / complex128 transpose-GEMM-
transpose product on the device (operands permuted to GEMM layout, made
contiguous, one 2-D matmul), and, for the contractions of at most 6.1e7
multiply-adds, against the oracle's restatement of the reference's five-loop
1M driver (t1m.h:432-557) on the same operands.  Inputs re/im i.i.d.
U[-1, 1], seeded.  Tolerances (relative Frobenius): float / complex<float>
2e-6 against the float64 product and 5e-6 against the oracle (both sides
round in float); double / complex<double> 1e-13 and 1e-12."""
import math

import numpy as np
import pytest
import torch

import t1m_amd
from oracle import oracle as orc
from t1m_amd.suite import SUITE
from util import rel_err, to_host

pytestmark = pytest.mark.gpu
DTYPES = {"f32": torch.float32, "f64": torch.float64, "c64": torch.complex64,
          "c128": torch.complex128}
WIDE = {torch.float32: torch.float64, torch.float64: torch.float64,
        torch.complex64: torch.complex128, torch.complex128: torch.complex128}
ORACLE_MAX_MNK = 61_000_000


def _colmajor(shape, dt, gen):
    rev = list(reversed(shape))
    rdt = torch.float64 if dt in (torch.float64, torch.complex128) else torch.float32
    x = torch.rand(rev, generator=gen, device="cuda", dtype=rdt) * 2 - 1
    if dt.is_complex:
        x = torch.complex(x, torch.rand(rev, generator=gen, device="cuda", dtype=rdt) * 2 - 1)
    return x.to(dt).permute(*reversed(range(len(shape))))


def _ttgt(a, la, b, lb, lc, wide):
    """C = A.B over labels through one 2-D matmul in `wide` precision."""
    mk = [ch for ch in lc if ch in la]
    nk = [ch for ch in lc if ch in lb]
    kk = [ch for ch in la if ch in lb]
    ext = dict(zip(la, a.shape))
    ext.update(zip(lb, b.shape))
    M = math.prod(ext[ch] for ch in mk)
    N = math.prod(ext[ch] for ch in nk)
    K = math.prod(ext[ch] for ch in kk)
    at = a.permute(*[la.index(ch) for ch in mk + kk]).to(wide).contiguous().reshape(M, K)
    bt = b.permute(*[lb.index(ch) for ch in kk + nk]).to(wide).contiguous().reshape(K, N)
    c = (at @ bt).reshape([ext[ch] for ch in mk + nk])
    return c.permute(*[(mk + nk).index(ch) for ch in lc])


def _mnk(dims_a, la, lb, dims_c, lc):
    ext = dict(zip(la, dims_a))
    ext.update(zip(lc, dims_c))
    return math.prod(ext[ch] for ch in set(la) | set(lb))


@pytest.mark.parametrize("name", list(DTYPES))
def test_reference_suite_all_24_contractions(name):
    dt = DTYPES[name]
    low = dt in (torch.float32, torch.complex64)
    tol_wide, tol_orc = (2e-6, 5e-6) if low else (1e-13, 1e-12)
    gen = torch.Generator(device="cuda").manual_seed(24)
    for i, (dims_a, la, dims_b, lb, dims_c, lc) in enumerate(SUITE):
        case = f"{name} #{i} {la},{lb}->{lc}"
        print(f"  [suite] {case}", flush=True)
        a, b = _colmajor(dims_a, dt, gen), _colmajor(dims_b, dt, gen)
        c = torch.zeros(list(reversed(dims_c)), dtype=dt, device="cuda")
        c = c.permute(*reversed(range(len(dims_c))))
        t1m_amd.contract(a, la, b, lb, c, lc)
        ref = _ttgt(a, la, b, lb, lc, WIDE[dt])
        e = float(torch.linalg.vector_norm(c.to(WIDE[dt]) - ref) / torch.linalg.vector_norm(ref))
        assert e < tol_wide, (case, e)
        if _mnk(dims_a, la, lb, dims_c, lc) <= ORACLE_MAX_MNK:
            A, B = to_host(a), to_host(b)
            want = np.zeros(dims_c, dtype=A.dtype, order="F")
            orc.contract(np.asfortranarray(A), la, np.asfortranarray(B), lb, want, lc)
            eo = rel_err(to_host(c), want)
            assert eo < tol_orc, (case, eo)
        del a, b, c, ref
