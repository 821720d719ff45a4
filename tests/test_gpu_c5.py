"""The batched + differentiation workload at its configured size: 512
complex<float> MPS transfer contractions 'asb,btc->astc' (physical dimension
4, bond dimensions (Dl, Db, Dr) drawn uniformly from {16, 32, ..., 256} with
seed 5, re/im i.i.d. U[-1, 1]) in one persistent graph launch, then the VJP
of L = sum |C|^2 (grad_C = 2C, torch's convention) as two more persistent
launches: grad_A = grad_C . conj(B), grad_B = conj(A) . grad_C.

Every instance is checked at full size against complex128 products of the
same operands in GEMM layout (A as (a s) x b, B as b x (t c), C as
(a s) x (t c)); the smallest instances also against the oracle's label loops
and complex128 central differences (h = 1e-6) of the loss.  Tolerances
(relative Frobenius, per instance): 2e-6 forward, 1e-5 gradients."""
import random

import numpy as np
import pytest
import torch

import t1m_amd
from oracle import oracle as orc
from util import rel_err, to_host

pytestmark = pytest.mark.gpu
INSTANCES = 512
TOL_FWD, TOL_GRAD = 2e-6, 1e-5


def _bond_dims(seed=5, count=INSTANCES):
    """(Dl, Db, Dr) per instance, each uniform over {16, 32, ..., 256}."""
    rnd = random.Random(seed)
    return [tuple(rnd.choice(range(16, 257, 16)) for _ in range(3)) for _ in range(count)]


def _cmat(shape, gen):
    """Column-major complex<float> device tensor, re/im i.i.d. U[-1, 1]."""
    rev = list(reversed(shape))
    z = torch.complex(torch.rand(rev, generator=gen, device="cuda") * 2 - 1,
                      torch.rand(rev, generator=gen, device="cuda") * 2 - 1)
    return z.permute(*reversed(range(len(shape))))


def _gemm(x, rows):
    """A logical tensor as a complex128 matrix: its first `rows` modes (last
    fastest, i.e. row-major) by the rest."""
    x = x.to(torch.complex128).contiguous()
    r = int(np.prod(x.shape[:rows]))
    return x.reshape(r, -1)


def _rel(got, want):
    return float(torch.linalg.vector_norm(got.to(torch.complex128) - want) /
                 torch.linalg.vector_norm(want))


def test_c5_mps512_forward_and_vjp_full_size():
    dims = _bond_dims()
    assert len(dims) == INSTANCES and min(min(d) for d in dims) == 16 and \
        max(max(d) for d in dims) == 256
    gen = torch.Generator(device="cuda").manual_seed(5)
    A = [_cmat((dl, 4, db), gen) for dl, db, dr in dims]
    B = [_cmat((db, 4, dr), gen) for dl, db, dr in dims]
    C = [torch.zeros((dr, 4, 4, dl), dtype=torch.complex64, device="cuda").permute(3, 2, 1, 0)
         for dl, db, dr in dims]
    fwd = t1m_amd.Batch(A, "asb", B, "btc", C, "astc")
    fwd.launch()
    torch.cuda.synchronize()
    G = [2 * c for c in C]                                  # dL/dC* for L = sum |C|^2
    GA = [torch.zeros_like(a) for a in A]
    GB = [torch.zeros_like(b) for b in B]
    vja = t1m_amd.Batch(G, "astc", B, "btc", GA, "asb", conj_b=True)
    vjb = t1m_amd.Batch(A, "asb", G, "astc", GB, "btc", conj_a=True)
    vja.launch()
    vjb.launch()
    torch.cuda.synchronize()
    for bt in (fwd, vja, vjb):
        bt.close()

    worst = {"fwd": 0.0, "ga": 0.0, "gb": 0.0}
    for i, (dl, db, dr) in enumerate(dims):
        a, b = _gemm(A[i], 2), _gemm(B[i], 1)                # (a s) x b, b x (t c)
        c_ref = a @ b                                         # (a s) x (t c)
        e = _rel(_gemm(C[i], 2), c_ref)
        assert e < TOL_FWD, (i, dims[i], e)
        g = _gemm(G[i], 2)
        ga_ref = g @ b.conj().transpose(0, 1).contiguous()     # (a s) x b
        gb_ref = a.conj().transpose(0, 1).contiguous() @ g     # b x (t c)
        ea, eb = _rel(_gemm(GA[i], 2), ga_ref), _rel(_gemm(GB[i], 1), gb_ref)
        assert ea < TOL_GRAD, (i, dims[i], ea)
        assert eb < TOL_GRAD, (i, dims[i], eb)
        worst["fwd"] = max(worst["fwd"], e)
        worst["ga"] = max(worst["ga"], ea)
        worst["gb"] = max(worst["gb"], eb)
    print(f"  c5 512 instances: worst rel err fwd={worst['fwd']:.2e} "
          f"grad_A={worst['ga']:.2e} grad_B={worst['gb']:.2e}", flush=True)

    # The smallest instances: the oracle's label loops (complex128) for C and
    # both gradients, and central differences of the loss.
    small = sorted(range(len(dims)), key=lambda i: np.prod(dims[i]))[:3]
    rng = np.random.default_rng(5)
    for i in small:
        Ah = to_host(A[i]).astype(np.complex128)
        Bh = to_host(B[i]).astype(np.complex128)
        Cw = np.zeros(C[i].shape, np.complex128, order="F")
        orc.einsum(Ah, "asb", Bh, "btc", Cw, "astc")
        assert rel_err(to_host(C[i]), Cw) < TOL_FWD, i
        gaw = np.zeros_like(Ah, order="F")
        gbw = np.zeros_like(Bh, order="F")
        orc.einsum(2 * Cw, "astc", Bh, "btc", gaw, "asb", conj_b=True)
        orc.einsum(Ah, "asb", 2 * Cw, "astc", gbw, "btc", conj_a=True)
        ga, gb = to_host(GA[i]), to_host(GB[i])
        assert rel_err(ga, gaw) < TOL_GRAD, i
        assert rel_err(gb, gbw) < TOL_GRAD, i

        def loss(Ax, Bx):
            return float(np.sum(np.abs(np.einsum("asb,btc->astc", Ax, Bx)) ** 2))

        h = 1e-6
        for _ in range(2):
            ia = tuple(int(rng.integers(0, s)) for s in Ah.shape)
            ib = tuple(int(rng.integers(0, s)) for s in Bh.shape)
            for unit, part in ((1.0, "re"), (1j, "im")):
                Ap, Am = Ah.copy(), Ah.copy()
                Ap[ia] += unit * h
                Am[ia] -= unit * h
                fd = (loss(Ap, Bh) - loss(Am, Bh)) / (2 * h)
                an = ga[ia].real if part == "re" else ga[ia].imag
                assert abs(fd - an) <= 1e-4 * max(1.0, abs(fd)), (i, ia, part, fd, an)
                Bp, Bm = Bh.copy(), Bh.copy()
                Bp[ib] += unit * h
                Bm[ib] -= unit * h
                fd = (loss(Ah, Bp) - loss(Ah, Bm)) / (2 * h)
                an = gb[ib].real if part == "re" else gb[ib].imag
                assert abs(fd - an) <= 1e-4 * max(1.0, abs(fd)), (i, ib, part, fd, an)
