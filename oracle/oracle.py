"""ctypes front end of the CPU oracle (oracle/t1m_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker and CPU timing baseline.  The
product (t1m_amd/) never imports it.

Each wrapper names the reference lines (include/t1m/t1m.h) that the C function
restates.  Arrays are numpy arrays; tensors carry explicit element strides, so
Fortran-ordered arrays are the reference's col-major tensors (t1m.h:42-50).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# T1M_ORACLE_LIB selects another build (the sanitizer build).
_LIB_PATH = os.environ.get("T1M_ORACLE_LIB") or os.path.join(_HERE, "_build", "libt1m_oracle.so")
MAX_RANK = 32

DTYPES = {np.dtype(np.float32): 0, np.dtype(np.float64): 1,
          np.dtype(np.complex64): 2, np.dtype(np.complex128): 3}

# BLIS haswell default block sizes {MR, NR, KP, MC, KC, NC}; KP = 4 is fixed by
# the reference (mappings.h:122).  The sgemm sizes serve float/complex<float>,
# the dgemm sizes double/complex<double> (mappings.h:129-142).
BLOCK_SIZES = {
    "s": (6, 16, 4, 168, 256, 4080),
    "d": (6, 8, 4, 72, 256, 4080),
}

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        i64p = ctypes.POINTER(ctypes.c_int64)
        _lib.orc_bundles.argtypes = [ctypes.c_char_p] * 3 + [u64p, i64p]
        _lib.orc_scatter_len.restype = ctypes.c_uint64
        _lib.orc_scatter_len.argtypes = [u64p, ctypes.c_int64, u64p]
        _lib.orc_scatter.argtypes = [u64p, ctypes.c_int64, u64p, u64p, u64p]
        _lib.orc_block_scatter_len.restype = ctypes.c_uint64
        _lib.orc_block_scatter_len.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        _lib.orc_block_scatter.argtypes = [u64p, ctypes.c_uint64, ctypes.c_uint64, u64p]
        view = [ctypes.c_uint64, u64p, ctypes.c_uint64, u64p, ctypes.c_uint64] * 2
        _lib.orc_pack_block.argtypes = ([ctypes.c_int, ctypes.c_int] + view +
                                        [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p])
        _lib.orc_unpack.argtypes = [ctypes.c_int] + view + [ctypes.c_void_p, ctypes.c_void_p]
        tens = [ctypes.c_void_p, ctypes.c_int32, u64p, u64p, ctypes.c_char_p]
        _lib.orc_contract.argtypes = ([ctypes.c_int, ctypes.c_double] + tens * 3 + [u64p])
        _lib.orc_contract_threads.argtypes = ([ctypes.c_int, ctypes.c_double] + tens * 3 +
                                              [u64p, ctypes.c_int])
        etens = [ctypes.c_void_p, i64p, i64p, ctypes.c_char_p]
        _lib.orc_einsum.argtypes = ([ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_int] + etens * 3)
    return _lib


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


def _p64(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def _pi64(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def bundles(la: str, lb: str, lc: str) -> dict:
    """Index bundles AX, AY, BX, BY, CX, CY (t1m.h:90-136).  A label missing
    from the searched string yields 2**64-1 (std::string::npos)."""
    out = np.zeros(6 * MAX_RANK, dtype=np.uint64)
    cnt = np.zeros(6, dtype=np.int64)
    rc = lib().orc_bundles(la.encode(), lb.encode(), lc.encode(), _p64(out), _pi64(cnt))
    if rc:
        raise ValueError("label string too long")
    names = ["AX", "AY", "BX", "BY", "CX", "CY"]
    return {n: [int(x) for x in out[i * MAX_RANK:i * MAX_RANK + cnt[i]]]
            for i, n in enumerate(names)}


def col_major_strides(dims) -> list:
    """t1m.h:42-50."""
    s, acc = [], 1
    for d in dims:
        s.append(acc)
        acc *= d
    return s


def scatter(indices, dims, strides) -> np.ndarray:
    """Scatter vector of a bundle (t1m.h:141-174): first bundle mode fastest."""
    idx, d, s = _u64(indices), _u64(dims), _u64(strides)
    n = lib().orc_scatter_len(_p64(idx), len(idx), _p64(d))
    out = np.zeros(max(int(n), 1), dtype=np.uint64)
    lib().orc_scatter(_p64(idx), len(idx), _p64(d), _p64(s), _p64(out))
    return out[:int(n)] if len(idx) else out


def block_scatter(scat, b: int) -> np.ndarray:
    """Block scatter vector (t1m.h:179-205)."""
    sc = _u64(scat)
    n = lib().orc_block_scatter_len(len(sc), b)
    out = np.zeros(max(int(n), 1), dtype=np.uint64)
    lib().orc_block_scatter(_p64(sc), len(sc), b, _p64(out))
    return out[:int(n)]


class Layout:
    """matrix_layout (t1m.h:246-268): scatter and block-scatter vectors."""

    def __init__(self, dims, strides, X, Y, br, bc):
        self.rs = scatter(X, dims, strides)
        self.cs = scatter(Y, dims, strides)
        self.rbs = block_scatter(self.rs, br)
        self.cbs = block_scatter(self.cs, bc)
        self.br, self.bc = br, bc

    def args(self):
        return [self.br, _p64(self.rbs), len(self.rbs), _p64(self.rs), len(self.rs),
                self.bc, _p64(self.cbs), len(self.cbs), _p64(self.cs), len(self.cs)]


def pack_block(layout: Layout, label: str, length: int, src: np.ndarray,
               dest: np.ndarray) -> None:
    """pack_block<T, A|B> (t1m.h:341-404) of a whole layout into dest (the real
    panel buffer; the caller zero-fills it, as t1m.h:518/527 do)."""
    dt = DTYPES[src.dtype]
    lib().orc_pack_block(dt, 1 if label == "A" else 0, *layout.args(), length,
                         src.ctypes.data, dest.ctypes.data)


def unpack(layout: Layout, src: np.ndarray, dest: np.ndarray) -> None:
    """unpack (t1m.h:406-430): dest[idx] += tile; src is the real tile."""
    dt = DTYPES[dest.dtype]
    lib().orc_unpack(dt, *layout.args(), src.ctypes.data, dest.ctypes.data)


def _estrides(arr: np.ndarray) -> list:
    return [st // arr.itemsize for st in arr.strides]


def contract(a: np.ndarray, la: str, b: np.ndarray, lb: str, c: np.ndarray,
             lc: str, alpha: float = 1.0, block_sizes=None, threads: int = 1) -> None:
    """Five-loop 1M/real contraction C += alpha*A*B (t1m.h:432-557) with the
    reference's packing and a plain C micro-kernel.  Arrays may have any
    non-negative strides.  threads > 1 splits loop 3 over OpenMP threads."""
    dt = DTYPES[a.dtype]
    if block_sizes is None:
        block_sizes = BLOCK_SIZES["s" if dt in (0, 2) else "d"]
    bs = _u64(block_sizes)
    keep = []

    def t(arr):
        d, s = _u64(arr.shape), _u64(_estrides(arr))
        keep.extend([d, s])
        return [arr.ctypes.data, arr.ndim, _p64(d), _p64(s)]

    rc = lib().orc_contract_threads(dt, float(alpha), *t(a), la.encode(), *t(b), lb.encode(),
                                    *t(c), lc.encode(), _p64(bs), int(threads))
    if rc:
        raise RuntimeError(f"orc_contract failed with {rc}")


def einsum(a: np.ndarray, la: str, b: np.ndarray, lb: str, c: np.ndarray,
           lc: str, alpha=1.0, conj_a: bool = False, conj_b: bool = False) -> None:
    """Naive label loops, long-double accumulation: C += alpha * A . op(B)."""
    dt = DTYPES[a.dtype]
    keep = []

    def t(arr):
        d = np.ascontiguousarray(np.asarray(arr.shape, dtype=np.int64))
        s = np.ascontiguousarray(np.asarray(_estrides(arr), dtype=np.int64))
        keep.extend([d, s])
        return [arr.ctypes.data, _pi64(d), _pi64(s)]

    al = complex(alpha)
    rc = lib().orc_einsum(dt, al.real, al.imag, int(conj_a), int(conj_b),
                          *t(a), la.encode(), *t(b), lb.encode(), *t(c), lc.encode())
    if rc:
        raise RuntimeError(f"orc_einsum failed with {rc}")
