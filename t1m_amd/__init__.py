"""t1m on MI355X: transposition-free tensor contraction on gfx950.

Python mirror of the reference's C++ interface (include/t1m/t1m.h):
``contract(a, "abc", b, "bd", c, "acd")`` accumulates ``C += A . B`` over
einsum-style labels (t1m.h:560-577), here on device tensors through the C ABI
of ``libt1m_hip.so`` (include/t1m_hip.h).  torch supplies device memory and
streams only; every contraction runs in the library's own HIP kernels.  There
is no fallback: if the library is missing, import-time loading fails.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

_HERE = os.path.dirname(os.path.abspath(__file__))
# T1M_LIB selects another build of the library (the sanitizer build of
# tests/run_sanitized.sh); by default the in-tree libt1m_hip.so.
LIB_PATH = os.environ.get("T1M_LIB") or os.path.join(_HERE, "libt1m_hip.so")
MAX_RANK = 16

FLOAT32, FLOAT64, COMPLEX64, COMPLEX128 = 0, 1, 2, 3
PRECISIONS = {"native": 0, "bf16x3": 1, "simple": 2}
KERNELS = {"simple": 0, "c64_f32mfma_144x144": 1, "c64_f32mfma_64x64": 2,
           "c128_f64mfma_64x64": 3, "c128_f64mfma_32x32": 4,
           "c64_bf16x3_144x144": 5, "c64_bf16x3_64x64": 6, "f32mfma_128x128": 7,
           "f32mfma_64x64": 8, "f64mfma_64x64": 9, "f64mfma_32x32": 10,
           "c64_f32mfma_128x16": 11, "c128_f64mfma_128x16": 12,
           "c128_f64mfma_64x64w16": 13}

_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)


class TensorDesc(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int32), ("rank", ctypes.c_int32),
                ("dims", _i64p), ("strides", _i64p), ("data", ctypes.c_void_p)]


class Options(ctypes.Structure):
    _fields_ = [("alpha_re", ctypes.c_double), ("alpha_im", ctypes.c_double),
                ("precision", ctypes.c_int32), ("conj_a", ctypes.c_int32),
                ("conj_b", ctypes.c_int32), ("synchronize", ctypes.c_int32),
                ("stream", ctypes.c_void_p), ("autotune", ctypes.c_int32),
                ("deterministic", ctypes.c_int32)]


class View(ctypes.Structure):
    _fields_ = [("br", ctypes.c_int64), ("bc", ctypes.c_int64),
                ("rs", ctypes.c_void_p), ("nrs", ctypes.c_int64),
                ("cs", ctypes.c_void_p), ("ncs", ctypes.c_int64),
                ("rbs", ctypes.c_void_p), ("nrbs", ctypes.c_int64),
                ("cbs", ctypes.c_void_p), ("ncbs", ctypes.c_int64)]


class T1MError(RuntimeError):
    """A failing status from libt1m_hip.so, with its message."""

    def __init__(self, status: int, message: str):
        super().__init__(f"t1m status {status}: {message}")
        self.status = status


# Exported symbols of include/t1m_hip.h.
EXPORTS = ["t1m_options_init", "t1m_contract", "t1m_contract_noalloc", "t1m_contract_vjp",
           "t1m_batch_create", "t1m_batch_launch", "t1m_batch_destroy", "t1m_batch_describe",
           "t1m_index_bundles", "t1m_index_maps", "t1m_pack_block", "t1m_unpack",
           "t1m_last_error", "t1m_version", "t1m_plan_cache_size",
           "t1m_plan_cache_clear", "t1m_autotune_cache_path", "t1m_plan_kernel"]

_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libt1m_hip.so (after torch, so one HIP runtime serves both)."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        import torch  # noqa: F401  (binds libamdhip64.so.7 first)
    except ImportError:
        pass
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with `make -C t1m_amd` "
                      "or __graft_entry__.build()")
    L = ctypes.CDLL(path)
    P = ctypes.POINTER
    L.t1m_options_init.argtypes = [P(Options)]
    L.t1m_options_init.restype = None
    L.t1m_contract.argtypes = [P(TensorDesc), ctypes.c_char_p] * 3 + [P(Options)]
    L.t1m_contract_noalloc.argtypes = [P(TensorDesc), ctypes.c_char_p] * 3 + [P(Options)]
    L.t1m_contract_vjp.argtypes = ([P(TensorDesc), ctypes.c_char_p] * 3 +
                                   [P(TensorDesc), P(TensorDesc), P(Options)])
    L.t1m_batch_create.argtypes = ([ctypes.c_int64] + [P(TensorDesc), ctypes.c_char_p] * 3 +
                                   [P(Options), P(ctypes.c_void_p)])
    L.t1m_batch_launch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.t1m_batch_destroy.argtypes = [ctypes.c_void_p]
    L.t1m_batch_describe.argtypes = [ctypes.c_void_p]
    L.t1m_batch_describe.restype = ctypes.c_char_p
    L.t1m_index_bundles.argtypes = [ctypes.c_char_p] * 3 + [_i64p, _i64p]
    L.t1m_index_maps.argtypes = [P(TensorDesc), _i64p, ctypes.c_int32, ctypes.c_int64,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.t1m_pack_block.argtypes = [ctypes.c_int32, ctypes.c_int32, P(View), ctypes.c_int64,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.t1m_unpack.argtypes = [ctypes.c_int32, P(View), ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_void_p]
    L.t1m_last_error.restype = ctypes.c_char_p
    L.t1m_version.restype = ctypes.c_char_p
    L.t1m_plan_cache_size.restype = ctypes.c_int64
    L.t1m_plan_cache_clear.restype = None
    L.t1m_autotune_cache_path.restype = ctypes.c_char_p
    L.t1m_plan_kernel.argtypes = [P(TensorDesc), ctypes.c_char_p] * 3 + [P(Options)]
    L.t1m_plan_kernel.restype = ctypes.c_char_p
    _lib = L
    return L


def _check(status: int) -> None:
    if status != 0:
        raise T1MError(status, load().t1m_last_error().decode())


def version() -> str:
    return load().t1m_version().decode()


_DTYPE_CODES = None


def _dtype_code(t) -> int:
    global _DTYPE_CODES
    if _DTYPE_CODES is None:
        import torch
        _DTYPE_CODES = {torch.float32: FLOAT32, torch.float64: FLOAT64,
                        torch.complex64: COMPLEX64, torch.complex128: COMPLEX128}
    code = _DTYPE_CODES.get(t.dtype)
    if code is None:
        raise TypeError(f"unsupported dtype {t.dtype}")
    return code


class _Desc:
    """Keeps the ctypes arrays of one tensor descriptor alive."""

    __slots__ = ("dims", "strides", "desc")

    def __init__(self, t, dtype: Optional[int] = None, dims=None, strides=None, ptr=None):
        dims = t.shape if dims is None else dims
        strides = t.stride() if strides is None else strides
        n = len(dims)
        arr = _I64_ARRAYS[n] if n < len(_I64_ARRAYS) else ctypes.c_int64 * max(n, 1)
        self.dims = arr(*dims)
        self.strides = arr(*strides)
        self.desc = TensorDesc(_dtype_code(t) if dtype is None else dtype, n,
                               self.dims, self.strides,
                               t.data_ptr() if ptr is None else ptr)


_I64_ARRAYS = [ctypes.c_int64 * max(n, 1) for n in range(33)]


def _options(alpha=1.0, precision="native", conj_a=False, conj_b=False,
             stream=None, synchronize=True, autotune=False,
             deterministic=False) -> Options:
    # Field by field (the defaults of t1m_options_init), without a library call.
    al = complex(alpha)
    return Options(al.real, al.imag,
                   PRECISIONS[precision] if isinstance(precision, str) else int(precision),
                   int(bool(conj_a)), int(bool(conj_b)), int(bool(synchronize)),
                   _stream_handle(stream), int(bool(autotune)), int(bool(deterministic)))


_RAW_STREAM = None


def _stream_handle(stream) -> Optional[int]:
    global _RAW_STREAM
    if stream is None:
        if _RAW_STREAM is None:
            try:
                import torch
                if torch.cuda.is_available():
                    # torch's raw current-stream query (the public accessor
                    # builds a Python Stream object per call)
                    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
                    _RAW_STREAM = ((lambda: raw(torch.cuda.current_device())) if raw else
                                   (lambda: torch.cuda.current_stream().cuda_stream))
                else:
                    _RAW_STREAM = lambda: None
            except ImportError:
                _RAW_STREAM = lambda: None
        return _RAW_STREAM() or None
    if isinstance(stream, int):
        return stream or None
    return stream.cuda_stream or None


def _operand(t):
    """An input tensor as (tensor over plain memory, conjugated?).

    torch marks some views lazily: a conj view (``x.conj()``, and the
    grad_output autograd hands to a backward through ``C.conj()``) shares x's
    memory with a conjugation bit, and a neg view holds the negated value.
    The library reads raw memory, so a conj bit becomes the call's conj flag
    (its memory is passed as is) and a neg bit is materialised."""
    if t.is_neg():
        t = t.resolve_neg()
    if t.is_conj():
        return t.conj(), True     # clears the bit: same memory, unconjugated
    return t, False


def _output(t, what: str):
    """An output tensor (accumulated in place) must hold its plain value."""
    if t.is_conj() or t.is_neg():
        raise ValueError(f"{what} is a lazy conj/neg view of another tensor; the "
                         "library would add into the unconjugated memory "
                         "(pass a resolved tensor)")
    return t


def contract(a, la: str, b, lb: str, c, lc: str, alpha=1.0, precision="native",
             conj_a: bool = False, conj_b: bool = False, stream=None,
             synchronize: bool = True, autotune: bool = False,
             deterministic: bool = False):
    """C += alpha * A . B over labels (t1m::contract, t1m.h:560-577).

    a, b, c are torch tensors of one dtype (float32/64, complex64/128) with any
    strides; device tensors run without copies.  Lazy conj views of a or b are
    read through the conj flags (no copy); c must not be a conj or neg view.
    autotune=True times the candidate tile shapes the first time a shape is
    seen; deterministic=True forbids splitting K (bitwise-reproducible
    results).  Returns c."""
    a, ca = _operand(a)
    b, cb = _operand(b)
    _output(c, "c")
    da, db, dc = _Desc(a), _Desc(b), _Desc(c)
    o = _options(alpha, precision, bool(conj_a) ^ ca, bool(conj_b) ^ cb, stream,
                 synchronize, autotune, deterministic)
    _check(load().t1m_contract(ctypes.byref(da.desc), la.encode(), ctypes.byref(db.desc),
                               lb.encode(), ctypes.byref(dc.desc), lc.encode(),
                               ctypes.byref(o)))
    return c


def contract_noalloc(a, la: str, b, lb: str, c, lc: str, alpha=1.0, precision="native",
                     conj_a: bool = False, conj_b: bool = False, stream=None,
                     synchronize: bool = True, deterministic: bool = False):
    """contract() without device allocation (t1m_contract_noalloc): device
    tensors only, no plan cache (the index maps are rebuilt per call in a
    static per-GPU array of NOALLOC_SCRATCH_BYTES), no autotuning.  Calls
    that overlap on one GPU race on that array.  Returns c."""
    a, ca = _operand(a)
    b, cb = _operand(b)
    _output(c, "c")
    da, db, dc = _Desc(a), _Desc(b), _Desc(c)
    o = _options(alpha, precision, bool(conj_a) ^ ca, bool(conj_b) ^ cb, stream,
                 synchronize, False, deterministic)
    _check(load().t1m_contract_noalloc(ctypes.byref(da.desc), la.encode(),
                                       ctypes.byref(db.desc), lb.encode(),
                                       ctypes.byref(dc.desc), lc.encode(), ctypes.byref(o)))
    return c


NOALLOC_SCRATCH_BYTES = 8 << 20     # T1M_NOALLOC_SCRATCH_BYTES


def contract_vjp(a, la: str, b, lb: str, grad_c, lc: str, grad_a=None, grad_b=None,
                 alpha=1.0, precision="native", stream=None, synchronize: bool = True):
    """grad_a += conj(alpha) grad_c . conj(B);  grad_b += conj(alpha) conj(A) . grad_c.

    Gradients are with respect to the values of a and b as torch sees them:
    a lazy conj view of a or b is passed as its memory with a conj flag, a
    lazy conj view of grad_c is materialised, and grad_a / grad_b must not be
    conj or neg views."""
    a, ca = _operand(a)
    b, cb = _operand(b)
    if grad_c.is_conj() or grad_c.is_neg():
        grad_c = grad_c.resolve_conj().resolve_neg()
    if grad_a is not None:
        _output(grad_a, "grad_a")
    if grad_b is not None:
        _output(grad_b, "grad_b")
    da, db, dg = _Desc(a), _Desc(b), _Desc(grad_c)
    ga = _Desc(grad_a) if grad_a is not None else None
    gb = _Desc(grad_b) if grad_b is not None else None
    o = _options(alpha, precision, ca, cb, stream, synchronize)
    _check(load().t1m_contract_vjp(
        ctypes.byref(da.desc), la.encode(), ctypes.byref(db.desc), lb.encode(),
        ctypes.byref(dg.desc), lc.encode(),
        ctypes.byref(ga.desc) if ga else None, ctypes.byref(gb.desc) if gb else None,
        ctypes.byref(o)))
    return grad_a, grad_b


class Batch:
    """Many contractions with one label triple, captured in one HIP graph."""

    def __init__(self, a: Sequence, la: str, b: Sequence, lb: str, c: Sequence, lc: str,
                 alpha=1.0, precision="native", conj_a=False, conj_b=False,
                 stream=None, synchronize: bool = True, autotune: bool = False):
        n = len(a)
        if not (len(b) == n and len(c) == n):
            raise ValueError("a, b and c must have the same length")
        a, ca = self._operands(a)
        b, cb = self._operands(b)
        c = [_output(z, "c") for z in c]
        # The captured graph holds these tensors' device addresses: keep the
        # tensors alive (and their memory out of torch's allocator) until
        # close().  C-ABI callers of t1m_batch_create keep the same rule.
        self._tensors = (a, b, c)
        self._keep = [(_Desc(x), _Desc(y), _Desc(z)) for x, y, z in zip(a, b, c)]
        arr = lambda k: (TensorDesc * max(n, 1))(*[d[k].desc for d in self._keep])
        self._a, self._b, self._c = arr(0), arr(1), arr(2)
        o = _options(alpha, precision, bool(conj_a) ^ ca, bool(conj_b) ^ cb, stream,
                     synchronize, autotune)
        h = ctypes.c_void_p()
        _check(load().t1m_batch_create(n, self._a, la.encode(), self._b, lb.encode(),
                                       self._c, lc.encode(), ctypes.byref(o),
                                       ctypes.byref(h)))
        self._h = h

    @staticmethod
    def _operands(ts):
        """Inputs of every instance, and one conj flag for the whole batch:
        when all instances are lazy conj views their memory is read with the
        flag set; when only some are, those are materialised."""
        ops = [_operand(t) for t in ts]
        if ops and all(cj for _, cj in ops):
            return [t for t, _ in ops], True
        return [t.conj().resolve_conj() if cj else t for t, cj in ops], False

    def describe(self) -> str:
        """Tiling mode and persistent launches of the captured graph."""
        return load().t1m_batch_describe(self._h).decode() if getattr(self, "_h", None) else ""

    def launch(self, stream=None) -> None:
        if not getattr(self, "_h", None):
            raise RuntimeError("launch() after close()")
        _check(load().t1m_batch_launch(self._h, _stream_handle(stream)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            load().t1m_batch_destroy(self._h)
            self._h = None
        self._tensors = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def index_bundles(la: str, lb: str, lc: str) -> dict:
    """get_index_bundles (t1m.h:90-136); -1 marks std::string::npos."""
    out = (ctypes.c_int64 * (6 * MAX_RANK))()
    cnt = (ctypes.c_int64 * 6)()
    _check(load().t1m_index_bundles(la.encode(), lb.encode(), lc.encode(), out, cnt))
    names = ["AX", "AY", "BX", "BY", "CX", "CY"]
    return {n: [int(out[i * MAX_RANK + j]) for j in range(cnt[i])] for i, n in enumerate(names)}


def index_maps(dims: Sequence[int], strides: Optional[Sequence[int]], modes: Sequence[int],
               block: int, device="cuda"):
    """Scatter and block-scatter vectors built on the device (t1m.h:141-205).
    Returns two int64 torch tensors holding the uint64 bit patterns."""
    import torch
    n = 1
    for x in modes:
        n *= int(dims[x])
    scat = torch.empty(max(n, 1), dtype=torch.int64, device=device)
    nb = -(-n // block) if block > 0 else 0
    bscat = torch.empty(max(nb, 1), dtype=torch.int64, device=device)
    d = (ctypes.c_int64 * max(len(dims), 1))(*dims)
    s = (ctypes.c_int64 * max(len(dims), 1))(*(strides or [0] * len(dims)))
    desc = TensorDesc(COMPLEX64, len(dims), d, s if strides is not None else None, None)
    m = (ctypes.c_int64 * max(len(modes), 1))(*modes)
    _check(load().t1m_index_maps(ctypes.byref(desc), m, len(modes), block,
                                 scat.data_ptr(), bscat.data_ptr(), _stream_handle(None)))
    torch.cuda.synchronize()
    return scat[:n] if modes else scat, bscat[:nb]


def _view(br, bc, rs, cs, rbs, cbs) -> View:
    return View(br, bc, rs.data_ptr(), rs.numel(), cs.data_ptr(), cs.numel(),
                rbs.data_ptr(), rbs.numel(), cbs.data_ptr(), cbs.numel())


def pack_block(dtype: int, label: str, br: int, bc: int, rs, cs, rbs, cbs, length: int,
               src, dest) -> None:
    """pack_block<T, A|B> on the device (t1m.h:341-404).  rs/cs/rbs/cbs are
    int64 device tensors (uint64 bits); dest is pre-zeroed by the caller."""
    import torch
    v = _view(br, bc, rs, cs, rbs, cbs)
    _check(load().t1m_pack_block(dtype, 0 if label == "A" else 1, ctypes.byref(v), length,
                                 src.data_ptr(), dest.data_ptr(), _stream_handle(None)))
    torch.cuda.synchronize()


def unpack(dtype: int, br: int, bc: int, rs, cs, rbs, cbs, src, dest) -> None:
    """unpack on the device (t1m.h:406-430): dest[idx] += tile."""
    import torch
    v = _view(br, bc, rs, cs, rbs, cbs)
    _check(load().t1m_unpack(dtype, ctypes.byref(v), src.data_ptr(), dest.data_ptr(),
                             _stream_handle(None)))
    torch.cuda.synchronize()


def plan_kernel(a, la: str, b, lb: str, c, lc: str, precision="native",
                autotune: bool = False) -> str:
    """Name of the kernel the plan for this contraction uses."""
    da, db, dc = _Desc(a), _Desc(b), _Desc(c)
    o = _options(precision=precision, autotune=autotune)
    r = load().t1m_plan_kernel(ctypes.byref(da.desc), la.encode(), ctypes.byref(db.desc),
                               lb.encode(), ctypes.byref(dc.desc), lc.encode(),
                               ctypes.byref(o))
    if r is None:
        raise T1MError(-1, load().t1m_last_error().decode())
    return r.decode()


def autotune_cache_path() -> str:
    return load().t1m_autotune_cache_path().decode()


def plan_cache_size() -> int:
    return int(load().t1m_plan_cache_size())


def plan_cache_clear() -> None:
    load().t1m_plan_cache_clear()


from .autograd import ContractFunction, contract_autograd  # noqa: E402,F401
