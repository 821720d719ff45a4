"""Benchmark: complex tensor contraction on one MI355X.

Default workload (the headline): complex<float>
A[48,48,48,48] 'abcd' x B[48,48,48,48] 'dbef' -> C[48,48,48,48] 'acef'
(m = n = k = 2304, col-major, no operand GEMM-contiguous), re/im i.i.d.
U[-1, 1], through t1m_contract on device-resident tensors with the plan
(index maps) cached.  One step = one contraction call (C += A.B) on the
FP32-MFMA path.  The same inputs then run on the BF16x3 path, and both
paths' relative Frobenius errors are measured on log-uniform-magnitude inputs
against a complex128 reference.

Timing: `value` is flop x steps over the host wall time of `steps` calls
issued back to back between two synchronizations; `roofline.kernel_ms` is
the GPU time of that same region (HIP events on the launch stream before
the first and after the last call) per launch, and `roofline.achieved` the
flops over it; bracketed_ms_median / _min come from a pass just before, each
call bracketed by its own events (see timed_region): single launches with the
GPU idle between them, so they can exceed kernel_ms.

The same line carries the other timed configurations as sub-objects: "c3"
(complex<double> asb,btc->astc 256x4x256 on FP64 MFMA) and "c5" (512 MPS
transfer contractions in one persistent graph launch, plus their VJP),
each with its value, roofline, errors and CPU baseline, timed over 10x and
1/4 of the headline's steps (each sub-object reports its own).

Prints one JSON line.  --gpus N with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp32": 157.3, "fp64": 78.6, "bf16": 2516.6}   # MI355X dense TFLOP/s
METRIC = "complex contraction TFLOP/s, % MFMA peak; BF16x3 vs FP32 at rel err <=1e-5"

CONFIGS = {
    "c2": dict(la="abcd", lb="dbef", lc="acef", dtype="complex64",
               dims={ch: 48 for ch in "abcdef"}, peak="fp32",
               workload="complex<float> abcd,dbef->acef 48^4 (m=n=k=2304), FP32 MFMA 1M"),
    "c3": dict(la="asb", lb="btc", lc="astc", dtype="complex128",
               dims={"a": 256, "s": 4, "b": 256, "t": 4, "c": 256}, peak="fp64",
               workload="complex<double> asb,btc->astc 256x4x256 (m=n=1024,k=256), FP64 MFMA 1M"),
}


def flops(cfg):
    m = n = k = 1
    for ch, d in cfg["dims"].items():
        if ch in cfg["la"] and ch in cfg["lb"]:
            k *= d
        elif ch in cfg["la"]:
            m *= d
        else:
            n *= d
    return 8 * m * n * k, (m, n, k)


def make_inputs(cfg, kind, seed, torch):
    """Col-major device tensors (strides (1, d0, d0*d1, ...))."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    dt = getattr(torch, cfg["dtype"])
    rdt = torch.float32 if dt == torch.complex64 else torch.float64

    def one(labels):
        shape = [cfg["dims"][ch] for ch in labels]
        rev = list(reversed(shape))
        if kind == "uniform":
            re = torch.rand(rev, generator=g, device="cuda", dtype=rdt) * 2 - 1
            im = torch.rand(rev, generator=g, device="cuda", dtype=rdt) * 2 - 1
            z = torch.complex(re, im)
        elif kind == "normal":
            scale = 1.0 / (cfg["dims"].get("b", 16) ** 0.5)
            z = torch.complex(torch.randn(rev, generator=g, device="cuda", dtype=rdt),
                              torch.randn(rev, generator=g, device="cuda", dtype=rdt)) * scale
        else:  # log-uniform magnitude over [1e-3, 1e3], uniform phase
            r = 10.0 ** (torch.rand(rev, generator=g, device="cuda", dtype=rdt) * 6 - 3)
            th = torch.rand(rev, generator=g, device="cuda", dtype=rdt) * (2 * torch.pi)
            z = torch.polar(r, th)
        z = z.to(dt).contiguous()
        return z.permute(*reversed(range(len(shape))))

    return one(cfg["la"]), one(cfg["lb"]), one(cfg["lc"]) * 0


def timed_region(torch, stream, fn, steps, warmup):
    """`warmup` untimed calls, then two passes of `steps` calls each:
    (1) every call bracketed by HIP events on the launch stream, giving each
    launch's duration on its own (median / min, and warmer clocks); (2) the
    timed region proper -- synchronize, host clock and a HIP event on the
    launch stream, `steps` calls back to back, an event, synchronize.  The
    two events of (2) give the GPU time of the region, so its average launch
    duration (`kernel_ms`, the roofline's denominator) comes from the same
    launches as the wall time and never exceeds it.  Returns (wall seconds,
    region kernel ms per launch, per-launch ms of pass 1)."""
    for _ in range(max(warmup, 1)):
        fn()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    for s_, e_ in evs:
        s_.record(stream)
        fn()
        e_.record(stream)
    torch.cuda.synchronize()
    r0, r1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    r0.record(stream)
    for _ in range(steps):
        fn()
    r1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, r0.elapsed_time(r1) / steps, [s_.elapsed_time(e_) for s_, e_ in evs]


def cold_call_ms(t1m, torch, a, b, c, cfg, precision):
    """One synchronous call with an empty plan cache: label parsing, the
    device index maps (one launch and a wait) and the contraction -- what
    the reference pays on every call (benchmark.cpp:68-75 times bundles,
    scatter vectors and workspace allocation inside each repeat)."""
    t1m.plan_cache_clear()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t1m.contract(a, cfg["la"], b, cfg["lb"], c, cfg["lc"], precision=precision)
    return (time.perf_counter() - t0) * 1e3


def host_staged(t1m, torch, a, b, c, cfg, fl, calls=3):
    """The same contraction on host (CPU) tensors: each call copies A, B and
    C to the device, contracts, copies C back and waits (PCIe-inclusive; a
    convenience path, not the metric).  Best of `calls` after one warm call."""
    ah, bh, ch = a.cpu(), b.cpu(), c.cpu()
    t1m.contract(ah, cfg["la"], bh, cfg["lb"], ch, cfg["lc"])
    best = 1e30
    for _ in range(calls):
        t0 = time.perf_counter()
        t1m.contract(ah, cfg["la"], bh, cfg["lb"], ch, cfg["lc"])
        best = min(best, time.perf_counter() - t0)
    es = a.element_size()
    moved = (a.numel() + b.numel() + 2 * c.numel()) * es
    return {"ms_per_call": round(best * 1e3, 3), "value": round(fl / best / 1e12, 3),
            "unit": "TFLOP/s", "pcie_bytes_per_call": moved,
            "pcie_GB_per_s": round(moved / best / 1e9, 2)}


def time_path(t1m, torch, a, b, c, cfg, precision, steps, warmup):
    stream = torch.cuda.current_stream()
    cold = cold_call_ms(t1m, torch, a, b, c, cfg, precision)
    # the next call autotunes the tile shape
    t1m.contract(a, cfg["la"], b, cfg["lb"], c, cfg["lc"], precision=precision,
                 synchronize=False, autotune=True)
    wall, kern_ms, per_launch = timed_region(
        torch, stream,
        lambda: t1m.contract(a, cfg["la"], b, cfg["lb"], c, cfg["lc"], precision=precision,
                             synchronize=False), steps, warmup)
    kernel = t1m.plan_kernel(a, cfg["la"], b, cfg["lb"], c, cfg["lc"], precision=precision)
    return wall, kern_ms, per_launch, kernel, cold


def rel_errors(t1m, torch, cfg, seed):
    """Relative Frobenius error against a complex128 product.  For complex128
    data the reference runs on the CPU: the vendor's FP64 GEMM on the GPU
    can sum in the same order as the kernel and agree bit for bit, which
    measures nothing."""
    a, b, _ = make_inputs(cfg, "logmag", seed, torch)
    on = "cpu" if cfg["dtype"] == "complex128" else a.device
    ref = torch.einsum(f"{cfg['la']},{cfg['lb']}->{cfg['lc']}",
                       a.to(device=on, dtype=torch.complex128),
                       b.to(device=on, dtype=torch.complex128)).to(a.device)
    out = {}
    precs = ["native", "bf16x3"] if cfg["dtype"] == "complex64" else ["native"]
    for p in precs:
        c = torch.zeros_like(ref, dtype=a.dtype)
        t1m.contract(a, cfg["la"], b, cfg["lb"], c, cfg["lc"], precision=p)
        out[p] = (torch.linalg.vector_norm(c.to(torch.complex128) - ref) /
                  torch.linalg.vector_norm(ref)).item()
    return out


def cpu_threads():
    """Host threads for the all-cores baseline: OMP_NUM_THREADS if set, else
    the CPUs this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


CPU_KERNEL_NOTE = ("BLIS-sized blocks; a plain C micro-kernel built with -ffp-contract=off, "
                   "so no FMA and no BLIS assembly micro-kernel: the reference's BLIS "
                   "haswell kernel would be several times faster per core")


def cpu_baseline(cfg, seconds_hint=15.0, threads=1):
    """The oracle's five-loop 1M restatement (one thread: the reference is
    serial; or loop 3 over `threads` OpenMP threads) on a bounded slice of
    the same workload: B's and C's last free mode cut to `cut` values, which
    scales the flops linearly (8 m n' k)."""
    import numpy as np
    from oracle import oracle as orc
    rng = np.random.default_rng(7)
    dt = np.complex64 if cfg["dtype"] == "complex64" else np.complex128
    lastb = [ch for ch in cfg["lc"] if ch not in cfg["la"]][-1]
    dims = dict(cfg["dims"])

    def arr(labels, d):
        shape = [d[ch] for ch in labels]
        x = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
        return np.asfortranarray(x.astype(dt))

    full, _ = flops(cfg)
    cut = 1
    while True:
        d = dict(dims)
        d[lastb] = cut
        A, B, C = arr(cfg["la"], d), arr(cfg["lb"], d), arr(cfg["lc"], d) * 0
        t0 = time.perf_counter()
        orc.contract(A, cfg["la"], B, cfg["lb"], C, cfg["lc"], threads=threads)
        dt_s = time.perf_counter() - t0
        fl = full * cut / dims[lastb]
        if dt_s * 2 > seconds_hint or cut * 2 > dims[lastb]:
            break
        cut *= 2
    return {"value": round(fl / dt_s / 1e12, 6), "unit": "TFLOP/s", "cores": threads,
            "kind": "port",
            "sample": (f"oracle five-loop 1M contraction ({CPU_KERNEL_NOTE}), "
                       f"{cfg['la']},{cfg['lb']}->{cfg['lc']} with "
                       f"'{lastb}' cut to {cut} of {dims[lastb]} ({fl:.3e} flop, "
                       f"{dt_s:.2f} s, {threads} thread{'s' if threads > 1 else ''})")}


def _ttgt_views(cfg, a, b):
    """Permutations and shapes that turn A and B into GEMM operands
    (A: m-modes x k-modes, B: k-modes x n-modes, torch row-major)."""
    la, lb, lc = cfg["la"], cfg["lb"], cfg["lc"]
    mk = [ch for ch in lc if ch in la]                 # free A modes, C order
    nk = [ch for ch in lc if ch in lb]                 # free B modes, C order
    kk = [ch for ch in la if ch in lb]                 # contracted
    pa = [la.index(ch) for ch in mk + kk]
    pb = [lb.index(ch) for ch in kk + nk]
    d = cfg["dims"]
    M = int(math.prod(d[ch] for ch in mk)); N = int(math.prod(d[ch] for ch in nk))
    K = int(math.prod(d[ch] for ch in kk))
    pc = [(mk + nk).index(ch) for ch in lc]             # (m.., n..) -> C's order
    return pa, pb, pc, mk, nk, M, N, K


def library_gemm(cfg, torch, a, b, c, steps):
    """Ceiling comparison, never in the product path: the vendor GEMM
    (torch.matmul -> hipBLASLt / rocBLAS) on operands already permuted to GEMM
    layout, and the full transpose-transpose-GEMM-transpose (TTGT) route that
    a transposing library takes for the same contraction."""
    pa, pb, pc, mk, nk, M, N, K = _ttgt_views(cfg, a, b)
    d = cfg["dims"]
    fl = 8 * M * N * K

    def ttgt():
        at = a.permute(*pa).reshape(M, K)
        bt = b.permute(*pb).reshape(K, N)
        ct = torch.matmul(at, bt).reshape([d[ch] for ch in mk + nk])
        c.add_(ct.permute(*pc))

    at = a.permute(*pa).contiguous().reshape(M, K)
    bt = b.permute(*pb).contiguous().reshape(K, N)
    out = torch.empty(M, N, dtype=a.dtype, device=a.device)

    def timed(fn):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    t_gemm = timed(lambda: torch.matmul(at, bt, out=out))
    t_ttgt = timed(ttgt)
    return {"gemm_only": round(fl / t_gemm / 1e12, 3), "ttgt": round(fl / t_ttgt / 1e12, 3),
            "unit": "TFLOP/s",
            "note": "torch.matmul on pre-permuted contiguous operands (gemm_only), and "
                    "permute + matmul + permute-add into C (ttgt); comparison only"}


def hiptensor_contract(cfg, torch, t1m, a, b, c_like, steps):
    """Ceiling comparison, never in the product path: AMD's tensor-contraction
    library (libhiptensor from ROCm) on the same strided device operands,
    C += A.B (alpha = beta = 1, D = C), default algorithm, timed with HIP
    events over `steps` calls.  Its result on a zeroed C is compared with
    this library's, so the two compute the same contraction."""
    import ctypes
    try:
        lib = ctypes.CDLL("libhiptensor.so")
    except OSError:
        try:
            lib = ctypes.CDLL("/opt/rocm/lib/libhiptensor.so")
        except OSError as e:
            return {"status": f"unavailable ({e})"}
    vp, i32, u32, u64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint64
    P = ctypes.POINTER
    lib.hiptensorCreate.argtypes = [P(vp)]
    lib.hiptensorCreateTensorDescriptor.argtypes = [vp, P(vp), u32, P(ctypes.c_int64),
                                                    P(ctypes.c_int64), i32, u32]
    lib.hiptensorCreateContraction.argtypes = [vp, P(vp), vp, P(i32), i32, vp, P(i32), i32,
                                               vp, P(i32), i32, vp, P(i32), u32]
    lib.hiptensorCreatePlanPreference.argtypes = [vp, P(vp), i32, i32]
    lib.hiptensorEstimateWorkspaceSize.argtypes = [vp, vp, vp, i32, P(u64)]
    lib.hiptensorCreatePlan.argtypes = [vp, P(vp), vp, vp, u64]
    lib.hiptensorContract.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, u64, vp]
    for f in ("hiptensorDestroyPlan", "hiptensorDestroyPlanPreference",
              "hiptensorDestroyOperationDescriptor", "hiptensorDestroyTensorDescriptor",
              "hiptensorDestroy"):
        getattr(lib, f).argtypes = [vp]
    cplx64 = cfg["dtype"] == "complex64"
    dtype_code = 4 if cplx64 else 5                       # HIPTENSOR_C_32F / C_64F
    computes = [1 << 11, 1 << 2] if cplx64 else [1 << 12, 1 << 4]   # C32F, 32F / C64F, 64F
    made = []

    def ck(st, what):
        if st != 0:
            raise RuntimeError(f"{what}: hiptensor status {st}")

    handle = vp()
    ck(lib.hiptensorCreate(ctypes.byref(handle)), "create")
    try:
        def desc(t):
            n = t.dim()
            lens = (ctypes.c_int64 * max(n, 1))(*t.shape)
            strides = (ctypes.c_int64 * max(n, 1))(*t.stride())
            d = vp()
            ck(lib.hiptensorCreateTensorDescriptor(handle, ctypes.byref(d), n, lens, strides,
                                                   dtype_code, 16), "tensor descriptor")
            made.append(("hiptensorDestroyTensorDescriptor", d))
            return d

        def modes(labels):
            return (i32 * max(len(labels), 1))(*[ord(ch) for ch in labels])

        c = torch.zeros_like(c_like)
        da, db, dc = desc(a), desc(b), desc(c)
        ma, mb, mc = modes(cfg["la"]), modes(cfg["lb"]), modes(cfg["lc"])
        op, status = vp(), None
        for comp in computes:
            status = lib.hiptensorCreateContraction(handle, ctypes.byref(op), da, ma, 1, db, mb, 1,
                                                    dc, mc, 1, dc, mc, comp)
            if status == 0:
                break
        ck(status, "contraction descriptor")
        made.append(("hiptensorDestroyOperationDescriptor", op))
        pref = vp()
        ck(lib.hiptensorCreatePlanPreference(handle, ctypes.byref(pref), -1, 0), "plan preference")
        made.append(("hiptensorDestroyPlanPreference", pref))
        ws = u64(0)
        ck(lib.hiptensorEstimateWorkspaceSize(handle, op, pref, 2, ctypes.byref(ws)), "workspace")
        plan = vp()
        ck(lib.hiptensorCreatePlan(handle, ctypes.byref(plan), op, pref, ws.value), "plan")
        made.append(("hiptensorDestroyPlan", plan))
        work = torch.empty(max(int(ws.value), 1), dtype=torch.uint8, device=a.device)
        one = (ctypes.c_float * 2)(1.0, 0.0) if cplx64 else (ctypes.c_double * 2)(1.0, 0.0)
        stream = torch.cuda.current_stream()

        def call():
            ck(lib.hiptensorContract(handle, plan, ctypes.cast(one, vp), a.data_ptr(), b.data_ptr(),
                                     ctypes.cast(one, vp), c.data_ptr(), c.data_ptr(),
                                     work.data_ptr(), ws.value, stream.cuda_stream), "contract")

        call()
        mine = torch.zeros_like(c_like)
        t1m.contract(a, cfg["la"], b, cfg["lb"], mine, cfg["lc"])
        torch.cuda.synchronize()
        rel = float(torch.linalg.vector_norm(c - mine) / torch.linalg.vector_norm(mine))
        for _ in range(2):
            call()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            call()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        fl, _ = flops(cfg)
        return {"value": round(fl / (ms * 1e-3) / 1e12, 3), "unit": "TFLOP/s", "ms": round(ms, 4),
                "workspace_bytes": int(ws.value), "rel_diff_vs_t1m": float(f"{rel:.3e}"),
                "note": "libhiptensor contraction, default algorithm, same operands; comparison only"}
    except RuntimeError as e:
        return {"status": str(e)}
    finally:
        for f, h in reversed(made):
            getattr(lib, f)(h)
        lib.hiptensorDestroy(handle)


def cpu_einsum(cfg, torch):
    """torch.einsum on the host, all threads, the full workload (one call
    after a warm-up): the library route on the CPU."""
    d = cfg["dims"]
    dt = getattr(torch, cfg["dtype"])
    g = torch.Generator().manual_seed(3)

    def one(labels):
        shape = [d[ch] for ch in labels]
        return torch.complex(torch.rand(shape, generator=g), torch.rand(shape, generator=g)).to(dt)

    a, b = one(cfg["la"]), one(cfg["lb"])
    eq = f"{cfg['la']},{cfg['lb']}->{cfg['lc']}"
    torch.einsum(eq, a, b)
    t0 = time.perf_counter()
    torch.einsum(eq, a, b)
    dt_s = time.perf_counter() - t0
    fl, _ = flops(cfg)
    return {"value": round(fl / dt_s / 1e12, 4), "unit": "TFLOP/s",
            "threads": torch.get_num_threads(), "seconds": round(dt_s, 3)}


def traffic_bytes(config, kernel):
    """HBM bytes per launch measured for this kernel on this workload by the
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised in profiles/."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f)["bytes_per_launch"].get(f"{config}|{kernel}")
    except (OSError, ValueError, KeyError):
        return None


def c5_result(args, torch, t1m, cpu_seconds=15.0):
    """512 complex<float> MPS transfer contractions 'asb,btc->astc' (physical
    dimension 4, bond dimensions uniform over {16, 32, ..., 256}), one
    persistent graph launch, then the VJP of L = sum |C|^2 (grad_c = 2C):
    grad_A = grad_c . conj(B), grad_B = conj(A) . grad_c, two more launches."""
    import random
    rnd = random.Random(5)
    g = torch.Generator(device="cuda").manual_seed(5)
    dims = [tuple(rnd.choice(range(16, 257, 16)) for _ in range(3)) for _ in range(512)]

    def cmat(shape):
        rev = list(reversed(shape))
        z = torch.complex(torch.rand(rev, generator=g, device="cuda") * 2 - 1,
                          torch.rand(rev, generator=g, device="cuda") * 2 - 1)
        return z.permute(*reversed(range(len(shape))))

    A = [cmat((dl, 4, db)) for dl, db, dr in dims]
    B = [cmat((db, 4, dr)) for dl, db, dr in dims]
    C = [cmat((dl, 4, 4, dr)) * 0 for dl, db, dr in dims]
    G = [cmat((dl, 4, 4, dr)) for dl, db, dr in dims]
    GA = [torch.zeros_like(a) for a in A]
    GB = [torch.zeros_like(b) for b in B]
    fwd_flop = sum(128 * dl * db * dr for dl, db, dr in dims)
    stream = torch.cuda.current_stream()
    t1m.plan_cache_clear()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fwd = t1m.Batch(A, "asb", B, "btc", C, "astc", stream=stream, synchronize=False)
    build_s = time.perf_counter() - t0
    # The graphs that are timed: tiling modes chosen by the batch autotuner
    # (timed on scratch outputs at creation, outside the timed region).
    t0 = time.perf_counter()
    fwd.close()
    fwd = t1m.Batch(A, "asb", B, "btc", C, "astc", stream=stream, synchronize=False,
                    autotune=True)
    tune_s = time.perf_counter() - t0
    vja = t1m.Batch(G, "astc", B, "btc", GA, "asb", conj_b=True, stream=stream,
                    synchronize=False, autotune=True)
    vjb = t1m.Batch(A, "asb", G, "astc", GB, "btc", conj_a=True, stream=stream,
                    synchronize=False, autotune=True)

    wall_f, kf, per_f = timed_region(torch, stream, lambda: fwd.launch(stream), args.steps,
                                     args.warmup)
    wall_v, kv, _ = timed_region(torch, stream, lambda: (vja.launch(stream), vjb.launch(stream)),
                                 args.steps, args.warmup)
    per_f = sorted(per_f)
    out = {
        "value": round(fwd_flop * args.steps / wall_f / 1e12, 3), "unit": "TFLOP/s",
        "ms_per_step": round(wall_f / args.steps * 1e3, 4), "steps": args.steps,
        "workload": "512 complex<float> MPS transfer contractions asb,btc->astc, phys 4, "
                    "bonds in {16..256} (seed 5), one persistent graph launch",
        "flop_per_step": fwd_flop, "instances": len(dims),
        "cold_build_ms": round(build_s * 1e3, 1),
        "tuned_build_ms": round(tune_s * 1e3, 1),
        "tiling": {"forward": fwd.describe(), "grad_a": vja.describe(), "grad_b": vjb.describe()},
        "roofline": {"bound": "mfma", "achieved": round(fwd_flop / (kf * 1e-3) / 1e12, 3),
                     "peak": PEAK["fp32"], "unit": "TFLOP/s",
                     "frac": round(fwd_flop / (kf * 1e-3) / 1e12 / PEAK["fp32"], 4),
                     "traffic": traffic_bytes("c5", "c64_f32mfma_64x64_batched_fwd"),
                     "kernel_ms": round(kf, 4), "bracketed_ms_median": round(per_f[len(per_f) // 2], 4)},
        "vjp": {"value": round(2 * fwd_flop * args.steps / wall_v / 1e12, 3), "unit": "TFLOP/s",
                "ms_per_step": round(wall_v / args.steps * 1e3, 4), "kernel_ms": round(kv, 4),
                "flop_per_step": 2 * fwd_flop},
    }
    for bt in (fwd, vja, vjb):
        bt.close()
    if not args.no_errors:
        out["rel_err_vs_fp64"] = c5_errors(torch, t1m, A, B, dims, stream)
    if not args.no_cpu:
        out["cpu_baseline"] = c5_cpu_baseline(dims, seconds_hint=cpu_seconds)
    return out


def bench_c5(args, torch, t1m):
    r = c5_result(args, torch, t1m)
    line = {
        "metric": METRIC, "value": r["value"], "unit": "TFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded bond dims, re/im i.i.d. U[-1,1]; device-resident)",
        "config": {"workload": r["workload"], "flop_per_step": r["flop_per_step"],
                   "instances": r["instances"], "parallelism": "single GPU"},
    }
    line.update({k: v for k, v in r.items() if k not in ("value", "unit", "ms_per_step",
                                                          "workload", "flop_per_step",
                                                          "instances")})
    print(json.dumps(line), flush=True)


def c5_errors(torch, t1m, A, B, dims, stream, count=512):
    """Relative Frobenius error of a fresh persistent batch over the first
    `count` instances (all 512 by default) against complex128 products of
    the same operands (permuted to GEMM layout, contiguous: (a s) x b times
    b x (t c))."""
    sel = range(min(count, len(dims)))
    C = [torch.zeros((dims[i][2], 4, 4, dims[i][0]), dtype=torch.complex64,
                     device="cuda").permute(3, 2, 1, 0) for i in sel]
    bt = t1m.Batch([A[i] for i in sel], "asb", [B[i] for i in sel], "btc", C, "astc",
                   stream=stream, synchronize=True)
    bt.launch(stream)
    bt.close()
    num = den = 0.0
    for j, i in enumerate(sel):
        dl, db, dr = dims[i]
        a = A[i].to(torch.complex128).contiguous().reshape(dl * 4, db)
        b = B[i].to(torch.complex128).contiguous().reshape(db, 4 * dr)
        ref = (a @ b).reshape(dl, 4, 4, dr)
        num += float((C[j].to(torch.complex128) - ref).abs().pow(2).sum())
        den += float(ref.abs().pow(2).sum())
    return {"fp32_batched": float(f"{math.sqrt(num / den):.3e}"), "instances": len(sel)}


def c5_cpu_baseline(dims, seconds_hint=15.0):
    """The oracle's five-loop 1M restatement, one thread, over the workload's
    instances in order (same seeded bond dimensions, fresh U[-1,1] operands)
    until about `seconds_hint` seconds of CPU work."""
    import numpy as np
    from oracle import oracle as orc
    rng = np.random.default_rng(7)

    def arr(shape):
        x = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
        return np.asfortranarray(x.astype(np.complex64))

    fl = 0
    spent = 0.0
    n = 0
    for dl, db, dr in dims:
        a, b = arr((dl, 4, db)), arr((db, 4, dr))
        c = np.zeros((dl, 4, 4, dr), dtype=np.complex64, order="F")
        t0 = time.perf_counter()
        orc.contract(a, "asb", b, "btc", c, "astc")
        spent += time.perf_counter() - t0
        fl += 128 * dl * db * dr
        n += 1
        if spent > seconds_hint:
            break
    return {"value": round(fl / spent / 1e12, 6), "unit": "TFLOP/s", "cores": 1, "kind": "port",
            "sample": (f"oracle five-loop 1M contraction ({CPU_KERNEL_NOTE}), the first "
                       f"{n} of {len(dims)} instances asb,btc->astc ({fl:.3e} flop, "
                       f"{spent:.2f} s, 1 thread)")}


def bench_maps(args, torch, t1m):
    """HBM rates of the device index-map and panel kernels at bandwidth-sized
    inputs (algorithmic bytes per launch / mean launch time, HIP events on
    the launch stream):
      k_scatter_maps  scatter vector of all modes of a 64^4 tensor (8 B written per entry)
      k_block_scatter its block-scatter vector, block 6 (8 B read per entry)
      k_pack_block    1M A panels of the 48^4 'abcd' operand, rows (a,c) x cols (b,d),
                      MR = 6, KP = 4 (8 B read + 16 B written per complex element)
      k_unpack        the 48^4 'acef' C tile, rows (a,c) x cols (e,f), leading dimension
                      2 x 2304 (8 B tile read + 8 B C read + 8 B C written per element)"""
    import ctypes
    L = t1m.load()
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    out = {}

    def rate(name, fn, nbytes):
        _, ms, _ = timed_region(torch, stream, fn, args.steps, args.warmup)
        out[name] = {"GB/s": round(nbytes / (ms * 1e-3) / 1e9, 1), "us": round(ms * 1e3, 2),
                     "bytes": int(nbytes), "frac_of_hbm_peak": round(nbytes / (ms * 1e-3) / 8e12, 4)}

    def desc(dims):
        d = (ctypes.c_int64 * len(dims))(*dims)
        return t1m.TensorDesc(t1m.COMPLEX64, len(dims), d, None, None), d

    # scatter / block scatter over 64^4
    big, keep = desc([64] * 4)
    n = 64 ** 4
    modes = (ctypes.c_int64 * 4)(0, 1, 2, 3)
    scat = torch.empty(n, dtype=torch.int64, device="cuda")
    bs = torch.empty(-(-n // 6), dtype=torch.int64, device="cuda")
    rate("k_scatter_maps", lambda: L.t1m_index_maps(ctypes.byref(big), modes, 4, 0,
                                                    scat.data_ptr(), None, sh), 8 * n)
    rate("k_scatter_maps+k_block_scatter",
         lambda: L.t1m_index_maps(ctypes.byref(big), modes, 4, 6, scat.data_ptr(),
                                  bs.data_ptr(), sh), 8 * n + 8 * n + 8 * bs.numel())

    # 1M A panels of the 48^4 operand
    dims = [48] * 4
    rows, cols = (0, 2), (1, 3)
    br, bc = 6, 4
    rs, rbs = t1m.index_maps(dims, None, rows, br)
    cs, cbs = t1m.index_maps(dims, None, cols, bc)
    m = k = 48 * 48
    src = torch.complex(torch.rand(48 ** 4, device="cuda"), torch.rand(48 ** 4, device="cuda"))
    length = 2 * k
    dest = torch.zeros(-(-2 * m // br) * br * length + 64, dtype=torch.float32, device="cuda")
    v = t1m._view(br, bc, rs, cs, rbs, cbs)
    rate("k_pack_block (1M A, complex64)",
         lambda: L.t1m_pack_block(t1m.COMPLEX64, 0, ctypes.byref(v), length, src.data_ptr(),
                                  dest.data_ptr(), sh), 24 * m * k)

    # unpack a 2304 x 2304 complex tile into the 48^4 C
    rs2, rbs2 = t1m.index_maps(dims, None, (0, 1), m)
    cs2, cbs2 = t1m.index_maps(dims, None, (2, 3), k)
    tile = torch.rand(2 * m * k, device="cuda")
    cdst = torch.zeros(48 ** 4, dtype=torch.complex64, device="cuda")
    v2 = t1m._view(2 * m, k, rs2, cs2, rbs2, cbs2)
    rate("k_unpack (complex64)",
         lambda: L.t1m_unpack(t1m.COMPLEX64, ctypes.byref(v2), tile.data_ptr(), cdst.data_ptr(),
                              sh), 24 * m * k)
    torch.cuda.synchronize()
    print(json.dumps({"metric": "index-map / panel kernels, HBM GB/s (algorithmic bytes)",
                      "unit": "GB/s", "peak": 8000, "steps": args.steps, "warmup": args.warmup,
                      "kernels": out}), flush=True)
    del keep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 2000 steps of the 0.8 ms c2 launch: the GPU is busy for seconds (not
    # milliseconds) beside the CPU baselines, so a device sampler sees it.
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS) + ["c5", "maps"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-errors", action="store_true", help="skip the error report")
    ap.add_argument("--no-library", action="store_true",
                    help="skip the vendor GEMM and hipTensor comparisons")
    ap.add_argument("--only", action="store_true",
                    help="with the default c2: skip the c3 and c5 sub-objects")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {args.gpus}); "
              "multi-GPU sharding is design-only.", file=sys.stderr)
        sys.exit(2)

    import torch
    import t1m_amd as t1m

    if args.config == "c5":
        bench_c5(args, torch, t1m)
        return
    if args.config == "maps":
        bench_maps(args, torch, t1m)
        return
    line = config_result(args.config, args, torch, t1m, headline=True)
    if args.config == "c2" and not args.only:
        # the other timed configurations, so one default run reports them all
        # with step counts scaled to similar GPU time per region: c3's
        # launches take 1/16 of c2's, c5's forward graph 2.3 times as long
        line["c3"] = config_result("c3", sub_steps(args, 10), torch, t1m, headline=False)
        line["c5"] = c5_result(sub_steps(args, 0.25), torch, t1m, cpu_seconds=6.0)
    print(json.dumps(line), flush=True)


def sub_steps(args, factor):
    """A copy of args whose step count is scaled by factor (at least 5)."""
    sub = argparse.Namespace(**vars(args))
    sub.steps = max(5, int(round(args.steps * factor)))
    return sub


def config_result(name, args, torch, t1m, headline):
    """One single-contraction configuration: FP32 / FP64 MFMA path timed
    (and BF16x3 for complex<float>), errors against complex128, comparison
    libraries and the CPU baseline.  `headline` gives the full JSON line
    (else a sub-object for another configuration)."""
    cfg = CONFIGS[name]
    fl, (m, n, k) = flops(cfg)
    a, b, c = make_inputs(cfg, "uniform" if cfg["dtype"] == "complex64" else "normal", 2, torch)

    wall, kern_ms, per_launch, kname, cold = time_path(t1m, torch, a, b, c, cfg, "native",
                                                       args.steps, args.warmup)
    ms_step = wall / args.steps * 1e3
    value = fl * args.steps / wall / 1e12
    per_launch = sorted(per_launch)
    achieved = fl / (kern_ms * 1e-3) / 1e12
    peak = PEAK[cfg["peak"]]
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak,
            "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
            "traffic": traffic_bytes(name, kname),
            "kernel_ms": round(kern_ms, 4), "kernel": kname,
            "bracketed_ms_median": round(per_launch[len(per_launch) // 2], 4),
            "bracketed_ms_min": round(per_launch[0], 4)}
    dtype = "fp32" if cfg["dtype"] == "complex64" else "fp64"
    if headline:
        line = {
            "metric": METRIC, "value": round(value, 3), "unit": "TFLOP/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": dtype,
            "data": "synthetic (seeded, re/im i.i.d. U[-1,1]; device-resident)",
            "config": {"workload": cfg["workload"],
                       "labels": f"{cfg['la']},{cfg['lb']}->{cfg['lc']}",
                       "m": m, "n": n, "k": k, "flop_per_step": fl,
                       "parallelism": "single GPU"},
            "roofline": roof,
        }
    else:
        line = {"value": round(value, 3), "unit": "TFLOP/s", "ms_per_step": round(ms_step, 4),
                "steps": args.steps, "dtype": dtype, "workload": cfg["workload"],
                "flop_per_step": fl, "roofline": roof}
    line["cold_call_ms"] = round(cold, 3)
    if headline:
        line["host_staged"] = host_staged(t1m, torch, a, b, c, cfg, fl)
    if cfg["dtype"] == "complex64":
        c2 = torch.zeros_like(c)
        wall_b, kb, _, kname_b, _ = time_path(t1m, torch, a, b, c2, cfg, "bf16x3", args.steps,
                                              args.warmup)
        line["bf16x3"] = {"value": round(fl * args.steps / wall_b / 1e12, 3), "unit": "TFLOP/s",
                          "kernel_ms": round(kb, 4), "kernel": kname_b,
                          "speedup_vs_fp32": round(kern_ms / kb, 3),
                          "frac_of_fp32_peak": round(fl / (kb * 1e-3) / 1e12 / peak, 4),
                          "frac_of_bf16_peak_3_products":
                              round(3 * fl / (kb * 1e-3) / 1e12 / PEAK["bf16"], 4),
                          "traffic": traffic_bytes(name, kname_b)}
    if not args.no_errors:
        errs = rel_errors(t1m, torch, cfg, 4)
        line["rel_err_vs_fp64"] = {p: float(f"{e:.3e}") for p, e in errs.items()}
    if not args.no_library:
        # the vendor GEMM over as many steps as this path; hipTensor (5 ms a call at c2) over
        # at most 200
        line["library_gemm"] = library_gemm(cfg, torch, a, b, torch.zeros_like(c), args.steps)
        line["hiptensor"] = hiptensor_contract(cfg, torch, t1m, a, b, c, min(args.steps, 200))
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(cfg, seconds_hint=10.0 if headline else 5.0)
        nt = cpu_threads()
        if nt > 1 and headline:
            line["cpu_baseline_all_cores"] = cpu_baseline(cfg, seconds_hint=3.0, threads=nt)
        if headline:
            line["cpu_torch_einsum"] = cpu_einsum(cfg, torch)
    return line


if __name__ == "__main__":
    main()
