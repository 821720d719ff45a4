"""The reference's benchmark suite on the MI355X.

Runs the 24 contractions of the reference's benchmark (benchmark/benchmark.cpp:
87-187) for float, double, complex<float> and complex<double>
(benchmark.cpp:189-194) with its method: C zeroed before each of 10 repeats,
the whole contract call timed, minimum in microseconds (benchmark.cpp:7,
64-79).  Here the tensors are device-resident and the call returns after the
result is visible.  Prints the reference's CSV (`labels,min`) with extra
columns: TFLOP/s, the kernel used, and the first (cold) call, which also
builds the plan (the reference builds bundles, scatter vectors and its
workspace inside every call, benchmark.cpp:68-75).

    python -m t1m_amd.suite [--dtypes f32,f64,c64,c128] [--repeats 10] [--autotune]
"""
from __future__ import annotations

import argparse
import sys
import time

# (dims_a, labels_a, dims_b, labels_b, dims_c, labels_c), benchmark.cpp:90-181
SUITE = [
    ((24, 16, 16, 24, 16), "efbad", (12, 16), "cf", (24, 16, 12, 16, 24), "abcde"),
    ((24, 16, 16, 24, 16), "efcad", (12, 16), "bf", (24, 12, 16, 16, 24), "abcde"),
    ((36, 36, 36, 36), "dbea", (36, 12), "ec", (36, 36, 12, 36), "abcd"),
    ((24, 16, 16, 24, 24), "ecbfa", (24, 12), "fd", (24, 16, 16, 12, 24), "abcde"),
    ((36, 36, 36, 36), "deca", (12, 36), "be", (36, 12, 36, 36), "abcd"),
    ((156, 156, 156), "bda", (156, 12), "dc", (156, 156, 12), "abc"),
    ((36, 36, 36, 36), "ebad", (12, 36), "ce", (36, 36, 12, 36), "abcd"),
    ((12, 8, 12, 12), "dega", (12, 8, 8, 8), "gfbc", (12, 8, 8, 12, 8, 8), "abcdef"),
    ((12, 8, 12, 8), "dfgb", (12, 8, 12, 8), "geac", (12, 8, 8, 12, 8, 8), "abcdef"),
    ((12, 8, 12, 8), "degb", (12, 8, 12, 8), "gfac", (12, 8, 8, 12, 8, 8), "abcdef"),
    ((12, 8, 12, 8), "degc", (12, 8, 12, 8), "gfab", (12, 8, 8, 12, 8, 8), "abcdef"),
    ((156, 148, 156), "dca", (12, 156), "bd", (156, 12, 148), "abc"),
    ((36, 36), "ea", (36, 36, 36, 36), "ebcd", (36, 36, 36, 36), "abcd"),
    ((36, 36), "eb", (36, 36, 36, 36), "aecd", (36, 36, 36, 36), "abcd"),
    ((36, 36), "ec", (36, 36, 36, 36), "abed", (36, 36, 36, 36), "abcd"),
    ((36, 36, 36, 36), "adec", (36, 36, 36), "ebd", (36, 36, 36), "abc"),
    ((156, 156, 156), "cad", (156, 156, 148), "dcb", (156, 148), "ab"),
    ((156, 148, 156), "acd", (156, 148, 148), "dbc", (156, 148), "ab"),
    ((156, 148, 148), "acd", (148, 156), "db", (156, 156, 148), "abc"),
    ((156, 148, 148), "adc", (156, 148), "bd", (156, 156, 148), "abc"),
    ((2568, 2568), "ac", (2568, 2560), "cb", (2568, 2560), "ab"),
    ((36, 36, 36, 36), "aebf", (36, 36, 36, 36), "fdec", (36, 36, 36, 36), "abcd"),
    ((36, 36, 36, 36), "eafd", (36, 36, 36, 36), "fbec", (36, 36, 36, 36), "abcd"),
    ((36, 36, 36, 36), "aebf", (36, 36, 36, 36), "dfce", (36, 36, 36, 36), "abcd"),
]


def flops(dims_a, la, lb, dims_c, lc, complex_):
    ext = dict(zip(la, dims_a))
    ext.update(zip(lc, dims_c))
    m = n = k = 1
    for ch in set(la) | set(lb):
        if ch in la and ch in lb:
            k *= ext[ch]
        elif ch in la:
            m *= ext[ch]
        else:
            n *= ext[ch]
    return (8 if complex_ else 2) * m * n * k


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--dtypes", default="f32,f64,c64,c128")
    ap.add_argument("--repeats", type=int, default=10)
    ap.add_argument("--autotune", action="store_true")
    args = ap.parse_args(argv)

    import torch
    import t1m_amd as t1m

    dts = {"f32": torch.float32, "f64": torch.float64, "c64": torch.complex64,
           "c128": torch.complex128}
    for name in args.dtypes.split(","):
        dt = dts[name]
        print(f"# {name}")
        print("contraction,min,tflops,kernel,first_call", flush=True)
        for dims_a, la, dims_b, lb, dims_c, lc in SUITE:
            def colmajor(shape):
                t = torch.rand(list(reversed(shape)), dtype=dt, device="cuda")
                return t.permute(*reversed(range(len(shape))))
            a, b, c = colmajor(dims_a), colmajor(dims_b), colmajor(dims_c)
            best = first = float("inf")
            for r in range(args.repeats):
                c.zero_()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                t1m.contract(a, la, b, lb, c, lc, autotune=args.autotune)
                dt_s = time.perf_counter() - t0
                if r == 0:
                    first = dt_s
                best = min(best, dt_s)
            us = best * 1e6
            fl = flops(dims_a, la, lb, dims_c, lc, dt.is_complex)
            kern = t1m.plan_kernel(a, la, b, lb, c, lc)
            print(f"{la}-{lb}-{lc},{us:.0f},{fl / best / 1e12:.2f},{kern},{first * 1e6:.0f}",
                  flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
