"""CPU tests of the C ABI: the library loads, exports every symbol that
include/t1m_hip.h declares, and its host-side logic (bundles, validation,
error reporting) behaves -- no kernel is launched here."""
import ctypes
import os
import random
import re
import string

import pytest
import torch

import t1m_amd
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "t1m_hip.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(t1m_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    lib = t1m_amd.load()
    names = declared_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(t1m_amd.EXPORTS) == names


def test_version():
    assert "gfx950" in t1m_amd.version()


def test_index_bundles_kats(kats):
    for case in kats["bundles"]:
        got = t1m_amd.index_bundles(case["la"], case["lb"], case["lc"])
        for key in ("AX", "AY", "BX", "BY", "CX", "CY"):
            assert got[key] == case[key], (case["source"], key)


def test_index_bundles_match_oracle():
    """Including the reference's multiset behaviour on repeated labels and
    npos for labels missing from C."""
    rng = random.Random(3)
    for _ in range(300):
        la = "".join(rng.choice("abcdefg") for _ in range(rng.randint(0, 6)))
        lb = "".join(rng.choice("abcdefg") for _ in range(rng.randint(0, 6)))
        lc = "".join(rng.sample(string.ascii_lowercase[:8], rng.randint(0, 8)))
        got = t1m_amd.index_bundles(la, lb, lc)
        want = orc.bundles(la, lb, lc)
        for k in want:
            want[k] = [-1 if v == 2 ** 64 - 1 else v for v in want[k]]
        assert got == want, (la, lb, lc)


def _t(shape, dtype=torch.complex64):
    return torch.zeros(shape, dtype=dtype)


@pytest.mark.parametrize("labels,shapes,status", [
    (("aa", "bc", "ac"), ((2, 2), (3, 4), (2, 4)), 1),     # repeated label
    (("ab", "bc", "abc"), ((2, 3), (4, 5), (2, 3, 5)), 1), # batch label b: extents
    (("ab", "cd", "ac"), ((2, 3), (3, 4), (2, 3)), 1),     # b of A nowhere
    (("ab", "bc", "ad"), ((2, 3), (3, 4), (2, 4)), 1),     # d of C nowhere
    (("ab", "bc", "ac"), ((2, 3), (5, 4), (2, 4)), 1),     # contracted extent
    (("ab", "bc", "ac"), ((2, 3), (3, 4), (7, 4)), 1),     # free extent
    (("abc", "bc", "a"), ((2, 3), (3, 4), (2,)), 1),       # rank vs labels
])
def test_validation_errors(labels, shapes, status):
    a, b, c = (_t(s) for s in shapes)
    with pytest.raises(t1m_amd.T1MError) as e:
        t1m_amd.contract(a, labels[0], b, labels[1], c, labels[2])
    assert e.value.status == status
    assert t1m_amd.load().t1m_last_error().decode()


def test_dtype_errors():
    a, b = _t((2, 3)), _t((3, 4))
    c = _t((2, 4), torch.complex128)
    with pytest.raises(t1m_amd.T1MError) as e:
        t1m_amd.contract(a, "ab", b, "bc", c, "ac")
    assert e.value.status == 1
    a2, b2, c2 = (_t(s, torch.complex128) for s in ((2, 3), (3, 4), (2, 4)))
    with pytest.raises(t1m_amd.T1MError) as e:
        t1m_amd.contract(a2, "ab", b2, "bc", c2, "ac", precision="bf16x3")
    assert e.value.status == 4


def test_zero_extent_is_a_no_op_without_device():
    # K = 0: C += 0 returns before any device work.
    a, b, c = _t((2, 0)), _t((0, 4)), _t((2, 4))
    t1m_amd.contract(a, "ab", b, "bc", c, "ac")


def test_bad_options_and_null_arguments():
    lib = t1m_amd.load()
    assert lib.t1m_contract(None, b"a", None, b"a", None, b"", None) == 1
    assert lib.t1m_index_bundles(None, b"", b"", None, None) == 1
    assert lib.t1m_batch_destroy(None) == 0
    assert lib.t1m_batch_launch(None, None) == 1


def test_view_and_extent_validation():
    # Rejected on the host before any device work.
    lib = t1m_amd.load()
    one = ctypes.c_void_p(16)          # never dereferenced: validation fails first
    bad_counts = t1m_amd.View(4, 4, one, -1, one, 2, one, 1, one, 1)
    assert lib.t1m_pack_block(t1m_amd.FLOAT32, 0, ctypes.byref(bad_counts), 8, one, one, None) == 1
    assert b"view" in lib.t1m_last_error()
    missing = t1m_amd.View(4, 4, one, 3, one, 2, None, 1, one, 1)
    assert lib.t1m_unpack(t1m_amd.FLOAT32, ctypes.byref(missing), one, one, None) == 1
    no_blocks = t1m_amd.View(4, 4, one, 3, one, 2, one, 0, one, 1)
    assert lib.t1m_pack_block(t1m_amd.FLOAT32, 1, ctypes.byref(no_blocks), 8, one, one, None) == 1
    bad_ld = t1m_amd.View(0, 4, one, 3, one, 2, one, 1, one, 1)
    assert lib.t1m_unpack(t1m_amd.FLOAT32, ctypes.byref(bad_ld), one, one, None) == 1
    # pack_block reads the block stride of every block: 9 rows at br = 4 need
    # ceil(9/4) = 3 row-block entries, 5 columns at bc = 2 need 3 column ones.
    for dt in (t1m_amd.FLOAT32, t1m_amd.COMPLEX64):
        for lab in (0, 1):
            short_rbs = t1m_amd.View(4, 2, one, 9, one, 5, one, 2, one, 3)
            assert lib.t1m_pack_block(dt, lab, ctypes.byref(short_rbs), 8, one, one, None) == 1
            assert b"view" in lib.t1m_last_error()
            short_cbs = t1m_amd.View(4, 2, one, 9, one, 5, one, 3, one, 2)
            assert lib.t1m_pack_block(dt, lab, ctypes.byref(short_cbs), 8, one, one, None) == 1
    dims = (ctypes.c_int64 * 2)(3, -2)
    t = t1m_amd.TensorDesc(t1m_amd.FLOAT32, 2, dims, None, None)
    modes = (ctypes.c_int64 * 2)(0, 1)
    assert lib.t1m_index_maps(ctypes.byref(t), modes, 2, 4, one, None, None) == 1
    assert b"negative extent" in lib.t1m_last_error()


def test_product_code_never_uses_the_oracle():
    # The oracle is test infrastructure: the package, its native sources and
    # the public headers neither import, link nor load it, so no product path
    # can fall back to the CPU restatement.
    pattern = re.compile(r"oracle", re.I)
    roots = [os.path.join(ROOT, "t1m_amd"), os.path.join(ROOT, "include")]
    hits = []
    for root in roots:
        for dirpath, _, files in os.walk(root):
            if "_build" in dirpath or "__pycache__" in dirpath:
                continue
            for f in files:
                if f.endswith((".py", ".hip", ".h", ".cpp", ".hpp")) or f == "Makefile":
                    with open(os.path.join(dirpath, f), errors="replace") as fh:
                        if pattern.search(fh.read()):
                            hits.append(os.path.join(dirpath, f))
    assert not hits, hits
    # bench.py touches it only inside its CPU-baseline functions
    with open(os.path.join(ROOT, "bench.py")) as fh:
        src = fh.read()
    for m in re.finditer(r"from oracle import", src):
        head = src[:m.start()]
        fn = head[head.rfind("\ndef "):].split("(")[0].strip()
        assert fn in ("def cpu_baseline", "def c5_cpu_baseline"), fn


def test_missing_library_fails_loudly():
    # No fallback: without the built HIP library the package refuses to run.
    code = ("import t1m_amd, sys\n"
            "try:\n"
            "    t1m_amd.load('/nonexistent/libt1m_hip.so')\n"
            "except OSError as e:\n"
            "    print('refused:', e); sys.exit(0)\n"
            "sys.exit(1)\n")
    import subprocess
    import sys
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and "is missing" in r.stdout, (r.stdout, r.stderr)
