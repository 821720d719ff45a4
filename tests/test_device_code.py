"""Properties of the built gfx950 device code, read from the library's own
code objects (this repository's build, unbundled from t1m_amd/libt1m_hip.so):

* no flat memory operation anywhere: every buffer a kernel reaches through a
  pointer is global memory, and a flat load or store counts in lgkmcnt as
  well as vmcnt, so each wait for an LDS operand would also wait for the
  global loads in flight (the persistent batch kernels, whose per-instance
  pointers come from device memory, lost 2-3 % that way);
* no register spills or scratch in the kernels the headline workloads run
  (2304^3 complex<float> on the 144 x 144 FP32 tiles and the BF16x3 tiles,
  the FP64 16-wave 64 x 64 tiles, the persistent FP32 64 x 64 batch tiles),
  but for the bounded spills listed below."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "t1m_amd", "libt1m_hip.so")
LLVM = "/opt/rocm/llvm/bin"
OBJCOPY, BUNDLER, OBJDUMP, READELF = (os.path.join(LLVM, t) for t in (
    "llvm-objcopy", "clang-offload-bundler", "llvm-objdump", "llvm-readelf"))
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"

# Mangled-name prefixes of the headline kernels (every gather-direction and
# pair variant of each).
HEADLINE = {
    "FP32 144x144, 12 waves": "_ZN7t1m_hip12_GLOBAL__N_113k_contract_1mIfiLi144ELi144ELi16ELi3ELi4E",
    "FP32 144x144, 64-bit offsets": "_ZN7t1m_hip12_GLOBAL__N_113k_contract_1mIflLi144ELi144ELi16ELi2ELi4E",
    "FP64 64x64 16 waves": "_ZN7t1m_hip12_GLOBAL__N_113k_contract_1mIdiLi64ELi64ELi16ELi8ELi2E",
    "FP32 64x64 batch": "_ZN7t1m_hip12_GLOBAL__N_121k_contract_1m_batchedIfiLi64ELi64ELi8ELi2ELi2E",
    "BF16x3 144x144 staging waves": "_ZN7t1m_hip12_GLOBAL__N_125k_contract_bf16x3_pair_ws",
}


@pytest.fixture(scope="module")
def code_objects(tmp_path_factory):
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    missing = [t for t in (OBJCOPY, BUNDLER, OBJDUMP, READELF) if not os.path.exists(t)]
    if missing:
        pytest.skip(f"ROCm LLVM tools missing: {missing}")
    d = tmp_path_factory.mktemp("code_objects")
    fat = d / "fatbin"
    subprocess.run([OBJCOPY, "--dump-section", f".hip_fatbin={fat}", LIB, str(d / "copy.so")],
                   check=True, capture_output=True)
    data = fat.read_bytes()
    # One offload bundle per translation unit, concatenated.
    magic = b"__CLANG_OFFLOAD_BUNDLE__"
    starts = [m.start() for m in re.finditer(re.escape(magic), data)]
    assert starts, "no offload bundle in .hip_fatbin"
    objs = []
    for i, s in enumerate(starts):
        end = starts[i + 1] if i + 1 < len(starts) else len(data)
        b = d / f"bundle{i}"
        b.write_bytes(data[s:end])
        co = d / f"gfx950_{i}.o"
        subprocess.run([BUNDLER, "--unbundle", "--type=o", f"--targets={TARGET}",
                        f"--input={b}", f"--output={co}"], check=True, capture_output=True)
        if co.stat().st_size:
            objs.append(co)
    return objs


def test_no_flat_memory_operations(code_objects):
    bad = []
    kernels = 0
    for co in code_objects:
        text = subprocess.run([OBJDUMP, "-d", str(co)], check=True, capture_output=True,
                              text=True).stdout
        current = None
        for line in text.splitlines():
            m = re.match(r"^[0-9a-f]+ <(.+)>:$", line)
            if m:
                current = m.group(1)
                kernels += 1
                continue
            if re.search(r"\bflat_(load|store|atomic)", line):
                bad.append((current, line.strip()))
    assert kernels > 100, kernels
    assert not bad, f"{len(bad)} flat memory operations, first: {bad[:3]}"


def kernel_metadata(co):
    """(name, vgpr_spill_count, private_segment_fixed_size) per kernel of a
    code object, from its AMDGPU metadata note."""
    text = subprocess.run([READELF, "--notes", str(co)], check=True, capture_output=True,
                          text=True).stdout
    out, cur = [], None
    for line in text.splitlines():
        m = re.match(r"^\s*\.name:\s+(\S+)$", line)
        if m:
            cur = {"name": m.group(1)}
            out.append(cur)
            continue
        m = re.match(r"^\s*\.(vgpr_spill_count|private_segment_fixed_size):\s+(\d+)$", line)
        if m and cur is not None:
            cur[m.group(1)] = int(m.group(2))
    return out


# The allowed exceptions: the 12-wave 144 x 144 tiles with two-element
# gathers keep the two-stage prefetch although it makes them spill 2-12
# registers, because they are faster with it (48^4 abcd,dbef / bdac,dbef /
# abcd,ebdf / bdac,ebdf: 795 / 786 / 795 / 803 -> 779 / 775 / 770 / 779 us).
# (mangled-name prefix: most spilled registers, most scratch bytes)
_W12 = "_ZN7t1m_hip12_GLOBAL__N_113k_contract_1mIfiLi144ELi144ELi16ELi3ELi4E"
ALLOWED_SPILL = {
    _W12 + "Lb1ELb1ELb1E": (5, 24),
    _W12 + "Lb1ELb0ELb1E": (12, 52),
    _W12 + "Lb0ELb1ELb1E": (11, 48),
    _W12 + "Lb0ELb0ELb1E": (2, 12),
}


def test_headline_kernels_do_not_spill(code_objects):
    meta = [k for co in code_objects for k in kernel_metadata(co)]
    for what, prefix in HEADLINE.items():
        ks = [k for k in meta if k["name"].startswith(prefix)]
        assert ks, f"{what}: no kernel named {prefix}*"
        for k in ks:
            spill, scratch = next((v for p, v in ALLOWED_SPILL.items()
                                   if k["name"].startswith(p)), (0, 0))
            assert k.get("vgpr_spill_count", 0) <= spill, (what, k)
            assert k.get("private_segment_fixed_size", 0) <= scratch, (what, k)
