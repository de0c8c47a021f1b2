"""The built library's kernels: none spills vector registers or uses scratch memory, and the
force kernel the autotuner picks for the 100k workload keeps its occupancy.

A kernel that spills needs a per-wave scratch allocation, which caps how many of its waves
can run (a force kernel held to 72 VGPRs that spilled 3 of them ran 6x slower).  Scalar
register spills go to VGPR lanes and are allowed.  The figures come from the AMDGPU
metadata of the library's gfx950 code object (llvm-objcopy, clang-offload-bundler and
llvm-readelf of the ROCm LLVM)."""

import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "cellflow_amd", "libcellflow.so")
LLVM = "/opt/rocm/llvm/bin"
TOOLS = ("llvm-objcopy", "clang-offload-bundler", "llvm-readelf")

# the autotuner's pick for the 100k workload: variant 4, 27-cell stencil, 8 lanes, tile sort
# (with and without fused edges); 72 VGPRs give 7 waves per SIMD
OCCUPANCY = {"_ZN2cf12k_force_listILb1ELi1ELi8ELb1ELb1EEEvNS_9ForceArgsE": 72,
             "_ZN2cf12k_force_listILb0ELi1ELi8ELb1ELb1EEEvNS_9ForceArgsE": 72}


MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def kernel_metadata(tmp):
    """Kernel name -> resource fields, over every offload bundle of the library (one per
    source file with kernels)."""
    fat = os.path.join(tmp, "fat.bin")
    subprocess.run([os.path.join(LLVM, "llvm-objcopy"), f"--dump-section=.hip_fatbin={fat}",
                    LIB, os.path.join(tmp, "lib.so")], check=True)
    data = open(fat, "rb").read()
    starts = [i for i in range(len(data)) if data.startswith(MAGIC, i)]
    notes = ""
    for n, (b0, b1) in enumerate(zip(starts, starts[1:] + [len(data)])):
        part, co = os.path.join(tmp, f"bundle{n}.bin"), os.path.join(tmp, f"co{n}.o")
        with open(part, "wb") as f:
            f.write(data[b0:b1])
        subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o",
                        "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--input={part}",
                        f"--output={co}"], check=True)
        notes += subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes", co],
                                check=True, capture_output=True, text=True).stdout
    out = {}
    for block in re.split(r"\n\s+- \.", notes):
        m = re.search(r"\.name:\s+(\S+)", block)
        if not m:
            continue

        def field(f):
            v = re.search(r"\." + f + r":\s+(\d+)", block)
            return int(v.group(1)) if v else None

        out[m.group(1)] = {f: field(f) for f in ("vgpr_count", "vgpr_spill_count",
                                                 "private_segment_fixed_size")}
    return out


@pytest.fixture(scope="module")
def meta(tmp_path_factory):
    if not os.path.exists(LIB):
        pytest.skip("libcellflow.so not built")
    if not all(os.path.exists(os.path.join(LLVM, t)) for t in TOOLS):
        pytest.skip("ROCm LLVM tools not found")
    return kernel_metadata(str(tmp_path_factory.mktemp("co")))


def test_no_kernel_spills_or_uses_scratch(meta):
    assert len(meta) > 50, "expected the library's kernels in the code objects"
    for k in ("_ZN2cf13k_shadow_stepENS_7AdjArgsE", "_ZN2cf9k_edges_gENS_8EdgeArgsE"):
        assert k in meta, k  # kernels of each source file are covered
    bad = {k: v for k, v in meta.items()
           if v["vgpr_spill_count"] or v["private_segment_fixed_size"]}
    assert not bad, bad


def test_force_kernel_occupancy(meta):
    for k, limit in OCCUPANCY.items():
        assert k in meta, k
        assert meta[k]["vgpr_count"] <= limit, (k, meta[k])
