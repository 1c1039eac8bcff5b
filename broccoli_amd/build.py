"""Builds broccoli_amd/libbpg.so (the C-ABI library) for gfx950 with hipcc.

Each source in csrc/ is compiled to an object under broccoli_amd/_build/ (in
parallel), then linked into one shared library.  Every call compiles again.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libbpg.so"

ARCH = "gfx950"
# Packed FP32 (v_pk_*_f32) is switched off for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          f"-I{ROOT / 'include'}", f"-I{CSRC}"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build broccoli_amd)")


# per-kernel register use of the device code, as the compiler reports it
# (name, VGPRs, AGPRs, scratch bytes per lane, spilled VGPRs): the tests check
# that no contraction kernel spills
RESOURCES = BUILD / "kernel_resources.tsv"


def _resources(stderr: str) -> list[tuple[str, int, int, int, int]]:
    import re
    rows = []
    for part in stderr.split("Function Name: ")[1:]:
        name = part.split()[0]

        def get(key):
            m = re.search(key + r": (\d+)", part)
            return int(m.group(1)) if m else -1
        rows.append((name, get("VGPRs"), get("AGPRs"), get(r"ScratchSize \[bytes/lane\]"), get("VGPRs Spill")))
    return rows


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    extra = []
    if src.name.startswith("bpg_contract_bf16"):
        # keep MFMA accumulators in architectural VGPRs: the epilogue reads
        # them directly instead of through v_accvgpr_read
        extra = ["-mllvm", "-amdgpu-mfma-vgpr-form"]
    cmd = [hipcc(), f"--offload-arch={ARCH}", *COMMON, *DEVICE_FLAGS, *extra,
           "-Rpass-analysis=kernel-resource-usage", "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd = [hipcc(), *COMMON, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-x", "c++",
               "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    _compile.resources[src.name] = _resources(res.stderr) if src.suffix == ".hip" else []
    return obj


_compile.resources = {}


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    with open(RESOURCES, "w") as f:
        f.write("source\tkernel\tvgprs\tagprs\tscratch_bytes\tvgpr_spills\n")
        for name in sorted(_compile.resources):
            for row in _compile.resources[name]:
                f.write("\t".join([name, *map(str, row)]) + "\n")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    build_driver()
    build_cli()
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_driver() -> Path:
    """The C++ program that drives the BROCCOLI_LIB drop-in with the
    reference CLI's call sequence (tests/cpp/dropin_driver.cpp)."""
    src = ROOT / "tests" / "cpp" / "dropin_driver.cpp"
    exe = src.with_suffix("")
    if not src.exists():
        return exe
    cmd = ["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(src), "-o", str(exe),
           f"-L{PKG}", "-lbpg", "-Wl,-rpath,$ORIGIN/../../broccoli_amd", "-Wl,-rpath-link,/opt/rocm/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return exe


CLI = PKG / "RandomiseGroupLevel"


def build_cli() -> Path:
    """The RandomiseGroupLevel command line front end (cli/), linked against
    libbpg.so and zlib."""
    srcs = [PKG / "cli" / "randomise_group_level.cpp", PKG / "cli" / "nifti_io.cpp"]
    tmp = CLI.with_suffix(".tmp")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", f"-I{PKG / 'cli'}", *map(str, srcs),
           "-o", str(tmp), f"-L{PKG}", "-lbpg", "-lz", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"CLI build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
