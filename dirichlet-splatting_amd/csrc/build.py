"""Build libdsplat_hip.so for gfx950 in-tree: every csrc/*.hip -> object -> one shared library.

    python dirichlet-splatting_amd/csrc/build.py [--jobs N]

Objects go to dirichlet-splatting_amd/_build/, the library to
dirichlet-splatting_amd/dsplat/_lib/libdsplat_hip.so, where dsplat._lib loads it.
"""

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

CSRC = Path(__file__).resolve().parent
ROOT = CSRC.parent
BUILD = ROOT / "_build"
LIB = ROOT / "dsplat" / "_lib" / "libdsplat_hip.so"
ARCH = os.environ.get("DSPLAT_ARCH", "gfx950")

FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # Packed FP32 VALU forms are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    # MFMA results stay in ArchVGPRs: no accumulator-to-VGPR copies after each tile and
    # one register file for the occupancy budget.
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall",
    "-Wno-unused-function",
]


def _compile(source: Path) -> Path:
    obj = BUILD / (source.stem + ".o")
    cmd = ["hipcc", *FLAGS, "-c", str(source), "-o", str(obj)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"hipcc failed on {source.name}:\n{result.stderr}")
    if result.stderr.strip():
        noise = [ln for ln in result.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if noise:
            print(f"[{source.name}]\n" + "\n".join(noise), file=sys.stderr)
    return obj


def build(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as pool:
        objects = list(pool.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objects), "-o", str(tmp)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"link failed:\n{result.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    parser = argparse.ArgumentParser()
    parser.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    print(build(parser.parse_args().jobs))
