"""Build libpm.so (HIP, gfx950) in-tree with hipcc.

    python -m pmgpu.build            (from the package directory)

Each ``csrc/*.hip`` file is compiled to an object in parallel, then linked into
``pmgpu/libpm.so`` next to this file.  The build always recompiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
SRC_DIR = PKG_DIR.parent / "csrc"
INCLUDE_DIR = PKG_DIR.parent.parent / "include"
LIB_PATH = PKG_DIR / "libpm.so"
BUILD_DIR = PKG_DIR.parent / "_build"

ARCH = "gfx950"
CXXFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-munsafe-fp-atomics",
    # packed-FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE_DIR}",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build libpm")
    return exe


def _compile(src: Path, obj: Path) -> tuple[Path, int, str]:
    cmd = [hipcc(), *CXXFLAGS, "-c", str(src), "-o", str(obj)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(line for line in (p.stdout + p.stderr).splitlines()
                    if "packed-fp32-ops" not in line)
    return src, p.returncode, out


def build(verbose: bool = False) -> Path:
    BUILD_DIR.mkdir(exist_ok=True)
    sources = sorted(SRC_DIR.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources under {SRC_DIR}")
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        futs = []
        for s in sources:
            o = BUILD_DIR / (s.stem + ".o")
            objs.append(o)
            futs.append(ex.submit(_compile, s, o))
        errors = []
        for f in futs:
            src, rc, out = f.result()
            if verbose and out.strip():
                print(out)
            if rc != 0:
                errors.append(f"{src.name}:\n{out}")
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-o", str(tmp), *map(str, objs)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
