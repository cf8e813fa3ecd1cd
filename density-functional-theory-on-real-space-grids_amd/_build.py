"""Build ``libcdft_hip.so`` in-tree from ``csrc/*.hip`` with hipcc for gfx950.

One shared library, C ABI only (``include/cdft_hip.h``). Packed FP32 instructions are disabled on
the device side (``-packed-fp32-ops``); hipcc's host pass warns that it does not know the feature.
"""

from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

PACKAGE_DIR = pathlib.Path(__file__).resolve().parent
CSRC = PACKAGE_DIR / "csrc"
LIBRARY = PACKAGE_DIR / "libcdft_hip.so"
ARCH = "gfx950"


def hipcc() -> str:
    for candidate in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if candidate and pathlib.Path(candidate).exists():
            return candidate
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def sources() -> list[pathlib.Path]:
    return sorted(CSRC.glob("*.hip"))


def build(verbose: bool = False, jobs: int = 8) -> pathlib.Path:
    """Compile every ``csrc/*.hip`` to an object (in parallel) and link the shared library."""
    compiler = hipcc()
    objdir = PACKAGE_DIR / "_build"
    objdir.mkdir(exist_ok=True)
    common = [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", str(CSRC),
    ]
    procs = []
    objects = []
    for src in sources():
        obj = objdir / (src.stem + ".o")
        objects.append(obj)
        cmd = [compiler, *common, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        if len(procs) >= jobs:
            _drain(procs)
    _drain(procs)
    tmp = LIBRARY.with_suffix(".so.tmp")
    link = [compiler, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objects), "-o", str(tmp),
            "-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(link), flush=True)
    out = subprocess.run(link, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if out.returncode != 0:
        raise RuntimeError(f"link failed:\n{out.stdout}")
    os.replace(tmp, LIBRARY)
    return LIBRARY


def _drain(procs) -> None:
    errors = []
    for src, proc in procs:
        text, _ = proc.communicate()
        if proc.returncode != 0:
            errors.append(f"{src.name}:\n{text}")
        elif text and "error" in text.lower():
            errors.append(f"{src.name}:\n{text}")
    procs.clear()
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
