"""Build recipes for the native libraries of flurrypp_amd.

* ``libflurry_hip.so``  - the drop-in FR hot path (HIP, gfx950), built with hipcc.
* ``libflurry_host.so`` - host set-up (mesh, connectivity, operators, metrics), g++.

Both land in ``flurrypp_amd/lib/`` so that they travel with the source tree.
Run ``python -m flurrypp_amd.build`` to rebuild.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "lib"
INC = ROOT / "include"
CSRC = PKG / "csrc"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FHIP_ARCH", "gfx950")

# Packed-FP32 instructions are kept out of every device code object.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_host() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libflurry_host.so"
    srcs = [str(CSRC / "host" / "fr_basis.cpp"), str(CSRC / "host" / "fr_mesh.cpp")]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", f"-I{INC}", *srcs, "-o", str(out)])
    return out


def build_hip() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libflurry_hip.so"
    dev = CSRC / "device"
    srcs = [str(dev / f) for f in ("fhip_api.hip", "fhip_staged.hip", "fhip_fused.hip")]
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, f"-I{INC}", f"-I{dev}", *srcs, "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_host()
    build_hip()


def ensure_built() -> None:
    """Build whichever library is missing (a fresh checkout holds none); a
    library that exists is used as it is."""
    if not (LIB / "libflurry_host.so").exists():
        build_host()
    if not (LIB / "libflurry_hip.so").exists():
        build_hip()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"host": build_host, "hip": build_hip, "all": build_all}[w]()
