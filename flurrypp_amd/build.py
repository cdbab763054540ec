"""Build recipes for the native libraries of flurrypp_amd.

* ``libflurry_hip.so``  - the drop-in FR hot path (HIP, gfx950), built with hipcc.
* ``libflurry_host.so`` - host set-up (mesh, connectivity, operators, metrics), g++.

Both land in ``flurrypp_amd/lib/`` so that they travel with the source tree.
Run ``python -m flurrypp_amd.build`` to rebuild.
"""
from __future__ import annotations

import hashlib
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
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp"]


def _inputs(kind: str) -> list[Path]:
    """Every file a library is built from: its sources, the headers they may
    include and this recipe."""
    sub = CSRC / ("device" if kind == "hip" else "host")
    files = sorted(p for p in sub.iterdir() if p.suffix in (".hip", ".cpp", ".hpp", ".h"))
    return files + sorted(INC.glob("*.h")) + [Path(__file__)]


def source_hash(kind: str) -> str:
    h = hashlib.sha256()
    for p in _inputs(kind):
        h.update(p.name.encode())
        h.update(p.read_bytes())
    return h.hexdigest()


def _stamp(out: Path) -> Path:
    return out.with_name(out.name + ".srchash")


def build_host() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libflurry_host.so"
    srcs = [str(CSRC / "host" / "fr_basis.cpp"), str(CSRC / "host" / "fr_mesh.cpp")]
    digest = source_hash("host")
    _run(["g++", *HOST_FLAGS, f"-I{INC}", *srcs, "-o", str(out)])
    _stamp(out).write_text(digest + "\n")
    return out


def build_hip() -> Path:
    """Each source to an object (in parallel: each holds its own kernels and
    their launches), then one shared library."""
    LIB.mkdir(exist_ok=True)
    out = LIB / "libflurry_hip.so"
    dev = CSRC / "device"
    obj = LIB / "obj"
    obj.mkdir(exist_ok=True)
    names = ("fhip_api", "fhip_staged", "fhip_fused")
    digest = source_hash("hip")
    procs = []
    for n in names:
        cmd = [HIPCC, *HIP_FLAGS, f"-I{INC}", f"-I{dev}", "-c", str(dev / f"{n}.hip"), "-o", str(obj / f"{n}.o")]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    bad = [n for n, p in zip(names, procs) if p.wait() != 0]
    if bad:
        raise subprocess.CalledProcessError(1, f"hipcc {' '.join(bad)}")
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, "-shared", *(str(obj / f"{n}.o") for n in names), "-o", str(tmp)])
    os.replace(tmp, out)
    _stamp(out).write_text(digest + "\n")
    return out


def is_current(kind: str) -> bool:
    """True when the library exists and was built from the sources as they
    are now (the hash stored beside it at build time matches)."""
    out = LIB / ("libflurry_hip.so" if kind == "hip" else "libflurry_host.so")
    st = _stamp(out)
    return out.exists() and st.exists() and st.read_text().strip() == source_hash(kind)


def build_all() -> None:
    build_host()
    build_hip()


def ensure_built() -> None:
    """Build whichever library is missing or older than its sources (a fresh
    checkout holds none), so that no run uses a stale library."""
    if not is_current("host"):
        build_host()
    if not is_current("hip"):
        build_hip()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"host": build_host, "hip": build_hip, "all": build_all}[w]()
