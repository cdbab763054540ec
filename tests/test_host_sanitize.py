"""libflurry_host's sources under AddressSanitizer, UBSan and LeakSanitizer
(CPU only): tests/host_sanitize/driver.cpp calls every entry point of the
host API (operators and P-multigrid operators at orders 0-5, Cartesian and
array meshes in 2-D and 3-D with periodic and mixed boundaries, jitter, face
lists, metric terms, curved 9-node quadrangles and 27-node hexahedra, a Gmsh
file read back and a malformed one rejected) in a sanitized build."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_library_clean_under_sanitizers(tmp_path):
    exe = tmp_path / "fhost_san"
    host = ROOT / "flurrypp_amd" / "csrc" / "host"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                    "-fopenmp", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "host_sanitize" / "driver.cpp"),
                    str(host / "fr_basis.cpp"), str(host / "fr_mesh.cpp"), "-o", str(exe)], check=True)
    env = dict(os.environ, UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", ASAN_OPTIONS="detect_leaks=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([str(exe), str(tmp_path)], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert "ok" in r.stdout
