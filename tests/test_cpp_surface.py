"""The header-only call surface (include/bulkx/bulk.hpp) and the C++ driver
facade (include/bulkx/algorithm.hpp): they compile on their own (CPU), and
the test programs built from tests/cpp/*.hip pass on the GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_bin")


@pytest.mark.parametrize("header", ["bulkx/bulk.hpp", "bulkx/algorithm.hpp"])
def test_header_compiles_alone(tmp_path, header):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    src = tmp_path / "t.hip"
    src.write_text(f"#include <{header}>\nint main() {{ return 0; }}\n")
    r = subprocess.run([hipcc, "-std=c++17", "--offload-arch=gfx950", "-fsyntax-only",
                        f"-I{os.path.join(ROOT, 'include')}", str(src)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]


@pytest.mark.gpu
@pytest.mark.parametrize("program", ["surface_test", "heap_test", "generic_test"])
def test_surface_program_passes(program):
    exe = os.path.join(BIN, program)
    assert os.path.exists(exe), f"tests/cpp/_bin/{program} is not built (run the build first)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "ALL PASS" in r.stdout
