"""Build the native library (libbulkx.so) for gfx950 with hipcc.

Every ``.hip`` / ``.cpp`` file under ``bulk_amd/csrc`` is compiled to an
object in ``_build/`` (in parallel) and linked into ``bulk_amd/lib/libbulkx.so``.
The library is built in-tree so it travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "bulk_amd", "csrc")
BUILD = os.path.join(ROOT, "_build")
LIBDIR = os.path.join(ROOT, "bulk_amd", "lib")
LIB = os.path.join(LIBDIR, "libbulkx.so")

ARCH = os.environ.get("BULKX_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# -packed-fp32-ops keeps hipcc from emitting v_pk_{add,mul,fma}_f32.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{os.path.join(ROOT, 'include')}",
]


def _sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, name))
    return out


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build_library(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = _sources()
    workers = min(8, max(1, os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"[bulk_amd.build] built {LIB} from {len(srcs)} sources", flush=True)
    return LIB


CEILING_SRC = os.path.join(ROOT, "bulk_amd", "hbm_ceiling.hip")
CEILING = os.path.join(LIBDIR, "hbm_ceiling")


def build_ceiling(verbose: bool = True) -> str:
    """Compile the HBM ceiling program (read / copy / read-two-write-an-eighth
    kernels) that the benchmarks quote their roofline fractions against."""
    os.makedirs(LIBDIR, exist_ok=True)
    cmd = [HIPCC, *[c for c in COMMON if c != "-fPIC"], CEILING_SRC, "-o", CEILING]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"[bulk_amd.build] built {CEILING}", flush=True)
    return CEILING


CPP_TESTS = os.path.join(ROOT, "tests", "cpp")
CPP_BIN = os.path.join(CPP_TESTS, "_bin")


def build_cpp_tests(verbose: bool = True) -> list[str]:
    """Compile the C++ programs under tests/cpp (users of the header-only
    call surface and of the driver facade) against the built library."""
    os.makedirs(CPP_BIN, exist_ok=True)
    out = []
    for name in sorted(os.listdir(CPP_TESTS)):
        if not name.endswith(".hip"):
            continue
        exe = os.path.join(CPP_BIN, name[:-4])
        cmd = [HIPCC, *[c for c in COMMON if c != "-fPIC"], os.path.join(CPP_TESTS, name),
               f"-L{LIBDIR}", "-lbulkx", "-Wl,-rpath,$ORIGIN/../../../bulk_amd/lib", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        out.append(exe)
    if verbose:
        print(f"[bulk_amd.build] built {len(out)} C++ test program(s) in {CPP_BIN}", flush=True)
    return out


if __name__ == "__main__":
    build_library()
    build_ceiling()
    build_cpp_tests()
    sys.exit(0)
