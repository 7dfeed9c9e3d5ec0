"""In-tree build of the detray_amd shared library (HIP, gfx950) and of the
oracle used by the tests.

``build_library()`` compiles ``libdetray_amd.so`` with hipcc for gfx950:
the C ABI, the device kernels and the host-side detector emitters.
``build_oracle()`` compiles the FP64 CPU restatement into
``oracle/liboracle.so`` with g++ (test infrastructure).
``build_engine_host()`` compiles the device engine's headers for the host
(hipcc --offload-host-only) into ``tests/native/libengine_host.so``, which
CPU tests drive (test infrastructure).
"""

from __future__ import annotations

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
LIB = PKG / "libdetray_amd.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"
ENGINE_HOST_SRC = ROOT / "tests" / "native" / "engine_host.cpp"
ENGINE_HOST_LIB = ROOT / "tests" / "native" / "libengine_host.so"
ADAPTER_TEST_SRC = ROOT / "tests" / "native" / "adapter_test.cpp"
ADAPTER_TEST_BIN = ROOT / "tests" / "native" / "adapter_test"

ARCH = os.environ.get("DETRAY_AMD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    # keep FP64 arithmetic in source order (no FMA contraction) so the
    # device path rounds like the oracle
    "-ffp-contract=off",
    # packed FP32 instructions disabled
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # export the C ABI only
    "-fvisibility=hidden",
    "-fvisibility-inlines-hidden",
    "-Wall",
]

# Flags of the kernels' translation units.  Machine-level loop-invariant code
# motion is off: the propagation loop body holds every phase of propagate()
# (stepping, navigation, actors), and hoisting the constants and addresses of
# rarely taken phases (the polynomial coefficients of atan2/log/sincos of the
# parameter updater, say) out of the loop kept them live in registers through
# all of it, which spilled the track state to scratch memory (the mixed
# kernel at two waves per SIMD: 824 -> 400 bytes of scratch per lane).
KERNEL_FLAGS = os.environ.get("DETRAY_AMD_KERNEL_FLAGS", "-mllvm -disable-machine-licm").split()


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def _sources() -> list[Path]:
    """The C ABI and host code, then one translation unit per group of
    kernel instantiations (compiled in parallel)."""
    return [CSRC / "capi.hip", CSRC / "host" / "builders.cpp", CSRC / "host" / "json_io.cpp",
            *sorted((CSRC / "kernels").glob("k_*.hip"))]


def _jobs() -> int:
    return max(1, min(int(os.environ.get("MAX_JOBS", os.cpu_count() or 1)), 16))


def build_library(out: Path = LIB) -> Path:
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    inc = [f"-I{INCLUDE}", f"-I{CSRC}"]
    hipcc = _hipcc()

    def compile_one(src: Path) -> Path:
        obj = objdir / (src.stem + ".o")
        lang = [] if src.suffix == ".hip" else ["-x", "hip"]
        extra = KERNEL_FLAGS if src.name.startswith("k_") else []
        _run([hipcc, *HIP_FLAGS, *extra, *inc, "-c", *lang, str(src), "-o", str(obj)])
        return obj

    with ThreadPoolExecutor(max_workers=_jobs()) as pool:
        objs = list(pool.map(compile_one, _sources()))
    tmp = out.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_oracle(out: Path = ORACLE_LIB) -> Path:
    cxx = os.environ.get("CXX", "g++")
    tmp = out.with_suffix(".so.tmp")
    cmd = [cxx, "-std=c++20", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           f"-I{INCLUDE}", str(ORACLE_DIR / "oracle.cpp"), "-o", str(tmp)]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_engine_host(out: Path = ENGINE_HOST_LIB) -> Path:
    """The engine (device/engine.hpp) compiled for the host, with the same
    contraction rule as the kernels' FP64 arithmetic."""
    tmp = out.with_suffix(".so.tmp")
    _run([_hipcc(), "-x", "hip", "--offload-host-only", "-O2", "-std=c++20", "-fPIC", "-shared",
          "-ffp-contract=off", f"-I{INCLUDE}", f"-I{CSRC}", str(ENGINE_HOST_SRC), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_adapter_test(out: Path = ADAPTER_TEST_BIN) -> Path:
    """The compiled test of the C++20 adapter (include/detray_amd.hpp): a
    host program (g++ -std=c++20) linked against libdetray_amd.so, the
    oracle and the HIP runtime (device buffers of the GPU checks)."""
    cxx = os.environ.get("CXX", "g++")
    rocm = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
    tmp = out.with_suffix(".tmp")
    _run([cxx, "-std=c++20", "-O2", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__", f"-I{INCLUDE}", f"-I{CSRC}",
          f"-I{ORACLE_DIR}", f"-I{rocm / 'include'}", str(ADAPTER_TEST_SRC), "-o", str(tmp),
          f"-L{PKG}", "-ldetray_amd", f"-L{ORACLE_DIR}", "-loracle", f"-L{rocm / 'lib'}", "-lamdhip64",
          f"-Wl,-rpath,{PKG}", f"-Wl,-rpath,{ORACLE_DIR}", f"-Wl,-rpath,{rocm / 'lib'}",
          "-Wl,-rpath,$ORIGIN/../../detray_amd", "-Wl,-rpath,$ORIGIN/../../oracle"])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_library()
    build_oracle()
    build_engine_host()
    build_adapter_test()


if __name__ == "__main__":
    build_all()
    print(f"built {LIB} and {ORACLE_LIB}")
