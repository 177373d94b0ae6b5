"""In-tree build of the native libraries (no JIT cache: the .so files travel with the tree).

    libmxmesh.so   host set-up, g++ -O3 -ffp-contract=off
    libmxeuler.so  HIP kernels and the C ABI, hipcc for gfx950
    libmxmlp.so    the warm-start network (BF16 matrix-core kernels), hipcc for gfx950
    oracle/_build/liboracle.so   the CPU oracle (test infrastructure), g++ -O3 -ffp-contract=off
    tests/_build/adapter_check   a test program of the C++ adapter include/mxeuler.hpp, g++
    tests/_build/exact_check     a test program of the host exact sums (csrc/device/exact.hpp), g++
    oracle/_ref/json_dump        the reference runtime's JSON library as a checker (oracle/build_ref.sh;
                                 built only where the reference tree is)

Every build passes -ffp-contract=off, as the reference builds host and device code
(CMakeLists.txt:45-55), so that no product is fused into an FMA.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "autan_amd"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MXEULER_ARCH", "gfx950")

CXXFLAGS = ["-std=c++20", "-O3", "-ffp-contract=off", "-fPIC", "-shared", "-Wall", "-Wextra"]
HIPFLAGS = [
    f"--offload-arch={ARCH}",
    "-std=c++20",
    "-O3",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-rdc",
    "-fPIC",
    "-shared",
    # Packed FP32 instructions stay off: the compiler would otherwise pair float operations.
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, cwd=ROOT)


def build_mesh() -> Path:
    out = PKG / "libmxmesh.so"
    src = sorted((PKG / "csrc" / "mesh").glob("*.cpp"))
    _run(["g++", *CXXFLAGS, "-o", out, *src])
    return out


def build_device() -> Path:
    out = PKG / "libmxeuler.so"
    _run([HIPCC, *HIPFLAGS, "-o", out, PKG / "csrc" / "device" / "stepper.hip"])
    return out


def build_mlp() -> Path:
    out = PKG / "libmxmlp.so"
    _run([HIPCC, *HIPFLAGS, "-o", out, PKG / "csrc" / "mlp" / "mlp.hip"])
    return out


def build_oracle() -> Path:
    out = ROOT / "oracle" / "_build" / "liboracle.so"
    out.parent.mkdir(parents=True, exist_ok=True)
    flags = [f for f in CXXFLAGS if f != "-Wextra"]
    _run(["g++", *flags, "-fopenmp", "-o", out, ROOT / "oracle" / "euler_oracle.cpp"])
    return out


def build_adapter_check() -> Path:
    """The C++ adapter's test program, linked against the in-tree libmxeuler.so."""
    out = ROOT / "tests" / "_build" / "adapter_check"
    out.parent.mkdir(parents=True, exist_ok=True)
    _run(["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", "-Werror", "-I", ROOT / "include", "-o", out,
          ROOT / "tests" / "cpp" / "adapter_check.cpp", f"-L{PKG}", "-l:libmxeuler.so",
          "-Wl,-rpath,$ORIGIN/../../autan_amd"])
    return out


def build_exact_check() -> Path:
    """The host exact sums' test program (the header the HIP library's host code includes)."""
    out = ROOT / "tests" / "_build" / "exact_check"
    out.parent.mkdir(parents=True, exist_ok=True)
    _run(["g++", "-std=c++20", "-O2", "-ffp-contract=off", "-Wall", "-Wextra", "-Werror", "-o", out,
          ROOT / "tests" / "cpp" / "exact_check.cpp"])
    return out


def build_ref() -> None:
    _run(["sh", ROOT / "oracle" / "build_ref.sh"])


def build_all() -> None:
    if shutil.which("g++") is None:
        raise RuntimeError("g++ not found")
    build_mesh()
    build_oracle()
    build_device()
    build_mlp()
    build_adapter_check()
    build_exact_check()
    build_ref()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"mesh": build_mesh, "device": build_device, "oracle": build_oracle, "mlp": build_mlp,
         "adapter": build_adapter_check, "exact": build_exact_check, "ref": build_ref, "all": build_all}[w]()
