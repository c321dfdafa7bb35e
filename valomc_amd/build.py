"""Build the HIP library (gfx950) and the oracle in-tree.

    python -m valomc_amd.build

Outputs valomc_amd/libvalomc.so (the product), valomc_amd/bin/MC2D, MC3D
(the drop-in command-line executables) and valomc_amd/bin/line_probe (the
memory-ceiling probe bench.py runs).  ``build_oracle`` compiles the
test oracle (and, where the reference tree exists, the reference engine)
through oracle/Makefile.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOSTSRC = os.path.join(HERE, "host")
LIB = os.path.join(HERE, "libvalomc.so")
BIN = os.path.join(HERE, "bin")
ARCH = os.environ.get("VMC_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# -packed-fp32-ops: no v_pk_*_f32 code generation; -munsafe-fp-atomics: FP64
# atomics become global_atomic_add_f64 on coarse-grained device memory;
# relaxed FP32 divide/sqrt only affects the single-precision transport path.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


# host side of the library: preparation, launch engine + C entry points, server
HOST_SOURCES = ("vmc_prepare.cpp", "vmc_engine.cpp", "vmc_server.cpp")


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=cwd)


def build_library() -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    k_obj = os.path.join(objdir, "vmc_kernels.o")
    _run([HIPCC, *HIP_FLAGS, *inc, "-c", os.path.join(CSRC, "vmc_kernels.hip"), "-o", k_obj])
    objs = [k_obj]
    for name in HOST_SOURCES:
        obj = os.path.join(objdir, name.replace(".cpp", ".o"))
        _run([HIPCC, *HIP_FLAGS, *inc, "-x", "hip", "-c", os.path.join(CSRC, name), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp,
          f"-Wl,--version-script={os.path.join(CSRC, 'exports.map')}"])
    os.replace(tmp, LIB)
    return LIB


def build_cli() -> None:
    """MC2D/MC3D executables over the C ABI (reference CLI syntax)."""
    if not os.path.isdir(HOSTSRC):
        return
    os.makedirs(BIN, exist_ok=True)
    for name in ("MC3D", "MC2D"):
        src = os.path.join(HOSTSRC, f"{name}_main.cpp")
        if not os.path.exists(src):
            continue
        _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", HOSTSRC,
              src, "-o", os.path.join(BIN, name), f"-L{HERE}", "-lvalomc",
              f"-Wl,-rpath,{HERE}"])


def build_probe() -> None:
    """The random-line probe (valomc_amd/bin/line_probe): memory ceilings
    for the walk's access pattern, run by `bench.py --line-probe`."""
    os.makedirs(BIN, exist_ok=True)
    _run([HIPCC, *HIP_FLAGS, os.path.join(CSRC, "line_probe.hip"), "-o", os.path.join(BIN, "line_probe")])


def build_oracle() -> None:
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-j8"])


def build_all() -> None:
    build_library()
    build_cli()
    build_probe()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "lib": build_library, "cli": build_cli, "probe": build_probe, "oracle": build_oracle}[w]()
