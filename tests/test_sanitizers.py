"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer (no
GPU needed), as the reference keeps ASan targets (cpp/3d/makefile:8,
cpp/2d/makefile:13).

The library's host sources (ErrorChecks/Init pre-pass on up to 16 threads:
face matching, Morton order, light sources, element records; the device-free
entry points and the error paths) and the oracle are compiled with
-fsanitize=address,undefined (host side only; the GPU kernels are the
regular build's) and linked into tests/sanitize_driver.cpp, which runs them
on golden cases and on the benchmark meshes, the 2M-tet frequency-domain box
among them.  Any sanitizer report fails the run.  The test program also reports
the device element order the pre-pass builds, checked here: whole sectors,
element and slot maps inverse to each other, nothing pointing at a padding
slot, and on the grid meshes every cell inside one sector.
"""
import os
import subprocess

import numpy as np
import pytest

import cases
from valomc_amd import meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
HIPCC = "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/llvm/bin/clang++"
# sanitizers for host code only: each flag behind -Xarch_host
HOST_SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]
CPU_SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    kernels = os.path.join(ROOT, "valomc_amd", "_build", "vmc_kernels.o")
    if not (os.path.exists(HIPCC) and os.path.exists(CLANG) and os.path.exists(kernels)):
        pytest.skip("ROCm compilers or the library build are not here")
    d = tmp_path_factory.mktemp("san")
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "valomc_amd", "csrc"), "-I",
           os.path.join(ROOT, "oracle")]
    objs = []
    for name in ("vmc_prepare", "vmc_engine", "vmc_server"):
        o = str(d / (name + ".o"))
        subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
                               "-Wno-unused-command-line-argument", *HOST_SAN, *inc, "-x", "hip", "-c",
                               os.path.join(ROOT, "valomc_amd", "csrc", name + ".cpp"), "-o", o],
                              stderr=subprocess.DEVNULL)
        objs.append(o)
    o = str(d / "vmc_oracle.o")
    subprocess.check_call([CLANG, "-O1", "-g", "-std=c++17", *CPU_SAN, "-fopenmp", *inc, "-c",
                           os.path.join(ROOT, "oracle", "vmc_oracle.cpp"), "-o", o])
    objs.append(o)
    o = str(d / "sanitize_driver.o")   # includes the library's internal header: HIP host code
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-Wno-unused-command-line-argument",
                           *HOST_SAN, *inc, "-x", "hip", "-c", os.path.join(ROOT, "tests", "sanitize_driver.cpp"),
                           "-o", o], stderr=subprocess.DEVNULL)
    objs.append(o)
    exe = str(d / "sanitize_driver")
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", *HOST_SAN, "-fopenmp", *objs, kernels, "-o", exe],
                          stderr=subprocess.DEVNULL)
    supp = str(d / "lsan.supp")
    open(supp, "w").write("leak:libomp.so\n")   # the OpenMP runtime's one-time allocation at start-up
    return exe, supp


def _run(driver, p, tmp, nphoton):
    exe, supp = driver
    cm = lambda a, t: np.asfortranarray(np.asarray(a, dtype=t)).ravel(order="F")
    for k, t in (("H", np.int64), ("BH", np.int64), ("r", np.float64), ("mua", np.float64), ("mus", np.float64),
                 ("g", np.float64), ("n", np.float64), ("BCn", np.float64)):
        cm(getattr(p, k), t).tofile(os.path.join(tmp, k))
    np.asarray(p.BCType, dtype=np.uint8).tofile(os.path.join(tmp, "BCType"))
    inten = np.ones(p.Nb) if p.BCIntensity is None else np.asarray(p.BCIntensity, dtype=np.float64)
    inten.tofile(os.path.join(tmp, "BCIntensity"))
    env = {**os.environ, "ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0", "UBSAN_OPTIONS": "halt_on_error=1",
           "LSAN_OPTIONS": f"suppressions={supp}", "OMP_NUM_THREADS": "4"}
    r = subprocess.run([exe, tmp, str(p.dim), str(p.Ne), str(p.Nb), str(p.Nr), str(nphoton)], capture_output=True,
                       text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    out = {}
    for ln in r.stdout.splitlines():
        k, *v = ln.split()
        out.setdefault(k, []).append(v)
    return out


def _check(out, nphoton, expect_sources=True, grid=False):
    assert out["check"] == [["0"]]
    st, slots, used, mismatch, to_padding, split = (int(x) for x in out["order"][0])
    assert st == 0 and slots % 8 == 0 and mismatch == 0 and to_padding == 0, out["order"]
    if grid:   # whole cells per sector: a 3D grid takes 4 slots per 3 tets, a 2D one fills its sectors
        assert split == 0 and (4 * used == 3 * slots or (used + 7) // 8 * 8 == slots), out["order"]
    if expect_sources:
        assert int(out["nsources"][0][0]) > 0
        assert out["sources"][0] == ["0", out["nsources"][0][0]]
    assert out["bad_h"] == [["11"]]                      # 1 + INCONSISTENT_H
    # BuildNeighbourhoods alone: with BH the pre-pass's table; without it a
    # 3D mesh leaves exactly its boundary faces unmatched and a 2D one is
    # MISSING_BOUNDARY (1 + 15)
    sb, same, total, st, unmatched = (int(x) for x in out["neighbours_only"][0])
    assert sb == 0 and same == total
    assert st == 16 or (st == 0 and unmatched == int(out["nsources"][0][2]))
    if nphoton >= 0:
        for mode in out["oracle"]:
            assert mode[1] == "0" and abs(float(mode[5])) < 1e-9, mode


@pytest.mark.parametrize("name", ["p3_cosinic_hetero_n", "p3_iso_mirror", "p3_vacuum_nonabsorbing",
                                  "p2_square_cosinic", "p2_mirror_hetero"])
def test_golden_cases_under_sanitizers(driver, tmp_path, name):
    p = cases.from_arrays(dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)))
    _check(_run(driver, p, str(tmp_path), 5000), 5000)


def test_benchmark_meshes_under_sanitizers(driver, tmp_path):
    """The host preparation of the 384k-tet box, the 2,035,200-tet FD box
    (16 host threads) and the serving pool's 2D and 3D meshes."""
    box = meshgen.sphere_box_3d(nphoton=1, seed=1234)
    # the pencil source needs its position, which the test program does not pass:
    # use the top face's direct light instead
    box.BCType[:] = ord("a")
    top = np.all(np.abs(np.asarray(box.r)[np.asarray(box.BH)][:, :, 2] - 20.0) < 1e-9, axis=1)
    box.BCType[top] = ord("l")
    _check(_run(driver, box, str(tmp_path), 300), 300, grid=True)
    fd = meshgen.fd_box_3d(nphoton=1, seed=7)
    _check(_run(driver, fd, str(tmp_path), -1), -1, grid=True)
    for q in meshgen.serving_pool(seed=2024, scale=0.05):
        _check(_run(driver, q, str(tmp_path), 500), 500, grid=True)
