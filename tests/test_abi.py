"""C ABI checks that need no GPU: the library loads, exports exactly what
include/valomc.h declares, and its host pre-pass (ErrorChecks + Init
without the device) reproduces the reference's neighbour tables and error
statuses from the golden vectors."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import cases
import valomc_amd as vmc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _declared():
    txt = open(os.path.join(ROOT, "include", "valomc.h")).read()
    return set(re.findall(r"\b(vmc_[a-z_0-9]+)\s*\(", txt))


def test_library_exports_every_declared_symbol():
    vmc.load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", vmc.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (vmc_\w+)", out))
    declared = _declared()
    assert declared, "no declarations parsed"
    assert declared <= exported, declared - exported
    assert exported <= declared, exported - declared


C_CALLER = r"""
#include "valomc.h"
#include <stdio.h>
int main(void) {
  vmc_problem_desc d = {0};
  vmc_run_desc run = {0};
  vmc_server_desc sd = {0};
  vmc_output out = {0};
  vmc_stats stats = {0};
  run.precision = VMC_FP32; run.tally = VMC_TALLY_FIXED64; run.first_photon = 0;
  sd.policy = VMC_SERVE_SHORTEST;
  (void)d; (void)out; (void)stats;
  printf("%s|%s|%d|%d|%d\n", vmc_version(), vmc_status_string(VMC_ERR_OUT_OF_MEMORY),
         vmc_problem_create(NULL, NULL), run.tally + (int)run.first_photon, sd.policy);
  return 0;
}
"""


def test_header_serves_c_callers(tmp_path):
    """include/valomc.h is plain C (the MEX gateway and FFI callers see C):
    a C99 program that uses its types and constants compiles warning-free
    and links to the library; its calls that need no GPU answer."""
    vmc.load_library()
    src, exe = tmp_path / "caller.c", tmp_path / "caller"
    src.write_text(C_CALLER)
    lib_dir = os.path.dirname(vmc.LIB_PATH)
    cc = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                         "-L", lib_dir, "-lvalomc", f"-Wl,-rpath,{lib_dir}", "-o", str(exe)],
                        capture_output=True, text=True)
    assert cc.returncode == 0, cc.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    version, oom, st, tally, policy = out.stdout.strip().split("|")
    assert version.startswith("valomc") and "Out of memory" in oom and int(st) == 18
    assert int(tally) == 1 and int(policy) == 0


def test_status_strings():
    lib = vmc.load_library()
    assert lib.vmc_status_string(0) == b"Success"
    assert lib.vmc_status_string(1) == b"No lightsource"          # cpp/Errors.hpp:27-28
    assert b"GaussianSigma" not in lib.vmc_status_string(15)
    assert lib.vmc_status_string(17).startswith(b"A light source")
    assert lib.vmc_version().startswith(b"valomc")


def _desc(p, keep):
    def ptr(a, dtype):
        if a is None:
            return None
        arr = np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel(order="F"))
        keep.append(arr)
        return arr.ctypes.data

    d = vmc.ProblemDesc()
    d.dim, d.Ne, d.Nb, d.Nr = p.dim, p.Ne, p.Nb, p.Nr
    d.H, d.HN, d.BH, d.r = ptr(p.H, np.int64), ptr(p.HN, np.int64), ptr(p.BH, np.int64), ptr(p.r, np.float64)
    d.mua, d.mus, d.g, d.n = (ptr(a, np.float64) for a in (p.mua, p.mus, p.g, p.n))
    d.BCType, d.BCn = ptr(p.BCType, np.uint8), ptr(p.BCn, np.float64)
    d.BCLNormal = ptr(p.BCLNormal, np.float64)
    d.BCLightDirectionType = ptr(p.BCLDirType, np.uint8)
    d.BCIntensity = ptr(p.BCIntensity, np.float64)
    d.GaussianSigma = ptr(p.GaussianSigma, np.float64)
    return d


def _check(p):
    lib = vmc.load_library()
    lib.vmc_problem_check.argtypes = [C.POINTER(vmc.ProblemDesc), C.c_void_p, C.POINTER(C.c_int64)]
    keep = []
    d = _desc(p, keep)
    hn = np.zeros(p.Ne * (p.dim + 1), dtype=np.int64)
    ns = C.c_int64(0)
    st = lib.vmc_problem_check(C.byref(d), hn.ctypes.data, C.byref(ns))
    return st, hn.reshape(p.dim + 1, p.Ne).T, ns.value


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_host_prepass_neighbour_table_matches_reference(name):
    d = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    p = cases.from_arrays(d)
    st, hn, ns = _check(p)
    assert st == 0
    ref = d["out_T1_HN"]
    assert np.array_equal(hn, ref)
    assert ns > 0


@pytest.mark.parametrize("name", sorted(cases.ERROR_CASES))
def test_host_prepass_error_status_matches_reference(name):
    d = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    st, _, _ = _check(cases.from_arrays(d))
    assert st == int(d["out_status"])


def _neighbours_only(dim, H, BH, Nr):
    lib = vmc.load_library()
    lib.vmc_build_neighbours.argtypes = [C.c_int32, C.c_int64, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p,
                                         C.c_void_p]
    H = np.asfortranarray(np.asarray(H, dtype=np.int64))
    ne = H.shape[0]
    bh = None if BH is None else np.asfortranarray(np.asarray(BH, dtype=np.int64))
    nb = 0 if bh is None else bh.shape[0]
    hn = np.zeros(ne * (dim + 1), dtype=np.int64)
    st = lib.vmc_build_neighbours(dim, ne, nb, Nr, H.ctypes.data, None if bh is None else bh.ctypes.data,
                                  hn.ctypes.data)
    return st, hn.reshape(dim + 1, ne).T


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_build_neighbours_alone_matches_reference(name):
    """vmc_build_neighbours, BuildNeighbourhoods() without the rest of
    ErrorChecks()/Init(): with BH it writes the reference's table; without
    BH (createBH3mex's use) a 3D boundary face matches nothing and stays
    INT64_MAX while interior faces keep their neighbours, and a 2D edge
    without a neighbour is the reference's MISSING_BOUNDARY."""
    d = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    p = cases.from_arrays(d)
    ref = d["out_T1_HN"]
    st, hn = _neighbours_only(p.dim, p.H, p.BH, p.Nr)
    assert st == 0 and np.array_equal(hn, ref)
    st, hn = _neighbours_only(p.dim, p.H, None, p.Nr)
    if p.dim == 3:
        assert st == 0
        assert np.array_equal(hn, np.where(ref < 0, np.iinfo(np.int64).max, ref))
    else:
        assert st == 1 + vmc.MCERRORS.index("MISSING_BOUNDARY")


def test_build_neighbours_alone_errors():
    d = dict(np.load(os.path.join(GOLDEN, "err_duplicate_neighbour_3d.npz"), allow_pickle=False))
    p = cases.from_arrays(d)
    st, _ = _neighbours_only(3, p.H, p.BH, p.Nr)
    assert st == int(d["out_status"])   # the reference's duplicate-neighbour error
    p = cases.p3_single_cube()
    H = np.asarray(p.H).copy()
    H[0, 0] = p.Nr   # a node id past the node count: INCONSISTENT_H
    assert _neighbours_only(3, H, p.BH, p.Nr)[0] == 1 + vmc.MCERRORS.index("INCONSISTENT_H")
    BH = np.asarray(p.BH).copy()
    BH[0, 0] = -1
    assert _neighbours_only(3, p.H, BH, p.Nr)[0] == 1 + vmc.MCERRORS.index("INCONSISTENT_BH")
    assert _neighbours_only(4, p.H, None, p.Nr)[0] == vmc.STATUS_INVALID_ARGUMENT


def test_undefined_direction_is_rejected():
    p = cases.p3_given_L_I()
    p.BCLDirType = None          # 'L' and 'I' then have no direction (MC3D.hpp:1112-1192)
    st, _, _ = _check(p)
    assert st == vmc.STATUS_UNDEFINED_DIRECTION


def test_neighbours_on_config_meshes():
    """Sort-based face matching on the 40 mm benchmark box: every interior
    face is shared by exactly two tets, every boundary triangle is used
    once, and the table is symmetric."""
    from valomc_amd import meshgen as mg
    p = mg.sphere_box_3d(nphoton=1, side_mm=12)
    st, hn, ns = _check(p)
    assert st == 0 and ns == 1
    inner = hn >= 0
    el = np.nonzero(inner)
    nbr = hn[inner]
    # symmetry: the neighbour lists this element back
    assert np.all(np.any(hn[nbr] == el[0][:, None], axis=1))
    bnd = -1 - hn[~inner]
    assert np.array_equal(np.sort(bnd), np.arange(p.Nb))


def test_brute_force_neighbours_small():
    """Pin the face matcher against an O(Ne^2) vertex-set comparison."""
    from valomc_amd import meshgen as mg
    r, H, BH = mg.grid_mesh(np.arange(3) + 0.5, np.arange(2) + 0.5, np.arange(2) + 0.5)
    p = mg.homogeneous(3, r, H, BH, 0.1, 1, 0.9, 1.4)
    p.BCType[0] = ord("l")
    st, hn, _ = _check(p)
    assert st == 0
    faces = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
    for e in range(p.Ne):
        for f, fv in enumerate(faces):
            key = set(H[e, list(fv)])
            others = [j for j in range(p.Ne) if j != e and key <= set(H[j])]
            bnds = [-1 - b for b in range(p.Nb) if key == set(BH[b])]
            cand = others + bnds
            assert len(cand) == 1
            assert hn[e, f] == cand[0]


def test_size_limits_are_refused_before_any_array_is_read():
    """Counts outside what the device layout addresses (2^28 elements,
    int32 boundary indices) or that no mesh has are refused with
    VMC_ERR_INVALID_ARGUMENT before the library reads a single array, so
    the descriptor can point at one-element buffers."""
    lib = vmc.load_library()
    lib.vmc_problem_check.argtypes = [C.POINTER(vmc.ProblemDesc), C.c_void_p, C.POINTER(C.c_int64)]
    one_i, one_f, one_c = np.zeros(1, np.int64), np.zeros(1), np.zeros(1, np.uint8)
    base = dict(dim=3, Ne=6, Nb=12, Nr=8)
    for bad in (dict(Ne=2 ** 28), dict(Nb=2 ** 31 - 1), dict(dim=4), dict(dim=1), dict(Ne=0), dict(Nb=-1),
                dict(Nr=0)):
        d = vmc.ProblemDesc()
        for k, v in {**base, **bad}.items():
            setattr(d, k, v)
        d.H = d.BH = one_i.ctypes.data
        d.r = d.mua = d.mus = d.g = d.n = d.BCn = one_f.ctypes.data
        d.BCType = one_c.ctypes.data
        assert lib.vmc_problem_check(C.byref(d), None, None) == vmc.STATUS_INVALID_ARGUMENT, bad
    d = vmc.ProblemDesc()
    d.dim, d.Ne, d.Nb, d.Nr = 3, 6, 12, 8   # counts fine, arrays missing
    assert lib.vmc_problem_check(C.byref(d), None, None) == vmc.STATUS_INVALID_ARGUMENT


def test_binding_refuses_arrays_of_the_wrong_length():
    """The library reads Ne, Nb, Nr (times the dimension) values from every
    pointer it is given, so the binding checks each array's length first
    and raises instead of handing over a short buffer."""
    vmc.load_library()
    p = cases.p3_single_cube()
    p.mua = p.mua[:-1]
    with pytest.raises(ValueError, match="mua"):
        vmc.Problem(p)
    p = cases.p3_single_cube()
    p.BCLNormal = np.zeros((p.Nb - 1, 3))
    with pytest.raises(ValueError, match="BCLNormal"):
        vmc.Problem(p)
    p = cases.p2_doc_quad()
    p.BCType = p.BCType[:-1]
    with pytest.raises(ValueError, match="BCType"):
        vmc.Problem(p)


def test_run_desc_layout_matches_the_header():
    """The Python mirror of vmc_run_desc has the C struct's size and field
    offsets (a C program prints them)."""
    import tempfile
    src = r"""
#include "valomc.h"
#include <stddef.h>
#include <stdio.h>
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(vmc_run_desc), offsetof(vmc_run_desc, model),
         offsetof(vmc_run_desc, first_photon), offsetof(vmc_run_desc, weight0),
         offsetof(vmc_run_desc, c0), sizeof(vmc_batch_item));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(c, "w").write(src)
        subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        got = [int(x) for x in subprocess.check_output([exe]).split()]
    R = vmc.RunDesc
    want = [C.sizeof(R), R.model.offset, R.first_photon.offset, R.weight0.offset, R.c0.offset,
            C.sizeof(vmc.BatchItem)]
    assert got == want, (got, want)
