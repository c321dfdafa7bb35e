"""The MEX gateways (valomc_amd/host/MC3Dmex_gpu.cpp, MC2Dmex_gpu.cpp,
createBH3mex_gpu.cpp) run without MATLAB against a mock of the MEX API
(tests/mex_mock/mex.h): a small
driver feeds each gateway the right-hand sides ValoMC.m passes, with their
MATLAB classes (int64 meshes, int8 boundary types, int64 photon count,
uint64 [seed, flag]), and writes back what mexFunction returns.

* no GPU needed: the argument count, classes and sizes are checked the way
  the reference gateways' Convert_mxArray and ErrorChecks do, and the
  errors come back as MATLAB errors;
* on the GPU: both gateways, with the progress bar (create + run with a
  waitbar callback) and without (vmc_monte_carlo), return the oracle's
  results photon for photon on the same streams, the neighbour table and
  the seed; pressing Cancel stops the run and keeps the planned-N
  normalisation.
"""
import os
import subprocess

import numpy as np
import pytest

import cases
import pyoracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _golden(name):
    return cases.from_arrays(dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)))


@pytest.fixture(scope="module")
def drivers(tmp_path_factory):
    import valomc_amd
    valomc_amd.load_library()
    d = tmp_path_factory.mktemp("mex")
    exe = {}
    for dim in (2, 3):
        out = str(d / f"drv{dim}")
        subprocess.check_call(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                               "-I", os.path.join(ROOT, "tests", "mex_mock"), "-I",
                               os.path.join(ROOT, "valomc_amd", "host"),
                               os.path.join(ROOT, "tests", "mex_mock", "driver.cpp"),
                               os.path.join(ROOT, "valomc_amd", "host", f"MC{dim}Dmex_gpu.cpp"), "-o", out,
                               "-L" + os.path.join(ROOT, "valomc_amd"), "-lvalomc",
                               "-Wl,-rpath," + os.path.join(ROOT, "valomc_amd")])
        exe[dim] = out
    return exe


@pytest.fixture(scope="module")
def bh3_driver(tmp_path_factory):
    import valomc_amd
    valomc_amd.load_library()
    out = str(tmp_path_factory.mktemp("mexbh") / "drvbh3")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                           "-I", os.path.join(ROOT, "tests", "mex_mock"),
                           os.path.join(ROOT, "tests", "mex_mock", "driver.cpp"),
                           os.path.join(ROOT, "valomc_amd", "host", "createBH3mex_gpu.cpp"), "-o", out,
                           "-L" + os.path.join(ROOT, "valomc_amd"), "-lvalomc",
                           "-Wl,-rpath," + os.path.join(ROOT, "valomc_amd")])
    return out


def _rhs(p, nphoton, seed, disable_pbar=1, overrides=None):
    """The gateway's right-hand sides for problem p, in argument order, as
    (class, array-or-None) with ValoMC.m's classes."""
    nb, dim = p.Nb, p.dim
    args = [("int64", np.asarray(p.H, dtype=np.int64)),
            ("int64", None if p.HN is None else np.asarray(p.HN, dtype=np.int64)),
            ("int64", np.asarray(p.BH, dtype=np.int64)),
            ("double", np.asarray(p.r, dtype=np.float64)),
            ("int8", np.asarray(p.BCType, dtype=np.int8).reshape(nb, 1)),
            ("double", None if p.BCIntensity is None else np.asarray(p.BCIntensity, dtype=np.float64).reshape(nb, 1)),
            ("int8", None if p.BCLDirType is None else np.asarray(p.BCLDirType, dtype=np.int8).reshape(nb, 1)),
            ("double", None if p.BCLNormal is None else np.asarray(p.BCLNormal, dtype=np.float64).reshape(nb, dim)),
            ("double", np.asarray(p.BCn, dtype=np.float64).reshape(nb, 1)),
            ("double", np.asarray(p.mua, dtype=np.float64).reshape(-1, 1)),
            ("double", np.asarray(p.mus, dtype=np.float64).reshape(-1, 1)),
            ("double", np.asarray(p.g, dtype=np.float64).reshape(-1, 1)),
            ("double", np.asarray(p.n, dtype=np.float64).reshape(-1, 1)),
            ("double", np.array([[p.f]])), ("double", np.array([[p.phase0]])),
            ("int64", np.array([[nphoton]], dtype=np.int64))]
    if dim == 2:
        args.append(("double", None if p.GaussianSigma is None
                     else np.asarray(p.GaussianSigma, dtype=np.float64).reshape(nb, 1)))
    args += [("int64", np.array([[disable_pbar]], dtype=np.int64)),
             ("uint64", np.array([[seed, 1]], dtype=np.uint64))]
    for k, v in (overrides or {}).items():
        args[k] = v
    return args


def _call(exe, args, tmp, nlhs=6, env=None):
    lines = []
    for k, (cls, a) in enumerate(args):
        if a is None or a.size == 0:
            lines.append(f"{cls} 0 0 -")
            continue
        a2 = a if a.ndim == 2 else a.reshape(-1, 1)
        path = os.path.join(tmp, f"in{k}.bin")
        np.asfortranarray(a2).ravel(order="F").tofile(path)
        lines.append(f"{cls} {a2.shape[0]} {a2.shape[1]} {path}")
    man = os.path.join(tmp, "manifest.txt")
    open(man, "w").write("\n".join(lines) + "\n")
    r = subprocess.run([exe, man, tmp, str(nlhs)], capture_output=True, text=True, timeout=300,
                       env={**os.environ, **(env or {})})
    outs, evals, err = {}, [], None
    for ln in r.stdout.splitlines():
        if ln.startswith("error "):
            err = ln[6:]
        elif ln.startswith("eval "):
            evals.append(ln[5:])
        elif ln.startswith("out "):
            _, k, cls, m, n, cplx = ln.split()
            m, n = int(m), int(n)
            dt = {"double": np.float64, "int64": np.int64, "uint64": np.uint64}[cls]
            raw = np.fromfile(os.path.join(tmp, f"out{k}.bin"), dtype=dt)
            if cplx == "1":
                raw = raw[:m * n] + 1j * raw[m * n:]
            outs[int(k)] = raw.reshape(n, m).T
    return r.returncode, outs, err, evals, r.stderr


def test_gateway_checks_arguments_like_the_reference(drivers, tmp_path):
    p = _golden("p3_single_cube")
    # argument count: the reference's syntax message
    rc, _, err, _, _ = _call(drivers[3], _rhs(p, 1000, 5)[:17], str(tmp_path))
    assert rc == 3 and err.startswith("Syntax:") and "MC3Dmex(H, HN, BH" in err
    rc, _, err, _, _ = _call(drivers[2], _rhs(_golden("p2_doc_quad"), 1000, 5)[:18], str(tmp_path))
    assert rc == 3 and "GaussianSigma, disablepbar, rndseed" in err
    # classes: H must be int64, as Convert_mxArray insists
    rc, _, err, _, _ = _call(drivers[3], _rhs(p, 1000, 5, overrides={0: ("double", np.asarray(p.H, float))}),
                             str(tmp_path))
    assert rc == 3 and "type error" in err and "H" in err
    # sizes: a short absorption array is the reference's SIZE_MISMATCH_MUA
    rc, _, err, _, _ = _call(drivers[3], _rhs(p, 1000, 5, overrides={9: ("double", np.asarray(p.mua)[:-1].reshape(-1, 1))}),
                             str(tmp_path))
    assert rc == 3 and "Absorption array is not equal in size" in err
    # an input error the library finds (no light source) comes back with the reference's text
    q = _golden("p3_single_cube")
    q.BCType = np.full(q.Nb, ord("a"), dtype=np.uint8)
    rc, _, err, _, _ = _call(drivers[3], _rhs(q, 1000, 5), str(tmp_path))
    assert rc == 3 and "No lightsource" in err


def _device():
    import valomc_amd
    try:
        valomc_amd.device_info()
        return True
    except Exception:
        return False


def test_gateway_without_a_gpu_reports_it(drivers, tmp_path):
    if _device():
        pytest.skip("a GPU is present")
    rc, _, err, _, _ = _call(drivers[3], _rhs(_golden("p3_single_cube"), 1000, 5), str(tmp_path))
    assert rc == 3 and "No GPU device" in err


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a - b)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["p3_direct_abs_fd", "p3_pencil_sphere", "p2_iso_gauss_fd", "p2_given_directions"])
@pytest.mark.parametrize("bar", [0, 1])
def test_gateway_matches_oracle(drivers, tmp_path, name, bar):
    p = _golden(name)
    n, seed = 20000, 31
    print(f"gateway MC{p.dim}Dmex on {name}, progress bar {'off' if bar else 'on'}", flush=True)
    rc, out, err, evals, log = _call(drivers[p.dim], _rhs(p, n, seed, disable_pbar=bar), str(tmp_path))
    assert rc == 0, (err, log[-2000:])
    q = _golden(name)
    q.Nphoton, q.seed = n, seed
    ref = pyoracle.run(q, 4, pyoracle.RNG_PHILOX)
    assert out[0].shape == (p.Ne, 1) and out[1].shape == (p.Nb, 1) and out[2].shape == (p.Nb, 1)
    assert rel_l2(out[0][:, 0].real, ref.ER) < 1e-9
    if p.f > 0:
        assert rel_l2(out[0][:, 0].imag, ref.EI) < 1e-9
        assert rel_l2(out[1][:, 0].imag, ref.EBI) < 1e-9
    assert rel_l2(out[1][:, 0].real, ref.EBR) < 1e-9
    if p.dim == 2:
        assert rel_l2(out[2][:, 0].real, ref.DEBR) < 1e-9
    assert out[3].shape == (1, 1) and out[3][0, 0] > 0
    assert int(out[4][0, 0]) == seed
    assert np.array_equal(out[5], ref.HN)
    if bar == 0:
        assert any("waitbar(0,'Please wait..'" in e for e in evals) and evals[-1] == "delete(mcwaitbar);"
    else:
        assert not any("waitbar" in e for e in evals)


@pytest.mark.gpu
def test_gateway_cancel_keeps_planned_normalisation(drivers, tmp_path):
    p = _golden("p2_square_cosinic")
    n = 200000
    rc, full, err, _, _ = _call(drivers[2], _rhs(p, n, 3, disable_pbar=0), str(tmp_path), nlhs=5)
    assert rc == 0, err
    rc, part, err, evals, log = _call(drivers[2], _rhs(p, n, 3, disable_pbar=0), str(tmp_path), nlhs=5,
                                      env={"VMC_MOCK_CANCEL_AFTER": "4"})
    assert rc == 0, err
    assert "ABORTED" in log
    # the fifth progress call sees Cancel: 5 of the 20 photon ranges ran and
    # the tallies are still divided by the planned N
    ratio = part[0].real.sum() / full[0].real.sum()
    assert 0.2 < ratio < 0.3, ratio


FACES3 = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]   # MC3D.hpp:1013, createBH3mex.cpp:50-77


@pytest.mark.parametrize("name", ["p3_pencil_sphere", "p3_cosinic_hetero_n", "p3_single_cube", "p3_iso_mirror"])
def test_create_bh3_gateway_lists_unshared_faces(bh3_driver, tmp_path, name):
    """createBH3mex_gpu: [BH] = createBH3mex(H) with 1-based int64 H gives
    the faces no other tetrahedron shares, 1-based, element-major in the
    reference's face and vertex order -- here the faces the reference's own
    neighbour table (written with the golden's BH) sends to the boundary,
    in the order of its element-major scan."""
    d = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    p = cases.from_arrays(d)
    H = np.asarray(p.H, dtype=np.int64)
    hn = d["out_T1_HN"]
    expect = [[H[e, v] + 1 for v in FACES3[f]] for e in range(p.Ne) for f in range(4) if hn[e, f] < 0]
    rc, out, err, _, _ = _call(bh3_driver, [("int64", H + 1)], str(tmp_path), nlhs=1)
    assert rc == 0, err
    assert out[0].dtype == np.int64 and out[0].shape == (len(expect), 3)
    assert out[0].tolist() == expect
    # the same set as the golden's boundary triangles
    assert sorted(map(sorted, out[0].tolist())) == sorted(sorted(int(v) + 1 for v in row) for row in p.BH)


def test_create_bh3_gateway_on_a_grid_by_brute_force(bh3_driver, tmp_path):
    """On a 4 x 3 x 3 voxel grid (createGridMesh numbering, six tets per
    voxel): every face counted once among all tetrahedra is a boundary
    face, in element-major order, two per voxel side on the box's faces."""
    from valomc_amd import meshgen as mg
    _, H, _ = mg.grid_mesh(np.arange(4) + 0.5, np.arange(3) + 0.5, np.arange(3) + 0.5)
    H = np.asarray(H, dtype=np.int64)
    count = {}
    for e in range(H.shape[0]):
        for f in FACES3:
            k = frozenset(H[e, list(f)].tolist())
            count[k] = count.get(k, 0) + 1
    expect = [[H[e, v] + 1 for v in f] for e in range(H.shape[0]) for f in FACES3
              if count[frozenset(H[e, list(f)].tolist())] == 1]
    rc, out, err, _, _ = _call(bh3_driver, [("int64", H + 1)], str(tmp_path), nlhs=1)
    assert rc == 0, err
    assert out[0].tolist() == expect and len(expect) == 2 * 2 * (4 * 3 + 4 * 3 + 3 * 3)


def test_create_bh3_gateway_checks_arguments(bh3_driver, tmp_path):
    H = np.asarray(cases.p3_single_cube().H, dtype=np.int64) + 1
    rc, _, err, _, _ = _call(bh3_driver, [("int64", H), ("int64", H)], str(tmp_path), nlhs=1)
    assert rc == 3 and err.startswith("Syntax:") and "createBH3mex(H)" in err
    rc, _, err, _, _ = _call(bh3_driver, [("double", H.astype(float))], str(tmp_path), nlhs=1)
    assert rc == 3 and "int64" in err
