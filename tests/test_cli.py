"""The MC2D / MC3D executables (valomc_amd/bin, built from valomc_amd/host)
read the text problems that ValoMC.m exports (ValoMC.m:200-258 for 2D,
293-336 for 3D) as the reference's executables do (MC3D.cpp:97-179,
MC2D.cpp:125-206, fileio.hpp:6-106), and write the solution in their format
(MC3D.cpp:186-201, MC2D.cpp:212-226).  They run on the drop-in class API of
valomc_amd/host/MC3D.hpp / MC2D.hpp, so these tests cover that too.

Without a GPU, ``--check`` parses a file and runs ErrorChecks() + Init()
(host only) and prints a summary, which is checked against the case: files
written with the MATLAB exporter's formatting and tag spellings
(BCLIntesity, BCLightDirectionType), and files with the 3D reader's own
spellings (BCIntensity, BCLDirType), in the exporter's order and in another
order.  Where the reference build is present, its own executables read the
same export back to the golden results, which pins the exporter.  On the GPU the executables' solutions must equal the library's
results for the same problem and seed bit for bit (fixed-point tallies,
17-digit output)."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
BIN = os.path.join(ROOT, "valomc_amd", "bin")
CASES = ["p3_direct_abs_fd", "p3_relative_cos_intensity", "p3_given_L_I", "p2_iso_gauss_fd",
         "p2_given_directions", "p2_square_cosinic"]


def _golden(name):
    d = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    return cases.from_arrays(d), d


def export_text(p, path, style):
    """Write problem p as ValoMC.m does (style "matlab": its formats and tag
    spellings, optional sections in its order), or with full precision and
    the 3D reader's spellings, in the exporter's order ("exact": what the
    reference's own executables read back exactly) or with the optional
    sections reversed ("reader")."""
    d3 = p.dim == 3
    full = style in ("reader", "exact")
    fx = (lambda v: f"{v:.17g}") if full else (lambda v: f"{v:18.10f}")
    lines = [f"{p.Ne} {p.Nb} {p.Nr} {p.Nphoton}"]
    if d3 or full:
        lines.append(f"{fx(p.f)} {fx(p.phase0)} {p.seed} 1")
    else:
        lines.append(f"{p.f:e} {p.phase0:e} {p.seed} 1")
    lines.append("0 0 0" if d3 else "0 0")
    lines.append("H")
    lines += [" ".join(str(int(v)) for v in row) for row in np.asarray(p.H)]
    lines.append("BH")
    lines += [" ".join(str(int(v)) for v in row) for row in np.asarray(p.BH)]
    lines.append("r")
    lines += [" ".join(fx(v) for v in row) for row in np.asarray(p.r)]
    lines.append("mua mus g n")
    lines += [" ".join(fx(v) for v in row) for row in zip(p.mua, p.mus, p.g, p.n)]
    lines.append("BCType")
    lines += [chr(int(c)) for c in p.BCType]
    opt = []
    if p.BCn is not None:
        opt.append(["BCn"] + [fx(v) for v in p.BCn])
    if p.BCLNormal is not None:
        opt.append(["BCLightDirection"] + [" ".join(fx(v) for v in row) for row in np.asarray(p.BCLNormal)])
    if p.BCLDirType is not None:
        tag = "BCLDirType" if full else "BCLightDirectionType"
        opt.append([tag] + [chr(int(c)) for c in p.BCLDirType])
    if p.BCIntensity is not None:
        tag = "BCIntensity" if full else "BCLIntesity"
        opt.append([tag] + [fx(v) for v in p.BCIntensity])
    if p.GaussianSigma is not None:
        opt.append(["GaussianSigma"] + [fx(v) for v in p.GaussianSigma])
    if style == "reader":
        opt.reverse()
    for sec in opt:
        lines += sec
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def _exe(p):
    exe = os.path.join(BIN, "MC3D" if p.dim == 3 else "MC2D")
    assert os.path.exists(exe), f"{exe} is not built: run `python -m valomc_amd.build`"
    return exe


@pytest.mark.parametrize("style", ["matlab", "exact", "reader"])
@pytest.mark.parametrize("name", CASES)
def test_executable_reads_exported_problem(tmp_path, name, style):
    p, d = _golden(name)
    fin = str(tmp_path / "in.txt")
    export_text(p, fin, style)
    out = subprocess.run([_exe(p), fin, str(tmp_path / "out.txt"), "--check"], capture_output=True,
                         text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    line = [ln for ln in out.stdout.splitlines() if ln.startswith("CHECK ")]
    assert line, out.stdout
    got = dict(kv.split("=") for kv in line[0].split()[1:])
    assert int(got["dim"]) == p.dim
    assert (int(got["Ne"]), int(got["Nb"]), int(got["Nr"])) == (p.Ne, p.Nb, p.Nr)
    assert int(got["Nphoton"]) == p.Nphoton and int(got["seed"]) == p.seed
    assert math.isclose(float(got["f"]), p.f, rel_tol=1e-9)
    phase0 = p.phase0 + (2 * math.pi * math.ceil(-p.phase0 / (2 * math.pi)) if p.phase0 < 0 else 0.0)
    assert math.isclose(float(got["phase0"]), phase0, rel_tol=1e-8, abs_tol=1e-9)
    # the neighbour table Init() built against the reference's own
    hn = np.asarray(d["out_T1_HN"], dtype=np.int64).ravel()
    assert int(got["hnsum"]) == int(np.fmod(hn, 1000003).sum())
    assert int(got["nonabsorbing_bc"]) == int(np.sum(np.asarray(p.BCType) != ord("a")))
    ndir = 0 if p.BCLDirType is None else int(np.sum(np.asarray(p.BCLDirType) != ord("n")))
    assert int(got["directed"]) == ndir
    for key, arr in (("bcn", p.BCn), ("bcl", p.BCLNormal), ("intensity", p.BCIntensity),
                     ("sigma", p.GaussianSigma)):
        assert int(got[key]) == int(arr is not None), key


@pytest.mark.parametrize("name", ["p3_direct_abs_fd", "p2_iso_gauss_fd"])
def test_reference_executable_reads_the_export(tmp_path, name):
    """The export is what the reference's own executables read: the serial
    reference build (oracle/_ref) run on it reproduces the golden serial
    results to its print precision (%18.10e in 3D, %e in 2D)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "MC3D_serial" if name.startswith("p3") else "MC2D_serial")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref is built only where the reference sources are present")
    p, d = _golden(name)
    fin, fout = str(tmp_path / "in.txt"), str(tmp_path / "out.txt")
    export_text(p, fin, "exact")
    out = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout
    tol = 1e-9 if p.dim == 3 else 1e-5
    for k, a in zip(("ER", "EI", "EBR", "EBI"), _read_solution(fout, p.Ne, p.Nb)):
        b = d["out_T1_" + k]
        assert np.max(np.abs(a - b)) <= tol * max(np.max(np.abs(b)), 1e-300), k


def test_executable_reports_reading_errors(tmp_path):
    p, _ = _golden("p3_direct_abs_fd")
    fin = str(tmp_path / "in.txt")
    export_text(p, fin, "matlab")
    text = open(fin).read().replace("mua mus g n", "mua mus g")
    open(fin, "w").write(text)
    out = subprocess.run([_exe(p), fin, str(tmp_path / "out.txt"), "--check"], capture_output=True,
                         text=True, timeout=60)
    assert out.returncode == 1 and "Reading error" in out.stdout


def _read_solution(path, ne, nb):
    rows = open(path).read().split("\n")
    assert re.match(r"^\d+ 0$", rows[0]), rows[0]
    vals = np.array([[float(x) for x in r.split()] for r in rows[1:1 + ne + nb]])
    return vals[:ne, 0], vals[:ne, 1], vals[ne:, 0], vals[ne:, 1]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["p3_direct_abs_fd", "p3_given_L_I", "p2_iso_gauss_fd", "p2_given_directions"])
def test_executable_solution_equals_library(tmp_path, name):
    import valomc_amd as vmc
    p, _ = _golden(name)
    fin, fout = str(tmp_path / "in.txt"), str(tmp_path / "out.txt")
    export_text(p, fin, "reader")
    print(f"case {name}", flush=True)
    out = subprocess.run([_exe(p), fin, fout, "-s", "--fixed", "--precise"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    ER, EI, EBR, EBI = _read_solution(fout, p.Ne, p.Nb)
    with vmc.Problem(p) as prob:
        r = prob.run(p.Nphoton, seed=p.seed, f=p.f, phase0=p.phase0, precision=vmc.FP64,
                     tally=vmc.TALLY_FIXED64)
    for a, b, k in ((ER, r.ER, "ER"), (EI, r.EI, "EI"), (EBR, r.EBR, "EBR"), (EBI, r.EBI, "EBI")):
        assert np.array_equal(a, b, equal_nan=True), k
