"""ValoMC text problem/solution files.

``write_problem`` produces the input file of the external executables the
way ValoMC.m exports it (ValoMC.m:200-258 for 2D, 293-336 for 3D: counts,
constants, a grid-size line, then tagged sections).  ``tags="reader"``
uses the spellings the reference readers accept (BCLDirType in 3D,
BCIntensity); ``tags="exporter"`` the ones the MATLAB exporter writes
(BCLightDirectionType, BCLIntesity).  ``read_solution`` parses an output
file (MC3D.cpp:186-201, MC2D.cpp:212-226).
"""
from __future__ import annotations

import numpy as np


def write_problem(path: str, p, tags: str = "reader", seed_given: bool = True) -> None:
    dim = p.dim
    fmt = "%18.10f" if dim == 3 else "%e"
    with open(path, "w") as f:
        f.write(f"{p.Ne} {p.Nb} {p.Nr} {int(p.Nphoton)}\n")
        if dim == 3:
            f.write(f"{p.f:18.10f} {p.phase0:18.10f} {int(p.seed)} {1 if seed_given else 0}\n")
            f.write("0 0 0\n")
        else:
            f.write(f"{p.f:e} {p.phase0:e} {int(p.seed)} {1 if seed_given else 0}\n")
            f.write("0 0\n")
        f.write("H\n")
        np.savetxt(f, np.asarray(p.H, dtype=np.int64), fmt="%d")
        f.write("BH\n")
        np.savetxt(f, np.asarray(p.BH, dtype=np.int64), fmt="%d")
        f.write("r\n")
        np.savetxt(f, np.asarray(p.r), fmt=fmt)
        f.write("mua mus g n\n")
        np.savetxt(f, np.stack([p.mua, p.mus, p.g, p.n], axis=1), fmt=fmt)
        f.write("BCType\n")
        f.write("".join(chr(c) + "\n" for c in np.asarray(p.BCType, dtype=np.uint8)))
        f.write("BCn\n")
        np.savetxt(f, np.asarray(p.BCn).reshape(-1, 1), fmt=fmt)
        if getattr(p, "BCLNormal", None) is not None:
            f.write("BCLightDirection\n")
            np.savetxt(f, np.asarray(p.BCLNormal), fmt=fmt)
        if getattr(p, "BCLDirType", None) is not None:
            if tags == "exporter" or dim == 2:
                f.write("BCLightDirectionType\n")
            else:
                f.write("BCLDirType\n")
            f.write("".join(chr(c) + "\n" for c in np.asarray(p.BCLDirType, dtype=np.uint8)))
        if getattr(p, "BCIntensity", None) is not None:
            f.write("BCLIntesity\n" if tags == "exporter" else "BCIntensity\n")
            np.savetxt(f, np.asarray(p.BCIntensity).reshape(-1, 1), fmt=fmt)
        if dim == 2 and getattr(p, "GaussianSigma", None) is not None:
            f.write("GaussianSigma\n")
            np.savetxt(f, np.asarray(p.GaussianSigma).reshape(-1, 1), fmt=fmt)


def read_solution(path: str, Ne: int, Nb: int):
    """Returns (seconds, ER, EI, EBR, EBI)."""
    with open(path) as f:
        head = f.readline().split()
        data = np.loadtxt(f, ndmin=2)
    assert data.shape[0] == Ne + Nb, (data.shape, Ne, Nb)
    return int(head[0]), data[:Ne, 0], data[:Ne, 1], data[Ne:, 0], data[Ne:, 1]
