"""ctypes access to the test oracle -- TEST INFRASTRUCTURE ONLY.

Loads oracle/liboracle.so (the CPU restatement) and, when built,
oracle/_ref/libref{2,3}d.so (the reference engine compiled from its own
sources).  Only tests/, the smoke check and bench.py's cpu_baseline leg use
this module; the product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
RNG_MT19937 = 0
RNG_PHILOX = 1


class _Problem(C.Structure):
    _fields_ = [
        ("Ne", C.c_int64), ("Nb", C.c_int64), ("Nr", C.c_int64),
        ("H", C.c_void_p), ("HN", C.c_void_p), ("BH", C.c_void_p), ("r", C.c_void_p),
        ("mua", C.c_void_p), ("mus", C.c_void_p), ("g", C.c_void_p), ("n", C.c_void_p),
        ("BCType", C.c_void_p), ("BCn", C.c_void_p), ("BCLNormal", C.c_void_p),
        ("BCLDirType", C.c_void_p), ("BCIntensity", C.c_void_p), ("GaussianSigma", C.c_void_p),
        ("f", C.c_double), ("phase0", C.c_double), ("Nphoton", C.c_int64), ("seed", C.c_uint64),
        ("weight0", C.c_double), ("chance", C.c_double), ("c0", C.c_double),
        ("stream_tag", C.c_uint32), ("pad", C.c_uint32), ("first_photon", C.c_uint64),
    ]


# the reference's model constants (MC3D.hpp:178-183)
WEIGHT0, CHANCE, C0 = 0.001, 0.1, 2.99792458e11


class _Result(C.Structure):
    _fields_ = [
        ("ER", C.c_void_p), ("EI", C.c_void_p), ("EBR", C.c_void_p), ("EBI", C.c_void_p),
        ("DEBR", C.c_void_p), ("DEBI", C.c_void_p), ("HN", C.c_void_p),
        ("loss", C.c_int64), ("ledger", C.c_double * 5), ("seconds", C.c_double),
    ]


@dataclass
class Result:
    ER: np.ndarray
    EI: np.ndarray
    EBR: np.ndarray
    EBI: np.ndarray
    DEBR: np.ndarray
    DEBI: np.ndarray
    HN: np.ndarray
    loss: int
    ledger: np.ndarray
    seconds: float


class OracleError(RuntimeError):
    def __init__(self, status: int):
        super().__init__(f"oracle status {status}")
        self.status = status


_libs = {}


def _load(name: str):
    if name not in _libs:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        lib = C.CDLL(path)
        _libs[name] = lib
    return _libs[name]


def have_reference() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libref3d.so"))


def _colmajor(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype)).ravel(order="F")


def _marshal(p, stream_tag=0, first_photon=0):
    keep = []

    def ptr(a, dtype):
        if a is None:
            return None
        arr = np.ascontiguousarray(_colmajor(a, dtype))
        keep.append(arr)
        return arr.ctypes.data

    s = _Problem()
    s.Ne, s.Nb, s.Nr = p.Ne, p.Nb, p.Nr
    s.H = ptr(p.H, np.int64)
    s.HN = ptr(p.HN, np.int64)
    s.BH = ptr(p.BH, np.int64)
    s.r = ptr(p.r, np.float64)
    s.mua, s.mus = ptr(p.mua, np.float64), ptr(p.mus, np.float64)
    s.g, s.n = ptr(p.g, np.float64), ptr(p.n, np.float64)
    s.BCType = ptr(p.BCType, np.uint8)
    s.BCn = ptr(p.BCn, np.float64)
    s.BCLNormal = ptr(p.BCLNormal, np.float64)
    s.BCLDirType = ptr(p.BCLDirType, np.uint8)
    s.BCIntensity = ptr(p.BCIntensity, np.float64)
    s.GaussianSigma = ptr(p.GaussianSigma, np.float64)
    s.f, s.phase0 = float(p.f), float(p.phase0)
    s.Nphoton, s.seed = int(p.Nphoton), int(p.seed) & 0xFFFFFFFFFFFFFFFF
    s.weight0 = float(getattr(p, "weight0", WEIGHT0))
    s.chance = float(getattr(p, "chance", CHANCE))
    s.c0 = float(getattr(p, "c0", C0))
    s.stream_tag, s.first_photon = int(stream_tag), int(first_photon)
    return s, keep


def _call(fn, p, *args, stream_tag=0, first_photon=0) -> Result:
    s, keep = _marshal(p, stream_tag, first_photon)
    nv = p.dim + 1
    out = dict(ER=np.zeros(p.Ne), EI=np.zeros(p.Ne), EBR=np.zeros(p.Nb), EBI=np.zeros(p.Nb),
               DEBR=np.zeros(p.Nb), DEBI=np.zeros(p.Nb), HN=np.zeros(p.Ne * nv, dtype=np.int64))
    res = _Result()
    for k, v in out.items():
        setattr(res, k, v.ctypes.data)
    st = fn(C.byref(s), *args, C.byref(res))
    if st != 0:
        raise OracleError(st)
    return Result(ER=out["ER"], EI=out["EI"], EBR=out["EBR"], EBI=out["EBI"], DEBR=out["DEBR"],
                  DEBI=out["DEBI"], HN=out["HN"].reshape(nv, p.Ne).T.copy(), loss=int(res.loss),
                  ledger=np.array(list(res.ledger)), seconds=float(res.seconds))


def run(p, nthreads: int = 1, rng_mode: int = RNG_MT19937, stream_tag: int = 0,
        first_photon: int = 0) -> Result:
    """Run the CPU restatement.  ``p`` may carry the model constants
    ``weight0``, ``chance`` and ``c0`` (the reference's defaults otherwise);
    in Philox mode photon i of the run draws from the stream (p.seed,
    stream_tag, first_photon + i), as the GPU's run with that tag and first
    photon does."""
    lib = _load("liboracle.so")
    fn = lib.oracle_run3d if p.dim == 3 else lib.oracle_run2d
    fn.restype = C.c_int
    return _call(fn, p, C.c_int(nthreads), C.c_int(rng_mode), stream_tag=stream_tag, first_photon=first_photon)


def run_reference(p, nthreads: int = 1) -> Result:
    """Run the reference engine itself (oracle/_ref, built from its sources)."""
    lib = _load(os.path.join("_ref", "libref3d.so" if p.dim == 3 else "libref2d.so"))
    fn = lib.ref3d_run if p.dim == 3 else lib.ref2d_run
    fn.restype = C.c_int
    return _call(fn, p, C.c_int(nthreads))


def philox(ctr, key):
    lib = _load("liboracle.so")
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.oracle_philox4x32_10(c, k, o)
    return list(o)


def xoshiro(state, n):
    """n outputs of the photon stream's xoshiro128** step from `state`."""
    lib = _load("liboracle.so")
    s = (C.c_uint32 * 4)(*state)
    o = (C.c_uint32 * n)()
    lib.oracle_xoshiro128ss(s, C.c_int(n), o)
    return list(o)
