"""MI355X photon-transport engine with ValoMC's MC3D/MC2D interface.

Python binding of the C ABI in include/valomc.h (libvalomc.so, built
in-tree by ``python -m valomc_amd.build``).  ``MC3D``/``MC2D`` mirror the
reference engine classes (cpp/3d/MC3D.hpp:51-173, cpp/2d/MC2D.hpp:59-180):
set the public arrays, call ``ErrorChecks()``, ``Init()`` and
``MonteCarlo()``, read ``ER, EI, EBR, EBI, DEBR, DEBI`` and ``loss``.
There is no CPU fallback: without the HIP library or a GPU the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvalomc.so")

FP64 = 0
FP32 = 1
TALLY_FP64 = 0
TALLY_FIXED64 = 1
SERVE_SHORTEST = 0   # server policy: least estimated remaining device time first
SERVE_FAIR = 1       # server policy: groups in turn, equal shares of a launch

# reference mcerror order (cpp/Errors.hpp:4-21); status = 1 + code
MCERRORS = (
    "NO_LIGHTSOURCE", "SIZE_MISMATCH_G", "SIZE_MISMATCH_MUA", "SIZE_MISMATCH_MUS",
    "SIZE_MISMATCH_N", "SIZE_MISMATCH_BCN", "MISSING_BCN", "SIZE_MISMATCH_BCTYPE",
    "SIZE_MISMATCH_LIGHT_DIRECTION", "SIZE_MISMATCH_LIGHT_DIRECTION_TYPE", "INCONSISTENT_H",
    "INCONSISTENT_BH", "INCONSISTENT_MESH", "INCONSISTENT_MESH_DUPLICATE_NEIGHBORS",
    "NO_GAUSSIAN_SIGMA", "MISSING_BOUNDARY",
)
STATUS_UNDEFINED_DIRECTION = 17
STATUS_INVALID_ARGUMENT = 18
STATUS_ABORTED = 103
STATUS_TALLY_RANGE = 104


class ProblemDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("Ne", C.c_int64), ("Nb", C.c_int64), ("Nr", C.c_int64),
        ("H", C.c_void_p), ("HN", C.c_void_p), ("BH", C.c_void_p), ("r", C.c_void_p),
        ("mua", C.c_void_p), ("mus", C.c_void_p), ("g", C.c_void_p), ("n", C.c_void_p),
        ("BCType", C.c_void_p), ("BCn", C.c_void_p), ("BCLNormal", C.c_void_p),
        ("BCLightDirectionType", C.c_void_p), ("BCIntensity", C.c_void_p),
        ("GaussianSigma", C.c_void_p),
    ]


class RunDesc(C.Structure):
    _fields_ = [
        ("f", C.c_double), ("phase0", C.c_double), ("Nphoton", C.c_int64), ("seed", C.c_uint64),
        ("precision", C.c_int32), ("tally", C.c_int32), ("block_size", C.c_int32),
        ("blocks_per_cu", C.c_int32), ("replicas", C.c_int32), ("refill_min", C.c_int32),
        ("stream_tag", C.c_uint32), ("lds_slots", C.c_int32), ("direct_flush", C.c_int32),
        ("chunk", C.c_int32), ("cache_epoch", C.c_int32), ("model", C.c_int32),
        ("first_photon", C.c_uint64), ("weight0", C.c_double), ("chance", C.c_double), ("c0", C.c_double),
    ]


# the reference's model constants (MC3D.hpp:178-183): roulette threshold,
# roulette survival chance, speed of light in mm/s
WEIGHT0, CHANCE, C0 = 0.001, 0.1, 2.99792458e11


class Output(C.Structure):
    _fields_ = [("ER", C.c_void_p), ("EI", C.c_void_p), ("EBR", C.c_void_p), ("EBI", C.c_void_p),
                ("DEBR", C.c_void_p), ("DEBI", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [
        ("photons", C.c_int64), ("loss", C.c_int64), ("steps", C.c_int64),
        ("kernel_ms", C.c_double), ("total_ms", C.c_double), ("ledger", C.c_double * 5),
        ("rechecks", C.c_int64), ("grid_blocks", C.c_int32), ("block_size", C.c_int32),
        ("lds_mesh", C.c_int32),
    ]


PROGRESS_FN = C.CFUNCTYPE(C.c_int, C.c_double, C.c_void_p)


class VmcError(RuntimeError):
    """Non-zero status from the C ABI."""

    def __init__(self, status: int, lib=None):
        msg = _status_string(status)
        if lib is not None and status >= 100:
            detail = lib.vmc_last_error().decode()
            if detail:
                msg += f" ({detail})"
        super().__init__(f"status {status}: {msg}")
        self.status = status
        self.mcerror = MCERRORS[status - 1] if 1 <= status <= 16 else None


_lib = None


def load_library():
    """Load libvalomc.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} not built: run `python -m valomc_amd.build`")
        lib = C.CDLL(LIB_PATH)
        lib.vmc_problem_create.argtypes = [C.POINTER(ProblemDesc), C.POINTER(C.c_void_p)]
        lib.vmc_problem_run.argtypes = [C.c_void_p, C.POINTER(RunDesc), C.POINTER(Output),
                                        C.POINTER(Stats), C.c_void_p, C.c_void_p]
        lib.vmc_problem_run_noalloc.argtypes = lib.vmc_problem_run.argtypes
        lib.vmc_problem_get_hn.argtypes = [C.c_void_p, C.c_void_p]
        lib.vmc_problem_slots.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        lib.vmc_problem_set_medium.argtypes = [C.c_void_p] + [C.c_void_p] * 4
        lib.vmc_problem_destroy.argtypes = [C.c_void_p]
        lib.vmc_batch_run.argtypes = [C.c_void_p, C.c_int32, C.POINTER(Stats)]
        lib.vmc_batch_autotune.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.POINTER(RunDesc),
                                           C.POINTER(C.c_int32)]
        lib.vmc_batch_plan_create.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]
        lib.vmc_batch_plan_run.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(Stats)]
        lib.vmc_batch_plan_destroy.argtypes = [C.c_void_p]
        lib.vmc_batch_plan_destroy.restype = None
        lib.vmc_problem_autotune.argtypes = [C.c_void_p, C.POINTER(RunDesc), C.POINTER(RunDesc), C.c_int64,
                                             C.POINTER(C.c_int32)]
        lib.vmc_problem_create_variant.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]
        lib.vmc_server_create.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        lib.vmc_server_submit.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]
        lib.vmc_server_poll.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_int32)]
        lib.vmc_server_wait.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(Stats)]
        lib.vmc_server_counters.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        lib.vmc_server_destroy.argtypes = [C.c_void_p]
        lib.vmc_server_destroy.restype = None
        lib.vmc_problem_destroy.restype = None
        lib.vmc_status_string.restype = C.c_char_p
        lib.vmc_last_error.restype = C.c_char_p
        lib.vmc_version.restype = C.c_char_p
        lib.vmc_device_info.argtypes = [C.c_char_p, C.c_int32, C.POINTER(C.c_int32)]
        _lib = lib
    return _lib


def _status_string(st: int) -> str:
    try:
        return load_library().vmc_status_string(st).decode()
    except FileNotFoundError:
        return "unknown"


def device_info() -> str:
    """vmc_device_info: the GPU's name, CU count and memory."""
    lib = load_library()
    buf = C.create_string_buffer(256)
    ncu = C.c_int32(0)
    st = lib.vmc_device_info(buf, 256, C.byref(ncu))
    if st:
        raise VmcError(st, lib)
    return buf.value.decode()


def _colmajor(a, dtype):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel(order="F"))


def _sized(a, dtype, n: int, name: str):
    """``a`` as a column-major buffer of exactly ``n`` values; the library
    reads that many, so a shorter or longer array is refused here."""
    if isinstance(a, (str, bytes)):
        a = np.frombuffer(a.encode() if isinstance(a, str) else a, dtype=np.uint8)
    arr = _colmajor(a, dtype)
    if arr.size != n:
        raise ValueError(f"{name} has {arr.size} values, expected {n}")
    return arr


@dataclass
class RunResult:
    """Outputs of one solve (normalised as the reference leaves them, in the
    caller's element order) and its statistics (vmc_stats)."""
    ER: np.ndarray
    EI: np.ndarray
    EBR: np.ndarray
    EBI: np.ndarray
    DEBR: np.ndarray
    DEBI: np.ndarray
    loss: int
    steps: int
    kernel_ms: float
    total_ms: float
    ledger: np.ndarray
    photons: int
    grid_blocks: int
    block_size: int
    status: int = 0
    rechecks: int = 0
    lds_mesh: bool = False   # the element records were held in LDS (small mesh)


def make_run_desc(nphoton: int, seed: int = 5489, f: float = 0.0, phase0: float = 0.0,
                  precision: int = FP64, tally: int = TALLY_FP64, block_size: int = 0,
                  blocks_per_cu: int = 0, replicas: int = 0, refill_min: int = 0, stream_tag: int = 0,
                  lds_slots: int = 0, direct_flush: int = 0, chunk: int = 0,
                  cache_epoch: int = 0, first_photon: int = 0, weight0: Optional[float] = None,
                  chance: Optional[float] = None, c0: Optional[float] = None) -> RunDesc:
    """vmc_run_desc with the C ABI's field meanings (include/valomc.h).
    ``weight0``, ``chance`` and ``c0`` are the model constants (MC3D::weight0,
    chance, c0); giving any of them sets ``model`` = 1 with the reference's
    defaults for the others."""
    rd = RunDesc()
    rd.f, rd.phase0, rd.Nphoton, rd.seed = float(f), float(phase0), int(nphoton), int(seed) & (2**64 - 1)
    rd.precision, rd.tally = int(precision), int(tally)
    rd.block_size, rd.blocks_per_cu, rd.replicas = int(block_size), int(blocks_per_cu), int(replicas)
    rd.refill_min, rd.stream_tag, rd.lds_slots = int(refill_min), int(stream_tag), int(lds_slots)
    rd.direct_flush, rd.chunk, rd.cache_epoch = int(direct_flush), int(chunk), int(cache_epoch)
    rd.first_photon = int(first_photon)
    if weight0 is not None or chance is not None or c0 is not None:
        rd.model = 1
        rd.weight0 = WEIGHT0 if weight0 is None else float(weight0)
        rd.chance = CHANCE if chance is None else float(chance)
        rd.c0 = C0 if c0 is None else float(c0)
    return rd


class VariantDesc(C.Structure):
    _fields_ = [("mua", C.c_void_p), ("mus", C.c_void_p), ("g", C.c_void_p), ("n", C.c_void_p),
                ("BCType", C.c_void_p), ("BCLNormal", C.c_void_p), ("BCLightDirectionType", C.c_void_p),
                ("BCIntensity", C.c_void_p), ("GaussianSigma", C.c_void_p)]


class BatchItem(C.Structure):
    _fields_ = [("prob", C.c_void_p), ("run", RunDesc), ("out", Output), ("loss", C.c_int64),
                ("status", C.c_int32), ("reserved", C.c_int32), ("ledger", C.c_double * 5)]


class Problem:
    """A mesh + medium + boundary set resident on the device
    (vmc_problem_create ... vmc_problem_destroy)."""

    def __init__(self, p):
        """`p` has the fields of ``meshgen.Problem`` (row-per-element arrays)."""
        self.lib = load_library()
        self.dim, self.Ne, self.Nb, self.Nr = int(p.dim), int(p.Ne), int(p.Nb), int(p.Nr)
        if self.dim not in (2, 3):
            raise ValueError(f"dim must be 2 or 3, not {self.dim}")
        Ne, Nb, Nr, nv = self.Ne, self.Nb, self.Nr, self.dim + 1
        keep = []

        def ptr(a, dtype, n, name):
            if a is None:
                return None
            arr = _sized(a, dtype, n, name)
            keep.append(arr)
            return arr.ctypes.data

        d = ProblemDesc()
        d.dim, d.Ne, d.Nb, d.Nr = self.dim, Ne, Nb, Nr
        d.H, d.HN = ptr(p.H, np.int64, Ne * nv, "H"), ptr(p.HN, np.int64, Ne * nv, "HN")
        d.BH, d.r = ptr(p.BH, np.int64, Nb * self.dim, "BH"), ptr(p.r, np.float64, Nr * self.dim, "r")
        d.mua, d.mus = ptr(p.mua, np.float64, Ne, "mua"), ptr(p.mus, np.float64, Ne, "mus")
        d.g, d.n = ptr(p.g, np.float64, Ne, "g"), ptr(p.n, np.float64, Ne, "n")
        d.BCType, d.BCn = ptr(p.BCType, np.uint8, Nb, "BCType"), ptr(p.BCn, np.float64, Nb, "BCn")
        d.BCLNormal = ptr(getattr(p, "BCLNormal", None), np.float64, Nb * self.dim, "BCLNormal")
        d.BCLightDirectionType = ptr(getattr(p, "BCLDirType", None), np.uint8, Nb, "BCLightDirectionType")
        d.BCIntensity = ptr(getattr(p, "BCIntensity", None), np.float64, Nb, "BCIntensity")
        d.GaussianSigma = ptr(getattr(p, "GaussianSigma", None), np.float64, Nb, "GaussianSigma")
        h = C.c_void_p()
        st = self.lib.vmc_problem_create(C.byref(d), C.byref(h))
        if st:
            raise VmcError(st, self.lib)
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self.lib.vmc_problem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def variant(self, mua=None, mus=None, g=None, n=None, BCType=None, BCLNormal=None, BCLDirType=None,
                BCIntensity=None, GaussianSigma=None) -> "Problem":
        """vmc_problem_create_variant: same mesh, other light sources and/or
        optical parameters (None keeps this problem's)."""
        keep = []
        Ne, Nb = self.Ne, self.Nb

        def ptr(a, dtype, n, name):
            if a is None:
                return None
            arr = _sized(a, dtype, n, name)
            keep.append(arr)
            return arr.ctypes.data

        d = VariantDesc(ptr(mua, np.float64, Ne, "mua"), ptr(mus, np.float64, Ne, "mus"),
                        ptr(g, np.float64, Ne, "g"), ptr(n, np.float64, Ne, "n"),
                        ptr(BCType, np.uint8, Nb, "BCType"), ptr(BCLNormal, np.float64, Nb * self.dim, "BCLNormal"),
                        ptr(BCLDirType, np.uint8, Nb, "BCLightDirectionType"),
                        ptr(BCIntensity, np.float64, Nb, "BCIntensity"),
                        ptr(GaussianSigma, np.float64, Nb, "GaussianSigma"))
        h = C.c_void_p()
        st = self.lib.vmc_problem_create_variant(self._h, C.byref(d), C.byref(h))
        if st:
            raise VmcError(st, self.lib)
        v = Problem.__new__(Problem)
        v.lib, v.dim, v.Ne, v.Nb, v.Nr = self.lib, self.dim, self.Ne, self.Nb, self.Nr
        v._h = h
        return v

    def _outputs(self):
        out = dict(ER=np.zeros(self.Ne), EI=np.zeros(self.Ne), EBR=np.zeros(self.Nb),
                   EBI=np.zeros(self.Nb), DEBR=np.zeros(self.Nb), DEBI=np.zeros(self.Nb))
        o = Output(*[out[k].ctypes.data for k in ("ER", "EI", "EBR", "EBI", "DEBR", "DEBI")])
        return out, o

    def hn(self) -> np.ndarray:
        """vmc_problem_get_hn: the neighbour table Init() built, [Ne, dim+1]."""
        nv = self.dim + 1
        out = np.zeros(self.Ne * nv, dtype=np.int64)
        st = self.lib.vmc_problem_get_hn(self._h, out.ctypes.data)
        if st:
            raise VmcError(st, self.lib)
        return out.reshape(nv, self.Ne).T.copy()

    @property
    def slots(self) -> int:
        """vmc_problem_slots: device element slots (padding included) that
        size the element records and the tally replicas."""
        n = C.c_int64(0)
        st = self.lib.vmc_problem_slots(self._h, C.byref(n))
        if st:
            raise VmcError(st, self.lib)
        return int(n.value)

    def set_medium(self, mua=None, mus=None, g=None, n=None):
        """vmc_problem_set_medium: new optical parameters per element (None
        keeps the current ones) on the resident mesh."""
        arrs = [None if a is None else _sized(a, np.float64, self.Ne, name)
                for a, name in ((mua, "mua"), (mus, "mus"), (g, "g"), (n, "n"))]
        st = self.lib.vmc_problem_set_medium(self._h, *[None if a is None else a.ctypes.data for a in arrs])
        if st:
            raise VmcError(st, self.lib)

    def run(self, nphoton: int, seed: int = 5489, f: float = 0.0, phase0: float = 0.0,
            precision: int = FP64, tally: int = TALLY_FP64, block_size: int = 0,
            blocks_per_cu: int = 0, replicas: int = 0, refill_min: int = 0, stream_tag: int = 0,
            lds_slots: int = 0, direct_flush: int = 0, chunk: int = 0, cache_epoch: int = 0,
            first_photon: int = 0, weight0: Optional[float] = None, chance: Optional[float] = None,
            c0: Optional[float] = None,
            progress: Optional[Callable[[float], bool]] = None, check: bool = True,
            allocate: bool = True) -> RunResult:
        """vmc_problem_run, MonteCarlo() of the reference (MC3D.hpp:1641-1822):
        transport ``nphoton`` photons (stream ids first_photon ..
        first_photon + nphoton - 1 of (seed, stream_tag)) and return the
        normalised ER/EI/EBR/EBI/DEBR/DEBI in the caller's element order with
        the run's statistics.  ``precision`` FP64 or FP32 (FP32 transport,
        FP64 tallies); ``tally`` TALLY_FP64 or TALLY_FIXED64 (bit-reproducible);
        ``weight0``, ``chance``, ``c0`` are the model constants (None: the
        reference's 0.001, 0.1, 2.99792458e11); the remaining keywords are
        the launch knobs of vmc_run_desc (0 = default), which never change
        fixed-point results.  ``progress(percent)``
        returning False aborts (status STATUS_ABORTED, outputs normalised by
        the photons planned, as the reference does); ``check`` raises VmcError
        on any other non-zero status.  ``allocate=False`` calls
        vmc_problem_run_noalloc (the library's static device scratch instead
        of device allocations; needs an earlier run of the same precision)."""
        rd = make_run_desc(nphoton, seed=seed, f=f, phase0=phase0, precision=precision, tally=tally,
                           block_size=block_size, blocks_per_cu=blocks_per_cu, replicas=replicas,
                           refill_min=refill_min, stream_tag=stream_tag, lds_slots=lds_slots,
                           direct_flush=direct_flush, chunk=chunk, cache_epoch=cache_epoch,
                           first_photon=first_photon, weight0=weight0, chance=chance, c0=c0)
        out, o = self._outputs()
        s = Stats()
        cb = None
        if progress is not None:
            cb = PROGRESS_FN(lambda pct, user: 1 if progress(pct) else 0)
        fn = self.lib.vmc_problem_run if allocate else self.lib.vmc_problem_run_noalloc
        st = fn(self._h, C.byref(rd), C.byref(o), C.byref(s), C.cast(cb, C.c_void_p) if cb else None, None)
        if st and (check and st != STATUS_ABORTED):
            raise VmcError(st, self.lib)
        return RunResult(status=st, loss=int(s.loss), steps=int(s.steps), kernel_ms=float(s.kernel_ms),
                         total_ms=float(s.total_ms), ledger=np.array(list(s.ledger)),
                         photons=int(s.photons), grid_blocks=int(s.grid_blocks),
                         block_size=int(s.block_size), rechecks=int(s.rechecks),
                         lds_mesh=bool(s.lds_mesh), **out)

    def autotune(self, nphoton: int, trial_photons: int = 0, **run_kwargs) -> dict:
        """vmc_problem_autotune: the launch shape and tally strategy that run
        ``run(nphoton, **run_kwargs)`` fastest, as keyword arguments for
        ``run`` (block_size, refill_min, chunk, replicas, lds_slots, cache_epoch), plus
        ``trials``, the number of timed trial runs."""
        rd = make_run_desc(nphoton, **run_kwargs)
        tuned = RunDesc()
        n = C.c_int32(0)
        st = self.lib.vmc_problem_autotune(self._h, C.byref(rd), C.byref(tuned), C.c_int64(int(trial_photons)),
                                           C.byref(n))
        if st:
            raise VmcError(st, self.lib)
        return dict(block_size=tuned.block_size, refill_min=tuned.refill_min, chunk=tuned.chunk,
                    replicas=tuned.replicas, lds_slots=tuned.lds_slots, cache_epoch=tuned.cache_epoch,
                trials=int(n.value))


def _batch_array(items, common):
    arr = (BatchItem * len(items))()
    outs = []
    for k, (prob, kw) in enumerate(items):
        kw = {**common, **kw}
        n = kw.pop("nphoton")
        arr[k].prob = prob._h
        arr[k].run = make_run_desc(n, **kw)
        out, o = prob._outputs()
        arr[k].out = o
        outs.append(out)
    return arr, outs


def _batch_results(arr, outs, s):
    return [RunResult(status=int(arr[k].status), loss=int(arr[k].loss), steps=int(s.steps),
                      kernel_ms=float(s.kernel_ms), total_ms=float(s.total_ms),
                      ledger=np.array(list(arr[k].ledger)), photons=int(arr[k].run.Nphoton),
                      grid_blocks=int(s.grid_blocks), block_size=int(s.block_size),
                      rechecks=int(s.rechecks), **out) for k, out in enumerate(outs)]


def batch_run(items, **common):
    """vmc_batch_run: ``items`` is a list of ``(Problem, run_kwargs)``; keyword
    arguments in ``common`` apply to every item (precision, tally, tuning).
    Returns one RunResult per item (``steps`` and ``kernel_ms`` are the
    batch's; ``ledger`` and ``loss`` are the item's own) and the batch's
    vmc_stats, whose ledger is the sum of the items'."""
    lib = load_library()
    arr, outs = _batch_array(items, common)
    s = Stats()
    st = lib.vmc_batch_run(arr, len(items), C.byref(s))
    if st:
        raise VmcError(st, lib)
    return _batch_results(arr, outs, s), s


def batch_autotune(items, trial_fraction: float = 0.0, **common) -> dict:
    """vmc_batch_autotune: launch knobs for ``batch_run(items, **common)``
    timed on the same batch with ``trial_fraction`` of its photons (0 = 2 %);
    returns keyword arguments (block_size, refill_min, chunk, replicas,
    lds_slots, cache_epoch) to add to ``common``, plus ``trials``."""
    lib = load_library()
    arr, _ = _batch_array(items, common)
    tuned = RunDesc()
    n = C.c_int32(0)
    st = lib.vmc_batch_autotune(arr, len(items), C.c_double(float(trial_fraction)), C.byref(tuned), C.byref(n))
    if st:
        raise VmcError(st, lib)
    return dict(block_size=tuned.block_size, refill_min=tuned.refill_min, chunk=tuned.chunk,
                replicas=tuned.replicas, lds_slots=tuned.lds_slots, cache_epoch=tuned.cache_epoch,
                trials=int(n.value))


class BatchPlan:
    """vmc_batch_plan_*: a batch recorded once as a HIP graph and replayed.
    ``run(items)`` replays it for ``items`` (default: the ones it was built
    from), which must name the same problems and photon counts in the same
    order and may change seeds, stream tags and phases; media changed with
    ``Problem.set_medium`` are picked up.  Returns what ``batch_run``
    returns."""

    def __init__(self, items, **common):
        self.lib = load_library()
        self._items, self._common = list(items), dict(common)
        arr, _ = _batch_array(self._items, self._common)
        h = C.c_void_p()
        st = self.lib.vmc_batch_plan_create(arr, len(self._items), C.byref(h))
        if st:
            raise VmcError(st, self.lib)
        self._h = h

    def run(self, items=None):
        """vmc_batch_plan_run: replay the recorded batch, with the items'
        current seeds and media (or new (problem, run keywords) items of the
        same shapes); returns (results, stats)."""
        items = self._items if items is None else list(items)
        arr, outs = _batch_array(items, self._common)
        s = Stats()
        st = self.lib.vmc_batch_plan_run(self._h, arr, len(items), C.byref(s))
        if st:
            raise VmcError(st, self.lib)
        return _batch_results(arr, outs, s), s

    def close(self):
        if getattr(self, "_h", None):
            self.lib.vmc_batch_plan_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServerDesc(C.Structure):
    _fields_ = [("precision", C.c_int32), ("tally", C.c_int32), ("arena_bytes", C.c_int64),
                ("launch_photons", C.c_int64), ("min_slice", C.c_int64), ("policy", C.c_int32),
                ("reserved", C.c_int32 * 7)]


class Server:
    """vmc_server_*: continuous batching of forward solves on resident
    problems.  ``submit`` returns a ticket; ``wait`` returns its RunResult
    (``kernel_ms`` = launches it took part in, ``total_ms`` = submit to
    completion)."""

    def __init__(self, precision: int = FP32, tally: int = TALLY_FP64, arena_bytes: int = 0,
                 launch_photons: int = 0, min_slice: int = 0, policy: int = SERVE_SHORTEST):
        self.lib = load_library()
        d = ServerDesc(int(precision), int(tally), int(arena_bytes), int(launch_photons), int(min_slice),
                       int(policy))
        h = C.c_void_p()
        st = self.lib.vmc_server_create(C.byref(d), C.byref(h))
        if st:
            raise VmcError(st, self.lib)
        self._h = h
        self._pending = {}

    def submit(self, prob: "Problem", nphoton: int, **kw) -> int:
        """vmc_server_submit: queue a solve (Problem.run's keywords); returns
        its ticket."""
        rd = make_run_desc(nphoton, **kw)
        out, o = prob._outputs()
        t = C.c_uint64()
        st = self.lib.vmc_server_submit(self._h, prob._h, C.byref(rd), C.byref(o), C.byref(t))
        if st:
            raise VmcError(st, self.lib)
        self._pending[t.value] = (out, o, int(nphoton))
        return t.value

    def done(self, ticket: int) -> bool:
        """vmc_server_poll: whether the ticket's outputs are written."""
        d = C.c_int32()
        st = self.lib.vmc_server_poll(self._h, C.c_uint64(ticket), C.byref(d))
        if st:
            raise VmcError(st, self.lib)
        return bool(d.value)

    def wait(self, ticket: int) -> RunResult:
        """vmc_server_wait: block until the ticket is done; its RunResult."""
        s = Stats()
        st = self.lib.vmc_server_wait(self._h, C.c_uint64(ticket), C.byref(s))
        out, _, n = self._pending.pop(ticket)
        if st:
            raise VmcError(st, self.lib)
        return RunResult(status=st, loss=int(s.loss), steps=0, kernel_ms=float(s.kernel_ms),
                         total_ms=float(s.total_ms), ledger=np.array(list(s.ledger)), photons=n, grid_blocks=0,
                         block_size=0, **out)

    def counters(self):
        """vmc_server_counters: launches, photons, busy_ms, waiting, active."""
        l, p, b = C.c_int64(), C.c_int64(), C.c_double()
        w, a = C.c_int32(), C.c_int32()
        self.lib.vmc_server_counters(self._h, C.byref(l), C.byref(p), C.byref(b), C.byref(w), C.byref(a))
        return dict(launches=l.value, photons=p.value, busy_ms=b.value, waiting=w.value, active=a.value)

    def close(self):
        if getattr(self, "_h", None):
            self.lib.vmc_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _EngineBase:
    """Public-member mirror of MC3D/MC2D (MC3D.hpp:107-172)."""

    dim = 3

    def __init__(self):
        self.H = self.HN = self.BH = self.r = None
        self.mua = self.mus = self.g = self.n = None
        self.BCType = self.BCLNormal = self.BCn = self.BCIntensity = None
        self.BCLightDirectionType = None
        self.GaussianSigma = None
        self.f = 0.0
        self.phase0 = 0.0
        self.Nphoton = 1
        self.seed = 5489
        self.loss = 0
        self.weight0, self.chance, self.c0 = 0.001, 0.1, 2.99792458e11
        self.precision = FP64
        self.tally = TALLY_FP64
        self.ER = self.EI = self.EBR = self.EBI = self.DEBR = self.DEBI = None
        self._problem: Optional[Problem] = None
        self.stats: Optional[RunResult] = None

    def _as_problem(self):
        from types import SimpleNamespace
        H = np.asarray(self.H)
        BCT = self.BCType
        if isinstance(BCT, (str, bytes)):
            BCT = np.frombuffer(BCT.encode() if isinstance(BCT, str) else BCT, dtype=np.uint8)
        ldt = self.BCLightDirectionType
        if isinstance(ldt, (str, bytes)):
            ldt = np.frombuffer(ldt.encode() if isinstance(ldt, str) else ldt, dtype=np.uint8)
        return SimpleNamespace(dim=self.dim, Ne=H.shape[0], Nb=np.asarray(self.BH).shape[0],
                               Nr=np.asarray(self.r).shape[0], H=H, HN=self.HN, BH=self.BH, r=self.r,
                               mua=self.mua, mus=self.mus, g=self.g, n=self.n, BCType=BCT,
                               BCn=self.BCn, BCLNormal=self.BCLNormal, BCLDirType=ldt,
                               BCIntensity=self.BCIntensity, GaussianSigma=self.GaussianSigma)

    def ErrorChecks(self):
        """Validation happens with Init() on the device side; kept for API parity."""
        return None

    def Init(self):
        """ErrorChecks() + Init(): the members become a resident problem."""
        if self._problem is not None:
            self._problem.close()
        self._problem = Problem(self._as_problem())
        self.HN = self._problem.hn()

    def MonteCarlo(self, progress: Optional[Callable[[float], bool]] = None, finalchecks=None):
        """MonteCarlo(progress, finalchecks): fills ER/EI/EBR/EBI/DEBR/DEBI and
        loss as the reference's members."""
        if self._problem is None:
            self.Init()
        res = self._problem.run(self.Nphoton, seed=self.seed, f=self.f, phase0=self.phase0,
                                precision=self.precision, tally=self.tally, weight0=self.weight0,
                                chance=self.chance, c0=self.c0, progress=progress, check=False)
        if res.status not in (0, STATUS_ABORTED):
            raise VmcError(res.status, self._problem.lib)
        for k in ("ER", "EI", "EBR", "EBI", "DEBR", "DEBI"):
            setattr(self, k, getattr(res, k))
        self.loss = res.loss
        self.stats = res
        if finalchecks is not None:
            finalchecks(int(self.Nphoton) if res.status == 0 else -1, int(self.Nphoton))
        return res


class MC3D(_EngineBase):
    """The reference's class MC3D (cpp/3d/MC3D.hpp:108-172) over the C ABI."""
    dim = 3


class MC2D(_EngineBase):
    """The reference's class MC2D (cpp/2d/MC2D.hpp:113-180) over the C ABI."""
    dim = 2
