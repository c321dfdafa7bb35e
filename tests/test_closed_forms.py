"""Transport against closed forms, independent of the reference's code.

Each case has an answer fixed by the physics, so a mistake shared by the
oracle and the GPU path (a dropped term, a wrong sign or a transposed
Fresnel formula) fails here even though the two would agree with each other.
The same cases run on the oracle (CPU) and on the GPU path (FP64 and FP32).

* Beer-Lambert: a non-scattering slab (mus = 0) with matched refractive
  indices, lit straight down by a direct source over its top face.  Every
  photon crosses the slab once: the exitance through the bottom is exactly
  e^(-mua L) of the launched energy, the absorbed energy 1 - e^(-mua L), and
  nothing leaves through the other faces.  Deterministic, so the bound is
  rounding.
* Frequency domain: the same slab at 5 GHz with phase0 = 0.3.  A photon
  reaches the bottom with weight e^(-mua L) and phase phase0 + k L
  (k = 2 pi f n / c0), so the bottom's EBR + i EBI sums to
  e^(-mua L) e^(-i (phase0 + k L)), and the element fluence ER + i EI,
  integrated over the slab, to e^(-i phase0) (1 - e^(-(mua + i k) L)) /
  (mua + i k); in 2D DEBR + i DEBI equals EBR + i EBI at normal exit.
* Henyey-Greenstein at g = +1 and g = -1 (3D): g = 1 scatters straight
  ahead, so with scattering on the slab still transmits exactly
  e^(-mua L); g = -1 reverses the direction at every event, which is the
  rod model with backscattering rate mus: without absorption the slab
  transmits 1 / (1 + mus L) and reflects the rest (statistical, 5 sigma).
* Conservation: mua = 0 with scattering and a refractive-index jump: no
  absorption, no roulette, so all launched energy leaves through the
  boundary, each photon with weight 1.
* Fresnel: a non-absorbing, non-scattering slab of index 1.5 in air lit at
  normal incidence and at 30 degrees inside the slab.  A photon bounces
  between the two parallel faces, each reflecting with the unpolarised
  Fresnel reflectance R at that angle, so the bottom receives T / (1 - R^2)
  and the top R T / (1 - R^2).  Statistical (binomial), checked to 5 sigma.
  In 2D, DEBR is EBR divided by the exit cosine.
"""
import math

import numpy as np
import pytest

import pyoracle
from valomc_amd import meshgen


def _tri_areas(r, BH):
    t = r[BH]
    if t.shape[2] == 2:   # 2D: edge lengths
        return np.linalg.norm(t[:, 1] - t[:, 0], axis=1)
    return 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)


def _volumes(r, H):
    t = r[H]
    if t.shape[1] == 3:   # 2D: triangle areas
        a, b = t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]
        return 0.5 * np.abs(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0])
    return np.abs(np.einsum("ij,ij->i", t[:, 1] - t[:, 0], np.cross(t[:, 2] - t[:, 0], t[:, 3] - t[:, 0]))) / 6


def slab(dim, thick, half, mua, mus, g, n_in, n_out, src_half=None, direction=None, nphoton=100000, seed=11):
    """A slab of `thick` 1 mm cells, `2 half` cells wide, top face at
    z = thick (y in 2D); a direct ('l') source on the top-face elements
    whose centres lie within src_half of the axis, straight down or along
    `direction` (absolute)."""
    c = np.arange(-half, half) + 0.5
    z = np.arange(thick) + 0.5
    r, H, BH = meshgen.grid_mesh(c, c, z, h=1.0) if dim == 3 else meshgen.grid_mesh(c, z, h=1.0)
    p = meshgen.homogeneous(dim, r, H, BH, mua, mus, g, n_in, n_out)
    up = dim - 1
    b = r[BH]
    top = np.all(b[:, :, up] == thick, axis=1)
    cen = b.mean(axis=1)
    src = top & np.all(np.abs(cen[:, :up]) < (src_half if src_half is not None else half), axis=1)
    p.BCType[src] = ord("l")
    if direction is not None:
        p.BCLDirType = np.full(p.Nb, ord("n"), dtype=np.uint8)
        p.BCLDirType[src] = ord("a")
        p.BCLNormal = np.zeros((p.Nb, dim))
        p.BCLNormal[src] = direction
    p.Nphoton, p.seed = nphoton, seed
    faces = dict(top=top, bottom=np.all(b[:, :, up] == 0, axis=1))
    faces["side"] = ~(faces["top"] | faces["bottom"])
    return p, faces


def exitance(p, res, faces):
    e = res.EBR * _tri_areas(np.asarray(p.r), np.asarray(p.BH))
    return {k: float(e[m].sum()) for k, m in faces.items()}


def absorbed(p, res):
    return float(np.sum(res.ER * np.asarray(p.mua) * _volumes(np.asarray(p.r), np.asarray(p.H))))


def fresnel_unpolarised(n1, n2, cos_i):
    """Reflectance for unpolarised light from n1 into n2 (Fresnel's
    equations; total internal reflection gives 1)."""
    sin_t = n1 / n2 * math.sqrt(max(0.0, 1.0 - cos_i * cos_i))
    if sin_t >= 1.0:
        return 1.0
    cos_t = math.sqrt(1.0 - sin_t * sin_t)
    rs = (n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t)
    rp = (n2 * cos_i - n1 * cos_t) / (n2 * cos_i + n1 * cos_t)
    return 0.5 * (rs * rs + rp * rp)


def test_fresnel_helper_special_values():
    assert fresnel_unpolarised(1.5, 1.0, 1.0) == pytest.approx(0.04)
    assert fresnel_unpolarised(1.0, 1.5, 1.0) == pytest.approx(0.04)
    assert fresnel_unpolarised(1.5, 1.0, math.cos(math.asin(1 / 1.5) + 1e-6)) == 1.0
    # Brewster angle: the p-wave is not reflected
    tb = math.atan(1.5)
    assert fresnel_unpolarised(1.0, 1.5, math.cos(tb)) == pytest.approx(
        0.5 * ((math.cos(tb) - 1.5 * math.cos(math.asin(math.sin(tb) / 1.5)))
               / (math.cos(tb) + 1.5 * math.cos(math.asin(math.sin(tb) / 1.5)))) ** 2)


# ---- the three cases, each run by a given engine -------------------------

def check_beer_lambert(run, dim, tol):
    mua, thick = 0.1, 5
    p, faces = slab(dim, thick, 4, mua, 0.0, 0.9, 1.3, 1.3, nphoton=20000)
    res = run(p)
    ex = exitance(p, res, faces)
    t = math.exp(-mua * thick)
    assert res.loss == 0
    assert abs(ex["bottom"] - t) <= tol, ex
    assert ex["top"] == 0.0 and ex["side"] == 0.0, ex
    assert abs(absorbed(p, res) - (1.0 - t)) <= tol


C0 = 2.99792458e11   # mm/s


def check_fd_beer_lambert(run, dim, tol):
    mua, thick, n = 0.1, 5, 1.3
    p, faces = slab(dim, thick, 4, mua, 0.0, 0.9, n, n, nphoton=20000)
    p.f, p.phase0 = 5e9, 0.3
    res = run(p)
    k = 2.0 * math.pi * p.f / C0 * n
    area = _tri_areas(np.asarray(p.r), np.asarray(p.BH))
    bot = faces["bottom"]
    got = complex(np.sum((res.EBR * area)[bot]), np.sum((res.EBI * area)[bot]))
    want = math.exp(-mua * thick) * complex(math.cos(p.phase0 + k * thick), -math.sin(p.phase0 + k * thick))
    assert res.loss == 0
    assert abs(got - want) <= tol, (got, want)
    vol = _volumes(np.asarray(p.r), np.asarray(p.H))
    flu = complex(np.sum(res.ER * vol), np.sum(res.EI * vol))
    z = complex(mua, k)
    want_flu = complex(math.cos(p.phase0), -math.sin(p.phase0)) * (1.0 - np.exp(-z * thick)) / z
    assert abs(flu - want_flu) <= tol * abs(want_flu), (flu, want_flu)
    if dim == 2:
        d = complex(np.sum((res.DEBR * area)[bot]), np.sum((res.DEBI * area)[bot]))
        assert abs(d - got) <= tol, (d, got)


def check_forward_scattering(run, tol):
    mua, thick = 0.1, 5
    p, faces = slab(3, thick, 4, mua, 1.0, 1.0, 1.3, 1.3, nphoton=20000)
    res = run(p)
    ex = exitance(p, res, faces)
    assert res.loss == 0
    assert abs(ex["bottom"] - math.exp(-mua * thick)) <= tol, ex
    assert ex["top"] == 0.0 and ex["side"] == 0.0, ex


def check_rod_backscatter(run, nphoton):
    mus, thick = 0.4, 5
    p, faces = slab(3, thick, 4, 0.0, mus, -1.0, 1.3, 1.3, nphoton=nphoton)
    res = run(p)
    ex = exitance(p, res, faces)
    t = 1.0 / (1.0 + mus * thick)
    sigma = math.sqrt(t * (1.0 - t) / nphoton)
    assert res.loss == 0
    assert abs(ex["bottom"] - t) < 5 * sigma, (ex, t, sigma)
    assert abs(ex["top"] - (1.0 - t)) < 5 * sigma, (ex, 1.0 - t)
    assert ex["side"] == 0.0, ex


def check_conservation(run, dim, tol):
    p, faces = slab(dim, 4, 3, 0.0, 1.0, 0.8, 1.4, 1.0, src_half=1, nphoton=20000)
    res = run(p)
    ex = exitance(p, res, faces)
    assert res.loss == 0
    assert abs(sum(ex.values()) - 1.0) <= tol, ex
    assert min(ex.values()) > 0.0, ex


def check_fresnel(run, dim, nphoton, cos_i):
    n_in = 1.5
    direction = None
    if cos_i < 1.0:
        s = math.sqrt(1.0 - cos_i * cos_i)
        direction = (s, 0.0, -cos_i) if dim == 3 else (s, -cos_i)
    # one cell thick and wide, so a photon drifting sideways between
    # reflections meets a side face only after many of them
    p, faces = slab(dim, 1, 12, 0.0, 0.0, 0.9, n_in, 1.0, src_half=2, direction=direction, nphoton=nphoton)
    res = run(p)
    ex = exitance(p, res, faces)
    R = fresnel_unpolarised(n_in, 1.0, cos_i)
    T = 1.0 - R
    want_top = R * T / (1.0 - R * R)
    sigma = math.sqrt(want_top * (1.0 - want_top) / nphoton)
    assert res.loss == 0
    assert abs(ex["top"] - want_top) < 5 * sigma, (ex, want_top, sigma)
    assert abs(ex["bottom"] - T / (1.0 - R * R)) < 5 * sigma, (ex, T / (1.0 - R * R))
    assert ex["side"] < 1e-6, ex
    if dim == 2:
        d = res.DEBR * _tri_areas(np.asarray(p.r), np.asarray(p.BH))
        for k in ("top", "bottom"):
            # FP32 directions carry ~1e-7 relative error in the cosine
            assert abs(float(d[faces[k]].sum()) * cos_i - ex[k]) <= 1e-6, k


def _oracle(p):
    return pyoracle.run(p, nthreads=1, rng_mode=pyoracle.RNG_PHILOX)


@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_beer_lambert(dim):
    check_beer_lambert(_oracle, dim, 1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_frequency_domain_slab(dim):
    check_fd_beer_lambert(_oracle, dim, 1e-12)


def test_oracle_forward_scattering():
    check_forward_scattering(_oracle, 1e-12)


def test_oracle_rod_backscatter():
    check_rod_backscatter(_oracle, 100000)


@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_conserves_energy_without_absorption(dim):
    check_conservation(_oracle, dim, 1e-12)


@pytest.mark.parametrize("dim,cos_i", [(3, 1.0), (3, math.cos(math.radians(30))), (2, 1.0),
                                       (2, math.cos(math.radians(30)))])
def test_oracle_fresnel_slab(dim, cos_i):
    check_fresnel(_oracle, dim, 200000, cos_i)


# ---- the GPU path --------------------------------------------------------

@pytest.fixture(scope="module")
def vmc():
    import valomc_amd
    return valomc_amd


def _gpu(vmc, precision):
    def run(p):
        with vmc.Problem(p) as prob:
            return prob.run(p.Nphoton, seed=p.seed, precision=precision, f=p.f, phase0=p.phase0)
    return run


PREC = [("fp64", 1e-12), ("fp32", 2e-6)]


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", PREC)
@pytest.mark.parametrize("dim", [2, 3])
def test_gpu_beer_lambert(vmc, dim, prec, tol):
    check_beer_lambert(_gpu(vmc, getattr(vmc, prec.upper())), dim, tol)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
@pytest.mark.parametrize("dim", [2, 3])
def test_gpu_frequency_domain_slab(vmc, dim, prec, tol):
    check_fd_beer_lambert(_gpu(vmc, getattr(vmc, prec.upper())), dim, tol)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", PREC)
def test_gpu_forward_scattering(vmc, prec, tol):
    check_forward_scattering(_gpu(vmc, getattr(vmc, prec.upper())), tol)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_gpu_rod_backscatter(vmc, prec):
    check_rod_backscatter(_gpu(vmc, getattr(vmc, prec.upper())), 2000000)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", PREC)
@pytest.mark.parametrize("dim", [2, 3])
def test_gpu_conserves_energy_without_absorption(vmc, dim, prec, tol):
    check_conservation(_gpu(vmc, getattr(vmc, prec.upper())), dim, tol)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("dim,cos_i", [(3, 1.0), (3, math.cos(math.radians(30))), (2, 1.0),
                                       (2, math.cos(math.radians(30)))])
def test_gpu_fresnel_slab(vmc, dim, cos_i, prec):
    check_fresnel(_gpu(vmc, getattr(vmc, prec.upper())), dim, 2000000, cos_i)
