"""The oracle's volume jitter (volumeIntegration.h:83,120) against closed
forms.  A jittered segment starts at t0 + ss*u1 + step*u2 (ss the field's
step min(spacing)/2, step = ss / volumeSamplingRate, u1 and u2 uniform), so
the number of samples a slab of length L gets is N = floor((L - off)/step) + 1
(0 when off > L), and a constant slab of LUT opacity a gives
O = 1 - (1 - a)^(N step / unitDistance).

* Expectation: the frame's mean opacity over many pixels (independent PCG32
  sequences) equals E_u1,u2[O] computed here by quadrature, within 4 sigma.
* Limit: as the sampling step shrinks (volumeSamplingRate grows) the
  jitter-averaged opacity tends to the continuous emission-absorption
  integral over the slab shortened by the first jitter, which is of the
  field's own step whatever the rate: E_u[1 - exp(-sigma (L - ss u))]
  = 1 - exp(-sigma L) (exp(sigma ss) - 1) / (sigma ss), sigma =
  -ln(1 - a) / unitDistance; with the first jitter dropped it would be the
  textbook 1 - exp(-sigma L), which the limit misses by about sigma ss / 2.
* Jitter off is unchanged, and jitter on is deterministic for a given pixel
  and frame.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from visrtx_amd import scenes as S

W, H = 64, 64


def _slab(n=16, spacing=1.0 / 15.0, alpha=0.05, unit=1.0 / 15.0):
    """A constant field over [0,1]^3 (the LUT's alpha constant)."""
    data = np.full((n, n, n), 0.5, np.float32)
    vol = S.VolumeDesc(data=data, origin=(0.0, 0.0, 0.0), spacing=(spacing,) * 3,
                       color=np.array([[1.0, 1.0, 1.0], [1.0, 1.0, 1.0]], np.float32),
                       opacity=np.array([alpha, alpha], np.float32), value_range=(0.0, 1.0),
                       unit_distance=unit, id=3)
    return S.SceneDesc(volumes=[vol], world_id=0), vol


# a camera straight down the z axis with a 2-degree field of view: every ray
# crosses the slab 0 <= z <= 1 with length 1/|dz| (within 2e-4 of 1)
CAM = S.CameraDesc((0.5, 0.5, -2.0), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0), math.radians(2.0), 1.0)


def _lengths():
    o, _ = orc.camera_ray(CAM, W, H, 0, 0)
    L = np.zeros((H, W))
    for y in range(H):
        for x in range(W):
            _, d = orc.camera_ray(CAM, W, H, x, y)
            L[y, x] = 1.0 / abs(float(d[2]))
    return L


def _expected(L, step, ss, alpha, unit, q=400):
    """E over u1, u2 ~ U[0,1) of 1 - (1-a)^(N step/unit), midpoint quadrature."""
    u = (np.arange(q) + 0.5) / q
    off = ss * u[:, None] + step * u[None, :]
    e = np.zeros_like(L)
    for idx, l in np.ndenumerate(L):
        N = np.where(off <= l, np.floor((l - off) / step) + 1.0, 0.0)
        e[idx] = np.mean(1.0 - (1.0 - alpha) ** (N * step / unit))
    return e


def _render(sc, **rp):
    return orc.OracleScene(sc).render(CAM, W, H, background=(0.0, 0.0, 0.0, 0.0), **rp)


# sampling rates for which the slab (30 field steps long) is not a whole
# number of sampling steps, so no sample sits on the exit
@pytest.mark.parametrize("rate", [1.15, 0.45])
def test_jitter_mean_equals_the_expectation_over_offsets(rate):
    alpha, unit, sp = 0.05, 1.0 / 15.0, 1.0 / 15.0
    sc, vol = _slab(alpha=alpha, unit=unit, spacing=sp)
    o = _render(sc, volumeSamplingRate=rate, volumeJitter=True)
    op = o["color"][..., 3].astype(np.float64)
    ss = float(np.float32(sp / 2))
    step = float(np.float32(ss * float(np.float32(1.0 / rate))))
    L = _lengths()
    e = _expected(L, step, ss, alpha, unit)
    # the per-pixel values take a few discrete levels (N varies by one or
    # two); their spread sets the tolerance of the mean
    sigma = op.std() / math.sqrt(op.size)
    assert abs(op.mean() - e.mean()) <= 4 * sigma + 1e-6, (op.mean(), e.mean(), sigma)
    assert len(np.unique(op)) >= 2
    # without jitter every ray starts on the entry: the most samples
    off = _render(sc, volumeSamplingRate=rate)["color"][..., 3]
    n0 = np.floor(L / step) + 1
    np.testing.assert_allclose(off, 1 - (1 - alpha) ** (n0 * step / unit), rtol=2e-6)
    assert op.mean() < off.mean()          # jitter drops samples near the entry


def test_jitter_limit_is_the_shortened_continuous_integral():
    alpha, unit, sp = 0.02, 1.0 / 15.0, 1.0 / 15.0
    sc, vol = _slab(alpha=alpha, unit=unit, spacing=sp)
    sigma = -math.log(1 - alpha) / unit
    ss = float(np.float32(sp / 2))
    L = _lengths()
    limit = np.mean(1 - np.exp(-sigma * L) * (math.exp(sigma * ss) - 1) / (sigma * ss))
    textbook = np.mean(1 - np.exp(-sigma * L))
    errs = []
    for rate in (1.15, 4.15, 10.0):
        op = _render(sc, volumeSamplingRate=rate, volumeJitter=True)["color"][..., 3]
        step = ss / rate
        errs.append(abs(op.mean() - limit))
        # what is left is the discretisation of the remaining length: at most
        # one sampling step of optical depth
        assert errs[-1] <= sigma * step + 1e-3, (rate, op.mean(), limit)
    assert errs[-1] < errs[0]
    assert abs(limit - textbook) > 10 * errs[-1]   # the shortening is real


def test_jitter_is_deterministic_and_keyed_by_pixel():
    sc, _ = _slab()
    a = _render(sc, volumeSamplingRate=1.0, volumeJitter=True)["color"]
    b = _render(sc, volumeSamplingRate=1.0, volumeJitter=True)["color"]
    np.testing.assert_array_equal(a, b)
    assert len(np.unique(a[..., 3])) > 1   # pixels draw different offsets
