"""The oracle's macro-cell majorant grid and Woodcock distance sampling
(UniformGrid.cu:44-258 layout; volumeIntegration.h:167-228, dda.h:43-122).

The grid is pinned by an independent numpy restatement of its rule; Woodcock
tracking by its expectation: the probability that a ray crosses a
heterogeneous slab without a real collision is exp(-sum(alpha_i * L_i) /
step), whatever the (larger) majorants of the cells it crosses."""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from visrtx_amd import scenes as S


def grid_reference(data, lut, vr):
    """numpy restatement of the conservative majorant grid: per 16^3-voxel
    cell, min / max over its voxels plus a one-voxel halo (NaN ignored), then
    the largest LUT alpha over the entries that range classifies to (the
    fp32 clamp / scale / floor of classification)."""
    nz, ny, nx = data.shape
    m = [(n + 15) // 16 for n in (nx, ny, nz)]
    out = np.zeros((m[2], m[1], m[0]), np.float32)
    f = np.float32
    lo_v, hi_v = f(vr[0]), f(vr[1])
    inv = f(1.0) / f(hi_v - lo_v)

    def entry(v):
        cl = min(max(f(v), lo_v), hi_v)
        x = min(max(f(f(f(cl - lo_v) * inv) * f(256.0)) - f(0.5), f(0.0)), f(255.0))
        return min(int(x), 254)

    for cz in range(m[2]):
        for cy in range(m[1]):
            for cx in range(m[0]):
                blk = data[max(16 * cz - 1, 0):min(16 * cz + 17, nz - 1) + 1,
                           max(16 * cy - 1, 0):min(16 * cy + 17, ny - 1) + 1,
                           max(16 * cx - 1, 0):min(16 * cx + 17, nx - 1) + 1]
                fin = blk[~np.isnan(blk)]
                if fin.size == 0:
                    continue
                e0, e1 = entry(fin.min()), entry(fin.max())
                i0, i1 = min(e0, e1), max(e0, e1) + 1
                out[cz, cy, cx] = lut.reshape(-1, 4)[i0:i1 + 1, 3].max()
    return out


def test_oracle_majorant_grid_matches_restatement():
    rng = np.random.default_rng(3)
    data = rng.random((40, 33, 50), dtype=np.float32)
    data[data < 0.6] = 0.0
    data[:18, :18, :18] = 0.0
    data[tuple(rng.integers(0, s, 5) for s in data.shape)] = np.nan
    vol = S.VolumeDesc(data, color=S.COLORMAP5, opacity=np.array([0.0, 0.0, 0.3, 1.0],
                                                                 np.float32))
    lut = orc.tf_discretize(vol.color, vol.opacity, vol.value_range)
    np.testing.assert_array_equal(orc.majorant_grid(vol), grid_reference(data, lut, (0.0, 1.0)))


def _slab_rays(n, rng, x0=-0.5):
    """Rays along +x through the slab at random (y, z) inside it."""
    org = np.stack([np.full(n, x0), rng.uniform(0.1, 0.9, n), rng.uniform(0.1, 0.9, n)],
                   1).astype(np.float32)
    dirs = np.tile(np.array([1.0, 0.0, 0.0], np.float32), (n, 1))
    tmm = np.tile(np.array([0.0, 1e30], np.float32), (n, 1))
    return org, dirs, tmm


@pytest.mark.parametrize("split", [False, True])
def test_woodcock_transmittance_expectation(split):
    """P(no real collision) = exp(-sum(alpha L) / step).  With a split field the
    cells straddling the split carry the larger alpha as majorant, so null
    collisions must be rejected at the right rate."""
    n = 64
    sp = 1.0 / (n - 1)
    data = np.full((n, n, n), 0.5, np.float32)
    a1, a2 = 0.02, 0.08
    opacity = np.array([a1, a1, a2, a2], np.float32)        # value 0.5 -> a1, 0.9 -> a2
    if split:
        data[:, :, n // 2:] = 0.9
    vol = S.VolumeDesc(data, spacing=(sp, sp, sp), color=S.COLORMAP5, opacity=opacity)
    lut = orc.tf_discretize(vol.color, vol.opacity, vol.value_range).reshape(-1, 4)
    alpha = lambda v: float(np.interp(v * 256 - 0.5, np.arange(256), lut[:, 3]))
    step = sp / 2
    rng = np.random.default_rng(1)
    org, dirs, tmm = _slab_rays(40000, rng)
    t, alb, tr = orc.sample_distance(vol, org, dirs, tmm, seed=5)
    if split:
        half = (n // 2) * sp                                 # voxel n/2 starts the 0.9 region
        # trilinear ramp over one voxel between the two values
        xs = np.linspace(0.0, 1.0, 20001)
        vals = np.where(xs <= half - sp, 0.5, np.where(xs >= half, 0.9,
                        0.5 + 0.4 * (xs - (half - sp)) / sp))
        tau = np.trapezoid([alpha(v) for v in vals], xs) / step
    else:
        tau = alpha(0.5) * 1.0 / step
    p = math.exp(-tau)
    est = float((tr == 1.0).mean())
    sigma = math.sqrt(p * (1 - p) / len(tr))
    assert abs(est - p) < 4 * sigma + 1e-3, (est, p, sigma)
    hit = tr == 0.0
    assert (t[hit] >= 0.5 - 1e-5).all() and (t[hit] <= 1.5 + 1e-5).all()
    assert (t[~hit] == 1e30).all()
    if not split:
        # free paths in the homogeneous slab: exponential with rate
        # sigma = alpha / step, truncated at the slab length 1
        sig, x = tau, t[hit] - 0.5
        mean = 1.0 / sig - math.exp(-sig) / (1.0 - math.exp(-sig))
        var = 1.0 / sig ** 2 - math.exp(-sig) / (1.0 - math.exp(-sig)) ** 2
        assert abs(x.mean() - mean) < 4.0 * math.sqrt(var / len(x)), (x.mean(), mean)


def test_woodcock_skips_empty_cells_and_misses():
    """Rays through zero-majorant cells only, or missing the box, never
    collide and consume no collisions."""
    n = 48
    data = np.zeros((n, n, n), np.float32)
    vol = S.VolumeDesc(data, spacing=(1 / (n - 1),) * 3, color=S.COLORMAP5,
                       opacity=np.array([0.0, 0.0, 0.5, 0.5], np.float32))
    assert (orc.majorant_grid(vol) == 0).all()
    rng = np.random.default_rng(2)
    org, dirs, tmm = _slab_rays(500, rng)
    org[:100, 1] = 5.0                                        # these miss the box
    t, alb, tr = orc.sample_distance(vol, org, dirs, tmm, seed=1)
    assert (tr == 1.0).all() and (t == 1e30).all() and (alb == 0).all()
