"""Differentiable volume: HIP forward vs the FP64 oracle march, HIP adjoint vs
FP64 central finite differences of the oracle's loss (h = 1e-3), and a short
Adam run through torch.autograd.

The finite-difference check uses a LUT that is linear in every channel (no
interior knots, so no kinks for the differences to straddle) and disables
early termination, as the gradient-check mode requires.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from visrtx_amd import scenes as S

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _setup(n, views, res, lut, early, seed=3, tiled=True, adjoint="suffix"):
    from visrtx_amd import diffvol as D
    dens = S.gen_noise_volume(n, seed=seed)
    sp = 1.0 / (n - 1)
    spec = D.GridSpec(dims=(n, n, n), spacing=(sp, sp, sp), unit_distance=sp, step=0.5 * sp,
                      early_stop=early, image=(res, res) if tiled else None, adjoint=adjoint)
    cams = D.orbit_cameras(views, seed=seed, fovy=45.0)
    org, dirs = D.make_rays(cams, res, res)
    td = torch.from_numpy(dens).cuda()
    tl = torch.from_numpy(lut).cuda()
    return D, spec, td, tl, org, dirs, dens


def test_forward_matches_fp64_oracle(require_gpu):
    lut = orc.tf_discretize(S.COLORMAP5, np.array([0.0, 0.3], np.float32))
    D, spec, td, tl, org, dirs, dens = _setup(32, 3, 48, lut, early=True)
    print(f"[diffvol] forward 32^3, {org.shape[0]} rays", flush=True)
    rgba = D.render(td, tl, org, dirs, spec)
    torch.cuda.synchronize()
    ref, amb = orc.diffvol_forward(spec.dims, spec.origin, spec.spacing, dens, lut,
                                   spec.value_range, spec.unit_distance, spec.step,
                                   org.cpu().numpy(), dirs.cpu().numpy(), early_stop=True)
    g = rgba.cpu().numpy()
    ok = amb == 0
    assert ok.mean() > 0.99
    np.testing.assert_allclose(g[ok], ref[ok], rtol=0, atol=1e-4)
    assert ref[:, 3].max() > 0.5


@pytest.mark.parametrize("adjoint", ["suffix", "replay"])
def test_adjoint_vs_fp64_central_differences(require_gpu, adjoint):
    """1000 parameters (600 voxels, 400 LUT scalars): relative error <= 1e-3,
    for both backward schedules."""
    from visrtx_amd import diffvol as D
    lut = D.linear_lut(alpha_max=0.5)
    D, spec, td, tl, org, dirs, dens = _setup(24, 2, 32, lut, early=False, seed=4, adjoint=adjoint)
    print(f"[diffvol] FD check, adjoint={adjoint}", flush=True)
    n = org.shape[0]
    rng = np.random.default_rng(0)
    w = rng.normal(size=(n, 4))
    td.requires_grad_(True)
    tl.requires_grad_(True)
    rgba = D.render(td, tl, org, dirs, spec)
    (rgba * torch.from_numpy(w).float().cuda()).sum().backward()
    torch.cuda.synchronize()
    gd = td.grad.cpu().numpy().reshape(-1)
    gl = tl.grad.cpu().numpy().reshape(-1)
    # the same gradients with rays treated as unstructured (no tile mapping)
    td.grad = None
    tl.grad = None
    spec_u = D.GridSpec(**{**spec.__dict__, "image": None})
    rgba_u = D.render(td, tl, org, dirs, spec_u)
    (rgba_u * torch.from_numpy(w).float().cuda()).sum().backward()
    torch.cuda.synchronize()
    np.testing.assert_allclose(td.grad.cpu().numpy().reshape(-1), gd, rtol=1e-4,
                               atol=1e-6 * np.abs(gd).max())
    np.testing.assert_allclose(tl.grad.cpu().numpy().reshape(-1), gl, rtol=1e-4,
                               atol=1e-6 * np.abs(gl).max())
    o, d = org.cpu().numpy(), dirs.cpu().numpy()
    args = (spec.dims, spec.origin, spec.spacing)
    dd = dens.astype(np.float64).reshape(-1).copy()
    ll = lut.astype(np.float64).reshape(-1).copy()
    h = 1e-3

    def loss(den, lt):
        return orc.diffvol_loss(*args, den, lt, spec.value_range, spec.unit_distance, spec.step,
                                o, d, w, early_stop=False)

    # voxels that rays actually sample (others have zero gradient on both sides)
    vox = rng.choice(np.nonzero(np.abs(gd) > 0)[0], 600, replace=False)
    lutp = rng.choice(np.arange(4, 1020), 400, replace=False)   # avoid the clamped end entries
    fd, gg = [], []
    for i in vox:
        dp = dd.copy(); dp[i] += h
        dm = dd.copy(); dm[i] -= h
        fd.append((loss(dp, ll) - loss(dm, ll)) / (2 * h))
        gg.append(gd[i])
    for i in lutp:
        lp = ll.copy(); lp[i] += h
        lm = ll.copy(); lm[i] -= h
        fd.append((loss(dd, lp) - loss(dd, lm)) / (2 * h))
        gg.append(gl[i])
    fd, gg = np.array(fd), np.array(gg)
    norm_rel = np.linalg.norm(gg - fd) / np.linalg.norm(fd)
    max_rel = np.abs(gg - fd).max() / np.abs(fd).max()
    print(f"[diffvol] FD check: normwise rel err {norm_rel:.2e}, max err / max|fd| {max_rel:.2e}")
    assert norm_rel <= 1e-3
    assert max_rel <= 1e-3


def test_adam_reduces_loss(require_gpu):
    from visrtx_amd import diffvol as D
    gt_lut = orc.tf_discretize(S.COLORMAP5, np.array([0.0, 0.3], np.float32))
    D, spec, td_gt, tl_gt, org, dirs, dens = _setup(32, 4, 32, gt_lut, early=True, seed=6)
    with torch.no_grad():
        target = D.render(td_gt, tl_gt, org, dirs, spec)
    dens0 = torch.full_like(td_gt, 0.5).requires_grad_(True)
    lut0 = torch.from_numpy(D.linear_lut(0.2)).cuda().requires_grad_(True)
    opt = torch.optim.Adam([dens0, lut0], lr=1e-2)
    losses = []
    for _ in range(30):
        opt.zero_grad()
        loss = ((D.render(dens0, lut0, org, dirs, spec) - target) ** 2).mean()
        loss.backward()
        opt.step()
        with torch.no_grad():
            lut0.clamp_(0.0, 0.999)
        losses.append(loss.item())
    print(f"[diffvol] loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    assert losses[-1] < 0.5 * losses[0]


def _grads(D, spec, td, tl, org, dirs, w):
    td = td.detach().clone().requires_grad_(True)
    tl = tl.detach().clone().requires_grad_(True)
    rgba = D.render(td, tl, org, dirs, spec)
    (rgba * w).sum().backward()
    torch.cuda.synchronize()
    return td.grad.cpu().numpy().reshape(-1), tl.grad.cpu().numpy().reshape(-1)


def _close(a, b, tol):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30) <= tol


def test_suffix_and_replay_adjoints_agree(require_gpu):
    """Both schedules give the same gradients (early termination on, a 40x40
    image that leaves partial workgroup tiles, rays that miss the grid), and
    the tiled and untiled ray layouts agree."""
    from visrtx_amd import diffvol as D
    lut = orc.tf_discretize(S.COLORMAP5, np.array([0.0, 0.6], np.float32))
    D, spec, td, tl, org, dirs, dens = _setup(32, 3, 40, lut, early=True, seed=8)
    print(f"[diffvol] suffix vs replay, {org.shape[0]} rays", flush=True)
    w = torch.from_numpy(np.random.default_rng(1).normal(size=(org.shape[0], 4))).float().cuda()
    gd_s, gl_s = _grads(D, spec, td, tl, org, dirs, w)
    gd_r, gl_r = _grads(D, D.GridSpec(**{**spec.__dict__, "adjoint": "replay"}), td, tl, org, dirs, w)
    gd_u, gl_u = _grads(D, D.GridSpec(**{**spec.__dict__, "image": None}), td, tl, org, dirs, w)
    assert np.abs(gd_s).max() > 0 and np.abs(gl_s).max() > 0
    assert _close(gd_s, gd_r, 1e-5) and _close(gl_s, gl_r, 1e-5)
    assert _close(gd_u, gd_s, 1e-6) and _close(gl_u, gl_s, 1e-6)


def test_suffix_adjoint_fixed_point_scale(require_gpu):
    """The suffix schedule sums gradients in 64-bit fixed point whose scale
    follows the workgroup's output gradients: scaling the loss by 1e-12 or
    1e+12 scales the gradients by the same factor, with no loss of digits."""
    from visrtx_amd import diffvol as D
    lut = D.linear_lut(alpha_max=0.5)
    D, spec, td, tl, org, dirs, dens = _setup(24, 2, 32, lut, early=True, seed=9)
    w = torch.from_numpy(np.random.default_rng(2).normal(size=(org.shape[0], 4))).float().cuda()
    gd, gl = _grads(D, spec, td, tl, org, dirs, w)
    for f in (1e-12, 1e12):
        print(f"[diffvol] fixed-point scale, loss x {f:g}", flush=True)
        gdf, glf = _grads(D, spec, td, tl, org, dirs, w * f)
        assert _close(gdf / f, gd, 1e-6) and _close(glf / f, gl, 1e-6)
