"""Parity at the sizes the benchmark measures, one test per benched workload:

* 512^3 volume, 1920x1080, 0.5-voxel step (volumeSamplingRate 1) against the
  FP64 oracle march (volumeIntegration.h:64-165, Raycast_ptx.cu:60-179);
* the batched scene (64 views at 512^2, 256 instances of 16 meshes, 9.4M
  triangles): every view of the autotuned persistent launch equals the
  per-view ANARI frame, two views equal the oracle;
* the differentiable volume at 128^3 with the colormap LUT and 32 views at
  256^2: the hand-written adjoint against FP64 central differences of the
  oracle's loss on 1,000 parameters;
* 10M spheres, AO renderer at 16 spp, 1920x1080 (AmbientOcclusion_ptx.cu:
  86-220) against the oracle.

Bars as in the reduced-size tests: ids exact, depth and FLOAT32 colour
within 1e-4 (volume colour within 1e-4 absolute on pixels the oracle does not
flag), sRGB within one code value, adjoint relative error <= 1e-3.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from visrtx_amd import anari as A
from visrtx_amd import scenes as S

pytestmark = [pytest.mark.gpu, pytest.mark.slow, pytest.mark.timeout(600)]

CHANNELS = ["color", "depth", "primitiveId", "objectId", "instanceId"]


def _render(sc, cam, w, h, kind="raycast", color=A.FLOAT32_VEC4, **rp):
    dev = A.Device()
    s = S.build_anari(dev, sc)
    c = S.make_camera(dev, s, cam)
    r = S.make_renderer(dev, s, kind, **rp)
    f = S.make_frame(dev, s, s.world, c, r, w, h, color=color)
    dev.render(f)
    dev.ready(f)
    out = {ch: dev.map_frame(f, "channel." + ch)[0] for ch in CHANNELS}
    out["volumeSamples"] = dev.get_property(f, "volumeSamples", A.UINT64)
    out["rays"] = dev.get_property(f, "raysTraced", A.UINT64)
    out["errors"] = dev.errors()
    s.release()
    dev.close()
    return out


def volume_tie_bounds(vol, step_over_unit):
    """Largest colour change (any rgb channel of the raycast frame) that a
    flagged pixel's doubtful decision can make, from the opacity increment of
    one sample, w = T (1 - (1 - a)^(step/unitDistance)) <= T (1 - tau_min),
    with a the LUT's largest opacity (volumeIntegration.h:64-103):

    * an early-termination tie: the samples taken while the opacity is within
      the oracle's 1e-5 band around 0.99 add at most 2e-5 plus one last
      sample, at transmittance T <= 0.01 + 1e-5;
    * a segment-end tie: one sample more or less, at T <= 1.

    The frame's rgb is C * O (Raycast_ptx.cu:159), so a change dC = dO = w
    moves it by at most 2 w."""
    tau_min = (1.0 - float(np.max(vol.opacity))) ** step_over_unit
    w_one = 1.0 - tau_min
    stop = 2.0 * (2e-5 + (0.01 + 1e-5) * w_one)
    exit_ = 2.0 * w_one
    return stop, exit_


def test_volume_512_full_size(require_gpu):
    """The benched 512^3 frame, every pixel checked: colour within 1e-4 of
    the FP64 march on pixels the oracle does not flag and on at least 99% of
    all pixels; a flagged pixel (its fp32 early-termination or segment-end
    decision may go either way) within the colour that decision can change;
    ids exact, entry depth within 1e-5, and the GPU's sample count within
    0.001% of the oracle's."""
    sc, cam = S.volume_workload()
    print("[fullsize] volume 512^3 at 1920x1080, volumeSamplingRate 1", flush=True)
    g = _render(sc, cam, 1920, 1080, volumeSamplingRate=1.0)
    assert not g["errors"], g["errors"]
    print("[fullsize] FP64 oracle march", flush=True)
    o = orc.OracleScene(sc).render(cam, 1920, 1080, volumeSamplingRate=1.0)
    amb = o["ambiguous"]
    ok = amb == 0
    dc = np.abs(g["color"] - o["color"]).max(axis=-1)
    vol = sc.volumes[0]
    step = 0.5 * min(vol.spacing)                 # volumeSamplingRate 1: half a voxel
    b_stop, b_exit = volume_tie_bounds(vol, step / vol.unit_distance)
    bound = np.where(amb & orc.AMB_TIE_STOP, b_stop, 0.0) + \
        np.where(amb & orc.AMB_TIE_T, b_exit, 0.0) + 1e-4
    print(f"[fullsize] unflagged {ok.mean():.5f}, max |dc| unflagged {dc[ok].max():.2e}, "
          f"all pixels within 1e-4 {(dc <= 1e-4).mean():.5f}, flagged max |dc| "
          f"{dc[~ok].max() if (~ok).any() else 0.0:.2e} (bounds stop {b_stop:.2e}, "
          f"exit {b_exit:.2e})", flush=True)
    np.testing.assert_allclose(g["color"][ok], o["color"][ok], rtol=0, atol=1e-4)
    assert (dc <= 1e-4).mean() >= 0.99
    assert np.all(dc <= bound), np.argwhere(dc > bound)[:10]
    for ch in ("primitiveId", "objectId", "instanceId"):
        np.testing.assert_array_equal(g[ch], o[ch])
    np.testing.assert_allclose(g["depth"], o["depth"], rtol=1e-5)
    ref_samples = int(o["stats"][3])       # the oracle's volume sample count
    print(f"[fullsize] samples: GPU {g['volumeSamples']}, oracle {ref_samples}", flush=True)
    assert abs(int(g["volumeSamples"]) - ref_samples) <= 1e-5 * ref_samples
    assert (g["depth"] < 1e29).mean() > 0.3


def test_batched_64_views_full_size(require_gpu):
    """The benched batched scene: the autotuned persistent launch (sRGB, as
    benched) gives every channel of all 64 views bit-identical to the
    per-view ANARI frames; views 0 and 37 equal the oracle (ids exact,
    depth 1e-4, colour one code value)."""
    torch = pytest.importorskip("torch")
    from visrtx_amd import batch as B
    sc, cams = S.batched_workload()
    W = H = 512
    dev = A.Device()
    s = S.build_anari(dev, sc)
    r = S.make_renderer(dev, s, "raycast")
    bf = B.BatchFrame(B.scene_of(dev, s.world, wide=True), B.renderer_of(dev, r), cams, W, H,
                      srgb=True)
    print("[fullsize] batched: autotune + one launch of 64 views at 512^2", flush=True)
    bf.autotune()
    bf.depth.fill_(-1.0)
    bf.render()
    torch.cuda.synchronize()
    got = bf.channels()
    print(f"[fullsize] tile {bf.b.tile_w}x{bf.b.tile_h} block {bf.b.block_threads} stack "
          f"{bf.renderer.stack_lds} width {bf.renderer.node_width}", flush=True)
    for v, cam in enumerate(cams):
        c = S.make_camera(dev, s, cam)
        f = S.make_frame(dev, s, s.world, c, r, W, H, color=A.UFIXED8_RGBA_SRGB)
        dev.render(f)
        dev.ready(f)
        for ch in CHANNELS:
            ref = dev.map_frame(f, "channel." + ch)[0]
            np.testing.assert_array_equal(got[ch][v], ref, err_msg=f"view {v}: {ch}")
        s.drop(f)
        s.drop(c)
    assert not dev.errors(), dev.errors()
    bf = None
    s.release()
    dev.close()
    print("[fullsize] oracle: views 0 and 37", flush=True)
    osc = orc.OracleScene(sc)
    for v in (0, 37):
        o = osc.render(cams[v], W, H, color_format=orc.ORC_COLOR_SRGB8)
        for ch in ("primitiveId", "objectId", "instanceId"):
            np.testing.assert_array_equal(got[ch][v], o[ch], err_msg=f"view {v}: {ch}")
        np.testing.assert_allclose(got["depth"][v], o["depth"], rtol=1e-4, atol=0)
        a = got["color"][v].view(np.uint8).astype(np.int32)
        b = o["color"].view(np.uint8).astype(np.int32)
        assert np.abs(a - b).max() <= 1
    assert len(np.unique(got["instanceId"])) > 200


def _kink_values(lut: np.ndarray) -> np.ndarray:
    """Field values where the loss has a kink in the density: the texel
    centres where the LUT's slope changes, and the two clamps (the LUT is
    fetched at texel coordinate value*256 - 0.5, colorMapHelpers / a14)."""
    d2 = np.abs(np.diff(lut.reshape(256, 4), n=2, axis=0)).max(axis=1)
    j = np.nonzero(d2 > 1e-7)[0] + 1
    return np.concatenate([(j + 0.5) / 256.0, [0.5 / 256.0, 255.5 / 256.0]])


def test_adjoint_128_fd_full_size(require_gpu):
    """The benched C5 configuration (128^3 seeded noise grid, the colormap
    LUT, 32 views at 256^2, 2,097,152 rays) with early stop off: the GPU
    gradient of sum(w * rgba), w non-zero on 8,192 rays, against FP64
    central differences (h = 1e-3) of the oracle's loss over those rays, on
    700 voxels away from the LUT's kinks and 300 LUT scalars."""
    torch = pytest.importorskip("torch")
    from visrtx_amd import diffvol as D
    n = 128
    sp = 1.0 / (n - 1)
    dens = S.gen_noise_volume(n, seed=11)
    lut = D.colormap_lut(S.COLORMAP5, np.array([0.0, 0.3]))
    spec = D.GridSpec(dims=(n, n, n), spacing=(sp, sp, sp), unit_distance=sp, step=0.5 * sp,
                      early_stop=False, image=(256, 256))
    cams = D.orbit_cameras(32, seed=7, fovy=45.0)
    org, dirs = D.make_rays(cams, 256, 256)
    nr = org.shape[0]
    rng = np.random.default_rng(12)
    pick = np.sort(rng.choice(nr, 8192, replace=False))
    w = np.zeros((nr, 4))
    w[pick] = rng.normal(size=(len(pick), 4))
    td = torch.from_numpy(dens).cuda().requires_grad_(True)
    tl = torch.from_numpy(lut).cuda().requires_grad_(True)
    print(f"[fullsize] adjoint 128^3, {nr} rays, 8192 weighted", flush=True)
    rgba = D.render(td, tl, org, dirs, spec)
    (rgba * torch.from_numpy(w).float().cuda()).sum().backward()
    torch.cuda.synchronize()
    gd = td.grad.cpu().numpy().reshape(-1)
    gl = tl.grad.cpu().numpy().reshape(-1)
    o = org.cpu().numpy()[pick]
    d = dirs.cpu().numpy()[pick]
    wp = w[pick]
    args = (spec.dims, spec.origin, spec.spacing)
    dd = dens.astype(np.float64).reshape(-1).copy()
    ll = lut.astype(np.float64).reshape(-1).copy()
    h = 1e-3

    def loss(den, lt, sel=None):
        oo, dv, ww = (o, d, wp) if sel is None else (o[sel], d[sel], wp[sel])
        return orc.diffvol_loss(*args, den, lt, spec.value_range, spec.unit_distance,
                                spec.step, oo, dv, ww, early_stop=False)

    # voxels whose 3x3x3 neighbourhood (every value a sample that reads the
    # voxel interpolates) stays 2h away from every kink value
    kinks = _kink_values(lut)
    vol = dens
    cand = np.nonzero(np.abs(gd) > 0)[0]
    rng.shuffle(cand)
    vox = []
    for i in cand:
        z, y, x = np.unravel_index(i, vol.shape)
        if min(x, y, z) < 1 or max(x, y, z) > n - 2:
            continue
        nb = vol[z - 1:z + 2, y - 1:y + 2, x - 1:x + 2]
        lo, hi = nb.min() - 2 * h, nb.max() + 2 * h
        if np.any((kinks >= lo) & (kinks <= hi)):
            continue
        vox.append(i)
        if len(vox) == 700:
            break
    assert len(vox) == 700
    # rays that can sample a voxel: slab test against the box of the 8 cells
    # around it (a ray's other samples do not change, so the difference is
    # taken over these rays only)
    org_w = np.asarray(spec.origin, np.float64)
    inv = 1.0 / np.where(np.abs(d) < 1e-30, 1e-30, d.astype(np.float64))

    def rays_near(i):
        z, y, x = np.unravel_index(i, vol.shape)
        c = org_w + np.array([x, y, z], np.float64) * sp
        lo, hi = c - 1.01 * sp, c + 1.01 * sp
        t0 = (lo - o) * inv
        t1 = (hi - o) * inv
        tn = np.minimum(t0, t1).max(axis=1)
        tf = np.maximum(t0, t1).min(axis=1)
        return np.nonzero((tn <= tf) & (tf >= 0))[0]

    print("[fullsize] FP64 central differences: 700 voxels", flush=True)
    fd, gg = [], []
    for i in vox:
        sel = rays_near(i)
        dp = dd.copy(); dp[i] += h
        dm = dd.copy(); dm[i] -= h
        fd.append((loss(dp, ll, sel) - loss(dm, ll, sel)) / (2 * h) if len(sel) else 0.0)
        gg.append(gd[i])
    print("[fullsize] FP64 central differences: 300 LUT scalars", flush=True)
    lutp = rng.choice(np.arange(4, 1020), 300, replace=False)
    for i in lutp:
        lp = ll.copy(); lp[i] += h
        lm = ll.copy(); lm[i] -= h
        fd.append((loss(dd, lp) - loss(dd, lm)) / (2 * h))
        gg.append(gl[i])
    fd, gg = np.array(fd), np.array(gg)
    for name, sl in (("voxels", slice(0, 700)), ("lut", slice(700, None)), ("all", slice(None))):
        a, b = gg[sl], fd[sl]
        norm_rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        max_rel = np.abs(a - b).max() / np.abs(b).max()
        print(f"[fullsize] FD {name}: normwise {norm_rel:.2e}, max {max_rel:.2e}", flush=True)
        assert norm_rel <= 1e-3 and max_rel <= 1e-3, name


def test_ao_10m_spheres_16spp_full_size(require_gpu):
    """The benched AO frame: 10M spheres, 1920x1080, 16 spp, one AO ray per
    hit, occlusion distance 0.02, against the oracle (same PCG32 streams)."""
    sc, cam = S.spheres_workload()
    rp = dict(pixelSamples=16, ambientSamples=1, ambientOcclusionDistance=0.02)
    print("[fullsize] AO 16 spp, 10M spheres, 1920x1080", flush=True)
    g = _render(sc, cam, 1920, 1080, kind="ao", **rp)
    assert not g["errors"], g["errors"]
    print("[fullsize] oracle AO", flush=True)
    o = orc.OracleScene(sc).render(cam, 1920, 1080, kind="ao", **rp)
    for ch in ("primitiveId", "objectId", "instanceId"):
        np.testing.assert_array_equal(g[ch], o[ch])
    np.testing.assert_allclose(g["depth"], o["depth"], rtol=1e-4, atol=0)
    np.testing.assert_allclose(g["color"], o["color"], rtol=1e-4, atol=1e-4)
    print(f"[fullsize] AO colour bit-identical fraction {np.mean(g['color'] == o['color']):.6f}",
          flush=True)
    assert g["rays"] > 33_000_000
