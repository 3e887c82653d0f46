"""Benchmark of the MI355X ANARI device on the synthetic workloads.

Headline (``value``): primary closest-hit Mrays/s of the raycast renderer on
the 10M-sphere scene at 1920x1080, 1 spp -- one step is one ANARI frame
(camera commit, which resets accumulation, then anariRenderFrame) with the
scene already resident in HBM.  Also reported under ``extra``: LBVH build ms
of that scene, AO renderer Mrays/s (16 spp), the 131k-triangle heightfield
raycast, and the 512^3 volume march in Gsamples/s.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json

import numpy as np
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD public spec


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def device_sync():
    from visrtx_amd import _native
    _native.check(_native.hip().vrx_device_sync(), "device sync")
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except ImportError:
        pass


def load_counts():
    with open(os.path.join(ROOT, "tests", "golden", "workload_counts.json")) as f:
        return json.load(f)


def load_traffic(name):
    p = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


class Workload:
    """An ANARI scene + frame on the device."""

    def __init__(self, dev, sc, cam, w, h, kind="raycast", color=None, ids=True, **rp):
        from visrtx_amd import anari as A
        from visrtx_amd import scenes as S
        self.dev, self.A = dev, A
        self.s = S.build_anari(dev, sc)
        self.cam = S.make_camera(dev, self.s, cam)
        self.r = S.make_renderer(dev, self.s, kind, **rp)
        self.f = S.make_frame(dev, self.s, self.s.world, self.cam, self.r, w, h,
                              color=color if color is not None else A.UFIXED8_RGBA_SRGB,
                              depth=True, ids=ids)
        self.w, self.h = w, h

    def step(self):
        # a camera commit resets accumulation (Frame.cu:574-588), so every step
        # renders a full frame even for the single-shot raycast renderer
        self.dev.commit(self.cam)
        self.dev.render(self.f)

    def duration(self):
        return self.dev.get_property(self.f, "duration", self.A.FLOAT32)

    def rays(self):
        return self.dev.get_property(self.f, "raysTraced", self.A.UINT64)

    def samples(self):
        return self.dev.get_property(self.f, "volumeSamples", self.A.UINT64)

    def tune(self, max_frames=12):
        """Untimed frames until the frame's launch autotuner (node width x LDS
        stack depth, one candidate per frame) has settled."""
        # the property describes the latest render, so a parameter change shows
        # only after one frame has been rendered with it
        self.step()
        for _ in range(max_frames):
            if self.dev.get_property(self.f, "tunedNodeWidth", self.A.INT32):
                break
            self.step()
        device_sync()

    def tuned(self):
        return (self.dev.get_property(self.f, "tunedNodeWidth", self.A.INT32),
                self.dev.get_property(self.f, "tunedStackDepth", self.A.INT32))

    def pin_other_width(self):
        """Renderer parameters for the node width the autotuner did not pick,
        at the tuned LDS stack depth; returns that width."""
        w, sd = self.tuned()
        other = 2 if w == 4 else 4
        self.dev.set(self.r, "nodeWidth", other, self.A.INT32)
        self.dev.set(self.r, "stackDepthLDS", sd, self.A.INT32)
        self.dev.commit(self.r)
        return other

    def timed(self, steps, warmup):
        self.tune()
        for _ in range(warmup):
            self.step()
        device_sync()
        durs = []
        t0 = time.perf_counter()
        for _ in range(steps):
            self.step()
            durs.append(self.duration())   # waits on the frame's end event
        device_sync()
        wall = time.perf_counter() - t0
        return wall / steps, durs

    def release(self):
        self.s.release()


def bench_lbvh(dev, work, reps=5):
    """LBVH build time of the world (hipEvents around the BLAS build kernels)."""
    from visrtx_amd import anari as A
    ms = []
    geom = work.s.geoms[0]
    for _ in range(reps):
        dev.commit(geom)
        dev.get_property(work.s.world, "bounds", A.FLOAT32_BOX3)
        ms.append(1000.0 * dev.get_property(work.s.world, "bvhBuildDuration", A.FLOAT32))
    return statistics.median(ms), ms


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """HBM bandwidth of a plain streaming copy kernel (read + write bytes / time)."""
    import ctypes as C
    from visrtx_amd import _native
    L = _native.hip()
    src, dst = C.c_void_p(), C.c_void_p()
    _native.check(L.vrx_malloc(C.byref(src), nbytes))
    _native.check(L.vrx_malloc(C.byref(dst), nbytes))
    _native.check(L.vrx_fill_u32(src, 1, nbytes // 4, None))
    e0, e1 = C.c_void_p(), C.c_void_p()
    L.vrx_event_create(C.byref(e0))
    L.vrx_event_create(C.byref(e1))
    _native.check(L.vrx_copy_kernel(dst, src, nbytes, None))
    L.vrx_event_record(e0, None)
    for _ in range(reps):
        _native.check(L.vrx_copy_kernel(dst, src, nbytes, None))
    L.vrx_event_record(e1, None)
    L.vrx_event_sync(e1)
    ms = C.c_float()
    L.vrx_event_elapsed_ms(e0, e1, C.byref(ms))
    L.vrx_event_destroy(e0)
    L.vrx_event_destroy(e1)
    L.vrx_free(src)
    L.vrx_free(dst)
    return 2.0 * nbytes * reps / (ms.value / 1000.0) / 1e9


def bench_diffvol(iters=200, log_every=50):
    """128^3 density + 256-entry LUT, 32 views at 256x256: forward-only and
    forward+adjoint times, then Adam (lr 1e-2) from a flat grid and a grey LUT."""
    import torch
    from visrtx_amd import diffvol as D
    from visrtx_amd import scenes as S
    n = 128
    sp = 1.0 / (n - 1)
    spec = D.GridSpec(dims=(n, n, n), spacing=(sp, sp, sp), unit_distance=sp, step=0.5 * sp,
                      early_stop=True, image=(256, 256))
    gt = torch.from_numpy(S.gen_noise_volume(n, seed=11)).cuda()
    gt_lut = torch.from_numpy(D.colormap_lut(S.COLORMAP5, np.array([0.0, 0.3]))).cuda()
    cams = D.orbit_cameras(32, seed=7, fovy=45.0)
    org, dirs = D.make_rays(cams, 256, 256)
    with torch.no_grad():
        target = D.render(gt, gt_lut, org, dirs, spec)
    dens = torch.full_like(gt, 0.5).requires_grad_(True)
    lut = torch.from_numpy(D.linear_lut(0.2)).cuda().requires_grad_(True)

    def timeit(fn, reps):
        """median over reps of one call each, bracketed by HIP events"""
        fn()
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return statistics.median(ts)

    def fwd():
        with torch.no_grad():
            D.render(dens, lut, org, dirs, spec)

    def fwdbwd():
        loss = ((D.render(dens, lut, org, dirs, spec) - target) ** 2).mean()
        loss.backward()

    # the integral alone: forward, then backward from a fixed upstream gradient
    # (the MSE gradient at the start point); no loss kernels in the timing
    with torch.no_grad():
        rgba0 = D.render(dens, lut, org, dirs, spec)
        g_up = (2.0 / rgba0.numel()) * (rgba0 - target)

    def fwdbwd_integral():
        dens.grad = None
        lut.grad = None
        D.render(dens, lut, org, dirs, spec).backward(g_up)

    fwd_ms = timeit(fwd, 20)
    fbi_ms = timeit(fwdbwd_integral, 20)
    fb_ms = timeit(fwdbwd, 20)
    # the same with the backward replayed from the transmittance checkpoints
    spec_suffix = spec
    spec = D.GridSpec(**{**spec_suffix.__dict__, "adjoint": "replay"})
    fbi_replay_ms = timeit(fwdbwd_integral, 20)
    fb_replay_ms = timeit(fwdbwd, 20)
    spec = spec_suffix
    dens.grad = None
    lut.grad = None
    opt = torch.optim.Adam([dens, lut], lr=1e-2)
    dens.grad = None
    lut.grad = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    first = last = None
    # per iteration: forward (the render) and backward (MSE loss, its
    # gradient and the adjoint) bracketed by events on the current stream
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(iters)]
    for it in range(iters):
        opt.zero_grad()
        ev[it][0].record()
        rgba = D.render(dens, lut, org, dirs, spec)
        ev[it][1].record()
        loss = ((rgba - target) ** 2).mean()
        loss.backward()
        ev[it][2].record()
        opt.step()
        with torch.no_grad():
            lut.clamp_(0.0, 0.999)
        if it == 0:
            first = loss.item()
        if log_every and (it + 1) % log_every == 0:
            log(f"diffvol iter {it + 1}: loss {loss.item():.4e}")
    torch.cuda.synchronize()
    iter_ms = (time.perf_counter() - t0) * 1000.0 / iters
    per_it = [(a.elapsed_time(b), b.elapsed_time(c)) for a, b, c in ev]
    ratio_mean = statistics.mean((f + b) / f for f, b in per_it)
    with torch.no_grad():
        last = ((D.render(dens, lut, org, dirs, spec) - target) ** 2).mean().item()
        # the same timings at the state Adam reached (denser field, rays cross
        # transfer-function intervals at almost every sample)
        rgba1 = D.render(dens, lut, org, dirs, spec)
        g_up = (2.0 / rgba1.numel()) * (rgba1 - target)
    fwd_ms_final = timeit(fwd, 10)
    fbi_ms_final = timeit(fwdbwd_integral, 10)
    return {"fwd_ms": round(fwd_ms, 3), "adjoint": "suffix",
            "fwd_bwd_integral_ms": round(fbi_ms, 3),
            "fwd_bwd_integral_over_fwd": round(fbi_ms / fwd_ms, 3),
            "fwd_bwd_with_mse_loss_ms": round(fb_ms, 3),
            "fwd_bwd_with_mse_loss_over_fwd": round(fb_ms / fwd_ms, 3),
            "replay_adjoint": {"fwd_bwd_integral_ms": round(fbi_replay_ms, 3),
                               "fwd_bwd_integral_over_fwd": round(fbi_replay_ms / fwd_ms, 3),
                               "fwd_bwd_with_mse_loss_ms": round(fb_replay_ms, 3)},
            "after_adam": {"fwd_ms": round(fwd_ms_final, 3),
                           "fwd_bwd_integral_ms": round(fbi_ms_final, 3),
                           "fwd_bwd_integral_over_fwd": round(fbi_ms_final / fwd_ms_final, 3)},
            "adam_ms_per_iter": round(iter_ms, 3),
            "adam_fwd_bwd_over_fwd_mean": round(ratio_mean, 3),
            "adam_fwd_ms_mean": round(statistics.mean(f for f, _ in per_it), 3),
            "adam_bwd_ms_mean": round(statistics.mean(b for _, b in per_it), 3),
            "iters": iters, "mse_first": first, "mse_final": last, "rays": int(org.shape[0])}


def bench_woodcock(dev, work, cam, w, h, reps=10):
    """Woodcock distance sampling (vrx_volume_sample_distance) of one camera
    ray per pixel through the world's volume, over its majorant grid: the
    volume record (field, transfer-function LUT, grid) is the one the ANARI
    device built for the world."""
    import ctypes as C
    import torch
    from visrtx_amd import _native as N
    from visrtx_amd import batch as B
    L = N.hip()
    scene = B.scene_of(dev, work.s.world)
    v = N.vrx_volume()
    N.check(L.vrx_memcpy_sync(C.c_void_p(C.addressof(v)), C.c_void_p(scene.d_volumes),
                              C.c_size_t(C.sizeof(v)), N.VRX_D2H), "volume record")
    if not v.d_majorant:
        raise RuntimeError("the world attached no majorant grid to the volume")
    cells = L.vrx_volume_grid_cells(C.byref(v))
    grid = np.empty(cells, np.float32)
    N.check(L.vrx_memcpy_sync(C.c_void_p(grid.ctypes.data), C.c_void_p(v.d_majorant),
                              C.c_size_t(grid.nbytes), N.VRX_D2H), "majorant grid")
    n = w * h
    cs = B.camera_struct(cam)
    org = torch.tensor(list(cs.pos), device="cuda").expand(n, 3).contiguous()
    ys, xs = torch.meshgrid(torch.arange(h, device="cuda"), torch.arange(w, device="cuda"),
                            indexing="ij")
    sx = (xs.reshape(-1).float() + 0.5) / w
    sy = (ys.reshape(-1).float() + 0.5) / h
    d00, du, dv = (torch.tensor(list(x), device="cuda") for x in (cs.d00, cs.du, cs.dv))
    dirs = d00[None] + sx[:, None] * du[None] + sy[:, None] * dv[None]
    dirs = (dirs / dirs.norm(dim=1, keepdim=True)).contiguous()
    tmm = torch.tensor([0.0, 1e30], device="cuda").expand(n, 2).contiguous()
    t = torch.empty(n, device="cuda")
    alb = torch.empty((n, 3), device="cuda")
    tr = torch.empty(n, device="cuda")

    def run(seed):
        N.check(L.vrx_volume_sample_distance(
            C.byref(v), n, C.c_void_p(org.data_ptr()), C.c_void_p(dirs.data_ptr()),
            C.c_void_p(tmm.data_ptr()), C.c_uint64(seed), C.c_void_p(t.data_ptr()),
            C.c_void_p(alb.data_ptr()), C.c_void_p(tr.data_ptr()), None), "sample_distance")

    run(0)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for k in range(reps):
        run(k + 1)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return {"rays": n, "ms": round(ms, 4), "mrays_s": round(n / (ms / 1000) / 1e6, 2),
            "collision_fraction": round(float((tr == 0).float().mean()), 4),
            "empty_cell_fraction": round(float((grid == 0).mean()), 4),
            "lut": "built by the device's transferFunction1D volume"}


def bench_batched(dev, reps=5):
    """64 orbit views at 512x512 of 256 instances of 16 meshes (10k-2M
    triangles): one autotuned persistent launch (eager and HIP-graph replay)
    against 64 separate per-view ANARI frames rendered back to back."""
    import torch
    from visrtx_amd import anari as A
    from visrtx_amd import batch as B
    from visrtx_amd import scenes as S
    t = time.perf_counter()
    sc, cams = S.batched_workload()
    tris = sum(g.num_prims for g in sc.geoms)
    log(f"batched scene: {len(sc.geoms)} meshes, {tris} triangles, {len(sc.instances)} "
        f"instances, generated in {time.perf_counter() - t:.1f}s")
    s = S.build_anari(dev, sc)
    r = S.make_renderer(dev, s, "raycast")
    scene = B.scene_of(dev, s.world, wide=True)
    out = {"meshes": len(sc.geoms), "triangles": int(tris), "instances": len(sc.instances),
           "views": len(cams), "res": 512,
           "build_ms": round(1000.0 * dev.get_property(s.world, "bvhBuildDuration", A.FLOAT32),
                             3)}
    W = H = 512
    # plain: one ANARI frame per view
    frames = []
    for cam in cams:
        c = S.make_camera(dev, s, cam)
        frames.append((c, S.make_frame(dev, s, s.world, c, r, W, H, color=A.UFIXED8_RGBA_SRGB)))

    def plain():
        for c, f in frames:
            dev.commit(c)
            dev.render(f)
        for _, f in frames:
            dev.ready(f)

    plain()
    device_sync()
    t = time.perf_counter()
    for _ in range(reps):
        plain()
    device_sync()
    plain_ms = (time.perf_counter() - t) * 1000.0 / reps
    kern_ms = sum(dev.get_property(f, "duration", A.FLOAT32) for _, f in frames) * 1000.0
    for c, f in frames:
        s.drop(f)
        s.drop(c)

    bf = B.BatchFrame(scene, B.renderer_of(dev, r), cams, W, H, srgb=True)
    best = bf.autotune()
    log(f"batched autotune: {best:.3f} ms, tile {bf.b.tile_w}x{bf.b.tile_h}, block "
        f"{bf.b.block_threads}, LDS stack {bf.renderer.stack_lds}, node width "
        f"{bf.renderer.node_width}")
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        bf.render(stream)
        stream.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t = time.perf_counter()
        e0.record(stream)
        for _ in range(reps):
            bf.render(stream)
        e1.record(stream)
        stream.synchronize()
        eager_ms = (time.perf_counter() - t) * 1000.0 / reps
        eager_ev = e0.elapsed_time(e1) / reps
        bf.capture(stream)
        bf.replay(stream)
        stream.synchronize()
        t = time.perf_counter()
        e0.record(stream)
        for _ in range(reps):
            bf.replay(stream)
        e1.record(stream)
        stream.synchronize()
        graph_ms = (time.perf_counter() - t) * 1000.0 / reps
        graph_ev = e0.elapsed_time(e1) / reps
    bf.release_graph()
    rays = len(cams) * W * H
    out.update({"rays": rays, "plain_ms": round(plain_ms, 3),
                "plain_sum_of_kernels_ms": round(kern_ms, 3),
                "batched_eager_ms": round(eager_ms, 3), "batched_eager_event_ms": round(eager_ev, 3),
                "batched_graph_ms": round(graph_ms, 3), "batched_graph_event_ms": round(graph_ev, 3),
                "speedup_vs_plain": round(plain_ms / graph_ms, 3),
                "batched_mrays_s": round(rays / (graph_ms / 1000.0) / 1e6, 2),
                "plain_mrays_s": round(rays / (plain_ms / 1000.0) / 1e6, 2),
                "autotuned": {"tile": [bf.b.tile_w, bf.b.tile_h], "block": bf.b.block_threads,
                              "stack_lds": bf.renderer.stack_lds,
                              "node_width": bf.renderer.node_width}})
    del bf
    s.release()
    return out


def cpu_baseline(threads):
    """The oracle (C/OpenMP port of the same path) on the host cores: the
    headline 10M-sphere frame at full size (1920x1080 primary raycast rays,
    same camera), median of 5 after a warm-up; plus the oracle's figures for
    the 131k-triangle heightfield (full 1024x768 frame) and for the 512^3
    volume march (FP64, a 480x270 frame of the same camera, Gsamples/s)."""
    from oracle import oracle as orc
    from visrtx_amd import scenes as S

    def timed(fn, reps):
        fn()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
        return statistics.median(ts)

    sc, cam = S.spheres_workload()
    t0 = time.perf_counter()
    o = orc.OracleScene(sc)
    build_s = time.perf_counter() - t0
    w, h = 1920, 1080
    med = timed(lambda: o.render(cam, w, h, "raycast", threads=threads), 5)
    o.close()
    del sc
    sc1, cam1 = S.heightfield_workload()
    o1 = orc.OracleScene(sc1)
    med1 = timed(lambda: o1.render(cam1, 1024, 768, "raycast", threads=threads), 5)
    o1.close()
    sc3, cam3 = S.volume_workload(width=480, height=270)
    o3 = orc.OracleScene(sc3)
    box = {}

    def vol():
        box["out"] = o3.render(cam3, 480, 270, "raycast", threads=threads, volumeSamplingRate=1.0)

    med3 = timed(vol, 3)
    samples3 = int(box["out"]["stats"][3])
    o3.close()
    base = {"value": round(w * h / med / 1e6, 3), "unit": "Mrays/s", "cores": threads,
            "kind": "port",
            "sample": f"10M-sphere scene, the full {w}x{h} primary raycast frame (same camera), "
                      f"oracle LBVH + BVH2 traversal in C/OpenMP, median of 5; oracle LBVH "
                      f"build {build_s * 1000:.0f} ms"}
    others = {"heightfield_1024x768_mrays_s": round(1024 * 768 / med1 / 1e6, 3),
              "volume512_480x270_gsamples_s": round(samples3 / med3 / 1e9, 4),
              "volume512_480x270_samples": samples3, "cores": threads,
              "kind": "port (FP64 march for the volume)"}
    return base, others


SECTIONS = ("headline", "ao", "direct", "instances", "heightfield", "volume", "sparse",
            "batched", "diffvol")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only-headline", action="store_true")
    ap.add_argument("--sections", default=",".join(SECTIONS),
                    help="comma-separated subset of " + ",".join(SECTIONS) +
                         " (profiling one kernel); the JSON value needs 'headline'")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this device renders on one GPU; multi-GPU sharding is "
                                   "described but not built (one GPU per process); run with --gpus 1"}))
        sys.exit(2)
    sec = {"headline"} if args.only_headline else set(args.sections.split(","))
    unknown = sec - set(SECTIONS)
    if unknown:
        ap.error(f"unknown sections {sorted(unknown)}")

    from visrtx_amd import anari as A
    from visrtx_amd import scenes as S
    counts = load_counts()

    dev = A.Device()
    extra = {}
    extra["copy_kernel_gbs"] = round(copy_bandwidth(), 1)
    log(f"copy kernel bandwidth {extra['copy_kernel_gbs']} GB/s")

    W, H = 1920, 1080
    value = sec_per_step = roofline = None
    rays = W * H
    if "headline" in sec or "ao" in sec or "direct" in sec:
        # ---- headline: 10M spheres, 1920x1080 raycast ---------------------
        log("generating 10M spheres")
        sc, cam = S.spheres_workload(width=W, height=H)
        work = Workload(dev, sc, cam, W, H, "raycast")
        log("building scene")
        t = time.perf_counter()
        dev.get_property(work.s.world, "bounds", A.FLOAT32_BOX3)
        log(f"scene build (upload + LBVH) {time.perf_counter() - t:.2f}s")
    if "headline" in sec:
        lbvh_ms, lbvh_all = bench_lbvh(dev, work)
        extra["lbvh_build_ms_10m_spheres"] = round(lbvh_ms, 3)
        # algorithmic build bytes per primitive: AABB read 24, (key, index)
        # write 8, 16 per radix pass (4 passes here), Karras node 32, refit 100
        bpp = 24 + 8 + 16 * 4 + 32 + 100
        gbs_b = 10_000_000 * bpp / (lbvh_ms / 1000) / 1e9
        extra["lbvh_roofline"] = {"bound": "hbm", "algorithmic_bytes_per_prim": bpp,
                                  "achieved_gbs": round(gbs_b, 1),
                                  "frac": round(gbs_b / HBM_PEAK_GBS, 4)}
        log(f"LBVH build {lbvh_ms:.3f} ms (all {['%.3f' % x for x in lbvh_all]})")
        sec_per_step, durs = work.timed(args.steps, args.warmup)
        value = rays / sec_per_step / 1e6
        kern_ms = statistics.mean(durs) * 1000.0
        tw, tsd = work.tuned()
        log(f"primary raycast: {sec_per_step * 1000:.3f} ms/step wall, kernel {kern_ms:.3f} ms "
            f"-> {value:.1f} Mrays/s (kernel-only {rays / (kern_ms / 1000) / 1e6:.1f}); tuned "
            f"node width {tw}, LDS stack {tsd}")
        extra["tuned_launch"] = {"node_width": tw, "stack_lds": tsd}
        extra["primary_kernel_mrays_s"] = round(rays / (kern_ms / 1000) / 1e6, 2)
        c2 = counts["spheres10m_1920x1080"]
        bytes_per_ray = 32 + 16 + 64 * c2["nodes_per_ray"] + 16 * c2["prims_per_ray"]
        alg_bytes = rays * bytes_per_ray
        achieved = alg_bytes / (kern_ms / 1000) / 1e9
        traffic = load_traffic("render_spheres10m_1920x1080")
        # HBM bytes of the kernel instance the autotuner picked (both node
        # widths are profiled)
        inst = f"k_render<1, false, 0, {'true' if tw == 4 else 'false'}>"
        tb = [v for k, v in (traffic or {}).get("per_kernel", {}).items() if inst in k]
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": tb[0] if tb else None,
                    "kernel": f"k_render<sphere,single-level,raycast,{tw}-wide nodes>",
                    "stack_lds": tsd,
                    "algorithmic_bytes_per_ray": round(bytes_per_ray, 2),
                    "avg_kernel_ms": round(kern_ms, 4)}
        extra["widen_ms_10m_spheres"] = round(
            1000.0 * dev.get_property(work.s.world, "bvhWidenDuration", A.FLOAT32), 3)
        # the same frames with the node width the autotuner did not pick
        ow = work.pin_other_width()
        sps2, durs2 = work.timed(args.steps, args.warmup)
        k2 = statistics.mean(durs2) * 1000.0
        extra[f"primary_kernel_mrays_s_width{ow}"] = round(rays / (k2 / 1000) / 1e6, 2)
        extra[f"primary_wall_mrays_s_width{ow}"] = round(rays / sps2 / 1e6, 2)
        log(f"primary raycast, {ow}-wide nodes: kernel {k2:.3f} ms -> "
            f"{extra[f'primary_kernel_mrays_s_width{ow}']} Mrays/s (4-wide widen "
            f"{extra['widen_ms_10m_spheres']} ms)")
    if "ao" in sec:
        # AO renderer, 16 spp
        ao = Workload(dev, sc, cam, W, H, "ao", ids=True, pixelSamples=16, ambientSamples=1,
                      ambientOcclusionDistance=0.02)
        sps, ad = ao.timed(3, 1)
        nr = ao.rays()
        extra["ao16_mrays_s"] = round(nr / sps / 1e6, 2)
        extra["ao16_rays_per_frame"] = int(nr)
        extra["ao16_ms_per_frame"] = round(sps * 1000, 3)
        extra["ao16_tuned_launch"] = dict(zip(("node_width", "stack_lds"), ao.tuned()))
        log(f"AO 16spp: {sps * 1000:.2f} ms/frame, {nr} rays -> {extra['ao16_mrays_s']} Mrays/s "
            f"(tuned {extra['ao16_tuned_launch']})")
        ow = ao.pin_other_width()
        sps2, _ = ao.timed(3, 1)
        extra[f"ao16_mrays_s_width{ow}"] = round(ao.rays() / sps2 / 1e6, 2)
        log(f"AO 16spp, {ow}-wide nodes: {sps2 * 1000:.2f} ms/frame -> "
            f"{extra[f'ao16_mrays_s_width{ow}']} Mrays/s")
        ao.release()
    if "direct" in sec:
        # directLight renderer on the same spheres: a directional and a point
        # light (one shadow ray each per hit), one AO ray, 4 spp
        sc.lights = [S.LightDesc("directional", direction=(0.3, -1.0, 0.5), irradiance=2.0),
                     S.LightDesc("point", position=(0.5, 1.5, -0.5), intensity=1.0)]
        dl = Workload(dev, sc, cam, W, H, "directLight", ids=True, pixelSamples=4,
                      ambientSamples=1, ambientOcclusionDistance=0.02, ambientRadiance=0.2)
        sps, _ = dl.timed(3, 1)
        nr = dl.rays()
        extra["direct_light_4spp"] = {
            "mrays_s": round(nr / sps / 1e6, 2), "rays_per_frame": int(nr),
            "ms_per_frame": round(sps * 1000, 3),
            "tuned_launch": dict(zip(("node_width", "stack_lds"), dl.tuned())),
            "lights": "directional + point", "ambient_samples": 1}
        log(f"directLight 4spp: {sps * 1000:.2f} ms/frame, {nr} rays -> "
            f"{extra['direct_light_4spp']['mrays_s']} Mrays/s")
        dl.release()
        sc.lights = []
    if "headline" in sec or "ao" in sec or "direct" in sec:
        work.release()
        del sc

    if "instances" in sec:
        # one Instance with a 1M-element transform array over a sphere group,
        # 1920x1080 primary raycast (two-level traversal)
        sci, cami = S.instanced_spheres_workload(1_000_000, width=W, height=H)
        t = time.perf_counter()
        iw = Workload(dev, sci, cami, W, H, "raycast")
        dev.get_property(iw.s.world, "bounds", A.FLOAT32_BOX3)
        build_s = time.perf_counter() - t
        sps, di = iw.timed(10, 2)
        extra["instances_1m_transform_array"] = {
            "scene_build_s": round(build_s, 3),
            "bvh_build_ms": round(1000.0 * dev.get_property(iw.s.world, "bvhBuildDuration",
                                                             A.FLOAT32), 3),
            "mrays_s": round(W * H / sps / 1e6, 2),
            "kernel_ms": round(statistics.mean(di) * 1000, 4),
            "tuned_launch": dict(zip(("node_width", "stack_lds"), iw.tuned()))}
        log(f"1M instances: build {build_s:.2f}s, "
            f"{extra['instances_1m_transform_array']['mrays_s']} Mrays/s")
        iw.release()
        del sci
    if "heightfield" in sec:
        sc1, cam1 = S.heightfield_workload()
        hf = Workload(dev, sc1, cam1, 1024, 768, "raycast")
        sps, d1 = hf.timed(20, 3)
        extra["heightfield_mrays_s"] = round(1024 * 768 / sps / 1e6, 2)
        extra["heightfield_kernel_ms"] = round(statistics.mean(d1) * 1000, 4)
        tw, tsd = hf.tuned()
        extra["heightfield_tuned_launch"] = {
            "node_width": tw, "stack_lds": tsd,
            "long_ray_steps": dev.get_property(hf.f, "tunedLongRaySteps", A.INT32)}
        ow = hf.pin_other_width()
        _, d12 = hf.timed(20, 3)
        extra[f"heightfield_kernel_ms_width{ow}"] = round(statistics.mean(d12) * 1000, 4)
        # the tuned launch without the long-ray pass
        dev.set(hf.r, "nodeWidth", tw, A.INT32)
        dev.set(hf.r, "longRaySteps", 0, A.INT32)
        dev.commit(hf.r)
        _, d13 = hf.timed(20, 3)
        extra["heightfield_kernel_ms_no_long_ray_pass"] = round(statistics.mean(d13) * 1000, 4)
        # roofline of the tuned kernel: algorithmic traversal bytes per ray
        # (ray in, hit out, 64-B nodes and 48-B triangles the oracle visits)
        c1 = counts["heightfield_1024x768"]
        bpr1 = 32 + 16 + 64 * c1["nodes_per_ray"] + 48 * c1["prims_per_ray"]
        gbs1 = 1024 * 768 * bpr1 / (extra["heightfield_kernel_ms"] / 1000) / 1e9
        extra["heightfield_roofline"] = {"bound": "hbm", "algorithmic_bytes_per_ray": round(bpr1, 2),
                                         "achieved_gbs": round(gbs1, 1),
                                         "frac": round(gbs1 / HBM_PEAK_GBS, 4)}
        log(f"heightfield: {sps * 1000:.3f} ms/step -> {extra['heightfield_mrays_s']} Mrays/s")
        hf.release()

    if "volume" in sec:
        log("generating 512^3 volume")
        sc3, cam3 = S.volume_workload()
        vol = Workload(dev, sc3, cam3, W, H, "raycast", volumeSamplingRate=1.0)
        sps, d3 = vol.timed(5, 1)
        ns = vol.samples()
        kms = statistics.mean(d3) * 1000
        ref_ns = counts["volume512_1920x1080"]["samples"]
        extra["volume_gsamples_s"] = round(ref_ns / (kms / 1000) / 1e9, 2)
        extra["volume_gpu_samples"] = int(ns)
        extra["volume_oracle_samples"] = int(ref_ns)
        extra["volume_kernel_ms"] = round(kms, 3)
        extra["volume_hbm_frac_32B_per_sample"] = round(ref_ns * 32 / (kms / 1000) / 1e9 /
                                                        HBM_PEAK_GBS, 4)
        vt = load_traffic("render_volume512_1920x1080")
        extra["volume_traffic_bytes_per_launch"] = vt["hbm_bytes_per_launch"] if vt else None
        # the profiled HBM bytes over this run's kernel time: what the
        # launch really moves (neighbouring samples share cache lines, so the
        # algorithmic fraction above can exceed 1)
        extra["volume_hbm_frac_profiled_traffic"] = (
            round(vt["hbm_bytes_per_launch"] / (kms / 1000) / 1e9 / HBM_PEAK_GBS, 4) if vt else None)
        log(f"volume: kernel {kms:.3f} ms, {ns} samples (oracle {ref_ns}) -> "
            f"{extra['volume_gsamples_s']} Gsamples/s")
        vol.release()
        del sc3
    if "sparse" in sec:
        # a sparse 512^3 field (12 blobs): the march with and without skipping
        # of empty 16^3 macro cells (identical frames)
        log("generating sparse 512^3 volume")
        sc4, cam4 = S.sparse_volume_workload(512, width=W, height=H)
        res = {}
        for skip in (False, True):
            sv = Workload(dev, sc4, cam4, W, H, "raycast", volumeSamplingRate=1.0,
                          emptySpaceSkipping=skip)
            _, d4 = sv.timed(5, 1)
            res[skip] = (statistics.mean(d4) * 1000, sv.samples())
            if skip:
                extra["woodcock_sparse_512"] = bench_woodcock(dev, sv, cam4, W, H)
                log(f"woodcock: {extra['woodcock_sparse_512']['mrays_s']} Mrays/s")
            sv.release()
        full = res[False][1]
        extra["volume_sparse_512"] = {
            "kernel_ms_no_skip": round(res[False][0], 3),
            "kernel_ms_skip": round(res[True][0], 3),
            "speedup": round(res[False][0] / res[True][0], 2),
            "samples_no_skip": int(full), "samples_skip": int(res[True][1]),
            "effective_gsamples_s_skip": round(full / (res[True][0] / 1000) / 1e9, 2)}
        log(f"sparse volume: {res[False][0]:.3f} ms -> {res[True][0]:.3f} ms with skipping")
        del sc4

    if "batched" in sec:
        log("batched 64 views")
        bt = bench_batched(dev)
        # roofline of the graph-replayed launch: traversal bytes per ray from
        # the oracle's counts on views 0 and 37 (BLAS nodes and triangles)
        c4 = counts["batched_views0_37_512"]
        bpr4 = 32 + 16 + 64 * c4["nodes_per_ray"] + 48 * c4["prims_per_ray"]
        gbs4 = bt["rays"] * bpr4 / (bt["batched_graph_event_ms"] / 1000) / 1e9
        bt["roofline"] = {"bound": "hbm", "algorithmic_bytes_per_ray": round(bpr4, 1),
                          "achieved_gbs": round(gbs4, 1), "frac": round(gbs4 / HBM_PEAK_GBS, 4),
                          "counts": "oracle, views 0 and 37"}
        extra["batched_64x512"] = bt
        log(f"batched: plain {bt['plain_ms']} ms, graph {bt['batched_graph_ms']} ms "
            f"-> {bt['speedup_vs_plain']}x ({bt['batched_mrays_s']} Mrays/s)")

    dev.close()

    if "diffvol" in sec:
        log("differentiable volume")
        dv = bench_diffvol()
        extra["diffvol_128"] = dv
        log(f"diffvol: fwd {dv['fwd_ms']} ms, fwd+bwd {dv['fwd_bwd_integral_ms']} ms "
            f"({dv['fwd_bwd_integral_over_fwd']}x; with the MSE loss "
            f"{dv['fwd_bwd_with_mse_loss_ms']} ms = {dv['fwd_bwd_with_mse_loss_over_fwd']}x; "
            f"checkpoint-replay adjoint {dv['replay_adjoint']['fwd_bwd_integral_ms']} ms), "
            f"Adam {dv['adam_ms_per_iter']} ms/iter, "
            f"MSE {dv['mse_first']:.3e} -> {dv['mse_final']:.3e}")

    cpu = None
    if not args.no_cpu_baseline:
        # the host cores this GPU's share of the machine has (16 per GPU),
        # or fewer where the host has fewer
        threads = min(os.cpu_count() or 1, 16)
        log(f"cpu baseline with {threads} threads")
        cpu, cpu_other = cpu_baseline(threads)
        extra["cpu_oracle_other"] = cpu_other
        log(f"cpu baseline: {cpu['value']} Mrays/s; heightfield "
            f"{cpu_other['heightfield_1024x768_mrays_s']} Mrays/s; volume "
            f"{cpu_other['volume512_480x270_gsamples_s']} Gsamples/s")

    out = {
        "metric": "Mrays/s closest-hit, volume Gsamples/s, fwd+bwd ms/iter on 1 MI355X",
        "value": round(value, 2) if value is not None else None,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec_per_step * 1000, 4) if sec_per_step else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded PCG32 spheres)",
        "config": {"workload": "10M spheres (centers U[0,1]^3, radii log-uniform [1e-4,3e-3]), "
                               "1920x1080 raycast renderer, 1 spp, primary closest-hit",
                   "rays_per_step": rays, "renderer": "raycast",
                   "channels": "color(SRGB8), depth, primitiveId, objectId, instanceId"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
