"""The C++ ANARI client (visrtx_amd/examples/spheres_app.cpp) end to end on the
GPU: it loads the device library by name like a visrtx client, renders, and
its primitiveId checksum over hit pixels must equal the oracle's for the same
spheres (regenerated here with the app's xorshift64* generator)."""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc
from visrtx_amd import scenes as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "visrtx_amd", "examples", "spheres_app")
M64 = (1 << 64) - 1


def _app_spheres(n):
    """The app's generator: xorshift64* (12, 25, 27), top 24 bits / 2^24."""
    s = 0x9E3779B97F4A7C15
    out = np.empty(4 * n, np.float32)
    for i in range(4 * n):
        s ^= s >> 12
        s ^= (s << 25) & M64
        s ^= s >> 27
        out[i] = np.float32(((s * 2685821657736338717) & M64) >> 40) * np.float32(1.0 / 16777216.0)
    v = out.reshape(n, 4)
    pos = np.ascontiguousarray(v[:, :3])
    rad = (np.float32(0.002) + np.float32(0.01) * v[:, 3]).astype(np.float32)
    return pos, rad


def test_cpp_client_matches_oracle(require_gpu, tmp_path):
    n, W, H = 20000, 320, 240
    out = tmp_path / "spheres.ppm"
    print(f"[app] {APP} {n} {W}x{H}", flush=True)
    r = subprocess.run([APP, str(n), str(W), str(H), str(out)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    kv = dict(f.split("=", 1) for f in r.stdout.split())
    assert kv["errors"] == "0" and int(kv["hits"]) > 0
    assert out.stat().st_size == len(f"P6\n{W} {H}\n255\n") + 3 * W * H

    pos, rad = _app_spheres(n)
    sc = S.SceneDesc(geoms=[S.GeomDesc("sphere", pos, radius=rad, surface_id=7,
                                       color=(0.9, 0.6, 0.3))], world_geoms=[0])
    cam = S.CameraDesc((0.5, 0.5, -1.2), (0.0, 0.0, 1.0), (0, 1, 0), math.pi / 3.0, W / H)
    o = orc.OracleScene(sc).render(cam, W, H, background=(0.1, 0.1, 0.1, 1.0))
    hit = (o["depth"] < 1e29).reshape(-1)
    ids = o["primitiveId"].reshape(-1)[hit]
    cs = 0
    for v in ids.tolist():
        cs = (cs * 1099511628211 + v) & M64
    assert int(kv["hits"]) == int(hit.sum())
    assert kv["checksum"] == f"{cs:016x}"


def test_cpp_client_direct_light(require_gpu, tmp_path):
    """The same client with the directLight renderer and a directional light
    created through the C API: its sRGB image against the oracle's, within
    one code value."""
    n, W, H = 20000, 160, 120
    out = tmp_path / "lit.ppm"
    print(f"[app] {APP} {n} {W}x{H} directLight", flush=True)
    r = subprocess.run([APP, str(n), str(W), str(H), str(out), "directLight"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    kv = dict(f.split("=", 1) for f in r.stdout.split())
    assert kv["errors"] == "0"
    data = out.read_bytes()
    head = f"P6\n{W} {H}\n255\n".encode()
    assert data.startswith(head)
    img = np.frombuffer(data[len(head):], np.uint8).reshape(H, W, 3)[::-1]   # bottom row first
    pos, rad = _app_spheres(n)
    sc = S.SceneDesc(geoms=[S.GeomDesc("sphere", pos, radius=rad, surface_id=7,
                                       color=(0.9, 0.6, 0.3))], world_geoms=[0],
                     lights=[S.LightDesc("directional", direction=(0.3, -1.0, 0.5),
                                         irradiance=2.0)])
    cam = S.CameraDesc((0.5, 0.5, -1.2), (0.0, 0.0, 1.0), (0, 1, 0), math.pi / 3.0, W / H)
    o = orc.OracleScene(sc).render(cam, W, H, kind="directLight", color_format=orc.ORC_COLOR_SRGB8,
                                   background=(0.1, 0.1, 0.1, 1.0), ambientRadiance=0.2)
    ref = o["color"].view(np.uint8).reshape(H, W, 4)[..., :3]
    assert np.abs(img.astype(np.int32) - ref.astype(np.int32)).max() <= 1
    assert img.std() > 5
