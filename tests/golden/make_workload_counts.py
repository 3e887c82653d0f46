"""Writes tests/golden/workload_counts.json: per-workload traversal and sample
counts from the CPU oracle, which define the algorithmic bytes of the
roofline figures (BVH2 internal nodes visited and primitives tested per ray
on the oracle's LBVH, near child first; FP64 volume sample counts).

Run from the repository root:  python tests/golden/make_workload_counts.py
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402
from visrtx_amd import scenes as S  # noqa: E402


def traversal_counts(sc, cam, w, h):
    o = orc.OracleScene(sc)
    t = time.time()
    out = o.render(cam, w, h, "raycast")
    st = out["stats"]
    rays = int(st[2])
    return {"width": w, "height": h, "rays": rays, "internal_nodes": int(st[0]),
            "prims_tested": int(st[1]), "nodes_per_ray": st[0] / rays,
            "prims_per_ray": st[1] / rays, "hit_fraction": float((out["depth"] < 1e29).mean()),
            "oracle_seconds": time.time() - t}


def main():
    res = {}
    sc, cam = S.heightfield_workload()
    res["heightfield_1024x768"] = traversal_counts(sc, cam, 1024, 768)
    print(res["heightfield_1024x768"], flush=True)
    sc, cam = S.spheres_workload()
    res["spheres10m_1920x1080"] = traversal_counts(sc, cam, 1920, 1080)
    print(res["spheres10m_1920x1080"], flush=True)
    # the batched scene: views 0 and 37 of the 64 (BLAS nodes and triangles;
    # instance-level nodes and transforms are not counted)
    sc, cams = S.batched_workload()
    o = orc.OracleScene(sc)
    t = time.time()
    nodes = prims = rays = 0
    for v in (0, 37):
        st = o.render(cams[v], 512, 512, "raycast")["stats"]
        nodes, prims, rays = nodes + int(st[0]), prims + int(st[1]), rays + int(st[2])
    res["batched_views0_37_512"] = {"width": 512, "height": 512, "views": [0, 37], "rays": rays,
                                    "internal_nodes": nodes, "prims_tested": prims,
                                    "nodes_per_ray": nodes / rays, "prims_per_ray": prims / rays,
                                    "oracle_seconds": time.time() - t}
    print(res["batched_views0_37_512"], flush=True)
    o.close()
    sc, cam = S.volume_workload()
    t = time.time()
    _, ns = orc.volume_image(sc.volumes[0], cam, 1920, 1080, inv_sampling_rate=1.0,
                             early_stop=True)
    res["volume512_1920x1080"] = {"width": 1920, "height": 1080, "samples": int(ns),
                                  "oracle_seconds": time.time() - t}
    print(res["volume512_1920x1080"], flush=True)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "workload_counts.json"),
              "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
