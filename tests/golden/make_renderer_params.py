"""Writes renderer_params.json: the renderer subtypes and their parameters
(name, type, default) as the reference device declares them in
devices/rtx/visrtx_device.json.  That file is a CMake template: lines that
are generator expressions are dropped before it is read as JSON.  Run in a
container that has the reference checkout:

    python tests/golden/make_renderer_params.py /root/reference
"""
import json
import os
import re
import sys


def main(ref):
    text = open(os.path.join(ref, "devices/rtx/visrtx_device.json")).read()
    text = "\n".join(ln for ln in text.split("\n") if not ln.strip().startswith("$<"))
    text = re.sub(r",(\s*[\]\}])", r"\1", text)
    spec = json.loads(text)
    out = {}
    for obj in spec["objects"]:
        if obj["type"] != "ANARI_RENDERER":
            continue
        out[obj["name"]] = [{"name": p["name"], "type": p["types"][0], "default": p.get("default")}
                            for p in obj.get("parameters", [])]
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "renderer_params.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
