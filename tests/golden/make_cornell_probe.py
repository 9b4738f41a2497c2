"""Writes tests/golden/cornell_probe.json: the Cornell probe workload's sizes,
SHA-256 of its compiled f32 positions and BVH nodes, and 64 sampled pixels of
the f64 brute-force probe (the oracle) at 256x256.

    python tests/golden/make_cornell_probe.py
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT / "luce-render_amd"), str(ROOT)]

from luce_render import render, scenes  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    sc = render.compile_scene(scenes.cornell_probe_set())
    f64 = O.ray_frame(sc.cameras[0].matrix, False, 36.0, 30.0, 256, 256)
    face, dist, margin = O.brute_force(sc.positions, sc.indices, sc.faces, frame=f64, width=256, height=256,
                                       threads=8)
    rng = np.random.default_rng(5)
    pix = sorted(rng.choice(256 * 256, 64, replace=False).tolist())
    out = {
        "triangles": int(sc.triangle_count()),
        "polygons": int(sc.faces.max()) + 1,
        "positions_sha256": hashlib.sha256(np.ascontiguousarray(sc.positions).tobytes()).hexdigest(),
        "nodes_sha256": hashlib.sha256(np.ascontiguousarray(sc.nodes).tobytes()).hexdigest(),
        "node_count": int(len(sc.nodes)),
        "hit_pixels": int((face >= 0).sum()),
        "samples": [{"pixel": p, "face": int(face[p]), "distance": float(dist[p]), "margin": float(min(margin[p], 1e30))}
                    for p in pix],
    }
    (Path(__file__).resolve().parent / "cornell_probe.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
