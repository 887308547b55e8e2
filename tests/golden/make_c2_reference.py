#!/usr/bin/env python3
"""Writes tests/golden/c2_reference_crops.npz: the FP64 CPU oracle at 4096
spp on 16 fixed 32x32 crops of the procedural 1080p scene (5.24 M triangles,
env map + 4 triangle lights, depth 6), shot 0.

This is the converged reference that relMSE is measured against:
    relMSE = mean over pixels and RGB of (x - ref)^2 / (ref^2 + 1e-2)

The scene comes from the library's seeded generator (master seed
0xA5A52026), so the GPU side regenerates the identical scene files.
Run from the repository root after building: python tests/golden/make_c2_reference.py
(about 2-4 minutes on 8 cores).
"""
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import asuna_amd as A  # noqa: E402
from oracle import oracle as O  # noqa: E402

SPP = 4096
CROP = 32


def crop_origins(w, h):
    """4x4 grid of crop origins, centred in the cells of a 4x4 tiling."""
    out = []
    for j in range(4):
        for i in range(4):
            out.append((int((i + 0.5) * w / 4) - CROP // 2, int((j + 0.5) * h / 4) - CROP // 2))
    return out


def main():
    threads = int(os.environ.get("THREADS", os.cpu_count() or 1))
    sdir = os.path.join(tempfile.gettempdir(), "asuna_bench_c2_s0")
    path = os.path.join(sdir, "spheres.json")
    if not os.path.exists(path):
        path = A.generate_scene(1, sdir)
    scene = A.HostScene(path)
    w, h = scene.film
    cam, pc = scene.shot(0)
    osc = O.OracleScene(scene.desc)
    images = np.zeros((9, h, w, 4), np.float32)
    origins = crop_origins(w, h)
    ref = np.zeros((len(origins), CROP, CROP, 3), np.float32)
    t0 = time.time()
    for k, (x0, y0) in enumerate(origins):
        osc.render(cam, pc, w, h, 0, SPP, images, crop=(x0, y0, x0 + CROP, y0 + CROP), precision=1,
                   threads=threads)
        ref[k] = images[0, y0:y0 + CROP, x0:x0 + CROP, :3]
        print(f"crop {k} at ({x0},{y0}) done, {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "c2_reference_crops.npz"), ref=ref,
                        origins=np.array(origins, np.int32), crop=np.int32(CROP), spp=np.int32(SPP),
                        width=np.int32(w), height=np.int32(h), triangles=np.int64(scene.state.num_triangles))


if __name__ == "__main__":
    main()
