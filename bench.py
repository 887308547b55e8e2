#!/usr/bin/env python3
"""Benchmark of the render hot path on one MI355X.

Workload (default): the procedural 1080p scene -- 8x8 grid of level-6
icospheres (81,920 triangles each) displaced by 6-octave fBm, ground quad,
2048x1024 env map (sky gradient + 0.5 deg sun disk) + 4 triangle lights;
pbr_metalness_roughness / dielectric materials; 1920x1080, 16 spp, max depth 6,
FP32.  The scene is generated with the fixed master seed, written as real
JSON + OBJ + EXR files and loaded through the library's loader.

One step = one full offline render of the shot: 16 sample-frames of every
pixel (33.2 M paths), from ray generation to the accumulated radiance/AOV
images in HBM.  Scene upload and BVH build happen before the timed region.

Prints one JSON line:
  value        Mpaths/s over the timed steps (camera samples per second)
  roofline     HBM roofline of the closest-hit traversal kernel: algorithmic
               bytes per launch (node visits x 64 B + triangle tests x 48 B +
               68 B ray/hit I/O per ray, counted in a calibration render) over
               its average launch time (HIP events on the render stream)
  cpu_baseline the CPU oracle (the same estimator, FP32, own SAH BVH) on the
               host cores over a bounded sample of the same workload
"""
import argparse
import json
import os
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "Mpaths/s & Mrays/s at 1920x1080; p50/p99 job latency at equal relMSE"
HBM_PEAK_GBS = 8000.0
NODE_BYTES, TRI_BYTES, RAY_IO_BYTES = 64, 48, 68


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", choices=["c2", "c1", "c3"], default="c2")
    ap.add_argument("--scale", type=int, default=0, help="shrink the procedural scene (tests)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=2)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--queue-paths", type=int, default=0)
    ap.add_argument("--scene-dir", default="")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 256 MiB device-to-device copy first (calibrates rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding across GPUs is designed but not built)", file=sys.stderr)
        sys.exit(2)

    import asuna_amd as A

    if args.copy_probe:
        import torch  # plumbing only: a plain copy kernel of known size
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: {src.numel() * 4} bytes read + written")
        del src, dst

    cfg_id = {"c1": 0, "c2": 1, "c3": 2}[args.config]
    sdir = args.scene_dir or os.path.join(tempfile.gettempdir(), f"asuna_bench_{args.config}_s{args.scale}")
    t0 = time.time()
    scene_path = os.path.join(sdir, {0: "box.json", 1: "spheres.json", 2: "multishot.json"}[cfg_id])
    if not os.path.exists(scene_path):
        scene_path = A.generate_scene(cfg_id, sdir, scale=args.scale)
    t_gen = time.time() - t0
    t0 = time.time()
    scene = A.HostScene(scene_path)
    t_load = time.time() - t0
    w, h = scene.film
    nshots = scene.state.num_shots if cfg_id == 2 else 1
    cams, pcs = zip(*[scene.shot(i) for i in range(nshots)])
    spp = pcs[0].spp
    log(f"[bench] scene {scene_path}: {scene.state.num_triangles} tris, {w}x{h}, {spp} spp, "
        f"depth {pcs[0].maxPathDepth}, gen {t_gen:.1f}s load {t_load:.1f}s")

    r = A.Renderer(0)
    t0 = time.time()
    r.upload(scene)
    t_up = time.time() - t0
    binfo = r.build()
    log(f"[bench] upload {t_up:.2f}s, BVH {binfo}")
    r.set_film(w, h)
    dch = scene.state.depth_out_channel

    def step():
        if cfg_id == 2:
            return r.render_shots(list(cams), list(pcs), spp, queue_paths=args.queue_paths)
        r.set_film(w, h)
        return r.render(cams[0], pcs[0], 0, spp, queue_paths=args.queue_paths, depth_channel=dch)

    # calibration: traversal work per ray (not timed)
    cal = r.render(cams[0], pcs[0], 0, 1, queue_paths=args.queue_paths, count_traversal=True, depth_channel=dch)
    ext_rays = max(1, cal["extension_rays"])
    nodes_per_ray = cal["ext_node_visits"] / ext_rays
    tris_per_ray = cal["ext_tri_tests"] / ext_rays
    bytes_per_ray = nodes_per_ray * NODE_BYTES + tris_per_ray * TRI_BYTES + RAY_IO_BYTES
    log(f"[bench] calibration: {nodes_per_ray:.2f} nodes, {tris_per_ray:.2f} tris per closest-hit ray "
        f"-> {bytes_per_ray:.0f} B/ray")

    for i in range(args.warmup):
        step()
        log(f"[bench] warmup {i} done")
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    stats = []
    for i in range(args.steps):
        stats.append(step())
    hip.hipDeviceSynchronize()
    elapsed = time.perf_counter() - t0
    paths = sum(s["paths"] for s in stats)
    rays = sum(s["extension_rays"] + s["shadow_rays"] for s in stats)
    ms_per_step = elapsed * 1e3 / args.steps
    mpaths = paths / elapsed / 1e6
    mrays = rays / elapsed / 1e6
    log(f"[bench] {args.steps} steps in {elapsed:.3f}s: {mpaths:.1f} Mpaths/s, {mrays:.1f} Mrays/s")

    # per-kernel timing pass (HIP events around every launch; not the timed steps)
    if cfg_id == 2:
        tp = r.render_shots(list(cams), list(pcs), spp, queue_paths=args.queue_paths, timing=True)
    else:
        r.set_film(w, h)
        tp = r.render(cams[0], pcs[0], 0, spp, queue_paths=args.queue_paths, timing=True, depth_channel=dch)
    ext_ms_per_launch = tp["extend_ms"] / max(1, tp["extend_launches"])
    rays_per_launch = tp["extension_rays"] / max(1, tp["extend_launches"])
    bytes_per_launch = rays_per_launch * bytes_per_ray
    achieved = bytes_per_launch / (ext_ms_per_launch * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(HERE, "profiles", f"traffic_extend_{args.config}.json")
    if os.path.exists(tfile) and args.scale == 0:
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "kExtend (closest-hit traversal)", "avg_launch_ms": round(ext_ms_per_launch, 4),
                "bytes_per_ray": round(bytes_per_ray, 1), "nodes_per_ray": round(nodes_per_ray, 3),
                "tris_per_ray": round(tris_per_ray, 3), "rays_per_launch": round(rays_per_launch)}

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        osc = O.OracleScene(scene.desc)
        nf = args.cpu_frames
        t0 = time.perf_counter()
        _, ost = osc.render(cams[0], pcs[0], w, h, 0, nf, threads=threads)
        ct = time.perf_counter() - t0
        cpu = {"value": round(ost["paths"] / ct / 1e6, 4), "unit": "Mpaths/s", "cores": threads, "kind": "port",
               "sample": f"frames 0..{nf - 1} of shot 0 at {w}x{h} ({ost['paths']} paths, {ct:.1f} s), "
                         f"CPU oracle FP32 with its own binned-SAH BVH, {threads} threads",
               "mrays_per_s": round((ost["extension_rays"] + ost["shadow_rays"]) / ct / 1e6, 3)}
        log(f"[bench] cpu baseline {cpu}")

    workload = {0: "box 512x512, 64 spp, depth 8 (36 tris)",
                1: "procedural 1920x1080, 5.24M tris (8x8 level-6 fBm icospheres), env 2048x1024 + 4 tri "
                   "lights, 16 spp, depth 6",
                2: "64 shots of the box scene, 512x512, 32 spp, depth 6, one batched launch"}[cfg_id]
    if args.scale:
        workload += f" (reduced scale {args.scale})"
    out = {"metric": METRIC, "value": round(mpaths, 2), "unit": "Mpaths/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded procedural scene files, master seed 0xA5A52026)",
           "config": {"workload": workload, "width": w, "height": h, "spp": spp,
                      "max_path_depth": pcs[0].maxPathDepth, "triangles": scene.state.num_triangles,
                      "shots": nshots, "bvh": binfo},
           "mrays_per_s": round(mrays, 1),
           "rays_per_path": round(rays / max(1, paths), 3),
           "roofline": roofline, "cpu_baseline": cpu,
           "gpu_over_cpu": round(mpaths / cpu["value"], 1) if cpu else None}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
