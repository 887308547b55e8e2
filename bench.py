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
  roofline     HBM roofline of the dominant traversal kernel: on the default
               scheme the fused trace launch (shadow rays of one bounce and
               extension rays of the next), otherwise the closest-hit kernel.
               achieved = the HBM bytes per launch that rocprofv3 counts for
               that kernel in a counter pass of this run (one child render
               under --pmc, L2-to-memory read and write requests priced by
               request size) over the kernel's average launch time (HIP events
               on the render stream), frac = achieved / 8 TB/s; beside it the
               VALU issue and lane utilisation of the same dispatches.  The
               algorithmic bytes per launch (node visits x node bytes: 64 B for
               the quantized 4-wide node, 112 B for the float 4-wide node, 64 B
               binary; + triangle tests x 48 B + ray I/O, 68 B per extension
               ray, 48 B per shadow ray; counted in a calibration render of the
               same batches) are reported as algorithmic_bytes: the L2 and the
               MALL serve most of them, so over the launch time they exceed the
               HBM peak
  cpu_baseline the CPU oracle (the same estimator, FP32, own SAH BVH) on the
               host cores over a bounded sample of the same workload
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "Mpaths/s & Mrays/s at 1920x1080; p50/p99 job latency at equal relMSE"
HBM_PEAK_GBS = 8000.0
# bytes a node visit reads: binary node 64 B; float 4-wide node 7 x 16 B (6
# box planes + child refs) of its 128 B line; quantized 4-wide node 64 B
NODE_BYTES = {(2, 64): 64, (4, 128): 112, (4, 64): 64}
TRI_BYTES, RAY_IO_BYTES = 48, 68
L2_BYTES = 4 << 20  # L2 of one XCD
# shadow ray I/O: origin+tmax, direction+path id, and the pending NEE radiance
SHADOW_IO_BYTES = 48
# counters of the in-run pass: L2-to-memory (EA) read and write requests by
# size, VALU issue and lane activity, GPU cycles (within one pass's limits:
# 4 TCC, 4 SQ, 1 GRBM)
PMC_COUNTERS = ["TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_32B_sum", "TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum",
                "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_WAVES", "GRBM_GUI_ACTIVE"]
COPY_PROBE_BYTES = 256 << 20
# U-Net FLOPs per full-resolution pixel: 2 x 9 x cin x cout over every 3x3
# convolution with the network's real channel counts, each level weighted by
# its pixel share (1 / 4^level); the kernels multiply 32-padded channels,
# 252,288 FLOPs per pixel
UNET_FLOP_PER_PIXEL = 181_440
BF16_PEAK_TFLOPS = 2500.0
XCDS, SIMDS_PER_CU = 8, 4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def copy_probe():
    """A plain device-to-device copy of COPY_PROBE_BYTES through the HIP
    runtime (the runtime's copy kernel), before anything else runs on the
    device: the counter pass checks its request counts against the bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(COPY_PROBE_BYTES)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(COPY_PROBE_BYTES)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(COPY_PROBE_BYTES)) == 0
    assert hip.hipMemcpy(dst, src, ctypes.c_size_t(COPY_PROBE_BYTES), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"[bench] copy probe: {COPY_PROBE_BYTES} bytes read + written")


def counter_pass(args, qp, bpc, lds, trav, kernel_prefix):
    """Runs one step of the same render (same batches and launch shape) in a
    child process under rocprofv3 --pmc and returns the per-dispatch HBM
    bytes and VALU activity of the kernel whose name contains
    `kernel_prefix`.  Read requests are priced at 128 B (32 B for RDREQ_32B),
    writes at 64 B for WRREQ_64B and 32 B otherwise; the child first copies
    COPY_PROBE_BYTES, whose counted bytes must equal what it moved.  Returns
    None (with a log line) when rocprofv3 is missing or the pass fails."""
    import csv
    import glob
    import shutil
    import subprocess
    exe = shutil.which("rocprofv3")
    if not exe:
        log("[bench] counter pass skipped: rocprofv3 not found")
        return None
    d = tempfile.mkdtemp(prefix="asuna_pmc_")
    child = [sys.executable, os.path.abspath(__file__), "--config", args.config, "--steps", "1", "--warmup", "0",
             "--no-cpu-baseline", "--no-pmc", "--no-autotune", "--copy-probe", "--queue-paths", str(qp),
             "--blocks-per-cu", str(bpc), "--lds-stack", str(lds), "--traversal", str(trav),
             "--branching", str(args.branching), "--quality", str(args.quality),
             "--node-format", str(args.node_format), "--scale", str(args.scale)]
    if args.scene_dir:
        child += ["--scene-dir", args.scene_dir]
    if args.no_graphs:
        child.append("--no-graphs")
    cmd = ["timeout", "-s", "KILL", "180", exe, "--pmc", *PMC_COUNTERS, "--output-format", "csv", "-d", d,
           "-o", "pmc", "--", *child]
    t0 = time.time()
    try:
        res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=240)
    except (OSError, subprocess.TimeoutExpired) as e:
        log(f"[bench] counter pass failed: {e}")
        return None
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if res.returncode != 0 or not files:
        log(f"[bench] counter pass failed (rc {res.returncode}): {res.stderr[-400:]}")
        return None
    per = {}
    with open(files[0]) as f:
        for row in csv.DictReader(f):
            key = int(row["Dispatch_Id"])
            e = per.setdefault(key, {"name": row["Kernel_Name"]})
            e[row["Counter_Name"]] = float(row["Counter_Value"])

    def nbytes(e):
        rd32 = e.get("TCC_EA0_RDREQ_32B_sum", 0.0)
        rd = (e.get("TCC_EA0_RDREQ_sum", 0.0) - rd32) * 128 + rd32 * 32
        w64 = e.get("TCC_EA0_WRREQ_64B_sum", 0.0)
        wr = w64 * 64 + (e.get("TCC_EA0_WRREQ_sum", 0.0) - w64) * 32
        return rd, wr

    order = sorted(per)
    copies = [per[k] for k in order if "copyBuffer" in per[k]["name"]]
    copy_ratio = None
    if copies:
        rd, wr = nbytes(copies[0])  # the probe is the process's first copy
        copy_ratio = [round(rd / COPY_PROBE_BYTES, 4), round(wr / COPY_PROBE_BYTES, 4)]
    ks = [per[k] for k in order if kernel_prefix in per[k]["name"]]
    if not ks:
        log(f"[bench] counter pass: no {kernel_prefix} dispatch found")
        return None
    rds, wrs = zip(*[nbytes(e) for e in ks])
    n = len(ks)
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    ncu = ctypes.c_int()
    hip.hipDeviceGetAttribute(ctypes.byref(ncu), 63, 0)  # hipDeviceAttributeMultiprocessorCount
    simds = max(1, ncu.value) * SIMDS_PER_CU
    insts = sum(e.get("SQ_INSTS_VALU", 0.0) for e in ks)
    active = sum(e.get("SQ_ACTIVE_INST_VALU", 0.0) for e in ks)
    lanes = sum(e.get("SQ_THREAD_CYCLES_VALU", 0.0) for e in ks)
    cycles = sum(e.get("GRBM_GUI_ACTIVE", 0.0) for e in ks) / XCDS  # GRBM counts every XCD's cycles
    out = {"read_bytes": sum(rds) / n, "write_bytes": sum(wrs) / n, "dispatches": n, "copy_ratio": copy_ratio,
           # a wave64 VALU instruction occupies a 16-lane SIMD for 4 cycles
           "valu_issue": round(insts * 4 / (simds * cycles), 4) if cycles else None,
           "lane_util": round(lanes / (64 * active), 4) if active else None}
    out["hbm_bytes"] = out["read_bytes"] + out["write_bytes"]
    log(f"[bench] counter pass ({time.time() - t0:.1f}s, {n} dispatches): {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", choices=["c2", "c1", "c3", "c4", "c4-infer", "c5"], default="c2",
                    help="c4-infer: only the denoiser's 1080p inference on the procedural scene's 4-spp film")
    ap.add_argument("--train-iters", type=int, default=4000, help="c4: denoiser training steps")
    ap.add_argument("--train-scenes", type=int, default=200, help="c4: seeded random training scenes (10 views each)")
    ap.add_argument("--jobs", type=int, default=2000, help="c5: jobs per arrival rate")
    ap.add_argument("--rates", default="10,50,200", help="c5: Poisson arrival rates (jobs/s)")
    ap.add_argument("--serve-policy", type=int, default=0,
                    help="c5: server batch fill order (0 oldest first across scenes, 1 round robin, 2 oldest first)")
    ap.add_argument("--serve-queue", type=int, default=1 << 25, help="c5: server paths per batch")
    ap.add_argument("--jobs-log", default="", help="c5: write one JSON line per job (both columns) to this file")
    ap.add_argument("--matched", default="",
                    help="c5 at equal relMSE: plain spp -> denoised spp map (e.g. 1:1,4:1,16:2,64:64p, from c4; "
                         "a p suffix: no denoised spp reaches that relMSE, render plain); the server renders every "
                         "job at the mapped spp (denoised unless p), the one-at-a-time baseline renders the plain spp")
    ap.add_argument("--scale", type=int, default=0, help="shrink the procedural scene (tests)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=16)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--queue-paths", type=int, default=0)
    ap.add_argument("--scene-dir", default="")
    ap.add_argument("--branching", type=int, choices=[2, 4], default=4)
    ap.add_argument("--quality", type=int, choices=[0, 1, 2], default=2,
                    help="BVH builder: 0 LBVH, 1 LBVH + SAH leaf collapse, 2 PLOC + SAH leaf collapse")
    ap.add_argument("--node-format", type=int, choices=[0, 1, 2], default=0,
                    help="4-wide nodes: 0 default (float), 1 float 128 B, 2 quantized 64 B")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--blocks-per-cu", type=int, default=0, help="persistent traversal blocks per CU when not autotuned")
    ap.add_argument("--lds-stack", type=int, choices=[0, 8, 12], default=0,
                    help="closest-hit stack entries in LDS when not autotuned")
    ap.add_argument("--traversal", type=int, choices=[0, 1, 2], default=0,
                    help="ray fetch of the traversal kernels when not autotuned: 0 per-XCD queue slices, "
                         "1 queue order, 2 per-lane refill")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 256 MiB device-to-device copy first (checks the request sizes of the counter pass)")
    ap.add_argument("--no-pmc", action="store_true",
                    help="skip the rocprofv3 counter pass that measures the traversal kernel's HBM traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding across GPUs is designed but not built)", file=sys.stderr)
        sys.exit(2)

    import asuna_amd as A

    if args.copy_probe:
        copy_probe()
    if args.config == "c5":
        return run_serving(A, args)
    if args.config == "c4":
        return run_denoiser(A, args)
    if args.config == "c4-infer":
        return run_denoiser_inference(A, args)

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
    binfo = r.build(branching=args.branching, node_format=args.node_format, quality=args.quality)
    log(f"[bench] upload {t_up:.2f}s, BVH {binfo}")
    r.set_film(w, h)
    dch = scene.state.depth_out_channel

    tune = None
    qp, bpc, lds, trav = args.queue_paths, args.blocks_per_cu, args.lds_stack, args.traversal
    if not args.no_autotune and not args.queue_paths:
        # probe with the frames of a whole step when they fit in 2^25 paths, so that every queue
        # size is tried on the batch split the timed steps will use
        probe = max(1, min(spp, (1 << 25) // (w * h * nshots)))
        opts, tune = r.autotune(list(cams), list(pcs), probe_frames=probe, traversal=args.traversal)
        tune["probe_frames"] = probe
        qp, bpc, lds, trav = opts.queue_paths, opts.blocks_per_cu, opts.lds_stack, opts.traversal
        log(f"[bench] autotune: {tune}")
    graphs = not args.no_graphs

    def step():
        if cfg_id == 2:
            return r.render_shots(list(cams), list(pcs), spp, queue_paths=qp, blocks_per_cu=bpc, graphs=graphs,
                                  lds_stack=lds, traversal=trav)
        r.set_film(w, h)
        return r.render(cams[0], pcs[0], 0, spp, queue_paths=qp, blocks_per_cu=bpc, depth_channel=dch,
                        graphs=graphs, lds_stack=lds, traversal=trav)

    # calibration: traversal work of the same batches as a timed step (counting kernels, not timed)
    if cfg_id == 2:
        cal = r.render_shots(list(cams), list(pcs), spp, queue_paths=qp, blocks_per_cu=bpc, count_traversal=True,
                             lds_stack=lds, traversal=trav)
    else:
        r.set_film(w, h)
        cal = r.render(cams[0], pcs[0], 0, spp, queue_paths=qp, blocks_per_cu=bpc, count_traversal=True,
                       depth_channel=dch, lds_stack=lds, traversal=trav)
    ext_rays = max(1, cal["extension_rays"])
    nodes_per_ray = cal["ext_node_visits"] / ext_rays
    tris_per_ray = cal["ext_tri_tests"] / ext_rays
    node_bytes = NODE_BYTES[(binfo["branching"], binfo["node_bytes"])]
    bytes_per_ray = nodes_per_ray * node_bytes + tris_per_ray * TRI_BYTES + RAY_IO_BYTES
    log(f"[bench] calibration: {nodes_per_ray:.2f} nodes, {tris_per_ray:.2f} tris per closest-hit ray "
        f"-> {bytes_per_ray:.0f} B/ray")

    for i in range(args.warmup):
        step()
        log(f"[bench] warmup {i} done")
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    stats, lat = [], []
    for i in range(args.steps):
        ts = time.perf_counter()
        stats.append(step())  # returns after the render's last kernel has finished
        lat.append((time.perf_counter() - ts) * 1e3)
    hip.hipDeviceSynchronize()
    elapsed = time.perf_counter() - t0
    paths = sum(s["paths"] for s in stats)
    rays = sum(s["extension_rays"] + s["shadow_rays"] for s in stats)
    ms_per_step = elapsed * 1e3 / args.steps
    mpaths = paths / elapsed / 1e6
    mrays = rays / elapsed / 1e6
    log(f"[bench] {args.steps} steps in {elapsed:.3f}s: {mpaths:.1f} Mpaths/s, {mrays:.1f} Mrays/s")

    # relMSE of the last timed render against the FP64 oracle at 4096 spp on
    # 16 fixed 32x32 crops: mean of (x - ref)^2 / (ref^2 + 1e-2) over pixels and RGB
    relmse = None
    gold = os.path.join(HERE, "tests", "golden", "c2_reference_crops.npz")
    if cfg_id == 1 and args.scale == 0 and os.path.exists(gold):
        g = np.load(gold)
        img = r.read(0)
        c = int(g["crop"])
        x = np.stack([img[y0:y0 + c, x0:x0 + c, :3] for x0, y0 in g["origins"]])
        ref = g["ref"]
        relmse = float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))
        log(f"[bench] relMSE vs FP64 oracle at {int(g['spp'])} spp (16 crops of {c}x{c}): {relmse:.5f}")

    # per-kernel timing pass (HIP events around every launch; not the timed steps)
    if cfg_id == 2:
        tp = r.render_shots(list(cams), list(pcs), spp, queue_paths=qp, blocks_per_cu=bpc, timing=True, lds_stack=lds,
                            traversal=trav)
    else:
        r.set_film(w, h)
        tp = r.render(cams[0], pcs[0], 0, spp, queue_paths=qp, blocks_per_cu=bpc, timing=True, depth_channel=dch,
                      lds_stack=lds, traversal=trav)
    tree = f"bvh{binfo['branching']}" + ("q" if binfo["branching"] == 4 and binfo["node_bytes"] == 64 else "")
    fused = tp["trace_launches"] > 0 and cal["trace_launches"] == tp["trace_launches"]
    if fused:
        # kTraceDyn: the shadow rays of bounce k and the extension rays of bounce k + 1
        n = cal["trace_launches"]
        ext_b = (cal["trace_ext_node_visits"] * node_bytes + cal["trace_ext_tri_tests"] * TRI_BYTES
                 + cal["trace_ext_rays"] * RAY_IO_BYTES)
        sh_b = (cal["trace_shadow_node_visits"] * node_bytes + cal["trace_shadow_tri_tests"] * TRI_BYTES
                + cal["trace_shadow_rays"] * SHADOW_IO_BYTES)
        bytes_per_launch = (ext_b + sh_b) / n
        ms_per_launch = tp["trace_ms"] / n
        kernel, tkey = "kTraceDyn (fused shadow + next-bounce closest-hit traversal)", "trace"
        per_launch = {"ext_rays_per_launch": round(cal["trace_ext_rays"] / n),
                      "shadow_rays_per_launch": round(cal["trace_shadow_rays"] / n),
                      "ext_nodes_per_ray": round(cal["trace_ext_node_visits"] / max(1, cal["trace_ext_rays"]), 3),
                      "ext_tris_per_ray": round(cal["trace_ext_tri_tests"] / max(1, cal["trace_ext_rays"]), 3),
                      "shadow_nodes_per_ray": round(cal["trace_shadow_node_visits"] /
                                                    max(1, cal["trace_shadow_rays"]), 3),
                      "shadow_tris_per_ray": round(cal["trace_shadow_tri_tests"] / max(1, cal["trace_shadow_rays"]), 3)}
    else:
        ms_per_launch = tp["extend_ms"] / max(1, tp["extend_launches"])
        rays_per_launch = tp["extension_rays"] / max(1, tp["extend_launches"])
        bytes_per_launch = rays_per_launch * bytes_per_ray
        kernel, tkey = "kExtend (closest-hit traversal)", "extend"
        per_launch = {"rays_per_launch": round(rays_per_launch)}
    algo_gbs = bytes_per_launch / (ms_per_launch * 1e-3) / 1e9
    pmc = None
    if not args.no_pmc:
        pmc = counter_pass(args, qp, bpc, lds, trav, "kTraceDyn<false" if fused else "kExtend")
    roofline = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": None, "kernel": kernel, "avg_launch_ms": round(ms_per_launch, 4),
                "algorithmic_bytes": round(bytes_per_launch), "algorithmic_gbs": round(algo_gbs, 1),
                "node_bytes": node_bytes, "bytes_per_ray": round(bytes_per_ray, 1),
                "nodes_per_ray": round(nodes_per_ray, 3), "tris_per_ray": round(tris_per_ray, 3), **per_launch}
    if pmc:
        achieved = pmc["hbm_bytes"] / (ms_per_launch * 1e-3) / 1e9
        roofline.update({"achieved": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": round(pmc["hbm_bytes"]), "traffic_read": round(pmc["read_bytes"]),
                         "traffic_write": round(pmc["write_bytes"]),
                         "traffic_source": "rocprofv3 --pmc pass of this run: TCC_EA0_RDREQ x 128 B "
                                           "(- RDREQ_32B x 96 B) + TCC_EA0_WRREQ_64B x 64 B + the other WRREQ x 32 B, "
                                           "per dispatch of the kernel",
                         "copy_probe_ratio": pmc["copy_ratio"],
                         "valu_issue_frac": pmc["valu_issue"], "valu_lane_util": pmc["lane_util"],
                         "pmc_dispatches": pmc["dispatches"],
                         "note": "the kernel sits nearer its VALU issue roof (valu_issue_frac) than the HBM roof "
                                 "(frac); algorithmic_gbs above the HBM peak is cache service"})
    tree_bytes = binfo["num_wide_nodes"] * 128 + binfo["num_triangles"] * TRI_BYTES if binfo["branching"] == 4 \
        else binfo["num_nodes"] * 64 + binfo["num_triangles"] * TRI_BYTES
    if tree_bytes < L2_BYTES:
        roofline["note"] = (f"tree and triangles ({tree_bytes} B) fit in one XCD's 4 MiB L2: the algorithmic bytes "
                            f"are served from cache and the kernel is bound by issue, not HBM")

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        threads = args.cpu_threads or os.cpu_count() or 1
        osc = O.OracleScene(scene.desc)
        nf = args.cpu_frames
        t0 = time.perf_counter()
        _, ost = osc.render(cams[0], pcs[0], w, h, 0, nf, threads=threads)
        ct = time.perf_counter() - t0
        cpu = {"value": round(ost["paths"] / ct / 1e6, 4), "unit": "Mpaths/s", "cores": threads, "kind": "port",
               "sample": f"frames 0..{nf - 1} of shot 0 at {w}x{h} ({ost['paths']} paths, {ct:.1f} s), "
                         f"CPU oracle FP32 with its own binned-SAH BVH, {threads} threads",
               "nproc": os.cpu_count(), "cpu_model": cpu_model(),
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
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2), "p99": round(float(np.percentile(lat, 99)), 2),
                          "job": "one full render of the workload, submit to result in HBM"},
           "relmse": relmse,
           "relmse_ref": "FP64 CPU oracle, 4096 spp, 16 fixed 32x32 crops" if relmse is not None else None,
           "autotune": tune, "graphs": graphs,
           "batches_per_step": stats[-1]["batches"],
           "rays_per_path": round(rays / max(1, paths), 3),
           "roofline": roofline, "cpu_baseline": cpu,
           "gpu_over_cpu": round(mpaths / cpu["value"], 1) if cpu else None}
    print(json.dumps(out), flush=True)


def run_denoiser(A, args):
    """Learned denoiser on its training workload: 200 seeded random scenes
    (1-20 displaced meshes with materials from the plugin set, 1-4 rect
    lights, a random sky or constant env), 10 random 256x256 views each
    (2,000 crops), every view rendered at 4 spp (and 1, 16 and 64 spp)
    with its albedo / normal / depth planes, stored as FP16, against a
    4096-spp target.  The U-Net is trained on the GPU (Adam, batches of 8
    random 128x128 windows, half of them at the 4-spp input, the rest at
    1, 16 or 64 spp), then run on the 1080p procedural scene.  Reports
    training throughput, 1080p inference time and relMSE against the FP64
    oracle at 4096 spp (16 fixed 32x32 crops) for plain and denoised renders
    at 1 to 256 spp, and the equal-relMSE map: for each plain spp, the
    fewest denoised spp that reach its relMSE."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(0xA5A52026)
    t0 = time.time()
    nsc, views, res, spp_ref = args.train_scenes, 10, 256, 4096
    in_spp = (4, 1, 16, 64)

    def gen(k):
        d = os.path.join(tempfile.gettempdir(), f"asuna_train_{k}")
        p = os.path.join(d, "training.json")
        if not os.path.exists(p):
            p = A.generate_scene(3, d, seed=0xA5A52026 + k)
        return p

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        paths = list(ex.map(gen, range(nsc)))
    log(f"[den] {nsc} training scenes generated in {time.time() - t0:.1f}s")
    X = np.zeros((len(in_spp), nsc * views, res, res, 10), np.float16)
    Y = np.zeros((nsc * views, res, res, 3), np.float32)
    t1 = time.time()
    r = A.Renderer(0)
    for k, p in enumerate(paths):
        s = A.HostScene(p)
        st = s.state
        _, pc = s.shot(0)
        r.upload(s)
        r.build()
        r.set_film(res, res)
        cams = []
        for _ in range(views):
            ang = rng.uniform(0, 2 * np.pi)
            dist = rng.uniform(3.0, 7.0)
            eye = (dist * np.cos(ang), rng.uniform(0.5, 4.0), dist * np.sin(ang))
            look = (rng.uniform(-1.5, 1.5), rng.uniform(0.0, 0.8), rng.uniform(-1.5, 1.5))
            cams.append(A.camera_lookat(res, res, float(rng.uniform(30, 65)), eye, look))
        pcs = [pc] * views
        for i, spp in enumerate(in_spp):
            r.render_shots(cams, pcs, spp, depth_channel=st.depth_out_channel)
            for v in range(views):
                X[i, k * views + v] = np.concatenate(
                    [r.read_shot(v, 0)[..., :3], r.read_shot(v, pc.diffuseOutChannel + 1)[..., :3],
                     r.read_shot(v, pc.normalOutChannel + 1)[..., :3],
                     r.read_shot(v, st.depth_out_channel + 1)[..., :1]], -1)
        r.render_shots(cams, pcs, spp_ref)
        for v in range(views):
            Y[k * views + v] = r.read_shot(v, 0)[..., :3]
        s.close()
        if (k + 1) % 20 == 0:
            log(f"[den] rendered {k + 1}/{nsc} scenes ({time.time() - t1:.0f}s)")
    r.close()
    t_data = time.time() - t1
    n = len(Y)
    den = A.Denoiser(0, seed=1)
    nb, crop = 8, 128
    t2 = time.perf_counter()
    losses = []
    for it in range(args.train_iters):
        idx = rng.choice(n, nb, replace=False)
        lvl = np.where(rng.random(nb) < 0.5, 0, rng.integers(1, len(in_spp), nb))
        oy, ox = rng.integers(0, res - crop + 1, 2)
        xb = np.stack([X[l, i, oy:oy + crop, ox:ox + crop] for l, i in zip(lvl, idx)]).astype(np.float32)
        yb = Y[idx, oy:oy + crop, ox:ox + crop]
        if it % 2:
            xb, yb = xb[:, :, ::-1], yb[:, :, ::-1]
        lr = 2e-3 if it < 0.7 * args.train_iters else 4e-4
        losses.append(den.train_step(np.ascontiguousarray(xb), np.ascontiguousarray(yb), lr=lr))
        if it % 500 == 0:
            log(f"[den] step {it}: loss {np.mean(losses[-50:]):.4f}")
    t_train = time.perf_counter() - t2
    # held-out check on the last scene's views is not separate: the data set
    # is the training set; the judge of the network is the 1080p scene below
    sdir = os.path.join(tempfile.gettempdir(), "asuna_bench_c2_s0")
    p = os.path.join(sdir, "spheres.json")
    if not os.path.exists(p):
        p = A.generate_scene(1, sdir)
    s = A.HostScene(p)
    w, h = s.film
    cam, pc = s.shot(0)
    st = s.state
    r = A.Renderer(0)
    r.upload(s)
    r.build()
    g = np.load(os.path.join(HERE, "tests", "golden", "c2_reference_crops.npz"))
    c = int(g["crop"])

    def relmse(img):
        x = np.stack([img[y0:y0 + c, x0:x0 + c, :3] for x0, y0 in g["origins"]])
        return float(np.mean((x - g["ref"]) ** 2 / (g["ref"] ** 2 + 1e-2)))

    out = {}
    for spp in (1, 4, 16, 64, 256):
        r.set_film(w, h)
        st_r = r.render(cam, pc, 0, spp, depth_channel=st.depth_out_channel)
        out[f"plain_{spp}spp"] = {"relmse": round(relmse(r.read(0)), 5), "render_ms": round(st_r["render_ms"], 2)}
        if spp <= 64:
            den.denoise_film(r, pc.diffuseOutChannel, pc.normalOutChannel, st.depth_out_channel, 7)  # warm
            ms = min(den.denoise_film(r, pc.diffuseOutChannel, pc.normalOutChannel, st.depth_out_channel, 7)
                     for _ in range(3))
            out[f"denoised_{spp}spp"] = {"relmse": round(relmse(r.read(7)), 5), "denoise_ms": round(ms, 2),
                                         "total_ms": round(ms + st_r["render_ms"], 2)}
    # equal relMSE: for each plain spp, the fewest denoised spp whose relMSE is no higher
    matched = {}
    for sp in (1, 4, 16, 64, 256):
        ok = [d for d in (1, 4, 16, 64) if out[f"denoised_{d}spp"]["relmse"] <= out[f"plain_{sp}spp"]["relmse"]]
        if ok:
            d = min(ok)
            matched[str(sp)] = {"denoised_spp": d, "plain_ms": out[f"plain_{sp}spp"]["render_ms"],
                                "denoised_ms": out[f"denoised_{d}spp"]["total_ms"]}
    flop = UNET_FLOP_PER_PIXEL * 1920 * 1088
    res_line = {"metric": "denoiser: relMSE vs FP64 4096-spp oracle and 1080p inference time",
                "config": {"workload": f"U-Net 32/48/64/96/128, BF16 MFMA; trained on {n} 256x256 crops of {nsc} "
                                       "seeded random scenes (4-spp inputs, plus 1/16/64 spp, FP16; 4096-spp "
                                       "targets); inference on the 1920x1080 procedural scene (5.24M triangles)",
                           "train_iters": args.train_iters, "batch": f"{nb}x{crop}x{crop}"},
                "data_s": round(t_data, 1), "train_step_ms": round(t_train / args.train_iters * 1e3, 2),
                "train_loss_first_last": [round(float(np.mean(losses[:50])), 4), round(float(np.mean(losses[-50:])), 4)],
                "inference_tflops": round(flop / (out["denoised_4spp"]["denoise_ms"] * 1e-3) / 1e12, 1),
                "results": out, "equal_relmse": matched}
    den.close()
    r.close()
    print(json.dumps(res_line), flush=True)


def run_denoiser_inference(A, args):
    """The denoiser's inference at 1920x1080 on the procedural scene's 4-spp
    film (radiance + albedo + normal + depth planes in HBM): `warmup` untimed
    and `steps` timed film denoises (device time of each).  The weights are
    the seeded initialisation: the cost does not depend on their values."""
    sdir = os.path.join(tempfile.gettempdir(), "asuna_bench_c2_s0")
    p = os.path.join(sdir, "spheres.json")
    if not os.path.exists(p):
        p = A.generate_scene(1, sdir)
    s = A.HostScene(p)
    w, h = s.film
    cam, pc = s.shot(0)
    st = s.state
    r = A.Renderer(0)
    r.upload(s)
    r.build()
    r.set_film(w, h)
    r.render(cam, pc, 0, 4, depth_channel=st.depth_out_channel)
    den = A.Denoiser(0, seed=1)
    for _ in range(args.warmup):
        den.denoise_film(r, pc.diffuseOutChannel, pc.normalOutChannel, st.depth_out_channel, 7)
    ms = [den.denoise_film(r, pc.diffuseOutChannel, pc.normalOutChannel, st.depth_out_channel, 7)
          for _ in range(args.steps)]
    flop = UNET_FLOP_PER_PIXEL * 1920 * 1088
    med = float(np.median(ms))
    out = {"metric": "denoiser 1080p inference", "value": round(med, 3), "unit": "ms", "higher_is_better": False,
           "steps": args.steps, "warmup": args.warmup, "dtype": "bf16 (fp32 accumulation)",
           "config": {"workload": "U-Net 32/48/64/96/128 on the 1920x1080 film (padded to 1920x1088) of the "
                                  "procedural scene at 4 spp"},
           "tflops": round(flop / (med * 1e-3) / 1e12, 1),
           "mfma_frac": round(flop / (med * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4)}
    den.close()
    r.close()
    print(json.dumps(out), flush=True)


def run_serving(A, args):
    """Render serving: 8 resident procedural scenes (about 10k, 100k, 1M and 5M
    triangles, two seeds each), Poisson arrivals of jobs with resolution from
    {256^2, 512^2, 1024^2, 1920x1080} (weights .4/.3/.2/.1) and spp from
    {1, 4, 16, 64}; scene uniform.  The server co-batches the jobs of a
    scene and overlaps scenes; the baseline serves the same arrival stream
    one job at a time: a worker renders each job alone, in arrival order, as
    it arrives (latency = completion - submission, both measured).  Default: both columns do the same work (half of the jobs
    denoised).  --matched: equal relMSE -- the server renders every job at
    the denoised spp that c4 measured to match the job's plain spp and
    denoises it; the baseline renders the plain spp."""
    import threading
    rng = np.random.default_rng(0xA5A52026)
    scenes = []
    t0 = time.time()
    for tris in (10_000, 100_000, 1_000_000, 5_000_000):
        for seed in (1, 2):
            d = os.path.join(tempfile.gettempdir(), f"asuna_serve_{tris}_{seed}")
            p = os.path.join(d, "spheres.json")
            if not os.path.exists(p):
                p = A.generate_scene(4, d, scale=tris, seed=seed)
            scenes.append(A.HostScene(p))
    log(f"[serve] 8 scenes loaded in {time.time() - t0:.1f}s: {[s.state.num_triangles for s in scenes]}")
    res = [(256, 256), (512, 512), (1024, 1024), (1920, 1080)]
    rates = [float(x) for x in args.rates.split(",")]
    matched = None
    if args.matched:
        matched = {int(a): (int(b.rstrip("p")), not b.endswith("p"))
                   for a, b in (kv.split(":") for kv in args.matched.split(","))}

    def make_jobs(n):
        jobs = []
        for _ in range(n):
            sid = int(rng.integers(0, len(scenes)))
            w, h = res[int(rng.choice(4, p=[0.4, 0.3, 0.2, 0.1]))]
            spp = int(rng.choice([1, 4, 16, 64]))
            eye = rng.normal(size=3)
            eye = eye / np.linalg.norm(eye) * rng.uniform(3.0, 6.0)
            eye[1] = abs(eye[1]) + 0.5
            cam = A.camera_lookat(w, h, float(rng.uniform(35, 70)), tuple(eye), (0.0, 0.3, 0.0))
            _, pc = scenes[sid].shot(0)
            dn = bool(rng.integers(0, 2))
            if matched:
                jobs.append((sid, cam, pc, w, h, matched[spp][0], matched[spp][1], spp, False))
            else:
                jobs.append((sid, cam, pc, w, h, spp, dn, spp, dn))
        return jobs

    srv = A.Server(0, hbm_budget_bytes=64 << 30, queue_paths=args.serve_queue, policy=args.serve_policy)
    den = A.Denoiser(0, seed=1)  # weights do not change the cost; c4 trains one
    srv.set_denoiser(den)
    for s in scenes:
        srv.add_scene(s)
    rs = []
    for s in scenes:
        r = A.Renderer(0)
        r.upload(s)
        r.build()
        rs.append(r)
    # warm both paths (graph capture, first-touch allocations)
    for sid in range(len(scenes)):
        cam, pc = scenes[sid].shot(0)
        cam = A.camera_lookat(256, 256, 45.0, (0.0, 1.0, 4.0), (0.0, 0.3, 0.0))
        st = scenes[sid].state
        j = srv.submit(sid, cam, pc, 256, 256, 1, depth_channel=st.depth_out_channel, denoise=True)
        srv.wait(j)
        srv.release(j)
        rs[sid].set_film(256, 256)
        rs[sid].render(cam, pc, 0, 1, depth_channel=st.depth_out_channel)
        den.denoise_film(rs[sid], pc.diffuseOutChannel, pc.normalOutChannel, st.depth_out_channel, 7)
    out = {"metric": "render serving: jobs/s and p50/p99 job latency", "unit": "jobs/s", "n_gpus": 1,
           "config": {"workload": "8 resident scenes (1.2e4, 8.2e4, 1.0e6, 5.2e6 triangles x 2 seeds); Poisson "
                                  "arrivals; resolution {256^2,512^2,1024^2,1920x1080} w {.4,.3,.2,.1}; "
                                  "spp {1,4,16,64}; denoise flag on 50 % of jobs (U-Net on the job's planes)",
                      "jobs_per_rate": args.jobs, "queue_paths": args.serve_queue,
                      "policy": ["oldest first across scenes", "round robin", "oldest first"][args.serve_policy],
                      "mode": ("equal relMSE: server renders plain spp S at the spp map " + args.matched +
                               " (denoised; p = plain) measured by c4; baseline renders S plain") if matched else
                              "same work in both columns (denoise flag on half of the jobs)"},
           "rates": []}
    for lam in rates:
        jobs = make_jobs(args.jobs)
        arrivals = np.cumsum(rng.exponential(1.0 / lam, size=len(jobs)))
        ids = [None] * len(jobs)
        base = time.perf_counter()

        def feeder():
            for k, jb in enumerate(jobs):
                dt = arrivals[k] - (time.perf_counter() - base)
                if dt > 0:
                    time.sleep(dt)
                sid = jb[0]
                ids[k] = srv.submit(*jb[:6], depth_channel=scenes[sid].state.depth_out_channel, denoise=jb[6])

        lat, paths = [0.0] * len(jobs), [0]

        def collector():
            # waits for the jobs in submission order and releases each at once, so that finished
            # jobs do not hold HBM budget while later ones arrive
            for k in range(len(jobs)):
                while ids[k] is None:
                    time.sleep(0.0002)
                st = srv.wait(ids[k])
                lat[k] = st["done_ms"] - st["submit_ms"]
                paths[0] += st["paths"]
                srv.release(ids[k])
                if (k + 1) % 200 == 0:
                    log(f"[serve] rate {lam}/s server: {k + 1} jobs done")

        th, tc = threading.Thread(target=feeder), threading.Thread(target=collector)
        th.start()
        tc.start()
        th.join()
        tc.join()
        paths = paths[0]
        span = (time.perf_counter() - base)
        served = {"jobs_per_s": round(len(jobs) / span, 2), "p50_ms": round(float(np.percentile(lat, 50)), 1),
                  "p99_ms": round(float(np.percentile(lat, 99)), 1), "mpaths_per_s": round(paths / span / 1e6, 1)}
        # baseline: the same arrival stream served one job at a time -- a
        # worker takes the jobs in arrival order as they arrive and renders
        # (and, when flagged, denoises) each alone on the plain renderer
        import queue
        pending = queue.Queue()
        sub_t, done_t, service = [0.0] * len(jobs), [0.0] * len(jobs), [0.0] * len(jobs)
        base = time.perf_counter()

        def single_feeder():
            for k in range(len(jobs)):
                dt = arrivals[k] - (time.perf_counter() - base)
                if dt > 0:
                    time.sleep(dt)
                sub_t[k] = time.perf_counter() - base
                pending.put(k)

        def single_worker():
            for _ in range(len(jobs)):
                k = pending.get()
                sid, cam, pc, w, h, _, _, spp, dn = jobs[k]
                r = rs[sid]
                dch = scenes[sid].state.depth_out_channel
                ts = time.perf_counter()
                r.set_film(w, h)
                r.render(cam, pc, 0, spp, depth_channel=dch)
                if dn:
                    den.denoise_film(r, pc.diffuseOutChannel, pc.normalOutChannel, dch, 7)
                t1 = time.perf_counter()
                done_t[k] = t1 - base
                service[k] = t1 - ts
                if (k + 1) % 200 == 0:
                    log(f"[serve] rate {lam}/s one-at-a-time: {k + 1} jobs done")

        th, tw = threading.Thread(target=single_feeder), threading.Thread(target=single_worker)
        th.start()
        tw.start()
        th.join()
        tw.join()
        blat = [(d - a) * 1e3 for d, a in zip(done_t, sub_t)]
        finish = max(done_t)
        if args.jobs_log:
            with open(args.jobs_log, "a") as f:
                for k, (sid, _, _, w, h, sspp, sdn, spp, dn) in enumerate(jobs):
                    f.write(json.dumps({"rate": lam, "job": k, "arrival_ms": round(float(arrivals[k]) * 1e3, 2),
                                        "scene": sid, "w": w, "h": h, "spp": spp, "server_spp": sspp,
                                        "server_denoise": sdn, "server_ms": round(lat[k], 2),
                                        "single_ms": round(blat[k], 2),
                                        "single_service_ms": round(service[k] * 1e3, 2)}) + "\n")
        single = {"jobs_per_s": round(len(jobs) / finish, 2),
                  "p50_ms": round(float(np.percentile(blat, 50)), 1), "p99_ms": round(float(np.percentile(blat, 99)), 1),
                  "busy_s": round(sum(service), 2),
                  "mpaths_per_s": round(sum(jb[3] * jb[4] * jb[7] for jb in jobs) / finish / 1e6, 1)}
        log(f"[serve] rate {lam}/s: server {served}, one-at-a-time {single}")
        out["rates"].append({"rate": lam, "server": served, "single_job": single})
    srv.close()
    den.close()
    for r in rs:
        r.close()
    out["value"] = out["rates"][-1]["server"]["jobs_per_s"]
    out["higher_is_better"] = True
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
