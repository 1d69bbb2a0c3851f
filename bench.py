"""Benchmark of the brick-DDA traversal on one MI355X.

Workload (N=1): the endless-world shape — 13 x 13 streamed chunks of 128 x 256 x 128 voxels in a
256 x 32 x 256-brick ring grid (seeded 4-octave terrain + 400 trees per chunk), a 1920 x 1080 pinhole
camera 17 voxels above the ground, and one sun ray (elevation 35 deg) per G-buffer hit.  A step is
the primary pass (gbuffer.comp) followed by the sun-shadow pass (shadow_trace.comp) over the whole
image, with the world resident in HBM.  value = traced rays (primary + sun) / s over the timed
steps, in Grays/s.

Also printed in the same JSON line: the roofline of the dominant kernel (algorithmic bytes of the
reference visit sequence, counted by the kernels' counting variant in an untimed pass, over the
kernel time from HIP events on the launch stream) and the CPU baseline (the project's C oracle with
OpenMP on the same frame, which also checks the frame's hits against the GPU's).  The frame is
recorded once as a hipGraph (vf_graph_begin / vf_graph_end) and replayed per step.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "voxelforge-rustrewrite_amd"))

import numpy as np  # noqa: E402
import torch  # noqa: E402

import vftrace as vf  # noqa: E402

METRIC = "Grays/s brick-DDA (hit-exact); 1080p GI ms/frame at fixed MSE vs 64spp"
HBM_PEAK_GBS = 8000.0
W, H = 1920, 1080
SUN_AZ, SUN_EL = 32.0, 35.0
HIT_BYTES = 20          # vf_hit record
GBPOS_BYTES, GBN_BYTES, SHADOW_BYTES = 16, 4, 4


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_world(args):
    t0 = time.time()
    world = vf.GenWorld.generate(2, chunks=args.chunks, trees_per_chunk=args.trees, seed=args.seed)
    log(f"world generated in {time.time() - t0:.1f}s: {world.stats()}")
    return world


def pick_camera(world, seed):
    """Seeded pose near the centre chunk, 17 voxels (1.7 m) above the ground, in air, looking away
    from the sun (yaw = the anti-sun azimuth +- 0.5 rad) so that the G-buffer hits face the sun and
    each gets its sun ray, as the workload intends (a hit facing away is shadowed without a ray)."""
    anti_sun_yaw = math.atan2(-math.cos(math.radians(SUN_AZ)), -math.sin(math.radians(SUN_AZ)))
    rng = np.random.default_rng(seed)
    for _ in range(10000):
        x = math.floor(rng.uniform(16, 112)) + 0.37
        z = math.floor(rng.uniform(16, 112)) + 0.63
        y = world.ground(int(math.floor(x)), int(math.floor(z))) + 17.21
        fx, fy, fz = int(math.floor(x)), int(math.floor(y)), int(math.floor(z))
        clear = all(world.voxel(fx + dx, fy + dy, fz + dz) == 0
                    for dx in (-2, 0, 2) for dy in (-2, 0, 2) for dz in (-2, 0, 2))
        if clear:
            yaw = anti_sun_yaw + float(rng.uniform(-0.5, 0.5))
            pitch = float(rng.uniform(-0.15, 0.05))
            return (x, y, z), yaw, pitch
    raise RuntimeError("no camera position in air found")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--chunks", type=int, default=13)
    ap.add_argument("--trees", type=int, default=400)
    ap.add_argument("--seed", type=int, default=1337)
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="minimum CPU-baseline time (repeats the frame)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="launch the passes eagerly instead of replaying a graph")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="run a plain device copy of this many MiB first (calibrates rocprof FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: rays are independent and the path would shard "
                                   "by image tiles with the world replicated per GPU (see DESIGN.md); "
                                   f"--gpus {args.gpus} is not supported"}))
        sys.exit(2)
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a HIP device")
    dev = torch.device("cuda:0")
    torch.cuda.init()

    if args.copy_mb:
        a = torch.ones(args.copy_mb << 18, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.copy_mb} MiB read + {args.copy_mb} MiB written, 3 times")
        del a, b

    world = build_world(args)
    arrays_wb = world.desc().world_bricks
    win = tuple(world.desc().world_win)
    ctx = vf.Context(0)
    stream = torch.cuda.Stream()
    ctx.set_stream(stream)
    t0 = time.time()
    ctx.upload(world)
    log(f"upload + distance field + sky-top: {time.time() - t0:.2f}s")
    scene = vf.default_scene(tuple(arrays_wb), win, sun=(SUN_AZ, SUN_EL))
    ctx.set_scene(scene)
    pos, yaw, pitch = pick_camera(world, args.seed)
    cam_pos, right, up, fwd = vf.fly_camera(pos, yaw, pitch)
    pu = vf.make_pass(cam_pos, right, up, fwd, 70.0, 0)
    log(f"camera {pos} yaw {yaw:.3f} pitch {pitch:.3f}")

    n = W * H
    hits = torch.empty(n * HIT_BYTES, dtype=torch.uint8, device=dev)
    gbp = torch.empty(n * 4, dtype=torch.float32, device=dev)
    gbn = torch.empty(n, dtype=torch.int32, device=dev)
    sh = torch.empty(n, dtype=torch.int32, device=dev)

    # untimed counting pass: the algorithmic fetch bytes of the reference visit sequence
    ctx.set_counting(True)
    ctx.counters(reset=True)
    ctx.trace_primary(pu, W, H, hits, gbp, gbn)
    cp = ctx.counters(reset=True)
    ctx.trace_shadow(W, H, gbp, gbn, sh)
    cs = ctx.counters(reset=True)
    ctx.set_counting(False)
    torch.cuda.synchronize()
    prim_bytes = cp["bytes"] + n * (HIT_BYTES + GBPOS_BYTES + GBN_BYTES)
    shad_bytes = cs["bytes"] + n * (GBPOS_BYTES + GBN_BYTES + SHADOW_BYTES)
    rays_per_step = cp["rays"] + cs["rays"]
    log(f"per step: {cp['rays']} primary + {cs['rays']} sun rays; fetch bytes/ray primary "
        f"{cp['bytes'] / max(cp['rays'], 1):.1f}, sun {cs['bytes'] / max(cs['rays'], 1):.1f}")

    # the frame (primary pass, then the sun pass) recorded once as a graph and replayed per step;
    # one-kernel graphs time each kernel alone
    class Eager:   # --no-graph: the same calls, launched directly each time
        def __init__(self, fn):
            self.fn = fn

        def close(self):
            pass

    def record(fn):
        if args.no_graph:
            return Eager(fn)
        ctx.graph_begin()
        fn()
        return ctx.graph_end()

    def launch(g):
        if isinstance(g, Eager):
            g.fn()
        else:
            ctx.graph_launch(g)

    g_step = record(lambda: (ctx.trace_primary(pu, W, H, hits, gbp, gbn), ctx.trace_shadow(W, H, gbp, gbn, sh)))
    g_prim = record(lambda: ctx.trace_primary(pu, W, H, hits, gbp, gbn))
    g_shad = record(lambda: ctx.trace_shadow(W, H, gbp, gbn, sh))

    def timed_loop(graph, k):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(k):
            launch(graph)
        e1.record(stream)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, e0.elapsed_time(e1) / k

    for _ in range(args.warmup):
        launch(g_step)
    torch.cuda.synchronize()
    wall, step_ms_ev = timed_loop(g_step, args.steps)
    ms_per_step = wall * 1e3 / args.steps
    value = rays_per_step * args.steps / wall / 1e9
    _, prim_ms1 = timed_loop(g_prim, args.steps)
    _, shad_ms1 = timed_loop(g_shad, args.steps)
    prim_ms = np.array([prim_ms1])
    shad_ms = np.array([shad_ms1])
    log(f"step {ms_per_step:.4f} ms wall ({step_ms_ev:.4f} ms events); k_primary {prim_ms1:.4f} ms, "
        f"k_shadow {shad_ms1:.4f} ms")

    prim_gbs = prim_bytes / (prim_ms.mean() * 1e-3) / 1e9
    shad_gbs = shad_bytes / (shad_ms.mean() * 1e-3) / 1e9
    dominant = "k_primary" if prim_ms.mean() >= shad_ms.mean() else "k_shadow"
    ach = prim_gbs if dominant == "k_primary" else shad_gbs
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_endless_1080p.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get(dominant)
        except (OSError, ValueError):
            traffic = None

    result = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"endless-world shape: {args.chunks}x{args.chunks} chunks of 128x256x128 voxels in a "
                        f"{arrays_wb[0]}x{arrays_wb[1]}x{arrays_wb[2]}-brick ring grid, {args.trees} trees/chunk; "
                        f"{W}x{H} primary rays + 1 sun ray (el {SUN_EL}) per hit",
            "rays_per_step": int(rays_per_step),
            "primary_rays": int(cp["rays"]),
            "sun_rays": int(cs["rays"]),
        },
        "primary_grays": round(cp["rays"] / (prim_ms.mean() * 1e-3) / 1e9, 4),
        "sun_grays": round(cs["rays"] / (shad_ms.mean() * 1e-3) / 1e9, 4),
        "kernel_ms": {"k_primary": round(float(prim_ms.mean()), 5), "k_shadow": round(float(shad_ms.mean()), 5)},
        "algorithmic_bytes_per_launch": {"k_primary": int(prim_bytes), "k_shadow": int(shad_bytes)},
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": round(ach, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 5),
            "traffic": traffic,
        },
    }

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(world, win, arrays_wb, cam_pos, right, up, fwd, pu.tan_half_fov, scene,
                                              hits, args.cpu_seconds)
    for g in (g_step, g_prim, g_shad):
        g.close()
    ctx.close()
    print(json.dumps(result), flush=True)


def cpu_baseline(world, win, wb, cam_pos, right, up, fwd, thf, scene, gpu_hits, min_seconds):
    """The C oracle (OpenMP over the host cores) on the same workload: the whole frame's primary and
    sun rays, repeated until min_seconds have passed.  Its hits double as a hit-exact check of the
    GPU's whole frame."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    cores = len(os.sched_getaffinity(0))
    cores = min(cores, 16)
    t0 = time.time()
    ow = O.OracleWorld(world.arrays(), build_skytop=False)
    log(f"oracle world built in {time.time() - t0:.1f}s")
    cam = {"pos": cam_pos, "right": right, "up": up, "fwd": fwd, "tan_half_fov": thf}
    sun = np.array(scene.sun_dir, np.float32)
    rays, frames = 0, 0
    t0 = time.perf_counter()
    while True:
        oh, opos, onrm, cp = ow.pass_primary_rows(cam, W, H, 0, 1, H, threads=cores)
        osh, _, cs = ow.pass_shadow(sun, W, H, opos, onrm, threads=cores)
        rays += cp["rays"] + cs["rays"]
        frames += 1
        dt = time.perf_counter() - t0
        if dt >= min_seconds:
            break
    g = vf.hits_to_numpy(gpu_hits)
    mism = int(((g.view(np.uint8).reshape(-1, 20) != oh.view(np.uint8).reshape(-1, 20)).any(1)).sum())
    return {
        "value": round(rays / dt / 1e9, 6),
        "unit": "Grays/s",
        "cores": cores,
        "kind": "port",
        "sample": f"the whole 1920x1080 frame ({cp['rays']} primary + {cs['rays']} sun rays) x {frames} repeats; "
                  f"the project's C oracle (FP32 restatement of traceVolumeS) with OpenMP; lavapipe baseline "
                  f"unavailable (no cargo / shaderc / Vulkan loader on the box)",
        "seconds": round(dt, 3),
        "hit_exact_check": {"rays": int(g.size), "mismatches": mism},
    }


if __name__ == "__main__":
    main()
