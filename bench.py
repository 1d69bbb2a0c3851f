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
OpenMP on the same frame, which also checks the frame's hits against the GPU's).  Each step launches
the frame's pass eagerly on the context's stream (--graph-steps: replays one recorded hipGraph per
step instead).  The temporal passes that consume the traced output (motion vectors, GI accumulation,
shadow history, a-trous filter) are timed at 1920 x 1080 in the same run.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "voxelforge-rustrewrite_amd"))
# Kernel arguments in device memory (a ROCm runtime option, read when the runtime starts): the packet
# processor then reads a launch's arguments from HBM instead of over PCIe, which shortens the dispatch
# of each eager launch (measured: 100.0-100.8 -> 97.7-98.2 us per step).  A renderer sets it the same way.
RUNTIME_ENV = {"HIP_FORCE_DEV_KERNARG": os.environ.setdefault("HIP_FORCE_DEV_KERNARG", "1")}

import numpy as np  # noqa: E402
import torch  # noqa: E402

import vftrace as vf  # noqa: E402
from vftrace import bodies as B  # noqa: E402

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
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the open-forest, fixed-world, GI-batch, learned-cache and uneven-batch lines")
    ap.add_argument("--extras", default="temporal,dist,pack,forest,other",
                    help="comma list of the extra measurements: temporal (the passes after the trace), dist (the "
                         "brick distance field over the whole world), pack (the world packed on the device), forest "
                         "(open forest with bodies), other (all of the fixed world's lines: fixed = primary + sun "
                         "rays at 640x360, gibatch = the GI batch, cache = the learned GI cache, batch = probes + "
                         "16 views in one launch; each may be named on its own)")
    ap.add_argument("--no-graph", action="store_true",
                    help="record no graphs at all (also the kernel-duration loop launches eagerly)")
    ap.add_argument("--graph-steps", action="store_true",
                    help="replay one recorded graph per step instead of launching the pass eagerly")
    ap.add_argument("--cache-frames", type=int, default=256, help="online-training frames before the GI cache is judged")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="run a plain device copy of this many MiB first (calibrates rocprof FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU: rays are independent and the path would shard "
                                   "by image tiles with the world replicated per GPU; "
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
    pos, yaw, pitch = pick_camera(world, args.seed)
    cam_pos, right, up, fwd = vf.fly_camera(pos, yaw, pitch)
    pu = vf.make_pass(cam_pos, right, up, fwd, 70.0, 0)
    log(f"camera {pos} yaw {yaw:.3f} pitch {pitch:.3f}")
    ctx = vf.Context(0)
    stream = torch.cuda.Stream()
    ctx.set_stream(stream)
    torch.cuda.set_stream(stream)   # torch's own work on the buffers stays ordered with the library's launches
    t0 = time.time()
    ctx.upload(world)
    log(f"upload + distance field + sky-top: {time.time() - t0:.2f}s")
    scene = vf.default_scene(tuple(arrays_wb), win, sun=(SUN_AZ, SUN_EL))
    ctx.set_scene(scene)

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

    # each step launches the frame's pass (primary and sun rays, one fused kernel); the comparison lines
    # launch the two passes, or each kernel alone, the same way
    class Eager:   # the same calls, launched directly each time
        def __init__(self, fn):
            self.fn = fn

        def close(self):
            pass

    def record(fn, graph=None):
        # steps are launched eagerly (a renderer launches each frame's pass once; consecutive eager
        # launches queue back to back, a one-kernel graph replay adds ~8 us per step); the
        # kernel-duration loop below records its back-to-back launches as one graph
        if args.no_graph or not (args.graph_steps if graph is None else graph):
            return Eager(fn)
        ctx.graph_begin()
        fn()
        return ctx.graph_end()

    def launch(g):
        if isinstance(g, Eager):
            g.fn()
        else:
            ctx.graph_launch(g)

    # the step: primary and sun rays fused in one kernel (same outputs as the two passes)
    g_step = record(lambda: ctx.trace_primary_shadow(pu, W, H, hits, gbp, gbn, sh))
    g_two = record(lambda: (ctx.trace_primary(pu, W, H, hits, gbp, gbn), ctx.trace_shadow(W, H, gbp, gbn, sh)))
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
    _, two_ms = timed_loop(g_two, args.steps)
    _, prim_ms1 = timed_loop(g_prim, args.steps)
    _, shad_ms1 = timed_loop(g_shad, args.steps)
    prim_ms = np.array([prim_ms1])
    shad_ms = np.array([shad_ms1])
    log(f"step (fused) {ms_per_step:.4f} ms wall ({step_ms_ev:.4f} ms events); two passes {two_ms:.4f} ms; "
        f"k_primary {prim_ms1:.4f} ms, k_shadow {shad_ms1:.4f} ms")

    # the dominant (only) kernel of the step: k_primary_shadow64 (the fused kernel launched one 8x8-pixel
    # wave per workgroup for a camera in air); its algorithmic bytes = the fetches of
    # both rays + the primary outputs (hit, gbPos, gbNormal) + shadowRaw (no G-buffer read-back).
    # Its launch duration: 20 launches recorded back to back in one graph, timed with HIP events on
    # the launch stream (per-replay timing would add the graph-launch gap to every kernel).
    fused_bytes = cp["bytes"] + cs["bytes"] + n * (HIT_BYTES + GBPOS_BYTES + GBN_BYTES + SHADOW_BYTES)
    dominant = "k_primary_shadow64"
    reps = 20
    g_rep = record(lambda: [ctx.trace_primary_shadow(pu, W, H, hits, gbp, gbn, sh) for _ in range(reps)], graph=True)
    _, rep_ms = timed_loop(g_rep, max(1, args.steps // reps))
    kernel_ms = rep_ms / reps
    g_rep.close()
    ach = fused_bytes / (kernel_ms * 1e-3) / 1e9
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
        "runtime_env": RUNTIME_ENV,
        "config": {
            "workload": f"endless-world shape: {args.chunks}x{args.chunks} chunks of 128x256x128 voxels in a "
                        f"{arrays_wb[0]}x{arrays_wb[1]}x{arrays_wb[2]}-brick ring grid, {args.trees} trees/chunk; "
                        f"{W}x{H} primary rays + 1 sun ray (el {SUN_EL}) per hit",
            "rays_per_step": int(rays_per_step),
            "primary_rays": int(cp["rays"]),
            "sun_rays": int(cs["rays"]),
        },
        "median_primary_t": round(float(gbp.view(-1, 4)[:, 3][gbp.view(-1, 4)[:, 3] >= 0].median().item()), 2),
        "primary_grays": round(cp["rays"] / (prim_ms.mean() * 1e-3) / 1e9, 4),
        "sun_grays": round(cs["rays"] / (shad_ms.mean() * 1e-3) / 1e9, 4),
        "kernel_ms": {dominant: round(float(kernel_ms), 5), "step_replay": round(float(step_ms_ev), 5), "k_primary": round(float(prim_ms.mean()), 5),
                      "k_shadow": round(float(shad_ms.mean()), 5), "two_pass_step": round(float(two_ms), 5)},
        "algorithmic_bytes_per_launch": {dominant: int(fused_bytes), "k_primary": int(prim_bytes),
                                         "k_shadow": int(shad_bytes)},
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": round(ach, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            # what binds the kernel instead (no MFMA work exists on this path): its SQ counters show the
            # VALU issue slots filled for the launch's whole duration
            "binding": "VALU issue (SQ_ACTIVE_INST_VALU x 4 / SIMDs = the launch's cycles; "
                       "profiles/k_primary_shadow_endless_1080p.md)",
        },
    }

    if not args.no_cpu:
        # one more (untimed) fused launch that also writes the sun rays' first-segment hits, for the check
        shh = torch.empty(n * HIT_BYTES, dtype=torch.uint8, device=dev)
        ctx.trace_primary_shadow(pu, W, H, hits, gbp, gbn, sh, shh)
        torch.cuda.synchronize()
        result["cpu_baseline"] = cpu_baseline(world, win, arrays_wb, cam_pos, right, up, fwd, pu.tan_half_fov, scene,
                                              hits, gbn, sh, shh, args.cpu_seconds)
        del shh
    for g in (g_step, g_two, g_prim, g_shad):
        g.close()
    if not args.no_extra:
        extras = set(args.extras.split(","))
        if "temporal" in extras:
            result["temporal_1080p"] = temporal_config(ctx, stream, args, pu)
        if "dist" in extras:
            result["brickdist_full"] = dist_config(ctx, world)
        if "pack" in extras:
            result["world_pack"] = pack_config(ctx, world, scene)
        if "forest" in extras:
            result["open_forest_1080p"] = open_forest_config(ctx, stream, args, hits, gbp, gbn, sh)
        parts = {"fixed", "gibatch", "cache", "batch"} if "other" in extras else extras & {"fixed", "gibatch", "cache", "batch"}
        if parts:
            result["other_configs"] = other_configs(ctx, args, parts)
    ctx.close()
    print(json.dumps(result), flush=True)


def dist_config(ctx, world, reps=20):
    """record_distance_field over the whole headline world (both fields), timed with HIP events around its
    passes (vf_last_kernel_ms).  Bytes: the passes' own traffic, 15 B per brick (normal field: grid 4 B read,
    the x, y, z passes' reads and writes 5 B; dry field: occupancy 1 B + 5 B), and the same plus the
    reference's dry-occupancy scan of 512 atlas bytes per detailed brick, which this build replaces by a
    per-slot flag kept with the occupancy masks."""
    d = world.desc()
    wb, win = tuple(d.world_bricks), tuple(d.world_win)
    lo = [win[0], 0, win[2]]
    hi = [win[0] + wb[0] - 1, wb[1] - 1, win[2] + wb[2] - 1]
    ctx.build_brickdist(lo, hi)
    ms = []
    ctx.set_timing(True)
    for _ in range(reps):
        ctx.build_brickdist(lo, hi)
        ms.append(ctx.last_kernel_ms())
    ctx.set_timing(False)
    t = float(np.median(ms))
    n = wb[0] * wb[1] * wb[2]
    detailed = int((np.asarray(world.arrays()["grid"], np.uint32) >= 256).sum())
    own, scan = 15 * n, 15 * n + 512 * detailed
    res = {"bricks": n, "detailed_bricks": detailed, "ms": round(t, 5), "launches": 2,
           "pass_bytes": own, "pass_GB/s": round(own / (t * 1e-3) / 1e9, 1), "pass_frac": round(own / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "with_reference_scan_bytes": scan,
           "with_reference_scan_frac": round(scan / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    log(f"distance field: {res}")
    return res


def pack_config(ctx, world, scene):
    """The headline world uploaded twice: its host-packed arrays (vf_world_upload: grid and atlas as the
    reference's CPU pack leaves them), then the reference's CPU-side form (grid of brick references, the
    bricks in another order) packed on the device (vf_world_pack); the two device worlds must be equal.
    Wall times include the host-to-device copies and the whole derived-data build."""
    a = world.arrays()
    grid, atlas = np.asarray(a["grid"], np.uint32), np.asarray(a["atlas"], np.uint8).reshape(-1, 512)
    rng = np.random.default_rng(5)
    perm = rng.permutation(atlas.shape[0])
    src = grid.copy()
    det = grid >= 256
    src[det] = 256 + np.argsort(perm)[grid[det] - 256]
    bricks = np.ascontiguousarray(atlas[perm])
    out = {"bricks": int(atlas.shape[0]), "atlas_mb": round(atlas.nbytes / 1e6, 1)}
    t0 = time.perf_counter()
    ctx.upload_arrays(a)
    out["upload_host_packed_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    g_up = ctx.world_read(0)[0]
    t0 = time.perf_counter()
    slot = ctx.pack_world(a["world_bricks"], a["world_win"], src, bricks, a["palette"], atlas_slots=atlas.shape[0] + 4096)
    out["pack_on_device_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    g_pk = ctx.world_read(0)[0]
    out["same_grid"] = bool(np.array_equal(g_up, g_pk) and np.array_equal(slot, perm))
    wb, win = tuple(a["world_bricks"]), tuple(a["world_win"])
    if win[3]:   # ring world: two streaming steps over the four fullest chunk rows (the first grows the buffers)
        ring_w = wb[0] // 16
        order = np.argsort(-(src.reshape(wb[2], wb[1], wb[0]) >= 256).reshape(ring_w, -1).sum(1), kind="stable")
        first, src1, b1, s1 = sync_step(ctx, a, src, bricks, slot, (int(order[0]), int(order[1])))
        out["chunk_sync"], _, _, _ = sync_step(ctx, a, src1, b1, s1, (int(order[2]), int(order[3])))
        out["chunk_sync"]["first_call_ms"] = first["sync_ms"]
        out["chunk_sync"]["consistent"] &= first["consistent"]
    ctx.set_scene(scene)
    log(f"world pack: {out}")
    return out


def sync_step(ctx, a, src, bricks, slot, rows):
    """One streaming step of the endless world (vf_world_sync_chunks: GpuWorld::sync_chunks): the ring's
    row of chunks rows[0] is evicted, the slots of the bricks only it names released, and refilled with
    the content of row rows[1] as new bricks (copies, no slot yet).  Wall time of the call: the chunk
    cells and the named bricks to the device, slot allocation, grid/atlas/mask/coarse writes, the
    distance field over each chunk's box +-16 bricks, cell words and sky-top, the host bookkeeping.
    Returns the report and the world after the step (source grid, bricks, slots)."""
    wb, win = tuple(a["world_bricks"]), tuple(a["world_win"])
    cb, ring_w = 16, wb[0] // 16
    g3 = src.reshape(wb[2], wb[1], wb[0])
    tz_dst, tz_src = rows[0] * cb, rows[1] * cb
    cz_dst = next(c for c in range(win[2] // cb, win[2] // cb + ring_w) if (c * cb) & (wb[2] - 1) == tz_dst)
    dst_rows = g3[tz_dst:tz_dst + cb]
    rest = np.concatenate([g3[:tz_dst].reshape(-1), g3[tz_dst + cb:].reshape(-1)])
    gone = np.setdiff1d(np.unique(dst_rows[dst_rows >= 256]), rest[rest >= 256]) - 256   # named only there
    n_b = bricks.shape[0]
    src_rows = g3[tz_src:tz_src + cb]
    named = np.unique(src_rows[src_rows >= 256]) - 256
    new_id = np.full(n_b, -1, np.int64)
    new_id[named] = n_b + np.arange(named.size)
    all_bricks = np.concatenate([bricks, bricks[named]])
    gslot = np.concatenate([slot, np.full(named.size, -1, np.int32)]).astype(np.int32)
    released = gslot[gone].copy()
    gslot[gone] = -1
    new_rows = src_rows.copy()
    det = new_rows >= 256
    new_rows[det] = 256 + new_id[new_rows[det] - 256]
    chunks = []
    for i in range(ring_w):
        cx = win[0] // cb + i
        tx = (cx * cb) & (wb[0] - 1)
        chunks.append((cx, cz_dst, np.ascontiguousarray(new_rows[:, :, tx:tx + cb]).reshape(-1)))
    used0, free0, _ = ctx.atlas_state()
    ctx.sync()
    t0 = time.perf_counter()
    out_slot = ctx.sync_chunks(chunks, all_bricks, gslot, released)
    ms = (time.perf_counter() - t0) * 1e3
    used, n_free, _ = ctx.atlas_state()
    fresh = out_slot[n_b:]
    ok = bool((fresh >= 0).all() and len(set(fresh.tolist())) == fresh.size and
              used == used0 + max(0, named.size - gone.size - free0) and n_free == max(0, free0 + gone.size - named.size))
    cells = ring_w * cb * wb[1] * cb
    src2 = src.copy()
    src2.reshape(g3.shape)[tz_dst:tz_dst + cb] = new_rows
    rep = {"chunks": ring_w, "cells": cells, "bricks_in": int(named.size), "slots_released": int(gone.size),
           "sync_ms": round(ms, 2), "consistent": ok,
           "host_to_device_mb": round((cells * 4 + named.size * 516) / 1e6, 2)}
    return rep, src2, all_bricks, out_slot


def open_forest_config(ctx, stream, args, hits, gbp, gbn, sh):
    """The endless-world shape with a tenth of the trees (40 per chunk): at 400 per chunk the crowns
    (radius 10-30 from 10 voxels up) merge into a leaf layer that the 1.7 m camera sees a few voxels
    away, here the primary rays cross open forest.  The same frame (primary + sun rays, fused) without
    and with the seven bodies of the reference demo scene (two crates, three spheres, two pillars, spun
    and tilted) 30 to 110 voxels ahead, traced through their BVH by every primary and sun ray."""
    dev = torch.device("cuda:0")
    world = vf.GenWorld.generate(2, chunks=args.chunks, trees_per_chunk=max(args.trees // 10, 1), seed=args.seed)
    pos, yaw, pitch = pick_camera(world, args.seed)
    cam_pos, right, up, fwd = vf.fly_camera(pos, yaw, pitch)
    pu = vf.make_pass(cam_pos, right, up, fwd, 70.0, 0)
    vols, _ = B.demo_bodies((0.0, 0.0), lambda x, z: 0.0)
    brng = np.random.default_rng(args.seed)
    poses = []
    for k, v in enumerate(vols):
        d = 30.0 + 13.0 * k
        c = cam_pos + fwd * d + right * ((-1.0) ** k * 0.35 * d) + up * float(brng.uniform(-6.0, 6.0))
        poses.append(B.pose(c, v.shape, float(brng.uniform(0.0, 6.28)), float(brng.uniform(-0.5, 0.5))))
    arrays, hd, grid = B.attach(world.arrays(), vols, poses)
    ctx.upload_arrays(arrays)
    ctx.objects_upload(hd, grid)
    d = world.desc()
    scene = vf.default_scene(tuple(d.world_bricks), tuple(d.world_win), sun=(SUN_AZ, SUN_EL))
    out = {"trees_per_chunk": max(args.trees // 10, 1), "camera": [round(float(x), 3) for x in cam_pos]}
    n = W * H
    body = torch.empty(n, dtype=torch.int32, device=dev)
    for label, n_obj in (("no_bodies", 0), ("bodies_7", len(hd))):
        scene.num_objects = n_obj
        ctx.set_scene(scene)
        ctx.set_counting(True)
        ctx.counters(reset=True)
        ctx.trace_primary_shadow(pu, W, H, hits, gbp, gbn, sh)
        c = ctx.counters(reset=True)
        ctx.set_counting(False)
        ctx.trace_gbuffer(pu, W, H, hits, gbp, gbn, body)
        torch.cuda.synchronize()
        t = gbp.view(-1, 4)[:, 3]
        ms = graph_ms(ctx, stream, lambda: ctx.trace_primary_shadow(pu, W, H, hits, gbp, gbn, sh), args.steps)
        out[label] = {"rays": int(c["rays"]), "ms": round(ms, 5), "grays": round(c["rays"] / (ms * 1e-3) / 1e9, 4),
                      "fetch_bytes_per_ray": round(c["bytes"] / max(c["rays"], 1), 2),
                      "median_primary_t": round(float(t[t >= 0].median().item()), 2),
                      "body_pixels": int((body >= 0).sum().item())}
        log(f"open forest, {label}: {out[label]}")
    scene.num_objects = 0
    ctx.set_scene(scene)
    ctx.objects_upload(hd[:0], grid[:0])
    return out


# Algorithmic bytes per unit of the temporal passes (each input texel the pass needs read once, every
# output written once): gbuffer motion per pixel = gbPos 16 + gbMotion 4 + gbDelta 8; gi.comp
# accumulation per half-res texel = this frame's gbPos 2x16 + gbNormal 2x4 (representative and rotating
# sample) + motion 4 + delta 8 + traced radiance 16 + last frame's gbPos 16 + gbNormal 4 + history 8 +
# moments 16 + outputs 8 + 16 + 16; lighting.comp history per pixel = gbPos 16 + gbNormal 4 + motion 4 +
# delta 8 + sample 4 + last frame's gbPos 16 + gbNormal 4 + history 16 + output 16; an a-trous level per
# half-res texel = gbPos 16 + gbNormal 4 + guide 16 + its input (history 8 + moments 16 at level 0,
# the previous level 8 after) + output 8.
TEMPORAL_BYTES = {"k_motion": 28, "k_gi_accumulate": 152, "k_light_accumulate": 88, "k_atrous_3": 68 + 52 * 2}


def temporal_config(ctx, stream, args, pu_prev):
    """The temporal passes of one 1920x1080 frame of the endless world after the camera moved 0.3 voxels
    and turned 0.01 rad: motion vectors, gi.comp accumulation (Catmull-Rom on), lighting.comp history,
    3 a-trous levels, on inputs traced by the library (G-buffer of both frames, half-res GI, cone sun
    ray).  Each kernel: 20 launches recorded back to back in a graph, timed with HIP events."""
    dev = torch.device("cuda:0")
    n = W * H
    hw, hh = (W + 1) // 2, (H + 1) // 2
    pos = [pu_prev.cam_pos[i] + 0.3 * pu_prev.cam_fwd[i] for i in range(3)]
    yaw = math.atan2(pu_prev.cam_fwd[0], pu_prev.cam_fwd[2]) + 0.01
    pitch = math.asin(max(-1.0, min(1.0, pu_prev.cam_fwd[1])))
    cam_pos, right, up, fwd = vf.fly_camera(pos, yaw, pitch)
    pu = vf.set_prev_camera(vf.make_pass(cam_pos, right, up, fwd, 70.0, 1), pu_prev)
    z = lambda k, dt: torch.zeros(k, dtype=dt, device=dev)
    hits = z(n * HIT_BYTES, torch.uint8)
    gpos, gn = [z(n * 4, torch.float32) for _ in range(2)], [z(n, torch.int32) for _ in range(2)]
    mot, dl = z(n, torch.int32), z(n * 4, torch.int16)
    rad, sun = z(hw * hh * 4, torch.float32), z(n, torch.float32)
    hist = [z(hw * hh * 4, torch.int16) for _ in range(2)]
    mom = [z(hw * hh * 4, torch.float32) for _ in range(2)]
    lh = [z(n * 4, torch.float32) for _ in range(2)]
    guide, f0, f1 = z(hw * hh * 4, torch.int32), z(hw * hh * 4, torch.int16), z(hw * hh * 4, torch.int16)
    ctx.trace_gbuffer(pu_prev, W, H, hits, gpos[1], gn[1], None)
    ctx.trace_gbuffer(pu, W, H, hits, gpos[0], gn[0], None)
    ctx.gbuffer_motion(pu, W, H, gpos[0], mot, dl)
    ctx.trace_gi(pu, W, H, gpos[0], gn[0], rad, None, half_res=True, gi_reach=48.0)
    ctx.trace_sun_cone(pu, W, H, gpos[0], gn[0], sun)
    # a converged history: accumulate a few frames first (same inputs)
    for _ in range(6):
        ctx.gi_accumulate(pu, W, H, gpos[0], gn[0], gpos[1], gn[1], mot, dl, rad, hist[1], mom[1], hist[0], mom[0], guide)
        ctx.gi_accumulate(pu, W, H, gpos[0], gn[0], gpos[1], gn[1], mot, dl, rad, hist[0], mom[0], hist[1], mom[1], guide)
        ctx.light_accumulate(pu, W, H, gpos[0], gn[0], gpos[1], gn[1], mot, dl, sun, lh[1], lh[0])
        ctx.light_accumulate(pu, W, H, gpos[0], gn[0], gpos[1], gn[1], mot, dl, sun, lh[0], lh[1])
    torch.cuda.synchronize()
    calls = {
        "k_motion": (lambda: ctx.gbuffer_motion(pu, W, H, gpos[0], mot, dl), n),
        "k_gi_accumulate": (lambda: ctx.gi_accumulate(pu, W, H, gpos[0], gn[0], gpos[1], gn[1], mot, dl, rad, hist[1],
                                                      mom[1], hist[0], mom[0], guide), hw * hh),
        "k_light_accumulate": (lambda: ctx.light_accumulate(pu, W, H, gpos[0], gn[0], gpos[1], gn[1], mot, dl, sun,
                                                            lh[1], lh[0]), n),
        "k_atrous_3": (lambda: ctx.gi_atrous(W, H, 3, gpos[0], gn[0], mom[0], hist[0], guide, f0, f1), hw * hh),
    }
    reps = 20
    out = {}
    total = 0.0
    for name, (fn, units) in calls.items():
        if args.no_graph:   # counter runs: plain launches
            ms = eager_ms(stream, fn, reps * max(1, args.steps // reps))
        else:
            ms = graph_ms(ctx, stream, lambda: [fn() for _ in range(reps)], max(1, args.steps // reps)) / reps
        nbytes = TEMPORAL_BYTES[name] * units
        gbs = nbytes / (ms * 1e-3) / 1e9
        total += ms
        out[name] = {"us": round(ms * 1e3, 3), "algorithmic_bytes": int(nbytes), "GB/s": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4)}
    out["all_us"] = round(total * 1e3, 3)
    cnt = hist[0].view(torch.float16).view(-1, 4)[:, 3].float()
    out["history_reused"] = round(float((cnt > 1).float().mean().item()), 4)
    log(f"temporal passes 1080p: {out}")
    return out


def eager_ms(stream, fn, k):
    """Average time of fn over k back-to-back eager calls (HIP events on the launch stream)."""
    for _ in range(3):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(k):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / k


def graph_ms(ctx, stream, fn, k):
    """Average time of one replay of fn's recorded graph over k back-to-back replays (HIP events)."""
    ctx.graph_begin()
    fn()
    g = ctx.graph_end()
    for _ in range(3):
        ctx.graph_launch(g)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(k):
        ctx.graph_launch(g)
    e1.record(stream)
    torch.cuda.synchronize()
    g.close()
    return e0.elapsed_time(e1) / k


def other_configs(ctx, args, parts):
    """The fixed world (256x128x256 voxels): primary + sun rays at 640x360 ("fixed"; with the fused kernel's
    time at 320x180 and 1280x720 beside it), the GI batch at 1920x1080 ("gibatch": 1 GI ray per pixel, reach
    640 voxels, 1 sun ray per GI hit; autotuned queue), the learned GI cache at 1920x1080 ("cache"), and one
    launch of a 32x16x32 DDGI probe grid (256 rays each) plus 16 views of mixed sizes ("batch")."""
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream()
    ctx.set_stream(stream)
    prev = torch.cuda.current_stream()
    torch.cuda.set_stream(stream)
    out = {}
    world = vf.GenWorld.generate(1, seed=args.seed)
    ctx.upload(world)
    scene = vf.default_scene((32, 16, 32))
    ctx.set_scene(scene)
    cam_pos, right, up, fwd = vf.fly_camera([37.3, 96.4, 41.7], 0.8, -0.35)
    if "fixed" in parts:
        out.update(fixed_world_config(ctx, stream, args, cam_pos, right, up, fwd))
    if parts & {"gibatch", "cache"}:
        gibatch_and_cache(ctx, stream, args, cam_pos, right, up, fwd, parts, out)
    if "batch" in parts:
        out["uneven_batch"] = uneven_batch_config(ctx, stream, args)
    ctx.set_stream(prev)
    torch.cuda.set_stream(prev)
    return out


def fixed_world_config(ctx, stream, args, cam_pos, right, up, fwd):
    """Primary + sun rays on the fixed world at 640x360 (the fused pass launched eagerly each step, as the
    headline's frames are, so the launch is included), and the same frame at 320x180 and 1280x720 with the
    fused kernel alone (20 back-to-back launches in one graph): its time hardly changes from 320x180 to
    640x360, the latency of the frame's longest wavefronts.  These launches take the small-frame traversal
    (vf_set_traversal); the 640x360 frame is also timed with the plain one."""
    dev = torch.device("cuda:0")
    out = {}
    pu = vf.make_pass(cam_pos, right, up, fwd, 70.0, 0)
    scaling = {}
    for (w1, h1) in ((640, 360), (320, 180), (1280, 720)):
        n1 = w1 * h1
        hits = torch.empty(n1 * HIT_BYTES, dtype=torch.uint8, device=dev)
        gbp = torch.empty(n1 * 4, dtype=torch.float32, device=dev)
        gbn = torch.empty(n1, dtype=torch.int32, device=dev)
        sh = torch.empty(n1, dtype=torch.int32, device=dev)
        ctx.set_counting(True)
        ctx.counters(reset=True)
        ctx.trace_primary_shadow(pu, w1, h1, hits, gbp, gbn, sh)
        c1 = ctx.counters(reset=True)
        ctx.set_counting(False)
        fn = lambda: ctx.trace_primary_shadow(pu, w1, h1, hits, gbp, gbn, sh)   # noqa: E731
        if (w1, h1) == (640, 360):
            ms = eager_ms(stream, fn, args.steps)
            out["fixed_world_640x360"] = {"rays": int(c1["rays"]), "ms": round(ms, 5),
                                          "grays": round(c1["rays"] / (ms * 1e-3) / 1e9, 4)}
        reps = 20
        ms_k = graph_ms(ctx, stream, lambda: [fn() for _ in range(reps)], max(1, args.steps // reps)) / reps
        scaling[f"{w1}x{h1}"] = {"rays": int(c1["rays"]), "kernel_us": round(ms_k * 1e3, 2),
                                 "grays": round(c1["rays"] / (ms_k * 1e-3) / 1e9, 4)}
        if (w1, h1) == (640, 360):   # the same frame with the plain traversal (vf_set_traversal 1)
            ctx.set_traversal(1)
            ms_p = graph_ms(ctx, stream, lambda: [fn() for _ in range(reps)], max(1, args.steps // reps)) / reps
            ctx.set_traversal(0)
            scaling["640x360_plain_traversal"] = {"rays": int(c1["rays"]), "kernel_us": round(ms_p * 1e3, 2),
                                                  "grays": round(c1["rays"] / (ms_p * 1e-3) / 1e9, 4)}
    out["fixed_world_frame_sizes"] = scaling
    return out


def gibatch_and_cache(ctx, stream, args, cam_pos, right, up, fwd, parts, out):
    dev = torch.device("cuda:0")
    # GI batch at 1080p
    w3, h3 = 1920, 1080
    n3 = w3 * h3
    pu3 = vf.make_pass(cam_pos, right, up, fwd, 70.0, 4)
    hits3 = torch.empty(n3 * HIT_BYTES, dtype=torch.uint8, device=dev)
    gbp3 = torch.empty(n3 * 4, dtype=torch.float32, device=dev)
    gbn3 = torch.empty(n3, dtype=torch.int32, device=dev)
    ctx.trace_primary(pu3, w3, h3, hits3, gbp3, gbn3)
    if "cache" in parts:
        out["learned_gi_cache_1080p"] = gi_cache_config(ctx, stream, args, pu3, w3, h3, gbp3, gbn3)
    if "gibatch" not in parts:
        return
    gi = torch.empty(n3 * HIT_BYTES, dtype=torch.uint8, device=dev)
    sun = torch.empty(n3 * HIT_BYTES, dtype=torch.uint8, device=dev)
    best, table = ctx.autotune_gi_batch(pu3, w3, h3, 640.0, gbp3, gbn3, gi, sun)
    ctx.set_counting(True)
    ctx.counters(reset=True)
    ctx.trace_gi_batch(pu3, w3, h3, 640.0, gbp3, gbn3, gi, sun, best)
    c3 = ctx.counters(reset=True)
    ctx.set_counting(False)
    ms3 = graph_ms(ctx, stream, lambda: ctx.trace_gi_batch(pu3, w3, h3, 640.0, gbp3, gbn3, gi, sun, best), 50)
    rad = torch.empty(n3 * 4, dtype=torch.float32, device=dev)
    ms3p = graph_ms(ctx, stream, lambda: ctx.trace_gi(pu3, w3, h3, gbp3, gbn3, rad, gi, half_res=False, gi_reach=640.0), 20)
    out["gi_batch_1080p"] = {"rays": int(c3["rays"]), "ms": round(ms3, 5), "grays": round(c3["rays"] / (ms3 * 1e-3) / 1e9, 4),
                             "queue": {"tile": [best.tile_w, best.tile_h], "chunk": best.chunk, "sort": best.sort,
                                       "refill": best.refill},
                             "autotune_candidates": len(table),
                             "plain_gi_pass_ms": round(ms3p, 5)}


def uneven_batch_config(ctx, stream, args):
    """One launch of a 32x16x32 DDGI probe grid (256 rays each, 2 m = 20 voxel spacing) plus 16 seeded views
    drawn from {256^2, 512x288, 960x540, 1024^2} on the fixed world (autotuned mode), against the same
    requests as sequential plain launches (one graph each)."""
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(args.seed)
    sizes = [(256, 256), (512, 288), (960, 540), (1024, 1024)]
    views = []
    for k in range(16):
        vw, vh = sizes[int(rng.integers(0, 4))]
        pos = [float(rng.uniform(20, 236)), float(rng.uniform(75, 120)), float(rng.uniform(20, 236))]
        vp, vr, vu, vfw = vf.fly_camera(pos, float(rng.uniform(0, 2 * math.pi)), float(rng.uniform(-0.7, 0.1)))
        views.append((vw, vh, vf.make_pass(vp, vr, vu, vfw, 70.0, 0)))
    views.sort(key=lambda v: v[0] * v[1])        # smallest views first: they finish first
    outs = [torch.empty(vw * vh * HIT_BYTES, dtype=torch.uint8, device=dev) for vw, vh, _ in views]
    dims, nr = (32, 16, 32), 256
    origin = (128.0 - 310.0 + 0.5, 64.0 - 150.0 + 0.5, 128.0 - 310.0 + 0.5)
    probe_out = torch.empty(int(np.prod(dims)) * nr * HIT_BYTES, dtype=torch.uint8, device=dev)
    reqs = [ctx.view_request(vp, vw, vh, o) for (vw, vh, vp), o in zip(views, outs)]
    reqs.append(ctx.probe_request(origin, 20.0, dims, nr, 7, 256.0, probe_out))
    bbest, btable = ctx.autotune_batch(reqs)
    total_rays = sum(vw * vh for vw, vh, _ in views) + int(np.prod(dims)) * nr
    ms5 = graph_ms(ctx, stream, lambda: ctx.trace_batch(reqs, bbest), 20)
    ctx.trace_batch(reqs, bbest)
    ctx.sync()
    times = ctx.batch_times(len(reqs))
    view_t = np.array(times[:16])

    def plain():
        for (vw, vh, vp), o in zip(views, outs):
            ctx.trace_primary(vp, vw, vh, o)
        ctx.trace_probes(origin, 20.0, dims, nr, 7, 256.0, probe_out)

    ms5p = graph_ms(ctx, stream, plain, 20)
    return {"rays": total_rays, "ms": round(ms5, 5), "grays": round(total_rays / (ms5 * 1e-3) / 1e9, 4),
            "view_latency_ms_p50": round(float(np.percentile(view_t, 50)), 5),
            "view_latency_ms_p99": round(float(np.percentile(view_t, 99)), 5),
            "probes_done_ms": round(float(times[16]), 5),
            "queue": {"tile": [bbest.tile_w, bbest.tile_h], "chunk": bbest.chunk, "refill": bbest.refill,
                      "traversal": bbest.traversal},
            "sequential_plain_launches_ms": round(ms5p, 5),
            "view_sizes": [[vw, vh] for vw, vh, _ in views]}


def _errors(x, ref, mask):
    """MSE and relMSE (squared error / (ref^2 + 0.01)) over the non-sky pixels' RGB."""
    d = (x[mask, :3] - ref[mask, :3]).double()
    r = ref[mask, :3].double()
    return float((d * d).mean()), float((d * d / (r * r + 1e-2)).mean())


def _reference(ctx, pu, W, H, gbp, gbn, bounces):
    """64-spp traced GI (samples 0..63 of salt 0x5EED, 8 per launch)."""
    ref = torch.zeros(W * H * 4, dtype=torch.float32, device=gbp.device)
    part = torch.zeros_like(ref)
    for c in range(8):
        ctx.trace_gi_path(pu, W, H, gbp, gbn, part, frame=c, spp=8, salt=0x5EED, bounces=bounces)
        ref += part
    return ref / 8.0


def gi_cache_config(ctx, stream, args, pu, W, H, gbp, gbn):
    """Learned GI radiance cache on the fixed world at 1920x1080: traced GI at 4 spp against the cached GI
    (online training on 65,536 traced samples per frame in 4 Adam steps, then the render), both judged by
    MSE / relMSE against a 64-spp traced reference of the same frame, for the two-bounce GI of the reference's
    giPath and for three-bounce GI (the cache's three-bounce samples: with terminate 1 each training target
    traces two bounces after bounce 1).  Static camera: each cache variant first trains for --cache-frames
    frames; the timed frame includes its sample tracing, training and render.  Moving camera: every method
    runs a 64-frame fly-through (2 voxels per frame) from a cold cache and is judged on its last 32 frames."""
    dev = torch.device("cuda:0")
    n = W * H
    mask = (gbp.view(-1, 4)[:, 3] >= 0.0)
    frame = args.cache_frames - 1
    res = {"reference": "64-spp traced GI (samples 0..63 of salt 0x5EED) with the method's bounce count",
           "pixels_with_gi": int(mask.sum().item())}
    for bounces in (2, 3):
        refv = _reference(ctx, pu, W, H, gbp, gbn, bounces).view(-1, 4)
        sfx = "" if bounces == 2 else "_3bounce"
        plain = {}
        for spp in (1, 4):
            o = torch.zeros(n * 4, dtype=torch.float32, device=dev)
            ms = graph_ms(ctx, stream, lambda: ctx.trace_gi_path(pu, W, H, gbp, gbn, o, frame=frame, spp=spp,
                                                                 salt=0x2f6b9a31, bounces=bounces), 20)
            mse, rel = _errors(o.view(-1, 4), refv, mask)
            plain[spp] = (ms, mse)
            res[f"traced_{spp}spp{sfx}"] = {"ms_per_frame": round(ms, 4), "mse": round(mse, 6), "relmse": round(rel, 5)}
        for term in (0, 1):
            g = vf.GiCache(ctx, terminate=term, bounces=bounces)
            o = torch.zeros(n * 4, dtype=torch.float32, device=dev)
            for f in range(args.cache_frames):
                g.frame(pu, W, H, gbp, gbn, o, f)
            st = g.stats()
            mse, rel = _errors(o.view(-1, 4), refv, mask)
            ms = graph_ms(ctx, stream, lambda: g.frame(pu, W, H, gbp, gbn, o, frame), 20)
            key = ("cache_at_gbuffer_point" if term == 0 else "cache_after_bounce1") + sfx
            res[key] = {"terminate": term, "bounces": bounces, "ms_per_frame": round(ms, 4), "mse": round(mse, 6),
                        "relmse": round(rel, 5), "speedup_vs_traced_4spp": round(plain[4][0] / ms, 3),
                        "mse_vs_traced_4spp": round(mse / plain[4][1], 4),
                        "frames_trained": args.cache_frames, "fallback": st["fallback"], "queries": st["n_queries"],
                        "running_error": round(st["running_error"], 5)}
            if term == 0 and bounces == 2:
                # matrix-core rate: the training step and the inference on their own
                samples = torch.zeros(65536 * 64, dtype=torch.uint8, device=dev)
                g.samples(pu, W, H, gbp, gbn, frame, samples)
                ms_t = graph_ms(ctx, stream, lambda: g.train(samples, 65536), 20)
                q = _random_queries(n, dev)
                h = torch.zeros(n * 4, dtype=torch.int16, device=dev)
                ms_i = graph_ms(ctx, stream, lambda: g.infer(q, n, h), 20)
                fl_t, fl_i = 65536 * 3 * FWD_FLOPS, n * FWD_FLOPS
                res["mfma"] = {"peak_tflops_bf16_dense": 2500.0,
                               "train_65536": {"ms": round(ms_t, 4), "tflops": round(fl_t / ms_t / 1e9, 2),
                                               "frac": round(fl_t / ms_t / 1e9 / 2500.0, 5)},
                               f"infer_{n}": {"ms": round(ms_i, 4), "tflops": round(fl_i / ms_i / 1e9, 2),
                                              "frac": round(fl_i / ms_i / 1e9 / 2500.0, 5)},
                               "flops_per_query": FWD_FLOPS}
            g.close()
    res["moving_camera"] = gi_cache_moving(ctx, stream, W, H)
    log(f"GI cache: traced 4spp {res['traced_4spp']}, cache@P {res['cache_at_gbuffer_point']}, "
        f"cache@bounce1 {res['cache_after_bounce1']}, 3-bounce cache@bounce1 {res['cache_after_bounce1_3bounce']}")
    return res


def gi_cache_moving(ctx, stream, W, H, frames=64, judged=32, step=2.0):
    """The GI methods on a fly-through of the fixed world: the camera moves `step` voxels per frame along its
    horizontal heading (yaw 0.8, pitch -0.35 from (37.3, 96.4, 41.7)); each frame traces its G-buffer, then
    every method renders it (caches from cold, warm-up frames included).  Reported per method: mean device
    ms per frame and mean MSE / relMSE against the frame's 64-spp reference over the last `judged` frames."""
    dev = torch.device("cuda:0")
    n = W * H
    hits = torch.empty(n * HIT_BYTES, dtype=torch.uint8, device=dev)
    gbp = torch.empty(n * 4, dtype=torch.float32, device=dev)
    gbn = torch.empty(n, dtype=torch.int32, device=dev)
    yaw, pitch = 0.8, -0.35
    p0 = np.array([37.3, 96.4, 41.7], np.float64)
    head = np.array([math.sin(yaw), 0.0, math.cos(yaw)])
    methods = {}
    for bounces in (2, 3):
        sfx = "" if bounces == 2 else "_3bounce"
        methods[f"traced_4spp{sfx}"] = ("traced", bounces, None)
        for term in (0, 1):
            key = ("cache_at_gbuffer_point" if term == 0 else "cache_after_bounce1") + sfx
            methods[key] = ("cache", bounces, vf.GiCache(ctx, terminate=term, bounces=bounces))
    outs = {k: torch.zeros(n * 4, dtype=torch.float32, device=dev) for k in methods}
    acc = {k: {"ms": 0.0, "mse": 0.0, "relmse": 0.0, "fallback_frames": 0} for k in methods}
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for f in range(frames):
        pos, right, up, fwd = vf.fly_camera(p0 + head * (step * f), yaw, pitch)
        pu = vf.make_pass(pos, right, up, fwd, 70.0, f)
        ctx.trace_primary(pu, W, H, hits, gbp, gbn)
        judge = f >= frames - judged
        refs = {b: _reference(ctx, pu, W, H, gbp, gbn, b).view(-1, 4) for b in (2, 3)} if judge else None
        mask = gbp.view(-1, 4)[:, 3] >= 0.0
        for k, (kind, bounces, g) in methods.items():
            o = outs[k]
            e0.record(stream)
            if kind == "traced":
                ctx.trace_gi_path(pu, W, H, gbp, gbn, o, frame=f, spp=4, salt=0x2f6b9a31, bounces=bounces)
            else:
                g.frame(pu, W, H, gbp, gbn, o, f)
            e1.record(stream)
            torch.cuda.synchronize()
            if not judge:
                continue
            acc[k]["ms"] += e0.elapsed_time(e1)
            mse, rel = _errors(o.view(-1, 4), refs[bounces], mask)
            acc[k]["mse"] += mse
            acc[k]["relmse"] += rel
            if g is not None:
                acc[k]["fallback_frames"] += int(g.stats()["fallback"])
    out = {"frames": frames, "judged_frames": judged, "voxels_per_frame": step}
    for k, (kind, bounces, g) in methods.items():
        a = acc[k]
        out[k] = {"bounces": bounces, "ms_per_frame": round(a["ms"] / judged, 4), "mse": round(a["mse"] / judged, 6),
                  "relmse": round(a["relmse"] / judged, 5)}
        if g is not None:
            out[k]["fallback_frames"] = a["fallback_frames"]
            g.close()
    for k in methods:
        if k.startswith("cache"):
            base = out["traced_4spp" + ("_3bounce" if k.endswith("_3bounce") else "")]
            out[k]["speedup_vs_traced_4spp"] = round(base["ms_per_frame"] / out[k]["ms_per_frame"], 3)
            out[k]["mse_vs_traced_4spp"] = round(out[k]["mse"] / base["mse"], 4)
    return out


# forward FLOPs per query: 2 * (96*64 + 4*64*64 + 64*3)
FWD_FLOPS = 2 * (96 * 64 + 4 * 64 * 64 + 64 * 3)


def _random_queries(n, dev):
    """n query records (64 B each) with positions in the fixed world's box and unit normal / view
    vectors: the inference cost does not depend on the values."""
    g = torch.Generator(device="cpu").manual_seed(3)
    rec = torch.zeros(n, 16, dtype=torch.float32)
    rec[:, 0:3] = torch.rand(n, 3, generator=g) * torch.tensor([256.0, 128.0, 256.0])
    nv = torch.randn(n, 3, generator=g)
    rec[:, 4:7] = nv / nv.norm(dim=1, keepdim=True)
    vv = torch.randn(n, 3, generator=g)
    rec[:, 8:11] = vv / vv.norm(dim=1, keepdim=True)
    rec[:, 12:15] = 1.0
    return rec.to(dev).view(torch.uint8).view(-1)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(world, win, wb, cam_pos, right, up, fwd, thf, scene, gpu_hits, gpu_gbn, gpu_sh, gpu_shh, min_seconds):
    """The C oracle (OpenMP over the host cores this process may use, at most the 16 a one-GPU share of
    the machine gets) on the same workload: the whole frame's primary and sun rays, repeated until
    min_seconds have passed.  Its outputs double as a bit-exact check of the GPU's whole frame: primary
    hits, gbNormal (foliage alpha included), shadowRaw and the sun rays' first-segment hits."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    affinity = len(os.sched_getaffinity(0))
    cores = min(affinity, 16)
    t0 = time.time()
    ow = O.OracleWorld(world.arrays(), build_skytop=False)
    log(f"oracle world built in {time.time() - t0:.1f}s")
    cam = {"pos": cam_pos, "right": right, "up": up, "fwd": fwd, "tan_half_fov": thf}
    sun = np.array(scene.sun_dir, np.float32)
    rays, frames = 0, 0
    t0 = time.perf_counter()
    while True:
        oh, opos, onrm, cp = ow.pass_primary_rows(cam, W, H, 0, 1, H, threads=cores)
        osh, oshh, cs = ow.pass_shadow(sun, W, H, opos, onrm, threads=cores)
        rays += cp["rays"] + cs["rays"]
        frames += 1
        dt = time.perf_counter() - t0
        if dt >= min_seconds:
            break
    g = vf.hits_to_numpy(gpu_hits)
    mism = int(((g.view(np.uint8).reshape(-1, 20) != oh.view(np.uint8).reshape(-1, 20)).any(1)).sum())
    gn = gpu_gbn.cpu().numpy().view(np.uint32)
    nrm_mism = int((gn != onrm).sum())
    want_sh = (osh[:, 0].astype(np.float16).view(np.uint16).astype(np.uint32) |
               (osh[:, 1].astype(np.float16).view(np.uint16).astype(np.uint32) << 16))
    sh_mism = int((gpu_sh.cpu().numpy().view(np.uint32) != want_sh).sum())
    gs = vf.hits_to_numpy(gpu_shh)
    shh_mism = int(((gs.view(np.uint8).reshape(-1, 20) != oshh.view(np.uint8).reshape(-1, 20)).any(1)).sum())
    nb = onrm
    N = np.stack([((nb >> (8 * k)) & 255).astype(np.float32) / 255 * 2 - 1 for k in range(3)], 1)
    back_lit = int(((opos[:, 3] >= 0) & ((nb >> 24) == 204) & ((N * sun).sum(1) <= 0)).sum())
    return {
        "value": round(rays / dt / 1e9, 6),
        "unit": "Grays/s",
        "cores": cores,
        "cpu_model": _cpu_model(),
        "nproc": os.cpu_count(),
        "affinity_cpus": affinity,
        "kind": "port",
        "sample": f"the whole 1920x1080 frame ({cp['rays']} primary + {cs['rays']} sun rays) x {frames} repeats; "
                  f"the project's C oracle (FP32 restatement of traceVolumeS) with OpenMP on {cores} threads (the "
                  f"CPU share of one GPU, of {affinity} usable); lavapipe baseline unavailable (no cargo / shaderc / "
                  f"Vulkan loader on the box)",
        "seconds": round(dt, 3),
        "hit_exact_check": {"rays": int(g.size), "mismatches": mism, "gb_normal_mismatches": nrm_mism,
                            "sun_rays": int(cs["rays"]), "back_lit_foliage_rays": back_lit,
                            "sun_mismatches": sh_mism, "sun_hit_mismatches": shh_mism},
    }


if __name__ == "__main__":
    main()
