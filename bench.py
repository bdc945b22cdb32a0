#!/usr/bin/env python3
"""Benchmark: Msamples/s shaded at 2x1080x1200 MSAA4 over a 262,144-triangle scene, with
p50/p99 frame latency.

One step = one stereo frame: both eyes (1080x1200, 4 subsamples each, barrel distortion
k1=0.22 k2=0.24) rendered through the full GPU path (frustum setup, block + tile
traversal, intersect, resolve, remap) over the seeded room scene (1,024 meshes x 256
triangles, 4 point + 1 directional light). Inputs are resident on the GPU before timing.

Prints one JSON line (see the driver contract in the repository docs).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import hvvr_amd as H  # noqa: E402

W, HGT, A, EYES, MESHES = 1080, 1200, 4, 2, 1024
PEAK_FP32 = 157.3e12  # MI355X dense FP32 vector FLOP/s
PEAK_HBM = 8.0e12  # MI355X HBM3E bytes/s
STAGES = H.STAGES


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, views, budget_s=20.0, threads=None):
    """The reference's CPU stage beside a scalar restatement of its GPU kernels, timed on
    a bounded sample of the stereo frame (oracle/cpu_baseline.cpp): per task of 2 blocks
    (traversal.cpp:249-254), frustum setup, the AVX2 block and tile traversal
    (traverse_avx.h:104-527) with FTZ/DAZ, then the scalar C intersect and resolve of the
    blocks' samples, on a thread pool. The reference sizes its pool by
    hardware_concurrency (raycaster.cpp:33-37); this process uses the CPUs it is given
    (OMP_NUM_THREADS, else its affinity mask), and reports both."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    available = len(os.sched_getaffinity(0))
    threads = threads or min(available, int(os.environ.get("OMP_NUM_THREADS") or available))
    st = O.cpu_baseline_run(scene, views, A, threads=threads, budget_s=budget_s)
    value = st.subsamples / st.seconds / 1e6
    busy = st.frustumSeconds + st.traverseSeconds + st.intersectSeconds + st.resolveSeconds
    split = {k: round(getattr(st, k + "Seconds") / busy, 3) for k in ("frustum", "traverse", "intersect", "resolve")}
    blocks = sum(v.block_count for v in views)
    return {"value": round(value, 4), "unit": "Msamples/s", "cores": int(st.threads), "kind": "port",
            "cpu_model": cpu_model(), "hardware_concurrency": int(st.hardwareConcurrency),
            "cpus_available": available, "stage_share": split,
            "sample": f"{st.blocks} block renders ({blocks} blocks per stereo frame, 8192 samples x {A} subsamples "
                      f"each, tasks of 2 blocks in a seeded order over both eyes) in {st.seconds:.1f} s on "
                      f"{st.threads} threads: frustum setup + AVX2 block/tile traversal (FTZ/DAZ) + scalar C "
                      f"intersect + resolve; scene refit excluded"}


def copy_calibration(mib):
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4}))


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE passes, each scaled by the same counter's reading for a
    plain copy of known size), or None if the summary has no entry for it."""
    import re

    path = os.path.join(ROOT, "profiles", "pmc_traffic_stereo_1080x1200_a4.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"```json\n(.*?)\n```", open(path).read(), re.S)
    if not m:
        return None
    table = json.loads(m.group(1))
    entry = table.get(kernel)
    if entry is None:  # the summary names template kernels in full: match the name before '<'
        entry = next((v for k, v in table.items() if k.split("<")[0] == kernel.split("<")[0]), None)
    return None if entry is None else int(entry["read_bytes"] + entry["write_bytes"])


def measured_counters(kernel):
    """The committed SQ/L2 counter summary's row for `kernel` (l2_hit, valu_busy,
    valu_util, salu_share, ...), or {} if it has none."""
    path = os.path.join(ROOT, "profiles", "pmc_counters_stereo_1080x1200_a4.md")
    if not os.path.exists(path):
        return {}
    head = None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if cells and cells[0] == "kernel":
            head = cells
        elif head and cells and cells[0].split("<")[0] == kernel.split("<")[0] and len(cells) == len(head):
            return {h: float(c) for h, c in zip(head[1:], cells[1:])}
    return {}


def roofline(stage_ms, survivors, entries, visits, tested, kernel_ms=None):
    """Roofline of the dominant kernel, per launch (one launch per eye).
    tile_lists: algorithmic bytes = 128 B per BVH node visit + 4 B per list entry written,
    against HBM; the node reads mostly hit L2, so beside it stand the HBM bytes the
    counters measured (`traffic`, its fraction of 8 TB/s) and the L2 hit rate: the kernel
    is bound by the issue and dependency latency of its stack walk, not by either;
    intersect: FP32 VALU flops = N_pair (80 + 16 A) with N_pair the (sample, survivor) pairs
    the kernel tested (the tile-cull survivors less those its occlusion skip left out),
    against the FP32 vector figure. `kernel_ms`: the standard tile-list kernel's own average
    duration over the timed frames (HIP events its dispatch records), used for tile_lists
    when given; otherwise the stage's events."""
    dom = STAGES[int(np.argmax(stage_ms))]
    t = stage_ms[STAGES.index(dom)] * 1e-3
    timed_live = dom == "tile_lists" and kernel_ms is not None
    if timed_live:
        t = kernel_ms * 1e-3
    if dom == "intersect":
        name = "intersectKernel<4>"
        flops = tested * (80 + 16 * A)
        ach = flops / t
        pmc = measured_counters(name)
        return {"kernel": name, "bound": "valu", "achieved": round(ach / 1e12, 3),
                "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s", "frac": round(ach / PEAK_FP32, 4),
                "traffic": measured_traffic(name), "valu_busy_pct": pmc.get("valu_busy"),
                "valu_lane_util_pct": pmc.get("valu_util"),
                "note": f"FP32 VALU work (no matrix cores on this path) against the 157.3 TFLOP/s FP32 vector "
                        f"figure (78.6 without packed FP32, which the build disables); F = N_pair*(80+16A), "
                        f"N_pair={int(tested)} (sample, survivor) pairs tested per launch of the {int(survivors)} "
                        f"tile-cull survivor pairs (the rest a wave's rays cannot reach in front of their hits)"}
    bytes_ = 128 * visits + 4 * entries
    name = {"tile_lists": "tileListKernel", "resolve": "resolveKernel<4>"}.get(dom, dom)
    traffic = measured_traffic(name)
    pmc = measured_counters(name)
    return {"kernel": name, "bound": "hbm", "achieved": round(bytes_ / t / 1e9, 3), "peak": PEAK_HBM / 1e9,
            "unit": "GB/s", "frac": round(bytes_ / t / PEAK_HBM, 5), "traffic": traffic,
            "kernel_ms": round(t * 1e3, 4),
            "kernel_ms_source": "timed frames, dispatch events" if timed_live else "per-stage events",
            "hbm_counter_frac": None if traffic is None else round(traffic / t / PEAK_HBM, 5),
            "l2_hit_pct": pmc.get("l2_hit"), "salu_share": pmc.get("salu_share"),
            "valu_busy_pct": pmc.get("valu_busy"),
            "note": f"issue- and latency-bound stack walk; algorithmic bytes = 128 B x {int(visits)} node visits + "
                    f"4 B x {int(entries)} list entries per launch, mostly L2 hits (l2_hit_pct); the HBM bytes "
                    f"the counters measured are `traffic` (hbm_counter_frac of 8 TB/s)"}


def serving_bench(args):
    """Serving mix: Poisson arrivals of view requests over resident scenes of 100k, 262k, 1M
    and 2M triangles (the room generator at 391, 1,024, 4,096 and 8,192 meshes); each
    request picks a scene, a resolution log-uniform in area between 256x256 and a
    2x1080x1200 stereo pair (six size classes; the stereo class is two eye views submitted
    together, done when both are), an MSAA rate (1/2/4/8) and, for half of them, polar
    foveation (the feb2017 pattern reconstructed at that size). Admission by a device
    budget, FP16 shading. Reports completed requests/s and submit-to-image latency."""
    rng = np.random.default_rng(args.seed)
    sizes = [(256, 256), (360, 400), (540, 600), (760, 840), (1080, 1200)]
    classes = sizes + [(1080, 1200)]  # the last class: a stereo pair
    rates = [1, 2, 4, 8]
    t0 = time.perf_counter()
    scenes = [H.HostScene(H.SCENE_ROOM, k, n) for k, n in enumerate((391, 1024, 4096, 8192))]
    screens = [[H.HostView(H.VIEW_PRESET_ROOM_EYE, e, w, h) for e in range(2)] for w, h in sizes]
    fov = H.HostFoveated(H.HostFoveated.FEB2017)
    log(f"serving setup: scenes {[len(s.triShade) for s in scenes]}, {time.perf_counter() - t0:.1f} s")
    ctx = H.GpuContext(0)
    gscenes = [H.GpuScene(ctx, s) for s in scenes]
    for g in gscenes:
        g.update()
    ctx.synchronize()
    log(f"scenes resident ({time.perf_counter() - t0:.1f} s)")
    srv = H.Server(ctx, int(args.budget_gib * (1 << 30)), args.slots, shading_bits=16)
    shape = {}
    for i in range(len(sizes)):
        for e in range(2):
            for a in rates:
                shape[(i, e, a, False)] = srv.add_shape(screens[i][e], a)
                shape[(i, e, a, True)] = srv.add_foveated_shape(fov, screens[i][e], a)
    log(f"{len(shape)} shapes ({time.perf_counter() - t0:.1f} s)")
    n = args.requests
    arrival = np.cumsum(rng.exponential(1.0 / args.rate, n))
    areas = np.log([w * h * (2 if c == len(sizes) else 1) for c, (w, h) in enumerate(classes)])
    pick = rng.uniform(areas[0], areas[-1], n)
    cls = np.abs(pick[:, None] - areas[None, :]).argmin(1)
    reqs = [(int(rng.integers(4)), int(cls[k]), rates[int(rng.integers(4))], bool(rng.random() < 0.5))
            for k in range(n)]

    def views_of(c):
        return [(len(sizes) - 1, 0), (len(sizes) - 1, 1)] if c == len(sizes) else [(c, 0)]

    def submit(k):
        sc, c, a, fv = reqs[k]
        eye = None
        if fv:
            y, p = np.radians(rng.uniform(-20, 20, 2))
            eye = np.array([np.sin(y) * np.cos(p), np.sin(p), -np.cos(y) * np.cos(p)], np.float32)
        return [srv.submit(gscenes[sc], shape[(i, e, a, fv)], screens[i][e].camera_to_world, eye)[0]
                for i, e in views_of(c)]

    # warm-up: the densest scene first sizes each shape's triangle lists (views created
    # later start at that size), then one request per scene at once leaves four views of
    # every shape in the pool
    for key in sorted(shape):
        m = screens[key[0]][key[1]].camera_to_world
        srv.submit(gscenes[-1], shape[key], m, None)
        srv.drain()
        for g in gscenes:
            srv.submit(g, shape[key], m, None)
        srv.drain()
    warm = srv.stats()
    log(f"warm-up: {warm.viewsCreated} views, {warm.listGrowths} list growths, "
        f"{warm.peakDeviceBytes / (1 << 30):.2f} GiB ({time.perf_counter() - t0:.1f} s)")
    ids = []
    t_start = time.perf_counter()
    k = 0
    views_submitted = views_completed = 0
    next_log = 1.0
    while True:
        now = time.perf_counter() - t_start
        while k < n and arrival[k] <= now:
            ids.append(submit(k))
            views_submitted += len(ids[-1])
            k += 1
        views_completed += srv.poll()
        if k == n and views_completed == views_submitted:
            break
        if now > next_log:
            st = srv.stats()
            log(f"t={now:.1f} s: requests submitted {k}, views completed {views_completed}, in flight "
                f"{st.inFlight}, queued {st.queued}")
            next_log += 1.0
    elapsed = time.perf_counter() - t_start
    st = srv.stats()
    lat = []
    for rids in ids:
        ls = [srv.done(r) for r in rids]
        assert all(d for d, _ in ls)
        lat.append(max(l for _, l in ls))
    lat = np.array(lat)
    def cls_of(r):
        sc, c, a, fv = r
        return f"{['100k', '262k', '1M', '2M'][sc]}{' foveated' if fv else ''}"

    by_class = {}
    for r, l in zip(reqs, lat):
        by_class.setdefault(cls_of(r), []).append(l)
    by_class = {k: {"n": len(v), "p50_ms": round(float(np.percentile(v, 50)), 2),
                    "p99_ms": round(float(np.percentile(v, 99)), 2)} for k, v in sorted(by_class.items())}
    subsamples = 0
    for sc, c, a, fv in reqs:
        for i, e in views_of(c):
            w, h = sizes[i]
            subsamples += (fov.valid_sample_count if fv else w * h) * a
    out = {
        "metric": "serving mix: completed requests/s and submit-to-image latency (ms)",
        "value": round(n / elapsed, 2), "unit": "requests/s", "n_gpus": 1, "higher_is_better": True,
        "requests": n, "offered_rate": args.rate, "elapsed_s": round(elapsed, 3),
        "msamples_per_s": round(subsamples / elapsed / 1e6, 1),
        "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "mean_ms": round(float(lat.mean()), 3), "latency_by_scene": by_class,
        "slowest": [{"request": int(k), "arrival_s": round(float(arrival[k]), 3), "class": cls_of(reqs[k]),
                     "size_class": int(reqs[k][1]), "msaa": int(reqs[k][2]), "ms": round(float(lat[k]), 2)}
                    for k in np.argsort(lat)[::-1][:5]],
        "peak_in_flight": st.peakInFlight, "views_created": st.viewsCreated, "views_evicted": st.viewsEvicted,
        "list_growths": st.listGrowths, "failed": st.failed,
        "peak_device_gib": round(st.peakDeviceBytes / (1 << 30), 2), "dtype": "f32 intersect, f16 shade",
        "data": "synthetic: seeded room scenes of 100k/262k/1M/2M triangles",
        "config": {"workload": "serving mix", "slots": args.slots, "budget_gib": args.budget_gib,
                   "sizes": sizes, "stereo_class": "2x1080x1200", "msaa": rates, "foveated_fraction": 0.5},
    }
    srv.close()
    print(json.dumps(out))


def batch_bench(args):
    """Batched multi-view frames with a per-frame BVH refit: 64 views of the 1M-triangle
    room scene (4,096 meshes x 256 triangles), resolutions log-uniform in area between
    128x128 and 1920x1080, MSAA 1/2/4/8 uniformly, FP16 shading; every frame displaces the
    vertices of 10% of the scene by N(0, 0.02 m) and refits. The batch runs as one
    captured HIP graph (argument upload, refit, the kernel chain) with an autotuned launch
    shape; the same frames rendered view by view (the plain pipeline: one view's kernels
    after another's, on one stream) are timed beside it."""
    rng = np.random.default_rng(args.seed)
    t0 = time.perf_counter()
    scene = H.HostScene(H.SCENE_ROOM, 0, 4096)
    lo, hi = np.log(128 * 128), np.log(1920 * 1080)
    aspect = 1920 / 1080
    specs = []
    for _ in range(args.views):
        area = np.exp(rng.uniform(lo, hi))
        h = int(round(np.sqrt(area / aspect)))
        w = int(round(h * aspect))
        specs.append((max(w, 8), max(h, 8), int(rng.choice([1, 2, 4, 8]))))
    hviews = [H.HostView(H.VIEW_PRESET_ROOM_EYE, k & 1, w, h) for k, (w, h, _) in enumerate(specs)]
    ctx = H.GpuContext(0)
    gs = H.GpuScene(ctx, scene)
    gviews = []
    for hv, (_, _, a) in zip(hviews, specs):
        gv = H.GpuView(ctx, hv, a)
        gv.set_shading_precision(16)
        gviews.append(gv)
    subsamples = sum(hv.valid_sample_count * a for hv, (_, _, a) in zip(hviews, specs))
    log(f"batch setup: {len(scene.triShade)} triangles, {len(gviews)} views, {subsamples / 1e6:.1f} M subsamples "
        f"per frame ({time.perf_counter() - t0:.1f} s)")
    verts = scene.verts.copy()
    n_move = len(verts) // 10

    def animate():
        first = int(rng.integers(0, len(verts) - n_move))
        moved = verts[first:first + n_move].copy()
        moved["pos"] += rng.normal(0, 0.02, size=(n_move, 3)).astype(np.float32)
        gs.set_vertices(first, moved)

    def cams(k):
        out = []
        for i, hv in enumerate(hviews):
            m = np.array(hv.camera_to_world, np.float32).reshape(4, 4).copy()
            m[3, :3] += np.float32(0.01 * ((k + i) % 7))
            out.append(m.reshape(16))
        return out

    batch = H.GpuBatch(ctx, gviews)
    batch.set_refit(True)
    # size every view's triangle lists for this scene, then tune the launch shape
    for _ in range(2):
        batch.render(gs, cams(0))
        batch.synchronize()
        for gv in gviews:
            need, cap, of = gv.list_demand()
            if of and need > cap:
                gv.reserve_lists(need + need // 4)
    shape, tuned_ms = batch.autotune(gs, cams(0), frames_per_trial=3)
    log(f"autotuned shape {shape.as_dict()}: {tuned_ms:.2f} ms ({time.perf_counter() - t0:.1f} s)")
    batch.set_graph(True)
    for k in range(args.warmup):
        animate()
        batch.render(gs, cams(k))
        batch.synchronize()
    lat, gpu = [], []
    t_start = time.perf_counter()
    for k in range(args.steps):
        animate()
        f0 = time.perf_counter()
        batch.render(gs, cams(k))
        gpu.append(batch.synchronize())
        lat.append((time.perf_counter() - f0) * 1e3)
        log(f"batched frame {k}: {lat[-1]:.2f} ms")
    elapsed = time.perf_counter() - t_start
    for gv in gviews:
        need, cap, of = gv.list_demand()
        if of:
            raise SystemExit("a batched frame overflowed its lists or a traversal stack")
    batched_rate = subsamples * args.steps / (sum(lat) * 1e-3) / 1e6
    # the plain pipeline over the same frames: refit, then each view's own launches
    plain = []
    for k in range(args.steps):
        animate()
        ctx.synchronize()
        f0 = time.perf_counter()
        gs.update()
        for gv, m in zip(gviews, cams(k)):
            gv.render(gs, m)
        ctx.synchronize()
        plain.append((time.perf_counter() - f0) * 1e3)
        log(f"plain frame {k}: {plain[-1]:.2f} ms")
    plain_rate = subsamples * args.steps / (sum(plain) * 1e-3) / 1e6
    out = {
        "metric": "batched multi-view Msamples/s (64 views, 1M tris, refit per frame, graph replay)",
        "value": round(batched_rate, 1), "unit": "Msamples/s", "n_gpus": 1, "higher_is_better": True,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "gpu_ms_per_frame": round(float(np.median(gpu)), 3),
        "plain_msamples_per_s": round(plain_rate, 1), "plain_ms_per_frame": round(float(np.median(plain)), 3),
        "speedup_vs_plain": round(batched_rate / plain_rate, 3),
        "launch_shape": shape.as_dict(), "subsamples_per_frame": int(subsamples),
        "dtype": "f32 intersect, f16 shade", "data": "synthetic: seeded room scene, 4,096 meshes",
        "config": {"workload": "batched multi-view + refit", "views": args.views,
                   "sizes": [[w, h, a] for w, h, a in specs]},
    }
    batch.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--pipeline", choices=["plain", "batched", "graph"], default="batched",
                    help="batched (the default): both eyes in one submission (hvvr_gpu_render_views, what "
                         "Raycaster::render issues for its cameras), overlapping on separate streams; plain: "
                         "one view after the other on one stream; graph: both eyes as one batch (each kernel "
                         "launched once for both), captured as a HIP graph and replayed, launch shape autotuned")
    ap.add_argument("--shading", type=int, choices=[32, 16], default=32,
                    help="shading precision: 32 = FP32 (the oracle's arithmetic), 16 = packed-FP16 colour terms")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (5 launches), the reference "
                         "workload for scaling FETCH_SIZE / WRITE_SIZE counters to bytes")
    ap.add_argument("--workload", choices=["stereo", "serving", "batch"], default="stereo",
                    help="stereo: the headline stereo frame; serving: the request-serving mix; batch: 64 views "
                         "per launch with a per-frame refit")
    ap.add_argument("--views", type=int, default=64, help="batch: views per frame")
    ap.add_argument("--requests", type=int, default=2000, help="serving: requests in the timed run")
    ap.add_argument("--rate", type=float, default=2000.0, help="serving: Poisson arrival rate (requests/s)")
    ap.add_argument("--slots", type=int, default=8, help="serving: requests on the device at once")
    ap.add_argument("--budget-gib", type=float, default=64.0, help="serving: device budget for views")
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU: the render path shards by view "
                                   "(eye / request) with no per-frame exchange, so N>1 is N independent "
                                   "copies; run with --gpus 1"}))
        sys.exit(2)
    if args.workload == "serving":
        serving_bench(args)
        return
    if args.workload == "batch":
        batch_bench(args)
        return

    t0 = time.perf_counter()
    scene = H.HostScene(H.SCENE_ROOM, 0, MESHES)
    views = [H.HostView(H.VIEW_PRESET_ROOM_EYE, e, W, HGT) for e in range(EYES)]
    log(f"setup: {len(scene.triShade)} triangles, {views[0].tile_count} tiles/eye "
        f"({time.perf_counter() - t0:.1f} s)")
    ctx = H.GpuContext(0)
    gs = H.GpuScene(ctx, scene)
    gvs = [H.GpuView(ctx, v, A) for v in views]
    for gv in gvs:
        gv.set_shading_precision(args.shading)
    batch, shape = None, None
    if args.pipeline == "graph":
        batch = H.GpuBatch(ctx, gvs)
        shape, _ = batch.autotune(gs, frames_per_trial=5)
        batch.set_graph(True)
        log(f"autotuned launch shape {shape.as_dict()}")

    def frame():
        if args.pipeline == "graph":
            batch.render(gs)
        elif args.pipeline == "batched":
            H.render_views(gs, gvs)
        else:
            for gv in gvs:
                gv.render(gs)

    def sync():
        if batch is not None:
            batch.synchronize()
        ctx.synchronize()

    for _ in range(args.warmup):
        frame()
    sync()

    # timed region: K stereo frames, each synchronised so its latency is measured too; in
    # the plain and batched pipelines the standard tile-list kernel's dispatch records HIP
    # events (profiling level 2: no packet added to the stream), read after each frame's
    # synchronise. Per-stage events (level 1) put a marker between every two kernels and
    # slow the frame ~10%, so the stage times come from a pass after the timed region.
    ctx.set_profiling(2 if batch is None else 0)
    lat = []
    kernel_ms = []
    sync()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        f0 = time.perf_counter()
        frame()
        sync()
        lat.append(time.perf_counter() - f0)
        if batch is None:
            for gv in gvs:
                kernel_ms.append(gv.tile_list_kernel_ms())
    sync()
    elapsed = time.perf_counter() - t_start
    ctx.set_profiling(False)
    samples_per_frame = EYES * W * HGT * A
    value = samples_per_frame * args.steps / elapsed / 1e6
    lat_ms = np.array(lat) * 1e3
    # per-stage times: the same views, one frame after the other in the same submission,
    # per-stage events on (a graph replay records none)
    stage_ms = np.zeros(len(STAGES))
    ctx.set_profiling(1)
    for _ in range(5):
        if batch is None:
            frame()
        else:
            for gv in gvs:
                gv.render(gs)
        ctx.synchronize()
        for gv in gvs:
            stage_ms += gv.stage_times()
    ctx.set_profiling(0)
    stage_ms /= 5 * EYES  # average per launch (one launch per eye)
    stats = [gv.stats() for gv in gvs]
    survivors = sum(s.survivorPairs for s in stats) / EYES
    entries = sum(s.listEntries for s in stats) / EYES
    for s in stats:
        if s.overflow:
            raise SystemExit("traversal budget overflow")

    visits = sum(s.nodeVisits for s in stats) / EYES
    tested = sum(s.testedPairs for s in stats) / EYES
    roof = roofline(stage_ms, survivors, entries, visits, tested,
                    float(np.mean(kernel_ms)) if kernel_ms else None)

    out = {
        "metric": "Msamples/s shaded at 2x1080x1200 MSAA4, 262k tris; p99 frame latency (ms)",
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 4),
        "p99_ms": round(float(np.percentile(lat_ms, 99)), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.shading == 32 else "f32 intersect, f16 shade",
        "data": "synthetic: seeded room scene (1,024 meshes x 256 triangles), grid samples with barrel distortion",
        "config": {"workload": f"stereo 2x1080x1200, MSAA 4, 262,144 triangles, {args.pipeline} pipeline",
                   "subsamples_per_frame": samples_per_frame, "pipeline": args.pipeline,
                   "shading_bits": args.shading, "launch_shape": shape.as_dict() if shape else None},
        "stage_ms": {k: round(float(v), 4) for k, v in zip(STAGES, stage_ms)},
        "list_entries_per_eye": int(entries),
        "survivor_pairs_per_eye": int(survivors),
        "tested_pairs_per_eye": int(tested),
        "node_visits_per_eye": int(visits),
        "node_tests_per_s": round(visits * 4 / (stage_ms[STAGES.index("tile_lists")] * 1e-3), 1),
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scene, views, args.cpu_budget)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
