"""Benchmark: nuclei segmentation (HistologicalEntities::segmentNuclei) on MI355X.

Workload (default): the full segmentNuclei chain on one 4096x4096 BGR uint8 synthetic H&E tile
(seeded generator nsc_synth_he, seed 2026), default chain parameters, one GPU.  A step is one
chain call on that tile with the input already resident in HBM.

Prints ONE JSON line: segmented Mpix/s (input pixels / wall time), p50/p99 tile latency, the
roofline of the imreconstruct<uchar> kernel inside the chain (3 B/px algorithmic: read marker,
read mask, write output) and the CPU baseline (the oracle chain on the host cores, bounded
sample).

--workload recon times imreconstruct<uchar> alone (8-connectivity) on the 4096x4096 grayscale
ellipse tile (background 220, ellipses U[60,120], noise sigma 8, marker = mask - 10 saturating):
the profiling workload for the reconstruction kernel.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size S] [--no-cpu] [--workload chain|recon]
"""
import argparse
import ctypes
import json
import os
import statistics
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# The served step runs 8 server slots, each with its own stream and three side streams; HIP maps a
# process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default), where kernels of
# unrelated streams wait behind each other: with 4 queues a kernel trace of the served step never
# shows more than 4 kernels running at once.  16 queues measured 5% more served throughput than 8
# and 8 about 7% more than 4 (the plain path is unchanged).  --hw-queues sets it (0 keeps the
# environment's value); it must be set before the HIP runtime starts (before torch touches the GPU),
# so it is read from the command line here, ahead of the argument parser.
def _hw_queues_arg(argv):
    for i, a in enumerate(argv):
        if a == "--hw-queues" and i + 1 < len(argv):
            return int(argv[i + 1])
        if a.startswith("--hw-queues="):
            return int(a.split("=", 1)[1])
    return 16


_HWQ = _hw_queues_arg(sys.argv[1:])
if _HWQ > 0:
    os.environ["GPU_MAX_HW_QUEUES"] = str(min(_HWQ, 32))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def host_cpus():
    """Cores this process may use: the affinity set, capped by a cgroup CPU quota when one is set."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) // int(period)))
    except Exception:
        pass
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    usable = min(aff, quota) if quota else aff
    return {"nproc": os.cpu_count(), "affinity": aff, "cgroup_quota_cpus": quota, "usable": usable, "model": model}


def oracle_cxxflags():
    try:
        for line in open(os.path.join(ROOT, "oracle", "Makefile")):
            if line.startswith("CXXFLAGS"):
                return "g++ " + line.split("=", 1)[1].strip()
    except Exception:
        pass
    return None


def cpu_baseline(size, tiles_per_thread=2):
    """Oracle chain (single-threaded restatement of the reference CPU path, which is itself
    single-threaded per tile) on the host cores: one thread per usable core, each running
    tiles_per_thread whole tiles (the reference scales out one tile per process), plus the
    single-core latency of one tile.  Bounded sample: about (1 + tiles_per_thread) tile times."""
    import numpy as np
    from nscale_amd.synth import synth_he
    from tests._lib import SegParams, oracle, ptr

    o = oracle()
    cpus = host_cpus()
    cores = cpus["usable"]
    tiles = [synth_he(size, size, 2026 + k) for k in range(min(cores, 8))]
    p = SegParams()
    o.orc_seg_params_default(ctypes.byref(p))

    def run(k, out, reps=1):
        lab = np.empty((size, size), np.int32)
        t = time.perf_counter()
        for r in range(reps):
            o.orc_segment_nuclei(ptr(tiles[(k + r) % len(tiles)]), size, size, ctypes.byref(p), ptr(lab), None, None)
        out[k] = time.perf_counter() - t

    lat = [0.0]
    run(0, lat)  # single-core latency of one tile (the process is otherwise idle)
    one_core = size * size / lat[0] / 1e6
    times = [0.0] * cores
    ths = [threading.Thread(target=run, args=(k, times, tiles_per_thread)) for k in range(cores)]
    t0 = time.perf_counter()
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    wall = time.perf_counter() - t0
    agg = cores * tiles_per_thread * size * size / wall / 1e6
    return {"value": round(agg, 3), "unit": "Mpix/s", "cores": cores, "kind": "port",
            "sample": "oracle segmentNuclei chain on %dx%d synthetic H&E tiles: %d threads x %d tiles each "
                      "(%.1f s wall); single-core %.3f Mpix/s (%.2f s/tile)"
                      % (size, size, cores, tiles_per_thread, wall, one_core, lat[0]),
            "one_core_mpix_s": round(one_core, 3), "tiles_per_thread": tiles_per_thread,
            "nproc": cpus["nproc"], "affinity_cpus": cpus["affinity"], "cgroup_quota_cpus": cpus["cgroup_quota_cpus"],
            "cpu_model": cpus["model"], "compiler": oracle_cxxflags()}


def bench_recon(args):
    import torch
    from nscale_amd import gpu
    from nscale_amd.synth import synth_mask, synth_marker

    S = args.size
    mask_h = synth_mask(S, S, 2026)
    mask = torch.from_numpy(mask_h).cuda()
    marker = torch.from_numpy(synth_marker(mask_h)).cuda()
    for _ in range(args.warmup):
        out = gpu.imreconstruct(marker, mask, 8)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record()
        out = gpu.imreconstruct(marker, mask, 8)
        b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = sorted(a.elapsed_time(b) for a, b in evs)
    avg = sum(ms) / len(ms)
    achieved = 3 * S * S / (avg * 1e-3) / 1e9
    out = {"metric": "imreconstruct<uchar> Mpix/s", "value": round(args.steps * S * S / wall / 1e6, 3),
           "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(wall * 1e3 / args.steps, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u8", "data": "synthetic",
           "config": {"workload": "imreconstruct<uchar> conn 8, %dx%d ellipse mask, marker = mask - 10" % (S, S)},
           "p50_ms": round(statistics.median(ms), 4),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": measured_traffic("recon_c0") if S == 4096 else None, "avg_ms": round(avg, 4)}}
    print(json.dumps(out), flush=True)


def measured_traffic(key):
    """HBM bytes per imreconstruct call measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    calibrated on a plain copy) for the 4096x4096 workloads; None if not recorded."""
    tf = os.path.join(ROOT, "profiles", "recon_u8_4096_traffic.json")
    try:
        return json.load(open(tf))[key]["bytes_per_call"]
    except Exception:
        return None


def checksum(lab):
    """Order-sensitive digest of a label image (int64 sums of labels and position-weighted labels)."""
    import torch
    v = lab.view(-1).to(torch.int64)
    w = torch.arange(v.numel(), device=v.device, dtype=torch.int64) % 1000003
    return int(v.sum().item()), int((v * w).sum().item())


def pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return float("nan")
    return xs[min(len(xs) - 1, int(round(q / 100.0 * (len(xs) - 1))))]


def serve_sizes(n, seed):
    from nscale_amd.synth import serve_sizes as sizes
    return sizes(n, seed)


def bench_serve(args):
    """Open-loop Poisson arrivals of tiles of the served size distribution, swept over rates; each
    rate reports achieved Mpix/s and p50/p99 request latency (submit to completion)."""
    import threading
    import numpy as np
    import torch
    from nscale_amd import gpu
    from nscale_amd.synth import synth_he

    sizes = serve_sizes(16, 2026)
    tiles = [torch.from_numpy(synth_he(h, w, 3000 + k)).cuda() for k, (h, w) in enumerate(sizes)]
    mx = max(max(h, w) for h, w in sizes)
    srv = gpu.Server(slots=args.slots, hbm_budget=args.budget_gb << 30, max_w=4096, max_h=max(8192, mx))
    expect = []
    for t in tiles:  # plain-path reference results (and warm-up)
        _, lab, _, _, _ = gpu.segment_nuclei(t)
        expect.append(checksum(lab))
    torch.cuda.synchronize()
    for tk in [srv.submit(tiles[k % len(tiles)]) for k in range(args.slots)]:
        srv.wait(tk)
    rng = np.random.default_rng(7)
    levels, mismatches = [], 0
    for rate in [float(r) for r in args.rates.split(",")]:
        n = args.requests
        gaps = rng.exponential(1.0 / rate, n)
        pick = rng.integers(0, len(tiles), n)
        tickets = [None] * n
        results = [None] * n

        def waiter():
            for i in range(n):
                while tickets[i] is None:
                    time.sleep(0.0002)
                results[i] = srv.wait(tickets[i])

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        due = t0
        for i in range(n):
            due += gaps[i]
            d = due - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[i] = srv.submit(tiles[pick[i]])
        th.join()
        wall = time.perf_counter() - t0
        px = sum(tiles[k].shape[0] * tiles[k].shape[1] for k in pick)
        lat = [r[5].total_ms for r in results]
        for i, r in enumerate(results):
            if checksum(r[2]) != expect[pick[i]]:
                mismatches += 1
        levels.append({"rate": rate, "requests": n, "mpix_s": round(px / wall / 1e6, 2),
                       "p50_ms": round(pct(lat, 50), 2), "p99_ms": round(pct(lat, 99), 2)})
        print(json.dumps(levels[-1]), file=sys.stderr, flush=True)
    srv.close()
    best = max(levels, key=lambda l: l["mpix_s"])
    print(json.dumps({"metric": "served segmentNuclei Mpix/s + p50/p99 request latency (open-loop Poisson)",
                      "value": best["mpix_s"], "unit": "Mpix/s", "n_gpus": 1, "higher_is_better": True,
                      "dtype": "u8", "data": "synthetic", "bit_exact_vs_plain": mismatches == 0,
                      "config": {"workload": "served chain, tile sides log-uniform [512,4096] x aspect U[1,2]",
                                 "slots": args.slots, "budget_gb": args.budget_gb, "distinct_tiles": len(tiles),
                                 "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))},
                      "levels": levels}), flush=True)
    if mismatches:
        sys.exit(1)


def bench_batched(args):
    """Plain vs batched reconstruction on n mask/marker pairs of uneven sizes (short side
    log-uniform in [512, 4096], long side x U[1, 2]), connectivity 4 or 8 at random: uint8
    imreconstruct, then float32 h-minima (h = 1) on the negated distance maps of the thresholded
    masks.  Pairs are crops of two 4096x8192 ellipse masks (the mask generator is the slow part)."""
    import numpy as np
    import torch
    from nscale_amd import gpu
    from nscale_amd.synth import synth_mask

    rng = np.random.default_rng(2026)
    bases = [torch.from_numpy(synth_mask(4096, 8192, 2026 + k)).cuda() for k in range(2)]
    sizes = serve_sizes(args.pairs, 2026)
    pairs, conns = [], []
    for k, (h, w) in enumerate(sizes):
        b = bases[k % 2] if h <= w else bases[k % 2].t()
        y0 = int(rng.integers(0, b.shape[0] - h + 1))
        x0 = int(rng.integers(0, b.shape[1] - w + 1))
        mask = b[y0:y0 + h, x0:x0 + w].contiguous()
        marker = (mask.to(torch.int16) - 10).clamp_(min=0).to(torch.uint8)
        pairs.append((marker, mask))
        conns.append(int(rng.choice([4, 8])))
    del bases
    px = sum(int(m.numel()) for _, m in pairs)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, r

    reps = max(1, args.steps)
    tuned = gpu.autotune_recon(pairs, conns)  # schedule for this batch (used by both paths below)
    t_plain, outs_p = timed(lambda: [gpu.imreconstruct(mk, m, c) for (mk, m), c in zip(pairs, conns)], reps)
    t_batch, outs_b = timed(lambda: gpu.imreconstruct_batched(pairs, conns), reps)
    exact_u8 = all(torch.equal(a, b) for a, b in zip(outs_p, outs_b))
    # the same batch as a replayed hipGraph (16 window passes per replay)
    graph = gpu.ReconGraph(pairs, conns, passes=16)
    t_graph, g_passes = timed(graph.run, reps)
    exact_u8 = exact_u8 and all(torch.equal(a, b) for a, b in zip(graph.outs, outs_b))
    graph.close()
    del outs_p, outs_b, graph
    # float32 h-minima on -EDT(mask < 150)
    imgs = []
    for _, m in pairs:
        bw = (m < 150).to(torch.uint8) * 255
        imgs.append(-gpu.dist_transform(bw))
    del pairs
    t_hplain, hp = timed(lambda: [gpu.imhmin(im, 1.0, c) for im, c in zip(imgs, conns)], reps)
    t_hbatch, hb = timed(lambda: gpu.imhmin_batched(imgs, 1.0, conns), reps)
    exact_f32 = all(torch.equal(a, b) for a, b in zip(hp, hb))
    achieved = 3.0 * px / t_batch / 1e9
    out = {"metric": "batched vs plain imreconstruct Mpix/s (bit-exact)", "value": round(px / t_batch / 1e6, 1),
           "unit": "Mpix/s", "n_gpus": 1, "steps": reps, "higher_is_better": True, "dtype": "u8", "data": "synthetic",
           "config": {"workload": "%d mask/marker pairs, sides log-uniform [512,4096] x aspect U[1,2], conn 4/8" % len(sizes),
                      "megapixels": round(px / 1e6, 1)},
           "plain_mpix_s": round(px / t_plain / 1e6, 1), "batched_vs_plain": round(t_plain / t_batch, 2),
           "autotuned": tuned,
           "graph": {"mpix_s": round(px / t_graph / 1e6, 1), "passes": g_passes, "passes_per_replay": 16},
           "bit_exact": exact_u8 and exact_f32,
           "hminima_f32": {"batched_mpix_s": round(px / t_hbatch / 1e6, 1), "plain_mpix_s": round(px / t_hplain / 1e6, 1),
                           "batched_vs_plain": round(t_hplain / t_hbatch, 2)},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                        "kernel": "batched imreconstruct<uchar>, 3 B/px algorithmic"}}
    print(json.dumps(out), flush=True)
    if not (exact_u8 and exact_f32):
        sys.exit(1)


def bench_slide(args):
    """Whole-slide scale: a 100000 x 80000 procedural H&E slide (stain intensity +-15% over a
    2000-px gradient-noise field) cut into 4096 x 4096 tiles with 64-px overlap (stride 4032:
    25 x 20 = 500 tiles), each tile rendered on the device and streamed through the server, which
    returns canonical int32 labels plus per-label area and bounding box; each tile's nuclei are then
    stitched into one slide-wide set (a nucleus is kept by the tile whose core holds its centre)."""
    import torch
    from nscale_amd import gpu
    from nscale_amd.slide import Stitcher

    SW, SH, T, OV = 100000, 80000, 4096, 64
    stride = T - OV
    xs = list(range(0, max(1, SW - OV), stride))
    ys = list(range(0, max(1, SH - OV), stride))
    tiles = [(x, y, min(T, SW - x), min(T, SH - y)) for y in ys for x in xs]
    if args.slide_tiles:
        tiles = tiles[:args.slide_tiles]
    srv = gpu.Server(slots=args.slots, hbm_budget=args.budget_gb << 30, max_w=T, max_h=T)
    ring = [torch.empty(T, T, 3, dtype=torch.uint8, device="cuda") for _ in range(2 * args.slots)]
    # warm-up: one tile per slot
    for tk in [srv.submit(gpu.synth_slide_tile(0, 0, T, T, 2026, ring[k])) for k in range(args.slots)]:
        srv.wait(tk)
    torch.cuda.synchronize()
    inflight = []
    nuclei, px, errors, lat = 0, 0, 0, []
    ws0 = gpu.watershed_stats()
    t0 = time.perf_counter()
    runs = []

    stitch = Stitcher(SW, SH, OV)

    def collect(entry):
        nonlocal nuclei, errors
        tk, xy = entry
        st, n, lab, bbox, area, res = srv.wait(tk)
        nuclei += n
        lat.append(res.total_ms)
        runs.append((round(res.run_ms, 1), round(res.queue_ms, 1), xy))
        errors += res.error != 0
        if st == 0:
            stitch.add(xy[0], xy[1], lab, bbox, area)

    for i, (x, y, w, h) in enumerate(tiles):
        if len(inflight) == len(ring):
            collect(inflight.pop(0))
        buf = ring[i % len(ring)].view(-1)[:h * w * 3].view(h, w, 3)
        gpu.synth_slide_tile(x, y, w, h, 2026, buf)
        inflight.append((srv.submit(buf, canonical=True, max_labels=65536), (x, y)))
        px += w * h
    for e in inflight:
        collect(e)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    srv.close()
    out = {"metric": "whole-slide segmented Mpix/s (tiles incl. overlap) + p50/p99 tile latency",
           "value": round(px / wall / 1e6, 1), "unit": "Mpix/s", "n_gpus": 1, "higher_is_better": True, "dtype": "u8",
           "data": "synthetic (procedural slide)",
           "config": {"workload": "100000x80000 slide, 4096 tiles, 64-px overlap, served with canonical labels, area, "
                                  "bbox, stitched", "tiles": len(tiles), "slots": args.slots,
                                  "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))},
           "seconds": round(wall, 2), "nuclei_in_tiles": nuclei, "nuclei_stitched": stitch.next_id - 1,
           "stitched_nuclei_touching_a_seam_edge": stitch.truncated, "errors": errors,
           "watershed_sequential_fallbacks": gpu.watershed_stats()[1] - ws0[1],
           "p50_ms": round(pct(lat, 50), 1), "p99_ms": round(pct(lat, 99), 1),
           "slowest_runs_ms_queue_ms_xy": sorted(runs, reverse=True)[:4]}
    print(json.dumps(out), flush=True)
    if errors:
        sys.exit(1)


def bench_normalize(args):
    """Normalization::normalization (Reinhard colour transfer) of the 4096x4096 synthetic H&E tile
    to the target statistics of segment/test/src/nu-segmentTest.cpp:185-187, input resident in HBM,
    output to a second device buffer.  Algorithmic bytes: 3 B/px read + 3 B/px written."""
    import torch
    from nscale_amd import gpu
    from nscale_amd.synth import synth_he

    S = args.size
    mean_t, std_t = [-0.632356, -0.0516004, 0.0376543], [0.26235, 0.0514831, 0.0114217]
    bgr = torch.from_numpy(synth_he(S, S, 2026)).cuda()
    out = torch.empty_like(bgr)
    for _ in range(args.warmup):
        gpu.normalization(bgr, mean_t, std_t, out=out)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record()
        gpu.normalization(bgr, mean_t, std_t, out=out)
        b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = sorted(a.elapsed_time(b) for a, b in evs)
    avg = sum(ms) / len(ms)
    achieved = 6 * S * S / (avg * 1e-3) / 1e9
    print(json.dumps({"metric": "normalization Mpix/s", "value": round(args.steps * S * S / wall / 1e6, 3),
                      "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": round(wall * 1e3 / args.steps, 4), "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
                      "config": {"workload": "Normalization::normalization %dx%d H&E tile" % (S, S)},
                      "p50_ms": round(statistics.median(ms), 4),
                      "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                                   "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                                   "avg_ms": round(avg, 4)}}), flush=True)


def bench_copy(args):
    """Calibration leg for HBM byte counters: a plain device copy of a known size (64 MiB)."""
    import torch
    n = 64 << 20
    a = torch.randint(0, 255, (n,), dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(args.warmup):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    print(json.dumps({"metric": "device copy GB/s", "value": round(2 * n * args.steps / wall / 1e9, 1), "unit": "GB/s",
                      "bytes_read_per_call": n, "bytes_written_per_call": n, "steps": args.steps}), flush=True)


def served_chain(args, S, bgr, params, plain_ck, slat):
    """B tiles per step through the server; returns (Mpix/s, wall s); mismatches vs the plain path
    are left in served_chain.mismatches."""
    import torch
    from nscale_amd import gpu
    from nscale_amd.synth import synth_he
    B = args.batch
    tiles = [bgr] + [torch.from_numpy(synth_he(S, S, 2026 + k)).cuda() for k in range(1, min(B, args.distinct))]
    srv = gpu.Server(slots=args.slots, hbm_budget=args.budget_gb << 30, max_w=S, max_h=S)
    expect = {}
    for k, t in enumerate(tiles):  # warm every slot; reference checksums from the plain path
        if k == 0:
            expect[0] = plain_ck
        else:
            st_k, lab_k, _, _, _ = gpu.segment_nuclei(t, params=params)
            expect[k] = checksum(lab_k)
    for tk in [srv.submit(tiles[k % len(tiles)]) for k in range(max(B, args.slots))]:
        srv.wait(tk)
    torch.cuda.synchronize()
    mismatches = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tks = [(k % len(tiles), srv.submit(tiles[k % len(tiles)])) for k in range(B)]
        for k, tk in tks:
            st_s, n_s, lab_s, _, _, res = srv.wait(tk)
            slat.append(res.total_ms)
            if checksum(lab_s) != expect[k]:
                mismatches += 1
    torch.cuda.synchronize()
    swall = time.perf_counter() - t0
    srv.close()
    served_chain.mismatches = mismatches
    return args.steps * B * S * S / swall / 1e6, swall


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--cpu-size", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("chain", "recon", "copy", "serve", "batched", "slide", "normalize"),
                    default="chain")
    ap.add_argument("--slide-tiles", type=int, default=0, help="slide: process only the first N tiles (0 = all)")
    ap.add_argument("--pairs", type=int, default=256, help="batched: mask/marker pairs")
    ap.add_argument("--batch", type=int, default=16, help="tiles per served step")
    ap.add_argument("--distinct", type=int, default=8, help="distinct synthetic tiles in a served batch")
    ap.add_argument("--slots", type=int, default=8, help="server slots (concurrent chains)")
    ap.add_argument("--budget-gb", type=int, default=200, help="server HBM budget")
    ap.add_argument("--rates", default="10,50,100,200,500,1000,2000", help="serve: arrival rates, req/s")
    ap.add_argument("--requests", type=int, default=500, help="serve: requests per rate")
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="GPU_MAX_HW_QUEUES for this process (0 = keep the environment's value)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the segmentation path has no multi-GPU component in this build "
              "(tiles shard across GPUs with no exchange); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from nscale_amd import gpu
    from nscale_amd.synth import synth_he

    if args.workload == "recon":
        return bench_recon(args)
    if args.workload == "copy":
        return bench_copy(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "slide":
        return bench_slide(args)
    if args.workload == "normalize":
        return bench_normalize(args)
    S = args.size
    bgr = torch.from_numpy(synth_he(S, S, 2026)).cuda()
    ws = gpu.Workspace(S, S)
    params = gpu.default_params()
    for _ in range(args.warmup):
        st, lab, cc, _, _ = gpu.segment_nuclei(bgr, params=params, ws=ws)
    torch.cuda.synchronize()

    # ---- plain path: one tile per chain call, one stream ----
    lat, recon_ms = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        st, lab, cc, _, ms = gpu.segment_nuclei(bgr, params=params, ws=ws, timing=True)
        lat.append((time.perf_counter() - t) * 1e3)  # the call returns after its final sync
        recon_ms.append(ms.get(26, float("nan")))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    plain_mpix = args.steps * S * S / wall / 1e6
    plain_ck = checksum(lab)
    lat.sort()
    rt = sum(recon_ms) / len(recon_ms)
    alg_bytes = 3 * S * S
    achieved = alg_bytes / (rt * 1e-3) / 1e9
    traffic = measured_traffic("recon_in_chain") if S == 4096 else None
    del ws

    # ---- served path: a batch of B tiles per step through the continuous-batching server ----
    B = args.batch
    slat, mismatches = [], 0
    if B > 0:
        served_mpix, swall = served_chain(args, S, bgr, params, plain_ck, slat)
        mismatches = served_chain.mismatches
    else:  # plain path only (profiling)
        served_mpix, swall = plain_mpix, wall
    out = {
        "metric": "segmented Mpix/s (batched vs plain, bit-exact) + p50/p99 tile latency",
        "value": round(served_mpix, 3), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(swall * 1e3 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": ("segmentNuclei chain on %dx%d BGR uint8 synthetic H&E tiles; a step = %d tiles "
                                "through the continuous-batching server (%d slots)" % (S, S, B, args.slots)) if B else
                               "segmentNuclei chain, one %dx%d BGR uint8 synthetic H&E tile per call (plain path)" % (S, S),
                   "tile": [S, S], "batch": B, "slots": args.slots, "status": st, "nuclei": cc,
                   "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))},
        "plain_mpix_s": round(plain_mpix, 3), "served_vs_plain": round(served_mpix / plain_mpix, 3),
        "bit_exact_served_vs_plain": mismatches == 0,
        "p50_ms": round(pct(slat or lat, 50), 3), "p99_ms": round(pct(slat or lat, 99), 3),
        "plain_p50_ms": round(pct(lat, 50), 3), "plain_p99_ms": round(pct(lat, 99), 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "imreconstruct<uchar>(rc_open, rc, 8) inside the chain, 3 B/px algorithmic",
                     "avg_ms": round(rt, 4)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_size)
    print(json.dumps(out), flush=True)
    if mismatches:
        print("served results differ from the plain path on %d requests" % mismatches, file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
