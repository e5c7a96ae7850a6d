"""Benchmark of the per-grain tracking hot path on one MI355X.

A step is one frame pair of the 6000x4000 12-bit workload with 2,000 disk grains
(lognormal radius, median 20 px, clipped to 8-40 px): precomputation, integer ZNCC
search +-20 px x 21 rotation steps, rescue +-40 px for grains with NCC < 0.7, subpixel
refinement, and the 9 result columns copied back to the host. Frames are resident in
HBM before timing; each step re-uploads the grains' input records (144 B each) as the
pair's input. Prints one JSON line.

The timed steps run without stage timers; the stage split comes from as many steps again with
HIP events around each stage. Then the same pairs run with every kernel of the exact search on
one stream, so that the search kernel's launch durations (HIP events) do not overlap: the
roofline is computed from those. Then the pair captured as one HIP graph and replayed; frames
from host memory (a 6000x4000 frame copied per pair from pageable memory before the pair, or from
pinned memory while the pair runs: the PCIe-inclusive rate); the plain
form, the exact-vs-plain pick comparison on configs 2 and
3, the rescue-zone variant of the pair (300 grains moved 21-35 px, beyond the main zone: the
rescue stage at its own +-40 px geometry) and the oracle's CPU baseline on a sample of grains
(one warm-up, median of three).

--sequence runs the batching form instead: a 16-frame 6000x4000 12-bit sequence with
10,000 disk grains of radius U(6, 60) px moved per frame by U(-5, 5) px and U(-0.03, 0.03)
rad (accumulated), tracked pair after pair with the reference frame swapped in between
(search +-8 px x 9 steps, subpixel). All frames are resident in HBM; a pair's latency runs
from the reference swap to the 9 result columns on the host. It prints one JSON line with
grains/s and the p50/p99 pair latency (pairs of launches, and each pair replaying one HIP
graph), and checks a sample of grains of every pair against the oracle.
"""
import argparse
import json
import os
import sys
import time
import types

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

METRIC = "grains/s per frame pair, adjoint-LM+MFMA vs plain search+Powell, ≤1e-3 px"
FP64_PEAK_TFLOPS = 78.6
INT8_PEAK_TOPS = 5000.0
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


TRAFFIC_FILE = "profiles/exact_lm_cfg2_6000x4000_2000grains_hbm_bytes.txt"


def hbm_bytes_per_launch(kernel):
    """HBM bytes per launch of `kernel` from the byte-counter summary, or None."""
    try:
        for line in open(os.path.join(ROOT, TRAFFIC_FILE)):
            if line.startswith("#") or kernel not in line:
                continue
            return float(line.split()[-1]) * 1e6
    except OSError:
        pass
    return None


def pick_mismatches(pk, tr, w, load=False):
    """Main-stage integer picks (dx, dy, drot) of the exact and plain searches on all grains
    of workload w; returns the number of grains whose picks differ in any bit."""
    if load:
        tr.set_image(0, w.I0, w.bits)
        tr.set_image(1, w.I1, w.bits)
        tr.set_image_indices(0, 1)
        tr.set_grains(w.grains, w.pattern_xy, w.pattern_begin)
    z = w.params.search_zone
    res = {}
    for m in (pk.SEARCH_EXACT, pk.SEARCH_PLAIN):
        tr.put_grains(w.grains)
        tr.track_pair(pk.track_params(z, rescue_level=0, subpixel=0, search_method=m))
        res[m] = tr.get_grains()
    a, b = res[pk.SEARCH_EXACT], res[pk.SEARCH_PLAIN]
    diff = np.zeros(len(a), dtype=bool)
    for c in ("dx", "dy", "drot"):
        diff |= a[c].view(np.uint64) != b[c].view(np.uint64)
    steps = len(ptref_steps(z))
    return {"grains": int(len(a)), "rotation_steps": steps, "mismatches": int(diff.sum()),
            "max_ncc_diff": float(np.abs(a["NCC"] - b["NCC"]).max())}


def ptref_steps(z):
    import ptref
    return ptref.rotation_steps(z.inc_rot, z.num_rot)


def copy_probe(nbytes):
    """One device-to-device copy of nbytes: a kernel of known HBM traffic (nbytes read,
    nbytes written) whose FETCH_SIZE / WRITE_SIZE counters scale the search kernel's."""
    import torch

    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda:0").fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes}), flush=True)


def run_sequence(args, pk, synth, ptref):
    import torch

    t0 = time.time()
    S = synth.sequence(n_frames=args.seq_frames, n_grains=args.seq_grains)
    H, W = S.frames[0].shape
    log(f"sequence generated in {time.time() - t0:.1f}s: {S.G} grains, {len(S.pattern_xy)} pattern px, "
        f"{len(S.frames)} frames")
    dev = [torch.from_numpy(f).to("cuda:0") for f in S.frames]  # every frame resident in HBM
    torch.cuda.synchronize()
    z = S.params.search_zone
    prm = pk.track_params(z, z, z, NCC_min=0.0, rescue_level=0, subpixel=1, search_method=pk.SEARCH_EXACT,
                          subpix=pk.default_subpix(pk.SUBPIX_LM))
    tr = pk.Tracker(0)
    if args.serial_search:
        tr.set_option(pk.OPT_SEARCH_STREAMS, 1)
    # radius buckets launched by decreasing work: this grain set's work grows with the radius
    tr.set_option(pk.OPT_BUCKET_ORDER, args.seq_bucket_order)
    npairs = len(S.frames) - 1

    def attach(slot, k):
        tr.set_image_device(slot, dev[k].data_ptr(), W, H, W, S.bits)

    def one_pass(lat, keep):
        attach(0, 0)
        attach(1, 1)
        tr.set_image_indices(0, 1)
        tr.set_grains(S.grains, S.pattern_xy, S.pattern_begin)
        tr.sync()
        outs = []
        for k in range(npairs):
            t = time.perf_counter()
            if k > 0:
                tr.change_reference()
                attach((k + 1) % 2, k + 1)
            tr.track_pair(prm)
            out = tr.get_grains()
            lat.append(time.perf_counter() - t)
            if keep:
                outs.append(out)
        return outs

    one_pass([], False)  # warm-up: exact-search plans, code objects
    lat = []
    outs = None
    f0 = tr.exact_fallbacks()
    t1 = time.perf_counter()
    for r in range(args.seq_repeats):
        o = one_pass(lat, r == 0)
        outs = outs or o
    el = time.perf_counter() - t1
    f1 = tr.exact_fallbacks()
    # the stage split from one more pass, with HIP events around each stage
    tr.set_timing(True)
    tr.reset_stage_times()
    one_pass([], False)
    ms, launches = tr.stage_times()
    tr.set_timing(False)
    stage_names = ["precompute", "search_main", "search_rescue", "search_super", "subpixel", "remap", "search_kernel"]
    stage_ms = {n: ms[i] / npairs for i, n in enumerate(stage_names) if launches[i]}
    log("stage ms/pair: " + ", ".join(f"{k} {v:.2f}" for k, v in stage_ms.items()) +
        f"; grains handed to the plain kernel: {(f1[1] - f0[1]) / len(lat):.0f} per pair")
    lat_ms = 1e3 * np.array(lat)
    value = S.G * len(lat) / lat_ms.sum() * 1e3
    log(f"sequence: {len(lat)} pairs, {value:.0f} grains/s, pair latency p50 {np.percentile(lat_ms, 50):.2f} ms "
        f"p99 {np.percentile(lat_ms, 99):.2f} ms ({el:.1f}s with the set-up of each pass)")
    # the same passes with each pair replaying one HIP graph (captured after two passes' pairs
    # of launches settle the buffers), the records checked against the pairs of launches
    graph = None
    if args.graph_steps > 0:
        tr.set_option(pk.OPT_GRAPH, 1)
        one_pass([], False)
        glat = []
        gouts = one_pass(glat, True)
        for r in range(1, args.seq_repeats):
            one_pass(glat, False)
        captures, glaunches, nodes = tr.graph_stats()
        tr.set_option(pk.OPT_GRAPH, 0)
        glat_ms = 1e3 * np.array(glat)
        same = all(a.tobytes() == b.tobytes() for a, b in zip(gouts, outs))
        graph = {"grains_per_s": S.G * len(glat) / glat_ms.sum() * 1e3,
                 "latency_ms": {"p50": float(np.percentile(glat_ms, 50)), "p99": float(np.percentile(glat_ms, 99))},
                 "nodes_per_pair": nodes, "captures": captures, "graph_launches": glaunches,
                 "records_equal_launches": same}
        log(f"sequence, graph replay: p50 {graph['latency_ms']['p50']:.2f} ms p99 {graph['latency_ms']['p99']:.2f} ms, "
            f"{nodes} nodes per pair, {captures} captures, records equal to launches: {same}")
    tr.close()

    # oracle check on a sample of every pair, from the GPU's own input records of that pair
    check = {}
    if args.seq_check_grains > 0:
        rng = np.random.default_rng(0)
        sample = np.sort(rng.choice(S.G, size=min(args.seq_check_grains, S.G), replace=False))
        tp = pk.track_params(z, z, z, NCC_min=0.0, rescue_level=0, subpixel=1, search_method=pk.SEARCH_PLAIN,
                             subpix=pk.default_subpix(pk.SUBPIX_LM, subpix_tol=1e-13))
        gin = S.grains.copy()
        epx, erad, bad_int = [], [], 0
        t2 = time.perf_counter()
        for k in range(npairs):
            Wk = ptref.Workload(S.frames[k], S.frames[k + 1], gin, S.pattern_xy, S.pattern_begin, S.bits)
            ot, _, _ = ptref.track_pair(Wk, tp, nthreads=min(16, os.cpu_count() or 1), subset=sample)
            g = outs[k][sample]
            o = ot  # the oracle returns the sample's rows
            epx.append(np.hypot(g["dx"] - o["dx"], g["dy"] - o["dy"]))
            erad.append(np.abs(g["drot"] - o["drot"]))
            # an integer pick that differs moves the result by about a pixel or a step
            bad_int += int(((np.abs(g["dx"] - o["dx"]) > 0.5) | (np.abs(g["dy"] - o["dy"]) > 0.5)).sum())
            gin = ptref.change_reference(outs[k])
        epx, erad = np.concatenate(epx), np.concatenate(erad)
        check = {"grains_per_pair": int(len(sample)), "pairs": npairs, "max_px": float(epx.max()),
                 "max_rad": float(erad.max()), "over_1e-3_px": int((epx > 1e-3).sum()),
                 "over_1e-5_rad": int((erad > 1e-5).sum()), "over_half_px": bad_int,
                 "seconds": time.perf_counter() - t2}
        log(f"oracle check: {len(sample)} grains x {npairs} pairs, max {epx.max():.2e} px {erad.max():.2e} rad, "
            f"{(epx > 1e-3).sum()} over 1e-3 px")
    P = np.diff(S.pattern_begin).astype(np.float64)
    R = len(ptref.rotation_steps(z.inc_rot, z.num_rot))
    Sh = (z.left + z.right + 1) * (z.up + z.down + 1)
    result = {
        "metric": "grains/s over a frame sequence (batching form)", "value": value, "unit": "grains/s",
        "n_gpus": 1, "steps": len(lat), "warmup": npairs, "ms_per_step": float(lat_ms.mean()),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "i8", "data": "synthetic",
        "config": {"workload": f"{len(S.frames)}-frame 6000x4000 12-bit sequence, {S.G} disk grains r ~ U(6, 60) "
                               "px (overlapping), per-frame motion U(-5, 5) px, U(-0.03, 0.03) rad accumulated, "
                               "consecutive pairs with reference swap, search +-8 px x 9 steps, subpixel LM",
                   "grains": int(S.G), "pattern_px": int(P.sum()), "rotation_steps": R, "shifts": Sh,
                   "pairs_timed": len(lat), "bucket_order": args.seq_bucket_order},
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99)),
                       "max": float(lat_ms.max()), "min": float(lat_ms.min()),
                       "per_pair_first_pass": [round(float(x), 2) for x in lat_ms[:npairs]]},
        "stage_ms": stage_ms,
        "graph": graph,
        "plain_kernel_grains_per_pair": (f1[1] - f0[1]) / len(lat),
        "algorithmic_search_flops_per_pair": 2.0 * R * Sh * P.sum(),
        "oracle_check": check,
    }
    print(json.dumps(result), flush=True)


def run_remap(args):
    """The distortion-correction remap of a 6000x4000 12-bit frame through the config-4 lens
    (k1 = -2e-9, k2 = 1e-16 about the centre), frame resident in HBM: one JSON line with the
    kernel's time per frame and its HBM rate against the 4 bytes per pixel it must move."""
    import ptref

    pk = ptref.package()
    from ptracker_amd import synth

    w = synth.workload(2, n_grains=200)
    H, W = w.I0.shape
    X = np.array([0.5 * W, 0.5 * H, -2e-9, 1e-16, 0, 0, 0, 0, 0.5 * W, 0.5 * H, 0, 0, 0, 0], dtype=np.float64)
    tr = pk.Tracker(0)
    tr.set_image(2, w.I0, w.bits)
    for _ in range(args.warmup):
        tr.remap(2, 3, X)
    tr.sync()
    tr.set_timing(True)
    tr.reset_stage_times()
    t = time.perf_counter()
    for _ in range(args.steps):
        tr.remap(2, 3, X)
    tr.sync()
    el = time.perf_counter() - t
    ms, _ = tr.stage_times()
    tr.close()
    k_ms = ms[pk.T_REMAP] / args.steps
    nbytes = 4.0 * W * H
    achieved = nbytes / (k_ms * 1e-3) / 1e9
    print(json.dumps({"metric": "remap frames/s", "value": args.steps / el, "unit": "frames/s", "frame": f"{W}x{H}",
                      "kernel_ms": k_ms, "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                                                      "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                                                      "algorithmic_bytes": nbytes}}), flush=True)


def rescue_variant(pk, synth, args):
    """The pair with 300 of its 2,000 grains moved 21-35 px along one axis (beyond the +-20 px
    main zone, inside the +-40 px rescue zone): the fast form's rescue stage at its own
    geometry, timed, and its picks against the plain form's."""
    w = synth.workload(2, n_grains=args.grains, far=(max(1, 300 * args.grains // 2000), 21.0, 35.0))
    tr = pk.Tracker(0)
    tr.set_image(0, w.I0, w.bits)
    tr.set_image(1, w.I1, w.bits)
    tr.set_image_indices(0, 1)
    tr.set_grains(w.grains, w.pattern_xy, w.pattern_begin)
    z = w.params
    out = {}
    for name, sm, sp in (("exact_lm", pk.SEARCH_EXACT, pk.SUBPIX_LM), ("plain_powell", pk.SEARCH_PLAIN, pk.SUBPIX_POWELL)):
        prm = pk.track_params(z.search_zone, z.search_zone_rescue, z.search_zone_super_rescue, NCC_min=0.7,
                              rescue_level=1, subpixel=1, search_method=sm, subpix=pk.default_subpix(sp))
        steps = 3 if sm == pk.SEARCH_EXACT else 1
        tr.put_grains(w.grains)
        tr.track_pair(prm)
        tr.sync()
        f0 = tr.exact_fallbacks()
        tr.set_timing(True)
        tr.reset_stage_times()
        t = time.perf_counter()
        for _ in range(steps):
            tr.put_grains(w.grains)
            tr.track_pair(prm)
            g = tr.get_grains()
        el = time.perf_counter() - t
        ms, _ = tr.stage_times()
        tr.set_timing(False)
        nres, _ = tr.queue_counts()
        f1 = tr.exact_fallbacks()
        out[name] = {"ms_per_pair": 1e3 * el / steps, "grains_per_s": w.G * steps / el, "rescued": int(nres),
                     "stage_ms": {"search_main": ms[pk.T_SEARCH_MAIN] / steps,
                                  "search_rescue": ms[pk.T_SEARCH_RESCUE] / steps,
                                  "subpixel": ms[pk.T_SUBPIXEL] / steps},
                     "exact_fallbacks": {"searches": int(f1[0] - f0[0]), "grains": int(f1[1] - f0[1])},
                     "_g": g}
    a, b = out["exact_lm"].pop("_g"), out["plain_powell"].pop("_g")
    tr.close()
    # integer picks after the rescue: the subpixel stage moves them, so compare the searches alone
    tr = pk.Tracker(0)
    tr.set_image(0, w.I0, w.bits)
    tr.set_image(1, w.I1, w.bits)
    tr.set_image_indices(0, 1)
    tr.set_grains(w.grains, w.pattern_xy, w.pattern_begin)
    picks = {}
    for sm in (pk.SEARCH_EXACT, pk.SEARCH_PLAIN):
        tr.put_grains(w.grains)
        tr.track_pair(pk.track_params(z.search_zone, z.search_zone_rescue, z.search_zone_super_rescue, NCC_min=0.7,
                                      rescue_level=1, subpixel=0, search_method=sm))
        picks[sm] = tr.get_grains()
    tr.close()
    diff = np.zeros(w.G, dtype=bool)
    for c in ("dx", "dy", "drot"):
        diff |= picks[pk.SEARCH_EXACT][c].view(np.uint64) != picks[pk.SEARCH_PLAIN][c].view(np.uint64)
    out["pick_mismatches_exact_vs_plain"] = int(diff.sum())
    out["grains"] = int(w.G)
    out["moved_beyond_main_zone"] = int(len(w.truth["far"]))
    log(f"rescue variant: exact+lm {out['exact_lm']['ms_per_pair']:.2f} ms/pair (rescue stage "
        f"{out['exact_lm']['stage_ms']['search_rescue']:.2f} ms, {out['exact_lm']['rescued']} grains), plain+powell "
        f"{out['plain_powell']['ms_per_pair']:.2f} ms/pair (rescue {out['plain_powell']['stage_ms']['search_rescue']:.2f}"
        f" ms); pick mismatches {out['pick_mismatches_exact_vs_plain']}")
    return out


def graph_section(pk, tr, w, prm, step, ref, args):
    """The same pairs with the pair captured once as a HIP graph and replayed (PTG_OPT_GRAPH):
    ms/pair, the graph's nodes (the pair's launches and copies) and whether the records equal
    those of the pairs run as launches."""
    tr.set_option(pk.OPT_GRAPH, 1)
    for _ in range(3):  # launches twice (new buffers change the key once), then the capture
        step(prm)
    tr.sync()
    t0 = time.perf_counter()
    for _ in range(args.graph_steps):
        got = step(prm)
    tr.sync()
    el = time.perf_counter() - t0
    captures, launches, nodes = tr.graph_stats()
    tr.set_option(pk.OPT_GRAPH, 0)
    same = got.tobytes() == ref.tobytes()
    ms = 1e3 * el / args.graph_steps
    log(f"graph replay: {ms:.2f} ms/pair, {nodes} nodes per pair, records equal to launches: {same}")
    return {"ms_per_pair": ms, "grains_per_s": w.G * args.graph_steps / el, "steps": args.graph_steps,
            "nodes_per_pair": nodes, "captures": captures, "graph_launches": launches,
            "records_equal_launches": same}


def ingest_section(pk, tr, w, prm, step, ref, args):
    """Frames from host memory: one 6000x4000 uint16 frame copied into a slot the pair does
    not use, (a) from pageable memory with the synchronous ptg_set_image, (b) from pinned
    memory with ptg_set_image_async (the copy alone, waited for), then the pair with a new
    frame per step, (c) copied synchronously before the pair and (d) copied while the pair
    runs (ptg_set_image_async queued before ptg_track_pair): the PCIe-inclusive rate of a
    sequence whose frames come from the host."""
    n = args.ingest_steps
    H, W = w.I1.shape
    nbytes = 2.0 * H * W
    page = np.ascontiguousarray(w.I1)
    pinned = [tr.host_frame(H, W), tr.host_frame(H, W)]
    for p in pinned:
        p[...] = w.I1
    tr.set_image(2, page, w.bits)
    tr.set_image(3, page, w.bits)
    tr.sync()

    def timed(body):
        t0 = time.perf_counter()
        for i in range(n):
            body(i)
        tr.sync()
        return (time.perf_counter() - t0) / n

    def sync_copy(i):
        tr.set_image(2 + (i & 1), page, w.bits)

    def async_copy(i):
        tr.set_image_async(2 + (i & 1), pinned[i & 1], w.bits)
        tr.wait_image(2 + (i & 1))

    def pair_sync_copy(i):
        tr.set_image(2 + (i & 1), page, w.bits)
        step(prm)

    got = []

    def pair_overlapped_copy(i):
        tr.set_image_async(2 + (i & 1), pinned[i & 1], w.bits)
        got.append(step(prm))

    t_page = timed(sync_copy)
    t_pin = timed(async_copy)
    t_ps = timed(pair_sync_copy)
    t_po = timed(pair_overlapped_copy)
    same = all(g.tobytes() == ref.tobytes() for g in got)
    log(f"frame copy: pageable {1e3 * t_page:.2f} ms ({nbytes / t_page / 1e9:.1f} GB/s), pinned "
        f"{1e3 * t_pin:.2f} ms ({nbytes / t_pin / 1e9:.1f} GB/s); pair + copy {1e3 * t_ps:.2f} ms, pair with the "
        f"copy overlapped {1e3 * t_po:.2f} ms; records equal: {same}")
    return {"frame_bytes": int(nbytes), "steps": n,
            "pageable_sync_copy_ms": 1e3 * t_page, "pageable_GBps": nbytes / t_page / 1e9,
            "pinned_async_copy_ms": 1e3 * t_pin, "pinned_GBps": nbytes / t_pin / 1e9,
            "pair_after_sync_copy_ms": 1e3 * t_ps, "pair_with_overlapped_copy_ms": 1e3 * t_po,
            "grains_per_s_frames_from_host": w.G / t_po, "records_equal": same}


def autotune_section(pk, tr, w, prm, step, args):
    """The exact search's autotuner on this pair: per-pair wall times (median, mean, largest)
    with the built-in launch shapes and with the tuned ones, in the same process, and the
    main-stage search span the tuner measured for both."""
    def times(n):
        for _ in range(2):
            step(prm)
        tr.sync()
        ts = []
        for _ in range(n):
            t = time.perf_counter()
            step(prm)
            ts.append(1e3 * (time.perf_counter() - t))
        tr.sync()
        ts = np.array(ts)
        return {"median": float(np.median(ts)), "mean": float(ts.mean()), "max": float(ts.max())}
    n = max(10, args.steps)
    tr.clear_tuning()
    heur = times(n)
    tr.put_grains(w.grains)  # tuned on the state every timed pair starts from
    t5 = time.perf_counter()
    n_t, ms_h, ms_t = tr.autotune_search(w.params.search_zone)
    sec = time.perf_counter() - t5
    tuned = times(n)
    tr.clear_tuning()
    out = {"searches_run": n_t, "seconds": sec, "search_span_ms_heuristic_shapes": ms_h,
           "search_span_ms_tuned_shapes": ms_t, "pair_ms_heuristic_shapes": heur, "pair_ms_tuned_shapes": tuned}
    log(f"autotune: {n_t} searches in {sec:.1f} s; search span {ms_h:.3f} -> {ms_t:.3f} ms; pair median "
        f"{heur['median']:.3f} -> {tuned['median']:.3f} ms, mean {heur['mean']:.3f} -> {tuned['mean']:.3f} ms")
    return out


def depth_variants(pk, w, args, flops):
    """The same pair requantised to 8 bits (v >> 4: one offset-INT8 digit plane) and spread
    over 16 bits (v * 16 + U{0..15}: two offset-byte planes, three template digits): the fast
    form's pair time, its main-stage search kernel on one stream (HIP events) against the
    int8 peak, and its integer picks against the plain form's."""
    out = {}
    rng = np.random.default_rng(16)
    for bits in (8, 16):
        if bits == 8:
            I0, I1 = (w.I0 >> 4).astype(np.uint16), (w.I1 >> 4).astype(np.uint16)
        else:
            I0 = (w.I0.astype(np.uint32) * 16 + rng.integers(0, 16, w.I0.shape)).astype(np.uint16)
            I1 = (w.I1.astype(np.uint32) * 16 + rng.integers(0, 16, w.I1.shape)).astype(np.uint16)
        wd = types.SimpleNamespace(I0=I0, I1=I1, bits=bits, grains=w.grains, pattern_xy=w.pattern_xy,
                                   pattern_begin=w.pattern_begin, params=w.params)
        tr = pk.Tracker(0)
        tr.set_image(0, I0, bits)
        tr.set_image(1, I1, bits)
        tr.set_image_indices(0, 1)
        tr.set_grains(w.grains, w.pattern_xy, w.pattern_begin)
        z = w.params
        prm = pk.track_params(z.search_zone, z.search_zone_rescue, z.search_zone_super_rescue, NCC_min=0.7,
                              rescue_level=1, subpixel=1, search_method=pk.SEARCH_EXACT,
                              subpix=pk.default_subpix(pk.SUBPIX_LM))
        steps = max(1, args.steps)
        f0 = tr.exact_fallbacks()
        for serial in (False, True):
            tr.set_option(pk.OPT_SEARCH_STREAMS, 1 if serial else 2)
            tr.put_grains(w.grains)
            tr.track_pair(prm)
            tr.sync()
            tr.set_timing(True)
            tr.reset_stage_times()
            t = time.perf_counter()
            for _ in range(steps):
                tr.put_grains(w.grains)
                tr.track_pair(prm)
                tr.get_grains()
            tr.sync()
            el = time.perf_counter() - t
            ms, _ = tr.stage_times()
            tr.set_timing(False)
            if not serial:
                rec = {"ms_per_pair": 1e3 * el / steps, "grains_per_s": w.G * steps / el,
                       "stage_ms": {"search_main": ms[pk.T_SEARCH_MAIN] / steps, "subpixel": ms[pk.T_SUBPIXEL] / steps}}
            else:
                k_ms = ms[pk.T_SEARCH_KERNEL] / steps
                ach = flops / (k_ms * 1e-3) / 1e12
                rec["search_kernel"] = {"ms_per_pair_one_stream": k_ms, "achieved_tops": ach,
                                        "frac_of_int8_peak": ach / INT8_PEAK_TOPS}
        f1 = tr.exact_fallbacks()
        rec["exact_fallbacks"] = {"searches": int(f1[0] - f0[0]), "grains": int(f1[1] - f0[1])}
        rec["picks_vs_plain"] = pick_mismatches(pk, tr, wd)
        tr.close()
        out[f"{bits}bit"] = rec
        log(f"{bits}-bit frames: exact+lm {rec['ms_per_pair']:.2f} ms/pair, search kernel "
            f"{rec['search_kernel']['ms_per_pair_one_stream']:.2f} ms ({rec['search_kernel']['achieved_tops']:.0f} TOPS), "
            f"pick mismatches vs plain {rec['picks_vs_plain']['mismatches']}, fallbacks {rec['exact_fallbacks']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--method", choices=["auto", "exact", "plain"], default="auto")
    ap.add_argument("--subpix", choices=["auto", "lm", "powell"], default="auto")
    ap.add_argument("--grains", type=int, default=2000)
    ap.add_argument("--plain-steps", type=int, default=2, help="steps of the plain search+Powell form (0: skip)")
    ap.add_argument("--cpu-grains", type=int, default=500, help="grains in the CPU-baseline sample (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (default: OMP_NUM_THREADS, else every CPU the process may use)")
    ap.add_argument("--no-rescue-variant", action="store_true",
                    help="skip the rescue-zone variant of the pair (300 grains moved beyond the main zone)")
    ap.add_argument("--autotune", action="store_true",
                    help="also autotune the exact search's launch shapes and time the tuned pairs")
    ap.add_argument("--no-depths", action="store_true",
                    help="skip the pair requantised to 8 bits and spread over 16 bits")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the exact-vs-plain pick comparison (configs 2 and 3) after the timed run")
    ap.add_argument("--graph-steps", type=int, default=10,
                    help="pairs timed with the pair captured as a HIP graph and replayed (0: skip)")
    ap.add_argument("--ingest-steps", type=int, default=10,
                    help="pairs timed with a frame copied from host memory per pair, synchronously and "
                         "overlapped with the pair (0: skip)")
    ap.add_argument("--sequence", action="store_true", help="the batching form over a frame sequence")
    ap.add_argument("--seq-frames", type=int, default=16)
    ap.add_argument("--seq-grains", type=int, default=10000)
    ap.add_argument("--seq-repeats", type=int, default=2, help="timed passes over the sequence")
    ap.add_argument("--seq-check-grains", type=int, default=48, help="grains per pair checked against the oracle")
    ap.add_argument("--seq-bucket-order", type=int, default=1, choices=[0, 1],
                    help="PTG_OPT_BUCKET_ORDER of the batching form (1: radius buckets by decreasing work)")
    ap.add_argument("--serial-search", action="store_true",
                    help="run every kernel of the exact search on one stream throughout (for profiler traces "
                         "and counters whose per-kernel figures must not overlap)")
    ap.add_argument("--remap", action="store_true",
                    help="only time the distortion-correction remap of a 6000x4000 12-bit frame (the lens of "
                         "config 4) and exit")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy this many bytes device to device and exit (calibrates rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    if args.remap:
        run_remap(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the path shards grains across GPUs "
              "without collectives in the hot loop, and this build runs on a single device.", file=sys.stderr)
        sys.exit(2)

    import ptref

    pk = ptref.package()
    from ptracker_amd import synth

    if args.sequence:
        run_sequence(args, pk, synth, ptref)
        return

    t0 = time.time()
    w = synth.workload(2, n_grains=args.grains)
    log(f"workload generated in {time.time() - t0:.1f}s: {w.G} grains, {len(w.pattern_xy)} pattern px")

    exact_ok = True
    method = args.method
    subpix = args.subpix
    tr = pk.Tracker(0)
    if args.serial_search:
        tr.set_option(pk.OPT_SEARCH_STREAMS, 1)
    tr.set_image(0, w.I0, w.bits)
    tr.set_image(1, w.I1, w.bits)
    tr.set_image_indices(0, 1)
    tr.set_grains(w.grains, w.pattern_xy, w.pattern_begin)
    if method == "auto":
        try:
            tr.search(pk.STAGE_MAIN, w.params.search_zone, 0.7, 0.5, pk.SEARCH_EXACT)
            method = "exact"
        except pk.PtgError:
            exact_ok = False
            method = "plain"
    if subpix == "auto":
        subpix = "lm" if method == "exact" else "powell"
    search_method = pk.SEARCH_EXACT if method == "exact" else pk.SEARCH_PLAIN
    subpix_method = pk.SUBPIX_LM if subpix == "lm" else pk.SUBPIX_POWELL
    prm = pk.track_params(w.params.search_zone, w.params.search_zone_rescue, w.params.search_zone_super_rescue,
                          NCC_min=0.7, rescue_level=1, subpixel=1, search_method=search_method,
                          subpix=pk.default_subpix(subpix_method))

    def step(p):
        tr.put_grains(w.grains)
        tr.track_pair(p)
        return tr.get_grains()


    for _ in range(args.warmup):
        step(prm)
    tr.sync()
    # the timed pairs, with no stage timers (their events are not part of the work)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        out = step(prm)
    tr.sync()
    el = time.perf_counter() - t1
    # the stage split from as many pairs again, with HIP events around each stage
    tr.set_timing(True)
    tr.reset_stage_times()
    for _ in range(args.steps):
        step(prm)
    tr.sync()
    ms, launches = tr.stage_times()
    tr.set_timing(False)
    nres, _ = tr.queue_counts()
    evals = tr.subpix_evals()
    value = w.G * args.steps / el
    ms_per_step = 1e3 * el / args.steps
    log(f"{method}+{subpix}: {ms_per_step:.2f} ms/pair, {value:.0f} grains/s, rescued {nres}, "
        f"subpixel evals {evals}")
    stage_names = ["precompute", "search_main", "search_rescue", "search_super", "subpixel", "remap", "search_kernel"]
    stage_ms = {n: (ms[i] / args.steps) for i, n in enumerate(stage_names) if launches[i]}

    # dominant kernel: the main-stage candidate-scoring kernel
    unmasked = ~w.grains["masked"].astype(bool)
    P = np.diff(w.pattern_begin)[unmasked].astype(np.float64)
    z = w.params.search_zone
    R = len(ptref.rotation_steps(z.inc_rot, z.num_rot))
    S = (z.left + z.right + 1) * (z.up + z.down + 1)
    flops = 2.0 * R * S * P.sum()
    # the scoring kernel's GPU time per pair (one launch for the plain form, one per
    # pattern-radius bucket for the exact form), HIP events on the library stream
    k_ms = ms[pk.T_SEARCH_KERNEL] / args.steps
    k_launches = launches[pk.T_SEARCH_KERNEL] / args.steps
    achieved = flops / (k_ms * 1e-3) / 1e12
    if method == "exact":
        roof = {"bound": "mfma", "achieved": achieved, "peak": INT8_PEAK_TOPS, "unit": "TFLOP/s",
                "frac": achieved / INT8_PEAK_TOPS, "traffic": None}
    else:
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None}
    roof["kernel"] = "k_search_exact (MFMA)" if method == "exact" else "k_search_plain (FP64)"
    if method == "exact":
        # the launches' own durations without overlap: the same pairs with every kernel of the
        # exact search in order on the library stream (the radius buckets' launches otherwise
        # run two at a time, and each event-timed duration then includes CU time shared with
        # the other); the achieved rate is taken from these
        tr.set_option(pk.OPT_SEARCH_STREAMS, 1)
        step(prm)
        tr.sync()
        tr.set_timing(True)
        tr.reset_stage_times()
        t4 = time.perf_counter()
        for _ in range(args.steps):
            step(prm)
        tr.sync()
        ser_el = time.perf_counter() - t4
        ms_s, launches_s = tr.stage_times()
        tr.set_timing(False)
        tr.set_option(pk.OPT_SEARCH_STREAMS, 1 if args.serial_search else 2)
        k_ms_overlapped = k_ms
        k_ms = ms_s[pk.T_SEARCH_KERNEL] / args.steps
        achieved = flops / (k_ms * 1e-3) / 1e12
        roof["achieved"] = achieved
        roof["frac"] = achieved / INT8_PEAK_TOPS
        roof["timing"] = ("HIP events around each k_search_exact launch with every kernel of the search "
                          "on one stream (no overlap)")
        roof["serial_ms_per_pair"] = 1e3 * ser_el / args.steps
        roof["kernel_ms_per_pair_overlapped"] = k_ms_overlapped
    # HBM bytes per launch of the search kernel, from the committed byte-counter summary
    # (FETCH_SIZE x 2 + WRITE_SIZE, calibrated on a device copy of known size)
    if method == "exact":
        tb = hbm_bytes_per_launch("k_search_exact")
        if tb is not None:
            roof["traffic"] = tb
            roof["traffic_source"] = TRAFFIC_FILE
    roof["algorithmic_flops_per_pair"] = flops
    roof["kernel_ms_per_pair"] = k_ms
    roof["launches_per_pair"] = k_launches
    if method == "exact":
        # the radius buckets' launches run two at a time on two streams, so their summed
        # durations exceed the time the main stage takes: the same flops over the stage's
        # own span (HIP events on the library stream around the whole search)
        st_ms = ms[pk.T_SEARCH_MAIN] / args.steps
        roof["search_stage_ms_per_pair"] = st_ms
        roof["achieved_over_stage_span"] = flops / (st_ms * 1e-3) / 1e12

    result = {
        "metric": METRIC, "value": value, "unit": "grains/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "i8" if method == "exact" else "f64", "data": "synthetic",
        "config": {"workload": "6000x4000 12-bit frame pair, 2000 disk grains (lognormal r, median 20, 8-40 px), "
                               "search +-20 px x 21 rotation steps, rescue +-40 px below NCC 0.7, subpixel",
                   "grains": int(w.G), "pattern_px": int(len(w.pattern_xy)), "search": method,
                   "subpixel": subpix, "rotation_steps": R, "shifts": S, "rescued": int(nres),
                   "subpixel_evals": int(evals)},
        "stage_ms": stage_ms,
        "roofline": roof,
    }
    if method == "exact" and args.autotune:
        result["autotune"] = autotune_section(pk, tr, w, prm, step, args)
    if args.graph_steps > 0:
        result["graph"] = graph_section(pk, tr, w, prm, step, out, args)
    if args.ingest_steps > 0:
        result["frames_from_host"] = ingest_section(pk, tr, w, prm, step, out, args)

    # the plain form (FP64 two-pass search + Powell) on the same workload
    if args.plain_steps > 0 and not (method == "plain" and subpix == "powell"):
        pp = pk.track_params(w.params.search_zone, w.params.search_zone_rescue, w.params.search_zone_super_rescue,
                             NCC_min=0.7, rescue_level=1, subpixel=1, search_method=pk.SEARCH_PLAIN,
                             subpix=pk.default_subpix(pk.SUBPIX_POWELL))
        step(pp)
        tr.sync()
        t2 = time.perf_counter()
        for _ in range(args.plain_steps):
            po = step(pp)
        tr.sync()
        pe = time.perf_counter() - t2
        pv = w.G * args.plain_steps / pe
        dev = float(max(np.abs(out["dx"] - po["dx"]).max(), np.abs(out["dy"] - po["dy"]).max()))
        result["plain_gpu"] = {"value": pv, "unit": "grains/s", "steps": args.plain_steps,
                               "speedup": value / pv, "max_px_diff_vs_plain": dev}
        log(f"plain search+Powell: {1e3 * pe / args.plain_steps:.2f} ms/pair, {pv:.0f} grains/s")
    # precision form: integer picks of the exact MFMA search against the plain FP64 search
    # (bit-identical to the reference's arithmetic) on every grain, main stage only
    if method == "exact" and not args.no_precision:
        prec = {"cfg2_21_steps_12bit": pick_mismatches(pk, tr, w)}
        tr.close()
        w3 = synth.workload(3, n_grains=args.grains)
        t3 = pk.Tracker(0)
        prec["cfg3_40_steps_12bit"] = pick_mismatches(pk, t3, w3, load=True)
        w3.I0, w3.I1, w3.bits = (w3.I0 >> 4).astype(np.uint16), (w3.I1 >> 4).astype(np.uint16), 8
        prec["cfg3_40_steps_8bit"] = pick_mismatches(pk, t3, w3, load=True)
        t3.close()
        result["precision"] = prec
        log("exact vs plain integer picks: " + ", ".join(f"{k}: {v['mismatches']} of {v['grains']}" for k, v in prec.items()))
    else:
        tr.close()

    if method == "exact" and not args.no_depths:
        result["bit_depths"] = depth_variants(pk, w, args, flops)

    if not args.no_rescue_variant:
        result["rescue_variant"] = rescue_variant(pk, synth, args)

    if args.cpu_grains > 0:
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or \
            len(os.sched_getaffinity(0))
        rng = np.random.default_rng(0)
        sample = np.sort(rng.choice(w.G, size=min(args.cpu_grains, w.G), replace=False))
        W = ptref.Workload(w.I0, w.I1, w.grains, w.pattern_xy, w.pattern_begin, w.bits)
        cp = pk.track_params(w.params.search_zone, w.params.search_zone_rescue, w.params.search_zone_super_rescue,
                             NCC_min=0.7, rescue_level=1, subpixel=1, search_method=pk.SEARCH_PLAIN,
                             subpix=pk.default_subpix(pk.SUBPIX_POWELL))
        # one warm-up run, then the median of three
        runs = []
        for r in range(4):
            t3 = time.perf_counter()
            ptref.track_pair(W, cp, nthreads=threads, subset=sample)
            if r > 0:
                runs.append(time.perf_counter() - t3)
        ce = float(np.median(runs))
        cpu_model = ""
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        # the fast form's error against the FP64 oracle (the oracle's Powell at a tight
        # tolerance, restarted until it no longer improves) on the same sample, untimed
        tp = pk.track_params(w.params.search_zone, w.params.search_zone_rescue, w.params.search_zone_super_rescue,
                             NCC_min=0.7, rescue_level=1, subpixel=1, search_method=pk.SEARCH_PLAIN,
                             subpix=pk.default_subpix(pk.SUBPIX_LM, subpix_tol=1e-13))
        ot, _, _ = ptref.track_pair(W, tp, nthreads=threads, subset=sample)
        fs = out[sample]
        epx = np.hypot(fs["dx"] - ot["dx"], fs["dy"] - ot["dy"])
        erad = np.abs(fs["drot"] - ot["drot"])
        encc = np.abs(fs["NCC_subpix"] - ot["NCC_subpix"])
        result["fp64_oracle_error"] = {
            "grains": int(len(sample)), "max_px": float(epx.max()), "max_rad": float(erad.max()),
            "max_ncc_subpix": float(encc.max()), "over_1e-3_px": int((epx > 1e-3).sum()),
            "over_1e-5_rad": int((erad > 1e-5).sum())}
        log(f"fast form vs FP64 oracle on {len(sample)} grains: max {epx.max():.2e} px, {erad.max():.2e} rad, "
            f"{(epx > 1e-3).sum()} grains over 1e-3 px, {(erad > 1e-5).sum()} over 1e-5 rad")
        result["cpu_baseline"] = {
            "value": len(sample) / ce, "unit": "grains/s", "cores": threads, "kind": "port",
            "sample": f"{len(sample)} grains drawn from the same pair (seed 0), whole per-pair path (precompute, "
                      f"search, rescue, Powell) through the oracle (gcc -O3 -fopenmp -ffp-contract=off) with OpenMP "
                      f"schedule(dynamic) on {threads} threads (the CPUs this process may use: "
                      f"{len(os.sched_getaffinity(0))} of {os.cpu_count()} online); one warm-up, median of 3 runs "
                      f"{', '.join(f'{x:.2f}' for x in runs)} s; {cpu_model}"}
        log(f"cpu baseline: {len(sample) / ce:.1f} grains/s on {threads} threads (median {ce:.2f}s)")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
