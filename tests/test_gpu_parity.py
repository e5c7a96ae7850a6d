"""GPU parity: libptgpu (through its C ABI) against the oracle on the same seeded inputs.

Bars: integer picks, NCC and NCC_rescue of the plain search bit-identical; precompute
(mean0, C0C0) bit-identical; sampler, undistor and remap bit-identical; subpixel
(u, v) within 1e-3 px, theta within 1e-5 rad and NCC_subpix within 1e-6 of the
oracle's Powell."""
import math
import sys

import numpy as np
import pytest

import ptref

pk = ptref.package()
from ptracker_amd import synth  # noqa: E402

pytestmark = pytest.mark.gpu

TOL_PX, TOL_RAD, TOL_NCC = 1e-3, 1e-5, 1e-6
INT_COLS = ("dx", "dy", "drot", "upix", "vpix", "rot_inc", "NCC", "NCC_rescue")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits_equal(a, b):
    return np.array_equal(np.asarray(a, dtype=np.float64).view(np.uint64), np.asarray(b, dtype=np.float64).view(np.uint64))


def oracle_workload(w, grains=None):
    return ptref.Workload(w.I0, w.I1, w.grains if grains is None else grains, w.pattern_xy, w.pattern_begin, w.bits)


def load(tracker, w, grains=None):
    tracker.set_image(0, w.I0, w.bits)
    tracker.set_image(1, w.I1, w.bits)
    tracker.set_image_indices(0, 1)
    tracker.set_grains(w.grains if grains is None else grains, w.pattern_xy, w.pattern_begin)


@pytest.fixture(scope="module")
def cfg1():
    return synth.workload(1)


@pytest.fixture(scope="module")
def cfg2s():
    # config-2 texture, motion and search at a reduced frame and grain count
    return synth.workload(2, n_grains=48, size=(1100, 900), seed=2024)


def test_sampler_bit_exact(tracker):
    rng = np.random.default_rng(1)
    f = rng.integers(0, 4096, size=(300, 400)).astype(np.uint16)
    say("sampler 300x400, 20000 points")
    tracker.set_image(2, f, 12)
    xy = np.stack([rng.uniform(1, 396.99, 20000), rng.uniform(1, 296.99, 20000)], axis=1)
    g = tracker.interp_cubic(2, xy)
    # the oracle's sampler is pinned to the reference's compiled interpol.cpp by the CPU tests
    assert bits_equal(g, ptref.interp_cubic(f, xy))


def test_precompute_bit_exact(tracker, cfg2s):
    say("precompute cfg2 small")
    load(tracker, cfg2s)
    tracker.precompute()
    got = tracker.get_grains()
    ref, _ = ptref.precompute(oracle_workload(cfg2s))
    assert bits_equal(got["mean0"], ref["mean0"])
    assert bits_equal(got["C0C0"], ref["C0C0"])


def _oracle_main(w, z, NCC_min=0.7):
    W = oracle_workload(w)
    gr, p0 = ptref.precompute(W)
    gr["NCC"] = gr["NCC_rescue"] = gr["NCC_subpix"] = 0.0
    gr["dx_prev"], gr["dy_prev"], gr["drot_prev"] = gr["dx"], gr["dy"], gr["drot"]
    queued = []
    for g in range(len(gr)):
        if gr[g]["masked"]:
            continue
        if ptref.follow_pattern(W, gr, p0, g, z, pk.STAGE_MAIN, NCC_min, 0.5):
            queued.append(g)
    return gr, queued


@pytest.mark.parametrize("which", ["cfg1", "cfg2s"])
def test_main_search_plain_bit_exact(tracker, which, request):
    w = request.getfixturevalue(which)
    z = w.params.search_zone
    say(f"main search plain {which}: G={w.G} zone={z.left} steps num_rot={z.num_rot} inc={z.inc_rot}")
    load(tracker, w)
    tracker.begin_pair()
    tracker.search(pk.STAGE_MAIN, z, 0.7, 0.5, pk.SEARCH_PLAIN)
    got = tracker.get_grains()
    ref, queued = _oracle_main(w, z)
    for col in INT_COLS:
        assert bits_equal(got[col], ref[col]), col
    assert tracker.queue_counts()[0] == len(queued)


def test_rescue_and_super_rescue_bit_exact(tracker):
    # motion of up to 6 px against a +-2 px main zone forces both rescue stages
    w = synth.workload(2, n_grains=24, size=(800, 700), max_disp=6.0, seed=77)
    z, zr, zs = pk.zone(2, num_rot=2, inc_rot=0.02), pk.zone(4, num_rot=2, inc_rot=0.02), pk.zone(8, num_rot=1, inc_rot=0.05)
    prm = pk.track_params(z, zr, zs, NCC_min=0.7, NCC_min_super=0.6, rescue_level=2, subpixel=0)
    say("rescue + super-rescue, 24 grains 800x700")
    load(tracker, w)
    tracker.track_pair(prm)
    got = tracker.get_grains()
    nres, nsup = tracker.queue_counts()
    ref, counts, _ = ptref.track_pair(oracle_workload(w), prm, nthreads=4)
    assert (nres, nsup) == counts
    assert nres > 0 and nsup > 0
    for col in INT_COLS:
        assert bits_equal(got[col], ref[col]), col


def test_masked_grains_are_skipped(tracker, cfg1):
    grains = cfg1.grains.copy()
    grains["masked"][[0, 5, 9]] = 1
    grains["dx"][5] = 0.3
    say("masked grains cfg1")
    load(tracker, cfg1, grains)
    tracker.track_pair(cfg1.params)
    got = tracker.get_grains()
    assert np.all(got["NCC"][[0, 5, 9]] == 0.0) and got["dx"][5] == 0.3 and np.all(got["upix"][[0, 5, 9]] == 0.0)
    ref, _, _ = ptref.track_pair(oracle_workload(cfg1, grains), cfg1.params, nthreads=4)
    for col in ("NCC", "NCC_rescue"):
        assert bits_equal(got[col], ref[col]), col


def test_objective_matches_oracle(tracker, cfg2s):
    say("objective cfg2 small")
    load(tracker, cfg2s)
    W = oracle_workload(cfg2s)
    gr, p0 = ptref.precompute(W)
    rng = np.random.default_rng(5)
    ig = rng.integers(0, cfg2s.G, 40)
    X = np.stack([rng.uniform(-1, 1, 40), rng.uniform(-1, 1, 40), rng.uniform(-0.05, 0.05, 40)], axis=1)
    f = tracker.objective(ig, X)
    fo = np.array([ptref.ncc_objective(W, gr, p0, g, x) for g, x in zip(ig, X)])
    assert np.allclose(f, fo, rtol=0, atol=1e-10)


def test_gradient_against_central_differences(tracker, cfg2s):
    say("gradient vs central differences")
    load(tracker, cfg2s)
    W = oracle_workload(cfg2s)
    gr, p0 = ptref.precompute(W)
    rng = np.random.default_rng(6)
    ig = rng.integers(0, cfg2s.G, 12)
    X = np.stack([rng.uniform(-0.6, 0.6, 12), rng.uniform(-0.6, 0.6, 12), rng.uniform(-0.03, 0.03, 12)], axis=1)
    f, g = tracker.objective_grad(ig, X)
    h = np.array([1e-3, 1e-3, 1e-5])
    for k in range(3):
        Xp, Xm = X.copy(), X.copy()
        Xp[:, k] += h[k]
        Xm[:, k] -= h[k]
        fd = (tracker.objective(ig, Xp) - tracker.objective(ig, Xm)) / (2 * h[k])
        # the reference's numeric Jacobian is a difference of its own objective
        fdo = np.array([(ptref.ncc_objective(W, gr, p0, gi, xp) - ptref.ncc_objective(W, gr, p0, gi, xm)) / (2 * h[k])
                        for gi, xp, xm in zip(ig, Xp, Xm)])
        scale = np.abs(g[:, k]).max()
        assert np.allclose(g[:, k], fd, rtol=0, atol=2e-3 * scale), k
        assert np.allclose(g[:, k], fdo, rtol=0, atol=2e-3 * scale), k


@pytest.mark.parametrize("method", ["powell", "lm"])
def test_track_pair_cfg1_against_oracle(tracker, cfg1, method):
    m = pk.SUBPIX_POWELL if method == "powell" else pk.SUBPIX_LM
    prm = pk.track_params(cfg1.params.search_zone, rescue_level=0, subpixel=1, subpix=pk.default_subpix(m))
    say(f"track_pair cfg1 subpixel={method}")
    load(tracker, cfg1)
    tracker.track_pair(prm)
    got = tracker.get_grains()
    # Powell parity: the oracle's Powell at the reference tolerance; LM: the FP64 oracle,
    # the oracle's Powell at a tight tolerance restarted to convergence (same basin)
    oprm = cfg1.params if method == "powell" else pk.track_params(
        cfg1.params.search_zone, rescue_level=0, subpixel=1, subpix=pk.default_subpix(pk.SUBPIX_LM, subpix_tol=1e-13))
    ref, _, _ = ptref.track_pair(oracle_workload(cfg1), oprm, nthreads=4)
    assert bits_equal(got["NCC"], ref["NCC"]) and bits_equal(got["NCC_rescue"], ref["NCC_rescue"])
    assert np.abs(got["dx"] - ref["dx"]).max() < TOL_PX
    assert np.abs(got["dy"] - ref["dy"]).max() < TOL_PX
    assert np.abs(got["drot"] - ref["drot"]).max() < TOL_RAD
    assert np.abs(got["NCC_subpix"] - ref["NCC_subpix"]).max() < TOL_NCC
    for col in ("upix", "vpix", "rot_inc"):
        assert np.abs(got[col] - ref[col]).max() < (TOL_RAD if col == "rot_inc" else TOL_PX), col


def test_track_pair_cfg2_small_lm_against_fp64_oracle(tracker, cfg2s):
    say("track_pair cfg2 small, exact search + LM")
    p = cfg2s.params
    prm = pk.track_params(p.search_zone, p.search_zone_rescue, p.search_zone_super_rescue, rescue_level=1,
                          search_method=pk.SEARCH_EXACT, subpix=pk.default_subpix(pk.SUBPIX_LM))
    load(tracker, cfg2s)
    tracker.track_pair(prm)
    got = tracker.get_grains()
    oprm = pk.track_params(p.search_zone, p.search_zone_rescue, p.search_zone_super_rescue, rescue_level=1,
                           subpix=pk.default_subpix(pk.SUBPIX_LM, subpix_tol=1e-13))
    ref, _, _ = ptref.track_pair(oracle_workload(cfg2s), oprm, nthreads=8)
    assert np.abs(got["NCC"] - ref["NCC"]).max() < 1e-12
    assert np.abs(got["dx"] - ref["dx"]).max() < TOL_PX and np.abs(got["dy"] - ref["dy"]).max() < TOL_PX
    assert np.abs(got["drot"] - ref["drot"]).max() < TOL_RAD
    assert np.abs(got["NCC_subpix"] - ref["NCC_subpix"]).max() < TOL_NCC


def test_track_pair_cfg2_small_against_oracle(tracker, cfg2s):
    say("track_pair cfg2 small, Powell")
    load(tracker, cfg2s)
    tracker.track_pair(cfg2s.params)
    got = tracker.get_grains()
    ref, counts, _ = ptref.track_pair(oracle_workload(cfg2s), cfg2s.params, nthreads=8)
    assert tracker.queue_counts()[0] == counts[0]
    assert bits_equal(got["NCC"], ref["NCC"])
    assert np.abs(got["dx"] - ref["dx"]).max() < TOL_PX and np.abs(got["dy"] - ref["dy"]).max() < TOL_PX
    assert np.abs(got["drot"] - ref["drot"]).max() < TOL_RAD
    assert np.abs(got["NCC_subpix"] - ref["NCC_subpix"]).max() < TOL_NCC


def test_undistort_points_and_remap_bit_exact(tracker):
    rng = np.random.default_rng(9)
    H, W = 260, 330
    f = rng.integers(0, 4096, size=(H, W)).astype(np.uint16)
    X = np.array([-1.0, -1.0, -2e-6, 3e-12, 0.0, 1e-6, -2e-6, 0.0, -1.0, -1.0, 0.01, 0.002, -0.02, 0.001])
    say("remap 330x260")
    tracker.set_image(2, f, 12)
    tracker.remap(2, 3, X)
    tracker.sync()
    # read image 3 back through the sampler at integer nodes
    ys, xs = np.mgrid[2:H - 3, 2:W - 3]
    nodes = np.stack([xs.ravel(), ys.ravel()], axis=1).astype(np.float64)
    got = tracker.interp_cubic(3, nodes).reshape(ys.shape)
    ref = ptref.remap(f, 12, X)
    assert np.array_equal(got, ref[2:H - 3, 2:W - 3].astype(np.float64))
    Xp = X.copy()
    Xp[[0, 8]], Xp[[1, 9]] = W / 2, H / 2
    pts = np.stack([rng.uniform(0, W, 500), rng.uniform(0, H, 500)], axis=1)
    u = tracker.undistort_points(Xp, pts)
    uo = np.array([ptref.undistor(Xp, x, y) for x, y in pts])
    assert bits_equal(u, uo)


def test_remap_config4_lens_full_frame_bit_exact(tracker):
    # the config-4 lens (k1 = -2e-9, k2 = 1e-16 about the centre) on a whole 6000x4000 12-bit
    # frame: every pixel of the GPU remap equals the oracle's (itself pinned by the quadratic
    # closed form and the round trip of the CPU tests)
    w = synth.workload(2, n_grains=200, seed=1002)
    H, W = w.I0.shape
    X = np.array([0.5 * W, 0.5 * H, -2e-9, 1e-16, 0, 0, 0, 0, 0.5 * W, 0.5 * H, 0, 0, 0, 0], dtype=np.float64)
    say("remap 6000x4000, config-4 lens")
    tracker.set_image(2, w.I0, 12)
    tracker.remap(2, 3, X)
    got = tracker.get_image(3, W, H)
    ref = ptref.remap(w.I0, 12, X)
    assert np.array_equal(got, ref)
    # and the remap undoes the lens: the config-4 frame remapped is the config-2 frame
    # within the round trip of two bicubic resamplings
    seen = synth.distort_frame(w.I0, X, 12)
    tracker.set_image(2, seen, 12)
    tracker.remap(2, 3, X)
    back = tracker.get_image(3, W, H).astype(np.float64)
    err = np.abs(back - w.I0)[50:-50, 50:-50].mean()
    uncorrected = np.abs(seen.astype(np.float64) - w.I0)[50:-50, 50:-50].mean()
    say(f"round trip: mean |error| {err:.3f}, uncorrected {uncorrected:.1f} gray levels")
    assert err < 2.0 and err < uncorrected / 20


def test_change_reference_sequence(tracker):
    # two consecutive pairs with the reference swapped in between (src/p-tracker.cpp:587-611)
    w1 = synth.workload(1, n_grains=10, size=(300, 300), seed=500)
    say("sequence of two pairs with reference change")
    load(tracker, w1)
    prm = pk.track_params(w1.params.search_zone, rescue_level=0, subpixel=0)
    tracker.track_pair(prm)
    g1 = tracker.get_grains()
    ref1, _, _ = ptref.track_pair(oracle_workload(w1), prm, nthreads=2)
    assert bits_equal(g1["dx"], ref1["dx"])
    tracker.change_reference()
    tracker.set_image(0, w1.I0, w1.bits)  # becomes the current frame after the swap
    tracker.track_pair(prm)
    g2 = tracker.get_grains()
    ref1c = ptref.change_reference(ref1)
    W2 = ptref.Workload(w1.I1, w1.I0, ref1c, w1.pattern_xy, w1.pattern_begin, w1.bits)
    ref2, _, _ = ptref.track_pair(W2, prm, nthreads=2)
    for col in INT_COLS:
        assert bits_equal(g2[col], ref2[col]), col
    assert np.array_equal(g2["refcoord_xpix"], ref2["refcoord_xpix"])


# ---- exact-integer MFMA search -------------------------------------------------------------
@pytest.mark.parametrize("which", ["cfg1", "cfg2s"])
def test_main_search_exact_picks_equal_oracle(tracker, which, request):
    w = request.getfixturevalue(which)
    z = w.params.search_zone
    say(f"main search exact {which}: G={w.G}")
    load(tracker, w)
    tracker.begin_pair()
    tracker.search(pk.STAGE_MAIN, z, 0.7, 0.5, pk.SEARCH_EXACT)
    got = tracker.get_grains()
    ref, queued = _oracle_main(w, z)
    for col in ("dx", "dy", "drot", "upix", "vpix", "rot_inc"):
        assert bits_equal(got[col], ref[col]), col
    assert np.abs(got["NCC"] - ref["NCC"]).max() < 1e-12
    assert tracker.queue_counts()[0] == len(queued)


def test_rescue_exact_picks_equal_oracle(tracker):
    w = synth.workload(2, n_grains=24, size=(800, 700), max_disp=6.0, seed=77)
    z, zr, zs = pk.zone(2, num_rot=2, inc_rot=0.02), pk.zone(4, num_rot=2, inc_rot=0.02), pk.zone(8, num_rot=1, inc_rot=0.05)
    prm = pk.track_params(z, zr, zs, NCC_min=0.7, NCC_min_super=0.6, rescue_level=2, subpixel=0,
                          search_method=pk.SEARCH_EXACT)
    say("rescue + super-rescue exact")
    load(tracker, w)
    tracker.track_pair(prm)
    got = tracker.get_grains()
    ref, counts, _ = ptref.track_pair(oracle_workload(w), prm, nthreads=4)
    assert tracker.queue_counts() == counts
    for col in ("dx", "dy", "drot"):
        assert bits_equal(got[col], ref[col]), col
    assert np.abs(got["NCC_rescue"] - ref["NCC_rescue"]).max() < 1e-12


@pytest.mark.parametrize("bits", [12, 8, 16])
def test_cfg3_exact_40_steps_picks_equal_oracle(tracker, bits):
    # config 3: +-20 px x 20 num_rot at inc 0.005, which is 40 steps in the reference's
    # loop; 12-bit frames, their 8-bit requantisation (v >> 4: one offset-INT8 plane) and
    # their spread over 16 bits (8 M-tiles of 5 steps). Zero argmax mismatches.
    w = synth.workload(3, n_grains=20, size=(1000, 800), seed=3033)
    if bits == 8:
        w.I0 = (w.I0 >> 4).astype(np.uint16)
        w.I1 = (w.I1 >> 4).astype(np.uint16)
        w.bits = 8
    elif bits == 16:
        w.I0, w.I1, w.bits = to16(w.I0, 330), to16(w.I1, 331), 16
    z = w.params.search_zone
    assert len(ptref.rotation_steps(z.inc_rot, z.num_rot)) == 40
    prm = pk.track_params(z, rescue_level=0, subpixel=0, search_method=pk.SEARCH_EXACT)
    say(f"config 3, 40 steps, {bits}-bit, exact")
    load(tracker, w)
    tracker.track_pair(prm)
    got = tracker.get_grains()
    oprm = pk.track_params(z, rescue_level=0, subpixel=0)
    ref, _, _ = ptref.track_pair(oracle_workload(w), oprm, nthreads=8)
    mism = sum(int(not bits_equal(got[c][i:i + 1], ref[c][i:i + 1])) for c in ("dx", "dy", "drot")
               for i in range(w.G))
    assert mism == 0
    assert np.abs(got["NCC"] - ref["NCC"]).max() < 1e-12


def to16(img, seed):
    """12-bit frame data spread over the whole 16-bit range: v * 16 + U{0, ..., 15}."""
    rng = np.random.default_rng(seed)
    return (img.astype(np.uint32) * 16 + rng.integers(0, 16, img.shape)).astype(np.uint16)


@pytest.mark.parametrize("which", ["cfg1", "cfg2s"])
def test_exact_16bit_frames_picks_equal_oracle(tracker, which, request):
    # 16-bit frames on the matrix cores (the window as two offset bytes, the template as three
    # 7-bit digits, num and d1 in int64): config 1's 8-bit data declared 16-bit, and the
    # config-2 texture spread over 0 .. 65535; no grain leaves the exact kernels
    w = request.getfixturevalue(which)
    I0, I1 = (w.I0, w.I1) if which == "cfg1" else (to16(w.I0, 160), to16(w.I1, 161))
    say(f"exact search on 16-bit frames, {which}: G={w.G}, values up to {int(max(I0.max(), I1.max()))}")
    tracker.set_image(0, I0, 16)
    tracker.set_image(1, I1, 16)
    tracker.set_image_indices(0, 1)
    tracker.set_grains(w.grains, w.pattern_xy, w.pattern_begin)
    n0 = tracker.exact_fallbacks()
    z = w.params.search_zone
    tracker.track_pair(pk.track_params(z, rescue_level=0, subpixel=0, search_method=pk.SEARCH_EXACT))
    got = tracker.get_grains()
    assert tracker.exact_fallbacks() == n0
    W = ptref.Workload(I0, I1, w.grains, w.pattern_xy, w.pattern_begin, 16)
    ref, _, _ = ptref.track_pair(W, pk.track_params(z, rescue_level=0, subpixel=0), nthreads=8)
    for col in ("dx", "dy", "drot", "upix", "vpix", "rot_inc"):
        assert bits_equal(got[col], ref[col]), col
    assert np.abs(got["NCC"] - ref["NCC"]).max() < 1e-12


def test_exact_16bit_rescue_and_super_rescue_equal_oracle(tracker):
    # the rescue stages on 16-bit frames, with disks and squares (the squares' base mask is
    # their middle rotation step, from the base-mask GEMM on the offset byte planes)
    w = synth.workload(2, n_grains=24, size=(800, 700), max_disp=6.0, seed=77)
    I0, I1 = to16(w.I0, 770), to16(w.I1, 771)
    pats = []
    for g in range(w.G):
        P = w.pattern_xy[w.pattern_begin[g]:w.pattern_begin[g + 1]]
        if g % 3 == 1:
            r = math.isqrt(int((P.astype(np.int64) ** 2).sum(1).max())) // 2
            ys, xs = np.mgrid[-r:r + 1, -r:r + 1]
            P = np.stack([xs.ravel(), ys.ravel()], 1)
        pats.append(P.astype(np.int32))
    pxy = np.concatenate(pats).astype(np.int32)
    pb = np.concatenate([[0], np.cumsum([len(p) for p in pats])]).astype(np.int64)
    z, zr, zs = pk.zone(2, num_rot=2, inc_rot=0.02), pk.zone(4, num_rot=2, inc_rot=0.02), pk.zone(8, num_rot=1, inc_rot=0.05)
    prm = pk.track_params(z, zr, zs, NCC_min=0.7, NCC_min_super=0.6, rescue_level=2, subpixel=0,
                          search_method=pk.SEARCH_EXACT)
    say("rescue + super-rescue exact, 16-bit frames, disks and squares")
    tracker.set_image(0, I0, 16)
    tracker.set_image(1, I1, 16)
    tracker.set_image_indices(0, 1)
    tracker.set_grains(w.grains, pxy, pb)
    n0 = tracker.exact_fallbacks()
    tracker.track_pair(prm)
    got = tracker.get_grains()
    assert tracker.exact_fallbacks() == n0
    ref, counts, _ = ptref.track_pair(ptref.Workload(I0, I1, w.grains, pxy, pb, 16), prm, nthreads=4)
    assert counts[0] > 0 and counts[1] > 0
    assert tracker.queue_counts() == counts
    for col in ("dx", "dy", "drot"):
        assert bits_equal(got[col], ref[col]), col
    assert np.abs(got["NCC"] - ref["NCC"]).max() < 1e-12
    assert np.abs(got["NCC_rescue"] - ref["NCC_rescue"]).max() < 1e-12


# ---- frame sequences (batching form) -------------------------------------------------------
def _sequence_vs_oracle(tracker, S, prm, oprm, pairs):
    """Tracks pairs 0 .. pairs-1 of sequence S on the GPU (reference frame swapped after
    each pair) and with the oracle from its own previous output; yields both per pair."""
    tracker.set_image(0, S.frames[0], S.bits)
    tracker.set_image(1, S.frames[1], S.bits)
    tracker.set_image_indices(0, 1)
    tracker.set_grains(S.grains, S.pattern_xy, S.pattern_begin)
    gin = S.grains.copy()
    for k in range(pairs):
        if k > 0:
            tracker.change_reference()
            tracker.set_image((k + 1) % 2, S.frames[k + 1], S.bits)
        tracker.track_pair(prm)
        got = tracker.get_grains()
        Wk = ptref.Workload(S.frames[k], S.frames[k + 1], gin, S.pattern_xy, S.pattern_begin, S.bits)
        ref, _, _ = ptref.track_pair(Wk, oprm, nthreads=8)
        yield k, got, ref
        gin = ptref.change_reference(ref)


@pytest.mark.parametrize("case", ["r50-64_s8", "r78-82_s2"])
def test_sequence_large_radii_exact_picks_equal_oracle(tracker, case):
    # patterns of radius 50-64 at +-8 px and 78-82 at +-2 px (near the largest the MFMA
    # kernel's LDS takes): the middle-step kernel stages its six digit planes in two
    # passes; three pairs with the reference swapped in between, so refrot != 0 from the
    # second pair on
    if case == "r50-64_s8":
        S = synth.sequence(n_frames=4, n_grains=8, size=(700, 600), seed=505, radius=(50.0, 64.0), margin=150)
        z = S.params.search_zone
    else:
        S = synth.sequence(n_frames=4, n_grains=4, size=(700, 600), seed=506, radius=(78.0, 82.0), margin=150,
                           max_step=1.5, max_rot=0.01)
        z = pk.zone(2, num_rot=1, inc_rot=0.01)
    say(f"sequence {case}: G={S.G}, P max {np.diff(S.pattern_begin).max()}")
    prm = pk.track_params(z, rescue_level=0, subpixel=0, search_method=pk.SEARCH_EXACT)
    oprm = pk.track_params(z, rescue_level=0, subpixel=0)
    n0 = tracker.exact_fallbacks()
    for k, got, ref in _sequence_vs_oracle(tracker, S, prm, oprm, 3):
        for col in ("dx", "dy", "drot", "upix", "vpix", "rot_inc"):
            assert bits_equal(got[col], ref[col]), (k, col)
        assert np.abs(got["NCC"] - ref["NCC"]).max() < 1e-12
        assert bits_equal(got["refrot"], ref["refrot"])
    assert tracker.exact_fallbacks() == n0


def test_sequence_beyond_exact_limits_falls_back_to_plain(tracker):
    # radius 80.5-82 at +-8 px: the MFMA kernel's window and staging exceed the CU's LDS,
    # so the search runs the plain form with the same picks
    S = synth.sequence(n_frames=2, n_grains=3, size=(700, 600), seed=507, radius=(80.5, 82.0), margin=150,
                       max_step=3.0, max_rot=0.01)
    z = pk.zone(8, num_rot=1, inc_rot=0.01)
    say(f"sequence beyond exact limits: G={S.G}")
    prm = pk.track_params(z, rescue_level=0, subpixel=0, search_method=pk.SEARCH_EXACT)
    oprm = pk.track_params(z, rescue_level=0, subpixel=0)
    n0 = tracker.exact_fallbacks()
    for k, got, ref in _sequence_vs_oracle(tracker, S, prm, oprm, 1):
        for col in INT_COLS:
            assert bits_equal(got[col], ref[col]), col
    assert tracker.exact_fallbacks()[0] == n0[0] + 1


@pytest.mark.parametrize("method", [pk.SEARCH_PLAIN, pk.SEARCH_EXACT])
def test_patterns_beyond_the_lds_equal_oracle(tracker, method):
    # radius 140-160 (about 70,000 pattern pixels): the rotated offsets alone exceed the LDS,
    # so the search runs the plain kernel that reads the frame through the caches (the
    # exact form hands the whole stage to it); picks bit-identical, and the LM subpixel
    # within tolerance of the FP64 oracle
    S = synth.sequence(n_frames=2, n_grains=3, size=(1000, 900), seed=511, radius=(140.0, 160.0), margin=200,
                       max_step=2.0, max_rot=0.01)
    z = pk.zone(2, num_rot=1, inc_rot=0.01)
    say(f"patterns beyond the LDS: G={S.G}, P max {np.diff(S.pattern_begin).max()}, method {method}")
    prm = pk.track_params(z, rescue_level=0, subpixel=0, search_method=method)
    oprm = pk.track_params(z, rescue_level=0, subpixel=0)
    for k, got, ref in _sequence_vs_oracle(tracker, S, prm, oprm, 1):
        for col in INT_COLS:
            assert bits_equal(got[col], ref[col]), col
        assert not (got["status"] & 1).any()  # every window inside the frame
    prm = pk.track_params(z, rescue_level=0, subpixel=1, search_method=method,
                          subpix=pk.default_subpix(pk.SUBPIX_LM))
    oprm = pk.track_params(z, rescue_level=0, subpixel=1, subpix=pk.default_subpix(pk.SUBPIX_LM, subpix_tol=1e-13))
    for k, got, ref in _sequence_vs_oracle(tracker, S, prm, oprm, 1):
        assert np.abs(got["dx"] - ref["dx"]).max() < TOL_PX and np.abs(got["dy"] - ref["dy"]).max() < TOL_PX
        assert np.abs(got["drot"] - ref["drot"]).max() < TOL_RAD
        assert np.abs(got["NCC_subpix"] - ref["NCC_subpix"]).max() < TOL_NCC


def test_sequence_exact_lm_within_tolerance_of_fp64_oracle(tracker):
    # the batching workload's shape at a small size: radii 6-60, +-8 px x 9 steps, LM
    # subpixel, three pairs; every grain within 1e-3 px / 1e-5 rad of the FP64 minimizer
    S = synth.sequence(n_frames=4, n_grains=40, size=(1000, 800), seed=508)
    z = S.params.search_zone
    say(f"sequence LM: G={S.G}")
    prm = pk.track_params(z, rescue_level=0, subpixel=1, search_method=pk.SEARCH_EXACT,
                          subpix=pk.default_subpix(pk.SUBPIX_LM))
    oprm = pk.track_params(z, rescue_level=0, subpixel=1, subpix=pk.default_subpix(pk.SUBPIX_LM, subpix_tol=1e-13))
    for k, got, ref in _sequence_vs_oracle(tracker, S, prm, oprm, 3):
        epx = np.hypot(got["dx"] - ref["dx"], got["dy"] - ref["dy"])
        assert epx.max() < TOL_PX, (k, epx.max())
        assert np.abs(got["drot"] - ref["drot"]).max() < TOL_RAD
        assert np.abs(got["NCC_subpix"] - ref["NCC_subpix"]).max() < TOL_NCC


def test_exact_overflowing_grains_go_to_plain_kernel(tracker):
    # at a reference rotation of 0.35 rad the rounded rotated disks of radius 40+ differ
    # from the middle step's in hundreds of cells per step: those grains are searched by
    # the plain kernel within the stage, the small ones by the MFMA kernel; every pick is
    # the reference's
    S = synth.sequence(n_frames=2, n_grains=24, size=(900, 800), seed=509, radius=(8.0, 56.0))
    gr = S.grains.copy()
    gr["refrot"] = 0.35
    z = S.params.search_zone
    say(f"exact search at refrot 0.35: G={S.G}")
    tracker.set_image(0, S.frames[0], S.bits)
    tracker.set_image(1, S.frames[1], S.bits)
    tracker.set_image_indices(0, 1)
    tracker.set_grains(gr, S.pattern_xy, S.pattern_begin)
    n0 = tracker.exact_fallbacks()
    tracker.track_pair(pk.track_params(z, rescue_level=0, subpixel=0, search_method=pk.SEARCH_EXACT))
    got = tracker.get_grains()
    n1 = tracker.exact_fallbacks()
    assert n1[0] == n0[0]
    assert 0 < n1[1] - n0[1] < S.G
    W = ptref.Workload(S.frames[0], S.frames[1], gr, S.pattern_xy, S.pattern_begin, S.bits)
    ref, _, _ = ptref.track_pair(W, pk.track_params(z, rescue_level=0, subpixel=0), nthreads=8)
    for col in ("dx", "dy", "drot", "upix", "vpix", "rot_inc"):
        assert bits_equal(got[col], ref[col]), col
    assert np.abs(got["NCC"] - ref["NCC"]).max() < 1e-12


@pytest.mark.parametrize("bits", [12, 8])
def test_exact_hands_values_beyond_the_declared_depth_to_plain_kernel(tracker, bits):
    # frames declared 12-bit (8-bit) whose right half holds 16x larger values (data stored
    # shifted up in wider words): the exact kernels find window values above 4095 (255)
    # while staging and send those grains to the plain kernel (disks through the prefix-sum
    # base mask, squares through the base-mask GEMM); every pick is the reference's
    w = synth.workload(2, n_grains=24, size=(900, 700), seed=313)
    I0, I1 = w.I0.copy(), w.I1.copy()
    if bits == 8:
        I0 >>= 4
        I1 >>= 4
    half = I0.shape[1] // 2
    I0[:, half:] *= 16
    I1[:, half:] *= 16
    pats = []
    for g in range(w.G):
        P = w.pattern_xy[w.pattern_begin[g]:w.pattern_begin[g + 1]]
        if g % 2:
            r = math.isqrt(int((P.astype(np.int64) ** 2).sum(1).max())) // 2
            ys, xs = np.mgrid[-r:r + 1, -r:r + 1]
            P = np.stack([xs.ravel(), ys.ravel()], 1)
        pats.append(P.astype(np.int32))
    pxy = np.concatenate(pats).astype(np.int32)
    pb = np.concatenate([[0], np.cumsum([len(p) for p in pats])]).astype(np.int64)
    right = w.grains["refcoord_xpix"] - 20 - w.truth["radii"] > half
    left = w.grains["refcoord_xpix"] + 80 < half
    assert right.any() and left.any()
    z = w.params.search_zone
    say(f"exact search, {bits}-bit frames with {bits + 4}-bit values in the right half: G={w.G}")
    tracker.set_image(0, I0, bits)
    tracker.set_image(1, I1, bits)
    tracker.set_image_indices(0, 1)
    tracker.set_grains(w.grains, pxy, pb)
    n0 = tracker.exact_fallbacks()
    tracker.track_pair(pk.track_params(z, rescue_level=0, subpixel=0, search_method=pk.SEARCH_EXACT))
    got = tracker.get_grains()
    n1 = tracker.exact_fallbacks()
    assert n1[0] == n0[0]
    assert int(right.sum()) <= n1[1] - n0[1] <= w.G - int(left.sum())
    W = ptref.Workload(I0, I1, w.grains, pxy, pb, 16)
    ref, _, _ = ptref.track_pair(W, pk.track_params(z, rescue_level=0, subpixel=0), nthreads=8)
    for col in ("dx", "dy", "drot", "upix", "vpix", "rot_inc"):
        assert bits_equal(got[col], ref[col]), col
    assert np.abs(got["NCC"] - ref["NCC"]).max() < 1e-12


@pytest.mark.parametrize("method", [pk.SEARCH_PLAIN, pk.SEARCH_EXACT])
def test_huge_zone_tiled_plain_search_picks_equal_oracle(tracker, method):
    # +-150 px: the window (341 x 341 cells) exceeds the LDS, so the plain kernel walks the
    # shifts in tiles; the exact form hands the whole search to it
    S = synth.sequence(n_frames=2, n_grains=3, size=(900, 900), seed=510, radius=(15.0, 20.0), margin=200,
                       max_step=3.0, max_rot=0.01)
    z = pk.zone(150, num_rot=0, inc_rot=1.0)
    say(f"+-150 px zone, method {method}")
    tracker.set_image(0, S.frames[0], S.bits)
    tracker.set_image(1, S.frames[1], S.bits)
    tracker.set_image_indices(0, 1)
    tracker.set_grains(S.grains, S.pattern_xy, S.pattern_begin)
    tracker.track_pair(pk.track_params(z, rescue_level=0, subpixel=0, search_method=method))
    got = tracker.get_grains()
    W = ptref.Workload(S.frames[0], S.frames[1], S.grains, S.pattern_xy, S.pattern_begin, S.bits)
    ref, _, _ = ptref.track_pair(W, pk.track_params(z, rescue_level=0, subpixel=0), nthreads=8)
    for col in INT_COLS:
        assert bits_equal(got[col], ref[col]), col
    assert (got["NCC"] > 0.8).all()  # found (the grains also rotate by up to 0.01 rad)


def _full_disk(P):
    """True when the offsets are every cell of x^2 + y^2 <= max once (the library then takes
    the pattern itself as the base mask of the Delta lists)."""
    d2 = (P.astype(np.int64) ** 2).sum(1)
    m2 = int(d2.max())
    r = math.isqrt(m2)
    cells = sum(2 * math.isqrt(m2 - y * y) + 1 for y in range(-r, r + 1))
    return cells == len(P) and len({(int(x), int(y)) for x, y in P}) == len(P)


def test_exact_base_mask_for_disks_and_other_patterns(tracker):
    # at a reference rotation of 0.2 rad: full digital disks (base mask = the pattern),
    # squares, rings, a disk with a cell missing and one with a cell counted twice (base mask
    # = the middle step's rounded rotation); every pick is the reference's
    S = synth.sequence(n_frames=2, n_grains=20, size=(800, 700), seed=511, radius=(10.0, 24.0))
    pats, kinds = [], []
    for g in range(S.G):
        P = S.pattern_xy[S.pattern_begin[g]:S.pattern_begin[g + 1]]
        r = math.isqrt(int((P.astype(np.int64) ** 2).sum(1).max()))
        k = g % 5
        if k == 1:
            ys, xs = np.mgrid[-(r // 2):r // 2 + 1, -(r // 2):r // 2 + 1]
            P = np.stack([xs.ravel(), ys.ravel()], 1)
        elif k == 2:
            d2 = (P.astype(np.int64) ** 2).sum(1)
            P = P[d2 > (r // 2) ** 2]
        elif k == 3:
            P = P[1:]
        elif k == 4:
            P = np.concatenate([P, P[len(P) // 2:len(P) // 2 + 1]])
        pats.append(P.astype(np.int32))
        kinds.append(_full_disk(P))
    assert any(kinds) and not all(kinds)
    pxy = np.concatenate(pats).astype(np.int32)
    pb = np.concatenate([[0], np.cumsum([len(p) for p in pats])]).astype(np.int64)
    gr = S.grains.copy()
    gr["refrot"] = 0.2
    z = S.params.search_zone
    say(f"exact search, base masks, refrot 0.2: G={S.G}")
    tracker.set_image(0, S.frames[0], S.bits)
    tracker.set_image(1, S.frames[1], S.bits)
    tracker.set_image_indices(0, 1)
    tracker.set_grains(gr, pxy, pb)
    tracker.track_pair(pk.track_params(z, rescue_level=0, subpixel=0, search_method=pk.SEARCH_EXACT))
    got = tracker.get_grains()
    W = ptref.Workload(S.frames[0], S.frames[1], gr, pxy, pb, S.bits)
    ref, _, _ = ptref.track_pair(W, pk.track_params(z, rescue_level=0, subpixel=0), nthreads=8)
    for col in ("dx", "dy", "drot", "upix", "vpix", "rot_inc"):
        assert bits_equal(got[col], ref[col]), col
    assert np.abs(got["NCC"] - ref["NCC"]).max() < 1e-12


def test_exact_and_plain_agree_on_clamped_windows(tracker):
    # grains moved against the frame edges: their windows (and patterns) read beyond the
    # frame, which both forms clamp to the edge pixel and report as PTG_STATUS_CLAMPED; the
    # exact form (disk and square base masks) picks as the plain one
    w = synth.workload(2, n_grains=24, size=(700, 600), seed=2077)
    gr = w.grains.copy()
    H, W = w.I0.shape
    edge = np.zeros(w.G, dtype=bool)
    for g in range(0, w.G, 3):
        r = int(gr["radius_pix"][g])
        side = (g // 3) % 4
        if side == 0:
            gr["refcoord_xpix"][g] = r // 2
        elif side == 1:
            gr["refcoord_ypix"][g] = H - 1 - r // 2
        elif side == 2:
            gr["refcoord_xpix"][g] = W - 1 - r // 3
        else:
            gr["refcoord_ypix"][g] = 2
        edge[g] = True
    pats = []
    for g in range(w.G):
        P = w.pattern_xy[w.pattern_begin[g]:w.pattern_begin[g + 1]]
        if g % 2 == 1:
            h = math.isqrt(int((P.astype(np.int64) ** 2).sum(1).max())) // 2
            ys, xs = np.mgrid[-h:h + 1, -h:h + 1]
            P = np.stack([xs.ravel(), ys.ravel()], 1)
        pats.append(P.astype(np.int32))
    pxy = np.concatenate(pats).astype(np.int32)
    pb = np.concatenate([[0], np.cumsum([len(p) for p in pats])]).astype(np.int64)
    z = pk.zone(12, num_rot=3, inc_rot=0.02)
    say(f"clamped windows, exact vs plain: G={w.G}")
    res = {}
    for m in (pk.SEARCH_PLAIN, pk.SEARCH_EXACT):
        tracker.set_image(0, w.I0, w.bits)
        tracker.set_image(1, w.I1, w.bits)
        tracker.set_image_indices(0, 1)
        tracker.set_grains(gr, pxy, pb)
        tracker.track_pair(pk.track_params(z, rescue_level=0, subpixel=0, search_method=m))
        res[m] = tracker.get_grains()
    a, b = res[pk.SEARCH_EXACT], res[pk.SEARCH_PLAIN]
    for col in ("dx", "dy", "drot", "upix", "vpix", "rot_inc"):
        assert bits_equal(a[col], b[col]), col
    assert np.abs(a["NCC"] - b["NCC"]).max() < 1e-12
    clamped = (a["status"] & 1) != 0
    assert clamped[edge].all() and not clamped[~edge].any()
    assert ((b["status"] & 1) != 0).tolist() == clamped.tolist()


def _same(a, b):
    """Bit-equal, or both NaN (the NaN's sign differs between the host's and the GPU's 0/0)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(both_nan | (a.view(np.uint64) == b.view(np.uint64))))


def test_zero_grains_is_a_no_op(tracker, cfg2s):
    # the reference's loops over an empty grain list do nothing; so does every stage here
    say("zero grains")
    tracker.set_image(0, cfg2s.I0, cfg2s.bits)
    tracker.set_image(1, cfg2s.I1, cfg2s.bits)
    tracker.set_image_indices(0, 1)
    tracker.set_grains(cfg2s.grains[:0], np.zeros((0, 2), np.int32), np.zeros(1, np.int64))
    for m in (pk.SEARCH_PLAIN, pk.SEARCH_EXACT):
        tracker.track_pair(pk.track_params(cfg2s.params.search_zone, rescue_level=2, subpixel=1, search_method=m,
                                           subpix=pk.default_subpix(pk.SUBPIX_LM)))
    assert len(tracker.get_grains()) == 0
    tracker.change_reference()


@pytest.mark.parametrize("method", [pk.SEARCH_PLAIN, pk.SEARCH_EXACT])
def test_degenerate_grains_follow_the_reference(tracker, method):
    # a flat reference patch (C0C0 = 0), a flat current window (C1C1 = 0 for every
    # candidate), a one-pixel pattern and a masked grain beside ordinary ones: an undefined
    # ZNCC never wins (best_NCC stays 0, the grain is queued for the rescues and keeps its
    # position) and the subpixel stage leaves an undefined objective alone, as in the reference
    w = synth.workload(2, n_grains=12, size=(900, 800), seed=2090)
    I0, I1 = w.I0.copy(), w.I1.copy()
    z, zr = pk.zone(8, num_rot=2, inc_rot=0.02), pk.zone(12, num_rot=2, inc_rot=0.02)
    gr = w.grains.copy()
    x, y, r = int(gr["refcoord_xpix"][0]), int(gr["refcoord_ypix"][0]), int(gr["radius_pix"][0]) + 2
    I0[y - r:y + r + 1, x - r:x + r + 1] = 1000
    x, y, r = int(gr["refcoord_xpix"][1]), int(gr["refcoord_ypix"][1]), int(gr["radius_pix"][1]) + 35
    I1[y - r:y + r + 1, x - r:x + r + 1] = 1500
    gr["masked"][3] = 1
    pats = [w.pattern_xy[w.pattern_begin[g]:w.pattern_begin[g + 1]] for g in range(w.G)]
    pats[2] = np.zeros((1, 2), np.int32)
    pxy = np.concatenate(pats).astype(np.int32)
    pb = np.concatenate([[0], np.cumsum([len(p) for p in pats])]).astype(np.int64)
    say(f"degenerate grains, method {method}")
    tracker.set_image(0, I0, w.bits)
    tracker.set_image(1, I1, w.bits)
    tracker.set_image_indices(0, 1)
    for sub in (pk.SUBPIX_POWELL, pk.SUBPIX_LM):
        tracker.set_grains(gr, pxy, pb)
        prm = pk.track_params(z, zr, zr, NCC_min=0.7, NCC_min_super=0.5, rescue_level=2, subpixel=1,
                              search_method=method, subpix=pk.default_subpix(sub))
        tracker.track_pair(prm)
        got = tracker.get_grains()
        oprm = prm if sub == pk.SUBPIX_POWELL else pk.track_params(
            z, zr, zr, NCC_min=0.7, NCC_min_super=0.5, rescue_level=2, subpixel=1,
            subpix=pk.default_subpix(pk.SUBPIX_LM, subpix_tol=1e-13))
        ref, counts, _ = ptref.track_pair(ptref.Workload(I0, I1, gr, pxy, pb, w.bits), oprm, nthreads=4)
        for g in (0, 1, 2, 3):
            for col in ("dx", "dy", "drot", "NCC", "NCC_rescue", "NCC_subpix"):
                assert _same(got[col][g], ref[col][g]), (g, col, got[col][g], ref[col][g])
        assert got["NCC"][0] == 0.0 and got["NCC"][2] == 0.0  # best_NCC never left its start
        # every grain's integer stages pick as the reference; the subpixel results of the
        # well-tracked grains away from the flattened patches agree to the tolerance (a
        # window partly flat, or a grain lost beyond the rescue zones, leaves a landscape
        # where the two optimizers may settle in different minima)
        for col in ("NCC", "NCC_rescue"):
            if method == pk.SEARCH_PLAIN:  # the reference's arithmetic: bit for bit
                assert _same(got[col], ref[col]), col
            else:  # exact integer moments: the same picks, NCC within FP64 rounding
                assert np.abs(got[col] - ref[col]).max() < 1e-12, col
        cx, cy = gr["refcoord_xpix"].astype(float), gr["refcoord_ypix"].astype(float)
        far = np.ones(w.G, dtype=bool)
        for g, pad in ((0, 2), (1, 35)):
            reach = gr["radius_pix"] + gr["radius_pix"][g] + pad + 12 + 16
            far &= np.maximum(np.abs(cx - cx[g]), np.abs(cy - cy[g])) > reach
        found = (ref["NCC"] >= 0.7) | (ref["NCC_rescue"] > 0.5)  # the integer search placed the grain
        ok = np.nonzero(far & found & (ref["NCC_subpix"] > 0.9))[0]
        assert len(ok) > 0
        assert np.abs(got["dx"][ok] - ref["dx"][ok]).max() < TOL_PX
        assert np.abs(got["dy"][ok] - ref["dy"][ok]).max() < TOL_PX
