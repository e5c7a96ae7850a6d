"""The exact search's autotuner (ptg_autotune_search): it times launch shapes per radius
bucket and keeps the fastest. The grains and the fallback counters are left as they were,
and searches with the tuned shapes give the same picks and NCC as with the built-in shapes,
and as the oracle. The radius buckets' launch order (PTG_OPT_BUCKET_ORDER, which the autotuner
also picks) leaves every record as it was."""
import sys

import numpy as np
import pytest

import ptref

pk = ptref.package()
from ptracker_amd import synth  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits_equal(a, b):
    return np.array_equal(np.asarray(a, dtype=np.float64).view(np.uint64), np.asarray(b, dtype=np.float64).view(np.uint64))


def test_autotuned_shapes_keep_the_picks(tracker):
    w = synth.workload(2, n_grains=48, size=(1100, 900), seed=2024)
    z = w.params.search_zone
    prm = pk.track_params(z, rescue_level=0, subpixel=0, search_method=pk.SEARCH_EXACT)
    tracker.set_image(0, w.I0, w.bits)
    tracker.set_image(1, w.I1, w.bits)
    tracker.set_image_indices(0, 1)
    tracker.set_grains(w.grains, w.pattern_xy, w.pattern_begin)
    tracker.clear_tuning()
    tracker.track_pair(prm)
    heuristic = tracker.get_grains()
    tracker.put_grains(w.grains)
    tracker.precompute()  # mean0, C0C0 (the autotuner's searches compute them too)
    g0 = tracker.get_grains()
    f0 = tracker.exact_fallbacks()
    say("autotuning the exact search's launch shapes, 48 grains, +-20 px x 21 steps")
    n, before, after = tracker.autotune_search(z)
    say(f"  {n} searches; the search's GPU span {before:.3f} ms with the heuristic's shapes, {after:.3f} ms tuned")
    assert n > 10 and after > 0.0 and before > 0.0
    g1 = tracker.get_grains()
    for col in g0.dtype.names:
        assert np.array_equal(g0[col], g1[col]), col
    assert tracker.exact_fallbacks() == f0
    tracker.track_pair(prm)
    tuned = tracker.get_grains()
    assert tracker.exact_fallbacks() == f0
    ref, _, _ = ptref.track_pair(ptref.Workload(w.I0, w.I1, w.grains, w.pattern_xy, w.pattern_begin, w.bits),
                                 pk.track_params(z, rescue_level=0, subpixel=0), nthreads=8)
    for col in ("dx", "dy", "drot", "upix", "vpix", "rot_inc", "NCC"):
        assert bits_equal(tuned[col], heuristic[col]), col
    for col in ("dx", "dy", "drot", "upix", "vpix", "rot_inc"):
        assert bits_equal(tuned[col], ref[col]), col
    assert np.abs(tuned["NCC"] - ref["NCC"]).max() < 1e-12
    tracker.clear_tuning()
    tracker.set_option(pk.OPT_BUCKET_ORDER, 0)


@pytest.mark.parametrize("devq", [0, 1])
def test_bucket_orders_give_the_same_records(devq):
    # grains over five radius buckets, 60 of them moved beyond the +-20 px zone into the
    # rescue stage (its queue read back, or kept on the device and served by the full plan)
    w = synth.workload(2, n_grains=400, size=(2400, 2000), seed=31)
    rng = np.random.default_rng(3)
    moved = rng.choice(w.G, size=60, replace=False)
    g = w.grains.copy()
    g["dx"][moved] += rng.choice([-1.0, 1.0], 60) * rng.uniform(22.0, 30.0, 60)
    prm = pk.track_params(w.params.search_zone, w.params.search_zone_rescue, w.params.search_zone_super_rescue,
                          NCC_min=0.7, rescue_level=1, subpixel=1, search_method=pk.SEARCH_EXACT,
                          subpix=pk.default_subpix(pk.SUBPIX_LM))
    out = {}
    for order in (0, 1):
        with pk.Tracker(0) as t:
            t.set_option(pk.OPT_BUCKET_ORDER, order)
            t.set_option(pk.OPT_DEVICE_QUEUES, devq)
            t.set_image(0, w.I0, w.bits)
            t.set_image(1, w.I1, w.bits)
            t.set_image_indices(0, 1)
            t.set_grains(g, w.pattern_xy, w.pattern_begin)
            say(f"bucket order {order}, device queues {devq}")
            t.track_pair(prm)
            out[order] = t.get_grains()
            counts = t.queue_counts()
    assert counts[0] > 0
    assert out[0].tobytes() == out[1].tobytes()
