"""A pair as one HIP graph (PTG_OPT_GRAPH) against the same pairs run as launches.

With the option on, the first pair with a new key runs as launches, the next is captured and
later pairs replay the graph on whatever frames the context holds. Every replay is checked
against the records of the same pair run as launches, bit for bit, with the program's own
eager work and synchronisations between the replays:
  - a frame sequence with reference swaps (a new current frame uploaded before each pair), the
    exact search and the LM subpixel stage;
  - one pair with rescue and super-rescue stages replayed several times from the same input
    records, both search forms (the rescue queues then stay on the device).
"""
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


def same_records(a, b):
    return a.tobytes() == b.tobytes()


def test_graph_sequence_equals_launches():
    S = synth.sequence(n_frames=7, n_grains=60, size=(900, 800), seed=616, radius=(8.0, 30.0), margin=120)
    prm = pk.track_params(S.params.search_zone, S.params.search_zone, S.params.search_zone, NCC_min=0.0,
                          rescue_level=0, subpixel=1, search_method=pk.SEARCH_EXACT,
                          subpix=pk.default_subpix(pk.SUBPIX_LM))
    npairs = len(S.frames) - 1

    def run(graph):
        outs = []
        with pk.Tracker(0) as t:
            t.set_option(pk.OPT_GRAPH, 1 if graph else 0)
            t.set_image(0, S.frames[0], S.bits)
            t.set_image(1, S.frames[1], S.bits)
            t.set_image_indices(0, 1)
            t.set_grains(S.grains, S.pattern_xy, S.pattern_begin)
            for k in range(npairs):
                if k > 0:
                    t.change_reference()
                    t.set_image((k + 1) % 2, S.frames[k + 1], S.bits)
                t.track_pair(prm)
                outs.append(t.get_grains())
            stats = t.graph_stats()
        return outs, stats

    say(f"sequence {npairs} pairs, {S.G} grains: launches")
    ref, _ = run(False)
    say("sequence: graph")
    got, (captures, launches, nodes) = run(True)
    say(f"  captures {captures}, graph launches {launches}, nodes {nodes}")
    assert captures == 1 and launches >= 2 and nodes > 0
    for k, (a, b) in enumerate(zip(got, ref)):
        assert same_records(a, b), f"pair {k}"


@pytest.mark.parametrize("method", [pk.SEARCH_PLAIN, pk.SEARCH_EXACT])
def test_graph_rescue_pair_replays_equal_launches(method):
    # motion of up to 6 px against a +-2 px main zone: both rescue stages have grains
    w = synth.workload(2, n_grains=24, size=(800, 700), max_disp=6.0, seed=77)
    z, zr, zs = pk.zone(2, num_rot=2, inc_rot=0.02), pk.zone(4, num_rot=2, inc_rot=0.02), pk.zone(8, num_rot=1, inc_rot=0.05)
    prm = pk.track_params(z, zr, zs, NCC_min=0.7, NCC_min_super=0.6, rescue_level=2, subpixel=1,
                          search_method=method, subpix=pk.default_subpix(pk.SUBPIX_LM))
    with pk.Tracker(0) as t:
        t.set_image(0, w.I0, w.bits)
        t.set_image(1, w.I1, w.bits)
        t.set_image_indices(0, 1)
        t.set_grains(w.grains, w.pattern_xy, w.pattern_begin)
        t.track_pair(prm)
        ref = t.get_grains()
        counts = t.queue_counts()
        say(f"method {method}: launches, queues {counts}")
        assert counts[0] > 0 and counts[1] > 0
        t.set_option(pk.OPT_GRAPH, 1)
        for r in range(5):
            t.put_grains(w.grains)
            t.track_pair(prm)
            got = t.get_grains()
            assert same_records(got, ref), f"run {r}"
            assert list(t.queue_counts()) == list(counts)
            # eager work between replays: an objective batch on the records just tracked
            f = t.objective(np.arange(4, dtype=np.int32), np.zeros((4, 3)))
            assert np.all(np.isfinite(f))
        captures, launches, nodes = t.graph_stats()
        say(f"  captures {captures}, graph launches {launches}, nodes {nodes}")
        assert captures == 1 and launches >= 3


def test_graph_sequence_10000_grains_equals_oracle():
    # the batching workload at its full size (6000x4000, 10,000 disk grains of radius 6-60 px,
    # +-8 px x 9 steps, LM subpixel) over 4 pairs with reference swaps: the first two run as
    # launches (the key settles), the third is captured, the fourth replays it; a sample of every
    # pair's grains against the oracle run from the GPU's own input records of that pair
    S = synth.sequence(n_frames=5, n_grains=10000, seed=1005)
    z = S.params.search_zone
    prm = pk.track_params(z, z, z, NCC_min=0.0, rescue_level=0, subpixel=1, search_method=pk.SEARCH_EXACT,
                          subpix=pk.default_subpix(pk.SUBPIX_LM))
    oprm = pk.track_params(z, z, z, NCC_min=0.0, rescue_level=0, subpixel=1, search_method=pk.SEARCH_PLAIN,
                           subpix=pk.default_subpix(pk.SUBPIX_LM, subpix_tol=1e-13))
    sample = np.sort(np.random.default_rng(0).choice(S.G, size=48, replace=False))
    outs = []
    with pk.Tracker(0) as t:
        t.set_option(pk.OPT_GRAPH, 1)
        t.set_image(0, S.frames[0], S.bits)
        t.set_image(1, S.frames[1], S.bits)
        t.set_image_indices(0, 1)
        t.set_grains(S.grains, S.pattern_xy, S.pattern_begin)
        for k in range(4):
            if k > 0:
                t.change_reference()
                t.set_image((k + 1) % 2, S.frames[k + 1], S.bits)
            t.track_pair(prm)
            outs.append(t.get_grains())
            say(f"pair {k}: graph stats {t.graph_stats()}")
        captures, launches, nodes = t.graph_stats()
    assert captures == 1 and launches == 2
    # Integer picks: the NCC of the search equal to the oracle's (within the exact form's ulps).
    # Subpixel: within 1e-3 px / 1e-5 rad / NCC_subpix 1e-6 of the oracle's FP64 minimizer for
    # grains the search matched well (NCC >= 0.7). A grain matched poorly (heavily overlapped by
    # larger grains drawn over it) can have an objective with several minima, and two minimizers
    # may end in different ones: there the GPU's minimum must be at least as good as the oracle's.
    gin = S.grains
    other_minima = 0
    for k in range(4):
        W = ptref.Workload(S.frames[k], S.frames[k + 1], gin, S.pattern_xy, S.pattern_begin, S.bits)
        ref, _, _ = ptref.track_pair(W, oprm, nthreads=8, subset=sample)
        got = outs[k][sample]
        assert np.abs(got["NCC"] - ref["NCC"]).max() < 1e-12
        epx = np.hypot(got["dx"] - ref["dx"], got["dy"] - ref["dy"])
        erad = np.abs(got["drot"] - ref["drot"])
        encc = np.abs(got["NCC_subpix"] - ref["NCC_subpix"])
        close = (epx < 1e-3) & (erad < 1e-5) & (encc < 1e-6)
        good = ref["NCC"] >= 0.7
        say(f"pair {k}: {good.sum()} well-matched grains within {epx[good].max():.2e} px, {erad[good].max():.2e} rad; "
            f"{int((~close).sum())} grains in another minimum")
        assert close[good].all()
        assert (got["NCC_subpix"][~close] >= ref["NCC_subpix"][~close] - 1e-6).all()
        other_minima += int((~close).sum())
        gin = ptref.change_reference(outs[k])
    assert other_minima <= 2
