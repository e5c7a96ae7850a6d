"""GPU parity at full size: the 6000 x 4000 12-bit frame pair with 2,000 grains (+-20 px x 21
rotation steps, rescue zone +-40 px), the workload the benchmark runs, against the oracle's
results stored in tests/golden/cfg2_full_oracle.npz (written by make_cfg2_full_fixture.py there).

Bars as in the reduced-size tests: integer picks bit-identical for both search forms, NCC and
NCC_rescue bit-identical for the plain form and within 1e-12 for the exact one; the subpixel
(u, v) within 1e-3 px, theta within 1e-5 rad, NCC_subpix within 1e-6 of the oracle (for the LM
form, of the objective's optimum; for the Powell form, of the oracle's Powell)."""
import os
import sys

import numpy as np
import pytest

import ptref

pk = ptref.package()
from ptracker_amd import synth  # noqa: E402

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from make_cfg2_full_fixture import digest  # noqa: E402

pytestmark = pytest.mark.gpu

TOL_PX, TOL_RAD, TOL_NCC = 1e-3, 1e-5, 1e-6
FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cfg2_full_oracle.npz")


def say(msg):
    print(msg, flush=True)


def bits_equal(a, b):
    return np.array_equal(np.asarray(a, dtype=np.float64).view(np.uint64), np.asarray(b, dtype=np.float64).view(np.uint64))


@pytest.fixture(scope="module")
def full():
    say("generating the 2,000-grain 6000 x 4000 pair")
    fx = dict(np.load(FIXTURE))
    w = synth.workload(2)
    assert digest(w) == str(fx["digest"]), "the generated pair differs from the one the fixture was made from"
    return w, fx


def run(tracker, w, method, subpixel, subpix_method=pk.SUBPIX_LM):
    p = w.params
    prm = pk.track_params(p.search_zone, p.search_zone_rescue, p.search_zone_super_rescue, NCC_min=0.7,
                          rescue_level=1, subpixel=subpixel, search_method=method,
                          subpix=pk.default_subpix(subpix_method))
    tracker.set_image(0, w.I0, w.bits)
    tracker.set_image(1, w.I1, w.bits)
    tracker.set_image_indices(0, 1)
    tracker.set_grains(w.grains, w.pattern_xy, w.pattern_begin)
    tracker.track_pair(prm)
    return tracker.get_grains()


@pytest.mark.parametrize("method,roles", [("plain", 1), ("exact", 1), ("exact", 2)])
def test_full_size_search_and_rescue_equal_oracle(tracker, full, method, roles):
    # roles: the exact MFMA kernel's waves multiply and score their own columns (1, the default)
    # or producer waves hand each pass to consumer waves (2)
    w, fx = full
    say(f"full size, {method} search + rescue, wave roles {roles}")
    tracker.set_option(pk.OPT_SEARCH_ROLES, roles)
    try:
        got = run(tracker, w, pk.SEARCH_EXACT if method == "exact" else pk.SEARCH_PLAIN, 0)
    finally:
        tracker.set_option(pk.OPT_SEARCH_ROLES, 1)
    assert list(tracker.queue_counts()) == list(fx["counts"])
    for c in ("dx", "dy", "drot", "upix", "vpix", "rot_inc"):
        bad = np.flatnonzero(np.ascontiguousarray(got[c]).view(np.uint64) != fx["search_" + c].view(np.uint64))
        assert bad.size == 0, (c, bad[:8])
    for c in ("NCC", "NCC_rescue"):
        if method == "plain":
            assert bits_equal(got[c], fx["search_" + c]), c
        else:
            assert np.abs(got[c] - fx["search_" + c]).max() < 1e-12, c


@pytest.mark.parametrize("form", ["exact_lm", "plain_powell"])
def test_full_size_subpixel_within_tolerance_of_oracle(tracker, full, form):
    w, fx = full
    say(f"full size, {form}")
    # the LM form against the objective's optimum; the Powell form against the oracle's Powell
    # (its tolerance is on the objective's value, so it stops short of the optimum by up to
    # a few 1e-5 rad)
    if form == "exact_lm":
        got, key = run(tracker, w, pk.SEARCH_EXACT, 1, pk.SUBPIX_LM), "sub_"
    else:
        got, key = run(tracker, w, pk.SEARCH_PLAIN, 1, pk.SUBPIX_POWELL), "powell_"
    du = np.abs(got["dx"] - fx[key + "dx"]).max()
    dv = np.abs(got["dy"] - fx[key + "dy"]).max()
    dr = np.abs(got["drot"] - fx[key + "drot"]).max()
    dn = np.abs(got["NCC_subpix"] - fx[key + "NCC_subpix"]).max()
    say(f"  max |du| {du:.3g} px, |dv| {dv:.3g} px, |dtheta| {dr:.3g} rad, |dNCC| {dn:.3g}")
    assert du < TOL_PX and dv < TOL_PX and dr < TOL_RAD and dn < TOL_NCC


@pytest.mark.parametrize("bits", [12, 8])
@pytest.mark.parametrize("method", ["plain", "exact"])
def test_full_size_40_steps_search_equal_oracle(tracker, method, bits):
    # the 2,000-grain pair searched over +-20 px x 40 rotation steps (num_rot 20 at inc 0.005),
    # on 12-bit frames and on their 8-bit requantisation
    from make_cfg3_full_fixture import COLS, params, workload

    fx = np.load(os.path.join(os.path.dirname(FIXTURE), "cfg3_full_oracle.npz"))
    say(f"full size, 40 steps, {bits}-bit, {method} search")
    w = workload(bits)
    assert digest(w) == str(fx[f"digest_{bits}"]), "the generated pair differs from the one the fixture was made from"
    prm = params(w)
    prm.search_method = pk.SEARCH_EXACT if method == "exact" else pk.SEARCH_PLAIN
    tracker.set_image(0, w.I0, w.bits)
    tracker.set_image(1, w.I1, w.bits)
    tracker.set_image_indices(0, 1)
    tracker.set_grains(w.grains, w.pattern_xy, w.pattern_begin)
    tracker.track_pair(prm)
    got = tracker.get_grains()
    for c in COLS[:3]:
        bad = np.flatnonzero(np.ascontiguousarray(got[c]).view(np.uint64) != fx[f"{c}_{bits}"].view(np.uint64))
        assert bad.size == 0, (c, bad[:8])
    if method == "plain":
        assert bits_equal(got["NCC"], fx[f"NCC_{bits}"])
    else:
        assert np.abs(got["NCC"] - fx[f"NCC_{bits}"]).max() < 1e-12


@pytest.fixture(scope="module")
def rescue_variant():
    from make_cfg2_rescue_fixture import FAR

    say("generating the rescue variant of the 2,000-grain pair (300 grains moved 21-35 px)")
    fx = dict(np.load(os.path.join(os.path.dirname(FIXTURE), "cfg2_rescue_oracle.npz")))
    w = synth.workload(2, far=FAR)
    assert digest(w) == str(fx["digest"]), "the generated pair differs from the one the fixture was made from"
    return w, fx


@pytest.mark.parametrize("method,devq", [("plain", 0), ("exact", 0), ("exact", 1)])
def test_full_size_rescue_zone_equals_oracle(tracker, rescue_variant, method, devq):
    # the rescue stage at its own geometry: +-40 px x 21 steps (137,781 candidates per grain)
    # for the grains the +-20 px main search misses; picks bit-identical to the oracle's, NCC
    # and NCC_rescue bit-identical for the plain form and within 1e-12 for the exact one; the
    # rescue queue read back by the host, or kept on the device (devq)
    w, fx = rescue_variant
    n0 = tracker.exact_fallbacks()
    tracker.set_option(pk.OPT_DEVICE_QUEUES, devq)
    try:
        got = run(tracker, w, pk.SEARCH_EXACT if method == "exact" else pk.SEARCH_PLAIN, 0)
    finally:
        tracker.set_option(pk.OPT_DEVICE_QUEUES, 0)
    n1 = tracker.exact_fallbacks()
    counts = list(tracker.queue_counts())
    say(f"full size rescue variant, {method}, device queue {devq}: queues {counts}, exact-form fallbacks (searches, grains) "
        f"{n1[0] - n0[0]}, {n1[1] - n0[1]}")
    assert counts == list(fx["counts"]) and counts[0] >= 250
    for c in ("dx", "dy", "drot", "upix", "vpix", "rot_inc"):
        bad = np.flatnonzero(np.ascontiguousarray(got[c]).view(np.uint64) != fx[c].view(np.uint64))
        assert bad.size == 0, (c, bad[:8])
    for c in ("NCC", "NCC_rescue"):
        if method == "plain":
            assert bits_equal(got[c], fx[c]), c
        else:
            assert np.abs(got[c] - fx[c]).max() < 1e-12, c
    if method == "exact":
        assert n1[0] == n0[0]  # neither stage left the exact form
