"""Device rotation tables (the default) against the host's.

The reference forms cos/sin(refrot + drot + step) with the host C library
(src/p-tracker.cpp:636-637, 687-689) and rounds the rotated offsets (:615-621). The device
tables use the device library's cosine and sine. Each (grain, step) with a rotated coordinate
near a rounding boundary is put on a guard list, which the host checks against its own values
after the call. The tests here:
  - whole pairs (rescue and super-rescue, both search forms, a sequence with reference swaps,
    so refrot is not zero) give the same records bit for bit on both tables, and with the
    rescue queues kept on the device or read back by the host;
  - with every (grain, step) on the guard list, the host finds no offset that differs, and no
    device cosine or sine more than 2^-50 from its own (the bound the guard's margin assumes);
  - a forced redo (the host's tables after the device's) gives the same records again.
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


def load(t, w):
    t.set_image(0, w.I0, w.bits)
    t.set_image(1, w.I1, w.bits)
    t.set_image_indices(0, 1)
    t.set_grains(w.grains, w.pattern_xy, w.pattern_begin)


@pytest.fixture(scope="module")
def rescue_case():
    # motion of up to 6 px against a +-2 px main zone: both rescue stages run
    w = synth.workload(2, n_grains=24, size=(800, 700), max_disp=6.0, seed=77)
    z, zr, zs = pk.zone(2, num_rot=2, inc_rot=0.02), pk.zone(4, num_rot=2, inc_rot=0.02), pk.zone(8, num_rot=1, inc_rot=0.05)
    return w, z, zr, zs


def _pair(w, prm, opts):
    with pk.Tracker(0) as t:
        for k, v in opts.items():
            t.set_option(k, v)
        load(t, w)
        t.track_pair(prm)
        out = t.get_grains()
        stats = t.rot_guard_stats()
    return out, stats


@pytest.mark.parametrize("method", [pk.SEARCH_PLAIN, pk.SEARCH_EXACT])
def test_device_tables_equal_host_tables(rescue_case, method):
    w, z, zr, zs = rescue_case
    prm = pk.track_params(z, zr, zs, NCC_min=0.7, NCC_min_super=0.6, rescue_level=2, subpixel=1,
                          search_method=method, subpix=pk.default_subpix(pk.SUBPIX_LM))
    say(f"rotation tables, method {method}: host tables")
    host, _ = _pair(w, prm, {pk.OPT_ROT_TABLES: 1})
    say("device tables")
    dev, (checked, redone, _) = _pair(w, prm, {})
    assert same_records(dev, host)
    assert redone == 0
    say("device tables, rescue queues kept on the device")
    dq, _ = _pair(w, prm, {pk.OPT_DEVICE_QUEUES: 1})
    assert same_records(dq, host)
    say("device tables, every (grain, step) checked by the host")
    allg, (checked, redone, max_diff) = _pair(w, prm, {pk.OPT_ROT_GUARD_TEST: 1})
    say(f"  checked {checked}, redone {redone}, max |device - host| {max_diff:.3g}")
    assert same_records(allg, host)
    # the precomputation's G entries and every step of every stage's grains
    assert checked >= w.G * (1 + 5)
    assert redone == 0
    assert max_diff <= 2.0 ** -50
    say("device tables, forced redo on the host's")
    red, (checked, redone, _) = _pair(w, prm, {pk.OPT_ROT_GUARD_TEST: 2})
    assert redone == 1
    assert same_records(red, host)


def test_device_tables_sequence_with_reference_swaps():
    # refrot accumulates over the swaps, so the tables' angles are not the steps alone
    S = synth.sequence(n_frames=4, n_grains=40, size=(900, 800), seed=515, radius=(8.0, 30.0), margin=120)
    H, W = S.frames[0].shape
    prm = pk.track_params(S.params.search_zone, S.params.search_zone, S.params.search_zone, NCC_min=0.0,
                          rescue_level=0, subpixel=1, search_method=pk.SEARCH_EXACT,
                          subpix=pk.default_subpix(pk.SUBPIX_LM))

    def run(opts):
        outs = []
        with pk.Tracker(0) as t:
            for k, v in opts.items():
                t.set_option(k, v)
            t.set_image(0, S.frames[0], S.bits)
            t.set_image(1, S.frames[1], S.bits)
            t.set_image_indices(0, 1)
            t.set_grains(S.grains, S.pattern_xy, S.pattern_begin)
            for k in range(3):
                if k > 0:
                    t.change_reference()
                    t.set_image((k + 1) % 2, S.frames[k + 1], S.bits)
                t.track_pair(prm)
                outs.append(t.get_grains())
            stats = t.rot_guard_stats()
        return outs, stats

    say("sequence 3 pairs, host tables")
    host, _ = run({pk.OPT_ROT_TABLES: 1})
    say("sequence 3 pairs, device tables, every (grain, step) checked")
    dev, (checked, redone, max_diff) = run({pk.OPT_ROT_GUARD_TEST: 1})
    say(f"  checked {checked}, redone {redone}, max |device - host| {max_diff:.3g}")
    assert any(np.any(o["refrot"] != 0.0) for o in host[1:])
    for a, b in zip(dev, host):
        assert same_records(a, b)
    assert redone == 0 and max_diff <= 2.0 ** -50
