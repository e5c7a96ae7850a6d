"""Frame copies that overlap the context's work (ptg_set_image_async) against the
synchronous copy (ptg_set_image).

The reference reads each frame into image[i] between pairs (src/p-tracker.cpp:426, 446). With
ptg_set_image_async the next frame goes into a third slot while the running pair tracks the
other two, and the first call that reads the slot waits for its copy on the GPU. Checked:
  - a frame sequence with reference swaps, each next frame copied from pinned memory during the
    previous pair, gives the records of the same sequence with synchronous copies, bit for bit;
  - ordering: a read of the slot (ptg_get_image), a synchronous copy into it, a remap from it
    and a reallocation of it (a larger frame) all come after the overlapped copy.
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


def test_overlapped_copies_give_the_synchronous_records():
    S = synth.sequence(n_frames=7, n_grains=60, size=(900, 800), seed=717, radius=(8.0, 30.0), margin=120)
    prm = pk.track_params(S.params.search_zone, S.params.search_zone, S.params.search_zone, NCC_min=0.0,
                          rescue_level=0, subpixel=1, search_method=pk.SEARCH_EXACT,
                          subpix=pk.default_subpix(pk.SUBPIX_LM))
    npairs = len(S.frames) - 1

    say(f"sequence {npairs} pairs, {S.G} grains: synchronous copies")
    ref = []
    with pk.Tracker(0) as t:
        t.set_image(0, S.frames[0], S.bits)
        t.set_image(1, S.frames[1], S.bits)
        t.set_image_indices(0, 1)
        t.set_grains(S.grains, S.pattern_xy, S.pattern_begin)
        for k in range(npairs):
            if k > 0:
                t.change_reference()
                t.set_image((k + 1) % 2, S.frames[k + 1], S.bits)
            t.track_pair(prm)
            ref.append(t.get_grains())

    say("overlapped copies: frame k + 2 into the free slot while pair (k, k + 1) runs")
    got = []
    H, W = S.frames[0].shape
    with pk.Tracker(0) as t:
        pinned = [t.host_frame(H, W) for _ in range(3)]
        slot_of = {0: 0, 1: 1}
        for f in (0, 1):
            pinned[f][...] = S.frames[f]
            t.set_image_async(f, pinned[f], S.bits)
        t.set_image_indices(0, 1)
        t.set_grains(S.grains, S.pattern_xy, S.pattern_begin)
        for k in range(npairs):
            if k > 0:
                t.change_reference()
                t.set_image_indices(slot_of[k], slot_of[k + 1])
            if k + 2 < len(S.frames):
                spare = ({0, 1, 2} - {slot_of[k], slot_of[k + 1]}).pop()
                t.wait_image(spare)  # its pinned buffer is rewritten below
                pinned[spare][...] = S.frames[k + 2]
                t.set_image_async(spare, pinned[spare], S.bits)
                slot_of[k + 2] = spare
            t.track_pair(prm)
            got.append(t.get_grains())
    for k, (a, b) in enumerate(zip(got, ref)):
        assert a.tobytes() == b.tobytes(), f"pair {k}"


def test_overlapped_copy_orders_later_calls():
    rng = np.random.default_rng(5)
    H, W = 600, 700
    a = rng.integers(0, 4096, (H, W), dtype=np.uint16)
    b = rng.integers(0, 4096, (H, W), dtype=np.uint16)
    with pk.Tracker(0) as t:
        pinned = t.host_frame(H, W)
        pinned[...] = a
        say("copy from pinned memory, then read back")
        t.set_image_async(2, pinned, 12)
        assert np.array_equal(t.get_image(2, W, H), a)
        say("pageable source changed as soon as the call returns")
        src = b.copy()
        t.set_image_async(2, src, 12)
        src[...] = 0
        assert np.array_equal(t.get_image(2, W, H), b)
        say("pageable source, then a synchronous copy into the same slot")
        t.set_image_async(2, np.ascontiguousarray(b), 12)
        t.set_image(2, a, 12)
        assert np.array_equal(t.get_image(2, W, H), a)
        say("a remap from a slot whose copy is queued")
        t.set_image_async(3, np.ascontiguousarray(b), 12)
        ident = np.zeros(14)
        ident[[0, 1, 8, 9]] = -1.0  # no correction; centres at the frame centre
        t.remap(3, 1, ident)
        assert np.array_equal(t.get_image(1, W, H), b)
        say("a larger frame into a slot with a queued copy (the slot's memory is replaced)")
        big = rng.integers(0, 4096, (H + 100, W + 100), dtype=np.uint16)
        t.set_image_async(3, np.ascontiguousarray(a), 12)
        t.set_image_async(3, big, 12)
        assert np.array_equal(t.get_image(3, W + 100, H + 100), big)
        t.wait_image(3)
        t.sync()


def test_pitched_frames_copy_row_by_row():
    # a host frame with padded rows (pitch > dimx) takes the pitched copy; contiguous rows one
    # linear copy: both give the frame, synchronous or overlapped, and back
    rng = np.random.default_rng(8)
    H, W, pad = 300, 410, 37
    host = rng.integers(0, 65536, (H, W + pad), dtype=np.uint16)
    with pk.Tracker(0) as t:
        for name, fn in (("ptg_set_image", t.lib.ptg_set_image), ("ptg_set_image_async", t.lib.ptg_set_image_async)):
            say(f"{name}, pitch {W + pad} for {W} columns")
            t._ck(fn(t.h, 1, host.ctypes.data_as(pk.c_void_p), W, H, W + pad, 16), name)
            assert np.array_equal(t.get_image(1, W, H), host[:, :W])
            back = np.zeros((H, W + 5), np.uint16)
            t._ck(t.lib.ptg_get_image(t.h, 1, back.ctypes.data_as(pk.c_void_p), W + 5), "ptg_get_image")
            assert np.array_equal(back[:, :W], host[:, :W]) and not back[:, W:].any()
        t.sync()
