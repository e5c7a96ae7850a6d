"""GPU test of ptg_ncc_objective_batch_noalloc, the objective entry point that allocates
nothing: through the C ABI, two calls in a row (the second smaller) give the values of
ptg_ncc_objective_batch bit for bit (the same kernel on the same points); sizes above the
static staging array and missing reference moments are refused with an error."""
import ctypes
import sys

import numpy as np
import pytest

import ptref

pk = ptref.package()
from ptracker_amd import synth  # noqa: E402

pytestmark = pytest.mark.gpu
MAX_EVALS = 16384


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _call(lib, h, fn, ig, X, dmax=3.0, stiff=0.0):
    f = np.full(len(ig), np.nan)
    rc = getattr(lib, fn)(h, len(ig), _p(ig), _p(X), dmax, stiff, _p(f))
    return rc, f


def test_noalloc_objective_equals_allocating_batch():
    w = synth.workload(2, n_grains=40, size=(900, 700), seed=77)
    rng = np.random.default_rng(11)
    with pk.Tracker(0) as t:
        lib, h = t.lib, t.h
        t.set_image(0, w.I0, w.bits)
        t.set_image(1, w.I1, w.bits)
        t.set_image_indices(0, 1)
        t.set_grains(w.grains, w.pattern_xy, w.pattern_begin)
        ig0 = np.arange(3, dtype=np.int32)
        X0 = np.zeros((3, 3))
        rc, _ = _call(lib, h, "ptg_ncc_objective_batch_noalloc", ig0, X0)
        assert rc == pk.PTG_ERR_STATE  # moments not computed yet: refused, nothing staged
        t.precompute()
        for n in (3000, 517):  # two calls in a row, the second with a smaller input
            print(f"noalloc objective, {n} points", flush=True)
            sys.stdout.flush()
            ig = rng.integers(0, w.G, n).astype(np.int32)
            X = np.ascontiguousarray(np.stack([rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
                                               rng.uniform(-0.02, 0.02, n)], axis=1))
            rc, got = _call(lib, h, "ptg_ncc_objective_batch_noalloc", ig, X)
            assert rc == pk.PTG_OK, lib.ptg_last_error(h)
            rc, ref = _call(lib, h, "ptg_ncc_objective_batch", ig, X)
            assert rc == pk.PTG_OK
            assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
            assert np.all(np.isfinite(got))
        big = np.zeros(MAX_EVALS + 1, dtype=np.int32)
        rc, _ = _call(lib, h, "ptg_ncc_objective_batch_noalloc", big, np.zeros((MAX_EVALS + 1, 3)))
        assert rc == pk.PTG_ERR_LIMIT
        assert b"at most 16384" in lib.ptg_last_error(h)
