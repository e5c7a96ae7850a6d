"""GPU parity on exact ties: the near-tie fixture (tests/golden/near_tie_oracle.npz, two
candidates per grain whose exact NCC are equal while the reference's FP64 values order them
either way, or are equal) searched by both forms. Picks and NCC must be the oracle's bit for
bit; the exact form resolves the ties by re-scoring its near-tie sets in the reference's
arithmetic, without handing a grain to the plain kernel."""
import os
import sys

import numpy as np
import pytest

import ptref

pk = ptref.package()
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
import make_near_tie_fixture as ntf  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", [pk.SEARCH_PLAIN, pk.SEARCH_EXACT])
def test_near_ties_resolve_as_the_reference(tracker, method):
    fx = dict(np.load(os.path.join(GOLDEN, "near_tie_oracle.npz")))
    specs = [tuple(int(v) for v in row) for row in fx["specs"]]
    I0, I1 = ntf.near_tie_pair(specs)
    pxy, pb = ntf.patterns()
    assert ntf.digest(I0, I1, pxy, pb) == str(fx["digest"])
    print(f"near-tie fixture, method {method}", flush=True)
    tracker.set_image(0, I0, 12)
    tracker.set_image(1, I1, 12)
    tracker.set_image_indices(0, 1)
    tracker.set_grains(ntf.grains(), pxy, pb)
    n0 = tracker.exact_fallbacks()
    tracker.track_pair(pk.track_params(ntf.zone(), rescue_level=0, subpixel=0, search_method=method))
    got = tracker.get_grains()
    n1 = tracker.exact_fallbacks()
    for c in ntf.COLS:
        assert np.array_equal(got[c].view(np.uint64), fx[c].view(np.uint64)), (c, got[c], fx[c])
    assert n1 == n0
