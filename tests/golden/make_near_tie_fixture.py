"""Writes near_tie_oracle.npz: four grains whose two best candidates tie exactly in exact
arithmetic while the reference's FP64 two-pass NCC orders them either way.

For each grain the current frame holds two copies of the grain's reference patch (plus the same
small noise), 20 px apart (shifts u = -10 and u = +10, v = 0; the reference visits u = -10 first,
src/p-tracker.cpp:695-696), one copy mapped by the gray-level change v -> a v + c with a > 0.
ZNCC is invariant under such a change, so the two candidates tie exactly; with a = 3 the
reference's means, differences and sums round differently for the two copies, so its NCC values
differ in the last bits (an integer offset alone mostly rounds alike). Per grain:
  grain 0 (disk r = 8)      a = 3 and c chosen so that the later copy's NCC is larger: the
                            reference picks u = +10;
  grain 1 (square 13 x 13)  the same, through the exact search's base-mask GEMM path;
  grain 2 (disk r = 8)      a = 1, c = 0: identical copies, equal NCC: the earlier copy wins;
  grain 3 (square 13 x 13)  a = 3 and c chosen so that the earlier copy's NCC is larger: u = -10.
The frames are rebuilt from near_tie_pair(specs) (seeded); the file stores the specs, a SHA-256 of
the frames and patterns, the oracle's main-search columns and both candidates' NCC per grain.

Run from the repository root:  python tests/golden/make_near_tie_fixture.py
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import ptref  # noqa: E402

pk = ptref.package()

W, H = 320, 120
CENTRES = [(40, 60), (120, 60), (200, 60), (280, 60)]
SQUARE = (False, True, False, True)
WANT = ("later", "later", "equal", "earlier")  # the copy whose reference NCC is larger
HALF = 9     # half-width of the copied patch
OFF = 10     # the copies sit at u = -OFF and u = +OFF
NOISE = 15   # the copies' common noise, uniform in [-NOISE, NOISE]
COLS = ("dx", "dy", "drot", "upix", "vpix", "rot_inc", "NCC", "NCC_rescue")


def zone():
    return pk.zone(20, num_rot=1, inc_rot=0.01)


def patterns():
    pats = []
    for sq in SQUARE:
        if sq:
            ys, xs = np.mgrid[-6:7, -6:7]
            pats.append(np.stack([xs.ravel(), ys.ravel()], 1).astype(np.int32))
        else:
            pats.append(ptref.circ_pattern(8).astype(np.int32))
    pxy = np.concatenate(pats).astype(np.int32)
    pb = np.concatenate([[0], np.cumsum([len(p) for p in pats])]).astype(np.int64)
    return pxy, pb


def grains():
    g = np.zeros(len(CENTRES), dtype=pk.GRAIN_DTYPE)
    g["refcoord_xpix"] = [c[0] for c in CENTRES]
    g["refcoord_ypix"] = [c[1] for c in CENTRES]
    g["radius_pix"] = 8.0
    return g


def near_tie_pair(specs):
    """The 12-bit frames (I0, I1); specs[g] = (a, c, which): copy `which` (0: u = -OFF, 1: u = +OFF)
    of grain g's patch is mapped by v -> a v + c."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(4321)

    def texture():
        t = gaussian_filter(rng.normal(size=(H, W)), 1.5)
        return np.clip(np.round(500 + t / t.std() * 150), 0, 1000)

    I0 = texture()
    I1 = texture()
    for (cx, cy), (a, c, which) in zip(CENTRES, specs):
        patch = I0[cy - HALF:cy + HALF + 1, cx - HALF:cx + HALF + 1]
        noisy = patch + rng.integers(-NOISE, NOISE + 1, size=patch.shape)
        for q, u in enumerate((-OFF, OFF)):
            I1[cy - HALF:cy + HALF + 1, cx + u - HALF:cx + u + HALF + 1] = a * noisy + c if q == which else noisy
    assert I1.max() <= 4095 and I1.min() >= 0
    return I0.astype(np.uint16), I1.astype(np.uint16)


def digest(I0, I1, pxy, pb):
    h = hashlib.sha256()
    for a in (I0, I1, pxy, pb):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def pair_ncc(I0, I1, g):
    """The reference's NCC of grain g's two copies (step 0: u = -OFF and u = +OFF, v = 0)."""
    pxy, pb = patterns()
    gr = grains()
    w = ptref.Workload(I0, I1, gr, pxy, pb, 12)
    gp, p0 = ptref.precompute(w)
    z = zone()
    v = ptref.candidate_ncc(w, gp, p0, g, z)
    V = z.up + z.down + 1
    at = lambda u: v[(u + z.left) * V + z.up]  # noqa: E731
    return at(-OFF), at(OFF), v


def choose_specs():
    specs = [(1, 0, 1)] * len(CENTRES)
    for g, want in enumerate(WANT):
        if want == "equal":
            continue
        found = None
        for which in (1, 0):
            for c in range(0, 900):
                trial = list(specs)
                trial[g] = (3, c, which)
                try:
                    I0, I1 = near_tie_pair(trial)
                except AssertionError:
                    break
                a, b, _ = pair_ncc(I0, I1, g)
                if (want == "later" and b > a) or (want == "earlier" and a > b):
                    found = (3, c, which)
                    break
            if found:
                break
        if not found:
            raise RuntimeError(f"no gray-level change found for grain {g}")
        specs[g] = found
    return specs


def main():
    specs = choose_specs()
    I0, I1 = near_tie_pair(specs)
    pxy, pb = patterns()
    gr = grains()
    w = ptref.Workload(I0, I1, gr, pxy, pb, 12)
    prm = pk.track_params(zone(), rescue_level=0, subpixel=0)
    got, _, _ = ptref.track_pair(w, prm, nthreads=4)
    out = {"digest": np.array(digest(I0, I1, pxy, pb)), "specs": np.asarray(specs, dtype=np.int32)}
    for c in COLS:
        out[c] = np.asarray(got[c], dtype=np.float64)
    pairs = np.array([pair_ncc(I0, I1, g)[:2] for g in range(len(CENTRES))])
    out["pair_ncc"] = pairs
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "near_tie_oracle.npz")
    np.savez(path, **out)
    print(f"specs {specs}; picks upix {got['upix']}; pair NCC\n{pairs!r}")


if __name__ == "__main__":
    main()
