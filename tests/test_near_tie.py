"""CPU checks behind the exact search's near-tie handling: the near-tie fixture (two candidates
per grain that tie in exact arithmetic while the reference's FP64 NCC orders them either way,
tests/golden/make_near_tie_fixture.py) is what the oracle computes, and the error bound the
exact search's window is built on holds for the reference's arithmetic against exact values."""
import math
import os
import sys
from decimal import Decimal, getcontext

import numpy as np
import pytest

import ptref

pk = ptref.package()
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
import make_near_tie_fixture as ntf  # noqa: E402

U = 2.0 ** -53


def load_fixture():
    fx = dict(np.load(os.path.join(GOLDEN, "near_tie_oracle.npz")))
    specs = [tuple(int(v) for v in row) for row in fx["specs"]]
    I0, I1 = ntf.near_tie_pair(specs)
    pxy, pb = ntf.patterns()
    assert ntf.digest(I0, I1, pxy, pb) == str(fx["digest"])
    return fx, I0, I1, pxy, pb


def exact_moments(I0, I1, gr, pat, u, v):
    """num = N A - Sa B, d0 = N Saa - Sa^2, d1 = N C - B^2 as Python integers for the unrotated
    pattern at shift (u, v)."""
    x0, y0 = int(gr["refcoord_xpix"]), int(gr["refcoord_ypix"])
    a = [int(I0[y0 + y, x0 + x]) for x, y in pat]
    b = [int(I1[y0 + v + y, x0 + u + x]) for x, y in pat]
    N = len(pat)
    Sa, Saa, B, C, A = sum(a), sum(t * t for t in a), sum(b), sum(t * t for t in b), sum(p * q for p, q in zip(a, b))
    return N * A - Sa * B, N * Saa - Sa * Sa, N * C - B * B, N, Sa, B


def test_near_tie_fixture_is_the_oracles():
    fx, I0, I1, pxy, pb = load_fixture()
    gr = ntf.grains()
    w = ptref.Workload(I0, I1, gr, pxy, pb, 12)
    got, _, _ = ptref.track_pair(w, pk.track_params(ntf.zone(), rescue_level=0, subpixel=0), nthreads=2)
    for c in ntf.COLS:
        assert np.array_equal(got[c].view(np.uint64), fx[c].view(np.uint64)), c
    assert list(got["upix"]) == [10.0, 10.0, -10.0, -10.0] and list(got["vpix"]) == [0.0] * 4
    for g in range(len(gr)):
        a, b, _ = ntf.pair_ncc(I0, I1, g)
        assert (a, b) == tuple(fx["pair_ncc"][g])
        want = ntf.WANT[g]
        assert (b > a) if want == "later" else (a > b) if want == "earlier" else (a == b)
        # the two copies tie exactly: equal num^2 / d1 with num > 0, so an argmax in exact
        # arithmetic (ties to the earlier candidate) would pick u = -10 on every grain
        pat = pxy[pb[g]:pb[g + 1]]
        n1, d0, d11, *_ = exact_moments(I0, I1, gr[g], pat, -ntf.OFF, 0)
        n2, _, d12, *_ = exact_moments(I0, I1, gr[g], pat, ntf.OFF, 0)
        assert n1 > 0 and n2 > 0 and n1 * n1 * d12 == n2 * n2 * d11


def ncc_error_bound(N, R0, R1):
    # the bound the exact search's window uses (its derivation is with the kernel's
    # near_window_lo); here it is checked against exact arithmetic
    gN = N * U / (1.0 - N * U)
    sn = math.sqrt(N)
    return 2.0 * (2.0 * U * (R0 + R1) + 7.0 * U + 2.0 * gN) + 8.0 * U * U * (R0 + sn) * (R1 + sn)


@pytest.mark.parametrize("case", ["textured", "high_mean_low_variance", "large_pattern", "sixteen_bit"])
def test_reference_ncc_error_within_the_search_window_bound(case):
    # every candidate's reference NCC (the oracle's two-pass FP64 sums) against its exact value
    # num / sqrt(d0 d1): within half the bound (the bound is doubled for margin), and the FP64
    # value the exact search forms from the exact integers within 4u (6u for 16-bit frames,
    # whose num and d1 reach 2^61 and are rounded to FP64 first)
    getcontext().prec = 60
    rng = np.random.default_rng({"textured": 1, "high_mean_low_variance": 2, "large_pattern": 3,
                                 "sixteen_bit": 4}[case])
    H, W = 90, 90
    bits = 12
    if case == "textured":
        r, I0 = 6, rng.integers(0, 4096, size=(H, W))
        I1 = np.clip(np.roll(I0, (1, 2), (0, 1)) + rng.integers(-40, 41, size=(H, W)), 0, 4095)
    elif case == "high_mean_low_variance":
        r, I0 = 6, 4000 + rng.integers(0, 3, size=(H, W))
        I1 = 4000 + rng.integers(0, 3, size=(H, W))
    elif case == "large_pattern":
        r, I0 = 25, 3000 + rng.integers(0, 50, size=(H, W))
        I1 = np.clip(np.roll(I0, (2, -1), (0, 1)) + rng.integers(-20, 21, size=(H, W)), 0, 4095)
    else:
        r, bits, I0 = 12, 16, rng.integers(0, 65536, size=(H, W))
        I1 = np.clip(np.roll(I0, (1, -2), (0, 1)) + rng.integers(-600, 601, size=(H, W)), 0, 65535)
    I0, I1 = I0.astype(np.uint16), I1.astype(np.uint16)
    pat = ptref.circ_pattern(r).astype(np.int32)
    gr = np.zeros(1, dtype=pk.GRAIN_DTYPE)
    gr["refcoord_xpix"], gr["refcoord_ypix"] = W // 2, H // 2
    w = ptref.Workload(I0, I1, gr, pat, np.array([0, len(pat)], dtype=np.int64), bits)
    gp, p0 = ptref.precompute(w)
    z = pk.zone(8, num_rot=0, inc_rot=1.0)
    fp = ptref.candidate_ncc(w, gp, p0, 0, z)
    k, worst = 0, 0.0
    for u in range(-8, 9):
        for v in range(-8, 9):
            num, d0, d1, N, Sa, B = exact_moments(I0, I1, gr[0], pat, u, v)
            if d1 == 0:
                assert math.isnan(fp[k])
                k += 1
                continue
            rho = Decimal(num) / (Decimal(d0) * Decimal(d1)).sqrt()
            R0, R1 = Sa / math.sqrt(d0), B / math.sqrt(d1)
            err = abs(Decimal(fp[k]) - rho)
            tau = ncc_error_bound(N, R0, R1)
            assert err <= Decimal(tau / 2), (u, v, float(err), tau)
            worst = max(worst, float(err) / tau)
            q = float(num) / math.sqrt(float(d0) * float(d1))
            assert abs(Decimal(q) - rho) <= Decimal((6 if bits == 16 else 4) * U) * abs(rho) + Decimal(1e-300)
            k += 1
    print(f"{case}: largest error / bound {worst:.3g}")
    assert worst > 0.0  # the reference's rounding is visible, and the bound covers it
