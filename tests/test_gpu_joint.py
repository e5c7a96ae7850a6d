"""GPU checks of the differentiation form's joint refinement (ptg_joint_eval,
ptg_refine_joint_distortion): the objective against the FP64 oracle (ptref_joint_objective, whose
value at K1 = K2 = 0 is the reference's NCC_to_minimize_xyR), its analytic gradient over
(Tx, Ty, theta, K1, K2) against central differences (h = 1e-3 px, 1e-5 rad, and the lens steps
that move a pixel at the frame corner by 1e-3 px) and against MINPACK's forward-difference
Jacobian of the oracle, and the joint Levenberg-Marquardt on config 4 (the config-2 frames seen
through k1 = -2e-9, k2 = 1e-16): lens and scene positions against the ground truth, and the
solution a stationary point of the oracle's objective."""
import math
import os
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


def load_and_track(tracker, w):
    tracker.set_image(0, w.I0, w.bits)
    tracker.set_image(1, w.I1, w.bits)
    tracker.set_image_indices(0, 1)
    tracker.set_grains(w.grains, w.pattern_xy, w.pattern_begin)
    z = w.params
    prm = pk.track_params(z.search_zone, z.search_zone_rescue, z.search_zone_super_rescue, NCC_min=0.7,
                          rescue_level=1, subpixel=1, search_method=pk.SEARCH_EXACT,
                          subpix=pk.default_subpix(pk.SUBPIX_LM))
    tracker.track_pair(prm)
    return tracker.get_grains()


@pytest.fixture(scope="module")
def small():
    # 1600 x 1200, the lens 10x / 100x stronger than config 4's
    return synth.workload(4, n_grains=60, size=(1600, 1200), seed=404, lens=(-2e-8, 1e-14))


def _points(w, got, rng, k):
    ok = np.flatnonzero(got["NCC_subpix"] > 0.8)[:k]
    P = np.stack([got["dx"][ok], got["dy"][ok], got["drot"][ok]], 1)
    P = P + np.stack([rng.uniform(-0.3, 0.3, len(ok)), rng.uniform(-0.3, 0.3, len(ok)),
                      rng.uniform(-0.01, 0.01, len(ok))], 1)
    return ok.astype(np.int32), np.ascontiguousarray(P)


def test_joint_objective_matches_oracle(tracker, small):
    w = small
    say("joint objective vs oracle, 1600x1200")
    got = load_and_track(tracker, w)
    rng = np.random.default_rng(5)
    ig, P = _points(w, got, rng, 16)
    W = ptref.Workload(w.I0, w.I1, w.grains, w.pattern_xy, w.pattern_begin, w.bits)
    gp, p0 = ptref.precompute(W)
    X = w.truth["lens"]
    for lens in ([0.0, 0.0, X[0], X[1]], [X[2], X[3], X[0], X[1]], [0.5 * X[2], 2.0 * X[3], X[0], X[1]]):
        f = tracker.joint_eval(ig, P, lens, derivs=False)
        fo = ptref.joint_objective_batch(W, gp, p0, ig, P, lens)
        say(f"  lens {lens[:2]}: max |f - f_oracle| {np.abs(f - fo).max():.2e}")
        assert np.abs(f - fo).max() < 1e-9
    # without a lens the joint objective is the reference's objective at X = P - (dx, dy, drot)
    f0 = tracker.joint_eval(ig, P, [0.0, 0.0, X[0], X[1]], derivs=False)
    for k, g in enumerate(ig):
        r = gp[g]
        Xr = P[k] - np.array([r["dx"], r["dy"], r["drot"]])
        assert abs(f0[k] - ptref.ncc_objective(W, gp, p0, int(g), Xr)) < 1e-9


def test_joint_gradient_against_differences(tracker, small):
    w = small
    say("joint gradient vs central differences and MINPACK forward differences")
    got = load_and_track(tracker, w)
    rng = np.random.default_rng(6)
    ig, P = _points(w, got, rng, 8)
    X = w.truth["lens"]
    lens = np.array([X[2], X[3], X[0], X[1]])
    f, g, h = tracker.joint_eval(ig, P, lens)
    rc = math.hypot(X[0], X[1])
    steps = [1e-3, 1e-3, 1e-5, 1e-3 / rc ** 3, 1e-3 / rc ** 5]  # 1e-3 px at the corner for the lens

    def f_at(q, val):
        Pq, Lq = P.copy(), lens.copy()
        if q < 3:
            Pq[:, q] += val
        else:
            Lq[q - 3] += val
        return tracker.joint_eval(ig, Pq, Lq, derivs=False)

    for q in range(5):
        cd = (f_at(q, steps[q]) - f_at(q, -steps[q])) / (2 * steps[q])
        scale = np.abs(g[:, q]).max()
        err = np.abs(cd - g[:, q]).max()
        say(f"  d/dp{q}: max |analytic - central| {err:.3e} of max |grad| {scale:.3e}")
        assert err < 2e-3 * scale + 1e-12
    # MINPACK's forward differences (step sqrt(eps) |x|, the reference's lmdif Jacobian) of the
    # oracle's objective, grain by grain
    W = ptref.Workload(w.I0, w.I1, w.grains, w.pattern_xy, w.pattern_begin, w.bits)
    gp, p0 = ptref.precompute(W)
    for k in range(3):
        x0 = np.concatenate([P[k], lens[:2]])

        def fun(x):
            return np.array([ptref.joint_objective(W, gp, p0, int(ig[k]), x[:3], [x[3], x[4], lens[2], lens[3]])])

        J = ptref.lmdif_jacobian(fun, x0)[0]
        rel = np.abs(J - g[k]) / (np.abs(g[k]) + 1e-6 * np.abs(g[k]).max())
        say(f"  grain {ig[k]}: MINPACK forward differences vs analytic, relative {rel.max():.2e}")
        assert np.all(np.abs(J - g[k]) < 2e-2 * np.abs(g[k]) + 1e-4 * np.abs(g[k]).max())
    # the Gauss-Newton matrix is symmetric positive semi-definite
    for k in range(len(ig)):
        M = np.zeros((5, 5))
        t = 0
        for a in range(5):
            for b in range(a, 5):
                M[a, b] = M[b, a] = h[k, t]
                t += 1
        s = np.sqrt(np.diag(M))
        assert np.linalg.eigvalsh(M / np.outer(s, s)).min() > -1e-9


def test_config4_joint_refinement_against_truth_and_oracle(tracker):
    say("config 4: 6000x4000 pair through k1 = -2e-9, k2 = 1e-16, 2000 grains")
    w = synth.workload(4)
    X = w.truth["lens"]
    g0 = load_and_track(tracker, w)
    k1, k2, fi, fe, it, ng = tracker.refine_joint_distortion(0.0, 0.0, X[0], X[1], 0.7, 200)
    g1 = tracker.get_grains()
    say(f"  k1 {k1:.4e} (truth {X[2]:.1e}), k2 {k2:.4e} (truth {X[3]:.1e}); objective {fi:.6f} -> {fe:.6f} in "
        f"{it} iterations over {ng} grains")
    assert ng > 1900 and fe < fi
    assert abs(k1 / X[2] - 1) < 0.2 and abs(k2 / X[3] - 1) < 0.2

    # scene displacement of every grain's centre against the truth, with and without the lens
    def U(x, y, a, b):
        dx, dy = x - X[0], y - X[1]
        r2 = dx * dx + dy * dy
        K = a * r2 + b * r2 * r2
        return x + dx * K, y + dy * K

    ok = g1["NCC_subpix"] > 0.7
    x0, y0 = w.grains["refcoord_xpix"] + w.grains["dx"], w.grains["refcoord_ypix"] + w.grains["dy"]
    errs = {}
    for name, gg, kk in (("no lens", g0, (0.0, 0.0)), ("joint", g1, (k1, k2))):
        wrx, wry = U(gg["refcoord_xpix"].astype(float), gg["refcoord_ypix"].astype(float), *kk)
        wcx, wcy = U(x0, y0, *kk)
        c, s = np.cos(gg["drot"]), np.sin(gg["drot"])
        ax, ay = wcx - wrx, wcy - wry
        mx = wrx + gg["dx"] + c * ax + s * ay - wcx
        my = wry + gg["dy"] - s * ax + c * ay - wcy
        e = np.hypot(mx - w.truth["disp"][:, 0], my - w.truth["disp"][:, 1])[ok]
        errs[name] = e
        say(f"  {name}: scene displacement error median {np.median(e):.4f} px, p99 {np.percentile(e, 99):.4f} px")
    assert np.median(errs["joint"]) < 0.8 * np.median(errs["no lens"])

    # a stationary point of the FP64 oracle's objective: along each lens parameter the oracle's
    # sum over the grains is smallest at the solution (the vertex of the parabola through
    # +-2% lies within 10% of the step), and for a sample of grains along each of its own
    # parameters likewise
    W = ptref.Workload(w.I0, w.I1, w.grains, w.pattern_xy, w.pattern_begin, w.bits)
    gp, p0 = ptref.precompute(W)
    lst = np.flatnonzero(~g1["masked"].astype(bool) & (g1["NCC_subpix"] >= 0.7)).astype(np.int32)
    P = np.ascontiguousarray(np.stack([g1["dx"][lst], g1["dy"][lst], g1["drot"][lst]], 1))
    threads = min(16, os.cpu_count() or 1)
    L = np.array([k1, k2, X[0], X[1]])
    F0 = ptref.joint_objective_batch(W, gp, p0, lst, P, L, threads).sum()
    for q, v in ((0, k1), (1, k2)):
        d = 0.02 * v
        Lp, Lm = L.copy(), L.copy()
        Lp[q] += d
        Lm[q] -= d
        Fp = ptref.joint_objective_batch(W, gp, p0, lst, P, Lp, threads).sum()
        Fm = ptref.joint_objective_batch(W, gp, p0, lst, P, Lm, threads).sum()
        off = 0.5 * (Fm - Fp) / (Fm - 2 * F0 + Fp)
        say(f"  oracle along K{q + 1}: F(-) - F0 {Fm - F0:.3e}, F(+) - F0 {Fp - F0:.3e}, vertex at {off:+.3f} steps")
        assert Fm > F0 and Fp > F0 and abs(off) < 0.1
    sample = np.arange(0, len(lst), 97)
    hs = (0.01, 0.01, 1e-3)
    for q in range(3):
        Pp, Pm = P[sample].copy(), P[sample].copy()
        Pp[:, q] += hs[q]
        Pm[:, q] -= hs[q]
        f0 = ptref.joint_objective_batch(W, gp, p0, lst[sample], P[sample], L, threads)
        fp = ptref.joint_objective_batch(W, gp, p0, lst[sample], Pp, L, threads)
        fm = ptref.joint_objective_batch(W, gp, p0, lst[sample], Pm, L, threads)
        off = 0.5 * (fm - fp) / (fm - 2 * f0 + fp)
        say(f"  oracle along p{q}: largest vertex offset {np.abs(off).max():.3f} steps over {len(sample)} grains")
        assert np.all(np.abs(off) < 0.1)
