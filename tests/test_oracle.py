"""Pins the oracle (CPU restatement of the reference path) against facts it did not
produce: counts and values measured from the reference's loops, the Gauss
circle counts, the reference's own compiled sampler (oracle/_ref), exact rational
brute force of the ZNCC argmax, closed forms and invariants. A dropped term, a wrong
sign or index, or a transposed operand in the oracle fails at least one of these."""
import ctypes
import math
from fractions import Fraction

import numpy as np
import pytest

import ptref

pk = ptref.package()


# ---- rotation steps, src/p-tracker.cpp:684-686 -----------------------------------------
def test_rotation_steps_config2():
    s = ptref.rotation_steps(0.01, 10)
    assert len(s) == 21
    assert s[-1] == 0.09999999999999998
    assert s[0] == -0.1


def test_rotation_steps_config3_has_40_not_41():
    s = ptref.rotation_steps(0.005, 20)
    assert len(s) == 40
    assert s[-1] == 0.09500000000000003


def test_rotation_steps_rotations_zero():
    # `rotations 0` sets inc_rot = 1.0, num_rot = 0 (src/p-tracker.cpp:1191-1201): one step at -0.0
    s = ptref.rotation_steps(1.0, 0)
    assert len(s) == 1 and s[0] == 0.0 and math.copysign(1.0, s[0]) == -1.0


def test_rotation_steps_reject_nonpositive_increment():
    with pytest.raises(ValueError):
        ptref.rotation_steps(0.0, 3)


# ---- patterns, src/p-tracker.cpp:1369-1433 -------------------------------------------------
@pytest.mark.parametrize("r,n", [(1, 5), (2, 13), (8, 197), (12, 441), (20, 1257), (40, 5025)])
def test_circ_pattern_gauss_circle_counts(r, n):
    # N(r) = #{(x, y): x^2 + y^2 <= r^2}, OEIS A000328
    p = ptref.circ_pattern(r)
    assert len(p) == n
    assert len(p) == sum(1 for x in range(-r, r + 1) for y in range(-r, r + 1) if x * x + y * y <= r * r)


def test_circ_pattern_row_major_order():
    p = ptref.circ_pattern(3)
    keys = [(int(dy), int(dx)) for dx, dy in p]
    assert keys == sorted(keys)  # dy outer, dx inner
    assert tuple(p[0]) == (0, -3) and tuple(p[-1]) == (0, 3)


def test_rect_and_ring_patterns():
    assert len(ptref.rect_pattern(2, 1)) == 15
    ring = ptref.ring_pattern(2, 3)
    assert len(ring) == len(ptref.circ_pattern(3)) - sum(
        1 for x in range(-3, 4) for y in range(-3, 4) if x * x + y * y < 4)


# ---- rotation of offsets, src/p-tracker.cpp:615-621 ---------------------------------------
def test_rotate_identity_and_quarter_turns():
    p = ptref.circ_pattern(6)
    assert np.array_equal(ptref.rotate_pattern(p, 1.0, 0.0), p)
    # y points down: matrix [[c, s], [-s, c]]; a quarter turn maps (x, y) -> (y, -x)
    q = ptref.rotate_pattern(p, math.cos(math.pi / 2), math.sin(math.pi / 2))
    assert np.array_equal(q, np.stack([p[:, 1], -p[:, 0]], axis=1))
    h = ptref.rotate_pattern(p, math.cos(math.pi), math.sin(math.pi))
    assert np.array_equal(h, -p)


def test_rotate_rounds_half_away_from_zero():
    p = np.array([[1, 0], [-1, 0], [3, 0], [-3, 0], [5, 0]], dtype=np.int32)
    q = ptref.rotate_pattern(p, 0.5, 0.0)
    assert q[:, 0].tolist() == [1, -1, 2, -2, 3]  # 0.5, -0.5, 1.5, -1.5, 2.5


def test_rotated_patterns_are_multisets_as_measured():
    # measured from the reference's loop: up to 12 of 1,257 pixels (r = 20) and 40 of
    # 5,025 (r = 40) land on an already used offset over the config-2 steps
    steps = ptref.rotation_steps(0.01, 10)
    for r, expected in ((20, 12), (40, 40)):
        p = ptref.circ_pattern(r)
        worst = 0
        for d in steps:
            q = ptref.rotate_pattern(p, math.cos(d), math.sin(d))
            worst = max(worst, len(q) - len({tuple(v) for v in q}))
        assert worst == expected


@pytest.mark.parametrize("r,identical", [(8, 12), (12, 8), (20, 4), (40, 2)])
def test_adjacent_identical_rotation_steps_as_measured(r, identical):
    steps = ptref.rotation_steps(0.01, 10)
    p = ptref.circ_pattern(r)
    rot = [ptref.rotate_pattern(p, math.cos(d), math.sin(d)) for d in steps]
    assert sum(np.array_equal(rot[k], rot[k + 1]) for k in range(len(steps) - 1)) == identical


# ---- bicubic sampler, src/interpol.cpp:42-86 -----------------------------------------------
def _frame(rng, H=40, W=50, hi=4096):
    return rng.integers(0, hi, size=(H, W)).astype(np.uint16)


def test_sampler_matches_reference_compiled_code():
    if not ptref.ref_interp_available():
        pytest.skip("oracle/_ref not built (the reference tree is absent)")
    rng = np.random.default_rng(7)
    f = _frame(rng)
    xy = np.stack([rng.uniform(1, 46.99, 4000), rng.uniform(1, 36.99, 4000)], axis=1)
    xy[:50] = np.round(xy[:50])  # integer nodes too
    a = ptref.interp_cubic(f, xy)
    b = ptref.ref_interp_cubic(f, xy)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_sampler_exact_at_nodes():
    rng = np.random.default_rng(8)
    f = _frame(rng)
    ys, xs = np.mgrid[2:30, 2:40]
    xy = np.stack([xs.ravel(), ys.ravel()], axis=1).astype(np.float64)
    assert np.array_equal(ptref.interp_cubic(f, xy), f[ys.ravel(), xs.ravel()].astype(np.float64))


def test_sampler_reproduces_quadratics():
    # Keys' a = -1/2 cubic convolution reproduces polynomials of degree <= 2
    ys, xs = np.mgrid[0:30, 0:30].astype(np.float64)
    poly = lambda x, y: 100 + 3 * x + 2 * y + 0.5 * x * x + x * y + 0.25 * y * y  # noqa: E731
    f = np.round(4 * poly(xs, ys)).astype(np.uint16)  # integer: 4 * poly
    rng = np.random.default_rng(9)
    pts = np.stack([rng.uniform(2, 26, 500), rng.uniform(2, 26, 500)], axis=1)
    v = ptref.interp_cubic(f, pts)
    assert np.allclose(v, 4 * poly(pts[:, 0], pts[:, 1]), rtol=0, atol=1e-9)


def test_sampler_is_not_transposed():
    f = np.zeros((20, 30), dtype=np.uint16)
    f[5, 12] = 1000  # row y = 5, column x = 12
    assert ptref.interp_cubic(f, [[12.0, 5.0]])[0] == 1000.0
    assert ptref.interp_cubic(f, [[5.0, 12.0]])[0] == 0.0


# ---- integer ZNCC search, src/p-tracker.cpp:659-743 ----------------------------------------
def _tiny_workload(rng, W=48, H=44, r=4, shift=(2, -1), noise=True, hi=256):
    I0 = rng.integers(0, hi, size=(H, W)).astype(np.uint16)
    I1 = np.roll(I0, shift=(shift[1], shift[0]), axis=(0, 1))
    if noise:
        I1 = np.clip(I1.astype(np.int64) + rng.integers(-3, 4, size=I1.shape), 0, hi - 1).astype(np.uint16)
    pat = ptref.circ_pattern(r)
    g = np.zeros(1, dtype=pk.GRAIN_DTYPE)
    g["refcoord_xpix"], g["refcoord_ypix"] = W // 2, H // 2
    return ptref.Workload(I0, I1, g, pat, np.array([0, len(pat)], dtype=np.int64), bits=8)


def _exact_pick(w, grains, p0, z):
    """Exact rational argmax over the reference's candidate order (first max wins, only
    values > 0 count), from integer sums."""
    g = grains[0]
    I0 = w.I0.astype(np.int64)
    I1 = w.I1.astype(np.int64)
    pat = w.pattern(0)
    N = len(pat)
    a = np.array([I0[g["refcoord_ypix"] + y, g["refcoord_xpix"] + x] for x, y in p0[:N]])
    Sa, Saa = int(a.sum()), int((a * a).sum())
    d0 = N * Saa - Sa * Sa
    gdx, gdy = int(round(g["dx"])), int(round(g["dy"]))
    best = None
    pick = (0, 0, 0.0)
    for drot in ptref.rotation_steps(z.inc_rot, z.num_rot):
        tr = g["refrot"] + g["drot"] + drot
        p1 = ptref.rotate_pattern(pat, math.cos(tr), math.sin(tr))
        for u in range(-z.left, z.right + 1):
            for v in range(-z.up, z.down + 1):
                xs = g["refcoord_xpix"] + gdx + u + p1[:, 0]
                ys = g["refcoord_ypix"] + gdy + v + p1[:, 1]
                b = I1[ys, xs]
                Sb, Sbb, Sab = int(b.sum()), int((b * b).sum()), int((a * b).sum())
                num = N * Sab - Sa * Sb
                d1 = N * Sbb - Sb * Sb
                if num <= 0 or d0 == 0 or d1 == 0:
                    continue
                key = Fraction(num * num, d1)  # NCC^2 * d0, monotone in NCC for NCC > 0
                if best is None or key > best:
                    best, pick = key, (u, v, float(drot))
    ncc = math.sqrt(best / d0) if best is not None else 0.0
    return pick, ncc


@pytest.mark.parametrize("seed", range(6))
def test_search_pick_equals_exact_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    w = _tiny_workload(rng, shift=(int(rng.integers(-3, 4)), int(rng.integers(-3, 4))))
    z = pk.zone(4, num_rot=2, inc_rot=0.05)
    gr, p0 = ptref.precompute(w)
    pick, ncc = _exact_pick(w, gr, p0, z)
    ptref.follow_pattern(w, gr, p0, 0, z, pk.STAGE_MAIN, 0.0, 0.0)
    assert (int(round(gr[0]["dx"])), int(round(gr[0]["dy"])), gr[0]["drot"]) == pick
    assert abs(gr[0]["NCC"] - ncc) < 1e-12


def test_search_ncc_equals_pearson_correlation():
    rng = np.random.default_rng(3)
    w = _tiny_workload(rng)
    z = pk.zone(3, num_rot=0, inc_rot=1.0)
    gr, p0 = ptref.precompute(w)
    ncc = ptref.candidate_ncc(w, gr, p0, 0, z)
    g = gr[0]
    pat = w.pattern(0)
    a = np.array([w.I0[g["refcoord_ypix"] + y, g["refcoord_xpix"] + x] for x, y in pat], dtype=np.float64)
    k = 0
    for u in range(-3, 4):
        for v in range(-3, 4):
            b = np.array([w.I1[g["refcoord_ypix"] + v + y, g["refcoord_xpix"] + u + x] for x, y in pat], dtype=np.float64)
            assert abs(ncc[k] - np.corrcoef(a, b)[0, 1]) < 1e-12
            k += 1


def test_search_finds_integer_translation_with_ncc_one():
    rng = np.random.default_rng(4)
    w = _tiny_workload(rng, shift=(3, -2), noise=False)
    z = pk.zone(4, num_rot=0, inc_rot=1.0)
    gr, p0 = ptref.precompute(w)
    ptref.follow_pattern(w, gr, p0, 0, z, pk.STAGE_MAIN, 0.7, 0.5)
    assert (gr[0]["dx"], gr[0]["dy"], gr[0]["NCC"]) == (3.0, -2.0, 1.0)


def test_search_ncc_invariant_under_affine_gray_change():
    rng = np.random.default_rng(5)
    w = _tiny_workload(rng, hi=200)
    z = pk.zone(3, num_rot=1, inc_rot=0.1)
    gr, p0 = ptref.precompute(w)
    a = ptref.candidate_ncc(w, gr, p0, 0, z)
    w2 = ptref.Workload(w.I0, (3 * w.I1.astype(np.int64) + 17).astype(np.uint16), w.grains, w.pattern_xy,
                        w.pattern_begin)
    b = ptref.candidate_ncc(w2, gr, p0, 0, z)
    assert np.allclose(a, b, rtol=0, atol=1e-12)


def test_exact_ties_pick_the_first_rotation_step():
    # a radius-2 pattern rotated by at most 0.03 rad rounds to the same offsets for every
    # step: the NCCs tie exactly and the first step (-0.03) wins, as `best < NCC`. The
    # accumulated loop stops after 6 steps here (-0.03 + 6 * 0.01 > 0.03 in double).
    rng = np.random.default_rng(6)
    w = _tiny_workload(rng, r=2, shift=(1, 1), noise=False)
    z = pk.zone(2, num_rot=3, inc_rot=0.01)
    steps = ptref.rotation_steps(0.01, 3)
    assert len(steps) == 6
    gr, p0 = ptref.precompute(w)
    ncc = ptref.candidate_ncc(w, gr, p0, 0, z).reshape(len(steps), -1)
    assert all(np.array_equal(ncc[0], ncc[k]) for k in range(len(steps)))
    ptref.follow_pattern(w, gr, p0, 0, z, pk.STAGE_MAIN, 0.7, 0.5)
    assert gr[0]["drot"] == ptref.rotation_steps(0.01, 3)[0]
    assert (gr[0]["dx"], gr[0]["dy"]) == (1.0, 1.0)


def test_main_stage_below_threshold_keeps_motion_and_queues():
    rng = np.random.default_rng(11)
    w = _tiny_workload(rng, shift=(0, 0))
    w.I1[:] = rng.integers(0, 256, size=w.I1.shape)  # unrelated frame
    z = pk.zone(2, num_rot=0, inc_rot=1.0)
    gr, p0 = ptref.precompute(w)
    gr["dx"], gr["drot"] = 0.25, 0.125
    q = ptref.follow_pattern(w, gr, p0, 0, z, pk.STAGE_MAIN, 0.99, 0.5)
    assert q == 1
    assert gr[0]["dx"] == 0.25 and gr[0]["drot"] == 0.125 and gr[0]["upix"] == 0.0
    assert gr[0]["NCC"] == gr[0]["NCC_rescue"] < 0.99


def test_super_rescue_updates_only_above_threshold():
    rng = np.random.default_rng(12)
    w = _tiny_workload(rng, shift=(1, 0), noise=False)
    z = pk.zone(2, num_rot=0, inc_rot=1.0)
    gr, p0 = ptref.precompute(w)
    g2 = gr.copy()
    ptref.follow_pattern(w, gr, p0, 0, z, pk.STAGE_SUPER, 0.7, 1.0)  # best = 1.0, not > 1.0
    assert gr[0]["dx"] == 0.0 and gr[0]["NCC_rescue"] == 1.0
    ptref.follow_pattern(w, g2, p0, 0, z, pk.STAGE_SUPER, 0.7, 0.5)
    assert g2[0]["dx"] == 1.0


# ---- subpixel objective, src/p-tracker.cpp:1018-1088 ---------------------------------------
def test_objective_at_zero_equals_integer_candidate():
    rng = np.random.default_rng(13)
    w = _tiny_workload(rng, shift=(2, 1))
    gr, p0 = ptref.precompute(w)
    gr["dx"], gr["dy"] = 2.0, 1.0
    f = ptref.ncc_objective(w, gr, p0, 0, [0.0, 0.0, 0.0])
    z = pk.zone(0, num_rot=0, inc_rot=1.0)
    ncc = ptref.candidate_ncc(w, gr, p0, 0, z)[0]
    assert abs(f - (1.0 - ncc)) < 1e-14


def test_objective_overflow_weight():
    rng = np.random.default_rng(14)
    w = _tiny_workload(rng, W=60, H=60, shift=(0, 0), noise=False)
    gr, p0 = ptref.precompute(w)
    X = [3.0, 4.0, 0.0]  # |(X0, X1)| = 5 > 3
    f1 = ptref.ncc_objective(w, gr, p0, 0, X, dmax=3.0, stiffness=0.0)
    f2 = ptref.ncc_objective(w, gr, p0, 0, X, dmax=3.0, stiffness=2.0)
    assert abs(f2 - f1 * (2.0 * (5.0 - 3.0) + 1.0)) < 1e-14


def _smooth_pair(sx, sy, th, W=80, H=80, r=12):
    """Analytic smooth texture: the pair differs by a known subpixel motion."""
    rng = np.random.default_rng(21)
    cx, cy = W / 2, H / 2
    bl = rng.uniform(-r, r, (30, 2))
    amp = rng.uniform(800, 2000, 30)

    def tex(x, y):
        v = np.full_like(x, 1000.0)
        for (bx, by), a in zip(bl, amp):
            v += a * np.exp(-((x - bx) ** 2 + (y - by) ** 2) / (2 * 2.5 ** 2))
        return v

    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    I0 = np.round(tex(xs - cx, ys - cy)).astype(np.uint16)
    px, py = xs - cx - sx, ys - cy - sy
    c, s = math.cos(th), math.sin(th)
    I1 = np.round(tex(c * px - s * py, s * px + c * py)).astype(np.uint16)
    pat = ptref.circ_pattern(r)
    g = np.zeros(1, dtype=pk.GRAIN_DTYPE)
    g["refcoord_xpix"], g["refcoord_ypix"] = int(cx), int(cy)
    return ptref.Workload(I0, I1, g, pat, np.array([0, len(pat)], dtype=np.int64), bits=12)


def test_powell_subpixel_recovers_known_motion():
    w = _smooth_pair(0.37, -0.21, 0.013)
    gr, p0 = ptref.precompute(w)
    ne = ptref.subpixel(w, gr, p0, 0, pk.default_subpix())
    assert ne > 10
    assert abs(gr[0]["dx"] - 0.37) < 0.01 and abs(gr[0]["dy"] + 0.21) < 0.01 and abs(gr[0]["drot"] - 0.013) < 1e-3
    assert gr[0]["NCC_subpix"] > 0.999


def test_powell_minimizes_quadratic_and_rosenbrock():
    f1 = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)(
        lambda x, u: (x[0] - 1) ** 2 + 10 * (x[1] + 2) ** 2 + 0.5 * (x[2] - 0.3) ** 2 + 0.3 * (x[0] - 1) * (x[1] + 2))
    ptref.lib().ptref_powell.argtypes = [type(f1), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                         ctypes.c_int32, ctypes.c_void_p]
    ptref.lib().ptref_powell.restype = ctypes.c_double
    x = np.zeros(3)
    dx = np.full(3, 0.01)
    ne = ctypes.c_int64(0)
    fmin = ptref.lib().ptref_powell(f1, None, x.ctypes.data, dx.ctypes.data, 1e-12, 200, ctypes.byref(ne))
    assert np.allclose(x, [1, -2, 0.3], atol=1e-5) and fmin < 1e-10
    f2 = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)(
        lambda x, u: 100 * (x[1] - x[0] ** 2) ** 2 + (1 - x[0]) ** 2 + (x[2] - 2) ** 2)
    x = np.array([-1.2, 1.0, 0.0])
    dx = np.full(3, 0.1)
    ptref.lib().ptref_powell(f2, None, x.ctypes.data, dx.ctypes.data, 1e-14, 500, ctypes.byref(ne))
    assert np.allclose(x, [1, 1, 2], atol=1e-4)


def test_fp64_minimizer_restarts_never_worse_and_reach_a_narrow_valley_minimum():
    # a narrow, rotated valley (condition number 1e6) where one Powell run stops early
    F = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)
    f = F(lambda x, u: (x[0] + x[1] - 1) ** 2 + 1e-6 * (x[0] - x[1] - 0.5) ** 2 + 1e3 * (x[2] - 0.01 * x[0]) ** 2)
    lib = ptref.lib()
    lib.ptref_powell.argtypes = [F, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int32,
                                 ctypes.c_void_p]
    lib.ptref_powell.restype = ctypes.c_double
    lib.ptref_minimize_fp64.argtypes = [F, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                        ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]
    lib.ptref_minimize_fp64.restype = ctypes.c_double
    dx = np.full(3, 0.01)
    ne = ctypes.c_int64(0)
    x1 = np.zeros(3)
    f1 = lib.ptref_powell(f, None, x1.ctypes.data, dx.ctypes.data, 1e-13, 200, ctypes.byref(ne))
    x2 = np.zeros(3)
    f2 = lib.ptref_minimize_fp64(f, None, x2.ctypes.data, dx.ctypes.data, 1e-13, 200, 20, ctypes.byref(ne))
    xs = np.array([0.75, 0.25, 0.0075])  # the closed-form minimizer
    assert f2 <= f1
    assert np.abs(x2 - xs).max() < 1e-5 and f2 < 1e-14


def test_fp64_minimizer_continues_where_one_powell_run_stalls():
    # grain 1656 of the seeded 2000-grain 6000x4000 pair: one Powell run at a tight
    # tolerance stops 2e-8 above the minimum in a flat valley (3.6e-5 rad off); the
    # restarted run reaches a point where the objective's central-difference gradient
    # vanishes
    from ptracker_amd import synth
    w = synth.workload(2)
    W = ptref.Workload(w.I0, w.I1, w.grains, w.pattern_xy, w.pattern_begin, w.bits)
    z = w.params
    g = 1656
    ti = pk.track_params(z.search_zone, z.search_zone_rescue, z.search_zone_super_rescue, NCC_min=0.7,
                         rescue_level=1, subpixel=0)
    oi, _, _ = ptref.track_pair(W, ti, nthreads=1, subset=np.array([g]))
    pat = np.ascontiguousarray(W.pattern(g), dtype=np.int32)
    sub = ptref.Workload(w.I0, w.I1, oi, pat, np.array([0, len(pat)], dtype=np.int64), w.bits)
    gr, p0 = ptref.precompute(sub, oi)
    one, rst = gr.copy(), gr.copy()
    ptref.subpixel(sub, one, p0, 0, pk.default_subpix(pk.SUBPIX_POWELL, subpix_tol=1e-13))
    ptref.subpixel(sub, rst, p0, 0, pk.default_subpix(pk.SUBPIX_LM, subpix_tol=1e-13))
    f_one, f_rst = 1 - one["NCC_subpix"][0], 1 - rst["NCC_subpix"][0]
    assert f_rst < f_one - 1e-9
    X = np.array([rst[c][0] - gr[c][0] for c in ("dx", "dy", "drot")])
    h = 1e-6
    grad = [(ptref.ncc_objective(sub, gr, p0, 0, X + h * e) - ptref.ncc_objective(sub, gr, p0, 0, X - h * e)) / (2 * h)
            for e in np.eye(3)]
    assert np.abs(grad).max() < 1e-6


# ---- undistor, src/distortion_correction.cpp:20-61 -----------------------------------------
def _params(**kw):
    X = np.zeros(14)
    X[0], X[1], X[8], X[9] = 300.0, 200.0, 300.0, 200.0
    for k, v in kw.items():
        X[{"K1": 2, "K2": 3, "K3": 4, "P1": 5, "P2": 6, "P3": 7, "t1": 10, "a1": 11, "t2": 12, "a2": 13}[k]] = v
    return X


def test_undistor_identity_and_radial_closed_form():
    assert ptref.undistor(_params(), 123.25, 77.5) == (123.25, 77.5)
    K1 = 1e-6
    xu, yu = ptref.undistor(_params(K1=K1), 400.0, 250.0)
    dx, dy = 100.0, 50.0
    r2 = dx * dx + dy * dy
    assert abs(xu - (400.0 + dx * K1 * r2)) < 1e-9 and abs(yu - (250.0 + dy * K1 * r2)) < 1e-9


def test_undistor_tangential_and_parallax_closed_form():
    P1, P2 = 2e-6, -1e-6
    xu, yu = ptref.undistor(_params(P1=P1, P2=P2), 310.0, 220.0)
    dx, dy = 10.0, 20.0
    r2 = dx * dx + dy * dy
    assert abs(xu - (310.0 + P1 * (r2 + 2 * dx * dx) + 2 * P2 * dx * dy)) < 1e-10
    assert abs(yu - (220.0 + 2 * P1 * dx * dy + P2 * (r2 + 2 * dy * dy))) < 1e-10
    # parallax1 along theta1 = 0 scales x about Xcp by (1 + a1)
    xu, yu = ptref.undistor(_params(a1=0.01), 310.0, 220.0)
    assert abs(xu - (300.0 + 1.01 * 10.0)) < 1e-10 and abs(yu - 220.0) < 1e-10


def test_distort_point_inverts_undistor():
    X = _params(K1=-2e-7, K2=1e-13, P1=1e-6, P3=1e-7, a2=0.002, t2=0.01)
    for xu, yu in [(0.0, 0.0), (599.0, 399.0), (123.4, 256.7), (300.0, 200.0)]:
        xd, yd = ptref.distort_point(X, xu, yu)
        x2, y2 = ptref.undistor(X, xd, yd)
        assert abs(x2 - xu) < 1e-9 and abs(y2 - yu) < 1e-9


def test_joint_objective_without_lens_is_the_reference_objective():
    # the differentiation form's objective at K1 = K2 = 0 is NCC_to_minimize_xyR at
    # X = P - (dx, dy, drot), bit for bit (the lens maps are then the identity; undistor and
    # its inverse are pinned separately), and its batch is the single evaluations
    from ptracker_amd import synth

    w = synth.workload(2, n_grains=12, size=(700, 600), seed=41)
    W = ptref.Workload(w.I0, w.I1, w.grains, w.pattern_xy, w.pattern_begin, w.bits)
    gp, p0 = ptref.precompute(W)
    rng = np.random.default_rng(3)
    glist = np.arange(12, dtype=np.int32)
    P = np.stack([gp["dx"] + rng.uniform(-1, 1, 12), gp["dy"] + rng.uniform(-1, 1, 12),
                  gp["drot"] + rng.uniform(-0.02, 0.02, 12)], 1)
    lens = [0.0, 0.0, 350.0, 300.0]
    f = ptref.joint_objective_batch(W, gp, p0, glist, P, lens, 2)
    for g in range(12):
        X = P[g] - np.array([gp[g]["dx"], gp[g]["dy"], gp[g]["drot"]])
        fr = ptref.ncc_objective(W, gp, p0, g, X)
        assert f[g] == ptref.joint_objective(W, gp, p0, g, P[g], lens)
        assert abs(f[g] - fr) < 1e-12


def test_joint_objective_lens_moves_samples_as_undistor():
    # with a lens, moving the scene point of each pattern pixel by T = U(q + t) - U(q) for a
    # whole-pixel t (the motion that maps each pixel onto pixel q + t) is, for a pure
    # translation of the frame content by t, a perfect match: 1 - NCC vanishes up to the
    # rounding of the lens round trip
    rng = np.random.default_rng(8)
    I0 = rng.integers(100, 4000, size=(200, 220)).astype(np.uint16)
    t = (3, -2)
    I1 = np.roll(I0, (t[1], t[0]), (0, 1))
    pat = ptref.circ_pattern(6).astype(np.int32)
    gr = np.zeros(1, dtype=pk.GRAIN_DTYPE)
    gr["refcoord_xpix"], gr["refcoord_ypix"] = 110, 100
    W = ptref.Workload(I0, I1, gr, pat, np.array([0, len(pat)], dtype=np.int64), 12)
    gp, p0 = ptref.precompute(W)
    X = np.zeros(14)
    X[0], X[1], X[2], X[3], X[8], X[9] = 90.0, 80.0, 0.0, 0.0, 90.0, 80.0
    assert ptref.joint_objective(W, gp, p0, 0, [t[0], t[1], 0.0], [0.0, 0.0, 90.0, 80.0]) < 1e-12
    # a lens with the scene motion of the reference point's image: every pixel moves by t to
    # within the lens's local deformation, far below a pixel over the pattern
    k1 = 1e-8
    lens = [k1, 0.0, 90.0, 80.0]
    X[2] = k1
    ux, uy = ptref.undistor(X, 110.0, 100.0)
    vx, vy = ptref.undistor(X, 110.0 + t[0], 100.0 + t[1])
    f = ptref.joint_objective(W, gp, p0, 0, [vx - ux, vy - uy, 0.0], lens)
    fwrong = ptref.joint_objective(W, gp, p0, 0, [t[0], t[1], 0.0], [1e-5, 0.0, 90.0, 80.0])
    assert f < 1e-4 < fwrong


def test_remap_identity():
    rng = np.random.default_rng(30)
    f = rng.integers(0, 4096, size=(40, 50)).astype(np.uint16)
    X = np.zeros(14)
    X[0] = X[1] = X[8] = X[9] = -1.0  # centres default to the frame centre
    assert np.array_equal(ptref.remap(f, 12, X), f)


def _radial_inverse(c, K1, K2, xu, yu):
    """The distorted point whose radial undistor image is (xu, yu): along the ray from the
    centre, r_u = r_d (1 + K1 r_d^2 + K2 r_d^4) solved for r_d by bisection (independent of
    the oracle's two-dimensional Newton iteration)."""
    dx, dy = xu - c[0], yu - c[1]
    ru = np.hypot(dx, dy)
    lo, hi = np.zeros_like(ru), 2.0 * ru + 1.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        below = mid * (1.0 + K1 * mid ** 2 + K2 * mid ** 4) < ru
        lo, hi = np.where(below, mid, lo), np.where(below, hi, mid)
    s = np.where(ru > 0, 0.5 * (lo + hi) / np.where(ru > 0, ru, 1.0), 1.0)
    return c[0] + dx * s, c[1] + dy * s


def test_remap_samples_a_quadratic_through_the_inverse_lens():
    # Catmull-Rom reproduces polynomials of degree 2, so on a quadratic frame every pixel
    # whose 16 taps lie inside the frame must equal the quadratic at the distorted point
    # that undistor maps onto it, rounded; an off-centre lens and an asymmetric quadratic
    # on a non-square frame make a warp in the wrong direction (undistor instead of its
    # inverse) or with x and y exchanged fail
    H, W = 48, 64
    c, K1, K2 = (27.3, 21.7), 4e-5, 1e-9
    X = np.array([c[0], c[1], K1, K2, 0, 0, 0, 0, c[0], c[1], 0, 0, 0, 0], dtype=np.float64)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)

    def q(x, y):
        return 300.0 + 7.0 * x + 11.0 * y + 2.0 * x * x + 3.0 * x * y + y * y

    f = q(xx, yy).astype(np.uint16)
    assert f.max() < 65535
    dst = ptref.remap(f, 16, X).astype(np.float64)
    xd, yd = _radial_inverse(c, K1, K2, xx, yy)
    inside = (np.floor(xd) >= 1) & (np.floor(xd) <= W - 3) & (np.floor(yd) >= 1) & (np.floor(yd) <= H - 3)
    assert inside.sum() > 0.9 * inside.size
    e = q(xd, yd)
    clear = np.abs(np.abs(e - np.floor(e)) - 0.5) > 1e-6  # away from a rounding tie
    assert np.array_equal(dst[inside & clear], np.round(e[inside & clear]))
    assert np.abs(dst - e)[inside].max() <= 0.5 + 1e-6
    # the same check would reject the wrong direction and exchanged axes on nearly every pixel
    r2 = (xx - c[0]) ** 2 + (yy - c[1]) ** 2
    k = K1 * r2 + K2 * r2 * r2
    wrong = np.round(q(xx + (xx - c[0]) * k, yy + (yy - c[1]) * k))
    assert np.mean(wrong[inside] != dst[inside]) > 0.9
    swapped = np.round(q(yd * W / H, xd * H / W))
    assert np.mean(swapped[inside] != dst[inside]) > 0.9


def test_remap_undoes_the_lens_on_a_smooth_frame():
    # a smooth 12-bit frame seen through a radial lens (synth.distort_frame: pixel d shows the
    # scene at undistor(d), sampled by an independent numpy Catmull-Rom) and remapped back:
    # away from the border the round trip of two bicubic resamplings returns the frame
    # within 1 gray level on average and 40 at most (of 4095), while remapping with the
    # opposite lens leaves errors of over 100 on average
    from scipy.ndimage import gaussian_filter
    from ptracker_amd import synth

    rng = np.random.default_rng(5)
    H, W = 160, 200
    img = gaussian_filter(rng.normal(size=(H, W)), 3.0)
    img = np.clip(2048 + img / img.std() * 600, 0, 4095).round().astype(np.uint16)
    X = np.array([100.0, 80.0, -2e-6, 0, 0, 0, 0, 0, 100.0, 80.0, 0, 0, 0, 0])
    seen = synth.distort_frame(img, X, 12)
    back = ptref.remap(seen, 12, X)
    err = np.abs(back.astype(np.float64) - img)[12:-12, 12:-12]
    assert err.mean() < 1.0 and err.max() <= 40
    Xw = X.copy()
    Xw[2] = -X[2]
    errw = np.abs(ptref.remap(seen, 12, Xw).astype(np.float64) - img)[12:-12, 12:-12]
    assert errw.mean() > 100.0


# ---- the per-pair stage sequence, src/p-tracker.cpp:472-581 ---------------------------------
def test_track_pair_masked_grains_and_increments():
    from ptracker_amd import synth
    w = synth.workload(1, n_grains=12, size=(256, 256))
    grains = w.grains.copy()
    grains["masked"][3] = 1
    grains["dx"][3] = 0.4
    W = ptref.Workload(w.I0, w.I1, grains, w.pattern_xy, w.pattern_begin, w.bits)
    gr, counts, ne = ptref.track_pair(W, w.params, nthreads=2)
    assert gr["NCC"][3] == 0.0 and gr["NCC_subpix"][3] == 0.0 and gr["dx"][3] == 0.4 and gr["upix"][3] == 0.0
    others = np.arange(12) != 3
    assert np.all(gr["NCC"][others] > 0.8)
    assert np.array_equal(gr["upix"], gr["dx"] - gr["dx_prev"])
    assert np.array_equal(gr["dx_prev"], grains["dx"])


def test_track_pair_rescue_recovers_large_motion():
    from ptracker_amd import synth
    w = synth.workload(2, n_grains=8, size=(700, 600), max_disp=6.0, seed=77)
    grains = w.grains.copy()
    W = ptref.Workload(w.I0, w.I1, grains, w.pattern_xy, w.pattern_begin, w.bits)
    z = pk.zone(2, num_rot=0, inc_rot=1.0)  # too small for +-6 px motion
    zr = pk.zone(8, num_rot=0, inc_rot=1.0)
    prm = pk.track_params(z, zr, zr, rescue_level=1, subpixel=0)
    gr, counts, _ = ptref.track_pair(W, prm, nthreads=2)
    assert counts[0] > 0
    rescued = gr["NCC"] < 0.7
    assert rescued.sum() == counts[0]
    assert np.all(gr["NCC_rescue"][rescued] > gr["NCC"][rescued])
    assert np.all(np.abs(gr["dx"] - np.round(w.truth["disp"][:, 0])) <= 1.0)


def test_change_reference_matches_closed_form():
    g = np.zeros(2, dtype=pk.GRAIN_DTYPE)
    g["refcoord_xpix"] = [10, 20]
    g["dx"] = [2.6, -0.4]
    g["refrot"] = [0.1, 0.0]
    g["drot"] = [0.05, -0.02]
    h = ptref.change_reference(g)
    assert h["refcoord_xpix"].tolist() == [13, 20]
    assert np.allclose(h["dx"], [-0.4, -0.4]) and np.allclose(h["refrot"], [0.15, -0.02]) and np.all(h["drot"] == 0)


# ---- distortion calibration: the equiprojectivity criterion ---------------------------------
def _lens(k1=-2e-9, k2=1e-16, W=6000, H=4000):
    return np.array([W / 2, H / 2, k1, k2, 0, 0, 0, 0, W / 2, H / 2, 0, 0, 0, 0], dtype=np.float64)


def _rigid_views(X, ref, shifts):
    """Tracked displacements of reference pixels ref when the scene moves rigidly by each
    shift behind the lens X: the pixel showing undistor(ref) + t, minus ref."""
    disp = np.zeros((len(shifts), len(ref), 2))
    for k, t in enumerate(shifts):
        for g, (x, y) in enumerate(ref):
            ux, uy = ptref.undistor(X, x, y)
            dx, dy = ptref.distort_point(X, ux + t[0], uy + t[1])
            disp[k, g] = (dx - x, dy - y)
    return disp


def test_precompute_pairs_matches_brute_force():
    ref = np.random.default_rng(3).integers(0, 800, (40, 2)).astype(np.float64)
    exp = [(i, j) for i in range(40) for j in range(i + 1, 40)
           if 100.0 <= math.sqrt((ref[j, 0] - ref[i, 0]) ** 2 + (ref[j, 1] - ref[i, 1]) ** 2) <= 400.0]
    assert [tuple(p) for p in ptref.precompute_pairs(ref, 100.0, 400.0)] == exp


def test_equiproj_uniform_scale_closed_form():
    # no correction (identity parameters) and a uniform scale s about c: every residual is
    # (s - 1) |B1 - A1|
    X = _lens(0.0, 0.0)
    rng = np.random.default_rng(4)
    ref = rng.uniform(500, 5500, (30, 2))
    s, c = 1.0 + 3e-4, np.array([2500.0, 1800.0])
    disp = ((s - 1.0) * (ref - c))[None]
    pairs = ptref.precompute_pairs(ref, 0.0, 1e9)
    r = ptref.equiproj_residuals(X, ref, disp, np.ones((1, 30)), pairs)
    exp = (s - 1.0) * np.hypot(*(ref[pairs[:, 1]] - ref[pairs[:, 0]]).T)
    assert np.abs(r - exp).max() < 1e-9


def test_equiproj_vanishes_for_rigid_motion_at_the_true_lens_only():
    X = _lens()
    rng = np.random.default_rng(5)
    ref = np.round(rng.uniform(300, 5700, (25, 2)))
    disp = _rigid_views(X, ref, [(36.0, 11.0), (-29.0, 33.0)])
    pairs = ptref.precompute_pairs(ref, 0.0, 1000.0)
    ncc = np.ones((2, 25))
    assert np.abs(ptref.equiproj_residuals(X, ref, disp, ncc, pairs)).max() < 1e-8
    assert np.abs(ptref.equiproj_residuals(_lens(0.0, 0.0), ref, disp, ncc, pairs)).max() > 1e-2
    # a grain below NCC_min drops its pairs
    ncc[1, 3] = 0.5
    r = ptref.equiproj_residuals(_lens(0.0, 0.0), ref, disp, ncc, pairs, NCC_min=0.7).reshape(2, -1)
    drop = (pairs[:, 0] == 3) | (pairs[:, 1] == 3)
    assert (r[1, drop] == 0.0).all() and (r[0, drop] != 0.0).all()


def test_equiproj_residuals_with_given_first_positions():
    # the POWELL_ITER form: with the first positions set to the undistorted reference pixels it
    # is the standard form bit for bit; with them set to the reference pixels it differs by
    # moving only the tracked points, which a hand evaluation of one pair confirms
    X = _lens()
    rng = np.random.default_rng(6)
    ref = np.round(rng.uniform(300, 5700, (20, 2)))
    disp = _rigid_views(X, ref, [(36.0, 11.0), (-29.0, 33.0)])
    pairs = ptref.precompute_pairs(ref, 0.0, 2000.0)
    ncc = np.ones((2, 20))
    ncc[0, 4] = 0.2
    und = np.array([ptref.undistor(X, x, y) for x, y in ref])
    std = ptref.equiproj_residuals(X, ref, disp, ncc, pairs)
    assert np.array_equal(ptref.equiproj_residuals_orig(X, und, ref, disp, ncc, pairs), std)
    r = ptref.equiproj_residuals_orig(X, ref, ref, disp, ncc, pairs).reshape(2, -1)
    k, l = 1, 7
    i, j = pairs[l]
    a2 = np.array(ptref.undistor(X, *(ref[i] + disp[k][i])))
    b2 = np.array(ptref.undistor(X, *(ref[j] + disp[k][j])))
    e = (ref[j] - ref[i]) / np.hypot(*(ref[j] - ref[i]))
    assert abs(r[k, l] - e @ ((b2 - ref[j]) - (a2 - ref[i]))) < 1e-9
    drop = (pairs[:, 0] == 4) | (pairs[:, 1] == 4)
    assert (r[0, drop] == 0.0).all()


def test_minpack_least_squares_recovers_the_lens():
    # the reference's LEVMAR path is MINPACK lmdif (forward-difference Jacobian); scipy's
    # leastsq is that routine. On scaled K1, K2 it recovers the lens from rigid views.
    from scipy.optimize import leastsq

    X = _lens()
    rng = np.random.default_rng(6)
    ref = np.round(rng.uniform(300, 5700, (60, 2)))
    disp = _rigid_views(X, ref, [(36.0, 11.0), (-29.0, 33.0), (22.0, -38.0)])
    pairs = ptref.precompute_pairs(ref, 0.0, 1000.0)
    ncc = np.ones((3, 60))

    def fun(q):
        Y = X.copy()
        Y[2], Y[3] = q[0] * 1e-9, q[1] * 1e-16
        return ptref.equiproj_residuals(Y, ref, disp, ncc, pairs)

    q, _ = leastsq(fun, [0.0, 0.0], ftol=np.finfo(float).eps, xtol=np.finfo(float).eps, gtol=0.0)
    assert abs(q[0] * 1e-9 - X[2]) < 1e-6 * abs(X[2]) and abs(q[1] * 1e-16 - X[3]) < 1e-5 * abs(X[3])


def test_full_size_fixture_is_the_oracle_on_the_generated_pair():
    # the stored full-size results (tests/golden/cfg2_full_oracle.npz) belong to the pair the
    # generator makes now, and the oracle reproduces them bit for bit on a sample of grains
    import os
    import sys

    from ptracker_amd import synth

    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    from make_cfg2_full_fixture import SEARCH_COLS, SUB_COLS, digest, params

    fx = np.load(os.path.join(here, "golden", "cfg2_full_oracle.npz"))
    w = synth.workload(2)
    assert digest(w) == str(fx["digest"])
    W = ptref.Workload(w.I0, w.I1, w.grains, w.pattern_xy, w.pattern_begin, w.bits)
    sample = np.arange(0, w.G, 250)
    gs, _, _ = ptref.track_pair(W, params(w, 0), nthreads=8, subset=sample)
    for c in SEARCH_COLS:
        assert np.array_equal(np.asarray(gs[c], dtype=np.float64).view(np.uint64),
                              fx["search_" + c][sample].view(np.uint64)), c
    gl, _, _ = ptref.track_pair(W, params(w, 1, pk.default_subpix(pk.SUBPIX_LM, subpix_tol=1e-13)), nthreads=8,
                                subset=sample)
    for c in SUB_COLS:
        assert np.array_equal(np.asarray(gl[c], dtype=np.float64).view(np.uint64),
                              fx["sub_" + c][sample].view(np.uint64)), c
    gp, _, _ = ptref.track_pair(W, params(w, 1, pk.default_subpix(pk.SUBPIX_POWELL)), nthreads=8, subset=sample)
    for c in SUB_COLS:
        assert np.array_equal(np.asarray(gp[c], dtype=np.float64).view(np.uint64),
                              fx["powell_" + c][sample].view(np.uint64)), c


def test_rescue_variant_fixture_is_the_oracle_on_the_generated_pair():
    # tests/golden/cfg2_rescue_oracle.npz: the pair with 300 grains moved beyond the main
    # zone; the oracle reproduces the stored columns on a sample of rescued and other grains,
    # and the rescued grains' picks lie beyond +-20 px (the rescue zone's own geometry)
    import os
    import sys

    from ptracker_amd import synth

    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    from make_cfg2_full_fixture import SEARCH_COLS, digest, params
    from make_cfg2_rescue_fixture import FAR

    fx = np.load(os.path.join(here, "golden", "cfg2_rescue_oracle.npz"))
    w = synth.workload(2, far=FAR)
    assert digest(w) == str(fx["digest"])
    assert np.array_equal(fx["far"], w.truth["far"])
    rescued = np.flatnonzero(fx["NCC"] < 0.7)
    assert len(rescued) == fx["counts"][0] >= 250
    moved = np.maximum(np.abs(fx["upix"][rescued]), np.abs(fx["vpix"][rescued]))
    assert np.mean(moved > 20.5) > 0.9
    W = ptref.Workload(w.I0, w.I1, w.grains, w.pattern_xy, w.pattern_begin, w.bits)
    sample = np.union1d(rescued[::40], np.arange(1, w.G, 400))
    gs, _, _ = ptref.track_pair(W, params(w, 0), nthreads=8, subset=sample)
    for c in SEARCH_COLS:
        assert np.array_equal(np.asarray(gs[c], dtype=np.float64).view(np.uint64), fx[c][sample].view(np.uint64)), c


@pytest.mark.parametrize("bits", [12, 8])
def test_full_size_40_step_fixture_is_the_oracle_on_the_generated_pair(bits):
    import os
    import sys

    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    from make_cfg2_full_fixture import digest
    from make_cfg3_full_fixture import COLS, params, workload

    fx = np.load(os.path.join(here, "golden", "cfg3_full_oracle.npz"))
    w = workload(bits)
    assert digest(w) == str(fx[f"digest_{bits}"])
    W = ptref.Workload(w.I0, w.I1, w.grains, w.pattern_xy, w.pattern_begin, w.bits)
    sample = np.arange(7, w.G, 250)
    g, _, _ = ptref.track_pair(W, params(w), nthreads=8, subset=sample)
    for c in COLS:
        assert np.array_equal(np.asarray(g[c], dtype=np.float64).view(np.uint64),
                              fx[f"{c}_{bits}"][sample].view(np.uint64)), c
