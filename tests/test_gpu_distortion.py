"""GPU: the distortion calibration (src/distortion_correction.cpp:135-479) through the C
ABI. The equiprojectivity residuals against the oracle, their analytic Jacobian against
central differences and against MINPACK's forward-difference Jacobian (the reference's
lmdif path), the reduced normal equations, the Levenberg-Marquardt fit against scipy's
MINPACK fit and the lens, and the whole differentiation form on distorted frames: grains
tracked with the LM subpixel stage, then K1, K2 fitted."""
import sys

import numpy as np
import pytest

import ptref

pk = ptref.package()
from ptracker_amd import synth  # noqa: E402

pytestmark = pytest.mark.gpu

sys.path.insert(0, __file__.rsplit("/", 1)[0])
from test_oracle import _lens, _rigid_views  # noqa: E402


def say(msg):
    print(msg, flush=True)


def bits_equal(a, b):
    return np.array_equal(np.asarray(a, dtype=np.float64).view(np.uint64), np.asarray(b, dtype=np.float64).view(np.uint64))


@pytest.fixture(scope="module")
def views():
    X = _lens()
    rng = np.random.default_rng(21)
    ref = np.round(rng.uniform(300, 5700, (70, 2)))
    disp = _rigid_views(X, ref, [(36.0, 11.0), (-29.0, 33.0), (22.0, -38.0)])
    disp += rng.normal(0, 0.01, disp.shape)  # tracking noise
    ncc = rng.uniform(0.75, 1.0, (3, 70))
    ncc[2, 5] = 0.3  # drops the pairs of grain 5 in image 2
    pairs = ptref.precompute_pairs(ref, 0.0, 1000.0)
    return X, ref, disp, ncc, pairs


def every_term():
    # all 14 parameters away from zero, so every path of the derivative is exercised
    return np.array([2950.0, 2040.0, -2e-9, 1e-16, 3e-24, 2e-8, -1e-8, 3e-8, 3010.0, 1985.0, 0.02, 1e-3, -0.015,
                     -2e-3])


def test_residuals_equal_oracle(tracker, views):
    X, ref, disp, ncc, pairs = views
    say(f"equiproj residuals: {len(pairs)} pairs x 3 images")
    tracker.equiproj_set_data(ref, disp, ncc, pairs, NCC_min=0.7)
    for P in (X, every_term()):
        f, jtr, jtj, res, jac = tracker.equiproj_eval(P, full=True)
        exp = ptref.equiproj_residuals(P, ref, disp, ncc, pairs, NCC_min=0.7)
        assert bits_equal(res, exp)
        assert abs(f - (exp ** 2).sum()) <= 1e-12 * (exp ** 2).sum()
        # the reduced sums are those of the returned residuals and Jacobian
        assert np.allclose(jtr, jac.T @ res, rtol=1e-11, atol=0.0)
        assert np.allclose(jtj, jac.T @ jac, rtol=1e-11, atol=1e-300)
    assert (res.reshape(3, -1)[2, (pairs[:, 0] == 5) | (pairs[:, 1] == 5)] == 0.0).all()


def test_jacobian_against_central_differences(tracker, views):
    _, ref, disp, ncc, pairs = views
    P = every_term()
    tracker.equiproj_set_data(ref, disp, ncc, pairs, NCC_min=0.7)
    _, _, _, res, jac = tracker.equiproj_eval(P, full=True)
    # steps between truncation and rounding error: 0.1 px for the distortion centre, 1e-4
    # rad for the parallax angles, about 1e-3 relative for the coefficients
    h = np.array([0.1, 0.1, 2e-12, 1e-19, 3e-27, 2e-10, 1e-10, 3e-10, 1e-3, 1e-3, 1e-4, 1e-5, 1e-4, 2e-5])
    for j in range(14):
        Pp, Pm = P.copy(), P.copy()
        Pp[j] += h[j]
        Pm[j] -= h[j]
        cd = (ptref.equiproj_residuals(Pp, ref, disp, ncc, pairs, 0.7) -
              ptref.equiproj_residuals(Pm, ref, disp, ncc, pairs, 0.7)) / (2 * h[j])
        if j in (8, 9):
            # the parallax centre moves every undistorted point alike: residuals do not
            # depend on it (central differences are rounding noise)
            assert np.abs(jac[:, j]).max() <= 1e-9 * np.abs(jac).max() and np.abs(cd).max() < 1e-6
            continue
        scale = np.abs(cd).max()
        assert np.abs(jac[:, j] - cd).max() < 1e-6 * scale, (j, np.abs(jac[:, j] - cd).max() / scale)


def test_residuals_with_given_first_positions_equal_oracle(tracker, views):
    # the POWELL_ITER form (src/distortion_correction.cpp:316-354): A1, B1 given, only the
    # tracked positions undistorted; residuals bit-identical to the oracle's, the Jacobian
    # (the tracked points' terms only) against central differences, and back to the
    # standard form when the origins are cleared
    X, ref, disp, ncc, pairs = views
    P = every_term()
    rng = np.random.default_rng(22)
    orig = ref + rng.normal(0, 2.0, ref.shape)
    tracker.equiproj_set_data(ref, disp, ncc, pairs, NCC_min=0.7)
    tracker.equiproj_set_origins(orig)
    f, jtr, jtj, res, jac = tracker.equiproj_eval(P, full=True)
    exp = ptref.equiproj_residuals_orig(P, orig, ref, disp, ncc, pairs, NCC_min=0.7)
    assert bits_equal(res, exp)
    assert np.allclose(jtr, jac.T @ res, rtol=1e-11, atol=0.0)
    h = np.array([0.1, 0.1, 2e-12, 1e-19, 3e-27, 2e-10, 1e-10, 3e-10, 1e-3, 1e-3, 1e-4, 1e-5, 1e-4, 2e-5])
    for j in range(14):
        Pp, Pm = P.copy(), P.copy()
        Pp[j] += h[j]
        Pm[j] -= h[j]
        cd = (ptref.equiproj_residuals_orig(Pp, orig, ref, disp, ncc, pairs, 0.7) -
              ptref.equiproj_residuals_orig(Pm, orig, ref, disp, ncc, pairs, 0.7)) / (2 * h[j])
        if j in (8, 9):
            # a shift of the parallax centre moves both tracked points alike: no dependence
            assert np.abs(jac[:, j]).max() <= 1e-9 * np.abs(jac).max() and np.abs(cd).max() < 1e-6
            continue
        scale = np.abs(cd).max()
        assert np.abs(jac[:, j] - cd).max() < 1e-6 * scale, (j, np.abs(jac[:, j] - cd).max() / scale)
    tracker.equiproj_set_origins(None)
    _, _, _, res2, _ = tracker.equiproj_eval(P, full=True)
    assert bits_equal(res2, ptref.equiproj_residuals(P, ref, disp, ncc, pairs, NCC_min=0.7))


def test_jacobian_against_minpack_forward_differences(tracker, views):
    _, ref, disp, ncc, pairs = views
    P = every_term()
    tracker.equiproj_set_data(ref, disp, ncc, pairs, NCC_min=0.7)
    _, _, _, _, jac = tracker.equiproj_eval(P, full=True)
    fun = lambda x: ptref.equiproj_residuals(x, ref, disp, ncc, pairs, 0.7)  # noqa: E731
    Jf = ptref.lmdif_jacobian(fun, P)
    h = np.array([0.1, 0.1, 2e-12, 1e-19, 3e-27, 2e-10, 1e-10, 3e-10, 1e-3, 1e-3, 1e-4, 1e-5, 1e-4, 2e-5])
    for j in range(14):
        if j in (8, 9):  # columns that are zero (see the central-difference test)
            continue
        scale = np.abs(Jf[:, j]).max()
        err = np.abs(jac[:, j] - Jf[:, j]).max() / scale
        if j <= 4:
            # centre and radial terms: MINPACK's columns hold to about 1e-4
            assert err < 1e-3, (j, err)
        else:
            # tangential and parallax terms: MINPACK's step sqrt(eps) |x| is below the
            # residuals' rounding, its columns are off by 1e-3 to 1e-1; the analytic column
            # is the one that agrees with well-stepped central differences
            Pp, Pm = P.copy(), P.copy()
            Pp[j] += h[j]
            Pm[j] -= h[j]
            cd = (fun(Pp) - fun(Pm)) / (2 * h[j])
            assert np.abs(jac[:, j] - cd).max() < 1e-2 * np.abs(Jf[:, j] - cd).max(), j


def test_calibration_matches_minpack_fit_and_lens(tracker, views):
    from scipy.optimize import leastsq

    X, ref, disp, ncc, pairs = views
    tracker.equiproj_set_data(ref, disp, ncc, pairs, NCC_min=0.7)
    start = X.copy()
    start[2] = start[3] = 0.0
    free = np.zeros(14, dtype=np.int32)
    free[[2, 3]] = 1
    got, f0, f1, it = tracker.calibrate_distortion(start, free, max_iter=100)
    say(f"LM: {it} iterations, criterion {f0:.3e} -> {f1:.3e}, K1 {got[2]:.6e}, K2 {got[3]:.6e}")
    assert f1 < 1e-2 * f0  # down to the tracking-noise floor
    assert (got[[0, 1] + list(range(4, 14))] == start[[0, 1] + list(range(4, 14))]).all()

    def fun(q):
        Y = start.copy()
        Y[2], Y[3] = q[0] * 1e-9, q[1] * 1e-16
        return ptref.equiproj_residuals(Y, ref, disp, ncc, pairs, 0.7)

    q, _ = leastsq(fun, [0.0, 0.0], ftol=np.finfo(float).eps, xtol=np.finfo(float).eps, gtol=0.0)
    assert f1 <= (fun(q) ** 2).sum() * (1.0 + 1e-9)
    assert abs(got[2] - q[0] * 1e-9) < 1e-6 * abs(got[2]) and abs(got[3] - q[1] * 1e-16) < 1e-5 * abs(got[3])
    # 0.01 px of tracking noise on 70 grains leaves the fit about 0.5% from the lens
    assert abs(got[2] - X[2]) < 2e-2 * abs(X[2]) and abs(got[3] - X[3]) < 2e-2 * abs(X[3])


def test_differentiation_form_on_distorted_frames(tracker):
    # the config-4 lens on a small scene seen at four imposed shifts: track each image from
    # the first (search +-8 px, LM subpixel), then fit K1, K2 by the equiprojectivity
    # criterion; compared with the lens and with MINPACK on the same tracked positions
    from scipy.optimize import leastsq

    C = synth.calibration_set(n_grains=90, size=(1600, 1200), seed=1044, k1=-2e-8, k2=1e-14)
    say(f"calibration frames: {C.G} grains, {len(C.frames)} images")
    z = pk.zone(8, num_rot=0, inc_rot=1.0)
    prm = pk.track_params(z, rescue_level=0, subpixel=1, search_method=pk.SEARCH_EXACT,
                          subpix=pk.default_subpix(pk.SUBPIX_LM))
    tracker.set_image(0, C.frames[0], C.bits)
    disp, ncc = [], []
    for k in range(1, len(C.frames)):
        gr = C.grains.copy()
        gr["dx"], gr["dy"] = C.shifts[k]
        tracker.set_image(1, C.frames[k], C.bits)
        tracker.set_image_indices(0, 1)
        tracker.set_grains(gr, C.pattern_xy, C.pattern_begin)
        tracker.track_pair(prm)
        out = tracker.get_grains()
        disp.append(np.stack([out["dx"], out["dy"]], axis=1))
        ncc.append(out["NCC_subpix"])
    disp, ncc = np.array(disp), np.array(ncc)
    assert (ncc > 0.9).mean() > 0.9
    ref = np.stack([C.grains["refcoord_xpix"], C.grains["refcoord_ypix"]], axis=1).astype(np.float64)
    pairs = ptref.precompute_pairs(ref, 0.0, 1000.0)
    tracker.equiproj_set_data(ref, disp, ncc, pairs, NCC_min=0.7)
    start = C.params14.copy()
    start[2] = start[3] = 0.0
    free = np.zeros(14, dtype=np.int32)
    free[[2, 3]] = 1
    got, f0, f1, it = tracker.calibrate_distortion(start, free, max_iter=100)
    say(f"K1 {got[2]:.5e} (lens {C.params14[2]:.1e}), K2 {got[3]:.5e} (lens {C.params14[3]:.1e}), {it} iterations")

    def fun(q):
        Y = start.copy()
        Y[2], Y[3] = q[0] * 1e-8, q[1] * 1e-14
        return ptref.equiproj_residuals(Y, ref, disp, ncc, pairs, 0.7)

    q, _ = leastsq(fun, [0.0, 0.0], ftol=np.finfo(float).eps, xtol=np.finfo(float).eps, gtol=0.0)
    assert abs(got[2] - q[0] * 1e-8) < 1e-6 * abs(got[2]) and abs(got[3] - q[1] * 1e-14) < 1e-5 * abs(got[3])
    # K1 and K2 trade off over a small field; the fitted correction itself matches the lens
    # wherever there are grains (the lens moves those points by up to 5 px)
    xs = np.linspace(ref[:, 0].min(), ref[:, 0].max(), 9)
    ys = np.linspace(ref[:, 1].min(), ref[:, 1].max(), 9)
    diff, size = 0.0, 0.0
    for x in xs:
        for y in ys:
            fx, fy = ptref.undistor(got, x, y)
            lx, ly = ptref.undistor(C.params14, x, y)
            diff = max(diff, np.hypot(fx - lx, fy - ly))
            size = max(size, np.hypot(lx - x, ly - y))
    say(f"fitted vs lens correction: max {diff:.3f} px over the grains' region (lens: up to {size:.2f} px)")
    assert size > 3.0 and diff < 0.15
    assert abs(got[2] - C.params14[2]) < 0.1 * abs(C.params14[2])


def _calibration_program(tmp_path, algo, extra=""):
    """Writes the calibration set as TIFF frames with a command file (procedure
    correction_distortion, optim_algo algo) and runs the tracking program on them; returns the
    set, the frames as the program decodes them and the log's entries."""
    import os
    import subprocess

    from test_host_io import write_tiff

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "p-tracker_amd", "host", "ptracker_gpu")
    io = os.path.join(root, "p-tracker_amd", "host", "ptracker_io_check")
    C = synth.calibration_set(n_grains=60, size=(1400, 1000), seed=1045, k1=-2e-8, k2=1e-14)
    for k, f in enumerate(C.frames):
        write_tiff(tmp_path / f"cal{k}.tif", f.astype(np.int64) << 4, 1, 16)  # 12-bit values in 16-bit samples
    with open(tmp_path / "start.txt", "w") as fs:
        fs.write(f"{C.G}\n")
        for g in C.grains:
            fs.write(f"{int(g['refcoord_xpix'])} {int(g['refcoord_ypix'])} 0 {float(g['radius_pix'])!r}\n")
    shifts = " ".join(f"{x:g} {y:g}" for x, y in C.shifts[1:])
    (tmp_path / "commands.txt").write_text(f"""procedure correction_distortion
image_name cal%d.tif
grains_to_follow start.txt
rotations 0
subpixel 1
rescue_level 0
search_zone.left 8
search_zone.right 8
search_zone.up 8
search_zone.down 8
NCC_min 0.7
pattern grain_diameter 1.0 4
image_numbers_corrDisto {len(C.frames)} {' '.join(str(k) for k in range(len(C.frames)))}
imposedDisplApprox_corrDisto {shifts}
equiProj_dist_range 0 1000
optim_algo {algo}
K1_corrDistor 0
K2_corrDistor 0
""" + extra + "".join(f"d{n} 0\n" for n in ("xc_corrDistor", "yc_corrDistor", "K3_corrDistor", "P1_corrDistor",
                                             "P2_corrDistor", "P3_corrDistor", "xc_corrParallax", "yc_corrParallax",
                                             "angle1_corrParallax", "A1_corrParallax", "angle2_corrParallax",
                                             "A2_corrParallax")))
    say(f"program: correction_distortion, {algo}")
    r = subprocess.run([exe, "--lm", "commands.txt"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    log = {k.strip(): v for k, v in (l.rsplit(" ", 1) for l in (tmp_path / "corrDisto_log.txt").read_text()
                                      .splitlines() if l.strip())}
    frames = []
    for k in range(len(C.frames)):
        subprocess.run([io, "tiff", str(tmp_path / f"cal{k}.tif"), str(tmp_path / "f.raw")], check=True)
        b = (tmp_path / "f.raw").read_bytes()
        l1 = b.index(b"\n")
        l2 = b.index(b"\n", l1 + 1)
        w, h = map(int, b[:l1].split())
        frames.append(np.frombuffer(b[l2 + 1:], dtype="<u2").reshape(h, w).copy())
    return C, frames, log


def _calibration_data(tracker, C, frames):
    """The program's tracking steps through the C ABI: reference pixels, displacements,
    NCC_subpix and pairs."""
    pats = [synth.circ_pattern(max(int(r), 4)) for r in C.grains["radius_pix"]]
    pxy = np.ascontiguousarray(np.concatenate(pats), dtype=np.int32)
    pb = np.zeros(C.G + 1, dtype=np.int64)
    pb[1:] = np.cumsum([len(p) for p in pats])
    z = pk.zone(8, num_rot=0, inc_rot=1.0)
    prm = pk.track_params(z, rescue_level=0, subpixel=1, search_method=pk.SEARCH_EXACT,
                          subpix=pk.default_subpix(pk.SUBPIX_LM))
    tracker.set_image(0, frames[0], 16)
    disp, ncc = [], []
    for k in range(1, len(frames)):
        gr = C.grains.copy()
        gr["dx"], gr["dy"] = C.shifts[k]
        tracker.set_image(1, frames[k], 16)
        tracker.set_image_indices(0, 1)
        tracker.set_grains(gr, pxy, pb)
        tracker.track_pair(prm)
        out = tracker.get_grains()
        disp.append(np.stack([out["dx"], out["dy"]], axis=1))
        ncc.append(out["NCC_subpix"])
    ref = np.stack([C.grains["refcoord_xpix"], C.grains["refcoord_ypix"]], axis=1).astype(np.float64)
    return ref, np.array(disp), np.array(ncc), ptref.precompute_pairs(ref, 0.0, 1000.0)


def _lm_fit(tracker, C, ref, disp, ncc, pairs):
    tracker.equiproj_set_data(ref, disp, ncc, pairs, 0.7)
    free = np.zeros(14, dtype=np.int32)
    free[[2, 3]] = 1
    start = C.params14.copy()
    start[2] = start[3] = 0.0
    got, _, f_end, _ = tracker.calibrate_distortion(start, free, max_iter=200)
    return got, f_end


def _expected_field(X, dimx, dimy):
    """errors.txt and distobox.txt as the reference writes them (src/distortion_correction.cpp:
    515-608: default stream formatting, that of printf's %g), from the oracle's undistor."""
    g = lambda v: "%g" % v  # noqa: E731
    dxd, dyd = max(1.0, dimx / 50.0), max(1.0, dimy / 50.0)
    lines, box = [], []
    errs = []
    xd = 0.0
    while xd < dimx:
        yd = 0.0
        while yd < dimy:
            xu, yu = ptref.undistor(X, xd, yd)
            errs.append((xu - xd, yu - yd))
            lines.append(f"{g(xd)} {g(yd)} {g(xu - xd)} {g(yu - yd)}")
            yd += dyd
        lines.append("")
        xd += dxd
    pts = []
    xd = 0.0
    while xd < dimx:
        pts.append((xd, 0.0))
        xd += dxd
    yd = 0.0
    while yd < dimy:
        pts.append((dimx - 1.0, yd))
        yd += dyd
    xd = dimx - 1.0
    while xd >= 0:
        pts.append((xd, dimy - 1.0))
        xd -= dxd
    yd = dimy - 1.0
    while yd >= 0:
        pts.append((0.0, yd))
        yd -= dyd
    pts.append((0.0, 0.0))
    for x, y in pts:
        xu, yu = ptref.undistor(X, x, y)
        box.append(f"{g(x)} {g(y)} {g(xu - x)} {g(yu - y)}")
    e = np.array(errs)
    return "\n".join(lines) + "\n", "\n".join(box) + "\n", np.abs(e[:, 0]).max(), np.abs(e[:, 1]).max(), \
        np.hypot(e[:, 0], e[:, 1]).max()


def test_program_correction_distortion(tracker, tmp_path):
    # procedure correction_distortion of the tracking program on TIFF frames, LEVMAR: the
    # same fit as through the C ABI on the same decoded frames; the correction field files
    # byte for byte as the reference writes them from the oracle's undistor
    C, frames, log = _calibration_program(tmp_path, "LEVMAR")
    K1, K2 = float(log["K1_corrDistor"]), float(log["K2_corrDistor"])
    assert float(log["Final distorsion index"]) < float(log["Initial distorsion index"])
    assert (tmp_path / "dic_out_3.txt").exists()
    ref, disp, ncc, pairs = _calibration_data(tracker, C, frames)
    got, _ = _lm_fit(tracker, C, ref, disp, ncc, pairs)
    say(f"program K1 {K1:.6e} K2 {K2:.6e}; C ABI K1 {got[2]:.6e} K2 {got[3]:.6e}")
    assert abs(K1 - got[2]) <= 1e-9 * abs(got[2]) and abs(K2 - got[3]) <= 1e-9 * abs(got[3])
    assert abs(K1 - C.params14[2]) < 0.2 * abs(C.params14[2])
    # the held parameters are the set's own (centres at the frame centre, the rest 0); K1, K2
    # the fitted doubles (the log prints 16 digits)
    X = C.params14.copy()
    X[2], X[3] = got[2], got[3]
    for q, name in ((0, "xc_corrDistor"), (1, "yc_corrDistor"), (4, "K3_corrDistor"), (8, "xc_corrParallax"),
                    (11, "A1_corrParallax")):
        assert float(log[name]) == X[q], name
    errors, box, xm, ym, em = _expected_field(X, frames[0].shape[1], frames[0].shape[0])
    assert (tmp_path / "errors.txt").read_text() == errors
    assert (tmp_path / "distobox.txt").read_text() == box
    assert abs(float(log["max |xu - xd|"]) - xm) <= 1e-12 * xm and abs(float(log["max |yu - yd|"]) - ym) <= 1e-12 * ym
    assert abs(float(log["errormax"]) - em) <= 1e-12 * em


def test_program_correction_distortion_powell(tracker, tmp_path):
    # optim_algo POWELL: Powell's method from the perturbations reaches the minimum the
    # Levenberg-Marquardt fit finds on the same data
    C, frames, log = _calibration_program(tmp_path, "POWELL")
    K1, K2 = float(log["K1_corrDistor"]), float(log["K2_corrDistor"])
    ref, disp, ncc, pairs = _calibration_data(tracker, C, frames)
    got, f_lm = _lm_fit(tracker, C, ref, disp, ncc, pairs)
    f_pw = float(log["Final distorsion index"]) * (len(frames) - 1)
    say(f"Powell K1 {K1:.6e} K2 {K2:.6e}, f {f_pw:.6e}; LM K1 {got[2]:.6e} K2 {got[3]:.6e}, f {f_lm:.6e}")
    assert abs(K1 - got[2]) <= 1e-4 * abs(got[2]) and abs(K2 - got[3]) <= 1e-4 * abs(got[3])
    assert f_pw <= f_lm * (1 + 1e-8)
    assert (tmp_path / "errors.txt").exists() and (tmp_path / "distobox.txt").exists()


def test_program_correction_distortion_powell_iter(tracker, tmp_path):
    # optim_algo POWELL_ITER with 3 runs: each run's index and parameters (convergence.txt)
    # against the same loop in Python, each run's minimum found by MINPACK (scipy leastsq) on
    # the oracle's residuals of that form, the first positions re-centred between runs
    from scipy.optimize import leastsq

    C, frames, log = _calibration_program(tmp_path, "POWELL_ITER", extra="powell_num_iter 3\n")
    rows = [l.split() for l in (tmp_path / "convergence.txt").read_text().splitlines()]
    assert len(rows) == 3 and all(len(r) == 16 for r in rows)
    ref, disp0, ncc, pairs = _calibration_data(tracker, C, frames)
    nimg = len(frames) - 1
    X = C.params14.copy()
    X[2] = X[3] = 0.0
    orig = ref.copy()
    disp = disp0.copy()
    for it, row in enumerate(rows, start=1):
        def fun(k, orig=orig, disp=disp):
            P = X.copy()
            P[2], P[3] = k[0] * 1e-8, k[1] * 1e-14
            return ptref.equiproj_residuals_orig(P, orig, ref, disp, ncc, pairs, 0.7)
        k, _ = leastsq(fun, [X[2] / 1e-8, X[3] / 1e-14], xtol=1e-14, ftol=1e-15)
        X[2], X[3] = k[0] * 1e-8, k[1] * 1e-14
        idx = (fun(k) ** 2).sum() / nimg
        got = np.array([float(v) for v in row])
        say(f"run {it}: index {got[1]:.9e} (MINPACK {idx:.9e}), K1 {got[4]:.6e} ({X[2]:.6e}), K2 {got[5]:.6e} ({X[3]:.6e})")
        assert int(got[0]) == it
        assert abs(got[1] - idx) <= 1e-8 * idx
        assert abs(got[4] - X[2]) <= 1e-4 * abs(X[2]) and abs(got[5] - X[3]) <= 1e-4 * abs(X[3])
        # the program continues from its own parameters
        X[2], X[3] = got[4], got[5]
        und = np.array([ptref.undistor(X, x, y) for x, y in ref])
        disp = disp0 + (und - ref)[None, :, :]
        orig = und
    f_end = (ptref.equiproj_residuals_orig(X, orig, ref, disp, ncc, pairs, 0.7) ** 2).sum() / nimg
    assert abs(float(log["Final distorsion index"]) - f_end) <= 1e-8 * f_end
