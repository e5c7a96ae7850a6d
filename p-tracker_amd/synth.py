"""Seeded synthetic frame pairs for the tracking path.

Disk grains carry a speckle of Gaussian blobs over a base level on a flat background;
the second frame re-renders every grain translated and rotated by Catmull-Rom
resampling of the first, plus Gaussian noise. Grains are placed by rejection sampling
with a gap wide enough that no two grains touch in either frame.

Workloads (seed = 1000 + number):
  1  512x512, 8-bit values, 64 grains r ~ U(12, 20), 40 blobs (sigma 1.5, amplitude
     U(60, 200)) over base 40, background 10; motion U(-3, 3) px, U(-0.05, 0.05) rad,
     noise sigma 2; search +-5 px, no rotation steps, subpixel on.
  2  6000x4000, 12-bit values, 2000 grains, radius lognormal (median 20, sigma_log 0.4)
     clipped to [8, 40]; 0.1 blobs per px^2 of grain area, sigma U(1, 3); motion
     U(-15, 15) px, U(-0.1, 0.1) rad; noise sigma 8; search +-20 px x 21 steps of 0.01
     rad, rescue +-40 px for NCC < 0.7, subpixel on.
  3  the frames of 2; search +-20 px with 20 steps each side of 0.005 rad (the
     reference loop yields 40 steps); also an 8-bit variant (values >> 4).
  4  the frames of 2 warped by radial distortion k1 = -2e-9, k2 = 1e-16 about the frame
     centre (the scene, grain motion included, lives in undistorted coordinates; the grain
     records sit at the distorted images of the grain centres).
The grain radius gives the pattern as `pattern grain_diameter 1.0 1`
(src/p-tracker.cpp:1662-1667): a disk of radius int(radius).
"""
from __future__ import annotations

import math

import numpy as np

from . import GRAIN_DTYPE, SEARCH_PLAIN, SUBPIX_POWELL, default_subpix, track_params, zone


def circ_pattern(r: int) -> np.ndarray:
    """make_circ_pattern, src/p-tracker.cpp:1386-1408: dy outer, dx inner, |(dx,dy)| <= r."""
    d = np.arange(-r, r + 1)
    dy, dx = np.meshgrid(d, d, indexing="ij")
    keep = np.sqrt(dx.astype(np.float64) ** 2 + dy.astype(np.float64) ** 2) <= r
    return np.stack([dx[keep], dy[keep]], axis=1).astype(np.int32)


def cubic_sample(img: np.ndarray, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Catmull-Rom sample of img at (x, y), taps clamped to the frame."""
    H, W = img.shape
    ix = np.floor(x).astype(np.int64)
    iy = np.floor(y).astype(np.int64)
    tx = x - ix
    ty = y - iy

    def w(t):
        t2, t3 = t * t, t * t * t
        return np.stack([-0.5 * t3 + t2 - 0.5 * t, 1.5 * t3 - 2.5 * t2 + 1.0, -1.5 * t3 + 2.0 * t2 + 0.5 * t,
                         0.5 * t3 - 0.5 * t2])

    wx, wy = w(tx), w(ty)
    out = np.zeros_like(x, dtype=np.float64)
    for j in range(4):
        yy = np.clip(iy - 1 + j, 0, H - 1)
        row = np.zeros_like(out)
        for i in range(4):
            xx = np.clip(ix - 1 + i, 0, W - 1)
            row += wx[i] * img[yy, xx]
        out += wy[j] * row
    return out


def place_grains(rng, n, W, H, radii, margin, gap):
    """Rejection sampling of real-valued centres; disks at least `gap` apart."""
    order = np.argsort(-radii)
    cell = 2 * radii.max() + gap
    gw, gh = int(W // cell) + 1, int(H // cell) + 1
    buckets = {}
    centres = np.zeros((n, 2))
    for k in order:
        r = radii[k]
        for _ in range(20000):
            cx = rng.uniform(margin, W - 1 - margin)
            cy = rng.uniform(margin, H - 1 - margin)
            bx, by = int(cx // cell), int(cy // cell)
            ok = True
            for nx in range(bx - 1, bx + 2):
                for ny in range(by - 1, by + 2):
                    for j in buckets.get((nx, ny), ()):
                        if (centres[j, 0] - cx) ** 2 + (centres[j, 1] - cy) ** 2 < (radii[j] + r + gap) ** 2:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                centres[k] = (cx, cy)
                buckets.setdefault((bx, by), []).append(k)
                break
        else:
            raise RuntimeError("could not place all grains")
    return centres


def render_pair(rng, W, H, centres, radii, disp, rot, bits, background, base, amp_range, blobs_per_grain,
                sigma_range, noise_sigma):
    maxv = float((1 << bits) - 1)
    G = len(radii)
    field = np.zeros(H * W, dtype=np.float64)
    inside = np.zeros((H, W), dtype=bool)
    # blobs, rendered within 3.5 sigma of their centre
    bl_g, bl_x, bl_y, bl_a, bl_s = [], [], [], [], []
    for g in range(G):
        nb = blobs_per_grain[g]
        rr = radii[g] * np.sqrt(rng.uniform(0, 1, nb))
        th = rng.uniform(0, 2 * np.pi, nb)
        bl_x.append(centres[g, 0] + rr * np.cos(th))
        bl_y.append(centres[g, 1] + rr * np.sin(th))
        bl_a.append(rng.uniform(*amp_range, nb))
        bl_s.append(rng.uniform(*sigma_range, nb) if sigma_range[0] != sigma_range[1] else np.full(nb, sigma_range[0]))
        bl_g.append(np.full(nb, g))
    bx, by, ba, bs = (np.concatenate(v) for v in (bl_x, bl_y, bl_a, bl_s))
    R = int(math.ceil(3.5 * bs.max()))
    off = np.arange(-R, R + 1)
    oy, ox = np.meshgrid(off, off, indexing="ij")
    ox, oy = ox.ravel(), oy.ravel()
    chunk = max(1, 4_000_000 // len(ox))
    for s in range(0, len(bx), chunk):
        cx, cy, a, sg = bx[s:s + chunk], by[s:s + chunk], ba[s:s + chunk], bs[s:s + chunk]
        px = np.round(cx)[:, None].astype(np.int64) + ox[None, :]
        py = np.round(cy)[:, None].astype(np.int64) + oy[None, :]
        ok = (px >= 0) & (px < W) & (py >= 0) & (py < H)
        val = a[:, None] * np.exp(-((px - cx[:, None]) ** 2 + (py - cy[:, None]) ** 2) / (2 * sg[:, None] ** 2))
        field += np.bincount((py * W + px)[ok], weights=val[ok], minlength=H * W)
    field = field.reshape(H, W)
    # disks of frame 1
    for g in range(G):
        r = radii[g]
        x0, x1 = int(np.floor(centres[g, 0] - r)), int(np.ceil(centres[g, 0] + r))
        y0, y1 = int(np.floor(centres[g, 1] - r)), int(np.ceil(centres[g, 1] + r))
        yy, xx = np.mgrid[y0:y1 + 1, x0:x1 + 1]
        inside[y0:y1 + 1, x0:x1 + 1] |= (xx - centres[g, 0]) ** 2 + (yy - centres[g, 1]) ** 2 <= r * r
    f1 = np.where(inside, base + field, background)
    I0 = np.clip(np.round(f1), 0, maxv).astype(np.uint16)
    return I0, resample_grains(rng, I0, centres, radii, disp, rot, bits, background, noise_sigma)


def resample_grains(rng, I0, centres, radii, disp, rot, bits, background, noise_sigma):
    """The second frame: each grain of I0 moved by disp and rot (Catmull-Rom resampling
    about its centre) on a flat background, plus Gaussian noise."""
    H, W = I0.shape
    G = len(radii)
    maxv = float((1 << bits) - 1)
    f2 = np.full((H, W), float(background))
    I0f = I0.astype(np.float64)
    for g in range(G):
        r = radii[g]
        c = centres[g] + disp[g]
        x0, x1 = int(np.floor(c[0] - r)), int(np.ceil(c[0] + r))
        y0, y1 = int(np.floor(c[1] - r)), int(np.ceil(c[1] + r))
        yy, xx = np.mgrid[y0:y1 + 1, x0:x1 + 1]
        m = (xx - c[0]) ** 2 + (yy - c[1]) ** 2 <= r * r
        px, py = xx[m] - c[0], yy[m] - c[1]
        ct, st = math.cos(rot[g]), math.sin(rot[g])
        # the tracker maps pattern offset (x, y) to (c x + s y, -s x + c y); invert it
        qx = centres[g, 0] + ct * px - st * py
        qy = centres[g, 1] + st * px + ct * py
        f2[yy[m], xx[m]] = cubic_sample(I0f, qx, qy)
    f2 += rng.normal(0.0, noise_sigma, size=f2.shape)
    return np.clip(np.round(f2), 0, maxv).astype(np.uint16)


def make_grains(centres, radii):
    G = len(radii)
    gr = np.zeros(G, dtype=GRAIN_DTYPE)
    fl = np.floor(centres)
    gr["refcoord_xpix"] = fl[:, 0].astype(np.int32)
    gr["refcoord_ypix"] = fl[:, 1].astype(np.int32)
    gr["radius_pix"] = radii
    # simple start format: the fractional part of the centre goes into dx (src/p-tracker.cpp:1265-1270)
    gr["dx"] = centres[:, 0] - fl[:, 0]
    gr["dy"] = centres[:, 1] - fl[:, 1]
    gr["upix"] = gr["dx"]
    gr["vpix"] = gr["dy"]
    return gr


def make_patterns(radii, reduction=1.0, min_radius=1):
    """`pattern grain_diameter reduction min_radius` (src/p-tracker.cpp:1658-1668)."""
    pats = []
    for r in radii:
        rad = int(r * reduction)
        if rad < min_radius:
            rad = int(min_radius)
        pats.append(circ_pattern(rad))
    pb = np.zeros(len(pats) + 1, dtype=np.int64)
    pb[1:] = np.cumsum([len(p) for p in pats])
    return np.ascontiguousarray(np.concatenate(pats), dtype=np.int32), pb


class SynthWorkload:
    def __init__(self, I0, I1, grains, pattern_xy, pattern_begin, bits, truth, params, name):
        self.I0, self.I1, self.grains = I0, I1, grains
        self.pattern_xy, self.pattern_begin = pattern_xy, pattern_begin
        self.bits, self.truth, self.params, self.name = bits, truth, params, name

    @property
    def G(self):
        return len(self.grains)


def workload(cfg: int, seed: int | None = None, n_grains: int | None = None, size=None,
             search_method=SEARCH_PLAIN, subpix_method=SUBPIX_POWELL, max_disp=None, far=None,
             lens=None) -> SynthWorkload:
    """Builds workload `cfg` (1-4). n_grains / size shrink it for tests; max_disp overrides
    the displacement range (to force rescues, for example). far = (n, lo, hi), configs 2-4:
    n grains drawn at random move by U(lo, hi) px, either sign, along one axis drawn at
    random (the other keeps its U(-15, 15)), e.g. beyond the main search zone but inside
    the rescue zone; truth["far"] lists them. lens = (k1, k2) replaces config 4's radial lens."""
    seed = 1000 + cfg if seed is None else seed
    rng = np.random.default_rng(seed)
    if cfg == 1:
        W, H = size or (512, 512)
        G = n_grains or 64
        bits = 8
        radii = rng.uniform(12, 20, G)
        md = 3.0 if max_disp is None else max_disp
        margin = 40
        centres = place_grains(rng, G, W, H, radii, margin, gap=2 * md + 4)
        disp = rng.uniform(-md, md, (G, 2))
        rot = rng.uniform(-0.05, 0.05, G)
        I0, I1 = render_pair(rng, W, H, centres, radii, disp, rot, bits, background=10, base=40,
                             amp_range=(60, 200), blobs_per_grain=np.full(G, 40), sigma_range=(1.5, 1.5),
                             noise_sigma=2.0)
        z = zone(5, num_rot=0, inc_rot=1.0)  # `rotations 0` (src/p-tracker.cpp:1191-1201)
        prm = track_params(z, zone(5, num_rot=0, inc_rot=1.0), zone(5, num_rot=0, inc_rot=1.0), rescue_level=0,
                           subpixel=1, search_method=search_method, subpix=default_subpix(subpix_method))
    elif cfg in (2, 3, 4):
        W, H = size or (6000, 4000)
        G = n_grains or 2000
        bits = 12
        radii = np.clip(np.exp(rng.normal(np.log(20.0), 0.4, G)), 8.0, 40.0)
        md = 15.0 if max_disp is None else max_disp
        margin = min(110, (min(W, H) - 1) // 4)
        centres = place_grains(rng, G, W, H, radii, margin, gap=2 * md + 4)
        disp = rng.uniform(-md, md, (G, 2))
        rot = rng.uniform(-0.1, 0.1, G)
        far_idx = np.zeros(0, dtype=np.int64)
        if far is not None:
            nf, lo, hi = far
            frng = np.random.default_rng(seed + 7919)  # leaves the main stream's draws unchanged
            far_idx = np.sort(frng.choice(G, size=nf, replace=False))
            axis = frng.integers(0, 2, nf)
            mag = frng.uniform(lo, hi, nf) * np.where(frng.uniform(0, 1, nf) < 0.5, -1.0, 1.0)
            disp[far_idx, axis] = mag
        nblobs = np.maximum(1, np.round(0.1 * np.pi * radii ** 2)).astype(int)
        I0, I1 = render_pair(rng, W, H, centres, radii, disp, rot, bits, background=160, base=640,
                             amp_range=(960, 3200), blobs_per_grain=nblobs, sigma_range=(1.0, 3.0),
                             noise_sigma=8.0)
        if cfg == 3:
            z = zone(20, num_rot=20, inc_rot=0.005)
        else:
            z = zone(20, num_rot=10, inc_rot=0.01)
        zr = zone(40, num_rot=z.num_rot, inc_rot=z.inc_rot)
        prm = track_params(z, zr, zr, NCC_min=0.7, rescue_level=1, subpixel=1, search_method=search_method,
                           subpix=default_subpix(subpix_method))
        lens14 = None
        if cfg == 4:
            k1, k2 = (-2e-9, 1e-16) if lens is None else lens
            lens14 = np.array([0.5 * W, 0.5 * H, k1, k2, 0, 0, 0, 0, 0.5 * W, 0.5 * H, 0, 0, 0, 0], dtype=np.float64)
            I0 = distort_frame(I0, lens14, bits)
            I1 = distort_frame(I1, lens14, bits)
    else:
        raise ValueError(f"unknown workload {cfg}")
    # the grains sit where the frames show them: through the lens for config 4 (the scene's
    # grain centres are undistorted coordinates)
    seen = centres if cfg != 4 else np.array([_distort_point(lens14, c[0], c[1]) for c in centres])
    grains = make_grains(seen, radii)
    pxy, pb = make_patterns(radii)
    truth = {"centres": centres, "disp": disp, "rot": rot, "radii": radii}
    if cfg in (2, 3, 4):
        truth["far"] = far_idx
    if cfg == 4:
        truth["lens"] = lens14  # undistor's 14 parameters; centres, disp, rot are undistorted coordinates
    return SynthWorkload(I0, I1, grains, pxy, pb, bits, truth, prm, f"cfg{cfg}")


def distort_frame(img, X, bits):
    """The frame as seen through the lens: the pixel at distorted position d shows the
    undistorted scene at undistor(d) (src/distortion_correction.cpp:20-61, radial part)."""
    H, W = img.shape
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    dx, dy = xx - X[0], yy - X[1]
    r2 = dx * dx + dy * dy
    k = X[2] * r2 + X[3] * r2 * r2
    xu, yu = xx + dx * k, yy + dy * k
    out = cubic_sample(img.astype(np.float64), xu, yu)
    return np.clip(np.round(out), 0, (1 << bits) - 1).astype(np.uint16)


class SynthSequence:
    """A frame sequence: frames[k] (uint16, H x W), the grain records of frame 0, the
    patterns, and the true centre and rotation of every grain in every frame."""

    def __init__(self, frames, grains, pattern_xy, pattern_begin, bits, centres, rots, params, name):
        self.frames, self.grains = frames, grains
        self.pattern_xy, self.pattern_begin = pattern_xy, pattern_begin
        self.bits, self.centres, self.rots, self.params, self.name = bits, centres, rots, params, name

    @property
    def G(self):
        return len(self.grains)


def texture_sheet(rng, W, H, bits, sigma=1.5, mean_frac=0.5, std_frac=0.17):
    """Gaussian-filtered white noise over the frame, scaled to the bit depth: the speckle
    each grain carries in its own body frame (grain g shows the sheet about its frame-0
    centre)."""
    from scipy.ndimage import gaussian_filter

    maxv = float((1 << bits) - 1)
    s = gaussian_filter(rng.standard_normal((H, W)).astype(np.float32), sigma)
    s *= np.float32(std_frac * maxv / max(float(s.std()), 1e-12))
    s += np.float32(mean_frac * maxv)
    return s


def _bilinear(img, x, y):
    x0 = np.floor(x).astype(np.int64)
    y0 = np.floor(y).astype(np.int64)
    tx, ty = (x - x0).astype(np.float32), (y - y0).astype(np.float32)
    H, W = img.shape
    x0 = np.clip(x0, 0, W - 2)
    y0 = np.clip(y0, 0, H - 2)
    a, b = img[y0, x0], img[y0, x0 + 1]
    c, d = img[y0 + 1, x0], img[y0 + 1, x0 + 1]
    return (a + (b - a) * tx) * (1 - ty) + (c + (d - c) * tx) * ty


def render_sequence_frame(sheet, centres0, centres, rots, radii, order, bits, background, noise_sigma, rng):
    """One frame: every grain drawn as an opaque disk at its centre and rotation, in
    `order` (later disks cover earlier ones). Frame pixel x of grain g shows the sheet at
    centres0[g] + R(rot) (x - centres[g]), R(t) = [[cos t, -sin t], [sin t, cos t]], which
    is the motion the tracker's rotation (c x + s y, -s x + c y) undoes."""
    H, W = sheet.shape
    f = np.full((H, W), np.float32(background), dtype=np.float32)
    for g in order:
        r = radii[g]
        cx, cy = centres[g]
        x0, x1 = max(0, int(np.floor(cx - r))), min(W - 1, int(np.ceil(cx + r)))
        y0, y1 = max(0, int(np.floor(cy - r))), min(H - 1, int(np.ceil(cy + r)))
        yy, xx = np.mgrid[y0:y1 + 1, x0:x1 + 1]
        px, py = xx - cx, yy - cy
        m = px * px + py * py <= r * r
        px, py = px[m], py[m]
        ct, st = math.cos(rots[g]), math.sin(rots[g])
        qx = centres0[g, 0] + ct * px - st * py
        qy = centres0[g, 1] + st * px + ct * py
        f[yy[m], xx[m]] = _bilinear(sheet, qx, qy)
    f += rng.normal(0.0, noise_sigma, size=f.shape).astype(np.float32)
    return np.clip(np.round(f), 0, (1 << bits) - 1).astype(np.uint16)


def sequence(n_frames=16, n_grains=10000, size=(6000, 4000), seed=1005, radius=(6.0, 60.0), max_step=5.0,
             max_rot=0.03, margin=160, bits=12, noise_sigma=8.0, search_method=SEARCH_PLAIN,
             subpix_method=SUBPIX_POWELL) -> SynthSequence:
    """Workload 5, the batching form over a sequence: n_frames frames of W x H with
    n_grains disks of radius U(radius), each moved per frame by U(-max_step, max_step) px
    and U(-max_rot, max_rot) rad, accumulated. At the full size the disks cover 1.74 times
    the frame area (mean area pi E[r^2] = 4184 px^2), so they cannot be disjoint: they are
    drawn largest first, each over the larger ones, in every frame. Tracker: consecutive
    pairs (the reference swaps the reference frame after each pair), search +-8 px x 9
    steps of 0.01 rad (num_rot 4), no rescue, subpixel on."""
    rng = np.random.default_rng(seed)
    W, H = size
    G = n_grains
    radii = rng.uniform(radius[0], radius[1], G)
    m = min(margin, (min(W, H) - 1) // 4)
    c0 = np.column_stack([rng.uniform(m, W - 1 - m, G), rng.uniform(m, H - 1 - m, G)])
    steps = rng.uniform(-max_step, max_step, (n_frames - 1, G, 2))
    rsteps = rng.uniform(-max_rot, max_rot, (n_frames - 1, G))
    centres = np.concatenate([c0[None], c0[None] + np.cumsum(steps, axis=0)], axis=0)
    rots = np.concatenate([np.zeros((1, G)), np.cumsum(rsteps, axis=0)], axis=0)
    sheet = texture_sheet(rng, W, H, bits)
    order = np.argsort(-radii, kind="stable")
    frames = [render_sequence_frame(sheet, c0, centres[k], rots[k], radii, order, bits, 160.0, noise_sigma, rng)
              for k in range(n_frames)]
    grains = make_grains(c0, radii)
    pxy, pb = make_patterns(radii)
    z = zone(8, num_rot=4, inc_rot=0.01)
    prm = track_params(z, z, z, NCC_min=0.0, rescue_level=0, subpixel=1, search_method=search_method,
                       subpix=default_subpix(subpix_method))
    return SynthSequence(frames, grains, pxy, pb, bits, centres, rots, prm, "cfg5")


class CalibrationSet:
    """Frames of one static scene at imposed rigid shifts, seen through a lens: frames[k]
    shows the scene moved by shifts[k] (shifts[0] = 0), each warped by the radial
    distortion params14 (undistor of a frame pixel gives its scene position)."""

    def __init__(self, frames, grains, pattern_xy, pattern_begin, bits, shifts, params14, truth):
        self.frames, self.grains = frames, grains
        self.pattern_xy, self.pattern_begin = pattern_xy, pattern_begin
        self.bits, self.shifts, self.params14, self.truth = bits, shifts, params14, truth

    @property
    def G(self):
        return len(self.grains)


def calibration_set(n_grains=2000, size=(6000, 4000), seed=1004, shifts=((0, 0), (36, 11), (-29, 33), (22, -38)),
                    k1=-2e-9, k2=1e-16, margin=240, noise_sigma=8.0) -> CalibrationSet:
    """Workload 4 for the distortion calibration (the reference's correction_distortion,
    src/distortion_correction.cpp:135-479): the config-2 grains and texture, the whole
    scene translated by each shift (the imposed displacements), every frame warped by
    radial distortion k1, k2 about the frame centre. Grains start with dx, dy = 0 at
    integer reference pixels; the tracker gets the shift as its starting displacement."""
    rng = np.random.default_rng(seed)
    W, H = size
    G = n_grains
    bits = 12
    radii = np.clip(np.exp(rng.normal(np.log(20.0), 0.4, G)), 8.0, 40.0)
    m = min(margin, (min(W, H) - 1) // 4)
    centres = place_grains(rng, G, W, H, radii, m, gap=4)
    nblobs = np.maximum(1, np.round(0.1 * np.pi * radii ** 2)).astype(int)
    zero = np.zeros((G, 2))
    I0, _ = render_pair(rng, W, H, centres, radii, zero, np.zeros(G), bits, background=160, base=640,
                        amp_range=(960, 3200), blobs_per_grain=nblobs, sigma_range=(1.0, 3.0), noise_sigma=noise_sigma)
    X = np.array([0.5 * W, 0.5 * H, k1, k2, 0, 0, 0, 0, 0.5 * W, 0.5 * H, 0, 0, 0, 0], dtype=np.float64)
    frames = []
    for t in shifts:
        moved = resample_grains(rng, I0, centres, radii, np.tile(np.asarray(t, dtype=np.float64), (G, 1)),
                                np.zeros(G), bits, 160, noise_sigma)
        frames.append(distort_frame(moved, X, bits))
    # reference pixels: the distorted images of the grain centres, rounded
    gr = np.zeros(G, dtype=GRAIN_DTYPE)
    cd = np.array([_distort_point(X, c[0], c[1]) for c in centres])
    gr["refcoord_xpix"] = np.round(cd[:, 0]).astype(np.int32)
    gr["refcoord_ypix"] = np.round(cd[:, 1]).astype(np.int32)
    gr["radius_pix"] = radii
    pxy, pb = make_patterns(np.maximum(4.0, radii - 2.0))
    return CalibrationSet(frames, gr, pxy, pb, bits, np.asarray(shifts, dtype=np.float64), X,
                          {"centres": centres, "radii": radii})


def _distort_point(X, xu, yu):
    """The frame pixel whose undistor image (radial part) is (xu, yu): Newton iterations."""
    x, y = xu, yu
    for _ in range(30):
        dx, dy = x - X[0], y - X[1]
        r2 = dx * dx + dy * dy
        k = X[2] * r2 + X[3] * r2 * r2
        fx, fy = x + dx * k - xu, y + dy * k - yu
        kp = X[2] + 2.0 * X[3] * r2
        j11 = 1.0 + k + 2.0 * dx * dx * kp
        j22 = 1.0 + k + 2.0 * dy * dy * kp
        j12 = 2.0 * dx * dy * kp
        det = j11 * j22 - j12 * j12
        x -= (j22 * fx - j12 * fy) / det
        y -= (j11 * fy - j12 * fx) / det
    return x, y
