"""ctypes binding of the oracle (libptref.so) and of oracle/_ref (the reference's own
compiled sampler).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg, always as the checker of libptgpu, never as the product.
"""
from __future__ import annotations

import ctypes
import math
import importlib.util
import os
import sys
from ctypes import c_double, c_int32, c_int64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libptref.so")
REF_INTERP = os.path.join(HERE, "_ref", "libref_interp.so")


def package():
    """The ptgpu Python package (directory p-tracker_amd), imported as `ptracker_amd`."""
    if "ptracker_amd" in sys.modules:
        return sys.modules["ptracker_amd"]
    pdir = os.path.join(ROOT, "p-tracker_amd")
    spec = importlib.util.spec_from_file_location("ptracker_amd", os.path.join(pdir, "__init__.py"),
                                                  submodule_search_locations=[pdir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ptracker_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"oracle not built: {LIB} (make -C oracle)")
        L = ctypes.CDLL(LIB)
        vp = c_void_p
        sig = {
            "ptref_rotation_steps": ([c_double, c_int32, vp, c_int32], ctypes.c_int),
            "ptref_make_circ_pattern": ([c_int32, vp], ctypes.c_int),
            "ptref_make_rect_pattern": ([c_int32, c_int32, vp], ctypes.c_int),
            "ptref_make_ring_pattern": ([c_int32, c_int32, vp], ctypes.c_int),
            "ptref_rotate_pattern": ([c_int32, vp, c_double, c_double, vp], None),
            "ptref_precompute": ([vp, c_int64, vp, c_int32, vp, vp], None),
            "ptref_follow_pattern": ([vp, vp, c_int64, vp, c_int32, vp, vp, vp, ctypes.c_int, c_double, c_double, vp],
                                     ctypes.c_int),
            "ptref_candidate_ncc": ([vp, vp, c_int64, vp, c_int32, vp, vp, vp, vp, c_int64], ctypes.c_int),
            "ptref_interp_cubic": ([vp, c_int64, c_double, c_double], c_double),
            "ptref_interp_cubic_batch": ([vp, c_int64, c_int32, vp, vp], None),
            "ptref_ncc_objective": ([vp, vp, c_int64, vp, c_int32, vp, vp, c_double, c_double], c_double),
            "ptref_joint_objective": ([vp, vp, c_int64, vp, c_int32, vp, vp, vp], c_double),
            "ptref_joint_objective_batch": ([vp, vp, c_int64, vp, vp, vp, c_int32, vp, vp, vp, vp, c_int32], None),
            "ptref_subpixel": ([vp, vp, c_int64, vp, c_int32, vp, vp, vp], None),
            "ptref_track_pair": ([vp, vp, c_int64, c_int32, vp, vp, vp, vp, c_int32, vp, vp], ctypes.c_int),
            "ptref_change_reference": ([c_int32, vp], None),
            "ptref_undistor": ([vp, c_double, c_double, vp, vp], None),
            "ptref_remap": ([vp, c_int32, c_int32, c_int64, c_int32, vp, vp, c_int64], None),
            "ptref_distort_point": ([vp, c_double, c_double, vp, vp], None),
            "ptref_precompute_pairs": ([c_int32, vp, c_double, c_double, vp, c_int64], c_int64),
            "ptref_equiproj_residuals": ([vp, c_int32, vp, c_int32, vp, vp, c_int32, vp, c_double, vp], None),
            "ptref_equiproj_residuals_orig": ([vp, c_int32, vp, vp, c_int32, vp, vp, c_int32, vp, c_double, vp],
                                              None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(c_void_p)


def _rec(grains, g):
    """Pointer to record g of a grain array."""
    return c_void_p(int(grains.ctypes.data) + int(g) * grains.itemsize)


def rotation_steps(inc_rot, num_rot):
    n = lib().ptref_rotation_steps(inc_rot, num_rot, None, 0)
    if n < 0:
        raise ValueError("inc_rot must be > 0")
    out = np.zeros(max(n, 1), dtype=np.float64)
    lib().ptref_rotation_steps(inc_rot, num_rot, _p(out), n)
    return out[:n]


def circ_pattern(r):
    n = lib().ptref_make_circ_pattern(r, None)
    out = np.zeros((n, 2), dtype=np.int32)
    lib().ptref_make_circ_pattern(r, _p(out))
    return out


def rect_pattern(hw, hh):
    n = lib().ptref_make_rect_pattern(hw, hh, None)
    out = np.zeros((n, 2), dtype=np.int32)
    lib().ptref_make_rect_pattern(hw, hh, _p(out))
    return out


def ring_pattern(rin, rout):
    n = lib().ptref_make_ring_pattern(rin, rout, None)
    out = np.zeros((max(n, 1), 2), dtype=np.int32)
    lib().ptref_make_ring_pattern(rin, rout, _p(out))
    return out[:n]


def rotate_pattern(xy, c, s):
    xy = np.ascontiguousarray(xy, dtype=np.int32)
    out = np.zeros_like(xy)
    lib().ptref_rotate_pattern(len(xy), _p(xy), c, s, _p(out))
    return out


def interp_cubic(frame, xy):
    f = np.ascontiguousarray(frame, dtype=np.uint16)
    xy = np.ascontiguousarray(np.asarray(xy, dtype=np.float64).reshape(-1, 2))
    out = np.zeros(len(xy))
    lib().ptref_interp_cubic_batch(_p(f), f.shape[1], len(xy), _p(xy), _p(out))
    return out


def ref_interp_available():
    return os.path.exists(REF_INTERP)


def ref_interp_cubic(frame, xy):
    """The reference's own image_interpCubic (src/interpol.cpp), compiled into oracle/_ref."""
    L = ctypes.CDLL(REF_INTERP)
    f = np.ascontiguousarray(frame, dtype=np.uint16)
    xy = np.ascontiguousarray(np.asarray(xy, dtype=np.float64).reshape(-1, 2))
    out = np.zeros(len(xy))
    L.refimpl_interp_cubic.argtypes = [c_void_p, ctypes.c_int, ctypes.c_int, c_int64, ctypes.c_int, c_void_p,
                                       c_void_p]
    L.refimpl_interp_cubic(_p(f), f.shape[1], f.shape[0], f.shape[1], len(xy), _p(xy), _p(out))
    return out


class Workload:
    """Frames + grain records + patterns in the layout both libraries take."""

    def __init__(self, I0, I1, grains, pattern_xy, pattern_begin, bits=16):
        self.I0 = np.ascontiguousarray(I0, dtype=np.uint16)
        self.I1 = np.ascontiguousarray(I1, dtype=np.uint16)
        self.grains = np.ascontiguousarray(grains, dtype=package().GRAIN_DTYPE)
        self.pattern_xy = np.ascontiguousarray(pattern_xy, dtype=np.int32)
        self.pattern_begin = np.ascontiguousarray(pattern_begin, dtype=np.int64)
        self.bits = bits

    @property
    def pitch(self):
        return self.I0.shape[1]

    def pattern(self, g):
        b, e = self.pattern_begin[g], self.pattern_begin[g + 1]
        return self.pattern_xy[b:e]


def precompute(w: Workload, grains=None):
    """do_precomputations on every grain; returns (grains, p0_xy)."""
    gr = (w.grains if grains is None else grains).copy()
    p0 = np.zeros_like(w.pattern_xy)
    for g in range(len(gr)):
        b, e = w.pattern_begin[g], w.pattern_begin[g + 1]
        lib().ptref_precompute(_p(w.I0), w.pitch, _rec(gr, g), int(e - b),
                               _p(w.pattern_xy[b:e]), _p(p0[b:e]))
    return gr, p0


def follow_pattern(w: Workload, grains, p0, g, z, stage, NCC_min=0.7, NCC_min_super=0.5):
    b, e = w.pattern_begin[g], w.pattern_begin[g + 1]
    scratch = np.zeros((e - b, 2), dtype=np.int32)
    return lib().ptref_follow_pattern(_p(w.I0), _p(w.I1), w.pitch, _rec(grains, g),
                                      int(e - b), _p(w.pattern_xy[b:e]), _p(p0[b:e]), ctypes.byref(z), stage,
                                      NCC_min, NCC_min_super, _p(scratch))


def candidate_ncc(w: Workload, grains, p0, g, z):
    b, e = w.pattern_begin[g], w.pattern_begin[g + 1]
    n = lib().ptref_candidate_ncc(_p(w.I0), _p(w.I1), w.pitch, _rec(grains, g),
                                  int(e - b), _p(w.pattern_xy[b:e]), _p(p0[b:e]), ctypes.byref(z), None, 0)
    out = np.zeros(n)
    lib().ptref_candidate_ncc(_p(w.I0), _p(w.I1), w.pitch, _rec(grains, g),
                              int(e - b), _p(w.pattern_xy[b:e]), _p(p0[b:e]), ctypes.byref(z), _p(out), n)
    return out


def ncc_objective(w: Workload, grains, p0, g, X, dmax=3.0, stiffness=0.0):
    b, e = w.pattern_begin[g], w.pattern_begin[g + 1]
    X = np.ascontiguousarray(X, dtype=np.float64)
    return lib().ptref_ncc_objective(_p(w.I0), _p(w.I1), w.pitch, _rec(grains, g),
                                     int(e - b), _p(p0[b:e]), _p(X), dmax, stiffness)


def joint_objective(w: Workload, grains, p0, g, P, lens):
    """The differentiation form's joint objective of grain g: P = (Tx, Ty, theta) in undistorted
    coordinates, lens = (K1, K2, xc, yc); see ptref_joint_objective."""
    b, e = w.pattern_begin[g], w.pattern_begin[g + 1]
    P = np.ascontiguousarray(P, dtype=np.float64)
    L = np.ascontiguousarray(lens, dtype=np.float64)
    return lib().ptref_joint_objective(_p(w.I0), _p(w.I1), w.pitch, _rec(grains, g), int(e - b), _p(p0[b:e]),
                                       _p(P), _p(L))


def joint_objective_batch(w: Workload, grains, p0, glist, P, lens, nthreads=0):
    """joint_objective for grains glist at P (len(glist) x 3), OpenMP over grains."""
    gl = np.ascontiguousarray(glist, dtype=np.int32)
    Pa = np.ascontiguousarray(np.asarray(P, dtype=np.float64).reshape(-1, 3))
    L = np.ascontiguousarray(lens, dtype=np.float64)
    gr = np.ascontiguousarray(grains)
    f = np.zeros(len(gl))
    lib().ptref_joint_objective_batch(_p(w.I0), _p(w.I1), w.pitch, _p(gr), _p(p0), _p(w.pattern_begin), len(gl),
                                      _p(gl), _p(Pa), _p(L), _p(f), nthreads)
    return f


def subpixel(w: Workload, grains, p0, g, prm):
    b, e = w.pattern_begin[g], w.pattern_begin[g + 1]
    ne = c_int64(0)
    lib().ptref_subpixel(_p(w.I0), _p(w.I1), w.pitch, _rec(grains, g),
                         int(e - b), _p(p0[b:e]), ctypes.byref(prm), ctypes.byref(ne))
    return ne.value


def track_pair(w: Workload, prm, grains=None, nthreads=0, subset=None):
    """One frame pair through the oracle. Returns (grains, (n_rescued, n_super), n_evals).
    `subset` restricts the run to those grain indices (used to bound CPU-baseline time)."""
    gr = (w.grains if grains is None else grains).copy()
    pxy, pb = w.pattern_xy, w.pattern_begin
    if subset is not None:
        subset = np.asarray(subset)
        gr = gr[subset].copy()
        parts = [w.pattern(g) for g in subset]
        pxy = np.ascontiguousarray(np.concatenate(parts), dtype=np.int32)
        pb = np.zeros(len(subset) + 1, dtype=np.int64)
        pb[1:] = np.cumsum([len(p) for p in parts])
    counts = np.zeros(2, dtype=np.int32)
    ne = c_int64(0)
    lib().ptref_track_pair(_p(w.I0), _p(w.I1), w.pitch, len(gr), _p(gr), _p(pxy), _p(pb), ctypes.byref(prm), nthreads,
                           _p(counts), ctypes.byref(ne))
    return gr, (int(counts[0]), int(counts[1])), ne.value


def change_reference(grains):
    gr = grains.copy()
    lib().ptref_change_reference(len(gr), _p(gr))
    return gr


def undistor(X14, xd, yd):
    X = np.ascontiguousarray(X14, dtype=np.float64)
    xu, yu = c_double(), c_double()
    lib().ptref_undistor(_p(X), xd, yd, ctypes.byref(xu), ctypes.byref(yu))
    return xu.value, yu.value


def distort_point(X14, xu, yu):
    X = np.ascontiguousarray(X14, dtype=np.float64)
    xd, yd = c_double(), c_double()
    lib().ptref_distort_point(_p(X), xu, yu, ctypes.byref(xd), ctypes.byref(yd))
    return xd.value, yd.value


def remap(src, bits, X14):
    s = np.ascontiguousarray(src, dtype=np.uint16)
    X = np.ascontiguousarray(X14, dtype=np.float64)
    dst = np.zeros_like(s)
    lib().ptref_remap(_p(s), s.shape[1], s.shape[0], s.shape[1], bits, _p(X), _p(dst), s.shape[1])
    return dst


def precompute_pairs(ref_xy, dmin, dmax):
    """Grain pairs (i < j) at reference distance in [dmin, dmax] (npairs x 2)."""
    ref = np.ascontiguousarray(np.asarray(ref_xy, dtype=np.float64).reshape(-1, 2))
    n = lib().ptref_precompute_pairs(len(ref), _p(ref), dmin, dmax, None, 0)
    out = np.zeros((n, 2), dtype=np.int32)
    lib().ptref_precompute_pairs(len(ref), _p(ref), dmin, dmax, _p(out), n)
    return out


def equiproj_residuals(X14, ref_xy, disp, ncc, pairs, NCC_min=0.7):
    """The equiprojectivity residuals (nimg * npairs, image major)."""
    X = np.ascontiguousarray(X14, dtype=np.float64)
    ref = np.ascontiguousarray(np.asarray(ref_xy, dtype=np.float64).reshape(-1, 2))
    G = len(ref)
    d = np.ascontiguousarray(np.asarray(disp, dtype=np.float64).reshape(-1, G, 2))
    n = np.ascontiguousarray(np.asarray(ncc, dtype=np.float64).reshape(-1, G))
    pr = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    res = np.zeros(len(d) * len(pr))
    lib().ptref_equiproj_residuals(_p(X), G, _p(ref), len(d), _p(d), _p(n), len(pr), _p(pr), NCC_min, _p(res))
    return res


def equiproj_residuals_orig(X14, orig_xy, ref_xy, disp, ncc, pairs, NCC_min=0.7):
    """The POWELL_ITER residuals: the first positions given (orig_xy), the tracked ones undistorted."""
    X = np.ascontiguousarray(X14, dtype=np.float64)
    ref = np.ascontiguousarray(np.asarray(ref_xy, dtype=np.float64).reshape(-1, 2))
    org = np.ascontiguousarray(np.asarray(orig_xy, dtype=np.float64).reshape(-1, 2))
    G = len(ref)
    d = np.ascontiguousarray(np.asarray(disp, dtype=np.float64).reshape(-1, G, 2))
    n = np.ascontiguousarray(np.asarray(ncc, dtype=np.float64).reshape(-1, G))
    pr = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    res = np.zeros(len(d) * len(pr))
    lib().ptref_equiproj_residuals_orig(_p(X), G, _p(org), _p(ref), len(d), _p(d), _p(n), len(pr), _p(pr), NCC_min,
                                        _p(res))
    return res


def lmdif_jacobian(fun, x, epsfcn=0.0):
    """MINPACK fdjac2's forward-difference Jacobian (lmdif, the reference's LEVMAR path):
    column j steps x_j by h = sqrt(max(epsfcn, DBL_EPSILON)) |x_j| (h = that root when x_j
    is 0)."""
    eps = math.sqrt(max(epsfcn, np.finfo(np.float64).eps))
    f0 = fun(x)
    J = np.zeros((len(f0), len(x)))
    for j in range(len(x)):
        h = eps * abs(x[j])
        if h == 0.0:
            h = eps
        xp = np.array(x, dtype=np.float64)
        xp[j] = x[j] + h
        J[:, j] = (fun(xp) - f0) / h
    return J
