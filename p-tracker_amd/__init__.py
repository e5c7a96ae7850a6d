"""ptgpu — Python binding of libptgpu, the MI355X (gfx950) per-grain tracking library.

The library is C ABI (``include/ptgpu.h``); this module mirrors its structs with ctypes
and wraps a context in :class:`Tracker`. Loading fails loudly when ``libptgpu.so`` has
not been built: there is no CPU fallback.

Reference types mirrored here: ``search_zone_type`` (src/search_zone_type.hpp:20-27)
as :class:`SearchZone`, the scalar part of ``grain_type_2D``
(src/grain_type_2D.hpp:23-56) as :class:`Grain`.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libptgpu.so")

PTG_OK = 0
OPT_SEARCH_STREAMS = 0
OPT_ROT_TABLES = 1  # 0: device tables with the host-checked guard (default), 1: host tables
OPT_ROT_GUARD_TEST = 2  # 1: guard every (grain, step); 2: also redo every guarded call
OPT_DEVICE_QUEUES = 3  # 1: rescue queues stay on the device (with device tables), 0: read back (default)
OPT_GRAPH = 4  # 1: track_pair captured once as a HIP graph and replayed
OPT_SEARCH_ROLES = 5  # exact search waves: 1 each wave both (default), 2 producer/consumer
OPT_BUCKET_ORDER = 6  # exact search bucket launches: 0 by increasing radius (default), 1 by decreasing work
PTG_ERR_ARG, PTG_ERR_HIP, PTG_ERR_STATE, PTG_ERR_NOMEM, PTG_ERR_LIMIT = -1, -2, -3, -4, -5
PTG_STATUS_CLAMPED = 1
STAGE_MAIN, STAGE_RESCUE, STAGE_SUPER = 0, 1, 2
SEARCH_PLAIN, SEARCH_EXACT = 0, 1
SUBPIX_POWELL, SUBPIX_LM = 0, 1
(T_PRECOMPUTE, T_SEARCH_MAIN, T_SEARCH_RESCUE, T_SEARCH_SUPER, T_SUBPIXEL, T_REMAP,
 T_SEARCH_KERNEL, T_COUNT) = range(8)


class SearchZone(ctypes.Structure):
    _fields_ = [("left", c_int32), ("right", c_int32), ("up", c_int32), ("down", c_int32),
                ("num_rot", c_int32), ("inc_rot", c_double)]


class Grain(ctypes.Structure):
    _fields_ = [("refcoord_xpix", c_int32), ("refcoord_ypix", c_int32),
                ("refrot", c_double), ("radius_pix", c_double),
                ("dx", c_double), ("dy", c_double), ("drot", c_double),
                ("upix", c_double), ("vpix", c_double), ("rot_inc", c_double),
                ("NCC", c_double), ("NCC_rescue", c_double), ("NCC_subpix", c_double),
                ("dx_prev", c_double), ("dy_prev", c_double), ("drot_prev", c_double),
                ("mean0", c_double), ("C0C0", c_double),
                ("masked", c_int32), ("status", c_int32)]


class SubpixParams(ctypes.Structure):
    _fields_ = [("subpix_tol", c_double), ("initial_direction", c_double * 3),
                ("subpix_displacement_max", c_double), ("overflow_stiffness", c_double),
                ("method", c_int32), ("max_iter", c_int32)]


class TrackParams(ctypes.Structure):
    _fields_ = [("search_zone", SearchZone), ("search_zone_rescue", SearchZone),
                ("search_zone_super_rescue", SearchZone),
                ("NCC_min", c_double), ("NCC_min_super", c_double),
                ("rescue_level", c_int32), ("subpixel", c_int32), ("search_method", c_int32),
                ("reserved", c_int32), ("subpix", SubpixParams)]


# numpy view of an array of Grain records
GRAIN_DTYPE = np.dtype({
    "names": [f[0] for f in Grain._fields_],
    "formats": ["<i4", "<i4"] + ["<f8"] * 16 + ["<i4", "<i4"],
    "offsets": [getattr(Grain, f[0]).offset for f in Grain._fields_],
    "itemsize": ctypes.sizeof(Grain),
})
OUTPUT_COLUMNS = ("dx", "dy", "drot", "upix", "vpix", "rot_inc", "NCC", "NCC_rescue", "NCC_subpix")


def default_subpix(method=SUBPIX_POWELL, **kw) -> SubpixParams:
    """Subpixel defaults of src/p-tracker.hpp:122-127."""
    p = SubpixParams()
    p.subpix_tol = kw.get("subpix_tol", 1e-8)
    d = kw.get("initial_direction", (0.01, 0.01, 0.01))
    for i in range(3):
        p.initial_direction[i] = d[i]
    p.subpix_displacement_max = kw.get("subpix_displacement_max", 3.0)
    p.overflow_stiffness = kw.get("overflow_stiffness", 0.0)
    p.method = method
    p.max_iter = kw.get("max_iter", 0)
    return p


def zone(left, right=None, up=None, down=None, num_rot=0, inc_rot=1.0) -> SearchZone:
    right = left if right is None else right
    up = left if up is None else up
    down = up if down is None else down
    return SearchZone(left, right, up, down, num_rot, inc_rot)


def track_params(search_zone, search_zone_rescue=None, search_zone_super_rescue=None, NCC_min=0.7,
                 NCC_min_super=0.5, rescue_level=2, subpixel=1, search_method=SEARCH_PLAIN,
                 subpix=None) -> TrackParams:
    """Tracking defaults of src/p-tracker.hpp:109-127."""
    p = TrackParams()
    p.search_zone = search_zone
    p.search_zone_rescue = search_zone_rescue or search_zone
    p.search_zone_super_rescue = search_zone_super_rescue or p.search_zone_rescue
    p.NCC_min, p.NCC_min_super = NCC_min, NCC_min_super
    p.rescue_level, p.subpixel, p.search_method = rescue_level, subpixel, search_method
    p.subpix = subpix or default_subpix()
    return p


_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libptgpu.so; raises when it is missing (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libptgpu.so not built at {path}: run __graft_entry__.build() (make -C p-tracker_amd)")
    lib = ctypes.CDLL(path)
    vp = c_void_p
    sig = {
        "ptg_abi_version": ([], c_int),
        "ptg_create": ([POINTER(c_void_p), c_int], c_int),
        "ptg_destroy": ([vp], c_int),
        "ptg_last_error": ([vp], c_char_p),
        "ptg_set_image": ([vp, c_int, vp, c_int32, c_int32, c_int64, c_int32], c_int),
        "ptg_set_image_device": ([vp, c_int, vp, c_int32, c_int32, c_int64, c_int32], c_int),
        "ptg_set_image_indices": ([vp, c_int, c_int], c_int),
        "ptg_set_image_async": ([vp, c_int, vp, c_int32, c_int32, c_int64, c_int32], c_int),
        "ptg_wait_image": ([vp, c_int], c_int),
        "ptg_host_alloc": ([vp, ctypes.c_size_t, POINTER(c_void_p)], c_int),
        "ptg_host_free": ([vp, vp], c_int),
        "ptg_get_image": ([vp, c_int, vp, c_int64], c_int),
        "ptg_set_grains": ([vp, c_int32, vp, vp, vp], c_int),
        "ptg_put_grains": ([vp, vp], c_int),
        "ptg_get_grains": ([vp, vp], c_int),
        "ptg_begin_pair": ([vp], c_int),
        "ptg_precompute": ([vp], c_int),
        "ptg_search": ([vp, c_int, POINTER(SearchZone), c_double, c_double, c_int], c_int),
        "ptg_subpixel": ([vp, POINTER(SubpixParams)], c_int),
        "ptg_end_pair": ([vp], c_int),
        "ptg_change_reference": ([vp], c_int),
        "ptg_track_pair": ([vp, POINTER(TrackParams)], c_int),
        "ptg_get_queue_counts": ([vp, POINTER(c_int32), POINTER(c_int32)], c_int),
        "ptg_sync": ([vp], c_int),
        "ptg_ncc_objective": ([vp, c_int32, vp, c_double, c_double, POINTER(c_double)], c_int),
        "ptg_ncc_objective_batch": ([vp, c_int32, vp, vp, c_double, c_double, vp], c_int),
        "ptg_ncc_objective_batch_noalloc": ([vp, c_int32, vp, vp, c_double, c_double, vp], c_int),
        "ptg_ncc_objective_grad_batch": ([vp, c_int32, vp, vp, vp, vp], c_int),
        "ptg_interp_cubic": ([vp, c_int, c_int32, vp, vp], c_int),
        "ptg_undistort_points": ([vp, vp, c_int32, vp, vp], c_int),
        "ptg_remap": ([vp, c_int, c_int, vp], c_int),
        "ptg_set_timing": ([vp, c_int], c_int),
        "ptg_set_option": ([vp, c_int, c_int64], c_int),
        "ptg_autotune_search": ([vp, POINTER(SearchZone), POINTER(c_int32), POINTER(c_double), POINTER(c_double)],
                                c_int),
        "ptg_clear_tuning": ([vp], c_int),
        "ptg_get_stage_times": ([vp, vp, vp, c_int], c_int),
        "ptg_reset_stage_times": ([vp], c_int),
        "ptg_get_subpix_evals": ([vp, POINTER(c_int64)], c_int),
        "ptg_get_subpix_evals_per_grain": ([vp, vp], c_int),
        "ptg_get_exact_fallbacks": ([vp, POINTER(c_int64), POINTER(c_int64)], c_int),
        "ptg_get_rot_guard_stats": ([vp, POINTER(c_int64), POINTER(c_int64), POINTER(c_double)], c_int),
        "ptg_get_graph_stats": ([vp, POINTER(c_int64), POINTER(c_int64), POINTER(c_int64)], c_int),
        "ptg_equiproj_set_data": ([vp, c_int32, vp, c_int32, vp, vp, c_int32, vp, c_double], c_int),
        "ptg_equiproj_set_origins": ([vp, vp], c_int),
        "ptg_equiproj_eval": ([vp, vp, vp, vp, vp, vp, vp], c_int),
        "ptg_calibrate_distortion": ([vp, vp, vp, c_int32, vp, vp, vp], c_int),
        "ptg_joint_eval": ([vp, c_int32, vp, vp, vp, vp, vp, vp], c_int),
        "ptg_refine_joint_distortion": ([vp, vp, c_double, c_int32, vp, vp, vp, vp], c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def exported_symbols():
    """The ptg_* names the binding expects; tests compare them with include/ptgpu.h."""
    return ["ptg_abi_version", "ptg_create", "ptg_destroy", "ptg_last_error", "ptg_set_image",
            "ptg_set_image_device", "ptg_set_image_indices", "ptg_set_image_async", "ptg_wait_image",
            "ptg_host_alloc", "ptg_host_free", "ptg_get_image", "ptg_set_grains", "ptg_put_grains",
            "ptg_get_grains", "ptg_begin_pair", "ptg_precompute", "ptg_search", "ptg_subpixel",
            "ptg_end_pair", "ptg_change_reference", "ptg_track_pair", "ptg_get_queue_counts",
            "ptg_sync", "ptg_ncc_objective", "ptg_ncc_objective_batch", "ptg_ncc_objective_batch_noalloc",
            "ptg_ncc_objective_grad_batch", "ptg_interp_cubic", "ptg_undistort_points", "ptg_remap",
            "ptg_set_timing", "ptg_set_option", "ptg_autotune_search", "ptg_clear_tuning", "ptg_get_stage_times", "ptg_reset_stage_times", "ptg_get_subpix_evals",
            "ptg_get_subpix_evals_per_grain", "ptg_get_exact_fallbacks", "ptg_get_rot_guard_stats", "ptg_get_graph_stats", "ptg_equiproj_set_data", "ptg_equiproj_set_origins",
            "ptg_equiproj_eval", "ptg_calibrate_distortion", "ptg_joint_eval", "ptg_refine_joint_distortion"]


class PtgError(RuntimeError):
    pass


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(c_void_p)


class Tracker:
    """One libptgpu context (one HIP stream on one device)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = c_void_p()
        rc = self.lib.ptg_create(ctypes.byref(h), device)
        if rc != PTG_OK:
            msg = self.lib.ptg_last_error(None)
            raise PtgError(f"ptg_create(device={device}) failed with {rc}: {msg.decode() if msg else ''}")
        self.h = h
        self.G = 0
        self._pinned = []       # (address, array) of ptg_host_alloc buffers
        self._in_copy = {}      # image slot -> the host array its overlapped copy reads

    def close(self):
        if self.h:
            self.lib.ptg_sync(self.h)
            self._in_copy.clear()
            for addr, _ in self._pinned:
                self.lib.ptg_host_free(self.h, c_void_p(addr))
            self._pinned = []
            self.lib.ptg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ck(self, rc, what):
        if rc != PTG_OK:
            msg = self.lib.ptg_last_error(self.h)
            raise PtgError(f"{what} -> {rc}: {msg.decode() if msg else ''}")

    def set_image(self, index: int, frame: np.ndarray, bits: int = 16):
        f = np.ascontiguousarray(frame, dtype=np.uint16)
        self._ck(self.lib.ptg_set_image(self.h, index, _ptr(f), f.shape[1], f.shape[0], f.shape[1], bits),
                 "ptg_set_image")

    def set_image_device(self, index: int, dev_ptr: int, dimx: int, dimy: int, pitch: int, bits: int = 16):
        self._ck(self.lib.ptg_set_image_device(self.h, index, c_void_p(dev_ptr), dimx, dimy, pitch, bits),
                 "ptg_set_image_device")

    def host_frame(self, height: int, width: int) -> np.ndarray:
        """A (height, width) uint16 array in pinned host memory (ptg_host_alloc), for
        set_image_async; released by close()."""
        addr = c_void_p()
        n = int(height) * int(width)
        self._ck(self.lib.ptg_host_alloc(self.h, 2 * n, ctypes.byref(addr)), "ptg_host_alloc")
        arr = np.ctypeslib.as_array((ctypes.c_uint16 * n).from_address(addr.value)).reshape(height, width)
        self._pinned.append((addr.value, arr))
        return arr

    def set_image_async(self, index: int, frame: np.ndarray, bits: int = 16):
        """ptg_set_image_async: a `frame` (C-contiguous uint16) in pinned memory (host_frame)
        must stay unchanged until wait_image(index) or sync(), and is kept alive until then;
        from other memory the call returns once the copy has read it."""
        if frame.dtype != np.uint16 or not frame.flags["C_CONTIGUOUS"]:
            raise PtgError("set_image_async needs a C-contiguous uint16 array")
        self._ck(self.lib.ptg_set_image_async(self.h, index, _ptr(frame), frame.shape[1], frame.shape[0],
                                              frame.shape[1], bits), "ptg_set_image_async")
        self._in_copy[index] = frame

    def wait_image(self, index: int):
        self._ck(self.lib.ptg_wait_image(self.h, index), "ptg_wait_image")
        self._in_copy.pop(index, None)

    def set_image_indices(self, iref: int, icur: int):
        self._ck(self.lib.ptg_set_image_indices(self.h, iref, icur), "ptg_set_image_indices")

    def set_grains(self, grains: np.ndarray, pattern_xy: np.ndarray, pattern_begin: np.ndarray):
        g = np.ascontiguousarray(grains, dtype=GRAIN_DTYPE)
        pxy = np.ascontiguousarray(pattern_xy, dtype=np.int32)
        pb = np.ascontiguousarray(pattern_begin, dtype=np.int64)
        self._ck(self.lib.ptg_set_grains(self.h, len(g), _ptr(g), _ptr(pxy), _ptr(pb)), "ptg_set_grains")
        self.G = len(g)

    def put_grains(self, grains: np.ndarray):
        g = np.ascontiguousarray(grains, dtype=GRAIN_DTYPE)
        assert len(g) == self.G
        self._ck(self.lib.ptg_put_grains(self.h, _ptr(g)), "ptg_put_grains")

    def get_grains(self) -> np.ndarray:
        out = np.zeros(self.G, dtype=GRAIN_DTYPE)
        self._ck(self.lib.ptg_get_grains(self.h, _ptr(out)), "ptg_get_grains")
        return out

    def begin_pair(self):
        self._ck(self.lib.ptg_begin_pair(self.h), "ptg_begin_pair")

    def precompute(self):
        self._ck(self.lib.ptg_precompute(self.h), "ptg_precompute")

    def search(self, stage: int, z: SearchZone, NCC_min=0.7, NCC_min_super=0.5, method=SEARCH_PLAIN):
        self._ck(self.lib.ptg_search(self.h, stage, ctypes.byref(z), NCC_min, NCC_min_super, method), "ptg_search")

    def subpixel(self, prm: SubpixParams):
        self._ck(self.lib.ptg_subpixel(self.h, ctypes.byref(prm)), "ptg_subpixel")

    def end_pair(self):
        self._ck(self.lib.ptg_end_pair(self.h), "ptg_end_pair")

    def change_reference(self):
        self._ck(self.lib.ptg_change_reference(self.h), "ptg_change_reference")

    def track_pair(self, prm: TrackParams):
        self._ck(self.lib.ptg_track_pair(self.h, ctypes.byref(prm)), "ptg_track_pair")

    def queue_counts(self):
        a, b = c_int32(), c_int32()
        self._ck(self.lib.ptg_get_queue_counts(self.h, ctypes.byref(a), ctypes.byref(b)), "ptg_get_queue_counts")
        return a.value, b.value

    def sync(self):
        self._ck(self.lib.ptg_sync(self.h), "ptg_sync")
        self._in_copy.clear()

    def objective(self, igrain, X, dmax=3.0, stiffness=0.0):
        ig = np.ascontiguousarray(np.atleast_1d(igrain), dtype=np.int32)
        Xa = np.ascontiguousarray(np.asarray(X, dtype=np.float64).reshape(-1, 3))
        f = np.zeros(len(ig), dtype=np.float64)
        self._ck(self.lib.ptg_ncc_objective_batch(self.h, len(ig), _ptr(ig), _ptr(Xa), dmax, stiffness, _ptr(f)),
                 "ptg_ncc_objective_batch")
        return f

    def objective_noalloc(self, igrain, X, dmax=3.0, stiffness=0.0):
        """As objective(), staged in the library's static device array (at most 16384 points;
        the reference moments must be computed: precompute())."""
        ig = np.ascontiguousarray(np.atleast_1d(igrain), dtype=np.int32)
        Xa = np.ascontiguousarray(np.asarray(X, dtype=np.float64).reshape(-1, 3))
        f = np.zeros(len(ig), dtype=np.float64)
        self._ck(self.lib.ptg_ncc_objective_batch_noalloc(self.h, len(ig), _ptr(ig), _ptr(Xa), dmax, stiffness,
                                                          _ptr(f)), "ptg_ncc_objective_batch_noalloc")
        return f

    def objective_grad(self, igrain, X):
        ig = np.ascontiguousarray(np.atleast_1d(igrain), dtype=np.int32)
        Xa = np.ascontiguousarray(np.asarray(X, dtype=np.float64).reshape(-1, 3))
        f = np.zeros(len(ig), dtype=np.float64)
        g = np.zeros((len(ig), 3), dtype=np.float64)
        self._ck(self.lib.ptg_ncc_objective_grad_batch(self.h, len(ig), _ptr(ig), _ptr(Xa), _ptr(f), _ptr(g)),
                 "ptg_ncc_objective_grad_batch")
        return f, g

    def interp_cubic(self, index: int, xy: np.ndarray) -> np.ndarray:
        xy = np.ascontiguousarray(np.asarray(xy, dtype=np.float64).reshape(-1, 2))
        out = np.zeros(len(xy), dtype=np.float64)
        self._ck(self.lib.ptg_interp_cubic(self.h, index, len(xy), _ptr(xy), _ptr(out)), "ptg_interp_cubic")
        return out

    def undistort_points(self, params14, xy):
        p = np.ascontiguousarray(params14, dtype=np.float64)
        xy = np.ascontiguousarray(np.asarray(xy, dtype=np.float64).reshape(-1, 2))
        out = np.zeros_like(xy)
        self._ck(self.lib.ptg_undistort_points(self.h, _ptr(p), len(xy), _ptr(xy), _ptr(out)),
                 "ptg_undistort_points")
        return out

    def equiproj_set_data(self, ref_xy, disp, ncc, pairs, NCC_min=0.7):
        """ref_xy (G x 2), disp (nimg x G x 2), ncc (nimg x G), pairs (npairs x 2)."""
        ref = np.ascontiguousarray(np.asarray(ref_xy, dtype=np.float64).reshape(-1, 2))
        G = len(ref)
        d = np.ascontiguousarray(np.asarray(disp, dtype=np.float64).reshape(-1, G, 2))
        n = np.ascontiguousarray(np.asarray(ncc, dtype=np.float64).reshape(-1, G))
        pr = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
        self._eq_m = len(d) * len(pr)
        self._ck(self.lib.ptg_equiproj_set_data(self.h, G, _ptr(ref), len(d), _ptr(d), _ptr(n), len(pr), _ptr(pr),
                                                NCC_min), "ptg_equiproj_set_data")

    def equiproj_set_origins(self, orig_xy=None):
        """The POWELL_ITER form: given first positions (G x 2), or None for the standard form."""
        if orig_xy is None:
            self._ck(self.lib.ptg_equiproj_set_origins(self.h, None), "ptg_equiproj_set_origins")
            return
        o = np.ascontiguousarray(np.asarray(orig_xy, dtype=np.float64).reshape(-1, 2))
        self._ck(self.lib.ptg_equiproj_set_origins(self.h, _ptr(o)), "ptg_equiproj_set_origins")

    def equiproj_eval(self, params14, full=False):
        """(f, J^T r, J^T J) and, with full, (residuals, Jacobian)."""
        p = np.ascontiguousarray(params14, dtype=np.float64)
        f = np.zeros(1)
        jtr = np.zeros(14)
        jtj = np.zeros((14, 14))
        res = np.zeros(self._eq_m) if full else None
        jac = np.zeros((self._eq_m, 14)) if full else None
        self._ck(self.lib.ptg_equiproj_eval(self.h, _ptr(p), _ptr(f), _ptr(jtr), _ptr(jtj),
                                            _ptr(res) if full else None, _ptr(jac) if full else None),
                 "ptg_equiproj_eval")
        return (f[0], jtr, jtj, res, jac) if full else (f[0], jtr, jtj)

    def calibrate_distortion(self, params14, free, max_iter=100):
        """Returns (params14, f_ini, f_end, iterations)."""
        p = np.ascontiguousarray(params14, dtype=np.float64).copy()
        fr = np.ascontiguousarray(np.asarray(free, dtype=np.int32).reshape(14))
        fi, fe, it = c_double(), c_double(), c_int32()
        self._ck(self.lib.ptg_calibrate_distortion(self.h, _ptr(p), _ptr(fr), max_iter, ctypes.byref(fi),
                                                   ctypes.byref(fe), ctypes.byref(it)), "ptg_calibrate_distortion")
        return p, fi.value, fe.value, it.value

    def get_image(self, index: int, dimx: int, dimy: int) -> np.ndarray:
        out = np.zeros((dimy, dimx), dtype=np.uint16)
        self._ck(self.lib.ptg_get_image(self.h, index, _ptr(out), dimx), "ptg_get_image")
        return out

    def joint_eval(self, igrain, P, lens, derivs=True):
        """The joint objective of grains igrain at P (n x 3: Tx, Ty, theta) with the radial lens
        (k1, k2, xc, yc): f, and with derivs df/d(Tx, Ty, theta, k1, k2) (n x 5) and the
        Gauss-Newton matrices' upper triangles (n x 15)."""
        ig = np.ascontiguousarray(np.atleast_1d(igrain), dtype=np.int32)
        Pa = np.ascontiguousarray(np.asarray(P, dtype=np.float64).reshape(-1, 3))
        L = np.ascontiguousarray(lens, dtype=np.float64)
        f = np.zeros(len(ig))
        g = np.zeros((len(ig), 5)) if derivs else None
        h = np.zeros((len(ig), 15)) if derivs else None
        self._ck(self.lib.ptg_joint_eval(self.h, len(ig), _ptr(ig), _ptr(Pa), _ptr(L), _ptr(f),
                                         _ptr(g) if derivs else None, _ptr(h) if derivs else None), "ptg_joint_eval")
        return (f, g, h) if derivs else f

    def refine_joint_distortion(self, k1, k2, xc, yc, ncc_min=0.7, max_iter=50):
        """Joint LM over the grains' motion and (k1, k2); returns (k1, k2, f_ini, f_end, iters, ngrains)."""
        L = np.array([k1, k2, xc, yc], dtype=np.float64)
        fi, fe = c_double(), c_double()
        it, ng = c_int32(), c_int32()
        self._ck(self.lib.ptg_refine_joint_distortion(self.h, _ptr(L), ncc_min, max_iter, ctypes.byref(fi),
                                                      ctypes.byref(fe), ctypes.byref(it), ctypes.byref(ng)),
                 "ptg_refine_joint_distortion")
        return float(L[0]), float(L[1]), fi.value, fe.value, it.value, ng.value

    def remap(self, src: int, dst: int, params14):
        p = np.ascontiguousarray(params14, dtype=np.float64)
        self._ck(self.lib.ptg_remap(self.h, src, dst, _ptr(p)), "ptg_remap")

    def set_option(self, option: int, value: int):
        self._ck(self.lib.ptg_set_option(self.h, option, value), "ptg_set_option")

    def autotune_search(self, z: SearchZone):
        """Times the exact search's launch shapes per radius bucket of a main-stage search over
        z and keeps the fastest; returns (searches run, kernel ms with the heuristic's shapes,
        kernel ms with the chosen ones)."""
        n, b, a = c_int32(), c_double(), c_double()
        self._ck(self.lib.ptg_autotune_search(self.h, ctypes.byref(z), ctypes.byref(n), ctypes.byref(b),
                                              ctypes.byref(a)), "ptg_autotune_search")
        return n.value, b.value, a.value

    def clear_tuning(self):
        self._ck(self.lib.ptg_clear_tuning(self.h), "ptg_clear_tuning")

    def set_timing(self, on: bool):
        self._ck(self.lib.ptg_set_timing(self.h, 1 if on else 0), "ptg_set_timing")

    def stage_times(self):
        ms = np.zeros(T_COUNT, dtype=np.float64)
        n = np.zeros(T_COUNT, dtype=np.int64)
        self._ck(self.lib.ptg_get_stage_times(self.h, _ptr(ms), _ptr(n), T_COUNT), "ptg_get_stage_times")
        return ms, n

    def reset_stage_times(self):
        self._ck(self.lib.ptg_reset_stage_times(self.h), "ptg_reset_stage_times")

    def subpix_evals_per_grain(self) -> np.ndarray:
        out = np.zeros(self.G, dtype=np.int32)
        self._ck(self.lib.ptg_get_subpix_evals_per_grain(self.h, _ptr(out)), "ptg_get_subpix_evals_per_grain")
        return out

    def graph_stats(self):
        """(pairs captured as HIP graphs, graph launches, nodes of the current graph)."""
        n, m, k = c_int64(), c_int64(), c_int64()
        self._ck(self.lib.ptg_get_graph_stats(self.h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(k)),
                 "ptg_get_graph_stats")
        return n.value, m.value, k.value

    def rot_guard_stats(self):
        """(entries the host compared, calls redone on the host's rotation tables, largest
        |device - host| cosine or sine among the compared entries)."""
        n, r, d = c_int64(), c_int64(), c_double()
        self._ck(self.lib.ptg_get_rot_guard_stats(self.h, ctypes.byref(n), ctypes.byref(r), ctypes.byref(d)),
                 "ptg_get_rot_guard_stats")
        return n.value, r.value, d.value

    def exact_fallbacks(self):
        """(searches, grains) handed from the exact form to the plain kernel."""
        n, g = c_int64(), c_int64()
        self._ck(self.lib.ptg_get_exact_fallbacks(self.h, ctypes.byref(n), ctypes.byref(g)), "ptg_get_exact_fallbacks")
        return n.value, g.value

    def subpix_evals(self) -> int:
        e = c_int64()
        self._ck(self.lib.ptg_get_subpix_evals(self.h, ctypes.byref(e)), "ptg_get_subpix_evals")
        return e.value
