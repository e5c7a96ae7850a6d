"""The C++ host side (p-tracker_amd/host/ptracker_gpu.hpp, reference-named class over the
C ABI) drives one frame pair; its 9 result columns are checked against the oracle."""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

import ptref

pk = ptref.package()
from ptracker_amd import synth  # noqa: E402

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "p-tracker_amd", "host", "ptracker_gpu_check")


def _zone_bytes(z):
    return struct.pack("<5i4xd", z.left, z.right, z.up, z.down, z.num_rot, z.inc_rot)


@pytest.mark.parametrize("method", [pk.SEARCH_PLAIN, pk.SEARCH_EXACT])
def test_cpp_host_track_pair(tmp_path, method):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (make -C p-tracker_amd)")
    w = synth.workload(1, n_grains=16, size=(320, 320), seed=909)
    print(f"C++ host driver, method {method}", flush=True)
    p = w.params
    buf = bytearray()
    buf += struct.pack("<8i", w.I0.shape[1], w.I0.shape[0], w.bits, w.G, method, pk.SUBPIX_POWELL, 0, 1)
    buf += _zone_bytes(p.search_zone) + _zone_bytes(p.search_zone_rescue) + _zone_bytes(p.search_zone_super_rescue)
    buf += struct.pack("<2d", 0.7, 0.5)
    buf += w.I0.astype("<u2").tobytes() + w.I1.astype("<u2").tobytes()
    for g in range(w.G):
        pat = w.pattern_xy[w.pattern_begin[g]:w.pattern_begin[g + 1]]
        r = w.grains[g]
        buf += struct.pack("<4i5d", int(r["refcoord_xpix"]), int(r["refcoord_ypix"]), int(r["masked"]), len(pat),
                           r["refrot"], r["radius_pix"], r["dx"], r["dy"], r["drot"])
        buf += pat.astype("<i4").tobytes()
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    fin.write_bytes(bytes(buf))
    subprocess.run([EXE, str(fin), str(fout)], check=True, timeout=120)
    raw = fout.read_bytes()
    cols = np.frombuffer(raw[:w.G * 72], dtype="<f8").reshape(w.G, 9)
    ref, _, _ = ptref.track_pair(ptref.Workload(w.I0, w.I1, w.grains, w.pattern_xy, w.pattern_begin, w.bits),
                                 p, nthreads=2)
    names = ["dx", "dy", "drot", "upix", "vpix", "rot_inc", "NCC", "NCC_rescue", "NCC_subpix"]
    tol = {"dx": 1e-3, "dy": 1e-3, "drot": 1e-5, "upix": 1e-3, "vpix": 1e-3, "rot_inc": 1e-5, "NCC": 1e-12,
           "NCC_rescue": 1e-12, "NCC_subpix": 1e-6}
    for k, n in enumerate(names):
        assert np.abs(cols[:, k] - ref[n]).max() < tol[n], n
    f0 = struct.unpack("<d", raw[w.G * 72 + 4:w.G * 72 + 12])[0]
    assert 0.0 <= f0 < 1.0
