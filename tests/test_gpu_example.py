"""End to end: the tracking program (p-tracker_amd/host/ptracker_gpu) on the reference's
example (two 380x380 RGB TIFF frames, 9 grains, zones +-6 / +-30 / +-200 px, 13 / 139 /
139 rotation steps, rescue level 2, Powell subpixel). Its dic_out files are checked
against the oracle run on the same decoded frames: the reference ships no expected
output for this example."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import ptref

pk = ptref.package()
from ptracker_amd import synth  # noqa: E402

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "p-tracker_amd", "host", "ptracker_gpu")
IO = os.path.join(ROOT, "p-tracker_amd", "host", "ptracker_io_check")
EXAMPLE = os.path.join(ROOT, "tests", "golden", "example")
COLS = ["refcoord_xpix", "refcoord_ypix", "refrot", "radius_pix", "dx", "dy", "drot", "upix", "vpix", "rot_inc",
        "NCC", "NCC_rescue", "NCC_subpix"]


def frame(path, tmp):
    subprocess.run([IO, "tiff", path, str(tmp / "f.raw")], check=True, timeout=60)
    b = (tmp / "f.raw").read_bytes()
    l1 = b.index(b"\n")
    l2 = b.index(b"\n", l1 + 1)
    w, h = map(int, b[:l1].split())
    return np.frombuffer(b[l2 + 1:], dtype="<u2").reshape(h, w).copy()


def read_dic(path):
    lines = open(path).read().split("\n")
    n = int(lines[0])
    rows = [l.split() for l in lines[1:1 + n]]
    trailer = [l for l in lines[1 + n:] if l.startswith("#")]
    return rows, trailer


def start_grains():
    rows = [l.split() for l in open(os.path.join(EXAMPLE, "recentered.txt")).read().split("\n")[1:] if l.strip()]
    g = np.zeros(len(rows), dtype=pk.GRAIN_DTYPE)
    for i, (x, y, rot, r) in enumerate(rows):
        x, y = float(x), float(y)
        g[i]["refcoord_xpix"], g[i]["refcoord_ypix"] = int(np.floor(x)), int(np.floor(y))
        g[i]["dx"] = g[i]["upix"] = x - np.floor(x)
        g[i]["dy"] = g[i]["vpix"] = y - np.floor(y)
        g[i]["refrot"], g[i]["radius_pix"] = float(rot), float(r)
    return g


@pytest.mark.parametrize("subpix", ["powell", "lm"])
def test_example_program_matches_oracle(tmp_path, subpix):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (make -C p-tracker_amd)")
    for f in ("TEST01.tiff", "TEST02.tiff", "recentered.txt", "commands.txt"):
        shutil.copyfile(os.path.join(EXAMPLE, f), tmp_path / f)
    print(f"example program, subpixel {subpix}", flush=True)
    args = [EXE, "commands.txt"] + (["--lm"] if subpix == "lm" else [])
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    I1, I2 = frame(tmp_path / "TEST01.tiff", tmp_path), frame(tmp_path / "TEST02.tiff", tmp_path)
    g = start_grains()
    p = synth.circ_pattern(10)
    pxy = np.ascontiguousarray(np.concatenate([p] * len(g)), dtype=np.int32)
    pb = np.arange(len(g) + 1, dtype=np.int64) * len(p)
    z = pk.zone(6, num_rot=6, inc_rot=0.0454)
    zr = pk.zone(30, num_rot=69, inc_rot=0.0454)
    zs = pk.zone(200, num_rot=69, inc_rot=0.045)
    sp = pk.default_subpix(pk.SUBPIX_LM, subpix_tol=1e-13) if subpix == "lm" else pk.default_subpix(pk.SUBPIX_POWELL)
    prm = pk.track_params(z, zr, zs, NCC_min=0.8, NCC_min_super=0.7, rescue_level=2, subpixel=1, subpix=sp)
    # images 1 -> 1, then 1 -> 2 from the first pair's grains (iref 1, ibeg 1, iend 2)
    o1, counts1, _ = ptref.track_pair(ptref.Workload(I1, I1, g, pxy, pb, 16), prm, nthreads=8)
    o2, counts, _ = ptref.track_pair(ptref.Workload(I1, I2, o1, pxy, pb, 16), prm, nthreads=8)
    assert counts[0] == 1  # one grain goes to the rescue search
    # the 16-bit frames' main searches (+-6 px x 13 steps) run on the exact MFMA kernels with
    # no grain handed over; the rescue stages (139 steps, beyond the exact form's 40 for 16-bit
    # frames) run wholly on the plain form, one search per stage with a non-empty queue
    fb = [l for l in r.stdout.splitlines() if l.startswith("plain-form fallbacks of the exact search:")]
    assert len(fb) == 1, r.stdout
    nsearch, ngrain = (int(t) for t in fb[0].split(":")[1].replace("searches", "").replace("grains", "").split(","))
    assert ngrain == 0 and nsearch == sum(int(c[0] > 0) + int(c[1] > 0) for c in (counts1, counts)), fb[0]
    tol = {"dx": 1e-3, "dy": 1e-3, "drot": 1e-5, "upix": 1e-3, "vpix": 1e-3, "rot_inc": 1e-5, "NCC": 1e-12,
           "NCC_rescue": 1e-12, "NCC_subpix": 1e-6}
    for num, ref in ((1, o1), (2, o2)):
        rows, trailer = read_dic(tmp_path / f"dic_out_{num}.txt")
        assert len(rows) == 9 and all(len(x) == 13 for x in rows)
        assert trailer == ["#_FROM_image: TEST01.tiff", "#___dateTime: dateTime unknown",
                           f"#___TO_image: TEST0{num}.tiff", "#___dateTime: dateTime unknown"]
        for k, c in enumerate(COLS):
            got = np.array([float(x[k]) for x in rows])
            if k < 4:
                assert np.array_equal(got, ref[c].astype(np.float64)), (num, c)
            else:
                assert np.abs(got - ref[c]).max() < tol[c], (num, c, np.abs(got - ref[c]).max())
