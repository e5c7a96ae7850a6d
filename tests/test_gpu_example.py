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


@pytest.mark.parametrize("iraz", [0, 3])
def test_program_frames_read_ahead_equal_serial(tmp_path, iraz):
    # a 7-image sequence of RGB TIFF files: the program with its frames decoded ahead, copied
    # during the previous pair and its files written during the next, against the reference's
    # serial order; every dic_out file byte for byte (iraz 0: the reference swapped after each
    # image; 3: after every third)
    from test_host_io import write_tiff

    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (make -C p-tracker_amd)")
    S = synth.sequence(n_frames=7, n_grains=24, size=(520, 440), seed=909, radius=(10.0, 10.0), max_step=2.0,
                       max_rot=0.02, margin=70)
    for k, f in enumerate(S.frames):
        v = (f >> 4).astype(np.uint8)
        write_tiff(tmp_path / f"f{k + 1:02d}.tiff", np.stack([v, v, v], axis=2), 2, 8)
    lines = [str(S.G), ""] + [f"{g['refcoord_xpix'] + g['dx']:.3f} {g['refcoord_ypix'] + g['dy']:.3f} 0 10"
                              for g in S.grains]
    (tmp_path / "grains.txt").write_text("\n".join(lines) + "\n")
    cmd = open(os.path.join(EXAMPLE, "commands.txt")).read()
    for key, val in (("image_name", "f%02d.tiff"), ("grains_to_follow", "grains.txt"), ("iref", "1"),
                     ("ibeg", "2"), ("iend", "7"), ("iraz", str(iraz)), ("rescue_level", "1"),
                     ("search_zone.left", "4"), ("search_zone.right", "4"), ("search_zone.up", "4"),
                     ("search_zone.down", "4"), ("search_zone.num_rot", "2"), ("search_zone.inc_rot", "0.01"),
                     ("search_zone_rescue.left", "8"), ("search_zone_rescue.right", "8"),
                     ("search_zone_rescue.up", "8"), ("search_zone_rescue.down", "8"),
                     ("search_zone_rescue.num_rot", "2"), ("search_zone_rescue.inc_rot", "0.01")):
        out = []
        for line in cmd.split("\n"):
            if line.split() and line.split()[0] == key:
                line = f"{key} {val}"
            out.append(line)
        cmd = "\n".join(out)
    (tmp_path / "commands.txt").write_text(cmd)
    runs = {}
    for mode in ("serial", "ahead"):
        d = tmp_path / mode
        d.mkdir()
        for f in os.listdir(tmp_path):
            if f.endswith((".tiff", ".txt")):
                shutil.copyfile(tmp_path / f, d / f)
        args = [EXE, "commands.txt", "--lm"] + (["--serial-frames"] if mode == "serial" else ["--frames-ahead", "3"])
        print(f"program on 7 images, iraz {iraz}: {mode}", flush=True)
        r = subprocess.run(args, cwd=d, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert "tracked 24 grains over 6 images" in r.stdout, r.stdout
        runs[mode] = {num: (d / f"dic_out_{num}.txt").read_bytes() for num in range(2, 8)}
    for num in range(2, 8):
        assert runs["ahead"][num] == runs["serial"][num], num
    rows, trailer = read_dic_bytes(runs["ahead"][7])
    assert len(rows) == 24 and trailer[2] == "#___TO_image: f07.tiff"


def read_dic_bytes(b):
    lines = b.decode().split("\n")
    n = int(lines[0])
    return [l.split() for l in lines[1:1 + n]], [l for l in lines[1 + n:] if l.startswith("#")]
