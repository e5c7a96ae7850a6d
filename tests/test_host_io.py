"""The host file side of the tracking program (p-tracker_amd/host/ptracker_io.hpp,
tiff_reader.hpp) through its CPU driver: frame ingest against libtiff's own RGBA decoder,
the dic_out writer against the format the reference writes (src/p-tracker.cpp:1313-1367),
grain-file reading, the command-file grammar and defaults, patterns and the point grid.
No GPU."""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

import ptref

pk = ptref.package()
from ptracker_amd import synth  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "p-tracker_amd", "host", "ptracker_io_check")
EXAMPLE = os.path.join(ROOT, "tests", "golden", "example")
LIBTIFF = "/opt/conda/lib/libtiff.so"


def run(*args):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (make -C p-tracker_amd)")
    subprocess.run([EXE, *map(str, args)], check=True, timeout=60)


def ours(path, tmp_path):
    out = tmp_path / "frame.raw"
    run("tiff", path, out)
    b = out.read_bytes()
    l1 = b.index(b"\n")
    l2 = b.index(b"\n", l1 + 1)
    w, h = map(int, b[:l1].split())
    return np.frombuffer(b[l2 + 1:], dtype="<u2").reshape(h, w), b[l1 + 1:l2].decode()


def gray_of_rgba8(R, G, B):
    # the reference's conversion, src/p-tracker.cpp:148-186, in its operation order
    fact = 1.0 / 255.0
    r, g, b = R.astype(np.float64) * fact, G.astype(np.float64) * fact, B.astype(np.float64) * fact
    return (65535.0 * ((0.299 * r + 0.587 * g) + 0.114 * b)).astype(np.uint16)


def libtiff_gray(path, w, h):
    """TIFFReadRGBAImage (the reference's reader) through the installed libtiff."""
    try:
        lib = ctypes.CDLL(LIBTIFF)
    except OSError:
        pytest.skip("libtiff is not installed here")
    lib.TIFFOpen.restype = ctypes.c_void_p
    lib.TIFFOpen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    lib.TIFFReadRGBAImage.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int]
    lib.TIFFClose.argtypes = [ctypes.c_void_p]
    tif = lib.TIFFOpen(str(path).encode(), b"r")
    assert tif
    r = (ctypes.c_uint32 * (w * h))()
    assert lib.TIFFReadRGBAImage(tif, w, h, r, 0) == 1
    lib.TIFFClose(tif)
    a = np.frombuffer(r, dtype=np.uint32).reshape(h, w)[::-1]  # the raster is bottom-up
    return gray_of_rgba8(a & 0xff, (a >> 8) & 0xff, (a >> 16) & 0xff)


def write_tiff(path, arr, photometric, bps, big=False, planar=False, predictor=False):
    """A minimal uncompressed strip TIFF (one strip); arr is H x W or H x W x S."""
    import struct

    bo = ">" if big else "<"
    h, w = arr.shape[:2]
    spp = 1 if arr.ndim == 2 else arr.shape[2]
    a = arr.astype(np.uint16 if bps == 16 else np.uint8)
    if predictor:
        d = a.reshape(h, w, spp).astype(np.int64)
        d[:, 1:] = d[:, 1:] - d[:, :-1]
        a = (d % (1 << bps)).astype(a.dtype).reshape(a.shape)
    if planar and spp > 1:
        planes = [np.ascontiguousarray(a[..., s]).astype(bo + ("u2" if bps == 16 else "u1")).tobytes() for s in range(spp)]
    else:
        planes = [a.astype(bo + ("u2" if bps == 16 else "u1")).tobytes()]
    entries = [(256, 4, [w]), (257, 4, [h]), (258, 3, [bps] * spp), (259, 3, [1]), (262, 3, [photometric]),
               (273, 4, [0] * len(planes)), (277, 3, [spp]), (278, 4, [h]), (279, 4, [len(p) for p in planes]),
               (284, 3, [2 if planar and spp > 1 else 1])]
    if predictor:
        entries.append((317, 3, [2]))
    n = len(entries)
    head = 8 + 2 + 12 * n + 4
    extra = bytearray()
    offs = {}
    for tag, typ, vals in entries:
        size = (2 if typ == 3 else 4) * len(vals)
        if size > 4:
            offs[tag] = head + len(extra)
            extra += b"".join(struct.pack(bo + ("H" if typ == 3 else "I"), v) for v in vals)
            if len(extra) % 2:
                extra += b"\0"
    data_off = head + len(extra)
    strip_offs = []
    o = data_off
    for p in planes:
        strip_offs.append(o)
        o += len(p)
    out = bytearray((b"MM" if big else b"II") + struct.pack(bo + "HI", 42, 8) + struct.pack(bo + "H", n))
    for tag, typ, vals in entries:
        if tag == 273:
            vals = strip_offs
        size = (2 if typ == 3 else 4) * len(vals)
        if size > 4:
            if tag == 273:  # rewrite the offsets table in place
                off = offs[tag] - head
                extra[off:off + 4 * len(vals)] = b"".join(struct.pack(bo + "I", v) for v in vals)
            out += struct.pack(bo + "HHII", tag, typ, len(vals), offs[tag])
        else:
            raw = b"".join(struct.pack(bo + ("H" if typ == 3 else "I"), v) for v in vals)
            out += struct.pack(bo + "HHI", tag, typ, len(vals)) + raw.ljust(4, b"\0")
    out += struct.pack(bo + "I", 0) + extra + b"".join(planes)
    open(path, "wb").write(bytes(out))


# ---- frame ingest ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["TEST01.tiff", "TEST02.tiff"])
def test_example_frames_equal_libtiff(tmp_path, name):
    # big-endian RGB, LZW with horizontal prediction, 4 strips, an ICC profile
    path = os.path.join(EXAMPLE, name)
    g, dt = ours(path, tmp_path)
    assert g.shape == (380, 380)
    assert dt == "dateTime unknown"
    assert np.array_equal(g, libtiff_gray(path, 380, 380))


@pytest.mark.parametrize("mode,compression,predictor", [
    ("L", "raw", 0), ("L", "tiff_lzw", 0), ("L", "packbits", 0), ("L", "tiff_deflate", 0),
    ("RGB", "raw", 0), ("RGB", "tiff_lzw", 0), ("RGB", "packbits", 0), ("RGB", "tiff_deflate", 0),
    ("P", "tiff_lzw", 0), ("I;16", "raw", 0), ("I;16", "tiff_lzw", 0),
    ("L", "tiff_lzw", 1), ("RGB", "tiff_adobe_deflate", 1), ("I;16", "tiff_lzw", 1), ("I;16", "tiff_adobe_deflate", 1)])
def test_pil_written_frames_equal_libtiff(tmp_path, mode, compression, predictor):
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(len(mode) * 100 + len(compression) * 10 + predictor)
    h, w = 53, 71
    if mode == "RGB":
        im = Image.fromarray(rng.integers(0, 256, (h, w, 3), dtype=np.uint8), "RGB")
    elif mode == "I;16":
        im = Image.fromarray(rng.integers(0, 65536, (h, w), dtype=np.uint16))
    elif mode == "P":
        im = Image.fromarray(rng.integers(0, 256, (h, w, 3), dtype=np.uint8), "RGB").quantize(200)
    else:
        im = Image.fromarray(rng.integers(0, 256, (h, w), dtype=np.uint8), "L")
    path = tmp_path / "f.tif"
    info = {306: "2024:05:06 07:08:09"}
    if predictor:
        info[317] = 2
    im.save(path, compression=compression, tiffinfo=info)
    g, dt = ours(path, tmp_path)
    assert dt == "2024:05:06 07:08:09"
    assert np.array_equal(g, libtiff_gray(path, w, h))


@pytest.mark.parametrize("case", ["rgb16", "gray16_big", "white_is_zero8", "planar_rgb8", "predictor_rgb8",
                                  "predictor_gray16"])
def test_hand_written_frames_equal_libtiff(tmp_path, case):
    # (a predictor tag on uncompressed strips is ignored, as libtiff does)
    rng = np.random.default_rng(7)
    h, w = 37, 45
    path = tmp_path / "f.tif"
    if case == "rgb16":
        write_tiff(path, rng.integers(0, 65536, (h, w, 3)), 2, 16)
    elif case == "gray16_big":
        write_tiff(path, rng.integers(0, 65536, (h, w)), 1, 16, big=True)
    elif case == "white_is_zero8":
        write_tiff(path, rng.integers(0, 256, (h, w)), 0, 8)
    elif case == "planar_rgb8":
        write_tiff(path, rng.integers(0, 256, (h, w, 3)), 2, 8, planar=True)
    elif case == "predictor_rgb8":
        write_tiff(path, rng.integers(0, 256, (h, w, 3)), 2, 8, predictor=True)
    else:
        write_tiff(path, rng.integers(0, 65536, (h, w)), 1, 16, big=True, predictor=True)
    g, _ = ours(path, tmp_path)
    assert np.array_equal(g, libtiff_gray(path, w, h))


def test_full_white_is_65534_and_rows_top_down(tmp_path):
    # 0.299 + 0.587 + 0.114 = 0.9999999999999999 in double (closed form); row 0 of the
    # frame is the file's first row
    a = np.full((4, 5, 3), 255)
    a[0] = 0
    write_tiff(tmp_path / "w.tif", a, 2, 8)
    g, _ = ours(tmp_path / "w.tif", tmp_path)
    assert (g[0] == 0).all() and (g[1:] == 65534).all()
    assert 0.299 + 0.587 + 0.114 == 0.9999999999999999


# ---- grain files ----------------------------------------------------------------------------
def dic_text(rows, trailer=("a", "b", "c", "d")):
    """The dic_out layout: count, rows of 2 ints (width 6, left) and 11 doubles
    (%23.15e), a blank line and the 4 trailer lines."""
    out = [f"{len(rows)}\n"]
    for r in rows:
        out.append(f"{int(r[0]):<6d} {int(r[1]):<6d} " + " ".join(f"{v:23.15e}" for v in r[2:]) + "\n")
    out.append("\n")
    out.append(f"#_FROM_image: {trailer[0]}\n#___dateTime: {trailer[1]}\n#___TO_image: {trailer[2]}\n"
               f"#___dateTime: {trailer[3]}\n")
    return "".join(out)


def random_rows(rng, n):
    rows = []
    for i in range(n):
        ints = [int(rng.integers(-50, 7000)), int(rng.integers(-50, 5000))]
        vals = [float(v) for v in rng.normal(0, 10, 11) * 10.0 ** rng.integers(-12, 6, 11)]
        rows.append(ints + vals)
    rows[0][2:5] = [-0.0, 0.0, 1e-300]
    rows[1][2:5] = [123456789.0, -1.5, 2.2250738585072014e-308]  # (istream >> rejects denormals)
    return rows


def test_dic_out_writer_matches_reference_layout(tmp_path):
    rng = np.random.default_rng(11)
    rows = random_rows(rng, 40)
    src = tmp_path / "in.txt"
    src.write_text(f"{len(rows)}\n" + "".join(" ".join(repr(v) for v in r) + "\n" for r in rows))
    run("dic", src, tmp_path / "out.txt", 1, 0)
    assert (tmp_path / "out.txt").read_text() == dic_text(rows)


def test_dic_out_restart_round_trip(tmp_path):
    rows = random_rows(np.random.default_rng(12), 25)
    (tmp_path / "a.txt").write_text(dic_text(rows))
    run("dic", tmp_path / "a.txt", tmp_path / "b.txt", 1, 0)
    run("dic", tmp_path / "b.txt", tmp_path / "c.txt", 1, 0)
    b, c = (tmp_path / "b.txt").read_text(), (tmp_path / "c.txt").read_text()
    assert b == c == dic_text([[r[0], r[1]] + [float(f"{v:.15e}") for v in r[2:]] for r in rows])


def test_simple_start_format(tmp_path):
    # real coordinates: the integer part is the reference pixel, the fraction an initial
    # displacement dx = upix (src/p-tracker.cpp:1265-1270); simpleVersion writes x y rot r
    (tmp_path / "s.txt").write_text("3\n10.25 20.5 0.1 12\n-3.25 7 0 5.5\n100 200.75 -0.2 30\n")
    run("dic", tmp_path / "s.txt", tmp_path / "o.txt", 0, 0)
    lines = (tmp_path / "o.txt").read_text().split("\n")
    r = [l.split() for l in lines[1:4]]
    assert [int(x[0]) for x in r] == [10, -4, 100] and [int(x[1]) for x in r] == [20, 7, 200]
    assert [float(x[4]) for x in r] == [0.25, 0.75, 0.0]  # dx
    assert [float(x[5]) for x in r] == [0.5, 0.0, 0.75]   # dy
    assert [float(x[7]) for x in r] == [0.25, 0.75, 0.0]  # upix
    run("dic", tmp_path / "s.txt", tmp_path / "p.txt", 0, 1)
    assert (tmp_path / "p.txt").read_text() == "3\n10.25 20.5 0.1 12\n-3.25 7 0 5.5\n100 200.75 -0.2 30\n"


# ---- command files --------------------------------------------------------------------------
def parse(tmp_path, text, files=None):
    for k, v in (files or {}).items():
        (tmp_path / k).write_text(v)
    (tmp_path / "cmd.txt").write_text(text)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        run("parse", "cmd.txt", "out.txt")
    finally:
        os.chdir(cwd)
    out = {}
    grains = []
    unknown = []
    for line in (tmp_path / "out.txt").read_text().splitlines():
        if line.startswith("Unknown token: "):
            unknown.append(line.split(": ", 1)[1])
        elif line.startswith("grain "):
            grains.append(line.split()[2:])
        else:
            k, _, v = line.partition(" ")
            out[k] = v
    return out, grains, unknown


def test_command_file_defaults(tmp_path):
    # the reference's defaults (src/p-tracker.hpp:67-152, src/p-tracker.cpp:20-52)
    s, g, unk = parse(tmp_path, "# nothing but a comment\n")
    assert s["procedure"] == "particle_tracking" and s["rotations"] == "1" and s["rescue_level"] == "2"
    assert float(s["NCC_min"]) == 0.7 and float(s["NCC_min_super"]) == 0.5 and float(s["subpix_tol"]) == 1e-8
    assert [float(x) for x in s["directions"].split()] == [0.01, 0.01, 0.01]
    assert float(s["subpix_displacement_max"]) == 3.0 and float(s["overflow_stiffness"]) == 0.0
    assert s["image_interpolator"] == "cubic" and s["wanted_num_threads"] == "4" and s["optim_algo"] == "0"
    assert [float(x) for x in s["disto"].split()] == [-1, -1, 0, 0, 0, 0, 0, 0, -1, -1, 0, 0, 0, 0]
    assert g == [] and unk == []


def test_command_file_keywords_and_order(tmp_path):
    text = """# comment line
image_name img%04d.tif   ! inline comment
grains_to_follow start.txt
iref 3
ibeg 4
iend 9
iinc 2
iraz 5
rescue_level 1
NCC_min 0.85
NCC_min_super 0.6
subpixel 0
subpix_tol 1e-10
initial_direction_dx 0.02
initial_direction_dy 0.03
initial_direction_dr 0.004
overflow_stiffness 2.5
subpix_displacement_max 4
wanted_num_threads -6
optim_algo LEVMAR
search_zone.left 7
search_zone.right 8
search_zone.up 9
search_zone.down 10
search_zone.inc_rot 0.01
search_zone.num_rot 10
search_zone_rescue.left 40
search_zone_rescue.num_rot 3
search_zone_super_rescue.down 50
search_zone.bogus 3
pattern circ 5
targetRadiusPattern 12
pattern grain_diameter 0.5 3
targetRadiusPattern -1
mask_rect 0 15 0 25
K1_corrDistor -2e-9
dK2_corrDistor 1e-18
frobnicate
"""
    start = "3\n10.5 20.25 0.0 12\n30 40 0.5 4\n50.75 60 0 12\n"
    s, g, unk = parse(tmp_path, text, {"start.txt": start})
    assert s["image_name"] == "img%04d.tif"
    assert (s["iref"], s["ibeg"], s["iend"], s["iinc"], s["iraz"]) == ("3", "4", "9", "2", "5")
    assert s["rescue_level"] == "1" and float(s["NCC_min"]) == 0.85 and float(s["NCC_min_super"]) == 0.6
    assert s["subpixel"] == "0" and float(s["subpix_tol"]) == 1e-10
    assert [float(x) for x in s["directions"].split()] == [0.02, 0.03, 0.004]
    assert float(s["overflow_stiffness"]) == 2.5 and float(s["subpix_displacement_max"]) == 4.0
    assert s["wanted_num_threads"] == "6" and s["optim_algo"] == "2"
    assert s["search_zone"].split() == ["7", "8", "9", "10", "10", "0.01"]
    assert s["search_zone_rescue"].split()[:5] == ["40", "0", "0", "0", "3"]
    assert s["search_zone_super_rescue"].split()[3] == "50"
    d = [float(x) for x in s["disto"].split()]
    assert d[2] == -2e-9
    # an unknown keyword leaves its argument to be read as the next token, as in the reference
    assert unk == ["search_zone.bogus", "3", "frobnicate"]
    # patterns apply in file order: circ 5 for all, then radius-12 grains get int(12 * 0.5) = 6
    sizes = [int(x[14]) for x in g]
    n5, n6 = len(synth.circ_pattern(5)), len(synth.circ_pattern(6))
    assert sizes == [n6, n5, n6]
    assert [x[15:19] for x in g] == [["0", "-6", "0", "6"], ["0", "-5", "0", "5"], ["0", "-6", "0", "6"]]
    # mask_rect 0..15 x 0..25 holds only the first grain (10, 20)
    assert [x[13] for x in g] == ["1", "0", "0"]


def test_rotations_zero_searches_one_step(tmp_path):
    s, _, _ = parse(tmp_path, "rotations 0\nsearch_zone.num_rot 7\nsearch_zone.inc_rot 0.02\n")
    for z in ("search_zone", "search_zone_rescue", "search_zone_super_rescue"):
        v = s[z].split()
        assert v[4] == "0" and float(v[5]) == 1.0
    steps = ptref.rotation_steps(1.0, 0)
    assert len(steps) == 1 and math.copysign(1.0, steps[0]) == -1.0  # one step at -0.0


def test_ring_rect_and_file_patterns(tmp_path):
    start = "2\n100 100 0 10\n200 200 0 10\n"
    s, g, _ = parse(tmp_path, "start_file start.txt\npattern ring 3 6\n", {"start.txt": start})
    ring = [(dx, dy) for dy in range(-6, 7) for dx in range(-6, 7) if 3 <= math.sqrt(dx * dx + dy * dy) <= 6]
    assert int(g[0][14]) == len(ring) and g[0][15:19] == ["0", "-6", "0", "6"]
    s, g, _ = parse(tmp_path, "start_file start.txt\npattern rect 4 2\n", {"start.txt": start})
    assert int(g[0][14]) == 9 * 5 and g[0][15:19] == ["-4", "-2", "4", "2"]
    pat = "2\n3\n0 0\n1 0\n0 1\n2\n-1 -1\n5 6\n"
    s, g, _ = parse(tmp_path, "start_file start.txt\npattern file pat.txt\n", {"start.txt": start, "pat.txt": pat})
    assert [x[14:] for x in g] == [["3", "0", "0", "0", "1"], ["2", "-1", "-1", "5", "6"]]


def test_restart_file_reads_dic_out(tmp_path):
    rows = random_rows(np.random.default_rng(13), 4)
    s, g, _ = parse(tmp_path, "restart_file prev.txt\n", {"prev.txt": dic_text(rows)})
    assert len(g) == 4
    for r, x in zip(rows, g):
        assert int(x[0]) == r[0] and int(x[1]) == r[1]
        assert [float(v) for v in x[2:13]] == [float(f"{v:.15e}") for v in r[2:13]]


def test_make_grid_uses_c_library_rand(tmp_path):
    run("grid", 100, 1000, 50, 500, 4, 3, 5, tmp_path / "g.txt")
    got = [tuple(map(int, l.split())) for l in (tmp_path / "g.txt").read_text().splitlines()]
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)  # the C library's default seed
    RAND_MAX = 2147483647
    dx, dy = int(math.floor(900 / 3)), int(math.floor(450 / 2))

    def rnd(v):  # std::round, half away from zero
        return int(math.floor(abs(v) + 0.5)) * (1 if v >= 0 else -1)

    exp = []
    for iy in range(3):
        for ix in range(4):
            angle = libc.rand() / RAND_MAX * (2.0 * math.pi)
            alea = libc.rand() / RAND_MAX * 5
            exp.append((int(100 + ix * dx + rnd(alea * math.cos(angle))), int(50 + iy * dy + rnd(alea * math.sin(angle)))))
    assert got == exp
