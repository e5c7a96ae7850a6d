"""Writes cfg3_full_oracle.npz, the oracle's main-search results on the full-size 40-step case:
the 6000 x 4000 frames with 2,000 disk grains of ptracker_amd.synth.workload(3) (seed 1003),
+-20 px shifts and num_rot 20 at inc 0.005, which the reference's rotation loop makes 40 steps;
once on the 12-bit frames and once on their 8-bit requantisation (v >> 4).

Stored arrays, for b in (12, 8):
  digest_<b>  SHA-256 of the generated frames, patterns and grain records;
  <col>_<b>   dx, dy, drot, NCC after the main search (no rescue, no subpixel stage).

The oracle is the CPU restatement under oracle/ (run with 8 threads, about four minutes). Run from
the repository root:  python tests/golden/make_cfg3_full_fixture.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_cfg2_full_fixture import digest, ptref, pk  # noqa: E402
from ptracker_amd import synth  # noqa: E402

COLS = ("dx", "dy", "drot", "NCC")


def workload(bits):
    w = synth.workload(3)
    if bits == 8:
        w.I0 = (w.I0 >> 4).astype(np.uint16)
        w.I1 = (w.I1 >> 4).astype(np.uint16)
        w.bits = 8
    return w


def params(w):
    return pk.track_params(w.params.search_zone, rescue_level=0, subpixel=0, search_method=pk.SEARCH_PLAIN)


def main():
    out = {}
    for bits in (12, 8):
        w = workload(bits)
        assert len(ptref.rotation_steps(w.params.search_zone.inc_rot, w.params.search_zone.num_rot)) == 40
        W = ptref.Workload(w.I0, w.I1, w.grains, w.pattern_xy, w.pattern_begin, w.bits)
        t = time.time()
        g, _, _ = ptref.track_pair(W, params(w), nthreads=8)
        print(f"{bits}-bit: {time.time() - t:.1f} s", flush=True)
        out[f"digest_{bits}"] = np.array(digest(w))
        for c in COLS:
            out[f"{c}_{bits}"] = np.asarray(g[c], dtype=np.float64)
    path = os.path.join(HERE, "cfg3_full_oracle.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
