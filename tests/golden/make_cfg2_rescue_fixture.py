"""Writes cfg2_rescue_oracle.npz, the oracle's results on the config-2 rescue variant: the
6000 x 4000 12-bit pair of ptracker_amd.synth.workload(2, far=FAR) (seed 1002), where 300 grains
move 21-35 px along one axis, beyond the +-20 px main search zone and inside the +-40 px rescue
zone (21 rotation steps, NCC_min 0.7, rescue level 1, no subpixel stage).

Stored arrays: digest (SHA-256 of the generated frames, patterns and grain records, as
make_cfg2_full_fixture.digest), far (the moved grains), counts = (rescued, super-rescue queue),
and the columns after the main search and the rescue (dx, dy, drot, upix, vpix, rot_inc, NCC,
NCC_rescue).

The oracle is the CPU restatement under oracle/ (8 threads, a few minutes). Run from the
repository root:  python tests/golden/make_cfg2_rescue_fixture.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)
import ptref  # noqa: E402

pk = ptref.package()
from ptracker_amd import synth  # noqa: E402
from make_cfg2_full_fixture import SEARCH_COLS, digest, params  # noqa: E402

FAR = (300, 21.0, 35.0)


def main():
    w = synth.workload(2, far=FAR)
    W = ptref.Workload(w.I0, w.I1, w.grains, w.pattern_xy, w.pattern_begin, w.bits)
    t = time.time()
    gs, counts, _ = ptref.track_pair(W, params(w, 0), nthreads=8)
    print(f"search + rescue: {time.time() - t:.1f} s, counts {tuple(counts)}", flush=True)
    out = {"digest": np.array(digest(w)), "far": w.truth["far"].astype(np.int32),
           "counts": np.asarray(counts, dtype=np.int32)}
    for c in SEARCH_COLS:
        out[c] = np.asarray(gs[c], dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "cfg2_rescue_oracle.npz"), **out)


if __name__ == "__main__":
    main()
