"""Copies the reference's example data (two 380x380 RGB TIFF frames and the start file of
its 9 grains) into tests/golden/example/, the inputs of the end-to-end test of the
tracking program. Data only: the reference ships no expected output for them, so the
test compares the program's dic_out files with the oracle on the same decoded frames.

Run from the repository root in a checkout next to the reference:
    python tests/golden/make_example_fixtures.py /root/reference
"""
import os
import shutil
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "example")
os.makedirs(out, exist_ok=True)
for name in ("TEST01.tiff", "TEST02.tiff"):
    shutil.copyfile(os.path.join(ref, "examples", "ImageData", name), os.path.join(out, name))
# the example's command file names this start file recentered.txt
shutil.copyfile(os.path.join(ref, "examples", "ParticleImageTracking", "recentered.data"),
                os.path.join(out, "recentered.txt"))
print("wrote", sorted(os.listdir(out)))
