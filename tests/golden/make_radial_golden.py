"""Writes radial_ref.jsonl: the reference's own calculate_radial (chebyshev x
quadratic / cosine cutoff) evaluated by oracle/_ref/ref_radial, which the
oracle Makefile builds from the reference's radial.cpp and cutoffs.cpp.

Run from the repository root:  python tests/golden/make_radial_golden.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle  # noqa: E402

if not pyoracle.build_ref():
    sys.exit("reference sources not present; fixture left unchanged")
exe = os.path.join(ROOT, "oracle", "_ref", "ref_radial")
out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
with open(os.path.join(HERE, "radial_ref.jsonl"), "w") as f:
    f.write(out)
print(f"wrote {len(out.splitlines())} records")
