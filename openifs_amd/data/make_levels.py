"""Write the hybrid vertical-coordinate tables used by the synthetic atmospheres.

Run once in a checkout that has the OpenIFS sources next to it:

    python openifs_amd/data/make_levels.py /path/to/openifs

L137: the A (Pa) and B coefficients of the 138 half levels are read verbatim
from the single-column-model namelist NAMVV1 of the reference
(ifs-test/tests/scm/common/namelists:90-368, DVALH(0:137) then DVBH(0:137)).

L91: the reference holds no 91-level table.  This project defines one by
resampling L137 in half-level index space: half level j of L91 (j = 0..91)
sits at fractional L137 index x = j*137/91, and A, B are interpolated
linearly in x between the neighbouring L137 half levels.  Both ends (model
top p = 0 and surface A = 0, B = 1) are kept exactly, A stays >= 0 and the
half-level pressures stay strictly increasing for any surface pressure.
"""
import os
import re
import sys

import numpy as np


def read_l137(ref_root):
    path = os.path.join(ref_root, "ifs-test", "tests", "scm", "common", "namelists")
    with open(path) as f:
        text = f.read()
    block = text[text.index("&NAMVV1"):]
    block = block[: block.index("/")]
    a_part = block[block.index("DVALH(0)=") + len("DVALH(0)="): block.index("DVBH(0)=")]
    b_part = block[block.index("DVBH(0)=") + len("DVBH(0)="):]
    num = re.compile(r"[-+]?\d+\.?\d*(?:[eEdD][-+]?\d+)?")
    a = np.array([float(x.replace("D", "E").replace("d", "e")) for x in num.findall(a_part)])
    b = np.array([float(x.replace("D", "E").replace("d", "e")) for x in num.findall(b_part)])
    if a.size != 138 or b.size != 138:
        raise SystemExit(f"expected 138 A and B values, got {a.size}, {b.size}")
    return a, b


def resample(a, b, nlev):
    n_src = a.size - 1
    x = np.arange(nlev + 1) * (n_src / nlev)
    i0 = np.minimum(np.floor(x).astype(int), n_src - 1)
    w = x - i0
    a_new = (1 - w) * a[i0] + w * a[i0 + 1]
    b_new = (1 - w) * b[i0] + w * b[i0 + 1]
    a_new[0], b_new[0] = a[0], b[0]
    a_new[-1], b_new[-1] = a[-1], b[-1]
    return a_new, b_new


def write(path, a, b, source):
    with open(path, "w") as f:
        f.write(f"# hybrid half-level coefficients, {a.size - 1} levels; p = A + B*ps\n# {source}\n")
        for j, (aa, bb) in enumerate(zip(a, b)):
            f.write(f"{j:4d} {aa:.10f} {bb:.10f}\n")


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    a, b = read_l137(ref_root)
    write(os.path.join(here, "l137_ab.txt"), a, b, "ifs-test/tests/scm/common/namelists:90-368 (NAMVV1)")
    a91, b91 = resample(a, b, 91)
    write(os.path.join(here, "l91_ab.txt"), a91, b91, "project-defined: L137 resampled linearly in half-level index")


if __name__ == "__main__":
    main()
