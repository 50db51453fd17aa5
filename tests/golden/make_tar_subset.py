"""Writes tests/golden/tar_subset_2000.dat: the 2,000 particles nearest the
origin of the reference's input/tar.dat (a 15,000-particle snapshot at
t = 1.0e4 s with tag-0 and tag-1 material), in the same ASCII format.
Runs only where the reference tree is mounted; the output file is the fixture.
"""
import os
import sys

import numpy as np

SRC = "/root/reference/input/tar.dat"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tar_subset_2000.dat")


def main():
    with open(SRC) as f:
        lines = f.read().splitlines()
    time_line, n = lines[0], int(lines[1])
    rows = lines[2:2 + n]
    pos = np.array([[float(v) for v in r.split()[3:6]] for r in rows])
    keep = np.sort(np.argsort(np.linalg.norm(pos, axis=1), kind="stable")[:2000])
    with open(OUT, "w") as f:
        f.write(time_line + "\n")
        f.write("%d\n" % len(keep))
        for k in keep:
            f.write(rows[k] + "\n")
    print("wrote", OUT, file=sys.stderr)


if __name__ == "__main__":
    main()
