"""Extract the small data fixtures the tests use from the reference's test cases
(testCases/*, data files only).  Run in a container that has the reference mounted:

    python tests/golden/make_fixtures.py /root/reference

Writes, next to this script:
  ghia_re100_centerline.csv   y, u/U of Ghia, Ghia & Shin (1982), Re=100 (cavityFlow/Ghia.dat)
  cavity_re100_u_x05.csv      y, u of the reference's own 100x100 cavity run (cavityFlow/u_x05.csv)
  poiseuille_u.csv            y, u of the reference's own 100x20 channel run (poiseuilleFlow/u.csv)
  cylinder.stl.gz             walls STL of flowAroundCylinder (one patch, 3664 triangles), gzip'ed
  backStepBoxes.stl.gz        walls STL of backStepFlow (patches Top, Bottom), gzip'ed
"""
import csv
import gzip
import os
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
here = os.path.dirname(os.path.abspath(__file__))


def read_csv_cols(path, cols):
    with open(path) as fh:
        rows = list(csv.DictReader(fh))
    return [[float(r[c]) for c in cols] for r in rows]


def write(name, header, rows):
    with open(os.path.join(here, name), "w") as fh:
        fh.write(",".join(header) + "\n")
        for r in rows:
            fh.write(",".join(f"{x:.9g}" for x in r) + "\n")


ghia = []
with open(os.path.join(ref, "testCases/cavityFlow/Ghia.dat")) as fh:
    next(fh)
    for line in fh:
        a, b = line.replace(",", " ").split()
        ghia.append([float(a), float(b)])
write("ghia_re100_centerline.csv", ["y", "u"], ghia)
write("cavity_re100_u_x05.csv", ["y", "u"],
      [[p, u] for u, p in read_csv_cols(os.path.join(ref, "testCases/cavityFlow/u_x05.csv"), ["u", "Points:1"])])
write("poiseuille_u.csv", ["y", "u"],
      [[p, u] for u, p in read_csv_cols(os.path.join(ref, "testCases/poiseuilleFlow/u.csv"), ["u", "Points:1"])])
for src, name in (("testCases/flowAroundCylinder/cylinder.stl", "cylinder.stl.gz"),
                  ("testCases/backStepFlow/backStepBoxes.stl", "backStepBoxes.stl.gz")):
    with open(os.path.join(ref, src), "rb") as fh, gzip.GzipFile(os.path.join(here, name), "wb", 9, mtime=0) as out:
        out.write(fh.read())
print("fixtures written to", here)
