"""Copies the data files of the reference's SmoothBump case (subsonic
inviscid flow over a smooth bump, two 48x48x1 blocks joined by an
interface) from the reference tree into this folder, keeping its layout:
the system/*.md settings, the layout and interface map, the two boundary
files and the two grid files.  They are inputs only; the expected result is
the reference's own report (tests/Report.txt: relative change in entropy
7.883E-07 after the case's 3000 LU-SGS iterations).

Run from the repository root:  python tests/golden/SmoothBump/copy_case.py
"""
import os
import shutil

SRC = "/root/reference/tests/SmoothBump"
DST = os.path.dirname(os.path.abspath(__file__))
FILES = ["system/control.md", "system/fvscheme.md", "system/flow.md",
         "system/mesh/layout/layout.md", "system/mesh/layout/mapping.txt",
         "system/mesh/bc/bc_00.md", "system/mesh/bc/bc_01.md",
         "system/mesh/gridfiles/grid_00.txt", "system/mesh/gridfiles/grid_01.txt"]

if __name__ == "__main__":
    for f in FILES:
        os.makedirs(os.path.dirname(os.path.join(DST, f)), exist_ok=True)
        shutil.copyfile(os.path.join(SRC, f), os.path.join(DST, f))
        print("copied", f)
