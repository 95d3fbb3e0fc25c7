"""Copies the data files of three of the reference's test cases from the
reference tree into tests/golden/<case>/, keeping their layout: the
system/*.md settings, the layout and interface map, the boundary files and
the grid files.  They are inputs only; the expected results are the
reference's own report (tests/Report.txt):

  SmoothBump  subsonic inviscid flow over a smooth bump, two 48x48x1
              blocks: relative change in entropy 7.883E-07
  Lfp         laminar flat plate, two blocks: drag coefficient 1.329E-03
  Tfp         turbulent (SST) flat plate, two blocks: drag coefficient
              2.872E-03, restarted (control.md "Restart level" 2) from the
              state its time_directories/0002 holds, also copied

The restart logs of the Tfp and SmoothBump dumps are copied too; they pin
the restart-log writer's line formats.

Run from the repository root:  python tests/golden/copy_cases.py
"""
import os
import shutil

SRC = "/root/reference/tests"
DST = os.path.dirname(os.path.abspath(__file__))
FILES = ["system/control.md", "system/fvscheme.md", "system/flow.md",
         "system/mesh/layout/layout.md", "system/mesh/layout/mapping.txt",
         "system/mesh/bc/bc_00.md", "system/mesh/bc/bc_01.md",
         "system/mesh/gridfiles/grid_00.txt", "system/mesh/gridfiles/grid_01.txt"]

RESTART = {"Tfp": ["system/output_control.md", "time_directories/0002/process_00.dat",
                   "time_directories/0002/process_01.dat",
                   "time_directories/0002/restart/process_00", "time_directories/0002/restart/process_01"],
           "SmoothBump": ["time_directories/0010/restart/process_00",
                          "time_directories/0010/restart/process_01"]}

if __name__ == "__main__":
    for case in ("SmoothBump", "Lfp", "Tfp"):
        for f in FILES + RESTART.get(case, []):
            dst = os.path.join(DST, case, f)
            os.makedirs(os.path.dirname(dst), exist_ok=True)
            shutil.copyfile(os.path.join(SRC, case, f), dst)
        print("copied", case)
