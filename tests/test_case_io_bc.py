"""Boundary-file reading of the turbulence values (read_bc.f90:16-202):
FIX_tkl under k-kL, whose omega values are neither filled nor read."""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "fest-3d_amd"))

from festgpu import case_io, cases as C  # noqa: E402

BC = "#imin\n - FIX_tkl 0.5\n - FIX_tw 9.0\n - FIX_tk 0.25\n#imax\n#jmin\n#jmax\n#kmin\n#kmax\n"


def test_k_kl_reads_fixed_kl_and_ignores_omega(tmp_path):
    case, _, _ = C.sst_bump(8, 6, 4, turb="kkl")
    p = tmp_path / "bc.md"
    p.write_text(BC)
    r = case_io.read_bc(str(p), case)
    assert r.fix_tkl == (0.5,) + (case.tkl_inf,) * 5
    assert r.fix_tk == (0.25,) + (case.tk_inf,) * 5
    assert r.fix_tw == (0.0,) * 6


def test_sst_reads_omega(tmp_path):
    case, _, _ = C.sst_bump(8, 6, 4, turb="sst")
    p = tmp_path / "bc.md"
    p.write_text(BC)
    r = case_io.read_bc(str(p), case)
    assert r.fix_tw == (9.0,) + (case.tw_inf,) * 5
