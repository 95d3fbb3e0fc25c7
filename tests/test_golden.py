"""The committed reference fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the amdflang build of the reference's
modules): one residual evaluation on seeded inputs, and (lusgs_*.npz) one
'implicit' time step, the residual evaluation followed by the reference's
LU-SGS update.  The oracle must reproduce them bit for bit on the CPU; the
HIP path, called through the C ABI, within 1e-12 relative on the GPU."""
import glob
import os

import numpy as np
import pytest

from binding import Oracle
from festgpu import cases as C

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ALL = sorted(glob.glob(os.path.join(HERE, "*.npz")))
LUSGS = [p for p in ALL if os.path.basename(p).startswith("lusgs_")]
FIXTURES = [p for p in ALL if p not in LUSGS]


def load(path):
    z = np.load(path, allow_pickle=False)
    case = C.case_from_json(str(z["meta"]))
    return case, z


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_oracle_reproduces_reference_fixture(path):
    case, z = load(path)
    o = Oracle(case)
    o.set_nodes(z["in_nodes"])
    assert np.array_equal(o.arr("nodes"), z["in_nodes"])
    o.arr("dist")[...] = z["in_dist"]
    o.set_state(z["in_qp"])
    o.arr("temp")[...] = z["in_temp"]
    assert o.residual(int(z["iter"])) == 0
    for name in ("qp", "F", "G", "H", "res"):
        assert np.array_equal(o.arr(name), z["out_" + name]), name


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_gpu_reproduces_reference_fixture(path):
    from festgpu.native import GpuBlock
    case, z = load(path)
    print("golden", os.path.basename(path), flush=True)
    g = GpuBlock(case)
    g.set_grid(z["in_nodes"], wall_dist=(case.turb == "sst"))
    if case.turb == "sst":
        assert np.array_equal(g.download("dist"), z["in_dist"])
    g.upload_state(z["in_qp"], z["in_temp"])
    g.residual(int(z["iter"]))
    out = g.download("qp", "F", "G", "H", "residue")
    for name, ref in (("qp", "qp"), ("F", "F"), ("G", "G"), ("H", "H"), ("residue", "res")):
        b = z["out_" + ref]
        err = np.abs(out[name] - b).max() / np.abs(b).max()
        assert err <= 1e-12, (name, err)


@pytest.mark.parametrize("path", LUSGS, ids=[os.path.basename(p)[:-4] for p in LUSGS])
def test_oracle_reproduces_reference_lusgs_step(path):
    case, z = load(path)
    o = Oracle(case)
    o.set_nodes(z["in_nodes"])
    o.arr("dist")[...] = z["in_dist"]
    o.set_state(z["in_qp"])
    o.arr("temp")[...] = z["in_temp"]
    assert o.step(int(z["iter"])) == 0
    assert np.array_equal(o.arr("res"), z["out_res"])
    assert np.array_equal(o.arr("dt"), z["in_dt"])
    assert np.array_equal(o.arr("qp")[:, 3:-3, 3:-3, 3:-3], z["out_qp"][:, 3:-3, 3:-3, 3:-3])


@pytest.mark.gpu
@pytest.mark.parametrize("path", LUSGS, ids=[os.path.basename(p)[:-4] for p in LUSGS])
def test_gpu_reproduces_reference_lusgs_step(path):
    from festgpu.native import GpuBlock
    case, z = load(path)
    print("golden", os.path.basename(path), flush=True)
    g = GpuBlock(case)
    g.set_grid(z["in_nodes"], wall_dist=(case.turb == "sst"))
    g.upload_state(z["in_qp"], z["in_temp"])
    g.get_next_solution(int(z["iter"]))
    out = g.download("qp", "residue", "dt")
    for name, ref in (("residue", z["out_res"]), ("dt", z["in_dt"]),
                      ("qp", z["out_qp"][:, 3:-3, 3:-3, 3:-3])):
        got = out[name][:, 3:-3, 3:-3, 3:-3] if name == "qp" else out[name]
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err <= 1e-12, (name, err)
