"""ctypes bindings of the checkers: the C restatement (oracle/liboracle.so)
and the amdflang build of the reference modules (oracle/_ref/libfestref.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libfestref.so")

ARRAY_IDS = ["qp", "temp", "mu", "mut", "f1", "dist", "vol", "cx", "cy", "cz",
             "iA", "inx", "iny", "inz", "jA", "jnx", "jny", "jnz", "kA", "knx", "kny", "knz",
             "xl", "xr", "yl", "yr", "zl", "zr", "F", "G", "H", "res", "dt", "ustore", "rstore",
             "gx", "gy", "gz", "nodes"]
AID = {n: i for i, n in enumerate(ARRAY_IDS)}


class CCase(ctypes.Structure):
    _fields_ = [("imx", ctypes.c_int32), ("jmx", ctypes.c_int32), ("kmx", ctypes.c_int32),
                ("n_var", ctypes.c_int32), ("flux", ctypes.c_int32), ("recon", ctypes.c_int32),
                ("turb", ctypes.c_int32), ("time_local", ctypes.c_int32),
                ("time_accuracy", ctypes.c_int32), ("accur", ctypes.c_int32),
                ("lim", ctypes.c_int32 * 3), ("tlim", ctypes.c_int32 * 3),
                ("mu_sutherland", ctypes.c_int32), ("bc_id", ctypes.c_int32 * 6),
                ("global_time_step", ctypes.c_double), ("CFL", ctypes.c_double),
                ("rho_inf", ctypes.c_double), ("u_inf", ctypes.c_double), ("v_inf", ctypes.c_double),
                ("w_inf", ctypes.c_double), ("p_inf", ctypes.c_double), ("tk_inf", ctypes.c_double),
                ("tw_inf", ctypes.c_double), ("vel_mag", ctypes.c_double), ("MInf", ctypes.c_double),
                ("gm", ctypes.c_double), ("R_gas", ctypes.c_double), ("mu_ref", ctypes.c_double),
                ("T_ref", ctypes.c_double), ("S_ref", ctypes.c_double), ("Pr", ctypes.c_double),
                ("tPr", ctypes.c_double),
                ("fix_rho", ctypes.c_double * 6), ("fix_u", ctypes.c_double * 6),
                ("fix_v", ctypes.c_double * 6), ("fix_w", ctypes.c_double * 6),
                ("fix_p", ctypes.c_double * 6), ("fix_tk", ctypes.c_double * 6),
                ("fix_tw", ctypes.c_double * 6), ("fix_twall", ctypes.c_double * 6),
                ("fix_tp", ctypes.c_double * 6), ("tv_inf", ctypes.c_double),
                ("fix_tv", ctypes.c_double * 6), ("transition", ctypes.c_int32), ("tu_inf", ctypes.c_double),
                ("tkl_inf", ctypes.c_double), ("fix_tkl", ctypes.c_double * 6),
                ("pbs", ctypes.c_int32 * 3), ("muscl_limiter", ctypes.c_int32),
                ("tgm_inf", ctypes.c_double), ("fix_tgm", ctypes.c_double * 6)]


def to_ccase(case) -> CCase:
    from festgpu.cases import FLUX, RECON, TURB, TIME_ACC
    c = CCase()
    c.imx, c.jmx, c.kmx, c.n_var = case.imx, case.jmx, case.kmx, case.n_var
    c.flux, c.recon, c.turb = FLUX[case.flux], RECON[case.recon], TURB[case.turb]
    c.time_local, c.time_accuracy, c.accur = case.time_local, TIME_ACC[case.time_accuracy], case.accur
    c.lim[:] = list(case.lim)
    c.tlim[:] = list(case.tlim)
    c.mu_sutherland = case.mu_sutherland
    c.bc_id[:] = list(case.bc_id)
    for f in ("global_time_step", "CFL", "rho_inf", "u_inf", "v_inf", "w_inf", "p_inf", "tk_inf",
              "tw_inf", "vel_mag", "MInf", "gm", "R_gas", "mu_ref", "T_ref", "S_ref", "Pr", "tPr"):
        setattr(c, f, float(getattr(case, f)))
    c.tv_inf = float(case.tv_inf)
    c.transition = {"none": 0, "bc": 1, "lctm2015": 2}[case.transition]
    c.tgm_inf = float(case.tgm_inf)
    c.tu_inf = float(case.tu_inf)
    c.tkl_inf = float(case.tkl_inf)
    c.pbs[:] = [int(v) for v in case.pbs]
    c.muscl_limiter = {"koren": 0, "van_albada": 1}[case.limiter]
    for f in ("fix_rho", "fix_u", "fix_v", "fix_w", "fix_p", "fix_tk", "fix_tw", "fix_twall", "fix_tp", "fix_tv", "fix_tkl",
              "fix_tgm"):
        getattr(c, f)[:] = [float(v) for v in getattr(case, f)]
    return c


def array_shape(case, name):
    """numpy (C-order) shape of each array in the Fortran layout."""
    imx, jmx, kmx, nv = case.imx, case.jmx, case.kmx, case.n_var
    ng = case.n_grad
    cell = (kmx + 5, jmx + 5, imx + 5)
    shapes = {
        "qp": (nv,) + cell, "ustore": (nv,) + cell,
        "temp": cell, "mu": cell, "mut": cell, "f1": cell, "dist": cell,
        "vol": cell, "cx": cell, "cy": cell, "cz": cell,
        "iA": (kmx + 5, jmx + 5, imx + 6), "jA": (kmx + 5, jmx + 6, imx + 5), "kA": (kmx + 6, jmx + 5, imx + 5),
        "xl": (nv, kmx - 1, jmx - 1, imx + 2), "yl": (nv, kmx - 1, jmx + 2, imx - 1),
        "zl": (nv, kmx + 2, jmx - 1, imx - 1),
        "F": (nv, kmx - 1, jmx - 1, imx), "G": (nv, kmx - 1, jmx, imx - 1), "H": (nv, kmx, jmx - 1, imx - 1),
        "res": (nv, kmx - 1, jmx - 1, imx - 1), "rstore": (nv, kmx - 1, jmx - 1, imx - 1),
        "dt": (kmx - 1, jmx - 1, imx - 1),
        "gx": (ng, kmx + 1, jmx + 1, imx + 1),
        "nodes": (3, kmx + 6, jmx + 6, imx + 6),
    }
    for a, b in (("inx", "iA"), ("iny", "iA"), ("inz", "iA"), ("jnx", "jA"), ("jny", "jA"), ("jnz", "jA"),
                 ("knx", "kA"), ("kny", "kA"), ("knz", "kA"), ("xr", "xl"), ("yr", "yl"), ("zr", "zl"),
                 ("gy", "gx"), ("gz", "gx")):
        shapes[a] = shapes[b]
    return shapes[name]


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run make -C oracle)")
    return ctypes.CDLL(path)


class CLink(ctypes.Structure):
    """orc_link (oracle/fest_oracle.h), one interface link."""
    _fields_ = [(n, ctypes.c_int32) for n in
                ("block", "face", "from_block", "from_face", "lo1", "hi1", "lo2", "hi2", "dir_switch")]


def _links(links):
    arr = (CLink * max(len(links), 1))()
    for n, L in enumerate(links):
        arr[n] = CLink(*[int(v) for v in L])
    return arr


def group_exchange(blocks, links):
    """The interface exchange over Oracle blocks (src/interface1.f90)."""
    hs = (ctypes.c_void_p * len(blocks))(*[o.h for o in blocks])
    return Oracle.lib().orc_exchange(hs, len(blocks), _links(links), len(links))


def group_wall_dist(blocks):
    """The wall distance of every Oracle block to the no-slip surface nodes
    of all blocks (src/wall/wall.f90:55-111, src/wall/wall_dist.f90)."""
    arr = (ctypes.c_void_p * len(blocks))(*[o.h for o in blocks])
    return Oracle.lib().orc_group_wall_dist(arr, len(blocks))


def group_step(blocks, links, current_iter):
    """One time step of every block, the exchange opening each residual."""
    hs = (ctypes.c_void_p * len(blocks))(*[o.h for o in blocks])
    return Oracle.lib().orc_group_step(hs, len(blocks), _links(links), len(links), int(current_iter))


def group_step_mt(blocks, links, current_iter, nthreads):
    """group_step with the blocks of each stage spread over `nthreads`
    threads (the GIL is released for the whole call)."""
    hs = (ctypes.c_void_p * len(blocks))(*[o.h for o in blocks])
    return Oracle.lib().orc_group_step_mt(hs, len(blocks), _links(links), len(links), int(current_iter),
                                          int(nthreads))


class Oracle:
    """One block of the C restatement."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            L = _load(ORACLE_SO)
            L.orc_create.restype = ctypes.c_void_p
            L.orc_create.argtypes = [ctypes.POINTER(CCase)]
            L.orc_destroy.argtypes = [ctypes.c_void_p]
            L.orc_array.restype = ctypes.POINTER(ctypes.c_double)
            L.orc_array.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]
            for f in ("orc_geometry", "orc_wall_dist", "orc_time_step"):
                getattr(L, f).argtypes = [ctypes.c_void_p]
                getattr(L, f).restype = ctypes.c_int
            for f in ("orc_residual", "orc_step"):
                getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_int]
                getattr(L, f).restype = ctypes.c_int
            L.orc_stage.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
            L.orc_stage.restype = ctypes.c_int
            dp = ctypes.POINTER(ctypes.c_double)
            L.orc_face_flux.argtypes = [ctypes.c_void_p, dp, dp] + [ctypes.c_double] * 4 + [ctypes.c_int, dp]
            L.orc_temp_refresh.argtypes = [ctypes.c_void_p]
            L.orc_resnorm.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
            L.orc_exchange.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.POINTER(CLink),
                                       ctypes.c_int]
            L.orc_exchange.restype = ctypes.c_int
            L.orc_group_step.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.POINTER(CLink),
                                         ctypes.c_int, ctypes.c_int]
            L.orc_group_step.restype = ctypes.c_int
            L.orc_group_step_mt.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.POINTER(CLink),
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int]
            L.orc_group_step_mt.restype = ctypes.c_int
            cls._lib = L
        return cls._lib

    def __init__(self, case):
        self.case = case
        self._c = to_ccase(case)
        self.h = self.lib().orc_create(ctypes.byref(self._c))
        if not self.h:
            raise MemoryError("orc_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib().orc_destroy(self.h)
            self.h = None

    def arr(self, name) -> np.ndarray:
        n = ctypes.c_size_t()
        p = self.lib().orc_array(self.h, AID[name], ctypes.byref(n))
        a = np.ctypeslib.as_array(p, shape=(n.value,))
        return a.reshape(array_shape(self.case, name))

    def set_nodes(self, nodes):
        self.arr("nodes")[...] = nodes
        st = self.lib().orc_geometry(self.h)
        if st:
            raise RuntimeError(f"orc_geometry failed: {st}")

    def wall_dist(self):
        return self.lib().orc_wall_dist(self.h)

    def set_state(self, q):
        self.arr("qp")[...] = q

    def temp_refresh(self):
        self.lib().orc_temp_refresh(self.h)

    def residual(self, it=1):
        return self.lib().orc_residual(self.h, it)

    def time_step(self):
        return self.lib().orc_time_step(self.h)

    def stage(self, it, mask):
        return self.lib().orc_stage(self.h, it, mask)

    def face_flux(self, qL, qR, n, A=1.0, zero=1):
        nv = self.case.n_var
        a = np.ascontiguousarray(qL, dtype=np.float64)
        b = np.ascontiguousarray(qR, dtype=np.float64)
        out = np.zeros(nv)
        dp = ctypes.POINTER(ctypes.c_double)
        self.lib().orc_face_flux(self.h, a.ctypes.data_as(dp), b.ctypes.data_as(dp), float(n[0]), float(n[1]),
                                 float(n[2]), float(A), int(zero), out.ctypes.data_as(dp))
        return out

    def step(self, it):
        return self.lib().orc_step(self.h, it)

    def resnorm(self):
        out = (ctypes.c_double * 16)()
        self.lib().orc_resnorm(self.h, out)
        return np.array(out[: self.case.n_var + 1])


def ref_available() -> bool:
    return os.path.exists(REF_SO)


_REF_SCRIPT = r"""
import ctypes, sys, numpy as np, pickle
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[2])
from binding import CCase, to_ccase, array_shape, AID, REF_SO
job = pickle.load(open(sys.argv[3], 'rb'))
case = job['case']
L = ctypes.CDLL(REF_SO)
c = to_ccase(case)
L.ref_init(ctypes.byref(c))
dp = ctypes.POINTER(ctypes.c_double)
def ptr(a): return a.ctypes.data_as(dp)
nodes = np.ascontiguousarray(job['nodes'], dtype=np.float64)
L.ref_set_nodes(ptr(nodes))
if job.get('wall_dist'):
    path = job['tmpfile'].encode()
    L.ref_wall_dist(ctypes.c_char_p(path), ctypes.c_int(len(path)))
elif job.get('dist') is not None:
    d = np.ascontiguousarray(job['dist'], dtype=np.float64); L.ref_set_dist(ptr(d))
q = np.ascontiguousarray(job['qp'], dtype=np.float64)
t = np.ascontiguousarray(job['temp'], dtype=np.float64)
L.ref_set_state(ptr(q), ptr(t))
if job.get('stage', 0) > 0:
    L.ref_stage(ctypes.c_int(job.get('iter', 1)), ctypes.c_int(job['stage']))
if job.get('lusgs_dt') is not None:
    d = np.ascontiguousarray(job['lusgs_dt'], dtype=np.float64); L.ref_lusgs(ptr(d))
out = {}
for name in job['want']:
    shp = array_shape(case, name)
    a = np.zeros(int(np.prod(shp)))
    L.ref_get(ctypes.c_int(AID[name]), ptr(a))
    out[name] = a.reshape(shp)
pickle.dump(out, open(sys.argv[4], 'wb'))
"""


def run_reference(case, nodes, qp, temp, stage, want, it=1, dist=None, wall_dist=False, tmpdir="/tmp",
                  lusgs_dt=None):
    """Runs the reference modules in a fresh process (their module state,
    e.g. the sticky ppm_flag of boundary_state_reconstruction, persists
    within one process) and returns the requested arrays."""
    import pickle
    import tempfile
    pkg = os.path.join(os.path.dirname(HERE), "fest-3d_amd")
    with tempfile.TemporaryDirectory(dir=tmpdir) as td:
        jin, jout = os.path.join(td, "in.pkl"), os.path.join(td, "out.pkl")
        script = os.path.join(td, "run.py")
        with open(script, "w") as f:
            f.write(_REF_SCRIPT)
        job = dict(case=case, nodes=nodes, qp=qp, temp=temp, stage=stage, want=list(want), iter=it,
                   dist=dist, wall_dist=wall_dist, tmpfile=os.path.join(td, "surfnode.dat"),
                   lusgs_dt=lusgs_dt)
        with open(jin, "wb") as f:
            pickle.dump(job, f)
        r = subprocess.run([sys.executable, script, HERE, pkg, jin, jout], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference run failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        with open(jout, "rb") as f:
            return pickle.load(f)


_REF_TIME_SCRIPT = r"""
import ctypes, sys, time, numpy as np, pickle
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[2])
from binding import to_ccase, REF_SO
job = pickle.load(open(sys.argv[3], 'rb'))
case = job['case']
L = ctypes.CDLL(REF_SO)
c = to_ccase(case)
L.ref_init(ctypes.byref(c))
dp = ctypes.POINTER(ctypes.c_double)
def ptr(a): return a.ctypes.data_as(dp)
nodes = np.ascontiguousarray(job['nodes'], dtype=np.float64)
L.ref_set_nodes(ptr(nodes))
d = np.ascontiguousarray(job['dist'], dtype=np.float64); L.ref_set_dist(ptr(d))
q = np.ascontiguousarray(job['qp'], dtype=np.float64)
t = np.ascontiguousarray(job['temp'], dtype=np.float64)
L.ref_set_state(ptr(q), ptr(t))
L.ref_stage(ctypes.c_int(1), ctypes.c_int(5))   # first touch of every array
t0 = time.perf_counter()
n = 0
while True:
    L.ref_stage(ctypes.c_int(1), ctypes.c_int(5))
    n += 1
    if time.perf_counter() - t0 > job['budget'] or n >= job['reps']:
        break
pickle.dump({'seconds': (time.perf_counter() - t0) / n, 'reps': n}, open(sys.argv[4], 'wb'))
"""


def time_reference_residual(case, nodes, qp, temp, dist, reps=3, budget=10.0, tmpdir="/tmp"):
    """Seconds per residual evaluation of the reference's own modules
    (ref_stage(.., 5): ghost cells, face states, fluxes, gradients,
    viscosity, viscous fluxes, residue, source), run in a child process on
    one core; None when the reference build is absent."""
    if not ref_available():
        return None
    import pickle
    import tempfile
    pkg = os.path.join(os.path.dirname(HERE), "fest-3d_amd")
    with tempfile.TemporaryDirectory(dir=tmpdir) as td:
        jin, jout = os.path.join(td, "in.pkl"), os.path.join(td, "out.pkl")
        script = os.path.join(td, "run.py")
        with open(script, "w") as f:
            f.write(_REF_TIME_SCRIPT)
        with open(jin, "wb") as f:
            pickle.dump(dict(case=case, nodes=nodes, qp=qp, temp=temp, dist=dist, reps=reps, budget=budget), f)
        r = subprocess.run([sys.executable, script, HERE, pkg, jin, jout], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference timing failed ({r.returncode}):\n{r.stderr[-2000:]}")
        with open(jout, "rb") as f:
            return pickle.load(f)
