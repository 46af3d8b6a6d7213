"""ctypes bindings of the oracle (oracle/liboracle.so) and fixture helpers.

Test infrastructure: only tests, the smoke check and the benchmark's CPU-baseline
leg import this module.
"""
import ctypes as C
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
VW, TW = 36, 18
SOLVERS = {"jc": 0, "gs": 1, "cg": 2, "mg": 3}

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(REPO, "oracle", "liboracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], check=True)
        _lib = C.CDLL(path)
        _declare(_lib)
    return _lib


PI = C.POINTER(C.c_int)
PD = C.POINTER(C.c_double)
PB = C.POINTER(C.c_byte)


class Mesh(C.Structure):
    _fields_ = [("L", C.c_int), ("lev", C.c_int), ("nel", C.c_int), ("nno", C.c_int),
                ("neq", C.c_int), ("neqT", C.c_int), ("gnox", C.c_int), ("gnoz", C.c_int),
                ("Mcode", PI), ("ien", PI), ("node", PI), ("id", PI), ("lm", PI),
                ("heat_id", PI), ("heat_lm", PI), ("v_bc", PB), ("v_bc_val", PD),
                ("T_bc", PB), ("T_bc_val", PD), ("X", PD), ("dx", PD), ("dz", PD),
                ("nh", C.c_int), ("h_ni", PI), ("h_ci", PI)]


class MgLevel(C.Structure):
    _fields_ = [("mesh", C.POINTER(Mesh)), ("Kncol", PI), ("Kcol", PI), ("Kval", PD), ("Kdiag", PD),
                ("interp_n", PI), ("interp_eq", PI), ("nproj", C.c_int), ("proj_l", PI), ("proj_u", PI)]


class Ctrl(C.Structure):
    _fields_ = [("Ra", C.c_double), ("A", C.c_double), ("refT", C.c_double), ("lm_jump", C.c_double),
                ("accuracy", C.c_double), ("citcom_accu", C.c_double), ("theta", C.c_double),
                ("Q", C.c_double), ("max_citcom_cycle", C.c_int), ("max_cg_cycle", C.c_int),
                ("cg_precondition", C.c_int), ("check_v_only", C.c_int), ("solver", C.c_int),
                ("auto_accuracy", C.c_int), ("schur_divide", C.c_int), ("mg_cycle", C.c_int),
                ("mg_finest_iteration", C.c_int), ("mg_coarse_iteration", C.c_int)]


class StokesStats(C.Structure):
    _fields_ = [("uzawa_cycles", C.c_int), ("dv", C.c_double), ("dp", C.c_double),
                ("inner_iterations", C.c_longlong), ("spmv", C.c_longlong), ("accuracy_used", C.c_double)]


class HeatStats(C.Structure):
    _fields_ = [("cg_cycles", C.c_int), ("init_cg_cycles", C.c_int), ("dt", C.c_double)]


def _declare(L):
    L.mclo_buoyancy.argtypes = [C.POINTER(Mesh), PD, C.c_double, PD]
    L.mclo_viscosity.argtypes = [C.POINTER(Mesh), PD, C.c_double, C.c_double, C.c_double, PD]
    L.mclo_assemble_stokes.argtypes = [C.POINTER(Mesh), PD, PD, PI, PI, PD, PD, PI, PI, PD, PD]
    L.mclo_stokes_solve.argtypes = [C.POINTER(Mesh), C.POINTER(Ctrl), PI, PI, PD, PD, PI, PI, PD, PD,
                                    PD, PD, C.POINTER(MgLevel), C.c_int, C.POINTER(StokesStats)]
    L.mclo_v_to_nodes.argtypes = [C.POINTER(Mesh), PD, PD]
    L.mclo_compute_stress.argtypes = [C.POINTER(Mesh), PD, PD, PD]
    L.mclo_advection_dt.argtypes = [C.POINTER(Mesh), PD]
    L.mclo_advection_dt.restype = C.c_double
    L.mclo_artificial_diffusivity.argtypes = [C.POINTER(Mesh), PD, C.c_double, PD]
    L.mclo_assemble_heat.argtypes = [C.POINTER(Mesh), C.POINTER(Ctrl), PD, PD, PI, PI, PD, PI, PI, PD, PD]
    L.mclo_update_temperature.argtypes = [C.POINTER(Mesh), C.POINTER(Ctrl), PD, PD, PD, PD, C.c_int, PD,
                                          C.POINTER(HeatStats), PD, PI, PI, PD, PI, PI, PD, PD]
    L.mclo_rows_prod.argtypes = [C.c_int, C.c_int, PI, PI, PD, PD, PD]


def _p(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


class MeshArrays:
    """Holds the numpy arrays of one mesh level and the ctypes struct over them."""

    def __init__(self, d):
        self.d = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in d.items()}
        g = self.d
        self.struct = Mesh(
            g["L"], g["lev"], g["nel"], g["nno"], g["neq"], g["neqT"], g["gnox"], g["gnoz"],
            _p(g["Mcode"], PI), _p(g["ien"], PI), _p(g["node"], PI), _p(g["id"], PI), _p(g["lm"], PI),
            _p(g["heat_id"], PI), _p(g["heat_lm"], PI), _p(g["v_bc"], PB), _p(g["v_bc_val"], PD),
            _p(g["T_bc"], PB), _p(g["T_bc_val"], PD), _p(g["X"], PD), _p(g["dx"], PD), _p(g["dz"], PD),
            g["nh"], _p(g["h_ni"], PI), _p(g["h_ci"], PI))

    def __getitem__(self, k):
        return self.d[k]


def load_fixture(name):
    with np.load(os.path.join(GOLDEN, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def fixture_keys(fx):
    txt = bytes(fx["input_keys"]).decode()
    out = {}
    for line in txt.splitlines():
        k, v = line.split("=", 1)
        out[k] = v.strip().strip('"')
    return out


def mesh_from_fixture(fx, lev=None):
    L = int(fx["max_level"][0])
    if lev is None:
        lev = L - 1
    p = f"L{lev}."
    nel = int(fx[p + "nel"][0])
    nno = int(fx[p + "nno"][0])
    nh = int(fx[p + "hangle.nh"][0])
    d = dict(
        L=L, lev=lev, nel=nel, nno=nno, neq=int(fx[p + "neq"][0]), neqT=int(fx[p + "heat.neq"][0]),
        gnox=int(fx["gnox"][0]), gnoz=int(fx["gnoz"][0]),
        Mcode=fx[p + "Mcode"].astype(np.int32),
        ien=np.stack([fx[p + f"ien{a}"] for a in range(4)]).astype(np.int32),
        node=fx[p + "node"].astype(np.int32),
        id=np.stack([fx[p + f"id{k}"] for k in range(2)]).astype(np.int32),
        lm=np.stack([np.stack([fx[p + f"lm{k}_{a}"] for a in range(4)]) for k in range(2)]).astype(np.int32),
        heat_id=fx[p + "heat.id"].astype(np.int32),
        heat_lm=np.stack([fx[p + f"heat.lm{a}"] for a in range(4)]).astype(np.int32),
        v_bc=np.stack([fx[p + f"v_bc{k}"] for k in range(2)]).astype(np.int8),
        v_bc_val=np.stack([fx[p + f"v_bc_val{k}"] for k in range(2)]).astype(np.float64),
        T_bc=fx[p + "T_bc"].astype(np.int8), T_bc_val=fx[p + "T_bc_val"].astype(np.float64),
        X=np.stack([fx[p + f"X{k}"] for k in range(2)]).astype(np.float64),
        dx=fx["dx"].astype(np.float64), dz=fx["dz"].astype(np.float64),
        nh=nh,
        h_ni=(fx[p + "hangle.ni"].astype(np.int32) if nh else np.zeros(1, np.int32)),
        h_ci=(np.stack([fx[p + f"hangle.ci{k}"] for k in range(2)]).astype(np.int32) if nh
              else np.zeros(2, np.int32)),
    )
    return MeshArrays(d)


def mesh_from_host_level(lv):
    """MeshArrays over one level of the host layer's mesh (Model.mesh.level(k)),
    for meshes that no fixture holds (the 256x256 benchmark mesh)."""
    nh = int(lv["nh"])
    return MeshArrays(dict(L=lv["L"], lev=lv["lev"], nel=lv["nel"], nno=lv["nno"], neq=lv["neq"],
                           neqT=lv["neqT"], gnox=lv["gnox"], gnoz=lv["gnoz"], Mcode=lv["Mcode"], ien=lv["ien"],
                           node=lv["node"], id=lv["id"], lm=lv["lm"], heat_id=lv["heat_id"],
                           heat_lm=lv["heat_lm"], v_bc=lv["v_bc"], v_bc_val=lv["v_bc_val"], T_bc=lv["T_bc"],
                           T_bc_val=lv["T_bc_val"], X=lv["X"], dx=lv["dx"], dz=lv["dz"], nh=nh,
                           h_ni=lv["h_ni"] if nh else np.zeros(1, np.int32),
                           h_ci=lv["h_ci"] if nh else np.zeros(2, np.int32)))


def ctrl_from_control(c, schur_divide=None, max_citcom_cycle=None):
    """Oracle controls from the library's mcl_control (Input.control())."""
    return Ctrl(c.Ra, c.A, c.refT, c.lm_jump, c.accuracy, c.citcom_accu, c.theta, c.Q,
                c.max_citcom_cycle if max_citcom_cycle is None else max_citcom_cycle, c.max_cg_cycle,
                c.cg_precondition, c.check_v_only, c.solver, c.auto_accuracy,
                c.schur if schur_divide is None else schur_divide, c.mg_cycle, c.mg_finest_iteration,
                c.mg_coarse_iteration)


def mg_from_fixture(fx, ctrl, T_levels):
    """The oracle's multigrid hierarchy for a fixture with solver="mg": every
    level min_level..max_level-1 assembled from its own temperature
    T_levels[lev] (the reference assembles K on every level from that level's
    T, Build_matrix.c:4-40), with the transfer maps the harness dumped
    (interp[lev+1] and project[lev] under the coarse level's name).
    Returns (levels, ctypes array of MgLevel, count, assembled systems)."""
    L = int(fx["max_level"][0])
    lo = int(fx["min_level"][0])
    keep, systems = [], {}
    arr = (MgLevel * (L - lo))()
    for i, lev in enumerate(range(lo, L)):
        m = mesh_from_fixture(fx, lev)
        o = Oracle(m, ctrl)
        T = np.ascontiguousarray(T_levels[lev], np.float64)
        A = o.assemble(o.viscosity(T), o.buoyancy(T))
        systems[lev] = A
        keep += [m, A]
        g = arr[i]
        g.mesh = C.pointer(m.struct)
        g.Kncol, g.Kcol, g.Kval, g.Kdiag = (_p(A["Kncol"], PI), _p(A["Kcol"], PI), _p(A["Kval"], PD),
                                            _p(A["Kdiag"], PD))
        if lev > lo:
            n = np.ascontiguousarray(fx[f"L{lev - 1}.interp.local_eq_n"], np.int32)
            eq = np.ascontiguousarray(fx[f"L{lev - 1}.interp.local_eq"], np.int32)
            keep += [n, eq]
            g.interp_n, g.interp_eq = _p(n, PI), _p(eq, PI)
        if lev < L - 1:
            pl = np.ascontiguousarray(fx[f"L{lev}.project.local_l"], np.int32)
            pu = np.ascontiguousarray(fx[f"L{lev}.project.local_u"], np.int32)
            keep += [pl, pu]
            g.nproj, g.proj_l, g.proj_u = pl.size, _p(pl, PI), _p(pu, PI)
    return keep, arr, L - lo, systems


def ctrl_from_keys(k, schur_divide=0):
    f = lambda n, dflt: float(k.get(n, dflt))  # noqa: E731
    i = lambda n, dflt: int(k.get(n, dflt))  # noqa: E731
    return Ctrl(f("Rayleigh", 1e7), f("A", 0.0), f("refT", 0.5), f("lower_mantle_jump", 1.0),
                f("accuracy", 1e-3), f("citcom_solver_accu", 1e-3), f("theta", 0.0), f("Q", 0.0),
                i("max_citcom_cycle", 375), i("max_cg_cycle", 200), i("cg_precondition", 0),
                i("check_v_only", 0), SOLVERS[k.get("solver", "jc")], i("auto_accuracy", 0), schur_divide,
                i("mg_cycle", 2), i("mg_finest_iteration", 4), i("mg_coarse_iteration", 4))


def csr_from_slots(ncol, col, val, W):
    n = ncol.size
    ptr = np.zeros(n + 1, np.int64)
    ptr[1:] = np.cumsum(ncol)
    mask = np.arange(W)[None, :] < ncol[:, None]
    return ptr, col.reshape(n, W)[mask], val.reshape(n, W)[mask]


class Oracle:
    """Convenience wrapper: one mesh level, FP64, reference-ordered arithmetic."""

    def __init__(self, mesh, ctrl):
        self.m = mesh
        self.c = ctrl
        self.L = lib()

    def buoyancy(self, T):
        out = np.empty(self.m["nno"])
        self.L.mclo_buoyancy(C.byref(self.m.struct), _p(np.ascontiguousarray(T, np.float64), PD),
                             self.c.Ra, _p(out, PD))
        return out

    def viscosity(self, T):
        out = np.empty(self.m["nno"])
        self.L.mclo_viscosity(C.byref(self.m.struct), _p(np.ascontiguousarray(T, np.float64), PD),
                              self.c.A, self.c.refT, self.c.lm_jump, _p(out, PD))
        return out

    def assemble(self, vis, buoy):
        n = self.m["neq"]
        r = dict(Kncol=np.zeros(n, np.int32), Kcol=np.zeros(n * VW, np.int32), Kval=np.zeros(n * VW),
                 Kdiag=np.zeros(n), Gncol=np.zeros(n, np.int32), Gcol=np.zeros(n * VW, np.int32),
                 Gval=np.zeros(n * VW), F=np.zeros(n))
        self.L.mclo_assemble_stokes(C.byref(self.m.struct), _p(vis, PD), _p(buoy, PD),
                                    _p(r["Kncol"], PI), _p(r["Kcol"], PI), _p(r["Kval"], PD), _p(r["Kdiag"], PD),
                                    _p(r["Gncol"], PI), _p(r["Gcol"], PI), _p(r["Gval"], PD), _p(r["F"], PD))
        return r

    def solve(self, A, d=None, p=None, mg=None):
        d = np.zeros(self.m["neq"]) if d is None else np.ascontiguousarray(d, np.float64).copy()
        p = np.zeros(self.m["nel"]) if p is None else np.ascontiguousarray(p, np.float64).copy()
        st = StokesStats()
        mgp, nl = (None, 0) if mg is None else mg
        self.L.mclo_stokes_solve(C.byref(self.m.struct), C.byref(self.c), _p(A["Kncol"], PI), _p(A["Kcol"], PI),
                                 _p(A["Kval"], PD), _p(A["Kdiag"], PD), _p(A["Gncol"], PI), _p(A["Gcol"], PI),
                                 _p(A["Gval"], PD), _p(A["F"], PD), _p(d, PD), _p(p, PD), mgp, nl, C.byref(st))
        return d, p, st

    def v_to_nodes(self, d):
        V = np.empty(2 * self.m["nno"])
        self.L.mclo_v_to_nodes(C.byref(self.m.struct), _p(d, PD), _p(V, PD))
        return V.reshape(2, -1)

    def stress(self, vis, V):
        out = np.empty(self.m["nno"])
        self.L.mclo_compute_stress(C.byref(self.m.struct), _p(vis, PD),
                                   _p(np.ascontiguousarray(V, np.float64), PD), _p(out, PD))
        return out

    def update_temperature(self, V, T, hd, hdd, first_call, model_time, want=False):
        n, nel = self.m["neqT"], self.m["nel"]
        T = np.ascontiguousarray(T, np.float64).copy()
        hd = np.zeros(n) if hd is None else hd.copy()
        hdd = np.zeros(n) if hdd is None else hdd.copy()
        mt = C.c_double(model_time)
        st = HeatStats()
        extra = None
        if want:
            extra = dict(tau=np.zeros(2 * nel), Kncol=np.zeros(n, np.int32), Kcol=np.zeros(n * TW, np.int32),
                         Kval=np.zeros(n * TW), Cncol=np.zeros(n, np.int32), Ccol=np.zeros(n * TW, np.int32),
                         Cval=np.zeros(n * TW), F=np.zeros(n))
        e = extra or {}
        self.L.mclo_update_temperature(C.byref(self.m.struct), C.byref(self.c),
                                       _p(np.ascontiguousarray(V, np.float64), PD), _p(T, PD), _p(hd, PD),
                                       _p(hdd, PD), int(first_call), C.byref(mt), C.byref(st),
                                       _p(e.get("tau"), PD), _p(e.get("Kncol"), PI), _p(e.get("Kcol"), PI),
                                       _p(e.get("Kval"), PD), _p(e.get("Cncol"), PI), _p(e.get("Ccol"), PI),
                                       _p(e.get("Cval"), PD), _p(e.get("F"), PD))
        return T, hd, hdd, mt.value, st, extra
