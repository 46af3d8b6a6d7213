"""Python mirror of the reference driver around the HIP hot path.

``Input`` reads the reference's input format (through the C host layer),
``Mesh`` holds the quadtree mesh arrays of every level, ``Context`` owns one
libmcl_hip context (device-resident fields), and ``Model`` runs the per-step
sequence of the reference's main loop (Main.c:28-69):

    buoyancy_force, viscosity, velocity_pressure_KGF_matrix,
    solve_velocity_pressure, [output], update_temperature
"""
import ctypes as C
import os
import time

import numpy as np

from . import _native as N


def _err(buf):
    return buf.value.decode(errors="replace")


class MclError(RuntimeError):
    pass


class Input:
    """The reference's run-time keys (Input.c:3-98) plus this build's keys."""

    def __init__(self, struct):
        self.s = struct

    @classmethod
    def from_file(cls, path):
        L = N.host_lib()
        s = N.HostInput()
        err = C.create_string_buffer(512)
        if L.mclh_read_input(os.fsencode(path), C.byref(s), err, 512):
            raise MclError(_err(err))
        return cls(s)

    @classmethod
    def from_keys(cls, keys):
        L = N.host_lib()
        s = N.HostInput()
        L.mclh_default_input(C.byref(s))
        for k, v in keys.items():
            if isinstance(v, str):
                v = v.strip().strip('"').strip("'")
            if L.mclh_set_key(C.byref(s), str(k).encode(), str(v).encode()):
                raise MclError(f"unknown key {k}")
        if s.min_level >= s.max_level:
            raise MclError("min_level must be smaller than max_level")
        return cls(s)

    def __getattr__(self, k):
        v = getattr(self.s, k)
        return v.decode() if isinstance(v, bytes) else v

    def control(self):
        c = N.Control()
        N.host_lib().mclh_fill_control(C.byref(self.s), C.byref(c))
        return c


def _arr(ptr, n, dtype):
    if n <= 0 or not ptr:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Mesh:
    """Quadtree mesh of every level (finest last), built by the C host layer."""

    def __init__(self, inp, depth=None, _ptr=None):
        """depth=None: the input's develop_mode recipe; depth=k: uniform
        leaves at quadtree depth k (the base mesh of an adaptive run)."""
        self.inp = inp
        L = N.host_lib()
        if _ptr is not None:
            self.ptr = _ptr
        else:
            self.ptr = C.POINTER(N.HostMesh)()
            err = C.create_string_buffer(512)
            rc = (L.mclh_build_mesh(C.byref(inp.s), C.byref(self.ptr), err, 512) if depth is None else
                  L.mclh_build_mesh_depth(C.byref(inp.s), int(depth), C.byref(self.ptr), err, 512))
            if rc:
                raise MclError(_err(err))
        m = self.ptr.contents
        self.L, self.min_level, self.nlev = m.L, m.min_level, m.nlev
        self.gnox, self.gnoz = m.gnox, m.gnoz
        self.dx = np.array(m.dx[: m.L])
        self.dz = np.array(m.dz[: m.L])

    def __del__(self):
        if getattr(self, "ptr", None):
            N.host_lib().mclh_free_mesh(self.ptr)
            self.ptr = None

    def adapt(self, T, min_depth, max_depth, refine_thr=0.05, coarsen_thr=0.01):
        """New mesh refined where |grad T| h > refine_thr and coarsened where
        it is < coarsen_thr, 2:1 balanced; returns (mesh, T on its finest
        nodes, refined leaves, coarsened families)."""
        L = N.host_lib()
        T = np.ascontiguousarray(T, np.float64)
        ptr = C.POINTER(N.HostMesh)()
        tn = N.PD()
        nr, nc = C.c_int(), C.c_int()
        err = C.create_string_buffer(512)
        if L.mclh_adapt_mesh(C.byref(self.inp.s), self.ptr, T.ctypes.data_as(N.PD), int(min_depth), int(max_depth),
                             float(refine_thr), float(coarsen_thr), C.byref(ptr), C.byref(tn), C.byref(nr),
                             C.byref(nc), err, 512):
            raise MclError(_err(err))
        new = Mesh(self.inp, _ptr=ptr)
        nno = new.level_struct(new.nlev - 1).nno
        Tn = np.ctypeslib.as_array(tn, shape=(nno,)).copy()
        L.mclh_free_array(C.cast(tn, C.c_void_p))
        return new, Tn, nr.value, nc.value

    def level_struct(self, k):
        return self.ptr.contents.lev[k]

    def level(self, k=-1):
        """numpy copies of one level's arrays in the reference layouts."""
        if k < 0:
            k += self.nlev
        l = self.level_struct(k)
        nel, nno, neq, neqT, nh = l.nel, l.nno, l.neq, l.neqT, l.nh
        d = dict(L=l.L, lev=l.lev, nel=nel, nno=nno, neq=neq, neqT=neqT, gnox=l.gnox, gnoz=l.gnoz, nh=nh,
                 Mcode=_arr(l.Mcode, nel, np.int32), ien=_arr(l.ien, 4 * nel, np.int32).reshape(4, nel),
                 node=_arr(l.node, nno, np.int32), id=_arr(l.id, 2 * nno, np.int32).reshape(2, nno),
                 lm=_arr(l.lm, 8 * nel, np.int32).reshape(2, 4, nel), heat_id=_arr(l.heat_id, nno, np.int32),
                 heat_lm=_arr(l.heat_lm, 4 * nel, np.int32).reshape(4, nel),
                 veq_node=_arr(l.veq_node, neq, np.int32), veq_dof=_arr(l.veq_dof, neq, np.int32),
                 heat_eq_node=_arr(l.heat_eq_node, neqT, np.int32),
                 v_bc=_arr(l.v_bc, 2 * nno, np.int8).reshape(2, nno),
                 v_bc_val=_arr(l.v_bc_val, 2 * nno, np.float64).reshape(2, nno),
                 T_bc=_arr(l.T_bc, nno, np.int8), T_bc_val=_arr(l.T_bc_val, nno, np.float64),
                 X=_arr(l.X, 2 * nno, np.float64).reshape(2, nno), area=_arr(l.area, nel, np.float64),
                 p2n_nel=_arr(l.p2n_nel, nno, np.int32), p2n_el=_arr(l.p2n_el, 4 * nno, np.int32).reshape(4, nno),
                 h_node=_arr(l.h_node, nh, np.int32), h_ni=_arr(l.h_ni, nh, np.int32),
                 h_cnode=_arr(l.h_cnode, 2 * nh, np.int32).reshape(2, nh),
                 h_ci=_arr(l.h_ci, 2 * nh, np.int32).reshape(2, nh),
                 dx=self.dx.copy(), dz=self.dz.copy())
        if l.interp_n:
            d["interp_n"] = _arr(l.interp_n, neq, np.int32)
            d["interp_eq"] = _arr(l.interp_eq, 4 * neq, np.int32).reshape(4, neq)
            d["proj_l"] = _arr(l.proj_l, l.nproj, np.int32)
            d["proj_u"] = _arr(l.proj_u, l.nproj, np.int32)
        return d

    def views(self):
        arr = (N.MeshView * self.nlev)()
        for k in range(self.nlev):
            N.host_lib().mclh_fill_view(self.ptr, k, C.byref(arr[k]))
        return arr

    def initial_temperature(self, k=-1):
        if k < 0:
            k += self.nlev
        T = np.zeros(self.level_struct(k).nno)
        err = C.create_string_buffer(512)
        if N.host_lib().mclh_initial_temperature(C.byref(self.inp.s), self.ptr, k,
                                                 T.ctypes.data_as(N.PD), err, 512):
            raise MclError(_err(err))
        return T

    def diagnostics(self, T, V):
        vr, nu = C.c_double(), C.c_double()
        N.host_lib().mclh_diagnostics(C.byref(self.inp.s), C.byref(self.level_struct(self.nlev - 1)),
                                      np.ascontiguousarray(T, np.float64).ctypes.data_as(N.PD),
                                      np.ascontiguousarray(V, np.float64).ctypes.data_as(N.PD),
                                      C.byref(vr), C.byref(nu))
        return vr.value, nu.value


class Context:
    """One libmcl_hip context: device-resident mesh and fields."""

    def __init__(self, device=0):
        self.L = N.hip_lib()
        self.h = C.c_void_p()
        rc = self.L.mcl_create(device, C.byref(self.h))
        if rc:
            raise MclError(f"mcl_create failed ({rc}): no usable GPU")
        self.levels = {}

    def close(self):
        if self.h:
            self.L.mcl_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        if rc:
            raise MclError(f"mcl error {rc}: {self.L.mcl_last_error(self.h).decode(errors='replace')}")

    def upload(self, views, sizes):
        """views: ctypes array of MeshView; sizes: {lev: dict(nel, nno, neq, neqT)}."""
        self._keep = views
        self._ck(self.L.mcl_mesh_upload(self.h, views, len(views)))
        self.levels = sizes

    def set_control(self, ctrl):
        self.ctrl = ctrl
        self._ck(self.L.mcl_set_control(self.h, C.byref(ctrl)))

    _sizes = {0: "nno", 1: "nno", 2: "nno", 3: "2nno", 4: "neq", 5: "nel", 6: "neq", 7: "neq", 8: "nno",
              9: "neqT", 10: "neqT", 11: "neqT", 12: "2nel"}

    def _count(self, lev, field):
        s = self._sizes[field]
        sz = self.levels[lev]
        return 2 * sz[s[1:]] if s.startswith("2") else sz[s]

    def set_field(self, lev, field, values):
        f = N.FIELD[field] if isinstance(field, str) else field
        a = np.ascontiguousarray(values, np.float64).ravel()
        self._ck(self.L.mcl_set_field(self.h, lev, f, a.ctypes.data_as(N.PD), a.size))

    def get_field(self, lev, field):
        f = N.FIELD[field] if isinstance(field, str) else field
        n = self._count(lev, f)
        a = np.empty(n)
        self._ck(self.L.mcl_get_field(self.h, lev, f, a.ctypes.data_as(N.PD), n))
        return a

    def get_matrix(self, lev, which, width=36):
        n = self.levels[lev]["neqT" if which in ("HEAT_K", "HEAT_C") else "neq"]
        ncol = np.zeros(n, np.int32)
        col = np.zeros(n * width, np.int32)
        val = np.zeros(n * width)
        self._ck(self.L.mcl_get_matrix(self.h, lev, N.MATRIX[which], width, ncol.ctypes.data_as(N.PI),
                                       col.ctypes.data_as(N.PI), val.ctypes.data_as(N.PD)))
        return ncol, col, val

    def set_model_time(self, t):
        self._ck(self.L.mcl_set_model_time(self.h, float(t)))

    def model_time(self):
        t = C.c_double()
        self._ck(self.L.mcl_get_model_time(self.h, C.byref(t)))
        return t.value

    def buoyancy_force(self):
        self._ck(self.L.mcl_buoyancy_force(self.h))

    def viscosity(self):
        self._ck(self.L.mcl_viscosity(self.h))

    def velocity_pressure_KGF_matrix(self):
        self._ck(self.L.mcl_velocity_pressure_KGF_matrix(self.h))

    def solve_velocity_pressure(self):
        st = N.StokesStats()
        self._ck(self.L.mcl_solve_velocity_pressure(self.h, C.byref(st)))
        return st

    def solve_velocity_pressure_noalloc(self):
        """The allocation-free form (work space in the library's static
        device scratch array; FP64 precision only)."""
        st = N.StokesStats()
        self._ck(self.L.mcl_solve_velocity_pressure_noalloc(self.h, C.byref(st)))
        return st

    def update_temperature(self):
        st = N.HeatStats()
        self._ck(self.L.mcl_update_temperature(self.h, C.byref(st)))
        return st

    def synchronize(self):
        self._ck(self.L.mcl_synchronize(self.h))


class Model:
    """The reference's main loop (Main.c:28-69) over the device hot path."""

    def __init__(self, inp, device=0, T0=None, ctrl_overrides=None, net=None, mesh=None):
        """net: optional learned-guess network (learn.Net) on the finest
        grid; before every solve it proposes (d, p) and the candidate with the
        smaller momentum residual (network or current fields) starts the
        Uzawa iteration, which always runs to the same tolerance."""
        self.inp = inp
        self.net = net
        # learned-guess gate: consecutive steps on which the previous solution
        # beat the network, and steps since the network was last consulted
        self._net_losses = 0
        self._net_idle = 0
        self.device = device
        self.ctrl_overrides = ctrl_overrides
        self.mesh = mesh if mesh is not None else Mesh(inp)
        self._setup(T0)
        self.timestep = 0
        self.history = []

    def _setup(self, T0, model_time=0.0):
        inp = self.inp
        self.ctx = Context(self.device)
        sizes = {}
        for k in range(self.mesh.nlev):
            l = self.mesh.level_struct(k)
            sizes[l.lev] = dict(nel=l.nel, nno=l.nno, neq=l.neq, neqT=l.neqT)
        self.ctx.upload(self.mesh.views(), sizes)
        ctrl = inp.control()
        for k, v in (self.ctrl_overrides or {}).items():
            setattr(ctrl, k, v)
        self.ctx.set_control(ctrl)
        self.finest = self.mesh.level_struct(self.mesh.nlev - 1).lev
        # initial_temperature (Main.c:55-56) on every level
        for k in range(self.mesh.nlev):
            lev = self.mesh.level_struct(k).lev
            T = self.mesh.initial_temperature(k) if (T0 is None or lev != self.finest) else T0
            self.ctx.set_field(lev, "T", T)
        self.ctx.set_model_time(model_time)

    def remesh(self, min_depth, max_depth, refine_thr=0.05, coarsen_thr=0.01):
        """Adapt the quadtree to the current temperature (the mesh, the
        device context and its fields are rebuilt; T is transferred, the
        solution restarts from zero; model time is kept).  When the element
        list comes out unchanged, the current mesh and context are kept, with
        their solution as the next warm start."""
        T = self.field("T")
        t = self.ctx.model_time()
        new, Tn, nr, nc = self.mesh.adapt(T, min_depth, max_depth, refine_thr, coarsen_thr)
        if np.array_equal(new.level(-1)["Mcode"], self.mesh.level(-1)["Mcode"]):
            # the same element list (e.g. families coarsened and then refined
            # again by the 2:1 balance): keep the mesh, context and warm start
            return 0, 0
        self.ctx.close()
        self.mesh = new
        self._setup(Tn, model_time=t)
        return nr, nc

    def step(self):
        c = self.ctx
        t0 = time.perf_counter()
        c.buoyancy_force()
        c.viscosity()
        c.velocity_pressure_KGF_matrix()
        guess = None
        if self.net is not None and self._consult_net():
            guess = self.net.guess(c)
            self._net_losses = 0 if guess[0] else self._net_losses + 1
        st = c.solve_velocity_pressure()
        if self.inp.write_output and self.inp.save_spacing > 0 and self.timestep % self.inp.save_spacing == 0:
            self.write_output()
        hs = c.update_temperature()
        rec = dict(step=self.timestep, uzawa=st.uzawa_cycles, converged=st.converged, dv=st.dv, dp=st.dp,
                   inner=st.inner_iterations, solve_ms=st.device_ms, heat_cg=hs.cg_cycles, dt=hs.dt,
                   model_time=hs.model_time, res_mom=st.res_momentum, res_cont=st.res_continuity,
                   wall_s=time.perf_counter() - t0)
        if guess is not None:
            rec.update(guess_accepted=guess[0], guess_res_net=guess[1], guess_res_prev=guess[2])
        self.history.append(rec)
        self.timestep += 1
        return rec

    # after this many consecutive losses the network is consulted only every
    # NET_RECHECK steps (a consultation costs a network pass and two probe
    # solves; once the flow settles the previous step's solution wins)
    NET_LOSSES = 2
    NET_RECHECK = 8

    def _consult_net(self):
        if self._net_losses < self.NET_LOSSES:
            return True
        self._net_idle += 1
        if self._net_idle >= self.NET_RECHECK:
            self._net_idle = 0
            return True
        return False

    def field(self, name):
        return self.ctx.get_field(self.finest, name)

    def write_output(self):
        l = self.mesh.level_struct(self.mesh.nlev - 1)
        T, S, Vis, V = (self.field("T"), self.field("STRESS"), self.field("VIS"), self.field("V"))
        err = C.create_string_buffer(512)
        if N.host_lib().mclh_write_output(C.byref(self.inp.s), C.byref(l), self.timestep,
                                          *(a.ctypes.data_as(N.PD) for a in (T, S, Vis, V)), err, 512):
            raise MclError(_err(err))

    def diagnostics(self):
        return self.mesh.diagnostics(self.field("T"), self.field("V"))
