"""The oracle (oracle/mclo.c) against the reference's own golden dumps.

The oracle restates the reference's arithmetic in order; built with the
reference's flags it must reproduce every dumped array bit for bit: buoyancy,
viscosity, K/G/F/Kdiag, the Uzawa solution and cycle count, nodal velocity and
stress, SUPG tau, heat K/C/F, dt, model time and the new temperature, for
uniform, variable-viscosity, plain-CG and hanging-node meshes.
"""
import numpy as np
import pytest

from tests.cases import STEP_CASES
from tests.oracle_lib import (TW, VW, Oracle, csr_from_slots, ctrl_from_keys, fixture_keys, load_fixture,
                              mesh_from_fixture)


def _oracle(name):
    fx = load_fixture(name)
    m = mesh_from_fixture(fx)
    return fx, m, Oracle(m, ctrl_from_keys(fixture_keys(fx)))


@pytest.mark.parametrize("name", STEP_CASES)
def test_oracle_first_step_bitwise(name):
    fx, m, o = _oracle(name)
    lev = m["lev"]
    T0 = fx[f"s0.L{lev}.T.init"]
    b = o.buoyancy(T0)
    v = o.viscosity(T0)
    assert np.array_equal(b, fx[f"s0.L{lev}.buoyancy"])
    assert np.array_equal(v, fx[f"s0.L{lev}.vis"])
    A = o.assemble(v, b)
    for mat in ("K", "G"):
        ptr, col, val = csr_from_slots(A[mat + "ncol"], A[mat + "col"], A[mat + "val"], VW)
        assert np.array_equal(ptr, fx[f"s0.L{lev}.{mat}.ptr"])
        assert np.array_equal(col, fx[f"s0.L{lev}.{mat}.col"])
        assert np.array_equal(val, fx[f"s0.L{lev}.{mat}.val"])
    assert np.array_equal(A["F"], fx[f"s0.L{lev}.F"])
    assert np.array_equal(A["Kdiag"], fx[f"s0.L{lev}.Kdiag"])
    d, p, st = o.solve(A)
    assert st.uzawa_cycles == fx["stdout.uzawa_cycles"][0]
    assert np.array_equal(d, fx["s0.d"]) and np.array_equal(p, fx["s0.p"])
    V = o.v_to_nodes(d)
    assert np.array_equal(V[0], fx["s0.V0"]) and np.array_equal(V[1], fx["s0.V1"])
    assert np.array_equal(o.stress(v, V), fx["s0.stress"])
    T, hd, hdd, mt, hs, ex = o.update_temperature(V, T0, None, None, True, 0.0, want=True)
    assert hs.init_cg_cycles == fx["stdout.heat_cg_cycles"][0]
    assert hs.cg_cycles == fx["stdout.heat_cg_cycles"][1]
    assert np.array_equal(T, fx["s0.T"])
    assert hs.dt == fx["s0.dt"][0] and mt == fx["s0.model_time"][0]
    assert np.array_equal(ex["tau"], np.concatenate([fx["s0.heat.tau0"], fx["s0.heat.tau1"]]))
    for mat in ("K", "C"):
        _, col, val = csr_from_slots(ex[mat + "ncol"], ex[mat + "col"], ex[mat + "val"], TW)
        assert np.array_equal(col, fx[f"s0.heat.{mat}.col"])
        assert np.array_equal(val, fx[f"s0.heat.{mat}.val"])
    assert np.array_equal(ex["F"], fx["s0.heat.F"])
    assert np.array_equal(hd, fx["s0.heat.d"]) and np.array_equal(hdd, fx["s0.heat.ddot"])


@pytest.mark.parametrize("name", ["c1_64", "c2s_32", "plain_cg_16"])
def test_oracle_second_step_bitwise(name):
    fx, m, o = _oracle(name)
    lev = m["lev"]
    T = fx[f"s0.L{lev}.T.init"]
    hd = hdd = None
    mt = 0.0
    for s in range(2):
        b = o.buoyancy(T)
        v = o.viscosity(T)
        d, p, st = o.solve(o.assemble(v, b))
        assert st.uzawa_cycles == fx["stdout.uzawa_cycles"][s]
        T, hd, hdd, mt, hs, _ = o.update_temperature(o.v_to_nodes(d), T, hd, hdd, s == 0, mt)
        assert np.array_equal(d, fx[f"s{s}.d"])
        assert np.array_equal(T, fx[f"s{s}.T"])
        assert mt == fx[f"s{s}.model_time"][0]


def test_oracle_catches_planted_mistakes():
    """A plausible slip in the restatement (a transposed element block, a
    dropped boundary term, a wrong sign) must fail the golden comparison."""
    fx, m, o = _oracle("c2s_32")
    lev = m["lev"]
    T0 = fx[f"s0.L{lev}.T.init"]
    b = o.buoyancy(T0)
    v = o.viscosity(T0)
    A = o.assemble(v, b)
    ptr, col, val = csr_from_slots(A["Kncol"], A["Kcol"], A["Kval"], VW)
    gold = fx[f"s0.L{lev}.K.val"]
    assert np.array_equal(val, gold)
    # viscosity sampled at the wrong Gauss point order changes K
    v2 = v.copy()
    v2[1:] = v[:-1]
    A2 = o.assemble(v2, b)
    assert not np.array_equal(csr_from_slots(A2["Kncol"], A2["Kcol"], A2["Kval"], VW)[2], gold)
    # a sign slip in the buoyancy load
    A3 = o.assemble(v, -b)
    assert not np.array_equal(A3["F"], fx[f"s0.L{lev}.F"])
