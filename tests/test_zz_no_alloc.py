"""mcl_solve_velocity_pressure_noalloc (work space in the library's static
device array) against mcl_solve_velocity_pressure (work space allocated with
the context), through the C-ABI on the MI355X.

Two calls in a row, the second on a smaller mesh, so that the second call
starts from the first call's leftovers in the shared array: both must give
exactly what the allocating entry point gives on an identical fresh context
(same kernels, same launch shape, same summation order: d, p, nodal velocity
and stress bit for bit, the same Uzawa and inner iteration counts).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = dict(nprox=1, nproz=1, model_length=1.0, model_height=1.0, Rayleigh=1e6, A=math.log(1000.0), refT=0.0,
            T_ini=4, T_seed=20261202, T_mantle_perturb=0.01, Q=0.0, theta=0.5, develop_mode=0, solver="cg",
            cg_precondition=1, accuracy=1e-8, citcom_solver_accu=1e-6, max_citcom_cycle=2000, max_cg_cycle=200,
            schur_mode=1, top_v_bc="SV", bot_v_bc="SV", lef_v_bc="VS", rig_v_bc="VS", write_output=0)


def _prepared(side):
    from mantle_convection_lab_amd import Input, Model
    level = int(math.log2(side)) + 1
    m = Model(Input.from_keys(dict(KEYS, level=level, mg_min_level=level - 1)))
    m.ctx.buoyancy_force()
    m.ctx.viscosity()
    m.ctx.velocity_pressure_KGF_matrix()
    return m


def _fields(m):
    return {f: m.field(f) for f in ("D", "P", "V", "STRESS")}


@pytest.mark.parametrize("sides", [(64, 32)])
def test_noalloc_solve_matches_allocating_solve(sides):
    for side in sides:
        print(f"[no-alloc] side {side}: allocating entry point", flush=True)
        a = _prepared(side)
        sa = a.ctx.solve_velocity_pressure()
        print(f"[no-alloc] side {side}: allocation-free entry point", flush=True)
        b = _prepared(side)
        sb = b.ctx.solve_velocity_pressure_noalloc()
        print(f"  uzawa {sa.uzawa_cycles} / {sb.uzawa_cycles}, inner {sa.inner_iterations} / {sb.inner_iterations}, "
              f"device ms {sa.device_ms:.2f} / {sb.device_ms:.2f}", flush=True)
        assert sa.converged and sb.converged
        assert (sa.uzawa_cycles, sa.inner_iterations, sa.spmv) == (sb.uzawa_cycles, sb.inner_iterations, sb.spmv)
        assert sa.dv == sb.dv and sa.dp == sb.dp
        assert sa.res_momentum == sb.res_momentum and sa.res_continuity == sb.res_continuity
        assert sb.device_ms > 0
        fa, fb = _fields(a), _fields(b)
        for f in fa:
            assert np.array_equal(fa[f], fb[f]), f
        a.ctx.close()
        b.ctx.close()


def test_noalloc_rejects_mixed_precision():
    from mantle_convection_lab_amd import Input, Model
    from mantle_convection_lab_amd.model import MclError
    print("[no-alloc] precision=1 is rejected", flush=True)
    m = Model(Input.from_keys(dict(KEYS, level=6, mg_min_level=2, precision=1, warm_start=1)))
    m.ctx.buoyancy_force()
    m.ctx.viscosity()
    m.ctx.velocity_pressure_KGF_matrix()
    with pytest.raises(MclError, match="FP64 Uzawa kernel only"):
        m.ctx.solve_velocity_pressure_noalloc()
    m.ctx.close()
