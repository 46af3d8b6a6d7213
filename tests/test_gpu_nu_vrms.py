"""Vrms and top Nusselt number of the mixed-precision path (warm-started,
FP64 CG + FP32 multigrid inner solves) against the FP64 path over ten time
steps of a C2-shaped case (64x64, Ra=1e6, 1000x viscosity contrast, seeded
initial temperature) on the MI355X.

The FP64 side is the reference's Uzawa/Jacobi-PCG arithmetic (the path that
matches the reference to 1e-10 after one step) with its inner solves run to
convergence (accuracy 1e-20), so that both paths solve the same discrete
problem to the same Uzawa tolerance (dv, dp < 1e-6).  Bar: Vrms and Nu agree
to 1e-6 relative after every step, and T to 1e-6 relative L2 at the end.
The same bar holds over the first two steps of the 256x256 benchmark
configuration.
"""
import math

import pytest

from tests.cases import rel_l2

pytestmark = pytest.mark.gpu

KEYS = dict(level=7, mg_min_level=2, nprox=1, nproz=1, model_length=1.0, model_height=1.0, Rayleigh=1e6,
            A=math.log(1000.0), refT=0.0, T_ini=4, T_seed=20261202, T_mantle_perturb=0.01, Q=0.0, theta=0.5,
            develop_mode=0, solver="cg", cg_precondition=1, accuracy=1e-8, citcom_solver_accu=1e-6,
            max_citcom_cycle=1000, max_cg_cycle=200, schur_mode=1, top_v_bc="SV", bot_v_bc="SV",
            lef_v_bc="VS", rig_v_bc="VS", write_output=0)


def _model(**over):
    from mantle_convection_lab_amd import Input, Model
    k = dict(KEYS)
    k.update(over)
    return Model(Input.from_keys(k))


def test_nu_vrms_mixed_vs_fp64():
    print("[gpu-nu-vrms] 64x64 C2-shaped, 10 steps", flush=True)
    a = _model(precision=0, mg_min_level=6, accuracy=1e-20, max_cg_cycle=100000)
    b = _model(precision=1, warm_start=1)
    worst_v = worst_n = 0.0
    for s in range(10):
        ra = a.step()
        rb = b.step()
        assert ra["converged"] and rb["converged"]
        va, na = a.diagnostics()
        vb, nb = b.diagnostics()
        ev, en = abs(vb - va) / abs(va), abs(nb - na) / abs(na)
        worst_v, worst_n = max(worst_v, ev), max(worst_n, en)
        print(f"  step {s}: fp64 uzawa={ra['uzawa']} mixed uzawa={rb['uzawa']}  Vrms {va:.10g} vs {vb:.10g} "
              f"({ev:.1e})  Nu {na:.10g} vs {nb:.10g} ({en:.1e})", flush=True)
    et = rel_l2(b.field("T"), a.field("T"))
    print(f"  worst Vrms {worst_v:.2e}, worst Nu {worst_n:.2e}, T rel L2 {et:.2e}", flush=True)
    assert worst_v < 1e-6
    assert worst_n < 1e-6
    assert et < 1e-6


def test_nu_vrms_mixed_vs_fp64_256():
    """The benchmark's own configuration (256x256, Ra=1e6, 1000x contrast,
    seeded perturbation) over its first two time steps: the warm-started mixed
    path against the FP64 path with converged inner solves (each FP64 step
    takes about 10 s)."""
    import bench
    from mantle_convection_lab_amd import Input, Model
    print("[gpu-nu-vrms] 256x256 benchmark configuration, 2 steps", flush=True)
    ka = bench.path_keys("fp64")
    ka.update(accuracy=1e-20, max_cg_cycle=100000)
    a = Model(Input.from_keys(ka))
    b = Model(Input.from_keys(bench.path_keys("mixed")))
    for s in range(2):
        ra = a.step()
        rb = b.step()
        assert ra["converged"] and rb["converged"]
        va, na = a.diagnostics()
        vb, nb = b.diagnostics()
        ev, en = abs(vb - va) / abs(va), abs(nb - na) / abs(na)
        print(f"  step {s}: fp64 uzawa={ra['uzawa']} mixed uzawa={rb['uzawa']}  Vrms {va:.12g} vs {vb:.12g} "
              f"({ev:.1e})  Nu {na:.12g} vs {nb:.12g} ({en:.1e})", flush=True)
        assert ev < 1e-6
        assert en < 1e-6
    et = rel_l2(b.field("T"), a.field("T"))
    print(f"  T rel L2 {et:.2e}", flush=True)
    assert et < 1e-6
    a.ctx.close()
    b.ctx.close()
