"""Adaptive run on the MI355X: an aspect-2 box (base 32 x 16) at Ra=1e6 with
a large smooth perturbation T = (1-z) + 0.5 cos(pi x) sin(pi z), so that
|grad T| h crosses the refinement threshold in part of the box only, remeshed
every 2 steps with up to 2 extra levels.
After the first remesh with hanging nodes, the mixed-precision solve (inner
CG on the Galerkin-constrained space) must still converge to dv, dp < 1e-6
in every step, the temperature must stay within the initial field's range up to 1e-2
(the transfer interpolates, SUPG is not monotone), and the
final velocities must satisfy the hanging-node constraint exactly.  A remesh
that changes nothing keeps the context and its warm start.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_amr_run_converges():
    from mantle_convection_lab_amd import amr
    print("[gpu-amr] base 32x16, 8 steps, remesh every 2", flush=True)
    logs = []
    out = amr.run(steps=8, every=2, base_nx=32, extra=2, refine_thr=0.08, coarsen_thr=0.005, Ra=1e6,
                  contrast=1e3, T_ini=0, T_mantle_perturb=0.5,
                  log=lambda s: (logs.append(s), print(s, flush=True)))
    rs = out["recs"]
    assert all(r["converged"] for r in rs)
    assert max(r["nh"] for r in rs) > 0, "no hanging nodes were created"
    assert np.isfinite(out["vrms"]) and out["vrms"] > 0


def test_amr_constraint_holds():
    from mantle_convection_lab_amd import amr
    from mantle_convection_lab_amd.model import Input, Mesh, Model
    keys, base = amr.c3_keys(32, 2, Ra=1e6, contrast=1e3, T_ini=0, T_mantle_perturb=0.5)
    inp = Input.from_keys(keys)
    m = Model(inp, mesh=Mesh(inp, depth=base))
    T0 = m.field("T")
    lo, hi = T0.min(), T0.max()
    m.step()
    m.remesh(base, base + 2, 0.08, 0.005)
    r = m.step()
    lv = m.mesh.level(-1)
    assert lv["nh"] > 0 and r["converged"]
    T = m.field("T")
    assert T.min() > lo - 1e-2 and T.max() < hi + 1e-2
    # hanging velocities are the means of their edge end points
    V = m.field("V").reshape(2, -1)
    for k in range(lv["nh"]):
        h, c0, c1 = lv["h_ni"][k], lv["h_ci"][0][k], lv["h_ci"][1][k]
        for d in range(2):
            assert abs(V[d][h] - 0.5 * (V[d][c0] + V[d][c1])) <= 1e-12 * (np.abs(V).max() + 1e-300)


def test_unchanged_remesh_keeps_context(monkeypatch):
    """A remesh that refines and coarsens nothing keeps the mesh, the device
    context and the solution, so the next solve is still warm-started (the
    mixed solve's augmented Lagrangian is switched off: it lowers the cold
    first solve's Uzawa count, the yardstick of the warm start here)."""
    monkeypatch.setenv("MCL_AL_GAMMA", "0")
    from mantle_convection_lab_amd import amr
    from mantle_convection_lab_amd.model import Input, Mesh, Model
    keys, base = amr.c3_keys(32, 2, Ra=1e5, contrast=10.0)
    inp = Input.from_keys(keys)
    m = Model(inp, mesh=Mesh(inp, depth=base))
    r0 = m.step()
    r1 = m.step()
    ctx, mesh = m.ctx, m.mesh
    D = m.field("D")
    assert m.remesh(base, base + 2, 1e30, 0.0) == (0, 0)
    assert m.ctx is ctx and m.mesh is mesh
    assert np.array_equal(m.field("D"), D)
    r2 = m.step()
    print(f"[gpu-amr] uzawa {r0['uzawa']}, {r1['uzawa']}, after the unchanged remesh {r2['uzawa']}", flush=True)
    assert r2["converged"] and r2["uzawa"] < r0["uzawa"]


def test_c3_full_scale_with_hanging_nodes():
    """The adaptive workload at its own size: aspect-2 box, base 128 x 64,
    three extra levels, Ra = 1e7, viscosity contrast 1e4, mixed precision,
    started from a developed temperature (boundary layers and 32 plumes, on
    the base mesh adapted to it: about 0.31M velocity equations and 30k
    hanging nodes) and remeshed once more after 10 steps.  Every solve must
    converge with the momentum residual within 1e-6, the mesh must stay at
    0.3-0.6M velocity equations with hanging nodes, and the final velocity
    must satisfy the hanging-node constraint."""
    from mantle_convection_lab_amd import amr
    print("[gpu-amr] C3 full scale: base 128x64 +3 levels, Ra=1e7, contrast 1e4, 12 steps", flush=True)
    out = amr.run(steps=12, every=10, base_nx=128, extra=3, developed=True, log=lambda s: print(s, flush=True),
                  keep_model=True)
    rs = out["recs"]
    assert all(r["converged"] and r["res_mom"] <= 1e-6 for r in rs)
    assert all(3e5 <= r["neq"] <= 6e5 and r["nh"] > 1000 for r in rs)
    m = out["model"]
    lv = m.mesh.level(-1)
    V = m.field("V").reshape(2, -1)
    h, c0, c1 = lv["h_ni"], lv["h_ci"][0], lv["h_ci"][1]
    gap = np.abs(V[:, h] - 0.5 * (V[:, c0] + V[:, c1])).max()
    print(f"  {lv['neq']} velocity equations, {lv['nh']} hanging nodes, constraint gap {gap:.2e}", flush=True)
    assert gap <= 1e-12 * np.abs(V).max()
    m.ctx.close()
