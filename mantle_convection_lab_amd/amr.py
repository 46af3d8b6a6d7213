"""Adaptive quadtree runs (the reference's adaptive_mesh_refine, Tree.c:95-99,
is empty; this is new): every `every` steps the mesh is refined where
|grad T| h > refine_thr (up to `extra` levels below the base depth), coarsened
where it is < coarsen_thr, 2:1 balanced, T is transferred, and the device
context is rebuilt.  Solves use the mixed-precision path, whose inner solves
work on the Galerkin-constrained space of the hanging nodes.

    python -m mantle_convection_lab_amd.amr [--steps 200] [--every 10] [--base-nx 128] [--extra 3]
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from .model import Input, Mesh, Model


def c3_keys(base_nx=128, extra=3, Ra=1e7, contrast=1e4, seed=20261202, **over):
    """Aspect-ratio-2 box of two square blocks, base mesh base_nx x base_nx/2."""
    base_depth = int(round(math.log2(base_nx // 2)))
    k = dict(level=base_depth + extra + 1, mg_min_level=2, nprox=2, nproz=1, model_length=2.0, model_height=1.0,
             Rayleigh=Ra, A=math.log(contrast), refT=0.0, T_ini=4, T_seed=seed, T_mantle_perturb=0.01, Q=0.0,
             theta=0.5, develop_mode=0, solver="cg", cg_precondition=1, accuracy=1e-8, citcom_solver_accu=1e-6,
             max_citcom_cycle=2000, max_cg_cycle=200, schur_mode=1, precision=1, warm_start=1,
             top_v_bc="SV", bot_v_bc="SV", lef_v_bc="VS", rig_v_bc="VS", write_output=0)
    k.update(over)
    return k, base_depth


def developed_temperature(X, height=1.0, length=2.0, delta=0.01, nplumes=32, width=0.01, amp=0.3):
    """A synthetic developed convective temperature on nodes X ([2][nno]):
    thermal boundary layers of thickness delta at the bottom (T = 1) and top
    (T = 0) around an interior at 1/2, and nplumes alternating hot and cold
    plumes of half-width `width` evenly spaced across the box,
        T = 1/2 + (erfc(z / delta) - erfc((H - z) / delta)) / 2
              + sum_k (-1)^k amp exp(-((x - x_k) / width)^2) sin(pi z / H),
    clipped to [0, 1].  Its gradients reach tens at the walls and plume
    edges, so the refinement criterion |grad T| h > 0.05 fires down to the
    finest allowed level over a large part of a 128 x 64 base mesh (about
    0.3M velocity equations with three extra levels), which a conductive
    start does not do for a long time."""
    from scipy.special import erfc
    x, z = X[0], X[1]
    T = 0.5 + 0.5 * (erfc(z / delta) - erfc((height - z) / delta))
    for k in range(nplumes):
        xk = (k + 0.5) * length / nplumes
        T = T + (1 if k % 2 == 0 else -1) * amp * np.exp(-((x - xk) / width) ** 2) * np.sin(np.pi * z / height)
    T = np.clip(T, 0.0, 1.0)
    T[np.isclose(z, 0.0)] = 1.0
    T[np.isclose(z, height)] = 0.0
    return T


def developed_mesh(inp, base, extra, refine_thr=0.05, coarsen_thr=0.01, log=None):
    """The base mesh adapted to developed_temperature() until nothing
    changes (each adaptation refines a leaf by one level at most); returns
    (mesh, T on its finest nodes)."""
    length = float(inp.length)
    mesh = Mesh(inp, depth=base)
    for _ in range(extra + 1):
        T = developed_temperature(mesh.level(-1)["X"], length=length)
        new, _, nr, nc = mesh.adapt(T, base, base + extra, refine_thr, coarsen_thr)
        if nr == 0 and nc == 0:
            break
        mesh = new
        if log:
            lv = mesh.level_struct(mesh.nlev - 1)
            log(f"[amr] initial adaptation: refined {nr} coarsened {nc} -> {lv.nel} elements, {lv.neq} velocity "
                f"dofs, {lv.nh} hanging nodes")
    return mesh, developed_temperature(mesh.level(-1)["X"], length=length)


def run(steps=200, every=10, base_nx=128, extra=3, refine_thr=0.05, coarsen_thr=0.01, log=print,
        developed=False, remesh_first=False, keep_model=False, **over):
    """developed: start from developed_temperature() on the base mesh adapted
    to it (developed_mesh) instead of the input's initial temperature;
    remesh_first: adapt the mesh before the first step too."""
    keys, base = c3_keys(base_nx, extra, **over)
    inp = Input.from_keys(keys)
    t0 = time.perf_counter()
    if developed:
        mesh, T0 = developed_mesh(inp, base, extra, refine_thr, coarsen_thr, log)
    else:
        mesh, T0 = Mesh(inp, depth=base), None
    model = Model(inp, mesh=mesh, T0=T0)
    recs = []
    for s in range(steps):
        if (s > 0 or remesh_first) and every > 0 and s % every == 0:
            nr, nc = model.remesh(base, base + extra, refine_thr, coarsen_thr)
            lv = model.mesh.level_struct(model.mesh.nlev - 1)
            log(f"[amr] step {s}: remesh refined {nr} coarsened {nc} -> {lv.nel} elements, {lv.neq} velocity "
                f"dofs, {lv.nh} hanging nodes")
        r = model.step()
        lv = model.mesh.level_struct(model.mesh.nlev - 1)
        r.update(nel=lv.nel, neq=lv.neq, nh=lv.nh)
        recs.append(r)
        if log:
            log(f"[amr] step {s}: uzawa={r['uzawa']} conv={r['converged']} inner={r['inner']} "
                f"solve={r['solve_ms']:.1f} ms neq={lv.neq} nh={lv.nh}")
    vr, nu = model.diagnostics()
    out = dict(steps=steps, wall_s=time.perf_counter() - t0, vrms=vr, nu_top=nu,
               converged=sum(1 for r in recs if r["converged"]), recs=recs)
    if keep_model:
        out["model"] = model
    else:
        model.ctx.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--base-nx", type=int, default=128)
    ap.add_argument("--extra", type=int, default=3)
    ap.add_argument("--developed", action="store_true",
                    help="start from the synthetic developed temperature on the mesh adapted to it")
    a = ap.parse_args(argv)
    out = run(a.steps, a.every, a.base_nx, a.extra, log=lambda m: print(m, file=sys.stderr, flush=True),
              developed=a.developed)
    rs = out.pop("recs")
    out["uzawa"] = [r["uzawa"] for r in rs]
    out["neq"] = [r["neq"] for r in rs]
    out["solve_ms"] = [round(r["solve_ms"], 2) for r in rs]
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
