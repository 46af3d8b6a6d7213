"""Converged FP64 runs of the oracle, stored as fixtures for the GPU tests.

The GPU box cannot afford the oracle's converged solves inside a test (about
5 s per 64x64 step on one core), so this script runs them here and commits
the results.  A CPU test (tests/test_oracle_runs.py) re-runs the first step
of each with the live oracle and checks the fixture bit for bit.

c2s_64_conv: the benchmark's workload at 64x64 (Ra=1e6, eta=exp(-ln(1e3) T),
seeded initial temperature, free-slip box), 20 time steps.  Each step is the
reference's sequence (buoyancy, viscosity, K/G/F, Uzawa solve from d = p = 0,
nodal velocity, SUPG step) with the Uzawa iteration in its convergent form
(z = r / BAB) and the inner Jacobi-PCG run to convergence (accuracy 1e-20, no
practical cycle cap), so every solve is the FP64 solution of the discrete
system to the Uzawa tolerance dv, dp < 1e-6.  Stored per step: Uzawa cycles,
Vrms of the step's velocity and the top Nusselt number of its new
temperature (host layer's mclh_diagnostics), dt, model time; the fields
T (after steps 0, 4, 9, 19), V, d and p of the last step.

    python tests/golden/make_oracle_runs.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

CONVERGED = dict(accuracy="1e-20", max_cg_cycle="100000", max_citcom_cycle="5000")
STEPS = 20
KEEP_T = (0, 4, 9, 19)


def converged_oracle(name="c2s_64"):
    from tests.oracle_lib import Oracle, ctrl_from_keys, fixture_keys, load_fixture, mesh_from_fixture
    fx = load_fixture(name)
    keys = fixture_keys(fx)
    keys.update(CONVERGED)
    m = mesh_from_fixture(fx)
    return fx, keys, m, Oracle(m, ctrl_from_keys(keys, schur_divide=1))


def host_mesh(keys):
    from mantle_convection_lab_amd import Input, Mesh
    k = dict(keys)
    k.pop("old_T_timestep", None)
    k.update(T_ini="0", write_output="0")
    return Mesh(Input.from_keys(k))


def run(steps=STEPS, name="c2s_64"):
    fx, keys, m, o = converged_oracle(name)
    mesh = host_mesh(keys)
    lev = m["lev"]
    T = fx[f"s0.L{lev}.T.init"].copy()
    hd = hdd = None
    mt = 0.0
    out = {"T.init": T.copy()}
    rec = {k: [] for k in ("uzawa", "vrms", "nu", "dt", "model_time")}
    for s in range(steps):
        A = o.assemble(o.viscosity(T), o.buoyancy(T))
        d, p, st = o.solve(A)
        V = o.v_to_nodes(d)
        T, hd, hdd, mt, hs, _ = o.update_temperature(V, T, hd, hdd, s == 0, mt)
        vr, nu = mesh.diagnostics(T, V.ravel())
        for k, v in (("uzawa", st.uzawa_cycles), ("vrms", vr), ("nu", nu), ("dt", hs.dt), ("model_time", mt)):
            rec[k].append(v)
        if s in KEEP_T:
            out[f"T.s{s}"] = T.copy()
        print(f"step {s}: uzawa={st.uzawa_cycles} inner={st.inner_iterations} vrms={vr:.12g} nu={nu:.12g}",
              flush=True)
    out.update(d=d, p=p, V=V.ravel().copy())
    for k, v in rec.items():
        out[k] = np.array(v, dtype=np.int32 if k == "uzawa" else np.float64)
    out["input_keys"] = np.frombuffer("\n".join(f"{k}={v}" for k, v in keys.items()).encode(), dtype=np.uint8)
    return out


def main():
    out = run()
    path = os.path.join(HERE, "c2s_64_conv.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
