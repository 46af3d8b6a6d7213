"""Whole-run parity of the FP64 path (MI355X) against the reference's own C1
run: 64x64, Ra=1e4, isoviscous, T = (1-z) + 0.01 cos(pi x) sin(pi z), 500 time
steps (golden fixture c1_64_run, written by the reference built from its
sources: dt and model time after every step, T, V, p every 50 steps).

The reference's own run is fragile: its Uzawa iteration hits the 2000-cycle
cap without converging at steps 218, 226, 230, 418, 423-425, 428-432 and 438
(the bit-exact CPU oracle reproduces the fixture and shows it), and where such
a solve stops depends on the last bit of every dot product.  With
exact_dots=1 the device sums every dot product of the solver in the
reference's sequential order; the remaining differences are last-ulp ones of
the device's transcendental functions in the SUPG step.  Bars: the same
13 solves hit the cap; dt within 1e-6 relative at every converged step (1e-3 at
the capped ones, whose iterates are not solutions) and model time within 1e-6;
T and V within 1e-6 and p within 1e-3 relative L2 at every snapshot (the
stopping rule bounds the pressure update, dp < 1e-6, not its error, which
converges slowly in the near-checkerboard modes, so a last-ulp difference
shows in p at the 1e-4 level); Vrms and
the top Nusselt number of the last step within 1e-6 relative.  (With the default compensated parallel sums the run follows the
reference to 1e-8 in T until the capped solves, then separates at the level
of the Uzawa tolerance: Vrms 1e-5, Nu 5e-5 after 500 steps.)
"""
import numpy as np
import pytest

from tests.cases import case_input, rel_l2

pytestmark = pytest.mark.gpu


def _run(exact, steps=500):
    from mantle_convection_lab_amd import Model
    fx, inp = case_input("c1_64_run")
    m = Model(inp, ctrl_overrides=dict(exact_dots=exact))
    out = dict(dt=[], time=[], snaps={}, caps=[])
    for t in range(steps):
        r = m.step()
        if not r["converged"]:
            out["caps"].append(t)
        out["dt"].append(r["dt"])
        out["time"].append(r["model_time"])
        if f"s{t}.T" in fx:
            out["snaps"][t] = (m.field("T"), m.field("V"), m.field("P"), r["uzawa"])
            T, V, p = out["snaps"][t][:3]
            Vr = np.concatenate([fx[f"s{t}.V0"], fx[f"s{t}.V1"]])
            print(f"  step {t}: uzawa={r['uzawa']} T {rel_l2(T, fx[f's{t}.T']):.1e} V {rel_l2(V, Vr):.1e} "
                  f"p {rel_l2(p, fx[f's{t}.p']):.1e} capped so far {out['caps']}", flush=True)
    out["diag"] = m.diagnostics()
    ref = m.mesh.diagnostics(fx["s499.T"], np.concatenate([fx["s499.V0"], fx["s499.V1"]]))
    return fx, out, ref


def test_c1_500_steps_match_reference():
    print("[gpu-run-c1] 64x64 Ra=1e4, 500 steps, reference-order dot products", flush=True)
    fx, out, ref = _run(1)
    print(f"  capped solves: {out['caps']}", flush=True)
    print(f"  Vrms {out['diag'][0]!r} vs {ref[0]!r}; Nu {out['diag'][1]!r} vs {ref[1]!r}", flush=True)
    caps = [218, 226, 230, 418, 423, 424, 425, 428, 429, 430, 431, 432, 438]
    assert out["caps"] == caps
    edt = [abs(out["dt"][t] - fx[f"s{t}.dt"][0]) / fx[f"s{t}.dt"][0] for t in range(500)]
    etm = [abs(out["time"][t] - fx[f"s{t}.model_time"][0]) / fx[f"s{t}.model_time"][0] for t in range(500)]
    print(f"  worst dt: converged steps {max(e for t, e in enumerate(edt) if t not in caps):.1e}, "
          f"capped steps {max(edt[t] for t in caps):.1e}; worst model time {max(etm):.1e}", flush=True)
    for t in range(500):
        # a capped solve's iterate is not a solution; its dt follows that iterate's velocity
        assert edt[t] <= (1e-3 if t in caps else 1e-6), t
        assert etm[t] <= 1e-6, t
    for t, (T, V, p, uz) in out["snaps"].items():
        assert rel_l2(T, fx[f"s{t}.T"]) < 1e-6, t
        assert rel_l2(V, np.concatenate([fx[f"s{t}.V0"], fx[f"s{t}.V1"]])) < 1e-6, t
        assert rel_l2(p, fx[f"s{t}.p"]) < 1e-3, t
    assert abs(out["diag"][0] - ref[0]) < 1e-6 * ref[0]
    assert abs(out["diag"][1] - ref[1]) < 1e-6 * ref[1]
