"""Launch-shape autotuner of the mixed-precision Stokes solve.

For each mesh side it times the solve under every combination of
  * MCL_MIXED_BLOCKS workgroups of the persistent mixed solve (128, 192,
                     256 = one per CU of the MI355X),
  * MCL_MG_FUSE      V-cycle phase fusion (1: one sweep per phase, 2: two),
  * MCL_MG_SWEEPS    Jacobi sweeps per level and leg (1, 2),
each candidate in a fresh process (the library reads these settings when it
builds a context's halo lists and launches its solves), on the first `steps`
time steps of a seeded instance of that side (Ra=1e6, contrast 1e3; the same
tolerance every time, so fewer sweeps may need more iterations).  The score is
the summed device time of the solves.  The best candidate per side is written
to weights/tuning.json; `apply(side)` exports it before a context is created
(variables the caller has set itself are left alone).  The table is advisory:
on the MI355X the tuner picks 256 workgroups, two-sweep fusion and two sweeps
at 128x128, 256x256 and 512x512, which are the library's defaults there, so
nothing applies it by default.  (The FP64 kernels' own knob, MCL_FP64_BLOCKS,
was swept by hand: 129 workgroups at 256x256, the default, 32.6 us per inner
iteration against 33.3-38.2 us for 192, 256 and 512.)

    python -m mantle_convection_lab_amd.tune [--sides 128 256 512] [--steps 3]
"""
import argparse
import itertools
import json
import os
import subprocess
import sys

KNOBS = ("MCL_MIXED_BLOCKS", "MCL_MG_FUSE", "MCL_MG_SWEEPS")
CANDIDATES = [dict(zip(KNOBS, v)) for v in itertools.product((128, 192, 256), (1, 2), (1, 2))]
DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "tuning.json")


def measure(side, steps, seed=20261202):
    """Solve time (ms) and iteration counts of the first `steps` steps of one
    seeded instance, in this process with its current environment."""
    from .model import Input, Model
    from .learn import workload_keys
    m = Model(Input.from_keys(workload_keys(side, 1e6, 1e3, seed=seed)))
    ms, uz, inner = 0.0, 0, 0
    for _ in range(steps):
        r = m.step()
        if not r["converged"]:
            raise RuntimeError("solve did not converge")
        ms += r["solve_ms"]
        uz += r["uzawa"]
        inner += r["inner"]
    m.ctx.close()
    return dict(ms=ms, uzawa=uz, inner=inner)


def tune_side(side, steps, log=print):
    rows = []
    for cand in CANDIDATES:
        env = dict(os.environ)
        env.update({k: str(v) for k, v in cand.items()})
        cmd = [sys.executable, "-m", "mantle_convection_lab_amd.tune", "--measure", str(side), "--steps", str(steps)]
        r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900,
                           cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
        if r.returncode != 0:
            log(f"[tune] {side}: {cand} failed: {r.stderr.strip()[-200:]}")
            continue
        res = json.loads(r.stdout.strip().splitlines()[-1])
        res.update(cand)
        rows.append(res)
        log(f"[tune] {side}: {cand} -> {res['ms']:.1f} ms ({res['uzawa']} Uzawa, {res['inner']} inner iterations)")
    if not rows:
        raise RuntimeError(f"no candidate ran for side {side}")
    best = min(rows, key=lambda r: r["ms"])
    return {k: best[k] for k in KNOBS}, rows


def apply(side, path=DEFAULT_PATH):
    """Export the tuned settings of `side` (if tuned) unless already set."""
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        tab = json.load(f)
    best = tab.get("best", {}).get(str(side))
    if not best:
        return {}
    done = {}
    for k, v in best.items():
        if k in KNOBS and k not in os.environ:
            os.environ[k] = str(v)
            done[k] = v
    return done


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--sides", type=int, nargs="+", default=[128, 256, 512])
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--out", default=DEFAULT_PATH)
    ap.add_argument("--measure", type=int, default=0, help=argparse.SUPPRESS)
    a = ap.parse_args(argv)
    if a.measure:
        print(json.dumps(measure(a.measure, a.steps)), flush=True)
        return 0
    log = lambda m: print(m, file=sys.stderr, flush=True)
    out = {"steps": a.steps, "knobs": list(KNOBS), "best": {}, "table": {}}
    for side in a.sides:
        best, rows = tune_side(side, a.steps, log)
        out["best"][str(side)] = best
        out["table"][str(side)] = rows
        log(f"[tune] {side}: best {best}")
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out["best"]))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
