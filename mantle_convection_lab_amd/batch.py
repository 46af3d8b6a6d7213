"""Batched independent instances (new: the reference advances one instance per
job).  N convection instances of uneven mesh size, Rayleigh number and
viscosity contrast, all drawn from one seed, advance together: each round
runs one time step of every instance (buoyancy, viscosity, assembly, the
mixed-precision Uzawa solve, the SUPG step), each on its own device context
with its fields resident in HBM.

The first round runs instance by instance through the ordinary entry points;
there, instances on meshes of side >= `learn_min` consult the learned guess
(the 256x256 U-Net weights applied fully convolutionally at the instance's
side; the gate keeps the previous solution whenever the network's guess has
the larger Schur residual).  After it, a batch (mcl_batch_create) holds each
instance's time step as a captured HIP graph, and every further round
replays all graphs on `lanes` concurrent streams (mcl_batch_step), each
persistent kernel capped at its lane's share of the device; these rounds
start every solve from the previous step's solution (the learned guess is
consulted at the cold start only; on the 256x256 benchmark the previous
step's solution wins from its third step on).  lanes=0 keeps the
instance-by-instance rounds throughout.

Instances are drawn with numpy's PCG64 generator: side = the power of two
nearest (in log space) to 10^U[log10 32, log10 512], Ra = 10^U[4, 7],
contrast = 10^U[0, 3], perturbation seed from the same stream.

    python -m mantle_convection_lab_amd.batch [--instances 96] [--steps 5] [--seed 20261202]
"""
import argparse
import json
import math
import statistics
import sys
import time

import numpy as np

import ctypes as C

from . import _native as N
from .model import Input, MclError, Model
from .learn import Net, default_weights_path, workload_keys

SIDE_MIN, SIDE_MAX = 32, 512


def draw_instances(n=96, seed=20261202):
    """[(side, Ra, contrast, perturbation seed)] of the batched workload."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        s = 10 ** rng.uniform(math.log10(SIDE_MIN), math.log10(SIDE_MAX))
        side = int(2 ** round(math.log2(s)))
        Ra = float(10 ** rng.uniform(4, 7))
        contrast = float(10 ** rng.uniform(0, 3))
        pseed = int(rng.integers(1, 2**40))
        out.append((side, Ra, contrast, pseed))
    return out


def dofs_per_step(model):
    """velocity + pressure + temperature unknowns of the finest level"""
    lv = model.mesh.level_struct(model.mesh.nlev - 1)
    return lv.neq + lv.nel + lv.neqT


class Batch:
    """The instances of one batched run, each a Model on the same device."""

    def __init__(self, specs, device=0, learn_min=128, weights=None, log=None):
        self.specs = list(specs)
        self.nets = {}
        self.models = []
        w = weights or default_weights_path(256)
        for side, Ra, ctr, pseed in self.specs:
            net = None
            if learn_min and side >= learn_min:
                if side not in self.nets:
                    nn = Net(side, side, 1)
                    nn.load(w)
                    self.nets[side] = nn
                net = self.nets[side]
            self.models.append(Model(Input.from_keys(workload_keys(side, Ra, ctr, seed=pseed)), device=device,
                                     net=net))
            if log:
                log(f"[batch] instance {len(self.models) - 1}: {side}x{side} Ra={Ra:.3g} contrast={ctr:.3g}")
        self.dofs = [dofs_per_step(m) for m in self.models]

    def step(self):
        """One time step of every instance through the ordinary entry points;
        returns their step records (wall_s = the instance's step time, host
        calls included)."""
        return [m.step() for m in self.models]

    def capture(self, lanes=4):
        """Capture every instance's step as a graph (after its first step)."""
        L = N.hip_lib()
        arr = (C.c_void_p * len(self.models))(*[m.ctx.h.value for m in self.models])
        self.bh = C.c_void_p()
        rc = L.mcl_batch_create(arr, len(self.models), int(lanes), C.byref(self.bh))
        if rc:
            msgs = [m.ctx.L.mcl_last_error(m.ctx.h).decode(errors="replace") for m in self.models]
            raise MclError(f"mcl_batch_create failed ({rc}): {[x for x in msgs if x][:3]}")
        self.lanes = lanes

    def graph_step(self):
        """One time step of every instance by graph replay on the lanes;
        records like step()'s, with wall_s = the instance's device time on
        its lane."""
        L = N.hip_lib()
        n = len(self.models)
        st = (N.StokesStats * n)()
        hs = (N.HeatStats * n)()
        ms = (C.c_double * n)()
        rc = L.mcl_batch_step(self.bh, None, st, hs, ms)
        if rc:
            raise MclError(f"mcl_batch_step failed ({rc}): {L.mcl_batch_last_error(self.bh).decode()}")
        out = []
        for i, m in enumerate(self.models):
            r = dict(step=m.timestep, uzawa=st[i].uzawa_cycles, converged=st[i].converged, dv=st[i].dv,
                     dp=st[i].dp, inner=st[i].inner_iterations, solve_ms=None, heat_cg=hs[i].cg_cycles, dt=hs[i].dt,
                     model_time=hs[i].model_time, res_mom=st[i].res_momentum, res_cont=st[i].res_continuity,
                     wall_s=ms[i] * 1e-3)
            m.history.append(r)
            m.timestep += 1
            out.append(r)
        return out

    def close(self):
        if getattr(self, "bh", None):
            N.hip_lib().mcl_batch_destroy(self.bh)
            self.bh = None
        for m in self.models:
            m.ctx.close()
        for n in self.nets.values():
            n.close()


def run(n=96, steps=5, seed=20261202, learn_min=128, lanes=4, log=print):
    """The batched workload: round 0 instance by instance, then graph rounds
    on `lanes` streams (lanes=0: every round instance by instance).
    Throughput counts every round; the capture time is reported apart."""
    specs = draw_instances(n, seed)
    t0 = time.perf_counter()
    b = Batch(specs, learn_min=learn_min, log=log)
    setup_s = time.perf_counter() - t0
    recs = []
    rounds = []
    wall = 0.0
    capture_s = 0.0
    for s in range(steps):
        t = time.perf_counter()
        if s == 0 or lanes <= 0:
            rs = b.step()
        else:
            if not getattr(b, "bh", None):
                tc = time.perf_counter()
                b.capture(lanes)
                capture_s = time.perf_counter() - tc
                t = time.perf_counter()
            rs = b.graph_step()
        round_s = time.perf_counter() - t
        rounds.append(round_s)
        wall += round_s
        recs.append(rs)
        if log:
            log(f"[batch] round {s}: {sum(r['converged'] for r in rs)}/{len(rs)} converged, "
                f"uzawa {sum(r['uzawa'] for r in rs)}, {round_s:.2f} s")
    # per-instance step time: the host wall time of an instance-by-instance
    # step, the device time of its graph on its lane in a graph round
    step_ms = sorted(r["wall_s"] * 1e3 for rs in recs for r in rs)
    q = statistics.quantiles(step_ms, n=100) if len(step_ms) >= 2 else step_ms * 99
    out = dict(instances=n, steps=steps, seed=seed, generator="numpy PCG64", setup_s=setup_s, wall_s=wall,
               lanes=lanes, capture_s=capture_s,
               round_s=[round(x, 4) for x in rounds],
               dof_updates_per_s=sum(b.dofs) * steps / wall, p50_step_ms=q[49], p99_step_ms=q[98],
               converged=sum(r["converged"] for rs in recs for r in rs), solves=n * steps,
               sides={str(k): sum(1 for s in specs if s[0] == k) for k in sorted({s[0] for s in specs})},
               learned_instances=sum(1 for s in specs if learn_min and s[0] >= learn_min))
    b.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=96)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--learn-min", type=int, default=128, help="smallest side that consults the network (0: none)")
    ap.add_argument("--lanes", type=int, default=4, help="concurrent streams of the graph rounds (0: none)")
    a = ap.parse_args(argv)
    out = run(a.instances, a.steps, a.seed, a.learn_min, a.lanes, log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
