"""Benchmark: Stokes solves/s to a 1e-6 Uzawa tolerance, 256x256, Ra=1e6:
learned guess + mixed precision against the plain FP64 path.

Workload (one "step" = one pass of the hot path): 256x256 unit box, Ra=1e6,
eta = exp(-ln(1e3) T), free-slip walls, T=1 bottom / T=0 top, insulating sides,
initial T = (1-z) + 0.01 (2u-1) sin(pi z) with u from SplitMix64(seed + node).
Each step runs buoyancy, viscosity, K/G/F assembly, the Uzawa Stokes solve to
dv, dp < 1e-6 (Solver.c:131,202-204), and the SUPG temperature step, all
resident on the GPU.  A solve counts as a Stokes solve only if it also leaves
a relative momentum residual |F - K u - G p| / |F| <= 1e-6 (RES_BAR).

Paths (the Uzawa outer iteration, z = r / BAB, and its stopping rule are the
same in all):
  learned    the previous step's solution or the U-Net guess, whichever has the
             smaller initial Schur residual, then FP64 CG inner solves
             preconditioned by the FP32 multigrid V-cycle ("mixed": the same
             without the network)
  fp64       the plain FP64 method: Jacobi-PCG inner solves in FP64 (the
             reference's solver="cg"), run to convergence (accuracy 1e-20) so
             that its solves meet the residual bar, warm-started from the
             previous step like the learned path's fallback
  fp64-caps  the same with the reference's own inner caps (accuracy 1e-8,
             200 cycles) from a zero guess every step: its solves stop at
             dv, dp < 1e-6 but leave residuals far above the bar (reported,
             not the denominator)

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--caps-steps K2]
                    [--path learned|mixed|fp64] [--no-accuracy] [--no-cpu]
                    [--copy-calibration]

Prints one JSON line.  value = Stokes solves per second of the primary path
over its K timed steps (time steps 0..K-1 of a fresh instance after W warmup
steps on a throwaway instance; inputs already in HBM; wall clock between two
device synchronisations).  The accuracy leg then replays the primary path
untimed in lockstep with the fp64 path and reports, per step, the relative
errors of Vrms and the top Nusselt number, and at the end those of T, u and p
(p without its constant mode, which the closed box leaves free); the replay
is checked to reproduce the timed run bit for bit.  The fp64 path's solves/s
(the denominator of speedup_vs_fp64) is measured in that leg, each step
between two synchronisations.  Then the fp64-caps path, the roofline of the
primary path's dominant kernel (the persistent Uzawa kernel) from its
HIP-event durations, and a CPU baseline (the FP64 oracle on one core) on a
bounded sample of the same workload.
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

SIDE = 256
WORKLOAD = dict(level=9, mg_min_level=8, nprox=1, nproz=1, model_length=1.0, model_height=1.0, Rayleigh=1e6,
                A=math.log(1000.0), refT=0.0, lower_mantle_jump=1.0, T_ini=4, T_seed=20261202,
                T_mantle_perturb=0.01, Q=0.0, theta=0.5, develop_mode=0, solver="cg", cg_precondition=1,
                accuracy=1e-8, citcom_solver_accu=1e-6, max_citcom_cycle=2000, max_cg_cycle=200,
                top_v_bc="SV", bot_v_bc="SV", lef_v_bc="VS", rig_v_bc="VS", top_T_bc="T", bot_T_bc="T",
                lef_T_bc="F", rig_T_bc="F", top_T_bc_val=0.0, bot_T_bc_val=1.0, write_output=0,
                schur_mode=1)  # z = r / BAB: the reference's multiply (Q1) does not converge here
HBM_PEAK_GBS = 8000.0
RES_BAR = 1e-6  # relative momentum residual a counted Stokes solve must meet
FP64_CONVERGED = dict(accuracy=1e-20, max_cg_cycle=100000, warm_start=1)
MG_SWEEPS = 2  # Jacobi sweeps per level and leg of the V-cycle (the library default)


def path_keys(path):
    """Input keys of a path: "learned"/"mixed", "fp64" (converged inner
    solves, warm start) or "fp64-caps" (the reference's inner caps, zero
    guess)."""
    k = dict(WORKLOAD)
    if path in ("learned", "mixed"):
        k.update(precision=1, warm_start=1, mg_min_level=2)
    elif path == "fp64":
        k.update(FP64_CONVERGED)
    elif path != "fp64-caps":
        raise ValueError(path)
    return k


# ---------------------------------------------------------------- roofline

def level_sizes(model):
    """(neq, nnz_K) of every multigrid level, finest last."""
    out = []
    for k in range(model.mesh.nlev):
        lv = model.mesh.level_struct(k)
        ncol, _, _ = model.ctx.get_matrix(lv.lev, "K", 36)
        out.append((lv.neq, int(ncol.sum()), lv.nel))
    return out


def fp64_bytes(rec, sizes, w=8, iw=4):
    """Algorithmic bytes of k_uzawa_fp64 for one solve: per inner Jacobi-PCG
    iteration nnz_K (w+iw) + 11 n_v w; per Uzawa iteration
    2 nnz_G (w+iw) + 7 n_v w + 12 n_p w with nnz_G = 4 n_v."""
    n_v, nnz, n_p = sizes[-1]
    b_cg = nnz * (w + iw) + 11 * n_v * w
    b_uz = 2 * (4 * n_v) * (w + iw) + 7 * n_v * w + 12 * n_p * w
    return rec["inner"] * b_cg + rec["uzawa"] * b_uz


def mixed_bytes(rec, sizes, pre=MG_SWEEPS):
    """Algorithmic bytes of k_uzawa_mixed for one solve: what the algorithm
    must move at the least, not what the kernel moves.  Per inner iteration:
    FP64 part  nnz_K 12 (K d' with 8-byte values, 4-byte columns)
               + n_v (4 + 8 + 8 + 8)        (read z32, d; write d', q)
               + n_v (4*8 + 2*8 + 3*4)      (x, r, d', q read; x, r written;
                                             b32, x32 written, 1/diag read);
    V-cycle    per level l above the coarsest (n_l rows, nnz_l entries, FP32
               values and int32 columns): the level's K once per leg, 2 nnz_l 8
               (a leg's `pre` sweeps and its residual reuse the rows it read),
               the down leg reading b and 1/diag and writing x, 12 n_l, the up
               leg reading x, b, 1/diag and writing x, 16 n_l, and the transfers
               writing the coarse right-hand side and reading the coarse
               correction, 8 n_(l-1);
    plus the Uzawa iterations as in fp64_bytes.  (`pre` does not enter: more
    sweeps per leg reuse the same rows.)"""
    n_v, nnz, n_p = sizes[-1]
    b_fp64 = nnz * 12 + n_v * 28 + n_v * 60
    b_v = 0
    for k in range(1, len(sizes)):
        n, z, _ = sizes[k]
        b_v += 2 * z * 8 + 28 * n + 8 * sizes[k - 1][0]
    b_uz = 2 * (4 * n_v) * 12 + 7 * n_v * 8 + 12 * n_p * 8
    return rec["inner"] * (b_fp64 + b_v) + rec["uzawa"] * b_uz


def pmc_traffic(kernel, recs, bytes_per_launch):
    """HBM bytes per launch of `kernel`: the committed rocprofv3 PMC summary
    gives the measured traffic of a profiled launch (FETCH_SIZE + WRITE_SIZE
    scaled by the copy-kernel factors) over that launch's algorithmic bytes;
    the ratio is applied to this run's mean algorithmic bytes per launch.
    None without a summary."""
    path = os.path.join(REPO, "profiles", f"{kernel}_256_pmc.txt")
    if not os.path.exists(path) or not recs:
        return None
    ratio = None
    for line in open(path):
        if line.startswith("traffic_over_algorithmic:"):
            ratio = float(line.split(":", 1)[1])
    if ratio is None:
        return None
    return ratio * bytes_per_launch


# ---------------------------------------------------------------- paths

def counted(r):
    """A Stokes solve: dv, dp below the tolerance and the momentum residual
    within RES_BAR."""
    return bool(r["converged"]) and r["res_mom"] <= RES_BAR


def load_net(path, log):
    """(net, path, note): the trained network for the learned path."""
    if path != "learned":
        return None, path, ""
    from mantle_convection_lab_amd import learn
    wp = learn.default_weights_path(SIDE)
    if not os.path.exists(wp):
        return None, "mixed", f"no trained weights at {os.path.relpath(wp, REPO)}: previous-step guess only"
    net = learn.Net(SIDE, SIDE, 1)
    net.load(wp)
    return net, path, ""


def make_model(path, net=None):
    from mantle_convection_lab_amd import Input, Model
    return Model(Input.from_keys(path_keys(path)), device=0, net=net)


def warm_up(path, net, warmup, log):
    """Warmup steps on a throwaway instance of the same workload (code paths,
    caches, allocations); timed runs then start from step 0 of a fresh
    instance, so every path is measured over the same time steps."""
    warm = []
    if warmup > 0:
        tw = make_model(path, net)
        for i in range(warmup):
            r = tw.step()
            warm.append(r)
            log(f"[bench] {path} warmup {i}: uzawa={r['uzawa']} inner={r['inner']} solve={r['solve_ms']:.1f} ms")
        tw.ctx.close()
    return warm


def step_log(path, i, r, net):
    g = f" guess={'net' if r.get('guess_accepted') else 'prev'}" if "guess_accepted" in r else \
        (" guess=skip" if net is not None else "")
    return (f"[bench] {path} step {i}: uzawa={r['uzawa']} conv={r['converged']} inner={r['inner']} "
            f"solve={r['solve_ms']:.2f} ms res={r['res_mom']:.1e}/{r['res_cont']:.1e}{g}")


def run_timed(path, net, steps, warmup, log, note=""):
    """K steps of one path between two device synchronisations."""
    warm = warm_up(path, net, warmup, log)
    model = make_model(path, net)
    model.ctx.synchronize()
    t0 = time.perf_counter()
    recs = []
    for i in range(steps):
        r = model.step()
        recs.append(r)
        log(step_log(path, i, r, net))
    model.ctx.synchronize()
    wall = time.perf_counter() - t0
    return finish(model, path, recs, warm, wall, note)


def finish(model, path, recs, warm, wall, note=""):
    sizes = level_sizes(model)
    vr, nu = model.diagnostics()
    neqT = model.mesh.level_struct(model.mesh.nlev - 1).neqT
    return dict(path=path, note=note, recs=recs, warm=warm, wall=wall, sizes=sizes, neqT=neqT, vrms=vr, nu_top=nu,
                model=model)


def rel(a, b):
    import numpy as np
    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (nb if nb > 0 else 1.0)


def nodal_pressure(model, p):
    """Element pressures averaged onto the nodes (each node takes the mean of
    its elements): this removes the Q1-P0 pair's checkerboard pressure mode,
    which the Schur complement barely sees, so the Uzawa iteration leaves it
    wherever the iteration history put it."""
    import numpy as np
    ien = model.mesh.level(-1)["ien"]
    s = np.zeros(int(ien.max()) + 1)
    c = np.zeros_like(s)
    for a in range(4):
        np.add.at(s, ien[a], p)
        np.add.at(c, ien[a], 1.0)
    return s / np.maximum(c, 1.0)


def accuracy_leg(prim_path, net, steps, prim_final, log):
    """Replay the primary path (untimed) in lockstep with the fp64 path
    (converged inner solves): per-step errors of Vrms and Nu, final errors of
    T, u (nodal velocity) and p (mean removed; also averaged onto the nodes,
    without the checkerboard mode).  The fp64 path's steps are
    each timed between two synchronisations (a step ends with the heat
    step's statistics read back), which gives the denominator."""
    import numpy as np
    warm_up("fp64", None, 1, log)
    a = make_model(prim_path, net)
    b = make_model("fp64")
    b.ctx.synchronize()
    ev, en, walls, brecs = [], [], [], []
    for i in range(steps):
        a.step()
        t0 = time.perf_counter()
        rb = b.step()
        walls.append(time.perf_counter() - t0)
        brecs.append(rb)
        va, na = a.diagnostics()
        vb, nb = b.diagnostics()
        ev.append(abs(va - vb) / abs(vb))
        en.append(abs(na - nb) / abs(nb))
        log(step_log("fp64", i, rb, None) + f" | {prim_path} Vrms err {ev[-1]:.1e} Nu err {en[-1]:.1e}")
    Ta, Tb = a.field("T"), b.field("T")
    Va, Vb = a.field("V"), b.field("V")
    pa, pb = a.field("P"), b.field("P")
    na_, nb_ = nodal_pressure(a, pa), nodal_pressure(b, pb)
    replay_same = all(np.array_equal(a.field(f), prim_final[f]) for f in prim_final)
    acc = dict(reference="fp64 path: FP64 Uzawa + Jacobi-PCG inner solves converged (accuracy 1e-20), warm start",
               steps=steps, vrms_rel_err_max=max(ev), nu_rel_err_max=max(en), vrms_rel_err=ev, nu_rel_err=en,
               T_rel_l2=rel(Ta, Tb), u_rel_l2=rel(Va, Vb), p_rel_l2=rel(pa - pa.mean(), pb - pb.mean()),
               p_nodal_rel_l2=rel(na_ - na_.mean(), nb_ - nb_.mean()), replay_bitwise=replay_same)
    res = dict(path="fp64", note="", recs=brecs, warm=[], wall=sum(walls), sizes=level_sizes(b),
               neqT=b.mesh.level_struct(b.mesh.nlev - 1).neqT, vrms=None, nu_top=None, model=b)
    res["vrms"], res["nu_top"] = b.diagnostics()
    a.ctx.close()
    return acc, res


def summary(res, steps):
    recs = res["recs"]
    conv = sum(1 for r in recs if counted(r))
    solve_ms = [r["solve_ms"] for r in recs]
    nb = res["sizes"][-1]
    rm = [r["res_mom"] for r in recs]
    return dict(solves_per_s=conv / res["wall"] if res["wall"] > 0 else 0.0, counted_solves=conv,
                dvdp_converged=sum(1 for r in recs if r["converged"]), steps=steps, wall_s=res["wall"],
                median_solve_ms=statistics.median(solve_ms) if solve_ms else None,
                uzawa_cycles=[r["uzawa"] for r in recs], inner_iterations=[r["inner"] for r in recs],
                res_momentum_min=min(rm) if rm else None, res_momentum_max=max(rm) if rm else None,
                dof_updates_per_s=(nb[0] + nb[2] + res["neqT"]) * steps / res["wall"],
                vrms=res["vrms"], nu_top=res["nu_top"])


def roofline(res, kernel):
    recs = res["recs"]
    fn = fp64_bytes if kernel == "k_uzawa_fp64" else mixed_bytes
    bytes_per = [fn(r, res["sizes"]) for r in recs]
    ms = [r["solve_ms"] for r in recs]
    if not ms:
        return None
    achieved = sum(bytes_per) / (sum(ms) * 1e-3) / 1e9
    traffic = pmc_traffic(kernel, recs, sum(bytes_per) / len(bytes_per))
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
            "avg_launch_ms": sum(ms) / len(ms), "bytes_per_launch": sum(bytes_per) / len(bytes_per)}


# ---------------------------------------------------------------- CPU baseline

def cpu_baseline(full_inner, full_cycles, sample_cycles=12):
    """The FP64 oracle on the host cores (OpenMP over its row products and
    vector updates; its sums keep their sequential order) on the first solve
    of the same workload with the fp64 path's settings (converged inner
    solves), bounded to `sample_cycles` Uzawa iterations and scaled by inner
    iterations to the `full_inner` the fp64 GPU path needed for that solve:
    assembly time + time per inner iteration x full_inner.  With the
    reference/oracle time ratio that oracle/calibrate.py measured where the
    reference builds (one core each), the same figure is also stated for the
    reference."""
    from mantle_convection_lab_amd import Input, Mesh
    from tests.oracle_lib import Oracle, ctrl_from_control, mesh_from_host_level
    threads = int(os.environ.get("OMP_NUM_THREADS") or min(16, os.cpu_count() or 1))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    inp = Input.from_keys(path_keys("fp64"))
    mesh = Mesh(inp)
    o = Oracle(mesh_from_host_level(mesh.level(-1)), ctrl_from_control(inp.control(), max_citcom_cycle=sample_cycles))
    T0 = mesh.initial_temperature(-1)
    t = time.perf_counter()
    A = o.assemble(o.viscosity(T0), o.buoyancy(T0))
    t_asm = time.perf_counter() - t
    t = time.perf_counter()
    d, p, st = o.solve(A)
    t_solve = time.perf_counter() - t
    per_inner = t_solve / max(st.inner_iterations, 1)
    est = t_asm + per_inner * full_inner
    out = dict(value=1.0 / est, unit="Stokes solves/s", cores=threads, kind="port",
               sample=f"oracle FP64 (C restatement, {threads} OpenMP threads) on the first 256x256 solve with the "
                      f"fp64 path's settings: assembly {t_asm:.2f} s + {st.uzawa_cycles} Uzawa iterations "
                      f"({st.inner_iterations} inner PCG iterations) in {t_solve:.1f} s, scaled by inner "
                      f"iterations to the {full_inner} ({full_cycles} Uzawa) the fp64 GPU path needed for that "
                      f"solve: {est:.0f} s per solve")
    cal = os.path.join(REPO, "oracle", "calibration.json")
    if os.path.exists(cal):
        with open(cal) as f:
            c = json.load(f)
        ratio = c["reference_over_oracle"]
        out["calibration"] = dict(reference_over_oracle_one_core=ratio, host=c.get("host"),
                                  note="reference and oracle, one core each, time per inner iteration on the "
                                       "same 256x256 solve (oracle/calibrate.py)")
        if threads == 1:
            out["reference_equivalent_value"] = out["value"] / ratio
    return out


def copy_calibration():
    """A plain device copy of known size, for scaling the PMC byte counters."""
    import torch
    n = 1 << 28  # 1 GiB of float32
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4,
                      "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--caps-steps", type=int, default=None,
                    help="timed steps of the fp64-caps path (default: --steps; 0 skips it)")
    ap.add_argument("--path", default="learned", choices=["learned", "mixed", "fp64"])
    ap.add_argument("--no-accuracy", action="store_true", help="skip the accuracy leg (and the fp64 denominator)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--workload", default="c2", choices=["c2", "batched"],
                    help="c2: the 256x256 metric above (default); batched: the 96-instance workload, one line of "
                         "aggregate DOF-updates/s and per-instance step-time percentiles")
    ap.add_argument("--lanes", type=int, default=4, help="batched: concurrent streams of the graph rounds")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(how the path would shard is described, not built)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return

    def log(s):
        print(s, file=sys.stderr, flush=True)

    if args.workload == "batched":
        batched_line(args, log)
        return
    log(f"[bench] {SIDE}x{SIDE}, primary path {args.path}")
    net, ppath, note = load_net(args.path, log)
    prim = run_timed(ppath, net, args.steps, args.warmup, log, note)
    ps = summary(prim, args.steps)
    kernel = "k_uzawa_fp64" if ppath == "fp64" else "k_uzawa_mixed"
    rl = roofline(prim, kernel)
    final = {f: prim["model"].field(f) for f in ("T", "D", "P")}
    prim["model"].ctx.close()
    acc, fp = None, None
    if args.path != "fp64" and not args.no_accuracy and args.steps > 0:
        log("[bench] accuracy leg: primary path replayed next to the fp64 path (converged inner solves)")
        acc, fpr = accuracy_leg(ppath, net, args.steps, final, log)
        fp = summary(fpr, args.steps)
        fp["roofline"] = roofline(fpr, "k_uzawa_fp64")
        fpr["model"].ctx.close()
    elif args.path == "fp64":
        fp = ps
    caps = None
    caps_steps = args.steps if args.caps_steps is None else args.caps_steps
    if caps_steps > 0:
        cr = run_timed("fp64-caps", None, caps_steps, min(args.warmup, 1), log)
        caps = summary(cr, caps_steps)
        caps["roofline"] = roofline(cr, "k_uzawa_fp64")
        caps["solves_per_s_dvdp_only"] = caps["dvdp_converged"] / cr["wall"] if cr["wall"] > 0 else 0.0
        cr["model"].ctx.close()
    if net is not None:
        net.close()
    out = {
        "metric": "Stokes solves/s at 1e-6 resid, 256² Ra=1e6: learned+FP32 vs plain FP64",
        "value": ps["solves_per_s"],
        "unit": "Stokes solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": prim["wall"] / max(args.steps, 1) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if ppath == "fp64" else "f64+f32",
        "data": "synthetic (seeded initial temperature; network weights trained on solver snapshots)",
        "config": {"workload": f"{SIDE}x{SIDE} box, Ra=1e6, eta=exp(-ln(1e3) T), Uzawa dv,dp<1e-6 and "
                               f"|F-Ku-Gp|/|F|<={RES_BAR:g}, path={ppath}",
                   "side": SIDE, "Ra": 1e6, "viscosity_contrast": 1e3, "path": ppath, "res_bar": RES_BAR,
                   "note": prim["note"], "primary": ps, "fp64": fp if fp is not ps else "primary",
                   "fp64_reference_caps": caps,
                   "speedup_vs_fp64": (ps["solves_per_s"] / fp["solves_per_s"])
                   if fp and fp is not ps and fp["solves_per_s"] > 0 else None},
        "accuracy": acc,
        "roofline": rl,
    }
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle on the host cores, bounded sample)...")
        # the fp64 GPU path's first solve (342 793 inner, 310 Uzawa iterations
        # measured on the MI355X when this run has no fp64 leg)
        first = (fp["inner_iterations"][0], fp["uzawa_cycles"][0]) if fp and fp["inner_iterations"] else (342793, 310)
        out["cpu_baseline"] = cpu_baseline(*first)
    print(json.dumps(out), flush=True)


def batched_line(args, log):
    """The batched workload: 96 seeded instances (sides 32-512, Ra 1e4-1e7,
    contrast 1-1e3), `steps` rounds (default 5): round 0 instance by instance
    with the learned guess on sides >= 128, then graph rounds on `lanes`
    streams.  value = aggregate DOF-updates/s (velocity + pressure +
    temperature unknowns of every instance per round, over the rounds' wall
    time); the sequential form of the same rounds (lanes 0) is measured
    after it for comparison unless --no-accuracy is given."""
    from mantle_convection_lab_amd import batch
    steps = args.steps if args.steps != 100 else 5
    log(f"[bench] batched workload: 96 instances, {steps} rounds, lanes {args.lanes}")
    out = batch.run(96, steps, lanes=args.lanes, log=log)
    seq = None
    if not args.no_accuracy:
        log("[bench] the same rounds instance by instance (lanes 0)")
        seq = batch.run(96, steps, lanes=0, log=log)
    line = {"metric": "aggregate DOF-updates/s, 96 batched instances", "value": out["dof_updates_per_s"],
            "unit": "DOF-updates/s", "n_gpus": 1, "steps": steps, "warmup": 0,
            "ms_per_step": out["wall_s"] / max(steps, 1) * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic (seeded instances)",
            "config": {"workload": "96 seeded instances, sides 32-512, Ra 1e4-1e7, contrast 1-1e3, mixed "
                                   "precision, learned guess on sides >= 128 at the cold start",
                       "batched": out, "sequential": seq,
                       "speedup_vs_sequential": (out["dof_updates_per_s"] / seq["dof_updates_per_s"]) if seq else None}}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
