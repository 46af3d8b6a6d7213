"""Benchmark: Stokes solves/s to a 1e-6 Uzawa tolerance, 256x256, Ra=1e6:
learned guess + mixed precision against the plain FP64 path.

Workload (one "step" = one pass of the hot path): 256x256 unit box, Ra=1e6,
eta = exp(-ln(1e3) T), free-slip walls, T=1 bottom / T=0 top, insulating sides,
initial T = (1-z) + 0.01 (2u-1) sin(pi z) with u from SplitMix64(seed + node).
Each step runs buoyancy, viscosity, K/G/F assembly, the Uzawa Stokes solve to
dv, dp < 1e-6 (Solver.c:131,202-204), and the SUPG temperature step, all
resident on the GPU.

Paths (the Uzawa outer iteration and its stopping rule are the same in both):
  fp64     the plain method: zero initial guess every step, Jacobi-PCG inner
           solves in FP64 (the reference's solver="cg" with its caps)
  learned  the previous step's solution or the U-Net guess, whichever has the
           smaller momentum residual, then FP64 CG inner solves preconditioned
           by the FP32 multigrid V-cycle; "mixed" is the same without the net

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--fp64-steps K2]
                    [--path learned|mixed|fp64] [--no-cpu] [--copy-calibration]

Prints one JSON line: value = converged Stokes solves per second of the primary
path over its K timed steps (default 100: the workload's whole run; time steps
0..K-1 of a fresh instance after W
warmup steps on a throwaway instance; inputs already in HBM; wall clock between
two device synchronisations), the plain FP64 path measured the same way in the
same run over the same time steps, the roofline of the primary path's dominant kernel (the persistent Uzawa
kernel) from its HIP-event durations, and a CPU baseline (the FP64 oracle on
one core) on a bounded sample of the same workload.
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
MG_SWEEPS = 2  # Jacobi sweeps per level and leg of the V-cycle (the library default)


def path_keys(path):
    k = dict(WORKLOAD)
    if path in ("learned", "mixed"):
        k.update(precision=1, warm_start=1, mg_min_level=2)
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
    """Algorithmic bytes of k_uzawa_mixed for one solve.  Per inner iteration:
    FP64 part  nnz_K 12 (K d' with 8-byte values, 4-byte columns)
               + n_v (4 + 8 + 8 + 8)        (read z32, d; write d', q)
               + n_v (4*8 + 2*8 + 3*4)      (x, r, d', q read; x, r written;
                                             b32, x32 written, 1/diag read);
    V-cycle    per level l with n_l rows and nnz_l entries (FP32 values,
               int32 columns): 2 pre passes over K32 (sweeps 2..pre, the
               residual, the post-sweeps) at nnz_l 8 + 3 n_l 4 each, plus the
               first sweep 3 n_l 4, restriction and prolongation 6 n_l 4
               (levels below the finest counted the same way);
    plus the Uzawa iterations as in fp64_bytes."""
    n_v, nnz, n_p = sizes[-1]
    b_fp64 = nnz * 12 + n_v * 28 + n_v * 60
    b_v = 0
    for (n, z, _) in sizes[1:] if len(sizes) > 1 else sizes:
        b_v += 2 * pre * (z * 8 + 3 * n * 4) + 3 * n * 4 + 6 * n * 4
    b_uz = 2 * (4 * n_v) * 12 + 7 * n_v * 8 + 12 * n_p * 8
    return rec["inner"] * (b_fp64 + b_v) + rec["uzawa"] * b_uz


def pmc_traffic(kernel, recs):
    """HBM bytes per launch of `kernel`: the committed rocprofv3 PMC summary
    gives the traffic per inner iteration of a profiled launch (FETCH_SIZE +
    WRITE_SIZE scaled by the copy-kernel factors); it is scaled to this run's
    mean inner iterations per launch.  None without a summary."""
    path = os.path.join(REPO, "profiles", f"{kernel}_256_pmc.txt")
    if not os.path.exists(path) or not recs:
        return None
    per_inner = None
    for line in open(path):
        if line.startswith("traffic_bytes_per_inner_iteration:"):
            per_inner = float(line.split(":", 1)[1])
    if per_inner is None:
        return None
    return per_inner * sum(r["inner"] for r in recs) / len(recs)


# ---------------------------------------------------------------- paths

def run_path(path, steps, warmup, log):
    from mantle_convection_lab_amd import Input, Model
    net = None
    note = ""
    if path == "learned":
        from mantle_convection_lab_amd import learn
        wp = learn.default_weights_path(SIDE)
        if os.path.exists(wp):
            net = learn.Net(SIDE, SIDE, 1)
            net.load(wp)
        else:
            note = f"no trained weights at {os.path.relpath(wp, REPO)}: previous-step guess only"
            path = "mixed"
    # warmup on a throwaway instance of the same workload (code paths, caches,
    # allocations); the timed run then starts from step 0 on a fresh instance,
    # so every path is timed over the same time steps
    warm = []
    if warmup > 0:
        tw = Model(Input.from_keys(path_keys(path)), device=0, net=net)
        for i in range(warmup):
            r = tw.step()
            warm.append(r)
            log(f"[bench] {path} warmup {i}: uzawa={r['uzawa']} inner={r['inner']} solve={r['solve_ms']:.1f} ms")
        tw.ctx.close()
        del tw
    model = Model(Input.from_keys(path_keys(path)), device=0, net=net)
    model.ctx.synchronize()
    t0 = time.perf_counter()
    recs = []
    for i in range(steps):
        r = model.step()
        recs.append(r)
        g = f" guess={'net' if r.get('guess_accepted') else 'prev'}" if "guess_accepted" in r else \
            (" guess=skip" if net is not None else "")
        log(f"[bench] {path} step {i}: uzawa={r['uzawa']} conv={r['converged']} inner={r['inner']} "
            f"solve={r['solve_ms']:.2f} ms res={r['res_mom']:.1e}/{r['res_cont']:.1e}{g}")
    model.ctx.synchronize()
    wall = time.perf_counter() - t0
    sizes = level_sizes(model)
    vr, nu = model.diagnostics()
    neqT = model.mesh.level_struct(model.mesh.nlev - 1).neqT
    return dict(path=path, note=note, recs=recs, warm=warm, wall=wall, sizes=sizes, neqT=neqT, vrms=vr, nu_top=nu,
                model=model)


def summary(res, steps):
    recs = res["recs"]
    conv = sum(1 for r in recs if r["converged"])
    solve_ms = [r["solve_ms"] for r in recs]
    nb = res["sizes"][-1]
    return dict(solves_per_s=conv / res["wall"] if res["wall"] > 0 else 0.0, converged_solves=conv,
                steps=steps, wall_s=res["wall"], median_solve_ms=statistics.median(solve_ms) if solve_ms else None,
                uzawa_cycles=[r["uzawa"] for r in recs], inner_iterations=[r["inner"] for r in recs],
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
    traffic = pmc_traffic(kernel, recs)
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
            "avg_launch_ms": sum(ms) / len(ms), "bytes_per_launch": sum(bytes_per) / len(bytes_per)}


# ---------------------------------------------------------------- CPU baseline

def cpu_baseline(full_cycles, sample_cycles=24):
    """The FP64 oracle (one core) on the first solve of the same workload,
    bounded to `sample_cycles` Uzawa iterations and scaled to `full_cycles`
    (the plain FP64 GPU path's count for that solve): assembly time + time
    per Uzawa iteration x full_cycles."""
    import numpy as np
    from mantle_convection_lab_amd import Input, Mesh
    from tests.oracle_lib import Ctrl, MeshArrays, Oracle
    keys = path_keys("fp64")
    inp = Input.from_keys(keys)
    mesh = Mesh(inp)
    lv = mesh.level(-1)
    m = MeshArrays(dict(L=lv["L"], lev=lv["lev"], nel=lv["nel"], nno=lv["nno"], neq=lv["neq"], neqT=lv["neqT"],
                        gnox=lv["gnox"], gnoz=lv["gnoz"], Mcode=lv["Mcode"], ien=lv["ien"], node=lv["node"],
                        id=lv["id"], lm=lv["lm"], heat_id=lv["heat_id"], heat_lm=lv["heat_lm"], v_bc=lv["v_bc"],
                        v_bc_val=lv["v_bc_val"], T_bc=lv["T_bc"], T_bc_val=lv["T_bc_val"], X=lv["X"],
                        dx=lv["dx"], dz=lv["dz"], nh=0, h_ni=np.zeros(1, np.int32), h_ci=np.zeros(2, np.int32)))
    c = inp.control()
    o = Oracle(m, Ctrl(c.Ra, c.A, c.refT, c.lm_jump, c.accuracy, c.citcom_accu, c.theta, c.Q, sample_cycles,
                       c.max_cg_cycle, c.cg_precondition, c.check_v_only, c.solver, c.auto_accuracy,
                       keys["schur_mode"], c.mg_cycle, c.mg_finest_iteration, c.mg_coarse_iteration))
    T0 = mesh.initial_temperature(-1)
    t = time.perf_counter()
    A = o.assemble(o.viscosity(T0), o.buoyancy(T0))
    t_asm = time.perf_counter() - t
    t = time.perf_counter()
    d, p, st = o.solve(A)
    t_solve = time.perf_counter() - t
    per_cycle = t_solve / max(st.uzawa_cycles, 1)
    est = t_asm + per_cycle * full_cycles
    return dict(value=1.0 / est, unit="Stokes solves/s", cores=1, kind="port",
                sample=f"oracle FP64 (C restatement, one core) on the first 256x256 solve: assembly "
                       f"{t_asm:.2f} s + {st.uzawa_cycles} Uzawa iterations ({st.inner_iterations} inner PCG "
                       f"iterations) in {t_solve:.1f} s, scaled to the {full_cycles} Uzawa iterations the "
                       f"plain FP64 path needs for that solve: {est:.0f} s per solve")


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
    ap.add_argument("--fp64-steps", type=int, default=None,
                    help="timed steps of the plain FP64 comparison (default: --steps, after the same warmup)")
    ap.add_argument("--path", default="learned", choices=["learned", "mixed", "fp64"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
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

    log(f"[bench] {SIDE}x{SIDE}, primary path {args.path}")
    prim = run_path(args.path, args.steps, args.warmup, log)
    ps = summary(prim, args.steps)
    kernel = "k_uzawa_fp64" if prim["path"] == "fp64" else "k_uzawa_mixed"
    rl = roofline(prim, kernel)
    fp = None
    first_cycles = None
    fp_steps = args.steps if args.fp64_steps is None else args.fp64_steps
    if args.path != "fp64" and fp_steps > 0:
        prim["model"] = None
        fpr = run_path("fp64", fp_steps, args.warmup, log)
        fp = summary(fpr, fp_steps)
        fp["roofline"] = roofline(fpr, "k_uzawa_fp64")
        first_cycles = fpr["recs"][0]["uzawa"] if fpr["recs"] else None
        fpr["model"] = None
    elif args.path == "fp64":
        fp = ps
        first_cycles = prim["recs"][0]["uzawa"] if prim["recs"] else None
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
        "dtype": "f64" if prim["path"] == "fp64" else "f64+f32",
        "data": "synthetic (seeded initial temperature; network weights trained on solver snapshots)",
        "config": {"workload": f"{SIDE}x{SIDE} box, Ra=1e6, eta=exp(-ln(1e3) T), Uzawa dv,dp<1e-6, "
                               f"path={prim['path']}",
                   "side": SIDE, "Ra": 1e6, "viscosity_contrast": 1e3, "path": prim["path"],
                   "note": prim["note"], "primary": ps, "fp64_plain": fp,
                   "speedup_vs_fp64_plain": (ps["solves_per_s"] / fp["solves_per_s"])
                   if fp and fp["solves_per_s"] > 0 and fp is not ps else None},
        "roofline": rl,
    }
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle, 1 core, bounded sample)...")
        out["cpu_baseline"] = cpu_baseline(first_cycles or 338)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
