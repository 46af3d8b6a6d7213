"""Generate the golden fixtures under tests/golden/ from the real reference.

Runs only where the reference sources exist (they are compiled, unmodified, by
``oracle/Makefile``'s ``ref`` target into ``oracle/_ref/mcl_ref_harness``).  Each
case writes a reference input file (the reference's own getpar format,
Input.c:3-98), runs the harness, and stores the dumped arrays as a compressed
``.npz`` fixture together with the reference's stdout (Uzawa counts and dv/dp as
printed by Solver.c:224-225, heat CG counts from Update_temperature.c:1254).

    python tests/golden/make_golden.py [case ...]
"""
import os
import re
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(REPO, "oracle", "_ref", "mcl_ref_harness")
sys.path.insert(0, REPO)

from tests.golden.seeded import seeded_temperature  # noqa: E402

# Keys shared by every case: free-slip box, T=1 bottom / T=0 top, insulating
# sides, Jacobi-PCG inner solver (the C1 mapping of the baseline configuration).
BASE_KEYS = dict(
    nprox=1, nproz=1, model_length=1.0, model_height=1.0,
    T_mantle_perturb=0.01, T_wave_number=1, Q=0.0, theta=0.5,
    develop_mode=0, AMR=0, solver='"cg"', cg_precondition=1, Tcg_precondition=1,
    accuracy=1e-8, citcom_solver_accu=1e-6, max_citcom_cycle=2000,
    top_v_bc='"SV"', bot_v_bc='"SV"', lef_v_bc='"VS"', rig_v_bc='"VS"',
    top_T_bc="T", bot_T_bc="T", lef_T_bc="F", rig_T_bc="F",
    top_T_bc_val=0.0, bot_T_bc_val=1.0, lef_T_bc_val=0.0, rig_T_bc_val=0.0,
    use_hash=0, A=0.0, refT=0.5, lower_mantle_jump=1.0,
    save_spacing=100000, data_dir='"."',
)

# name -> (keys, full_steps, snapshot_every, seeded_T[, extras])
CASES = {
    # C1: 64x64, Ra=1e4, isoviscous, T=(1-z)+0.01 cos(pi x) sin(pi z); first steps
    "c1_64": (dict(level=7, mg_min_level=6, Rayleigh=1e4, T_ini=0, max_timestep=2), 2, 1, False),
    # C1 full run (500 Stokes solves), fields every 50 steps
    "c1_64_run": (dict(level=7, mg_min_level=6, Rayleigh=1e4, T_ini=0, max_timestep=499), 0, 50, False),
    # C2-shaped at 32x32: Ra=1e6, eta=exp(-ln(1e3) T), seeded T; Uzawa capped
    "c2s_32": (dict(level=6, mg_min_level=5, Rayleigh=1e6, A=6.907755278982137, refT=0.0,
                    T_ini=-1, max_timestep=1, max_citcom_cycle=40), 2, 0, True),
    # same at 64x64, one step, capped
    "c2s_64": (dict(level=7, mg_min_level=6, Rayleigh=1e6, A=6.907755278982137, refT=0.0,
                    T_ini=-1, max_timestep=0, max_citcom_cycle=25), 1, 0, True),
    # plain (unpreconditioned) CG inner solver and heat solver, 16x16
    "plain_cg_16": (dict(level=5, mg_min_level=4, Rayleigh=1e5, T_ini=0, cg_precondition=0,
                         max_timestep=1, max_citcom_cycle=60), 2, 0, False),
    # refined meshes with hanging nodes (develop_mode recipes, Tree.c:110-275), capped solves
    "dev12_32": (dict(level=6, mg_min_level=5, Rayleigh=1e4, T_ini=0, develop_mode=12,
                      max_timestep=0, max_citcom_cycle=8), 1, 0, False),
    "dev8_16": (dict(level=5, mg_min_level=4, Rayleigh=1e4, T_ini=0, develop_mode=8,
                     max_timestep=0, max_citcom_cycle=6), 1, 0, False),
    "dev11_32": (dict(level=6, mg_min_level=5, Rayleigh=1e4, T_ini=0, develop_mode=11,
                      max_timestep=0, max_citcom_cycle=4), 1, 0, False),
    "dev13_64": (dict(level=7, mg_min_level=6, Rayleigh=1e4, T_ini=0, develop_mode=13,
                      max_timestep=0, max_citcom_cycle=2, max_cg_cycle=10), 1, 0, False),
    "dev14_16": (dict(level=5, mg_min_level=4, Rayleigh=1e4, T_ini=0, develop_mode=14,
                      max_timestep=0, max_citcom_cycle=6), 1, 0, False),
    # multigrid level maps (4 levels), capped solve
    "mg_16": (dict(level=5, mg_min_level=1, Rayleigh=1e4, T_ini=0, solver='"mg"', mg_cycle=2,
                   max_timestep=0, max_citcom_cycle=3), 1, 0, False),
    # the reference's other inner solvers (Solver.c:82-89): Jacobi (its default
    # "jc") and Gauss-Seidel, 1000 sweeps per inner solve, two full steps
    "jc_16": (dict(level=5, mg_min_level=4, Rayleigh=1e4, T_ini=0, solver='"jc"', max_timestep=1,
                   max_citcom_cycle=12), 2, 0, False),
    "gs_16": (dict(level=5, mg_min_level=4, Rayleigh=1e4, T_ini=0, solver='"gs"', max_timestep=1,
                   max_citcom_cycle=12), 2, 0, False),
    # stand-alone multigrid with variable viscosity, three levels; the second
    # step shows the coarse levels keeping the initial temperature (a T_ini=-1
    # restart fills only the finest level, Initial_condition.c:171-195, so the
    # analytic initial temperature is used)
    "mg_32": (dict(level=6, mg_min_level=3, Rayleigh=1e5, A=2.302585092994046, refT=0.5, T_ini=0,
                   solver='"mg"', mg_cycle=2, max_timestep=1, max_citcom_cycle=8), 2, 0, False),
    # auto_accuracy (Solver.c:41-51): inner tolerance from |F|
    "auto_16": (dict(level=5, mg_min_level=4, Rayleigh=1e5, T_ini=0, auto_accuracy=1, accuracy=1e-3,
                     max_timestep=0, max_citcom_cycle=30), 1, 0, False),
    # the reference's output files (Output.c:31-156, written at Main.c:62-63)
    # for steps 0 and 1, and a restart (T_ini=-1, Initial_condition.c:171-195)
    # from the temperature file of step 1
    "out_16": (dict(level=5, mg_min_level=4, Rayleigh=1e4, T_ini=0, max_timestep=1, save_spacing=1,
                    max_citcom_cycle=10), 2, 0, False, dict(write_output=True)),
    "restart_16": (dict(level=5, mg_min_level=4, Rayleigh=1e4, T_ini=-1, max_timestep=0, max_citcom_cycle=10),
                   1, 0, False, dict(restart_from=("out_16", 1))),
    # 256x256 setup only (integer structures); solves cut to one iteration
    "setup_256": (dict(level=9, mg_min_level=8, Rayleigh=1e6, T_ini=0, max_timestep=0,
                       max_citcom_cycle=1, max_cg_cycle=1), 0, 0, False),
}

C1_ONLY_KEEP = ("c1_64_run", "setup_256")


def write_input(path, keys):
    with open(path, "w") as f:
        for k, v in keys.items():
            f.write(f"{k}={v}\n")


def read_dump(path):
    out = {}
    with open(path, "rb") as f:
        while True:
            hdr = f.read(56)
            if len(hdr) < 56:
                break
            name = hdr[:48].split(b"\0", 1)[0].decode()
            kind = chr(hdr[48])
            (count,) = struct.unpack("<q", f.read(8))
            dt = {"d": "<f8", "i": "<i4", "c": "i1"}[kind]
            arr = np.frombuffer(f.read(count * np.dtype(dt).itemsize), dtype=dt).copy()
            out[name] = arr
    return out


def parse_stdout(text):
    uz = [(int(a), float(b), float(c)) for a, b, c in
          re.findall(r"^(\d+) dv = ([-+0-9.eE]+) dp = ([-+0-9.eE]+)$", text, re.M)]
    tcg = [int(x) for x in re.findall(r"^Temperature CG cycle = (\d+)$", text, re.M)]
    return uz, tcg


OUTPUT_FILES = ("coord", "temperature", "stress", "viscosity", "velocity")


def run_case(name):
    keys = dict(BASE_KEYS)
    ck, full_steps, snap, seeded = CASES[name][:4]
    extras = CASES[name][4] if len(CASES[name]) > 4 else {}
    keys.update(ck)
    with tempfile.TemporaryDirectory() as td:
        if "restart_from" in extras:
            src, step = extras["restart_from"]
            with np.load(os.path.join(HERE, src + ".npz")) as z:
                text = bytes(z[f"out.s{step}.temperature"])
            tdir = os.path.join(td, "oldT", "%07d" % step)
            os.makedirs(tdir)
            with open(os.path.join(tdir, "temperature.0.%07d" % step), "wb") as f:
                f.write(text)
            keys["old_T_file"] = '"%s"' % os.path.join(td, "oldT")
            keys["old_T_timestep"] = step
        if seeded:
            level = keys["level"]
            side = 2 ** (level - 1)
            tdir = os.path.join(td, "oldT", "0000000")
            os.makedirs(tdir)
            T = seeded_temperature(side, side)
            with open(os.path.join(tdir, "temperature.0.0000000"), "w") as f:
                for v in T:
                    f.write("%.17g\n" % v)
            keys["old_T_file"] = '"%s"' % os.path.join(td, "oldT")
            keys["old_T_timestep"] = 0
        inp = os.path.join(td, "case.input")
        write_input(inp, keys)
        dump = os.path.join(td, "dump.bin")
        res = subprocess.run([HARNESS, inp, dump, str(full_steps), str(snap),
                              "1" if extras.get("write_output") else "0"], cwd=td,
                             capture_output=True, text=True, timeout=3600)
        if res.returncode != 0:
            raise RuntimeError(f"{name}: harness failed\n{res.stderr[-2000:]}")
        data = read_dump(dump)
        if extras.get("write_output"):
            # the files exactly as the reference wrote them (data_dir ".")
            for step in range(int(keys["max_timestep"]) + 1):
                for f in OUTPUT_FILES:
                    path = os.path.join(td, "%07d" % step, "%s.0.%07d" % (f, step))
                    with open(path, "rb") as fh:
                        data[f"out.s{step}.{f}"] = np.frombuffer(fh.read(), dtype=np.uint8)
    uz, tcg = parse_stdout(res.stdout)
    data["stdout.uzawa_cycles"] = np.array([u[0] for u in uz], dtype=np.int32)
    data["stdout.uzawa_dv"] = np.array([u[1] for u in uz])
    data["stdout.uzawa_dp"] = np.array([u[2] for u in uz])
    data["stdout.heat_cg_cycles"] = np.array(tcg, dtype=np.int32)
    # the input keys, as text, so tests can rebuild the same case
    data["input_keys"] = np.frombuffer(
        "\n".join(f"{k}={v}" for k, v in keys.items() if k not in ("old_T_file",)).encode(), dtype=np.uint8)
    data["seeded_T"] = np.array([1 if seeded else 0], dtype=np.int32)
    if name == "setup_256":
        # integer structures only, as digests (the arrays themselves are large)
        import hashlib
        keep = {}
        for k, v in data.items():
            if k.startswith("L8.") and v.dtype.kind in "iu" or k in ("gnox", "gnoz", "input_keys", "seeded_T"):
                keep[k + ".sha256"] = np.frombuffer(hashlib.sha256(np.ascontiguousarray(v).tobytes()).digest(),
                                                    dtype=np.uint8)
                keep[k + ".len"] = np.array([v.size], dtype=np.int64)
        data = keep
    out = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(out, **data)
    print(f"{name}: {len(data)} arrays, {os.path.getsize(out)} bytes, uzawa={[u[0] for u in uz]}, tcg={tcg[:6]}")


def main(argv):
    if not os.path.exists(HARNESS):
        raise SystemExit("build the reference harness first: make -C oracle ref")
    names = argv or list(CASES)
    for n in names:
        run_case(n)


if __name__ == "__main__":
    main(sys.argv[1:])
