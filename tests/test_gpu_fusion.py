"""The mixed solve's V-cycle variants compute the same numbers (MI355X).

The fused phases of the grid levels repeat, on each workgroup's halo rows,
exactly the operations the rows' owners perform, in the same order; the
point-to-point waits and the fence-free grid barriers only change when data
becomes visible, not what is computed.  So the solution must be bit-identical
between the two-sweep fused V-cycle (default), the one-sweep fused form
(MCL_MG_FUSE=1), the unfused sequence (MCL_MG_FUSE=0) and grid barriers in
place of every neighbour wait (MCL_NBR_SYNC=0), over several warm-started
steps of the 64x64 variable-viscosity case (64 workgroups of 128 rows; every
halo fits the LDS buffer, and the coarse grid levels run the one-row-per-
thread two-sweep forms), and over the cold first solve at 256x256 (256
workgroups; the finest level runs the one-sweep fused phases, 128x128 the
general two-sweep phases, 64x64 and 32x32 the one-row forms).
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = dict(level=7, mg_min_level=2, nprox=1, nproz=1, model_length=1.0, model_height=1.0, Rayleigh=1e6,
            A=math.log(1000.0), refT=0.0, T_ini=4, T_mantle_perturb=0.01, Q=0.0, theta=0.5, develop_mode=0,
            solver="cg", cg_precondition=1, accuracy=1e-8, citcom_solver_accu=1e-6, max_citcom_cycle=500,
            max_cg_cycle=200, schur_mode=1, top_v_bc="SV", bot_v_bc="SV", lef_v_bc="VS", rig_v_bc="VS",
            write_output=0, precision=1, warm_start=1)
ENV = ("MCL_MG_FUSE", "MCL_NBR_SYNC")


def _run(env, steps=3, keys=KEYS):
    from mantle_convection_lab_amd import Input, Model
    saved = {k: os.environ.get(k) for k in ENV}
    try:
        for k in ENV:
            os.environ.pop(k, None)
        os.environ.update(env)
        m = Model(Input.from_keys(keys))
        recs = [m.step() for _ in range(steps)]
        out = (m.field("D").copy(), m.field("P").copy(), m.field("T").copy(),
               [(r["uzawa"], r["inner"]) for r in recs])
        m.ctx.close()
        return out
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def test_fused_unfused_and_barrier_variants_bit_identical():
    print("[gpu-fusion] 64x64, 3 steps per variant", flush=True)
    base = _run({})
    print(f"  default: (uzawa, inner) per step {base[3]}", flush=True)
    for env in ({"MCL_MG_FUSE": "1"}, {"MCL_MG_FUSE": "0"}, {"MCL_NBR_SYNC": "0"}):
        got = _run(env)
        print(f"  {env}: {got[3]}", flush=True)
        assert got[3] == base[3]
        for a, b in zip(got[:3], base[:3]):
            assert np.array_equal(a, b)


def test_variants_bit_identical_at_256():
    keys = dict(KEYS, level=9)
    print("[gpu-fusion] 256x256, first solve per variant", flush=True)
    base = _run({}, steps=1, keys=keys)
    print(f"  default: (uzawa, inner) {base[3]}", flush=True)
    for env in ({"MCL_MG_FUSE": "1"}, {"MCL_NBR_SYNC": "0"}):
        got = _run(env, steps=1, keys=keys)
        print(f"  {env}: {got[3]}", flush=True)
        assert got[3] == base[3]
        for a, b in zip(got[:3], base[:3]):
            assert np.array_equal(a, b)
