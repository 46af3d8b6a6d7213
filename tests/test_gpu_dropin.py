"""The drop-in itself: the reference's own unmodified Main.c, parser, set-up
and output code, with the C shim of INTEGRATION.md in place of its hot-path
files, linked against libmcl_hip.so (oracle/_ref/mcl_ref_dropin, built by
`make -C oracle dropin` where the reference sources are).  Run on reference
input files on the MI355X, it must print what the unmodified reference printed
for them (fixtures: Uzawa cycle counts and heat CG counts equal, dv and dp
within 1e-6 relative of the printed 10 digits) and, through the reference's
own output(), write the files the reference wrote (every value within the
"%e" format's 7 significant digits).  Cases: the C1 input (64x64, three
steps), the reference's default inner solver (Jacobi) and Gauss-Seidel, the
stand-alone multigrid, and a run that writes output files.

Skipped where the binary was not built (no reference sources) or the image's
MPICH library is missing.
"""
import os
import subprocess

import numpy as np
import pytest

from tests.golden.make_golden import OUTPUT_FILES, parse_stdout
from tests.oracle_lib import fixture_keys, load_fixture
from tests.test_gpu_driver import _write_input

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(REPO, "oracle", "_ref", "mcl_ref_dropin")
MPI_LIB = "/opt/conda/lib/libmpi.so.12"

need = pytest.mark.skipif(not (os.path.exists(DROPIN) and os.path.exists(MPI_LIB)),
                          reason="drop-in binary not built here (needs the reference sources) or MPICH missing")


def _run(name, tmp_path):
    fx = load_fixture(name)
    keys = fixture_keys(fx)
    keys["data_dir"] = str(tmp_path / "out")
    inp = tmp_path / "case.input"
    _write_input(inp, keys)
    print(f"[gpu-dropin] case={name}", flush=True)
    r = subprocess.run([DROPIN, str(inp)], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    print(r.stdout[-1500:], r.stderr[-800:], flush=True)
    assert r.returncode == 0 and "citcom solver" in r.stderr, r.stderr[-2000:]
    uz, tcg = parse_stdout(r.stdout)
    assert [u[0] for u in uz] == list(fx["stdout.uzawa_cycles"])
    assert tcg == list(fx["stdout.heat_cg_cycles"])
    for (n, dv, dp), rdv, rdp in zip(uz, fx["stdout.uzawa_dv"], fx["stdout.uzawa_dp"]):
        assert abs(dv - rdv) <= 1e-6 * abs(rdv) and abs(dp - rdp) <= 1e-6 * abs(rdp)
    return fx


@need
@pytest.mark.parametrize("name", ["c1_64", "jc_16", "gs_16", "mg_32"])
def test_dropin_prints_what_the_reference_printed(name, tmp_path):
    _run(name, tmp_path)


@need
def test_dropin_writes_what_the_reference_wrote(tmp_path):
    fx = _run("out_16", tmp_path)
    for step in (0, 1):
        for f in OUTPUT_FILES:
            mine = (tmp_path / "out" / ("%07d" % step) / ("%s.0.%07d" % (f, step))).read_bytes()
            ref = bytes(fx[f"out.s{step}.{f}"])
            a = np.array([float(x) for x in mine.split()])
            b = np.array([float(x) for x in ref.split()])
            assert a.shape == b.shape and mine.count(b"\n") == ref.count(b"\n"), (step, f)
            scale = np.max(np.abs(b)) if b.size else 1.0
            assert np.all(np.abs(a - b) <= 1e-6 * np.abs(b) + 1e-9 * scale), (step, f)
