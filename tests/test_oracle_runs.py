"""The committed converged oracle run (tests/golden/c2s_64_conv.npz, made by
tests/golden/make_oracle_runs.py) is the oracle's own output: its first step
is recomputed here with the live oracle and must agree bit for bit."""
import numpy as np

from tests.golden.make_oracle_runs import run
from tests.oracle_lib import load_fixture


def test_converged_run_fixture_is_the_oracles():
    ref = load_fixture("c2s_64_conv")
    out = run(steps=1)
    assert out["uzawa"][0] == ref["uzawa"][0]
    assert np.array_equal(out["T.s0"], ref["T.s0"])
    assert out["vrms"][0] == ref["vrms"][0] and out["nu"][0] == ref["nu"][0]
    # every solve of the run converged to the Uzawa tolerance well inside the cap
    assert ref["uzawa"].max() < 5000 and len(ref["vrms"]) == 20
