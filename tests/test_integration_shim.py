"""The C shim of INTEGRATION.md compiles against the reference's own headers
(global_variables.h, function_def.h) and the library's include/mcl.h, with
the reference's flags (gcc -O2 -fcommon, the image's MPICH headers).  Skipped
where the reference sources or the MPI headers are absent."""
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
MPI_INC = "/opt/conda/include"


def _shim_source():
    text = open(os.path.join(REPO, "INTEGRATION.md")).read()
    sec = text.split("## C shim a maintainer adds to the reference", 1)[1]
    return re.search(r"```c\n(.*?)```", sec, re.S).group(1)


@pytest.mark.skipif(not (os.path.exists(os.path.join(REF, "global_variables.h")) and
                         os.path.exists(os.path.join(MPI_INC, "mpi.h"))),
                    reason="reference headers or MPI headers not present")
def test_shim_compiles_against_reference_headers(tmp_path):
    src = tmp_path / "Mcl_bridge.c"
    src.write_text(_shim_source())
    r = subprocess.run(["gcc", "-O2", "-fcommon", "-Wall", "-Werror=implicit-function-declaration",
                        "-Werror=incompatible-pointer-types", "-Werror=int-conversion", "-I" + REF, "-I" + MPI_INC,
                        "-I" + os.path.join(REPO, "include"), "-c", str(src), "-o", str(tmp_path / "b.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # every hot-path function of the reference's Main.c loop is defined by the shim
    r = subprocess.run(["nm", str(tmp_path / "b.o")], capture_output=True, text=True, check=True)
    defined = {line.split()[-1] for line in r.stdout.splitlines() if " T " in line}
    for f in ("buoyancy_force", "viscosity", "velocity_pressure_KGF_matrix", "solve_velocity_pressure",
              "update_temperature", "mcl_bridge_setup", "mcl_bridge_set_T"):
        assert f in defined, f
