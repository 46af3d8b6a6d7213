"""Build the native libraries of mantle_convection_lab_amd in place.

    python -m mantle_convection_lab_amd.build

libmcl_hip.so   hipcc, gfx950: the hot-path kernels and the C-ABI of include/mcl.h
libmcl_host.so  gcc: the C host layer (input parser, quadtree mesh, outputs)
mcl_driver      gcc: the C driver executable (the reference's Main.c loop)
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(REPO, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MCL_OFFLOAD_ARCH", "gfx950")

# The packed-FP32 target feature is disabled for all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value",
             "-Wno-unused-result", "-I" + INC]

# (source, extra flags): -ffp-contract=off keeps the FP64 parity kernels free of
# FMA contraction, so they round like the reference's x86-64 build.
HIP_UNITS = [("mcl_fp64.hip", ["-ffp-contract=off"]), ("mcl_mixed.hip", []), ("mcl_unet.hip", []),
             ("mcl_net.cpp", ["-x", "hip"]), ("mcl_api.cpp", ["-x", "hip"])]


def run(cmd, quiet=False):
    if not quiet:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip():
    """Compiles the units side by side (one hipcc each), then links."""
    from concurrent.futures import ThreadPoolExecutor
    out = os.path.join(PKG, "libmcl_hip.so")
    jobs = []
    for src, extra in HIP_UNITS:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        jobs.append(([HIPCC] + HIP_FLAGS + extra + ["-c", path, "-o", obj], obj))
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(workers) as ex:
        list(ex.map(lambda j: run(j[0]), jobs))
    run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", out] + [o for _, o in jobs])
    return out


def build_host():
    out = os.path.join(PKG, "libmcl_host.so")
    src = os.path.join(CSRC, "mcl_host.c")
    run(["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-I" + INC, "-o", out, src, "-lm"])
    return out


def build_driver():
    drv = os.path.join(CSRC, "mcl_driver.c")
    if not os.path.exists(drv):
        return None
    out = os.path.join(PKG, "mcl_driver")
    run(["gcc", "-O2", "-Wall", "-I" + INC, "-o", out, drv, os.path.join(CSRC, "mcl_host.c"), "-L" + PKG,
         "-lmcl_hip", "-Wl,-rpath,$ORIGIN", "-lm"])
    return out


def build_all():
    build_host()
    build_hip()
    build_driver()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
