"""The C-ABI library loads and exports every entry point include/mcl.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(REPO, "include", "mcl.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mcl_[a-z_A-Z]+)\s*\(", src)))


def test_all_declared_symbols_exported():
    from mantle_convection_lab_amd import hip_lib
    lib = hip_lib()
    names = _declared()
    assert len(names) >= 17
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.mcl_abi_version() == 1


def test_create_without_device_fails_cleanly():
    """mcl_create reports a device error (no GPU here, or an invalid index)."""
    from mantle_convection_lab_amd import hip_lib
    lib = hip_lib()
    h = ctypes.c_void_p()
    rc = lib.mcl_create(999, ctypes.byref(h))
    assert rc != 0 and not h.value


def test_null_arguments_rejected():
    from mantle_convection_lab_amd import hip_lib
    lib = hip_lib()
    assert lib.mcl_create(0, None) != 0
    assert lib.mcl_mesh_upload(None, None, 0) != 0
    assert lib.mcl_solve_velocity_pressure(None, None) != 0


def test_binding_struct_layouts_match_c():
    """Every ctypes mirror of a C struct has the C struct's size (a field
    added on one side only would shift the others and corrupt memory)."""
    import ctypes as C
    from mantle_convection_lab_amd import _native as N
    L = N.host_lib()
    L.mclh_sizeof.argtypes = [C.c_int]
    L.mclh_sizeof.restype = C.c_long
    for which, st in enumerate([N.HostInput, N.Control, N.MeshView, N.StokesStats, N.HeatStats, N.HostLevel,
                                N.HostMesh]):
        assert C.sizeof(st) == L.mclh_sizeof(which), st.__name__
