"""ctypes bindings of the native libraries (built in place by build.py).

libmcl_hip.so   the C-ABI of include/mcl.h (HIP kernels, gfx950)
libmcl_host.so  the C host layer (mcl_host.h): input format, mesh, initial T
Loading libmcl_hip.so fails loudly if it has not been built: there is no
fallback implementation of the hot path.
"""
import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))

PI = C.POINTER(C.c_int)
PD = C.POINTER(C.c_double)
PB = C.POINTER(C.c_byte)

MCL_SOLVER = {"jc": 0, "gs": 1, "cg": 2, "mg": 3}
FIELD = dict(T=0, VIS=1, BUOYANCY=2, V=3, D=4, P=5, F=6, KDIAG=7, STRESS=8, HEAT_D=9, HEAT_DDOT=10,
             HEAT_F=11, TAU=12)
MATRIX = dict(K=0, G=1, HEAT_K=2, HEAT_C=3)


class MeshView(C.Structure):
    _fields_ = [("L", C.c_int), ("lev", C.c_int), ("nel", C.c_int), ("nno", C.c_int), ("neq", C.c_int),
                ("neqT", C.c_int), ("gnox", C.c_int), ("gnoz", C.c_int),
                ("Mcode", PI), ("ien", PI), ("node", PI), ("id", PI), ("lm", PI), ("heat_id", PI),
                ("heat_lm", PI), ("v_bc", PB), ("v_bc_val", PD), ("T_bc", PB), ("T_bc_val", PD), ("X", PD),
                ("dx", PD), ("dz", PD), ("nh", C.c_int), ("h_ni", PI), ("h_ci", PI),
                ("interp_n", PI), ("interp_eq", PI), ("nproj", C.c_int), ("proj_l", PI), ("proj_u", PI)]


class Control(C.Structure):
    _fields_ = [("Ra", C.c_double), ("A", C.c_double), ("refT", C.c_double), ("lm_jump", C.c_double),
                ("accuracy", C.c_double), ("citcom_accu", C.c_double), ("theta", C.c_double), ("Q", C.c_double),
                ("max_citcom_cycle", C.c_int), ("max_cg_cycle", C.c_int), ("cg_precondition", C.c_int),
                ("check_v_only", C.c_int), ("solver", C.c_int), ("auto_accuracy", C.c_int), ("zero_P", C.c_int),
                ("AMR", C.c_int), ("mg_cycle", C.c_int), ("mg_finest_iteration", C.c_int),
                ("mg_coarse_iteration", C.c_int), ("schur", C.c_int), ("precision", C.c_int),
                ("warm_start", C.c_int), ("inner_rtol", C.c_double), ("exact_dots", C.c_int)]


class StokesStats(C.Structure):
    _fields_ = [("uzawa_cycles", C.c_int), ("converged", C.c_int), ("dv", C.c_double), ("dp", C.c_double),
                ("inner_iterations", C.c_longlong), ("spmv", C.c_longlong), ("accuracy_used", C.c_double),
                ("res_momentum", C.c_double), ("res_continuity", C.c_double), ("device_ms", C.c_double)]


class HeatStats(C.Structure):
    _fields_ = [("cg_cycles", C.c_int), ("init_cg_cycles", C.c_int), ("dt", C.c_double),
                ("model_time", C.c_double), ("device_ms", C.c_double)]


class HostInput(C.Structure):
    _fields_ = [("data_dir", C.c_char * 256), ("solver", C.c_char * 8), ("max_level", C.c_int),
                ("min_level", C.c_int), ("npx", C.c_int), ("npz", C.c_int), ("length", C.c_double),
                ("height", C.c_double), ("top_v_bc", C.c_char * 2), ("bot_v_bc", C.c_char * 2),
                ("lef_v_bc", C.c_char * 2), ("rig_v_bc", C.c_char * 2), ("top_v_bc_val", C.c_double * 2),
                ("bot_v_bc_val", C.c_double * 2), ("lef_v_bc_val", C.c_double * 2),
                ("rig_v_bc_val", C.c_double * 2), ("top_T_bc", C.c_char), ("bot_T_bc", C.c_char),
                ("lef_T_bc", C.c_char), ("rig_T_bc", C.c_char), ("top_T_bc_val", C.c_double),
                ("bot_T_bc_val", C.c_double), ("lef_T_bc_val", C.c_double), ("rig_T_bc_val", C.c_double),
                ("T_ini", C.c_int), ("T_mantle", C.c_double), ("T_mantle_perturb", C.c_double),
                ("T_wave_number", C.c_int), ("Ra", C.c_double), ("use_hash", C.c_int), ("A", C.c_double),
                ("refT", C.c_double), ("lm_jump", C.c_double), ("highest_vis", C.c_double),
                ("lowest_vis", C.c_double), ("use_dense_matrix", C.c_int), ("mg_coarse_iteration", C.c_int),
                ("mg_finest_iteration", C.c_int), ("max_timestep", C.c_int), ("save_spacing", C.c_int),
                ("Q", C.c_double), ("theta", C.c_double), ("AMR", C.c_int), ("mpi_KF_method", C.c_int),
                ("cg_precondition", C.c_int), ("Tcg_precondition", C.c_int), ("accuracy", C.c_double),
                ("auto_accuracy", C.c_int), ("citcom_accu", C.c_double), ("dev_mode", C.c_int),
                ("check_v_only", C.c_int), ("max_citcom_cycle", C.c_int), ("mg_cycle", C.c_int),
                ("max_cg_cycle", C.c_int), ("old_T_file", C.c_char * 1000), ("old_T_level", C.c_int),
                ("old_T_timestep", C.c_int), ("lg_mul", C.c_int), ("grain_damage", C.c_int),
                ("show_convg", C.c_int), ("zero_P", C.c_int), ("schur", C.c_int), ("precision", C.c_int),
                ("warm_start", C.c_int), ("inner_rtol", C.c_double), ("exact_dots", C.c_int),
                ("T_seed", C.c_ulonglong),
                ("device", C.c_int), ("write_output", C.c_int)]


class HostLevel(C.Structure):
    _fields_ = [("L", C.c_int), ("lev", C.c_int), ("nel", C.c_int), ("nno", C.c_int), ("neq", C.c_int),
                ("neqT", C.c_int), ("gnox", C.c_int), ("gnoz", C.c_int),
                ("Mcode", PI), ("ien", PI), ("node", PI), ("id", PI), ("lm", PI), ("heat_id", PI),
                ("heat_lm", PI), ("veq_node", PI), ("veq_dof", PI), ("heat_eq_node", PI),
                ("v_bc", PB), ("v_bc_val", PD), ("T_bc", PB), ("T_bc_val", PD), ("X", PD), ("area", PD),
                ("p2n_nel", PI), ("p2n_el", PI), ("nh", C.c_int), ("h_node", PI), ("h_ni", PI),
                ("h_cnode", PI), ("h_ci", PI), ("interp_n", PI), ("interp_eq", PI), ("nproj", C.c_int),
                ("proj_l", PI), ("proj_u", PI)]


class HostMesh(C.Structure):
    _fields_ = [("L", C.c_int), ("min_level", C.c_int), ("nlev", C.c_int), ("gnox", C.c_int),
                ("gnoz", C.c_int), ("dx", C.c_double * 32), ("dz", C.c_double * 32),
                ("lev", C.POINTER(HostLevel))]


_hip = None
_host = None


class NativeMissing(RuntimeError):
    pass


def hip_lib():
    """libmcl_hip.so; raises NativeMissing if it has not been built."""
    global _hip
    if _hip is None:
        path = os.path.join(PKG, "libmcl_hip.so")
        if not os.path.exists(path):
            raise NativeMissing(f"{path} is missing: run `python -m mantle_convection_lab_amd.build` "
                                "(there is no CPU fallback for the hot path)")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.mcl_abi_version.restype = C.c_int
        L.mcl_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.mcl_destroy.argtypes = [vp]
        L.mcl_destroy.restype = None
        L.mcl_last_error.argtypes = [vp]
        L.mcl_last_error.restype = C.c_char_p
        L.mcl_mesh_upload.argtypes = [vp, C.POINTER(MeshView), C.c_int]
        L.mcl_set_control.argtypes = [vp, C.POINTER(Control)]
        L.mcl_set_field.argtypes = [vp, C.c_int, C.c_int, PD, C.c_size_t]
        L.mcl_get_field.argtypes = [vp, C.c_int, C.c_int, PD, C.c_size_t]
        L.mcl_get_matrix.argtypes = [vp, C.c_int, C.c_int, C.c_int, PI, PI, PD]
        L.mcl_set_model_time.argtypes = [vp, C.c_double]
        L.mcl_get_model_time.argtypes = [vp, PD]
        for n in ("mcl_buoyancy_force", "mcl_viscosity", "mcl_velocity_pressure_KGF_matrix", "mcl_synchronize"):
            getattr(L, n).argtypes = [vp]
        L.mcl_solve_velocity_pressure.argtypes = [vp, C.POINTER(StokesStats)]
        L.mcl_solve_velocity_pressure_noalloc.argtypes = [vp, C.POINTER(StokesStats)]
        L.mcl_update_temperature.argtypes = [vp, C.POINTER(HeatStats)]
        L.mcl_set_launch_cap.argtypes = [vp, C.c_int]
        L.mcl_batch_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.POINTER(vp)]
        L.mcl_batch_destroy.argtypes = [vp]
        L.mcl_batch_destroy.restype = None
        L.mcl_batch_last_error.argtypes = [vp]
        L.mcl_batch_last_error.restype = C.c_char_p
        L.mcl_batch_step.argtypes = [vp, PI, C.POINTER(StokesStats), C.POINTER(HeatStats), PD]
        _hip = L
    return _hip


def host_lib():
    global _host
    if _host is None:
        path = os.path.join(PKG, "libmcl_host.so")
        if not os.path.exists(path):
            raise NativeMissing(f"{path} is missing: run `python -m mantle_convection_lab_amd.build`")
        L = C.CDLL(path)
        L.mclh_read_input.argtypes = [C.c_char_p, C.POINTER(HostInput), C.c_char_p, C.c_int]
        L.mclh_default_input.argtypes = [C.POINTER(HostInput)]
        L.mclh_default_input.restype = None
        L.mclh_set_key.argtypes = [C.POINTER(HostInput), C.c_char_p, C.c_char_p]
        L.mclh_build_mesh.argtypes = [C.POINTER(HostInput), C.POINTER(C.POINTER(HostMesh)), C.c_char_p, C.c_int]
        L.mclh_free_mesh.argtypes = [C.POINTER(HostMesh)]
        L.mclh_free_mesh.restype = None
        L.mclh_build_mesh_depth.argtypes = [C.POINTER(HostInput), C.c_int, C.POINTER(C.POINTER(HostMesh)),
                                            C.c_char_p, C.c_int]
        L.mclh_adapt_mesh.argtypes = [C.POINTER(HostInput), C.POINTER(HostMesh), PD, C.c_int, C.c_int, C.c_double,
                                      C.c_double, C.POINTER(C.POINTER(HostMesh)), C.POINTER(PD),
                                      C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_char_p, C.c_int]
        L.mclh_free_array.argtypes = [C.c_void_p]
        L.mclh_free_array.restype = None
        L.mclh_initial_temperature.argtypes = [C.POINTER(HostInput), C.POINTER(HostMesh), C.c_int, PD,
                                               C.c_char_p, C.c_int]
        L.mclh_fill_view.argtypes = [C.POINTER(HostMesh), C.c_int, C.POINTER(MeshView)]
        L.mclh_fill_view.restype = None
        L.mclh_fill_control.argtypes = [C.POINTER(HostInput), C.POINTER(Control)]
        L.mclh_fill_control.restype = None
        L.mclh_seeded_unit.argtypes = [C.c_ulonglong, C.c_longlong]
        L.mclh_seeded_unit.restype = C.c_double
        L.mclh_write_output.argtypes = [C.POINTER(HostInput), C.POINTER(HostLevel), C.c_int, PD, PD, PD, PD,
                                        C.c_char_p, C.c_int]
        L.mclh_diagnostics.argtypes = [C.POINTER(HostInput), C.POINTER(HostLevel), PD, PD, PD, PD]
        _host = L
    return _host


def declare_net(L):
    """Declarations of the learned-guess entry points of mcl.h."""
    vp = C.c_void_p
    PF = C.POINTER(C.c_float)
    L.mcl_net_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.POINTER(vp)]
    L.mcl_net_destroy.argtypes = [vp]
    L.mcl_net_destroy.restype = None
    L.mcl_net_last_error.argtypes = [vp]
    L.mcl_net_last_error.restype = C.c_char_p
    L.mcl_net_num_params.argtypes = [vp]
    L.mcl_net_num_params.restype = C.c_longlong
    for f in ("mcl_net_get_params", "mcl_net_set_params", "mcl_net_get_grads"):
        getattr(L, f).argtypes = [vp, PF]
    L.mcl_net_forward.argtypes = [vp, PF, PF]
    L.mcl_net_train_step.argtypes = [vp, PF, PF, C.c_float, PF]
    L.mcl_net_capture.argtypes = [vp, vp]
    L.mcl_net_dataset_size.argtypes = [vp]
    L.mcl_net_dataset_size.restype = C.c_longlong
    L.mcl_net_train.argtypes = [vp, C.c_int, C.c_float, C.c_ulonglong, PF]
    L.mcl_stokes_guess.argtypes = [vp, vp, C.POINTER(C.c_int), PD, PD]
    L.mcl_net_get_scale.argtypes = [vp, C.c_int]
    L.mcl_net_get_scale.restype = C.c_double
    L.mcl_net_set_scale.argtypes = [vp, C.c_int, C.c_double]
    return L
