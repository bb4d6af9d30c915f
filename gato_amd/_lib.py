"""ctypes binding of libgato_mi355x.so (the C-ABI declared in include/gato_mi355x.h).

The library is built in-tree by ``gato_amd.build.build()`` (or ``__graft_entry__.build()``).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libgato_mi355x.so"

PLANT_IDS = {"iiwa14": 0, "indy7": 1}
# plant models of the simulator: the solvers' two and indy7 with its URDF centres of mass (GATO_PLANT_INDY7_URDF)
SIM_MODELS = {"iiwa14": 0, "indy7": 1, "indy7_urdf": 2}
MODE_PLAIN, MODE_FUSED, MODE_GRAPH = 0, 1, 2
FORCE_EST_STATE_LEN = 30  # estimate[6] momentum[6] smoothed[6] rotation[9] radius confidence updates
EXIT_PCG0, EXIT_KKT = 0, 1


class GatoParams(C.Structure):
    _fields_ = [("dt", C.c_float), ("max_sqp_iters", C.c_uint32), ("kkt_tol", C.c_float),
                ("max_pcg_iters", C.c_uint32), ("pcg_tol", C.c_float), ("solve_ratio", C.c_float),
                ("mu", C.c_float), ("q_cost", C.c_float), ("qd_cost", C.c_float), ("u_cost", C.c_float),
                ("N_cost", C.c_float), ("q_lim_cost", C.c_float), ("vel_lim_cost", C.c_float),
                ("ctrl_lim_cost", C.c_float), ("rho", C.c_float), ("mode", C.c_int32), ("exit_rule", C.c_int32),
                ("terminal_mode", C.c_int32), ("hess_mode", C.c_int32)]


_ip = C.POINTER(C.c_int32)
_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)


class GatoStats(C.Structure):
    _fields_ = [("solve_time_us", C.c_double), ("n_iters", C.c_int32), ("ls_num_iters", C.c_int32),
                ("sqp_iters", _ip), ("kkt_converged", _ip), ("initial_merit", _fp), ("final_merit", _fp),
                ("pcg_iters", _ip), ("ls_min_merit", _fp), ("ls_step_size", _fp), ("kkt_residual", _fp),
                ("status", _ip)]


# Per-problem status bits (gato_stats.status)
STATUS_KKT_CONVERGED, STATUS_ITER_CAP, STATUS_PCG_AT_CAP, STATUS_LS_FAILED, STATUS_NONFINITE = 1, 2, 4, 8, 16


# Every symbol declared in include/gato_mi355x.h with its ctypes signature.
_VP = C.c_void_p
SIGNATURES = {
    "gato_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.c_int32, C.c_int32, C.POINTER(GatoParams)]),
    "gato_destroy": (C.c_int, [_VP]),
    "gato_solve": (C.c_int, [_VP, _VP, _VP, _VP, C.c_float, _VP, C.POINTER(GatoStats)]),
    "gato_solve_noalloc": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.POINTER(GatoParams), _VP, _VP, _VP, C.c_float,
                                     _VP, C.POINTER(GatoStats)]),
    "gato_set_joint_cost": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "gato_create_f64": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.c_int32, C.c_int32, C.POINTER(GatoParams)]),
    "gato_destroy_f64": (C.c_int, [_VP]),
    "gato_solve_f64": (C.c_int, [_VP, _VP, _VP, _VP, C.c_double, _VP, C.POINTER(GatoStats)]),
    "gato_last_merits_f64": (C.c_int, [_VP, _dp, _dp]),
    "gato_set_joint_cost_f64": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "gato_set_f_ext_batch_f64": (C.c_int, [_VP, _VP, _VP]),
    "gato_set_batch_f64": (C.c_int, [_VP, C.c_int32, _dp, C.c_int32]),
    "gato_reset_f64": (C.c_int, [_VP, C.c_int32]),
    "gato_set_rho_adaptation_f64": (C.c_int, [_VP, C.c_int32]),
    "gato_get_state_f64": (C.c_int, [_VP, _VP, _dp, _dp]),
    "gato_sim_forward_f64": (C.c_int, [_VP, _VP, _VP, _VP, C.c_double, _VP]),
    "gato_set_f_ext_batch": (C.c_int, [_VP, _VP, _VP]),
    "gato_set_rho_penalty_batch": (C.c_int, [_VP, _fp, C.c_int32]),
    "gato_set_drho_batch": (C.c_int, [_VP, _fp, C.c_int32]),
    "gato_set_mu_batch": (C.c_int, [_VP, _fp]),
    "gato_set_pcg_tol_batch": (C.c_int, [_VP, _fp]),
    "gato_reset_dual": (C.c_int, [_VP]),
    "gato_reset_rho": (C.c_int, [_VP]),
    "gato_set_rho_adaptation": (C.c_int, [_VP, C.c_int32]),
    "gato_set_mode": (C.c_int, [_VP, C.c_int32, C.c_int32]),
    "gato_get_state": (C.c_int, [_VP, _VP, _fp, _fp]),
    "gato_set_state": (C.c_int, [_VP, _VP, _fp, _fp]),
    "gato_sim_forward": (C.c_int, [_VP, _VP, _VP, _VP, C.c_float, _VP]),
    "gato_plant_dims": (C.c_int, [C.c_int32, _ip, _ip]),
    "gato_last_kernel_time": (C.c_int, [_VP, C.POINTER(C.c_double), _ip]),
    "gato_backward": (C.c_int, [_VP, _VP, _VP, _VP, C.c_float, _VP, _VP, _VP, _VP, _VP, _VP, _ip]),
    "gato_set_backward_params": (C.c_int, [_VP, C.c_int32, C.c_float, C.c_int32]),
    "gato_get_backward_params": (C.c_int, [_VP, _ip, C.POINTER(C.c_float), _ip]),
    "gato_set_backward_implicit": (C.c_int, [_VP, C.c_int32, C.c_float]),
    "gato_ee_positions": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "gato_select_hypothesis": (C.c_int, [_VP, _VP, _VP, _VP, C.c_float, _VP, _VP, _ip, _VP]),
    "gato_simulate_rk4": (C.c_int, [_VP, _VP, _VP, _VP, C.c_float, C.c_int32, _VP]),
    "gato_window_reference": (C.c_int, [_VP, _VP, _VP, C.c_int32, C.c_int32, _VP]),
    "gato_simulate_plant": (C.c_int, [C.c_int32, C.c_int32, _VP, _VP, C.c_int32, _VP, C.c_double, C.c_int32, C.c_int32,
                                      C.c_double, _VP]),
    "gato_figure8_path": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_double, C.c_double, _dp, C.c_double, _VP]),
    "gato_force_est_create": (C.c_int, [C.POINTER(_VP), C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float,
                                        C.c_uint64]),
    "gato_force_est_destroy": (C.c_int, [_VP]),
    "gato_force_est_reset": (C.c_int, [_VP, C.c_float, _VP]),
    "gato_force_est_generate": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "gato_force_est_update": (C.c_int, [_VP, _VP, _VP, _VP, C.c_float, C.c_float, _dp, _VP]),
    "gato_force_est_state": (C.c_int, [_VP, _dp, _VP]),
    "gato_set_launch_shape": (C.c_int, [_VP, C.c_int32]),
    "gato_get_launch_shape": (C.c_int, [_VP, _ip, _ip]),
    "gato_autotune": (C.c_int, [_VP, _VP, _VP, _VP, C.c_float, _VP, _ip]),
    "gato_mixed_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, _ip, _ip, _ip, C.POINTER(GatoParams)]),
    "gato_mixed_destroy": (C.c_int, [_VP]),
    "gato_mixed_class": (C.c_int, [_VP, C.c_int32, C.POINTER(C.c_void_p)]),
    "gato_mixed_set_autotune": (C.c_int, [_VP, C.c_int32]),
    "gato_mixed_set_mode": (C.c_int, [_VP, C.c_int32, C.c_int32]),
    "gato_mixed_get_mode": (C.c_int, [_VP, _ip, _ip]),
    "gato_mixed_solve": (C.c_int, [_VP, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_float,
                                   _VP, C.POINTER(GatoStats)]),
    "gato_mixed_last_time": (C.c_int, [_VP, C.POINTER(C.c_double)]),
    "gato_debug_phase": (C.c_int, [_VP, C.c_int32, _VP, _VP, _VP, C.c_float, _VP]),
    "gato_debug_copy": (C.c_int, [_VP, C.c_char_p, _VP, C.c_int64, C.POINTER(C.c_int64)]),
    "gato_last_error": (C.c_char_p, []),
}

_lib = None


class GatoError(RuntimeError):
    pass


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise GatoError(f"{LIB_PATH} is missing: build it with gato_amd.build.build() (hipcc, gfx950)")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc, what):
    if rc != 0:
        msg = load().gato_last_error().decode(errors="replace")
        raise GatoError(f"{what} failed ({rc}): {msg}")


def plant_dims(plant):
    nq, nx = C.c_int32(), C.c_int32()
    check(load().gato_plant_dims(PLANT_IDS[plant], C.byref(nq), C.byref(nx)), "gato_plant_dims")
    return nq.value, nx.value
