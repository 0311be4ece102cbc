"""ctypes mirror of include/lbflow_hip.h (the C ABI of the HIP engine).

The descriptor layout here must match the header byte for byte; the CPU test
suite checks every struct size and a few field offsets against a compiled
probe of the header.
"""
import ctypes as C

ABI_VERSION = 1
MAX_Q = 27
MAX_SCALARS = 4

# error codes
OK, EINVAL, EUNSUPPORTED, EDEVICE, ENOMEM, ESTATE, ENONFINITE, ETIMEOUT = 0, -1, -2, -3, -4, -5, -6, -7

# velocity sets: `VelocitySet` discriminants (src/velocity_set.rs:18-25)
D2Q5, D2Q9, D3Q15, D3Q19, D3Q27 = 0, 1, 2, 3, 4
VELOCITY_SET_NAMES = {D2Q5: "D2Q5", D2Q9: "D2Q9", D3Q15: "D3Q15", D3Q19: "D3Q19", D3Q27: "D3Q27"}
Q_OF = {D2Q9: 9, D3Q15: 15, D3Q19: 19, D3Q27: 27}
DIM_OF = {D2Q9: 2, D3Q15: 3, D3Q19: 3, D3Q27: 3}

# collision operators (src/lib.rs:37-42)
BGK, TRT, MRT = 0, 1, 2

# faces (src/lib.rs:17-25), also the order of application
WEST, EAST, SOUTH, NORTH, BOTTOM, TOP = range(6)
FACE_NAMES = ["West", "East", "South", "North", "Bottom", "Top"]

# momentum face conditions (src/momentum/bc.rs:11-26)
BC_UNSET, BC_NO_SLIP, BC_BOUNCE_BACK, BC_ANTI_BOUNCE_BACK, BC_ZOU_HE, BC_PERIODIC = range(6)
KNOWN_DENSITY, KNOWN_UX, KNOWN_UY, KNOWN_UZ = 1, 2, 4, 8

# scalar face conditions (src/passive_scalar/bc.rs:9-30)
(SBC_UNSET, SBC_ANTI_BOUNCE_BACK, SBC_ANTI_BB_NO_FLUX, SBC_BB_NO_FLUX,
 SBC_ZEROTH_ORDER_NO_FLUX, SBC_SECOND_ORDER_NO_FLUX, SBC_PERIODIC) = range(7)
INNER_ANTI_BOUNCE_BACK, INNER_BOUNCE_BACK = 0, 1

# storage precision
F64, F32, F32_SHIFTED, F16_SHIFTED, BF16_SHIFTED = range(5)
PRECISION_NAMES = {F64: "f64", F32: "f32", F32_SHIFTED: "f32s", F16_SHIFTED: "f16s",
                   BF16_SHIFTED: "bf16s"}
PRECISION_BY_NAME = {v: k for k, v in PRECISION_NAMES.items()}
BYTES_OF = {F64: 8, F32: 4, F32_SHIFTED: 4, F16_SHIFTED: 2, BF16_SHIFTED: 2}

# streaming schemes
STREAM_AUTO, STREAM_TWO_ARRAY, STREAM_AA, STREAM_TWO_STEP = 0, 1, 2, 3

SOURCE_NONE, SOURCE_UNIFORM, SOURCE_FIELD = 0, 1, 2

CASE_FREEZE_MOMENTUM = 1
CASE_ZOU_HE_3D = 2
STEP_RESIDUALS = 1
STEP_STOP_ON_CONVERGENCE = 2

REDUCE_MEAN_DENSITY, REDUCE_TOTAL_MASS, REDUCE_SCALAR_TOTAL, REDUCE_ADSORBED, REDUCE_NONFINITE = range(5)

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


class FaceBC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("known", C.c_uint32), ("density", C.c_double),
                ("velocity", C.c_double * 3)]


class Collision(C.Structure):
    _fields_ = [("op", C.c_int32), ("reserved", C.c_int32), ("tau", C.c_double),
                ("omega_plus", C.c_double), ("omega_minus", C.c_double),
                ("relaxation", C.c_double * MAX_Q)]


class ScalarFaceBC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("unknown_populations_only", C.c_int32),
                ("scalar_value", C.c_double)]


class ScalarDesc(C.Structure):
    _fields_ = [("collision", Collision), ("velocity_set", C.c_int32), ("inner_bc", C.c_int32),
                ("faces", ScalarFaceBC * 6), ("initial_value", _dp),
                ("initial_value_uniform", C.c_double), ("source_kind", C.c_int32),
                ("has_adsorption", C.c_int32), ("source_uniform", C.c_double),
                ("source_field", _dp), ("adsorption_rate", C.c_double),
                ("desorption_rate", C.c_double), ("q_max", C.c_double),
                ("initial_q_ads", C.c_double)]


class CaseDesc(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("velocity_set", C.c_int32), ("n", C.c_uint32 * 3),
                ("flags", C.c_uint32), ("collision", Collision), ("node_type", _u8p),
                ("faces", FaceBC * 6), ("force_kind", C.c_int32), ("precision", C.c_int32),
                ("force", C.c_double * 3), ("force_field", _dp), ("initial_density", _dp),
                ("initial_density_uniform", C.c_double), ("initial_velocity", _dp),
                ("initial_velocity_uniform", C.c_double * 3), ("n_scalars", C.c_uint32),
                ("streaming", C.c_int32), ("scalars", C.POINTER(ScalarDesc)),
                ("shift_scale", C.c_double), ("multiphase", C.c_int32), ("reserved", C.c_int32),
                ("g_function", C.c_double), ("wall_density", C.c_double)]


# C-ABI entry points of the product library: (name, restype, argtypes).
# The CPU suite checks that the built library exports every one of these and
# that this list equals the set of functions declared in include/*.h.
_vp = C.c_void_p
ENTRY_POINTS = [
    ("lbf_create", C.c_int, [C.POINTER(CaseDesc), C.POINTER(_vp)]),
    ("lbf_destroy", None, [_vp]),
    ("lbf_initialize_nodes", C.c_int, [_vp]),
    ("lbf_main_steps", C.c_int, [_vp, C.c_uint32, C.c_uint32]),
    ("lbf_synchronize", C.c_int, [_vp]),
    ("lbf_run", C.c_int, [_vp, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("lbf_residuals", C.c_int, [_vp, _dp, _dp, _dp]),
    ("lbf_time_step", C.c_int, [_vp, C.POINTER(C.c_uint64)]),
    ("lbf_reduce", C.c_int, [_vp, C.c_int32, _dp]),
    ("lbf_get_fields", C.c_int, [_vp, _dp, _dp]),
    ("lbf_get_scalar", C.c_int, [_vp, C.c_uint32, _dp, _dp]),
    ("lbf_get_pops", C.c_int, [_vp, C.c_int32, _dp]),
    ("lbf_get_pops_noalloc", C.c_int, [_vp, C.c_int32, _dp]),
    ("lbf_set_pops", C.c_int, [_vp, C.c_int32, _dp]),
    ("lbf_get_link_tables", C.c_int, [_vp, _u8p, C.POINTER(C.c_uint32), _u8p]),
    ("lbf_device_bytes", C.c_int, [C.POINTER(CaseDesc), C.POINTER(C.c_uint64)]),
    ("lbf_last_error", C.c_char_p, []),
]


# ---- serving tier (include/lbflow_serve.h) ---------------------------------
SERVE_KEEP_POPS = 1
REQ_QUEUED, REQ_PREPARING, REQ_RUNNING, REQ_DONE, REQ_FAILED = range(5)


class ServerConfig(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("max_slots", C.c_uint32),
                ("hbm_budget_bytes", C.c_uint64), ("max_steps_per_round", C.c_uint32),
                ("autotune", C.c_int32), ("grid", C.c_uint32), ("tile", C.c_uint32)]


class RequestInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("error", C.c_int32), ("nodes", C.c_uint64),
                ("steps", C.c_uint64), ("device_bytes", C.c_uint64), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("queued", C.c_uint64), ("resident", C.c_uint64), ("rounds", C.c_uint64),
                ("graphs_built", C.c_uint64), ("hbm_in_use", C.c_uint64),
                ("hbm_peak", C.c_uint64), ("hbm_budget", C.c_uint64),
                ("node_updates", C.c_double), ("first_submit", C.c_double),
                ("last_done", C.c_double), ("gpu_seconds", C.c_double),
                ("p50_latency", C.c_double), ("p99_latency", C.c_double),
                ("mean_latency", C.c_double), ("cases_per_second", C.c_double),
                ("mlups", C.c_double), ("prep_seconds", C.c_double),
                ("retire_seconds", C.c_double), ("tuned_grid", C.c_uint32),
                ("tuned_tile", C.c_uint32)]


# benchmark and tuning hooks (include/lbflow_tuning.h; no reference counterpart)
TUNING_ENTRY_POINTS = [
    ("lbfx_copy", C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    ("lbfx_sweep", C.c_int, [_vp, C.c_int, C.c_longlong, C.c_int, C.c_longlong, _vp]),
    ("lbfx_case_stream", _vp, [_vp]),
    ("lbfx_set_block_group", C.c_int, [_vp, C.c_int]),
    ("lbfx_set_plane_pad", C.c_int, [C.c_longlong]),
    ("lbfx_mrt_matrices", C.c_int, [C.c_int, _dp, _dp]),
]

SERVER_ENTRY_POINTS = [
    ("lbf_server_create", C.c_int, [C.POINTER(ServerConfig), C.POINTER(_vp)]),
    ("lbf_server_destroy", None, [_vp]),
    ("lbf_server_submit", C.c_int, [_vp, C.POINTER(CaseDesc), C.c_uint64, C.c_uint32,
                                    C.POINTER(C.c_uint64)]),
    ("lbf_server_info", C.c_int, [_vp, C.c_uint64, C.POINTER(RequestInfo)]),
    ("lbf_server_wait", C.c_int, [_vp, C.c_uint64, C.c_double]),
    ("lbf_server_result", C.c_int, [_vp, C.c_uint64, _dp, _dp, _dp, C.c_int32]),
    ("lbf_server_drain", C.c_int, [_vp, C.c_double]),
    ("lbf_server_get_stats", C.c_int, [_vp, C.POINTER(ServerStats)]),
    ("lbf_server_now", C.c_double, [_vp]),
]


def declare(lib, entry_points, prefix_from="lbf_", prefix_to="lbf_"):
    """Attach restype/argtypes to the functions of a loaded CDLL."""
    for name, res, args in entry_points:
        fn = getattr(lib, name.replace(prefix_from, prefix_to, 1))
        fn.restype = res
        fn.argtypes = args
    return lib
