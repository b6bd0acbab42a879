"""ctypes binding of include/skt.h (libskt.so, built from skewtcloudsim_amd/csrc).

The library is the product: there is no Python or CPU fallback. Loading fails
loudly when the shared object is missing or does not export the ABI.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libskt.so")

# skt_field ids (include/skt.h)
FIELDS = [
    "QV", "QW", "QC", "QR", "QS", "QI", "POTTEMP", "VELX", "VELY", "VELZ", "PRESSURE",
    "DENSITY", "DENSITY_OLD", "PRESSURE_PROJ",
    "QRS", "QGR", "QGS", "QGI", "GROUND_P", "GROUND_TIME", "GROUND_T", "CLOUD_COVER",
    "ISEN_TEMP", "ISEN_VAPOR", "DEFAULT_VELX", "DEFAULT_VELZ", "DEFAULT_PRESSURE",
]
FIELD = {n: i for i, n in enumerate(FIELDS)}
SKY_FIELDS = FIELDS[:14]
GROUND_FIELDS = FIELDS[14:22]
PROFILE_FIELDS = FIELDS[22:]
NEUMANN, DIRICHLET, CUSTOM = 0, 1, 2
PCG_REFERENCE, PCG_STANDARD = 0, 1
PARAM = {n: i for i, n in enumerate(["POTTEMP", "QV", "QW", "QC", "QR", "QS", "QI", "WIND", "PGROUND", "PRESSURE"])}
DIRECTION = {n: i for i, n in enumerate(["LEFT", "RIGHT", "UP", "DOWN", "FORWARD", "BACKWARD"])}


class Config(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32), ("dx", ctypes.c_float),
        ("device", ctypes.c_int32), ("pcg_mode", ctypes.c_int32), ("pcg_tol", ctypes.c_float),
        ("pcg_max_iter", ctypes.c_int32), ("diffusion_sweeps", ctypes.c_int32), ("max_substeps", ctypes.c_int32),
        ("maxabs_mode", ctypes.c_int32), ("noise_seed", ctypes.c_uint32), ("init_noise", ctypes.c_int32),
        ("max_blocks", ctypes.c_int32), ("tracer_format", ctypes.c_int32),
        ("tracer_scale_log2", ctypes.c_int32 * 5), ("pcg_refine", ctypes.c_int32), ("pcg_refine_tol", ctypes.c_float),
        ("pcg_precon", ctypes.c_int32),
    ]


class Bounds(ctypes.Structure):
    _fields_ = [("sides", ctypes.c_int32), ("up", ctypes.c_int32), ("ground", ctypes.c_int32)]


class Params(ctypes.Structure):
    _fields_ = [
        ("time", ctypes.c_float), ("day", ctypes.c_int32), ("longitude", ctypes.c_float),
        ("sun_strength", ctypes.c_float), ("pause_diurnal", ctypes.c_int32), ("ground_changed", ctypes.c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("pcg_iterations", ctypes.c_int32), ("pcg_residual", ctypes.c_float), ("pcg_residual0", ctypes.c_float),
        ("n_advect", ctypes.c_int32), ("substeps", ctypes.c_int32 * 16), ("pcg_passes", ctypes.c_int32),
        ("pcg_iterations_total", ctypes.c_int32), ("pcg_residual64", ctypes.c_double), ("pcg_rhs64", ctypes.c_double),
    ]


class Diagnostics(ctypes.Structure):
    _fields_ = [
        ("max_updraft", ctypes.c_float), ("cloud_top_level", ctypes.c_int32), ("cloud_top_height", ctypes.c_float),
        ("pad", ctypes.c_int32), ("surface_rain", ctypes.c_double),
    ]


class Timing(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_int64 * 9), ("total_ms", ctypes.c_double * 9)]


TIMING_CLASSES = ["diffusion_sweep", "ppm_pass", "pcg_spmv", "pcg_update", "microphysics", "maxvel", "precip_fall",
                  "other", "multigrid"]


# skt_sounding_table columns, in declaration order (radioSondeData, Game/readTable.h:8-15)
SOUNDING_COLUMNS = ["altitude", "pressure", "dew_point", "temperature", "wind_speed", "wind_dir"]
DWD_REFERENCE_DATE_INDEX = 161
SKT_ERR_ARG = 1


class SoundingSpec(ctypes.Structure):
    _fields_ = [
        ("surface_theta", ctypes.c_float), ("surface_pressure", ctypes.c_float), ("lapse_rate", ctypes.c_float),
        ("tropopause_height", ctypes.c_float), ("surface_qv", ctypes.c_float), ("surface_rh", ctypes.c_float),
        ("qv_scale_height", ctypes.c_float), ("shear_u", ctypes.c_float), ("shear_height", ctypes.c_float),
        ("n_bubbles", ctypes.c_int32), ("bubble_amp", ctypes.c_float * 8), ("bubble_radius", ctypes.c_float * 8),
        ("bubble_x", ctypes.c_float * 8), ("bubble_y", ctypes.c_float * 8), ("bubble_z", ctypes.c_float * 8),
    ]


class SoundingTable(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int32), ("capacity", ctypes.c_int32)] + [
        (n, ctypes.POINTER(ctypes.c_float)) for n in SOUNDING_COLUMNS]


class BatchOptions(ctypes.Structure):
    _fields_ = [("use_graph", ctypes.c_int32), ("advection_substep_cap", ctypes.c_int32),
                ("launch_mode", ctypes.c_int32)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("step_ms", ctypes.c_double * 3), ("tuned_step_ms", ctypes.c_double), ("bodies", ctypes.c_int32),
                ("persistent_bodies", ctypes.c_int32)]


class ServerOptions(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("dt", ctypes.c_float), ("hbm_budget_bytes", ctypes.c_int64),
                ("max_active", ctypes.c_int32), ("cloud_threshold", ctypes.c_float), ("member_config", Config),
                ("batch", BatchOptions)]


class JobRequest(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32), ("dx", ctypes.c_float),
                ("sounding", SoundingSpec), ("sim_seconds", ctypes.c_float), ("report_every", ctypes.c_float)]


class JobEvent(ctypes.Structure):
    _fields_ = [("job", ctypes.c_int64), ("kind", ctypes.c_int32), ("steps", ctypes.c_int32),
                ("sim_time", ctypes.c_float), ("pad", ctypes.c_int32), ("wall_s", ctypes.c_double),
                ("submitted_s", ctypes.c_double), ("diag", Diagnostics)]


class ServerStatus(ctypes.Structure):
    _fields_ = [("active", ctypes.c_int32), ("queued", ctypes.c_int32), ("bytes_in_use", ctypes.c_int64),
                ("budget_bytes", ctypes.c_int64), ("steps", ctypes.c_int64), ("completed", ctypes.c_int64),
                ("cell_steps", ctypes.c_double), ("busy_s", ctypes.c_double), ("admit_s", ctypes.c_double),
                ("replan_s", ctypes.c_double)]


EVENT_KINDS = {0: "admitted", 1: "report", 2: "done"}

_P = ctypes.c_void_p
_F = ctypes.c_float
_I = ctypes.c_int32

# name -> (restype, argtypes); every entry point declared in include/skt.h
SIGNATURES = {
    "skt_config_default": (None, [ctypes.POINTER(Config), _I, _I, _I, _F]),
    "skt_env_create": (_I, [ctypes.POINTER(Config), ctypes.POINTER(_P)]),
    "skt_env_destroy": (_I, [_P]),
    "skt_env_last_error": (ctypes.c_char_p, [_P]),
    "skt_env_set_stream": (_I, [_P, _P]),
    "skt_env_synchronize": (_I, [_P]),
    "skt_env_set_ground_height": (_I, [_P, _P]),
    "skt_env_init": (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "skt_env_update": (_I, [_P, _F, _F]),
    "skt_env_micro_physics_ground": (_I, [_P, _F, _F, _F]),
    "skt_env_micro_physics_sky": (_I, [_P, _F, _F]),
    "skt_env_diffuse": (_I, [_P, _I, _I, _F]),
    "skt_env_advect_ground_water": (_I, [_P, _F, _F]),
    "skt_env_set_temps_at_ground": (_I, [_P, _F, _F]),
    "skt_env_advect_ppm": (_I, [_P, _I, _I, Bounds, _F]),
    "skt_env_advect_precip": (_I, [_P, _I, _I, _F]),
    "skt_env_pressure_project": (_I, [_P, _F]),
    "skt_env_calculate_pressure_project": (_I, [_P, _F]),
    "skt_env_editor_data": (_I, [_P]),
    "skt_env_irradiance": (_I, [_P, ctypes.POINTER(_F)]),
    "skt_env_ground_coverage_factor": (_I, [_P]),
    "skt_env_update_ground_temps": (_I, [_P, _F, _F, _F]),
    "skt_env_calculate_buoyancy": (_I, [_P, _F]),
    "skt_env_init_density": (_I, [_P]),
    "skt_env_calculate_new_pressure": (_I, [_P]),
    "skt_env_get_param_array": (_I, [_P, _I, _I, ctypes.POINTER(_P)]),
    "skt_env_get_field_ptr": (_I, [_P, _I, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_int64)]),
    "skt_env_copy_field_to_host": (_I, [_P, _I, _P]),
    "skt_env_copy_field_from_host": (_I, [_P, _I, _P]),
    "skt_env_get_field_format": (_I, [_P, _I, ctypes.POINTER(_I), ctypes.POINTER(_F)]),
    "skt_env_set_params": (_I, [_P, ctypes.POINTER(Params)]),
    "skt_env_get_params": (_I, [_P, ctypes.POINTER(Params)]),
    "skt_env_get_stats": (_I, [_P, ctypes.POINTER(Stats)]),
    "skt_env_diagnostics": (_I, [_P, _F, ctypes.POINTER(Diagnostics)]),
    "skt_terrain_generate": (_I, [_I, _I, _I, _I, _I, _F, _P, _P]),
    "skt_terrain_generate_noalloc": (_I, [_I, _I, _I, _I, _I, _F, _P, _P]),
    "skt_env_apply_brush": (_I, [_P, _I, _F, _P, _F, _F, _F, _F, _P, _I]),
    "skt_env_get_ground_height": (_I, [_P, _P]),
    "skt_env_apply_selection": (_I, [_P, _I, _P, _P, _F, _P, _I]),
    "skt_env_reset_parameter": (_I, [_P, _I]),
    "skt_env_compute_masks": (_I, [_P, _P]),
    "skt_env_phase_mask": (_I, [_P, _P]),
    "skt_mphys_region_set": (_I, [_P, _P, _P]),
    "skt_mphys_region_clear": (_I, [_P]),
    "skt_mphys_region_get": (_I, [_P, _P]),
    "skt_env_set_timing": (_I, [_P, _I]),
    "skt_env_set_timing_stride": (_I, [_P, _I]),
    "skt_env_get_timing": (_I, [_P, ctypes.POINTER(Timing)]),
    "skt_env_copy_probe": (_I, [_P, _I, ctypes.POINTER(_F), ctypes.POINTER(ctypes.c_double)]),
    "skt_device_synchronize": (_I, [_I]),
    "skt_sounding_spec_random": (_I, [ctypes.POINTER(SoundingSpec), ctypes.c_uint64, _I, _I, _I, _F, _I]),
    "skt_sounding_generate": (_I, [ctypes.POINTER(SoundingSpec), _I, _I, _I, _F, _P, _P, _P, _P, _P, _P, _P]),
    "skt_input_last_error": (ctypes.c_char_p, []),
    "skt_sounding_read_knmi": (_I, [ctypes.c_char_p, ctypes.POINTER(SoundingTable)]),
    "skt_sounding_parse_knmi": (_I, [ctypes.c_char_p, ctypes.c_int64, ctypes.POINTER(SoundingTable)]),
    "skt_sounding_read_dwd": (_I, [ctypes.c_char_p, _I, ctypes.POINTER(SoundingTable)]),
    "skt_sounding_parse_dwd": (_I, [ctypes.c_char_p, ctypes.c_int64, ctypes.c_char_p, ctypes.c_int64, _I,
                                    ctypes.POINTER(SoundingTable)]),
    "skt_sounding_init_fields": (_I, [ctypes.POINTER(SoundingTable), _I, _I, _I, _F, _P, _P, _P, _P, _P, _P, _P]),
    "skt_env_init_sounding": (_I, [_P, ctypes.POINTER(SoundingTable)]),
    "skt_batch_options_default": (None, [ctypes.POINTER(BatchOptions)]),
    "skt_batch_create": (_I, [_I, ctypes.POINTER(_P)]),
    "skt_batch_destroy": (_I, [_P]),
    "skt_batch_last_error": (ctypes.c_char_p, [_P]),
    "skt_batch_set_options": (_I, [_P, ctypes.POINTER(BatchOptions)]),
    "skt_batch_add": (_I, [_P, _P, ctypes.POINTER(_I)]),
    "skt_batch_remove": (_I, [_P, _I]),
    "skt_batch_size": (_I, [_P, ctypes.POINTER(_I)]),
    "skt_batch_step": (_I, [_P, _F, _F]),
    "skt_batch_synchronize": (_I, [_P]),
    "skt_batch_autotune": (_I, [_P, _F, _F, ctypes.POINTER(BatchTuning)]),
    "skt_batch_profile": (_I, [_P, _F, _F, ctypes.c_char_p, ctypes.c_int64]),
    "skt_env_footprint": (_I, [ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_int64)]),
    "skt_env_device_bytes": (_I, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "skt_server_options_default": (None, [ctypes.POINTER(ServerOptions)]),
    "skt_server_create": (_I, [ctypes.POINTER(ServerOptions), ctypes.POINTER(_P)]),
    "skt_server_destroy": (_I, [_P]),
    "skt_server_last_error": (ctypes.c_char_p, [_P]),
    "skt_server_submit": (_I, [_P, ctypes.POINTER(JobRequest), ctypes.POINTER(ctypes.c_int64)]),
    "skt_server_step": (_I, [_P]),
    "skt_server_poll": (_I, [_P, ctypes.POINTER(JobEvent), _I, ctypes.POINTER(_I)]),
    "skt_server_status_get": (_I, [_P, ctypes.POINTER(ServerStatus)]),
}

_lib = None


def load_library(path=None):
    """Loads libskt.so and binds every ABI symbol; raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} not found: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)  # AttributeError if the ABI is incomplete
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
