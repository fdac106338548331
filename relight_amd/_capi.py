"""ctypes declarations of include/relight.h (librelight_amd.so).

The library is the product: there is no Python or CPU fallback. Importing this module raises
if the HIP library has not been built.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librelight_amd.so")

RL_SUCCESS = 0
RL_INVALID_CALL = 1
RL_NOT_IMPLEMENTED = 2
RL_OUT_OF_MEMORY = 3
RL_REJECTED = 4
RL_DEVICE_ERROR = 5
RL_INVALID_ARGUMENT = 6

RL_LIGHT_POINT = 0
RL_LIGHT_SPOT = 1
RL_LIGHT_DIRECTIONAL = 2
RL_LIGHT_AREA = 3

RL_SAMPLE_UNIFORM = 0
RL_SAMPLE_COSINE = 1

RL_TRACE_ANY = 0
RL_TRACE_CLOSEST = 1
RL_NOALLOC_MAX_SEGMENTS = 65536

STATUS_NAMES = {
    0: "Success", 1: "InvalidCall", 2: "NotImplemented", 3: "OutOfMemory",
    4: "Rejected", 5: "DeviceError", 6: "InvalidArgument", 7: "NotConverged",
}


class rl_light_desc(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("casts_shadow", C.c_int32),
        ("position", C.c_float * 3),
        ("direction", C.c_float * 3),
        ("color", C.c_float * 3),
        ("intensity", C.c_float),
        ("has_attenuation", C.c_int32),
        ("radius", C.c_float),
        ("att_constant", C.c_float),
        ("att_linear", C.c_float),
        ("att_quadratic", C.c_float),
        ("spot_cutoff", C.c_float),
        ("spot_exponent", C.c_float),
        ("area_positions", C.POINTER(C.c_float)),
        ("area_indices", C.POINTER(C.c_uint32)),
        ("area_vertex_count", C.c_uint32),
        ("area_index_count", C.c_uint32),
    ]


class rl_ao_settings(C.Structure):
    _fields_ = [
        ("samples", C.c_int32),
        ("occluded_fraction", C.c_float),
        ("max_distance", C.c_float),
        ("exponent", C.c_float),
        ("seed", C.c_uint32),
        ("sampling", C.c_int32),
    ]


class rl_bake_iterator(C.Structure):
    _fields_ = [("kind", C.c_int32), ("first", C.c_int32), ("step", C.c_int32)]


rl_progress_fn = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.c_int32, C.c_int32)


class rl_progress(C.Structure):
    _fields_ = [("notify", rl_progress_fn), ("user", C.c_void_p), ("updates", C.c_int32)]


RL_ITERATE_LUMELS = 0
RL_ITERATE_FACES = 1


class rl_hit(C.Structure):
    _fields_ = [
        ("t", C.c_float),
        ("triangle", C.c_int32),
        ("mesh", C.c_int32),
        ("face", C.c_int32),
        ("u", C.c_float),
        ("v", C.c_float),
    ]


RL_PHOTON_THRESHOLD = 0
RL_PHOTON_RUSSIAN_ROULETTE = 1


class rl_indirect_settings(C.Structure):
    _fields_ = [
        ("photon_pass_count", C.c_int32),
        ("photon_bounce_count", C.c_int32),
        ("photon_energy_threshold", C.c_float),
        ("photon_max_distance", C.c_float),
        ("final_gather_samples", C.c_int32),
        ("final_gather_distance", C.c_float),
        ("final_gather_radius", C.c_int32),
        ("sky_color", C.c_float * 3),
        ("ambient_color", C.c_float * 3),
        ("seed", C.c_uint32),
        ("sampling", C.c_int32),
        ("photons_per_light", C.c_int32),
        ("termination", C.c_int32),
        ("gather_mode", C.c_int32),
        ("knn_k", C.c_int32),
        ("knn_radius", C.c_float),
    ]


RL_FORMAT_RAW_HDR = 0
RL_FORMAT_TGA_DOUBLE_LDR = 1
RL_FORMAT_TGA_RGBM = 2

RL_GATHER_TEXEL_DISC = 0
RL_GATHER_KNN = 1


RL_KERNEL_DIRECT = 0
RL_KERNEL_AO_COUNT = 1
RL_KERNEL_AO_FINALIZE = 2
RL_KERNEL_TRACE = 3
RL_KERNEL_PHOTON_TRACE = 4
RL_KERNEL_PHOTON_STORE = 5
RL_KERNEL_PHOTON_GATHER = 6
RL_KERNEL_FINAL_GATHER = 7
RL_KERNEL_FORM_FACTORS = 8
RL_KERNEL_FF_GEMM_ACCURATE = 9
RL_KERNEL_FF_GEMM_FAST = 10
RL_KERNEL_SLOTS = 16
KERNEL_NAMES = {0: "direct", 1: "ao_count", 2: "ao_finalize", 3: "trace", 4: "photon_trace",
                5: "photon_store", 6: "photon_gather", 7: "final_gather", 8: "form_factors",
                9: "ff_gemm_accurate", 10: "ff_gemm_fast", 11: "photon_grid"}
RL_NOT_CONVERGED = 7


RL_AO_KERNEL_CHUNKED = 0
RL_AO_KERNEL_REFILL = 1
RL_FG_KERNEL_AUTO = 0
RL_FG_KERNEL_WAVE = 1
RL_FG_KERNEL_LANE = 2
RL_DIRECT_KERNEL_AUTO = 0
RL_DIRECT_KERNEL_LUMEL = 1
RL_DIRECT_KERNEL_PAIRS = 2


class rl_tuning(C.Structure):
    _fields_ = [("ao_kernel", C.c_int32), ("refill_min", C.c_uint32), ("ao_chunk_lumels", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("fg_kernel", C.c_int32), ("direct_kernel", C.c_int32),
                ("local_trees", C.c_int32), ("occluder_hints", C.c_int32)]


class rl_solve_info(C.Structure):
    _fields_ = [("outer_iterations", C.c_int32), ("sweeps", C.c_int32), ("residual", C.c_double),
                ("gpu_ms", C.c_double)]


class rl_stats(C.Structure):
    _fields_ = [
        ("rays", C.c_uint64),
        ("texels", C.c_uint64),
        ("gpu_ms", C.c_double),
        ("wall_ms", C.c_double),
        ("launches", C.c_int32),
        ("pad", C.c_int32),
        ("kernel_ms", C.c_double * RL_KERNEL_SLOTS),
        ("kernel_launches", C.c_int32 * RL_KERNEL_SLOTS),
    ]


RL_JOB_DIRECT_AO = 0
RL_JOB_INDIRECT = 1
RL_JOB_RADIOSITY = 2
RL_JOB_FINAL_GATHER = 3
RL_JOB_FULL_BAKE = 4
RL_JOB_QUEUED = 0
RL_JOB_RUNNING = 1
RL_JOB_DONE = 2
RL_JOB_FAILED = 3
RL_JOB_CANCELLED = 4
JOB_STATE_NAMES = {0: "queued", 1: "running", 2: "done", 3: "failed", 4: "cancelled"}


class rl_bake_item(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("lightmap", C.c_void_p), ("first_lumel", C.c_uint32),
                ("lumel_count", C.c_uint32), ("direct", C.c_int32),
                ("ao", C.POINTER(rl_ao_settings))]


class rl_server_config(C.Structure):
    _fields_ = [("memory_budget", C.c_uint64), ("max_active", C.c_int32), ("quantum_rays", C.c_uint64)]


class rl_job_desc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("scene", C.c_void_p),
        ("lightmaps", C.POINTER(C.c_void_p)),
        ("num_lightmaps", C.c_int32),
        ("direct", C.c_int32),
        ("ao", C.POINTER(rl_ao_settings)),
        ("indirect", C.POINTER(rl_indirect_settings)),
        ("radiosity", C.c_void_p),
        ("rays_per_patch", C.c_uint32),
        ("ff_seed", C.c_uint32),
        ("ff_max_distance", C.c_float),
        ("emission", C.POINTER(C.c_float)),
        ("reflectance", C.POINTER(C.c_float)),
        ("num_rhs", C.c_uint32),
        ("tolerance", C.c_float),
        ("out_radiosity", C.POINTER(C.c_float)),
    ]


class rl_job_status(C.Structure):
    _fields_ = [("state", C.c_int32), ("status", C.c_int32), ("submit_ms", C.c_double),
                ("admit_ms", C.c_double), ("done_ms", C.c_double), ("rays", C.c_uint64),
                ("footprint", C.c_uint64)]


class rl_server_stats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("cancelled", C.c_uint64), ("rejected", C.c_uint64), ("steps", C.c_uint64),
                ("rays", C.c_uint64), ("busy_ms", C.c_double), ("peak_memory", C.c_uint64)]


_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_F = C.POINTER(C.c_float)
_U32 = C.POINTER(C.c_uint32)
_I32 = C.POINTER(C.c_int32)

# (name, restype, argtypes) for every entry point of include/relight.h.
SIGNATURES = [
    ("rl_last_error", C.c_char_p, []),
    ("rl_create", C.c_int32, [_PP]),
    ("rl_destroy", C.c_int32, [_P]),
    ("rl_scene_create", C.c_int32, [_P, _PP]),
    ("rl_scene_destroy", C.c_int32, [_P]),
    ("rl_scene_begin", C.c_int32, [_P]),
    ("rl_scene_add_mesh", C.c_int32, [_P, _F, _F, _F, C.c_uint32, _U32, C.c_uint32, _F, _F, _I32]),
    ("rl_texture_create", C.c_int32, [_P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_uint8), _PP]),
    ("rl_texture_destroy", C.c_int32, [_P]),
    ("rl_scene_add_mesh_textured", C.c_int32, [_P, _F, _F, _F, _F, C.c_uint32, _U32, C.c_uint32, _F, _F, _P,
                                               _I32]),
    ("rl_scene_add_light", C.c_int32, [_P, C.POINTER(rl_light_desc), _I32]),
    ("rl_scene_end", C.c_int32, [_P]),
    ("rl_scene_info", C.c_int32, [_P, _U32, _U32, _I32]),
    ("rl_mesh_generate_uv", C.c_int32, [_F, _F, _F, C.c_uint32, _U32, C.c_uint32, C.c_float, C.c_float, C.c_float,
                                        _U32, _U32, _F, _F, _F, _F, _U32]),
    ("rl_lightmap_create", C.c_int32, [_P, C.c_int32, C.c_int32, _PP]),
    ("rl_lightmap_destroy", C.c_int32, [_P]),
    ("rl_lightmap_add_mesh", C.c_int32, [_P, _P, C.c_int32]),
    ("rl_lightmap_set_seed", C.c_int32, [_P, C.c_uint32]),
    ("rl_lightmap_valid_count", C.c_int32, [_P, _U32]),
    ("rl_lightmap_read_lumels", C.c_int32, [_P, _U32, _F, _F]),
    ("rl_lightmap_read_hdr", C.c_int32, [_P, _F]),
    ("rl_bake_direct", C.c_int32, [_P, _P, C.c_int32]),
    ("rl_bake_ambient_occlusion", C.c_int32, [_P, _P, C.c_int32, C.POINTER(rl_ao_settings)]),
    ("rl_bake_batch", C.c_int32, [_P, _P, C.POINTER(C.c_void_p), C.c_int32, C.c_int32,
                                  C.POINTER(rl_ao_settings)]),
    ("rl_pass_create", C.c_int32, [_P, _P, C.POINTER(C.c_void_p), C.c_int32, C.c_int32,
                                   C.POINTER(rl_ao_settings), _PP]),
    ("rl_pass_launch", C.c_int32, [_P, _P]),
    ("rl_pass_destroy", C.c_int32, [_P]),
    ("rl_trace_segments", C.c_int32, [_P, _P, _F, _F, C.c_uint32, C.c_int32,
                                      C.POINTER(C.c_uint8), C.POINTER(rl_hit)]),
    ("rl_trace_segments_noalloc", C.c_int32, [_P, _P, _F, _F, C.c_uint32, C.c_int32,
                                              C.POINTER(C.c_uint8), C.POINTER(rl_hit)]),
    ("rl_photonmap_create", C.c_int32, [_P, C.c_int32, C.c_int32, _PP]),
    ("rl_photonmap_add_mesh", C.c_int32, [_P, _P, C.c_int32]),
    ("rl_photonmap_read", C.c_int32, [_P, _F, _U32, _F]),
    ("rl_emit_photons", C.c_int32, [_P, _P, C.POINTER(rl_indirect_settings)]),
    ("rl_bake_indirect", C.c_int32, [_P, _P, C.c_int32, C.POINTER(rl_indirect_settings)]),
    ("rl_bake_direct_iter", C.c_int32, [_P, _P, C.c_int32, C.POINTER(rl_bake_iterator), C.POINTER(rl_progress)]),
    ("rl_bake_ambient_occlusion_iter", C.c_int32, [_P, _P, C.c_int32, C.POINTER(rl_ao_settings),
                                                   C.POINTER(rl_bake_iterator), C.POINTER(rl_progress)]),
    ("rl_bake_indirect_iter", C.c_int32, [_P, _P, C.c_int32, C.POINTER(rl_indirect_settings),
                                          C.POINTER(rl_bake_iterator), C.POINTER(rl_progress)]),
    ("rl_get_progress", C.c_int32, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("rl_radiosity_create", C.c_int32, [_P, _P, _F, C.c_uint32, _U32, _PP]),
    ("rl_radiosity_destroy", C.c_int32, [_P]),
    ("rl_radiosity_build_form_factors", C.c_int32, [_P, _P, C.c_uint32, C.c_uint32, C.c_float]),
    ("rl_radiosity_read_form_factors", C.c_int32, [_P, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint16)]),
    ("rl_radiosity_solve", C.c_int32, [_P, _P, _F, _F, C.c_uint32, C.c_float, C.c_int32, C.c_int32, _F,
                                       C.POINTER(rl_solve_info)]),
    ("rl_set_tuning", C.c_int32, [_P, C.POINTER(rl_tuning)]),
    ("rl_get_tuning", C.c_int32, [_P, C.POINTER(rl_tuning)]),
    ("rl_get_stats", C.c_int32, [_P, C.POINTER(rl_stats)]),
    ("rl_bake_items", C.c_int32, [_P, C.POINTER(rl_bake_item), C.c_int32]),
    ("rl_obj_load", C.c_int32, [C.c_char_p, _U32, _U32, _F, _F, _F, _U32]),
    ("rl_scene_add_obj", C.c_int32, [_P, C.c_char_p, _F, _F, _I32]),
    ("rl_bake_scene", C.c_int32, [_P, _P, C.c_int32, C.POINTER(rl_ao_settings), C.POINTER(rl_indirect_settings)]),
    ("rl_lightmap_expand", C.c_int32, [_P]),
    ("rl_lightmap_blur", C.c_int32, [_P]),
    ("rl_lightmap_encode", C.c_int32, [_P, C.c_int32, C.POINTER(C.c_uint8)]),
    ("rl_lightmap_save", C.c_int32, [_P, C.c_char_p, C.c_int32]),
    ("rl_server_create", C.c_int32, [_P, C.POINTER(rl_server_config), _PP]),
    ("rl_server_submit", C.c_int32, [_P, C.POINTER(rl_job_desc), C.POINTER(C.c_uint64)]),
    ("rl_server_poll", C.c_int32, [_P, C.c_uint64, C.POINTER(rl_job_status)]),
    ("rl_server_wait", C.c_int32, [_P, C.c_uint64, C.c_double, C.POINTER(rl_job_status)]),
    ("rl_server_cancel", C.c_int32, [_P, C.c_uint64]),
    ("rl_server_stats_get", C.c_int32, [_P, C.POINTER(rl_server_stats)]),
    ("rl_server_destroy", C.c_int32, [_P]),
]

_lib = None


def load():
    """Loads librelight_amd.so; raises RuntimeError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"librelight_amd.so is missing at {LIB_PATH}: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
