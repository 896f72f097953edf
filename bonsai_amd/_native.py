"""ctypes binding of libbonsai_gpu.so (include/bonsai_gpu.h, include/bonsai_scenes.h).

The library is built in-tree by ``bonsai_amd.build``; importing this module
loads it and fails loudly if it is missing.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libbonsai_gpu.so"

BG_OK = 0
BG_ERR_INVALID = -1
BG_ERR_HIP = -2
BG_ERR_NOMEM = -3
BG_ERR_UNSUPPORTED = -4
BG_CLOSEST = 0
BG_ANY = 1
BG_COUNT = 2
BG_PROFILE_COUNT = 1
BG_PROFILE_TRAVERSAL_TIME = 2
BG_NOALLOC_SCRATCH_BYTES = 128 << 20
BG_MISS = 0xFFFFFFFF
BG_MAT_MATTE = 0
BG_MAT_CONDUCTOR = 1
BG_MAT_PLASTIC = 2


class bg_ray(C.Structure):
    _fields_ = [("o", C.c_float * 3), ("t_max", C.c_float), ("d", C.c_float * 3), ("time", C.c_float)]


class bg_hit(C.Structure):
    _fields_ = [("prim_id", C.c_uint32), ("t", C.c_float), ("b1", C.c_float), ("b2", C.c_float)]


class bg_surface(C.Structure):
    _fields_ = [("p", C.c_float * 3), ("t", C.c_float), ("p_error", C.c_float * 3), ("n", C.c_float * 3),
                ("ns", C.c_float * 3), ("dpdu", C.c_float * 3), ("uv", C.c_float * 2), ("pad", C.c_float * 2)]


class bg_bvh_node(C.Structure):
    _fields_ = [("bmin", C.c_float * 3), ("bmax", C.c_float * 3), ("offset", C.c_uint32), ("meta", C.c_uint32)]


class bg_material(C.Structure):
    _fields_ = [("type", C.c_uint32), ("reflectance", C.c_float * 3), ("alpha", C.c_float),
                ("eta", C.c_float * 3), ("k", C.c_float * 3)]


class bg_area_light(C.Structure):
    _fields_ = [("prim_id", C.c_uint32), ("radiance", C.c_float * 3), ("double_sided", C.c_uint32)]


class bg_scene_desc(C.Structure):
    _fields_ = [
        ("n_tris", C.c_uint64),
        ("tri_positions", C.c_void_p),
        ("tri_reverse_orientation", C.c_void_p),
        ("tri_material", C.c_void_p),
        ("n_spheres", C.c_uint64),
        ("sphere_object_to_world", C.c_void_p),
        ("sphere_world_to_object", C.c_void_p),
        ("sphere_params", C.c_void_p),
        ("sphere_reverse_orientation", C.c_void_p),
        ("sphere_material", C.c_void_p),
        ("n_materials", C.c_uint32),
        ("materials", C.c_void_p),
        ("n_lights", C.c_uint32),
        ("lights", C.c_void_p),
    ]


class bg_scene_info(C.Structure):
    _fields_ = [("n_prims", C.c_uint64), ("n_nodes", C.c_uint64), ("max_depth", C.c_uint32),
                ("max_leaf", C.c_uint32), ("build_ms", C.c_double), ("device_bytes", C.c_uint64)]


class bg_stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("total_ms", C.c_double), ("rays", C.c_uint64), ("launches", C.c_uint64),
                ("nodes_visited", C.c_uint64), ("tris_tested", C.c_uint64), ("spheres_tested", C.c_uint64),
                ("samples", C.c_uint64), ("shadow_rays", C.c_uint64), ("train_steps", C.c_uint64),
                ("traverse_ms", C.c_double)]


class bg_launch_config(C.Structure):
    _fields_ = [("block", C.c_uint32), ("stack", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("variant", C.c_uint32),
                ("compact", C.c_uint32)]


class bg_camera_desc(C.Structure):
    _fields_ = [
        ("camera_to_world", C.c_float * 16),
        ("camera_to_world_inv", C.c_float * 16),
        ("fov", C.c_float),
        ("near_clip", C.c_float),
        ("far_clip", C.c_float),
        ("lens_radius", C.c_float),
        ("focal_distance", C.c_float),
        ("shutter_open", C.c_float),
        ("shutter_close", C.c_float),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
    ]


class bg_render_desc(C.Structure):
    _fields_ = [
        ("camera", bg_camera_desc),
        ("spp", C.c_uint32),
        ("first_sample", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("rr_threshold", C.c_float),
        ("mode", C.c_uint32),
        ("seed", C.c_uint64),
        ("x0", C.c_uint32),
        ("y0", C.c_uint32),
        ("x1", C.c_uint32),
        ("y1", C.c_uint32),
        ("paths_per_pass", C.c_uint32),
    ]


class bg_guiding_desc(C.Structure):
    _fields_ = [("select_prob", C.c_float), ("train", C.c_uint32), ("train_paths", C.c_uint32),
                ("batch", C.c_uint32), ("steps", C.c_uint32), ("lr", C.c_float), ("replay", C.c_uint32),
                ("lobes", C.c_uint32), ("bounces", C.c_uint32), ("rr_start", C.c_uint32)]


class bg_direct_desc(C.Structure):
    _fields_ = [("integrator", C.c_uint32), ("strategy", C.c_uint32)]


class bg_batch_item(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", bg_camera_desc), ("seed", C.c_uint64)]


class bg_batch_desc(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("max_depth", C.c_uint32), ("rr_threshold", C.c_float), ("mode", C.c_uint32),
                ("paths_per_pass", C.c_uint32), ("use_graph", C.c_uint32), ("autotune", C.c_uint32),
                ("schedule", C.c_uint32)]


BG_MODE_REF = 0
BG_MODE_CORRECTED = 1
BG_SCHEDULE_WAVEFRONT = 0
BG_SCHEDULE_PERSISTENT = 1
BG_FILM_F64 = 1
BG_INTEGRATOR_DIRECT = 1
BG_INTEGRATOR_WHITTED = 2
BG_LIGHT_SAMPLE_ALL = 0
BG_LIGHT_SAMPLE_ONE = 1
BG_OBJ_NORMALS = 1
BG_OBJ_UVS = 2
BG_GUIDE_PARAMS = 19240
BG_GUIDE_OUTPUTS = 40

P = C.c_void_p
# name -> (restype, argtypes); the list every header declaration must appear in.
SIGNATURES = {
    "bg_abi_version": (C.c_int, []),
    "bg_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "bg_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "bg_destroy": (None, [P]),
    "bg_last_error": (C.c_char_p, [P]),
    "bg_stream": (P, [P]),
    "bg_scene_create": (C.c_int, [P, C.POINTER(bg_scene_desc), C.POINTER(P)]),
    "bg_scene_destroy": (None, [P]),
    "bg_scene_info_get": (C.c_int, [P, C.POINTER(bg_scene_info)]),
    "bg_scene_bvh": (C.c_int, [P, P, P]),
    "bg_build_bvh": (C.c_int, [C.POINTER(bg_scene_desc), P, C.POINTER(C.c_uint64), P, C.POINTER(C.c_uint64),
                               C.POINTER(C.c_uint32)]),
    "bg_intersect": (C.c_int, [P, P, P, C.c_size_t, P, C.c_uint32, C.POINTER(bg_stats)]),
    "bg_surface_interactions": (C.c_int, [P, P, P, P, C.c_size_t, P]),
    "bg_intersect_device": (C.c_int, [P, P, P, C.c_size_t, P, C.c_uint32, C.POINTER(bg_stats)]),
    "bg_intersect_device_noalloc": (C.c_int, [P, P, P, C.c_size_t, P, C.c_uint32, C.POINTER(bg_stats)]),
    "bg_set_launch_config": (C.c_int, [P, C.POINTER(bg_launch_config)]),
    "bg_get_launch_config": (C.c_int, [P, C.POINTER(bg_launch_config)]),
    "bg_set_profiling": (C.c_int, [P, C.c_uint32]),
    "bg_camera_rays_device": (C.c_int, [P, C.POINTER(bg_camera_desc), C.c_float, C.c_float, P]),
    "bg_camera_rays": (C.c_int, [P, C.POINTER(bg_camera_desc), C.c_float, C.c_float, P]),
    "bg_look_at": (C.c_int, [P, P, P, P, P]),
    "bg_film_create": (C.c_int, [P, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(P)]),
    "bg_film_destroy": (None, [P]),
    "bg_film_clear": (C.c_int, [P]),
    "bg_film_read": (C.c_int, [P, P, P]),
    "bg_render_film": (C.c_int, [P, P, C.POINTER(bg_render_desc), P, C.POINTER(bg_stats)]),
    "bg_render": (C.c_int, [P, P, C.POINTER(bg_render_desc), P, C.POINTER(bg_stats)]),
    "bg_film_load": (C.c_int, [P, P]),
    "bg_render_film_direct": (C.c_int, [P, P, C.POINTER(bg_render_desc), C.POINTER(bg_direct_desc), P,
                                        C.POINTER(bg_stats)]),
    "bg_film_write_image": (C.c_int, [P, C.c_float, C.c_float, P, C.c_char_p]),
    "bg_film_resolve_device": (C.c_int, [P, C.c_float, C.c_float, P, C.POINTER(C.c_double)]),
    "bg_exr_write": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, P]),
    "bg_exr_encode": (C.c_int, [C.c_uint32, C.c_uint32, P, P, C.POINTER(C.c_uint64)]),
    "bg_guide_create": (C.c_int, [P, P, C.c_uint64, C.POINTER(P)]),
    "bg_guide_destroy": (None, [P]),
    "bg_guide_set_params": (C.c_int, [P, P]),
    "bg_guide_get_params": (C.c_int, [P, P, P, P]),
    "bg_guide_set_bounds": (C.c_int, [P, P, P]),
    "bg_guide_init_params": (C.c_int, [C.c_uint64, P]),
    "bg_guide_snapshot": (C.c_int, [P, P, C.POINTER(C.c_uint64)]),
    "bg_guide_restore": (C.c_int, [P, P, C.c_uint64]),
    "bg_guide_eval": (C.c_int, [P, P, C.c_uint64, P]),
    "bg_guide_train": (C.c_int, [P, P, P, C.c_uint64, C.c_float, C.POINTER(C.c_double), P]),
    "bg_guide_eval_device": (C.c_int, [P, P, C.c_uint64, P, C.POINTER(C.c_double)]),
    "bg_guide_train_device": (C.c_int, [P, P, P, C.c_uint64, C.c_float, C.POINTER(C.c_double)]),
    "bg_render_film_guided": (C.c_int, [P, P, C.POINTER(bg_render_desc), P, C.POINTER(bg_guiding_desc), P,
                                        C.POINTER(bg_stats)]),
    "bg_guide_records": (C.c_int, [P, P, P, C.c_uint64, C.POINTER(C.c_uint64)]),
    "bg_batch_create": (C.c_int, [P, P, C.c_uint32, C.POINTER(P)]),
    "bg_batch_destroy": (None, [P]),
    "bg_render_batch": (C.c_int, [P, P, C.POINTER(bg_batch_desc), C.POINTER(bg_stats)]),
    "bg_batch_film_read": (C.c_int, [P, C.c_uint32, P]),
    "bg_batch_launch_config": (C.c_int, [P, C.POINTER(bg_launch_config)]),
    "bg_obj_read": (C.c_int, [C.c_char_p, P, P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), P, P, P]),
    "bgx_icosphere_tri_count": (C.c_uint64, [C.c_int]),
    "bgx_icosphere": (C.c_int, [C.c_int, P, C.c_float, C.c_float, C.c_float, C.c_uint64, P]),
    "bgx_rays_incoherent": (C.c_int, [C.c_uint64, P, P, C.c_uint64, P]),
    "bgx_uniform": (C.c_float, [C.c_uint64, C.c_uint64]),
}

_lib = None


def lib() -> C.CDLL:
    """Loads the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m bonsai_amd.build` first")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class BonsaiError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"bonsai_gpu error {code}: {msg}")
        self.code = code


def check(rc: int, ctx=None) -> None:
    if rc != BG_OK:
        msg = lib().bg_last_error(ctx).decode() if ctx else ""
        raise BonsaiError(rc, msg)
