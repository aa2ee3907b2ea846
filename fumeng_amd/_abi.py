"""ctypes mirror of include/fm_gpu.h (structure layouts and constants)."""
import ctypes as C

FM_ABI_VERSION = 3
FM_ACCEL_EMBREE = 3
FM_BUILDER_HOST = 0
FM_BUILDER_DEVICE = 1
BUILDERS = {"host": FM_BUILDER_HOST, "device": FM_BUILDER_DEVICE}
TRACE_STACKS = {"auto": 0, "full": 1, "short": 2}

FM_OK = 0
FM_ERR_INVALID = -1
FM_ERR_OOM = -2
FM_ERR_HIP = -3
FM_ERR_VERSION = -4
FM_ERR_IO = -5
FM_ERR_PARSE = -6
FM_ERR_UNSUPPORTED = -7
FM_ERR_NO_DEVICE = -8

FM_TEX_CONSTANT, FM_TEX_IMAGE = 0, 1
FM_WRAP_REPEAT, FM_WRAP_CLAMP = 0, 1
FM_SAMPLER_LINEAR, FM_SAMPLER_NEAREST = 0, 1
FM_IMAGE_F32_LINEAR, FM_IMAGE_U8_SRGB = 0, 1
FM_MAT_LAMBERT, FM_MAT_DISNEY, FM_MAT_SPECULAR_TRANSMISSION, FM_MAT_SPECULAR_REFLECTION = 0, 1, 2, 3
DISNEY_PARAMS = ["basecolor", "metallic", "specular", "specularTint", "roughness", "anisotropic", "sheen",
                 "sheenTint", "clearcoat", "clearcoatGloss", "specTrans", "specTransRoughness", "diffTrans",
                 "flatness", "ior"]
FM_D_NPARAM = 15
FM_SHAPE_TRIANGLE_MESH, FM_SHAPE_SPHERE, FM_SHAPE_RECTANGLE = 0, 1, 2
FM_LIGHT_ENV, FM_LIGHT_DIRECTIONAL = 0, 1
FM_CAMERA_PIN_HOLE, FM_CAMERA_THIN_LENS = 0, 1
FM_ACCEL_BVH, FM_ACCEL_NAIVE, FM_ACCEL_BVH_REFERENCE_TREE = 0, 1, 2
FM_MODE_MEGAKERNEL, FM_MODE_WAVEFRONT = 0, 1
FM_NOALLOC_MAX_PIXELS = 262144   # include/fm_gpu.h
SHEAR_AXIS = {"truncated": 0, "float": 1}
FM_RENDER_TIME_TRACE = 1
FM_RENDER_GRAPH = 2
FM_BVH_WIDTH_AUTO = 0
FM_DPARAM_PER_MATERIAL = 6
DPARAM_NAMES = ["base_r", "base_g", "base_b", "roughness", "metallic", "specular"]

F3 = C.c_float * 3


class fm_texture_desc(C.Structure):
    _fields_ = [("type", C.c_int32), ("image", C.c_int32), ("sampler", C.c_int32),
                ("wrap_u", C.c_int32), ("wrap_v", C.c_int32), ("value", F3)]


class fm_image_desc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("format", C.c_int32), ("pad_", C.c_int32),
                ("data", C.c_void_p)]


class fm_material_desc(C.Structure):
    _fields_ = [("type", C.c_int32), ("thin", C.c_int32), ("tex", fm_texture_desc * FM_D_NPARAM),
                ("reflect_color", F3), ("refract_color", F3), ("eta_i", C.c_float), ("eta_t", C.c_float),
                ("cond_eta_i", F3), ("cond_eta_t", F3), ("cond_k", F3)]


class fm_mesh_desc(C.Structure):
    _fields_ = [("num_positions", C.c_int64), ("positions", C.c_void_p),
                ("num_normals", C.c_int64), ("normals", C.c_void_p),
                ("num_texcoords", C.c_int64), ("texcoords", C.c_void_p),
                ("num_triangles", C.c_int64), ("indices", C.c_void_p)]


class fm_object_desc(C.Structure):
    _fields_ = [("shape", C.c_int32), ("material", C.c_int32), ("mesh", C.c_int32), ("radius", C.c_float),
                ("translation", F3), ("rotation", F3), ("scale", F3),
                ("pa", F3), ("pb", F3), ("pc", F3), ("pd", F3), ("emittance", F3)]


class fm_light_desc(C.Structure):
    _fields_ = [("type", C.c_int32), ("image", C.c_int32), ("rotation", F3), ("factor", C.c_float),
                ("radiance", F3), ("direction", F3)]


class fm_camera_desc(C.Structure):
    _fields_ = [("type", C.c_int32), ("pos", F3), ("look_at", F3), ("up", F3), ("fov", C.c_float),
                ("aspect", C.c_float), ("focal_distance", C.c_float), ("lens_radius", C.c_float)]


class fm_renderer_desc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("thread_count", C.c_int32),
                ("depth", C.c_int32), ("direct_loop", C.c_int32), ("rr_depth", C.c_int32), ("rr_coef", C.c_float),
                ("output", C.c_char * 256)]


FM_POST_TONEMAPPING, FM_POST_LINEAR_TO_SRGB, FM_POST_SRGB_TO_LINEAR, FM_POST_OIDN_DENOISE = 1, 2, 3, 4
FM_MAX_POST_OPS = 16
POST_TYPES = {"tonemapping": FM_POST_TONEMAPPING, "linear_to_srgb": FM_POST_LINEAR_TO_SRGB,
              "srgb_to_linear": FM_POST_SRGB_TO_LINEAR, "oidn_denoise": FM_POST_OIDN_DENOISE}


class fm_post_op(C.Structure):
    _fields_ = [("type", C.c_int32), ("lum", C.c_float)]


class fm_tune_trial(C.Structure):
    _fields_ = [("mode", C.c_int32), ("block_size", C.c_int32), ("grid_blocks", C.c_int32), ("chunk", C.c_int32),
                ("kernel_ms", C.c_double), ("bvh_width", C.c_int32), ("pad_", C.c_int32), ("film_hash", C.c_uint64),
                ("film_mean", C.c_double)]


class fm_tune_result(C.Structure):
    _fields_ = [("best", fm_tune_trial), ("baseline", fm_tune_trial), ("num_trials", C.c_int32),
                ("trial_spp", C.c_int32)]


class fm_scene_desc(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("accelerator", C.c_int32),
                ("camera", fm_camera_desc), ("renderer", fm_renderer_desc),
                ("num_images", C.c_int32), ("images", C.POINTER(fm_image_desc)),
                ("num_materials", C.c_int32), ("materials", C.POINTER(fm_material_desc)),
                ("num_meshes", C.c_int32), ("meshes", C.POINTER(fm_mesh_desc)),
                ("num_objects", C.c_int32), ("objects", C.POINTER(fm_object_desc)),
                ("num_lights", C.c_int32), ("lights", C.POINTER(fm_light_desc)),
                ("num_post_ops", C.c_int32), ("post_ops", C.POINTER(fm_post_op))]


class fm_scene_info(C.Structure):
    _fields_ = [("num_geometries", C.c_int64), ("num_bvh_nodes", C.c_int64), ("bvh_depth", C.c_int32),
                ("num_lights", C.c_int32), ("num_materials", C.c_int32), ("has_env", C.c_int32),
                ("device_bytes", C.c_int64), ("renderer", fm_renderer_desc), ("camera", fm_camera_desc),
                ("env_width", C.c_int32), ("env_height", C.c_int32), ("trav_width", C.c_int32),
                ("trav_stack", C.c_int32), ("num_trav_nodes", C.c_int64), ("trav_node_bytes", C.c_int32),
                ("build_threads", C.c_int32), ("build_ms_geometry", C.c_double), ("build_ms_ref_bvh", C.c_double),
                ("build_ms_trav_tree", C.c_double), ("build_ms_tables", C.c_double), ("build_ms_total", C.c_double),
                ("upload_ms", C.c_double), ("host_bytes", C.c_int64), ("trav_hash", C.c_uint64),
                ("trav_builder", C.c_int32), ("build_iterations", C.c_int32), ("build_ms_device", C.c_double)]


class fm_render_params(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("depth", C.c_int32),
                ("direct_loop", C.c_int32), ("rr_depth", C.c_int32), ("rr_coef", C.c_float), ("mode", C.c_int32),
                ("seed", C.c_uint64), ("sample_offset", C.c_int32), ("block_size", C.c_int32),
                ("grid_blocks", C.c_int32), ("chunk", C.c_int32), ("accumulate", C.c_int32), ("flags", C.c_int32)]


class fm_film(C.Structure):
    _fields_ = [("color", C.c_void_p), ("albedo", C.c_void_p), ("normal", C.c_void_p),
                ("color_sum", C.c_void_p), ("sumsq", C.c_void_p)]


class fm_render_stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("paths", C.c_int64), ("rays", C.c_int64),
                ("node_visits", C.c_int64), ("prim_tests", C.c_int64), ("trace_ms", C.c_double),
                ("trace_block", C.c_int32), ("trace_short_stack", C.c_int32), ("sphere_tests", C.c_int64),
                ("rect_tests", C.c_int64), ("passes", C.c_int32), ("iterations", C.c_int32)]


class fm_job(C.Structure):
    _fields_ = [("camera", fm_camera_desc), ("scene_camera", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
                ("spp", C.c_int32), ("depth", C.c_int32), ("sample_offset", C.c_int32), ("seed", C.c_uint64),
                ("color", C.c_void_p), ("albedo", C.c_void_p), ("normal", C.c_void_p), ("sumsq", C.c_void_p)]


class fm_batch_stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("paths", C.c_int64), ("rays", C.c_int64), ("passes", C.c_int32),
                ("units", C.c_int32)]


class fm_ray(C.Structure):
    _fields_ = [("ori", F3), ("t_min", C.c_float), ("dir", F3), ("t_max", C.c_float)]


class fm_hit(C.Structure):
    _fields_ = [("geom", C.c_int32), ("t", C.c_float), ("pos", F3), ("ng", F3), ("ns", F3), ("ss", F3),
                ("wo", F3), ("uv", C.c_float * 2), ("material", C.c_int32), ("light", C.c_int32)]


class fm_bsdf_query(C.Structure):
    _fields_ = [("material", C.c_int32), ("pad_", C.c_int32), ("ng", F3), ("ns", F3), ("ss", F3), ("wo", F3),
                ("uv", C.c_float * 2), ("wi", F3), ("sample", F3)]


class fm_bsdf_result(C.Structure):
    _fields_ = [("f", F3), ("pdf", C.c_float), ("sf", F3), ("spdf", C.c_float), ("swi", F3),
                ("sdelta", C.c_int32), ("albedo", F3), ("pad_", C.c_float)]


class fm_light_query(C.Structure):
    _fields_ = [("light", C.c_int32), ("pad_", C.c_int32), ("pos", F3), ("sample", F3)]


class fm_light_result(C.Structure):
    _fields_ = [("ref_pos", F3), ("light_pos", F3), ("wi", F3), ("radiance", F3), ("pdf", C.c_float),
                ("dist", C.c_float)]


# numpy dtypes equal to the record structs, for zero-copy batches
def _np_dtype(struct):
    import numpy as np
    return np.dtype(struct)


class fm_dnode(C.Structure):
    """Two-wide traversal node (fm_layout.h): child boxes, child refs (>= 0
    node index, < 0 leaf ~((start << 3) | count))."""
    _fields_ = [("b", C.c_float * 12), ("child", C.c_int32 * 2), ("pad", C.c_int32 * 2)]


EXPORTS = {
    "fm_init": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "fm_shutdown": (None, [C.c_void_p]),
    "fm_last_error": (C.c_char_p, [C.c_void_p]),
    "fm_scene_create": (C.c_int, [C.c_void_p, C.POINTER(fm_scene_desc), C.POINTER(C.c_void_p)]),
    "fm_scene_load": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "fm_scene_destroy": (None, [C.c_void_p]),
    "fm_scene_get_info": (C.c_int, [C.c_void_p, C.POINTER(fm_scene_info)]),
    "fm_render": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(fm_render_params), C.POINTER(fm_film),
                            C.POINTER(fm_render_stats)]),
    "fm_film_read": (C.c_int, [C.c_void_p, C.POINTER(fm_film)]),
    "fm_film_read_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "fm_set_path_budget": (C.c_int, [C.c_void_p, C.c_int64]),
    "fm_set_shear_axis": (C.c_int, [C.c_void_p, C.c_int32]),
    "fm_render_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(fm_render_params), C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.POINTER(fm_render_stats)]),
    "fm_render_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(fm_job), C.c_int, C.POINTER(fm_render_params),
                                  C.POINTER(C.c_double), C.POINTER(fm_batch_stats)]),
    "fm_set_counting": (C.c_int, [C.c_void_p, C.c_int]),
    "fm_set_bvh_width": (C.c_int, [C.c_void_p, C.c_int32]),
    "fm_set_bvh_builder": (C.c_int, [C.c_void_p, C.c_int32]),
    "fm_set_trace_stack": (C.c_int, [C.c_void_p, C.c_int32]),
    "fm_scene_traversal": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64]),
    "fm_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(fm_render_params), C.c_int32,
                              C.POINTER(fm_tune_result), C.c_void_p, C.c_int32]),
    "fm_scene_post_ops": (C.c_int, [C.c_void_p, C.POINTER(fm_post_op), C.c_int32, C.POINTER(C.c_int32)]),
    "fm_post_process": (C.c_int, [C.c_void_p, C.POINTER(fm_post_op), C.c_int32, C.c_void_p, C.c_int64, C.c_void_p,
                                  C.c_void_p]),
    "fm_host_scene_info": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(fm_scene_info)]),
    "fm_params_get": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "fm_params_set": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "fm_params_set_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "fm_render_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(fm_render_params), C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.POINTER(fm_render_stats)]),
    "fm_render_adjoint_device": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(fm_render_params), C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.POINTER(fm_render_stats)]),
    "fm_trace": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_void_p]),
    "fm_camera_rays": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "fm_bsdf_eval": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "fm_light_sample": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "fm_env_eval": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "fm_host_build_bvh": (C.c_int, [C.POINTER(fm_scene_desc), C.POINTER(C.c_int64), C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.POINTER(C.c_int64), C.c_void_p]),
    "fm_desc_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(C.POINTER(fm_scene_desc)),
                               C.c_char_p, C.c_int64]),
    "fm_desc_free": (None, [C.c_void_p]),
}
