"""fumeng_amd: MI355X path-tracing backend for Fumeng scenes.

The product is the C ABI in include/fm_gpu.h (libfumeng_gpu.so, HIP kernels
for gfx950).  This module binds it with ctypes and mirrors the reference's
driver interface: SceneLoader(path).load() / .draw_frame()
(create/scene_loader.h:293-307).  There is no CPU fallback: without the built
library or a GPU every call raises.
"""
import ctypes as C
import os
import weakref

import numpy as np

from . import _abi as A
from .scene import SceneDesc  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfumeng_gpu.so")
_lib = None


class FumengError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("fumeng error %d: %s" % (code, msg))
        self.code = code


def lib():
    """The loaded C ABI library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FumengError(A.FM_ERR_NO_DEVICE, "libfumeng_gpu.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in A.EXPORTS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc, ctx=None):
    if rc != A.FM_OK:
        msg = lib().fm_last_error(ctx)
        raise FumengError(rc, msg.decode() if msg else "")


def _ptr(a):
    return a.ctypes.data if a is not None else None


class Context:
    """One HIP device: stream, device film, scratch (fm_init / fm_shutdown)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().fm_init(device, C.byref(h)))
        self.h = h
        self.device = device
        self._scenes = weakref.WeakSet()

    def close(self):
        if self.h:
            for s in list(self._scenes):   # fm_shutdown releases them
                s.h = None
            lib().fm_shutdown(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_bvh_width(self, width):
        """Traversal tree width (2, 4 or 8) of the scenes created afterwards;
        0 (the default) takes 2 while the two-wide tree and the primitive
        records fit in the device's 256 MB last-level cache, else 4."""
        _check(lib().fm_set_bvh_width(self.h, int(width)), self.h)

    def set_bvh_builder(self, builder):
        """Traversal tree builder of the scenes created afterwards: "host"
        (binned SAH on the host threads, the default) or "device" (PLOC on
        the GPU, two-wide trees)."""
        b = A.BUILDERS[builder] if isinstance(builder, str) else int(builder)
        _check(lib().fm_set_bvh_builder(self.h, b), self.h)

    def set_shear_axis(self, rule="truncated"):
        """Triangle-test axis rule of scenes created afterwards
        (fm_set_shear_axis): "truncated" (abs resolved to the integer
        overload, the default) or "float"."""
        _check(lib().fm_set_shear_axis(self.h, A.SHEAR_AXIS[rule]), self.h)

    def set_path_budget(self, nbytes=0):
        """Bytes of wavefront path state per pass (fm_set_path_budget; 0 = default)."""
        _check(lib().fm_set_path_budget(self.h, int(nbytes)), self.h)

    def set_trace_stack(self, form):
        """Stack form of the wavefront trace kernel: "auto" (default), "full"
        (all entries in LDS) or "short" (16 in LDS, the rest spilled to
        HBM).  Results do not depend on it."""
        _check(lib().fm_set_trace_stack(self.h, A.TRACE_STACKS[form] if isinstance(form, str) else int(form)), self.h)

    def post_process(self, ops, rgb):
        """The reference's output step on the device (fm_post_process):
        returns (post-processed float image, uint8 image).  ops: list of
        {"type": "tonemapping" | "linear_to_srgb" | "srgb_to_linear", "lum": ...}."""
        rgb = np.ascontiguousarray(rgb, np.float32)
        arr = (A.fm_post_op * max(1, len(ops)))()
        for i, o in enumerate(ops):
            arr[i].type = A.POST_TYPES[o["type"]] if isinstance(o, dict) else int(o)
            arr[i].lum = float(o.get("lum", 1.0)) if isinstance(o, dict) else 1.0
        out = np.empty_like(rgb)
        out8 = np.empty(rgb.shape, np.uint8)
        _check(lib().fm_post_process(self.h, arr, len(ops), _ptr(rgb), rgb.size // 3, _ptr(out), _ptr(out8)), self.h)
        return out, out8

    def scene(self, desc):
        """Upload a SceneDesc (fm_scene_create)."""
        d, keep = desc.to_c()
        h = C.c_void_p()
        _check(lib().fm_scene_create(self.h, C.byref(d), C.byref(h)), self.h)
        return self._track(Scene(self, h))

    def _track(self, scene):
        self._scenes.add(scene)
        return scene

    def load(self, json_path):
        """Parse a scene JSON with the reference loader semantics (fm_scene_load)."""
        h = C.c_void_p()
        _check(lib().fm_scene_load(self.h, json_path.encode(), C.byref(h)), self.h)
        return self._track(Scene(self, h))


class Scene:
    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h
        self.params_generation = 0   # counts set_params / set_params_device calls

    def close(self):
        if self.h and self.ctx.h:
            lib().fm_scene_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        i = A.fm_scene_info()
        _check(lib().fm_scene_get_info(self.h, C.byref(i)))
        return i

    def autotune(self, trial_spp=16, mode=A.FM_MODE_WAVEFRONT, width=0, height=0, depth=0, seed=1):
        """fm_autotune: returns {"best": {...}, "baseline": {...}, "trials": [...]}
        with bvh_width / block_size / grid_blocks / chunk / kernel_ms and the
        trial's film_hash / film_mean per configuration; the scene is left at
        best's tree width; pass best's block_size, grid_blocks and chunk to
        render()."""
        p = A.fm_render_params(width=width, height=height, depth=depth, rr_depth=-1, mode=mode, seed=seed)
        res = A.fm_tune_result()
        cap = 256
        arr = (A.fm_tune_trial * cap)()
        _check(lib().fm_autotune(self.ctx.h, self.h, C.byref(p), int(trial_spp), C.byref(res), arr, cap), self.ctx.h)
        conv = lambda t: {"bvh_width": t.bvh_width, "block_size": t.block_size, "grid_blocks": t.grid_blocks,
                          "chunk": t.chunk, "kernel_ms": t.kernel_ms, "film_hash": t.film_hash,
                          "film_mean": t.film_mean}
        return {"best": conv(res.best), "baseline": conv(res.baseline), "trial_spp": res.trial_spp,
                "num_trials": res.num_trials, "trials": [conv(arr[i]) for i in range(min(res.num_trials, cap))]}

    def traversal(self):
        """(nodes, slot_geometry): the device's traversal tree (fm_dnode
        records for two-wide trees, raw bytes for four-wide ones) and the
        geometry index at each slot."""
        i = self.info()
        nb = int(i.num_trav_nodes) * int(i.trav_node_bytes)
        raw = np.zeros(max(nb, 1), np.uint8)
        slots = np.zeros(max(int(i.num_geometries), 1), np.int32)
        _check(lib().fm_scene_traversal(self.h, raw.ctypes.data, nb, slots.ctypes.data, int(i.num_geometries)), self.ctx.h)
        raw, slots = raw[:nb], slots[:int(i.num_geometries)]
        if i.trav_width == 2:
            return raw.view(np.dtype(A.fm_dnode)), slots
        return raw, slots

    def post_ops(self):
        """The scene's post-process list as dicts ({"type": ..., "lum": ...})."""
        arr = (A.fm_post_op * A.FM_MAX_POST_OPS)()
        n = C.c_int32()
        _check(lib().fm_scene_post_ops(self.h, arr, A.FM_MAX_POST_OPS, C.byref(n)))
        names = {v: k for k, v in A.POST_TYPES.items()}
        return [dict({"type": names[arr[i].type]}, **({"lum": arr[i].lum} if arr[i].type == A.FM_POST_TONEMAPPING else {}))
                for i in range(n.value)]

    def render(self, width=0, height=0, spp=0, depth=0, seed=1, rr_depth=-1, rr_coef=0.0, direct_loop=0,
               mode=A.FM_MODE_MEGAKERNEL, sample_offset=0, block_size=0, grid_blocks=0, chunk=0, accumulate=False,
               read=True, time_trace=False, graph=False):
        """Returns (color, albedo, normal, sumsq, stats); arrays are (H, W, 3), row 0 = top."""
        p = A.fm_render_params(width=width, height=height, spp=spp, depth=depth, direct_loop=direct_loop,
                               rr_depth=rr_depth, rr_coef=rr_coef, mode=mode, seed=seed, sample_offset=sample_offset,
                               block_size=block_size, grid_blocks=grid_blocks, chunk=chunk,
                               accumulate=1 if accumulate else 0,
                               flags=(A.FM_RENDER_TIME_TRACE if time_trace else 0) | (A.FM_RENDER_GRAPH if graph else 0))
        info = self.info()
        W = width or info.renderer.width
        H = height or info.renderer.height
        self.last_render_shape = (H, W)
        st = A.fm_render_stats()
        film = A.fm_film()
        out = None
        if read:
            out = [np.zeros((H, W, 3), np.float32) for _ in range(3)] + [np.zeros((H, W, 3), np.float64)]
            film.color, film.albedo, film.normal, film.sumsq = (o.ctypes.data for o in out)
        _check(lib().fm_render(self.ctx.h, self.h, C.byref(p), C.byref(film), C.byref(st)), self.ctx.h)
        if read:
            return out[0], out[1], out[2], out[3], st
        return None, None, None, None, st

    def render_batch(self, jobs, direct_loop=0, rr_depth=-1, rr_coef=0.0, block_size=0, grid_blocks=0, chunk=0,
                     graph=True, read=True):
        """Renders a list of jobs in one wavefront run (fm_render_batch).  A
        job is a dict with optional keys camera (a JSON camera dict; absent =
        the scene's camera), width, height, spp, depth, seed, sample_offset.
        Returns (films, done_ms, stats); films[i] = (color, albedo, normal,
        sumsq) as (H, W, 3) arrays, or None when read is False."""
        from .scene import camera_to_c
        info = self.info()
        n = len(jobs)
        cj = (A.fm_job * max(1, n))()
        films = []
        for i, j in enumerate(jobs):
            x = cj[i]
            if j.get("camera") is not None:
                camera_to_c(j["camera"], x.camera)
            else:
                x.scene_camera = 1
            x.width, x.height = int(j.get("width", 0)), int(j.get("height", 0))
            x.spp, x.depth = int(j.get("spp", 0)), int(j.get("depth", 0))
            x.seed, x.sample_offset = int(j.get("seed", 1)), int(j.get("sample_offset", 0))
            if read:
                W = x.width if x.width > 0 else info.renderer.width
                H = x.height if x.height > 0 else info.renderer.height
                f = [np.zeros((H, W, 3), np.float32) for _ in range(3)] + [np.zeros((H, W, 3), np.float64)]
                x.color, x.albedo, x.normal, x.sumsq = (o.ctypes.data for o in f)
                films.append(tuple(f))
            else:
                films.append(None)
        p = A.fm_render_params(direct_loop=direct_loop, rr_depth=rr_depth, rr_coef=rr_coef, block_size=block_size,
                               grid_blocks=grid_blocks, chunk=chunk, flags=A.FM_RENDER_GRAPH if graph else 0)
        done = (C.c_double * max(1, n))()
        st = A.fm_batch_stats()
        _check(lib().fm_render_batch(self.ctx.h, self.h, cj, n, C.byref(p), done, C.byref(st)), self.ctx.h)
        return films, np.array(done[:n]), st

    # ---- differentiable parameters -------------------------------------------
    def env_shape(self):
        """(H, W) of the first env light's image, or None."""
        i = self.info()
        return (i.env_height, i.env_width) if i.has_env and i.env_width > 0 else None

    def params(self):
        """(material_params (num_materials, 6) float32 [base r g b, roughness,
        metallic, specular], NaN where not a constant Disney value;
        env_texels (H, W, 3) float32 or None)."""
        nm = self.info().num_materials
        mp = np.zeros((nm, A.FM_DPARAM_PER_MATERIAL), np.float32)
        env = None
        if self.env_shape() is not None:
            env = np.zeros(self.env_shape() + (3,), np.float32)
        _check(lib().fm_params_get(self.ctx.h, self.h, _ptr(mp), _ptr(env) if env is not None else None), self.ctx.h)
        return mp, env

    def set_params(self, material_params=None, env_texels=None):
        mp = None if material_params is None else np.ascontiguousarray(material_params, np.float32)
        env = None if env_texels is None else np.ascontiguousarray(env_texels, np.float32)
        _check(lib().fm_params_set(self.ctx.h, self.h, _ptr(mp) if mp is not None else None,
                                   _ptr(env) if env is not None else None), self.ctx.h)
        self.params_generation += 1

    def set_params_device(self, material_params=0, env_texels=0):
        """fm_params_set_device: device addresses (ints; 0 = leave as is) of
        float32 arrays on the context's GPU, (num_materials, 6) and (H, W, 3),
        ready when called."""
        _check(lib().fm_params_set_device(self.ctx.h, self.h, material_params or None, env_texels or None), self.ctx.h)
        self.params_generation += 1

    def film_read_device(self, color=0, albedo=0, normal=0):
        """fm_film_read_device: the last render's means into device arrays
        (addresses of float32 (H, W, 3) arrays on the context's GPU; 0 = skip)."""
        _check(lib().fm_film_read_device(self.ctx.h, color or None, albedo or None, normal or None), self.ctx.h)

    def adjoint_device(self, dl_dimg, g_material, g_env, spp=0, seed=1, sample_offset=0, depth=0, width=0, height=0,
                       rr_depth=-1, rr_coef=0.0, direct_loop=0, grid_blocks=0):
        """fm_render_adjoint_device: dl_dimg (float32 (H, W, 3)), g_material
        (float64 (num_materials, 6)) and g_env (float64 (H, W, 3)) are device
        addresses on the context's GPU (0 = no output); returns the stats."""
        p = A.fm_render_params(width=width, height=height, spp=spp, depth=depth, seed=seed, sample_offset=sample_offset,
                               rr_depth=rr_depth, rr_coef=rr_coef, direct_loop=direct_loop, grid_blocks=grid_blocks)
        st = A.fm_render_stats()
        _check(lib().fm_render_adjoint_device(self.ctx.h, self.h, C.byref(p), dl_dimg, g_material or None,
                                              g_env or None, C.byref(st)), self.ctx.h)
        return st

    def adjoint(self, dl_dimg, spp=0, seed=1, sample_offset=0, depth=0, width=0, height=0, rr_depth=-1, rr_coef=0.0,
                direct_loop=0, grid_blocks=0):
        """dL/d(parameters) for dl_dimg = dL/d(image) (H, W, 3); returns
        (g_material (num_materials, 6) float64, g_env (H, W, 3) float64 or
        None, stats)."""
        info = self.info()
        W = width or info.renderer.width
        H = height or info.renderer.height
        g = np.ascontiguousarray(dl_dimg, np.float32).reshape(H, W, 3)
        gm = np.zeros((info.num_materials, A.FM_DPARAM_PER_MATERIAL), np.float64)
        ge = None
        if self.env_shape() is not None:
            ge = np.zeros(self.env_shape() + (3,), np.float64)
        p = A.fm_render_params(width=width, height=height, spp=spp, depth=depth, seed=seed, sample_offset=sample_offset,
                               rr_depth=rr_depth, rr_coef=rr_coef, direct_loop=direct_loop, grid_blocks=grid_blocks)
        st = A.fm_render_stats()
        _check(lib().fm_render_adjoint(self.ctx.h, self.h, C.byref(p), _ptr(g), _ptr(gm),
                                       _ptr(ge) if ge is not None else None, C.byref(st)), self.ctx.h)
        return gm, ge, st

    def trace(self, rays, any_hit=False):
        """rays: (N, 8) float32 [ox oy oz tmin dx dy dz tmax]; returns fm_hit records."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        out = np.zeros(rays.shape[0], dtype=np.dtype(A.fm_hit))
        _check(lib().fm_trace(self.ctx.h, self.h, _ptr(rays), rays.shape[0], 1 if any_hit else 0, _ptr(out)),
               self.ctx.h)
        return out

    def camera_rays(self, ndc_lens):
        q = np.ascontiguousarray(ndc_lens, dtype=np.float32).reshape(-1, 4)
        out = np.zeros(q.shape[0], dtype=np.dtype(A.fm_ray))
        _check(lib().fm_camera_rays(self.ctx.h, self.h, _ptr(q), q.shape[0], _ptr(out)), self.ctx.h)
        return out

    def bsdf(self, queries):
        q = np.ascontiguousarray(queries, dtype=np.dtype(A.fm_bsdf_query))
        out = np.zeros(q.shape[0], dtype=np.dtype(A.fm_bsdf_result))
        _check(lib().fm_bsdf_eval(self.ctx.h, self.h, _ptr(q), q.shape[0], _ptr(out)), self.ctx.h)
        return out

    def light_sample(self, queries):
        q = np.ascontiguousarray(queries, dtype=np.dtype(A.fm_light_query))
        out = np.zeros(q.shape[0], dtype=np.dtype(A.fm_light_result))
        _check(lib().fm_light_sample(self.ctx.h, self.h, _ptr(q), q.shape[0], _ptr(out)), self.ctx.h)
        return out

    def env_eval(self, dirs):
        d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        pdf = np.zeros(d.shape[0], np.float32)
        rgb = np.zeros((d.shape[0], 3), np.float32)
        _check(lib().fm_env_eval(self.ctx.h, self.h, _ptr(d), d.shape[0], _ptr(pdf), _ptr(rgb)), self.ctx.h)
        return pdf, rgb


def host_bvh(desc):
    """Reference BVH topology built on the host (no GPU needed)."""
    d, keep = desc.to_c()
    nn, ng = C.c_int64(), C.c_int64()
    _check(lib().fm_host_build_bvh(C.byref(d), C.byref(nn), None, None, None, None, C.byref(ng), None))
    boxes = np.zeros((nn.value, 6), np.float32)
    leaf = np.zeros(nn.value, np.int32)
    a = np.zeros(nn.value, np.int64)
    b = np.zeros(nn.value, np.int64)
    order = np.zeros(ng.value, np.int64)
    _check(lib().fm_host_build_bvh(C.byref(d), C.byref(nn), _ptr(boxes), _ptr(leaf), _ptr(a), _ptr(b),
                                   C.byref(ng), _ptr(order)))
    return boxes, leaf, a, b, order


def host_scene_info(desc, bvh_width=0):
    """fm_scene_info of the whole host build of a SceneDesc (no GPU needed)."""
    d, keep = desc.to_c()
    i = A.fm_scene_info()
    _check(lib().fm_host_scene_info(C.byref(d), int(bvh_width), C.byref(i)))
    return i


class SceneLoader:
    """Mirror of the reference's SceneLoader (create/scene_loader.h:278-307):
    load(path) then draw_frame() renders with the scene's renderer block and
    returns the mean colour, albedo and normal films (row 0 at the top)."""

    def __init__(self, device=0):
        self.ctx = Context(device)
        self.scene = None

    def load(self, path):
        self.scene = self.ctx.load(os.path.abspath(path))
        return self

    def draw_frame(self, seed=1, mode=A.FM_MODE_WAVEFRONT):
        """Renders in the wavefront form unless told otherwise (the same
        estimator as the megakernel, 3.4x its throughput)."""
        color, albedo, normal, _, st = self.scene.render(seed=seed, mode=mode)
        return color, albedo, normal, st
