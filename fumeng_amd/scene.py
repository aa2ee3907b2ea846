"""Scene descriptions in Python: the in-memory form of what the reference's
SceneLoader extracts from a scene JSON (scene_loader.h:42-302), convertible to
the C descriptor (fm_scene_desc) and writable as reference-loadable files
(scene JSON + OBJ + Radiance .hdr + TGA)."""
import ctypes as C
import json
import os

import numpy as np

from . import _abi as A


def const_tex(v, wrap_u="repeat", wrap_v="repeat"):
    v = np.broadcast_to(np.asarray(v, dtype=np.float32), (3,))
    return {"type": "constant", "value": [float(x) for x in v], "wrap_u": wrap_u, "wrap_v": wrap_v}


def image_tex(image, sampler="linear", wrap_u="repeat", wrap_v="repeat"):
    return {"type": "image", "image": int(image), "sampler": sampler, "wrap_u": wrap_u, "wrap_v": wrap_v}


DISNEY_DEFAULTS = {k: 0.0 for k in A.DISNEY_PARAMS}
DISNEY_DEFAULTS.update({"specular": 0.5, "ior": 1.5})  # core/material.h:36-76; basecolor black


def camera_to_c(c, out):
    """Fills an fm_camera_desc from a camera dict (the JSON camera block)."""
    out.type = A.FM_CAMERA_THIN_LENS if c["type"] == "thin_lens" else A.FM_CAMERA_PIN_HOLE
    for k in ("pos", "look_at", "up"):
        getattr(out, k)[:] = [float(x) for x in c[k]]
    out.fov, out.aspect = c["fov"], c["aspect"]
    out.focal_distance, out.lens_radius = c["focal_distance"], c.get("lens_radius", 0.0)
    return out


class SceneDesc:
    """Objects are kept in reference order; a triangle mesh expands to one
    geometry per face (triangle.cpp:312-329)."""

    def __init__(self):
        self.camera = {"type": "pin_hole", "pos": [0, 0, 0], "look_at": [0, 0, 0], "up": [0, 0, 1], "fov": 60.0,
                       "aspect": 1.0, "focal_distance": 1.0, "lens_radius": 0.0}
        self.renderer = {"width": 1, "height": 1, "spp": 10, "thread_count": 8, "depth": 10, "direct_loop": 1,
                         "rr_depth": 10, "rr_coef": 0.85, "output": "result"}
        self.accelerator = "bvh"
        self.post_process = []  # [{"type": "tonemapping", "lum": 1.0}, {"type": "linear_to_srgb"}, ...]
        self.images = []      # float32 HxWx3 linear (hdr) or uint8 HxWx3 sRGB
        self.materials = []   # dicts
        self.material_names = []
        self.meshes = []      # dicts of arrays
        self.objects = []
        self.lights = []

    # ---- builders ---------------------------------------------------------
    def add_image(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.ndim == 3 and arr.shape[2] == 3 and arr.dtype in (np.float32, np.uint8)
        self.images.append(arr)
        return len(self.images) - 1

    def add_lambert(self, basecolor=1.0, name=None):
        tex = basecolor if isinstance(basecolor, dict) else const_tex(basecolor)
        self.materials.append({"type": "lambert_diffuse", "basecolor": tex})
        self.material_names.append(name or "mat%d" % len(self.materials))
        return len(self.materials) - 1

    def add_dielectric(self, eta_i=1.0, eta_t=1.5, reflect_color=1.0, refract_color=1.0, name=None):
        self.materials.append({"type": "specular_transmission", "eta_i": float(eta_i), "eta_t": float(eta_t),
                               "reflect_color": np.broadcast_to(np.float32(reflect_color), (3,)).tolist(),
                               "refract_color": np.broadcast_to(np.float32(refract_color), (3,)).tolist()})
        self.material_names.append(name or "mat%d" % len(self.materials))
        return len(self.materials) - 1

    def add_conductor(self, color=1.0, eta_i=1.0, eta_t=(0.2, 0.92, 1.1), k=(3.9, 2.45, 2.14), name=None):
        """A conductor mirror (the reference's CreateSpecularReflection:
        colour, per-channel indices eta_i / eta_t and absorption k).  The
        scene JSON has no such material type: the reference loader, and this
        package's, warn and skip it; it exists for scenes built in code."""
        f3 = lambda v: np.broadcast_to(np.float32(v), (3,)).tolist()
        self.materials.append({"type": "specular_reflection", "reflect_color": f3(color), "eta_i": f3(eta_i),
                               "eta_t": f3(eta_t), "k": f3(k)})
        self.material_names.append(name or "mat%d" % len(self.materials))
        return len(self.materials) - 1

    def add_disney(self, thin=False, name=None, **params):
        m = {"type": "disney", "thin": bool(thin)}
        for k in A.DISNEY_PARAMS:
            v = params.get(k, None)
            if v is None:
                v = [0.0, 0.0, 0.0] if k == "basecolor" else DISNEY_DEFAULTS[k]
            m[k] = v if isinstance(v, dict) else const_tex(v)
        self.materials.append(m)
        self.material_names.append(name or "mat%d" % len(self.materials))
        return len(self.materials) - 1

    def add_mesh(self, positions, indices, normals=None, texcoords=None, normal_idx=None, tex_idx=None):
        """indices: (T,3) position indices; normal_idx/tex_idx default to the
        same indices when normals/texcoords are given, else -1."""
        positions = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        idx = np.asarray(indices, dtype=np.int32).reshape(-1, 3)
        full = np.full((idx.shape[0], 9), -1, dtype=np.int32)
        full[:, 0::3] = idx
        if normals is not None:
            normals = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
            full[:, 1::3] = idx if normal_idx is None else np.asarray(normal_idx, np.int32).reshape(-1, 3)
        if texcoords is not None:
            texcoords = np.ascontiguousarray(texcoords, dtype=np.float32).reshape(-1, 2)
            full[:, 2::3] = idx if tex_idx is None else np.asarray(tex_idx, np.int32).reshape(-1, 3)
        self.meshes.append({"positions": positions,
                            "normals": normals if normals is not None else np.zeros((0, 3), np.float32),
                            "texcoords": texcoords if texcoords is not None else np.zeros((0, 2), np.float32),
                            "indices": np.ascontiguousarray(full)})
        return len(self.meshes) - 1

    def _obj(self, shape, material=-1, emittance=0.0, **kw):
        o = {"shape": shape, "material": int(material), "emittance": np.broadcast_to(np.float32(emittance), (3,)).tolist(),
             "translation": [0.0, 0.0, 0.0], "rotation": [0.0, 0.0, 0.0], "scale": [1.0, 1.0, 1.0], "radius": 1.0,
             "mesh": -1}
        for k, v in kw.items():
            o[k] = v
        self.objects.append(o)
        return len(self.objects) - 1

    def add_rectangle(self, pa, pb, pc, pd, material=-1, emittance=0.0):
        return self._obj("rectangle", material, emittance, pa=list(map(float, pa)), pb=list(map(float, pb)),
                         pc=list(map(float, pc)), pd=list(map(float, pd)))

    def add_sphere(self, radius, translation=(0, 0, 0), rotation=(0, 0, 0), scale=(1, 1, 1), material=-1, emittance=0.0):
        return self._obj("sphere", material, emittance, radius=float(radius), translation=list(map(float, translation)),
                         rotation=list(map(float, rotation)), scale=list(map(float, scale)))

    def add_mesh_object(self, mesh, translation=(0, 0, 0), rotation=(0, 0, 0), scale=(1, 1, 1), material=-1,
                        emittance=0.0):
        return self._obj("triangle_mesh", material, emittance, mesh=int(mesh), translation=list(map(float, translation)),
                         rotation=list(map(float, rotation)), scale=list(map(float, scale)))

    def add_env(self, image, rotation=(0, 0, 0), factor=1.0):
        self.lights.append({"type": "env", "image": int(image), "rotation": list(map(float, rotation)),
                            "factor": float(factor)})

    def add_directional(self, radiance, direction):
        self.lights.append({"type": "directional", "radiance": list(map(float, np.broadcast_to(radiance, (3,)))),
                            "direction": list(map(float, direction))})

    def num_geometries(self):
        n = 0
        for o in self.objects:
            n += self.meshes[o["mesh"]]["indices"].shape[0] if o["shape"] == "triangle_mesh" else 1
        return n

    # ---- C descriptor -----------------------------------------------------
    def to_c(self):
        """Returns (fm_scene_desc, keepalive); keep both alive while in use."""
        keep = []
        d = A.fm_scene_desc()
        d.abi_version = A.FM_ABI_VERSION
        d.accelerator = {"naive": A.FM_ACCEL_NAIVE, "reference_tree": A.FM_ACCEL_BVH_REFERENCE_TREE,
                         "embree": A.FM_ACCEL_EMBREE}.get(self.accelerator, A.FM_ACCEL_BVH)
        camera_to_c(self.camera, d.camera)
        r = self.renderer
        for k in ("width", "height", "spp", "thread_count", "depth", "direct_loop", "rr_depth"):
            setattr(d.renderer, k, int(r[k]))
        d.renderer.rr_coef = r["rr_coef"]
        d.renderer.output = r["output"].encode()[:255]

        imgs = (A.fm_image_desc * max(1, len(self.images)))()
        for i, im in enumerate(self.images):
            imgs[i].width, imgs[i].height = im.shape[1], im.shape[0]
            imgs[i].format = A.FM_IMAGE_U8_SRGB if im.dtype == np.uint8 else A.FM_IMAGE_F32_LINEAR
            imgs[i].data = im.ctypes.data
            keep.append(im)
        mats = (A.fm_material_desc * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            _fill_material(mats[i], m)
        meshes = (A.fm_mesh_desc * max(1, len(self.meshes)))()
        for i, m in enumerate(self.meshes):
            for key in ("positions", "normals", "texcoords", "indices"):
                keep.append(m[key])
            meshes[i].num_positions = m["positions"].shape[0]
            meshes[i].positions = m["positions"].ctypes.data
            meshes[i].num_normals = m["normals"].shape[0]
            meshes[i].normals = m["normals"].ctypes.data if m["normals"].size else None
            meshes[i].num_texcoords = m["texcoords"].shape[0]
            meshes[i].texcoords = m["texcoords"].ctypes.data if m["texcoords"].size else None
            meshes[i].num_triangles = m["indices"].shape[0]
            meshes[i].indices = m["indices"].ctypes.data
        objs = (A.fm_object_desc * max(1, len(self.objects)))()
        shapes = {"triangle_mesh": A.FM_SHAPE_TRIANGLE_MESH, "sphere": A.FM_SHAPE_SPHERE, "rectangle": A.FM_SHAPE_RECTANGLE}
        for i, o in enumerate(self.objects):
            x = objs[i]
            x.shape = shapes[o["shape"]]
            x.material, x.mesh, x.radius = o["material"], o["mesh"], o["radius"]
            x.translation[:], x.rotation[:], x.scale[:] = o["translation"], o["rotation"], o["scale"]
            for k in ("pa", "pb", "pc", "pd"):
                if k in o:
                    getattr(x, k)[:] = o[k]
            x.emittance[:] = o["emittance"]
        lights = (A.fm_light_desc * max(1, len(self.lights)))()
        for i, l in enumerate(self.lights):
            if l["type"] == "env":
                lights[i].type, lights[i].image = A.FM_LIGHT_ENV, l["image"]
                lights[i].rotation[:] = l["rotation"]
                lights[i].factor = l["factor"]
            else:
                lights[i].type = A.FM_LIGHT_DIRECTIONAL
                lights[i].radiance[:] = l["radiance"]
                lights[i].direction[:] = l["direction"]
        posts = (A.fm_post_op * max(1, len(self.post_process)))()
        for i, pp in enumerate(self.post_process):
            posts[i].type = A.POST_TYPES[pp["type"]]
            posts[i].lum = float(pp.get("lum", 1.0))
        keep += [imgs, mats, meshes, objs, lights, posts]
        d.num_images, d.images = len(self.images), C.cast(imgs, C.POINTER(A.fm_image_desc))
        d.num_materials, d.materials = len(self.materials), C.cast(mats, C.POINTER(A.fm_material_desc))
        d.num_meshes, d.meshes = len(self.meshes), C.cast(meshes, C.POINTER(A.fm_mesh_desc))
        d.num_objects, d.objects = len(self.objects), C.cast(objs, C.POINTER(A.fm_object_desc))
        d.num_lights, d.lights = len(self.lights), C.cast(lights, C.POINTER(A.fm_light_desc))
        d.num_post_ops, d.post_ops = len(self.post_process), C.cast(posts, C.POINTER(A.fm_post_op))
        return d, keep

    # ---- files -------------------------------------------------------------
    def write(self, directory, name="scene.json"):
        """Writes a reference-loadable scene: JSON + one OBJ per mesh +
        images (.hdr for float images, .tga for 8-bit)."""
        os.makedirs(directory, exist_ok=True)
        img_files = []
        for i, im in enumerate(self.images):
            if im.dtype == np.float32:
                fn = "image%d.hdr" % i
                write_hdr(os.path.join(directory, fn), im)
            else:
                fn = "image%d.tga" % i
                write_tga(os.path.join(directory, fn), im)
            img_files.append(fn)

        def tex_json(t):
            if t["type"] == "constant":
                return {"type": "constant", "value": [_f(x) for x in t["value"]], "wrap_u": t["wrap_u"],
                        "wrap_v": t["wrap_v"]}
            return {"type": "image", "path": img_files[t["image"]], "sampler": t["sampler"], "wrap_u": t["wrap_u"],
                    "wrap_v": t["wrap_v"]}

        mats = []
        for m, nm in zip(self.materials, self.material_names):
            j = {"type": m["type"], "name": nm}
            if m["type"] == "lambert_diffuse":
                j["basecolor"] = tex_json(m["basecolor"])
            elif m["type"] == "disney":
                for k in A.DISNEY_PARAMS:
                    j[k] = tex_json(m[k])
                j["thin"] = m["thin"]
            elif m["type"] == "specular_reflection":
                j.update({"reflect_color": [_f(x) for x in m["reflect_color"]], "eta_i": [_f(x) for x in m["eta_i"]],
                          "eta_t": [_f(x) for x in m["eta_t"]], "k": [_f(x) for x in m["k"]]})
            else:
                j.update({"eta_i": _f(m["eta_i"]), "eta_t": _f(m["eta_t"]),
                          "reflect_color": [_f(x) for x in m["reflect_color"]],
                          "refract_color": [_f(x) for x in m["refract_color"]]})
            mats.append(j)
        objs = []
        for oi, o in enumerate(self.objects):
            sh = {"type": o["shape"]}
            if o["shape"] == "triangle_mesh":
                fn = "mesh%d.obj" % o["mesh"]
                if not os.path.exists(os.path.join(directory, fn)):
                    write_obj(os.path.join(directory, fn), self.meshes[o["mesh"]])
                sh.update({"path": fn, "translation": o["translation"], "rotation": o["rotation"], "scale": o["scale"]})
            elif o["shape"] == "sphere":
                sh.update({"radius": _f(o["radius"]), "translation": o["translation"], "rotation": o["rotation"],
                           "scale": o["scale"]})
            else:
                sh.update({k: [_f(x) for x in o[k]] for k in ("pa", "pb", "pc", "pd")})
            j = {"shape": sh, "emittance": [_f(x) for x in o["emittance"]]}
            j["material"] = self.material_names[o["material"]] if o["material"] >= 0 else "default_material"
            objs.append(j)
        lights = []
        for l in self.lights:
            if l["type"] == "env":
                lights.append({"type": "env", "path": img_files[l["image"]], "rotation": l["rotation"],
                               "factor": _f(l["factor"])})
            else:
                lights.append({"type": "directional", "radiance": l["radiance"], "direction": l["direction"]})
        cam = dict(self.camera)
        cam["pos"], cam["look_at"], cam["up"] = ([_f(x) for x in cam[k]] for k in ("pos", "look_at", "up"))
        if cam["type"] != "thin_lens":
            cam.pop("lens_radius", None)
        doc = {"camera": cam, "renderer": dict(self.renderer, type="path_tracer"),
               "scene": {"accelerator": {"type": self.accelerator}, "materials": mats, "render_objects": objs,
                         "lights": lights}}
        if self.post_process:
            doc["post_process"] = [dict(pp, **({"lum": _f(pp["lum"])} if "lum" in pp else {})) for pp in self.post_process]
        path = os.path.join(directory, name)
        with open(path, "w") as f:
            json.dump(doc, f, indent=1)
        return path


def _f(x):
    """float32 value written with enough digits to round-trip."""
    return float("%.9g" % np.float32(x))


def _fill_material(x, m):
    def tex(dst, t):
        if t["type"] == "constant":
            dst.type = A.FM_TEX_CONSTANT
            dst.value[:] = [float(v) for v in t["value"]]
        else:
            dst.type = A.FM_TEX_IMAGE
            dst.image = t["image"]
            dst.sampler = A.FM_SAMPLER_NEAREST if t["sampler"] == "nearest" else A.FM_SAMPLER_LINEAR
        dst.wrap_u = A.FM_WRAP_REPEAT if t["wrap_u"] == "repeat" else A.FM_WRAP_CLAMP
        dst.wrap_v = A.FM_WRAP_REPEAT if t["wrap_v"] == "repeat" else A.FM_WRAP_CLAMP

    for k in range(A.FM_D_NPARAM):
        tex(x.tex[k], const_tex(0.0))
    if m["type"] == "lambert_diffuse":
        x.type = A.FM_MAT_LAMBERT
        tex(x.tex[0], m["basecolor"])
    elif m["type"] == "disney":
        x.type = A.FM_MAT_DISNEY
        x.thin = 1 if m["thin"] else 0
        for k, key in enumerate(A.DISNEY_PARAMS):
            tex(x.tex[k], m[key])
    elif m["type"] == "specular_reflection":
        x.type = A.FM_MAT_SPECULAR_REFLECTION
        x.reflect_color[:] = m["reflect_color"]
        x.cond_eta_i[:] = m["eta_i"]
        x.cond_eta_t[:] = m["eta_t"]
        x.cond_k[:] = m["k"]
    else:
        x.type = A.FM_MAT_SPECULAR_TRANSMISSION
        x.eta_i, x.eta_t = m["eta_i"], m["eta_t"]
        x.reflect_color[:] = m["reflect_color"]
        x.refract_color[:] = m["refract_color"]


# ---- file writers ---------------------------------------------------------
def rgbe_encode(img):
    """Radiance RGBE of float RGB (flat records)."""
    img = np.asarray(img, dtype=np.float32)
    v = img.max(axis=2).astype(np.float64)
    m, e = np.frexp(v)
    scale = np.where(v < 1e-32, 0.0, m * 256.0 / np.where(v < 1e-32, 1.0, v))
    out = np.zeros(img.shape[:2] + (4,), np.uint8)
    out[..., :3] = np.clip(img * scale[..., None], 0, 255).astype(np.uint8)
    out[..., 3] = np.where(v < 1e-32, 0, e + 128).astype(np.uint8)
    return out


def rgbe_decode(rgbe):
    """mantissa * 2^(e - 136), e == 0 -> 0 (the stb_image .hdr conversion)."""
    e = rgbe[..., 3].astype(np.int32)
    f = np.where(e == 0, 0.0, np.ldexp(np.float32(1.0), e - 136)).astype(np.float32)
    return (rgbe[..., :3].astype(np.float32) * f[..., None]).astype(np.float32)


def write_hdr(path, img):
    h, w = img.shape[:2]
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n" % (h, w))
        f.write(rgbe_encode(img).tobytes())


def write_tga(path, img):
    h, w = img.shape[:2]
    hdr = bytearray(18)
    hdr[2] = 2
    hdr[12:14] = w.to_bytes(2, "little")
    hdr[14:16] = h.to_bytes(2, "little")
    hdr[16] = 24
    hdr[17] = 0x20  # top-left origin
    with open(path, "wb") as f:
        f.write(bytes(hdr))
        f.write(np.ascontiguousarray(img[..., ::-1]).tobytes())


def write_obj(path, mesh):
    """OBJ text: v / vn / vt records (%.9g round-trips float32) and faces
    with 1-based v/vt/vn indices, written a block of rows at a time."""
    p = mesh["positions"]
    idx = np.asarray(mesh["indices"], np.int64) + 1
    with open(path, "w") as f:
        if len(p):
            np.savetxt(f, np.asarray(p, np.float64), fmt="v %.9g %.9g %.9g")
        if mesh["normals"].size:
            np.savetxt(f, np.asarray(mesh["normals"], np.float64), fmt="vn %.9g %.9g %.9g")
        if mesh["texcoords"].size:
            np.savetxt(f, np.asarray(mesh["texcoords"], np.float64), fmt="vt %.9g %.9g")
        if not len(idx):
            return
        v, n, t = idx[:, 0::3], idx[:, 1::3], idx[:, 2::3]
        has_n, has_t = (n > 0).all(1), (t > 0).all(1)
        for hn, ht, fmt in ((True, True, "%d/%d/%d"), (True, False, "%d//%d"), (False, True, "%d/%d"),
                            (False, False, "%d")):
            rows = np.flatnonzero((has_n == hn) & (has_t == ht))
            if not len(rows):
                continue
            cols = [v[rows]] + ([t[rows]] if ht else []) + ([n[rows]] if hn else [])
            k = len(cols)
            arr = np.stack(cols, 2).reshape(len(rows), 3 * k)   # per corner: v[, vt][, vn]
            if len(rows) == len(idx):
                np.savetxt(f, arr, fmt="f " + " ".join([fmt] * 3))
            else:   # mixed faces keep their order: write them one by one
                for r, a in zip(rows, arr):
                    f.write("f " + " ".join([fmt] * 3) % tuple(a) + "\n")
