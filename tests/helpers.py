"""Shared inputs for the tests: seeded scenes, rays and BSDF queries."""
import numpy as np

from fumeng_amd import _abi as A
from fumeng_amd import scenes
from fumeng_amd.scene import SceneDesc


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def rays_toward(origins, targets, tmin=0.0, tmax=3.4028235e38):
    o = np.asarray(origins, np.float32)
    d = unit(np.asarray(targets, np.float64) - o).astype(np.float32)
    r = np.zeros((o.shape[0], 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def random_sphere_dirs(rng, n):
    return unit(rng.standard_normal((n, 3)))


def mesh_scene(freq=12, seed=5, with_box=True, with_spheres=True, accelerator="bvh"):
    """Displaced icosphere inside the box scene: triangles, rectangles and
    spheres in one BVH."""
    s = scenes.box_scene(width=48, height=48, spp=8, accelerator=accelerator) if with_box else SceneDesc()
    if not with_spheres:
        s.objects = [o for o in s.objects if o["shape"] != "sphere"]
    pos, tris = scenes.displaced_icosphere(freq, seed)
    pos = pos * 0.35 + np.array([0.0, 0.2, 0.1], np.float32)
    m = s.add_mesh(pos, tris)
    d = s.add_disney(basecolor=[0.6, 0.5, 0.4], roughness=0.4, metallic=0.2, clearcoat=0.5, sheen=0.3)
    s.add_mesh_object(m, material=d, rotation=(10.0, 20.0, 30.0))
    return s


def scene_rays(rng, n, spread=1.2):
    """Rays from random points inside the box toward random targets."""
    o = rng.uniform(-0.95, 0.95, (n, 3))
    t = rng.uniform(-spread, spread, (n, 3))
    return rays_toward(o, t)


def frame_from_normal(n):
    n = unit(n)
    a = np.where(np.abs(n[:, :1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    ss = unit(np.cross(a, n))
    return n, ss


def bsdf_queries(rng, n, material, grazing=False):
    q = np.zeros(n, dtype=np.dtype(A.fm_bsdf_query))
    nrm = random_sphere_dirs(rng, n)
    nrm, ss = frame_from_normal(nrm)
    wo = random_sphere_dirs(rng, n)
    wi = random_sphere_dirs(rng, n)
    if not grazing:
        # keep away from the shading horizon where the reference terms are singular
        cos_o = np.sum(wo * nrm, 1)
        wo = np.where(np.abs(cos_o)[:, None] < 0.05, unit(wo + 0.2 * np.sign(cos_o + 1e-9)[:, None] * nrm), wo)
    q["material"] = material
    q["ng"] = nrm
    q["ns"] = nrm
    q["ss"] = ss
    q["wo"] = wo
    q["wi"] = wi
    q["uv"] = rng.random((n, 2))
    q["sample"] = rng.random((n, 3)) * 0.999999
    return q


def convergence_scene(name):
    """Scenes of the FP64 image-convergence fixtures (tests/golden)."""
    if name == "box32":
        return scenes.box_scene(width=32, height=32, spp=64)
    if name == "ico48x27":
        return scenes.icosphere_scene(freq=24, width=48, height=27, spp=64, env_size=(256, 128))
    raise KeyError(name)


def px_scene(name):
    """Full-size frames of the pixel-subset FP64 references (tests/golden)."""
    if name == "ico1080":
        return scenes.icosphere_scene()
    if name == "stress1080":
        return scenes.stress_scene()
    raise KeyError(name)


def scene_signature(s):
    """A number that changes when the generated geometry, images, materials
    or renderer settings change."""
    acc = 0.0
    for k, m in enumerate(s.meshes):
        for key in sorted(m):
            v = m[key]
            if isinstance(v, np.ndarray) and v.size:
                acc += float(np.sum(v.astype(np.float64) * (1.0 + 1e-3 * (k % 97))))
    for im in s.images:
        acc += float(np.sum(im.astype(np.float64))) * 1e-3
    acc += float(len(s.objects)) * 7.0 + float(len(s.materials)) * 11.0 + float(len(s.lights)) * 13.0
    for key in ("width", "height", "depth", "rr_depth", "direct_loop"):
        acc += float(s.renderer.get(key, 0)) * 0.5
    return acc


def load_reference(name, prec="fp64"):
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "%s_%s.npz" % (prec, name))
    with np.load(path) as f:
        return {k: f[k] for k in f.files}


def convergence_stats(mean, sumsq, n, ref):
    """relMSE = mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2);
    per-pixel z = (x - r) / sqrt(s_x^2 / n + s_r^2 / n_r) with both sample
    variances from the sums of squares (channels with zero variance on both
    sides and equal means count as z = 0)."""
    nr = float(ref["spp"])
    x = np.asarray(mean, np.float64)
    q = np.asarray(sumsq, np.float64)
    # the reference estimator has no NaN guard: a path whose direction has an
    # exact zero component can pick up a NaN hit; such pixels are excluded
    finite = np.all(np.isfinite(x), axis=-1) & np.all(np.isfinite(ref["mean"]), axis=-1)
    x, q = x[finite], q[finite]
    r, rq = ref["mean"][finite], ref["sumsq"][finite]
    vx = np.maximum(q / n - x * x, 0.0) * n / max(n - 1.0, 1.0)
    vr = np.maximum(rq / nr - r * r, 0.0) * nr / (nr - 1.0)
    se = np.sqrt(vx / n + vr / nr)
    diff = x - r
    z = np.where(se > 0, diff / np.where(se > 0, se, 1.0), np.where(np.abs(diff) <= 1e-12, 0.0, np.inf))
    rel = float(np.mean(diff * diff / (r * r + 1e-2)))
    return {"relmse": rel, "mean_z": float(np.mean(z[np.isfinite(z)])), "frac_z3": float(np.mean(np.abs(z) > 3.0)),
            "max_abs_z": float(np.max(np.abs(z))), "nonfinite_pixels": int(np.sum(~finite)), "z": z}


def two_planes_scene(depth=3, rough=0.4, metal=0.3, width=16, height=16):
    """A Disney floor and a Disney wall under a sky env with a soft sun: a
    smooth test bed for parameter derivatives (no small geometry whose
    visibility jumps under finite differences)."""
    s = SceneDesc()
    s.camera = {"type": "pin_hole", "pos": [0.0, -3.0, 1.2], "look_at": [0.0, 0.0, 0.0], "up": [0.0, 0.0, 1.0],
                "fov": 40.0, "aspect": float(width) / height, "focal_distance": 1.0, "lens_radius": 0.0}
    s.renderer.update(width=width, height=height, spp=8, depth=depth, rr_depth=10)
    m0 = s.add_disney(basecolor=[0.6, 0.5, 0.3], roughness=rough, metallic=metal, specular=0.5, clearcoat=0.3, sheen=0.2)
    m1 = s.add_disney(basecolor=[0.3, 0.6, 0.4], roughness=0.5, metallic=0.1, specular=0.5)
    floor = s.add_mesh(np.array([[-2, -2, 0], [2, -2, 0], [2, 2, 0], [-2, 2, 0]], np.float32), np.array([[0, 1, 2], [0, 2, 3]]))
    wall = s.add_mesh(np.array([[-2, 1, 0], [2, 1, 0], [2, 1, 3], [-2, 1, 3]], np.float32), np.array([[0, 2, 1], [0, 3, 2]]))
    s.add_mesh_object(floor, material=m0)
    s.add_mesh_object(wall, material=m1)
    s.add_env(s.add_image(scenes.sky_env(64, 32, 3, sun_peak=50.0, sun_sigma_deg=8.0)))
    return s


def mesh_light_scene(accelerator="embree", width=16, height=16, spp=64, depth=3):
    """A Lambert floor (z = 0) under an emissive triangle mesh facing down:
    a 3 x 2 grid of quads of unequal sizes (12 faces), so that sampling the
    mesh by area differs from sampling each face as its own light."""
    s = SceneDesc()
    s.accelerator = accelerator
    s.camera = {"type": "pin_hole", "pos": [0.3, -2.5, 2.0], "look_at": [0, 0, 0.2], "up": [0, 0, 1], "fov": 45.0,
                "aspect": width / height, "focal_distance": 1.0, "lens_radius": 0.0}
    s.renderer.update(width=width, height=height, spp=spp, depth=depth, rr_depth=10)
    floor = s.add_lambert(0.6)
    s.add_rectangle((-3, -3, 0), (3, -3, 0), (3, 3, 0), (-3, 3, 0), floor)
    xs = np.array([-0.8, -0.5, 0.1, 0.9], np.float32)
    ys = np.array([-0.4, 0.0, 0.7], np.float32)
    pos = np.array([[x, y, 1.2] for y in ys for x in xs], np.float32)
    idx = []
    for j in range(len(ys) - 1):
        for i in range(len(xs) - 1):
            a, b, c, d = j * 4 + i, j * 4 + i + 1, (j + 1) * 4 + i + 1, (j + 1) * 4 + i
            idx += [[a, c, b], [a, d, c]]      # wound so that cross(v0 - v2, v1 - v2) points down
    m = s.add_mesh(pos, np.array(idx))
    s.add_mesh_object(m, emittance=[6.0, 5.0, 4.0])
    return s


def crn_sample_agreement(g, o, width, height, spp, seed, mode, tol=1e-3, threads=8):
    """Common random numbers, sample by sample: the fraction of (pixel,
    sample) radiances of the GPU render `g` that equal the oracle's `o` to
    `tol` relative (both NaN counts as equal).  Each sample is rendered as a
    1-spp frame at its own sample offset, so the figure does not depend on
    spp the way a per-pixel mean does."""
    ok = []
    for k in range(spp):
        cg = g.render(width=width, height=height, spp=1, seed=seed, mode=mode, sample_offset=k)[0]
        co = o.render(0, 0, width, height, spp=1, seed=seed, sample_offset=k, width=width, height=height,
                      threads=threads)[0]
        same = np.abs(cg - co) <= tol * np.maximum(1.0, np.abs(co))
        ok.append(np.all(same | (np.isnan(cg) & np.isnan(co)), axis=2))
    return float(np.mean(ok))
