"""Seeded procedural scenes for the benchmark configurations.

box_scene       open 2x2x2 box, 5 Lambert walls, 0.5x0.5 ceiling light,
                Lambert + dielectric spheres (512x512, 64 spp, depth 8, rr 3)
displaced_icosphere / icosphere_scene
                geodesic icosphere at frequency f (20 f^2 triangles) displaced
                by 6-octave value-noise fbm, ground plane, 64 random Disney
                materials, analytic sky + sun HDR env, thin-lens camera
                (f = 224: 1,003,520 triangles, 1920x1080, 256 spp, depth 12)
All values are float32 and fully determined by the seed.
"""
import numpy as np

from .scene import SceneDesc, image_tex, rgbe_decode, rgbe_encode


def box_scene(width=512, height=512, spp=64, depth=8, rr_depth=3, rr_coef=0.85, accelerator="bvh"):
    s = SceneDesc()
    s.camera = {"type": "pin_hole", "pos": [0.0, -3.75, 0.0], "look_at": [0.0, 0.0, 0.0], "up": [0.0, 0.0, 1.0],
                "fov": 40.0, "aspect": 1.0, "focal_distance": 1.0, "lens_radius": 0.0}
    s.renderer = {"width": width, "height": height, "spp": spp, "thread_count": 8, "depth": depth, "direct_loop": 1,
                  "rr_depth": rr_depth, "rr_coef": rr_coef, "output": "box"}
    s.accelerator = accelerator
    white = s.add_lambert(0.73, name="white")
    red = s.add_lambert([0.65, 0.05, 0.05], name="red")
    green = s.add_lambert([0.12, 0.45, 0.15], name="green")
    glass = s.add_dielectric(1.0, 1.5, name="glass")
    # normals point into the box: normalize((pb - pa) x (pc - pa))
    s.add_rectangle((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), white)   # floor, +z
    s.add_rectangle((-1, -1, 1), (-1, 1, 1), (1, 1, 1), (1, -1, 1), white)       # ceiling, -z
    s.add_rectangle((-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1), white)       # back, -y
    s.add_rectangle((1, -1, -1), (1, -1, 1), (1, 1, 1), (1, 1, -1), red)         # x = +1, -x (image left)
    s.add_rectangle((-1, -1, -1), (-1, 1, -1), (-1, 1, 1), (-1, -1, 1), green)   # x = -1, +x (image right)
    z = 1.0 - 1e-3
    s.add_rectangle((-0.25, -0.25, z), (-0.25, 0.25, z), (0.25, 0.25, z), (0.25, -0.25, z), white,
                    emittance=[17.0, 12.0, 4.0])
    s.add_sphere(0.3, translation=(0.4, 0.3, -0.7), material=white)
    s.add_sphere(0.3, translation=(-0.4, -0.2, -0.7), material=glass)
    return s


# ---- noise -------------------------------------------------------------------
def _hash_u32(*xs):
    h = np.uint32(0x9E3779B9)
    out = np.full(np.broadcast(*xs).shape, h, dtype=np.uint32)
    for x in xs:
        v = np.asarray(x).astype(np.int64).astype(np.uint32)
        out = (out ^ v) * np.uint32(0x85EBCA6B)
        out ^= out >> np.uint32(13)
        out = out * np.uint32(0xC2B2AE35)
        out ^= out >> np.uint32(16)
    return out


def value_noise(p, seed, octave):
    """3D lattice value noise in [-1, 1] with smoothstep interpolation."""
    f = np.floor(p)
    i = f.astype(np.int64)
    t = p - f
    t = t * t * (3.0 - 2.0 * t)
    acc = 0.0
    for dx in (0, 1):
        wx = t[:, 0] if dx else 1.0 - t[:, 0]
        for dy in (0, 1):
            wy = t[:, 1] if dy else 1.0 - t[:, 1]
            for dz in (0, 1):
                wz = t[:, 2] if dz else 1.0 - t[:, 2]
                h = _hash_u32(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, np.int64(octave), np.int64(seed))
                val = (h.astype(np.float64) / 4294967295.0) * 2.0 - 1.0
                acc = acc + wx * wy * wz * val
    return acc


def fbm(p, seed, octaves=6, lacunarity=2.0, gain=0.5, base_freq=2.0):
    amp, freq, total, norm = 1.0, base_freq, 0.0, 0.0
    for o in range(octaves):
        total = total + amp * value_noise(p * freq, seed, o)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


# ---- icosphere -----------------------------------------------------------------
def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
                  [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v, f


def geodesic_sphere(freq):
    """Class-I geodesic subdivision: 20 * freq^2 triangles, shared vertices."""
    v, faces = _icosahedron()
    # lattice points (i, j) with i + j <= freq on each face: p = (A (f-i-j) + B i + C j) / f
    ii, jj = np.meshgrid(np.arange(freq + 1), np.arange(freq + 1), indexing="ij")
    m = (ii + jj) <= freq
    ii, jj = ii[m], jj[m]
    kk = freq - ii - jj
    A, B, Cc = v[faces[:, 0]], v[faces[:, 1]], v[faces[:, 2]]
    pts = (A[:, None, :] * kk[None, :, None] + B[:, None, :] * ii[None, :, None] + Cc[:, None, :] * jj[None, :, None])
    # integer key of a lattice point: the barycentric weights attached to global corner ids
    # (points on shared edges / corners get the same key from every face)
    npf = ii.size
    w = np.stack([np.broadcast_to(kk, (20, npf)), np.broadcast_to(ii, (20, npf)), np.broadcast_to(jj, (20, npf))], -1)
    cid = np.broadcast_to(faces[:, None, :], (20, npf, 3))
    wk = np.where(w > 0, cid, 12)            # corner id, 12 = unused
    wv = np.where(w > 0, w, 0)
    order = np.argsort(wk, axis=-1)
    wk = np.take_along_axis(wk, order, -1)
    wv = np.take_along_axis(wv, order, -1)
    key = ((wk[..., 0] * 13 + wk[..., 1]) * 13 + wk[..., 2]) * (freq + 1) ** 2 + wv[..., 0] * (freq + 1) + wv[..., 1]
    key = key.reshape(-1)
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    verts = pts.reshape(-1, 3)[first] / freq
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    # triangles of the lattice
    lid = -np.ones((freq + 1, freq + 1), dtype=np.int64)
    lid[ii, jj] = np.arange(npf)
    tris = []
    for i in range(freq):
        for j in range(freq - i):
            tris.append((lid[i, j], lid[i + 1, j], lid[i, j + 1]))
            if i + j < freq - 1:
                tris.append((lid[i + 1, j], lid[i + 1, j + 1], lid[i, j + 1]))
    tris = np.array(tris, dtype=np.int64)
    all_tris = (tris[None, :, :] + (np.arange(20) * npf)[:, None, None]).reshape(-1, 3)
    return verts, inv[all_tris].astype(np.int32)


def displaced_icosphere(freq, seed, amplitude=0.15, radius=1.0):
    verts, tris = geodesic_sphere(freq)
    d = fbm(verts * 1.0, seed)
    d = d / max(1e-9, np.abs(d).max())
    pos = verts * (radius + amplitude * d)[:, None]
    return pos.astype(np.float32), tris


def sky_env(width, height, seed, sun_theta_deg=45.0, sun_phi_deg=60.0, sun_sigma_deg=1.0, sun_peak=5e4, sigma=0.3):
    """z-up lat-long float RGB: sky gradient + Gaussian sun lobe, log-normal
    texel noise, quantised through RGBE like a .hdr file."""
    rng = np.random.default_rng(seed)
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    th = np.pi * v[:, None]
    ph = 2 * np.pi * u[None, :]
    x, y, z = np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th) * np.ones_like(ph)
    zen = np.array([0.25, 0.45, 0.95]) * 0.9
    hor = np.array([0.95, 0.95, 1.0]) * 1.2
    gnd = np.array([0.35, 0.30, 0.25]) * 0.4
    zc = np.clip(z, 0, 1)[..., None]
    sky = np.where(z[..., None] >= 0, hor * (1 - np.sqrt(zc)) + zen * np.sqrt(zc), gnd)
    st, sp = np.radians(sun_theta_deg), np.radians(sun_phi_deg)
    sd = np.array([np.sin(st) * np.cos(sp), np.sin(st) * np.sin(sp), np.cos(st)])
    cosang = np.clip(x * sd[0] + y * sd[1] + z * sd[2], -1, 1)
    ang = np.arccos(cosang)
    sun = sun_peak * np.exp(-0.5 * (ang / np.radians(sun_sigma_deg)) ** 2)
    img = sky + sun[..., None] * np.array([1.0, 0.95, 0.88])
    img = img * np.exp(sigma * rng.standard_normal((height, width)))[..., None]
    return rgbe_decode(rgbe_encode(img.astype(np.float32)))


def random_disney(s, n, seed, prefix="d"):
    rng = np.random.default_rng(seed)
    ids = []
    for k in range(n):
        r = rng.random(8)
        ids.append(s.add_disney(name="%s%d" % (prefix, k), basecolor=[float(r[0]), float(r[1]), float(r[2])],
                                roughness=float(r[3]), metallic=float(r[4]), specular=float(r[5]),
                                clearcoat=float(r[6]), sheen=float(r[7])))
    return ids


def icosphere_scene(freq=224, width=1920, height=1080, spp=256, depth=12, seed=2, env_size=(2048, 1024),
                    n_materials=64):
    s = SceneDesc()
    s.camera = {"type": "thin_lens", "pos": [0.0, -4.0, 0.3], "look_at": [0.0, 0.0, 0.0], "up": [0.0, 0.0, 1.0],
                "fov": 40.0, "aspect": 16.0 / 9.0, "focal_distance": 3.0, "lens_radius": 0.05}
    # rr keys absent in the JSON: defaults rr_depth 10, rr_coef 0.85
    s.renderer = {"width": width, "height": height, "spp": spp, "thread_count": 8, "depth": depth, "direct_loop": 1,
                  "rr_depth": 10, "rr_coef": 0.85, "output": "icosphere"}
    mats = random_disney(s, n_materials, seed)
    ground_mat = s.add_lambert(0.4, name="ground")
    pos, tris = displaced_icosphere(freq, seed)
    tid = np.arange(tris.shape[0])
    mid = (_hash_u32(tid, np.int64(seed)) % np.uint32(n_materials)).astype(np.int64)
    for k in range(n_materials):
        sel = tris[mid == k]
        if sel.shape[0] == 0:
            continue
        used, local = np.unique(sel.reshape(-1), return_inverse=True)
        m = s.add_mesh(pos[used], local.reshape(-1, 3))
        s.add_mesh_object(m, material=mats[k])
    g = s.add_mesh(np.array([[-10, -10, -1.2], [10, -10, -1.2], [10, 10, -1.2], [-10, 10, -1.2]], np.float32),
                   np.array([[0, 1, 2], [0, 2, 3]], np.int32))
    s.add_mesh_object(g, material=ground_mat)
    env = s.add_image(sky_env(env_size[0], env_size[1], seed))
    s.add_env(env)
    return s


def batch_jobs(n=64, seed=4, radius=4.0):
    """Uneven renders of one scene: seeded camera positions on a sphere of
    the given radius looking at the origin (pin-hole, 40 degrees), resolution
    r x r with r ~ U{128..1024}, spp ~ U{16..256}, depth ~ U{4..16}.  Job
    dicts for Scene.render_batch."""
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        v = rng.standard_normal(3)
        v /= np.linalg.norm(v)
        up = [0.0, 0.0, 1.0] if abs(v[2]) < 0.99 else [0.0, 1.0, 0.0]
        r = int(rng.integers(128, 1025))
        spp = int(rng.integers(16, 257))
        depth = int(rng.integers(4, 17))
        cam = {"type": "pin_hole", "pos": (radius * v).tolist(), "look_at": [0.0, 0.0, 0.0], "up": up, "fov": 40.0,
               "aspect": 1.0, "focal_distance": 1.0, "lens_radius": 0.0}
        jobs.append({"camera": cam, "width": r, "height": r, "spp": spp, "depth": depth, "seed": 1000 + i})
    return jobs


def inverse_scene(freq=71, width=256, height=256, spp=32, depth=6, seed=5, env_size=(512, 256), n_materials=64):
    """Inverse-rendering workload: the displaced-icosphere geometry at the
    given frequency (71: 100,820 triangles + ground), n_materials Disney
    materials, an env map of env_size, a square image, depth 6."""
    s = icosphere_scene(freq=freq, width=width, height=height, spp=spp, depth=depth, seed=seed, env_size=env_size,
                        n_materials=n_materials)
    s.camera = dict(s.camera, aspect=float(width) / float(height))
    return s


def perturb_params(material_params, env_texels, seed=5, amount=0.25, env_sigma=0.5):
    """A seeded start point away from the given parameters: Disney values
    moved by U(-amount, amount) and clipped to [0.02, 0.98] (NaN entries
    kept), env texels scaled by exp(env_sigma N(0, 1)) on an 8 x 4 grid of
    blocks (smooth, so the image changes visibly)."""
    rng = np.random.default_rng(seed)
    mp = np.array(material_params, np.float32, copy=True)
    ok = np.isfinite(mp)
    mp[ok] = np.clip(mp[ok] + rng.uniform(-amount, amount, mp.shape)[ok], 0.02, 0.98).astype(np.float32)
    env = None
    if env_texels is not None:
        H, W = env_texels.shape[:2]
        blocks = np.exp(env_sigma * rng.standard_normal((4, 8))).astype(np.float32)
        scale = np.kron(blocks, np.ones((H // 4 + 1, W // 8 + 1), np.float32))[:H, :W]
        env = (env_texels * scale[:, :, None]).astype(np.float32)
    return mp, env


def noise_texture(size, seed, cells=6, octaves=2, block=4):
    """Seeded smooth colour texture: fBm of value noise with smoothstep
    interpolation, computed at size / block and replicated block x block,
    as uint8 sRGB RGB (size x size x 3)."""
    rng = np.random.default_rng(seed)
    full = size
    size = max(1, size // block)
    img = np.zeros((size, size, 3), np.float32)
    amp, total = 1.0, 0.0
    for o in range(octaves):
        n = cells * 2 ** o
        g = rng.random((n + 1, n + 1, 3)).astype(np.float32)
        x = np.arange(size, dtype=np.float32) * (n / size)
        i0 = x.astype(np.int32)
        f = x - i0
        f = (f * f * (3.0 - 2.0 * f))[:, None]
        rows = g[i0] * (1.0 - f)[:, :, None] + g[i0 + 1] * f[:, :, None]          # (size, n + 1, 3)
        img += amp * (rows[:, i0] * (1.0 - f.T)[:, :, None] + rows[:, i0 + 1] * f.T[:, :, None])
        total += amp
        amp *= 0.5
    img /= total
    lo = rng.random(3).astype(np.float32) * 0.3
    hi = 0.5 + rng.random(3).astype(np.float32) * 0.5
    small = np.clip((lo + (hi - lo) * img) * 255.0 + 0.5, 0, 255).astype(np.uint8)
    if size * block != full or block == 1:
        return np.ascontiguousarray(np.resize(np.repeat(np.repeat(small, block, 0), block, 1), (full, full, 3)))
    return np.ascontiguousarray(np.repeat(np.repeat(small, block, 0), block, 1))


def stress_scene(n_spheres=400, freq=35, n_textures=200, tex_size=2048, width=1920, height=1080, spp=128, depth=10,
                 seed=3, n_lights=10, ground=True):
    """Large-geometry stress scene: n_spheres displaced geodesic spheres
    (freq 35: 24,500 triangles each, 9.8 M in total) with centres uniform
    in a 40 x 40 area at z = r, r ~ U[0.3, 1.2], displaced by 0.15 r;
    n_lights 2 x 2 rectangle lights at z = 10 facing down with grey radiance
    U[5, 20]; one Disney material per sphere drawn like the icosphere's,
    the even spheres' baseColor an n_textures-image set of tex_size^2
    seeded colour noise textures (spherical UVs); a 60 x 60 Lambert ground;
    pin-hole camera at (0, -30, 15) looking at the origin, fov 50;
    1920 x 1080, 128 spp, depth 10 (Russian roulette inactive)."""
    rng = np.random.default_rng(seed)
    s = SceneDesc()
    s.camera = {"type": "pin_hole", "pos": [0.0, -30.0, 15.0], "look_at": [0.0, 0.0, 0.0], "up": [0.0, 0.0, 1.0],
                "fov": 50.0, "aspect": float(width) / height, "focal_distance": 1.0, "lens_radius": 0.0}
    s.renderer = {"width": width, "height": height, "spp": spp, "thread_count": 8, "depth": depth, "direct_loop": 1,
                  "rr_depth": 10, "rr_coef": 0.85, "output": "stress"}
    verts, tris = geodesic_sphere(freq)
    uv = np.stack([0.5 + np.arctan2(verts[:, 1], verts[:, 0]) / (2.0 * np.pi),
                   np.arccos(np.clip(verts[:, 2], -1.0, 1.0)) / np.pi], 1).astype(np.float32)
    tex_ids = [s.add_image(noise_texture(tex_size, seed * 1000 + t)) for t in range(n_textures)]
    for k in range(n_spheres):
        r = float(rng.uniform(0.3, 1.2))
        cx, cy = rng.uniform(-20.0, 20.0, 2)
        d = fbm(verts * 1.0, seed * 7919 + k)
        d = d / max(1e-9, np.abs(d).max())
        pos = (verts * (r * (1.0 + 0.15 * d))[:, None] + np.array([cx, cy, r])).astype(np.float32)
        m = s.add_mesh(pos, tris, texcoords=uv)
        q = rng.random(8)
        base = [float(q[0]), float(q[1]), float(q[2])]
        if k % 2 == 0 and k // 2 < n_textures:
            base = image_tex(tex_ids[k // 2])
        mat = s.add_disney(name="sphere%d" % k, basecolor=base, roughness=float(q[3]), metallic=float(q[4]),
                           specular=float(q[5]), clearcoat=float(q[6]), sheen=float(q[7]))
        s.add_mesh_object(m, material=mat)
    if ground:
        gm = s.add_lambert(0.5, name="ground")
        g = s.add_mesh(np.array([[-30, -30, 0], [30, -30, 0], [30, 30, 0], [-30, 30, 0]], np.float32),
                       np.array([[0, 1, 2], [0, 2, 3]], np.int32))
        s.add_mesh_object(g, material=gm)
    black = s.add_lambert(0.0, name="light")
    for li in range(n_lights):
        x, y = rng.uniform(-18.0, 18.0, 2)
        L = float(rng.uniform(5.0, 20.0))
        s.add_rectangle((x - 1, y - 1, 10.0), (x - 1, y + 1, 10.0), (x + 1, y + 1, 10.0), (x + 1, y - 1, 10.0),
                        material=black, emittance=L)
    return s
