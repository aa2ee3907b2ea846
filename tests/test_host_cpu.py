"""Host-side checks that need no GPU: the C ABI library loads and exports
every declared symbol, the ctypes mirror has the C layouts, the scene JSON
loader follows the reference loader's semantics, and the product's restated
BVH builder reproduces the reference topology."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import fumeng_amd
import oracle
from fumeng_amd import _abi as A
from fumeng_amd import scenes
from fumeng_amd.scene import SceneDesc, rgbe_decode, rgbe_encode

from helpers import mesh_scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fm_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fm_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = C.CDLL(fumeng_amd.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(A.EXPORTS), set(names) ^ set(A.EXPORTS)


def test_struct_layouts_match_c(tmp_path):
    structs = ["fm_texture_desc", "fm_image_desc", "fm_material_desc", "fm_mesh_desc", "fm_object_desc",
               "fm_light_desc", "fm_camera_desc", "fm_renderer_desc", "fm_scene_desc", "fm_scene_info",
               "fm_render_params", "fm_film", "fm_render_stats", "fm_ray", "fm_hit", "fm_bsdf_query",
               "fm_bsdf_result", "fm_light_query", "fm_light_result", "fm_job", "fm_batch_stats", "fm_post_op",
               "fm_tune_trial", "fm_tune_result"]
    src = ['#include "%s"' % HEADER, "#include <stdio.h>", "#include <stddef.h>", "int main(void){"]
    for s in structs:
        src.append('printf("%s %%zu\\n", sizeof(%s));' % (s, s))
        for f, _ in getattr(A, s)._fields_:
            src.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (s, f, s, f))
    src.append("return 0;}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", str(c), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    for s in structs:
        T = getattr(A, s)
        assert int(out[s]) == C.sizeof(T), s
        for f, _ in T._fields_:
            assert int(out[s + "." + f]) == getattr(T, f).offset, s + "." + f


def test_init_reports_missing_device_or_succeeds():
    h = C.c_void_p()
    rc = fumeng_amd.lib().fm_init(0, C.byref(h))
    assert rc in (A.FM_OK, A.FM_ERR_NO_DEVICE)
    if rc == A.FM_OK:
        fumeng_amd.lib().fm_shutdown(h)
    else:
        assert b"device" in fumeng_amd.lib().fm_last_error(None)


@pytest.mark.parametrize("freq", [6, 14])
def test_host_builder_equals_reference_topology(freq):
    s = mesh_scene(freq=freq)
    ours = fumeng_amd.host_bvh(s)
    ref = oracle.OracleScene(s).bvh()
    for a, b in zip(ours, ref):
        assert np.array_equal(a, b)


def test_automatic_width_follows_the_last_level_cache():
    """FM_BVH_WIDTH_AUTO (the default): a scene whose two-wide tree and
    primitive records (64 B per node, 48 B per slot) fit in 256 MiB keeps
    the two-wide tree; a 5.8 M-triangle grid, whose records alone take
    277 MB, gets the four-wide one.  Either way the tree is the one that
    width gives when asked for."""
    small = mesh_scene(freq=20)
    a, b = fumeng_amd.host_scene_info(small, 0), fumeng_amd.host_scene_info(small, 2)
    assert a.trav_width == 2 and a.trav_hash == b.trav_hash
    assert b.num_trav_nodes * 64 + b.num_geometries * 48 < (256 << 20)
    n = 1700
    g = np.arange(n + 1, dtype=np.float32)
    x, y = np.meshgrid(g, g, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), np.zeros(x.size, np.float32)], 1).astype(np.float32)
    k = (np.arange(n)[:, None] * (n + 1) + np.arange(n)[None, :]).ravel()
    tris = np.concatenate([np.stack([k, k + n + 1, k + 1], 1), np.stack([k + 1, k + n + 1, k + n + 2], 1)]).astype(np.int32)
    big = SceneDesc()
    big.camera = dict(small.camera)
    big.renderer = dict(small.renderer)
    big.add_mesh_object(big.add_mesh(pos, tris), material=big.add_lambert(0.5))
    c = fumeng_amd.host_scene_info(big, 0)
    assert c.num_geometries == 2 * n * n and c.num_geometries * 48 > (256 << 20)
    assert c.trav_width == 4 and c.trav_node_bytes == 128


@pytest.mark.parametrize("name", ["mesh", "box", "stress_small"])
def test_host_build_of_both_traversal_widths(name):
    """The whole host build (no device) for every tree width: a two-wide
    traversal pushes at most one entry per level (stack = depth - 1); the
    four- and eight-wide trees merge nodes (fewer of them, each of 128 or
    256 bytes); the four-wide stack bound stays within three entries per
    two binary levels, the eight-wide one within seven per level."""
    s = {"mesh": lambda: mesh_scene(freq=20), "box": lambda: scenes.box_scene(width=16, height=16),
         "stress_small": lambda: scenes.stress_scene(n_spheres=6, freq=6, n_textures=2, tex_size=32, width=16,
                                                     height=16)}[name]()
    i2 = fumeng_amd.host_scene_info(s, 2)
    i4 = fumeng_amd.host_scene_info(s, 4)
    i8 = fumeng_amd.host_scene_info(s, 8)
    assert i2.device_bytes == 0 and i2.num_geometries == i4.num_geometries == i8.num_geometries > 0
    assert (i2.trav_width, i2.trav_node_bytes, i4.trav_width, i4.trav_node_bytes) == (2, 64, 4, 128)
    assert (i8.trav_width, i8.trav_node_bytes) == (8, 256)
    assert i2.num_trav_nodes == i2.num_geometries - 1 or i2.num_trav_nodes < i2.num_geometries
    assert 0 < i8.num_trav_nodes <= i4.num_trav_nodes <= i2.num_trav_nodes
    depth = i2.trav_stack + 1
    assert 1 <= i4.trav_stack <= 3 * ((depth + 1) // 2)
    assert 1 <= i8.trav_stack <= min(64, 7 * (depth - 1))   # every wide node on a path spans >= 1 binary level
    with pytest.raises(fumeng_amd.FumengError):
        fumeng_amd.host_scene_info(s, 3)


def _load_desc(path):
    h = C.c_void_p()
    dp = C.POINTER(A.fm_scene_desc)()
    err = C.create_string_buffer(512)
    rc = fumeng_amd.lib().fm_desc_load(path.encode(), C.byref(h), C.byref(dp), err, 512)
    return rc, h, dp, err.value.decode()


def test_json_round_trip_reproduces_descriptor(tmp_path):
    s = mesh_scene(freq=5)
    s.images.append(np.clip(np.random.default_rng(1).random((8, 16, 3)) * 3, 0, 3).astype(np.float32))
    s.images[-1] = rgbe_decode(rgbe_encode(s.images[-1]))
    s.add_env(len(s.images) - 1, rotation=(0, 0, 30), factor=1.5)
    tex = np.random.default_rng(2).integers(0, 256, (4, 5, 3), dtype=np.uint8)
    ti = s.add_image(tex)
    s.add_lambert({"type": "image", "image": ti, "sampler": "nearest", "wrap_u": "clamp", "wrap_v": "repeat"})
    path = s.write(str(tmp_path))
    rc, h, dp, err = _load_desc(path)
    assert rc == A.FM_OK, err
    d = dp.contents
    try:
        assert d.num_objects == len(s.objects)
        assert d.num_materials == len(s.materials)
        assert d.num_lights == 1 and d.lights[0].type == A.FM_LIGHT_ENV
        assert abs(d.lights[0].factor - 1.5) < 1e-7
        # the mesh re-read from OBJ is bit-identical (9 significant digits round-trip float32)
        m = d.meshes[0]
        pos = np.ctypeslib.as_array(C.cast(m.positions, C.POINTER(C.c_float)), (m.num_positions * 3,))
        assert np.array_equal(pos, s.meshes[0]["positions"].ravel())
        # images are numbered in load order: material textures first, then the env
        assert d.num_images == 2 and d.lights[0].image == 1
        im = d.images[1]
        assert im.format == A.FM_IMAGE_F32_LINEAR
        px = np.ctypeslib.as_array(C.cast(im.data, C.POINTER(C.c_float)), (im.width * im.height * 3,))
        assert np.array_equal(px, s.images[0].ravel())   # the .hdr decodes to the same texels
        im = d.images[0]
        assert im.format == A.FM_IMAGE_U8_SRGB
        px = np.ctypeslib.as_array(C.cast(im.data, C.POINTER(C.c_uint8)), (im.width * im.height * 3,))
        assert np.array_equal(px, tex.ravel())
        last = d.materials[d.num_materials - 1]
        assert last.tex[0].type == A.FM_TEX_IMAGE and last.tex[0].sampler == A.FM_SAMPLER_NEAREST
        assert last.tex[0].wrap_u == A.FM_WRAP_CLAMP and last.tex[0].wrap_v == A.FM_WRAP_REPEAT
        # same BVH topology from the parsed descriptor as from the in-memory one
        nn, ng = C.c_int64(), C.c_int64()
        assert fumeng_amd.lib().fm_host_build_bvh(dp, C.byref(nn), None, None, None, None, C.byref(ng), None) == 0
        assert ng.value == s.num_geometries()
    finally:
        fumeng_amd.lib().fm_desc_free(h)


def test_loader_defaults_and_error_semantics(tmp_path):
    """config.h LoadValue: missing keys keep defaults, wrong types log and keep
    the default, vec3 broadcasts a single number; missing camera / renderer /
    scene blocks are errors (scene_loader.h:62,83,144)."""
    doc = """{
      "camera": {"type": "pin_hole", "pos": [0, -3, 0], "fov": [45], "up": [2]},
      "renderer": {"width": 8, "spp": "many"},
      "scene": {
        "materials": [{"type": "disney", "name": "d", "basecolor": {"type": "constant", "value": [0.5]}},
                      {"type": "mystery", "name": "m"}],
        "render_objects": [{"shape": {"type": "sphere", "radius": 0.5}, "material": "d", "emittance": 2},
                           {"shape": {"type": "rectangle", "pa": [0,0,0], "pb": [1,0,0], "pc": [1,1,0], "pd": [0,1,0]},
                            "material": "nope"},
                           {"shape": {"type": "cone"}, "material": "m"}],
        "lights": [{"type": "directional", "radiance": [1]}]
      }
    }"""
    p = tmp_path / "s.json"
    p.write_text(doc)
    rc, h, dp, err = _load_desc(str(p))
    assert rc == A.FM_OK, err
    d = dp.contents
    try:
        assert d.camera.fov == 60.0                 # array given for a real: default kept
        assert list(d.camera.up) == [2.0, 2.0, 2.0]  # 1 number broadcasts
        assert d.renderer.width == 8 and d.renderer.height == 1 and d.renderer.spp == 10
        assert d.renderer.depth == 10 and d.renderer.rr_depth == 10 and abs(d.renderer.rr_coef - 0.85) < 1e-7
        assert d.num_objects == 2                   # the unknown shape adds no geometry
        assert list(d.objects[0].emittance) == [0.0, 0.0, 0.0]   # bare number: not an array, default kept
        assert d.objects[1].material == -1          # unknown name -> default_material
        m = d.materials[0]
        assert m.type == A.FM_MAT_DISNEY and list(m.tex[0].value) == [0.5, 0.5, 0.5]
        assert abs(m.tex[A.DISNEY_PARAMS.index("specular")].value[0] - 0.5) < 1e-7
        assert abs(m.tex[A.DISNEY_PARAMS.index("ior")].value[0] - 1.5) < 1e-7
        assert d.lights[0].type == A.FM_LIGHT_DIRECTIONAL and list(d.lights[0].direction) == [1.0, 0.0, 0.0]
    finally:
        fumeng_amd.lib().fm_desc_free(h)
    for broken in ['{"renderer": {}, "scene": {}}', '{"camera": {}, "scene": {}}', '{"camera": {}, "renderer": {}}',
                   '{"camera": {"type": "fisheye"}, "renderer": {}, "scene": {}}', '{"camera": ']:
        p.write_text(broken)
        rc, h, dp, err = _load_desc(str(p))
        assert rc in (A.FM_ERR_PARSE,), (broken, rc, err)
    rc, h, dp, err = _load_desc(str(tmp_path / "missing.json"))
    assert rc == A.FM_ERR_IO


def test_rgbe_roundtrip_matches_radiance_decoding():
    img = np.array([[[1.0, 0.5, 0.25], [5e4, 3.0, 0.0]], [[0.0, 0.0, 0.0], [1e-3, 2e-3, 4e-3]]], np.float32)
    dec = rgbe_decode(rgbe_encode(img))
    assert np.all(dec[1, 0] == 0)
    # one shared exponent: every channel is within 1/128 of the pixel's largest channel
    err = np.abs(dec - img)
    assert np.all(err <= img.max(axis=2, keepdims=True) / 128.0)
    assert np.all(dec <= img)   # mantissas truncate


def test_geodesic_sphere_counts_and_closure():
    v, t = scenes.geodesic_sphere(7)
    assert t.shape[0] == 20 * 49 and v.shape[0] == 10 * 49 + 2
    # closed 2-manifold: every edge shared by exactly two triangles
    e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    assert np.all(counts == 2)


def test_loader_reads_the_post_process_list(tmp_path, capfd):
    s = mesh_scene(freq=4)
    s.post_process = [{"type": "tonemapping", "lum": 1.75}, {"type": "linear_to_srgb"}, {"type": "oidn_denoise"},
                      {"type": "srgb_to_linear"}]
    path = s.write(str(tmp_path))
    import json
    doc = json.load(open(path))
    doc["post_process"].insert(1, {"type": "sharpen"})        # unknown: skipped with a warning
    json.dump(doc, open(path, "w"))
    rc, h, dp, err = _load_desc(path)
    assert rc == 0, err
    d = dp.contents
    got = [(d.post_ops[i].type, d.post_ops[i].lum) for i in range(d.num_post_ops)]
    assert got == [(A.FM_POST_TONEMAPPING, np.float32(1.75)), (A.FM_POST_LINEAR_TO_SRGB, 1.0),
                   (A.FM_POST_OIDN_DENOISE, 1.0), (A.FM_POST_SRGB_TO_LINEAR, 1.0)]
    fumeng_amd.lib().fm_desc_free(h)
    assert "sharpen" in capfd.readouterr().err
    info = fumeng_amd.host_scene_info(s)       # the builder accepts the list
    assert info.num_geometries > 0


def test_threaded_builds_equal_sequential_builds(monkeypatch):
    """The builders split the top levels over threads: the reference
    topology (node array and leaf order) and the traversal tree with its
    slot records are the same as a one-thread build, on a scene large enough
    for the threaded levels (81,920 triangles)."""
    s = mesh_scene(freq=64, with_box=False, with_spheres=False)
    par = [fumeng_amd.host_scene_info(s, w) for w in (2, 4, 8)]
    ref_par = fumeng_amd.host_bvh(s)
    monkeypatch.setenv("FUMENG_BUILD_THREADS", "1")
    seq = [fumeng_amd.host_scene_info(s, w) for w in (2, 4, 8)]
    ref_seq = fumeng_amd.host_bvh(s)
    assert seq[0].build_threads == 1 and par[0].build_threads >= 1
    for a, b in zip(par, seq):
        assert a.trav_hash == b.trav_hash and a.num_trav_nodes == b.num_trav_nodes and a.trav_stack == b.trav_stack
    for a, b in zip(ref_par, ref_seq):
        assert np.array_equal(a, b)
    assert np.array_equal(ref_seq[4], oracle.OracleScene(s).bvh()[4])   # leaf order = the oracle's
