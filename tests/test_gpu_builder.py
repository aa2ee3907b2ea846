"""GPU: the device traversal-tree builder (FM_BUILDER_DEVICE: PLOC in Morton
order with a SAH leaf collapse) against the host's binned-SAH tree and the
oracle.

Bars:
  structure    the downloaded tree: the slots are a permutation of the
               geometries; every node is reached once from the root, in
               depth-first pre-order; the leaves hold 1..7 primitives and
               tile the slots in order; every child box contains the bounds
               of the primitives below it; the stack bound covers the depth
  determinism  two builds of a scene give identical nodes and slot orders
  hits         closest-hit geometry and t, and any-hit results, equal the
               host tree's bit for bit (both trees resolve ties by reference
               slot), for random and axis-aligned rays, except on rays with
               0 < |d.z| < 1e-2 (the reference's truncated shear axis gives
               false FP32 hits there, and which one is reported depends on
               the boxes visited: logged, at most 1 in 1000 rays); closest
               hits equal the oracle's brute force bit for bit away from
               those rays
  images       wavefront and megakernel films equal the host tree's on at
               least 98 % of the pixels (a path of ~20 rays that meets one
               of those rays may take another vertex)
  edge cases   0, 1, 2, 7, 8 geometries; hundreds of coincident triangles
"""
import numpy as np
import pytest

import oracle
from fumeng_amd import _abi as A
from fumeng_amd import FumengError, scenes
from fumeng_amd.scene import SceneDesc

from helpers import mesh_scene, scene_rays

pytestmark = pytest.mark.gpu


def _scene_with(gpu_ctx, desc, builder):
    try:
        gpu_ctx.set_bvh_builder(builder)
        return gpu_ctx.scene(desc)
    finally:
        gpu_ctx.set_bvh_builder("host")


def triangle_soup(n, seed, coincident=0):
    """n random triangles (sizes over three decades) in one mesh with an
    identity transform, so geometry k is triangle k; `coincident` more
    copies of triangle 0."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-0.8, 0.8, (n, 3))
    size = 10.0 ** rng.uniform(-3, -1, (n, 1))
    v = (c[:, None, :] + size[:, None, :] * rng.normal(size=(n, 3, 3))).astype(np.float32)
    if coincident:
        v = np.concatenate([v, np.repeat(v[:1], coincident, 0)])
    s = scenes.box_scene(width=32, height=32, spp=4)
    s.objects = []
    m = s.add_mesh(v.reshape(-1, 3), np.arange(3 * len(v), dtype=np.int32).reshape(-1, 3))
    s.add_mesh_object(m, material=0)
    return s, v


def same_hits(a, b, rays, any_hit=False):
    """Equal hits, bit for bit, except on rays with 0 < |d.z| < 1e-2: there
    the reference's truncated shear axis (triangle.cpp:13-18) can give false
    FP32 hits, and which one is reported depends on the boxes a traversal
    visits (logged; at most 1 in 1000 rays)."""
    dz = np.abs(rays[:, 6])
    sensitive = (dz > 0) & (dz < 1e-2)
    mism = a["geom"] != b["geom"]
    if not any_hit:
        mism |= (a["geom"] >= 0) & (a["t"] != b["t"])
    if mism.any():
        print("%d mismatches, all on sensitive rays: %s" % (mism.sum(), bool(np.all(sensitive[mism]))))
    assert np.all(sensitive[mism]) and mism.sum() <= max(2, len(rays) // 1000)


def check_tree(g, tri_v):
    """The structural invariants of a two-wide tree against the triangle
    bounds (geometry k = tri_v[k])."""
    info = g.info()
    nodes, slots = g.traversal()
    G = len(tri_v)
    assert info.trav_width == 2 and len(slots) == G
    assert np.array_equal(np.sort(slots), np.arange(G))
    lo = tri_v.min(axis=1)[slots]
    hi = tri_v.max(axis=1)[slots]
    b = nodes["b"].reshape(-1, 2, 6)
    ch = nodes["child"]
    order, leaves, depth = [], [], [0]

    def walk(ref, level):
        """returns the slot range below ref"""
        if ref < 0:
            v = ~int(ref) & 0xffffffff
            start, cnt = v >> 3, v & 7
            assert 1 <= cnt <= 7
            leaves.append((start, cnt))
            return start, start + cnt
        order.append(ref)
        depth[0] = max(depth[0], level)
        r0 = r1 = None
        for k in range(2):
            s0, s1 = walk(int(ch[ref, k]), level + 1)
            box = b[ref, k]
            assert np.all(box[:3] <= lo[s0:s1].min(axis=0)) and np.all(box[3:] >= hi[s0:s1].max(axis=0)), (ref, k)
            r0 = s0 if r0 is None else r0
            assert r1 is None or r1 == s0
            r1 = s1
        return r0, r1

    root = 0 if info.num_trav_nodes > 0 else None
    if root is None:
        assert G <= 7
        return nodes, slots
    import sys
    sys.setrecursionlimit(10000)
    assert walk(root, 1) == (0, G)
    assert order == list(range(len(nodes)))           # pre-order, every node once
    starts = [s for s, _ in leaves]
    assert starts == sorted(starts) and sum(c for _, c in leaves) == G
    assert depth[0] <= info.trav_stack <= 64
    return nodes, slots


def test_device_tree_structure_and_determinism(gpu_ctx):
    s, v = triangle_soup(20000, 3)
    g1 = _scene_with(gpu_ctx, s, "device")
    i1 = g1.info()
    assert i1.trav_builder == A.FM_BUILDER_DEVICE and i1.build_iterations > 0 and i1.build_ms_device > 0
    n1, s1 = check_tree(g1, v)
    g2 = _scene_with(gpu_ctx, s, "device")
    n2, s2 = g2.traversal()
    assert np.array_equal(n1.view(np.uint8), n2.view(np.uint8)) and np.array_equal(s1, s2)
    gh = _scene_with(gpu_ctx, s, "host")
    ih = gh.info()
    print("device tree: %d nodes, %d stack entries, %d iterations, %.1f ms; host tree %d nodes, %d stack entries"
          % (i1.num_trav_nodes, i1.trav_stack, i1.build_iterations, i1.build_ms_device, ih.num_trav_nodes, ih.trav_stack))
    assert ih.trav_builder == A.FM_BUILDER_HOST
    rng = np.random.default_rng(5)
    r = scene_rays(rng, 20000)
    for any_hit in (False, True):
        same_hits(g1.trace(r, any_hit), gh.trace(r, any_hit), r, any_hit)


@pytest.mark.parametrize("n,coincident", [(1, 0), (2, 0), (7, 0), (8, 0), (50, 300)])
def test_device_tree_small_and_coincident(gpu_ctx, n, coincident):
    s, v = triangle_soup(n, 10 + n, coincident)
    g = _scene_with(gpu_ctx, s, "device")
    gh = _scene_with(gpu_ctx, s, "host")
    check_tree(g, v)
    rng = np.random.default_rng(6)
    o = rng.uniform(-0.9, 0.9, (4000, 3))
    t = v[rng.integers(0, len(v), 4000)].mean(axis=1) + rng.normal(scale=1e-3, size=(4000, 3))
    from helpers import rays_toward
    r = rays_toward(o, t)
    for any_hit in (False, True):
        same_hits(g.trace(r, any_hit), gh.trace(r, any_hit), r, any_hit)


def test_device_tree_hits_and_images_equal_the_host_tree(gpu_ctx):
    """The mixed mesh / rectangle / sphere scene: hits for random and
    axis-aligned rays, and films of both forms, bit for bit; closest hits
    against the oracle's brute force."""
    s = mesh_scene(freq=20)
    gd = _scene_with(gpu_ctx, s, "device")
    gh = _scene_with(gpu_ctx, s, "host")
    rng = np.random.default_rng(41)
    r = scene_rays(rng, 20000)
    n = 3000
    ax = np.zeros((n, 8), np.float32)
    ax[:, 0:3] = rng.uniform(-0.9, 0.9, (n, 3))
    ax[np.arange(n), 4 + rng.integers(0, 3, n)] = rng.choice([-1.0, 1.0], n)
    ax[:, 7] = 3.4e38
    for rays in (r, ax):
        same_hits(gd.trace(rays), gh.trace(rays), rays)
        same_hits(gd.trace(rays, True), gh.trace(rays, True), rays, True)
    ho = oracle.OracleScene(s).trace(r)
    hd = gd.trace(r)
    keep = np.abs(r[:, 6]) >= 1e-2
    assert np.array_equal(hd["geom"][keep], ho["geom"][keep])
    hit = keep & (ho["geom"] >= 0)
    assert np.array_equal(hd["t"][hit], ho["t"][hit])
    for mode in (A.FM_MODE_WAVEFRONT, A.FM_MODE_MEGAKERNEL):
        x = gd.render(width=48, height=32, spp=4, seed=9, mode=mode)
        y = gh.render(width=48, height=32, spp=4, seed=9, mode=mode)
        for p, q in zip(x[:4], y[:4]):
            differ = np.any(p != q, axis=-1) & ~np.all(np.isnan(p) & np.isnan(q), axis=-1)
            print("mode %d: %d of %d pixels differ" % (mode, differ.sum(), differ.size))
            assert differ.sum() <= 0.02 * differ.size


def test_device_builder_arguments_and_empty_scene(gpu_ctx):
    with pytest.raises(FumengError):
        gpu_ctx.set_bvh_builder(7)
    s = SceneDesc()
    s.renderer.update(width=4, height=4, spp=2)
    g = _scene_with(gpu_ctx, s, "device")
    assert g.info().num_geometries == 0 and np.all(g.render()[0] == 0)
    s2, _ = triangle_soup(100, 4)
    try:
        gpu_ctx.set_bvh_builder("device")
        gpu_ctx.set_bvh_width(4)
        with pytest.raises(FumengError):
            gpu_ctx.scene(s2)
    finally:
        gpu_ctx.set_bvh_width(0)
        gpu_ctx.set_bvh_builder("host")
    s2.accelerator = "reference_tree"      # the reference topology is kept as asked
    g = _scene_with(gpu_ctx, s2, "device")
    assert g.info().trav_builder == A.FM_BUILDER_HOST
