"""Benchmark: Mpaths/s of the FP32 path tracer (wavefront form by default,
the megakernel timed beside it) on the displaced
icosphere scene (1,003,520 triangles + ground, 64 Disney materials,
2048x1024 HDR env, thin lens, 1920x1080, 256 spp, max depth 12).

A step is one full render of that frame with the scene already resident in
HBM.  Prints one JSON line (rank 0, N = 1) with the throughput, the rocprof-
comparable roofline of the render kernel and the CPU baseline (the oracle's
FP32 restatement of the reference path tracer on a bounded pixel sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def host_threads():
    """The host cores this process may use: its CPU affinity, or the thread
    count the host grants it (OMP_NUM_THREADS) if that is lower."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(cap))) if cap.isdigit() and int(cap) > 0 else max(1, n)


def cpu_baseline(desc, W, H, spp, seconds_target=15.0):
    """Oracle FP32 path tracer on a centred crop, on every host core the
    process may use (host_threads)."""
    import oracle
    threads = host_threads()
    o = oracle.OracleScene(desc)
    # calibrate on a tiny crop, then size the sample to about seconds_target
    cw, ch, cs = 16, 16, 4
    t0 = time.perf_counter()
    o.render(W // 2 - cw // 2, H // 2 - ch // 2, cw, ch, spp=cs, seed=11, threads=threads)
    dt = time.perf_counter() - t0
    rate = cw * ch * cs / max(dt, 1e-6)
    n = int(rate * seconds_target)
    side = int(max(16, min(256, (n / 16) ** 0.5)))
    t0 = time.perf_counter()
    o.render(W // 2 - side // 2, H // 2 - side // 2, side, side, spp=16, seed=12, threads=threads)
    dt = time.perf_counter() - t0
    paths = side * side * 16
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"value": paths / dt / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": "%dx%d centred crop of the same frame at 16 spp (%d paths, %.1f s; %d threads, %s cores in "
                      "the process's affinity)" % (side, side, paths, dt, threads, visible)}


# Algorithmic FP32 operations (add, sub, mul, div, sqrt, min, max; a fused
# multiply-add counts 2; compares and selects 0) of the reference formulas
# the traversal and shading run, counted from the restated code:
FLOPS_RAY_SETUP = 3        # 1 / dir (bvh_aggregate.cpp:101-123)
FLOPS_BOX = 24             # slab test of one child box: 6 sub, 6 mul, 6 min/max, 6 for the t_min / t_max clamps (bbox.h:59-69)
FLOPS_TRIANGLE = 48        # watertight test: 9 translate, 3 div shear, 12 shear, 9 edges, 2 det, 3 scale, 1 div, 6 t, 3 bc
                           # (triangle.cpp:25-94)
FLOPS_RECTANGLE = 2 * FLOPS_TRIANGLE   # two triangle tests when the first rejects (rectangle.cpp:42-85)
FLOPS_SPHERE = 70          # world-to-object point (27) and vector (15), quadratic setup (18), root (10) (sphere.cpp:91-129)
FLOPS_VERTEX = 280         # one path vertex of the box scene: hit fill 20, BSDF frame 41, rectangle light sample 45,
                           # Lambert eval + pdf 30, two MIS weights 40, two BSDF samples 80, RR + throughput 24


def box_leg(fumeng_amd, ctx, steps=3):
    """The open-box workload: 5 Lambert walls, a 0.5x0.5 area light, a
    Lambert and a glass sphere; pin-hole 40 degrees, 512x512, 64 spp,
    depth 8, Russian roulette at every vertex.  Both forms, plus the VALU
    roofline of the megakernel (the plain form): algorithmic FLOPs per ray
    from the op-count table above and the counting variant's node visits
    and primitive tests by kind, over the kernel's HIP-event duration.
    Vertices are taken as rays / 3 (one shadow, one MisBSDF and one
    continuation ray per vertex with one light)."""
    from fumeng_amd import scenes, _abi
    desc = scenes.box_scene()
    W, H, spp = desc.renderer["width"], desc.renderer["height"], desc.renderer["spp"]
    sc = ctx.scene(desc)
    res = {"workload": "open box, 5 Lambert walls + 0.5x0.5 light + Lambert and glass spheres, pin-hole 40 deg, "
                       "512x512, 64 spp, depth 8, rr_depth 3"}
    for name, mode in (("megakernel", _abi.FM_MODE_MEGAKERNEL), ("wavefront", _abi.FM_MODE_WAVEFRONT)):
        sc.render(seed=3, spp=4, read=False, mode=mode)   # warm-up
        ms, rays = [], 0
        t0 = time.perf_counter()
        for i in range(steps):
            _, _, _, _, st = sc.render(seed=1 + i, read=False, mode=mode)
            ms.append(st.kernel_ms)
            rays += st.rays
        wall = time.perf_counter() - t0
        res[name] = {"mpaths_per_s": W * H * spp * steps / wall / 1e6, "mrays_per_s": rays / wall / 1e6,
                     "ms_per_frame": wall * 1e3 / steps, "kernel_ms": float(sum(ms) / len(ms)), "rays": rays // steps}
        log("box %s: %.1f Mpaths/s, %.1f ms per frame" % (name, res[name]["mpaths_per_s"], res[name]["ms_per_frame"]))
    fumeng_amd.lib().fm_set_counting(ctx.h, 1)
    _, _, _, _, cst = sc.render(seed=2, spp=8, read=False, mode=_abi.FM_MODE_MEGAKERNEL)
    fumeng_amd.lib().fm_set_counting(ctx.h, 0)
    sc.close()
    r = float(cst.rays)
    tris = cst.prim_tests - cst.sphere_tests - cst.rect_tests
    per_ray = (FLOPS_RAY_SETUP + 2 * FLOPS_BOX * cst.node_visits / r + FLOPS_TRIANGLE * tris / r
               + FLOPS_RECTANGLE * cst.rect_tests / r + FLOPS_SPHERE * cst.sphere_tests / r + FLOPS_VERTEX / 3.0)
    mk = res["megakernel"]
    achieved = per_ray * mk["rays"] / (mk["kernel_ms"] * 1e-3) / 1e12
    res["roofline_valu"] = {"bound": "valu", "kernel": "k_render_mega<false>", "achieved": achieved,
                            "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                            "flops_per_ray": per_ray, "nodes_per_ray": cst.node_visits / r,
                            "rect_tests_per_ray": cst.rect_tests / r, "sphere_tests_per_ray": cst.sphere_tests / r,
                            "triangle_tests_per_ray": tris / r}
    log("box megakernel VALU roofline: %.0f FLOP per ray, %.2f TFLOP/s = %.4f of %.1f"
        % (per_ray, achieved, achieved / FP32_PEAK_TFLOPS, FP32_PEAK_TFLOPS))
    return res


def batch_leg(fumeng_amd, ctx):
    """64 uneven renders of the box scene (cameras on a radius-4 sphere
    looking at the origin, r x r with r ~ U{128..1024}, spp ~ U{16..256},
    depth ~ U{4..16}, about 4.1e9 paths) in one wavefront run with the
    captured pass graph; aggregate Mpaths/s and per-job completion times."""
    import numpy as np
    from fumeng_amd import scenes
    jobs = scenes.batch_jobs(64, seed=4)
    box = ctx.scene(scenes.box_scene())
    # warm-up: code objects, and the path-state buffers grown to this
    # batch's pass size (a one-time allocation of up to the context's budget)
    box.render_batch(jobs, read=False, graph=True)
    t0 = time.perf_counter()
    _, done, st = box.render_batch(jobs, read=False, graph=True)
    wall = time.perf_counter() - t0
    box.close()
    res = {"workload": "64 uneven jobs of the box scene", "jobs": len(jobs), "paths": st.paths, "rays": st.rays,
           "seconds": wall, "mpaths_per_s": st.paths / wall / 1e6, "mrays_per_s": st.rays / wall / 1e6,
           "kernel_ms": st.kernel_ms, "passes": st.passes, "units": st.units,
           "job_done_ms_p50": float(np.percentile(done, 50)), "job_done_ms_p99": float(np.percentile(done, 99))}
    log("batch: %d jobs, %.3g paths in %.2f s = %.1f Mpaths/s, %d passes, p50 %.0f ms, p99 %.0f ms"
        % (len(jobs), st.paths, wall, res["mpaths_per_s"], st.passes, res["job_done_ms_p50"], res["job_done_ms_p99"]))
    return res


def adjoint_leg(fumeng_amd, ctx, iters=6):
    """Inverse rendering: the displaced icosphere at frequency 71 (100,822
    geometries), 64 Disney materials, 512x256 env, 256x256, depth 6.  Adam
    on all 384 material parameters and 393,216 env texel values from a
    perturbed start towards a target rendered from the hidden values; one
    iteration = set parameters (env tables rebuilt in FP64 on the GPU) +
    32-spp forward render + 32-spp path-replay adjoint + optimiser step, with
    the parameters, image and gradients as tensors in HBM."""
    import numpy as np
    import torch
    from fumeng_amd import scenes, _abi
    from fumeng_amd.torch_ext import render
    sc = ctx.scene(scenes.inverse_scene())
    mp, env = sc.params()
    dev = torch.device("cuda", ctx.device)
    target = torch.from_numpy(sc.render(spp=256, seed=5, mode=_abi.FM_MODE_WAVEFRONT)[0]).to(dev)
    start_mp, start_env = scenes.perturb_params(mp, env, seed=5)
    mask = torch.from_numpy(np.isfinite(mp).astype(np.float32)).to(dev)
    theta = torch.tensor(np.nan_to_num(start_mp, nan=0.0), device=dev, requires_grad=True)
    env_t = torch.tensor(start_env, device=dev, requires_grad=True)
    opt = torch.optim.Adam([{"params": [theta], "lr": 0.03}, {"params": [env_t], "lr": 0.03 * float(env.mean())}])
    nan_rows = torch.full_like(theta, float("nan"))
    times, losses, fwd_ms, adj_ms = [], [], [], []
    orig_render, orig_adjoint = sc.render, sc.adjoint_device

    def timed_render(**kw):
        out = orig_render(**kw)
        fwd_ms.append(out[4].kernel_ms)
        return out

    def timed_adjoint(*a, **kw):
        out = orig_adjoint(*a, **kw)
        adj_ms.append(out.kernel_ms)
        return out
    sc.render, sc.adjoint_device = timed_render, timed_adjoint
    for it in range(iters + 1):
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        opt.zero_grad()
        params = torch.where(mask > 0, theta, nan_rows)
        img = render(sc, params, env_t, spp=32, seed=100 + it, spp_adjoint=32, seed_adjoint=900 + it)
        loss = torch.mean((img - target) ** 2)
        loss.backward()
        theta.grad = torch.nan_to_num(theta.grad) * mask
        opt.step()
        with torch.no_grad():
            theta.clamp_(0.0, 1.0)
            env_t.clamp_(min=0.0)
        torch.cuda.synchronize(dev)
        if it > 0:                       # the first iteration is the warm-up
            times.append(time.perf_counter() - t0)
        losses.append(float(loss.detach()))
    sc.close()
    res = {"workload": "inverse rendering, 100,822 geometries, 64 Disney materials, 512x256 env, 256x256, "
                       "32 + 32 spp, depth 6", "params": int(mask.sum().item()) + env.size,
           "ms_per_iter": 1e3 * float(np.mean(times)), "forward_kernel_ms": float(np.mean(fwd_ms[1:])),
           "adjoint_kernel_ms": float(np.mean(adj_ms[1:])), "iters": iters,
           "loss_first": losses[0], "loss_last": losses[-1]}
    log("adjoint: %.1f ms/iter (forward kernel %.1f ms, adjoint kernel %.1f ms), loss %.4g -> %.4g"
        % (res["ms_per_iter"], res["forward_kernel_ms"], res["adjoint_kernel_ms"], losses[0], losses[-1]))
    return res


def adjoint_child():
    """The adjoint leg in a child process that imports torch before the
    library: its tensors live in HBM, which needs torch's HIP runtime to be
    the process's only one (this process loaded the library's first)."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--adjoint-only"], stdout=subprocess.PIPE, text=True,
                       timeout=900)
    if r.returncode != 0:
        raise RuntimeError("adjoint child exited with %d" % r.returncode)
    return json.loads(r.stdout.strip().splitlines()[-1])["adjoint"]


def stress_leg(fumeng_amd, ctx):
    """Large-geometry stress scene: 400 displaced spheres (9.8 M triangles),
    200 sRGB 2048^2 textures, 10 rectangle lights, ground; 1920x1080, 128
    spp, depth 10; one wavefront frame (scene generation and upload are
    reported, not timed into the throughput).  The scene is built twice,
    with the device's traversal-tree builder and with the host's, and a
    frame is rendered from each."""
    from fumeng_amd import scenes, _abi
    t0 = time.perf_counter()
    desc = scenes.stress_scene()
    gen = time.perf_counter() - t0
    W, H, spp = desc.renderer["width"], desc.renderer["height"], desc.renderer["spp"]

    def build_and_frame(builder):
        ctx.set_bvh_builder(builder)
        try:
            t0 = time.perf_counter()
            sc = ctx.scene(desc)
            build = time.perf_counter() - t0
        finally:
            ctx.set_bvh_builder("host")
        info = sc.info()
        sc.render(seed=2, spp=1, read=False, mode=_abi.FM_MODE_WAVEFRONT)        # warm-up
        t0 = time.perf_counter()
        _, _, _, _, st = sc.render(seed=1, read=False, mode=_abi.FM_MODE_WAVEFRONT, time_trace=True)
        dt = time.perf_counter() - t0
        sc.close()
        return info, build, st, dt

    # the device-built tree of the same scene: build time against frame time
    dinfo, dbuild, dst, ddt = build_and_frame("device")
    device = {"build_upload_s": dbuild, "device_build_ms": dinfo.build_ms_device, "iterations": dinfo.build_iterations,
              "host_total_ms": dinfo.build_ms_total, "trav_nodes": dinfo.num_trav_nodes, "trav_stack": dinfo.trav_stack,
              "ms_per_frame": ddt * 1e3, "trace_ms": dst.trace_ms, "mpaths_per_s": W * H * spp / ddt / 1e6}
    log("stress, device-built tree: built in %.1f s (host %.0f ms, device %.0f ms, %d iterations), %d nodes, "
        "%d stack entries; %.0f ms per frame (traversal %.0f ms)"
        % (dbuild, dinfo.build_ms_total, dinfo.build_ms_device, dinfo.build_iterations, dinfo.num_trav_nodes,
           dinfo.trav_stack, ddt * 1e3, dst.trace_ms))
    info, build, st, dt = build_and_frame("host")
    files = None
    try:
        files = stress_from_files(desc, info)
    except Exception as e:
        log("stress from files failed: %s" % e)
    del desc
    res = {"workload": "stress: 400 displaced spheres (9,800,012 geometries), 200 textures 2048^2, 10 rect lights, "
                       "1920x1080, 128 spp, depth 10", "mpaths_per_s": W * H * spp / dt / 1e6,
           "mrays_per_s": st.rays / dt / 1e6, "ms_per_frame": dt * 1e3, "trace_ms": st.trace_ms,
           "generate_s": gen, "build_upload_s": build, "device_gb": info.device_bytes / 1e9,
           "build_report": {"threads": info.build_threads, "geometry_ms": info.build_ms_geometry,
                            "reference_bvh_ms": info.build_ms_ref_bvh, "traversal_tree_ms": info.build_ms_trav_tree,
                            "tables_ms": info.build_ms_tables, "host_total_ms": info.build_ms_total,
                            "upload_ms": info.upload_ms, "host_gb": info.host_bytes / 1e9},
           "trav_width": info.trav_width, "trav_nodes": info.num_trav_nodes, "trav_stack": info.trav_stack,
           "device_builder": device,
           "from_files": files}
    log("stress: %.1f Mpaths/s, %.1f Mrays/s (%.0f ms per frame, traversal %.0f ms), %d-wide tree, built in %.1f s, "
        "%.2f GB in HBM" % (res["mpaths_per_s"], res["mrays_per_s"], dt * 1e3, st.trace_ms, info.trav_width, build,
                           res["device_gb"]))
    log("stress build on %d threads: geometry %.0f ms, reference BVH %.0f ms, traversal tree %.0f ms, tables %.0f ms, "
        "upload %.0f ms; host copy %.2f GB" % (info.build_threads, info.build_ms_geometry, info.build_ms_ref_bvh,
                                               info.build_ms_trav_tree, info.build_ms_tables, info.upload_ms,
                                               info.host_bytes / 1e9))
    return res


LOAD_CHILD = r"""
import json, resource, sys, time
sys.path.insert(0, sys.argv[1])
import fumeng_amd
t0 = time.perf_counter()
ctx = fumeng_amd.Context(0)
t1 = time.perf_counter()
sc = ctx.load(sys.argv[2])
t2 = time.perf_counter()
i = sc.info()
print(json.dumps({"load_build_upload_s": t2 - t1, "context_s": t1 - t0, "geometries": i.num_geometries,
                  "trav_hash": str(i.trav_hash), "device_bytes": i.device_bytes,
                  "build_ms_total": i.build_ms_total, "upload_ms": i.upload_ms,
                  "peak_host_gb": resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1e6}))
sc.close()
ctx.close()
"""


def stress_from_files(desc, info):
    """The stress scene end to end from files: SceneDesc.write (scene JSON,
    400 OBJ meshes, 200 TGA textures) into a temporary directory, then a
    child process loads it with fm_scene_load (the project's JSON / OBJ /
    TGA readers with the reference loader's semantics) and builds and
    uploads it; its wall time and peak host memory are reported, and its
    traversal hash (tree and slot records) must equal the in-memory build's."""
    import shutil
    import subprocess
    import tempfile
    d = tempfile.mkdtemp(prefix="fm_stress_")
    try:
        t0 = time.perf_counter()
        path = desc.write(d)
        write_s = time.perf_counter() - t0
        size = sum(os.path.getsize(os.path.join(d, f)) for f in os.listdir(d))
        r = subprocess.run([sys.executable, "-c", LOAD_CHILD, ROOT, path], stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(r.stderr[-2000:])
        res = json.loads(r.stdout.strip().splitlines()[-1])
    finally:
        shutil.rmtree(d, ignore_errors=True)
    res.update({"write_s": write_s, "files_gb": size / 1e9, "files": len(os.listdir(d)) if os.path.exists(d) else None,
                "same_scene": res["trav_hash"] == str(info.trav_hash) and res["geometries"] == info.num_geometries})
    res.pop("files")
    log("stress from files: %.2f GB written in %.1f s; loaded, built and uploaded in %.1f s by a child process "
        "(peak host memory %.1f GB); same traversal tree and slots as the in-memory build: %s"
        % (res["files_gb"], write_s, res["load_build_upload_s"], res["peak_host_gb"], res["same_scene"]))
    return res


def profiled(tag, kernel):
    """A figure of `kernel` on the default workload from the committed
    rocprofv3 summaries under profiles/: lines "<tag>: <kernel> <value>".
    bench-traffic = HBM bytes per launch (FETCH_SIZE + WRITE_SIZE scaled by a
    1 GiB copy), bench-l2-hit = L2 hit rate (TCC_HIT / (TCC_HIT + TCC_MISS)),
    bench-valu-busy = share of cycles with VALU instructions issued,
    bench-launch-us = average launch duration in the counter run's trace.
    Read from the default workload's summary only (other workloads run the
    same kernels at other sizes).  None when it does not name the kernel."""
    path = os.path.join(ROOT, "profiles", "wavefront_icosphere_1080p_256spp.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(tag + ":"):
            name, _, value = line[len(tag) + 1:].strip().rpartition(" ")
            if name == kernel:
                try:
                    return float(value)
                except ValueError:
                    return None
    return None


def copy_probe(nbytes):
    """A plain elementwise copy (y = -x over nbytes of float32): its
    FETCH_SIZE / WRITE_SIZE counters should read nbytes each, which gives the
    scale factor of those counters on this device."""
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=256)
    ap.add_argument("--freq", type=int, default=224)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--mode", choices=["megakernel", "wavefront"], default="wavefront")
    ap.add_argument("--bvh-width", type=int, choices=[0, 2, 4, 8], default=0,
                    help="traversal tree width; 0 (default): two-wide while the tree and primitive records fit in "
                         "the 256 MB last-level cache, else four-wide")
    ap.add_argument("--bvh-builder", choices=["host", "device"], default="host",
                    help="traversal tree builder: binned SAH on the host threads or PLOC on the GPU")
    ap.add_argument("--no-stress", dest="stress", action="store_false",
                    help="skip the 9.8 M-triangle textured stress scene leg")
    ap.add_argument("--no-adjoint", dest="adjoint", action="store_false",
                    help="skip the inverse-rendering (adjoint) leg")
    ap.add_argument("--no-batch", dest="batch", action="store_false",
                    help="skip the batched-jobs leg (64 uneven renders of the box scene)")
    ap.add_argument("--no-box", dest="box", action="store_false",
                    help="skip the open-box leg (512x512, 64 spp, both forms, VALU roofline)")
    ap.add_argument("--no-compare", dest="compare", action="store_false",
                    help="skip the one-step run of the other form")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the inverse-rendering leg and print its result (the main run starts it as a child)")
    ap.add_argument("--box-only", action="store_true",
                    help="only run the open-box leg and print its JSON (profiling of that workload)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (a frame shards by "
              "sample range; see the design notes)", file=sys.stderr)
        sys.exit(2)

    if args.adjoint_only:
        import torch
        torch.ones(1, device="cuda")   # torch's HIP runtime first: the library then shares it
        import fumeng_amd
        ctx = fumeng_amd.Context(0)
        print(json.dumps({"adjoint": adjoint_leg(fumeng_amd, ctx)}))
        ctx.close()
        return
    import fumeng_amd
    from fumeng_amd import scenes
    from fumeng_amd import _abi as _abi_mod
    if args.box_only:
        ctx = fumeng_amd.Context(0)
        print(json.dumps(box_leg(fumeng_amd, ctx)))
        ctx.close()
        return

    t0 = time.perf_counter()
    desc = scenes.icosphere_scene(freq=args.freq, spp=args.spp)
    W, H, spp = desc.renderer["width"], desc.renderer["height"], desc.renderer["spp"]
    log("scene generated: %d geometries in %.1f s" % (desc.num_geometries(), time.perf_counter() - t0))
    ctx = fumeng_amd.Context(0)
    ctx.set_bvh_width(args.bvh_width)
    ctx.set_bvh_builder(args.bvh_builder)
    t0 = time.perf_counter()
    scene = ctx.scene(desc)
    info = scene.info()
    log("scene built and uploaded in %.1f s: %d BVH nodes, depth %d; traversal tree %d wide (%s builder%s), %d nodes, "
        "%d stack entries; %.1f MB in HBM"
        % (time.perf_counter() - t0, info.num_bvh_nodes, info.bvh_depth, info.trav_width, args.bvh_builder,
           ", %.1f ms on the device, %d iterations" % (info.build_ms_device, info.build_iterations)
           if info.trav_builder == _abi_mod.FM_BUILDER_DEVICE else "", info.num_trav_nodes,
           info.trav_stack, info.device_bytes / 1e6))

    from fumeng_amd import _abi
    mode = _abi.FM_MODE_WAVEFRONT if args.mode == "wavefront" else _abi.FM_MODE_MEGAKERNEL
    kw = dict(seed=1, block_size=args.block, grid_blocks=args.grid, chunk=args.chunk, read=False, mode=mode,
              time_trace=True)
    for i in range(args.warmup):
        _, _, _, _, st = scene.render(**kw)
        log("warmup %d: %.1f ms" % (i, st.kernel_ms))
    kernel_ms, trace_ms, rays, paths = 0.0, 0.0, 0, 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        _, _, _, _, st = scene.render(**kw)
        kernel_ms += st.kernel_ms
        trace_ms += st.trace_ms
        rays += st.rays
        paths += st.paths
        log("step %d: %.1f ms (traversal launches %.1f ms), %.1f Mrays, %d pass(es) of %d iterations"
            % (i, st.kernel_ms, st.trace_ms, st.rays / 1e6, st.passes, st.iterations))
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    mpaths = W * H * spp * args.steps / elapsed / 1e6
    mrays = rays / elapsed / 1e6

    # The other form on the same frame (one step): the wavefront runs the
    # megakernel's estimator with the same random numbers, so the two films
    # have equal error and their throughputs compare directly.
    other = None
    if args.compare:
        omode = _abi.FM_MODE_MEGAKERNEL if mode == _abi.FM_MODE_WAVEFRONT else _abi.FM_MODE_WAVEFRONT
        t1 = time.perf_counter()
        _, _, _, _, ost = scene.render(seed=1, read=False, mode=omode)
        odt = time.perf_counter() - t1
        other = {"mode": "megakernel" if omode == _abi.FM_MODE_MEGAKERNEL else "wavefront",
                 "mpaths_per_s": W * H * spp / odt / 1e6, "mrays_per_s": ost.rays / odt / 1e6,
                 "ms_per_step": odt * 1e3, "rays": ost.rays}
        log("%s: %.1f ms, %.1f Mpaths/s" % (other["mode"], odt * 1e3, other["mpaths_per_s"]))

    # Roofline of the traversal kernel (the wavefront's trace launches, or the
    # whole megakernel): algorithmic bytes per ray = the node and primitive
    # records the traversal must fetch (64 B per visited two-wide node or
    # 128 B per four-wide node, 48 B per tested primitive), counted by the
    # counting variant of the same kernels
    # on a 1-spp pass, times the rays of the timed steps, over the summed
    # duration of those launches (HIP events around each launch).
    roofline = None
    try:
        fumeng_amd.lib().fm_set_counting(ctx.h, 1)
        _, _, _, _, cst = scene.render(seed=2, spp=1, read=False, mode=mode)
        fumeng_amd.lib().fm_set_counting(ctx.h, 0)
        if cst.node_visits > 0 and cst.rays > 0 and trace_ms > 0:
            bytes_per_ray = (cst.node_visits * float(info.trav_node_bytes) + cst.prim_tests * 48.0) / cst.rays
            achieved = bytes_per_ray * rays / (trace_ms * 1e-3) / 1e9
            default_run = (args.freq, args.spp, args.block, args.grid, args.chunk) == (224, 256, 0, 0, 0)
            kname = (("k_wf_trace<false, %d, %s>" % (info.trav_width, "true" if st.trace_short_stack else "false"))
                     if mode == _abi.FM_MODE_WAVEFRONT else "k_render_mega<false>")
            traffic = profiled("bench-traffic", kname) if default_run else None
            traffic = int(traffic) if traffic is not None else None
            launch_ms = trace_ms / args.steps / max(1, st.passes * st.iterations)   # trace launches per frame
            # `achieved` / `frac` rate the bytes the traversal requests; the
            # measured HBM rate (counter traffic per launch over the launch
            # time) and the L2 hit rate say how much of that reaches HBM
            measured = None
            if traffic is not None:
                hbm_gbs = traffic / (launch_ms * 1e-3) / 1e9
                measured = {"hbm_gbs": hbm_gbs, "hbm_frac": hbm_gbs / HBM_PEAK_GBS,
                            "traffic_over_requested": traffic / (bytes_per_ray * rays / args.steps /
                                                                 max(1, st.passes * st.iterations)),
                            "l2_hit": profiled("bench-l2-hit", kname),
                            "valu_busy": profiled("bench-valu-busy", kname)}
            roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kname,
                        "achieved_is": "requested (algorithmic) bytes per second",
                        "measured": measured,
                        "bytes_per_ray": bytes_per_ray, "kernel_ms_per_step": trace_ms / args.steps,
                        "nodes_per_ray": cst.node_visits / cst.rays, "prims_per_ray": cst.prim_tests / cst.rays,
                        "node_bytes": info.trav_node_bytes}
    except Exception as e:  # counting is diagnostic; the throughput stands without it
        log("counting pass failed: %s" % e)

    box = None
    if args.box:
        try:
            box = box_leg(fumeng_amd, ctx)
        except Exception as e:
            log("box leg failed: %s" % e)

    batch = None
    if args.batch:
        try:
            batch = batch_leg(fumeng_amd, ctx)
        except Exception as e:
            log("batch leg failed: %s" % e)

    stress = None
    if args.stress:
        try:
            stress = stress_leg(fumeng_amd, ctx)
        except Exception as e:
            log("stress leg failed: %s" % e)

    adjoint = None
    if args.adjoint:
        try:
            adjoint = adjoint_child()
        except Exception as e:
            log("adjoint leg failed: %s" % e)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(desc, W, H, spp)
        except Exception as e:
            log("cpu baseline failed: %s" % e)

    out = {"metric": "Mpaths/s", "value": mpaths, "unit": "Mpaths/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scene)",
           "config": {"workload": "displaced icosphere %d tris + ground, 64 Disney materials, 2048x1024 HDR env, "
                                  "thin lens, %dx%d, %d spp, depth 12, %s" % (20 * args.freq ** 2, W, H, spp, args.mode),
                      "mode": args.mode, "width": W, "height": H, "spp": spp, "depth": 12,
                      "bvh_width": info.trav_width, "bvh_builder": args.bvh_builder},
           "mrays_per_s": mrays, "paths": paths, "rays": rays, "other_form": other, "box": box, "batch": batch, "adjoint": adjoint,
           "stress": stress, "roofline": roofline,
           "roofline_valu": box["roofline_valu"] if box else None, "cpu_baseline": cpu}
    print(json.dumps(out))
    scene.close()
    ctx.close()


if __name__ == "__main__":
    main()
