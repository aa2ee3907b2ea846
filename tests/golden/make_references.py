"""Writes the converged reference renders the image-convergence tests
compare against (tests/golden/fp64_*.npz and fp32_*.npz).

Each fixture is the oracle's restatement of the reference path tracer at
16384 samples per pixel, in FP64 (the reference's USE_FLOAT=OFF build) and in
FP32 (its USE_FLOAT build, the precision the GPU computes in), keeping
the per-pixel mean and sum of squares, so that a test can form per-pixel
z-scores that include the reference's own Monte-Carlo error.  The scenes
come from fumeng_amd.scenes with the arguments recorded in SCENES; the
fixture stores a signature of the generated scene so a changed generator is
detected instead of silently compared.

Run from the repository root:  python tests/golden/make_references.py
[fp64_box32 fp32_ico48x27 ...]  (no argument: all four)
(about 2 minutes on 8 host cores).  The FP32 renders use another seed than
the FP64 ones, so the two are independent estimates.
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle  # noqa: E402
from helpers import convergence_scene, scene_signature  # noqa: E402

SPP = 16384
SEEDS = {"fp64": 1001, "fp32": 1002}
NAMES = ("box32", "ico48x27")


def main():
    threads = min(16, os.cpu_count() or 1)
    only = sys.argv[1:]
    for prec, seed in SEEDS.items():
        for name in NAMES:
            if only and "%s_%s" % (prec, name) not in only:
                continue
            s = convergence_scene(name)
            W, H = s.renderer["width"], s.renderer["height"]
            t0 = time.time()
            o = oracle.OracleScene(s, fp64=prec == "fp64")
            c, a, n, q = o.render(0, 0, W, H, spp=SPP, seed=seed, threads=threads)
            path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "%s_%s.npz" % (prec, name))
            np.savez_compressed(path, mean=c, sumsq=q, albedo=a, normal=n, spp=np.int64(SPP), seed=np.int64(seed),
                                signature=np.float64(scene_signature(s)))
            print("%s %s: %dx%d x %d spp in %.1f s -> %s" % (prec, name, W, H, SPP, time.time() - t0, path))


if __name__ == "__main__":
    main()
