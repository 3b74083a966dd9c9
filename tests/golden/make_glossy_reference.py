"""Writes tests/golden/glossy_cornell_ref.npz: the converged plain image of the glossy-indirect
Cornell scene (scenes.glossy_cornell, 512x512, PathTracingMIS, stratified) that judges the error of
the plain and guided renders (bench.py guiding leg, tests/test_gpu_guide.py).

It is rendered on the GPU by the plain tracer as `pieces` independent renders of `spp` samples per
pixel each (seeds 1000000 + k, disjoint from every seed the benchmark and the tests use), resolved
per piece and accumulated in FP64. Stored:
  mean      (512, 512, 3) float16  mean of the resolved pieces (rounded through float32; the float16
                                   rounding adds about 1e-8 to a relMSE, the mean's own noise 8e-5)
  var_mean  (512, 512, 3) float16  variance of that mean (sample variance of the pieces / pieces):
                                   the reference's own per-pixel noise, subtracted by error measures
  spp, pieces, seed0, scene        what was rendered

Run on a GPU: python tests/golden/make_glossy_reference.py [--pieces 64] [--spp 16384] [--out PATH]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from raytracer_amd import rtg, scenes  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pieces", type=int, default=64)
    ap.add_argument("--spp", type=int, default=16384)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--out", default=os.path.join(HERE, "glossy_cornell_ref.npz"))
    a = ap.parse_args()
    sd = scenes.glossy_cornell(width=a.size, height=a.size)
    ctx = rtg.Context(0)
    g = ctx.upload(sd)
    s1 = np.zeros((a.size, a.size, 3), np.float64)
    s2 = np.zeros_like(s1)
    seed0 = 1000000
    t0 = time.perf_counter()
    for k in range(a.pieces):
        out, st = g.render(sd.camera, dict(sd.render, spp=a.spp, seed=seed0 + k))
        img = rtg.resolve(out).astype(np.float64)
        s1 += img
        s2 += img * img
        print(f"piece {k + 1}/{a.pieces}: {st.samples} samples, {st.milliseconds / 1e3:.1f} s "
              f"(total {time.perf_counter() - t0:.0f} s)", flush=True)
    n = a.pieces
    mean = s1 / n
    var_mean = np.maximum(s2 / n - mean * mean, 0.0) * n / (n - 1) / n
    np.savez_compressed(a.out, mean=mean.astype(np.float32).astype(np.float16), var_mean=var_mean.astype(np.float16),
                        spp=np.int64(a.spp), pieces=np.int64(n), seed0=np.int64(seed0),
                        scene=np.array(f"glossy_cornell {a.size}x{a.size}"))
    rel_noise = float(np.mean(var_mean / (mean * mean + 1e-2)))
    print(f"wrote {a.out}: {n} x {a.spp} spp, relMSE of the reference itself {rel_noise:.3e}", flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
