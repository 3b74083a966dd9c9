"""Writes the committed fixtures under tests/golden/ from the CPU oracle.

  cornell64_ref.npy   Cornell box (same scene and camera as the benchmark scene, 64x64 film),
                      65536 spp stratified, FP64 accumulation (the high-spp reference image).
  bsdf_vectors.npz    BSDF eval/pdf inputs and outputs for the five test materials.
  trace_cornell.npz   rays and closest-hit ids / distances for the Cornell scene.

Run: python tests/golden/make_golden.py [--threads N]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as O  # noqa: E402
from raytracer_amd import rtg, scenes  # noqa: E402

GOLDEN_MATS = [
    scenes.diffuse((0.73, 0.5, 0.2)),
    scenes.conductor(0.15),
    scenes.conductor(0.4, dist=rtg.DIST_BECKMANN),
    scenes.dielectric(0.1),
    scenes.dielectric(0.3, dist=rtg.DIST_BECKMANN),
]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count())
    ap.add_argument("--skip-image", action="store_true")
    a = ap.parse_args()
    rng = np.random.default_rng(1234)
    # BSDF vectors
    n = 256
    wi = rng.normal(size=(len(GOLDEN_MATS), n, 3)); wi /= np.linalg.norm(wi, axis=-1, keepdims=True)
    wo = rng.normal(size=(len(GOLDEN_MATS), n, 3)); wo /= np.linalg.norm(wo, axis=-1, keepdims=True)
    f = np.zeros((len(GOLDEN_MATS), n, 3), np.float32); pdf = np.zeros((len(GOLDEN_MATS), n), np.float32)
    for i, m in enumerate(GOLDEN_MATS):
        o = O.OracleScene(scenes.SceneDesc(materials=[m]))
        f[i], pdf[i] = o.bsdf_eval(0, wi[i], wo[i])
    np.savez_compressed(os.path.join(HERE, "bsdf_vectors.npz"), wi=wi.astype(np.float32), wo=wo.astype(np.float32),
                        f=f, pdf=pdf)
    # trace vectors
    sd = scenes.cornell()
    o = O.OracleScene(sd)
    m = 4096
    org = rng.uniform([-0.45, 0.05, -0.45], [0.45, 0.95, 0.45], (m, 3))
    d = rng.normal(size=(m, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((m, 8), np.float32); rays[:, :3] = org; rays[:, 3:6] = d; rays[:, 6] = 1e-4; rays[:, 7] = 1e8
    prim, t, uv, p, nrm = o.trace(rays, brute=True)
    np.savez_compressed(os.path.join(HERE, "trace_cornell.npz"), rays=rays, prim=prim, t=t, uv=uv)
    if a.skip_image:
        return
    cam = dict(sd.camera, width=64, height=64)
    t0 = time.time()
    rgb, w, st, _ = o.render(camera=cam, spp=65536, threads=a.threads)
    img = (rgb / w[..., None]).astype(np.float32)
    np.save(os.path.join(HERE, "cornell64_ref.npy"), img)
    print(f"reference image: {time.time() - t0:.1f} s, mean {img.mean(axis=(0, 1))}, nan samples {st[1]}")


if __name__ == "__main__":
    main()
