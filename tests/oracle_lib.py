"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker / CPU baseline, never by the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
_lib = None
P = C.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        sig = {
            "orc_scene_new": (P, []),
            "orc_scene_free": (None, [P]),
            "orc_add_material": (C.c_int, [P, P]),
            "orc_add_mesh": (C.c_int, [P, P, C.c_int, P, P, P, C.c_int, P, C.c_int, P]),
            "orc_add_sphere": (C.c_int, [P, P, C.c_float, C.c_int]),
            "orc_set_env": (C.c_int, [P, P, C.c_int, C.c_int, C.c_float]),
            "orc_finalize": (C.c_int, [P, C.c_int]),
            "orc_trace": (None, [P, P, C.c_int, C.c_int, C.c_int, P, P, P, P, P]),
            "orc_bsdf_eval": (None, [P, C.c_int, P, P, C.c_int, P, P]),
            "orc_bsdf_sample": (None, [P, C.c_int, P, P, C.c_int, P, P, P, P, P]),
            "orc_mf": (None, [C.c_int, P, P, C.c_int, P, P, P]),
            "orc_mf_sample": (None, [C.c_int, P, P, C.c_int, P]),
            "orc_fresnel_conductor": (None, [P, P, P, C.c_int, P]),
            "orc_fresnel_dielectric": (None, [C.c_float, C.c_float, P, C.c_int, P, P, P]),
            "orc_warp": (None, [C.c_int, P, C.c_int, P]),
            "orc_frame": (None, [P, C.c_int, P, P]),
            "orc_pcg32": (None, [C.c_uint64, C.c_uint64, C.c_int, P]),
            "orc_pcg32_real": (None, [C.c_uint64, C.c_uint64, C.c_int, P]),
            "orc_stratified": (C.c_int, [C.c_int, C.c_int, C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_int, P, P]),
            "orc_cdf": (None, [P, C.c_int, P, C.c_int, P, P, P]),
            "orc_camera_rays": (None, [P, C.c_int, C.c_int, P, C.c_int, P, P]),
            "orc_light_sample": (None, [P, C.c_int, P, P, C.c_int, P]),
            "orc_light_pdf": (None, [P, C.c_int, P, C.c_int, P]),
            "orc_env_le": (None, [P, P, C.c_int, P]),
            "orc_light_cdf": (None, [P, P]),
            "orc_li": (None, [P, P, P, C.c_int, P, C.c_int, C.c_int, C.c_int, P]),
            "orc_render": (C.c_int, [P, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_int, C.c_uint64, P, P, P, P]),
            "orc_bvh_info": (C.c_int, [P, P]),
            "orc_filter_value": (C.c_float, [P, C.c_float, C.c_float]),
            "orc_set_medium": (C.c_int, [P, P, C.c_int, C.c_int, C.c_int, P, C.c_float, C.c_float]),
            "orc_set_integrator": (None, [P, C.c_int]),
            "orc_set_medium_homogeneous": (C.c_int, [P, P, P, C.c_float, C.c_float]),
            "orc_tonemap": (None, [P, C.c_int, C.c_int, C.c_int, C.c_float, P]),
            "orc_add_texture": (C.c_int, [P, P, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_float]),
            "orc_texture_eval": (None, [P, C.c_int, P, C.c_int, P]),
            "orc_material_at": (None, [P, P, C.c_int, P, P]),
            "orc_medium_sigma": (None, [P, P, C.c_int, P]),
            "orc_medium_track": (None, [P, C.c_int, P, C.c_int, P, C.c_int, P]),
            "orc_hg": (None, [C.c_float, C.c_int, P, P, C.c_int, P]),
            "gor_num_params": (C.c_int, []),
            "gor_bf16": (C.c_double, [C.c_double]),
            "gor_encode": (None, [P, C.c_int, P]),
            "gor_forward": (None, [P, C.c_int, P, C.c_int, P, P]),
            "gor_mix_logpdf": (None, [P, P, C.c_int, P]),
            "gor_loss_grad": (C.c_double, [P, C.c_int, P, C.c_int, P]),
        }
        for k, (r, a) in sig.items():
            f = getattr(lib, k)
            f.restype = r
            f.argtypes = a
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a if shape is None else a.reshape(shape)


def filter_params(rp):
    """(type, radius x, radius y, param 0, param 1) from render parameters' "filter" entry."""
    f = rp.get("filter") or {}
    return f32([f.get("type", 0), *f.get("radius", (0.5, 0.5)), *f.get("param", (0.0, 0.0))])


def hg_eval(g, wi, wo):
    wi = f32(wi, (-1, 3)); wo = f32(wo, (-1, 3)); out = np.zeros(wi.shape[0], np.float32)
    load().orc_hg(float(g), 0, _p(wi), _p(wo), wi.shape[0], _p(out))
    return out


def hg_sample(g, wi, u):
    """-> (n, 4): sampled direction, pdf."""
    wi = f32(wi, (-1, 3)); u = f32(u, (-1, 2)); out = np.zeros((wi.shape[0], 4), np.float32)
    load().orc_hg(float(g), 1, _p(wi), _p(u), wi.shape[0], _p(out))
    return out


def filter_value(f, x, y):
    """The oracle's table lookup for filter entry f (see filter_params) at offset (x, y)."""
    p = filter_params({"filter": f})
    return float(load().orc_filter_value(_p(p), float(x), float(y)))


def material_params(m):
    p = np.zeros(20, np.float32)
    p[0] = m["type"]
    p[1:4] = m.get("albedo", (0.5, 0.5, 0.5))
    p[4] = m.get("alpha", 0.2)
    p[5:8] = m.get("eta", (0, 0, 0))
    p[8:11] = m.get("k", (0, 0, 0))
    p[11] = m.get("eta_int", 1.5046)
    p[12] = m.get("eta_ext", 1.000277)
    p[13] = m.get("distribution", 0)
    p[14:17] = m.get("transmitted", (1, 1, 1))
    p[17] = m.get("tex_albedo", 0)
    p[18] = m.get("tex_alpha", 0)
    p[19] = m.get("tex_transmitted", 0)
    return p


class OracleScene:
    """The scene restated for the CPU oracle (two-level reference BVH)."""

    def __init__(self, sd):
        self.lib = load()
        self.h = self.lib.orc_scene_new()
        for t in getattr(sd, "textures", None) or []:
            rgb = f32(t["rgb"])
            self.lib.orc_add_texture(self.h, _p(rgb), rgb.shape[1], rgb.shape[0], float(t.get("uv_scale", 1.0)),
                                     int(t.get("invert_x", 0)), int(t.get("invert_y", 0)),
                                     float(t.get("value_scale", 1.0)))
        for m in sd.materials:
            mp = material_params(m)
            self.lib.orc_add_material(self.h, _p(mp))
        for m in sd.meshes:
            v = f32(m.vertices, (-1, 3)); idx = np.ascontiguousarray(m.indices, np.int32).reshape(-1, 3)
            n = None if m.normals is None else f32(m.normals, (-1, 3))
            uv = None if m.uvs is None else f32(m.uvs, (-1, 2))
            tm = None if m.tri_material is None else np.ascontiguousarray(m.tri_material, np.int32)
            e = None if m.emission is None else f32(m.emission, (3,))
            self.lib.orc_add_mesh(self.h, _p(v), v.shape[0], _p(n), _p(uv), _p(idx), idx.shape[0], _p(tm),
                                  int(m.material), _p(e))
        for s in sd.spheres:
            M = f32(s.to_world, (16,))
            self.lib.orc_add_sphere(self.h, _p(M), float(s.radius), int(s.material))
        if sd.env is not None:
            env = f32(sd.env)
            self.lib.orc_set_env(self.h, _p(env), env.shape[1], env.shape[0], float(sd.env_scale))
        # the oracle numbers strategies as the reference's enum (Integrator.h:18-22: all_lights,
        # one_light_uniform, one_light_weighted); the C ABI's 0 is the reference default (uniform)
        strategy = {0: 1, 1: 1, 2: 2, 3: 0}[int(sd.light_strategy)]
        self.nprim = self.lib.orc_finalize(self.h, strategy)
        med = getattr(sd, "medium", None)
        if med is not None and med.get("homogeneous"):
            sa, ss = f32(med["sigma_a"], (3,)), f32(med["sigma_s"], (3,))
            self.lib.orc_set_medium_homogeneous(self.h, _p(sa), _p(ss), float(med.get("g", 0.0)),
                                                float(med.get("scale", 1.0)))
        elif med is not None:
            g = f32(med["density"])
            alb = f32(med.get("albedo", (0.0, 0.0, 0.0)), (3,))
            self.lib.orc_set_medium(self.h, _p(g), g.shape[2], g.shape[1], g.shape[0], _p(alb), float(med.get("g", 0.0)),
                                    float(med.get("scale", 1.0)))
        self.lib.orc_set_integrator(self.h, int((sd.render or {}).get("integrator", 0)))
        self.sd = sd

    def __del__(self):
        try:
            self.lib.orc_scene_free(self.h)
        except Exception:
            pass

    def texture_eval(self, ti, uv):
        uv = f32(uv, (-1, 2)); out = np.zeros((uv.shape[0], 3), np.float32)
        self.lib.orc_texture_eval(self.h, int(ti), _p(uv), uv.shape[0], _p(out))
        return out

    def material_at(self, rays):
        """Material at the closest hit with textures evaluated -> (n x 8: albedo/reflectance rgb, alpha,
        transmitted rgb, 0), (n x 2 texture uv)."""
        rays = f32(rays, (-1, 8)); n = rays.shape[0]
        out = np.zeros((n, 8), np.float32); uv = np.zeros((n, 2), np.float32)
        self.lib.orc_material_at(self.h, _p(rays), n, _p(out), _p(uv))
        return out, uv

    def medium_sigma(self, pts):
        pts = f32(pts, (-1, 3)); out = np.zeros(pts.shape[0], np.float32)
        self.lib.orc_medium_sigma(self.h, _p(pts), pts.shape[0], _p(out))
        return out

    def medium_track(self, op, rays, u):
        """op 0 distance sampling -> (scattered, t, draws used, weight rgb); op 1 transmittance ->
        (tr, draws used, 0, tr rgb)."""
        rays = f32(rays, (-1, 8)); u = f32(u); n = rays.shape[0]
        u = u.reshape(n, -1); out = np.zeros((n, 6), np.float32)
        self.lib.orc_medium_track(self.h, int(op), _p(rays), n, _p(u), u.shape[1], _p(out))
        return out

    def trace(self, rays, any_hit=False, brute=False):
        rays = f32(rays, (-1, 8)); n = rays.shape[0]
        prim = np.zeros(n, np.int32); t = np.zeros(n, np.float32); uv = np.zeros((n, 2), np.float32)
        p = np.zeros((n, 3), np.float32); nrm = np.zeros((n, 3), np.float32)
        self.lib.orc_trace(self.h, _p(rays), n, int(any_hit), int(brute), _p(prim), _p(t), _p(uv), _p(p), _p(nrm))
        return prim, t, uv, p, nrm

    def bsdf_eval(self, mat, wi, wo):
        wi = f32(wi, (-1, 3)); wo = f32(wo, (-1, 3)); n = wi.shape[0]
        f = np.zeros((n, 3), np.float32); pdf = np.zeros(n, np.float32)
        self.lib.orc_bsdf_eval(self.h, mat, _p(wi), _p(wo), n, _p(f), _p(pdf))
        return f, pdf

    def bsdf_sample(self, mat, wi, u):
        wi = f32(wi, (-1, 3)); u = f32(u, (-1, 2)); n = wi.shape[0]
        wo = np.zeros((n, 3), np.float32); w = np.zeros((n, 3), np.float32); pdf = np.zeros(n, np.float32)
        eta = np.zeros(n, np.float32); typ = np.zeros(n, np.int32)
        self.lib.orc_bsdf_sample(self.h, mat, _p(wi), _p(u), n, _p(wo), _p(w), _p(pdf), _p(eta), _p(typ))
        return wo, w, pdf, eta, typ

    def light_sample(self, li, p, u):
        p = f32(p, (-1, 3)); u = f32(u, (-1, 2)); n = p.shape[0]
        out = np.zeros((n, 14), np.float32)
        self.lib.orc_light_sample(self.h, li, _p(p), _p(u), n, _p(out))
        return out

    def light_pdf(self, li, q):
        q = f32(q, (-1, 12)); n = q.shape[0]
        out = np.zeros(n, np.float32)
        self.lib.orc_light_pdf(self.h, li, _p(q), n, _p(out))
        return out

    def env_le(self, dirs):
        d = f32(dirs, (-1, 3)); n = d.shape[0]
        out = np.zeros((n, 3), np.float32)
        self.lib.orc_env_le(self.h, _p(d), n, _p(out))
        return out

    def light_cdf(self, nlights):
        out = np.zeros(nlights + 1, np.float32)
        self.lib.orc_light_cdf(self.h, _p(out))
        return out

    def li(self, ray, u2, u1, min_depth=3, max_depth=1000):
        ray = f32(ray, (8,)); u2 = f32(u2, (-1, 2)); u1 = f32(u1, (-1,))
        L = np.zeros(3, np.float32)
        self.lib.orc_li(self.h, _p(ray), _p(u2), u2.shape[0], _p(u1), u1.shape[0], min_depth, max_depth, _p(L))
        return L

    def render(self, camera=None, spp=None, threads=1, blocks=(0, 0, 1), sampler=1, min_depth=None, max_depth=None,
               seed_offset=0):
        """Reference schedule (32x32 blocks seeded (blockX, blockY), LIFO job pop). blocks = (first, end,
        stride) over the blocks in creation order; returns (rgb sums, weights, [rays, nans], blocks run)."""
        sd = self.sd
        cam = dict(sd.camera if camera is None else camera)
        rp = sd.render
        W, H = int(cam["width"]), int(cam["height"])
        cv = camera_vector(cam)
        rgb = np.zeros((H, W, 3), np.float64); wgt = np.zeros((H, W), np.float64)
        stats = np.zeros(2, np.uint64)
        nblocks = self.lib.orc_render(self.h, _p(cv), W, H, int(spp or rp["spp"]), int(sampler),
                                      int(rp["min_depth"] if min_depth is None else min_depth),
                                      int(rp["max_depth"] if max_depth is None else max_depth), int(threads),
                                      int(blocks[0]), int(blocks[1]), int(blocks[2] if len(blocks) > 2 else 1),
                                      int(seed_offset), _p(rgb), _p(wgt), _p(stats), _p(filter_params(rp)))
        return rgb, wgt, stats, nblocks


def camera_vector(cam):
    """origin, target, up, fov, near, far, then a flag and the row-major camera-to-world matrix when
    the camera carries one ("to_world": a scene file's toWorld, e.g. a lookAt with a mirroring scale)."""
    v = [*cam["origin"], *cam["target"], *cam["up"], cam["fov"], cam.get("near", 1e-4), cam.get("far", 1e4)]
    M = cam.get("to_world")
    v += [1.0, *np.asarray(M, np.float64).reshape(16)] if M is not None else [0.0] * 17
    return f32(v)


def camera_rays(camera, pxy):
    lib = load()
    cam = camera
    cv = camera_vector(cam)
    pxy = f32(pxy, (-1, 2)); n = pxy.shape[0]
    rays = np.zeros((n, 8), np.float32); mats = np.zeros(32, np.float32)
    lib.orc_camera_rays(_p(cv), int(cam["width"]), int(cam["height"]), _p(pxy), n, _p(rays), _p(mats))
    return rays, mats


def pcg32(initstate, initseq, n):
    out = np.zeros(n, np.uint32)
    load().orc_pcg32(initstate, initseq, n, _p(out))
    return out


def pcg32_real(initstate, initseq, n):
    out = np.zeros(n, np.float32)
    load().orc_pcg32_real(initstate, initseq, n, _p(out))
    return out


def stratified(spp, dim, bx, by, npix, ndraw2, ndraw1):
    n = spp
    i = 0
    while i < n and i * i < n:
        i += 1
    nn = i * i if i * i != n else n
    o2 = np.zeros((npix, nn, ndraw2, 2), np.float32); o1 = np.zeros((npix, nn, ndraw1), np.float32)
    r = load().orc_stratified(spp, dim, bx, by, npix, ndraw2, ndraw1, _p(o2), _p(o1))
    assert r == nn
    return o2, o1


def warp(which, u):
    u = f32(u, (-1, 2)); n = u.shape[0]
    out = np.zeros((n, 2 if which == 0 else 3), np.float32)
    load().orc_warp(which, _p(u), n, _p(out))
    return out


def mf(dist, wh, alpha, v=None):
    wh = f32(wh, (-1, 3)); n = wh.shape[0]; alpha = f32(np.broadcast_to(alpha, (n,)))
    D = np.zeros(n, np.float32); G1 = np.zeros(n, np.float32) if v is not None else None
    vv = None if v is None else f32(v, (-1, 3))
    load().orc_mf(dist, _p(wh), _p(alpha), n, _p(D), _p(G1), _p(vv))
    return D, G1


def mf_sample(dist, u, alpha):
    u = f32(u, (-1, 2)); n = u.shape[0]; alpha = f32(np.broadcast_to(alpha, (n,)))
    out = np.zeros((n, 3), np.float32)
    load().orc_mf_sample(dist, _p(u), _p(alpha), n, _p(out))
    return out


def fresnel_conductor(eta, k, cosi):
    cosi = f32(cosi, (-1,)); n = cosi.shape[0]
    out = np.zeros((n, 3), np.float32)
    e = f32(eta, (3,)); kk = f32(k, (3,))
    load().orc_fresnel_conductor(_p(e), _p(kk), _p(cosi), n, _p(out))
    return out


def fresnel_dielectric(eta_ext, eta_int, cosi):
    cosi = f32(cosi, (-1,)); n = cosi.shape[0]
    fr = np.zeros(n, np.float32); ct = np.zeros(n, np.float32); re = np.zeros(n, np.float32)
    load().orc_fresnel_dielectric(eta_ext, eta_int, _p(cosi), n, _p(fr), _p(ct), _p(re))
    return fr, ct, re


def frame(n):
    n = f32(n, (-1, 3)); k = n.shape[0]
    s = np.zeros((k, 3), np.float32); t = np.zeros((k, 3), np.float32)
    load().orc_frame(_p(n), k, _p(s), _p(t))
    return s, t


def cdf(w, u):
    w = f32(w, (-1,)); u = f32(u, (-1,)); n = u.shape[0]
    idx = np.zeros(n, np.int32); ur = np.zeros(n, np.float32); pdf = np.zeros(n, np.float32)
    load().orc_cdf(_p(w), w.shape[0], _p(u), n, _p(idx), _p(ur), _p(pdf))
    return idx, ur, pdf


# ---------------------------------------------------------------- guiding network (guide_oracle.cpp)
GUIDE_PARAMS = 20200
GUIDE_SLICES = {  # name -> (offset, shape) in the FP32 parameter vector
    "W1": (0, (64, 144)), "b1": (9216, (64,)), "W2": (9280, (64, 64)), "b2": (13376, (64,)),
    "W3": (13440, (64, 64)), "b3": (17536, (64,)), "W4": (17600, (40, 64)), "b4": (20160, (40,)),
}


def guide_init(seed, scale=1.0):
    """Random parameters (Xavier-uniform weights, small biases) for tests."""
    rng = np.random.default_rng(seed)
    p = np.zeros(GUIDE_PARAMS, np.float32)
    for name, (off, shp) in GUIDE_SLICES.items():
        n = int(np.prod(shp))
        if name[0] == "W":
            lim = scale * np.sqrt(6.0 / (shp[0] + shp[1]))
            p[off:off + n] = rng.uniform(-lim, lim, n)
        else:
            p[off:off + n] = rng.uniform(-0.05, 0.05, n)
    return p


def guide_encode(in6):
    lib = load()
    in6 = f32(in6, (-1, 6)); n = in6.shape[0]
    out = np.zeros((n, 144), np.float64)
    lib.gor_encode(_p(in6), n, _p(out))
    return out


def guide_forward(params, in6, emulate_bf16=True):
    lib = load()
    params = f32(params); in6 = f32(in6, (-1, 6)); n = in6.shape[0]
    out = np.zeros((n, 40)); mix = np.zeros((n, 40))
    lib.gor_forward(_p(params), int(emulate_bf16), _p(in6), n, _p(out), _p(mix))
    return out, mix


def guide_mix_logpdf(mix, w):
    lib = load()
    mix = np.ascontiguousarray(mix, np.float64).reshape(-1, 40); w = np.ascontiguousarray(w, np.float64).reshape(-1, 3)
    out = np.zeros(w.shape[0])
    lib.gor_mix_logpdf(_p(mix), _p(w), w.shape[0], _p(out))
    return out


def guide_loss_grad(params, samples, emulate_bf16=True):
    lib = load()
    params = f32(params); samples = f32(samples, (-1, 10))
    g = np.zeros(GUIDE_PARAMS)
    loss = lib.gor_loss_grad(_p(params), int(emulate_bf16), _p(samples), samples.shape[0], _p(g))
    return loss, g


def tonemap(rgbw, mode, exposure=0.5):
    """The oracle's tone mappers (core/main.cpp:185-290) -> H x W x 3 uint8."""
    f = f32(rgbw)
    H, W = f.shape[0], f.shape[1]
    out = np.zeros((H, W, 3), np.uint8)
    load().orc_tonemap(_p(f), W, H, int(mode), float(exposure), _p(out))
    return out
