"""ctypes binding of librtg.so (include/rtg.h).

Loading fails loudly when the HIP library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librtg.so")

RTG_OK = 0
ERRORS = {-1: "INVALID", -2: "HIP", -3: "OOM", -4: "PARSE", -5: "NOT_READY", -6: "CANCELLED"}

MAT_DIFFUSE, MAT_ROUGH_CONDUCTOR, MAT_ROUGH_DIELECTRIC, MAT_CONDUCTOR, MAT_DIELECTRIC = 0, 1, 2, 3, 4
DIST_BECKMANN, DIST_GGX = 0, 2
FILTER_BOX, FILTER_GAUSSIAN, FILTER_MITCHELL = 0, 1, 2
# 0 = the reference's default strategy (one_light_uniform, integrator/Integrator.cpp:14)
LIGHTS_DEFAULT, LIGHTS_UNIFORM, LIGHTS_WEIGHTED, LIGHTS_ALL = 0, 1, 2, 3
SAMPLER_RANDOM, SAMPLER_STRATIFIED = 0, 1
INTEGRATOR_PATH_MIS, INTEGRATOR_VOLPATH = 0, 1
GUIDING_OFF, GUIDING_LEARNED, GUIDING_EXIT = 0, 1, 2
TONEMAP_CLAMP, TONEMAP_GAMMA, TONEMAP_REINHARD = 0, 1, 2


class rtg_material(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("distribution", C.c_int32), ("albedo", C.c_float * 3), ("alpha", C.c_float),
        ("eta", C.c_float * 3), ("k", C.c_float * 3), ("eta_int", C.c_float), ("eta_ext", C.c_float),
        ("transmitted", C.c_float * 3), ("tex_albedo", C.c_int32), ("tex_alpha", C.c_int32),
        ("tex_transmitted", C.c_int32),
    ]


class rtg_texture_desc(C.Structure):
    _fields_ = [
        ("rgb", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32), ("uv_scale", C.c_float),
        ("invert_x", C.c_int32), ("invert_y", C.c_int32), ("value_scale", C.c_float),
    ]


class rtg_mesh_desc(C.Structure):
    _fields_ = [
        ("vertices", C.c_void_p), ("normals", C.c_void_p), ("uvs", C.c_void_p), ("num_vertices", C.c_int32),
        ("indices", C.c_void_p), ("num_triangles", C.c_int32), ("tri_material", C.c_void_p),
        ("material", C.c_int32), ("emission", C.c_void_p),
    ]


class rtg_sphere_desc(C.Structure):
    _fields_ = [("to_world", C.c_float * 16), ("radius", C.c_float), ("material", C.c_int32)]


class rtg_scene_desc(C.Structure):
    _fields_ = [
        ("materials", C.c_void_p), ("num_materials", C.c_int32), ("meshes", C.c_void_p), ("num_meshes", C.c_int32),
        ("spheres", C.c_void_p), ("num_spheres", C.c_int32), ("env_rgb", C.c_void_p), ("env_width", C.c_int32),
        ("env_height", C.c_int32), ("env_scale", C.c_float), ("light_strategy", C.c_int32),
        ("medium_density", C.c_void_p), ("medium_res", C.c_int32 * 3), ("medium_albedo", C.c_float * 3),
        ("medium_g", C.c_float), ("medium_scale", C.c_float), ("textures", C.c_void_p), ("num_textures", C.c_int32),
        ("medium_homogeneous", C.c_int32), ("medium_sigma_a", C.c_float * 3), ("medium_sigma_s", C.c_float * 3),
    ]


class rtg_render_desc(C.Structure):
    _fields_ = [
        ("origin", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3), ("fov", C.c_float),
        ("near_clip", C.c_float), ("far_clip", C.c_float), ("width", C.c_int32), ("height", C.c_int32),
        ("spp", C.c_int32), ("sampler", C.c_int32), ("dimension", C.c_int32), ("min_depth", C.c_int32),
        ("max_depth", C.c_int32), ("seed", C.c_uint64), ("guiding", C.c_int32),
        ("use_camera_matrix", C.c_int32), ("camera_to_world", C.c_float * 16),
        ("filter", C.c_int32), ("filter_radius", C.c_float * 2), ("filter_param", C.c_float * 2),
        ("integrator", C.c_int32),
    ]


class rtg_stats(C.Structure):
    _fields_ = [
        ("samples", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
        ("nan_samples", C.c_uint64), ("iterations", C.c_uint64), ("milliseconds", C.c_double),
        ("nodes_visited", C.c_uint64), ("prims_tested", C.c_uint64), ("shadow_nodes_visited", C.c_uint64),
        ("shadow_prims_tested", C.c_uint64), ("stage_ms", C.c_double * 8), ("stage_launches", C.c_uint64 * 8),
        ("guide_train_steps", C.c_uint64), ("guide_samples", C.c_uint64), ("guide_loss", C.c_double),
        ("guide_train_ms", C.c_double), ("guide_queries", C.c_uint64), ("stack_spills", C.c_uint64),
    ]


class rtg_scene_info(C.Structure):
    _fields_ = [
        ("triangles", C.c_int32), ("spheres", C.c_int32), ("materials", C.c_int32), ("lights", C.c_int32),
        ("env_width", C.c_int32), ("env_height", C.c_int32), ("bbox_min", C.c_float * 3), ("bbox_max", C.c_float * 3),
        ("medium_res", C.c_int32 * 3), ("medium_albedo", C.c_float * 3), ("medium_g", C.c_float),
        ("medium_scale", C.c_float), ("medium_sigma_max", C.c_float), ("textures", C.c_int32),
        ("medium_kind", C.c_int32), ("medium_sigma_a", C.c_float * 3), ("medium_sigma_s", C.c_float * 3),
        ("traversal_stack", C.c_int32),
    ]


class rtg_server_stats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64), ("admitted", C.c_uint64), ("completed", C.c_uint64), ("cancelled", C.c_uint64),
        ("iterations", C.c_uint64), ("samples", C.c_uint64), ("film_bytes", C.c_uint64),
        ("film_bytes_peak", C.c_uint64), ("running", C.c_int32), ("pending", C.c_int32),
        ("max_running", C.c_int32), ("reserved", C.c_int32), ("graph_launches", C.c_uint64),
    ]


class rtg_tuning(C.Structure):
    _fields_ = [("pool_slots", C.c_int32), ("trace_waves", C.c_int32), ("trace_refill", C.c_int32),
                ("graphs", C.c_int32)]


STAGES = ("regen", "extend", "guide", "shade_diffuse", "shade_conductor", "shade_dielectric", "shade_miss", "shadow")
PROFILE_EVENTS, PROFILE_COUNT = 1, 2


# exported entry points: name -> (restype, argtypes)
P = C.c_void_p
SIGNATURES = {
    "rtg_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "rtg_ctx_destroy": (C.c_int, [P]),
    "rtg_last_error": (C.c_char_p, [P]),
    "rtg_ctx_set_pool": (C.c_int, [P, C.c_int32]),
    "rtg_ctx_set_profiling": (C.c_int, [P, C.c_int32]),
    "rtg_ctx_get_tuning": (C.c_int, [P, C.POINTER(rtg_tuning)]),
    "rtg_ctx_set_tuning": (C.c_int, [P, C.POINTER(rtg_tuning)]),
    "rtg_autotune": (C.c_int, [P, P, C.POINTER(rtg_render_desc), C.c_uint64, C.POINTER(rtg_tuning),
                               C.POINTER(C.c_double)]),
    "rtg_scene_upload": (C.c_int, [P, C.POINTER(rtg_scene_desc), C.POINTER(P)]),
    "rtg_scene_load_xml": (C.c_int, [P, C.c_char_p, C.POINTER(P), C.POINTER(rtg_render_desc)]),
    "rtg_scene_release": (C.c_int, [P, P]),
    "rtg_scene_bytes": (C.c_int, [P, C.POINTER(C.c_uint64)]),
    "rtg_scene_query": (C.c_int, [P, C.POINTER(rtg_scene_info)]),
    "rtg_scene_material": (C.c_int, [P, C.c_int32, C.POINTER(rtg_material)]),
    "rtg_render": (C.c_int, [P, P, C.POINTER(rtg_render_desc), P, C.POINTER(rtg_stats)]),
    "rtg_render_device": (C.c_int, [P, P, C.POINTER(rtg_render_desc), P, C.POINTER(rtg_stats)]),
    "rtg_trace": (C.c_int, [P, P, P, C.c_int32, C.c_int32, P, P, P, P, P]),
    "rtg_trace_spills": (C.c_int, [P, C.POINTER(C.c_uint64)]),
    "rtg_trace_device_noalloc": (C.c_int, [P, P, P, C.c_int32, C.c_int32, P, P, P, P, P]),
    "rtg_sampler_batch": (C.c_int, [P, C.POINTER(rtg_render_desc), P, C.c_int32, C.c_int32, C.c_int32, P, P]),
    "rtg_bsdf_batch": (C.c_int, [P, C.POINTER(rtg_material), C.c_int32, P, P, C.c_int32, P, P, P, P]),
    "rtg_medium_batch": (C.c_int, [P, P, C.c_int32, P, C.c_int32, P, C.c_int32, P]),
    "rtg_light_batch": (C.c_int, [P, P, C.c_int32, C.c_int32, P, C.c_int32, P, P]),
    "rtg_camera_rays": (C.c_int, [P, P, P, C.c_int32, P]),
    "rtg_tonemap": (C.c_int, [P, P, C.c_int32, C.c_int32, C.c_int32, C.c_float, P]),
    "rtg_spectrum_to_rgb": (C.c_int, [P, P, C.c_int32, P]),
    "rtg_server_create": (C.c_int, [P, C.c_uint64, C.POINTER(P)]),
    "rtg_server_destroy": (C.c_int, [P]),
    "rtg_submit": (C.c_int, [P, P, C.POINTER(rtg_render_desc), C.POINTER(C.c_int64)]),
    "rtg_server_step": (C.c_int, [P, C.c_double]),
    "rtg_poll": (C.c_int, [P, C.c_int64, P, C.POINTER(C.c_double)]),
    "rtg_wait": (C.c_int, [P, C.c_int64, C.c_double, P, C.POINTER(C.c_double)]),
    "rtg_server_start": (C.c_int, [P]),
    "rtg_cancel": (C.c_int, [P, C.c_int64]),
    "rtg_server_query": (C.c_int, [P, C.POINTER(rtg_server_stats)]),
    "rtg_guide_eval": (C.c_int, [P, P, P, C.c_int32, P]),
    "rtg_guide_grad": (C.c_int, [P, P, P, C.c_int32, P, P]),
    "rtg_guide_train": (C.c_int, [P, P, P, C.c_int32, C.c_int32, C.c_float, P, P]),
    "rtg_guide_sample": (C.c_int, [P, P, P, C.c_int32, P, P]),
    "rtg_guide_params": (C.c_int, [P, P]),
}
GUIDE_PARAMS = 20200

_lib = None


def load():
    """Load librtg.so; raises if it is absent (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"librtg.so not found at {LIB_PATH}; run __graft_entry__.build() first")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


class RtgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def _ptr(a):
    return None if a is None else a.ctypes.data


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def material_struct(m):
    s = rtg_material()
    s.type = int(m["type"])
    s.distribution = int(m.get("distribution", DIST_BECKMANN))
    s.albedo[:] = [float(x) for x in m.get("albedo", (0.5, 0.5, 0.5))]
    s.alpha = float(m.get("alpha", 0.2))
    s.eta[:] = [float(x) for x in m.get("eta", (0.0, 0.0, 0.0))]
    s.k[:] = [float(x) for x in m.get("k", (0.0, 0.0, 0.0))]
    s.eta_int = float(m.get("eta_int", 1.5046))
    s.eta_ext = float(m.get("eta_ext", 1.000277))
    s.transmitted[:] = [float(x) for x in m.get("transmitted", (1.0, 1.0, 1.0))]
    # image textures: 1-based indices into the scene's texture list, 0 = the constant above
    s.tex_albedo = int(m.get("tex_albedo", 0))
    s.tex_alpha = int(m.get("tex_alpha", 0))
    s.tex_transmitted = int(m.get("tex_transmitted", 0))
    return s


def render_struct(cam, rp):
    d = rtg_render_desc()
    d.origin[:] = [float(x) for x in cam["origin"]]
    d.target[:] = [float(x) for x in cam["target"]]
    d.up[:] = [float(x) for x in cam["up"]]
    d.fov = float(cam["fov"])
    d.near_clip = float(cam.get("near", 1e-4))
    d.far_clip = float(cam.get("far", 1e4))
    d.width = int(cam["width"])
    d.height = int(cam["height"])
    d.spp = int(rp["spp"])
    d.sampler = int(rp.get("sampler", SAMPLER_STRATIFIED))
    d.dimension = int(rp.get("dimension", 4))
    d.min_depth = int(rp.get("min_depth", 3))
    d.max_depth = int(rp.get("max_depth", 1000))
    d.seed = int(rp.get("seed", 1))
    d.guiding = int(rp.get("guiding", 0))
    if cam.get("to_world") is not None:
        d.use_camera_matrix = 1
        d.camera_to_world[:] = [float(x) for x in np.asarray(cam["to_world"], np.float64).reshape(16)]
    f = rp.get("filter")
    if f:
        d.filter = int(f.get("type", FILTER_BOX))
        d.filter_radius[:] = [float(x) for x in f.get("radius", (0.5, 0.5))]
        d.filter_param[:] = [float(x) for x in f.get("param", (0.0, 0.0))]
    d.integrator = int(rp.get("integrator", INTEGRATOR_PATH_MIS))
    return d


def filter_desc(kind="box", radius=(0.5, 0.5), alpha=2.0, omega=None, B=1.0 / 3, C_=1.0 / 3):
    """Render-parameter entry for a reconstruction filter, with the reference's defaults
    (GaussianFilter: alpha 1/(2 0.5^2), omega = radius x; MitchellNetravali: B = C = 1/3)."""
    if kind == "box":
        return dict(type=FILTER_BOX, radius=tuple(radius), param=(0.0, 0.0))
    if kind == "gaussian":
        return dict(type=FILTER_GAUSSIAN, radius=tuple(radius), param=(alpha, radius[0] if omega is None else omega))
    if kind == "mitchell":
        return dict(type=FILTER_MITCHELL, radius=tuple(radius), param=(B, C_))
    raise ValueError(kind)


class Context:
    """One GPU. Owns the wavefront path pool, the scene table and the job table."""

    def __init__(self, device=0):
        self.lib = load()
        h = P()
        rc = self.lib.rtg_ctx_create(device, C.byref(h))
        if rc != RTG_OK:
            raise RtgError(rc, "rtg_ctx_create failed (no gfx950 device?)")
        self.h = h
        self._servers = []

    def check(self, rc):
        if rc != RTG_OK:
            raise RtgError(rc, self.lib.rtg_last_error(self.h).decode())

    def sampler_draws(self, camera, params, pixel_sample, n2, n1):
        """Draws of the renderer's sampler (rtg_sampler_batch) for (pixel, sample) pairs ->
        (n x n2 x 2, n x n1)."""
        d = render_struct(camera, params)
        ps = np.ascontiguousarray(pixel_sample, dtype=np.uint32).reshape(-1, 2)
        n = ps.shape[0]
        o2 = np.zeros((n, n2, 2), np.float32); o1 = np.zeros((n, n1), np.float32)
        self.check(self.lib.rtg_sampler_batch(self.h, C.byref(d), _ptr(ps), n, n2, n1, _ptr(o2), _ptr(o1)))
        return o2, o1

    def trace_spills(self):
        """Stack pushes of this context's last trace that spilled past the on-chip entries."""
        v = C.c_uint64(0)
        self.check(self.lib.rtg_trace_spills(self.h, C.byref(v)))
        return v.value

    def close(self):
        if self.h:
            for srv in self._servers:  # rtg_ctx_destroy destroys a live server
                srv.h = None
            self.lib.rtg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_xml(self, path):
        """Scene file in the reference's XML schema -> (Scene, camera dict, render params dict)."""
        h = P()
        d = rtg_render_desc()
        self.check(self.lib.rtg_scene_load_xml(self.h, str(path).encode(), C.byref(h), C.byref(d)))
        cam = dict(origin=(0, 0, 0), target=(0, 0, 1), up=(0, 1, 0), fov=d.fov, near=d.near_clip, far=d.far_clip,
                   width=d.width, height=d.height)
        if d.use_camera_matrix:
            cam["to_world"] = np.array(d.camera_to_world[:], np.float64).reshape(4, 4)
        rp = dict(spp=d.spp, sampler=d.sampler, dimension=d.dimension, min_depth=d.min_depth,
                  max_depth=d.max_depth, seed=d.seed, integrator=d.integrator,
                  filter=dict(type=d.filter, radius=tuple(d.filter_radius[:]), param=tuple(d.filter_param[:])))
        return Scene.from_handle(self, h), cam, rp

    def tonemap(self, rgbw, mode=TONEMAP_CLAMP, exposure=0.5):
        """The reference's displayed / saved 8-bit image of a film (rtg_tonemap) -> H x W x 3 uint8."""
        f = _f32(rgbw)
        H, W = f.shape[0], f.shape[1]
        out = np.zeros((H, W, 3), np.uint8)
        self.check(self.lib.rtg_tonemap(self.h, _ptr(f), W, H, int(mode), float(exposure), _ptr(out)))
        return out

    def tuning(self):
        t = rtg_tuning()
        self.check(self.lib.rtg_ctx_get_tuning(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def set_tuning(self, **kw):
        t = self.tuning()
        t.update(kw)
        s = rtg_tuning(**{k: int(v) for k, v in t.items()})
        self.check(self.lib.rtg_ctx_set_tuning(self.h, C.byref(s)))

    def autotune(self, scene, camera, params, trial_samples=0):
        """rtg_autotune -> (chosen configuration dict, its samples/s in the trial renders)."""
        d = render_struct(camera, params)
        t = rtg_tuning()
        r = C.c_double(0.0)
        self.check(self.lib.rtg_autotune(self.h, scene.h, C.byref(d), C.c_uint64(int(trial_samples)), C.byref(t),
                                         C.byref(r)))
        return {k: getattr(t, k) for k, _ in t._fields_}, r.value

    def camera_rays(self, camera, pxy):
        """Camera tap (rtg_camera_rays): film positions (n, 2) -> rays (n, 8) of the camera dict."""
        d = render_struct(camera, {"spp": 1})
        pxy = _f32(pxy, (-1, 2)); n = pxy.shape[0]
        out = np.zeros((n, 8), np.float32)
        self.check(self.lib.rtg_camera_rays(self.h, C.byref(d), _ptr(pxy), n, _ptr(out)))
        return out

    def set_pool(self, slots):
        self.check(self.lib.rtg_ctx_set_pool(self.h, int(slots)))

    def set_profiling(self, flags):
        self.check(self.lib.rtg_ctx_set_profiling(self.h, int(flags)))

    def upload(self, scene):
        return Scene(self, scene)

    def bsdf_eval(self, material, wi, wo):
        wi = _f32(wi, (-1, 3)); wo = _f32(wo, (-1, 3))
        n = wi.shape[0]
        f = np.zeros((n, 3), np.float32); pdf = np.zeros(n, np.float32)
        m = material_struct(material)
        self.check(self.lib.rtg_bsdf_batch(self.h, C.byref(m), 0, _ptr(wi), _ptr(wo), n, None, _ptr(f), _ptr(pdf), None))
        return f, pdf

    def bsdf_density(self, material, wi, wo):
        """Sampling density of bsdf_sample at wo, its weight there and its reported pdf."""
        wi = _f32(wi, (-1, 3)); wo = _f32(wo, (-1, 3))
        n = wi.shape[0]
        f = np.zeros((n, 3), np.float32); ps = np.zeros(n, np.float32); pref = np.zeros(n, np.float32)
        m = material_struct(material)
        self.check(self.lib.rtg_bsdf_batch(self.h, C.byref(m), 2, _ptr(wi), _ptr(wo), n, None, _ptr(f), _ptr(ps),
                                           _ptr(pref)))
        return ps, f, pref

    def guide_eval(self, params, in6):
        params = _f32(params, (GUIDE_PARAMS,)); in6 = _f32(in6, (-1, 6))
        out = np.zeros((in6.shape[0], 40), np.float32)
        self.check(self.lib.rtg_guide_eval(self.h, _ptr(params), _ptr(in6), in6.shape[0], _ptr(out)))
        return out

    def guide_grad(self, params, samples):
        params = _f32(params, (GUIDE_PARAMS,)); samples = _f32(samples, (-1, 10))
        loss = np.zeros(1, np.float32); grad = np.zeros(GUIDE_PARAMS, np.float32)
        self.check(self.lib.rtg_guide_grad(self.h, _ptr(params), _ptr(samples), samples.shape[0], _ptr(loss),
                                           _ptr(grad)))
        return float(loss[0]), grad

    def guide_train(self, params, samples, steps, lr=1e-3):
        params = _f32(params, (GUIDE_PARAMS,)); samples = _f32(samples, (-1, 10))
        out = np.zeros(GUIDE_PARAMS, np.float32); losses = np.zeros(max(1, steps), np.float32)
        self.check(self.lib.rtg_guide_train(self.h, _ptr(params), _ptr(samples), samples.shape[0], int(steps),
                                            float(lr), _ptr(out), _ptr(losses)))
        return out, losses[:steps]

    def guide_params(self):
        """The network trained by the last guided render on this context."""
        out = np.zeros(GUIDE_PARAMS, np.float32)
        self.check(self.lib.rtg_guide_params(self.h, _ptr(out)))
        return out

    def guide_sample(self, mix, u):
        mix = _f32(mix, (-1, 40)); u = _f32(u, (-1, 3))
        n = mix.shape[0]
        d = np.zeros((n, 3), np.float32); p = np.zeros(n, np.float32)
        self.check(self.lib.rtg_guide_sample(self.h, _ptr(mix), _ptr(u), n, _ptr(d), _ptr(p)))
        return d, p

    def bsdf_sample(self, material, wi, u):
        wi = _f32(wi, (-1, 3)); u = _f32(u, (-1, 2))
        n = wi.shape[0]
        wo = np.zeros((n, 3), np.float32); f = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32); eta = np.zeros(n, np.float32)
        m = material_struct(material)
        self.check(self.lib.rtg_bsdf_batch(self.h, C.byref(m), 1, _ptr(wi), _ptr(u), n, _ptr(wo), _ptr(f),
                                           _ptr(pdf), _ptr(eta)))
        return wo, f, pdf, eta

    def exit_mis_sample(self, material, wi, u):
        """Rough-dielectric exit proposals, u = (technique, ux, uy) -> (wo, MIS-weighted estimate,
        reported pdf, MIS factor)."""
        wi = _f32(wi, (-1, 3)); u = _f32(u, (-1, 3))
        n = wi.shape[0]
        wo = np.zeros((n, 3), np.float32); f = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32); k = np.zeros(n, np.float32)
        m = material_struct(material)
        self.check(self.lib.rtg_bsdf_batch(self.h, C.byref(m), 3, _ptr(wi), _ptr(u), n, _ptr(wo), _ptr(f),
                                           _ptr(pdf), _ptr(k)))
        return wo, f, pdf, k


class Scene:
    """A resident scene (BVH + tables in HBM)."""

    @classmethod
    def from_handle(cls, ctx, h):
        obj = cls.__new__(cls)
        obj.ctx = ctx
        obj._keep = []
        obj.h = h
        return obj

    def __init__(self, ctx, sd):
        self.ctx = ctx
        self._keep = []
        keep = self._keep
        mats = (rtg_material * len(sd.materials))(*[material_struct(m) for m in sd.materials])
        keep.append(mats)
        meshes = (rtg_mesh_desc * max(1, len(sd.meshes)))()
        for i, m in enumerate(sd.meshes):
            v = _f32(m.vertices, (-1, 3)); idx = np.ascontiguousarray(m.indices, dtype=np.int32).reshape(-1, 3)
            keep += [v, idx]
            md = meshes[i]
            md.vertices = _ptr(v); md.num_vertices = v.shape[0]
            md.indices = _ptr(idx); md.num_triangles = idx.shape[0]
            md.normals = None; md.uvs = None; md.tri_material = None; md.emission = None
            if m.normals is not None:
                n = _f32(m.normals, (-1, 3)); keep.append(n); md.normals = _ptr(n)
            if m.uvs is not None:
                uv = _f32(m.uvs, (-1, 2)); keep.append(uv); md.uvs = _ptr(uv)
            if m.tri_material is not None:
                tm = np.ascontiguousarray(m.tri_material, dtype=np.int32); keep.append(tm); md.tri_material = _ptr(tm)
            md.material = int(m.material)
            if m.emission is not None:
                e = _f32(m.emission, (3,)); keep.append(e); md.emission = _ptr(e)
        keep.append(meshes)
        sph = (rtg_sphere_desc * max(1, len(sd.spheres)))()
        for i, s in enumerate(sd.spheres):
            sph[i].to_world[:] = [float(x) for x in np.asarray(s.to_world, np.float32).reshape(16)]
            sph[i].radius = float(s.radius)
            sph[i].material = int(s.material)
        keep.append(sph)
        desc = rtg_scene_desc()
        desc.materials = C.cast(mats, C.c_void_p); desc.num_materials = len(sd.materials)
        desc.meshes = C.cast(meshes, C.c_void_p); desc.num_meshes = len(sd.meshes)
        desc.spheres = C.cast(sph, C.c_void_p); desc.num_spheres = len(sd.spheres)
        desc.env_rgb = None; desc.env_width = 0; desc.env_height = 0; desc.env_scale = 1.0
        if sd.env is not None:
            env = _f32(sd.env); keep.append(env)
            desc.env_rgb = _ptr(env); desc.env_height, desc.env_width = env.shape[0], env.shape[1]
            desc.env_scale = float(sd.env_scale)
        desc.light_strategy = int(sd.light_strategy)
        texs = getattr(sd, "textures", None) or []
        if texs:
            tarr = (rtg_texture_desc * len(texs))(); keep.append(tarr)
            for i, t in enumerate(texs):
                rgb = _f32(t["rgb"]); keep.append(rgb)
                if rgb.ndim != 3 or rgb.shape[2] != 3:
                    raise ValueError("texture rgb must be (height, width, 3)")
                tarr[i].rgb = _ptr(rgb); tarr[i].height, tarr[i].width = rgb.shape[0], rgb.shape[1]
                tarr[i].uv_scale = float(t.get("uv_scale", 1.0)); tarr[i].value_scale = float(t.get("value_scale", 1.0))
                tarr[i].invert_x = int(t.get("invert_x", 0)); tarr[i].invert_y = int(t.get("invert_y", 0))
            desc.textures = C.cast(tarr, C.c_void_p); desc.num_textures = len(texs)
        med = getattr(sd, "medium", None)
        if med is not None and med.get("homogeneous"):
            desc.medium_homogeneous = 1
            for k in range(3):
                desc.medium_sigma_a[k] = float(med["sigma_a"][k])
                desc.medium_sigma_s[k] = float(med["sigma_s"][k])
            desc.medium_g = float(med.get("g", 0.0))
            desc.medium_scale = float(med.get("scale", 1.0))
        elif med is not None:
            grid = _f32(med["density"]); keep.append(grid)
            if grid.ndim != 3:
                raise ValueError("medium density must be a (z, y, x) array")
            desc.medium_density = _ptr(grid)
            desc.medium_res[0], desc.medium_res[1], desc.medium_res[2] = grid.shape[2], grid.shape[1], grid.shape[0]
            for k in range(3):
                desc.medium_albedo[k] = float(med.get("albedo", (0.0, 0.0, 0.0))[k])
            desc.medium_g = float(med.get("g", 0.0))
            desc.medium_scale = float(med.get("scale", 1.0))
        h = P()
        ctx.check(ctx.lib.rtg_scene_upload(ctx.h, C.byref(desc), C.byref(h)))
        self.h = h
        self._keep = []  # the library copied everything

    def release(self):
        if self.h:
            self.ctx.check(self.ctx.lib.rtg_scene_release(self.ctx.h, self.h))
            self.h = None

    def material(self, index):
        """Material `index` as uploaded (rtg_scene_material) -> dict of its fields."""
        m = rtg_material()
        self.ctx.check(self.ctx.lib.rtg_scene_material(self.h, int(index), C.byref(m)))
        return {k: (list(getattr(m, k)) if k in ("albedo", "eta", "k", "transmitted") else getattr(m, k))
                for k, _ in m._fields_}

    def info(self):
        i = rtg_scene_info()
        self.ctx.check(self.ctx.lib.rtg_scene_query(self.h, C.byref(i)))
        return {k: (list(getattr(i, k)) if k.startswith("bbox") or k in ("medium_res", "medium_albedo", "medium_sigma_a",
                                                                          "medium_sigma_s") else getattr(i, k))
                for k, _ in i._fields_}

    def bytes(self):
        b = C.c_uint64()
        self.ctx.check(self.ctx.lib.rtg_scene_bytes(self.h, C.byref(b)))
        return b.value

    def trace(self, rays, any_hit=False):
        rays = _f32(rays, (-1, 8))
        n = rays.shape[0]
        prim = np.zeros(n, np.int32); t = np.zeros(n, np.float32); uv = np.zeros((n, 2), np.float32)
        p = np.zeros((n, 3), np.float32); nrm = np.zeros((n, 3), np.float32)
        self.ctx.check(self.ctx.lib.rtg_trace(self.ctx.h, self.h, _ptr(rays), n, int(any_hit), _ptr(prim), _ptr(t),
                                              _ptr(uv), _ptr(p), _ptr(nrm)))
        return prim, t, uv, p, nrm

    def medium(self, op, x, u=None):
        """Camera-medium tap (rtg_medium_batch): op 0 density at points (n, 3); op 1 delta tracking /
        op 2 ratio tracking on rays (n, 8) with draws u (n, k); op 3 phase eval on (n, 6) = (wi, wo);
        op 4 phase sample of wi (n, 3) with u (n, 2)."""
        w_in = {0: 3, 1: 8, 2: 8, 3: 6, 4: 3}[op]
        w_out = {0: 1, 1: 3, 2: 3, 3: 1, 4: 4}[op]
        x = _f32(x, (-1, w_in)); n = x.shape[0]
        k = 0
        if u is not None:
            u = _f32(u).reshape(n, -1); k = u.shape[1]
        out = np.zeros((n, w_out), np.float32)
        self.ctx.check(self.ctx.lib.rtg_medium_batch(self.ctx.h, self.h, int(op), _ptr(x), n,
                                                     None if u is None else _ptr(u), k, _ptr(out)))
        return out[:, 0] if w_out == 1 else out

    def light(self, op, x, light=0, u=None):
        """Light tap (rtg_light_batch): op 0 sample light `light` from points (n, 3) with draws u (n, 2)
        -> (n, 14); op 1 its pdf for (n, 12) = (from, light point, light normal, direction) -> (n,);
        op 2 environment radiance of directions (n, 3) -> (n, 3); op 3 light choice for draws (n,)
        -> (n, 3) = (index, rescaled draw, probability)."""
        w_in = {0: 3, 1: 12, 2: 3, 3: 1}[op]
        w_out = {0: 14, 1: 1, 2: 3, 3: 3}[op]
        x = _f32(x, (-1, w_in)); n = x.shape[0]
        if u is not None:
            u = _f32(u, (n, 2))
        out = np.zeros((n, w_out), np.float32)
        self.ctx.check(self.ctx.lib.rtg_light_batch(self.ctx.h, self.h, int(op), int(light), _ptr(x), n,
                                                    None if u is None else _ptr(u), _ptr(out)))
        return out[:, 0] if w_out == 1 else out

    def render(self, camera, params):
        d = render_struct(camera, params)
        W, H = d.width, d.height
        out = np.zeros((H, W, 4), np.float32)
        st = rtg_stats()
        self.ctx.check(self.ctx.lib.rtg_render(self.ctx.h, self.h, C.byref(d), _ptr(out), C.byref(st)))
        return out, st

    def render_device(self, camera, params, out_ptr):
        d = render_struct(camera, params)
        st = rtg_stats()
        self.ctx.check(self.ctx.lib.rtg_render_device(self.ctx.h, self.h, C.byref(d), C.c_void_p(out_ptr),
                                                      C.byref(st)))
        return st


class Server:
    """Continuous-batching serving loop on one context (rtg_server_* in include/rtg.h).

    submit/poll/cancel may be called from other threads while step() runs: ctypes releases the
    GIL during the call and the library guards its request table separately from device work."""

    def __init__(self, ctx, hbm_budget_bytes=0):
        self.ctx = ctx
        self.lib = ctx.lib
        h = P()
        ctx.check(self.lib.rtg_server_create(ctx.h, C.c_uint64(int(hbm_budget_bytes)), C.byref(h)))
        self.h = h
        self.shapes = {}
        ctx._servers.append(self)

    def close(self):
        if self.h:
            self.lib.rtg_server_destroy(self.h)
            self.h = None
        if self in self.ctx._servers:
            self.ctx._servers.remove(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, scene, camera, params):
        d = render_struct(camera, params)
        jid = C.c_int64()
        self.ctx.check(self.lib.rtg_submit(self.h, scene.h, C.byref(d), C.byref(jid)))
        self.shapes[jid.value] = (d.height, d.width)
        return jid.value

    def step(self, max_ms=0.0):
        self.ctx.check(self.lib.rtg_server_step(self.h, C.c_double(float(max_ms))))

    def poll(self, jid, want_image=True):
        """-> (status, image or None, latency_ms); status is "done", "running" or "cancelled"."""
        H, W = self.shapes[jid]
        out = np.zeros((H, W, 4), np.float32) if want_image else None
        lat = C.c_double(0.0)
        rc = self.lib.rtg_poll(self.h, C.c_int64(jid), _ptr(out), C.byref(lat))
        if rc == -5:
            return "running", None, None
        if rc == -6:
            return "cancelled", None, lat.value
        self.ctx.check(rc)
        return "done", out, lat.value

    def start(self):
        """Run the loop on the server's own thread (rtg_server_start); use wait() afterwards."""
        self.ctx.check(self.lib.rtg_server_start(self.h))

    def wait(self, jid, timeout_ms=0.0, want_image=True):
        """Block until the request ends (rtg_wait) -> (status, image or None, latency_ms); status
        "running" only when timeout_ms > 0 passed first."""
        H, W = self.shapes[jid]
        out = np.zeros((H, W, 4), np.float32) if want_image else None
        lat = C.c_double(0.0)
        rc = self.lib.rtg_wait(self.h, C.c_int64(jid), C.c_double(float(timeout_ms)), _ptr(out), C.byref(lat))
        if rc == -5:
            return "running", None, None
        if rc == -6:
            return "cancelled", None, lat.value
        self.ctx.check(rc)
        return "done", out, lat.value

    def cancel(self, jid):
        self.ctx.check(self.lib.rtg_cancel(self.h, C.c_int64(jid)))

    def stats(self):
        st = rtg_server_stats()
        self.ctx.check(self.lib.rtg_server_query(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_ if k != "reserved"}


def spectrum_to_rgb(lam, values):
    """Sampled spectrum (wavelengths in nm, values) -> linear sRGB (rtg_spectrum_to_rgb), the
    conversion the scene loader applies to a named metal's IOR spectra (core/Color.cpp:423-443)."""
    lib = load()
    la, va = _f32(lam).ravel(), _f32(values).ravel()
    if la.size != va.size:
        raise ValueError("spectrum_to_rgb: wavelengths and values differ in length")
    out = np.zeros(3, np.float32)
    code = lib.rtg_spectrum_to_rgb(_ptr(la), _ptr(va), int(la.size), _ptr(out))
    if code != 0:
        raise RtgError(code, "rtg_spectrum_to_rgb needs at least two samples")
    return out


def resolve(rgbw):
    """Divide accumulated radiance by the filter weight (core/Screen.cpp:113-122)."""
    w = rgbw[..., 3:4]
    return np.where(w > 0, rgbw[..., :3] / np.maximum(w, 1e-30), 0.0).astype(np.float32)
