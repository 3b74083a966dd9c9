"""Host-side scene descriptions and the procedural benchmark scenes.

A SceneDesc is what the reference builds from its XML + OBJ inputs: triangle
meshes (geometricShape/Mesh.h), spheres (geometricShape/Sphere.h), constant-texture
materials (Diffuse / RoughConductor / RoughDielectric), mesh emitters
(light/MeshLight.cpp) and a lat-long environment map (light/EnvironmentLight.cpp).
Light order for the selection CDF: emissive meshes in mesh order, then the
environment light.

Procedural scenes:
  cornell()        Cornell-box, 256x256, 16 spp stratified, depth 8, RR after 3, seed 1.
  displaced_sphere() level-8 icosphere (1,310,720 tris) displaced by 4-octave Perlin noise,
                   rough GGX conductor patches, Lambertian ground, sky env map, 1024^2, 64 spp,
                   depth 6.
  glossy_cornell() Cornell variant: rough dielectric spheres, rough conductor walls, a small
                   light hidden behind a baffle; 512^2.
  vol_cornell()    the Cornell box over [0,1]^3 filled by a 256^3 fBm camera medium (albedo 0.9,
                   Henyey-Greenstein g 0.3), volumetric integrator, 512^2, 128 spp, depth 16.
"""
import functools
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import rtg as R

# RoughConductor.cpp:27-28 (gold, as given in the reference's comment)
GOLD_ETA = (0.22273, 0.42791, 1.2597)
GOLD_K = (3.0325, 2.3345, 1.7531)


@dataclass
class MeshDesc:
    vertices: np.ndarray
    indices: np.ndarray
    material: int = 0
    normals: Optional[np.ndarray] = None
    uvs: Optional[np.ndarray] = None
    tri_material: Optional[np.ndarray] = None
    emission: Optional[tuple] = None


@dataclass
class SphereDesc:
    to_world: np.ndarray
    radius: float
    material: int


@dataclass
class SceneDesc:
    materials: List[dict] = field(default_factory=list)
    meshes: List[MeshDesc] = field(default_factory=list)
    spheres: List[SphereDesc] = field(default_factory=list)
    env: Optional[np.ndarray] = None  # H x W x 3
    env_scale: float = 1.0
    light_strategy: int = R.LIGHTS_UNIFORM
    camera: dict = field(default_factory=dict)
    render: dict = field(default_factory=dict)
    # camera medium over the world box [0,1]^3: density (z, y, x) float32, albedo (3), g, scale
    medium: Optional[dict] = None
    # image textures: dicts of rgb (H x W x 3 linear), uv_scale, invert_x, invert_y, value_scale;
    # materials refer to them by 1-based index (tex_albedo, tex_alpha, tex_transmitted)
    textures: List[dict] = field(default_factory=list)

    def add_material(self, **m):
        self.materials.append(m)
        return len(self.materials) - 1

    def num_triangles(self):
        return int(sum(m.indices.shape[0] for m in self.meshes))


def diffuse(albedo):
    return dict(type=R.MAT_DIFFUSE, albedo=tuple(albedo))


def conductor(alpha, eta=GOLD_ETA, k=GOLD_K, dist=R.DIST_GGX, reflectance=(1.0, 1.0, 1.0)):
    return dict(type=R.MAT_ROUGH_CONDUCTOR, alpha=alpha, eta=tuple(eta), k=tuple(k), distribution=dist,
                albedo=tuple(reflectance))


def dielectric(alpha, eta_int=1.5, eta_ext=1.000277, dist=R.DIST_GGX):
    return dict(type=R.MAT_ROUGH_DIELECTRIC, alpha=alpha, eta_int=eta_int, eta_ext=eta_ext, distribution=dist,
                albedo=(1.0, 1.0, 1.0), transmitted=(1.0, 1.0, 1.0))


def smooth_conductor(eta=GOLD_ETA, k=GOLD_K, reflectance=(1.0, 1.0, 1.0)):
    """Perfectly specular conductor (bsdf/Conductor.cpp)."""
    return dict(type=R.MAT_CONDUCTOR, eta=tuple(eta), k=tuple(k), albedo=tuple(reflectance))


def smooth_dielectric(eta_int=1.5046, eta_ext=1.000277, reflectance=(1.0, 1.0, 1.0), transmitted=(1.0, 1.0, 1.0)):
    """Perfectly specular dielectric (bsdf/Dielectric.cpp)."""
    return dict(type=R.MAT_DIELECTRIC, eta_int=eta_int, eta_ext=eta_ext, albedo=tuple(reflectance),
                transmitted=tuple(transmitted))


def quad(corner, e1, e2, normal):
    """Two triangles spanning corner + [0,1]e1 + [0,1]e2, wound so (p1-p0)x(p2-p0) follows `normal`."""
    c = np.asarray(corner, np.float64); a = np.asarray(e1, np.float64); b = np.asarray(e2, np.float64)
    v = np.array([c, c + a, c + a + b, c + b], np.float64)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    if np.dot(np.cross(a, b), np.asarray(normal, np.float64)) < 0:
        idx = idx[:, ::-1].copy()
    return v.astype(np.float32), idx


def box_mesh(center, size, angle_deg):
    """Closed box (12 triangles, outward normals) rotated about +y (geometricShape/Cube.cpp:18-66)."""
    sx, sy, sz = [s * 0.5 for s in size]
    a = np.radians(angle_deg)
    Rm = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
    corners = np.array([[x, y, z] for x in (-sx, sx) for y in (-sy, sy) for z in (-sz, sz)])
    faces = [  # (indices into corners of a face in CCW order seen from outside)
        (0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    verts = (corners @ Rm.T) + np.asarray(center)
    idx = []
    for f in faces:
        p = verts[list(f)]
        n = np.cross(p[1] - p[0], p[2] - p[0])
        outward = np.dot(n, p.mean(0) - np.asarray(center)) > 0
        tri = [(f[0], f[1], f[2]), (f[0], f[2], f[3])]
        if not outward:
            tri = [(t[0], t[2], t[1]) for t in tri]
        idx += tri
    return verts.astype(np.float32), np.array(idx, np.int32)


def _cornell_walls(sd, white, red, green, left_m=None, right_m=None, back_m=None):
    L = -0.5
    walls = [
        (quad((L, 0, L), (1, 0, 0), (0, 0, 1), (0, 1, 0)), white),            # floor
        (quad((L, 1, L), (1, 0, 0), (0, 0, 1), (0, -1, 0)), white),           # ceiling
        (quad((L, 0, L), (1, 0, 0), (0, 1, 0), (0, 0, 1)), back_m if back_m is not None else white),  # back
        (quad((L, 0, L), (0, 1, 0), (0, 0, 1), (1, 0, 0)), left_m if left_m is not None else red),    # left
        (quad((0.5, 0, L), (0, 1, 0), (0, 0, 1), (-1, 0, 0)), right_m if right_m is not None else green),  # right
    ]
    v = []; idx = []; tm = []; off = 0
    for (vv, ii), m in walls:
        v.append(vv); idx.append(ii + off); tm += [m] * len(ii); off += len(vv)
    sd.meshes.append(MeshDesc(np.concatenate(v), np.concatenate(idx), material=white,
                              tri_material=np.array(tm, np.int32)))


def cornell(width=256, height=256, spp=16, max_depth=8, min_depth=3, seed=1):
    """Cornell box: 5 Lambertian quads (10 tris), two boxes rotated +-18 deg (24 tris), a 0.25x0.25
    ceiling light of radiance (17,12,4) facing down, NoriPerspective fov 40."""
    sd = SceneDesc()
    white = sd.add_material(**diffuse((0.73, 0.73, 0.73)))
    red = sd.add_material(**diffuse((0.65, 0.05, 0.05)))
    green = sd.add_material(**diffuse((0.12, 0.45, 0.15)))
    black = sd.add_material(**diffuse((0.0, 0.0, 0.0)))
    _cornell_walls(sd, white, red, green)
    v, i = box_mesh((0.17, 0.15, 0.12), (0.3, 0.3, 0.3), -18.0)
    sd.meshes.append(MeshDesc(v, i, material=white))
    v, i = box_mesh((-0.16, 0.3, -0.15), (0.3, 0.6, 0.3), 18.0)
    sd.meshes.append(MeshDesc(v, i, material=white))
    lv, li = quad((-0.125, 0.999, -0.125), (0.25, 0, 0), (0, 0, 0.25), (0, -1, 0))
    sd.meshes.append(MeshDesc(lv, li, material=black, emission=(17.0, 12.0, 4.0)))
    sd.camera = dict(origin=(0.0, 0.5, 1.95), target=(0.0, 0.5, 0.0), up=(0.0, 1.0, 0.0), fov=40.0,
                     width=width, height=height, near=1e-4, far=1e4)
    sd.render = dict(spp=spp, sampler=R.SAMPLER_STRATIFIED, dimension=4, min_depth=min_depth, max_depth=max_depth,
                     seed=seed)
    return sd


# ---------------------------------------------------------------- Perlin noise (seeded)
class Pcg32:
    """PCG32 (tools/Rng.h:68-141) for procedural parameters."""
    MULT = 0x5851f42d4c957f2d
    MASK = (1 << 64) - 1

    def __init__(self, initstate, initseq=1):
        self.state = 0
        self.inc = ((initseq << 1) | 1) & self.MASK
        self.next_uint()
        self.state = (self.state + initstate) & self.MASK
        self.next_uint()

    def next_uint(self):
        old = self.state
        self.state = (old * self.MULT + self.inc) & self.MASK
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    def next_bounded(self, bound):
        threshold = ((~bound + 1) & 0xFFFFFFFF) % bound
        while True:
            r = self.next_uint()
            if r >= threshold:
                return r % bound

    def next_real(self):
        u = (self.next_uint() >> 9) | 0x3F800000
        return float(np.frombuffer(np.uint32(u).tobytes(), np.float32)[0]) - 1.0


def perlin_perm(seed):
    rng = Pcg32(seed)
    p = list(range(256))
    for i in range(255, 0, -1):  # Fisher-Yates
        j = rng.next_bounded(i + 1)
        p[i], p[j] = p[j], p[i]
    return np.array(p + p, np.int64)


def perlin3(x, y, z, perm):
    """Improved Perlin noise (Perlin 2002) on arrays; the permutation is seeded instead of fixed."""
    xi = np.floor(x).astype(np.int64) & 255; yi = np.floor(y).astype(np.int64) & 255
    zi = np.floor(z).astype(np.int64) & 255
    xf = x - np.floor(x); yf = y - np.floor(y); zf = z - np.floor(z)
    fade = lambda t: t * t * t * (t * (t * 6 - 15) + 10)
    u, v, w = fade(xf), fade(yf), fade(zf)

    def grad(h, x, y, z):
        h = h & 15
        a = np.where(h < 8, x, y)
        b = np.where(h < 4, y, np.where((h == 12) | (h == 14), x, z))
        return np.where(h & 1, -a, a) + np.where(h & 2, -b, b)

    A = perm[xi] + yi; AA = perm[A] + zi; AB = perm[A + 1] + zi
    B = perm[xi + 1] + yi; BA = perm[B] + zi; BB = perm[B + 1] + zi
    lerp = lambda t, a, b: a + t * (b - a)
    return lerp(w,
                lerp(v, lerp(u, grad(perm[AA], xf, yf, zf), grad(perm[BA], xf - 1, yf, zf)),
                     lerp(u, grad(perm[AB], xf, yf - 1, zf), grad(perm[BB], xf - 1, yf - 1, zf))),
                lerp(v, lerp(u, grad(perm[AA + 1], xf, yf, zf - 1), grad(perm[BA + 1], xf - 1, yf, zf - 1)),
                     lerp(u, grad(perm[AB + 1], xf, yf - 1, zf - 1), grad(perm[BB + 1], xf - 1, yf - 1, zf - 1))))


def icosphere(level):
    """Unit icosphere: 20 * 4^level triangles; children of face f are faces 4f..4f+3."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (len(v) + 1) + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a = uk // (len(v) + 1); b = uk % (len(v) + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = inv.reshape(3, -1) + len(v)  # m[0]=mid(0,1), m[1]=mid(1,2), m[2]=mid(2,0)
        v = np.concatenate([v, mid])
        n = len(f)
        nf = np.empty((n, 4, 3), np.int64)
        nf[:, 0] = np.stack([f[:, 0], m[0], m[2]], 1)
        nf[:, 1] = np.stack([m[0], f[:, 1], m[1]], 1)
        nf[:, 2] = np.stack([m[2], m[1], f[:, 2]], 1)
        nf[:, 3] = np.stack([m[0], m[1], m[2]], 1)
        f = nf.reshape(-1, 3)
    return v, f


def sky_env(width=512, height=256):
    """Procedural sky gradient, lat-long (row 0 = +y): zenith blue to bright horizon, dim ground,
    plus a soft warm sun lobe so the env-light sampler has structure to follow."""
    th = (np.arange(height) + 0.5) / height * np.pi
    ph = (np.arange(width) + 0.5) / width * 2 * np.pi
    cy = np.cos(th)[:, None, None]                       # H x 1 x 1
    zen = np.array([0.25, 0.45, 1.0]); hor = np.array([1.0, 0.95, 0.85]); gnd = np.array([0.25, 0.22, 0.2])
    up = np.sqrt(np.clip(cy, 0, 1))
    sky = hor * (1 - up) + zen * up                      # H x 1 x 3
    ground = gnd * (0.5 + 0.5 * (1 + cy))
    col = np.where(cy > 0, sky, ground) * 1.5            # H x 1 x 3
    env = np.repeat(col, width, axis=1)                  # H x W x 3
    d = np.stack([np.sin(ph)[None, :] * np.sin(th)[:, None], np.cos(th)[:, None] * np.ones((1, width)),
                  -np.cos(ph)[None, :] * np.sin(th)[:, None]], -1)
    sun = np.array([0.4, 0.8, 0.45]); sun /= np.linalg.norm(sun)
    lobe = np.exp(40.0 * (d @ sun - 1.0))[..., None] * np.array([40.0, 36.0, 30.0])
    return np.ascontiguousarray((env + lobe).astype(np.float32))


def displaced_sphere(width=1024, height=1024, spp=64, max_depth=6, min_depth=3, level=8, seed=2, patch=1024):
    """Large-mesh BVH stress scene: icosphere level 8 (1,310,720 triangles) radially displaced by
    4-octave Perlin noise (amplitude 0.05, seeded permutation), GGX conductor with alpha ~ U[0.05, 0.4]
    per `patch` consecutive triangles, Lambertian ground plane, sky env map 512x256."""
    sd = SceneDesc()
    ground = sd.add_material(**diffuse((0.5, 0.5, 0.5)))  # Diffuse.cpp:17 default albedo
    v, f = icosphere(level)
    perm = perlin_perm(seed)
    p = v * 4.0
    disp = np.zeros(len(v))
    amp = 1.0; freq = 1.0
    for _ in range(4):
        disp += amp * perlin3(p[:, 0] * freq, p[:, 1] * freq, p[:, 2] * freq, perm)
        amp *= 0.5; freq *= 2.0
    center = np.array([0.0, 1.1, 0.0])
    verts = (v * (1.0 + 0.05 * disp)[:, None] + center).astype(np.float32)
    ntri = len(f)
    npatch = (ntri + patch - 1) // patch
    rng = Pcg32(seed)
    base = len(sd.materials)
    for _ in range(npatch):
        a = 0.05 + 0.35 * rng.next_real()
        sd.add_material(**conductor(a))
    tri_mat = (base + np.arange(ntri) // patch).astype(np.int32)
    sd.meshes.append(MeshDesc(verts, f.astype(np.int32), material=base, tri_material=tri_mat))
    gv, gi = quad((-10, 0, -10), (20, 0, 0), (0, 0, 20), (0, 1, 0))
    sd.meshes.append(MeshDesc(gv, gi, material=ground))
    sd.env = sky_env(512, 256)
    sd.camera = dict(origin=(0.0, 1.9, 4.3), target=(0.0, 1.05, 0.0), up=(0.0, 1.0, 0.0), fov=40.0,
                     width=width, height=height, near=1e-4, far=1e4)
    sd.render = dict(spp=spp, sampler=R.SAMPLER_STRATIFIED, dimension=4, min_depth=min_depth, max_depth=max_depth,
                     seed=seed)
    return sd


@functools.lru_cache(maxsize=2)
def fbm_density(res=256, octaves=5, seed=3, freq=4.0, gain=6.0, bias=1.0, lo=0.0, hi=4.0):
    """Density grid for the volumetric scene: clip(bias + gain * fBm(x), lo, hi) at the grid
    vertices x = (i, j, k) / res, fBm = sum over octaves of 0.5^o Perlin(freq 2^o x) with a seeded
    permutation. Returns a (z, y, x) float32 array (read-only, cached). One z-slice per task on a
    thread pool."""
    from concurrent.futures import ThreadPoolExecutor

    perm = perlin_perm(seed)
    c = np.arange(res, dtype=np.float64) / res
    y, x = np.meshgrid(c, c, indexing="ij")
    out = np.empty((res, res, res), np.float32)

    def slab(k):
        z = np.full_like(x, c[k])
        acc = np.zeros_like(x)
        amp, f = 1.0, freq
        for _ in range(octaves):
            acc += amp * perlin3(x * f, y * f, z * f, perm)
            amp *= 0.5; f *= 2.0
        out[k] = np.clip(bias + gain * acc, lo, hi)

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        list(ex.map(slab, range(res)))
    out.setflags(write=False)  # cached: shared by every caller
    return out


def vol_cornell(width=512, height=512, spp=128, max_depth=16, min_depth=3, seed=1, res=256, albedo=0.9, g=0.3):
    """Heterogeneous volumetric scene: the Cornell box moved to span [0,1]^3 (x and z shifted by
    0.5), filled by a camera medium over that box (a camera medium's grid spans the world box
    [0,1]^3): res^3 fBm density (5 octaves, seed 3, clipped to [0,4]), single-scattering albedo 0.9,
    Henyey-Greenstein g = 0.3; rendered by the volumetric integrator."""
    sd = cornell(width, height, spp, max_depth, min_depth, seed)
    off = np.array([0.5, 0.0, 0.5], np.float32)
    for m in sd.meshes:
        m.vertices = (np.asarray(m.vertices, np.float32) + off).astype(np.float32)
    cam = dict(sd.camera)
    cam["origin"] = tuple(float(a + b) for a, b in zip(cam["origin"], off))
    cam["target"] = tuple(float(a + b) for a, b in zip(cam["target"], off))
    sd.camera = cam
    sd.medium = dict(density=fbm_density(res), albedo=(albedo, albedo, albedo), g=g, scale=1.0)
    sd.render = dict(sd.render, integrator=R.INTEGRATOR_VOLPATH)
    return sd


def two_light_cornell(width=256, height=256, spp=64, max_depth=8, min_depth=3, seed=9, strategy=R.LIGHTS_ALL):
    """Cornell box with a second emitter (a small blue panel on the left wall facing +x) and a dim
    sky seen through the open front: three lights for the light-selection strategies."""
    sd = cornell(width, height, spp, max_depth, min_depth, seed)
    black = sd.add_material(**diffuse((0.0, 0.0, 0.0)))
    v, i = quad((-0.499, 0.55, -0.2), (0, 0.2, 0), (0, 0, 0.2), (1, 0, 0))
    sd.meshes.append(MeshDesc(v, i, material=black, emission=(2.0, 4.0, 8.0)))
    sd.env = np.full((8, 16, 3), 0.05, np.float32)
    sd.light_strategy = strategy
    return sd


def checker_texture(n=16, cells=4, lo=(0.1, 0.1, 0.1), hi=(0.9, 0.8, 0.6), smooth=False):
    """An n x n RGB texture: a checkerboard of `cells` x `cells` squares, or (smooth) a linear ramp
    lo + (hi - lo) * (i + j) / (2 (n - 1)) over texel column i and row j."""
    lo, hi = np.array(lo, np.float32), np.array(hi, np.float32)
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    if smooth:
        t = ((i + j) / (2.0 * (n - 1)))[..., None]
    else:
        t = (((i * cells // n) + (j * cells // n)) % 2)[..., None].astype(np.float32)
    return np.ascontiguousarray((lo + (hi - lo) * t).astype(np.float32))


def textured_cornell(width=256, height=256, spp=64, max_depth=8, min_depth=3, seed=7):
    """Cornell box with image textures: a checkerboard albedo on the back wall (uvs over the quad),
    a smooth ramp albedo on the short box (barycentric uvs), a rough-conductor sphere whose
    roughness comes from a texture (spherical uvs), and a ScaleTexture-style value scale."""
    sd = cornell(width, height, spp, max_depth, min_depth, seed)
    sd.textures.append(dict(rgb=checker_texture(32, 8), uv_scale=1.0))
    sd.textures.append(dict(rgb=checker_texture(16, smooth=True), uv_scale=2.0, value_scale=0.9))
    sd.textures.append(dict(rgb=checker_texture(8, 2, lo=(0.15, 0.15, 0.15), hi=(0.5, 0.5, 0.5)), invert_y=1))
    back = sd.add_material(**dict(diffuse((0.5, 0.5, 0.5)), tex_albedo=1))
    ramp = sd.add_material(**dict(diffuse((0.5, 0.5, 0.5)), tex_albedo=2))
    rough = sd.add_material(**dict(conductor(0.2), tex_alpha=3))
    # the walls (floor, ceiling, back, left, right; 4 vertices and 2 triangles each) get per-quad
    # uvs; the back wall's triangles the checkerboard material
    walls = sd.meshes[0]
    walls.uvs = np.tile(np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32), (5, 1))
    tm = np.array(walls.tri_material, np.int32)
    tm[4:6] = back
    walls.tri_material = tm
    sd.meshes[1].material = ramp  # short box: no uvs, its barycentrics are the uvs
    sd.spheres.append(SphereDesc(np.array([[1, 0, 0, -0.2], [0, 1, 0, 0.75], [0, 0, 1, 0.1], [0, 0, 0, 1]], np.float32),
                                 0.12, rough))
    return sd


def glossy_cornell(width=512, height=512, spp=64, max_depth=1000, min_depth=3, seed=5):
    """Glossy-indirect scene: rough dielectric spheres (IOR 1.5, alpha 0.1), rough GGX conductor walls
    (alpha 0.15), a white diffuse floor and ceiling, and a 0.05x0.05 light 0.1 below the ceiling that
    faces up, hidden from the room by a 0.2x0.2 baffle just beneath it: the room is lit by the bright
    patch of ceiling above the light (illumination is indirect; next-event estimation from the room
    cannot reach the light)."""
    sd = SceneDesc()
    white = sd.add_material(**diffuse((0.73, 0.73, 0.73)))
    wall_l = sd.add_material(**conductor(0.15, eta=(0.2, 0.92, 1.1), k=(3.9, 2.45, 2.14)))   # copper-like
    wall_r = sd.add_material(**conductor(0.15, eta=(1.66, 0.88, 0.52), k=(9.2, 6.3, 4.8)))   # aluminium-like
    wall_b = sd.add_material(**conductor(0.15))                                              # gold
    glass = sd.add_material(**dielectric(0.1, eta_int=1.5, eta_ext=1.000277))
    black = sd.add_material(**diffuse((0.0, 0.0, 0.0)))
    _cornell_walls(sd, white, wall_l, wall_r, left_m=wall_l, right_m=wall_r, back_m=wall_b)
    lv, li = quad((-0.025, 0.9, -0.025), (0.05, 0, 0), (0, 0, 0.05), (0, 1, 0))  # light, facing the ceiling
    sd.meshes.append(MeshDesc(lv, li, material=black, emission=(425.0, 300.0, 100.0)))
    bv, bi = quad((-0.1, 0.895, -0.1), (0.2, 0, 0), (0, 0, 0.2), (0, -1, 0))  # baffle under the light
    sd.meshes.append(MeshDesc(bv, bi, material=white))
    for c, r in (((-0.2, 0.18, -0.1), 0.18), ((0.22, 0.15, 0.15), 0.15)):
        M = np.eye(4, dtype=np.float32); M[:3, 3] = c
        sd.spheres.append(SphereDesc(M, r, glass))
    sd.camera = dict(origin=(0.0, 0.5, 1.95), target=(0.0, 0.5, 0.0), up=(0.0, 1.0, 0.0), fov=40.0,
                     width=width, height=height, near=1e-4, far=1e4)
    sd.render = dict(spp=spp, sampler=R.SAMPLER_STRATIFIED, dimension=4, min_depth=min_depth, max_depth=max_depth,
                     seed=seed)
    return sd


def specular_cornell(width=256, height=256, spp=64, max_depth=16, min_depth=3, seed=6):
    """Cornell box with perfectly specular surfaces: the tall box a smooth gold mirror, a smooth
    glass sphere (IOR 1.5) in front of the short box, the ceiling light of cornell()."""
    sd = cornell(width=width, height=height, spp=spp, max_depth=max_depth, min_depth=min_depth, seed=seed)
    mirror = sd.add_material(**smooth_conductor())
    glass = sd.add_material(**smooth_dielectric(eta_int=1.5))
    sd.meshes[2].material = mirror   # the tall box
    M = np.eye(4, dtype=np.float32); M[:3, 3] = (0.2, 0.43, 0.15)
    sd.spheres.append(SphereDesc(M, 0.12, glass))
    return sd


def indirect_cornell(width=512, height=512, spp=64, max_depth=1000, min_depth=3, seed=5):
    """The glossy-indirect scene's light and baffle in an all-Lambertian Cornell box (white floor,
    ceiling and back wall, red and green side walls, no spheres): the same indirect-only lighting
    from a small bright patch of ceiling, without the rough dielectric and conductor paths whose
    rare large sample weights dominate any error measure on the glossy scene."""
    sd = SceneDesc()
    white = sd.add_material(**diffuse((0.73, 0.73, 0.73)))
    red = sd.add_material(**diffuse((0.65, 0.05, 0.05)))
    green = sd.add_material(**diffuse((0.12, 0.45, 0.15)))
    black = sd.add_material(**diffuse((0.0, 0.0, 0.0)))
    _cornell_walls(sd, white, red, green)
    lv, li = quad((-0.025, 0.9, -0.025), (0.05, 0, 0), (0, 0, 0.05), (0, 1, 0))
    sd.meshes.append(MeshDesc(lv, li, material=black, emission=(425.0, 300.0, 100.0)))
    bv, bi = quad((-0.1, 0.895, -0.1), (0.2, 0, 0), (0, 0, 0.2), (0, -1, 0))
    sd.meshes.append(MeshDesc(bv, bi, material=white))
    sd.camera = dict(origin=(0.0, 0.5, 1.95), target=(0.0, 0.5, 0.0), up=(0.0, 1.0, 0.0), fov=40.0,
                     width=width, height=height, near=1e-4, far=1e4)
    sd.render = dict(spp=spp, sampler=R.SAMPLER_STRATIFIED, dimension=4, min_depth=min_depth, max_depth=max_depth,
                     seed=seed)
    return sd


# ---------------------------------------------------------------- scene files
def write_rgbe(path, rgb):
    """Radiance RGBE (.hdr) image from an (H, W, 3) float array (flat scanlines, -Y H +X W)."""
    rgb = np.asarray(rgb, np.float64)
    H, W = rgb.shape[:2]
    m = rgb.max(axis=2)
    e = np.zeros((H, W), np.int32)
    mant = np.zeros((H, W), np.float64)
    nz = m > 1e-32
    mant[nz], e[nz] = np.frexp(m[nz])
    px = np.zeros((H, W, 4), np.uint8)
    scale = np.where(nz, mant * 256.0 / np.where(nz, m, 1.0), 0.0)
    px[..., :3] = np.clip(np.floor(rgb * scale[..., None]), 0, 255).astype(np.uint8)
    px[..., 3] = np.where(nz, e + 128, 0).astype(np.uint8)
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n")
        f.write(f"-Y {H} +X {W}\n".encode())
        f.write(px.tobytes())


def _fmt(v):
    return " ".join(repr(float(x)) for x in v)


def write_vol(path, density, for_reference_reader=True):
    """Mitsuba VOL v3 grid file (float32, one channel, bounding box [0,1]^3) of a (z, y, x)
    density array. The reference's GridVolume reader skips 4 bytes after the 48-byte header
    (volume/GridVolume.cpp:56-60); with for_reference_reader one padding value is written there
    so that it reads the grid as given."""
    d = np.ascontiguousarray(density, np.float32)
    nz, ny, nx = d.shape
    with open(path, "wb") as f:
        f.write(b"VOL" + bytes([3]))
        f.write(np.array([1, nx, ny, nz, 1], np.int32).tobytes())
        f.write(np.array([0, 0, 0, 1, 1, 1], np.float32).tobytes())
        if for_reference_reader:
            f.write(np.zeros(1, np.float32).tobytes())
        f.write(d.tobytes())


def _medium_xml(med, directory):
    import os

    write_vol(os.path.join(directory, "medium.vol"), med["density"])
    return ['  <medium type="Heterogeneous">',
            '    <volume name="volume" type="GridVolume"><string name="path" value="medium.vol"/></volume>',
            f'    <color name="albedo" value="{_fmt(med.get("albedo", (0.0, 0.0, 0.0)))}"/>',
            f'    <double name="scale" value="{float(med.get("scale", 1.0))!r}"/>',
            '    <phase name="phaseFunction" type="HGPhaseFunction">',
            f'      <double name="g" value="{float(med.get("g", 0.0))!r}"/>', "    </phase>", "  </medium>"]


def _filter_xml(f):
    if not f:
        return ['    <filter type="BoxFilter"/>']
    kind = {R.FILTER_BOX: "BoxFilter", R.FILTER_GAUSSIAN: "GaussianFilter", R.FILTER_MITCHELL: "MitchellNetravaliFilter"}
    rx, ry = f.get("radius", (0.5, 0.5))
    out = [f'    <filter type="{kind[f["type"]]}">', f'      <double name="radiusX" value="{float(rx)!r}"/>',
           f'      <double name="radiusY" value="{float(ry)!r}"/>']
    p0, p1 = f.get("param", (0.0, 0.0))
    if f["type"] == R.FILTER_GAUSSIAN:
        out += [f'      <double name="alpha" value="{float(p0)!r}"/>', f'      <double name="omega" value="{float(p1)!r}"/>']
    if f["type"] == R.FILTER_MITCHELL:
        out += [f'      <double name="B" value="{float(p0)!r}"/>', f'      <double name="C" value="{float(p1)!r}"/>']
    return out + ["    </filter>"]


def write_xml(sd, directory, name="scene.xml"):
    """Write a SceneDesc as a scene file in the reference's XML schema (one OBJ per mesh and
    material, a MeshLight on emitting meshes, Sphere shapes with a toWorld matrix, an RGBE
    environment map). Returns the path of the XML file."""
    import os

    os.makedirs(directory, exist_ok=True)
    cam, rp = sd.camera, sd.render
    out = ['<?xml version="1.0" encoding="utf-8"?>', "<scene>"]
    strat = {R.LIGHTS_DEFAULT: "one_light_uniform", R.LIGHTS_UNIFORM: "one_light_uniform",
             R.LIGHTS_WEIGHTED: "one_light_weighted", R.LIGHTS_ALL: "all_lights"}[sd.light_strategy]
    integ = "VolPathTracing" if rp.get("integrator", R.INTEGRATOR_PATH_MIS) == R.INTEGRATOR_VOLPATH else "PathTracingMIS"
    out += [f'  <integrator type="{integ}">',
            f'    <integer name="minDepth" value="{int(rp.get("min_depth", 3))}"/>',
            f'    <integer name="maxDepth" value="{int(rp.get("max_depth", 1000))}"/>',
            f'    <string name="strategy" value="{strat}"/>', "  </integrator>"]
    stype = "StratifiedSampler" if rp.get("sampler", R.SAMPLER_STRATIFIED) == R.SAMPLER_STRATIFIED else "RandomSampler"
    out += [f'  <sampler type="{stype}">', f'    <integer name="sampleNumber" value="{int(rp["spp"])}"/>',
            f'    <integer name="dimension" value="{int(rp.get("dimension", 4))}"/>', "  </sampler>"]
    out += ['  <camera type="NoriPerspective">', '    <transform name="toWorld">', "      <lookAt>",
            f"        <origin>{_fmt(cam['origin'])}</origin>", f"        <lookAt>{_fmt(cam['target'])}</lookAt>",
            f"        <up>{_fmt(cam['up'])}</up>", "      </lookAt>", "    </transform>",
            f'    <double name="fov" value="{float(cam["fov"])!r}"/>',
            f'    <double name="nearClip" value="{float(cam.get("near", 1e-4))!r}"/>',
            f'    <double name="farClip" value="{float(cam.get("far", 1e4))!r}"/>',
            f'    <integer name="sizeX" value="{int(cam["width"])}"/>',
            f'    <integer name="sizeY" value="{int(cam["height"])}"/>',
            *_filter_xml(rp.get("filter")), "  </camera>"]
    if sd.medium is not None:
        out += _medium_xml(sd.medium, directory)

    def image(name, ti, ind):  # ImageTexture (in a ScaleTexture when it has a value scale)
        t = sd.textures[ti - 1]
        fn = f"tex{ti}.hdr"
        write_rgbe(os.path.join(directory, fn), t["rgb"])
        img = [f'<texture name="{{}}" type="ImageTexture"><string name="path" value="{fn}"/>'
               f'<double name="scale" value="{float(t.get("uv_scale", 1.0))!r}"/>'
               f'<bool name="invertX" value="{"true" if t.get("invert_x") else "false"}"/>'
               f'<bool name="invertY" value="{"true" if t.get("invert_y") else "false"}"/></texture>']
        vs = float(t.get("value_scale", 1.0))
        if vs == 1.0:
            return [ind + "  " + img[0].format(name)]
        return [f'{ind}  <texture name="{name}" type="ScaleTexture"><double name="scale" value="{vs!r}"/>'
                + img[0].format("texture") + "</texture>"]

    def bsdf(i, ind="      "):
        m = sd.materials[i]
        dist = "ggx" if m.get("distribution", R.DIST_BECKMANN) == R.DIST_GGX else "beckmann"
        if m["type"] == R.MAT_DIFFUSE:
            if m.get("tex_albedo", 0):
                return [f'{ind}<bsdf type="Diffuse">'] + image("albedo", m["tex_albedo"], ind) + [f"{ind}</bsdf>"]
            return [f'{ind}<bsdf type="Diffuse">', f'{ind}  <color name="albedo" value="{_fmt(m["albedo"])}"/>',
                    f"{ind}</bsdf>"]

        def tex(name, key):  # constant textures (texture/ConstantTexture.cpp:6-10) or images
            tk = {"albedo": "tex_albedo", "transmitted": "tex_transmitted"}[key]
            if m.get(tk, 0):
                return image(name, m[tk], ind)
            c = tuple(float(x) for x in m.get(key, (1.0, 1.0, 1.0)))
            if c == (1.0, 1.0, 1.0):
                return []
            return [f'{ind}  <texture name="{name}" type="ConstantTexture"><color name="albedo" value="{_fmt(c)}"/>'
                    "</texture>"]

        def alpha():
            if m.get("tex_alpha", 0):
                return image("alpha", m["tex_alpha"], ind)
            return [f'{ind}  <double name="alpha" value="{float(m["alpha"])!r}"/>']

        if m["type"] == R.MAT_CONDUCTOR:
            return ([f'{ind}<bsdf type="Conductor">', f'{ind}  <color name="eta" value="{_fmt(m["eta"])}"/>',
                     f'{ind}  <color name="absorption" value="{_fmt(m["k"])}"/>']
                    + tex("reflectanceTexture", "albedo") + [f"{ind}</bsdf>"])
        if m["type"] == R.MAT_DIELECTRIC:
            return ([f'{ind}<bsdf type="Dielectric">', f'{ind}  <double name="etaInt" value="{float(m["eta_int"])!r}"/>',
                     f'{ind}  <double name="etaExt" value="{float(m["eta_ext"])!r}"/>']
                    + tex("reflectanceTexture", "albedo") + tex("transmittedTexture", "transmitted") + [f"{ind}</bsdf>"])
        if m["type"] == R.MAT_ROUGH_CONDUCTOR:
            return ([f'{ind}<bsdf type="RoughConductor">'] + alpha() + [
                     f'{ind}  <string name="distribution" value="{dist}"/>',
                     f'{ind}  <color name="eta" value="{_fmt(m["eta"])}"/>',
                     f'{ind}  <color name="absorption" value="{_fmt(m["k"])}"/>']
                    + tex("reflectanceTexture", "albedo") + [f"{ind}</bsdf>"])
        return ([f'{ind}<bsdf type="RoughDielectric">'] + alpha() + [
                 f'{ind}  <string name="distribution" value="{dist}"/>',
                 f'{ind}  <double name="etaInt" value="{float(m["eta_int"])!r}"/>',
                 f'{ind}  <double name="etaExt" value="{float(m["eta_ext"])!r}"/>']
                + tex("reflectanceTexture", "albedo") + tex("transmittedTexture", "transmitted") + [f"{ind}</bsdf>"])

    out.append("  <objects>")
    k = 0
    for mi, mesh in enumerate(sd.meshes):
        idx = np.asarray(mesh.indices, np.int64).reshape(-1, 3)
        tm = (np.full(len(idx), mesh.material, np.int64) if mesh.tri_material is None
              else np.asarray(mesh.tri_material, np.int64))
        for mat in dict.fromkeys(tm.tolist()):  # materials in order of first use
            sel = idx[tm == mat]
            used, first, inv = np.unique(sel.reshape(-1), return_index=True, return_inverse=True)
            # number vertices in order of first use (the file then reads like the mesh)
            order = np.argsort(first)
            remap = np.empty(len(used), np.int64)
            remap[order] = np.arange(len(used))
            tri = remap[inv].reshape(-1, 3)
            verts = np.asarray(mesh.vertices, np.float64).reshape(-1, 3)[used[order]]
            fn = f"mesh{k}.obj"
            k += 1
            with open(os.path.join(directory, fn), "w") as f:
                for v in verts.tolist():
                    f.write(f"v {v[0]!r} {v[1]!r} {v[2]!r}\n")
                has_n = mesh.normals is not None
                has_t = mesh.uvs is not None
                if has_n:
                    for v in np.asarray(mesh.normals, np.float64).reshape(-1, 3)[used[order]].tolist():
                        f.write(f"vn {v[0]!r} {v[1]!r} {v[2]!r}\n")
                if has_t:
                    for v in np.asarray(mesh.uvs, np.float32).reshape(-1, 2)[used[order]].astype(np.float64).tolist():
                        f.write(f"vt {v[0]!r} {v[1]!r}\n")
                for t in tri + 1:
                    a, b, c = t
                    if has_n and has_t:
                        f.write(f"f {a}/{a}/{a} {b}/{b}/{b} {c}/{c}/{c}\n")
                    elif has_t:
                        f.write(f"f {a}/{a} {b}/{b} {c}/{c}\n")
                    elif has_n:
                        f.write(f"f {a}//{a} {b}//{b} {c}//{c}\n")
                    else:
                        f.write(f"f {a} {b} {c}\n")
            out += ['    <shape type="Mesh">', f'      <string name="path" value="{fn}"/>']
            out += bsdf(int(mat))
            if mesh.emission is not None:
                out += ['      <light type="MeshLight">', f'        <color name="radiance" value="{_fmt(mesh.emission)}"/>',
                        "      </light>"]
            out.append("    </shape>")
    for s in sd.spheres:
        M = np.asarray(s.to_world, np.float64).reshape(16)
        out += ['    <shape type="Sphere">', f'      <double name="radius" value="{float(s.radius)!r}"/>',
                '      <transform name="toWorld">', f'        <matrix value="{_fmt(M)}"/>', "      </transform>"]
        out += bsdf(int(s.material))
        out.append("    </shape>")
    out.append("  </objects>")
    if sd.env is not None:
        write_rgbe(os.path.join(directory, "env.hdr"), sd.env)
        out += ["  <lights>", '    <light type="EnvironmentLight">', '      <string name="path" value="env.hdr"/>',
                f'      <double name="scale" value="{float(sd.env_scale)!r}"/>', "    </light>", "  </lights>"]
    out.append("</scene>")
    path = os.path.join(directory, name)
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    return path
