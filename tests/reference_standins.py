"""Stand-in meshes for the reference's own scene files (tests/golden/reference_scenes/*.xml), whose
OBJ files are not part of the reference's repository, and the scene descriptions those files
describe with them (for the oracle and for comparing what the loader built).

cbox_fog.xml: the Cornell box in its 0..556 units (floor, ceiling with a hole, two side walls,
back wall, two boxes) and a luminaire the file lifts by 50 above the hole, a homogeneous fog,
VolPathTracing, RandomSampler, Mitchell filter of radius 2, a mirrored lookAt camera.
testCbox.xml: a unit Cornell box with two spheres (scaled and moved by the file's transforms), a
named-metal gold conductor, VolPathTracing without a medium."""
import os

import numpy as np

from raytracer_amd import rtg, scenes

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SCENES = os.path.join(HERE, "golden", "reference_scenes")


def quad(corner, e1, e2, normal):
    return scenes.quad(corner, e1, e2, normal)


def box(center, size, inward=False):
    """Axis-aligned box (12 triangles), normals outward."""
    c = np.asarray(center, np.float64)
    h = np.asarray(size, np.float64) / 2
    V, I = [], []
    for ax in range(3):
        for sgn in (-1, 1):
            n = np.zeros(3); n[ax] = sgn
            a1, a2 = (ax + 1) % 3, (ax + 2) % 3
            e1 = np.zeros(3); e1[a1] = 2 * h[a1]
            e2 = np.zeros(3); e2[a2] = 2 * h[a2]
            corner = c.copy(); corner[ax] += sgn * h[ax]; corner[a1] -= h[a1]; corner[a2] -= h[a2]
            v, i = quad(tuple(corner), tuple(e1), tuple(e2), tuple(-n if inward else n))
            I.append(np.asarray(i) + sum(len(x) for x in V))
            V.append(np.asarray(v))
    return np.concatenate(V).astype(np.float32), np.concatenate(I).astype(np.int32)


def merge(*parts):
    V, I, off = [], [], 0
    for v, i in parts:
        V.append(np.asarray(v, np.float32)); I.append(np.asarray(i, np.int32) + off); off += len(v)
    return np.concatenate(V), np.concatenate(I)


def write_obj(path, v, i):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        for p in v:
            f.write(f"v {p[0]:.9g} {p[1]:.9g} {p[2]:.9g}\n")
        for t in i:
            f.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")


# ------------------------------------------------------------------ cbox_fog.xml
def cbox_fog_meshes():
    """name -> (vertices, indices) in the OBJ's own space, in the file's shape order."""
    W, D, Hc = 556.0, 559.2, 548.8
    hole = (213.0, 343.0, 227.0, 332.0)  # x0, x1, z0, z1
    x0, x1, z0, z1 = hole
    down = (0, -1, 0)
    ceiling = merge(quad((0, Hc, 0), (W, 0, 0), (0, 0, z0), down),
                    quad((0, Hc, z1), (W, 0, 0), (0, 0, D - z1), down),
                    quad((0, Hc, z0), (x0, 0, 0), (0, 0, z1 - z0), down),
                    quad((x1, Hc, z0), (W - x1, 0, 0), (0, 0, z1 - z0), down))
    return {
        "cbox_floor": quad((0, 0, 0), (W, 0, 0), (0, 0, D), (0, 1, 0)),
        "ceiling_with_hole": ceiling,
        "cbox_greenwall": quad((0, 0, 0), (0, Hc, 0), (0, 0, D), (1, 0, 0)),
        "cbox_redwall": quad((W, 0, 0), (0, Hc, 0), (0, 0, D), (-1, 0, 0)),
        "cbox_back": quad((0, 0, D), (W, 0, 0), (0, Hc, 0), (0, 0, -1)),
        "cbox_smallbox": box((185.0, 82.5, 169.0), (165.0, 165.0, 165.0)),
        "cbox_largebox": box((368.0, 165.0, 351.0), (165.0, 330.0, 165.0)),
        "cbox_luminaire": quad((x0, Hc - 0.1, z0), (x1 - x0, 0, 0), (0, 0, z1 - z0), down),
    }


def cbox_fog_materials():
    """(albedo per shape in file order, emission of the luminaire)."""
    alb = [(0.75, 0.75, 0.75), (0.75, 0.75, 0.75), (0.117, 0.373, 0.100), (0.61, 0.62, 0.06), (0.75, 0.75, 0.75),
           (0.75, 0.75, 0.75), (0.75, 0.75, 0.75), (0.5, 0.5, 0.5)]
    return alb, (100.0, 100.0, 100.0)


def lookat_matrix(origin, target, up):
    d = np.subtract(target, origin).astype(np.float64); d /= np.linalg.norm(d)
    u = np.asarray(up, np.float64) / np.linalg.norm(up)
    left = np.cross(u, d); left /= np.linalg.norm(left)
    nu = np.cross(d, left); nu /= np.linalg.norm(nu)
    M = np.eye(4)
    M[:3, 0], M[:3, 1], M[:3, 2], M[:3, 3] = left, nu, d, origin
    return M


def setup_cbox_fog(directory):
    """Copies cbox_fog.xml verbatim into `directory` with stand-ins for its meshes; -> path."""
    for name, (v, i) in cbox_fog_meshes().items():
        write_obj(os.path.join(directory, "data", "cbox", name + ".obj"), v, i)
    dst = os.path.join(directory, "cbox_fog.xml")
    with open(os.path.join(REF_SCENES, "cbox_fog.xml"), "rb") as f, open(dst, "wb") as g:
        g.write(f.read())
    return dst


def cbox_fog_scene(width=512, height=512, spp=256, min_depth=3, max_depth=1000):
    """The scene cbox_fog.xml describes with the stand-ins, as a SceneDesc (for the oracle)."""
    sd = scenes.SceneDesc()
    alb, Le = cbox_fog_materials()
    for k, (name, (v, i)) in enumerate(cbox_fog_meshes().items()):
        m = sd.add_material(**scenes.diffuse(alb[k]))
        v = np.asarray(v, np.float32)
        if name == "cbox_luminaire":
            v = (v + np.array([0, 50, 0], np.float32)).astype(np.float32)
            sd.meshes.append(scenes.MeshDesc(v, i, material=m, emission=Le))
        else:
            sd.meshes.append(scenes.MeshDesc(v, i, material=m))
    sd.medium = dict(homogeneous=True, sigma_a=(0.002, 0.002, 0.002), sigma_s=(0.0005, 0.0005, 0.0005), g=0.0,
                     scale=1.0)
    M = lookat_matrix((278, 273, -800), (278, 273, -799), (0, 1, 0)) @ np.diag([-1.0, 1, 1, 1])
    sd.camera = dict(origin=(278, 273, -800), target=(278, 273, -799), up=(0, 1, 0), fov=39.3077, width=width,
                     height=height, near=1e-4, far=1e4, to_world=M)
    sd.render = dict(spp=spp, sampler=rtg.SAMPLER_RANDOM, dimension=4, min_depth=min_depth, max_depth=max_depth,
                     seed=1, integrator=rtg.INTEGRATOR_VOLPATH,
                     filter=dict(type=rtg.FILTER_MITCHELL, radius=(2.0, 2.0), param=(1 / 3, 1 / 3)))
    return sd


# ------------------------------------------------------------------ testCbox.xml
def test_cbox_meshes():
    """name -> (vertices, indices) in the OBJ's own space."""
    x0, x1, y1, zb, zf = -1.02, 1.0, 1.59, -1.04, 0.99
    up, down = (0, 1, 0), (0, -1, 0)
    walls = merge(quad((x0, 0, zb), (x1 - x0, 0, 0), (0, 0, zf - zb), up),          # floor
                  quad((x0, y1, zb), (x1 - x0, 0, 0), (0, 0, zf - zb), down),       # ceiling
                  quad((x0, 0, zb), (x1 - x0, 0, 0), (0, y1, 0), (0, 0, 1)))        # back
    sv, si = scenes.icosphere(2)
    return {
        "mitsuba2": ((np.asarray(sv) + [0, 1.0, 0]).astype(np.float32), si),
        "mitsubaInterior": ((0.6 * np.asarray(sv) + [0, 1.0, 0]).astype(np.float32), si),
        "cbox/walls2": walls,
        "cbox/rightwall2": quad((x1, 0, zb), (0, y1, 0), (0, 0, zf - zb), (-1, 0, 0)),
        "cbox/leftwall2": quad((x0, 0, zb), (0, y1, 0), (0, 0, zf - zb), (1, 0, 0)),
        "cbox/light": quad((-0.24, y1 - 0.01, -0.22), (0.47, 0, 0), (0, 0, 0.38), down),
    }


def setup_test_cbox(directory):
    """testCbox.xml verbatim in `directory` with stand-in meshes and the gold spectra; -> path."""
    for name, (v, i) in test_cbox_meshes().items():
        write_obj(os.path.join(directory, "data", name + ".obj"), v, i)
    for f in ("Au.eta.spd", "Au.k.spd"):
        os.makedirs(os.path.join(directory, "data", "IOR"), exist_ok=True)
        with open(os.path.join(REF_SCENES, "data", "IOR", f), "rb") as a, \
                open(os.path.join(directory, "data", "IOR", f), "wb") as b:
            b.write(a.read())
    dst = os.path.join(directory, "testCbox.xml")
    with open(os.path.join(REF_SCENES, "testCbox.xml"), "rb") as f, open(dst, "wb") as g:
        g.write(f.read())
    return dst


def gold_eta_k():
    """eta and k of the file's "gold" from the reference's Au spectra (tools/Common.cpp:56-110)."""
    out = []
    for f in ("Au.eta.spd", "Au.k.spd"):
        lam, val = [], []
        for line in open(os.path.join(REF_SCENES, "data", "IOR", f)):
            p = line.split()
            if len(p) >= 2 and not line.lstrip().startswith("#"):
                try:
                    lam.append(float(p[0])); val.append(float(p[1]))
                except ValueError:
                    lam.append(0.0); val.append(0.0)
        out.append(rtg.spectrum_to_rgb(lam, val))
    return out
