"""C ABI and host logic on the CPU (no GPU calls)."""
import ctypes
import os
import re
import subprocess

import numpy as np

from raytracer_amd import rtg, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "rtg.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(rtg_\w+)\s*\(", hdr, re.M)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 18
    lib = ctypes.CDLL(rtg.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", rtg.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (rtg_\w+)", out))
    assert set(syms) <= exported
    # the Python binding declares a signature for each of them
    assert set(syms) <= set(rtg.SIGNATURES)


def test_struct_layouts_match_header(tmp_path):
    """The ctypes structures have the sizes and field offsets the C compiler gives include/rtg.h."""
    structs = {"rtg_material": rtg.rtg_material, "rtg_mesh_desc": rtg.rtg_mesh_desc,
               "rtg_sphere_desc": rtg.rtg_sphere_desc, "rtg_scene_desc": rtg.rtg_scene_desc,
               "rtg_render_desc": rtg.rtg_render_desc, "rtg_stats": rtg.rtg_stats,
               "rtg_server_stats": rtg.rtg_server_stats, "rtg_scene_info": rtg.rtg_scene_info}
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "rtg.h"', "int main(void) {"]
    for name, cls in structs.items():
        src.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for fname, _ in cls._fields_:
            src.append(f'printf("{name}.{fname} %zu\\n", offsetof({name}, {fname}));')
    src.append("return 0; }")
    c = tmp_path / "sizes.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.splitlines())
    for name, cls in structs.items():
        assert int(out[name]) == ctypes.sizeof(cls), name
        for fname, _ in cls._fields_:
            assert int(out[f"{name}.{fname}"]) == getattr(cls, fname).offset, (name, fname)


def test_icosphere_counts_and_closure():
    v, f = scenes.icosphere(3)
    assert f.shape == (20 * 4 ** 3, 3) and v.shape == (10 * 4 ** 3 + 2, 3)
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0)
    # closed manifold: every edge shared by exactly two faces
    e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    assert np.all(counts == 2)


def test_procedural_scenes_deterministic():
    a = scenes.displaced_sphere(level=4)
    b = scenes.displaced_sphere(level=4)
    assert np.array_equal(a.meshes[0].vertices, b.meshes[0].vertices)
    assert [m.get("alpha") for m in a.materials] == [m.get("alpha") for m in b.materials]
    alphas = np.array([m["alpha"] for m in a.materials[1:]])
    assert alphas.min() >= 0.05 and alphas.max() <= 0.4
    c = scenes.cornell()
    assert c.num_triangles() == 10 + 24 + 2
    assert c.meshes[-1].emission == (17.0, 12.0, 4.0)


def test_cornell_normals_face_inside():
    c = scenes.cornell()
    walls = c.meshes[0]
    p = walls.vertices[walls.indices]
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    centre = np.array([0.0, 0.5, 0.0])
    assert np.all(np.einsum("ij,ij->i", n, centre - p[:, 0]) > 0)
