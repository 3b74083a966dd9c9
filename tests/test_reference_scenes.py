"""The reference renderer's own scene files, loaded verbatim (tests/golden/reference_scenes: byte
copies of RaytracerNew2/Scenes/cbox_fog.xml, testCbox.xml and veach-MIS.xml) by rtg_scene_load_xml,
on the CPU (librtg's host sources against the HIP stand-in, in a child process as in
tests/test_xml_host.py). The files' meshes are not in the reference's repository: stand-ins are
written beside the copy (tests/reference_standins.py). What the loader built is compared with what
the file says, element by element (core/XMLLoader.cpp:33-904):
  * camera: NoriPerspective fov and size; toWorld = lookAt x scale(-1, 1, 1), the file's elements
    composed by left multiplication (XMLLoader.cpp:845-890); the Mitchell filter of radius 2 with
    the default B = C = 1/3;
  * sampler RandomSampler and its sampleNumber; integrator VolPathTracing;
  * the top-level Homogeneous medium's sigmaA / sigmaS (cbox_fog.xml);
  * shapes: triangle counts and the bounding box of the stand-ins after each shape's toWorld, one
    MeshLight; materials: every Diffuse albedo of the file, the default albedo of a Diffuse without
    one, the gold Conductor's eta / k from the reference's Au spectra (testCbox.xml);
  * XML comments (including commented-out scene blocks), a UTF-8 byte-order mark, the
    <parameters> block, attribute spacing like value ="..." are accepted;
  * veach-MIS.xml (DirectLightingMIS, DistantLight, Microfacet: outside this renderer) fails
    with RTG_ERR_PARSE and a message naming its integrator, and the context stays usable.
tests/test_gpu_xml.py renders cbox_fog.xml on the GPU against the oracle."""
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)


@pytest.fixture(scope="module")
def mock_lib():
    from test_serving_host import build_mock_lib

    return build_mock_lib()


def _run(mock_lib, case, tmp_path):
    code = textwrap.dedent(f"""
        import sys
        sys.path.insert(0, {ROOT!r}); sys.path.insert(0, {HERE!r})
        from raytracer_amd import rtg
        rtg.LIB_PATH = {mock_lib!r}
        import test_reference_scenes as T
        ctx = rtg.Context(0)
        T.{case}(ctx, {str(tmp_path)!r})
        ctx.close()
        print("CASE OK", flush=True)
    """)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, MALLOC_CHECK_="3"))
    assert r.returncode == 0 and "CASE OK" in r.stdout, (r.stdout[-2000:], r.stderr[-3000:])


def _bbox(sd):
    v = np.concatenate([np.asarray(m.vertices, np.float32) for m in sd.meshes])
    return v.min(0), v.max(0)


def _materials(sc, n):
    return [sc.material(i) for i in range(n)]


# ------------------------------------------------------------------ cases (run in the child)
def case_cbox_fog(ctx, tmp):
    from raytracer_amd import rtg
    import reference_standins as RS

    path = RS.setup_cbox_fog(tmp)
    raw = open(path, "rb").read()
    assert raw.startswith(b"\xef\xbb\xbf") and b"<parameters>" in raw  # loaded as the reference wrote it
    sc, cam, rp = ctx.load_xml(path)
    ref = RS.cbox_fog_scene()
    info = sc.info()
    assert info["triangles"] == sum(len(m.indices) for m in ref.meshes)
    assert info["lights"] == 1 and info["spheres"] == 0
    lo, hi = _bbox(ref)
    assert np.allclose(info["bbox_min"], lo) and np.allclose(info["bbox_max"], hi)  # luminaire lifted by 50
    assert info["medium_kind"] == 2
    assert np.allclose(info["medium_sigma_a"], 0.002) and np.allclose(info["medium_sigma_s"], 0.0005)
    assert cam["width"] == 512 and cam["height"] == 512 and np.isclose(cam["fov"], 39.3077)
    assert np.allclose(cam["to_world"], ref.camera["to_world"], atol=1e-6)
    assert rp["spp"] == 256 and rp["sampler"] == rtg.SAMPLER_RANDOM
    assert rp["integrator"] == rtg.INTEGRATOR_VOLPATH
    f = rp["filter"]
    assert f["type"] == rtg.FILTER_MITCHELL and np.allclose(f["radius"], 2.0)
    assert all(p <= 0 or np.isclose(p, 1 / 3) for p in f["param"])  # unset -> the defaults B = C = 1/3
    mats = _materials(sc, info["materials"])
    albedos = sorted(tuple(np.round(m["albedo"], 5)) for m in mats)
    alb, _ = RS.cbox_fog_materials()
    assert albedos == sorted(tuple(np.round(a, 5)) for a in alb)
    assert all(m["type"] == rtg.MAT_DIFFUSE for m in mats)


def case_test_cbox(ctx, tmp):
    from raytracer_amd import rtg
    import reference_standins as RS

    path = RS.setup_test_cbox(tmp)
    sc, cam, rp = ctx.load_xml(path)
    info = sc.info()
    meshes = RS.test_cbox_meshes()
    assert info["triangles"] == sum(len(i) for _, i in meshes.values())
    assert info["lights"] == 1 and info["medium_kind"] == 0
    # the two spheres: toWorld = translate(0.2, 0, -0.5) x scale(0.3) on the stand-in
    pts = [np.asarray(meshes[k][0], np.float64) * 0.3 + [0.2, 0, -0.5] for k in ("mitsuba2", "mitsubaInterior")]
    pts += [np.asarray(meshes[k][0], np.float64) for k in ("cbox/walls2", "cbox/rightwall2", "cbox/leftwall2",
                                                               "cbox/light")]
    allp = np.concatenate(pts)
    assert np.allclose(info["bbox_min"], allp.min(0), atol=1e-5) and np.allclose(info["bbox_max"], allp.max(0), atol=1e-5)
    M = RS.lookat_matrix((0, 0.999769, 3.41159), (0, 0.89305, 2.41198), (0, 1, 0)) @ np.diag([-1.0, 1, 1, 1])
    assert np.allclose(cam["to_world"], M, atol=1e-6)
    assert cam["width"] == 800 and cam["height"] == 600 and np.isclose(cam["fov"], 27.7856)
    assert rp["spp"] == 64 and rp["sampler"] == rtg.SAMPLER_RANDOM and rp["integrator"] == rtg.INTEGRATOR_VOLPATH
    assert rp["filter"]["type"] == rtg.FILTER_MITCHELL and np.allclose(rp["filter"]["radius"], 2.0)
    mats = _materials(sc, info["materials"])
    gold = [m for m in mats if m["type"] == rtg.MAT_CONDUCTOR]
    assert len(gold) == 1
    eta, k = RS.gold_eta_k()
    assert np.allclose(gold[0]["eta"], eta, rtol=1e-6) and np.allclose(gold[0]["k"], k, rtol=1e-6)
    diff = sorted(tuple(np.round(m["albedo"], 5)) for m in mats if m["type"] == rtg.MAT_DIFFUSE)
    expect = sorted([(0.2, 0.8, 0.8), (0.725, 0.71, 0.68), (0.161, 0.133, 0.427), (0.63, 0.065, 0.05), (0.5, 0.5, 0.5)])
    assert diff == expect, diff


def case_veach_mis_rejected(ctx, tmp):
    import shutil
    from raytracer_amd import rtg
    import reference_standins as RS

    dst = os.path.join(tmp, "veach-MIS.xml")
    shutil.copyfile(os.path.join(RS.REF_SCENES, "veach-MIS.xml"), dst)
    try:
        ctx.load_xml(dst)
    except rtg.RtgError as e:
        assert e.code == -4, e  # RTG_ERR_PARSE
        msg = str(e)
        assert "DirectLightingMIS" in msg, msg  # the first element outside the GPU path
    else:
        raise AssertionError("veach-MIS.xml should not load")
    # the context stays usable
    sc, cam, rp = ctx.load_xml(RS.setup_cbox_fog(tmp))
    assert sc.info()["lights"] == 1


# ------------------------------------------------------------------ tests (parent)
@pytest.mark.parametrize("case", ["case_cbox_fog", "case_test_cbox", "case_veach_mis_rejected"])
def test_reference_scene_file(mock_lib, case, tmp_path):
    _run(mock_lib, case, tmp_path)
