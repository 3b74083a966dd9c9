"""Serving loop (continuous batching on the shared path pool): a request renders the same image
alone (rtg_render without a server, same seed) and inside a mix of scenes — plain and volumetric
(VolPathTracing) — resolutions and spp that arrive while others run; HBM-budget admission;
cancellation; concurrent submission from another thread; the server's own loop thread with
rtg_wait; rtg_render and rtg_render_device on a serving context joining the mix.

Tolerance: the film accumulates with float atomics, so only the order of additions may differ
between a request alone and in the mix: allclose(rtol=1e-5, atol=1e-6) on (rgb sum, weight)."""
import threading
import time

import numpy as np
import pytest
# torch (device buffers for the caller-owned film case) is imported before librtg creates a
# context: torch ships its own HIP runtime, which then serves both; imported after, it finds no GPU
import torch

from raytracer_amd import rtg, scenes

pytestmark = [pytest.mark.gpu]


@pytest.fixture(scope="module")
def ctx():
    c = rtg.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def resident(ctx):
    out = {}
    for name, sd in (("cornell", scenes.cornell()), ("glossy", scenes.glossy_cornell()),
                     ("displaced", scenes.displaced_sphere(level=5)), ("vol", scenes.vol_cornell(res=32))):
        out[name] = (sd, ctx.upload(sd))
    return out


# (scene, width, height, spp, seed)
MIX = [
    ("cornell", 64, 48, 16, 5),
    ("glossy", 96, 64, 64, 6),
    ("displaced", 128, 128, 4, 7),
    ("cornell", 32, 32, 256, 8),
    ("glossy", 128, 96, 16, 9),
    ("displaced", 48, 80, 64, 10),
    ("cornell", 200, 120, 4, 11),
]


def _args(resident, name, w, h, spp, seed):
    sd, sc = resident[name]
    return sc, dict(sd.camera, width=w, height=h), dict(sd.render, spp=spp, seed=seed)


def _alone(resident):
    imgs = []
    for name, w, h, spp, seed in MIX:
        sc, cam, rp = _args(resident, name, w, h, spp, seed)
        imgs.append(sc.render(cam, rp)[0])
    return imgs


def test_same_alone_and_in_mix(ctx, resident):
    print("[serving] alone vs staggered mix of 7 requests, 3 scenes", flush=True)
    alone = _alone(resident)
    with rtg.Server(ctx) as srv:
        # first three arrive, the loop runs a little, the rest arrive while those are in flight
        ids = [srv.submit(*_args(resident, *MIX[i])) for i in range(3)]
        srv.step(max_ms=1.0)
        ids += [srv.submit(*_args(resident, *MIX[i])) for i in range(3, len(MIX))]
        srv.step()
        st = srv.stats()
        for i, jid in enumerate(ids):
            status, img, lat = srv.poll(jid)
            assert status == "done", (i, status)
            assert lat > 0
            assert np.allclose(img, alone[i], rtol=1e-5, atol=1e-6), (i, np.abs(img - alone[i]).max())
            assert img[..., 3].sum() > 0
    print(f"  stats {st}", flush=True)
    assert st["completed"] == len(MIX) and st["cancelled"] == 0
    assert st["max_running"] >= 4  # the requests shared iterations
    assert st["film_bytes"] == 0
    assert st["samples"] == sum(w * h * spp for _, w, h, spp, _ in MIX)


def test_render_on_a_serving_context_joins_the_mix(ctx, resident):
    """rtg_render while a server exists becomes a request of the server (its caller runs the loop
    when no server thread does) and returns the image it gives without the server; a guided
    render, which the server does not take, fails; one server per context."""
    sc, cam, rp = _args(resident, *MIX[0])
    vsc, vcam, vrp = _args(resident, "vol", 40, 32, 16, 3)
    before = sc.render(cam, rp)[0]
    vbefore = vsc.render(vcam, vrp)[0]
    srv = rtg.Server(ctx)
    try:
        during, st = sc.render(cam, rp)
        assert st.samples == cam["width"] * cam["height"] * rp["spp"]
        assert np.allclose(before, during, rtol=1e-5, atol=1e-6)
        assert np.allclose(vbefore, vsc.render(vcam, vrp)[0], rtol=1e-5, atol=1e-6)
        with pytest.raises(rtg.RtgError):
            sc.render(cam, dict(rp, guiding=1))
        with pytest.raises(rtg.RtgError):
            rtg.Server(ctx)  # one server per context
        assert srv.stats()["completed"] == 2
    finally:
        srv.close()
    after = sc.render(cam, rp)[0]
    assert np.allclose(before, after, rtol=1e-5, atol=1e-6)


def test_volumetric_requests_share_the_pool(ctx, resident):
    """VolPathTracing requests run in mixed iterations next to PathTracingMIS requests of other
    scenes; each equals its own render alone."""
    reqs = [("vol", 48, 40, 16, 51), ("cornell", 64, 48, 16, 52), ("vol", 32, 32, 64, 53),
            ("displaced", 64, 64, 4, 54), ("glossy", 48, 48, 16, 55)]
    alone = []
    for r in reqs:
        sc, cam, rp = _args(resident, *r)
        alone.append(sc.render(cam, rp)[0])
    with rtg.Server(ctx) as srv:
        ids = [srv.submit(*_args(resident, *r)) for r in reqs]
        srv.step()
        st = srv.stats()
        for i, jid in enumerate(ids):
            status, img, _ = srv.poll(jid)
            assert status == "done", (i, status)
            assert np.allclose(img, alone[i], rtol=1e-5, atol=1e-6), (reqs[i], np.abs(img - alone[i]).max())
    assert st["completed"] == len(reqs) and st["max_running"] >= 3


def server_thread_and_wait(ctx, resident, device_film=True):
    """rtg_server_start: the server's thread runs the loop; rtg_wait blocks until each request is
    done; (device_film) a request accumulates into a caller-owned device film through
    rtg_render_device; a wait with a short timeout on a long request reports it still running; the
    loop sleeps when idle. Returns the server's final stats."""
    reqs = [("cornell", 64, 64, 16, 61), ("vol", 32, 32, 16, 62), ("glossy", 48, 32, 64, 63)]
    alone = []
    for r in reqs:
        sc, cam, rp = _args(resident, *r)
        alone.append(sc.render(cam, rp)[0])
    with rtg.Server(ctx) as srv:
        srv.start()
        with pytest.raises(rtg.RtgError):
            srv.step()  # the server thread owns the loop
        ids = [srv.submit(*_args(resident, *r)) for r in reqs]
        for i, jid in enumerate(ids):
            status, img, lat = srv.wait(jid)
            assert status == "done" and lat > 0
            assert np.allclose(img, alone[i], rtol=1e-5, atol=1e-6), reqs[i]
        if device_film:
            # a device film of the caller's (not cleared: the render adds into it)
            sc, cam, rp = _args(resident, *reqs[0])
            film = torch.full((cam["height"] * cam["width"] * 4,), 1.0, dtype=torch.float32, device="cuda")
            st = sc.render_device(cam, rp, film.data_ptr())
            torch.cuda.synchronize()
            got = film.cpu().numpy().reshape(cam["height"], cam["width"], 4) - 1.0
            assert st.samples == cam["width"] * cam["height"] * rp["spp"]
            assert np.allclose(got, alone[0], rtol=1e-5, atol=1e-5)
        big = srv.submit(*_args(resident, "displaced", 512, 512, 256, 64))
        status, _, _ = srv.wait(big, timeout_ms=1.0, want_image=False)
        assert status == "running"
        srv.cancel(big)
        assert srv.wait(big)[0] == "cancelled"
        stt = srv.stats()
        assert stt["running"] == 0 and stt["pending"] == 0
    return stt


def test_server_thread_wait_and_device_film(ctx, resident):
    stt = server_thread_and_wait(ctx, resident)
    assert stt["graph_launches"] > 0  # iteration batches replayed from hipGraphs


def test_budget_admission_serialises_films(ctx, resident):
    """A budget that holds one 128x128 film admits the requests one at a time; results unchanged."""
    reqs = [("cornell", 128, 128, 16, 21), ("glossy", 128, 128, 16, 22), ("displaced", 64, 64, 16, 23)]
    alone = []
    for r in reqs:
        sc, cam, rp = _args(resident, *r)
        alone.append(sc.render(cam, rp)[0])
    budget = 128 * 128 * 16
    with rtg.Server(ctx, hbm_budget_bytes=budget) as srv:
        ids = [srv.submit(*_args(resident, *r)) for r in reqs]
        # a request whose film alone exceeds the budget is refused at admission
        big = srv.submit(*_args(resident, "cornell", 256, 256, 4, 24))
        srv.step()
        st = srv.stats()
        for i, jid in enumerate(ids):
            status, img, _ = srv.poll(jid)
            assert status == "done"
            assert np.allclose(img, alone[i], rtol=1e-5, atol=1e-6)
        assert srv.poll(big)[0] == "cancelled"
    assert st["film_bytes_peak"] <= budget
    assert st["max_running"] == 1
    assert st["completed"] == 3 and st["cancelled"] == 1


def test_cancel_drains_and_others_finish(ctx, resident):
    small = ("cornell", 64, 64, 16, 31)
    sc, cam, rp = _args(resident, *small)
    ref = sc.render(cam, rp)[0]
    with rtg.Server(ctx) as srv:
        big = srv.submit(*_args(resident, "displaced", 512, 512, 256, 32))  # 67M samples
        srv.step(max_ms=5.0)
        jid = srv.submit(*_args(resident, *small))
        srv.cancel(big)
        srv.step()
        assert srv.poll(big)[0] == "cancelled"
        status, img, _ = srv.poll(jid)
        assert status == "done"
        assert np.allclose(img, ref, rtol=1e-5, atol=1e-6)
        st = srv.stats()
    assert st["cancelled"] == 1 and st["completed"] == 1 and st["running"] == 0


def test_submit_from_another_thread_while_stepping(ctx, resident):
    """rtg_submit does not wait for a running step: a producer thread submits at intervals while
    the main thread steps; every request finishes with its alone image."""
    reqs = [("cornell", 64, 64, 64, 40 + i) if i % 2 == 0 else ("glossy", 64, 64, 64, 40 + i) for i in range(6)]
    alone = []
    for r in reqs:
        sc, cam, rp = _args(resident, *r)
        alone.append(sc.render(cam, rp)[0])
    with rtg.Server(ctx) as srv:
        ids = []

        def producer():
            for r in reqs:
                ids.append(srv.submit(*_args(resident, *r)))
                time.sleep(0.004)

        t = threading.Thread(target=producer)
        t.start()
        t_end = time.time() + 60
        while (t.is_alive() or srv.stats()["running"] + srv.stats()["pending"] > 0) and time.time() < t_end:
            srv.step(max_ms=2.0)
        t.join()
        srv.step()
        for i, jid in enumerate(ids):
            status, img, _ = srv.poll(jid)
            assert status == "done"
            assert np.allclose(img, alone[i], rtol=1e-5, atol=1e-6)
