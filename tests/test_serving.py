"""Serving loop (include/hvvr_server.h): resident scenes, view shapes admitted by their
device footprint, requests overlapping on the device on their own streams, requests
submitted while others run admitted as slots and memory free up.

GPU: every image of a mixed run (two scenes, grid shapes of two sizes and MSAA rates, a
polar foveated shape; more requests than slots; a budget that forces evictions) equals the
same request served alone, bit for bit, and equals the plain render of the same view."""

import numpy as np
import pytest


def _camera(base, k):
    m = np.array(base, np.float32).reshape(4, 4).copy()
    m[3, :3] += np.float32(0.05) * np.float32(k) * np.array([1.0, 0.3, -0.5], np.float32)
    return m.reshape(16)


@pytest.fixture(scope="module")
def shapes(H):
    hf = H.HostView(H.VIEW_PRESET_HEIGHTFIELD, 0, 320, 240)
    room = H.HostView(H.VIEW_PRESET_ROOM_EYE, 1, 270, 300)
    fov = H.HostFoveated(H.HostFoveated.DEFAULT)
    return hf, room, fov


def _requests(H, scene1, scene2, shapes):
    hf, room, fov = shapes
    reqs = []
    for k in range(12):
        kind = k % 3
        if kind == 0:
            reqs.append(("hf", 0, _camera(hf.camera_to_world, k), None))
        elif kind == 1:
            reqs.append(("room", 1, _camera(room.camera_to_world, k), None))
        else:
            g = np.array([np.sin(0.05 * k), 0.02 * k, -1.0], np.float32)
            reqs.append(("fov", 1, _camera(room.camera_to_world, k), g / np.linalg.norm(g)))
    return reqs


def _serve(H, ctx, gscenes, shapes, reqs, budget, slots, one_at_a_time=False):
    hf, room, fov = shapes
    srv = H.Server(ctx, budget, slots)
    sid = {"hf": srv.add_shape(hf, 4), "room": srv.add_shape(room, 2), "fov": srv.add_foveated_shape(fov, room, 4)}
    out = []
    for k, (shape, scene, m, eye) in enumerate(reqs):
        rid, img = srv.submit(gscenes[scene], sid[shape], m, eye)
        out.append((rid, img))
        if one_at_a_time:
            srv.drain()
        elif k == 5:
            srv.poll()  # the first requests are on the device while the rest arrive
    srv.drain()
    for rid, _ in out:
        done, lat = srv.done(rid)
        assert done and lat > 0
    st = srv.stats()
    sizes = {k: srv.shape_bytes(v) for k, v in sid.items()}
    srv.close()
    return [img for _, img in out], st, sizes


@pytest.mark.gpu
def test_mixed_requests_match_solo(H, ctx, scene1, scene2, shapes):
    gscenes = [H.GpuScene(ctx, scene1), H.GpuScene(ctx, scene2)]
    reqs = _requests(H, scene1, scene2, shapes)
    # probe the shape sizes, then a budget that holds the two largest views at once
    _, _, sizes = _serve(H, ctx, gscenes, shapes, reqs[:3], 1 << 40, 1)
    big = sorted(sizes.values())
    budget = big[-1] + big[-2] + 1
    mixed, st, _ = _serve(H, ctx, gscenes, shapes, reqs, budget, 4)
    assert st.completed == len(reqs) and st.queued == 0 and st.inFlight == 0
    assert st.peakInFlight >= 2, "requests should overlap on the device"
    assert st.viewsEvicted > 0, "the budget should force evictions"
    assert st.peakDeviceBytes <= budget
    assert st.p99Ms >= st.p50Ms > 0
    solo, st1, _ = _serve(H, ctx, gscenes, shapes, reqs, 1 << 40, 1, one_at_a_time=True)
    assert st1.peakInFlight == 1
    for k, (a, b) in enumerate(zip(mixed, solo)):
        assert (a != 0).mean() > 0.5, f"request {k} rendered nothing"
        assert np.array_equal(a, b), f"request {k} differs between the mix and alone"
    # the served grid image is the plain render of the same view
    hf = shapes[0]
    gv = H.GpuView(ctx, hf, 4)
    gv.render(gscenes[0], reqs[0][2])
    assert np.array_equal(gv.read_image(), mixed[0])


@pytest.mark.gpu
def test_shape_beyond_budget_is_rejected(H, ctx, shapes):
    srv = H.Server(ctx, 1 << 20, 2)
    with pytest.raises(H.HvvrError, match="budget"):
        srv.add_shape(shapes[1], 4)
    srv.close()


@pytest.mark.gpu
def test_view_footprint_estimate_is_exact(H, ctx, shapes):
    hf, room, fov = shapes
    srv = H.Server(ctx, 1 << 40, 1)
    s = srv.add_shape(room, 2)
    gv = H.GpuView(ctx, room, 2)
    b = ctypes_bytes(H, gv)
    assert srv.shape_bytes(s) == b
    srv.close()


def ctypes_bytes(H, gv):
    import ctypes
    out = ctypes.c_uint64()
    H._check_gpu(gv.lib.hvvr_gpu_view_device_bytes(gv.handle, ctypes.byref(out)))
    return out.value
