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


@pytest.mark.gpu
def test_served_images_match_oracle(H, O, ctx, scene1, scene2, shapes):
    # a grid request on each scene and a foveated one, served together: the grid images
    # are within 1/255 of the oracle's frame; the foveated image equals a fresh foveated
    # view's render of the same frame (no history: every request is a first frame), whose
    # samples match the oracle bit for bit and whose reconstruction stages match it
    from conftest import rgba_delta
    from test_foveated import _check_samples, _stage_checks

    hf, room, fov = shapes
    gscenes = [H.GpuScene(ctx, scene1), H.GpuScene(ctx, scene2)]
    reqs = _requests(H, scene1, scene2, shapes)[:3]
    images, st, _ = _serve(H, ctx, gscenes, shapes, reqs, 1 << 40, 4)
    assert st.completed == 3 and st.failed == 0
    for (shape, sc, m, _), img, view, A, scene in zip(reqs[:2], images[:2], (hf, room), (4, 2), (scene1, scene2)):
        ref = O.render_view(scene, view, A, camera_to_world=m)
        d = rgba_delta(img, ref["image"])
        assert d.max() <= 1, f"served {shape} image differs from the oracle by {d.max()}/255"
    _, sc, m, eye = reqs[2]
    fr = H.foveated_frame(m, eye, room.sample_to_camera, room.sample_bounds, H.foveated_history())
    gv = H.GpuFoveatedView(ctx, fov, room.width, room.height, 4)
    gv.render_foveated(gscenes[sc], fr)
    assert np.array_equal(gv.read_image(), images[2]), "served foveated image differs from a direct render"
    ref = O.render_beams(scene2, fov.beams, fov.tile_frusta, fov.block_frusta, 4, fr["eye_to_world"], (0.0, 1.0),
                         fr["cull_planes"])
    _check_samples(O, gv, ref, 4)
    _stage_checks(O, gv, fov, fr, np.zeros((room.height, room.width), np.uint32))


@pytest.mark.gpu
def test_serving_mix_four_scenes_matches_oracle(H, O, ctx):
    # the serving mix at reduced sizes: the four resident room scenes (100k, 262k, 1M and
    # 2M triangles), FP16 shading, grid requests of several sizes and MSAA 1-8 and polar
    # foveated ones, all in flight together. Grid images are within 2/255 of the FP64
    # oracle with ids and masks bit for bit behind them (a fresh view of the same request
    # reads back the GBuffer); foveated images equal a direct render whose samples match
    # the oracle bit for bit.
    from concurrent.futures import ThreadPoolExecutor

    from conftest import rgba_delta

    meshes = [391, 1024, 4096, 8192]
    hscenes = [H.HostScene(H.SCENE_ROOM, 0, n) for n in meshes]
    gscenes = [H.GpuScene(ctx, s) for s in hscenes]
    grid = [(H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, 200, 150), 1), (H.HostView(H.VIEW_PRESET_ROOM_EYE, 1, 160, 160), 8),
            (H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, 256, 144), 4), (H.HostView(H.VIEW_PRESET_ROOM_EYE, 1, 128, 96), 2)]
    fov = H.HostFoveated(H.HostFoveated.DEFAULT)
    screen = H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, 270, 300)
    srv = H.Server(ctx, 8 << 30, 6, shading_bits=16)
    gsid = [srv.add_shape(v, a) for v, a in grid]
    fsid = srv.add_foveated_shape(fov, screen, 2)
    reqs = []  # (scene, grid index or -1, camera, eye)
    for k in range(8):
        sc = k % 4
        m = _camera(grid[k % 4][0].camera_to_world, k)
        if k % 2 == 0:
            reqs.append((sc, (k // 2) % 4, m, None))
        else:
            g = np.array([0.1 * np.sin(k), 0.05 * k, -1.0], np.float32)
            reqs.append((sc, -1, _camera(screen.camera_to_world, k), g / np.linalg.norm(g)))
    out = []
    for sc, gi, m, eye in reqs:
        out.append(srv.submit(gscenes[sc], gsid[gi] if gi >= 0 else fsid, m, eye))
    srv.drain()
    st = srv.stats()
    assert st.completed == len(reqs) and st.failed == 0
    assert st.peakInFlight >= 2
    for rid, _ in out:
        assert srv.done(rid)[0]
    srv.close()

    def oracle(k):
        sc, gi, m, eye = reqs[k]
        if gi >= 0:
            return O.render_view(hscenes[sc], grid[gi][0], grid[gi][1], camera_to_world=m, precision=1)
        fr = H.foveated_frame(m, eye, screen.sample_to_camera, screen.sample_bounds, H.foveated_history())
        return fr, O.render_beams(hscenes[sc], fov.beams, fov.tile_frusta, fov.block_frusta, 2, fr["eye_to_world"],
                                  (0.0, 1.0), fr["cull_planes"], precision=1)

    with ThreadPoolExecutor(max_workers=8) as pool:
        refs = list(pool.map(oracle, range(len(reqs))))
    for k, ((sc, gi, m, eye), (_, img), ref) in enumerate(zip(reqs, out, refs)):
        what = f"request {k} ({meshes[sc]} meshes, {'grid %d' % gi if gi >= 0 else 'foveated'})"
        if gi >= 0:
            view, a = grid[gi]
            d = rgba_delta(img, ref["image"])
            assert d.max() <= 2, f"{what}: image differs from the FP64 oracle by {d.max()}/255"
            gv = H.GpuView(ctx, view, a)
            gv.set_shading_precision(16)
            gv.reserve_lists(int(ref["counts"].sum()))
            gv.render(gscenes[sc], m)
            assert np.array_equal(gv.read_image(), img), f"{what}: served image differs from a direct render"
            tri, mask = gv.read_gbuffer()
        else:
            fr, ref = ref
            gv = H.GpuFoveatedView(ctx, fov, screen.width, screen.height, 2)
            gv.set_shading_precision(16)
            gv.reserve_lists(max(int(ref["counts"].sum()), 1))
            gv.render_foveated(gscenes[sc], fr)
            assert np.array_equal(gv.read_image(), img), f"{what}: served image differs from a direct render"
            counts, lists = gv.read_tile_lists()
            assert np.array_equal(counts, ref["counts"]) and np.array_equal(lists, ref["lists"]), what
            tri, mask = gv.read_gbuffer()
        rtri, rmask = O.gbuffer_slots(ref["tri_ids"])
        assert np.array_equal(tri, rtri) and np.array_equal(mask, rmask), f"{what}: ids or masks differ"
        gv.close()
    for g in gscenes:
        g.close()


@pytest.mark.gpu
def test_device_grid_shapes_serve_like_host_shapes(H, ctx, scene2):
    # a grid shape whose views set up their samples on the device serves the same images
    # as the same camera's host-built shape, and its footprint estimate is exact
    w, h = 270, 300
    host = H.HostView(H.VIEW_PRESET_ROOM_EYE, 1, w, h)
    vp, barrel, lens = H.preset_params(H.VIEW_PRESET_ROOM_EYE, 1, w, h)
    gs = H.GpuScene(ctx, scene2)
    srv = H.Server(ctx, 8 << 30, 3)
    a = srv.add_shape(host, 4)
    b = srv.add_grid_shape(w, h, 4, vp, barrel, lens)
    reqs = [(_camera(host.camera_to_world, k), k % 2) for k in range(6)]
    out = [srv.submit(gs, (a, b)[which], m, None) for m, which in reqs]
    out += [srv.submit(gs, b if which == 0 else a, m, None) for m, which in reqs]
    srv.drain()
    st = srv.stats()
    assert st.completed == 12 and st.failed == 0
    for k in range(6):
        assert np.array_equal(out[k][1], out[6 + k][1]), f"request {k}: device and host shapes differ"
        assert (out[k][1] != 0).mean() > 0.9
    gv = H.GpuDeviceSetupView(ctx, w, h, 4, viewport=vp, barrel=barrel, lens=lens)
    assert srv.shape_bytes(b) == ctypes_bytes(H, gv)
    gv.close()
    srv.close()


@pytest.mark.gpu
def test_list_growth_stays_within_the_budget(H, ctx):
    # The 1M-triangle room at 360x400 outgrows a new view's list storage. With no room in
    # the budget for the growth the request fails with a budget error; with room once an
    # idle view of another shape is evicted, it grows, renders, and equals a plain view.
    dense = H.GpuScene(ctx, H.HostScene(H.SCENE_ROOM, 0, 4096))
    view = H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, 360, 400)
    small = H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, 64, 64)
    m = view.camera_to_world
    gv = H.GpuView(ctx, view, 1)
    gv.render(dense)
    need, cap, overflowed = gv.list_demand()
    assert overflowed and need > cap
    growth = ((need + need // 4) - cap) * 4
    probe = H.Server(ctx, 1 << 40, 1)
    big, lit = probe.shape_bytes(probe.add_shape(view, 1)), probe.shape_bytes(probe.add_shape(small, 1))
    probe.close()

    srv = H.Server(ctx, big + growth // 2, 1)
    rid, _ = srv.submit(dense, srv.add_shape(view, 1), m)
    srv.drain()
    with pytest.raises(H.HvvrError, match="budget"):
        srv.done(rid)
    assert srv.stats().peakDeviceBytes <= big + growth // 2
    srv.close()

    budget = big + growth + lit // 2
    srv = H.Server(ctx, budget, 1)
    s_small, s_big = srv.add_shape(small, 1), srv.add_shape(view, 1)
    r0, _ = srv.submit(dense, s_small, m)
    srv.drain()
    r1, img = srv.submit(dense, s_big, m)
    srv.drain()
    assert srv.done(r0)[0] and srv.done(r1)[0]
    st = srv.stats()
    assert st.viewsEvicted >= 1 and st.peakDeviceBytes <= budget
    srv.close()
    gv.reserve_lists(need)
    gv.render(dense)
    assert np.array_equal(img, gv.read_image())
