"""The camera sample setup on the device (hvvr_gpu_view_create_grid / _create_from_samples;
Camera::setSamples, camera.cpp:21-64; naiveXYCluster, sample_clustering.h:17-49;
sample_hierarchy.cpp:46-190).

GPU: the beams, tile and block frusta, remap and sample bounds built on the device equal
the host setup's bit for bit: the full 1080x1200 barrel-distorted eye, the 640x480
heightfield camera, an odd size with padding, split-colour samples with a thin lens given
as raw samples; a spherical section agrees to within the last bits of sinf / cosf. A view
built on the device renders the host-built view's image."""

import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same_setup(dev, host, what):
    for k in ("beams", "tile_frusta", "block_frusta", "remap"):
        a, b = getattr(dev.view, k), getattr(host, k)
        assert a.shape == b.shape, f"{what}: {k} shape"
        assert a.tobytes() == b.tobytes(), f"{what}: {k} differs"
    assert np.array_equal(dev.view.sample_bounds, np.asarray(host.sample_bounds, np.float32).reshape(4))
    assert (dev.view.sample_count, dev.view.valid_sample_count, dev.view.tile_count) == \
        (host.sample_count, host.valid_sample_count, host.tile_count)


@pytest.mark.parametrize("preset,eye,w,h", [("room", 0, 1080, 1200), ("room", 1, 333, 211), ("hf", 0, 640, 480)])
def test_grid_setup_equals_host(H, ctx, preset, eye, w, h):
    kind = H.VIEW_PRESET_ROOM_EYE if preset == "room" else H.VIEW_PRESET_HEIGHTFIELD
    t0 = time.perf_counter()
    host = H.HostView(kind, eye, w, h)
    t1 = time.perf_counter()
    vp, barrel, lens = H.preset_params(kind, eye, w, h)
    dev = H.GpuDeviceSetupView(ctx, w, h, 4, viewport=vp, barrel=barrel, lens=lens,
                               camera_to_world=host.camera_to_world)
    t2 = time.perf_counter()
    print(f"{preset} {w}x{h}: host setup {1e3 * (t1 - t0):.1f} ms, device setup {1e3 * (t2 - t1):.1f} ms", flush=True)
    _same_setup(dev, host, f"{preset} {w}x{h}")
    dev.close()


def test_raw_samples_split_and_lens_equal_host(H, ctx):
    w, h, split = 270, 300, 3
    base = H.grid_samples(w, h, 0.22, 0.24)
    samples = np.repeat(base, split)
    for c in range(split):
        samples["position"][c::split, 0] += np.float32((c - 1) / (split * w))
    vp = (-0.8, -0.9, 0.8, 0.9)
    host = H.HostView.from_samples(samples, w, h, vp, split=split, lens_radius=0.01, focal_distance=3.0)
    dev = H.GpuDeviceSetupView(ctx, w, h, 2, viewport=vp, samples=samples, split=split, lens=(0.01, 3.0))
    _same_setup(dev, host, "split 3 with a thin lens")
    dev.close()


def test_spherical_setup_close_to_host(H, ctx):
    w, h = 320, 160
    vp = (-1.0, -0.5, 1.0, 0.5)
    host = H.HostView.from_samples(H.grid_samples(w, h), w, h, vp, spherical=(150.0, 75.0))
    dev = H.GpuDeviceSetupView(ctx, w, h, 1, viewport=vp, spherical=(150.0, 75.0))
    assert np.array_equal(dev.view.remap, host.remap)
    for k in ("beams", "tile_frusta", "block_frusta"):
        a = getattr(dev.view, k).view(np.float32).reshape(-1)
        b = getattr(host, k).view(np.float32).reshape(-1)
        assert np.abs(a - b).max() <= 4e-7, k
    dev.close()


def test_device_built_view_renders_like_host_view(H, ctx, scene2):
    w, h = 540, 600
    host = H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, w, h)
    vp, barrel, lens = H.preset_params(H.VIEW_PRESET_ROOM_EYE, 0, w, h)
    dev = H.GpuDeviceSetupView(ctx, w, h, 4, viewport=vp, barrel=barrel, lens=lens,
                               camera_to_world=host.camera_to_world)
    gs = H.GpuScene(ctx, scene2)
    ref = H.GpuView(ctx, host, 4)
    ref.render(gs)
    dev.render(gs)
    img = dev.read_image()
    assert (img != 0).mean() > 0.9
    assert np.array_equal(img, ref.read_image())
    assert np.array_equal(dev.read_gbuffer()[0], ref.read_gbuffer()[0])
    ref.close()
    dev.close()
