"""Python bindings for the MI355X-native HVVR render path.

Two in-tree shared libraries are loaded with ctypes:

* ``lib/libhvvr_gpu.so``  - gfx950 HIP kernels behind the C ABI of ``include/hvvr_gpu.h``;
* ``lib/libhvvr.so``      - host C++ (BVH builder, scene flattening, sample setup, seeded
  scene generators, the ``hvvr::Raycaster`` C++ API) with the C interface of
  ``include/hvvr_host.h``.

Arrays cross the boundary as numpy arrays with the reference's byte layouts (see
``DTYPES``). Every GPU entry point raises :class:`HvvrError` on failure; nothing here
falls back to a CPU implementation.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, byref, c_char_p, c_float, c_int, c_uint32, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")

__all__ = [
    "HvvrError", "DTYPES", "load_gpu", "load_host", "HostScene", "HostView", "GpuContext", "GpuScene",
    "GpuView", "render_views", "subsample_offsets", "lens_table", "grid_samples", "transform_matrix",
]


class HvvrError(RuntimeError):
    pass


# ---------------------------------------------------------------------------------------
# struct layouts (byte-identical to the reference's)
# ---------------------------------------------------------------------------------------
DTYPES = {
    "vertex": np.dtype([("pos", "<f4", 3), ("normal", "<u2", 4), ("uv", "<f4", 2), ("pad", "<u4")]),
    "tri_shade": np.dtype([("indices", "<u4", 3), ("material", "<u4")]),
    "tri_intersect": np.dtype([("v0", "<f4", 3), ("edge0", "<f4", 3), ("edge1", "<f4", 3)]),
    "node": np.dtype([("bounds", "<f4", (6, 4)), ("leafMask", "<u4"), ("data", "<u4", 5), ("pad", "<u4", 2)]),
    "material": np.dtype([("emissive", "<f4", 4), ("diffuse", "<f4", 4), ("specular", "<f4", 4),
                          ("glossiness", "<f4"), ("opacity", "<f4"), ("textureIDsAndShadingModel", "<u8")]),
    "light": np.dtype([("type", "<u4"), ("body", "<f4", 13)]),
    "beam": np.dtype([("centerRay", "<f4", 3), ("du", "<f4", 3), ("dv", "<f4", 3)]),
    "sfrustum": np.dtype([("origins", "<f4", (4, 3)), ("directions", "<f4", (4, 3))]),
    "sample": np.dtype([("pixel", "<u4", 2), ("position", "<f4", 2), ("minorAxis", "<f4", 2),
                        ("majorAxisLength", "<f4")]),
}
assert DTYPES["vertex"].itemsize == 32 and DTYPES["node"].itemsize == 128
assert DTYPES["material"].itemsize == 64 and DTYPES["light"].itemsize == 56
assert DTYPES["beam"].itemsize == 36 and DTYPES["sfrustum"].itemsize == 96 and DTYPES["sample"].itemsize == 28

# per-stage timings returned by GpuView.stage_times (hvvr_gpu_view_stage_times)
STAGES = ["frustum_setup", "block_traverse", "tile_lists", "tile_lists_fallback", "intersect", "resolve", "remap"]

_gpu = None
_host = None


def _ptr(a):
    return a.ctypes.data_as(c_void_p) if a is not None else None


def load_gpu():
    """Load lib/libhvvr_gpu.so (raises if it has not been built)."""
    global _gpu
    if _gpu is None:
        path = os.path.join(LIB_DIR, "libhvvr_gpu.so")
        if not os.path.exists(path):
            raise HvvrError(f"{path} is missing: run __graft_entry__.build() (or make -C hvvr_amd)")
        lib = ctypes.CDLL(path)
        lib.hvvr_gpu_last_error.restype = c_char_p
        lib.hvvr_gpu_version.restype = c_char_p
        _gpu = lib
    return _gpu


def load_host():
    """Load lib/libhvvr.so (host C++ library; depends on libhvvr_gpu.so)."""
    global _host
    if _host is None:
        load_gpu()
        path = os.path.join(LIB_DIR, "libhvvr.so")
        if not os.path.exists(path):
            raise HvvrError(f"{path} is missing: run __graft_entry__.build() (or make -C hvvr_amd)")
        lib = ctypes.CDLL(path)
        lib.hvvr_host_last_error.restype = c_char_p
        _host = lib
    return _host


def _check_gpu(rc):
    if rc < 0:
        msg = load_gpu().hvvr_gpu_last_error().decode()
        raise HvvrError(f"hvvr_gpu error {rc}: {msg}")
    return rc


def _check_host(rc):
    if rc < 0:
        msg = load_host().hvvr_host_last_error().decode()
        raise HvvrError(f"hvvr_host error {rc}: {msg}")
    return rc


# ---------------------------------------------------------------------------------------
# host-side building blocks
# ---------------------------------------------------------------------------------------
SCENE_HEIGHTFIELD_BOXES = 1
SCENE_ROOM = 2
VIEW_PRESET_HEIGHTFIELD = 1
VIEW_PRESET_ROOM_EYE = 2

_SCENE_ARRAYS = [("verts", "vertex", 0), ("vertexToTransform", np.uint32, 0), ("triShade", "tri_shade", 1),
                 ("nodes", "node", 2), ("materials", "material", 3), ("transforms", (np.float32, 16), 4),
                 ("lights", "light", 5)]


def _dt(spec):
    if isinstance(spec, str):
        return DTYPES[spec], ()
    if isinstance(spec, tuple):
        return np.dtype(spec[0]), (spec[1],)
    return np.dtype(spec), ()


class HostScene:
    """A flattened scene (BVH built, models merged) as numpy arrays."""

    def __init__(self, kind=SCENE_HEIGHTFIELD_BOXES, seed=0, mesh_count=0, _handle=None):
        lib = load_host()
        h = c_void_p()
        if _handle is None:
            _check_host(lib.hvvr_host_scene_create(c_int(kind), c_uint64(seed), c_uint32(mesh_count), byref(h)))
        else:
            h = _handle
        counts = np.zeros(7, np.uint64)
        _check_host(lib.hvvr_host_scene_counts(h, _ptr(counts)))
        n_verts, n_tris, n_nodes, n_mats, n_xf, n_lights, n_models = (int(c) for c in counts)
        sizes = [n_verts, n_verts, n_tris, n_nodes, n_mats, n_xf, n_lights]
        for (name, spec, _), n in zip(_SCENE_ARRAYS, sizes):
            dt, extra = _dt(spec)
            arr = np.zeros((n,) + extra, dt)
            if n:
                idx = [a[0] for a in _SCENE_ARRAYS].index(name)
                _check_host(lib.hvvr_host_scene_copy(h, c_int(idx), _ptr(arr), c_uint64(arr.nbytes)))
            setattr(self, name, arr)
        self.model_count = n_models
        lib.hvvr_host_scene_destroy(h)

    @classmethod
    def from_bin(cls, path):
        """A scene file in the reference's binary format (include/hvvr/model_import.h)."""
        h = c_void_p()
        _check_host(load_host().hvvr_host_scene_load_bin(os.fsencode(path), byref(h)))
        return cls(_handle=h)

    @staticmethod
    def save_bin(path, kind=SCENE_HEIGHTFIELD_BOXES, seed=0, mesh_count=0):
        """Writes the seeded scene of HostScene(kind, seed, mesh_count) as a .bin file."""
        _check_host(load_host().hvvr_host_scene_save_bin(c_int(kind), c_uint64(seed), c_uint32(mesh_count),
                                                         os.fsencode(path)))

    @classmethod
    def from_meshes(cls, meshes, materials, lights=None):
        """meshes: list of (verts[vertex dtype], tri_indices[n,3] local, tri_material[n], transform16 or None)."""
        lib = load_host()
        vstart = np.zeros(len(meshes) + 1, np.uint32)
        tstart = np.zeros(len(meshes) + 1, np.uint32)
        for i, m in enumerate(meshes):
            vstart[i + 1] = vstart[i] + len(m[0])
            tstart[i + 1] = tstart[i] + len(m[1])
        verts = np.concatenate([m[0] for m in meshes]).astype(DTYPES["vertex"])
        tris = np.ascontiguousarray(np.concatenate([np.asarray(m[1], np.uint32) for m in meshes]).reshape(-1, 3))
        mats = np.ascontiguousarray(np.concatenate([np.asarray(m[2], np.uint32) for m in meshes]))
        xf = np.stack([np.asarray(m[3] if m[3] is not None else np.eye(4, dtype=np.float32).reshape(16),
                                  np.float32).reshape(16) for m in meshes])
        materials = np.ascontiguousarray(materials, DTYPES["material"])
        lights = np.zeros(0, DTYPES["light"]) if lights is None else np.ascontiguousarray(lights, DTYPES["light"])
        h = c_void_p()
        _check_host(lib.hvvr_host_scene_create_from_meshes(
            c_uint32(len(meshes)), _ptr(vstart), _ptr(verts), _ptr(tstart), _ptr(tris), _ptr(mats), _ptr(xf),
            c_uint32(len(materials)), _ptr(materials), c_uint32(len(lights)), _ptr(lights), byref(h)))
        return cls(_handle=h)


class HostView:
    """Camera-space sample hierarchy: beams, tile/block frusta, remap."""

    def __init__(self, preset=VIEW_PRESET_HEIGHTFIELD, eye=0, width=0, height=0, _handle=None):
        lib = load_host()
        h = c_void_p()
        if _handle is None:
            _check_host(lib.hvvr_host_view_create_preset(c_int(preset), c_uint32(eye), c_uint32(width),
                                                         c_uint32(height), byref(h)))
        else:
            h = _handle
        counts = np.zeros(7, np.uint64)
        _check_host(lib.hvvr_host_view_counts(h, _ptr(counts)))
        (self.valid_sample_count, self.sample_count, self.block_count, self.width, self.height, self.split,
         self.tile_count) = (int(c) for c in counts)
        self.beams = np.zeros(self.sample_count, DTYPES["beam"])
        self.tile_frusta = np.zeros(self.tile_count, DTYPES["sfrustum"])
        self.block_frusta = np.zeros(self.block_count, DTYPES["sfrustum"])
        self.remap = np.zeros(self.width * self.height * self.split, np.int32)
        for idx, arr in enumerate([self.beams, self.tile_frusta, self.block_frusta, self.remap]):
            if arr.size:
                _check_host(lib.hvvr_host_view_copy(h, c_int(idx), _ptr(arr), c_uint64(arr.nbytes)))
        self.camera_to_world = np.zeros(16, np.float32)
        self.lens = np.zeros(2, np.float32)
        _check_host(lib.hvvr_host_view_camera(h, _ptr(self.camera_to_world), _ptr(self.lens)))
        # sample space: sampleToCamera (column-major 3x3) and the sample-position bounds
        self.sample_to_camera = np.zeros(9, np.float32)
        self.sample_bounds = np.zeros(4, np.float32)
        _check_host(lib.hvvr_host_view_sample_space(h, _ptr(self.sample_to_camera), _ptr(self.sample_bounds)))
        self.handle = h  # kept for the serving loop (hvvr_server_add_shape)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                load_host().hvvr_host_view_destroy(h)
            except Exception:
                pass
            self.handle = None

    @classmethod
    def from_samples(cls, samples, width, height, viewport, split=1, lens_radius=0.0, focal_distance=1.0,
                     spherical=None):
        """spherical = (fovXDegrees, fovYDegrees): rays from a spherical section
        (Camera::setSphericalWarpSettings) instead of the viewport's perspective."""
        lib = load_host()
        samples = np.ascontiguousarray(samples, DTYPES["sample"])
        vp = np.asarray(viewport, np.float32)
        fx, fy = spherical if spherical is not None else (0.0, 0.0)
        h = c_void_p()
        _check_host(lib.hvvr_host_view_create_spherical(
            _ptr(samples), c_uint32(len(samples)), c_uint32(split), c_uint32(width), c_uint32(height), _ptr(vp),
            c_float(fx), c_float(fy), c_float(lens_radius), c_float(focal_distance), byref(h)))
        return cls(_handle=h)


class HostFoveated:
    """Eye-space polar foveated samples of a FoveatedSamplePattern (raycaster_spec.h:29-47):
    clustered beams, eye-space tile/block frusta and each sample's polar texel."""

    # (degreeTrackingError, maxFOVDegrees, marSlope, fovealMARDegrees)
    DEFAULT = (0.5, 110.0, 0.022, 1.0 / 60.0)
    FEB2017 = (7.0, 90.0, 0.015, 1.0 / 20.0)

    def __init__(self, pattern=DEFAULT):
        lib = load_host()
        h = c_void_p()
        pat = np.asarray(pattern, np.float32)
        _check_host(lib.hvvr_host_foveated_create(_ptr(pat), byref(h)))
        try:
            counts = np.zeros(6, np.uint64)
            _check_host(lib.hvvr_host_foveated_counts(h, _ptr(counts)))
            (self.valid_sample_count, self.sample_count, self.block_count, self.samples_per_ring, self.ring_count,
             self.tile_count) = (int(c) for c in counts)
            self.beams = np.zeros(self.sample_count, DTYPES["beam"])
            self.tile_frusta = np.zeros(self.tile_count, DTYPES["sfrustum"])
            self.block_frusta = np.zeros(self.block_count, DTYPES["sfrustum"])
            self.polar_remap = np.zeros((self.valid_sample_count, 2), np.uint32)
            for idx, arr in enumerate([self.beams, self.tile_frusta, self.block_frusta, self.polar_remap]):
                _check_host(lib.hvvr_host_foveated_copy(h, c_int(idx), _ptr(arr), c_uint64(arr.nbytes)))
            self.emap = np.zeros(6, np.float32)
            mx = c_float(0.0)
            _check_host(lib.hvvr_host_foveated_params(h, _ptr(self.emap), byref(mx)))
            self.max_eccentricity = float(mx.value)
        except Exception:
            lib.hvvr_host_foveated_destroy(h)
            raise
        self.handle = h  # kept for the serving loop (hvvr_server_add_foveated_shape)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                load_host().hvvr_host_foveated_destroy(h)
            except Exception:
                pass
            self.handle = None


def foveated_history():
    """Previous-frame state of a foveated camera: worldToEyePrevious (zero) and
    eyePreviousToSamplePrevious (identity), camera.h:121-122."""
    h = np.zeros(25, np.float32)
    h[16] = h[20] = h[24] = 1.0
    return h


def foveated_frame(camera_to_world, eye_dir, sample_to_camera, sample_bounds, history):
    """One frame's foveated transforms (render.cpp:454-488); advances `history` in place."""
    lib = load_host()
    m = np.ascontiguousarray(camera_to_world, np.float32).reshape(16)
    e = np.asarray(eye_dir, np.float32)
    s2c = np.ascontiguousarray(sample_to_camera, np.float32).reshape(9)
    b = np.ascontiguousarray(sample_bounds, np.float32).reshape(4)
    out = np.zeros(57, np.float32)
    assert history.dtype == np.float32 and history.size == 25 and history.flags.c_contiguous
    _check_host(lib.hvvr_host_foveated_frame(_ptr(m), _ptr(e), _ptr(s2c), _ptr(b), _ptr(history), _ptr(out)))
    return dict(eye_to_world=out[:16].copy(), cull_planes=out[16:32].copy(), sample_to_eye=out[32:41].copy(),
                eye_to_sample_previous=out[41:57].copy())


def preset_params(preset, eye=0, width=0, height=0):
    """(viewport lower x y upper x y, (barrel k1, k2), (lens radius, focal distance)) of a view preset"""
    p = np.zeros(8, np.float32)
    _check_host(load_host().hvvr_host_view_preset_params(c_int(preset), c_uint32(eye), c_uint32(width),
                                                          c_uint32(height), _ptr(p)))
    return p[:4].copy(), (float(p[4]), float(p[5])), (float(p[6]), float(p[7]))


def subsample_offsets(msaa_rate, spin=0.0):
    out = np.zeros((msaa_rate, 2), np.float32)
    _check_host(load_host().hvvr_host_subsample_offsets(c_uint32(msaa_rate), c_float(spin), _ptr(out)))
    return out


def lens_table(msaa_rate):
    """Depth-of-field lens-position table: [4 rows][msaa_rate][128 samples] x (u, v)."""
    out = np.zeros((4 * 128 * msaa_rate, 2), np.float32)
    _check_host(load_host().hvvr_host_lens_table(c_uint32(msaa_rate), _ptr(out)))
    return out


def grid_samples(width, height, k1=0.0, k2=0.0):
    out = np.zeros(width * height, DTYPES["sample"])
    _check_host(load_host().hvvr_host_grid_samples(c_uint32(width), c_uint32(height), c_float(k1), c_float(k2),
                                                   _ptr(out)))
    return out


def transform_matrix(translation, rotation_xyzw, scale=1.0):
    out = np.zeros(16, np.float32)
    t = np.asarray(translation, np.float32)
    r = np.asarray(rotation_xyzw, np.float32)
    _check_host(load_host().hvvr_host_transform_matrix(_ptr(t), _ptr(r), c_float(scale), _ptr(out)))
    return out


# ---------------------------------------------------------------------------------------
# GPU C ABI
# ---------------------------------------------------------------------------------------
class _SceneDesc(ctypes.Structure):
    _fields_ = [("vertices", c_void_p), ("vertexToTransform", c_void_p), ("vertexCount", c_uint32),
                ("triangles", c_void_p), ("triangleCount", c_uint32), ("nodes", c_void_p), ("nodeCount", c_uint32),
                ("materials", c_void_p), ("materialCount", c_uint32), ("transforms", c_void_p),
                ("transformCount", c_uint32)]


class _TextureDesc(ctypes.Structure):
    _fields_ = [("format", c_uint32), ("width", c_uint32), ("height", c_uint32), ("strideElements", c_uint32),
                ("data", c_void_p)]


# TextureFormat (include/hvvr/raycaster.h) -> (code, numpy dtype of one channel, channels)
TEXTURE_FORMATS = {"srgb8": (1, np.uint8, 4), "unorm8": (2, np.uint8, 4), "unorm16": (3, np.uint16, 4),
                   "float4": (4, np.float32, 4), "half4": (5, np.float16, 4), "r11g11b10": (6, np.uint32, 1),
                   "r32": (7, np.float32, 1)}


class _ViewDesc(ctypes.Structure):
    _fields_ = [("msaaRate", c_uint32), ("sampleCount", c_uint32), ("validSampleCount", c_uint32),
                ("beams", c_void_p), ("tileFrusta", c_void_p), ("blockFrusta", c_void_p), ("remap", c_void_p),
                ("width", c_uint32), ("height", c_uint32), ("split", c_uint32), ("lensRadius", c_float),
                ("focalDistance", c_float)]


class ViewStats(ctypes.Structure):
    _fields_ = [("listEntries", c_uint64), ("survivorPairs", c_uint64), ("occupiedTiles", c_uint32),
                ("tileCount", c_uint32), ("maxTileStack", c_uint32), ("overflow", c_uint32),
                ("nodeVisits", c_uint64), ("largeTiles", c_uint32), ("hugeTiles", c_uint32),
                ("giantTiles", c_uint32), ("walkTiles", c_uint32), ("testedPairs", c_uint64),
                ("colossalTiles", c_uint32)]


class GpuContext:
    def __init__(self, device=0):
        self.lib = load_gpu()
        self.handle = c_void_p()
        _check_gpu(self.lib.hvvr_gpu_context_create(c_int(device), byref(self.handle)))

    def synchronize(self):
        _check_gpu(self.lib.hvvr_gpu_context_synchronize(self.handle))

    def set_profiling(self, enable):
        """True / 1: per-stage events and the tile-list kernel's dispatch timing; 2: only the
        dispatch timing (no packet added to the stream); False / 0: off."""
        level = (1 if enable else 0) if isinstance(enable, bool) else int(enable)
        _check_gpu(self.lib.hvvr_gpu_context_set_profiling(self.handle, c_int(level)))

    def close(self):
        if self.handle:
            self.lib.hvvr_gpu_context_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuScene:
    def __init__(self, ctx: GpuContext, scene: HostScene):
        self.ctx = ctx
        self.lib = ctx.lib
        self.scene = scene
        self._keep = [np.ascontiguousarray(a) for a in (scene.verts, scene.vertexToTransform, scene.triShade,
                                                        scene.nodes, scene.materials, scene.transforms)]
        v, vt, ts, nd, mt, xf = self._keep
        d = _SceneDesc(_ptr(v), _ptr(vt), len(v), _ptr(ts), len(ts), _ptr(nd), len(nd), _ptr(mt), len(mt), _ptr(xf),
                       len(xf))
        self.handle = c_void_p()
        _check_gpu(self.lib.hvvr_gpu_scene_create(ctx.handle, byref(d), byref(self.handle)))
        self.set_lights(scene.lights)

    def set_lights(self, lights):
        lights = np.ascontiguousarray(lights, DTYPES["light"])
        dropped = c_uint32(0)
        _check_gpu(self.lib.hvvr_gpu_scene_set_lights(self.handle, _ptr(lights), c_uint32(len(lights)),
                                                      byref(dropped)))
        return dropped.value

    def set_transforms(self, transforms):
        t = np.ascontiguousarray(transforms, np.float32).reshape(-1, 16)
        _check_gpu(self.lib.hvvr_gpu_scene_set_transforms(self.handle, _ptr(t), c_uint32(len(t))))

    def set_vertices(self, first, verts):
        v = np.ascontiguousarray(verts, DTYPES["vertex"])
        _check_gpu(self.lib.hvvr_gpu_scene_set_vertices(self.handle, c_uint32(first), c_uint32(len(v)), _ptr(v)))

    def update(self):
        _check_gpu(self.lib.hvvr_gpu_scene_update(self.handle))

    def read_nodes(self):
        out = np.zeros(len(self.scene.nodes), DTYPES["node"])
        _check_gpu(self.lib.hvvr_gpu_scene_read_nodes(self.handle, _ptr(out)))
        return out

    def read_triangles(self):
        out = np.zeros(len(self.scene.triShade), DTYPES["tri_intersect"])
        _check_gpu(self.lib.hvvr_gpu_scene_read_triangles(self.handle, _ptr(out)))
        return out

    def read_vertices(self):
        out = np.zeros(len(self.scene.verts), DTYPES["vertex"])
        _check_gpu(self.lib.hvvr_gpu_scene_read_vertices(self.handle, _ptr(out)))
        return out

    def set_textures(self, textures):
        """Replace the scene's textures with [(format, texels), ...], format a key of
        TEXTURE_FORMATS and texels [h, w, 4] ([h, w] for r11g11b10 and r32); texture i is
        the index materials name. Mip chains are built on the device."""
        keep, descs = [], (_TextureDesc * max(len(textures), 1))()
        for i, (fmt, data) in enumerate(textures):
            code, dt, ch = TEXTURE_FORMATS[fmt]
            a = np.ascontiguousarray(data, dt)
            if a.ndim != (3 if ch > 1 else 2) or (ch > 1 and a.shape[2] != ch):
                raise ValueError(f"texture {i}: texels of format {fmt} are [h, w{', 4' if ch > 1 else ''}]")
            keep.append(a)
            descs[i] = _TextureDesc(code, a.shape[1], a.shape[0], a.shape[1], _ptr(a))
        _check_gpu(self.lib.hvvr_gpu_scene_set_textures(self.handle, c_uint32(len(textures)), descs))

    def texture_info(self, index):
        """(width, height, levels, stored kind, chain bytes) of texture `index`"""
        info = (c_uint32 * 4)()
        nbytes = c_uint64(0)
        _check_gpu(self.lib.hvvr_gpu_scene_texture_info(self.handle, c_uint32(index), info, byref(nbytes)))
        return info[0], info[1], info[2], info[3], nbytes.value

    def read_texture(self, index):
        """the stored mip chain of texture `index` as bytes"""
        out = np.zeros(self.texture_info(index)[4], np.uint8)
        _check_gpu(self.lib.hvvr_gpu_scene_read_texture(self.handle, c_uint32(index), _ptr(out)))
        return out

    def close(self):
        if self.handle:
            self.lib.hvvr_gpu_scene_destroy(self.handle)
            self.handle = c_void_p()


def render_views(scene, views, cameras=None, jitters=None):
    """Several views of one scene in one submission (hvvr_gpu_render_views): independent
    views run concurrently on the context's view streams."""
    n = len(views)
    if n == 0:
        return
    lib = views[0].lib
    ms = np.stack([np.ascontiguousarray(v.view.camera_to_world if cameras is None or cameras[i] is None
                                        else cameras[i], np.float32).reshape(16) for i, v in enumerate(views)])
    js = None if jitters is None else np.ascontiguousarray(jitters, np.float32).reshape(n, 2)
    handles = (c_void_p * n)(*[v.handle.value for v in views])
    _check_gpu(lib.hvvr_gpu_render_views(scene.handle, c_uint32(n), handles, _ptr(ms), _ptr(js)))


class LaunchShape(ctypes.Structure):
    _fields_ = [("largeGroups", c_uint32), ("hugeGroups", c_uint32), ("giantGroups", c_uint32),
                ("walkGroups", c_uint32), ("xcdSwizzle", c_uint32), ("streams", c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GpuBatch:
    """Many views of one scene, each kernel of the frame launched once for all of them
    (hvvr_gpu_batch_*); optionally captured as a HIP graph and replayed."""

    def __init__(self, ctx: GpuContext, views):
        self.ctx = ctx
        self.lib = ctx.lib
        self.views = list(views)
        handles = (c_void_p * len(self.views))(*[v.handle.value for v in self.views])
        self.handle = c_void_p()
        _check_gpu(self.lib.hvvr_gpu_batch_create(ctx.handle, handles, c_uint32(len(self.views)),
                                                  byref(self.handle)))

    def set_graph(self, enable):
        _check_gpu(self.lib.hvvr_gpu_batch_set_graph(self.handle, c_int(1 if enable else 0)))

    def set_refit(self, every_frame):
        _check_gpu(self.lib.hvvr_gpu_batch_set_refit(self.handle, c_int(1 if every_frame else 0)))

    def set_launch_shape(self, shape: LaunchShape):
        _check_gpu(self.lib.hvvr_gpu_batch_set_launch_shape(self.handle, byref(shape)))

    def launch_shape(self):
        sh = LaunchShape()
        _check_gpu(self.lib.hvvr_gpu_batch_get_launch_shape(self.handle, byref(sh)))
        return sh

    def _cameras(self, cameras):
        return np.ascontiguousarray(np.stack([
            np.asarray(v.view.camera_to_world if cameras is None or cameras[i] is None else cameras[i],
                       np.float32).reshape(16) for i, v in enumerate(self.views)]), np.float32)

    def render(self, scene: GpuScene, cameras=None):
        m = self._cameras(cameras)
        _check_gpu(self.lib.hvvr_gpu_batch_render(self.handle, scene.handle, _ptr(m)))

    def synchronize(self):
        ms = ctypes.c_float()
        _check_gpu(self.lib.hvvr_gpu_batch_synchronize(self.handle, byref(ms)))
        return ms.value

    def autotune(self, scene: GpuScene, cameras=None, frames_per_trial=5):
        m = self._cameras(cameras)
        best, ms = LaunchShape(), ctypes.c_float()
        _check_gpu(self.lib.hvvr_gpu_batch_autotune(self.handle, scene.handle, _ptr(m), c_uint32(frames_per_trial),
                                                    byref(best), byref(ms)))
        return best, ms.value

    def close(self):
        if self.handle:
            self.lib.hvvr_gpu_batch_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _FovDesc(ctypes.Structure):
    _fields_ = [("polarRemap", c_void_p), ("samplesPerRing", c_uint32), ("ringCount", c_uint32),
                ("eccentricityMap", c_float * 6), ("maxEccentricity", c_float), ("taaAlpha", c_float),
                ("taaStddevMultiplier", c_float), ("contrastEnhancement", c_int), ("contrastFe", c_float)]


class _FovFrame(ctypes.Structure):
    _fields_ = [("eyeToWorld", c_float * 16), ("cullPlanes", c_float * 16), ("sampleToEye", c_float * 9),
                ("eyeToSamplePrevious", c_float * 16)]


FOVEATED_BUFFERS = {"polar_color": 0, "polar_depth": 1, "moment1": 2, "moment2": 3, "filtered": 4, "previous": 5,
                    "blur1": 6, "blur2": 7, "tmax": 8}


class GpuView:
    def __init__(self, ctx: GpuContext, view: HostView, msaa_rate=1):
        self.ctx = ctx
        self.lib = ctx.lib
        self.view = view
        self.msaa_rate = msaa_rate
        self._keep = [np.ascontiguousarray(a) for a in (view.beams, view.tile_frusta, view.block_frusta, view.remap)]
        b, tf, bf, rm = self._keep
        d = _ViewDesc(msaa_rate, view.sample_count, view.valid_sample_count, _ptr(b), _ptr(tf), _ptr(bf), _ptr(rm),
                      view.width, view.height, view.split, float(view.lens[0]), float(view.lens[1]))
        self.handle = c_void_p()
        _check_gpu(self.lib.hvvr_gpu_view_create(ctx.handle, byref(d), byref(self.handle)))

    def render(self, scene: GpuScene, camera_to_world=None, jitter=(0.0, 0.0)):
        m = np.ascontiguousarray(self.view.camera_to_world if camera_to_world is None else camera_to_world,
                                 np.float32).reshape(16)
        j = np.asarray(jitter, np.float32)
        _check_gpu(self.lib.hvvr_gpu_view_render(self.handle, scene.handle, _ptr(m), _ptr(j)))

    def render_shared_scratch(self, scene: GpuScene, camera_to_world=None, jitter=(0.0, 0.0)):
        """render() with the tile-list scratch in the library's per-GPU shared array
        (hvvr_gpu_view_render_shared_scratch): allocates nothing; calls must not overlap."""
        m = np.ascontiguousarray(self.view.camera_to_world if camera_to_world is None else camera_to_world,
                                 np.float32).reshape(16)
        j = np.asarray(jitter, np.float32)
        _check_gpu(self.lib.hvvr_gpu_view_render_shared_scratch(self.handle, scene.handle, _ptr(m), _ptr(j)))

    def set_traversal_stack_limit(self, entries):
        _check_gpu(self.lib.hvvr_gpu_view_set_traversal_stack_limit(self.handle, c_uint32(entries)))

    def set_shading_precision(self, bits):
        _check_gpu(self.lib.hvvr_gpu_view_set_shading_precision(self.handle, c_int(bits)))

    def list_demand(self):
        """(entries the last frame's triangle lists needed, entries the view holds, overflowed)"""
        need, cap, of = c_uint64(), c_uint64(), c_int()
        _check_gpu(self.lib.hvvr_gpu_view_list_demand(self.handle, byref(need), byref(cap), byref(of)))
        return need.value, cap.value, bool(of.value)

    def reserve_lists(self, entries):
        _check_gpu(self.lib.hvvr_gpu_view_reserve_lists(self.handle, c_uint64(entries)))

    def set_hit_output(self, enable):
        _check_gpu(self.lib.hvvr_gpu_view_set_hit_output(self.handle, c_int(1 if enable else 0)))

    def read_hits(self):
        """[sample_count, msaa_rate, 4] float32: (t, u, v, w) of each subsample's nearest hit."""
        out = np.zeros((self.view.sample_count, self.msaa_rate, 4), np.float32)
        _check_gpu(self.lib.hvvr_gpu_view_read_hits(self.handle, _ptr(out)))
        return out

    def set_jitter_spin(self, enable):
        _check_gpu(self.lib.hvvr_gpu_view_set_jitter_spin(self.handle, c_int(1 if enable else 0)))

    def read_image(self):
        out = np.zeros((self.view.height, self.view.width), np.uint32)
        _check_gpu(self.lib.hvvr_gpu_view_read_image(self.handle, _ptr(out), c_uint32(self.view.width)))
        return out

    def read_sample_colors(self):
        out = np.zeros(self.view.sample_count, np.uint32)
        _check_gpu(self.lib.hvvr_gpu_view_read_sample_colors(self.handle, _ptr(out)))
        return out

    def read_gbuffer(self):
        n = self.view.sample_count * self.msaa_rate
        tri = np.zeros(n, np.uint32)
        mask = np.zeros(n, np.uint32)
        _check_gpu(self.lib.hvvr_gpu_view_read_gbuffer(self.handle, _ptr(tri), _ptr(mask)))
        return tri.reshape(-1, self.msaa_rate), mask.reshape(-1, self.msaa_rate)

    def read_tile_lists(self):
        counts = np.zeros(self.view.tile_count, np.uint32)
        total = c_uint64(0)
        _check_gpu(self.lib.hvvr_gpu_view_read_tile_lists(self.handle, _ptr(counts), None, c_uint64(0),
                                                          byref(total)))
        lists = np.zeros(max(total.value, 1), np.uint32)
        _check_gpu(self.lib.hvvr_gpu_view_read_tile_lists(self.handle, _ptr(counts), _ptr(lists),
                                                          c_uint64(len(lists)), byref(total)))
        return counts, lists[: total.value]

    def stats(self):
        st = ViewStats()
        _check_gpu(self.lib.hvvr_gpu_view_get_stats(self.handle, byref(st)))
        return st

    def stage_times(self):
        out = np.zeros(len(STAGES), np.float32)
        n = _check_gpu(self.lib.hvvr_gpu_view_stage_times(self.handle, _ptr(out), c_int(len(STAGES))))
        return out[:n]

    def tile_list_kernel_ms(self):
        """GPU ms of the standard tile-list kernel in the last profiled frame (dispatch events)."""
        ms = ctypes.c_float(0.0)
        _check_gpu(self.lib.hvvr_gpu_view_tile_list_kernel_time(self.handle, byref(ms)))
        return float(ms.value)

    def close(self):
        if self.handle:
            self.lib.hvvr_gpu_view_destroy(self.handle)
            self.handle = c_void_p()


class _SetupDesc(ctypes.Structure):
    _fields_ = [("msaaRate", c_uint32), ("width", c_uint32), ("height", c_uint32), ("split", c_uint32),
                ("sampleToCamera", c_float * 9), ("lensRadius", c_float), ("focalDistance", c_float),
                ("fovXDegrees", c_float), ("fovYDegrees", c_float)]


class _DeviceSetupShape:
    """The host-side description of a view whose sample hierarchy the device built: its
    counts, and its beams, frusta and remap read back (for comparisons)."""

    def __init__(self, gv, sample_to_camera, lens, camera_to_world, bounds):
        shape = np.zeros(8, np.uint32)
        _check_gpu(gv.lib.hvvr_gpu_view_get_shape(gv.handle, _ptr(shape)))
        (_, self.sample_count, self.valid_sample_count, self.tile_count, self.block_count, self.width, self.height,
         self.split) = (int(x) for x in shape)
        self.beams = np.zeros(self.sample_count, DTYPES["beam"])
        self.tile_frusta = np.zeros(self.tile_count, DTYPES["sfrustum"])
        self.block_frusta = np.zeros(self.block_count, DTYPES["sfrustum"])
        self.remap = np.zeros(self.width * self.height * self.split, np.int32)
        _check_gpu(gv.lib.hvvr_gpu_view_read_sample_setup(gv.handle, _ptr(self.beams), _ptr(self.tile_frusta),
                                                          _ptr(self.block_frusta), _ptr(self.remap)))
        self.sample_to_camera = np.asarray(sample_to_camera, np.float32)
        self.lens = np.asarray(lens, np.float32)
        self.camera_to_world = np.asarray(camera_to_world, np.float32).reshape(16)
        self.sample_bounds = bounds


class GpuDeviceSetupView(GpuView):
    """A view whose sample hierarchy is built on the device (hvvr_gpu_view_create_grid /
    _create_from_samples): Camera::setSamples without the host setup."""

    def __init__(self, ctx: GpuContext, width, height, msaa_rate=1, viewport=(-1.0, -1.0, 1.0, 1.0), samples=None,
                 split=1, barrel=(0.0, 0.0), lens=(0.0, 1.0), spherical=None, camera_to_world=None):
        self.ctx = ctx
        self.lib = ctx.lib
        self.msaa_rate = msaa_rate
        s2c = np.zeros(9, np.float32)
        _check_host(load_host().hvvr_host_sample_to_camera(_ptr(np.asarray(viewport, np.float32)), _ptr(s2c)))
        fx, fy = spherical if spherical is not None else (0.0, 0.0)
        d = _SetupDesc(msaa_rate, width, height, split, (c_float * 9)(*[float(x) for x in s2c]), float(lens[0]),
                       float(lens[1]), float(fx), float(fy))
        self.handle = c_void_p()
        bounds = np.zeros(4, np.float32)
        if samples is None:
            _check_gpu(self.lib.hvvr_gpu_view_create_grid(ctx.handle, byref(d), c_float(barrel[0]), c_float(barrel[1]),
                                                          byref(self.handle), _ptr(bounds)))
        else:
            samples = np.ascontiguousarray(samples, DTYPES["sample"])
            _check_gpu(self.lib.hvvr_gpu_view_create_from_samples(ctx.handle, byref(d), _ptr(samples),
                                                                  c_uint32(len(samples)), byref(self.handle),
                                                                  _ptr(bounds)))
        m = np.eye(4, dtype=np.float32).reshape(16) if camera_to_world is None else camera_to_world
        self.view = _DeviceSetupShape(self, s2c, lens, m, bounds)


class _FoveatedViewShape:
    def __init__(self, fov, width, height, lens):
        self.sample_count, self.valid_sample_count = fov.sample_count, fov.valid_sample_count
        self.tile_count, self.block_count = fov.tile_count, fov.block_count
        self.width, self.height, self.split = width, height, 1
        self.beams, self.tile_frusta, self.block_frusta = fov.beams, fov.tile_frusta, fov.block_frusta
        self.camera_to_world = np.eye(4, dtype=np.float32).reshape(16)
        self.lens = np.asarray(lens, np.float32)


class GpuFoveatedView(GpuView):
    """A polar foveated view: eye-space polar samples traced each frame and reconstructed
    into a width x height screen image (hvvr_gpu_view_set_foveated)."""

    def __init__(self, ctx: GpuContext, fov: HostFoveated, width, height, msaa_rate=1, lens_radius=0.0,
                 focal_distance=1.0, taa_alpha=0.05, taa_stddev=1.0, contrast=True, contrast_fe=1.0):
        self.ctx = ctx
        self.lib = ctx.lib
        self.fov = fov
        self.view = _FoveatedViewShape(fov, width, height, (lens_radius, focal_distance))
        self.msaa_rate = msaa_rate
        self._keep = [np.ascontiguousarray(a) for a in (fov.beams, fov.tile_frusta, fov.block_frusta)]
        b, tf, bf = self._keep
        d = _ViewDesc(msaa_rate, fov.sample_count, fov.valid_sample_count, _ptr(b), _ptr(tf), _ptr(bf), None, width,
                      height, 1, float(lens_radius), float(focal_distance))
        self.handle = c_void_p()
        _check_gpu(self.lib.hvvr_gpu_view_create(ctx.handle, byref(d), byref(self.handle)))
        remap = np.ascontiguousarray(fov.polar_remap, np.uint32)
        fd = _FovDesc(_ptr(remap), fov.samples_per_ring, fov.ring_count, (c_float * 6)(*fov.emap),
                      fov.max_eccentricity, taa_alpha, taa_stddev, 1 if contrast else 0, contrast_fe)
        _check_gpu(self.lib.hvvr_gpu_view_set_foveated(self.handle, byref(fd)))

    def render_foveated(self, scene: GpuScene, frame, jitter=(0.0, 0.0)):
        fr = _FovFrame((c_float * 16)(*frame["eye_to_world"]), (c_float * 16)(*frame["cull_planes"]),
                       (c_float * 9)(*frame["sample_to_eye"]), (c_float * 16)(*frame["eye_to_sample_previous"]))
        j = np.asarray(jitter, np.float32)
        _check_gpu(self.lib.hvvr_gpu_view_render_foveated(self.handle, scene.handle, byref(fr), _ptr(j)))

    def read_foveated(self, name):
        f, v = self.fov, self.view
        polar = (f.ring_count, f.samples_per_ring)
        shape, dt = {"polar_color": (polar, np.uint32), "polar_depth": (polar, np.float32),
                     "moment1": (polar, np.uint64), "moment2": (polar, np.uint64),
                     "tmax": ((f.sample_count,), np.float32)}.get(name, ((v.height, v.width), np.uint32))
        out = np.zeros(shape, dt)
        _check_gpu(self.lib.hvvr_gpu_view_read_foveated(self.handle, c_int(FOVEATED_BUFFERS[name]), _ptr(out),
                                                        c_uint64(out.nbytes)))
        return out


# ---------------------------------------------------------------------------------------
# serving loop (include/hvvr_server.h)
# ---------------------------------------------------------------------------------------
class _ServerDesc(ctypes.Structure):
    _fields_ = [("deviceBudgetBytes", c_uint64), ("maxInFlight", c_uint32), ("shadingPrecision", c_uint32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", c_uint64), ("completed", c_uint64), ("queued", c_uint64), ("inFlight", c_uint64),
                ("deviceBytes", c_uint64), ("peakDeviceBytes", c_uint64), ("viewsCreated", c_uint32),
                ("viewsEvicted", c_uint32), ("peakInFlight", c_uint32), ("p50Ms", ctypes.c_double),
                ("p99Ms", ctypes.c_double), ("meanMs", ctypes.c_double), ("failed", c_uint64),
                ("listGrowths", c_uint32)]


class Server:
    """Frame-request serving loop: resident scenes, view shapes admitted by HBM footprint,
    requests overlapping on the device (hvvr_server_*)."""

    def __init__(self, ctx: GpuContext, budget_bytes, max_in_flight=8, shading_bits=32):
        self.ctx = ctx
        self.lib = load_host()
        self.lib.hvvr_server_create.argtypes = [c_void_p, c_void_p, c_void_p]
        self.handle = c_void_p()
        d = _ServerDesc(int(budget_bytes), int(max_in_flight), int(shading_bits))
        _check_host(self.lib.hvvr_server_create(ctx.handle, byref(d), byref(self.handle)))
        self.shapes = []
        self._images = {}

    def add_shape(self, view: HostView, msaa_rate):
        sid = c_uint32()
        _check_host(self.lib.hvvr_server_add_shape(self.handle, view.handle, c_uint32(msaa_rate), byref(sid)))
        self.shapes.append((view.height, view.width))
        return sid.value

    def add_grid_shape(self, width, height, msaa_rate, viewport, barrel=(0.0, 0.0), lens=(0.0, 1.0)):
        """A grid shape whose views set up their samples on the device (hvvr_server_add_grid_shape)."""
        s2c = np.zeros(9, np.float32)
        _check_host(load_host().hvvr_host_sample_to_camera(_ptr(np.asarray(viewport, np.float32)), _ptr(s2c)))
        d = _SetupDesc(msaa_rate, width, height, 1, (c_float * 9)(*[float(x) for x in s2c]), float(lens[0]),
                       float(lens[1]), 0.0, 0.0)
        sid = c_uint32()
        _check_host(self.lib.hvvr_server_add_grid_shape(self.handle, byref(d), c_float(barrel[0]), c_float(barrel[1]),
                                                        byref(sid)))
        self.shapes.append((height, width))
        return sid.value

    def add_foveated_shape(self, fov: HostFoveated, screen: HostView, msaa_rate):
        sid = c_uint32()
        _check_host(self.lib.hvvr_server_add_foveated_shape(self.handle, fov.handle, screen.handle,
                                                            c_uint32(msaa_rate), byref(sid)))
        self.shapes.append((screen.height, screen.width))
        return sid.value

    def shape_bytes(self, shape):
        b = c_uint64()
        _check_host(self.lib.hvvr_server_shape_bytes(self.handle, c_uint32(shape), byref(b)))
        return b.value

    def submit(self, scene: GpuScene, shape, camera_to_world, eye_dir=None):
        """Queues a frame; returns (request id, the image array it will be written to)."""
        img = np.zeros(self.shapes[shape], np.uint32)
        m = np.ascontiguousarray(camera_to_world, np.float32).reshape(16)
        e = None if eye_dir is None else np.ascontiguousarray(eye_dir, np.float32)
        rid = c_uint64()
        _check_host(self.lib.hvvr_server_submit(self.handle, scene.handle, c_uint32(shape), _ptr(m),
                                                _ptr(e) if e is not None else None, _ptr(img), byref(rid)))
        self._images[rid.value] = img
        return rid.value, img

    def poll(self):
        return _check_host(self.lib.hvvr_server_poll(self.handle))

    def drain(self):
        _check_host(self.lib.hvvr_server_drain(self.handle))

    def done(self, rid):
        lat = ctypes.c_double()
        r = _check_host(self.lib.hvvr_server_request_done(self.handle, c_uint64(rid), byref(lat)))
        return (r == 1), lat.value

    def stats(self):
        st = ServerStats()
        _check_host(self.lib.hvvr_server_get_stats(self.handle, byref(st)))
        return st

    def close(self):
        if self.handle:
            self.lib.hvvr_server_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
