"""ctypes shim over libprt.so (include/prt.h).

This is the Python face of the render entry point: ``Scene(desc)`` is what
``SceneEnv::Load`` + ``CreateEntList`` set up (scene_env.h:357,
scene.cpp:38), ``Scene.render`` is ``Render::Run`` (render.h:1125-1455) and
``Scene.trace`` is ``intersect_scene`` over a ray batch (scene.cpp:448).
Everything runs in the HIP library; there is no CPU fallback — a missing or
unloadable library raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from .scene import (CDesc, HIT_DTYPE, PrtBatchOpts, PrtDumpCallback, PrtDumpOpts, PrtErptOpts, PrtMltOpts,
                    PrtRenderOpts, PrtSceneDesc, PrtStats, RAY_DTYPE, SceneDesc)

HERE = os.path.dirname(os.path.abspath(__file__))
# PRT_LIBRARY selects another build of libprt.so (e.g. a tuning variant)
LIB_PATH = os.environ.get("PRT_LIBRARY") or os.path.join(HERE, "_lib", "libprt.so")

_lib = None


class PrtError(RuntimeError):
    pass


def lib():
    """Load libprt.so (built by ``python -m prender_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PrtError(f"{LIB_PATH} is missing: build it with `python -m prender_amd.build`")
    # libprt and PyTorch both need libamdhip64.so.7, and the first one loaded
    # serves the whole process; load torch's first (when torch is installed)
    # so that tensors and libprt share one HIP runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    L.prt_abi_version.restype = C.c_int
    L.prt_last_error.restype = C.c_char_p
    L.prt_device_info.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.prt_scene_create.restype = C.c_void_p
    L.prt_scene_create.argtypes = [C.POINTER(PrtSceneDesc)]
    L.prt_scene_destroy.argtypes = [C.c_void_p]
    L.prt_scene_info.argtypes = [C.c_void_p] + [C.POINTER(C.c_int64)] * 4
    L.prt_trace_closest.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(PrtStats)]
    L.prt_trace_closest_device.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(PrtStats)]
    L.prt_trace_closest_device_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                                   C.POINTER(PrtStats)]
    L.prt_render.argtypes = [C.c_void_p, C.POINTER(PrtRenderOpts), C.c_void_p, C.c_void_p, C.POINTER(PrtStats)]
    L.prt_render_device.argtypes = [C.c_void_p, C.POINTER(PrtRenderOpts), C.c_void_p, C.c_void_p,
                                    C.POINTER(PrtStats)]
    L.prt_render_dump.argtypes = [C.c_void_p, C.POINTER(PrtRenderOpts), C.POINTER(PrtDumpOpts), C.c_void_p,
                                  C.c_void_p, C.POINTER(PrtStats)]
    L.prt_scene_file_image_dump.argtypes = [C.c_void_p]
    L.prt_render_erpt.argtypes = [C.c_void_p, C.POINTER(PrtErptOpts), C.c_void_p, C.POINTER(PrtStats)]
    L.prt_render_erpt_device.argtypes = [C.c_void_p, C.POINTER(PrtErptOpts), C.c_void_p, C.POINTER(PrtStats)]
    L.prt_render_mlt.argtypes = [C.c_void_p, C.POINTER(PrtMltOpts), C.c_void_p, C.POINTER(PrtStats)]
    L.prt_render_mlt_device.argtypes = [C.c_void_p, C.POINTER(PrtMltOpts), C.c_void_p, C.POINTER(PrtStats)]
    L.prt_render_batch_device.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(PrtBatchOpts),
                                          C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(PrtStats)]
    L.prt_tune_batch.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(PrtBatchOpts), C.POINTER(C.c_void_p),
                                 C.POINTER(C.c_void_p), C.POINTER(PrtBatchOpts), C.POINTER(C.c_double)]
    L.prt_set_face_albedo_device.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.prt_set_emission.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
    L.prt_hold_rr_albedo.argtypes = [C.c_void_p, C.c_int]
    L.prt_adjoint_device.argtypes = [C.c_void_p, C.POINTER(PrtRenderOpts), C.c_void_p, C.c_int, C.c_void_p,
                                     C.c_int, C.POINTER(C.c_double), C.POINTER(PrtStats)]
    L.prt_scene_file_load.restype = C.c_void_p
    L.prt_scene_file_load.argtypes = [C.c_char_p]
    L.prt_scene_file_desc.restype = C.POINTER(PrtSceneDesc)
    L.prt_scene_file_desc.argtypes = [C.c_void_p]
    L.prt_scene_file_output.restype = C.c_char_p
    L.prt_scene_file_output.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.prt_scene_file_free.argtypes = [C.c_void_p]
    L.prt_scene_file_mlt.argtypes = [C.c_void_p, C.POINTER(PrtMltOpts)]
    L.prt_eval_expression.restype = C.c_double
    L.prt_eval_expression.argtypes = [C.c_char_p, C.POINTER(C.c_int)]
    L.prt_image_to_rgb8.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_void_p]
    L.prt_write_ppm.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.c_int]
    L.prt_write_bmp.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.c_int, C.c_double]
    L.prt_write_hdr.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.c_int]
    _lib = L
    return L


EXPORTED = ["prt_abi_version", "prt_last_error", "prt_device_info", "prt_scene_create", "prt_scene_destroy",
            "prt_scene_info", "prt_trace_closest", "prt_trace_closest_device", "prt_trace_closest_device_noalloc",
            "prt_render", "prt_render_dump", "prt_scene_file_image_dump", "prt_render_device",
            "prt_render_mlt", "prt_render_mlt_device", "prt_render_erpt", "prt_render_erpt_device", "prt_render_batch_device", "prt_tune_batch", "prt_set_face_albedo_device", "prt_set_emission",
            "prt_hold_rr_albedo", "prt_adjoint_device", "prt_scene_file_load", "prt_scene_file_desc",
            "prt_scene_file_output", "prt_scene_file_mlt", "prt_scene_file_free", "prt_eval_expression", "prt_image_to_rgb8",
            "prt_write_ppm", "prt_write_bmp", "prt_write_hdr"]


def check(rc: int, what: str):
    if rc != 0:
        raise PrtError(f"{what} failed ({rc}): {lib().prt_last_error().decode()}")


def device_info():
    name = C.create_string_buffer(64)
    cu, ma, mi = C.c_int(), C.c_int(), C.c_int()
    check(lib().prt_device_info(name, 64, C.byref(cu), C.byref(ma), C.byref(mi)), "prt_device_info")
    return {"arch": name.value.decode(), "cu_count": cu.value, "major": ma.value, "minor": mi.value}


def render_opts(seed: int = 0, samples: Optional[int] = None, batch_pixels: int = 0, max_depth: int = 0,
                timing: bool = False, image_f64: bool = False, engine: int = 0) -> PrtRenderOpts:
    o = PrtRenderOpts()
    o.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    o.samples_override = int(samples or 0)
    o.batch_pixels = int(batch_pixels)
    o.max_depth = int(max_depth)
    o.timing = 1 if timing else 0
    o.image_f64 = 1 if image_f64 else 0
    o.engine = int(engine)
    return o


def mlt_opts(seed: int = 0, mutation: float = 0.0, pre_sample: float = 0.0, p_large: float = 0.0,
             mutations_per_chain: int = 0, samples: Optional[int] = None, chains: int = 0, sample_depths: int = 0,
             max_depth: int = 0, image_f64: bool = False) -> PrtMltOpts:
    o = PrtMltOpts()
    o.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    o.mutation = float(mutation)
    o.pre_sample = float(pre_sample)
    o.p_large = float(p_large)
    o.mutations_per_chain = int(mutations_per_chain)
    o.samples_override = int(samples or 0)
    o.chains = int(chains)
    o.sample_depths = int(sample_depths)
    o.max_depth = int(max_depth)
    o.image_f64 = 1 if image_f64 else 0
    return o


def erpt_opts(seed: int = 0, mutation: float = 0.0, mutate_distance: float = 0.0, samples: Optional[int] = None,
              sample_depths: int = 0, max_depth: int = 0, image_f64: bool = False) -> PrtErptOpts:
    o = PrtErptOpts()
    o.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    o.mutation = float(mutation)
    o.mutate_distance = float(mutate_distance)
    o.samples_override = int(samples or 0)
    o.sample_depths = int(sample_depths)
    o.max_depth = int(max_depth)
    o.image_f64 = 1 if image_f64 else 0
    return o


class Scene:
    """A scene uploaded to the current HIP device."""

    def __init__(self, desc: SceneDesc):
        self.desc = desc
        self._c = CDesc(desc)
        h = lib().prt_scene_create(self._c.ptr())
        if not h:
            raise PrtError("prt_scene_create failed: " + lib().prt_last_error().decode())
        self.h = C.c_void_p(h)

    def close(self):
        if getattr(self, "h", None):
            lib().prt_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def width(self):
        return self.desc.camera.width

    @property
    def height(self):
        return self.desc.camera.height

    def info(self):
        v = [C.c_int64() for _ in range(4)]
        check(lib().prt_scene_info(self.h, *[C.byref(x) for x in v]), "prt_scene_info")
        return {"nodes": v[0].value, "tris": v[1].value, "lights": v[2].value, "bvh_bytes": v[3].value}

    # ---- closest hit
    def trace(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(rays.shape[0], dtype=HIT_DTYPE)
        st = PrtStats()
        check(lib().prt_trace_closest(self.h, rays.ctypes.data, rays.shape[0], hits.ctypes.data, C.byref(st)),
              "prt_trace_closest")
        return hits, st.as_dict()

    def trace_device(self, rays_ptr: int, n: int, hits_ptr: int):
        st = PrtStats()
        check(lib().prt_trace_closest_device(self.h, C.c_void_p(rays_ptr), n, C.c_void_p(hits_ptr), C.byref(st)),
              "prt_trace_closest_device")
        return st.as_dict()

    def trace_device_noalloc(self, rays_ptr: int, n: int, hits_ptr: int, stats: bool = True):
        """prt_trace_closest_device_noalloc: the same query, allocating nothing."""
        st = PrtStats()
        check(lib().prt_trace_closest_device_noalloc(self.h, C.c_void_p(rays_ptr), n, C.c_void_p(hits_ptr),
                                                     C.byref(st) if stats else None),
              "prt_trace_closest_device_noalloc")
        return st.as_dict() if stats else None

    # ---- render
    def render(self, seed: int = 0, samples: Optional[int] = None, variance: bool = True, **kw):
        """Returns (rgb[H,W,3], var[H,W,3] or None, stats); rows in the
        reference's image order (row 0 = top, render.h:1234)."""
        H, W = self.height, self.width
        rgb = np.zeros((H, W, 3), dtype=np.float32)
        var = np.zeros((H, W, 3), dtype=np.float32) if variance else None
        st = PrtStats()
        o = render_opts(seed, samples, **kw)
        check(lib().prt_render(self.h, C.byref(o), rgb.ctypes.data, var.ctypes.data if variance else None,
                               C.byref(st)), "prt_render")
        return rgb, var, st.as_dict()

    def render_dump(self, seconds: float, callback=None, directory: Optional[str] = None,
                    gamma_offset: float = 0.0, seed: int = 0, samples: Optional[int] = None, **kw):
        """prt_render with progress images (Render::Run's periodic OutputBmp):
        ``callback(rgb[H,W,3] copy, index, fraction_done)`` and/or
        <directory>/output_W%06d.bmp between batches once ``seconds`` have
        passed.  Returns (rgb, var, stats) like render()."""
        H, W = self.height, self.width
        rgb = np.zeros((H, W, 3), dtype=np.float32)
        var = np.zeros((H, W, 3), dtype=np.float32)
        errors = []

        def trampoline(user, ptr, w, h, index, fraction):
            try:
                img = np.ctypeslib.as_array(ptr, shape=(h * w * 3,)).reshape(h, w, 3).copy()
                callback(img, int(index), float(fraction))
            except Exception as e:  # an exception must not unwind through the C frames
                errors.append(e)

        d = PrtDumpOpts()
        d.seconds = float(seconds)
        d.directory = os.fsencode(directory) if directory else None
        d.gamma_offset = float(gamma_offset)
        cb = PrtDumpCallback(trampoline) if callback is not None else PrtDumpCallback()
        d.callback = cb
        st = PrtStats()
        o = render_opts(seed, samples, **kw)
        check(lib().prt_render_dump(self.h, C.byref(o), C.byref(d), rgb.ctypes.data, var.ctypes.data, C.byref(st)),
              "prt_render_dump")
        if errors:
            raise errors[0]
        return rgb, var, st.as_dict()

    def render_device(self, rgb_ptr: int, var_ptr: Optional[int] = None, seed: int = 0,
                      samples: Optional[int] = None, **kw):
        st = PrtStats()
        o = render_opts(seed, samples, **kw)
        check(lib().prt_render_device(self.h, C.byref(o), C.c_void_p(rgb_ptr),
                                      C.c_void_p(var_ptr) if var_ptr else None, C.byref(st)),
              "prt_render_device")
        return st.as_dict()

    # ---- Metropolis light transport (Render::RunMLT, render.h:360-780)
    def render_mlt(self, seed: int = 0, **kw):
        """Returns (rgb[H,W,3], stats) of a primary-sample-space MLT render;
        keyword arguments are the prt_mlt_opts fields (mlt_opts)."""
        H, W = self.height, self.width
        rgb = np.zeros((H, W, 3), dtype=np.float32)
        st = PrtStats()
        o = mlt_opts(seed, **kw)
        check(lib().prt_render_mlt(self.h, C.byref(o), rgb.ctypes.data, C.byref(st)), "prt_render_mlt")
        return rgb, st.as_dict()

    def render_mlt_device(self, rgb_ptr: int, seed: int = 0, **kw):
        st = PrtStats()
        o = mlt_opts(seed, **kw)
        check(lib().prt_render_mlt_device(self.h, C.byref(o), C.c_void_p(rgb_ptr), C.byref(st)),
              "prt_render_mlt_device")
        return st.as_dict()

    # ---- energy redistribution path tracing (Render::RunERPT, render.h:786-1123)
    def render_erpt(self, seed: int = 0, **kw):
        """Returns (rgb[H,W,3], stats); keyword arguments are the
        prt_erpt_opts fields (erpt_opts)."""
        H, W = self.height, self.width
        rgb = np.zeros((H, W, 3), dtype=np.float32)
        st = PrtStats()
        o = erpt_opts(seed, **kw)
        check(lib().prt_render_erpt(self.h, C.byref(o), rgb.ctypes.data, C.byref(st)), "prt_render_erpt")
        return rgb, st.as_dict()

    def render_erpt_device(self, rgb_ptr: int, seed: int = 0, **kw):
        st = PrtStats()
        o = erpt_opts(seed, **kw)
        check(lib().prt_render_erpt_device(self.h, C.byref(o), C.c_void_p(rgb_ptr), C.byref(st)),
              "prt_render_erpt_device")
        return st.as_dict()

    # ---- differentiable parameters (per-face albedo, light emission)
    def set_face_albedo_device(self, entity: int, albedo_ptr: int):
        check(lib().prt_set_face_albedo_device(self.h, entity, C.c_void_p(albedo_ptr)), "prt_set_face_albedo_device")

    def set_emission(self, entity: int, emission):
        e = (C.c_double * 3)(*[float(x) for x in emission])
        check(lib().prt_set_emission(self.h, entity, e), "prt_set_emission")

    def hold_rr_albedo(self, hold: bool = True):
        """Freeze (or release) the albedo that Russian-roulette probabilities use."""
        check(lib().prt_hold_rr_albedo(self.h, 1 if hold else 0), "prt_hold_rr_albedo")

    def adjoint_device(self, dimage_ptr: int, entity: int, grad_albedo_ptr: Optional[int], light: int,
                       seed: int = 0, samples: Optional[int] = None, **kw):
        """Path-replay adjoint: writes d loss/d albedo (n_faces*3 floats) of
        `entity` to grad_albedo_ptr and returns (d loss/d emission [3] of
        `light`, stats) for the adjoint image at dimage_ptr (W*H*3 floats)."""
        st = PrtStats()
        ge = (C.c_double * 3)()
        o = render_opts(seed, samples, **kw)
        check(lib().prt_adjoint_device(self.h, C.byref(o), C.c_void_p(dimage_ptr), entity,
                                       C.c_void_p(grad_albedo_ptr) if grad_albedo_ptr else None, light, ge,
                                       C.byref(st)), "prt_adjoint_device")
        return np.array(ge[:]), st.as_dict()


# ---- batched scenes
def batch_opts(seed: int = 0, tile_samples: int = 0, inflight_tiles: int = 0, stack_lds: int = 0,
               trav_blocks_per_cu: int = 0, use_graph: bool = False, tile_stride: int = 0,
               max_depth: int = 0, timing: bool = False, persistent: int = 0,
               persist_waves: int = 0) -> PrtBatchOpts:
    o = PrtBatchOpts()
    o.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    o.tile_samples = int(tile_samples)
    o.inflight_tiles = int(inflight_tiles)
    o.stack_lds = int(stack_lds)
    o.trav_blocks_per_cu = int(trav_blocks_per_cu)
    o.use_graph = 1 if use_graph else 0
    o.tile_stride = int(tile_stride)
    o.max_depth = int(max_depth)
    o.timing = 1 if timing else 0
    o.persistent = int(persistent)  # 0 / False tile queue, 1 / True one persistent launch, 2 split by NEE
    o.persist_waves = int(persist_waves)  # persistent kernel's waves per SIMD (0 = the mode's default)
    return o


def _batch_args(scenes, rgb_ptrs, var_ptrs):
    n = len(scenes)
    if n == 0 or len(rgb_ptrs) != n or (var_ptrs is not None and len(var_ptrs) != n):
        raise PrtError("one output buffer per scene is required")
    hs = (C.c_void_p * n)(*[s.h.value for s in scenes])
    rg = (C.c_void_p * n)(*[int(p) for p in rgb_ptrs])
    vr = (C.c_void_p * n)(*[int(p) if p else None for p in var_ptrs]) if var_ptrs is not None else None
    return n, hs, rg, vr


def render_batch_device(scenes, rgb_ptrs, var_ptrs=None, opts: Optional[PrtBatchOpts] = None, **kw):
    """Renders every scene of `scenes` (Scene objects on one device) in one
    batched call into the device images rgb_ptrs[i] (W_i*H_i*3 floats)."""
    n, hs, rg, vr = _batch_args(scenes, rgb_ptrs, var_ptrs)
    o = opts if opts is not None else batch_opts(**kw)
    st = PrtStats()
    check(lib().prt_render_batch_device(hs, n, C.byref(o), rg, vr, C.byref(st)), "prt_render_batch_device")
    return st.as_dict()


def tune_batch(scenes, rgb_ptrs, var_ptrs=None, base: Optional[PrtBatchOpts] = None, **kw):
    """Runs the batch tuner; returns (best PrtBatchOpts, ns per camera sample)."""
    n, hs, rg, vr = _batch_args(scenes, rgb_ptrs, var_ptrs)
    b = base if base is not None else batch_opts(**kw)
    best = PrtBatchOpts()
    ns = C.c_double()
    check(lib().prt_tune_batch(hs, n, C.byref(b), rg, vr, C.byref(best), C.byref(ns)), "prt_tune_batch")
    return best, ns.value


# ---- scene expressions and image files (host-side, no device needed)
def eval_expression(text: str) -> float:
    """A numeric scene-file field as the reader evaluates it (Expr.cpp)."""
    st = C.c_int(0)
    v = lib().prt_eval_expression(text.encode(), C.byref(st))
    check(st.value, "prt_eval_expression")
    return v


def _image(rgb: np.ndarray):
    a = np.ascontiguousarray(rgb, dtype=np.float32)
    if a.ndim != 3 or a.shape[2] != 3:
        raise PrtError("expected an (H, W, 3) image")
    return a, a.shape[1], a.shape[0]


def to_rgb8(rgb: np.ndarray, gamma: float = 0.0) -> np.ndarray:
    """8-bit image with ppm.h's to_int (gamma 0 = 1/2.2)."""
    a, w, h = _image(rgb)
    out = np.empty((h, w, 3), dtype=np.uint8)
    check(lib().prt_image_to_rgb8(a.ctypes.data, w, h, float(gamma), out.ctypes.data), "prt_image_to_rgb8")
    return out


def write_ppm(path: str, rgb: np.ndarray):
    a, w, h = _image(rgb)
    check(lib().prt_write_ppm(os.fsencode(path), a.ctypes.data, w, h), "prt_write_ppm")


def write_bmp(path: str, rgb: np.ndarray, gamma_offset: float = 0.0):
    a, w, h = _image(rgb)
    check(lib().prt_write_bmp(os.fsencode(path), a.ctypes.data, w, h, float(gamma_offset)), "prt_write_bmp")


def write_hdr(path: str, rgb: np.ndarray):
    a, w, h = _image(rgb)
    check(lib().prt_write_hdr(os.fsencode(path), a.ctypes.data, w, h), "prt_write_hdr")
