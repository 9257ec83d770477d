"""ctypes binding of the FP64 CPU oracle (oracle/prt_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
product package (prender_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional, Tuple

import numpy as np

from prender_amd.scene import (CDesc, HIT_DTYPE, PrtHit, PrtRay, PrtSceneDesc, RAY_DTYPE,
                               SceneDesc)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++, no GPU needed)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE] + (["-B"] if force else []))
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_scene_create.restype = C.c_void_p
        L.orc_scene_create.argtypes = [C.POINTER(PrtSceneDesc)]
        L.orc_scene_destroy.argtypes = [C.c_void_p]
        L.orc_last_error.restype = C.c_char_p
        L.orc_trace_closest.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int,
                                        C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_int]
        L.orc_render.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                 C.c_void_p, C.POINTER(C.c_uint64), C.c_int]
        L.orc_radiance_sample.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_int, C.c_int,
                                          C.POINTER(C.c_double)]
        L.orc_radiance_ray.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.c_int, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_uint64)]
        L.orc_camera_rays.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_qbvh_info.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int64), C.POINTER(C.c_int)]
        L.orc_light_order.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int]
        L.orc_set_max_vertices.argtypes = [C.c_void_p, C.c_int]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_uint32)]
        L.orc_adjoint.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                  C.POINTER(C.c_double), C.c_int]
        L.orc_set_face_albedo.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_set_emission.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
        L.orc_mlt_mutate0.restype = C.c_uint32
        L.orc_mlt_mutate0.argtypes = [C.c_uint32, C.c_double]
        L.orc_render_erpt.argtypes = [C.c_void_p, C.c_uint64, C.c_double, C.c_double, C.c_int, C.c_int,
                                      C.c_void_p, C.POINTER(C.c_uint64), C.c_int]
        L.orc_render_mlt.argtypes = [C.c_void_p, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_int64,
                                     C.c_int, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_uint64), C.c_int]
        _lib = L
    return _lib


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (C.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [int(x) for x in o]


class OracleScene:
    def __init__(self, desc: SceneDesc):
        self.desc = desc
        self._c = CDesc(desc)
        h = lib().orc_scene_create(self._c.ptr())
        if not h:
            raise ValueError("oracle scene create failed: " + lib().orc_last_error().decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_max_vertices(self, cap: int):
        assert lib().orc_set_max_vertices(self.h, cap) == 0

    def trace(self, rays: np.ndarray, mode: str = "qbvh", threads: int = 0):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(rays.shape[0], dtype=HIT_DTYPE)
        nodes, tris = C.c_uint64(0), C.c_uint64(0)
        rc = lib().orc_trace_closest(self.h, rays.ctypes.data, rays.shape[0], hits.ctypes.data,
                                     0 if mode == "qbvh" else 1, C.byref(nodes), C.byref(tris), threads)
        assert rc == 0
        return hits, int(nodes.value), int(tris.value)

    def render(self, seed: int, samples: Optional[int] = None, rows: Optional[Tuple[int, int]] = None,
               threads: int = 0):
        W, H = self.desc.camera.width, self.desc.camera.height
        rgb = np.zeros((H, W, 3), dtype=np.float64)
        var = np.zeros((H, W, 3), dtype=np.float64)
        stats = (C.c_uint64 * 7)()
        r0, r1 = rows if rows is not None else (0, H)
        rc = lib().orc_render(self.h, seed, samples or 0, r0, r1, rgb.ctypes.data, var.ctypes.data,
                              stats, threads)
        assert rc == 0
        keys = ["primary", "extension", "shadow", "pickup", "nodes", "tris", "truncated"]
        return rgb, var, dict(zip(keys, [int(x) for x in stats]))

    def render_mlt(self, seed: int, mutation: float = 0.0, pre_sample: float = 0.0, p_large: float = 0.0,
                   mutations_per_chain: int = 0, samples: Optional[int] = None, chains: int = 0,
                   sample_depths: int = 0, threads: int = 0):
        """Metropolis render as prt_render_mlt runs it; returns (rgb[H,W,3]
        float64, stats)."""
        W, H = self.desc.camera.width, self.desc.camera.height
        rgb = np.zeros((H, W, 3), dtype=np.float64)
        stats = (C.c_uint64 * 6)()
        rc = lib().orc_render_mlt(self.h, seed, mutation, pre_sample, p_large, mutations_per_chain, samples or 0,
                                  chains, sample_depths, rgb.ctypes.data, stats, threads)
        assert rc == 0, rc
        keys = ["seeds", "chains", "steps", "proposals", "accepts", "lum_total"]
        return rgb, dict(zip(keys, [int(x) for x in stats]))

    def render_erpt(self, seed: int, mutation: float = 0.0, mutate_distance: float = 0.0,
                    samples: Optional[int] = None, sample_depths: int = 0, threads: int = 0):
        """Energy redistribution render as prt_render_erpt runs it; returns
        (rgb[H,W,3] float64, stats)."""
        W, H = self.desc.camera.width, self.desc.camera.height
        rgb = np.zeros((H, W, 3), dtype=np.float64)
        stats = (C.c_uint64 * 4)()
        rc = lib().orc_render_erpt(self.h, seed, mutation, mutate_distance, samples or 0, sample_depths,
                                   rgb.ctypes.data, stats, threads)
        assert rc == 0, rc
        keys = ["initial", "chains", "proposals", "accepts"]
        return rgb, dict(zip(keys, [int(x) for x in stats]))

    def radiance_sample(self, seed: int, x: int, y: int, sample: int):
        out = (C.c_double * 3)()
        lib().orc_radiance_sample(self.h, seed, x, y, sample, out)
        return np.array(out[:])

    def radiance_ray(self, org, direction, depth: int = 0, n: int = 1, seed: int = 1):
        """radiance() of one ray entering at path depth ``depth``, over n
        independent paths: (mean[3], variance of the mean[3], stats)."""
        o = (C.c_double * 3)(*[float(x) for x in org])
        d = (C.c_double * 3)(*[float(x) for x in direction])
        mean, var = (C.c_double * 3)(), (C.c_double * 3)()
        stats = (C.c_uint64 * 7)()
        rc = lib().orc_radiance_ray(self.h, seed, o, d, depth, n, mean, var, stats)
        assert rc == 0, rc
        keys = ["primary", "extension", "shadow", "pickup", "nodes", "tris", "truncated"]
        return np.array(mean[:]), np.array(var[:]), dict(zip(keys, [int(x) for x in stats]))

    def camera_rays(self) -> np.ndarray:
        W, H = self.desc.camera.width, self.desc.camera.height
        ss = self.desc.sampling.supersamples
        out = np.zeros(W * H * ss * ss, dtype=RAY_DTYPE)
        lib().orc_camera_rays(self.h, out.ctypes.data)
        return out

    def qbvh_info(self, entity: int):
        n, r, ok = C.c_int64(), C.c_int64(), C.c_int()
        assert lib().orc_qbvh_info(self.h, entity, C.byref(n), C.byref(r), C.byref(ok)) == 0
        return int(n.value), int(r.value), bool(ok.value)

    def light_order(self):
        buf = (C.c_int32 * 256)()
        n = lib().orc_light_order(self.h, buf, 256)
        return [int(buf[i]) for i in range(n)]

    def set_face_albedo(self, entity: int, albedo: np.ndarray):
        a = np.ascontiguousarray(albedo, dtype=np.float32)
        assert lib().orc_set_face_albedo(self.h, entity, a.ctypes.data) == 0

    def set_emission(self, entity: int, emission):
        e = (C.c_double * 3)(*[float(x) for x in emission])
        assert lib().orc_set_emission(self.h, entity, e) == 0

    def adjoint(self, seed: int, dimage: np.ndarray, entity: int, light: int, samples: Optional[int] = None,
                threads: int = 0):
        """(d loss/d albedo [n_faces, 3], d loss/d emission [3]) for the
        adjoint image dimage [H, W, 3]."""
        d = np.ascontiguousarray(dimage, dtype=np.float64)
        nf = self.desc.entities[entity].indices.shape[0] if entity >= 0 else 0
        g = np.zeros((nf, 3), dtype=np.float64)
        ge = (C.c_double * 3)()
        rc = lib().orc_adjoint(self.h, seed, samples or 0, d.ctypes.data, entity, g.ctypes.data, light, ge, threads)
        assert rc == 0, rc
        return g, np.array(ge[:])


def mlt_mutate0(word: int, r: float) -> int:
    """Kelemen's Mutate0 (mlt.h:42-54) of one primary-sample word."""
    return int(lib().orc_mlt_mutate0(word, r))
