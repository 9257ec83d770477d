"""Scene description of the in-scope SceneEnv subset, and its C-ABI marshalling.

The field names and defaults follow prender's SceneEnv / Material / Entity
(proj/scene_env.h:262-350 defaults, 1392-1520 MATERIAL keywords, 1544-1850
OBJECT sphere/uvplane/objfile; proj/material.h:54-240; proj/entity.h:29-70).
A ``SceneDesc`` is what ``SceneEnv::Load`` would have produced from a scene
file; ``to_c`` turns it into the ``prt_scene_desc`` struct of include/prt.h.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Union

import numpy as np

REFLECTION_DIFFUSE = 0
REFLECTION_SPECULAR = 1
REFLECTION_REFRACTION = 2

ENTITY_SPHERE = 1
ENTITY_UVPLANE = 3
ENTITY_POLYGON = 5

# ---------------------------------------------------------------- C structs
# Mirrors of include/prt.h (layout checked by tests/test_abi.py).
D3 = C.c_double * 3


class PrtMaterial(C.Structure):
    _fields_ = [("color", D3), ("emission", D3), ("roughness", C.c_double),
                ("refractive_index", C.c_double), ("reflection_type", C.c_int32),
                ("pad_", C.c_int32)]


class PrtEntity(C.Structure):
    _fields_ = [("type", C.c_int32), ("material", C.c_int32),
                ("normal_vector_inverse", C.c_int32), ("n_vertices", C.c_int32),
                ("p0", D3), ("p1", D3), ("p2", D3), ("radius", C.c_double),
                ("vertices", C.POINTER(C.c_float)), ("indices", C.POINTER(C.c_int32)),
                ("face_albedo", C.POINTER(C.c_float)), ("n_faces", C.c_int64)]


class PrtCamera(C.Structure):
    _fields_ = [("camera_position", D3), ("camera_dir", D3), ("camera_up", D3),
                ("world_screen_width", C.c_double), ("world_screen_height", C.c_double),
                ("world_screen_dist", C.c_double), ("camera_angle", C.c_double),
                ("image_width", C.c_int32), ("image_height", C.c_int32)]


class PrtSampling(C.Structure):
    _fields_ = [("samples", C.c_int32), ("supersamples", C.c_int32), ("depth", C.c_int32),
                ("depth_limit", C.c_int32), ("next_event_estimation", C.c_int32),
                ("pad_", C.c_int32), ("luminance_cutoff", C.c_double),
                ("sensor_response", C.c_double), ("background_color", D3),
                ("environment_light", D3)]


class PrtSceneDesc(C.Structure):
    _fields_ = [("camera", PrtCamera), ("sampling", PrtSampling),
                ("materials", C.POINTER(PrtMaterial)), ("n_materials", C.c_int32),
                ("n_entities", C.c_int32), ("entities", C.POINTER(PrtEntity))]


class PrtRay(C.Structure):
    _fields_ = [("org", D3), ("dir", D3)]


class PrtHit(C.Structure):
    _fields_ = [("t", C.c_double), ("u", C.c_double), ("v", C.c_double),
                ("entity", C.c_int32), ("prim", C.c_int32)]


RAY_DTYPE = np.dtype([("org", "<f8", 3), ("dir", "<f8", 3)])
HIT_DTYPE = np.dtype([("t", "<f8"), ("u", "<f8"), ("v", "<f8"), ("entity", "<i4"), ("prim", "<i4")])


class PrtStats(C.Structure):
    _fields_ = [("camera_samples", C.c_uint64), ("primary_rays", C.c_uint64),
                ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("pickup_hits", C.c_uint64), ("node_visits", C.c_uint64),
                ("tri_tests", C.c_uint64), ("truncated_paths", C.c_uint64),
                ("bytes_algorithmic", C.c_uint64), ("ms_total", C.c_double),
                ("ms_extend", C.c_double), ("ms_shade", C.c_double),
                ("ms_shadow", C.c_double), ("ms_other", C.c_double),
                ("launches_extend", C.c_int64), ("shadow_node_visits", C.c_uint64),
                ("shadow_tri_tests", C.c_uint64), ("bytes_extend", C.c_uint64),
                ("bytes_shadow", C.c_uint64), ("ms_extend_primary", C.c_double),
                ("ms_nee", C.c_double), ("mixed_scene_passes", C.c_uint64),
                ("ms_traversal", C.c_double), ("mlt_proposals", C.c_uint64),
                ("mlt_accepts", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PrtRenderOpts(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("samples_override", C.c_int32),
                ("batch_pixels", C.c_int32), ("max_depth", C.c_int32),
                ("reserved_", C.c_int32), ("timing", C.c_int32), ("image_f64", C.c_int32),
                ("engine", C.c_int32)]


# void callback(void* user, const float* rgb, int32_t width, int32_t height, int32_t index, double fraction_done)
PrtDumpCallback = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_float), C.c_int32, C.c_int32, C.c_int32, C.c_double)


class PrtDumpOpts(C.Structure):
    _fields_ = [("seconds", C.c_double), ("directory", C.c_char_p), ("gamma_offset", C.c_double),
                ("callback", PrtDumpCallback), ("user", C.c_void_p)]


class PrtMltOpts(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("mutation", C.c_double), ("pre_sample", C.c_double),
                ("p_large", C.c_double), ("mutations_per_chain", C.c_int64), ("samples_override", C.c_int32),
                ("chains", C.c_int32), ("sample_depths", C.c_int32), ("max_depth", C.c_int32),
                ("image_f64", C.c_int32), ("pad_", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad_"}


class PrtErptOpts(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("mutation", C.c_double), ("mutate_distance", C.c_double),
                ("samples_override", C.c_int32), ("sample_depths", C.c_int32), ("max_depth", C.c_int32),
                ("image_f64", C.c_int32)]


class PrtBatchOpts(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("tile_samples", C.c_int32), ("inflight_tiles", C.c_int32),
                ("stack_lds", C.c_int32), ("trav_blocks_per_cu", C.c_int32), ("use_graph", C.c_int32),
                ("tile_stride", C.c_int32), ("max_depth", C.c_int32), ("timing", C.c_int32),
                ("persistent", C.c_int32), ("persist_waves", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# ---------------------------------------------------------------- Python side
Vec = Sequence[float]


@dataclass
class Material:
    """Material keywords (scene_env.h:1392-1491).  ``reflection`` follows the
    keyword: "diffuse" sets roughness 1, "specular" roughness 0,
    "reflection" is glass (IOR default 1.5).  With ``reflection=None`` the
    material is DIFFUSE with roughness 0 (a mirror, material.h:72,96)."""
    color: Vec = (0.0, 0.0, 0.0)
    emission: Vec = (0.0, 0.0, 0.0)
    reflection: Optional[str] = "diffuse"
    roughness: Optional[float] = None
    refractive_index: float = 1.5

    def resolved(self):
        rtype, rough = REFLECTION_DIFFUSE, 0.0
        if self.reflection == "diffuse":
            rtype, rough = REFLECTION_DIFFUSE, 1.0
        elif self.reflection == "specular":
            rtype, rough = REFLECTION_SPECULAR, 0.0
        elif self.reflection == "reflection":
            rtype = REFLECTION_REFRACTION
        elif self.reflection is not None:
            raise ValueError(f"reflection type {self.reflection!r} is outside the in-scope subset")
        if self.roughness is not None:
            rough = float(self.roughness)
        return rtype, rough

    def is_light(self) -> bool:
        """Material::isLight (material.h:195-210)."""
        rtype, _ = self.resolved()
        c = np.asarray(self.color, dtype=np.float64)
        e = np.asarray(self.emission, dtype=np.float64)
        return rtype == REFLECTION_DIFFUSE and float(np.sqrt(c @ c)) < 0.001 and float(np.sqrt(e @ e)) > 0.001


@dataclass
class Sphere:
    position: Vec
    radius: float
    material: int
    normal_vector_inverse: int = 1


@dataclass
class UVPlane:
    org: Vec
    U: Vec
    V: Vec
    material: int
    normal_vector_inverse: int = 1


@dataclass
class Polygon:
    vertices: np.ndarray          # (nv, 3) float32
    indices: np.ndarray           # (nf, 3) int32, 0-based
    material: int
    face_albedo: Optional[np.ndarray] = None  # (nf, 3) float32
    normal_vector_inverse: int = 1

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, dtype=np.float32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(self.indices, dtype=np.int32).reshape(-1, 3)
        if self.face_albedo is not None:
            self.face_albedo = np.ascontiguousarray(self.face_albedo, dtype=np.float32).reshape(-1, 3)


Entity = Union[Sphere, UVPlane, Polygon]


@dataclass
class Camera:
    """SceneEnv camera defaults (scene_env.h:326-343)."""
    position: Vec = (50.0, 52.0, 220.0)
    dir: Vec = (0.0, -0.04, -1.0)
    up: Vec = (0.0, 1.0, 0.0)
    world_screen_width: float = 30.0
    world_screen_height: float = 30.0
    world_screen_dist: float = 40.0
    camera_angle: Optional[float] = None  # CAMERA_ANGLE (vertical FOV, degrees)
    width: int = 320
    height: int = 240


@dataclass
class Sampling:
    """Sampling defaults (scene_env.h:221, 329-349)."""
    samples: int = 8
    supersamples: int = 2
    depth: int = 8
    depth_limit: int = 64
    next_event_estimation: int = 0
    luminance_cutoff: float = 0.025
    sensor_response: float = 1.0
    background: Vec = (0.0, 0.0, 0.0)
    environment_light: Vec = (0.0, 0.0, 0.0)


@dataclass
class SceneDesc:
    camera: Camera = field(default_factory=Camera)
    sampling: Sampling = field(default_factory=Sampling)
    materials: List[Material] = field(default_factory=list)
    entities: List[Entity] = field(default_factory=list)
    name: str = "scene"

    @property
    def spp(self) -> int:
        return self.sampling.samples * self.sampling.supersamples ** 2

    @property
    def n_triangles(self) -> int:
        return int(sum(e.indices.shape[0] for e in self.entities if isinstance(e, Polygon)))

    def light_entities(self) -> List[int]:
        return [i for i, e in enumerate(self.entities) if self.materials[e.material].is_light()]


def _d3(v) -> D3:
    a = [float(x) for x in v]
    if len(a) != 3:
        raise ValueError("expected 3 components")
    return D3(*a)


class CDesc:
    """A prt_scene_desc plus the buffers it points into (kept alive here)."""

    def __init__(self, desc: SceneDesc):
        self.keep = []
        mats = (PrtMaterial * max(1, len(desc.materials)))()
        for i, m in enumerate(desc.materials):
            rtype, rough = m.resolved()
            mats[i].color = _d3(m.color)
            mats[i].emission = _d3(m.emission)
            mats[i].roughness = rough
            mats[i].refractive_index = float(m.refractive_index)
            mats[i].reflection_type = rtype
        ents = (PrtEntity * max(1, len(desc.entities)))()
        for i, e in enumerate(desc.entities):
            pe = ents[i]
            pe.material = int(e.material)
            pe.normal_vector_inverse = int(e.normal_vector_inverse)
            if isinstance(e, Sphere):
                pe.type = ENTITY_SPHERE
                pe.p0 = _d3(e.position)
                pe.radius = float(e.radius)
            elif isinstance(e, UVPlane):
                pe.type = ENTITY_UVPLANE
                pe.p0 = _d3(e.org)
                pe.p1 = _d3(e.U)
                pe.p2 = _d3(e.V)
            elif isinstance(e, Polygon):
                pe.type = ENTITY_POLYGON
                self.keep += [e.vertices, e.indices]
                pe.vertices = e.vertices.ctypes.data_as(C.POINTER(C.c_float))
                pe.indices = e.indices.ctypes.data_as(C.POINTER(C.c_int32))
                pe.n_vertices = int(e.vertices.shape[0])
                pe.n_faces = int(e.indices.shape[0])
                if e.face_albedo is not None:
                    self.keep.append(e.face_albedo)
                    pe.face_albedo = e.face_albedo.ctypes.data_as(C.POINTER(C.c_float))
            else:
                raise TypeError(f"unknown entity {type(e)}")
        d = PrtSceneDesc()
        cam = desc.camera
        d.camera.camera_position = _d3(cam.position)
        d.camera.camera_dir = _d3(cam.dir)
        d.camera.camera_up = _d3(cam.up)
        d.camera.world_screen_width = float(cam.world_screen_width)
        d.camera.world_screen_height = float(cam.world_screen_height)
        d.camera.world_screen_dist = float(cam.world_screen_dist)
        d.camera.camera_angle = float(cam.camera_angle) if cam.camera_angle else 0.0
        d.camera.image_width = int(cam.width)
        d.camera.image_height = int(cam.height)
        s = desc.sampling
        d.sampling.samples = int(s.samples)
        d.sampling.supersamples = int(s.supersamples)
        d.sampling.depth = int(s.depth)
        d.sampling.depth_limit = int(s.depth_limit)
        d.sampling.next_event_estimation = int(s.next_event_estimation)
        d.sampling.luminance_cutoff = float(s.luminance_cutoff)
        d.sampling.sensor_response = float(s.sensor_response)
        d.sampling.background_color = _d3(s.background)
        d.sampling.environment_light = _d3(s.environment_light)
        d.materials = mats
        d.n_materials = len(desc.materials)
        d.entities = ents
        d.n_entities = len(desc.entities)
        self.keep += [mats, ents]
        self.desc = d

    def ptr(self):
        return C.byref(self.desc)


def rays_to_array(org: np.ndarray, dirs: np.ndarray) -> np.ndarray:
    org = np.asarray(org, dtype=np.float64).reshape(-1, 3)
    dirs = np.asarray(dirs, dtype=np.float64).reshape(-1, 3)
    r = np.empty(org.shape[0], dtype=RAY_DTYPE)
    r["org"] = org
    r["dir"] = dirs
    return r
