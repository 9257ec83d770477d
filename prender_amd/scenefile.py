"""Scene files of the reference (SceneEnv::Load + CreateEntList,
scene_env.h:357-1990) through the library's reader (prt_scene_file_load):
the parsed prt_scene_desc is copied into the Python description used by
``prt.Scene`` and the oracle."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import prt
from .scene import (ENTITY_POLYGON, ENTITY_SPHERE, ENTITY_UVPLANE, REFLECTION_REFRACTION, REFLECTION_SPECULAR,
                    Camera, Material, Polygon, Sampling, SceneDesc, Sphere, UVPlane)


def _v(a):
    return tuple(float(x) for x in a)


def load(path: str) -> SceneDesc:
    """Reads a scene file; raises prt.PrtError with the reader's message."""
    L = prt.lib()
    h = L.prt_scene_file_load(os.fsencode(path))
    if not h:
        raise prt.PrtError(f"prt_scene_file_load failed: {L.prt_last_error().decode(errors='replace')}")
    try:
        d = L.prt_scene_file_desc(h).contents
        g = C.c_double(0.0)
        output = L.prt_scene_file_output(h, C.byref(g)).decode(errors="replace")
        cam = d.camera
        camera = Camera(position=_v(cam.camera_position), dir=_v(cam.camera_dir), up=_v(cam.camera_up),
                        world_screen_width=cam.world_screen_width, world_screen_height=cam.world_screen_height,
                        world_screen_dist=cam.world_screen_dist,
                        camera_angle=cam.camera_angle if cam.camera_angle > 0 else None,
                        width=cam.image_width, height=cam.image_height)
        s = d.sampling
        sampling = Sampling(samples=s.samples, supersamples=s.supersamples, depth=s.depth,
                            depth_limit=s.depth_limit, next_event_estimation=s.next_event_estimation,
                            luminance_cutoff=s.luminance_cutoff, sensor_response=s.sensor_response,
                            background=_v(s.background_color), environment_light=_v(s.environment_light))
        mats = []
        for i in range(d.n_materials):
            m = d.materials[i]
            refl = {REFLECTION_SPECULAR: "specular", REFLECTION_REFRACTION: "reflection"}.get(m.reflection_type)
            mats.append(Material(color=_v(m.color), emission=_v(m.emission), reflection=refl,
                                 roughness=float(m.roughness), refractive_index=float(m.refractive_index)))
        ents = []
        for i in range(d.n_entities):
            e = d.entities[i]
            if e.type == ENTITY_SPHERE:
                ents.append(Sphere(_v(e.p0), float(e.radius), e.material, e.normal_vector_inverse))
            elif e.type == ENTITY_UVPLANE:
                ents.append(UVPlane(_v(e.p0), _v(e.p1), _v(e.p2), e.material, e.normal_vector_inverse))
            elif e.type == ENTITY_POLYGON:
                v = np.ctypeslib.as_array(e.vertices, shape=(e.n_vertices * 3,)).reshape(-1, 3).copy()
                f = np.ctypeslib.as_array(e.indices, shape=(e.n_faces * 3,)).reshape(-1, 3).copy()
                alb = None
                if e.face_albedo:
                    alb = np.ctypeslib.as_array(e.face_albedo, shape=(e.n_faces * 3,)).reshape(-1, 3).copy()
                ents.append(Polygon(v, f, e.material, alb, e.normal_vector_inverse))
        desc = SceneDesc(camera=camera, sampling=sampling, materials=mats, entities=ents,
                         name=os.path.splitext(os.path.basename(path))[0])
        desc.output = output
        desc.gamma_offset = g.value
        desc.image_dump = int(L.prt_scene_file_image_dump(h))  # IMAGEDUMP seconds
        return desc
    finally:
        L.prt_scene_file_free(h)


def mlt_settings(path: str):
    """(mode, {"mutation", "pre_sample"}) of a scene file: mode 2 = ERPT
    (energyRedistributionPathTracing), 1 = MLT (metropolisTransport), 0 =
    path tracing, with the reference's defaults (scene_env.h:1946-1968);
    pass the dict to ``prt.Scene.render_mlt`` (or its mutation to
    ``render_erpt``)."""
    L = prt.lib()
    h = L.prt_scene_file_load(os.fsencode(path))
    if not h:
        raise prt.PrtError(f"prt_scene_file_load failed: {L.prt_last_error().decode(errors='replace')}")
    try:
        o = prt.PrtMltOpts()
        mode = L.prt_scene_file_mlt(h, C.byref(o))
        return int(mode), {"mutation": o.mutation, "pre_sample": o.pre_sample}
    finally:
        L.prt_scene_file_free(h)
