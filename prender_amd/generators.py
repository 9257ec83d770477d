"""Seeded scene generators for the benchmark configurations.

Every generator is deterministic in its ``seed`` and returns a ``SceneDesc``
whose fields carry prender's scene-file semantics (proj/scene_env.h):
``SAMPLING``/``SUPERSAMPLING`` keep the reference's fixed 2x2 sub-pixel grid,
``DEPTH``/``DEPTH_LIMIT`` its soft depth cap, ``luminance_cutoff 0`` disables
the per-sample clamp (render.h:1375-1384), and an a x b area light is a
``uvplane`` with |U| = sqrt(a), |V| = sqrt(b) because UVPlane::intersect bounds
the plane coordinates by u_length after dividing by it (plane.h:232-238).
"""
from __future__ import annotations

import math
from typing import List, Optional, Tuple

import numpy as np

from .scene import (Camera, Material, Polygon, SceneDesc, Sampling, Sphere, UVPlane,
                    rays_to_array)

# ---------------------------------------------------------------- meshes


def icosphere(level: int) -> Tuple[np.ndarray, np.ndarray]:
    """Unit icosphere by repeated 4-way midpoint subdivision: 20*4^level faces."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        e.sort(axis=1)
        key = e[:, 0] * (v.shape[0] + 1) + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a = uniq // (v.shape[0] + 1)
        b = uniq % (v.shape[0] + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid], axis=0)
        nf = f.shape[0]
        m01 = base + inv[:nf]
        m12 = base + inv[nf:2 * nf]
        m20 = base + inv[2 * nf:]
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], 1),
            np.stack([f[:, 1], m12, m01], 1),
            np.stack([f[:, 2], m20, m12], 1),
            np.stack([m01, m12, m20], 1)], axis=0)
    return v, f


def value_noise(p: np.ndarray, seed: int, octaves: int = 4, base_freq: float = 2.0) -> np.ndarray:
    """Fractal 3-D value noise in [-1, 1]: lattice values from a seeded
    permutation table, smoothstep-trilinear interpolation, octaves at
    frequency x2 and amplitude x0.5, normalised by the amplitude sum."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(256).astype(np.int64)
    perm = np.concatenate([perm, perm])
    vals = rng.uniform(-1.0, 1.0, 256)

    def lattice(ix, iy, iz):
        return vals[perm[perm[perm[ix & 255] + (iy & 255)] + (iz & 255)]]

    total = np.zeros(p.shape[0])
    amp, freq, norm = 1.0, base_freq, 0.0
    for _ in range(octaves):
        q = p * freq
        i0 = np.floor(q).astype(np.int64)
        fr = q - i0
        s = fr * fr * (3.0 - 2.0 * fr)
        acc = np.zeros(p.shape[0])
        for dx in (0, 1):
            wx = s[:, 0] if dx else 1.0 - s[:, 0]
            for dy in (0, 1):
                wy = s[:, 1] if dy else 1.0 - s[:, 1]
                for dz in (0, 1):
                    wz = s[:, 2] if dz else 1.0 - s[:, 2]
                    acc += wx * wy * wz * lattice(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz)
        total += amp * acc
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def displaced_icosphere(level: int, radius: float, center, seed: int, amplitude: float = 0.1):
    """Level-`level` icosphere radially displaced by 4-octave value noise of
    amplitude ``amplitude * radius``; float32 vertices, int32 indices."""
    v, f = icosphere(level)
    n = value_noise(v, seed)
    r = radius * (1.0 + amplitude * n)
    verts = v * r[:, None] + np.asarray(center, dtype=np.float64)[None, :]
    return verts.astype(np.float32), f.astype(np.int32)


def displaced_latlong(n_tri_target: int, radius: float, center, seed: int, amplitude: float = 0.1):
    """Latitude/longitude sphere with about ``n_tri_target`` triangles,
    displaced like ``displaced_icosphere`` (used where a mesh of a given size is
    needed between icosphere levels)."""
    nt = max(3, int(round(math.sqrt(n_tri_target / 4.0))))
    nphi = 2 * nt
    th = np.linspace(0.0, math.pi, nt + 1)
    ph = np.linspace(0.0, 2 * math.pi, nphi, endpoint=False)
    T, P = np.meshgrid(th[1:-1], ph, indexing="ij")
    body = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    v = np.concatenate([[[0, 1, 0]], body, [[0, -1, 0]]], axis=0)
    faces = []
    ring = lambda i, j: 1 + i * nphi + (j % nphi)
    for j in range(nphi):
        faces.append([0, ring(0, j + 1), ring(0, j)])
    for i in range(nt - 2):
        for j in range(nphi):
            a, b, c, d = ring(i, j), ring(i, j + 1), ring(i + 1, j + 1), ring(i + 1, j)
            faces.append([a, b, c])
            faces.append([c, d, a])
    last = v.shape[0] - 1
    for j in range(nphi):
        faces.append([last, ring(nt - 2, j), ring(nt - 2, j + 1)])
    f = np.asarray(faces, dtype=np.int64)
    n = value_noise(v, seed)
    r = radius * (1.0 + amplitude * n)
    verts = v * r[:, None] + np.asarray(center, dtype=np.float64)[None, :]
    return verts.astype(np.float32), f.astype(np.int32)


def quad_mesh(quads: List[np.ndarray]) -> Tuple[np.ndarray, np.ndarray]:
    """Quads (4 corners each) -> triangles, split as ObjLoader does a 4-vertex
    face: (0,1,2) + (2,3,0) (ObjLoader.cpp:435-455)."""
    verts, faces = [], []
    for q in quads:
        b = len(verts)
        verts.extend(np.asarray(q, dtype=np.float64))
        faces.append([b, b + 1, b + 2])
        faces.append([b + 2, b + 3, b])
    return np.asarray(verts, dtype=np.float32), np.asarray(faces, dtype=np.int32)


def box_quads(lo, hi) -> List[np.ndarray]:
    """Six outward-facing faces of an axis-aligned box."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    return [
        np.array([[x0, y0, z0], [x1, y0, z0], [x1, y0, z1], [x0, y0, z1]]),  # bottom
        np.array([[x0, y1, z0], [x0, y1, z1], [x1, y1, z1], [x1, y1, z0]]),  # top
        np.array([[x0, y0, z0], [x0, y1, z0], [x1, y1, z0], [x1, y0, z0]]),  # back
        np.array([[x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]]),  # front
        np.array([[x0, y0, z0], [x0, y0, z1], [x0, y1, z1], [x0, y1, z0]]),  # left
        np.array([[x1, y0, z0], [x1, y1, z0], [x1, y1, z1], [x1, y0, z1]]),  # right
    ]


# ---------------------------------------------------------------- configs

GREY = (0.75, 0.75, 0.75)
RED = (0.75, 0.25, 0.25)
BLUE = (0.25, 0.25, 0.75)


def cornell_spheres(width: int = 320, height: int = 240, samples: int = 4, supersamples: int = 2,
                    nee: int = 0, light_y: float = 76.0) -> SceneDesc:
    """Sphere-only Cornell box (edupt layout of proj/test/CornellBox.txt): six
    radius-1e5 diffuse walls, a mirror and a glass sphere (r=16.5, IOR 1.5) and
    a radius-5 sphere light of radiance 36 centred below the ceiling
    (y + 5 <= 81.6).  Camera (50,52,220) looking (0,-0.04,-1), CAMERA_ANGLE 30."""
    mats = [
        Material(color=RED),                                    # 0 left
        Material(color=BLUE),                                   # 1 right
        Material(color=GREY),                                   # 2 back / floor / ceiling
        Material(color=(0.0, 0.0, 0.0)),                        # 3 front (black)
        Material(color=(0.99, 0.99, 0.99), reflection="specular"),   # 4 mirror
        Material(color=(0.99, 0.99, 0.99), reflection="reflection", refractive_index=1.5),  # 5 glass
        Material(color=(0.0, 0.0, 0.0), emission=(36.0, 36.0, 36.0)),  # 6 light
    ]
    ents = [
        Sphere((1e5 + 1, 40.8, 81.6), 1e5, 0),
        Sphere((-1e5 + 99, 40.8, 81.6), 1e5, 1),
        Sphere((50, 40.8, 1e5), 1e5, 2),
        Sphere((50, 40.8, -1e5 + 250), 1e5, 3),
        Sphere((50, 1e5, 81.6), 1e5, 2),
        Sphere((50, -1e5 + 81.6, 81.6), 1e5, 2),
        Sphere((27, 16.5, 47), 16.5, 4),
        Sphere((77, 16.5, 78), 16.5, 5),
        Sphere((50, light_y, 81.6), 5.0, 6),
    ]
    cam = Camera(position=(50.0, 52.0, 220.0), dir=(0.0, -0.04, -1.0), up=(0.0, 1.0, 0.0),
                 world_screen_dist=40.0, camera_angle=30.0, width=width, height=height)
    smp = Sampling(samples=samples, supersamples=supersamples, depth=5, depth_limit=8,
                   next_event_estimation=nee, luminance_cutoff=0.0, sensor_response=1.0)
    return SceneDesc(camera=cam, sampling=smp, materials=mats, entities=ents, name="cfg1")


def config1(**kw) -> SceneDesc:
    """Plain sphere Cornell box: 320x240, SAMPLING 4 x SUPERSAMPLING 2 (16 spp),
    DEPTH 5 / DEPTH_LIMIT 8, no NEE."""
    return cornell_spheres(**kw)


def quad_box(with_front: bool = True):
    """Unit Cornell box as 12 quads (24 triangles): every wall is two quads;
    x in [-1,1], y in [0,2], z in [-1,3.5].  Per-face albedo gives the red left,
    blue right and grey other walls."""
    x0, x1, y0, y1, z0, z1 = -1.0, 1.0, 0.0, 2.0, -1.0, 3.5
    zm, xm = 0.5 * (z0 + z1), 0.0
    quads, cols = [], []

    def add(q, c):
        quads.append(np.asarray(q, dtype=np.float64))
        cols.append(c)
    # floor / ceiling split along z
    for za, zb in ((z0, zm), (zm, z1)):
        add([[x0, y0, za], [x1, y0, za], [x1, y0, zb], [x0, y0, zb]], GREY)
        add([[x0, y1, za], [x0, y1, zb], [x1, y1, zb], [x1, y1, za]], GREY)
    # back / front split along x
    for xa, xb in ((x0, xm), (xm, x1)):
        add([[xa, y0, z0], [xa, y1, z0], [xb, y1, z0], [xb, y0, z0]], GREY)
        if with_front:
            add([[xa, y0, z1], [xb, y0, z1], [xb, y1, z1], [xa, y1, z1]], GREY)
    # left / right split along z
    for za, zb in ((z0, zm), (zm, z1)):
        add([[x0, y0, za], [x0, y0, zb], [x0, y1, zb], [x0, y1, za]], RED)
        add([[x1, y0, za], [x1, y1, za], [x1, y1, zb], [x1, y0, zb]], BLUE)
    v, f = quad_mesh(quads)
    alb = np.repeat(np.asarray(cols, dtype=np.float32), 2, axis=0)
    return v, f, alb


def quad_box_scene(mesh_v: np.ndarray, mesh_f: np.ndarray, width: int, height: int, samples: int,
                   mesh_albedo: Optional[np.ndarray] = None, light_emission: float = 60.0,
                   name: str = "cfg2", mesh_color=(0.8, 0.8, 0.8)) -> SceneDesc:
    bv, bf, balb = quad_box()
    mats = [
        Material(color=GREY),                                       # 0 box (per-face albedo)
        Material(color=mesh_color),                                 # 1 procedural mesh
        Material(color=(0.0, 0.0, 0.0), emission=(light_emission,) * 3),  # 2 area light
    ]
    # 0.25 x 0.25 emitter: |U| = |V| = 0.5, centred under the ceiling, facing down
    light = UVPlane(org=(-0.125, 2.0 - 1e-3, -0.125), U=(0.5, 0.0, 0.0), V=(0.0, 0.0, 0.5), material=2)
    ents = [Polygon(bv, bf, 0, face_albedo=balb), Polygon(mesh_v, mesh_f, 1, face_albedo=mesh_albedo), light]
    cam = Camera(position=(0.0, 1.0, 3.3), dir=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0),
                 world_screen_dist=1.0, camera_angle=40.0, width=width, height=height)
    smp = Sampling(samples=samples, supersamples=2, depth=5, depth_limit=8, next_event_estimation=1,
                   luminance_cutoff=0.0, sensor_response=1.0)
    return SceneDesc(camera=cam, sampling=smp, materials=mats, entities=ents, name=name)


def config2(width: int = 512, height: int = 512, samples: int = 16, level: int = 6, seed: int = 2) -> SceneDesc:
    """Triangle Cornell box (24 tris) + 0.25^2 area light + a level-6 icosphere
    (81,920 tris) displaced by 4-octave value noise (amplitude 0.1 r), albedo 0.8;
    512x512, 64 spp (SAMPLING 16 x SUPERSAMPLING 2), DEPTH 5 / DEPTH_LIMIT 8, NEE."""
    v, f = displaced_icosphere(level, 0.5, (0.0, 0.6, 0.0), seed)
    return quad_box_scene(v, f, width, height, samples)


def config3(width: int = 1024, height: int = 768, samples: int = 16, seed: int = 3,
            scale: float = 1.0) -> SceneDesc:
    """About 2M triangles: 320 displaced icospheres (8 x L6, 40 x L5, 80 x L4,
    192 x L3) and 80 boxes, sizes log-uniform, placed uniformly on a 100x100
    ground mesh; 70/15/15 diffuse/mirror/glass with albedo U[0.2,0.9]; one
    1x1 and eight 0.25x0.25 area lights of distinct power; 1024x768, 64 spp,
    DEPTH 5 / DEPTH_LIMIT 10.  ``scale`` < 1 shrinks the instance counts for tests."""
    rng = np.random.default_rng(seed)
    levels = [6] * 8 + [5] * 40 + [4] * 80 + [3] * 192
    kinds = [("ico", L) for L in levels] + [("box", 0)] * 80
    if scale < 1.0:
        keep = max(4, int(len(kinds) * scale))
        kinds = [k for k in kinds if k[1] <= 4][:keep]
    order = rng.permutation(len(kinds))
    kinds = [kinds[i] for i in order]
    mats: List[Material] = [Material(color=(0.6, 0.6, 0.6))]  # ground
    ents = []
    # ground: 20x20 grid of quads over [-50,50]^2
    g = np.linspace(-50.0, 50.0, 21)
    quads = []
    for i in range(20):
        for j in range(20):
            quads.append(np.array([[g[i], 0, g[j]], [g[i], 0, g[j + 1]], [g[i + 1], 0, g[j + 1]], [g[i + 1], 0, g[j]]]))
    gv, gf = quad_mesh(quads)
    ents.append(Polygon(gv, gf, 0))
    for k, (kind, L) in enumerate(kinds):
        r = float(np.exp(rng.uniform(math.log(0.6), math.log(3.0))))
        x, z = rng.uniform(-48.0, 48.0, 2)
        u = rng.uniform()
        alb = tuple(float(a) for a in rng.uniform(0.2, 0.9, 3))
        if u < 0.70:
            mats.append(Material(color=alb))
        elif u < 0.85:
            mats.append(Material(color=alb, reflection="specular"))
        else:
            mats.append(Material(color=alb, reflection="reflection", refractive_index=1.5))
        mid = len(mats) - 1
        if kind == "ico":
            v, f = displaced_icosphere(L, r, (x, r * 1.0, z), seed * 1000 + k)
        else:
            h = r * rng.uniform(0.8, 2.0)
            v, f = quad_mesh(box_quads((x - r, 0.0, z - r), (x + r, h, z + r)))
        ents.append(Polygon(v, f, mid))
    # lights: one 1x1 (|U|=|V|=1) and eight 0.25x0.25 (|U|=|V|=0.5), power
    # area*|E| distinct by far more than the reference's sort jitter
    big = len(mats)
    mats.append(Material(color=(0, 0, 0), emission=(3000.0,) * 3))
    ents.append(UVPlane(org=(-0.5, 40.0, -0.5), U=(1.0, 0, 0), V=(0, 0, 1.0), material=big))
    for i in range(8):
        ang = 2 * math.pi * i / 8
        e = 4000.0 + 400.0 * i
        mats.append(Material(color=(0, 0, 0), emission=(e, e, e)))
        ents.append(UVPlane(org=(30 * math.cos(ang), 25.0, 30 * math.sin(ang)), U=(0.5, 0, 0), V=(0, 0, 0.5),
                            material=len(mats) - 1))
    cam = Camera(position=(0.0, 30.0, 80.0), dir=(0.0, -0.38, -1.0), up=(0.0, 1.0, 0.0),
                 world_screen_dist=1.0, camera_angle=45.0, width=width, height=height)
    smp = Sampling(samples=samples, supersamples=2, depth=5, depth_limit=10, next_event_estimation=1,
                   luminance_cutoff=0.0, sensor_response=1.0)
    return SceneDesc(camera=cam, sampling=smp, materials=mats, entities=ents, name="cfg3")


def config4(n_scenes: int = 64, seed: int = 4, tri_range=(5_000, 500_000),
            resolutions=(128, 256, 384, 512), spps=(8, 16, 32, 64)) -> List[SceneDesc]:
    """64 scenes from the sphere-box and quad-box generators, each with one
    displaced mesh of log-uniform size in [5k, 500k] triangles, a resolution
    from {128,256,384,512}^2 and spp from {8,16,32,64} (SUPERSAMPLING 2)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_scenes):
        ntri = int(np.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1]))))
        res = int(rng.choice(resolutions))
        spp = int(rng.choice(spps))
        samples = max(1, spp // 4)
        if i % 2 == 0:
            v, f = displaced_latlong(ntri, 0.5, (0.0, 0.6, 0.0), seed * 100 + i)
            sc = quad_box_scene(v, f, res, res, samples, name=f"cfg4_{i:02d}")
        else:
            sc = cornell_spheres(width=res, height=res, samples=samples, nee=0)
            v, f = displaced_latlong(ntri, 14.0, (50.0, 14.0, 100.0), seed * 100 + i)
            sc.materials.append(Material(color=(0.25, 0.75, 0.25)))
            sc.entities.append(Polygon(v, f, len(sc.materials) - 1))
            sc.name = f"cfg4_{i:02d}"
        out.append(sc)
    return out


def config5(width: int = 256, height: int = 256, samples: int = 8, seed: int = 5, target: bool = False,
            light_emission: float = 60.0) -> SceneDesc:
    """Inverse-rendering scene: the quad box + level-6 displaced icosphere with
    81,920 per-face RGB albedos (U[0.1,0.9] seeded for the target, 0.5 for the
    initial guess) and the area light's emission as a parameter; 256x256.
    Forward and adjoint run at 32 spp (SAMPLING 8 x SUPERSAMPLING 2)."""
    v, f = displaced_icosphere(6, 0.5, (0.0, 0.6, 0.0), 2)
    if target:
        alb = np.random.default_rng(seed).uniform(0.1, 0.9, (f.shape[0], 3)).astype(np.float32)
    else:
        alb = np.full((f.shape[0], 3), 0.5, dtype=np.float32)
    return quad_box_scene(v, f, width, height, samples, mesh_albedo=alb, light_emission=light_emission,
                          name="cfg5")


# ---------------------------------------------------------------- ray batches

def scene_aabb(desc: SceneDesc):
    lo = np.full(3, np.inf)
    hi = np.full(3, -np.inf)
    for e in desc.entities:
        if isinstance(e, Polygon):
            lo = np.minimum(lo, e.vertices.min(0))
            hi = np.maximum(hi, e.vertices.max(0))
    return lo, hi


def camera_basis(desc: SceneDesc):
    """CameraScreenPram (render.h:100-146) in numpy (double)."""
    cam = desc.camera
    d = np.asarray(cam.dir, dtype=np.float64)
    d = d / np.linalg.norm(d)
    wsw, wsh = cam.world_screen_width, cam.world_screen_height
    if cam.camera_angle:
        wsw = 2.0 * cam.world_screen_dist * math.tan(cam.camera_angle * math.pi / 180.0 / 2.0)
        wsh = wsw
    sw = wsw * cam.width / cam.height
    sx = np.cross(d, np.asarray(cam.up, dtype=np.float64))
    sx = sx / np.linalg.norm(sx) * sw
    sy = np.cross(sx, d)
    sy = sy / np.linalg.norm(sy) * wsh
    center = np.asarray(cam.position, dtype=np.float64) + d * cam.world_screen_dist
    return np.asarray(cam.position, dtype=np.float64), center, sx, sy


def camera_rays(desc: SceneDesc) -> np.ndarray:
    """The unique camera rays of a frame: one per sub-pixel of the fixed
    SSxSS grid, r = (s + 0.5) / SS (render.h:1248-1269), in Run's order
    (y, x, sy, sx) with y = 0 the bottom image row."""
    pos, center, sx, sy = camera_basis(desc)
    W, H, ss = desc.camera.width, desc.camera.height, desc.sampling.supersamples
    y, x, syi, sxi = np.meshgrid(np.arange(H), np.arange(W), np.arange(ss), np.arange(ss), indexing="ij")
    r1 = ((sxi + 0.5) / ss + x) / W - 0.5
    r2 = ((syi + 0.5) / ss + y) / H - 0.5
    sp = center + sx * r1.reshape(-1, 1) + sy * r2.reshape(-1, 1)
    d = sp - pos
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return rays_to_array(np.broadcast_to(pos, d.shape), d)


def hit_parity_rays(desc: SceneDesc, n: int = 1 << 20, seed: int = 0x5EED) -> np.ndarray:
    """Seeded closest-hit batch: half camera rays through uniform sub-pixel
    positions, half with origins uniform in the mesh AABB and directions
    uniform on the sphere."""
    rng = np.random.default_rng(seed)
    n_cam = n // 2
    n_rnd = n - n_cam
    pos, center, sx, sy = camera_basis(desc)
    W, H = desc.camera.width, desc.camera.height
    px = rng.uniform(0, W, n_cam)
    py = rng.uniform(0, H, n_cam)
    sp = center[None, :] + sx[None, :] * (px / W - 0.5)[:, None] + sy[None, :] * (py / H - 0.5)[:, None]
    d1 = sp - pos[None, :]
    d1 /= np.linalg.norm(d1, axis=1, keepdims=True)
    o1 = np.repeat(pos[None, :], n_cam, axis=0)
    lo, hi = scene_aabb(desc)
    o2 = lo + (hi - lo) * rng.uniform(size=(n_rnd, 3))
    z = rng.uniform(-1, 1, n_rnd)
    ph = rng.uniform(0, 2 * math.pi, n_rnd)
    s = np.sqrt(1 - z * z)
    d2 = np.stack([s * np.cos(ph), s * np.sin(ph), z], 1)
    return rays_to_array(np.concatenate([o1, o2]), np.concatenate([d1, d2]))
