"""Scene descriptions and the seeded generators for the benchmark configurations.

A ``SceneData`` holds the arrays of a ``bg_scene_desc``: a non-indexed
triangle soup (triangle.rs:96), spheres with their object_to_world
transform and its inverse (sphere.rs:38-65), materials and per-primitive
diffuse area lights (diffuse.rs:98-123).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N


@dataclass
class Camera:
    """PerspectiveCameraOptions (cameras/perspective.rs:36-46) with a static look_at transform."""

    pos: tuple
    look: tuple
    up: tuple
    fov: float
    width: int
    height: int
    near: float = 0.01
    far: float = 1000.0
    lens_radius: float = 0.0
    focal_distance: float = 1e6
    shutter_open: float = 0.0
    shutter_close: float = 1.0

    def matrices(self) -> tuple[np.ndarray, np.ndarray]:
        m = np.zeros(16, np.float32)
        mi = np.zeros(16, np.float32)
        p = np.asarray(self.pos, np.float32)
        lk = np.asarray(self.look, np.float32)
        u = np.asarray(self.up, np.float32)
        N.check(N.lib().bg_look_at(p.ctypes.data, lk.ctypes.data, u.ctypes.data, m.ctypes.data, mi.ctypes.data))
        return m, mi

    def desc(self) -> N.bg_camera_desc:
        m, mi = self.matrices()
        d = N.bg_camera_desc()
        d.camera_to_world[:] = [float(x) for x in m]
        d.camera_to_world_inv[:] = [float(x) for x in mi]
        d.fov = self.fov
        d.near_clip = self.near
        d.far_clip = self.far
        d.lens_radius = self.lens_radius
        d.focal_distance = self.focal_distance
        d.shutter_open = self.shutter_open
        d.shutter_close = self.shutter_close
        d.width = self.width
        d.height = self.height
        return d


@dataclass
class SceneData:
    tri_positions: np.ndarray = field(default_factory=lambda: np.zeros((0, 9), np.float32))
    tri_reverse: np.ndarray | None = None
    tri_material: np.ndarray | None = None
    sphere_o2w: np.ndarray = field(default_factory=lambda: np.zeros((0, 16), np.float32))
    sphere_w2o: np.ndarray = field(default_factory=lambda: np.zeros((0, 16), np.float32))
    sphere_params: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.float32))
    sphere_reverse: np.ndarray | None = None
    sphere_material: np.ndarray | None = None
    materials: list = field(default_factory=list)   # bg_material
    lights: list = field(default_factory=list)      # bg_area_light
    camera: Camera | None = None

    @property
    def n_tris(self) -> int:
        return int(self.tri_positions.shape[0])

    @property
    def n_spheres(self) -> int:
        return int(self.sphere_params.shape[0])

    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        lo = np.full(3, np.inf, np.float32)
        hi = np.full(3, -np.inf, np.float32)
        if self.n_tris:
            v = self.tri_positions.reshape(-1, 3)
            lo = np.minimum(lo, v.min(0))
            hi = np.maximum(hi, v.max(0))
        for i in range(self.n_spheres):
            c = self.sphere_o2w[i].reshape(4, 4)[:3, 3]
            r = self.sphere_params[i, 0]
            lo = np.minimum(lo, c - r)
            hi = np.maximum(hi, c + r)
        return lo.astype(np.float32), hi.astype(np.float32)

    def desc(self) -> tuple[N.bg_scene_desc, list]:
        """Returns the C descriptor and the arrays it points into (keep them alive)."""
        keep = []

        def arr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return a.ctypes.data

        d = N.bg_scene_desc()
        d.n_tris = self.n_tris
        d.tri_positions = arr(self.tri_positions, np.float32)
        d.tri_reverse_orientation = arr(self.tri_reverse, np.uint8)
        d.tri_material = arr(self.tri_material, np.uint32)
        d.n_spheres = self.n_spheres
        d.sphere_object_to_world = arr(self.sphere_o2w, np.float32)
        d.sphere_world_to_object = arr(self.sphere_w2o, np.float32)
        d.sphere_params = arr(self.sphere_params, np.float32)
        d.sphere_reverse_orientation = arr(self.sphere_reverse, np.uint8)
        d.sphere_material = arr(self.sphere_material, np.uint32)
        if self.materials:
            mats = (N.bg_material * len(self.materials))(*self.materials)
            keep.append(mats)
            d.n_materials = len(self.materials)
            d.materials = C.addressof(mats)
        if self.lights:
            ls = (N.bg_area_light * len(self.lights))(*self.lights)
            keep.append(ls)
            d.n_lights = len(self.lights)
            d.lights = C.addressof(ls)
        return d, keep


def load_obj(path: str, object_to_world=None, world_to_object=None) -> dict:
    """OBJ::read (io/obj.rs:16-124) through the library: {"positions": (n, 9) world-space triangles,
    "normals": (n, 9) or None, "uvs": (n, 6) or None}. object_to_world and its inverse are 4x4
    row-major (identity by default)."""
    m = np.eye(4, dtype=np.float32) if object_to_world is None else np.ascontiguousarray(object_to_world, np.float32)
    mi = (np.linalg.inv(m.astype(np.float64)).astype(np.float32) if world_to_object is None
          else np.ascontiguousarray(world_to_object, np.float32))
    L = N.lib()
    n = C.c_uint64(0)
    has = C.c_uint32(0)
    N.check(L.bg_obj_read(str(path).encode(), m.ctypes.data, mi.ctypes.data, C.byref(n), C.byref(has), None, None,
                          None))
    pos = np.empty((n.value, 9), np.float32)
    nrm = np.empty((n.value, 9), np.float32) if has.value & N.BG_OBJ_NORMALS else None
    uv = np.empty((n.value, 6), np.float32) if has.value & N.BG_OBJ_UVS else None
    N.check(L.bg_obj_read(str(path).encode(), m.ctypes.data, mi.ctypes.data, C.byref(n), C.byref(has),
                          pos.ctypes.data, None if nrm is None else nrm.ctypes.data,
                          None if uv is None else uv.ctypes.data))
    return {"positions": pos, "normals": nrm, "uvs": uv}


def matte(r, g, b) -> N.bg_material:
    m = N.bg_material()
    m.type = N.BG_MAT_MATTE
    m.reflectance[:] = [r, g, b]
    return m


# Gold-like conductor constants (RGB eta, k) recorded for the conductor materials.
GOLD_ETA = (0.143119, 0.374957, 1.44248)
GOLD_K = (3.98316, 2.38572, 1.60322)


def conductor(alpha, eta=GOLD_ETA, k=GOLD_K, r=(1.0, 1.0, 1.0)) -> N.bg_material:
    m = N.bg_material()
    m.type = N.BG_MAT_CONDUCTOR
    m.reflectance[:] = list(r)
    m.alpha = alpha
    m.eta[:] = list(eta)
    m.k[:] = list(k)
    return m


def plastic(kd, ks, roughness, remap_roughness=True) -> N.bg_material:
    """PlasticMaterial (materials/plastic.rs:24-52): Lambertian kd + Trowbridge-Reitz ks with a dielectric
    Fresnel term; roughness is remapped to alpha (microfacet.rs:45-53) when remap_roughness."""
    m = N.bg_material()
    m.type = N.BG_MAT_PLASTIC
    m.reflectance[:] = list(kd)
    m.eta[:] = list(ks)
    m.alpha = roughness
    m.k[:] = [1.0 if remap_roughness else 0.0, 0.0, 0.0]
    return m


def area_light(prim_id, radiance, double_sided=False) -> N.bg_area_light:
    a = N.bg_area_light()
    a.prim_id = prim_id
    a.radiance[:] = list(radiance)
    a.double_sided = 1 if double_sided else 0
    return a


def translate(c) -> tuple[np.ndarray, np.ndarray]:
    """Transform::translate (transform.rs:87-97): m and m_inverse, row-major."""
    m = np.eye(4, dtype=np.float32)
    mi = np.eye(4, dtype=np.float32)
    m[:3, 3] = np.asarray(c, np.float32)
    mi[:3, 3] = -np.asarray(c, np.float32)
    return m.reshape(16), mi.reshape(16)


def icosphere(level: int, center, radius: float, amplitude: float, frequency: float, seed: int) -> np.ndarray:
    n = int(N.lib().bgx_icosphere_tri_count(level))
    out = np.empty((n, 9), np.float32)
    c = np.asarray(center, np.float32)
    N.check(N.lib().bgx_icosphere(level, c.ctypes.data, radius, amplitude, frequency, seed, out.ctypes.data))
    return out


def incoherent_rays(n: int, lo, hi, seed: int) -> np.ndarray:
    """(n, 8) float32 rays: origins uniform in [lo, hi], directions uniform on the sphere."""
    out = np.empty((n, 8), np.float32)
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    N.check(N.lib().bgx_rays_incoherent(n, lo.ctypes.data, hi.ctypes.data, seed, out.ctypes.data))
    return out


def add_spheres(s: SceneData, centers, radii, material_ids=None) -> None:
    o2w, w2o, prm = [], [], []
    for c, r in zip(centers, radii):
        m, mi = translate(c)
        o2w.append(m)
        w2o.append(mi)
        prm.append([r, -r, r, 360.0])
    s.sphere_o2w = np.concatenate([s.sphere_o2w, np.asarray(o2w, np.float32).reshape(-1, 16)])
    s.sphere_w2o = np.concatenate([s.sphere_w2o, np.asarray(w2o, np.float32).reshape(-1, 16)])
    s.sphere_params = np.concatenate([s.sphere_params, np.asarray(prm, np.float32).reshape(-1, 4)])
    if material_ids is not None:
        prev = s.sphere_material if s.sphere_material is not None else np.zeros(0, np.uint32)
        s.sphere_material = np.concatenate([prev, np.asarray(material_ids, np.uint32)])


def quad(p0, p1, p2, p3) -> np.ndarray:
    """Two triangles (p0, p1, p2), (p0, p2, p3)."""
    return np.array([[*p0, *p1, *p2], [*p0, *p2, *p3]], np.float32)


def cornell_box(double_sided_light: bool = False, width: int = 256, height: int = 256,
                with_plastic: bool = False, sphere_light: bool = False) -> SceneData:
    """Small Cornell-style box in [0,1]^3: 5 walls as 10 triangles, a 0.25 x 0.25
    quad light (2 triangles, radiance 17) under the ceiling facing down, two
    r = 0.3 spheres (matte albedo 0.75, conductor alpha 0.2). `with_plastic`
    makes the matte sphere plastic (kd (0.5, 0.25, 0.1), ks 0.4, roughness 0.1
    remapped) and the floor glossy plastic (kd 0.6, ks 0.3, alpha 0.3) with its
    winding reversed so that its normal faces into the box (the reference's
    dielectric Fresnel term has a pole for light from behind the normal).
    `sphere_light` adds a third area light: a black r = 0.08 sphere at
    (0.5, 0.78, 0.45) emitting (8, 6, 4), with the quad light's sidedness."""
    walls = [
        quad((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)),        # floor y=0
        quad((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0)),        # ceiling y=1
        quad((0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)),        # back z=1
        quad((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0)),        # left x=0
        quad((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1)),        # right x=1
    ]
    light = quad((0.375, 0.999, 0.375), (0.625, 0.999, 0.375), (0.625, 0.999, 0.625), (0.375, 0.999, 0.625))
    tris = np.concatenate(walls + [light]).astype(np.float32)
    s = SceneData(tri_positions=tris)
    # materials: 0 white, 1 red, 2 green, 3 black (light), 4 sphere matte, 5 conductor
    s.materials = [matte(0.73, 0.73, 0.73), matte(0.65, 0.05, 0.05), matte(0.12, 0.45, 0.15),
                   matte(0.0, 0.0, 0.0), matte(0.75, 0.75, 0.75), conductor(0.2)]
    s.tri_material = np.array([0, 0, 0, 0, 0, 0, 1, 1, 2, 2, 3, 3], np.uint32)
    if with_plastic:
        tris[0:2] = quad((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0))
        s.tri_positions = tris
        s.materials[4] = plastic((0.5, 0.25, 0.1), (0.4, 0.4, 0.4), 0.1)
        s.materials.append(plastic((0.6, 0.6, 0.6), (0.3, 0.3, 0.3), 0.3, remap_roughness=False))
        s.tri_material[0:2] = 6
    add_spheres(s, [(0.3, 0.3, 0.6), (0.7, 0.3, 0.35)], [0.3, 0.3], [4, 5])
    s.lights = [area_light(10, (17.0, 17.0, 17.0), double_sided_light),
                area_light(11, (17.0, 17.0, 17.0), double_sided_light)]
    if sphere_light:
        add_spheres(s, [(0.5, 0.78, 0.45)], [0.08], [3])
        s.lights.append(area_light(s.n_tris + 2, (8.0, 6.0, 4.0), double_sided_light))
    s.camera = Camera(pos=(0.5, 0.5, -1.4), look=(0.5, 0.5, 0.5), up=(0.0, 1.0, 0.0), fov=40.0,
                      width=width, height=height)
    return s


def intersection_scene(level: int = 8, seed: int = 7) -> SceneData:
    """Seeded noise-displaced icosphere (20 * 4^level triangles; level 8 -> 1,310,720)
    of radius 1 at the origin, displacement amplitude 0.15, frequency 4."""
    return SceneData(tri_positions=icosphere(level, (0.0, 0.0, 0.0), 1.0, 0.15, 4.0, seed))


def _uniform(seed: int, index: int) -> float:
    return float(N.lib().bgx_uniform(seed, index))


def path_tracer_scene(width: int = 1920, height: int = 1080, seed: int = 2024, small: bool = False) -> SceneData:
    """The plain path-tracer configuration: 16 seeded displaced icospheres (4 at level 7, 8 at level 6,
    4 at level 5: 2,048,000 triangles) and 256 spheres of radius log-uniform in [0.01, 0.5] scattered
    over a ground quad, four double-sided quad area lights overhead, and per object either a matte
    material (albedo U[0.1, 0.9]) or a Trowbridge-Reitz conductor (alpha U[0.05, 0.8]), half each.
    `small` lowers every icosphere by three levels (for tests)."""
    levels = [7] * 4 + [6] * 8 + [5] * 4
    if small:
        levels = [max(1, l - 3) for l in levels]
    k = 0

    def u():
        nonlocal k
        k += 1
        return _uniform(seed, k)

    s = SceneData()
    s.materials = []
    tris, tri_mat = [], []
    # ground quad (matte 0.5) spanning the scene
    g = quad((-14, 0, -14), (14, 0, -14), (14, 0, 14), (-14, 0, 14))
    s.materials.append(matte(0.5, 0.5, 0.5))
    tris.append(g)
    tri_mat += [0, 0]
    for i, lv in enumerate(levels):
        r = 0.6 + 1.4 * u()
        c = (-10 + 20 * u(), r * 0.9, -10 + 20 * u())
        m = icosphere(lv, c, r, 0.12, 3.0, seed * 100 + i)
        if u() < 0.5:
            s.materials.append(matte(0.1 + 0.8 * u(), 0.1 + 0.8 * u(), 0.1 + 0.8 * u()))
        else:
            s.materials.append(conductor(0.05 + 0.75 * u()))
        tris.append(m)
        tri_mat += [len(s.materials) - 1] * m.shape[0]
    # four quad lights (double-sided), 2 x 2 m at y = 9
    light_base = sum(t.shape[0] for t in tris)
    for (lx, lz) in [(-6, -6), (6, -6), (-6, 6), (6, 6)]:
        tris.append(quad((lx - 1, 9, lz - 1), (lx + 1, 9, lz - 1), (lx + 1, 9, lz + 1), (lx - 1, 9, lz + 1)))
        s.materials.append(matte(0.0, 0.0, 0.0))
        tri_mat += [len(s.materials) - 1] * 2
    s.tri_positions = np.concatenate(tris).astype(np.float32)
    s.tri_material = np.asarray(tri_mat, np.uint32)
    s.lights = [area_light(light_base + j, (12.0, 12.0, 11.0), True) for j in range(8)]
    centers, radii, smat = [], [], []
    for i in range(256):
        r = float(np.exp(np.log(0.01) + (np.log(0.5) - np.log(0.01)) * u()))
        centers.append((-12 + 24 * u(), r, -12 + 24 * u()))
        radii.append(r)
        if u() < 0.5:
            s.materials.append(matte(0.1 + 0.8 * u(), 0.1 + 0.8 * u(), 0.1 + 0.8 * u()))
        else:
            s.materials.append(conductor(0.05 + 0.75 * u()))
        smat.append(len(s.materials) - 1)
    add_spheres(s, centers, radii, smat)
    s.camera = Camera(pos=(0.0, 7.0, -22.0), look=(0.0, 1.0, 0.0), up=(0.0, 1.0, 0.0), fov=50.0,
                      width=width, height=height)
    return s


def grid(origin, eu, ev, h: float) -> np.ndarray:
    """The rectangle origin + [0,1] eu + [0,1] ev as a grid of cells no larger than h, two triangles
    per cell (winding as quad())."""
    o, eu, ev = (np.asarray(x, np.float64) for x in (origin, eu, ev))
    nu = max(1, int(np.ceil(np.linalg.norm(eu) / h)))
    nv = max(1, int(np.ceil(np.linalg.norm(ev) / h)))
    iu, iv = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    iu, iv = iu.ravel()[:, None], iv.ravel()[:, None]
    p = lambda a, b: o + (a / nu) * eu + (b / nv) * ev  # noqa: E731
    p0, p1, p2, p3 = p(iu, iv), p(iu + 1, iv), p(iu + 1, iv + 1), p(iu, iv + 1)
    t = np.stack([np.concatenate([p0, p1, p2], 1), np.concatenate([p0, p2, p3], 1)], 1)
    return t.reshape(-1, 9).astype(np.float32)


def box_faces(lo, hi, h: float, skip=()) -> list[np.ndarray]:
    """The six faces of an axis-aligned box as grids (names in `skip` left out: -x, +x, -y, +y, -z, +z)."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    faces = {
        "-x": ((x0, y0, z0), (0, y1 - y0, 0), (0, 0, z1 - z0)),
        "+x": ((x1, y0, z0), (0, 0, z1 - z0), (0, y1 - y0, 0)),
        "-y": ((x0, y0, z0), (0, 0, z1 - z0), (x1 - x0, 0, 0)),
        "+y": ((x0, y1, z0), (x1 - x0, 0, 0), (0, 0, z1 - z0)),
        "-z": ((x0, y0, z0), (x1 - x0, 0, 0), (0, y1 - y0, 0)),
        "+z": ((x0, y0, z1), (0, y1 - y0, 0), (x1 - x0, 0, 0)),
    }
    return [grid(*faces[k], h) for k in faces if k not in skip]


AL_ETA = (1.657, 0.880, 0.521)
AL_K = (9.224, 6.270, 4.837)


def door_gap_room(width: int = 1920, height: int = 1080, cell: float = 0.0179, seed: int = 5,
                  L: float = 3000.0, wall_scale: float = 1.0, matte_walls: bool = False,
                  light_up: bool = False) -> SceneData:
    """Hard indirect interior: a 5 x 3 x 6 m room lit only through a 5 cm gap beside a nearly closed
    door from a neighbouring 2.9 x 3 x 6 m room whose ceiling holds the single quad area light
    (1 x 1 m, radiance L, facing down). The camera room's walls are glossy Trowbridge-Reitz
    conductors (alpha 0.1, 0.2, 0.3), its floor and ceiling matte; two displaced icospheres stand in
    it. Every surface is tessellated into cells of at most `cell` metres (0.0179 -> 1,513,386
    triangles; larger cells give small test scenes with the same lighting). `wall_scale` scales the
    glossy walls' reflectance; `matte_walls` makes them Lambertian instead. `light_up` turns the
    light toward the light room's ceiling, 15 cm below it."""
    s = SceneData()
    ws = wall_scale

    def wall(alpha, r):
        return matte(*r) if matte_walls else conductor(alpha, AL_ETA, AL_K, tuple(ws * c for c in r))

    s.materials = [matte(0.6, 0.6, 0.58),                               # 0 floor / ceiling (camera room)
                   wall(0.1, (0.75, 0.72, 0.68)),                       # 1 glossy wall
                   wall(0.2, (0.7, 0.75, 0.72)),                        # 2 glossy wall
                   wall(0.3, (0.72, 0.7, 0.75)),                        # 3 glossy wall / partition
                   matte(0.8, 0.78, 0.75),                              # 4 light room
                   matte(0.45, 0.3, 0.2),                               # 5 door
                   matte(0.0, 0.0, 0.0),                                # 6 light
                   conductor(0.15),                                      # 7 gold sculpture
                   matte(0.3, 0.5, 0.7)]                                # 8 matte sculpture
    parts, mats = [], []

    def add(tris, m):
        parts.append(tris)
        mats.append(np.full(len(tris), m, np.uint32))

    # camera room x in [0, 5], y in [0, 3], z in [0, 6]
    add(grid((0, 0, 0), (0, 0, 6), (5, 0, 0), cell), 0)                # floor
    add(grid((0, 3, 0), (5, 0, 0), (0, 0, 6), cell), 0)                # ceiling
    add(grid((0, 0, 0), (0, 3, 0), (0, 0, 6), cell), 1)                # x = 0 wall
    add(grid((0, 0, 0), (5, 0, 0), (0, 3, 0), cell), 2)                # z = 0 wall
    add(grid((0, 0, 6), (0, 3, 0), (5, 0, 0), cell), 2)                # z = 6 wall
    # partition x in [5, 5.1] with a door opening z in [2.5, 3.5], y in [0, 2.1]
    for x in (5.0, 5.1):
        m = 3 if x == 5.0 else 4
        add(grid((x, 0, 0), (0, 0, 2.5), (0, 3, 0), cell), m)
        add(grid((x, 0, 3.5), (0, 0, 2.5), (0, 3, 0), cell), m)
        add(grid((x, 2.1, 2.5), (0, 0, 1.0), (0, 0.9, 0), cell), m)
    add(grid((5.0, 0, 2.5), (0.1, 0, 0), (0, 2.1, 0), cell), 3)        # jambs and lintel
    add(grid((5.0, 0, 3.5), (0.1, 0, 0), (0, 2.1, 0), cell), 3)
    add(grid((5.0, 2.1, 2.5), (0.1, 0, 0), (0, 0, 1.0), cell), 3)
    # door panel leaves the gap z in [2.5, 2.55]
    for f in box_faces((5.02, 0.0, 2.55), (5.08, 2.09, 3.5), cell, skip=("-y",)):
        add(f, 5)
    # light room x in [5.1, 8]
    add(grid((5.1, 0, 0), (0, 0, 6), (2.9, 0, 0), cell), 4)
    add(grid((5.1, 3, 0), (2.9, 0, 0), (0, 0, 6), cell), 4)
    add(grid((8, 0, 0), (0, 0, 6), (0, 3, 0), cell), 4)
    add(grid((5.1, 0, 0), (2.9, 0, 0), (0, 3, 0), cell), 4)
    add(grid((5.1, 0, 6), (0, 3, 0), (2.9, 0, 0), cell), 4)
    # sculptures
    lv = 6 if cell < 0.05 else 2
    add(icosphere(lv, (2.2, 0.6, 3.8), 0.6, 0.1, 3.0, seed), 7)
    add(icosphere(lv, (3.4, 0.45, 1.6), 0.45, 0.1, 3.0, seed + 1), 8)
    base = sum(len(t) for t in parts)
    if light_up:  # facing the light room's ceiling: the camera room gets no direct light at all
        add(quad((6.0, 2.85, 2.5), (6.0, 2.85, 3.5), (7.0, 2.85, 3.5), (7.0, 2.85, 2.5)), 6)
    else:
        add(quad((6.0, 2.95, 2.5), (7.0, 2.95, 2.5), (7.0, 2.95, 3.5), (6.0, 2.95, 3.5)), 6)
    s.tri_positions = np.concatenate(parts).astype(np.float32)
    s.tri_material = np.concatenate(mats)
    s.lights = [area_light(base, (L, L, L)), area_light(base + 1, (L, L, L))]
    s.camera = Camera(pos=(0.6, 1.6, 0.8), look=(5.0, 0.9, 3.3), up=(0.0, 1.0, 0.0), fov=62.0,
                      width=width, height=height)
    return s


def batch_scene(index: int, seed: int = 77, scale: float = 1.0) -> SceneData:
    """Scene `index` of the seeded batch: a triangle budget log-uniform in [1k, 500k] (times `scale`)
    spent on displaced icospheres (largest level first), a matte ground quad, one double-sided
    quad light overhead, per object a matte (albedo U[0.1, 0.9]) or Trowbridge-Reitz conductor
    (alpha U[0.05, 0.8]) material, and a square film of side uniform in [128, 1024]
    (times sqrt(scale)), viewed from a seeded position."""
    k = 0

    def u():
        nonlocal k
        k += 1
        return _uniform(seed * 1000 + index, k)

    budget = int(np.exp(np.log(1e3) + (np.log(5e5) - np.log(1e3)) * u()) * scale)
    side = max(16, int((128 + (1024 - 128) * u()) * np.sqrt(scale)))
    s = SceneData()
    s.materials = [matte(0.5, 0.5, 0.5)]
    tris = [quad((-6, 0, -6), (6, 0, -6), (6, 0, 6), (-6, 0, 6))]
    mat = [0, 0]
    j = 0
    while budget >= 80:
        lv = 1
        while lv < 7 and 20 * 4 ** (lv + 1) <= budget:
            lv += 1
        r = 0.4 + 0.8 * u()
        c = (-4 + 8 * u(), r, -4 + 8 * u())
        m = icosphere(lv, c, r, 0.1, 3.0, seed * 7919 + index * 31 + j)
        if u() < 0.5:
            s.materials.append(matte(0.1 + 0.8 * u(), 0.1 + 0.8 * u(), 0.1 + 0.8 * u()))
        else:
            s.materials.append(conductor(0.05 + 0.75 * u()))
        tris.append(m)
        mat += [len(s.materials) - 1] * m.shape[0]
        budget -= m.shape[0]
        j += 1
    base = sum(t.shape[0] for t in tris)
    tris.append(quad((-1.5, 7, -1.5), (1.5, 7, -1.5), (1.5, 7, 1.5), (-1.5, 7, 1.5)))
    s.materials.append(matte(0.0, 0.0, 0.0))
    mat += [len(s.materials) - 1] * 2
    s.tri_positions = np.concatenate(tris).astype(np.float32)
    s.tri_material = np.asarray(mat, np.uint32)
    s.lights = [area_light(base, (20.0, 20.0, 18.0), True), area_light(base + 1, (20.0, 20.0, 18.0), True)]
    a = 2 * np.pi * u()
    s.camera = Camera(pos=(9 * np.cos(a), 4 + 3 * u(), 9 * np.sin(a)), look=(0.0, 0.8, 0.0), up=(0.0, 1.0, 0.0),
                      fov=45.0, width=side, height=side)
    return s


def intersection_camera(width: int = 4096, height: int = 4096) -> Camera:
    return Camera(pos=(0.0, 0.3, -3.2), look=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=45.0,
                  width=width, height=height)
