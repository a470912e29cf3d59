"""Scene descriptions for the wavefront path tracer, as flat numpy arrays.

A scene is a dict with
  vertices      (V, 3) float32
  indices       (T, 3) uint32
  normals       (V, 3) float32 (all zero = flat shading)
  uvs           (V, 2) float32
  material_ids  (T,)   int32
  materials     (M, 11) float32 rows {type, albedo rgb, emission rgb,
                two_sided, roughness, metallic, eta}
  camera        (8,)   float32 {origin xyz, look_at xyz, vertical fov deg, aspect}
Material types follow MaterialType (src/material/material.h:6-13):
0 Lambert, 1 ThinLambert, 2 GGX, 3 Dielectric, 4 Interface, 5 Transparent.
No coordinate is -0.0 (every array goes through `+ 0.0`).
"""

import numpy as np

LAMBERT, THIN_LAMBERT, GGX, DIELECTRIC, INTERFACE, TRANSPARENT = range(6)


def material(kind, albedo=(0.8, 0.8, 0.8), emission=(0, 0, 0), two_sided=0, roughness=0.5, metallic=0.0,
             eta=1.5):
    return [kind, *albedo, *emission, two_sided, roughness, metallic, eta]


def _pack(tris, mat_ids, materials, camera, normals=None):
    tris = np.asarray(tris, np.float32).reshape(-1, 9)
    t = len(tris)
    verts = tris.reshape(-1, 3) + np.float32(0.0)
    return {
        "vertices": np.ascontiguousarray(verts, np.float32),
        "indices": np.arange(3 * t, dtype=np.uint32).reshape(t, 3),
        "normals": np.zeros_like(verts) if normals is None else normals,
        "uvs": np.zeros((3 * t, 2), np.float32),
        "material_ids": np.asarray(mat_ids, np.int32),
        "materials": np.asarray(materials, np.float32),
        "camera": np.asarray(camera, np.float32),
    }


def cornell(width=256, height=256):
    """The built-in Cornell box (make_cornell_box, src/scene/cornell_box.h:31-56)
    with its CLI camera (src/app/render_runner.cpp:85-90): 36 triangles, one
    quad area light (15, 14, 12)."""
    f = np.float32
    materials = [
        material(LAMBERT, (0.78, 0.78, 0.74)),
        material(LAMBERT, (0.65, 0.05, 0.05)),
        material(LAMBERT, (0.12, 0.45, 0.15)),
        material(LAMBERT, (0, 0, 0), emission=(15.0, 14.0, 12.0)),
    ]
    white, red, green, light = 0, 1, 2, 3
    tris, mats = [], []

    def quad(p0, p1, p2, p3, m):
        tris.append([*p0, *p1, *p2])
        tris.append([*p0, *p2, *p3])
        mats.extend([m, m])

    def box(lo, hi, m):
        x0, y0, z0 = lo
        x1, y1, z1 = hi
        p000, p001, p010, p011 = (x0, y0, z0), (x0, y0, z1), (x0, y1, z0), (x0, y1, z1)
        p100, p101, p110, p111 = (x1, y0, z0), (x1, y0, z1), (x1, y1, z0), (x1, y1, z1)
        quad(p000, p100, p110, p010, m)
        quad(p101, p001, p011, p111, m)
        quad(p001, p000, p010, p011, m)
        quad(p100, p101, p111, p110, m)
        quad(p010, p110, p111, p011, m)
        quad(p001, p101, p100, p000, m)

    quad((-1, 0, 2), (1, 0, 2), (1, 0, 0), (-1, 0, 0), white)
    quad((-1, 2, 0), (1, 2, 0), (1, 2, 2), (-1, 2, 2), white)
    quad((-1, 0, 0), (1, 0, 0), (1, 2, 0), (-1, 2, 0), white)
    quad((-1, 0, 2), (-1, 0, 0), (-1, 2, 0), (-1, 2, 2), red)
    quad((1, 0, 0), (1, 0, 2), (1, 2, 2), (1, 2, 0), green)
    box((f(-0.65), 0, f(0.35)), (f(-0.15), f(0.8), f(0.9)), white)
    box((f(0.25), 0, f(0.75)), (f(0.75), f(1.2), f(1.35)), white)
    quad((f(-0.35), f(1.99), f(0.75)), (f(0.35), f(1.99), f(0.75)), (f(0.35), f(1.99), f(1.35)),
         (f(-0.35), f(1.99), f(1.35)), light)
    aspect = np.float32(width) / np.float32(height)
    return _pack(tris, mats, materials, [0, 1, f(3.35), 0, 1, f(0.9), 40.0, aspect])


def mixed_materials(width=40, height=40):
    """Every material type and both emitter kinds in the Cornell room: a thin
    Lambert left wall, a transparent sheet and an interface panel (pass-through
    surfaces, also for shadow rays), a metallic GGX box, a rough dielectric-free
    GGX box, a glass box, the single-sided ceiling light and a two-sided
    emitting panel hanging in the room."""
    f = np.float32
    materials = [
        material(LAMBERT, (0.78, 0.78, 0.74)),                                          # 0 white
        material(THIN_LAMBERT, (0.65, 0.10, 0.08)),                                     # 1 thin red wall
        material(LAMBERT, (0.12, 0.45, 0.15)),                                          # 2 green
        material(LAMBERT, (0, 0, 0), emission=(15.0, 14.0, 12.0)),                      # 3 ceiling light
        material(LAMBERT, (0, 0, 0), emission=(2.0, 3.0, 6.0), two_sided=1),            # 4 two-sided panel
        material(TRANSPARENT, (0.6, 0.8, 0.9)),                                         # 5 tinted sheet
        material(INTERFACE, (1.0, 1.0, 1.0)),                                           # 6 interface panel
        material(GGX, (0.9, 0.7, 0.4), roughness=f(0.2), metallic=f(1.0)),              # 7 gold-ish metal
        material(GGX, (0.3, 0.5, 0.8), roughness=f(0.45), metallic=f(0.0)),             # 8 rough plastic
        material(DIELECTRIC, (1.0, 1.0, 1.0), roughness=0.0, eta=f(1.5)),               # 9 glass
    ]
    tris, mats = [], []

    def quad(p0, p1, p2, p3, m):
        tris.append([*p0, *p1, *p2])
        tris.append([*p0, *p2, *p3])
        mats.extend([m, m])

    def box(lo, hi, m):
        x0, y0, z0 = lo
        x1, y1, z1 = hi
        p000, p001, p010, p011 = (x0, y0, z0), (x0, y0, z1), (x0, y1, z0), (x0, y1, z1)
        p100, p101, p110, p111 = (x1, y0, z0), (x1, y0, z1), (x1, y1, z0), (x1, y1, z1)
        quad(p000, p100, p110, p010, m)
        quad(p101, p001, p011, p111, m)
        quad(p001, p000, p010, p011, m)
        quad(p100, p101, p111, p110, m)
        quad(p010, p110, p111, p011, m)
        quad(p001, p101, p100, p000, m)

    quad((-1, 0, 2), (1, 0, 2), (1, 0, 0), (-1, 0, 0), 0)
    quad((-1, 2, 0), (1, 2, 0), (1, 2, 2), (-1, 2, 2), 0)
    quad((-1, 0, 0), (1, 0, 0), (1, 2, 0), (-1, 2, 0), 0)
    quad((-1, 0, 2), (-1, 0, 0), (-1, 2, 0), (-1, 2, 2), 1)
    quad((1, 0, 0), (1, 0, 2), (1, 2, 2), (1, 2, 0), 2)
    box((f(-0.7), 0, f(0.3)), (f(-0.3), f(0.7), f(0.7)), 7)
    box((f(0.3), 0, f(0.9)), (f(0.7), f(0.5), f(1.3)), 9)
    box((f(-0.2), 0, f(0.15)), (f(0.2), f(0.3), f(0.45)), 8)
    quad((f(-0.9), f(0.05), f(1.6)), (f(-0.1), f(0.05), f(1.6)), (f(-0.1), f(1.1), f(1.6)),
         (f(-0.9), f(1.1), f(1.6)), 5)
    quad((f(0.1), f(0.6), f(1.5)), (f(0.9), f(0.6), f(1.5)), (f(0.9), f(1.4), f(1.5)), (f(0.1), f(1.4), f(1.5)), 6)
    quad((f(-0.35), f(1.99), f(0.75)), (f(0.35), f(1.99), f(0.75)), (f(0.35), f(1.99), f(1.35)),
         (f(-0.35), f(1.99), f(1.35)), 3)
    quad((f(0.2), f(1.2), f(0.4)), (f(0.6), f(1.2), f(0.4)), (f(0.6), f(1.6), f(0.4)), (f(0.2), f(1.6), f(0.4)), 4)
    aspect = np.float32(width) / np.float32(height)
    return _pack(tris, mats, materials, [0, 1, f(3.35), 0, 1, f(0.9), 40.0, aspect])


def probe(kind, width=256, height=256):
    """The reference CLI's light demo scenes (make_point_light_scene,
    make_distant_light_scene, make_uniform_env_scene, make_image_env_scene,
    src/app/render_runner.cpp:99-138): a 2.4 x 2.4 grey Lambert quad at y = 0
    seen from (0, 0.8, 2.8), lit by one scene light: "point" (25 at
    (0, 1.6, 0.2)), "distant" (1.2 from (-0.4, 1, 0.3)), "uniform_env"
    (radiance (0.15, 0.18, 0.25)) or "image_env" (a 2x2 lat-long image scaled
    by 0.35)."""
    f = np.float32
    materials = [material(LAMBERT, (0.75, 0.75, 0.75))]
    q = [(f(-1.2), 0, f(1.2)), (f(1.2), 0, f(1.2)), (f(1.2), 0, f(-1.2)), (f(-1.2), 0, f(-1.2))]
    tris = [[*q[0], *q[1], *q[2]], [*q[0], *q[2], *q[3]]]
    aspect = np.float32(width) / np.float32(height)
    d = _pack(tris, [0, 0], materials, [0, f(0.8), f(2.8), 0, 0, 0, 35.0, aspect])
    lights = {
        "point": {"kind": "point", "vector": (0.0, 1.6, 0.2), "color": (25.0, 25.0, 25.0)},
        "distant": {"kind": "distant", "vector": (-0.4, 1.0, 0.3), "color": (1.2, 1.2, 1.2)},
        "uniform_env": {"kind": "uniform_infinite", "color": (0.15, 0.18, 0.25)},
        "image_env": {"kind": "image_infinite", "color": (0.35, 0.35, 0.35),
                      "image": np.asarray([[[1.4, 1.0, 0.7], [0.2, 0.3, 0.8]],
                                           [[0.1, 0.2, 0.5], [0.8, 1.1, 1.5]]], np.float32)},
    }
    d["lights"] = [lights[kind]]
    return d


def many_lights(width=40, height=40, n_lights=256, seed=0x4C5561):
    """A many-light scene for the light samplers: the Cornell room without its
    ceiling lamp, a Lambert box, a GGX box and a glass box, lit by n_lights
    small emissive triangles (one-sided or two-sided, emission log-uniform
    over two decades in random colours) scattered under the ceiling and along
    the back wall."""
    f = np.float32
    rng = np.random.default_rng(seed)
    materials = [
        material(LAMBERT, (0.78, 0.78, 0.74)),
        material(LAMBERT, (0.65, 0.05, 0.05)),
        material(LAMBERT, (0.12, 0.45, 0.15)),
        material(GGX, (0.9, 0.7, 0.4), roughness=f(0.25), metallic=f(1.0)),
        material(DIELECTRIC, (1.0, 1.0, 1.0), roughness=0.0, eta=f(1.5)),
    ]
    tris, mats = [], []

    def quad(p0, p1, p2, p3, m):
        tris.append([*p0, *p1, *p2])
        tris.append([*p0, *p2, *p3])
        mats.extend([m, m])

    def box(lo, hi, m):
        x0, y0, z0 = lo
        x1, y1, z1 = hi
        p000, p001, p010, p011 = (x0, y0, z0), (x0, y0, z1), (x0, y1, z0), (x0, y1, z1)
        p100, p101, p110, p111 = (x1, y0, z0), (x1, y0, z1), (x1, y1, z0), (x1, y1, z1)
        quad(p000, p100, p110, p010, m)
        quad(p101, p001, p011, p111, m)
        quad(p001, p000, p010, p011, m)
        quad(p100, p101, p111, p110, m)
        quad(p010, p110, p111, p011, m)
        quad(p001, p101, p100, p000, m)

    quad((-1, 0, 2), (1, 0, 2), (1, 0, 0), (-1, 0, 0), 0)
    quad((-1, 2, 0), (1, 2, 0), (1, 2, 2), (-1, 2, 2), 0)
    quad((-1, 0, 0), (1, 0, 0), (1, 2, 0), (-1, 2, 0), 0)
    quad((-1, 0, 2), (-1, 0, 0), (-1, 2, 0), (-1, 2, 2), 1)
    quad((1, 0, 0), (1, 0, 2), (1, 2, 2), (1, 2, 0), 2)
    box((f(-0.65), 0, f(0.35)), (f(-0.15), f(0.8), f(0.9)), 0)
    box((f(0.25), 0, f(0.75)), (f(0.75), f(0.6), f(1.35)), 3)
    box((f(-0.1), 0, f(1.2)), (f(0.2), f(0.3), f(1.5)), 4)
    for k in range(n_lights):
        on_ceiling = k % 3 != 0
        if on_ceiling:  # facing down, just under the ceiling
            c = np.array([rng.uniform(-0.9, 0.9), 1.98, rng.uniform(0.1, 1.9)])
            e1, e2 = np.array([1.0, 0, 0]), np.array([0, 0, 1.0])
        else:  # facing the camera, just in front of the back wall
            c = np.array([rng.uniform(-0.9, 0.9), rng.uniform(0.2, 1.8), 0.02])
            e1, e2 = np.array([0, 1.0, 0]), np.array([1.0, 0, 0])
        s = rng.uniform(0.02, 0.08)
        a = rng.uniform(0, 2 * np.pi)
        d1 = np.cos(a) * e1 + np.sin(a) * e2
        d2 = -np.sin(a) * e1 + np.cos(a) * e2
        p0, p1, p2 = c - s * d1 - s * d2, c + s * d1 - s * d2, c + s * d2
        tris.append([*p0, *p1, *p2])
        color = rng.uniform(0.2, 1.0, 3)
        power = np.exp(rng.uniform(np.log(2.0), np.log(200.0)))
        materials.append(material(LAMBERT, (0, 0, 0), emission=tuple(color * power),
                                  two_sided=int(rng.uniform() < 0.25)))
        mats.append(len(materials) - 1)
    aspect = np.float32(width) / np.float32(height)
    return _pack(np.asarray(tris, np.float32), mats, materials, [0, 1, f(3.35), 0, 1, f(0.9), 40.0, aspect])


# ---- instanced scenes ------------------------------------------------------------

def instanced_room(width=40, height=40, pass_through=True):
    """Meshes placed by affine instances (an instanced scene dict: the mesh
    arrays concatenated with "vertex_offsets" / "triangle_offsets", mesh-local
    indices, and "inst_mesh", "inst_m", "inst_override"):
      mesh 0  the Cornell room (walls + ceiling light), identity instance
      mesh 1  a unit geodesic sphere (freq 4, 320 triangles) with smooth
              normals: uniformly scaled (plastic), non-uniformly scaled and
              rotated (metal override) and mirrored (glass override)
      mesh 2  a unit cube (flat): a rotated block, a thin two-sided emitting
              slab (emitter override: an instanced light, non-uniform scale)
              and, with pass_through, a tinted transparent sheet in front
    """
    from .workloads import affine

    f = np.float32
    materials = [
        material(LAMBERT, (0.78, 0.78, 0.74)),                                  # 0 white
        material(LAMBERT, (0.65, 0.05, 0.05)),                                  # 1 red
        material(LAMBERT, (0.12, 0.45, 0.15)),                                  # 2 green
        material(LAMBERT, (0, 0, 0), emission=(15.0, 14.0, 12.0)),              # 3 ceiling light
        material(GGX, (0.3, 0.5, 0.8), roughness=f(0.45), metallic=f(0.0)),     # 4 plastic
        material(GGX, (0.9, 0.7, 0.4), roughness=f(0.2), metallic=f(1.0)),      # 5 metal
        material(DIELECTRIC, (1.0, 1.0, 1.0), roughness=0.0, eta=f(1.5)),       # 6 glass
        material(LAMBERT, (0, 0, 0), emission=(2.0, 3.0, 6.0), two_sided=1),    # 7 two-sided emitter
        material(TRANSPARENT, (0.6, 0.8, 0.9)),                                 # 8 tinted sheet
    ]
    room, room_m = [], []

    def quad(p0, p1, p2, p3, m):
        room.append([*p0, *p1, *p2])
        room.append([*p0, *p2, *p3])
        room_m.extend([m, m])

    quad((-1, 0, 2), (1, 0, 2), (1, 0, 0), (-1, 0, 0), 0)
    quad((-1, 2, 0), (1, 2, 0), (1, 2, 2), (-1, 2, 2), 0)
    quad((-1, 0, 0), (1, 0, 0), (1, 2, 0), (-1, 2, 0), 0)
    quad((-1, 0, 2), (-1, 0, 0), (-1, 2, 0), (-1, 2, 2), 1)
    quad((1, 0, 0), (1, 0, 2), (1, 2, 2), (1, 2, 0), 2)
    quad((f(-0.35), f(1.99), f(0.75)), (f(0.35), f(1.99), f(0.75)), (f(0.35), f(1.99), f(1.35)),
         (f(-0.35), f(1.99), f(1.35)), 3)
    room_v = np.asarray(room, np.float32).reshape(-1, 3)
    room_i = np.arange(len(room_v), dtype=np.uint32).reshape(-1, 3)
    sv, si = geodesic_sphere(4)
    sv = sv.astype(np.float32)
    c = np.array([[x, y, z] for z in (-0.5, 0.5) for y in (-0.5, 0.5) for x in (-0.5, 0.5)], np.float32)
    cube_i = np.array([[0, 2, 3], [0, 3, 1], [4, 5, 7], [4, 7, 6], [0, 1, 5], [0, 5, 4],
                       [2, 6, 7], [2, 7, 3], [0, 4, 6], [0, 6, 2], [1, 3, 7], [1, 7, 5]], np.uint32)
    verts = np.concatenate([room_v, sv, c]) + np.float32(0.0)
    normals = np.concatenate([np.zeros_like(room_v), sv, np.zeros_like(c)]) + np.float32(0.0)
    indices = np.concatenate([room_i, si.astype(np.uint32), cube_i])
    mat_ids = np.concatenate([room_m, np.full(len(si), 4), np.zeros(len(cube_i))]).astype(np.int32)
    vo = np.cumsum([0, len(room_v), len(sv), len(c)]).astype(np.int32)
    to = np.cumsum([0, len(room_i), len(si), len(cube_i)]).astype(np.int32)

    def ry(deg):
        a = np.radians(deg) / 2
        return (np.cos(a), 0.0, np.sin(a), 0.0)

    def rx(deg):
        a = np.radians(deg) / 2
        return (np.cos(a), np.sin(a), 0.0, 0.0)

    inst = [
        (0, affine(), -1),
        (1, affine(ry(30), (0.3, 0.3, 0.3), (-0.45, 0.3, 0.6)), -1),
        (1, affine(ry(-20), (0.35, 0.2, 0.25), (0.45, 0.22, 1.1)), 5),
        (1, affine(None, (-0.18, 0.18, 0.18), (0.0, 0.9, 1.2)), 6),
        (2, affine(ry(45), (0.3, 0.7, 0.3), (-0.5, 0.35, 1.35)), -1),
        (2, affine(rx(30), (0.25, 0.02, 0.25), (0.5, 1.3, 0.5)), 7),
    ]
    if pass_through:
        inst.append((2, affine(None, (0.8, 0.8, 0.01), (0.0, 0.5, 1.7)), 8))
    aspect = np.float32(width) / np.float32(height)
    return {
        "vertices": np.ascontiguousarray(verts, np.float32),
        "indices": np.ascontiguousarray(indices, np.uint32),
        "normals": np.ascontiguousarray(normals, np.float32),
        "uvs": np.zeros((len(verts), 2), np.float32),
        "material_ids": mat_ids,
        "materials": np.asarray(materials, np.float32),
        "camera": np.asarray([0, 1, f(3.35), 0, 1, f(0.9), 40.0, aspect], np.float32),
        "vertex_offsets": vo,
        "triangle_offsets": to,
        "inst_mesh": np.array([k for k, _, _ in inst], np.int32),
        "inst_m": np.stack([m for _, m, _ in inst]).astype(np.float32),
        "inst_override": np.array([o for _, _, o in inst], np.int32),
    }


def instanced_field(width=1024, height=1024, n_meshes=4, freq=64, grid=8, seed=0x4C5560):
    """A field of grid x grid instanced displaced icospheres over a ground plane
    (an instanced scene dict, see instanced_room): n_meshes distinct spheres
    (freq -> 20 freq^2 triangles each, smooth normals, the 48-material
    Lambert / GGX / dielectric palette of displaced_icosphere), each instance a
    random rotation, uniform scale U(0.35, 0.6) and position on the grid, every
    fourth with a palette material override; mesh 0 holds the 12 x 12 ground
    plane and a 4 x 4 ceiling light at y = 6 (emission (6, 6, 5.5)) under an
    identity instance. Camera (0, 5, 9) -> (0, -0.5, 0), vertical fov 40."""
    from .workloads import affine

    spheres = [displaced_icosphere(freq=freq, seed=seed + k, width=width, height=height) for k in range(n_meshes)]
    palette = spheres[0]["materials"][:-1]  # drop the sphere scene's own light material
    ground, light = len(palette), len(palette) + 1
    materials = np.concatenate([palette, np.asarray([material(LAMBERT, (0.7, 0.7, 0.68)),
                                                     material(LAMBERT, (0, 0, 0), emission=(6.0, 6.0, 5.5))],
                                                    np.float32)])
    g0 = np.array([[-6, -1.05, -6], [6, -1.05, -6], [6, -1.05, 6], [-6, -1.05, 6]], np.float32)
    l0 = np.array([[-2, 6, -2], [2, 6, -2], [2, 6, 2], [-2, 6, 2]], np.float32)
    mesh0_v = np.concatenate([g0, l0])
    mesh0_i = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7]], np.uint32)  # ground up, light down
    verts, normals, indices, mat_ids = [mesh0_v], [np.zeros_like(mesh0_v)], [mesh0_i], [[ground, ground, light, light]]
    for sp in spheres:  # each sphere mesh without its light quad
        nv = len(sp["vertices"]) - 4
        verts.append(sp["vertices"][:nv])
        normals.append(sp["normals"][:nv])
        indices.append(sp["indices"][:-2])
        mat_ids.append(sp["material_ids"][:-2])
    vo = np.cumsum([0] + [len(v) for v in verts]).astype(np.int32)
    to = np.cumsum([0] + [len(i) for i in indices]).astype(np.int32)
    rng = np.random.default_rng(seed)
    inst_mesh, inst_m, inst_ov = [0], [affine()], [-1]
    step = 10.0 / grid
    for gz in range(grid):
        for gx in range(grid):
            k = len(inst_mesh) - 1
            sc = rng.uniform(0.35, 0.6)
            pos = (-5 + step * (gx + 0.5) + rng.uniform(-0.1, 0.1), -1.05 + sc, -5 + step * (gz + 0.5))
            inst_mesh.append(1 + int(rng.integers(0, n_meshes)))
            inst_m.append(affine(rng.standard_normal(4), (sc, sc, sc), pos))
            inst_ov.append(int(rng.integers(0, len(palette))) if k % 4 == 3 else -1)
    aspect = np.float32(width) / np.float32(height)
    v = np.concatenate(verts)
    return {
        "vertices": np.ascontiguousarray(v, np.float32) + np.float32(0.0),
        "indices": np.ascontiguousarray(np.concatenate(indices), np.uint32),
        "normals": np.ascontiguousarray(np.concatenate(normals), np.float32) + np.float32(0.0),
        "uvs": np.zeros((len(v), 2), np.float32),
        "material_ids": np.concatenate(mat_ids).astype(np.int32),
        "materials": materials.astype(np.float32),
        "camera": np.asarray([0, 5, 9, 0, -0.5, 0, 40.0, aspect], np.float32),
        "vertex_offsets": vo,
        "triangle_offsets": to,
        "inst_mesh": np.asarray(inst_mesh, np.int32),
        "inst_m": np.stack(inst_m).astype(np.float32),
        "inst_override": np.asarray(inst_ov, np.int32),
    }


def bunny(obj_path, width=512, height=512, loader=None):
    """The reference's "bunny" scene (make_stanford_bunny_scene,
    src/scene/stanford_bunny.h:15-56, camera of make_bunny_scene_from_config,
    src/app/render_runner.cpp:148-154) as an instanced scene dict: mesh 0 holds
    the grey floor quad at y = 0 and the ceiling light quad at y = 0.5 under an
    identity instance; mesh 1 is the OBJ imported with default options and
    ground_to_y (its instance grounds it at y = 0), shaded with the default
    material turned ThinLambert(0.8). `loader(path, **options)` returns the
    imported mesh dict (lux_amd.obj.load by default)."""
    if loader is None:
        from .obj import load as loader
    m = loader(obj_path, scene_material_count=2, ground_to_y=True)
    materials = np.asarray([material(LAMBERT, (0.5, 0.5, 0.5)),
                            material(LAMBERT, (0, 0, 0), emission=(15.0, 14.0, 12.0)),
                            material(THIN_LAMBERT, (0.8, 0.8, 0.8))], np.float32)
    f = np.float32
    quads = [[(-1, 0, 1), (1, 0, 1), (1, 0, -1), (-1, 0, -1)],
             [(f(-0.3), f(0.5), f(-0.3)), (f(0.3), f(0.5), f(-0.3)), (f(0.3), f(0.5), f(0.3)),
              (f(-0.3), f(0.5), f(0.3))]]
    tris = []
    for q in quads:  # add_quad: (p0, p1, p2), (p0, p2, p3)
        tris += [[q[0], q[1], q[2]], [q[0], q[2], q[3]]]
    mesh0_v = np.asarray(tris, np.float32).reshape(-1, 3)
    mesh0_i = np.arange(len(mesh0_v), dtype=np.uint32).reshape(-1, 3)
    verts = np.concatenate([mesh0_v, m["vertices"]])
    aspect = np.float32(width) / np.float32(height)
    return {
        "vertices": np.ascontiguousarray(verts, np.float32),
        "indices": np.ascontiguousarray(np.concatenate([mesh0_i, m["indices"]]), np.uint32),
        "normals": np.ascontiguousarray(np.concatenate([np.zeros_like(mesh0_v), m["normals"]]), np.float32),
        "uvs": np.ascontiguousarray(np.concatenate([np.zeros((len(mesh0_v), 2), np.float32), m["uvs"]])),
        "material_ids": np.concatenate([[0, 0, 1, 1], m["material_ids"]]).astype(np.int32),
        "materials": materials,
        "camera": np.asarray([0, f(0.3), f(0.5), f(-0.02), f(0.11), 0, 40.0, aspect], np.float32),
        "vertex_offsets": np.array([0, len(mesh0_v), len(verts)], np.int32),
        "triangle_offsets": np.array([0, len(mesh0_i), len(mesh0_i) + len(m["indices"])], np.int32),
        "inst_mesh": np.array([0, 1], np.int32),
        "inst_m": np.stack([np.eye(4, dtype=np.float32), m["object_to_world"]]).astype(np.float32),
        "inst_inv": np.stack([np.eye(4, dtype=np.float32), m["world_to_object"]]).astype(np.float32),
        "inst_override": np.array([-1, -1], np.int32),
    }


def flattened(view, materials, camera):
    """A one-mesh scene dict holding an instanced scene's shading view (world
    corners, world corner normals, materials after the override: the
    renderer's Scene.triangles()), for comparing one-level and two-level
    traversal on the same geometry."""
    t = len(view["material"])
    return {
        "vertices": np.ascontiguousarray(view["corners"].reshape(-1, 3), np.float32),
        "indices": np.arange(3 * t, dtype=np.uint32).reshape(t, 3),
        "normals": np.ascontiguousarray(view["normals"].reshape(-1, 3), np.float32),
        "uvs": np.zeros((3 * t, 2), np.float32),
        "material_ids": np.ascontiguousarray(view["material"], np.int32),
        "materials": np.asarray(materials, np.float32),
        "camera": np.asarray(camera, np.float32),
    }


# ---- seeded displaced icosphere -----------------------------------------------

def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    return v, f


def geodesic_sphere(freq):
    """Unit geodesic icosphere of frequency `freq`: 20 freq^2 triangles,
    10 freq^2 + 2 shared vertices."""
    base, faces = _icosahedron()
    n = freq
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    keep = ii + jj <= n
    gi, gj = ii[keep], jj[keep]  # grid points of one face
    local = -np.ones((n + 1, n + 1), np.int64)
    local[gi, gj] = np.arange(len(gi))
    # triangles of one face in local grid numbering
    up_i, up_j = np.nonzero((ii + jj) <= n - 1)
    up = np.stack([local[up_i, up_j], local[up_i + 1, up_j], local[up_i, up_j + 1]], 1)
    dn_i, dn_j = np.nonzero((ii + jj) <= n - 2)
    dn = np.stack([local[dn_i + 1, dn_j], local[dn_i + 1, dn_j + 1], local[dn_i, dn_j + 1]], 1)
    tri_local = np.concatenate([up, dn])
    pts, tris = [], []
    for fi, (a, b, c) in enumerate(faces):
        A, B, C = base[a], base[b], base[c]
        p = A + (B - A) * (gi / n)[:, None] + (C - A) * (gj / n)[:, None]
        pts.append(p / np.linalg.norm(p, axis=1, keepdims=True))
        tris.append(tri_local + fi * len(gi))
    pts = np.concatenate(pts)
    tris = np.concatenate(tris)
    key = np.round(pts * 1e9).astype(np.int64)
    _, first, inverse = np.unique(key, axis=0, return_index=True, return_inverse=True)
    return pts[first], inverse.reshape(-1)[tris]


def _lattice_hash(ix, iy, iz, seed):
    h = (ix.astype(np.uint64) * np.uint64(0x8DA6B343) ^ iy.astype(np.uint64) * np.uint64(0xD8163841)
         ^ iz.astype(np.uint64) * np.uint64(0xCB1AB31F) ^ np.uint64(seed))
    h ^= h >> np.uint64(13)
    h *= np.uint64(0x5BD1E995)
    h &= np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(15)
    return (h & np.uint64(0xFFFFFF)).astype(np.float64) / float(0xFFFFFF) * 2.0 - 1.0


def value_noise(p, seed, octaves=3, base_freq=4.0, lacunarity=2.0, gain=0.5):
    """Trilinear value noise with smoothstep weights, summed over octaves and
    normalised to [-1, 1]."""
    total = np.zeros(len(p))
    amp, freq, norm = 1.0, base_freq, 0.0
    for o in range(octaves):
        q = p * freq + 1000.0
        i0 = np.floor(q).astype(np.int64)
        f = q - i0
        w = f * f * (3 - 2 * f)
        acc = np.zeros(len(p))
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    v = _lattice_hash(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, seed + o)
                    wx = w[:, 0] if dx else 1 - w[:, 0]
                    wy = w[:, 1] if dy else 1 - w[:, 1]
                    wz = w[:, 2] if dz else 1 - w[:, 2]
                    acc += v * wx * wy * wz
        total += amp * acc
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


def displaced_icosphere(freq=208, seed=0x4C5559, width=1024, height=1024, palette=48, amplitude=0.05):
    """Seeded icosphere (20 freq^2 triangles; freq 208 -> 865 280) with radial
    displacement amplitude * noise(p), smooth vertex normals, a seeded-hash mix of
    Lambert / GGX (roughness U(0.05, 0.5)) / dielectric (IOR 1.5) materials and
    one 1x1 quad area light at y = 2.5 facing down, emission (15, 14, 12);
    camera (0, 0, 3.5) -> origin, vertical fov 40 degrees."""
    pts, tris = geodesic_sphere(freq)
    r = 1.0 + amplitude * value_noise(pts, seed)
    verts = pts * r[:, None]
    v0, v1, v2 = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    fn = np.cross(v1 - v0, v2 - v0)
    # make faces point outwards
    flip = np.einsum("ij,ij->i", fn, v0 + v1 + v2) < 0
    tris[flip] = tris[flip][:, [0, 2, 1]]
    v0, v1, v2 = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    fn = np.cross(v1 - v0, v2 - v0)
    vn = np.zeros_like(verts)
    for k in range(3):
        np.add.at(vn, tris[:, k], fn)
    vn /= np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-30)

    rng = np.random.default_rng(seed)
    mats = []
    for k in range(palette):
        kind = k % 3
        if kind == 0:
            mats.append(material(LAMBERT, tuple(rng.uniform(0.2, 0.8, 3))))
        elif kind == 1:
            mats.append(material(GGX, tuple(rng.uniform(0.2, 0.8, 3)), roughness=rng.uniform(0.05, 0.5)))
        else:
            mats.append(material(DIELECTRIC, (1.0, 1.0, 1.0), roughness=0.0, eta=1.5))
    light_mat = len(mats)
    mats.append(material(LAMBERT, (0, 0, 0), emission=(15.0, 14.0, 12.0)))
    h = (np.arange(len(tris), dtype=np.uint64) * np.uint64(0x9E3779B1) ^ np.uint64(seed)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x7FEB352D)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(15)
    mat_ids = (h % np.uint64(palette)).astype(np.int32)

    # light quad at y = 2.5, normal (0, -1, 0)
    lq = np.array([[-0.5, 2.5, -0.5], [0.5, 2.5, -0.5], [0.5, 2.5, 0.5], [-0.5, 2.5, 0.5]], np.float64)
    n_v = len(verts)
    verts = np.concatenate([verts, lq])
    vn = np.concatenate([vn, np.zeros((4, 3))])
    tris = np.concatenate([tris, np.array([[n_v, n_v + 1, n_v + 2], [n_v, n_v + 2, n_v + 3]])])
    mat_ids = np.concatenate([mat_ids, [light_mat, light_mat]]).astype(np.int32)
    aspect = np.float32(width) / np.float32(height)
    return {
        "vertices": np.ascontiguousarray(verts.astype(np.float32) + np.float32(0.0)),
        "indices": np.ascontiguousarray(tris.astype(np.uint32)),
        "normals": np.ascontiguousarray(vn.astype(np.float32) + np.float32(0.0)),
        "uvs": np.zeros((len(verts), 2), np.float32),
        "material_ids": mat_ids,
        "materials": np.asarray(mats, np.float32),
        "camera": np.asarray([0, 0, 3.5, 0, 0, 0, 40.0, aspect], np.float32),
    }


def icosphere_freq_for_triangles(t):
    return max(7, int(round((t / 20.0) ** 0.5)))


def batched_mix(count=64, seed=0x4C555A):
    """The multi-scene mix: triangle counts log-uniform in [1e3, 5e5], image sides
    in {128, 256, 512, 1024}, spp log-uniform in [16, 256], depth 6. Returns scene
    specs (generate each with displaced_icosphere)."""
    specs = []
    for i in range(count):
        rng = np.random.default_rng(seed + i)
        t = float(np.exp(rng.uniform(np.log(1e3), np.log(5e5))))
        side = int(rng.choice([128, 256, 512, 1024]))
        spp = int(round(float(np.exp(rng.uniform(np.log(16), np.log(256))))))
        specs.append({"seed": seed + i, "freq": icosphere_freq_for_triangles(t), "width": side, "height": side,
                      "spp": spp, "depth": 6})
    return specs
