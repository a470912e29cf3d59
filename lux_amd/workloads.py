"""Seeded synthetic inputs of the LBVH microbenchmark (triangle soups, rays,
vertex jitter, overlap query boxes). Pure numpy; float32 throughout.

Triangle soup: v0 ~ U[0,1]^3, v1 = v0 + L1*u1, v2 = v0 + L2*u2 with edge
lengths L ~ U(0.001, 0.01) and u uniform on the unit sphere (normalised
Gaussian). Rays: origins ~ U[0,1]^3, directions uniform on the sphere,
t_min = 0, t_max = +inf, time = 0.
"""

import numpy as np

SEED_SOUP = 0x4C5558  # refit jitter uses +1, rays +2, query soup +3


def _unit(rng, n):
    g = rng.standard_normal((n, 3))
    return g / np.linalg.norm(g, axis=1, keepdims=True)


def soup(n, seed=SEED_SOUP, edge=(0.001, 0.01)):
    rng = np.random.default_rng(seed)
    v0 = rng.random((n, 3))
    l1 = rng.uniform(edge[0], edge[1], n)[:, None]
    l2 = rng.uniform(edge[0], edge[1], n)[:, None]
    v1 = v0 + l1 * _unit(rng, n)
    v2 = v0 + l2 * _unit(rng, n)
    return np.concatenate([v0, v1, v2], axis=1).astype(np.float32)


def clustered_soup(n, seed=7, grid=6):
    """Soup whose vertices snap to a coarse grid, so many centroids (and Morton
    codes) coincide and some triangles are exact duplicates (exact-t ties)."""
    rng = np.random.default_rng(seed)
    cells = rng.integers(0, grid, size=(n, 3)) / float(grid)
    jit = rng.integers(0, 3, size=(n, 6)) * 0.01
    v0 = cells
    v1 = cells + np.stack([jit[:, 0] + 0.02, jit[:, 1], jit[:, 2]], axis=1)
    v2 = cells + np.stack([jit[:, 3], jit[:, 4] + 0.02, jit[:, 5] + 0.005], axis=1)
    return np.concatenate([v0, v1, v2], axis=1).astype(np.float32)


def rays(n, seed=SEED_SOUP + 2, t_min=0.0, t_max=np.inf):
    rng = np.random.default_rng(seed)
    o = rng.random((n, 3))
    d = _unit(rng, n)
    r = np.zeros((n, 9), np.float32)
    r[:, 0:3] = o
    r[:, 3:6] = d
    r[:, 6] = t_min
    r[:, 7] = t_max
    return r


def axis_rays(n, seed=11):
    """Rays along the coordinate axes (direction components exactly 0), which
    take the parallel-slab branch of the box test."""
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 9), np.float32)
    r[:, 0:3] = rng.random((n, 3))
    axis = rng.integers(0, 3, n)
    sign = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    r[np.arange(n), 3 + axis] = sign
    r[:, 7] = np.inf
    return r


def jitter(tris, sigma=1e-3, seed=SEED_SOUP + 1):
    rng = np.random.default_rng(seed)
    return (tris.astype(np.float64) + rng.normal(0.0, sigma, tris.shape)).astype(np.float32)


def triangle_boxes(tris):
    """AABB of each triangle: exact componentwise min/max of its 3 vertices."""
    t = tris.reshape(-1, 3, 3)
    return np.concatenate([t.min(axis=1), t.max(axis=1)], axis=1).astype(np.float32)


def c2_inputs(n_tris=1 << 20, n_rays=1 << 22, n_queries=100_000, seed=SEED_SOUP):
    """The full LBVH microbenchmark input set (1M soup, jittered copy, 4M rays,
    100k query boxes)."""
    tris = soup(n_tris, seed)
    return {
        "tris": tris,
        "boxes": triangle_boxes(tris),
        "tris_jittered": jitter(tris, seed=seed + 1),
        "rays": rays(n_rays, seed=seed + 2),
        "query_boxes": triangle_boxes(soup(n_queries, seed + 3)),
    }


# ---- two-level scenes (meshes + affine instances) ------------------------------
def affine(rotation=None, scale=(1.0, 1.0, 1.0), translation=(0.0, 0.0, 0.0)):
    """Row-major 4x4 object_to_world = T * R * S (float32; row 3 = 0 0 0 1).
    `rotation` is a quaternion (w, x, y, z), normalised here."""
    m = np.eye(4, dtype=np.float64)
    if rotation is not None:
        w, x, y, z = np.asarray(rotation, np.float64) / np.linalg.norm(rotation)
        m[:3, :3] = [[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]]
    m[:3, :3] = m[:3, :3] * np.asarray(scale, np.float64)[None, :]
    m[:3, 3] = translation
    return m.astype(np.float32)


def soup_mesh(n, seed, edge=(0.02, 0.1)):
    """A soup of n triangles in [0,1]^3 as an indexed mesh (vertices, indices)."""
    t = soup(n, seed=seed, edge=edge)
    return t.reshape(-1, 3), np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)


def instances_random(n, n_meshes, seed, extent=3.0, scale=(0.4, 1.6)):
    """n instances over n_meshes meshes: random rotation, scale and translation."""
    rng = np.random.default_rng(seed)
    mesh = rng.integers(0, n_meshes, n).astype(np.int32)
    m = np.stack([affine(rng.standard_normal(4), rng.uniform(scale[0], scale[1], 3),
                         rng.uniform(-extent, extent, 3)) for _ in range(n)])
    return mesh, m


def scene_rays(n, seed, extent=4.0, aim=3.0):
    """Rays from random points of [-extent, extent]^3 towards random points of
    [-aim, aim]^3 (unit directions, t_min 0, t_max +inf)."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(-extent, extent, (n, 3))
    d = rng.uniform(-aim, aim, (n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 9), np.float32)
    r[:, 0:3] = o
    r[:, 3:6] = d
    r[:, 7] = np.inf
    return r


def two_level_inputs(n_meshes=4, tris_per_mesh=1 << 16, n_instances=64, n_rays=1 << 22, seed=SEED_SOUP + 5):
    """Instanced microbench: n_meshes triangle soups of tris_per_mesh triangles
    (edges U(0.002, 0.02) in the unit cube), n_instances random affine
    instances in [-3, 3]^3 (scales U(0.4, 1.6)), and n_rays rays aimed into the
    instanced region."""
    meshes = [soup_mesh(tris_per_mesh, seed + 10 * k, edge=(0.002, 0.02)) for k in range(n_meshes)]
    inst_mesh, inst_m = instances_random(n_instances, n_meshes, seed + 1)
    return {"meshes": meshes, "inst_mesh": inst_mesh, "inst_m": inst_m, "rays": scene_rays(n_rays, seed + 2)}
