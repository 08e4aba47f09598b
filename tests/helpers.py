"""Shared test inputs: seeded ray sets aimed at a scene, including rays
through triangle vertices and edges (the watertightness cases)."""
import numpy as np


def scene_bounds(scene):
    lo, hi = np.full(3, np.inf), np.full(3, -np.inf)
    for inst in scene.instances:
        m = scene.meshes[inst["blas"]]
        T = np.asarray(inst["transform"], np.float64).reshape(3, 4)
        p = m.positions.astype(np.float64) @ T[:, :3].T + T[:, 3]
        lo = np.minimum(lo, p.min(0))
        hi = np.maximum(hi, p.max(0))
    return lo, hi


def world_vertices(scene, k, inst_index=None):
    inst = scene.instances[k if inst_index is None else inst_index]
    m = scene.meshes[inst["blas"]]
    T = np.asarray(inst["transform"], np.float64).reshape(3, 4)
    return m.positions.astype(np.float64) @ T[:, :3].T + T[:, 3], m.indices


def random_rays(scene, n, seed=0, edge_fraction=0.3):
    """n rays (n, 8): origin, tmin, direction, tmax.  A share of them aim at
    triangle vertices and edge points exactly."""
    rng = np.random.default_rng(seed)
    lo, hi = scene_bounds(scene)
    c, r = (lo + hi) / 2, np.linalg.norm(hi - lo) / 2 + 1.0
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    origins = c + u * r * rng.uniform(0.3, 1.5, size=(n, 1))
    targets = lo + (hi - lo) * rng.uniform(size=(n, 3))
    ne = int(n * edge_fraction)
    if ne:
        k = rng.integers(0, len(scene.instances), ne)
        for j in range(ne):
            P, I = world_vertices(scene, k[j])
            t = I[rng.integers(0, I.shape[0])]
            a, b = P[t[0]], P[t[1]]
            w = rng.choice([0.0, 0.5, rng.uniform()])
            targets[j] = a * (1 - w) + b * w
    d = targets - origins
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = origins
    rays[:, 3] = 0.0
    rays[:, 4:7] = d
    rays[:, 7] = 1e30
    # some bounded rays
    m = rng.uniform(size=n) < 0.2
    rays[m, 7] = rng.uniform(0.5, 2 * r, size=m.sum())
    return rays


def rel_mse(img, ref, eps=1e-2):
    """Relative MSE of an image against a reference (per channel, mean)."""
    img = np.asarray(img, np.float64)[..., :3]
    ref = np.asarray(ref, np.float64)[..., :3]
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + eps)))
