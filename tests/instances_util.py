"""Two-level test scenes: BLAS triangle sets, instance transforms and rays.

"Exact" transforms (a power-of-two uniform scale, a cyclic permutation of the
axes, sign flips and a translation on a 2^-6 grid) map the dyadic triangles
and rays used here without rounding, and the watertight test commutes with
them exactly: a ray traced through such an instance gives the same bits as
the world ray against the transformed triangles.  General transforms are
checked against an FP64 intersection instead.
"""
from __future__ import annotations

import numpy as np

from appearance_amd.raytrace import INSTANCE_DTYPE

GRID = 1.0 / 64.0


def dyadic_tris(n: int, seed: int, spread: float = 4.0, size: float = 0.5) -> np.ndarray:
    rng = np.random.default_rng(seed)
    c = np.round(rng.uniform(-spread, spread, (n, 3)) / GRID) * GRID
    e1 = np.round(rng.normal(size=(n, 3)) * size / GRID) * GRID
    e2 = np.round(rng.normal(size=(n, 3)) * size / GRID) * GRID
    return np.concatenate([c, c + e1, c + e2], axis=1).astype(np.float32)


def dyadic_rays(n: int, seed: int, spread: float = 14.0, aim: float = 6.0, center=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Rays from grid origins in [-spread, spread]^3 towards grid points in
    center + [-aim, aim]^3 (the scene), with grid directions whose three
    |components| differ (no ties in the dominant-axis choice) and are nonzero."""
    rng = np.random.default_rng(seed)
    c = np.round(np.asarray(center, np.float64) / GRID) * GRID
    o = np.round(rng.uniform(-spread, spread, (n, 3)) / GRID) * GRID
    d = np.zeros((n, 3))
    for i in range(n):
        while True:
            v = c + np.round(rng.uniform(-aim, aim, 3) / GRID) * GRID - o[i]
            a = np.abs(v)
            if a.min() > 0 and len(set(a.tolist())) == 3:
                break
            o[i] = np.round(rng.uniform(-spread, spread, 3) / GRID) * GRID
        d[i] = v
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 4:7] = d
    r[:, 7] = 1e30
    return r


CYCLIC = [np.eye(3), np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]], float),
          np.array([[0, 1, 0], [0, 0, 1], [1, 0, 0]], float)]


def exact_transform(rng) -> np.ndarray:
    """(3, 4) object-to-world: s * D * P | t with s in {1/2, 1, 2}, P cyclic, D = diag(+-1)."""
    s = float(rng.choice([0.5, 1.0, 2.0]))
    P = CYCLIC[int(rng.integers(3))]
    D = np.diag(rng.choice([-1.0, 1.0], 3))
    t = np.round(rng.uniform(-6, 6, 3) / GRID) * GRID
    m = np.zeros((3, 4))
    m[:, :3] = s * D @ P
    m[:, 3] = t
    return m


def general_transform(rng, scale=(0.3, 3.0), shift=6.0) -> np.ndarray:
    """(3, 4) object-to-world: a random rotation times a non-uniform scale, plus a translation."""
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q *= np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    m = np.zeros((3, 4))
    m[:, :3] = q @ np.diag(rng.uniform(*scale, 3))
    m[:, 3] = rng.uniform(-shift, shift, 3)
    return m


def place(tris: np.ndarray, m: np.ndarray, dtype=np.float64) -> np.ndarray:
    """Triangles (n, 9) through a (3, 4) transform, evaluated in `dtype`."""
    v = tris.reshape(-1, 3).astype(dtype)
    A = m[:, :3].astype(np.float32).astype(dtype)
    t = m[:, 3].astype(np.float32).astype(dtype)
    return (v @ A.T + t).reshape(-1, 9)


def make_instances(transforms, blas_ids, blas_sizes) -> np.ndarray:
    """apt_instance records; custom_index = running prim base (a flattened soup's order)."""
    out = np.zeros(len(transforms), INSTANCE_DTYPE)
    base = 0
    for k, (m, b) in enumerate(zip(transforms, blas_ids)):
        out[k]["transform"] = np.asarray(m, np.float32)
        out[k]["blas"] = b
        out[k]["custom_index"] = base
        base += int(blas_sizes[b])
    return out


def world_soup(blas_tris, inst: np.ndarray, dtype=np.float32) -> np.ndarray:
    """The instances' triangles placed in world space, in custom-index order."""
    parts = [place(blas_tris[int(i["blas"])], i["transform"].astype(np.float64), np.float64)
             for i in inst]
    return np.concatenate(parts, 0).astype(dtype)


def fp64_closest(tris64: np.ndarray, rays: np.ndarray, chunk: int = 128, eps: float = 1e-5):
    """Closest hit of each ray in float64 (Moller-Trumbore): t, prim, and whether
    the ray is ambiguous at float32 precision -- it passes within eps (in
    barycentrics) of a triangle's edge, or its two nearest hits lie within eps
    (relative) of each other."""
    tris64 = np.asarray(tris64, np.float64).reshape(-1, 9)
    v0, v1, v2 = tris64[:, 0:3], tris64[:, 3:6], tris64[:, 6:9]
    e1, e2 = v1 - v0, v2 - v0
    n = rays.shape[0]
    t_out = np.full(n, np.inf)
    p_out = np.full(n, 0xFFFFFFFF, np.uint64)
    amb = np.zeros(n, bool)
    for s in range(0, n, chunk):
        r = rays[s:s + chunk].astype(np.float64)
        o, d = r[:, None, 0:3], r[:, None, 4:7]
        pv = np.cross(d, e2[None])
        det = np.einsum("rnk,nk->rn", pv, e1)
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = 1.0 / det
            tv = o - v0[None]
            u = np.einsum("rnk,rnk->rn", tv, pv) * inv
            qv = np.cross(tv, e1[None])
            v = np.einsum("rnk,rnk->rn", np.broadcast_to(d, qv.shape), qv) * inv
            t = np.einsum("rnk,nk->rn", qv, e2) * inv
        w = 1.0 - u - v
        in_t = (t > r[:, 3:4]) & (t < r[:, 7:8])
        ok = (det != 0) & (u >= 0) & (v >= 0) & (w >= 0) & in_t
        near = (det != 0) & (u >= -eps) & (v >= -eps) & (w >= -eps) & in_t
        edge = near & (np.minimum(np.minimum(u, v), w) < eps)
        tt = np.where(ok, t, np.inf)
        order = np.argsort(tt, axis=1)[:, :2]
        rows = np.arange(tt.shape[0])
        t1, t2 = tt[rows, order[:, 0]], tt[rows, order[:, 1]]
        hit = np.isfinite(t1)
        t_out[s:s + chunk] = t1
        p_out[s:s + chunk] = np.where(hit, order[:, 0], 0xFFFFFFFF)
        with np.errstate(invalid="ignore"):
            close = hit & ((t2 - t1) <= eps * np.maximum(t1, 1e-30))
        amb[s:s + chunk] = edge.any(axis=1) | close
    return t_out, p_out.astype(np.uint32), amb
