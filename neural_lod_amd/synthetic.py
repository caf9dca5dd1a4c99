"""Seeded synthetic workloads (meshes, queries, labels) for tests and benchmarks.

All generators are deterministic in their ``seed`` and return float32 numpy
arrays laid out as the C ABI expects (AoS rows).
"""
from __future__ import annotations

import numpy as np


def blob_mesh(n_triangles: int = 200_000, n_blobs: int = 64, sigma: float = 0.05,
              half_extent: float = 0.01, seed: int = 0) -> np.ndarray:
    """Triangle soup float32[n][9] of small triangles clustered in Gaussian blobs.

    Blob centres ~ U[0.1, 0.9]^3; each triangle centre ~ N(blob, sigma^2 I)
    clipped to [0, 1]^3 with its blob drawn uniformly; vertices are
    centre + U[-half_extent, half_extent]^3.
    """
    rng = np.random.default_rng(seed)
    blobs = rng.uniform(0.1, 0.9, size=(n_blobs, 3))
    which = rng.integers(0, n_blobs, size=n_triangles)
    centres = np.clip(blobs[which] + rng.normal(0.0, sigma, size=(n_triangles, 3)), 0.0, 1.0)
    offs = rng.uniform(-half_extent, half_extent, size=(n_triangles, 3, 3))
    tris = (centres[:, None, :] + offs).astype(np.float32)
    return np.ascontiguousarray(tris.reshape(n_triangles, 9))


def box_mesh(lo=(0.2, 0.25, 0.3), hi=(0.7, 0.8, 0.75)) -> np.ndarray:
    """Closed axis-aligned box as 12 triangles, outward-facing."""
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    c = np.array([[lo[0] if (i & 1) == 0 else hi[0],
                   lo[1] if (i & 2) == 0 else hi[1],
                   lo[2] if (i & 4) == 0 else hi[2]] for i in range(8)], np.float32)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    tris = []
    for a, b, cc, d in quads:
        tris.append(np.concatenate([c[a], c[b], c[cc]]))
        tris.append(np.concatenate([c[a], c[cc], c[d]]))
    return np.ascontiguousarray(np.stack(tris).astype(np.float32))


def sphere_mesh(n_theta: int = 24, n_phi: int = 48, radius: float = 0.4,
                centre=(0.5, 0.5, 0.5)) -> np.ndarray:
    """Closed UV sphere (watertight triangle soup)."""
    ctr = np.asarray(centre, np.float64)

    def p(i, j):
        th = np.pi * i / n_theta
        ph = 2 * np.pi * (j % n_phi) / n_phi
        return ctr + radius * np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])

    tris = []
    for i in range(n_theta):
        for j in range(n_phi):
            a, b, c, d = p(i, j), p(i + 1, j), p(i + 1, j + 1), p(i, j + 1)
            if i != 0:
                tris.append(np.concatenate([a, b, d]))
            if i != n_theta - 1:
                tris.append(np.concatenate([b, c, d]))
    return np.ascontiguousarray(np.stack(tris).astype(np.float32))


def visibility_queries(n: int, seed: int = 1) -> np.ndarray:
    """Inputs float32[n][6]: voxel-local position U[0,1]^3 then a direction
    uniform on the sphere mapped to [0,1]^3 as (d + 1) / 2."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, 1.0, size=(n, 3))
    z = rng.uniform(-1.0, 1.0, size=n)
    phi = rng.uniform(0.0, 2 * np.pi, size=n)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    return np.ascontiguousarray(np.concatenate([pos, (d + 1.0) / 2.0], axis=1).astype(np.float32))


def throughput_queries(n: int, n_lods: int = 5, seed: int = 2) -> np.ndarray:
    """Inputs float32[n][10]: position U[0,1]^3, outgoing and incoming
    directions mapped to [0,1]^3, and a LoD level in {0..n_lods-1} as a scalar."""
    rng = np.random.default_rng(seed)

    def dirs():
        z = rng.uniform(-1.0, 1.0, size=n)
        phi = rng.uniform(0.0, 2 * np.pi, size=n)
        r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
        return (np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1) + 1.0) / 2.0

    pos = rng.uniform(0.0, 1.0, size=(n, 3))
    lod = rng.integers(0, n_lods, size=(n, 1)).astype(np.float64)
    return np.ascontiguousarray(np.concatenate([pos, dirs(), dirs(), lod], axis=1).astype(np.float32))


def sphere_visibility_labels(queries: np.ndarray, centre=(0.5, 0.5, 0.5), radius: float = 0.25,
                             max_dist: float = 0.35) -> np.ndarray:
    """Analytic visibility targets for visibility queries [pos(3), (dir+1)/2 (3)]:
    1.0 where the ray from pos along dir hits the sphere within max_dist
    (a stand-in for "hit before leaving the voxel"), else 0.0.  float32 [n][1]."""
    q = np.asarray(queries, np.float64)
    o = q[:, :3] - np.asarray(centre, np.float64)
    d = q[:, 3:6] * 2.0 - 1.0
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    b = (o * d).sum(axis=1)
    c = (o * o).sum(axis=1) - radius * radius
    disc = b * b - c
    sq = np.sqrt(np.maximum(disc, 0.0))
    t0, t1 = -b - sq, -b + sq
    t = np.where(t0 > 0.0, t0, t1)  # nearest hit in front of the origin (inside: the exit)
    hit = (disc >= 0.0) & (t > 0.0) & (t <= max_dist)
    return hit.astype(np.float32)[:, None]
