"""Synthetic grids and environment fields (host-side input generation).

* ``rect_grid``: W x H lattice with the neighbour slot order of
  GridFactory::createCellsRectPeriodic (QHG4/tools_io/GridFactory.cpp:1024-1068:
  slot 0 = +x, 1 = +y, 2 = -x, 3 = -y).  Non-periodic borders keep the slot
  and store -1 while the neighbour count is decremented, as
  GridFactory::createCellsRect does (:966-1018).
* ``ico_grid``: vertex grid of a recursively subdivided icosahedron,
  10*4^n + 2 cells (12 pentagon vertices with 5 neighbours, the rest 6).
  Neighbours are listed counter-clockwise (seen from outside) starting at
  the one closest to local north; slot 5 of pentagons is -1.
* ``value_noise``: seeded multi-octave 3-D value noise evaluated at unit
  vectors, in [0, 1).
"""
import numpy as np

MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def rect_grid(width, height, periodic=False):
    n = width * height
    i = np.arange(n)
    x, y = i % width, i // width
    nbr = np.stack([
        y * width + (x + 1) % width,
        ((y + 1) % height) * width + x,
        y * width + (x - 1 + width) % width,
        ((y - 1 + height) % height) * width + x,
    ], axis=1).astype(np.int32)
    nn = np.full(n, 4, dtype=np.uint8)
    if not periodic:
        right, left = x == width - 1, x == 0
        top, bottom = y == height - 1, y == 0
        for mask, slot in ((right, 0), (left, 2), (top, 1), (bottom, 3)):
            nbr[mask, slot] = -1
            nn[mask] -= 1
    return {"num_cells": n, "connectivity": 4, "neighbors": np.ascontiguousarray(nbr),
            "num_neighbors": nn, "cell_ids": i.astype(np.int64)}


def _icosahedron():
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def ico_vertices_faces(level):
    v, f = _icosahedron()
    for _ in range(level):
        nv = len(v)
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * nv + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk // nv, uk % nv
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        nf = len(f)
        m01, m12, m20 = inv[:nf] + nv, inv[nf:2 * nf] + nv, inv[2 * nf:] + nv
        v = np.concatenate([v, mid])
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v, f


def ico_grid(level):
    v, f = ico_vertices_faces(level)
    n = len(v)
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    e.sort(axis=1)
    e = np.unique(e[:, 0] * n + e[:, 1])
    a, b = e // n, e % n
    src = np.concatenate([a, b])
    dst = np.concatenate([b, a])
    # local frame: north = z projected on the tangent plane (x near the poles)
    p = v[src]
    ref = np.where((np.abs(p[:, 2]) > 0.999)[:, None], np.array([1.0, 0, 0]), np.array([0, 0, 1.0]))
    north = ref - (ref * p).sum(1, keepdims=True) * p
    north /= np.linalg.norm(north, axis=1, keepdims=True)
    east = np.cross(north, p)
    d = v[dst] - p
    ang = np.arctan2((d * east).sum(1), (d * north).sum(1))
    ang = np.where(ang < 0, ang + 2 * np.pi, ang)
    order = np.lexsort((ang, src))
    src, dst = src[order], dst[order]
    cnt = np.bincount(src, minlength=n)
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    slot = np.arange(len(src)) - start[src]
    nbr = np.full((n, 6), -1, dtype=np.int32)
    nbr[src, slot] = dst
    lat = np.degrees(np.arcsin(np.clip(v[:, 2], -1, 1)))
    lon = np.degrees(np.arctan2(v[:, 1], v[:, 0]))
    return {"num_cells": n, "connectivity": 6, "neighbors": np.ascontiguousarray(nbr),
            "num_neighbors": cnt.astype(np.uint8), "cell_ids": np.arange(n, dtype=np.int64),
            "xyz": v, "lon": lon, "lat": lat}


def _mix64(x):
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def _lattice_value(ix, iy, iz, salt):
    with np.errstate(over="ignore"):
        h = (ix.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
             ^ iy.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)
             ^ iz.astype(np.uint64) * np.uint64(0x165667B19E3779F9)
             ^ np.uint64(salt))
        h = _mix64(h)
    return (h >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def value_noise(xyz, seed, octaves=6, base_freq=2.0):
    """Seeded multi-octave value noise at points xyz (n x 3), result in [0,1)."""
    total = np.zeros(len(xyz))
    norm = 0.0
    for o in range(octaves):
        freq = base_freq * (2 ** o)
        amp = 0.5 ** o
        p = xyz * freq + 1000.0
        i0 = np.floor(p).astype(np.int64)
        fr = p - i0
        s = fr * fr * (3 - 2 * fr)
        salt = (int(seed) * 0x100000001B3 + o * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF
        acc = np.zeros(len(xyz))
        for dx in (0, 1):
            wx = s[:, 0] if dx else 1 - s[:, 0]
            for dy in (0, 1):
                wy = s[:, 1] if dy else 1 - s[:, 1]
                for dz in (0, 1):
                    wz = s[:, 2] if dz else 1 - s[:, 2]
                    acc += wx * wy * wz * _lattice_value(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, salt)
        total += amp * acc
        norm += amp
    return total / norm


def bfs_patch(grid, start, ncells, allowed):
    """First ``ncells`` cells reached by BFS from ``start`` through allowed cells."""
    nbr = grid["neighbors"]
    seen = np.zeros(grid["num_cells"], dtype=bool)
    out = [start]
    seen[start] = True
    head = 0
    while head < len(out) and len(out) < ncells:
        c = out[head]
        head += 1
        for nb in nbr[c]:
            if nb >= 0 and not seen[nb] and allowed[nb]:
                seen[nb] = True
                out.append(int(nb))
                if len(out) >= ncells:
                    break
    return np.array(out, dtype=np.int32)


def bfs_levels_patch(grid, start, ncells, allowed):
    """``ncells`` allowed cells nearest ``start`` by hop count: whole BFS
    levels, the last one cut in cell-index order (vectorised; for large
    patches)."""
    nbr = grid["neighbors"]
    seen = np.zeros(grid["num_cells"], dtype=bool)
    seen[start] = True
    out = [np.array([start], dtype=np.int32)]
    total = 1
    frontier = out[0]
    while total < ncells and frontier.size:
        cand = nbr[frontier].ravel()
        cand = np.unique(cand[cand >= 0])
        cand = cand[~seen[cand] & allowed[cand]]
        seen[cand] = True
        take = cand[:ncells - total].astype(np.int32)
        out.append(take)
        total += take.size
        frontier = cand
    return np.concatenate(out)
