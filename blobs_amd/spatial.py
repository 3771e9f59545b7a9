"""SpatialHash (blobs/src/spatial.rs) with its queries on the GPU.

`SpatialHash(cell_size)` keeps the reference's bookkeeping on the host --
points by id, and per cell (floor(x / cell_size), floor(y / cell_size)) as
i32 the ids filed there (spatial.rs:30-138), including the entry an
insert_with_id of an existing id leaves in the old cell -- and answers
`query` / `query_many` with bg_spatial_hash_query_host: the memberships and
queries go to device memory, one launch sequence hashes the occupied cells,
files the memberships into per-cell runs and evaluates every query
(spatial.rs:155-195).  A query's hits come in the reference's neighbour-cell
order and, inside a cell, in insertion order (the reference's HashSet order
is unspecified).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

POINT = np.dtype([("id", np.uint64), ("position", np.float32, 2), ("radius", np.float32), ("cell", np.int32, 2),
                  ("pad", np.uint32)])
QUERY = np.dtype([("position", np.float32, 2), ("radius", np.float32), ("pad", np.uint32)])
HIT = np.dtype([("id", np.uint64), ("position", np.float32, 2), ("radius", np.float32), ("point", np.uint32)])
assert POINT.itemsize == 32 and QUERY.itemsize == 16 and HIT.itemsize == 24


def _cast_i32(f: float) -> int:
    """Rust `f32 as i32`: saturating, NaN -> 0."""
    if f != f:
        return 0
    if f >= 2147483648.0:
        return 2147483647
    if f < -2147483648.0:
        return -2147483648
    return int(f)


def query_points(cell_size: float, points: np.ndarray, queries: np.ndarray, hit_cap: int | None = None):
    """The GPU query over explicit memberships (POINT records) for QUERY
    records, from host arrays (bg_spatial_hash_query_host): (offsets
    [n_queries + 1], hits [offsets[-1]] as HIT records)."""
    lib = L.load()
    points = np.ascontiguousarray(points, POINT)
    queries = np.ascontiguousarray(queries, QUERY)
    offsets = np.zeros(queries.shape[0] + 1, np.uint32)
    cap = 1 << 16 if hit_cap is None else int(hit_cap)
    total = C.c_uint64()
    for _ in range(2):
        hits = np.zeros(max(cap, 1), HIT)
        code = lib.bg_spatial_hash_query_host(float(cell_size), points.ctypes.data, points.shape[0],
                                              queries.ctypes.data, queries.shape[0], offsets.ctypes.data,
                                              hits.ctypes.data, cap, C.byref(total))
        if code == L.BG_ERR_CAPACITY and hit_cap is None:
            cap = int(total.value)  # offsets are complete: size the buffer once more
            continue
        L.check(code, None)
        break
    return offsets.astype(np.int64), hits[: int(total.value)]


class SpatialHash:
    """spatial.rs:30-196 (the HashMap grid)."""

    def __init__(self, cell_size: float):  # spatial.rs:43-53
        self.cell_size = np.float32(cell_size)
        self.next_id = 0
        self.points: dict[int, tuple] = {}
        self.grid: dict[tuple, dict] = {}

    def get_cell_coords(self, position) -> tuple:  # :55-60
        x = np.floor(np.float32(position[0]) / self.cell_size)
        y = np.floor(np.float32(position[1]) / self.cell_size)
        return (_cast_i32(float(x)), _cast_i32(float(y)))

    def insert(self, position, radius) -> int:  # :62-67
        i = self.next_id
        self.next_id += 1
        self.insert_with_id(i, position, radius)
        return i

    def insert_with_id(self, i: int, position, radius) -> None:  # :69-87
        p = (int(i), (np.float32(position[0]), np.float32(position[1])), np.float32(radius))
        self.grid.setdefault(self.get_cell_coords(p[1]), {})[p[0]] = None
        self.points[p[0]] = p

    def remove(self, i: int):  # :89-106
        p = self.points.pop(int(i), None)
        if p is not None:
            cell = self.grid.get(self.get_cell_coords(p[1]))
            if cell is not None:
                cell.pop(p[0], None)
        return p

    def move_point(self, i: int, offset):  # :108-138
        p = self.points.get(int(i))
        if p is None:
            return None
        new = (np.float32(p[1][0] + np.float32(offset[0])), np.float32(p[1][1] + np.float32(offset[1])))
        oc, nc = self.get_cell_coords(p[1]), self.get_cell_coords(new)
        if oc != nc:
            cell = self.grid.get(oc)
            if cell is not None:
                cell.pop(p[0], None)
            self.grid.setdefault(nc, {})[p[0]] = None
        self.points[p[0]] = (p[0], new, p[2])
        return ()

    def memberships(self) -> np.ndarray:
        recs = [(p[0], p[1], p[2], c) for c, ids in self.grid.items() for p in (self.points[i] for i in ids)]
        out = np.zeros(len(recs), POINT)
        for k, (i, pos, r, c) in enumerate(recs):
            out[k] = (i, pos, r, c, 0)
        return out

    def query_many(self, positions, radii):
        """SpatialHash::query for every (position, radius): list of HIT arrays."""
        positions = np.asarray(positions, np.float32).reshape(-1, 2)
        q = np.zeros(positions.shape[0], QUERY)
        q["position"] = positions
        q["radius"] = np.broadcast_to(np.asarray(radii, np.float32), positions.shape[:1])
        off, hits = query_points(self.cell_size, self.memberships(), q)
        return [hits[off[k]:off[k + 1]] for k in range(q.shape[0])]

    def query(self, position, query_radius):  # :155-195
        return self.query_many([position], [query_radius])[0]
