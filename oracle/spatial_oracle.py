"""Restatement of blobs' SpatialHash (blobs/src/spatial.rs) in Python.

TEST INFRASTRUCTURE ONLY -- the checker for the GPU SpatialHash query.  Only
tests/ import it; nothing in blobs_amd/ does.

f32 arithmetic as the Rust source: get_cell_coords is
(x / cell_size).floor() as i32 (a saturating cast, NaN -> 0); query compares
(p - q).length_squared() = dx*dx + dy*dy with (query_radius + p.radius)^2.
The grid's HashSet iteration order is unspecified in the reference; here a
cell keeps its ids in insertion order.
"""
from __future__ import annotations

import numpy as np

F = np.float32


def sat_i32(f) -> int:
    """Rust `f32 as i32`: saturating, NaN -> 0."""
    f = float(f)
    if f != f:
        return 0
    if f >= 2147483648.0:
        return 2147483647
    if f < -2147483648.0:
        return -2147483648
    return int(f)


def wrap_i32(v: int) -> int:
    return ((v + 2 ** 31) % 2 ** 32) - 2 ** 31


class SpatialHashOracle:
    """spatial.rs:30-196 (the HashMap grid, not the `use-grid` feature)."""

    def __init__(self, cell_size: float):  # spatial.rs:43-53
        self.cell_size = F(cell_size)
        self.next_id = 0
        self.points: dict[int, tuple] = {}       # id -> (id, (x, y), radius)
        self.grid: dict[tuple, dict] = {}        # cell -> ids (dict as an insertion-ordered set)

    def get_cell_coords(self, position) -> tuple:  # :55-60
        x, y = F(position[0]), F(position[1])
        return (sat_i32(np.floor(x / self.cell_size)), sat_i32(np.floor(y / self.cell_size)))

    def insert(self, position, radius) -> int:  # :62-67
        i = self.next_id
        self.next_id += 1
        self.insert_with_id(i, position, radius)
        return i

    def insert_with_id(self, i, position, radius):  # :69-87
        p = (int(i), (F(position[0]), F(position[1])), F(radius))
        self.grid.setdefault(self.get_cell_coords(p[1]), {})[p[0]] = None
        self.points[p[0]] = p

    def remove(self, i):  # :89-106
        p = self.points.pop(int(i), None)
        if p is not None:
            cell = self.grid.get(self.get_cell_coords(p[1]))
            if cell is not None:
                cell.pop(p[0], None)
        return p

    def move_point(self, i, offset):  # :108-138
        p = self.points.get(int(i))
        if p is None:
            return None
        old = p[1]
        new = (F(old[0] + F(offset[0])), F(old[1] + F(offset[1])))
        oc, nc = self.get_cell_coords(old), self.get_cell_coords(new)
        if oc != nc:
            cell = self.grid.get(oc)
            if cell is not None:
                cell.pop(p[0], None)
            self.grid.setdefault(nc, {})[p[0]] = None
        self.points[p[0]] = (p[0], new, p[2])
        return ()

    @staticmethod
    def neighbor_cells(c):  # :140-153
        x, y = c
        return [(wrap_i32(x + dx), wrap_i32(y + dy)) for dx in (-1, 0, 1) for dy in (-1, 0, 1)]

    def query(self, position, query_radius) -> list:  # :155-195
        q = (F(position[0]), F(position[1]))
        r = F(query_radius)
        out = []
        for c in self.neighbor_cells(self.get_cell_coords(q)):
            for i in self.grid.get(c, {}):
                p = self.points[i]
                dist = r + p[2]
                dx, dy = p[1][0] - q[0], p[1][1] - q[1]
                if dx * dx + dy * dy <= dist * dist:
                    out.append(p)
        return out

    def memberships(self) -> list:
        """(id, position, radius, cell) of every id in every cell, cell by cell
        in the grid's insertion order, ids in insertion order."""
        out = []
        for c, ids in self.grid.items():
            for i in ids:
                p = self.points[i]
                out.append((p[0], p[1], p[2], c))
        return out
