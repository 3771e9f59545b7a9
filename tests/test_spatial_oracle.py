"""Pins the SpatialHash restatement (oracle/spatial_oracle.py) before it
checks the GPU query.  The reference's own tests of spatial.rs
(blobs/src/tests.rs:14-90) are restated first; the rest is what the
mathematics fixes: with query_radius + max radius <= cell_size the 3x3
cells hold every point within range, so the query equals a brute-force
radius search; a point two cells away is never returned whatever the
radius; the cell cast saturates."""
import numpy as np

from oracle.spatial_oracle import SpatialHashOracle, sat_i32


def create_spatial_hash():  # tests.rs:14-23
    h = SpatialHashOracle(100.0)
    for k in range(5):
        h.insert((50.0 + 100 * k, 50.0 + 100 * k), 1.0)
    return h


def test_reference_get_cell_coords():  # tests.rs:25-33
    h = create_spatial_hash()
    for k in range(5):
        assert h.get_cell_coords((50.0 + 100 * k, 50.0 + 100 * k)) == (k, k)


def test_reference_insert():  # tests.rs:35-52
    h = SpatialHashOracle(100.0)
    i = h.insert((50.5, -25.5), 1.0)
    assert i == 0 and h.next_id == 1
    assert h.grid[h.get_cell_coords((50.5, -25.5))]
    p = h.points[i]
    assert p[0] == 0 and abs(p[1][0] - 50.5) < 1e-6 and abs(p[1][1] + 25.5) < 1e-6


def test_reference_insert_and_query():  # tests.rs:54-64
    h = SpatialHashOracle(1.0)
    p1 = h.insert((0.0, 0.0), 0.0)
    p2 = h.insert((2.0, 2.0), 0.0)
    res = h.query((1.0, 1.0), 1.5)
    assert len(res) == 2 and {r[0] for r in res} == {p1, p2}


def test_reference_move_point():  # tests.rs:66-78
    h = SpatialHashOracle(1.0)
    p = h.insert((0.0, 0.0), 0.0)
    assert h.move_point(p, (2.0, 2.0)) is not None
    res = h.query((1.0, 1.0), 1.5)
    assert len(res) == 1 and res[0][0] == p and res[0][1] == (2.0, 2.0)


def test_reference_remove():  # tests.rs:80-89
    h = SpatialHashOracle(1.0)
    p = h.insert((0.0, 0.0), 0.0)
    h.remove(p)
    assert h.query((0.0, 0.0), 1.5) == []


def test_query_equals_brute_force_within_one_cell():
    rng = np.random.default_rng(3)
    h = SpatialHashOracle(2.0)
    pts = rng.uniform(-20, 20, (600, 2)).astype(np.float32)
    rad = rng.uniform(0.0, 0.5, 600).astype(np.float32)
    for p, r in zip(pts, rad):
        h.insert(p, r)
    for q in rng.uniform(-22, 22, (200, 2)).astype(np.float32):
        qr = np.float32(1.5)  # qr + max radius <= cell size
        got = sorted(r[0] for r in h.query(q, qr))
        d2 = ((pts - q) ** 2).sum(1, dtype=np.float32)
        want = sorted(np.flatnonzero(d2 <= (qr + rad) ** 2).tolist())
        assert got == want


def test_points_two_cells_away_are_never_returned():
    h = SpatialHashOracle(1.0)
    far = h.insert((2.5, 0.5), 0.0)  # cell (2, 0); the query's cell is (0, 0)
    assert h.query((0.5, 0.5), 100.0) == []
    assert [r[0] for r in h.query((1.5, 0.5), 100.0)] == [far]


def test_cell_cast_saturates_and_maps_nan_to_zero():
    assert sat_i32(np.float32(3e9)) == 2147483647 and sat_i32(np.float32(-3e9)) == -2147483648
    assert sat_i32(np.float32(np.nan)) == 0
    h = SpatialHashOracle(1.0)
    assert h.get_cell_coords((np.nan, -0.5)) == (0, -1)


def test_reinserting_an_id_keeps_its_old_cell_entry():
    """insert_with_id on an existing id (spatial.rs:69-87) files it under the
    new cell and leaves the old cell's entry: a query near both returns it twice."""
    h = SpatialHashOracle(1.0)
    h.insert_with_id(7, (0.5, 0.5), 0.1)
    h.insert_with_id(7, (1.5, 0.5), 0.1)
    assert [r[0] for r in h.query((1.0, 0.5), 1.0)] == [7, 7]
