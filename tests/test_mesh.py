"""Pins of the oracle's restatement of the outline step of
world::updateChunkMesh (world.cpp:722-895): MarchingSquares::value and
FindPerimeter (physics_math.cpp:1866-1965), simplify (Ramer-Douglas-Peucker,
physics_math.cpp:1766-1843) and the loop scan with its edgeSeen marks.

The reference's sources need SDL headers (physics_math.hpp includes
engine/core/sdl_wrapper.h), so they are not built here; the expected values
below are worked out by hand from the marching-squares table and the
geometry, and the invariants hold for any mask."""
import numpy as np

from oracle import pyoracle as O

N, E, S, W = (0, 1), (1, 0), (0, -1), (-1, 0)


def mask(rows):
    return np.array([[c == "#" for c in r] for r in rows], np.uint8)


def test_vertex_value_table():
    # vertex (x, y) sums cells (x-1, y-1):1, (x, y-1):2, (x-1, y):4, (x, y):8
    m = mask(["#.", ".."])
    assert O.ms_value(m, 0, 0) == 8 and O.ms_value(m, 1, 0) == 4
    assert O.ms_value(m, 0, 1) == 2 and O.ms_value(m, 1, 1) == 1
    assert O.ms_value(m, 2, 2) == 0 and O.ms_value(m, -1, 0) == 0
    assert O.ms_value(mask(["##", "##"]), 1, 1) == 15


def test_find_perimeter_unit_cell_and_square():
    # a lone cell from its top-left vertex: down, right, up, left (y up in directions)
    assert O.find_perimeter(mask(["#.", ".."]), 0, 0) == [S, E, N, W]
    # runs are merged: a 2x2 block gives four runs of two
    sq = mask(["##..", "##..", "....", "...."])
    assert O.find_perimeter(sq, 0, 0) == [(0, -2), (2, 0), (0, 2), (-2, 0)]
    # off a perimeter (value 0 or 15): no runs
    assert O.find_perimeter(sq, 3, 3) == []
    full = mask(["###", "###", "###"])
    assert O.find_perimeter(full, 1, 1) == []


def test_find_perimeter_l_shape_and_closure():
    L = mask(["#..", "#..", "###"])
    runs = O.find_perimeter(L, 0, 0)
    assert runs == [(0, -3), (3, 0), (0, 1), (-2, 0), (0, 2), (-1, 0)]
    # every traced loop closes and alternates axis between runs
    rng = np.random.default_rng(1)
    for _ in range(200):
        m = (rng.random((9, 11)) < 0.5).astype(np.uint8)
        ys, xs = np.nonzero(m)
        for x, y in zip(xs[:5], ys[:5]):
            r = O.find_perimeter(m, int(x), int(y))
            if not r:
                continue
            assert sum(d[0] for d in r) == 0 and sum(d[1] for d in r) == 0
            assert all((a[0] == 0) != (b[0] == 0) for a, b in zip(r, r[1:]))


def test_find_perimeter_saddles_follow_the_previous_direction():
    # two cells touching at a corner: the shared vertex is a saddle (6 or 9);
    # arriving from the north the walk turns west, so each cell is its own loop
    d = mask(["#.", ".#"])
    a = O.find_perimeter(d, 0, 0)
    assert a == [S, E, N, W]
    b = O.find_perimeter(d, 1, 1)
    assert b == [S, E, N, W]
    d2 = mask([".#", "#."])
    assert O.find_perimeter(d2, 1, 0) == [S, E, N, W]


def test_simplify_keeps_corners_and_drops_collinear_points():
    # a 4x3 rectangle walked vertex by vertex: only its corners stay
    pts = [(0, 1), (0, 2), (0, 3), (1, 3), (2, 3), (3, 3), (4, 3), (4, 2), (4, 1), (4, 0), (3, 0), (2, 0), (1, 0), (0, 0)]
    assert O.simplify(pts) == [(0.0, 1.0), (0.0, 3.0), (4.0, 3.0), (4.0, 0.0), (0.0, 0.0)]
    # points within distance 1 of the chord go, farther ones stay; exactly 1 goes (max <= tolerance)
    assert O.simplify([(0, 0), (5, 1), (10, 0)]) == [(0.0, 0.0), (10.0, 0.0)]
    assert O.simplify([(0, 0), (5, 2), (10, 0)]) == [(0.0, 0.0), (5.0, 2.0), (10.0, 0.0)]
    # the first of equally far points splits (a strict > scan)
    assert O.simplify([(0, 0), (3, 3), (6, 3), (9, 0)]) == [(0.0, 0.0), (3.0, 3.0), (6.0, 3.0), (9.0, 0.0)]
    # two points or fewer are returned as they are
    assert O.simplify([(0, 0), (7, 7)]) == [(0.0, 0.0), (7.0, 7.0)]
    # the distance is to the segment, not the line: a point beyond an end
    assert O.simplify([(0, 0), (-3, 0), (1, 0)]) == [(0.0, 0.0), (-3.0, 0.0), (1.0, 0.0)]


def test_mesh_of_a_ring_has_an_outline_and_a_hole():
    ring = mask(["......", ".####.", ".#..#.", ".#..#.", ".####.", "......"])
    loops = O.mesh_mask(ring)
    assert len(loops) == 2
    (sx, sy, hole, v), (sx2, sy2, hole2, v2) = loops
    assert (sx, sy, hole) == (1, 1, False) and sorted(v) == [(1.0, 1.0), (1.0, 5.0), (5.0, 1.0), (5.0, 5.0)]
    assert (sx2, sy2, hole2) == (4, 2, True) and sorted(v2) == [(2.0, 2.0), (2.0, 4.0), (4.0, 2.0), (4.0, 4.0)]


def test_mesh_drops_tiny_loops_and_marks_seen_cells():
    # a lone cell simplifies to 2 points and is dropped ("1x1 that breaks everything")
    assert O.mesh_mask(mask(["....", ".#..", "....", "...."])) == []
    # a 2x2 block: cell (1,1) has its three right/lower neighbours set, so the
    # scan starts at (2,1); the walk is W1 S2 E2 N2 W1 with run ends (1,1),
    # (1,3), (3,3), (3,1), (2,1), and simplify drops (3,1), 0.894 from the
    # chord (3,3)-(2,1).  Reversed (TPPLPoly order) it is counter-clockwise.
    m22 = mask(["....", ".##.", ".##.", "...."])
    assert O.find_perimeter(m22, 2, 1) == [W, (0, -2), (2, 0), (0, 2), W]
    (sx, sy, hole, v), = O.mesh_mask(m22)
    assert (sx, sy, hole) == (2, 1, False) and v == [(1.0, 1.0), (1.0, 3.0), (3.0, 3.0), (2.0, 1.0)]
    # a filled 16x16 chunk: the top row's cells have all three neighbours set
    # up to (15,0), where the scan starts; the walk W15 S16 E16 N16 W1 ends
    # its runs at (0,0), (0,16), (16,16), (16,0), (15,0), and simplify drops
    # the corner (16,0): 0.998 from the chord (16,16)-(15,0).  The walk marks
    # the other border cells seen, so it is the only loop.
    full = np.ones((16, 16), np.uint8)
    assert O.find_perimeter(full, 15, 0) == [(-15, 0), (0, -16), (16, 0), (0, 16), W]
    loops = O.mesh_mask(full)
    assert [(a, b, h) for a, b, h, _ in loops] == [(15, 0, False)]
    assert loops[0][3] == [(0.0, 0.0), (0.0, 16.0), (16.0, 16.0), (15.0, 0.0)]


def test_mesh_invariants_on_random_masks():
    """Every kept loop has 3+ integer vertices from its own perimeter, and a
    loop is a hole exactly when its vertex list runs clockwise in the
    reference's reversed order (negative shoelace sum)."""
    rng = np.random.default_rng(7)
    for _ in range(40):
        m = (rng.random((40, 33)) < rng.uniform(0.3, 0.7)).astype(np.uint8)
        for sx, sy, hole, v in O.mesh_mask(m):
            assert len(v) >= 3 and all(a == int(a) and b == int(b) for a, b in v)
            assert m[sy, sx] == 1
            p = v[::-1]
            area = sum(p[i][0] * p[(i + 1) % len(p)][1] - p[i][1] * p[(i + 1) % len(p)][0] for i in range(len(p)))
            assert hole == (area < 0)
            # vertices lie on the perimeter's vertex set
            full = O.find_perimeter(m, sx, sy)
            x, y, pts = sx, sy, set()
            for dx, dy in full:
                x, y = x + dx, y - dy
                pts.add((float(x), float(y)))
            assert set(v) <= pts
