"""Pins of the oracle's geometry restatement (STL parse, bbox/grid rule, ray parity, labels).

The reference ships no tests or fixtures, so each check is a closed form, a hand-built known
answer or an independent computation that a plausible mistake (a dropped nudge, an exclusive bound,
a vertex-3 bbox, an nz + 1) would fail.
"""
import os

import numpy as np
import pytest

from tests import oracle_ctypes as O


def tri(v1, v2, v3, n=None):
    v1, v2, v3 = map(np.asarray, (v1, v2, v3))
    if n is None:
        n = np.cross(v2 - v1, v3 - v1)
        n = n / np.linalg.norm(n)
    return np.concatenate([v1, v2, v3, n]).astype(np.float64)


def write_stl(path, tris, extra_lines=()):
    with open(path, "w") as fp:
        fp.write("solid t\n")
        for t in tris:
            fp.write("facet normal %r %r %r\n outer loop\n" % tuple(float(x) for x in t[9:12]))
            for v in range(3):
                fp.write("  vertex %r %r %r\n" % tuple(float(x) for x in t[3 * v:3 * v + 3]))
            fp.write(" endloop\nendfacet\n")
        for line in extra_lines:
            fp.write(line + "\n")
        fp.write("endsolid t\n")


def test_stl_parse_roundtrip_and_malformed(tmp_path):
    rng = np.random.default_rng(1)
    tris = [tri(*rng.normal(size=(3, 3))) for _ in range(5)]
    p = tmp_path / "a.stl"
    bad = ["facet normal 0 0 1", "outer loop", "vertex 1 2", "vertex 1 2 3", "vertex 1 2 3",
           "endloop", "endfacet"]
    write_stl(p, tris, bad)
    got = O.read_stl(p)
    # repr() of a float round-trips; the facet with a 2-value vertex line is dropped
    assert got.shape == (5, 12)
    assert np.array_equal(got, np.array(tris))


def test_bbox_excludes_vertex3_in_y_and_z():
    # y and z extremes only at vertex 3: the reference bbox ignores them (ComplexGeometry.cpp:147-151)
    t = np.array([tri([0, 0, 0], [1, 0, 0], [2, 5, -7])])
    bb = O.bbox(t)
    assert bb.tolist() == [0.0, 2.0, 0.0, 0.0, 0.0, 0.0]


def test_grid_dims_rule():
    # nx = ceil(dx/h)+1, ny = ceil(dy/h)+1, nz = ceil(dz/h) (no +1)
    assert O.grid_dims([0, 10, 0, 4.5, 0, 3.2], 1.0) == (11, 6, 4)
    assert O.grid_dims([0, 80, -9.97859, 9.97859, -9.97859, 9.97859], 1.0) == (81, 21, 20)


def square_z(z, lo=0.0, hi=10.0):
    # two triangles over [lo,hi]^2 at height z, split along the diagonal x = y
    a, b, c, d = [lo, lo, z], [hi, lo, z], [hi, hi, z], [lo, hi, z]
    return np.array([tri(a, b, c), tri(a, c, d)])


def test_ray_single_hit_and_miss():
    t = square_z(5.0)[:1]  # triangle (0,0),(10,0),(10,10): inside where y <= x
    assert O.count_crossings(t, [6, 2, 0], [6, 2, 20]) == 1
    assert O.count_crossings(t, [2, 6, 0], [2, 6, 20]) == 0   # y > x: outside
    assert O.count_crossings(t, [6, 2, 6], [6, 2, 20]) == 0   # plane below the start: t < 0
    assert O.count_crossings(t, [6, 2, 0], [6, 2, 4]) == 0    # plane beyond the end: t > 1


def test_ray_inclusive_bounds_and_shared_edge():
    s = square_z(5.0)
    # a ray through the shared diagonal is counted by both triangles (inclusive s, t bounds)
    assert O.count_crossings(s, [3, 3, 0], [3, 3, 20]) == 2
    # interior point of one triangle only
    assert O.count_crossings(s, [7, 2, 0], [7, 2, 20]) == 1
    # t = 0 and t = 1 are inclusive
    assert O.count_crossings(s[:1], [6, 2, 5], [6, 2, 20]) == 1
    assert O.count_crossings(s[:1], [6, 2, 0], [6, 2, 5]) == 1


def test_vertical_facet_never_counted():
    t = np.array([tri([5, 0, 0], [5, 10, 0], [5, 0, 10])])  # plane x = 5, file normal (1,0,0)
    assert O.count_crossings(t, [5, 1, -1], [5, 1, 20]) == 0


def test_file_normal_sets_the_ray_parameter():
    # the crossing parameter uses the FILE normal: a file normal inconsistent with the facet moves
    # the t test but the inside test re-projects on the true plane
    t = square_z(5.0)[:1].copy()
    t[0, 9:12] = [0.0, 0.0, 1.0]
    assert O.count_crossings(t, [6, 2, 0], [6, 2, 20]) == 1
    t[0, 9:12] = [1.0, 0.0, 1e-17]  # |d . n| < 1e-16: skipped
    assert O.count_crossings(t, [6, 2, 0], [6, 2, 20]) == 0


def polygon_inside(y, z, R, nseg):
    # point-in-convex-polygon for the tube's ring (vertex angles (k + 1/2) * 2 pi / nseg)
    th = (np.arange(nseg) + 0.5) * 2 * np.pi / nseg
    vy, vz = R * np.cos(th), R * np.sin(th)
    inside = np.ones(np.broadcast(y, z).shape, dtype=bool)
    margin = np.full(inside.shape, np.inf)
    for k in range(nseg):
        k1 = (k + 1) % nseg
        ey, ez = vy[k1] - vy[k], vz[k1] - vz[k]
        cr = ey * (z - vz[k]) - ez * (y - vy[k])  # > 0 on the inner side (counter-clockwise ring)
        inside &= cr > 0
        margin = np.minimum(margin, np.abs(cr) / np.hypot(ey, ez))
    return inside, margin


@pytest.fixture(scope="module")
def tube_tri():
    from pv_dec_original_amd.workloads import tube_mesh
    return tube_mesh(10.0, 80.0, 48)


def test_tube_labels_match_analytic_prism(tube_tri):
    bb = O.bbox(tube_tri)
    dims = O.grid_dims(bb, 1.0)
    assert dims == (81, 21, 20)
    origin = [bb[0], bb[2], bb[4]]
    lab = O.label_points(tube_tri, bb, dims, origin, 1.0)
    k, j, i = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    y = origin[1] + j * 1.0
    z = origin[2] + k * 1.0 + 0.5  # ray start z + h/2
    y = np.where(y == bb[2], y + 0.02, y)  # nudge of the min-y column
    inside, margin = polygon_inside(y, z, 10.0, 48)
    ok = margin > 1e-9
    assert ok.sum() > 0.99 * ok.size
    assert np.array_equal(lab[ok].astype(bool), inside[ok])
    # x = 0 and x = 80 columns are nudged inside the side facets' x range: fluid at both ends
    assert lab[:, :, 0].sum() == lab[:, :, 40].sum() == lab[:, :, 80].sum() > 250


def test_binned_labels_equal_brute_force(tube_tri):
    from pv_dec_original_amd.workloads import ybif_mesh
    cases = [(tube_tri, None)]
    small = ybif_mesh(rp=6.0, rd=4.25, angle=30.0, xb=40.0, y0=16.0, z0=16.0, xcap0=-2.5,
                      xcap1=65.5, box=(64, 32, 32))
    cases.append((small, ((0.0, 0.0, 0.0), (64, 32, 32))))
    rng = np.random.default_rng(7)
    rand = np.array([tri(*rng.uniform(0, 12, size=(3, 3))) for _ in range(60)])
    cases.append((rand, None))
    for t, grid in cases:
        bb = O.bbox(t)
        if grid is None:
            dims = O.grid_dims(bb, 1.0)
            origin = [bb[0], bb[2], bb[4]]
        else:
            origin, dims = grid
            for a in range(3):
                bb[2 * a] = min(bb[2 * a], origin[a])
                bb[2 * a + 1] = max(bb[2 * a + 1], origin[a] + dims[a] - 1)
        a = O.label_points(t, bb, dims, origin, 1.0)
        b = O.label_points(t, bb, dims, origin, 1.0, binned=True)
        assert np.array_equal(a, b)


def test_binary_stl_round_trip_and_reference_normalisation(tmp_path):
    """Binary STL (GeometryReader.cpp:119-176): vertices are the file's float32 values, normals
    are normalised in float32 arithmetic (checked against numpy float32 ops), the library reader
    matches the oracle bit for bit, and a truncated file is an error."""
    import struct
    from pv_dec_original_amd import geometry as G
    from pv_dec_original_amd.workloads import vessel_mesh
    tri = vessel_mesh(11)[:500].copy()
    tri[0, 9:12] = (3.0, 4.0, 0.0)      # not unit length
    tri[1, 9:12] = (0.0, 0.0, 0.0)      # zero normal -> NaN, as the reference computes it
    path = tmp_path / "v.stl"
    G.write_stl_binary(path, tri)
    assert G.is_binary_stl(path) and path.stat().st_size == 84 + 50 * len(tri)
    got = G.read_stl_binary(path)
    ref = O.read_stl_binary(path)
    assert got.shape == (500, 12) and np.array_equal(got, ref, equal_nan=True)
    assert np.array_equal(got[:, :9], tri[:, :9].astype(np.float32).astype(np.float64))
    n = tri[:, 9:12].astype(np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        norm = np.sqrt(n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1] + n[:, 2] * n[:, 2])
        want = (n / norm[:, None]).astype(np.float64)
    assert np.array_equal(got[:, 9:12], want, equal_nan=True)
    assert np.allclose(got[0, 9:12], [0.6, 0.8, 0.0], atol=1e-7) and np.isnan(got[1, 9:12]).all()
    # Geometry reads binary files by content
    assert np.array_equal(G.read_stl(path), got, equal_nan=True)
    # truncated: the library refuses, the oracle keeps the complete facets
    data = path.read_bytes()
    (tmp_path / "t.stl").write_bytes(data[:84 + 50 * 10 + 20])
    with pytest.raises(Exception):
        G.read_stl_binary(tmp_path / "t.stl")
    assert len(O.read_stl_binary(tmp_path / "t.stl")) == 10
    assert struct.unpack("<I", data[80:84])[0] == 500


def _tilted_box(lo, hi, theta):
    """Closed box [lo, hi] rotated by theta about the x axis: 12 triangles with outward normals,
    and the 6 half-spaces n . p <= d."""
    c, s = np.cos(theta), np.sin(theta)
    R = np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    corner = lambda b: R @ np.array([hi[0] if b & 1 else lo[0], hi[1] if b & 2 else lo[1], hi[2] if b & 4 else lo[2]])
    quads = [(0, 2, 6, 4, -1, 0), (1, 5, 7, 3, 1, 0), (0, 4, 5, 1, -1, 1), (2, 3, 7, 6, 1, 1),
             (0, 1, 3, 2, -1, 2), (4, 6, 7, 5, 1, 2)]
    tris, planes = [], []
    for a, b, cc, d, sign, ax in quads:
        n = R @ (sign * np.eye(3)[ax])
        P = [corner(v) for v in (a, b, cc, d)]
        tris += [tri(P[0], P[1], P[2], n), tri(P[0], P[2], P[3], n)]
        planes.append((n, n @ P[0]))
    return np.array(tris), planes


def test_link_fractions_closed_form_tilted_box():
    """Wall-link fractions (ComplexGeometry.cpp:550-608) on a convex box tilted 0.3 rad about x:
    a link from an inside node crosses the surface once, at t* = min over the faces it leaves
    through of (d - n.p) / (n.c), so q = t* |c|.  Faces whose normal has no z part are never
    hit (the point-in-triangle test projects along z), so links leaving through the x faces
    keep -1; so do links to points that are not solid and directions >= ndirs."""
    tris, planes = _tilted_box((0.37, -3.29, -2.71), (9.61, 3.83, 2.47), 0.3)
    h = 1.0
    origin = np.array([-1.0, -5.0, -5.0])
    dims = (13, 11, 11)
    k, j, i = np.meshgrid(*[np.arange(n) for n in dims[::-1]], indexing="ij")
    P = np.stack([origin[0] + i * h, origin[1] + j * h, origin[2] + k * h], -1)
    sd = np.stack([P @ n - d for n, d in planes], -1)
    assert np.abs(sd).min() > 1e-3  # no grid point on a face
    fluid = (sd < 0).all(-1)
    orc = O.OracleSim(fluid, origin, h, 19, [], tris, 0.8, 8e-6, normals=False)
    box, q = orc.link_fractions(tris, 19)
    assert len(box) > 50
    c, _, _ = O.lattice(19)
    nx, ny, nz = dims
    fl = fluid.reshape(-1)
    hit = 0
    for w, b in enumerate(box):
        ii, jj, kk = b % nx, (b // nx) % ny, b // (nx * ny)
        p = origin + np.array([ii, jj, kk]) * h
        for d in range(19):
            a = np.array([ii, jj, kk]) + c[d]
            if d == 18 or fl[a[0] + nx * (a[1] + ny * a[2])]:
                assert q[w, d] == -1.0
                continue
            cd = c[d].astype(float) * h
            ts = [((dd - n @ p) / (n @ cd), f) for f, (n, dd) in enumerate(planes) if n @ cd > 0]
            t, f = min(ts)
            if f < 2:  # x faces: vertical facets
                assert q[w, d] == -1.0
            else:
                assert q[w, d] == pytest.approx(t * np.linalg.norm(c[d]), abs=1e-12), (b, d)
                hit += 1
    assert hit > 100
    # the same nodes with only the first 6 directions: the diagonal columns stay -1
    _, q6 = orc.link_fractions(tris, 6)
    assert (q6[:, 6:] == -1).all() and np.array_equal(q6[:, :6], q[:, :6])
