"""The oracle against what the mathematics fixes, independently of the
reference's fixtures: isoparametric weights reproduce vertices, linear fields
and the point itself; cell volumes of unit cells and their affine images;
OBB axes of a rotated box; hole-map flood fill against breadth-first search;
coincident-point representatives; AMR linear interpolation of a linear field.
A dropped term, a wrong sign or a transposed index in the restatement fails
one of these."""
from collections import deque

import numpy as np
import pytest

from oracle.oracle import lib, ptr

# reference vertex orders (math.c:149-294, cellVolume.f90:20-109): tet 4,
# pyramid 5 (quad base then apex), prism 6 (two triangles), hex 8
UNIT = {
    4: np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float),
    5: np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 0.5, 1]], float),
    6: np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]], float),
    8: np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float),
}
UNIT_VOL = {4: 1 / 6, 5: 1 / 3, 6: 1 / 2, 8: 1.0}


def weights(xv: np.ndarray, xp: np.ndarray) -> np.ndarray:
    L = lib()
    a = np.zeros((8, 3))
    a[: xv.shape[0]] = xv
    f = np.zeros(8)
    L.orc_nodal_weights(ptr(a), ptr(np.ascontiguousarray(xp, float)), ptr(f), xv.shape[0])
    return f[: xv.shape[0]]


def volume(xv: np.ndarray) -> float:
    a = np.zeros((8, 3))
    a[: xv.shape[0]] = xv
    return lib().orc_cell_volume(ptr(a), xv.shape[0])


def affine(rng):
    A = np.eye(3) + 0.3 * rng.uniform(-1, 1, (3, 3))
    return A, rng.uniform(-2, 2, 3)


def interior_point(nv, rng):
    """A point strictly inside the unit reference cell."""
    while True:
        p = rng.uniform(0.05, 0.95, 3)
        if nv == 4 and p.sum() < 0.95:
            return p
        if nv == 6 and p[0] + p[1] < 0.95:
            return p
        if nv == 5 and abs(p[0] - 0.5) < 0.45 * (1 - p[2]) and abs(p[1] - 0.5) < 0.45 * (1 - p[2]):
            return p
        if nv == 8:
            return p


@pytest.mark.parametrize("nv", [4, 5, 6, 8])
def test_weights_at_vertices_are_unit_vectors(nv):
    rng = np.random.default_rng(nv)
    A, b = affine(rng)
    xv = UNIT[nv] @ A.T + b
    for k in range(nv):
        w = weights(xv, xv[k])
        e = np.zeros(nv)
        e[k] = 1
        assert np.allclose(w, e, atol=1e-10), (k, w)


@pytest.mark.parametrize("nv", [4, 5, 6, 8])
def test_weights_reproduce_point_and_linear_fields(nv):
    rng = np.random.default_rng(10 + nv)
    for _ in range(50):
        A, b = affine(rng)
        xv = UNIT[nv] @ A.T + b
        xp = interior_point(nv, rng) @ A.T + b
        w = weights(xv, xp)
        assert abs(w.sum() - 1.0) < 1e-12
        assert np.allclose(w @ xv, xp, atol=1e-12, rtol=0)
        c, g = rng.uniform(-1, 1), rng.uniform(-1, 1, 3)
        assert abs(w @ (c + xv @ g) - (c + xp @ g)) < 1e-12


def test_hex_centre_weights_are_one_eighth():
    rng = np.random.default_rng(3)
    A, b = affine(rng)
    xv = UNIT[8] @ A.T + b
    w = weights(xv, np.full(3, 0.5) @ A.T + b)
    assert np.allclose(w, 0.125, atol=1e-13)


@pytest.mark.parametrize("nv", [4, 5, 6, 8])
def test_volumes_of_unit_cells_and_affine_images(nv):
    v = volume(UNIT[nv])
    assert abs(abs(v) - UNIT_VOL[nv]) < 1e-14
    rng = np.random.default_rng(20 + nv)
    for _ in range(10):
        A, b = affine(rng)
        va = volume(UNIT[nv] @ A.T + b)
        assert abs(va - np.linalg.det(A) * v) < 1e-13 * max(1.0, abs(va))


def test_hex_volume_positive_in_reference_order():
    assert volume(UNIT[8]) > 0


def test_obb_of_rotated_box():
    rng = np.random.default_rng(4)
    L = lib()
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    half = np.array([3.0, 2.0, 0.5])
    c = rng.uniform(-1, 1, 3)
    g = np.stack(np.meshgrid(*[np.linspace(-h, h, 7) for h in half], indexing="ij"), -1).reshape(-1, 3)
    x = np.ascontiguousarray(g @ Q.T + c)
    xc, dxc, vec = np.zeros(3), np.zeros(3), np.zeros((3, 3))
    L.orc_find_obb(ptr(x), ptr(xc), ptr(dxc), ptr(vec), x.shape[0])
    assert np.allclose(xc, c, atol=1e-12)
    # each OBB axis is one of the box axes; the extents follow it (x 1.01)
    for k in range(3):
        dots = np.abs(Q.T @ vec[k])
        a = int(np.argmax(dots))
        assert dots[a] > 1 - 1e-10
        assert abs(dxc[k] - 1.01 * half[a]) < 1e-10


def test_obb_intersection_separated_and_overlapping():
    L = lib()
    I3 = np.ascontiguousarray(np.eye(3))
    d = np.ones(3)
    for shift, want in ((1.9, 1), (2.1, 0), (0.0, 1)):
        xa, xb = np.zeros(3), np.array([shift, 0.3, -0.2])
        got = L.orc_obb_intersect(ptr(I3), ptr(xa), ptr(d), ptr(I3), ptr(xb), ptr(d))
        assert got == want, (shift, got)


def bfs_fill(sam, nx, isym):
    """Breadth-first flood from the boundary shell through 0-cells with the
    isym stencil (isym=3: no z neighbours, 1: no x, 2: no y, 0: all six)."""
    n0, n1, n2 = nx
    s = sam.reshape(n2, n1, n0).copy()
    steps = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    if isym == 1:
        steps = [d for d in steps if d[0] == 0]
    elif isym == 2:
        steps = [d for d in steps if d[1] == 0]
    elif isym == 3:
        steps = [d for d in steps if d[2] == 0]
    out = np.ones_like(s, dtype=np.int32)  # 1 = not reached from outside (wall or hole)
    q = deque()
    for k in range(n2):
        for j in range(n1):
            for i in range(n0):
                if (i in (0, n0 - 1) or j in (0, n1 - 1) or k in (0, n2 - 1)):
                    out[k, j, i] = 0
                    q.append((i, j, k))
    while q:
        i, j, k = q.popleft()
        for di, dj, dk in steps:
            a, b, c = i + di, j + dj, k + dk
            if 0 <= a < n0 and 0 <= b < n1 and 0 <= c < n2 and s[c, b, a] == 0 and out[c, b, a] == 1:
                out[c, b, a] = 0
                q.append((a, b, c))
    return out.reshape(-1)


@pytest.mark.parametrize("isym", [0, 1, 2, 3])
def test_hole_fill_matches_breadth_first_search(isym):
    """fillHoleMap's in-place sweeps reach the fixed point a breadth-first
    flood reaches: a cell ends inside the hole iff it is a wall cell or a
    0-cell that no path of 0-cells joins to the boundary shell."""
    L = lib()
    rng = np.random.default_rng(30 + isym)
    for _ in range(6):
        nx = rng.integers(5, 12, 3).astype(np.int32)
        n = int(np.prod(nx))
        sam = np.zeros(n, np.int32)
        sam[rng.uniform(size=n) < 0.35] = 2  # wall-marked cells
        want = bfs_fill(sam, nx, isym)
        got = sam.copy()
        L.orc_fill_hole_map(ptr(got), ptr(nx), isym)
        assert np.array_equal((got != 0).astype(np.int32), want)


def test_coincident_points_representatives():
    """Every point's representative is an earlier (or the same) point at the
    same position with the same mesh tag; each distinct (position, tag) keeps
    exactly one representative of itself."""
    L = lib()
    rng = np.random.default_rng(5)
    base = rng.uniform(-1, 1, (300, 3))
    pick = rng.integers(0, 300, 2000)
    x = np.ascontiguousarray(base[pick])
    tag = rng.integers(1, 3, 2000).astype(np.int32)
    rtag = rng.uniform(0, 1, 2000)
    itag = np.zeros(2000, np.int32)
    r0 = rtag.copy()
    L.orc_unique_nodes_octree(ptr(x), ptr(tag), ptr(rtag), ptr(itag), 2000)
    assert np.all(itag <= np.arange(2000))
    assert np.array_equal(x[itag], x) and np.array_equal(tag[itag], tag)
    groups = {}
    for i in range(2000):
        groups.setdefault((pick[i], tag[i]), []).append(i)
    nrep = sum(1 for i in range(2000) if itag[i] == i)
    assert nrep == len(groups)
    for g in groups.values():
        assert itag[g[0]] == g[0]  # the first of each group represents itself
        assert rtag[g].max() <= r0[g].max() + 0 and rtag[g[0]] == r0[g].max()


def test_amr_interpolation_reproduces_linear_fields():
    """Cell-centred values of a linear function, interpolated by the patch
    stencils (linCartInterp.C), give the function at the receptor wherever
    the stencil is not clamped at a patch edge."""
    from oracle.oracle import Oracle
    from tioga_amd import workloads as W

    grid = W.amr_hierarchy((-1.0, -1.0, -1.0), 0.25, [((0, 0, 0), (8, 8, 8)), ((4, 4, 4), (12, 12, 12))],
                           patch=4, nf=1, shuffle_seed=7)
    body = W.all_receptors(W.latlong_shell(10, 6, 3, 0.2, 0.45, 1.1, tag=1, center=(0.05, -0.1, 0.02)))
    g = np.array([0.7, -1.3, 0.4])
    c0 = 0.25
    pq = []
    for gi in grid.patch_order():
        dims = grid.dims(int(gi))
        xlo, dx = grid.rdata[gi, :3], grid.rdata[gi, 3:]
        nf = grid.nf
        ii = [np.arange(-nf, d + nf) for d in dims]
        K, J, I = np.meshgrid(ii[2], ii[1], ii[0], indexing="ij")
        xc = np.stack([xlo[0] + (I + 0.5) * dx[0], xlo[1] + (J + 0.5) * dx[1], xlo[2] + (K + 0.5) * dx[2]], -1)
        pq.append(np.ascontiguousarray((c0 + xc.reshape(-1, 3) @ g)))
    o = Oracle([body])
    o.register_amr(grid, pq)
    o.profile()
    assert o.connect_amr() == 0
    q = np.zeros(body.nnodes)
    q = o.data_update_amr([q], 1, 0)[0]
    ijk = o.amr_array("rec_ijk").reshape(-1, 8, 3)
    info = o.amr_array("rec_info").reshape(-1, 3)
    rid = o.array(1, "receptorIdCart")
    unclamped = np.array([len({tuple(c) for c in cells}) == 8 for cells in ijk])
    assert unclamped.sum() > 50
    nodes = rid[info[unclamped, 1]]
    want = c0 + body.x[nodes] @ g
    assert np.allclose(q[nodes], want, atol=1e-12)
    o.close()
