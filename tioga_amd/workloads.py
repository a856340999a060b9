"""Seeded synthetic overset meshes for the connectivity workloads.

Every generator returns :class:`Block` objects in TIOGA's registration
layout (tioga.h:121-123): FP64 AoS coordinates, 1-based connectivity per cell
type (codetypes.h:36 ``BASE 1``), 1-based wall / overset-boundary node lists.

Workloads (sizes as stated; all coordinates FP64):

* ``hex48_pair``  background 48^3 hex cube [-1,1]^3 (tag 2) + lat-long spherical
  hex shell, r 0.25..0.6, 64 azimuth x 32 polar x 24 radial cells, radial
  stretching 1.1 (tag 1).  Pole nodes are kept per azimuth (coincident
  copies), which exercises the coincident-query dedupe.
* ``hex256_mixed_pair``  background 256^3 hex cube with interior nodes jittered by
  uniform +-10% of the spacing (tag 2) + a mixed-element shell (about 3.9M
  cells: 40% tet, 30% prism, 5% pyramid, 25% hex by cell count) between an
  ellipsoid wall with semi-axes (0.3, 0.2, 0.2) and a sphere of radius 0.6
  (tag 1).  Poles are merged, so pole rows are prisms.
* ``tiny``     the hex48_pair construction at a size the oracle finishes in a
  fraction of a second (parity cases).
* ``shells64_hex192``  batching: 64 hex/prism spherical shells (tags 1..64) whose
  cell counts are log-uniform in [1e4, 1e6], wall radius uniform in
  [0.02, 0.08], outer radius 2.5 x the wall radius, centres uniform in
  [-0.8, 0.8]^3, uniformly random rotations (Shoemake's quaternion method),
  seed 3; background 192^3 hex cube [-1,1]^3 (tag 65).  Each shell is the
  mixed_shell construction with a spherical wall and only hex and prism
  layers (its polar rows are prisms).
* ``serve16_hex256``  serving: the 256^3 jittered background of hex256_mixed_pair (tag 17)
  and 16 hex/prism spherical shells (tags 1..16) of log-uniform size in
  [1e4, 2e6] cells, wall radius uniform in [0.02, 0.08], centres uniform in
  [-0.6, 0.6]^3, seed 4; requests move one shell rigidly (rotation by an
  angle uniform in [0, 5] degrees about a uniformly random axis, about the
  shell's centroid, then a shift of uniformly random direction and length
  uniform in [0, 0.02]) and arrive as a Poisson stream.
* ``amr3_mixed_shell``  AMR Cartesian background: 3 levels (refinement ratio 2) of
  4096 patches of 16^3 cells with one ghost layer on [-1,1]^3 (level 0: the
  whole cube at h = 1/64, 512 patches; level 1: [-1,1]x[-0.75,0.75]x
  [-0.5,0.5] at h/2, 1536 patches; level 2: [-0.5,0.5]^2x[-0.25,0.25] at
  h/4, 2048 patches), pdegree 0, 8 FP64 fields per patch.  Receptors are the
  nodes of the mixed-element near-body shell of hex256_mixed_pair (the construction
  named for this workload; hex48_pair's shell is all-hex), every one of them
  marked unresolved by a node resolution of BIGVALUE, so every node takes its
  value from the patches.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np


@dataclass
class Block:
    tag: int
    x: np.ndarray                      # (nnodes, 3) float64
    types: List[Tuple[int, np.ndarray]]  # [(nv, conn (nc, nv) int32, 1-based)]
    wbc: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    obc: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    nres: Optional[np.ndarray] = None
    cres: Optional[np.ndarray] = None
    cell_gid: Optional[np.ndarray] = None  # uint64 global ids (registerGridData, tioga.h:121-123)
    node_gid: Optional[np.ndarray] = None

    @property
    def nnodes(self) -> int:
        return int(self.x.shape[0])

    @property
    def ncells(self) -> int:
        return int(sum(c.shape[0] for _, c in self.types))

    def normalized(self) -> "Block":
        """Contiguous arrays of the exact dtypes the C ABI binds."""
        return Block(
            self.tag,
            np.ascontiguousarray(self.x, dtype=np.float64),
            [(int(nv), np.ascontiguousarray(c, dtype=np.int32)) for nv, c in self.types],
            np.ascontiguousarray(self.wbc, dtype=np.int32),
            np.ascontiguousarray(self.obc, dtype=np.int32),
            None if self.nres is None else np.ascontiguousarray(self.nres, dtype=np.float64),
            None if self.cres is None else np.ascontiguousarray(self.cres, dtype=np.float64),
            None if self.cell_gid is None else np.ascontiguousarray(self.cell_gid, dtype=np.uint64),
            None if self.node_gid is None else np.ascontiguousarray(self.node_gid, dtype=np.uint64),
        )


def signed_volume(x: np.ndarray, nv: int, conn: np.ndarray) -> np.ndarray:
    """Cell volume by the divergence theorem, the face split of
    cellVolume.f90:67-91 (vectorised; used only to check orientation)."""
    faces = {
        4: [(1, 2, 3), (1, 4, 2), (2, 4, 3), (1, 3, 4)],
        5: [(1, 2, 3, 4), (1, 5, 2), (2, 5, 3), (4, 3, 5), (1, 4, 5)],
        6: [(1, 2, 3), (1, 4, 5, 2), (2, 5, 6, 3), (1, 3, 6, 4), (4, 6, 5)],
        8: [(1, 2, 3, 4), (1, 5, 6, 2), (2, 6, 7, 3), (3, 7, 8, 4), (1, 4, 8, 5), (5, 8, 7, 6)],
    }[nv]
    v = x[conn - 1]

    def bp(a, b, c):
        return np.einsum("ij,ij->i", a, np.cross(b, c))

    vol = np.zeros(conn.shape[0])
    for f in faces:
        p = [v[:, k - 1] for k in f]
        if len(f) == 3:
            vol -= 0.5 * bp(p[0], p[1], p[2])
        else:
            vol -= 0.25 * (bp(p[0], p[1], p[2]) + bp(p[0], p[2], p[3]) + bp(p[0], p[1], p[3]) + bp(p[1], p[2], p[3]))
    return vol / 3.0


def cart_box(n: int, tag: int, lo: float = -1.0, hi: float = 1.0, jitter: float = 0.0,
             seed: int = 2) -> Block:
    """n^3 hex cells on [lo,hi]^3; node id 1 + i + (n+1)(j + (n+1)k); hex
    vertices (i,j,k),(i+1,j,k),(i+1,j+1,k),(i,j+1,k) then the same at k+1
    (the vertex order of poisson_utils.f90:77-88).  Interior nodes are moved
    by an independent uniform +-jitter*h per coordinate."""
    m = n + 1
    g = lo + (hi - lo) * np.arange(m, dtype=np.float64) / n
    k, j, i = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    x = np.empty((m * m * m, 3))
    x[:, 0] = g[i.ravel()]
    x[:, 1] = g[j.ravel()]
    x[:, 2] = g[k.ravel()]
    del k, j, i
    if jitter > 0.0:
        h = (hi - lo) / n
        rng = np.random.default_rng(seed)
        idx = np.arange(m)
        inner = (idx > 0) & (idx < n)
        mask = (inner[:, None, None] & inner[None, :, None] & inner[None, None, :]).ravel()
        d = rng.uniform(-jitter * h, jitter * h, size=(int(mask.sum()), 3))
        x[mask] += d
    c = np.arange(n)
    kk, jj, ii = np.meshgrid(c, c, c, indexing="ij")
    base = (ii + m * (jj + m * kk)).ravel().astype(np.int64)
    del kk, jj, ii
    off = np.array([0, 1, 1 + m, m, m * m, 1 + m * m, 1 + m + m * m, m + m * m], np.int64)
    conn = (base[:, None] + off[None, :] + 1).astype(np.int32)
    return Block(tag, x, [(8, conn)])


def partition(b: Block, cell_mask: np.ndarray) -> Block:
    """The partition of block b holding the cells where cell_mask (one flag
    per cell, in type order) is set, as a rank of a partitioned mesh holds
    it: the nodes those cells use, in ascending original order, the
    connectivity renumbered to them, the wall and overset-boundary nodes and
    user resolutions and global ids that fall in the partition.  Nodes on a
    cut belong to every partition that touches them."""
    cell_mask = np.asarray(cell_mask, bool)
    used = np.zeros(b.nnodes, bool)
    c0 = 0
    for nv, c in b.types:
        m = cell_mask[c0:c0 + c.shape[0]]
        used[(c[m] - 1).ravel()] = True
        c0 += c.shape[0]
    remap = np.full(b.nnodes, -1, np.int64)
    keep = np.nonzero(used)[0]
    remap[keep] = np.arange(keep.size)
    types, c0 = [], 0
    for nv, c in b.types:
        m = cell_mask[c0:c0 + c.shape[0]]
        c0 += c.shape[0]
        if m.any():
            types.append((nv, (remap[c[m] - 1] + 1).astype(np.int32)))

    def nodes(lst):
        lst = np.asarray(lst, np.int64) - 1
        lst = lst[used[lst]]
        return (remap[lst] + 1).astype(np.int32)

    return Block(b.tag, b.x[keep].copy(), types, nodes(b.wbc), nodes(b.obc),
                 None if b.nres is None else b.nres[keep].copy(),
                 None if b.cres is None else b.cres[cell_mask].copy(),
                 None if b.cell_gid is None else b.cell_gid[cell_mask].copy(),
                 None if b.node_gid is None else b.node_gid[keep].copy())


def _radii(r0: float, r1: float, nr: int, stretch: float) -> np.ndarray:
    k = np.arange(nr + 1, dtype=np.float64)
    if stretch == 1.0:
        return r0 + (r1 - r0) * k / nr
    dr0 = (r1 - r0) * (stretch - 1.0) / (stretch ** nr - 1.0)
    return r0 + dr0 * (stretch ** k - 1.0) / (stretch - 1.0)


def latlong_shell(nphi: int, ntheta: int, nr: int, r0: float, r1: float, stretch: float, tag: int,
                  center=(0.0, 0.0, 0.0)) -> Block:
    """Hex shell, phi_i = 2 pi i / nphi (seam merged), theta_j = pi j / ntheta,
    r_k geometric from r0 to r1.  Pole nodes are kept per azimuth.  Node id
    1 + i + nphi (j + (ntheta+1) k).  Cell vertex order runs (r, theta, phi),
    a right-handed frame, so cell volumes are positive.  wbc = k=0 nodes,
    obc = k=nr nodes."""
    r = _radii(r0, r1, nr, stretch)
    th = np.pi * np.arange(ntheta + 1) / ntheta
    ph = 2.0 * np.pi * np.arange(nphi) / nphi
    K, J, I = np.meshgrid(np.arange(nr + 1), np.arange(ntheta + 1), np.arange(nphi), indexing="ij")
    rr, tt, pp = r[K.ravel()], th[J.ravel()], ph[I.ravel()]
    x = np.stack([rr * np.sin(tt) * np.cos(pp), rr * np.sin(tt) * np.sin(pp), rr * np.cos(tt)], axis=1)
    x += np.asarray(center, np.float64)[None, :]

    def nid(i, j, k):
        return 1 + (i % nphi) + nphi * (j + (ntheta + 1) * k)

    K, J, I = np.meshgrid(np.arange(nr), np.arange(ntheta), np.arange(nphi), indexing="ij")
    k, j, i = K.ravel(), J.ravel(), I.ravel()
    conn = np.stack([nid(i, j, k), nid(i, j, k + 1), nid(i, j + 1, k + 1), nid(i, j + 1, k),
                     nid(i + 1, j, k), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i + 1, j + 1, k)],
                    axis=1).astype(np.int32)
    nlayer = nphi * (ntheta + 1)
    wbc = (np.arange(nlayer) + 1).astype(np.int32)
    obc = (np.arange(nlayer) + 1 + nr * nlayer).astype(np.int32)
    return Block(tag, x, [(8, conn)], wbc, obc)


# Kuhn split of a hex (local vertices 0..7 in TIOGA order) into 6 positive
# tets sharing the 0-6 diagonal; the split is face-conforming between
# neighbours of one structured grid.
_KUHN = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]
# hex -> 2 prisms cut through the 0-2 / 4-6 diagonal plane
_PRISMS = [(0, 1, 2, 4, 5, 6), (0, 2, 3, 4, 6, 7)]
# hex -> 3 pyramids with apex at vertex 0 on the faces not containing it
_PYRS = [(1, 5, 6, 2, 0), (2, 6, 7, 3, 0), (4, 7, 6, 5, 0)]


def _orient(x: np.ndarray, nv: int, conn: np.ndarray) -> np.ndarray:
    """Reorder cells with negative volume so every cell is positively oriented."""
    vol = signed_volume(x, nv, conn)
    neg = vol < 0
    if neg.any():
        c = conn.copy()
        if nv == 4:
            c[neg] = c[neg][:, [1, 0, 2, 3]]
        elif nv == 5:
            c[neg] = c[neg][:, [0, 3, 2, 1, 4]]
        elif nv == 6:
            c[neg] = c[neg][:, [3, 4, 5, 0, 1, 2]]
        else:
            c[neg] = c[neg][:, [4, 5, 6, 7, 0, 1, 2, 3]]
        conn = c
    return conn


def mixed_shell(nphi: int, ntheta: int, layers=(15, 9, 1, 4), tag: int = 1,
                wall_axes=(0.3, 0.2, 0.2), r_outer: float = 0.6, stretch: float = 1.08) -> Block:
    """Mixed-element shell from an ellipsoid wall to a sphere.  Radial
    layers, wall outward: hex, prism (2 per hex), pyramid (3 per hex), tet (6
    per hex, Kuhn).  Poles are merged, so the two polar rows of every layer are
    prisms.  With layers (15, 9, 1, 4) the cell counts are 25/30/5/40 % hex/
    prism/pyramid/tet."""
    lh, lp, ly, lt = layers
    nr = lh + lp + ly + lt
    t = _radii(0.0, 1.0, nr, stretch)
    th = np.pi * np.arange(1, ntheta) / ntheta           # interior rings only
    ph = 2.0 * np.pi * np.arange(nphi) / nphi
    a, b, c = wall_axes
    # ring nodes per layer: (ntheta-1)*nphi, plus 2 poles
    TT, PP = np.meshgrid(th, ph, indexing="ij")
    dirs = np.stack([np.sin(TT) * np.cos(PP), np.sin(TT) * np.sin(PP), np.cos(TT)], -1).reshape(-1, 3)
    dirs = np.concatenate([dirs, [[0.0, 0.0, 1.0], [0.0, 0.0, -1.0]]], 0)
    inner = dirs * np.array([a, b, c])[None, :]
    outer = dirs * r_outer
    nl = dirs.shape[0]
    x = (inner[None, :, :] * (1.0 - t)[:, None, None] + outer[None, :, :] * t[:, None, None]).reshape(-1, 3)
    north, south = nl - 2, nl - 1

    def nid(i, j, k):
        # j in 0..ntheta (0 = north pole, ntheta = south pole)
        ring = np.where(j == 0, north, np.where(j == ntheta, south, (j - 1) * nphi + (i % nphi)))
        return 1 + ring + nl * k

    out = {4: [], 5: [], 6: [], 8: []}
    for k in range(nr):
        J, I = np.meshgrid(np.arange(ntheta), np.arange(nphi), indexing="ij")
        j, i = J.ravel(), I.ravel()
        kk = np.full_like(j, k)
        hexv = np.stack([nid(i, j, kk), nid(i, j, kk + 1), nid(i, j + 1, kk + 1), nid(i, j + 1, kk),
                         nid(i + 1, j, kk), nid(i + 1, j, kk + 1), nid(i + 1, j + 1, kk + 1),
                         nid(i + 1, j + 1, kk)], 1)
        pole = (j == 0) | (j == ntheta - 1)
        # polar rows: the hex has a collapsed edge -> prism with distinct nodes
        pn = hexv[pole & (j == 0)]
        if pn.size:   # north row: hex vertices 0,4 and 1,5 coincide (j = 0)
            out[6].append(pn[:, [0, 3, 7, 1, 2, 6]])
        ps = hexv[pole & (j == ntheta - 1)]
        if ps.size:   # south row: vertices 3 and 7 coincide (j+1 = ntheta)
            out[6].append(ps[:, [0, 3, 4, 1, 2, 5]])
        hv = hexv[~pole]
        if k < lh:
            out[8].append(hv)
        elif k < lh + lp:
            out[6].extend(hv[:, list(p)] for p in _PRISMS)
        elif k < lh + lp + ly:
            out[5].extend(hv[:, list(p)] for p in _PYRS)
        else:
            out[4].extend(hv[:, list(p)] for p in _KUHN)
    types = []
    for nv in (4, 5, 6, 8):
        if out[nv]:
            conn = np.concatenate(out[nv], 0).astype(np.int32)
            types.append((nv, _orient(x, nv, conn)))
    wbc = (np.arange(nl) + 1).astype(np.int32)
    obc = (np.arange(nl) + 1 + nr * nl).astype(np.int32)
    return Block(tag, x, types, wbc, obc)


def fields(nnodes: int, nvar: int, seed: int) -> np.ndarray:
    """nvar FP64 nodal fields, uniform[-1,1], row layout (node-major)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, size=nnodes * nvar)


def hex48_pair() -> List[Block]:
    return [latlong_shell(64, 32, 24, 0.25, 0.6, 1.1, tag=1), cart_box(48, tag=2)]


def tiny(nbox: int = 10, nphi: int = 16, ntheta: int = 8, nr: int = 6) -> List[Block]:
    return [latlong_shell(nphi, ntheta, nr, 0.25, 0.6, 1.1, tag=1), cart_box(nbox, tag=2)]


def hex256_mixed_pair(nbox: int = 256, nphi: int = 360, ntheta: int = 180) -> List[Block]:
    return [mixed_shell(nphi, ntheta, tag=1), cart_box(nbox, tag=2, jitter=0.1, seed=2)]


def rigid_motion(x: np.ndarray, angle_deg: float, axis, shift) -> np.ndarray:
    """Rotate about the centroid by angle_deg around axis, then translate."""
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    th = np.deg2rad(angle_deg)
    kx = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    R = np.eye(3) + np.sin(th) * kx + (1 - np.cos(th)) * kx @ kx
    c = x.mean(axis=0)
    return (x - c) @ R.T + c + np.asarray(shift, np.float64)


BIGVALUE = 1.0e15  # codetypes.h:41


@dataclass
class AmrGrid:
    """A Cartesian patch hierarchy in the layout tioga_register_amr_global_data_
    reads (tiogaInterface.C:132-136, CartGrid.C:22-70): per grid 11 ints
    (global id, level, proc, porder, local id, ilo[3], ihi[3]) and 6 doubles
    (xlo[3], dx[3]).  One rank: proc 0, and patch ``local id`` is registered
    with ``tioga_register_amr_local_data_(local id, global id, ...)``."""
    nf: int
    idata: np.ndarray   # (ngrids, 11) int32
    rdata: np.ndarray   # (ngrids, 6) float64
    qnode: np.ndarray = field(default_factory=lambda: np.zeros(1))
    qstride: int = 1

    @property
    def ngrids(self) -> int:
        return int(self.idata.shape[0])

    def dims(self, g: int) -> np.ndarray:
        return self.idata[g, 8:11] - self.idata[g, 5:8] + 1

    def d3nf(self, g: int) -> int:
        return int(np.prod(self.dims(g) + 2 * self.nf))

    def patch_order(self) -> np.ndarray:
        """Global id of each local patch (local ids 0..ngrids-1)."""
        out = np.zeros(self.ngrids, np.int64)
        out[self.idata[:, 4]] = np.arange(self.ngrids)
        return out

    def patch_fields(self, nvar: int, seed: int) -> List[np.ndarray]:
        """Per local patch: nvar fields uniform[-1,1], field n at
        n*d3nf (get_amr_index_xyz.c:27-45), ghosts included."""
        rng = np.random.default_rng(seed)
        return [rng.uniform(-1.0, 1.0, self.d3nf(int(g)) * nvar) for g in self.patch_order()]


def amr_hierarchy(origin, h0: float, level_boxes, patch: int = 16, nf: int = 1,
                  shuffle_seed: Optional[int] = None, local_seed: Optional[int] = None) -> AmrGrid:
    """Patches of ``patch``^3 cells tiling one box per level.  Level l has
    spacing h0/2^l; ``level_boxes[l]`` = (lo[3], hi[3]) in level-l cell indices
    (hi exclusive).  Grids are listed level by level, (k, j, i) within a level,
    unless ``shuffle_seed`` permutes the list; ``local_seed`` permutes the
    local ids."""
    rows, reals = [], []
    for l, (lo, hi) in enumerate(level_boxes):
        h = h0 / 2 ** l
        lo, hi = np.asarray(lo), np.asarray(hi)
        for k in range(lo[2], hi[2], patch):
            for j in range(lo[1], hi[1], patch):
                for i in range(lo[0], hi[0], patch):
                    a = np.array([i, j, k])
                    b = np.minimum(a + patch, hi) - 1
                    rows.append([0, l, 0, 0, 0, *a, *b])
                    reals.append([origin[0] + a[0] * h, origin[1] + a[1] * h, origin[2] + a[2] * h, h, h, h])
    idata = np.array(rows, np.int32)
    rdata = np.array(reals, np.float64)
    if shuffle_seed is not None:
        p = np.random.default_rng(shuffle_seed).permutation(len(rows))
        idata, rdata = idata[p], rdata[p]
    n = idata.shape[0]
    idata[:, 0] = np.arange(n)
    lid = np.arange(n) if local_seed is None else np.random.default_rng(local_seed).permutation(n)
    idata[:, 4] = lid
    return AmrGrid(nf, np.ascontiguousarray(idata), np.ascontiguousarray(rdata))


def all_receptors(b: Block) -> Block:
    """The block with every node marked unresolved (node resolution
    BIGVALUE, cell resolution 1), so that every node becomes a receptor of
    the Cartesian background (cartOps.C:73-97)."""
    b.nres = np.full(b.nnodes, BIGVALUE)
    b.cres = np.ones(b.ncells)
    return b


def amr3_grid() -> AmrGrid:
    """The patch hierarchy of the AMR gather workload."""
    return amr_hierarchy((-1.0, -1.0, -1.0), 2.0 / 128,
                         [((0, 0, 0), (128, 128, 128)),
                          ((0, 32, 64), (256, 224, 192)),
                          ((128, 128, 192), (384, 384, 320))], patch=16, nf=1)


def amr3_mixed_shell():
    """(near-body blocks, AmrGrid) of the AMR gather workload."""
    return [all_receptors(mixed_shell(360, 180, tag=1))], amr3_grid()


def shoemake_rotation(u) -> np.ndarray:
    """Uniformly random rotation from three uniform[0,1) numbers (K. Shoemake,
    "Uniform random rotations", Graphics Gems III, 1992)."""
    u1, u2, u3 = u
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    w, x, y, z = (a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3))
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def sphere_shell(ncells: float, r_wall: float, tag: int, center, rot) -> Block:
    """Hex/prism shell of about ncells cells around a sphere of radius r_wall
    out to 2.5 r_wall: nphi = 2 ntheta azimuths, ntheta polar rows and
    ntheta/2 radial layers (half hex, half prism), rotated by rot and moved
    to center."""
    nth = max(4, int(round(ncells ** (1.0 / 3.0))))
    nr = max(2, nth // 2)
    lh = nr // 2
    b = mixed_shell(2 * nth, nth, layers=(lh, nr - lh, 0, 0), tag=tag, wall_axes=(r_wall, r_wall, r_wall),
                    r_outer=2.5 * r_wall, stretch=1.05)
    b.x = b.x @ np.asarray(rot).T + np.asarray(center, np.float64)[None, :]
    return b


def shells64_hex192(nshell: int = 64, nbox: int = 192, seed: int = 3, lo: float = 1e4, hi: float = 1e6) -> List[Block]:
    rng = np.random.default_rng(seed)
    blocks = []
    for t in range(1, nshell + 1):
        n = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        r = float(rng.uniform(0.02, 0.08))
        c = rng.uniform(-0.8, 0.8, 3)
        rot = shoemake_rotation(rng.uniform(size=3))
        blocks.append(sphere_shell(n, r, t, c, rot))
    blocks.append(cart_box(nbox, tag=nshell + 1))
    return blocks


def rigid_apply(x: np.ndarray, rot, center, shift) -> np.ndarray:
    """x' = R (x - c) + c + t per node, in the operation order the serving
    loop uses (each row of R applied left to right, then + c, then + t)."""
    R = np.asarray(rot, np.float64).reshape(3, 3)
    c = np.asarray(center, np.float64)
    t = np.asarray(shift, np.float64)
    d = [x[:, k] - c[k] for k in range(3)]
    out = np.empty_like(x)
    for j in range(3):
        out[:, j] = R[j, 0] * d[0] + R[j, 1] * d[1] + R[j, 2] * d[2] + c[j] + t[j]
    return out


def axis_angle(axis, angle_deg) -> np.ndarray:
    """Rotation matrix for angle_deg about axis (Rodrigues)."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    th = np.deg2rad(angle_deg)
    k = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(th) * k + (1 - np.cos(th)) * k @ k


def serve16_shells(nshell: int = 16, seed: int = 4, lo: float = 1e4, hi: float = 2e6) -> List[Block]:
    rng = np.random.default_rng(seed)
    out = []
    for t in range(1, nshell + 1):
        n = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        r = float(rng.uniform(0.02, 0.08))
        c = rng.uniform(-0.6, 0.6, 3)
        out.append(sphere_shell(n, r, t, c, shoemake_rotation(rng.uniform(size=3))))
    return out


def motion_requests(n: int, nshell: int, seed: int):
    """n requests: (shell tag, rotation matrix, shift), rotation uniform in
    [0, 5] degrees about a uniformly random axis, shift of uniformly random
    direction and length uniform in [0, 0.02]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        tag = int(rng.integers(1, nshell + 1))
        ax = rng.normal(size=3)
        R = axis_angle(ax, rng.uniform(0, 5))
        d = rng.normal(size=3)
        t = d / np.linalg.norm(d) * rng.uniform(0, 0.02)
        out.append((tag, R, t))
    return out
