"""Parity cases shared by the golden-fixture script and the tests.

Each case is a list of blocks (tioga_amd.workloads.Block), the nodal fields
for dataUpdate and the reference's runtime parameters (isym, nfringe,
mexclude: tioga.h:106-107 defaults 3, 1, 3; interptype 0 = row, 1 = column,
tiogaInterface.C:178-190), plus optional rigid motions of blocks between
connectivity steps.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from tioga_amd import workloads as W


@dataclass
class Case:
    name: str
    blocks: list
    nvar: int
    interptype: int = 0
    isym: int = 3
    nfringe: int = 1
    mexclude: int = 3
    moves: List[List[Optional[np.ndarray]]] = field(default_factory=list)
    reduce: bool = False  # tioga::reduce_fringes after each connectivity
    ranks: Optional[List[int]] = None  # MPI rank of each block (multi-rank cases)
    # performConnectivityHighOrder and a point update after each step: 1
    # without callbacks, 2 with tests/cxx/ho_callbacks.c (two single-type
    # blocks, block 0 the receptor mesh, block 1 the donor mesh)
    points: int = 0

    @property
    def qs(self) -> List[np.ndarray]:
        return [W.fields(b.nnodes, self.nvar, seed=11 + i) for i, b in enumerate(self.blocks)]


def _user_res(blocks):
    out = []
    for b in blocks:
        rng = np.random.default_rng(100 + b.tag)
        b.nres = rng.uniform(1e-4, 2e-3, b.nnodes)
        b.cres = rng.uniform(1e-4, 2e-3, b.ncells)
        out.append(b)
    return out


def _moves(blocks, nsteps, seed):
    rng = np.random.default_rng(seed)
    steps = []
    x = blocks[0].x.copy()
    for _ in range(nsteps):
        ax = rng.normal(size=3)
        x = W.rigid_motion(x, rng.uniform(0, 5), ax, rng.uniform(-0.02, 0.02, 3))
        steps.append([x.copy()] + [None] * (len(blocks) - 1))
    return steps


def three_blocks():
    a = W.latlong_shell(16, 8, 5, 0.2, 0.45, 1.1, tag=1, center=(-0.15, 0.05, 0.0))
    b = W.latlong_shell(12, 8, 5, 0.2, 0.4, 1.05, tag=2, center=(0.2, -0.05, 0.03))
    return [a, b, W.cart_box(12, tag=3)]


def nested_boxes(shell: bool = False):
    """A coarse box with a finer jittered box inside it (outer faces of the
    fine box overset boundaries): every coarse node well inside the fine box
    finds a finer donor, so the coarse mesh holds a thick receptor region
    that reduce_fringes thins; optionally a walled shell inside both."""
    coarse = W.cart_box(10, tag=1)
    n = 20
    fine = W.cart_box(n, tag=2, lo=-0.61, hi=0.59, jitter=0.05, seed=4)
    m = n + 1
    k, j, i = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    face = ((i == 0) | (i == n) | (j == 0) | (j == n) | (k == 0) | (k == n)).ravel()
    fine.obc = (np.nonzero(face)[0] + 1).astype(np.int32)
    out = [coarse, fine]
    if shell:
        out.append(W.latlong_shell(12, 8, 5, 0.12, 0.3, 1.1, tag=3, center=(0.02, -0.03, 0.01)))
    return out


def _position_gids(blocks, base: int = 1 << 33):
    """Node global ids from positions (coordinates rounded to 1e-9): nodes
    at the same place, in one block (the shell's per-azimuth pole copies)
    or in two (the seam of a split box), share an id, as a global node
    numbering gives them; ids above 2^32.  Cell ids (tag << 40) + 3*cell + 1."""
    ids = {}
    for b in blocks:
        key = np.round(b.x * 1e9).astype(np.int64)
        g = np.empty(b.nnodes, np.uint64)
        for i, k in enumerate(map(tuple, key)):
            g[i] = ids.setdefault(k, base + len(ids))
        b.node_gid = g
        b.cell_gid = (np.uint64(b.tag) << np.uint64(40)) + np.uint64(3) * np.arange(b.ncells, dtype=np.uint64) + \
            np.uint64(1)
    return blocks


def split_box(n: int, cut: int, tags=(2, 3)):
    """The n^3 box of cart_box cut at cell column i = cut into two blocks of
    their own tags; the nodes of the cut plane belong to both (abutting
    partitions as separate meshes)."""
    full = W.cart_box(n, tag=tags[0])
    conn = full.types[0][1] - 1
    m = n + 1
    ci = np.arange(n ** 3) % n
    out = []
    for tag, sel in zip(tags, (ci < cut, ci >= cut)):
        c = conn[sel]
        used = np.unique(c)
        remap = np.full(m ** 3, -1, np.int64)
        remap[used] = np.arange(used.size)
        out.append(W.Block(tag, full.x[used].copy(), [(8, (remap[c] + 1).astype(np.int32))]))
    return out


def ranked_blocks(three: bool = False):
    """Partitioned meshes on several MPI ranks, each rank holding at most
    one block per mesh tag.  Two ranks: a lat-long shell (tag 1) on rank 0
    and a 12^3 box (tag 2) cut at cell column 5 into a partition on each rank
    (the cut nodes are sent by both, coincident queries of one mesh).  Three
    ranks: an off-centre shell (tag 1) cut radially into an inner partition
    with the wall (rank 2) and an outer one without wall nodes (rank 0), so
    the wall bounds of tag 1 are reduced with the zeros the wall-less
    partition reports (MeshBlock.C:640-647, holeMap.C:110-112); a second
    shell (tag 3) on rank 1; the box cut in three, one partition per rank.
    The blocks are listed out of rank order (each rank registers its own in
    list order)."""
    if not three:
        shell = W.latlong_shell(12, 8, 5, 0.2, 0.45, 1.1, tag=1, center=(0.02, 0.05, 0.01))
        box = W.cart_box(12, tag=2)
        ci = np.arange(box.ncells) % 12
        parts = [W.partition(box, ci < 5), W.partition(box, ci >= 5)]
        return [parts[1], shell, parts[0]], [1, 0, 0]
    s1 = W.latlong_shell(12, 8, 6, 0.12, 0.3, 1.1, tag=1, center=(0.42, 0.1, -0.05))
    kr = np.arange(s1.ncells) // (12 * 8)  # radial layer of each cell
    s3 = W.latlong_shell(10, 6, 4, 0.1, 0.25, 1.05, tag=3, center=(-0.35, -0.2, 0.1))
    box = W.cart_box(14, tag=2)
    ci = np.arange(box.ncells) % 14
    b0, b1, b2 = (W.partition(box, m) for m in (ci < 4, (ci >= 4) & (ci < 9), ci >= 9))
    inner, outer = W.partition(s1, kr < 3), W.partition(s1, kr >= 3)
    return [b1, outer, s3, b0, inner, b2], [1, 0, 1, 0, 2, 2]


def build(name: str) -> Case:
    if name == "tiny":
        return Case(name, W.tiny(), nvar=5)
    if name == "tiny_isym0":
        return Case(name, W.tiny(), nvar=2, interptype=1, isym=0, nfringe=2, mexclude=1)
    if name == "tiny_isym1":
        return Case(name, W.tiny(nbox=9, nphi=12, ntheta=6, nr=5), nvar=1, isym=1, nfringe=0, mexclude=0)
    if name == "tiny_isym2":
        return Case(name, W.tiny(nbox=11, nphi=10, ntheta=8, nr=4), nvar=3, isym=2, nfringe=3, mexclude=2)
    if name == "mixed":
        return Case(name, [W.mixed_shell(12, 8, layers=(2, 2, 1, 2)),
                           W.cart_box(12, tag=2, jitter=0.1, seed=7)], nvar=3)
    if name == "userres":
        return Case(name, _user_res(W.tiny()), nvar=4)
    if name == "moving":
        blocks = W.tiny()
        return Case(name, blocks, nvar=2, moves=_moves(blocks, 2, seed=9))
    if name == "three":
        return Case(name, three_blocks(), nvar=3)
    if name == "bigtag":  # a mesh tag above 127: receptor iblanks of -200
        blocks = W.tiny()
        blocks[1].tag = 200
        return Case(name, blocks, nvar=2)
    if name == "hex48_pair":
        return Case(name, W.hex48_pair(), nvar=5)
    # reduce_fringes (tioga.C:737-885) after the connectivity: fringe depths
    # nfringe + 1 = 2, 3 and 4 sweeps, mixed cell types, three blocks, and a
    # moving body reduced at every step
    if name == "tiny_rf":
        return Case(name, W.tiny(), nvar=3, reduce=True)
    if name == "isym0_rf":
        return Case(name, W.tiny(), nvar=2, interptype=1, isym=0, nfringe=2, mexclude=1, reduce=True)
    if name == "mixed_rf":
        return Case(name, [W.mixed_shell(12, 8, layers=(2, 2, 1, 2)),
                           W.cart_box(12, tag=2, jitter=0.1, seed=7)], nvar=2, nfringe=3, reduce=True)
    if name == "three_rf":
        return Case(name, three_blocks(), nvar=2, nfringe=2, reduce=True)
    if name == "moving_rf":
        blocks = W.tiny()
        return Case(name, blocks, nvar=2, moves=_moves(blocks, 2, seed=21), reduce=True)
    # overset-boundary bands (mexclude) thicker than the nfringe + 1 sweeps,
    # so the reduction drops receptors and cancels their donors
    if name == "nested_rf":
        return Case(name, nested_boxes(), nvar=2, nfringe=0, mexclude=5, reduce=True)
    if name == "nested3_rf":
        return Case(name, nested_boxes(shell=True), nvar=3, nfringe=0, mexclude=4, reduce=True)
    # registerGridData with global ids (tioga.h:121-123): the reference built
    # with TIOGA_HAS_NODEGID dedupes queries by node id (search.C:48-74) and
    # getReceptorInfo packs the donor cells' 64-bit ids (tioga.C:572-671)
    if name == "tiny_gid":  # pole copies share node ids
        return Case(name, _position_gids(W.tiny()), nvar=3)
    if name == "seam_gid":  # a box split in two meshes: the seam nodes of both share ids
        blocks = [W.latlong_shell(12, 8, 5, 0.2, 0.45, 1.1, tag=1, center=(0.02, 0.05, 0.01))] + split_box(12, 6)
        return Case(name, _position_gids(blocks), nvar=2)
    if name == "seam_gid_rf":
        blocks = [W.latlong_shell(12, 8, 5, 0.2, 0.45, 1.1, tag=1, center=(0.02, 0.05, 0.01))] + split_box(12, 6)
        return Case(name, _position_gids(blocks), nvar=2, nfringe=2, reduce=True)
    # the point path (performConnectivityHighOrder without callbacks, then
    # dataUpdate at points): receptor nodes searched again as points
    if name == "points":
        return Case(name, W.tiny(), nvar=3, points=1)
    if name == "points_three":
        return Case(name, three_blocks(), nvar=2, interptype=1, points=1)
    if name == "points_cb":  # high-order callbacks
        return Case(name, W.tiny(), nvar=3, points=2)
    if name == "points_cb_jitter":  # callbacks; overlapping donor-cell boxes in a jittered box
        return Case(name, [W.tiny()[0], W.cart_box(10, tag=2, jitter=0.15, seed=8)], nvar=2, nfringe=2, points=2)
    # several MPI ranks (the reference under mpiexec; the engine's virtual ranks)
    if name == "ranks2":
        blocks, ranks = ranked_blocks()
        return Case(name, blocks, nvar=3, ranks=ranks)
    if name == "ranks3":
        blocks, ranks = ranked_blocks(three=True)
        return Case(name, blocks, nvar=2, ranks=ranks)
    if name == "ranks3_rf":
        blocks, ranks = ranked_blocks(three=True)
        return Case(name, blocks, nvar=2, nfringe=2, reduce=True, ranks=ranks)
    if name == "ranks3_gid":
        blocks, ranks = ranked_blocks(three=True)
        return Case(name, _position_gids(blocks), nvar=2, ranks=ranks)
    if name == "ranks4_apart":  # a fourth rank whose block overlaps no other (exchangeDonors.C:36)
        blocks, ranks = ranked_blocks(three=True)
        far = W.cart_box(4, tag=4, lo=1.3, hi=1.6)
        return Case(name, blocks + [far], nvar=2, nfringe=2, reduce=True, ranks=ranks + [3])
    raise KeyError(name)


SMALL_CASES = ["tiny", "tiny_isym0", "tiny_isym1", "tiny_isym2", "mixed", "userres", "moving", "three", "bigtag"]
REDUCE_CASES = ["tiny_rf", "isym0_rf", "mixed_rf", "three_rf", "moving_rf", "nested_rf", "nested3_rf"]
GID_CASES = ["tiny_gid", "seam_gid", "seam_gid_rf"]
ALL_CASES = SMALL_CASES + ["hex48_pair"] + REDUCE_CASES + GID_CASES
# the point path: pinned by the reference directly (the oracle restates the
# node path only)
POINT_CASES = ["points", "points_three", "points_cb", "points_cb_jitter"]
# several ranks: pinned by the reference run under mpiexec (the oracle
# restates the single-rank path only)
RANK_CASES = ["ranks2", "ranks3", "ranks3_rf", "ranks3_gid", "ranks4_apart"]


# ---------------------------------------------------------------------------
# AMR Cartesian background, gather direction: near-body blocks receive from
# Cartesian patches (tioga.C:180-334 on one rank, pdegree 0)
# ---------------------------------------------------------------------------

@dataclass
class AmrCase:
    name: str
    blocks: list
    grid: object           # tioga_amd.workloads.AmrGrid
    nvar: int
    interptype: int = 0
    # both directions (performConnectivity for the hole maps, then the
    # reference's performConnectivityAMR sequence: one near-body block)
    full: bool = False

    @property
    def qs(self) -> List[np.ndarray]:
        return [W.fields(b.nnodes, self.nvar, seed=31 + i) for i, b in enumerate(self.blocks)]

    @property
    def patch_q(self) -> List[np.ndarray]:
        return self.grid.patch_fields(self.nvar, seed=41)


def build_amr(name: str) -> AmrCase:
    if name == "amr_fringe":
        # receptors: the outer fringe of a lat-long shell (default resolutions);
        # three levels around the body, grid list and local ids permuted
        grid = W.amr_hierarchy((-1.0, -1.0, -1.0), 0.25,
                               [((0, 0, 0), (8, 8, 8)), ((2, 2, 2), (14, 14, 14)), ((8, 8, 8), (24, 24, 24))],
                               patch=4, nf=1, shuffle_seed=1, local_seed=2)
        return AmrCase(name, [W.latlong_shell(16, 8, 6, 0.25, 0.6, 1.1, tag=1)], grid, nvar=5)
    if name == "amr_all":
        # every node of a mixed-element shell a receptor, plus the fringe of a
        # second body; two ghost layers, column layout; the patches stop at
        # z = 0.4, so receptors above find no donor
        grid = W.amr_hierarchy((-1.0, -1.0, -1.0), 0.2,
                               [((0, 0, 0), (10, 10, 7)), ((4, 4, 4), (16, 16, 14)),
                                ((12, 12, 12), (28, 28, 28))],
                               patch=4, nf=2, shuffle_seed=3, local_seed=4)
        a = W.all_receptors(W.mixed_shell(12, 8, layers=(2, 2, 1, 2), tag=1))
        b = W.latlong_shell(12, 6, 4, 0.1, 0.3, 1.05, tag=2, center=(0.3, -0.35, 0.2))
        return AmrCase(name, [a, b], grid, nvar=3, interptype=1)
    if name == "amr_faces":
        # receptors exactly on cell faces, patch faces and the refined region's
        # boundary: ties between patches, clamped stencils, upper-face points
        grid = W.amr_hierarchy((-1.0, -1.0, -1.0), 0.25, [((0, 0, 0), (8, 8, 8)), ((4, 4, 4), (12, 12, 12))],
                               patch=4, nf=1, shuffle_seed=5, local_seed=6)
        return AmrCase(name, [W.all_receptors(W.cart_box(8, tag=1, lo=-0.5, hi=0.5))], grid, nvar=2)
    # both directions: patch cells inside the near-body block take values from it
    if name == "amr_cart":
        # the shell of amr_fringe: its wall cuts holes in the patches
        grid = W.amr_hierarchy((-1.0, -1.0, -1.0), 0.25,
                               [((0, 0, 0), (8, 8, 8)), ((2, 2, 2), (14, 14, 14)), ((8, 8, 8), (24, 24, 24))],
                               patch=4, nf=1, shuffle_seed=1, local_seed=2)
        return AmrCase(name, [W.latlong_shell(16, 8, 6, 0.25, 0.6, 1.1, tag=1)], grid, nvar=5, full=True)
    if name == "amr_cart_col":
        # a mixed-element shell off the origin, two ghost layers, column layout
        grid = W.amr_hierarchy((-1.0, -1.0, -1.0), 0.2,
                               [((0, 0, 0), (10, 10, 10)), ((4, 4, 4), (16, 16, 16)),
                                ((12, 12, 12), (28, 28, 28))],
                               patch=4, nf=2, shuffle_seed=3, local_seed=4)
        return AmrCase(name, [W.mixed_shell(12, 8, layers=(2, 2, 1, 2), tag=1)], grid, nvar=3, interptype=1,
                       full=True)
    raise KeyError(name)


AMR_CASES = ["amr_fringe", "amr_all", "amr_faces"]
# both directions: pinned by the reference directly (the oracle restates the
# gather direction only)
AMR_FULL_CASES = ["amr_cart", "amr_cart_col"]
