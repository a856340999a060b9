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
    if name == "config1":
        return Case(name, W.config1(), nvar=5)
    raise KeyError(name)


SMALL_CASES = ["tiny", "tiny_isym0", "tiny_isym1", "tiny_isym2", "mixed", "userres", "moving", "three"]
ALL_CASES = SMALL_CASES + ["config1"]
