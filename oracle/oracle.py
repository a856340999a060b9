"""TEST INFRASTRUCTURE ONLY -- Python handle on the oracle and the compiled
reference.

* :class:`Oracle` drives ``oracle/liboracle.so`` (the C restatement of the
  reference's single-rank connectivity path, tioga_oracle.c) through ctypes.
* :func:`write_case` / :func:`read_result` speak the file format of
  ``oracle/ref_driver.cpp``, which runs the reference compiled from its own
  sources into ``oracle/_ref``.

Only tests/, the smoke check and bench.py's cpu_baseline leg import this
module; the engine in tioga_amd never does.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import struct
import subprocess
from typing import Dict, List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")
REF_DRIVER_GID = os.path.join(HERE, "_ref", "gid", "ref_driver")  # built with TIOGA_HAS_NODEGID

_P = C.c_void_p
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = _P
        L.orc_destroy.argtypes = [_P]
        L.orc_register_grid.argtypes = [_P, C.c_int, C.c_int, _P, _P, C.c_int, C.c_int, _P, _P,
                                        C.c_int, _P, _P, _P]
        L.orc_set_resolutions.argtypes = [_P, C.c_int, _P, _P]
        L.orc_set_cell_iblank.argtypes = [_P, C.c_int, _P]
        L.orc_set_params.argtypes = [_P, C.c_int, C.c_int, C.c_int]
        L.orc_profile.argtypes = [_P]
        L.orc_perform_connectivity.argtypes = [_P]
        L.orc_reduce_fringes.argtypes = [_P]
        L.orc_register_solution.argtypes = [_P, C.c_int, _P]
        L.orc_data_update.argtypes = [_P, C.c_int, C.c_int]
        L.orc_get_donor_count.argtypes = [_P, C.c_int, _P, _P]
        L.orc_get_donor_info.argtypes = [_P, C.c_int, _P, _P, _P, _P]
        L.orc_block_array.argtypes = [_P, C.c_int, C.c_char_p, _P]
        L.orc_block_array.restype = C.c_long
        L.orc_hole_map.argtypes = [_P, C.c_int, _P, _P, _P]
        L.orc_hole_map.restype = C.c_long
        L.orc_search_points.argtypes = [_P, C.c_int, _P, C.c_int, _P]
        L.orc_search_points.restype = C.c_double
        L.orc_search_points_obb.argtypes = [_P, C.c_int, _P, C.c_int, _P, C.c_int, _P, C.c_int, _P]
        L.orc_search_points_obb.restype = C.c_double
        L.orc_set_gids.argtypes = [_P, C.c_int, _P, _P]
        L.orc_unique_nodes_gid.argtypes = [_P, _P, _P, C.c_int]
        L.orc_get_receptor_info.argtypes = [_P, _P]
        L.orc_get_receptor_info.restype = C.c_long
        L.orc_nodal_weights.argtypes = [_P, _P, _P, C.c_int]
        L.orc_cell_volume.argtypes = [_P, C.c_int]
        L.orc_cell_volume.restype = C.c_double
        L.orc_median.argtypes = [_P, _P, C.c_int]
        L.orc_find_obb.argtypes = [_P, _P, _P, _P, C.c_int]
        L.orc_obb_intersect.argtypes = [_P, _P, _P, _P, _P, _P]
        L.orc_obb_intersect.restype = C.c_int
        L.orc_fill_hole_map.argtypes = [_P, _P, C.c_int]
        L.orc_check_hole_map.argtypes = [_P, _P, _P, _P]
        L.orc_check_hole_map.restype = C.c_int
        L.orc_unique_nodes_octree.argtypes = [_P, _P, _P, _P, C.c_int]
        L.orc_build_adt.argtypes = [C.c_int, _P, _P, _P, _P]
        L.orc_weight_aborts.restype = C.c_int
        L.orc_amr_global_data.argtypes = [_P, C.c_int, C.c_int, _P, _P, _P, C.c_int, C.c_int]
        L.orc_amr_patch_count.argtypes = [_P, C.c_int]
        L.orc_amr_local_data.argtypes = [_P, C.c_int, C.c_int, _P, _P]
        L.orc_amr_search.argtypes = [_P, _P, C.c_int, _P]
        L.orc_perform_connectivity_amr.argtypes = [_P]
        L.orc_perform_connectivity_amr.restype = C.c_int
        L.orc_data_update_amr.argtypes = [_P, C.c_int, C.c_int]
        L.orc_amr_array.argtypes = [_P, C.c_char_p, _P]
        L.orc_amr_array.restype = C.c_long
        _lib = L
    return _lib


def ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_P)


_I32_NAMES = {"isearch", "tagsearch", "donorId", "xtag", "elementList", "adtIntegers", "iblank_cell",
              "receptorIdCart", "donorIdCart"}


class Oracle:
    """Single-rank TIOGA restated in C; mirrors the reference's call order
    (tioga.h:121-258): register, profile, performConnectivity,
    registerSolution, dataUpdate."""

    def __init__(self, blocks, isym: int = 3, nfringe: int = 1, mexclude: int = 3):
        self.L = lib()
        self.h = self.L.orc_create()
        self.blocks = [b.normalized() for b in blocks]
        self.iblank = []
        self.iblank_cell = []
        self._keep = []
        self.L.orc_set_params(self.h, isym, nfringe, mexclude)
        for b in self.blocks:
            ib = np.ones(b.nnodes, np.int32)
            ibc = np.ones(max(b.ncells, 1), np.int32)
            nv = np.array([t[0] for t in b.types], np.int32)
            nc = np.array([t[1].shape[0] for t in b.types], np.int32)
            conns = [t[1] for t in b.types]
            cptr = (C.c_void_p * len(conns))(*[c.ctypes.data for c in conns])
            self._keep += [nv, nc, conns, cptr]
            self.L.orc_register_grid(self.h, b.tag, b.nnodes, ptr(b.x), ptr(ib), len(b.wbc), len(b.obc),
                                     ptr(b.wbc), ptr(b.obc), len(b.types), ptr(nv), ptr(nc),
                                     C.cast(cptr, _P))
            if b.nres is not None:
                self.L.orc_set_resolutions(self.h, b.tag, ptr(b.nres), ptr(b.cres))
            self.L.orc_set_cell_iblank(self.h, b.tag, ptr(ibc))
            if b.cell_gid is not None or b.node_gid is not None:
                self.L.orc_set_gids(self.h, b.tag, ptr(b.cell_gid), ptr(b.node_gid))
            self.iblank.append(ib)
            self.iblank_cell.append(ibc)

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_coords(self, bi: int, x: np.ndarray) -> None:
        """Move a block in place (the reference keeps the user's pointer)."""
        self.blocks[bi].x[...] = x

    def profile(self):
        self.L.orc_profile(self.h)

    def connect(self):
        self.L.orc_perform_connectivity(self.h)

    def reduce_fringes(self):
        self.L.orc_reduce_fringes(self.h)

    def data_update(self, qs: Sequence[np.ndarray], nvar: int, interptype: int = 0) -> List[np.ndarray]:
        out = [np.ascontiguousarray(q, dtype=np.float64).copy() for q in qs]
        for b, q in zip(self.blocks, out):
            self.L.orc_register_solution(self.h, b.tag, ptr(q))
        self.L.orc_data_update(self.h, nvar, interptype)
        return out

    def donor_info(self, btag: int):
        d = np.zeros(1, np.int32)
        f = np.zeros(1, np.int32)
        self.L.orc_get_donor_count(self.h, btag, ptr(d), ptr(f))
        dc, fc = int(d[0]), int(f[0])
        rec = np.zeros(4 * dc + 1, np.int32)
        ind = np.zeros(fc + 1, np.int32)
        frac = np.zeros(fc + 1, np.float64)
        if dc > 0:
            self.L.orc_get_donor_info(self.h, btag, ptr(rec), ptr(ind), ptr(frac), ptr(d))
        return rec[:4 * dc], ind[:fc], frac[:fc]

    def array(self, btag: int, name: str) -> np.ndarray:
        n = self.L.orc_block_array(self.h, btag, name.encode(), None)
        dt = np.int32 if name in _I32_NAMES else np.float64
        out = np.zeros(max(n, 0), dt)
        if n > 0:
            self.L.orc_block_array(self.h, btag, name.encode(), ptr(out))
        return out

    def search_points_obb(self, btag: int, xobb: np.ndarray, xq: np.ndarray, nthreads: int = 1):
        """Donors of xq against the tree MeshBlock::search builds for the
        query set xobb (no dedupe): (donor[nq], tree size, seconds)."""
        xobb = np.ascontiguousarray(xobb, np.float64).reshape(-1, 3)
        xq = np.ascontiguousarray(xq, np.float64).reshape(-1, 3)
        donor = np.zeros(xq.shape[0], np.int32)
        ne = C.c_int(0)
        t = self.L.orc_search_points_obb(self.h, btag, ptr(xobb), xobb.shape[0], ptr(xq), xq.shape[0], ptr(donor),
                                         int(nthreads), C.byref(ne))
        return donor, int(ne.value), float(t)

    def receptor_info(self) -> np.ndarray:
        n = self.L.orc_get_receptor_info(self.h, None)
        out = np.zeros(max(n, 0), np.int32)
        if n > 0:
            self.L.orc_get_receptor_info(self.h, ptr(out))
        return out

    def hole_map(self, mtag: int):
        nx = np.zeros(3, np.int32)
        ext = np.zeros(6)
        n = self.L.orc_hole_map(self.h, mtag, ptr(nx), ptr(ext), None)
        if n <= 0:
            return None
        sam = np.zeros(n, np.int32)
        self.L.orc_hole_map(self.h, mtag, ptr(nx), ptr(ext), ptr(sam))
        return nx, ext, sam

    # ---- AMR Cartesian background, gather direction ----
    def register_amr(self, grid, patch_q: Sequence[np.ndarray]) -> None:
        idata = np.ascontiguousarray(grid.idata, np.int32)
        rdata = np.ascontiguousarray(grid.rdata, np.float64)
        qnode = np.ascontiguousarray(grid.qnode, np.float64)
        self._keep += [idata, rdata, qnode]
        self.L.orc_amr_global_data(self.h, grid.nf, grid.qstride, ptr(qnode), ptr(idata), ptr(rdata), grid.ngrids,
                                   qnode.size)
        self.L.orc_amr_patch_count(self.h, len(patch_q))
        order = grid.patch_order()
        for lp, q in enumerate(patch_q):
            g = int(order[lp])
            ibl = np.ones(grid.d3nf(g), np.int32)
            q = np.ascontiguousarray(q, np.float64)
            self._keep += [ibl, q]
            self.L.orc_amr_local_data(self.h, lp, g, ptr(ibl), ptr(q))

    def connect_amr(self) -> int:
        return int(self.L.orc_perform_connectivity_amr(self.h))

    def amr_search(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        out = np.zeros(x.shape[0], np.int32)
        self.L.orc_amr_search(self.h, ptr(x), x.shape[0], ptr(out))
        return out

    def amr_array(self, name: str) -> np.ndarray:
        n = self.L.orc_amr_array(self.h, name.encode(), None)
        out = np.zeros(max(n, 0), np.float64 if name == "rec_w" else np.int32)
        if n > 0:
            self.L.orc_amr_array(self.h, name.encode(), ptr(out))
        return out

    def data_update_amr(self, qs: Sequence[np.ndarray], nvar: int, interptype: int = 0) -> List[np.ndarray]:
        out = [np.ascontiguousarray(q, dtype=np.float64).copy() for q in qs]
        for b, q in zip(self.blocks, out):
            self.L.orc_register_solution(self.h, b.tag, ptr(q))
        self.L.orc_data_update_amr(self.h, nvar, interptype)
        return out

    def amr_results(self, qs=None, nvar: int = 0, interptype: int = 0) -> Dict[str, np.ndarray]:
        """The arrays the reference driver dumps for the AMR gather path."""
        r: Dict[str, np.ndarray] = {}
        for i, b in enumerate(self.blocks):
            p = f"b{i}."
            r[p + "iblank"] = self.iblank[i].copy()
            r[p + "iblank_cell"] = self.iblank_cell[i][:b.ncells].copy()
            r[p + "nodeRes"] = self.array(b.tag, "nodeRes")
            for nm in ("receptorIdCart", "donorIdCart", "rxyzCart"):
                r[p + nm] = self.array(b.tag, nm)
            r[p + "ncart"] = np.array([r[p + "receptorIdCart"].size], np.int32)
        for nm in ("rec_patch", "rec_info", "rec_ijk", "rec_w"):
            r["amr." + nm] = self.amr_array(nm)
        if qs is not None:
            for i, q in enumerate(self.data_update_amr(qs, nvar, interptype)):
                r[f"b{i}.q"] = q
        return r

    def results(self, qs: Optional[Sequence[np.ndarray]] = None, nvar: int = 0,
                interptype: int = 0) -> Dict[str, np.ndarray]:
        """The arrays the reference driver dumps, under the same names."""
        r: Dict[str, np.ndarray] = {}
        for i, b in enumerate(self.blocks):
            p = f"b{i}."
            r[p + "iblank"] = self.iblank[i].copy()
            r[p + "iblank_cell"] = self.iblank_cell[i][:b.ncells].copy()
            for nm in ("nodeRes", "cellRes", "obb", "xsearch", "res_search", "isearch", "tagsearch",
                       "donorId", "xtag", "elementList", "adtIntegers", "adtReals", "adtExtents"):
                a = self.array(b.tag, nm)
                if a.size or nm != "adtExtents":
                    r[p + nm] = a
            r[p + "nsearch"] = np.array([r[p + "donorId"].size], np.int32)
            rec, ind, frac = self.donor_info(b.tag)
            r[p + "dcount"] = np.array([rec.size // 4], np.int32)
            r[p + "receptors"], r[p + "indices"], r[p + "frac"] = rec, ind, frac
        mt = max(b.tag for b in self.blocks)
        for m in range(1, mt + 1):
            hm = self.hole_map(m)
            r[f"hm{m}.existWall"] = np.array([0 if hm is None else 1], np.int32)
            if hm is not None:
                r[f"hm{m}.nx"], r[f"hm{m}.extents"], r[f"hm{m}.sam"] = hm
        if any(b.cell_gid is not None for b in self.blocks):
            r["receptor_info"] = self.receptor_info()
        if qs is not None:
            for i, q in enumerate(self.data_update(qs, nvar, interptype)):
                r[f"b{i}.q"] = q
        return r


# --------------------------------------------------------------------------
# case / result files of oracle/ref_driver.cpp
# --------------------------------------------------------------------------

def write_case(path: str, blocks, qs: Sequence[np.ndarray], nvar: int, interptype: int = 0,
               isym: int = 3, nfringe: int = 1, mexclude: int = 3,
               moves: Sequence[Sequence[Optional[np.ndarray]]] = (), amr=None, reduce: bool = False,
               ranks: Optional[Sequence[int]] = None, points: int = 0, amr_full: bool = False) -> None:
    """amr = (AmrGrid, patch fields per local patch) selects the AMR gather
    mode of the driver (format version 2); reduce runs reduce_fringes after
    each step's connectivity (format version 3); ranks (the MPI rank of each
    block) selects the multi-rank mode (format version 5); points runs the
    point path after each step (format version 6); amr_full runs both
    directions of the AMR path (format version 7)."""
    gids = any(b.cell_gid is not None or b.node_gid is not None for b in blocks)
    version = 7 if amr_full else (6 if points else (5 if ranks is not None else (4 if gids else (3 if reduce else (1 if amr is None else 2)))))
    with open(path, "wb") as f:
        w = lambda fmt, *v: f.write(struct.pack("<" + fmt, *v))
        w("9i", 0x41474954, version, len(blocks), isym, nfringe, mexclude, nvar, interptype,
          len(moves))
        for b, q in zip(blocks, qs):
            b = b.normalized()
            has = b.nres is not None
            w("6i", b.tag, b.nnodes, len(b.wbc), len(b.obc), len(b.types), int(has))
            f.write(b.x.tobytes())
            f.write(b.wbc.tobytes())
            f.write(b.obc.tobytes())
            for nv, c in b.types:
                w("2i", nv, c.shape[0])
                f.write(c.tobytes())
            if has:
                f.write(b.nres.tobytes())
                f.write(b.cres.tobytes())
            f.write(np.ascontiguousarray(q, np.float64).tobytes())
        if version >= 2 and amr is None:
            w("i", 0)
        if amr is not None:
            grid, pq = amr
            w("i", 1)
            w("5i", grid.nf, grid.qstride, grid.ngrids, grid.qnode.size, len(pq))
            f.write(np.ascontiguousarray(grid.qnode, np.float64).tobytes())
            f.write(np.ascontiguousarray(grid.idata, np.int32).tobytes())
            f.write(np.ascontiguousarray(grid.rdata, np.float64).tobytes())
            order = grid.patch_order()
            for lp, q in enumerate(pq):
                w("i", int(order[lp]))
                f.write(np.ascontiguousarray(q, np.float64).tobytes())
        if version >= 3:
            w("i", int(reduce))
        if version >= 4:
            w("i", int(gids))
            for b in blocks if gids else ():
                b = b.normalized()
                f.write(b.cell_gid.tobytes())
                f.write(b.node_gid.tobytes())
        if version >= 5:
            w("i", int(ranks is not None))
            if ranks is not None:
                f.write(np.asarray(ranks, np.int32).tobytes())
        if version >= 6:
            w("i", int(points))  # 1: the point path, 2: with the test callbacks
        if version >= 7:
            w("i", int(amr_full))
        for step in moves:
            for xb in step:
                if xb is None:
                    w("i", 0)
                else:
                    w("i", 1)
                    f.write(np.ascontiguousarray(xb, np.float64).tobytes())


def read_result(path: str) -> Dict[str, np.ndarray]:
    out: Dict[str, np.ndarray] = {}
    with open(path, "rb") as f:
        data = f.read()
    o = 0
    while o < len(data):
        (l,) = struct.unpack_from("<i", data, o)
        o += 4
        name = data[o:o + l].decode()
        o += l
        dt, n = struct.unpack_from("<iq", data, o)
        o += 12
        dtype = np.int32 if dt == 0 else np.float64
        nb = n * (4 if dt == 0 else 8)
        out[name] = np.frombuffer(data[o:o + nb], dtype=dtype).copy()
        o += nb
    return out


MPIEXEC = "/opt/conda/bin/mpiexec"


def run_reference(case_path: str, out_path: str, timeout: float = 600.0,
                  cwd: Optional[str] = None, nranks: int = 0, gids: Optional[bool] = None):
    """Run the compiled reference (oracle/_ref/ref_driver) on one rank
    (in cwd: the reference writes a cartGrid.dat there for AMR cases).

    nranks > 0 (a version-5 case) runs it under mpiexec on that many ranks
    and returns (results, packets): every rank's arrays merged, each name
    carrying its rank as "r<k>." after the step and reduce_fringes prefixes,
    and packets[k] = the sendRecvPackets calls of rank k (parse_packets).
    gids selects the global-id build (default: a version 4 case)."""
    env = dict(os.environ)
    env.setdefault("HYDRA_LAUNCHER", "fork")
    with open(case_path, "rb") as f:
        version = struct.unpack("<2i", f.read(8))[1]
    if gids is None:
        gids = version == 4  # version 5 and 6 cases name it
    drv = REF_DRIVER_GID if gids else REF_DRIVER
    if nranks <= 0:
        subprocess.run([drv, case_path, out_path], check=True, timeout=timeout, env=env, cwd=cwd,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        return read_result(out_path)
    subprocess.run([MPIEXEC, "-n", str(nranks), drv, case_path, out_path], check=True, timeout=timeout,
                   env=env, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    res: Dict[str, np.ndarray] = {}
    packets = []
    for k in range(nranks):
        for name, a in read_result(f"{out_path}.r{k}").items():
            res[re.sub(r"^((?:s\d+\.)?(?:rf\.)?)", rf"\1r{k}.", name)] = a
        packets.append(parse_packets(f"{out_path}.r{k}.pk"))
    return res, packets


def parse_packets(path: str):
    """The packet log of one rank (ref_driver.cpp) as a list with one entry
    per parallelComm::sendRecvPackets call, each {dest rank: (int32 payload,
    float64 payload, nints, nreals)}: the counts come from the call's first
    group of sends, the payloads from its second, in destination order."""
    with open(path, "rb") as f:
        data = f.read()
    groups, cur, o = [], [], 0
    while o < len(data):
        (kind,) = struct.unpack_from("<i", data, o)
        o += 4
        if kind == 2:
            groups.append(cur)
            cur = []
            continue
        dest, tag, dt, n = struct.unpack_from("<4i", data, o)
        o += 16
        nb = n * (8 if dt == 1 else 4)
        cur.append((dest, tag, np.frombuffer(data[o:o + nb], np.float64 if dt == 1 else np.int32).copy()))
        o += nb
    if cur or len(groups) % 2:
        raise ValueError(f"{path}: unpaired send groups")
    calls = []
    for g in range(0, len(groups), 2):
        counts, payload = groups[g], iter(groups[g + 1])
        call = {}
        for dest, _tag, c in counts:
            ni, nr = int(c[0]), int(c[1])
            ints = next(payload)[2] if ni > 0 else np.zeros(0, np.int32)
            reals = next(payload)[2] if nr > 0 else np.zeros(0, np.float64)
            call[int(dest)] = (ints, reals, ni, nr)
        calls.append(call)
    return calls
