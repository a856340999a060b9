"""Python handle on libtioga_amd.so through its C ABI (include/tioga_amd.h).

:class:`Tioga` mirrors the reference's tiogaInterface call sequence
(tiogaInterface.C:36-328): init, registergrid_data_mb, set*, preprocess_grids,
performconnectivity, registersolution, dataupdate_mb, getdonorcount/info,
delete.  Arrays are numpy arrays whose memory the library keeps by pointer,
as the reference keeps the solver's arrays, so the handle holds references to
them.  Every call checks the library's error status and raises on failure;
there is no fallback path: without the built library or a GPU, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtioga_amd.so")

_P = C.c_void_p
_lib = None


class TiogaError(RuntimeError):
    pass


def load():
    """Load the in-tree library (build it with tioga_amd.build first)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TiogaError(f"{LIB_PATH} is missing: run `python -m tioga_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.tioga_init_.argtypes = [C.c_int]
        for name in ("tioga_preprocess_grids_", "tioga_performconnectivity_", "tioga_delete_",
                     "tioga_reduce_fringes_"):
            getattr(L, name).argtypes = []
        L.tioga_registersolution_.argtypes = [_P, _P]
        L.tioga_dataupdate_mb_.argtypes = [_P, C.c_char_p]
        L.tioga_getdonorcount_.argtypes = [_P, _P, _P]
        L.tioga_getdonorinfo_.argtypes = [_P, _P, _P, _P, _P]
        L.tioga_setsymmetry_.argtypes = [_P]
        L.tioga_setnfringe_.argtypes = [_P]
        L.tioga_setmexclude_.argtypes = [_P]
        L.tioga_setresolutions_multi_.argtypes = [_P, _P, _P]
        L.tioga_setcelliblank_.argtypes = [_P]
        L.tioga_amd_status.argtypes = [C.c_char_p, C.c_int]
        L.tioga_amd_status.restype = C.c_int
        L.tioga_amd_clear_status.argtypes = []
        L.tioga_amd_block_array.argtypes = [C.c_int, C.c_char_p, _P]
        L.tioga_amd_block_array.restype = C.c_long
        L.tioga_amd_hole_map.argtypes = [C.c_int, _P, _P, _P]
        L.tioga_amd_hole_map.restype = C.c_long
        L.tioga_amd_time_search.argtypes = [C.c_int, _P]
        L.tioga_amd_time_search.restype = C.c_double
        L.tioga_amd_time_dataupdate.argtypes = [C.c_int, C.c_int, C.c_int, _P]
        L.tioga_amd_time_dataupdate.restype = C.c_double
        L.tioga_amd_phase_times.argtypes = [_P]
        L.tioga_amd_query_set_create.argtypes = [C.c_int, _P, C.c_long]
        L.tioga_amd_query_set_create.restype = C.c_int
        L.tioga_amd_query_set_search.argtypes = [C.c_int, _P]
        L.tioga_amd_time_query_sets.argtypes = [C.c_int, _P]
        L.tioga_amd_time_query_sets.restype = C.c_double
        L.tioga_amd_query_set_clear.argtypes = []
        L.tioga_amd_last_search_kernel_ms.argtypes = []
        L.tioga_amd_last_search_kernel_ms.restype = C.c_double
        L.tioga_amd_search_stats.argtypes = [_P]
        L.tioga_amd_search_bins.argtypes = [C.c_int]
        L.tioga_amd_search_bins.restype = C.c_int
        L.tioga_amd_search_stats.restype = C.c_int
        L.tioga_amd_query_set_nelem.argtypes = [C.c_int]
        L.tioga_amd_query_set_nelem.restype = C.c_int
        L.tioga_amd_register_gids.argtypes = [C.c_int, _P, _P]
        L.tioga_amd_receptor_info.argtypes = [_P]
        L.tioga_amd_receptor_info.restype = C.c_long
        L.tioga_amd_query_set_donors.argtypes = [C.c_int, _P]
        L.tioga_amd_query_set_extents.argtypes = [C.c_int, _P]
        L.tioga_amd_query_set_extents.restype = C.c_int
        L.tioga_amd_preprocess_block.argtypes = [C.c_int]
        L.tioga_amd_search_tuning.argtypes = [_P, _P]
        L.tioga_amd_search_tuning.restype = C.c_int
        L.tioga_amd_search_capacity.argtypes = [C.c_long]
        L.tioga_amd_search_capacity.restype = C.c_long
        L.tioga_amd_weight_aborts.restype = C.c_int
        L.tioga_register_amr_global_data_.argtypes = [_P] * 7
        L.tioga_register_amr_patch_count_.argtypes = [_P]
        L.tioga_register_amr_local_data_.argtypes = [_P] * 4
        L.tioga_performconnectivity_amr_.argtypes = []
        L.tioga_amd_time_dataupdate_amr.argtypes = [C.c_int, C.c_int, C.c_int, _P]
        L.tioga_amd_time_dataupdate_amr.restype = C.c_double
        L.tioga_amd_set_batching.argtypes = [C.c_int]
        L.tioga_amd_serve_open.argtypes = [C.c_int, C.c_int, C.c_double]
        L.tioga_amd_serve_open.restype = C.c_int
        L.tioga_amd_serve_submit.argtypes = [C.c_int, _P, _P, _P]
        L.tioga_amd_serve_submit.restype = C.c_long
        L.tioga_amd_serve_wait.argtypes = [C.c_long, _P, C.c_int, _P, _P, _P, C.c_int, _P, _P]
        L.tioga_amd_serve_wait.restype = C.c_int
        L.tioga_amd_serve_open2.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int]
        L.tioga_amd_serve_open2.restype = C.c_int
        L.tioga_amd_serve_wait2.argtypes = [C.c_long, _P, C.c_int, _P, _P, _P, C.c_int, _P, _P]
        L.tioga_amd_serve_wait2.restype = C.c_int
        L.tioga_amd_serve_close.argtypes = []
        L.tioga_amd_search_graph.argtypes = [C.c_int]
        L.tioga_amd_set_rank.argtypes = [C.c_int]
        L.tioga_amd_rank.restype = C.c_int
        L.tioga_amd_points_array.argtypes = [C.c_int, C.c_char_p, _P]
        L.tioga_amd_points_array.restype = C.c_long
        L.tioga_amd_record_packets.argtypes = [C.c_int]
        L.tioga_amd_packet.argtypes = [C.c_char_p, C.c_int, C.c_int, _P, _P, _P]
        L.tioga_amd_packet.restype = C.c_int
        L.tioga_amd_amr_array.argtypes = [C.c_int, C.c_char_p, _P]
        L.tioga_amd_amr_array.restype = C.c_long
        _lib = L
    return _lib


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_P)


def _ip(v: int):
    return C.byref(C.c_int(v))


_I32 = {"isearch", "tagsearch", "donorId", "xtag", "elementList", "adtIntegers", "iblank_cell"}
_AMR_F64 = {"rxyzCart", "rec_w", "interpCart_w"}
PHASES = ("holemap", "boxes", "searchdata", "search_prep", "search_kernel", "donors", "total", "nsearch",
          "prep_obb", "prep_filter", "prep_adt", "prep_dedupe", "obb_stepped")


class Tioga:
    """One overset system on the current GPU (the C ABI keeps one global
    instance, as the reference's `tg`)."""

    def __init__(self):
        self.L = load()
        self.L.tioga_amd_clear_status()
        self.L.tioga_init_(0)
        self._check()
        self._keep: List[object] = []
        self.blocks = []
        self.ranks: List[int] = []  # virtual rank of each registered block
        self.patch_q: List[np.ndarray] = []    # AMR patches' solutions and iblanks, as registered
        self.patch_ibl: List[np.ndarray] = []
        self.iblank: List[np.ndarray] = []
        self.iblank_cell: List[np.ndarray] = []
        self.q: Dict[int, np.ndarray] = {}

    def _check(self):
        buf = C.create_string_buffer(1024)
        if self.L.tioga_amd_status(buf, 1024):
            msg = buf.value.decode()
            self.L.tioga_amd_clear_status()
            raise TiogaError(msg)

    def close(self):
        if self.L is not None:
            self.L.tioga_delete_()
            self._check()
            self.L = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_rank(self, rank: int) -> None:
        """Virtual rank that the per-block calls which follow act on."""
        self.L.tioga_amd_set_rank(int(rank))
        self._check()

    def record_packets(self, on: bool = True) -> None:
        """Keep what the exchanges between virtual ranks send (tioga_amd_packet)."""
        self.L.tioga_amd_record_packets(1 if on else 0)
        self._check()

    def packet(self, exchange: str, src: int, dst: int):
        """(ints, reals, (nints, nreals)) that rank src sends rank dst in one
        exchange of the last connectivity, or None when they do not exchange."""
        cnt = np.zeros(4, np.int64)
        if self.L.tioga_amd_packet(exchange.encode(), src, dst, None, None, _p(cnt)) != 0:
            self._check()
            return None
        ints = np.zeros(int(cnt[2]), np.int32)
        reals = np.zeros(int(cnt[3]), np.float64)
        self.L.tioga_amd_packet(exchange.encode(), src, dst, _p(ints), _p(reals), _p(cnt))
        self._check()
        return ints, reals, (int(cnt[0]), int(cnt[1]))

    def packets(self, exchange: str, name: Optional[str] = None) -> Dict[str, np.ndarray]:
        """Every packet of one exchange as pk.<name>.<src>.<dst>.{ints,reals,counts}."""
        out: Dict[str, np.ndarray] = {}
        nr = max(self.ranks) + 1 if self.ranks else 1
        for src in range(nr):
            for dst in range(nr):
                got = self.packet(exchange, src, dst)
                if got is None:
                    continue
                p = f"pk.{name or exchange}.{src}.{dst}."
                out[p + "ints"], out[p + "reals"] = got[0], got[1]
                out[p + "counts"] = np.array(got[2], np.int32)
        return out

    # ---- registration (tiogaInterface.C:108-130) ----
    def register_grid(self, block, bid: Optional[int] = None, rank: int = 0) -> None:
        """Register a block on virtual rank `rank` (slot bid, by default the
        next free slot of that rank)."""
        self.set_rank(rank)
        b = block.normalized()
        ib = np.ones(b.nnodes, np.int32)
        args = []
        for nv, c in b.types:
            nvp, ncp = C.c_int(nv), C.c_int(c.shape[0])
            self._keep += [nvp, ncp, c]
            args += [C.byref(nvp), C.byref(ncp), _p(c)]
        bid = self.ranks.count(rank) + 1 if bid is None else bid
        self._keep += [b, ib]
        self.L.tioga_registergrid_data_mb_(_ip(bid), _ip(b.tag), _ip(b.nnodes), _p(b.x), _p(ib),
                                           _ip(len(b.wbc)), _ip(len(b.obc)), _p(b.wbc), _p(b.obc),
                                           _ip(len(b.types)), *args)
        self._check()
        if b.nres is not None:
            self.L.tioga_setresolutions_multi_(_ip(b.tag), _p(b.nres), _p(b.cres))
            self._check()
        if b.cell_gid is not None or b.node_gid is not None:
            self.L.tioga_amd_register_gids(b.tag, _p(b.cell_gid), _p(b.node_gid))
            self._check()
        self.blocks.append(b)
        self.ranks.append(rank)
        self.iblank.append(ib)

    def set_params(self, isym: int = 3, nfringe: int = 1, mexclude: int = 3):
        self.L.tioga_setsymmetry_(_ip(isym))
        self.L.tioga_setnfringe_(_ip(nfringe))
        self.L.tioga_setmexclude_(_ip(mexclude))
        self._check()

    def set_batching(self, on: bool) -> None:
        """Batched form (all blocks per stage) or block by block."""
        self.L.tioga_amd_set_batching(1 if on else 0)
        self._check()

    def set_coords(self, bi: int, x: np.ndarray) -> None:
        """Move block bi in place (the library keeps the user's pointer)."""
        self.blocks[bi].x[...] = x

    def preprocess_block(self, btag: int):
        """Preprocess after only block btag moved (tioga_amd_preprocess_block)."""
        self.L.tioga_amd_preprocess_block(btag)
        self._check()

    def preprocess(self):
        self.L.tioga_preprocess_grids_()
        self._check()

    def connect(self):
        self.L.tioga_performconnectivity_()
        self._check()

    def connect_points(self):
        """tioga_performconnectivity_highorder_: the receptor nodes searched
        again as points (no high-order callbacks)."""
        self.L.tioga_performconnectivity_highorder_()
        self._check()

    def data_update_points(self, nvar: int, interptype: int = 0):
        """dataUpdate(nvar, itype, at_points=1) (tioga_amd_dataupdate_points)."""
        self.L.tioga_amd_dataupdate_points(_ip(nvar), b"row" if interptype == 0 else b"column")
        self._check()

    def points_array(self, btag: int, name: str) -> np.ndarray:
        n = self.L.tioga_amd_points_array(btag, name.encode(), None)
        self._check()
        out = np.zeros(max(n, 0), np.float64 if name in ("rxyz", "interp2_weights") else np.int32)
        if n > 0:
            self.L.tioga_amd_points_array(btag, name.encode(), _p(out))
            self._check()
        return out

    def points_results(self, qs, nvar: int, interptype: int = 0) -> Dict[str, np.ndarray]:
        """The point path's arrays as the reference driver dumps them (hp.*):
        per block the picked points, the point search arrays and the point
        donor records, then q and iblank after a point update of fresh
        copies of qs."""
        r: Dict[str, np.ndarray] = {}
        for i, b in enumerate(self.blocks):
            self.set_rank(self.ranks[i])
            p = f"b{i}."
            for nm in ("ntotalPoints", "rxyz"):
                r[p + nm] = self.points_array(b.tag, nm)
            for nm in ("xsearch", "res_search", "isearch", "tagsearch", "donorId", "xtag"):
                a = self.array(b.tag, nm)
                if a.size:
                    r[p + nm] = a
            r[p + "nsearch"] = np.array([self.array(b.tag, "donorId").size], np.int32)
            for nm in ("ninterp2", "interp2_inode", "interp2_weights", "interp2_rinfo"):
                r[p + nm] = self.points_array(b.tag, nm)
        out = [np.ascontiguousarray(q, np.float64).copy() for q in qs]
        for b, rk, q in zip(self.blocks, self.ranks, out):
            self.set_rank(rk)
            self.register_solution(b.tag, q)
        self.data_update_points(nvar, interptype)
        for i, b in enumerate(self.blocks):
            self.set_rank(self.ranks[i])
            r[f"b{i}.q"] = out[i]
            r[f"b{i}.iblank"] = self.iblank[i].copy()
            r[f"b{i}.iblank_cell"] = self.array(b.tag, "iblank_cell")
        self.set_rank(0)
        return r

    def reduce_fringes(self):
        """tioga_reduce_fringes_ (tiogaInterface.C:302-305): thin the fringes
        of the last connectivity to nfringe + 1 layers; updates the
        registered iblank arrays."""
        self.L.tioga_reduce_fringes_()
        self._check()

    def register_solution(self, btag: int, q: np.ndarray):
        self.q[btag] = q
        self.L.tioga_registersolution_(_ip(btag), _p(q))
        self._check()

    def data_update(self, nvar: int, interptype: int = 0):
        self.L.tioga_dataupdate_mb_(_ip(nvar), b"row" if interptype == 0 else b"column")
        self._check()

    def data_update_copies(self, qs: Sequence[np.ndarray], nvar: int, interptype: int = 0):
        out = [np.ascontiguousarray(q, np.float64).copy() for q in qs]
        for b, r, q in zip(self.blocks, self.ranks, out):
            self.set_rank(r)
            self.register_solution(b.tag, q)
        self.data_update(nvar, interptype)
        return out

    def donor_info(self, btag: int):
        d = np.zeros(1, np.int32)
        f = np.zeros(1, np.int32)
        self.L.tioga_getdonorcount_(_ip(btag), _p(d), _p(f))
        self._check()
        dc, fc = int(d[0]), int(f[0])
        rec = np.zeros(4 * dc + 1, np.int32)
        ind = np.zeros(fc + 1, np.int32)
        frac = np.zeros(fc + 1, np.float64)
        if dc > 0:
            self.L.tioga_getdonorinfo_(_ip(btag), _p(rec), _p(ind), _p(frac), _p(d))
            self._check()
        return rec[:4 * dc], ind[:fc], frac[:fc]

    # ---- AMR Cartesian background (tiogaInterface.C:132-148, 165-168) ----
    def register_amr(self, grid, patch_q: Sequence[np.ndarray]) -> None:
        """Register the patch hierarchy; patch_q[local id] holds that patch's
        fields (kept by pointer, as the reference keeps them)."""
        idata = np.ascontiguousarray(grid.idata, np.int32)
        rdata = np.ascontiguousarray(grid.rdata, np.float64)
        qnode = np.ascontiguousarray(grid.qnode, np.float64)
        self._keep += [idata, rdata, qnode]
        self.L.tioga_register_amr_global_data_(_ip(grid.nf), _ip(grid.qstride), _p(qnode), _p(idata), _p(rdata),
                                               _ip(grid.ngrids), _ip(qnode.size))
        self._check()
        self.L.tioga_register_amr_patch_count_(_ip(len(patch_q)))
        self._check()
        order = grid.patch_order()
        for lp, q in enumerate(patch_q):
            g = int(order[lp])
            ibl = np.ones(grid.d3nf(g), np.int32)
            q = np.ascontiguousarray(q, np.float64)
            self._keep += [ibl, q]
            self.patch_q.append(q)
            self.patch_ibl.append(ibl)
            self.L.tioga_register_amr_local_data_(_ip(lp), _ip(g), _p(ibl), _p(q))
            self._check()

    def connect_amr(self):
        self.L.tioga_performconnectivity_amr_()
        self._check()

    def amr_array(self, btag: int, name: str) -> np.ndarray:
        n = self.L.tioga_amd_amr_array(btag, name.encode(), None)
        self._check()
        out = np.zeros(max(n, 0), np.float64 if name in _AMR_F64 else np.int32)
        if n > 0:
            self.L.tioga_amd_amr_array(btag, name.encode(), _p(out))
            self._check()
        return out

    def time_data_update_amr(self, nvar: int, interptype: int, nrep: int):
        n = C.c_longlong(0)
        ms = self.L.tioga_amd_time_dataupdate_amr(nvar, interptype, nrep, C.byref(n))
        self._check()
        return ms, n.value

    def amr_results(self, qs=None, nvar: int = 0, interptype: int = 0, full: bool = False) -> Dict[str, np.ndarray]:
        """The arrays the reference driver dumps for the AMR path; full adds
        the Cartesian-receptor direction: the near-body search of the patch
        cells, interpListCart, every patch's iblank and, after the update,
        every patch's q."""
        r: Dict[str, np.ndarray] = {}
        for i, b in enumerate(self.blocks):
            p = f"b{i}."
            r[p + "iblank"] = self.iblank[i].copy()
            r[p + "iblank_cell"] = self.array(b.tag, "iblank_cell")
            r[p + "nodeRes"] = self.array(b.tag, "nodeRes")
            for nm in ("receptorIdCart", "donorIdCart", "rxyzCart"):
                r[p + nm] = self.amr_array(b.tag, nm)
            r[p + "ncart"] = np.array([r[p + "receptorIdCart"].size], np.int32)
            if full:
                for nm in ("xsearch", "isearch", "donorId", "xtag"):
                    a = self.array(b.tag, nm)
                    if a.size:
                        r[p + nm] = a
                r[p + "nsearch"] = np.array([self.array(b.tag, "donorId").size], np.int32)
                for nm in ("interpCart_rinfo", "interpCart_inode", "interpCart_w"):
                    r[p + nm] = self.amr_array(b.tag, nm)
                r[p + "ninterpCart"] = np.array([r[p + "interpCart_rinfo"].size // 3], np.int32)
        for nm in ("rec_patch", "rec_info", "rec_ijk", "rec_w"):
            r["amr." + nm] = self.amr_array(0, nm)
        if full:
            for lp, ibl in enumerate(self.patch_ibl):
                r[f"amr.patch{lp}.iblank"] = ibl.copy()
        if qs is not None:
            for i, q in enumerate(self.data_update_copies(qs, nvar, interptype)):
                r[f"b{i}.q"] = q
            if full:
                for lp, q in enumerate(self.patch_q):
                    r[f"amr.patch{lp}.q"] = q.copy()
        return r

    # ---- serving: moving blocks over the resident background ----
    def serve_open(self, bg_tag: int, nslots: int = 4, budget_gb: float = 64.0, resident_tree: bool = False) -> None:
        self.L.tioga_amd_serve_open2(bg_tag, nslots, budget_gb, 1 if resident_tree else 0)
        self._check()

    def serve_submit(self, shell_tag: int, rot, center, shift) -> int:
        r = np.ascontiguousarray(rot, np.float64).reshape(9)
        c = np.ascontiguousarray(center, np.float64).reshape(3)
        t = np.ascontiguousarray(shift, np.float64).reshape(3)
        rid = self.L.tioga_amd_serve_submit(shell_tag, _p(r), _p(c), _p(t))
        self._check()
        return int(rid)

    def serve_wait(self, rid: int, nshell: int, cap: int = 1 << 22) -> Dict[str, object]:
        ib = np.zeros(max(nshell, 1), np.int32)
        nbg = np.zeros(1, np.int32)
        bn = np.zeros(cap, np.int32)
        bv = np.zeros(cap, np.int32)
        tm = np.zeros(4)
        st = np.zeros(4, np.int64)
        self.L.tioga_amd_serve_wait2(rid, _p(ib), nshell, _p(nbg), _p(bn), _p(bv), cap, _p(tm), _p(st))
        self._check()
        k = min(int(nbg[0]), cap)
        return {"shell_iblank": ib[:nshell], "bg_nodes": bn[:k].copy(), "bg_vals": bv[:k].copy(),
                "t_submit": tm[0], "t_start": tm[1], "t_done": tm[2], "connectivity_ms": tm[3],
                "nsearch": int(st[0]), "tieband": int(st[1]), "est_bytes": int(st[2]), "peak_bytes": int(st[3])}

    def serve_close(self) -> None:
        self.L.tioga_amd_serve_close()
        self._check()

    # ---- extensions ----
    def array(self, btag: int, name: str) -> np.ndarray:
        n = self.L.tioga_amd_block_array(btag, name.encode(), None)
        self._check()
        out = np.zeros(max(n, 0), np.int32 if name in _I32 else np.float64)
        if n > 0:
            self.L.tioga_amd_block_array(btag, name.encode(), _p(out))
            self._check()
        return out

    def hole_map(self, mtag: int):
        nx = np.zeros(3, np.int32)
        ext = np.zeros(6)
        n = self.L.tioga_amd_hole_map(mtag, _p(nx), _p(ext), None)
        self._check()
        if n <= 0:
            return None
        sam = np.zeros(n, np.int32)
        self.L.tioga_amd_hole_map(mtag, _p(nx), _p(ext), _p(sam))
        self._check()
        return nx, ext, sam

    def time_search(self, nrep: int):
        n = C.c_longlong(0)
        ms = self.L.tioga_amd_time_search(nrep, C.byref(n))
        self._check()
        return ms, n.value

    def time_data_update(self, nvar: int, interptype: int, nrep: int):
        n = C.c_longlong(0)
        ms = self.L.tioga_amd_time_dataupdate(nvar, interptype, nrep, C.byref(n))
        self._check()
        return ms, n.value

    def query_set(self, btag: int, xq: np.ndarray) -> int:
        xq = np.ascontiguousarray(xq, np.float64).reshape(-1, 3)
        i = self.L.tioga_amd_query_set_create(btag, _p(xq), xq.shape[0])
        self._check()
        return i

    def query_search(self, qid: int, nq: int) -> np.ndarray:
        out = np.zeros(nq, np.int32)
        self.L.tioga_amd_query_set_search(qid, _p(out))
        self._check()
        return out

    def query_nelem(self, qid: int) -> int:
        n = self.L.tioga_amd_query_set_nelem(qid)
        self._check()
        return int(n)

    def receptor_info(self) -> np.ndarray:
        n = self.L.tioga_amd_receptor_info(None)
        self._check()
        out = np.zeros(max(n, 0), np.int32)
        if n > 0:
            self.L.tioga_amd_receptor_info(_p(out))
            self._check()
        return out

    def query_clear(self) -> None:
        self.L.tioga_amd_query_set_clear()
        self._check()

    def query_donors(self, qid: int, nq: int) -> np.ndarray:
        """Donors as the last pass over the set left them (no new search)."""
        out = np.zeros(nq, np.int32)
        self.L.tioga_amd_query_set_donors(qid, _p(out))
        self._check()
        return out

    def query_extents(self, qid: int) -> np.ndarray:
        ext = np.zeros(6)
        self.L.tioga_amd_query_set_extents(qid, _p(ext))
        self._check()
        return ext

    def search_tuning(self, morton_bits: Optional[int] = None, seg: Optional[int] = None,
                      eval_grid: Optional[int] = None, bins: Optional[int] = None) -> Dict[str, int]:
        """Sets the given search parameters; returns all four (bins: elements
        per bin of the candidate bin grids, 0 = the tree walk)."""
        nb = self.L.tioga_amd_search_bins(-1 if bins is None else int(bins))
        self._check()
        cur = np.zeros(3, np.int32)
        self.L.tioga_amd_search_tuning(None, _p(cur))
        self._check()
        want = cur.copy()
        for i, v in enumerate((morton_bits, seg, eval_grid)):
            if v is not None:
                want[i] = v
        if not np.array_equal(want, cur):
            self.L.tioga_amd_search_tuning(_p(want), _p(cur))
            self._check()
        return {"morton_bits": int(cur[0]), "seg": int(cur[1]), "eval_grid": int(cur[2]), "bins": int(nb)}

    def search_graph(self, on: bool = True) -> None:
        """Timed query-set passes replay their search launch from a HIP graph
        (on) or launch its kernels one by one."""
        self.L.tioga_amd_search_graph(1 if on else 0)
        self._check()

    def search_capacity(self, cap: int = 0) -> int:
        r = self.L.tioga_amd_search_capacity(int(cap))
        self._check()
        return int(r)

    def time_query_sets(self, nrep: int):
        """(ms per pass, queries per pass, ms per pass in the search kernel alone)"""
        n = C.c_longlong(0)
        ms = self.L.tioga_amd_time_query_sets(nrep, C.byref(n))
        self._check()
        return ms, n.value, self.L.tioga_amd_last_search_kernel_ms()

    def search_stats(self) -> Dict[str, int]:
        """Counts of the last search launch: candidates listed, listing work
        items, queries that searched a tree, queries, and the candidates by
        test outcome."""
        out = np.zeros(10, np.int64)
        self.L.tioga_amd_search_stats(_p(out))
        self._check()
        return {"candidates": int(out[0]), "listing_items": int(out[1]), "walking_queries": int(out[2]),
                "queries": int(out[3]), "rejected": int(out[4]), "accepted_near_zero": int(out[5]),
                "accepted": int(out[6]), "outside_element_box": int(out[7]), "after_first_hit": int(out[8]),
                "listing": "bins" if out[9] else "walk"}

    def phase_times(self) -> Dict[str, float]:
        out = np.zeros(len(PHASES))
        self.L.tioga_amd_phase_times(_p(out))
        self._check()
        return dict(zip(PHASES, out.tolist()))

    def weight_aborts(self) -> int:
        return int(self.L.tioga_amd_weight_aborts())

    def results(self, qs=None, nvar: int = 0, interptype: int = 0) -> Dict[str, np.ndarray]:
        """Every array the reference driver dumps, under the same names; with
        blocks on several virtual ranks, each rank's arrays under "r<rank>."
        as the multi-rank driver's files merge them."""
        ranks = sorted(set(self.ranks))
        if ranks != [0]:
            r: Dict[str, np.ndarray] = {}
            for k in ranks:
                for nm, a in self._results_of_rank(k).items():
                    r[f"r{k}." + nm] = a
            if qs is not None:
                for i, q in enumerate(self.data_update_copies(qs, nvar, interptype)):
                    r[f"r{self.ranks[i]}.b{i}.q"] = q
            self.set_rank(0)
            return r
        r = self._results_of_rank(0)
        if qs is not None:
            for i, q in enumerate(self.data_update_copies(qs, nvar, interptype)):
                r[f"b{i}.q"] = q
        return r

    def _results_of_rank(self, rank: int) -> Dict[str, np.ndarray]:
        self.set_rank(rank)
        r: Dict[str, np.ndarray] = {}
        for i, b in enumerate(self.blocks):
            if self.ranks[i] != rank:
                continue
            p = f"b{i}."
            r[p + "iblank"] = self.iblank[i].copy()
            r[p + "iblank_cell"] = self.array(b.tag, "iblank_cell")
            for nm in ("nodeRes", "cellRes", "obb", "xsearch", "res_search", "isearch", "tagsearch",
                       "donorId", "xtag", "elementList", "adtIntegers", "adtReals", "adtExtents"):
                a = self.array(b.tag, nm)
                if a.size or nm != "adtExtents":
                    r[p + nm] = a
            r[p + "nsearch"] = np.array([r[p + "donorId"].size], np.int32)
            rec, ind, frac = self.donor_info(b.tag)
            r[p + "dcount"] = np.array([rec.size // 4], np.int32)
            r[p + "receptors"], r[p + "indices"], r[p + "frac"] = rec, ind, frac
        if any(b.cell_gid is not None for b in self.blocks):
            r["receptor_info"] = self.receptor_info()
        mt = max(b.tag for b in self.blocks)
        for m in range(1, mt + 1):
            hm = self.hole_map(m)
            r[f"hm{m}.existWall"] = np.array([0 if hm is None else 1], np.int32)
            if hm is not None:
                r[f"hm{m}.nx"], r[f"hm{m}.extents"], r[f"hm{m}.sam"] = hm
        return r
