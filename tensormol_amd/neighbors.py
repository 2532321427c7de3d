"""Drop-in neighbour-list API backed by the HIP kernels.

Same classes, methods, argument meaning and output formats as
TensorMol/ForceModifiers/Neighbors.py (NeighborList :23-201, NeighborListSet
:203-471, NeighborListSetWithImages :473-489) and MolEmb.Make_NListNaive
(C_API/MolEmb.cpp:1180-1247).  Sets are bit-exact (same float64 predicate);
rows are returned sorted by atom index where the reference's order followed
its x-sorted sweep (every reference consumer re-sorts or sums them).

These entry points copy host arrays in and out, as the reference API does.
The engine's own path keeps the lists resident on the device.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .engine import nl_threshold

_P = _lib.ptr


def _dev():
    if not torch.cuda.is_available():
        raise RuntimeError("tensormol_amd neighbour lists run on the GPU (HIP); no device found")
    return torch.device("cuda")


class _Packed:
    """Packed device copy of padded molecules (x (B, maxN, 3), nnz (B,))."""

    def __init__(self, x, nnz, ele=None, nreal=None):
        x = np.asarray(x, dtype=np.float64)
        if x.ndim == 2:
            x = x[None]
        nnz = np.atleast_1d(np.asarray(nnz, dtype=np.int64))
        self.B = x.shape[0]
        self.nnz = nnz
        self.ms = np.concatenate([[0], np.cumsum(nnz)]).astype(np.int64)
        self.N = int(self.ms[-1])
        pos = np.concatenate([x[m, :nnz[m]] for m in range(self.B)]) if self.N else np.zeros((0, 3))
        self.pos_np = np.ascontiguousarray(pos)
        self.mol_np = np.repeat(np.arange(self.B, dtype=np.int32), nnz)
        dev = _dev()
        self.dev = dev
        self.pos = torch.tensor(self.pos_np, device=dev)
        self.mol = torch.tensor(self.mol_np, device=dev)
        self.mol_start = torch.tensor(self.ms, dtype=torch.int32, device=dev)
        self.nreal = None if nreal is None else torch.tensor(np.atleast_1d(nreal), dtype=torch.int32,
                                                             device=dev)
        self.nreal_np = nnz if nreal is None else np.atleast_1d(np.asarray(nreal, dtype=np.int64))
        if ele is not None:
            ele = np.asarray(ele)
            if ele.ndim == 1:
                ele = ele[None]
            self.Z = np.concatenate([ele[m, :nnz[m]] for m in range(self.B)]).astype(np.int64)
        else:
            self.Z = None

    def local(self, gi):
        m = self.mol_np[gi]
        return m, gi - self.ms[m]


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _pairs(pk, rc, half):
    """(P,3) int64 [m, i_local, j_local] of all pairs within rc; centres limited
    to the first nreal atoms of each molecule."""
    L = _lib.lib()
    st = _stream()
    thr = nl_threshold(rc)
    count = torch.zeros(max(1, pk.N), dtype=torch.int32, device=pk.dev)
    flags = torch.zeros(8, dtype=torch.int32, device=pk.dev)
    args = (pk.N, _P(pk.pos), _P(pk.mol), _P(pk.mol_start), _P(pk.nreal), None, None, None, None,
            None, None, None, 1, 1, 1, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0, float(rc), thr, int(half))
    L.tm_pairs(*args, _P(count), None, None, _P(flags), st)
    cnt = count[:pk.N].cpu().numpy().astype(np.int64)
    start = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    out = torch.zeros(max(1, int(start[-1])), dtype=torch.int32, device=pk.dev)
    st_t = torch.tensor(start[:-1] if pk.N else np.zeros(1, np.int64), device=pk.dev)
    L.tm_pairs(*args, _P(count), _P(st_t), _P(out), _P(flags), st)
    j = out.cpu().numpy()[: start[-1]].astype(np.int64)
    i = np.repeat(np.arange(pk.N), cnt)
    m = pk.mol_np[i].astype(np.int64)
    return np.stack([m, i - pk.ms[m], j - pk.ms[m]], axis=1)


def Make_NListNaive(x, rc, nreal, DoPerms):
    """MolEmb.Make_NListNaive (MolEmb.cpp:1180-1247): list of nreal lists of
    neighbour indices (sorted).  x: (N, 3) float64."""
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 2 and x.shape[0] == 3 and x.shape[1] != 3:
        x = x.T
    pk = _Packed(x, [x.shape[0]], nreal=[int(nreal)])
    p = _pairs(pk, float(rc), half=0)
    rows = [[] for _ in range(int(nreal))]
    if int(DoPerms) == 1:
        for _, i, j in p:
            rows[i].append(int(j))
    else:
        # each unordered pair once, under min(i, j), for pairs with a real atom
        for _, i, j in p:
            a, b = (i, j) if i < j else (j, i)
            if i < j or j >= int(nreal):
                rows[a].append(int(b))
        rows = [sorted(r) for r in rows]
    return rows


class NeighborListSet:
    """NeighborListSet(x_ (nmol, maxN, 3), nnz_ (nmol,), DoTriples_, DoPerms_, ele_ (nmol, maxN),
    alg_, sort_) (Neighbors.py:203-246)."""

    def __init__(self, x_, nnz_, DoTriples_=False, DoPerms_=False, ele_=None, alg_=None, sort_=False):
        self.x = np.asarray(x_, dtype=np.float64).copy()
        self.nnz = np.asarray(nnz_).copy()
        self.nreal = np.asarray(nnz_).copy()
        self.nmol = self.x.shape[0]
        self.maxnatom = self.x.shape[1]
        self.ele = ele_
        self.sort = sort_
        self.DoTriples = DoTriples_
        self.DoPerms = DoPerms_
        self.alg = 0 if alg_ is None else alg_
        self.pairs = None
        self.triples = None
        self.UpdateInterval = 1
        self.UpdateCounter = 0

    def _packed(self):
        return _Packed(self.x, self.nnz, self.ele, self.nreal)

    def Update(self, x_, rcut_pairs=5.0, rcut_triples=5.0):
        if self.UpdateCounter == 0:
            self.UpdateCounter += 1
            self.x = np.asarray(x_, dtype=np.float64).copy()
            if self.DoTriples:
                self.pairs, self.triples = self.buildPairsAndTriples(rcut_pairs, rcut_triples)
            else:
                self.pairs = self.buildPairs(rcut_pairs)
        elif self.UpdateCounter < self.UpdateInterval:
            self.UpdateCounter += 1
        else:
            self.UpdateCounter = 0

    def buildPairs(self, rcut=5.0):
        """(P, 3) uint64 [m, i, j]: full list when DoPerms, else each pair once
        under min(i, j) (Neighbors.py:262-282, MolEmb.cpp:1220-1228)."""
        pk = self._packed()
        p = _pairs(pk, float(rcut), half=0 if self.DoPerms else 1)
        if not self.DoPerms and np.any(self.nreal != self.nnz):
            # pairs whose lower atom is an image but whose upper atom is real
            # are stored under the lower index (MolEmb.cpp:1222-1228)
            extra = _pairs(_Packed(self.x, self.nnz, self.ele, None), float(rcut), half=1)
            nr = self.nreal[extra[:, 0]]
            keep = (extra[:, 1] >= nr) & (extra[:, 2] < nr)
            p = np.concatenate([p, extra[keep]])
        p = p[np.lexsort((p[:, 2], p[:, 1], p[:, 0]))]
        return p.astype(np.uint64)

    def buildPairsAndTriples(self, rcut_pairs=5.0, rcut_triples=5.0):
        """(trp (P,3), trt (T,4)) uint64 [m, i, j(, k)] with the lower-Z atom first
        (Neighbors.py:117-201, 284-321); rows sorted by index."""
        trpE, trtE, _, _ = self._with_ele(rcut_pairs, rcut_triples, None)
        return trpE[:, :3].astype(np.uint64), trtE[:, :4].astype(np.uint64)

    def buildPairsWithBothEleIndex(self, rcut=5.0, ele=None, sort_=False):
        """(P, 5) [m, i, j, l_i, l_j] (Neighbors.py:323-342)."""
        trp = self.buildPairs(rcut).astype(np.int64)
        Zs = np.asarray(self.ele)
        eles = np.asarray(ele).reshape(-1)
        lut = {int(e): a for a, e in enumerate(eles)}
        li = np.array([lut[int(Zs[m, i])] for m, i in trp[:, :2]], dtype=np.int64).reshape(-1, 1)
        lj = np.array([lut[int(Zs[m, j])] for m, j in trp[:, [0, 2]]], dtype=np.int64).reshape(-1, 1)
        out = np.concatenate([trp, li, lj], axis=1)
        if sort_:
            swap = out[:, 3] > out[:, 4]
            out[swap] = out[swap][:, [0, 2, 1, 4, 3]]
        return out.astype(np.uint64)

    def buildPairsAndTriplesWithEleIndex(self, rcut_pairs=5.0, rcut_triples=5.0, ele=None, elep=None):
        """trpE (P,4) [m,i,j,l] sorted (m,i,l,j); trtE (T,5) [m,i,j,k,l] sorted
        (m,i,l,k,j); mil_jk (T,4) float64 [m,i,l,rank]; jk_max (Neighbors.py:344-423)."""
        eles = None if ele is None else [int(e) for e in np.asarray(ele).reshape(-1)]
        trpE, trtE, mil_jk, jk_max = self._with_ele(rcut_pairs, rcut_triples, eles)
        return trpE.astype(np.uint64), trtE.astype(np.uint64), mil_jk, jk_max

    def buildPairsAndTriplesWithEleIndexPeriodic(self, rcut_pairs=5.0, rcut_triples=5.0, ele=None,
                                                 elep=None):
        """As above plus mil_j (P,4) [m,i,l,rank] (Neighbors.py:426-467)."""
        trpE, trtE, mil_jk, _ = self.buildPairsAndTriplesWithEleIndex(rcut_pairs, rcut_triples, ele,
                                                                      elep)
        t = trpE.astype(np.int64)
        mil_j = np.zeros((len(t), 4))
        if len(t):
            key = t[:, [0, 1, 3]]
            new = np.ones(len(t), dtype=bool)
            new[1:] = np.any(key[1:] != key[:-1], axis=1)
            start = np.maximum.accumulate(np.where(new, np.arange(len(t)), 0))
            mil_j[:, :3] = key
            mil_j[:, 3] = np.arange(len(t)) - start
        return trpE, trtE, mil_j, mil_jk

    # ------------------------------------------------------------------
    def _with_ele(self, rr, ra, eles):
        from .engine import BPEEEngine  # noqa: F401  (library load check)
        pk = self._packed()
        if pk.Z is None:
            raise ValueError("element types (ele_) are needed for triples")
        if eles is None:
            eles = sorted(set(int(z) for z in pk.Z))
        lut = {e: a for a, e in enumerate(eles)}
        elem = np.array([lut[int(z)] for z in pk.Z], dtype=np.int32)
        nele = len(eles)
        L = _lib.lib()
        st = _stream()
        dev = pk.dev
        ldr, lda = 256, 128
        i32 = dict(dtype=torch.int32, device=dev)
        N = pk.N
        el = torch.tensor(elem, device=dev)
        nbr = torch.zeros(max(1, N * ldr), **i32)
        ncnt = torch.zeros(max(1, N), **i32)
        rseg = torch.zeros(max(1, N * (nele + 1)), **i32)
        ang = torch.zeros(max(1, N * lda), **i32)
        acnt = torch.zeros(max(1, N), **i32)
        aseg = torch.zeros(max(1, N * (nele + 1)), **i32)
        rev = torch.zeros(max(1, N * ldr), **i32)
        flags = torch.zeros(8, **i32)
        L.tm_nl_build(N, _P(pk.pos), _P(el), _P(pk.mol), _P(pk.mol_start), None, None, None, None,
                      None, None, None, 1, 1, 1, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0, nele, float(rr),
                      nl_threshold(rr),
                      nl_threshold(ra), ldr, lda, _P(nbr), _P(ncnt), _P(rseg), _P(ang), _P(acnt),
                      _P(aseg), _P(rev), _P(flags), st)
        f = flags.cpu().numpy()
        if f[0] or f[1]:
            raise RuntimeError(f"neighbour rows exceed {ldr} (pairs) or {lda} (triples) entries")
        ncnt_h = ncnt.cpu().numpy()[:N]
        acnt_h = acnt.cpu().numpy()[:N].astype(np.int64)
        nbr_h = nbr.cpu().numpy()[: N * ldr].reshape(N, ldr) if N else np.zeros((0, ldr), np.int32)
        # centres limited to the first nreal atoms of each molecule
        loc = np.arange(N) - pk.ms[pk.mol_np]
        real = loc < pk.nreal_np[pk.mol_np]
        rows = []
        for i in np.nonzero(real)[0]:
            m = pk.mol_np[i]
            js = nbr_h[i, : ncnt_h[i]]
            rows.append(np.stack([np.full(len(js), m), np.full(len(js), loc[i]), js - pk.ms[m],
                                  elem[js]], axis=1))
        trpE = np.concatenate(rows).astype(np.int64) if rows else np.zeros((0, 4), np.int64)
        ntrip = acnt_h * (acnt_h - 1) // 2
        ntrip[~real] = 0
        tstart = np.concatenate([[0], np.cumsum(ntrip)]).astype(np.int64)
        total = int(tstart[-1])
        out = torch.zeros(max(1, total) * 4, **i32)
        acnt_eff = torch.tensor(np.where(real, acnt_h, 0).astype(np.int32), device=dev)
        ts = torch.tensor(tstart[:-1] if N else np.zeros(1, np.int64), device=dev)
        L.tm_triples(N, ldr, lda, nele, _P(nbr), _P(ang), _P(acnt_eff), _P(aseg), _P(ts), _P(out), st)
        o = out.cpu().numpy().reshape(-1, 4)[:total].astype(np.int64)
        ci = np.repeat(np.arange(N), ntrip)
        cm = pk.mol_np[ci].astype(np.int64)
        trtE = np.stack([cm, ci - pk.ms[cm], o[:, 0] - pk.ms[cm], o[:, 1] - pk.ms[cm], o[:, 2]], axis=1) \
            if total else np.zeros((0, 5), np.int64)
        mil_jk = np.zeros((total, 4))
        if total:
            mil_jk[:, 0] = cm
            mil_jk[:, 1] = ci - pk.ms[cm]
            mil_jk[:, 2] = o[:, 2]
            mil_jk[:, 3] = o[:, 3]
        jk_max = float(mil_jk[:, 3].max()) if total else 0
        return trpE, trtE, mil_jk, jk_max


class NeighborListSetWithImages(NeighborListSet):
    """Neighbour lists whose centres are the first nreal atoms of each molecule
    (Neighbors.py:473-489)."""

    def __init__(self, x_, nnz_, nreal_, DoTriples_=False, DoPerms_=False, ele_=None, alg_=None,
                 sort_=False):
        NeighborListSet.__init__(self, x_, nnz_, DoTriples_, DoPerms_, ele_, alg_, sort_)
        self.nreal = np.atleast_1d(np.asarray(nreal_))


class NeighborList(NeighborListSet):
    """Single-molecule NeighborList (Neighbors.py:23-201)."""

    def __init__(self, x_, DoTriples_=False, DoPerms_=False, ele_=None, alg_=None, sort_=False):
        x = np.asarray(x_, dtype=np.float64)
        ele = None if ele_ is None else np.asarray(ele_)[None]
        NeighborListSet.__init__(self, x[None], np.array([x.shape[0]]), DoTriples_, DoPerms_, ele,
                                 alg_, sort_)
        self.natom = x.shape[0]
        self.npairs = None
        self.ntriples = None

    def Update(self, x_, rcut_pairs=5.0, rcut_triples=5.0, molind_=None, nreal_=None):
        self.x = np.asarray(x_, dtype=np.float64)[None].copy()
        if nreal_ is not None:
            self.nreal = np.array([int(nreal_)])
        if self.DoTriples:
            p, t = NeighborListSet.buildPairsAndTriples(self, rcut_pairs, rcut_triples)
            if molind_ is None:
                p, t = p[:, 1:], t[:, 1:]
            else:
                p[:, 0] = molind_
                t[:, 0] = molind_
            self.pairs, self.triples = p, t
            self.npairs, self.ntriples = len(p), len(t)
        else:
            p = NeighborListSet.buildPairs(self, rcut_pairs)
            if molind_ is None:
                p = p[:, 1:]
            else:
                p[:, 0] = molind_
            self.pairs = p
            self.npairs = len(p)
