"""Device-resident BP+EE energy/force engine over packed atoms.

One evaluation is a fixed sequence of HIP launches on the caller's current
stream (no host synchronisation except the optional overflow check):

  cell grid (large systems) -> neighbour rows (4.6 A + 3.1 A sub-list, mirror
  slots) -> descriptors G -> per-element energy and charge MLPs with their input
  gradients -> neutral charges -> DSF-elu Coulomb + vdW (phi_i, direct
  gradient) -> charge upstream u_i = phi_i - mean(phi) -> descriptor adjoint
  with dE/dG = dE_i/dG + u_i dq_i/dG -> mirror-slot force gather ->
  per-molecule FP64 reductions.

This is the data flow of MolInstance_DirectBP_EE_ChargeEncode_Update_vdw_DSF_
elu_Normalize_Dropout.evaluate (TensorMol/TFNetworks/TFMolInstanceDirect.py:
5684-5711) with tf.gradients (:5761) replaced by the hand-written adjoint.
Memory layout in HBM: positions float64 [N,3]; descriptors float32 [N,D];
neighbour rows int32 ELL [N, ldr]; per-atom scalars float32/float64 [N].
"""
import ctypes as _ctypes
import math
import struct

import numpy as np
import torch

from . import _lib
from .params import PARAMS, BOHRPERA, JOULEPERHARTREE, TWO_OVER_SQRTPI
from .model import BPEEModel

_P = _lib.ptr

# largest table (float4 entries, 48 KiB) the tabulated Coulomb sweep accepts
CL_TABLE_ENTRIES = 3072


def nl_threshold(rc):
    """First float64 s for which sqrt(s) + 1e-13 < rc is false (MolEmb.cpp:1214-1218).

    sqrt and + are correctly rounded and monotone, so the reference predicate
    equals s < threshold for s = (dx*dx + dy*dy) + dz*dz."""
    def as_d(k):
        return struct.unpack("<d", struct.pack("<q", k))[0]

    def pred(s):
        return math.sqrt(s) + 0.0000000000001 < rc

    lo = 0
    hi = struct.unpack("<q", struct.pack("<d", 4.0 * rc * rc + 1.0))[0]
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if pred(as_d(mid)):
            lo = mid
        else:
            hi = mid
    return as_d(hi)


def dsf_constants(dsf_alpha, elu_width, ee_off):
    """[alpha', R_sr, R_lr, ZZ, YY, elu_shift, elu_alpha] in bohr units.

    elu_shift = DSF(R_sr), elu_alpha = DSF_Gradient(R_sr) (Util.py:172-192 as
    used at TFMolInstanceDirect.py:4371-4372); ZZ and YY as in
    TFCoulombEluSRDSFLR (RawSymFunc.py:1338-1340)."""
    a = dsf_alpha / BOHRPERA
    R = elu_width * BOHRPERA
    Rc = ee_off * BOHRPERA
    XX = a * Rc
    ZZ = math.erfc(XX) / Rc
    YY = TWO_OVER_SQRTPI * a * math.exp(-XX * XX) / Rc
    shift = math.erfc(a * R) / R - ZZ + (R - Rc) * (ZZ / Rc + YY)
    grad = -((math.erfc(a * R) / R / R + TWO_OVER_SQRTPI * a * math.exp(-a * R * a * R) / R)
             - (ZZ / Rc + YY))
    return np.array([a, R, Rc, ZZ, YY, shift, grad], dtype=np.float64)


class Workspace:
    """Device buffers and static topology of one packed batch."""
    pass


class BPEEEngine:
    """Energies, charges and forces of a BPEEModel on one GPU.

    precision: "fp32" (plain FP32 VALU GEMMs, every intermediate in HBM),
    "bf16" / "fp16" (fused MFMA kernel, FP32 accumulation) or "bf16x3"
    (MFMA with hi/lo operand splitting, near-FP32 accuracy).
    """

    def __init__(self, model: BPEEModel, device="cuda", precision=None, ldr=128, lda=48,
                 check_overflow=True, coulomb_tables=True):
        self.model = model
        self.device = torch.device(device)
        self._dev_index = self.device.index   # None: the current device at each call
        self.precision = precision or PARAMS["MLPPrecision"]
        self.ldr, self.lda = int(ldr), int(lda)
        self.check_overflow = check_overflow
        self.lib = _lib.lib()
        sym = model.sym
        self.nele = sym.nele
        self.D = sym.D
        self.s_r = nl_threshold(sym.rr)
        self.s_a = nl_threshold(sym.ra)
        self.sym_rs_r = np.ascontiguousarray(sym.rs_r, dtype=np.float32)
        self.sym_rs_a = np.ascontiguousarray(sym.rs_a, dtype=np.float32)
        self.sym_th = np.ascontiguousarray(sym.thetas, dtype=np.float32)
        self.dsf = dsf_constants(model.dsf_alpha, model.elu_width, model.ee_off)
        c6, r0 = model.vdw_tables()
        self.c6 = np.ascontiguousarray(c6, dtype=np.float64)
        self.r0 = np.ascontiguousarray(r0, dtype=np.float64)
        self.vdw_width = model.poly_width * BOHRPERA
        self.hid = np.array(model.hidden, dtype=np.int32)
        # One large system on a cell grid takes the Coulomb sweep with
        # tabulated pair functions (256 intervals over the cutoff; up to four
        # elements, 44 KiB of tables).
        self.coulomb_tables = bool(coulomb_tables)
        ntab = 1 + self.nele * (self.nele + 1) // 2
        self.cl_nint = 256
        if ntab * self.cl_nint > CL_TABLE_ENTRIES:
            self.coulomb_tables = False
        self._cl_tables = {}
        self._pack_f32()
        if self.precision != "fp32":
            from . import mlp_mfma
            self.mfma = mlp_mfma.MFMAWeights(self, self.precision)

    # ------------------------------------------------------------ weights
    def _pack_f32(self):
        self.w_f32 = {}
        for tag, nets in (("E", self.model.energy_nets), ("Q", self.model.charge_nets)):
            for a, e in enumerate(self.model.eles):
                parts = []
                for (W, b) in nets[e]:
                    parts.append(np.asarray(W, np.float32).ravel())
                    parts.append(np.asarray(b, np.float32).ravel())
                flat = np.concatenate(parts)
                self.w_f32[(tag, a)] = torch.tensor(flat, device=self.device)

    # ---------------------------------------------------------- topology
    def prepare(self, Z, mol_start, boxes=None, margin=4.0, pos0=None):
        """Allocate device buffers for a packed batch (Z (N,), mol_start (B+1,),
        optional boxes (B,3) orthorhombic edges in A; 0 or None = open)."""
        ms = np.asarray(mol_start, dtype=np.int64)
        ws = self.alloc(len(Z), len(ms) - 1)
        self.set_topology(ws, Z, ms, boxes, margin=margin, pos0=pos0)
        return ws

    def alloc(self, ncap, bcap):
        """Device buffers for up to ncap atoms in up to bcap molecules."""
        dev = self.device
        ws = Workspace()
        ws.ncap, ws.bcap = int(ncap), int(bcap)
        ws.epoch, ws.argc = 0, {}   # layout version and the launch arguments cached for it
        N, B = max(1, ws.ncap), max(1, ws.bcap)
        i32 = dict(dtype=torch.int32, device=dev)
        f32 = dict(dtype=torch.float32, device=dev)
        f64 = dict(dtype=torch.float64, device=dev)
        ws.elem = ws.buf_elem = torch.zeros(N, **i32)
        ws.mol = ws.buf_mol = torch.zeros(N, **i32)
        ws.mol_start = ws.buf_mol_start = torch.zeros(B + 1, **i32)
        ws.perm = ws.buf_perm = torch.zeros(N, **i32)
        ws.nbr = torch.zeros(N * self.ldr, **i32)
        ws.ncnt = torch.zeros(N, **i32)
        ws.rseg = torch.zeros(N * (self.nele + 1), **i32)
        ws.ang = torch.zeros(N * self.lda, **i32)
        ws.acnt = torch.zeros(N, **i32)
        ws.aseg = torch.zeros(N * (self.nele + 1), **i32)
        ws.rev = torch.zeros(N * self.ldr, **i32)
        ws.flags = torch.zeros(8, **i32)
        ws.red = torch.zeros(((N + 1023) // 1024) * 8 + 8, **f64)
        ws.G = torch.zeros(N, self.D, **f32)
        ws.Y2 = torch.zeros(2 * N, **f32)
        ws.J2 = torch.zeros(2 * N * self.D, **f32)
        ws.qmean = torch.zeros(B, **f64)
        ws.phimean = torch.zeros(B, **f64)
        ws.xq = torch.zeros(N, 4, **f32)
        ws.xqc = None
        ws.elc = None
        ws.q = torch.zeros(N, **f64)
        ws.qx = torch.zeros(N, 3, **f64)
        ws.phi = torch.zeros(N, **f64)
        ws.gradc = torch.zeros(N, 3, **f64)
        ws.eatom = torch.zeros(N, 2, **f64)
        ws.u = torch.zeros(N, **f32)
        ws.wslot = torch.zeros(N * self.ldr * 4, **f32)
        ws.grad = torch.zeros(N, 3, **f64)
        ws.force = torch.zeros(N, 3, **f64)
        ws.Ebp = torch.zeros(B, **f64)
        ws.Eext = torch.zeros(B, 2, **f64)
        ws.dipole = torch.zeros(B, 3, **f64)
        ws.mlp_ws = None
        ws.use_cells = False
        return ws

    def set_topology(self, ws, Z, mol_start, boxes=None, margin=4.0, pos0=None, stager=None):
        """Point a workspace at a new packed batch (N <= ncap, B <= bcap).
        Only small host-built index arrays are copied; no buffer is reallocated
        unless a cell grid is (re)built for a single large system.  With an
        open PinnedStager the arrays join its batch: the workspace then points
        at their device copies, which are valid for work queued after the
        stager's flush()."""
        up = (lambda a, out=None: stager.put(a, out)) if stager is not None else _plain_upload(self.device)
        dev = self.device
        Z = np.asarray(Z)
        ms = np.asarray(mol_start, dtype=np.int64)
        N, B = len(Z), len(ms) - 1
        if N > ws.ncap or B > ws.bcap:
            raise ValueError(f"batch of {N} atoms / {B} molecules exceeds the workspace "
                             f"({ws.ncap} / {ws.bcap})")
        ws.N, ws.B = N, B
        elem = self.model.element_index(Z)
        ws.elem_np = elem
        ws.mol_np = np.repeat(np.arange(B, dtype=np.int32), np.diff(ms))
        ws.ms_np = ms
        perm = np.argsort(elem.astype(np.uint8), kind="stable").astype(np.int32)   # radix sort
        counts = np.bincount(elem, minlength=self.nele)
        ws.eoff = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        if stager is not None:
            # per-batch index arrays are views of the stager's device arena,
            # valid for the work queued after its flush
            if N:
                ws.elem = up(elem.astype(np.int32))
                ws.mol = up(ws.mol_np)
                ws.perm = up(perm)
            ws.mol_start = up(ms.astype(np.int32))
        else:
            ws.elem, ws.mol, ws.perm = ws.buf_elem, ws.buf_mol, ws.buf_perm
            ws.mol_start = ws.buf_mol_start
            if N:
                up(elem, ws.elem[:N])
                up(ws.mol_np, ws.mol[:N])
                up(perm, ws.perm[:N])
            up(ms.astype(np.int32), ws.mol_start[:B + 1])
        ws.box = None
        ws.periodic = 0
        ws.boxes_np = None
        if boxes is not None:
            bx = np.asarray(boxes, dtype=np.float64).reshape(B, 3)
            if np.any(bx > 0):
                if np.any(bx <= 0):
                    raise ValueError("mixed open and periodic molecules in one batch")
                if np.any(bx < 2.0 * max(self.model.sym.rr, self.model.ee_off)):
                    raise ValueError("box edge below twice the largest cutoff: minimum image is ambiguous")
                ws.box = torch.tensor(bx, device=dev)
                ws.periodic = 1
                ws.boxes_np = bx
        # network outputs laid out for this N: y[net*N + atom], dG[(net*N + atom)*D + f]
        ws.eps, ws.qraw = ws.Y2[:N], ws.Y2[N:2 * N]
        ws.JE = ws.J2[:N * self.D].view(N, self.D)
        ws.JQ = ws.J2[N * self.D:2 * N * self.D].view(N, self.D)
        # cell grid for one large system
        ws.use_cells = (B == 1 and N > 256)
        if ws.use_cells:
            self._setup_grid(ws, pos0, margin)
            ws.xqc = torch.zeros(max(1, N), 4, dtype=torch.float32, device=dev)
            ws.elc = torch.zeros(max(1, N), dtype=torch.int32, device=dev)
        if self.precision == "fp32":
            hmax = int(self.hid.max())
            maxrows = int(counts.max()) if N else 0
            nh = len(self.hid)
            need = max(1, (2 * nh + 2) * maxrows * hmax)
            if ws.mlp_ws is None or ws.mlp_ws.numel() < need:
                ws.mlp_ws = torch.zeros(need, dtype=torch.float32, device=dev)
        else:
            ws.mlp_ws = self.mfma.workspace(ws, up)
        # every cached launch argument list of this workspace is stale now
        ws.epoch = getattr(ws, "epoch", 0) + 1
        ws.argc = {}
        return ws

    def _setup_grid(self, ws, pos0, margin):
        """Cell grid with edges of about half the short cutoff (>= rr/2)."""
        target = 0.5 * self.model.sym.rr
        if ws.periodic:
            L = ws.boxes_np[0]
            nc = np.maximum(1, np.floor(L / target)).astype(int)
            cs = L / nc
            origin = np.zeros(3)
        else:
            if pos0 is None:
                raise ValueError("open large systems need pos0 to size the cell grid")
            p = np.asarray(pos0)
            lo = p.min(0) - margin
            hi = p.max(0) + margin
            nc = np.maximum(1, np.floor((hi - lo) / target)).astype(int)
            cs = (hi - lo) / nc
            origin = lo
        ws.grid_nc = [int(v) for v in nc]
        ws.grid_cs = [float(v) for v in cs]
        ws.grid_origin = [float(v) for v in origin]
        ncell = int(np.prod(nc))
        i32 = dict(dtype=torch.int32, device=self.device)
        ws.cell_of = torch.zeros(ws.N, **i32)
        ws.cell_count = torch.zeros(ncell, **i32)
        ws.cell_start = torch.zeros(ncell + 1, **i32)
        ws.cell_fill = torch.zeros(max(ncell, ws.N), **i32)   # counters, then sort staging
        ws.cell_atoms = torch.zeros(ws.N, **i32)
        ws.cell_pos = torch.zeros(ws.N, **i32)
        ws.posc = torch.zeros(ws.N, 3, dtype=torch.float64, device=self.device)
        ws.molc = torch.zeros(ws.N, **i32) if ws.B > 1 else None

    def _grid_args(self, ws):
        """(cell_start, cell_atoms, cell_of, cell_pos, posc, molc, nx, ny, nz, cx, cy, cz,
        ox, oy, oz)."""
        if not ws.use_cells:
            return (None, None, None, None, None, None, 1, 1, 1, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0)
        n, c, o = ws.grid_nc, ws.grid_cs, ws.grid_origin
        return (_P(ws.cell_start), _P(ws.cell_atoms), _P(ws.cell_of), _P(ws.cell_pos), _P(ws.posc),
                _P(ws.molc), n[0], n[1], n[2], c[0], c[1], c[2], o[0], o[1], o[2])

    # ---------------------------------------------------------- stages
    def _call(self, ws, name, key, build):
        """Call entry point `name` with the argument list build() returns.
        The list (pointers already converted) is kept on the workspace and
        reused while the workspace's layout and `key` (whatever else the
        arguments depend on: positions, stream, settings) are unchanged."""
        k = (ws.epoch, key)
        e = ws.argc.get(name)
        if e is None or e[0] != k:
            e = (k, tuple(build()))
            ws.argc[name] = e
        getattr(self.lib, name.split("#")[0])(*e[1])

    def _raw_stream(self):
        idx = self._dev_index if self._dev_index is not None else torch.cuda.current_device()
        return torch._C._cuda_getCurrentRawStream(idx)

    def _stream(self):
        """The current stream of the engine's device as a ctypes pointer (read
        without resolving a device object: it is asked for at every launch)."""
        idx = self._dev_index if self._dev_index is not None else torch.cuda.current_device()
        return _ctypes.c_void_p(torch._C._cuda_getCurrentRawStream(idx))

    def build_lists(self, ws, pos, clear_flags=True):
        """Cell grid and neighbour rows for positions pos.  The status flags
        (row overflow, non-finite coordinates) are cleared first unless
        clear_flags is False: steps captured together into one graph then
        accumulate them, and a check after the replay sees every step's."""
        key = (pos.data_ptr(), self._raw_stream())
        ws.rev_ok = False   # the mirror slots are filled by the next descriptor forward
        if clear_flags:
            self._call(ws, "tm_zero_i32", key, lambda: (8, _P(ws.flags), self._stream()))
        if ws.use_cells:
            def cell_args():
                o, c, n = ws.grid_origin, ws.grid_cs, ws.grid_nc
                return (ws.N, _P(pos), _P(ws.mol), o[0], o[1], o[2], c[0], c[1], c[2], n[0], n[1],
                        n[2], ws.periodic, _P(ws.cell_of), _P(ws.cell_count), _P(ws.cell_start),
                        _P(ws.cell_fill), _P(ws.cell_atoms), _P(ws.cell_pos), _P(ws.posc),
                        _P(ws.molc), _P(ws.flags), self._stream())
            self._call(ws, "tm_cell_build", key, cell_args)
        self._call(ws, "tm_nl_build", key, lambda: (
            ws.N, _P(pos), _P(ws.elem), _P(ws.mol), _P(ws.mol_start), _P(ws.box),
            *self._grid_args(ws), ws.periodic, self.nele, self.model.sym.rr, self.s_r, self.s_a,
            self.ldr, self.lda, _P(ws.nbr), _P(ws.ncnt), _P(ws.rseg), _P(ws.ang), _P(ws.acnt),
            _P(ws.aseg), None, _P(ws.flags), self._stream()))   # rev: filled by tm_sym_forward

    def check_flags(self, ws):
        """Read the status flags (a synchronisation) and raise on any overflow
        or non-finite coordinate; returns them otherwise."""
        f = ws.flags.cpu().numpy()
        if f[3]:
            raise FloatingPointError("non-finite coordinates")
        if f[0] & 4:
            raise RuntimeError(f"Coulomb sweep needed more candidate ranges than it holds "
                               f"(flags {f[:4]}); use larger cells")
        if f[0] or f[1]:
            raise NeighbourOverflow(f"neighbour row overflow (flags {f[:4]}); raise ldr/lda")
        return f

    def descriptors(self, ws, pos):
        sym = self.model.sym
        self._call(ws, "tm_sym_forward", (pos.data_ptr(), self._raw_stream()), lambda: (
            ws.N, _P(pos), _P(ws.box), _P(ws.mol), _P(ws.elem), _P(ws.nbr), _P(ws.ncnt),
            _P(ws.rseg), _P(ws.ang), _P(ws.acnt), _P(ws.aseg), self.ldr, self.lda, self.nele,
            sym.nr_r, sym.eta, sym.zeta, sym.rr, sym.ra, _lib.fptr(self.sym_rs_r),
            _lib.fptr(self.sym_rs_a), _lib.fptr(self.sym_th), _P(ws.G), self.D, _P(ws.rev),
            _P(ws.flags), self._stream()))
        ws.rev_ok = True

    def networks(self, ws):
        if self.precision == "fp32":
            st = self._stream()
            hid = _lib.fptr(self.hid)
            for tag, y, J in (("E", ws.eps, ws.JE), ("Q", ws.qraw, ws.JQ)):
                for a in range(self.nele):
                    M = int(ws.eoff[a + 1] - ws.eoff[a])
                    if M == 0:
                        continue
                    perm = ws.perm[int(ws.eoff[a]):]
                    self.lib.tm_mlp_f32(M, self.D, len(self.hid), hid, _P(ws.G), self.D, _P(perm),
                                        _P(self.w_f32[(tag, a)]), float(self.model.alpha), _P(y),
                                        _P(J), _P(ws.mlp_ws), st)
        else:
            self.mfma.run(ws)

    def _cl_table(self):
        """(device tables, 1/h, off/h) of the Coulomb kernel and the
        element-pair vdW energies (float4 cubic pieces, built on the host in
        float64) for the current AddEcc setting."""
        add_ecc = int(bool(PARAMS["AddEcc"]))
        t = self._cl_tables.get(add_ecc)
        if t is None:
            ntab = 1 + self.nele * (self.nele + 1) // 2
            host = np.zeros(4 * ntab * self.cl_nint, dtype=np.float32)
            scale = np.zeros(2, dtype=np.float64)
            self.lib.tm_coulomb_table(self.dsf.ctypes.data_as(_lib.P), self.model.ee_off,
                                      self.vdw_width, add_ecc, int(self.model.has_vdw), self.nele,
                                      self.c6.ctypes.data_as(_lib.P),
                                      self.r0.ctypes.data_as(_lib.P), self.cl_nint,
                                      host.ctypes.data_as(_lib.P), scale.ctypes.data_as(_lib.P))
            t = (torch.tensor(host, device=self.device), float(scale[0]), float(scale[1]))
            self._cl_tables[add_ecc] = t
        return t

    def electrostatics(self, ws, pos, timer=None):
        """Charges (neutral per molecule), the Coulomb / vdW sweep and the
        molecule means of the potential.  timer: optional pair of callables run
        on the host just before and just after the sweep's launch is queued
        (to record timing events on the launch stream)."""
        key = (pos.data_ptr(), self._raw_stream())
        st = self._stream
        # One system: the mean of the raw charges is formed inside the charges
        # launch (each block sums it).  A batch of many small molecules takes
        # one block per molecule in a reduction launch first instead, since a
        # block would otherwise step through every molecule it touches.
        ms = None
        if ws.B == 1:
            ms = ws.mol_start
        else:
            self._call(ws, "tm_mol_sum_f32", key, lambda: (
                ws.B, _P(ws.mol_start), 1, _P(ws.qraw), 1, _P(ws.qmean), 1, _P(ws.red), ws.N, st()))
        if ws.use_cells:   # charges written straight into cell order for the sweep
            self._call(ws, "tm_charges", key, lambda: (
                ws.N, _P(ws.mol), _P(ms), _P(ws.qmean), _P(ws.qraw), _P(pos), None, _P(ws.q),
                _P(ws.qx), _P(ws.cell_pos), _P(ws.elem), _P(ws.xqc), _P(ws.elc), st()))
        else:
            self._call(ws, "tm_charges", key, lambda: (
                ws.N, _P(ws.mol), _P(ms), _P(ws.qmean), _P(ws.qraw), _P(pos), _P(ws.xq), _P(ws.q),
                _P(ws.qx), None, None, None, None, st()))
        table = self._cl_table() if (ws.use_cells and self.coulomb_tables) else None
        if timer is not None:
            timer[0]()
        if table is not None:
            def tab_args():
                g = self._grid_args(ws)
                return (ws.N, _P(ws.box), g[0], g[1], g[2], g[6], g[7], g[8], g[9], g[10], g[11],
                        g[12], g[13], g[14], ws.periodic, self.model.ee_off, self.nele,
                        _P(table[0]), self.cl_nint, table[1], table[2], _P(ws.xqc), _P(ws.elc),
                        _P(ws.phi), _P(ws.gradc), _P(ws.eatom), _P(ws.flags), st())
            self._call(ws, "tm_coulomb_tab", key + (table[0].data_ptr(),), tab_args)
        else:
            add_ecc = int(bool(PARAMS["AddEcc"]))

            def coul_args():
                g = self._grid_args(ws)
                return (ws.N, None if ws.use_cells else _P(ws.xq), _P(ws.elem), _P(ws.mol),
                        _P(ws.mol_start), _P(ws.box), g[0], g[1], g[2], g[3], g[5], g[6], g[7],
                        g[8], g[9], g[10], g[11], g[12], g[13], g[14], ws.periodic,
                        self.dsf.ctypes.data_as(_lib.P), self.model.ee_off, self.vdw_width,
                        add_ecc, int(self.model.has_vdw), self.nele,
                        self.c6.ctypes.data_as(_lib.P), self.r0.ctypes.data_as(_lib.P),
                        _P(ws.xqc), _P(ws.elc), _P(ws.phi), _P(ws.gradc), _P(ws.eatom),
                        _P(ws.flags), st())
            self._call(ws, "tm_coulomb", key + (add_ecc,), coul_args)
        if timer is not None:
            timer[1]()
        self._call(ws, "tm_mol_sum_f64", key, lambda: (
            ws.B, _P(ws.mol_start), 1, _P(ws.phi), 1, _P(ws.phimean), 1, _P(ws.red), ws.N, st()))

    def _need_rev(self, ws, pos):
        """The force gather reads the mirror slots of the current lists, which
        the descriptor forward fills."""
        if not getattr(ws, "rev_ok", False):
            self.descriptors(ws, pos)

    def adjoint(self, ws, pos, vv=None):
        """Descriptor adjoint and force gather.  vv = (dt, m, v, a, ke, x)
        (device float64 arrays, ke may be None) also advances a Velocity Verlet
        state in the same launch as the gather: the kick of this step with the
        new forces, then the drift of the next (x is updated in place)."""
        self._need_rev(ws, pos)
        key = (pos.data_ptr(), self._raw_stream())
        st = self._stream
        sym = self.model.sym
        self._call(ws, "tm_sym_adjoint#energy", key, lambda: (
            ws.N, _P(pos), _P(ws.box), _P(ws.mol), _P(ws.elem), _P(ws.nbr), _P(ws.ncnt),
            _P(ws.rseg), _P(ws.ang), _P(ws.acnt), _P(ws.aseg), self.ldr, self.lda, self.nele,
            sym.nr_r, sym.eta, sym.zeta, sym.rr, sym.ra, _lib.fptr(self.sym_rs_r),
            _lib.fptr(self.sym_rs_a), _lib.fptr(self.sym_th), _P(ws.JE), _P(ws.JQ), None,
            _P(ws.phi), _P(ws.phimean), self.D, _P(ws.wslot), st()))
        if vv is None:
            self._call(ws, "tm_force_gather#energy", key, lambda: (
                ws.N, self.ldr, _P(ws.ncnt), _P(ws.rev), _P(ws.wslot), _P(ws.gradc),
                JOULEPERHARTREE, _P(ws.grad), _P(ws.force), st()))
        else:
            dt, m, v, a, ke, x = vv
            vkey = key + (float(dt),) + tuple(0 if t is None else t.data_ptr() for t in (m, v, a, ke, x))
            self._call(ws, "tm_force_gather_vv", vkey, lambda: (
                ws.N, self.ldr, _P(ws.ncnt), _P(ws.rev), _P(ws.wslot), _P(ws.gradc),
                JOULEPERHARTREE, _P(ws.grad), _P(ws.force), float(dt), _P(m), _P(v), _P(a),
                _P(ke), _P(x), st()))

    def descriptor_jvp(self, ws, pos, du, out):
        """out[:N, :D] = dG/dx . du (float32) for per-atom displacements du
        (cuda float64 [N, 3]); lists and positions of the last build."""
        sym = self.model.sym
        assert du.dtype == torch.float64 and du.is_contiguous() and du.shape[0] >= ws.N
        assert out.dtype == torch.float32 and out.is_contiguous() and out.shape[1] == self.D
        self.lib.tm_sym_jvp(ws.N, _P(pos), _P(ws.box), _P(ws.mol), _P(ws.elem), _P(ws.nbr),
                            _P(ws.ncnt), _P(ws.rseg), _P(ws.ang), _P(ws.acnt), _P(ws.aseg),
                            self.ldr, self.lda, self.nele, sym.nr_r, sym.eta, sym.zeta, sym.rr,
                            sym.ra, _lib.fptr(self.sym_rs_r), _lib.fptr(self.sym_rs_a),
                            _lib.fptr(self.sym_th), _P(du), _P(out), self.D, self._stream())

    def descriptor_vjp(self, ws, pos, gE, grad):
        """grad[:N] = (dG/dx)^T gE: the per-atom gradient of sum_i gE_i . G_i
        (gE cuda float32 [N, D]); the BP-only adjoint (no charge branch)."""
        self._need_rev(ws, pos)
        L, st = self.lib, self._stream()
        sym = self.model.sym
        assert gE.dtype == torch.float32 and gE.is_contiguous() and gE.shape[1] == self.D
        L.tm_sym_adjoint(ws.N, _P(pos), _P(ws.box), _P(ws.mol), _P(ws.elem), _P(ws.nbr),
                         _P(ws.ncnt), _P(ws.rseg), _P(ws.ang), _P(ws.acnt), _P(ws.aseg), self.ldr,
                         self.lda, self.nele, sym.nr_r, sym.eta, sym.zeta, sym.rr, sym.ra,
                         _lib.fptr(self.sym_rs_r), _lib.fptr(self.sym_rs_a), _lib.fptr(self.sym_th),
                         _P(gE), None, None, None, None, self.D, _P(ws.wslot), st)
        L.tm_force_gather(ws.N, self.ldr, _P(ws.ncnt), _P(ws.rev), _P(ws.wslot), None, 1.0,
                          _P(grad), None, st)

    def reductions(self, ws):
        """Per-molecule BP energy, Coulomb/vdW energies and dipole in one launch."""
        key = (ws.eps.data_ptr(), ws.eatom.data_ptr(), ws.qx.data_ptr(), ws.Ebp.data_ptr(),
               ws.Eext.data_ptr(), ws.dipole.data_ptr())
        arr = getattr(ws, "_sums_args", None)
        if arr is None or arr[0] != key:
            c = _lib.ctypes
            srcs = ((ws.eps, 0, 1, 1, ws.Ebp), (ws.eatom, 1, 2, 2, ws.Eext), (ws.qx, 1, 3, 3, ws.dipole))
            arr = (key, (c.c_void_p * 3)(*[t.data_ptr() for t, *_ in srcs]),
                   (c.c_int * 3)(*[f for _, f, _, _, _ in srcs]),
                   (c.c_int * 3)(*[k for _, _, k, _, _ in srcs]),
                   (c.c_int * 3)(*[ld for _, _, _, ld, _ in srcs]),
                   (c.c_void_p * 3)(*[o.data_ptr() for *_, o in srcs]))
            ws._sums_args = arr
        _, ins, f64, cols, ld, outs = arr
        self.lib.tm_mol_sums(ws.B, _P(ws.mol_start), ws.N, 3, ins, f64, cols, ld, outs, _P(ws.red),
                             self._stream())

    def run(self, ws, pos, forces=True, check=None):
        """Enqueue one full evaluation for positions pos (cuda float64 [N,3]).
        With check (default: the engine's check_overflow) the status flags are
        read once the evaluation is queued (a synchronisation), so an overflow
        raised by any stage (rows, mirror slots, the Coulomb sweep's ranges)
        or non-finite coordinates raise here."""
        assert pos.dtype == torch.float64 and pos.is_contiguous() and pos.shape[0] >= ws.N
        self.build_lists(ws, pos)
        self.descriptors(ws, pos)
        self.networks(ws)
        self.electrostatics(ws, pos)
        if forces:
            self.adjoint(ws, pos)
        self.reductions(ws)
        if self.check_overflow if check is None else check:
            self.check_flags(ws)

    def results(self, ws):
        """Host copies of the outputs of the last run (float64 numpy)."""
        Ebp = ws.Ebp[:ws.B].cpu().numpy()
        Eext = ws.Eext[:ws.B].cpu().numpy()
        return dict(Etotal=Ebp + Eext[:, 0] + Eext[:, 1], Ebp=Ebp, Ecc=Eext[:, 0],
                    Evdw=Eext[:, 1], Ebp_atom=ws.eps.double().cpu().numpy(),
                    charge=ws.q[:ws.N].cpu().numpy(), dipole=ws.dipole[:ws.B].cpu().numpy(),
                    gradient=ws.grad[:ws.N].cpu().numpy(), force=ws.force[:ws.N].cpu().numpy(),
                    G=ws.G[:ws.N].cpu().numpy())

    def energy_tensor(self, ws):
        """Per-molecule total energy as a device tensor (no host copy)."""
        return ws.Ebp[:ws.B] + ws.Eext[:ws.B, 0] + ws.Eext[:ws.B, 1]


class NeighbourOverflow(RuntimeError):
    """A neighbour row (4.6 A list or 3.1 A sub-list) exceeded its stride, so
    the rows were truncated and the evaluation is not valid."""


def _plain_upload(device):
    def up(a, out=None):
        t = torch.from_numpy(np.ascontiguousarray(a))
        if out is None:
            return t.to(device)
        out.copy_(t)
        return out
    return up


def ctypes_stream(s):
    import ctypes
    return ctypes.c_void_p(s.cuda_stream)
