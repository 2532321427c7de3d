"""Host side of the fused MFMA MLP kernel: weight packing and tile lists.

Weights of every (net, element) pair are packed once into a blob whose layout
tm_mlp_mfma_pack_info reports: six fragment arrays (forward W1^T, W2^T, W3^T;
backward W3, W2, W1) in the operand order of v_mfma_f32_16x16x32 (A operand:
lane l holds row 16*t + l%16, columns 32*s + 8*(l//16) .. +7), followed by the
float32 biases and output layer.  The split modes store a hi and a lo plane:
bf16x3 for every array, bf16x2 / fp16x2 / bf16x3f for the three forward arrays
only.

Modes (operands of v_mfma_f32_16x16x32, FP32 accumulation throughout):
  bf16     weights and activations rounded once to BF16
  fp16     the same in FP16
  bf16x3   hi/lo BF16 split of weights and activations in every GEMM (3 MFMAs)
  bf16x2   forward GEMMs with hi/lo split weights (2 MFMAs), BF16 activations;
           input-gradient GEMMs as bf16
  fp16x2   the same with FP16 operands
  bf16x3f  forward GEMMs split in weights and activations (3 MFMAs),
           input-gradient GEMMs as bf16
"""
import ctypes

import numpy as np
import torch

from . import _lib

MODES = {"bf16": 0, "fp16": 1, "bf16x3": 2, "bf16x2": 3, "fp16x2": 4, "bf16x3f": 5}
TILE_ROWS = {0: 64, 1: 64, 2: 32, 3: 64, 4: 64, 5: 64}
HALF = {0: "bf16", 1: "fp16", 2: "bf16", 3: "bf16", 4: "fp16", 5: "bf16"}   # operand type


def frag_pack(M):
    """[R, C] matrix -> A-operand fragments [R/16][C/32][64 lanes][8]."""
    R, C = M.shape
    return M.reshape(R // 16, 16, C // 32, 4, 8).transpose(0, 2, 3, 1, 4).reshape(-1)


def _to16(x, mode):
    t = torch.tensor(np.ascontiguousarray(x), dtype=torch.float32)
    if HALF[mode] == "fp16":
        h = t.to(torch.float16)
        return h.view(torch.int16).numpy(), h.to(torch.float32).numpy()
    h = t.to(torch.bfloat16)
    return h.view(torch.int16).numpy(), h.to(torch.float32).numpy()


NXCD = 8   # MI355X: workgroups are dispatched round-robin over 8 XCDs, each with its own L2


def xcd_order(til):
    """Tile table in launch order: XCD x (workgroups x, x + 8, ...) takes the
    x-th contiguous run of the weight-blob-sorted tiles, so each XCD's L2
    holds the weights of one or two (net, element) blobs instead of all of
    them.  Padding entries (segment -1) make the grid a multiple of 8; their
    workgroups exit at once."""
    n = len(til)
    if n == 0:
        return til
    bounds = (np.arange(NXCD + 1) * n) // NXCD      # near-equal contiguous runs
    c = int(np.diff(bounds).max())
    out = np.full((c * NXCD, 2), -1, dtype=np.int32)
    out[:, 1] = 0
    j = np.arange(n)
    x = np.searchsorted(bounds, j, side="right") - 1
    t = j - bounds[x]
    out[t * NXCD + x] = til
    return out


class MFMAWeights:
    def __init__(self, engine, precision):
        if precision not in MODES:
            raise ValueError(f"unknown MLP precision {precision!r}; use fp32 or one of "
                             f"{sorted(MODES)}")
        self.engine = engine
        self.mode = MODES[precision]
        self.mt = TILE_ROWS[self.mode]
        dev = engine.device
        self.ncu = (torch.cuda.get_device_properties(dev).multi_processor_count
                    if dev.type == "cuda" and torch.cuda.is_available() else 256)
        model = engine.model
        self.D = model.sym.D
        if len(model.hidden) != 3 or len(set(model.hidden)) != 1:
            raise ValueError("the MFMA path needs three equal hidden layers")
        self.H = int(model.hidden[0])
        info = (ctypes.c_long * 10)()
        engine.lib.tm_mlp_mfma_pack_info(self.D, self.H, 3, self.mode, info)
        self.info = list(info)
        self.blobs = []
        for tag, nets in (("E", model.energy_nets), ("Q", model.charge_nets)):
            for e in model.eles:
                self.blobs.append(self._blob(nets[e]))
        dev = engine.device
        self.blob_t = [torch.tensor(b, device=dev) for b in self.blobs]
        self.blob_ptrs = torch.tensor([t.data_ptr() for t in self.blob_t], dtype=torch.int64,
                                      device=dev)

    def _blob(self, layers):
        (W1, b1), (W2, b2), (W3, b3), (Wo, bo) = layers
        nbytes, fsec = self.info[0], self.info[1]
        offs = self.info[2:8]
        fwd_planes, bwd_planes = self.info[8], self.info[9]
        blob = np.zeros(nbytes, dtype=np.uint8)
        el = blob[:fsec].view(np.int16)
        for k, (off, M) in enumerate(zip(offs, (W1.T, W2.T, W3.T, W3, W2, W1))):
            planes = fwd_planes if k < 3 else bwd_planes
            hi_bits, hi_val = _to16(M, self.mode)
            packed = frag_pack(hi_bits.reshape(M.shape))
            n = packed.size
            el[off:off + n] = packed
            if planes == 2:
                lo_bits, _ = _to16(np.asarray(M, np.float64) - hi_val, self.mode)
                el[off + n:off + 2 * n] = frag_pack(lo_bits.reshape(M.shape))
        H = self.H
        fl = blob[fsec:].view(np.float32)
        fl[0:H] = b1
        fl[H:2 * H] = b2
        fl[2 * H:3 * H] = b3
        fl[3 * H:4 * H] = np.asarray(Wo).reshape(-1)
        fl[4 * H] = np.asarray(bo).reshape(-1)[0]
        return blob

    def workspace(self, ws, upload=None):
        """Segment and tile tables for one packed batch (rows are element-sorted):
        seg[s] = (net, net * nele + element, first row, rows), tiles[t] = (seg, first row).

        64-row tiles fill whole rounds of the device (one workgroup per CU); the
        remaining rows go to 32-row tiles in a second launch, so the last round
        is not half empty.  bf16x3 uses 32-row tiles throughout."""
        nele = self.engine.nele
        counts = np.diff(ws.eoff).astype(np.int64)
        segs, t64 = [], []
        for net in range(2):
            for a in range(nele):
                M = int(counts[a])
                if M == 0:
                    continue
                s = len(segs)
                segs.append((net, net * nele + a, int(ws.eoff[a]), M))
                r0 = np.arange(0, M, 64, dtype=np.int32)
                t64.append(np.stack([np.full_like(r0, s), r0], axis=1))
        seg = np.array(segs, dtype=np.int32).reshape(-1, 4)
        t64 = np.concatenate(t64).astype(np.int32) if t64 else np.zeros((0, 2), np.int32)
        full = 0 if self.mt == 32 else (len(t64) // self.ncu) * self.ncu
        head, rest = t64[:full], t64[full:]
        # each remaining 64-row tile becomes two 32-row tiles (the second only if rows remain)
        halves = np.repeat(rest, 2, axis=0)
        halves[1::2, 1] += 32
        rows = seg[halves[:, 0], 3] if len(halves) else np.zeros(0, np.int32)
        t32 = halves[halves[:, 1] < rows] if len(halves) else halves
        t64x, t32x = xcd_order(head), xcd_order(t32)
        if upload is None:
            dev = self.engine.device
            upload = lambda a, out=None: torch.from_numpy(a).to(dev)
        return dict(seg=upload(seg), tiles64=upload(t64x), n64=len(t64x),
                    tiles32=upload(t32x), n32=len(t32x))

    def run(self, ws):
        eng = self.engine
        m = ws.mlp_ws
        for rows, key, n in ((64, "tiles64", m["n64"]), (32, "tiles32", m["n32"])):
            if n == 0:
                continue
            eng.lib.tm_mlp_mfma(ws.N, self.D, self.H, n, _lib.ptr(m[key]), _lib.ptr(m["seg"]),
                                self.D, _lib.ptr(ws.G), _lib.ptr(ws.perm), _lib.ptr(self.blob_ptrs),
                                float(eng.model.alpha), self.mode, rows, _lib.ptr(ws.Y2),
                                _lib.ptr(ws.J2), eng._stream())
