"""Weights-only import of TensorFlow 1.x checkpoints of TensorMol BP+EE networks.

The reference saves a trained network as a TF checkpoint (TFMolInstance.py:
98-124, TFInstance.py:246-259) next to a pickled instance object.  The pickle
is never loaded here.  This module reads the checkpoint's tensor bundle, the
format of `tf.train.Saver` (V2):

  <prefix>.index                 an SSTable (the LevelDB table layout): data
                                 blocks of prefix-compressed (key, value)
                                 entries, an index block pointing at them and a
                                 48-byte footer ending in the table magic
                                 number.  Key "" holds a BundleHeaderProto
                                 (num_shards, endianness); every other key is a
                                 tensor name whose value is a BundleEntryProto
                                 (dtype, shape, shard_id, offset, size, crc32c).
  <prefix>.data-SSSSS-of-NNNNN   raw little-endian tensor bytes, one file per
                                 shard.

and maps the variables of the BP+EE graph (TFMolInstanceDirect.py:5179-5201
for the energy nets, 5235-5264 for the charge nets) onto BPEEModel:

  EnergyNet/<Z>_hidden1/{weights, biaseslayer0}
  EnergyNet/<Z>_hidden<k>/{weights, biaseslayer<k-1>}        k = 2 .. L
  EnergyNet/<Z>_regression_linear/{weights, biases}
  DipoleNet/<Z>_hidden1_charge/{weights, biases}
  DipoleNet/<Z>_hidden<k>_charge/{weights, biases}
  DipoleNet/<Z>_regression_linear_charge/{weights, biases}

Weights are [fan_in, fan_out] in both, as BPEEModel stores them.  Settings that
live only in the pickled instance (activation sharpness, DSF alpha, vdW switch)
are arguments of model_from_checkpoint.  Optimizer slots and other variables
are ignored.
"""
import os
import struct

import numpy as np

TABLE_MAGIC = 0xDB4775248B80FB57
FOOTER_LEN = 48
# DataType enum values of the float types a BP+EE checkpoint holds
_DTYPES = {1: np.dtype("<f4"), 2: np.dtype("<f8"), 3: np.dtype("<i4"), 9: np.dtype("<i8")}

# ---------------------------------------------------------------- CRC32C
_CRC_TABLE = []
for _n in range(256):
    _c = _n
    for _ in range(8):
        _c = (_c >> 1) ^ 0x82F63B78 if _c & 1 else _c >> 1
    _CRC_TABLE.append(_c)


def crc32c(data):
    """CRC-32C (Castagnoli polynomial, reflected), as used by the bundle."""
    crc = 0xFFFFFFFF
    tab = _CRC_TABLE
    for byte in bytes(data):
        crc = tab[(crc ^ byte) & 0xFF] ^ (crc >> 8)
    return crc ^ 0xFFFFFFFF


def crc32c_mask(crc):
    """The masked form a LevelDB-family table stores next to data."""
    return ((((crc >> 15) | (crc << 17)) & 0xFFFFFFFF) + 0xA282EAD8) & 0xFFFFFFFF


# ---------------------------------------------------------------- wire formats
def _varint(buf, pos):
    """(value, next position) of the base-128 varint at pos."""
    shift = 0
    val = 0
    while True:
        if pos >= len(buf):
            raise ValueError("truncated varint")
        b = buf[pos]
        pos += 1
        val |= (b & 0x7F) << shift
        if not b & 0x80:
            return val, pos
        shift += 7
        if shift > 63:
            raise ValueError("varint longer than 64 bits")


def _proto_fields(buf):
    """Yield (field number, wire type, value) of a serialized protobuf message;
    length-delimited values come back as bytes, fixed32/64 as ints."""
    pos = 0
    while pos < len(buf):
        key, pos = _varint(buf, pos)
        field, wt = key >> 3, key & 7
        if wt == 0:
            val, pos = _varint(buf, pos)
        elif wt == 1:
            val = struct.unpack_from("<Q", buf, pos)[0]
            pos += 8
        elif wt == 2:
            n, pos = _varint(buf, pos)
            val = bytes(buf[pos:pos + n])
            if len(val) != n:
                raise ValueError("truncated length-delimited field")
            pos += n
        elif wt == 5:
            val = struct.unpack_from("<I", buf, pos)[0]
            pos += 4
        else:
            raise ValueError(f"unsupported protobuf wire type {wt}")
        yield field, wt, val


def _block(buf, offset, size):
    """(key, value) entries of the table block at [offset, offset + size); the
    block is followed by a 1-byte compression type and a 4-byte masked CRC."""
    if offset + size + 5 > len(buf):
        raise ValueError("block handle points past the end of the index")
    ctype = buf[offset + size]
    if ctype != 0:
        raise ValueError(f"compressed table block (type {ctype}) is not supported")
    data = buf[offset:offset + size]
    stored = struct.unpack_from("<I", buf, offset + size + 1)[0]
    if crc32c_mask(crc32c(buf[offset:offset + size + 1])) != stored:
        raise ValueError("table block checksum mismatch")
    nres = struct.unpack_from("<I", data, size - 4)[0]
    end = size - 4 - 4 * nres
    if end < 0:
        raise ValueError("bad restart array")
    pos, key = 0, b""
    out = []
    while pos < end:
        shared, pos = _varint(data, pos)
        nonshared, pos = _varint(data, pos)
        vlen, pos = _varint(data, pos)
        key = key[:shared] + bytes(data[pos:pos + nonshared])
        pos += nonshared
        out.append((key, bytes(data[pos:pos + vlen])))
        pos += vlen
    return out


def _handle(buf, pos=0):
    off, pos = _varint(buf, pos)
    size, pos = _varint(buf, pos)
    return off, size, pos


def read_index(path):
    """{key: value bytes} of an SSTable file (all data blocks)."""
    with open(path, "rb") as f:
        buf = f.read()
    if len(buf) < FOOTER_LEN:
        raise ValueError(f"{path}: too short for a table footer")
    foot = buf[-FOOTER_LEN:]
    magic = struct.unpack_from("<Q", foot, FOOTER_LEN - 8)[0]
    if magic != TABLE_MAGIC:
        raise ValueError(f"{path}: not a table file (magic {magic:#x})")
    _, _, p = _handle(foot)              # metaindex handle (unused)
    ioff, isize, _ = _handle(foot, p)    # index block handle
    entries = {}
    for _, hv in _block(buf, ioff, isize):
        doff, dsize, _ = _handle(hv)
        for k, v in _block(buf, doff, dsize):
            entries[k] = v
    return entries


def _entry(value):
    """BundleEntryProto -> dict(dtype, shape, shard, offset, size, crc)."""
    e = dict(dtype=0, shape=(), shard=0, offset=0, size=0, crc=None, sliced=False)
    for field, _, val in _proto_fields(value):
        if field == 1:
            e["dtype"] = val
        elif field == 2:
            dims = []
            for f2, _, v2 in _proto_fields(val):
                if f2 == 2:   # TensorShapeProto.dim
                    size = 0
                    for f3, _, v3 in _proto_fields(v2):
                        if f3 == 1:
                            size = v3 - (1 << 64) if v3 >= 1 << 63 else v3
                    dims.append(size)
            e["shape"] = tuple(dims)
        elif field == 3:
            e["shard"] = val
        elif field == 4:
            e["offset"] = val
        elif field == 5:
            e["size"] = val
        elif field == 6:
            e["crc"] = val
        elif field == 7:
            e["sliced"] = True
    return e


def read_bundle(prefix, names=None, verify=True):
    """{tensor name: numpy array} of a V2 checkpoint (float32 / float64 /
    int32 / int64 tensors; others are skipped).  names: optional iterable
    restricting which tensors are read.  verify: check each tensor's CRC-32C
    (stored plain or masked)."""
    idx = read_index(prefix + ".index")
    header = idx.pop(b"", None)
    nshards, big = 1, False
    if header is not None:
        for field, _, val in _proto_fields(header):
            if field == 1:
                nshards = val
            elif field == 2:
                big = val == 1
    if big:
        raise ValueError("big-endian bundles are not supported")
    want = None if names is None else set(names)
    files = {}
    out = {}
    try:
        for key, value in idx.items():
            name = key.decode()
            if want is not None and name not in want:
                continue
            e = _entry(value)
            dt = _DTYPES.get(e["dtype"])
            if dt is None or e["sliced"]:
                continue
            n = int(np.prod(e["shape"])) if e["shape"] else 1
            if n * dt.itemsize != e["size"]:
                raise ValueError(f"{name}: {e['size']} bytes for shape {e['shape']} of {dt}")
            f = files.get(e["shard"])
            if f is None:
                path = f"{prefix}.data-{e['shard']:05d}-of-{nshards:05d}"
                f = files[e["shard"]] = open(path, "rb")
            f.seek(e["offset"])
            raw = f.read(e["size"])
            if len(raw) != e["size"]:
                raise ValueError(f"{name}: data shard ends early")
            if verify and e["crc"] is not None:
                c = crc32c(raw)
                if e["crc"] not in (c, crc32c_mask(c)):
                    raise ValueError(f"{name}: checksum mismatch")
            out[name] = np.frombuffer(raw, dtype=dt).reshape(e["shape"]).copy()
    finally:
        for f in files.values():
            f.close()
    return out


# ---------------------------------------------------------------- BP+EE mapping
def bpee_variable_names(Z, nhidden):
    """[(weights name, biases name)] per layer, energy net then charge net, for
    element Z with nhidden hidden layers (TFMolInstanceDirect.py:5179-5264)."""
    E = [(f"EnergyNet/{Z}_hidden{k + 1}/weights", f"EnergyNet/{Z}_hidden{k + 1}/biaseslayer{k}")
         for k in range(nhidden)]
    E.append((f"EnergyNet/{Z}_regression_linear/weights", f"EnergyNet/{Z}_regression_linear/biases"))
    Q = [(f"DipoleNet/{Z}_hidden{k + 1}_charge/weights", f"DipoleNet/{Z}_hidden{k + 1}_charge/biases")
         for k in range(nhidden)]
    Q.append((f"DipoleNet/{Z}_regression_linear_charge/weights",
              f"DipoleNet/{Z}_regression_linear_charge/biases"))
    return E, Q


def find_last_checkpoint(train_dir, name):
    """Prefix of the highest-step `<name>-chk-<step>` bundle in train_dir, or
    None (the reference's FindLastCheckpoint, TFInstance.py:251-258, which
    looks for the `.meta` graph file; the bundle's `.index` is what is read)."""
    if not os.path.isdir(train_dir):
        return None
    head = name + "-chk-"
    steps = []
    for f in os.listdir(train_dir):
        if f.startswith(head) and f.endswith(".index") and f[len(head):-6].isdigit():
            steps.append(int(f[len(head):-6]))
    return os.path.join(train_dir, f"{head}{max(steps)}") if steps else None


def checkpoint_elements(names):
    """Elements with an energy net among the variable names, ascending."""
    out = set()
    for n in names:
        if n.startswith("EnergyNet/") and n.endswith("_hidden1/weights"):
            z = n[len("EnergyNet/"):-len("_hidden1/weights")]
            if z.isdigit():
                out.add(int(z))
    return sorted(out)


def model_from_checkpoint(prefix, eles=None, **model_kwargs):
    """BPEEModel whose energy and charge nets are the checkpoint's variables.
    eles: the network's elements (the pickled instance's `eles`; default: the
    elements that have an energy net in the checkpoint, ascending); model_kwargs
    go to BPEEModel (activation_alpha, dsf_alpha, has_vdw, ...).  The hidden
    layer sizes are read from the weights; the descriptor width must match the
    one BPEEModel derives from eles."""
    from .model import BPEEModel
    vars_ = read_bundle(prefix)
    eles = checkpoint_elements(vars_) if eles is None else [int(e) for e in eles]
    if not eles:
        raise KeyError(f"{prefix}: no EnergyNet/<Z>_hidden1/weights variable")
    nh = 0
    while f"EnergyNet/{eles[0]}_hidden{nh + 1}/weights" in vars_:
        nh += 1
    if nh == 0:
        raise KeyError(f"{prefix}: no EnergyNet/{eles[0]}_hidden1/weights variable")
    energy, charge = {}, {}
    for Z in eles:
        for nets, names in zip((energy, charge), bpee_variable_names(Z, nh)):
            layers = []
            for wn, bn in names:
                if wn not in vars_ or bn not in vars_:
                    raise KeyError(f"{prefix}: missing variable {wn if wn not in vars_ else bn}")
                W = np.asarray(vars_[wn], dtype=np.float64)
                b = np.asarray(vars_[bn], dtype=np.float64).reshape(-1)
                if W.ndim != 2 or W.shape[1] != b.shape[0]:
                    raise ValueError(f"{wn}: shape {W.shape} does not match {bn} {b.shape}")
                layers.append((W, b))
            nets[Z] = layers
    hidden = [energy[eles[0]][k][0].shape[1] for k in range(nh)]
    model = BPEEModel(eles, hidden, energy_nets=energy, charge_nets=charge, **model_kwargs)
    D = model.sym.D
    for Z in eles:
        for nets in (energy, charge):
            if nets[Z][0][0].shape[0] != D:
                raise ValueError(f"element {Z}: first layer takes {nets[Z][0][0].shape[0]} inputs, "
                                 f"the descriptor has {D}")
            for (W0, _), (W1, _) in zip(nets[Z][:-1], nets[Z][1:]):
                if W0.shape[1] != W1.shape[0]:
                    raise ValueError(f"element {Z}: layer sizes do not chain")
    return model
