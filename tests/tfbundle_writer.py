"""Writer of TF V2 checkpoint bundles for the reader's tests: an SSTable index
(data blocks with prefix-compressed keys and restart points, an index block,
an empty metaindex block, the 48-byte footer with the table magic) and raw
data shards.  Written from the table and bundle layouts the reader documents;
test infrastructure only."""
import struct

import numpy as np

from tensormol_amd.tfbundle import TABLE_MAGIC, crc32c, crc32c_mask

_DT = {np.dtype("float32"): 1, np.dtype("float64"): 2, np.dtype("int32"): 3, np.dtype("int64"): 9}


def varint(v):
    out = bytearray()
    while True:
        b = v & 0x7F
        v >>= 7
        if v:
            out.append(b | 0x80)
        else:
            out.append(b)
            return bytes(out)


def _field(num, wt, payload):
    key = varint((num << 3) | wt)
    if wt == 0:
        return key + varint(payload)
    if wt == 2:
        return key + varint(len(payload)) + payload
    if wt == 5:
        return key + struct.pack("<I", payload)
    raise ValueError(wt)


def entry_proto(dtype, shape, shard, offset, size, crc):
    dims = b"".join(_field(2, 2, _field(1, 0, int(d))) for d in shape)
    return (_field(1, 0, dtype) + _field(2, 2, dims) + _field(3, 0, shard) + _field(4, 0, offset)
            + _field(5, 0, size) + _field(6, 5, crc))


def header_proto(nshards):
    return _field(1, 0, nshards) + _field(2, 0, 0)


class _BlockBuilder:
    def __init__(self, restart_interval=4):
        self.buf = bytearray()
        self.restarts = [0]
        self.count = 0
        self.last = b""
        self.interval = restart_interval

    def add(self, key, value):
        shared = 0
        if self.count < self.interval:
            n = min(len(key), len(self.last))
            while shared < n and key[shared] == self.last[shared]:
                shared += 1
        else:
            self.restarts.append(len(self.buf))
            self.count = 0
        self.buf += varint(shared) + varint(len(key) - shared) + varint(len(value))
        self.buf += key[shared:] + value
        self.last = key
        self.count += 1

    def finish(self):
        return bytes(self.buf) + b"".join(struct.pack("<I", r) for r in self.restarts) + \
            struct.pack("<I", len(self.restarts))


def _write_block(out, contents):
    off = len(out)
    out += contents
    ctype = b"\x00"
    out += ctype + struct.pack("<I", crc32c_mask(crc32c(contents + ctype)))
    return off, len(contents)


def write_table(path, items, block_entries=3):
    """SSTable of sorted (key bytes, value bytes) items, block_entries per data block."""
    items = sorted(items)
    out = bytearray()
    index = _BlockBuilder(restart_interval=1)
    for i in range(0, len(items), block_entries):
        b = _BlockBuilder()
        chunk = items[i:i + block_entries]
        for k, v in chunk:
            b.add(k, v)
        off, size = _write_block(out, b.finish())
        index.add(chunk[-1][0], varint(off) + varint(size))
    meta_off, meta_size = _write_block(out, _BlockBuilder().finish())
    idx_off, idx_size = _write_block(out, index.finish())
    handles = varint(meta_off) + varint(meta_size) + varint(idx_off) + varint(idx_size)
    out += handles + b"\x00" * (40 - len(handles)) + struct.pack("<Q", TABLE_MAGIC)
    with open(path, "wb") as f:
        f.write(out)


def write_bundle(prefix, tensors, nshards=1, masked_crc=True, block_entries=3):
    """Bundle of {name: numpy array}; tensors are dealt to shards round-robin."""
    shards = [bytearray() for _ in range(nshards)]
    items = [(b"", header_proto(nshards))]
    for i, (name, arr) in enumerate(sorted(tensors.items())):
        arr = np.array(arr, order="C")   # keeps 0-d tensors 0-d
        raw = arr.astype(arr.dtype.newbyteorder("<")).tobytes()
        s = i % nshards
        off = len(shards[s])
        shards[s] += raw
        crc = crc32c(raw)
        items.append((name.encode(), entry_proto(_DT[arr.dtype], arr.shape, s, off, len(raw),
                                                 crc32c_mask(crc) if masked_crc else crc)))
    write_table(prefix + ".index", items, block_entries)
    for s, data in enumerate(shards):
        with open(f"{prefix}.data-{s:05d}-of-{nshards:05d}", "wb") as f:
            f.write(bytes(data))
