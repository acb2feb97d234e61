"""HDF5 files of a simulation without h5py: the layout h5py writes by default.

The reference saves every result through ``SimArchivist.save_h5(fname, keyz, valz)``
(simulation_utilities/sim_archive.py:13-24): one ``h5py.File(fname, 'w')`` with one
``create_dataset(key, data=val)`` per key.  Its analysis tools read them back
(``SimInfo``, analysis/extract_sim_info.py:29-133):
* ``{sim_name}_sim_info.hdf5``: vref_vs_tau, pop_vs_tau, atomic_nums, atomic_masses
  (pyvibdmc.py:918-922);
* ``wfns/{sim_name}_wfn_{t}ts.hdf5``: coords, desc_wts [, parent_wts] (pyvibdmc.py:861-869);
* ``{sim_name}_training_{t}ts.hdf5``: coords, pots (pyvibdmc.py:800-804, 840-844).

``write`` produces the file h5py's defaults produce for such a call, byte for byte (checked
against the reference's sample data): superblock version 0, a root group with a symbol
table (a version-1 B-tree leaf, a local heap, one symbol-table node), version-1 object
headers of 256 message bytes (dataspace, datatype, fill value, contiguous layout,
modification time, padding), and the library's allocation order: metadata packed into
2048-byte blocks, raw data of 2048 bytes or more at the end of the file, smaller raw data
in a 2048-byte small-data block, the file cut at its last used byte.

``read`` reads that layout (and the same layout with several B-tree levels or compact
storage) into a dict of NumPy arrays; it raises on anything else.
"""
import struct
import time

import numpy as np

__all__ = ['write', 'read', 'save_h5']

UNDEF = 0xFFFFFFFFFFFFFFFF
SIGNATURE = b'\x89HDF\r\n\x1a\n'
META_BLOCK = 2048     # metadata aggregation block
SMALL_BLOCK = 2048    # small raw-data aggregation block
OHDR_DATA = 256       # message bytes of a dataset object header
GROUP_LEAF_K = 4      # symbol-table node holds 2K entries
GROUP_INTERNAL_K = 16
HEAP_DATA = 88        # initial local-heap data segment
BTREE_NODE = 24 + (2 * GROUP_INTERNAL_K + 1) * 8 + 2 * GROUP_INTERNAL_K * 8
SNOD_SIZE = 8 + 2 * GROUP_LEAF_K * 40


def _pad8(n):
    return (n + 7) // 8 * 8


def _datatype(dt):
    """Datatype message body (class and version, bit field, size, properties), padded."""
    dt = np.dtype(dt)
    if dt.kind == 'f' and dt.itemsize == 8:
        body = bytes([0x11, 0x20, 0x3F, 0x00]) + struct.pack('<IHHBBBBI', 8, 0, 64, 52, 11, 0, 52, 1023)
    elif dt.kind == 'f' and dt.itemsize == 4:
        body = bytes([0x11, 0x20, 0x1F, 0x00]) + struct.pack('<IHHBBBBI', 4, 0, 32, 23, 8, 0, 23, 127)
    elif dt.kind in 'iu':
        signed = 0x08 if dt.kind == 'i' else 0x00
        body = bytes([0x10, signed, 0x00, 0x00]) + struct.pack('<IHH', dt.itemsize, 0, 8 * dt.itemsize)
    else:
        raise TypeError(f'unsupported dtype {dt}')
    return body + b'\0' * (_pad8(len(body)) - len(body))


def _message(mtype, body, flags=0):
    return struct.pack('<HHB3x', mtype, len(body), flags) + body


class _Allocator:
    """The allocation order of the library's metadata and small-data aggregators."""

    def __init__(self):
        self.eoa = META_BLOCK
        self.meta = [96, META_BLOCK]        # superblock occupies [0, 96) of the first block
        self.small = None
        self.end = 96

    def meta_alloc(self, size):
        if self.meta[0] + size > self.meta[1]:
            self.meta = [self.eoa, self.eoa + META_BLOCK]
            self.eoa += META_BLOCK
        a = self.meta[0]
        self.meta[0] += size
        self.end = max(self.end, a + size)
        return a

    def raw_alloc(self, size):
        if size == 0:
            return UNDEF
        if size >= SMALL_BLOCK:
            a = self.eoa
            self.eoa += size
        else:
            if self.small is None or self.small[0] + size > self.small[1]:
                self.small = [self.eoa, self.eoa + SMALL_BLOCK]
                self.eoa += SMALL_BLOCK
            a = self.small[0]
            self.small[0] += size
        self.end = max(self.end, a + size)
        return a


def write(fname, keys, values, mtime=None):
    """Write datasets ``keys[i] = values[i]`` (NumPy arrays of float64/32 or integers) in the
    order given, as ``h5py.File(fname, 'w')`` + ``create_dataset`` would.  ``mtime``: the
    modification time stored in each object header (seconds; default: now)."""
    keys = [str(k) for k in keys]
    arrays = [np.ascontiguousarray(np.asarray(v)) for v in values]
    if len(keys) != len(arrays) or len(set(keys)) != len(keys):
        raise ValueError('keys must be unique and match the values')
    if not 0 < len(keys) <= 2 * GROUP_LEAF_K:
        raise ValueError(f'between 1 and {2 * GROUP_LEAF_K} datasets per file')
    mt = int(time.time()) if mtime is None else int(mtime)
    # local heap: "" then each name, NUL-terminated and padded to 8 bytes, in insertion order
    heap = bytearray(8)
    name_off = {}
    for k in keys:
        name_off[k] = len(heap)
        enc = k.encode() + b'\0'
        heap += enc + b'\0' * (_pad8(len(enc)) - len(enc))
    rem = HEAP_DATA - len(heap)
    if rem < 0 or rem == 8:  # a free block needs 16 bytes
        raise ValueError('dataset names too long for the initial local heap')
    free_off = len(heap) if rem else UNDEF
    if free_off != UNDEF:
        heap += struct.pack('<QQ', 1, HEAP_DATA - free_off)
    heap += b'\0' * (HEAP_DATA - len(heap))

    al = _Allocator()
    root_oh = al.meta_alloc(16 + 24)
    btree = al.meta_alloc(BTREE_NODE)
    heap_hdr = al.meta_alloc(32 + HEAP_DATA)
    snod = None
    headers, raws = {}, {}
    for k, a in zip(keys, arrays):
        headers[k] = al.meta_alloc(16 + OHDR_DATA)
        if snod is None:
            snod = al.meta_alloc(SNOD_SIZE)
        raws[k] = al.raw_alloc(a.nbytes)
    out = bytearray(al.end)

    # superblock version 0 and the root group's symbol-table entry
    out[0:96] = (SIGNATURE + bytes([0, 0, 0, 0, 0, 8, 8, 0]) +
                 struct.pack('<HHI', GROUP_LEAF_K, GROUP_INTERNAL_K, 0) +
                 struct.pack('<QQQQ', 0, UNDEF, al.end, UNDEF) +
                 struct.pack('<QQII', 0, root_oh, 1, 0) + struct.pack('<QQ', btree, heap_hdr))
    # root object header: one symbol-table message
    out[root_oh:root_oh + 40] = (struct.pack('<BBHII4x', 1, 0, 1, 1, 24) +
                                 _message(0x11, struct.pack('<QQ', btree, heap_hdr)))
    # B-tree leaf: key "" , child SNOD, key = the largest name
    last = max(keys)
    node = b'TREE' + struct.pack('<BBHQQ', 0, 0, 1, UNDEF, UNDEF) + struct.pack('<QQQ', 0, snod, name_off[last])
    out[btree:btree + len(node)] = node
    out[heap_hdr:heap_hdr + 32] = b'HEAP' + struct.pack('<B3xQQQ', 0, HEAP_DATA, free_off, heap_hdr + 32)
    out[heap_hdr + 32:heap_hdr + 32 + HEAP_DATA] = heap
    # symbol-table node: entries sorted by name
    ents = b''.join(struct.pack('<QQII16x', name_off[k], headers[k], 0, 0) for k in sorted(keys))
    out[snod:snod + 8 + len(ents)] = b'SNOD' + struct.pack('<BBH', 1, 0, len(keys)) + ents
    # dataset object headers and raw data
    for k, a in zip(keys, arrays):
        dims = a.shape
        space = struct.pack('<BBBB4x', 1, len(dims), 1, 0) + struct.pack(f'<{2 * len(dims)}Q', *dims, *dims)
        layout = struct.pack('<BBQQ', 3, 1, raws[k], a.nbytes)
        msgs = (_message(0x01, space) + _message(0x03, _datatype(a.dtype), 1) +
                _message(0x05, bytes([2, 2, 2, 1, 0, 0, 0, 0]), 1) +
                _message(0x08, layout + b'\0' * (_pad8(len(layout)) - len(layout))) +
                _message(0x12, struct.pack('<B3xI', 1, mt)))
        nil = OHDR_DATA - len(msgs) - 8
        if nil < 0:
            raise ValueError('too many dimensions for one object header')
        msgs += _message(0x00, b'\0' * nil)
        h = headers[k]
        out[h:h + 16 + OHDR_DATA] = struct.pack('<BBHII4x', 1, 0, 6, 1, OHDR_DATA) + msgs
        if a.nbytes:
            out[raws[k]:raws[k] + a.nbytes] = a.astype(a.dtype.newbyteorder('<'), copy=False).tobytes()
    with open(fname, 'wb') as fh:
        fh.write(bytes(out))


def save_h5(fname, keyz, valz):
    """SimArchivist.save_h5 (sim_archive.py:13-24)."""
    write(fname, keyz, valz)


# ------------------------------------------------------------------------------ reader
class _File:
    def __init__(self, buf):
        self.b = buf

    def u(self, fmt, off):
        return struct.unpack_from('<' + fmt, self.b, off)

    def cstr(self, off):
        end = self.b.index(b'\0', off)
        return self.b[off:end].decode()

    def messages(self, addr):
        ver = self.b[addr]
        if ver != 1:
            raise ValueError(f'object header version {ver} is not supported')
        nmsg, _, size = self.u('HII', addr + 2)
        out, queue = [], [(addr + 16, size)]
        while queue and len(out) < nmsg:
            p, size = queue.pop(0)
            end = p + size
            while p < end and len(out) < nmsg:
                t, sz, fl = self.u('HHB', p)
                body = self.b[p + 8:p + 8 + sz]
                if t == 0x10:  # continuation
                    queue.append(struct.unpack('<QQ', body[:16]))
                out.append((t, body))
                p += 8 + sz
        return out

    def group(self, btree, heap_data):
        """name -> object header address, from a version-1 group B-tree."""
        if self.b[btree:btree + 4] != b'TREE':
            raise ValueError('corrupt group B-tree')
        level, used = self.b[btree + 5], self.u('H', btree + 6)[0]
        out = {}
        for i in range(used):
            child = self.u('Q', btree + 24 + 16 * i + 8)[0]
            if level > 0:
                out.update(self.group(child, heap_data))
                continue
            if self.b[child:child + 4] != b'SNOD':
                raise ValueError('corrupt symbol-table node')
            for e in range(self.u('H', child + 6)[0]):
                lno, oh = self.u('QQ', child + 8 + 40 * e)
                out[self.cstr(heap_data + lno)] = oh
        return out

    def dataset(self, addr):
        shape = dtype = None
        data_addr = size = None
        compact = None
        for t, body in self.messages(addr):
            if t == 0x01:
                ver, rank = body[0], body[1]
                off = 8 if ver == 1 else 4
                shape = struct.unpack_from(f'<{rank}Q', body, off)
            elif t == 0x03:
                cls, bits0, size_ = body[0] & 0x0F, body[1], struct.unpack_from('<I', body, 4)[0]
                order = '>' if bits0 & 1 else '<'
                if cls == 1:
                    dtype = np.dtype(f'{order}f{size_}')
                elif cls == 0:
                    dtype = np.dtype(f'{order}{"i" if bits0 & 0x08 else "u"}{size_}')
                else:
                    raise ValueError(f'datatype class {cls} is not supported')
            elif t == 0x08:
                if body[0] != 3:
                    raise ValueError(f'layout message version {body[0]} is not supported')
                if body[1] == 1:
                    data_addr, size = struct.unpack_from('<QQ', body, 2)
                elif body[1] == 0:
                    n = struct.unpack_from('<H', body, 2)[0]
                    compact = body[4:4 + n]
                else:
                    raise ValueError('chunked datasets are not supported')
        if shape is None or dtype is None:
            raise ValueError('dataset without dataspace or datatype')
        count = int(np.prod(shape)) if len(shape) else 1
        if compact is not None:
            raw = compact
        elif data_addr == UNDEF or count == 0:
            raw = b'\0' * (count * dtype.itemsize)
        else:
            raw = self.b[data_addr:data_addr + count * dtype.itemsize]
        return np.frombuffer(raw, dtype=dtype, count=count).reshape(shape).astype(dtype.newbyteorder('='))


def read(fname):
    """Every dataset of the root group of ``fname`` as a NumPy array."""
    with open(fname, 'rb') as fh:
        f = _File(fh.read())
    if f.b[:8] != SIGNATURE:
        raise ValueError(f'{fname} is not an HDF5 file')
    if f.b[8] != 0:
        raise ValueError(f'superblock version {f.b[8]} is not supported')
    root = f.u('Q', 56 + 8)[0]
    stab = [body for t, body in f.messages(root) if t == 0x11]
    if not stab:
        raise ValueError('root group without a symbol table')
    btree, heap = struct.unpack('<QQ', stab[0][:16])
    heap_data = f.u('Q', heap + 24)[0]
    return {name: f.dataset(addr) for name, addr in sorted(f.group(btree, heap_data).items())}
