"""HDF5 distribution files without h5py.

The reference writes and restarts from HDF5 files made with h5py
(dgfs1D/std/distribution.py:36-41, dgfs1D/bi/distribution.py:36-42,
restart readers dgfs1D/std/std.py:235-247, dgfs1D/bi/bi.py:222-236, and
dgfs1D/error.py:9-14): one flat dataset per species (``coeff``, or
``coeff0`` and ``coeff1``) with scalar attributes time, K, Ne and Nv.  h5py
is not part of this stack, so this module reads and writes that shape of file
from the published HDF5 file-format specification (version 3.0 of the
specification; superblock, object-header, B-tree, heap and message layouts as
named below).

Writer: the layout h5py writes by default (library format "earliest"):
superblock version 0, a root group with a symbol table (version-1 B-tree,
one symbol-table node, a local heap of the names), and per dataset a
version-1 object header with dataspace, datatype, fill-value, contiguous
layout and attribute messages, the raw data stored contiguously behind it.

Reader: files of both of h5py's formats -- superblock 0/1 with symbol-table
groups and version-1 object headers, and superblock 2/3 with version-2
object headers and compact link messages -- for contiguous or compact
datasets of IEEE floats or two's-complement integers and scalar or simple
dataspaces.  Anything else (chunked or filtered storage, dense link or
attribute storage, other types) raises ``ValueError`` naming what it found.
"""
import struct

import numpy as np

SIGNATURE = b'\x89HDF\r\n\x1a\n'
UNDEF = 0xFFFFFFFFFFFFFFFF

# message types (format spec, IV.A.2)
NIL, DATASPACE, LINKINFO, DATATYPE, FILL_OLD, FILL, LINK, LAYOUT = 0, 1, 2, 3, 4, 5, 6, 8
FILTERS, ATTRIBUTE, CONTINUATION, SYMTABLE, MTIME, ATTRINFO = 11, 12, 16, 17, 18, 21

# group leaf / internal node K of the writer (the library's defaults)
LEAF_K, INTERNAL_K = 4, 16


def _pad8(n):
    return (n + 7) & ~7


class _Reader:
    def __init__(self, data):
        self.d = data
        base = None
        for off in (0, 512, 1024, 2048, 4096, 8192):
            if data[off:off + 8] == SIGNATURE:
                base = off
                break
        if base is None:
            raise ValueError('not an HDF5 file (no signature)')
        self.sb = base
        ver = data[base + 8]
        if ver in (0, 1):
            so, sl = data[base + 13], data[base + 14]
            p = base + 24 + (4 if ver == 1 else 0)
            self._sizes(so, sl)
            self.base = self.u(p, 8)
            # root group symbol table entry after four addresses
            ent = p + 32
            self.root = self.u(ent + 8, 8)
        elif ver in (2, 3):
            so, sl = data[base + 9], data[base + 10]
            self._sizes(so, sl)
            self.base = self.u(base + 12, 8)
            self.root = self.u(base + 12 + 24, 8)
        else:
            raise ValueError('unsupported superblock version %d' % ver)

    def _sizes(self, so, sl):
        if so != 8 or sl != 8:
            raise ValueError('unsupported offset/length sizes %d/%d' % (so, sl))

    def u(self, off, n):
        return int.from_bytes(self.d[off:off + n], 'little')

    def addr(self, a):
        return self.base + a

    # -- object headers ---------------------------------------------------
    def messages(self, a):
        """[(type, bytes)] of the object header at file address a."""
        p = self.addr(a)
        if self.d[p:p + 4] == b'OHDR':
            return self._messages_v2(p)
        if self.d[p] != 1:
            raise ValueError('unknown object header version %d at %d' % (self.d[p], a))
        nmsg, size = self.u(p + 2, 2), self.u(p + 8, 4)
        out, blocks = [], [(p + 16, size)]
        while blocks and len(out) < nmsg:
            q, n = blocks.pop(0)
            end = q + n
            while q + 8 <= end and len(out) < nmsg:
                t, s = self.u(q, 2), self.u(q + 2, 2)
                body = self.d[q + 8:q + 8 + s]
                out.append((t, body))
                if t == CONTINUATION:
                    blocks.append((self.addr(int.from_bytes(body[:8], 'little')),
                                   int.from_bytes(body[8:16], 'little')))
                q += 8 + s
        return out

    def _messages_v2(self, p):
        flags = self.d[p + 5]
        q = p + 6
        if flags & 0x20:
            q += 16
        if flags & 0x10:
            q += 4
        w = 1 << (flags & 3)
        size = self.u(q, w)
        q += w
        out, blocks = [], [(q, size)]
        tracked = bool(flags & 0x04)
        hdr = 6 if tracked else 4
        while blocks:
            q, n = blocks.pop(0)
            end = q + n
            while q + hdr <= end:
                t, s = self.d[q], self.u(q + 1, 2)
                body = self.d[q + hdr:q + hdr + s]
                out.append((t, body))
                if t == CONTINUATION:
                    c = self.addr(int.from_bytes(body[:8], 'little'))
                    ln = int.from_bytes(body[8:16], 'little')
                    if self.d[c:c + 4] != b'OCHK':
                        raise ValueError('bad continuation block at %d' % c)
                    blocks.append((c + 4, ln - 8))            # signature, checksum
                q += hdr + s
        return out

    # -- groups -----------------------------------------------------------
    def links(self, a):
        """{name: object header address} of the group at address a."""
        msgs = self.messages(a)
        out = {}
        for t, b in msgs:
            if t == SYMTABLE:
                bt, heap = int.from_bytes(b[:8], 'little'), int.from_bytes(b[8:16], 'little')
                names = self._heap(heap)
                self._btree(bt, names, out)
            elif t == LINK:
                name, target = self._link(b)
                if target is not None:
                    out[name] = target
            elif t == LINKINFO:
                fheap = int.from_bytes(b[2 + (8 if b[1] & 1 else 0):][:8], 'little')
                if fheap != UNDEF:
                    raise ValueError('dense link storage (fractal heap) is not supported')
        return out

    def _heap(self, a):
        p = self.addr(a)
        if self.d[p:p + 4] != b'HEAP':
            raise ValueError('bad local heap at %d' % a)
        seg = self.addr(self.u(p + 24, 8))

        def name(off):
            q = seg + off
            return self.d[q:self.d.index(b'\0', q)].decode()
        return name

    def _btree(self, a, names, out):
        p = self.addr(a)
        if self.d[p:p + 4] != b'TREE' or self.d[p + 4] != 0:
            raise ValueError('bad group B-tree node at %d' % a)
        level, used = self.d[p + 5], self.u(p + 6, 2)
        q = p + 24
        for i in range(used):
            child = self.u(q + 8, 8)
            if level > 0:
                self._btree(child, names, out)
            else:
                self._snod(child, names, out)
            q += 16

    def _snod(self, a, names, out):
        p = self.addr(a)
        if self.d[p:p + 4] != b'SNOD':
            raise ValueError('bad symbol table node at %d' % a)
        for i in range(self.u(p + 6, 2)):
            e = p + 8 + 40 * i
            out[names(self.u(e, 8))] = self.u(e + 8, 8)

    def _link(self, b):
        flags = b[1]
        q = 2
        ltype = 0
        if flags & 0x08:
            ltype = b[q]
            q += 1
        if flags & 0x04:
            q += 8
        if flags & 0x10:
            q += 1
        w = 1 << (flags & 3)
        n = int.from_bytes(b[q:q + w], 'little')
        q += w
        name = bytes(b[q:q + n]).decode()
        q += n
        if ltype != 0:                       # soft / external links are not followed
            return name, None
        return name, int.from_bytes(b[q:q + 8], 'little')

    # -- datasets ---------------------------------------------------------
    def dataset(self, a):
        """(ndarray, {attribute: value}) of the dataset at address a."""
        shape = dtype = layout = None
        attrs = {}
        for t, b in self.messages(a):
            if t == DATASPACE:
                shape = _dataspace(b)
            elif t == DATATYPE:
                dtype = _datatype(b)
            elif t == LAYOUT:
                layout = b
            elif t == FILTERS:
                raise ValueError('filtered (compressed) datasets are not supported')
            elif t == ATTRIBUTE:
                k, v = self._attribute(b)
                attrs[k] = v
            elif t == ATTRINFO:
                fheap = int.from_bytes(b[2 + (2 if b[1] & 1 else 0):][:8], 'little')
                if fheap != UNDEF:
                    raise ValueError('dense attribute storage is not supported')
        if shape is None or dtype is None or layout is None:
            raise ValueError('object at %d is not a dataset' % a)
        count = int(np.prod(shape, dtype=np.int64))
        ver, cls = layout[0], layout[1]
        if ver not in (3, 4):
            raise ValueError('unsupported layout message version %d' % ver)
        if cls == 0:                                         # compact
            n = int.from_bytes(layout[2:4], 'little')
            raw = bytes(layout[4:4 + n])
            arr = np.frombuffer(raw, dtype=dtype, count=count)
        elif cls == 1:                                       # contiguous
            at = int.from_bytes(layout[2:10], 'little')
            if at == UNDEF:                                  # never written: default fill
                arr = np.zeros(count, dtype=dtype)
            else:
                arr = np.frombuffer(self.d, dtype=dtype, count=count, offset=self.addr(at))
        else:
            raise ValueError('chunked datasets are not supported')
        return arr.reshape(shape).astype(dtype.newbyteorder('='), copy=True), attrs

    def _attribute(self, b):
        ver = b[0]
        nn, nt, ns = (int.from_bytes(b[i:i + 2], 'little') for i in (2, 4, 6))
        q = 8 if ver in (1, 2) else 9
        pad = _pad8 if ver == 1 else (lambda n: n)
        name = bytes(b[q:q + nn]).split(b'\0')[0].decode()
        q += pad(nn)
        dtype = _datatype(b[q:q + nt])
        q += pad(nt)
        shape = _dataspace(b[q:q + ns])
        q += pad(ns)
        count = int(np.prod(shape, dtype=np.int64))
        v = np.frombuffer(bytes(b[q:q + count * dtype.itemsize]), dtype=dtype, count=count)
        v = v.astype(dtype.newbyteorder('=')).reshape(shape)
        return name, (v[()] if shape == () else v)


def _dataspace(b):
    ver, rank, flags = b[0], b[1], b[2]
    if ver == 1:
        q = 8
    elif ver == 2:
        if b[3] == 2:
            raise ValueError('null dataspaces are not supported')
        q = 4
    else:
        raise ValueError('unsupported dataspace version %d' % ver)
    return tuple(int.from_bytes(b[q + 8 * i:q + 8 * i + 8], 'little') for i in range(rank))


def _datatype(b):
    cls, bits, size = b[0] & 0x0F, b[1], int.from_bytes(b[4:8], 'little')
    order = '>' if bits & 1 else '<'
    if cls == 0:
        return np.dtype('%s%s%d' % (order, 'i' if bits & 0x08 else 'u', size))
    if cls == 1:
        if size not in (2, 4, 8):
            raise ValueError('unsupported float size %d' % size)
        return np.dtype('%sf%d' % (order, size))
    raise ValueError('unsupported datatype class %d' % cls)


def read(path):
    """{dataset name: (ndarray, {attribute: value})} of the root group."""
    with open(path, 'rb') as fh:
        data = fh.read()
    r = _Reader(data)
    return {name: r.dataset(a) for name, a in sorted(r.links(r.root).items())}


# -- writer ----------------------------------------------------------------
def _msg(t, body, flags=0):
    body = bytes(body)
    body += b'\0' * (_pad8(len(body)) - len(body))
    return struct.pack('<HHB3x', t, len(body), flags) + body


def _dtype_msg(dt):
    dt = np.dtype(dt)
    if dt.kind == 'f':
        if dt.itemsize == 8:
            sign, eloc, esz, msz, bias = 63, 52, 11, 52, 1023
        elif dt.itemsize == 4:
            sign, eloc, esz, msz, bias = 31, 23, 8, 23, 127
        else:
            raise ValueError('unsupported float size %d' % dt.itemsize)
        # class 1, version 1; little endian, implied msb mantissa normalisation
        return (struct.pack('<B3BI', 0x11, 0x20, sign, 0, dt.itemsize)
                + struct.pack('<HHBBBBI', 0, 8 * dt.itemsize, eloc, esz, 0, msz, bias))
    if dt.kind in 'iu':
        return (struct.pack('<B3BI', 0x10, 0x08 if dt.kind == 'i' else 0, 0, 0, dt.itemsize)
                + struct.pack('<HH', 0, 8 * dt.itemsize))
    raise ValueError('unsupported dtype %s' % dt)


def _space_msg(shape):
    # version 1; maximum dimensions stored (equal to the dimensions)
    b = struct.pack('<BBB5x', 1, len(shape), 1 if shape else 0)
    for n in shape:
        b += struct.pack('<Q', n)
    for n in shape:
        b += struct.pack('<Q', n)
    return b


def _attr_msg(name, value):
    v = np.asarray(value)
    if v.dtype.kind == 'f':
        v = v.astype('<f8')
    elif v.dtype.kind in 'iub':
        v = v.astype('<i8')
    else:
        raise ValueError('unsupported attribute %s of type %s' % (name, v.dtype))
    nm = name.encode() + b'\0'
    dt, sp = _dtype_msg(v.dtype), _space_msg(v.shape)

    def padded(x):
        return x + b'\0' * (_pad8(len(x)) - len(x))
    return _msg(ATTRIBUTE, struct.pack('<BxHHH', 1, len(nm), len(dt), len(sp))
                + padded(nm) + padded(dt) + padded(sp) + v.tobytes())


def _object_header(msgs):
    body = b''.join(msgs)
    return struct.pack('<BxHII4x', 1, len(msgs), 1, len(body)) + body


def write(path, datasets):
    """Write {name: (ndarray, {attribute: scalar})} as datasets of the root
    group, in the layout h5py writes by default (module docstring)."""
    names = sorted(datasets)
    if not 1 <= len(names) <= 2 * LEAF_K:
        raise ValueError('between 1 and %d datasets per file' % (2 * LEAF_K))
    # local heap data: "" at 0 (the root's own name), then the names
    heapdata, offs = b'\0' * 8, {}
    for n in names:
        offs[n] = len(heapdata)
        e = n.encode() + b'\0'
        heapdata += e + b'\0' * (_pad8(len(e)) - len(e))
    SB, SYMENT = 56, 40
    root_hdr = _object_header([_msg(SYMTABLE, b'\0' * 16)])   # addresses patched below
    a_root = SB + SYMENT
    a_btree = a_root + len(root_hdr)
    btree_size = 24 + (2 * INTERNAL_K + 1) * 8 + 2 * INTERNAL_K * 8
    a_heap = a_btree + btree_size
    a_heapdata = a_heap + 32
    a_snod = a_heapdata + len(heapdata)
    snod_size = 8 + 2 * LEAF_K * SYMENT
    p = a_snod + snod_size
    objs, addrs = [], {}
    for n in names:
        arr, attrs = datasets[n]
        arr = np.ascontiguousarray(arr)
        if arr.dtype.byteorder == '>':
            arr = arr.astype(arr.dtype.newbyteorder('<'))
        # header first with a placeholder data address; its size does not depend on it
        def header(at, arr=arr, attrs=attrs):
            msgs = [_msg(DATASPACE, _space_msg(arr.shape)),
                    _msg(DATATYPE, _dtype_msg(arr.dtype), flags=1),
                    # fill value v2: allocate late, write fill if set, default value
                    _msg(FILL, struct.pack('<BBBB', 2, 2, 2, 1) + struct.pack('<I', 0), flags=1),
                    _msg(LAYOUT, struct.pack('<BBQQ', 3, 1, at, arr.nbytes))]
            msgs += [_attr_msg(k, attrs[k]) for k in sorted(attrs)]
            return _object_header(msgs)
        hlen = len(header(0))
        addrs[n] = p
        at = _pad8(p + hlen)
        objs.append((p, header(at)))
        objs.append((at, arr.tobytes()))
        p = _pad8(at + arr.nbytes)
    eof = p
    out = bytearray(eof)
    # superblock version 0
    out[0:8] = SIGNATURE
    out[8:16] = bytes([0, 0, 0, 0, 0, 8, 8, 0])
    out[16:24] = struct.pack('<HHI', LEAF_K, INTERNAL_K, 0)
    out[24:56] = struct.pack('<QQQQ', 0, UNDEF, eof, UNDEF)
    # root symbol table entry: cached B-tree and heap addresses
    out[56:96] = struct.pack('<QQI4xQQ', 0, a_root, 1, a_btree, a_heap)
    root_hdr = _object_header([_msg(SYMTABLE, struct.pack('<QQ', a_btree, a_heap))])
    out[a_root:a_root + len(root_hdr)] = root_hdr
    # one leaf B-tree node: key 0 (""), child SNOD, key of the last name
    bt = b'TREE' + struct.pack('<BBHQQ', 0, 0, 1, UNDEF, UNDEF)
    bt += struct.pack('<QQQ', 0, a_snod, offs[names[-1]])
    out[a_btree:a_btree + len(bt)] = bt
    # no free block: the free-list head holds the library's null offset, 1
    out[a_heap:a_heap + 32] = b'HEAP' + struct.pack('<B3xQQQ', 0, len(heapdata), 1, a_heapdata)
    out[a_heapdata:a_heapdata + len(heapdata)] = heapdata
    sn = b'SNOD' + struct.pack('<BxH', 1, len(names))
    for n in names:
        sn += struct.pack('<QQI4x16x', offs[n], addrs[n], 0)
    out[a_snod:a_snod + len(sn)] = sn
    for at, blob in objs:
        out[at:at + len(blob)] = blob
    with open(path, 'wb') as fh:
        fh.write(out)


def is_hdf5(path):
    with open(path, 'rb') as fh:
        head = fh.read(8)
    return head == SIGNATURE
