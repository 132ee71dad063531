"""A reader for the subset of the HDF5 file format that Sedona's data files
use (atomic data `data/*.hdf5`, model and plot files `*.h5`), in plain
Python + numpy: the image has no HDF5 library.

Covered (HDF5 File Format Specification, version 2.0 structures of the
"earliest" file layout that h5py and the HDF5 C library write by default):
superblock versions 0 and 1; version-1 object headers with continuation
blocks; groups by symbol table (version-1 B-tree of group nodes, symbol
nodes, local heap); datasets with a version 1 or 2 dataspace, fixed-point,
floating-point and fixed-length string datatypes (either byte order), and
compact, contiguous or chunked storage (chunk B-tree, deflate and shuffle
filters); attributes of those types.  Links other than hard links,
variable-length and compound types are not read.

    f = H5File(path)
    f.keys()                 # root group member names
    f["1/0/level_E"]         # numpy array (a 0-d array for a scalar)
    "1/0/n_lines" in f
    f.attrs("1")             # {"n_ions": ..., ...}
"""
import zlib

import numpy as np

_SIG = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF


class H5Error(ValueError):
    pass


class _Node:
    """A group or dataset: its object header's messages by type."""

    def __init__(self, f, addr):
        self.f, self.addr = f, addr
        self.msgs = f._read_object_header(addr)

    def msg(self, mtype):
        for t, data in self.msgs:
            if t == mtype:
                return data
        return None


class H5File:
    def __init__(self, path):
        with open(path, "rb") as fh:
            self.buf = fh.read()
        b = self.buf
        if b[:8] != _SIG:
            raise H5Error("%s: not an HDF5 file" % path)
        ver = b[8]
        if ver not in (0, 1):
            raise H5Error("%s: superblock version %d is not supported" % (path, ver))
        self.so, self.sl = b[13], b[14]
        if self.so != 8 or self.sl != 8:
            raise H5Error("%s: only 8-byte offsets and lengths are supported" % path)
        p = 24 + (4 if ver == 1 else 0)
        self.base = self._u(p, 8)
        p += 32                                   # base, free space, end of file, driver block
        # root group symbol table entry: name offset, object header address
        self.root = self._u(p + 8, 8)
        self.path = path

    # -- primitive reads --------------------------------------------------
    def _u(self, p, n):
        return int.from_bytes(self.buf[p:p + n], "little")

    def _addr(self, a):
        return self.base + a

    # -- object headers -----------------------------------------------------
    def _read_object_header(self, addr):
        b = self.buf
        p = self._addr(addr)
        if b[p] != 1:
            raise H5Error("object header version %d at %d is not supported" % (b[p], addr))
        nmsg = self._u(p + 2, 2)
        size = self._u(p + 8, 4)
        blocks = [(p + 16, size)]
        msgs = []
        while blocks and len(msgs) < nmsg:
            q, n = blocks.pop(0)
            end = q + n
            while q + 8 <= end and len(msgs) < nmsg:
                mtype, msize = self._u(q, 2), self._u(q + 2, 2)
                data = b[q + 8:q + 8 + msize]
                if mtype == 0x0010:               # continuation
                    blocks.append((self._addr(int.from_bytes(data[0:8], "little")),
                                   int.from_bytes(data[8:16], "little")))
                msgs.append((mtype, data))
                q += 8 + msize
        return msgs

    # -- groups ---------------------------------------------------------------
    def _heap_data(self, heap_addr):
        p = self._addr(heap_addr)
        if self.buf[p:p + 4] != b"HEAP":
            raise H5Error("local heap signature missing at %d" % heap_addr)
        return self._addr(self._u(p + 24, 8))

    def _name(self, heap_data, off):
        p = heap_data + off
        e = self.buf.index(b"\x00", p)
        return self.buf[p:e].decode()

    def _group_members(self, node):
        st = node.msg(0x0011)
        if st is None:
            raise H5Error("object at %d is not a group (no symbol table)" % node.addr)
        btree = int.from_bytes(st[0:8], "little")
        heap = self._heap_data(int.from_bytes(st[8:16], "little"))
        out = {}
        self._walk_group_btree(btree, heap, out)
        return out

    def _walk_group_btree(self, addr, heap, out):
        b = self.buf
        p = self._addr(addr)
        if b[p:p + 4] != b"TREE":
            raise H5Error("B-tree signature missing at %d" % addr)
        ntype, level, used = b[p + 4], b[p + 5], self._u(p + 6, 2)
        if ntype != 0:
            raise H5Error("group B-tree node of type %d" % ntype)
        q = p + 24
        for _ in range(used):
            q += 8                                 # key (heap offset)
            child = self._u(q, 8)
            q += 8
            if level > 0:
                self._walk_group_btree(child, heap, out)
            else:
                self._read_snod(child, heap, out)

    def _read_snod(self, addr, heap, out):
        b = self.buf
        p = self._addr(addr)
        if b[p:p + 4] != b"SNOD":
            raise H5Error("symbol node signature missing at %d" % addr)
        n = self._u(p + 6, 2)
        q = p + 8
        for _ in range(n):
            name = self._name(heap, self._u(q, 8))
            out[name] = self._u(q + 8, 8)
            q += 40

    def _resolve(self, path):
        addr = self.root
        for part in [s for s in path.strip("/").split("/") if s]:
            members = self._group_members(_Node(self, addr))
            if part not in members:
                raise KeyError(path)
            addr = members[part]
        return _Node(self, addr)

    # -- public ---------------------------------------------------------------
    def keys(self, path="/"):
        return sorted(self._group_members(self._resolve(path)))

    def is_group(self, path):
        return self._resolve(path).msg(0x0011) is not None

    def __contains__(self, path):
        try:
            self._resolve(path)
            return True
        except KeyError:
            return False

    def __getitem__(self, path):
        node = self._resolve(path)
        if node.msg(0x0011) is not None:
            raise H5Error("%s is a group" % path)
        return self._read_dataset(node, path)

    # -- datasets -------------------------------------------------------------
    def _dataspace(self, d):
        ver = d[0]
        if ver == 1:
            rank, flags = d[1], d[2]
            p = 8
        elif ver == 2:
            rank, flags = d[1], d[2]
            if d[3] == 2:                          # null dataspace
                return None
            p = 4
        else:
            raise H5Error("dataspace version %d" % ver)
        return tuple(int.from_bytes(d[p + 8 * i:p + 8 * i + 8], "little") for i in range(rank))

    def _dtype(self, d):
        cls, ver = d[0] & 0x0F, d[0] >> 4
        bits = d[1] | (d[2] << 8) | (d[3] << 16)
        size = int.from_bytes(d[4:8], "little")
        order = ">" if bits & 1 else "<"
        if cls == 0:
            return np.dtype("%s%s%d" % (order, "i" if bits & 8 else "u", size))
        if cls == 1:
            if size not in (2, 4, 8):
                raise H5Error("floating-point size %d" % size)
            return np.dtype("%sf%d" % (order, size))
        if cls == 3:
            return np.dtype("S%d" % size)
        raise H5Error("datatype class %d (version %d) is not supported" % (cls, ver))

    def _filters(self, node):
        d = node.msg(0x000B)
        if d is None:
            return []
        ver, n = d[0], d[1]
        p = 8 if ver == 1 else 2
        out = []
        for _ in range(n):
            fid = int.from_bytes(d[p:p + 2], "little")
            if ver == 1 or fid >= 256:
                nlen = int.from_bytes(d[p + 2:p + 4], "little")
            else:
                nlen = 0
            nvals = int.from_bytes(d[p + 6:p + 8], "little")
            p += 8
            if nlen:
                p += (nlen + 7) // 8 * 8 if ver == 1 else nlen
            vals = [int.from_bytes(d[p + 4 * i:p + 4 * i + 4], "little") for i in range(nvals)]
            p += 4 * nvals
            if ver == 1 and nvals % 2:
                p += 4
            out.append((fid, vals))
        return out

    def _unfilter(self, raw, filters, mask, itemsize):
        for k in range(len(filters) - 1, -1, -1):
            if mask & (1 << k):
                continue
            fid, vals = filters[k]
            if fid == 1:                           # deflate
                raw = zlib.decompress(raw)
            elif fid == 2:                         # shuffle
                a = np.frombuffer(raw, np.uint8)
                n = len(a) // itemsize
                raw = a[:n * itemsize].reshape(itemsize, n).T.tobytes() + a[n * itemsize:].tobytes()
            else:
                raise H5Error("filter %d is not supported" % fid)
        return raw

    def _read_dataset(self, node, path):
        shape = self._dataspace(node.msg(0x0001))
        dt = self._dtype(node.msg(0x0003))
        lay = node.msg(0x0008)
        if shape is None:
            return np.zeros(0, dt)
        n = int(np.prod(shape)) if shape else 1
        nbytes = n * dt.itemsize
        ver = lay[0]
        if ver != 3:
            raise H5Error("%s: data layout version %d is not supported" % (path, ver))
        cls = lay[1]
        if cls == 0:                               # compact
            size = int.from_bytes(lay[2:4], "little")
            raw = lay[4:4 + size]
        elif cls == 1:                             # contiguous
            addr = int.from_bytes(lay[2:10], "little")
            if addr == UNDEF:
                raw = bytes(nbytes)
            else:
                p = self._addr(addr)
                raw = self.buf[p:p + nbytes]
        elif cls == 2:                             # chunked
            rank = lay[2]
            btree = int.from_bytes(lay[3:11], "little")
            cdims = [int.from_bytes(lay[11 + 4 * i:15 + 4 * i], "little") for i in range(rank)]
            return self._read_chunked(node, path, shape, dt, btree, cdims[:-1])
        else:
            raise H5Error("%s: layout class %d" % (path, cls))
        a = np.frombuffer(raw[:nbytes], dt).copy()
        return a.reshape(shape) if shape else a.reshape(())

    def _read_chunked(self, node, path, shape, dt, btree, cdims):
        out = np.zeros(shape, dt)
        filters = self._filters(node)
        rank = len(shape)
        csize = int(np.prod(cdims)) * dt.itemsize
        for offs, addr, size, mask in self._chunk_entries(btree, rank):
            raw = self.buf[self._addr(addr):self._addr(addr) + size]
            raw = self._unfilter(raw, filters, mask, dt.itemsize)
            chunk = np.frombuffer(raw[:csize], dt).reshape(cdims)
            sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offs, cdims, shape))
            out[sl] = chunk[tuple(slice(0, s.stop - s.start) for s in sl)]
        return out

    def _chunk_entries(self, addr, rank):
        b = self.buf
        p = self._addr(addr)
        if b[p:p + 4] != b"TREE" or b[p + 4] != 1:
            raise H5Error("chunk B-tree node missing at %d" % addr)
        level, used = b[p + 5], self._u(p + 6, 2)
        ksize = 8 + 8 * (rank + 1)
        q = p + 24
        out = []
        for _ in range(used):
            size, mask = self._u(q, 4), self._u(q + 4, 4)
            offs = [self._u(q + 8 + 8 * i, 8) for i in range(rank)]
            child = self._u(q + ksize, 8)
            q += ksize + 8
            if level > 0:
                out.extend(self._chunk_entries(child, rank))
            else:
                out.append((offs, child, size, mask))
        return out

    def attrs(self, path="/"):
        """The attributes of a group or dataset whose datatype this reader
        covers, as numpy arrays (a 0-d array for a scalar)."""
        node = self._resolve(path)
        out = {}
        for t, d in node.msgs:
            if t != 0x000C:
                continue
            ver = d[0]
            nsz, tsz, ssz = (int.from_bytes(d[k:k + 2], "little") for k in (2, 4, 6))
            if ver == 1:
                pad = lambda v: (v + 7) // 8 * 8
                p = 8
            elif ver in (2, 3):
                pad = lambda v: v
                p = 8 if ver == 2 else 9
            else:
                raise H5Error("attribute message version %d" % ver)
            name = d[p:p + nsz].split(b"\x00")[0].decode()
            p += pad(nsz)
            dt = self._dtype(d[p:p + tsz])
            p += pad(tsz)
            shape = self._dataspace(d[p:p + ssz])
            p += pad(ssz)
            if shape is None:
                out[name] = np.zeros(0, dt)
                continue
            n = int(np.prod(shape)) if shape else 1
            a = np.frombuffer(d[p:p + n * dt.itemsize], dt).copy()
            out[name] = a.reshape(shape) if shape else a.reshape(())
        return out

    def walk(self, path="/"):
        """(path, shape, dtype) of every dataset under path."""
        out = []
        for k in self.keys(path):
            sub = (path.rstrip("/") + "/" + k).lstrip("/")
            if self.is_group(sub):
                out.extend(self.walk(sub))
            else:
                a = self[sub]
                out.append((sub, a.shape, a.dtype))
        return out
