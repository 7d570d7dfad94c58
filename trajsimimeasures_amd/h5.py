"""A reader for the subset of HDF5 that the trajectory dataset files use.

The dataset files (written by preprocessing/pkl2h5py.py:16-27 with h5py's
defaults, read by utilities/data_loader.py:52-99 and
query/cpp/utils/dataset.hpp:48-113) hold numeric attributes on the root
group, a ``/trajs_len`` dataset and one small dataset per trajectory under
``/trajs_merc`` (and ``/trajs_ts``).  h5py is not a dependency of this
package, so this module reads those files itself, from the published HDF5
file format (specification version 3): superblock versions 0-3, version 1
and 2 object headers with continuation blocks, old-style groups (symbol
table message, version 1 group B-tree, local heap, symbol table nodes),
new-style groups whose links are stored compactly in link messages,
dataspace, datatype (fixed and floating point of either byte order),
attribute, and contiguous, compact or chunked data layout messages (chunks
indexed by the version 1 B-tree, through the deflate, shuffle and
Fletcher-32 filters).  Anything else (other chunk indexes or filters, dense
link storage, variable-length or compound types, external files) raises
``H5FormatError`` naming it, never a silent misread.

Data are read through a read-only memory map, so opening a file with a
million trajectories costs only the group index, and each dataset read is
one ``numpy.frombuffer`` slice of the map (chunked datasets: one per chunk).
"""
from __future__ import annotations

import mmap
import zlib

import numpy as np

__all__ = ["H5File", "H5Dataset", "H5FormatError"]

_SIG = b"\x89HDF\r\n\x1a\n"
_UNDEF = 0xFFFFFFFFFFFFFFFF


class H5FormatError(ValueError):
    """The file is not HDF5, or uses a feature this reader does not read."""


class _Obj:
    """Parsed object header: its messages as (type, bytes) in file order."""

    __slots__ = ("addr", "msgs")

    def __init__(self, addr, msgs):
        self.addr = addr
        self.msgs = msgs

    def first(self, mtype):
        for t, d in self.msgs:
            if t == mtype:
                return d
        return None

    def all(self, mtype):
        return [d for t, d in self.msgs if t == mtype]


class H5Dataset:
    """A dataset: shape, numpy dtype, and its bytes in the file."""

    def __init__(self, f: "H5File", obj: _Obj, name: str):
        self.name = name
        self._f = f
        self.shape = f._dataspace(obj.first(0x0001), name)
        self.dtype = f._datatype(obj.first(0x0003), name)
        self._layout = f._layout(obj.first(0x0008), name)
        self._filters = f._filters(obj.first(0x000B), name)
        self.attrs = f._attributes(obj)

    @property
    def size(self) -> int:
        return int(np.prod(self.shape, dtype=np.int64)) if self.shape else 1

    def read(self, count: int | None = None) -> np.ndarray:
        """The whole dataset, or its first ``count`` rows along axis 0."""
        shape = tuple(self.shape)
        if count is not None and shape:
            shape = (max(0, min(int(count), shape[0])),) + shape[1:]
        n = int(np.prod(shape, dtype=np.int64)) if shape else 1
        kind, a, b = self._layout
        if n == 0:
            return np.zeros(shape, self.dtype.newbyteorder("="))
        nbytes = n * self.dtype.itemsize
        if kind == "chunked":
            return self._read_chunked()[:shape[0]] if shape else self._read_chunked()
        if self._filters:
            raise H5FormatError(f"{self.name}: filters on a {kind} dataset")
        if kind == "compact":
            raw = a[:nbytes]
        else:
            if a == _UNDEF:  # storage never written: HDF5's fill value, zero by default
                return np.zeros(shape, self.dtype.newbyteorder("="))
            if nbytes > b:
                raise H5FormatError(f"{self.name}: contiguous storage of {b} bytes is shorter than its shape")
            raw = self._f._bytes(a, nbytes)
        if len(raw) < nbytes:
            raise H5FormatError(f"{self.name}: data runs past the end of the file")
        arr = np.frombuffer(raw, dtype=self.dtype, count=n).reshape(shape)
        return arr.astype(self.dtype.newbyteorder("="), copy=True)

    def _read_chunked(self) -> np.ndarray:
        f = self._f
        so = f._so
        _, btree, cdims = self._layout
        out = np.zeros(self.shape, self.dtype.newbyteorder("="))
        if btree == _UNDEF or out.size == 0:
            return out
        rank = len(self.shape)
        if len(cdims) != rank:
            raise H5FormatError(f"{self.name}: chunk rank {len(cdims)} for a rank-{rank} dataset")
        ksize = 8 + 8 * (rank + 1)  # chunk size, filter mask, offsets (one more for the element)
        per = int(np.prod(cdims, dtype=np.int64))
        stack, seen = [btree], set()
        while stack:
            a = stack.pop()
            if a in seen:
                raise H5FormatError(f"{self.name}: chunk B-tree cycle at {a:#x}")
            seen.add(a)
            node = f._bytes(a, 8 + 2 * so)
            if node[:4] != b"TREE" or node[4] != 1:
                raise H5FormatError(f"{self.name}: bad chunk B-tree node at {a:#x}")
            level, used = node[5], f._uint(node, 6, 2)
            body = f._bytes(a + 8 + 2 * so, used * (ksize + so) + ksize)
            for i in range(used):
                k = i * (ksize + so)
                child = f._uint(body, k + ksize, so)
                if level > 0:
                    stack.append(child)
                    continue
                csize, mask = f._uint(body, k, 4), f._uint(body, k + 4, 4)
                offs = [f._uint(body, k + 8 + 8 * j, 8) for j in range(rank)]
                raw = self._unfilter(bytes(f._bytes(child, csize)), mask)
                if len(raw) < per * self.dtype.itemsize:
                    raise H5FormatError(f"{self.name}: chunk at {child:#x} holds {len(raw)} bytes")
                chunk = np.frombuffer(raw, self.dtype, count=per).reshape(cdims)
                dst = tuple(slice(o, min(o + c, n)) for o, c, n in zip(offs, cdims, self.shape))
                if any(d.start >= d.stop for d in dst):
                    continue
                out[dst] = chunk[tuple(slice(0, d.stop - d.start) for d in dst)]
        return out

    def _unfilter(self, raw: bytes, mask: int) -> bytes:
        for i in reversed(range(len(self._filters))):
            if mask & (1 << i):
                continue  # this chunk skipped the filter
            fid = self._filters[i]
            if fid == 1:
                raw = zlib.decompress(raw)
            elif fid == 2:
                w = self.dtype.itemsize
                b = np.frombuffer(raw, np.uint8)
                m = len(b) // w
                raw = b[:m * w].reshape(w, m).T.tobytes() + bytes(b[m * w:])
            elif fid == 3:
                raw = raw[:-4]  # Fletcher-32 checksum (not verified)
            else:
                raise H5FormatError(f"{self.name}: filter {fid}")
        return raw

    def __getitem__(self, key):
        if key == () or key is Ellipsis:
            return self.read()
        if isinstance(key, slice) and key.start in (None, 0) and key.step in (None, 1):
            return self.read(key.stop)
        return self.read()[key]


class H5File:
    """Read-only HDF5 file.  ``f.attrs``: root attributes; ``f[path]``: a
    dataset; ``f.keys(group)``: member names of a group."""

    def __init__(self, path: str):
        self.path = path
        self._fh = open(path, "rb")
        try:
            self._mm = mmap.mmap(self._fh.fileno(), 0, access=mmap.ACCESS_READ)
        except ValueError as e:  # empty file
            self._fh.close()
            raise H5FormatError(f"{path}: not an HDF5 file ({e})") from None
        self._groups = {}  # path -> {name: object header address}
        try:
            self._superblock()
            root = self._object(self._root_addr)
            self.attrs = self._attributes(root)
            self._root = root
        except Exception:
            self.close()
            raise

    # -- file plumbing -------------------------------------------------------
    def close(self):
        if getattr(self, "_mm", None) is not None:
            self._mm.close()
            self._mm = None
        if not self._fh.closed:
            self._fh.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _bytes(self, addr, n):
        a = self._base + addr
        if addr == _UNDEF or a < 0 or a + n > len(self._mm):
            raise H5FormatError(f"{self.path}: address {addr:#x}+{n} outside the file")
        return self._mm[a:a + n]

    def _uint(self, buf, off, size):
        return int.from_bytes(buf[off:off + size], "little")

    # -- superblock ----------------------------------------------------------
    def _superblock(self):
        mm = self._mm
        base = None
        off = 0
        while off + 8 <= len(mm):  # signature at 0, 512, 1024, 2048, ...
            if mm[off:off + 8] == _SIG:
                base = off
                break
            off = 512 if off == 0 else off * 2
        if base is None:
            raise H5FormatError(f"{self.path}: no HDF5 signature")
        ver = mm[base + 8]
        if ver in (0, 1):
            self._so, self._sl = mm[base + 13], mm[base + 14]
            so = self._so
            # base, free-space, end-of-file and driver addresses, then the
            # root group's symbol table entry (link name offset, header address)
            p = base + 24 + (4 if ver == 1 else 0) + 4 * so
            self._root_addr = self._uint(mm, p + so, so)
        elif ver in (2, 3):
            self._so, self._sl = mm[base + 9], mm[base + 10]
            so = self._so
            # base, superblock extension, end-of-file, root object header addresses
            self._root_addr = self._uint(mm, base + 12 + 3 * so, so)
        else:
            raise H5FormatError(f"{self.path}: superblock version {ver}")
        if self._so not in (2, 4, 8) or self._sl not in (2, 4, 8):
            raise H5FormatError(f"{self.path}: offset/length sizes {self._so}/{self._sl}")
        # addresses count from the superblock (offset 0 unless a user block precedes it)
        self._base = base

    # -- object headers ------------------------------------------------------
    def _object(self, addr) -> _Obj:
        head = self._bytes(addr, 16)
        if head[:4] == b"OHDR":
            return _Obj(addr, self._object_v2(addr))
        if head[0] != 1:
            raise H5FormatError(f"{self.path}: object header version {head[0]} at {addr:#x}")
        nmsg = self._uint(head, 2, 2)
        size = self._uint(head, 8, 4)
        msgs = []
        blocks = [(addr + 16, size)]
        seen = set()
        while blocks and len(msgs) < nmsg:
            a, n = blocks.pop(0)
            if a in seen:
                raise H5FormatError(f"{self.path}: object header continuation cycle at {a:#x}")
            seen.add(a)
            buf = self._bytes(a, n)
            p = 0
            while p + 8 <= n and len(msgs) < nmsg:
                mtype = self._uint(buf, p, 2)
                msize = self._uint(buf, p + 2, 2)
                flags = buf[p + 4]
                data = buf[p + 8:p + 8 + msize]
                p += 8 + msize
                if flags & 0x02:
                    raise H5FormatError(f"{self.path}: shared message (type {mtype:#x}) at {addr:#x}")
                if mtype == 0x0010:
                    blocks.append((self._uint(data, 0, self._so), self._uint(data, self._so, self._sl)))
                msgs.append((mtype, data))
        return _Obj(addr, msgs)

    def _object_v2(self, addr):
        head = self._bytes(addr, 64)
        if head[4] != 2:
            raise H5FormatError(f"{self.path}: OHDR version {head[4]}")
        flags = head[5]
        p = 6 + (16 if flags & 0x20 else 0) + (4 if flags & 0x10 else 0)
        szw = 1 << (flags & 0x03)
        size = self._uint(head, p, szw)
        p += szw
        track = bool(flags & 0x04)
        msgs = []
        blocks = [(addr + p, size)]
        seen = set()
        while blocks:
            a, n = blocks.pop(0)
            if a in seen:
                raise H5FormatError(f"{self.path}: object header continuation cycle at {a:#x}")
            seen.add(a)
            buf = self._bytes(a, n)
            q = 0
            hdr = 4 + (2 if track else 0)
            while q + hdr <= n:
                mtype = buf[q]
                msize = self._uint(buf, q + 1, 2)
                mflags = buf[q + 3]
                data = buf[q + hdr:q + hdr + msize]
                q += hdr + msize
                if mflags & 0x02:
                    raise H5FormatError(f"{self.path}: shared message (type {mtype:#x}) at {addr:#x}")
                if mtype == 0x0010:
                    ca, cl = self._uint(data, 0, self._so), self._uint(data, self._so, self._sl)
                    if self._bytes(ca, 4) != b"OCHK":
                        raise H5FormatError(f"{self.path}: bad continuation block at {ca:#x}")
                    blocks.append((ca + 4, cl - 8))  # signature and checksum
                msgs.append((mtype, data))
        return msgs

    # -- messages ------------------------------------------------------------
    def _dataspace(self, d, name):
        if d is None:
            raise H5FormatError(f"{name}: no dataspace message")
        ver, rank, flags = d[0], d[1], d[2]
        if ver == 1:
            p = 8
        elif ver == 2:
            if d[3] == 2:  # null dataspace
                return (0,)
            p = 4
        else:
            raise H5FormatError(f"{name}: dataspace version {ver}")
        return tuple(self._uint(d, p + i * self._sl, self._sl) for i in range(rank))

    def _datatype(self, d, name):
        if d is None:
            raise H5FormatError(f"{name}: no datatype message")
        cls = d[0] & 0x0F
        bits = d[1]
        size = self._uint(d, 4, 4)
        order = ">" if bits & 0x01 else "<"
        if cls == 0:
            kind = "i" if bits & 0x08 else "u"
            if size not in (1, 2, 4, 8):
                raise H5FormatError(f"{name}: integer of {size} bytes")
            return np.dtype(f"{order}{kind}{size}")
        if cls == 1:
            if bits & 0x40 or size not in (2, 4, 8):
                raise H5FormatError(f"{name}: floating point of {size} bytes (flags {bits:#x})")
            return np.dtype(f"{order}f{size}")
        raise H5FormatError(f"{name}: datatype class {cls} (only fixed and floating point are read)")

    def _layout(self, d, name):
        if d is None:
            raise H5FormatError(f"{name}: no data layout message")
        ver = d[0]
        if ver == 3 or ver == 4:
            cls = d[1]
            if cls == 0:
                n = self._uint(d, 2, 2)
                return ("compact", bytes(d[4:4 + n]), n)
            if cls == 1:
                return ("contiguous", self._uint(d, 2, self._so), self._uint(d, 2 + self._so, self._sl))
            if cls == 2 and ver == 3:  # chunks indexed by a version 1 B-tree
                r = d[2]
                a = self._uint(d, 3, self._so)
                dims = [self._uint(d, 3 + self._so + 4 * i, 4) for i in range(r)]
                return ("chunked", a, tuple(dims[:-1]))  # the last is the element size
            raise H5FormatError(f"{name}: {'chunked (version 4 index)' if cls == 2 else 'virtual'} data layout")
        if ver in (1, 2):
            rank, cls = d[1], d[2]
            if cls == 1:
                a = self._uint(d, 8, self._so)
                dims = [self._uint(d, 8 + self._so + 4 * i, 4) for i in range(rank)]
                return ("contiguous", a, int(np.prod(dims)) if dims else 0)
            if cls == 0:
                p = 8 + 4 * rank
                n = self._uint(d, p, 4)
                return ("compact", bytes(d[p + 4:p + 4 + n]), n)
            if cls == 2:
                a = self._uint(d, 8, self._so)
                dims = [self._uint(d, 8 + self._so + 4 * i, 4) for i in range(rank)]
                return ("chunked", a, tuple(dims[:-1]))
            raise H5FormatError(f"{name}: data layout class {cls}")
        raise H5FormatError(f"{name}: data layout version {ver}")

    def _filters(self, d, name):
        if d is None:
            return []
        ver, n = d[0], d[1]
        out = []
        if ver == 1:
            p = 8
            for _ in range(n):
                fid, nlen, nval = self._uint(d, p, 2), self._uint(d, p + 2, 2), self._uint(d, p + 6, 2)
                p += 8 + ((nlen + 7) & ~7) + 4 * nval + (4 if nval % 2 else 0)
                out.append(fid)
        elif ver == 2:
            p = 2
            for _ in range(n):
                fid = self._uint(d, p, 2)
                p += 2
                nlen = 0
                if fid >= 256:
                    nlen = self._uint(d, p, 2)
                    p += 2
                nval = self._uint(d, p + 2, 2)
                p += 4 + nlen + 4 * nval
                out.append(fid)
        else:
            raise H5FormatError(f"{name}: filter pipeline version {ver}")
        return out

    def _attributes(self, obj):
        out = {}
        for d in obj.all(0x000C):
            ver = d[0]
            nsz, tsz, ssz = self._uint(d, 2, 2), self._uint(d, 4, 2), self._uint(d, 6, 2)
            if ver == 1:
                pad = lambda n: (n + 7) & ~7
                p = 8
                name = bytes(d[p:p + nsz]).split(b"\0", 1)[0].decode()
                p += pad(nsz)
                t = d[p:p + tsz]
                p += pad(tsz)
                s = d[p:p + ssz]
                p += pad(ssz)
            elif ver in (2, 3):
                p = 8 + (1 if ver == 3 else 0)
                name = bytes(d[p:p + nsz]).split(b"\0", 1)[0].decode()
                p += nsz
                t = d[p:p + tsz]
                p += tsz
                s = d[p:p + ssz]
                p += ssz
            else:
                raise H5FormatError(f"attribute message version {ver}")
            try:
                dt = self._datatype(t, "attribute " + name)
            except H5FormatError:
                continue  # attributes of other types (strings, ...) are not needed here
            shape = self._dataspace(s, "attribute " + name)
            n = int(np.prod(shape, dtype=np.int64)) if shape else 1
            arr = np.frombuffer(bytes(d[p:p + n * dt.itemsize]), dtype=dt, count=n)
            arr = arr.astype(dt.newbyteorder("="))
            out[name] = arr.reshape(shape) if shape else arr[0]
        return out

    # -- groups --------------------------------------------------------------
    def _members(self, obj: _Obj, path: str):
        stab = obj.first(0x0011)
        if stab is not None:
            return self._symbol_table(self._uint(stab, 0, self._so), self._uint(stab, self._so, self._so), path)
        info = obj.first(0x0002)
        if info is not None:
            fl = info[1]
            p = 2 + (8 if fl & 0x01 else 0)
            if self._uint(info, p, self._so) != (1 << (8 * self._so)) - 1:
                raise H5FormatError(f"{path}: dense link storage (fractal heap)")
        out = {}
        for d in obj.all(0x0006):
            fl = d[1]
            p = 2
            ltype = 0
            if fl & 0x08:
                ltype = d[p]
                p += 1
            if fl & 0x04:
                p += 8
            if fl & 0x10:
                p += 1
            lw = 1 << (fl & 0x03)
            ln = self._uint(d, p, lw)
            p += lw
            name = bytes(d[p:p + ln]).decode()
            p += ln
            if ltype != 0:
                continue  # soft and external links are not followed
            out[name] = self._uint(d, p, self._so)
        return out

    def _symbol_table(self, btree, heap, path):
        h = self._bytes(heap, 8 + 2 * self._sl + self._so)
        if h[:4] != b"HEAP":
            raise H5FormatError(f"{path}: bad local heap at {heap:#x}")
        hsize = self._uint(h, 8, self._sl)
        hdata = self._bytes(self._uint(h, 8 + 2 * self._sl, self._so), hsize)
        out = {}
        stack = [btree]
        seen = set()
        while stack:
            a = stack.pop()
            if a in seen:
                raise H5FormatError(f"{path}: group B-tree cycle at {a:#x}")
            seen.add(a)
            node = self._bytes(a, 8 + 2 * self._so)
            if node[:4] != b"TREE" or node[4] != 0:
                raise H5FormatError(f"{path}: bad group B-tree node at {a:#x}")
            level, used = node[5], self._uint(node, 6, 2)
            p = 8 + 2 * self._so
            body = self._bytes(a + p, used * (self._sl + self._so) + self._sl)
            children = [self._uint(body, self._sl + i * (self._sl + self._so), self._so) for i in range(used)]
            if level > 0:
                stack.extend(reversed(children))
                continue
            for c in children:
                snod = self._bytes(c, 8)
                if snod[:4] != b"SNOD":
                    raise H5FormatError(f"{path}: bad symbol table node at {c:#x}")
                nsym = self._uint(snod, 6, 2)
                ent = 2 * self._so + 24
                ents = self._bytes(c + 8, nsym * ent)
                for i in range(nsym):
                    noff = self._uint(ents, i * ent, self._so)
                    oaddr = self._uint(ents, i * ent + self._so, self._so)
                    end = hdata.find(b"\0", noff)
                    out[bytes(hdata[noff:end]).decode()] = oaddr
        return out

    def _group(self, path: str):
        path = "/" + path.strip("/")
        if path in self._groups:
            return self._groups[path]
        if path == "/":
            m = self._members(self._root, "/")
        else:
            parent, _, name = path.rpartition("/")
            pm = self._group(parent or "/")
            if name not in pm:
                raise KeyError(path)
            m = self._members(self._object(pm[name]), path)
        self._groups[path] = m
        return m

    def keys(self, group: str = "/"):
        return list(self._group(group).keys())

    def __contains__(self, path):
        parent, _, name = ("/" + path.strip("/")).rpartition("/")
        try:
            return name in self._group(parent or "/")
        except KeyError:
            return False

    def __getitem__(self, path: str) -> H5Dataset:
        parent, _, name = ("/" + path.strip("/")).rpartition("/")
        g = self._group(parent or "/")
        if name not in g:
            raise KeyError(path)
        obj = self._object(g[name])
        if obj.first(0x0008) is None:
            raise H5FormatError(f"{path}: not a dataset")
        return H5Dataset(self, obj, path)
