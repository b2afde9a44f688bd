"""A small reader (and a writer for particle-bundle dumps) of the HDF5 files
the reference exchanges, for hosts without the HDF5 library.

The reference reads its geometry from an HDF5 file written with h5py
(R/script/geom/geometry.py:721-738; read in R/src/base/init.F90:105-310)
and dumps particle bundles as HDF5 (DumpParticleBundleHDF5 /
LoadParticleBundleHDF5, R/src/pic/particle/ParticleBundle.f90:117-227).
Both are plain HDF5: a root group holding scalar attributes and datasets
of fixed-point or IEEE floating-point numbers.  This module follows the
published HDF5 file format specification (superblock versions 0-3, object
header versions 1-2, symbol-table and compact link storage, contiguous,
compact and chunked layouts with the deflate and shuffle filters) as far as
such files need, and writes bundle dumps with superblock 0, version-1
object headers and contiguous datasets, the layout the HDF5 library itself
produces by default.

Arrays come back in C order with the dataset's dimensions.  A Fortran
program that writes A(n1, n2) stores dimensions (n2, n1), so
``File.read_fortran`` returns the transpose, indexed as the Fortran code
indexes it.
"""
import struct
import zlib

import numpy as np

from . import restart

SIG = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF


class _Buf:
    def __init__(self, data):
        self.d = data

    def u(self, off, n):
        return int.from_bytes(self.d[off:off + n], "little")


class File:
    """Read-only view of an HDF5 file: ``attrs`` of the root group, the
    root group's datasets through ``keys`` / ``read``."""

    def __init__(self, path):
        with open(path, "rb") as f:
            self.b = _Buf(f.read())
        d = self.b.d
        base = None
        for cand in (0, 512, 1024, 2048, 4096, 8192):
            if d[cand:cand + 8] == SIG:
                base = cand
                break
        if base is None:
            raise ValueError("not an HDF5 file")
        ver = d[base + 8]
        if ver in (0, 1):
            self.so, self.sl = d[base + 13], d[base + 14]
            p = base + 24 + (4 if ver == 1 else 0)
            self.base = self.b.u(p, self.so)
            p += 4 * self.so
            root_oh = self.b.u(p + self.so, self.so)     # symbol table entry: name offset, object header
        elif ver in (2, 3):
            self.so, self.sl = d[base + 9], d[base + 10]
            p = base + 12
            self.base = self.b.u(p, self.so)
            root_oh = self.b.u(p + 3 * self.so, self.so)
        else:
            raise ValueError(f"superblock version {ver} not supported")
        if self.so != 8 or self.sl != 8:
            raise ValueError("only 8-byte offsets and lengths are supported")
        self.root = self._object(root_oh)
        self.attrs = self.root["attrs"]
        self.links = self._group_links(self.root)

    # ---------------------------------------------------- object headers
    def _messages(self, addr):
        d, b = self.b.d, self.b
        a = self.base + addr
        msgs = []
        if d[a:a + 4] == b"OHDR":
            flags = d[a + 5]
            p = a + 6
            if flags & 0x20:
                p += 16
            if flags & 0x10:
                p += 4
            sz = 1 << (flags & 3)
            chunk0 = b.u(p, sz)
            p += sz
            blocks = [(p, chunk0)]
            while blocks:
                start, length = blocks.pop(0)
                q, end = start, start + length - (0 if start == p else 0)
                end = start + length
                if d[start:start + 4] == b"OCHK":
                    q = start + 4
                    end = start + length - 4
                else:
                    end = start + length
                while q + 4 <= end:
                    mtype, msize, mflags = d[q], b.u(q + 1, 2), d[q + 3]
                    q += 4
                    if flags & 0x04:
                        q += 2
                    if mtype == 0x10:
                        blocks.append((self.base + b.u(q, 8), b.u(q + 8, 8)))
                    elif mtype != 0:
                        msgs.append((mtype, q, msize))
                    q += msize
            return msgs
        if d[a] != 1:
            raise ValueError(f"object header version {d[a]} not supported")
        nmsg = b.u(a + 2, 2)
        blocks = [(a + 16, b.u(a + 8, 4))]
        while blocks and len(msgs) < nmsg:
            q, length = blocks.pop(0)
            end = q + length
            while q + 8 <= end and len(msgs) < nmsg:
                mtype, msize = b.u(q, 2), b.u(q + 2, 2)
                q += 8
                if mtype == 0x10:
                    blocks.append((self.base + b.u(q, 8), b.u(q + 8, 8)))
                msgs.append((mtype, q, msize))
                q += msize
        return msgs

    def _object(self, addr):
        obj = {"addr": addr, "attrs": {}, "msgs": []}
        for mtype, q, size in self._messages(addr):
            obj["msgs"].append((mtype, q, size))
            if mtype == 0x0C:
                name, value = self._attribute(q)
                obj["attrs"][name] = value
            elif mtype == 0x01:
                obj["shape"] = self._dataspace(q)
            elif mtype == 0x03:
                obj["dtype"] = self._datatype(q)
            elif mtype == 0x08:
                obj["layout"] = self._layout(q)
            elif mtype == 0x0B:
                obj["filters"] = self._filters(q)
        return obj

    # ---------------------------------------------------------- messages
    def _dataspace(self, q):
        d, b = self.b.d, self.b
        ver, rank, flags = d[q], d[q + 1], d[q + 2]
        if ver == 1:
            p = q + 8
        elif ver == 2:
            if d[q + 3] == 2:          # null dataspace
                return None
            p = q + 4
        else:
            raise ValueError(f"dataspace version {ver}")
        return tuple(b.u(p + 8 * i, 8) for i in range(rank))

    def _datatype(self, q):
        d, b = self.b.d, self.b
        cls = d[q] & 0x0F
        bits0 = d[q + 1]
        size = b.u(q + 4, 4)
        order = ">" if bits0 & 1 else "<"
        if cls == 0:
            signed = bool(bits0 & 0x08)
            return np.dtype(f"{order}{'i' if signed else 'u'}{size}")
        if cls == 1:
            return np.dtype(f"{order}f{size}")
        if cls == 3:
            return np.dtype(f"S{size}")
        raise ValueError(f"datatype class {cls} not supported")

    def _layout(self, q):
        d, b = self.b.d, self.b
        ver = d[q]
        if ver in (1, 2):
            rank, cls = d[q + 1], d[q + 2]
            p = q + 8
            if cls in (1, 2):
                addr = b.u(p, 8)
                p += 8
            dims = [b.u(p + 4 * i, 4) for i in range(rank)]
            p += 4 * rank
            if cls == 0:
                size = b.u(p, 4)
                return ("compact", p + 4, size)
            if cls == 1:
                return ("contiguous", addr, None)
            return ("chunked", addr, dims)
        if ver in (3, 4):
            cls = d[q + 1]
            if cls == 0:
                size = b.u(q + 2, 2)
                return ("compact", q + 4, size)
            if cls == 1:
                return ("contiguous", b.u(q + 2, 8), b.u(q + 10, 8))
            if cls == 2 and ver == 3:
                rank = d[q + 2]
                addr = b.u(q + 3, 8)
                dims = [b.u(q + 11 + 4 * i, 4) for i in range(rank)]
                return ("chunked", addr, dims)
        raise ValueError(f"data layout version {ver} class not supported")

    def _filters(self, q):
        d, b = self.b.d, self.b
        ver, n = d[q], d[q + 1]
        p = q + (8 if ver == 1 else 2)
        out = []
        for _ in range(n):
            fid = b.u(p, 2)
            if ver == 1 or fid >= 256:
                nlen = b.u(p + 2, 2)
            else:
                nlen = 0
            nval = b.u(p + 6, 2)
            p += 8
            if nlen:
                p += (nlen + 7) // 8 * 8 if ver == 1 else nlen
            vals = [b.u(p + 4 * i, 4) for i in range(nval)]
            p += 4 * nval
            if ver == 1 and nval % 2:
                p += 4
            out.append((fid, vals))
        return out

    def _attribute(self, q):
        d, b = self.b.d, self.b
        ver = d[q]
        nsize, tsize, ssize = b.u(q + 2, 2), b.u(q + 4, 2), b.u(q + 6, 2)
        if ver == 1:
            p = q + 8
            pad = lambda n: (n + 7) // 8 * 8          # noqa: E731
            name = d[p:p + nsize].split(b"\0")[0].decode()
            p += pad(nsize)
            dt = self._datatype(p)
            p += pad(tsize)
            shape = self._dataspace(p)
            p += pad(ssize)
        elif ver in (2, 3):
            p = q + 8 + (1 if ver == 3 else 0)
            name = d[p:p + nsize].split(b"\0")[0].decode()
            p += nsize
            dt = self._datatype(p)
            p += tsize
            shape = self._dataspace(p)
            p += ssize
        else:
            raise ValueError(f"attribute version {ver}")
        count = int(np.prod(shape)) if shape else 1
        arr = np.frombuffer(d, dtype=dt, count=count, offset=p)
        val = arr.reshape(shape) if shape else arr[0]
        if shape in ((), (1,), None) or count == 1:
            val = arr[0].item() if dt.kind != "S" else arr[0]
        return name, val

    # ------------------------------------------------------------- groups
    def _group_links(self, obj):
        links = {}
        for mtype, q, size in obj["msgs"]:
            if mtype == 0x11:              # symbol table: v1 B-tree + local heap
                btree, heap = self.b.u(q, 8), self.b.u(q + 8, 8)
                names = self._heap(heap)
                self._btree_group(btree, names, links)
            elif mtype == 0x06:            # link message (compact storage)
                name, addr = self._link(q)
                if addr is not None:
                    links[name] = addr
        return links

    def _heap(self, addr):
        d, b = self.b.d, self.b
        a = self.base + addr
        if d[a:a + 4] != b"HEAP":
            raise ValueError("bad local heap")
        size, data = b.u(a + 8, 8), b.u(a + 24, 8)
        seg = d[self.base + data:self.base + data + size]
        return lambda off: seg[off:seg.index(b"\0", off)].decode()

    def _btree_group(self, addr, names, links):
        d, b = self.b.d, self.b
        a = self.base + addr
        if d[a:a + 4] != b"TREE" or d[a + 4] != 0:
            raise ValueError("bad group B-tree")
        level, used = d[a + 5], b.u(a + 6, 2)
        p = a + 24 + 8                     # past the first key
        for _ in range(used):
            child = b.u(p, 8)
            p += 16                        # child address, next key
            if level > 0:
                self._btree_group(child, names, links)
            else:
                s = self.base + child
                if d[s:s + 4] != b"SNOD":
                    raise ValueError("bad symbol table node")
                n = b.u(s + 6, 2)
                for k in range(n):
                    e = s + 8 + 40 * k
                    links[names(b.u(e, 8))] = b.u(e + 8, 8)

    def _link(self, q):
        d, b = self.b.d, self.b
        flags = d[q + 1]
        p = q + 2
        ltype = 0
        if flags & 0x08:
            ltype = d[p]
            p += 1
        if flags & 0x04:
            p += 8
        if flags & 0x10:
            p += 1
        lsize = 1 << (flags & 3)
        nlen = b.u(p, lsize)
        p += lsize
        name = d[p:p + nlen].decode()
        p += nlen
        if ltype != 0:
            return name, None
        return name, b.u(p, 8)

    # ----------------------------------------------------------- datasets
    def keys(self):
        return list(self.links)

    def read(self, name):
        obj = self._object(self.links[name])
        shape, dt, lay = obj["shape"], obj["dtype"], obj["layout"]
        count = int(np.prod(shape)) if shape else 1
        d = self.b.d
        kind = lay[0]
        if kind == "compact":
            arr = np.frombuffer(d, dtype=dt, count=count, offset=lay[1])
        elif kind == "contiguous":
            if lay[1] == UNDEF:
                arr = np.zeros(count, dtype=dt)
            else:
                arr = np.frombuffer(d, dtype=dt, count=count, offset=self.base + lay[1])
        else:
            arr = self._chunked(lay[1], lay[2], shape, dt, obj.get("filters", []))
            return arr.astype(dt.newbyteorder("="))
        return arr.reshape(shape).astype(dt.newbyteorder("="))

    def read_fortran(self, name):
        """The dataset as the Fortran program that wrote or reads it indexes
        it (dimensions reversed)."""
        return self.read(name).T

    def _chunked(self, addr, cdims, shape, dt, filters):
        rank = len(shape)
        cshape = tuple(cdims[:rank])
        out = np.zeros(shape, dtype=dt)
        b, d = self.b, self.b.d

        def walk(node):
            a = self.base + node
            if d[a:a + 4] != b"TREE" or d[a + 4] != 1:
                raise ValueError("bad chunk B-tree")
            level, used = d[a + 5], b.u(a + 6, 2)
            p = a + 24
            ksize = 8 + 8 * (rank + 1)
            for _ in range(used):
                csize, mask = b.u(p, 4), b.u(p + 4, 4)
                offs = [b.u(p + 8 + 8 * i, 8) for i in range(rank)]
                child = b.u(p + ksize, 8)
                p += ksize + 8
                if level > 0:
                    walk(child)
                    continue
                raw = d[self.base + child:self.base + child + csize]
                for i, (fid, vals) in enumerate(reversed(filters)):
                    if mask & (1 << (len(filters) - 1 - i)):
                        continue
                    if fid == 1:
                        raw = zlib.decompress(raw)
                    elif fid == 2:
                        n = len(raw) // dt.itemsize
                        raw = np.frombuffer(raw, np.uint8).reshape(dt.itemsize, n).T.tobytes()
                    else:
                        raise ValueError(f"filter {fid} not supported")
                chunk = np.frombuffer(raw, dtype=dt, count=int(np.prod(cshape))).reshape(cshape)
                sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offs, cshape, shape))
                out[sl] = chunk[tuple(slice(0, x.stop - x.start) for x in sl)]

        walk(addr)
        return out


# -------------------------------------------------------------- geometry
def read_geometry(path):
    """The reference's geometry input (written by R/script/geom/geometry.py:721-738,
    read by R/src/base/init.F90:105-310): Nz, Nr and the Fortran-indexed arrays
    cell_material (Nz-1, Nr-1), epsilon_r, load (Nz-1, Nr-1), point_type and
    volume (the vacuum fraction of each node's control volume, Nz x Nr),
    hedge_type (Nz, Nr-1), vedge_type (Nz-1, Nr)."""
    f = File(path)
    out = {"Nz": int(f.attrs["Nz"]), "Nr": int(f.attrs["Nr"])}
    for k in ("cell_material", "epsilon_r", "load", "point_type", "volume", "hedge_type", "vedge_type"):
        out[k] = f.read_fortran(k)
    return out


# ---------------------------------------------------------- bundle dumps
_BUNDLE_ATTRS = ("NPar", "size", "Charge", "Mass", "Weight", "XFactor", "VFactor", "LXScaled", "LVScaled")


def read_bundle(path):
    """LoadParticleBundleHDF5, ParticleBundle.f90:178-227: the root
    attributes and the size-long datasets Z R Vz Vr Vt Az Ar WQ."""
    f = File(path)
    a = f.attrs
    npar, size = int(a["NPar"]), int(a["size"])
    po = np.empty((8, size))
    for k, name in enumerate(restart.FIELDS):
        v = f.read(name).reshape(-1)
        if v.size != size:
            raise ValueError(f"dataset {name} holds {v.size} values, size is {size}")
        po[k] = v
    return restart.BundleDump(npar, size, float(a["Charge"]), float(a["Mass"]), float(a["Weight"]),
                              float(a["XFactor"]), float(a["VFactor"]), bool(a["LXScaled"]), bool(a["LVScaled"]), po)


class _Writer:
    def __init__(self):
        self.buf = bytearray()

    def alloc(self, n, align=8):
        while len(self.buf) % align:
            self.buf += b"\0"
        at = len(self.buf)
        self.buf += b"\0" * n
        return at

    def put(self, at, data):
        self.buf[at:at + len(data)] = data


def _dt_msg(dt):
    dt = np.dtype(dt)
    if dt.kind == "f":
        # class 1, version 1; little-endian IEEE: sign at bit size*8-1
        bits = dt.itemsize * 8
        if bits == 64:
            props = struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)
            b0, b1 = 0x20, 0x3F
        else:
            props = struct.pack("<HHBBBBI", 0, 32, 23, 8, 0, 23, 127)
            b0, b1 = 0x20, 0x1F
        return bytes([0x11, b0, b1, 0]) + struct.pack("<I", dt.itemsize) + props
    if dt.kind in "iu":
        b0 = 0x08 if dt.kind == "i" else 0
        return bytes([0x10, b0, 0, 0]) + struct.pack("<I", dt.itemsize) + struct.pack("<HH", 0, dt.itemsize * 8)
    raise ValueError(f"dtype {dt} not supported")


def _ds_msg(shape):
    rank = len(shape)
    return bytes([1, rank, 0, 0, 0, 0, 0, 0]) + b"".join(struct.pack("<Q", n) for n in shape)


def _msg(mtype, body, flags=0):
    body = bytes(body)
    body += b"\0" * ((8 - len(body) % 8) % 8)
    return struct.pack("<HHB3x", mtype, len(body), flags) + body


def _attr_msg(name, value):
    if isinstance(value, (bool, np.bool_)):
        arr = np.array([1 if value else 0], dtype="<i4")
    elif isinstance(value, (int, np.integer)):
        arr = np.array([value], dtype="<i4")
    else:
        arr = np.array([value], dtype="<f8")
    nm = name.encode() + b"\0"
    dtm = _dt_msg(arr.dtype)
    dsm = bytes([1, 0, 0, 0, 0, 0, 0, 0])            # scalar dataspace
    pad = lambda x: x + b"\0" * ((8 - len(x) % 8) % 8)  # noqa: E731
    body = struct.pack("<BBHHH", 1, 0, len(nm), len(dtm), len(dsm)) + pad(nm) + pad(dtm) + pad(dsm) + arr.tobytes()
    return _msg(0x0C, body)


def _ohdr(msgs):
    body = b"".join(msgs)
    return struct.pack("<BBHII", 1, 0, len(msgs), 1, len(body)) + b"\0" * 4 + body


def _write_flat(path, datasets, attrs):
    """A superblock-0 file whose root group (one symbol table node, at most
    eight links) holds contiguous datasets (name -> C-order array) and
    scalar attributes (name, int | bool | float)."""
    if len(datasets) > 8:
        raise ValueError("at most eight datasets in the root group")
    w = _Writer()
    sb = w.alloc(96)                                   # superblock with the root symbol table entry
    # local heap of the root group: names of the datasets
    names = list(datasets)
    heap_data = bytearray(b"\0" * 8)                   # offset 0 holds the empty name
    name_off = {}
    for n in names:
        name_off[n] = len(heap_data)
        heap_data += n.encode() + b"\0"
        while len(heap_data) % 8:
            heap_data += b"\0"
    heap = w.alloc(32)
    heap_seg = w.alloc(len(heap_data))
    w.put(heap, b"HEAP" + bytes([0, 0, 0, 0]) + struct.pack("<QQQ", len(heap_data), UNDEF, heap_seg))
    w.put(heap_seg, bytes(heap_data))
    # datasets: object header + contiguous data
    ds_addr = {}
    for n in names:
        arr = np.ascontiguousarray(datasets[n])
        arr = arr.astype(arr.dtype.newbyteorder("<"), copy=False)
        data = arr.tobytes()
        data_at = w.alloc(len(data))
        w.put(data_at, data)
        fill = _msg(0x05, bytes([2, 2, 2, 0]))        # fill value: late allocation, written if set, undefined
        layout = _msg(0x08, bytes([3, 1]) + struct.pack("<QQ", data_at, len(data)))
        oh = _ohdr([_msg(0x01, _ds_msg(arr.shape)), _msg(0x03, _dt_msg(arr.dtype), flags=1), fill, layout])
        at = w.alloc(len(oh))
        w.put(at, oh)
        ds_addr[n] = at
    # symbol table node with all entries (sorted by name, as the B-tree requires)
    order = sorted(names)
    snod = w.alloc(8 + 40 * 8)
    body = b"SNOD" + bytes([1, 0]) + struct.pack("<H", len(order))
    for n in order:
        body += struct.pack("<QQI4x16x", name_off[n], ds_addr[n], 0)
    w.put(snod, body)
    # group B-tree (one leaf level node pointing at the symbol table node)
    tree = w.alloc(24 + 8 + 8 + 8 + 2 * 16 * 16)
    w.put(tree, b"TREE" + bytes([0, 0]) + struct.pack("<HQQ", 1, UNDEF, UNDEF) +
          struct.pack("<QQQ", 0, snod, name_off[order[-1]]))
    # root group object header: symbol table message and the attributes
    root = _ohdr([_msg(0x11, struct.pack("<QQ", tree, heap))] + [_attr_msg(k, v) for k, v in attrs])
    root_at = w.alloc(len(root))
    w.put(root_at, root)
    eof = len(w.buf)
    sbb = (SIG + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + struct.pack("<HHI", 4, 16, 0) +
           struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF) +
           struct.pack("<QQI4x", 0, root_at, 1) + struct.pack("<QQ", tree, heap))
    w.put(sb, sbb)
    with open(path, "wb") as f:
        f.write(bytes(w.buf))


def write_bundle(path, dump):
    """DumpParticleBundleHDF5, ParticleBundle.f90:117-148: root attributes
    NPar, size (32-bit integers), Charge, Mass, Weight, XFactor, VFactor
    (doubles), LXScaled, LVScaled (32-bit integers, 1 = true) and datasets
    Z R Vz Vr Vt Az Ar WQ of `size` doubles, in a superblock-0 file with one
    symbol-table root group."""
    po = np.ascontiguousarray(dump.po, dtype="<f8")
    if po.shape != (8, dump.size):
        raise ValueError("po must be (8, size)")
    a = dump
    attrs = [("NPar", int(a.npar)), ("size", int(a.size)), ("Charge", float(a.charge)), ("Mass", float(a.mass)),
             ("Weight", float(a.weight)), ("XFactor", float(a.xfactor)), ("VFactor", float(a.vfactor)),
             ("LXScaled", bool(a.lxscaled)), ("LVScaled", bool(a.lvscaled))]
    _write_flat(path, {n: po[k] for k, n in enumerate(restart.FIELDS)}, attrs)


_GEOM_FIELDS = (("cell_material", "<i4", -1, -1), ("epsilon_r", "<f8", -1, -1), ("load", "<i4", -1, -1),
                ("point_type", "<i4", 0, 0), ("volume", "<f8", 0, 0), ("hedge_type", "<i4", 0, -1),
                ("vedge_type", "<i4", -1, 0))


def write_geometry(path, geom):
    """A geometry input as the reference reads it (R/src/base/init.F90:105-310,
    written there by R/script/geom/geometry.py:721-738): root attributes Nz,
    Nr and the datasets of ``read_geometry``, given as Fortran-indexed arrays
    (cell_material holds 0-based indices into the materials list) and stored
    with their dimensions reversed, as a Fortran reader expects."""
    nz, nr = int(geom["Nz"]), int(geom["Nr"])
    ds = {}
    for name, dt, dz, dr in _GEOM_FIELDS:
        a = np.asarray(geom[name])
        if a.shape != (nz + dz, nr + dr):
            raise ValueError(f"{name} is {a.shape}, expected {(nz + dz, nr + dr)}")
        ds[name] = np.ascontiguousarray(a.T).astype(dt)
    _write_flat(path, ds, [("Nz", nz), ("Nr", nr)])
