"""A small HDF5 writer for the drivers' checkpoint files, used when h5py is not
importable.

The reference writes its checkpoints with h5py: groups `Checkpoints`,
`Scalar_Data`, `Parameters` and `Bifurcation` of numeric datasets
(Main.py:301-321, 1021-1040), read back by Plot_Tools.py.  This module writes
the same tree in the current HDF5 file format (as h5py does with
libver="latest"):

* superblock version 2, 8-byte offsets and lengths;
* every object a version-2 object header ("OHDR") with a Jenkins lookup3
  checksum (the checksum HDF5 uses for its metadata);
* groups hold their links compactly: a Link Info message (no fractal heap,
  no name index), a Group Info message and one hard-link message per child;
* datasets: a dataspace (scalar or simple), a datatype, a fill-value message
  and a contiguous data layout pointing at the raw data (no storage for an
  empty dataset).

Values map as h5py maps them: floats to little-endian IEEE float64, integers
to little-endian int64, booleans to h5py's enum of int8 {FALSE: 0, TRUE: 1},
sequences to arrays of those.  Strings become fixed-length UTF-8 strings
(h5py would write variable-length ones, which need a global heap; h5py reads
these back as bytes).
"""
import struct

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
_M = 0xFFFFFFFF


def _rot(x, k):
    return ((x << k) | (x >> (32 - k))) & _M


def lookup3(data, initval=0):
    """Bob Jenkins' lookup3 hashlittle over bytes, as HDF5 checksums its
    metadata (byte-wise form: the value does not depend on alignment)."""
    k = bytes(data)
    n = len(k)
    a = b = c = (0xDEADBEEF + n + initval) & _M
    i = 0
    while n > 12:
        a = (a + int.from_bytes(k[i:i + 4], "little")) & _M
        b = (b + int.from_bytes(k[i + 4:i + 8], "little")) & _M
        c = (c + int.from_bytes(k[i + 8:i + 12], "little")) & _M
        a = (a - c) & _M; a ^= _rot(c, 4); c = (c + b) & _M
        b = (b - a) & _M; b ^= _rot(a, 6); a = (a + c) & _M
        c = (c - b) & _M; c ^= _rot(b, 8); b = (b + a) & _M
        a = (a - c) & _M; a ^= _rot(c, 16); c = (c + b) & _M
        b = (b - a) & _M; b ^= _rot(a, 19); a = (a + c) & _M
        c = (c - b) & _M; c ^= _rot(b, 4); b = (b + a) & _M
        n -= 12
        i += 12
    if n == 0:
        return c
    tail = k[i:i + n] + bytes(12 - n)
    a = (a + int.from_bytes(tail[0:4], "little")) & _M
    b = (b + int.from_bytes(tail[4:8], "little")) & _M
    c = (c + int.from_bytes(tail[8:12], "little")) & _M
    c ^= b; c = (c - _rot(b, 14)) & _M
    a ^= c; a = (a - _rot(c, 11)) & _M
    b ^= a; b = (b - _rot(a, 25)) & _M
    c ^= b; c = (c - _rot(b, 16)) & _M
    a ^= c; a = (a - _rot(c, 4)) & _M
    b ^= a; b = (b - _rot(a, 14)) & _M
    c ^= b; c = (c - _rot(b, 24)) & _M
    return c


# ---------------------------------------------------------------------------
# messages
# ---------------------------------------------------------------------------
def _msg(mtype, body, flags=0):
    return struct.pack("<BHB", mtype, len(body), flags) + body


def _object_header(messages):
    body = b"".join(messages)
    head = b"OHDR" + struct.pack("<BBI", 2, 0x02, len(body))   # 4-byte chunk size
    blob = head + body
    return blob + struct.pack("<I", lookup3(blob))


def _dataspace(shape):
    if shape == ():
        return _msg(0x01, struct.pack("<BBBB", 2, 0, 0, 0))
    return _msg(0x01, struct.pack("<BBBB", 2, len(shape), 0, 1)
                + b"".join(struct.pack("<Q", int(d)) for d in shape))


def _float64_type():
    # class 1 (floating point), version 1; little endian, implied mantissa
    # normalisation, sign at bit 63; offset 0, precision 64, exponent at 52
    # (11 bits), mantissa at 0 (52 bits), bias 1023
    return (bytes([0x11, 0x20, 63, 0]) + struct.pack("<I", 8)
            + struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023))


def _int_type(size, signed=True):
    return (bytes([0x10, 0x08 if signed else 0x00, 0, 0]) + struct.pack("<I", size)
            + struct.pack("<HH", 0, 8 * size))


def _bool_type():
    # h5py's boolean: enum (class 8, version 1) of int8 with FALSE = 0, TRUE = 1;
    # names null-terminated and padded to multiples of eight bytes
    names = b"FALSE\0\0\0" + b"TRUE\0\0\0\0"
    return (bytes([0x18, 2, 0, 0]) + struct.pack("<I", 1) + _int_type(1) + names
            + bytes([0, 1]))


def _string_type(size):
    # class 3 (string), version 1: null padded, UTF-8
    return bytes([0x13, 0x01 | (1 << 4), 0, 0]) + struct.pack("<I", size)


def _fill_value():
    # version 3: space allocated late, fill value written if set, none set
    return _msg(0x05, struct.pack("<BB", 3, 2 | (2 << 2)))


def _layout(addr, size):
    return _msg(0x08, struct.pack("<BBQQ", 3, 1, addr, size))


def _encode(value):
    """(datatype message body, shape, raw little-endian bytes) of a value."""
    if isinstance(value, (str, bytes)):
        raw = value.encode("utf-8") if isinstance(value, str) else bytes(value)
        size = max(len(raw), 1)
        return _string_type(size), (), raw.ljust(size, b"\0")
    arr = np.asarray(value)
    if arr.dtype == object:
        raise TypeError("cannot store an object array in HDF5")
    if arr.dtype == np.bool_:
        return _bool_type(), arr.shape, arr.astype(np.int8).tobytes()
    if np.issubdtype(arr.dtype, np.integer):
        return _int_type(8), arr.shape, arr.astype("<i8").tobytes()
    if np.issubdtype(arr.dtype, np.floating):
        return _float64_type(), arr.shape, arr.astype("<f8").tobytes()
    raise TypeError("unsupported value type %s" % arr.dtype)


def _link(name, addr):
    raw = name.encode("ascii")
    if not 0 < len(raw) < 256:
        raise ValueError("link names must be 1-255 ASCII characters")
    # version 1, flags 0: one-byte name length, hard link, ASCII
    return _msg(0x06, struct.pack("<BBB", 1, 0, len(raw)) + raw + struct.pack("<Q", addr))


class _File:
    def __init__(self):
        self.blobs = [bytes(48)]   # the superblock, written last
        self.end = 48

    def put(self, blob, align=8):
        pad = (-self.end) % align
        if pad:
            self.blobs.append(bytes(pad))
            self.end += pad
        addr = self.end
        self.blobs.append(blob)
        self.end += len(blob)
        return addr

    def dataset(self, value):
        dtype, shape, raw = _encode(value)
        addr = self.put(raw) if raw else UNDEF
        return self.put(_object_header([_dataspace(shape), _msg(0x03, dtype, 1), _fill_value(),
                                        _layout(addr, len(raw))]))

    def group(self, tree):
        links = []
        for name, value in tree.items():
            addr = self.group(value) if isinstance(value, dict) else self.dataset(value)
            links.append(_link(name, addr))
        info = _msg(0x02, struct.pack("<BBQQ", 0, 0, UNDEF, UNDEF))
        return self.put(_object_header([info, _msg(0x0A, struct.pack("<BB", 0, 0))] + links))


def write(path, tree):
    """Write `tree` (a dict: names to sub-dicts (groups) or values (datasets))
    as an HDF5 file at `path`."""
    f = _File()
    root = f.group(tree)
    sb = (b"\x89HDF\r\n\x1a\n" + struct.pack("<BBBB", 2, 8, 8, 0)
          + struct.pack("<QQQQ", 0, UNDEF, f.end, root))
    f.blobs[0] = sb + struct.pack("<I", lookup3(sb))
    with open(path, "wb") as fh:
        for blob in f.blobs:
            fh.write(blob)
    return path


# ---------------------------------------------------------------------------
# reading back what write() produces (and files of the same layout)
# ---------------------------------------------------------------------------
def _u(data, off, n):
    return int.from_bytes(data[off:off + n], "little")


def _decode_type(body):
    cls, size = body[0] & 0x0F, _u(body, 4, 4)
    if cls == 1 and size in (4, 8):
        return np.dtype("<f%d" % size)
    if cls == 0:
        return np.dtype(("<i%d" if body[1] & 0x08 else "<u%d") % size)
    if cls == 8 and size == 1 and _u(body, 1, 2) == 2 and b"FALSE" in body and b"TRUE" in body:
        return np.dtype(np.bool_)
    if cls == 3:
        return np.dtype("S%d" % size)
    raise ValueError("unsupported HDF5 datatype class %d" % cls)


def _read_object(data, addr):
    if data[addr:addr + 4] != b"OHDR" or data[addr + 4] != 2:
        raise ValueError("expected a version-2 object header at %d" % addr)
    flags = data[addr + 5]
    p = addr + 6
    if flags & 0x20:
        p += 16                      # access, modification, change, birth times
    if flags & 0x10:
        p += 4                       # attribute phase change values
    nsz = 1 << (flags & 0x03)
    chunk = _u(data, p, nsz)
    p += nsz
    end = p + chunk
    if lookup3(data[addr:end]) != _u(data, end, 4):
        raise ValueError("object header checksum mismatch at %d" % addr)
    links, space, dtype, layout = {}, None, None, None
    while p + 4 <= end:
        mtype, msize = data[p], _u(data, p + 1, 2)
        p += 4 + (2 if flags & 0x04 else 0)
        body = data[p:p + msize]
        p += msize
        if mtype == 0x06:
            lflags = body[1]
            q = 2 + (1 if lflags & 0x08 else 0) + (8 if lflags & 0x04 else 0)
            q += 1 if lflags & 0x10 else 0
            nlen_size = 1 << (lflags & 0x03)
            nlen = _u(body, q, nlen_size)
            q += nlen_size
            name = body[q:q + nlen].decode("utf-8")
            if lflags & 0x08 and body[2] != 0:
                raise ValueError("only hard links are supported")
            links[name] = _u(body, q + nlen, 8)
        elif mtype == 0x01:
            rank, stype = body[1], body[3]
            space = () if stype == 0 else tuple(_u(body, 4 + 8 * i, 8) for i in range(rank))
        elif mtype == 0x03:
            dtype = _decode_type(body)
        elif mtype == 0x08:
            # versions 3 and 4: contiguous (address, size) or compact (inline)
            if body[0] not in (3, 4) or body[1] not in (0, 1):
                raise ValueError("only contiguous and compact layouts are supported")
            if body[1] == 1:
                layout = (_u(body, 2, 8), _u(body, 10, 8))
            else:
                layout = body[4:4 + _u(body, 2, 2)]
    if dtype is None:
        return {name: _read_object(data, a) for name, a in links.items()}
    if isinstance(layout, tuple):
        a, n = layout
        raw = b"" if n == 0 else data[a:a + n]
    else:
        raw = layout
    arr = np.frombuffer(raw, dtype=np.int8 if dtype == np.bool_ else dtype).copy()
    arr = arr.astype(np.bool_) if dtype == np.bool_ else arr
    return arr.reshape(space)


def read(path):
    """The tree of a file in the layout write() produces: nested dicts of
    NumPy arrays (0-d for scalars)."""
    with open(path, "rb") as fh:
        data = fh.read()
    # superblock versions 2 and 3 share this layout
    if data[:8] != b"\x89HDF\r\n\x1a\n" or data[8] not in (2, 3) or data[9] != 8 or data[10] != 8:
        raise ValueError("not an HDF5 file with a version-2 or -3 superblock")
    if lookup3(data[:44]) != _u(data, 44, 4):
        raise ValueError("superblock checksum mismatch")
    return _read_object(data, _u(data, 36, 8))
