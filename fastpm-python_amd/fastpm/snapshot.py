"""Snapshots in the bigfile directory layout (reference: fastpm/core.py:65-89,
which writes through ``bigfile.FileMPI``).

A snapshot is a directory; every block is a sub-directory named after the
block (``Header``, ``1/Position``, ...) holding

* ``header``: ``DTYPE: <f4`` / ``NMEMB: 3`` / ``NFILE: 1`` lines, then one
  line per data file, ``%06X: <rows> : <byte sum> : <folded sum>``, where the
  byte sum is the sum of the file's bytes modulo 2^32 and the folded sum is
  the System V 16-bit fold of it;
* the data files ``000000``, ...: rows in C order, ``NMEMB`` items of
  ``DTYPE`` each, in the byte order the dtype names;
* ``attr-v2`` (when the block has attributes): one line per attribute,
  ``<name> <dtype> <nmemb> <hex>``, the value's bytes as upper-case hex
  digits; strings are ``S1`` arrays of their characters.

bigfile itself is not importable here, so the layout follows the library's
published on-disk format and byte compatibility with it is not checked by a
test; the reader below reads what the writer writes (and any single- or
multi-file block of plain numeric dtypes in that layout).  The writer runs on
the host: device arrays are copied back first.
"""
import os

import numpy as np

HEADER = "header"
ATTR = "attr-v2"


def _dtype_str(dt):
    dt = np.dtype(dt)
    if dt.kind == "S":
        return "S1"
    s = dt.str
    return ("<" + s[1:]) if s[0] == "|" else s


def _fold(s):
    r = (s & 0xFFFF) + ((s & 0xFFFFFFFF) >> 16)
    return (r & 0xFFFF) + (r >> 16)


def byte_sum(buf):
    """Sum of the bytes of `buf` modulo 2^32 (the data files' checksum)."""
    b = np.frombuffer(memoryview(buf).cast("B"), dtype=np.uint8)
    return int(b.sum(dtype=np.uint64) & 0xFFFFFFFF)


def _encode_attr(value):
    if isinstance(value, (bytes, str)):
        raw = value.encode() if isinstance(value, str) else value
        return "S1", len(raw), raw
    a = np.atleast_1d(np.asarray(value))
    if a.dtype.kind in "U":
        raw = "".join(a.ravel().tolist()).encode()
        return "S1", len(raw), raw
    if a.dtype.kind == "b":
        a = a.astype("i1")
    if a.dtype.kind not in "iufc":
        raise TypeError("attribute of dtype %s" % a.dtype)
    a = np.ascontiguousarray(a.ravel())
    return _dtype_str(a.dtype), a.size, a.tobytes()


def _decode_attr(dtype, nmemb, hexdata):
    raw = bytes.fromhex(hexdata)
    if dtype == "S1":
        return raw.decode()
    a = np.frombuffer(raw, dtype=np.dtype(dtype), count=nmemb)
    return a.copy()


class Block:
    """One block of a snapshot directory (data, attributes, or both)."""

    def __init__(self, root, name):
        self.path = os.path.join(root, *name.split("/"))
        self.name = name

    def write(self, array=None, attrs=None, nfile=1):
        """Writes the block: `array` ([N] or [N, nmemb], any numeric dtype)
        split into `nfile` files of about equal row counts, or a data-less
        block (``DTYPE: <i8``, ``NFILE: 0``) when `array` is None."""
        os.makedirs(self.path, exist_ok=True)
        lines = []
        if array is None:
            dtype, nmemb, sizes, sums = "<i8", 1, [], []
        else:
            a = np.ascontiguousarray(array)
            if a.ndim == 1:
                a = a.reshape(-1, 1)
            elif a.ndim != 2:
                raise ValueError("a block holds [N] or [N, nmemb] arrays, got %s" % (a.shape,))
            dtype, nmemb = _dtype_str(a.dtype), a.shape[1]
            nfile = max(1, int(nfile))
            cuts = [len(a) * i // nfile for i in range(nfile + 1)]
            sizes, sums = [], []
            for i in range(nfile):
                part = a[cuts[i]:cuts[i + 1]]
                buf = part.tobytes()
                with open(os.path.join(self.path, "%06X" % i), "wb") as f:
                    f.write(buf)
                sizes.append(len(part))
                sums.append(byte_sum(buf))
        lines.append("DTYPE: %s" % dtype)
        lines.append("NMEMB: %d" % nmemb)
        lines.append("NFILE: %d" % len(sizes))
        for i, (n, s) in enumerate(zip(sizes, sums)):
            lines.append("%06X: %d : %d : %d" % (i, n, s, _fold(s)))
        with open(os.path.join(self.path, HEADER), "w") as f:
            f.write("\n".join(lines) + "\n")
        if attrs:
            out = []
            for key, value in attrs.items():
                if " " in key:
                    raise ValueError("attribute names cannot hold spaces: %r" % key)
                dt, n, raw = _encode_attr(value)
                out.append("%s %s %d %s" % (key, dt, n, raw.hex().upper()))
            with open(os.path.join(self.path, ATTR), "w") as f:
                f.write("\n".join(out) + "\n")

    def header(self):
        """(dtype, nmemb, [(rows, byte sum, folded sum)] per file)."""
        with open(os.path.join(self.path, HEADER)) as f:
            lines = [ln.strip() for ln in f if ln.strip()]
        kv = {}
        files = []
        for ln in lines:
            key, _, rest = ln.partition(":")
            if key in ("DTYPE", "NMEMB", "NFILE"):
                kv[key] = rest.strip()
            else:
                n, s, c = (int(v) for v in rest.split(":"))
                files.append((n, s, c))
        if len(files) != int(kv["NFILE"]):
            raise ValueError("%s: NFILE %s but %d file lines" % (self.name, kv["NFILE"], len(files)))
        return kv["DTYPE"], int(kv["NMEMB"]), files

    def read(self, check=True):
        """The block's rows as an [N, nmemb] array ([N] when nmemb is 1);
        with `check`, every file's size and byte sum against the header."""
        dtype, nmemb, files = self.header()
        dt = np.dtype(dtype)
        parts = []
        for i, (n, s, c) in enumerate(files):
            with open(os.path.join(self.path, "%06X" % i), "rb") as f:
                buf = f.read()
            if check:
                if len(buf) != n * nmemb * dt.itemsize:
                    raise ValueError("%s/%06X: %d bytes, header says %d rows" % (
                        self.name, i, len(buf), n))
                got = byte_sum(buf)
                if got != s or _fold(got) != c:
                    raise ValueError("%s/%06X: checksum %d, header says %d" % (self.name, i, got, s))
            parts.append(np.frombuffer(buf, dtype=dt).reshape(-1, nmemb))
        a = np.concatenate(parts) if parts else np.zeros((0, nmemb), dtype=dt)
        return a[:, 0].copy() if nmemb == 1 else a

    def attrs(self):
        path = os.path.join(self.path, ATTR)
        out = {}
        if not os.path.exists(path):
            return out
        with open(path) as f:
            for ln in f:
                parts = ln.split()
                if not parts:
                    continue
                name, dt, n = parts[0], parts[1], int(parts[2])
                out[name] = _decode_attr(dt, n, parts[3] if len(parts) > 3 else "")
        return out


def write_snapshot(path, header_attrs, blocks, nfile=1):
    """A snapshot directory: the ``Header`` block's attributes and data
    blocks {name: array}."""
    os.makedirs(path, exist_ok=True)
    Block(path, "Header").write(None, header_attrs)
    for name, arr in blocks.items():
        Block(path, name).write(arr, nfile=nfile)


def read_snapshot(path, names):
    """(Header attributes, {name: array}) of a snapshot directory."""
    head = Block(path, "Header").attrs()
    return head, {name: Block(path, name).read() for name in names}
