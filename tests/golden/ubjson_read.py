"""Minimal UBJSON (Universal Binary JSON, draft 12) decoder.

Used only to read the reference's test fixtures (``*.ubjson`` files written by
nlohmann::json::to_ubjson) as plain data; it executes nothing from the file.
Supports the value markers Z N T F i U I l L d D C S H [ { and the optimised
container forms with ``$`` (type) and ``#`` (count).
"""
import struct

_INT = {b"i": (">b", 1), b"U": (">B", 1), b"I": (">h", 2), b"l": (">i", 4),
        b"L": (">q", 8), b"d": (">f", 4), b"D": (">d", 8)}


class _Reader:
    def __init__(self, buf):
        self.buf = buf
        self.pos = 0

    def take(self, n):
        out = self.buf[self.pos:self.pos + n]
        if len(out) != n:
            raise ValueError("truncated UBJSON")
        self.pos += n
        return out

    def marker(self):
        m = self.take(1)
        while m == b"N":  # no-op
            m = self.take(1)
        return m

    def length(self):
        m = self.marker()
        fmt, n = _INT[m]
        return struct.unpack(fmt, self.take(n))[0]

    def string(self):
        return self.take(self.length()).decode("utf-8")

    def value(self, m=None):
        if m is None:
            m = self.marker()
        if m == b"Z":
            return None
        if m == b"T":
            return True
        if m == b"F":
            return False
        if m in _INT:
            fmt, n = _INT[m]
            return struct.unpack(fmt, self.take(n))[0]
        if m == b"C":
            return self.take(1).decode("utf-8")
        if m == b"S" or m == b"H":
            return self.string()
        if m == b"[":
            return self.container(array=True)
        if m == b"{":
            return self.container(array=False)
        raise ValueError("unknown UBJSON marker %r at %d" % (m, self.pos))

    def container(self, array):
        typ = None
        count = None
        m = self.marker()
        if m == b"$":
            typ = self.take(1)
            m = self.marker()
        if m == b"#":
            count = self.length()
        else:
            self.pos -= 1
        if array:
            out = []
            if count is not None:
                if typ is not None and typ in _INT:
                    fmt, n = _INT[typ]
                    raw = self.take(n * count)
                    return list(struct.unpack(">" + fmt[1] * count, raw))
                for _ in range(count):
                    out.append(self.value(typ))
                return out
            while True:
                m = self.marker()
                if m == b"]":
                    return out
                out.append(self.value(m))
        out = {}
        if count is not None:
            for _ in range(count):
                k = self.string()
                out[k] = self.value(typ)
            return out
        while True:
            m = self.marker()
            if m == b"}":
                return out
            self.pos -= 1
            k = self.string()
            out[k] = self.value()


def load(path):
    with open(path, "rb") as fh:
        return _Reader(fh.read()).value()
