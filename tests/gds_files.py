"""Small GDSII stream files written for tests (known coordinates), and the encoder of the
format's 8-byte real (excess-64 base-16 exponent, 56-bit mantissa)."""
import math
import struct


def gdsii_real_bytes(v: float) -> bytes:
    if v == 0:
        return bytes(8)
    sign = 0x80 if v < 0 else 0
    v = abs(v)
    e = 64
    while v >= 1.0:
        v /= 16.0
        e += 1
    while v < 1.0 / 16.0:
        v *= 16.0
        e -= 1
    m = int(round(v * 2 ** 56))
    if m >= 2 ** 56:  # rounding carried into the next hex digit
        m //= 16
        e += 1
    return bytes([sign | e]) + m.to_bytes(7, "big")


def _rec(rtype: int, dtype: int, body: bytes = b"") -> bytes:
    if len(body) % 2:
        body += b"\0"
    return struct.pack(">HBB", 4 + len(body), rtype, dtype) + body


def _xy(points):
    return _rec(0x10, 3, b"".join(struct.pack(">ii", x, y) for x, y in points))


def write_gds(path, cells, db_in_user=1e-3, db_in_m=1e-9):
    """cells: list of (name, elements); element = ("boundary"|"box"|"path", [(x, y) int points
    without the closing point]). Boundaries and boxes are written closed, as GDSII requires."""
    out = _rec(0x00, 2, struct.pack(">h", 600)) + _rec(0x01, 2, bytes(24)) + _rec(0x02, 6, b"LIB")
    out += _rec(0x03, 5, gdsii_real_bytes(db_in_user) + gdsii_real_bytes(db_in_m))
    for name, elems in cells:
        out += _rec(0x05, 2, bytes(24)) + _rec(0x06, 6, name.encode())
        for kind, pts in elems:
            if kind == "path":
                out += _rec(0x09, 0) + _rec(0x0D, 2, struct.pack(">h", 1)) + _rec(0x0E, 2, struct.pack(">h", 0))
                out += _rec(0x0F, 3, struct.pack(">i", 100)) + _xy(pts) + _rec(0x11, 0)
            else:
                rt = 0x08 if kind == "boundary" else 0x2D
                out += _rec(rt, 0) + _rec(0x0D, 2, struct.pack(">h", 1))
                out += _rec(0x0E if kind == "boundary" else 0x2E, 2, struct.pack(">h", 0))
                out += _xy(list(pts) + [pts[0]]) + _rec(0x11, 0)
        out += _rec(0x07, 0)
    out += _rec(0x04, 0)
    with open(path, "wb") as f:
        f.write(out)


def synthetic_cells():
    """One rectangle, one trapezoid with a sloped edge, one box (first cell), a path (not a
    polygon) and a second cell (not read by the sampler). Database unit 1 nm."""
    rect = [(1000, -500), (1600, -500), (1600, 700), (1000, 700)]
    trap = [(2000, -400), (2600, -400), (2300, 600), (2000, 600)]
    box = [(3000, -200), (3300, -200), (3300, 200), (3000, 200)]
    path = [(0, 0), (5000, 0)]
    other = [(0, 0), (100, 0), (100, 100)]
    return [("TOP", [("boundary", rect), ("boundary", trap), ("box", box), ("path", path)]),
            ("SECOND", [("boundary", other)])]
