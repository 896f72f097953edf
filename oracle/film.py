"""Oracle (test infrastructure only): Film::write_image (src/base/film.rs:156-202,
with merge_film_tile's RGB -> XYZ step, film.rs:138-154, and the matrices of
spectrum.rs:34-44) in NumPy float32 arithmetic, and an OpenEXR reader written
from the published OpenEXR file layout, for checking the library's GPU resolve
and its EXR writer.

Pinned by tests/test_film_exr.py: the XYZ round trip of the reference's
matrices is the identity to float32 rounding (their product is I to 1e-6),
a pixel with weight w and sum w * c resolves to c, negative channels clamp to
0 and NaN to 0 (f32::max), weight-0 pixels pass through unnormalised, and the
reader decodes a file assembled byte by byte in the test from the format's
field definitions."""
from __future__ import annotations

import struct

import numpy as np

f32 = np.float32
_TO_XYZ = [[0.412453, 0.357580, 0.180423], [0.212671, 0.715160, 0.072169], [0.019334, 0.119193, 0.950227]]
_TO_RGB = [[3.240479, -1.537150, -0.498535], [-0.969256, 1.875991, 0.041556], [0.055648, -0.204043, 1.057311]]


def _mat3(m, a, b, c):
    # m[0] * a + m[1] * b + m[2] * c, left to right, each step rounded to float32
    # (spectrum.rs:34-44); the source's "x - k * y" equals "x + (-k) * y" exactly
    # in IEEE arithmetic, so the rows are kept with signed coefficients
    return (f32(m[0]) * a + f32(m[1]) * b) + f32(m[2]) * c


def write_image(sums: np.ndarray, splat_scale: float = 1.0, scale: float = 1.0) -> np.ndarray:
    """sums: (h, w, 4) float32 (r, g, b, weight). Returns the (h, w, 3) float32 image."""
    s = np.asarray(sums, np.float32)
    r, g, b, w = s[..., 0], s[..., 1], s[..., 2], s[..., 3]
    with np.errstate(all="ignore"):
        xyz = [f32(0) + _mat3(_TO_XYZ[k], r, g, b) for k in range(3)]  # merge into Pixel::default()
        rgb = [_mat3(_TO_RGB[k], *xyz) for k in range(3)]
        nz = w != 0
        inv = np.where(nz, f32(1) / np.where(nz, w, f32(1)), f32(1))
        rgb = [np.where(nz, np.fmax(c * inv, f32(0)), c) for c in rgb]
        zero = np.zeros_like(r)
        splat = [_mat3(_TO_RGB[k], zero, zero, zero) for k in range(3)]
        rgb = [(c + f32(splat_scale) * sp) * f32(scale) for c, sp in zip(rgb, splat)]
    return np.stack(rgb, axis=-1).astype(np.float32)


# ----------------------------------------------------------------- OpenEXR
class ExrError(ValueError):
    pass


_PIXEL_TYPES = {0: ("<u4", 4), 1: ("<f2", 2), 2: ("<f4", 4)}


def _cstr(buf: bytes, at: int) -> tuple[str, int]:
    end = buf.index(b"\0", at)
    return buf[at:end].decode("latin-1"), end + 1


def read_exr(data: bytes) -> dict:
    """Decodes a single-part scanline OpenEXR file without compression. Returns
    {"header": {name: (type, raw bytes)}, "channels": [(name, pixel type, xs, ys)],
    "width", "height", "pixels": {channel: (h, w) array}}."""
    if len(data) < 8:
        raise ExrError("short file")
    magic, version = struct.unpack_from("<II", data, 0)
    if magic != 20000630:
        raise ExrError("not an OpenEXR file")
    if version & 0xFF != 2:
        raise ExrError("unsupported version")
    if version & ~0xFF:
        raise ExrError("tiled, long-name, deep or multi-part files are not supported")
    at = 8
    header = {}
    while True:
        if data[at] == 0:
            at += 1
            break
        name, at = _cstr(data, at)
        typ, at = _cstr(data, at)
        (size,) = struct.unpack_from("<i", data, at)
        at += 4
        header[name] = (typ, data[at:at + size])
        at += size
    for req in ("channels", "compression", "dataWindow", "displayWindow", "lineOrder", "pixelAspectRatio",
                "screenWindowCenter", "screenWindowWidth"):
        if req not in header:
            raise ExrError(f"missing required attribute {req}")
    if header["compression"][1][0] != 0:
        raise ExrError("compressed files are not supported")
    chl = header["channels"][1]
    channels = []
    k = 0
    while chl[k] != 0:
        cname, k = _cstr(chl, k)
        ptype, plinear, xs, ys = struct.unpack_from("<iB3xii", chl, k)
        k += 16
        channels.append((cname, ptype, xs, ys))
    if [c[0] for c in channels] != sorted(c[0] for c in channels):
        raise ExrError("channel list not sorted")
    if any(c[2] != 1 or c[3] != 1 for c in channels):
        raise ExrError("subsampled channels are not supported")
    x0, y0, x1, y1 = struct.unpack("<iiii", header["dataWindow"][1])
    width, height = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from(f"<{height}Q", data, at)
    pixels = {c[0]: np.empty((height, width), _PIXEL_TYPES[c[1]][0]) for c in channels}
    line_bytes = sum(_PIXEL_TYPES[c[1]][1] for c in channels) * width
    for off in offsets:
        y, size = struct.unpack_from("<ii", data, off)
        if size != line_bytes:
            raise ExrError("unexpected scanline size")
        p = off + 8
        row = y - y0
        for cname, ptype, _, _ in channels:
            dt, nb = _PIXEL_TYPES[ptype]
            pixels[cname][row] = np.frombuffer(data, dt, width, p)
            p += nb * width
    return {"header": header, "channels": channels, "width": width, "height": height, "pixels": pixels}
