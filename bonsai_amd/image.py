"""Image::write (src/io/image.rs:47-53) for film output: RGB float32 images as
single-part scanline OpenEXR files (FLOAT channels B, G, R, no compression),
written by the library's C++ encoder."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


def _rgb(image: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(image, np.float32)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("image must have shape (height, width, 3)")
    return a


def exr_bytes(image: np.ndarray) -> bytes:
    """The OpenEXR file image of an (h, w, 3) float32 array, top row first."""
    a = _rgb(image)
    h, w = a.shape[:2]
    size = C.c_uint64(0)
    N.check(N.lib().bg_exr_encode(w, h, a.ctypes.data, None, C.byref(size)))
    out = (C.c_ubyte * size.value)()
    N.check(N.lib().bg_exr_encode(w, h, a.ctypes.data, out, C.byref(size)))
    return bytes(out)


def write_exr(path: str, image: np.ndarray) -> None:
    a = _rgb(image)
    h, w = a.shape[:2]
    N.check(N.lib().bg_exr_write(str(path).encode(), w, h, a.ctypes.data))
