"""Film::write_image (film.rs:156-202) and Image::write (io/image.rs:47-53):
the oracle's float32 restatement of the resolve pinned by closed forms, its
OpenEXR reader pinned by a file assembled here byte by byte from the format's
field definitions, and the library's EXR encoder (host code, no device) against
both."""
import struct
import sys
from pathlib import Path

import numpy as np
import pytest

from bonsai_amd import _native as N
from bonsai_amd.image import exr_bytes, write_exr

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import film as F  # noqa: E402

XYZ = np.array([[0.412453, 0.357580, 0.180423], [0.212671, 0.715160, 0.072169], [0.019334, 0.119193, 0.950227]])
RGB = np.array([[3.240479, -1.537150, -0.498535], [-0.969256, 1.875991, 0.041556], [0.055648, -0.204043, 1.057311]])


def sums_of(colors, weights):
    c = np.asarray(colors, np.float32)
    w = np.asarray(weights, np.float32)
    out = np.zeros(c.shape[:-1] + (4,), np.float32)
    out[..., :3] = c * w[..., None]
    out[..., 3] = w
    return out


def test_reference_matrices_are_inverse():
    # the resolve relies on the two matrices of spectrum.rs:34-44 cancelling; a transposed or
    # mistyped row breaks this
    np.testing.assert_allclose(RGB @ XYZ, np.eye(3), atol=2e-6)


def test_weighted_mean_is_recovered():
    rng = np.random.default_rng(1)
    c = rng.uniform(0.0, 5.0, (7, 9, 3)).astype(np.float32)
    w = rng.integers(1, 64, (7, 9)).astype(np.float32)
    img = F.write_image(sums_of(c, w))
    assert img.dtype == np.float32
    # the round trip's float32 cancellation costs a few ulps of the pixel's brightest channel
    tol = 2e-6 * c.max(axis=-1, keepdims=True) * 8
    assert (np.abs(img - c) <= tol).all()
    # the float32 steps agree with the same arithmetic in float64 to rounding
    s = sums_of(c, w).astype(np.float64)
    ref = np.maximum((s[..., :3] @ XYZ.T) @ RGB.T / s[..., 3:], 0)
    assert (np.abs(img - ref) <= tol).all()


def test_clamp_nan_zero_weight_and_scale():
    s = np.zeros((1, 5, 4), np.float32)
    s[0, 0] = [-1.0, -2.0, -3.0, 2.0]            # negative: clamped to 0 (weight non-zero)
    s[0, 1] = [np.nan, 1.0, 1.0, 1.0]            # NaN propagates through XYZ; f32::max makes it 0
    s[0, 2] = [0.5, 0.25, 0.125, 0.0]            # weight 0: the XYZ round trip, not normalised
    s[0, 3] = [-0.5, 0.25, 0.125, 0.0]           # weight 0: no clamp either
    s[0, 4] = [3.0, 3.0, 3.0, 3.0]
    img = F.write_image(s)
    assert (img[0, 0] == 0).all()
    assert (img[0, 1] == 0).all()
    np.testing.assert_allclose(img[0, 2], [0.5, 0.25, 0.125], rtol=1e-5)
    assert img[0, 3, 0] < 0
    np.testing.assert_allclose(img[0, 4], 1.0, rtol=1e-5)
    np.testing.assert_array_equal(F.write_image(s, scale=2.0), img * np.float32(2.0))
    # the splat is zero, so any finite splat_scale leaves the image unchanged; an infinite one
    # multiplies zero into NaN, as the reference's arithmetic would
    np.testing.assert_array_equal(F.write_image(s, splat_scale=7.0), img)
    assert np.isnan(F.write_image(s, splat_scale=np.inf)).all()


def hand_made_exr(width, height, rgb):
    """An OpenEXR file written field by field from the format definition: magic, version,
    attributes (name\\0 type\\0 size value) in name order, a zero byte, the scanline offset table,
    and per scanline (y, byte count, B row, G row, R row)."""
    def attr(name, typ, value):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(value)) + value

    chlist = b"".join(n + b"\0" + struct.pack("<iB3xii", 2, 0, 1, 1) for n in (b"B", b"G", b"R")) + b"\0"
    head = struct.pack("<II", 20000630, 2)
    head += attr("channels", "chlist", chlist)
    head += attr("compression", "compression", b"\0")
    head += attr("dataWindow", "box2i", struct.pack("<4i", 0, 0, width - 1, height - 1))
    head += attr("displayWindow", "box2i", struct.pack("<4i", 0, 0, width - 1, height - 1))
    head += attr("lineOrder", "lineOrder", b"\0")
    head += attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    head += attr("screenWindowCenter", "v2f", struct.pack("<2f", 0.0, 0.0))
    head += attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    head += b"\0"
    line = 8 + 12 * width
    first = len(head) + 8 * height
    table = b"".join(struct.pack("<Q", first + y * line) for y in range(height))
    body = b""
    for y in range(height):
        body += struct.pack("<ii", y, 12 * width)
        for c in (2, 1, 0):
            body += np.asarray(rgb[y, :, c], "<f4").tobytes()
    return head + table + body


def test_reader_decodes_hand_made_file():
    rgb = np.array([[[1.0, 2.0, 3.0], [0.5, -0.25, 1e-30]]], np.float32)
    data = hand_made_exr(2, 1, rgb)
    assert data[:4] == bytes([0x76, 0x2F, 0x31, 0x01])  # the format's magic bytes
    e = F.read_exr(data)
    assert (e["width"], e["height"]) == (2, 1)
    assert [c[0] for c in e["channels"]] == ["B", "G", "R"] and all(c[1] == 2 for c in e["channels"])
    np.testing.assert_array_equal(e["pixels"]["R"], rgb[..., 0])
    np.testing.assert_array_equal(e["pixels"]["G"], rgb[..., 1])
    np.testing.assert_array_equal(e["pixels"]["B"], rgb[..., 2])


@pytest.mark.parametrize("shape", [(1, 1), (1, 2), (3, 5), (17, 31)])
def test_encoder_matches_hand_made_file(shape):
    h, w = shape
    rgb = np.random.default_rng(h * 100 + w).standard_normal((h, w, 3)).astype(np.float32)
    assert exr_bytes(rgb) == hand_made_exr(w, h, rgb)


def test_encoder_keeps_every_bit():
    rgb = np.random.default_rng(3).standard_normal((6, 4, 3)).astype(np.float32)
    rgb[0, 0] = [np.nan, np.inf, -np.inf]
    rgb[0, 1] = [-0.0, 1e-45, np.finfo(np.float32).max]
    e = F.read_exr(exr_bytes(rgb))
    back = np.stack([e["pixels"][c] for c in "RGB"], axis=-1)
    np.testing.assert_array_equal(back.view(np.uint32), rgb.view(np.uint32))


def test_write_to_file(tmp_path):
    rgb = np.random.default_rng(4).uniform(0, 2, (9, 13, 3)).astype(np.float32)
    p = tmp_path / "result.exr"
    write_exr(str(p), rgb)
    assert p.read_bytes() == exr_bytes(rgb)
    assert not (tmp_path / "result.exr.part").exists()


def test_bad_arguments():
    with pytest.raises(N.BonsaiError):
        write_exr("/nonexistent-dir/x.exr", np.zeros((2, 2, 3), np.float32))
    lib = N.lib()
    buf = np.zeros((1, 1, 3), np.float32)
    assert lib.bg_exr_write(b"/tmp/never.exr", 0, 1, buf.ctypes.data) == N.BG_ERR_INVALID
    assert lib.bg_exr_write(b"/tmp/never.exr", 1, 1, None) == N.BG_ERR_INVALID
    import ctypes as C
    size = C.c_uint64(3)
    out = (C.c_ubyte * 3)()
    assert lib.bg_exr_encode(1, 1, buf.ctypes.data, out, C.byref(size)) == N.BG_ERR_INVALID
    assert size.value == len(exr_bytes(buf))
