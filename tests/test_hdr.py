"""Radiance RGBE (.hdr) decoding (hvx_hdr_decode): known texel values, the
float -> RGBE -> float round trip, flat and run-length encoded scanlines
decoding alike, the flip-on-load row order, and malformed files. The files
are written here by a small encoder of the published format (the reference's
.hdr assets are not present)."""
import numpy as np
import pytest

from hellovulkan_amd import abi


def rgbe(img):
    """Ward's RGBE encoding of an (h, w, 3) float image."""
    v = img.max(axis=2)
    m, e = np.frexp(v)
    scale = np.where(v > 1e-32, m * 256.0 / np.where(v > 1e-32, v, 1.0), 0.0)
    out = np.zeros(img.shape[:2] + (4,), np.uint8)
    out[..., :3] = np.clip(np.floor(img * scale[..., None]), 0, 255).astype(np.uint8)
    out[..., 3] = np.where(v > 1e-32, e + 128, 0).astype(np.uint8)
    return out


def header(w, h, fmt="32-bit_rle_rgbe", magic="#?RADIANCE"):
    return f"{magic}\n# made by the test\nFORMAT={fmt}\n\n-Y {h} +X {w}\n".encode()


def flat_file(q):
    h, w, _ = q.shape
    return header(w, h) + q.tobytes()


def rle_scanline(row):
    w = row.shape[0]
    out = bytearray([2, 2, w >> 8, w & 255])
    for ch in range(4):
        plane = row[:, ch]
        x = 0
        while x < w:
            run = 1
            while x + run < w and run < 127 and plane[x + run] == plane[x]:
                run += 1
            if run >= 3:
                out += bytes([128 + run, plane[x]])
                x += run
            else:
                lit = []
                while x < w and len(lit) < 128:
                    if x + 2 < w and plane[x] == plane[x + 1] == plane[x + 2]:
                        break
                    lit.append(plane[x])
                    x += 1
                out += bytes([len(lit)]) + bytes(lit)
    return bytes(out)


def rle_file(q):
    h, w, _ = q.shape
    return header(w, h) + b"".join(rle_scanline(q[y]) for y in range(h))


def expected(q):
    f = np.where(q[..., 3:] > 0, np.ldexp(1.0, q[..., 3:].astype(np.int64) - 136), 0.0)
    rgb = (q[..., :3] * f).astype(np.float32)
    return np.concatenate([rgb, np.ones(q.shape[:2] + (1,), np.float32)], axis=2)


def test_known_texels():
    q = np.array([[[128, 64, 32, 129], [0, 0, 0, 0], [255, 255, 255, 128 + 20], [1, 2, 3, 1]]], np.uint8)
    img = abi.hdr_decode(flat_file(q), flip_vertically=False)
    assert img.shape == (1, 4, 4)
    assert np.array_equal(img[0, 0], [1.0, 0.5, 0.25, 1.0])           # 128 * 2^-7
    assert np.array_equal(img[0, 1], [0.0, 0.0, 0.0, 1.0])            # e == 0
    assert np.array_equal(img[0, 2], [255 * 2.0 ** 12] * 3 + [1.0])   # 255 * 2^(148-136)
    assert img[0, 3, 0] == np.float32(2.0 ** -135)


def test_round_trip_relative_error():
    rng = np.random.default_rng(3)
    img = np.exp(rng.normal(0, 4, (37, 61, 3))).astype(np.float32)  # HDR spread over ~e^+-12
    out = abi.hdr_decode(rle_file(rgbe(img)), flip_vertically=False)
    rel = np.abs(out[..., :3] - img) / img.max(axis=2, keepdims=True)
    assert rel.max() < 2.0 ** -7  # mantissas are 8 bits of the largest channel


@pytest.mark.parametrize("w,h", [(64, 9), (8, 3), (300, 5), (7, 4)])
def test_rle_and_flat_agree(w, h):
    rng = np.random.default_rng(w)
    img = np.exp(rng.normal(0, 2, (h, w, 3))).astype(np.float32)
    img[:, : w // 3] = 1.5  # runs
    q = rgbe(img)
    a = abi.hdr_decode(flat_file(q), flip_vertically=False)
    b = abi.hdr_decode(rle_file(q) if w >= 8 else flat_file(q), flip_vertically=False)
    assert np.array_equal(a, b)
    assert np.array_equal(a, expected(q))


def test_flip_on_load():
    q = rgbe(np.arange(5 * 9 * 3, dtype=np.float32).reshape(5, 9, 3) + 1.0)
    up = abi.hdr_decode(rle_file(q), flip_vertically=True)
    down = abi.hdr_decode(rle_file(q), flip_vertically=False)
    assert np.array_equal(up, down[::-1])


def test_rejects_malformed():
    q = rgbe(np.ones((4, 16, 3), np.float32))
    for data, status in [(b"P6\n1 1\n255\n", abi.HVX_ERR_INVALID_ARG),
                         (header(16, 4, fmt="32-bit_rle_xyze") + q.tobytes(), abi.HVX_ERR_UNSUPPORTED),
                         (b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n+Y 4 +X 16\n" + q.tobytes(), abi.HVX_ERR_UNSUPPORTED),
                         (flat_file(q)[:-5], abi.HVX_ERR_INVALID_ARG),
                         (rle_file(q)[:-3], abi.HVX_ERR_INVALID_ARG)]:
        with pytest.raises(abi.HvxError) as e:
            abi.hdr_decode(data)
        assert e.value.status == status
