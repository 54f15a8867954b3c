"""Radiance RGBE (.hdr) reading and writing.

The reference writes its renders with stbi_write_hdr (core/ImageUtils.cpp:51-75)
and its tests read them back with stbi_loadf (src/bin/unittests/RenderTests.cpp:7-33).
This module implements the public RGBE format (flat and new-style run-length
scanlines): encoding takes frexp of the largest component and truncates each
component times 256 / 2^e; decoding multiplies the byte by 2^(e - 136).
"""
import numpy as np


def encode_rgbe(rgb):
    """(..., 3+) float -> (..., 4) uint8 RGBE."""
    rgb = np.asarray(rgb, np.float32)[..., :3]
    m = rgb.max(axis=-1)
    out = np.zeros(rgb.shape[:-1] + (4,), np.uint8)
    nz = m >= 1e-32
    mant, ex = np.frexp(m[nz])
    sc = (mant * np.float32(256.0) / m[nz]).astype(np.float32)
    for c in range(3):
        out[..., c][nz] = np.clip(rgb[..., c][nz] * sc, 0, 255).astype(np.uint8)
    out[..., 3][nz] = (ex + 128).astype(np.uint8)
    return out


def decode_rgbe(rgbe):
    """(..., 4) uint8 RGBE -> (..., 3) float32."""
    rgbe = np.asarray(rgbe).astype(np.int32)
    e = rgbe[..., 3]
    scale = np.where(e > 0, np.ldexp(1.0, e - (128 + 8)), 0.0).astype(np.float32)
    return (rgbe[..., :3].astype(np.float32) * scale[..., None]).astype(np.float32)


def read_hdr_rgbe(path):
    """Return the (H, W, 4) uint8 RGBE bytes of a Radiance file."""
    with open(path, "rb") as f:
        data = f.read()
    header_end = data.find(b"\n\n")
    if header_end < 0:
        raise ValueError("not a Radiance HDR file")
    pos = header_end + 2
    eol = data.find(b"\n", pos)
    dims = data[pos:eol].split()
    pos = eol + 1
    if dims[0] != b"-Y" or dims[2] != b"+X":
        raise ValueError("unsupported orientation")
    H, W = int(dims[1]), int(dims[3])
    img = np.zeros((H, W, 4), np.uint8)
    buf = np.frombuffer(data, np.uint8)
    for y in range(H):
        if 8 <= W < 32768 and buf[pos] == 2 and buf[pos + 1] == 2 and (buf[pos + 2] & 0x80) == 0:
            if (int(buf[pos + 2]) << 8 | int(buf[pos + 3])) != W:
                raise ValueError("bad scanline width")
            pos += 4
            for c in range(4):
                x = 0
                while x < W:
                    n = int(buf[pos])
                    pos += 1
                    if n > 128:
                        n -= 128
                        img[y, x:x + n, c] = buf[pos]
                        pos += 1
                    else:
                        img[y, x:x + n, c] = buf[pos:pos + n]
                        pos += n
                    x += n
        else:
            img[y] = buf[pos:pos + 4 * W].reshape(W, 4)
            pos += 4 * W
    return img


def read_hdr(path):
    """Return an (H, W, 3) float32 image."""
    return decode_rgbe(read_hdr_rgbe(path))


def write_hdr(path, rgb):
    """Write an (H, W, 3+) float image as flat (uncompressed) RGBE."""
    q = encode_rgbe(rgb)
    H, W = q.shape[:2]
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n")
        f.write(f"-Y {H} +X {W}\n".encode())
        f.write(q.tobytes())
