"""Image files for the renderer's inputs and outputs, without third-party codecs:
PNG (8-bit RGB, written as captureImage writes screenshots, src/Application.cpp:370-378),
PFM (float RGB) and Radiance HDR (RGBE, read for environment maps, which the reference
loads with stbi_loadf, src/util/Image.cpp:15-19)."""
import re
import struct
import zlib

import numpy as np


def write_png(path, rgb8):
    """8-bit RGB PNG of an (h, w, 3) uint8 array (row 0 at the top)."""
    a = np.ascontiguousarray(rgb8, np.uint8)
    h, w, c = a.shape
    assert c == 3
    raw = b"".join(b"\x00" + a[y].tobytes() for y in range(h))

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)
    png = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)) +
           chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))
    with open(path, "wb") as f:
        f.write(png)


def read_png_rgb8(path):
    """Reads what write_png writes (8-bit RGB, filter type 0 rows)."""
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        tag = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        if tag == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert depth == 8 and ctype == 2
        elif tag == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 3).copy()


def write_pfm(path, rgb):
    """Little-endian PFM of an (h, w, 3) float array (row 0 at the top)."""
    a = np.ascontiguousarray(rgb, np.float32)
    h, w, _ = a.shape
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1.0\n".encode())
        f.write(a[::-1].astype("<f4").tobytes())


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        a = np.frombuffer(f.read(), "<f4" if scale < 0 else ">f4").reshape(h, w, 3)
    return a[::-1].astype(np.float32)


def _rgbe_to_float(rgbe):
    e = rgbe[..., 3].astype(np.int32)
    f = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0)  # 2^(e - 128 - 8)
    return (rgbe[..., :3].astype(np.float64) * f[..., None]).astype(np.float32)


def read_hdr(path):
    """Radiance RGBE (.hdr): (h, w, 3) float32, row 0 at the top; flat and new-style
    run-length encoded scanlines, -Y H +X W orientation."""
    data = open(path, "rb").read()
    pos = 0
    while True:
        end = data.index(b"\n", pos)
        line = data[pos:end].strip()
        pos = end + 1
        if not line:
            break
        if line.startswith(b"FORMAT=") and line != b"FORMAT=32-bit_rle_rgbe":
            raise ValueError("unsupported HDR format " + line.decode())
    end = data.index(b"\n", pos)
    m = re.match(rb"-Y (\d+) \+X (\d+)", data[pos:end].strip())
    if not m:
        raise ValueError("unsupported HDR orientation")
    h, w = int(m.group(1)), int(m.group(2))
    pos = end + 1
    out = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        if 8 <= w < 32768 and data[pos] == 2 and data[pos + 1] == 2 and (data[pos + 2] << 8 | data[pos + 3]) == w:
            pos += 4
            for c in range(4):
                x = 0
                while x < w:
                    n = data[pos]
                    pos += 1
                    if n > 128:
                        out[y, x:x + n - 128, c] = data[pos]
                        pos += 1
                        x += n - 128
                    else:
                        out[y, x:x + n, c] = np.frombuffer(data[pos:pos + n], np.uint8)
                        pos += n
                        x += n
        else:
            out[y] = np.frombuffer(data[pos:pos + 4 * w], np.uint8).reshape(w, 4)
            pos += 4 * w
    return _rgbe_to_float(out)


def write_hdr(path, rgb):
    """Flat (not run-length encoded) Radiance RGBE file of an (h, w, 3) float array."""
    a = np.asarray(rgb, np.float64)
    h, w, _ = a.shape
    mx = a.max(axis=2)
    m, e = np.frexp(mx)
    rgbe = np.zeros((h, w, 4), np.uint8)
    ok = mx > 1e-32
    scale = np.where(ok, m * 256.0 / np.where(ok, mx, 1.0), 0.0)
    rgbe[..., :3] = np.clip(np.floor(a * scale[..., None]), 0, 255).astype(np.uint8)
    rgbe[..., 3] = np.where(ok, e + 128, 0).astype(np.uint8)
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode())
        f.write(rgbe.tobytes())
