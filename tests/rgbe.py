"""Radiance RGBE (.hdr) writer for test inputs and a reader of such files, from the format's published
description (Ward, "Real pixels", Graphics Gems II): per pixel an 8-bit
mantissa per channel and a shared exponent e with value (m + 1/2) 2^(e - 136);
scanlines flat, or run-length encoded per channel (new-style: 2 2 hi lo header,
runs 128+n followed by a byte, literal spans n followed by n bytes)."""
import numpy as np


def to_rgbe(rgb):
    """(h, w, 3) float -> (h, w, 4) uint8."""
    rgb = np.asarray(rgb, np.float64)
    mx = rgb.max(axis=-1)
    out = np.zeros(rgb.shape[:-1] + (4,), np.uint8)
    ok = mx > 1e-32
    m, e = np.frexp(mx[ok])  # mx = m 2^e, m in [0.5, 1)
    scale = m * 256.0 / mx[ok]
    out[ok, :3] = np.clip(np.floor(rgb[ok] * scale[:, None]), 0, 255).astype(np.uint8)
    out[ok, 3] = (e + 128).astype(np.uint8)
    return out


def decode(rgbe):
    """(h, w, 4) uint8 -> (h, w, 3) float32 by the Radiance convention."""
    rgbe = np.asarray(rgbe)
    e = rgbe[..., 3].astype(np.int64)
    s = np.where(e == 0, 0.0, np.ldexp(1.0, e - 136))
    return np.where(e[..., None] == 0, 0.0, (rgbe[..., :3] + 0.5) * s[..., None]).astype(np.float32)


def _rle_channel(vals):
    out, i, n = bytearray(), 0, len(vals)
    while i < n:
        j = i
        while j < n and j - i < 127 and vals[j] == vals[i]:
            j += 1
        if j - i >= 3:
            out += bytes([128 + (j - i), vals[i]])
            i = j
            continue
        k = i
        while k < n and k - i < 128 and not (k + 2 < n and vals[k] == vals[k + 1] == vals[k + 2]):
            k += 1
        out += bytes([k - i]) + bytes(vals[i:k])
        i = k
    return bytes(out)


def write(path, rgbe, rle=True, exposure=None):
    """Write (h, w, 4) uint8 RGBE, first row at the top (-Y +X)."""
    h, w = rgbe.shape[:2]
    hdr = b"#?RADIANCE\n# test image\nFORMAT=32-bit_rle_rgbe\n"
    if exposure is not None:
        hdr += b"EXPOSURE=%g\n" % exposure
    hdr += b"\n-Y %d +X %d\n" % (h, w)
    body = bytearray()
    for y in range(h):
        row = rgbe[y]
        if rle and 8 <= w <= 0x7FFF:
            body += bytes([2, 2, w >> 8, w & 255])
            for c in range(4):
                body += _rle_channel(list(row[:, c]))
        else:
            body += row.tobytes()
    with open(path, "wb") as f:
        f.write(hdr + bytes(body))


def read(path):
    """(h, w, 4) uint8 RGBE of a '-Y h +X w' file, first row at the top, and its
    header lines; flat or new-style run-length-encoded scanlines."""
    data = open(path, "rb").read()
    pos, header = 0, []
    while True:
        end = data.index(b"\n", pos)
        line = data[pos:end]
        pos = end + 1
        if line.startswith(b"-Y") or line.startswith(b"+Y"):
            break
        header.append(line.decode("latin-1"))
    tok = line.split()
    assert tok[0] == b"-Y" and tok[2] == b"+X", line
    h, w = int(tok[1]), int(tok[3])
    out = np.zeros((h, w, 4), np.uint8)
    buf = np.frombuffer(data, np.uint8)
    for y in range(h):
        if 8 <= w <= 0x7FFF and buf[pos] == 2 and buf[pos + 1] == 2 and not (buf[pos + 2] & 0x80):
            assert (int(buf[pos + 2]) << 8) | int(buf[pos + 3]) == w
            pos += 4
            for c in range(4):
                x = 0
                while x < w:
                    n = int(buf[pos])
                    pos += 1
                    if n > 128:
                        out[y, x:x + n - 128, c] = buf[pos]
                        x += n - 128
                        pos += 1
                    else:
                        out[y, x:x + n, c] = buf[pos:pos + n]
                        x += n
                        pos += n
                assert x == w
        else:
            out[y] = buf[pos:pos + 4 * w].reshape(w, 4)
            pos += 4 * w
    return out, header
