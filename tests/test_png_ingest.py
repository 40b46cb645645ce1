"""PNG images and numbered PNG sequences as CLI input (row f3: frame ingest
without OpenCV; the reference opens any cv::VideoCapture source,
src/videoio.cpp:13-52, and picks a channel as Mat2Array does,
include/videoio.h:71-98).

The CLI's decoder (speckle-stacker_amd/cli/frame_source.h, through the
smip_frame_dump tool) is checked against PIL on files PIL writes, and against
the pixel arrays themselves on files written here with every row filter and
Adam7 interlacing (PIL writes neither on demand).  No GPU is used."""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DUMP = os.path.join(REPO, "speckle-stacker_amd", "lib", "smip_frame_dump")

MASK_R, MASK_G, MASK_B, MASK_I = 4, 2, 1, 7


def _dump(path, mask=MASK_I):
    if not os.path.exists(DUMP):
        pytest.skip("smip_frame_dump is not built")
    out = path + ".raw" if "%" not in path else os.path.join(os.path.dirname(path), "seq.raw")
    r = subprocess.run([DUMP, path, str(mask), out], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr.strip())
    n, X, Y, dt = map(int, r.stdout.split())
    dtype = {1: np.uint8, 2: np.uint16}[dt]
    return np.fromfile(out, dtype=dtype).reshape(n, Y, X)


def _want(rgb_or_grey, mask):
    """Mat2Array's channel rule: round(log2(mask)) over B,G,R, grey images give grey."""
    a = np.asarray(rgb_or_grey)
    if a.ndim == 2:
        return a
    c = min(int(round(np.log2(mask))), 2)
    return a[..., 2 - c]  # stored R,G,B; index c counts B,G,R


def test_pil_written_modes(tmp_path):
    from PIL import Image
    rng = np.random.default_rng(3)
    H, W = 37, 53
    rgb = rng.integers(0, 256, size=(H, W, 3), dtype=np.uint8)
    grey = rng.integers(0, 256, size=(H, W), dtype=np.uint8)
    cases = {
        "L": (Image.fromarray(grey, "L"), grey),
        "RGB": (Image.fromarray(rgb, "RGB"), rgb),
        "RGBA": (Image.fromarray(np.dstack([rgb, grey]), "RGBA"), rgb),
        "LA": (Image.fromarray(np.dstack([grey, grey[::-1]]), "LA"), grey),
        "P": (Image.fromarray(rgb, "RGB").quantize(61), None),
        "1": (Image.fromarray(grey > 128).convert("1"), None),
    }
    for name, (im, arr) in cases.items():
        path = str(tmp_path / f"{name}.png")
        im.save(path)
        ref = np.asarray(Image.open(path).convert("RGB")) if arr is None else arr
        if name == "1":
            ref = np.asarray(Image.open(path).convert("L"))
        for mask in (MASK_R, MASK_G, MASK_B, MASK_I):
            got = _dump(path, mask)
            assert got.shape == (1, H, W) and got.dtype == np.uint8, name
            assert np.array_equal(got[0], _want(ref, mask)), (name, mask)


def test_16bit_grey_gives_uint16(tmp_path):
    from PIL import Image
    rng = np.random.default_rng(4)
    a = rng.integers(0, 65536, size=(20, 31), dtype=np.uint16)
    path = str(tmp_path / "g16.png")
    Image.fromarray(a.astype(np.int32), "I").convert("I;16").save(path)
    want = np.asarray(Image.open(path)).astype(np.uint16)
    got = _dump(path)
    assert got.dtype == np.uint16 and np.array_equal(got[0], want)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def _filter_rows(raw_rows, bpp, rng):
    """Each scanline with a random filter type 0-4."""
    out = b""
    prev = bytes(len(raw_rows[0]))
    for row in raw_rows:
        ft = int(rng.integers(0, 5))
        f = bytearray(len(row))
        for i, x in enumerate(row):
            a = row[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][ft]
            f[i] = (x - pred) & 0xFF
        out += bytes([ft]) + bytes(f)
        prev = row
    return out


def _pack(samples, depth):
    """Big-endian packing of one row of samples at `depth` bits."""
    if depth == 8:
        return bytes(int(v) for v in samples)
    if depth == 16:
        return b"".join(struct.pack(">H", int(v)) for v in samples)
    bits = "".join(format(int(v), f"0{depth}b") for v in samples)
    bits += "0" * (-len(bits) % 8)
    return bytes(int(bits[k:k + 8], 2) for k in range(0, len(bits), 8))


def _chunk(t, body):
    return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body) & 0xFFFFFFFF)


def write_png(path, pix, depth, ctype, interlace, rng, plte=None):
    """pix: (H, W, spp) integer samples; every row filter and optional Adam7."""
    H, W, spp = pix.shape
    bpp = max(1, spp * depth // 8)
    passes = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)] \
        if interlace else [(0, 0, 1, 1)]
    data = b""
    for x0, y0, dx, dy in passes:
        sub = pix[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        rows = [_pack(sub[y].reshape(-1), depth) for y in range(sub.shape[0])]
        data += _filter_rows(rows, bpp, rng)
    png = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, depth, ctype, 0, 0, interlace))
    if plte is not None:
        png += _chunk(b"PLTE", bytes(int(v) for v in plte.reshape(-1)))
    png += _chunk(b"IDAT", zlib.compress(data, 6)) + _chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(png)


@pytest.mark.parametrize("interlace", [0, 1])
@pytest.mark.parametrize("ctype,depth", [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 2), (3, 8),
                                         (4, 8), (4, 16), (6, 8), (6, 16)])
def test_every_format_filter_and_interlace(tmp_path, ctype, depth, interlace):
    rng = np.random.default_rng(ctype * 100 + depth * 2 + interlace)
    H, W = 23, 19  # odd sizes leave partial Adam7 blocks
    spp = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    top = (1 << depth) - 1
    pix = rng.integers(0, top + 1, size=(H, W, spp), dtype=np.int64)
    plte = None
    if ctype == 3:
        plte = rng.integers(0, 256, size=(top + 1, 3), dtype=np.int64)
    path = str(tmp_path / "t.png")
    write_png(path, pix, depth, ctype, interlace, rng, plte)
    for mask in (MASK_R, MASK_G, MASK_B):
        got = _dump(path, mask)[0]
        if ctype == 3:
            want = _want(plte[pix[..., 0]], mask)
        elif ctype in (0, 4):  # grey: scaled to 8 bits below depth 8
            g = pix[..., 0]
            want = g * 255 // top if depth < 8 else g
        else:
            want = _want(pix[..., :3], mask)
        assert got.dtype == (np.uint16 if depth == 16 else np.uint8)
        assert np.array_equal(got.astype(np.int64), want), (ctype, depth, interlace, mask)
    if interlace == 0 and depth == 8 and ctype in (0, 2, 6):  # PIL reads these too
        from PIL import Image
        assert np.array_equal(_dump(path, MASK_R)[0], _want(np.asarray(Image.open(path).convert("RGB")), MASK_R))


def test_numbered_sequences(tmp_path):
    from PIL import Image
    rng = np.random.default_rng(8)
    frames = rng.integers(0, 256, size=(5, 12, 9), dtype=np.uint8)
    for k, f in enumerate(frames):
        Image.fromarray(f, "L").save(str(tmp_path / f"frame_{k + 3:04d}.png"))
    Image.fromarray(frames[0], "L").save(str(tmp_path / "frame_0009.png"))  # after a gap: not part of it
    # a printf pattern starts at the first existing index
    assert np.array_equal(_dump(str(tmp_path / "frame_%04d.png")), frames)
    # a numbered name continues from its number with the same padding
    assert np.array_equal(_dump(str(tmp_path / "frame_0005.png")), frames[2:])
    # a name without digits is one image
    Image.fromarray(frames[1], "L").save(str(tmp_path / "single.png"))
    assert np.array_equal(_dump(str(tmp_path / "single.png")), frames[1:2])
    # frames of another size in a sequence are refused
    Image.fromarray(np.zeros((4, 4), np.uint8), "L").save(str(tmp_path / "frame_0008.png"))
    with pytest.raises(RuntimeError, match="differs"):
        _dump(str(tmp_path / "frame_%04d.png"))


def test_corrupt_files_are_refused(tmp_path):
    from PIL import Image
    path = str(tmp_path / "c.png")
    Image.fromarray(np.zeros((8, 8), np.uint8), "L").save(path)
    data = bytearray(open(path, "rb").read())
    data[40] ^= 0xFF  # inside a chunk: its CRC no longer matches
    bad = str(tmp_path / "bad.png")
    open(bad, "wb").write(bytes(data))
    with pytest.raises(RuntimeError, match="CRC|PNG"):
        _dump(bad)
    open(bad, "wb").write(b"not a png at all")
    with pytest.raises(RuntimeError, match="signature"):
        _dump(bad)
