"""AddressSanitizer and UndefinedBehaviorSanitizer builds of the CPU code: the
oracle (all of its functions, through oracle/san_check.c) and the CLI's frame
decoders (speckle-stacker_amd/cli/frame_source.h, through tests/cpp/frame_dump.cpp;
the reference reads frames with cv::VideoCapture, src/videoio.cpp:13-52).

The decoders get the reference's GIF, PNG files written by PIL, and seeded
truncated, byte-mutated and hand-made malformed files.  A malformed file must end
in an ordinary error (exit status 1 and a message), never a sanitizer report.
The builds use g++/gcc only, so no GPU code is involved.  No GPU is used."""
import os
import shutil
import struct
import subprocess
import zlib

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(REPO, "oracle")
PKG = os.path.join(REPO, "speckle-stacker_amd")
GIF = os.path.join(REPO, "tests", "golden", "hu940ani.gif")

SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
ENV = dict(os.environ,
           ASAN_OPTIONS="exitcode=86:allocator_may_return_null=1:detect_leaks=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:exitcode=87")


def _have(tool):
    if shutil.which(tool) is None:
        pytest.skip(f"{tool} is not installed")


def _run(cmd, timeout=120):
    return subprocess.run(cmd, capture_output=True, text=True, env=ENV, timeout=timeout)


def _clean(r):
    """No sanitizer report and an ordinary exit status."""
    report = "AddressSanitizer" in r.stderr or "runtime error:" in r.stderr
    return not report and r.returncode in (0, 1)


def test_oracle_under_sanitizers(tmp_path):
    _have("gcc")
    exe = str(tmp_path / "san_check")
    subprocess.run(["gcc", "-std=c11", *SAN, "-Wall", "-Wextra", f"-I{ORACLE}", "-o", exe,
                    os.path.join(ORACLE, "san_check.c"), os.path.join(ORACLE, "smip_oracle.c"), "-lm"],
                   check=True, capture_output=True, text=True)
    r = _run([exe])
    assert r.returncode == 0, r.stderr[-3000:]
    assert "san_check: ok" in r.stdout


@pytest.fixture(scope="module")
def dump_san(tmp_path_factory):
    _have("g++")
    exe = str(tmp_path_factory.mktemp("san") / "frame_dump_san")
    r = subprocess.run(["g++", "-std=c++20", *SAN, f"-I{os.path.join(REPO, 'include')}", f"-I{PKG}", "-o", exe,
                        os.path.join(REPO, "tests", "cpp", "frame_dump.cpp"), "-lz"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("sanitizer build of frame_dump failed:\n" + r.stderr[-3000:])
    return exe


def _decode(exe, path, tmp_path, mask=7):
    r = _run([exe, path, str(mask), str(tmp_path / "out.raw")])
    assert _clean(r), f"{os.path.basename(path)}: status {r.returncode}\n{r.stderr[-3000:]}"
    return r


def _chunk(tag, data):
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def _png(width, height, depth, colour, raw, extra=b""):
    ihdr = struct.pack(">IIBBBBB", width, height, depth, colour, 0, 0, 0)
    return b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr) + extra + _chunk(b"IDAT", raw) + _chunk(b"IEND", b"")


def test_valid_inputs_decode_cleanly(dump_san, tmp_path):
    from PIL import Image
    r = _decode(dump_san, GIF, tmp_path, mask=4)
    assert r.returncode == 0 and int(r.stdout.split()[0]) > 0
    rng = np.random.default_rng(3)
    rgb = rng.integers(0, 256, (37, 53, 3), dtype=np.uint8)
    Image.fromarray(rgb).save(tmp_path / "rgb.png")
    grey16 = rng.integers(0, 65536, (29, 31), dtype=np.uint16)
    Image.fromarray(grey16).save(tmp_path / "g16.png")
    for name in ("rgb.png", "g16.png"):
        assert _decode(dump_san, str(tmp_path / name), tmp_path).returncode == 0


def test_truncated_and_mutated_gifs_fail_cleanly(dump_san, tmp_path):
    data = open(GIF, "rb").read()
    rng = np.random.default_rng(11)
    cases = [data[:n] for n in (0, 5, 6, 12, 13, 20, 800, 801, 4096, len(data) // 2, len(data) - 1)]
    for _ in range(40):
        b = bytearray(data)
        for pos in rng.integers(0, len(b), int(rng.integers(1, 6))):
            b[int(pos)] = int(rng.integers(0, 256))
        cases.append(bytes(b))
    # header area mutations: screen size, colour-table flags, first image descriptor
    for pos in range(6, 40):
        b = bytearray(data)
        b[pos] ^= 0xFF
        cases.append(bytes(b))
    for k, blob in enumerate(cases):
        p = tmp_path / f"m{k}.gif"
        p.write_bytes(blob)
        _decode(dump_san, str(p), tmp_path, mask=4)


def test_malformed_pngs_fail_cleanly(dump_san, tmp_path):
    from PIL import Image
    rng = np.random.default_rng(5)
    Image.fromarray(rng.integers(0, 256, (24, 40, 3), dtype=np.uint8)).save(tmp_path / "ok.png")
    good = (tmp_path / "ok.png").read_bytes()
    rows8 = b"".join(b"\x00" + bytes(range(16)) for _ in range(16))          # 16x16 grey, filter 0
    bad_filter = b"".join(b"\x05" + bytes(16) for _ in range(16))            # filter type 5
    cases = {
        "trunc": [good[:n] for n in (0, 7, 8, 20, 33, 40, len(good) // 2, len(good) - 4)],
        "crafted": [
            _png(16, 16, 8, 0, zlib.compress(rows8[: len(rows8) // 2])),       # IDAT short of the image
            _png(16, 16, 8, 0, zlib.compress(rows8 * 3)),                       # IDAT longer than the image
            _png(16, 16, 8, 0, zlib.compress(bad_filter)),                      # unknown row filter
            _png(16, 16, 8, 0, bytes(rng.integers(0, 256, 300, dtype=np.uint8))),  # not a zlib stream
            _png(16, 16, 8, 0, zlib.compress(rows8)[:-6]),                      # cut zlib stream
            _png(0, 16, 8, 0, zlib.compress(b"")),                              # zero width
            _png(1 << 20, 1 << 20, 8, 0, zlib.compress(rows8)),                 # huge extents, little data
            _png(16, 16, 7, 0, zlib.compress(rows8)),                           # bad bit depth
            _png(16, 16, 8, 5, zlib.compress(rows8)),                           # bad colour type
            _png(16, 16, 8, 3, zlib.compress(rows8)),                           # palette without PLTE
            _png(16, 16, 8, 3, zlib.compress(rows8), _chunk(b"PLTE", bytes(12))),  # indices past a 4-entry PLTE
            _png(16, 16, 16, 2, zlib.compress(rows8)),                          # 16-bit RGB, data for 8-bit grey
        ],
    }
    for _ in range(30):
        b = bytearray(good)
        for pos in rng.integers(8, len(b), int(rng.integers(1, 4))):
            b[int(pos)] = int(rng.integers(0, 256))
        cases["trunc"].append(bytes(b))
    for kind, blobs in cases.items():
        for k, blob in enumerate(blobs):
            p = tmp_path / f"{kind}{k}.png"
            p.write_bytes(blob)
            _decode(dump_san, str(p), tmp_path)
