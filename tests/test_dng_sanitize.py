"""The DNG reader (climage_amd/csrc/dng.cpp) parses untrusted files.  These
tests build it on the host with AddressSanitizer and UndefinedBehaviorSanitizer
(g++, no GPU code involved) and feed it the golden DNGs, hand-made malformed
files and seeded random mutations of the golden files; any out-of-bounds
access, overflow or undefined shift aborts the harness with a report.  The
malformed cases are also run through the library itself, where they must
come back as CLIM_EINVAL."""
import glob
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

import climage_amd as ca
from oracle import dng_oracle as D

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = sorted(glob.glob(os.path.join(HERE, "golden", "dng", "*.dng")))


def _walk_ifd0(buf):
    """(endian, [(entry offset, tag, type, count)]) of IFD0."""
    e = "<" if buf[:2] == b"II" else ">"
    off = struct.unpack(e + "I", buf[4:8])[0]
    n = struct.unpack(e + "H", buf[off:off + 2])[0]
    ents = []
    for i in range(n):
        p = off + 2 + 12 * i
        tag, typ, cnt = struct.unpack(e + "HHI", buf[p:p + 8])
        ents.append((p, tag, typ, cnt))
    return e, ents


def bad_huffman_table():
    """lj92_strip_nc2.dng with every Huffman code moved to length 1: more codes
    of one length than that length holds (the overflow of a lookup-table
    builder that checks only after writing)."""
    buf = bytearray(open(os.path.join(HERE, "golden", "dng", "lj92_strip_nc2.dng"), "rb").read())
    p = buf.index(b"\xff\xc4")
    counts = buf[p + 5:p + 21]
    tot = sum(counts)
    assert tot > 2
    buf[p + 5:p + 21] = bytes([tot] + [0] * 15)
    return bytes(buf)


def huge_tiles():
    """A tiled DNG whose TileWidth and TileLength claim 2^30 samples each."""
    raw = np.arange(64 * 48, dtype=np.uint16).reshape(48, 64)
    buf = bytearray(D.write_dng(raw, 2, bits=16, tile=(16, 16)))
    e, ents = _walk_ifd0(buf)
    n = 0
    for p, tag, typ, cnt in ents:
        if tag in (322, 323):
            buf[p + 2:p + 12] = struct.pack(e + "HII", 4, 1, 1 << 30)
            n += 1
    assert n == 2
    return bytes(buf)


def point_transform_too_large():
    """A lossless JPEG whose point transform equals the precision (the first
    predictor 2^(P - Pt - 1) would be a negative shift)."""
    buf = bytearray(open(os.path.join(HERE, "golden", "dng", "lj92_p16_nc4_pt2.dng"), "rb").read())
    p = buf.index(b"\xff\xda")
    ln = struct.unpack(">H", buf[p + 2:p + 4])[0]
    ns = buf[p + 4]
    al_at = p + 4 + 1 + 2 * ns + 2
    assert al_at < p + 2 + ln
    buf[al_at] = (buf[al_at] & 0xF0) | 15
    sof = buf.index(b"\xff\xc3")
    buf[sof + 4] = 12  # P = 12 <= Pt = 15
    return bytes(buf)


MALFORMED = {"bad_huffman_table": bad_huffman_table, "huge_tiles": huge_tiles,
             "point_transform": point_transform_too_large}


@pytest.mark.parametrize("name", sorted(MALFORMED))
def test_malformed_rejected_by_library(name):
    with pytest.raises(ca.ClimError) as ei:
        ca.dng_decode(MALFORMED[name]())
    assert ei.value.status == ca.EINVAL


def _mutants(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        src = bytearray(open(GOLD[i % len(GOLD)], "rb").read())
        k = int(rng.integers(1, 9))
        for _ in range(k):
            pos = int(rng.integers(0, len(src)))
            mode = int(rng.integers(0, 3))
            if mode == 0:
                src[pos] = int(rng.integers(0, 256))
            elif mode == 1:
                src[pos] ^= 1 << int(rng.integers(0, 8))
            else:
                src[pos] = 0xFF
        out.append(bytes(src))
    return out


def test_sanitizer_build_over_golden_malformed_and_mutated(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    exe = str(tmp_path / "dng_san")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "sanitize", "dng_harness.cpp"),
           os.path.join(ROOT, "climage_amd", "csrc", "dng.cpp"), os.path.join(ROOT, "climage_amd", "csrc", "colour.cpp"),
           "-o", exe, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    files = list(GOLD)
    for name, make in MALFORMED.items():
        p = tmp_path / f"{name}.dng"
        p.write_bytes(make())
        files.append(str(p))
    for i, m in enumerate(_mutants(7, 240)):
        p = tmp_path / f"mut{i:03d}.dng"
        p.write_bytes(m)
        files.append(str(p))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, *files], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, \
        r.stderr[-4000:]
    lines = dict(l.split(": ", 1) for l in r.stdout.splitlines())
    for g in GOLD:
        assert lines[g].startswith("0"), (g, lines[g])
    for name in MALFORMED:
        assert lines[str(tmp_path / f"{name}.dng")].startswith("1"), (name, lines[str(tmp_path / f"{name}.dng")])
