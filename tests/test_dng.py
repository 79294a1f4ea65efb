"""DNG ingest (read_dng_file, src/gls_image_tiff.cpp:255-458, and
unpackDNGMetadata, cpp/demosaic_utils.cpp:280-375) in the library's host code,
against the CPU oracle (oracle/dng_oracle.py) and what the specifications fix.

The reference tree holds no DNG file and its decoder depends on libtiff,
which is absent; the lossless JPEG codec is pinned by a stream assembled by
hand from ITU-T T.81 (H.1.2.2 / F.1.2.1) and by libjpeg-turbo (through PIL),
an independent decoder, on the oracle encoder's 8-bit streams; the
container by fixtures written with the oracle and decoded by both readers.
No GPU is needed."""
import glob
import io
import os

import numpy as np
import pytest

import climage_amd as ca
from oracle import dng_oracle as D

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dng")
FIXTURES = sorted(glob.glob(os.path.join(GOLD, "*.dng")))


def hand_stream():
    """P = 8, 3 x 2 samples, one component, predictor 1, Huffman codes
    cat 0 '00', 1 '01', 2 '100', 3 '101', 8 '110'.  Samples 130 126 126 /
    131 0 255: differences +2 (from 2^(P-1)), -4, 0 / +1 (from above), -131,
    +255; a negative difference d sends the low bits of d - 1 (d + 2^s - 1)."""
    bits = ("100" + "10") + ("101" + "011") + "00" + ("01" + "1") + ("110" + "01111100") + ("110" + "11111111")
    bits += "1" * (-len(bits) % 8)
    data = bytearray()
    for i in range(0, len(bits), 8):
        b = int(bits[i:i + 8], 2)
        data.append(b)
        if b == 0xFF:
            data.append(0)
    s = bytearray(b"\xff\xd8")
    s += b"\xff\xc3\x00\x0b\x08\x00\x02\x00\x03\x01\x01\x11\x00"
    counts = [0, 2, 3] + [0] * 13
    s += b"\xff\xc4\x00\x18\x00" + bytes(counts) + bytes([0, 1, 2, 3, 8])
    s += b"\xff\xda\x00\x08\x01\x01\x00\x01\x00\x00"
    s += data + b"\xff\xd9"
    return bytes(s), np.array([[130, 126, 126], [131, 0, 255]], np.uint16)


def test_hand_assembled_lossless_jpeg():
    stream, expected = hand_stream()
    np.testing.assert_array_equal(D.lj92_decode(stream), expected)
    from PIL import Image
    np.testing.assert_array_equal(np.array(Image.open(io.BytesIO(stream))), expected)
    dng = D.write_dng(expected, 2, bits=8, compression=7, lj92_nc=1, lj92_stream=stream)
    raw, info = ca.dng_decode(dng)
    np.testing.assert_array_equal(raw, expected)
    assert info.compression == 7 and info.width == 3 and info.height == 2


@pytest.mark.parametrize("pred", [1, 2, 3, 4, 5, 6, 7])
def test_oracle_encoder_against_libjpeg_turbo(pred):
    """8-bit, one component: libjpeg-turbo's lossless decoder (through PIL)
    reads the oracle encoder's streams, with and without restart intervals
    (two rows each)."""
    from PIL import Image
    rng = np.random.default_rng(pred)
    img = np.clip(np.cumsum(rng.integers(-9, 10, (20, 33)), axis=1) + 128, 0, 255).astype(np.int64)
    for rst in (0, 66):
        s = D.lj92_encode(img, 8, nc=1, predictor=pred, restart=rst)
        np.testing.assert_array_equal(np.array(Image.open(io.BytesIO(s))), img)
        np.testing.assert_array_equal(D.lj92_decode(s), img)


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_fixture_decodes_in_library_and_oracle(path):
    data = open(path, "rb").read()
    exp = np.load(path[:-4] + ".npy", allow_pickle=False)
    raw, info = ca.dng_decode(data)
    np.testing.assert_array_equal(raw, exp)
    ref, meta = D.read_dng(data)
    np.testing.assert_array_equal(ref, exp)
    assert info.bayerPattern == D.bayer_of_cfa(meta[D.CFAPATTERN])
    assert info.black_level == meta[D.BLACKLEVEL][0] and info.white_level == meta[D.WHITELEVEL][0]


def test_fixture_set_covers_the_variants():
    names = {os.path.basename(p)[:-4] for p in FIXTURES}
    assert len(names) >= 8
    infos = [ca.dng_decode(open(p, "rb").read())[1] for p in FIXTURES]
    assert {i.bits_per_sample for i in infos} >= {8, 10, 12, 14, 16}
    assert {i.compression for i in infos} == {1, 7} and any(i.tiled for i in infos)


@pytest.mark.parametrize("awb", [False, True])
def test_metadata_to_parameters(awb):
    data = open(os.path.join(GOLD, "lj92_tiles_be.dng"), "rb").read()
    raw, info = ca.dng_decode(data)
    ref_raw, meta = D.read_dng(data)
    p, exposure, hl, asn = ca.unpack_dng_metadata(info, raw, auto_white_balance=awb)
    exp, e_ref = D.unpack_dng_metadata(meta, ref_raw.astype(np.float64), auto_white_balance=awb)
    assert p.bayerPattern == exp["bayerPattern"]
    assert p.black_level == exp["black_level"] and p.white_level == exp["white_level"]
    np.testing.assert_allclose(exposure, e_ref, rtol=1e-6)
    np.testing.assert_allclose(p.exposure_multiplier, exp["exposure_multiplier"], rtol=1e-6)
    np.testing.assert_allclose(np.array(p.scale_mul[:]), exp["scale_mul"], rtol=2e-6)
    rc = np.array([[p.rgb_cam[i][j] for j in range(3)] for i in range(3)])
    np.testing.assert_allclose(rc, exp["rgb_cam"], rtol=1e-5, atol=1e-6)
    if not awb:
        np.testing.assert_allclose(asn, np.array(meta[D.ASSHOTNEUTRAL]), rtol=1e-6)


def test_cam_xyz_coeff_closed_forms():
    """rgb_cam maps camera white (1, 1, 1) to (1, 1, 1); a camera whose
    matrix is sRGB's own (cam_xyz = inverse(xyz_rgb)) gets rgb_cam = I and
    pre_mul = 1."""
    info = ca.DngInfo()
    info.width, info.height, info.bayerPattern = 64, 48, ca.RGGB
    info.black_level, info.white_level = 0.0, 65535.0
    for cm in (np.linalg.inv(D.XYZ_RGB), np.array([[0.6722, -0.0635, -0.0963], [-0.4287, 1.2460, 0.2028],
                                                   [-0.0908, 0.2162, 0.5668]])):
        for k in range(9):
            info.color_matrix1[k] = float(cm.ravel()[k])
        info.has_color_matrix1 = 1
        p, _, _, asn = ca.unpack_dng_metadata(info)
        rc = np.array([[p.rgb_cam[i][j] for j in range(3)] for i in range(3)])
        np.testing.assert_allclose(rc @ np.ones(3), np.ones(3), atol=2e-6)
    np.testing.assert_allclose(np.linalg.inv(D.XYZ_RGB) @ D.XYZ_RGB, np.eye(3), atol=1e-12)
    for k in range(9):
        info.color_matrix1[k] = float(np.linalg.inv(D.XYZ_RGB).ravel()[k])
    p, _, _, asn = ca.unpack_dng_metadata(info)
    rc = np.array([[p.rgb_cam[i][j] for j in range(3)] for i in range(3)])
    np.testing.assert_allclose(rc, np.eye(3), atol=2e-6)
    np.testing.assert_allclose(asn, np.ones(3), atol=2e-6)
    np.testing.assert_allclose(np.array(p.scale_mul[:]), 1.0, atol=1e-5)


def test_awb_recovers_a_colour_cast():
    """A neutral synthetic scene seen through channel gains (0.5, 1, 0.8):
    the automatic white balance finds gains close to their inverse."""
    from oracle import oracle as O
    raw = O.generate_bayer(512, 384, 2, 9, 10, 0.002, 1e-5, 0.0, 65535.0).astype(np.float64)
    gains = {0: 0.5, 1: 1.0, 2: 0.8}
    for (ox, oy), c in zip([(0, 0), (1, 0), (0, 1), (1, 1)], [0, 1, 1, 2]):
        raw[oy::2, ox::2] *= gains[c]
    raw = raw.astype(np.uint16)
    info = ca.DngInfo()
    info.width, info.height, info.bayerPattern = 512, 384, ca.RGGB
    info.black_level, info.white_level = 0.0, 65535.0
    for k in range(9):
        info.color_matrix1[k] = float(np.linalg.inv(D.XYZ_RGB).ravel()[k])
    info.has_color_matrix1 = 1
    p, _, _, asn = ca.unpack_dng_metadata(info, raw, auto_white_balance=True)
    cam_mul = 1 / asn
    np.testing.assert_allclose(cam_mul / cam_mul[1], [2.0, 1.0, 1.25], rtol=0.15)
    exp, _ = D.unpack_dng_metadata({D.COLORMATRIX1: list(np.linalg.inv(D.XYZ_RGB).ravel()), D.BLACKLEVEL: [0],
                                    D.WHITELEVEL: [65535], D.CFAPATTERN: b"\0\1\1\2"},
                                   raw.astype(np.float64), auto_white_balance=True)
    np.testing.assert_allclose(np.array(p.scale_mul[:]), exp["scale_mul"], rtol=1e-5)


@pytest.mark.parametrize("mutate,msg", [
    (lambda d: d[:6], "TIFF"),
    (lambda d: b"XX" + d[2:], "byte order"),
    (lambda d: d[:len(d) // 2], ""),
])
def test_bad_files_fail_cleanly(mutate, msg):
    data = open(os.path.join(GOLD, "lj92_strip_nc2.dng"), "rb").read()
    with pytest.raises(ca.ClimError) as e:
        ca.dng_decode(mutate(data))
    assert msg in str(e.value)


def test_unsupported_compression_is_rejected():
    data = bytearray(D.write_dng(np.zeros((8, 8), np.uint16), 2, bits=16))
    # compression tag value 1 -> 8 (deflate)
    raw, meta = D.read_dng(bytes(data))
    i = bytes(data).find(bytes([0x03, 0x01, 0x03, 0x00, 0x01, 0x00, 0x00, 0x00, 0x01, 0x00]))
    assert i > 0
    data[i + 8] = 8
    with pytest.raises(ca.ClimError) as e:
        ca.dng_decode(bytes(data))
    assert "compression 8" in str(e.value)


EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "climage_amd", "lib", "raw_pipeline")


@pytest.mark.parametrize("name,awb", [("lj92_tiles_be", 0), ("le14_preview_crop", 0), ("le16_strip", 1)])
def test_cpp_api_read_dng_and_unpack_metadata(tmp_path, name, awb):
    """gls::image<luma_pixel_16>::read_dng_file + unpackDNGMetadata through the
    C++ headers (climage_amd/examples/raw_pipeline.cpp --dng-params, no GPU)."""
    import subprocess
    path = os.path.join(GOLD, name + ".dng")
    r = subprocess.run([EXE, "--dng-params", path, str(tmp_path / "o"), str(awb)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    v = np.array(open(tmp_path / "o.params").read().split(), np.float64)
    data = open(path, "rb").read()
    ref_raw, meta = D.read_dng(data)
    exp, e_ref = D.unpack_dng_metadata(meta, ref_raw.astype(np.float64), auto_white_balance=bool(awb))
    assert (v[0], v[1]) == (ref_raw.shape[1], ref_raw.shape[0])
    assert v[2] == exp["bayerPattern"] and v[3] == exp["black_level"] and v[4] == exp["white_level"]
    np.testing.assert_allclose(v[5:7], [exp["exposure_multiplier"], e_ref], rtol=1e-6)
    np.testing.assert_allclose(v[7:11], exp["scale_mul"], rtol=1e-5)
    np.testing.assert_allclose(v[11:20].reshape(3, 3), exp["rgb_cam"], rtol=1e-5, atol=1e-6)
