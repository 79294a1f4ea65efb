"""CPU oracle for DNG ingest.  TEST INFRASTRUCTURE ONLY: used by tests/ and
tests/golden/make_dng.py, never by the library.

Restated from the published specifications, since the reference reads DNG
files through libtiff (a third-party dependency, not in the reference tree)
and Adobe's DNG SDK lossless JPEG codec (src/gls_dng_lossless_jpeg.cpp:3253):

* TIFF 6.0 / DNG 1.4 container: byte order, IFD chain, SubIFDs, strips and
  tiles, uncompressed samples of 8..16 bits packed MSB first (TIFF FillOrder
  1), compression 7 = lossless JPEG.  The raw IFD is chosen as
  src/gls_image_tiff.cpp:268-295 does (IFD0, or its first SubIFD whose
  NewSubFileType has bit 0 clear when IFD0 is a preview); tags are merged
  with IFD0 first (std::map::insert keeps the first value,
  src/gls_tiff_metadata.cpp:80); the crop is DefaultCropOrigin + ActiveArea
  (top, left) and DefaultCropSize (src/gls_image_tiff.cpp:318-337).
* Lossless JPEG, ITU-T T.81 Annex H (process 14): SOF3, DHT, SOS with
  predictor Ss in 1..7 and point transform Al, DRI restart intervals,
  Huffman tables built per Annex C, difference categories and additional
  bits per F.1.2.1 / H.1.2.2, byte stuffing; a restart interval is a whole
  number of sample rows (as in the reference's decoder,
  src/gls_dng_lossless_jpeg.cpp:387,1173).  The encoder (for fixtures) builds
  optimal code lengths limited to 16 bits per Annex K.2.
* unpackDNGMetadata, cam_xyz_coeff and autoWhiteBalance
  (cpp/demosaic_utils.cpp:25-52, 280-375, 857-1026) in numpy.
"""
import struct

import numpy as np

# TIFF / DNG tags
NEWSUBFILETYPE, IMAGEWIDTH, IMAGELENGTH, BITSPERSAMPLE, COMPRESSION = 254, 256, 257, 258, 259
PHOTOMETRIC, STRIPOFFSETS, ORIENTATION, SAMPLESPERPIXEL, ROWSPERSTRIP = 262, 273, 274, 277, 278
STRIPBYTECOUNTS, PLANARCONFIG, TILEWIDTH, TILELENGTH, TILEOFFSETS, TILEBYTECOUNTS = 279, 284, 322, 323, 324, 325
SUBIFDS, SAMPLEFORMAT = 330, 339
CFAREPEATPATTERNDIM, CFAPATTERN = 33421, 33422
DNGVERSION, BLACKLEVEL, WHITELEVEL = 50706, 50714, 50717
DEFAULTCROPORIGIN, DEFAULTCROPSIZE = 50719, 50720
COLORMATRIX1, COLORMATRIX2, ASSHOTNEUTRAL, BASELINEEXPOSURE, ACTIVEAREA = 50721, 50722, 50728, 50730, 50829

_TSIZE = {1: 1, 2: 1, 3: 2, 4: 4, 5: 8, 6: 1, 7: 1, 8: 2, 9: 4, 10: 8, 11: 4, 12: 8, 13: 4}
_TFMT = {3: "H", 4: "I", 6: "b", 8: "h", 9: "i", 11: "f", 12: "d", 13: "I"}


# ----------------------------------------------------------------- TIFF reader
def _decode_value(e, typ, cnt, data):
    if typ in (1, 7):
        return bytes(data)
    if typ == 2:
        return bytes(data).split(b"\0")[0].decode("latin-1")
    if typ in (5, 10):
        f = "I" if typ == 5 else "i"
        v = struct.unpack(e + f * (2 * cnt), data)
        return [v[2 * i] / v[2 * i + 1] if v[2 * i + 1] else 0.0 for i in range(cnt)]
    return list(struct.unpack(e + _TFMT[typ] * cnt, data))


def _read_ifd(buf, e, off):
    n = struct.unpack(e + "H", buf[off:off + 2])[0]
    tags = {}
    for i in range(n):
        p = off + 2 + 12 * i
        tag, typ, cnt = struct.unpack(e + "HHI", buf[p:p + 8])
        if typ not in _TSIZE:
            continue
        size = _TSIZE[typ] * cnt
        if size <= 4:
            data = buf[p + 8:p + 8 + size]
        else:
            vo = struct.unpack(e + "I", buf[p + 8:p + 12])[0]
            data = buf[vo:vo + size]
        tags[tag] = _decode_value(e, typ, cnt, data)
    nxt = struct.unpack(e + "I", buf[off + 2 + 12 * n:off + 6 + 12 * n])[0]
    return tags, nxt


def _first(tags, tag, default=None):
    v = tags.get(tag)
    if v is None:
        return default
    return v[0] if isinstance(v, (list, bytes)) else v


def _unpack_bits(data, bits, count):
    """count samples of `bits` bits, MSB first, from bytes."""
    if bits == 16:
        raise AssertionError("16-bit samples are unpacked with the byte order")
    a = np.frombuffer(data, np.uint8)
    bitsarr = np.unpackbits(a)[:count * bits].reshape(count, bits).astype(np.uint32)
    w = (1 << np.arange(bits - 1, -1, -1)).astype(np.uint32)
    return (bitsarr * w).sum(1).astype(np.uint16)


def read_dng(buf):
    """Returns (raw uint16 (H, W) after the crop, merged metadata dict)."""
    bo = buf[:2]
    e = "<" if bo == b"II" else ">"
    assert bo in (b"II", b"MM") and struct.unpack(e + "H", buf[2:4])[0] == 42
    ifd0, _ = _read_ifd(buf, e, struct.unpack(e + "I", buf[4:8])[0])
    meta = dict(ifd0)
    ifd = ifd0
    if _first(ifd0, NEWSUBFILETYPE, 0) & 1:
        for so in ifd0.get(SUBIFDS, []):
            sub, _ = _read_ifd(buf, e, so)
            if (_first(sub, NEWSUBFILETYPE, 0) & 1) == 0:
                ifd = sub
                for k, v in sub.items():
                    meta.setdefault(k, v)
                break
    assert _first(ifd, SAMPLEFORMAT, 1) == 1
    spp = _first(ifd, SAMPLESPERPIXEL, 1)
    bits = _first(ifd, BITSPERSAMPLE, 1)
    comp = _first(ifd, COMPRESSION, 1)
    width, height = _first(ifd, IMAGEWIDTH), _first(ifd, IMAGELENGTH)
    assert spp == 1
    full = np.zeros((height, width), np.uint16)
    if TILEWIDTH in ifd:
        tw, th = _first(ifd, TILEWIDTH), _first(ifd, TILELENGTH)
        offs, cnts = ifd[TILEOFFSETS], ifd[TILEBYTECOUNTS]
        tx = (width + tw - 1) // tw
        for t, (o, c) in enumerate(zip(offs, cnts)):
            x0, y0 = tw * (t % tx), th * (t // tx)
            tile = _decode_block(buf[o:o + c], e, comp, bits, tw, th)
            hh, ww = min(th, height - y0), min(tw, width - x0)
            full[y0:y0 + hh, x0:x0 + ww] = tile[:hh, :ww]
    else:
        offs, cnts = ifd[STRIPOFFSETS], ifd[STRIPBYTECOUNTS]
        rps = _first(ifd, ROWSPERSTRIP, height)
        if comp == 7:
            assert len(offs) == 1, "one strip expected for compressed DNG files"
        for s, (o, c) in enumerate(zip(offs, cnts)):
            y0 = s * rps
            hh = min(rps, height - y0) if comp != 7 else height
            full[y0:y0 + hh] = _decode_block(buf[o:o + c], e, comp, bits, width, hh)
    cw, ch = width, height
    if DEFAULTCROPSIZE in meta:
        cw, ch = int(meta[DEFAULTCROPSIZE][0]), int(meta[DEFAULTCROPSIZE][1])
    cx = cy = 0
    if DEFAULTCROPORIGIN in meta:
        cx, cy = int(meta[DEFAULTCROPORIGIN][0]), int(meta[DEFAULTCROPORIGIN][1])
    if ACTIVEAREA in meta:
        cx += int(meta[ACTIVEAREA][1])
        cy += int(meta[ACTIVEAREA][0])
    out = np.zeros((ch, cw), np.uint16)
    src = full[cy:cy + ch, cx:cx + cw]
    out[:src.shape[0], :src.shape[1]] = src
    meta.setdefault(ORIENTATION, [_first(ifd, ORIENTATION, 1)])
    return out, meta


def _decode_block(data, e, comp, bits, w, h):
    if comp == 7:
        samples = lj92_decode(data)
        return np.asarray(samples, np.uint16).reshape(-1)[:w * h].reshape(h, w)
    assert comp == 1
    if bits == 16:
        return np.frombuffer(data[:2 * w * h], e + "u2").astype(np.uint16).reshape(h, w)
    if bits == 8:
        return np.frombuffer(data[:w * h], np.uint8).astype(np.uint16).reshape(h, w)
    # rows are packed continuously (TIFF strips: each row starts on a byte
    # boundary only for bits * w % 8 == 0, which DNG writers use)
    return _unpack_bits(data, bits, w * h).reshape(h, w)


# ------------------------------------------------------------ lossless JPEG
def _huff_codes(bits, vals):
    """Annex C: canonical codes from the 16 counts and the symbol list."""
    codes, code, k = {}, 0, 0
    for length in range(1, 17):
        for _ in range(bits[length - 1]):
            codes[(length, code)] = vals[k]
            code += 1
            k += 1
        code <<= 1
    return codes


class _Bits:
    def __init__(self, data, pos):
        self.d, self.p, self.acc, self.n = data, pos, 0, 0

    def bit(self):
        if self.n == 0:
            b = self.d[self.p]
            self.p += 1
            if b == 0xFF:
                nxt = self.d[self.p]
                if nxt == 0x00:
                    self.p += 1
                else:
                    raise ValueError("marker inside entropy-coded data")
            self.acc, self.n = b, 8
        self.n -= 1
        return (self.acc >> self.n) & 1

    def bits(self, k):
        v = 0
        for _ in range(k):
            v = (v << 1) | self.bit()
        return v

    def restart(self):
        """skip to the next RSTm marker and past it"""
        self.n = 0
        while not (self.d[self.p] == 0xFF and 0xD0 <= self.d[self.p + 1] <= 0xD7):
            self.p += 1
        self.p += 2


def _predict(sel, ra, rb, rc):
    if sel == 1:
        return ra
    if sel == 2:
        return rb
    if sel == 3:
        return rc
    if sel == 4:
        return ra + rb - rc
    if sel == 5:
        return ra + ((rb - rc) >> 1)
    if sel == 6:
        return rb + ((ra - rc) >> 1)
    return (ra + rb) >> 1


def lj92_decode(data):
    """Returns the samples as a (Y, X * Nc) array (components interleaved)."""
    data = bytes(data)
    assert data[0:2] == b"\xff\xd8"
    p = 2
    tables, restart = {}, 0
    while True:
        while data[p] != 0xFF:
            p += 1
        m = data[p + 1]
        p += 2
        if m in (0xD8, 0x01) or 0xD0 <= m <= 0xD7:
            continue
        ln = struct.unpack(">H", data[p:p + 2])[0]
        seg = data[p + 2:p + ln]
        if m == 0xC3:
            P, Y, X, nf = seg[0], struct.unpack(">H", seg[1:3])[0], struct.unpack(">H", seg[3:5])[0], seg[5]
            comps = [seg[6 + 3 * i] for i in range(nf)]
        elif m == 0xC4:
            q = 0
            while q < len(seg):
                tc_th = seg[q]
                counts = list(seg[q + 1:q + 17])
                vals = list(seg[q + 17:q + 17 + sum(counts)])
                tables[tc_th & 15] = _huff_codes(counts, vals)
                q += 17 + sum(counts)
        elif m == 0xDD:
            restart = struct.unpack(">H", seg[0:2])[0]
        elif m == 0xDA:
            ns = seg[0]
            sel_tab = {seg[1 + 2 * i]: seg[2 + 2 * i] >> 4 for i in range(ns)}
            ss, al = seg[1 + 2 * ns], seg[3 + 2 * ns] & 15
            p += ln
            break
        elif m == 0xD9:
            raise ValueError("EOI before SOS")
        p += ln
    nc = len(comps)
    if restart and restart % X:
        raise ValueError("restart interval not a whole number of rows")
    tab = [tables[sel_tab[c]] for c in comps]
    out = np.zeros((Y, X * nc), np.int64)
    br = _Bits(data, p)
    mcus = 0
    first = True  # the row after a restart uses the first-row rule
    for y in range(Y):
        for x in range(X):
            if restart and mcus and mcus % restart == 0:
                br.restart()
                first = True
            for c in range(nc):
                code, length = 0, 0
                while True:
                    code = (code << 1) | br.bit()
                    length += 1
                    if (length, code) in tab[c]:
                        s = tab[c][(length, code)]
                        break
                    if length > 16:
                        raise ValueError("bad Huffman code")
                if s == 0:
                    diff = 0
                elif s == 16:
                    diff = 32768
                else:
                    v = br.bits(s)
                    diff = v if v >= (1 << (s - 1)) else v - ((1 << s) - 1)
                col = x * nc + c
                if first and x == 0:
                    pred = 1 << (P - al - 1)
                elif first:
                    pred = out[y, col - nc]
                elif x == 0:
                    pred = out[y - 1, col]
                else:
                    pred = _predict(ss, out[y, col - nc], out[y - 1, col], out[y - 1, col - nc])
                out[y, col] = (pred + diff) & 0xFFFF
            mcus += 1
        first = False
    return (out << al) & 0xFFFF


def _category(d):
    if d == 0:
        return 0
    if d == 32768 or d == -32768:
        return 16
    return int(abs(d)).bit_length()


def _code_lengths(freq):
    """Annex K.2: optimal Huffman code lengths limited to 16 bits, with a
    reserved symbol so that no code is all ones."""
    freq = dict(freq)
    freq[256] = 1
    syms = sorted(freq)
    size = {s: 0 for s in syms}
    others = {s: None for s in syms}
    f = {s: freq[s] for s in syms}
    while True:
        live = [s for s in syms if f[s] > 0]
        if len(live) < 2:
            break
        v1 = min(live, key=lambda s: (f[s], -s))
        live2 = [s for s in live if s != v1]
        v2 = min(live2, key=lambda s: (f[s], -s))
        f[v1] += f[v2]
        f[v2] = 0
        size[v1] += 1
        s = v1
        while others[s] is not None:
            s = others[s]
            size[s] += 1
        others[s] = v2
        size[v2] += 1
        s = v2
        while others[s] is not None:
            s = others[s]
            size[s] += 1
    bits = [0] * 33
    for s in syms:
        if size[s]:
            bits[size[s]] += 1
    i = 32
    while i > 16:
        while bits[i] > 0:
            j = i - 2
            while bits[j] == 0:
                j -= 1
            bits[i] -= 2
            bits[i - 1] += 1
            bits[j + 1] += 2
            bits[j] -= 1
        i -= 1
    while bits[i] == 0:
        i -= 1
    bits[i] -= 1  # remove the reserved symbol
    order = sorted((s for s in syms if s != 256 and size[s]), key=lambda s: (size[s], s))
    return bits[1:17], order


def lj92_encode(samples, precision, nc=1, predictor=1, restart=0, point_transform=0):
    """samples: (Y, X * nc) array; returns the SOF3 byte stream."""
    a = np.asarray(samples, np.int64)
    Y, XN = a.shape
    X = XN // nc
    assert restart % X == 0, "restart intervals are whole rows"
    v = a >> point_transform
    diffs = np.zeros_like(v)
    mcus = 0
    first = True
    for y in range(Y):
        for x in range(X):
            if restart and mcus and mcus % restart == 0:
                first = True
            for c in range(nc):
                col = x * nc + c
                if first and x == 0:
                    pred = 1 << (precision - point_transform - 1)
                elif first:
                    pred = v[y, col - nc]
                elif x == 0:
                    pred = v[y - 1, col]
                else:
                    pred = _predict(predictor, v[y, col - nc], v[y - 1, col], v[y - 1, col - nc])
                d = (int(v[y, col]) - int(pred)) & 0xFFFF
                diffs[y, col] = d - 0x10000 if d >= 0x8000 else d
            mcus += 1
        first = False
    freq = {}
    for d in diffs.ravel():
        s = _category(int(d))
        freq[s] = freq.get(s, 0) + 1
    counts, order = _code_lengths(freq)
    code, k, codes = 0, 0, {}
    for length in range(1, 17):
        for _ in range(counts[length - 1]):
            codes[order[k]] = (code, length)
            code += 1
            k += 1
        code <<= 1
    out = bytearray(b"\xff\xd8")
    sof = bytes([precision]) + struct.pack(">HH", Y, X) + bytes([nc]) + b"".join(
        bytes([c, 0x11, 0]) for c in range(nc))
    out += b"\xff\xc3" + struct.pack(">H", 2 + len(sof)) + sof
    dht = bytes([0x00]) + bytes(counts) + bytes(order)
    out += b"\xff\xc4" + struct.pack(">H", 2 + len(dht)) + dht
    if restart:
        out += b"\xff\xdd" + struct.pack(">HH", 4, restart)
    sos = bytes([nc]) + b"".join(bytes([c, 0x00]) for c in range(nc)) + bytes([predictor, 0, point_transform])
    out += b"\xff\xda" + struct.pack(">H", 2 + len(sos)) + sos
    acc, nacc, mcus, rst = 0, 0, 0, 0

    def flush_bits(final=False):
        nonlocal acc, nacc
        while nacc >= 8:
            b = (acc >> (nacc - 8)) & 0xFF
            out.append(b)
            if b == 0xFF:
                out.append(0)
            nacc -= 8
        if final and nacc:
            b = ((acc << (8 - nacc)) | ((1 << (8 - nacc)) - 1)) & 0xFF  # pad with ones
            out.append(b)
            if b == 0xFF:
                out.append(0)
            nacc = 0
        acc &= (1 << nacc) - 1

    for y in range(Y):
        for x in range(X):
            if restart and mcus and mcus % restart == 0:
                flush_bits(final=True)
                out += bytes([0xFF, 0xD0 + (rst & 7)])
                rst += 1
            for c in range(nc):
                d = int(diffs[y, x * nc + c])
                s = _category(d)
                cd, ln = codes[s]
                acc = (acc << ln) | cd
                nacc += ln
                if 0 < s < 16:
                    extra = d if d > 0 else d + (1 << s) - 1
                    acc = (acc << s) | extra
                    nacc += s
                flush_bits()
            mcus += 1
    flush_bits(final=True)
    out += b"\xff\xd9"
    return bytes(out)


# ----------------------------------------------------------------- TIFF writer
def _entry(e, tag, typ, values):
    if typ in (1, 7):
        data = bytes(values)
        cnt = len(data)
    elif typ == 2:
        data = values.encode("latin-1") + b"\0"
        cnt = len(data)
    elif typ in (5, 10):
        f = "I" if typ == 5 else "i"
        den = 10000
        data = b"".join(struct.pack(e + f + f, int(round(v * den)), den) for v in values)
        cnt = len(values)
    else:
        data = struct.pack(e + _TFMT[typ] * len(values), *values)
        cnt = len(values)
    return tag, typ, cnt, data


def write_dng(raw, bayer, black=0, white=65535, bits=16, compression=1, tile=None, big_endian=False,
              preview=False, crop_origin=None, crop_size=None, active_area=None, color_matrix=None,
              color_matrix2=None, as_shot_neutral=None, baseline_exposure=None, lj92_nc=2, predictor=1,
              restart=0, orientation=1, lj92_stream=None, rows_per_strip=None, point_transform=0):
    """A DNG holding `raw` (uint16 (H, W)).  bayer: 0 grbg, 1 gbrg, 2 rggb,
    3 bggr.  compression 1 (bits 8..16, MSB-first packing) or 7 (lossless
    JPEG, lj92_nc components per JPEG sample row).  tile: (tw, th) or None
    (one strip).  preview: IFD0 is an 8x8 preview, the raw is its SubIFD.
    lj92_stream: a ready lossless JPEG stream for the single strip."""
    e = ">" if big_endian else "<"
    raw = np.asarray(raw, np.uint16)
    h, w = raw.shape
    cfa = {0: [1, 0, 2, 1], 1: [1, 2, 0, 1], 2: [0, 1, 1, 2], 3: [2, 1, 1, 0]}[bayer]

    def block_bytes(blk):
        bh, bw = blk.shape
        if compression == 7:
            assert bw % lj92_nc == 0
            return lj92_encode(blk.reshape(bh, bw), bits, nc=lj92_nc, predictor=predictor, restart=restart,
                               point_transform=point_transform)
        if bits == 16:
            return blk.astype(e + "u2").tobytes()
        if bits == 8:
            return blk.astype(np.uint8).tobytes()
        v = blk.ravel().astype(np.uint32)
        bitarr = ((v[:, None] >> np.arange(bits - 1, -1, -1)) & 1).astype(np.uint8).ravel()
        return np.packbits(bitarr).tobytes()

    blocks = []
    if lj92_stream is not None:
        assert compression == 7 and not tile
        blocks.append(bytes(lj92_stream))
    elif tile:
        tw, th = tile
        for y0 in range(0, h, th):
            for x0 in range(0, w, tw):
                blk = np.zeros((th, tw), np.uint16)
                src = raw[y0:y0 + th, x0:x0 + tw]
                blk[:src.shape[0], :src.shape[1]] = src
                blocks.append(block_bytes(blk))
    elif rows_per_strip and compression == 1:
        for y0 in range(0, h, rows_per_strip):
            blocks.append(block_bytes(raw[y0:y0 + rows_per_strip]))
    else:
        blocks.append(block_bytes(raw))

    raw_tags = [(NEWSUBFILETYPE, 4, [0]), (IMAGEWIDTH, 4, [w]), (IMAGELENGTH, 4, [h]), (BITSPERSAMPLE, 3, [bits]),
                (COMPRESSION, 3, [compression]), (PHOTOMETRIC, 3, [32803]), (SAMPLESPERPIXEL, 3, [1]),
                (PLANARCONFIG, 3, [1]), (CFAREPEATPATTERNDIM, 3, [2, 2]), (CFAPATTERN, 1, cfa),
                (BLACKLEVEL, 4, [int(black)]), (WHITELEVEL, 4, [int(white)])]
    if crop_origin is not None:
        raw_tags.append((DEFAULTCROPORIGIN, 4, list(crop_origin)))
    if crop_size is not None:
        raw_tags.append((DEFAULTCROPSIZE, 4, list(crop_size)))
    if active_area is not None:
        raw_tags.append((ACTIVEAREA, 4, list(active_area)))
    if tile:
        raw_tags += [(TILEWIDTH, 4, [tile[0]]), (TILELENGTH, 4, [tile[1]])]
    meta_tags = [(DNGVERSION, 1, [1, 4, 0, 0]), (ORIENTATION, 3, [orientation])]
    if color_matrix is not None:
        meta_tags.append((COLORMATRIX1, 10, list(np.asarray(color_matrix).ravel())))
    if color_matrix2 is not None:
        meta_tags.append((COLORMATRIX2, 10, list(np.asarray(color_matrix2).ravel())))
    if as_shot_neutral is not None:
        meta_tags.append((ASSHOTNEUTRAL, 5, list(as_shot_neutral)))
    if baseline_exposure is not None:
        meta_tags.append((BASELINEEXPOSURE, 10, [baseline_exposure]))

    out = bytearray(b"II*\0" if not big_endian else b"MM\0*")
    out += b"\0\0\0\0"
    data_off = []
    for b in blocks:
        data_off.append(len(out))
        out += b
        if len(out) & 1:
            out += b"\0"
    pv = b""
    if preview:
        pv_off = len(out)
        pv = bytes(8 * 8)
        out += pv

    def put_ifd(tags, nxt=0):
        ents = []
        for tag, typ, vals in sorted(tags, key=lambda t: t[0]):
            ents.append(_entry(e, tag, typ, vals))
        # out-of-line data first, then the IFD
        ext = {}
        for tag, typ, cnt, data in ents:
            if len(data) > 4:
                if len(out) & 1:
                    out.append(0)
                ext[tag] = len(out)
                out.extend(data)
        if len(out) & 1:
            out.append(0)
        off = len(out)
        out.extend(struct.pack(e + "H", len(ents)))
        for tag, typ, cnt, data in ents:
            if tag in ext:
                out.extend(struct.pack(e + "HHII", tag, typ, cnt, ext[tag]))
            else:
                out.extend(struct.pack(e + "HHI", tag, typ, cnt) + data + b"\0" * (4 - len(data)))
        out.extend(struct.pack(e + "I", nxt))
        return off

    counts = [len(b) for b in blocks]
    if tile:
        raw_tags += [(TILEOFFSETS, 4, data_off), (TILEBYTECOUNTS, 4, counts)]
    else:
        rps = rows_per_strip if (rows_per_strip and compression == 1) else h
        raw_tags += [(STRIPOFFSETS, 4, data_off), (ROWSPERSTRIP, 4, [rps]), (STRIPBYTECOUNTS, 4, counts)]
    if preview:
        sub_off = put_ifd(raw_tags)
        ifd0 = [(NEWSUBFILETYPE, 4, [1]), (IMAGEWIDTH, 4, [8]), (IMAGELENGTH, 4, [8]), (BITSPERSAMPLE, 3, [8]),
                (COMPRESSION, 3, [1]), (PHOTOMETRIC, 3, [1]), (SAMPLESPERPIXEL, 3, [1]),
                (STRIPOFFSETS, 4, [pv_off]), (ROWSPERSTRIP, 4, [8]), (STRIPBYTECOUNTS, 4, [64]),
                (SUBIFDS, 4, [sub_off])] + meta_tags
        first = put_ifd(ifd0)
    else:
        first = put_ifd(raw_tags + meta_tags)
    out[4:8] = struct.pack(e + "I", first)
    return bytes(out)


# ----------------------------------------------------- metadata -> parameters
# sRGB -> XYZ D65 (cpp/demosaic.hpp:91-95)
XYZ_RGB = np.array([[0.4124564, 0.3575761, 0.1804375],
                    [0.2126729, 0.7151522, 0.0721750],
                    [0.0193339, 0.1191920, 0.9503041]])


def cam_xyz_coeff(cam_xyz):
    """cpp/demosaic_utils.cpp:25-52: (rgb_cam, pre_mul)."""
    rgb_cam = np.asarray(cam_xyz, np.float64) @ XYZ_RGB
    cam_white = rgb_cam @ np.ones(3)
    if np.any(cam_white <= 0.00001):
        raise ValueError("cam_xyz_coeff: non-positive camera white")
    pre_mul = 1.0 / cam_white
    return np.linalg.inv(np.diag(pre_mul) @ rgb_cam), pre_mul


def bayer_of_cfa(cfa):
    """cpp/demosaic_utils.cpp:302-305"""
    cfa = bytes(cfa[:4])
    return {b"\0\1\1\2": 2, b"\2\1\1\0": 3, b"\1\0\2\1": 0}.get(cfa, 1)


def unpack_dng_metadata(meta, raw=None, auto_white_balance=False):
    """cpp/demosaic_utils.cpp:280-375 without ColorChecker calibration: a dict
    of the DemosaicParameters fields it sets, and the exposure multiplier."""
    cm = meta.get(COLORMATRIX2) or meta.get(COLORMATRIX1)
    cam_xyz = np.asarray(cm, np.float64).reshape(3, 3)
    asn = meta.get(ASSHOTNEUTRAL)
    be = meta.get(BASELINEEXPOSURE, [0.0])[0]
    exposure = 2.0 ** be
    black = float(meta[BLACKLEVEL][0]) if BLACKLEVEL in meta else 0.0
    white = float(meta[WHITELEVEL][0]) if WHITELEVEL in meta else 65535.0
    bayer = bayer_of_cfa(meta[CFAPATTERN])
    rgb_cam, pre_mul = cam_xyz_coeff(cam_xyz)
    if asn is not None and not auto_white_balance:
        pre_mul = 1.0 / np.asarray(asn, np.float64)

    def scale(pm):
        return np.array([max(exposure, 1.0) * (pm[1 if c == 3 else c] / pm.min()) * 65535.0 / (white - black)
                         for c in range(4)])

    scale_mul = scale(pre_mul)
    if auto_white_balance:
        c2y = _cam_ycbcr(rgb_cam)
        cam_mul = auto_white_balance_gain(raw, c2y, scale_mul, white, black, bayer)[0]
        pre_mul = cam_mul
        scale_mul = scale(pre_mul)
    return {"bayerPattern": bayer, "black_level": black, "white_level": white,
            "exposure_multiplier": min(exposure, 1.0), "scale_mul": scale_mul, "rgb_cam": rgb_cam}, exposure


def _cam_ycbcr(rgb_cam):
    k = XYZ_RGB[1] @ rgb_cam
    KR, KG, KB = k
    return np.array([[KR, KG, KB],
                     [-0.5 * KR / (1 - KB), -0.5 * KG / (1 - KB), 0.5],
                     [0.5, -0.5 * KG / (1 - KR), -0.5 * KB / (1 - KR)]])


def _quads(raw, bayer):
    """(H/2, W/2, 3) of (R, (G + G2) / 2, B) with the C++ Bayer table of
    cpp/demosaic.hpp:82-88 (offsets (x, y) of R, G, B, G2)."""
    off = {0: [(1, 0), (0, 0), (0, 1), (1, 1)], 1: [(0, 1), (0, 0), (1, 0), (1, 1)],
           2: [(0, 0), (1, 0), (1, 1), (0, 1)], 3: [(1, 1), (1, 0), (0, 0), (0, 1)]}[bayer]
    r = raw.astype(np.float64)
    h2, w2 = raw.shape[0] // 2, raw.shape[1] // 2
    pl = [r[oy:oy + 2 * h2:2, ox:ox + 2 * w2:2] for ox, oy in off]
    return np.stack([pl[0], (pl[1] + pl[3]) / 2.0, pl[2]], -1)


def _awb_tile(raw, c2y, scale_mul3, white, black, bayer, frac):
    """autoWhiteBalanceKernel, cpp/demosaic_utils.cpp:864-955.  The histogram
    bin of a negative luma is round() cast to size_t in the reference, which
    is undefined; it is taken as 0 (the saturating conversion of ARM64, the
    reference's target)."""
    q = _quads(raw, bayer)
    rgb = scale_mul3 * (q - black) / 65535.0
    hl = ((rgb > 0.5) & ~(rgb > 1.0)).any(-1)
    rgb = np.where(rgb > 1.0, 1.0, rgb)
    yuv = rgb @ c2y.T
    M = yuv.reshape(-1, 3).mean(0)
    D = np.abs(yuv - M).reshape(-1, 3).mean(0)
    sel = (np.abs(yuv[..., 1] - (M[1] + np.copysign(D[1], M[1]))) < 1.5 * D[1]) & \
          (np.abs(yuv[..., 2] - (1.5 * M[2] + np.copysign(D[2], M[2]))) < 1.5 * D[2])
    if not sel.any():
        return np.ones(3), 0
    Y = yuv[..., 0][sel]
    rgbw = (q[sel] - black) / white
    ymax = max(Y.max(), 0.0)
    ent = np.clip(np.round(127 * Y), 0, 127).astype(int)
    top = int(np.clip(np.round(127 * ymax), 0, 127))
    if top == 0:
        return np.ones(3), 0  # (cpp/demosaic_utils.cpp:930-932)
    acc, cnt = np.zeros(3), 0
    total = sel.sum()
    for i in range(top, -1, -1):
        m = ent == i
        acc += rgbw[m].sum(0)
        cnt += int(m.sum())
        if cnt > frac * total:
            break
    g = ymax / (acc / cnt)
    return g / g[1], int(hl.sum())


def auto_white_balance_gain(raw, c2y, scale_mul4, white, black, bayer):
    """autoWhiteBalance, cpp/demosaic_utils.cpp:967-1026: the mean of the 4x4
    tiles' gains, normalised to green; and the highlight pixel fraction."""
    sm = np.asarray(scale_mul4[:3], np.float64) / scale_mul4[1]
    h, w = raw.shape
    tw, th = 2 * ((w // 4) // 2), 2 * ((h // 4) // 2)
    gain, hl = np.zeros(3), 0
    for ty in range(4):
        for tx in range(4):
            g, n = _awb_tile(raw[ty * th:(ty + 1) * th, tx * tw:(tx + 1) * tw], c2y, sm, white, black, bayer, 0.01)
            gain += g
            hl += n
    gain /= 16
    gain /= gain[1]
    return gain, hl / (w * h / 4)
