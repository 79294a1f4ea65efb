// dng.cpp - DNG ingest on the host: read_dng_file (src/gls_image_tiff.cpp:255-458)
// and unpackDNGMetadata / cam_xyz_coeff / autoWhiteBalance
// (cpp/demosaic_utils.cpp:25-52, 280-375, 857-1026).
//
// The reference reads the container with libtiff and decodes compressed raw
// data with Adobe's DNG SDK lossless JPEG codec
// (src/gls_dng_lossless_jpeg.cpp:3253).  Here both are written from the
// specifications: TIFF 6.0 / DNG 1.4 (IFDs, SubIFDs, strips, tiles, samples of
// 8..16 bits packed MSB first) and ITU-T T.81 Annex H (lossless JPEG, SOF3:
// Huffman tables per Annex C, predictors 1-7, point transform, restart
// intervals).  Tiles decode in parallel on host threads.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "climage.h"

namespace {

thread_local std::string g_dng_error;

clim_status dng_fail(clim_status s, const std::string& msg) {
    g_dng_error = msg;
    return s;
}

struct DngError {
    std::string msg;
};

[[noreturn]] void fail(const std::string& msg) { throw DngError{msg}; }

// largest raw image decoded (samples of the full, uncropped image): 1 Gi
// samples (2 GiB), far above any camera's mosaic
constexpr size_t kMaxSamples = size_t(1) << 30;

// ------------------------------------------------------------ TIFF container
enum : uint16_t {
    kNewSubFileType = 254, kImageWidth = 256, kImageLength = 257, kBitsPerSample = 258, kCompression = 259,
    kStripOffsets = 273, kOrientation = 274, kSamplesPerPixel = 277, kRowsPerStrip = 278, kStripByteCounts = 279,
    kTileWidth = 322, kTileLength = 323, kTileOffsets = 324, kTileByteCounts = 325, kSubIFDs = 330,
    kSampleFormat = 339, kCFAPattern = 33422, kBlackLevel = 50714, kWhiteLevel = 50717,
    kDefaultCropOrigin = 50719, kDefaultCropSize = 50720, kColorMatrix1 = 50721, kColorMatrix2 = 50722,
    kAsShotNeutral = 50728, kBaselineExposure = 50730, kActiveArea = 50829
};

struct Tag {
    uint16_t type = 0;
    std::vector<double> num;     // numeric values
    std::vector<uint8_t> bytes;  // BYTE / UNDEFINED / ASCII
};

struct Ifd {
    std::vector<std::pair<uint16_t, Tag>> tags;
    const Tag* find(uint16_t t) const {
        for (const auto& e : tags)
            if (e.first == t) return &e.second;
        return nullptr;
    }
    double get(uint16_t t, double dflt) const {
        const Tag* g = find(t);
        return g && !g->num.empty() ? g->num[0] : dflt;
    }
};

class Tiff {
   public:
    Tiff(const uint8_t* d, size_t n) : d_(d), n_(n) {
        if (n < 8) fail("not a TIFF file (too short)");
        if (d[0] == 'I' && d[1] == 'I') le_ = true;
        else if (d[0] == 'M' && d[1] == 'M') le_ = false;
        else fail("not a TIFF file (byte order)");
        if (u16(2) != 42) fail("not a TIFF file (magic)");
    }
    uint32_t first_ifd() const { return u32(4); }
    uint16_t u16(size_t o) const {
        need(o, 2);
        return le_ ? (uint16_t)(d_[o] | d_[o + 1] << 8) : (uint16_t)(d_[o] << 8 | d_[o + 1]);
    }
    uint32_t u32(size_t o) const {
        need(o, 4);
        return le_ ? (uint32_t)d_[o] | (uint32_t)d_[o + 1] << 8 | (uint32_t)d_[o + 2] << 16 | (uint32_t)d_[o + 3] << 24
                   : (uint32_t)d_[o] << 24 | (uint32_t)d_[o + 1] << 16 | (uint32_t)d_[o + 2] << 8 | (uint32_t)d_[o + 3];
    }
    void need(size_t o, size_t k) const {
        if (o > n_ || k > n_ - o) fail("truncated TIFF data");
    }
    bool little_endian() const { return le_; }
    const uint8_t* data() const { return d_; }
    size_t size() const { return n_; }

    Ifd read_ifd(uint32_t off) const {
        static const int kSize[14] = {0, 1, 1, 2, 4, 8, 1, 1, 2, 4, 8, 4, 8, 4};
        Ifd ifd;
        const uint16_t n = u16(off);
        for (uint32_t i = 0; i < n; i++) {
            const size_t p = off + 2 + 12 * (size_t)i;
            const uint16_t tag = u16(p), type = u16(p + 2);
            const uint32_t cnt = u32(p + 4);
            if (type == 0 || type > 13) continue;
            const uint64_t size = (uint64_t)kSize[type] * cnt;
            if (size > n_) fail("bad TIFF tag size");
            const size_t vo = size <= 4 ? p + 8 : u32(p + 8);
            need(vo, (size_t)size);
            Tag t;
            t.type = type;
            for (uint32_t k = 0; k < cnt; k++) {
                const size_t q = vo + (size_t)k * kSize[type];
                switch (type) {
                    case 1: case 2: case 7: t.bytes.push_back(d_[q]); break;
                    case 3: t.num.push_back(u16(q)); break;
                    case 4: case 13: t.num.push_back(u32(q)); break;
                    case 6: t.num.push_back((int8_t)d_[q]); break;
                    case 8: t.num.push_back((int16_t)u16(q)); break;
                    case 9: t.num.push_back((int32_t)u32(q)); break;
                    case 5: {
                        const uint32_t a = u32(q), b = u32(q + 4);
                        t.num.push_back(b ? (double)a / b : 0.0);
                        break;
                    }
                    case 10: {
                        const int32_t a = (int32_t)u32(q), b = (int32_t)u32(q + 4);
                        t.num.push_back(b ? (double)a / b : 0.0);
                        break;
                    }
                    case 11: {
                        const uint32_t u = u32(q);
                        float f;
                        std::memcpy(&f, &u, 4);
                        t.num.push_back(f);
                        break;
                    }
                    case 12: {
                        const uint64_t u = le_ ? (uint64_t)u32(q) | (uint64_t)u32(q + 4) << 32
                                               : (uint64_t)u32(q) << 32 | u32(q + 4);
                        double f;
                        std::memcpy(&f, &u, 8);
                        t.num.push_back(f);
                        break;
                    }
                }
            }
            if (type == 1 || type == 7)
                for (uint8_t b : t.bytes) t.num.push_back(b);
            ifd.tags.emplace_back(tag, std::move(t));
        }
        return ifd;
    }

   private:
    const uint8_t* d_;
    size_t n_;
    bool le_ = true;
};

// ------------------------------------------------------- lossless JPEG (T.81)
struct Huff {
    // Annex C canonical codes: per length the largest code and the index of
    // its first symbol; a 9-bit lookup table for the short codes
    int maxcode[18], valptr[17], mincode[17];
    uint8_t vals[256];
    uint16_t lut[512];  // (length << 8) | symbol, 0 = longer than 9 bits
};

void build_huff(const uint8_t counts[16], const uint8_t* vals, int nvals, Huff& h) {
    std::memcpy(h.vals, vals, nvals);
    std::memset(h.lut, 0, sizeof(h.lut));
    int code = 0, k = 0;
    for (int len = 1; len <= 16; len++) {
        h.valptr[len] = k;
        h.mincode[len] = code;
        // Annex C: the codes of one length must fit in `len` bits; checked
        // before any lookup entry of this length is written
        if (code + counts[len - 1] > (1 << len)) fail("lossless JPEG: bad Huffman table");
        if (k + counts[len - 1] > nvals) fail("lossless JPEG: Huffman table with fewer symbols than codes");
        for (int i = 0; i < counts[len - 1]; i++, k++, code++) {
            if (len <= 9) {
                const int shift = 9 - len;
                for (int j = 0; j < (1 << shift); j++) h.lut[(code << shift) | j] = (uint16_t)(len << 8 | h.vals[k]);
            }
        }
        h.maxcode[len] = counts[len - 1] ? code - 1 : -1;
        code <<= 1;
    }
    h.maxcode[17] = 0x7fffffff;
}

// MSB-first bits of the entropy-coded segment; 0xFF00 is a stuffed 0xFF; a
// marker ends the data (zeros are fed past it, as T.81 decoders do)
class BitReader {
   public:
    BitReader(const uint8_t* d, size_t n, size_t p) : d_(d), n_(n), p_(p) {}
    uint32_t peek(int k) {
        fill();
        return (uint32_t)(acc_ >> (nbits_ - k)) & ((1u << k) - 1);
    }
    void skip(int k) { nbits_ -= k; }
    uint32_t bits(int k) {
        if (k == 0) return 0;
        const uint32_t v = peek(k);
        skip(k);
        return v;
    }
    // drop the rest of the byte and the next RSTm marker
    void restart() {
        nbits_ = 0;
        acc_ = 0;
        marker_ = false;
        while (p_ + 1 < n_ && !(d_[p_] == 0xFF && d_[p_ + 1] >= 0xD0 && d_[p_ + 1] <= 0xD7)) p_++;
        if (p_ + 1 >= n_) fail("lossless JPEG: missing restart marker");
        p_ += 2;
    }

   private:
    void fill() {
        while (nbits_ <= 56) {
            uint8_t b = 0;
            if (!marker_ && p_ < n_) {
                b = d_[p_];
                if (b == 0xFF) {
                    if (p_ + 1 < n_ && d_[p_ + 1] == 0x00) {
                        p_ += 2;
                    } else {
                        marker_ = true;
                        b = 0;
                    }
                } else {
                    p_++;
                }
            }
            acc_ = acc_ << 8 | b;
            nbits_ += 8;
        }
    }
    const uint8_t* d_;
    size_t n_, p_;
    uint64_t acc_ = 0;
    int nbits_ = 0;
    bool marker_ = false;
};

int decode_symbol(BitReader& br, const Huff& h) {
    const uint32_t look = br.peek(9);
    const uint16_t e = h.lut[look];
    if (e) {
        br.skip(e >> 8);
        return e & 0xff;
    }
    int code = (int)br.peek(9);
    br.skip(9);
    for (int len = 10; len <= 16; len++) {
        code = code << 1 | (int)br.bits(1);
        if (h.maxcode[len] >= 0 && code <= h.maxcode[len]) return h.vals[h.valptr[len] + code - h.mincode[len]];
    }
    fail("lossless JPEG: bad Huffman code");
}

// Decodes one SOF3 stream into out (count samples, components interleaved).
void lj92_decode(const uint8_t* d, size_t n, uint16_t* out, size_t count) {
    if (n < 4 || d[0] != 0xFF || d[1] != 0xD8) fail("lossless JPEG: no SOI");
    size_t p = 2;
    Huff tables[4];
    bool have[4] = {false, false, false, false};
    int P = 0, Y = 0, X = 0, nc = 0, comp_id[4] = {0}, comp_tab[4] = {0}, pred = 1, al = 0, restart = 0;
    for (;;) {
        while (p < n && d[p] != 0xFF) p++;
        while (p < n && d[p] == 0xFF) p++;
        if (p >= n) fail("lossless JPEG: no SOS");
        const uint8_t m = d[p++];
        if (m == 0xD8 || m == 0x01 || (m >= 0xD0 && m <= 0xD7)) continue;
        if (m == 0xD9) fail("lossless JPEG: EOI before SOS");
        if (p + 2 > n) fail("lossless JPEG: truncated segment");
        const size_t len = (size_t)d[p] << 8 | d[p + 1];
        if (len < 2 || p + len > n) fail("lossless JPEG: truncated segment");
        const uint8_t* s = d + p + 2;
        const size_t sl = len - 2;
        if (m == 0xC3) {
            if (sl < 6) fail("lossless JPEG: short SOF3");
            P = s[0];
            Y = s[1] << 8 | s[2];
            X = s[3] << 8 | s[4];
            nc = s[5];
            if (nc < 1 || nc > 4 || sl < 6 + 3 * (size_t)nc) fail("lossless JPEG: bad SOF3");
            for (int c = 0; c < nc; c++) comp_id[c] = s[6 + 3 * c];
        } else if (m == 0xC0 || m == 0xC1 || m == 0xC2 || (m >= 0xC5 && m <= 0xCF && m != 0xC8 && m != 0xCC)) {
            fail("not a lossless (SOF3) JPEG stream");
        } else if (m == 0xC4) {
            size_t q = 0;
            while (q + 17 <= sl) {
                const int th = s[q] & 15;
                if (th > 3) fail("lossless JPEG: bad table id");
                int tot = 0;
                for (int i = 0; i < 16; i++) tot += s[q + 1 + i];
                if (tot > 256 || q + 17 + tot > sl) fail("lossless JPEG: bad DHT");
                build_huff(s + q + 1, s + q + 17, tot, tables[th]);
                have[th] = true;
                q += 17 + tot;
            }
        } else if (m == 0xDD) {
            if (sl < 2) fail("lossless JPEG: bad DRI");
            restart = s[0] << 8 | s[1];
        } else if (m == 0xDA) {
            const int ns = sl ? s[0] : 0;
            if (ns != nc || sl < 4 + 2 * (size_t)ns) fail("lossless JPEG: bad SOS");
            for (int i = 0; i < ns; i++) {
                int c = 0;
                while (c < nc && comp_id[c] != s[1 + 2 * i]) c++;
                if (c == nc) fail("lossless JPEG: SOS names an unknown component");
                comp_tab[c] = s[2 + 2 * i] >> 4;
                if (comp_tab[c] > 3 || !have[comp_tab[c]]) fail("lossless JPEG: missing Huffman table");
            }
            pred = s[1 + 2 * ns];
            al = s[3 + 2 * ns] & 15;
            p += len;
            break;
        }
        p += len;
    }
    if (P < 2 || P > 16 || X <= 0 || Y <= 0 || pred < 1 || pred > 7) fail("lossless JPEG: unsupported frame");
    // the first predictor 2^(P - Pt - 1) needs a point transform below P
    if (al >= P) fail("lossless JPEG: point transform not below the sample precision");
    // a restart interval is a whole number of rows (T.81 H.1.1; the reference's
    // decoder counts it in rows, src/gls_dng_lossless_jpeg.cpp:387,1173)
    if (restart && restart % X) fail("lossless JPEG: restart interval not a whole number of rows");
    if ((size_t)X * Y * nc < count) fail("lossless JPEG: fewer samples than the tile holds");
    BitReader br(d, n, p);
    const size_t row = (size_t)X * nc;
    std::vector<uint16_t> prev(row), cur(row);
    size_t written = 0;
    long mcus = 0;
    bool first = true;
    for (int y = 0; y < Y; y++) {
        for (int x = 0; x < X; x++) {
            if (restart && mcus && mcus % restart == 0) {
                br.restart();
                first = true;
            }
            for (int c = 0; c < nc; c++) {
                const int ssss = decode_symbol(br, tables[comp_tab[c]]);
                int diff;
                if (ssss == 0) diff = 0;
                else if (ssss == 16) diff = 32768;
                else if (ssss > 16) fail("lossless JPEG: bad difference category");
                else {
                    const int v = (int)br.bits(ssss);
                    diff = v >= (1 << (ssss - 1)) ? v : v - ((1 << ssss) - 1);
                }
                const size_t col = (size_t)x * nc + c;
                int pv;
                if (first && x == 0) pv = 1 << (P - al - 1);
                else if (first) pv = cur[col - nc];
                else if (x == 0) pv = prev[col];
                else {
                    const int ra = cur[col - nc], rb = prev[col], rc = prev[col - nc];
                    switch (pred) {
                        case 1: pv = ra; break;
                        case 2: pv = rb; break;
                        case 3: pv = rc; break;
                        case 4: pv = ra + rb - rc; break;
                        case 5: pv = ra + ((rb - rc) >> 1); break;
                        case 6: pv = rb + ((ra - rc) >> 1); break;
                        default: pv = (ra + rb) >> 1; break;
                    }
                }
                cur[col] = (uint16_t)((pv + diff) & 0xFFFF);
            }
            mcus++;
        }
        first = false;
        for (size_t i = 0; i < row && written < count; i++) out[written++] = (uint16_t)(cur[i] << al);
        std::swap(prev, cur);
    }
}

// ------------------------------------------------------------ the raw image
struct Decoded {
    clim_dng_info info{};
    int full_w = 0, full_h = 0, crop_x = 0, crop_y = 0;
};

struct RawIfd {
    Ifd ifd;
    Ifd merged;  // IFD0 tags, then the raw IFD's tags that IFD0 lacks
};

RawIfd find_raw_ifd(const Tiff& t) {
    RawIfd r;
    r.ifd = t.read_ifd(t.first_ifd());
    r.merged = r.ifd;
    // IFD0 a preview: the first SubIFD that is not (src/gls_image_tiff.cpp:268-295)
    if ((int)r.ifd.get(kNewSubFileType, 0) & 1) {
        const Tag* subs = r.ifd.find(kSubIFDs);
        if (subs) {
            for (double so : subs->num) {
                Ifd sub = t.read_ifd((uint32_t)so);
                if (((int)sub.get(kNewSubFileType, 0) & 1) == 0) {
                    for (auto& e : sub.tags)
                        if (!r.merged.find(e.first)) r.merged.tags.push_back(e);
                    r.ifd = std::move(sub);
                    break;
                }
            }
        }
    }
    return r;
}

void fill_info(const RawIfd& r, Decoded& dec) {
    const Ifd& ifd = r.ifd;
    const Ifd& m = r.merged;
    clim_dng_info& in = dec.info;
    if (ifd.get(kSampleFormat, 1) != 1) fail("can not read sample format other than uint");
    if (ifd.get(kSamplesPerPixel, 1) != 1) fail("DNG raw data with more than one sample per pixel is not supported");
    dec.full_w = (int)ifd.get(kImageWidth, 0);
    dec.full_h = (int)ifd.get(kImageLength, 0);
    if (dec.full_w <= 0 || dec.full_h <= 0) fail("DNG: missing image size");
    in.bits_per_sample = (int)ifd.get(kBitsPerSample, 1);
    in.compression = (int)ifd.get(kCompression, 1);
    in.tiled = ifd.find(kTileWidth) != nullptr;
    // crop (src/gls_image_tiff.cpp:318-337)
    in.width = dec.full_w;
    in.height = dec.full_h;
    if (const Tag* cs = m.find(kDefaultCropSize); cs && cs->num.size() >= 2) {
        in.width = (int)cs->num[0];
        in.height = (int)cs->num[1];
    }
    if (const Tag* co = m.find(kDefaultCropOrigin); co && co->num.size() >= 2) {
        dec.crop_x = (int)co->num[0];
        dec.crop_y = (int)co->num[1];
    }
    if (const Tag* aa = m.find(kActiveArea); aa && aa->num.size() >= 2) {
        dec.crop_x += (int)aa->num[1];
        dec.crop_y += (int)aa->num[0];
    }
    if (in.width <= 0 || in.height <= 0) fail("DNG: empty crop");
    in.orientation = (int)m.get(kOrientation, 1);
    in.black_level = (float)m.get(kBlackLevel, 0.0);
    in.white_level = (float)m.get(kWhiteLevel, 65535.0);
    // CFAPattern -> BayerPattern (cpp/demosaic_utils.cpp:302-305)
    in.bayerPattern = CLIM_GBRG;
    if (const Tag* cfa = m.find(kCFAPattern); cfa && cfa->bytes.size() >= 4) {
        const uint8_t* c = cfa->bytes.data();
        if (!std::memcmp(c, "\0\1\1\2", 4)) in.bayerPattern = CLIM_RGGB;
        else if (!std::memcmp(c, "\2\1\1\0", 4)) in.bayerPattern = CLIM_BGGR;
        else if (!std::memcmp(c, "\1\0\2\1", 4)) in.bayerPattern = CLIM_GRBG;
    }
    auto vec = [&](uint16_t tag, float* dst, int n, int32_t* has) {
        const Tag* t = m.find(tag);
        *has = t && (int)t->num.size() >= n;
        if (*has)
            for (int i = 0; i < n; i++) dst[i] = (float)t->num[i];
    };
    vec(kColorMatrix1, in.color_matrix1, 9, &in.has_color_matrix1);
    vec(kColorMatrix2, in.color_matrix2, 9, &in.has_color_matrix2);
    vec(kAsShotNeutral, in.as_shot_neutral, 3, &in.has_as_shot_neutral);
    in.baseline_exposure = (float)m.get(kBaselineExposure, 0.0);
}

// uncompressed samples: 16-bit in the file's byte order, else MSB-first packed
void unpack_block(const Tiff& t, size_t off, size_t bytes, int bits, uint16_t* out, size_t count) {
    t.need(off, bytes);
    const uint8_t* s = t.data() + off;
    if (bits == 16) {
        if (bytes < 2 * count) fail("DNG: short strip");
        for (size_t i = 0; i < count; i++)
            out[i] = t.little_endian() ? (uint16_t)(s[2 * i] | s[2 * i + 1] << 8) : (uint16_t)(s[2 * i] << 8 | s[2 * i + 1]);
        return;
    }
    if (bits < 1 || bits > 16) fail("DNG: unsupported BitsPerSample");
    if ((uint64_t)count * bits > (uint64_t)bytes * 8) fail("DNG: short strip");
    uint64_t acc = 0;
    int nacc = 0;
    size_t q = 0;
    for (size_t i = 0; i < count; i++) {
        while (nacc < bits) {
            acc = acc << 8 | s[q++];
            nacc += 8;
        }
        out[i] = (uint16_t)((acc >> (nacc - bits)) & ((1u << bits) - 1));
        nacc -= bits;
    }
}

void decode_pixels(const Tiff& t, const RawIfd& r, const Decoded& dec, uint16_t* pixels, int64_t pitch) {
    const Ifd& ifd = r.ifd;
    const int W = dec.full_w, H = dec.full_h;
    const int bits = dec.info.bits_per_sample, comp = dec.info.compression;
    if (comp != 1 && comp != 7) fail("DNG: compression " + std::to_string(comp) + " is not supported");
    const clim_dng_info& in = dec.info;
    // copies the block's rows [y0, y0 + h) x [x0, x0 + w) into the crop
    auto place = [&](const uint16_t* blk, int bw, int x0, int y0, int w, int h) {
        for (int y = 0; y < h; y++) {
            const int oy = y0 + y - dec.crop_y;
            if (oy < 0 || oy >= in.height) continue;
            for (int x = 0; x < w; x++) {
                const int ox = x0 + x - dec.crop_x;
                if (ox >= 0 && ox < in.width) pixels[(size_t)oy * pitch + ox] = blk[(size_t)y * bw + x];
            }
        }
    };
    auto decode_block = [&](size_t off, size_t bytes, int bw, int bh, uint16_t* buf) {
        if (comp == 7) {
            t.need(off, bytes);
            lj92_decode(t.data() + off, bytes, buf, (size_t)bw * bh);
        } else {
            unpack_block(t, off, bytes, bits, buf, (size_t)bw * bh);
        }
    };
    // sizes come from the file: bound them before anything is allocated from them
    if ((size_t)W * (size_t)H > kMaxSamples) fail("DNG: image larger than " + std::to_string(kMaxSamples) + " samples");
    for (int y = 0; y < in.height; y++) std::fill(pixels + (size_t)y * pitch, pixels + (size_t)y * pitch + in.width, 0);
    if (in.tiled) {
        const double twd = ifd.get(kTileWidth, 0), thd = ifd.get(kTileLength, 0);
        // a tile may overhang the image, but by less than one tile per axis
        if (!(twd >= 1 && thd >= 1 && twd <= 2.0 * W + 16 && thd <= 2.0 * H + 16))
            fail("DNG: tile size outside the image's bounds");
        const int tw = (int)twd, th = (int)thd;
        const Tag* to = ifd.find(kTileOffsets);
        const Tag* tc = ifd.find(kTileByteCounts);
        if (!to || !tc || to->num.size() != tc->num.size()) fail("DNG: bad tile layout");
        const int tx = (W + tw - 1) / tw, ty = (H + th - 1) / th;
        if ((int)to->num.size() < tx * ty) fail("DNG: missing tiles");
        const int ntiles = tx * ty;
        const int nthreads = std::max(1, std::min<int>(ntiles, std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
        // a worker's exception must not leave its thread (std::terminate):
        // each keeps its message, and out-of-memory is rethrown as such
        std::vector<std::string> errs(nthreads);
        std::vector<char> oom(nthreads, 0);
        std::vector<std::thread> pool;
        for (int k = 0; k < nthreads; k++) {
            pool.emplace_back([&, k]() {
                try {
                    std::vector<uint16_t> buf((size_t)tw * th);
                    for (int i = k; i < ntiles; i += nthreads) {
                        const int x0 = tw * (i % tx), y0 = th * (i / tx);
                        decode_block((size_t)to->num[i], (size_t)tc->num[i], tw, th, buf.data());
                        place(buf.data(), tw, x0, y0, std::min(tw, W - x0), std::min(th, H - y0));
                    }
                } catch (const DngError& e) {
                    errs[k] = e.msg;
                } catch (const std::bad_alloc&) {
                    oom[k] = 1;
                } catch (const std::exception& e) {
                    errs[k] = std::string("DNG: ") + e.what();
                }
            });
        }
        for (auto& th_ : pool) th_.join();
        for (char o : oom)
            if (o) throw std::bad_alloc();
        for (const auto& e : errs)
            if (!e.empty()) fail(e);
    } else {
        const Tag* so = ifd.find(kStripOffsets);
        const Tag* sc = ifd.find(kStripByteCounts);
        if (!so || !sc || so->num.size() != sc->num.size() || so->num.empty()) fail("DNG: bad strip layout");
        if (comp == 7 && so->num.size() != 1)
            fail("Only one TIFF strip expected for compressed DNG files.");
        const int rps = comp == 7 ? H : (int)std::min<double>(ifd.get(kRowsPerStrip, H), H);
        std::vector<uint16_t> buf((size_t)W * rps);
        for (size_t s = 0; s < so->num.size(); s++) {
            const int y0 = (int)s * rps;
            if (y0 >= H) break;
            const int h = std::min(rps, H - y0);
            decode_block((size_t)so->num[s], (size_t)sc->num[s], W, h, buf.data());
            place(buf.data(), W, 0, y0, W, h);
        }
    }
}

// ---------------------------------------------------------- metadata -> params
// sRGB -> XYZ D65 (cpp/demosaic.hpp:91-95)
const double kXyzRgb[3][3] = {{0.4124564, 0.3575761, 0.1804375},
                              {0.2126729, 0.7151522, 0.0721750},
                              {0.0193339, 0.1191920, 0.9503041}};

bool inverse3(const double m[3][3], double out[3][3]) {
    const double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    const double c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    const double c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    const double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
    if (det == 0.0) return false;
    const double inv[3][3] = {{c00, m[0][2] * m[2][1] - m[0][1] * m[2][2], m[0][1] * m[1][2] - m[0][2] * m[1][1]},
                              {c01, m[0][0] * m[2][2] - m[0][2] * m[2][0], m[0][2] * m[1][0] - m[0][0] * m[1][2]},
                              {c02, m[0][1] * m[2][0] - m[0][0] * m[2][1], m[0][0] * m[1][1] - m[0][1] * m[1][0]}};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[i][j] = inv[i][j] / det;
    return true;
}

// C++ Bayer table of cpp/demosaic.hpp:82-88: (x, y) of R, G, B, G2
const int kCppBayer[4][4][2] = {{{1, 0}, {0, 0}, {0, 1}, {1, 1}},
                                {{0, 1}, {0, 0}, {1, 0}, {1, 1}},
                                {{0, 0}, {1, 0}, {1, 1}, {0, 1}},
                                {{1, 1}, {1, 0}, {0, 0}, {0, 1}}};

struct AwbTile {
    double gain[3];
    long highlights;
};

// autoWhiteBalanceKernel (cpp/demosaic_utils.cpp:864-955) on one tile, in FP64
AwbTile awb_tile(const uint16_t* raw, int64_t pitch, int x0, int y0, int tw, int th, const double c2y[3][3],
                 const double sm[3], double white, double black, int bayer) {
    const int qw = tw / 2, qh = th / 2;
    const auto& o = kCppBayer[bayer];
    auto quad = [&](int qx, int qy, double q[3]) {
        const int x = x0 + 2 * qx, y = y0 + 2 * qy;
        auto px = [&](int c) { return (double)raw[(size_t)(y + o[c][1]) * pitch + x + o[c][0]]; };
        q[0] = px(0);
        q[1] = (px(1) + px(3)) / 2.0;
        q[2] = px(2);
    };
    std::vector<double> yuv((size_t)qw * qh * 3);
    double M[3] = {0, 0, 0};
    long hl = 0;
    for (int qy = 0; qy < qh; qy++)
        for (int qx = 0; qx < qw; qx++) {
            double q[3], rgb[3];
            quad(qx, qy, q);
            bool h = false;
            for (int c = 0; c < 3; c++) {
                rgb[c] = sm[c] * (q[c] - black) / 65535.0;
                if (rgb[c] > 1.0) rgb[c] = 1.0;
                else if (rgb[c] > 0.5) h = true;
            }
            hl += h;
            double* p = &yuv[((size_t)qy * qw + qx) * 3];
            for (int r = 0; r < 3; r++) {
                p[r] = c2y[r][0] * rgb[0] + c2y[r][1] * rgb[1] + c2y[r][2] * rgb[2];
                M[r] += p[r];
            }
        }
    const double n = (double)qw * qh;
    for (double& m : M) m /= n;
    double D[3] = {0, 0, 0};
    for (size_t i = 0; i < (size_t)qw * qh; i++)
        for (int r = 0; r < 3; r++) D[r] += std::fabs(yuv[3 * i + r] - M[r]);
    for (double& d : D) d /= n;
    double hist[128][3] = {};
    long hcount[128] = {};
    long selected = 0;
    double ymax = 0;
    for (int qy = 0; qy < qh; qy++)
        for (int qx = 0; qx < qw; qx++) {
            const double* p = &yuv[((size_t)qy * qw + qx) * 3];
            if (std::fabs(p[1] - (M[1] + std::copysign(D[1], M[1]))) < 1.5 * D[1] &&
                std::fabs(p[2] - (1.5 * M[2] + std::copysign(D[2], M[2]))) < 1.5 * D[2]) {
                double q[3];
                quad(qx, qy, q);
                if (p[0] > ymax) ymax = p[0];
                // negative luma: bin 0 (the reference's size_t cast of a
                // negative value is undefined; ARM64 saturates it to 0)
                const int e = (int)std::clamp(std::round(127.0 * p[0]), 0.0, 127.0);
                for (int c = 0; c < 3; c++) hist[e][c] += (q[c] - black) / white;
                hcount[e]++;
                selected++;
            }
        }
    AwbTile r{{1, 1, 1}, hl};
    const int top = (int)std::clamp(std::round(127.0 * ymax), 0.0, 127.0);
    if (top == 0 || selected == 0) return AwbTile{{1, 1, 1}, 0};  // no highlight count either (:930-932)
    double acc[3] = {0, 0, 0};
    long cnt = 0;
    for (int i = top; i >= 0; i--) {
        for (int c = 0; c < 3; c++) acc[c] += hist[i][c];
        cnt += hcount[i];
        if (cnt > 0.01 * selected) break;
    }
    double g[3];
    for (int c = 0; c < 3; c++) g[c] = ymax / (acc[c] / cnt);
    for (int c = 0; c < 3; c++) r.gain[c] = g[c] / g[1];
    return r;
}

}  // namespace

extern "C" const char* clim_dng_last_error(void) { return g_dng_error.c_str(); }

extern "C" clim_status clim_dng_decode(const uint8_t* data, int64_t size, clim_dng_info* info, uint16_t* pixels,
                                       int64_t pitch_elems) {
    if (!data || size <= 0 || !info) return dng_fail(CLIM_EINVAL, "clim_dng_decode: null argument");
    try {
        const Tiff t(data, (size_t)size);
        const RawIfd r = find_raw_ifd(t);
        Decoded dec;
        fill_info(r, dec);
        *info = dec.info;
        if (pixels) {
            if (pitch_elems < info->width) return dng_fail(CLIM_EINVAL, "clim_dng_decode: pitch below the width");
            decode_pixels(t, r, dec, pixels, pitch_elems);
        }
        return CLIM_OK;
    } catch (const DngError& e) {
        return dng_fail(CLIM_EINVAL, e.msg);
    } catch (const std::bad_alloc&) {
        return dng_fail(CLIM_ENOMEM, "clim_dng_decode: out of memory");
    } catch (const std::exception& e) {
        return dng_fail(CLIM_EINVAL, std::string("clim_dng_decode: ") + e.what());
    }
}

extern "C" clim_status clim_unpack_dng_metadata(const clim_dng_info* info, const uint16_t* raw, int64_t pitch_elems,
                                                int32_t auto_white_balance, clim_demosaic_params* p,
                                                float* exposure_out, float* highlights_out,
                                                float* as_shot_neutral_out) {
    if (!info || !p) return dng_fail(CLIM_EINVAL, "unpackDNGMetadata: null argument");
    if (auto_white_balance && (!raw || pitch_elems < info->width))
        return dng_fail(CLIM_EINVAL, "unpackDNGMetadata: automatic white balance needs the raw image");
    // ColorMatrix2 (usually D65) if present, else ColorMatrix1 (:287-288)
    const float* cm = info->has_color_matrix2 ? info->color_matrix2
                      : info->has_color_matrix1 ? info->color_matrix1 : nullptr;
    if (!cm) return dng_fail(CLIM_EINVAL, "unpackDNGMetadata: no ColorMatrix");
    const double exposure = std::pow(2.0, (double)info->baseline_exposure);
    const double black = info->black_level, white = info->white_level;
    if (!(white > black)) return dng_fail(CLIM_EINVAL, "unpackDNGMetadata: white level not above black");
    p->black_level = (float)black;
    p->white_level = (float)white;
    p->exposure_multiplier = (float)std::min(exposure, 1.0);
    p->bayerPattern = info->bayerPattern;
    // cam_xyz_coeff (:25-52): rgb_cam = inverse(diag(pre_mul) * cam_xyz * xyz_rgb)
    double rgb_cam[3][3], pre_mul[3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            rgb_cam[i][j] = 0;
            for (int k = 0; k < 3; k++) rgb_cam[i][j] += (double)cm[i * 3 + k] * kXyzRgb[k][j];
        }
    for (int i = 0; i < 3; i++) {
        const double cw = rgb_cam[i][0] + rgb_cam[i][1] + rgb_cam[i][2];
        if (!(cw > 0.00001)) return dng_fail(CLIM_ENUMERIC, "cam_xyz_coeff: non-positive camera white");
        pre_mul[i] = 1.0 / cw;
    }
    double scaled[3][3], cam_rgb[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) scaled[i][j] = pre_mul[i] * rgb_cam[i][j];
    if (!inverse3(scaled, cam_rgb)) return dng_fail(CLIM_ENUMERIC, "cam_xyz_coeff: singular matrix");
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) p->rgb_cam[i][j] = (float)cam_rgb[i][j];
    if (info->has_as_shot_neutral && !auto_white_balance)
        for (int c = 0; c < 3; c++) pre_mul[c] = 1.0 / info->as_shot_neutral[c];
    auto set_scale = [&]() {
        const double mn = std::min(pre_mul[0], std::min(pre_mul[1], pre_mul[2]));
        for (int c = 0; c < 4; c++)
            p->scale_mul[c] = (float)(std::max(exposure, 1.0) * (pre_mul[c == 3 ? 1 : c] / mn) * 65535.0 / (white - black));
    };
    set_scale();
    float highlights = 0;
    if (auto_white_balance) {
        // autoWhiteBalance (:967-1026): the mean of 4x4 tiles' gains
        float c2yf[9];
        float rc[9];
        for (int i = 0; i < 9; i++) rc[i] = p->rgb_cam[i / 3][i % 3];
        clim_cam_ycbcr(rc, c2yf);
        double c2y[3][3], sm[3];
        for (int i = 0; i < 9; i++) c2y[i / 3][i % 3] = c2yf[i];
        for (int c = 0; c < 3; c++) sm[c] = (double)p->scale_mul[c] / p->scale_mul[1];
        const int tw = 2 * ((info->width / 4) / 2), th = 2 * ((info->height / 4) / 2);
        if (tw < 2 || th < 2) return dng_fail(CLIM_EINVAL, "autoWhiteBalance: image too small");
        AwbTile tiles[16];
        std::vector<std::thread> pool;
        for (int k = 0; k < 16; k++)
            pool.emplace_back([&, k]() {
                tiles[k] = awb_tile(raw, pitch_elems, (k % 4) * tw, (k / 4) * th, tw, th, c2y, sm, white, black,
                                    info->bayerPattern);
            });
        for (auto& t : pool) t.join();
        double gain[3] = {0, 0, 0};
        long hl = 0;
        for (const auto& t : tiles) {
            for (int c = 0; c < 3; c++) gain[c] += t.gain[c];
            hl += t.highlights;
        }
        for (int c = 0; c < 3; c++) gain[c] /= 16.0;
        const double g1 = gain[1];
        for (int c = 0; c < 3; c++) pre_mul[c] = gain[c] / g1;
        highlights = (float)((double)hl / ((double)info->width * info->height / 4));
        set_scale();
    }
    if (exposure_out) *exposure_out = (float)exposure;
    if (highlights_out) *highlights_out = highlights;
    if (as_shot_neutral_out)
        for (int c = 0; c < 3; c++) as_shot_neutral_out[c] = (float)(1.0 / pre_mul[c]);
    return CLIM_OK;
}
