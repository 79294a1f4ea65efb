// Host-only driver of the DNG reader for the sanitizer build
// (tests/test_dng_sanitize.py): decodes every file named on the command line
// with clim_dng_decode (info, then pixels) and runs clim_unpack_dng_metadata
// with automatic white balance on the result.  Prints one status line per
// file; exits non-zero only through a sanitizer report.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "climage.h"

int main(int argc, char** argv) {
    for (int i = 1; i < argc; i++) {
        std::FILE* f = std::fopen(argv[i], "rb");
        if (!f) {
            std::printf("%s: unreadable\n", argv[i]);
            continue;
        }
        std::vector<uint8_t> d;
        uint8_t buf[1 << 16];
        size_t n;
        while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) d.insert(d.end(), buf, buf + n);
        std::fclose(f);
        clim_dng_info info;
        clim_status s = clim_dng_decode(d.data(), (int64_t)d.size(), &info, nullptr, 0);
        if (s == CLIM_OK && (int64_t)info.width * info.height <= (int64_t)1 << 26) {
            std::vector<uint16_t> px((size_t)info.width * info.height);
            s = clim_dng_decode(d.data(), (int64_t)d.size(), &info, px.data(), info.width);
            if (s == CLIM_OK) {
                clim_demosaic_params p{};
                float e, h, asn[3];
                s = clim_unpack_dng_metadata(&info, px.data(), info.width, 1, &p, &e, &h, asn);
            }
        }
        std::printf("%s: %d %s\n", argv[i], (int)s, s == CLIM_OK ? "" : clim_dng_last_error());
    }
    return 0;
}
