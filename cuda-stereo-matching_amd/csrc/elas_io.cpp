// elas_io.cpp -- caller-side image I/O of the ELAS path (host only, no device).
//
// PGM input follows the reference's loadPGM (src/Image.h:129-160): header
// tokens "P5", width, height, maxval (<= 255), each token possibly preceded
// by '#' comment lines, one whitespace byte after maxval, then width*height
// raw bytes.  The reference throws on a bad magic or maxval and reads a
// short file silently; here every malformed or truncated file is an error
// code.  PGM output follows savePGM (src/Image.h:162-169).  The reference
// shows D1 in a window (src/Main.cu:14-34); this library writes the maps
// instead, as float PFM and as the same 8-bit scaling the window uses.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "elas_hip.h"

namespace {

// One header token: skips whitespace, skips '#' lines at a token start
// (pnm_read, src/Image.h:110-127).
bool pnm_token(FILE* f, std::string& tok) {
    int c = fgetc(f);
    for (;;) {
        while (c != EOF && isspace(c)) c = fgetc(f);
        if (c != '#') break;
        while (c != EOF && c != '\n') c = fgetc(f);
    }
    tok.clear();
    while (c != EOF && !isspace(c)) {
        tok.push_back((char)c);
        if (tok.size() > 64) return false;
        c = fgetc(f);
    }
    // the whitespace byte that ended the token is consumed, as the
    // reference's ignore() does: after maxval it is the one byte between the
    // header and the data
    return !tok.empty();
}

bool parse_int(const std::string& s, long* v) {
    char* end = nullptr;
    const long x = strtol(s.c_str(), &end, 10);
    if (!end || *end != '\0') return false;
    *v = x;
    return true;
}

}  // namespace

extern "C" int elas_io_read_pgm(const char* path, uint8_t* dst, int64_t cap, int32_t* width, int32_t* height) {
    if (!path || !width || !height) return ELAS_ERR_INVALID_ARGUMENT;
    FILE* f = fopen(path, "rb");
    if (!f) return ELAS_ERR_IO;
    std::string t;
    long w = 0, h = 0, mx = 0;
    const bool ok = pnm_token(f, t) && t == "P5" && pnm_token(f, t) && parse_int(t, &w) && pnm_token(f, t) &&
                    parse_int(t, &h) && pnm_token(f, t) && parse_int(t, &mx);
    if (!ok || w < 1 || h < 1 || w > 65535 || h > 65535 || mx < 1 || mx > 255) {
        fclose(f);
        return ELAS_ERR_IO;
    }
    *width = (int32_t)w;
    *height = (int32_t)h;
    const int64_t n = (int64_t)w * h;
    if (!dst) {
        fclose(f);
        return ELAS_OK;
    }
    if (cap < n) {
        fclose(f);
        return ELAS_ERR_INVALID_ARGUMENT;
    }
    const size_t got = fread(dst, 1, (size_t)n, f);
    fclose(f);
    return got == (size_t)n ? ELAS_OK : ELAS_ERR_IO;
}

extern "C" int elas_io_write_pgm(const char* path, const uint8_t* src, int32_t width, int32_t height) {
    if (!path || !src || width < 1 || height < 1) return ELAS_ERR_INVALID_ARGUMENT;
    FILE* f = fopen(path, "wb");
    if (!f) return ELAS_ERR_IO;
    const size_t n = (size_t)width * height;
    const bool ok = fprintf(f, "P5\n%d %d\n255\n", width, height) > 0 && fwrite(src, 1, n, f) == n;
    return (fclose(f) == 0 && ok) ? ELAS_OK : ELAS_ERR_IO;
}

// Portable float map: "Pf", width height, scale -1 (little endian), rows
// from the bottom of the image to the top.
extern "C" int elas_io_write_pfm(const char* path, const float* src, int32_t width, int32_t height) {
    if (!path || !src || width < 1 || height < 1) return ELAS_ERR_INVALID_ARGUMENT;
    FILE* f = fopen(path, "wb");
    if (!f) return ELAS_ERR_IO;
    bool ok = fprintf(f, "Pf\n%d %d\n-1\n", width, height) > 0;
    std::vector<unsigned char> row(4 * (size_t)width);
    for (int32_t y = height - 1; ok && y >= 0; y--) {
        for (int32_t x = 0; x < width; x++) {
            uint32_t b;
            memcpy(&b, src + (size_t)y * width + x, 4);
            for (int k = 0; k < 4; k++) row[4 * (size_t)x + k] = (unsigned char)(b >> (8 * k));
        }
        ok = fwrite(row.data(), 1, row.size(), f) == row.size();
    }
    return (fclose(f) == 0 && ok) ? ELAS_OK : ELAS_ERR_IO;
}

// The 8-bit view of a disparity map that the reference displays
// (src/Main.cu:14-27): out = (uint8)max(255.0 * d / d_max, 0.0) with d_max
// the largest value in the map (at least 0).  A map without a positive value
// (where the reference divides by zero) gives all zeros.
extern "C" int elas_io_disparity_to_u8(const float* d, uint8_t* out, int64_t n) {
    if ((!d || !out) && n > 0) return ELAS_ERR_INVALID_ARGUMENT;
    float dmax = 0.0f;
    for (int64_t i = 0; i < n; i++)
        if (d[i] > dmax) dmax = d[i];
    for (int64_t i = 0; i < n; i++) {
        if (!(dmax > 0.0f)) {
            out[i] = 0;
            continue;
        }
        const double v = 255.0 * d[i] / dmax;
        out[i] = (uint8_t)(v > 0.0 ? v : 0.0);
    }
    return ELAS_OK;
}
