// Record-mode frame files of the drop-in class without OpenCV: the 8-bit
// density image coloured with viridis and written as an RGB PNG, the file
// the reference's writeStagesTo_png produces through cv::applyColorMap and
// cv::imwrite (src/wrappers.cuh:323-345).  The PNG holds one IDAT zlib
// stream, filter type 0 on every row (PNG specification, RFC 1950, RFC
// 1951): compressed by zlib's compress2 at level 1 (OpenCV's default for
// PNG) when the process can load libz.so.1 -- looked up at run time, so a
// caller links nothing more -- else made of stored (uncompressed) deflate
// blocks, which is also forced by FVM2D_PNG_STORED=1 in the environment.
#pragma once

#include <dlfcn.h>

#include <array>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fvm2d_viridis.hpp"

namespace fvm2d {

// CRC-32 of the PNG chunks (polynomial 0xedb88320, reflected).
inline uint32_t png_crc(const unsigned char *p, size_t n, uint32_t c = 0xffffffffu) {
    static const std::array<uint32_t, 256> table = [] {
        std::array<uint32_t, 256> t{};
        for (uint32_t k = 0; k < 256; k++) {
            uint32_t v = k;
            for (int j = 0; j < 8; j++) v = (v & 1u) ? 0xedb88320u ^ (v >> 1) : v >> 1;
            t[k] = v;
        }
        return t;
    }();
    for (size_t i = 0; i < n; i++) c = table[(c ^ p[i]) & 0xffu] ^ (c >> 8);
    return c;
}

inline void png_put32(std::vector<unsigned char> &b, uint32_t v) {
    b.push_back((unsigned char)(v >> 24));
    b.push_back((unsigned char)(v >> 16));
    b.push_back((unsigned char)(v >> 8));
    b.push_back((unsigned char)v);
}

inline void png_chunk(std::vector<unsigned char> &out, const char *type, const std::vector<unsigned char> &data) {
    png_put32(out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    png_put32(out, png_crc(out.data() + start, out.size() - start) ^ 0xffffffffu);
}

// zlib stream of raw by zlib's compress2 (level 1); false if libz cannot be
// loaded or fails.
inline bool png_zlib_compress(const std::vector<unsigned char> &raw, std::vector<unsigned char> &z) {
    using compress2_fn = int (*)(unsigned char *, unsigned long *, const unsigned char *, unsigned long, int);
    using bound_fn = unsigned long (*)(unsigned long);
    struct Zlib {
        compress2_fn compress2 = nullptr;
        bound_fn bound = nullptr;
        Zlib() {
            const char *e = std::getenv("FVM2D_PNG_STORED");
            if (e && *e && *e != '0') return;
            if (void *h = dlopen("libz.so.1", RTLD_NOW | RTLD_LOCAL)) {
                compress2 = reinterpret_cast<compress2_fn>(dlsym(h, "compress2"));
                bound = reinterpret_cast<bound_fn>(dlsym(h, "compressBound"));
            }
        }
    };
    static const Zlib zl;
    if (!zl.compress2 || !zl.bound) return false;
    unsigned long n = zl.bound((unsigned long)raw.size());
    z.resize(n);
    if (zl.compress2(z.data(), &n, raw.data(), (unsigned long)raw.size(), 1) != 0) return false;
    z.resize(n);
    return true;
}

// An 8-bit RGB image of w x h pixels (rgb: 3 bytes per pixel, rows top to
// bottom) as a PNG file; throws std::runtime_error if the file cannot be
// written.
inline void write_png_rgb(const std::string &path, int w, int h, const unsigned char *rgb) {
    if (w <= 0 || h <= 0) throw std::runtime_error("write_png_rgb: empty image");
    // raw scanlines: filter byte 0, then the row
    const size_t row = 3 * (size_t)w;
    std::vector<unsigned char> raw;
    raw.reserve((row + 1) * (size_t)h);
    for (int j = 0; j < h; j++) {
        raw.push_back(0);
        raw.insert(raw.end(), rgb + j * row, rgb + (j + 1) * row);
    }
    std::vector<unsigned char> z;
    if (!png_zlib_compress(raw, z)) {
        // zlib stream: header 0x78 0x01 (deflate, 32 KiB window), stored
        // blocks of at most 65535 bytes, Adler-32 of the raw data
        z = {0x78, 0x01};
        for (size_t off = 0; off < raw.size(); off += 65535) {
            const size_t len = raw.size() - off < 65535 ? raw.size() - off : 65535;
            z.push_back(off + len >= raw.size() ? 1 : 0);  // BFINAL, BTYPE = 00
            z.push_back((unsigned char)(len & 0xff));
            z.push_back((unsigned char)(len >> 8));
            z.push_back((unsigned char)(~len & 0xff));
            z.push_back((unsigned char)((~len >> 8) & 0xff));
            z.insert(z.end(), raw.begin() + off, raw.begin() + off + len);
        }
        uint32_t a = 1, b = 0;
        for (unsigned char c : raw) {
            a = (a + c) % 65521u;
            b = (b + a) % 65521u;
        }
        png_put32(z, (b << 16) | a);
    }

    std::vector<unsigned char> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<unsigned char> ihdr;
    png_put32(ihdr, (uint32_t)w);
    png_put32(ihdr, (uint32_t)h);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit, RGB, deflate, filter 0, no interlace
    png_chunk(out, "IHDR", ihdr);
    // a chunk holds at most 2^31 - 1 bytes: the zlib stream split over
    // consecutive IDAT chunks of 1 GiB
    const size_t kIdat = size_t(1) << 30;
    for (size_t off = 0; off < z.size(); off += kIdat) {
        const size_t n = z.size() - off < kIdat ? z.size() - off : kIdat;
        png_chunk(out, "IDAT", std::vector<unsigned char>(z.begin() + off, z.begin() + off + n));
    }
    png_chunk(out, "IEND", {});
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(out.data()), (std::streamsize)out.size());
    if (!f) throw std::runtime_error("cannot write " + path);
}

// A record-mode frame: the 8-bit density image (gray, w x h, rows top to
// bottom) through the viridis table, as <path>.
inline void write_frame_png(const std::string &path, int w, int h, const unsigned char *gray) {
    std::vector<unsigned char> rgb(3 * (size_t)w * (size_t)h);
    for (size_t i = 0; i < (size_t)w * (size_t)h; i++) {
        const unsigned char *c = kViridis + 3 * (size_t)gray[i];
        rgb[3 * i] = c[0];
        rgb[3 * i + 1] = c[1];
        rgb[3 * i + 2] = c[2];
    }
    write_png_rgb(path, w, h, rgb.data());
}

}  // namespace fvm2d
