// png.cpp -- RGBA8 images to PNG files, the headless stand-in for the reference's swapchain
// presentation (src/core/scanline/scanline_rasterizer.cpp:611-666 draws into the swapchain image).
// The encoder writes one IHDR (8-bit RGBA), the rows (filter 0) in one zlib stream of stored
// (uncompressed) deflate blocks, and IEND; CRC-32 (the PNG / zlib polynomial 0xEDB88320,
// reflected) and Adler-32 as the PNG and zlib specifications define them.
#include "scene.hpp"

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace spr {

namespace {

uint32_t crc32_update(uint32_t crc, const uint8_t* p, size_t n) {
    static uint32_t table[256];
    static bool ready = false;
    if (!ready) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        ready = true;
    }
    crc = ~crc;
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return ~crc;
}

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24)); v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8)); v.push_back((uint8_t)x);
}

void chunk(std::vector<uint8_t>& out, const char type[4], const std::vector<uint8_t>& data) {
    put_be32(out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    put_be32(out, crc32_update(0, out.data() + start, out.size() - start));
}

}  // namespace

std::vector<uint8_t> png_encode_rgba(const uint8_t* rgba, uint32_t w, uint32_t h) {
    if (w == 0 || h == 0 || w > (1u << 20) || h > (1u << 20)) throw std::runtime_error("png: bad image size");
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    std::vector<uint8_t> out(sig, sig + 8);
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, w);
    put_be32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8 bits, RGBA, deflate, adaptive filters, no interlace
    chunk(out, "IHDR", ihdr);
    // the filtered rows: a 0 (none) filter byte before each row
    const size_t row = (size_t)w * 4, raw_n = (row + 1) * h;
    std::vector<uint8_t> z;
    z.reserve(raw_n + raw_n / 65535 * 5 + 16);
    z.push_back(0x78);
    z.push_back(0x01);  // zlib header: deflate, 32K window, no dictionary (CMF*256 + FLG divisible by 31)
    uint32_t a = 1, b = 0;  // Adler-32 of the uncompressed bytes
    size_t done = 0;
    std::vector<uint8_t> block;
    block.reserve(65535);
    auto raw_at = [&](size_t i) -> uint8_t {
        const size_t y = i / (row + 1), x = i % (row + 1);
        return x == 0 ? 0 : rgba[y * row + x - 1];
    };
    while (done < raw_n) {
        const size_t n = raw_n - done < 65535 ? raw_n - done : 65535;
        z.push_back(done + n == raw_n ? 1 : 0);  // BFINAL, BTYPE 00 (stored)
        z.push_back((uint8_t)n); z.push_back((uint8_t)(n >> 8));
        z.push_back((uint8_t)~n); z.push_back((uint8_t)(~n >> 8));
        for (size_t i = 0; i < n; ++i) {
            const uint8_t c = raw_at(done + i);
            z.push_back(c);
            a = (a + c) % 65521u;
            b = (b + a) % 65521u;
        }
        done += n;
    }
    put_be32(z, (b << 16) | a);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    return out;
}

void png_write_rgba(const char* path, const uint8_t* rgba, uint32_t w, uint32_t h) {
    const std::vector<uint8_t> png = png_encode_rgba(rgba, w, h);
    FILE* f = std::fopen(path, "wb");
    if (!f) throw std::runtime_error(std::string("png: cannot open ") + path);
    const size_t n = std::fwrite(png.data(), 1, png.size(), f);
    const int rc = std::fclose(f);
    if (n != png.size() || rc != 0) throw std::runtime_error(std::string("png: write failed: ") + path);
}

}  // namespace spr
