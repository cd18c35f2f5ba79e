// gr_exr.cpp -- scanline OpenEXR output of a film (ExrWriter,
// src/render/exr.rs:13-200): uncompressed, three FLOAT channels in B, G, R
// order, increasing-Y line order with rows flipped so film row 0 is the bottom
// line, pixel value accum / weight, screenWindowWidth = width.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace gr {
namespace {

struct Buf {
    std::vector<uint8_t> b;
    void i32(int32_t v) { raw(&v, 4); }
    void u32(uint32_t v) { raw(&v, 4); }
    void u64(uint64_t v) { raw(&v, 8); }
    void f32(float v) { raw(&v, 4); }
    void u8(uint8_t v) { b.push_back(v); }
    void str(const char* s) { raw(s, std::strlen(s) + 1); }  // NUL-terminated
    void raw(const void* p, size_t n) {  // little-endian host (x86-64)
        const uint8_t* c = static_cast<const uint8_t*>(p);
        b.insert(b.end(), c, c + n);
    }
};

}  // namespace

// Encodes the whole file into `out`.  rgbw: width*height*4 floats, row 0
// first.
void encode_exr(const float* rgbw, uint32_t width, uint32_t height, std::vector<uint8_t>* out) {
    Buf h;
    h.i32(20000630);  // magic
    h.i32(2);         // version
    h.str("channels");
    h.str("chlist");
    h.i32(2 * 3 + 16 * 3 + 1);
    for (const char* ch : {"B", "G", "R"}) {
        h.str(ch);
        h.i32(2);  // FLOAT
        h.i32(0);  // pLinear + reserved
        h.i32(1);  // xSampling
        h.i32(1);  // ySampling
    }
    h.u8(0);
    h.str("compression");
    h.str("compression");
    h.i32(1);
    h.u8(0);  // NONE
    const int32_t win[4] = {0, 0, (int32_t)width - 1, (int32_t)height - 1};
    for (const char* name : {"dataWindow", "displayWindow"}) {
        h.str(name);
        h.str("box2i");
        h.i32(16);
        for (int32_t v : win) h.i32(v);
    }
    h.str("lineOrder");
    h.str("lineOrder");
    h.i32(1);
    h.u8(0);  // INCREASING_Y
    h.str("pixelAspectRatio");
    h.str("float");
    h.i32(4);
    h.f32(1.0f);
    h.str("screenWindowCenter");
    h.str("v2f");
    h.i32(8);
    h.f32(0.0f);
    h.f32(0.0f);
    h.str("screenWindowWidth");
    h.str("float");
    h.i32(4);
    h.f32((float)width);
    h.u8(0);  // end of header
    const size_t line = 4 + 4 + (size_t)width * 4 * 3;
    const size_t data_offset = h.b.size() + 8 * (size_t)height;
    for (uint32_t y = 0; y < height; ++y) h.u64((uint64_t)(data_offset + y * line));
    for (uint32_t y = 0; y < height; ++y) {
        h.i32((int32_t)y);
        h.u32((uint32_t)(line - 8));
        const float* row = rgbw + 4 * (size_t)(height - y - 1) * width;
        for (int c = 2; c >= 0; --c)  // B, G, R blocks
            for (uint32_t i = 0; i < width; ++i) h.f32(row[4 * i + c] / row[4 * i + 3]);
    }
    out->swap(h.b);
}

bool write_exr(const float* rgbw, uint32_t width, uint32_t height, const char* path) {
    std::vector<uint8_t> bytes;
    encode_exr(rgbw, width, height, &bytes);
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    const bool ok = std::fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
    return std::fclose(f) == 0 && ok;
}

}  // namespace gr
