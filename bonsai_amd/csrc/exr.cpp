// Image::write (src/io/image.rs:47-53) for the film's output: width x height
// RGB f32 pixels, row-major with the top row first, written as a single-part
// scanline OpenEXR file. The reference hands the buffer to the image crate
// (0.24.5, which drives the exr crate 1.6.3); that container is lossless for
// f32 data, so any valid OpenEXR file holding the same FLOAT values is the
// same image. This writer uses the simplest layout the OpenEXR format allows:
//
//   magic 20000630, version 2 with no flags (single-part scanline);
//   header attributes in name order: channels (chlist: B, G, R, each FLOAT,
//   sampling 1 x 1), compression (NO_COMPRESSION), dataWindow and
//   displayWindow ((0, 0) to (width - 1, height - 1)), lineOrder
//   (INCREASING_Y), pixelAspectRatio 1, screenWindowCenter (0, 0),
//   screenWindowWidth 1; a zero byte ends the header;
//   a line offset table of one uint64 file offset per scanline;
//   per scanline: int32 y, int32 byte count (12 * width), then the B, G and R
//   values of the line, each channel contiguous.
//
// Every field is little-endian, as the format requires.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bg_internal.h"

namespace {

struct Bytes {
    std::vector<unsigned char> b;
    void u8(unsigned v) { b.push_back((unsigned char)v); }
    void u32(uint32_t v) {
        for (int k = 0; k < 4; ++k) b.push_back((unsigned char)(v >> (8 * k)));
    }
    void i32(int32_t v) { u32((uint32_t)v); }
    void u64(uint64_t v) {
        for (int k = 0; k < 8; ++k) b.push_back((unsigned char)(v >> (8 * k)));
    }
    void f32(float v) {
        uint32_t u;
        std::memcpy(&u, &v, 4);
        u32(u);
    }
    void str(const char* s) {
        while (*s) b.push_back((unsigned char)*s++);
        b.push_back(0);
    }
    // attribute: name, type name, int32 size, value
    void attr(const char* name, const char* type, const Bytes& value) {
        str(name);
        str(type);
        i32((int32_t)value.b.size());
        b.insert(b.end(), value.b.begin(), value.b.end());
    }
};

Bytes box2i(int32_t x0, int32_t y0, int32_t x1, int32_t y1) {
    Bytes v;
    v.i32(x0);
    v.i32(y0);
    v.i32(x1);
    v.i32(y1);
    return v;
}

}  // namespace

namespace bg {

// The complete file image; shared by bg_exr_write and the film writer.
int exr_encode(uint32_t width, uint32_t height, const float* rgb, std::vector<unsigned char>* out) {
    if (width == 0 || height == 0 || !rgb || width > (uint32_t)INT32_MAX / 12 || height > (uint32_t)INT32_MAX)
        return BG_ERR_INVALID;
    Bytes f;
    f.u32(20000630u);  // magic
    f.u32(2u);         // version 2, single-part scanline, short names
    Bytes ch;
    for (const char* name : {"B", "G", "R"}) {
        ch.str(name);
        ch.i32(2);  // FLOAT
        ch.u8(0);   // pLinear
        ch.u8(0);
        ch.u8(0);
        ch.u8(0);   // reserved
        ch.i32(1);  // xSampling
        ch.i32(1);  // ySampling
    }
    ch.u8(0);
    f.attr("channels", "chlist", ch);
    Bytes comp;
    comp.u8(0);  // NO_COMPRESSION
    f.attr("compression", "compression", comp);
    f.attr("dataWindow", "box2i", box2i(0, 0, (int32_t)width - 1, (int32_t)height - 1));
    f.attr("displayWindow", "box2i", box2i(0, 0, (int32_t)width - 1, (int32_t)height - 1));
    Bytes lo;
    lo.u8(0);  // INCREASING_Y
    f.attr("lineOrder", "lineOrder", lo);
    Bytes par;
    par.f32(1.0f);
    f.attr("pixelAspectRatio", "float", par);
    Bytes swc;
    swc.f32(0.0f);
    swc.f32(0.0f);
    f.attr("screenWindowCenter", "v2f", swc);
    Bytes sww;
    sww.f32(1.0f);
    f.attr("screenWindowWidth", "float", sww);
    f.u8(0);  // end of header

    const uint64_t line_bytes = 12ull * width;
    const uint64_t table_at = f.b.size();
    const uint64_t first = table_at + 8ull * height;
    for (uint32_t y = 0; y < height; ++y) f.u64(first + (uint64_t)y * (8 + line_bytes));
    f.b.reserve(first + (uint64_t)height * (8 + line_bytes));
    for (uint32_t y = 0; y < height; ++y) {
        f.i32((int32_t)y);
        f.i32((int32_t)line_bytes);
        const float* row = rgb + 3ull * width * y;
        for (int c : {2, 1, 0})  // B, G, R
            for (uint32_t x = 0; x < width; ++x) f.f32(row[3ull * x + c]);
    }
    out->swap(f.b);
    return BG_OK;
}

int exr_write_file(const char* path, uint32_t width, uint32_t height, const float* rgb) {
    if (!path) return BG_ERR_INVALID;
    std::vector<unsigned char> bytes;
    const int rc = exr_encode(width, height, rgb, &bytes);
    if (rc) return rc;
    const std::string tmp = std::string(path) + ".part";
    FILE* fp = std::fopen(tmp.c_str(), "wb");
    if (!fp) return BG_ERR_INVALID;
    const bool ok = std::fwrite(bytes.data(), 1, bytes.size(), fp) == bytes.size();
    const bool closed = std::fclose(fp) == 0;
    if (!ok || !closed || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return BG_ERR_INVALID;
    }
    return BG_OK;
}

}  // namespace bg

extern "C" {

int bg_exr_write(const char* path, uint32_t width, uint32_t height, const float* rgb) {
    return bg::exr_write_file(path, width, height, rgb);
}

int bg_exr_encode(uint32_t width, uint32_t height, const float* rgb, unsigned char* out, uint64_t* size) {
    if (!size) return BG_ERR_INVALID;
    std::vector<unsigned char> bytes;
    const int rc = bg::exr_encode(width, height, rgb, &bytes);
    if (rc) return rc;
    const uint64_t cap = *size;
    *size = bytes.size();
    if (!out) return BG_OK;
    if (cap < bytes.size()) return BG_ERR_INVALID;
    std::memcpy(out, bytes.data(), bytes.size());
    return BG_OK;
}

}  // extern "C"
