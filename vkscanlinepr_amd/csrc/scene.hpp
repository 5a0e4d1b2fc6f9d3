// scene.hpp -- vector scenes on the host: the VGContainer model, its SoA flattening, the RVG
// text format, and the synthetic workload generators.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "paint.hpp"

namespace spr {

// The flattened SoA that VkScanlinePR's loadVG uploads (src/core/scanline/vk_vg_data.h:14-35,
// built at src/core/scanline/scanline_rasterizer.cpp:67-134).  This is the device input format.
// Beyond the reference: quadratic and arc curves (the CurveType values it declares but leaves
// TODO in its shaders), arc weights, and gradient paints.
struct Scene {
    std::vector<float> pos;            // 2 floats per point
    std::vector<uint32_t> pos_path;    // path of each point
    std::vector<uint32_t> curve_pos;   // first point of each curve
    std::vector<uint32_t> curve_type;  // 2 line, 3 quadratic, 4 cubic, 0x13 arc (points: type & 7)
    std::vector<uint32_t> curve_path;  // path of each curve, non-decreasing
    std::vector<uint32_t> rule;        // 0 nonzero, 1 even-odd
    std::vector<uint32_t> fill;        // RGBA8, R in the low byte
    std::vector<float> curve_weight;   // per curve (an arc's weight w > 0), or empty without arcs
    std::vector<uint32_t> path_paint;  // per path: kSolidPaint or a paint index, or empty
    std::vector<Paint> paints;
    std::vector<Stop> stops;
    uint32_t width = 0, height = 0;

    uint32_t n_points() const { return (uint32_t)pos_path.size(); }
    uint32_t n_curves() const { return (uint32_t)curve_type.size(); }
    uint32_t n_paths() const { return (uint32_t)rule.size(); }
};

// Throws std::runtime_error with a message on malformed or unsupported input.
Scene load_rvg(const std::string& path);
void write_rvg(const Scene& s, const std::string& path);
Scene generate(int kind, uint64_t seed, uint32_t n_paths, uint32_t canvas);
Scene subset(const Scene& s, uint32_t first, uint32_t count);
// Checks the invariants the pipeline relies on; returns an empty string when valid.
std::string validate(const float* pos, const uint32_t* pos_path, uint32_t P, const uint32_t* cpos,
                     const uint32_t* ctype, const uint32_t* cpath, uint32_t C,
                     const uint32_t* rule, uint32_t Np);
// The same for the arc weights and paints (all arrays may be null when unused).
std::string validate_extras(const uint32_t* ctype, uint32_t C, const float* cweight, uint32_t Np,
                            const uint32_t* path_paint, const Paint* paints, uint32_t n_paints,
                            const Stop* stops, uint32_t n_stops);

// Packs a float colour the way loadVG does (scanline_rasterizer.cpp:107-118): alpha times
// opacity, each channel times 255 truncated to a byte, and 0 when the alpha byte is 0.
uint32_t pack_fill(float r, float g, float b, float a, float opacity);

// PNG (8-bit RGBA, rows top first) of an image in host memory (png.cpp).
std::vector<uint8_t> png_encode_rgba(const uint8_t* rgba, uint32_t w, uint32_t h);
void png_write_rgba(const char* path, const uint8_t* rgba, uint32_t w, uint32_t h);

}  // namespace spr
