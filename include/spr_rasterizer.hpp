// spr_rasterizer.hpp -- C++ face of libspr.so with the reference's rasterizer method set.
//
// VkScanlinePR's plugin interface is Galaxysailing::VGRasterizer (src/core/rasterizer.h:9-20):
// initialize(window, w, h), render(), loadVG(shared_ptr<VGContainer>), setMVP(const glm::mat4&).
// spr::Rasterizer offers the same four operations on plain arrays, over the C ABI of spr.h, and
// turns every failure status into std::runtime_error as the reference does
// (src/core/scanline/scanline_rasterizer.cpp:219,224; src/core/vg/rvg.cpp:14).  Header only; link
// with -lspr.
#ifndef SPR_RASTERIZER_HPP_
#define SPR_RASTERIZER_HPP_

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "spr.h"

namespace spr {

class Rasterizer {
public:
    Rasterizer() = default;
    Rasterizer(const Rasterizer&) = delete;
    Rasterizer& operator=(const Rasterizer&) = delete;
    ~Rasterizer() { if (h_) spr_destroy(h_); }

    // VGRasterizer::initialize(void* window, uint32_t w, uint32_t h): there is no window; the
    // target is a w x h RGBA8 image in device memory.
    void initialize(uint32_t w, uint32_t h, uint32_t compose = SPR_COMPOSE_OVERWRITE,
                    uint32_t mode = SPR_MODE_PLAIN, int32_t device = -1) {
        spr_opts o{};
        o.compose = compose;
        o.mode = mode;
        o.device = device;
        if (h_) { spr_destroy(h_); h_ = nullptr; }
        if (spr_create(w, h, &o, &h_) != SPR_OK) throw std::runtime_error(spr_global_error());
        w_ = w;
        hgt_ = h;
    }

    // VGRasterizer::loadVG: the flattened SoA that ScanlineVGRasterizer::loadVG builds
    // (scanline_rasterizer.cpp:67-134), fill_info already packed to RGBA8.
    void loadPaths(const float* pos_xy, const uint32_t* pos_path_idx, uint32_t n_points,
                   const uint32_t* curve_pos_map, const uint32_t* curve_type, const uint32_t* curve_path_idx,
                   uint32_t n_curves, const uint32_t* fill_rule, const uint32_t* fill_info, uint32_t n_paths) {
        check(spr_load_paths(h(), pos_xy, pos_path_idx, n_points, curve_pos_map, curve_type, curve_path_idx,
                             n_curves, fill_rule, fill_info, n_paths));
    }
    void loadRvg(const std::string& path) { check(spr_load_rvg(h(), path.c_str())); }

    // VGRasterizer::setMVP(const glm::mat4& m): the four columns of glm's column-major matrix,
    // which the reference stores as TransPosIn rows m0..m3 (scanline_rasterizer.cpp:191-197);
    // `&m[0][0]` of a glm::mat4 is exactly this array.
    void setMVP(const float m_rows[16]) { check(spr_set_mvp(h(), m_rows)); }

    // VGRasterizer::render(): one frame into the device image; optionally read back.
    void render(uint8_t* rgba_out = nullptr) { check(spr_render_frame(h(), rgba_out)); }
    std::vector<uint8_t> renderToHost() {
        std::vector<uint8_t> img((size_t)4 * w_ * hgt_);
        render(img.data());
        return img;
    }

    spr_stats stats() const {
        spr_stats s{};
        check(spr_get_stats(h_, &s));
        return s;
    }
    void* imageDevicePtr() {
        void* p = nullptr;
        check(spr_image_device_ptr(h(), &p));
        return p;
    }
    spr_handle* handle() const { return h_; }

private:
    spr_handle* h() const {
        if (!h_) throw std::runtime_error("spr::Rasterizer used before initialize()");
        return h_;
    }
    void check(int rc) const {
        if (rc != SPR_OK) throw std::runtime_error(h_ ? spr_last_error(h_) : spr_global_error());
    }
    spr_handle* h_ = nullptr;
    uint32_t w_ = 0, hgt_ = 0;
};

}  // namespace spr

#endif  // SPR_RASTERIZER_HPP_
