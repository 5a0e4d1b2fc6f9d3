// paint.hpp -- gradient paints, shared by the host scene model and the composite kernel.  The
// layout is spr_paint's and spr_stop's (include/spr.h).
#pragma once

#include <stdint.h>

namespace spr {

constexpr uint32_t kPaintLinear = 1, kPaintRadial = 2;
constexpr uint32_t kSpreadPad = 0, kSpreadRepeat = 1, kSpreadReflect = 2, kSpreadTransparent = 3;
constexpr uint32_t kSolidPaint = 0xFFFFFFFFu;

struct Paint {
    uint32_t kind, spread, first_stop, n_stops;
    float opacity;
    float xform[6];  // scene -> gradient space: gx = x0*x + x1*y + x2, gy = x3*x + x4*y + x5
    float geom[5];   // linear: x1, y1, x2, y2, unused; radial: cx, cy, fx, fy, r
};
static_assert(sizeof(Paint) == 64, "paint layout");

struct Stop {
    float offset, r, g, b, a;
};
static_assert(sizeof(Stop) == 20, "stop layout");

}  // namespace spr
