// learned.cpp -- learned single-pass allocation (see learned.hpp).
#include "learned.hpp"

#include <stdexcept>

namespace spr {

void learned_on_load(LearnedState& L) { L.ready = false; }

void learned_frame(spr_handle*, LearnedState&) {
    throw std::runtime_error("learned allocation mode is not available in this build");
}

void learned_release(LearnedState& L) { L.ready = false; }

}  // namespace spr
