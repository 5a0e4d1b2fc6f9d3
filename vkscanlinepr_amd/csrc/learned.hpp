// learned.hpp -- learned single-pass allocation: per-path capacity prediction by a small MLP,
// predicted-allocation intersection write, overflow detection and exact fallback.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct spr_handle;

namespace spr {

struct LearnedState {
    bool ready = false;
};

void learned_on_load(LearnedState& L);
void learned_frame(spr_handle* H, LearnedState& L);
void learned_release(LearnedState& L);

}  // namespace spr
