// k_chain.hip -- intersection write (make_intersection_1.comp:217-447) for gfx950.
//
// Inside one monotone piece the reference interleaves x-grid and y-grid crossings, but every
// bisection for the next crossing on one axis starts from the previous crossing of the SAME axis
// (the t_min it starts from is that axis' pending value, :338-359, :402).  So each (piece, axis)
// is an independent chain of dependent solves, and the reference's output order is a merge of the
// two chains (ties to x, :338) followed by its equal-t OR-merge (:361-375).  This file keeps
// every float operation of the reference and only changes who computes what:
//
//  1. k_chain_prep: one thread per curve walks its pieces and writes one descriptor per piece
//     slot and per non-empty chain; chains are counted in length classes
//     (quarter octaves for cubic chains, so that the 64 chains of one wave have similar lengths).
//  2. k_chain_order: chains are laid out costliest class first (longest-processing-time order).
//  3. k_chain_solve: a persistent kernel.  Waves first take the long cubic chains, one chain per
//     wave: the 64 lanes evaluate the 63 midpoints of the next 6 bisection levels at once (lane l
//     is node l of the decision tree and recomputes exactly the midpoints on its path), a ballot
//     resolves the path in scalar registers, and 4 rounds give the reference's 24 iterations.
//     Then waves take the remaining chains 64 at a time, one chain per lane, in descending
//     length.  The first item of every wave is its global wave index (the first waves of a
//     launch sit on distinct SIMDs, so the costliest items do not share one), later items come
//     from an atomic counter; issue priority follows the item's cost.
//  4. k_piece_small: one lane per piece slot.  A piece of at most 32 entries is merged by its
//     lane with the reference's sequential walk, OR-merged and turned into fragment keys,
//     winding changes and bucket counts (gen_fragment: the segment from an entry to the next one
//     of its curve) in one streaming pass; the entries themselves are only stored for export.
//     Longer pieces are listed (lo lists) or queued (the heavy queue).
//  5. k_piece_fragments: one wave per listed piece stages its two chains in LDS, merges them with
//     its start in the reference's order (merge-path when both are non-decreasing, a
//     concatenation when both are non-increasing, the sequential walk otherwise), applies the
//     equal-t OR-merge there and produces the same fragment outputs: the light list beside the
//     heavy solve, the list of pieces that wait on heavy chains after it.  Pieces of more than
//     1024 entries (and short ones whose next piece waits on a heavy chain) go to k_piece_heavy,
//     which does the same with a whole workgroup per piece and more LDS.
#include "device_common.hpp"
#include "kernels.hpp"

namespace spr {

namespace {

constexpr int kBlock = 256;
// cubic chains of class 4 * 9 + 3 and up (896+ crossings) get a whole wave: a lane-solved chain of
// 900-1023 crossings outlasted the longest wave-solved chains
constexpr int kWaveClass = 4 * 9 + 3;
// below that, each launch also wave-solves its costliest cubic classes while they hold at most
// kWaveBudget chains in all and at least 2^5 crossings each: in a scene with few long chains
// (a 64-path scene, a glyph page) those chains alone set the launch's span, and a wave takes
// 0.64 us per crossing against about 1.2 us for a lane
constexpr int kWaveBudget = 512;
constexpr int kWaveMinClass = 4 * 5;
constexpr int kStage = 4096;       // piece values staged in LDS by the merging wave
constexpr int kClasses = kChainClasses;  // 0..63: cubic chains (4 per octave), 64..79: line chains
constexpr int kRounds = 4, kLevels = 6;  // wave solve: 4 rounds of 6 bisection levels
// first heavy cubic class: a quarter-octave class 4*l + sub holds lengths >= 2^l (l = 8: 256)
constexpr int kHeavyClass = 4 * 8;
static_assert(kHeavyCrossings == 1 << 8, "heavy class boundary");
static_assert(kFrag == 2, "chain descriptors keep only the sign of the grid step");
static_assert(kRounds * kLevels == kCubicIters, "wave solve must run the reference's iterations");

inline unsigned grid_for(int64_t n) {
    int64_t g = (n + kBlock - 1) / kBlock;
    return (unsigned)(g < 1 ? 1 : g);
}

// Length class of a chain of n >= 1 crossings: bisected chains (cubics, quadratics and arcs,
// called cubic chains below) by octave and by the two bits after the leading one (quarter
// octaves), line chains by octave.
__device__ inline int chain_class(int n, bool bisected) {
    const int l = 31 - __clz(n);
    if (!bisected) return 64 + (l > 15 ? 15 : l);
    if (l > 15) return 63;
    const int sub = l >= 2 ? (n >> (l - 2)) & 3 : (l == 1 ? (n & 1) << 1 : 0);
    return 4 * l + sub;
}

// Cost octave of a class (a line crossing costs about 2^-6 of a cubic one).
__device__ inline int class_cost_octave(int cls) { return cls < 64 ? cls >> 2 : cls - 64 - 6; }

// One thread per curve walks its monotone pieces once, in order: piece k runs from the end of
// piece k-1 (flagged end parameter and end point) to its own end, and its entries start after
// the 1 + nx + ny entries of each earlier piece.  Each of the curve's 5 piece slots gets a
// descriptor (base = -1: no piece), each non-empty chain a descriptor and every chain slot its
// class byte (0xFF: no chain).
__global__ __launch_bounds__(kBlock) void k_chain_prep(
    const float2* __restrict__ tpos, const uint32_t* __restrict__ cpos,
    const uint32_t* __restrict__ ctype, const float* __restrict__ cweight, const uint32_t* __restrict__ cpath,
    const int32_t* __restrict__ visible, const float4* __restrict__ cuts,
    const uint32_t* __restrict__ ncuts, const int32_t* __restrict__ offsets, uint32_t C, int w,
    int h, ChainDesc* __restrict__ chains, PieceDesc* __restrict__ pieces, uint8_t* __restrict__ chain_cls,
    int32_t* __restrict__ class_count, Budget budget, bool budgeted, const int2* __restrict__ pcanvas) {
    __shared__ int32_t hist[kClasses];
    if (threadIdx.x < kClasses) hist[threadIdx.x] = 0;
    __syncthreads();
    // a frame over its budget (or whose int32 total wrapped) writes nothing past any buffer: every
    // curve takes the culled route (no chain, no piece)
    bool skip = false;
    if (budgeted) {
        const int32_t total = offsets[C];
        skip = total < 0 || (int64_t)total > budget.entries;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            *budget.live = skip ? 0 : total;
            *budget.over = skip ? 1 : 0;
        }
    }
    // grid-stride over curves; the class histogram is flushed once per workgroup
    for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < C; c += gridDim.x * kBlock) {
        const uint32_t path = cpath[c];
        const uint32_t nc = !skip && visible_mask(visible[path]) ? ncuts[c] : 0xFFFFFFFFu;  // all slots empty
        int cw = w, ch = h;
        if (pcanvas && nc != 0xFFFFFFFFu) { const int2 cvs = pcanvas[path]; cw = cvs.x; ch = cvs.y; }
        const uint32_t type = ctype[c];
        CurvePts cv;
        float q[4] = {1.0f, 1.0f, 1.0f, 1.0f};
        int base0 = 0, limit = 0;
        if (nc != 0xFFFFFFFFu) {
            cv = load_curve(tpos, cpos[c], type, cweight, c);
            const float4 qq = cuts[c];
            q[0] = qq.x; q[1] = qq.y; q[2] = qq.z; q[3] = qq.w;
            base0 = offsets[c];
            limit = offsets[c + 1];  // the curve's range ends at the next curve's offset
        }
        float t0 = 0.0f, p0x = 0.0f, p0y = 0.0f;
        if (nc != 0xFFFFFFFFu) { p0x = cv.x[0]; p0y = cv.y[0]; }
        int off = 0;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            PieceDesc pd;
            pd.base = -1; pd.nx = pd.ny = pd.curve = 0; pd.t0 = pd.tend = 0.0f;
            uint32_t k0 = 0xFFu, k1 = 0xFFu;
            if (nc != 0xFFFFFFFFu && (uint32_t)k <= nc) {
                const float traw = (uint32_t)k < nc ? q[k] : 1.0f;
                const float p1x = cv.ex(traw), p1y = cv.ey(traw);
                const float t1 = flag_piece_end(traw, p1x);
                const PieceGrid G = piece_grid(p0x, p0y, p1x, p1y, cw, ch);
                const int base = base0 + off;
                if (base + G.nx + G.ny + 1 <= limit) {  // count and write passes agree
                    pd.base = base; pd.nx = G.nx; pd.ny = G.ny; pd.curve = (int32_t)c; pd.t0 = t0;
                    // the last piece's last segment runs to t = 1; the others run to the next
                    // piece's first entry
                    pd.tend = (uint32_t)k == nc ? 1.0f : -1.0f;
#pragma unroll
                    for (int side = 0; side < 2; ++side) {
                        const int n = side ? G.ny : G.nx;
                        if (n == 0) continue;
                        ChainDesc d;
                        d.n = n;
                        d.pos = (int32_t)cpos[c];
                        d.flags = (int16_t)(side | ((side ? G.dy : G.dx) < 0.0f ? 2 : 0));
                        d.out = base + 1 + (side ? G.nx : 0);
                        d.line0 = side ? G.y0 : G.x0;
                        d.w = type == kArc ? cweight[c] : 0.0f;
                        d.t0 = t0;
                        d.t1 = t1;
                        d.type = (int16_t)type;
                        chains[2 * (5 * (int64_t)c + k) + side] = d;
                        const uint32_t kc = (uint32_t)chain_class(n, type != kLine);
                        atomicAdd(&hist[kc], 1);
                        if (side) k1 = kc; else k0 = kc;
                    }
                }
                off += 1 + G.nx + G.ny;
                t0 = t1; p0x = p1x; p0y = p1y;
            }
            pieces[5 * (int64_t)c + k] = pd;
            reinterpret_cast<uint16_t*>(chain_cls)[5 * (int64_t)c + k] = (uint16_t)(k0 | (k1 << 8));
        }
    }
    __syncthreads();
    if (threadIdx.x < kClasses && hist[threadIdx.x]) atomicAdd(&class_count[threadIdx.x], hist[threadIdx.x]);
}

// Class order: by estimated cost in quarter octaves, costliest first (a line crossing costs
// about 2^-6 of a cubic one), cubic before line on ties.  Cost key: cubic class c -> 2(c+24)+1,
// line class 64+l -> 8l; the keys are distinct.  The cubic classes from kWaveClass up
// come first: the wave-solved chains.
__device__ inline int class_cost_key(int c) { return c < 64 ? 2 * (c + 24) + 1 : 8 * (c - 64); }

// Chains in class order: each workgroup computes the class offsets (a class's offset is the
// number of chains in the classes of higher cost key; the keys are distinct, and the wave-solved
// classes and the heavy ones come first, so their totals are also the ends of those groups;
// workgroup 0 stores them for the solve), owns a contiguous range of chain slots, counts its
// chains per class (from the one-byte classes written by k_chain_prep), takes one block of places
// per class with a single atomic, then writes the chain indices.
__global__ __launch_bounds__(kBlock) void k_chain_order(const uint8_t* __restrict__ chain_cls, int64_t n_slots,
                                                        const int32_t* __restrict__ class_count,
                                                        int32_t* __restrict__ class_fill,
                                                        int32_t* __restrict__ class_off, int32_t* __restrict__ n_long,
                                                        int32_t* __restrict__ light_begin,
                                                        int32_t* __restrict__ n_long_light,
                                                        int32_t* __restrict__ order) {
    __shared__ int32_t cnt[kClasses], base[kClasses], off[kClasses], total[kClasses];
    if (threadIdx.x < kClasses) {
        total[threadIdx.x] = class_count[threadIdx.x];
        cnt[threadIdx.x] = 0;
    }
    __syncthreads();
    if (threadIdx.x < kClasses) {
        const int key = class_cost_key(threadIdx.x);
        int32_t o = 0;
        for (int d = 0; d < kClasses; ++d) o += class_cost_key(d) > key ? total[d] : 0;
        off[threadIdx.x] = o;
        if (blockIdx.x == 0) class_off[threadIdx.x] = o;
    } else if (blockIdx.x == 0 && threadIdx.x == kClasses) {
        int32_t all = 0, nl = 0, lb = 0;
        for (int d = 0; d < kClasses; ++d) {
            all += total[d];
            if (d < 64 && d >= kWaveClass) nl += total[d];
            if (d < 64 && d >= kHeavyClass) lb += total[d];
        }
        // the heavy launch's range holds cubic classes only (every line key is below theirs)
        for (int c = kWaveClass - 1; c >= kHeavyClass && nl + total[c] <= kWaveBudget; --c) nl += total[c];
        // the light launch's range in key order: cubic class c comes right after line class
        // (c + 25) / 4 when (c + 25) % 4 == 0 (key 8l = 2c + 50); wave items stop at a line class
        int32_t nll = 0;
        for (int c = kHeavyClass - 1; c >= kWaveMinClass; --c) {
            if ((c + 25) % 4 == 0 && total[64 + (c + 25) / 4] > 0) break;
            if (nll + total[c] > kWaveBudget) break;
            nll += total[c];
        }
        class_off[kClasses] = all;
        *n_long = nl;
        *light_begin = lb;
        *n_long_light = nll;
    }
    const int64_t per = ((n_slots + gridDim.x - 1) / gridDim.x + kBlock - 1) / kBlock * kBlock;
    const int64_t b0 = (int64_t)blockIdx.x * per, b1 = b0 + per < n_slots ? b0 + per : n_slots;
    for (int64_t i = b0 + threadIdx.x; i < b1; i += kBlock) {
        const uint32_t k = chain_cls[i];
        if (k != 0xFFu) atomicAdd(&cnt[k], 1);
    }
    __syncthreads();
    if (threadIdx.x < kClasses) {
        base[threadIdx.x] = cnt[threadIdx.x] ? atomicAdd(&class_fill[threadIdx.x], cnt[threadIdx.x]) : 0;
        cnt[threadIdx.x] = 0;
    }
    __syncthreads();
    for (int64_t i = b0 + threadIdx.x; i < b1; i += kBlock) {
        const uint32_t k = chain_cls[i];
        if (k != 0xFFu) order[off[k] + base[k] + atomicAdd(&cnt[k], 1)] = (int32_t)i;
    }
}

// The chain's axis of its curve, in evaluation form (an arc's c1 = w * x1 and c3 = w).
__device__ inline CurveAxis load_axis(const float2* __restrict__ tpos, const uint32_t* __restrict__ cpos,
                                      const float* __restrict__ cweight, const ChainDesc& d) {
    const float2* p = tpos + d.pos;
    const uint32_t type = (uint32_t)d.type;
    const bool ys = chain_side(d) != 0;
    CurveAxis a;
    const int npts = (int)(type & 7u);
    a.c[0] = ys ? p[0].y : p[0].x;
    a.c[1] = ys ? p[1].y : p[1].x;
    a.c[2] = npts >= 3 ? (ys ? p[2].y : p[2].x) : 0.0f;
    a.c[3] = npts == 4 ? (ys ? p[3].y : p[3].x) : 0.0f;
    if (type == kArc) {
        a.c[1] = d.w * a.c[1];
        a.c[3] = d.w;
    }
    return a;
}

// Sequential chain (one lane): the reference's solve per crossing.
// The chain's values go to out (= chain_t + d.out) through a 16-float LDS slot of the lane
// (value k at stage[(k % 16) * kBlock]): every 16 crossings, at the same step in all lanes of
// the wave, each lane stores its 16 values back to back, so each cache line of a chain is
// written whole while it is resident instead of one value per solve (a lane's stores are
// otherwise hundreds of instructions apart, and the 64 lanes write 64 different chains).
template <uint32_t T>
__device__ inline void solve_chain_lane_t(const ChainDesc& d, const CurveAxis& a, float* __restrict__ out,
                                          uint8_t* __restrict__ mono, float* __restrict__ stage) {
    float tprev = d.t0, line = d.line0;
    uint32_t rises = 0, falls = 0;  // the chain's values go up / down somewhere
    const uint32_t side = (uint32_t)chain_side(d);
    ChainSolver<T> solver(a, chain_step(d), d.t1);
    for (int k = 0; k < d.n; ++k) {
        // the usual case in every lane: the straight-line form, no divergent branches
        const float ts = __all(solver.common(tprev)) ? solver.next_common(line, tprev, d.t1)
                                                     : solver.next(line, tprev, d.t1);
        const float v = bitsf((fbits(ts) & 0xFFFFFFFCu) | side);
        stage[(k & 15) * kBlock] = v;
        if ((k & 15) == 15) {
            float* dst = out + (k - 15);
#pragma unroll
            for (int j = 0; j < 16; ++j) dst[j] = stage[j * kBlock];
        } else if (k == d.n - 1) {
            float* dst = out + (k & ~15);
            for (int j = 0; j <= (k & 15); ++j) dst[j] = stage[j * kBlock];
        }
        // the y chain is preceded by the piece start in the merge, the x chain is not
        const float prev = k > 0 || side ? tprev : v;
        falls |= !(v >= prev) ? 1u : 0u;  // as the flags of the reference order: NaN breaks both
        rises |= !(v <= prev) ? 1u : 0u;
        tprev = v;
        line += chain_step(d);
    }
    *mono = (falls ? 0 : 1) | (rises ? 0 : 2);
}

__device__ inline void solve_chain_lane(const ChainDesc& d, const CurveAxis& a, float* __restrict__ out,
                                        uint8_t* __restrict__ mono, float* __restrict__ stage) {
    const uint32_t type = (uint32_t)d.type;
    if (type == kCubic) solve_chain_lane_t<kCubic>(d, a, out, mono, stage);
    else if (type == kLine) solve_chain_lane_t<kLine>(d, a, out, mono, stage);
    else if (type == kQuad) solve_chain_lane_t<kQuad>(d, a, out, mono, stage);
    else solve_chain_lane_t<kArc>(d, a, out, mono, stage);
}

// Cubic chain solved by one wave: 4 rounds of a 6-level speculative bisection per crossing.
// Lane l (1..63) is node l of the round's decision tree: its depth is floor(log2 l) and the bits
// of l below the leading one are its path (1 = right).  Each lane halves the round's interval
// along its own path, exactly as the sequential bisection would, and evaluates the midpoint of
// its node; the ballot D of "same sign as at t_min" fixes the path taken through the tree.  A
// depth-5 node (lanes 32..63) lies on that path when D holds, at its five ancestors, the turns
// that lead to it: one masked compare per lane and a second ballot name the node, whose
// interval, midpoint and value are then read from its lane.
struct WaveTree {
    bool on[kLevels - 1], go_right[kLevels - 1];
    unsigned long long anc_mask, anc_turn;
};

__device__ inline WaveTree wave_tree(int lane) {
    WaveTree W;
    const int depth = lane ? 31 - __clz(lane) : 0;
    W.anc_mask = 0ull;
    W.anc_turn = 1ull;  // inner nodes never match
#pragma unroll
    for (int lev = 0; lev < kLevels - 1; ++lev) {
        W.on[lev] = lev < depth;
        W.go_right[lev] = W.on[lev] && ((lane >> (depth - 1 - lev)) & 1);
    }
    if (depth == kLevels - 1) {
        W.anc_turn = 0ull;
#pragma unroll
        for (int lev = 0; lev < kLevels - 1; ++lev) {
            const int anc = lane >> (kLevels - 1 - lev);  // ancestor at depth lev
            W.anc_mask |= 1ull << anc;
            W.anc_turn |= (unsigned long long)((lane >> (kLevels - 2 - lev)) & 1) << anc;
        }
    }
    return W;
}

// The 24 levels of one crossing's bisection on [lo, hi]: tm and vtm of the last level.  A node
// turns right where the sign of v - c is s0 (MODE 0), or, with every value finite, where v < c
// for a start below the line (MODE 1) or v >= c for a start above it (MODE 2).
template <uint32_t T, int MODE>
__device__ __forceinline__ void wave_bisect(const CurveAxis& a, float c, float lo, float hi, uint32_t s0,
                                            const WaveTree& W, float& tm, float& vtm) {
#pragma unroll
    for (int round = 0; round < kRounds; ++round) {
        float llo = lo, lhi = hi;
#pragma unroll
        for (int lev = 0; lev < kLevels - 1; ++lev) {
            const float mid = (llo + lhi) * 0.5f;
            llo = W.go_right[lev] ? mid : llo;
            lhi = (W.on[lev] && !W.go_right[lev]) ? mid : lhi;
        }
        const float mym = (llo + lhi) * 0.5f;
        const float myv = eval_axis(T, mym, a.c[0], a.c[1], a.c[2], a.c[3]);
        // lane 0 is no node: its bit is never read
        const bool right = MODE == 1 ? myv < c : (MODE == 2 ? !(myv < c) : (fbits(myv - c) & 0x80000000u) == s0);
        // the interval this node leaves for the next round
        const float nlo = right ? mym : llo, nhi = right ? lhi : mym;
        const unsigned long long D = __ballot(right);
        const unsigned long long P = __ballot((D & W.anc_mask) == W.anc_turn);
        const int node = __builtin_amdgcn_readfirstlane(__builtin_ctzll(P));
        lo = bitsf((uint32_t)__builtin_amdgcn_readlane((int)fbits(nlo), node));
        hi = bitsf((uint32_t)__builtin_amdgcn_readlane((int)fbits(nhi), node));
        if (round == kRounds - 1) {
            tm = bitsf((uint32_t)__builtin_amdgcn_readlane((int)fbits(mym), node));
            vtm = bitsf((uint32_t)__builtin_amdgcn_readlane((int)fbits(myv), node));
        }
    }
}

// DIRECT (every control value finite and below 2^17 in magnitude): a crossing whose start lies
// on the side the chain comes from takes the single-compare rounds, and a crossing after one
// that did not fall back starts there without evaluating at t_min (see bisect_below).
template <uint32_t T, bool DIRECT>
__device__ inline void solve_chain_wave_t(const ChainDesc& d, const CurveAxis& a, float* __restrict__ out,
                                          uint8_t* __restrict__ mono) {
    const int lane = threadIdx.x & 63;
    const WaveTree W = wave_tree(lane);
    const uint32_t side = (uint32_t)chain_side(d);
    const bool neg = chain_step(d) < 0.0f;
    float tprev = d.t0, line = d.line0;
    bool up = true, down = true;
    float kept = 0.0f;
    bool known = false;  // this crossing starts on the side the chain comes from
    for (int k = 0; k < d.n; ++k) {
        const float c = line;
        const float t_min = tprev;
        const bool fb = DIRECT && fast_bracket(t_min, d.t1);
        bool run = true, from = true;
        uint32_t s0 = neg ? 0u : 0x80000000u;
        if (!(fb && known)) {
            // the same value in every lane: uniform branches below
            const float vt0 = eval_axis(T, t_min, a.c[0], a.c[1], a.c[2], a.c[3]);
            run = __builtin_amdgcn_readfirstlane(vt0 != c ? 1 : 0) != 0;
            from = __builtin_amdgcn_readfirstlane((neg ? vt0 > c : vt0 < c) ? 1 : 0) != 0;
            s0 = fbits(vt0 - c) & 0x80000000u;
        }
        float ts = t_min;
        known = fb;  // a start on the line itself is a crossing with no fallback
        if (run) {
            float tm = t_min, vtm = 0.0f;
            if (DIRECT && from && !neg) wave_bisect<T, DIRECT ? 1 : 0>(a, c, t_min, d.t1, s0, W, tm, vtm);
            else if (DIRECT && from) wave_bisect<T, DIRECT ? 2 : 0>(a, c, t_min, d.t1, s0, W, tm, vtm);
            else wave_bisect<T, 0>(a, c, t_min, d.t1, s0, W, tm, vtm);
            const bool fell = fabsf(vtm - c) > 1.0f;
            ts = fell ? t_min : tm;
            known = fb && from && !fell;
        }
        const float v = bitsf((fbits(ts) & 0xFFFFFFFCu) | side);
        // lane k % 64 keeps value k; 64 values are stored at once (coalesced)
        if (lane == (k & 63)) kept = v;
        if ((k & 63) == 63 || k == d.n - 1) {
            if (lane <= (k & 63)) out[k - (k & 63) + lane] = kept;
        }
        if (k > 0 || side) { up = up && v >= tprev; down = down && v <= tprev; }
        tprev = v;
        line += chain_step(d);
    }
    if (lane == 0) *mono = (up ? 1 : 0) | (down ? 2 : 0);
}

// Wave-solved chains are never lines (line chains take the lane classes).
__device__ inline void solve_chain_wave(const ChainDesc& d, const CurveAxis& a, float* __restrict__ out,
                                        uint8_t* __restrict__ mono) {
    const uint32_t type = (uint32_t)d.type;
    const bool direct = fast_axis(type, a);
    if (type == kCubic) {
        if (direct) solve_chain_wave_t<kCubic, true>(d, a, out, mono);
        else solve_chain_wave_t<kCubic, false>(d, a, out, mono);
    } else if (type == kQuad) {
        if (direct) solve_chain_wave_t<kQuad, true>(d, a, out, mono);
        else solve_chain_wave_t<kQuad, false>(d, a, out, mono);
    } else {
        solve_chain_wave_t<kArc, false>(d, a, out, mono);
    }
}

__global__ __launch_bounds__(kBlock) void k_chain_solve(const float2* __restrict__ tpos,
                                                        const uint32_t* __restrict__ cpos,
                                                        const float* __restrict__ cweight,
                                                        const ChainDesc* __restrict__ chains,
                                                        const int32_t* __restrict__ order,
                                                        const int32_t* __restrict__ class_off,
                                                        const int32_t* __restrict__ n_long,
                                                        const int32_t* __restrict__ light_begin, int heavy,
                                                        int32_t* __restrict__ work,
                                                        float* __restrict__ chain_t,
                                                        uint8_t* __restrict__ mono) {
    __shared__ float s_stage[16 * kBlock];
    float* stage = s_stage + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int32_t total = class_off[kClasses];
    const int32_t lb = *light_begin;
    // the launch's chains [r0, r1) in class order (heavy: [0, lb), light: [lb, total)): wave items
    // on the first nl, lane items on the rest
    const int32_t r0 = heavy ? 0 : lb, r1 = heavy ? lb : total;
    const int32_t nl = *n_long;
    const int32_t lane0 = r0 + nl, lane1 = r1;
    const int32_t items = nl + (lane1 - lane0 + 63) / 64;
    const int32_t n_waves = (int32_t)gridDim.x * (kBlock / 64);
    int32_t item = (int32_t)blockIdx.x * (kBlock / 64) + (int32_t)(threadIdx.x >> 6);
    // the wave's next item is taken one item ahead, so that the atomic's round trip overlaps the
    // loads of the current item instead of following its solve
    int32_t next = 0;
    if (lane == 0 && item < items) next = n_waves + atomicAdd(work, 1);
    while (item < items) {
        if (item < nl) {
            __builtin_amdgcn_s_setprio(3);
            const int32_t ci = order[r0 + item];
            const ChainDesc d = chains[ci];
            const CurveAxis a = load_axis(tpos, cpos, cweight, d);
            solve_chain_wave(d, a, chain_t + d.out, mono + ci);
            __builtin_amdgcn_s_setprio(0);
        } else {
            const int32_t k = lane0 + (item - nl) * 64 + lane;
            ChainDesc d;
            d.n = 0; d.type = (int16_t)kLine;
            int32_t ci = 0;
            if (k < lane1) {
                ci = order[k];
                d = chains[ci];
            }
            // issue priority follows the cost of the wave's chains, read from lane 0's (the item's
            // first chain: the longest run first and must not be starved by the short ones)
            const int n0 = __builtin_amdgcn_readfirstlane(d.n);
            const int b0 = __builtin_amdgcn_readfirstlane(d.type != (int16_t)kLine ? 1 : 0);
            const int oct = class_cost_octave(chain_class(n0, b0 != 0));
            if (oct >= 8) __builtin_amdgcn_s_setprio(3);
            else if (oct >= 6) __builtin_amdgcn_s_setprio(2);
            else if (oct >= 4) __builtin_amdgcn_s_setprio(1);
            if (k < lane1) {
                const CurveAxis a = load_axis(tpos, cpos, cweight, d);
                solve_chain_lane(d, a, chain_t + d.out, mono + ci, stage);
            }
            __builtin_amdgcn_s_setprio(0);
        }
        item = __shfl(next, 0);
        if (lane == 0 && item < items) next = n_waves + atomicAdd(work, 1);
    }
}

// Merge mode of a piece: 0 = both merged sequences non-decreasing (merge-path), 1 = both
// non-increasing (the reference merge takes the sequence with the smaller head to its end, then
// the other one: a concatenation), 2 = anything else (the sequential merge).  Sequences:
// nx>0 && ny>0: A = X, B = [t0, Y]; ny == 0: A = [t0, X]; nx == 0: B = [t0, Y].
__device__ inline uint8_t piece_mode(const PieceDesc& pd, int64_t slot, const uint8_t* __restrict__ mono,
                                     const float* __restrict__ chain_t) {
    uint32_t fa = 3u, fb = 3u;
    if (pd.nx > 0) fa = mono[2 * slot];
    if (pd.ny > 0) fb = mono[2 * slot + 1];
    if (pd.ny == 0 && pd.nx > 0) {
        const float x1 = chain_t[pd.base + 1];
        if (!(x1 >= pd.t0)) fa &= ~1u;
        if (!(x1 <= pd.t0)) fa &= ~2u;
    }
    if ((fa & fb & 1u) != 0) return 0;
    if ((fa & fb & 2u) != 0) return 1;
    return 2;
}

// Orders within one wave: LDS written by some lanes and then read by others.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Merges one piece held in LDS (X = xs[0..nx), Y = xs[nx..nx+ny)) into out[0..n) with nthr
// threads (a wave or a workgroup; lane = the thread's index among them), in the order of
// make_intersection_1.comp:310-359, ties to x.
__device__ inline void wave_merge(const PieceDesc& pd, uint8_t mode, const float* X, uint32_t* out, int lane,
                                  int nthr = 64) {
    const int nx = pd.nx, ny = pd.ny;
    const float* Y = X + nx;
    const bool head_in_a = ny == 0;
    const int na = nx + (head_in_a ? 1 : 0), nb = head_in_a ? 0 : ny + 1;
    const float t0 = pd.t0;
    // loads always from the staged chains (index clamped), the piece start selected after: the
    // accessors must not form a pointer to t0, which would turn every access into a flat load
    auto Aat = [&](int i) -> float {
        if (!head_in_a) return X[i];
        const float x = X[i > 0 ? i - 1 : 0];
        return i == 0 ? t0 : x;
    };
    auto Bat = [&](int j) -> float {
        const float y = Y[j > 0 ? j - 1 : 0];
        return j == 0 ? t0 : y;
    };
    const int n = na + nb;
    if (mode == 2) {
        if (lane == 0) {
            const float kX2 = bitsf(kTwoBits), kY2 = bitsf(kTwoBits | 1u);
            float tx = pd.t0, ty = pd.t0;
            int xi = 0, yi = 0;
            if (nx == 0) tx = kX2;
            else if (ny == 0) ty = kY2;
            else { tx = X[0]; xi = 1; }
            for (int i = 0; i < n; ++i) {
                float tmin;
                if (tx <= ty) { tmin = tx; tx = xi < nx ? X[xi++] : kX2; }
                else { tmin = ty; ty = yi < ny ? Y[yi++] : kY2; }
                out[i] = fbits(tmin);
            }
        }
    } else if (mode == 1) {
        const bool a_first = nb == 0 || (na > 0 && Aat(0) <= Bat(0));
        for (int o = lane; o < n; o += nthr) {
            float v;
            if (a_first) v = o < na ? Aat(o) : Bat(o - na);
            else v = o < nb ? Bat(o) : Aat(o - nb);
            out[o] = fbits(v);
        }
    } else {
        const int per = (n + nthr - 1) / nthr;
        const int o0 = lane * per;
        const int o1 = o0 + per < n ? o0 + per : n;
        if (o0 < n) {
            int lo = o0 > nb ? o0 - nb : 0, hi = o0 < na ? o0 : na;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (Aat(mid) <= Bat(o0 - 1 - mid)) lo = mid + 1; else hi = mid;
            }
            int i = lo, j = o0 - lo;
            for (int o = o0; o < o1; ++o) {
                const bool takeA = j >= nb || (i < na && Aat(i) <= Bat(j));
                const float v = takeA ? Aat(i) : Bat(j);
                i += takeA; j += !takeA;
                out[o] = fbits(v);
            }
        }
    }
}

// The reference's OR of equal t inside a piece (make_intersection_1.comp:361-375): within a run
// [s, e) of equal masked t, entry o ends up as the OR of raw values s..min(o+1, e-1).  Runs never
// leave the piece (its first entry starts a run).  `v` is the merged piece in LDS.
__device__ inline uint32_t or_merged(const uint32_t* v, int n, int o) {
    const uint32_t m = v[o] & 0xFFFFFFFCu;
    uint32_t acc = v[o];
    for (int s = o; s > 0 && (v[s - 1] & 0xFFFFFFFCu) == m; --s) acc |= v[s - 1];
    if (o + 1 < n && (v[o + 1] & 0xFFFFFFFCu) == m) acc |= v[o + 1];
    return acc;
}

constexpr int kWaveStage = 1024;  // pieces up to this many entries are merged by k_piece_fragments

// Masked parameter of the entry that follows the last entry of piece `slot`: t = 1 after the
// curve's last piece, else the first entry of the next piece.  The merges of
// make_intersection_1.comp:310-359 all start with the smaller of X[0] and the piece start when
// the piece has crossings on both axes (ties to x), and with the piece start otherwise; a
// piece that runs backwards in t (unsorted cuts) can therefore start with X[0] < start.
__device__ inline float piece_end_t(const PieceDesc* __restrict__ pieces, int64_t slot, const PieceDesc& pd,
                                    const float* __restrict__ chain_t) {
    if (pd.tend >= 0.0f) return pd.tend;
    const PieceDesc nd = pieces[slot + 1];
    if (nd.base < 0) return 1.0f;  // no next piece is written (counts and writes disagree)
    float first = nd.t0;
    if (nd.nx > 0 && nd.ny > 0) {
        const float x0 = chain_t[nd.base + 1];
        if (x0 <= nd.t0) first = x0;
    }
    return glmax(bitsf(fbits(first) & 0xFFFFFFFCu), 0.0f);
}

// What the merge and the fragments of one piece need besides its chains.  One lane gathers it
// (a chain of dependent loads), the wave reads it with meta_of_lane.
struct PieceMeta {
    PieceDesc pd;
    float t_end;        // parameter where the last entry's segment ends
    uint32_t path, bb;  // path, its first bucket
    int32_t w, h, mode;
    int2 pr;            // the path's row range
    CurvePts cv;
};

__device__ inline PieceMeta gather_meta(const PieceDesc* __restrict__ pieces, int64_t slot,
                                        const uint8_t* __restrict__ mono, const float* __restrict__ chain_t,
                                        const FragArgs& A) {
    PieceMeta m;
    m.pd = pieces[slot];
    m.mode = piece_mode(m.pd, slot, mono, chain_t);
    m.t_end = piece_end_t(pieces, slot, m.pd, chain_t);
    const uint32_t c = (uint32_t)m.pd.curve;
    m.path = A.cpath[c];
    m.w = A.w; m.h = A.h;
    if (A.pcanvas) { const int2 cw = A.pcanvas[m.path]; m.w = cw.x; m.h = cw.y; }
    m.cv = load_curve(A.tpos, A.cpos[c], A.ctype[c], A.cweight, c);
    m.bb = (uint32_t)A.bucket_base[m.path];
    m.pr = A.prow[m.path];
    return m;
}

__device__ inline int32_t rl_i(int32_t v, int k) { return __builtin_amdgcn_readlane(v, k); }
__device__ inline float rl_f(float v, int k) { return bitsf((uint32_t)__builtin_amdgcn_readlane((int)fbits(v), k)); }

// Lane k's meta, wave-uniform (k must be uniform).
__device__ inline PieceMeta meta_of_lane(const PieceMeta& m, int k) {
    PieceMeta r;
    r.pd.base = rl_i(m.pd.base, k); r.pd.nx = rl_i(m.pd.nx, k); r.pd.ny = rl_i(m.pd.ny, k);
    r.pd.curve = rl_i(m.pd.curve, k); r.pd.t0 = rl_f(m.pd.t0, k); r.pd.tend = rl_f(m.pd.tend, k);
    r.t_end = rl_f(m.t_end, k);
    r.path = (uint32_t)rl_i((int32_t)m.path, k); r.bb = (uint32_t)rl_i((int32_t)m.bb, k);
    r.w = rl_i(m.w, k); r.h = rl_i(m.h, k); r.mode = rl_i(m.mode, k);
    r.pr = make_int2(rl_i(m.pr.x, k), rl_i(m.pr.y, k));
#pragma unroll
    for (int i = 0; i < 4; ++i) { r.cv.x[i] = rl_f(m.cv.x[i], k); r.cv.y[i] = rl_f(m.cv.y[i], k); }
    r.cv.type = (uint32_t)rl_i((int32_t)m.cv.type, k);
    return r;
}

// An entry's fragment outputs; returns its global bucket.  Frames with narrow sort words store the
// bucket (in key) and the narrow word (bucket inside the path, x + 2 or 0 in the off-canvas
// bucket, winding change + 1), which is all their scatter needs; the others store the key, the
// winding change and the path.
__device__ inline uint32_t store_fragment(const FragArgs& A, int64_t i, uint32_t yx, int swn, uint32_t path,
                                          uint32_t bb, int2 pr) {
    const uint32_t loc = bucket_local((int32_t)yx, pr, A.err);
    const uint32_t bucket = bb + loc;
    if (A.nword) {
        const uint32_t xl = loc == (uint32_t)pr.y ? 0u : (yx & 0xFFFFu) - kNarrowX0;
        A.key[i] = (int32_t)bucket;
        A.nword[i] = (loc << 17) | (xl << 2) | (uint32_t)(swn + 1);
    } else {
        A.key[i] = (int32_t)yx;
        A.wn[i] = (int8_t)swn;
        A.epath[i] = path;
    }
    return bucket;
}

// gen_fragment for the entries of one piece, fin[0..n) = their final t bits: entry o's segment
// runs to entry o+1, the last one's to m.t_end.  Keys, winding changes and paths are written
// coalesced; bucket counts are aggregated over runs of the wave.  Every lane of the n_waves
// waves calls it; wave w takes the chunks w, w + n_waves, ...
__device__ inline void piece_fragments(const PieceMeta& m, const uint32_t* fin, int n, const FragArgs& A, int lane,
                                       int wave = 0, int n_waves = 1) {
    // lane l evaluates point o0 + l (point o is entry o's parameter, point n is t_end); entry o's
    // segment runs from point o to point o + 1, so a chunk of 64 points serves 63 entries
    for (int o0 = wave * 63; o0 < n; o0 += n_waves * 63) {
        const int o = o0 + lane;
        float t = 0.0f;
        uint32_t f = 0;
        if (o < n) { f = fin[o]; t = glmax(bitsf(f & 0xFFFFFFFCu), 0.0f); }
        else if (o == n) t = m.t_end;
        const float px = m.cv.ex(t), py = m.cv.ey(t);
        const float t1 = __shfl_down(t, 1), qx = __shfl_down(px, 1), qy = __shfl_down(py, 1);
        const bool valid = lane < 63 && o < n;
        uint32_t bucket = 0;
        if (valid) {
            uint32_t yx;
            int swn;
            segment_cell_pts(t < t1, px, py, qx, qy, m.w, m.h, yx, swn);
            const int64_t i = (int64_t)m.pd.base + o;
            bucket = store_fragment(A, i, yx, swn, m.path, m.bb, m.pr);
            if (A.inter) A.inter[i] = make_int2(m.pd.curve, (int32_t)f);
            if (A.chord) A.chord[i] = make_float4(px, py, qx, qy);
        }
        wave_count(A.hist, bucket, valid);
    }
}

constexpr int kLaneMax = 32;  // pieces of at most this many entries are finished by one lane

// Appends v of the lanes with `take` to a queue: one atomic per wave.  Every lane calls it.
__device__ inline void queue_append(int32_t* __restrict__ list, int32_t* __restrict__ count, int32_t v, bool take) {
    const int lane = threadIdx.x & 63;
    const unsigned long long m = __ballot(take);
    if (!m) return;
    const int leader = __ffsll((long long)m) - 1;
    int32_t base = 0;
    if (lane == leader) base = atomicAdd(count, (int32_t)__popcll(m));
    base = __shfl(base, leader);
    if (take) list[base + (int32_t)__popcll(m & ((1ull << lane) - 1ull))] = v;
}

// One lane per piece slot.  Pieces of at most kLaneMax entries: the lane stages its chains in
// LDS (entry i of lane l at i * 64 + l), walks the reference's sequential merge
// (make_intersection_1.comp:310-359, ties to x), OR-merges runs of equal masked t on the fly
// (:361-375: an entry is the OR of its run up to itself and the next entry when that is in the
// run) and emits each entry's fragment once the next entry is known.  Longer pieces of at most
// kWaveStage entries are finished one wave per piece later: listed when their chains are light
// (one lo list of up to 64 slots per group of 64 slots, no atomics), queued when they wait on
// heavy chains.  The rest go to the heavy queue (one workgroup per piece once every chain is
// solved).
__global__ __launch_bounds__(kBlock) void k_piece_small(const PieceDesc* __restrict__ pieces, int64_t n_slots,
                                                        const uint8_t* __restrict__ mono,
                                                        const float* __restrict__ chain_t,
                                                        const PieceLists L, const FragArgs A) {
    __shared__ float s_ch[kBlock / 64][kLaneMax - 1][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float* ch = &s_ch[wave][0][0];
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    const int64_t n_iter = (n_slots + stride - 1) / stride;
    for (int64_t it = 0; it < n_iter; ++it) {  // wave-uniform trip count
        const int64_t slot = it * stride + (int64_t)blockIdx.x * kBlock + threadIdx.x;
        PieceDesc pd;
        pd.base = -1; pd.nx = pd.ny = pd.curve = 0; pd.t0 = pd.tend = 0.0f;
        if (slot < n_slots) pd = pieces[slot];
        const int nx = pd.nx, ny = pd.ny, nv = nx + ny, n = pd.base >= 0 ? nv + 1 : 0;
        // light: the piece's chains and the next piece's first x crossing (its last segment
        // ends there when that piece has crossings on both axes) are all solved by the light
        // launch (chains of fewer than kHeavyCrossings crossings)
        const bool x_light = pd.base < 0 || pd.ny == 0 || pd.nx < kHeavyCrossings;  // as a next piece
        bool next_ok = __shfl_down(x_light ? 1 : 0, 1) != 0;
        if (lane == 63 && pd.base >= 0 && pd.tend < 0.0f && slot + 1 < n_slots) {
            const PieceDesc nd = pieces[slot + 1];
            next_ok = nd.base < 0 || nd.ny == 0 || nd.nx < kHeavyCrossings;
        }
        const bool light = nx < kHeavyCrossings && ny < kHeavyCrossings && (pd.tend >= 0.0f || next_ok);
        const bool mine = n > 0 && n <= kLaneMax && light;
        const bool to_lo = n > kLaneMax && n <= kWaveStage && light;
        const bool to_lo_h = n > kLaneMax && n <= kWaveStage && !light;
        const bool to_heavy = n > 0 && !mine && !to_lo && !to_lo_h;
        {
            const int64_t group = (slot - lane) / 64;  // this wave's 64 slots
            const unsigned long long ml = __ballot(to_lo);
            if (to_lo) L.lo[group * 64 + __popcll(ml & ((1ull << lane) - 1ull))] = (int32_t)slot;
            if (lane == 0 && group * 64 < n_slots) L.lo_cnt[group] = (int32_t)__popcll(ml);
        }
        // the pieces that wait on heavy chains come from a few long curves, in consecutive slots:
        // queued, so that the waves that finish them stride over them
        queue_append(L.lo_h, L.lo_h_cnt, (int32_t)slot, to_lo_h);
        queue_append(L.heavy, L.heavy_count, (int32_t)slot, to_heavy);
        float tend = 1.0f;
        uint32_t path = 0, bb = 0;
        int2 pr = make_int2(0, 0);
        int w = A.w, h = A.h;
        CurvePts cv;
#pragma unroll
        for (int k = 0; k < 4; ++k) cv.x[k] = cv.y[k] = 0.0f;
        cv.type = kLine;
        // chain values: when the wave's pieces lie in one window of the LDS size (consecutive
        // pieces are neighbours), the wave loads the window coalesced and value i of a piece is
        // at (base + 1 - lo) + i; otherwise each lane loads its own at i * 64 + lane
        int lo = mine && nv > 0 ? pd.base + 1 : 0x7FFFFFFF, hi = mine && nv > 0 ? pd.base + 1 + nv : 0;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) { lo = min(lo, __shfl_xor(lo, o)); hi = max(hi, __shfl_xor(hi, o)); }
        const bool coop = hi > lo && hi - lo <= 64 * (kLaneMax - 1);
        const int off0 = coop ? pd.base + 1 - lo : lane, step = coop ? 1 : 64;
        if (coop) {
            for (int i = lane; i < hi - lo; i += 64) ch[i] = chain_t[lo + i];
        } else if (mine) {
            for (int i = 0; i < nv; ++i) ch[i * 64 + lane] = chain_t[pd.base + 1 + i];
        }
        if (mine) {
            tend = piece_end_t(pieces, slot, pd, chain_t);
            const uint32_t c = (uint32_t)pd.curve;
            path = A.cpath[c];
            if (A.pcanvas) { const int2 cw = A.pcanvas[path]; w = cw.x; h = cw.y; }
            cv = load_curve(A.tpos, A.cpos[c], A.ctype[c], A.cweight, c);
            bb = (uint32_t)A.bucket_base[path];
            pr = A.prow[path];
        }
        wave_sync();  // the window was loaded by all lanes
        // the sequential walk: A = X (or [t0, X] without y crossings), B = [t0, Y]
        const float kX2 = bitsf(kTwoBits), kY2 = bitsf(kTwoBits | 1u);
        float tx = pd.t0, ty = pd.t0;
        int xi = 0, yi = 0;
        if (nx == 0) tx = kX2;
        else if (ny == 0) ty = kY2;
        else { tx = ch[off0]; xi = 1; }
        // one step of the walk with value selects only (a branch per side would make the
        // compiler address tx / ty and keep the walk state in scratch memory)
        auto next = [&]() -> uint32_t {
            const bool takex = tx <= ty;
            const float t = takex ? tx : ty;
            const bool more = takex ? xi < nx : yi < ny;
            const int at = takex ? xi : nx + yi;
            const float v = ch[off0 + (more ? at : 0) * step];
            const float repl = more ? v : (takex ? kX2 : kY2);
            tx = takex ? repl : tx;
            ty = takex ? ty : repl;
            xi += takex ? 1 : 0;
            yi += takex ? 0 : 1;
            return fbits(t);
        };
        // m0, m1: the current and next merged raw values; acc: OR of the current run up to m0
        uint32_t m0 = 0, m1 = 0, acc = 0, fin0 = 0;
        float t0 = 0.0f, px = 0.0f, py = 0.0f;  // the current entry's parameter and point
        if (mine) {
            m0 = next();
            m1 = n > 1 ? next() : 0u;
            acc = m0;
            fin0 = acc | ((n > 1 && (m1 & 0xFFFFFFFCu) == (m0 & 0xFFFFFFFCu)) ? m1 : 0u);
            t0 = glmax(bitsf(fin0 & 0xFFFFFFFCu), 0.0f);
            px = cv.ex(t0); py = cv.ey(t0);
        }
        int n_max = mine ? n : 0;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) n_max = max(n_max, __shfl_xor(n_max, o));
        // bucket counts: each lane counts its own run of equal buckets (consecutive entries of a
        // piece mostly share their cell row) and adds it when the bucket changes
        uint32_t run_b = 0;
        int32_t run_n = 0;
        for (int o = 0; o < n_max; ++o) {
            const bool valid = mine && o < n;
            if (valid) {
                // the next entry's final value (needs the entry after it)
                uint32_t fin1 = 0;
                float t1 = tend;
                if (o + 1 < n) {
                    const uint32_t m2 = o + 2 < n ? next() : 0u;
                    const bool same = (m1 & 0xFFFFFFFCu) == (m0 & 0xFFFFFFFCu);
                    const uint32_t acc1 = same ? (acc | m1) : m1;
                    fin1 = acc1 | ((o + 2 < n && (m2 & 0xFFFFFFFCu) == (m1 & 0xFFFFFFFCu)) ? m2 : 0u);
                    t1 = glmax(bitsf(fin1 & 0xFFFFFFFCu), 0.0f);
                    acc = acc1; m0 = m1; m1 = m2;
                }
                const float qx = cv.ex(t1), qy = cv.ey(t1);
                uint32_t yx;
                int swn;
                segment_cell_pts(t0 < t1, px, py, qx, qy, w, h, yx, swn);
                const int64_t i = (int64_t)pd.base + o;
                const uint32_t bucket = store_fragment(A, i, yx, swn, path, bb, pr);
                if (A.inter) A.inter[i] = make_int2(pd.curve, (int32_t)fin0);
                if (A.chord) A.chord[i] = make_float4(px, py, qx, qy);
                if (run_n > 0 && bucket != run_b) {
                    atomicAdd(&A.hist[run_b], run_n);
                    run_n = 0;
                }
                run_b = bucket;
                ++run_n;
                fin0 = fin1;
                t0 = t1; px = qx; py = qy;
            }
        }
        if (run_n > 0) atomicAdd(&A.hist[run_b], run_n);
        __builtin_amdgcn_wave_barrier();
    }
}

// One wave finishes `mine` pieces, slots list[k * stride] (k < mine <= 64): lane k gathers the
// meta of the k-th piece up front, so the wave pays that chain of loads once.  The chains of the
// next piece are loaded into registers while the current one is merged, OR-merged and turned
// into fragments in the wave's LDS (xin, vout: kWaveStage values each).
__device__ inline void wave_pieces(const int32_t* __restrict__ list, int stride, int mine,
                                   const PieceDesc* __restrict__ pieces, const uint8_t* __restrict__ mono,
                                   const float* __restrict__ chain_t, const FragArgs& A, float* xin, uint32_t* vout) {
    constexpr int kPre = kWaveStage / 64;
    uint32_t* fin = reinterpret_cast<uint32_t*>(xin);  // the staged chains are dead after the merge
    const int lane = threadIdx.x & 63;
    PieceMeta mm;
    if (lane < mine) mm = gather_meta(pieces, list[(int64_t)stride * lane], mono, chain_t, A);
    // the chains of piece k: values base + 1 .. base + nv, lane + 64 j held in pre[j]
    float pre[kPre];
    auto fetch = [&](int k) {
        const int base = rl_i(mm.pd.base, k), nv = rl_i(mm.pd.nx, k) + rl_i(mm.pd.ny, k);
#pragma unroll
        for (int j = 0; j < kPre; ++j) {
            const int i = lane + 64 * j;
            pre[j] = i < nv ? chain_t[base + 1 + i] : 0.0f;
        }
    };
    fetch(0);
    for (int k = 0; k < mine; ++k) {
        const PieceMeta m = meta_of_lane(mm, k);
        const int nv = m.pd.nx + m.pd.ny, n = nv + 1;
#pragma unroll
        for (int j = 0; j < kPre; ++j) {
            const int i = lane + 64 * j;
            if (i < nv) xin[i] = pre[j];
        }
        if (k + 1 < mine) fetch(k + 1);
        wave_sync();
        wave_merge(m.pd, (uint8_t)m.mode, xin, vout, lane);
        wave_sync();
        for (int o = lane; o < n; o += 64) fin[o] = or_merged(vout, n, o);
        wave_sync();
        piece_fragments(m, fin, n, A, lane);
        wave_sync();
    }
}

// One wave (a workgroup of its own) per quarter of a lo list (the lo pieces of 64 consecutive
// slots): wave w of list g takes the list's pieces w, w + 4, ...
constexpr int kLoQuarters = 4;
__global__ __launch_bounds__(64) void k_piece_fragments(const PieceDesc* __restrict__ pieces,
                                                        const int32_t* __restrict__ lo_list,
                                                        const int32_t* __restrict__ lo_cnt, int64_t n_groups,
                                                        const uint8_t* __restrict__ mono,
                                                        const float* __restrict__ chain_t, const FragArgs A) {
    __shared__ float xin[kWaveStage];
    __shared__ uint32_t vout[kWaveStage];
    const int64_t g = blockIdx.x / kLoQuarters;
    const int q = (int)(blockIdx.x % kLoQuarters);
    if (g >= n_groups) return;
    const int32_t cnt = lo_cnt[g];
    const int32_t mine = (cnt - q + kLoQuarters - 1) / kLoQuarters;
    if (mine <= 0) return;
    wave_pieces(lo_list + g * 64 + q, kLoQuarters, mine, pieces, mono, chain_t, A, xin, vout);
}

// One wave (a workgroup of its own) per gridDim.x-th of a queue: wave b takes the queue's pieces
// b, b + gridDim.x, ..., 64 at a time.
__global__ __launch_bounds__(64) void k_piece_waves(const PieceDesc* __restrict__ pieces,
                                                    const int32_t* __restrict__ queue,
                                                    const int32_t* __restrict__ queue_count,
                                                    const uint8_t* __restrict__ mono,
                                                    const float* __restrict__ chain_t, const FragArgs A) {
    __shared__ float xin[kWaveStage];
    __shared__ uint32_t vout[kWaveStage];
    const int32_t cnt = *queue_count;
    const int32_t W = (int32_t)gridDim.x;
    for (int32_t q0 = (int32_t)blockIdx.x; q0 < cnt; q0 += 64 * W) {
        const int32_t mine = min(64, (cnt - q0 + W - 1) / W);
        wave_pieces(queue + q0, W, mine, pieces, mono, chain_t, A, xin, vout);
    }
}

// One workgroup per piece of the heavy queue (pieces of more than kWaveStage entries, and short
// pieces whose last segment waits on a heavy chain of the next piece), its four waves together.  Every wave gathers the meta of the
// workgroup's pieces (lane k: piece blockIdx.x + k * gridDim.x of the queue).  Pieces up to
// kStage entries are staged in LDS; longer ones are merged into scratch[base..] in global memory
// and OR-merged into scratch[n_entries + base..].
__global__ __launch_bounds__(kBlock) void k_piece_heavy(const PieceDesc* __restrict__ pieces,
                                                        const int32_t* __restrict__ queue,
                                                        const int32_t* __restrict__ queue_count,
                                                        const uint8_t* __restrict__ mono,
                                                        const float* __restrict__ chain_t,
                                                        uint32_t* __restrict__ scratch, int64_t n_entries,
                                                        const FragArgs A) {
    __shared__ float st_in[kStage];
    __shared__ uint32_t st_out[kStage + 1];
    constexpr int kWaves = kBlock / 64;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int32_t cnt = *queue_count;
    const int32_t G = (int32_t)gridDim.x;
    for (int32_t q0 = (int32_t)blockIdx.x; q0 < cnt; q0 += 64 * G) {
        const int32_t mine = min(64, (cnt - q0 + G - 1) / G);
        PieceMeta mm;
        // the meta is gathered here, after every chain is solved (the queueing kernel may run
        // before the heavy chains are)
        if (lane < mine) mm = gather_meta(pieces, queue[q0 + lane * G], mono, chain_t, A);
        for (int k = 0; k < mine; ++k) {
            const PieceMeta m = meta_of_lane(mm, k);
            const int nv = m.pd.nx + m.pd.ny, n = nv + 1;
            // the two branches keep every pointer in one address space (no flat accesses)
            if (n <= kStage) {
                for (int i = tid; i < nv; i += kBlock) st_in[i] = chain_t[m.pd.base + 1 + i];
                __syncthreads();
                wave_merge(m.pd, (uint8_t)m.mode, st_in, st_out, tid, kBlock);
                __syncthreads();
                uint32_t* fin = reinterpret_cast<uint32_t*>(st_in);
                for (int o = tid; o < n; o += kBlock) fin[o] = or_merged(st_out, n, o);
                __syncthreads();
                piece_fragments(m, fin, n, A, lane, wave, kWaves);
            } else {
                uint32_t* vout = scratch + m.pd.base;
                wave_merge(m.pd, (uint8_t)m.mode, chain_t + m.pd.base + 1, vout, tid, kBlock);
                __threadfence_block();
                __syncthreads();
                uint32_t* fin = scratch + n_entries + m.pd.base;
                for (int o = tid; o < n; o += kBlock) fin[o] = or_merged(vout, n, o);
                __threadfence_block();
                __syncthreads();
                piece_fragments(m, fin, n, A, lane, wave, kWaves);
            }
            __syncthreads();
        }
    }
}

}  // namespace

hipError_t launch_chain_prep(const float2* tpos, const uint32_t* cpos, const uint32_t* ctype,
                             const float* cweight, const uint32_t* cpath, const int32_t* visible, const float4* cuts,
                             const uint32_t* ncuts, const int32_t* offsets, uint32_t C, int w, int h,
                             ChainDesc* chains, PieceDesc* pieces, uint8_t* chain_cls, int32_t* class_count,
                             int n_cu, hipStream_t s, const Budget* budget, const int2* pcanvas) {
    if (C == 0) return hipSuccess;
    unsigned g = grid_for((int64_t)C);
    if (g > (unsigned)n_cu * 8) g = (unsigned)n_cu * 8;
    const Budget b = budget ? *budget : Budget{0, nullptr, nullptr};
    k_chain_prep<<<g, kBlock, 0, s>>>(tpos, cpos, ctype, cweight, cpath, visible, cuts, ncuts, offsets, C, w, h, chains,
                                      pieces, chain_cls, class_count, b, budget != nullptr, pcanvas);
    return hipGetLastError();
}

hipError_t launch_chain_order(const uint8_t* chain_cls, uint32_t C, const int32_t* class_count,
                              int32_t* class_off, int32_t* class_fill, int32_t* n_long, int32_t* light_begin,
                              int32_t* n_long_light, int32_t* order, int n_cu, hipStream_t s) {
    if (C == 0) return hipSuccess;
    unsigned g = grid_for((int64_t)C * 10);
    if (g > (unsigned)n_cu * 4) g = (unsigned)n_cu * 4;
    k_chain_order<<<g, kBlock, 0, s>>>(chain_cls, (int64_t)C * 10, class_count, class_fill, class_off, n_long,
                                       light_begin, n_long_light, order);
    return hipGetLastError();
}

hipError_t launch_chain_solve(const float2* tpos, const uint32_t* cpos, const float* cweight,
                              const ChainDesc* chains,
                              const int32_t* order, const int32_t* class_off, const int32_t* n_long,
                              const int32_t* light_begin, int heavy, int32_t* work, float* chain_t, uint8_t* mono,
                              int n_cu, hipStream_t s) {
    // the heavy launch holds a few hundred long items: one wave per SIMD at most
    const unsigned grid = heavy ? (unsigned)n_cu : (unsigned)(n_cu * 8);
    k_chain_solve<<<grid, kBlock, 0, s>>>(tpos, cpos, cweight, chains, order, class_off, n_long, light_begin, heavy, work,
                                          chain_t, mono);
    return hipGetLastError();
}

hipError_t launch_piece_small(const PieceDesc* pieces, uint32_t C, const uint8_t* mono, const float* chain_t,
                              const PieceLists& L, const FragArgs& A, int n_cu, hipStream_t s) {
    if (C == 0) return hipSuccess;
    const int64_t slots = (int64_t)C * 5;
    int64_t g = (slots + kBlock - 1) / kBlock;
    if (g > (int64_t)n_cu * 20) g = (int64_t)n_cu * 20;
    k_piece_small<<<(unsigned)g, kBlock, 0, s>>>(pieces, slots, mono, chain_t, L, A);
    return hipGetLastError();
}

hipError_t launch_piece_lo(const PieceDesc* pieces, uint32_t C, const uint8_t* mono, const float* chain_t,
                           const int32_t* list, const int32_t* cnt, const FragArgs& A, hipStream_t s) {
    if (C == 0) return hipSuccess;
    const int64_t groups = ((int64_t)C * 5 + 63) / 64;
    k_piece_fragments<<<(unsigned)(groups * kLoQuarters), 64, 0, s>>>(pieces, list, cnt, groups, mono, chain_t, A);
    return hipGetLastError();
}

hipError_t launch_piece_waves(const PieceDesc* pieces, const uint8_t* mono, const float* chain_t,
                              const int32_t* queue, const int32_t* queue_count, const FragArgs& A, int n_cu,
                              hipStream_t s) {
    k_piece_waves<<<(unsigned)(n_cu * 16), 64, 0, s>>>(pieces, queue, queue_count, mono, chain_t, A);
    return hipGetLastError();
}

hipError_t launch_piece_heavy(const PieceDesc* pieces, const PieceLists& L, const uint8_t* mono,
                              const float* chain_t, uint32_t* scratch, int64_t n_entries, const FragArgs& A, int n_cu,
                              hipStream_t s) {
    k_piece_heavy<<<(unsigned)(n_cu * 4), kBlock, 0, s>>>(pieces, L.heavy, L.heavy_count, mono, chain_t, scratch,
                                                          n_entries, A);
    return hipGetLastError();
}

}  // namespace spr
