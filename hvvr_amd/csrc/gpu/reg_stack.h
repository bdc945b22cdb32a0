#pragma once
// Distance-sorted traversal stack held in VGPRs for one wave64 (tile traversal, the
// replay of traverse_avx.h:412-527). Wave-uniform operations only: every lane of the
// wave must be active when they run.

#include "wave_traverse.h"

namespace hvvr {

// The tile traversal's distance-sorted stack, held in VGPRs: the entry at depth d (0 =
// top, the closest) lives in lane d % 64 of register pair (t, s)[d / 64]; t = distance
// key, s = record index. Capacity 256.
struct RegStack {
    float t0, t1, t2, t3;
    uint32_t s0, s1, s2, s3;
    uint32_t size;
};

__device__ inline uint32_t readLane(uint32_t v, uint32_t l) {
    return uint32_t(__builtin_amdgcn_readlane(int(v), int(l)));
}

// Whole-wave lane shifts (DPP wave_shl:1 / wave_shr:1). Lane i takes lane i+1 (down) or
// i-1 (up); the lane with no source (63 / 0) keeps `fill`. A DPP read from an inactive
// lane yields the old value, so the shift must run with all 64 lanes active. Written as
// volatile asm, it stays where it is written (always wave-uniform code here); as a
// builtin under a select, the compiler may move it into a branch that masks lanes off.
// The s_nop gives the two wait states a DPP read needs after a VALU write of its source.
__device__ inline uint32_t laneShiftDown(uint32_t v, uint32_t fill) {
    uint32_t out = fill;
    asm volatile("s_nop 1\n\tv_mov_b32_dpp %0, %1 wave_shl:1 row_mask:0xf bank_mask:0xf" : "+v"(out) : "v"(v));
    return out;
}
__device__ inline uint32_t laneShiftUp(uint32_t v, uint32_t fill) {
    uint32_t out = fill;
    asm volatile("s_nop 1\n\tv_mov_b32_dpp %0, %1 wave_shr:1 row_mask:0xf bank_mask:0xf" : "+v"(out) : "v"(v));
    return out;
}

// Removes the top entry and returns its record: every entry moves up one depth (lane
// i takes lane i+1; lane 63 takes lane 0 of the next register). Wave-uniform.
__device__ inline uint32_t stackPop(RegStack& st) {
    const uint32_t top = uniform(st.s0);
    // carries are read before the registers they come from are shifted
    const uint32_t c1t = readLane(fbits(st.t1), 0), c1s = readLane(st.s1, 0);
    st.t0 = bitsf(laneShiftDown(fbits(st.t0), c1t));
    st.s0 = laneShiftDown(st.s0, c1s);
    if (st.size > 64u) {
        const uint32_t c2t = readLane(fbits(st.t2), 0), c2s = readLane(st.s2, 0);
        st.t1 = bitsf(laneShiftDown(fbits(st.t1), c2t));
        st.s1 = laneShiftDown(st.s1, c2s);
        if (st.size > 128u) {
            const uint32_t c3t = readLane(fbits(st.t3), 0), c3s = readLane(st.s3, 0);
            st.t2 = bitsf(laneShiftDown(fbits(st.t2), c3t));
            st.s2 = laneShiftDown(st.s2, c3s);
            if (st.size > 192u) {
                st.t3 = bitsf(laneShiftDown(fbits(st.t3), 0u));
                st.s3 = laneShiftDown(st.s3, 0u);
            }
        }
    }
    st.size--;
    return top;
}

// depth of the first entry, from the top, whose key is not smaller than `key` (the
// sentinel below the bottom never is)
__device__ inline uint32_t runDepth(const RegStack& st, float key, uint32_t lane) {
    uint64_t run = ~__ballot(lane < st.size && st.t0 < key);
    uint32_t r = run ? uint32_t(__builtin_ctzll(run)) : 64u;
    if (r == 64u && st.size > 64u) {
        run = ~__ballot(64u + lane < st.size && st.t1 < key);
        r = 64u + (run ? uint32_t(__builtin_ctzll(run)) : 64u);
        if (r == 128u && st.size > 128u) {
            run = ~__ballot(128u + lane < st.size && st.t2 < key);
            r = 128u + (run ? uint32_t(__builtin_ctzll(run)) : 64u);
            if (r == 192u && st.size > 192u) {
                run = ~__ballot(192u + lane < st.size && st.t3 < key);
                r = 192u + (run ? uint32_t(__builtin_ctzll(run)) : 64u);
            }
        }
    }
    return r;
}

// Inserts (key, rec) below the run of closer entries at the top: the entries from that
// depth down move one depth deeper (lane i takes lane i-1; lane 0 takes lane 63 of the
// register above). Wave-uniform; the caller keeps size < 256.
__device__ inline void stackInsert(RegStack& st, float key, uint32_t rec, uint32_t lane) {
    const uint32_t r = runDepth(st, key, lane);
    const uint32_t newSize = st.size + 1u;
    const uint32_t c0t = readLane(fbits(st.t0), 63), c0s = readLane(st.s0, 63);
    const uint32_t c1t = readLane(fbits(st.t1), 63), c1s = readLane(st.s1, 63);
    const uint32_t c2t = readLane(fbits(st.t2), 63), c2s = readLane(st.s2, 63);
    if (r < 64u) {
        const uint32_t d = lane;
        const float ut = bitsf(laneShiftUp(fbits(st.t0), 0u));
        const uint32_t us = laneShiftUp(st.s0, 0u);
        st.t0 = d < r ? st.t0 : (d == r ? key : ut);
        st.s0 = d < r ? st.s0 : (d == r ? rec : us);
    }
    if (newSize > 64u && r < 128u) {
        const uint32_t d = 64u + lane;
        const float ut = bitsf(laneShiftUp(fbits(st.t1), c0t));
        const uint32_t us = laneShiftUp(st.s1, c0s);
        st.t1 = d < r ? st.t1 : (d == r ? key : ut);
        st.s1 = d < r ? st.s1 : (d == r ? rec : us);
    }
    if (newSize > 128u && r < 192u) {
        const uint32_t d = 128u + lane;
        const float ut = bitsf(laneShiftUp(fbits(st.t2), c1t));
        const uint32_t us = laneShiftUp(st.s2, c1s);
        st.t2 = d < r ? st.t2 : (d == r ? key : ut);
        st.s2 = d < r ? st.s2 : (d == r ? rec : us);
    }
    if (newSize > 192u) {
        const uint32_t d = 192u + lane;
        const float ut = bitsf(laneShiftUp(fbits(st.t3), c2t));
        const uint32_t us = laneShiftUp(st.s3, c2s);
        st.t3 = d < r ? st.t3 : (d == r ? key : ut);
        st.s3 = d < r ? st.s3 : (d == r ? rec : us);
    }
    st.size = newSize;
}

} // namespace hvvr
