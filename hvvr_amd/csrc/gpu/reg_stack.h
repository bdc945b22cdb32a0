#pragma once
// Distance-sorted traversal stack held in one VGPR pair of a wave64 (the tile walk of
// traverse_avx.h:412-527, replayed over LDS records).
//
// Layout is bottom-up: stack position q (0 = bottom, size-1 = top, the closest entry)
// lives in lane q of (t, m); t = distance key, m = the entry's packed record meta.
// Capacity 64. A pop reads the top lane; an insert shifts the lanes above the insertion
// point up by one. Every operation is wave-uniform: all 64 lanes must be active.
// The lanes at and above `size` hold t = -inf (stackInit sets them, a pop restores its
// lane), so with finite keys an insert's search needs no mask of the live lanes: the
// entries whose key is not smaller than the new one are the bottom run, counted by one
// ballot and a popcount.

#include "wave_traverse.h"

namespace hvvr {

enum : uint32_t { REG_STACK_CAP = 64 };

struct RegStack {
    float t;
    uint32_t m;
    uint32_t size;
};

__device__ inline uint32_t readLane(uint32_t v, uint32_t l) {
    return uint32_t(__builtin_amdgcn_readlane(int(v), int(l)));
}

// v with lane `l` replaced by the wave-uniform value `x`
__device__ inline uint32_t writeLane(uint32_t v, uint32_t x, uint32_t l, uint32_t lane) {
    return lane == l ? x : v;
}

// lanes 0 .. n-1 (n <= 64), built in scalar registers: a ballot of `lane < n && c` compiles
// to a compare, a select and a second compare; ballot(c) & lanesBelowMask(n) does not
__device__ inline uint64_t lanesBelowMask(uint32_t n) {
    return n >= 64u ? ~0ull : (1ull << n) - 1ull;
}

// Whole-wave lane shift (DPP wave_shr:1): lane i takes lane i-1, lane 0 keeps `fill`. A DPP
// read from an inactive lane yields the old value, so the shift must run with all 64
// lanes active. Written as volatile asm, it stays where it is written (always
// wave-uniform code here); as a builtin under a select, the compiler may move it into a
// branch that masks lanes off. The s_nop gives the two wait states a DPP read needs
// after a VALU write of its source.
__device__ inline uint32_t laneShiftUp(uint32_t v, uint32_t fill) {
    uint32_t out = fill;
    asm volatile("s_nop 1\n\tv_mov_b32_dpp %0, %1 wave_shr:1 row_mask:0xf bank_mask:0xf" : "+v"(out) : "v"(v));
    return out;
}

// removes the top entry and returns its meta (size > 0)
__device__ inline uint32_t stackPop(RegStack& st, uint32_t lane) {
    st.size--;
    st.t = lane == st.size ? -INFINITY : st.t;
    return readLane(st.m, st.size);
}

__device__ inline float stackTopKey(const RegStack& st) {
    return bitsf(readLane(fbits(st.t), st.size - 1u));
}

// Inserts (key, meta) one above the highest entry whose key is not smaller, so the new
// entry lands below the run of closer entries at the top (traverse_avx.h:505-515; the
// sentinel below the bottom is never smaller). The caller keeps size < 64, the keys
// finite and the stack sorted (bottom to top, keys not increasing): the entries whose
// key is not smaller are then positions 0 .. p-1, and the -inf lanes are never counted.
__device__ inline void stackInsert(RegStack& st, float key, uint32_t meta, uint32_t lane) {
    const uint32_t p = uint32_t(__builtin_popcountll(ballot(!(st.t < key))));
    const float ut = bitsf(laneShiftUp(fbits(st.t), 0u));
    const uint32_t um = laneShiftUp(st.m, 0u);
    st.t = lane < p ? st.t : (lane == p ? key : ut);
    st.m = lane < p ? st.m : (lane == p ? meta : um);
    st.size++;
}

// ---------------------------------------------------------------------------------
// N-register variant (capacity 64 N) for tiles with deep stacks: position q lives in lane
// q % 64 of register pair (t, m)[q / 64], bottom-up as above. The arrays are indexed by
// compile-time constants only (unrolled loops), so they stay in VGPRs.
// ---------------------------------------------------------------------------------
template <int N>
struct RegStackN {
    float t[N];
    uint32_t m[N];
    uint32_t size;
};
using RegStack4 = RegStackN<4>;
using RegStack8 = RegStackN<8>;
using RegStack16 = RegStackN<16>;
enum : uint32_t { REG_STACK4_CAP = 256, REG_STACK8_CAP = 512, REG_STACK16_CAP = 1024 };

template <int N>
__device__ inline void stackInit(RegStackN<N>& st, float t, uint32_t m, uint32_t size, uint32_t) {
#pragma unroll
    for (int k = 0; k < N; k++) {
        st.t[k] = k == 0 ? t : 0.0f;
        st.m[k] = k == 0 ? m : 0u;
    }
    st.size = size;
}

// key and meta at stack position q (q < size)
template <int N>
__device__ inline void stackRead(const RegStackN<N>& st, uint32_t q, float& key, uint32_t& meta) {
    const uint32_t ln = q & 63u, r = q >> 6;
    key = 0.0f;
    meta = 0u;
#pragma unroll
    for (int k = 0; k < N; k++) {
        if (uint32_t(k) == r) {
            key = bitsf(readLane(fbits(st.t[k]), ln));
            meta = readLane(st.m[k], ln);
        }
    }
}

template <int N>
__device__ inline uint32_t stackPop(RegStackN<N>& st) {
    float key;
    uint32_t meta;
    stackRead(st, st.size - 1u, key, meta);
    st.size--;
    return meta;
}

template <int N>
__device__ inline uint32_t stackPop(RegStackN<N>& st, uint32_t) {
    return stackPop(st);
}

template <int N>
__device__ inline float stackTopKey(const RegStackN<N>& st) {
    float key;
    uint32_t meta;
    stackRead(st, st.size - 1u, key, meta);
    return key;
}

// one register's part of an insert at position p: entries at positions >= p move up by
// one (lane 0 takes lane 63 of the register below, `carryT`/`carryM`), position p takes
// the new entry
__device__ inline void shiftInsertReg(float& t, uint32_t& m, uint32_t base, uint32_t lane, uint32_t p, float key,
                                      uint32_t meta, uint32_t carryT, uint32_t carryM) {
    const float ut = bitsf(laneShiftUp(fbits(t), carryT));
    const uint32_t um = laneShiftUp(m, carryM);
    const uint32_t q = base + lane;
    t = q < p ? t : (q == p ? key : ut);
    m = q < p ? m : (q == p ? meta : um);
}

// Inserts (key, meta) one above the highest entry whose key is not smaller; the caller
// keeps size < 64 N.
template <int N>
__device__ inline void stackInsert(RegStackN<N>& st, float key, uint32_t meta, uint32_t lane) {
    uint32_t p = 0; // found positions are >= 1
#pragma unroll
    for (int reg = N - 1; reg >= 0; reg--) {
        if (p == 0u && 64u * uint32_t(reg) < st.size) {
            const uint64_t b = ballot(!(st.t[reg] < key)) & lanesBelowMask(st.size - 64u * uint32_t(reg));
            if (b)
                p = 64u * uint32_t(reg) + (64u - uint32_t(__builtin_clzll(b)));
        }
    }
    const uint32_t lo = p >> 6, hi = st.size >> 6; // registers touched (the new top is at position size)
    // highest register first, so every carry is read before its source register changes
#pragma unroll
    for (int reg = N - 1; reg >= 0; reg--) {
        if (lo <= uint32_t(reg) && hi >= uint32_t(reg)) {
            const int below = reg > 0 ? reg - 1 : 0;
            const uint32_t ct = reg > 0 ? readLane(fbits(st.t[below]), 63) : 0u;
            const uint32_t cm = reg > 0 ? readLane(st.m[below], 63) : 0u;
            shiftInsertReg(st.t[reg], st.m[reg], 64u * uint32_t(reg), lane, p, key, meta, ct, cm);
        }
    }
    st.size++;
}

// ---------------------------------------------------------------------------------
// LDS variant for the deepest stacks: position q at t[q] / m[q] of two LDS arrays owned
// by the caller (capacity = their length, checked by the caller), bottom-up as above.
// An insert searches the top 64 positions with one ballot per chunk, going down until a
// chunk holds an entry whose key is not smaller, then the wave moves the entries above
// the insertion point up by one, 64 at a time, highest chunk first.
// ---------------------------------------------------------------------------------
struct LdsStack {
    float* t;
    uint32_t* m;
    uint32_t size;
};

__device__ inline void stackBindLds(LdsStack& st, float* t, uint32_t* m) {
    st.t = t;
    st.m = m;
    st.size = 0;
}

// positions 0..size-1 from lanes 0..size-1 of (t, m); size <= 64
__device__ inline void stackInit(LdsStack& st, float t, uint32_t m, uint32_t size, uint32_t lane) {
    if (lane < size) {
        st.t[lane] = t;
        st.m[lane] = m;
    }
    waveLdsSync();
    st.size = size;
}

__device__ inline uint32_t stackPop(LdsStack& st) {
    st.size--;
    return uniform(st.m[st.size]);
}

__device__ inline float stackTopKey(const LdsStack& st) {
    return bitsf(uniform(fbits(st.t[st.size - 1u])));
}

// entry at position q (q < size)
__device__ inline uint32_t stackAt(const LdsStack& st, uint32_t q) {
    return uniform(st.m[q]);
}

// Inserts (key, meta) one above the highest entry whose key is not smaller (position 0
// when there is none); the caller keeps size below the arrays' length.
__device__ inline void stackInsert(LdsStack& st, float key, uint32_t meta, uint32_t lane) {
    int32_t p = 0;
    for (int32_t base = int32_t(st.size) - 64;; base -= 64) {
        const int32_t q = base + int32_t(lane);
        const bool in = q >= 0;
        const float t = st.t[in ? q : 0];
        const uint64_t b = ballot(in && !(t < key));
        if (b) {
            p = base + 64 - int32_t(__builtin_clzll(b));
            break;
        }
        if (base <= 0)
            break;
    }
    for (int32_t base = int32_t(st.size) - 64; base + 64 > p; base -= 64) {
        const int32_t q = base + int32_t(lane);
        const bool move = q >= p;
        float t = 0.0f;
        uint32_t m = 0u;
        if (move) {
            t = st.t[q];
            m = st.m[q];
        }
        waveLdsSync(); // every lane has read its entry before any lane overwrites one
        if (move) {
            st.t[q + 1] = t;
            st.m[q + 1] = m;
        }
        waveLdsSync();
    }
    if (lane == 0) {
        st.t[p] = key;
        st.m[p] = meta;
    }
    waveLdsSync();
    st.size++;
}

} // namespace hvvr
