// GPU unit probes of wave-level building blocks. The VGPR-resident traversal stacks
// (hvvr_amd/csrc/gpu/reg_stack.h): one wave runs a host-given sequence of inserts and
// pops, lane 0 records every popped record, and the final register contents are written
// out. The wave prefix sum and sum (wave_traverse.h): one wave per 64 values. Test
// infrastructure only.

#include <hip/hip_runtime.h>

#include "reg_stack.h"

namespace {

using namespace hvvr;

__device__ void probeInit(RegStack& st) {
    st.t = -INFINITY; // the empty lanes' key
    st.m = 0u;
    st.size = 0;
}
template <int N>
__device__ void probeInit(RegStackN<N>& st) {
#pragma unroll
    for (int k = 0; k < N; k++) {
        st.t[k] = 0.0f;
        st.m[k] = 0u;
    }
    st.size = 0;
}
__device__ void probeStore(const RegStack& st, uint32_t lane, float* finalT, uint32_t* finalS) {
    finalT[lane] = st.t;
    finalS[lane] = st.m;
}
template <int N>
__device__ void probeStore(const RegStackN<N>& st, uint32_t lane, float* finalT, uint32_t* finalS) {
#pragma unroll
    for (int k = 0; k < N; k++) {
        finalT[64 * k + lane] = st.t[k];
        finalS[64 * k + lane] = st.m[k];
    }
}

// ops[i]: 0 = pop (popped[i] = record), 1 = insert (keys[i], record i as the entry's meta).
// finalT/finalS[q] = key/meta at stack position q (0 = bottom).
template <class Stack, uint32_t Cap>
__global__ __launch_bounds__(64) void regStackProbe(const uint32_t* ops, const float* keys, uint32_t n,
                                                   uint32_t* popped, float* finalT, uint32_t* finalS,
                                                   uint32_t* finalSize) {
    const uint32_t lane = threadIdx.x;
    Stack st;
    probeInit(st);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t op = uniform(ops[i]);
        if (op == 0u) {
            if (st.size == 0u)
                continue;
            const uint32_t rec = stackPop(st, lane);
            if (lane == 0)
                popped[i] = rec;
        } else if (st.size < Cap) {
            stackInsert(st, bitsf(uniform(fbits(keys[i]))), i, lane);
        }
    }
    probeStore(st, lane, finalT, finalS);
    if (lane == 0)
        *finalSize = st.size;
}

// the same over the LDS stack (capacity LDS_PROBE_CAP)
enum : uint32_t { LDS_PROBE_CAP = 2048 };
__global__ __launch_bounds__(64) void ldsStackProbe(const uint32_t* ops, const float* keys, uint32_t n,
                                                   uint32_t* popped, float* finalT, uint32_t* finalS,
                                                   uint32_t* finalSize) {
    __shared__ float t[LDS_PROBE_CAP];
    __shared__ uint32_t m[LDS_PROBE_CAP];
    const uint32_t lane = threadIdx.x;
    LdsStack st;
    stackBindLds(st, t, m);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t op = uniform(ops[i]);
        if (op == 0u) {
            if (st.size == 0u)
                continue;
            const uint32_t rec = stackPop(st);
            if (lane == 0)
                popped[i] = rec;
        } else if (st.size < LDS_PROBE_CAP) {
            stackInsert(st, bitsf(uniform(fbits(keys[i]))), i, lane);
        }
    }
    for (uint32_t q = lane; q < st.size; q += 64) {
        finalT[q] = t[q];
        finalS[q] = m[q];
    }
    if (lane == 0)
        *finalSize = st.size;
}

// capacity 64 (RegStack), 256 (RegStack4), 512 (RegStack8), 1024 (RegStack16) or 2048 (LdsStack)
int probeRun(uint32_t capacity, const uint32_t* ops, const float* keys, uint32_t n, uint32_t* popped, float* finalT,
             uint32_t* finalS, uint32_t* finalSize) {
    if (capacity != REG_STACK_CAP && capacity != REG_STACK4_CAP && capacity != REG_STACK8_CAP &&
        capacity != REG_STACK16_CAP && capacity != LDS_PROBE_CAP)
        return -5;
    uint32_t *dOps = nullptr, *dPopped = nullptr, *dS = nullptr, *dSize = nullptr;
    float *dKeys = nullptr, *dT = nullptr;
    int rc = 0;
    if (hipMalloc(&dOps, n * 4) || hipMalloc(&dKeys, n * 4) || hipMalloc(&dPopped, n * 4) ||
        hipMalloc(&dT, capacity * 4) || hipMalloc(&dS, capacity * 4) || hipMalloc(&dSize, 4))
        rc = -1;
    if (!rc && (hipMemcpy(dOps, ops, n * 4, hipMemcpyHostToDevice) ||
                hipMemcpy(dKeys, keys, n * 4, hipMemcpyHostToDevice) || hipMemset(dPopped, 0xff, n * 4)))
        rc = -2;
    if (!rc) {
        if (capacity == REG_STACK_CAP)
            regStackProbe<RegStack, REG_STACK_CAP><<<1, 64>>>(dOps, dKeys, n, dPopped, dT, dS, dSize);
        else if (capacity == REG_STACK4_CAP)
            regStackProbe<RegStack4, REG_STACK4_CAP><<<1, 64>>>(dOps, dKeys, n, dPopped, dT, dS, dSize);
        else if (capacity == REG_STACK8_CAP)
            regStackProbe<RegStack8, REG_STACK8_CAP><<<1, 64>>>(dOps, dKeys, n, dPopped, dT, dS, dSize);
        else if (capacity == REG_STACK16_CAP)
            regStackProbe<RegStack16, REG_STACK16_CAP><<<1, 64>>>(dOps, dKeys, n, dPopped, dT, dS, dSize);
        else
            ldsStackProbe<<<1, 64>>>(dOps, dKeys, n, dPopped, dT, dS, dSize);
        if (hipDeviceSynchronize() != hipSuccess)
            rc = -3;
    }
    if (!rc && (hipMemcpy(popped, dPopped, n * 4, hipMemcpyDeviceToHost) ||
                hipMemcpy(finalT, dT, capacity * 4, hipMemcpyDeviceToHost) ||
                hipMemcpy(finalS, dS, capacity * 4, hipMemcpyDeviceToHost) ||
                hipMemcpy(finalSize, dSize, 4, hipMemcpyDeviceToHost)))
        rc = -4;
    (void)hipFree(dOps);
    (void)hipFree(dKeys);
    (void)hipFree(dPopped);
    (void)hipFree(dT);
    (void)hipFree(dS);
    (void)hipFree(dSize);
    return rc;
}

// one wave per 64 values: inclusive prefix sums within each group of 64, and the group's
// sum in every lane
__global__ __launch_bounds__(64) void waveScanProbe(const uint32_t* in, uint32_t* scan, uint32_t* sum) {
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    const uint32_t v = in[i];
    scan[i] = waveInclusiveScan(v);
    sum[i] = waveSum(v);
}

int scanRun(const uint32_t* in, uint32_t n, uint32_t* scan, uint32_t* sum) {
    if (n == 0 || n % 64u != 0u)
        return -5;
    uint32_t *dIn = nullptr, *dScan = nullptr, *dSum = nullptr;
    int rc = 0;
    if (hipMalloc(&dIn, n * 4) || hipMalloc(&dScan, n * 4) || hipMalloc(&dSum, n * 4))
        rc = -1;
    if (!rc && hipMemcpy(dIn, in, n * 4, hipMemcpyHostToDevice))
        rc = -2;
    if (!rc) {
        waveScanProbe<<<n / 64u, 64>>>(dIn, dScan, dSum);
        if (hipDeviceSynchronize() != hipSuccess)
            rc = -3;
    }
    if (!rc && (hipMemcpy(scan, dScan, n * 4, hipMemcpyDeviceToHost) || hipMemcpy(sum, dSum, n * 4, hipMemcpyDeviceToHost)))
        rc = -4;
    (void)hipFree(dIn);
    (void)hipFree(dScan);
    (void)hipFree(dSum);
    return rc;
}

} // namespace

// n a multiple of 64
extern "C" int wave_scan_probe_run(const uint32_t* in, uint32_t n, uint32_t* scan, uint32_t* sum) {
    return scanRun(in, n, scan, sum);
}

extern "C" int reg_stack_probe_run(uint32_t capacity, const uint32_t* ops, const float* keys, uint32_t n,
                                   uint32_t* popped, float* finalT, uint32_t* finalS, uint32_t* finalSize) {
    return probeRun(capacity, ops, keys, n, popped, finalT, finalS, finalSize);
}
