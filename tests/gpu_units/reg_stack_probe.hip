// GPU unit probe for the VGPR-resident traversal stack (hvvr_amd/csrc/gpu/reg_stack.h):
// one wave runs a host-given sequence of inserts and pops, lane 0 records every popped
// record, and the final register contents are written out. Test infrastructure only.

#include <hip/hip_runtime.h>

#include "reg_stack.h"

namespace {

// ops[i]: 0 = pop (popped[i] = record), 1 = insert (keys[i], record i as the entry's meta).
// finalT/finalS[q] = key/meta at stack position q (0 = bottom).
__global__ __launch_bounds__(64) void regStackProbe(const uint32_t* ops, const float* keys, uint32_t n,
                                                   uint32_t* popped, float* finalT, uint32_t* finalS,
                                                   uint32_t* finalSize) {
    using namespace hvvr;
    const uint32_t lane = threadIdx.x;
    RegStack st;
    st.t = 0.0f;
    st.m = 0u;
    st.size = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t op = uniform(ops[i]);
        if (op == 0u) {
            if (st.size == 0u)
                continue;
            const uint32_t rec = stackPop(st);
            if (lane == 0)
                popped[i] = rec;
        } else if (st.size < REG_STACK_CAP) {
            stackInsert(st, bitsf(uniform(fbits(keys[i]))), i, lane);
        }
    }
    finalT[lane] = st.t;
    finalS[lane] = st.m;
    if (lane == 0)
        *finalSize = st.size;
}

} // namespace

extern "C" int reg_stack_probe_run(const uint32_t* ops, const float* keys, uint32_t n, uint32_t* popped,
                                   float* finalT, uint32_t* finalS, uint32_t* finalSize) {
    uint32_t *dOps = nullptr, *dPopped = nullptr, *dS = nullptr, *dSize = nullptr;
    float *dKeys = nullptr, *dT = nullptr;
    int rc = 0;
    if (hipMalloc(&dOps, n * 4) || hipMalloc(&dKeys, n * 4) || hipMalloc(&dPopped, n * 4) ||
        hipMalloc(&dT, 64 * 4) || hipMalloc(&dS, 64 * 4) || hipMalloc(&dSize, 4))
        rc = -1;
    if (!rc && (hipMemcpy(dOps, ops, n * 4, hipMemcpyHostToDevice) ||
                hipMemcpy(dKeys, keys, n * 4, hipMemcpyHostToDevice) ||
                hipMemset(dPopped, 0xff, n * 4)))
        rc = -2;
    if (!rc) {
        regStackProbe<<<1, 64>>>(dOps, dKeys, n, dPopped, dT, dS, dSize);
        if (hipDeviceSynchronize() != hipSuccess)
            rc = -3;
    }
    if (!rc && (hipMemcpy(popped, dPopped, n * 4, hipMemcpyDeviceToHost) ||
                hipMemcpy(finalT, dT, 64 * 4, hipMemcpyDeviceToHost) ||
                hipMemcpy(finalS, dS, 64 * 4, hipMemcpyDeviceToHost) ||
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
