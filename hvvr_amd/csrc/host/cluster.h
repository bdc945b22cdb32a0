#pragma once
// naiveXYCluster (raycaster/sample_clustering.h:17-49): orders points so that every run
// of 8192 forms a compact block and every run of 128 inside it a compact tile. y strata
// of blocks, blocks sorted y-then-x, then rows of 8 tiles sorted x-then-y.
//
// The sorts are stable: the reference's std::sort leaves equal keys in an unspecified
// order, so ties follow the input order here.

#include "hvvr/types.h"

#include <algorithm>
#include <math.h>
#include <vector>

namespace hvvr {

// `pos(p)` returns the vector2 a point is clustered by
template <class T, class PosFn>
void naiveXYCluster(std::vector<T>& v, size_t k, PosFn pos) {
    auto yThenX = [&](const T& a, const T& b) {
        vector2 pa = pos(a), pb = pos(b);
        return (pa.y < pb.y) || (pa.y == pb.y && pa.x < pb.x);
    };
    auto xThenY = [&](const T& a, const T& b) {
        vector2 pa = pos(a), pb = pos(b);
        return (pa.x < pb.x) || (pa.x == pb.x && pa.y < pb.y);
    };
    auto sortRange = [&](size_t begin, size_t end, bool byY) {
        begin = std::min(begin, v.size());
        end = std::min(end, v.size());
        if (begin >= end)
            return;
        if (byY)
            std::stable_sort(v.begin() + begin, v.begin() + end, yThenX);
        else
            std::stable_sort(v.begin() + begin, v.begin() + end, xThenY);
    };
    size_t yStrata = (size_t)sqrt((double)k);
    if (yStrata == 0)
        yStrata = 1;
    size_t xStrata = (k + yStrata - 1) / yStrata;
    size_t blockYStrata = (size_t)sqrt((double)TILES_PER_BLOCK);
    size_t blockXStrata = (TILES_PER_BLOCK + blockYStrata - 1) / blockYStrata;
    sortRange(0, v.size(), true);
    for (size_t i = 0; i < yStrata; i++)
        sortRange(i * BLOCK_SIZE * xStrata, (i + 1) * BLOCK_SIZE * xStrata, false);
    for (size_t i = 0; i < k; i++)
        sortRange(i * BLOCK_SIZE, (i + 1) * BLOCK_SIZE, true);
    for (size_t i = 0; i < v.size(); i += TILE_SIZE * blockXStrata)
        sortRange(i, i + TILE_SIZE * blockXStrata, false);
}

} // namespace hvvr
