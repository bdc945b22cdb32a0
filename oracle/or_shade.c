/* Oracle: resolve (float and double), remap, lighting environment (test infrastructure
 * only). See or_shade_impl.h for the shading body and its citations. */
#include "or_dof.h"
#include <stdlib.h>

#define R float
#define SFX _f
#include "or_shade_impl.h"
#undef R
#undef SFX
#define R double
#define SFX _d
#include "or_shade_impl.h"
#undef R
#undef SFX

void or_layout_sizes(uint32_t out[10]) {
    out[0] = sizeof(o_vertex);
    out[1] = sizeof(o_tri);
    out[2] = sizeof(o_trishade);
    out[3] = sizeof(o_node);
    out[4] = sizeof(o_material);
    out[5] = sizeof(o_sfrustum);
    out[6] = sizeof(o_beam);
    out[7] = sizeof(o_frustum);
    out[8] = sizeof(o_stack);
    out[9] = sizeof(o_lightenv);
}

/* GPUSceneState::updateLighting (raycaster/gpu_scene_state.cu:47-81): lights past the
 * per-type maxima are dropped */
int or_lighting_env(const void* lights, uint32_t count, o_lightenv* env) {
    memset(env, 0, sizeof(*env));
    int dropped = 0;
    const unsigned char* p = (const unsigned char*)lights;
    for (uint32_t i = 0; i < count; i++) {
        uint32_t type;
        memcpy(&type, p + 56 * i, 4);
        const unsigned char* body = p + 56 * i + 4;
        if (type == 1) {
            if (env->dirCount < 1) memcpy(&env->dir[env->dirCount++], body, sizeof(o_ldir));
            else dropped++;
        } else if (type == 2) {
            if (env->pointCount < 4) memcpy(&env->point[env->pointCount++], body, sizeof(o_lpoint));
            else dropped++;
        } else if (type == 3) {
            if (env->spotCount < 4) memcpy(&env->spot[env->spotCount++], body, sizeof(o_lspot));
            else dropped++;
        } else {
            dropped++;
        }
    }
    return dropped;
}

int or_resolve_view(const o_beam* beams, uint32_t sampleCount, const float m[16], const ov2* offsets, uint32_t A,
                    const uint32_t* triIds, const o_tri* tris, const o_trishade* shade, const o_vertex* verts,
                    const o_material* materials, const o_lightenv* env, int precision, uint32_t* colorOut,
                    float* linearOut) {
    return or_resolve_view_dof(beams, sampleCount, m, offsets, A, triIds, tris, shade, verts, materials, env,
                               precision, NULL, NULL, colorOut, linearOut);
}

int or_resolve_view_dof(const o_beam* beams, uint32_t sampleCount, const float m[16], const ov2* offsets,
                        uint32_t A, const uint32_t* triIds, const o_tri* tris, const o_trishade* shade,
                        const o_vertex* verts, const o_material* materials, const o_lightenv* env, int precision,
                        const o_dof* dof, const uint32_t* counts, uint32_t* colorOut, float* linearOut) {
    if (A < 1 || A > 32) return -1;
    if (dof && (!(dof->radius > 0.0f) || !dof->lensTable || !counts)) return -1;
    ov3 cam = v3(m[12], m[13], m[14]);
    uint32_t rank = 0, rankTile = 0;
    dof_sample_ctx dc;
    memset(&dc, 0, sizeof(dc));
    for (uint32_t s = 0; s < sampleCount; s++) {
        uint32_t slotTri[32], slotMask[32];
        uint32_t n = or_emit_gbuffer(triIds + (uint64_t)s * A, A, slotTri, slotMask);
        ov3 dir = m3mul(m, beams[s].c), du = m3mul(m, beams[s].du), dv = m3mul(m, beams[s].dv);
        const dof_sample_ctx* pdc = NULL;
        if (dof) {
            uint32_t tile = s / 128;
            if (s % 128 == 0) {
                const o_beam* mid = &beams[(uint64_t)tile * 128 + 64];
                dof_tile_load(m, m3mul(m, mid->du), m3mul(m, mid->dv), dof->radius, dof->focalDistance, &dc.tile);
                dc.row = dof_row(counts, tile, &rank, &rankTile);
                dc.table = dof->lensTable;
            }
            dc.sample = s % 128;
            dc.lensToFocal = dof_lens_to_focal(m, dir, dof->focalDistance);
            pdc = &dc;
        }
        if (precision == 0) {
            float lin[3];
            colorOut[s] = shade_sample_f(dir, du, dv, cam, offsets, A, slotTri, slotMask, n, tris, shade, verts,
                                         materials, env, pdc, lin);
            if (linearOut) memcpy(linearOut + 3 * (uint64_t)s, lin, sizeof(lin));
        } else {
            double lin[3];
            colorOut[s] = shade_sample_d(dir, du, dv, cam, offsets, A, slotTri, slotMask, n, tris, shade, verts,
                                         materials, env, pdc, lin);
            if (linearOut)
                for (int c = 0; c < 3; c++) linearOut[3 * (uint64_t)s + c] = (float)lin[c];
        }
    }
    return 0;
}

/* RemapKernel (raycaster/remap.cu:35-77) */
void or_remap(const uint32_t* results, const int32_t* remap, uint32_t width, uint32_t height, uint32_t split,
              uint32_t stride, uint32_t* image) {
    for (uint32_t y = 0; y < height; y++)
        for (uint32_t x = 0; x < width; x++) {
            if (split == 1) {
                int32_t o = remap[(uint64_t)y * width + x];
                if (o >= 0) image[(uint64_t)y * stride + x] = results[o];
            } else {
                uint32_t c[3] = {0, 0, 0};
                for (uint32_t ch = 0; ch < split; ch++) {
                    int32_t o = remap[((uint64_t)y * width + x) * split + ch];
                    if (o >= 0) c[ch] = results[o];
                }
                if (split == 2)
                    image[(uint64_t)y * stride + x] = (c[1] & 0xff) | (c[0] & 0xff00) | (c[1] & 0xff0000) | 0xff000000u;
                else
                    image[(uint64_t)y * stride + x] = (c[0] & 0xff) | (c[1] & 0xff00) | (c[2] & 0xff0000) | 0xff000000u;
            }
        }
}
