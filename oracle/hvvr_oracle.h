#ifndef HVVR_ORACLE_H
#define HVVR_ORACLE_H
/*
 * CPU oracle for the HVVR raycasting hot path. TEST INFRASTRUCTURE ONLY: nothing in
 * the product (hvvr_amd/, include/) links, loads or calls this code; only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, to check results.
 *
 * It restates, in plain C99 and scalar form, the reference algorithms the GPU
 * kernels implement (citations are relative to the reference repository root,
 * libraries/hvvr/...):
 *   frustum setup      raycaster/traversal.h:38-47, raycaster/traverse_avx.cpp:22-211
 *   child AABB tests   raycaster/traverse_avx.h:104-217
 *   block traversal    raycaster/traverse_avx.h:293-410
 *   tile traversal     raycaster/traverse_avx.h:412-527, raycaster/traversal.cpp:99-211
 *   intersect          raycaster/intersect.cu:117-342, raycaster/prim_tests.h:18-245
 *   GBuffer compress   raycaster/intersect.cu:28-89, raycaster/sort.h:40-74
 *   resolve / shade    raycaster/resolve.cu:161-284,324-384, raycaster/shading.h:19-161,
 *                      raycaster/shading_helpers.h:89-200, shared/util_graphics.h:26-28
 *   clear / remap      raycaster/resolve.cu:527-542, raycaster/remap.cu:35-77
 *   foveated           raycaster/gpu_foveated.cu:37-333, remap.cu:127-141,
 *                      raycaster/foveated.h:86-111, resolve.cu:386-457 (t-max)
 *   refit              raycaster/gpu_scene_state.cu:88-375
 *
 * Arithmetic contract (shared with the GPU kernels, so results compare bit for bit):
 * compiled with -ffp-contract=off; fmaf() exactly where the reference fuses
 * (_mm256_fmadd_ps and the "n FMA" annotations of prim_tests.h); correctly rounded
 * division and sqrt; IEEE denormals. The reference's own binaries run with fast-math
 * and FTZ/DAZ on NVIDIA hardware and cannot be run here, so agreement with them is
 * not claimed bit for bit; each function is pinned instead by closed forms, brute
 * force and invariants (tests/test_oracle_pins.py and the tests named in each file).
 *
 * All layouts below are byte-identical to the product's (include/hvvr/types.h),
 * which a CPU test checks field by field.
 */
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y; } ov2;
typedef struct { float x, y, z; } ov3;

typedef struct { ov3 pos; uint16_t normal[4]; ov2 uv; uint32_t pad; } o_vertex;          /* 32 B */
typedef struct { ov3 v0, e0, e1; } o_tri;                                                /* 36 B */
typedef struct { uint32_t idx[3]; uint32_t material; } o_trishade;                       /* 16 B */
typedef struct { float b[6][4]; uint32_t leafMask; uint32_t data[5]; uint32_t pad[2]; } o_node; /* 128 B */
typedef struct { float emissive[4], diffuse[4], specular[4]; float glossiness, opacity; uint64_t tex; } o_material;
typedef struct { ov3 o[4]; ov3 d[4]; } o_sfrustum;                                       /* 96 B */
typedef struct { ov3 c, du, dv; } o_beam;                                                /* 36 B */
typedef struct {
    float planeNeg[4][4], planePos[4][4];
    float projLo[3], projHi[3];
    float edge[12][3], edgeMin[12], edgeMax[12];
    int32_t axis, degenerate;
} o_frustum;                                                                             /* 400 B */
typedef struct { float tMin; uint32_t nodeMask; } o_stack;                               /* 8 B */

typedef struct { ov3 Direction, Power; } o_ldir;
typedef struct { ov3 Position; float FalloffEnd; ov3 Color; float FalloffScale; } o_lpoint;
typedef struct { ov3 Position; float FalloffEnd; ov3 Direction; float FalloffScale; ov3 Color;
                 float CosOuterAngle, CosAngleScale; } o_lspot;
typedef struct {
    o_ldir dir[1]; o_lpoint point[4]; o_lspot spot[4];
    int32_t dirCount, pointCount, spotCount;
} o_lightenv;                                                                            /* 372 B */

/* sizes of the layouts, for the layout check */
void or_layout_sizes(uint32_t out[10]);

/* ---- frustum setup ---- */
/* in: n camera-space frusta; m: camera-to-world (column-major 4x4).
 * out: world-space frusta and the pre-swizzled test data. */
void or_frustum_setup(const o_sfrustum* in, uint32_t n, const float m[16], o_sfrustum* worldOut, o_frustum* preOut);
/* with screen cull planes (planes[16] = 4 x (normal, dist) in world space, or NULL; dist[0]
 * = +inf disables them): a frustum wholly outside one becomes degenerate
 * (render.cpp:271-340, the foveated path) */
void or_frustum_setup_cull(const o_sfrustum* in, uint32_t n, const float m[16], const float* planes,
                           o_sfrustum* worldOut, o_frustum* preOut);
/* AABB tests on one node: coarse mask, then refine of a previous mask */
uint32_t or_test_children(const o_frustum* f, const o_node* node);
uint32_t or_test_children_refine(const o_frustum* f, const o_node* node, uint32_t prev);

/* ---- traversal ---- */
/* block traversal + distance sort; returns the stack size (<= 64) */
uint32_t or_block_traverse(const o_node* nodes, const o_frustum* f, o_stack* sorted);
/* tile traversal from a block's sorted stack; returns the triangle count, or -1 when the
 * 4096-entry stack or the output capacity would overflow */
int64_t or_tile_traverse(const o_node* nodes, const o_stack* blockStack, uint32_t stackSize, const o_frustum* f,
                         uint32_t* out, uint64_t cap);
/* all tiles of a view in tile order: counts[tile], lists concatenated into out; returns
 * total or -1 on overflow */
int64_t or_build_tile_lists(const o_node* nodes, const o_frustum* blockPre, const o_frustum* tilePre,
                            uint32_t blockCount, uint32_t* counts, uint32_t* out, uint64_t cap);

/* ---- depth of field ---- */
/* Thin lens (ThinLens, raycaster/camera.h) and the lens-position table of
 * GPUCamera::initLookupTables (raycaster/gpu_camera.cu:59-88): 4 rows x 128 samples x A
 * subsamples, entry [row * 128 * A + subsample * 128 + sample]. A tile uses row
 * (rank of the tile among the tiles with a non-empty list) % 4, as the reference's
 * kernels index it with their compacted-tile block index. */
typedef struct {
    float radius, focalDistance;
    const ov2* lensTable;
} o_dof;

/* ---- intersect ---- */
/* Intersects one view. worldTiles: world-space tile frusta; lists as from
 * or_build_tile_lists. Outputs per subsample (sample*A + i): triangle id (0xffffffff =
 * miss) and depth (FLT_MAX on miss). survivorsOut: number of (tile, triangle) pairs
 * surviving the tile culls, summed over tiles. */
int or_intersect_view(const o_beam* beams, uint32_t sampleCount, const o_sfrustum* worldTiles, uint32_t tileCount,
                      const float m[16], const ov2* offsets, uint32_t A, const uint32_t* counts,
                      const uint32_t* lists, const o_tri* tris, uint32_t* triOut, float* depthOut,
                      uint64_t* survivorsOut);
/* or_intersect_view with depth of field: IntersectTriangleTile<true>, IntersectTriangleThreadDoF,
 * TileData<true>, GetSampleUVsDoF (prim_tests.h:247-392, tile_data.h:37-60,
 * gpu_samples.h:84-103). dof == NULL is the pinhole path. */
int or_intersect_view_dof(const o_beam* beams, uint32_t sampleCount, const o_sfrustum* worldTiles,
                          uint32_t tileCount, const float m[16], const ov2* offsets, uint32_t A,
                          const uint32_t* counts, const uint32_t* lists, const o_tri* tris, const o_dof* dof,
                          uint32_t* triOut, float* depthOut, uint64_t* survivorsOut);
/* The world-space depth-of-field rays the intersect tests, per subsample (sample*A + i):
 * origin = lens centre + lensUV.x*lensU + lensUV.y*lensV, direction = lensCenterToFocalCenter
 * + dirUV.x*lensU + dirUV.y*lensV (prim_tests.h:330-333), for tiles with a non-empty
 * list (zero elsewhere). For tests of the DoF path against brute force. */
int or_dof_rays(const o_beam* beams, uint32_t sampleCount, const float m[16], const ov2* offsets, uint32_t A,
                const o_dof* dof, const uint32_t* counts, ov3* originsOut, ov3* dirsOut);
/* Compresses one sample's A subsample ids into (tri, mask) slots; returns the slot count. */
uint32_t or_emit_gbuffer(const uint32_t* sampleTri, uint32_t A, uint32_t* slotTri, uint32_t* slotMask);

/* ---- resolve ---- */
/* Shades every sample from its subsample ids. colorOut: RGBA8 per sample;
 * linearOut (optional): pre-tonemap linear radiance / A (rgb) per sample.
 * precision: 0 = float (same operation order as the kernels), 1 = double. */
int or_resolve_view(const o_beam* beams, uint32_t sampleCount, const float m[16], const ov2* offsets, uint32_t A,
                    const uint32_t* triIds, const o_tri* tris, const o_trishade* shade, const o_vertex* verts,
                    const o_material* materials, const o_lightenv* env, int precision, uint32_t* colorOut,
                    float* linearOut);
/* or_resolve_view with depth of field (the DoF branch of ShadeMSAA, resolve.cu:161-284,
 * 324-384); counts: tile list lengths, for the lens-table row of each tile. */
int or_resolve_view_dof(const o_beam* beams, uint32_t sampleCount, const float m[16], const ov2* offsets,
                        uint32_t A, const uint32_t* triIds, const o_tri* tris, const o_trishade* shade,
                        const o_vertex* verts, const o_material* materials, const o_lightenv* env, int precision,
                        const o_dof* dof, const uint32_t* counts, uint32_t* colorOut, float* linearOut);
/* image[y*stride + x] = results[remap[...]] (split 1, 2 or 3) */
void or_remap(const uint32_t* results, const int32_t* remap, uint32_t width, uint32_t height, uint32_t split,
              uint32_t stride, uint32_t* image);

/* ---- foveated path (raycaster/gpu_foveated.cu, remap.cu:127-158, resolve.cu:386-457) ---- */
/* Per sample: distance along the view axis (-m column 2) of subsample 0's hit point from its
 * barycentrics (calcUVW at subsample 0), +inf on a miss or in a tile with an empty list. */
int or_tmax_view(const o_beam* beams, uint32_t sampleCount, const float m[16], const ov2* offsets, uint32_t A,
                 const uint32_t* triIds, const o_tri* tris, const o_dof* dof, const uint32_t* counts, float* tMaxOut);
/* Per subsample (sample*A + i, 3 floats): barycentrics (u, v, w) of the subsample's hit
 * triangle on that subsample's own ray (IntersectTriangleThread::calcUVW, prim_tests.h:233-245;
 * with depth of field IntersectTriangleThreadDoF::calcUVW, :366-392, FMAs as in its test),
 * zero on a miss or in a tile with an empty list. The ray parameter of the hit is the
 * depth or_intersect_view returns. */
int or_hits_view(const o_beam* beams, uint32_t sampleCount, const float m[16], const ov2* offsets, uint32_t A,
                 const uint32_t* triIds, const o_tri* tris, const o_dof* dof, const uint32_t* counts, float* uvwOut);
/* sRGB byte -> linear (util_graphics.h:16-18), and byte -> linear -> byte (truncated) */
void or_srgb_tables(float toLinear[256], uint8_t roundTrip[256]);
/* RemapPolarFoveatedKernel: polar[remap[i]] = sRGB round trip of results[i], depth = tMax[i] */
void or_polar_remap(const uint32_t* results, const float* tMax, const uint32_t* remap, uint32_t count,
                    uint32_t width, uint32_t* polarColor, float* polarDepth);
/* ComputeMoments: 3x3 mean / mean square (RGBA16 unorm, 4 x 16 bits) per polar texel */
void or_polar_moments(const uint32_t* polarColor, uint32_t w, uint32_t h, uint64_t* m1, uint64_t* m2);
typedef struct {
    const uint32_t* polarColor;
    const float* polarDepth;
    const uint64_t *moment1, *moment2;
    const uint32_t* previous; /* [height][width] */
    uint32_t polarW, polarH, width, height;
    float sampleToEye[9];
    float eyeToSamplePrevious[16];
    float emap[6];            /* m, w0, S, 1/log(1+m), 1/w0, 1/m */
    float maxEccentricity;
    float alpha, stddevMultiplier;
} o_reproject;
/* FoveatedPolarToScreenSpaceKernel: the temporally filtered screen image */
void or_foveated_reproject(const o_reproject* a, uint32_t* out);
/* SeparableFilterUsingEccentricity with step (stepX, stepY) */
void or_foveated_blur(const uint32_t* in, uint32_t* out, uint32_t w, uint32_t h, int stepX, int stepY,
                      const float sampleToEye[9]);
/* FinishConstrastEnhancement */
void or_foveated_contrast(const uint32_t* unfiltered, const uint32_t* filtered, uint32_t* out, uint32_t w, uint32_t h,
                          const float sampleToEye[9], float maxEccentricity, float fe);
/* EccentricityMap::apply / applyInverse (raycaster/foveated.h:86-111) */
float or_emap_apply(const float emap[6], float i);
float or_emap_apply_inverse(const float emap[6], float e);

/* ---- textures (or_texture.c; raycaster/texture.cu:78-272, shading.h:45-47, 99-127) ---- */
/* one texture's mip chain: level l of max(1, w >> l) x max(1, h >> l) texels at byte
 * offset[l] of the texel buffer; kind 0 = RGBA8 unorm, 1 = RGBA8 sRGB, 2 = RGBA32F */
typedef struct { uint32_t width, height, levels, kind; uint64_t offset[16]; } o_texinfo;  /* 144 B */
/* [0, 256): sRGB byte -> linear; [256, 512): mip-fraction thresholds 2^((j + 0.5) / 256) */
void or_texture_tables(float out[512]);
uint32_t or_texture_levels(uint32_t width, uint32_t height);
/* level offsets from 0 (each level 16-byte aligned); returns the chain's bytes */
uint64_t or_texture_layout(o_texinfo* t);
/* levels 1.. from level 0 (d_mipmap) */
void or_texture_build_mips(uint8_t* base, const o_texinfo* t);
/* tex2DGrad(uv, ddx, ddy) with wrap addressing, trilinear and up to 8 anisotropic probes */
void or_texture_sample(const uint8_t* base, const o_texinfo* t, float u, float v, float dxu, float dxv, float dyu,
                       float dyv, float out[4]);
/* level floor(lambda) and the mip fraction of lambda = log2(x) clamped to [0, levels - 1] */
void or_texture_lod(float x, uint32_t levels, uint32_t* level, float* f);
/* the textures later resolves read (offsets relative to `texels`); count 0 = none. The
 * pointers are kept, not copied. */
void or_set_textures(const uint8_t* texels, const o_texinfo* info, uint32_t count);
/* texture `id` of the current set at (uv, ddx, ddy); 0 = no such texture (out untouched) */
int or_texture_read(uint32_t id, float u, float v, float dxu, float dxv, float dyu, float dyv, float out[4]);
int or_textures_present(uint32_t id);

/* ---- refit ---- */
/* World-space vertices, precomputed triangles and node bounds from scratch. nodes:
 * topology (bounds ignored), nodesOut gets bounds filled. */
int or_refit(const o_vertex* verts, uint32_t vertexCount, const uint32_t* vertexToTransform, const float* transforms,
             const o_trishade* shade, uint32_t triCount, const o_node* nodes, uint32_t nodeCount,
             o_vertex* worldVerts, o_tri* trisOut, o_node* nodesOut);

/* lighting environment from a light list (LightUnion, 56 bytes each); returns the
 * number of lights dropped for exceeding the per-type maxima */
int or_lighting_env(const void* lights, uint32_t count, o_lightenv* env);

#ifdef __cplusplus
}
#endif
#endif
