/* Resolve / GGX shading body, instantiated for float (same operation order as the
 * GPU kernels) and double (the FP64 shading oracle). Included by or_shade.c with
 * R = float|double and SFX = _f|_d. Test infrastructure only.
 *
 *   ShadeMSAA                          raycaster/resolve.cu:161-284
 *   IntersectTriangleThread::calcUVW   raycaster/prim_tests.h:233-245
 *   IntersectTriangleThreadDoF::calcUVW raycaster/prim_tests.h:366-392 (FMAs as in its test())
 *   interpolate                        raycaster/shading.h:19-36
 *   GGXShade                           raycaster/shading.h:84-161
 *   GGX_SchlickFresnel_SchlickSmithGGX raycaster/shading_helpers.h:89-112
 *   PBR_BRDF1, light contributions     raycaster/shading_helpers.h:115-152
 *   ACESFilm, ToColor4Unorm8SRgb       raycaster/shading_helpers.h:158-200
 *   linearToSRgb                       shared/util_graphics.h:26-28
 *   UV derivatives for tex2DGrad       raycaster/resolve.cu:236-275 (barycentrics one unit
 *                                      along each subsample / focal-offset axis)
 *   textured material inputs           raycaster/shading.h:93-127 (or_texture.c samples)
 */
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

typedef struct { R x, y, z; } FN(rv3);
#define RV3 FN(rv3)

static inline RV3 FN(mk)(R x, R y, R z) { RV3 r = {x, y, z}; return r; }
static inline RV3 FN(up)(ov3 a) { return FN(mk)((R)a.x, (R)a.y, (R)a.z); }
static inline RV3 FN(add)(RV3 a, RV3 b) { return FN(mk)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline RV3 FN(sub)(RV3 a, RV3 b) { return FN(mk)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline RV3 FN(mul)(RV3 a, RV3 b) { return FN(mk)(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline RV3 FN(scl)(RV3 a, R s) { return FN(mk)(a.x * s, a.y * s, a.z * s); }
static inline R FN(dot)(RV3 a, RV3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline RV3 FN(crs)(RV3 a, RV3 b) {
    return FN(mk)(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline RV3 FN(nrm)(RV3 a) {
    R l = (R)sqrt((double)FN(dot)(a, a));
    if (sizeof(R) == 4) l = (R)sqrtf((float)FN(dot)(a, a));
    return FN(mk)(a.x / l, a.y / l, a.z / l);
}
static inline R FN(sat)(R x) { return x > (R)0 ? (x < (R)1 ? x : (R)1) : (R)0; }
static inline R FN(fm)(R a, R b, R c) {
    return sizeof(R) == 4 ? (R)fmaf((float)a, (float)b, (float)c) : (R)fma((double)a, (double)b, (double)c);
}
static inline R FN(pw)(R x, R y) { return sizeof(R) == 4 ? (R)powf((float)x, (float)y) : (R)pow((double)x, (double)y); }

static RV3 FN(ggx)(RV3 N, RV3 L, RV3 V, R roughness, RV3 F0) {
    R alpha = roughness * roughness;
    RV3 H = FN(nrm)(FN(add)(L, V));
    R alpha2 = alpha * alpha;
    R NdotH = FN(sat)(FN(dot)(N, H));
    R q = NdotH * NdotH * (alpha2 - (R)1) + (R)1;
    R D = alpha2 / ((R)3.141592653589f * (q * q));
    R LdotH = FN(sat)(FN(dot)(L, H));
    R x = (R)1 - LdotH;
    R x2 = x * x;
    R p5 = x2 * x2 * x;
    RV3 F = FN(add)(F0, FN(scl)(FN(sub)(FN(mk)(1, 1, 1), F0), p5));
    R NdotL = FN(sat)(FN(dot)(N, L));
    R NdotV = FN(sat)(FN(dot)(N, V));
    R k = alpha * (R)0.5f;
    R vis = ((R)1 / (NdotL * ((R)1 - k) + k)) * ((R)1 / (NdotV * ((R)1 - k) + k));
    RV3 r = FN(scl)(FN(scl)(F, D), vis);
    return FN(scl)(r, (R)1 / (R)4) ; /* F * D * Vis / 4 */
}

static RV3 FN(brdf)(RV3 L, RV3 N, RV3 V, RV3 base, R roughness, RV3 F0, RV3 lightColor) {
    RV3 spec = FN(ggx)(N, L, V, roughness, F0);
    R NdotL = FN(sat)(FN(dot)(N, L));
    RV3 diffuse = FN(scl)(FN(mul)(FN(mul)(base, lightColor), FN(sub)(FN(mk)(1, 1, 1), spec)), NdotL);
    RV3 glossy = FN(mul)(lightColor, spec);
    return FN(add)(diffuse, glossy);
}

/* a material's inputs to GGXShade: its constants, or where it names a texture of the
 * current set, the texture's value at the sample (shading.h:93-127) */
typedef struct { uint32_t model; RV3 emissiveColor, emissive, base, F0; R roughness; } FN(sin);

static FN(sin) FN(material_inputs)(const o_material* mat) {
    FN(sin) in;
    in.model = (uint32_t)(mat->tex & 0xffffu);
    in.emissiveColor = in.emissive = FN(mk)((R)mat->emissive[0], (R)mat->emissive[1], (R)mat->emissive[2]);
    in.base = FN(mk)((R)mat->diffuse[0], (R)mat->diffuse[1], (R)mat->diffuse[2]);
    in.F0 = FN(mk)((R)mat->specular[0], (R)mat->specular[1], (R)mat->specular[2]);
    in.roughness = (R)(1.0 - (double)mat->glossiness);
    return in;
}

static uint32_t FN(tex_id)(const o_material* mat, int shift) { return (uint32_t)(mat->tex >> shift) & 0xffffu; }

/* whether shading `mat` reads a texture of the current set */
static int FN(textured)(const o_material* mat) {
    if ((mat->tex & 0xffffu) == 2) return or_textures_present(FN(tex_id)(mat, 48));
    return or_textures_present(FN(tex_id)(mat, 48)) || or_textures_present(FN(tex_id)(mat, 32)) ||
           or_textures_present(FN(tex_id)(mat, 16));
}

static FN(sin) FN(surface_inputs)(const o_material* mat, R u, R v, R dxu, R dxv, R dyu, R dyv) {
    FN(sin) in = FN(material_inputs)(mat);
    float c[4];
#define READ(id) or_texture_read(id, (float)u, (float)v, (float)dxu, (float)dxv, (float)dyu, (float)dyv, c)
    if (in.model == 2) {
        if (READ(FN(tex_id)(mat, 48))) in.emissiveColor = FN(mk)((R)c[0], (R)c[1], (R)c[2]);
        return in;
    }
    if (READ(FN(tex_id)(mat, 48))) in.base = FN(mk)((R)c[0], (R)c[1], (R)c[2]);
    if (READ(FN(tex_id)(mat, 32))) in.F0 = FN(mk)((R)c[0], (R)c[1], (R)c[2]);
    if (READ(FN(tex_id)(mat, 16))) in.roughness = (R)1 - (R)c[0];
#undef READ
    return in;
}

/* GGXShade; returns rgb */
static RV3 FN(ggx_shade)(const FN(sin)* in, RV3 pos, RV3 nrmIn, RV3 camPos, const o_lightenv* env) {
    RV3 normal = FN(nrm)(nrmIn);
    if (in->model == 2) /* emissive */
        return in->emissiveColor;
    RV3 base = in->base;
    RV3 V = FN(nrm)(FN(sub)(camPos, pos));
    RV3 F0 = in->F0;
    R roughness = in->roughness;
    RV3 rad = in->emissive;
    rad = FN(add)(rad, FN(mul)(base, FN(mk)((R)0.15f, (R)0.15f, (R)0.165f)));
    for (int i = 0; i < env->dirCount && i < 1; i++) {
        RV3 L = FN(up)(env->dir[i].Direction);
        rad = FN(add)(rad, FN(brdf)(L, normal, V, base, roughness, F0, FN(up)(env->dir[i].Power)));
    }
    for (int i = 0; i < env->pointCount && i < 4; i++) {
        const o_lpoint* l = &env->point[i];
        RV3 delta = FN(sub)(FN(up)(l->Position), pos);
        R r = (R)sqrt((double)FN(dot)(delta, delta));
        if (sizeof(R) == 4) r = (R)sqrtf((float)FN(dot)(delta, delta));
        RV3 L = FN(nrm)(delta);
        R fdist = FN(sat)(((R)l->FalloffEnd - r) * (R)l->FalloffScale);
        RV3 lc = FN(scl)(FN(up)(l->Color), fdist);
        rad = FN(add)(rad, FN(brdf)(L, normal, V, base, roughness, F0, lc));
    }
    for (int i = 0; i < env->spotCount && i < 4; i++) {
        const o_lspot* l = &env->spot[i];
        RV3 delta = FN(sub)(FN(up)(l->Position), pos);
        R r = (R)sqrt((double)FN(dot)(delta, delta));
        if (sizeof(R) == 4) r = (R)sqrtf((float)FN(dot)(delta, delta));
        RV3 L = FN(nrm)(delta);
        R fdist = FN(sat)(((R)l->FalloffEnd - r) * (R)l->FalloffScale);
        RV3 lc = FN(scl)(FN(up)(l->Color), fdist);
        RV3 lit = FN(brdf)(L, normal, V, base, roughness, F0, lc);
        R cosA = FN(dot)(FN(up)(l->Direction), FN(scl)(L, (R)-1));
        R fang = FN(sat)((cosA - (R)l->CosOuterAngle) * (R)l->CosAngleScale);
        rad = FN(add)(rad, FN(scl)(lit, fang));
    }
    return FN(mk)(rad.x > 0 ? rad.x : 0, rad.y > 0 ? rad.y : 0, rad.z > 0 ? rad.z : 0);
}

static R FN(aces)(R x) {
    R a = (R)2.51f, b = (R)0.03f, c = (R)2.43f, d = (R)0.59f, e = (R)0.14f;
    return FN(sat)((x * (a * x + b)) / (x * (c * x + d) + e));
}
static R FN(srgb)(R x) {
    return (x <= (R)0.0031308f) ? (x * (R)12.92f) : (FN(pw)(x, (R)(1.0f / 2.4f)) * (R)1.055f - (R)0.055f);
}

/* barycentrics of triangle T at subsample offset (cx, cy) (IntersectTriangleThread::calcUVW)
 * or, with depth of field, at lens position (lx, ly) and focal-plane offset (dx, dy)
 * (IntersectTriangleTile<true>::setup + IntersectTriangleThreadDoF ctor + calcUVW) */
static void FN(bary)(const o_tri* T, RV3 cam, RV3 dir, RV3 du, RV3 dv, const dof_sample_ctx* dc, ov3 camPos, R lx, R ly,
                     R dx, R dy, R cx, R cy, R b[3]) {
    if (dc) {
        RV3 e0 = FN(up)(T->e0), e1 = FN(up)(T->e1);
        RV3 normal = FN(crs)(e0, e1);
        RV3 lu = FN(up)(dc->tile.lensU), lv = FN(up)(dc->tile.lensV), lcfc = FN(up)(dc->lensToFocal);
        RV3 v0ToLens = FN(sub)(FN(up)(camPos), FN(up)(T->v0));
        R dDA0 = FN(dot)(FN(scl)(lu, (R)-1), normal), dDA1 = FN(dot)(FN(scl)(lv, (R)-1), normal);
        R denomC = FN(dot)(FN(scl)(lcfc, (R)-1), normal);
        R denom = FN(fm)(dDA1, dy, FN(fm)(dDA0, dx, denomC));
        RV3 o = FN(mk)(FN(fm)(ly, lv.x, FN(fm)(lx, lu.x, v0ToLens.x)), FN(fm)(ly, lv.y, FN(fm)(lx, lu.y, v0ToLens.y)),
                       FN(fm)(ly, lv.z, FN(fm)(lx, lu.z, v0ToLens.z)));
        RV3 rd = FN(mk)(FN(fm)(dy, lv.x, FN(fm)(dx, lu.x, lcfc.x)), FN(fm)(dy, lv.y, FN(fm)(dx, lu.y, lcfc.y)),
                        FN(fm)(dy, lv.z, FN(fm)(dx, lu.z, lcfc.z)));
        RV3 nr = FN(scl)(rd, (R)-1);
        RV3 e = FN(mk)(FN(fm)(nr.y, o.z, -(nr.z * o.y)), FN(fm)(nr.z, o.x, -(nr.x * o.z)),
                       FN(fm)(nr.x, o.y, -(nr.y * o.x)));
        R v = FN(fm)(e1.z, e.z, FN(fm)(e1.y, e.y, e1.x * e.x));
        R w = -FN(fm)(e0.z, e.z, FN(fm)(e0.y, e.y, e0.x * e.x));
        R u = (denom - v) - w;
        R dinv = (R)1 / denom;
        b[0] = u * dinv;
        b[1] = v * dinv;
        b[2] = w * dinv;
    } else {
        RV3 e0 = FN(up)(T->e0), e1 = FN(up)(T->e1);
        RV3 v0ToRayOrigin = FN(sub)(cam, FN(up)(T->v0));
        RV3 normal = FN(crs)(e0, e1);
        RV3 ndir = FN(scl)(dir, (R)-1), ndu = FN(scl)(du, (R)-1), ndv = FN(scl)(dv, (R)-1);
        R denomC = FN(dot)(ndir, normal);
        RV3 eC = FN(crs)(ndir, v0ToRayOrigin);
        R vC = FN(dot)(e1, eC);
        R wC = -FN(dot)(e0, eC);
        R dD0 = FN(dot)(ndu, normal), dD1 = FN(dot)(ndv, normal);
        RV3 cu = FN(crs)(ndu, v0ToRayOrigin), cv = FN(crs)(ndv, v0ToRayOrigin);
        R dV0 = FN(dot)(e1, cu), dV1 = FN(dot)(e1, cv);
        R dW0 = -FN(dot)(e0, cu), dW1 = -FN(dot)(e0, cv);
        R denom = FN(fm)(dD1, cy, FN(fm)(dD0, cx, denomC));
        R v = FN(fm)(dV1, cy, FN(fm)(dV0, cx, vC));
        R w = FN(fm)(dW1, cy, FN(fm)(dW0, cx, wC));
        R u = (denom - v) - w;
        R dinv = (R)1 / denom;
        b[0] = u * dinv;
        b[1] = v * dinv;
        b[2] = w * dinv;
    }
}

/* shade one sample from its compressed slots; returns RGBA8, writes pre-tonemap rgb */
static uint32_t FN(shade_sample)(ov3 dirC, ov3 duC, ov3 dvC, ov3 camPos, const ov2* offsets, uint32_t A,
                                 const uint32_t* slotTri, const uint32_t* slotMask, uint32_t slots,
                                 const o_tri* tris, const o_trishade* shade, const o_vertex* verts,
                                 const o_material* materials, const o_lightenv* env, const dof_sample_ctx* dc,
                                 R outLinear[3]) {
    uint32_t maskAll = A < 32 ? ((1u << A) - 1u) : 0xffffffffu;
    RV3 result = FN(mk)(0, 0, 0);
    uint32_t combined = 0;
    RV3 dir = FN(up)(dirC), du = FN(up)(duC), dv = FN(up)(dvC), cam = FN(up)(camPos);
    for (uint32_t slot = 0; slot < slots && slot < A; slot++) {
        if (combined == maskAll) break;
        uint32_t tri = slotTri[slot];
        if (tri == 0xffffffffu) break;
        uint32_t mask = slotMask[slot];
        combined |= mask;
        R inv = (R)1 / (R)__builtin_popcount(mask);
        R cx = 0, cy = 0;
        R lx = 0, ly = 0, dx = 0, dy = 0; /* depth of field: centroid lens and direction offsets */
        for (uint32_t i = 0; i < A; i++) {
            if (!(mask & (1u << i))) continue;
            cx = cx + (R)offsets[i].x * inv;
            cy = cy + (R)offsets[i].y * inv;
            if (dc) {
                ov2 lensUV, dirUV;
                dof_sample_uvs(dc->table, dc->row, A, i, dc->sample, offsets[i], dc->tile.focalToLensScale, &lensUV,
                               &dirUV);
                lx = lx + (R)lensUV.x * inv;
                ly = ly + (R)lensUV.y * inv;
                dx = dx + (R)dirUV.x * inv;
                dy = dy + (R)dirUV.y * inv;
            }
        }
        const o_tri* T = &tris[tri];
        R b[3];
        FN(bary)(T, cam, dir, du, dv, dc, camPos, lx, ly, dx, dy, cx, cy, b);
        R bx = b[0], by = b[1], bz = b[2];
        const o_trishade* S = &shade[tri];
        const o_vertex* V0 = &verts[S->idx[0]];
        const o_vertex* V1 = &verts[S->idx[1]];
        const o_vertex* V2 = &verts[S->idx[2]];
        RV3 pos = FN(add)(FN(add)(FN(scl)(FN(up)(V0->pos), bx), FN(scl)(FN(up)(V1->pos), by)),
                          FN(scl)(FN(up)(V2->pos), bz));
        RV3 n0 = FN(mk)((R)h2f(V0->normal[0]), (R)h2f(V0->normal[1]), (R)h2f(V0->normal[2]));
        RV3 n1 = FN(mk)((R)h2f(V1->normal[0]), (R)h2f(V1->normal[1]), (R)h2f(V1->normal[2]));
        RV3 n2 = FN(mk)((R)h2f(V2->normal[0]), (R)h2f(V2->normal[1]), (R)h2f(V2->normal[2]));
        RV3 nrm = FN(add)(FN(add)(FN(scl)(n0, bx), FN(scl)(n1, by)), FN(scl)(n2, bz));
        const o_material* mat = &materials[S->material];
        FN(sin) in;
        if (FN(textured)(mat)) {
            /* UV derivatives (resolve.cu:236-275): barycentrics one unit along each subsample
             * axis from the centroid (depth of field: along each focal-offset axis) */
            R xb[3], yb[3];
            FN(bary)(T, cam, dir, du, dv, dc, camPos, lx, ly, dx + (R)1, dy, cx + (R)1, cy, xb);
            FN(bary)(T, cam, dir, du, dv, dc, camPos, lx, ly, dx, dy + (R)1, cx, cy + (R)1, yb);
            R uvx[3] = {(R)V0->uv.x, (R)V1->uv.x, (R)V2->uv.x}, uvy[3] = {(R)V0->uv.y, (R)V1->uv.y, (R)V2->uv.y};
            R u = uvx[0] * bx + uvx[1] * by + uvx[2] * bz, v = uvy[0] * bx + uvy[1] * by + uvy[2] * bz;
            R dxu = (uvx[0] * xb[0] + uvx[1] * xb[1] + uvx[2] * xb[2]) - u;
            R dxv = (uvy[0] * xb[0] + uvy[1] * xb[1] + uvy[2] * xb[2]) - v;
            R dyu = (uvx[0] * yb[0] + uvx[1] * yb[1] + uvx[2] * yb[2]) - u;
            R dyv = (uvy[0] * yb[0] + uvy[1] * yb[1] + uvy[2] * yb[2]) - v;
            in = FN(surface_inputs)(mat, u, v, dxu, dxv, dyu, dyv);
        } else {
            in = FN(material_inputs)(mat);
        }
        RV3 c = FN(ggx_shade)(&in, pos, nrm, cam, env);
        R weight = (R)__builtin_popcount(mask);
        result = FN(add)(result, FN(scl)(c, weight));
    }
    R invA = (R)1 / (R)A;
    result = FN(scl)(result, invA);
    if (outLinear) {
        outLinear[0] = result.x;
        outLinear[1] = result.y;
        outLinear[2] = result.z;
    }
    R r = FN(srgb)(FN(aces)(result.x)), g = FN(srgb)(FN(aces)(result.y)), b = FN(srgb)(FN(aces)(result.z));
    uint32_t ri = (uint32_t)(r * (R)255.0f), gi = (uint32_t)(g * (R)255.0f), bi = (uint32_t)(b * (R)255.0f);
    return ri | (gi << 8) | (bi << 16) | 0xff000000u;
}

#undef RV3
#undef FN
#undef CAT
#undef CAT2
