// mtgl_dev.hip - MI355X (gfx950) implementation of the mtgl_dev C-ABI seam.
//
// Per draw (include/mtgl_dev.h, mtgl_dev_draw):
//   K1 k_vertex   one thread per emitted vertex: attribute fetch with the
//                 reference's get_array_element rules, glColor4f sanitising,
//                 emit_vertex transform and lighting (src/gl_api.c:263-348,
//                 src/lighting.h:53-142).
//   K2 k_setup    one thread per assembled triangle: Sutherland-Hodgman clip
//                 (src/clipping.h), perspective divide, fan split, integer
//                 screen mapping, cull, bounding box, area, per-triangle LOD
//                 (src/raster.c:451-529, :729-958).  Writes up to 7 fan
//                 triangles per input triangle, ordered by (triangle, fan j).
//   K3 binner     scan of per-triangle tile counts, load-balanced expansion to
//                 (tile, triangle) pairs in submission order, stable LSD radix
//                 sort by tile (LDS digit histograms, wavefront ballot ranks),
//                 segment list of non-empty tiles.
//   K4 k_raster   persistent waves; one 8x8 tile per wave, one pixel per lane;
//                 the pixel's colour, depth and stencil stay in registers while
//                 the wave walks the tile's bin in submission order, applying
//                 the per-fragment pipeline of src/raster.c:532-724 exactly.
//   K5 k_clear    scissor-aware fill (src/gl_api.c:409-457).
//   K6 k_mip1     2x2 box filter of level 1 (src/textures.c:311-354).
//
// Bit-exactness: compiled with -ffp-contract=off and without fast-math, so
// every float expression keeps the reference's operation order; fp32
// division and sqrt are the correctly rounded sequences; float->int
// conversions follow x86 cvttss2si (INT_MIN for NaN / out of range).  The
// transcendental calls of the reference (log2f for the triangle LOD, expf for
// GL_EXP/GL_EXP2 fog, powf for spot and specular terms) are evaluated in fp64
// and rounded once; spot cutoff cosines arrive precomputed by the host C
// library.

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>

#include "GL/gl.h"
#include "mtgl_batch.h"
#include "mtgl_dev.h"
#include "mtgl_diff.h"

#define MTGL_API extern "C" __attribute__((visibility("default")))

namespace {

constexpr int TILE = 8;                 // 8x8 pixels = one wave64
constexpr int MAX_CLIP = 12;            // reference MAX_CLIP_VERTS: at most 10 fan triangles
constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 8;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;
constexpr int RADIX_BITS = 8;
constexpr int RADIX = 1 << RADIX_BITS;
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;
constexpr int SCAN1_THREADS = 1024;     // the one-workgroup scan
// 64-element rows per wave it holds in registers (128 bytes per lane), and
// the elements it takes
template <typename T> constexpr uint32_t SCAN1_ROWS = sizeof(T) == 4 ? 32u : 16u;
template <typename T> constexpr uint64_t SCAN1_MAX = (uint64_t)SCAN1_ROWS<T> * SCAN1_THREADS;
constexpr uint64_t SCAN2_MAX_BLOCKS = 256;   // scans of at most this many tiles take two launches
constexpr uint32_t HEAVY_BIN = 192;     // default: tiles with longer bins are rasterised first, in parts
constexpr uint32_t HEAVY_SPLIT = 8;     // parts of a heavy tile, TILE / HEAVY_SPLIT pixel rows each
constexpr uint32_t ITEM_TILE_BITS = 28; // raster work item = global tile | part << 28 (0: whole tile, 1..HEAVY_SPLIT)
constexpr int32_t PART_ROWS = TILE / (int32_t)HEAVY_SPLIT;
static_assert(TILE % HEAVY_SPLIT == 0 && HEAVY_SPLIT < (1u << (32 - ITEM_TILE_BITS)), "heavy-tile parts");
constexpr uint32_t MED_BIN = 32;        // whole tiles with longer bins join the front of the tile list
constexpr int SETUP_BLOCK = 64;         // triangles per block of the set-up kernel (one per lane)
constexpr uint32_t SETUP_SPLIT_CU = 4;  // set-up launches of at most this many blocks per CU: a wave per corner
constexpr uint32_t EXPAND_SMALL = 64;   // pairs one thread writes; triangles with more go to a wave each
constexpr uint32_t EXPAND_WAVE = 512;   // ... its own set-up block's wave up to this many, else one in k_expand_big
constexpr uint32_t STAT_SLOTS = 16384;  // per-block fragment statistics (raster grids are capped at this)

// ---------------------------------------------------------------------------
// records in HBM
// ---------------------------------------------------------------------------

// post-transform vertex, hot part (48 B)
struct __align__(16) VHot {
    float x, y, z, w;
    float r, g, b, a;
    float s, t, ez, pad;
};

// post-transform vertex, eye-space part used by per-fragment lighting (32 B)
struct __align__(16) VCold {
    float ep[3];
    float en[3];
    float pad[2];
};

// fan triangle after setup (176 B); an unclipped triangle t owns slot t, a
// clipped one owns a run of slots from the overflow pool after the first ntri
struct __align__(16) RTri {
    int32_t x0, y0, x1, y1, x2, y2;
    int32_t minX, minY, maxX, maxY;      // viewport and scissor clamped
    float area, inv_area;
    float z0, z1, z2;
    float lod;                           // the first 64 bytes are what coverage and the depth pass read
    float w0, w1, w2;                    // 1/w
    float c0[4], c1[4], c2[4];
    float uv0[2], uv1[2], uv2[2];
    float ez0, ez1, ez2;
    int32_t back;
    int32_t pad0[3];                     // all 176 bytes are what colour computation reads
};
static_assert(sizeof(RTri) == 176, "RTri layout");

// what the binner reads of a record (16 B, same slot as the record): the
// framebuffer-clamped tile box as the launch-wide key of its first tile, its
// width, its tile count and the row pitch of the record's scene; a box of at
// most 64 tiles carries the mask of the tiles the record can cover (bit k:
// tile k of the box, row-major), and ntiles counts those
struct __align__(16) RBox {
    uint32_t key0;                       // scene tile offset + ty0 * tiles_x + tx0
    uint32_t tw;                         // tiles per box row
    uint32_t ntiles;                     // pairs of the record
    uint32_t tiles_x;
    unsigned long long mask;             // (masked) tiles of the box with pairs
    uint32_t masked;
    uint32_t pad;
};
constexpr uint32_t BOX_MASK_TILES = 64;
constexpr int RT_COV_VEC4 = 4;           // 64 bytes: vertices, box, area, 1/area, z, lod
static_assert(offsetof(RTri, lod) == 60, "RTri layout");

struct __align__(16) RCold {
    float ep[9];
    float en[9];
    float pad[2];
};

// Counters of an Exec.  The first group belongs to one launch and is reset
// by the launch's first kernel; the second persists across launches.
struct Counters {
    unsigned long long npairs;    // (tile, record) pairs of the launch (the pair-count scan or k_expand)
    unsigned int nheavy;          // work items at the front of the tile list: heavy-tile parts, medium tiles
    unsigned int nlight;          // the other non-empty tiles (with nheavy: one 64-bit word for the list atomics)
    unsigned int unsupported;
    unsigned int noverflow;       // overflow slots requested by clipped triangles
    unsigned int overflow_full;   // bit 0: the clip pool was too small; bit 1: the pair arrays were
    unsigned int nbig;            // triangles queued for the one-wave pair expansion
    unsigned int pad0;
    // persistent
    unsigned int sticky;          // overflow_full of every launch since the host last cleared it
    unsigned int poison;          // 0, or the sequence number of the first launch that overflowed: every
                                  // later framebuffer write of this Exec is skipped until the host clears it
    unsigned long long poison_pairs;   // pairs that launch needed
    unsigned int poison_flags;    // its overflow_full
    unsigned int pad1;
    unsigned long long acc_tri_raster; // fan triangles and pairs of every launch that did not overflow
    unsigned long long acc_bin_pairs;
};

static_assert(offsetof(Counters, nheavy) % 8 == 0 && offsetof(Counters, nlight) == offsetof(Counters, nheavy) + 4,
              "nheavy and nlight form one 64-bit word");

// attribute stream resolved to a device pointer
struct DArr {
    const uint8_t *src;
    int64_t size;
    int64_t offset;
    int32_t stride;
    int32_t comps;
    uint32_t type;
    uint32_t enabled;
};

struct VertexParams {
    uint32_t count;               // vertex work items: indices, or vertex slots when `unique`
    uint32_t unique;              // slot i < count - 1 is vertex ubase + i, the last slot any other index
    int32_t ubase;
    int32_t first;
    uint32_t index_type;
    uint32_t n_vstates;
    DArr indices, position, color, texcoord, normal, state_index;
    float cur_color[4], cur_texcoord[2], cur_normal[3];
    const mtgl_vertex_state *vstates;
    VHot *out_hot;
    VCold *out_cold;              // null when no per-fragment lighting needs it
};

// One scene (one draw into one framebuffer) of a launch.  A launch covers
// 1..n scenes: the triangles of scene s are global triangles
// [tri_off, tri_off + ntri), its records the same slots of the record pool,
// clipped triangles take slots of the shared overflow pool after all scenes'
// triangles.  Pointers to per-triangle arrays are already offset to the scene.
struct SetupParams {
    uint32_t mode;
    uint32_t count;
    const VertexParams *vp;       // the scene's vertex streams and states (set-up transforms its own vertices)
    uint32_t cold;                // eye-space position and normal are kept (per-fragment lighting)
    uint32_t ntri;
    uint32_t tri_off;             // first global triangle of the scene
    uint32_t scene;
    uint32_t tile_off;            // first global tile of the scene
    int32_t fb_w, fb_h;
    int32_t tiles_x;
    int32_t tex_w, tex_h;         // 0 when the draw is untextured
    const mtgl_raster_state *rs;
    RTri *rt;                     // the scene's own slots
    RCold *rcold;
    RBox *box;
    RTri *rt_over;                // the shared overflow pool (after every scene's slots)
    RCold *rcold_over;
    RBox *box_over;
    uint32_t over_base;           // global slot of rt_over[0]
    uint32_t *n_rt;               // fan triangles kept per input triangle
    uint32_t *rec_base;           // first global record slot per input triangle
    unsigned long long *tri_tiles;  // (tile, record) pairs per input triangle
    uint32_t overflow_cap;        // slots of the overflow pool
    Counters *ctr;
};

// what the raster stage needs of a scene
enum : uint32_t {
    KIND_LAST_WRITER = 1u,        // no blending, full or empty colour mask, no alpha test on a texture
    KIND_STENCIL = 2u,            // stencil test enabled
    KIND_COLOR = 4u,              // colour written
};

struct SceneRaster {
    const mtgl_raster_state *rs;
    uint32_t *color;
    float *depth;
    uint8_t *stencil;
    uint32_t *writer;             // per pixel: 1 + record slot of the last passing fragment
    const uint32_t *tex;          // level 0 (null = untextured)
    const uint32_t *tex1;         // level 1 (null = none: samples read as white)
    int32_t tex_w, tex_h, tex1_w, tex1_h;
    int32_t fb_w, fb_h, tiles_x;
    uint32_t tile_off;            // first global tile of the scene
    uint32_t kind;
    uint32_t lod_mode;            // MTGL_LOD_*
    const uint32_t *chain;        // levels 2.. of the full chain, back to back (null: none)
    int32_t nlev;                 // levels of the chain, 0 and 1 included
    int32_t pad_;
};

struct RasterParams {
    const SceneRaster *scenes;
    const uint32_t *tile_pre;      // nscenes + 1 first global tiles
    const uint32_t *tile_scene;    // global tile -> scene (nscenes > 1)
    uint32_t nscenes;
    const RTri *rt;
    const RCold *rcold;
    const uint32_t *vals;          // per bin, in submission order: record slot
    const uint4 *list;             // work items (tile | part << 28, e0, e1, scene): nheavy from the front, nlight from the back
    uint32_t ntiles;
    uint32_t list_cap;             // items the list holds (HEAVY_SPLIT per tile)
    uint32_t part;                 // list items of this launch: 0 all, 1 the heavy ones, 2 the others
    unsigned long long *stats;     // per block: fragments, algorithmic bytes (summed when read)
    Counters *ctr;
};

// ---------------------------------------------------------------------------
// scalar helpers (reference arithmetic, see file header)
// ---------------------------------------------------------------------------

__device__ __forceinline__ int32_t cvtt(float f)
{
    return (f >= -2147483648.0f && f < 2147483648.0f) ? (int32_t)f : INT32_MIN;
}

__device__ __forceinline__ int32_t isub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }

__device__ __forceinline__ float clamp01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

struct Col { float r, g, b, a; };

__device__ __forceinline__ Col col_add(Col a, Col b) { return {a.r + b.r, a.g + b.g, a.b + b.b, a.a + b.a}; }
__device__ __forceinline__ Col col_mul(Col a, Col b) { return {a.r * b.r, a.g * b.g, a.b * b.b, a.a * b.a}; }
__device__ __forceinline__ Col col_scale(Col a, float s) { return {a.r * s, a.g * s, a.b * s, a.a * s}; }
__device__ __forceinline__ Col col_clamp(Col c) { return {clamp01(c.r), clamp01(c.g), clamp01(c.b), clamp01(c.a)}; }
__device__ __forceinline__ Col col_lerp(Col a, Col b, float t) { return col_add(col_scale(a, 1.0f - t), col_scale(b, t)); }
__device__ __forceinline__ Col col_of(const mtgl_color &c) { return {c.r, c.g, c.b, c.a}; }

// clamp then truncate x*255 (reference color_to_rgba32): fmaxf maps NaN to 0,
// which is also the byte x86 produces for a NaN channel, and the converted
// value always lies in [0, 255]
__device__ __forceinline__ uint32_t pack8(float v)
{
    float x = fminf(fmaxf(v, 0.0f), 1.0f);
    return (uint32_t)(x * 255.0f);
}

__device__ __forceinline__ uint32_t col_pack(Col c)
{
    return pack8(c.r) | (pack8(c.g) << 8) | (pack8(c.b) << 16) | (pack8(c.a) << 24);
}

// byte/255.0f, correctly rounded: one product by 1/255 and one FMA residual
// correction reproduce the IEEE quotient for all 256 byte values (checked
// exhaustively by the test suite)
__device__ __forceinline__ float u8_to_unit(uint32_t b)
{
    const float r = 1.0f / 255.0f;
    float fb = (float)b;
    float m = fb * r;
    float e = __builtin_fmaf(-m, 255.0f, fb);
    return __builtin_fmaf(e, r, m);
}

__device__ __forceinline__ Col col_unpack(uint32_t p)
{
    return {u8_to_unit(p & 0xFF), u8_to_unit((p >> 8) & 0xFF), u8_to_unit((p >> 16) & 0xFF), u8_to_unit(p >> 24)};
}

__device__ __forceinline__ float dot3(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__device__ __forceinline__ void normalize3(float *v)
{
    float len = sqrtf(dot3(v, v));
    if (len > 0.0f) {
        float s = 1.0f / len;
        v[0] = v[0] * s; v[1] = v[1] * s; v[2] = v[2] * s;
    }
}

// Out of line: inlined, their fp64 polynomial constants are hoisted into the
// raster kernels' loop preheaders, held across every item and spilled to
// scratch (per-lane stores to memory on every wave); as calls they cost a
// call only on the paths that evaluate them (fog EXP/EXP2, spot and specular
// powers, per-pixel LOD).
__device__ __attribute__((noinline)) float pow_cr(float b, float e) { return (float)pow((double)b, (double)e); }
__device__ __attribute__((noinline)) float exp_cr(float x) { return (float)exp((double)x); }
__device__ __attribute__((noinline)) float log2_cr(float x) { return (float)log2((double)x); }

// m * (x, y, z, w) with left-to-right row sums
__device__ __forceinline__ void xform(const float *m, float x, float y, float z, float w, float *o)
{
    o[0] = m[0] * x + m[4] * y + m[8] * z + m[12] * w;
    o[1] = m[1] * x + m[5] * y + m[9] * z + m[13] * w;
    o[2] = m[2] * x + m[6] * y + m[10] * z + m[14] * w;
    o[3] = m[3] * x + m[7] * y + m[11] * z + m[15] * w;
}

// compute_lighting (reference src/lighting.h:53-142) before its final clamp.
// OVERRIDE (differentiable shading): the lights' diffuse rgb come from
// dif_rgb (3 per light) and each light's att * N.L where it adds diffuse
// light (else 0) is stored in ndl_out.
// OVERRIDE: light i's diffuse rgb is dif_rgb[3i..3i+2] and ndl_out[i]
// receives att * N.L (0 where the light does not reach); ld_pre, if given,
// holds the normalised directions of the directional lights (normalize3 of
// their positions, the same arithmetic hoisted out of a per-vertex loop)
template <bool OVERRIDE>
__device__ Col lighting_sum(const mtgl_lighting *L, uint32_t flags, const float *eye_pos, const float *eye_normal,
                            Col m_amb, Col m_dif, Col m_spec, Col m_emi, float shininess, const float *dif_rgb,
                            float *ndl_out, const float *ld_pre = nullptr)
{
    Col result = col_add(m_emi, col_mul(m_amb, col_of(L->model_ambient)));
    result.a = m_dif.a;
    float N[3] = {eye_normal[0], eye_normal[1], eye_normal[2]};
    if (flags & MTGL_F_NORMALIZE) normalize3(N);
#pragma unroll 1
    for (int i = 0; i < MTGL_MAX_LIGHTS; i++) {
        const mtgl_light *lt = &L->lights[i];
        if (OVERRIDE) ndl_out[i] = 0.0f;
        if (!lt->enabled) continue;
        float Ld[3];
        float att = 1.0f;
        if (lt->position[3] == 0.0f) {
            if (ld_pre) {
                Ld[0] = ld_pre[3 * i]; Ld[1] = ld_pre[3 * i + 1]; Ld[2] = ld_pre[3 * i + 2];
            } else {
                Ld[0] = lt->position[0]; Ld[1] = lt->position[1]; Ld[2] = lt->position[2];
                normalize3(Ld);
            }
        } else {
            float to[3] = {lt->position[0] - eye_pos[0], lt->position[1] - eye_pos[1], lt->position[2] - eye_pos[2]};
            float dist = sqrtf(dot3(to, to));
            if (dist < 1e-6f) dist = 1e-6f;
            float s = 1.0f / dist;
            Ld[0] = to[0] * s; Ld[1] = to[1] * s; Ld[2] = to[2] * s;
            float den = lt->constant_attenuation + lt->linear_attenuation * dist + lt->quadratic_attenuation * dist * dist;
            if (den < 1e-6f) den = 1e-6f;
            att = 1.0f / den;
            if (lt->spot_cutoff < 180.0f) {
                float sd[3] = {lt->spot_direction[0], lt->spot_direction[1], lt->spot_direction[2]};
                normalize3(sd);
                float cos_angle = -dot3(Ld, sd);
                if (cos_angle < lt->spot_cos_cutoff) att = 0.0f;
                else att *= pow_cr(cos_angle, lt->spot_exponent);
            }
        }
        if (att <= 0.0f) continue;
        result = col_add(result, col_scale(col_mul(m_amb, col_of(lt->ambient)), att));
        float ndl = dot3(N, Ld);
        if (ndl > 0.0f) {
            Col ld = col_of(lt->diffuse);
            if (OVERRIDE) {
                ld.r = dif_rgb[3 * i];
                ld.g = dif_rgb[3 * i + 1];
                ld.b = dif_rgb[3 * i + 2];
                ndl_out[i] = att * ndl;
            }
            result = col_add(result, col_scale(col_mul(m_dif, ld), att * ndl));
            if (shininess > 0.0f) {
                float V[3] = {0.0f, 0.0f, 1.0f};
                if (L->local_viewer) {
                    V[0] = eye_pos[0] * -1.0f; V[1] = eye_pos[1] * -1.0f; V[2] = eye_pos[2] * -1.0f;
                    normalize3(V);
                }
                float H[3] = {Ld[0] + V[0], Ld[1] + V[1], Ld[2] + V[2]};
                normalize3(H);
                float ndh = dot3(N, H);
                if (ndh > 0.0f) {
                    float spec = pow_cr(ndh, shininess) * att;
                    result = col_add(result, col_scale(col_mul(m_spec, col_of(lt->specular)), spec));
                }
            }
        }
    }
    return result;
}

__device__ Col compute_lighting(const mtgl_lighting *L, uint32_t flags, const float *eye_pos, const float *eye_normal,
                                Col m_amb, Col m_dif, Col m_spec, Col m_emi, float shininess)
{
    return col_clamp(lighting_sum<false>(L, flags, eye_pos, eye_normal, m_amb, m_dif, m_spec, m_emi, shininess,
                                         nullptr, nullptr));
}

// ---------------------------------------------------------------------------
// K1: vertex fetch, transform and lighting
// ---------------------------------------------------------------------------

__device__ __forceinline__ float ld_f32(const uint8_t *p)
{
    if (((uintptr_t)p & 3u) == 0) return *(const float *)p;
    float v;
    __builtin_memcpy(&v, p, 4);
    return v;
}

// reference get_array_element (src/gl_api.c:1758-1797); elements outside the
// stream's source take the defaults
__device__ void fetch(const DArr &a, int32_t index, float *out, int n)
{
    int64_t elem = (int64_t)a.comps * (a.type == GL_FLOAT ? 4 : 1);
    int64_t at = a.offset + (int64_t)index * a.stride;
    if (!a.src || index < 0 || a.stride <= 0 || at < 0 || at + elem > a.size) {
        for (int i = 0; i < n; i++) out[i] = i < 3 ? 0.0f : 1.0f;
        return;
    }
    const uint8_t *p = a.src + at;
    for (int i = 0; i < n; i++) {
        if (i < a.comps) out[i] = a.type == GL_FLOAT ? ld_f32(p + 4 * i) : (float)p[i] / 255.0f;
        else out[i] = i == 3 ? 1.0f : 0.0f;
    }
}

// The scene of a launch that owns item x: the last s with pre[s] <= x (pre
// holds the n + 1 running starts, pre[0] = 0).
__device__ __forceinline__ uint32_t find_scene(const uint32_t *pre, uint32_t n, uint32_t x)
{
    uint32_t lo = 0, hi = n;   // pre[lo] <= x < pre[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

// Scene of block, tile or item x when a launch has several scenes: a lookup
// in a map the launch built once (k_scene_maps), one scalar load instead of a
// chain of dependent ones.
__device__ __forceinline__ uint32_t scene_of(const uint32_t *map, uint32_t nscenes, uint32_t x)
{
    return nscenes == 1u ? 0u : __builtin_amdgcn_readfirstlane(map[x]);
}

// The three maps of a multi-scene launch, one block per scene: vertex block,
// setup block and tile -> scene.
__global__ void __launch_bounds__(256) k_scene_maps(const uint32_t *vblock_pre, const uint32_t *sblock_pre,
                                                    const uint32_t *tile_pre, uint32_t *vmap, uint32_t *smap,
                                                    uint32_t *tmap)
{
    const uint32_t s = blockIdx.x;
    for (uint32_t i = vblock_pre[s] + threadIdx.x; i < vblock_pre[s + 1]; i += blockDim.x) vmap[i] = s;
    for (uint32_t i = sblock_pre[s] + threadIdx.x; i < sblock_pre[s + 1]; i += blockDim.x) smap[i] = s;
    for (uint32_t i = tile_pre[s] + threadIdx.x; i < tile_pre[s + 1]; i += blockDim.x) tmap[i] = s;
}

__device__ int32_t fetch_index_of(const DArr &a, uint32_t index_type, int32_t first, uint32_t i)
{
    if (index_type == 0) return (int32_t)((uint32_t)first + i);
    int64_t isz = index_type == GL_UNSIGNED_INT ? 4 : (index_type == GL_UNSIGNED_SHORT ? 2 : 1);
    int64_t at = a.offset + (int64_t)i * isz;
    if (!a.src || at < 0 || at + isz > a.size) return -1;
    const uint8_t *p = a.src + at;
    if (isz == 4) {
        uint32_t v;
        if (((uintptr_t)p & 3u) == 0) v = *(const uint32_t *)p; else __builtin_memcpy(&v, p, 4);
        return (int32_t)v;
    }
    if (isz == 2) {
        uint16_t v;
        if (((uintptr_t)p & 1u) == 0) v = *(const uint16_t *)p; else __builtin_memcpy(&v, p, 2);
        return (int32_t)v;
    }
    return (int32_t)p[0];
}

__device__ __forceinline__ int32_t fetch_index(const VertexParams &P, uint32_t i)
{
    return fetch_index_of(P.indices, P.index_type, P.first, i);
}

__device__ __forceinline__ float sanitize(float v, float nan_value)
{
    if (__builtin_isnan(v) || __builtin_isinf(v)) v = nan_value;
    return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
}

// Inclusive prefix sum over the 64 lanes of a wave with DPP row shifts and
// row broadcasts (VALU only): Hillis-Steele within each row of 16 lanes,
// then row 0's and row 1's totals carried into the rows above.
__device__ __forceinline__ uint32_t dpp_incl_scan(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);   // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);   // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xe, false);  // row_shr:4, lanes 4..15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xc, false);  // row_shr:8, lanes 8..15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
    return v;
}

// Inclusive prefix minimum (MAX false) or maximum over the 64 lanes, the
// same DPP pattern as dpp_incl_scan; lanes without a source keep the
// identity, and NaN operands are passed over (v_min / v_max return the other).
template <bool MAX>
__device__ __forceinline__ float dpp_incl_minmax(float v)
{
    const int id = __float_as_int(MAX ? -__builtin_inff() : __builtin_inff());
#define MINMAX_STEP(ctrl, rm, bm)                                                                          \
    {                                                                                                      \
        const float o = __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), ctrl, rm, bm, false)); \
        v = MAX ? fmaxf(v, o) : fminf(v, o);                                                              \
    }
    MINMAX_STEP(0x111, 0xf, 0xf)   // row_shr:1
    MINMAX_STEP(0x112, 0xf, 0xf)   // row_shr:2
    MINMAX_STEP(0x114, 0xf, 0xe)   // row_shr:4, lanes 4..15
    MINMAX_STEP(0x118, 0xf, 0xc)   // row_shr:8, lanes 8..15
    MINMAX_STEP(0x142, 0xa, 0xf)   // row_bcast:15 into rows 1, 3
    MINMAX_STEP(0x143, 0xc, 0xf)   // row_bcast:31 into rows 2, 3
#undef MINMAX_STEP
    return v;
}

__device__ __forceinline__ uint32_t dpp_wave_sum(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)dpp_incl_scan(v), 63);
}

// inclusive prefix sum over the 64 lanes of a wave, any arithmetic type
template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    return v;
}

// 32-bit values: DPP row shifts and broadcasts, no LDS permutes
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) { return dpp_incl_scan(v); }

// lane 63's value of v, to every lane
template <typename T>
__device__ __forceinline__ T wave_last(T v) { return __shfl(v, 63); }
__device__ __forceinline__ uint32_t wave_last(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// clip-space vertex carried through clipping (the object-space normal the
// reference also interpolates is never read after clipping and is dropped)
struct CV {
    float x, y, z, w;
    float r, g, b, a;
    float s, t, ez;
    float ep[3], en[3];
};

// emit_vertex (reference src/gl_api.c:263-348) of stream position i, whose
// array element is idx (-1: an element no array holds, which reads each
// array's defaults): attribute fetch with get_array_element's rules
// (src/gl_api.c:1758-1797), glColor4f sanitising, transform, Gouraud
// lighting, texture matrix.  ep / en are written when cold is set.
// vs0 is the draw's first vertex state; MULTI draws may switch states per
// vertex through state_index.
template <bool MULTI>
__device__ void vertex_at(const VertexParams &P, uint32_t i, int32_t idx, bool cold, CV &o,
                          const mtgl_vertex_state *vs0)
{
    float pos[4], col[4], tc[2], nrm[3];
    fetch(P.position, idx, pos, 4);
    if (P.position.comps == 2) pos[2] = 0.0f;
    if (P.color.enabled) {
        fetch(P.color, idx, col, 4);
        for (int k = 0; k < 4; k++) col[k] = sanitize(col[k], k == 3 ? 1.0f : 0.0f);
    } else {
        for (int k = 0; k < 4; k++) col[k] = P.cur_color[k];
    }
    if (P.texcoord.enabled) fetch(P.texcoord, idx, tc, 2);
    else { tc[0] = P.cur_texcoord[0]; tc[1] = P.cur_texcoord[1]; }
    if (P.normal.enabled) fetch(P.normal, idx, nrm, 3);
    else { nrm[0] = P.cur_normal[0]; nrm[1] = P.cur_normal[1]; nrm[2] = P.cur_normal[2]; }

    const mtgl_vertex_state *vs = vs0;
    if (MULTI && P.n_vstates > 1 && P.state_index.enabled) {
        uint32_t si = 0;
        const DArr &a = P.state_index;
        int64_t at = a.offset + (int64_t)i * a.stride;
        if (a.src && at >= 0 && at + 4 <= a.size) __builtin_memcpy(&si, a.src + at, 4);
        if (si >= P.n_vstates) si = P.n_vstates - 1;
        vs = &P.vstates[si];
    }

    float eye[4];
    xform(vs->modelview, pos[0], pos[1], pos[2], 1.0f, eye);
    float en4[4];
    xform(vs->normal_matrix, nrm[0], nrm[1], nrm[2], 0.0f, en4);
    float en[3] = {en4[0], en4[1], en4[2]};
    normalize3(en);
    float ep[3] = {eye[0], eye[1], eye[2]};

    Col vc = {col[0], col[1], col[2], col[3]};
    if ((vs->flags & MTGL_F_LIGHTING) && vs->shade_model != GL_PHONG) {
        const mtgl_material &f = vs->lighting.front;
        Col amb = col_of(f.ambient), dif = col_of(f.diffuse), spe = col_of(f.specular), emi = col_of(f.emission);
        uint32_t face = vs->color_material_face, mode = vs->color_material_mode;
        if ((vs->flags & MTGL_F_COLOR_MATERIAL) && (face == GL_FRONT || face == GL_FRONT_AND_BACK)) {
            Col k = col_clamp(vc);
            if (mode == GL_AMBIENT || mode == GL_AMBIENT_AND_DIFFUSE) amb = k;
            if (mode == GL_DIFFUSE || mode == GL_AMBIENT_AND_DIFFUSE) dif = k;
            if (mode == GL_SPECULAR) spe = k;
            if (mode == GL_EMISSION) emi = k;
        }
        vc = compute_lighting(&vs->lighting, vs->flags, ep, en, amb, dif, spe, emi, f.shininess);
    }

    float clip[4];
    xform(vs->projection, eye[0], eye[1], eye[2], eye[3], clip);

    float t4[4];
    xform(vs->texture, tc[0], tc[1], 0.0f, 1.0f, t4);
    float s_out, t_out;
    if (t4[3] != 0.0f && t4[3] != 1.0f) { s_out = t4[0] / t4[3]; t_out = t4[1] / t4[3]; }
    else { s_out = t4[0]; t_out = t4[1]; }

    o.x = clip[0]; o.y = clip[1]; o.z = clip[2]; o.w = clip[3];
    o.r = vc.r; o.g = vc.g; o.b = vc.b; o.a = vc.a;
    o.s = s_out; o.t = t_out; o.ez = -eye[2];
    for (int k = 0; k < 3; k++) {
        o.ep[k] = cold ? ep[k] : 0.0f;
        o.en[k] = cold ? en[k] : 0.0f;
    }
}

// One thread per vertex record of a point / line draw (the tile pipeline
// transforms its vertices inside set-up): records for prim_load.
__global__ void __launch_bounds__(256) k_vertex(const VertexParams *Ps, const uint32_t *block_pre,
                                                const uint32_t *block_scene, uint32_t nscenes)
{
    const uint32_t sc = scene_of(block_scene, nscenes, blockIdx.x);
    const VertexParams &P = Ps[sc];
    const uint32_t i = (blockIdx.x - block_pre[sc]) * blockDim.x + threadIdx.x;
    if (i >= P.count) return;
    // a slot past the addressed span stands for every index no array reaches:
    // -1 reads the defaults such an index reads
    const int32_t idx = P.unique ? (i + 1u < P.count ? P.ubase + (int32_t)i : -1) : fetch_index(P, i);
    CV o;
    vertex_at<true>(P, i, idx, P.out_cold != nullptr, o, P.vstates);
    VHot h;
    h.x = o.x; h.y = o.y; h.z = o.z; h.w = o.w;
    h.r = o.r; h.g = o.g; h.b = o.b; h.a = o.a;
    h.s = o.s; h.t = o.t; h.ez = o.ez; h.pad = 0.0f;
    P.out_hot[i] = h;
    if (P.out_cold) {
        VCold c;
        c.ep[0] = o.ep[0]; c.ep[1] = o.ep[1]; c.ep[2] = o.ep[2];
        c.en[0] = o.en[0]; c.en[1] = o.en[1]; c.en[2] = o.en[2];
        c.pad[0] = c.pad[1] = 0.0f;
        P.out_cold[i] = c;
    }
}

// ---------------------------------------------------------------------------
// K2: primitive assembly, clipping, set-up
// ---------------------------------------------------------------------------

// triangles of a primitive (reference flush_*, src/raster.c:961-1017, :1199-1231)
__host__ __device__ inline uint32_t tri_count(uint32_t mode, uint32_t n)
{
    switch (mode) {
    case GL_TRIANGLES: return n / 3;
    case GL_TRIANGLE_STRIP: case GL_TRIANGLE_FAN: case GL_POLYGON: return n >= 3 ? n - 2 : 0;
    case GL_QUADS: return (n / 4) * 2;
    case GL_QUAD_STRIP: return n >= 4 ? ((n - 4) / 2 + 1) * 2 : 0;
    default: return 0;
    }
}

__device__ __forceinline__ void tri_vertices(uint32_t mode, uint32_t t, uint32_t &a, uint32_t &b, uint32_t &c)
{
    switch (mode) {
    case GL_TRIANGLES: a = 3 * t; b = a + 1; c = a + 2; break;
    case GL_TRIANGLE_STRIP:
        if (t % 2 == 0) { a = t; b = t + 1; } else { a = t + 1; b = t; }
        c = t + 2;
        break;
    case GL_QUADS: {
        uint32_t q = 4 * (t / 2);
        if (t % 2 == 0) { a = q; b = q + 1; c = q + 2; } else { a = q; b = q + 2; c = q + 3; }
        break;
    }
    case GL_QUAD_STRIP: {
        uint32_t q = 2 * (t / 2);
        if (t % 2 == 0) { a = q; b = q + 1; c = q + 3; } else { a = q; b = q + 3; c = q + 2; }
        break;
    }
    default: a = 0; b = t + 1; c = t + 2; break;   // fan, polygon
    }
}

__device__ __forceinline__ float plane_dist(const CV &v, int plane)
{
    switch (plane) {
    case 0: return v.z + v.w;
    case 1: return v.w - v.z;
    case 2: return v.x + v.w;
    case 3: return v.w - v.x;
    case 4: return v.y + v.w;
    default: return v.w - v.y;
    }
}

__device__ __forceinline__ float lerpf(float a, float b, float t) { return a + t * (b - a); }

// vertex_lerp (reference src/graphics.h:465-475) followed by snap_to_plane
// (src/clipping.h:37-47)
__device__ void cv_lerp_snap(const CV &a, const CV &b, float t, int plane, bool cold, CV &o)
{
    o.x = lerpf(a.x, b.x, t); o.y = lerpf(a.y, b.y, t); o.z = lerpf(a.z, b.z, t); o.w = lerpf(a.w, b.w, t);
    o.r = a.r * (1.0f - t) + b.r * t;
    o.g = a.g * (1.0f - t) + b.g * t;
    o.b = a.b * (1.0f - t) + b.b * t;
    o.a = a.a * (1.0f - t) + b.a * t;
    o.s = lerpf(a.s, b.s, t); o.t = lerpf(a.t, b.t, t);
    o.ez = lerpf(a.ez, b.ez, t);
    if (cold) {
        for (int k = 0; k < 3; k++) { o.ep[k] = lerpf(a.ep[k], b.ep[k], t); o.en[k] = lerpf(a.en[k], b.en[k], t); }
    }
    switch (plane) {
    case 0: o.z = -o.w; break;
    case 1: o.z = o.w; break;
    case 2: o.x = -o.w; break;
    case 3: o.x = o.w; break;
    case 4: o.y = -o.w; break;
    default: o.y = o.w; break;
    }
}

// Sutherland-Hodgman against one plane (reference clip_polygon_plane_id,
// src/clipping.h:50-99)
__device__ int clip_plane(const CV *in, int n, CV *out, int plane, bool cold)
{
    if (n == 0) return 0;
    int m = 0;
    int pi = n - 1;
    float pd = plane_dist(in[pi], plane);
    for (int i = 0; i < n; i++) {
        float cd = plane_dist(in[i], plane);
        if (pd >= 0) {
            if (cd >= 0) {
                if (m < MAX_CLIP) out[m++] = in[i];
            } else {
                float den = pd - cd;
                if (fabsf(den) > 1e-10f && m < MAX_CLIP) { cv_lerp_snap(in[pi], in[i], pd / den, plane, cold, out[m]); m++; }
            }
        } else if (cd >= 0) {
            float den = pd - cd;
            if (fabsf(den) > 1e-10f && m < MAX_CLIP) { cv_lerp_snap(in[pi], in[i], pd / den, plane, cold, out[m]); m++; }
            if (m < MAX_CLIP) out[m++] = in[i];
        }
        pi = i;
        pd = cd;
    }
    return m;
}

// the transformed vertex of stream position i (set-up recomputes the
// vertices a triangle shares with its neighbours: the transform is a few
// hundred operations, cheaper than a launch and a round trip through HBM)
__device__ __forceinline__ void load_cv(const SetupParams &P, uint32_t i, bool cold, CV &o)
{
    const VertexParams &V = *P.vp;
    vertex_at<true>(V, i, fetch_index(V, i), cold, o, V.vstates);
}

__device__ __forceinline__ void perspective_divide(CV &v)
{
    if (fabsf(v.w) < 1e-6f) { v.x = 0.0f; v.y = 0.0f; v.z = 0.0f; v.w = 1.0f; return; }
    float iw = 1.0f / v.w;
    v.x *= iw; v.y *= iw; v.z *= iw; v.w = iw;
}

// one fan triangle (c[0], c[j], c[j+1]) after the divide; returns 1 when a
// record was written
// The rasteriser set-up of one filled fan triangle (src/raster.c:458-529):
// the box clamped to viewport and scissor, the area, the drop below half a
// pixel and the per-triangle LOD, into record r; box gets the box clamped to
// the framebuffer as well (pixels outside it have no observable effect).
// Returns 0 when the triangle produces no fragment.
__device__ int fan_record(const mtgl_raster_state *rs, int32_t fb_w, int32_t fb_h, int32_t tex_w, int32_t tex_h,
                          const CV &A, const CV &B, const CV &C, const int32_t *X, const int32_t *Y, int back, RTri &r,
                          int32_t *box)
{
    const int32_t vx = rs->viewport[0], vy = rs->viewport[1], vw = rs->viewport[2], vh = rs->viewport[3];
    int32_t minX = X[0] < X[1] ? (X[0] < X[2] ? X[0] : X[2]) : (X[1] < X[2] ? X[1] : X[2]);
    int32_t minY = Y[0] < Y[1] ? (Y[0] < Y[2] ? Y[0] : Y[2]) : (Y[1] < Y[2] ? Y[1] : Y[2]);
    int32_t maxX = X[0] > X[1] ? (X[0] > X[2] ? X[0] : X[2]) : (X[1] > X[2] ? X[1] : X[2]);
    int32_t maxY = Y[0] > Y[1] ? (Y[0] > Y[2] ? Y[0] : Y[2]) : (Y[1] > Y[2] ? Y[1] : Y[2]);
    if (minX < vx) minX = vx;
    if (minY < vy) minY = vy;
    if (maxX >= vx + vw) maxX = vx + vw - 1;
    if (maxY >= vy + vh) maxY = vy + vh - 1;
    if (rs->flags & MTGL_F_SCISSOR_TEST) {
        const int32_t sx = rs->scissor[0], sy = rs->scissor[1], sw = rs->scissor[2], sh = rs->scissor[3];
        if (minX < sx) minX = sx;
        if (minY < sy) minY = sy;
        if (maxX >= sx + sw) maxX = sx + sw - 1;
        if (maxY >= sy + sh) maxY = sy + sh - 1;
    }
    if (minX > maxX || minY > maxY) return 0;
    const float fx0 = (float)X[0], fy0 = (float)Y[0], fx1 = (float)X[1], fy1 = (float)Y[1];
    const float fx2 = (float)X[2], fy2 = (float)Y[2];
    float area = (fx2 - fx0) * (fy1 - fy0) - (fy2 - fy0) * (fx1 - fx0);
    if (fabsf(area) < 0.5f) return 0;
    box[0] = minX < 0 ? 0 : minX;
    box[1] = minY < 0 ? 0 : minY;
    box[2] = maxX >= fb_w ? fb_w - 1 : maxX;
    box[3] = maxY >= fb_h ? fb_h - 1 : maxY;
    if (box[0] > box[2] || box[1] > box[3]) return 0;

    float lod = 0.0f;
    if (tex_w > 0) {
        float screen_area = fabsf(area) * 0.5f;
        float du1 = (B.s - A.s) * (float)tex_w;
        float dv1 = (B.t - A.t) * (float)tex_h;
        float du2 = (C.s - A.s) * (float)tex_w;
        float dv2 = (C.t - A.t) * (float)tex_h;
        float texel_area = fabsf(du1 * dv2 - du2 * dv1) * 0.5f;
        if (screen_area > 0.0f) {
            float tpp = texel_area / screen_area;
            if (tpp > 0.0f) {
                lod = log2_cr(tpp) * 0.5f;
                if (lod < 0.0f) lod = 0.0f;
            }
        }
    }
    r.x0 = X[0]; r.y0 = Y[0]; r.x1 = X[1]; r.y1 = Y[1]; r.x2 = X[2]; r.y2 = Y[2];
    r.minX = minX; r.minY = minY; r.maxX = maxX; r.maxY = maxY;
    r.area = area;
    r.inv_area = 1.0f / area;
    r.z0 = A.z; r.z1 = B.z; r.z2 = C.z;
    r.w0 = A.w; r.w1 = B.w; r.w2 = C.w;
    r.c0[0] = A.r; r.c0[1] = A.g; r.c0[2] = A.b; r.c0[3] = A.a;
    r.c1[0] = B.r; r.c1[1] = B.g; r.c1[2] = B.b; r.c1[3] = B.a;
    r.c2[0] = C.r; r.c2[1] = C.g; r.c2[2] = C.b; r.c2[3] = C.a;
    r.uv0[0] = A.s; r.uv0[1] = A.t; r.uv1[0] = B.s; r.uv1[1] = B.t; r.uv2[0] = C.s; r.uv2[1] = C.t;
    r.ez0 = A.ez; r.ez1 = B.ez; r.ez2 = C.ez;
    r.lod = lod;
    r.back = back;
    r.pad0[0] = r.pad0[1] = r.pad0[2] = 0;
    return 1;
}

__device__ __forceinline__ void cold_record(const CV &A, const CV &B, const CV &C, RCold &q)
{
    for (int k = 0; k < 3; k++) {
        q.ep[k] = A.ep[k]; q.ep[3 + k] = B.ep[k]; q.ep[6 + k] = C.ep[k];
        q.en[k] = A.en[k]; q.en[3 + k] = B.en[k]; q.en[6 + k] = C.en[k];
    }
    q.pad[0] = q.pad[1] = 0.0f;
}

// The tiles of a record's box (tw x th tiles from (tx0, ty0), at most 64)
// that can hold a covered pixel: bit k for tile k, row-major.  The raster
// covers a pixel when, for each edge, A(x) = fl((x - fx) * dy) >= B(y) =
// fl((y - fy) * dx), formed exactly as in cover_chunk (the sign of the area
// folded into the deltas), within the record's pixel box and the
// framebuffer.  A is monotone in x and B in y (a correctly rounded product of
// an exact difference), so a tile holds a pixel passing edge k exactly when
// A at the tile's column end that maximises it is >= B at the row end that
// minimises it; a tile failing one edge this way holds no covered pixel.
__device__ __forceinline__ unsigned long long box_tile_mask(const int32_t *X, const int32_t *Y, float area,
                                                            int32_t minX, int32_t minY, int32_t maxX, int32_t maxY,
                                                            int32_t tx0, int32_t ty0, int32_t tw, int32_t th,
                                                            int32_t fb_w, int32_t fb_h, int32_t lo = 0,
                                                            int32_t hi = 64)
{
    const float fx0 = (float)X[0], fy0 = (float)Y[0], fx1 = (float)X[1], fy1 = (float)Y[1];
    const float fx2 = (float)X[2], fy2 = (float)Y[2];
    const float sg = area > 0.0f ? 1.0f : -1.0f;
    const float ex[3] = {fx1, fx2, fx0}, ey[3] = {fy1, fy2, fy0};
    const float dy[3] = {sg * (fy2 - fy1), sg * (fy0 - fy2), sg * (fy1 - fy0)};
    const float dx[3] = {sg * (fx2 - fx1), sg * (fx0 - fx2), sg * (fx1 - fx0)};
    const int32_t xl = minX, xh = min(maxX, fb_w - 1), yl = minY, yh = min(maxY, fb_h - 1);
    // A of tile column i (its end maximising it), B of tile row j (its end
    // minimising it), per edge
    auto col_a = [&](int32_t i, float *a) {
        const int32_t tx = tx0 + i;
        const float xa = (float)max(tx * TILE, xl), xb = (float)min(tx * TILE + TILE - 1, xh);
#pragma unroll
        for (int e = 0; e < 3; e++) a[e] = ((dy[e] >= 0.0f ? xb : xa) - ex[e]) * dy[e];
    };
    auto row_b = [&](int32_t j, float *b) {
        const int32_t ty = ty0 + j;
        const float ya = (float)max(ty * TILE, yl), yb = (float)min(ty * TILE + TILE - 1, yh);
#pragma unroll
        for (int e = 0; e < 3; e++) b[e] = ((dx[e] >= 0.0f ? ya : yb) - ey[e]) * dx[e];
    };
    // the box holds at most 64 tiles, so its shorter side at most 8: those
    // terms are formed once, the longer side's once per step (the steps
    // [lo, hi) of it)
    unsigned long long m = 0ull;
    if (th <= tw) {
        float b[8][3];
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (j < th) row_b(j, b[j]);
        for (int32_t i = lo; i < min(tw, hi); i++) {
            float a[3];
            col_a(i, a);
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const bool in = j < th && (a[0] >= b[j][0]) & (a[1] >= b[j][1]) & (a[2] >= b[j][2]);
                m |= in ? 1ull << (uint32_t)(j * tw + i) : 0ull;   // (j * tw + i < 64 when j < th)
            }
        }
    } else {
        float a[8][3];
#pragma unroll
        for (int i = 0; i < 8; i++)
            if (i < tw) col_a(i, a[i]);
        for (int32_t j = lo; j < min(th, hi); j++) {
            float b[3];
            row_b(j, b);
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const bool in = i < tw && (a[i][0] >= b[0]) & (a[i][1] >= b[1]) & (a[i][2] >= b[2]);
                m |= in ? 1ull << (uint32_t)(j * tw + i) : 0ull;
            }
        }
    }
    return m;
}

// What box_tile_mask needs of one record, handed from the set-up wave to
// the waves that form the mask when set-up defers it (want: a box of at
// most BOX_MASK_TILES tiles)
struct MaskIn {
    int32_t X[3], Y[3];
    float area;
    int32_t minX, minY, maxX, maxY;
    int32_t tx0, ty0, tw, th;
    int32_t want;
};

__device__ int setup_fan(const SetupParams &P, const mtgl_raster_state *rs, const CV &A, const CV &B, const CV &C,
                         bool cold, RTri *outp, RCold *coldp, RBox *boxp, int *tiles_out, MaskIn *defer = nullptr)
{
    const int32_t vx = rs->viewport[0], vy = rs->viewport[1], vw = rs->viewport[2], vh = rs->viewport[3];
    const CV *v[3] = {&A, &B, &C};
    int32_t X[3], Y[3];
    for (int k = 0; k < 3; k++) {
        X[k] = cvtt((v[k]->x + 1.0f) * 0.5f * (float)vw + (float)vx);
        Y[k] = cvtt((1.0f - v[k]->y) * 0.5f * (float)vh + (float)vy);
    }
    float sa = (float)isub(X[1], X[0]) * (float)isub(Y[2], Y[0]) - (float)isub(X[2], X[0]) * (float)isub(Y[1], Y[0]);
    if (rs->flags & MTGL_F_CULL_FACE) {
        int front = rs->front_face == GL_CCW ? (sa < 0) : (sa > 0);
        int cull = rs->cull_face_mode == GL_FRONT ? front : (rs->cull_face_mode == GL_BACK ? !front : 1);
        if (cull) return 0;
    }
    int back = rs->front_face == GL_CCW ? (sa >= 0) : (sa < 0);
    uint32_t pm = back ? rs->polygon_mode_back : rs->polygon_mode_front;
    if (pm == GL_POINT || pm == GL_LINE) { atomicOr(&P.ctr->unsupported, 1u); return 0; }

    RTri r;
    int32_t bb[4];
    if (!fan_record(rs, P.fb_w, P.fb_h, P.tex_w, P.tex_h, A, B, C, X, Y, back, r, bb)) return 0;
    *outp = r;
    const int32_t tx0 = bb[0] / TILE, ty0 = bb[1] / TILE, tx1 = bb[2] / TILE, ty1 = bb[3] / TILE;
    RBox bx;
    bx.key0 = P.tile_off + (uint32_t)(ty0 * P.tiles_x + tx0);
    bx.tw = (uint32_t)(tx1 - tx0 + 1);
    bx.ntiles = bx.tw * (uint32_t)(ty1 - ty0 + 1);
    bx.tiles_x = (uint32_t)P.tiles_x;
    bx.mask = 0ull;
    bx.masked = 0u;
    bx.pad = 0u;
    if (defer) {
        // the caller has the mask formed elsewhere and completes the box
        MaskIn &d = *defer;
        for (int k = 0; k < 3; k++) {
            d.X[k] = X[k];
            d.Y[k] = Y[k];
        }
        d.area = r.area;
        d.minX = r.minX;
        d.minY = r.minY;
        d.maxX = r.maxX;
        d.maxY = r.maxY;
        d.tx0 = tx0;
        d.ty0 = ty0;
        d.tw = (int32_t)bx.tw;
        d.th = ty1 - ty0 + 1;
        d.want = bx.ntiles <= BOX_MASK_TILES ? 1 : 0;
    } else if (bx.ntiles <= BOX_MASK_TILES) {
        // the tiles of a small box the record can cover: only those get pairs
        const unsigned long long m = box_tile_mask(X, Y, r.area, r.minX, r.minY, r.maxX, r.maxY, tx0, ty0,
                                                   (int32_t)bx.tw, ty1 - ty0 + 1, P.fb_w, P.fb_h);
        bx.mask = m;
        bx.masked = 1u;
        bx.ntiles = (uint32_t)__popcll(m);
    }
    *boxp = bx;
    if (cold) {
        RCold q;
        cold_record(A, B, C, q);
        *coldp = q;
    }
    *tiles_out = (int)bx.ntiles;
    return 1;
}

// slow path: Sutherland-Hodgman in the reference's plane order (near, far,
// left, right, bottom, top; src/clipping.h:106-126), divide, fan split of
// input triangle t into slots of the shared overflow pool.  Out of line: a
// few triangles of a draw take it, and its vertex arrays live in scratch.
// Returns the fan triangles kept; base gets the first slot, tiles their
// (tile, record) pairs.
__device__ __noinline__ uint32_t setup_clipped(const SetupParams &P, const mtgl_raster_state *rs, const CV &A,
                                               const CV &B, const CV &C, bool cold, uint32_t &base,
                                               unsigned long long &tiles)
{
    CV a[MAX_CLIP], b[MAX_CLIP];
    b[0] = A;
    b[1] = B;
    b[2] = C;
    int n = clip_plane(b, 3, a, 0, cold);
    n = n ? clip_plane(a, n, b, 1, cold) : 0;
    n = n ? clip_plane(b, n, a, 2, cold) : 0;
    n = n ? clip_plane(a, n, b, 3, cold) : 0;
    n = n ? clip_plane(b, n, a, 4, cold) : 0;
    n = n ? clip_plane(a, n, b, 5, cold) : 0;
    base = 0;
    tiles = 0;
    if (n < 3) return 0;
    base = atomicAdd(&P.ctr->noverflow, (uint32_t)(n - 2));
    if (base + (uint32_t)(n - 2) > P.overflow_cap) {
        atomicOr(&P.ctr->overflow_full, 1u);
        return 0;
    }
    uint32_t nout = 0;
    RTri *out = P.rt_over + base;
    RCold *cout = cold ? P.rcold_over + base : nullptr;
    for (int k = 0; k < n; k++) perspective_divide(b[k]);
    for (int j = 1; j + 1 < n; j++) {
        int nt = 0;
        RBox q;
        if (setup_fan(P, rs, b[0], b[j], b[j + 1], cold, out + nout, cout ? cout + nout : nullptr, &q, &nt)) {
            P.box_over[base + nout] = q;
            nout++;
            tiles += (unsigned long long)nt;
        }
    }
    return nout;
}

// The uniform inputs of one set-up block, copied to LDS once: every thread
// reads the matrices, lights and viewport many times over, and from LDS that
// costs no global round trip (the raster state's lighting block is not read
// by set-up and is left out).
struct SetupStage {
    SetupParams sp;
    VertexParams vp;
    mtgl_vertex_state vs;
    mtgl_raster_state rs;
};

__device__ __forceinline__ void stage_words(void *dst, const void *src, uint32_t bytes)
{
    uint32_t *d = static_cast<uint32_t *>(dst);
    const uint32_t *q = static_cast<const uint32_t *>(src);
    for (uint32_t k = threadIdx.x; k < bytes / 4u; k += blockDim.x) d[k] = q[k];
}

// Set-up of every input triangle, one thread each: triangles inside every
// frustum plane pass through the clipper unchanged (each plane re-emits the
// vertices in order), so they are set up directly into slot t; the others go
// through setup_clipped.  Every kept record's tiles are counted into the
// bins.  A block takes SETUP_BLOCK triangles.  With NW = 3 waves (launches
// of few blocks, which leave SIMDs idle) wave c runs the vertex pipeline of
// every triangle's corner c into LDS, three SIMDs at once, then wave 0 sets
// the triangles up; with NW = 1 (launches that fill the GPU) the one wave
// runs all three corners itself.  MULTI draws switch vertex states per
// vertex and read them from HBM.
template <bool MULTI, int NW>
__global__ void __launch_bounds__(NW * SETUP_BLOCK) k_setup(const SetupParams *Ss, const VertexParams *Vs,
                                                            const uint32_t *block_pre, const uint32_t *block_scene,
                                                            uint32_t nscenes, uint32_t *blk_kept,
                                                            unsigned long long *blk_pairs,
                                                            const mtgl_raster_state *rs1,
                                                            const mtgl_vertex_state *vs1)
{
    static_assert(NW == 1 || NW == 3, "one wave, or one per corner");
    __shared__ SetupStage st;
    __shared__ CV s_corner[NW == 3 ? 3 : 1][NW == 3 ? SETUP_BLOCK : 1];
    __shared__ MaskIn s_mi[NW == 3 ? SETUP_BLOCK : 1];                       // records whose mask waves 1, 2 form
    __shared__ unsigned long long s_mpart[NW == 3 ? 2 : 1][NW == 3 ? SETUP_BLOCK : 1];   // ... their halves
    const uint32_t sc = scene_of(block_scene, nscenes, blockIdx.x);
    // the scene's tables and the states they point to: one round of loads
    // when the launch passes the states of its one scene (rs1, vs1), else two
    stage_words(&st.sp, Ss + sc, sizeof(SetupParams));
    stage_words(&st.vp, Vs + sc, sizeof(VertexParams));
    if (rs1) {
        stage_words(&st.rs, rs1, offsetof(mtgl_raster_state, lighting));
        if (!MULTI) stage_words(&st.vs, vs1, sizeof(mtgl_vertex_state));
    }
    __syncthreads();
    if (!rs1) {
        stage_words(&st.rs, st.sp.rs, offsetof(mtgl_raster_state, lighting));
        if (!MULTI) stage_words(&st.vs, st.vp.vstates, sizeof(mtgl_vertex_state));
        __syncthreads();
    }
    const SetupParams &P = st.sp;
    const VertexParams &V = st.vp;
    const mtgl_vertex_state *vs0 = MULTI ? V.vstates : &st.vs;
    // the counters later kernels of the launch accumulate start at zero
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        P.ctr->nheavy = 0u;
        P.ctr->nlight = 0u;
        P.ctr->nbig = 0u;
    }
    const uint32_t l = threadIdx.x % SETUP_BLOCK, corner = threadIdx.x / SETUP_BLOCK;
    const uint32_t t = (blockIdx.x - block_pre[sc]) * SETUP_BLOCK + l;
    const bool act = t < P.ntri;
    const bool cold = P.cold != 0u;
    if (NW == 3) {
        if (act) {
            uint32_t iv[3];
            tri_vertices(P.mode, t, iv[0], iv[1], iv[2]);
            const uint32_t i = corner == 0u ? iv[0] : corner == 1u ? iv[1] : iv[2];
            CV o;
            vertex_at<MULTI>(V, i, fetch_index(V, i), cold, o, vs0);
            s_corner[NW == 3 ? corner : 0u][NW == 3 ? l : 0u] = o;
        }
        __syncthreads();
    }
    const mtgl_raster_state *rs = &st.rs;
    uint32_t nout = 0, rb = 0;
    unsigned long long tiles = 0;
    // NW = 3: the set-up wave leaves the tile masks of its unclipped records
    // to waves 1 and 2 (half of each box's longer side each) and completes
    // those boxes afterwards
    MaskIn mi;
    mi.want = 0;
    RBox bxd;
    if (corner == 0u && act) {
        CV v[3];
        if (NW == 3) {
            for (int k = 0; k < 3; k++) v[k] = s_corner[NW == 3 ? k : 0][NW == 3 ? l : 0u];
        } else {
            // the three indices first, so their loads overlap
            uint32_t ia, ib, ic;
            tri_vertices(P.mode, t, ia, ib, ic);
            const int32_t xa = fetch_index(V, ia), xb = fetch_index(V, ib), xc = fetch_index(V, ic);
            vertex_at<MULTI>(V, ia, xa, cold, v[0], vs0);
            vertex_at<MULTI>(V, ib, xb, cold, v[1], vs0);
            vertex_at<MULTI>(V, ic, xc, cold, v[2], vs0);
        }
        bool all_inside = true;
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int pl = 0; pl < 6; pl++)
                all_inside = all_inside && (plane_dist(v[k], pl) >= 0);
        rb = P.tri_off + t;
        if (all_inside) {
            for (int k = 0; k < 3; k++) perspective_divide(v[k]);
            int nt = 0;
            RBox bx;
            if (setup_fan(P, rs, v[0], v[1], v[2], cold, P.rt + t, cold ? P.rcold + t : nullptr, &bx, &nt,
                          NW == 3 ? &mi : nullptr)) {
                nout = 1;
                tiles = (unsigned long long)nt;
                if (mi.want) bxd = bx;
                else P.box[t] = bx;
            } else {
                mi.want = 0;
            }
        } else {
            uint32_t base;
            nout = setup_clipped(P, rs, v[0], v[1], v[2], cold, base, tiles);
            rb = P.over_base + base;
        }
    }
    if (NW == 3) {
        if (corner == 0u) {
            if (mi.want) s_mi[l] = mi;
            else s_mi[l].want = 0;
        }
        __syncthreads();
        if (corner != 0u) {
            const MaskIn &d = s_mi[l];
            unsigned long long m = 0ull;
            if (d.want) {
                const int32_t half = (max(d.tw, d.th) + 1) / 2;
                const int32_t lo = corner == 1u ? 0 : half, hi = corner == 1u ? half : 64;
                m = box_tile_mask(d.X, d.Y, d.area, d.minX, d.minY, d.maxX, d.maxY, d.tx0, d.ty0, d.tw, d.th,
                                  P.fb_w, P.fb_h, lo, hi);
            }
            s_mpart[corner - 1u][l] = m;
        }
        __syncthreads();
        if (corner != 0u) return;   // (whole waves)
        if (mi.want) {
            const unsigned long long m = s_mpart[0][l] | s_mpart[1][l];
            bxd.mask = m;
            bxd.masked = 1u;
            bxd.ntiles = (uint32_t)__popcll(m);
            P.box[t] = bxd;
            tiles = bxd.ntiles;
        }
    }
    if (act) {
        P.n_rt[t] = nout;
        P.rec_base[t] = rb;
        P.tri_tiles[t] = tiles;
    }
    // the block's kept fan triangles and pairs into its own slots (the pair
    // offsets of the launch are a scan of the block sums)
    const uint32_t kept = (uint32_t)__builtin_amdgcn_readlane((int)dpp_incl_scan(nout), 63);
    const unsigned long long pairs = wave_inclusive_scan(tiles);
    if (threadIdx.x == 63) {
        blk_kept[blockIdx.x] = kept;
        blk_pairs[blockIdx.x] = pairs;
    }
}

// ---------------------------------------------------------------------------
// scans (exclusive), one value per element; out may alias in
// ---------------------------------------------------------------------------


// block-wide exclusive scan of one value per thread (blockDim a multiple of
// 64, at most 1024; lds_waves holds one value per wave); returns the
// exclusive prefix and the block total
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T *lds_waves, T &total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    T inc = wave_inclusive_scan(v);
    if (lane == 63) lds_waves[wave] = inc;
    __syncthreads();
    T base = 0;
    total = 0;
    for (int w = 0; w < nwaves; w++) {
        T s = lds_waves[w];
        if (w < wave) base += s;
        total += s;
    }
    __syncthreads();
    return base + inc - v;
}

template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce(const T *in, uint64_t n, T *block_sums)
{
    __shared__ T waves[SCAN_THREADS / 64];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    T s = 0;
    for (int k = 0; k < SCAN_ITEMS; k++) {
        uint64_t i = base + (uint64_t)threadIdx.x * SCAN_ITEMS + k;
        if (i < n) s += in[i];
    }
    T total;
    block_exclusive_scan<T>(s, waves, total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// single workgroup: exclusive scan of the block sums in chunks, carrying the
// running total; writes the grand total to *total_out
template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_sums(T *sums, uint64_t nblocks, T *total_out)
{
    __shared__ T waves[SCAN_THREADS / 64];
    T carry = 0;
    for (uint64_t base = 0; base < nblocks; base += SCAN_THREADS) {
        uint64_t i = base + threadIdx.x;
        T v = i < nblocks ? sums[i] : (T)0;
        T total;
        T ex = block_exclusive_scan<T>(v, waves, total);
        if (i < nblocks) sums[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(const T *in, uint64_t n, const T *block_offsets, T *out,
                                                             T *clear)
{
    __shared__ T waves[SCAN_THREADS / 64];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    T vals[SCAN_ITEMS];
    T s = 0;
    for (int k = 0; k < SCAN_ITEMS; k++) {
        vals[k] = base + k < n ? in[base + k] : (T)0;
        s += vals[k];
    }
    // the input read for the last time: cleared for its next use when asked
    if (clear)
        for (int k = 0; k < SCAN_ITEMS; k++)
            if (base + k < n) clear[base + k] = (T)0;
    T total;
    T ex = block_exclusive_scan<T>(s, waves, total) + block_offsets[blockIdx.x];
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = ex;
        ex += vals[k];
    }
}

// k_scan_apply for short scans (at most SCAN2_MAX_BLOCKS tiles), with the
// middle launch folded in: each block adds up the sums of the blocks before
// it; the last block writes the grand total to *total_out.
template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply_short(const T *in, uint64_t n, const T *block_sums,
                                                                   T *out, T *total_out, T *clear)
{
    __shared__ T waves[SCAN_THREADS / 64];
    T part = 0;
    for (uint64_t j = threadIdx.x; j < blockIdx.x; j += SCAN_THREADS) part += block_sums[j];
    T offset;
    block_exclusive_scan<T>(part, waves, offset);
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    T vals[SCAN_ITEMS];
    T s = 0;
    for (int k = 0; k < SCAN_ITEMS; k++) {
        vals[k] = base + k < n ? in[base + k] : (T)0;
        s += vals[k];
    }
    if (clear)
        for (int k = 0; k < SCAN_ITEMS; k++)
            if (base + k < n) clear[base + k] = (T)0;
    T total;
    T ex = block_exclusive_scan<T>(s, waves, total) + offset;
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = ex;
        ex += vals[k];
    }
    if (total_out && threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) *total_out = offset + total;
}

// Scans of at most SCAN1_MAX elements in one launch of one workgroup: wave w
// owns the w-th run of `per` (at most SCAN1_ROWS) 64-element rows, loads them
// all at once (coalesced, one round trip), adds them up; the 16 wave totals
// are scanned in LDS and each wave scans its rows carrying its offset.  The
// grand total goes to *total_out.
template <typename T>
__global__ void __launch_bounds__(SCAN1_THREADS) k_scan_single(const T *in, uint64_t n, T *out, T *total_out,
                                                               T *clear)
{
    __shared__ T waves[SCAN1_THREADS / 64];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t rows = (uint32_t)((n + 63) / 64);
    const uint32_t per = (rows + SCAN1_THREADS / 64 - 1) / (SCAN1_THREADS / 64);
    const uint64_t base = (uint64_t)wave * per * 64;
    T v[SCAN1_ROWS<T>];
    T s = 0;
#pragma unroll
    for (uint32_t k = 0; k < SCAN1_ROWS<T>; k++) {
        const uint64_t i = base + (uint64_t)k * 64 + lane;
        v[k] = k < per && i < n ? in[i] : (T)0;
        s += v[k];
    }
    T total;
    T carry = block_exclusive_scan<T>(s, waves, total);
    // the exclusive scan of the lane totals gave lane 0 its wave's offset
    carry = __shfl(carry, 0);
#pragma unroll
    for (uint32_t k = 0; k < SCAN1_ROWS<T>; k++) {
        if (k < per) {   // (uniform)
            const uint64_t i = base + (uint64_t)k * 64 + lane;
            const T inc = wave_inclusive_scan(v[k]);
            if (i < n) {
                out[i] = carry + inc - v[k];
                if (clear) clear[i] = (T)0;
            }
            carry += wave_last(inc);
        }
    }
    if (total_out && threadIdx.x == 0) *total_out = total;
}

// K3a: expand triangles into (tile, fan triangle) pairs, in submission order
// ---------------------------------------------------------------------------


// (tile, record) pairs of fan record j of a triangle: tiles k0 <= k < k1 of
// its tile box (row-major), pair p0 + k onwards; pairs at or beyond cap are
// dropped (the host learns of it through the overflow flag).
__device__ __forceinline__ void expand_record(const RBox &q, uint32_t rec, uint64_t p0, uint32_t k0, uint32_t k1,
                                              uint32_t step, uint64_t cap, uint32_t *keys, uint32_t *vals)
{
    if (q.masked) {
        // the box's tiles with pairs, in box order: one thread takes them
        // all (step 1), or lane k0 of a wave takes tile k0 (step 64)
        const uint32_t tw = q.tw, base = q.key0, pitch = q.tiles_x;
        if (step == 1u) {
            uint64_t p = p0;
            for (unsigned long long m = q.mask; m; m &= m - 1ull, p++) {
                if (p >= cap) break;
                const uint32_t k = (uint32_t)__builtin_ctzll(m);
                keys[p] = base + (k / tw) * pitch + k % tw;
                vals[p] = rec;
            }
        } else if (k0 < 64u && ((q.mask >> k0) & 1ull)) {
            const uint64_t p = p0 + (uint64_t)__popcll(q.mask & ((1ull << k0) - 1ull));
            if (p < cap) {
                keys[p] = base + (k0 / tw) * pitch + k0 % tw;
                vals[p] = rec;
            }
        }
        return;
    }
    const uint32_t tw = q.tw, base = q.key0, pitch = q.tiles_x;
    uint32_t ky = k0 / tw, kx = k0 % tw;
    const uint32_t sy = step / tw, sx = step % tw;
    for (uint32_t k = k0; k < k1; k += step) {
        const uint64_t p = p0 + k;
        if (p >= cap) break;
        keys[p] = base + ky * pitch + kx;
        vals[p] = rec;
        kx += sx;
        ky += sy;
        if (kx >= tw) {
            kx -= tw;
            ky++;
        }
    }
}

// Pairs of every triangle, in (triangle, fan record, tile) order: one wave
// per set-up block, whose pairs start at blk_off (the scan of the set-up
// blocks' pair sums) and follow in triangle order.  A thread writes its
// triangle's pairs when there are at most EXPAND_SMALL, the block's wave
// writes those of a triangle with at most EXPAND_WAVE together, and longer
// triangles are queued for k_expand_big (one wave each) with their first
// pair in pair_base.
// With blk_off null (launches of at most EXPAND_FOLD set-up blocks) each
// block sums the pair counts of the blocks before it itself, and block 0
// sums them all into npairs: no scan launch.  Block 0 also ends the set-up's
// use of the per-launch counters: the clip pool, its overflow bits (kept in
// sticky) and the unsupported flags start the next launch at zero.
constexpr uint32_t EXPAND_FOLD = 2048;
__global__ void __launch_bounds__(SETUP_BLOCK) k_expand(const SetupParams *Ss, const uint32_t *block_pre,
                                                        const uint32_t *block_scene, uint32_t nscenes,
                                                        const unsigned long long *blk_off,
                                                        const unsigned long long *blk_pairs,
                                                        const unsigned long long *tri_tiles,
                                                        unsigned long long *pair_base, const uint32_t *n_rt,
                                                        const uint32_t *rec_base, const RBox *box, uint64_t cap,
                                                        uint32_t *keys, uint32_t *vals, uint32_t *big,
                                                        Counters *ctr, uint32_t seq, const uint32_t *blk_kept,
                                                        uint32_t nsblk, uint32_t ttot)
{
    const int lane = threadIdx.x;
    const uint32_t sc = scene_of(block_scene, nscenes, blockIdx.x);
    // one scene: its triangles are the launch's (no parameter loads)
    const uint32_t ntri = nscenes == 1u ? ttot : Ss[sc].ntri;
    const uint32_t tri_off = nscenes == 1u ? 0u : Ss[sc].tri_off;
    const uint32_t bpre = nscenes == 1u ? 0u : block_pre[sc];
    const uint32_t tl = (blockIdx.x - bpre) * SETUP_BLOCK + (uint32_t)lane;
    const uint32_t t = tri_off + tl;
    const bool act = tl < ntri;
    // the triangle's set-up outputs, loaded with the block sums below (the
    // one record of an unclipped triangle is its own slot t)
    const uint64_t np = act ? tri_tiles[t] : 0ull;
    const uint32_t nr = act ? n_rt[t] : 0u, rb = act ? rec_base[t] : 0u;
    const RBox q0 = act ? box[t] : RBox{0u, 1u, 0u, 0u};
    unsigned long long first = 0, sum_all = 0;
    if (blk_off) {
        first = blk_off[blockIdx.x];
    } else {
        const uint32_t upto = blockIdx.x == 0 ? nsblk : blockIdx.x;
        unsigned long long v = 0;
        for (uint32_t b = lane; b < upto; b += 64) v += blk_pairs[b];
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
        if (blockIdx.x == 0) sum_all = v;
        else first = v;
    }
    if (blockIdx.x == 0) {
        // fan triangles kept by k_setup, one count per set-up block (16 B
        // loads, several in flight: a batch launch has tens of thousands)
        uint32_t kept = 0;
        const uint4 *kv = reinterpret_cast<const uint4 *>(blk_kept);
#pragma unroll 8
        for (uint32_t b = lane; b < nsblk / 4u; b += 64) {
            const uint4 q = kv[b];
            kept += q.x + q.y + q.z + q.w;
        }
        for (uint32_t b = (nsblk & ~3u) + lane; b < nsblk; b += 64) kept += blk_kept[b];
        kept = (uint32_t)__builtin_amdgcn_readlane((int)dpp_incl_scan(kept), 63);
        if (lane == 0) {
            if (!blk_off) ctr->npairs = sum_all;
            // the launch's pair total against the capacity of the pair arrays:
            // a launch that overflowed (pairs, or the clip pool) poisons the
            // Exec with its sequence number, so that its raster stage and
            // every later framebuffer write are skipped until the host has
            // grown the arrays and run it again; a launch that fits adds to
            // the running counts
            // every counter this reads, in one round of loads (only this
            // lane of the launch writes the running counts and sticky)
            const unsigned long long total = blk_off ? ctr->npairs : sum_all;
            const uint32_t ovf = ctr->overflow_full, poison = ctr->poison, sticky = ctr->sticky;
            const unsigned long long acc_tri = ctr->acc_tri_raster, acc_pairs = ctr->acc_bin_pairs;
            const uint32_t over = ovf | (total > cap ? 2u : 0u);
            if (over) {
                if (atomicCAS(&ctr->poison, 0u, seq) == 0u) {
                    ctr->poison_pairs = total;
                    ctr->poison_flags = over;
                }
            } else if (poison == 0u) {   // (a launch after a poisoned one runs again: counted then)
                ctr->acc_tri_raster = acc_tri + kept;
                ctr->acc_bin_pairs = acc_pairs + total;
            }
            ctr->sticky = sticky | over;
            ctr->overflow_full = 0u;
            ctr->noverflow = 0u;
            ctr->unsupported = 0u;
        }
    }
    uint64_t p0 = first + wave_inclusive_scan(np) - np;
    // up to EXPAND_SMALL pairs a thread writes alone; up to EXPAND_WAVE the
    // whole wave writes them after that (a lane per tile); longer triangles
    // are queued for k_expand_big, one queue slot reservation per wave
    const bool coop = np > EXPAND_SMALL && np <= EXPAND_WAVE;
    const bool queued = np > EXPAND_WAVE;
    const uint64_t bal = __ballot(queued);
    if (bal) {
        uint32_t base = 0;
        if (lane == __builtin_ctzll(bal)) base = atomicAdd(&ctr->nbig, (uint32_t)__popcll(bal));
        base = __shfl(base, __builtin_ctzll(bal));
        if (queued) {
            big[base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = t;
            pair_base[t] = p0;
        }
    }
    if (np != 0u && !coop && !queued) {
        uint64_t p = p0;
        for (uint32_t j = 0; j < nr; j++) {
            const RBox q = j == 0u && rb == t ? q0 : box[rb + j];
            expand_record(q, rb + j, p, 0u, q.ntiles, 1u, cap, keys, vals);
            p += q.ntiles;
        }
    }
    for (uint64_t cb = __ballot(coop); cb; cb &= cb - 1ull) {   // (uniform)
        const int src = __builtin_ctzll(cb);
        const uint32_t rbs = (uint32_t)__builtin_amdgcn_readlane((int)rb, src);
        const uint32_t nrs = (uint32_t)__builtin_amdgcn_readlane((int)nr, src);
        uint64_t p = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(p0 >> 32), src) << 32) |
                     (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)p0, src);
        for (uint32_t j = 0; j < nrs; j++) {
            const RBox q = box[rbs + j];
            expand_record(q, rbs + j, p, (uint32_t)lane, q.ntiles, 64u, cap, keys, vals);
            p += q.ntiles;
        }
    }
}

// The queued triangles, one wave each (lanes stride over each record's tile
// box).  The sort and the tile ranges read only the pair total, so the
// capacity past it is left as it is.
__global__ void __launch_bounds__(64) k_expand_big(const unsigned long long *pair_base, const uint32_t *n_rt,
                                                   const uint32_t *rec_base, const RBox *box, uint64_t cap,
                                                   uint32_t *keys, uint32_t *vals,
                                                   const uint32_t *big, const Counters *ctr)
{
    const uint32_t lane = threadIdx.x;
    const uint32_t nbig = __builtin_amdgcn_readfirstlane(*(volatile const uint32_t *)&ctr->nbig);
    for (uint32_t b = blockIdx.x; b < nbig; b += gridDim.x) {
        const uint32_t t = __builtin_amdgcn_readfirstlane(big[b]);
        uint64_t p0 = pair_base[t];
        const uint32_t nr = n_rt[t], rb = rec_base[t];
        for (uint32_t j = 0; j < nr; j++) {
            const RBox q = box[rb + j];
            expand_record(q, rb + j, p0, lane, q.ntiles, 64u, cap, keys, vals);
            p0 += q.ntiles;
        }
    }
}

// ---------------------------------------------------------------------------
// stable LSD radix sort of (key, value) pairs by key
__device__ __forceinline__ bool exec_poisoned(const Counters *ctr)
{
    return __builtin_amdgcn_readfirstlane(*(volatile const uint32_t *)&ctr->poison) != 0u;
}

// ---------------------------------------------------------------------------
// stable LSD radix sort of (key, value) pairs by key
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(SORT_THREADS) k_sort_hist(const uint32_t *keys, uint32_t n_max,
                                                            const unsigned long long *n_dev, int shift,
                                                            uint32_t nblocks, uint32_t *counts)
{
    const uint32_t n = n_dev ? (uint32_t)min((unsigned long long)n_max, *n_dev) : n_max;
    __shared__ uint32_t hist[RADIX];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * SORT_TILE;
    // every key of the thread is loaded before the first count
    uint32_t kk[SORT_ITEMS];
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; r++) {
        const uint32_t i = base + r * SORT_THREADS + threadIdx.x;
        kk[r] = i < n ? keys[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; r++)
        if (base + r * SORT_THREADS + threadIdx.x < n) atomicAdd(&hist[(kk[r] >> shift) & (RADIX - 1)], 1u);
    __syncthreads();
    counts[threadIdx.x * nblocks + blockIdx.x] = hist[threadIdx.x];
}

__global__ void __launch_bounds__(SORT_THREADS) k_sort_scatter(const uint32_t *keys_in, const uint32_t *vals_in,
                                                               uint32_t n_max, const unsigned long long *n_dev,
                                                               int shift, uint32_t nblocks,
                                                               const uint32_t *offsets, uint32_t *keys_out,
                                                               uint32_t *vals_out)
{
    const uint32_t n = n_dev ? (uint32_t)min((unsigned long long)n_max, *n_dev) : n_max;
    if ((uint64_t)blockIdx.x * SORT_TILE >= n) return;   // beyond the keys: its counts are zero
    __shared__ uint32_t wave_hist[SORT_THREADS / 64][RADIX];
    __shared__ uint32_t running[RADIX];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    running[threadIdx.x] = offsets[threadIdx.x * nblocks + blockIdx.x];
    for (int w = 0; w < SORT_THREADS / 64; w++) wave_hist[w][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t base = blockIdx.x * SORT_TILE;
    // every key and value of the thread is loaded before the first round
    // (the rounds are separated by barriers, which loads do not cross)
    uint32_t kk[SORT_ITEMS], vv[SORT_ITEMS];
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; r++) {
        const uint32_t i = base + r * SORT_THREADS + threadIdx.x;
        kk[r] = i < n ? keys_in[i] : 0u;
        vv[r] = i < n ? vals_in[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; r++) {
        uint32_t i = base + r * SORT_THREADS + threadIdx.x;
        bool valid = i < n;
        uint32_t key = kk[r];
        uint32_t val = vv[r];
        uint32_t d = (key >> shift) & (RADIX - 1);
        // lanes of this wave holding the same digit
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < RADIX_BITS; b++) {
            uint64_t m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        uint32_t below = (uint32_t)__popcll(peers & lt_mask);
        if (valid && below == 0) wave_hist[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = running[d] + below;
            for (int w = 0; w < wave; w++) pos += wave_hist[w][d];
            if (pos < n) {
                keys_out[pos] = key;
                vals_out[pos] = val;
            }
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < SORT_THREADS / 64; w++) { add += wave_hist[w][threadIdx.x]; wave_hist[w][threadIdx.x] = 0; }
        running[threadIdx.x] += add;
        __syncthreads();
    }
}

// K3c: tile segments of the sorted pairs
// ---------------------------------------------------------------------------

// Tile ranges of the sorted pair keys, one thread per key: the first and
// last key of each run write its range; the run's tile is stamped with the
// draw's generation so no per-draw clear of the tile arrays is needed.
__global__ void __launch_bounds__(256) k_segments(const uint32_t *keys, uint32_t n_max,
                                                  const unsigned long long *n_dev, uint32_t ntiles,
                                                  uint32_t *tile_start, uint32_t *tile_end, uint32_t *tile_gen,
                                                  uint32_t gen)
{
    const uint32_t n = (uint32_t)min((unsigned long long)n_max, *n_dev);
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t k = keys[p];
    if (k >= ntiles) return;
    if (p == 0 || keys[p - 1] != k) tile_start[k] = p;
    if (p == n - 1 || keys[p + 1] != k) {
        tile_end[k] = p + 1;
        tile_gen[k] = gen;
    }
}

// The work items of the launch: HEAVY_SPLIT parts (TILE / HEAVY_SPLIT pixel
// rows each) of every tile whose bin holds more than heavy_bin entries and
// every tile with more than MED_BIN entries whole, from the front of `list`,
// and every other non-empty tile whole, from its back.  The raster kernels
// walk the list from the front and their blocks start in list order, so the
// long items start first.  The tile ranges are those k_segments stamped;
// each thread takes TILE_LIST_PER consecutive tiles, so a block's items keep
// tile order and one atomic (both list ends: nheavy and nlight are the low
// and high words of one aligned 64-bit counter) serves TILE_LIST_PER * 256
// tiles.
constexpr uint32_t TILE_LIST_PER = 4;
__global__ void __launch_bounds__(256) k_tile_list(const uint32_t *tile_start, const uint32_t *tile_end,
                                                   const uint32_t *tile_gen, uint32_t gen, uint32_t ntiles,
                                                   uint32_t heavy_bin, const uint32_t *tile_scene,
                                                   const uint32_t *tile_pre, uint32_t nscenes, uint4 *list,
                                                   uint32_t list_cap, Counters *ctr)
{
    __shared__ uint32_t s_waves[4];
    __shared__ uint32_t s_base[2];
    if (exec_poisoned(ctr)) return;
    const uint32_t t0 = (blockIdx.x * 256u + threadIdx.x) * TILE_LIST_PER;
    uint32_t st[TILE_LIST_PER], en[TILE_LIST_PER];
    uint32_t front = 0, back = 0;
#pragma unroll
    for (uint32_t q = 0; q < TILE_LIST_PER; q++) {
        const uint32_t t = t0 + q;
        const bool ne = t < ntiles && tile_gen[t] == gen;
        st[q] = ne ? tile_start[t] : 0u;
        en[q] = ne ? tile_end[t] : 0u;
        const uint32_t len = en[q] - st[q];
        front += len > heavy_bin ? HEAVY_SPLIT : len > MED_BIN ? 1u : 0u;
        back += len > 0u && len <= heavy_bin && len <= MED_BIN ? 1u : 0u;
    }
    // one scan of (front, back): a block's front items are at most
    // HEAVY_SPLIT * TILE_LIST_PER * 256 < 1 << 16
    static_assert(HEAVY_SPLIT * TILE_LIST_PER * 256 < (1u << 16), "packed list counts");
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan<uint32_t>(front | (back << 16), s_waves, tot);
    if (threadIdx.x == 0) {
        const uint32_t th = tot & 0xFFFFu, tl = tot >> 16;
        const unsigned long long hl =
            th | tl ? atomicAdd(reinterpret_cast<unsigned long long *>(&ctr->nheavy),
                                ((unsigned long long)tl << 32) | th)
                    : 0ull;
        s_base[0] = (uint32_t)hl;
        s_base[1] = (uint32_t)(hl >> 32);
    }
    __syncthreads();
    uint32_t of = s_base[0] + (ex & 0xFFFFu), ob = s_base[1] + (ex >> 16);
#pragma unroll
    for (uint32_t q = 0; q < TILE_LIST_PER; q++) {
        const uint32_t len = en[q] - st[q];
        if (len == 0u) continue;
        const uint32_t t = t0 + q;
        const uint32_t sc = nscenes == 1u ? 0u : tile_scene[t];
        // an item: scene-local tile | part << ITEM_TILE_BITS, bin start, bin end, scene
        const uint4 item = make_uint4(t - tile_pre[sc], st[q], en[q], sc);
        if (len > heavy_bin) {
            for (uint32_t h = 0; h < HEAVY_SPLIT; h++)
                if (of + h < list_cap)
                    list[of + h] = make_uint4(item.x | ((h + 1u) << ITEM_TILE_BITS), item.y, item.z, item.w);
            of += HEAVY_SPLIT;
        } else if (len > MED_BIN) {
            if (of < list_cap) list[of] = item;
            of++;
        } else {
            if (ob < list_cap) list[list_cap - 1u - ob] = item;
            ob++;
        }
    }
}

// The second walk of k_tile_bins over one wave's share [ws, we) of a
// bucket (positions in the bucket's order of pairs): each pair goes to its
// tile's cursor plus its rank among the equal tiles of its 64-pair chunk
// (ballots over the `shift` bits of the tile); the first lane of each tile
// then advances the cursor, after every lane has read it (one wave's LDS
// operations complete in order).  lk / val hold the share when it fits one
// round (pre), else load(c) fills them with the R chunks from position c.
template <uint32_t R, typename Load>
__device__ __forceinline__ void bins_scatter(uint32_t (&lk)[R], uint32_t (&val)[R], uint32_t ws, uint32_t we,
                                             bool pre, Load &&load, uint32_t nb, int shift, uint32_t *cur,
                                             uint32_t *vals_out)
{
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (uint32_t c = ws; c < we; c += 64u * R) {   // (uniform in the wave)
        if (!pre) load(c, true);
#pragma unroll
        for (uint32_t r = 0; r < R; r++) {
            if (c + r * 64u >= we) break;   // (uniform)
            const bool ok = lk[r] < nb;
            const uint32_t k = ok ? lk[r] : 0u;
            uint64_t peers = __ballot(ok);
            for (int bit = 0; bit < shift; bit++) {
                const uint64_t m = __ballot((k >> bit) & 1u);
                peers &= ((k >> bit) & 1u) ? m : ~m;
            }
            const uint32_t rank = (uint32_t)__popcll(peers & lt);
            const uint32_t base = ok ? cur[k] : 0u;
            if (ok) vals_out[base + rank] = val[r];
            if (ok && rank == 0u) cur[k] = base + (uint32_t)__popcll(peers);
        }
    }
}

// First level of the tile sort, one workgroup per LOCAL_TILE pairs: the
// block's pairs sorted by bucket (key >> shift), stably, within the block's
// own range of keys_out / vals_out, and per bucket the block's count and
// the start of its run within the block (rows of RADIX in counts and loc;
// zero rows for blocks past the pairs).  Each wave owns a contiguous quarter
// of the block: a first walk counts its pairs per bucket, a second ranks
// equal buckets within each 64-pair chunk by ballots and advances the wave's
// own cursors.
constexpr uint32_t LOCAL_TILE = 1024;
constexpr uint32_t LOCAL_PER_WAVE = LOCAL_TILE / (SORT_THREADS / 64);
constexpr uint32_t LOCAL_CHUNKS = LOCAL_PER_WAVE / 64;
static_assert(SORT_THREADS == RADIX && LOCAL_CHUNKS * 64 * (SORT_THREADS / 64) == LOCAL_TILE, "local sort shape");
__global__ void __launch_bounds__(SORT_THREADS) k_sort_local(const uint32_t *keys_in, const uint32_t *vals_in,
                                                             uint32_t n_max, const unsigned long long *n_dev,
                                                             int shift, uint32_t *counts, uint32_t *loc,
                                                             uint32_t *keys_out, uint32_t *vals_out)
{
    constexpr int W = SORT_THREADS / 64;
    __shared__ uint32_t cur[W][RADIX];
    __shared__ uint32_t s_waves[W];
    const uint32_t n = (uint32_t)min((unsigned long long)n_max, *n_dev);
    const uint32_t base = blockIdx.x * LOCAL_TILE;
    const uint32_t row = blockIdx.x * RADIX + threadIdx.x;
    if (base >= n) {
        counts[row] = 0u;
        loc[row] = 0u;
        return;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t kk[LOCAL_CHUNKS], vv[LOCAL_CHUNKS], d[LOCAL_CHUNKS];
#pragma unroll
    for (uint32_t r = 0; r < LOCAL_CHUNKS; r++) {
        const uint32_t i = base + (uint32_t)w * LOCAL_PER_WAVE + r * 64u + (uint32_t)lane;
        const bool valid = i < n;
        kk[r] = valid ? keys_in[i] : 0u;
        vv[r] = valid ? vals_in[i] : 0u;
        d[r] = valid ? (kk[r] >> shift) & (RADIX - 1) : (uint32_t)RADIX;
    }
#pragma unroll
    for (int v = 0; v < W; v++) cur[v][threadIdx.x] = 0u;
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < LOCAL_CHUNKS; r++)
        if (d[r] < (uint32_t)RADIX) atomicAdd(&cur[w][d[r]], 1u);
    __syncthreads();
    // bucket threadIdx.x: the block's count, its start, each wave's cursor
    uint32_t c[W], tot = 0;
#pragma unroll
    for (int v = 0; v < W; v++) {
        c[v] = cur[v][threadIdx.x];
        tot += c[v];
    }
    uint32_t all;
    const uint32_t ex = block_exclusive_scan<uint32_t>(tot, s_waves, all);
    counts[row] = tot;
    loc[row] = ex;
    uint32_t run = base + ex;
#pragma unroll
    for (int v = 0; v < W; v++) {
        cur[v][threadIdx.x] = run;
        run += c[v];
    }
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (uint32_t r = 0; r < LOCAL_CHUNKS; r++) {
        const bool ok = d[r] < (uint32_t)RADIX;
        const uint32_t k = ok ? d[r] : 0u;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int bit = 0; bit < RADIX_BITS; bit++) {
            const uint64_t m = __ballot((k >> bit) & 1u);
            peers &= ((k >> bit) & 1u) ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        const uint32_t at = ok ? cur[w][k] : 0u;
        if (ok) {
            keys_out[at + rank] = kk[r];
            vals_out[at + rank] = vv[r];
        }
        if (ok && rank == 0u) cur[w][k] = at + (uint32_t)__popcll(peers);
    }
}

// Second level of the tile sort when every tile key is below
// RADIX << BINS_MAX_SHIFT: one workgroup per bucket of 1 << shift
// consecutive tiles counts its tiles, sorts its pairs by tile, stably, into
// vals_out, stamps its non-empty tiles with the launch generation and
// appends their work items.  The bucket's pairs, in submission order, are
// either one contiguous run (offsets: the scanned, digit-major counts of a
// global first pass), or the runs k_sort_local left in each block, taken in
// block order (counts / loc: a row of RADIX per block, at most
// BINS_MAX_SEGS blocks); the bucket's place in vals_out is then the sum of
// every block's start of its run.
// Each wave owns a contiguous 1/BINS_WAVES of the bucket: a first walk counts
// its pairs per tile, so every wave knows where its pairs of each tile start;
// a second walk ranks equal tiles within each 64-pair chunk by ballots and
// advances the wave's own cursors, so no barrier is needed between chunks.
// A wave whose share fits BINS_ROUND chunks loads it once, for both walks;
// the list slots are reserved before the second walk and filled after it,
// so the atomics' round trip overlaps the walk.
constexpr int BINS_WAVES = 16;
constexpr int BINS_THREADS = 64 * BINS_WAVES;
constexpr int BINS_MAX_SHIFT = 10;
constexpr uint32_t BINS_ROUND = 8;
constexpr uint32_t BINS_MAX_SEGS = BINS_THREADS;
static_assert((1 << BINS_MAX_SHIFT) <= BINS_THREADS, "one tile of a bucket per thread");
__global__ void __launch_bounds__(BINS_THREADS) k_tile_bins(const uint32_t *keys, const uint32_t *vals,
                                                            uint32_t n_max, const unsigned long long *n_dev,
                                                            int shift, uint32_t nblocks,
                                                            const uint32_t *offsets, const uint32_t *counts,
                                                            const uint32_t *loc, uint32_t ntiles,
                                                            uint32_t *vals_out, uint32_t *tile_gen, uint32_t gen,
                                                            uint32_t heavy_bin, const uint32_t *tile_scene,
                                                            const uint32_t *tile_pre, uint32_t nscenes, uint4 *list,
                                                            uint32_t list_cap, Counters *ctr)
{
    __shared__ uint32_t cur[BINS_WAVES][1 << BINS_MAX_SHIFT];   // per wave and tile: count, then cursor
    __shared__ uint32_t seg_v[BINS_MAX_SEGS + 1];   // runs of the bucket: first position in bucket order
    __shared__ uint32_t seg_src[BINS_MAX_SEGS];     // ... and in keys / vals
    __shared__ uint32_t s_waves[BINS_WAVES];
    __shared__ uint32_t s_n[2][BINS_WAVES];
    __shared__ uint32_t s_base[2];
    const uint32_t bk = blockIdx.x;
    // (the loads are independent; nothing is written when poisoned)
    const uint32_t poison = __builtin_amdgcn_readfirstlane(*(volatile const uint32_t *)&ctr->poison);
    const uint32_t n = (uint32_t)min((unsigned long long)n_max, *n_dev);
    uint32_t a, len, nseg = 1u;
    if (!counts) {
        const uint32_t oa = offsets[bk * nblocks];
        const uint32_t oe = bk + 1u < (uint32_t)RADIX ? offsets[(bk + 1u) * nblocks] : 0xFFFFFFFFu;
        if (poison != 0u) return;
        a = min(oa, n);
        const uint32_t e = min(oe, n);
        len = e > a ? e - a : 0u;
        if (threadIdx.x == 0) {
            seg_v[0] = 0u;
            seg_src[0] = a;
            seg_v[1] = len;
        }
    } else {
        // thread i: block i's run of the bucket
        const uint32_t i = threadIdx.x;
        const uint32_t c = i < nblocks ? counts[i * RADIX + bk] : 0u;
        const uint32_t lo = i < nblocks ? loc[i * RADIX + bk] : 0u;
        if (poison != 0u) return;
        // one scan of (pairs, runs): a block holds at most LOCAL_TILE pairs,
        // so the bucket at most BINS_MAX_SEGS * LOCAL_TILE < 1 << 21
        uint32_t tot;
        const uint32_t packed = block_exclusive_scan<uint32_t>(c | (c ? 1u << 21 : 0u), s_waves, tot);
        len = tot & ((1u << 21) - 1u);
        nseg = tot >> 21;
        block_exclusive_scan<uint32_t>(lo, s_waves, a);   // a: the pairs of every earlier bucket
        if (c) {
            seg_v[packed >> 21] = packed & ((1u << 21) - 1u);
            seg_src[packed >> 21] = i * LOCAL_TILE + lo;
        }
        if (i == 0) seg_v[nseg] = len;
    }
    const uint32_t t0 = bk << shift;
    const uint32_t nb = min(1u << shift, ntiles - t0);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t share = ((len + BINS_WAVES - 1u) / BINS_WAVES + 63u) & ~63u;
    const uint32_t ws = min((uint32_t)w * share, len), we = min(ws + share, len);
    const bool one_round = share <= 64u * BINS_ROUND;   // (uniform)
    uint32_t lk[BINS_ROUND], val[BINS_ROUND];
    for (uint32_t k = threadIdx.x; k < nb; k += BINS_THREADS)
        for (int v = 0; v < BINS_WAVES; v++) cur[v][k] = 0u;
    __syncthreads();   // the runs are in place
    // positions [c, c + 64 * BINS_ROUND) of the bucket, through its runs: the
    // run of position c by a search every lane makes alike, then each lane
    // steps over the (few) run ends before its own positions
    auto load = [&](uint32_t c, bool with_vals) {
        uint32_t j = 0, hi = nseg;
        while (hi - j > 1u) {
            const uint32_t mid = (j + hi) >> 1;
            if (seg_v[mid] <= c) j = mid;
            else hi = mid;
        }
#pragma unroll
        for (uint32_t r = 0; r < BINS_ROUND; r++) {
            const uint32_t v = c + r * 64u + (uint32_t)lane;
            const bool in = v < we;
            uint32_t i = 0;
            if (in) {
                while (seg_v[j + 1u] <= v) j++;   // (seg_v[nseg] = len ends every walk)
                i = seg_src[j] + (v - seg_v[j]);
            }
            lk[r] = in ? keys[i] - t0 : 0xFFFFFFFFu;
            if (with_vals) val[r] = in ? vals[i] : 0u;
        }
    };
    if (one_round) load(ws, true);
    for (uint32_t c = ws; c < we; c += 64u * BINS_ROUND) {
        if (!one_round) load(c, false);
#pragma unroll
        for (uint32_t r = 0; r < BINS_ROUND; r++)
            if (lk[r] < nb) atomicAdd(&cur[w][lk[r]], 1u);
    }
    __syncthreads();
    // tile k0 of the bucket per thread: its range, and its list slots
    // reserved now and filled last
    const uint32_t k0 = threadIdx.x;
    uint32_t sum = 0;
    if (k0 < nb)
        for (int v = 0; v < BINS_WAVES; v++) sum += cur[v][k0];
    uint32_t total;
    uint32_t pos = a + block_exclusive_scan<uint32_t>(sum, s_waves, total);
    const uint32_t start = pos;
    if (k0 < nb)
        for (int v = 0; v < BINS_WAVES; v++) {
            const uint32_t cnt = cur[v][k0];
            cur[v][k0] = pos;
            pos += cnt;
        }
    const bool ne = k0 < nb && pos > start;
    const uint32_t blen = pos - start;
    const bool hv = ne && blen > heavy_bin;
    const bool md = ne && !hv && blen > MED_BIN;
    const bool ltl = ne && !hv && !md;
    const uint64_t bh = __ballot(hv), bm = __ballot(md), bl = __ballot(ltl);
    if (lane == 0) {
        s_n[0][w] = HEAVY_SPLIT * (uint32_t)__popcll(bh) + (uint32_t)__popcll(bm);   // front items
        s_n[1][w] = (uint32_t)__popcll(bl);
    }
    __syncthreads();
    uint32_t hb = 0, lb = 0;
    if (threadIdx.x == 0) {
        uint32_t th = 0, tl = 0;
        for (int v = 0; v < BINS_WAVES; v++) {
            const uint32_t x0 = s_n[0][v], x1 = s_n[1][v];
            s_n[0][v] = th;
            s_n[1][v] = tl;
            th += x0;
            tl += x1;
        }
        const unsigned long long hl =
            th | tl ? atomicAdd(reinterpret_cast<unsigned long long *>(&ctr->nheavy),
                                ((unsigned long long)tl << 32) | th)
                    : 0ull;
        hb = (uint32_t)hl;
        lb = (uint32_t)(hl >> 32);
    }
    __syncthreads();   // cursors and wave offsets in place
    bins_scatter(lk, val, ws, we, one_round, load, nb, shift, cur[w], vals_out);
    if (threadIdx.x == 0) {
        s_base[0] = hb;
        s_base[1] = lb;
    }
    __syncthreads();
    if (ne) {
        tile_gen[t0 + k0] = gen;
        const uint64_t below = (1ull << lane) - 1ull;
        const uint32_t t = t0 + k0;
        const uint32_t sc = nscenes == 1u ? 0u : tile_scene[t];
        const uint4 item = make_uint4(t - tile_pre[sc], start, pos, sc);
        const uint32_t of = s_base[0] + s_n[0][w] + HEAVY_SPLIT * (uint32_t)__popcll(bh & below) +
                            (uint32_t)__popcll(bm & below);
        if (hv) {
            for (uint32_t q = 0; q < HEAVY_SPLIT; q++)
                if (of + q < list_cap)
                    list[of + q] = make_uint4(item.x | ((q + 1u) << ITEM_TILE_BITS), item.y, item.z, item.w);
        }
        if (md && of < list_cap) list[of] = item;
        if (ltl) {
            const uint32_t o = s_base[1] + s_n[1][w] + (uint32_t)__popcll(bl & below);
            if (o < list_cap) list[list_cap - 1u - o] = item;
        }
    }
}


// ---------------------------------------------------------------------------
// K4: fine raster and per-fragment pipeline
// ---------------------------------------------------------------------------


// per-launch counters at zero: k_launch_begin before the point / line path,
// and before a tile launch when the previous one stopped before k_expand
// (which, with k_setup, otherwise leaves them at zero)
__device__ __forceinline__ void reset_counters(Counters *ctr)
{
    ctr->npairs = 0;
    ctr->nheavy = 0;
    ctr->nlight = 0;
    ctr->unsupported = 0;
    ctr->noverflow = 0;
    ctr->nbig = 0;
    ctr->sticky |= ctr->overflow_full;
    ctr->overflow_full = 0;
}

__global__ void __launch_bounds__(64) k_launch_begin(Counters *ctr)
{
    if (threadIdx.x == 0) reset_counters(ctr);
}

// ---------------------------------------------------------------------------
// K4: fine raster and per-fragment pipeline
// ---------------------------------------------------------------------------

__device__ __forceinline__ int compare_op(uint32_t func, float a, float b)
{
    switch (func) {
    case GL_NEVER: return 0;
    case GL_LESS: return a < b;
    case GL_EQUAL: return a == b;
    case GL_LEQUAL: return a <= b;
    case GL_GREATER: return a > b;
    case GL_NOTEQUAL: return a != b;
    case GL_GEQUAL: return a >= b;
    default: return 1;
    }
}

__device__ __forceinline__ int stencil_test(uint32_t func, int32_t ref, uint32_t mask, uint32_t s)
{
    int32_t r = (int32_t)((uint32_t)ref & mask);
    int32_t v = (int32_t)(s & mask);
    switch (func) {
    case GL_NEVER: return 0;
    case GL_LESS: return r < v;
    case GL_EQUAL: return r == v;
    case GL_LEQUAL: return r <= v;
    case GL_GREATER: return r > v;
    case GL_NOTEQUAL: return r != v;
    case GL_GEQUAL: return r >= v;
    default: return 1;
    }
}

// GL comparison functions as accept masks: GL_NEVER..GL_ALWAYS are
// 0x200 + (less ? 1 : 0) + (equal ? 2 : 0) + (greater ? 4 : 0); any other
// value behaves as GL_ALWAYS, as in compare_op
__device__ __forceinline__ uint32_t cmp_mask_of(uint32_t func)
{
    return (func >= GL_NEVER && func <= GL_ALWAYS) ? func - GL_NEVER : 7u;
}

// compare_op through an accept mask; unordered operands (NaN) pass only
// GL_NOTEQUAL and GL_ALWAYS, the two masks holding both less and greater
__device__ __forceinline__ bool cmp_f(uint32_t m, float a, float b)
{
    const bool lt = a < b, eq = a == b, gt = a > b;
    const bool un = !(lt | eq | gt);
    return (lt & ((m & 1u) != 0u)) | (eq & ((m & 2u) != 0u)) | (gt & ((m & 4u) != 0u)) | (un & ((m & 5u) == 5u));
}

// stencil_test through an accept mask
__device__ __forceinline__ bool cmp_stencil(uint32_t m, int32_t ref, uint32_t mask, uint32_t s)
{
    const int32_t r = (int32_t)((uint32_t)ref & mask);
    const int32_t v = (int32_t)(s & mask);
    return ((r < v) & ((m & 1u) != 0u)) | ((r == v) & ((m & 2u) != 0u)) | ((r > v) & ((m & 4u) != 0u));
}

__device__ __forceinline__ uint32_t stencil_op(uint32_t op, uint32_t s, int32_t ref)
{
    switch (op) {
    case GL_KEEP: return s;
    case GL_ZERO: return 0;
    case GL_REPLACE: return (uint32_t)ref & 0xFFu;
    case GL_INCR: return s < 255 ? s + 1 : 255;
    case GL_INCR_WRAP: return (s + 1) & 0xFFu;
    case GL_DECR: return s > 0 ? s - 1 : 0;
    case GL_DECR_WRAP: return (s - 1) & 0xFFu;
    case GL_INVERT: return (~s) & 0xFFu;
    default: return s;
    }
}

__device__ __forceinline__ Col blend_factor(uint32_t f, Col s, Col d)
{
    switch (f) {
    case GL_ZERO: return {0.0f, 0.0f, 0.0f, 0.0f};
    case GL_ONE: return {1.0f, 1.0f, 1.0f, 1.0f};
    case GL_SRC_COLOR: return s;
    case GL_ONE_MINUS_SRC_COLOR: return {1 - s.r, 1 - s.g, 1 - s.b, 1 - s.a};
    case GL_DST_COLOR: return d;
    case GL_ONE_MINUS_DST_COLOR: return {1 - d.r, 1 - d.g, 1 - d.b, 1 - d.a};
    case GL_SRC_ALPHA: return {s.a, s.a, s.a, s.a};
    case GL_ONE_MINUS_SRC_ALPHA: return {1 - s.a, 1 - s.a, 1 - s.a, 1 - s.a};
    case GL_DST_ALPHA: return {d.a, d.a, d.a, d.a};
    case GL_ONE_MINUS_DST_ALPHA: return {1 - d.a, 1 - d.a, 1 - d.a, 1 - d.a};
    case GL_SRC_ALPHA_SATURATE: {
        float k = s.a < (1 - d.a) ? s.a : (1 - d.a);
        return {k, k, k, 1.0f};
    }
    default: return {1.0f, 1.0f, 1.0f, 1.0f};
    }
}

struct TexLevel {
    const uint32_t *px;
    int32_t w, h;
    bool pow2;
};

// the draw's fragment state, copied once per block into (scalar) registers
struct RasterConsts {
    const mtgl_raster_state *rs;
    bool depth_on, stencil_on, blend_on, fog_on, alpha_on, lighting_on, textured, persp, flat, phong, unit_range;
    bool rep_s, rep_t, depth_mask, two_side, depth_write;
    uint32_t pass_bytes;   // bytes a passing, kept fragment writes: depth and colour
    uint32_t cmask;
    uint32_t depth_cmp, stencil_cmp, alpha_cmp;   // accept masks (cmp_mask_of)
    uint32_t stencil_mask, stencil_writemask, stencil_fail, stencil_zfail, stencil_zpass;
    int32_t stencil_ref;
    uint32_t blend_src, blend_dst, tex_env_mode, fog_mode, min_filter, mag_filter;
    float alpha_ref, fog_density, fog_start, fog_end;
    Col tex_env_color, fog_color;
    double depth_near, depth_far;
    TexLevel L0, L1;
    bool chain_on;                // full chain with a per-pixel LOD (extension)
    const uint32_t *chain;
    int32_t nlev;
    // How a pixel's depth test over fragments in submission order reduces
    // (no stencil): DEPTH_IND, each fragment's test reads only the depth
    // before the first (mask off, EQUAL, NEVER, ALWAYS, no depth test);
    // DEPTH_MIN / DEPTH_MAX, the depth a fragment is tested against is the
    // smallest / largest of the first depth and every earlier fragment's
    // (LESS, LEQUAL / GREATER, GEQUAL with writes); DEPTH_SERIAL otherwise.
    uint32_t depth_mode;
};

enum : uint32_t { DEPTH_IND = 0u, DEPTH_MIN = 1u, DEPTH_MAX = 2u, DEPTH_SERIAL = 3u };

__device__ __forceinline__ int32_t wrap_coord(int32_t x, int32_t w, bool repeat, bool pow2)
{
    if (repeat) {
        if (pow2) return x & (w - 1);
        return ((x % w) + w) % w;
    }
    return x < 0 ? 0 : (x >= w ? w - 1 : x);
}

// one level: bilinear with wrapped taps (reference bilinear_filter,
// get_texel_wrapped) or nearest with a clamped index
// the four texels (00, 10, 01, 11) and fractions of a bilinear tap at (u, v)
__device__ __forceinline__ void bilinear_taps(const TexLevel &L, bool rep_s, bool rep_t, float u, float v, size_t *at,
                                              float &fx, float &fy)
{
    float tx = u * (float)L.w - 0.5f;
    float ty = v * (float)L.h - 0.5f;
    int32_t x0 = cvtt(floorf(tx));
    int32_t y0 = cvtt(floorf(ty));
    int32_t x1 = x0 + 1, y1 = y0 + 1;
    fx = tx - (float)x0;
    fy = ty - (float)y0;
    int32_t wx0 = wrap_coord(x0, L.w, rep_s, L.pow2), wx1 = wrap_coord(x1, L.w, rep_s, L.pow2);
    int32_t wy0 = wrap_coord(y0, L.h, rep_t, L.pow2), wy1 = wrap_coord(y1, L.h, rep_t, L.pow2);
    at[0] = (size_t)wy0 * L.w + wx0;
    at[1] = (size_t)wy0 * L.w + wx1;
    at[2] = (size_t)wy1 * L.w + wx0;
    at[3] = (size_t)wy1 * L.w + wx1;
}

__device__ uint32_t sample_level(const TexLevel &L, bool rep_s, bool rep_t, float u, float v, bool linear, int &taps)
{
    if (linear) {
        size_t at[4];
        float fx, fy;
        bilinear_taps(L, rep_s, rep_t, u, v, at, fx, fy);
        uint32_t c00 = L.px[at[0]];
        uint32_t c10 = L.px[at[1]];
        uint32_t c01 = L.px[at[2]];
        uint32_t c11 = L.px[at[3]];
        taps += 4;
        Col top = col_lerp(col_unpack(c00), col_unpack(c10), fx);
        Col bot = col_lerp(col_unpack(c01), col_unpack(c11), fx);
        return col_pack(col_lerp(top, bot, fy));
    }
    const float tx = u * (float)L.w - 0.5f;
    const float ty = v * (float)L.h - 0.5f;
    int32_t x = cvtt(floorf(tx + 0.5f));
    int32_t y = cvtt(floorf(ty + 0.5f));
    if (x < 0) x = 0;
    if (x >= L.w) x = L.w - 1;
    if (y < 0) y = 0;
    if (y >= L.h) y = L.h - 1;
    taps += 1;
    return L.px[(size_t)y * L.w + x];
}

// reference texture_sample_lod (src/textures.c:457-557)
// texture coordinate wrap of texture_sample_lod: fraction for GL_REPEAT, else clamp to [0, 1]
__device__ __forceinline__ void wrap_uv(bool rep_s, bool rep_t, float &u, float &v)
{
    if (rep_s) { u = u - (float)cvtt(u); if (u < 0) u += 1.0f; }
    else { if (u < 0.0f) u = 0.0f; if (u > 1.0f) u = 1.0f; }
    if (rep_t) { v = v - (float)cvtt(v); if (v < 0) v += 1.0f; }
    else { if (v < 0.0f) v = 0.0f; if (v > 1.0f) v = 1.0f; }
}

__device__ uint32_t sample_texture(const RasterConsts &K, const TexLevel &L0, const TexLevel &L1, float u, float v,
                                   float lod, int &taps)
{
    const bool rep_s = K.rep_s, rep_t = K.rep_t;
    wrap_uv(rep_s, rep_t, u, v);
    uint32_t filter = lod > 0.0f ? K.min_filter : K.mag_filter;
    if (filter == GL_NEAREST_MIPMAP_NEAREST || filter == GL_LINEAR_MIPMAP_NEAREST) {
        if (lod >= 0.5f) {
            if (!L1.px) return 0xFFFFFFFFu;
            return sample_level(L1, rep_s, rep_t, u, v, filter == GL_LINEAR_MIPMAP_NEAREST, taps);
        }
        filter = filter == GL_NEAREST_MIPMAP_NEAREST ? GL_NEAREST : GL_LINEAR;
    } else if (filter == GL_NEAREST_MIPMAP_LINEAR || filter == GL_LINEAR_MIPMAP_LINEAR) {
        if (lod > 0.0f) {
            float cl = lod > 1.0f ? 1.0f : lod;
            bool lin = filter == GL_LINEAR_MIPMAP_LINEAR;
            uint32_t c0 = sample_level(L0, rep_s, rep_t, u, v, lin, taps);
            uint32_t c1 = L1.px ? sample_level(L1, rep_s, rep_t, u, v, lin, taps) : 0xFFFFFFFFu;
            return col_pack(col_lerp(col_unpack(c0), col_unpack(c1), cl));
        }
        filter = filter == GL_NEAREST_MIPMAP_LINEAR ? GL_NEAREST : GL_LINEAR;
    }
    return sample_level(L0, rep_s, rep_t, u, v, filter == GL_LINEAR, taps);
}


// per-pixel state a lane carries through its tile
struct PixelState {
    uint32_t c;          // packed colour
    float d;             // depth
    uint32_t s;          // stencil
    bool dirty_c, dirty_d, dirty_s;
};


__device__ __forceinline__ void stencil_write(const RasterConsts &K, PixelState &px, uint32_t op)
{
    uint32_t nv = stencil_op(op, px.s, K.stencil_ref);
    uint32_t m8 = K.stencil_writemask & 0xFFu;
    px.s = ((px.s & ~m8) | (nv & m8)) & 0xFFu;
    px.dirty_s = true;
}

// texture coordinates of a fragment (src/raster.c:618-633): perspective-correct
// unless the hint is GL_FASTEST
__device__ __forceinline__ void texcoord_at(const RTri *r, bool persp, float b0, float b1, float b2, float &u, float &v)
{
    if (persp) {
        const float u0w = r->uv0[0] * r->w0, v0w = r->uv0[1] * r->w0;
        const float u1w = r->uv1[0] * r->w1, v1w = r->uv1[1] * r->w1;
        const float u2w = r->uv2[0] * r->w2, v2w = r->uv2[1] * r->w2;
        float uow = b0 * u0w + b1 * u1w + b2 * u2w;
        float vow = b0 * v0w + b1 * v1w + b2 * v2w;
        float oow = b0 * r->w0 + b1 * r->w1 + b2 * r->w2;
        float w = 1.0f / oow;
        u = uow * w;
        v = vow * w;
    } else {
        u = b0 * r->uv0[0] + b1 * r->uv1[0] + b2 * r->uv2[0];
        v = b0 * r->uv0[1] + b1 * r->uv1[1] + b2 * r->uv2[1];
    }
}

// ---- full mip chain with a per-pixel LOD (MTGL_LOD_FULL_CHAIN, an extension) ----

// level k of the chain: 0 and 1 as in the reference's sampling, then the
// levels stored back to back after level 1, each halving the previous sides
__device__ __forceinline__ TexLevel chain_level(const RasterConsts &K, int k)
{
    if (k <= 0) return K.L0;
    if (k == 1) return K.L1;
    int32_t w = K.L1.w / 2, h = K.L1.h / 2;
    size_t off = 0;
    for (int j = 2; j < k; j++) {
        off += (size_t)w * (size_t)h;
        w /= 2;
        h /= 2;
    }
    return {K.chain + off, w, h, (w & (w - 1)) == 0 && (h & (h - 1)) == 0};
}

// log2 of the larger of |d(s, t)/dx| and |d(s, t)/dy| in texels at a pixel
// with weights b: the derivatives of the weights are the reference's edge
// function coefficients over the area; with perspective the coordinates are
// (sum b u/w) / (sum b 1/w), differentiated as a quotient
__device__ float pixel_lod(const RasterConsts &K, const RTri *r, float b0, float b1, float b2)
{
    const float ia = r->inv_area;
    const float x0 = (float)r->x0, y0 = (float)r->y0, x1 = (float)r->x1, y1 = (float)r->y1;
    const float x2 = (float)r->x2, y2 = (float)r->y2;
    const float d0x = (y2 - y1) * ia, d0y = -(x2 - x1) * ia;
    const float d1x = (y0 - y2) * ia, d1y = -(x0 - x2) * ia;
    const float d2x = (y1 - y0) * ia, d2y = -(x1 - x0) * ia;
    float ux, uy, vx, vy;
    if (K.persp) {
        const float a0 = r->w0, a1 = r->w1, a2 = r->w2;
        const float u0 = r->uv0[0] * a0, u1 = r->uv1[0] * a1, u2 = r->uv2[0] * a2;
        const float v0 = r->uv0[1] * a0, v1 = r->uv1[1] * a1, v2 = r->uv2[1] * a2;
        const float q = b0 * a0 + b1 * a1 + b2 * a2;
        const float pu = b0 * u0 + b1 * u1 + b2 * u2, pv = b0 * v0 + b1 * v1 + b2 * v2;
        const float qx = d0x * a0 + d1x * a1 + d2x * a2, qy = d0y * a0 + d1y * a1 + d2y * a2;
        const float iq2 = 1.0f / (q * q);
        ux = ((d0x * u0 + d1x * u1 + d2x * u2) * q - pu * qx) * iq2;
        uy = ((d0y * u0 + d1y * u1 + d2y * u2) * q - pu * qy) * iq2;
        vx = ((d0x * v0 + d1x * v1 + d2x * v2) * q - pv * qx) * iq2;
        vy = ((d0y * v0 + d1y * v1 + d2y * v2) * q - pv * qy) * iq2;
    } else {
        ux = d0x * r->uv0[0] + d1x * r->uv1[0] + d2x * r->uv2[0];
        uy = d0y * r->uv0[0] + d1y * r->uv1[0] + d2y * r->uv2[0];
        vx = d0x * r->uv0[1] + d1x * r->uv1[1] + d2x * r->uv2[1];
        vy = d0y * r->uv0[1] + d1y * r->uv1[1] + d2y * r->uv2[1];
    }
    const float tw = (float)K.L0.w, th = (float)K.L0.h;
    ux *= tw; uy *= tw; vx *= th; vy *= th;
    const float rho = fmaxf(sqrtf(ux * ux + vx * vx), sqrtf(uy * uy + vy * vy));
    return log2f(rho);
}

// GL 1.5 level selection over the chain; each level sampled and packed to
// RGBA8, two levels blended like the reference's level 0 / level 1 blend
__device__ uint32_t sample_chain(const RasterConsts &K, float u, float v, float lod, int &taps)
{
    const bool rep_s = K.rep_s, rep_t = K.rep_t;
    wrap_uv(rep_s, rep_t, u, v);
    if (!(lod > 0.0f)) return sample_level(K.L0, rep_s, rep_t, u, v, K.mag_filter == GL_LINEAR, taps);
    const uint32_t f = K.min_filter;
    if (f == GL_NEAREST || f == GL_LINEAR) return sample_level(K.L0, rep_s, rep_t, u, v, f == GL_LINEAR, taps);
    const bool lin = f == GL_LINEAR_MIPMAP_NEAREST || f == GL_LINEAR_MIPMAP_LINEAR;
    const int top = K.nlev - 1;
    if (f == GL_NEAREST_MIPMAP_NEAREST || f == GL_LINEAR_MIPMAP_NEAREST) {
        const float d = fminf(fmaxf(ceilf(lod + 0.5f) - 1.0f, 0.0f), (float)top);
        return sample_level(chain_level(K, (int)d), rep_s, rep_t, u, v, lin, taps);
    }
    if (lod >= (float)top) return sample_level(chain_level(K, top), rep_s, rep_t, u, v, lin, taps);
    const float fl = floorf(lod);
    const int d = (int)fl;
    const uint32_t c0 = sample_level(chain_level(K, d), rep_s, rep_t, u, v, lin, taps);
    const uint32_t c1 = sample_level(chain_level(K, d + 1), rep_s, rep_t, u, v, lin, taps);
    return col_pack(col_lerp(col_unpack(c0), col_unpack(c1), lod - fl));
}

// The part of one fragment's pipeline (src/raster.c:546-721) that does not
// read the framebuffer: window depth, colour through lighting, texturing and
// fog, whether the alpha test keeps it, and how many texels it read.  It
// depends only on the triangle and the pixel, so the fragments of a tile are
// computed in parallel and in any order; resolve_fragment then applies the
// stencil, depth, blend and mask steps in submission order.  meta = kept
// (bit 0) | texel taps << 1.
__device__ void fragment_source(const RasterConsts &K, const RTri *r, const RCold *q, float b0, float b1, float b2,
                                Col &c_out, float &depth_out, uint32_t &meta_out)
{
    const mtgl_raster_state *rs = K.rs;
    const float z = b0 * r->z0 + b1 * r->z1 + b2 * r->z2;
    if (K.unit_range) depth_out = (z + 1.0f) * 0.5f;   // (double)q * 1.0 + 0.0 rounds back to q
    else depth_out = (float)((double)((z + 1.0f) * 0.5f) * (K.depth_far - K.depth_near) + K.depth_near);

    Col c;
    if (K.flat) {
        c = {r->c2[0], r->c2[1], r->c2[2], r->c2[3]};
    } else {
        c.r = r->c0[0] * b0 + r->c1[0] * b1 + r->c2[0] * b2;
        c.g = r->c0[1] * b0 + r->c1[1] * b1 + r->c2[1] * b2;
        c.b = r->c0[2] * b0 + r->c1[2] * b1 + r->c2[2] * b2;
        c.a = r->c0[3] * b0 + r->c1[3] * b1 + r->c2[3] * b2;
    }
    meta_out = 1u;
    if (K.lighting_on) {
        const bool two = r->back && K.two_side;
        if (K.phong || two) {
            float ep[3], en[3];
            for (int k = 0; k < 3; k++) {
                ep[k] = q->ep[k] * b0 + q->ep[3 + k] * b1 + q->ep[6 + k] * b2;
                en[k] = q->en[k] * b0 + q->en[3 + k] * b1 + q->en[6 + k] * b2;
            }
            const mtgl_material *m = &rs->lighting.front;
            if (two) {
                for (int k = 0; k < 3; k++) en[k] = en[k] * -1.0f;
                m = &rs->lighting.back;
            }
            c = compute_lighting(&rs->lighting, rs->flags, ep, en, col_of(m->ambient), col_of(m->diffuse),
                                 col_of(m->specular), col_of(m->emission), m->shininess);
        }
    }
    if (K.textured) {
        float u, v;
        texcoord_at(r, K.persp, b0, b1, b2, u, v);
        int taps = 0;
        Col tc = col_unpack(K.chain_on ? sample_chain(K, u, v, pixel_lod(K, r, b0, b1, b2), taps)
                                       : sample_texture(K, K.L0, K.L1, u, v, r->lod, taps));
        if (K.alpha_on && !cmp_f(K.alpha_cmp, tc.a, K.alpha_ref)) {
            c_out = c;
            meta_out = (uint32_t)taps << 1;
            return;
        }
        meta_out = 1u | ((uint32_t)taps << 1);
        switch (K.tex_env_mode) {
        case GL_REPLACE: c = tc; break;
        case GL_DECAL:
            c = {c.r + (tc.r - c.r) * tc.a, c.g + (tc.g - c.g) * tc.a, c.b + (tc.b - c.b) * tc.a, c.a};
            break;
        case GL_BLEND: {
            Col e = K.tex_env_color;
            c = {c.r * (1.0f - tc.r) + e.r * tc.r, c.g * (1.0f - tc.g) + e.g * tc.g, c.b * (1.0f - tc.b) + e.b * tc.b,
                 c.a * tc.a};
            break;
        }
        case GL_ADD: c = {c.r + tc.r, c.g + tc.g, c.b + tc.b, c.a * tc.a}; break;
        default: c = col_mul(c, tc); break;
        }
    }
    if (K.fog_on) {
        float fc = b0 * r->ez0 + b1 * r->ez1 + b2 * r->ez2;
        float f;
        switch (K.fog_mode) {
        case GL_LINEAR: f = K.fog_end != K.fog_start ? (K.fog_end - fc) / (K.fog_end - K.fog_start) : 1.0f; break;
        case GL_EXP: f = exp_cr(-K.fog_density * fc); break;
        case GL_EXP2: { float d = K.fog_density * fc; f = exp_cr(-d * d); break; }
        default: f = 1.0f; break;
        }
        if (f < 0.0f) f = 0.0f;
        if (f > 1.0f) f = 1.0f;
        // color_lerp_rgb(fog_color, c, f): the alpha is the fog colour's
        Col fg = K.fog_color;
        c = {fg.r + (c.r - fg.r) * f, fg.g + (c.g - fg.g) * f, fg.b + (c.b - fg.b) * f, fg.a};
    }
    c_out = c;
}

// The framebuffer part of one fragment's pipeline, in the reference's order:
// stencil test, depth test, stencil update, alpha-test discard, depth write,
// blend, clamp and colour mask (src/raster.c:552-568, :651-721).  Byte
// counting follows the reference's reads and writes of each step.
// First half: stencil test, depth test, stencil update, the texel bytes the
// fragment read, the alpha-test discard (kept == false), depth write and the
// colour bytes.  Returns whether the colour step runs.
template <typename Bytes>
__device__ __forceinline__ bool depth_stencil_step(const RasterConsts &K, float depth,
                                                   uint32_t taps, bool kept, PixelState &px, Bytes &bytes)
{
    if (!K.stencil_on) {
        // no stencil: the depth test alone, written to select rather than branch
        bool pass = true;
        if (K.depth_on) {
            bytes += 4;
            pass = cmp_f(K.depth_cmp, depth, px.d);
        }
        if (pass) bytes += 4u * taps + (kept ? K.pass_bytes : 0u);
        const bool write = pass && kept;
        if (write && K.depth_write) {
            px.d = depth;
            px.dirty_d = true;
        }
        return write;
    }
    if (K.stencil_on) {
        bytes += 2;
        if (!cmp_stencil(K.stencil_cmp, K.stencil_ref, K.stencil_mask, px.s)) {
            stencil_write(K, px, K.stencil_fail);
            return false;
        }
    }
    if (K.depth_on) {
        bytes += 4;
        if (!cmp_f(K.depth_cmp, depth, px.d)) {
            if (K.stencil_on) stencil_write(K, px, K.stencil_zfail);
            return false;
        }
    }
    if (K.stencil_on) stencil_write(K, px, K.stencil_zpass);
    bytes += 4u * taps;
    if (!kept) return false;
    if (K.depth_on && K.depth_mask) {
        bytes += 4;
        px.d = depth;
        px.dirty_d = true;
    }
    if (K.cmask == 0xFu) bytes += K.blend_on ? 8 : 4;
    else if (K.cmask != 0u) bytes += 8;
    return true;
}

// Second half: blend with the pixel's colour, clamp, colour mask, pack.
__device__ __forceinline__ void color_step(const RasterConsts &K, Col c, PixelState &px)
{
    if (K.blend_on) {
        Col dst = col_unpack(px.c);
        Col sf = blend_factor(K.blend_src, c, dst);
        Col df = blend_factor(K.blend_dst, c, dst);
        c = col_clamp(col_add(col_mul(c, sf), col_mul(dst, df)));
    }
    c = col_clamp(c);
    if (K.cmask == 0xFu) {
        px.c = col_pack(c);
        px.dirty_c = true;
    } else if (K.cmask != 0u) {
        Col d = col_unpack(px.c);
        if (K.cmask & 1u) d.r = c.r;
        if (K.cmask & 2u) d.g = c.g;
        if (K.cmask & 4u) d.b = c.b;
        if (K.cmask & 8u) d.a = c.a;
        px.c = col_pack(d);
        px.dirty_c = true;
    }
}

// the reference's three edge functions at integer pixel (x, y)
// (src/raster.c:299-302, :534-540), evaluated in its operation order
__device__ __forceinline__ void edge_values(const RTri *r, int32_t x, int32_t y, float &w0e, float &w1e, float &w2e)
{
    const float px = (float)x, py = (float)y;
    const float fx0 = (float)r->x0, fy0 = (float)r->y0, fx1 = (float)r->x1, fy1 = (float)r->y1;
    const float fx2 = (float)r->x2, fy2 = (float)r->y2;
    w0e = (px - fx1) * (fy2 - fy1) - (py - fy1) * (fx2 - fx1);
    w1e = (px - fx2) * (fy0 - fy2) - (py - fy2) * (fx0 - fx2);
    w2e = (px - fx0) * (fy1 - fy0) - (py - fy0) * (fx1 - fx0);
}

constexpr int CHUNK = 64;          // bin entries per coverage pass: one bit each in a pixel's mask
constexpr int FCAP = 128;          // fragment colours held in LDS per batch (general kernel)

// what the in-order depth pass reads per bin entry: the edge-function terms
// in the reference's operation order, 1/area and the three window z
struct __align__(16) RCov {
    float fx1, fy1, d0y, d0x;      // w0e = (x - fx1) * d0y - (y - fy1) * d0x
    float fx2, fy2, d1y, d1x;      // w1e
    float fx0, fy0, d2y, d2x;      // w2e
    float inv_area, z0, z1, z2;
};

// field of a record held as float4 registers
template <int OFF>
__device__ __forceinline__ float rec_f(const float4 *q)
{
    constexpr int v = OFF / 16, c = (OFF % 16) / 4;
    return c == 0 ? q[v].x : (c == 1 ? q[v].y : (c == 2 ? q[v].z : q[v].w));
}
#define REC_F(q, field) rec_f<(int)offsetof(RTri, field)>(q)
#define REC_I(q, field) __float_as_int(rec_f<(int)offsetof(RTri, field)>(q))

// texels one fragment of a triangle with this LOD reads (the branch structure
// of sample_texture: the count depends on the triangle, not on the pixel)
__device__ __forceinline__ uint32_t texel_taps(const RasterConsts &K, float lod)
{
    if (!K.textured) return 0u;
    const uint32_t filter = lod > 0.0f ? K.min_filter : K.mag_filter;
    if (filter == GL_NEAREST_MIPMAP_NEAREST || filter == GL_LINEAR_MIPMAP_NEAREST) {
        if (lod >= 0.5f) return K.L1.px ? (filter == GL_LINEAR_MIPMAP_NEAREST ? 4u : 1u) : 0u;
        return filter == GL_LINEAR_MIPMAP_NEAREST ? 4u : 1u;
    }
    if (filter == GL_NEAREST_MIPMAP_LINEAR || filter == GL_LINEAR_MIPMAP_LINEAR) {
        const uint32_t per = filter == GL_LINEAR_MIPMAP_LINEAR ? 4u : 1u;
        return (lod > 0.0f && K.L1.px) ? 2u * per : per;
    }
    return filter == GL_LINEAR ? 4u : 1u;
}

// Orders this wave's LDS accesses before and after it.  The raster kernels
// run one wave per block, whose LDS operations complete in issue order, so
// no s_barrier (and no wait on outstanding global loads) is needed.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// The mask of the chunk's entries covering pixel (R, I): the AND of three
// ballots (each compare writes its lane mask straight to scalar registers).
// VIA_LDS: lane 0 stores it to the pixel's LDS slot and the pixel's lane
// reads it back once every pixel is done (one wave: LDS operations complete
// in order), 4 vector instructions per pixel; else the pixel's lane selects
// it from the scalar result (no LDS traffic, about twice the instructions:
// the general kernel, whose LDS is busier, keeps this form).
template <int R, int I, bool VIA_LDS>
__device__ __forceinline__ void cover_pixel(const float *a0, const float *a1, const float *a2, float b0, float b1,
                                            float b2, uint64_t *s_mask, int32_t &mlo, int32_t &mhi)
{
    const uint64_t bal = __builtin_amdgcn_ballot_w64(a0[I] >= b0) & __builtin_amdgcn_ballot_w64(a1[I] >= b1) &
                         __builtin_amdgcn_ballot_w64(a2[I] >= b2);
    if (VIA_LDS) {
        if (threadIdx.x == 0) s_mask[R * TILE + I] = bal;
    } else if ((int)threadIdx.x == R * TILE + I) {
        mlo = (int32_t)(uint32_t)bal;
        mhi = (int32_t)(uint32_t)(bal >> 32);
    }
}

template <int R, bool VIA_LDS>
__device__ __forceinline__ void cover_row(const float *a0, const float *a1, const float *a2, const float *b0,
                                          const float *b1, const float *b2, uint64_t *s_mask, int32_t &mlo,
                                          int32_t &mhi)
{
    cover_pixel<R, 0, VIA_LDS>(a0, a1, a2, b0[R], b1[R], b2[R], s_mask, mlo, mhi);
    cover_pixel<R, 1, VIA_LDS>(a0, a1, a2, b0[R], b1[R], b2[R], s_mask, mlo, mhi);
    cover_pixel<R, 2, VIA_LDS>(a0, a1, a2, b0[R], b1[R], b2[R], s_mask, mlo, mhi);
    cover_pixel<R, 3, VIA_LDS>(a0, a1, a2, b0[R], b1[R], b2[R], s_mask, mlo, mhi);
    cover_pixel<R, 4, VIA_LDS>(a0, a1, a2, b0[R], b1[R], b2[R], s_mask, mlo, mhi);
    cover_pixel<R, 5, VIA_LDS>(a0, a1, a2, b0[R], b1[R], b2[R], s_mask, mlo, mhi);
    cover_pixel<R, 6, VIA_LDS>(a0, a1, a2, b0[R], b1[R], b2[R], s_mask, mlo, mhi);
    cover_pixel<R, 7, VIA_LDS>(a0, a1, a2, b0[R], b1[R], b2[R], s_mask, mlo, mhi);
}


// Depth-pass terms of a bin entry from its first 64 record bytes.
__device__ __forceinline__ RCov make_cov(const float4 *q)
{
    RCov cv;
    const float fx0 = (float)REC_I(q, x0), fy0 = (float)REC_I(q, y0);
    const float fx1 = (float)REC_I(q, x1), fy1 = (float)REC_I(q, y1);
    const float fx2 = (float)REC_I(q, x2), fy2 = (float)REC_I(q, y2);
    cv.fx1 = fx1; cv.fy1 = fy1; cv.d0y = fy2 - fy1; cv.d0x = fx2 - fx1;
    cv.fx2 = fx2; cv.fy2 = fy2; cv.d1y = fy0 - fy2; cv.d1x = fx0 - fx2;
    cv.fx0 = fx0; cv.fy0 = fy0; cv.d2y = fy1 - fy0; cv.d2x = fx1 - fx0;
    cv.inv_area = REC_F(q, inv_area);
    cv.z0 = REC_F(q, z0); cv.z1 = REC_F(q, z1); cv.z2 = REC_F(q, z2);
    return cv;
}

// Coverage of one chunk, one lane per bin entry.  The edge functions are
// separable, w = A(x) - B(y) with A and B the reference's rounded products,
// so the lane forms 8 column and 8 row terms per edge and a pixel's test is
// three compares (A and B are finite integer-valued floats, so fl(A - B) >= 0
// exactly when A >= B).  One ballot per pixel gives the mask of entries
// covering it, kept by that pixel's lane.  Negative-area triangles are
// tested with A and B negated; the bounding box and the framebuffer edge
// fold in as -inf / +inf.
template <bool VIA_LDS>
__device__ __forceinline__ uint64_t cover_chunk(const RCov &cv, int32_t minX, int32_t minY, int32_t maxX, int32_t maxY,
                                                float area, bool valid, int32_t X0, int32_t Y0, int32_t fb_w,
                                                int32_t fb_h, int32_t r0, int32_t r1, uint64_t *s_mask)
{
    const float fx0 = cv.fx0, fy0 = cv.fy0, fx1 = cv.fx1, fy1 = cv.fy1, fx2 = cv.fx2, fy2 = cv.fy2;
    const bool live = valid && (area > 0.0f || area < 0.0f);
    const float sg = area > 0.0f ? 1.0f : -1.0f;
    // sg * fl(p * q) == fl(p * (sg * q)) for sg = +-1 (rounding to nearest is
    // symmetric), so the sign folds into the edge deltas once per lane
    const float d0y = sg * cv.d0y, d0x = sg * cv.d0x, d1y = sg * cv.d1y, d1x = sg * cv.d1x;
    const float d2y = sg * cv.d2y, d2x = sg * cv.d2x;
    float a0[TILE], a1[TILE], a2[TILE], b0[TILE], b1[TILE], b2[TILE];
#pragma unroll
    for (int i = 0; i < TILE; i++) {
        const int32_t xi = X0 + i, yi = Y0 + i;
        const float pxi = (float)xi, pyi = (float)yi;
        a0[i] = (pxi - fx1) * d0y;
        a1[i] = (pxi - fx2) * d1y;
        a2[i] = (pxi - fx0) * d2y;
        b0[i] = (pyi - fy1) * d0x;
        b1[i] = (pyi - fy2) * d1x;
        b2[i] = (pyi - fy0) * d2x;
        if (!(live && xi >= minX && xi <= maxX && xi < fb_w)) a0[i] = -__builtin_inff();
        if (!(yi >= minY && yi <= maxY && yi < fb_h)) b0[i] = __builtin_inff();
    }
    // only the rows [r0, r1) of the item (a part of a heavy tile, or all)
    int32_t mlo = 0, mhi = 0;
    if (r0 <= 0 && 0 < r1) cover_row<0, VIA_LDS>(a0, a1, a2, b0, b1, b2, s_mask, mlo, mhi);
    if (r0 <= 1 && 1 < r1) cover_row<1, VIA_LDS>(a0, a1, a2, b0, b1, b2, s_mask, mlo, mhi);
    if (r0 <= 2 && 2 < r1) cover_row<2, VIA_LDS>(a0, a1, a2, b0, b1, b2, s_mask, mlo, mhi);
    if (r0 <= 3 && 3 < r1) cover_row<3, VIA_LDS>(a0, a1, a2, b0, b1, b2, s_mask, mlo, mhi);
    if (r0 <= 4 && 4 < r1) cover_row<4, VIA_LDS>(a0, a1, a2, b0, b1, b2, s_mask, mlo, mhi);
    if (r0 <= 5 && 5 < r1) cover_row<5, VIA_LDS>(a0, a1, a2, b0, b1, b2, s_mask, mlo, mhi);
    if (r0 <= 6 && 6 < r1) cover_row<6, VIA_LDS>(a0, a1, a2, b0, b1, b2, s_mask, mlo, mhi);
    if (r0 <= 7 && 7 < r1) cover_row<7, VIA_LDS>(a0, a1, a2, b0, b1, b2, s_mask, mlo, mhi);
    if (!VIA_LDS) return ((uint64_t)(uint32_t)mhi << 32) | (uint32_t)mlo;
    wave_sync();
    const int32_t row = (int32_t)(threadIdx.x >> 3);
    const uint64_t m = row >= r0 && row < r1 ? s_mask[threadIdx.x] : 0ull;
    wave_sync();   // read before the next chunk's masks are stored
    return m;
}

// Depth and stencil of this lane's pixel for the entries in m, in
// submission order; returns the entries that passed (the ones whose colour
// step runs).  Valid when no alpha test reads the texture.
__device__ __forceinline__ uint64_t depth_pass(const RasterConsts &K, const RCov *s_cov, const uint32_t *s_tap,
                                               uint64_t m, float xf, float yf, PixelState &px,
                                               unsigned long long &bytes)
{
    uint64_t pass = 0;
    while (m) {
        const uint32_t k = (uint32_t)__builtin_ctzll(m);
        m &= m - 1;
        const RCov &cv = s_cov[k];
        const float w0e = (xf - cv.fx1) * cv.d0y - (yf - cv.fy1) * cv.d0x;
        const float w1e = (xf - cv.fx2) * cv.d1y - (yf - cv.fy2) * cv.d1x;
        const float w2e = (xf - cv.fx0) * cv.d2y - (yf - cv.fy0) * cv.d2x;
        const float bb0 = w0e * cv.inv_area, bb1 = w1e * cv.inv_area, bb2 = w2e * cv.inv_area;
        const float z = bb0 * cv.z0 + bb1 * cv.z1 + bb2 * cv.z2;
        float depth;
        if (K.unit_range) depth = (z + 1.0f) * 0.5f;
        else depth = (float)((double)((z + 1.0f) * 0.5f) * (K.depth_far - K.depth_near) + K.depth_near);
        if (depth_stencil_step(K, depth, s_tap[k], true, px, bytes)) pass |= 1ull << k;
    }
    return pass;
}

// The raster kernels run a persistent grid over a list of work items: items
// [0, nheavy) are parts (TILE / HEAVY_SPLIT pixel rows) of the tiles whose bins hold more
// than the heavy-bin threshold, four per tile, and whole tiles with more than
// MED_BIN entries, at the front of the list; the other items are whole
// non-empty tiles, from the back of the list, so the longest items start
// first.  Pass k of block b takes item k * gridDim.x + b
// (even k) or k * gridDim.x + gridDim.x - 1 - b (odd k).  Grids hold up to
// three times the resident blocks (six for the general kernel), so while
// items last the dispatcher hands
// each freed slot the next block: a slot busy with a long item takes no
// other.
__device__ __forceinline__ void list_counts(const RasterParams &P, uint32_t &nh, uint32_t &it0, uint32_t &nt)
{
    nh = __builtin_amdgcn_readfirstlane(*(volatile uint32_t *)&P.ctr->nheavy);
    nt = nh + __builtin_amdgcn_readfirstlane(*(volatile uint32_t *)&P.ctr->nlight);
    nt = min(nt, P.list_cap);
    nh = min(nh, nt);
    it0 = P.part == 2u ? nh : 0u;
    if (P.part == 1u) nt = nh;
}

// The dispatcher deals consecutive blocks of a grid to the 8 XCDs in turn,
// each with its own L2.  Blocks are renumbered so that runs of XCD_RUN
// consecutive list items go to blocks of one XCD, the runs to the XCDs in
// turn: neighbouring tiles, whose pixel rows share cache lines and whose
// bins share triangle records, then meet in one L2, and the long items at the
// front of the list spread over every XCD.  Grids that are not a multiple of
// NXCD * XCD_RUN blocks keep their numbering.
constexpr uint32_t NXCD = 8;
constexpr uint32_t XCD_RUN = 8;

__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t G)
{
    if (G % (NXCD * XCD_RUN) != 0u) return b;
    const uint32_t x = b % NXCD, i = b / NXCD;
    return ((i / XCD_RUN) * NXCD + x) * XCD_RUN + i % XCD_RUN;
}

__device__ __forceinline__ uint32_t list_item(uint32_t it0, uint32_t k, uint32_t b, uint32_t G)
{
    const uint32_t v = xcd_block(b, G);
    return it0 + k * G + ((k & 1u) ? G - 1u - v : v);
}


// list item -> scene, the scene-local tile, the part (0 whole tile, q = 1..HEAVY_SPLIT
// pixel rows [(q - 1) * PART_ROWS, q * PART_ROWS)) and the tile's bin [e0, e1)
__device__ __forceinline__ uint4 load_item(const RasterParams &P, uint32_t it, uint32_t nh)
{
    return P.list[it < nh ? it : P.list_cap - 1u - (it - nh)];
}

__device__ __forceinline__ void decode_item(uint4 v, uint32_t &sc, uint32_t &tile, uint32_t &e0, uint32_t &e1,
                                            uint32_t &quarter)
{
    const uint32_t g = __builtin_amdgcn_readfirstlane(v.x);
    quarter = g >> ITEM_TILE_BITS;
    tile = g & ((1u << ITEM_TILE_BITS) - 1u);
    e0 = __builtin_amdgcn_readfirstlane(v.y);
    e1 = __builtin_amdgcn_readfirstlane(v.z);
    sc = __builtin_amdgcn_readfirstlane(v.w);
}

// Items at the front of the list (heavy-tile parts and medium tiles) set
// the length of a launch: their waves take issue priority over the waves of
// light items that share their SIMD.
__device__ __forceinline__ void item_priority(uint32_t it, uint32_t nh)
{
    if (it < nh) __builtin_amdgcn_s_setprio(2);
    else __builtin_amdgcn_s_setprio(0);
}

__device__ __forceinline__ void item_tile(const RasterParams &P, uint32_t it, uint32_t nh, uint32_t &sc, uint32_t &tile,
                                          uint32_t &e0, uint32_t &e1, uint32_t &quarter)
{
    decode_item(load_item(P, it, nh), sc, tile, e0, e1, quarter);
}

__device__ __forceinline__ void raster_consts(const SceneRaster &P, RasterConsts &K)
{
    const mtgl_raster_state *rs = P.rs;
    K.rs = rs;
    const uint32_t flags = rs->flags;
    K.depth_on = flags & MTGL_F_DEPTH_TEST;
    K.stencil_on = flags & MTGL_F_STENCIL_TEST;
    K.blend_on = flags & MTGL_F_BLEND;
    K.fog_on = flags & MTGL_F_FOG;
    K.alpha_on = flags & MTGL_F_ALPHA_TEST;
    K.lighting_on = flags & MTGL_F_LIGHTING;
    K.cmask = rs->color_mask & 0xFu;
    K.textured = P.tex != nullptr;
    K.persp = rs->perspective_hint != GL_FASTEST;
    K.flat = rs->shade_model == GL_FLAT;
    K.phong = rs->shade_model == GL_PHONG;
    K.unit_range = rs->depth_near == 0.0 && rs->depth_far == 1.0;
    K.two_side = rs->lighting.two_side != 0;
    K.rep_s = rs->wrap_s == GL_REPEAT;
    K.rep_t = rs->wrap_t == GL_REPEAT;
    K.depth_mask = rs->depth_mask != 0;
    K.depth_write = K.depth_on && K.depth_mask;
    K.pass_bytes = (K.depth_write ? 4u : 0u) + (K.cmask == 0xFu ? (K.blend_on ? 8u : 4u) : (K.cmask ? 8u : 0u));
    K.depth_cmp = cmp_mask_of(rs->depth_func);
    K.stencil_cmp = cmp_mask_of(rs->stencil_func);
    K.stencil_mask = rs->stencil_mask;
    K.stencil_writemask = rs->stencil_writemask;
    K.stencil_fail = rs->stencil_fail;
    K.stencil_zfail = rs->stencil_zfail;
    K.stencil_zpass = rs->stencil_zpass;
    K.stencil_ref = rs->stencil_ref;
    K.blend_src = rs->blend_src;
    K.blend_dst = rs->blend_dst;
    K.alpha_cmp = cmp_mask_of(rs->alpha_func);
    K.alpha_ref = rs->alpha_ref;
    K.tex_env_mode = rs->tex_env_mode;
    K.tex_env_color = col_of(rs->tex_env_color);
    K.fog_mode = rs->fog_mode;
    K.fog_density = rs->fog_density;
    K.fog_start = rs->fog_start;
    K.fog_end = rs->fog_end;
    K.fog_color = col_of(rs->fog_color);
    K.min_filter = rs->min_filter;
    K.mag_filter = rs->mag_filter;
    K.depth_near = rs->depth_near;
    K.depth_far = rs->depth_far;
    K.L0 = {P.tex, P.tex_w, P.tex_h, P.tex_w > 0 && (P.tex_w & (P.tex_w - 1)) == 0 && (P.tex_h & (P.tex_h - 1)) == 0};
    K.L1 = {P.tex1, P.tex1_w, P.tex1_h,
            P.tex1_w > 0 && (P.tex1_w & (P.tex1_w - 1)) == 0 && (P.tex1_h & (P.tex1_h - 1)) == 0};
    K.chain_on = P.lod_mode == MTGL_LOD_FULL_CHAIN && P.tex != nullptr;
    K.chain = P.chain;
    K.nlev = P.nlev > 0 ? P.nlev : 1;
    if (K.stencil_on) K.depth_mode = DEPTH_SERIAL;
    else if (!K.depth_write) K.depth_mode = DEPTH_IND;
    else if (rs->depth_func == GL_LESS || rs->depth_func == GL_LEQUAL) K.depth_mode = DEPTH_MIN;
    else if (rs->depth_func == GL_GREATER || rs->depth_func == GL_GEQUAL) K.depth_mode = DEPTH_MAX;
    else if (rs->depth_func == GL_NOTEQUAL) K.depth_mode = DEPTH_SERIAL;
    else K.depth_mode = DEPTH_IND;
}

// Per-lane counts summed over the wave (bytes as two 32-bit halves) and
// added to this block's own slot: no two blocks of a launch share a slot and
// launches on the stream are ordered, so plain adds suffice (thousands of
// same-address atomics would queue at one memory channel and stall every
// load that maps there).
__device__ __forceinline__ void add_stats(const RasterParams &P, unsigned long long frags, unsigned long long bytes)
{
    const uint32_t f = dpp_wave_sum((uint32_t)frags);
    const unsigned long long blo = dpp_wave_sum((uint32_t)(bytes & 0xFFFFFFFFull));
    const unsigned long long bhi = dpp_wave_sum((uint32_t)(bytes >> 32));
    if (threadIdx.x == 0 && f && blockIdx.x < STAT_SLOTS) {
        unsigned long long *slot = P.stats + 2u * blockIdx.x;
        slot[0] += (unsigned long long)f;
        slot[1] += blo + (bhi << 32);
    }
}

constexpr int DCAP = 512;          // fragment depths held in LDS per batch (depth kernel at 4 waves per SIMD)

// K4a, the depth kernel of draws without blending, alpha test on a texture or
// a partial colour mask: when nothing blends, only the last fragment that
// passes a pixel can leave its colour there, so this kernel runs coverage and
// the in-order depth/stencil pass, writes depth and stencil, and records per
// pixel 1 + the record slot of that last fragment (0: none), then shades that
// one fragment per pixel (barycentrics, colour, lighting, texture, fog) from
// its record.  One 8x8 tile per one-wave block; the depth pass reads only the
// first 64 bytes of each record.  WPE waves per SIMD: with three the shading
// keeps its registers (one scene's launch is faster so); with four a few
// spill, which a launch of many small scenes gains back in latency hiding.
// Per chunk: coverage (cover_chunk); then batches of at most DCAP_W fragments
// in pixel-major order: the fragments' window depths one per lane, and per
// pixel lane the depth/stencil steps in submission order over its own
// fragments, which only load, compare and select.
template <bool STENCIL, int WPE, bool SIMPLE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE))) k_raster_depth(RasterParams P)
{
    // fragments per batch: at three waves per SIMD LDS has room for more
    constexpr int DCAP_W = WPE == 3 ? 2 * DCAP : DCAP;
    __shared__ RCov s_cov[CHUNK];
    __shared__ uint64_t s_mask[TILE * TILE];   // coverage masks of a chunk, per pixel
    __shared__ uint32_t s_rid[CHUNK];
    __shared__ uint32_t s_tap[CHUNK];
    __shared__ float s_dep[DCAP_W];
    __shared__ uint16_t s_slot[DCAP_W];   // taps << 12 | entry << 6 | pixel

    if (exec_poisoned(P.ctr)) return;   // the launch runs again once its arrays have grown
    RasterConsts K;
    const int lane = threadIdx.x;
    uint32_t nh, it0, nt;
    list_counts(P, nh, it0, nt);
    uint32_t frag_count = 0;
    unsigned long long frag_bytes = 0;
    // the next item's header is loaded while the current item is processed
    uint32_t it = list_item(it0, 0u, blockIdx.x, gridDim.x);
    uint4 hdr = it < nt ? load_item(P, it, nh) : make_uint4(0u, 0u, 0u, 0u);
    for (uint32_t pass = 0; it < nt; pass++) {
    const uint32_t it_next = list_item(it0, pass + 1u, blockIdx.x, gridDim.x);
    const uint4 hdr_next = it_next < nt ? load_item(P, it_next, nh) : make_uint4(0u, 0u, 0u, 0u);
    uint32_t sc, tile, e0, e1, quarter;
    decode_item(hdr, sc, tile, e0, e1, quarter);
    item_priority(it, nh);
    hdr = hdr_next;
    it = it_next;
    const SceneRaster &S = P.scenes[sc];
    const uint32_t kind = __builtin_amdgcn_readfirstlane(S.kind);
    if (!(kind & KIND_LAST_WRITER) || ((kind & KIND_STENCIL) != 0u) != STENCIL) continue;
    const int32_t r0 = quarter ? ((int32_t)quarter - 1) * PART_ROWS : 0, r1 = quarter ? r0 + PART_ROWS : TILE;
    // the scene's constants are read for every item (held across items they
    // spill scalar registers on every pass), and again for the shading below
    raster_consts(S, K);
    K.stencil_on = STENCIL;   // the instantiation matches the scene's state
    const int32_t X0 = (int32_t)(tile % (uint32_t)S.tiles_x) * TILE, Y0 = (int32_t)(tile / (uint32_t)S.tiles_x) * TILE;
    const int32_t x = X0 + (lane & 7), y = Y0 + (lane >> 3);
    const int32_t fb_w = S.fb_w, fb_h = S.fb_h;
    const bool in_fb = (lane >> 3) >= r0 && (lane >> 3) < r1 && x < fb_w && y < fb_h;   // this item's pixel
    const size_t pix = in_fb ? (size_t)y * fb_w + x : 0;
    PixelState px = {0u, 1.0f, 0u, false, false, false};
    if (in_fb) {
        if (K.depth_on) px.d = S.depth[pix];
        if (K.stencil_on) px.s = S.stencil[pix];
    }
    uint32_t writer = 0;
    // the next chunk's entries (record index, then the record's first 64
    // bytes) are loaded while the current one is processed
    float4 qn[RT_COV_VEC4];
    uint32_t ridn = 0;
    {
        const bool v = (uint32_t)lane < e1 - e0;
        const uint32_t e = v ? e0 + lane : e0;
        ridn = P.vals[e];
#pragma unroll
        for (int i = 0; i < RT_COV_VEC4; i++) qn[i] = reinterpret_cast<const float4 *>(P.rt + ridn)[i];
    }
    if (quarter != 0u && K.depth_mode != DEPTH_SERIAL) {
        // One pixel row of a heavy tile, the depth pass with one bin entry
        // per lane.  Per chunk of 64 entries in submission order and per
        // pixel of the row: each lane tests its entry's coverage and window
        // depth, and the fragment's test against the depth before it is
        // resolved across lanes: in DEPTH_MIN / DEPTH_MAX mode that depth is
        // the running minimum / maximum of the pixel's depth and every earlier
        // covered fragment's (a fragment that fails changes nothing; NaN
        // depths are passed over and fail), so an exclusive prefix min / max
        // gives every fragment's test exactly as the in-order loop does; in
        // DEPTH_IND mode every test reads the depth before the chunk.  The
        // pixel then holds the depth of its last passing fragment.
        const int32_t yrow = Y0 + r0;
        const float yf = (float)yrow;
        const int own0 = TILE * r0;   // lane of the row's first pixel
        for (uint32_t c0 = e0; c0 < e1; c0 += CHUNK) {
            const uint32_t cn = min((uint32_t)CHUNK, e1 - c0);
            const bool valid = (uint32_t)lane < cn;
            float4 q[RT_COV_VEC4];
#pragma unroll
            for (int i = 0; i < RT_COV_VEC4; i++) q[i] = qn[i];
            const uint32_t rid = ridn;
            if (c0 + CHUNK < e1) {
                const bool v = (uint32_t)lane < e1 - (c0 + CHUNK);
                const uint32_t e = v ? c0 + CHUNK + lane : c0 + CHUNK;
                ridn = P.vals[e];
#pragma unroll
                for (int i = 0; i < RT_COV_VEC4; i++) qn[i] = reinterpret_cast<const float4 *>(P.rt + ridn)[i];
            }
            const RCov cv = make_cov(q);
            const float area = REC_F(q, area);
            const int32_t minX = REC_I(q, minX), maxX = REC_I(q, maxX);
            const bool live = valid && (area > 0.0f || area < 0.0f) && yrow >= REC_I(q, minY) &&
                              yrow <= REC_I(q, maxY) && yrow < fb_h;
            const float sg = area > 0.0f ? 1.0f : -1.0f;
            const uint32_t taps = texel_taps(K, REC_F(q, lod));
            const uint32_t pass_add = 4u * taps + K.pass_bytes;
            for (int j = 0; j < TILE; j++) {
                const int32_t xj = X0 + j;
                if (xj >= fb_w) break;
                const float xf = (float)xj;
                const float w0e = (xf - cv.fx1) * cv.d0y - (yf - cv.fy1) * cv.d0x;
                const float w1e = (xf - cv.fx2) * cv.d1y - (yf - cv.fy2) * cv.d1x;
                const float w2e = (xf - cv.fx0) * cv.d2y - (yf - cv.fy0) * cv.d2x;
                const bool inside = live && xj >= minX && xj <= maxX && sg * w0e >= 0.0f && sg * w1e >= 0.0f &&
                                    sg * w2e >= 0.0f;
                const float bb0 = w0e * cv.inv_area, bb1 = w1e * cv.inv_area, bb2 = w2e * cv.inv_area;
                const float z = bb0 * cv.z0 + bb1 * cv.z1 + bb2 * cv.z2;
                float depth;
                if (K.unit_range) depth = (z + 1.0f) * 0.5f;
                else depth = (float)((double)((z + 1.0f) * 0.5f) * (K.depth_far - K.depth_near) + K.depth_near);
                // the pixel's depth before this chunk, from the lane that holds it
                const float dcur = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(px.d), own0 + j));
                bool pass;
                if (K.depth_mode == DEPTH_IND) {
                    pass = inside && (!K.depth_on || cmp_f(K.depth_cmp, depth, dcur));
                } else {
                    const bool mx = K.depth_mode == DEPTH_MAX;
                    const float idf = mx ? -__builtin_inff() : __builtin_inff();
                    const float incl = mx ? dpp_incl_minmax<true>(inside ? depth : idf)
                                          : dpp_incl_minmax<false>(inside ? depth : idf);
                    float excl = __shfl_up(incl, 1);
                    if (lane == 0) excl = idf;
                    const float thr = mx ? fmaxf(dcur, excl) : fminf(dcur, excl);
                    // a pixel holding NaN passes nothing: every test against it fails
                    pass = inside && !__builtin_isnan(dcur) && cmp_f(K.depth_cmp, depth, thr);
                }
                frag_count += inside ? 1u : 0u;
                frag_bytes += (inside && K.depth_on ? 4u : 0u) + (pass ? pass_add : 0u);
                const uint64_t pm = __ballot(pass);
                if (pm) {
                    const int lp = 63 - __builtin_clzll(pm);
                    const float dl = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(depth), lp));
                    const uint32_t wr = 1u + (uint32_t)__builtin_amdgcn_readlane((int)rid, lp);
                    if (lane == own0 + j) {
                        if (K.depth_write) {
                            px.d = dl;
                            px.dirty_d = true;
                        }
                        writer = wr;
                    }
                }
            }
        }
    } else
    for (uint32_t c0 = e0; c0 < e1; c0 += CHUNK) {
        const uint32_t cn = min((uint32_t)CHUNK, e1 - c0);
        const bool valid = (uint32_t)lane < cn;
        float4 q[RT_COV_VEC4];
#pragma unroll
        for (int i = 0; i < RT_COV_VEC4; i++) q[i] = qn[i];
        const uint32_t rid = ridn;
        if (c0 + CHUNK < e1) {
            const bool v = (uint32_t)lane < e1 - (c0 + CHUNK);
            const uint32_t e = v ? c0 + CHUNK + lane : c0 + CHUNK;
            ridn = P.vals[e];
#pragma unroll
            for (int i = 0; i < RT_COV_VEC4; i++) qn[i] = reinterpret_cast<const float4 *>(P.rt + ridn)[i];
        }
        const RCov cv = make_cov(q);
        wave_sync();   // the previous chunk's entries are no longer read
        if (valid) {
            s_cov[lane] = cv;
            s_rid[lane] = rid;
            s_tap[lane] = texel_taps(K, REC_F(q, lod));
        }
        const uint64_t m = cover_chunk<true>(cv, REC_I(q, minX), REC_I(q, minY), REC_I(q, maxX), REC_I(q, maxY),
                                       REC_F(q, area), valid, X0, Y0, fb_w, fb_h, r0, r1, s_mask);
        const uint32_t cnt = (uint32_t)__popcll(m);
        frag_count += cnt;
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int32_t)dpp_incl_scan(cnt), 63);
        uint64_t rem = m;
        int32_t last = -1;   // entry of this pixel's last passing fragment in the chunk
        bool first = true;
        while (total && (first || __builtin_amdgcn_ballot_w64(rem != 0ull))) {
            // each pixel takes its next L fragments, L the largest quota that
            // keeps the batch within DCAP_W (64 always fits)
            uint32_t L = 64;
            if (!first || total > (uint32_t)DCAP_W) {
                const uint32_t rc = (uint32_t)__popcll(rem);
                L = 1;
#pragma unroll
                for (uint32_t step = 32; step >= 1; step >>= 1) {
                    const uint32_t sum = (uint32_t)__builtin_amdgcn_readlane((int32_t)dpp_incl_scan(min(rc, L + step)), 63);
                    if (sum <= (uint32_t)DCAP_W) L += step;
                }
            }
            first = false;
            const uint32_t take = min((uint32_t)__popcll(rem), L);
            const uint32_t ij = dpp_incl_scan(take);
            const uint32_t off = ij - take;
            const uint32_t fr = min((uint32_t)__builtin_amdgcn_readlane((int32_t)ij, 63), (uint32_t)DCAP_W);
            for (uint32_t o = off, qn_ = 0; qn_ < take; qn_++, o++) {
                const uint32_t k = (uint32_t)__builtin_ctzll(rem);
                rem &= rem - 1;
                if (o < (uint32_t)DCAP_W) s_slot[o] = (uint16_t)((k << 6) | (uint32_t)lane);
            }
            wave_sync();
            // window depth of every fragment of the batch, one per lane
            for (uint32_t f0 = 0; f0 < fr; f0 += 64) {
                const uint32_t f = f0 + (uint32_t)lane;
                if (f < fr) {
                    const uint32_t sl = s_slot[f];
                    const uint32_t k = (sl >> 6) & 63u, j = sl & 63u;
                    const RCov &c = s_cov[k];
                    const float fxp = (float)(X0 + (int32_t)(j & 7u)), fyp = (float)(Y0 + (int32_t)(j >> 3));
                    const float w0e = (fxp - c.fx1) * c.d0y - (fyp - c.fy1) * c.d0x;
                    const float w1e = (fxp - c.fx2) * c.d1y - (fyp - c.fy2) * c.d1x;
                    const float w2e = (fxp - c.fx0) * c.d2y - (fyp - c.fy0) * c.d2x;
                    const float bb0 = w0e * c.inv_area, bb1 = w1e * c.inv_area, bb2 = w2e * c.inv_area;
                    const float z = bb0 * c.z0 + bb1 * c.z1 + bb2 * c.z2;
                    float depth;
                    if (K.unit_range) depth = (z + 1.0f) * 0.5f;
                    else depth = (float)((double)((z + 1.0f) * 0.5f) * (K.depth_far - K.depth_near) + K.depth_near);
                    s_dep[f] = depth;
                    s_slot[f] = (uint16_t)((s_tap[k] << 12) | sl);
                }
            }
            wave_sync();
            // per pixel, in submission order; the LDS reads of four steps are
            // issued together, ahead of the dependent compares
            const uint32_t nst = off < (uint32_t)DCAP_W ? min(take, (uint32_t)DCAP_W - off) : 0u;
            uint32_t bytes32 = 0;
            uint32_t qn_ = 0;
            for (; qn_ + 4u <= nst; qn_ += 4u) {
                float dd[4];
                uint32_t ss[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    dd[u] = s_dep[off + qn_ + u];
                    ss[u] = s_slot[off + qn_ + u];
                }
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (depth_stencil_step(K, dd[u], ss[u] >> 12, true, px, bytes32)) last = (int32_t)((ss[u] >> 6) & 63u);
            }
            for (; qn_ < nst; qn_++) {
                const uint32_t f = off + qn_;
                const float depth = s_dep[f];
                const uint32_t sl = s_slot[f];
                if (depth_stencil_step(K, depth, sl >> 12, true, px, bytes32)) last = (int32_t)((sl >> 6) & 63u);
            }
            frag_bytes += bytes32;   // a batch gives a pixel at most 64 fragments of at most 50 bytes
            wave_sync();   // slots and depths are reused by the next batch
        }
        if (last >= 0) writer = 1u + s_rid[last];
    }
    {
        // the colour of the pixel's last passing fragment, from its record
        // (its loads are issued before this item's stores, which they would
        // otherwise wait behind)
        // the constants again for the shading (a compiler barrier keeps the
        // loads here): those only the shading reads are not held through the
        // depth pass, which takes 18 VGPRs off the kernel
        __asm__ volatile("" ::: "memory");
        raster_consts(S, K);
        K.stencil_on = STENCIL;
        if (SIMPLE) {
            // no scene of the launch lights per fragment or samples the full
            // chain (the launch checked them): those paths compile away, 16
            // fewer VGPRs, more blocks resident
            K.phong = false;
            K.two_side = false;
            K.chain_on = false;
        }
        const bool shade = in_fb && (kind & KIND_COLOR) && writer;
        uint32_t packed = 0;
        if (shade) {
            const RTri *r = P.rt + (writer - 1u);
            float w0e, w1e, w2e;
            edge_values(r, x, y, w0e, w1e, w2e);
            const float inv = r->inv_area;
            Col c;
            float dep;
            uint32_t meta;
            fragment_source(K, r, P.rcold ? P.rcold + (writer - 1u) : nullptr, w0e * inv, w1e * inv, w2e * inv, c, dep,
                            meta);
            packed = col_pack(col_clamp(c));
        }
        if (in_fb) {
            if (px.dirty_d) S.depth[pix] = px.d;
            if (px.dirty_s) S.stencil[pix] = (uint8_t)px.s;
            if (S.writer) S.writer[pix] = writer;
            if (shade) S.color[pix] = packed;
        }
    }
    }
    add_stats(P, frag_count, frag_bytes);
}

// K4, general: blending, a partial colour mask, or an alpha test on a
// texture.  One 8x8 tile per one-wave block.  Per chunk of 64 bin entries:
// coverage (cover_chunk); unless the alpha test decides depth writes, the
// in-order depth/stencil pass; then colours for the fragments whose colour
// step runs (all covered ones under an alpha test on a texture), one
// fragment per lane in batches of at most FCAP, applied per pixel in
// submission order.  The colour step reads each fragment's record from
// global memory (L1/L2: a chunk's fragments share few records) rather than
// staging the chunk's 176-byte records in LDS: with 7.5 KB of LDS and at
// most 168 VGPRs (a few spill) three waves share a SIMD instead of two (the
// 4K overdraw stage took 1.41 ms so, 1.57 with the records in LDS).  PLAIN:
// every scene of the launch that this kernel draws has no texture and no
// per-fragment lighting (Phong or two-sided), so those paths compile away:
// 124 VGPRs, no spills and four waves per SIMD (that stage: 0.94 ms).
template <bool PLAIN>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PLAIN ? 4 : 3))) k_raster(RasterParams P)
{
    __shared__ RCov s_cov[CHUNK];
    __shared__ uint64_t s_mask[TILE * TILE];   // coverage masks of a chunk, per pixel
    __shared__ uint32_t s_rid[CHUNK];
    __shared__ uint32_t s_tap[CHUNK];
    __shared__ float4 s_col[FCAP];
    __shared__ float s_dep[FCAP];
    __shared__ uint16_t s_slot[FCAP];
    __shared__ uint8_t s_meta[FCAP];

    if (exec_poisoned(P.ctr)) return;   // the launch runs again once its arrays have grown
    const int lane = threadIdx.x;
    uint32_t nh, it0, nt;
    list_counts(P, nh, it0, nt);
    unsigned long long frag_count = 0, frag_bytes = 0;
    for (uint32_t pass = 0;; pass++) {
    const uint32_t it = list_item(it0, pass, blockIdx.x, gridDim.x);
    if (it >= nt) break;
    uint32_t sc, tile, e0, e1, quarter;
    item_tile(P, it, nh, sc, tile, e0, e1, quarter);
    item_priority(it, nh);
    const SceneRaster &S = P.scenes[sc];
    const uint32_t kind = __builtin_amdgcn_readfirstlane(S.kind);
    if (kind & KIND_LAST_WRITER) continue;
    const int32_t r0 = quarter ? ((int32_t)quarter - 1) * PART_ROWS : 0, r1 = quarter ? r0 + PART_ROWS : TILE;
    // the scene's constants are read again for every item: held across
    // items they spill scalar registers on every pass
    RasterConsts K;
    raster_consts(S, K);
    if (PLAIN) {
        // what these scenes hold already (the launch checked them), as
        // constants; phong and two_side only matter with lighting on
        K.textured = false;
        K.chain_on = false;
        K.phong = false;
        K.two_side = false;
    }
    const bool need_color = K.blend_on || (K.cmask != 0xFu && K.cmask != 0u);
    // the alpha test reads the texture, and its outcome gates the depth write
    const bool alpha_dep = K.alpha_on && K.textured;
    const int32_t X0 = (int32_t)(tile % (uint32_t)S.tiles_x) * TILE, Y0 = (int32_t)(tile / (uint32_t)S.tiles_x) * TILE;
    const int32_t x = X0 + (lane & 7), y = Y0 + (lane >> 3);
    const float xf = (float)x, yf = (float)y;
    const int32_t fb_w = S.fb_w, fb_h = S.fb_h;
    const bool in_fb = (lane >> 3) >= r0 && (lane >> 3) < r1 && x < fb_w && y < fb_h;   // this item's pixel
    const size_t pix = in_fb ? (size_t)y * fb_w + x : 0;
    PixelState px = {0u, 1.0f, 0u, false, false, false};
    if (in_fb) {
        if (need_color) px.c = S.color[pix];
        if (K.depth_on) px.d = S.depth[pix];
        if (K.stencil_on) px.s = S.stencil[pix];
    }

    // the next chunk's record indices are loaded while the current chunk is
    // processed (its records' first 64 bytes then take one memory wait, not
    // two)
    uint32_t ridn = (uint32_t)lane < e1 - e0 ? P.vals[e0 + lane] : 0u;
    for (uint32_t c0 = e0; c0 < e1; c0 += CHUNK) {
        const uint32_t cn = min((uint32_t)CHUNK, e1 - c0);
        const bool valid = (uint32_t)lane < cn;
        wave_sync();   // the previous chunk's records are no longer read
        float4 q[RT_COV_VEC4];
#pragma unroll
        for (int i = 0; i < RT_COV_VEC4; i++) q[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        uint32_t rid = 0;
        const uint32_t rid_here = ridn;
        if (c0 + CHUNK < e1) ridn = (uint32_t)lane < e1 - (c0 + CHUNK) ? P.vals[c0 + CHUNK + lane] : 0u;
        if (valid) {
            rid = rid_here;
            const float4 *src = reinterpret_cast<const float4 *>(&P.rt[rid]);
#pragma unroll
            for (int i = 0; i < RT_COV_VEC4; i++) q[i] = src[i];
        }
        const RCov cv = make_cov(q);
        if (valid) {
            s_cov[lane] = cv;
            s_rid[lane] = rid;
            s_tap[lane] = texel_taps(K, REC_F(q, lod));
        }
        const uint64_t m = cover_chunk<false>(cv, REC_I(q, minX), REC_I(q, minY), REC_I(q, maxX), REC_I(q, maxY),
                                       REC_F(q, area), valid, X0, Y0, fb_w, fb_h, r0, r1, s_mask);
        frag_count += (unsigned long long)__popcll(m);
        wave_sync();   // staged entries visible to every lane

        const uint64_t todo = alpha_dep ? m : depth_pass(K, s_cov, s_tap, m, xf, yf, px, frag_bytes);
        // batches: each pixel takes its next L fragments, L the largest quota
        // that keeps the batch within FCAP (L >= 1 fits since 64 <= FCAP)
        uint64_t rem = todo;
        while (__builtin_amdgcn_ballot_w64(rem != 0ull)) {
            const uint32_t rc = (uint32_t)__popcll(rem);
            uint32_t L = 1;
#pragma unroll
            for (uint32_t step = 32; step >= 1; step >>= 1)
                if (dpp_wave_sum(min(rc, L + step)) <= (uint32_t)FCAP) L += step;
            const uint32_t take = min(rc, L);
            const uint32_t ij = dpp_incl_scan(take);
            const uint32_t off = ij - take;
            const uint32_t fr = min((uint32_t)__builtin_amdgcn_readlane((int32_t)ij, 63), (uint32_t)FCAP);
            {
                uint32_t o = off;
                for (uint32_t qn = 0; qn < take; qn++) {
                    const uint32_t k = (uint32_t)__builtin_ctzll(rem);
                    rem &= rem - 1;
                    if (o < (uint32_t)FCAP) s_slot[o] = (uint16_t)((k << 6) | (uint32_t)lane);
                    o++;
                }
            }
            wave_sync();
            for (uint32_t f0 = 0; f0 < fr; f0 += 64) {
                const uint32_t f = f0 + (uint32_t)lane;
                if (f < fr) {
                    const uint32_t sl = s_slot[f];
                    const uint32_t k = sl >> 6, j = sl & 63u;
                    const RTri *r = P.rt + s_rid[k];
                    float w0e, w1e, w2e;
                    edge_values(r, X0 + (int32_t)(j & 7u), Y0 + (int32_t)(j >> 3), w0e, w1e, w2e);
                    const float inv = r->inv_area;
                    Col c;
                    float dep;
                    uint32_t meta;
                    fragment_source(K, r, P.rcold ? P.rcold + s_rid[k] : nullptr, w0e * inv, w1e * inv, w2e * inv, c,
                                    dep, meta);
                    s_col[f] = make_float4(c.r, c.g, c.b, c.a);
                    s_dep[f] = dep;
                    s_meta[f] = (uint8_t)meta;
                }
            }
            wave_sync();
            for (uint32_t qn = 0; qn < take && off + qn < (uint32_t)FCAP; qn++) {
                const uint32_t f = off + qn;
                const float4 cc = s_col[f];
                const uint32_t meta = s_meta[f];
                if (!alpha_dep || depth_stencil_step(K, s_dep[f], meta >> 1, (meta & 1u) != 0u, px, frag_bytes))
                    color_step(K, Col{cc.x, cc.y, cc.z, cc.w}, px);
            }
            wave_sync();   // slots and colours are reused by the next batch
        }
    }
    if (in_fb) {
        if (px.dirty_c) S.color[pix] = px.c;
        if (px.dirty_d) S.depth[pix] = px.d;
        if (px.dirty_s) S.stencil[pix] = (uint8_t)px.s;
    }
    }
    add_stats(P, frag_count, frag_bytes);
}

// ---------------------------------------------------------------------------
// K5 clear, K6 level-1 box filter
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// K8: points, lines and glPolygonMode LINE / POINT (reference
// src/raster.c:65-296, :776-898, :1019-1196).  One thread per primitive
// (point, segment or assembled triangle) walks the reference's loops twice:
// first counting its fragments, then, at its offset in the scanned counts,
// writing each fragment's pixel, depth and colour in the reference's order.
// A stable sort by pixel keeps that order per pixel; one thread per pixel
// then runs depth test, blend, depth write and masked colour write.
// ---------------------------------------------------------------------------

constexpr uint32_t PRIM_NO_PIXEL = 0xFFFFFFFFu;   // fragment key outside the framebuffer or scissor, or discarded
constexpr uint64_t PRIM_MAX_FRAGS = 1ull << 30;   // fragments one primitive may produce

struct PrimParams {
    SceneRaster S;                     // state, texture, framebuffer size
    uint32_t mode, count, units;
    uint32_t tri;                      // 1: units are assembled triangles (glPolygonMode LINE / POINT)
    const VHot *vhot;
    const VCold *vcold;                // null when no per-vertex lighting reads it
    unsigned long long *counts;        // count pass: fragments per unit
    const unsigned long long *base;    // emit pass: first fragment of each unit
    uint32_t *keys, *vals;
    float *dep;
    float4 *col;
    Counters *ctr;
};

__device__ __forceinline__ void prim_load(const PrimParams &P, uint32_t v, CV &o)
{
    const float4 *h = reinterpret_cast<const float4 *>(&P.vhot[v]);
    const float4 a = h[0], b = h[1], c = h[2];
    o.x = a.x; o.y = a.y; o.z = a.z; o.w = a.w;
    o.r = b.x; o.g = b.y; o.b = b.z; o.a = b.w;
    o.s = c.x; o.t = c.y; o.ez = c.z;
    for (int k = 0; k < 3; k++) {
        o.ep[k] = P.vcold ? P.vcold[v].ep[k] : 0.0f;
        o.en[k] = P.vcold ? P.vcold[v].en[k] : 0.0f;
    }
}

// (z + 1) * 0.5 * (far - near) + near in double, stored as float
__device__ __forceinline__ float prim_depth(const RasterConsts &K, float z)
{
    return (float)((double)((z + 1.0f) * 0.5f) * (K.depth_far - K.depth_near) + K.depth_near);
}

// fog of lines and points: coordinate -eye_z (the triangles use +eye_z),
// rgb towards the fragment, alpha of the fog colour
__device__ __forceinline__ Col prim_fog(const RasterConsts &K, Col c, float fog_coord)
{
    float f = 1.0f;
    if (K.fog_mode == GL_LINEAR) {
        if (K.fog_end != K.fog_start) f = (K.fog_end - fog_coord) / (K.fog_end - K.fog_start);
    } else if (K.fog_mode == GL_EXP) {
        f = expf(-K.fog_density * fog_coord);
    } else if (K.fog_mode == GL_EXP2) {
        const float d = K.fog_density * fog_coord;
        f = expf(-d * d);
    }
    if (f < 0.0f) f = 0.0f;
    if (f > 1.0f) f = 1.0f;
    const Col fg = K.fog_color;
    return {fg.r + (c.r - fg.r) * f, fg.g + (c.g - fg.g) * f, fg.b + (c.b - fg.b) * f, fg.a};
}

// pixel key of a fragment: bounds and scissor (raw scissor y against
// top-down rows, as the reference compares them)
__device__ __forceinline__ uint32_t prim_key(const PrimParams &P, const RasterConsts &K, int32_t px, int32_t py)
{
    if (px < 0 || px >= P.S.fb_w || py < 0 || py >= P.S.fb_h) return PRIM_NO_PIXEL;
    if (K.rs->flags & MTGL_F_SCISSOR_TEST) {
        const int32_t *sc = K.rs->scissor;
        if (px < sc[0] || px >= sc[0] + sc[2] || py < sc[1] || py >= sc[1] + sc[3]) return PRIM_NO_PIXEL;
    }
    return (uint32_t)py * (uint32_t)P.S.fb_w + (uint32_t)px;
}

__device__ __forceinline__ void prim_put(const PrimParams &P, uint64_t j, uint32_t key, float depth, Col c)
{
    P.keys[j] = key;
    P.vals[j] = (uint32_t)j;
    P.dep[j] = depth;
    P.col[j] = make_float4(c.r, c.g, c.b, c.a);
}

__device__ __forceinline__ int prim_abs(int32_t v) { return v < 0 ? (int32_t)(0u - (uint32_t)v) : v; }

// reference draw_line_full (src/raster.c:107-246); returns the fragments it
// produces: one per step and width offset.  Where the alpha test rejects a
// pixel the reference's loop continues without stepping and never ends;
// here the step's fragments are discarded and the line goes on.
template <bool EMIT>
__device__ uint64_t prim_line(const PrimParams &P, const RasterConsts &K, int32_t x0, int32_t y0, float z0, int32_t x1,
                              int32_t y1, float z1, Col c0, Col c1, float ez0, float ez1, float u0, float v0, float u1,
                              float v1, uint64_t base)
{
    const int32_t dx = isub(x1, x0), dy = isub(y1, y0);
    const int32_t adx = prim_abs(dx), ady = prim_abs(dy);
    int lw = (int)(K.rs->line_width + 0.5f);
    if (lw < 1) lw = 1;
    if (adx < 0 || ady < 0) return 0;   // a coordinate at INT_MIN: no clipped segment gets there
    const uint64_t steps = (uint64_t)(adx > ady ? adx : ady) + 1u;
    const uint64_t n = steps * (uint64_t)lw;
    if (n > PRIM_MAX_FRAGS) {
        if (!EMIT) atomicOr(&P.ctr->unsupported, 2u);
        return 0;
    }
    if (!EMIT) return n;
    const int32_t sx = dx < 0 ? -1 : 1, sy = dy < 0 ? -1 : 1;
    int32_t err = adx - ady;
    int32_t total = adx > ady ? adx : ady;
    if (total == 0) total = 1;
    const int half = lw / 2;
    const int ex = adx > ady ? 0 : 1, ey = adx > ady ? 1 : 0;
    int32_t cx = x0, cy = y0;
    for (uint64_t step = 0; step < steps; step++) {
        const float t = (float)(int32_t)step / (float)total;
        const float z = z0 + t * (z1 - z0);
        const float depth = prim_depth(K, z);
        Col c = col_lerp(c0, c1, t);
        bool keep = true;
        if (K.textured) {
            const float u = u0 + t * (u1 - u0), v = v0 + t * (v1 - v0);
            int taps = 0;
            const Col tc = col_unpack(sample_texture(K, K.L0, K.L1, u, v, 0.0f, taps));
            if (K.alpha_on && !cmp_f(K.alpha_cmp, tc.a, K.alpha_ref)) keep = false;
            else c = col_mul(c, tc);
        } else if (K.alpha_on && !cmp_f(K.alpha_cmp, c.a, K.alpha_ref)) {
            keep = false;
        }
        if (keep && K.fog_on) c = prim_fog(K, c, -(ez0 + t * (ez1 - ez0)));
        const uint64_t j = base + step * (uint64_t)lw;
        for (int w = 0; w < lw; w++) {
            const int o = lw == 1 ? 0 : w - half;
            const uint32_t key = keep ? prim_key(P, K, cx + o * ex, cy + o * ey) : PRIM_NO_PIXEL;
            prim_put(P, j + (uint64_t)w, key, depth, c);
        }
        const int32_t e2 = err * 2;
        if (e2 > -ady) { err -= ady; cx += sx; }
        if (e2 < adx) { err += adx; cy += sy; }
    }
    return n;
}

__device__ __forceinline__ int prim_outcode(const CV &v)
{
    int code = 0;
    if (v.x < -v.w) code |= 1; else if (v.x > v.w) code |= 2;
    if (v.y < -v.w) code |= 4; else if (v.y > v.w) code |= 8;
    if (v.z < -v.w) code |= 16; else if (v.z > v.w) code |= 32;
    return code;
}

// reference clip_line (src/clipping.h:152-230): each pass moves one endpoint
// onto a plane, so a finite segment needs at most 12; the bound only keeps a
// NaN-free loop finite by construction
__device__ bool prim_clip_line(CV &a, CV &b, bool cold)
{
    int c0 = prim_outcode(a), c1 = prim_outcode(b);
    for (int it = 0; it < 64; it++) {
        if (!(c0 | c1)) return true;
        if (c0 & c1) return false;
        const int out = c0 ? c0 : c1;
        const int plane = (out & 1) ? 2 : (out & 2) ? 3 : (out & 4) ? 4 : (out & 8) ? 5 : (out & 16) ? 0 : 1;
        const float d0 = plane_dist(a, plane), d1 = plane_dist(b, plane);
        const float den = d0 - d1;
        if (fabsf(den) < 1e-10f) return false;
        CV cl;
        cv_lerp_snap(a, b, d0 / den, plane, cold, cl);
        if (out == c0) { a = cl; c0 = prim_outcode(a); }
        else { b = cl; c1 = prim_outcode(b); }
    }
    return false;
}

__device__ __forceinline__ int32_t prim_sx(const mtgl_raster_state *rs, float x)
{
    return cvtt((x + 1.0f) * 0.5f * (float)rs->viewport[2] + (float)rs->viewport[0]);
}
__device__ __forceinline__ int32_t prim_sy(const mtgl_raster_state *rs, float y)
{
    return cvtt((1.0f - y) * 0.5f * (float)rs->viewport[3] + (float)rs->viewport[1]);
}

// reference draw_line_segment (src/raster.c:249-288)
template <bool EMIT>
__device__ uint64_t prim_segment(const PrimParams &P, const RasterConsts &K, uint32_t ia, uint32_t ib, uint64_t base)
{
    CV a, b;
    prim_load(P, ia, a);
    prim_load(P, ib, b);
    if (!prim_clip_line(a, b, P.vcold != nullptr)) return 0;
    float z0 = 0.0f, z1 = 0.0f;
    if (fabsf(a.w) >= 1e-6f) { const float iw = 1.0f / a.w; a.x *= iw; a.y *= iw; z0 = a.z * iw; }
    else { a.x = 0.0f; a.y = 0.0f; }
    if (fabsf(b.w) >= 1e-6f) { const float iw = 1.0f / b.w; b.x *= iw; b.y *= iw; z1 = b.z * iw; }
    else { b.x = 0.0f; b.y = 0.0f; }
    const mtgl_raster_state *rs = K.rs;
    return prim_line<EMIT>(P, K, prim_sx(rs, a.x), prim_sy(rs, a.y), z0, prim_sx(rs, b.x), prim_sy(rs, b.y), z1,
                           Col{a.r, a.g, a.b, a.a}, Col{b.r, b.g, b.b, b.a}, a.ez, b.ez, a.s, a.t, b.s, b.t, base);
}

// reference flush_points, one vertex (src/raster.c:1043-1155)
template <bool EMIT>
__device__ uint64_t prim_point(const PrimParams &P, const RasterConsts &K, uint32_t i, uint64_t base)
{
    CV q;
    prim_load(P, i, q);
    if (q.x < -q.w || q.x > q.w || q.y < -q.w || q.y > q.w || q.z < -q.w || q.z > q.w || q.w <= 0.0f) return 0;
    int ps = (int)(K.rs->point_size + 0.5f);
    if (ps < 1) ps = 1;
    const uint64_t n = (uint64_t)ps * (uint64_t)ps;
    if (!EMIT) return n;
    const float nx = q.x / q.w, ny = q.y / q.w, nz = q.z / q.w;
    const int32_t cx = prim_sx(K.rs, nx), cy = prim_sy(K.rs, ny);
    const float depth = prim_depth(K, nz);
    Col c = {q.r, q.g, q.b, q.a};
    bool keep = true;
    if (K.textured) {
        int taps = 0;
        const Col tc = col_unpack(sample_texture(K, K.L0, K.L1, q.s, q.t, 0.0f, taps));
        if (K.alpha_on && !cmp_f(K.alpha_cmp, tc.a, K.alpha_ref)) keep = false;
        else c = col_mul(c, tc);
    } else if (K.alpha_on && !cmp_f(K.alpha_cmp, c.a, K.alpha_ref)) {
        keep = false;
    }
    if (keep && K.fog_on) c = prim_fog(K, c, -q.ez);
    const int half = ps / 2;
    uint64_t j = base;
    for (int py = 0; py < ps; py++)
        for (int px = 0; px < ps; px++, j++)
            prim_put(P, j, keep ? prim_key(P, K, cx - half + px, cy - half + py) : PRIM_NO_PIXEL, depth, c);
    return n;
}

// Fragment values of a filled face in a draw that also draws faces as lines
// or points: bit 31 marks a triangle fragment (the full stencil / depth /
// alpha-test pipeline of src/raster.c:547-721 applies to it in the resolve),
// bit 30 a fragment the alpha test of its texel discards (after the stencil
// update); the low 30 bits index its depth and colour.
constexpr uint32_t PRIM_TRI_BIT = 0x80000000u;
constexpr uint32_t PRIM_KILL_BIT = 0x40000000u;
constexpr uint32_t PRIM_INDEX_MASK = 0x3FFFFFFFu;

// One filled fan triangle of such a draw, one thread: the reference's
// set-up and row-major pixel loop (src/raster.c:458-545) over the box
// clamped to the framebuffer, each covered pixel's depth and colour as the
// tile pipeline computes them (fragment_source), in the reference's pixel
// order.  Returns the covered pixels.
template <bool EMIT>
__device__ uint64_t prim_fill(const PrimParams &P, const RasterConsts &K, const CV &A, const CV &B, const CV &C,
                              const int32_t *X, const int32_t *Y, int back, uint64_t base)
{
    RTri r;
    int32_t bb[4];
    if (!fan_record(K.rs, P.S.fb_w, P.S.fb_h, P.S.tex_w, P.S.tex_h, A, B, C, X, Y, back, r, bb)) return 0;
    RCold q;
    const bool cold = P.vcold != nullptr;
    if (cold) cold_record(A, B, C, q);
    uint64_t n = 0;
    for (int32_t y = bb[1]; y <= bb[3]; y++) {
        for (int32_t x = bb[0]; x <= bb[2]; x++) {
            float e0, e1, e2;
            edge_values(&r, x, y, e0, e1, e2);
            const bool inside = (r.area > 0.0f && e0 >= 0.0f && e1 >= 0.0f && e2 >= 0.0f) ||
                                (r.area < 0.0f && e0 <= 0.0f && e1 <= 0.0f && e2 <= 0.0f);
            if (!inside) continue;
            if (EMIT) {
                Col c;
                float dep;
                uint32_t meta;
                fragment_source(K, &r, cold ? &q : nullptr, e0 * r.inv_area, e1 * r.inv_area, e2 * r.inv_area, c,
                                dep, meta);
                const uint64_t j = base + n;
                P.keys[j] = (uint32_t)y * (uint32_t)P.S.fb_w + (uint32_t)x;
                P.vals[j] = (uint32_t)j | PRIM_TRI_BIT | ((meta & 1u) ? 0u : PRIM_KILL_BIT);
                P.dep[j] = dep;
                P.col[j] = make_float4(c.r, c.g, c.b, c.a);
            }
            n++;
        }
    }
    return n;
}

// reference render_triangle with a LINE or POINT polygon mode
// (src/raster.c:901-958, draw_triangle_wireframe / draw_triangle_points
// :841-898): clip, divide, fan, cull, then per fan triangle three lines or
// three one-pixel points at the corners (no texture for points; GL_PHONG
// corners lit per vertex with the front material)
template <bool EMIT>
__device__ uint64_t prim_triangle(const PrimParams &P, const RasterConsts &K, uint32_t t, uint64_t base)
{
    const mtgl_raster_state *rs = K.rs;
    const bool cold = P.vcold != nullptr;
    uint32_t ia, ib, ic;
    tri_vertices(P.mode, t, ia, ib, ic);
    CV a[MAX_CLIP], b[MAX_CLIP];
    prim_load(P, ia, b[0]);
    prim_load(P, ib, b[1]);
    prim_load(P, ic, b[2]);
    int n = clip_plane(b, 3, a, 0, cold);
    n = n ? clip_plane(a, n, b, 1, cold) : 0;
    n = n ? clip_plane(b, n, a, 2, cold) : 0;
    n = n ? clip_plane(a, n, b, 3, cold) : 0;
    n = n ? clip_plane(b, n, a, 4, cold) : 0;
    n = n ? clip_plane(a, n, b, 5, cold) : 0;
    if (n < 3) return 0;
    for (int k = 0; k < n; k++) perspective_divide(b[k]);
    const bool phong = K.lighting_on && K.phong;
    const mtgl_material &m = rs->lighting.front;
    uint64_t total = 0;
    for (int j = 1; j + 1 < n; j++) {
        const CV *v[3] = {&b[0], &b[j], &b[j + 1]};
        int32_t X[3], Y[3];
        for (int k = 0; k < 3; k++) {
            X[k] = prim_sx(rs, v[k]->x);
            Y[k] = prim_sy(rs, v[k]->y);
        }
        const float sa = (float)isub(X[1], X[0]) * (float)isub(Y[2], Y[0]) -
                         (float)isub(X[2], X[0]) * (float)isub(Y[1], Y[0]);
        if (rs->flags & MTGL_F_CULL_FACE) {
            const int front = rs->front_face == GL_CCW ? (sa < 0) : (sa > 0);
            const int cull = rs->cull_face_mode == GL_FRONT ? front : (rs->cull_face_mode == GL_BACK ? !front : 1);
            if (cull) continue;
        }
        const int back = rs->front_face == GL_CCW ? (sa >= 0) : (sa < 0);
        const uint32_t pm = back ? rs->polygon_mode_back : rs->polygon_mode_front;
        if (pm != GL_LINE && pm != GL_POINT) {
            total += prim_fill<EMIT>(P, K, *v[0], *v[1], *v[2], X, Y, back, base + total);
            continue;
        }
        Col c[3];
        for (int k = 0; k < 3; k++) {
            c[k] = phong ? compute_lighting(&rs->lighting, rs->flags, v[k]->ep, v[k]->en, col_of(m.ambient),
                                            col_of(m.diffuse), col_of(m.specular), col_of(m.emission), m.shininess)
                         : Col{v[k]->r, v[k]->g, v[k]->b, v[k]->a};
        }
        if (pm == GL_LINE) {
            for (int k = 0; k < 3; k++) {
                const int q = (k + 1) % 3;
                total += prim_line<EMIT>(P, K, X[k], Y[k], v[k]->z, X[q], Y[q], v[q]->z, c[k], c[q], v[k]->ez,
                                         v[q]->ez, v[k]->s, v[k]->t, v[q]->s, v[q]->t, base + total);
            }
        } else {
            for (int k = 0; k < 3; k++) {
                if (EMIT) {
                    Col ck = c[k];
                    bool keep = !(K.alpha_on && !cmp_f(K.alpha_cmp, ck.a, K.alpha_ref));
                    if (keep && K.fog_on) ck = prim_fog(K, ck, -v[k]->ez);
                    prim_put(P, base + total, keep ? prim_key(P, K, X[k], Y[k]) : PRIM_NO_PIXEL,
                             prim_depth(K, v[k]->z), ck);
                }
                total += 1;
            }
        }
    }
    return total;
}

template <bool EMIT>
__global__ void __launch_bounds__(64) k_prim(PrimParams P)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= P.units) return;
    RasterConsts K;
    raster_consts(P.S, K);
    const uint64_t base = EMIT ? P.base[u] : 0;
    uint64_t n;
    if (P.tri) {
        n = prim_triangle<EMIT>(P, K, u, base);
    } else if (P.mode == GL_POINTS) {
        n = prim_point<EMIT>(P, K, u, base);
    } else {
        uint32_t ia, ib;
        if (P.mode == GL_LINES) { ia = 2 * u; ib = ia + 1; }
        else { ia = u; ib = u + 1 < P.count ? u + 1 : 0; }   // strip; the loop's closing segment wraps to vertex 0
        n = prim_segment<EMIT>(P, K, ia, ib, base);
    }
    if (!EMIT) P.counts[u] = n;
}

// One thread per pixel of the sorted fragments (the first of its run): the
// reference's per-pixel steps over the pixel's fragments in order.
__global__ void __launch_bounds__(256) k_prim_resolve(SceneRaster S, const uint32_t *keys, const uint32_t *vals,
                                                      uint32_t n, const float *dep, const float4 *col, uint32_t npix)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = keys[i];
    if (k >= npix || (i > 0 && keys[i - 1] == k)) return;
    RasterConsts K;
    raster_consts(S, K);
    PixelState px = {S.color[k], S.depth[k], K.stencil_on ? (uint32_t)S.stencil[k] : 0u, false, false, false};
    for (uint32_t j = i; j < n && keys[j] == k; j++) {
        const uint32_t v = vals[j], f = v & PRIM_INDEX_MASK;
        const float d = dep[f];
        if (v & PRIM_TRI_BIT) {
            // a filled face's fragment: stencil and depth tests, stencil
            // update, alpha-test discard, depth write, colour
            unsigned long long bytes = 0;
            if (!depth_stencil_step(K, d, 0u, (v & PRIM_KILL_BIT) == 0u, px, bytes)) continue;
            const float4 c = col[f];
            color_step(K, Col{c.x, c.y, c.z, c.w}, px);
            continue;
        }
        if (K.depth_on && !cmp_f(K.depth_cmp, d, px.d)) continue;
        if (K.depth_write) {
            px.d = d;
            px.dirty_d = true;
        }
        const float4 c = col[f];
        color_step(K, Col{c.x, c.y, c.z, c.w}, px);
    }
    if (px.dirty_d) S.depth[k] = px.d;
    if (px.dirty_c) S.color[k] = px.c;
    if (px.dirty_s) S.stencil[k] = (uint8_t)px.s;
}

// ---------------------------------------------------------------------------
// glDrawPixels (reference src/gl_api.c:1231-1373): each source pixel lands
// on its own framebuffer pixel, so one thread per pixel needs no ordering
// ---------------------------------------------------------------------------

struct PixDesc {
    int32_t x, y, w, h, fb_w, fb_h;
    uint32_t format, flags, alpha_func, depth_func, depth_mask, blend_src, blend_dst;
    float alpha_ref;
};

// the reference's drawpixels_alpha_test / drawpixels_depth_test switch
__device__ __forceinline__ bool dp_compare(uint32_t func, float a, float b)
{
    switch (func) {
    case GL_NEVER: return false;
    case GL_LESS: return a < b;
    case GL_EQUAL: return a == b;
    case GL_LEQUAL: return a <= b;
    case GL_GREATER: return a > b;
    case GL_NOTEQUAL: return a != b;
    case GL_GEQUAL: return a >= b;
    default: return true;
    }
}

// the reference's drawpixels_blend_factor (GL_CONSTANT_* and unknown: one)
__device__ __forceinline__ Col dp_factor(uint32_t f, Col s, Col d)
{
    switch (f) {
    case GL_ZERO: return {0.0f, 0.0f, 0.0f, 0.0f};
    case GL_SRC_COLOR: return s;
    case GL_ONE_MINUS_SRC_COLOR: return {1.0f - s.r, 1.0f - s.g, 1.0f - s.b, 1.0f - s.a};
    case GL_DST_COLOR: return d;
    case GL_ONE_MINUS_DST_COLOR: return {1.0f - d.r, 1.0f - d.g, 1.0f - d.b, 1.0f - d.a};
    case GL_SRC_ALPHA: return {s.a, s.a, s.a, s.a};
    case GL_ONE_MINUS_SRC_ALPHA: return {1.0f - s.a, 1.0f - s.a, 1.0f - s.a, 1.0f - s.a};
    case GL_DST_ALPHA: return {d.a, d.a, d.a, d.a};
    case GL_ONE_MINUS_DST_ALPHA: return {1.0f - d.a, 1.0f - d.a, 1.0f - d.a, 1.0f - d.a};
    case GL_SRC_ALPHA_SATURATE: {
        const float f1 = (s.a < (1.0f - d.a)) ? s.a : (1.0f - d.a);
        return {f1, f1, f1, 1.0f};
    }
    default: return {1.0f, 1.0f, 1.0f, 1.0f};
    }
}

__global__ void __launch_bounds__(256) k_draw_pixels(PixDesc P, const uint8_t *src, uint32_t *color, float *depth,
                                                     const Counters *ctr)
{
    if (exec_poisoned(ctr)) return;   // an earlier launch of the stream runs again first
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)P.w * (uint32_t)P.h) return;
    const int32_t row = (int32_t)(i / (uint32_t)P.w), col = (int32_t)(i % (uint32_t)P.w);
    const int32_t fb_y = P.fb_h - 1 - (P.y + row);
    const int32_t dst_x = P.x + col;
    if (fb_y < 0 || fb_y >= P.fb_h || dst_x < 0 || dst_x >= P.fb_w) return;
    uint32_t r, g, b, a = 255u;
    if (P.format == GL_RGBA) {
        const uint32_t v = reinterpret_cast<const uint32_t *>(src)[i];   // src is 256-byte aligned
        r = v & 0xFFu; g = (v >> 8) & 0xFFu; b = (v >> 16) & 0xFFu; a = v >> 24;
    } else if (P.format == GL_RGB) {
        const uint8_t *q = src + 3u * (size_t)i;
        r = q[0]; g = q[1]; b = q[2];
    } else if (P.format == GL_LUMINANCE) {
        r = g = b = src[i];
    } else {
        const uint8_t *q = src + 2u * (size_t)i;
        r = g = b = q[0];
        a = q[1];
    }
    const size_t pix = (size_t)fb_y * (size_t)P.fb_w + (size_t)dst_x;
    const bool depth_on = (P.flags & MTGL_F_DEPTH_TEST) != 0, blend_on = (P.flags & MTGL_F_BLEND) != 0;
    // both framebuffer reads are issued before either is needed
    const float stored = depth_on ? depth[pix] : 0.0f;
    const uint32_t dst = blend_on ? color[pix] : 0u;
    if ((P.flags & MTGL_F_ALPHA_TEST) && !dp_compare(P.alpha_func, (float)a / 255.0f, P.alpha_ref)) return;
    if (depth_on && !dp_compare(P.depth_func, 0.0f, stored)) return;
    Col c = {(float)r / 255.0f, (float)g / 255.0f, (float)b / 255.0f, (float)a / 255.0f};
    if (blend_on) {
        const Col d = col_unpack(dst);
        const Col sf = dp_factor(P.blend_src, c, d), df = dp_factor(P.blend_dst, c, d);
        c = col_clamp(col_add(col_mul(c, sf), col_mul(d, df)));
    }
    if (depth_on && P.depth_mask) depth[pix] = 0.0f;
    color[pix] = col_pack(c);
}

__global__ void __launch_bounds__(256) k_clear(uint32_t *color, float *depth, uint8_t *stencil, int32_t fb_w,
                                               int32_t x0, int32_t y0, int32_t w, int32_t h, uint32_t cv, float dv,
                                               uint8_t sv, const Counters *ctr)
{
    if (exec_poisoned(ctr)) return;   // an earlier launch of the stream runs again first
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)w * (uint64_t)h) return;
    int32_t y = y0 + (int32_t)(i / (uint64_t)w), x = x0 + (int32_t)(i % (uint64_t)w);
    size_t p = (size_t)y * fb_w + x;
    if (color) color[p] = cv;
    if (depth) depth[p] = dv;
    if (stencil) stencil[p] = sv;
}

// A clear of whole framebuffer rows: pixels [p0, p0 + n) of the planes,
// four per thread (16-byte colour and depth stores, one 4-byte stencil store
// where the four are inside the span; single stores at its ends).
__global__ void __launch_bounds__(256) k_clear_span(uint32_t *color, float *depth, uint8_t *stencil, uint64_t p0,
                                                    uint64_t n, uint32_t cv, float dv, uint8_t sv, const Counters *ctr)
{
    if (exec_poisoned(ctr)) return;   // an earlier launch of the stream runs again first
    const uint64_t q = (p0 & ~3ull) + 4ull * ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
    const uint64_t end = p0 + n;
    if (q >= end) return;
    if (q >= p0 && q + 4 <= end) {
        // streaming stores: the cleared lines are not kept in L2 for reading
        typedef unsigned int v4u __attribute__((ext_vector_type(4)));
        const v4u cvv = {cv, cv, cv, cv};
        const v4u dvv = {__float_as_uint(dv), __float_as_uint(dv), __float_as_uint(dv), __float_as_uint(dv)};
        if (color) __builtin_nontemporal_store(cvv, reinterpret_cast<v4u *>(color + q));
        if (depth) __builtin_nontemporal_store(dvv, reinterpret_cast<v4u *>(depth + q));
        if (stencil) *reinterpret_cast<uint32_t *>(stencil + q) = 0x01010101u * sv;
        return;
    }
    for (uint64_t p = q; p < q + 4; p++) {
        if (p < p0 || p >= end) continue;
        if (color) color[p] = cv;
        if (depth) depth[p] = dv;
        if (stencil) stencil[p] = sv;
    }
}

// glReadPixels packing (reference src/gl_api.c:1180-1229), one thread per
// output pixel of the band of window rows [r0, r0 + nrows): window row r is
// framebuffer row fb_h - 1 - (y + r); rows outside the framebuffer are zero
// bytes, columns outside it in rows that exist read as (0, 0, 0, 255).  bpp 4
// stores the packed pixel (its bytes are r, g, b, a in memory), bpp 3 the
// first three bytes.
__global__ void __launch_bounds__(256) k_read_pixels(const uint32_t *color, int32_t fb_w, int32_t fb_h, int32_t x,
                                                     int32_t y, int32_t w, int32_t r0, int32_t nrows, int32_t bpp,
                                                     uint8_t *out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)w * (uint64_t)nrows) return;
    const int64_t r = r0 + (int64_t)(i / (uint64_t)w), col = (int64_t)(i % (uint64_t)w);
    const int64_t fy = (int64_t)fb_h - 1 - ((int64_t)y + r), sx = (int64_t)x + col;
    uint32_t p = 0u;
    if (fy >= 0 && fy < fb_h) p = (sx >= 0 && sx < fb_w) ? color[fy * fb_w + sx] : 0xFF000000u;
    if (bpp == 4) {
        reinterpret_cast<uint32_t *>(out)[i] = p;
    } else {
        out[3 * i + 0] = (uint8_t)(p & 0xFFu);
        out[3 * i + 1] = (uint8_t)((p >> 8) & 0xFFu);
        out[3 * i + 2] = (uint8_t)((p >> 16) & 0xFFu);
    }
}

__global__ void __launch_bounds__(256) k_mip1(const uint32_t *src, int32_t w, uint32_t *dst, int32_t mw, int32_t mh)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mw * mh) return;
    int32_t y = i / mw, x = i % mw;
    size_t s = (size_t)(2 * y) * w + (size_t)(2 * x);
    Col k = col_unpack(src[s]);
    k = col_add(k, col_unpack(src[s + 1]));
    k = col_add(k, col_unpack(src[s + w]));
    k = col_add(k, col_unpack(src[s + w + 1]));
    dst[i] = col_pack(col_scale(k, 0.25f));
}

// A clear of part of one framebuffer, as the batched clear sees it.
struct ClearDesc {
    uint32_t *color;
    float *depth;
    uint8_t *stencil;
    int32_t fb_w, x0, y0, w, h;
    uint32_t cv;
    float dv;
    uint32_t sv;
};

constexpr uint32_t CLEAR_BLOCK_PX = 4096;   // pixels per block of the batched clear

// The clears of many framebuffers in one launch: block b clears pixels
// [CLEAR_BLOCK_PX * (b - block_pre[c]), +CLEAR_BLOCK_PX) of clear c =
// block_clear[b] (row-major over the clear's rectangle).
__global__ void __launch_bounds__(256) k_clear_multi(const ClearDesc *cd, const uint32_t *block_pre,
                                                     const uint32_t *block_clear)
{
    const uint32_t sc = __builtin_amdgcn_readfirstlane(block_clear[blockIdx.x]);
    const ClearDesc &c = cd[sc];
    // framebuffer sides are at most 16384, so a rectangle has < 2^28 pixels
    const uint32_t w = (uint32_t)c.w, n = w * (uint32_t)c.h;
    const uint32_t base = (blockIdx.x - block_pre[sc]) * CLEAR_BLOCK_PX;
    const bool full_rows = c.x0 == 0 && c.w == c.fb_w;
    uint32_t *color = c.color;
    float *depth = c.depth;
    uint8_t *stencil = c.stencil;
    for (uint32_t k = threadIdx.x; k < CLEAR_BLOCK_PX; k += blockDim.x) {
        const uint32_t i = base + k;
        if (i >= n) break;
        size_t p;
        if (full_rows) {
            p = (size_t)c.y0 * (size_t)c.fb_w + i;
        } else {
            const uint32_t y = i / w;
            p = (size_t)(c.y0 + (int32_t)y) * (size_t)c.fb_w + (size_t)(c.x0 + (int32_t)(i - y * w));
        }
        if (color) color[p] = c.cv;
        if (depth) depth[p] = c.dv;
        if (stencil) stencil[p] = (uint8_t)c.sv;
    }
}

// Fill kernel, used where a memset node cannot be (graph capture).
__global__ void __launch_bounds__(256) k_fill_u32(uint32_t *p, uint32_t n, uint32_t v)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}


// ---------------------------------------------------------------------------
// differentiable shading of captured views (include/mtgl_diff.h)
// ---------------------------------------------------------------------------

struct GPx {
    int32_t tri;
    float b[3];
    uint32_t texel[4];
    float fx, fy;
};
static_assert(sizeof(GPx) == sizeof(mtgl_gbuffer_px), "G-buffer pixel layout");

// one captured view as the kernels see it
struct DiffView {
    const GPx *gb;                 // w * h
    const uint32_t *elems;         // 3 per triangle: element of each vertex
    const float *pos;              // 3 per element, object space
    const float *nrm;              // 3 per element
    const mtgl_vertex_state *vs;
    uint32_t ntri;
    int32_t w, h;
    uint32_t pad;
};

// Per pixel of a context's latest draw (a single-scene launch of a
// last-writer kind): the triangle of the fragment that left its colour, that
// fragment's colour weights and its bilinear texture footprint, computed as
// the depth kernel's shading and sample_texture compute them.  Tiles the launch did not
// list (tile_gen != gen) and pixels no fragment reached get tri = -1 and
// their colour; fragments of clipped triangles (overflow slots) get -2.
__global__ void __launch_bounds__(256) k_gbuffer(const SceneRaster *scene, const RTri *rt, const uint32_t *writer,
                                                 const uint32_t *tile_gen, uint32_t gen, uint32_t over_base,
                                                 GPx *out, uint32_t *nclipped)
{
    const SceneRaster &S = scene[0];
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)S.fb_w * (uint32_t)S.fb_h) return;
    const int32_t x = (int32_t)(p % (uint32_t)S.fb_w), y = (int32_t)(p / (uint32_t)S.fb_w);
    const uint32_t tile = (uint32_t)(y / TILE) * (uint32_t)S.tiles_x + (uint32_t)(x / TILE);
    const uint32_t w = tile_gen[tile] == gen ? writer[p] : 0u;
    GPx g = {};
    if (w == 0u) {
        g.tri = -1;
        g.texel[0] = S.color[p];
        out[p] = g;
        return;
    }
    const uint32_t rec = w - 1u;
    if (rec >= over_base) {
        g.tri = -2;
        out[p] = g;
        atomicAdd(nclipped, 1u);
        return;
    }
    RasterConsts K;
    raster_consts(S, K);
    const RTri *r = rt + rec;
    float w0e, w1e, w2e;
    edge_values(r, x, y, w0e, w1e, w2e);
    const float inv = r->inv_area;
    const float b0 = w0e * inv, b1 = w1e * inv, b2 = w2e * inv;
    float u, v;
    texcoord_at(r, K.persp, b0, b1, b2, u, v);
    wrap_uv(K.rep_s, K.rep_t, u, v);
    size_t at[4];
    float fx, fy;
    bilinear_taps(K.L0, K.rep_s, K.rep_t, u, v, at, fx, fy);
    g.tri = (int32_t)rec;
    g.b[0] = b0;
    g.b[1] = b1;
    g.b[2] = b2;
    for (int k = 0; k < 4; k++) g.texel[k] = (uint32_t)at[k];
    g.fx = fx;
    g.fy = fy;
    out[p] = g;
}

// The draw's index stream, positions and normals as dense arrays.
__global__ void __launch_bounds__(256) k_diff_geom(const VertexParams *vp, uint32_t ntri, uint32_t nvert,
                                                   uint32_t *elems, float *pos, float *nrm)
{
    const VertexParams &P = *vp;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 3u * ntri) elems[i] = (uint32_t)fetch_index(P, i);
    if (i < nvert) {
        float p4[4], n3[3];
        fetch(P.position, (int32_t)i, p4, 4);
        if (P.position.comps == 2) p4[2] = 0.0f;
        if (P.normal.enabled) fetch(P.normal, (int32_t)i, n3, 3);
        else { n3[0] = P.cur_normal[0]; n3[1] = P.cur_normal[1]; n3[2] = P.cur_normal[2]; }
        for (int k = 0; k < 3; k++) {
            pos[3 * i + k] = p4[k];
            nrm[3 * i + k] = n3[k];
        }
    }
}

// eye-space position and normal of element e, and the material of its
// lighting with the vertex colour through GL_COLOR_MATERIAL, as k_vertex
// computes them
struct DiffVertex {
    float ep[3], en[3];
    Col vc;                       // sanitised vertex colour
    Col amb, dif, spe, emi;
};

__device__ __forceinline__ void diff_vertex(const DiffView &D, uint32_t e, const float *vcol, DiffVertex &o)
{
    const mtgl_vertex_state *vs = D.vs;
    const float *pp = D.pos + 3u * e, *nn = D.nrm + 3u * e;
    float eye[4];
    xform(vs->modelview, pp[0], pp[1], pp[2], 1.0f, eye);
    float en4[4];
    xform(vs->normal_matrix, nn[0], nn[1], nn[2], 0.0f, en4);
    o.en[0] = en4[0]; o.en[1] = en4[1]; o.en[2] = en4[2];
    normalize3(o.en);
    o.ep[0] = eye[0]; o.ep[1] = eye[1]; o.ep[2] = eye[2];
    const float *c = vcol + 4u * e;
    o.vc = {sanitize(c[0], 0.0f), sanitize(c[1], 0.0f), sanitize(c[2], 0.0f), sanitize(c[3], 1.0f)};
    const mtgl_material &f = vs->lighting.front;
    o.amb = col_of(f.ambient); o.dif = col_of(f.diffuse); o.spe = col_of(f.specular); o.emi = col_of(f.emission);
    const Col k = col_clamp(o.vc);
    const uint32_t mode = vs->color_material_mode;
    if (mode == GL_AMBIENT || mode == GL_AMBIENT_AND_DIFFUSE) o.amb = k;
    if (mode == GL_DIFFUSE || mode == GL_AMBIENT_AND_DIFFUSE) o.dif = k;
}

// lit colour of every (view, element): grid (elements, views).  With aux
// (the backward's), also the unclamped sum and each light's att * N.L, in
// three planes of views * nvert float4 each, for k_diff_lit_bwd.
__global__ void __launch_bounds__(256) k_diff_lit(const DiffView *views, uint32_t nvert, const float *vcol,
                                                  const float *light_rgb, float4 *lit, float4 *aux)
{
    __shared__ float s_ld[3 * MTGL_MAX_LIGHTS];
    const DiffView &D = views[blockIdx.y];
    if (threadIdx.x < MTGL_MAX_LIGHTS) {
        const mtgl_light &lt = D.vs->lighting.lights[threadIdx.x];
        float Ld[3] = {lt.position[0], lt.position[1], lt.position[2]};
        normalize3(Ld);
        for (int k = 0; k < 3; k++) s_ld[3 * threadIdx.x + k] = Ld[k];
    }
    __syncthreads();
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nvert) return;
    DiffVertex q;
    diff_vertex(D, e, vcol, q);
    float ndl[MTGL_MAX_LIGHTS];
    const Col u = lighting_sum<true>(&D.vs->lighting, D.vs->flags, q.ep, q.en, q.amb, q.dif, q.spe, q.emi,
                                     D.vs->lighting.front.shininess, light_rgb, ndl, s_ld);
    const Col r = col_clamp(u);
    const size_t i = (size_t)blockIdx.y * nvert + e;
    lit[i] = make_float4(r.r, r.g, r.b, r.a);
    if (aux) {
        const size_t plane = (size_t)gridDim.y * nvert;
        aux[i] = make_float4(u.r, u.g, u.b, u.a);
        aux[plane + i] = make_float4(ndl[0], ndl[1], ndl[2], ndl[3]);
        aux[2 * plane + i] = make_float4(ndl[4], ndl[5], ndl[6], ndl[7]);
    }
}

__device__ __forceinline__ Col col4(float4 v) { return {v.x, v.y, v.z, v.w}; }

// colour and texture sample of a covered pixel (GL_MODULATE operands)
__device__ __forceinline__ void diff_pixel(const GPx &g, const uint32_t *el, const float4 *litv, uint32_t nvert,
                                           const float4 *tex, Col &c, Col &tc)
{
    Col l[3];
    for (int k = 0; k < 3; k++) {
        const uint32_t e = el[k];
        l[k] = e < nvert ? col4(litv[e]) : Col{0.0f, 0.0f, 0.0f, 0.0f};
    }
    const float b0 = g.b[0], b1 = g.b[1], b2 = g.b[2];
    c.r = l[0].r * b0 + l[1].r * b1 + l[2].r * b2;
    c.g = l[0].g * b0 + l[1].g * b1 + l[2].g * b2;
    c.b = l[0].b * b0 + l[1].b * b1 + l[2].b * b2;
    c.a = l[0].a * b0 + l[1].a * b1 + l[2].a * b2;
    const Col top = col_lerp(col4(tex[g.texel[0]]), col4(tex[g.texel[1]]), g.fx);
    const Col bot = col_lerp(col4(tex[g.texel[2]]), col4(tex[g.texel[3]]), g.fx);
    tc = col_lerp(top, bot, g.fy);
}

// images: grid (pixels, views)
__global__ void __launch_bounds__(256) k_diff_shade(const DiffView *views, uint32_t nvert, const float4 *lit,
                                                    const float4 *tex, float4 *images)
{
    const DiffView &D = views[blockIdx.y];
    const uint32_t npx = (uint32_t)D.w * (uint32_t)D.h;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npx) return;
    const GPx g = D.gb[p];
    Col out;
    if (g.tri < 0) {
        out = col_unpack(g.texel[0]);
    } else {
        Col c, tc;
        diff_pixel(g, D.elems + 3u * (uint32_t)g.tri, lit + (size_t)blockIdx.y * nvert, nvert, tex, c, tc);
        out = col_clamp(col_mul(c, tc));
    }
    images[(size_t)blockIdx.y * npx + p] = make_float4(out.r, out.g, out.b, out.a);
}

__device__ __forceinline__ float clamp_grad(float v, float g) { return (v >= 0.0f && v <= 1.0f) ? g : 0.0f; }

// adjoint of k_diff_shade: dL/dlit per (view, element) into dlit, dL/dtex.
// Channels are independent through the shading, so a thread takes one
// (pixel, channel): the four lanes of a pixel add to the four consecutive
// floats of each target together, one L2 request per target and instruction.
// Grid (4 x pixels, views).
__global__ void __launch_bounds__(256) k_diff_shade_bwd(const DiffView *views, uint32_t nvert, const float *lit,
                                                        const float *tex, const float *grad, float *dlit,
                                                        float *g_tex)
{
    const DiffView &D = views[blockIdx.y];
    const uint32_t npx = (uint32_t)D.w * (uint32_t)D.h;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t p = t >> 2, ch = t & 3u;
    if (p >= npx) return;
    const GPx g = D.gb[p];
    if (g.tri < 0) return;
    const uint32_t *el = D.elems + 3u * (uint32_t)g.tri;
    const float *litv = lit + 4u * (size_t)blockIdx.y * nvert;
    // the operations of diff_pixel, one channel
    uint32_t e[3];
    float l[3];
    for (int k = 0; k < 3; k++) {
        e[k] = el[k];
        l[k] = e[k] < nvert ? litv[4u * e[k] + ch] : 0.0f;
    }
    const float c = l[0] * g.b[0] + l[1] * g.b[1] + l[2] * g.b[2];
    const float fx = g.fx, fy = g.fy;
    const float t0 = tex[4u * (size_t)g.texel[0] + ch], t1 = tex[4u * (size_t)g.texel[1] + ch];
    const float t2 = tex[4u * (size_t)g.texel[2] + ch], t3 = tex[4u * (size_t)g.texel[3] + ch];
    const float top = t0 * (1.0f - fx) + t1 * fx, bot = t2 * (1.0f - fx) + t3 * fx;
    const float tc = top * (1.0f - fy) + bot * fy;
    const float gm = clamp_grad(c * tc, grad[4u * ((size_t)blockIdx.y * npx + p) + ch]);
    const float dc = gm * tc, dt = gm * c;
    for (int k = 0; k < 3; k++)
        if (e[k] < nvert) unsafeAtomicAdd(dlit + 4u * ((size_t)blockIdx.y * nvert + e[k]) + ch, dc * g.b[k]);
    const float wt[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy, fx * fy};
    for (int k = 0; k < 4; k++) unsafeAtomicAdd(g_tex + 4u * (size_t)g.texel[k] + ch, dt * wt[k]);
}

// adjoint of k_diff_lit from the unclamped sums and N.L terms the
// backward's k_diff_lit stored (aux): a block holds DIFF_BWD_E elements x
// DIFF_BWD_G view groups (thread (x, y) takes element x over views y,
// y + DIFF_BWD_G, ...); dL/dvcol is summed over the groups in LDS and
// written, dL/dlight_rgb is reduced per block and added
constexpr int DIFF_BWD_E = 64, DIFF_BWD_G = 8;
__global__ void __launch_bounds__(DIFF_BWD_E * DIFF_BWD_G) k_diff_lit_bwd(const DiffView *views, uint32_t nviews,
                                                                          uint32_t nvert, const float *vcol,
                                                                          const float *light_rgb, const float *dlit,
                                                                          const float4 *aux, float *g_vcol,
                                                                          float *g_light)
{
    __shared__ float s_gl[3 * MTGL_MAX_LIGHTS];
    __shared__ float4 s_gv[DIFF_BWD_G][DIFF_BWD_E];
    const uint32_t tid = threadIdx.y * DIFF_BWD_E + threadIdx.x;
    if (tid < 3 * MTGL_MAX_LIGHTS) s_gl[tid] = 0.0f;
    __syncthreads();
    const uint32_t e = blockIdx.x * DIFF_BWD_E + threadIdx.x;
    const size_t plane = (size_t)nviews * nvert;
    Col gv = {0.0f, 0.0f, 0.0f, 0.0f};
    if (e < nvert) {
        float gl[3 * MTGL_MAX_LIGHTS];
        for (int i = 0; i < 3 * MTGL_MAX_LIGHTS; i++) gl[i] = 0.0f;
        bool any_light = false;
        const float *c = vcol + 4u * e;
        const Col vc = {sanitize(c[0], 0.0f), sanitize(c[1], 0.0f), sanitize(c[2], 0.0f), sanitize(c[3], 1.0f)};
        const Col k = col_clamp(vc);
        for (uint32_t v = threadIdx.y; v < nviews; v += DIFF_BWD_G) {
            const size_t i = (size_t)v * nvert + e;
            const float4 dl = reinterpret_cast<const float4 *>(dlit)[i];
            if (dl.x == 0.0f && dl.y == 0.0f && dl.z == 0.0f && dl.w == 0.0f) continue;
            const float4 u = aux[i], n03 = aux[plane + i], n47 = aux[2 * plane + i];
            const float ndl[MTGL_MAX_LIGHTS] = {n03.x, n03.y, n03.z, n03.w, n47.x, n47.y, n47.z, n47.w};
            const mtgl_vertex_state *vs = views[v].vs;
            const mtgl_lighting &L = vs->lighting;
            const Col g = {clamp_grad(u.x, dl.x), clamp_grad(u.y, dl.y), clamp_grad(u.z, dl.z), clamp_grad(u.w, dl.w)};
            const uint32_t mode = vs->color_material_mode;
            const bool amb_c = mode == GL_AMBIENT || mode == GL_AMBIENT_AND_DIFFUSE;
            const bool dif_c = mode == GL_DIFFUSE || mode == GL_AMBIENT_AND_DIFFUSE;
            const Col dif = dif_c ? k : col_of(L.front.diffuse);
            // dr/dk for the material colour k = clamp(vcol) (att = 1: directional lights)
            Col dk = {0.0f, 0.0f, 0.0f, dif_c ? 1.0f : 0.0f};
            if (amb_c) {
                dk.r += L.model_ambient.r;
                dk.g += L.model_ambient.g;
                dk.b += L.model_ambient.b;
            }
            for (int li = 0; li < MTGL_MAX_LIGHTS; li++) {
                const mtgl_light &lt = L.lights[li];
                if (!lt.enabled) continue;
                if (amb_c) {
                    dk.r += lt.ambient.r;
                    dk.g += lt.ambient.g;
                    dk.b += lt.ambient.b;
                    dk.a += lt.ambient.a;
                }
                const float n = ndl[li];
                if (n > 0.0f) {
                    if (dif_c) {
                        dk.r += light_rgb[3 * li] * n;
                        dk.g += light_rgb[3 * li + 1] * n;
                        dk.b += light_rgb[3 * li + 2] * n;
                        dk.a += lt.diffuse.a * n;
                    }
                    gl[3 * li] += g.r * dif.r * n;
                    gl[3 * li + 1] += g.g * dif.g * n;
                    gl[3 * li + 2] += g.b * dif.b * n;
                    any_light = true;
                }
            }
            const Col gk = col_mul(g, dk);
            gv.r += clamp_grad(vc.r, gk.r);
            gv.g += clamp_grad(vc.g, gk.g);
            gv.b += clamp_grad(vc.b, gk.b);
            gv.a += clamp_grad(vc.a, gk.a);
        }
        if (any_light)
            for (int i = 0; i < 3 * MTGL_MAX_LIGHTS; i++)
                if (gl[i] != 0.0f) atomicAdd(&s_gl[i], gl[i]);
    }
    s_gv[threadIdx.y][threadIdx.x] = make_float4(gv.r, gv.g, gv.b, gv.a);
    __syncthreads();
    if (threadIdx.y == 0 && e < nvert) {
        float4 t = s_gv[0][threadIdx.x];
        for (int g = 1; g < DIFF_BWD_G; g++) {
            const float4 u = s_gv[g][threadIdx.x];
            t.x += u.x;
            t.y += u.y;
            t.z += u.z;
            t.w += u.w;
        }
        reinterpret_cast<float4 *>(g_vcol)[e] = t;
    }
    if (tid < 3 * MTGL_MAX_LIGHTS && s_gl[tid] != 0.0f) unsafeAtomicAdd(g_light + tid, s_gl[tid]);
}

}  // namespace

// ===========================================================================
// host side
// ===========================================================================
//
// A launch runs the pipeline over 1..n scenes (one draw into one framebuffer
// each).  mtgl_dev_draw is a launch of one scene on the context's stream; a
// batch records the clears and draws of many contexts and runs round r (the
// r-th command of every context) as one launch, then replays all rounds from
// one captured hipGraph.

namespace {

struct DevTex {
    int32_t w = 0, h = 0, m1w = 0, m1h = 0;
    uint32_t *px = nullptr, *m1 = nullptr;
    uint32_t *chain = nullptr;     // levels 2.. back to back
    int32_t nlev = 0;              // levels, 0 and 1 included
};

struct DevBuf {
    uint8_t *data = nullptr;
    int64_t size = 0;
};

// grow-only device allocation
struct Arena {
    void *ptr = nullptr;
    size_t cap = 0;
};

// Stream, pipeline buffers and counters shared by the launches of one owner
// (a context for its own draws, a batch for its rounds).
struct Exec {
    hipStream_t stream = nullptr;
    int cu_count = 256;
    // blocks at once: depth<stencil>, depth<no stencil>, general, then the
    // depth kernels at four waves per SIMD (launches of several scenes)
    uint32_t resident[8] = {256, 256, 256, 256, 256, 256, 256, 256};
    Counters *ctr = nullptr;                       // device
    Counters *ctr_host = nullptr;                  // pinned mirror for readback
    uint8_t *pinned = nullptr;                     // staging for uploads that wait for their copy
    size_t pinned_cap = 0;
    // staging of launch tables and staged arrays: blocks taken in order and
    // released together once the stream has drained, so a launch never waits
    // for the previous one's copy, and a launch that must run again finds its
    // inputs where they were
    std::vector<std::pair<uint8_t *, size_t>> ring;
    size_t ring_blk = 0, ring_off = 0, ring_used = 0;
    hipEvent_t upload_done = nullptr;
    Arena vhot, vcold, rt, rcold, box, n_rt, rec_base, tri_tiles, pair_base, scan_tmp, big_list;
    Arena blk_kept, blk_pairs, blk_off;   // per set-up block: kept fan triangles, pairs, first pair
    uint32_t overflow_cap = 0;
    Arena keys[2], vals[2], sort_tmp, tile_list;   // (tile, record) pairs and their sort
    Arena tile_start, tile_end, tile_gen;          // per launch-wide tile: bin range, generation of the range
    uint32_t gen = 0;                              // launch generation stamped into tile_gen
    float pair_ratio = 4.0f;                       // pairs per triangle a launch is sized for (grows)
    uint32_t heavy_bin = HEAVY_BIN;                // bins longer than this go to the front of the tile list
    uint32_t grid_div = 1;                         // raster grids: resident blocks / grid_div
    unsigned long long *stat_slots = nullptr;      // 2 x STAT_SLOTS
    std::vector<unsigned long long> stat_host;
    mtgl_dev_stats stats = {};
    bool timing = false;
    bool ctr_clean = false;                        // per-launch counters at zero (k_expand's block 0 leaves them so)
    uint32_t timing_every = 1;     // time one fragment-stage launch in this many
    uint32_t timing_count = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending_events;
    std::vector<hipEvent_t> free_events;
};

// Device tables of one launch and what its kernels are launched with.  A
// batch keeps one per round so that a captured graph replays the same launch.
struct Launch {
    Arena params;                  // raster/vertex states, staged bytes, scene tables
    uint8_t *host = nullptr;       // their pinned staging (ring): scene 0's raster state, vertex states and
    uint8_t *host_rs = nullptr, *host_vs = nullptr, *host_st = nullptr;   // staged bytes
    uint32_t ns = 0;               // scenes with triangles
    uint32_t tri_total = 0, tiles_total = 0;
    uint32_t vblocks = 0, sblocks = 0;
    bool any_cold = false;
    bool multi_vs = false;         // some scene switches vertex states per vertex
    bool unsupported = false;
    uint32_t kinds = 0;            // bit (1 << kind) for each raster kind present
    bool plain = true;             // no scene for the general kernel has a texture or per-fragment lighting
    bool simple = true;            // no scene for the depth kernel lights per fragment or samples the full chain
    uint64_t pairs_cap = 0;        // pairs of the sizing run (frozen replays use it)
    const VertexParams *vp = nullptr;
    const uint32_t *vblock_pre = nullptr;
    const SetupParams *sp = nullptr;
    const uint32_t *sblock_pre = nullptr;
    const uint32_t *tri_pre = nullptr;
    const SceneRaster *sr = nullptr;
    const uint32_t *tile_pre = nullptr;
    // one scene: its raster state and first vertex state (set-up stages them
    // with its tables instead of after them)
    const mtgl_raster_state *rs1 = nullptr;
    const mtgl_vertex_state *vs1 = nullptr;
    Arena maps;                    // block/tile -> scene maps of a launch with several scenes
    const uint32_t *vmap = nullptr, *smap = nullptr, *tmap = nullptr;
    // the bytes the device copy of `params` holds (blocks up to PARAM_SHADOW_MAX):
    // a launch whose prepared bytes equal them does not upload them again
    std::vector<uint8_t> shadow;
};
constexpr size_t PARAM_SHADOW_MAX = 64u << 10;

}  // namespace

// what a differentiable capture needs to know of a context's latest draw
struct LastDraw {
    bool valid = false;
    uint32_t mode = 0, ntri = 0, n_vstates = 0;
    bool color_array = false;
    mtgl_raster_state rs = {};
    mtgl_vertex_state vs = {};
    uint32_t kind = 0;
    uint32_t gen = 0;              // tile generation of its launch
    uint32_t over_base = 0;        // first overflow record slot of its launch
    int32_t tex_w = 0, tex_h = 0;
    bool captured = false;         // its last-writer plane was recorded
};

// A clear, triangle draw or glDrawPixels queued on a context's stream that
// the host has not yet seen finish.  Until then its inputs stay where its
// launch staged them (the Exec's ring), so that it can run again when an
// earlier launch of the stream overflowed (see confirm()).
struct PendingOp {
    enum Kind : uint32_t { CLEAR, DRAW, PIXELS };
    uint32_t seq = 0;
    Kind kind = CLEAR;
    uint32_t buffers = 0;          // clear
    int32_t x0 = 0, y0 = 0, x1 = 0, y1 = 0;
    uint32_t color = 0;
    float depth = 0.0f;
    uint8_t stencil = 0;
    mtgl_draw draw = {};           // draw: raster state, vertex states and staged bytes in the ring
    uint32_t ntri = 0;
    mtgl_pixels px = {};           // glDrawPixels: source bytes in the ring
};

struct mtgl_dev {
    int device = 0;
    int32_t w = 0, h = 0, tiles_x = 0, tiles_y = 0;
    uint32_t *color = nullptr;
    float *depth = nullptr;
    uint8_t *stencil = nullptr;
    Arena writer;                  // per pixel: last passing fragment (depth kernel)
    Arena pixels;                  // glDrawPixels source bytes
    Arena readback;                // glReadPixels: packed rows of one band
    Arena prim_params;             // states, staged bytes and vertex parameters of a point / line draw
    Arena prim_dep, prim_col;      // fragments of a point / line draw: depth, colour
    std::unordered_map<uint32_t, DevBuf> segs;   // display-list segments by id
    DevTex tex[MTGL_MAX_OBJECTS + 1];
    DevBuf buf[MTGL_MAX_OBJECTS + 1];
    Exec ex;
    Launch one;                    // tables of the context's latest draw
    LastDraw last;
    mtgl_batch *pending = nullptr; // batch launched into this framebuffer, not yet joined
    bool dirty = false;            // work queued on ex.stream since the last batch launch joined it
    std::vector<PendingOp> ops;    // queued clears, draws and glDrawPixels not yet confirmed
    uint32_t seq = 0;              // sequence number of the latest of them (never 0)
    bool replaying = false;
    bool capture = false;          // draws record each pixel's last writer (mtgl_diff_arm)
};

static int confirm(mtgl_dev *d);

namespace {


int map_err(hipError_t e)
{
    if (e == hipSuccess) return MTGL_OK;
    if (e == hipErrorOutOfMemory) return MTGL_ERR_OUT_OF_MEMORY;
    return MTGL_ERR_DEVICE;
}

#define HIPC(expr)                                                                                  \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) {                                                                     \
            fprintf(stderr, "libmtgl: %s failed: %s (%s:%d)\n", #expr, hipGetErrorString(e_), __FILE__, \
                    __LINE__);                                                                      \
            return map_err(e_);                                                                     \
        }                                                                                           \
    } while (0)

// Grow a to at least bytes.  A frozen launch (a replay, or one being
// captured) may not allocate: it runs on buffers its sizing run sized.
int arena_reserve(Exec &x, Arena &a, size_t bytes, bool frozen = false)
{
    if (bytes <= a.cap) return MTGL_OK;
    if (frozen) {
        fprintf(stderr, "libmtgl: a replayed launch needs %zu bytes, its sizing run reserved %zu\n", bytes, a.cap);
        return MTGL_ERR_INVALID;
    }
    size_t cap = a.cap ? a.cap : 4096;
    while (cap < bytes) cap *= 2;
    HIPC(hipStreamSynchronize(x.stream));
    if (a.ptr) HIPC(hipFree(a.ptr));
    a.ptr = nullptr;
    a.cap = 0;
    HIPC(hipMalloc(&a.ptr, cap));
    a.cap = cap;
    return MTGL_OK;
}

void arena_free(Arena &a)
{
    if (a.ptr) (void)hipFree(a.ptr);
    a = Arena();
}

template <typename T>
T *as(Arena &a) { return reinterpret_cast<T *>(a.ptr); }

int pinned_reserve(Exec &x, size_t bytes)
{
    if (x.upload_done) HIPC(hipEventSynchronize(x.upload_done));
    if (bytes <= x.pinned_cap) return MTGL_OK;
    size_t cap = x.pinned_cap ? x.pinned_cap : (64u << 10);
    while (cap < bytes) cap *= 2;
    if (x.pinned) HIPC(hipHostFree(x.pinned));
    x.pinned = nullptr;
    x.pinned_cap = 0;
    HIPC(hipHostMalloc((void **)&x.pinned, cap, hipHostMallocDefault));
    x.pinned_cap = cap;
    return MTGL_OK;
}

unsigned blocks_for(uint64_t n, unsigned threads) { return (unsigned)((n + threads - 1) / threads); }

constexpr size_t RING_BLOCK = 8u << 20;
constexpr size_t RING_LIMIT = 256u << 20;   // staging a context holds before it waits for its queued work
constexpr size_t PIXELS_DIRECT = 1u << 20;   // glDrawPixels images at least this large skip the staging ring

// bytes of pinned staging from the ring (nullptr when the host is out of pinned memory)
uint8_t *ring_take(Exec &x, size_t bytes)
{
    bytes = (bytes + 255) & ~(size_t)255;
    while (x.ring_blk < x.ring.size()) {
        auto &b = x.ring[x.ring_blk];
        if (x.ring_off + bytes <= b.second) {
            uint8_t *p = b.first + x.ring_off;
            x.ring_off += bytes;
            x.ring_used += bytes;
            return p;
        }
        x.ring_blk++;
        x.ring_off = 0;
    }
    const size_t sz = std::max(bytes, RING_BLOCK);
    uint8_t *p = nullptr;
    if (hipHostMalloc((void **)&p, sz, hipHostMallocDefault) != hipSuccess || !p) return nullptr;
    x.ring.emplace_back(p, sz);
    x.ring_blk = x.ring.size() - 1;
    x.ring_off = bytes;
    x.ring_used += bytes;
    return p;
}

// every block free again: only once the stream has drained
void ring_release(Exec &x)
{
    x.ring_blk = 0;
    x.ring_off = 0;
    x.ring_used = 0;
}

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

// exclusive scan of n values of T; out may alias in; *total_out (when
// given) receives the sum
template <typename T>
int device_scan(Exec &x, const T *in, T *out, uint64_t n, bool frozen, T *total_out = nullptr, T *clear = nullptr)
{
    if (n == 0) return MTGL_OK;
    if (n <= SCAN1_MAX<T>) {
        hipLaunchKernelGGL(k_scan_single<T>, dim3(1), dim3(SCAN1_THREADS), 0, x.stream, in, n, out, total_out, clear);
        HIPC(hipGetLastError());
        return MTGL_OK;
    }
    uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    int rc = arena_reserve(x, x.scan_tmp, nb * sizeof(T), frozen);
    if (rc) return rc;
    T *sums = as<T>(x.scan_tmp);
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, x.stream, in, n, sums);
    if (nb <= SCAN2_MAX_BLOCKS) {
        hipLaunchKernelGGL(k_scan_apply_short<T>, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, x.stream, in, n,
                           (const T *)sums, out, total_out, clear);
        HIPC(hipGetLastError());
        return MTGL_OK;
    }
    hipLaunchKernelGGL(k_scan_sums<T>, dim3(1), dim3(SCAN_THREADS), 0, x.stream, sums, nb, total_out);
    hipLaunchKernelGGL(k_scan_apply<T>, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, x.stream, in, n,
                       (const T *)sums, out, clear);
    HIPC(hipGetLastError());
    return MTGL_OK;
}

// resolve an mtgl_array to a device stream
DArr resolve(const mtgl_dev *d, const mtgl_array &a, const uint8_t *staged_dev, uint64_t staged_bytes)
{
    DArr r = {};
    r.enabled = a.enabled;
    r.comps = a.size;
    r.type = a.type;
    r.stride = a.stride;
    r.offset = a.offset;
    if (!a.enabled) return r;
    if (a.buffer == 0) {
        r.src = staged_dev;
        r.size = (int64_t)staged_bytes;
    } else if (a.buffer & MTGL_SEGMENT_BIT) {
        const auto it = d->segs.find(a.buffer & ~MTGL_SEGMENT_BIT);
        if (it != d->segs.end()) {
            r.src = it->second.data;
            r.size = it->second.size;
        }
    } else if (a.buffer <= MTGL_MAX_OBJECTS && d->buf[a.buffer].data) {
        r.src = d->buf[a.buffer].data;
        r.size = d->buf[a.buffer].size;
    }
    return r;
}

hipEvent_t take_event(Exec &x)
{
    if (!x.free_events.empty()) {
        hipEvent_t e = x.free_events.back();
        x.free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Both events of a timed launch, or neither (an event taken alone goes back
// to the pool).
bool take_event_pair(Exec &x, hipEvent_t &a, hipEvent_t &b)
{
    a = take_event(x);
    b = take_event(x);
    if (a && b) return true;
    if (a) x.free_events.push_back(a);
    if (b) x.free_events.push_back(b);
    a = b = nullptr;
    return false;
}

int resolve_timing(Exec &x)
{
    for (auto &pr : x.pending_events) {
        HIPC(hipEventSynchronize(pr.second));
        float ms = 0.0f;
        HIPC(hipEventElapsedTime(&ms, pr.first, pr.second));
        x.stats.raster_ms += ms;
        x.stats.raster_launches++;
        x.free_events.push_back(pr.first);
        x.free_events.push_back(pr.second);
    }
    x.pending_events.clear();
    return MTGL_OK;
}

int exec_init(Exec &x, int device)
{
    if (hipDeviceGetAttribute(&x.cu_count, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        x.cu_count <= 0)
        x.cu_count = 256;
    // resident one-wave blocks of each raster kernel (registers and LDS bound)
    const void *kern[8] = {(const void *)k_raster_depth<true, 3, false>, (const void *)k_raster_depth<false, 3, false>,
                           (const void *)k_raster<false>, (const void *)k_raster_depth<true, 4, false>,
                           (const void *)k_raster_depth<false, 4, false>, (const void *)k_raster<true>,
                           (const void *)k_raster_depth<true, 3, true>, (const void *)k_raster_depth<false, 3, true>};
    for (int k = 0; k < 8; k++) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern[k], 64, 0) != hipSuccess || nb <= 0) nb = 2;
        x.resident[k] = std::min<uint32_t>((uint32_t)nb * (uint32_t)x.cu_count, STAT_SLOTS);
    }
    HIPC(hipStreamCreateWithFlags(&x.stream, hipStreamNonBlocking));
    HIPC(hipMalloc((void **)&x.ctr, sizeof(Counters)));
    HIPC(hipMalloc((void **)&x.stat_slots, 2 * STAT_SLOTS * sizeof(unsigned long long)));
    HIPC(hipHostMalloc((void **)&x.ctr_host, sizeof(Counters), hipHostMallocDefault));
    HIPC(hipEventCreateWithFlags(&x.upload_done, hipEventDisableTiming));
    HIPC(hipMemsetAsync(x.stat_slots, 0, 2 * STAT_SLOTS * sizeof(unsigned long long), x.stream));
    HIPC(hipMemsetAsync(x.ctr, 0, sizeof(Counters), x.stream));
    HIPC(hipStreamSynchronize(x.stream));
    return MTGL_OK;
}

void exec_free(Exec &x)
{
    if (x.stream) (void)hipStreamSynchronize(x.stream);
    Arena *arenas[] = {&x.vhot, &x.vcold, &x.rt, &x.rcold, &x.box, &x.n_rt, &x.rec_base, &x.tri_tiles, &x.pair_base,
                       &x.scan_tmp, &x.big_list, &x.blk_kept, &x.blk_pairs, &x.blk_off, &x.keys[0], &x.keys[1], &x.vals[0], &x.vals[1], &x.sort_tmp,
                       &x.tile_list, &x.tile_start, &x.tile_end, &x.tile_gen};
    for (Arena *a : arenas) arena_free(*a);
    for (auto &pr : x.pending_events) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    for (hipEvent_t e : x.free_events) (void)hipEventDestroy(e);
    x.pending_events.clear();
    x.free_events.clear();
    if (x.ctr) (void)hipFree(x.ctr);
    if (x.stat_slots) (void)hipFree(x.stat_slots);
    if (x.ctr_host) (void)hipHostFree(x.ctr_host);
    if (x.pinned) (void)hipHostFree(x.pinned);
    for (auto &b : x.ring) (void)hipHostFree(b.first);
    x.ring.clear();
    ring_release(x);
    if (x.upload_done) (void)hipEventDestroy(x.upload_done);
    if (x.stream) (void)hipStreamDestroy(x.stream);
    x.ctr = nullptr;
    x.stat_slots = nullptr;
    x.ctr_host = nullptr;
    x.pinned = nullptr;
    x.upload_done = nullptr;
    x.stream = nullptr;
}

int exec_stats(Exec &x, mtgl_dev_stats *out)
{
    x.stat_host.resize(2 * STAT_SLOTS);
    HIPC(hipMemcpyAsync(x.stat_host.data(), x.stat_slots, 2 * STAT_SLOTS * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, x.stream));
    HIPC(hipStreamSynchronize(x.stream));
    int rc = resolve_timing(x);
    if (rc) return rc;
    HIPC(hipMemcpy(x.ctr_host, x.ctr, sizeof(Counters), hipMemcpyDeviceToHost));
    *out = x.stats;   // fragments: those of point and line draws so far, plus the triangles' below
    out->triangles_raster += x.ctr_host->acc_tri_raster;
    out->bin_pairs += x.ctr_host->acc_bin_pairs;
    out->frag_bytes = 0;
    for (uint32_t i = 0; i < STAT_SLOTS; i++) {
        out->fragments += x.stat_host[2 * i];
        out->frag_bytes += x.stat_host[2 * i + 1];
    }
    return MTGL_OK;
}

int exec_stats_reset(Exec &x)
{
    HIPC(hipStreamSynchronize(x.stream));
    int rc = resolve_timing(x);
    if (rc) return rc;
    x.stats = mtgl_dev_stats();
    HIPC(hipMemsetAsync(x.ctr, 0, sizeof(Counters), x.stream));
    HIPC(hipMemsetAsync(x.stat_slots, 0, 2 * STAT_SLOTS * sizeof(unsigned long long), x.stream));
    HIPC(hipStreamSynchronize(x.stream));
    return MTGL_OK;
}

bool draw_mode_ok(uint32_t mode)
{
    switch (mode) {
    case GL_TRIANGLES: case GL_TRIANGLE_STRIP: case GL_TRIANGLE_FAN: case GL_QUADS: case GL_QUAD_STRIP: case GL_POLYGON:
        return true;
    default:
        return false;
    }
}

const DevTex *bound_texture(const mtgl_dev *d, const mtgl_raster_state *rs)
{
    return (rs->texture && rs->texture <= MTGL_MAX_OBJECTS && d->tex[rs->texture].px) ? &d->tex[rs->texture] : nullptr;
}

uint32_t raster_kind(const mtgl_raster_state *rs, const DevTex *tx)
{
    const uint32_t cm = rs->color_mask & 0xFu;
    // without blending, a partial colour mask or an alpha test on a texture,
    // only each pixel's last passing fragment leaves its colour
    const bool lw = !(rs->flags & MTGL_F_BLEND) && (cm == 0xFu || cm == 0u) && !((rs->flags & MTGL_F_ALPHA_TEST) && tx);
    uint32_t k = lw ? KIND_LAST_WRITER : 0u;
    if (rs->flags & MTGL_F_STENCIL_TEST) k |= KIND_STENCIL;
    if (cm) k |= KIND_COLOR;
    return k;
}

bool needs_cold(const mtgl_raster_state *rs)
{
    return (rs->flags & MTGL_F_LIGHTING) && (rs->shade_model == GL_PHONG || rs->lighting.two_side);
}

// one draw of one context, as a launch sees it
struct SceneIn {
    mtgl_dev *dev;
    const mtgl_draw *dr;
};

// Vertex work of a draw.  An indexed draw whose arrays address fewer vertices
// than it has indices transforms each addressed vertex once: slot k is vertex
// base + k for k < n - 1, and slot n - 1 stands for every index no array
// addresses (each array fetch of such an index reads the same defaults).
// Setup then finds a stream position's record through its index.  Other
// draws transform one vertex per stream position.
struct VSlots {
    uint32_t n;
    int32_t base;
    bool unique;
};

VSlots vertex_slots(const mtgl_dev *d, const mtgl_draw *dr)
{
    VSlots r = {dr->count, 0, false};
    if (dr->index_type == 0 || (dr->n_vstates > 1 && dr->state_index.enabled)) return r;
    // the indices each array reads in range (the bounds test of fetch) and their hull
    const uint8_t marker = 0;
    const mtgl_array *arrs[4] = {&dr->position, &dr->color, &dr->texcoord, &dr->normal};
    int64_t lo = INT64_MAX, hi = -1;
    for (int k = 0; k < 4; k++) {
        if (k > 0 && !arrs[k]->enabled) continue;
        const DArr a = resolve(d, *arrs[k], &marker, dr->staged_bytes);
        if (!a.src || a.stride <= 0) continue;
        const int64_t elem = (int64_t)a.comps * (a.type == GL_FLOAT ? 4 : 1);
        const int64_t room = a.size - elem - a.offset;   // at + elem <= size
        if (room < 0) continue;
        const int64_t ilo = a.offset >= 0 ? 0 : (-a.offset + a.stride - 1) / a.stride;
        const int64_t ihi = std::min<int64_t>(room / a.stride, 0x7FFFFFFF);
        if (ilo > ihi) continue;
        lo = std::min(lo, ilo);
        hi = std::max(hi, ihi);
    }
    const int64_t n = hi < lo ? 1 : hi - lo + 2;
    if (n > (int64_t)dr->count) return r;
    r.n = (uint32_t)n;
    r.base = hi < lo ? 0 : (int32_t)lo;
    r.unique = true;
    return r;
}

// Triangle and tile totals of a set of scenes and the buffers they need.
struct Sizes {
    uint64_t tri = 0, tiles = 0;
    uint64_t sblocks = 0;          // set-up blocks (each scene starts a block of its own)
    bool cold = false;
};

Sizes scene_sizes(const SceneIn *in, uint32_t n)
{
    Sizes z;
    for (uint32_t i = 0; i < n; i++) {
        const mtgl_draw *dr = in[i].dr;
        if (!draw_mode_ok(dr->mode)) continue;
        const uint32_t ntri = tri_count(dr->mode, dr->count);
        if (ntri == 0) continue;
        z.tri += ntri;
        z.sblocks += blocks_for(ntri, SETUP_BLOCK);
        z.tiles += (uint64_t)in[i].dev->tiles_x * (uint64_t)in[i].dev->tiles_y;
        z.cold |= needs_cold(dr->raster);
    }
    return z;
}

int reserve_scene_buffers(Exec &x, const Sizes &z)
{
    int rc;
    const size_t t = (size_t)z.tri;
    // record pool: one slot per triangle plus an overflow pool for clipped ones
    if (x.overflow_cap < 4096u) x.overflow_cap = 4096u;
    x.overflow_cap = std::max<uint32_t>(x.overflow_cap, std::min<uint32_t>((uint32_t)(z.tri / 8u), 1u << 24));
    if ((rc = arena_reserve(x, x.rt, (t + x.overflow_cap) * sizeof(RTri)))) return rc;
    if (z.cold && (rc = arena_reserve(x, x.rcold, (t + x.overflow_cap) * sizeof(RCold)))) return rc;
    if ((rc = arena_reserve(x, x.box, (t + x.overflow_cap) * sizeof(RBox)))) return rc;
    if ((rc = arena_reserve(x, x.n_rt, t * sizeof(uint32_t)))) return rc;
    if ((rc = arena_reserve(x, x.rec_base, t * sizeof(uint32_t)))) return rc;
    if ((rc = arena_reserve(x, x.big_list, t * sizeof(uint32_t)))) return rc;
    const size_t sb = (size_t)std::max<uint64_t>(z.sblocks, 1u);
    if ((rc = arena_reserve(x, x.blk_kept, sb * sizeof(uint32_t)))) return rc;
    if ((rc = arena_reserve(x, x.blk_pairs, sb * sizeof(unsigned long long)))) return rc;
    if ((rc = arena_reserve(x, x.blk_off, sb * sizeof(unsigned long long)))) return rc;
    if ((rc = arena_reserve(x, x.tri_tiles, (t + 1) * sizeof(unsigned long long)))) return rc;
    if ((rc = arena_reserve(x, x.pair_base, (t + 1) * sizeof(unsigned long long)))) return rc;
    return MTGL_OK;
}

// Upload the scenes' raster and vertex states, staged client arrays and the
// scene tables into L.params, after sizing the vertex and record buffers.
// Scenes without triangles or with an unsupported primitive are left out.
// The tables hold pointers into x's buffers: a later growth of those buffers
// needs the tables prepared again.
int prepare_launch(Exec &x, Launch &L, const SceneIn *in, uint32_t n)
{
    struct Plan {
        const SceneIn *s;
        uint32_t ntri;
        bool cold;
        const DevTex *tx;
        size_t off_rs, off_vs, off_st;
    };
    std::vector<Plan> plan;
    plan.reserve(n);
    size_t bytes = 0;
    const uint64_t pairs_cap = L.pairs_cap;
    const Arena params = L.params, maps = L.maps;
    std::vector<uint8_t> shadow;
    shadow.swap(L.shadow);
    L = Launch();
    L.params = params;
    L.maps = maps;
    L.pairs_cap = pairs_cap;
    L.shadow.swap(shadow);
    for (uint32_t i = 0; i < n; i++) {
        const mtgl_draw *dr = in[i].dr;
        if (!draw_mode_ok(dr->mode)) {
            L.unsupported = true;
            continue;
        }
        const uint32_t ntri = tri_count(dr->mode, dr->count);
        if (ntri == 0) continue;
        Plan p;
        p.s = &in[i];
        p.ntri = ntri;
        p.cold = needs_cold(dr->raster);
        p.tx = bound_texture(in[i].dev, dr->raster);
        p.off_rs = bytes;
        bytes += up256(sizeof(mtgl_raster_state));
        p.off_vs = bytes;
        bytes += up256((size_t)dr->n_vstates * sizeof(mtgl_vertex_state));
        p.off_st = bytes;
        bytes += up256((size_t)dr->staged_bytes);
        plan.push_back(p);
    }
    const uint32_t ns = (uint32_t)plan.size();
    L.ns = ns;
    if (ns == 0) return MTGL_OK;
    const size_t off_vp = bytes;
    bytes += up256(ns * sizeof(VertexParams));
    const size_t off_sp = bytes;
    bytes += up256(ns * sizeof(SetupParams));
    const size_t off_sr = bytes;
    bytes += up256(ns * sizeof(SceneRaster));
    const size_t off_pre = bytes;
    bytes += up256(4 * (size_t)(ns + 1) * sizeof(uint32_t));

    const Sizes z = scene_sizes(in, n);
    uint64_t sb = 0;
    for (const Plan &p : plan) sb += blocks_for(p.ntri, SETUP_BLOCK);
    if (z.tri > 0x0FFFFFFFull || z.tiles >= (1ull << ITEM_TILE_BITS) / HEAVY_SPLIT) return MTGL_ERR_OUT_OF_MEMORY;
    L.tri_total = (uint32_t)z.tri;
    L.tiles_total = (uint32_t)z.tiles;
    L.vblocks = 0;
    L.sblocks = (uint32_t)sb;
    L.any_cold = z.cold;
    L.plain = true;
    L.simple = true;

    int rc;
    if ((rc = reserve_scene_buffers(x, z))) return rc;
    for (const Plan &p : plan) {
        if ((raster_kind(p.s->dr->raster, p.tx) & KIND_LAST_WRITER) && p.s->dev->capture) {
            mtgl_dev *d = p.s->dev;
            if ((rc = arena_reserve(x, d->writer, (size_t)d->w * (size_t)d->h * 4))) return rc;
        }
    }
    const size_t params_cap = L.params.cap;
    if ((rc = arena_reserve(x, L.params, bytes))) return rc;
    if (L.params.cap != params_cap) L.shadow.clear();   // a new device block holds nothing of ours
    uint8_t *host = ring_take(x, bytes);
    if (!host) return MTGL_ERR_OUT_OF_MEMORY;
    L.host = host;
    L.host_rs = host + plan[0].off_rs;
    L.host_vs = host + plan[0].off_vs;
    L.host_st = host + plan[0].off_st;

    const uint32_t ttot = L.tri_total;
    const uint8_t *dev = (const uint8_t *)L.params.ptr;
    VertexParams *vp = reinterpret_cast<VertexParams *>(host + off_vp);
    SetupParams *sp = reinterpret_cast<SetupParams *>(host + off_sp);
    SceneRaster *sr = reinterpret_cast<SceneRaster *>(host + off_sr);
    uint32_t *pre = reinterpret_cast<uint32_t *>(host + off_pre);
    uint32_t *vbp = pre, *sbp = pre + (ns + 1), *tp = pre + 2 * (ns + 1), *tlp = pre + 3 * (ns + 1);
    uint32_t toff = 0, tileoff = 0, sbo = 0;
    for (uint32_t s = 0; s < ns; s++) {
        const Plan &p = plan[s];
        mtgl_dev *d = p.s->dev;
        const mtgl_draw *dr = p.s->dr;
        memcpy(host + p.off_rs, dr->raster, sizeof(mtgl_raster_state));
        memcpy(host + p.off_vs, dr->vstates, (size_t)dr->n_vstates * sizeof(mtgl_vertex_state));
        if (dr->staged_bytes) memcpy(host + p.off_st, dr->staged, (size_t)dr->staged_bytes);
        const mtgl_raster_state *rs_dev = (const mtgl_raster_state *)(dev + p.off_rs);
        const uint8_t *staged_dev = dev + p.off_st;
        const uint64_t stb = dr->staged_bytes;

        VertexParams v = {};
        v.count = dr->count;
        v.unique = 0;
        v.ubase = 0;
        v.first = dr->first;
        v.index_type = dr->index_type;
        v.n_vstates = dr->n_vstates;
        if (dr->n_vstates > 1 && dr->state_index.enabled) L.multi_vs = true;
        v.indices = resolve(d, dr->indices, staged_dev, stb);
        v.indices.enabled = dr->index_type != 0;
        v.position = resolve(d, dr->position, staged_dev, stb);
        v.color = resolve(d, dr->color, staged_dev, stb);
        v.texcoord = resolve(d, dr->texcoord, staged_dev, stb);
        v.normal = resolve(d, dr->normal, staged_dev, stb);
        v.state_index = resolve(d, dr->state_index, staged_dev, stb);
        memcpy(v.cur_color, dr->cur_color, sizeof(v.cur_color));
        memcpy(v.cur_texcoord, dr->cur_texcoord, sizeof(v.cur_texcoord));
        memcpy(v.cur_normal, dr->cur_normal, sizeof(v.cur_normal));
        v.vstates = (const mtgl_vertex_state *)(dev + p.off_vs);
        v.out_hot = nullptr;
        v.out_cold = nullptr;
        vp[s] = v;

        SetupParams q = {};
        q.mode = dr->mode;
        q.count = dr->count;
        q.vp = (const VertexParams *)(dev + off_vp) + s;
        q.cold = p.cold ? 1u : 0u;
        q.ntri = p.ntri;
        q.tri_off = toff;
        q.scene = s;
        q.tile_off = tileoff;
        q.fb_w = d->w;
        q.fb_h = d->h;
        q.tiles_x = d->tiles_x;
        q.tex_w = p.tx ? p.tx->w : 0;
        q.tex_h = p.tx ? p.tx->h : 0;
        q.rs = rs_dev;
        if (s == 0) {
            L.rs1 = rs_dev;
            L.vs1 = v.vstates;
        }
        q.rt = as<RTri>(x.rt) + toff;
        q.rcold = L.any_cold ? as<RCold>(x.rcold) + toff : nullptr;
        q.box = as<RBox>(x.box) + toff;
        q.rt_over = as<RTri>(x.rt) + ttot;
        q.rcold_over = L.any_cold ? as<RCold>(x.rcold) + ttot : nullptr;
        q.box_over = as<RBox>(x.box) + ttot;
        q.over_base = ttot;
        q.n_rt = as<uint32_t>(x.n_rt) + toff;
        q.rec_base = as<uint32_t>(x.rec_base) + toff;
        q.tri_tiles = as<unsigned long long>(x.tri_tiles) + toff;
        q.overflow_cap = x.overflow_cap;
        q.ctr = x.ctr;
        sp[s] = q;

        SceneRaster r = {};
        r.rs = rs_dev;
        r.color = d->color;
        r.depth = d->depth;
        r.stencil = d->stencil;
        r.kind = raster_kind(dr->raster, p.tx);
        r.writer = (r.kind & KIND_LAST_WRITER) && d->capture ? as<uint32_t>(d->writer) : nullptr;
        r.tex = p.tx ? p.tx->px : nullptr;
        r.tex1 = p.tx ? p.tx->m1 : nullptr;
        r.tex_w = p.tx ? p.tx->w : 0;
        r.tex_h = p.tx ? p.tx->h : 0;
        r.tex1_w = p.tx ? p.tx->m1w : 0;
        r.tex1_h = p.tx ? p.tx->m1h : 0;
        r.lod_mode = dr->raster->lod_mode;
        r.chain = p.tx ? p.tx->chain : nullptr;
        r.nlev = p.tx ? p.tx->nlev : 0;
        r.fb_w = d->w;
        r.fb_h = d->h;
        r.tiles_x = d->tiles_x;
        r.tile_off = tileoff;
        sr[s] = r;
        L.kinds |= 1u << r.kind;
        if (!(r.kind & KIND_LAST_WRITER) && (r.tex != nullptr || needs_cold(dr->raster))) L.plain = false;
        if ((r.kind & KIND_LAST_WRITER) && (needs_cold(dr->raster) || r.lod_mode == MTGL_LOD_FULL_CHAIN)) L.simple = false;

        vbp[s] = 0;
        sbp[s] = sbo;
        tp[s] = toff;
        tlp[s] = tileoff;
        toff += p.ntri;
        tileoff += (uint32_t)(d->tiles_x * d->tiles_y);
        sbo += blocks_for(p.ntri, SETUP_BLOCK);
    }
    vbp[ns] = 0;
    sbp[ns] = sbo;
    tp[ns] = toff;
    tlp[ns] = tileoff;
    // redundant uploads are skipped: the device block already holds these
    // bytes when the previous launch of L uploaded the same ones
    if (!(L.shadow.size() == bytes && memcmp(L.shadow.data(), host, bytes) == 0)) {
        HIPC(hipMemcpyAsync(L.params.ptr, host, bytes, hipMemcpyHostToDevice, x.stream));
        if (bytes <= PARAM_SHADOW_MAX) L.shadow.assign(host, host + bytes);
        else L.shadow.clear();
    }
    L.vp = (const VertexParams *)(dev + off_vp);
    L.sp = (const SetupParams *)(dev + off_sp);
    L.sr = (const SceneRaster *)(dev + off_sr);
    const uint32_t *pre_dev = (const uint32_t *)(dev + off_pre);
    L.vblock_pre = pre_dev;
    L.sblock_pre = pre_dev + (ns + 1);
    L.tri_pre = pre_dev + 2 * (ns + 1);
    L.tile_pre = pre_dev + 3 * (ns + 1);
    if (ns > 1) {
        const size_t nmap = (size_t)L.vblocks + L.sblocks + L.tiles_total;
        if ((rc = arena_reserve(x, L.maps, nmap * 4))) return rc;
        uint32_t *m = as<uint32_t>(L.maps);
        L.vmap = m;
        L.smap = m + L.vblocks;
        L.tmap = m + L.vblocks + L.sblocks;
        hipLaunchKernelGGL(k_scene_maps, dim3(ns), dim3(256), 0, x.stream, L.vblock_pre, L.sblock_pre, L.tile_pre,
                           m, m + L.vblocks, m + L.vblocks + L.sblocks);
        HIPC(hipGetLastError());
    }
    return MTGL_OK;
}

// The pipeline over the scenes of L, queued on x.stream without waiting for
// the GPU.  The pair arrays are sized from x.pair_ratio (a frozen run, a
// replay or one being captured into a graph, allocates nothing and uses the
// capacity of its sizing run); a launch that needs more pairs, or more clip
// slots than the pool holds, poisons the Exec with seq (k_bin_scatter) and
// writes nothing: the host learns of it when it next waits for the stream and
// runs the launch again with larger arrays.
int run_launch(Exec &x, Launch &L, bool frozen, uint32_t seq)
{
    if (L.ns == 0) return L.unsupported ? MTGL_ERR_UNSUPPORTED : MTGL_OK;
    int rc;
    const uint32_t ttot = L.tri_total, ntiles = L.tiles_total;
    if (!frozen) {
        const double want = std::min(2147483647.0, (double)x.pair_ratio * (double)ttot + 4096.0);
        for (int k = 0; k < 2; k++) {
            if ((rc = arena_reserve(x, x.keys[k], (size_t)want * 4))) return rc;
            if ((rc = arena_reserve(x, x.vals[k], (size_t)want * 4))) return rc;
        }
        size_t c = x.keys[0].cap;
        for (int k = 0; k < 2; k++) c = std::min(c, std::min(x.keys[k].cap, x.vals[k].cap));
        // the estimate, not the arenas' power-of-two size, bounds the launch
        // (the sort's grid and histograms scale with it)
        L.pairs_cap = std::min<uint64_t>(std::min<uint64_t>(c / 4, (uint64_t)want), 0x7FFFFFFFull);
    }
    const uint32_t cap = (uint32_t)L.pairs_cap;
    if ((rc = arena_reserve(x, x.tile_start, (size_t)ntiles * 4, frozen))) return rc;
    if ((rc = arena_reserve(x, x.tile_end, (size_t)ntiles * 4, frozen))) return rc;
    const uint32_t list_cap = HEAVY_SPLIT * ntiles;   // up to HEAVY_SPLIT items per tile
    if ((rc = arena_reserve(x, x.tile_list, (size_t)list_cap * sizeof(uint4), frozen))) return rc;
    const uint32_t nblocks = (cap + SORT_TILE - 1) / SORT_TILE;
    const uint32_t nlocal = (cap + LOCAL_TILE - 1) / LOCAL_TILE;
    // digit counts, or the local sort's bucket counts and runs' starts
    if ((rc = arena_reserve(x, x.sort_tmp, (size_t)2 * RADIX * std::max(nblocks, nlocal) * 4, frozen))) return rc;
    if (!frozen) {
        if (x.tile_gen.cap < (size_t)ntiles * 4 || x.gen == 0xFFFFFFFFu) {
            if ((rc = arena_reserve(x, x.tile_gen, (size_t)ntiles * 4))) return rc;
            HIPC(hipMemsetAsync(x.tile_gen.ptr, 0, x.tile_gen.cap, x.stream));
            x.gen = 0;
        }
        x.gen++;
    } else {
        // replays of different rounds share one generation: clear the stamps
        if ((rc = arena_reserve(x, x.tile_gen, (size_t)ntiles * 4, true))) return rc;
        hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(ntiles, 256)), dim3(256), 0, x.stream, as<uint32_t>(x.tile_gen),
                           ntiles, 0u);
    }

    // ---- K2: set-up, with the vertex transform
    unsigned long long *tri_tiles = as<unsigned long long>(x.tri_tiles);
    unsigned long long *blk_pairs = as<unsigned long long>(x.blk_pairs);
    // k_setup and k_expand reset the per-launch counters as they go; a
    // launch that stopped between them leaves them to k_launch_begin
    if (!x.ctr_clean) hipLaunchKernelGGL(k_launch_begin, dim3(1), dim3(64), 0, x.stream, x.ctr);
    x.ctr_clean = false;
    // a wave per corner while the blocks leave SIMDs idle
    const bool split = L.sblocks <= SETUP_SPLIT_CU * x.cu_count;
    auto setup_kernel = [&](auto kern, int nw) {
        hipLaunchKernelGGL(kern, dim3(L.sblocks), dim3(nw * SETUP_BLOCK), 0, x.stream, L.sp, L.vp, L.sblock_pre,
                           L.smap, L.ns, as<uint32_t>(x.blk_kept), blk_pairs, L.ns == 1u ? L.rs1 : nullptr,
                           L.ns == 1u ? L.vs1 : nullptr);
    };
    if (L.multi_vs) {
        if (split) setup_kernel(k_setup<true, 3>, 3);
        else setup_kernel(k_setup<true, 1>, 1);
    } else {
        if (split) setup_kernel(k_setup<false, 3>, 3);
        else setup_kernel(k_setup<false, 1>, 1);
    }
    HIPC(hipGetLastError());

    // ---- K3: pair offsets and total (read on the device: the host never
    // waits), expansion in submission order, stable sort by tile, tile
    // ranges and the work list
    unsigned long long *pair_base = as<unsigned long long>(x.pair_base);
    unsigned long long *blk_off = as<unsigned long long>(x.blk_off);
    const bool fold = L.sblocks <= EXPAND_FOLD;
    if (!fold && (rc = device_scan<unsigned long long>(x, blk_pairs, blk_off, L.sblocks, frozen, &x.ctr->npairs)))
        return rc;
    const unsigned long long *np = &x.ctr->npairs;
    uint32_t *keys = as<uint32_t>(x.keys[0]), *vals = as<uint32_t>(x.vals[0]);
    uint32_t *keys2 = as<uint32_t>(x.keys[1]), *vals2 = as<uint32_t>(x.vals[1]);
    hipLaunchKernelGGL(k_expand, dim3(L.sblocks), dim3(SETUP_BLOCK), 0, x.stream, L.sp, L.sblock_pre, L.smap, L.ns,
                       fold ? nullptr : (const unsigned long long *)blk_off, (const unsigned long long *)blk_pairs,
                       (const unsigned long long *)tri_tiles, pair_base,
                       (const uint32_t *)as<uint32_t>(x.n_rt), (const uint32_t *)as<uint32_t>(x.rec_base),
                       (const RBox *)as<RBox>(x.box), (uint64_t)cap, keys, vals, as<uint32_t>(x.big_list), x.ctr, seq,
                       (const uint32_t *)as<uint32_t>(x.blk_kept), L.sblocks, ttot);
    x.ctr_clean = true;
    // (the queue holds only triangles of over EXPAND_WAVE pairs: few waves)
    hipLaunchKernelGGL(k_expand_big, dim3(std::min<uint32_t>(std::max<uint32_t>(ttot, 1u), 256u)), dim3(64), 0,
                       x.stream, (const unsigned long long *)pair_base, (const uint32_t *)as<uint32_t>(x.n_rt),
                       (const uint32_t *)as<uint32_t>(x.rec_base), (const RBox *)as<RBox>(x.box), (uint64_t)cap,
                       keys, vals, (const uint32_t *)as<uint32_t>(x.big_list), (const Counters *)x.ctr);
    uint32_t *counts = as<uint32_t>(x.sort_tmp);
    // tile keys below RADIX << BINS_MAX_SHIFT: buckets of 1 << shift tiles
    // by the key's bits from `shift` up, either sorted within each block of
    // LOCAL_TILE pairs (up to BINS_MAX_SEGS blocks) or by one global pass,
    // then one workgroup per bucket sorts it by tile and emits the work list
    int shift = 0;
    while (shift <= BINS_MAX_SHIFT && ((ntiles - 1u) >> shift) >= (uint32_t)RADIX) shift++;
    if (ntiles > 0u && shift <= BINS_MAX_SHIFT && nlocal <= BINS_MAX_SEGS) {
        // each block sorted by bucket in place; the buckets gather their runs
        uint32_t *loc = counts + (size_t)RADIX * nlocal;
        hipLaunchKernelGGL(k_sort_local, dim3(nlocal), dim3(SORT_THREADS), 0, x.stream, (const uint32_t *)keys,
                           (const uint32_t *)vals, cap, np, shift, counts, loc, keys2, vals2);
        hipLaunchKernelGGL(k_tile_bins, dim3(((ntiles - 1u) >> shift) + 1u), dim3(BINS_THREADS), 0, x.stream,
                           (const uint32_t *)keys2, (const uint32_t *)vals2, cap, np, shift, nlocal, nullptr,
                           (const uint32_t *)counts, (const uint32_t *)loc, ntiles, vals, as<uint32_t>(x.tile_gen),
                           x.gen, x.heavy_bin, L.tmap, L.tile_pre, L.ns, as<uint4>(x.tile_list), list_cap, x.ctr);
    } else if (ntiles > 0u && shift <= BINS_MAX_SHIFT) {
        hipLaunchKernelGGL(k_sort_hist, dim3(nblocks), dim3(SORT_THREADS), 0, x.stream, (const uint32_t *)keys, cap, np,
                           shift, nblocks, counts);
        if ((rc = device_scan<uint32_t>(x, counts, counts, (uint64_t)RADIX * nblocks, frozen))) return rc;
        hipLaunchKernelGGL(k_sort_scatter, dim3(nblocks), dim3(SORT_THREADS), 0, x.stream, (const uint32_t *)keys,
                           (const uint32_t *)vals, cap, np, shift, nblocks, (const uint32_t *)counts, keys2, vals2);
        hipLaunchKernelGGL(k_tile_bins, dim3(((ntiles - 1u) >> shift) + 1u), dim3(BINS_THREADS), 0, x.stream,
                           (const uint32_t *)keys2, (const uint32_t *)vals2, cap, np, shift, nblocks,
                           (const uint32_t *)counts, nullptr, nullptr, ntiles, vals, as<uint32_t>(x.tile_gen), x.gen,
                           x.heavy_bin, L.tmap, L.tile_pre, L.ns, as<uint4>(x.tile_list), list_cap, x.ctr);
    } else {
        // stable LSD radix sort by tile: pairs keep submission order within a tile
        int key_bits = 1;
        while ((1u << key_bits) <= ntiles) key_bits++;
        for (int sh = 0; sh < key_bits; sh += RADIX_BITS) {
            hipLaunchKernelGGL(k_sort_hist, dim3(nblocks), dim3(SORT_THREADS), 0, x.stream, (const uint32_t *)keys, cap,
                               np, sh, nblocks, counts);
            if ((rc = device_scan<uint32_t>(x, counts, counts, (uint64_t)RADIX * nblocks, frozen))) return rc;
            hipLaunchKernelGGL(k_sort_scatter, dim3(nblocks), dim3(SORT_THREADS), 0, x.stream, (const uint32_t *)keys,
                               (const uint32_t *)vals, cap, np, sh, nblocks, (const uint32_t *)counts, keys2, vals2);
            std::swap(keys, keys2);
            std::swap(vals, vals2);
        }
        hipLaunchKernelGGL(k_segments, dim3(blocks_for(cap, 256)), dim3(256), 0, x.stream, (const uint32_t *)keys, cap,
                           np, ntiles, as<uint32_t>(x.tile_start), as<uint32_t>(x.tile_end), as<uint32_t>(x.tile_gen),
                           x.gen);
        hipLaunchKernelGGL(k_tile_list, dim3(blocks_for(ntiles, 256 * TILE_LIST_PER)), dim3(256), 0, x.stream,
                           (const uint32_t *)as<uint32_t>(x.tile_start), (const uint32_t *)as<uint32_t>(x.tile_end),
                           (const uint32_t *)as<uint32_t>(x.tile_gen), x.gen, ntiles, x.heavy_bin, L.tmap, L.tile_pre,
                           L.ns, as<uint4>(x.tile_list), list_cap, x.ctr);
    }
    HIPC(hipGetLastError());
    // raster grids: no more blocks than items, which the estimated pairs bound
    const uint64_t est_items = std::min<uint64_t>(cap, (uint64_t)((double)x.pair_ratio * (double)ttot) + 64u);

    // ---- K4
    RasterParams rp = {};
    rp.scenes = L.sr;
    rp.tile_pre = L.tile_pre;
    rp.tile_scene = L.tmap;
    rp.nscenes = L.ns;
    rp.rt = as<RTri>(x.rt);
    rp.rcold = L.any_cold ? as<RCold>(x.rcold) : nullptr;
    rp.vals = vals;
    rp.list = as<uint4>(x.tile_list);
    rp.ntiles = ntiles;
    rp.list_cap = list_cap;
    rp.part = 0;
    rp.ctr = x.ctr;
    rp.stats = x.stat_slots;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (x.timing && !frozen && x.timing_count++ % x.timing_every == 0u) (void)take_event_pair(x, ev0, ev1);
    // persistent grids over the work list (longest items first): at most one
    // block per item, at most as many as are resident at once
    const uint32_t nlist = (uint32_t)std::min<uint64_t>(est_items, (uint64_t)HEAVY_SPLIT * ntiles);
    auto grid = [&](int k) {
        // statistics take one slot per block: grids stay within STAT_SLOTS;
        // larger grids are whole runs of the XCD interleave (xcd_block)
        // (three times the resident blocks, six for the general kernel:
        // twice is 20% slower on the 1080p depth stage, four and six times
        // equal to three there; on the 4K overdraw stage six times is 1.5%
        // faster than three)
        const uint32_t over = k == 2 || k == 5 ? 6u : 3u;
        const uint32_t g = std::min<uint32_t>(over * x.resident[k] / x.grid_div, STAT_SLOTS);
        uint32_t n = std::max<uint32_t>(1u, std::min<uint32_t>(nlist, g));
        if (n >= 64u) n &= ~63u;
        return dim3(n);
    };
    const uint32_t LW = KIND_LAST_WRITER, ST = KIND_STENCIL, CO = KIND_COLOR;
    const uint32_t lw_stencil = (1u << (LW | ST)) | (1u << (LW | ST | CO));
    const uint32_t lw_plain = (1u << LW) | (1u << (LW | CO));
    const uint32_t general = (1u << 0) | (1u << ST) | (1u << CO) | (1u << (ST | CO));
    const bool many = L.ns > 1u;   // many small scenes: more waves per SIMD
    const void *fns[3];
    dim3 grids[3];
    int nl = 0;
    // one scene: three waves per SIMD, SIMPLE when it allows
    if (L.kinds & lw_stencil) {
        fns[nl] = many ? (const void *)k_raster_depth<true, 4, false>
                       : L.simple ? (const void *)k_raster_depth<true, 3, true> : (const void *)k_raster_depth<true, 3, false>;
        grids[nl++] = grid(many ? 3 : L.simple ? 6 : 0);
    }
    if (L.kinds & lw_plain) {
        fns[nl] = many ? (const void *)k_raster_depth<false, 4, false>
                       : L.simple ? (const void *)k_raster_depth<false, 3, true> : (const void *)k_raster_depth<false, 3, false>;
        grids[nl++] = grid(many ? 4 : L.simple ? 7 : 1);
    }
    if (L.kinds & general) {
        fns[nl] = L.plain ? (const void *)k_raster<true> : (const void *)k_raster<false>;
        grids[nl++] = grid(L.plain ? 5 : 2);
    }
    void *args[] = {&rp};
    if (ev0 && nl > 0) {
        // timed: the events are taken at the first kernel's start and the
        // last one's end, not around the launches' dispatch
        for (int i = 0; i < nl; i++)
            HIPC(hipExtLaunchKernel(fns[i], grids[i], dim3(64), args, 0, x.stream, i == 0 ? ev0 : nullptr,
                                    i == nl - 1 ? ev1 : nullptr, 0));
        x.pending_events.emplace_back(ev0, ev1);
    } else {
        for (int i = 0; i < nl; i++) HIPC(hipLaunchKernel(fns[i], grids[i], dim3(64), args, 0, x.stream));
        if (ev0) {
            x.free_events.push_back(ev0);
            x.free_events.push_back(ev1);
        }
    }
    return L.unsupported ? MTGL_ERR_UNSUPPORTED : MTGL_OK;
}

// ---- points, lines and glPolygonMode LINE / POINT ----

// 1: a point or line mode; 2: a triangle mode whose faces are both drawn as
// lines or points; 0: filled triangles (the tile pipeline); 3: a triangle
// mode with one face filled and the other drawn as lines or points (the
// point/line path with filled faces rasterised per triangle, so that every
// fragment of the draw is resolved per pixel in submission order)
int prim_class(const mtgl_draw *dr)
{
    switch (dr->mode) {
    case GL_POINTS: case GL_LINES: case GL_LINE_STRIP: case GL_LINE_LOOP: return 1;
    default: break;
    }
    const bool ff = dr->raster->polygon_mode_front == GL_FILL, bf = dr->raster->polygon_mode_back == GL_FILL;
    if (ff && bf) return 0;
    return (!ff && !bf) ? 2 : 3;
}

uint32_t prim_units(const mtgl_draw *dr, int cls)
{
    const uint32_t n = dr->count;
    if (cls >= 2) return tri_count(dr->mode, n);
    switch (dr->mode) {
    case GL_POINTS: return n;
    case GL_LINES: return n / 2;
    case GL_LINE_STRIP: return n >= 2 ? n - 1 : 0;
    default: return n >= 2 ? n : 0;   // GL_LINE_LOOP
    }
}

int run_prims(mtgl_dev *d, const mtgl_draw *dr, int cls)
{
    Exec &x = d->ex;
    const uint32_t units = prim_units(dr, cls);
    if (units == 0 || dr->count == 0) return MTGL_OK;
    const bool cold = needs_cold(dr->raster);
    const DevTex *tx = bound_texture(d, dr->raster);
    // parameters: raster state, vertex states, staged bytes, vertex parameters, a zero block prefix
    const size_t off_rs = 0;
    size_t bytes = up256(sizeof(mtgl_raster_state));
    const size_t off_vs = bytes;
    bytes += up256((size_t)dr->n_vstates * sizeof(mtgl_vertex_state));
    const size_t off_st = bytes;
    bytes += up256((size_t)dr->staged_bytes);
    const size_t off_vp = bytes;
    bytes += up256(sizeof(VertexParams));
    const size_t off_pre = bytes;
    bytes += 256;
    int rc;
    if ((rc = arena_reserve(x, d->prim_params, bytes))) return rc;
    if ((rc = pinned_reserve(x, bytes))) return rc;
    if ((rc = arena_reserve(x, x.vhot, (size_t)dr->count * sizeof(VHot)))) return rc;
    if (cold && (rc = arena_reserve(x, x.vcold, (size_t)dr->count * sizeof(VCold)))) return rc;
    if ((rc = arena_reserve(x, x.tri_tiles, ((size_t)units + 1) * 8))) return rc;
    if ((rc = arena_reserve(x, x.pair_base, ((size_t)units + 1) * 8))) return rc;
    uint8_t *host = x.pinned;
    const uint8_t *dev = (const uint8_t *)d->prim_params.ptr;
    memcpy(host + off_rs, dr->raster, sizeof(mtgl_raster_state));
    memcpy(host + off_vs, dr->vstates, (size_t)dr->n_vstates * sizeof(mtgl_vertex_state));
    if (dr->staged_bytes) memcpy(host + off_st, dr->staged, (size_t)dr->staged_bytes);
    const uint8_t *staged_dev = dev + off_st;
    VertexParams v = {};
    v.count = dr->count;
    v.first = dr->first;
    v.index_type = dr->index_type;
    v.n_vstates = dr->n_vstates;
    v.indices = resolve(d, dr->indices, staged_dev, dr->staged_bytes);
    v.indices.enabled = dr->index_type != 0;
    v.position = resolve(d, dr->position, staged_dev, dr->staged_bytes);
    v.color = resolve(d, dr->color, staged_dev, dr->staged_bytes);
    v.texcoord = resolve(d, dr->texcoord, staged_dev, dr->staged_bytes);
    v.normal = resolve(d, dr->normal, staged_dev, dr->staged_bytes);
    v.state_index = resolve(d, dr->state_index, staged_dev, dr->staged_bytes);
    memcpy(v.cur_color, dr->cur_color, sizeof(v.cur_color));
    memcpy(v.cur_texcoord, dr->cur_texcoord, sizeof(v.cur_texcoord));
    memcpy(v.cur_normal, dr->cur_normal, sizeof(v.cur_normal));
    v.vstates = (const mtgl_vertex_state *)(dev + off_vs);
    v.out_hot = as<VHot>(x.vhot);
    v.out_cold = cold ? as<VCold>(x.vcold) : nullptr;
    memcpy(host + off_vp, &v, sizeof(v));
    memset(host + off_pre, 0, 256);
    HIPC(hipMemcpyAsync(d->prim_params.ptr, host, bytes, hipMemcpyHostToDevice, x.stream));
    HIPC(hipEventRecord(x.upload_done, x.stream));

    unsigned long long *counts = as<unsigned long long>(x.tri_tiles);
    unsigned long long *base = as<unsigned long long>(x.pair_base);
    hipLaunchKernelGGL(k_launch_begin, dim3(1), dim3(64), 0, x.stream, x.ctr);
    hipLaunchKernelGGL(k_vertex, dim3(blocks_for(dr->count, 256)), dim3(256), 0, x.stream,
                       (const VertexParams *)(dev + off_vp), (const uint32_t *)(dev + off_pre), (const uint32_t *)nullptr,
                       1u);

    PrimParams P = {};
    SceneRaster &S = P.S;
    S.rs = (const mtgl_raster_state *)(dev + off_rs);
    S.color = d->color;
    S.depth = d->depth;
    S.stencil = d->stencil;
    S.tex = tx ? tx->px : nullptr;
    S.tex1 = tx ? tx->m1 : nullptr;
    S.tex_w = tx ? tx->w : 0;
    S.tex_h = tx ? tx->h : 0;
    S.tex1_w = tx ? tx->m1w : 0;
    S.tex1_h = tx ? tx->m1h : 0;
    S.lod_mode = dr->raster->lod_mode;
    S.chain = tx ? tx->chain : nullptr;
    S.nlev = tx ? tx->nlev : 0;
    S.fb_w = d->w;
    S.fb_h = d->h;
    S.tiles_x = d->tiles_x;
    P.mode = dr->mode;
    P.count = dr->count;
    P.units = units;
    P.tri = cls >= 2 ? 1u : 0u;
    P.vhot = as<VHot>(x.vhot);
    P.vcold = cold ? as<VCold>(x.vcold) : nullptr;
    P.counts = counts;
    P.base = base;
    P.ctr = x.ctr;
    hipLaunchKernelGGL(k_prim<false>, dim3(blocks_for(units, 64)), dim3(64), 0, x.stream, P);
    HIPC(hipGetLastError());
    if ((rc = device_scan<unsigned long long>(x, counts, base, (uint64_t)units, false, &x.ctr->npairs))) return rc;
    HIPC(hipMemcpyAsync(x.ctr_host, x.ctr, sizeof(Counters), hipMemcpyDeviceToHost, x.stream));
    HIPC(hipStreamSynchronize(x.stream));
    if (x.ctr_host->unsupported & 2u) return MTGL_ERR_OUT_OF_MEMORY;   // one primitive over PRIM_MAX_FRAGS
    const uint64_t total = x.ctr_host->npairs;
    x.stats.fragments += total;   // pixels the reference's point and line loops visit
    if (total == 0) return MTGL_OK;
    // fragment indices travel through the sort in the value's low bits
    if (total > (cls == 3 ? (uint64_t)PRIM_INDEX_MASK + 1u : 0x7FFFFFFFull)) return MTGL_ERR_OUT_OF_MEMORY;
    const uint32_t n = (uint32_t)total;
    for (int k = 0; k < 2; k++) {
        if ((rc = arena_reserve(x, x.keys[k], (size_t)n * 4))) return rc;
        if ((rc = arena_reserve(x, x.vals[k], (size_t)n * 4))) return rc;
    }
    if ((rc = arena_reserve(x, d->prim_dep, (size_t)n * 4))) return rc;
    if ((rc = arena_reserve(x, d->prim_col, (size_t)n * 16))) return rc;
    uint32_t *keys = as<uint32_t>(x.keys[0]), *vals = as<uint32_t>(x.vals[0]);
    uint32_t *keys2 = as<uint32_t>(x.keys[1]), *vals2 = as<uint32_t>(x.vals[1]);
    P.keys = keys;
    P.vals = vals;
    P.dep = as<float>(d->prim_dep);
    P.col = as<float4>(d->prim_col);
    hipLaunchKernelGGL(k_prim<true>, dim3(blocks_for(units, 64)), dim3(64), 0, x.stream, P);
    // stable LSD radix sort by pixel; the all-ones key of discarded fragments stays last
    const uint32_t npix = (uint32_t)d->w * (uint32_t)d->h;
    int key_bits = 1;
    while (key_bits < 32 && (1u << key_bits) <= npix) key_bits++;
    const uint32_t nblocks = (n + SORT_TILE - 1) / SORT_TILE;
    if ((rc = arena_reserve(x, x.sort_tmp, (size_t)RADIX * nblocks * 4))) return rc;
    uint32_t *sc = as<uint32_t>(x.sort_tmp);
    for (int shift = 0; shift < key_bits; shift += RADIX_BITS) {
        hipLaunchKernelGGL(k_sort_hist, dim3(nblocks), dim3(SORT_THREADS), 0, x.stream, (const uint32_t *)keys, n,
                           (const unsigned long long *)nullptr, shift, nblocks, sc);
        if ((rc = device_scan<uint32_t>(x, sc, sc, (uint64_t)RADIX * nblocks, false))) return rc;
        hipLaunchKernelGGL(k_sort_scatter, dim3(nblocks), dim3(SORT_THREADS), 0, x.stream, (const uint32_t *)keys,
                           (const uint32_t *)vals, n, (const unsigned long long *)nullptr, shift, nblocks,
                           (const uint32_t *)sc, keys2, vals2);
        std::swap(keys, keys2);
        std::swap(vals, vals2);
    }
    hipLaunchKernelGGL(k_prim_resolve, dim3(blocks_for(n, 256)), dim3(256), 0, x.stream, S, (const uint32_t *)keys,
                       (const uint32_t *)vals, n, (const float *)as<float>(d->prim_dep),
                       (const float4 *)as<float4>(d->prim_col), npix);
    HIPC(hipGetLastError());
    return MTGL_OK;
}

// prepare + sizing run, growing the clip overflow pool once when it was too small
int execute_scenes(Exec &x, Launch &L, const SceneIn *in, uint32_t n, uint32_t seq)
{
    int rc = prepare_launch(x, L, in, n);
    if (rc) return rc;
    return run_launch(x, L, false, seq);
}

// After a launch poisoned x (it needed more pairs or clip slots than x
// held): larger arrays for its next run.  ntri: the launch's triangles.
void exec_grow(Exec &x, const Counters &c, uint32_t ntri)
{
    if (c.poison_flags & 1u) x.overflow_cap = std::min<uint32_t>(std::max<uint32_t>(x.overflow_cap * 2u, 4096u), 1u << 26);
    if (c.poison_flags & 2u) {
        const double r = 1.25 * (double)c.poison_pairs / (double)std::max<uint32_t>(ntri, 1u) + 1.0;
        x.pair_ratio = (float)std::max<double>(x.pair_ratio, std::min(r, 2.0e9));
    }
}

// Reads x's counters back (the stream must have drained) and clears the
// poison it carries, returning it (0: none).
int exec_take_poison(Exec &x, uint32_t &poison)
{
    HIPC(hipMemcpy(x.ctr_host, x.ctr, sizeof(Counters), hipMemcpyDeviceToHost));
    poison = x.ctr_host->poison;
    if (poison) {
        HIPC(hipMemsetAsync((uint8_t *)x.ctr + offsetof(Counters, poison), 0, sizeof(uint32_t), x.stream));
        HIPC(hipStreamSynchronize(x.stream));
    }
    return MTGL_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------

namespace {

struct BatchCmd {
    mtgl_dev *dev = nullptr;
    bool is_clear = false;
    ClearDesc clear = {};
    mtgl_draw draw = {};                   // its pointers point at the copies below
    mtgl_raster_state raster = {};
    std::vector<mtgl_vertex_state> vstates;
    std::vector<uint8_t> staged;
};

struct BatchRound {
    std::vector<SceneIn> scenes;           // draws of the round, at most one per context
    std::vector<ClearDesc> clears;
    Launch launch;
    Arena clear_params;
    uint32_t clear_blocks = 0;
    const ClearDesc *clear_dev = nullptr;
    const uint32_t *clear_pre = nullptr;
    const uint32_t *clear_map = nullptr;   // block -> clear
};

}  // namespace

struct mtgl_batch {
    int device = 0;
    Exec ex;
    std::vector<BatchCmd *> cmds;
    std::vector<mtgl_dev *> devs;          // contexts in first-use order
    std::vector<BatchRound> rounds;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipEvent_t done = nullptr;             // end of the latest launch
    hipEvent_t join = nullptr;
    bool recorded = false;                 // mtgl_batch_end ran
    bool ready = false;                    // a graph is instantiated and every context is alive
    mtgl_batch_info_t info = {};
};

namespace {

thread_local mtgl_batch *g_recording = nullptr;
std::mutex g_batches_mu;
std::vector<mtgl_batch *> g_batches;   // live batches that hold contexts

// A context's buffers or objects are about to be reallocated or freed: the
// graphs of the batches that use it hold stale pointers and stop launching.
void invalidate_batches(mtgl_dev *d)
{
    std::lock_guard<std::mutex> lk(g_batches_mu);
    for (mtgl_batch *b : g_batches) {
        for (mtgl_dev *e : b->devs) {
            if (e == d) {
                if (b->ready) fprintf(stderr, "libmtgl: a context of a batch changed its objects; the batch must be rebuilt\n");
                b->ready = false;
                break;
            }
        }
    }
}

// Order the context's next work after the batch launch that rendered into it.
int dev_join(mtgl_dev *d)
{
    if (d->pending) {
        HIPC(hipStreamWaitEvent(d->ex.stream, d->pending->done, 0));
        d->pending = nullptr;
    }
    return MTGL_OK;
}

void batch_note_dev(mtgl_batch *b, mtgl_dev *d)
{
    for (mtgl_dev *e : b->devs)
        if (e == d) return;
    b->devs.push_back(d);
}

// Round r holds the r-th recorded command of every context, so each
// context's commands keep their order and a round touches each framebuffer
// at most once.
void batch_plan(mtgl_batch *b)
{
    b->rounds.clear();
    std::vector<std::vector<BatchCmd *>> per_dev(b->devs.size());
    for (BatchCmd *c : b->cmds) {
        size_t i = 0;
        while (b->devs[i] != c->dev) i++;
        per_dev[i].push_back(c);
    }
    size_t nr = 0;
    for (auto &v : per_dev) nr = std::max(nr, v.size());
    b->rounds.resize(nr);
    for (size_t r = 0; r < nr; r++) {
        for (auto &v : per_dev) {
            if (r >= v.size()) continue;
            BatchCmd *c = v[r];
            if (c->is_clear) b->rounds[r].clears.push_back(c->clear);
            else b->rounds[r].scenes.push_back({c->dev, &c->draw});
        }
    }
}

int batch_upload_clears(Exec &x, BatchRound &R)
{
    const uint32_t n = (uint32_t)R.clears.size();
    if (n == 0) return MTGL_OK;
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n; i++) acc += (uint64_t)R.clears[i].w * R.clears[i].h / CLEAR_BLOCK_PX + 1;
    if (acc > 0x7FFFFFFFull) return MTGL_ERR_OUT_OF_MEMORY;
    const size_t b_cd = up256(n * sizeof(ClearDesc)), b_pre = up256((n + 1) * sizeof(uint32_t));
    const size_t bytes = b_cd + b_pre + up256(acc * sizeof(uint32_t));
    int rc;
    if ((rc = arena_reserve(x, R.clear_params, bytes))) return rc;
    if ((rc = pinned_reserve(x, bytes))) return rc;
    memcpy(x.pinned, R.clears.data(), n * sizeof(ClearDesc));
    uint32_t *pre = reinterpret_cast<uint32_t *>(x.pinned + b_cd);
    uint32_t *map = reinterpret_cast<uint32_t *>(x.pinned + b_cd + b_pre);
    uint32_t nb = 0;
    for (uint32_t i = 0; i < n; i++) {
        pre[i] = nb;
        const uint32_t k = (uint32_t)blocks_for((uint64_t)R.clears[i].w * (uint64_t)R.clears[i].h, CLEAR_BLOCK_PX);
        for (uint32_t j = 0; j < k; j++) map[nb + j] = i;
        nb += k;
    }
    pre[n] = nb;
    R.clear_blocks = nb;
    HIPC(hipMemcpyAsync(R.clear_params.ptr, x.pinned, bytes, hipMemcpyHostToDevice, x.stream));
    HIPC(hipEventRecord(x.upload_done, x.stream));
    R.clear_dev = (const ClearDesc *)R.clear_params.ptr;
    R.clear_pre = (const uint32_t *)((const uint8_t *)R.clear_params.ptr + b_cd);
    R.clear_map = (const uint32_t *)((const uint8_t *)R.clear_params.ptr + b_cd + b_pre);
    return MTGL_OK;
}

// Every round in order.  The sizing run (frozen false) uploads the tables
// and sizes the buffers; frozen runs replay exactly those launches.
int batch_run(mtgl_batch *b, bool frozen)
{
    Exec &x = b->ex;
    for (BatchRound &R : b->rounds) {
        int rc;
        if (!R.clears.empty()) {
            if (!frozen && (rc = batch_upload_clears(x, R))) return rc;
            if (R.clear_blocks)
                hipLaunchKernelGGL(k_clear_multi, dim3(R.clear_blocks), dim3(256), 0, x.stream, R.clear_dev,
                                   R.clear_pre, R.clear_map);
            HIPC(hipGetLastError());
        }
        if (R.scenes.empty()) continue;
        // sequence number of a round's launch: its index + 1
        const uint32_t seq = (uint32_t)(&R - b->rounds.data()) + 1u;
        rc = frozen ? run_launch(x, R.launch, true, seq)
                    : execute_scenes(x, R.launch, R.scenes.data(), (uint32_t)R.scenes.size(), seq);
        if (rc && rc != MTGL_ERR_UNSUPPORTED) return rc;
    }
    return MTGL_OK;
}

// Save (restore false) or restore the colour, depth and stencil planes of
// the batch's contexts to or from mem, on the batch stream.
void planes_copy(mtgl_batch *b, uint8_t *mem, bool restore)
{
    for (mtgl_dev *d : b->devs) {
        const size_t n = (size_t)d->w * d->h;
        uint8_t *c = mem, *z = mem + 4 * n, *st = mem + 8 * n;
        const hipMemcpyKind k = hipMemcpyDeviceToDevice;
        if (!restore) {
            (void)hipMemcpyAsync(c, d->color, 4 * n, k, b->ex.stream);
            (void)hipMemcpyAsync(z, d->depth, 4 * n, k, b->ex.stream);
            (void)hipMemcpyAsync(st, d->stencil, n, k, b->ex.stream);
        } else {
            (void)hipMemcpyAsync(d->color, c, 4 * n, k, b->ex.stream);
            (void)hipMemcpyAsync(d->depth, z, 4 * n, k, b->ex.stream);
            (void)hipMemcpyAsync(d->stencil, st, n, k, b->ex.stream);
        }
        mem += up256(9 * n);
    }
}

// The largest round fixes the sizes of the shared buffers up front, so that
// the sizing run seldom has to grow them between rounds.
int batch_reserve(mtgl_batch *b)
{
    Sizes m;
    for (BatchRound &R : b->rounds) {
        const Sizes z = scene_sizes(R.scenes.data(), (uint32_t)R.scenes.size());
        m.tri = std::max(m.tri, z.tri);
        m.tiles = std::max(m.tiles, z.tiles);
        m.cold |= z.cold;
    }
    if (m.tri > 0x0FFFFFFFull) return MTGL_ERR_OUT_OF_MEMORY;
    return m.tri ? reserve_scene_buffers(b->ex, m) : MTGL_OK;
}

int batch_build_inner(mtgl_batch *b, int autotune)
{
    Exec &x = b->ex;
    int rc;
    if ((rc = batch_reserve(b))) return rc;
    // sizing runs: every round in order, until none needs larger arrays than
    // the earlier rounds grew them to
    for (int attempt = 0;; attempt++) {
        if ((rc = batch_run(b, false))) return rc;
        HIPC(hipStreamSynchronize(x.stream));
        uint32_t poison = 0;
        if ((rc = exec_take_poison(x, poison))) return rc;
        if (!poison) break;
        if (attempt >= 16) return MTGL_ERR_OUT_OF_MEMORY;
        exec_grow(x, *x.ctr_host, b->rounds[poison - 1].launch.tri_total);
        ring_release(x);
    }
    // a later round may have grown a buffer that an earlier round's tables
    // point into: prepare every round's tables again, at the final sizes
    size_t c = x.keys[0].cap;
    for (int k = 0; k < 2; k++) c = std::min(c, std::min(x.keys[k].cap, x.vals[k].cap));
    const uint64_t cap = std::min<uint64_t>(c / 4, 0x7FFFFFFFull);
    for (BatchRound &R : b->rounds) {
        if (R.scenes.empty()) continue;
        if ((rc = prepare_launch(x, R.launch, R.scenes.data(), (uint32_t)R.scenes.size()))) return rc;
        R.launch.pairs_cap = cap;
    }
    HIPC(hipStreamSynchronize(x.stream));
    ring_release(x);
    b->info.heavy_bin = x.heavy_bin;
    b->info.grid_div = x.grid_div;
    if (autotune) {
        // the heavy-bin threshold of the tile list and the raster grid size:
        // one warm and three timed frozen replays of the batch per candidate
        static const uint32_t cand_bin[] = {96u, 192u, 320u, 0xFFFFFFFFu, 192u, 192u};
        static const uint32_t cand_div[] = {1u, 1u, 1u, 1u, 2u, 3u};
        const int ncand = (int)(sizeof(cand_bin) / sizeof(cand_bin[0]));
        static_assert(sizeof(cand_bin) / sizeof(cand_bin[0]) <= MTGL_BATCH_TUNE_MAX, "tune table");
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPC(hipEventCreate(&e0));
        HIPC(hipEventCreate(&e1));
        float best = -1.0f;
        for (int c = 0; c < ncand && !rc; c++) {
            x.heavy_bin = cand_bin[c];
            x.grid_div = cand_div[c];
            if ((rc = batch_run(b, true))) break;
            if ((rc = map_err(hipEventRecord(e0, x.stream)))) break;
            for (int rep = 0; rep < 3 && !rc; rep++) rc = batch_run(b, true);
            if (rc || (rc = map_err(hipEventRecord(e1, x.stream))) || (rc = map_err(hipEventSynchronize(e1)))) break;
            float ms = 0.0f;
            if ((rc = map_err(hipEventElapsedTime(&ms, e0, e1)))) break;
            b->info.tune_ms[c] = ms / 3.0;
            if (best < 0.0f || ms < best) {
                best = ms;
                b->info.heavy_bin = cand_bin[c];
                b->info.grid_div = cand_div[c];
            }
        }
        b->info.tune_candidates = (uint32_t)ncand;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) return rc;
        x.heavy_bin = b->info.heavy_bin;
        x.grid_div = b->info.grid_div;
    }
    if (b->rounds.empty()) return MTGL_OK;   // nothing recorded: launches do nothing
    // capture the replay
    HIPC(hipStreamBeginCapture(x.stream, hipStreamCaptureModeThreadLocal));
    rc = batch_run(b, true);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(x.stream, &g);
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : map_err(e);
    }
    b->graph = g;
    HIPC(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
    return MTGL_OK;
}

// The sizing run and the tuning replays render into the contexts'
// framebuffers: their planes are saved first and restored after, so that
// building a batch leaves every framebuffer as it was.
int batch_build(mtgl_batch *b, int autotune)
{
    batch_plan(b);
    for (mtgl_dev *d : b->devs) {
        int rc = dev_join(d);
        if (rc) return rc;
        if ((rc = confirm(d))) return rc;
        HIPC(hipStreamSynchronize(d->ex.stream));
        d->dirty = false;
    }
    size_t total = 0;
    for (mtgl_dev *d : b->devs) total += up256((size_t)d->w * d->h * 9);
    void *snap = nullptr;
    HIPC(hipMalloc(&snap, std::max<size_t>(total, 256)));
    planes_copy(b, (uint8_t *)snap, false);
    const int rc = batch_build_inner(b, autotune);
    planes_copy(b, (uint8_t *)snap, true);
    const hipError_t e = hipStreamSynchronize(b->ex.stream);
    (void)hipFree(snap);
    if (rc) return rc;
    if (e != hipSuccess) return map_err(e);
    // counters count launches only from here on
    return exec_stats_reset(b->ex);
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================

MTGL_API const char *mtgl_dev_backend(void)
{
    return "hip:gfx950 (K1 vertex, K2 setup, K3 binner, K4 tile raster)";
}

// ---- the log of queued clears, draws and glDrawPixels (see PendingOp) ----

static uint32_t next_seq(mtgl_dev *d)
{
    if (++d->seq == 0u) d->seq = 1u;
    return d->seq;
}

static int launch_clear(mtgl_dev *d, const PendingOp &op)
{
    const int32_t w = op.x1 - op.x0, h = op.y1 - op.y0;
    const uint64_t n = (uint64_t)w * (uint64_t)h;
    uint32_t *c = (op.buffers & MTGL_BUF_COLOR) ? d->color : nullptr;
    float *z = (op.buffers & MTGL_BUF_DEPTH) ? d->depth : nullptr;
    uint8_t *st = (op.buffers & MTGL_BUF_STENCIL) ? d->stencil : nullptr;
    if (op.x0 == 0 && w == d->w) {
        // whole rows: one contiguous span of each plane
        const uint64_t p0 = (uint64_t)op.y0 * (uint64_t)d->w;
        const uint64_t quads = (p0 + n - (p0 & ~3ull) + 3) / 4;
        hipLaunchKernelGGL(k_clear_span, dim3(blocks_for(quads, 256)), dim3(256), 0, d->ex.stream, c, z, st, p0, n,
                           op.color, op.depth, op.stencil, (const Counters *)d->ex.ctr);
    } else {
        hipLaunchKernelGGL(k_clear, dim3(blocks_for(n, 256)), dim3(256), 0, d->ex.stream, c, z, st, d->w, op.x0, op.y0,
                           w, h, op.color, op.depth, op.stencil, (const Counters *)d->ex.ctr);
    }
    HIPC(hipGetLastError());
    d->dirty = true;
    return MTGL_OK;
}

// ---- glDrawPixels without device allocation (mtgl_dev_draw_pixels_static) ----
//
// The source image is staged in `scratch`, a fixed device array that each
// GPU the library runs on holds until the process ends.  Host code never
// passes the array's name to a kernel (on the host it is only a stand-in for
// the symbol): the device address of each GPU's copy is looked up once per
// device, with that device current, when the first context on it is created
// (scratch_setup), and kept in g_scratch.  One mutex per device serialises
// the calls, each of which waits for its kernel before returning, so two
// calls never use a GPU's copy at once.
static __device__ __align__(256) uint8_t scratch[MTGL_STATIC_PIXELS_MAX];

namespace {
constexpr int SCRATCH_DEVICES = 64;
std::mutex g_scratch_setup_mu;
std::mutex g_scratch_mu[SCRATCH_DEVICES];
void *g_scratch[SCRATCH_DEVICES];

int scratch_setup(int device)
{
    if (device < 0 || device >= SCRATCH_DEVICES) return MTGL_ERR_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_setup_mu);
    if (g_scratch[device]) return MTGL_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        fprintf(stderr, "libmtgl: hipGetSymbolAddress(scratch) failed on device %d: %s\n", device,
                hipGetErrorString(e));
        return MTGL_ERR_DEVICE;
    }
    g_scratch[device] = ptr;
    return MTGL_OK;
}
}  // namespace

MTGL_API int mtgl_dev_create(int32_t width, int32_t height, mtgl_dev **out)
{
    if (!out || width <= 0 || height <= 0 || width > MTGL_MAX_FB_DIM || height > MTGL_MAX_FB_DIM) return MTGL_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        fprintf(stderr, "libmtgl: no HIP device available\n");
        return MTGL_ERR_DEVICE;
    }
    mtgl_dev *d = new (std::nothrow) mtgl_dev();
    if (!d) return MTGL_ERR_OUT_OF_MEMORY;
    int rc = MTGL_OK;
    do {
        if (hipGetDevice(&d->device) != hipSuccess) { rc = MTGL_ERR_DEVICE; break; }
        if ((rc = scratch_setup(d->device)) != MTGL_OK) break;
        if ((rc = exec_init(d->ex, d->device)) != MTGL_OK) break;
        d->w = width;
        d->h = height;
        d->tiles_x = (width + TILE - 1) / TILE;
        d->tiles_y = (height + TILE - 1) / TILE;
        const size_t n = (size_t)width * (size_t)height;
        if ((rc = map_err(hipMalloc((void **)&d->color, n * 4))) != MTGL_OK) break;
        if ((rc = map_err(hipMalloc((void **)&d->depth, n * 4))) != MTGL_OK) break;
        if ((rc = map_err(hipMalloc((void **)&d->stencil, n))) != MTGL_OK) break;
        if ((rc = map_err(hipMemsetAsync(d->color, 0, n * 4, d->ex.stream))) != MTGL_OK) break;
        if ((rc = map_err(hipMemsetAsync(d->depth, 0, n * 4, d->ex.stream))) != MTGL_OK) break;
        if ((rc = map_err(hipMemsetAsync(d->stencil, 0, n, d->ex.stream))) != MTGL_OK) break;
        if ((rc = map_err(hipStreamSynchronize(d->ex.stream))) != MTGL_OK) break;
    } while (0);
    if (rc != MTGL_OK) {
        mtgl_dev_destroy(d);
        return rc;
    }
    *out = d;
    return MTGL_OK;
}

MTGL_API void mtgl_dev_destroy(mtgl_dev *d)
{
    if (!d) return;
    invalidate_batches(d);
    {
        // batches forget the context: their graphs no longer launch
        std::lock_guard<std::mutex> lk(g_batches_mu);
        for (mtgl_batch *b : g_batches) {
            auto it = std::find(b->devs.begin(), b->devs.end(), d);
            if (it != b->devs.end()) *it = nullptr;
        }
    }
    if (d->pending) (void)hipStreamSynchronize(d->pending->ex.stream);
    if (d->ex.stream) (void)hipStreamSynchronize(d->ex.stream);
    for (int i = 0; i <= MTGL_MAX_OBJECTS; i++) {
        if (d->tex[i].px) (void)hipFree(d->tex[i].px);
        if (d->tex[i].m1) (void)hipFree(d->tex[i].m1);
        if (d->tex[i].chain) (void)hipFree(d->tex[i].chain);
        if (d->buf[i].data) (void)hipFree(d->buf[i].data);
    }
    arena_free(d->writer);
    arena_free(d->pixels);
    arena_free(d->readback);
    arena_free(d->prim_params);
    arena_free(d->prim_dep);
    arena_free(d->prim_col);
    for (auto &kv : d->segs)
        if (kv.second.data) (void)hipFree(kv.second.data);
    d->segs.clear();
    arena_free(d->one.params);
    arena_free(d->one.maps);
    exec_free(d->ex);
    if (d->color) (void)hipFree(d->color);
    if (d->depth) (void)hipFree(d->depth);
    if (d->stencil) (void)hipFree(d->stencil);
    delete d;
}

MTGL_API int mtgl_dev_finish(mtgl_dev *d)
{
    if (!d) return MTGL_ERR_INVALID;
    int rc = dev_join(d);
    if (rc) return rc;
    if ((rc = confirm(d))) return rc;
    HIPC(hipStreamSynchronize(d->ex.stream));
    return MTGL_OK;
}

MTGL_API int mtgl_dev_clear(mtgl_dev *d, uint32_t buffers, int32_t x0, int32_t y0, int32_t x1, int32_t y1, uint32_t color,
                            float depth, uint8_t stencil)
{
    if (!d) return MTGL_ERR_INVALID;
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    if (x1 > d->w) x1 = d->w;
    if (y1 > d->h) y1 = d->h;
    if (x0 >= x1 || y0 >= y1 || !(buffers & (MTGL_BUF_COLOR | MTGL_BUF_DEPTH | MTGL_BUF_STENCIL))) return MTGL_OK;
    ClearDesc c;
    c.color = (buffers & MTGL_BUF_COLOR) ? d->color : nullptr;
    c.depth = (buffers & MTGL_BUF_DEPTH) ? d->depth : nullptr;
    c.stencil = (buffers & MTGL_BUF_STENCIL) ? d->stencil : nullptr;
    c.fb_w = d->w;
    c.x0 = x0;
    c.y0 = y0;
    c.w = x1 - x0;
    c.h = y1 - y0;
    c.cv = color;
    c.dv = depth;
    c.sv = stencil;
    if (g_recording) {
        BatchCmd *cmd = new (std::nothrow) BatchCmd();
        if (!cmd) return MTGL_ERR_OUT_OF_MEMORY;
        cmd->dev = d;
        cmd->is_clear = true;
        cmd->clear = c;
        g_recording->cmds.push_back(cmd);
        batch_note_dev(g_recording, d);
        return MTGL_OK;
    }
    int rc = dev_join(d);
    if (rc) return rc;
    PendingOp op;
    op.seq = next_seq(d);
    op.kind = PendingOp::CLEAR;
    op.buffers = buffers;
    op.x0 = x0;
    op.y0 = y0;
    op.x1 = x1;
    op.y1 = y1;
    op.color = color;
    op.depth = depth;
    op.stencil = stencil;
    if ((rc = launch_clear(d, op))) return rc;
    d->ops.push_back(op);
    return MTGL_OK;
}

static int plane_of(mtgl_dev *d, uint32_t which, uint8_t **base, size_t *esz)
{
    if (which == MTGL_BUF_COLOR) { *base = (uint8_t *)d->color; *esz = 4; return 1; }
    if (which == MTGL_BUF_DEPTH) { *base = (uint8_t *)d->depth; *esz = 4; return 1; }
    if (which == MTGL_BUF_STENCIL) { *base = d->stencil; *esz = 1; return 1; }
    return 0;
}

MTGL_API int mtgl_dev_read(mtgl_dev *d, uint32_t which, int32_t x, int32_t y, int32_t w, int32_t h, void *dst)
{
    uint8_t *base;
    size_t esz;
    if (!d || !dst || !plane_of(d, which, &base, &esz)) return MTGL_ERR_INVALID;
    if (x < 0 || y < 0 || w < 0 || h < 0 || x + w > d->w || y + h > d->h) return MTGL_ERR_INVALID;
    if (w == 0 || h == 0) return MTGL_OK;
    int rc = dev_join(d);
    if (rc) return rc;
    if ((rc = confirm(d))) return rc;
    HIPC(hipStreamSynchronize(d->ex.stream));
    HIPC(hipMemcpy2D(dst, (size_t)w * esz, base + ((size_t)y * d->w + x) * esz, (size_t)d->w * esz, (size_t)w * esz,
                     (size_t)h, hipMemcpyDeviceToHost));
    return MTGL_OK;
}

// Packed on the device in bands of rows that fit READBACK_BAND bytes (one row
// at least), each copied to the caller's memory as soon as it is packed.
constexpr size_t READBACK_BAND = 64u << 20;

MTGL_API int mtgl_dev_read_pixels(mtgl_dev *d, int32_t x, int32_t y, int32_t width, int32_t height, uint32_t format,
                                  void *dst)
{
    if (!d || !dst) return MTGL_ERR_INVALID;
    const int32_t bpp = format == GL_RGBA ? 4 : (format == GL_RGB ? 3 : 0);
    if (width <= 0 || height <= 0 || bpp == 0) return MTGL_OK;
    int rc = dev_join(d);
    if (rc) return rc;
    if ((rc = confirm(d))) return rc;
    const size_t row_bytes = (size_t)width * (size_t)bpp;
    const int32_t band = (int32_t)std::max<size_t>(1, std::min<size_t>((size_t)height, READBACK_BAND / row_bytes));
    if ((rc = arena_reserve(d->ex, d->readback, row_bytes * (size_t)band))) return rc;
    uint8_t *out = (uint8_t *)dst;
    for (int32_t r0 = 0; r0 < height; r0 += band) {
        const int32_t nr = std::min(band, height - r0);
        const uint64_t n = (uint64_t)width * (uint64_t)nr;
        hipLaunchKernelGGL(k_read_pixels, dim3((uint32_t)blocks_for(n, 256)), dim3(256), 0, d->ex.stream, d->color,
                           d->w, d->h, x, y, width, r0, nr, bpp, as<uint8_t>(d->readback));
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(out + (size_t)r0 * row_bytes, d->readback.ptr, row_bytes * (size_t)nr,
                            hipMemcpyDeviceToHost, d->ex.stream));
    }
    HIPC(hipStreamSynchronize(d->ex.stream));
    return MTGL_OK;
}

MTGL_API int mtgl_dev_present(mtgl_dev *d, void *dst, int64_t pitch)
{
    if (!d || !dst || pitch < (int64_t)d->w * 4) return MTGL_ERR_INVALID;
    int rc = dev_join(d);
    if (rc) return rc;
    if ((rc = confirm(d))) return rc;
    HIPC(hipStreamSynchronize(d->ex.stream));
    HIPC(hipMemcpy2D(dst, (size_t)pitch, d->color, (size_t)d->w * 4, (size_t)d->w * 4, (size_t)d->h,
                     hipMemcpyDeviceToHost));
    return MTGL_OK;
}

MTGL_API int mtgl_dev_write(mtgl_dev *d, uint32_t which, int32_t x, int32_t y, int32_t w, int32_t h, const void *src)
{
    uint8_t *base;
    size_t esz;
    if (!d || !src || !plane_of(d, which, &base, &esz)) return MTGL_ERR_INVALID;
    if (x < 0 || y < 0 || w < 0 || h < 0 || x + w > d->w || y + h > d->h) return MTGL_ERR_INVALID;
    if (w == 0 || h == 0) return MTGL_OK;
    int rc = dev_join(d);
    if (rc) return rc;
    if ((rc = confirm(d))) return rc;
    HIPC(hipStreamSynchronize(d->ex.stream));
    HIPC(hipMemcpy2D(base + ((size_t)y * d->w + x) * esz, (size_t)d->w * esz, src, (size_t)w * esz, (size_t)w * esz,
                     (size_t)h, hipMemcpyHostToDevice));
    return MTGL_OK;
}

MTGL_API int mtgl_dev_tex_image(mtgl_dev *d, uint32_t id, int32_t w, int32_t h, const uint32_t *texels)
{
    if (!d || id == 0 || id > MTGL_MAX_OBJECTS || w <= 0 || h <= 0 || w > MTGL_MAX_TEX_SIZE || h > MTGL_MAX_TEX_SIZE ||
        !texels)
        return MTGL_ERR_INVALID;
    int rc = dev_join(d);
    if (rc) return rc;
    if ((rc = confirm(d))) return rc;
    invalidate_batches(d);
    const size_t n = (size_t)w * (size_t)h;
    uint32_t *px = nullptr, *m1 = nullptr;
    int32_t mw = 0, mh = 0;
    HIPC(hipMalloc((void **)&px, n * 4));
    if (w >= 2 && h >= 2) {
        mw = w / 2;
        mh = h / 2;
        hipError_t e = hipMalloc((void **)&m1, (size_t)mw * mh * 4);
        if (e != hipSuccess) {
            (void)hipFree(px);
            return map_err(e);
        }
    }
    HIPC(hipMemcpy(px, texels, n * 4, hipMemcpyHostToDevice));
    if (m1) {
        hipLaunchKernelGGL(k_mip1, dim3(blocks_for((uint64_t)mw * mh, 256)), dim3(256), 0, d->ex.stream, px, w, m1, mw,
                           mh);
        HIPC(hipGetLastError());
    }
    // the rest of the chain (sampled only under MTGL_LOD_FULL_CHAIN): each
    // level the box filter of the one before, while both sides are >= 2
    int32_t nlev = m1 ? 2 : 1;
    size_t chain_n = 0;
    for (int32_t lw = mw, lh = mh; lw >= 2 && lh >= 2; lw /= 2, lh /= 2) chain_n += (size_t)(lw / 2) * (lh / 2);
    uint32_t *chain = nullptr;
    if (chain_n) {
        hipError_t e = hipMalloc((void **)&chain, chain_n * 4);
        if (e != hipSuccess) {
            (void)hipFree(px);
            if (m1) (void)hipFree(m1);
            return map_err(e);
        }
        const uint32_t *src = m1;
        uint32_t *dst = chain;
        for (int32_t lw = mw, lh = mh; lw >= 2 && lh >= 2; lw /= 2, lh /= 2) {
            hipLaunchKernelGGL(k_mip1, dim3(blocks_for((uint64_t)(lw / 2) * (lh / 2), 256)), dim3(256), 0,
                               d->ex.stream, src, lw, dst, lw / 2, lh / 2);
            src = dst;
            dst += (size_t)(lw / 2) * (lh / 2);
            nlev++;
        }
        HIPC(hipGetLastError());
    }
    HIPC(hipStreamSynchronize(d->ex.stream));
    DevTex &t = d->tex[id];
    if (t.px) HIPC(hipFree(t.px));
    if (t.m1) HIPC(hipFree(t.m1));
    if (t.chain) HIPC(hipFree(t.chain));
    t.px = px;
    t.w = w;
    t.h = h;
    t.m1 = m1;
    t.m1w = mw;
    t.m1h = mh;
    t.chain = chain;
    t.nlev = nlev;
    return MTGL_OK;
}

MTGL_API int mtgl_dev_tex_free(mtgl_dev *d, uint32_t id)
{
    if (!d || id == 0 || id > MTGL_MAX_OBJECTS) return MTGL_ERR_INVALID;
    int rc = dev_join(d);
    if (rc) return rc;
    if ((rc = confirm(d))) return rc;
    HIPC(hipStreamSynchronize(d->ex.stream));
    DevTex &t = d->tex[id];
    if (t.px || t.m1) invalidate_batches(d);
    if (t.px) HIPC(hipFree(t.px));
    if (t.m1) HIPC(hipFree(t.m1));
    if (t.chain) HIPC(hipFree(t.chain));
    t = DevTex();
    return MTGL_OK;
}

MTGL_API int mtgl_dev_buffer_data(mtgl_dev *d, uint32_t id, int64_t size, const void *data)
{
    if (!d || id == 0 || id > MTGL_MAX_OBJECTS) return MTGL_ERR_INVALID;
    int rc = dev_join(d);
    if (rc) return rc;
    if ((rc = confirm(d))) return rc;
    HIPC(hipStreamSynchronize(d->ex.stream));
    DevBuf &b = d->buf[id];
    if (b.data) invalidate_batches(d);
    if (size <= 0) {
        if (b.data) HIPC(hipFree(b.data));
        b = DevBuf();
        return MTGL_OK;
    }
    uint8_t *p = nullptr;
    HIPC(hipMalloc((void **)&p, (size_t)size));
    hipError_t e = data ? hipMemcpy(p, data, (size_t)size, hipMemcpyHostToDevice) : hipMemset(p, 0, (size_t)size);
    if (e != hipSuccess) {
        (void)hipFree(p);
        return map_err(e);
    }
    if (b.data) HIPC(hipFree(b.data));
    b.data = p;
    b.size = size;
    return MTGL_OK;
}

MTGL_API int mtgl_dev_buffer_sub_data(mtgl_dev *d, uint32_t id, int64_t offset, int64_t size, const void *data)
{
    if (!d || id == 0 || id > MTGL_MAX_OBJECTS || !data || offset < 0 || size < 0) return MTGL_ERR_INVALID;
    DevBuf &b = d->buf[id];
    if (!b.data || offset + size > b.size) return MTGL_ERR_INVALID;
    int rc = dev_join(d);
    if (rc) return rc;
    if ((rc = confirm(d))) return rc;
    HIPC(hipStreamSynchronize(d->ex.stream));
    HIPC(hipMemcpy(b.data + offset, data, (size_t)size, hipMemcpyHostToDevice));
    return MTGL_OK;
}

MTGL_API int mtgl_dev_buffer_free(mtgl_dev *d, uint32_t id)
{
    if (!d || id == 0 || id > MTGL_MAX_OBJECTS) return MTGL_ERR_INVALID;
    int rc = dev_join(d);
    if (rc) return rc;
    if ((rc = confirm(d))) return rc;
    HIPC(hipStreamSynchronize(d->ex.stream));
    DevBuf &b = d->buf[id];
    if (b.data) invalidate_batches(d);
    if (b.data) HIPC(hipFree(b.data));
    b = DevBuf();
    return MTGL_OK;
}

MTGL_API int mtgl_dev_segment_data(mtgl_dev *d, uint32_t id, int64_t size, const void *data)
{
    if (!d || id == 0 || id > MTGL_MAX_SEGMENT) return MTGL_ERR_INVALID;
    if (size <= 0) return mtgl_dev_segment_free(d, id);
    int rc = dev_join(d);
    if (rc) return rc;
    if ((rc = confirm(d))) return rc;
    HIPC(hipStreamSynchronize(d->ex.stream));
    DevBuf &b = d->segs[id];
    // a recorded batch holds the old allocation's address
    if (b.data) invalidate_batches(d);
    d->ex.stats.upload_bytes += (uint64_t)size;
    if (b.data && b.size == size) {
        if (data) HIPC(hipMemcpy(b.data, data, (size_t)size, hipMemcpyHostToDevice));
        return MTGL_OK;
    }
    uint8_t *p = nullptr;
    HIPC(hipMalloc((void **)&p, (size_t)size));
    hipError_t e = data ? hipMemcpy(p, data, (size_t)size, hipMemcpyHostToDevice) : hipMemset(p, 0, (size_t)size);
    if (e != hipSuccess) {
        (void)hipFree(p);
        return map_err(e);
    }
    if (b.data) HIPC(hipFree(b.data));
    b.data = p;
    b.size = size;
    return MTGL_OK;
}

MTGL_API int mtgl_dev_segment_free(mtgl_dev *d, uint32_t id)
{
    if (!d || id == 0 || id > MTGL_MAX_SEGMENT) return MTGL_ERR_INVALID;
    const auto it = d->segs.find(id);
    if (it == d->segs.end()) return MTGL_OK;
    int rc = dev_join(d);
    if (rc) return rc;
    if ((rc = confirm(d))) return rc;
    HIPC(hipStreamSynchronize(d->ex.stream));
    invalidate_batches(d);
    if (it->second.data) HIPC(hipFree(it->second.data));
    d->segs.erase(it);
    return MTGL_OK;
}

MTGL_API int mtgl_dev_draw(mtgl_dev *d, const mtgl_draw *dr)
{
    if (!d || !dr || !dr->raster || !dr->vstates || dr->n_vstates == 0) return MTGL_ERR_INVALID;
    const int cls = prim_class(dr);
    if (!g_recording) d->ex.stats.upload_bytes += dr->staged_bytes;
    if (cls != 0) {
        // a recorded batch replays filled triangles only
        if (g_recording) return MTGL_ERR_UNSUPPORTED;
        d->ex.stats.draws++;
        int rc = dev_join(d);
        if (rc) return rc;
        // the point and line path waits for its own counts: the queued work first
        if ((rc = confirm(d))) return rc;
        d->last.valid = false;
        rc = run_prims(d, dr, cls);
        d->dirty = true;
        return rc;
    }
    if (!draw_mode_ok(dr->mode)) return MTGL_ERR_UNSUPPORTED;
    const uint32_t ntri = tri_count(dr->mode, dr->count);
    if (g_recording) {
        BatchCmd *cmd = new (std::nothrow) BatchCmd();
        if (!cmd) return MTGL_ERR_OUT_OF_MEMORY;
        cmd->dev = d;
        cmd->draw = *dr;
        cmd->raster = *dr->raster;
        cmd->vstates.assign(dr->vstates, dr->vstates + dr->n_vstates);
        if (dr->staged_bytes) {
            const uint8_t *st = (const uint8_t *)dr->staged;
            cmd->staged.assign(st, st + dr->staged_bytes);
        }
        cmd->draw.raster = &cmd->raster;
        cmd->draw.vstates = cmd->vstates.data();
        cmd->draw.staged = cmd->staged.empty() ? nullptr : cmd->staged.data();
        g_recording->cmds.push_back(cmd);
        batch_note_dev(g_recording, d);
        return MTGL_OK;
    }
    d->ex.stats.draws++;
    if (ntri == 0) return MTGL_OK;
    d->ex.stats.triangles_in += ntri;
    int rc = dev_join(d);
    if (rc) return rc;
    if (d->ex.pending_events.size() > 4096) (void)resolve_timing(d->ex);
    const size_t staging = sizeof(mtgl_raster_state) + (size_t)dr->n_vstates * sizeof(mtgl_vertex_state) +
                           (size_t)dr->staged_bytes + 4096;
    if (d->ex.ring_used + staging > RING_LIMIT && (rc = confirm(d))) return rc;
    SceneIn in = {d, dr};
    d->one.pairs_cap = 0;
    d->last.valid = false;
    const uint32_t seq = next_seq(d);
    rc = execute_scenes(d->ex, d->one, &in, 1, seq);
    d->dirty = true;
    if ((rc == MTGL_OK || rc == MTGL_ERR_UNSUPPORTED) && d->one.ns == 1) {
        // logged with its inputs where the launch staged them
        PendingOp op;
        op.seq = seq;
        op.kind = PendingOp::DRAW;
        op.draw = *dr;
        op.draw.raster = (const mtgl_raster_state *)d->one.host_rs;
        op.draw.vstates = (const mtgl_vertex_state *)d->one.host_vs;
        op.draw.staged = dr->staged_bytes ? d->one.host_st : nullptr;
        op.ntri = ntri;
        d->ops.push_back(op);
    }
    if (rc == MTGL_OK && d->one.ns == 1) {
        LastDraw &l = d->last;
        const DevTex *tx = bound_texture(d, dr->raster);
        l.valid = true;
        l.mode = dr->mode;
        l.ntri = ntri;
        l.n_vstates = dr->n_vstates;
        l.color_array = dr->color.enabled != 0;
        l.rs = *dr->raster;
        l.vs = dr->vstates[0];
        l.kind = raster_kind(dr->raster, tx);
        l.gen = d->ex.gen;
        l.over_base = d->one.tri_total;
        l.tex_w = tx ? tx->w : 0;
        l.tex_h = tx ? tx->h : 0;
        l.captured = d->capture;
    }
    return rc;
}


static int launch_pixels(mtgl_dev *d, const mtgl_pixels &px, uint64_t bytes);

// the launch both glDrawPixels entry points share: src holds the image in HBM
static int draw_pixels_kernel(mtgl_dev *d, const mtgl_pixels *p, const uint8_t *src, uint64_t n, bool timed)
{
    Exec &x = d->ex;
    PixDesc P;
    P.x = p->x;
    P.y = p->y;
    P.w = p->width;
    P.h = p->height;
    P.fb_w = d->w;
    P.fb_h = d->h;
    P.format = p->format;
    P.flags = p->flags;
    P.alpha_func = p->alpha_func;
    P.depth_func = p->depth_func;
    P.depth_mask = p->depth_mask;
    P.blend_src = p->blend_src;
    P.blend_dst = p->blend_dst;
    P.alpha_ref = p->alpha_ref;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (timed && x.timing && x.timing_count++ % x.timing_every == 0u) (void)take_event_pair(x, ev0, ev1);
    const Counters *ctr = (const Counters *)x.ctr;
    void *args[] = {&P, &src, &d->color, &d->depth, &ctr};
    // a timed launch takes its events at the kernel's own start and end
    HIPC(hipExtLaunchKernel((const void *)k_draw_pixels, dim3(blocks_for(n, 256)), dim3(256), args, 0, x.stream, ev0,
                            ev1, 0));
    if (ev0) x.pending_events.emplace_back(ev0, ev1);
    d->dirty = true;
    return MTGL_OK;
}

// source bytes per pixel of a glDrawPixels format, 0 for formats the
// reference reads nothing of
static uint32_t draw_pixels_bpp(uint32_t format)
{
    switch (format) {
    case GL_RGBA: return 4;
    case GL_RGB: return 3;
    case GL_LUMINANCE: return 1;
    case GL_LUMINANCE_ALPHA: return 2;
    default: return 0;
    }
}

MTGL_API int mtgl_dev_draw_pixels_static(mtgl_dev *d, const mtgl_pixels *p)
{
    if (!d || !p || !p->pixels || p->width < 0 || p->height < 0) return MTGL_ERR_INVALID;
    const uint32_t bpp = draw_pixels_bpp(p->format);
    if (bpp == 0) return MTGL_OK;
    const uint64_t n = (uint64_t)p->width * (uint64_t)p->height;
    if (n == 0) return MTGL_OK;
    if (n > 0xFFFFFFFFull || n * bpp > (uint64_t)MTGL_STATIC_PIXELS_MAX) {
        fprintf(stderr, "libmtgl: mtgl_dev_draw_pixels_static: %llu source bytes exceed the %llu-byte buffer\n",
                (unsigned long long)(n * bpp), (unsigned long long)MTGL_STATIC_PIXELS_MAX);
        return MTGL_ERR_INVALID;
    }
    if (g_recording) return MTGL_ERR_UNSUPPORTED;
    if (d->device < 0 || d->device >= SCRATCH_DEVICES || !g_scratch[d->device]) return MTGL_ERR_DEVICE;
    int rc = dev_join(d);
    if (rc) return rc;
    if (!d->ops.empty()) {
        // the queued work must have taken effect; running an overflowed
        // launch again would allocate, which this entry point may not do
        HIPC(hipStreamSynchronize(d->ex.stream));
        HIPC(hipMemcpy(d->ex.ctr_host, d->ex.ctr, sizeof(Counters), hipMemcpyDeviceToHost));
        if (d->ex.ctr_host->poison) {
            fprintf(stderr, "libmtgl: mtgl_dev_draw_pixels_static: earlier draws need larger buffers; "
                            "call mtgl_dev_finish first\n");
            return MTGL_ERR_INVALID;
        }
        d->ops.clear();
        ring_release(d->ex);
    }
    uint8_t *buf = (uint8_t *)g_scratch[d->device];
    std::lock_guard<std::mutex> lock(g_scratch_mu[d->device]);
    HIPC(hipMemcpyAsync(buf, p->pixels, (size_t)(n * bpp), hipMemcpyHostToDevice, d->ex.stream));
    if ((rc = draw_pixels_kernel(d, p, buf, n, false))) return rc;
    HIPC(hipStreamSynchronize(d->ex.stream));
    return MTGL_OK;
}

MTGL_API int mtgl_dev_draw_pixels(mtgl_dev *d, const mtgl_pixels *p)
{
    if (!d || !p || !p->pixels || p->width < 0 || p->height < 0) return MTGL_ERR_INVALID;
    const uint32_t bpp = draw_pixels_bpp(p->format);
    if (bpp == 0) return MTGL_OK;   // the reference reads nothing of other formats
    const uint64_t n = (uint64_t)p->width * (uint64_t)p->height;
    if (n == 0) return MTGL_OK;
    if (n > 0xFFFFFFFFull) return MTGL_ERR_INVALID;
    // a recorded batch replays clears and draws only: out of order otherwise
    if (g_recording) return MTGL_ERR_UNSUPPORTED;
    int rc = dev_join(d);
    if (rc) return rc;
    Exec &x = d->ex;
    const size_t bytes = (size_t)n * bpp;
    if (bytes >= PIXELS_DIRECT) {
        // a large image is copied straight from the caller's memory once the
        // queued work has taken effect (the log is then empty, so the image
        // never needs to be drawn again): one wait instead of a host copy
        // into the staging ring
        if ((rc = confirm(d))) return rc;
        if ((rc = arena_reserve(x, d->pixels, bytes))) return rc;
        HIPC(hipMemcpy(d->pixels.ptr, p->pixels, bytes, hipMemcpyHostToDevice));
        return draw_pixels_kernel(d, p, (const uint8_t *)d->pixels.ptr, n, true);
    }
    if (x.ring_used + bytes > RING_LIMIT && (rc = confirm(d))) return rc;
    uint8_t *host = ring_take(x, bytes);
    if (!host) return MTGL_ERR_OUT_OF_MEMORY;
    memcpy(host, p->pixels, bytes);
    PendingOp op;
    op.seq = next_seq(d);
    op.kind = PendingOp::PIXELS;
    op.px = *p;
    op.px.pixels = host;
    if ((rc = launch_pixels(d, op.px, n * bpp))) return rc;
    d->ops.push_back(op);
    return MTGL_OK;
}

// (a staged glDrawPixels source, from its first run or a replay)
static int launch_pixels(mtgl_dev *d, const mtgl_pixels &px, uint64_t bytes)
{
    Exec &x = d->ex;
    int rc;
    if ((rc = arena_reserve(x, d->pixels, (size_t)bytes))) return rc;
    HIPC(hipMemcpyAsync(d->pixels.ptr, px.pixels, (size_t)bytes, hipMemcpyHostToDevice, x.stream));
    return draw_pixels_kernel(d, &px, (const uint8_t *)d->pixels.ptr, (uint64_t)px.width * (uint64_t)px.height,
                              !d->replaying);
}

// Waits for the context's queued work and makes sure every logged clear,
// draw and glDrawPixels took effect.  A launch that overflowed its arrays
// poisoned the stream (every later framebuffer write was skipped): the
// arrays grow, and that launch and every later logged op run again, in
// order, until none overflows.  Then the log and its staging are released.
static int confirm(mtgl_dev *d)
{
    if (d->ops.empty()) return MTGL_OK;
    Exec &x = d->ex;
    for (int attempt = 0;; attempt++) {
        HIPC(hipStreamSynchronize(x.stream));
        uint32_t poison = 0;
        int rc = exec_take_poison(x, poison);
        if (rc) return rc;
        if (!poison) break;
        size_t first = d->ops.size();
        for (size_t i = 0; i < d->ops.size(); i++)
            if (d->ops[i].seq == poison) first = i;
        if (first == d->ops.size() || attempt >= 16) return MTGL_ERR_OUT_OF_MEMORY;
        exec_grow(x, *x.ctr_host, d->ops[first].ntri);
        d->replaying = true;
        const bool timing = x.timing;
        x.timing = false;
        for (size_t i = first; i < d->ops.size() && !rc; i++) {
            const PendingOp &op = d->ops[i];
            if (op.kind == PendingOp::CLEAR) {
                rc = launch_clear(d, op);
            } else if (op.kind == PendingOp::DRAW) {
                SceneIn in = {d, &op.draw};
                rc = execute_scenes(x, d->one, &in, 1, op.seq);
                if (rc == MTGL_ERR_UNSUPPORTED) rc = MTGL_OK;
            } else {
                rc = launch_pixels(d, op.px, (uint64_t)op.px.width * op.px.height * draw_pixels_bpp(op.px.format));
            }
        }
        x.timing = timing;
        d->replaying = false;
        if (rc) return rc;
    }
    d->ops.clear();
    ring_release(x);
    return MTGL_OK;
}

MTGL_API int mtgl_dev_set_timing(mtgl_dev *d, int every)
{
    if (!d || every < 0) return MTGL_ERR_INVALID;
    d->ex.timing = every != 0;
    d->ex.timing_every = every ? (uint32_t)every : 1u;
    d->ex.timing_count = 0;
    return MTGL_OK;
}

MTGL_API int mtgl_dev_stats_get(mtgl_dev *d, mtgl_dev_stats *out)
{
    if (!d || !out) return MTGL_ERR_INVALID;
    int rc = confirm(d);
    if (rc) return rc;
    return exec_stats(d->ex, out);
}

MTGL_API int mtgl_dev_stats_reset(mtgl_dev *d)
{
    if (!d) return MTGL_ERR_INVALID;
    int rc = confirm(d);
    if (rc) return rc;
    return exec_stats_reset(d->ex);
}

// ---- batches

MTGL_API int mtgl_batch_create(mtgl_batch **out)
{
    if (!out) return MTGL_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MTGL_ERR_DEVICE;
    mtgl_batch *b = new (std::nothrow) mtgl_batch();
    if (!b) return MTGL_ERR_OUT_OF_MEMORY;
    int rc = MTGL_OK;
    if (hipGetDevice(&b->device) != hipSuccess) rc = MTGL_ERR_DEVICE;
    if (!rc) rc = exec_init(b->ex, b->device);
    if (!rc) rc = map_err(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
    if (!rc) rc = map_err(hipEventCreateWithFlags(&b->join, hipEventDisableTiming));
    if (rc) {
        mtgl_batch_destroy(b);
        return rc;
    }
    {
        std::lock_guard<std::mutex> lk(g_batches_mu);
        g_batches.push_back(b);
    }
    *out = b;
    return MTGL_OK;
}

MTGL_API void mtgl_batch_destroy(mtgl_batch *b)
{
    if (!b) return;
    if (g_recording == b) g_recording = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_batches_mu);
        g_batches.erase(std::remove(g_batches.begin(), g_batches.end(), b), g_batches.end());
    }
    if (b->ex.stream) (void)hipStreamSynchronize(b->ex.stream);
    for (mtgl_dev *d : b->devs)
        if (d && d->pending == b) d->pending = nullptr;   // the launch has finished
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (BatchRound &R : b->rounds) {
        arena_free(R.launch.params);
        arena_free(R.launch.maps);
        arena_free(R.clear_params);
    }
    for (BatchCmd *c : b->cmds) delete c;
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->join) (void)hipEventDestroy(b->join);
    exec_free(b->ex);
    delete b;
}

MTGL_API int mtgl_batch_begin(mtgl_batch *b)
{
    if (!b || g_recording || b->recorded || !b->cmds.empty()) return MTGL_ERR_INVALID;
    g_recording = b;
    return MTGL_OK;
}

MTGL_API int mtgl_batch_end(mtgl_batch *b, int autotune)
{
    if (!b || g_recording != b) return MTGL_ERR_INVALID;
    g_recording = nullptr;
    b->recorded = true;
    b->info.contexts = (uint32_t)b->devs.size();
    b->info.commands = (uint32_t)b->cmds.size();
    for (mtgl_dev *d : b->devs)
        if (!d) return MTGL_ERR_INVALID;   // a context was destroyed while recording
    int rc = batch_build(b, autotune);
    b->info.rounds = (uint32_t)b->rounds.size();
    uint64_t scenes = 0;
    for (BatchRound &R : b->rounds) scenes += R.launch.ns;
    b->info.scenes = (uint32_t)scenes;
    if (rc) return rc;
    b->ready = true;
    return MTGL_OK;
}

MTGL_API int mtgl_batch_launch(mtgl_batch *b)
{
    if (!b || !b->ready) return MTGL_ERR_INVALID;
    Exec &x = b->ex;
    // after the contexts' queued work (which has taken effect) ...
    for (mtgl_dev *d : b->devs) {
        int rc = confirm(d);
        if (rc) return rc;
        if (d->pending && d->pending != b) {
            HIPC(hipStreamWaitEvent(x.stream, d->pending->done, 0));
        }
        if (d->dirty) {
            HIPC(hipEventRecord(b->join, d->ex.stream));
            HIPC(hipStreamWaitEvent(x.stream, b->join, 0));
            d->dirty = false;
        }
    }
    if (b->graph_exec) HIPC(hipGraphLaunch(b->graph_exec, x.stream));
    HIPC(hipEventRecord(b->done, x.stream));
    // ... and before their next work (joined lazily)
    for (mtgl_dev *d : b->devs) d->pending = b;
    b->info.launches++;
    return MTGL_OK;
}

MTGL_API int mtgl_batch_sync(mtgl_batch *b)
{
    if (!b) return MTGL_ERR_INVALID;
    HIPC(hipStreamSynchronize(b->ex.stream));
    Counters *c = b->ex.ctr_host;
    HIPC(hipMemcpy(c, b->ex.ctr, sizeof(Counters), hipMemcpyDeviceToHost));
    const unsigned int over = c->sticky | c->overflow_full | c->poison;
    if (over) {
        // a replay produced more clipped triangles or pairs than the sizing
        // run (buffer objects changed since): its output is incomplete
        HIPC(hipMemsetAsync(b->ex.ctr, 0, sizeof(Counters), b->ex.stream));
        HIPC(hipStreamSynchronize(b->ex.stream));
        return MTGL_ERR_OUT_OF_MEMORY;
    }
    return MTGL_OK;
}

MTGL_API int mtgl_batch_info(mtgl_batch *b, mtgl_batch_info_t *out)
{
    if (!b || !out) return MTGL_ERR_INVALID;
    *out = b->info;
    return MTGL_OK;
}

MTGL_API int mtgl_batch_stats_get(mtgl_batch *b, mtgl_dev_stats *out)
{
    if (!b || !out) return MTGL_ERR_INVALID;
    return exec_stats(b->ex, out);
}

MTGL_API int mtgl_batch_stats_reset(mtgl_batch *b)
{
    if (!b) return MTGL_ERR_INVALID;
    return exec_stats_reset(b->ex);
}

// ---- differentiable shading

namespace {

struct DiffCapture {
    Arena gb, elems, pos, nrm, vs;
    uint32_t ntri = 0;
};

// the capture conditions of include/mtgl_diff.h
bool diff_capturable(const LastDraw &l, int32_t tex_w, int32_t tex_h)
{
    const mtgl_raster_state &rs = l.rs;
    const mtgl_vertex_state &vs = l.vs;
    if (!l.valid || !l.captured || l.mode != GL_TRIANGLES || l.n_vstates != 1 || !l.color_array) return false;
    if ((l.kind & (KIND_LAST_WRITER | KIND_COLOR)) != (KIND_LAST_WRITER | KIND_COLOR) || (rs.color_mask & 0xFu) != 0xFu)
        return false;
    if (rs.flags & (MTGL_F_BLEND | MTGL_F_ALPHA_TEST | MTGL_F_FOG)) return false;
    if (!rs.texture || l.tex_w != tex_w || l.tex_h != tex_h) return false;
    if (rs.min_filter != GL_LINEAR || rs.mag_filter != GL_LINEAR || rs.tex_env_mode != GL_MODULATE) return false;
    if (rs.shade_model == GL_FLAT || rs.shade_model == GL_PHONG || rs.lighting.two_side) return false;
    if (!(vs.flags & MTGL_F_LIGHTING) || !(vs.flags & MTGL_F_COLOR_MATERIAL) || vs.shade_model == GL_PHONG) return false;
    if (vs.color_material_face != GL_FRONT && vs.color_material_face != GL_FRONT_AND_BACK) return false;
    const uint32_t m = vs.color_material_mode;
    if (m != GL_AMBIENT && m != GL_DIFFUSE && m != GL_AMBIENT_AND_DIFFUSE) return false;
    for (int i = 0; i < MTGL_MAX_LIGHTS; i++)
        if (vs.lighting.lights[i].enabled && vs.lighting.lights[i].position[3] != 0.0f) return false;
    return true;
}

}  // namespace

struct mtgl_diff {
    int device = 0;
    uint32_t nvert = 0;
    int32_t tex_w = 0, tex_h = 0, w = 0, h = 0;
    std::vector<DiffCapture> views;
    Arena table, lit, counter;
    bool table_ok = false;
};

namespace {

int diff_table(mtgl_diff *df, hipStream_t st)
{
    if (df->table_ok) return MTGL_OK;
    std::vector<DiffView> t(df->views.size());
    for (size_t i = 0; i < t.size(); i++) {
        DiffCapture &c = df->views[i];
        t[i].gb = as<GPx>(c.gb);
        t[i].elems = as<uint32_t>(c.elems);
        t[i].pos = as<float>(c.pos);
        t[i].nrm = as<float>(c.nrm);
        t[i].vs = as<mtgl_vertex_state>(c.vs);
        t[i].ntri = c.ntri;
        t[i].w = df->w;
        t[i].h = df->h;
        t[i].pad = 0;
    }
    if (df->table.cap < t.size() * sizeof(DiffView)) {
        if (df->table.ptr) HIPC(hipFree(df->table.ptr));
        df->table = Arena();
        HIPC(hipMalloc(&df->table.ptr, std::max<size_t>(t.size() * sizeof(DiffView), 256)));
        df->table.cap = std::max<size_t>(t.size() * sizeof(DiffView), 256);
    }
    HIPC(hipMemcpy(df->table.ptr, t.data(), t.size() * sizeof(DiffView), hipMemcpyHostToDevice));
    const size_t lit_bytes = t.size() * (size_t)df->nvert * sizeof(float4);
    if (df->lit.cap < lit_bytes) {
        if (df->lit.ptr) HIPC(hipFree(df->lit.ptr));
        df->lit = Arena();
        HIPC(hipMalloc(&df->lit.ptr, lit_bytes));
        df->lit.cap = lit_bytes;
    }
    (void)st;
    df->table_ok = true;
    return MTGL_OK;
}

int fill_zero(void *p, size_t floats, hipStream_t st)
{
    for (size_t off = 0; off < floats; off += 0x40000000u) {
        const uint32_t n = (uint32_t)std::min<size_t>(floats - off, 0x40000000u);
        hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256)), dim3(256), 0, st, (uint32_t *)p + off, n, 0u);
    }
    HIPC(hipGetLastError());
    return MTGL_OK;
}

}  // namespace

MTGL_API int mtgl_diff_create(uint32_t vertices, int32_t tex_w, int32_t tex_h, mtgl_diff **out)
{
    if (!out || vertices == 0 || tex_w <= 0 || tex_h <= 0 || tex_w > MTGL_MAX_TEX_SIZE || tex_h > MTGL_MAX_TEX_SIZE)
        return MTGL_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MTGL_ERR_DEVICE;
    mtgl_diff *df = new (std::nothrow) mtgl_diff();
    if (!df) return MTGL_ERR_OUT_OF_MEMORY;
    if (hipGetDevice(&df->device) != hipSuccess || hipMalloc(&df->counter.ptr, 256) != hipSuccess) {
        delete df;
        return MTGL_ERR_DEVICE;
    }
    df->counter.cap = 256;
    df->nvert = vertices;
    df->tex_w = tex_w;
    df->tex_h = tex_h;
    *out = df;
    return MTGL_OK;
}

MTGL_API void mtgl_diff_destroy(mtgl_diff *df)
{
    if (!df) return;
    (void)hipDeviceSynchronize();
    for (DiffCapture &c : df->views) {
        arena_free(c.gb);
        arena_free(c.elems);
        arena_free(c.pos);
        arena_free(c.nrm);
        arena_free(c.vs);
    }
    arena_free(df->table);
    arena_free(df->lit);
    arena_free(df->counter);
    delete df;
}

MTGL_API int mtgl_diff_arm(mtgl_dev *d)
{
    if (!d) return MTGL_ERR_INVALID;
    d->capture = true;
    return MTGL_OK;
}

MTGL_API int mtgl_diff_add_view(mtgl_diff *df, mtgl_dev *d)
{
    if (!df || !d) return MTGL_ERR_INVALID;
    if (!diff_capturable(d->last, df->tex_w, df->tex_h)) return MTGL_ERR_UNSUPPORTED;
    if (!df->views.empty() && (d->w != df->w || d->h != df->h)) return MTGL_ERR_INVALID;
    int rc = dev_join(d);
    if (rc) return rc;
    if ((rc = confirm(d))) return rc;
    Exec &x = d->ex;
    const LastDraw &l = d->last;
    DiffCapture c;
    c.ntri = l.ntri;
    const size_t npx = (size_t)d->w * d->h;
    do {
        if ((rc = map_err(hipMalloc(&c.gb.ptr, npx * sizeof(GPx))))) break;
        if ((rc = map_err(hipMalloc(&c.elems.ptr, std::max<size_t>(3u * (size_t)l.ntri, 1) * 4)))) break;
        if ((rc = map_err(hipMalloc(&c.pos.ptr, 3u * (size_t)df->nvert * 4)))) break;
        if ((rc = map_err(hipMalloc(&c.nrm.ptr, 3u * (size_t)df->nvert * 4)))) break;
        if ((rc = map_err(hipMalloc(&c.vs.ptr, sizeof(mtgl_vertex_state))))) break;
        if ((rc = map_err(hipMemcpyAsync(c.vs.ptr, &l.vs, sizeof(mtgl_vertex_state), hipMemcpyHostToDevice,
                                         x.stream))))
            break;
        if ((rc = map_err(hipMemsetAsync(df->counter.ptr, 0, 4, x.stream)))) break;
        hipLaunchKernelGGL(k_gbuffer, dim3(blocks_for(npx, 256)), dim3(256), 0, x.stream, d->one.sr,
                           (const RTri *)as<RTri>(x.rt), (const uint32_t *)as<uint32_t>(d->writer),
                           (const uint32_t *)as<uint32_t>(x.tile_gen), l.gen, l.over_base, as<GPx>(c.gb),
                           as<uint32_t>(df->counter));
        const uint64_t n = std::max<uint64_t>(3u * (uint64_t)l.ntri, df->nvert);
        hipLaunchKernelGGL(k_diff_geom, dim3(blocks_for(n, 256)), dim3(256), 0, x.stream, d->one.vp, l.ntri,
                           df->nvert, as<uint32_t>(c.elems), as<float>(c.pos), as<float>(c.nrm));
        if ((rc = map_err(hipGetLastError()))) break;
        uint32_t clipped = 0;
        if ((rc = map_err(hipMemcpyAsync(&clipped, df->counter.ptr, 4, hipMemcpyDeviceToHost, x.stream)))) break;
        if ((rc = map_err(hipStreamSynchronize(x.stream)))) break;
        if (clipped) rc = MTGL_ERR_UNSUPPORTED;   // a triangle crossed a frustum plane
    } while (0);
    if (rc) {
        arena_free(c.gb);
        arena_free(c.elems);
        arena_free(c.pos);
        arena_free(c.nrm);
        arena_free(c.vs);
        return rc;
    }
    c.gb.cap = npx * sizeof(GPx);
    df->w = d->w;
    df->h = d->h;
    df->views.push_back(c);
    df->table_ok = false;
    return MTGL_OK;
}

MTGL_API int mtgl_diff_info(mtgl_diff *df, mtgl_diff_info_t *out)
{
    if (!df || !out) return MTGL_ERR_INVALID;
    out->views = (uint32_t)df->views.size();
    out->width = df->w;
    out->height = df->h;
    out->vertices = df->nvert;
    out->tex_w = df->tex_w;
    out->tex_h = df->tex_h;
    out->work_floats = (uint64_t)df->views.size() * df->nvert * 16u;
    return MTGL_OK;
}

MTGL_API int mtgl_diff_gbuffer(mtgl_diff *df, uint32_t view, mtgl_gbuffer_px *host_out)
{
    if (!df || !host_out || view >= df->views.size()) return MTGL_ERR_INVALID;
    HIPC(hipMemcpy(host_out, df->views[view].gb.ptr, (size_t)df->w * df->h * sizeof(GPx), hipMemcpyDeviceToHost));
    return MTGL_OK;
}

MTGL_API int mtgl_diff_forward(mtgl_diff *df, const float *vcol, const float *light_rgb, const float *tex,
                               float *images, void *stream)
{
    if (!df || !vcol || !light_rgb || !tex || !images || df->views.empty()) return MTGL_ERR_INVALID;
    hipStream_t st = (hipStream_t)stream;
    int rc = diff_table(df, st);
    if (rc) return rc;
    const uint32_t nv = (uint32_t)df->views.size(), npx = (uint32_t)df->w * (uint32_t)df->h;
    hipLaunchKernelGGL(k_diff_lit, dim3(blocks_for(df->nvert, 256), nv), dim3(256), 0, st,
                       (const DiffView *)df->table.ptr, df->nvert, vcol, light_rgb, as<float4>(df->lit),
                       (float4 *)nullptr);
    hipLaunchKernelGGL(k_diff_shade, dim3(blocks_for(npx, 256), nv), dim3(256), 0, st,
                       (const DiffView *)df->table.ptr, df->nvert, (const float4 *)as<float4>(df->lit),
                       (const float4 *)tex, (float4 *)images);
    HIPC(hipGetLastError());
    return MTGL_OK;
}

MTGL_API int mtgl_diff_backward(mtgl_diff *df, const float *vcol, const float *light_rgb, const float *tex,
                                const float *grad_images, float *g_vcol, float *g_light_rgb, float *g_tex,
                                float *work, void *stream)
{
    if (!df || !vcol || !light_rgb || !tex || !grad_images || !g_vcol || !g_light_rgb || !g_tex || !work ||
        df->views.empty())
        return MTGL_ERR_INVALID;
    hipStream_t st = (hipStream_t)stream;
    int rc = diff_table(df, st);
    if (rc) return rc;
    const uint32_t nv = (uint32_t)df->views.size(), npx = (uint32_t)df->w * (uint32_t)df->h;
    const DiffView *tab = (const DiffView *)df->table.ptr;
    // work: dL/dlit (views * nvert float4), then the unclamped lit sums and
    // the N.L terms of the current parameters (three planes of the same size)
    const size_t plane = (size_t)nv * df->nvert;
    float4 *aux = reinterpret_cast<float4 *>(work) + plane;
    // the lit colours of the current parameters (the forward's, recomputed)
    hipLaunchKernelGGL(k_diff_lit, dim3(blocks_for(df->nvert, 256), nv), dim3(256), 0, st, tab, df->nvert, vcol,
                       light_rgb, as<float4>(df->lit), aux);
    if ((rc = fill_zero(work, plane * 4u, st))) return rc;
    if ((rc = fill_zero(g_tex, (size_t)df->tex_w * df->tex_h * 4u, st))) return rc;
    if ((rc = fill_zero(g_light_rgb, 3u * MTGL_MAX_LIGHTS, st))) return rc;
    hipLaunchKernelGGL(k_diff_shade_bwd, dim3(blocks_for(4ull * npx, 256), nv), dim3(256), 0, st, tab, df->nvert,
                       (const float *)as<float>(df->lit), tex, grad_images, work, g_tex);
    hipLaunchKernelGGL(k_diff_lit_bwd, dim3(blocks_for(df->nvert, DIFF_BWD_E)), dim3(DIFF_BWD_E, DIFF_BWD_G), 0, st,
                       tab, nv, df->nvert, vcol, light_rgb, (const float *)work, (const float4 *)aux, g_vcol,
                       g_light_rgb);
    HIPC(hipGetLastError());
    return MTGL_OK;
}
