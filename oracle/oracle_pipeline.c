/*
 * oracle_pipeline.c - CPU restatement of the MyTinyGL per-draw pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it implements the
 * C-ABI seam of include/mtgl_dev.h on the host CPU by restating, operation for
 * operation, the reference renderer's per-vertex, clipping, set-up, per-fragment
 * and texture-sampling arithmetic.  It is linked (with the host GL layer) into
 * oracle/_build/liboracle_gl.so, which only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load.  The shipped library libmtgl.so never links
 * or loads it.
 *
 * Build strict: -std=c99 -O2 -ffp-contract=off, no fast-math (the reference's own
 * Makefile flags -O3 -march=native -ffast-math change depth values and are not a
 * valid oracle).
 *
 * Pinning (what checks each part):
 *   - whole frames, bit for bit, against oracle/_ref/libmtgl_ref.so (the
 *     reference sources compiled strict by oracle/Makefile) and against the
 *     committed SHA-256 fixtures under tests/golden/ made from it;
 *   - closed forms and invariants in tests/test_oracle_units.py (stencil and
 *     depth truth tables, blend factors, RGBA8 pack/unpack, mip1 box filter,
 *     clipping invariants, perspective vs affine interpolation, fill rule).
 * The libm calls (log2f, expf, powf, cosf) are the host C library's, the same
 * calls the reference makes.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "GL/gl.h"
#include "mtgl_batch.h"
#include "mtgl_dev.h"
#include "mtgl_diff.h"

#define OR_MAX_CLIP 12   /* reference MAX_CLIP_VERTS, src/clipping.h:16 */

typedef struct { float x, y, z, w; } or_v4;
typedef struct { float r, g, b, a; } or_col;

/* post-transform vertex (reference vertex_t, src/graphics.h:438-446) */
typedef struct {
    or_v4 pos;
    or_col color;
    float tc[2];
    float nrm[3];
    float ez;
    float ep[3];
    float en[3];
} or_vtx;

typedef struct {
    int32_t w, h;
    uint32_t *px;
    int32_t m1w, m1h;
    uint32_t *m1;       /* level 1, NULL when the level-0 image is smaller than 2x2 */
} or_tex;

typedef struct {
    uint8_t *data;
    int64_t size;
} or_buf;

/* last colour-writing fragment of a pixel, kept for draws a differentiable
 * capture accepts (stamp = the draw's) */
typedef struct {
    uint32_t stamp;
    int32_t tri;            /* -2: the fragment's triangle crossed a frustum plane */
    float b0, b1, b2, u, v;
} or_gbrec;

struct mtgl_dev {
    int32_t w, h;
    uint32_t *color;
    float *depth;
    uint8_t *stencil;
    or_tex tex[MTGL_MAX_OBJECTS + 1];
    or_buf buf[MTGL_MAX_OBJECTS + 1];
    or_buf *segs;          /* display-list segments by id (index 0 unused) */
    uint32_t nsegs;
    mtgl_dev_stats stats;
    /* differentiable capture of the latest draw */
    int armed;                       /* mtgl_diff_arm was called */
    int last_valid;
    int last_armed;
    mtgl_draw last_dr;
    mtgl_raster_state last_rs;
    mtgl_vertex_state last_vs;
    uint8_t *last_staged;
    uint32_t last_ntri;
    or_gbrec *gbr;
    uint32_t gb_stamp;
    int gb_on;
    int32_t cur_tri, tri_counter;
    int cur_clipped;
};

/* ------------------------------------------------------------------------- */
/* scalar helpers                                                              */
/* ------------------------------------------------------------------------- */

/* float -> int32 the way x86 cvttss2si does it: truncation, and INT_MIN for NaN
 * or out-of-range inputs.  The reference's (int32_t) casts compile to this on the
 * x86-64 hosts it runs on (src/raster.c:61-62, src/textures.c:385, :465, :529). */
static int32_t or_cvtt(float f)
{
    if (f >= -2147483648.0f && f < 2147483648.0f) return (int32_t)f;
    return INT32_MIN;
}

static int32_t or_isub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }

static or_col or_unpack(uint32_t p)
{
    or_col c = { (p & 0xFF) / 255.0f, ((p >> 8) & 0xFF) / 255.0f,
                 ((p >> 16) & 0xFF) / 255.0f, ((p >> 24) & 0xFF) / 255.0f };
    return c;
}

/* reference color_to_rgba32, src/graphics.h:337-348 */
static uint32_t or_pack(or_col c)
{
    float v[4] = { c.r, c.g, c.b, c.a };
    uint32_t out = 0;
    for (int k = 0; k < 4; k++) {
        float x = v[k] < 0.0f ? 0.0f : (v[k] > 1.0f ? 1.0f : v[k]);
        uint32_t byte = (uint32_t)or_cvtt(x * 255.0f) & 0xFFu;
        out |= byte << (8 * k);
    }
    return out;
}

static or_col or_clamp(or_col c)
{
    or_col r = { c.r < 0 ? 0 : (c.r > 1 ? 1 : c.r), c.g < 0 ? 0 : (c.g > 1 ? 1 : c.g),
                 c.b < 0 ? 0 : (c.b > 1 ? 1 : c.b), c.a < 0 ? 0 : (c.a > 1 ? 1 : c.a) };
    return r;
}

static or_col or_add(or_col a, or_col b) { or_col r = { a.r + b.r, a.g + b.g, a.b + b.b, a.a + b.a }; return r; }
static or_col or_mul(or_col a, or_col b) { or_col r = { a.r * b.r, a.g * b.g, a.b * b.b, a.a * b.a }; return r; }
static or_col or_scale(or_col a, float s) { or_col r = { a.r * s, a.g * s, a.b * s, a.a * s }; return r; }

/* reference color_lerp: a*(1-t) + b*t, src/graphics.h:293-295 */
static or_col or_lerp(or_col a, or_col b, float t) { return or_add(or_scale(a, 1.0f - t), or_scale(b, t)); }

static or_col or_from(mtgl_color c) { or_col r = { c.r, c.g, c.b, c.a }; return r; }

static float or_dot3(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* reference vec3_normalize: multiply by 1/len, src/graphics.h:54-60 */
static void or_normalize3(float *v)
{
    float len = sqrtf(or_dot3(v, v));
    if (len > 0.0f) {
        float s = 1.0f / len;
        v[0] = v[0] * s; v[1] = v[1] * s; v[2] = v[2] * s;
    }
}

/* reference mat4_mul_vec4, src/graphics.h:131-138 */
static or_v4 or_xform(const float *m, or_v4 v)
{
    or_v4 r;
    r.x = m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w;
    r.y = m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w;
    r.z = m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w;
    r.w = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w;
    return r;
}

/* reference mat4_normal_matrix, src/graphics.h:229-260 */
static void or_normal_matrix(const float *m, float *r)
{
    float a = m[0], b = m[1], c = m[2];
    float d = m[4], e = m[5], f = m[6];
    float g = m[8], h = m[9], i = m[10];
    float det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
    memset(r, 0, 16 * sizeof(float));
    if (fabsf(det) < 1e-10f) { r[0] = r[5] = r[10] = r[15] = 1.0f; return; }
    float inv = 1.0f / det;
    r[0] = (e * i - f * h) * inv;
    r[1] = (d * i - f * g) * -inv;
    r[2] = (d * h - e * g) * inv;
    r[4] = (b * i - c * h) * -inv;
    r[5] = (a * i - c * g) * inv;
    r[6] = (a * h - b * g) * -inv;
    r[8] = (b * f - c * e) * inv;
    r[9] = (a * f - c * d) * -inv;
    r[10] = (a * e - b * d) * inv;
    r[15] = 1.0f;
}

/* ------------------------------------------------------------------------- */
/* lighting (reference compute_lighting, src/lighting.h:53-142)               */
/* ------------------------------------------------------------------------- */

/* compute_lighting before its final clamp; with dif_rgb the lights' diffuse
 * rgb are taken from it (3 per light) and att * N.L of each light that adds
 * diffuse light is stored in ndl_out (else 0) - the differentiable form */
static or_col or_light_sum(const mtgl_lighting *L, uint32_t flags, const float *eye_pos, const float *eye_normal,
                           const mtgl_material *mat, const float *dif_rgb, float *ndl_out)
{
    or_col result = or_add(or_from(mat->emission), or_mul(or_from(mat->ambient), or_from(L->model_ambient)));
    result.a = mat->diffuse.a;
    float N[3] = { eye_normal[0], eye_normal[1], eye_normal[2] };
    if (flags & MTGL_F_NORMALIZE) or_normalize3(N);
    for (int i = 0; i < MTGL_MAX_LIGHTS; i++) {
        const mtgl_light *lt = &L->lights[i];
        if (ndl_out) ndl_out[i] = 0.0f;
        if (!lt->enabled) continue;
        float Ld[3];
        float att = 1.0f;
        if (lt->position[3] == 0.0f) {
            Ld[0] = lt->position[0]; Ld[1] = lt->position[1]; Ld[2] = lt->position[2];
            or_normalize3(Ld);
        } else {
            float to[3] = { lt->position[0] - eye_pos[0], lt->position[1] - eye_pos[1], lt->position[2] - eye_pos[2] };
            float dist = sqrtf(or_dot3(to, to));
            if (dist < 1e-6f) dist = 1e-6f;
            float s = 1.0f / dist;
            Ld[0] = to[0] * s; Ld[1] = to[1] * s; Ld[2] = to[2] * s;
            float den = lt->constant_attenuation + lt->linear_attenuation * dist +
                        lt->quadratic_attenuation * dist * dist;
            if (den < 1e-6f) den = 1e-6f;
            att = 1.0f / den;
            if (lt->spot_cutoff < 180.0f) {
                float sd[3] = { lt->spot_direction[0], lt->spot_direction[1], lt->spot_direction[2] };
                or_normalize3(sd);
                float cos_angle = -or_dot3(Ld, sd);
                float cos_cutoff = cosf(lt->spot_cutoff * 3.14159265f / 180.0f);
                if (cos_angle < cos_cutoff) att = 0.0f;
                else att *= powf(cos_angle, lt->spot_exponent);
            }
        }
        if (att <= 0.0f) continue;
        result = or_add(result, or_scale(or_mul(or_from(mat->ambient), or_from(lt->ambient)), att));
        float ndl = or_dot3(N, Ld);
        if (ndl > 0.0f) {
            or_col ld = or_from(lt->diffuse);
            if (dif_rgb) { ld.r = dif_rgb[3 * i]; ld.g = dif_rgb[3 * i + 1]; ld.b = dif_rgb[3 * i + 2]; }
            if (ndl_out) ndl_out[i] = att * ndl;
            result = or_add(result, or_scale(or_mul(or_from(mat->diffuse), ld), att * ndl));
            if (mat->shininess > 0.0f) {
                float V[3] = { 0.0f, 0.0f, 1.0f };
                if (L->local_viewer) {
                    V[0] = eye_pos[0] * -1.0f; V[1] = eye_pos[1] * -1.0f; V[2] = eye_pos[2] * -1.0f;
                    or_normalize3(V);
                }
                float H[3] = { Ld[0] + V[0], Ld[1] + V[1], Ld[2] + V[2] };
                or_normalize3(H);
                float ndh = or_dot3(N, H);
                if (ndh > 0.0f) {
                    float spec = powf(ndh, mat->shininess) * att;
                    result = or_add(result, or_scale(or_mul(or_from(mat->specular), or_from(lt->specular)), spec));
                }
            }
        }
    }
    return result;
}

static or_col or_light(const mtgl_lighting *L, uint32_t flags, const float *eye_pos, const float *eye_normal,
                       const mtgl_material *mat)
{
    return or_clamp(or_light_sum(L, flags, eye_pos, eye_normal, mat, NULL, NULL));
}

/* ------------------------------------------------------------------------- */
/* vertex fetch and T&L                                                        */
/* ------------------------------------------------------------------------- */

/* source bytes of a stream: the draw's staged block or a buffer object */
static const uint8_t *or_source(mtgl_dev *dev, const mtgl_draw *d, const mtgl_array *a, int64_t *size)
{
    if (a->buffer == 0) { *size = (int64_t)d->staged_bytes; return (const uint8_t *)d->staged; }
    if (a->buffer & MTGL_SEGMENT_BIT) {
        uint32_t id = a->buffer & ~MTGL_SEGMENT_BIT;
        if (id >= dev->nsegs || !dev->segs[id].data) { *size = 0; return NULL; }
        *size = dev->segs[id].size;
        return dev->segs[id].data;
    }
    if (a->buffer > MTGL_MAX_OBJECTS || !dev->buf[a->buffer].data) { *size = 0; return NULL; }
    *size = dev->buf[a->buffer].size;
    return dev->buf[a->buffer].data;
}

/* reference get_array_element, src/gl_api.c:1758-1797; an element outside its
 * source takes the defaults (the reference reads out of bounds there). */
static void or_fetch(mtgl_dev *dev, const mtgl_draw *d, const mtgl_array *a, int32_t index, float *out, int n)
{
    int64_t size;
    const uint8_t *src = or_source(dev, d, a, &size);
    int64_t elem = (int64_t)a->size * (a->type == GL_FLOAT ? 4 : 1);
    int64_t at = a->offset + (int64_t)index * a->stride;
    if (!src || index < 0 || a->stride <= 0 || at < 0 || at + elem > size) {
        for (int i = 0; i < n; i++) out[i] = i < 3 ? 0.0f : 1.0f;
        return;
    }
    const uint8_t *p = src + at;
    for (int i = 0; i < a->size && i < n; i++) {
        if (a->type == GL_FLOAT) { float v; memcpy(&v, p + 4 * i, 4); out[i] = v; }
        else out[i] = p[i] / 255.0f;
    }
    for (int i = a->size; i < n; i++) out[i] = i == 3 ? 1.0f : 0.0f;
}

static int32_t or_index(mtgl_dev *dev, const mtgl_draw *d, uint32_t i)
{
    if (d->index_type == 0) return (int32_t)((uint32_t)d->first + i);
    int64_t size;
    const uint8_t *src = or_source(dev, d, &d->indices, &size);
    int64_t isz = d->index_type == GL_UNSIGNED_INT ? 4 : (d->index_type == GL_UNSIGNED_SHORT ? 2 : 1);
    int64_t at = d->indices.offset + (int64_t)i * isz;
    if (!src || at < 0 || at + isz > size) return -1;
    if (isz == 4) { uint32_t v; memcpy(&v, src + at, 4); return (int32_t)v; }
    if (isz == 2) { uint16_t v; memcpy(&v, src + at, 2); return (int32_t)v; }
    return (int32_t)src[at];
}

/* glColor4f's sanitising (reference src/gl_api.c:708-722) */
static float or_color_component(float v, float nan_value)
{
    if (isnan(v) || isinf(v)) v = nan_value;
    return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
}

/* reference emit_vertex, src/gl_api.c:263-348 */
static void or_emit_vertex(const mtgl_vertex_state *vs, const float *obj, const float *color,
                           const float *texcoord, const float *normal, or_vtx *out)
{
    or_v4 v = { obj[0], obj[1], obj[2], 1.0f };
    or_v4 eye = or_xform(vs->modelview, v);
    out->ez = -eye.z;
    out->ep[0] = eye.x; out->ep[1] = eye.y; out->ep[2] = eye.z;

    float nm[16];
    or_normal_matrix(vs->modelview, nm);
    or_v4 n4 = or_xform(nm, (or_v4){ normal[0], normal[1], normal[2], 0.0f });
    out->en[0] = n4.x; out->en[1] = n4.y; out->en[2] = n4.z;
    or_normalize3(out->en);

    or_col vc = { color[0], color[1], color[2], color[3] };
    int lit = (vs->flags & MTGL_F_LIGHTING) != 0;
    if (lit && vs->shade_model != GL_PHONG) {
        mtgl_material front = vs->lighting.front;
        GLenum face = vs->color_material_face, mode = vs->color_material_mode;
        if ((vs->flags & MTGL_F_COLOR_MATERIAL) && (face == GL_FRONT || face == GL_FRONT_AND_BACK)) {
            or_col k = or_clamp(vc);
            mtgl_color kk = { k.r, k.g, k.b, k.a };
            if (mode == GL_AMBIENT || mode == GL_AMBIENT_AND_DIFFUSE) front.ambient = kk;
            if (mode == GL_DIFFUSE || mode == GL_AMBIENT_AND_DIFFUSE) front.diffuse = kk;
            if (mode == GL_SPECULAR) front.specular = kk;
            if (mode == GL_EMISSION) front.emission = kk;
        }
        vc = or_light(&vs->lighting, vs->flags, out->ep, out->en, &front);
    }
    out->color = vc;

    out->pos = or_xform(vs->projection, or_xform(vs->modelview, v));

    or_v4 t4 = or_xform(vs->texture, (or_v4){ texcoord[0], texcoord[1], 0.0f, 1.0f });
    if (t4.w != 0.0f && t4.w != 1.0f) { out->tc[0] = t4.x / t4.w; out->tc[1] = t4.y / t4.w; }
    else { out->tc[0] = t4.x; out->tc[1] = t4.y; }
    out->nrm[0] = normal[0]; out->nrm[1] = normal[1]; out->nrm[2] = normal[2];
}

/* ------------------------------------------------------------------------- */
/* clipping (reference src/clipping.h:27-126, vertex_lerp src/graphics.h:465) */
/* ------------------------------------------------------------------------- */

static float or_plane(const or_v4 *p, int plane)
{
    switch (plane) {
    case 0: return p->z + p->w;   /* near */
    case 1: return p->w - p->z;   /* far */
    case 2: return p->x + p->w;   /* left */
    case 3: return p->w - p->x;   /* right */
    case 4: return p->y + p->w;   /* bottom */
    default: return p->w - p->y;  /* top */
    }
}

static float or_lerpf(float a, float b, float t) { return a + t * (b - a); }

static or_vtx or_vtx_lerp(const or_vtx *a, const or_vtx *b, float t)
{
    or_vtx v;
    v.pos.x = or_lerpf(a->pos.x, b->pos.x, t);
    v.pos.y = or_lerpf(a->pos.y, b->pos.y, t);
    v.pos.z = or_lerpf(a->pos.z, b->pos.z, t);
    v.pos.w = or_lerpf(a->pos.w, b->pos.w, t);
    v.color = or_lerp(a->color, b->color, t);
    for (int k = 0; k < 2; k++) v.tc[k] = or_lerpf(a->tc[k], b->tc[k], t);
    for (int k = 0; k < 3; k++) v.nrm[k] = or_lerpf(a->nrm[k], b->nrm[k], t);
    v.ez = or_lerpf(a->ez, b->ez, t);
    for (int k = 0; k < 3; k++) v.ep[k] = or_lerpf(a->ep[k], b->ep[k], t);
    for (int k = 0; k < 3; k++) v.en[k] = or_lerpf(a->en[k], b->en[k], t);
    return v;
}

static void or_snap(or_vtx *v, int plane)
{
    switch (plane) {
    case 0: v->pos.z = -v->pos.w; break;
    case 1: v->pos.z = v->pos.w; break;
    case 2: v->pos.x = -v->pos.w; break;
    case 3: v->pos.x = v->pos.w; break;
    case 4: v->pos.y = -v->pos.w; break;
    default: v->pos.y = v->pos.w; break;
    }
}

static int or_clip_plane(const or_vtx *in, int n, or_vtx *out, int plane)
{
    if (n == 0) return 0;
    int m = 0;
    const or_vtx *prev = &in[n - 1];
    float pd = or_plane(&prev->pos, plane);
    for (int i = 0; i < n; i++) {
        const or_vtx *cur = &in[i];
        float cd = or_plane(&cur->pos, plane);
        if (pd >= 0) {
            if (cd >= 0) {
                if (m < OR_MAX_CLIP) out[m++] = *cur;
            } else {
                float den = pd - cd;
                if (fabsf(den) > 1e-10f && m < OR_MAX_CLIP) {
                    out[m] = or_vtx_lerp(prev, cur, pd / den);
                    or_snap(&out[m], plane);
                    m++;
                }
            }
        } else if (cd >= 0) {
            float den = pd - cd;
            if (fabsf(den) > 1e-10f && m < OR_MAX_CLIP) {
                out[m] = or_vtx_lerp(prev, cur, pd / den);
                or_snap(&out[m], plane);
                m++;
            }
            if (m < OR_MAX_CLIP) out[m++] = *cur;
        }
        prev = cur;
        pd = cd;
    }
    return m;
}

static int or_clip_triangle(const or_vtx *tri, or_vtx *out)
{
    or_vtx a[OR_MAX_CLIP], b[OR_MAX_CLIP];
    int n = or_clip_plane(tri, 3, a, 0);
    if (!n) return 0;
    n = or_clip_plane(a, n, b, 1);
    if (!n) return 0;
    n = or_clip_plane(b, n, a, 2);
    if (!n) return 0;
    n = or_clip_plane(a, n, b, 3);
    if (!n) return 0;
    n = or_clip_plane(b, n, a, 4);
    if (!n) return 0;
    return or_clip_plane(a, n, out, 5);
}

/* ------------------------------------------------------------------------- */
/* textures (reference src/textures.c:272-557)                                */
/* ------------------------------------------------------------------------- */

static uint32_t or_texel(const uint32_t *px, int32_t w, int32_t h, uint32_t wrap_s, uint32_t wrap_t,
                         int32_t x, int32_t y)
{
    if (wrap_s == GL_REPEAT) x = ((x % w) + w) % w;
    else { if (x < 0) x = 0; else if (x >= w) x = w - 1; }
    if (wrap_t == GL_REPEAT) y = ((y % h) + h) % h;
    else { if (y < 0) y = 0; else if (y >= h) y = h - 1; }
    return px[(size_t)y * (size_t)w + (size_t)x];
}

static uint32_t or_bilinear(uint32_t c00, uint32_t c10, uint32_t c01, uint32_t c11, float fx, float fy)
{
    or_col top = or_lerp(or_unpack(c00), or_unpack(c10), fx);
    or_col bot = or_lerp(or_unpack(c01), or_unpack(c11), fx);
    return or_pack(or_lerp(top, bot, fy));
}

/* one level, GL_LINEAR (bilinear, wrapped taps) or nearest (clamped index) */
static uint32_t or_sample_level(const uint32_t *px, int32_t w, int32_t h, uint32_t wrap_s, uint32_t wrap_t,
                                float u, float v, int linear, int *taps)
{
    float tx = u * w - 0.5f;
    float ty = v * h - 0.5f;
    if (linear) {
        int32_t x0 = or_cvtt(floorf(tx));
        int32_t y0 = or_cvtt(floorf(ty));
        int32_t x1 = x0 + 1, y1 = y0 + 1;
        float fx = tx - x0;
        float fy = ty - y0;
        *taps += 4;
        return or_bilinear(or_texel(px, w, h, wrap_s, wrap_t, x0, y0), or_texel(px, w, h, wrap_s, wrap_t, x1, y0),
                           or_texel(px, w, h, wrap_s, wrap_t, x0, y1), or_texel(px, w, h, wrap_s, wrap_t, x1, y1),
                           fx, fy);
    }
    int32_t x = or_cvtt(floorf(tx + 0.5f));
    int32_t y = or_cvtt(floorf(ty + 0.5f));
    if (x < 0) x = 0;
    if (x >= w) x = w - 1;
    if (y < 0) y = 0;
    if (y >= h) y = h - 1;
    *taps += 1;
    return px[(size_t)y * (size_t)w + (size_t)x];
}

/* reference texture_sample_mip1, src/textures.c:413-451: white without a level 1 */
static uint32_t or_sample_mip1(const or_tex *t, const mtgl_raster_state *rs, float u, float v, uint32_t filter, int *taps)
{
    if (!t->m1) return 0xFFFFFFFFu;
    int linear = filter == GL_LINEAR || filter == GL_LINEAR_MIPMAP_NEAREST || filter == GL_LINEAR_MIPMAP_LINEAR;
    return or_sample_level(t->m1, t->m1w, t->m1h, rs->wrap_s, rs->wrap_t, u, v, linear, taps);
}

/* reference texture_sample_lod, src/textures.c:457-557 */
static uint32_t or_sample(const or_tex *t, const mtgl_raster_state *rs, float u, float v, float lod, int *taps)
{
    if (rs->wrap_s == GL_REPEAT) { u = u - (float)or_cvtt(u); if (u < 0) u += 1.0f; }
    else { if (u < 0.0f) u = 0.0f; if (u > 1.0f) u = 1.0f; }
    if (rs->wrap_t == GL_REPEAT) { v = v - (float)or_cvtt(v); if (v < 0) v += 1.0f; }
    else { if (v < 0.0f) v = 0.0f; if (v > 1.0f) v = 1.0f; }

    uint32_t filter = lod > 0.0f ? rs->min_filter : rs->mag_filter;
    if (filter == GL_NEAREST_MIPMAP_NEAREST || filter == GL_LINEAR_MIPMAP_NEAREST) {
        if (lod >= 0.5f) return or_sample_mip1(t, rs, u, v, filter, taps);
        filter = filter == GL_NEAREST_MIPMAP_NEAREST ? GL_NEAREST : GL_LINEAR;
    } else if (filter == GL_NEAREST_MIPMAP_LINEAR || filter == GL_LINEAR_MIPMAP_LINEAR) {
        if (lod > 0.0f) {
            float cl = lod > 1.0f ? 1.0f : lod;
            int base_linear = filter == GL_LINEAR_MIPMAP_LINEAR;
            uint32_t c0 = or_sample_level(t->px, t->w, t->h, rs->wrap_s, rs->wrap_t, u, v, base_linear, taps);
            uint32_t c1 = or_sample_mip1(t, rs, u, v, filter, taps);
            return or_pack(or_lerp(or_unpack(c0), or_unpack(c1), cl));
        }
        filter = filter == GL_NEAREST_MIPMAP_LINEAR ? GL_NEAREST : GL_LINEAR;
    }
    return or_sample_level(t->px, t->w, t->h, rs->wrap_s, rs->wrap_t, u, v, filter == GL_LINEAR, taps);
}

/* ------------------------------------------------------------------------- */
/* per-fragment tests and operations (reference src/raster.c:20-45, :343-448) */
/* ------------------------------------------------------------------------- */

static int or_compare(uint32_t func, float a, float b)
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

static int or_stencil_test(uint32_t func, int32_t ref, uint32_t mask, uint8_t s)
{
    int32_t r = (int32_t)((uint32_t)ref & mask);
    int32_t v = (int32_t)((uint32_t)s & mask);
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

static uint8_t or_stencil_op(uint32_t op, uint8_t s, int32_t ref)
{
    switch (op) {
    case GL_KEEP: return s;
    case GL_ZERO: return 0;
    case GL_REPLACE: return (uint8_t)(ref & 0xFF);
    case GL_INCR: return s < 255 ? (uint8_t)(s + 1) : 255;
    case GL_INCR_WRAP: return (uint8_t)(s + 1);
    case GL_DECR: return s > 0 ? (uint8_t)(s - 1) : 0;
    case GL_DECR_WRAP: return (uint8_t)(s - 1);
    case GL_INVERT: return (uint8_t)~s;
    default: return s;
    }
}

static or_col or_blend_factor(uint32_t f, or_col s, or_col d)
{
    or_col r;
    switch (f) {
    case GL_ZERO: r = (or_col){0, 0, 0, 0}; break;
    case GL_ONE: r = (or_col){1, 1, 1, 1}; break;
    case GL_SRC_COLOR: r = s; break;
    case GL_ONE_MINUS_SRC_COLOR: r = (or_col){1 - s.r, 1 - s.g, 1 - s.b, 1 - s.a}; break;
    case GL_DST_COLOR: r = d; break;
    case GL_ONE_MINUS_DST_COLOR: r = (or_col){1 - d.r, 1 - d.g, 1 - d.b, 1 - d.a}; break;
    case GL_SRC_ALPHA: r = (or_col){s.a, s.a, s.a, s.a}; break;
    case GL_ONE_MINUS_SRC_ALPHA: r = (or_col){1 - s.a, 1 - s.a, 1 - s.a, 1 - s.a}; break;
    case GL_DST_ALPHA: r = (or_col){d.a, d.a, d.a, d.a}; break;
    case GL_ONE_MINUS_DST_ALPHA: r = (or_col){1 - d.a, 1 - d.a, 1 - d.a, 1 - d.a}; break;
    case GL_SRC_ALPHA_SATURATE: {
        float k = s.a < (1 - d.a) ? s.a : (1 - d.a);
        r = (or_col){k, k, k, 1};
        break;
    }
    default: r = (or_col){1, 1, 1, 1}; break;
    }
    return r;
}

static int or_in_fb(const mtgl_dev *dev, int32_t x, int32_t y)
{
    return x >= 0 && x < dev->w && y >= 0 && y < dev->h;
}

static uint8_t or_get_stencil(mtgl_dev *dev, int32_t x, int32_t y)
{
    return or_in_fb(dev, x, y) ? dev->stencil[(size_t)y * dev->w + x] : 0;
}

static void or_put_stencil_masked(mtgl_dev *dev, int32_t x, int32_t y, uint8_t nv, uint32_t wm)
{
    uint8_t old = or_get_stencil(dev, x, y);
    uint8_t m8 = (uint8_t)(wm & 0xFF);
    if (or_in_fb(dev, x, y)) dev->stencil[(size_t)y * dev->w + x] = (uint8_t)((old & ~m8) | (nv & m8));
}

/* reference write_pixel_masked, src/raster.c:20-45 */
static void or_write_color(mtgl_dev *dev, uint32_t mask, int32_t x, int32_t y, or_col c)
{
    if (!or_in_fb(dev, x, y)) return;
    uint32_t *p = &dev->color[(size_t)y * dev->w + x];
    if (mask == 0xF) { *p = or_pack(c); return; }
    if (mask == 0) return;
    or_col d = or_unpack(*p);
    if (mask & 1) d.r = c.r;
    if (mask & 2) d.g = c.g;
    if (mask & 4) d.b = c.b;
    if (mask & 8) d.a = c.a;
    *p = or_pack(d);
}

/* ------------------------------------------------------------------------- */
/* triangle rasterisation (reference rasterize_triangle_smooth,               */
/* src/raster.c:451-725)                                                       */
/* ------------------------------------------------------------------------- */

typedef struct {
    int32_t x, y;
    float z, winv;
    or_col c;
    float uv[2];
    float ez;
    float ep[3], en[3];
} or_rv;

static float or_edge(float ax, float ay, float bx, float by, float px, float py)
{
    return (px - ax) * (by - ay) - (py - ay) * (bx - ax);
}

static void or_raster(mtgl_dev *dev, const mtgl_raster_state *rs, const or_rv *v0, const or_rv *v1,
                      const or_rv *v2, int back)
{
    int32_t minX = v0->x < v1->x ? (v0->x < v2->x ? v0->x : v2->x) : (v1->x < v2->x ? v1->x : v2->x);
    int32_t minY = v0->y < v1->y ? (v0->y < v2->y ? v0->y : v2->y) : (v1->y < v2->y ? v1->y : v2->y);
    int32_t maxX = v0->x > v1->x ? (v0->x > v2->x ? v0->x : v2->x) : (v1->x > v2->x ? v1->x : v2->x);
    int32_t maxY = v0->y > v1->y ? (v0->y > v2->y ? v0->y : v2->y) : (v1->y > v2->y ? v1->y : v2->y);
    const int32_t vx = rs->viewport[0], vy = rs->viewport[1], vw = rs->viewport[2], vh = rs->viewport[3];
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
    if (minX > maxX || minY > maxY) return;

    const float fx0 = (float)v0->x, fy0 = (float)v0->y, fx1 = (float)v1->x, fy1 = (float)v1->y;
    const float fx2 = (float)v2->x, fy2 = (float)v2->y;
    float area = or_edge(fx0, fy0, fx1, fy1, fx2, fy2);
    if (fabsf(area) < 0.5f) return;
    float inv_area = 1.0f / area;
    dev->stats.triangles_raster++;

    const int depth_on = (rs->flags & MTGL_F_DEPTH_TEST) != 0;
    const int stencil_on = (rs->flags & MTGL_F_STENCIL_TEST) != 0;
    const int persp = rs->perspective_hint != GL_FASTEST;
    const or_tex *tex = (rs->texture && rs->texture <= MTGL_MAX_OBJECTS && dev->tex[rs->texture].px)
                            ? &dev->tex[rs->texture] : NULL;

    float u0w = v0->uv[0] * v0->winv, v0w = v0->uv[1] * v0->winv;
    float u1w = v1->uv[0] * v1->winv, v1w = v1->uv[1] * v1->winv;
    float u2w = v2->uv[0] * v2->winv, v2w = v2->uv[1] * v2->winv;

    float lod = 0.0f;
    if (tex) {
        float screen_area = fabsf(area) * 0.5f;
        float du1 = (v1->uv[0] - v0->uv[0]) * tex->w;
        float dv1 = (v1->uv[1] - v0->uv[1]) * tex->h;
        float du2 = (v2->uv[0] - v0->uv[0]) * tex->w;
        float dv2 = (v2->uv[1] - v0->uv[1]) * tex->h;
        float texel_area = fabsf(du1 * dv2 - du2 * dv1) * 0.5f;
        if (screen_area > 0.0f) {
            float tpp = texel_area / screen_area;
            if (tpp > 0.0f) {
                lod = log2f(tpp) * 0.5f;
                if (lod < 0.0f) lod = 0.0f;
            }
        }
    }

    const uint32_t cmask = rs->color_mask & 0xF;
    const int blend_on = (rs->flags & MTGL_F_BLEND) != 0;
    for (int32_t y = minY; y <= maxY; y++) {
        for (int32_t x = minX; x <= maxX; x++) {
            const float px = (float)x, py = (float)y;
            float e0 = or_edge(fx1, fy1, fx2, fy2, px, py);
            float e1 = or_edge(fx2, fy2, fx0, fy0, px, py);
            float e2 = or_edge(fx0, fy0, fx1, fy1, px, py);
            if (!((area > 0 && e0 >= 0 && e1 >= 0 && e2 >= 0) || (area < 0 && e0 <= 0 && e1 <= 0 && e2 <= 0)))
                continue;
            /* a pixel outside the framebuffer (viewport larger than it) reads the
             * defaults and every write to it is dropped (src/framebuffer.h:92-134):
             * it has no observable effect and is not counted */
            if (!or_in_fb(dev, x, y)) continue;
            dev->stats.fragments++;
            float b0 = e0 * inv_area, b1 = e1 * inv_area, b2 = e2 * inv_area;
            float z = b0 * v0->z + b1 * v1->z + b2 * v2->z;
            float depth = (float)((double)((z + 1.0f) * 0.5f) * (rs->depth_far - rs->depth_near) + rs->depth_near);

            uint8_t sv = 0;
            if (stencil_on) {
                dev->stats.frag_bytes += 2;
                sv = or_get_stencil(dev, x, y);
                if (!or_stencil_test(rs->stencil_func, rs->stencil_ref, rs->stencil_mask, sv)) {
                    or_put_stencil_masked(dev, x, y, or_stencil_op(rs->stencil_fail, sv, rs->stencil_ref),
                                          rs->stencil_writemask);
                    continue;
                }
            }
            if (depth_on) {
                dev->stats.frag_bytes += 4;
                float stored = or_in_fb(dev, x, y) ? dev->depth[(size_t)y * dev->w + x] : 1.0f;
                if (!or_compare(rs->depth_func, depth, stored)) {
                    if (stencil_on)
                        or_put_stencil_masked(dev, x, y, or_stencil_op(rs->stencil_zfail, sv, rs->stencil_ref),
                                              rs->stencil_writemask);
                    continue;
                }
            }
            if (stencil_on)
                or_put_stencil_masked(dev, x, y, or_stencil_op(rs->stencil_zpass, sv, rs->stencil_ref),
                                      rs->stencil_writemask);

            or_col c;
            if (rs->shade_model == GL_FLAT) {
                c = v2->c;
            } else {
                c.r = v0->c.r * b0 + v1->c.r * b1 + v2->c.r * b2;
                c.g = v0->c.g * b0 + v1->c.g * b1 + v2->c.g * b2;
                c.b = v0->c.b * b0 + v1->c.b * b1 + v2->c.b * b2;
                c.a = v0->c.a * b0 + v1->c.a * b1 + v2->c.a * b2;
            }
            if (rs->flags & MTGL_F_LIGHTING) {
                int phong = rs->shade_model == GL_PHONG;
                int two = back && rs->lighting.two_side;
                if (phong || two) {
                    float ep[3], en[3];
                    for (int k = 0; k < 3; k++) {
                        ep[k] = v0->ep[k] * b0 + v1->ep[k] * b1 + v2->ep[k] * b2;
                        en[k] = v0->en[k] * b0 + v1->en[k] * b1 + v2->en[k] * b2;
                    }
                    const mtgl_material *mat = &rs->lighting.front;
                    if (two) {
                        for (int k = 0; k < 3; k++) en[k] = en[k] * -1.0f;
                        mat = &rs->lighting.back;
                    }
                    c = or_light(&rs->lighting, rs->flags, ep, en, mat);
                }
            }
            float u = 0.0f, v = 0.0f;
            if (tex) {
                if (persp) {
                    float uow = b0 * u0w + b1 * u1w + b2 * u2w;
                    float vow = b0 * v0w + b1 * v1w + b2 * v2w;
                    float oow = b0 * v0->winv + b1 * v1->winv + b2 * v2->winv;
                    float w = 1.0f / oow;
                    u = uow * w;
                    v = vow * w;
                } else {
                    u = b0 * v0->uv[0] + b1 * v1->uv[0] + b2 * v2->uv[0];
                    v = b0 * v0->uv[1] + b1 * v1->uv[1] + b2 * v2->uv[1];
                }
                int taps = 0;
                or_col tc = or_unpack(or_sample(tex, rs, u, v, lod, &taps));
                dev->stats.frag_bytes += 4u * (uint64_t)taps;
                if ((rs->flags & MTGL_F_ALPHA_TEST) && !or_compare(rs->alpha_func, tc.a, rs->alpha_ref)) continue;
                switch (rs->tex_env_mode) {
                case GL_REPLACE: c = tc; break;
                case GL_DECAL:
                    c = (or_col){ c.r + (tc.r - c.r) * tc.a, c.g + (tc.g - c.g) * tc.a, c.b + (tc.b - c.b) * tc.a, c.a };
                    break;
                case GL_BLEND: {
                    or_col e = or_from(rs->tex_env_color);
                    c = (or_col){ c.r * (1.0f - tc.r) + e.r * tc.r, c.g * (1.0f - tc.g) + e.g * tc.g,
                                  c.b * (1.0f - tc.b) + e.b * tc.b, c.a * tc.a };
                    break;
                }
                case GL_ADD: c = (or_col){ c.r + tc.r, c.g + tc.g, c.b + tc.b, c.a * tc.a }; break;
                default: c = or_mul(c, tc); break;
                }
            }
            if (rs->flags & MTGL_F_FOG) {
                float fc = b0 * v0->ez + b1 * v1->ez + b2 * v2->ez;
                float f;
                switch (rs->fog_mode) {
                case GL_LINEAR:
                    f = rs->fog_end != rs->fog_start ? (rs->fog_end - fc) / (rs->fog_end - rs->fog_start) : 1.0f;
                    break;
                case GL_EXP: f = expf(-rs->fog_density * fc); break;
                case GL_EXP2: { float d = rs->fog_density * fc; f = expf(-d * d); break; }
                default: f = 1.0f; break;
                }
                if (f < 0.0f) f = 0.0f;
                if (f > 1.0f) f = 1.0f;
                /* reference color_lerp_rgb(fog_color, c, f): rgb towards the fragment,
                 * alpha taken from the fog colour (src/graphics.h:298-305) */
                or_col fg = or_from(rs->fog_color);
                c = (or_col){ fg.r + (c.r - fg.r) * f, fg.g + (c.g - fg.g) * f, fg.b + (c.b - fg.b) * f, fg.a };
            }
            if (depth_on && rs->depth_mask) {
                dev->stats.frag_bytes += 4;
                if (or_in_fb(dev, x, y)) dev->depth[(size_t)y * dev->w + x] = depth;
            }
            if (blend_on) {
                or_col dst = or_unpack(or_in_fb(dev, x, y) ? dev->color[(size_t)y * dev->w + x] : 0u);
                or_col sf = or_blend_factor(rs->blend_src, c, dst);
                or_col df = or_blend_factor(rs->blend_dst, c, dst);
                c = or_clamp(or_add(or_mul(c, sf), or_mul(dst, df)));
            }
            c = or_clamp(c);
            if (cmask) dev->stats.frag_bytes += (blend_on || cmask != 0xF) ? 8 : 4;
            or_write_color(dev, cmask, x, y, c);
            if (dev->gb_on) {
                or_gbrec *g = &dev->gbr[(size_t)y * dev->w + x];
                g->stamp = dev->gb_stamp;
                g->tri = dev->cur_clipped ? -2 : dev->cur_tri;
                g->b0 = b0; g->b1 = b1; g->b2 = b2;
                g->u = u; g->v = v;
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* points and lines (reference src/raster.c:65-296, :776-898, :1019-1196)      */
/* ------------------------------------------------------------------------- */

/* the per-pixel tail of write_line_pixel (src/raster.c:66-104), also that of
 * flush_points (:1123-1155) and draw_point_at_screen (:777-838), whose extra
 * alpha test and fog depend on the point only: bounds, scissor (raw scissor
 * y against top-down rows, as there), depth test, blend, depth write, masked
 * colour write.  No stencil. */
static void or_prim_pixel(mtgl_dev *dev, const mtgl_raster_state *rs, int32_t px, int32_t py, float depth, or_col c)
{
    if (!or_in_fb(dev, px, py)) return;
    if (rs->flags & MTGL_F_SCISSOR_TEST) {
        const int32_t sx = rs->scissor[0], sy = rs->scissor[1], sw = rs->scissor[2], sh = rs->scissor[3];
        if (px < sx || px >= sx + sw || py < sy || py >= sy + sh) return;
    }
    const size_t i = (size_t)py * dev->w + px;
    const int depth_on = (rs->flags & MTGL_F_DEPTH_TEST) != 0;
    if (depth_on && !or_compare(rs->depth_func, depth, dev->depth[i])) return;
    if (rs->flags & MTGL_F_BLEND) {
        or_col dst = or_unpack(dev->color[i]);
        or_col sf = or_blend_factor(rs->blend_src, c, dst), df = or_blend_factor(rs->blend_dst, c, dst);
        c = or_clamp(or_add(or_mul(c, sf), or_mul(dst, df)));
    }
    if (depth_on && rs->depth_mask) dev->depth[i] = depth;
    or_write_color(dev, rs->color_mask, px, py, c);
}

/* (z + 1) * 0.5 * (far - near) + near in double, stored as float */
static float or_prim_depth(const mtgl_raster_state *rs, float z)
{
    return (float)((double)((z + 1.0f) * 0.5f) * (rs->depth_far - rs->depth_near) + rs->depth_near);
}

/* fog of lines and points: the coordinate is the negated eye_z (the
 * triangles use eye_z itself), colour rgb towards the fragment, alpha of the
 * fog colour (color_lerp_rgb) */
static or_col or_prim_fog(const mtgl_raster_state *rs, or_col c, float fog_coord)
{
    float f = 1.0f;
    switch (rs->fog_mode) {
    case GL_LINEAR:
        if (rs->fog_end != rs->fog_start) f = (rs->fog_end - fog_coord) / (rs->fog_end - rs->fog_start);
        break;
    case GL_EXP: f = expf(-rs->fog_density * fog_coord); break;
    case GL_EXP2: { float d = rs->fog_density * fog_coord; f = expf(-d * d); break; }
    default: break;
    }
    if (f < 0.0f) f = 0.0f;
    if (f > 1.0f) f = 1.0f;
    or_col fg = or_from(rs->fog_color);
    or_col r = { fg.r + (c.r - fg.r) * f, fg.g + (c.g - fg.g) * f, fg.b + (c.b - fg.b) * f, fg.a };
    return r;
}

static const or_tex *or_prim_tex(const mtgl_dev *dev, const mtgl_raster_state *rs)
{
    return (rs->texture && rs->texture <= MTGL_MAX_OBJECTS && dev->tex[rs->texture].px) ? &dev->tex[rs->texture] : NULL;
}

/* reference draw_line_full (src/raster.c:107-246).  Where the alpha test
 * rejects a pixel the reference's loop `continue`s without stepping and never
 * ends; here the pixel is skipped and the line goes on. */
static void or_line(mtgl_dev *dev, const mtgl_raster_state *rs, int32_t x0, int32_t y0, float z0, int32_t x1,
                    int32_t y1, float z1, or_col c0, or_col c1, float ez0, float ez1, const float *uv0,
                    const float *uv1)
{
    int32_t dx = or_isub(x1, x0), dy = or_isub(y1, y0);
    int32_t adx = dx < 0 ? -dx : dx, ady = dy < 0 ? -dy : dy;
    int32_t sx = dx < 0 ? -1 : 1, sy = dy < 0 ? -1 : 1;
    int32_t err = adx - ady;
    int32_t total = adx > ady ? adx : ady;
    if (total == 0) total = 1;
    int32_t step = 0;
    int lw = (int)(rs->line_width + 0.5f);
    if (lw < 1) lw = 1;
    int half = lw / 2;
    int ex = adx > ady ? 0 : 1, ey = adx > ady ? 1 : 0;
    const or_tex *tex = (rs->flags & MTGL_F_TEXTURE_2D) ? or_prim_tex(dev, rs) : NULL;
    int32_t cx = x0, cy = y0;
    for (;;) {
        dev->stats.fragments += (uint64_t)lw;   /* pixels the loop visits */
        float t = (float)step / (float)total;
        float z = z0 + t * (z1 - z0);
        float depth = or_prim_depth(rs, z);
        or_col c = or_lerp(c0, c1, t);
        int keep = 1;
        if (tex) {
            float u = uv0[0] + t * (uv1[0] - uv0[0]);
            float v = uv0[1] + t * (uv1[1] - uv0[1]);
            int taps = 0;
            or_col tc = or_unpack(or_sample(tex, rs, u, v, 0.0f, &taps));
            if ((rs->flags & MTGL_F_ALPHA_TEST) && !or_compare(rs->alpha_func, tc.a, rs->alpha_ref)) keep = 0;
            else c = or_mul(c, tc);
        } else if ((rs->flags & MTGL_F_ALPHA_TEST) && !or_compare(rs->alpha_func, c.a, rs->alpha_ref)) {
            keep = 0;
        }
        if (keep) {
            if (rs->flags & MTGL_F_FOG) c = or_prim_fog(rs, c, -(ez0 + t * (ez1 - ez0)));
            if (lw == 1) {
                or_prim_pixel(dev, rs, cx, cy, depth, c);
            } else {
                for (int w = -half; w < lw - half; w++)
                    or_prim_pixel(dev, rs, cx + w * ex, cy + w * ey, depth, c);
            }
        }
        if (cx == x1 && cy == y1) break;
        int32_t e2 = err * 2;
        if (e2 > -ady) { err -= ady; cx += sx; }
        if (e2 < adx) { err += adx; cy += sy; }
        step++;
    }
}

static int or_outcode(const or_v4 *v)
{
    int code = 0;
    if (v->x < -v->w) code |= 1; else if (v->x > v->w) code |= 2;
    if (v->y < -v->w) code |= 4; else if (v->y > v->w) code |= 8;
    if (v->z < -v->w) code |= 16; else if (v->z > v->w) code |= 32;
    return code;
}

/* reference clip_line (src/clipping.h:152-230): Cohen-Sutherland on the
 * outcodes, planes tried left, right, bottom, top, near, far */
static int or_clip_line(or_vtx *v0, or_vtx *v1)
{
    int c0 = or_outcode(&v0->pos), c1 = or_outcode(&v1->pos);
    for (;;) {
        if (!(c0 | c1)) return 1;
        if (c0 & c1) return 0;
        int out = c0 ? c0 : c1;
        int plane = out & 1 ? 2 : out & 2 ? 3 : out & 4 ? 4 : out & 8 ? 5 : out & 16 ? 0 : 1;
        float d0 = or_plane(&v0->pos, plane), d1 = or_plane(&v1->pos, plane);
        float denom = d0 - d1;
        if (fabsf(denom) < 1e-10f) return 0;
        float t = d0 / denom;
        or_vtx cl = or_vtx_lerp(v0, v1, t);
        or_snap(&cl, plane);
        if (out == c0) { *v0 = cl; c0 = or_outcode(&v0->pos); }
        else { *v1 = cl; c1 = or_outcode(&v1->pos); }
    }
}

/* reference draw_line_segment (src/raster.c:249-288) */
static void or_line_segment(mtgl_dev *dev, const mtgl_raster_state *rs, const or_vtx *a, const or_vtx *b)
{
    or_vtx v0 = *a, v1 = *b;
    if (!or_clip_line(&v0, &v1)) return;
    float z0 = 0.0f, z1 = 0.0f;
    if (fabsf(v0.pos.w) >= 1e-6f) {
        float iw = 1.0f / v0.pos.w;
        v0.pos.x *= iw; v0.pos.y *= iw; z0 = v0.pos.z * iw;
    } else {
        v0.pos.x = 0.0f; v0.pos.y = 0.0f;
    }
    if (fabsf(v1.pos.w) >= 1e-6f) {
        float iw = 1.0f / v1.pos.w;
        v1.pos.x *= iw; v1.pos.y *= iw; z1 = v1.pos.z * iw;
    } else {
        v1.pos.x = 0.0f; v1.pos.y = 0.0f;
    }
    const int32_t vx = rs->viewport[0], vy = rs->viewport[1], vw = rs->viewport[2], vh = rs->viewport[3];
    int32_t x0 = or_cvtt((v0.pos.x + 1.0f) * 0.5f * (float)vw + (float)vx);
    int32_t y0 = or_cvtt((1.0f - v0.pos.y) * 0.5f * (float)vh + (float)vy);
    int32_t x1 = or_cvtt((v1.pos.x + 1.0f) * 0.5f * (float)vw + (float)vx);
    int32_t y1 = or_cvtt((1.0f - v1.pos.y) * 0.5f * (float)vh + (float)vy);
    or_line(dev, rs, x0, y0, z0, x1, y1, z1, v0.color, v1.color, v0.ez, v1.ez, v0.tc, v1.tc);
}

/* reference flush_points (src/raster.c:1019-1158) */
static void or_points(mtgl_dev *dev, const mtgl_raster_state *rs, const or_vtx *vt, uint32_t n)
{
    int ps = (int)(rs->point_size + 0.5f);
    if (ps < 1) ps = 1;
    int half = ps / 2;
    const or_tex *tex = (rs->flags & MTGL_F_TEXTURE_2D) ? or_prim_tex(dev, rs) : NULL;
    const int32_t vx = rs->viewport[0], vy = rs->viewport[1], vw = rs->viewport[2], vh = rs->viewport[3];
    for (uint32_t i = 0; i < n; i++) {
        or_v4 p = vt[i].pos;
        if (p.x < -p.w || p.x > p.w || p.y < -p.w || p.y > p.w || p.z < -p.w || p.z > p.w || p.w <= 0.0f) continue;
        dev->stats.fragments += (uint64_t)ps * (uint64_t)ps;
        float nx = p.x / p.w, ny = p.y / p.w, nz = p.z / p.w;
        int32_t cx = or_cvtt((nx + 1.0f) * 0.5f * (float)vw + (float)vx);
        int32_t cy = or_cvtt((1.0f - ny) * 0.5f * (float)vh + (float)vy);
        float depth = or_prim_depth(rs, nz);
        or_col c = vt[i].color;
        if (tex) {
            int taps = 0;
            or_col tc = or_unpack(or_sample(tex, rs, vt[i].tc[0], vt[i].tc[1], 0.0f, &taps));
            if ((rs->flags & MTGL_F_ALPHA_TEST) && !or_compare(rs->alpha_func, tc.a, rs->alpha_ref)) continue;
            c = or_mul(c, tc);
        } else if ((rs->flags & MTGL_F_ALPHA_TEST) && !or_compare(rs->alpha_func, c.a, rs->alpha_ref)) {
            continue;
        }
        if (rs->flags & MTGL_F_FOG) c = or_prim_fog(rs, c, -vt[i].ez);
        int32_t xs = cx - half, ys = cy - half;
        for (int32_t py = ys; py < ys + ps; py++)
            for (int32_t px = xs; px < xs + ps; px++) or_prim_pixel(dev, rs, px, py, depth, c);
    }
}

/* colours of a fan triangle's corners for glPolygonMode LINE / POINT: the
 * vertex colours, or per-vertex lighting with the front material under
 * GL_PHONG (reference src/raster.c:852-858, :880-886) */
static or_col or_corner_color(const mtgl_raster_state *rs, const or_vtx *v)
{
    if ((rs->flags & MTGL_F_LIGHTING) && rs->shade_model == GL_PHONG)
        return or_light(&rs->lighting, rs->flags, v->ep, v->en, &rs->lighting.front);
    return v->color;
}

/* reference draw_triangle_wireframe / draw_triangle_points (src/raster.c:841-898):
 * corners already divided (x, y, z = NDC), screen positions as for filling */
static void or_tri_prims(mtgl_dev *dev, const mtgl_raster_state *rs, const or_vtx *const *src, const int32_t *X,
                         const int32_t *Y, uint32_t mode)
{
    or_col col[3];
    for (int k = 0; k < 3; k++) col[k] = or_corner_color(rs, src[k]);
    if (mode == GL_LINE) {
        for (int k = 0; k < 3; k++) {
            int m = (k + 1) % 3;
            or_line(dev, rs, X[k], Y[k], src[k]->pos.z, X[m], Y[m], src[m]->pos.z, col[k], col[m], src[k]->ez,
                    src[m]->ez, src[k]->tc, src[m]->tc);
        }
        return;
    }
    for (int k = 0; k < 3; k++) {
        or_col c = col[k];
        dev->stats.fragments += 1;
        if ((rs->flags & MTGL_F_ALPHA_TEST) && !or_compare(rs->alpha_func, c.a, rs->alpha_ref)) continue;
        if (rs->flags & MTGL_F_FOG) c = or_prim_fog(rs, c, -src[k]->ez);
        or_prim_pixel(dev, rs, X[k], Y[k], or_prim_depth(rs, src[k]->pos.z), c);
    }
}

/* ------------------------------------------------------------------------- */
/* per-triangle: clip, divide, fan, cull (reference src/raster.c:729-774,     */
/* :901-958)                                                                   */
/* ------------------------------------------------------------------------- */

static void or_render_triangle(mtgl_dev *dev, const mtgl_raster_state *rs, const or_vtx *a, const or_vtx *b,
                               const or_vtx *c, int *unsupported)
{
    or_vtx tri[3] = { *a, *b, *c };
    or_vtx cl[OR_MAX_CLIP];
    dev->stats.triangles_in++;
    /* for a differentiable capture: the triangle's index in the draw, and
     * whether any vertex lies outside a frustum plane */
    dev->cur_tri = dev->tri_counter++;
    dev->cur_clipped = 0;
    for (int k = 0; k < 3; k++)
        for (int pl = 0; pl < 6; pl++)
            if (!(or_plane(&tri[k].pos, pl) >= 0.0f)) dev->cur_clipped = 1;
    int n = or_clip_triangle(tri, cl);
    if (n < 3) return;
    for (int j = 0; j < n; j++) {
        or_v4 *p = &cl[j].pos;
        if (fabsf(p->w) < 1e-6f) { p->x = 0; p->y = 0; p->z = 0; p->w = 1; continue; }
        float iw = 1.0f / p->w;
        p->x *= iw; p->y *= iw; p->z *= iw; p->w = iw;
    }
    const int32_t vx = rs->viewport[0], vy = rs->viewport[1], vw = rs->viewport[2], vh = rs->viewport[3];
    for (int j = 1; j + 1 < n; j++) {
        const or_vtx *src[3] = { &cl[0], &cl[j], &cl[j + 1] };
        or_rv rv[3];
        for (int k = 0; k < 3; k++) {
            const or_vtx *s = src[k];
            rv[k].x = or_cvtt((s->pos.x + 1.0f) * 0.5f * (float)vw + (float)vx);
            rv[k].y = or_cvtt((1.0f - s->pos.y) * 0.5f * (float)vh + (float)vy);
            rv[k].z = s->pos.z;
            rv[k].winv = s->pos.w;
            rv[k].c = s->color;
            rv[k].uv[0] = s->tc[0]; rv[k].uv[1] = s->tc[1];
            rv[k].ez = s->ez;
            memcpy(rv[k].ep, s->ep, sizeof(rv[k].ep));
            memcpy(rv[k].en, s->en, sizeof(rv[k].en));
        }
        float sa = (float)or_isub(rv[1].x, rv[0].x) * (float)or_isub(rv[2].y, rv[0].y) -
                   (float)or_isub(rv[2].x, rv[0].x) * (float)or_isub(rv[1].y, rv[0].y);
        if (rs->flags & MTGL_F_CULL_FACE) {
            int front = rs->front_face == GL_CCW ? (sa < 0) : (sa > 0);
            int cull = rs->cull_face_mode == GL_FRONT ? front : (rs->cull_face_mode == GL_BACK ? !front : 1);
            if (cull) continue;
        }
        int back = rs->front_face == GL_CCW ? (sa >= 0) : (sa < 0);
        uint32_t pm = back ? rs->polygon_mode_back : rs->polygon_mode_front;
        if (pm == GL_POINT || pm == GL_LINE) {
            const int32_t X[3] = { rv[0].x, rv[1].x, rv[2].x }, Y[3] = { rv[0].y, rv[1].y, rv[2].y };
            or_tri_prims(dev, rs, src, X, Y, pm);
            continue;
        }
        or_raster(dev, rs, &rv[0], &rv[1], &rv[2], back);
    }
}

/* ------------------------------------------------------------------------- */
/* C-ABI                                                                       */
/* ------------------------------------------------------------------------- */

int mtgl_dev_create(int32_t width, int32_t height, mtgl_dev **out)
{
    if (!out || width <= 0 || height <= 0 || width > MTGL_MAX_FB_DIM || height > MTGL_MAX_FB_DIM)
        return MTGL_ERR_INVALID;
    mtgl_dev *d = (mtgl_dev *)calloc(1, sizeof(mtgl_dev));
    if (!d) return MTGL_ERR_OUT_OF_MEMORY;
    size_t n = (size_t)width * (size_t)height;
    d->w = width;
    d->h = height;
    d->color = (uint32_t *)calloc(n, sizeof(uint32_t));
    d->depth = (float *)calloc(n, sizeof(float));
    d->stencil = (uint8_t *)calloc(n, 1);
    if (!d->color || !d->depth || !d->stencil) { mtgl_dev_destroy(d); return MTGL_ERR_OUT_OF_MEMORY; }
    *out = d;
    return MTGL_OK;
}

void mtgl_dev_destroy(mtgl_dev *d)
{
    if (!d) return;
    free(d->gbr);
    free(d->last_staged);
    for (int i = 0; i <= MTGL_MAX_OBJECTS; i++) {
        free(d->tex[i].px);
        free(d->tex[i].m1);
        free(d->buf[i].data);
    }
    for (uint32_t i = 0; i < d->nsegs; i++) free(d->segs[i].data);
    free(d->segs);
    free(d->color);
    free(d->depth);
    free(d->stencil);
    free(d);
}

/* Failure injection for the host layer's error tests (not part of the seam's
 * header): the next `count` calls of the seam that can fail on a device
 * (clear, draw, glDrawPixels, glReadPixels, texture and buffer uploads,
 * finish) return `rc` without doing anything. */
static int or_fail_rc, or_fail_count;

__attribute__((visibility("default"))) void oracle_inject_failure(int rc, int count)
{
    or_fail_rc = rc;
    or_fail_count = count;
}

static int or_injected(void)
{
    if (or_fail_count <= 0) return MTGL_OK;
    or_fail_count--;
    return or_fail_rc;
}

int mtgl_dev_finish(mtgl_dev *d)
{
    if (!d) return MTGL_ERR_INVALID;
    return or_injected();
}

static int or_record_clear(mtgl_dev *d, uint32_t buffers, int32_t x0, int32_t y0, int32_t x1, int32_t y1,
                           uint32_t color, float depth, uint8_t stencil);
static void or_diff_prepare(mtgl_dev *d, const mtgl_draw *dr);
static int or_record_draw(mtgl_dev *d, const mtgl_draw *dr);
static __thread mtgl_batch *or_recording;

int mtgl_dev_clear(mtgl_dev *d, uint32_t buffers, int32_t x0, int32_t y0, int32_t x1, int32_t y1,
                   uint32_t color, float depth, uint8_t stencil)
{
    if (!d) return MTGL_ERR_INVALID;
    {
        const int frc = or_injected();
        if (frc) return frc;
    }
    if (or_recording) return or_record_clear(d, buffers, x0, y0, x1, y1, color, depth, stencil);
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    if (x1 > d->w) x1 = d->w;
    if (y1 > d->h) y1 = d->h;
    for (int32_t y = y0; y < y1; y++) {
        for (int32_t x = x0; x < x1; x++) {
            size_t i = (size_t)y * d->w + x;
            if (buffers & MTGL_BUF_COLOR) d->color[i] = color;
            if (buffers & MTGL_BUF_DEPTH) d->depth[i] = depth;
            if (buffers & MTGL_BUF_STENCIL) d->stencil[i] = stencil;
        }
    }
    return MTGL_OK;
}

static int or_plane_ptr(mtgl_dev *d, uint32_t which, uint8_t **base, size_t *esz)
{
    if (which == MTGL_BUF_COLOR) { *base = (uint8_t *)d->color; *esz = 4; return 1; }
    if (which == MTGL_BUF_DEPTH) { *base = (uint8_t *)d->depth; *esz = 4; return 1; }
    if (which == MTGL_BUF_STENCIL) { *base = d->stencil; *esz = 1; return 1; }
    return 0;
}

int mtgl_dev_read(mtgl_dev *d, uint32_t which, int32_t x, int32_t y, int32_t w, int32_t h, void *dst)
{
    uint8_t *base; size_t esz;
    if (!d || !dst || !or_plane_ptr(d, which, &base, &esz)) return MTGL_ERR_INVALID;
    if (x < 0 || y < 0 || w < 0 || h < 0 || x + w > d->w || y + h > d->h) return MTGL_ERR_INVALID;
    for (int32_t r = 0; r < h; r++)
        memcpy((uint8_t *)dst + (size_t)r * w * esz, base + ((size_t)(y + r) * d->w + x) * esz, (size_t)w * esz);
    return MTGL_OK;
}

/* glReadPixels, reference src/gl_api.c:1180-1229: window row r is framebuffer
 * row h - 1 - (y + r); rows outside are zero bytes, pixels beside the
 * framebuffer in rows that exist are (0, 0, 0, 255); RGBA or RGB bytes. */
int mtgl_dev_read_pixels(mtgl_dev *d, int32_t x, int32_t y, int32_t width, int32_t height, uint32_t format,
                         void *dst)
{
    if (!d || !dst) return MTGL_ERR_INVALID;
    {
        const int frc = or_injected();
        if (frc) return frc;
    }
    const int bpp = format == GL_RGBA ? 4 : (format == GL_RGB ? 3 : 0);
    if (width <= 0 || height <= 0 || bpp == 0) return MTGL_OK;
    uint8_t *out = (uint8_t *)dst;
    for (int32_t row = 0; row < height; row++) {
        const int64_t fb_y = (int64_t)d->h - 1 - ((int64_t)y + row);
        uint8_t *o = out + (size_t)row * (size_t)width * (size_t)bpp;
        if (fb_y < 0 || fb_y >= d->h) {
            memset(o, 0, (size_t)width * (size_t)bpp);
            continue;
        }
        for (int32_t col = 0; col < width; col++) {
            const int64_t sx = (int64_t)x + col;
            uint8_t r = 0, g = 0, b = 0, a = 255;
            if (sx >= 0 && sx < d->w) {
                const uint32_t p = d->color[fb_y * d->w + sx];
                r = p & 0xFF;
                g = (p >> 8) & 0xFF;
                b = (p >> 16) & 0xFF;
                a = (p >> 24) & 0xFF;
            }
            o[(size_t)col * bpp + 0] = r;
            o[(size_t)col * bpp + 1] = g;
            o[(size_t)col * bpp + 2] = b;
            if (bpp == 4) o[(size_t)col * bpp + 3] = a;
        }
    }
    return MTGL_OK;
}

/* the colour plane, top row first, rows `pitch` bytes apart (the reference's
 * SDL glue uploads framebuffer.color with pitch width * 4,
 * include/mytinygl/sdl.h:74-84) */
int mtgl_dev_present(mtgl_dev *d, void *dst, int64_t pitch)
{
    if (!d || !dst || pitch < (int64_t)d->w * 4) return MTGL_ERR_INVALID;
    for (int32_t r = 0; r < d->h; r++)
        memcpy((uint8_t *)dst + (size_t)r * (size_t)pitch, d->color + (size_t)r * d->w, (size_t)d->w * 4);
    return MTGL_OK;
}

int mtgl_dev_write(mtgl_dev *d, uint32_t which, int32_t x, int32_t y, int32_t w, int32_t h, const void *src)
{
    uint8_t *base; size_t esz;
    if (!d || !src || !or_plane_ptr(d, which, &base, &esz)) return MTGL_ERR_INVALID;
    if (x < 0 || y < 0 || w < 0 || h < 0 || x + w > d->w || y + h > d->h) return MTGL_ERR_INVALID;
    for (int32_t r = 0; r < h; r++)
        memcpy(base + ((size_t)(y + r) * d->w + x) * esz, (const uint8_t *)src + (size_t)r * w * esz, (size_t)w * esz);
    return MTGL_OK;
}

/* glDrawPixels, reference src/gl_api.c:1284-1373, with the reference's own
 * compare switch (:1231-1258) and blend factors (:1260-1282) */
static int or_dp_compare(uint32_t func, float a, float b)
{
    switch (func) {
        case GL_NEVER:    return 0;
        case GL_LESS:     return a < b;
        case GL_EQUAL:    return a == b;
        case GL_LEQUAL:   return a <= b;
        case GL_GREATER:  return a > b;
        case GL_NOTEQUAL: return a != b;
        case GL_GEQUAL:   return a >= b;
        default:          return 1;
    }
}

static or_col or_dp_factor(uint32_t f, or_col s, or_col d)
{
    or_col one = { 1.0f, 1.0f, 1.0f, 1.0f };
    switch (f) {
        case GL_ZERO: { or_col z = { 0.0f, 0.0f, 0.0f, 0.0f }; return z; }
        case GL_SRC_COLOR: return s;
        case GL_ONE_MINUS_SRC_COLOR: { or_col c = { 1 - s.r, 1 - s.g, 1 - s.b, 1 - s.a }; return c; }
        case GL_DST_COLOR: return d;
        case GL_ONE_MINUS_DST_COLOR: { or_col c = { 1 - d.r, 1 - d.g, 1 - d.b, 1 - d.a }; return c; }
        case GL_SRC_ALPHA: { or_col c = { s.a, s.a, s.a, s.a }; return c; }
        case GL_ONE_MINUS_SRC_ALPHA: { or_col c = { 1 - s.a, 1 - s.a, 1 - s.a, 1 - s.a }; return c; }
        case GL_DST_ALPHA: { or_col c = { d.a, d.a, d.a, d.a }; return c; }
        case GL_ONE_MINUS_DST_ALPHA: { or_col c = { 1 - d.a, 1 - d.a, 1 - d.a, 1 - d.a }; return c; }
        case GL_SRC_ALPHA_SATURATE: {
            float k = (s.a < (1 - d.a)) ? s.a : (1 - d.a);
            or_col c = { k, k, k, 1.0f };
            return c;
        }
        default: return one;
    }
}

int mtgl_dev_draw_pixels(mtgl_dev *d, const mtgl_pixels *p)
{
    if (!d || !p || !p->pixels || p->width < 0 || p->height < 0) return MTGL_ERR_INVALID;
    if ((uint64_t)p->width * (uint64_t)p->height > 0xFFFFFFFFull) return MTGL_ERR_INVALID;
    if (or_recording) return MTGL_ERR_UNSUPPORTED;
    {
        const int frc = or_injected();
        if (frc) return frc;
    }
    const uint8_t *src = (const uint8_t *)p->pixels;
    int alpha_on = (p->flags & MTGL_F_ALPHA_TEST) != 0;
    int depth_on = (p->flags & MTGL_F_DEPTH_TEST) != 0;
    int blend_on = (p->flags & MTGL_F_BLEND) != 0;
    for (int32_t row = 0; row < p->height; row++) {
        int32_t fb_y = d->h - 1 - (p->y + row);
        if (fb_y < 0 || fb_y >= d->h) continue;
        for (int32_t col = 0; col < p->width; col++) {
            int32_t x = p->x + col;
            if (x < 0 || x >= d->w) continue;
            size_t i = (size_t)row * (size_t)p->width + (size_t)col;
            uint32_t r, g, b, a = 255;
            if (p->format == GL_RGBA) { r = src[4 * i]; g = src[4 * i + 1]; b = src[4 * i + 2]; a = src[4 * i + 3]; }
            else if (p->format == GL_RGB) { r = src[3 * i]; g = src[3 * i + 1]; b = src[3 * i + 2]; }
            else if (p->format == GL_LUMINANCE) { r = g = b = src[i]; }
            else if (p->format == GL_LUMINANCE_ALPHA) { r = g = b = src[2 * i]; a = src[2 * i + 1]; }
            else continue;
            if (alpha_on && !or_dp_compare(p->alpha_func, a / 255.0f, p->alpha_ref)) continue;
            size_t pix = (size_t)fb_y * d->w + x;
            if (depth_on && !or_dp_compare(p->depth_func, 0.0f, d->depth[pix])) continue;
            or_col c = { r / 255.0f, g / 255.0f, b / 255.0f, a / 255.0f };
            if (blend_on) {
                or_col dst = or_unpack(d->color[pix]);
                or_col sf = or_dp_factor(p->blend_src, c, dst), df = or_dp_factor(p->blend_dst, c, dst);
                or_col m = { c.r * sf.r + dst.r * df.r, c.g * sf.g + dst.g * df.g,
                             c.b * sf.b + dst.b * df.b, c.a * sf.a + dst.a * df.a };
                c = or_clamp(m);
            }
            if (depth_on && p->depth_mask) d->depth[pix] = 0.0f;
            d->color[pix] = or_pack(c);
        }
    }
    return MTGL_OK;
}

/* the no-allocation variant: the same pixels, refused when the source image
 * exceeds the fixed staging buffer of the device library */
int mtgl_dev_draw_pixels_static(mtgl_dev *d, const mtgl_pixels *p)
{
    if (!d || !p || !p->pixels || p->width < 0 || p->height < 0) return MTGL_ERR_INVALID;
    uint64_t bpp = 0;
    switch (p->format) {
    case GL_RGBA: bpp = 4; break;
    case GL_RGB: bpp = 3; break;
    case GL_LUMINANCE: bpp = 1; break;
    case GL_LUMINANCE_ALPHA: bpp = 2; break;
    default: return MTGL_OK;
    }
    const uint64_t n = (uint64_t)p->width * (uint64_t)p->height;
    if (n > 0xFFFFFFFFull || n * bpp > (uint64_t)MTGL_STATIC_PIXELS_MAX) return MTGL_ERR_INVALID;
    return mtgl_dev_draw_pixels(d, p);
}

/* level 0 plus the 2x2 box-filtered level 1: (((c00 + c10) + c01) + c11) * 0.25,
 * packed with truncation (reference texture_generate_mip1, src/textures.c:311-354) */
int mtgl_dev_tex_image(mtgl_dev *d, uint32_t id, int32_t w, int32_t h, const uint32_t *texels)
{
    if (!d || id == 0 || id > MTGL_MAX_OBJECTS || w <= 0 || h <= 0 || w > MTGL_MAX_TEX_SIZE ||
        h > MTGL_MAX_TEX_SIZE || !texels)
        return MTGL_ERR_INVALID;
    {
        const int frc = or_injected();
        if (frc) return frc;
    }
    or_tex *t = &d->tex[id];
    size_t n = (size_t)w * (size_t)h;
    uint32_t *px = (uint32_t *)malloc(n * sizeof(uint32_t));
    if (!px) return MTGL_ERR_OUT_OF_MEMORY;
    memcpy(px, texels, n * sizeof(uint32_t));
    uint32_t *m1 = NULL;
    int32_t mw = 0, mh = 0;
    if (w >= 2 && h >= 2) {
        mw = w / 2; mh = h / 2;
        m1 = (uint32_t *)malloc((size_t)mw * (size_t)mh * sizeof(uint32_t));
        if (!m1) { free(px); return MTGL_ERR_OUT_OF_MEMORY; }
        for (int32_t y = 0; y < mh; y++) {
            for (int32_t x = 0; x < mw; x++) {
                size_t s = (size_t)(2 * y) * w + (size_t)(2 * x);
                or_col k = or_unpack(px[s]);
                k = or_add(k, or_unpack(px[s + 1]));
                k = or_add(k, or_unpack(px[s + w]));
                k = or_add(k, or_unpack(px[s + w + 1]));
                m1[(size_t)y * mw + x] = or_pack(or_scale(k, 0.25f));
            }
        }
    }
    free(t->px);
    free(t->m1);
    t->px = px; t->w = w; t->h = h;
    t->m1 = m1; t->m1w = mw; t->m1h = mh;
    return MTGL_OK;
}

int mtgl_dev_tex_free(mtgl_dev *d, uint32_t id)
{
    if (!d || id == 0 || id > MTGL_MAX_OBJECTS) return MTGL_ERR_INVALID;
    free(d->tex[id].px);
    free(d->tex[id].m1);
    memset(&d->tex[id], 0, sizeof(d->tex[id]));
    return MTGL_OK;
}

int mtgl_dev_buffer_data(mtgl_dev *d, uint32_t id, int64_t size, const void *data)
{
    if (!d || id == 0 || id > MTGL_MAX_OBJECTS) return MTGL_ERR_INVALID;
    {
        const int frc = or_injected();
        if (frc) return frc;
    }
    if (size <= 0) { free(d->buf[id].data); d->buf[id].data = NULL; d->buf[id].size = 0; return MTGL_OK; }
    uint8_t *p = (uint8_t *)malloc((size_t)size);
    if (!p) return MTGL_ERR_OUT_OF_MEMORY;
    if (data) memcpy(p, data, (size_t)size);
    else memset(p, 0, (size_t)size);
    free(d->buf[id].data);
    d->buf[id].data = p;
    d->buf[id].size = size;
    return MTGL_OK;
}

int mtgl_dev_buffer_sub_data(mtgl_dev *d, uint32_t id, int64_t offset, int64_t size, const void *data)
{
    if (!d || id == 0 || id > MTGL_MAX_OBJECTS || !d->buf[id].data || !data || offset < 0 || size < 0 ||
        offset + size > d->buf[id].size)
        return MTGL_ERR_INVALID;
    {
        const int frc = or_injected();
        if (frc) return frc;
    }
    memcpy(d->buf[id].data + offset, data, (size_t)size);
    return MTGL_OK;
}

int mtgl_dev_buffer_free(mtgl_dev *d, uint32_t id)
{
    if (!d || id == 0 || id > MTGL_MAX_OBJECTS) return MTGL_ERR_INVALID;
    free(d->buf[id].data);
    d->buf[id].data = NULL;
    d->buf[id].size = 0;
    return MTGL_OK;
}

/* one draw: per-vertex T&L of every emitted vertex, then primitive assembly in
 * the reference's flush order (src/raster.c:961-1017, :1199-1231) */
int mtgl_dev_segment_data(mtgl_dev *d, uint32_t id, int64_t size, const void *data)
{
    if (!d || id == 0 || id > MTGL_MAX_SEGMENT) return MTGL_ERR_INVALID;
    if (size <= 0) return mtgl_dev_segment_free(d, id);
    if (id >= d->nsegs) {
        uint32_t n = d->nsegs ? d->nsegs : 16;
        while (n <= id) n *= 2;
        or_buf *p = (or_buf *)realloc(d->segs, (size_t)n * sizeof(or_buf));
        if (!p) return MTGL_ERR_OUT_OF_MEMORY;
        memset(p + d->nsegs, 0, (size_t)(n - d->nsegs) * sizeof(or_buf));
        d->segs = p;
        d->nsegs = n;
    }
    uint8_t *q = (uint8_t *)malloc((size_t)size);
    if (!q) return MTGL_ERR_OUT_OF_MEMORY;
    if (data) memcpy(q, data, (size_t)size); else memset(q, 0, (size_t)size);
    d->stats.upload_bytes += (uint64_t)size;
    free(d->segs[id].data);
    d->segs[id].data = q;
    d->segs[id].size = size;
    return MTGL_OK;
}

int mtgl_dev_segment_free(mtgl_dev *d, uint32_t id)
{
    if (!d || id == 0 || id > MTGL_MAX_SEGMENT) return MTGL_ERR_INVALID;
    if (id < d->nsegs) {
        free(d->segs[id].data);
        d->segs[id].data = NULL;
        d->segs[id].size = 0;
    }
    return MTGL_OK;
}

int mtgl_dev_draw(mtgl_dev *d, const mtgl_draw *dr)
{
    if (!d || !dr || !dr->raster || !dr->vstates || dr->n_vstates == 0) return MTGL_ERR_INVALID;
    {
        const int frc = or_injected();
        if (frc) return frc;
    }
    const uint32_t n = dr->count;
    switch (dr->mode) {
    case GL_POINTS: case GL_LINES: case GL_LINE_STRIP: case GL_LINE_LOOP:
    case GL_TRIANGLES: case GL_TRIANGLE_STRIP: case GL_TRIANGLE_FAN: case GL_QUADS:
    case GL_QUAD_STRIP: case GL_POLYGON:
        break;
    default:
        return MTGL_ERR_UNSUPPORTED;
    }
    if (or_recording) return or_record_draw(d, dr);
    or_diff_prepare(d, dr);
    d->stats.draws++;
    d->stats.upload_bytes += dr->staged_bytes;
    if (n == 0) return MTGL_OK;
    or_vtx *vt = (or_vtx *)malloc((size_t)n * sizeof(or_vtx));
    if (!vt) return MTGL_ERR_OUT_OF_MEMORY;
    for (uint32_t i = 0; i < n; i++) {
        int32_t idx = or_index(d, dr, i);
        float pos[4], col[4], tc[2], nrm[3];
        or_fetch(d, dr, &dr->position, idx, pos, 4);
        if (dr->color.enabled) {
            or_fetch(d, dr, &dr->color, idx, col, 4);
            for (int k = 0; k < 4; k++) col[k] = or_color_component(col[k], k == 3 ? 1.0f : 0.0f);
        } else {
            memcpy(col, dr->cur_color, sizeof(col));
        }
        if (dr->texcoord.enabled) or_fetch(d, dr, &dr->texcoord, idx, tc, 2);
        else memcpy(tc, dr->cur_texcoord, sizeof(tc));
        if (dr->normal.enabled) or_fetch(d, dr, &dr->normal, idx, nrm, 3);
        else memcpy(nrm, dr->cur_normal, sizeof(nrm));
        if (dr->position.size == 2) pos[2] = 0.0f;
        uint32_t si = 0;
        if (dr->n_vstates > 1 && dr->state_index.enabled) {
            int64_t size;
            const uint8_t *src = or_source(d, dr, &dr->state_index, &size);
            int64_t at = dr->state_index.offset + (int64_t)i * dr->state_index.stride;
            if (src && at >= 0 && at + 4 <= size) memcpy(&si, src + at, 4);
            if (si >= dr->n_vstates) si = dr->n_vstates - 1;
        }
        or_emit_vertex(&dr->vstates[si], pos, col, tc, nrm, &vt[i]);
    }
    const mtgl_raster_state *rs = dr->raster;
    int unsupported = 0;
    switch (dr->mode) {
    case GL_POINTS:
        or_points(d, rs, vt, n);
        break;
    case GL_LINES:
        for (uint32_t i = 0; i + 1 < n; i += 2) or_line_segment(d, rs, &vt[i], &vt[i + 1]);
        break;
    case GL_LINE_STRIP:
    case GL_LINE_LOOP:
        for (uint32_t i = 0; i + 1 < n; i++) or_line_segment(d, rs, &vt[i], &vt[i + 1]);
        if (dr->mode == GL_LINE_LOOP && n >= 2) or_line_segment(d, rs, &vt[n - 1], &vt[0]);
        break;
    case GL_TRIANGLES:
        for (uint32_t i = 0; i + 2 < n; i += 3) or_render_triangle(d, rs, &vt[i], &vt[i + 1], &vt[i + 2], &unsupported);
        break;
    case GL_TRIANGLE_STRIP:
        for (uint32_t i = 0; i + 2 < n; i++) {
            if (i % 2 == 0) or_render_triangle(d, rs, &vt[i], &vt[i + 1], &vt[i + 2], &unsupported);
            else or_render_triangle(d, rs, &vt[i + 1], &vt[i], &vt[i + 2], &unsupported);
        }
        break;
    case GL_TRIANGLE_FAN:
    case GL_POLYGON:
        for (uint32_t i = 1; i + 1 < n; i++) or_render_triangle(d, rs, &vt[0], &vt[i], &vt[i + 1], &unsupported);
        break;
    case GL_QUADS:
        for (uint32_t i = 0; i + 3 < n; i += 4) {
            or_render_triangle(d, rs, &vt[i], &vt[i + 1], &vt[i + 2], &unsupported);
            or_render_triangle(d, rs, &vt[i], &vt[i + 2], &vt[i + 3], &unsupported);
        }
        break;
    case GL_QUAD_STRIP:
        for (uint32_t i = 0; i + 3 < n; i += 2) {
            or_render_triangle(d, rs, &vt[i], &vt[i + 1], &vt[i + 3], &unsupported);
            or_render_triangle(d, rs, &vt[i], &vt[i + 3], &vt[i + 2], &unsupported);
        }
        break;
    default:
        break;
    }
    free(vt);
    return unsupported ? MTGL_ERR_UNSUPPORTED : MTGL_OK;
}

int mtgl_dev_set_timing(mtgl_dev *d, int every) { return d && every >= 0 ? MTGL_OK : MTGL_ERR_INVALID; }

int mtgl_dev_stats_get(mtgl_dev *d, mtgl_dev_stats *out)
{
    if (!d || !out) return MTGL_ERR_INVALID;
    *out = d->stats;
    return MTGL_OK;
}

int mtgl_dev_stats_reset(mtgl_dev *d)
{
    if (!d) return MTGL_ERR_INVALID;
    memset(&d->stats, 0, sizeof(d->stats));
    return MTGL_OK;
}

const char *mtgl_dev_backend(void) { return "oracle:cpu (strict C restatement)"; }

/* ------------------------------------------------------------------------- */
/* batches: the semantic restatement of include/mtgl_batch.h.  Recording copies */
/* each clear and draw; a launch executes them in recorded order, which keeps   */
/* every context's order (the only order the batch promises).                   */
/* ------------------------------------------------------------------------- */

typedef struct {
    mtgl_dev *dev;
    int is_clear;
    uint32_t buffers;
    int32_t x0, y0, x1, y1;
    uint32_t color;
    float depth;
    uint8_t stencil;
    mtgl_draw draw;
    mtgl_raster_state raster;
    mtgl_vertex_state *vstates;
    uint8_t *staged;
} or_cmd;

struct mtgl_batch {
    or_cmd *cmds;
    size_t n, cap;
    mtgl_dev **devs;
    size_t ndevs;
    int recorded, ready;
    mtgl_batch_info_t info;
    mtgl_dev_stats stats;
};

/* triangles the assembly loops of mtgl_dev_draw above produce for n vertices */
static uint32_t or_tri_count(uint32_t mode, uint32_t n)
{
    switch (mode) {
    case GL_TRIANGLES: return n / 3;
    case GL_TRIANGLE_STRIP: case GL_TRIANGLE_FAN: case GL_POLYGON: return n >= 3 ? n - 2 : 0;
    case GL_QUADS: return (n / 4) * 2;
    case GL_QUAD_STRIP: return n >= 4 ? ((n - 4) / 2 + 1) * 2 : 0;
    default: return 0;
    }
}

static or_cmd *or_batch_push(mtgl_batch *b, mtgl_dev *d)
{
    if (b->n == b->cap) {
        size_t cap = b->cap ? 2 * b->cap : 64;
        or_cmd *c = (or_cmd *)realloc(b->cmds, cap * sizeof(or_cmd));
        if (!c) return NULL;
        b->cmds = c;
        b->cap = cap;
    }
    size_t i;
    for (i = 0; i < b->ndevs; i++)
        if (b->devs[i] == d) break;
    if (i == b->ndevs) {
        mtgl_dev **nd = (mtgl_dev **)realloc(b->devs, (b->ndevs + 1) * sizeof(mtgl_dev *));
        if (!nd) return NULL;
        b->devs = nd;
        b->devs[b->ndevs++] = d;
    }
    or_cmd *c = &b->cmds[b->n++];
    memset(c, 0, sizeof(*c));
    c->dev = d;
    return c;
}

static int or_record_clear(mtgl_dev *d, uint32_t buffers, int32_t x0, int32_t y0, int32_t x1, int32_t y1,
                           uint32_t color, float depth, uint8_t stencil)
{
    or_cmd *c = or_batch_push(or_recording, d);
    if (!c) return MTGL_ERR_OUT_OF_MEMORY;
    c->is_clear = 1;
    c->buffers = buffers;
    c->x0 = x0;
    c->y0 = y0;
    c->x1 = x1;
    c->y1 = y1;
    c->color = color;
    c->depth = depth;
    c->stencil = stencil;
    return MTGL_OK;
}

static int or_record_draw(mtgl_dev *d, const mtgl_draw *dr)
{
    /* a batch replays filled triangles only: points, lines and polygon modes
     * other than GL_FILL are refused while recording */
    switch (dr->mode) {
    case GL_POINTS: case GL_LINES: case GL_LINE_STRIP: case GL_LINE_LOOP: return MTGL_ERR_UNSUPPORTED;
    default: break;
    }
    if (dr->raster->polygon_mode_front != GL_FILL || dr->raster->polygon_mode_back != GL_FILL)
        return MTGL_ERR_UNSUPPORTED;
    or_cmd *c = or_batch_push(or_recording, d);
    if (!c) return MTGL_ERR_OUT_OF_MEMORY;
    c->draw = *dr;
    c->raster = *dr->raster;
    c->vstates = (mtgl_vertex_state *)malloc((size_t)dr->n_vstates * sizeof(mtgl_vertex_state));
    if (!c->vstates) return MTGL_ERR_OUT_OF_MEMORY;
    memcpy(c->vstates, dr->vstates, (size_t)dr->n_vstates * sizeof(mtgl_vertex_state));
    if (dr->staged_bytes) {
        c->staged = (uint8_t *)malloc((size_t)dr->staged_bytes);
        if (!c->staged) return MTGL_ERR_OUT_OF_MEMORY;
        memcpy(c->staged, dr->staged, (size_t)dr->staged_bytes);
    }
    c->draw.raster = NULL;   /* set at launch: the command array may move */
    c->draw.vstates = c->vstates;
    c->draw.staged = c->staged;
    return MTGL_OK;
}

int mtgl_batch_create(mtgl_batch **out)
{
    if (!out) return MTGL_ERR_INVALID;
    *out = (mtgl_batch *)calloc(1, sizeof(mtgl_batch));
    return *out ? MTGL_OK : MTGL_ERR_OUT_OF_MEMORY;
}

void mtgl_batch_destroy(mtgl_batch *b)
{
    if (!b) return;
    if (or_recording == b) or_recording = NULL;
    for (size_t i = 0; i < b->n; i++) {
        free(b->cmds[i].vstates);
        free(b->cmds[i].staged);
    }
    free(b->cmds);
    free(b->devs);
    free(b);
}

int mtgl_batch_begin(mtgl_batch *b)
{
    if (!b || or_recording || b->recorded || b->n) return MTGL_ERR_INVALID;
    or_recording = b;
    return MTGL_OK;
}

int mtgl_batch_end(mtgl_batch *b, int autotune)
{
    (void)autotune;
    if (!b || or_recording != b) return MTGL_ERR_INVALID;
    or_recording = NULL;
    b->recorded = 1;
    b->ready = 1;
    b->info.contexts = (uint32_t)b->ndevs;
    b->info.commands = (uint32_t)b->n;
    /* rounds: the most commands of one context */
    uint32_t rounds = 0, scenes = 0;
    for (size_t i = 0; i < b->ndevs; i++) {
        uint32_t k = 0;
        for (size_t j = 0; j < b->n; j++)
            if (b->cmds[j].dev == b->devs[i]) k++;
        if (k > rounds) rounds = k;
    }
    for (size_t j = 0; j < b->n; j++)
        if (!b->cmds[j].is_clear && or_tri_count(b->cmds[j].draw.mode, b->cmds[j].draw.count)) scenes++;
    b->info.rounds = rounds;
    b->info.scenes = scenes;
    return MTGL_OK;
}

int mtgl_batch_launch(mtgl_batch *b)
{
    if (!b || !b->ready) return MTGL_ERR_INVALID;
    for (size_t j = 0; j < b->n; j++) {
        or_cmd *c = &b->cmds[j];
        if (c->is_clear) {
            mtgl_dev_clear(c->dev, c->buffers, c->x0, c->y0, c->x1, c->y1, c->color, c->depth, c->stencil);
        } else {
            mtgl_dev_stats s0 = c->dev->stats;
            c->draw.raster = &c->raster;   /* the command array may have moved since recording */
            (void)mtgl_dev_draw(c->dev, &c->draw);
            const mtgl_dev_stats *s1 = &c->dev->stats;
            b->stats.fragments += s1->fragments - s0.fragments;
            b->stats.frag_bytes += s1->frag_bytes - s0.frag_bytes;
            b->stats.triangles_in += s1->triangles_in - s0.triangles_in;
            b->stats.triangles_raster += s1->triangles_raster - s0.triangles_raster;
            b->stats.draws += s1->draws - s0.draws;
        }
    }
    b->info.launches++;
    return MTGL_OK;
}

int mtgl_batch_sync(mtgl_batch *b) { return b ? MTGL_OK : MTGL_ERR_INVALID; }

int mtgl_batch_info(mtgl_batch *b, mtgl_batch_info_t *out)
{
    if (!b || !out) return MTGL_ERR_INVALID;
    *out = b->info;
    return MTGL_OK;
}

int mtgl_batch_stats_get(mtgl_batch *b, mtgl_dev_stats *out)
{
    if (!b || !out) return MTGL_ERR_INVALID;
    *out = b->stats;
    return MTGL_OK;
}

int mtgl_batch_stats_reset(mtgl_batch *b)
{
    if (!b) return MTGL_ERR_INVALID;
    memset(&b->stats, 0, sizeof(b->stats));
    return MTGL_OK;
}

/* ------------------------------------------------------------------------- */
/* differentiable shading: the restatement of include/mtgl_diff.h, and the    */
/* CPU reference gradient (adjoint accumulated in double)                     */
/* ------------------------------------------------------------------------- */

/* the capture conditions of include/mtgl_diff.h, from the draw's own state */
static int or_diff_qualifies(const mtgl_dev *d, const mtgl_draw *dr)
{
    const mtgl_raster_state *rs = dr->raster;
    const mtgl_vertex_state *vs = &dr->vstates[0];
    if (dr->mode != GL_TRIANGLES || dr->n_vstates != 1 || !dr->color.enabled) return 0;
    if ((rs->color_mask & 0xFu) != 0xFu) return 0;
    if (rs->flags & (MTGL_F_BLEND | MTGL_F_ALPHA_TEST | MTGL_F_FOG)) return 0;
    if (!rs->texture || rs->texture > MTGL_MAX_OBJECTS || !d->tex[rs->texture].px) return 0;
    if (rs->min_filter != GL_LINEAR || rs->mag_filter != GL_LINEAR || rs->tex_env_mode != GL_MODULATE) return 0;
    if (rs->shade_model == GL_FLAT || rs->shade_model == GL_PHONG || rs->lighting.two_side) return 0;
    if (!(vs->flags & MTGL_F_LIGHTING) || !(vs->flags & MTGL_F_COLOR_MATERIAL) || vs->shade_model == GL_PHONG) return 0;
    if (vs->color_material_face != GL_FRONT && vs->color_material_face != GL_FRONT_AND_BACK) return 0;
    if (vs->color_material_mode != GL_AMBIENT && vs->color_material_mode != GL_DIFFUSE &&
        vs->color_material_mode != GL_AMBIENT_AND_DIFFUSE)
        return 0;
    for (int i = 0; i < MTGL_MAX_LIGHTS; i++)
        if (vs->lighting.lights[i].enabled && vs->lighting.lights[i].position[3] != 0.0f) return 0;
    return 1;
}

/* before a draw: keep a copy of it and turn per-pixel recording on when a
 * capture could take it */
static void or_diff_prepare(mtgl_dev *d, const mtgl_draw *dr)
{
    d->tri_counter = 0;
    d->last_valid = 0;
    d->gb_on = 0;
    if (!or_diff_qualifies(d, dr)) return;
    if (!d->gbr) {
        d->gbr = (or_gbrec *)calloc((size_t)d->w * d->h, sizeof(or_gbrec));
        if (!d->gbr) return;
    }
    free(d->last_staged);
    d->last_staged = NULL;
    if (dr->staged_bytes) {
        d->last_staged = (uint8_t *)malloc((size_t)dr->staged_bytes);
        if (!d->last_staged) return;
        memcpy(d->last_staged, dr->staged, (size_t)dr->staged_bytes);
    }
    d->last_dr = *dr;
    d->last_rs = *dr->raster;
    d->last_vs = dr->vstates[0];
    d->last_dr.raster = &d->last_rs;
    d->last_dr.vstates = &d->last_vs;
    d->last_dr.staged = d->last_staged;
    d->last_ntri = dr->count / 3;
    if (++d->gb_stamp == 0) {
        memset(d->gbr, 0, (size_t)d->w * d->h * sizeof(or_gbrec));
        d->gb_stamp = 1;
    }
    d->gb_on = 1;
    d->last_valid = 1;
    d->last_armed = d->armed;
}

typedef struct {
    mtgl_gbuffer_px *gb;
    uint32_t *elems;
    float *pos, *nrm;
    mtgl_vertex_state vs;
    uint32_t ntri;
} or_dview;

struct mtgl_diff {
    uint32_t nvert;
    int32_t tex_w, tex_h, w, h;
    uint32_t nviews;
    or_dview *views;
};

int mtgl_diff_create(uint32_t vertices, int32_t tex_w, int32_t tex_h, mtgl_diff **out)
{
    if (!out || vertices == 0 || tex_w <= 0 || tex_h <= 0 || tex_w > MTGL_MAX_TEX_SIZE || tex_h > MTGL_MAX_TEX_SIZE)
        return MTGL_ERR_INVALID;
    mtgl_diff *df = (mtgl_diff *)calloc(1, sizeof(mtgl_diff));
    if (!df) return MTGL_ERR_OUT_OF_MEMORY;
    df->nvert = vertices;
    df->tex_w = tex_w;
    df->tex_h = tex_h;
    *out = df;
    return MTGL_OK;
}

void mtgl_diff_destroy(mtgl_diff *df)
{
    if (!df) return;
    for (uint32_t i = 0; i < df->nviews; i++) {
        free(df->views[i].gb);
        free(df->views[i].elems);
        free(df->views[i].pos);
        free(df->views[i].nrm);
    }
    free(df->views);
    free(df);
}

/* mtgl_diff_arm: the draws that follow are captured (the oracle records the
 * writer of every pixel regardless; the flag mirrors the device's refusal of
 * a draw made before arming). */
int mtgl_diff_arm(mtgl_dev *d)
{
    if (!d) return MTGL_ERR_INVALID;
    d->armed = 1;
    return MTGL_OK;
}

int mtgl_diff_add_view(mtgl_diff *df, mtgl_dev *d)
{
    if (!df || !d) return MTGL_ERR_INVALID;
    if (!d->last_valid || !d->last_armed || !d->gbr) return MTGL_ERR_UNSUPPORTED;
    const or_tex *t = &d->tex[d->last_rs.texture];
    if (!t->px || t->w != df->tex_w || t->h != df->tex_h) return MTGL_ERR_UNSUPPORTED;
    if (df->nviews && (d->w != df->w || d->h != df->h)) return MTGL_ERR_INVALID;
    const size_t npx = (size_t)d->w * d->h;
    or_dview v;
    memset(&v, 0, sizeof(v));
    v.ntri = d->last_ntri;
    v.vs = d->last_vs;
    v.gb = (mtgl_gbuffer_px *)calloc(npx, sizeof(mtgl_gbuffer_px));
    v.elems = (uint32_t *)malloc((3 * (size_t)v.ntri + 1) * sizeof(uint32_t));
    v.pos = (float *)malloc(3 * (size_t)df->nvert * sizeof(float));
    v.nrm = (float *)malloc(3 * (size_t)df->nvert * sizeof(float));
    or_dview *nv = (or_dview *)realloc(df->views, (df->nviews + 1) * sizeof(or_dview));
    if (!v.gb || !v.elems || !v.pos || !v.nrm || !nv) {
        free(v.gb); free(v.elems); free(v.pos); free(v.nrm);
        if (nv) df->views = nv;
        return MTGL_ERR_OUT_OF_MEMORY;
    }
    df->views = nv;
    const mtgl_raster_state *rs = &d->last_rs;
    for (size_t p = 0; p < npx; p++) {
        const or_gbrec *r = &d->gbr[p];
        mtgl_gbuffer_px *g = &v.gb[p];
        if (r->stamp != d->gb_stamp) {
            g->tri = -1;
            g->texel[0] = d->color[p];
            continue;
        }
        if (r->tri == -2) {
            free(v.gb); free(v.elems); free(v.pos); free(v.nrm);
            return MTGL_ERR_UNSUPPORTED;
        }
        g->tri = r->tri;
        g->b[0] = r->b0; g->b[1] = r->b1; g->b[2] = r->b2;
        /* the wrap of or_sample and the bilinear taps of or_sample_level */
        float u = r->u, vv = r->v;
        if (rs->wrap_s == GL_REPEAT) { u = u - (float)or_cvtt(u); if (u < 0) u += 1.0f; }
        else { if (u < 0.0f) u = 0.0f; if (u > 1.0f) u = 1.0f; }
        if (rs->wrap_t == GL_REPEAT) { vv = vv - (float)or_cvtt(vv); if (vv < 0) vv += 1.0f; }
        else { if (vv < 0.0f) vv = 0.0f; if (vv > 1.0f) vv = 1.0f; }
        float tx = u * t->w - 0.5f, ty = vv * t->h - 0.5f;
        int32_t x0 = or_cvtt(floorf(tx)), y0 = or_cvtt(floorf(ty));
        g->fx = tx - x0;
        g->fy = ty - y0;
        int32_t xs[2] = { x0, x0 + 1 }, ys[2] = { y0, y0 + 1 };
        for (int k = 0; k < 2; k++) {
            if (rs->wrap_s == GL_REPEAT) xs[k] = ((xs[k] % t->w) + t->w) % t->w;
            else { if (xs[k] < 0) xs[k] = 0; else if (xs[k] >= t->w) xs[k] = t->w - 1; }
            if (rs->wrap_t == GL_REPEAT) ys[k] = ((ys[k] % t->h) + t->h) % t->h;
            else { if (ys[k] < 0) ys[k] = 0; else if (ys[k] >= t->h) ys[k] = t->h - 1; }
        }
        g->texel[0] = (uint32_t)(ys[0] * t->w + xs[0]);
        g->texel[1] = (uint32_t)(ys[0] * t->w + xs[1]);
        g->texel[2] = (uint32_t)(ys[1] * t->w + xs[0]);
        g->texel[3] = (uint32_t)(ys[1] * t->w + xs[1]);
    }
    for (uint32_t i = 0; i < 3 * v.ntri; i++) v.elems[i] = (uint32_t)or_index(d, &d->last_dr, i);
    for (uint32_t e = 0; e < df->nvert; e++) {
        float p4[4], n3[3];
        or_fetch(d, &d->last_dr, &d->last_dr.position, (int32_t)e, p4, 4);
        if (d->last_dr.position.size == 2) p4[2] = 0.0f;
        if (d->last_dr.normal.enabled) or_fetch(d, &d->last_dr, &d->last_dr.normal, (int32_t)e, n3, 3);
        else memcpy(n3, d->last_dr.cur_normal, sizeof(n3));
        memcpy(v.pos + 3 * (size_t)e, p4, 3 * sizeof(float));
        memcpy(v.nrm + 3 * (size_t)e, n3, 3 * sizeof(float));
    }
    df->views[df->nviews++] = v;
    df->w = d->w;
    df->h = d->h;
    return MTGL_OK;
}

int mtgl_diff_info(mtgl_diff *df, mtgl_diff_info_t *out)
{
    if (!df || !out) return MTGL_ERR_INVALID;
    out->views = df->nviews;
    out->width = df->w;
    out->height = df->h;
    out->vertices = df->nvert;
    out->tex_w = df->tex_w;
    out->tex_h = df->tex_h;
    out->work_floats = (uint64_t)df->nviews * df->nvert * 16u;
    return MTGL_OK;
}

int mtgl_diff_gbuffer(mtgl_diff *df, uint32_t view, mtgl_gbuffer_px *out)
{
    if (!df || !out || view >= df->nviews) return MTGL_ERR_INVALID;
    memcpy(out, df->views[view].gb, (size_t)df->w * df->h * sizeof(mtgl_gbuffer_px));
    return MTGL_OK;
}

/* eye-space vertex, sanitised colour and material as or_emit_vertex forms them */
typedef struct {
    float ep[3], en[3];
    or_col vc;
    mtgl_material mat;
} or_dvtx;

static void or_diff_vertex(const or_dview *v, uint32_t e, const float *vcol, or_dvtx *o)
{
    const mtgl_vertex_state *vs = &v->vs;
    const float *pp = v->pos + 3 * (size_t)e, *nn = v->nrm + 3 * (size_t)e;
    or_v4 eye = or_xform(vs->modelview, (or_v4){ pp[0], pp[1], pp[2], 1.0f });
    o->ep[0] = eye.x; o->ep[1] = eye.y; o->ep[2] = eye.z;
    float nm[16];
    or_normal_matrix(vs->modelview, nm);
    or_v4 n4 = or_xform(nm, (or_v4){ nn[0], nn[1], nn[2], 0.0f });
    o->en[0] = n4.x; o->en[1] = n4.y; o->en[2] = n4.z;
    or_normalize3(o->en);
    const float *c = vcol + 4 * (size_t)e;
    o->vc = (or_col){ or_color_component(c[0], 0.0f), or_color_component(c[1], 0.0f), or_color_component(c[2], 0.0f),
                      or_color_component(c[3], 1.0f) };
    o->mat = vs->lighting.front;
    or_col k = or_clamp(o->vc);
    mtgl_color kk = { k.r, k.g, k.b, k.a };
    if (vs->color_material_mode == GL_AMBIENT || vs->color_material_mode == GL_AMBIENT_AND_DIFFUSE) o->mat.ambient = kk;
    if (vs->color_material_mode == GL_DIFFUSE || vs->color_material_mode == GL_AMBIENT_AND_DIFFUSE) o->mat.diffuse = kk;
}

static or_col or_diff_lit(const or_dview *v, uint32_t e, const float *vcol, const float *light_rgb)
{
    or_dvtx q;
    or_diff_vertex(v, e, vcol, &q);
    return or_clamp(or_light_sum(&v->vs.lighting, v->vs.flags, q.ep, q.en, &q.mat, light_rgb, NULL));
}

static or_col or_tex4(const float *tex, uint32_t i)
{
    const float *t = tex + 4 * (size_t)i;
    or_col c = { t[0], t[1], t[2], t[3] };
    return c;
}

static void or_diff_pixel(const mtgl_gbuffer_px *g, const uint32_t *el, const or_col *lit, uint32_t nvert,
                          const float *tex, or_col *c, or_col *tc)
{
    or_col l[3];
    for (int k = 0; k < 3; k++) {
        or_col z = { 0.0f, 0.0f, 0.0f, 0.0f };
        l[k] = el[k] < nvert ? lit[el[k]] : z;
    }
    float b0 = g->b[0], b1 = g->b[1], b2 = g->b[2];
    c->r = l[0].r * b0 + l[1].r * b1 + l[2].r * b2;
    c->g = l[0].g * b0 + l[1].g * b1 + l[2].g * b2;
    c->b = l[0].b * b0 + l[1].b * b1 + l[2].b * b2;
    c->a = l[0].a * b0 + l[1].a * b1 + l[2].a * b2;
    or_col top = or_lerp(or_tex4(tex, g->texel[0]), or_tex4(tex, g->texel[1]), g->fx);
    or_col bot = or_lerp(or_tex4(tex, g->texel[2]), or_tex4(tex, g->texel[3]), g->fx);
    *tc = or_lerp(top, bot, g->fy);
}

int mtgl_diff_forward(mtgl_diff *df, const float *vcol, const float *light_rgb, const float *tex, float *images,
                      void *stream)
{
    (void)stream;
    if (!df || !vcol || !light_rgb || !tex || !images || !df->nviews) return MTGL_ERR_INVALID;
    or_col *lit = (or_col *)malloc((size_t)df->nvert * sizeof(or_col));
    if (!lit) return MTGL_ERR_OUT_OF_MEMORY;
    const size_t npx = (size_t)df->w * df->h;
    for (uint32_t vi = 0; vi < df->nviews; vi++) {
        const or_dview *v = &df->views[vi];
        for (uint32_t e = 0; e < df->nvert; e++) lit[e] = or_diff_lit(v, e, vcol, light_rgb);
        float *img = images + 4 * npx * vi;
        for (size_t p = 0; p < npx; p++) {
            const mtgl_gbuffer_px *g = &v->gb[p];
            or_col o;
            if (g->tri < 0) {
                o = or_unpack(g->texel[0]);
            } else {
                or_col c, tc;
                or_diff_pixel(g, v->elems + 3 * (size_t)g->tri, lit, df->nvert, tex, &c, &tc);
                o = or_clamp(or_mul(c, tc));
            }
            img[4 * p] = o.r; img[4 * p + 1] = o.g; img[4 * p + 2] = o.b; img[4 * p + 3] = o.a;
        }
    }
    free(lit);
    return MTGL_OK;
}

static float or_clamp_grad(float v, float g) { return (v >= 0.0f && v <= 1.0f) ? g : 0.0f; }

int mtgl_diff_backward(mtgl_diff *df, const float *vcol, const float *light_rgb, const float *tex,
                       const float *grad_images, float *g_vcol, float *g_light_rgb, float *g_tex, float *work,
                       void *stream)
{
    (void)stream;
    (void)work;
    if (!df || !vcol || !light_rgb || !tex || !grad_images || !g_vcol || !g_light_rgb || !g_tex || !df->nviews)
        return MTGL_ERR_INVALID;
    const size_t npx = (size_t)df->w * df->h, ntex = (size_t)df->tex_w * df->tex_h;
    or_col *lit = (or_col *)malloc((size_t)df->nvert * sizeof(or_col));
    double *dlit = (double *)malloc((size_t)df->nvert * 4 * sizeof(double));
    double *gt = (double *)calloc(ntex * 4, sizeof(double));
    double *gv = (double *)calloc((size_t)df->nvert * 4, sizeof(double));
    double gl[3 * MTGL_MAX_LIGHTS] = { 0 };
    if (!lit || !dlit || !gt || !gv) {
        free(lit); free(dlit); free(gt); free(gv);
        return MTGL_ERR_OUT_OF_MEMORY;
    }
    for (uint32_t vi = 0; vi < df->nviews; vi++) {
        const or_dview *v = &df->views[vi];
        for (uint32_t e = 0; e < df->nvert; e++) lit[e] = or_diff_lit(v, e, vcol, light_rgb);
        memset(dlit, 0, (size_t)df->nvert * 4 * sizeof(double));
        const float *gimg = grad_images + 4 * npx * vi;
        /* adjoint of the pixel: clamp, modulate, colour weights, bilinear taps */
        for (size_t p = 0; p < npx; p++) {
            const mtgl_gbuffer_px *g = &v->gb[p];
            if (g->tri < 0) continue;
            const uint32_t *el = v->elems + 3 * (size_t)g->tri;
            or_col c, tc;
            or_diff_pixel(g, el, lit, df->nvert, tex, &c, &tc);
            or_col prod = or_mul(c, tc);
            or_col gm = { or_clamp_grad(prod.r, gimg[4 * p]), or_clamp_grad(prod.g, gimg[4 * p + 1]),
                          or_clamp_grad(prod.b, gimg[4 * p + 2]), or_clamp_grad(prod.a, gimg[4 * p + 3]) };
            or_col dc = or_mul(gm, tc), dt = or_mul(gm, c);
            for (int k = 0; k < 3; k++) {
                if (el[k] >= df->nvert) continue;
                double *d = dlit + 4 * (size_t)el[k];
                float b = g->b[k];
                d[0] += dc.r * b; d[1] += dc.g * b; d[2] += dc.b * b; d[3] += dc.a * b;
            }
            float fx = g->fx, fy = g->fy;
            float wt[4] = { (1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy, fx * fy };
            for (int k = 0; k < 4; k++) {
                double *d = gt + 4 * (size_t)g->texel[k];
                d[0] += dt.r * wt[k]; d[1] += dt.g * wt[k]; d[2] += dt.b * wt[k]; d[3] += dt.a * wt[k];
            }
        }
        /* adjoint of the lighting: clamp, material colour = clamp(vertex colour), light diffuse */
        const mtgl_lighting *L = &v->vs.lighting;
        const uint32_t mode = v->vs.color_material_mode;
        const int amb_c = mode == GL_AMBIENT || mode == GL_AMBIENT_AND_DIFFUSE;
        const int dif_c = mode == GL_DIFFUSE || mode == GL_AMBIENT_AND_DIFFUSE;
        for (uint32_t e = 0; e < df->nvert; e++) {
            const double *dl = dlit + 4 * (size_t)e;
            if (dl[0] == 0.0 && dl[1] == 0.0 && dl[2] == 0.0 && dl[3] == 0.0) continue;
            or_dvtx q;
            or_diff_vertex(v, e, vcol, &q);
            float ndl[MTGL_MAX_LIGHTS];
            or_col r = or_light_sum(L, v->vs.flags, q.ep, q.en, &q.mat, light_rgb, ndl);
            double g[4] = { or_clamp_grad(r.r, 1.0f) * dl[0], or_clamp_grad(r.g, 1.0f) * dl[1],
                            or_clamp_grad(r.b, 1.0f) * dl[2], or_clamp_grad(r.a, 1.0f) * dl[3] };
            double dk[4] = { 0.0, 0.0, 0.0, dif_c ? 1.0 : 0.0 };
            if (amb_c) { dk[0] += L->model_ambient.r; dk[1] += L->model_ambient.g; dk[2] += L->model_ambient.b; }
            for (int i = 0; i < MTGL_MAX_LIGHTS; i++) {
                const mtgl_light *lt = &L->lights[i];
                if (!lt->enabled) continue;
                if (amb_c) { dk[0] += lt->ambient.r; dk[1] += lt->ambient.g; dk[2] += lt->ambient.b; dk[3] += lt->ambient.a; }
                if (ndl[i] > 0.0f) {
                    if (dif_c) {
                        dk[0] += (double)light_rgb[3 * i] * ndl[i];
                        dk[1] += (double)light_rgb[3 * i + 1] * ndl[i];
                        dk[2] += (double)light_rgb[3 * i + 2] * ndl[i];
                        dk[3] += (double)lt->diffuse.a * ndl[i];
                    }
                    gl[3 * i] += g[0] * q.mat.diffuse.r * ndl[i];
                    gl[3 * i + 1] += g[1] * q.mat.diffuse.g * ndl[i];
                    gl[3 * i + 2] += g[2] * q.mat.diffuse.b * ndl[i];
                }
            }
            const float vc[4] = { q.vc.r, q.vc.g, q.vc.b, q.vc.a };
            for (int ch = 0; ch < 4; ch++)
                if (vc[ch] >= 0.0f && vc[ch] <= 1.0f) gv[4 * (size_t)e + ch] += g[ch] * dk[ch];
        }
    }
    for (size_t i = 0; i < ntex * 4; i++) g_tex[i] = (float)gt[i];
    for (size_t i = 0; i < (size_t)df->nvert * 4; i++) g_vcol[i] = (float)gv[i];
    for (int i = 0; i < 3 * MTGL_MAX_LIGHTS; i++) g_light_rgb[i] = (float)gl[i];
    free(lit); free(dlit); free(gt); free(gv);
    return MTGL_OK;
}
