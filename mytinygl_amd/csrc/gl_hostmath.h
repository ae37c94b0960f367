/*
 * gl_hostmath.h - matrix and colour arithmetic of the GL host layer.
 *
 * The matrices built here feed the device pipeline bit for bit, so every
 * expression keeps the operand order of the reference's inline math
 * (src/graphics.h) and the file must be compiled without FMA contraction
 * (-ffp-contract=off) and without fast-math.  Matrices are column-major
 * float[16], as in OpenGL.
 */
#ifndef MTGL_GL_HOSTMATH_H
#define MTGL_GL_HOSTMATH_H

#include <math.h>
#include <string.h>
#include "mtgl_dev.h"

static inline void hm_identity(float *m)
{
    memset(m, 0, 16 * sizeof(float));
    m[0] = m[5] = m[10] = m[15] = 1.0f;
}

/* r = a * b; each element is the left-to-right sum over k of a[k][row]*b[col][k]
 * (reference mat4_mul, src/graphics.h:140-152).  r may alias neither input. */
static inline void hm_mul(float *r, const float *a, const float *b)
{
    for (int col = 0; col < 4; col++) {
        for (int row = 0; row < 4; row++) {
            float s = a[row] * b[col * 4];
            s = s + a[4 + row] * b[col * 4 + 1];
            s = s + a[8 + row] * b[col * 4 + 2];
            s = s + a[12 + row] * b[col * 4 + 3];
            r[col * 4 + row] = s;
        }
    }
}

/* m <- m * t, the update rule of every glTranslate/glRotate/... call
 * (reference src/gl_api.c:551-603). */
static inline void hm_post_multiply(float *m, const float *t)
{
    float r[16];
    hm_mul(r, m, t);
    memcpy(m, r, sizeof(r));
}

/* v' = m * v with the row sums ((m0 x + m4 y) + m8 z) + m12 w
 * (reference mat4_mul_vec4, src/graphics.h:131-138). */
static inline void hm_mul_vec4(const float *m, const float *v, float *out)
{
    float r[4];
    for (int row = 0; row < 4; row++) {
        float s = m[row] * v[0];
        s = s + m[4 + row] * v[1];
        s = s + m[8 + row] * v[2];
        s = s + m[12 + row] * v[3];
        r[row] = s;
    }
    out[0] = r[0]; out[1] = r[1]; out[2] = r[2]; out[3] = r[3];
}

static inline void hm_translate(float *m, float x, float y, float z)
{
    hm_identity(m);
    m[12] = x; m[13] = y; m[14] = z;
}

static inline void hm_scale(float *m, float x, float y, float z)
{
    hm_identity(m);
    m[0] = x; m[5] = y; m[10] = z;
}

/* Rotation about (x,y,z) by angle degrees, the glRotate matrix
 * (reference mat4_rotate, src/graphics.h:172-186): the axis is normalised by
 * division, cos/sin come from the C library. */
static inline void hm_rotate(float *m, float angle_deg, float x, float y, float z)
{
    float rad = angle_deg * 3.14159265358979f / 180.0f;
    float c = cosf(rad);
    float s = sinf(rad);
    float len = sqrtf(x * x + y * y + z * z);
    if (len == 0) { hm_identity(m); return; }
    x /= len; y /= len; z /= len;
    float ic = 1 - c;
    m[0] = x * x * ic + c;  m[1] = y * x * ic + z * s; m[2] = x * z * ic - y * s; m[3] = 0;
    m[4] = x * y * ic - z * s; m[5] = y * y * ic + c;  m[6] = y * z * ic + x * s; m[7] = 0;
    m[8] = x * z * ic + y * s; m[9] = y * z * ic - x * s; m[10] = z * z * ic + c; m[11] = 0;
    m[12] = 0; m[13] = 0; m[14] = 0; m[15] = 1;
}

/* glOrtho / glFrustum matrices, evaluated in float after the GLdouble arguments
 * are narrowed (reference mat4_ortho / mat4_frustum, src/graphics.h:188-224). */
static inline void hm_ortho(float *m, float l, float r, float b, float t, float n, float f)
{
    float rl = r - l, tb = t - b, fn = f - n;
    hm_identity(m);
    if (rl == 0.0f || tb == 0.0f || fn == 0.0f) return;
    m[0] = 2.0f / rl;
    m[5] = 2.0f / tb;
    m[10] = -2.0f / fn;
    m[12] = -(r + l) / rl;
    m[13] = -(t + b) / tb;
    m[14] = -(f + n) / fn;
}

static inline void hm_frustum(float *m, float l, float r, float b, float t, float n, float f)
{
    float rl = r - l, tb = t - b, fn = f - n;
    hm_identity(m);
    if (rl == 0.0f || tb == 0.0f || fn == 0.0f) return;
    m[0] = 2.0f * n / rl;
    m[5] = 2.0f * n / tb;
    m[8] = (r + l) / rl;
    m[9] = (t + b) / tb;
    m[10] = -(f + n) / fn;
    m[11] = -1;
    m[14] = -2.0f * f * n / fn;
    m[15] = 0;
}

/* Inverse transpose of the upper 3x3 (cofactors times 1/det), identity when
 * |det| < 1e-10 (reference mat4_normal_matrix, src/graphics.h:229-260). */
static inline void hm_normal_matrix(const float *mv, float *r)
{
    float a = mv[0], b = mv[1], c = mv[2];
    float d = mv[4], e = mv[5], f = mv[6];
    float g = mv[8], h = mv[9], i = mv[10];
    float det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
    if (fabsf(det) < 1e-10f) { hm_identity(r); return; }
    float inv = 1.0f / det;
    memset(r, 0, 16 * sizeof(float));
    r[0]  = (e * i - f * h) * inv;
    r[1]  = (d * i - f * g) * -inv;
    r[2]  = (d * h - e * g) * inv;
    r[4]  = (b * i - c * h) * -inv;
    r[5]  = (a * i - c * g) * inv;
    r[6]  = (a * h - b * g) * -inv;
    r[8]  = (b * f - c * e) * inv;
    r[9]  = (a * f - c * d) * -inv;
    r[10] = (a * e - b * d) * inv;
    r[15] = 1.0f;
}

static inline float hm_clamp01(float v)
{
    return v < 0 ? 0 : (v > 1 ? 1 : v);
}

static inline mtgl_color hm_color_clamp(mtgl_color c)
{
    mtgl_color r = { hm_clamp01(c.r), hm_clamp01(c.g), hm_clamp01(c.b), hm_clamp01(c.a) };
    return r;
}

/* Float -> RGBA8 with clamp then truncation (reference color_to_rgba32,
 * src/graphics.h:337-348).  The float->int conversion follows x86 cvttss2si
 * (NaN gives 0 after the byte truncation). */
static inline uint32_t hm_pack_rgba8(mtgl_color c)
{
    float v[4] = { c.r, c.g, c.b, c.a };
    uint32_t out = 0;
    for (int k = 0; k < 4; k++) {
        float x = v[k] < 0.0f ? 0.0f : (v[k] > 1.0f ? 1.0f : v[k]);
        float y = x * 255.0f;
        uint32_t byte = (y == y) ? ((uint32_t)(int32_t)y & 0xFFu) : 0u;
        out |= byte << (8 * k);
    }
    return out;
}

#endif /* MTGL_GL_HOSTMATH_H */
