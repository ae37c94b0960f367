/*
 * gl_arrays.c - buffer objects, vertex arrays and array draws of libmtgl.
 *
 * The reference draws arrays by replaying immediate mode one element at a time
 * (glDrawArrays / glDrawElements, src/gl_api.c:1799-1941).  Here a whole array
 * draw becomes one mtgl_draw: buffer-object data already lives in HBM, client
 * arrays are staged once, and the device fetches every element itself with the
 * reference's get_array_element rules (src/gl_api.c:1758-1797).  The slow
 * replay path is kept for the cases whose observable behaviour depends on it:
 * display-list compilation, calls inside glBegin/glEnd, invalid primitive modes
 * and vertices left over from an unmatched glBegin.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdint.h>

#include "gl_state.h"
#include "gl_hostmath.h"

#define MTGL_EXPORT __attribute__((visibility("default")))
#define CTX_OR_RETURN()        GLState *c = gls_current(); if (!c) return
#define CTX_OR_RETURN_V(val)   GLState *c = gls_current(); if (!c) return (val)

/* ------------------------------------------------------------------------- */
/* buffer objects (reference src/vbo.c:34-161, src/gl_api.c:1516-1616)        */
/* ------------------------------------------------------------------------- */

static gls_buffer *buffer_lookup(GLState *c, GLuint id)
{
    if (id == 0 || id > c->buffer_count) return NULL;
    gls_buffer *b = &c->buffers[id - 1];
    return b->allocated ? b : NULL;
}

MTGL_EXPORT void glGenBuffers(GLsizei n, GLuint *buffers)
{
    CTX_OR_RETURN();
    if (n < 0) { gls_set_error(c, GL_INVALID_VALUE); return; }
    for (GLsizei i = 0; i < n; i++) {
        GLuint id = 0;
        for (size_t k = 0; k < c->buffer_count; k++) {
            if (!c->buffers[k].allocated) { id = (GLuint)(k + 1); break; }
        }
        if (!id && c->buffer_count < MTGL_MAX_OBJECTS) id = (GLuint)++c->buffer_count;
        if (id) {
            gls_buffer *b = &c->buffers[id - 1];
            b->allocated = 1;
            b->shadow = NULL;
            b->size = 0;
            b->usage = 0;
        }
        buffers[i] = id;
    }
}

MTGL_EXPORT void glDeleteBuffers(GLsizei n, const GLuint *buffers)
{
    CTX_OR_RETURN();
    if (n < 0) { gls_set_error(c, GL_INVALID_VALUE); return; }
    for (GLsizei i = 0; i < n; i++) {
        if (buffers[i] == c->bound_array_buffer) c->bound_array_buffer = 0;
        if (buffers[i] == c->bound_element_buffer) c->bound_element_buffer = 0;
    }
    for (GLsizei i = 0; i < n; i++) {
        gls_buffer *b = buffer_lookup(c, buffers[i]);
        if (!b) continue;
        if (b->shadow) gls_report_device_error(c, mtgl_dev_buffer_free(c->dev, buffers[i]), "glDeleteBuffers");
        free(b->shadow);
        b->shadow = NULL;
        b->size = 0;
        b->allocated = 0;
    }
}

MTGL_EXPORT void glBindBuffer(GLenum target, GLuint buffer)
{
    CTX_OR_RETURN();
    if (target == GL_ARRAY_BUFFER) c->bound_array_buffer = buffer;
    else if (target == GL_ELEMENT_ARRAY_BUFFER) c->bound_element_buffer = buffer;
    else gls_set_error(c, GL_INVALID_ENUM);
}

static int bound_buffer_id(GLState *c, GLenum target, GLuint *id)
{
    if (target == GL_ARRAY_BUFFER) { *id = c->bound_array_buffer; return 1; }
    if (target == GL_ELEMENT_ARRAY_BUFFER) { *id = c->bound_element_buffer; return 1; }
    gls_set_error(c, GL_INVALID_ENUM);
    return 0;
}

MTGL_EXPORT void glBufferData(GLenum target, GLsizeiptr size, const GLvoid *data, GLenum usage)
{
    CTX_OR_RETURN();
    if (size < 0) { gls_set_error(c, GL_INVALID_VALUE); return; }
    GLuint id;
    if (!bound_buffer_id(c, target, &id)) return;
    gls_buffer *b = buffer_lookup(c, id);
    if (!b) return;
    b->usage = usage;
    if (size <= 0) {
        if (b->shadow) gls_report_device_error(c, mtgl_dev_buffer_free(c->dev, id), "glBufferData");
        free(b->shadow);
        b->shadow = NULL;
        b->size = 0;
        return;
    }
    uint8_t *shadow = (uint8_t *)malloc((size_t)size);
    if (!shadow) return;                     /* old contents kept, as in the reference */
    if (data) memcpy(shadow, data, (size_t)size);
    else memset(shadow, 0, (size_t)size);
    int rc = mtgl_dev_buffer_data(c->dev, id, (int64_t)size, shadow);
    if (rc != MTGL_OK) { free(shadow); gls_report_device_error(c, rc, "glBufferData"); return; }
    free(b->shadow);
    b->shadow = shadow;
    b->size = size;
}

MTGL_EXPORT void glBufferSubData(GLenum target, GLintptr offset, GLsizeiptr size, const GLvoid *data)
{
    CTX_OR_RETURN();
    if (offset < 0 || size < 0) { gls_set_error(c, GL_INVALID_VALUE); return; }
    GLuint id;
    if (!bound_buffer_id(c, target, &id)) return;
    gls_buffer *b = buffer_lookup(c, id);
    if (!b) { gls_set_error(c, GL_INVALID_OPERATION); return; }
    if (!b->shadow || !data || offset + size > b->size) { gls_set_error(c, GL_INVALID_VALUE); return; }
    memcpy(b->shadow + offset, data, (size_t)size);
    gls_report_device_error(c, mtgl_dev_buffer_sub_data(c->dev, id, (int64_t)offset, (int64_t)size, data),
                            "glBufferSubData");
}

MTGL_EXPORT GLboolean glIsBuffer(GLuint buffer)
{
    CTX_OR_RETURN_V(GL_FALSE);
    if (buffer == 0) return GL_FALSE;
    return buffer_lookup(c, buffer) ? GL_TRUE : GL_FALSE;
}

/* ------------------------------------------------------------------------- */
/* client arrays (reference src/gl_api.c:1620-1742)                           */
/* ------------------------------------------------------------------------- */

static uint32_t client_bit(GLenum array)
{
    switch (array) {
    case GL_VERTEX_ARRAY: return GLS_CLIENT_VERTEX;
    case GL_COLOR_ARRAY: return GLS_CLIENT_COLOR;
    case GL_TEXTURE_COORD_ARRAY: return GLS_CLIENT_TEXCOORD;
    case GL_NORMAL_ARRAY: return GLS_CLIENT_NORMAL;
    default: return 0;
    }
}

MTGL_EXPORT void glEnableClientState(GLenum array)
{
    CTX_OR_RETURN();
    uint32_t bit = client_bit(array);
    if (!bit) { gls_set_error(c, GL_INVALID_ENUM); return; }
    c->client_state |= bit;
}

MTGL_EXPORT void glDisableClientState(GLenum array)
{
    CTX_OR_RETURN();
    uint32_t bit = client_bit(array);
    if (!bit) { gls_set_error(c, GL_INVALID_ENUM); return; }
    c->client_state &= ~bit;
}

static void set_pointer(GLState *c, gls_array_ptr *a, GLint min_size, GLint max_size, GLint size,
                        GLenum type, GLsizei stride, const GLvoid *pointer)
{
    if (size < min_size || size > max_size) { gls_set_error(c, GL_INVALID_VALUE); return; }
    if (type != GL_FLOAT && type != GL_UNSIGNED_BYTE) { gls_set_error(c, GL_INVALID_ENUM); return; }
    if (stride < 0) { gls_set_error(c, GL_INVALID_VALUE); return; }
    a->size = size;
    a->type = type;
    a->stride = stride;
    a->pointer = pointer;
}

MTGL_EXPORT void glVertexPointer(GLint size, GLenum type, GLsizei stride, const GLvoid *pointer)
{
    CTX_OR_RETURN();
    set_pointer(c, &c->vertex_ptr, 2, 4, size, type, stride, pointer);
}

MTGL_EXPORT void glColorPointer(GLint size, GLenum type, GLsizei stride, const GLvoid *pointer)
{
    CTX_OR_RETURN();
    set_pointer(c, &c->color_ptr, 3, 4, size, type, stride, pointer);
}

MTGL_EXPORT void glTexCoordPointer(GLint size, GLenum type, GLsizei stride, const GLvoid *pointer)
{
    CTX_OR_RETURN();
    set_pointer(c, &c->texcoord_ptr, 1, 4, size, type, stride, pointer);
}

MTGL_EXPORT void glNormalPointer(GLenum type, GLsizei stride, const GLvoid *pointer)
{
    CTX_OR_RETURN();
    if (type != GL_FLOAT && type != GL_UNSIGNED_BYTE) { gls_set_error(c, GL_INVALID_ENUM); return; }
    if (stride < 0) { gls_set_error(c, GL_INVALID_VALUE); return; }
    c->normal_ptr.size = 3;
    c->normal_ptr.type = type;
    c->normal_ptr.stride = stride;
    c->normal_ptr.pointer = pointer;
}

/* Base address of an array at draw time: with an array buffer bound, the pointer
 * is an offset into that buffer (reference get_array_pointer, :1745-1755). */
static const uint8_t *array_base(GLState *c, const gls_array_ptr *a, int *from_buffer, size_t *avail)
{
    *from_buffer = 0;
    *avail = SIZE_MAX;
    if (c->bound_array_buffer) {
        gls_buffer *b = buffer_lookup(c, c->bound_array_buffer);
        if (b && b->shadow) {
            size_t off = (size_t)a->pointer;
            *from_buffer = 1;
            *avail = off < (size_t)b->size ? (size_t)b->size - off : 0;
            return b->shadow + off;
        }
        return NULL;
    }
    return (const uint8_t *)a->pointer;
}

static GLsizei array_stride(const gls_array_ptr *a)
{
    if (a->stride != 0) return a->stride;
    return a->size * (a->type == GL_FLOAT ? (GLsizei)sizeof(GLfloat) : (GLsizei)sizeof(GLubyte));
}

/* One element with the reference's defaults: missing components are 0 except
 * the 4th, which is 1; a negative index yields (0,0,0,1)
 * (reference get_array_element, :1758-1797).  An element that does not fit in
 * the bytes available behind `base` (a buffer object read past its end, which
 * the reference leaves undefined) also takes the defaults, as on the device. */
static void fetch_element(const gls_array_ptr *a, const uint8_t *base, size_t avail, GLint index,
                          float *out, int n)
{
    GLsizei stride = array_stride(a);
    size_t elem = (size_t)a->size * (a->type == GL_FLOAT ? 4u : 1u);
    if (!base || index < 0 || stride <= 0 ||
        (avail != SIZE_MAX && ((size_t)index * (size_t)stride > avail || avail - (size_t)index * (size_t)stride < elem))) {
        for (int i = 0; i < n; i++) out[i] = i < 3 ? 0.0f : 1.0f;
        return;
    }
    const uint8_t *p = base + (size_t)index * (size_t)stride;
    for (int i = 0; i < a->size && i < n; i++) {
        if (a->type == GL_FLOAT) { float v; memcpy(&v, p + 4 * i, 4); out[i] = v; }
        else out[i] = p[i] / 255.0f;
    }
    for (int i = a->size; i < n; i++) out[i] = i == 3 ? 1.0f : 0.0f;
}

/* ------------------------------------------------------------------------- */
/* array draws                                                                */
/* ------------------------------------------------------------------------- */

typedef struct {
    const uint8_t *base[4];    /* position, colour, texcoord, normal; NULL = not used */
    int from_buffer[4];
    size_t avail[4];           /* bytes readable behind base (SIZE_MAX: client memory) */
    const gls_array_ptr *ptr[4];
} draw_arrays_t;

static void gather_arrays(GLState *c, draw_arrays_t *da)
{
    static const uint32_t bits[4] = { GLS_CLIENT_VERTEX, GLS_CLIENT_COLOR, GLS_CLIENT_TEXCOORD, GLS_CLIENT_NORMAL };
    const gls_array_ptr *ptrs[4] = { &c->vertex_ptr, &c->color_ptr, &c->texcoord_ptr, &c->normal_ptr };
    for (int k = 0; k < 4; k++) {
        da->ptr[k] = ptrs[k];
        da->from_buffer[k] = 0;
        da->base[k] = NULL;
        da->avail[k] = 0;
        if (k == 0 || (c->client_state & bits[k]))
            da->base[k] = array_base(c, ptrs[k], &da->from_buffer[k], &da->avail[k]);
    }
}

/* The reference per-element replay (glBegin, glColor4f / glTexCoord2f / glNormal3f,
 * glVertex, glEnd), used where its side effects are observable. */
static void replay_elements(GLState *c, GLenum mode, const draw_arrays_t *da, GLsizei count,
                            GLint first, const void *indices, GLenum itype)
{
    glBegin(mode);
    for (GLsizei i = 0; i < count; i++) {
        GLint idx;
        if (!indices) idx = first + i;
        else if (itype == GL_UNSIGNED_SHORT) idx = (GLint)((const GLushort *)indices)[i];
        else if (itype == GL_UNSIGNED_INT) idx = (GLint)((const GLuint *)indices)[i];
        else idx = (GLint)((const GLubyte *)indices)[i];
        float v[4], col[4], t[2], n[3];
        fetch_element(da->ptr[0], da->base[0], da->avail[0], idx, v, 4);
        if (da->base[1]) { fetch_element(da->ptr[1], da->base[1], da->avail[1], idx, col, 4); glColor4f(col[0], col[1], col[2], col[3]); }
        if (da->base[2]) { fetch_element(da->ptr[2], da->base[2], da->avail[2], idx, t, 2); glTexCoord2f(t[0], t[1]); }
        if (da->base[3]) { fetch_element(da->ptr[3], da->base[3], da->avail[3], idx, n, 3); glNormal3f(n[0], n[1], n[2]); }
        if (da->ptr[0]->size == 2) glVertex2f(v[0], v[1]);
        else glVertex3f(v[0], v[1], v[2]);
        (void)c;
    }
    glEnd();
}

static int fast_path_ok(GLState *c, GLenum mode)
{
    return c->list_index == 0 && !(c->flags & GLS_INSIDE_BEGIN_END) && mode <= GL_POLYGON && c->imm_count == 0;
}

static uint8_t *stage_reserve(GLState *c, size_t bytes)
{
    if (bytes > c->stage_capacity) {
        size_t cap = c->stage_capacity ? c->stage_capacity : 4096;
        while (cap < bytes) cap *= 2;
        uint8_t *p = (uint8_t *)realloc(c->stage, cap);
        if (!p) return NULL;
        c->stage = p;
        c->stage_capacity = cap;
    }
    return c->stage;
}

/* Describe attribute k of the draw.  Buffer-object arrays point into the device
 * copy; client arrays are staged over the element window [lo, hi]. */
static int describe_array(GLState *c, const draw_arrays_t *da, int k, int64_t lo, int64_t hi,
                          size_t *stage_used, mtgl_array *out)
{
    memset(out, 0, sizeof(*out));
    if (!da->base[k]) return 1;
    const gls_array_ptr *a = da->ptr[k];
    GLsizei stride = array_stride(a);
    if (stride <= 0) return 1;                /* every element takes its defaults */
    out->enabled = 1;
    out->size = a->size;
    out->type = a->type;
    out->stride = stride;
    if (da->from_buffer[k]) {
        out->buffer = c->bound_array_buffer;
        out->offset = (int64_t)(size_t)a->pointer;
        return 1;
    }
    out->buffer = 0;
    if (lo > hi) { out->offset = 0; return 1; }   /* no element with a valid index */
    size_t elem = (size_t)a->size * (a->type == GL_FLOAT ? 4u : 1u);
    size_t bytes = (size_t)(hi - lo) * (size_t)stride + elem;
    size_t at = (*stage_used + 15) & ~(size_t)15;
    uint8_t *stage = stage_reserve(c, at + bytes);
    if (!stage) return 0;
    memcpy(stage + at, da->base[k] + (size_t)lo * (size_t)stride, bytes);
    out->offset = (int64_t)at - lo * (int64_t)stride;
    *stage_used = at + bytes;
    return 1;
}

/* Current values after the draw: the last element's colour (through glColor4f's
 * clamping), texcoord and normal, plus the GL_COLOR_MATERIAL side effect of that
 * colour on the materials. */
static void apply_last_element(GLState *c, const draw_arrays_t *da, GLint idx)
{
    float col[4], t[2], n[3];
    if (da->base[1]) {
        fetch_element(da->ptr[1], da->base[1], da->avail[1], idx, col, 4);
        float *dst = &c->current_color.r;
        for (int i = 0; i < 4; i++) {
            float v = col[i];
            if (isnan(v) || isinf(v)) v = i == 3 ? 1.0f : 0.0f;
            dst[i] = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
        }
    }
    if (da->base[2]) { fetch_element(da->ptr[2], da->base[2], da->avail[2], idx, t, 2); c->current_texcoord[0] = t[0]; c->current_texcoord[1] = t[1]; }
    if (da->base[3]) { fetch_element(da->ptr[3], da->base[3], da->avail[3], idx, n, 3); memcpy(c->current_normal, n, sizeof(n)); }
    float cc[4] = { c->current_color.r, c->current_color.g, c->current_color.b, c->current_color.a };
    gls_apply_color_material(c, cc);
}

static void issue_array_draw(GLState *c, GLenum mode, const draw_arrays_t *da, GLsizei count, GLint first,
                             GLenum itype, const mtgl_array *index_desc, int64_t lo, int64_t hi,
                             const void *client_indices, GLint last_index)
{
    mtgl_vertex_state vs;
    gls_snapshot_vertex_state(c, &vs);
    size_t used = 0;
    mtgl_draw d;
    memset(&d, 0, sizeof(d));
    d.mode = mode;
    d.count = (uint32_t)count;
    d.first = first;
    d.index_type = itype;
    if (itype) {
        d.indices = *index_desc;
        if (client_indices) {
            size_t isz = itype == GL_UNSIGNED_INT ? 4 : (itype == GL_UNSIGNED_SHORT ? 2 : 1);
            size_t bytes = isz * (size_t)count;
            uint8_t *stage = stage_reserve(c, bytes);
            if (!stage) { gls_set_error(c, GL_OUT_OF_MEMORY); return; }
            memcpy(stage, client_indices, bytes);
            used = bytes;
            d.indices.offset = 0;
            d.indices.buffer = 0;
        }
    }
    mtgl_array *outs[4] = { &d.position, &d.color, &d.texcoord, &d.normal };
    for (int k = 0; k < 4; k++) {
        if (!describe_array(c, da, k, lo, hi, &used, outs[k])) { gls_set_error(c, GL_OUT_OF_MEMORY); return; }
    }
    d.cur_color[0] = c->current_color.r; d.cur_color[1] = c->current_color.g;
    d.cur_color[2] = c->current_color.b; d.cur_color[3] = c->current_color.a;
    d.cur_texcoord[0] = c->current_texcoord[0]; d.cur_texcoord[1] = c->current_texcoord[1];
    memcpy(d.cur_normal, c->current_normal, sizeof(d.cur_normal));
    d.n_vstates = 1;
    d.vstates = &vs;
    d.staged = c->stage;
    d.staged_bytes = used;

    /* the rasteriser sees the state as left by the last vertex */
    apply_last_element(c, da, last_index);
    mtgl_raster_state rs;
    gls_snapshot_raster_state(c, &rs);
    d.raster = &rs;
    c->primitive_mode = mode;
    int rc = mtgl_dev_draw(c->dev, &d);
    gls_report_device_error(c, rc, "array draw");
    c->tl_dirty = 1;
}

MTGL_EXPORT void glDrawArrays(GLenum mode, GLint first, GLsizei count)
{
    CTX_OR_RETURN();
    if (count < 0) { gls_set_error(c, GL_INVALID_VALUE); return; }
    draw_arrays_t da;
    gather_arrays(c, &da);
    if (!(c->client_state & GLS_CLIENT_VERTEX) || !da.base[0]) return;
    if (!fast_path_ok(c, mode)) { replay_elements(c, mode, &da, count, first, NULL, 0); return; }
    if (count == 0) return;
    int64_t lo = first < 0 ? 0 : first;
    int64_t hi = (int64_t)first + count - 1;
    issue_array_draw(c, mode, &da, count, first, 0, NULL, lo, hi, NULL, (GLint)((int64_t)first + count - 1));
}

MTGL_EXPORT void glDrawElements(GLenum mode, GLsizei count, GLenum type, const GLvoid *indices)
{
    CTX_OR_RETURN();
    if (count < 0) { gls_set_error(c, GL_INVALID_VALUE); return; }
    if (type != GL_UNSIGNED_BYTE && type != GL_UNSIGNED_SHORT && type != GL_UNSIGNED_INT) {
        gls_set_error(c, GL_INVALID_ENUM); return;
    }
    draw_arrays_t da;
    gather_arrays(c, &da);
    const uint8_t *index_data = (const uint8_t *)indices;
    mtgl_array idesc;
    memset(&idesc, 0, sizeof(idesc));
    size_t isz = type == GL_UNSIGNED_INT ? 4 : (type == GL_UNSIGNED_SHORT ? 2 : 1);
    idesc.enabled = 1;
    idesc.size = 1;
    idesc.type = type;
    idesc.stride = (int32_t)isz;
    const void *client_indices = indices;
    if (c->bound_element_buffer) {
        gls_buffer *b = buffer_lookup(c, c->bound_element_buffer);
        if (!b || !b->shadow) return;
        uintptr_t off = (uintptr_t)indices;
        if (off >= (uintptr_t)b->size) { gls_set_error(c, GL_INVALID_VALUE); return; }
        index_data = b->shadow + off;
        idesc.buffer = c->bound_element_buffer;
        idesc.offset = (int64_t)off;
        client_indices = NULL;
    }
    if (!(c->client_state & GLS_CLIENT_VERTEX) || !da.base[0] || !index_data) return;
    if (!fast_path_ok(c, mode)) { replay_elements(c, mode, &da, count, 0, index_data, type); return; }
    if (count == 0) return;

    /* element window of the client arrays and the last element index */
    int any_client = 0;
    for (int k = 0; k < 4; k++) if (da.base[k] && !da.from_buffer[k]) any_client = 1;
    int64_t lo = INT64_MAX, hi = -1;
    GLint last = 0;
    for (GLsizei i = 0; i < count; i++) {
        GLint idx;
        if (type == GL_UNSIGNED_SHORT) { GLushort v; memcpy(&v, index_data + 2 * (size_t)i, 2); idx = (GLint)v; }
        else if (type == GL_UNSIGNED_INT) { GLuint v; memcpy(&v, index_data + 4 * (size_t)i, 4); idx = (GLint)v; }
        else idx = (GLint)index_data[i];
        if (i == count - 1) last = idx;
        if (!any_client) continue;
        if (idx >= 0) {
            if (idx < lo) lo = idx;
            if (idx > hi) hi = idx;
        }
    }
    if (!any_client) { lo = 0; hi = -1; }
    issue_array_draw(c, mode, &da, count, 0, type, &idesc, lo, hi, client_indices, last);
}
