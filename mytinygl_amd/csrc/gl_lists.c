/*
 * gl_lists.c - display lists of libmtgl.
 *
 * Recording and replay follow the reference (src/lists.c, src/gl_api.c:2268-2499):
 * recorded calls are stored with their arguments; GL_COMPILE records without
 * executing, GL_COMPILE_AND_EXECUTE does both; replay calls the same GL entry
 * points, so list geometry goes through the device pipeline with the matrices
 * current at replay time.  List ids are 1-based, a contiguous run of freed ids
 * is reused first, at most 1024 ids exist, nesting stops at 64 levels.
 *
 * Device-resident list geometry: the vertex records a replayed glBegin/glEnd
 * block produces (object-space attributes and a state index; the matrices
 * and lighting travel separately per draw) are kept in a device segment per
 * block.  A replay runs the recorded calls on the host, compares the records
 * with the segment's host copy and sends them only when they differ (a block
 * that inherits the current colour or normal from outside the list can change
 * between calls).  A block of vertex / colour / texcoord / normal calls only,
 * entered from an empty vertex buffer, makes records fixed by those of the
 * current colour, texcoord and normal at its glBegin that it does not set
 * before its first vertex; when these equal those of its previous replay, the
 * calls are skipped: the block's side effects (current
 * attributes, GL_COLOR_MATERIAL, the vertex-state snapshot) are set from
 * what that replay left, and the segment is drawn.
 */
#include <stdlib.h>
#include <string.h>

#include "gl_state.h"

#define MTGL_EXPORT __attribute__((visibility("default")))
#define CTX_OR_RETURN()        GLState *c = gls_current(); if (!c) return
#define CTX_OR_RETURN_V(val)   GLState *c = gls_current(); if (!c) return (val)

static gls_list *list_lookup(GLState *c, GLuint id)
{
    if (id == 0 || id > c->list_count) return NULL;
    gls_list *l = &c->lists[id - 1];
    return l->allocated ? l : NULL;
}

int gls_list_recording(GLState *c)
{
    return c && c->list_index != 0;
}

/* Append cmd to the list being compiled.  Returns 1 when the caller must not
 * execute the call (GL_COMPILE), 0 otherwise (reference record_cmd,
 * src/lists.c:185-197). */
int gls_list_record(GLState *c, const gls_list_cmd *cmd)
{
    if (!c || c->list_index == 0) return 0;
    gls_list *l = list_lookup(c, c->list_index);
    if (l) {
        if (l->count == l->capacity) {
            size_t cap = l->capacity ? l->capacity * 2 : 64;
            gls_list_cmd *p = (gls_list_cmd *)realloc(l->cmds, cap * sizeof(*p));
            if (p) { l->cmds = p; l->capacity = cap; }
        }
        if (l->count < l->capacity) l->cmds[l->count++] = *cmd;
    }
    return c->list_mode == GL_COMPILE;
}

static void list_reset(gls_list *l)
{
    l->cmds = NULL;
    l->count = 0;
    l->capacity = 0;
    l->valid = GL_FALSE;
    l->allocated = GL_TRUE;
    l->segs = NULL;
    l->nsegs = l->segs_capacity = 0;
    l->seg_hint = 0;
}

void gls_list_free_segments(GLState *c, gls_list *l)
{
    for (size_t i = 0; i < l->nsegs; i++) {
        if (c->dev) mtgl_dev_segment_free(c->dev, l->segs[i].id);
        free(l->segs[i].host);
    }
    free(l->segs);
    l->segs = NULL;
    l->nsegs = l->segs_capacity = 0;
    l->seg_hint = 0;
}

/* The segment of the block whose LOP_END (by_begin 0) or LOP_BEGIN
 * (by_begin 1) is command i.  A replay visits blocks in order, so the search
 * starts at the segment the previous block used. */
static gls_lseg *seg_find(gls_list *l, size_t i, int by_begin)
{
    for (size_t k = 0; k < l->nsegs; k++) {
        size_t j = l->seg_hint + k;
        if (j >= l->nsegs) j -= l->nsegs;
        if ((by_begin ? l->segs[j].begin : l->segs[j].cmd) == i) {
            l->seg_hint = j;
            return &l->segs[j];
        }
    }
    return NULL;
}

uint32_t gls_list_segment(GLState *c, const void *recs, size_t bytes)
{
    if (!c || c->replay_list == 0 || bytes == 0) return 0;
    gls_list *l = list_lookup(c, c->replay_list);
    if (!l) return 0;
    gls_lseg *s = seg_find(l, c->replay_cmd, 0);
    if (s && s->bytes == bytes && memcmp(s->host, recs, bytes) == 0) return s->id;
    if (!s) {
        if (c->seg_next >= MTGL_MAX_SEGMENT) return 0;
        if (l->nsegs == l->segs_capacity) {
            size_t cap = l->segs_capacity ? l->segs_capacity * 2 : 4;
            gls_lseg *p = (gls_lseg *)realloc(l->segs, cap * sizeof(*p));
            if (!p) return 0;
            l->segs = p;
            l->segs_capacity = cap;
        }
        s = &l->segs[l->nsegs++];
        memset(s, 0, sizeof(*s));
        s->cmd = c->replay_cmd;
        s->begin = SIZE_MAX;
        s->id = ++c->seg_next;
    }
    s->fast = 0;
    uint8_t *h = (uint8_t *)realloc(s->host, bytes);
    if (!h) { s->bytes = 0; return 0; }
    s->host = h;
    memcpy(s->host, recs, bytes);
    s->bytes = bytes;
    if (mtgl_dev_segment_data(c->dev, s->id, (int64_t)bytes, recs) != MTGL_OK) {
        s->bytes = 0;   /* compared unequal next time: sent again */
        return 0;
    }
    return s->id;
}

MTGL_EXPORT GLuint glGenLists(GLsizei range)
{
    CTX_OR_RETURN_V(0);
    if (range < 0) { gls_set_error(c, GL_INVALID_VALUE); return 0; }
    if (range == 0) return 0;
    /* first run of `range` free slots among the existing ones */
    size_t run = 0, start = 0;
    for (size_t i = 0; i < c->list_count; i++) {
        if (!c->lists[i].allocated) {
            if (run == 0) start = i;
            if (++run >= (size_t)range) {
                for (GLsizei k = 0; k < range; k++) list_reset(&c->lists[start + (size_t)k]);
                return (GLuint)(start + 1);
            }
        } else {
            run = 0;
        }
    }
    if (c->list_count + (size_t)range > GLS_MAX_LISTS) return 0;
    GLuint first = (GLuint)(c->list_count + 1);
    for (GLsizei k = 0; k < range; k++) list_reset(&c->lists[c->list_count + (size_t)k]);
    c->list_count += (size_t)range;
    return first;
}

MTGL_EXPORT void glDeleteLists(GLuint list, GLsizei range)
{
    CTX_OR_RETURN();
    if (range < 0) { gls_set_error(c, GL_INVALID_VALUE); return; }
    for (GLsizei i = 0; i < range; i++) {
        gls_list *l = list_lookup(c, list + (GLuint)i);
        if (!l) continue;
        gls_list_free_segments(c, l);
        free(l->cmds);
        l->cmds = NULL;
        l->count = l->capacity = 0;
        l->valid = GL_FALSE;
        l->allocated = GL_FALSE;
    }
}

MTGL_EXPORT void glNewList(GLuint list, GLenum mode)
{
    CTX_OR_RETURN();
    if (c->list_index != 0) { gls_set_error(c, GL_INVALID_OPERATION); return; }
    if (mode != GL_COMPILE && mode != GL_COMPILE_AND_EXECUTE) { gls_set_error(c, GL_INVALID_ENUM); return; }
    gls_list *l = list_lookup(c, list);
    if (!l) { gls_set_error(c, GL_INVALID_VALUE); return; }
    gls_list_free_segments(c, l);
    l->count = 0;
    l->valid = GL_FALSE;
    c->list_index = list;
    c->list_mode = mode;
}

MTGL_EXPORT void glEndList(void)
{
    CTX_OR_RETURN();
    if (c->list_index == 0) { gls_set_error(c, GL_INVALID_OPERATION); return; }
    gls_list *l = list_lookup(c, c->list_index);
    if (l) l->valid = GL_TRUE;
    c->list_index = 0;
    c->list_mode = 0;
}

/* Current colour, texcoord and normal: what a block's vertex records take
 * from outside the block. */
static void current_attribs(const GLState *c, float *a)
{
    a[0] = c->current_color.r; a[1] = c->current_color.g;
    a[2] = c->current_color.b; a[3] = c->current_color.a;
    a[4] = c->current_texcoord[0]; a[5] = c->current_texcoord[1];
    a[6] = c->current_normal[0]; a[7] = c->current_normal[1]; a[8] = c->current_normal[2];
}

/* Which of colour (bit 0), texcoord (1) and normal (2) the block set before
 * its first vertex, as ops of the block are replayed one by one. */
static uint32_t attrib_bit(gls_list_op op)
{
    return op == LOP_COLOR ? 1u : op == LOP_TEXCOORD ? 2u : op == LOP_NORMAL ? 4u : 0u;
}

/* Compare the current attributes with a remembered start, over the attributes
 * the block's records take from outside it. */
static int same_start(const float *a, const float *b, uint32_t mask)
{
    return (!(mask & 1u) || memcmp(a, b, 4 * sizeof(float)) == 0)
        && (!(mask & 2u) || memcmp(a + 4, b + 4, 2 * sizeof(float)) == 0)
        && (!(mask & 4u) || memcmp(a + 6, b + 6, 3 * sizeof(float)) == 0);
}

/* After a block's glEnd: whether a later replay may skip its calls.  ok says
 * that the block started from an empty vertex buffer outside glBegin/glEnd,
 * with no list being compiled, and held vertex / colour / texcoord / normal
 * calls only; then its records (all with state index 0) are fixed by the
 * attributes of `start` it did not set before its first vertex (those outside
 * `set`), and the segment holds the ones this replay made exactly when their
 * count is nverts. */
static void remember_block(GLState *c, gls_list *l, size_t b, size_t e, int ok,
                           const float *start, uint32_t set, size_t nverts)
{
    gls_lseg *s = seg_find(l, e, 0);
    if (!s) return;
    s->begin = b;
    s->fast = ok && nverts > 0 && s->bytes == nverts * sizeof(gls_imm_vertex);
    if (!s->fast) return;
    memcpy(s->start, start, sizeof(s->start));
    s->start_mask = 7u & ~set;
    current_attribs(c, s->end);
    memcpy(s->last_color, ((const gls_imm_vertex *)s->host)[nverts - 1].color, sizeof(s->last_color));
    s->nverts = nverts;
}

/* A remembered block replayed from the same start: the side effects of its
 * calls (primitive mode, the vertex-state snapshot of its first vertex, the
 * current attributes it leaves, GL_COLOR_MATERIAL's write of its last
 * vertex's colour) and the draw of its segment. */
static int replay_block(GLState *c, const gls_lseg *s, GLenum mode)
{
    if (!gls_begin_snapshot(c)) return 0;
    c->primitive_mode = mode;
    c->current_color = (mtgl_color){s->end[0], s->end[1], s->end[2], s->end[3]};
    c->current_texcoord[0] = s->end[4]; c->current_texcoord[1] = s->end[5];
    c->current_normal[0] = s->end[6]; c->current_normal[1] = s->end[7];
    c->current_normal[2] = s->end[8];
    gls_apply_color_material(c, s->last_color);
    gls_draw_records(c, s->nverts, s->id);
    return 1;
}

static void execute_list(GLState *c, GLuint id)
{
    gls_list *l = list_lookup(c, id);
    if (!l || !l->valid) return;
    /* the block being replayed through its calls */
    size_t blk = SIZE_MAX, blk_verts = 0;
    int blk_ok = 0;
    uint32_t blk_set = 0;
    float blk_start[9] = {0};
    for (size_t i = 0; i < l->count; i++) {
        /* copy: a nested call may grow or free this list */
        gls_list_cmd k = l->cmds[i];
        const GLfloat *f = k.f;
        if (k.op > LOP_NORMAL) blk_ok = 0;
        if (blk_verts == 0) blk_set |= attrib_bit(k.op);
        switch (k.op) {
        case LOP_BEGIN: {
            blk = SIZE_MAX;
            if (c->list_index == 0 && !(c->flags & GLS_INSIDE_BEGIN_END) && k.e[0] <= GL_POLYGON
                && c->imm_count == 0 && c->vstate_count == 0) {
                float cur[9];
                current_attribs(c, cur);
                const gls_lseg *s = seg_find(l, i, 1);
                if (s && s->fast && same_start(s->start, cur, s->start_mask)
                    && replay_block(c, s, k.e[0])) {
                    i = s->cmd;
                    break;
                }
                blk = i; blk_ok = 1; blk_verts = 0; blk_set = 0;
                memcpy(blk_start, cur, sizeof(cur));
            }
            glBegin(k.e[0]);
            break;
        }
        case LOP_END: {
            /* nested replays restore the outer block's marks */
            const GLuint pl = c->replay_list;
            const size_t pc = c->replay_cmd;
            c->replay_list = id;
            c->replay_cmd = i;
            glEnd();
            c->replay_list = pl;
            c->replay_cmd = pc;
            l = list_lookup(c, id);
            if (l) remember_block(c, l, blk, i, blk != SIZE_MAX && blk_ok, blk_start, blk_set, blk_verts);
            blk = SIZE_MAX;
            break;
        }
        case LOP_VERTEX:        glVertex3f(f[0], f[1], f[2]); blk_verts++; break;
        case LOP_COLOR:         glColor4f(f[0], f[1], f[2], f[3]); break;
        case LOP_TEXCOORD:      glTexCoord2f(f[0], f[1]); break;
        case LOP_NORMAL:        glNormal3f(f[0], f[1], f[2]); break;
        case LOP_TRANSLATEF:    glTranslatef(f[0], f[1], f[2]); break;
        case LOP_ROTATEF:       glRotatef(f[0], f[1], f[2], f[3]); break;
        case LOP_SCALEF:        glScalef(f[0], f[1], f[2]); break;
        case LOP_PUSH_MATRIX:   glPushMatrix(); break;
        case LOP_POP_MATRIX:    glPopMatrix(); break;
        case LOP_LOAD_IDENTITY: glLoadIdentity(); break;
        case LOP_MULT_MATRIXF:  glMultMatrixf(f); break;
        case LOP_LOAD_MATRIXF:  glLoadMatrixf(f); break;
        case LOP_MATRIX_MODE:   glMatrixMode(k.e[0]); break;
        case LOP_ENABLE:        glEnable(k.e[0]); break;
        case LOP_DISABLE:       glDisable(k.e[0]); break;
        case LOP_BIND_TEXTURE:  glBindTexture(k.e[0], k.e[1]); break;
        case LOP_BLEND_FUNC:    glBlendFunc(k.e[0], k.e[1]); break;
        case LOP_DEPTH_FUNC:    glDepthFunc(k.e[0]); break;
        case LOP_DEPTH_MASK:    glDepthMask((GLboolean)k.e[0]); break;
        case LOP_CULL_FACE:     glCullFace(k.e[0]); break;
        case LOP_FRONT_FACE:    glFrontFace(k.e[0]); break;
        case LOP_SHADE_MODEL:   glShadeModel(k.e[0]); break;
        case LOP_LIGHTF:        glLightf(k.e[0], k.e[1], f[0]); break;
        case LOP_LIGHTFV:       glLightfv(k.e[0], k.e[1], f); break;
        case LOP_MATERIALF:     glMaterialf(k.e[0], k.e[1], f[0]); break;
        case LOP_MATERIALFV:    glMaterialfv(k.e[0], k.e[1], f); break;
        case LOP_CALL_LIST:     glCallList(k.e[0]); break;
        case LOP_ORTHO:         glOrtho(k.d[0], k.d[1], k.d[2], k.d[3], k.d[4], k.d[5]); break;
        case LOP_FRUSTUM:       glFrustum(k.d[0], k.d[1], k.d[2], k.d[3], k.d[4], k.d[5]); break;
        }
        l = list_lookup(c, id);
        if (!l) return;
    }
}

MTGL_EXPORT void glCallList(GLuint list)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_CALL_LIST, { list } };
    if (gls_list_record(c, &cmd)) return;
    if (c->list_call_depth >= GLS_LIST_DEPTH) { gls_set_error(c, GL_STACK_OVERFLOW); return; }
    c->list_call_depth++;
    execute_list(c, list);
    c->list_call_depth--;
}

MTGL_EXPORT void glCallLists(GLsizei n, GLenum type, const GLvoid *lists)
{
    CTX_OR_RETURN();
    if (n < 0) { gls_set_error(c, GL_INVALID_VALUE); return; }
    for (GLsizei i = 0; i < n; i++) {
        GLuint offset;
        switch (type) {
        case GL_UNSIGNED_BYTE:  offset = ((const GLubyte *)lists)[i]; break;
        case GL_UNSIGNED_SHORT: offset = ((const GLushort *)lists)[i]; break;
        case GL_UNSIGNED_INT:   offset = ((const GLuint *)lists)[i]; break;
        default: gls_set_error(c, GL_INVALID_ENUM); return;
        }
        glCallList(c->list_base + offset);
    }
}

MTGL_EXPORT void glListBase(GLuint base)
{
    CTX_OR_RETURN();
    c->list_base = base;
}

MTGL_EXPORT GLboolean glIsList(GLuint list)
{
    CTX_OR_RETURN_V(GL_FALSE);
    gls_list *l = list_lookup(c, list);
    return (l && l->valid) ? GL_TRUE : GL_FALSE;
}
