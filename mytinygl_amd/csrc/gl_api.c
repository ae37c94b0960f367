/*
 * gl_api.c - GL 1.5 entry points of libmtgl (state machine, matrices, immediate
 * mode, textures, lighting, fog, pixel transfer, queries).
 *
 * Argument validation, sticky error codes and display-list recording follow the
 * reference entry points in src/gl_api.c (cited per function).  What changes is
 * where the work runs: primitives are not rasterised here but handed to the
 * device pipeline through mtgl_dev_draw() (include/mtgl_dev.h), and the pixel
 * planes live in HBM.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "gl_state.h"
#include "gl_hostmath.h"

#define MTGL_EXPORT __attribute__((visibility("default")))

static __thread GLState *g_ctx = NULL;

#define CTX_OR_RETURN()        GLState *c = g_ctx; if (!c) return
#define CTX_OR_RETURN_V(val)   GLState *c = g_ctx; if (!c) return (val)
#define TL_DIRTY(c)            ((c)->tl_dirty = 1)

GLState *gls_current(void) { return g_ctx; }

/* first error sticks until glGetError (reference gl_set_error, src/gl_api.c:30-35) */
void gls_set_error(GLState *c, GLenum err)
{
    if (c && c->error == GL_NO_ERROR) c->error = err;
}

/* A device failure becomes a sticky GL error: GL_OUT_OF_MEMORY for an
 * allocation failure, GL_INVALID_OPERATION for everything else, including a
 * call the device refuses (a point, line or glDrawPixels call while a batch is
 * recording), so that a C caller sees with glGetError that the call drew
 * nothing.  The refusal is also reported once per context on stderr. */
void gls_report_device_error(GLState *c, int rc, const char *what)
{
    if (rc == MTGL_OK) return;
    if (rc == MTGL_ERR_UNSUPPORTED) {
        if (!(c->warned & 1u)) {
            fprintf(stderr, "libmtgl: %s: the device pipeline refused this call (unsupported state); skipped\n",
                    what);
            c->warned |= 1u;
        }
    } else {
        fprintf(stderr, "libmtgl: device error %d in %s\n", rc, what);
    }
    gls_set_error(c, rc == MTGL_ERR_OUT_OF_MEMORY ? GL_OUT_OF_MEMORY : GL_INVALID_OPERATION);
}

/* ------------------------------------------------------------------------- */
/* context lifetime (reference src/gl_api.c:81-260)                           */
/* ------------------------------------------------------------------------- */

static void light_defaults(mtgl_light *l, int index)
{
    memset(l, 0, sizeof(*l));
    l->ambient = (mtgl_color){0, 0, 0, 1};
    l->diffuse = index == 0 ? (mtgl_color){1, 1, 1, 1} : (mtgl_color){0, 0, 0, 1};
    l->specular = l->diffuse;
    l->position[0] = 0; l->position[1] = 0; l->position[2] = 1; l->position[3] = 0;
    l->spot_direction[0] = 0; l->spot_direction[1] = 0; l->spot_direction[2] = -1;
    l->spot_exponent = 0;
    l->spot_cutoff = 180;
    l->spot_cos_cutoff = cosf(180.0f * 3.14159265f / 180.0f);
    l->constant_attenuation = 1;
    l->linear_attenuation = 0;
    l->quadratic_attenuation = 0;
    l->enabled = 0;
}

static void material_defaults(mtgl_material *m)
{
    memset(m, 0, sizeof(*m));
    m->ambient = (mtgl_color){0.2f, 0.2f, 0.2f, 1};
    m->diffuse = (mtgl_color){0.8f, 0.8f, 0.8f, 1};
    m->specular = (mtgl_color){0, 0, 0, 1};
    m->emission = (mtgl_color){0, 0, 0, 1};
    m->shininess = 0;
}

MTGL_EXPORT GLState *gl_create_context(int32_t width, int32_t height)
{
    if (width <= 0 || height <= 0 || width > MTGL_MAX_FB_DIM || height > MTGL_MAX_FB_DIM)
        return NULL;
    GLState *c = (GLState *)calloc(1, sizeof(GLState));
    if (!c) return NULL;
    int rc = mtgl_dev_create(width, height, &c->dev);
    if (rc != MTGL_OK) {
        fprintf(stderr, "libmtgl: gl_create_context(%d, %d) failed: device error %d (%s)\n",
                width, height, rc, mtgl_dev_backend());
        free(c);
        return NULL;
    }
    c->width = width;
    c->height = height;
    c->clear_color = (mtgl_color){0, 0, 0, 1};
    c->clear_depth = 1.0;
    c->viewport_w = width;
    c->viewport_h = height;
    c->current_color = (mtgl_color){1, 1, 1, 1};
    c->current_normal[2] = 1.0f;
    c->matrix_mode = GL_MODELVIEW;
    for (int i = 0; i < GLS_STACK_DEPTH; i++) {
        hm_identity(c->modelview[i]);
        hm_identity(c->projection[i]);
        hm_identity(c->texmat[i]);
    }
    c->blend_src = GL_ONE;
    c->blend_dst = GL_ZERO;
    c->cull_face_mode = GL_BACK;
    c->front_face = GL_CCW;
    c->depth_func = GL_LESS;
    c->depth_mask = GL_TRUE;
    c->alpha_func = GL_ALWAYS;
    c->scissor_w = width;
    c->scissor_h = height;
    c->stencil_func = GL_ALWAYS;
    c->stencil_mask = 0xFFFFFFFFu;
    c->stencil_fail = c->stencil_zfail = c->stencil_zpass = GL_KEEP;
    c->stencil_writemask = 0xFFFFFFFFu;
    c->depth_near = 0.0;
    c->depth_far = 1.0;
    for (int i = 0; i < 4; i++) c->color_mask[i] = GL_TRUE;
    c->line_width = 1.0f;
    c->point_size = 1.0f;
    c->polygon_mode_front = c->polygon_mode_back = GL_FILL;
    c->tex_env_mode = GL_MODULATE;
    c->perspective_hint = GL_DONT_CARE;
    c->raster_valid = GL_TRUE;
    c->fog_mode = GL_EXP;
    c->fog_density = 1.0f;
    c->fog_end = 1.0f;
    for (int i = 0; i < MTGL_MAX_LIGHTS; i++) light_defaults(&c->lights[i], i);
    material_defaults(&c->material_front);
    c->material_back = c->material_front;
    c->light_model_ambient = (mtgl_color){0.2f, 0.2f, 0.2f, 1};
    c->color_material_face = GL_FRONT_AND_BACK;
    c->color_material_mode = GL_AMBIENT_AND_DIFFUSE;
    c->shade_model = GL_SMOOTH;
    c->tl_dirty = 1;
    c->error = GL_NO_ERROR;
    return c;
}

MTGL_EXPORT void gl_destroy_context(GLState *c)
{
    if (!c) return;
    if (g_ctx == c) g_ctx = NULL;
    for (size_t i = 0; i < GLS_MAX_LISTS; i++) {
        free(c->lists[i].cmds);
        for (size_t k = 0; k < c->lists[i].nsegs; k++) free(c->lists[i].segs[k].host);
        free(c->lists[i].segs);   /* the device segments go with the device context */
    }
    for (size_t i = 0; i < MTGL_MAX_OBJECTS; i++) free(c->buffers[i].shadow);
    free(c->imm);
    free(c->vstates);
    free(c->stage);
    mtgl_dev_destroy(c->dev);
    free(c);
}

MTGL_EXPORT void gl_make_current(GLState *c) { g_ctx = c; }
MTGL_EXPORT GLState *gl_get_current_context(void) { return g_ctx; }
MTGL_EXPORT int32_t mtgl_context_width(GLState *c) { return c ? c->width : 0; }
MTGL_EXPORT int32_t mtgl_context_height(GLState *c) { return c ? c->height : 0; }
MTGL_EXPORT mtgl_dev *mtgl_context_device(GLState *c) { return c ? c->dev : NULL; }

MTGL_EXPORT void mtgl_texture_lod_mode(GLState *c, uint32_t mode)
{
    if (!c) return;
    if (mode != MTGL_LOD_REFERENCE && mode != MTGL_LOD_FULL_CHAIN) { gls_set_error(c, GL_INVALID_ENUM); return; }
    c->lod_mode = mode;
}

MTGL_EXPORT int mtgl_read_framebuffer(GLState *c, uint32_t *color, float *depth, uint8_t *stencil)
{
    if (!c) return MTGL_ERR_INVALID;
    int rc = MTGL_OK;
    if (color && rc == MTGL_OK) rc = mtgl_dev_read(c->dev, MTGL_BUF_COLOR, 0, 0, c->width, c->height, color);
    if (depth && rc == MTGL_OK) rc = mtgl_dev_read(c->dev, MTGL_BUF_DEPTH, 0, 0, c->width, c->height, depth);
    if (stencil && rc == MTGL_OK) rc = mtgl_dev_read(c->dev, MTGL_BUF_STENCIL, 0, 0, c->width, c->height, stencil);
    return rc;
}

MTGL_EXPORT int mtgl_present(GLState *c, void *pixels, int32_t pitch)
{
    if (!c || !pixels) return MTGL_ERR_INVALID;
    return mtgl_dev_present(c->dev, pixels, pitch);
}

/* ------------------------------------------------------------------------- */
/* state snapshots handed to the device                                        */
/* ------------------------------------------------------------------------- */

static void snapshot_lighting(GLState *c, mtgl_lighting *l)
{
    memcpy(l->lights, c->lights, sizeof(l->lights));
    l->front = c->material_front;
    l->back = c->material_back;
    l->model_ambient = c->light_model_ambient;
    l->local_viewer = c->light_model_local_viewer ? 1u : 0u;
    l->two_side = c->light_model_two_side ? 1u : 0u;
    l->pad_[0] = l->pad_[1] = 0;
}

void gls_snapshot_vertex_state(GLState *c, mtgl_vertex_state *vs)
{
    memset(vs, 0, sizeof(*vs));
    memcpy(vs->modelview, c->modelview[c->modelview_depth], sizeof(vs->modelview));
    memcpy(vs->projection, c->projection[c->projection_depth], sizeof(vs->projection));
    memcpy(vs->texture, c->texmat[c->texture_depth], sizeof(vs->texture));
    hm_normal_matrix(vs->modelview, vs->normal_matrix);
    vs->flags = c->flags & ~GLS_INSIDE_BEGIN_END;
    vs->shade_model = c->shade_model;
    vs->color_material_face = c->color_material_face;
    vs->color_material_mode = c->color_material_mode;
    snapshot_lighting(c, &vs->lighting);
}

void gls_snapshot_raster_state(GLState *c, mtgl_raster_state *rs)
{
    memset(rs, 0, sizeof(*rs));
    rs->flags = c->flags & ~GLS_INSIDE_BEGIN_END;
    rs->viewport[0] = c->viewport_x; rs->viewport[1] = c->viewport_y;
    rs->viewport[2] = c->viewport_w; rs->viewport[3] = c->viewport_h;
    rs->scissor[0] = c->scissor_x; rs->scissor[1] = c->scissor_y;
    rs->scissor[2] = c->scissor_w; rs->scissor[3] = c->scissor_h;
    rs->cull_face_mode = c->cull_face_mode;
    rs->front_face = c->front_face;
    rs->polygon_mode_front = c->polygon_mode_front;
    rs->polygon_mode_back = c->polygon_mode_back;
    rs->shade_model = c->shade_model;
    rs->depth_func = c->depth_func;
    rs->depth_mask = c->depth_mask ? 1u : 0u;
    rs->depth_near = c->depth_near;
    rs->depth_far = c->depth_far;
    rs->stencil_func = c->stencil_func;
    rs->stencil_ref = c->stencil_ref;
    rs->stencil_mask = c->stencil_mask;
    rs->stencil_fail = c->stencil_fail;
    rs->stencil_zfail = c->stencil_zfail;
    rs->stencil_zpass = c->stencil_zpass;
    rs->stencil_writemask = c->stencil_writemask;
    rs->alpha_func = c->alpha_func;
    rs->alpha_ref = c->alpha_ref;
    rs->blend_src = c->blend_src;
    rs->blend_dst = c->blend_dst;
    rs->color_mask = (c->color_mask[0] ? 1u : 0u) | (c->color_mask[1] ? 2u : 0u) |
                     (c->color_mask[2] ? 4u : 0u) | (c->color_mask[3] ? 8u : 0u);
    /* the draw samples a texture only when texturing is on, a texture is bound and
     * it holds pixels (reference src/raster.c:494-498, :509, :618) */
    rs->texture = 0;
    if ((c->flags & MTGL_F_TEXTURE_2D) && c->bound_texture != 0 &&
        c->bound_texture <= c->texture_count) {
        gls_texture *t = &c->textures[c->bound_texture - 1];
        if (t->allocated && t->width > 0 && t->height > 0) {
            rs->texture = c->bound_texture;
            rs->min_filter = t->min_filter;
            rs->mag_filter = t->mag_filter;
            rs->wrap_s = t->wrap_s;
            rs->wrap_t = t->wrap_t;
        }
    }
    rs->lod_mode = c->lod_mode;
    rs->tex_env_mode = c->tex_env_mode;
    rs->tex_env_color = c->tex_env_color;
    rs->perspective_hint = c->perspective_hint;
    rs->fog_mode = c->fog_mode;
    rs->fog_density = c->fog_density;
    rs->fog_start = c->fog_start;
    rs->fog_end = c->fog_end;
    rs->fog_color = c->fog_color;
    rs->line_width = c->line_width;
    rs->point_size = c->point_size;
    snapshot_lighting(c, &rs->lighting);
}

/* GL_COLOR_MATERIAL side effect of one vertex: the clamped colour overwrites the
 * selected material fields (reference emit_vertex, src/gl_api.c:285-312). */
void gls_apply_color_material(GLState *c, const float *color)
{
    if (!((c->flags & MTGL_F_LIGHTING) && (c->flags & MTGL_F_COLOR_MATERIAL))) return;
    mtgl_color k = hm_color_clamp((mtgl_color){color[0], color[1], color[2], color[3]});
    GLenum mode = c->color_material_mode, face = c->color_material_face;
    mtgl_material *mats[2] = {
        (face == GL_FRONT || face == GL_FRONT_AND_BACK) ? &c->material_front : NULL,
        (face == GL_BACK || face == GL_FRONT_AND_BACK) ? &c->material_back : NULL };
    for (int i = 0; i < 2; i++) {
        mtgl_material *m = mats[i];
        if (!m) continue;
        if (mode == GL_AMBIENT || mode == GL_AMBIENT_AND_DIFFUSE) m->ambient = k;
        if (mode == GL_DIFFUSE || mode == GL_AMBIENT_AND_DIFFUSE) m->diffuse = k;
        if (mode == GL_SPECULAR) m->specular = k;
        if (mode == GL_EMISSION) m->emission = k;
    }
}

/* ------------------------------------------------------------------------- */
/* enable / clear / viewport (reference src/gl_api.c:351-478)                 */
/* ------------------------------------------------------------------------- */

static uint32_t cap_flag(GLenum cap)
{
    switch (cap) {
    case GL_DEPTH_TEST:     return MTGL_F_DEPTH_TEST;
    case GL_CULL_FACE:      return MTGL_F_CULL_FACE;
    case GL_BLEND:          return MTGL_F_BLEND;
    case GL_TEXTURE_2D:     return MTGL_F_TEXTURE_2D;
    case GL_LIGHTING:       return MTGL_F_LIGHTING;
    case GL_FOG:            return MTGL_F_FOG;
    case GL_NORMALIZE:      return MTGL_F_NORMALIZE;
    case GL_COLOR_MATERIAL: return MTGL_F_COLOR_MATERIAL;
    case GL_ALPHA_TEST:     return MTGL_F_ALPHA_TEST;
    case GL_SCISSOR_TEST:   return MTGL_F_SCISSOR_TEST;
    case GL_STENCIL_TEST:   return MTGL_F_STENCIL_TEST;
    default:                return 0;
    }
}

static void set_cap(GLState *c, GLenum cap, int on)
{
    uint32_t f = cap_flag(cap);
    TL_DIRTY(c);
    if (f) {
        if (on) c->flags |= f; else c->flags &= ~f;
        return;
    }
    if (cap >= GL_LIGHT0 && cap <= GL_LIGHT7) {
        c->lights[cap - GL_LIGHT0].enabled = on ? 1u : 0u;
        return;
    }
    gls_set_error(c, GL_INVALID_ENUM);
}

MTGL_EXPORT void glEnable(GLenum cap)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_ENABLE, { cap } };
    if (gls_list_record(c, &cmd)) return;
    set_cap(c, cap, 1);
}

MTGL_EXPORT void glDisable(GLenum cap)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_DISABLE, { cap } };
    if (gls_list_record(c, &cmd)) return;
    set_cap(c, cap, 0);
}

/* Clear region: the scissor box clamped to the framebuffer when the scissor test
 * is on, else the whole framebuffer; colour mask and stencil write mask do not
 * apply (reference src/gl_api.c:409-457). */
MTGL_EXPORT void glClear(GLbitfield mask)
{
    CTX_OR_RETURN();
    int32_t x0 = 0, y0 = 0, x1 = c->width, y1 = c->height;
    if (c->flags & MTGL_F_SCISSOR_TEST) {
        x0 = c->scissor_x; y0 = c->scissor_y;
        x1 = c->scissor_x + c->scissor_w; y1 = c->scissor_y + c->scissor_h;
        if (x0 < 0) x0 = 0;
        if (y0 < 0) y0 = 0;
        if (x1 > c->width) x1 = c->width;
        if (y1 > c->height) y1 = c->height;
    }
    uint32_t bufs = 0;
    if (mask & GL_COLOR_BUFFER_BIT) bufs |= MTGL_BUF_COLOR;
    if (mask & GL_DEPTH_BUFFER_BIT) bufs |= MTGL_BUF_DEPTH;
    if (mask & GL_STENCIL_BUFFER_BIT) bufs |= MTGL_BUF_STENCIL;
    if (!bufs || x0 >= x1 || y0 >= y1) return;
    int rc = mtgl_dev_clear(c->dev, bufs, x0, y0, x1, y1, hm_pack_rgba8(c->clear_color),
                            (float)c->clear_depth, (uint8_t)(c->stencil_clear & 0xFF));
    gls_report_device_error(c, rc, "glClear");
}

MTGL_EXPORT void glClearColor(GLclampf r, GLclampf g, GLclampf b, GLclampf a)
{
    CTX_OR_RETURN();
    c->clear_color = (mtgl_color){r, g, b, a};
}

MTGL_EXPORT void glClearDepth(GLclampd depth)
{
    CTX_OR_RETURN();
    c->clear_depth = depth;
}

MTGL_EXPORT void glClearStencil(GLint s)
{
    CTX_OR_RETURN();
    c->stencil_clear = s;
}

MTGL_EXPORT void glViewport(GLint x, GLint y, GLsizei w, GLsizei h)
{
    CTX_OR_RETURN();
    c->viewport_x = x; c->viewport_y = y;
    c->viewport_w = w; c->viewport_h = h;
}

/* ------------------------------------------------------------------------- */
/* matrices (reference src/gl_api.c:480-610)                                  */
/* ------------------------------------------------------------------------- */

static float *top_matrix(GLState *c)
{
    switch (c->matrix_mode) {
    case GL_PROJECTION: return c->projection[c->projection_depth];
    case GL_TEXTURE:    return c->texmat[c->texture_depth];
    default:            return c->modelview[c->modelview_depth];
    }
}

static GLint *top_depth(GLState *c)
{
    switch (c->matrix_mode) {
    case GL_PROJECTION: return &c->projection_depth;
    case GL_TEXTURE:    return &c->texture_depth;
    default:            return &c->modelview_depth;
    }
}

static void mult_top(GLState *c, const float *t)
{
    hm_post_multiply(top_matrix(c), t);
    TL_DIRTY(c);
}

MTGL_EXPORT void glMatrixMode(GLenum mode)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_MATRIX_MODE, { mode } };
    if (gls_list_record(c, &cmd)) return;
    c->matrix_mode = mode;
}

MTGL_EXPORT void glLoadIdentity(void)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_LOAD_IDENTITY };
    if (gls_list_record(c, &cmd)) return;
    hm_identity(top_matrix(c));
    TL_DIRTY(c);
}

MTGL_EXPORT void glPushMatrix(void)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_PUSH_MATRIX };
    if (gls_list_record(c, &cmd)) return;
    GLint *d = top_depth(c);
    if (*d >= GLS_STACK_DEPTH - 1) { gls_set_error(c, GL_STACK_OVERFLOW); return; }
    float copy[16];
    memcpy(copy, top_matrix(c), sizeof(copy));
    (*d)++;
    memcpy(top_matrix(c), copy, sizeof(copy));
    TL_DIRTY(c);
}

MTGL_EXPORT void glPopMatrix(void)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_POP_MATRIX };
    if (gls_list_record(c, &cmd)) return;
    GLint *d = top_depth(c);
    if (*d <= 0) { gls_set_error(c, GL_STACK_UNDERFLOW); return; }
    (*d)--;
    TL_DIRTY(c);
}

MTGL_EXPORT void glOrtho(GLdouble l, GLdouble r, GLdouble b, GLdouble t, GLdouble n, GLdouble f)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_ORTHO };
    cmd.d[0] = l; cmd.d[1] = r; cmd.d[2] = b; cmd.d[3] = t; cmd.d[4] = n; cmd.d[5] = f;
    if (gls_list_record(c, &cmd)) return;
    float m[16];
    hm_ortho(m, (float)l, (float)r, (float)b, (float)t, (float)n, (float)f);
    mult_top(c, m);
}

MTGL_EXPORT void glFrustum(GLdouble l, GLdouble r, GLdouble b, GLdouble t, GLdouble n, GLdouble f)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_FRUSTUM };
    cmd.d[0] = l; cmd.d[1] = r; cmd.d[2] = b; cmd.d[3] = t; cmd.d[4] = n; cmd.d[5] = f;
    if (gls_list_record(c, &cmd)) return;
    float m[16];
    hm_frustum(m, (float)l, (float)r, (float)b, (float)t, (float)n, (float)f);
    mult_top(c, m);
}

MTGL_EXPORT void glTranslatef(GLfloat x, GLfloat y, GLfloat z)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_TRANSLATEF, {0}, { x, y, z } };
    if (gls_list_record(c, &cmd)) return;
    float m[16];
    hm_translate(m, x, y, z);
    mult_top(c, m);
}

MTGL_EXPORT void glRotatef(GLfloat angle, GLfloat x, GLfloat y, GLfloat z)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_ROTATEF, {0}, { angle, x, y, z } };
    if (gls_list_record(c, &cmd)) return;
    float m[16];
    hm_rotate(m, angle, x, y, z);
    mult_top(c, m);
}

MTGL_EXPORT void glScalef(GLfloat x, GLfloat y, GLfloat z)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_SCALEF, {0}, { x, y, z } };
    if (gls_list_record(c, &cmd)) return;
    float m[16];
    hm_scale(m, x, y, z);
    mult_top(c, m);
}

MTGL_EXPORT void glMultMatrixf(const GLfloat *m)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_MULT_MATRIXF };
    memcpy(cmd.f, m, 16 * sizeof(float));
    if (gls_list_record(c, &cmd)) return;
    mult_top(c, m);
}

MTGL_EXPORT void glLoadMatrixf(const GLfloat *m)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_LOAD_MATRIXF };
    memcpy(cmd.f, m, 16 * sizeof(float));
    if (gls_list_record(c, &cmd)) return;
    memcpy(top_matrix(c), m, 16 * sizeof(float));
    TL_DIRTY(c);
}

/* ------------------------------------------------------------------------- */
/* immediate mode (reference src/gl_api.c:614-752, emit_vertex :263-348)      */
/* ------------------------------------------------------------------------- */

MTGL_EXPORT void glBegin(GLenum mode)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_BEGIN, { mode } };
    if (gls_list_record(c, &cmd)) return;
    if (c->flags & GLS_INSIDE_BEGIN_END) { gls_set_error(c, GL_INVALID_OPERATION); return; }
    if (mode > GL_POLYGON) { gls_set_error(c, GL_INVALID_ENUM); return; }
    c->primitive_mode = mode;
    c->flags |= GLS_INSIDE_BEGIN_END;
}

MTGL_EXPORT void glEnd(void)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_END };
    if (gls_list_record(c, &cmd)) return;
    if (!(c->flags & GLS_INSIDE_BEGIN_END)) { gls_set_error(c, GL_INVALID_OPERATION); return; }
    c->flags &= ~GLS_INSIDE_BEGIN_END;
    gls_flush_immediate(c);
}

/* Buffer one vertex.  Like the reference, vertices are buffered even outside
 * glBegin/glEnd and are drawn by the next successful glEnd. */
static void emit_vertex(GLState *c, float x, float y, float z)
{
    if (c->tl_dirty || c->vstate_count == 0) {
        if (c->vstate_count == c->vstate_capacity) {
            size_t cap = c->vstate_capacity ? c->vstate_capacity * 2 : 4;
            mtgl_vertex_state *p = (mtgl_vertex_state *)realloc(c->vstates, cap * sizeof(*p));
            if (!p) { gls_set_error(c, GL_OUT_OF_MEMORY); return; }
            c->vstates = p; c->vstate_capacity = cap;
        }
        gls_snapshot_vertex_state(c, &c->vstates[c->vstate_count++]);
        c->tl_dirty = 0;
    }
    if (c->imm_count == c->imm_capacity) {
        size_t cap = c->imm_capacity ? c->imm_capacity * 2 : 64;
        gls_imm_vertex *p = (gls_imm_vertex *)realloc(c->imm, cap * sizeof(*p));
        if (!p) { gls_set_error(c, GL_OUT_OF_MEMORY); return; }
        c->imm = p; c->imm_capacity = cap;
    }
    gls_imm_vertex *v = &c->imm[c->imm_count++];
    v->pos[0] = x; v->pos[1] = y; v->pos[2] = z;
    v->color[0] = c->current_color.r; v->color[1] = c->current_color.g;
    v->color[2] = c->current_color.b; v->color[3] = c->current_color.a;
    v->texcoord[0] = c->current_texcoord[0]; v->texcoord[1] = c->current_texcoord[1];
    v->normal[0] = c->current_normal[0]; v->normal[1] = c->current_normal[1];
    v->normal[2] = c->current_normal[2];
    v->state = (uint32_t)(c->vstate_count - 1);
    gls_apply_color_material(c, v->color);
}

/* The draw of n vertex records: from display-list segment seg, or (seg 0)
 * the buffered records, themselves kept in a segment when a list block is
 * being replayed.  Resets the vertex buffer as glEnd does. */
void gls_draw_records(GLState *c, size_t n, uint32_t seg)
{
    if (n > 0) {
        mtgl_raster_state rs;
        gls_snapshot_raster_state(c, &rs);
        mtgl_draw d;
        memset(&d, 0, sizeof(d));
        const int32_t stride = (int32_t)sizeof(gls_imm_vertex);
        d.mode = c->primitive_mode;
        d.count = (uint32_t)n;
        d.first = 0;
        d.index_type = 0;
        d.position = (mtgl_array){1, 3, GL_FLOAT, stride, 0, 0, (int64_t)offsetof(gls_imm_vertex, pos)};
        d.color = (mtgl_array){1, 4, GL_FLOAT, stride, 0, 0, (int64_t)offsetof(gls_imm_vertex, color)};
        d.texcoord = (mtgl_array){1, 2, GL_FLOAT, stride, 0, 0, (int64_t)offsetof(gls_imm_vertex, texcoord)};
        d.normal = (mtgl_array){1, 3, GL_FLOAT, stride, 0, 0, (int64_t)offsetof(gls_imm_vertex, normal)};
        if (c->vstate_count > 1)
            d.state_index = (mtgl_array){1, 1, GL_UNSIGNED_INT, stride, 0, 0,
                                         (int64_t)offsetof(gls_imm_vertex, state)};
        d.n_vstates = (uint32_t)c->vstate_count;
        d.vstates = c->vstates;
        d.raster = &rs;
        const size_t bytes = n * sizeof(gls_imm_vertex);
        /* a replayed list block draws from its device segment */
        if (!seg) seg = gls_list_segment(c, c->imm, bytes);
        if (seg) {
            const uint32_t b = MTGL_SEGMENT_BIT | seg;
            d.position.buffer = d.color.buffer = d.texcoord.buffer = d.normal.buffer = b;
            if (c->vstate_count > 1) d.state_index.buffer = b;
        } else {
            d.staged = c->imm;
            d.staged_bytes = (uint64_t)bytes;
        }
        int rc = mtgl_dev_draw(c->dev, &d);
        gls_report_device_error(c, rc, "glEnd");
    }
    c->imm_count = 0;
    c->vstate_count = 0;
    c->tl_dirty = 1;
}

/* glEnd: hand the buffered vertices to the device as one draw whose attributes
 * are the staged vertex records. */
void gls_flush_immediate(GLState *c) { gls_draw_records(c, c->imm_count, 0); }

/* The vertex-state snapshot the first glVertex of a block takes (emit_vertex),
 * for a list block replayed without its calls. */
int gls_begin_snapshot(GLState *c)
{
    if (c->vstate_capacity == 0) {
        mtgl_vertex_state *p = (mtgl_vertex_state *)realloc(c->vstates, 4 * sizeof(*p));
        if (!p) return 0;
        c->vstates = p;
        c->vstate_capacity = 4;
    }
    gls_snapshot_vertex_state(c, &c->vstates[0]);
    c->vstate_count = 1;
    c->tl_dirty = 0;
    return 1;
}

MTGL_EXPORT void glVertex2f(GLfloat x, GLfloat y)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_VERTEX, {0}, { x, y, 0.0f } };
    if (gls_list_record(c, &cmd)) return;
    emit_vertex(c, x, y, 0.0f);
}

MTGL_EXPORT void glVertex3f(GLfloat x, GLfloat y, GLfloat z)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_VERTEX, {0}, { x, y, z } };
    if (gls_list_record(c, &cmd)) return;
    emit_vertex(c, x, y, z);
}

MTGL_EXPORT void glVertex2i(GLint x, GLint y) { glVertex2f((float)x, (float)y); }
MTGL_EXPORT void glVertex3i(GLint x, GLint y, GLint z) { glVertex3f((float)x, (float)y, (float)z); }

static float sanitize(float v, float replacement)
{
    if (isnan(v) || isinf(v)) v = replacement;
    if (v < 0.0f) v = 0.0f; else if (v > 1.0f) v = 1.0f;
    return v;
}

/* NaN/Inf components become 0 (alpha 1), then clamp to [0,1]
 * (reference src/gl_api.c:694-722). */
MTGL_EXPORT void glColor3f(GLfloat r, GLfloat g, GLfloat b)
{
    CTX_OR_RETURN();
    r = sanitize(r, 0.0f); g = sanitize(g, 0.0f); b = sanitize(b, 0.0f);
    gls_list_cmd cmd = { LOP_COLOR, {0}, { r, g, b, 1.0f } };
    if (gls_list_record(c, &cmd)) return;
    c->current_color = (mtgl_color){r, g, b, 1.0f};
}

MTGL_EXPORT void glColor4f(GLfloat r, GLfloat g, GLfloat b, GLfloat a)
{
    CTX_OR_RETURN();
    r = sanitize(r, 0.0f); g = sanitize(g, 0.0f); b = sanitize(b, 0.0f); a = sanitize(a, 1.0f);
    gls_list_cmd cmd = { LOP_COLOR, {0}, { r, g, b, a } };
    if (gls_list_record(c, &cmd)) return;
    c->current_color = (mtgl_color){r, g, b, a};
}

MTGL_EXPORT void glColor3ub(GLubyte r, GLubyte g, GLubyte b)
{
    CTX_OR_RETURN();
    float rf = r / 255.0f, gf = g / 255.0f, bf = b / 255.0f;
    gls_list_cmd cmd = { LOP_COLOR, {0}, { rf, gf, bf, 1.0f } };
    if (gls_list_record(c, &cmd)) return;
    c->current_color = (mtgl_color){rf, gf, bf, 1.0f};
}

MTGL_EXPORT void glColor4ub(GLubyte r, GLubyte g, GLubyte b, GLubyte a)
{
    CTX_OR_RETURN();
    float rf = r / 255.0f, gf = g / 255.0f, bf = b / 255.0f, af = a / 255.0f;
    gls_list_cmd cmd = { LOP_COLOR, {0}, { rf, gf, bf, af } };
    if (gls_list_record(c, &cmd)) return;
    c->current_color = (mtgl_color){rf, gf, bf, af};
}

MTGL_EXPORT void glTexCoord2f(GLfloat s, GLfloat t)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_TEXCOORD, {0}, { s, t } };
    if (gls_list_record(c, &cmd)) return;
    c->current_texcoord[0] = s; c->current_texcoord[1] = t;
}

MTGL_EXPORT void glNormal3f(GLfloat x, GLfloat y, GLfloat z)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_NORMAL, {0}, { x, y, z } };
    if (gls_list_record(c, &cmd)) return;
    c->current_normal[0] = x; c->current_normal[1] = y; c->current_normal[2] = z;
}

/* ------------------------------------------------------------------------- */
/* textures (reference src/gl_api.c:756-885, src/textures.c:58-269)           */
/* ------------------------------------------------------------------------- */

static gls_texture *texture_lookup(GLState *c, GLuint id)
{
    if (id == 0 || id > c->texture_count) return NULL;
    gls_texture *t = &c->textures[id - 1];
    return t->allocated ? t : NULL;
}

static void texture_reset(gls_texture *t)
{
    t->allocated = 1;
    t->width = t->height = 0;
    t->min_filter = GL_NEAREST;
    t->mag_filter = GL_NEAREST;
    t->wrap_s = GL_REPEAT;
    t->wrap_t = GL_REPEAT;
}

MTGL_EXPORT void glGenTextures(GLsizei n, GLuint *textures)
{
    CTX_OR_RETURN();
    if (n < 0) { gls_set_error(c, GL_INVALID_VALUE); return; }
    for (GLsizei i = 0; i < n; i++) {
        GLuint id = 0;
        for (size_t k = 0; k < c->texture_count; k++) {
            if (!c->textures[k].allocated) { texture_reset(&c->textures[k]); id = (GLuint)(k + 1); break; }
        }
        if (!id && c->texture_count < MTGL_MAX_OBJECTS) {
            texture_reset(&c->textures[c->texture_count]);
            id = (GLuint)++c->texture_count;
        }
        textures[i] = id;
    }
}

MTGL_EXPORT void glDeleteTextures(GLsizei n, const GLuint *textures)
{
    CTX_OR_RETURN();
    if (n < 0) { gls_set_error(c, GL_INVALID_VALUE); return; }
    for (GLsizei i = 0; i < n; i++) {
        GLuint id = textures[i];
        if (id == c->bound_texture) c->bound_texture = 0;
        gls_texture *t = texture_lookup(c, id);
        if (!t) continue;
        if (t->width > 0) gls_report_device_error(c, mtgl_dev_tex_free(c->dev, id), "glDeleteTextures");
        t->width = t->height = 0;
        t->allocated = 0;
    }
}

MTGL_EXPORT void glBindTexture(GLenum target, GLuint texture)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_BIND_TEXTURE, { target, texture } };
    if (gls_list_record(c, &cmd)) return;
    if (target != GL_TEXTURE_2D) { gls_set_error(c, GL_INVALID_ENUM); return; }
    c->bound_texture = texture;
}

/* Level 0 only, GL_UNSIGNED_BYTE only; formats are expanded to packed RGBA8 as in
 * the reference upload routines (src/textures.c:141-269); invalid sizes leave
 * the texture untouched without an error. */
MTGL_EXPORT void glTexImage2D(GLenum target, GLint level, GLint internalformat, GLsizei width, GLsizei height,
                              GLint border, GLenum format, GLenum type, const GLvoid *pixels)
{
    CTX_OR_RETURN();
    (void)internalformat; (void)border;
    if (target != GL_TEXTURE_2D) { gls_set_error(c, GL_INVALID_ENUM); return; }
    if (level != 0 || type != GL_UNSIGNED_BYTE) { gls_set_error(c, GL_INVALID_VALUE); return; }
    if (width < 0 || height < 0) { gls_set_error(c, GL_INVALID_VALUE); return; }
    gls_texture *t = texture_lookup(c, c->bound_texture);
    if (!t) return;
    int bpp;
    switch (format) {
    case GL_RGBA: bpp = 4; break;
    case GL_RGB: bpp = 3; break;
    case GL_LUMINANCE: bpp = 1; break;
    case GL_LUMINANCE_ALPHA: bpp = 2; break;
    default: gls_set_error(c, GL_INVALID_ENUM); return;
    }
    if (width <= 0 || height <= 0 || width > MTGL_MAX_TEX_SIZE || height > MTGL_MAX_TEX_SIZE) return;
    size_t count = (size_t)width * (size_t)height;
    uint32_t *texels = (uint32_t *)malloc(count * sizeof(uint32_t));
    if (!texels) { gls_set_error(c, GL_OUT_OF_MEMORY); return; }
    const uint8_t *p = (const uint8_t *)pixels;
    for (size_t i = 0; i < count; i++) {
        uint32_t r, g, b, a;
        if (!p) { r = g = b = a = 0; }
        else if (bpp == 4) { r = p[4 * i]; g = p[4 * i + 1]; b = p[4 * i + 2]; a = p[4 * i + 3]; }
        else if (bpp == 3) { r = p[3 * i]; g = p[3 * i + 1]; b = p[3 * i + 2]; a = 255; }
        else if (bpp == 1) { r = g = b = p[i]; a = 255; }
        else { r = g = b = p[2 * i]; a = p[2 * i + 1]; }
        texels[i] = (a << 24) | (b << 16) | (g << 8) | r;
    }
    int rc = mtgl_dev_tex_image(c->dev, c->bound_texture, width, height, texels);
    free(texels);
    if (rc != MTGL_OK) { gls_report_device_error(c, rc, "glTexImage2D"); return; }
    t->width = width;
    t->height = height;
}

MTGL_EXPORT void glTexParameteri(GLenum target, GLenum pname, GLint param)
{
    CTX_OR_RETURN();
    if (target != GL_TEXTURE_2D) { gls_set_error(c, GL_INVALID_ENUM); return; }
    gls_texture *t = texture_lookup(c, c->bound_texture);
    if (!t) return;
    switch (pname) {
    case GL_TEXTURE_MIN_FILTER:
        if (param != GL_NEAREST && param != GL_LINEAR && param != GL_NEAREST_MIPMAP_NEAREST &&
            param != GL_LINEAR_MIPMAP_NEAREST && param != GL_NEAREST_MIPMAP_LINEAR &&
            param != GL_LINEAR_MIPMAP_LINEAR) { gls_set_error(c, GL_INVALID_ENUM); return; }
        t->min_filter = (GLenum)param;
        break;
    case GL_TEXTURE_MAG_FILTER:
        if (param != GL_NEAREST && param != GL_LINEAR) { gls_set_error(c, GL_INVALID_ENUM); return; }
        t->mag_filter = (GLenum)param;
        break;
    case GL_TEXTURE_WRAP_S:
    case GL_TEXTURE_WRAP_T:
        if (param != GL_REPEAT && param != GL_CLAMP && param != GL_CLAMP_TO_EDGE) {
            gls_set_error(c, GL_INVALID_ENUM); return;
        }
        if (pname == GL_TEXTURE_WRAP_S) t->wrap_s = (GLenum)param; else t->wrap_t = (GLenum)param;
        break;
    default:
        gls_set_error(c, GL_INVALID_ENUM);
    }
}

MTGL_EXPORT GLboolean glIsTexture(GLuint texture)
{
    CTX_OR_RETURN_V(GL_FALSE);
    if (texture == 0) return GL_FALSE;
    return texture_lookup(c, texture) ? GL_TRUE : GL_FALSE;
}

MTGL_EXPORT void glTexEnvi(GLenum target, GLenum pname, GLint param)
{
    CTX_OR_RETURN();
    if (target != GL_TEXTURE_ENV || pname != GL_TEXTURE_ENV_MODE) { gls_set_error(c, GL_INVALID_ENUM); return; }
    switch (param) {
    case GL_MODULATE: case GL_DECAL: case GL_REPLACE: case GL_BLEND: case GL_ADD:
        c->tex_env_mode = (GLenum)param;
        break;
    default:
        gls_set_error(c, GL_INVALID_ENUM);
    }
}

MTGL_EXPORT void glTexEnvf(GLenum target, GLenum pname, GLfloat param)
{
    glTexEnvi(target, pname, (GLint)param);
}

MTGL_EXPORT void glTexEnvfv(GLenum target, GLenum pname, const GLfloat *params)
{
    CTX_OR_RETURN();
    if (!params) return;
    if (target != GL_TEXTURE_ENV) { gls_set_error(c, GL_INVALID_ENUM); return; }
    if (pname == GL_TEXTURE_ENV_MODE) glTexEnvi(target, pname, (GLint)params[0]);
    else if (pname == GL_TEXTURE_ENV_COLOR) c->tex_env_color = (mtgl_color){params[0], params[1], params[2], params[3]};
    else gls_set_error(c, GL_INVALID_ENUM);
}

/* ------------------------------------------------------------------------- */
/* misc state (reference src/gl_api.c:887-1176)                              */
/* ------------------------------------------------------------------------- */

MTGL_EXPORT void glFlush(void)
{
    CTX_OR_RETURN();
    (void)c;
}

MTGL_EXPORT void glFinish(void)
{
    CTX_OR_RETURN();
    gls_report_device_error(c, mtgl_dev_finish(c->dev), "glFinish");
}

MTGL_EXPORT GLenum glGetError(void)
{
    GLState *c = g_ctx;
    if (!c) return GL_NO_ERROR;
    GLenum e = c->error;
    c->error = GL_NO_ERROR;
    return e;
}

static int blend_factor_ok(GLenum f, int is_src)
{
    switch (f) {
    case GL_ZERO: case GL_ONE: case GL_SRC_COLOR: case GL_ONE_MINUS_SRC_COLOR:
    case GL_DST_COLOR: case GL_ONE_MINUS_DST_COLOR: case GL_SRC_ALPHA:
    case GL_ONE_MINUS_SRC_ALPHA: case GL_DST_ALPHA: case GL_ONE_MINUS_DST_ALPHA:
    case GL_CONSTANT_COLOR: case GL_ONE_MINUS_CONSTANT_COLOR: case GL_CONSTANT_ALPHA:
    case GL_ONE_MINUS_CONSTANT_ALPHA:
        return 1;
    case GL_SRC_ALPHA_SATURATE:
        return is_src;
    default:
        return 0;
    }
}

MTGL_EXPORT void glBlendFunc(GLenum sfactor, GLenum dfactor)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_BLEND_FUNC, { sfactor, dfactor } };
    if (gls_list_record(c, &cmd)) return;
    if (!blend_factor_ok(sfactor, 1) || !blend_factor_ok(dfactor, 0)) { gls_set_error(c, GL_INVALID_ENUM); return; }
    c->blend_src = sfactor;
    c->blend_dst = dfactor;
}

MTGL_EXPORT void glCullFace(GLenum mode)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_CULL_FACE, { mode } };
    if (gls_list_record(c, &cmd)) return;
    if (mode != GL_FRONT && mode != GL_BACK && mode != GL_FRONT_AND_BACK) { gls_set_error(c, GL_INVALID_ENUM); return; }
    c->cull_face_mode = mode;
}

MTGL_EXPORT void glFrontFace(GLenum mode)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_FRONT_FACE, { mode } };
    if (gls_list_record(c, &cmd)) return;
    if (mode != GL_CW && mode != GL_CCW) { gls_set_error(c, GL_INVALID_ENUM); return; }
    c->front_face = mode;
}

MTGL_EXPORT void glDepthFunc(GLenum func)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_DEPTH_FUNC, { func } };
    if (gls_list_record(c, &cmd)) return;
    if (func < GL_NEVER || func > GL_ALWAYS) { gls_set_error(c, GL_INVALID_ENUM); return; }
    c->depth_func = func;
}

MTGL_EXPORT void glDepthMask(GLboolean flag)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_DEPTH_MASK, { flag } };
    if (gls_list_record(c, &cmd)) return;
    c->depth_mask = flag;
}

MTGL_EXPORT void glAlphaFunc(GLenum func, GLclampf ref)
{
    CTX_OR_RETURN();
    if (func < GL_NEVER || func > GL_ALWAYS) { gls_set_error(c, GL_INVALID_ENUM); return; }
    c->alpha_func = func;
    if (ref < 0.0f) ref = 0.0f;
    if (ref > 1.0f) ref = 1.0f;
    c->alpha_ref = ref;
}

MTGL_EXPORT void glLineWidth(GLfloat width)
{
    CTX_OR_RETURN();
    if (width <= 0.0f || isnan(width) || isinf(width)) { gls_set_error(c, GL_INVALID_VALUE); return; }
    c->line_width = width;
}

MTGL_EXPORT void glPointSize(GLfloat size)
{
    CTX_OR_RETURN();
    if (size <= 0.0f || isnan(size) || isinf(size)) { gls_set_error(c, GL_INVALID_VALUE); return; }
    c->point_size = size;
}

MTGL_EXPORT void glPolygonMode(GLenum face, GLenum mode)
{
    CTX_OR_RETURN();
    if (face != GL_FRONT && face != GL_BACK && face != GL_FRONT_AND_BACK) { gls_set_error(c, GL_INVALID_ENUM); return; }
    if (mode != GL_POINT && mode != GL_LINE && mode != GL_FILL) { gls_set_error(c, GL_INVALID_ENUM); return; }
    if (face == GL_FRONT || face == GL_FRONT_AND_BACK) c->polygon_mode_front = mode;
    if (face == GL_BACK || face == GL_FRONT_AND_BACK) c->polygon_mode_back = mode;
}

MTGL_EXPORT void glScissor(GLint x, GLint y, GLsizei w, GLsizei h)
{
    CTX_OR_RETURN();
    if (w < 0 || h < 0) { gls_set_error(c, GL_INVALID_VALUE); return; }
    c->scissor_x = x; c->scissor_y = y; c->scissor_w = w; c->scissor_h = h;
}

MTGL_EXPORT void glDepthRange(GLclampd n, GLclampd f)
{
    CTX_OR_RETURN();
    if (n < 0.0) n = 0.0;
    if (n > 1.0) n = 1.0;
    if (f < 0.0) f = 0.0;
    if (f > 1.0) f = 1.0;
    c->depth_near = n;
    c->depth_far = f;
}

MTGL_EXPORT void glColorMask(GLboolean r, GLboolean g, GLboolean b, GLboolean a)
{
    CTX_OR_RETURN();
    c->color_mask[0] = r; c->color_mask[1] = g; c->color_mask[2] = b; c->color_mask[3] = a;
}

static int stencil_func_ok(GLenum f) { return f >= GL_NEVER && f <= GL_ALWAYS; }

static int stencil_op_ok(GLenum op)
{
    switch (op) {
    case GL_KEEP: case GL_ZERO: case GL_REPLACE: case GL_INCR: case GL_INCR_WRAP:
    case GL_DECR: case GL_DECR_WRAP: case GL_INVERT:
        return 1;
    default:
        return 0;
    }
}

MTGL_EXPORT void glStencilFunc(GLenum func, GLint ref, GLuint mask)
{
    CTX_OR_RETURN();
    if (!stencil_func_ok(func)) { gls_set_error(c, GL_INVALID_ENUM); return; }
    c->stencil_func = func;
    c->stencil_ref = ref;
    c->stencil_mask = mask;
}

MTGL_EXPORT void glStencilOp(GLenum sfail, GLenum dpfail, GLenum dppass)
{
    CTX_OR_RETURN();
    if (!stencil_op_ok(sfail) || !stencil_op_ok(dpfail) || !stencil_op_ok(dppass)) {
        gls_set_error(c, GL_INVALID_ENUM); return;
    }
    c->stencil_fail = sfail;
    c->stencil_zfail = dpfail;
    c->stencil_zpass = dppass;
}

MTGL_EXPORT void glStencilMask(GLuint mask)
{
    CTX_OR_RETURN();
    c->stencil_writemask = mask;
}

MTGL_EXPORT void glPixelStorei(GLenum pname, GLint param)
{
    CTX_OR_RETURN();
    if (pname == GL_PACK_ALIGNMENT || pname == GL_UNPACK_ALIGNMENT) {
        if (param != 1 && param != 2 && param != 4 && param != 8) gls_set_error(c, GL_INVALID_VALUE);
    } else {
        gls_set_error(c, GL_INVALID_ENUM);
    }
}

MTGL_EXPORT void glHint(GLenum target, GLenum mode)
{
    CTX_OR_RETURN();
    if (mode != GL_DONT_CARE && mode != GL_FASTEST && mode != GL_NICEST) { gls_set_error(c, GL_INVALID_ENUM); return; }
    switch (target) {
    case GL_PERSPECTIVE_CORRECTION_HINT: c->perspective_hint = mode; break;
    case GL_POINT_SMOOTH_HINT: case GL_LINE_SMOOTH_HINT: case GL_FOG_HINT: break;
    default: gls_set_error(c, GL_INVALID_ENUM);
    }
}

/* ------------------------------------------------------------------------- */
/* pixel transfer (reference src/gl_api.c:1180-1424)                          */
/* ------------------------------------------------------------------------- */

/* Rows are returned bottom-up (GL window convention); rows outside the
 * framebuffer read as zero bytes, pixels beside it in rows that exist as
 * (0, 0, 0, 255).  The device packs and flips the window and copies it to
 * `pixels` in one transfer (mtgl_dev_read_pixels).  Negative sizes write
 * nothing (the reference's memset would take them as huge sizes). */
MTGL_EXPORT void glReadPixels(GLint x, GLint y, GLsizei width, GLsizei height, GLenum format, GLenum type, GLvoid *pixels)
{
    CTX_OR_RETURN();
    if (type != GL_UNSIGNED_BYTE || !pixels) return;
    if (width <= 0 || height <= 0 || (format != GL_RGBA && format != GL_RGB)) return;
    int rc = mtgl_dev_read_pixels(c->dev, x, y, width, height, format, pixels);
    if (rc != MTGL_OK) gls_report_device_error(c, rc, "glReadPixels");
}

/* Pixels at the raster position with alpha test, depth test and blending
 * (reference src/gl_api.c:1284-1373).  Formats other than the four the
 * reference reads draw nothing, as there. */
MTGL_EXPORT void glDrawPixels(GLsizei width, GLsizei height, GLenum format, GLenum type, const GLvoid *pixels)
{
    CTX_OR_RETURN();
    if (type != GL_UNSIGNED_BYTE || !pixels || !c->raster_valid) return;
    if (width <= 0 || height <= 0) return;
    if (format != GL_RGBA && format != GL_RGB && format != GL_LUMINANCE && format != GL_LUMINANCE_ALPHA) return;
    mtgl_pixels p;
    memset(&p, 0, sizeof(p));
    p.x = c->raster_x;
    p.y = c->raster_y;
    p.width = width;
    p.height = height;
    p.format = format;
    p.flags = c->flags & (MTGL_F_ALPHA_TEST | MTGL_F_DEPTH_TEST | MTGL_F_BLEND);
    p.alpha_func = c->alpha_func;
    p.alpha_ref = c->alpha_ref;
    p.depth_func = c->depth_func;
    p.depth_mask = c->depth_mask ? 1u : 0u;
    p.blend_src = c->blend_src;
    p.blend_dst = c->blend_dst;
    p.pixels = pixels;
    gls_report_device_error(c, mtgl_dev_draw_pixels(c->dev, &p), "glDrawPixels");
}

/* window position of a raster position (reference src/gl_api.c:1375-1424) */
static void set_raster_pos(GLState *c, float x, float y, float z)
{
    float v[4] = { x, y, z, 1.0f }, eye[4], clip[4];
    hm_mul_vec4(c->modelview[c->modelview_depth], v, eye);
    hm_mul_vec4(c->projection[c->projection_depth], eye, clip);
    if (clip[3] <= 0.0f) { c->raster_valid = GL_FALSE; return; }
    float nx = clip[0] / clip[3];
    float ny = clip[1] / clip[3];
    float fx = (nx + 1.0f) * 0.5f * (float)c->viewport_w + (float)c->viewport_x;
    float fy = (1.0f - ny) * 0.5f * (float)c->viewport_h + (float)c->viewport_y;
    int32_t sx = (fx >= -2147483648.0f && fx < 2147483648.0f) ? (int32_t)fx : INT32_MIN;
    int32_t sy = (fy >= -2147483648.0f && fy < 2147483648.0f) ? (int32_t)fy : INT32_MIN;
    c->raster_x = sx;
    c->raster_y = c->height - 1 - sy;
    c->raster_valid = GL_TRUE;
}

MTGL_EXPORT void glRasterPos2i(GLint x, GLint y)
{
    CTX_OR_RETURN();
    set_raster_pos(c, (float)x, (float)y, 0.0f);
}

MTGL_EXPORT void glRasterPos2f(GLfloat x, GLfloat y) { glRasterPos3f(x, y, 0.0f); }

MTGL_EXPORT void glRasterPos3f(GLfloat x, GLfloat y, GLfloat z)
{
    CTX_OR_RETURN();
    set_raster_pos(c, x, y, z);
}

/* ------------------------------------------------------------------------- */
/* fog (reference src/gl_api.c:1426-1512)                                     */
/* ------------------------------------------------------------------------- */

static int fog_mode_ok(GLenum m) { return m == GL_LINEAR || m == GL_EXP || m == GL_EXP2; }

static void fog_param(GLState *c, GLenum pname, const GLfloat *p, int has_color)
{
    switch (pname) {
    case GL_FOG_MODE:
        if (!fog_mode_ok((GLenum)(int)p[0])) { gls_set_error(c, GL_INVALID_ENUM); return; }
        c->fog_mode = (GLenum)(int)p[0];
        break;
    case GL_FOG_DENSITY:
        if (p[0] < 0.0f) { gls_set_error(c, GL_INVALID_VALUE); return; }
        c->fog_density = p[0];
        break;
    case GL_FOG_START: c->fog_start = p[0]; break;
    case GL_FOG_END: c->fog_end = p[0]; break;
    case GL_FOG_COLOR:
        if (has_color) { c->fog_color = (mtgl_color){p[0], p[1], p[2], p[3]}; break; }
        /* fall through */
    default:
        gls_set_error(c, GL_INVALID_ENUM);
    }
}

MTGL_EXPORT void glFogi(GLenum pname, GLint param)
{
    CTX_OR_RETURN();
    if (pname != GL_FOG_MODE) { gls_set_error(c, GL_INVALID_ENUM); return; }
    if (!fog_mode_ok((GLenum)param)) { gls_set_error(c, GL_INVALID_ENUM); return; }
    c->fog_mode = (GLenum)param;
}

MTGL_EXPORT void glFogf(GLenum pname, GLfloat param)
{
    CTX_OR_RETURN();
    GLfloat p[4] = { param, param, param, param };
    fog_param(c, pname, p, 0);
}

MTGL_EXPORT void glFogfv(GLenum pname, const GLfloat *params)
{
    CTX_OR_RETURN();
    fog_param(c, pname, params, 1);
}

/* ------------------------------------------------------------------------- */
/* lighting (reference src/gl_api.c:1943-2262)                                */
/* ------------------------------------------------------------------------- */

/* values a light, material or light-model parameter takes: the caller's array
 * holds that many, no more (an unknown name reads one, then reports the error) */
static int light_param_count(GLenum pname)
{
    switch (pname) {
    case GL_AMBIENT: case GL_DIFFUSE: case GL_SPECULAR: case GL_POSITION: return 4;
    case GL_SPOT_DIRECTION: return 3;
    default: return 1;
    }
}

static int material_param_count(GLenum pname)
{
    switch (pname) {
    case GL_AMBIENT: case GL_DIFFUSE: case GL_SPECULAR: case GL_EMISSION: case GL_AMBIENT_AND_DIFFUSE: return 4;
    default: return 1;
    }
}

static int light_model_param_count(GLenum pname) { return pname == GL_LIGHT_MODEL_AMBIENT ? 4 : 1; }

MTGL_EXPORT void glLightfv(GLenum light, GLenum pname, const GLfloat *params)
{
    CTX_OR_RETURN();
    GLfloat p[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
    memcpy(p, params, (size_t)light_param_count(pname) * sizeof(float));
    gls_list_cmd cmd = { LOP_LIGHTFV, { light, pname } };
    memcpy(cmd.f, p, 4 * sizeof(float));
    if (gls_list_record(c, &cmd)) return;
    if (light < GL_LIGHT0 || light > GL_LIGHT7) { gls_set_error(c, GL_INVALID_ENUM); return; }
    mtgl_light *l = &c->lights[light - GL_LIGHT0];
    TL_DIRTY(c);
    switch (pname) {
    case GL_AMBIENT:  l->ambient = (mtgl_color){p[0], p[1], p[2], p[3]}; break;
    case GL_DIFFUSE:  l->diffuse = (mtgl_color){p[0], p[1], p[2], p[3]}; break;
    case GL_SPECULAR: l->specular = (mtgl_color){p[0], p[1], p[2], p[3]}; break;
    case GL_POSITION: {
        /* stored in eye space: transformed by the modelview current at this call */
        float v[4] = { p[0], p[1], p[2], p[3] };
        hm_mul_vec4(c->modelview[c->modelview_depth], v, l->position);
        break;
    }
    case GL_SPOT_DIRECTION: {
        float v[4] = { p[0], p[1], p[2], 0.0f }, out[4];
        hm_mul_vec4(c->modelview[c->modelview_depth], v, out);
        l->spot_direction[0] = out[0]; l->spot_direction[1] = out[1]; l->spot_direction[2] = out[2];
        break;
    }
    case GL_SPOT_EXPONENT: l->spot_exponent = p[0]; break;
    case GL_SPOT_CUTOFF:
        l->spot_cutoff = p[0];
        l->spot_cos_cutoff = cosf(p[0] * 3.14159265f / 180.0f);
        break;
    case GL_CONSTANT_ATTENUATION:  l->constant_attenuation = p[0]; break;
    case GL_LINEAR_ATTENUATION:    l->linear_attenuation = p[0]; break;
    case GL_QUADRATIC_ATTENUATION: l->quadratic_attenuation = p[0]; break;
    default: break;
    }
}

MTGL_EXPORT void glLightf(GLenum light, GLenum pname, GLfloat param)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_LIGHTF, { light, pname }, { param } };
    if (gls_list_record(c, &cmd)) return;
    GLfloat p[4] = { param, param, param, param };
    glLightfv(light, pname, p);
}

MTGL_EXPORT void glLighti(GLenum light, GLenum pname, GLint param) { glLightf(light, pname, (GLfloat)param); }

MTGL_EXPORT void glLightiv(GLenum light, GLenum pname, const GLint *params)
{
    GLfloat p[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
    for (int i = 0; i < light_param_count(pname); i++) p[i] = (GLfloat)params[i];
    glLightfv(light, pname, p);
}

MTGL_EXPORT void glMaterialfv(GLenum face, GLenum pname, const GLfloat *params)
{
    CTX_OR_RETURN();
    GLfloat p[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
    memcpy(p, params, (size_t)material_param_count(pname) * sizeof(float));
    gls_list_cmd cmd = { LOP_MATERIALFV, { face, pname } };
    memcpy(cmd.f, p, 4 * sizeof(float));
    if (gls_list_record(c, &cmd)) return;
    if (face != GL_FRONT && face != GL_BACK && face != GL_FRONT_AND_BACK) { gls_set_error(c, GL_INVALID_ENUM); return; }
    mtgl_material *mats[2] = {
        (face == GL_FRONT || face == GL_FRONT_AND_BACK) ? &c->material_front : NULL,
        (face == GL_BACK || face == GL_FRONT_AND_BACK) ? &c->material_back : NULL };
    mtgl_color k = { p[0], p[1], p[2], p[3] };
    TL_DIRTY(c);
    for (int i = 0; i < 2; i++) {
        mtgl_material *m = mats[i];
        if (!m) continue;
        switch (pname) {
        case GL_AMBIENT:   m->ambient = k; break;
        case GL_DIFFUSE:   m->diffuse = k; break;
        case GL_SPECULAR:  m->specular = k; break;
        case GL_EMISSION:  m->emission = k; break;
        case GL_SHININESS: m->shininess = p[0]; break;
        case GL_AMBIENT_AND_DIFFUSE: m->ambient = k; m->diffuse = k; break;
        default: gls_set_error(c, GL_INVALID_ENUM); return;
        }
    }
}

MTGL_EXPORT void glMaterialf(GLenum face, GLenum pname, GLfloat param)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_MATERIALF, { face, pname }, { param } };
    if (gls_list_record(c, &cmd)) return;
    GLfloat p[4] = { param, param, param, param };
    glMaterialfv(face, pname, p);
}

MTGL_EXPORT void glMateriali(GLenum face, GLenum pname, GLint param) { glMaterialf(face, pname, (GLfloat)param); }

MTGL_EXPORT void glMaterialiv(GLenum face, GLenum pname, const GLint *params)
{
    GLfloat p[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
    for (int i = 0; i < material_param_count(pname); i++) p[i] = (GLfloat)params[i];
    glMaterialfv(face, pname, p);
}

MTGL_EXPORT void glLightModelfv(GLenum pname, const GLfloat *p)
{
    CTX_OR_RETURN();
    TL_DIRTY(c);
    switch (pname) {
    case GL_LIGHT_MODEL_AMBIENT: c->light_model_ambient = (mtgl_color){p[0], p[1], p[2], p[3]}; break;
    case GL_LIGHT_MODEL_LOCAL_VIEWER: c->light_model_local_viewer = (GLboolean)(p[0] != 0.0f); break;
    case GL_LIGHT_MODEL_TWO_SIDE: c->light_model_two_side = (GLboolean)(p[0] != 0.0f); break;
    default: break;
    }
}

MTGL_EXPORT void glLightModelf(GLenum pname, GLfloat param)
{
    GLfloat p[4] = { param, param, param, param };
    glLightModelfv(pname, p);
}

MTGL_EXPORT void glLightModeli(GLenum pname, GLint param) { glLightModelf(pname, (GLfloat)param); }

MTGL_EXPORT void glLightModeliv(GLenum pname, const GLint *params)
{
    GLfloat p[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
    for (int i = 0; i < light_model_param_count(pname); i++) p[i] = (GLfloat)params[i];
    glLightModelfv(pname, p);
}

MTGL_EXPORT void glColorMaterial(GLenum face, GLenum mode)
{
    CTX_OR_RETURN();
    if (face != GL_FRONT && face != GL_BACK && face != GL_FRONT_AND_BACK) { gls_set_error(c, GL_INVALID_ENUM); return; }
    if (mode != GL_EMISSION && mode != GL_AMBIENT && mode != GL_DIFFUSE && mode != GL_SPECULAR &&
        mode != GL_AMBIENT_AND_DIFFUSE) { gls_set_error(c, GL_INVALID_ENUM); return; }
    c->color_material_face = face;
    c->color_material_mode = mode;
    TL_DIRTY(c);
}

MTGL_EXPORT void glShadeModel(GLenum mode)
{
    CTX_OR_RETURN();
    gls_list_cmd cmd = { LOP_SHADE_MODEL, { mode } };
    if (gls_list_record(c, &cmd)) return;
    if (mode != GL_FLAT && mode != GL_SMOOTH && mode != GL_PHONG) { gls_set_error(c, GL_INVALID_ENUM); return; }
    c->shade_model = mode;
    TL_DIRTY(c);
}

static void color4_out(GLfloat *p, mtgl_color k) { p[0] = k.r; p[1] = k.g; p[2] = k.b; p[3] = k.a; }

MTGL_EXPORT void glGetLightfv(GLenum light, GLenum pname, GLfloat *p)
{
    CTX_OR_RETURN();
    if (!p) return;
    if (light < GL_LIGHT0 || light > GL_LIGHT7) { gls_set_error(c, GL_INVALID_ENUM); return; }
    mtgl_light *l = &c->lights[light - GL_LIGHT0];
    switch (pname) {
    case GL_AMBIENT: color4_out(p, l->ambient); break;
    case GL_DIFFUSE: color4_out(p, l->diffuse); break;
    case GL_SPECULAR: color4_out(p, l->specular); break;
    case GL_POSITION: memcpy(p, l->position, 4 * sizeof(float)); break;
    case GL_SPOT_DIRECTION: memcpy(p, l->spot_direction, 3 * sizeof(float)); break;
    case GL_SPOT_EXPONENT: p[0] = l->spot_exponent; break;
    case GL_SPOT_CUTOFF: p[0] = l->spot_cutoff; break;
    case GL_CONSTANT_ATTENUATION: p[0] = l->constant_attenuation; break;
    case GL_LINEAR_ATTENUATION: p[0] = l->linear_attenuation; break;
    case GL_QUADRATIC_ATTENUATION: p[0] = l->quadratic_attenuation; break;
    default: gls_set_error(c, GL_INVALID_ENUM);
    }
}

MTGL_EXPORT void glGetMaterialfv(GLenum face, GLenum pname, GLfloat *p)
{
    CTX_OR_RETURN();
    if (!p) return;
    mtgl_material *m;
    if (face == GL_FRONT) m = &c->material_front;
    else if (face == GL_BACK) m = &c->material_back;
    else { gls_set_error(c, GL_INVALID_ENUM); return; }
    switch (pname) {
    case GL_AMBIENT: color4_out(p, m->ambient); break;
    case GL_DIFFUSE: color4_out(p, m->diffuse); break;
    case GL_SPECULAR: color4_out(p, m->specular); break;
    case GL_EMISSION: color4_out(p, m->emission); break;
    case GL_SHININESS: p[0] = m->shininess; break;
    default: gls_set_error(c, GL_INVALID_ENUM);
    }
}

/* ------------------------------------------------------------------------- */
/* queries (reference src/gl_api.c:2505-3189)                                 */
/* ------------------------------------------------------------------------- */

MTGL_EXPORT void glGetIntegerv(GLenum pname, GLint *p)
{
    CTX_OR_RETURN();
    if (!p) return;
    switch (pname) {
    case GL_VIEWPORT: p[0] = c->viewport_x; p[1] = c->viewport_y; p[2] = c->viewport_w; p[3] = c->viewport_h; break;
    case GL_MATRIX_MODE: p[0] = (GLint)c->matrix_mode; break;
    case GL_MODELVIEW_STACK_DEPTH: p[0] = c->modelview_depth + 1; break;
    case GL_PROJECTION_STACK_DEPTH: p[0] = c->projection_depth + 1; break;
    case GL_TEXTURE_STACK_DEPTH: p[0] = c->texture_depth + 1; break;
    case GL_SHADE_MODEL: p[0] = (GLint)c->shade_model; break;
    case GL_COLOR_MATERIAL_FACE: p[0] = (GLint)c->color_material_face; break;
    case GL_COLOR_MATERIAL_PARAMETER: p[0] = (GLint)c->color_material_mode; break;
    case GL_FOG_MODE: p[0] = (GLint)c->fog_mode; break;
    case GL_LIGHT_MODEL_LOCAL_VIEWER: p[0] = c->light_model_local_viewer; break;
    case GL_LIGHT_MODEL_TWO_SIDE: p[0] = c->light_model_two_side; break;
    case GL_CULL_FACE_MODE: p[0] = (GLint)c->cull_face_mode; break;
    case GL_FRONT_FACE: p[0] = (GLint)c->front_face; break;
    case GL_POLYGON_MODE: p[0] = (GLint)c->polygon_mode_front; p[1] = (GLint)c->polygon_mode_back; break;
    case GL_TEXTURE_BINDING_2D: p[0] = (GLint)c->bound_texture; break;
    case GL_TEXTURE_ENV_MODE: p[0] = (GLint)c->tex_env_mode; break;
    case GL_SCISSOR_BOX: p[0] = c->scissor_x; p[1] = c->scissor_y; p[2] = c->scissor_w; p[3] = c->scissor_h; break;
    case GL_ALPHA_TEST_FUNC: p[0] = (GLint)c->alpha_func; break;
    case GL_STENCIL_FUNC: p[0] = (GLint)c->stencil_func; break;
    case GL_STENCIL_VALUE_MASK: p[0] = (GLint)c->stencil_mask; break;
    case GL_STENCIL_REF: p[0] = c->stencil_ref; break;
    case GL_STENCIL_FAIL: p[0] = (GLint)c->stencil_fail; break;
    case GL_STENCIL_PASS_DEPTH_FAIL: p[0] = (GLint)c->stencil_zfail; break;
    case GL_STENCIL_PASS_DEPTH_PASS: p[0] = (GLint)c->stencil_zpass; break;
    case GL_STENCIL_WRITEMASK: p[0] = (GLint)c->stencil_writemask; break;
    case GL_STENCIL_CLEAR_VALUE: p[0] = c->stencil_clear; break;
    case GL_DEPTH_FUNC: p[0] = (GLint)c->depth_func; break;
    case GL_DEPTH_WRITEMASK: p[0] = c->depth_mask; break;
    case GL_BLEND_SRC: p[0] = (GLint)c->blend_src; break;
    case GL_BLEND_DST: p[0] = (GLint)c->blend_dst; break;
    case GL_COLOR_WRITEMASK: for (int i = 0; i < 4; i++) p[i] = c->color_mask[i]; break;
    case GL_UNPACK_ALIGNMENT: case GL_PACK_ALIGNMENT: p[0] = 4; break;
    case GL_PERSPECTIVE_CORRECTION_HINT: p[0] = (GLint)c->perspective_hint; break;
    case GL_FOG_HINT: case GL_LINE_SMOOTH_HINT: case GL_POINT_SMOOTH_HINT: case GL_POLYGON_SMOOTH_HINT:
        p[0] = GL_DONT_CARE; break;
    case GL_ARRAY_BUFFER_BINDING: p[0] = (GLint)c->bound_array_buffer; break;
    case GL_ELEMENT_ARRAY_BUFFER_BINDING: p[0] = (GLint)c->bound_element_buffer; break;
    case GL_VERTEX_ARRAY_SIZE: p[0] = c->vertex_ptr.size; break;
    case GL_VERTEX_ARRAY_TYPE: p[0] = (GLint)c->vertex_ptr.type; break;
    case GL_VERTEX_ARRAY_STRIDE: p[0] = c->vertex_ptr.stride; break;
    case GL_COLOR_ARRAY_SIZE: p[0] = c->color_ptr.size; break;
    case GL_COLOR_ARRAY_TYPE: p[0] = (GLint)c->color_ptr.type; break;
    case GL_COLOR_ARRAY_STRIDE: p[0] = c->color_ptr.stride; break;
    case GL_NORMAL_ARRAY_TYPE: p[0] = (GLint)c->normal_ptr.type; break;
    case GL_NORMAL_ARRAY_STRIDE: p[0] = c->normal_ptr.stride; break;
    case GL_TEXTURE_COORD_ARRAY_SIZE: p[0] = c->texcoord_ptr.size; break;
    case GL_TEXTURE_COORD_ARRAY_TYPE: p[0] = (GLint)c->texcoord_ptr.type; break;
    case GL_TEXTURE_COORD_ARRAY_STRIDE: p[0] = c->texcoord_ptr.stride; break;
    case GL_LIST_BASE: p[0] = (GLint)c->list_base; break;
    case GL_LIST_INDEX: p[0] = (GLint)c->list_index; break;
    case GL_LIST_MODE: p[0] = c->list_index ? (GLint)c->list_mode : 0; break;
    case GL_CURRENT_RASTER_POSITION: p[0] = c->raster_x; p[1] = c->raster_y; p[2] = 0; p[3] = 1; break;
    case GL_CURRENT_RASTER_POSITION_VALID: p[0] = c->raster_valid; break;
    case GL_RENDER_MODE: p[0] = 0x1C00; break;
    case GL_MAX_LIGHTS: p[0] = MTGL_MAX_LIGHTS; break;
    case GL_MAX_CLIP_PLANES: p[0] = 6; break;
    case GL_MAX_TEXTURE_SIZE: p[0] = MTGL_MAX_TEX_SIZE; break;
    case GL_MAX_3D_TEXTURE_SIZE: case GL_MAX_CUBE_MAP_TEXTURE_SIZE: p[0] = 0; break;
    case GL_MAX_PIXEL_MAP_TABLE: p[0] = 256; break;
    case GL_MAX_ATTRIB_STACK_DEPTH: case GL_MAX_CLIENT_ATTRIB_STACK_DEPTH: p[0] = 16; break;
    case GL_MAX_MODELVIEW_STACK_DEPTH: case GL_MAX_PROJECTION_STACK_DEPTH: case GL_MAX_TEXTURE_STACK_DEPTH:
        p[0] = GLS_STACK_DEPTH; break;
    case GL_MAX_NAME_STACK_DEPTH: p[0] = 64; break;
    case GL_MAX_VIEWPORT_DIMS: p[0] = MTGL_MAX_FB_DIM; p[1] = MTGL_MAX_FB_DIM; break;
    case GL_MAX_TEXTURE_UNITS: p[0] = 1; break;
    case GL_MAX_ELEMENTS_VERTICES: case GL_MAX_ELEMENTS_INDICES: p[0] = 65536; break;
    case GL_SUBPIXEL_BITS: p[0] = 4; break;
    case GL_INDEX_BITS: p[0] = 0; break;
    case GL_RED_BITS: case GL_GREEN_BITS: case GL_BLUE_BITS: case GL_ALPHA_BITS: p[0] = 8; break;
    case GL_DEPTH_BITS: p[0] = 32; break;
    case GL_STENCIL_BITS: p[0] = 8; break;
    case GL_ACCUM_RED_BITS: case GL_ACCUM_GREEN_BITS: case GL_ACCUM_BLUE_BITS: case GL_ACCUM_ALPHA_BITS:
        p[0] = 0; break;
    case GL_AUX_BUFFERS: p[0] = 0; break;
    case GL_DOUBLEBUFFER: p[0] = GL_TRUE; break;
    case GL_STEREO: p[0] = GL_FALSE; break;
    case GL_RGBA_MODE: p[0] = GL_TRUE; break;
    case GL_INDEX_MODE: p[0] = GL_FALSE; break;
    case GL_SAMPLE_BUFFERS: case GL_SAMPLES: p[0] = 0; break;
    default: gls_set_error(c, GL_INVALID_ENUM);
    }
}

MTGL_EXPORT void glGetFloatv(GLenum pname, GLfloat *p)
{
    CTX_OR_RETURN();
    if (!p) return;
    switch (pname) {
    case GL_MODELVIEW_MATRIX: memcpy(p, c->modelview[c->modelview_depth], 16 * sizeof(float)); break;
    case GL_PROJECTION_MATRIX: memcpy(p, c->projection[c->projection_depth], 16 * sizeof(float)); break;
    case GL_TEXTURE_MATRIX: memcpy(p, c->texmat[c->texture_depth], 16 * sizeof(float)); break;
    case GL_CURRENT_COLOR: case GL_CURRENT_RASTER_COLOR: color4_out(p, c->current_color); break;
    case GL_CURRENT_NORMAL: memcpy(p, c->current_normal, 3 * sizeof(float)); break;
    case GL_CURRENT_TEXTURE_COORDS: p[0] = c->current_texcoord[0]; p[1] = c->current_texcoord[1]; break;
    case GL_CURRENT_RASTER_POSITION: p[0] = (GLfloat)c->raster_x; p[1] = (GLfloat)c->raster_y; p[2] = 0; p[3] = 1; break;
    case GL_DEPTH_RANGE: p[0] = (GLfloat)c->depth_near; p[1] = (GLfloat)c->depth_far; break;
    case GL_VIEWPORT:
        p[0] = (GLfloat)c->viewport_x; p[1] = (GLfloat)c->viewport_y;
        p[2] = (GLfloat)c->viewport_w; p[3] = (GLfloat)c->viewport_h; break;
    case GL_DEPTH_CLEAR_VALUE: p[0] = (GLfloat)c->clear_depth; break;
    case GL_COLOR_CLEAR_VALUE: color4_out(p, c->clear_color); break;
    case GL_FOG_COLOR: color4_out(p, c->fog_color); break;
    case GL_FOG_DENSITY: p[0] = c->fog_density; break;
    case GL_FOG_START: p[0] = c->fog_start; break;
    case GL_FOG_END: p[0] = c->fog_end; break;
    case GL_LIGHT_MODEL_AMBIENT: color4_out(p, c->light_model_ambient); break;
    case GL_ALPHA_TEST_REF: p[0] = c->alpha_ref; break;
    case GL_BLEND_COLOR: p[0] = p[1] = p[2] = p[3] = 0.0f; break;
    case GL_POINT_SIZE: p[0] = c->point_size; break;
    case GL_POINT_SIZE_RANGE: p[0] = 1.0f; p[1] = 64.0f; break;
    case GL_POINT_SIZE_GRANULARITY: case GL_LINE_WIDTH_GRANULARITY: p[0] = 1.0f; break;
    case GL_LINE_WIDTH: p[0] = c->line_width; break;
    case GL_LINE_WIDTH_RANGE: p[0] = 1.0f; p[1] = 16.0f; break;
    case GL_POLYGON_OFFSET_FACTOR: case GL_POLYGON_OFFSET_UNITS: p[0] = 0.0f; break;
    case GL_TEXTURE_ENV_COLOR: color4_out(p, c->tex_env_color); break;
    case GL_SCISSOR_BOX:
        p[0] = (GLfloat)c->scissor_x; p[1] = (GLfloat)c->scissor_y;
        p[2] = (GLfloat)c->scissor_w; p[3] = (GLfloat)c->scissor_h; break;
    case GL_MAX_TEXTURE_LOD_BIAS: p[0] = 2.0f; break;
    default: {
        /* integer state converted; an unknown name reports GL_INVALID_ENUM once */
        GLint ip[4] = { 0 };
        GLenum saved = c->error;
        c->error = GL_NO_ERROR;
        glGetIntegerv(pname, ip);
        if (c->error == GL_NO_ERROR) {
            p[0] = (GLfloat)ip[0];
            c->error = saved;
        } else {
            c->error = saved;
            gls_set_error(c, GL_INVALID_ENUM);
        }
    }
    }
}

MTGL_EXPORT void glGetDoublev(GLenum pname, GLdouble *p)
{
    CTX_OR_RETURN();
    (void)c;
    if (!p) return;
    GLfloat f[16];
    glGetFloatv(pname, f);
    int n = 1;
    switch (pname) {
    case GL_MODELVIEW_MATRIX: case GL_PROJECTION_MATRIX: case GL_TEXTURE_MATRIX: n = 16; break;
    case GL_CURRENT_COLOR: case GL_CURRENT_RASTER_COLOR: case GL_CURRENT_RASTER_POSITION: case GL_VIEWPORT:
    case GL_SCISSOR_BOX: case GL_COLOR_CLEAR_VALUE: case GL_FOG_COLOR: case GL_LIGHT_MODEL_AMBIENT:
    case GL_BLEND_COLOR: case GL_TEXTURE_ENV_COLOR: case GL_COLOR_WRITEMASK: n = 4; break;
    case GL_CURRENT_NORMAL: n = 3; break;
    case GL_CURRENT_TEXTURE_COORDS: case GL_DEPTH_RANGE: case GL_POINT_SIZE_RANGE: case GL_LINE_WIDTH_RANGE:
    case GL_MAX_VIEWPORT_DIMS: case GL_POLYGON_MODE: n = 2; break;
    default: break;
    }
    for (int i = 0; i < n; i++) p[i] = (GLdouble)f[i];
}

MTGL_EXPORT void glGetBooleanv(GLenum pname, GLboolean *p)
{
    CTX_OR_RETURN();
    if (!p) return;
    uint32_t f = cap_flag(pname);
    if (f) { p[0] = (c->flags & f) ? GL_TRUE : GL_FALSE; return; }
    switch (pname) {
    case GL_DEPTH_WRITEMASK: p[0] = c->depth_mask; return;
    case GL_COLOR_WRITEMASK: for (int i = 0; i < 4; i++) p[i] = c->color_mask[i]; return;
    case GL_DOUBLEBUFFER: case GL_RGBA_MODE: p[0] = GL_TRUE; return;
    case GL_STEREO: case GL_INDEX_MODE: p[0] = GL_FALSE; return;
    case GL_CURRENT_RASTER_POSITION_VALID: p[0] = c->raster_valid; return;
    case GL_LIGHT_MODEL_LOCAL_VIEWER: p[0] = c->light_model_local_viewer; return;
    case GL_LIGHT_MODEL_TWO_SIDE: p[0] = c->light_model_two_side; return;
    case GL_VERTEX_ARRAY: p[0] = (c->client_state & GLS_CLIENT_VERTEX) ? GL_TRUE : GL_FALSE; return;
    case GL_COLOR_ARRAY: p[0] = (c->client_state & GLS_CLIENT_COLOR) ? GL_TRUE : GL_FALSE; return;
    case GL_NORMAL_ARRAY: p[0] = (c->client_state & GLS_CLIENT_NORMAL) ? GL_TRUE : GL_FALSE; return;
    case GL_TEXTURE_COORD_ARRAY: p[0] = (c->client_state & GLS_CLIENT_TEXCOORD) ? GL_TRUE : GL_FALSE; return;
    default: break;
    }
    if (pname >= GL_LIGHT0 && pname <= GL_LIGHT7) { p[0] = c->lights[pname - GL_LIGHT0].enabled ? GL_TRUE : GL_FALSE; return; }
    GLint ip[16] = { 0 };
    glGetIntegerv(pname, ip);
    p[0] = ip[0] != 0 ? GL_TRUE : GL_FALSE;
}

MTGL_EXPORT GLboolean glIsEnabled(GLenum cap)
{
    CTX_OR_RETURN_V(GL_FALSE);
    uint32_t f = cap_flag(cap);
    if (f) return (c->flags & f) ? GL_TRUE : GL_FALSE;
    if (cap >= GL_LIGHT0 && cap <= GL_LIGHT7) return c->lights[cap - GL_LIGHT0].enabled ? GL_TRUE : GL_FALSE;
    gls_set_error(c, GL_INVALID_ENUM);
    return GL_FALSE;
}

MTGL_EXPORT const GLubyte *glGetString(GLenum name)
{
    CTX_OR_RETURN_V(NULL);
    switch (name) {
    case GL_VENDOR: return (const GLubyte *)"mtgl-mi355x";
    case GL_RENDERER: return (const GLubyte *)"MyTinyGL pipeline on AMD Instinct MI355X (gfx950)";
    case GL_VERSION: return (const GLubyte *)"1.5 mtgl-mi355x";
    case GL_EXTENSIONS: return (const GLubyte *)"";
    default: gls_set_error(c, GL_INVALID_ENUM); return NULL;
    }
}
