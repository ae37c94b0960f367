/*
 * gl_state.h - host-side GL context of libmtgl.
 *
 * Holds the GL 1.5 state machine (the reference's GLState, src/mytinygl.h:86-225)
 * minus the pixel storage, which lives in HBM behind mtgl_dev.  Object stores keep
 * the reference's id rules: 1-based ids, the first freed slot is reused, hard caps
 * of 256 textures / 256 buffers / 1024 display lists.
 */
#ifndef MTGL_GL_STATE_H
#define MTGL_GL_STATE_H

#include <stddef.h>
#include <stdint.h>
#include "GL/gl.h"
#include "mtgl_dev.h"
#include "mtgl_context.h"

#define GLS_STACK_DEPTH   24    /* reference MAX_MATRIX_STACK_DEPTH */
#define GLS_LIST_DEPTH    64    /* reference MAX_LIST_CALL_DEPTH */
#define GLS_MAX_LISTS     1024  /* reference GL_MAX_LISTS */

#define GLS_INSIDE_BEGIN_END  (1u << 0)

#define GLS_CLIENT_VERTEX    (1u << 0)
#define GLS_CLIENT_COLOR     (1u << 1)
#define GLS_CLIENT_TEXCOORD  (1u << 2)
#define GLS_CLIENT_NORMAL    (1u << 3)

typedef struct {
    GLint size;
    GLenum type;
    GLsizei stride;
    const void *pointer;
} gls_array_ptr;

typedef struct {
    uint8_t allocated;
    int32_t width, height;        /* level 0 size; 0 when no pixels were uploaded */
    GLenum min_filter, mag_filter, wrap_s, wrap_t;
} gls_texture;

typedef struct {
    uint8_t allocated;
    uint8_t *shadow;              /* host copy of the device buffer (index scans, current values) */
    GLsizeiptr size;
    GLenum usage;
} gls_buffer;

/* Display-list opcodes (the reference's 31 recorded calls, src/lists.h:21-53). */
typedef enum {
    LOP_END = 0, LOP_BEGIN, LOP_VERTEX, LOP_COLOR, LOP_TEXCOORD, LOP_NORMAL,
    LOP_TRANSLATEF, LOP_ROTATEF, LOP_SCALEF, LOP_PUSH_MATRIX, LOP_POP_MATRIX,
    LOP_LOAD_IDENTITY, LOP_MULT_MATRIXF, LOP_LOAD_MATRIXF, LOP_MATRIX_MODE,
    LOP_ORTHO, LOP_FRUSTUM, LOP_ENABLE, LOP_DISABLE, LOP_BIND_TEXTURE,
    LOP_BLEND_FUNC, LOP_DEPTH_FUNC, LOP_DEPTH_MASK, LOP_CULL_FACE, LOP_FRONT_FACE,
    LOP_SHADE_MODEL, LOP_LIGHTF, LOP_LIGHTFV, LOP_MATERIALF, LOP_MATERIALFV,
    LOP_CALL_LIST
} gls_list_op;

typedef struct {
    gls_list_op op;
    GLenum e[3];                  /* enum / id operands in call order */
    GLfloat f[16];                /* float operands in call order */
    GLdouble d[6];                /* glOrtho / glFrustum operands */
} gls_list_cmd;

/* Device-resident vertex records of one glBegin/glEnd block of a list: the
 * records its last replay sent (host copy, to compare) and the device segment
 * holding them. */
typedef struct {
    size_t cmd;                   /* index of the block's LOP_END in the list */
    uint32_t id;                  /* device segment id */
    uint8_t *host;
    size_t bytes;
    /* A block of vertex / colour / texcoord / normal calls only, last
     * replayed through its calls from an empty vertex buffer: its records are
     * a function of the current colour, texcoord and normal at glBegin, so a
     * replay from the same values skips the calls and draws the segment. */
    int fast;
    size_t begin;                 /* index of the block's LOP_BEGIN */
    float start[9];               /* current colour, texcoord, normal at glBegin */
    uint32_t start_mask;          /* which of them (bits 0-2) the records take */
    float end[9];                 /* the same after the block */
    float last_color[4];          /* colour of the block's last vertex */
    size_t nverts;
} gls_lseg;

typedef struct {
    gls_list_cmd *cmds;
    size_t count, capacity;
    GLboolean valid;
    GLboolean allocated;
    gls_lseg *segs;
    size_t nsegs, segs_capacity;
    size_t seg_hint;              /* segment the next block of a replay most likely uses */
} gls_list;

/* One glVertex call as seen by the host: object-space attributes plus the index
 * of the transform/lighting snapshot current at that call. */
typedef struct {
    float pos[3];
    float color[4];
    float texcoord[2];
    float normal[3];
    uint32_t state;
} gls_imm_vertex;

struct GLState {
    mtgl_dev *dev;
    int32_t width, height;

    mtgl_color clear_color;
    GLdouble clear_depth;
    GLint stencil_clear;

    GLint viewport_x, viewport_y;
    GLsizei viewport_w, viewport_h;

    mtgl_color current_color;
    float current_texcoord[2];
    float current_normal[3];

    GLenum matrix_mode;
    float modelview[GLS_STACK_DEPTH][16];
    float projection[GLS_STACK_DEPTH][16];
    float texmat[GLS_STACK_DEPTH][16];
    GLint modelview_depth, projection_depth, texture_depth;

    GLenum primitive_mode;
    uint32_t flags;               /* MTGL_F_* plus GLS_INSIDE_BEGIN_END */

    GLenum blend_src, blend_dst;
    GLenum cull_face_mode, front_face;
    GLenum depth_func;
    GLboolean depth_mask;
    GLenum alpha_func;
    GLfloat alpha_ref;
    GLint scissor_x, scissor_y;
    GLsizei scissor_w, scissor_h;
    GLenum stencil_func;
    GLint stencil_ref;
    GLuint stencil_mask;
    GLenum stencil_fail, stencil_zfail, stencil_zpass;
    GLuint stencil_writemask;
    GLdouble depth_near, depth_far;
    GLboolean color_mask[4];
    GLfloat line_width, point_size;
    GLenum polygon_mode_front, polygon_mode_back;

    gls_texture textures[MTGL_MAX_OBJECTS];
    size_t texture_count;
    GLuint bound_texture;
    GLenum tex_env_mode;
    mtgl_color tex_env_color;
    GLenum perspective_hint;

    GLint raster_x, raster_y;
    GLboolean raster_valid;

    GLenum fog_mode;
    GLfloat fog_density, fog_start, fog_end;
    mtgl_color fog_color;

    mtgl_light lights[MTGL_MAX_LIGHTS];
    mtgl_material material_front, material_back;
    mtgl_color light_model_ambient;
    GLboolean light_model_local_viewer, light_model_two_side;
    GLenum color_material_face, color_material_mode;
    GLenum shade_model;

    gls_buffer buffers[MTGL_MAX_OBJECTS];
    size_t buffer_count;
    GLuint bound_array_buffer, bound_element_buffer;

    uint32_t client_state;
    gls_array_ptr vertex_ptr, color_ptr, texcoord_ptr, normal_ptr;

    gls_list lists[GLS_MAX_LISTS];
    size_t list_count;
    GLuint list_base, list_index;
    GLenum list_mode;
    GLuint list_call_depth;
    /* the list block whose glEnd is being replayed (replay_list 0: none) */
    GLuint replay_list;
    size_t replay_cmd;
    uint32_t seg_next;            /* next unused device segment id */

    /* vertices buffered since the last successful glEnd */
    gls_imm_vertex *imm;
    size_t imm_count, imm_capacity;
    /* transform/lighting snapshots referenced by imm[].state */
    mtgl_vertex_state *vstates;
    size_t vstate_count, vstate_capacity;
    int tl_dirty;                 /* T&L state changed since the last snapshot */

    /* reusable staging for client arrays */
    uint8_t *stage;
    size_t stage_capacity;

    GLenum error;
    uint32_t warned;              /* one-shot diagnostics already printed */
    uint32_t lod_mode;            /* MTGL_LOD_* (extension) */
};

/* internal helpers shared by the host translation units */
GLState *gls_current(void);
void gls_set_error(GLState *c, GLenum err);
void gls_snapshot_vertex_state(GLState *c, mtgl_vertex_state *vs);
void gls_snapshot_raster_state(GLState *c, mtgl_raster_state *rs);
void gls_apply_color_material(GLState *c, const float *color);
void gls_flush_immediate(GLState *c);
void gls_draw_records(GLState *c, size_t n, uint32_t seg);
int  gls_begin_snapshot(GLState *c);
void gls_report_device_error(GLState *c, int rc, const char *what);
int  gls_list_recording(GLState *c);
void gls_list_free_segments(GLState *c, gls_list *l);
/* the device segment that holds `bytes` of vertex records `recs` for the list
 * block being replayed, refreshed when they differ from its last replay; 0
 * when none applies */
uint32_t gls_list_segment(GLState *c, const void *recs, size_t bytes);
int  gls_list_record(GLState *c, const gls_list_cmd *cmd);

#endif /* MTGL_GL_STATE_H */
