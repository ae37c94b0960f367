/*
 * mtgl_dev.h - C-ABI seam between the GL 1.5 host layer and the MI355X pipeline.
 *
 * In the reference renderer every GL call runs synchronously on one CPU thread and
 * glEnd() hands the buffered vertices to flush_{triangles,...}(GLState*)
 * (reference src/gl_api.c:648-659, prototypes src/mytinygl.h:247-256).  Here that
 * seam is a plain C ABI: the host layer (gl_api.c, gl_lists.c, gl_buffers.c)
 * snapshots the GL state into the POD records below and calls mtgl_dev_draw();
 * the device side owns the colour/depth/stencil buffers, textures and buffer
 * objects in HBM and runs vertex T&L, clipping, setup, binning and the per-tile
 * fragment pipeline.
 *
 * Conventions shared by every entry point:
 *   - return MTGL_OK (0) or a positive MTGL_ERR_* code; nothing aborts;
 *   - pixel rows are top-down framebuffer rows (row 0 = top), colour texels and
 *     pixels are RGBA8 packed as (a<<24)|(b<<16)|(g<<8)|r
 *     (reference src/graphics.h:337-348);
 *   - work is queued on one stream per context, in call order; mtgl_dev_read*
 *     and mtgl_dev_finish wait for it.
 *
 * No torch or HIP type appears here: the header is included by plain C.
 */
#ifndef MTGL_DEV_H
#define MTGL_DEV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define MTGL_OK                0
#define MTGL_ERR_OUT_OF_MEMORY 1   /* mapped to GL_OUT_OF_MEMORY by the host layer */
#define MTGL_ERR_INVALID       2   /* bad argument (id, size, range) */
#define MTGL_ERR_DEVICE        3   /* no usable GPU or a HIP runtime failure */
#define MTGL_ERR_UNSUPPORTED   4   /* primitive class not implemented on the device */

/* ---- limits (reference src/mytinygl.h:21-23, src/textures.h:17-18, src/vbo.h:17,
 *      src/framebuffer.h:30) ---- */
#define MTGL_MAX_LIGHTS      8
#define MTGL_MAX_OBJECTS     256     /* texture ids and buffer ids are 1..256 */
#define MTGL_MAX_TEX_SIZE    2048
#define MTGL_MAX_FB_DIM      16384

/* ---- GLState.flags bits (reference src/mytinygl.h:26-37) ---- */
#define MTGL_F_DEPTH_TEST      (1u << 1)
#define MTGL_F_CULL_FACE       (1u << 2)
#define MTGL_F_BLEND           (1u << 3)
#define MTGL_F_TEXTURE_2D      (1u << 4)
#define MTGL_F_LIGHTING        (1u << 5)
#define MTGL_F_FOG             (1u << 6)
#define MTGL_F_NORMALIZE       (1u << 7)
#define MTGL_F_COLOR_MATERIAL  (1u << 8)
#define MTGL_F_ALPHA_TEST      (1u << 9)
#define MTGL_F_SCISSOR_TEST    (1u << 10)
#define MTGL_F_STENCIL_TEST    (1u << 11)

/* ---- texture level of detail (mtgl_raster_state.lod_mode) ----
 * MTGL_LOD_REFERENCE: the reference's sampling, level 0 and a level 1 with one
 * LOD per triangle (src/textures.c:457-557, src/raster.c:505-529).
 * MTGL_LOD_FULL_CHAIN, an extension with no counterpart in the reference: the
 * whole chain of box-filtered levels (each halving the previous one's sides,
 * built like level 1) and a LOD per pixel, log2 of the larger screen-space
 * derivative length of the texture coordinates in texels; GL 1.5 level
 * selection (nearest level at ceil(lod + 0.5) - 1, linear between floor(lod)
 * and the next).  Points and lines sample at LOD 0 in both modes. */
#define MTGL_LOD_REFERENCE   0u
#define MTGL_LOD_FULL_CHAIN  1u

/* ---- buffers selected by mtgl_dev_clear / mtgl_dev_read / mtgl_dev_write ---- */
#define MTGL_BUF_COLOR    1u
#define MTGL_BUF_DEPTH    2u
#define MTGL_BUF_STENCIL  4u

typedef struct { float r, g, b, a; } mtgl_color;

/* One light source in eye space (reference light_t, src/mytinygl.h:54-66).
 * spot_cos_cutoff is cosf(spot_cutoff * 3.14159265f / 180.0f) evaluated by the host
 * C library, exactly the value the reference computes per vertex
 * (src/lighting.h:103). */
typedef struct {
    mtgl_color ambient, diffuse, specular;
    float position[4];
    float spot_direction[3];
    float spot_exponent;
    float spot_cutoff;
    float spot_cos_cutoff;
    float constant_attenuation, linear_attenuation, quadratic_attenuation;
    uint32_t enabled;
    uint32_t pad_[2];
} mtgl_light;

/* Material (reference material_t, src/mytinygl.h:69-75). */
typedef struct {
    mtgl_color ambient, diffuse, specular, emission;
    float shininess;
    uint32_t pad_[3];
} mtgl_material;

/* Inputs of compute_lighting (reference src/lighting.h:53-142). */
typedef struct {
    mtgl_light lights[MTGL_MAX_LIGHTS];
    mtgl_material front, back;
    mtgl_color model_ambient;
    uint32_t local_viewer;
    uint32_t two_side;
    uint32_t pad_[2];
} mtgl_lighting;

/* Per-vertex transform-and-lighting state: everything emit_vertex reads
 * (reference src/gl_api.c:263-348).  normal_matrix is the inverse transpose of the
 * modelview's upper 3x3 as computed by the reference's mat4_normal_matrix
 * (src/graphics.h:229-260) -- a function of the modelview alone, evaluated once
 * per state by the host.  lighting.front/back hold the material before the draw's
 * first vertex; GL_COLOR_MATERIAL overrides are applied per vertex on the device. */
typedef struct {
    float modelview[16];
    float projection[16];
    float texture[16];
    float normal_matrix[16];
    uint32_t flags;                 /* MTGL_F_* */
    uint32_t shade_model;           /* GL_FLAT / GL_SMOOTH / GL_PHONG */
    uint32_t color_material_face;   /* GL_FRONT / GL_BACK / GL_FRONT_AND_BACK */
    uint32_t color_material_mode;   /* GL_AMBIENT ... GL_AMBIENT_AND_DIFFUSE */
    mtgl_lighting lighting;
} mtgl_vertex_state;

/* Per-draw rasterisation and fragment state: the GLState fields read by
 * render_triangle / rasterize_triangle_smooth at glEnd time
 * (reference src/raster.c:451-725, :901-958). */
typedef struct {
    uint32_t flags;                 /* MTGL_F_* */
    int32_t  viewport[4];           /* x, y, w, h (y in top-down rows) */
    int32_t  scissor[4];            /* x, y, w, h (y in top-down rows) */
    uint32_t cull_face_mode, front_face;
    uint32_t polygon_mode_front, polygon_mode_back;
    uint32_t shade_model;
    uint32_t depth_func, depth_mask;
    double   depth_near, depth_far;
    uint32_t stencil_func;
    int32_t  stencil_ref;
    uint32_t stencil_mask, stencil_fail, stencil_zfail, stencil_zpass, stencil_writemask;
    uint32_t alpha_func;
    float    alpha_ref;
    uint32_t blend_src, blend_dst;
    uint32_t color_mask;            /* bit0 r, bit1 g, bit2 b, bit3 a */
    uint32_t texture;               /* id of the texture sampled, 0 = untextured draw */
    uint32_t min_filter, mag_filter, wrap_s, wrap_t;
    uint32_t tex_env_mode;
    mtgl_color tex_env_color;
    uint32_t perspective_hint;      /* GL_PERSPECTIVE_CORRECTION_HINT value */
    uint32_t fog_mode;
    float    fog_density, fog_start, fog_end;
    mtgl_color fog_color;
    float    line_width, point_size;
    uint32_t lod_mode;              /* MTGL_LOD_REFERENCE or MTGL_LOD_FULL_CHAIN (extension) */
    uint32_t pad_;
    mtgl_lighting lighting;         /* lights/materials at glEnd (GL_PHONG, two-sided) */
} mtgl_raster_state;

/* One attribute or index stream.  Element i lives at byte offset + i*stride of its
 * source: the draw's staged block (buffer == 0) or device buffer object `buffer`. */
typedef struct {
    uint32_t enabled;    /* 0: use the draw's current value instead */
    int32_t  size;       /* components per element */
    uint32_t type;       /* GL_FLOAT / GL_UNSIGNED_BYTE (attributes);
                            GL_UNSIGNED_BYTE / _SHORT / _INT (indices, state ids) */
    int32_t  stride;     /* bytes between consecutive elements (> 0) */
    uint32_t buffer;     /* 0 = staged block, MTGL_SEGMENT_BIT | id = display-list segment,
                            else buffer object id */
    uint32_t pad_;
    int64_t  offset;     /* byte offset of element 0 (may be negative for a staged
                            window that starts at a later element) */
} mtgl_array;

/* One draw call: glDrawArrays, glDrawElements or a glBegin/glEnd block
 * (reference src/gl_api.c:1799-1941, :614-662). */
typedef struct {
    uint32_t mode;              /* GL_POINTS ... GL_POLYGON */
    uint32_t count;             /* vertices emitted */
    int32_t  first;             /* element of vertex 0 when index_type == 0 */
    uint32_t index_type;        /* 0: element of vertex i is first + i;
                                   else indices are read from `indices` */
    mtgl_array indices;
    mtgl_array position, color, texcoord, normal;
    mtgl_array state_index;     /* per-vertex u32 index into vstates when n_vstates > 1 */
    float cur_color[4];         /* current attributes for disabled arrays */
    float cur_texcoord[2];
    float cur_normal[3];
    uint32_t n_vstates;
    const mtgl_vertex_state *vstates;
    const mtgl_raster_state *raster;
    const void *staged;         /* host bytes copied by the call */
    uint64_t staged_bytes;
} mtgl_draw;

/* Work counters accumulated by the fragment stage (reset by mtgl_dev_stats_reset).
 * A fragment is one (triangle, pixel) pair passing the coverage test of
 * reference src/raster.c:539-540, or one pixel visited by the point and line
 * loops (src/raster.c:152-243, :1113-1155: steps x line width, point size
 * squared, one per corner point); frag_bytes counts the algorithmic HBM bytes
 * of the reference per-fragment path of triangles (stencil, depth, texel
 * taps, colour). */
typedef struct {
    uint64_t fragments;
    uint64_t frag_bytes;
    uint64_t triangles_in;      /* triangles assembled */
    uint64_t triangles_raster;  /* fan triangles reaching setup */
    uint64_t bin_pairs;         /* (tile, triangle) pairs binned */
    uint64_t draws;
    double   raster_ms;         /* summed duration of the fragment kernel launches
                                   (the draws' fragment stages and glDrawPixels),
                                   measured with events on the context stream when
                                   timing is enabled */
    uint64_t raster_launches;
    uint64_t upload_bytes;      /* host bytes draws staged and display-list segments sent */
} mtgl_dev_stats;

typedef struct mtgl_dev mtgl_dev;

/* Context lifetime: HBM colour (u32), depth (f32) and stencil (u8) planes of
 * width x height (reference framebuffer_init, src/framebuffer.h:32-58). */
int  mtgl_dev_create(int32_t width, int32_t height, mtgl_dev **out);
void mtgl_dev_destroy(mtgl_dev *dev);
/* Wait for all queued work (reference glFinish, src/gl_api.c:895-899). */
int  mtgl_dev_finish(mtgl_dev *dev);

/* glClear: fill [x0,x1) x [y0,y1) of the selected planes
 * (reference src/gl_api.c:409-457). */
int  mtgl_dev_clear(mtgl_dev *dev, uint32_t buffers, int32_t x0, int32_t y0, int32_t x1, int32_t y1,
                    uint32_t color, float depth, uint8_t stencil);

/* Copy a w x h window whose top-left is framebuffer pixel (x, y) to/from tightly
 * packed host memory (u32 colour, f32 depth or u8 stencil).  The window must lie
 * inside the framebuffer.  Replaces the reference's direct framebuffer reads in
 * glReadPixels (src/gl_api.c:1180-1229) and the framebuffer accessors
 * (src/framebuffer.h:92-134). */
int  mtgl_dev_read(mtgl_dev *dev, uint32_t buffer, int32_t x, int32_t y, int32_t w, int32_t h, void *dst);
int  mtgl_dev_write(mtgl_dev *dev, uint32_t buffer, int32_t x, int32_t y, int32_t w, int32_t h, const void *src);

/* glReadPixels (reference src/gl_api.c:1180-1229): a width x height window whose
 * bottom-left pixel is window pixel (x, y) in GL convention (y counts up from
 * the bottom row), written to host memory `dst` as tightly packed
 * GL_UNSIGNED_BYTE pixels, rows bottom-up.  format GL_RGBA writes r, g, b, a
 * (4 bytes), GL_RGB writes r, g, b (3 bytes).  Window rows outside the
 * framebuffer read as zero bytes; pixels left or right of it, in rows that
 * exist, read as (0, 0, 0, 255).  The window may extend past every edge.
 * width or height <= 0, or another format, writes nothing and returns
 * MTGL_OK (the reference's loops write nothing for them).  The pixels are
 * packed and flipped on the device and copied to `dst` in one transfer. */
int  mtgl_dev_read_pixels(mtgl_dev *dev, int32_t x, int32_t y, int32_t width, int32_t height, uint32_t format,
                          void *dst);

/* Present: the whole colour plane, top row first, each row width * 4 bytes of
 * packed RGBA8 (SDL_PIXELFORMAT_ABGR8888 on little-endian hosts) at dst +
 * row * pitch, pitch >= width * 4.  Replaces the reference's SDL glue, which
 * uploads GLState.framebuffer.color with framebuffer_get_pitch() as pitch
 * (include/mytinygl/sdl.h:74-84, src/framebuffer.h:136-143). */
int  mtgl_dev_present(mtgl_dev *dev, void *dst, int64_t pitch);

/* Texture level 0 of texture `id` (1..256) as packed RGBA8 texels; the device also
 * builds the 2x2 box-filtered level 1 (reference texture_upload_*,
 * src/textures.c:141-269, texture_generate_mip1 :311-354). */
int  mtgl_dev_tex_image(mtgl_dev *dev, uint32_t id, int32_t width, int32_t height, const uint32_t *texels);
int  mtgl_dev_tex_free(mtgl_dev *dev, uint32_t id);

/* Buffer objects (reference buffer_data / buffer_sub_data, src/vbo.c:120-161).
 * size <= 0 frees the storage; data may be NULL (contents left unspecified). */
int  mtgl_dev_buffer_data(mtgl_dev *dev, uint32_t id, int64_t size, const void *data);
int  mtgl_dev_buffer_sub_data(mtgl_dev *dev, uint32_t id, int64_t offset, int64_t size, const void *data);
int  mtgl_dev_buffer_free(mtgl_dev *dev, uint32_t id);

/* Display-list segments: device-resident copies of the vertex records a
 * display list's glBegin/glEnd block produced when it was last replayed
 * (reference lists replay their recorded calls, src/lists.c:185-425,
 * src/gl_api.c:2290-2455; the host layer keeps replaying them and sends the
 * records again only when they differ).  A draw addresses segment `id`
 * (1..MTGL_MAX_SEGMENT) through mtgl_array.buffer = MTGL_SEGMENT_BIT | id.
 * size <= 0 frees the segment. */
#define MTGL_SEGMENT_BIT  0x80000000u
#define MTGL_MAX_SEGMENT  0x00FFFFFFu
int  mtgl_dev_segment_data(mtgl_dev *dev, uint32_t id, int64_t size, const void *data);
int  mtgl_dev_segment_free(mtgl_dev *dev, uint32_t id);

/* Render one draw into the context's framebuffer: replaces glEnd's flush_*
 * dispatch (reference src/gl_api.c:648-659, flush_* at src/raster.c:961-1231)
 * and the array draws glDrawArrays / glDrawElements (src/gl_api.c:1799-1941). */
int  mtgl_dev_draw(mtgl_dev *dev, const mtgl_draw *draw);

/* glDrawPixels: width x height GL_UNSIGNED_BYTE pixels of `format` (GL_RGBA,
 * GL_RGB, GL_LUMINANCE or GL_LUMINANCE_ALPHA; rows bottom-up, tightly packed),
 * pixel (col, row) landing on window (x + col, y + row) in GL convention (y up
 * from the bottom row).  Alpha test, depth test (incoming depth 0) and
 * blending follow `flags` (MTGL_F_ALPHA_TEST, MTGL_F_DEPTH_TEST, MTGL_F_BLEND;
 * other bits are ignored); no scissor, stencil, fog or colour mask applies
 * (reference glDrawPixels, src/gl_api.c:1284-1373, with its own compare and
 * blend-factor helpers :1231-1282).  `pixels` is host memory, copied by the
 * call. */
typedef struct {
    int32_t  x, y;              /* raster position (reference raster_pos_x / _y) */
    int32_t  width, height;
    uint32_t format;
    uint32_t flags;
    uint32_t alpha_func;
    float    alpha_ref;
    uint32_t depth_func;
    uint32_t depth_mask;
    uint32_t blend_src, blend_dst;
    const void *pixels;
} mtgl_pixels;

int  mtgl_dev_draw_pixels(mtgl_dev *dev, const mtgl_pixels *pixels);

/* glDrawPixels that allocates nothing.  No counterpart in the reference: for
 * callers that cannot allocate device or pinned memory (code under stream
 * capture, serving loops with a fixed memory footprint).  Same arguments and
 * result as mtgl_dev_draw_pixels, but the source image is staged in a fixed
 * device buffer of MTGL_STATIC_PIXELS_MAX bytes that the library holds on
 * each GPU it runs on until the process ends (its address is looked up when
 * the first context on that GPU is created).  A source image larger than the
 * buffer (width * height * bytes per pixel) is refused with MTGL_ERR_INVALID
 * and draws nothing.  Calls on one GPU share its buffer: the library
 * serialises them and each returns after its kernel finished. */
#define MTGL_STATIC_PIXELS_MAX (32u << 20)
int  mtgl_dev_draw_pixels_static(mtgl_dev *dev, const mtgl_pixels *pixels);

/* Counters and timing; no counterpart in the reference (fragments passing
 * coverage, algorithmic bytes of the per-fragment pipeline, fragment-stage time).
 * mtgl_dev_set_timing(dev, every): 0 stops timing; every >= 1 brackets one
 * fragment-stage launch in `every` with events (each pair of events costs a
 * few microseconds of queue time, so a sample keeps the timed frames honest). */
int  mtgl_dev_set_timing(mtgl_dev *dev, int every);
int  mtgl_dev_stats_get(mtgl_dev *dev, mtgl_dev_stats *out);
int  mtgl_dev_stats_reset(mtgl_dev *dev);

/* Short text naming the implementation ("hip:gfx950 ...", "oracle:cpu"). */
const char *mtgl_dev_backend(void);

#ifdef __cplusplus
}
#endif

#endif /* MTGL_DEV_H */
