/*
 * mtgl_context.h - GL context lifetime and headless framebuffer access.
 *
 * Context management mirrors the reference (src/mytinygl.h:227-233):
 * gl_create_context allocates a context whose colour (RGBA8), depth (fp32) and
 * stencil (u8) planes live in HBM; the current context is per OS thread.  The
 * reference's SDL glue reads GLState.framebuffer.color directly
 * (include/mytinygl/sdl.h:74-84); with an HBM framebuffer that read becomes an
 * explicit download, mtgl_read_framebuffer().
 */
#ifndef MTGL_CONTEXT_H
#define MTGL_CONTEXT_H

#include <stdint.h>
#include "GL/gl.h"
#include "mtgl_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct GLState GLState;

GLState *gl_create_context(int32_t width, int32_t height);
void     gl_destroy_context(GLState *ctx);
void     gl_make_current(GLState *ctx);
GLState *gl_get_current_context(void);

/* Download whole planes (row 0 = top row) after all queued work finished.
 * Any of the pointers may be NULL.  Returns MTGL_OK or an MTGL_ERR_* code. */
int mtgl_read_framebuffer(GLState *ctx, uint32_t *color, float *depth, uint8_t *stencil);

/* Present: the colour plane, top row first, as packed RGBA8 rows
 * (SDL_PIXELFORMAT_ABGR8888) of width * 4 bytes at pixels + row * pitch, after
 * all queued work finished; pitch >= width * 4.  The reference's SDL glue
 * uploads GLState.framebuffer.color with pitch width * 4
 * (include/mytinygl/sdl.h:74-84); include/mytinygl/sdl.h here calls this on
 * the streaming texture's locked memory.  Returns MTGL_OK or MTGL_ERR_*. */
int mtgl_present(GLState *ctx, void *pixels, int32_t pitch);

/* Framebuffer size and the device context behind a GL context (for counters). */
int32_t   mtgl_context_width(GLState *ctx);
int32_t   mtgl_context_height(GLState *ctx);
mtgl_dev *mtgl_context_device(GLState *ctx);

/* Texture level of detail of the context's later draws: MTGL_LOD_REFERENCE
 * (the default) or MTGL_LOD_FULL_CHAIN (see mtgl_dev.h).  No counterpart in
 * the reference; for callers that want full mipmapping with a per-pixel LOD.
 * The CPU oracle renders the reference's sampling in either mode. */
void mtgl_texture_lod_mode(GLState *ctx, uint32_t mode);

#ifdef __cplusplus
}
#endif

#endif /* MTGL_CONTEXT_H */
