/*
 * ref_access.c - framebuffer accessor compiled into oracle/_ref/libmtgl_ref.so.
 *
 * TEST INFRASTRUCTURE ONLY.  The reference renderer exposes its colour plane
 * through glReadPixels but has no call that returns depth or stencil, which the
 * parity tests compare bit for bit.  This accessor reads the planes of a
 * reference GLState through the reference's own internal header
 * (src/mytinygl.h, framebuffer_t at src/framebuffer.h:19-25).
 */
#include <string.h>
#include "mytinygl.h"

int mtgl_ref_read_framebuffer(GLState *ctx, uint32_t *color, float *depth, uint8_t *stencil)
{
    if (!ctx) return 2;
    size_t n = (size_t)ctx->framebuffer.width * (size_t)ctx->framebuffer.height;
    if (color) memcpy(color, ctx->framebuffer.color, n * sizeof(uint32_t));
    if (depth) memcpy(depth, ctx->framebuffer.depth, n * sizeof(float));
    if (stencil) memcpy(stencil, ctx->framebuffer.stencil, n);
    return 0;
}
