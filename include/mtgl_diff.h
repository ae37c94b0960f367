/*
 * mtgl_diff.h - gradients of an image loss through the per-draw pipeline.
 *
 * No counterpart in the reference: MyTinyGL renders forward only.  This is
 * for callers that fit texels, vertex colours and light colours to target
 * images (inverse rendering) with the geometry held fixed, because
 * visibility is discontinuous in it.
 *
 * A view is captured from a context right after a draw rendered through the
 * ordinary path (mtgl_dev_draw): per pixel, the triangle that left its colour,
 * the weights of that triangle's three vertex colours and the bilinear
 * footprint of its texture lookup (a G-buffer; fragment lists are not kept),
 * plus the draw's vertex state, positions, normals and index stream.  Shading
 * is then recomputed from parameters:
 *
 *   lit_e   = compute_lighting of element e with vertex colour vcol[e] through
 *             GL_COLOR_MATERIAL and light i's diffuse rgb = light_rgb[i]
 *             (reference src/lighting.h:53-142, src/gl_api.c:263-348)
 *   pixel   = clamp((b0 lit_e0 + b1 lit_e1 + b2 lit_e2) * bilinear(tex))
 *             (reference src/raster.c:588, :663 GL_MODULATE, :720; bilinear_filter
 *             src/textures.c:294, :398)
 *
 * with fp32 texels (the ordinary path samples RGBA8 and re-quantises; with
 * tex = bytes / 255 the two agree within one RGBA8 step).  The backward pass
 * is the adjoint of that forward: fp32 atomics into texel and per-vertex
 * accumulators, then a per-vertex pass through the lighting.
 *
 * Captured draws must be GL_TRIANGLES with a colour array, smooth shading,
 * per-vertex lighting (not GL_PHONG, not two-sided) with GL_COLOR_MATERIAL in
 * GL_AMBIENT, GL_DIFFUSE or GL_AMBIENT_AND_DIFFUSE mode, directional lights,
 * a GL_LINEAR/GL_LINEAR texture of the diff's size in GL_MODULATE mode, no
 * fog, blending, partial colour mask or alpha test, and no triangle that
 * crosses a frustum plane; anything else is refused with
 * MTGL_ERR_UNSUPPORTED.  Every view has the same framebuffer size.
 *
 * Arrays are dense and row-major: vcol [vertices][4], light_rgb
 * [MTGL_MAX_LIGHTS][3], tex [tex_h][tex_w][4], images [views][height][width][4]
 * (top row first, as the framebuffer).  Pointers are device memory for
 * libmtgl.so, host memory for the CPU oracle; stream is a hipStream_t (NULL:
 * the default stream) and is ignored by the oracle.
 */
#ifndef MTGL_DIFF_H
#define MTGL_DIFF_H

#include <stddef.h>
#include <stdint.h>

#include "mtgl_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mtgl_diff mtgl_diff;

/* One pixel of a captured view (40 bytes). */
typedef struct {
    int32_t  tri;       /* triangle of the draw that left the pixel's colour; -1: none
                           (texel[0] then holds the pixel's packed colour) */
    float    b[3];      /* weights of the triangle's three vertex colours */
    uint32_t texel[4];  /* level-0 texels of the bilinear tap: (x0,y0) (x1,y0) (x0,y1) (x1,y1) */
    float    fx, fy;    /* bilinear fractions */
} mtgl_gbuffer_px;

typedef struct {
    uint32_t views;
    int32_t  width, height;
    uint32_t vertices;     /* elements addressed by vcol */
    int32_t  tex_w, tex_h;
    uint64_t work_floats;  /* floats of the backward workspace (views * vertices * 16) */
} mtgl_diff_info_t;

int  mtgl_diff_create(uint32_t vertices, int32_t tex_w, int32_t tex_h, mtgl_diff **out);
void mtgl_diff_destroy(mtgl_diff *diff);
/* Have the draws that follow on dev record each pixel's last passing fragment
 * (4 bytes per pixel of every tile a draw touches), which mtgl_diff_add_view
 * reads; draws on a device that was never armed skip that plane. */
int  mtgl_diff_arm(mtgl_dev *dev);
/* Capture the latest draw of dev as the next view (synchronises dev); the draw
 * must have been made after mtgl_diff_arm(dev), else MTGL_ERR_UNSUPPORTED. */
int  mtgl_diff_add_view(mtgl_diff *diff, mtgl_dev *dev);
int  mtgl_diff_info(mtgl_diff *diff, mtgl_diff_info_t *out);
/* Copy view's G-buffer (width * height pixels) to host memory. */
int  mtgl_diff_gbuffer(mtgl_diff *diff, uint32_t view, mtgl_gbuffer_px *host_out);
/* images = render(vcol, light_rgb, tex) for every view. */
int  mtgl_diff_forward(mtgl_diff *diff, const float *vcol, const float *light_rgb, const float *tex, float *images,
                       void *stream);
/* Given dL/dimages, write dL/dvcol, dL/dlight_rgb and dL/dtex (overwritten).
 * work: work_floats floats of scratch. */
int  mtgl_diff_backward(mtgl_diff *diff, const float *vcol, const float *light_rgb, const float *tex,
                        const float *grad_images, float *g_vcol, float *g_light_rgb, float *g_tex, float *work,
                        void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MTGL_DIFF_H */
