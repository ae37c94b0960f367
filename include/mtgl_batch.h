/*
 * mtgl_batch.h - many contexts' frames in one device launch.
 *
 * No counterpart in the reference: MyTinyGL renders each draw when glEnd,
 * glDrawArrays or glDrawElements is called (src/gl_api.c:648-659,
 * :1799-1941), one context at a time.  A batch is for callers that render
 * many independent small frames (one scene per context) and pay per-draw
 * launch and synchronisation costs for each of them.
 *
 * Use:
 *   mtgl_batch_create(&b);
 *   mtgl_batch_begin(b);            // this thread's clears and draws are recorded
 *   for each context: gl_make_current(ctx); glClear(...); glDrawElements(...);
 *   mtgl_batch_end(b, 1);           // size, autotune and capture; framebuffers unchanged
 *   mtgl_batch_launch(b);           // replay every recorded command, asynchronously
 *   glReadPixels / mtgl_read_framebuffer on any of the contexts (waits for the launch)
 *
 * Between begin and end, mtgl_dev_clear and mtgl_dev_draw on any context
 * issued by the calling thread are recorded (their arguments, raster state and
 * client arrays are copied) and return MTGL_OK without rendering; every other
 * call runs at once.  A launch replays the commands of each context in the
 * order they were recorded; commands of different contexts are independent.
 * Round r of a launch holds the r-th command of every context, so a batch of
 * one clear and one draw per context is two device launches for all of them.
 *
 * Buffer objects are read when the batch is launched, so glBufferSubData
 * between launches changes the geometry; textures and buffer objects must keep
 * their storage: glTexImage2D, glDeleteTextures, glBufferData or
 * glDeleteBuffers on a context of the batch, or destroying such a context,
 * retires the batch (mtgl_batch_launch then returns MTGL_ERR_INVALID).
 * mtgl_batch_sync returns MTGL_ERR_OUT_OF_MEMORY when a launch produced more
 * clipped triangles or (tile, triangle) pairs than the run that sized the batch
 * (geometry changed since): that launch's frames are incomplete.
 */
#ifndef MTGL_BATCH_H
#define MTGL_BATCH_H

#include <stdint.h>

#include "mtgl_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MTGL_BATCH_TUNE_MAX 8

typedef struct mtgl_batch mtgl_batch;

typedef struct {
    uint32_t contexts;          /* framebuffers the batch renders into */
    uint32_t commands;          /* recorded clears and draws */
    uint32_t rounds;            /* device launches per replay */
    uint32_t scenes;            /* draws with triangles */
    uint32_t launches;          /* mtgl_batch_launch calls so far */
    uint32_t heavy_bin;         /* chosen: bins longer than this are rasterised first */
    uint32_t grid_div;          /* chosen: raster grids hold resident blocks / grid_div */
    uint32_t tune_candidates;   /* entries of tune_ms used (0 without autotune) */
    double tune_ms[MTGL_BATCH_TUNE_MAX];   /* replay time of each candidate */
} mtgl_batch_info_t;

int  mtgl_batch_create(mtgl_batch **out);
void mtgl_batch_destroy(mtgl_batch *batch);
/* Start recording this thread's clears and draws.  A batch records once. */
int  mtgl_batch_begin(mtgl_batch *batch);
/* Stop recording; run the commands once to size the device buffers, with
 * autotune != 0 time replays over a few tile-list and grid settings, capture
 * the replay into a graph.  The framebuffers are left as they were. */
int  mtgl_batch_end(mtgl_batch *batch, int autotune);
/* Replay every recorded command (asynchronous, ordered after the contexts'
 * queued work and before their later work). */
int  mtgl_batch_launch(mtgl_batch *batch);
/* Wait for the batch's launches; reports incomplete launches (see above). */
int  mtgl_batch_sync(mtgl_batch *batch);
int  mtgl_batch_info(mtgl_batch *batch, mtgl_batch_info_t *out);
/* Fragment counters of the batch's launches (the fields of mtgl_dev_stats). */
int  mtgl_batch_stats_get(mtgl_batch *batch, mtgl_dev_stats *out);
int  mtgl_batch_stats_reset(mtgl_batch *batch);

#ifdef __cplusplus
}
#endif

#endif /* MTGL_BATCH_H */
