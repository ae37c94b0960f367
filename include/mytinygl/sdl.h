/*
 * mytinygl/sdl.h - SDL2 window glue for the MI355X library.
 *
 * Same helpers as the reference's header (include/mytinygl/sdl.h:19-98):
 * mtgl_init opens a window with an SDL_PIXELFORMAT_ABGR8888 streaming texture
 * and makes a new GL context current; mtgl_swap presents the frame;
 * mtgl_destroy tears everything down.  The reference's mtgl_swap hands
 * GLState.framebuffer.color (host memory) to SDL_UpdateTexture.  Here the
 * colour plane lives in HBM, so mtgl_swap locks the streaming texture and
 * mtgl_present() (include/mtgl_context.h) copies the plane straight into it
 * after the queued draws finish: one device-to-host copy per frame, no
 * intermediate host buffer.  Programs that include it link -lmtgl -lSDL2.
 */
#ifndef MYTINYGL_SDL_H
#define MYTINYGL_SDL_H

#include <stddef.h>

#include <SDL2/SDL.h>

#include "../mtgl_context.h"

static SDL_Window *mtgl_window = NULL;
static SDL_Renderer *mtgl_renderer = NULL;
static SDL_Texture *mtgl_texture = NULL;
static GLState *mtgl_ctx = NULL;

static inline void mtgl_sdl_teardown(void)
{
    if (mtgl_ctx) {
        gl_make_current(NULL);
        gl_destroy_context(mtgl_ctx);
        mtgl_ctx = NULL;
    }
    if (mtgl_texture) SDL_DestroyTexture(mtgl_texture);
    if (mtgl_renderer) SDL_DestroyRenderer(mtgl_renderer);
    if (mtgl_window) SDL_DestroyWindow(mtgl_window);
    mtgl_texture = NULL;
    mtgl_renderer = NULL;
    mtgl_window = NULL;
    SDL_Quit();
}

/* 0 on success, -1 when SDL or the GL context (no usable GPU) fails */
static inline int mtgl_init(const char *title, int32_t width, int32_t height)
{
    if (SDL_Init(SDL_INIT_VIDEO) < 0) return -1;
    mtgl_window = SDL_CreateWindow(title, SDL_WINDOWPOS_CENTERED, SDL_WINDOWPOS_CENTERED, width, height,
                                   SDL_WINDOW_SHOWN);
    if (mtgl_window)
        mtgl_renderer = SDL_CreateRenderer(mtgl_window, -1, SDL_RENDERER_ACCELERATED | SDL_RENDERER_PRESENTVSYNC);
    if (mtgl_renderer)
        mtgl_texture = SDL_CreateTexture(mtgl_renderer, SDL_PIXELFORMAT_ABGR8888, SDL_TEXTUREACCESS_STREAMING, width,
                                         height);
    if (mtgl_texture) mtgl_ctx = gl_create_context(width, height);
    if (!mtgl_ctx) {
        mtgl_sdl_teardown();
        return -1;
    }
    gl_make_current(mtgl_ctx);
    return 0;
}

static inline void mtgl_swap(void)
{
    void *pixels = NULL;
    int pitch = 0;
    if (SDL_LockTexture(mtgl_texture, NULL, &pixels, &pitch) == 0) {
        (void)mtgl_present(mtgl_ctx, pixels, pitch);
        SDL_UnlockTexture(mtgl_texture);
    }
    SDL_RenderCopy(mtgl_renderer, mtgl_texture, NULL, NULL);
    SDL_RenderPresent(mtgl_renderer);
}

static inline void mtgl_destroy(void)
{
    mtgl_sdl_teardown();
}

#endif /* MYTINYGL_SDL_H */
