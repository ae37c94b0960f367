"""A plain C program linked against the drop-in library.

tests/c/headless_frames.c uses only include/GL/gl.h and the context calls,
as the reference's testbed programs do (testbed/1.5-0-vbo-test.c:116-199),
and writes each frame's glReadPixels bytes.  tests/c/Makefile links it with
-lmtgl (libmtgl.so), with the CPU oracle and with the strict reference build.
On the CPU the oracle's output equals the reference's; on the GPU the
product's output, including its mtgl_present copy, equals the oracle's.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import REF_LIB, ROOT

CDIR = os.path.join(ROOT, "tests", "c")
W, H, FRAMES = 200, 150, 7


def _build(target):
    r = subprocess.run(["make", "-C", CDIR, target], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def _run(binary, out):
    r = subprocess.run([os.path.join(CDIR, "_build", binary), out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, f"{binary} exited {r.returncode}: {r.stdout}{r.stderr}"
    return open(out, "rb").read()


def _frames(blob):
    n = W * H * 4 * FRAMES
    return np.frombuffer(blob[:n], np.uint8).reshape(FRAMES, H, W, 4), blob[n:n + 4]


def test_oracle_program_matches_reference(tmp_path):
    if not os.path.exists(REF_LIB):
        pytest.skip("reference library not built (oracle/_ref/libmtgl_ref.so)")
    _build("oracle")
    _build("ref")
    a = _run("headless_oracle", str(tmp_path / "oracle.bin"))
    b = _run("headless_ref", str(tmp_path / "ref.bin"))
    assert len(b) == W * H * 4 * FRAMES + 4
    fa, ea = _frames(a)
    fb, eb = _frames(b)
    for k in range(FRAMES):
        assert np.array_equal(fa[k], fb[k]), f"frame {k}: {(fa[k] != fb[k]).any(-1).sum()} pixels differ"
    assert ea == eb
    # the frames are not trivial: several colours each, the textured one many
    assert all(len(np.unique(fb[k].reshape(-1, 4), axis=0)) >= 4 for k in range(FRAMES))
    assert len(np.unique(fb[5].reshape(-1, 4), axis=0)) > 500


def _present_rows(blob):
    n = W * H * 4 * FRAMES + 4
    pad = np.frombuffer(blob[n:], np.uint8).reshape(H, 1024)
    return pad


def test_oracle_present_is_the_flipped_readback(tmp_path):
    """mtgl_present's rows are the top-down colour plane: glReadPixels' last
    frame flipped, with the row padding left untouched."""
    _build("oracle")
    a = _run("headless_oracle", str(tmp_path / "oracle.bin"))
    fa, _ = _frames(a)
    pad = _present_rows(a)
    assert np.array_equal(pad[:, :W * 4].reshape(H, W, 4), fa[FRAMES - 1][::-1])
    assert (pad[:, W * 4:] == 0xCD).all()


@pytest.mark.gpu
def test_device_program_matches_oracle(tmp_path):
    _build("oracle")
    _build("mtgl")
    a = _run("headless_oracle", str(tmp_path / "oracle.bin"))
    print("running the program linked with -lmtgl", flush=True)
    b = _run("headless_mtgl", str(tmp_path / "mtgl.bin"))
    assert len(a) == len(b)
    fa, ea = _frames(a)
    fb, eb = _frames(b)
    assert ea == eb
    for k in range(FRAMES):
        d = np.abs(fa[k].astype(np.int16) - fb[k].astype(np.int16)).max(-1)
        hist = np.bincount(d.reshape(-1), minlength=256)
        print(f"frame {k}: max-diff histogram {dict((i, int(v)) for i, v in enumerate(hist) if v)}", flush=True)
        assert hist[2:].sum() == 0 and hist[1] <= max(1, W * H // 10000), f"frame {k} differs"
    assert np.array_equal(_present_rows(a), _present_rows(b))


SDL_STUB = """
/* declarations of the SDL2 calls the glue makes (signatures of SDL 2.x),
   enough to type-check include/mytinygl/sdl.h where SDL2 is not installed */
#include <stdint.h>
typedef struct SDL_Window SDL_Window;
typedef struct SDL_Renderer SDL_Renderer;
typedef struct SDL_Texture SDL_Texture;
typedef struct SDL_Rect { int x, y, w, h; } SDL_Rect;
typedef uint32_t Uint32;
#define SDL_INIT_VIDEO 0x20u
#define SDL_WINDOWPOS_CENTERED 0x2FFF0000u
#define SDL_WINDOW_SHOWN 0x4u
#define SDL_RENDERER_ACCELERATED 0x2u
#define SDL_RENDERER_PRESENTVSYNC 0x4u
#define SDL_PIXELFORMAT_ABGR8888 0x16762004u
#define SDL_TEXTUREACCESS_STREAMING 1
int SDL_Init(Uint32 flags);
void SDL_Quit(void);
SDL_Window *SDL_CreateWindow(const char *title, int x, int y, int w, int h, Uint32 flags);
void SDL_DestroyWindow(SDL_Window *window);
SDL_Renderer *SDL_CreateRenderer(SDL_Window *window, int index, Uint32 flags);
void SDL_DestroyRenderer(SDL_Renderer *renderer);
SDL_Texture *SDL_CreateTexture(SDL_Renderer *renderer, Uint32 format, int access, int w, int h);
void SDL_DestroyTexture(SDL_Texture *texture);
int SDL_LockTexture(SDL_Texture *texture, const SDL_Rect *rect, void **pixels, int *pitch);
void SDL_UnlockTexture(SDL_Texture *texture);
int SDL_RenderCopy(SDL_Renderer *renderer, SDL_Texture *texture, const SDL_Rect *src, const SDL_Rect *dst);
void SDL_RenderPresent(SDL_Renderer *renderer);
"""

SDL_USER = """
#include <GL/gl.h>
#include <mytinygl/sdl.h>
int main(void)
{
    if (mtgl_init("demo", 320, 240) != 0) return 1;
    glClear(GL_COLOR_BUFFER_BIT);
    mtgl_swap();
    mtgl_destroy();
    return 0;
}
"""


def test_sdl_glue_type_checks(tmp_path):
    """include/mytinygl/sdl.h compiles as C99 with -Wall -Werror against the
    SDL2 signatures it calls (SDL2 itself is not installed here)."""
    (tmp_path / "SDL2").mkdir()
    (tmp_path / "SDL2" / "SDL.h").write_text(SDL_STUB)
    (tmp_path / "demo.c").write_text(SDL_USER)
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I", str(tmp_path),
                        "-I", os.path.join(ROOT, "include"), str(tmp_path / "demo.c")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
