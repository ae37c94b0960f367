/*
 * headless_frames.c - a plain C program against the GL 1.5 drop-in API.
 *
 * It is written the way the reference's testbed programs use the library
 * (testbed/1.5-0-vbo-test.c:116-199, 1.0-6-zbuffer-test.c, 1.0-13-displaylist-
 * test.c), with the window replaced by glReadPixels: it creates a context
 * (src/mytinygl.h:227-233), renders a few frames -- a VBO cube drawn as
 * GL_QUADS with depth test, a client-array triangle, a display list of a lit
 * fan, a mipmapped textured quad with fog, blended quads with a stencil mask --
 * and writes every frame's glReadPixels(GL_RGBA) bytes, bottom row first, to
 * the file named by argv[1].  Built with -DWITH_PRESENT it also writes the
 * last frame through mtgl_present (top row first, padded pitch) after the
 * frames.  The same source links against libmtgl.so, the CPU oracle and the
 * reference build, whose outputs must be equal byte for byte.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <GL/gl.h>

typedef struct GLState GLState;
GLState *gl_create_context(int32_t width, int32_t height);
void gl_destroy_context(GLState *ctx);
void gl_make_current(GLState *ctx);
#ifdef WITH_PRESENT
int mtgl_present(GLState *ctx, void *pixels, int32_t pitch);
#endif

#define W 200
#define H 150

static const GLfloat cube[] = {
    /* x, y, z, r, g, b: four vertices per face */
    -1, -1, 1, 1, 0, 0,   1, -1, 1, 1, 0, 0,   1, 1, 1, 1, 0, 0,   -1, 1, 1, 1, 0, 0,
    -1, -1, -1, 0, 1, 0,  -1, 1, -1, 0, 1, 0,  1, 1, -1, 0, 1, 0,  1, -1, -1, 0, 1, 0,
    -1, 1, -1, 0, 0, 1,   -1, 1, 1, 0, 0, 1,   1, 1, 1, 0, 0, 1,   1, 1, -1, 0, 0, 1,
    -1, -1, -1, 1, 1, 0,  1, -1, -1, 1, 1, 0,  1, -1, 1, 1, 1, 0,  -1, -1, 1, 1, 1, 0,
    1, -1, -1, 1, 0, 1,   1, 1, -1, 1, 0, 1,   1, 1, 1, 1, 0, 1,   1, -1, 1, 1, 0, 1,
    -1, -1, -1, 0, 1, 1,  -1, -1, 1, 0, 1, 1,  -1, 1, 1, 0, 1, 1,  -1, 1, -1, 0, 1, 1,
};

static const GLfloat tri[] = {
    0.0f, 0.6f, 0.0f, 1, 0.5f, 0,
    -0.6f, -0.4f, 0.0f, 0, 0.5f, 1,
    0.6f, -0.4f, 0.0f, 0.5f, 1, 0,
};

static uint8_t frame_px[W * H * 4];

static int dump(FILE *f)
{
    memset(frame_px, 0xCD, sizeof(frame_px));
    glReadPixels(0, 0, W, H, GL_RGBA, GL_UNSIGNED_BYTE, frame_px);
    return fwrite(frame_px, 1, sizeof(frame_px), f) == sizeof(frame_px) ? 0 : -1;
}

static void projection(void)
{
    glMatrixMode(GL_PROJECTION);
    glLoadIdentity();
    const float aspect = (float)W / (float)H;
    glFrustum(-aspect * 0.1, aspect * 0.1, -0.1, 0.1, 0.1, 100.0);
    glMatrixMode(GL_MODELVIEW);
    glLoadIdentity();
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s OUT\n", argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "wb");
    if (!f) return 2;
    GLState *ctx = gl_create_context(W, H);
    if (!ctx) {
        fprintf(stderr, "gl_create_context failed\n");
        return 3;
    }
    gl_make_current(ctx);
    glViewport(0, 0, W, H);
    projection();
    glEnable(GL_DEPTH_TEST);
    glDepthFunc(GL_LESS);
    glClearColor(0.1f, 0.1f, 0.2f, 1.0f);

    GLuint vbo = 0;
    glGenBuffers(1, &vbo);
    glBindBuffer(GL_ARRAY_BUFFER, vbo);
    glBufferData(GL_ARRAY_BUFFER, sizeof(cube), cube, GL_STATIC_DRAW);
    glBindBuffer(GL_ARRAY_BUFFER, 0);

    /* frames 0-2: the rotating VBO cube */
    for (int k = 0; k < 3; k++) {
        glClear(GL_COLOR_BUFFER_BIT | GL_DEPTH_BUFFER_BIT);
        glLoadIdentity();
        glTranslatef(0.0f, 0.0f, -4.0f);
        glRotatef(20.0f + 37.0f * (float)k, 1.0f, 1.0f, 0.0f);
        glBindBuffer(GL_ARRAY_BUFFER, vbo);
        glEnableClientState(GL_VERTEX_ARRAY);
        glEnableClientState(GL_COLOR_ARRAY);
        glVertexPointer(3, GL_FLOAT, 6 * sizeof(GLfloat), (void *)0);
        glColorPointer(3, GL_FLOAT, 6 * sizeof(GLfloat), (void *)(3 * sizeof(GLfloat)));
        glDrawArrays(GL_QUADS, 0, 24);
        glDisableClientState(GL_VERTEX_ARRAY);
        glDisableClientState(GL_COLOR_ARRAY);
        glBindBuffer(GL_ARRAY_BUFFER, 0);
        if (dump(f)) return 4;
    }

    /* frame 3: a client-array triangle over the cube's last frame */
    glLoadIdentity();
    glTranslatef(0.0f, 0.0f, -2.0f);
    glRotatef(30.0f, 0.0f, 0.0f, 1.0f);
    glEnableClientState(GL_VERTEX_ARRAY);
    glEnableClientState(GL_COLOR_ARRAY);
    glVertexPointer(3, GL_FLOAT, 6 * sizeof(GLfloat), tri);
    glColorPointer(3, GL_FLOAT, 6 * sizeof(GLfloat), tri + 3);
    glDrawArrays(GL_TRIANGLES, 0, 3);
    glDisableClientState(GL_VERTEX_ARRAY);
    glDisableClientState(GL_COLOR_ARRAY);
    if (dump(f)) return 4;

    /* frame 4: a display list of a lit triangle fan, called twice */
    GLuint list = glGenLists(1);
    glNewList(list, GL_COMPILE);
    glBegin(GL_TRIANGLE_FAN);
    glNormal3f(0.0f, 0.0f, 1.0f);
    glVertex3f(0.0f, 0.0f, 0.0f);
    for (int i = 0; i <= 12; i++) {
        const float a = (float)i * 0.5235988f;
        const float c = 1.0f - a * a / 2.0f + a * a * a * a / 24.0f, s = a - a * a * a / 6.0f;
        glNormal3f(0.3f * c, 0.3f * s, 0.9f);
        glVertex3f(0.8f * c, 0.8f * s, 0.0f);
    }
    glEnd();
    glEndList();
    glClear(GL_COLOR_BUFFER_BIT | GL_DEPTH_BUFFER_BIT);
    glEnable(GL_LIGHTING);
    glEnable(GL_LIGHT0);
    glEnable(GL_COLOR_MATERIAL);
    glColor3f(0.9f, 0.6f, 0.2f);
    glLoadIdentity();
    glTranslatef(-0.6f, 0.0f, -3.0f);
    glCallList(list);
    glColor3f(0.2f, 0.6f, 0.9f);
    glLoadIdentity();
    glTranslatef(0.7f, 0.1f, -3.5f);
    glRotatef(25.0f, 0.0f, 1.0f, 0.0f);
    glCallList(list);
    glDisable(GL_LIGHTING);
    glDisable(GL_COLOR_MATERIAL);
    if (dump(f)) return 4;

    /* frame 5: a mipmapped, fogged, perspective textured quad */
    static uint32_t texels[64 * 64];
    for (int y = 0; y < 64; y++)
        for (int x = 0; x < 64; x++) {
            const uint32_t v = ((x / 8 + y / 8) & 1) ? 0xFF20E0F0u : 0xFF402010u;
            texels[y * 64 + x] = v ^ (uint32_t)((x * 7 + y * 13) & 0x1F);
        }
    GLuint tex = 0;
    glGenTextures(1, &tex);
    glBindTexture(GL_TEXTURE_2D, tex);
    glTexImage2D(GL_TEXTURE_2D, 0, GL_RGBA, 64, 64, 0, GL_RGBA, GL_UNSIGNED_BYTE, texels);
    glTexParameteri(GL_TEXTURE_2D, GL_TEXTURE_MIN_FILTER, GL_LINEAR_MIPMAP_LINEAR);
    glTexParameteri(GL_TEXTURE_2D, GL_TEXTURE_MAG_FILTER, GL_LINEAR);
    glEnable(GL_TEXTURE_2D);
    glEnable(GL_FOG);
    glFogi(GL_FOG_MODE, GL_LINEAR);
    glFogf(GL_FOG_START, 2.0f);
    glFogf(GL_FOG_END, 12.0f);
    glHint(GL_PERSPECTIVE_CORRECTION_HINT, GL_NICEST);
    glClear(GL_COLOR_BUFFER_BIT | GL_DEPTH_BUFFER_BIT);
    glLoadIdentity();
    glTranslatef(0.0f, -0.8f, -2.0f);
    glRotatef(-75.0f, 1.0f, 0.0f, 0.0f);
    glColor3f(1.0f, 1.0f, 1.0f);
    glBegin(GL_QUADS);
    glTexCoord2f(0.0f, 0.0f); glVertex3f(-2.0f, 0.0f, 0.0f);
    glTexCoord2f(6.0f, 0.0f); glVertex3f(2.0f, 0.0f, 0.0f);
    glTexCoord2f(6.0f, 12.0f); glVertex3f(2.0f, 10.0f, 0.0f);
    glTexCoord2f(0.0f, 12.0f); glVertex3f(-2.0f, 10.0f, 0.0f);
    glEnd();
    glDisable(GL_FOG);
    glDisable(GL_TEXTURE_2D);
    if (dump(f)) return 4;

    /* frame 6: blended quads through a stencil mask */
    glClearStencil(0);
    glClear(GL_COLOR_BUFFER_BIT | GL_DEPTH_BUFFER_BIT | GL_STENCIL_BUFFER_BIT);
    glDisable(GL_DEPTH_TEST);
    glEnable(GL_STENCIL_TEST);
    glStencilFunc(GL_ALWAYS, 1, 0xFF);
    glStencilOp(GL_KEEP, GL_KEEP, GL_REPLACE);
    glColorMask(GL_FALSE, GL_FALSE, GL_FALSE, GL_FALSE);
    glLoadIdentity();
    glTranslatef(0.0f, 0.0f, -2.0f);
    glBegin(GL_TRIANGLES);
    glVertex3f(-0.9f, -0.6f, 0.0f); glVertex3f(0.9f, -0.6f, 0.0f); glVertex3f(0.0f, 0.8f, 0.0f);
    glEnd();
    glColorMask(GL_TRUE, GL_TRUE, GL_TRUE, GL_TRUE);
    glStencilFunc(GL_EQUAL, 1, 0xFF);
    glStencilOp(GL_KEEP, GL_KEEP, GL_KEEP);
    glEnable(GL_BLEND);
    glBlendFunc(GL_SRC_ALPHA, GL_ONE_MINUS_SRC_ALPHA);
    for (int i = 0; i < 4; i++) {
        glColor4f(0.25f * (float)i, 1.0f - 0.2f * (float)i, 0.5f, 0.4f + 0.1f * (float)i);
        const float o = -0.6f + 0.35f * (float)i;
        glBegin(GL_QUADS);
        glVertex3f(o, -0.5f, 0.0f); glVertex3f(o + 0.6f, -0.5f, 0.0f);
        glVertex3f(o + 0.6f, 0.5f, 0.0f); glVertex3f(o, 0.5f, 0.0f);
        glEnd();
    }
    glDisable(GL_BLEND);
    glDisable(GL_STENCIL_TEST);
    if (dump(f)) return 4;

    const GLenum err = glGetError();
    if (fwrite(&err, sizeof(err), 1, f) != 1) return 4;
#ifdef WITH_PRESENT
    /* the last frame again through the present path, rows padded to 1024 bytes */
    static uint8_t present_px[H * 1024];
    memset(present_px, 0xCD, sizeof(present_px));
    if (mtgl_present(ctx, present_px, 1024) != 0) return 5;
    if (fwrite(present_px, 1, sizeof(present_px), f) != sizeof(present_px)) return 4;
#endif
    glDeleteLists(list, 1);
    glDeleteTextures(1, &tex);
    glDeleteBuffers(1, &vbo);
    gl_make_current(NULL);
    gl_destroy_context(ctx);
    fclose(f);
    printf("headless_frames: 7 frames of %dx%d written (glGetError %#x)\n", W, H, (unsigned)err);
    return 0;
}
