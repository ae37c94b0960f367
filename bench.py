#!/usr/bin/env python3
"""Throughput of the MI355X per-draw pipeline on the 1080p textured-sphere frame.

One step = one frame through the GL API of libmtgl.so: glClear of colour and
depth, then glDrawElements of the 128x128 grid sphere (32,768 triangles,
98,304 indices) from HBM-resident buffer objects with a 256x256
GL_LINEAR_MIPMAP_LINEAR texture, perspective-correct interpolation, linear fog
and one light, at 1920x1080.  Inputs (buffers, texture) are uploaded before the
timed region.

Reported: Mfragments/s (fragments = (triangle, pixel) pairs passing coverage,
counted by the device), frames/s, the roofline of the fragment kernel
(algorithmic bytes per launch / kernel time from HIP events on the context's
stream) and a CPU baseline: the reference renderer compiled strict, one core,
on a bounded sample of the same frame.

"batched": the 512-scene workload (one context per scene) issued per draw and
replayed from one mtgl_batch (a captured graph of two launches over all
scenes), with its own bounded CPU sample.

"differentiable": forward + L2 loss + adjoint over 64 captured views
(mytinygl_amd/diff.py), with the oracle's CPU restatement timed on two views.

"draw_pixels": glDrawPixels of a 1080p RGBA image with depth test and
blending, per call and per kernel, with the oracle timed on five calls.

"lines_points": 100,000 width-2 lines and 100,000 blended size-3 points per
1080p frame, with the oracle timed on two frames.

"display_lists": 64 replays of an icosphere display list per 1080p frame,
against display-list-free draws from client arrays and from a buffer object.

"configs": the icosphere at 640x480 and the 200,000-triangle overdraw
workload at 3840x2160 (blending, stencil, culling), each next to the
reference on one core.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import threading
import time

import numpy as np

try:
    # torch's HIP runtime must be the process's first: loaded after libmtgl.so
    # it finds no GPU, loaded before it serves both
    import torch  # noqa: F401
except ImportError:
    torch = None

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mytinygl_amd.gl import GLLibrary, ENUMS as E  # noqa: E402
from mytinygl_amd import scenes as S  # noqa: E402

HBM_PEAK_GBS = 8000.0
W, H = 1920, 1080
WORKLOAD = "textured_sphere_1920x1080"
TIMING_EVERY = 8   # fragment-stage launches per one timed with HIP events


def frame(g, n):
    g.glClear(E["GL_COLOR_BUFFER_BIT"] | E["GL_DEPTH_BUFFER_BIT"])
    S.draw_textured_sphere(g, n)


LDS_PEAK_TBS = 128 * 256 * 2.4e9 / 1e12   # 128 B/clk per CU, 256 CUs, 2.4 GHz


def profile_figures():
    """Figures of the committed rocprofv3 summary of the fragment kernel on
    this workload (PMC counters: HBM bytes per launch from FETCH_SIZE and
    WRITE_SIZE scaled by the copy-kernel calibration, LDS bytes and the traced
    kernel time, VALU busy and lane use, occupancy), {} when absent."""
    path = os.path.join(ROOT, "profiles", "fragment_stage_textured_sphere_1080p.md")
    if not os.path.exists(path):
        return {}
    txt = open(path).read()
    out = {}
    for key in ("traffic_bytes_per_launch", "lds_bytes_per_launch", "lds_kernel_us", "valu_busy",
                "valu_lane_utilization", "occupancy"):
        m = re.search(rf"^{key}:\s*([0-9.eE+]+)", txt, re.M)
        if m:
            out[key] = float(m.group(1))
    return out


def cpu_library():
    """(library, loader kind, baseline kind) of the CPU baselines: the
    reference renderer built strict from its own sources under oracle/_ref
    when present, else the oracle restatement."""
    ref = os.path.join(ROOT, "oracle", "_ref", "libmtgl_ref.so")
    if os.path.exists(ref):
        return ref, "reference", "reference"
    port = os.path.join(ROOT, "oracle", "_build", "liboracle_gl.so")
    return (port, "oracle", "port") if os.path.exists(port) else (None, None, None)


def cpu_baseline(budget_s: float):
    """Frames of the same workload on the reference renderer (strict build, one
    thread) until budget_s elapses; falls back to the oracle restatement when
    the reference library is not present."""
    ref = os.path.join(ROOT, "oracle", "_ref", "libmtgl_ref.so")
    kind = "reference"
    if not os.path.exists(ref):
        ref = os.path.join(ROOT, "oracle", "_build", "liboracle_gl.so")
        kind = "port"
    if not os.path.exists(ref):
        return None
    g = GLLibrary(ref, "reference" if kind == "reference" else "oracle")
    ctx = g.create_context(W, H)
    g.glClearColor(0.0, 0.0, 0.0, 1.0)
    n = S.setup_textured_sphere(g, W, H)
    frame(g, n)                       # warm-up (lazy level-1 build in the reference)
    frames = 0
    t0 = time.perf_counter()
    while True:
        frame(g, n)
        frames += 1
        el = time.perf_counter() - t0
        if el >= budget_s or frames >= 200:
            break
    g.destroy_context(ctx)
    g.release()
    return {"frames": frames, "seconds": el, "kind": kind}


def configs_bench(frames: dict, cpu_budget_s: float):
    """The other two single-frame workloads: the icosphere at 640x480 (1,280
    triangles, one light, depth test; about 51k fragments) and the overdraw
    workload at 3840x2160 (200,000 random triangles, back faces culled,
    SRC_ALPHA blending, stencil GREATER/INCR, depth test; about 108.7M
    fragments).  One frame = clear of every plane + one glDrawArrays from
    client arrays (staged per draw, as the reference reads them).  The
    reference (strict build, one thread) renders the same frames for a
    bounded time."""
    cases = {"icosphere_640x480": (S.setup_icosphere, S.draw_icosphere, 640, 480),
             "overdraw_3840x2160": (S.setup_overdraw, S.draw_overdraw, 3840, 2160)}
    ref = os.path.join(ROOT, "oracle", "_ref", "libmtgl_ref.so")
    kind = "reference"
    if not os.path.exists(ref):
        ref, kind = os.path.join(ROOT, "oracle", "_build", "liboracle_gl.so"), "port"
    bits = E["GL_COLOR_BUFFER_BIT"] | E["GL_DEPTH_BUFFER_BIT"] | E["GL_STENCIL_BUFFER_BIT"]
    out = {}
    for name, (setup, draw, w, h) in cases.items():
        if name not in frames:
            continue
        g = GLLibrary(os.path.join(ROOT, "mytinygl_amd", "libmtgl.so"), "mtgl")
        ctx = g.create_context(w, h)
        g.glClearColor(0.0, 0.0, 0.0, 1.0)
        n = setup(g, w, h)
        for _ in range(2):
            g.glClear(bits)
            draw(g, n)
        g.glFinish()
        g.stats_reset(ctx)
        nf = frames[name]
        g.set_timing(ctx, True, every=TIMING_EVERY if nf >= 10 * TIMING_EVERY else 1)
        t0 = time.perf_counter()
        for _ in range(nf):
            g.glClear(bits)
            draw(g, n)
        g.glFinish()
        sec = time.perf_counter() - t0
        st = g.stats(ctx)
        g.set_timing(ctx, False)
        g.destroy_context(ctx)
        g.release()
        frags = st.fragments / nf
        rec = {"frames": nf, "ms_per_frame": round(sec * 1e3 / nf, 4), "fragments_per_frame": int(frags),
               "mfragments_s": round(frags * nf / sec / 1e6, 1),
               "fragment_stage_ms": round(st.raster_ms / max(1, st.raster_launches), 5),
               "algorithmic_bytes_per_frame": int(st.frag_bytes / nf)}
        if st.raster_ms > 0:
            rec["fragment_stage_gb_s"] = round(st.frag_bytes / nf / (rec["fragment_stage_ms"] * 1e-3) / 1e9, 1)
        if cpu_budget_s > 0 and os.path.exists(ref):
            c = GLLibrary(ref, "reference" if kind == "reference" else "oracle")
            cctx = c.create_context(w, h)
            c.glClearColor(0.0, 0.0, 0.0, 1.0)
            cn = setup(c, w, h)
            cf, t0 = 0, time.perf_counter()
            while True:
                c.glClear(bits)
                draw(c, cn)
                cf += 1
                el = time.perf_counter() - t0
                if el >= cpu_budget_s or cf >= 200:
                    break
            c.destroy_context(cctx)
            c.release()
            rec["cpu_baseline"] = {"ms_per_frame": round(el * 1e3 / cf, 2), "mfragments_s": round(frags * cf / el / 1e6, 2),
                                   "cores": 1, "kind": kind,
                                   "sample": f"{cf} frames of the same workload in {el:.1f} s, one thread"}
            rec["speedup_vs_cpu"] = round((el / cf) / (sec / nf), 1)
        out[name] = rec
    return out


def batch_bench(launches: int, per_draw_reps: int, cpu_seconds: float):
    """The 512-scene batched workload: the same frames (one clear and one
    glDrawElements per context) issued per draw, context after context, and
    replayed from one batch.  Times are whole 512-scene frame sets with the
    GPU drained on both sides."""
    g = GLLibrary(os.path.join(ROOT, "mytinygl_amd", "libmtgl.so"), "mtgl")
    specs = S.batch_specs()
    run = S.BatchRun(g, S.batch_items(specs))
    try:
        run.frames_immediate()
        run.finish()
        for ctx in run.ctxs:
            g.stats_reset(ctx)
        run.frames_immediate()
        run.finish()
        frags_scene = [int(g.stats(ctx).fragments) for ctx in run.ctxs]
        frags = sum(frags_scene)
        t0 = time.perf_counter()
        for _ in range(per_draw_reps):
            run.frames_immediate()
        run.finish()
        per_draw_ms = (time.perf_counter() - t0) * 1e3 / per_draw_reps
        b = run.record(autotune=True)
        info = b.info()
        for _ in range(3):
            b.launch()
        if b.sync() != 0:
            raise RuntimeError("batched launch incomplete")
        b.stats_reset()
        b.launch()
        b.sync()
        frags_batch = int(b.stats().fragments)
        t0 = time.perf_counter()
        for _ in range(launches):
            b.launch()
        rc = b.sync()
        batched_ms = (time.perf_counter() - t0) * 1e3 / launches
        if rc != 0:
            raise RuntimeError("batched launch incomplete")
    finally:
        run.close()
        g.release()
    out = {
        "workload": "512 scenes, one context each: square framebuffers 64..1024 px (log-uniform), grid spheres "
                    "of 98..48,672 triangles (log-uniform), every other one textured 64..512 px "
                    "GL_LINEAR_MIPMAP_LINEAR; one clear + one glDrawElements per scene",
        "scenes": len(specs), "triangles": int(sum(sp["triangles"] for sp in specs)),
        "pixels": int(sum(sp["w"] * sp["h"] for sp in specs)), "fragments": frags,
        "fragments_batched": frags_batch,
        "per_draw_ms": round(per_draw_ms, 3), "batched_ms": round(batched_ms, 3),
        "speedup": round(per_draw_ms / batched_ms, 2),
        "mfragments_s_per_draw": round(frags / per_draw_ms / 1e3, 2),
        "mfragments_s_batched": round(frags / batched_ms / 1e3, 2),
        "scene_frames_s_batched": round(len(specs) / batched_ms * 1e3, 1),
        "launch": {"rounds": info.rounds, "heavy_bin": info.heavy_bin, "grid_div": info.grid_div,
                   "tune_ms": [round(info.tune_ms[i], 3) for i in range(info.tune_candidates)]},
    }
    if cpu_seconds > 0:
        out["cpu_baseline"] = batch_cpu_baseline(specs, frags_scene, cpu_seconds)
    return out


def batch_cpu_baseline(specs, frags_scene, budget_s: float):
    """The reference renderer (strict build) on the batched workload with one
    context per host thread, as the reference allows (its current context is
    thread-local): the scenes are dealt to the threads by fragment count and
    set up, then every thread renders one frame of each of its scenes, all
    threads at once; the time runs from the start of the first frame to the
    end of the last.  The threads spend their time inside the library, which
    ctypes calls without the interpreter lock.  budget_s > 0 enables it."""
    ref = os.path.join(ROOT, "oracle", "_ref", "libmtgl_ref.so")
    kind = "reference"
    if not os.path.exists(ref):
        ref = os.path.join(ROOT, "oracle", "_build", "liboracle_gl.so")
        kind = "port"
    if not os.path.exists(ref) or budget_s <= 0:
        return None
    g = GLLibrary(ref, kind if kind == "reference" else "oracle")
    nthreads = max(1, min(len(os.sched_getaffinity(0)), 16))
    parts = [[] for _ in range(nthreads)]
    loads = [0] * nthreads
    for i in sorted(range(len(specs)), key=lambda i: -frags_scene[i]):
        k = loads.index(min(loads))
        parts[k].append(i)
        loads[k] += frags_scene[i] + 1
    start, done = threading.Barrier(nthreads + 1), threading.Barrier(nthreads + 1)
    errors = []

    def work(k):
        ctxs = []
        try:
            for i in parts[k]:
                ctx = g.create_context(specs[i]["w"], specs[i]["h"])
                ctxs.append((ctx, S.setup_batch_scene(g, specs[i])))
        except Exception as e:   # reported after the barriers, so no thread waits forever
            errors.append(e)
        start.wait()
        try:
            for ctx, n in ctxs:
                g.gl_make_current(ctx)
                S.frame_batch_scene(g, n)
        except Exception as e:
            errors.append(e)
        done.wait()
        for ctx, _ in ctxs:
            g.destroy_context(ctx)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(nthreads)]
    for t in threads:
        t.start()
    start.wait()
    t0 = time.perf_counter()
    done.wait()
    el = time.perf_counter() - t0
    for t in threads:
        t.join()
    g.release()
    if errors:
        raise errors[0]
    frags = int(sum(frags_scene))
    return {"value": round(frags / el / 1e6, 4), "unit": "Mfragments/s", "cores": nthreads, "kind": kind,
            "cpu": cpu_model(),
            "sample": f"all {len(specs)} scenes ({frags} fragments), one context per thread on {nthreads} "
                      f"threads, one frame each in {el:.2f} s"}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def diff_bench(steps: int, warmup: int, cpu_views: int):
    """The differentiable workload: 64 views of 256x256 of a 100,352-triangle
    grid sphere, 1024x1024 fp32 RGBA texture, 8 lights.  One step = forward
    render of every view, L2 loss against target renders, backward (adjoint)
    to texels, vertex colours and light colours, through the
    torch.autograd.Function of mytinygl_amd/diff.py."""
    import torch
    from mytinygl_amd import diff as D
    g = GLLibrary(os.path.join(ROOT, "mytinygl_amd", "libmtgl.so"), "mtgl")
    P = D.diff_params()
    t0 = time.perf_counter()
    V = D.DiffViews(g, P, 256, 256)
    capture_s = time.perf_counter() - t0
    dev = torch.device("cuda:0")

    def t(a):
        return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)

    T = D.perturbed(P, seed=13)
    with torch.no_grad():
        target = D.render(t(T["vcol"]), t(T["light_rgb"]), t(T["tex"]), V)
    vc, lr, tx = (t(P[k]).requires_grad_() for k in ("vcol", "light_rgb", "tex"))

    def step():
        for p in (vc, lr, tx):
            p.grad = None
        img = D.render(vc, lr, tx, V)
        loss = ((img - target) ** 2).sum()
        loss.backward()

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t0) * 1e3 / steps
    with torch.no_grad():
        for _ in range(warmup):
            D.render(vc, lr, tx, V)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            D.render(vc, lr, tx, V)
        torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - t0) * 1e3 / steps
    covered = int(sum((V.gbuffer(v)["tri"] >= 0).sum() for v in range(V.nviews)))
    V.close()
    g.release()
    out = {
        "workload": "64 views of 256x256 (Fibonacci sphere, distance 3), 224x224 grid sphere (100,352 triangles), "
                    "1024x1024 fp32 RGBA texture, 8 directional lights, bilinear; L2 loss; adjoint to texels, "
                    "vertex colours and light colours",
        "views": 64, "triangles": len(P["idx"]) // 3, "covered_pixels": covered,
        "capture_s": round(capture_s, 3), "step_ms": round(step_ms, 3), "forward_ms": round(fwd_ms, 3),
        "views_per_s": round(64 / step_ms * 1e3, 1),
        "mpixels_s": round(covered / step_ms / 1e3, 2),
    }
    if cpu_views > 0:
        out["cpu_baseline"] = diff_cpu_baseline(P, T, cpu_views)
        if out["cpu_baseline"]:
            out["speedup_vs_cpu"] = round(out["cpu_baseline"]["step_ms"] / step_ms, 1)
    return out


def diff_cpu_baseline(P, T, nviews: int):
    """The oracle's differentiable path (CPU restatement, one thread): forward,
    loss gradient and backward over the first nviews views, scaled to 64."""
    from mytinygl_amd import diff as D
    lib = os.path.join(ROOT, "oracle", "_build", "liboracle_gl.so")
    if not os.path.exists(lib):
        return None
    g = GLLibrary(lib, "oracle")
    Ps = dict(P, eyes=P["eyes"][:nviews])
    V = D.DiffViews(g, Ps, 256, 256)
    target = V.forward_np(T["vcol"], T["light_rgb"], T["tex"])
    t0 = time.perf_counter()
    img = V.forward_np(P["vcol"], P["light_rgb"], P["tex"])
    V.backward_np(P["vcol"], P["light_rgb"], P["tex"], 2.0 * (img - target))
    sec = time.perf_counter() - t0
    V.close()
    g.release()
    return {"step_ms": round(sec * 1e3 * 64 / nviews, 1), "unit": "ms per 64-view step", "cores": 1, "kind": "port",
            "cpu": cpu_model(), "extrapolated": True, "views_timed": nviews,
            "sample": f"{nviews} of the 64 views, forward + backward in {sec:.2f} s, one thread, scaled to 64"}


def pixels_bench(reps: int, cpu_reps: int):
    """glDrawPixels of a 1920x1080 GL_RGBA image over a 1080p framebuffer with
    the depth test (GL_LEQUAL, depth writes on) and blending (GL_SRC_ALPHA,
    GL_ONE_MINUS_SRC_ALPHA) enabled.  Each call copies the host image to the
    device (8.3 MB over PCIe) and runs k_draw_pixels; the kernel's own time
    comes from events on the context's stream.  Algorithmic bytes per pixel:
    4 source + 4 depth read + 4 depth write + 4 colour read + 4 colour write."""
    rng = np.random.default_rng(S.SEED_BASE + 8)
    img = rng.integers(0, 256, size=W * H * 4, dtype=np.uint8)

    def setup(g):
        g.glMatrixMode(E["GL_PROJECTION"])
        g.glLoadIdentity()
        g.glOrtho(0.0, float(W), 0.0, float(H), -1.0, 1.0)
        g.glMatrixMode(E["GL_MODELVIEW"])
        g.glLoadIdentity()
        g.glRasterPos2i(0, 0)
        g.glEnable(E["GL_DEPTH_TEST"])
        g.glDepthFunc(E["GL_LEQUAL"])
        g.glEnable(E["GL_BLEND"])
        g.glBlendFunc(E["GL_SRC_ALPHA"], E["GL_ONE_MINUS_SRC_ALPHA"])
        g.glClearColor(0.2, 0.3, 0.4, 1.0)
        g.glClear(E["GL_COLOR_BUFFER_BIT"] | E["GL_DEPTH_BUFFER_BIT"])

    g = GLLibrary(os.path.join(ROOT, "mytinygl_amd", "libmtgl.so"), "mtgl")
    ctx = g.create_context(W, H)
    setup(g)
    ptr = g.keep(img)
    for _ in range(3):
        g.glDrawPixels(W, H, E["GL_RGBA"], E["GL_UNSIGNED_BYTE"], ptr)
    g.glFinish()
    g.stats_reset(ctx)
    g.set_timing(ctx, True)
    t0 = time.perf_counter()
    for _ in range(reps):
        g.glDrawPixels(W, H, E["GL_RGBA"], E["GL_UNSIGNED_BYTE"], ptr)
    g.glFinish()
    sec = time.perf_counter() - t0
    st = g.stats(ctx)
    g.set_timing(ctx, False)
    g.destroy_context(ctx)
    g.release()
    k_ms = st.raster_ms / max(1, st.raster_launches)
    px = W * H
    achieved = 20.0 * px / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
    out = {"workload": "glDrawPixels 1920x1080 GL_RGBA, depth test GL_LEQUAL with writes, blend SRC_ALPHA / "
                       "ONE_MINUS_SRC_ALPHA, 1080p framebuffer",
           "calls": reps, "ms_per_call": round(sec * 1e3 / reps, 4),
           "mpixels_s_per_call": round(px * reps / sec / 1e6, 1),
           "kernel_ms": round(k_ms, 5), "kernel_launches": int(st.raster_launches),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 5), "bytes_per_launch": 20 * px,
                        "kernel": "k_draw_pixels"},
           "note": "ms_per_call includes the host-to-device copy of the image (pageable host memory -> HBM, "
                   "after the context's queued work: images of 1 MB and more are not staged); the kernel "
                   "reads an image already in HBM"}
    lib, lk, kind = cpu_library()
    if cpu_reps > 0 and lib:
        o = GLLibrary(lib, lk)
        octx = o.create_context(W, H)
        setup(o)
        optr = o.keep(img)
        t0 = time.perf_counter()
        for _ in range(cpu_reps):
            o.glDrawPixels(W, H, E["GL_RGBA"], E["GL_UNSIGNED_BYTE"], optr)
        csec = time.perf_counter() - t0
        o.destroy_context(octx)
        o.release()
        out["cpu_baseline"] = {"ms_per_call": round(csec * 1e3 / cpu_reps, 3), "cores": 1, "kind": kind,
                               "cpu": cpu_model(),
                               "sample": f"{cpu_reps} calls of the same workload, one thread"}
    return out


def lines_bench(frames: int, cpu_frames: int, n_lines: int = 100_000, n_points: int = 100_000):
    """Points and lines at 1080p: one frame = clear, glDrawArrays(GL_LINES) of
    n_lines width-2 segments and glDrawArrays(GL_POINTS) of n_points blended
    size-3 points from a buffer object, depth test on.  Fragments = the
    pixels the reference's line and point loops visit (steps x width, size
    squared), counted by the device."""
    def run(g, nframes):
        ctx = g.create_context(W, H)
        st = S.setup_lines_points(g, W, H, S.SEED_BASE + 9, n_lines, n_points)

        def frame():
            g.glClear(E["GL_COLOR_BUFFER_BIT"] | E["GL_DEPTH_BUFFER_BIT"])
            S.draw_lines_points(g, st)

        frame()
        g.glFinish()
        frags = 0
        if g.kind != "reference":   # (the reference keeps no counters)
            g.stats_reset(ctx)
            frame()
            g.glFinish()
            frags = g.stats(ctx).fragments
        t0 = time.perf_counter()
        for _ in range(nframes):
            frame()
        g.glFinish()
        sec = time.perf_counter() - t0
        g.destroy_context(ctx)
        g.release()
        return frags, sec

    g = GLLibrary(os.path.join(ROOT, "mytinygl_amd", "libmtgl.so"), "mtgl")
    frags, sec = run(g, frames)
    out = {"workload": f"{n_lines} GL_LINES of width 2 and {n_points} blended GL_POINTS of size 3 at 1920x1080, "
                       "depth test, one buffer object",
           "frames": frames, "ms_per_frame": round(sec * 1e3 / frames, 4), "fragments_per_frame": int(frags),
           "mfragments_s": round(frags * frames / sec / 1e6, 1)}
    lib, lk, kind = cpu_library()
    if cpu_frames > 0 and lib:
        _, osec = run(GLLibrary(lib, lk), cpu_frames)
        out["cpu_baseline"] = {"ms_per_frame": round(osec * 1e3 / cpu_frames, 2), "cores": 1, "kind": kind,
                               "cpu": cpu_model(), "mfragments_s": round(frags * cpu_frames / osec / 1e6, 2),
                               "sample": f"{cpu_frames} frames of the same workload, one thread"}
    return out


def lists_bench(frames: int, cpu_frames: int, copies: int = 64):
    """Display lists at 1080p: an icosphere (subdivision 3, 1,280 triangles,
    3,840 vertices with per-vertex colours and normals, one light) drawn
    `copies` times per frame under different modelviews, three ways: replays
    of a display list holding it as one glBegin/glEnd block (the records stay
    in a device segment after the first replay), glDrawArrays from client
    arrays (staged and sent on every draw), and glDrawArrays from a buffer
    object (resident)."""
    pos, tri = S.icosphere(3)
    rng = np.random.default_rng(S.SEED_BASE + 10)
    vcol = rng.random((len(pos), 4)).astype(np.float32)
    idx = tri.reshape(-1)
    inter = np.concatenate([pos[idx].astype(np.float32), vcol[idx], pos[idx].astype(np.float32)], 1)   # x y z r g b a nx ny nz
    inter = np.ascontiguousarray(inter, np.float32)
    nvtx = len(idx)

    def run(g, nframes, how):
        ctx = g.create_context(W, H)
        S._perspective(g, W, H, dist=-12.0)
        g.glEnable(E["GL_DEPTH_TEST"])
        g.glEnable(E["GL_LIGHTING"])
        g.glEnable(E["GL_LIGHT0"])
        g.glEnable(E["GL_COLOR_MATERIAL"])
        g.glLightfv(E["GL_LIGHT0"], E["GL_POSITION"], g.keep(np.array([1.0, 1.0, 1.0, 0.0], np.float32)))
        lid = g.glGenLists(1)
        ptr = g.keep(inter)
        if how == "list":
            g.glNewList(lid, E["GL_COMPILE"])
            g.glBegin(E["GL_TRIANGLES"])
            for v in inter:
                g.glColor4f(*v[3:7])
                g.glNormal3f(*v[7:10])
                g.glVertex3f(*v[0:3])
            g.glEnd()
            g.glEndList()
        else:
            off = 0
            if how == "vbo":
                ids = np.zeros(1, np.uint32)
                g.glGenBuffers(1, ids)
                g.glBindBuffer(E["GL_ARRAY_BUFFER"], int(ids[0]))
                g.glBufferData(E["GL_ARRAY_BUFFER"], inter.nbytes, ptr, E["GL_STATIC_DRAW"])
            for arr in ("GL_VERTEX_ARRAY", "GL_COLOR_ARRAY", "GL_NORMAL_ARRAY"):
                g.glEnableClientState(E[arr])
            # byte addresses: the buffer offset 0, or the client array's address
            # (kept alive by keep); the attributes sit at +0, +12 and +28
            base = 0 if how == "vbo" else int(ptr.ctypes.data)
            g.glVertexPointer(3, E["GL_FLOAT"], 40, base)
            g.glColorPointer(4, E["GL_FLOAT"], 40, base + 12)
            g.glNormalPointer(E["GL_FLOAT"], 40, base + 28)
        side = int(np.ceil(np.sqrt(copies)))

        def frame():
            g.glClear(E["GL_COLOR_BUFFER_BIT"] | E["GL_DEPTH_BUFFER_BIT"])
            for k in range(copies):
                g.glPushMatrix()
                g.glTranslatef(2.2 * (k % side - side / 2 + 0.5), 1.4 * (k // side - side / 2 + 0.5), 0.0)
                g.glScalef(0.6, 0.6, 0.6)
                if how == "list":
                    g.glCallList(lid)
                else:
                    g.glDrawArrays(E["GL_TRIANGLES"], 0, nvtx)
                g.glPopMatrix()

        frame()
        g.glFinish()
        counted = g.kind != "reference"   # (the reference keeps no counters)
        if counted:
            g.stats_reset(ctx)
        t0 = time.perf_counter()
        for _ in range(nframes):
            frame()
        g.glFinish()
        sec = time.perf_counter() - t0
        st = g.stats(ctx) if counted else None
        g.destroy_context(ctx)
        g.release()
        return (sec * 1e3 / nframes, st.upload_bytes / nframes if st else 0.0,
                st.fragments / nframes if st else 0.0)

    lib = os.path.join(ROOT, "mytinygl_amd", "libmtgl.so")
    res = {how: run(GLLibrary(lib, "mtgl"), frames, how) for how in ("list", "client_arrays", "vbo")}
    out = {"workload": f"{copies} icospheres (1,280 triangles, 3,840 vertices, one light) per 1080p frame: display "
                       "list replays vs glDrawArrays from client arrays vs from a buffer object",
           "frames": frames, "fragments_per_frame": int(res["list"][2])}
    for how, (ms, up, _) in res.items():
        out[f"ms_per_frame_{how}"] = round(ms, 3)
        out[f"upload_bytes_per_frame_{how}"] = int(up)
    olib, lk, kind = cpu_library()
    if cpu_frames > 0 and olib:
        ms_cpu, _, _ = run(GLLibrary(olib, lk), cpu_frames, "list")
        out["cpu_baseline"] = {"ms_per_frame": round(ms_cpu, 2), "cores": 1, "kind": kind, "cpu": cpu_model(),
                               "sample": f"{cpu_frames} frames of the list workload, one thread"}
    return out


def copy_calibration(nbytes: int, reps: int = 5) -> None:
    """Plain device-to-device copies of nbytes (torch's elementwise copy
    kernel): the reference the HBM counters FETCH_SIZE / WRITE_SIZE are
    scaled by, since that copy moves exactly nbytes in and nbytes out."""
    import torch
    n = nbytes // 4
    x = torch.arange(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-scene batched leg")
    ap.add_argument("--batch-launches", type=int, default=50)
    ap.add_argument("--batch-per-draw", type=int, default=5)
    ap.add_argument("--only-batch", action="store_true", help="run only the batched leg (profiling)")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable leg")
    ap.add_argument("--only-diff", action="store_true", help="run only the differentiable leg (profiling)")
    ap.add_argument("--diff-steps", type=int, default=10)
    ap.add_argument("--no-pixels", action="store_true", help="skip the glDrawPixels leg")
    ap.add_argument("--only-pixels", action="store_true", help="run only the glDrawPixels leg (profiling)")
    ap.add_argument("--pixels-calls", type=int, default=50)
    ap.add_argument("--no-lines", action="store_true", help="skip the points-and-lines leg")
    ap.add_argument("--only-lines", action="store_true", help="run only the points-and-lines leg (profiling)")
    ap.add_argument("--lines-frames", type=int, default=20)
    ap.add_argument("--no-lists", action="store_true", help="skip the display-list leg")
    ap.add_argument("--only-lists", action="store_true", help="run only the display-list leg (profiling)")
    ap.add_argument("--lists-frames", type=int, default=8)
    ap.add_argument("--no-configs", action="store_true", help="skip the icosphere 640x480 and overdraw 4K legs")
    ap.add_argument("--only-configs", action="store_true", help="run only those legs (profiling)")
    ap.add_argument("--only-config", choices=["icosphere_640x480", "overdraw_3840x2160"],
                    help="run only this one of those legs, 3 frames (profiling)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run plain device copies of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU split is a design note, not an implemented path)", file=sys.stderr)
        sys.exit(2)

    if args.only_batch:
        print(json.dumps(batch_bench(args.batch_launches, args.batch_per_draw, 0.0)), flush=True)
        return
    if args.only_diff:
        print(json.dumps(diff_bench(args.diff_steps, 2, 0)), flush=True)
        return
    if args.only_pixels:
        print(json.dumps(pixels_bench(args.pixels_calls, 0)), flush=True)
        return
    if args.only_lines:
        print(json.dumps(lines_bench(args.lines_frames, 0)), flush=True)
        return
    if args.only_lists:
        print(json.dumps(lists_bench(args.lists_frames, 0)), flush=True)
        return
    if args.only_config:
        print(json.dumps(configs_bench({args.only_config: 3}, 0.0)), flush=True)
        return
    if args.only_configs:
        print(json.dumps(configs_bench({"icosphere_640x480": 200, "overdraw_3840x2160": 3}, 0.0)), flush=True)
        return

    g = GLLibrary(os.path.join(ROOT, "mytinygl_amd", "libmtgl.so"), "mtgl")
    ctx = g.create_context(W, H)
    g.glClearColor(0.0, 0.0, 0.0, 1.0)
    n = S.setup_textured_sphere(g, W, H)
    for _ in range(args.warmup):
        frame(g, n)
    g.glFinish()

    # fragments and algorithmic bytes of one frame, counted by the device
    g.stats_reset(ctx)
    frame(g, n)
    st = g.stats(ctx)
    frags_per_frame = st.fragments
    bytes_per_frame = st.frag_bytes

    g.stats_reset(ctx)
    # HIP events bracket one fragment-stage launch in TIMING_EVERY of the
    # timed frames (a pair of events on every launch would add queue time)
    g.set_timing(ctx, True, every=TIMING_EVERY)
    g.glFinish()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame(g, n)
    g.glFinish()
    t1 = time.perf_counter()
    st = g.stats(ctx)
    g.set_timing(ctx, False)
    g.destroy_context(ctx)
    g.release()

    sec = t1 - t0
    ms_per_step = sec * 1e3 / args.steps
    fps = args.steps / sec
    mfrag = frags_per_frame * fps / 1e6
    k_ms = st.raster_ms / max(1, st.raster_launches)
    bytes_per_launch = st.frag_bytes / args.steps   # one fragment-stage launch per frame
    achieved = bytes_per_launch / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0

    prof = profile_figures()
    lds_fig = valu_fig = None
    if "lds_bytes_per_launch" in prof and prof.get("lds_kernel_us"):
        lds_tbs = prof["lds_bytes_per_launch"] / (prof["lds_kernel_us"] * 1e-6) / 1e12
        lds_fig = {"achieved": round(lds_tbs, 3), "peak": round(LDS_PEAK_TBS, 1), "unit": "TB/s",
                   "frac": round(lds_tbs / LDS_PEAK_TBS, 4),
                   "source": "rocprofv3 SQ_INSTS_LDS_LOAD/STORE_BANDWIDTH x 64 B over the traced kernel time "
                             "(profiles/fragment_stage_textured_sphere_1080p.md)"}
    if "valu_busy" in prof:
        valu_fig = {"busy": prof["valu_busy"], "lane_utilization": prof.get("valu_lane_utilization"),
                    "occupancy": prof.get("occupancy"),
                    "source": "rocprofv3 SQ_ACTIVE_INST_VALU, SQ_THREAD_CYCLES_VALU, SQ_WAVE_CYCLES, GRBM_GUI_ACTIVE"}

    cpu = None if args.no_cpu else cpu_baseline(args.cpu_seconds)
    cpu_obj = None
    if cpu:
        cpu_fps = cpu["frames"] / cpu["seconds"]
        cpu_obj = {"value": round(frags_per_frame * cpu_fps / 1e6, 4), "unit": "Mfragments/s", "cores": 1,
                   "kind": cpu["kind"], "cpu": cpu_model(),
                   "sample": f"{cpu['frames']} frames of the same 1080p textured-sphere workload in "
                             f"{cpu['seconds']:.1f} s, one thread ({os.cpu_count()} host CPUs visible)",
                   "frames_per_s": round(cpu_fps, 4)}

    batched = None if args.no_batch else batch_bench(args.batch_launches, args.batch_per_draw,
                                                      0.0 if args.no_cpu else args.cpu_seconds / 2)
    differentiable = None if args.no_diff else diff_bench(args.diff_steps, 2, 0 if args.no_cpu else 2)
    pixels = None if args.no_pixels else pixels_bench(args.pixels_calls, 0 if args.no_cpu else 5)
    lines = None if args.no_lines else lines_bench(args.lines_frames, 0 if args.no_cpu else 2)
    lists = None if args.no_lists else lists_bench(args.lists_frames, 0 if args.no_cpu else 1)
    configs = None if args.no_configs else configs_bench({"icosphere_640x480": 200, "overdraw_3840x2160": 5},
                                                         0.0 if args.no_cpu else 5.0)

    out = {
        "metric": "Mfragments/s at 1080p (textured sphere, one GPU)",
        "value": round(mfrag, 3),
        "unit": "Mfragments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "frames_per_s": round(fps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded grid sphere and value-noise texture)",
        "config": {"workload": WORKLOAD, "width": W, "height": H, "triangles": n // 3,
                   "fragments_per_frame": int(frags_per_frame), "algorithmic_bytes_per_frame": int(bytes_per_frame),
                   "texture": "256x256 RGBA8, GL_LINEAR_MIPMAP_LINEAR/GL_LINEAR, GL_REPEAT",
                   "fog": "GL_LINEAR 1..3"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": prof.get("traffic_bytes_per_launch"),
                     "lds": lds_fig, "valu": valu_fig,
                     "kernel": "fragment stage (k_raster_depth: coverage, in-order depth pass, shading of the last passing fragment)", "kernel_ms": round(k_ms, 5),
                     "launches_timed": int(st.raster_launches), "timed_one_in": TIMING_EVERY,
                     "bytes_per_launch": int(bytes_per_launch)},
        "cpu_baseline": cpu_obj,
        "batched": batched,
        "differentiable": differentiable,
        "draw_pixels": pixels,
        "lines_points": lines,
        "display_lists": lists,
        "configs": configs,
    }
    if cpu_obj:
        out["speedup_vs_cpu"] = round(mfrag / cpu_obj["value"], 1) if cpu_obj["value"] else None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
