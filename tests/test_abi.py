"""The C ABI of libmtgl.so without a GPU: the library loads, exports every
function declared in include/*.h, its enumerants equal the reference's, and
with no device it refuses to create a context instead of falling back to a
CPU path."""
import ctypes as C
import os
import re

import pytest

from conftest import MTGL_LIB, ORACLE_LIB, ROOT
from mytinygl_amd.gl import ENUMS, GL_FUNCTIONS

INCLUDE = os.path.join(ROOT, "include")
REF_HEADER = "/root/reference/include/GL/gl.h"


def declared_functions():
    names = set()
    for dirpath, _, files in os.walk(INCLUDE):
        if os.path.basename(dirpath) == "mytinygl":
            continue   # header-only SDL glue: static inline helpers, nothing exported
        for fn in files:
            if not fn.endswith(".h"):
                continue
            text = open(os.path.join(dirpath, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            text = re.sub(r"//[^\n]*", "", text)
            text = re.sub(r"#define[^\n]*", "", text)
            text = re.sub(r"typedef[^;]*;", "", text)
            for m in re.finditer(r"\b((?:gl|mtgl_|gl_)\w*)\s*\([^;{]*\)\s*;", text):
                names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_gl_api():
    names = declared_functions()
    assert len([n for n in names if n.startswith("gl") and not n.startswith("gl_")]) >= 100
    assert {"gl_create_context", "gl_destroy_context", "gl_make_current", "gl_get_current_context"} <= set(names)
    assert {"mtgl_dev_draw", "mtgl_dev_create", "mtgl_dev_clear"} & set(names)


@pytest.mark.parametrize("path", [MTGL_LIB, ORACLE_LIB], ids=["libmtgl", "oracle"])
def test_exports_every_declared_function(path):
    if not os.path.exists(path):
        pytest.fail(f"{path} not built")
    lib = C.CDLL(path, mode=os.RTLD_LOCAL | os.RTLD_NOW)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"{os.path.basename(path)} lacks {missing}"


def test_binding_covers_header():
    declared = {n for n in declared_functions() if n.startswith("gl") and not n.startswith("gl_")}
    assert set(GL_FUNCTIONS) <= declared


@pytest.mark.skipif(not os.path.exists(REF_HEADER), reason="reference header not present")
def test_enums_match_reference_header():
    ref = {}
    for m in re.finditer(r"#define\s+(GL_\w+)\s+(0x[0-9A-Fa-f]+|\d+)\b", open(REF_HEADER).read()):
        ref.setdefault(m.group(1), int(m.group(2), 0))
    common = set(ref) & set(ENUMS)
    assert len(common) >= 0.95 * len(ref)
    diff = {k: (ENUMS[k], ref[k]) for k in common if ENUMS[k] != ref[k]}
    assert not diff


def test_no_device_no_context():
    """Without a usable GPU, gl_create_context returns NULL (no CPU fallback);
    with one, it returns a context that can be destroyed."""
    lib = C.CDLL(MTGL_LIB, mode=os.RTLD_LOCAL | os.RTLD_NOW)
    lib.gl_create_context.restype = C.c_void_p
    lib.gl_create_context.argtypes = [C.c_int32, C.c_int32]
    lib.gl_destroy_context.argtypes = [C.c_void_p]
    lib.glGetError.restype = C.c_uint
    have_gpu = os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK)
    ctx = lib.gl_create_context(16, 16)
    if ctx:
        assert have_gpu
        lib.gl_destroy_context(C.c_void_p(ctx))
    else:
        assert lib.glGetError() == 0   # no current context: a silent no-op
