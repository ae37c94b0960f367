"""The GL call scripts of tests/api_scripts.py on libmtgl.so, against the
fixtures recorded from the reference (tests/golden/golden.json): error
codes, state queries, display-list and object bookkeeping, and glReadPixels
(src/gl_api.c:1180-1229) of windows inside, across and outside the
framebuffer, read back through the device's packing kernel."""
import json
import os

import pytest

from conftest import ROOT
import api_scripts

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))


@pytest.mark.parametrize("name", sorted(GOLDEN["api"]))
def test_api_script_device(mtgl, oracle, name):
    got = json.loads(json.dumps(api_scripts.run_script(mtgl, name)))
    assert got == GOLDEN["api"][name]
    assert got == json.loads(json.dumps(api_scripts.run_script(oracle, name)))
