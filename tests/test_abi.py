"""The C ABI library loads and exports every entry point include/*.h declares;
without a GPU the calls fail loudly (no CPU fallback)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header, prefix):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", text)))


def test_rbhip_exports_every_declared_symbol():
    from openmm_rigidbody_plugin_amd import rbhip
    lib = ctypes.CDLL(rbhip.LIB_PATH)
    names = declared(os.path.join(ROOT, "include", "rbhip.h"), "rbhip_")
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares a signature for each of them
    bound = set(rbhip.exported_symbols())
    assert set(names) <= bound, set(names) - bound


def test_plugin_exports_registration_symbols():
    from openmm_rigidbody_plugin_amd import build
    if not os.path.exists(build.PLUGIN):
        pytest.skip("C++ plugin layer not built")
    header = os.path.join(ROOT, "include", "rigidbody_hip_plugin.h")
    lib = ctypes.CDLL(build.PLUGIN)
    names = declared(header, "") if os.path.exists(header) else []
    names = [n for n in names if n.startswith(("register", "rbplugin_"))]
    assert names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_headers_are_plain_c():
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        assert 'extern "C"' in text
        assert "torch" not in text and "std::" not in text


def test_no_gpu_means_loud_failure():
    from openmm_rigidbody_plugin_amd import rbhip
    if rbhip.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(rbhip.RbhipError):
        rbhip.Context([1.0, 1.0], [1, 1], 0.002)
    assert rbhip.lib().rbhip_version() == 1
