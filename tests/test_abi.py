"""The C ABI library loads and exports every entry point include/*.h declares;
without a GPU the calls fail loudly (no CPU fallback)."""
import ctypes
import glob
import os
import re

import numpy as np
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


def test_estimate_counts_bodies_whatever_the_labels():
    # host-only: the body count and the byte estimate depend on which atoms
    # share a positive label, not on the labels' values (labels <= 0 are
    # free atoms); compact labels are counted directly, labels spread wider
    # than 4 n + 1024 go through the renumbering
    from openmm_rigidbody_plugin_amd import rbhip
    lib = rbhip.lib()

    def estimate(labels):
        bi = np.ascontiguousarray(labels, dtype=np.int32)
        n, nb = ctypes.c_size_t(), ctypes.c_int()
        assert lib.rbhip_estimate_bytes(bi.size, bi.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 0,
                                        ctypes.byref(n), ctypes.byref(nb)) == 0
        return n.value, nb.value

    rng = np.random.default_rng(3)
    sizes = rng.choice([1, 2, 3, 6, 20], size=300)
    labels = np.repeat(np.arange(1, 301), sizes)
    labels = np.concatenate([labels, np.zeros(150, int), -np.arange(1, 40)])
    labels = labels[rng.permutation(labels.size)]
    base = estimate(labels)
    assert base[1] == 300
    body = labels > 0
    for relabel in (lambda v: v + 1000, lambda v: v * 7_000_000 // 300 + 5, lambda v: 2_000_000_000 - v):
        other = labels.copy()
        other[body] = relabel(labels[body])
        assert estimate(other) == base
    # every atom free, and one body holding every atom
    assert estimate(np.zeros(10, int))[1] == 0
    assert estimate(np.full(10, 123456789))[1] == 1


def test_graph_cache_logic():
    """The keyed graph cache (csrc/rbhip_graph_cache.h) on the CPU with fake
    handles: replay only under the capture key, destroy only after the
    graph's stream was waited for (tests/cpp/TestGraphCache.cpp)."""
    import subprocess
    cpp = os.path.join(ROOT, "tests", "cpp")
    subprocess.run(["make", "-C", cpp, "_bin/TestGraphCache"], check=True, capture_output=True)
    r = subprocess.run([os.path.join(cpp, "_bin", "TestGraphCache")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("Done")


def test_shape_choice_logic():
    """The choice of shapes that bodies of one size class share
    (csrc/rbhip_shapes.h) on the CPU: sign variants with rounding noise
    share, noise above the tolerance or too many shapes do not
    (tests/cpp/TestShapes.cpp)."""
    import subprocess
    cpp = os.path.join(ROOT, "tests", "cpp")
    subprocess.run(["make", "-C", cpp, "_bin/TestShapes"], check=True, capture_output=True)
    r = subprocess.run([os.path.join(cpp, "_bin", "TestShapes")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("Done")
