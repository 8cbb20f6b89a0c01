"""CPU-only checks of the boundary and the host logic (no GPU calls)."""
from __future__ import annotations

import ctypes
import glob
import os
import re

import numpy as np
import pytest

from lie_cliffalg_analog_svd_amd import _lib, profiles, rng
from lie_cliffalg_analog_svd_amd.handle import pack_offsets

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    syms = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(lsvd_[a-z0-9_]+)\s*\(", text):
            syms.add(m.group(1))
    return syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = _declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in sorted(syms) if not hasattr(lib, s)]
    assert not missing, f"declared but not exported: {missing}"


def test_python_binding_covers_every_symbol():
    assert _declared_symbols() == set(_lib.SIGNATURES), (
        sorted(_declared_symbols() ^ set(_lib.SIGNATURES)))


def test_status_strings():
    lib = _lib.lib()
    assert lib.lsvd_status_string(0) == b"ok"
    assert lib.lsvd_status_string(1) == b"shape error"
    assert lib.lsvd_version().decode().count(".") == 2


def test_pack_offsets_layout():
    L = pack_offsets([5, 3, 4], [2, 6, 4])
    assert list(L["a_off"]) == [0, 10, 28]
    assert list(L["u_off"]) == [0, 10, 19]      # m*k: 10, 9, 16
    assert list(L["s_off"]) == [0, 2, 5]
    assert list(L["vt_off"]) == [0, 4, 22]      # k*n: 4, 18, 16
    assert (L["a_len"], L["u_len"], L["s_len"], L["vt_len"]) == (44, 35, 9, 38)


def test_rng_is_deterministic_and_standard():
    a = rng.normal(123, 200000)
    assert np.array_equal(a, rng.normal(123, 200000))
    assert abs(a.mean()) < 0.01 and abs(a.std() - 1.0) < 0.01
    u = rng.uniform(5, 100000)
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01
    q = rng.haar(9, 17)
    assert np.linalg.norm(q.T @ q - np.eye(17)) < 1e-13


def test_deterministic_profiles_match_reference_formulas():
    """src/lib.rs:270-317 entries spot-checked from the Rust text."""
    j = profiles.generate(8, profiles.Profile.JordanDefective, 0).a
    assert j[0, 0] == 1.0 and j[3, 3] == 1.0 - 0.5 * 3 / 8 and j[2, 3] == 20.0 and j[4, 6] == -3.0 and j[1, 3] == 0
    s = profiles.generate(10, profiles.Profile.SparseStructured, 0).a
    assert s[8, 8] == 4.0 + 1 * 0.25 and s[0, 1] == -1.0 and s[1, 0] == 0.75 and s[0, 8] == 0.15 and s[9, 1] == -0.10
    nd = profiles.generate(4, profiles.Profile.NearlyDiagonal, 0)
    assert nd.a[1, 1] == 1.0 + 3 / 4 and nd.a[0, 1] == 1e-6 * np.sin(17.0)
    k = profiles.generate(6, profiles.Profile.KronStructured, 0).a
    assert k[4, 4] == 1.0 and k[5, 5] == 1.01 and k[0, 0] == 1.6 * (1.6 + 0.09 * 1)
    assert profiles.generate(1, profiles.Profile.KronStructured, 0).a[0, 0] == 2.0


def test_shape_errors_are_assertion_errors():
    """The reference panics (assert!) on shape misuse; the Python mirror
    raises ShapeError, an AssertionError, before touching the GPU."""
    from lie_cliffalg_analog_svd_amd import LieSvdMicro, LieSvdSmall
    with pytest.raises(AssertionError):
        LieSvdSmall.solve(np.zeros((3, 4)))
    with pytest.raises(AssertionError):
        LieSvdMicro.solve(np.zeros((5, 5)))
    with pytest.raises(AssertionError):
        LieSvdSmall.solve_rectangular(np.zeros((0, 3)))


def test_product_does_not_import_oracle():
    """The product package never routes through the CPU oracle."""
    pkg = os.path.join(ROOT, "lie_cliffalg_analog_svd_amd")
    for f in glob.glob(os.path.join(pkg, "**", "*.py"), recursive=True):
        text = open(f).read()
        assert "import oracle" not in text and "from oracle" not in text, f
    for f in glob.glob(os.path.join(pkg, "csrc", "*")):
        assert "lsvd_oracle" not in open(f).read(), f


def test_missing_library_fails_loudly():
    """With the shared library absent the binding raises ImportError naming
    the build command; there is no CPU path to fall back to."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from lie_cliffalg_analog_svd_amd import _lib\n"
            "_lib.LIB_PATH = '/nonexistent/liblsvd_amd.so'\n"
            "try:\n    _lib.lib()\nexcept ImportError as e:\n    print('IMPORT_ERROR', 'no CPU fallback' in str(e))\n"
            % ROOT)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert "IMPORT_ERROR True" in out.stdout, out.stdout + out.stderr


def test_solvers_raise_without_a_gpu():
    """Where no GPU is visible the solvers raise (handle creation fails)
    instead of computing anything on the host."""
    import ctypes
    h = ctypes.c_void_p()
    if _lib.lib().lsvd_create(0, ctypes.byref(h)) == _lib.LSVD_OK:
        _lib.lib().lsvd_destroy(h)
        pytest.skip("a GPU is visible")
    from lie_cliffalg_analog_svd_amd import LieSvdSmall, LieSvdTensor, Handle
    with pytest.raises(_lib.LsvdError):
        Handle()
    with pytest.raises(_lib.LsvdError):
        LieSvdSmall.solve(np.eye(3))
    with pytest.raises(_lib.LsvdError):
        LieSvdTensor.hosvd3(np.ones((2, 2, 2)))


def _chunks(layers, npl, max_nodes):
    lib = _lib.lib()
    nch = lib.lsvd_sweep_graph_chunks(layers, npl, max_nodes, None, None, 0)
    first = (ctypes.c_int32 * max(nch, 1))()
    count = (ctypes.c_int32 * max(nch, 1))()
    assert lib.lsvd_sweep_graph_chunks(layers, npl, max_nodes, first, count, nch) == nch
    return [(first[c], count[c]) for c in range(nch)]


@pytest.mark.parametrize("n", [2, 3, 16, 129, 255, 256, 257, 511, 512, 1023, 1024, 2047, 2048, 4096])
def test_sweep_graph_chunks_bound_and_cover(n):
    """Phase-JADE / joint-SVD sweeps are captured as graphs of at most
    lsvd_max_graph_nodes() kernel nodes (two per layer, one stop test at the
    end): the chunks cover every layer once, in order, and none exceeds the
    bound (a 2047-node sweep graph, n = 1024, ended in a host-side fault
    under the kernel tracer)."""
    layers = n - 1 if n % 2 == 0 else n
    cap = _lib.lib().lsvd_max_graph_nodes()
    assert 2 < cap <= 512
    ch = _chunks(layers, 2, cap)
    nxt = 0
    for c, (f, k) in enumerate(ch):
        assert f == nxt and k >= 1
        nodes = 2 * k + (1 if c == len(ch) - 1 else 0)
        assert nodes <= cap
        nxt = f + k
    assert nxt == layers
    if n <= 256:
        assert len(ch) == 1  # n <= 256 keeps the whole sweep in one graph
    assert _chunks(0, 2, cap) == []
    assert _chunks(5, 3, 7) == [(0, 2), (2, 2), (4, 1)]
