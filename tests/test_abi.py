"""The C-ABI libraries load and export every entry point the public headers
declare; the plugin library registers the aligner types.  CPU only (no
compute calls)."""
import ctypes
import os
import re

from kd_slam2_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kdh_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    names = _declared("kdh.h") + _declared("kdh_host.h") + _declared("kdh_io.h")
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_plugin_registers_types():
    A = _lib.adapter()
    buf = ctypes.create_string_buffer(512)
    assert A.kdha_registered_types(buf, 512) == 0
    types = buf.value.decode().split(",")
    for t in ("ICPHIP3D", "CTICPHIP3D", "ICPTerminationCriteriaScore3D"):
        assert t in types


def test_params_defaults_match_reference():
    """ICPParams_ / CTICPParams_ defaults (tree_aligner_base_.h:12-21, cticp_.h:17)."""
    import math

    from kd_slam2_amd import default_params

    p = default_params()
    assert abs(p.kernel_threshold - 0.1) < 1e-7 and p.damping == 1.0 and p.max_error == 1.0
    assert abs(p.min_normal_cos - math.cos(math.radians(25))) < 1e-7
    assert p.min_good_measurements == 10 and p.max_iterations == 20 and p.velocity_damping == 100.0


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libkdhip.so"))
    try:
        _lib.lib()
    except RuntimeError as e:
        assert "no CPU fallback" in str(e)
    else:
        raise AssertionError("loading a missing library must raise")
