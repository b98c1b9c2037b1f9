"""The C-ABI library loads and exports every entry point declared in include/bh.h (no GPU)."""
import ctypes
import os
import re

from masterthesislucas_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "bh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bh_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "bh_create" in syms and "bh_forces" in syms and len(syms) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_config_default_without_gpu():
    lib = _lib.load()
    cfg = _lib.BHConfig()
    assert lib.bh_config_default(ctypes.byref(cfg)) == 0
    assert cfg.max_children == 16 and cfg.theta == 0.5 and cfg.code_bits == 32
    assert cfg.tree_mode == _lib.TREE_QUADTREE and cfg.kernel == _lib.KERNEL_TSNE


def test_create_rejects_bad_config_without_touching_gpu():
    lib = _lib.load()
    cfg = _lib.BHConfig()
    lib.bh_config_default(ctypes.byref(cfg))
    h = ctypes.c_void_p()
    cfg.n_max = 0
    assert lib.bh_create(ctypes.byref(h), ctypes.byref(cfg)) == 1
    cfg.n_max = 10
    cfg.max_children = 0
    assert lib.bh_create(ctypes.byref(h), ctypes.byref(cfg)) == 1
    cfg.max_children = 16
    cfg.code_bits = 40
    assert lib.bh_create(ctypes.byref(h), ctypes.byref(cfg)) == 1
