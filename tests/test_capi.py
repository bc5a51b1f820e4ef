"""The C-ABI library loads and exports every entry point include/elas_hip.h
declares; the parameter struct matches Elas::parameters (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    src = open(os.path.join(ROOT, "include", "elas_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(elas_\w+)\s*\(", src)))


def test_header_declares_the_reference_boundary():
    names = _declared_functions()
    for n in ("elas_params_default", "elas_hip_create", "elas_hip_process", "elas_hip_process_batch",
              "elas_hip_process_device", "elas_hip_debug_fetch", "elas_hip_destroy"):
        assert n in names


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.load_library()
    for name in _declared_functions():
        assert hasattr(lib, name), name
    assert set(pkg.SIGNATURES) == set(_declared_functions())


def test_parameter_struct_layout(pkg, oracle):
    # same order/types as Elas::parameters: 23 fields, bools are 1 byte
    assert [f for f, _ in pkg.ElasParams._fields_] == [f for f, _ in oracle.Params._fields_]
    assert ctypes.sizeof(pkg.ElasParams) == ctypes.sizeof(oracle.Params) == 80
    assert pkg.ElasParams.add_corners.offset == 32 and pkg.ElasParams.grid_size.offset == 36
    assert pkg.ElasParams.subsampling.offset == 79 - 0 or pkg.ElasParams.subsampling.offset == 79


def test_effective_defaults_match_kernel_macros(pkg, oracle):
    """KernelsCommon.h:8-27 values, post-processing off."""
    d = pkg.default_params().as_dict()
    exp = oracle.params()
    for k, _ in oracle.Params._fields_:
        assert d[k] == getattr(exp, k), k
    assert d["support_threshold"] == pytest.approx(0.85) and d["gamma"] == 3.0 and d["sradius"] == 2.0


def test_declared_defaults_match_reference_struct(pkg):
    """The dead initialisers of src/Elas.h:42-66."""
    d = pkg.declared_params().as_dict()
    assert d["support_threshold"] == pytest.approx(0.95) and d["gamma"] == 5.0 and d["sradius"] == 3.0
    assert d["speckle_size"] == 200 and d["ipol_gap_width"] == 5000 and d["filter_median"]
    assert d["add_corners"] and not d["subsampling"]


def test_create_without_gpu_fails_loudly(pkg):
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(pkg.ElasError) as e:
        pkg.Elas()
    assert e.value.code == 5
