"""The C-ABI library without a GPU: it loads, exports every entry point that
include/lb3d_gpu.h declares, and rejects invalid parameters before touching
the device (no compute calls here)."""
import ctypes
import os
import re

import pytest

from lb3d_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "lb3d_gpu.h")).read()
    return sorted(set(re.findall(r"\b(lb3d_gpu_\w+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(gpu.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = gpu.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name
    assert L.lb3d_gpu_abi_version() == 1


@pytest.mark.parametrize("kw,msg", [
    (dict(ncomp=3, boundary=gpu.LB3D_BC_LEES_EDWARDS), "surfactant"),
    (dict(ncomp=4), "ncomp"),
    (dict(bdist=7), "bdist"),
    (dict(precision=7), "precision"),
    (dict(tau=(0.0, 1.0, 1.0)), "tau"),
    (dict(inv_fluid=11), "inv_fluid"),
    (dict(inv_fluid=12, inv_type=2), "inv_type"),
    (dict(inv_fluid=12, boundary=gpu.LB3D_BC_LEES_EDWARDS), "Zou-He"),
])
def test_invalid_parameters_rejected(kw, msg):
    with pytest.raises(gpu.LB3DError) as e:
        gpu.Lattice(8, 8, 8, **kw)
    assert e.value.code == 1 and msg in str(e.value)


@pytest.mark.parametrize("kw,msg", [
    (dict(inv_fluid=12), "invasion"),
    (dict(bdist=7), "bdist"),
    (dict(boundary=gpu.LB3D_BC_LEES_EDWARDS), "periodic"),
    (dict(precision=gpu.LB3D_FP32_DEV), "share"),
])
def test_batch_invalid_parameters_rejected(kw, msg):
    """The batch takes periodic lattices of one form (ncomp, precision) with
    bdist 0-6 and no invasion; the checks run before any device call."""
    with pytest.raises(gpu.LB3DError) as e:
        gpu.Batch([(8, 8, 8), (6, 6, 6)], per_lattice=[{}, kw], ncomp=2)
    assert e.value.code == 1 and msg in str(e.value)


def test_too_large_lattice_rejected():
    with pytest.raises(gpu.LB3DError):
        gpu.Lattice(2048, 2048, 512, ncomp=3)


def test_params_struct_layout():
    # the Fortran bind(C) type and the C struct must agree on this size
    assert ctypes.sizeof(gpu.Params) == 4 * 8 + 8 * (3 + 3 + 3 + 3 + 3 + 3) + 4 * 6 + 8 * 3 + 8 * 6 + 4 * 2 + 8 * 3


def test_sanity_struct_layout():
    # lb3d_sanity: 24 doubles, 27 + 1 int32, 4 int64 (the Fortran bind(C) type mirrors it)
    assert ctypes.sizeof(gpu.Sanity) == 24 * 8 + 28 * 4 + 4 * 8


def test_batch_entry_points_reject_a_null_batch():
    """Plan and group calls on a NULL batch return LB3D_EINVAL (1) and say
    why, without a device call."""
    L = gpu.lib()
    assert L.lb3d_gpu_batch_set_groups(None, 1 << 20) == 1
    assert b"NULL" in L.lb3d_gpu_last_error()
    assert L.lb3d_gpu_batch_set_plan(None, 0, 64, 0) == 1
    assert L.lb3d_gpu_batch_autotune(None, 1, None, 0) == 1
