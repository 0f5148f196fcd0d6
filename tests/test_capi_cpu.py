"""CPU checks of the C ABI library: it loads, and exports every symbol
include/femh.h declares (no device work is launched here)."""
import ctypes
import os
import re

import numpy as np
import pytest

import fem_amd
from fem_amd import femh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "femh.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(femh_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = fem_amd.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(femh.EXPORTS)


def test_abi_version_and_default_opts():
    L = fem_amd.lib()
    assert L.femh_abi_version() == 3
    o = femh.make_opts((8, 8, 8))
    assert o.nu == 0.3 and o.tol == 1e-10 and o.precond == 1 and o.nu1 == 2
    # the library's own defaults for the fields added in version 2
    d = femh.Opts()
    L.femh_default_opts(ctypes.byref(d))
    assert (d.structure, d.autotune, d.streams) == (0, 0, 1)
    assert d.batch_threads == 0  # version 3: femh_homogenize_batch keeps its 8 host threads
    o = femh.make_opts((8, 8, 8), structure="cubic", autotune=True, streams=3)
    assert (o.structure, o.autotune, o.streams) == (1, 1, 3)
    with pytest.raises(KeyError):
        femh.make_opts((8, 8, 8), structure="hexagonal")


def test_opts_struct_layout_matches_header():
    # the Python mirror must have the same field order as femh_opts
    src = open(os.path.join(ROOT, "include", "femh.h")).read()
    body = src[src.index("typedef struct femh_opts {"):src.index("} femh_opts;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for line in body.splitlines()[1:]:
        line = line.strip().rstrip(";")
        if not line:
            continue
        typ, rest = line.split(None, 1)
        fields += [f.strip() for f in rest.split(",")]
    assert fields == [f[0] for f in femh.Opts._fields_]


def test_stats_struct_layout_matches_header():
    src = open(os.path.join(ROOT, "include", "femh.h")).read()
    start = src.index("typedef struct")
    start = src.index("typedef struct", src.index("} femh_opts;"))
    body = src[start:src.index("} femh_stats;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for line in body.splitlines()[1:]:
        line = line.strip().rstrip(";")
        if not line:
            continue
        typ, rest = line.split(None, 1)
        fields += [re.sub(r"\[.*\]", "", f).strip() for f in rest.split(",")]
    assert fields == [f[0] for f in femh.Stats._fields_]


def test_host_element_table_is_symmetric_psd():
    K = femh.element_stiffness(1.0, 0.3)
    assert np.abs(K - K.T).max() < 1e-15
    assert np.linalg.eigvalsh(K).min() > -1e-13
    assert femh.lib().femh_element_stiffness(1.0, 0.7, K.ctypes.data) == femh.BAD_ARG


def test_host_facade_library_exports_reference_classes():
    """The C++ facade is built and exports the reference's class names."""
    import subprocess
    lib = os.path.join(ROOT, "fem_amd", "libfemh_host.so")
    assert os.path.exists(lib)
    syms = subprocess.run(["nm", "-DC", "--defined-only", lib], capture_output=True, text=True).stdout
    for name in ("femh::FEM3DFun::setParam", "femh::FEM3DFun::df",
                 "femh::NumericalCoarseningUsingMultigrid::computeCoarsenedElasticityTensorAndParameters3D",
                 "femh::MaterialOptimizer::run3D", "femh::gradientDescent", "femh::lineSearch",
                 "femh::check_df", "femh::PiecewiseConstantCubic3D"):
        assert name in syms, name


def test_noalloc_entry_point_rejects_bad_arguments_before_device_work():
    """Argument checks of femh_homogenize_device_noalloc run on the host
    before any device call: an oversized grid or a solver other than FP64
    Jacobi-PCG returns BAD_ARG with a message (null pointers stand in for the
    device buffers, which are never touched)."""
    L = fem_amd.lib()
    buf = ctypes.create_string_buffer(256)
    fake = ctypes.c_void_p(16)
    o = femh.make_opts((64, 64, 64), precond="jacobi")
    st = L.femh_homogenize_device_noalloc(ctypes.byref(o), fake, fake, None, None, buf, 256)
    assert st == femh.BAD_ARG and b"32768" in buf.value
    o = femh.make_opts((8, 8, 8), precond="mg")
    st = L.femh_homogenize_device_noalloc(ctypes.byref(o), fake, fake, None, None, buf, 256)
    assert st == femh.BAD_ARG and b"Jacobi" in buf.value
    st = L.femh_homogenize_device_noalloc(None, fake, fake, None, None, buf, 256)
    assert st == femh.BAD_ARG


def test_device_loop_entry_points_reject_bad_arguments_on_the_host():
    """femh_gradient_device and femh_descent_trial check their arguments on
    the host: without a context (or with a null pointer) they return BAD_ARG
    and launch nothing."""
    L = fem_amd.lib()
    fake = ctypes.c_void_p(16)
    h = np.ones(4)
    G = np.zeros(36)
    assert L.femh_gradient_device(None, G.ctypes.data, None, fake, None, None) == femh.BAD_ARG
    assert L.femh_descent_trial(None, 4, fake, fake, h.ctypes.data, 1e-3, 1.0, fake, None, None) == femh.BAD_ARG
