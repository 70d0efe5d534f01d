"""The C-ABI library loads without a GPU and exports every function that
include/prt_hip.h declares; host-only entry points behave."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest


def test_header_declares_the_abi(prt):
    names = prt.declared_symbols()
    assert "prt_hip_create" in names and "prt_hip_render" in names and len(names) >= 20


def test_library_exports_every_declared_symbol(prt):
    lib = prt.load_library()
    missing = [n for n in prt.declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"
    assert lib.prt_hip_abi_version() == 4


def test_library_exports_only_the_c_abi(prt):
    """Helpers, kernels' host stubs and template instantiations stay internal:
    the dynamic symbol table holds the prt_hip_* entry points (and the HIP
    toolchain's compilation-unit ids) and nothing else."""
    if shutil.which("nm") is None:
        pytest.skip("nm not available")
    out = subprocess.run(["nm", "-D", "--defined-only", prt.LIB_PATH], check=True, capture_output=True,
                         text=True).stdout
    names = [ln.split()[-1] for ln in out.splitlines() if ln.strip()]
    extra = [n for n in names if not n.startswith("prt_hip_") and not n.startswith("__hip_cuid_")]
    assert not extra, f"unexpected exports: {extra}"
    declared = set(prt.declared_symbols())
    exported = {n for n in names if n.startswith("prt_hip_")}
    assert exported == declared, f"exported but not declared: {sorted(exported - declared)}"


def test_no_context_without_gpu_is_an_error_not_a_crash(prt):
    lib = prt.load_library()
    ctx = ctypes.c_void_p()
    rc = lib.prt_hip_create(-1, ctypes.byref(ctx))
    assert rc != 0 and not ctx.value
    assert lib.prt_hip_last_error(None) == b"null context"
    assert lib.prt_hip_destroy(None) == 0


def test_camera_from_view_matches_oracle(prt, orc, scenes):
    for name in scenes.CONFIGS:
        v = scenes.CONFIGS[name]["view"]
        for aspect in (1.0, 16 / 9):
            a = prt.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], aspect)
            b = orc.camera_from_view(v["pos"], v["angle_x"], v["angle_y"], v["fov_y"], aspect)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_oracle_library_is_separate(prt):
    # the product library must not depend on the oracle
    with open(prt.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"liboracle" not in blob and b"orc_render" not in blob
