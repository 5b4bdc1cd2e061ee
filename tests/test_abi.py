"""The C ABI library loads and exports every symbol include/ffx.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import subprocess

import ffxpy


def test_library_exports_every_declared_symbol():
    lib = ffxpy.load()
    names = ffxpy.header_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version_and_error_string():
    lib = ffxpy.load()
    assert lib.ffx_abi_version() == 1
    assert isinstance(lib.ffx_last_error(), bytes)


def test_binding_signatures_cover_header():
    # every declared function has a ctypes signature in the binding
    assert set(ffxpy.header_symbols()) <= set(ffxpy._SIGS)


def test_exported_symbols_are_c_abi():
    out = subprocess.run(["nm", "-D", "--defined-only", ffxpy.LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    for n in ffxpy.header_symbols():
        assert n in exported, n  # unmangled: extern "C"


def test_null_arguments_fail_cleanly_without_gpu():
    lib = ffxpy.load()
    # argument validation happens before any HIP call
    assert lib.ffx_search_slots_async(None, None, 0, 5, 100, None, None) == -22
    assert lib.ffx_sieve_slab_async(None, 3, 100, None, None) == -22
    assert b"segLo" in lib.ffx_last_error()
    assert lib.ffx_compact_async(None, None, 0, 5, None, None, None) == -22
    assert lib.ffx_sieve_plan_destroy(None) == 0
    assert lib.ffx_search_plan_destroy(None) == 0


def test_cli_fails_loudly_on_bad_flags():
    p = subprocess.run([ffxpy.FF_SIEVE, "--bogus", "5", "100"], capture_output=True, text=True)
    assert p.returncode == 1 and "unknown option" in p.stderr
    p = subprocess.run([ffxpy.FF_SIEVE, "--no-gpu", "5", "100"], capture_output=True, text=True)
    assert p.returncode == 1


REF_HEADERS = "/root/reference/source"


def test_reference_adapter_compiles_against_reference_headers():
    # the reference-side adapter type-checks against the reference's own
    # SievePoolOps / GpuRecord declarations and exports the three getters
    import pytest
    if not os.path.exists(os.path.join(REF_HEADERS, "sieve_slab_engine.h")):
        pytest.skip("reference headers absent")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = os.path.join(root, "ff-gpu_amd", "bindings", "reference_adapter.cpp")
    p = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I" + os.path.join(root, "include"), "-I" + REF_HEADERS, src], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    so = os.path.join(os.path.dirname(ffxpy.LIB_PATH), "libffx_refadapter.so")
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True).stdout
    for n in ("SievePoolGet_gfx950", "SieveSlabEngineGetLaunchFn_gfx950", "SearchKernelGetLaunchFn_gfx950"):
        assert n in out, n
