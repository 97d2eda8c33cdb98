"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

import trinity_amd as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(trn_[a-z0-9_]+)\s*\(", src))
    return names


def test_header_declares_the_bound_symbols():
    assert declared_symbols() == set(T.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(T.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_exports_are_plain_c():
    out = subprocess.run(["nm", "-D", "--defined-only", T.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for n in declared_symbols():
        assert n in exported, n   # unmangled extern "C" names


def test_version_and_error_strings_without_gpu():
    assert "gfx950" in T.version()
    assert isinstance(T.lib().trn_last_error(), bytes)


def test_null_arguments_are_rejected_not_crashed():
    L = T.lib()
    assert L.trn_ctx_create(0, None) == T.TRN_ERR_ARG
    assert L.trn_mesh_info(None, None, None, None, None) == T.TRN_ERR_ARG
    assert b"null" in L.trn_last_error()


def test_header_compiles_as_c():
    cc = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-x", "c",
                         os.path.join(ROOT, "include", "trinity_amd.h")], capture_output=True,
                        text=True)
    assert cc.returncode == 0, cc.stderr


@pytest.mark.skipif(not os.path.isdir("/root/reference/include"),
                    reason="reference headers absent")
def test_dropin_sources_compile_against_reference_headers():
    dropin = os.path.join(ROOT, "trinity_amd", "dropin")
    for src in ("metric.cpp", "smoothing.cpp", "trn_dropin.cpp"):
        r = subprocess.run(["g++", "-std=gnu++14", "-fopenmp", "-fsyntax-only", "-w",
                            "-I/root/reference/include", "-I" + os.path.join(ROOT, "include"),
                            "-I" + dropin, "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                            os.path.join(dropin, src)], capture_output=True, text=True)
        assert r.returncode == 0, (src, r.stderr[-2000:])


def test_product_package_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, "trinity_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "liboracle" not in text, f
