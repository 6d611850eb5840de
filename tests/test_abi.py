"""C-ABI boundary checks that need no GPU: the library loads, exports every function
declared in include/flashgemm.h with C linkage, the header is valid C and C++, and
the reference-style C drivers compile and link against it."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "flashgemm.h")
LIB = os.path.join(ROOT, "flashgemm_amd", "libflashgemm.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*([A-Za-z_][A-Za-z_0-9]*)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_reference_symbols():
    names = declared_functions()
    for ref in ("flashgemm_set_thread_num", "flashgemm_multi_f32f32f32", "flashgemm_multi_bf16bf16f32",
                "flashgemm_multi_int8uint8int32", "LibShalom_set_thread_nums", "LibShalom_sgemm_mp"):
        assert ref in names


@pytest.fixture(scope="module")
def built_lib():
    if not os.path.exists(LIB):
        pytest.skip("libflashgemm.so not built (run __graft_entry__.build())")
    return ctypes.CDLL(LIB)


def test_library_exports_every_declared_symbol(built_lib):
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    assert "flashgemm_thread_num" in exported  # the extern int


def test_thread_num_global_is_shared(built_lib):
    built_lib.flashgemm_set_thread_num(7)
    assert ctypes.c_int.in_dll(built_lib, "flashgemm_thread_num").value == 7
    built_lib.flashgemm_set_thread_num(1)


def test_status_strings(built_lib):
    built_lib.fg_status_string.restype = ctypes.c_char_p
    assert built_lib.fg_status_string(0) == b"ok"
    assert built_lib.fg_status_string(-2).startswith(b"reference layout")


@pytest.mark.parametrize("lang,compiler", [("c", "gcc"), ("c++", "g++")])
def test_header_compiles(tmp_path, lang, compiler):
    src = tmp_path / ("t.c" if lang == "c" else "t.cpp")
    src.write_text('#include "flashgemm.h"\nint main(void){fg_chain_opts o; fg_chain_opts_default(&o);'
                   ' return o.layout;}\n')
    subprocess.run([compiler, "-std=c11" if lang == "c" else "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                    "-I", os.path.join(ROOT, "include"), str(src)], check=True)


def test_drivers_link(tmp_path, built_lib):
    drivers = os.path.join(ROOT, "drivers")
    srcs = sorted(f for f in os.listdir(drivers) if f.endswith(".c"))
    assert srcs
    for f in srcs:
        exe = tmp_path / f.replace(".c", "")
        subprocess.run(["gcc", "-O2", "-std=gnu11", "-Wall", "-I", os.path.join(ROOT, "include"),
                        os.path.join(drivers, f), "-o", str(exe), "-L", os.path.dirname(LIB), "-lflashgemm",
                        f"-Wl,-rpath,{os.path.dirname(LIB)}", "-lm"], check=True)


def test_packed_weight_bytes(built_lib):
    """Packed weights are whole 1 KiB MFMA fragments (2 KiB for the split-bf16 form); a
    shape that fills every fragment packs to exactly its own bytes, a ragged one is
    zero padded up, and invalid arguments size to 0 (no device needed)."""
    import flashgemm_amd as fg
    for dtype, mode, esz in ((fg.DTYPE_INT8, fg.F32_MFMA, 1), (fg.DTYPE_BF16, fg.F32_MFMA, 2),
                             (fg.DTYPE_F32, fg.F32_MFMA, 4), (fg.DTYPE_F32, fg.F32_BF16X3, 4)):
        assert fg.packed_weight_bytes(dtype, 512, 512, mode) == 512 * 512 * esz
        assert fg.packed_weight_bytes(dtype, 2048, 1024, mode) == 2048 * 1024 * esz
        for M, K in ((1, 1), (33, 17), (130, 1000)):
            got = fg.packed_weight_bytes(dtype, M, K, mode)
            assert got >= M * K * esz and got % 1024 == 0
        assert fg.packed_weight_bytes(dtype, 0, 64, mode) == 0
        assert fg.packed_weight_bytes(dtype, 64, -1, mode) == 0
    assert fg.packed_weight_bytes(7, 64, 64) == 0
