"""AddressSanitizer and UndefinedBehaviorSanitizer builds of host code (CPU only, gcc):
the oracle's C restatement driven by tests/sanitize/oracle_harness.c over ragged shapes,
both layouts and 1..4 reference threads. Any sanitizer report fails the test."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_oracle_under_asan_ubsan(tmp_path):
    exe = tmp_path / "oracle_san"
    flags = ["-O1", "-g", "-std=c11", "-fopenmp", "-ffp-contract=off", "-mfma", "-fno-omit-frame-pointer",
             "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
    r = subprocess.run(["gcc", *flags, "-I", ORACLE, os.path.join(ORACLE, "oracle.c"),
                        os.path.join(ROOT, "tests", "sanitize", "oracle_harness.c"), "-o", str(exe), "-lm"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr[-3000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-3000:]
    assert "sgemm tail" in r.stdout


HOST_SOURCES = ["api", "grouped", "serving", "autotune"]  # the library's host-only sources
KERNEL_OBJECTS = ["gemm_kernels", "chain_kernel", "f64_kernels", "im2col", "noalloc"]


@pytest.mark.skipif(shutil.which("gcc") is None or not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="gcc or hipcc not found")
def test_host_abi_under_asan_ubsan(tmp_path):
    """The library's host sources built with the sanitizers (host pass only; the kernel
    objects are the normal build's), linked into tests/sanitize/host_harness.c: the
    planner over random shapes and options, autotune table files good and malformed,
    and the other calls that need no GPU."""
    import concurrent.futures as cf
    hipcc = "/opt/rocm/bin/hipcc"
    build = os.path.join(ROOT, "flashgemm_amd", "_build")
    if not all(os.path.exists(os.path.join(build, n + ".o")) for n in KERNEL_OBJECTS):
        pytest.skip("the library's kernel objects are not built (python -m flashgemm_amd.build)")
    san = ["-Xarch_host", "-fsanitize=address,undefined", "-Xarch_host", "-fno-sanitize-recover=all"]

    def compile_one(name):
        return subprocess.run([hipcc, "-O1", "-g", "-std=c++17", "-fPIC", *san, "-Xclang", "-target-feature",
                               "-Xclang", "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"), "-c",
                               os.path.join(ROOT, "flashgemm_amd", "csrc", name + ".cpp"), "-o",
                               str(tmp_path / (name + ".o"))], capture_output=True, text=True)

    with cf.ThreadPoolExecutor(4) as ex:
        for r in ex.map(compile_one, HOST_SOURCES):
            assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run(["gcc", "-O1", "-g", "-c", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "sanitize", "host_harness.c"), "-o", str(tmp_path / "h.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    exe = tmp_path / "host_san"
    objs = [str(tmp_path / "h.o")] + [str(tmp_path / (n + ".o")) for n in HOST_SOURCES] + \
        [os.path.join(build, n + ".o") for n in KERNEL_OBJECTS]
    r = subprocess.run([hipcc, *san[:2], *objs, "-o", str(exe), "-lpthread"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0, (r.stdout[-1000:], r.stderr[-3000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-3000:]
    assert "0 failed checks" in r.stdout
