"""Host-side sanitizer runs of the C ABI (no GPU).

The ABI translation unit (argument checks, error reporting, dispatch) is
rebuilt with AddressSanitizer + UndefinedBehaviorSanitizer, and separately
with ThreadSanitizer, on the host side only (device code is compiled as
usual), linked with the library's kernel objects, and driven by
tests/sanitize/abi_harness.cpp: every entry point refuses malformed calls
(NULL buffers, negative and INT64_MAX ray counts, overflowing image sizes,
bad modes, launch shapes, tolerances and checkpoint intervals) with
BH_EINVAL and its own thread-local message, with no sanitizer report.

The harness hands host addresses to entry points that must refuse them, so
it runs only where no GPU is visible.
"""

from __future__ import annotations

import os
import shutil
import subprocess

import pytest

from blackhorizon_amd import build as bh_build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "tests", "sanitize", "abi_harness.cpp")
CLANG = "/opt/rocm/llvm/bin/clang++"
ROCM_LIB = "/opt/rocm/lib"

SANITIZERS = {
    "asan_ubsan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
    "tsan": ["-fsanitize=thread"],
}

pytestmark = pytest.mark.skipif(
    os.path.exists("/dev/kfd") or not os.path.exists(CLANG)
    or shutil.which(bh_build.hipcc()) is None and not os.path.exists(bh_build.hipcc()),
    reason="host sanitizer harness runs only without a visible GPU, with hipcc and clang++")


def _kernel_objects() -> list[str]:
    objdir = os.path.join(bh_build.PKG, "_build")
    objs = [os.path.join(objdir, u + ".o") for u in ("bh_precise", "bh_fast")]
    if not all(os.path.exists(o) for o in objs):
        bh_build.build()
    return objs


def _run(cmd: list[str], **kw) -> subprocess.CompletedProcess:
    return subprocess.run(cmd, capture_output=True, text=True, timeout=600, **kw)


@pytest.fixture(scope="module", params=sorted(SANITIZERS))
def harness(request, tmp_path_factory):
    name = request.param
    flags = SANITIZERS[name]
    out = tmp_path_factory.mktemp(name)
    abi_obj = str(out / "bh_abi.o")
    # sanitizer flags apply to the host pass only: each one follows -Xarch_host
    host_flags = []
    for f in flags:
        host_flags += ["-Xarch_host", f]
    cmd = [bh_build.hipcc(), *bh_build.COMMON, *bh_build.UNITS["bh_abi"], "-O1", "-g",
           "-fno-omit-frame-pointer", *host_flags, "-c",
           os.path.join(bh_build.CSRC, "bh_abi.hip"), "-o", abi_obj]
    proc = _run(cmd)
    assert proc.returncode == 0, proc.stderr
    exe = str(out / "abi_harness")
    cmd = [CLANG, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", *flags,
           "-I", os.path.join(ROOT, "include"), HARNESS, abi_obj, *_kernel_objects(),
           "-L", ROCM_LIB, "-lamdhip64", f"-Wl,-rpath,{ROCM_LIB}", "-pthread", "-o", exe]
    proc = _run(cmd)
    assert proc.returncode == 0, proc.stderr
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:exitcode=23"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["TSAN_OPTIONS"] = "halt_on_error=1:exitcode=66"
    return name, exe, env


@pytest.mark.parametrize("mode", ["cases", "threads"])
def test_abi_refusals_under_sanitizer(harness, mode):
    name, exe, env = harness
    proc = _run([exe, mode], env=env)
    report = proc.stdout[-4000:] + proc.stderr[-4000:]
    assert proc.returncode == 0, f"{name} {mode}:\n{report}"
    assert "Sanitizer" not in proc.stderr, f"{name} {mode}:\n{report}"
    assert "0 failures" in proc.stdout
