"""The library's host code under sanitizers (CPU only; no kernel runs).

tests/host_check.cpp drives the host-side paths -- scene building on many
threads at once, cell typing, tile files (and a corrupted one), the pool
build's tree construction on up to 16 threads with its worklists and coarse
plans, and the serving engine's range allocator against a bitmap model -- in a
build of the library's host code with AddressSanitizer + UndefinedBehavior-
Sanitizer, and in one with ThreadSanitizer.  Every sanitizer report fails the
test.  The scenes are the benchmark's: the adaptive 256^3 problem, the
256x256x512 vortex tree and 40 problems of the batched pool (its largest
among them).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import cirrus_amd as ca
from cirrus_amd.workloads import ADAPTIVE256, PLAIN64, VORTEX256, batch_scene_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "cirrus_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CLANG = "/opt/rocm/llvm/bin/clang++"
POOL = list(range(0, 256, 7)) + [242, 218, 96]

pytestmark = pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)), reason="no ROCm toolchain")


def scene_file(path):
    kws = [ADAPTIVE256, PLAIN64, VORTEX256] + [batch_scene_params(q)[0] for q in POOL]
    with open(path, "wb") as f:
        for kw in kws:
            f.write(bytes(ca.scene_params(**kw)))
    return len(kws)


def build(tmp, sanitizer):
    """The host code of the library plus the driver, instrumented; returns the binary."""
    san = [f"-fsanitize={s}" for s in sanitizer.split(",")]
    flags = ["-O1", "-g", "-fno-omit-frame-pointer", "-std=c++17", "-ffp-contract=off", "-pthread",
             "-I", os.path.join(ROOT, "include"), "-I", CSRC]
    xhost = [a for s in san for a in ("-Xarch_host", s)]  # the sanitizers instrument host code only
    objs = []
    for src, lang in [("tree.cpp", "c++"), ("tilefile.cpp", "c++")]:
        o = os.path.join(tmp, src + ".o")
        subprocess.run([HIPCC, "-x", lang] + flags + xhost + ["-c", os.path.join(CSRC, src), "-o", o], check=True)
        objs.append(o)
    o = os.path.join(tmp, "mgpcg.o")
    # device code is built as well (the object must carry its code object) but never runs
    subprocess.run([HIPCC, "--offload-arch=gfx950"] + flags + xhost + ["-c", os.path.join(CSRC, "mgpcg.hip"), "-o", o],
                   check=True)
    objs.append(o)
    o = os.path.join(tmp, "host_check.o")
    subprocess.run([HIPCC, "-x", "c++"] + flags + xhost + ["-c", os.path.join(ROOT, "tests", "host_check.cpp"), "-o", o],
                   check=True)
    objs.append(o)
    exe = os.path.join(tmp, "host_check")
    subprocess.run([CLANG] + san + ["-pthread", "-o", exe] + objs +
                   ["-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"], check=True)
    return exe


@pytest.mark.parametrize("sanitizer", ["address,undefined", "thread"])
def test_host_code_is_clean_under_sanitizers(tmp_path, sanitizer):
    tmp = str(tmp_path)
    exe = build(tmp, sanitizer)
    ns = scene_file(os.path.join(tmp, "scenes.bin"))
    tiles = os.path.join(tmp, "tiles")
    os.makedirs(tiles)
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:exitcode=24",
               TSAN_OPTIONS="halt_on_error=1:exitcode=25:report_signal_unsafe=0",
               HIP_VISIBLE_DEVICES="", ROCR_VISIBLE_DEVICES="")
    r = subprocess.run([exe, os.path.join(tmp, "scenes.bin"), "8", tiles], capture_output=True, text=True, env=env,
                       timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-6000:]
    assert f"host_check: {ns} scenes" in r.stdout and "ok" in r.stdout, out[-3000:]
    for word in ("ERROR: AddressSanitizer", "runtime error:", "WARNING: ThreadSanitizer"):
        assert word not in out, out[-6000:]
    shutil.rmtree(tiles, ignore_errors=True)
