"""The served loop's host logic (asv_serve.cpp, asv_stream.cpp) under
ThreadSanitizer on the CPU: tests/tsan/serve_tsan.cpp links the real sources
against host stand-ins for the device calls and runs a mixed workload (four
workers, evictions and adoption of idle contexts, a malformed request, two
cancellations, a failing allocation) three times; every stream that is not
meant to fail must give the checksum it gives alone, and TSan must report
nothing.  Built with ROCm's clang, whose TSan runtime intercepts
pthread_cond_clockwait (GCC 11's libtsan does not, and reports false races
around std::condition_variable::wait_for)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.mark.skipif(not os.path.exists(CLANG), reason="ROCm clang not installed")
def test_served_loop_has_no_data_races(tmp_path):
    csrc = os.path.join(ROOT, "anisotropic_supervoxel_amd", "csrc")
    exe = str(tmp_path / "serve_tsan")
    cmd = [CLANG, "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           os.path.join(csrc, "asv_serve.cpp"), os.path.join(csrc, "asv_stream.cpp"), os.path.join(csrc, "asv_synth.cpp"),
           os.path.join(ROOT, "tests", "tsan", "serve_tsan.cpp"), "-o", exe, "-lpthread"]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 second_deadlock_stack=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    out = r.stdout + r.stderr
    assert "ThreadSanitizer" not in out, out[-6000:]
    assert r.returncode == 0 and "serve harness ok" in r.stdout, out[-4000:]
    shutil.rmtree(tmp_path, ignore_errors=True)
