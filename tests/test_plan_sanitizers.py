"""The host-side dedispersion planner (``dsx_dedisp_plan``) under
AddressSanitizer and UndefinedBehaviorSanitizer: ``tests/native/plan_sanitize.cpp``
drives every plan mode over the node's 16-subband 256-DM shape, odd and tiny
shapes and the error paths, checking that every term base lies inside the
LDS the plan reserves.  Built with the host compiler from the planner's own
source (no device code), run on the CPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "dsa110-rt_amd", "csrc")
ROCM = "/opt/rocm"


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists(os.path.join(ROCM, "lib", "libamdhip64.so")),
                    reason="needs g++ and the ROCm runtime library")
def test_planner_clean_under_asan_ubsan(tmp_path):
    exe = tmp_path / "plan_sanitize"
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
           os.path.join(CSRC, "dedisp_plan.cpp"), os.path.join(CSRC, "capi.cpp"),
           os.path.join(ROOT, "tests", "native", "plan_sanitize.cpp"),
           f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib", "-o", str(exe)]
    out = subprocess.run(cmd, capture_output=True, text=True)
    assert out.returncode == 0, out.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    assert run.returncode == 0, run.stdout[-2000:] + run.stderr[-4000:]
    assert "bad 0" in run.stdout
