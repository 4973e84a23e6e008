"""The library's host code (index lowering, planning, batch and server) built
with AddressSanitizer and UndefinedBehaviorSanitizer into a test executable
(muscle.jl_amd/csrc/sanitize_driver.cpp, `make -C muscle.jl_amd/csrc sanitize`).

Without a GPU it checks the lowering on 3,000 random contractions through
muscle_hip_describe (fused extents, the class of every label) and that invalid
descriptors end in an error status.  With a GPU it also runs random
contractions of every dtype, a batch in all run modes and a server with
contraction-tree jobs, each against a naive host loop over every label
assignment.  A sanitizer report aborts the executable (non-zero status).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "muscle.jl_amd", "lib", "sanitize_driver")


def _run(args, env_extra, timeout):
    if not os.path.exists(DRIVER):
        pytest.skip("sanitize_driver is not built (make -C muscle.jl_amd/csrc sanitize)")
    env = dict(os.environ, **env_extra)
    r = subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=timeout, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "ok: 0 failures" in r.stdout, out[-4000:]
    return out


def test_host_sanitizers_lowering():
    out = _run([], {}, 300)
    assert "3000 random contractions checked" in out


@pytest.mark.gpu
def test_host_sanitizers_gpu_paths():
    # the HIP runtime keeps some host allocations until exit: no leak report here
    out = _run(["--gpu"], {"ASAN_OPTIONS": "detect_leaks=0"}, 600)
    assert "gpu: 80 contractions" in out and "server: 12 jobs checked" in out
