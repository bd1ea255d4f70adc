"""Host-side FFT planning (macromax_amd/csrc/mmx_plan.hpp) under the address and undefined-behaviour sanitizers.

The planning header is plain C++; tests/host_planning_test.cpp checks its axis factorisation, Bluestein and four-step
choices, fast division, k-vectors and tile shapes against independent statements of what they must satisfy. This test
builds it with g++ and runs it (about two seconds on one core).
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which('g++') is None, reason='g++ not installed')
def test_host_planning_sanitized(tmp_path):
    exe = str(tmp_path / 'host_planning_test')
    cmd = ['g++', '-std=c++17', '-O1', '-g', '-Wall', '-Wextra', '-Werror', '-fsanitize=address,undefined',
           '-fno-sanitize-recover=all', '-fno-omit-frame-pointer', '-o', exe,
           os.path.join(HERE, 'host_planning_test.cpp')]
    build = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS='detect_leaks=1:abort_on_error=0', UBSAN_OPTIONS='print_stacktrace=1')
    run = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-4000:])
    assert 'all checks passed' in run.stdout
