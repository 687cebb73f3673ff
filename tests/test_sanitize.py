"""The library's host code (edge-list reader, host CSR build) and the CPU
oracle built with AddressSanitizer and UndefinedBehaviorSanitizer (host
compilers only) and driven through their C interfaces by
tests/sanitize/san_main.cpp: malformed and edge-case edge lists, random
multigraphs with tombstones against a direct CSR construction, and every
oracle routine on random graphs including the empty and one-node graphs."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="no host compilers")
def test_host_code_under_asan_ubsan(tmp_path):
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "sanitize")], check=True)
    exe = os.path.join(ROOT, "_build", "san", "san_main")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    p = subprocess.run([exe, str(tmp_path), os.path.join(HERE, "golden", "wiki_vote_sample.txt")],
                       capture_output=True, text=True, env=env, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ERROR: AddressSanitizer" not in p.stderr and "runtime error" not in p.stderr, p.stderr
    assert "all checks passed" in p.stdout
