"""Host code under AddressSanitizer and UBSan.

`make -C oracle sanitize` builds, with g++/gcc and -fsanitize=address,undefined,
the product's host mesh layer (generator, SU2/topology writer and reader,
block builder, memory-order permutations) together with the oracle, and
oracle/sanitize_run.cpp drives them over three meshes (single and
multi-block, structured and scrambled numbering) and every scheme of the
reference's factories. A sanitizer finding aborts the run with its report.
Host code only; nothing here runs on a GPU.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-j8", "sanitize"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([os.path.join(ROOT, "oracle", "_san", "sanitize_run"), str(tmp_path)], env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and "sanitize ok" in r.stdout, r.stderr[-4000:]
