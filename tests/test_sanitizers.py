"""AddressSanitizer / UBSan builds on the CPU (no GPU needed).

* The Schwarz host analysis (aboria_amd/csrc/schwarz_analysis.hpp, the code
  the library runs before its device factorisation) at the size of the
  512-problem batch (about 8 million 2D points, 64 per bucket) and of the
  400,056-point cell list, with the invariants every domain list must hold.
* The oracle (oracle/oracle_capi.cpp) through every entry point on small
  inputs in 1-3 dimensions.

Both are built with g++ only (no device code), -fsanitize=address,undefined
and -fno-sanitize-recover, so any finding fails the run.
"""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLAGS = ["-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-fno-omit-frame-pointer"]


def _build_and_run(srcs, extra, timeout):
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "a.out")
        subprocess.run(["g++"] + FLAGS + extra + srcs + ["-o", exe], check=True)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
                   OMP_NUM_THREADS="4")
        r = subprocess.run([exe], capture_output=True, text=True, timeout=timeout, env=env)
        print(r.stdout, r.stderr)
        assert r.returncode == 0, r.stderr[-4000:]
        assert "ok" in r.stdout


def test_schwarz_analysis_asan():
    _build_and_run([os.path.join(ROOT, "tests", "cpp", "asan_schwarz_analysis.cpp")], [], 600)


def test_oracle_asan():
    _build_and_run([os.path.join(ROOT, "tests", "cpp", "asan_oracle.cpp"),
                    os.path.join(ROOT, "oracle", "oracle_capi.cpp")],
                   ["-fopenmp", "-ffp-contract=off", "-I" + os.path.join(ROOT, "oracle")], 600)
