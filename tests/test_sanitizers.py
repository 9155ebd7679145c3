"""The host grid pipeline (libvirs_host.so) and the CPU oracle under
AddressSanitizer and UndefinedBehaviorSanitizer.

The reference's Debug build compiles its host code with
-fsanitize=address,undefined (CMakeLists.txt:50).  Here the CPU test suite
runs again in a child process against sanitized builds of both libraries
(host code only: the HIP library is the ordinary build, and no GPU is
touched); any sanitizer report aborts the child and fails this test."""
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    path = subprocess.check_output(["gcc", f"-print-file-name={name}"], text=True).strip()
    if not os.path.isabs(path) or not os.path.exists(path):
        pytest.skip(f"{name} not found by gcc")
    return path


def test_cpu_suite_under_asan_and_ubsan(tmp_path):
    if os.environ.get("VIRS_SANITIZED"):
        pytest.skip("this is the sanitized run itself")
    hip = os.path.join(ROOT, "virs_amd", "lib", "libvirs_hip.so")
    if not os.path.exists(hip):
        pytest.skip("libvirs_hip.so not built")
    libdir = tmp_path / "lib"
    subprocess.run([sys.executable, "-m", "virs_amd.build", "--sanitize", str(libdir)], cwd=ROOT, check=True,
                   capture_output=True)
    os.symlink(hip, libdir / "libvirs_hip.so")
    subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "SANITIZE=1"], check=True,
                   capture_output=True)
    oracle_so = os.path.join(ROOT, "oracle", "_build_san", "libvirs_oracle.so")
    env = dict(os.environ,
               LD_PRELOAD=f"{_runtime('libasan.so')}:{_runtime('libubsan.so')}",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               VIRS_LIB_DIR=str(libdir), VIRS_ORACLE_SO=oracle_so, VIRS_SANITIZED="1")
    # the sanitized builds are the ones the child loads
    probe = ("import virs_amd._lib as L; from oracle import oracle as O; L.host(); O.lib(); "
             "print(open('/proc/self/maps').read())")
    maps = subprocess.run([sys.executable, "-c", probe], env=env, cwd=ROOT, capture_output=True, text=True,
                          timeout=300)
    assert maps.returncode == 0, maps.stderr[-2000:]
    assert str(libdir / "libvirs_host.so") in maps.stdout
    assert oracle_so in maps.stdout
    assert "libasan" in maps.stdout
    r = subprocess.run([sys.executable, "-m", "pytest", "tests", "-x", "-q", "-m", "not gpu",
                        "-p", "no:cacheprovider"], env=env, cwd=ROOT, capture_output=True, text=True,
                       timeout=1200)
    tail = r.stdout[-3000:] + r.stderr[-3000:]
    assert r.returncode == 0, tail
    assert "AddressSanitizer" not in tail and "runtime error" not in tail, tail
    m = re.search(r"(\d+) passed", r.stdout)
    assert m and int(m.group(1)) >= 80, tail
    print(r.stdout.strip().splitlines()[-1], flush=True)
