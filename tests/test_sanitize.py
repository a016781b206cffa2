"""AddressSanitizer and UndefinedBehaviorSanitizer builds of the host code.

- The CPU oracle (oracle/bs_oracle.c, the checker and the CPU baseline),
  built with gcc -fsanitize=address,undefined and run by
  tests/sanitize/oracle_driver.c over random wakes and edge cases.
- The engine library's host code (the C ABI, argument parsing, option
  validation, workspace sizing, object set-up and tear-down), built with
  hipcc with the sanitizers on the host side only (device code is compiled
  as usual), run by tests/sanitize/abi_driver.c: without a GPU every entry
  point must return an error code with a message.

Leaks are checked too (the HIP runtime's own allocations at exit are
suppressed).  Any report aborts the driver and fails the test.
"""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

from conftest import gpu_available

ROOT = Path(__file__).resolve().parents[1]
SAN = ROOT / "tests" / "sanitize"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))


def _run(cmd, env=None, timeout=600):
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=timeout)
    return r.returncode, r.stdout + r.stderr


def _env(tmp):
    supp = tmp / "lsan.supp"
    supp.write_text("leak:libamdhip64\nleak:libhsa-runtime64\nleak:libdrm\nleak:libnuma\n")
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:halt_on_error=1"
    env["LSAN_OPTIONS"] = f"suppressions={supp}"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["OMP_NUM_THREADS"] = "2"
    return env


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_under_sanitizers(tmp_path):
    exe = tmp_path / "oracle_driver"
    rc, out = _run(["gcc", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-ffp-contract=off", "-fopenmp",
                    str(ROOT / "oracle" / "bs_oracle.c"), str(SAN / "oracle_driver.c"),
                    "-o", str(exe), "-lquadmath", "-lm"])
    assert rc == 0, out
    rc, out = _run([str(exe)], env=_env(tmp_path))
    assert rc == 0 and "oracle driver: ok" in out, out


@pytest.mark.skipif(not (ROCM / "bin" / "hipcc").exists(), reason="needs hipcc")
@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error paths")
def test_abi_host_code_under_sanitizers(tmp_path):
    sys.path.insert(0, str(ROOT / "winds-master_amd"))
    from winds import build
    host_san = []
    for f in ("-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=all",
              "-fno-omit-frame-pointer"):
        host_san += ["-Xarch_host", f]
    flags = [f for f in build.common_flags() if f != "-O3"] + ["-O1", "-g"] + host_san
    procs, objs = [], []
    for src in build.SOURCES:
        obj = tmp_path / (Path(src).stem + ".o")
        objs.append(str(obj))
        procs.append(subprocess.Popen([build.hipcc(), *flags, "-x", "hip", "-c",
                                       str(build.CSRC / src), "-o", str(obj)],
                                      stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))
    for p in procs:
        _, err = p.communicate(timeout=900)
        assert p.returncode == 0, build._quiet(err)
    lib = tmp_path / "libwinds_san.so"
    rc, out = _run([build.hipcc(), "-shared", "-fPIC", f"--offload-arch={build.ARCH}",
                    "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                    *objs, "-o", str(lib)])
    assert rc == 0, out
    exe = tmp_path / "abi_driver"
    rc, out = _run([str(ROCM / "llvm" / "bin" / "clang"), "-O1", "-g",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    f"-I{ROOT / 'include'}", str(SAN / "abi_driver.c"), str(lib),
                    f"-Wl,-rpath,{tmp_path}", f"-L{ROCM / 'lib'}", "-lamdhip64",
                    f"-Wl,-rpath,{ROCM / 'lib'}", "-lm", "-o", str(exe)])
    assert rc == 0, out
    rc, out = _run([str(exe)], env=_env(tmp_path))
    assert rc == 0 and "abi driver: ok" in out, out
