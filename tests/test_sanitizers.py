"""The host code and the oracle under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only; GPU
sanitizers are not part of this toolchain): the library's host translation units are rebuilt with
hipcc -Xarch_host -fsanitize=address,undefined (device code as always) and the oracle with gcc's
sanitizers, each into a temporary directory, and the CPU suite runs against each build in a child
process with the matching sanitizer runtime preloaded. Any report fails the run (UBSan does not
recover; ASan aborts)."""
import glob
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _child_suite(env, tmp_path, name):
    log = tmp_path / f"{name}.log"
    cmd = [sys.executable, "-m", "pytest", os.path.join(ROOT, "tests"), "-m", "not gpu", "-q", "-x",
           "-p", "no:cacheprovider", "--ignore", os.path.join(ROOT, "tests", "test_sanitizers.py")]
    with open(log, "w") as f:
        rc = subprocess.run(cmd, cwd=ROOT, env=env, stdout=f, stderr=subprocess.STDOUT, timeout=900).returncode
    reports = sorted(glob.glob(str(tmp_path / f"{name}_san*")))
    text = "".join(open(p).read()[:4000] for p in reports)
    assert rc == 0 and not reports, f"{name}: rc {rc}\n{text}\n{open(log).read()[-3000:]}"


def _env(tmp_path, name, preload, **extra):
    env = dict(os.environ)
    env.update(LD_PRELOAD=preload, LABINIUS_NO_TORCH="1",
               ASAN_OPTIONS=f"detect_leaks=0:log_path={tmp_path}/{name}_san_asan",
               UBSAN_OPTIONS=f"print_stacktrace=1:log_path={tmp_path}/{name}_san_ubsan", **extra)
    return env


@pytest.mark.slow
def test_host_library_under_asan_and_ubsan(tmp_path):
    from labinius_amd import build
    cc = build.hipcc()
    rt = subprocess.run([cc, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True).stdout
    rt = rt.strip()
    if not rt or not os.path.exists(rt):
        rt = next(iter(glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")), "")
    if not rt:
        pytest.skip("no clang ASan runtime in this toolchain")
    lib = build.build(sanitize=True, out_path=str(tmp_path / "liblabinius_san.so"))
    _child_suite(_env(tmp_path, "host", rt, LABINIUS_LIB=lib), tmp_path, "host")


@pytest.mark.slow
def test_oracle_under_asan_and_ubsan(tmp_path):
    gcc = shutil.which("gcc")
    if not gcc:
        pytest.skip("no gcc")
    so = str(tmp_path / "liboracle_san.so")
    subprocess.run([gcc, "-O1", "-g", "-std=c11", "-fPIC", "-ffp-contract=off", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-shared", "-o", so,
                    os.path.join(ROOT, "oracle", "oracle.c"), "-lm"], check=True)
    rts = [subprocess.run([gcc, f"-print-file-name={n}"], capture_output=True, text=True).stdout.strip()
           for n in ("libasan.so", "libubsan.so")]
    _child_suite(_env(tmp_path, "oracle", ":".join(rts), ORACLE_LIB=so), tmp_path, "oracle")
