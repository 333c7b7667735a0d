"""Builds the HIP/C++ library of the prover in-tree: labinius_amd/liblabinius.so (gfx950).

The kernels are compiled for gfx950 only, with packed FP32 disabled; host translation units
keep floating-point contraction off (the challenge sampler's norm test is bit-sensitive).
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblabinius.so")
HOST_SOURCES = ["ring.cpp", "blake3.cpp", "protocol.cpp", "prover.cpp", "dropped.cpp", "batch.cpp", "switch.cpp", "wire.cpp", "verify.cpp",
                "noalloc.cpp", "pool.cpp", "server.cpp"]
DEVICE_SOURCES = ["kernels.hip"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False, sanitize=False, out_path=None):
    """sanitize: the host code with AddressSanitizer and UndefinedBehaviorSanitizer (device code as
    always), into `out_path`; such a library runs with the clang ASan runtime preloaded."""
    cc = hipcc()
    common = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
              "-I", os.path.join(HERE, "..", "include")]
    if sanitize:
        common = ["-O1", "-g"] + common[1:] + ["-Xarch_host", "-fsanitize=address", "-Xarch_host",
                                               "-fsanitize=undefined", "-Xarch_host", "-fno-sanitize-recover=undefined",
                                               "-Xarch_host", "-fno-omit-frame-pointer"]
    tmp = tempfile.mkdtemp(prefix="lbn_build_")
    objs = []
    jobs = []
    try:
        for src in HOST_SOURCES:
            obj = os.path.join(tmp, src + ".o")
            cmd = [cc, *common, "-x", "hip", "--offload-arch=gfx950", "-c", os.path.join(CSRC, src), "-o", obj,
                   "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
            jobs.append((src, cmd))
            objs.append(obj)
        for src in DEVICE_SOURCES:
            obj = os.path.join(tmp, src + ".o")
            cmd = [cc, *common, "--offload-arch=gfx950", "-c", os.path.join(CSRC, src), "-o", obj,
                   "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
            jobs.append((src, cmd))
            objs.append(obj)
        procs = [(src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
                 for src, cmd in jobs]
        failed = []
        for src, p in procs:
            out, _ = p.communicate()
            lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln and ln.strip()
                     and "is not a recognized feature" not in ln]
            if p.returncode != 0:
                failed.append(src)
                sys.stderr.write(out)
            elif verbose and lines:
                sys.stderr.write("\n".join(lines) + "\n")
        if failed:
            raise RuntimeError("compilation failed: " + ", ".join(failed))
        lib_path = out_path or LIB
        tmp_lib = lib_path + ".tmp"
        link = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-shared-libsan"] if sanitize else []
        subprocess.run([cc, "-shared", "--offload-arch=gfx950", *link, "-o", tmp_lib, *objs], check=True)
        os.replace(tmp_lib, lib_path)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return lib_path


if __name__ == "__main__":
    print(build(verbose=True))
