"""Builds the in-tree native libraries of cucm_amd.

  libcucm.so          HIP kernels (gfx950) + host engine + C ABI (include/cucm.h)
  libcucm_datagen.so  seeded synthetic input generators (host, OpenMP)
  libcucm_probe.so    measurement kernels (gather ceiling, copy calibration)
  bin/cucm_compress, bin/cucm_decompress   drop-in command-line tools

Run:  python -m cucm_amd.build   (or __graft_entry__.build()).
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CUCM_ARCH", "gfx950")

LIB = os.path.join(HERE, "libcucm.so")
BIN = os.path.join(HERE, "bin")
CLIS = ["cucm_compress", "cucm_decompress"]
DATAGEN = os.path.join(HERE, "libcucm_datagen.so")

HIP_SOURCES = ["cucm_kernels.hip", "cucm_engine.cpp", "cucm_host_model.cpp", "cucm_serve.cpp",
               "cucm_stream.cpp", "cucm_noalloc.hip"]
# measurement kernels (bench.py's ceilings), a library of their own
PROBE_DIR = os.path.join(HERE, "probe")
PROBE_SOURCES = ["gather_probe.hip"]
PROBE_LIB = os.path.join(HERE, "libcucm_probe.so")
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    # packed FP32 instructions stay off (no float math on this path anyway)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC, "-x", "hip"] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        jobs.append((cmd, obj))
    probe_jobs = []
    for src in PROBE_SOURCES:
        obj = os.path.join(OBJ, "probe_" + os.path.splitext(src)[0] + ".o")
        probe_jobs.append(([HIPCC] + HIP_FLAGS + ["-I" + PROBE_DIR, "-c", os.path.join(PROBE_DIR, src),
                           "-o", obj], obj))
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(jobs) + len(probe_jobs) + 1) as ex:
        futs = [ex.submit(_run, cmd) for cmd, _ in jobs + probe_jobs]
        dg = ex.submit(_run, ["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c11",
                              "-o", DATAGEN + ".tmp", os.path.join(CSRC, "datagen.c"), "-lm"])
        for f in futs:
            out = f.result()
            if verbose and out.strip():
                print(out)
        dg.result()
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-o", LIB + ".tmp"] + [o for _, o in jobs])
    os.replace(LIB + ".tmp", LIB)
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-o", PROBE_LIB + ".tmp"] +
         [o for _, o in probe_jobs])
    os.replace(PROBE_LIB + ".tmp", PROBE_LIB)
    os.replace(DATAGEN + ".tmp", DATAGEN)
    # drop-in command-line tools: host C++ over the C ABI
    os.makedirs(BIN, exist_ok=True)
    for tool in CLIS:
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"),
              os.path.join(HERE, "cli", tool + ".cpp"), "-o", os.path.join(BIN, tool),
              "-L" + HERE, "-lcucm", "-Wl,-rpath,$ORIGIN/.."])
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB, DATAGEN)
