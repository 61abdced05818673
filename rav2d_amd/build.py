"""Build the native libraries in-tree.

    python -m rav2d_amd.build            # everything
    python -m rav2d_amd.build --no-oracle

Outputs (next to the sources, git-ignored, shipped with the tree):
  rav2d_amd/librav2d_gpu.so    gfx950 HIP kernels + the C ABI of include/rav2d_gpu.h
  rav2d_amd/librav2d_synth.so  seeded synthetic workloads (tests, benchmark)
  oracle/liboracle.so          CPU oracle (test infrastructure)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")

HIP_SOURCES = ["runtime.hip", "deblock.hip", "cdef.hip", "lr.hip", "grain.hip", "batch.hip", "serve.hip", "probe.hip", "dsp.hip"]
ARCH = "gfx950"


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=cwd)


def build_gpu(jobs=8):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    out = os.path.join(HERE, "librav2d_gpu.so")
    objs = []
    flags = [
        "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
        # keep the compiler from selecting packed-fp32 instructions
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument", "-Wno-missing-braces",
    ]
    procs = []
    for src in HIP_SOURCES:
        obj = os.path.join(HERE, "_build", src.replace(".hip", ".o"))
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        objs.append(obj)
        cmd = [hipcc] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    for p in procs:
        if p.wait() != 0:
            raise SystemExit("hipcc failed")
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", out] + objs)
    return out


def build_synth():
    out = os.path.join(HERE, "librav2d_synth.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wno-missing-braces",
          os.path.join(CSRC, "synth.cpp"), "-o", out])
    return out


def build_oracle():
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-B"])


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    build_synth()
    if "--no-oracle" not in argv:
        build_oracle()
    build_gpu()


if __name__ == "__main__":
    main()
