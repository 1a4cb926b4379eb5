"""Build recipe for libbsmps.so (gfx950).  Run ``python -m bosonsupremacy_amd.build``.

Each .hip translation unit is compiled with hipcc for gfx950 only, with
-ffp-contract=off (the hafnian and the Gamma epilogue reproduce NumPy's complex128
arithmetic bit for bit, so no expression may be fused beyond the explicit fma calls) and
with packed FP32 disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB = os.path.join(HERE, "libbsmps.so")
SOURCES = ["context.hip", "hafnian.hip", "kron.hip", "sampling.hip", "mlp.hip"]
OFFLOAD_ARCH = "gfx950"

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={OFFLOAD_ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]
# The split-form contraction keeps its accumulators across a two-stage k loop; left to its
# heuristics the register allocator puts them in VGPRs for one stage and AGPRs for the other
# and copies all 64 across twice per iteration.  The VGPR form of the MFMA operands avoids it.
SOURCE_FLAGS = {"sampling.hip": ["-mllvm", "--amdgpu-mfma-vgpr-form"]}


def _hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def build(verbose=False, extra_flags=()):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [_hipcc(), *HIPCC_FLAGS, *SOURCE_FLAGS.get(src, []), *extra_flags, "-I", INCLUDE, "-c",
               os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((src, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        for src, text in failed:
            sys.stderr.write(f"--- {src} ---\n{text}\n")
        raise RuntimeError("hipcc failed for " + ", ".join(s for s, _ in failed))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={OFFLOAD_ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
