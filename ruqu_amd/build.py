"""Builds libruqu_gpu.so (HIP kernels for gfx950 + the C++ host runtime).

Every source under ruqu_amd/csrc is compiled with hipcc for gfx950 and linked
into ruqu_amd/libruqu_gpu.so.  Objects go to ruqu_amd/_build/.  Runs without a
GPU (hipcc cross-compiles).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libruqu_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-std=c++17",
    "-O3",
    "-fPIC",
    "-ffp-contract=off",
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith(".hip") or name.endswith(".cpp"):
            out.append(os.path.join(CSRC, name))
    return out


# Per-file extra flags.  The fused kernel dispatches on runtime register bits
# through switches whose bodies index a register array with constants;
# SimplifyCFG's common-instruction sinking would merge those bodies into
# pointer phis and force the array into scratch memory.  Its op dispatch is
# wave-uniform (scalar op records): left unstructurized, the switches compile
# to plain scalar branch trees instead of predicate-flag chains.  Its
# matrix-core phases accumulate in ordinary VGPRs: with accumulation registers
# the amplitudes of every phase would live in them after the phase and the
# VALU phases pay 32 extra moves each way.
EXTRA = {
    "kernels_fused.hip": ["-mllvm", "-sink-common-insts=false",
                          "-mllvm", "-structurizecfg-skip-uniform-regions=true",
                          "-mllvm", "-amdgpu-mfma-vgpr-form"],
}


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + EXTRA.get(os.path.basename(src), []) + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    failed = False
    for src, _obj, rc, log in results:
        if rc != 0:
            failed = True
            sys.stderr.write("hipcc failed for %s:\n%s\n" % (src, log))
        elif verbose and log.strip():
            sys.stderr.write(log)
    if failed:
        raise RuntimeError("libruqu_gpu build failed")
    objs = [o for _s, o, _rc, _l in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("libruqu_gpu link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
