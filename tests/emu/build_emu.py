"""Builds tests/emu/libbatch_emu.so: the batched kernel's source compiled for
the host with the thread-emulation shim (test infrastructure, CPU only)."""
import os
import re
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(ROOT, "ruqu_amd", "csrc", "kernels_batch.hip")
OUT = os.path.join(HERE, "libbatch_emu.so")


def build():
    s = open(SRC).read()
    s = s.replace("#include <hip/hip_runtime.h>", "")
    s = s.replace('#include "batch.h"', '#include "%s"' % os.path.join(ROOT, "ruqu_amd", "csrc", "batch.h"))
    s = re.sub(r"extern __shared__ float2 (\w+)\[\];", r"float2* \1 = (float2*)emu::lds;", s)
    s = s.replace('asm("v_mov_b32 %0, %1" : "=v"(r) : "v"(x));', "r = x;")  # register copy
    s = s.replace("__shared__ ", "static ")  # block-shared = one copy (blocks run one at a time)
    inc = os.path.join(HERE, "batch_kernel_src.inc")
    with open(inc, "w") as f:
        f.write(s)
    subprocess.run(["g++", "-std=c++20", "-O2", "-fPIC", "-shared", "-pthread", "-I", HERE,
                    os.path.join(HERE, "batch_emu.cpp"), "-o", OUT], check=True)
    return OUT


if __name__ == "__main__":
    print(build())
