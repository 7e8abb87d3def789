"""Build recipe of the gfx950 library (in-tree, no JIT cache).

    python -m padenti_amd.build      # -> padenti_amd/libpadenti_hip.so

hipcc cross-compiles for gfx950 without a GPU.  Flags:
  -ffp-contract=off        every float op rounds as written (bit parity
                           with the CPU restatement of the reference)
  -fno-fast-math           IEEE division / sqrt, no reassociation
  -Xclang -target-feature -Xclang -packed-fp32-ops
                           no packed FP32 instructions are generated
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libpadenti_hip.so")
SOURCES = [os.path.join(HERE, "csrc", "pd_capi.hip"), os.path.join(HERE, "csrc", "pd_synth.cpp")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"),
]


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + HIP_FLAGS + SOURCES + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
    sys.exit(0)
