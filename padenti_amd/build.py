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
SOURCES = [os.path.join(HERE, "csrc", n) for n in ("pd_capi.hip", "pd_ingest.hip", "pd_synth.cpp", "pd_png.cpp")]
LIBS = ["-lz"]  # PNG streams
# example of a separately compiled feature hook (include/padenti/pd_plugin.hpp)
PLUGIN_LIB = os.path.join(HERE, "libpd_example_plugin.so")
PLUGIN_SOURCES = [os.path.join(HERE, "examples", "depth_plugin.hip")]
# one hook written for any element type / channel count / FeatType, ten
# instantiations (every FeatType, up to 16 channels) in two plugin libraries
# (a plugin library holds one translation unit of the engine)
GENERIC_LIBS = {os.path.join(HERE, "libpd_channel_diff_%s.so" % k): [os.path.join(HERE, "examples",
                                                                               "channel_diff_%s.hip" % k)]
                for k in ("a", "b")}
# an auto-context hook (reads the tree being trained and the node-id map)
GENERIC_LIBS[os.path.join(HERE, "libpd_context_plugin.so")] = [os.path.join(HERE, "examples", "context_plugin.hip")]
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
    "-fvisibility=hidden",
    "-fvisibility-inlines-hidden",
    "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _compile(sources, out, verbose, libs=()):
    """Objects side by side (one hipcc per source), then one link."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    flags = [f for f in HIP_FLAGS if f != "-shared"]
    objs = [os.path.join(objdir, os.path.basename(out) + "." + os.path.basename(s) + ".o") for s in sources]
    with ThreadPoolExecutor(len(sources)) as ex:
        for j in [ex.submit(_run, [HIPCC] + flags + ["-c", s, "-o", o], verbose) for s, o in zip(sources, objs)]:
            j.result()
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + list(libs) + ["-o", tmp], verbose)
    os.replace(tmp, out)


DEMO_SRC = os.path.join(ROOT, "tests", "cpp", "api_demo.cpp")
DEMO_BIN = os.path.join(ROOT, "tests", "cpp", "api_demo")


def build_demo(verbose: bool = False) -> str:
    """C++ drop-in API demo (plain g++, links the library and the plugin)."""
    cmd = ["g++", "-std=c++11", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), DEMO_SRC,
           "-L" + HERE, "-lpadenti_hip", "-lpd_example_plugin", "-Wl,-rpath,$ORIGIN/../../padenti_amd",
           "-Wl,-rpath-link,/opt/rocm/lib", "-o", DEMO_BIN]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return DEMO_BIN


def build(verbose: bool = False, plugin: bool = True) -> str:
    if plugin:
        # the two libraries compile independently: run them side by side
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(2 + len(GENERIC_LIBS)) as ex:
            jobs = [ex.submit(_compile, SOURCES, LIB, verbose, LIBS), ex.submit(_compile, PLUGIN_SOURCES, PLUGIN_LIB, verbose)]
            jobs += [ex.submit(_compile, src, lib, verbose) for lib, src in GENERIC_LIBS.items()]
            for j in jobs:
                j.result()
        build_demo(verbose)
    else:
        _compile(SOURCES, LIB, verbose, LIBS)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
    sys.exit(0)
