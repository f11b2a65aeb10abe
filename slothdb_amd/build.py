"""Builds the native libraries in-tree.

  slothdb_amd/lib/libslothdb_gpu.so      the GROUP BY operator (HIP, gfx950) + C ABI
  slothdb_amd/lib/libslothdb_datagen.so  seeded synthetic inputs (host C++)
  slothdb_amd/lib/libslothdb_gpu_op.so   C++20 PhysicalGpuHashAggregate over the C ABI
  slothdb_amd/lib/test_operator          its C++ test program (tests/cpp/test_operator.cpp)
  slothdb_amd/lib/test_fixed_point       host test of the exact double SUM arithmetic

Run as `python -m slothdb_amd.build` or through __graft_entry__.build().
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
CSRC = os.path.join(HERE, "csrc")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


GEN = os.path.join(LIB, "gen")


def gen_selector_builtin():
    """The committed trained selector (slothdb_amd/selector_weights.json) as a
    C array the library compiles in: AUTO plans consult it by default."""
    import json
    with open(os.path.join(HERE, "selector_weights.json")) as f:
        params = json.load(f)["params"]
    os.makedirs(GEN, exist_ok=True)
    body = ",\n".join(", ".join(repr(float(v)) + "f" for v in params[i:i + 8]) for i in range(0, len(params), 8))
    text = ("// generated by slothdb_amd/build.py from slothdb_amd/selector_weights.json\n"
            f"static const float kSelectorBuiltin[{len(params)}] = {{\n{body}\n}};\n")
    path = os.path.join(GEN, "selector_builtin.inc")
    with open(path + ".tmp", "w") as f:
        f.write(text)
    os.replace(path + ".tmp", path)
    return path


def build_gpu():
    os.makedirs(LIB, exist_ok=True)
    gen_selector_builtin()
    out = os.path.join(LIB, "libslothdb_gpu.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-I", GEN, "-o", tmp, os.path.join(CSRC, "gbagg.hip")])
    os.replace(tmp, out)
    return out


def build_datagen():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libslothdb_datagen.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-o", tmp,
          os.path.join(CSRC, "datagen.cpp")])
    os.replace(tmp, out)
    return out


def build_host():
    """The C++20 operator wrapper and its test program, linked to the HIP library."""
    os.makedirs(LIB, exist_ok=True)
    host = os.path.join(HERE, "host")
    so = os.path.join(LIB, "libslothdb_gpu_op.so")
    _run(["g++", "-std=c++20", "-O2", "-Wall", "-fPIC", "-shared", "-o", so + ".tmp",
          os.path.join(host, "gpu_hash_aggregate.cpp"), "-L" + LIB, "-lslothdb_gpu",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(so + ".tmp", so)
    exe = os.path.join(LIB, "test_operator")
    _run(["g++", "-std=c++20", "-O2", "-Wall", "-o", exe + ".tmp",
          os.path.join(ROOT, "tests", "cpp", "test_operator.cpp"), "-L" + LIB, "-lslothdb_gpu_op",
          "-lslothdb_gpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(exe + ".tmp", exe)
    return so


def build_fixed_point_test():
    """Host-side test program of the exact double SUM arithmetic
    (tests/cpp/test_fixed_point.cpp), which includes the HIP device header."""
    os.makedirs(LIB, exist_ok=True)
    exe = os.path.join(LIB, "test_fixed_point")
    _run([HIPCC, "-x", "hip", f"--offload-arch={ARCH}", "-O2", "-std=c++17",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
          "-o", exe + ".tmp", os.path.join(ROOT, "tests", "cpp", "test_fixed_point.cpp")])
    os.replace(exe + ".tmp", exe)
    return exe


def build_oracle():
    """The CPU oracle (test infrastructure) and, with the reference tree
    present, oracle/_ref built from the reference's own aggregator sources."""
    make = shutil.which("make") or "make"
    return _run([make, "-s", "-B", "-C", os.path.join(ROOT, "oracle")])


def build_engine_test():
    """With the reference tree present: oracle/_ref/test_engine_operator, the
    engine-side operator over the reference's own DataChunk / Vector / Value
    sources, linked against the library built by build_gpu()."""
    make = shutil.which("make") or "make"
    return _run([make, "-s", "-C", os.path.join(ROOT, "oracle"), "engine"])


def build_all():
    build_datagen()
    build_oracle()
    build_gpu()
    build_host()
    build_fixed_point_test()
    build_engine_test()


if __name__ == "__main__":
    build_all()
    print("built", file=sys.stderr)
