"""Build recipe for the native parts of mshqc_amd (in-tree, no JIT cache).

* ``libmshqc_hip.so``  -- HIP kernels + C ABI (hipcc, gfx950 only)
* ``_mshqc*.so``       -- pybind11 module with the drop-in C++ classes
                          (g++, links libmshqc_hip.so with an $ORIGIN rpath)

Run ``python -m mshqc_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import glob
import os
import re
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIP_DIR = os.path.join(HERE, "csrc", "hip")
HOST_DIR = os.path.join(HERE, "csrc", "host")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_LIB = os.path.join(HERE, "libmshqc_hip.so")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # Keep hipcc from emitting packed FP32 instructions on its own.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-Wno-unused-value",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


# Sources whose kernels stream operands into LDS with direct-to-LDS loads and
# wait for them with fixed vector-memory counts (s_waitcnt vmcnt(N)).  Their
# kernels must compile without scratch: a spilling build adds vector-memory
# traffic and spill sequences those waits were not written for, and one such
# build (20 bytes per lane in the five-slice pair kernel) returned wrong
# amplitudes.  The build fails instead of shipping one.
NO_SCRATCH_SOURCES = ("mp2_bf16.hip",)


def scratch_sizes(remarks: str) -> dict:
    """{kernel: scratch bytes per lane} from hipcc's kernel-resource-usage remarks."""
    out, name = {}, None
    for line in remarks.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            continue
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name:
            out[name] = int(m.group(1))
    return out


def build_hip(jobs: int = 8) -> str:
    """Compile every .hip source to an object in parallel, then link."""
    srcs = sorted(glob.glob(os.path.join(HIP_DIR, "*.hip")))
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    procs = []
    objs = []

    def finish(item):
        src, proc = item
        _, err = proc.communicate()
        sys.stderr.write("\n".join(l for l in err.splitlines() if "remark:" not in l) + "\n")
        if proc.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}")
        if os.path.basename(src) in NO_SCRATCH_SOURCES:
            bad = {k: v for k, v in scratch_sizes(err).items() if v > 0}
            if bad:
                raise RuntimeError(f"{os.path.basename(src)}: kernels use scratch {bad}; the DMA "
                                   "pipelines need spill-free kernels")

    for s in srcs:
        o = os.path.join(objdir, os.path.basename(s) + ".o")
        objs.append(o)
        cmd = [hipcc, *HIPCC_FLAGS, "-c", s, "-o", o]
        if os.path.basename(s) in NO_SCRATCH_SOURCES:
            cmd.append("-Rpass-analysis=kernel-resource-usage")
        print("+", " ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stderr=subprocess.PIPE, text=True)))
        if len(procs) >= jobs:
            finish(procs.pop(0))
    for item in procs:
        finish(item)
    tmp = HIP_LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def _ext_suffix() -> str:
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def build_host() -> str:
    """Compile the pybind11 module `_mshqc` against libmshqc_hip.so."""
    import pybind11

    srcs = sorted(glob.glob(os.path.join(HOST_DIR, "*.cc")))
    if not srcs:
        return ""
    out = os.path.join(HERE, "_mshqc" + _ext_suffix())
    tmp = out + ".tmp"
    cmd = [
        "g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall",
        "-Wno-unused-function",
        f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
        f"-I{os.path.join(ROOT, 'include')}", f"-I{HOST_DIR}",
        *srcs, "-o", tmp,
        f"-L{HERE}", "-l:libmshqc_hip.so", "-Wl,-rpath,$ORIGIN",
    ]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_hip()
    build_host()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"hip": build_hip, "host": build_host, "all": build_all}[what]()
