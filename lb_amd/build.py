"""In-tree build of every native component (no JIT cache, nothing installed).

  lb_amd/liblbm_core_hip.so   HIP collide-stream core, gfx950 (include/lbm_core.h)
  lb_amd/liblbm_host.so       C++ host side: OFF voxelizer, boundary mesher,
                              mesh/VTK writers, solver driver (include/lbm_host.h)
  oracle/liblbm_oracle.so     CPU oracle (test infrastructure)
  oracle/libvoxel_oracle.so   brute-force voxelizer oracle (test infrastructure)
  tests/native/libcollide_host.so  the core's collision header compiled for the
                              host, checked bit-for-bit against the oracle

Usage: python -m lb_amd.build [core|host|oracle|tests|all]
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("LBM_OFFLOAD_ARCH", "gfx950")
# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = ["-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             # matrix-core results in ordinary VGPRs: the closure kernels read
             # every result with VALU instructions, and accumulators in AGPRs
             # cost one v_accvgpr_read per value (128 per wave in the step)
             "-mllvm", "-amdgpu-mfma-vgpr-form=true"]
CXX = shutil.which("g++") or "g++"
CC = shutil.which("gcc") or "gcc"


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... (exit {r.returncode})")
    # hipcc's host pass does not know the packed-fp32 feature, and the link
    # step does not use the code-generation option; those warnings are harmless
    noise = [ln for ln in (r.stderr or "").splitlines()
             if "packed-fp32-ops" not in ln and "amdgpu-mfma-vgpr-form" not in ln and ln.strip()]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


CORE_UNITS = ["lbm_core.hip", "closure.hip", "voxel.hip", "noalloc.hip"]


def build_core():
    objs = []
    jobs = []
    for u in CORE_UNITS:
        src = os.path.join(CSRC, u)
        obj = os.path.join(CSRC, u.replace(".hip", ".o"))
        objs.append(obj)
        jobs.append([HIPCC, *HIP_FLAGS, "-c", "-o", obj, src])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    out = os.path.join(HERE, "liblbm_core_hip.so")
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", out, *objs])
    return out


HOST_SOURCES = ["geometry_off.cpp", "geometry_boundary.cpp", "outputers.cpp", "solver.cpp", "capi.cpp"]


def build_host():
    srcs = [os.path.join(CSRC, "host", s) for s in HOST_SOURCES]
    srcs = [s for s in srcs if os.path.exists(s)]
    if not srcs:
        return None
    out = os.path.join(HERE, "liblbm_host.so")
    # links the HIP core (its C ABI); found next to this library at run time
    _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-Wall",
          f"-I{os.path.join(ROOT, 'include')}", "-o", out, *srcs, f"-L{HERE}", "-llbm_core_hip",
          "-Wl,-rpath,$ORIGIN"])
    return out


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), f"CC={CC}"])


def build_tests():
    nat = os.path.join(ROOT, "tests", "native")
    src = os.path.join(nat, "collide_host.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(nat, "libcollide_host.so")
    _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
          f"-I{CSRC}", f"-I{os.path.join(ROOT, 'include')}", "-o", out, src])
    return out


def build_all():
    build_oracle()
    build_tests()
    build_core()
    build_host()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"core": build_core, "host": build_host, "oracle": build_oracle, "tests": build_tests,
     "all": build_all}[what]()
