"""Build libskgemm.so (HIP, gfx950 only) in-tree with hipcc.

Run as a script or call build().  Always recompiles every source (no
incremental state), in parallel, then links.  Output:
stream-k-gemm_amd/libskgemm.so.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libskgemm.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# gfx950 only; packed FP32 instructions disabled on the device side.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-Wno-unused-command-line-argument", "-Wno-unknown-warning-option",
          "-Wno-unused-value", "-Wno-unused-result"]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-munsafe-fp-atomics",
          "-mllvm", "-amdgpu-mfma-vgpr-form"]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, name))
    return out


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON
    if src.endswith(".hip"):
        cmd += ["-x", "hip"] + DEVICE
    else:
        # Host translation units: any device pass hipcc runs targets gfx950.
        cmd += ["--offload-arch=gfx950", "-D__HIP_PLATFORM_AMD__", "-pthread"]
    cmd += ["-I", CSRC, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(jobs, 1)) as ex:
        results = list(ex.map(_compile, srcs))
    objs = []
    for src, obj, r in results:
        if verbose and r.stderr.strip():
            sys.stderr.write(r.stderr)
        if r.returncode != 0:
            raise RuntimeError("hipcc failed on %s:\n%s" % (src, r.stderr))
        objs.append(obj)
    tmp = LIB + ".tmp"
    # --no-undefined: a kernel template whose host stub was not emitted
    # must fail the build, not the first dlopen.
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-pthread",
           "-Wl,--no-undefined", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
