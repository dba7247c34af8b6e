"""Build libtdscha_hip.so for gfx950 in-tree with hipcc (no JIT cache, no site-packages).

    python -m tdscha_amd._build        (from python-sscha_amd/)
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(PKG_DIR), "csrc")
INCLUDE = os.path.join(os.path.dirname(os.path.dirname(PKG_DIR)), "include")
LIB_DIR = os.path.join(PKG_DIR, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libtdscha_hip.so")
TORCH_OPS_PATH = os.path.join(LIB_DIR, "libtdscha_torch.so")

SOURCES = ["tdscha_hip.cpp", "rotate.hip", "gemm.hip", "gemm_bf16.hip", "gemm_bf16_rp.hip", "operator.hip",
           "lanczos.hip", "project.hip", "symmetry.hip", "perturb.hip", "images.hip",
           "projector.hip", "mfma_stream.hip"]
ARCH = os.environ.get("TDSCHA_OFFLOAD_ARCH", "gfx950")
# Packed-FP32 instructions are disabled for every device pass.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    obj_dir = os.path.join(LIB_DIR, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    cc = hipcc()
    common = ["-O3", "-fPIC", "-std=c++17", "-I", INCLUDE, "-I", CSRC, "--offload-arch=" + ARCH,
              "-Wall", "-Wno-unused-result"]

    def compile_one(src):
        path = os.path.join(CSRC, src)
        obj = os.path.join(obj_dir, src + ".o")
        lang = ["-x", "hip"]   # tdscha_hip.cpp holds a device array (the no-allocation scratch)
        cmd = [cc] + lang + common + DEVICE_FLAGS + ["-c", path, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
        if verbose:
            msg = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
            if msg.strip():
                print(msg)
        return obj

    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB_PATH + ".tmp"
    cmd = [cc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB_PATH)
    build_torch_ops(cc, verbose)
    return LIB_PATH


def build_torch_ops(cc, verbose=False):
    """The PyTorch-ROCm operators (csrc/torch_ops.cpp) as lib/libtdscha_torch.so, linked
    against libtdscha_hip.so and torch's own libraries (loaded with torch.ops.load_library)."""
    import torch
    tdir = os.path.dirname(torch.__file__)
    inc = [os.path.join(tdir, "include"), os.path.join(tdir, "include", "torch", "csrc", "api",
                                                         "include")]
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    src = os.path.join(CSRC, "torch_ops.cpp")
    tmp = TORCH_OPS_PATH + ".tmp"
    cmd = [cc, "-O2", "-fPIC", "-shared", "-std=c++17", "-D__HIP_PLATFORM_AMD__=1",
           "-DUSE_ROCM=1", "-D_GLIBCXX_USE_CXX11_ABI=%d" % abi, "-Wno-unused-result"]
    for d in inc:
        cmd += ["-I", d]
    cmd += [src, "-o", tmp, "-L", LIB_DIR, "-ltdscha_hip", "-L", os.path.join(tdir, "lib"),
            "-lc10", "-lc10_hip", "-ltorch_cpu", "-ltorch_hip", "-ltorch",
            "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(tdir, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("torch operator build failed: %s\n%s\n%s" % (" ".join(cmd),
                                                                          r.stdout, r.stderr))
    if verbose and r.stderr.strip():
        print(r.stderr)
    os.replace(tmp, TORCH_OPS_PATH)
    return TORCH_OPS_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
