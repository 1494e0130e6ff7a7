"""Build libdfea_hip.so in-tree with hipcc for gfx950 (no torch extension machinery, no hipify)."""
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
REPO = os.path.dirname(PKG_ROOT)
SOURCES = [os.path.join(PKG_ROOT, "csrc", "dfea_kernels.hip")]
INCLUDE = os.path.join(REPO, "include")
OUT = os.path.join(_HERE, "libdfea_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions stay off (they have given silently wrong results on these boxes)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the first kernel arguments arrive in SGPRs (no kernarg-segment load before the descriptor load): measured
    # 20.9 -> 20.3 us per 400x400 FP64 product, 11.9 -> 11.3 us FP32
    "-mllvm", "-amdgpu-kernarg-preload-count=8",
]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-I" + INCLUDE, *SOURCES, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
