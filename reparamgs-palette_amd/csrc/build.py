"""Build librgp.so (the gfx950 HIP kernels + C ABI) in-tree with hipcc.

Flags:
  --offload-arch=gfx950                      CDNA4 only
  -ffp-contract=off                          every FP32 op rounded on its own
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE division and sqrt
  -fno-gpu-flush-denormals-to-zero           IEEE denormals
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 (v_pk_*_f32)
  -munsafe-fp-atomics                        hardware float atomics
"""
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT_DIR = os.path.join(PKG, "lib")
OUT = os.path.join(OUT_DIR, "librgp.so")
SOURCES = ["rgp_capi.hip", "rgp_geometry.hip", "rgp_binning.hip", "rgp_raster.hip", "rgp_loss.hip", "rgp_bary.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose=False, jobs=4):
    os.makedirs(OUT_DIR, exist_ok=True)
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    t0 = time.time()
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC] + FLAGS + ["-c", os.path.join(HERE, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            time.sleep(0.05)
    failed = False
    for src, p in procs:
        out = p.communicate()[0].decode(errors="replace")
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"hipcc failed on {src}:\n{out}\n")
        elif out.strip() and verbose:
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("librgp build failed")
    tmp = OUT + ".tmp"
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} in {time.time() - t0:.1f}s")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
