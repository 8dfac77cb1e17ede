"""Builds the HIP library ``_build/libbvp_mirk.so`` for gfx950 with hipcc."""
import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "_build"
LIB = OUT / "libbvp_mirk.so"
SOURCES = ["bvp_api.hip", "bvp_host.cpp", "mirk_wave.hip", "wave_m1.hip", "wave_m2.hip", "wave_m2v.hip", "wave_m2_bratu.hip", "wave_m2_singpert.hip", "wave_m3.hip", "wave_m4.hip", "wave_m6_mirk4.hip", "wave_m6_mirk6.hip", "wave_ge_m12.hip", "wave_ge_m34.hip",
           "wave_f32_m2.hip", "wave_f32_m2v.hip", "wave_f32_bratu.hip", "wave_f32_singpert.hip", "wave_f32_m134.hip",
           "block_dispatch.hip", "block_m8.hip", "block_m16.hip", "block_m32.hip", "mlp.hip",
           "firk_dispatch.hip", "firk_m1.hip", "firk_m2.hip", "firk_m3.hip", "firk_m4.hip", "firk_m6a.hip", "firk_m6b.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BVP_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off: some of their forms have given wrong results on this chip
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]


def _compile(src, obj):
    cmd = [HIPCC, *FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return r.stderr


def build(verbose=False):
    OUT.mkdir(exist_ok=True)
    objs = []
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        futs = {}
        for s in SOURCES:
            obj = OUT / (pathlib.Path(s).stem + ".o")
            objs.append(obj)
            futs[s] = ex.submit(_compile, s, obj)
        for s, f in futs.items():
            log = f.result()
            if verbose and log.strip():
                sys.stderr.write(f"[{s}]\n{log}")
    tmp = OUT / "libbvp_mirk.so.tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
