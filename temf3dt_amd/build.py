"""Build libtemfetd.so (hand-written gfx950 HIP kernels + C++ host runtime).

Usage: python -m temf3dt_amd.build [--jobs N]

Compiles every source under temf3dt_amd/csrc with hipcc for gfx950 into
object files under temf3dt_amd/_build and links temf3dt_amd/libtemfetd.so
in-tree (the library travels with the source tree; nothing is installed).
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtemfetd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TF_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 forms are kept out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", os.path.join(HERE, "..", "include"),
]
HOST_ONLY = ["-fopenmp"]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name == "host_probe.cpp":
            continue
        if name.endswith(".hip") or name.endswith(".cpp"):
            out.append(os.path.join(CSRC, name))
    return out


def compile_one(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd[1:1] = HOST_ONLY
    else:
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(jobs=8, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(compile_one, srcs))
    objs = [o for o, _ in results]
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fopenmp", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_host_probe()
    build_fortran()
    return LIB


HOST_LIB = os.path.join(HERE, "libtemfetd_host.so")


def build_host_probe():
    """Host-only library over topology.cpp + symbolic.cpp (no GPU needed),
    used to check and size the symbolic analysis on the CPU."""
    srcs = [os.path.join(CSRC, n) for n in ("host_probe.cpp", "topology.cpp", "symbolic.cpp")]
    tmp = HOST_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-o", tmp] + srcs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host probe build failed:\n{r.stderr}")
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")
FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(HERE, "fetd_amd")
FIO_CHECK = os.path.join(HERE, "fetd_io_check")


def build_fortran():
    """Fortran driver `fetd_amd -f ROOT` over the C ABI (ISO_C_BINDING), and
    `fetd_io_check`, the Fortran reader/writer alone (no GPU), for tests.
    Skipped (returns None) when no Fortran compiler is installed."""
    if not os.path.exists(FLANG):
        return None
    mods = os.path.join(BUILD, "fmod")
    os.makedirs(mods, exist_ok=True)
    objs = {}
    for name in ("fetd_io", "temfetd", "fetd_main", "fetd_io_check"):
        obj = os.path.join(BUILD, name + ".f.o")
        cmd = [FLANG, "-O2", "-J", mods, "-c", os.path.join(FORTRAN, name + ".f90"), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed for {name}:\n{r.stdout}\n{r.stderr}")
        objs[name] = obj
    for exe, parts, libs in ((DRIVER, ("fetd_main", "fetd_io", "temfetd"), ["-L", HERE, "-ltemfetd",
                                                                          "-Wl,-rpath,$ORIGIN"]),
                             (FIO_CHECK, ("fetd_io_check", "fetd_io"), [])):
        cmd = [FLANG, "-o", exe + ".tmp"] + [objs[p] for p in parts] + libs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link of {exe} failed:\n{r.stdout}\n{r.stderr}")
        os.replace(exe + ".tmp", exe)
    return DRIVER


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.v))


if __name__ == "__main__":
    main()
