"""Build the MI355X library giem2g_amd/_lib/libgiem2g_amd.so with hipcc.

Usage: python -m giem2g_amd.build [-j N] [--verbose]
Sources: giem2g_amd/csrc/*.hip; target gfx950; objects are compiled in
parallel into giem2g_amd/_lib/obj and linked into one shared library.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libgiem2g_amd.so")
ARCH = os.environ.get("GIEM2G_ARCH", "gfx950")

HIPFLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 instructions (v_pk_*_f32) in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(jobs=4, verbose=False, extra=()):
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()

    def one(src):
        obj = os.path.join(objdir, os.path.basename(src)[:-4] + ".o")
        cmd = [cc, *HIPFLAGS, *extra, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-6000:]))
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, sources()))
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-6000:])
    os.replace(tmp, LIB)
    return LIB


FSRC = os.path.join(HERE, "fortran")
HOST = os.path.join(LIBDIR, "giem2g_amd_host")


def fortran_compiler():
    for c in (shutil.which("amdflang"), "/opt/rocm/bin/amdflang", shutil.which("flang")):
        if c and os.path.exists(c):
            return c
    return None


def build_fortran(verbose=False):
    """Fortran driver giem2g_amd/_lib/giem2g_amd_host over the C ABI (ISO_C_BINDING);
    returns its path, or None when no Fortran compiler is installed."""
    fc = fortran_compiler()
    if fc is None:
        return None
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for f in ("g2g_json.f90", "giem2g_amd_api.f90", "giem2g_amd_host.f90"):
        obj = os.path.join(objdir, f[:-4] + ".o")
        cmd = [fc, "-O2", "-module-dir", objdir, "-I", objdir, "-c", os.path.join(FSRC, f), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("Fortran compile failed for %s:\n%s" % (f, r.stderr[-4000:]))
        objs.append(obj)
    cmd = [fc, "-o", HOST, *objs, "-L" + LIBDIR, "-lgiem2g_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("Fortran link failed:\n" + r.stderr[-4000:])
    return HOST


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=4)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    print(build(a.j, a.verbose))
    print(build_fortran(a.verbose))


if __name__ == "__main__":
    sys.exit(main())
