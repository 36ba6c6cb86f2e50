"""Build libplasma_amd.so in-tree: HIP kernels for gfx950 with hipcc, the C
host library with gcc, linked into plasma_amd/lib/libplasma_amd.so.

Run ``python -m plasma_amd.build`` (or ``__graft_entry__.build()``).  Every
call recompiles all sources; objects go to plasma_amd/_build/.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libplasma_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("PLASMA_OFFLOAD_ARCH", "gfx950")

INCLUDES = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
            "-I" + os.path.join(CSRC, "kernels")]
HIPFLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
            # packed FP32 instructions stay disabled (see the kernels' notes)
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument"]
CFLAGS = ["-O2", "-fPIC", "-std=c11", "-Wall", "-Wno-unused-function",
          "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include")]


def _sources():
    hip, c = [], []
    for d, _, files in os.walk(CSRC):
        for f in sorted(files):
            p = os.path.join(d, f)
            if f.endswith(".hip"):
                hip.append(p)
            elif f.endswith(".c"):
                c.append(p)
    return sorted(hip), sorted(c)


def _obj(src):
    rel = os.path.relpath(src, CSRC).replace(os.sep, "_")
    return os.path.join(OBJ, rel + ".o")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    hip, c = _sources()
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    jobs = []
    for s in hip:
        jobs.append([hipcc] + HIPFLAGS + INCLUDES + ["-c", s, "-o", _obj(s)])
    for s in c:
        jobs.append(["gcc"] + CFLAGS + INCLUDES + ["-c", s, "-o", _obj(s)])
    nproc = min(16, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(nproc) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    objs = [_obj(s) for s in hip + c]
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp]
         + objs + ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-ldl",
                   "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
