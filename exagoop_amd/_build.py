"""Build recipe for libexagoop_amd.so (HIP, gfx950).

Compiles the HIP sources under csrc/ with hipcc into an in-tree shared
library.  FP contraction is off so the kernels follow the reference's
arithmetic order; packed FP32 ops are disabled for all HIP code.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libexagoop_amd.so")
SOURCES = ["mpm_kernels.hip", "mpm_sort.hip", "mpm_partile.hip", "mpm_ctx.hip", "mpm_mlp.hip",
           "mpm_learn.hip", "mpm_batch.hip", "mpm_thermal.hip", "mpm_noalloc.hip", "mpm_io.cpp",
           "mpm_geometry.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    for cand in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.sep not in cand or os.path.exists(cand):
            return cand
    return "hipcc"


RUN = os.path.join(HERE, "mpm_run")


def build_driver():
    """The command-line time-loop driver (csrc/mpm_run.cpp) over the C++
    facade, linked against the in-tree library it sits beside."""
    tmp = RUN + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(HERE, "..", "include"),
           "-I" + CSRC, os.path.join(CSRC, "mpm_run.cpp"), "-o", tmp, "-L" + HERE,
           "-lexagoop_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed on mpm_run.cpp:\n" + r.stdout + r.stderr)
    os.replace(tmp, RUN)
    return RUN


def build(verbose=False, jobs=None):
    """Compile libexagoop_amd.so in place; raises on failure.  Each source is
    compiled to an object in parallel (the kernels of one source call no
    device code of another), then the objects are linked."""
    import concurrent.futures
    import shutil
    import tempfile

    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmpdir = tempfile.mkdtemp(prefix="exagoop_amd_build_")
    try:
        objs = [os.path.join(tmpdir, os.path.basename(s) + ".o") for s in srcs]
        cflags = [f for f in HIPCC_FLAGS if f != "-shared"]

        def compile_one(pair):
            src, obj = pair
            r = subprocess.run([hipcc()] + cflags + ["-c", src, "-o", obj],
                               capture_output=True, text=True)
            return src, r

        n = jobs or min(len(srcs), os.cpu_count() or 1)
        with concurrent.futures.ThreadPoolExecutor(max_workers=n) as ex:
            for src, r in ex.map(compile_one, zip(srcs, objs)):
                if r.returncode != 0:
                    raise RuntimeError("hipcc failed on %s:\n%s%s" % (src, r.stdout, r.stderr))
                if verbose and r.stderr:
                    print(r.stderr)
        tmp = LIB + ".tmp"
        r = subprocess.run([hipcc()] + HIPCC_FLAGS + objs + ["-o", tmp], capture_output=True,
                           text=True)
        if r.returncode != 0:
            raise RuntimeError("hipcc link failed:\n" + r.stdout + r.stderr)
        os.replace(tmp, LIB)
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    build_driver()
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
