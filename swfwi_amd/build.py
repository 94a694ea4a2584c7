"""Build libswfwi_hip.so in-tree: every HIP/C++ source under csrc/ compiled
for gfx950 with hipcc, objects in swfwi_amd/_build, the library next to this
file.  `python -m swfwi_amd.build` or __graft_entry__.build()."""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libswfwi_hip.so")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("SWFWI_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
    # The reference's arithmetic is the spec: no contraction of a*b+c into
    # FMAs unless a kernel writes fmaf itself.
    "-ffp-contract=off",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of every kernel on this hardware
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [hipcc()] + COMMON + ["-I", os.path.join(HERE, "..", "include"), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 2) // 2), 8)
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, srcs))
    objs = [o for o, _ in results]
    if verbose:
        for (o, err) in results:
            if err.strip():
                print(err, file=sys.stderr)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    build_apps(jobs)
    return OUT


APPS = {  # binary -> (source, extra defines)
    "fm-damp": ("fm_damp.cpp", []),
    "essfwi-damp": ("fwi_damp.cpp", []),
    "fwi-damp": ("fwi_damp.cpp", ["-DFWI_PER_SHOT"]),
    "enfwi-damp": ("enfwi_damp.cpp", []),
}


def _build_app(item):
    name, (src, defs) = item
    bindir = os.path.join(HERE, "bin")
    out = os.path.join(bindir, name)
    cmd = [hipcc(), "-O2", "-std=c++17", "-ffp-contract=off", f"--offload-arch={ARCH}"] + defs + [
        os.path.join(HERE, "apps", src), "-o", out + ".tmp", "-L", HERE, "-lswfwi_hip",
        "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"app build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(out + ".tmp", out)
    return out


def build_apps(jobs=None):
    """The reference's command-line tools over the library (swfwi_amd/bin)."""
    os.makedirs(os.path.join(HERE, "bin"), exist_ok=True)
    with cf.ThreadPoolExecutor(jobs or 3) as ex:
        return list(ex.map(_build_app, APPS.items()))


if __name__ == "__main__":
    print(build(verbose=True))
