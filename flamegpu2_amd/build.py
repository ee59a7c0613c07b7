"""Build recipe for the gfx950 libraries (in-tree, hipcc only).

    python -m flamegpu2_amd.build            # everything
    python -m flamegpu2_amd.build libfgx     # one target

Outputs go to flamegpu2_amd/lib/.  FP32 is strict IEEE (no FMA contraction,
correctly rounded divide and sqrt, denormals kept) so that agent state can be
compared bit-for-bit with a serial CPU implementation.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "lib", "obj")
INC_C = os.path.join(ROOT, "include")
INC_CPP = os.path.join(PKG, "include")
EXAMPLES = os.path.join(PKG, "examples")
TEST_MODELS = os.path.join(ROOT, "tests", "models")
MODEL_DIRS = [EXAMPLES, TEST_MODELS]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
    f"-I{INC_C}", f"-I{INC_CPP}",
]

# target name -> (sources, link libs)
LIBFGX_SRCS = ["fgx_core.hip", "fgx_prims.hip"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n  " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile(src, obj, extra=()):
    os.makedirs(os.path.dirname(obj), exist_ok=True)
    _run([HIPCC, *HIP_FLAGS, *extra, "-c", src, "-o", obj])
    return obj


def build_libfgx(jobs=8):
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, os.path.join(CSRC, s), os.path.join(OBJ, s.replace(".hip", ".o")))
                for s in LIBFGX_SRCS]
        for f in futs:
            objs.append(f.result())
    out = os.path.join(LIB, "libfgx.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def model_targets():
    """Model libraries: every <dir>/<name>.hip of MODEL_DIRS, built with FGX_MODEL_ABI -> lib/libfgxm_<name>.so."""
    names = []
    for d in MODEL_DIRS:
        if os.path.isdir(d):
            names += [f[:-4] for f in os.listdir(d) if f.endswith(".hip")]
    return sorted(names)


def _model_src(name):
    for d in MODEL_DIRS:
        p = os.path.join(d, name + ".hip")
        if os.path.exists(p):
            return p
    raise FileNotFoundError(name)


def build_model(name):
    src = _model_src(name)
    obj = os.path.join(OBJ, "model_" + name + ".o")
    _compile(src, obj, extra=["-DFGX_MODEL_ABI", "-fvisibility=hidden", "-fvisibility-inlines-hidden"])
    out = os.path.join(LIB, f"libfgxm_{name}.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, obj, f"-L{LIB}", "-lfgx",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_all(jobs=8):
    os.makedirs(LIB, exist_ok=True)
    outs = [build_libfgx(jobs)]
    names = model_targets()
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(names) or 1))) as ex:
        outs += list(ex.map(build_model, names))
    return outs


if __name__ == "__main__":
    targets = sys.argv[1:]
    if not targets:
        for o in build_all():
            print(o)
    else:
        for t in targets:
            print(build_libfgx() if t == "libfgx" else build_model(t))
