"""Build the in-tree shared libraries of the package.

  libzdm_mesh.so          host mesh service and HDF5 initial-condition files (g++)
  libzdm_<model>.so       HIP time step for gfx950, one per model directory under models/

Models are compiled exactly as the reference compiles its solver on the generated headers:
the model directory (Model.H + Numerics.H [+ ChorinOps.H]) is put on the include path.
Device code is compiled with -ffp-contract=off so that it rounds like the host.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
MODELS = ["sme0_1d", "sme0", "sme0_visc", "sme1", "sme1m", "sme2", "vam"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), r.stdout[-4000:]))
    return r.stdout


def build_mesh():
    out = os.path.join(HERE, "libzdm_mesh.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-I" + INCLUDE,
          os.path.join(HERE, "csrc", "mesh.cpp"), os.path.join(HERE, "csrc", "h5io.cpp"), "-ldl", "-o", out])
    return out


def build_model(model, extra_src=None):
    mdir = os.path.join(HERE, "models", model)
    out = os.path.join(HERE, "libzdm_%s.so" % model)
    src = [os.path.join(HERE, "csrc", "zdm_solver.hip")]
    if extra_src:
        src += extra_src
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-I" + INCLUDE, "-I" + mdir, "-I" + os.path.join(HERE, "csrc"),
                                "-I" + os.path.join(HERE, "models"),
                                '-DZDM_MODEL_NAME="%s"' % model] + src + ["-o", tmp])
    os.replace(tmp, out)
    return out


def build_cli(model):
    """zdm_cli_<model>: the C++ host (include/zoomy/) over libzdm_<model>.so, found next to it."""
    out = os.path.join(HERE, "zdm_cli_%s" % model)
    _run(["g++", "-O2", "-std=c++17", "-I" + INCLUDE, os.path.join(HERE, "host", "main.cpp"),
          "-L" + HERE, "-l:libzdm_%s.so" % model, "-l:libzdm_mesh.so", "-Wl,-rpath,$ORIGIN",
          "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_all(models=None, jobs=4):
    models = models or MODELS
    outs = [build_mesh()]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        outs += list(ex.map(build_model, models))
    outs += [build_cli(m) for m in models]
    return outs


if __name__ == "__main__":
    for o in build_all(sys.argv[1:] or None):
        print("built", o)
