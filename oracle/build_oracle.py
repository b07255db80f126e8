"""Build the CPU oracle (test infrastructure): liboracle_<model>.so, one per model directory.

The oracle is compiled with g++ and -ffp-contract=off so that its rounding matches the
device build of the same model headers.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
MODELS_DIR = os.path.join(ROOT, "zoomy-dmplex_amd", "models")
MODELS = ["sme0_1d", "sme0", "sme0_visc", "sme1", "sme1m", "sme2", "vam"]


TANGENT_MODELS = ["sme1m", "vam"]


def build(model, tangent=False):
    """liboracle_<model>.so, or with tangent=True liboracle_<model>_tl.so: the same source on
    one-direction dual numbers (the tangent-linear reference run)."""
    out = os.path.join(HERE, "liboracle_%s%s.so" % (model, "_tl" if tangent else ""))
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(MODELS_DIR, model),
           "-I" + MODELS_DIR, "-I" + HERE] + (["-DORACLE_TANGENT"] if tangent else []) + \
          [os.path.join(HERE, "oracle.cpp"), "-o", out + ".tmp"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed for %s:\n%s" % (model, r.stdout[-4000:]))
    os.replace(out + ".tmp", out)
    return out


def build_all(models=None):
    models = models or MODELS
    return [build(m) for m in models] + [build(m, True) for m in TANGENT_MODELS if m in models]


if __name__ == "__main__":
    for o in build_all(sys.argv[1:] or None):
        print("built", o)
