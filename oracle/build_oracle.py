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
MODELS = ["sme0_1d", "sme0", "sme1", "sme2", "vam"]


def build(model):
    out = os.path.join(HERE, "liboracle_%s.so" % model)
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(MODELS_DIR, model),
           "-I" + HERE, os.path.join(HERE, "oracle.cpp"), "-o", out + ".tmp"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed for %s:\n%s" % (model, r.stdout[-4000:]))
    os.replace(out + ".tmp", out)
    return out


def build_all(models=None):
    return [build(m) for m in (models or MODELS)]


if __name__ == "__main__":
    for o in build_all(sys.argv[1:] or None):
        print("built", o)
