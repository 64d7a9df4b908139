"""Build the CPU oracle (test infrastructure) with g++ -O3 -fopenmp.

Outputs oracle/liboracle.so.  When /root/reference is present, also builds
oracle/_ref/libcsim_ref.so from the reference's own csim sources (see
oracle/ref_csim.mk) to pin the oracle's gate kernels; the GPU box only uses
the prebuilt files.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "liboracle.so")
REF_ROOT = "/root/reference"


def build_oracle():
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
           os.path.join(HERE, "oracle.cpp"), "-o", OUT + ".tmp"]
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


def build_ref():
    if not os.path.isdir(os.path.join(REF_ROOT, "src", "csim")):
        return None
    mk = os.path.join(HERE, "ref_csim.mk")
    if not os.path.exists(mk):
        return None
    r = subprocess.run(["make", "-s", "-f", mk, f"REF={REF_ROOT}", f"OUT={os.path.join(HERE, '_ref')}"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        print("reference csim build failed:\n" + r.stderr[-2000:], file=sys.stderr)
        return None
    return os.path.join(HERE, "_ref", "libcsim_ref.so")


def build():
    o = build_oracle()
    r = build_ref()
    return o, r


if __name__ == "__main__":
    print(build())
