"""Build recipe of the oracle's C++ restatement (test infrastructure only).

    python oracle/build.py      ->  oracle/liboracle.so

The reference's own engine cannot be compiled here (it needs Boost, Rcpp and R
headers that the image does not have), so there is no oracle/_ref build; the
restatement is pinned by the checks under tests/ instead.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "ccd_oracle.cpp")
OUT = os.path.join(HERE, "liboracle.so")


def build(verbose=False):
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", OUT + ".tmp", SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
