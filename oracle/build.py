"""Build recipe for the oracle: compiles oracle.cpp into _build/liboracle.so.

The reference's own knn_rms cannot be compiled here (it needs GROMACS,
Boost.program_options, Berkeley DB and ARPACK, none of which are present), so
there is no oracle/_ref build; the oracle is the C++ restatement alone.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(verbose=False):
    out_dir = os.path.join(HERE, "_build")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "liboracle.so")
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-fopenmp", "-shared", "-fPIC", "-std=c++17",
           os.path.join(HERE, "oracle.cpp"), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
