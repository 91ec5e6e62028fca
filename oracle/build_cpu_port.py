"""Builds oracle/build/libngt_cpu_port.so from oracle/cpu_port.cpp (the
FP32 OpenMP CPU restatement used as checker and CPU baseline; test
infrastructure only).  ``-march=x86-64-v3`` (AVX2 + FMA) instead of
``native`` so that the library built here also runs on other x86 hosts."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_port.cpp")
OUT_DIR = os.path.join(HERE, "build")
LIB = os.path.join(OUT_DIR, "libngt_cpu_port.so")
FLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c++17"]


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = ["g++"] + FLAGS + [SRC, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
