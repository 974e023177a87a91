"""Builds the oracle's C restatement (oracle/libcpu_nuts.so) with gcc.

Test / baseline infrastructure only.  x86-64-v3 (AVX2 + FMA) so the library
built here also runs on the GPU box's host CPU."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "libcpu_nuts.so")


def build():
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-ffp-contract=fast", "-fopenmp", "-fPIC",
           "-shared", "-std=gnu11", os.path.join(HERE, "cpu_nuts.c"), "-o", OUT + ".tmp", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("gcc failed building oracle/libcpu_nuts.so:\n" + r.stderr)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build())
