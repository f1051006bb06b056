"""Builds the oracle's C restatement into oracle/liborc.so (test infrastructure).

The reference implementation is Rust; no Rust toolchain exists here, so there
is no oracle/_ref build — the restatement is pinned by the reference's own
hand vectors, closed forms and brute force instead.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build():
    out = os.path.join(HERE, "liborc.so")
    tmp = out + ".tmp"
    cmd = ["gcc", "-O3", "-std=c11", "-fPIC", "-fopenmp", "-shared", "-Wall",
           os.path.join(HERE, "phast_oracle.c"), "-o", tmp]
    print("[build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build()
