"""Builds the oracle's C restatement into ``oracle/lib/libvk_oracle.so`` (test infrastructure).

The reference is a Rust crate and no Rust toolchain exists here, so no ``oracle/_ref``
build of the reference itself is possible; the C restatement is the oracle.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(verbose: bool = False) -> str:
    out_dir = os.path.join(HERE, "lib")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libvk_oracle.so")
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
           "-std=c11", "-Wall", "-Wextra", "-o", out + ".tmp",
           os.path.join(HERE, "vk_oracle.c"), "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    cpu = os.path.join(HERE, "cpu_best.c")
    if os.path.exists(cpu):
        out2 = os.path.join(out_dir, "libvk_cpu_best.so")
        cmd = ["gcc", "-O3", "-fPIC", "-shared", "-march=x86-64-v4", "-mavx512vbmi2",
               "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-std=c11", "-Wall",
               "-o", out2 + ".tmp", cpu, "-lm"]
        if verbose:
            print(" ".join(cmd))
        subprocess.run(cmd, check=True)
        os.replace(out2 + ".tmp", out2)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
