"""Build recipe of the oracle (test infrastructure): gt_oracle.c -> liboracle.so.

-ffp-contract=off keeps the FP32 arithmetic free of fused multiply-adds,
as in the HIP build, so both evaluate the shared primitives bit for bit.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(verbose=False):
    src = os.path.join(HERE, "gt_oracle.c")
    out = os.path.join(HERE, "liboracle.so")
    cmd = ["gcc", "-O2", "-std=gnu99", "-fPIC", "-shared", "-ffp-contract=off", "-pthread", "-Wall",
           "-Wno-unused-function", "-o", out + ".tmp", src, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
