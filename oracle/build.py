"""Build the CPU oracle (test infrastructure only) with gcc.

Output: oracle/liboracle.so.  -ffp-contract=off and no fast-math keep every
FP64/FP32 operation separately rounded, matching the HIP kernels' order.

The reference's own native path (src/flux/cgal/aabb_wrapper.cpp) needs CGAL
and Boost headers that this image does not have, so there is no
oracle/_ref build; see the oracle header comment for what pins each part.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build(verbose=False):
    srcs = [os.path.join(HERE, f) for f in ("flux_oracle.c", "pcc_oracle.c")]
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=c99", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off",
           "-fno-fast-math", "-Wall", "-Wno-unused-function", *srcs, "-o", tmp, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip() and verbose:
        print(r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
