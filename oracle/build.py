"""Build the CPU oracle (test infrastructure only) with gcc: oracle/liboracle.so."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build() -> str:
    src = os.path.join(HERE, "rsr_oracle.c")
    tmp = LIB + ".tmp"
    # binary32 arithmetic exactly as written: no FMA contraction, no fast math, no
    # reassociation; the 8-lane loops may be vectorised (AVX2; each lane's operations stay
    # the same IEEE operations), as the reference's f32x8 evaluator is SIMD
    subprocess.run(["gcc", "-O3", "-mavx2", "-std=gnu11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                    "-fexcess-precision=standard", "-mno-fma", "-o", tmp, src, "-lm"], check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
