"""Build the CPU oracle (test infrastructure) into oracle/build/.

    python oracle/build_oracle.py

libipm_oracle.so   -O2 -ffp-contract=off, the checker used by tests/ and smoke()
libipm_oracle_fast.so  -O3 -march=x86-64-v2 -ffp-contract=off, timed by bench.py's
                   cpu_baseline leg (same source, same rounding: contraction off)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "build")
SRC = os.path.join(HERE, "ipm_oracle.c")


def build():
    os.makedirs(OUT, exist_ok=True)
    common = ["gcc", "-shared", "-fPIC", "-std=c99", "-ffp-contract=off", "-fno-fast-math", "-Wall"]
    for name, opt in (("libipm_oracle.so", ["-O2"]), ("libipm_oracle_fast.so", ["-O3", "-march=x86-64-v2"])):
        tmp = os.path.join(OUT, name + ".tmp")
        cmd = common + opt + [SRC, "-o", tmp, "-lm"]
        print("+", " ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, os.path.join(OUT, name))


if __name__ == "__main__":
    sys.exit(build())
