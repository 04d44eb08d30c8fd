"""Compiles oracle/nblist_ref.c into oracle/_build/liboracle.so (gcc, FP32 exact)."""

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent


def build() -> Path:
    out = HERE / "_build"
    out.mkdir(exist_ok=True)
    lib = out / "liboracle.so"
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC", str(HERE / "nblist_ref.c"), "-o", str(lib), "-lm"]
    subprocess.run(cmd, check=True)
    return lib


if __name__ == "__main__":
    print(build())
