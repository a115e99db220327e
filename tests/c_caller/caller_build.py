"""Compiles tests/c_caller/fpe_caller.c against include/fpe_engine.h and links
it to fpe_option_amd/libfpe_engine.so (output: tests/_build/fpe_caller)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIBDIR = os.path.join(ROOT, "fpe_option_amd")
OUT = os.path.join(ROOT, "tests", "_build", "fpe_caller")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    subprocess.run(["gcc", "-std=c99", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "fpe_caller.c"), "-o", OUT, "-L", LIBDIR, "-lfpe_engine",
                    f"-Wl,-rpath,{LIBDIR}"], check=True)
    return OUT


if __name__ == "__main__":
    print(build())
