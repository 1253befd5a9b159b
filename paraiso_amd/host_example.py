"""Builds the C++ host example: a program that includes the generated ``Life.hpp``
class and drives the Life machine through the C ABI (``include/paraiso_hip.h``), as the
reference's ``examples/Life/main.cpp`` drives its generated class.  Host code only, so
g++ compiles it; it links the runtime library and loads the machine with dlopen."""
from __future__ import annotations

import os
import subprocess

from . import annotation as A
from .generator import GEN_DIR, INCLUDE_DIR, LIB_DIR, Language, Setup, generate

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCE = os.path.join(ROOT, "tests", "host", "life_main.cpp")
OUT_DIR = os.path.join(GEN_DIR, "host")
SIZE = (256, 192)


def executable() -> str:
    return os.path.join(OUT_DIR, "life_main")


def build() -> str:
    from .schemes.life import life_om
    os.makedirs(OUT_DIR, exist_ok=True)
    setup = Setup(SIZE, Language.HIP, boundary=(A.Condition.Cyclic,) * 2)
    for name, text in generate(setup, life_om()):
        if name.endswith(".hpp"):
            with open(os.path.join(OUT_DIR, name), "w") as f:
                f.write(text)
    exe = executable()
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + OUT_DIR, "-I" + INCLUDE_DIR, SOURCE,
                    "-o", exe, "-L" + LIB_DIR, "-lparaiso_hip", "-ldl",
                    "-Wl,-rpath,$ORIGIN/../../lib"], check=True)
    return exe
