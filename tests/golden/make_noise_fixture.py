"""Writes tests/golden/noise_fixture.json: deviates of the compiled C++ standard library's
std::mt19937_64 + std::normal_distribution<double> (GCC libstdc++ of this image), seeded as
the reference's deterministic gaussian_noise_r seeds a cell (cell_index * 1023), for the
restatement in oracle/lattice_noise.py to be checked against bitwise.

Run from the repository root: python tests/golden/make_noise_fixture.py (needs g++)."""
import json
import os
import subprocess
import tempfile

CELLS = [0, 1, 2, 37, 63]
DRAWS = 24
SIGMA = 0.05

SRC = r"""
#include <cstdio>
#include <random>
int main() {
  const unsigned long long cells[] = {%s};
  for (unsigned long long c : cells) {
    std::mt19937_64 re;
    re.seed(c * 1023);
    std::normal_distribution<double> g(0.0, %r);
    for (int i = 0; i < %d; i++) std::printf("%%a\n", g(re));
  }
  std::mt19937_64 d;
  d.discard(9999);
  std::printf("%%llu\n", (unsigned long long)d());
}
"""


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    with tempfile.TemporaryDirectory() as t:
        src = os.path.join(t, "n.cpp")
        exe = os.path.join(t, "n")
        open(src, "w").write(SRC % (", ".join(str(c) for c in CELLS), SIGMA, DRAWS))
        subprocess.run(["g++", "-O2", "-o", exe, src], check=True)
        lines = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
        ver = subprocess.run(["g++", "--version"], check=True, capture_output=True, text=True).stdout.splitlines()[0]
    dev = lines[:-1]
    out = {"generator": "std::mt19937_64 seeded cell * 1023, std::normal_distribution<double>(0, %g)" % SIGMA,
           "compiler": ver, "sigma": SIGMA, "mt19937_64_10000th_default_seed": int(lines[-1]),
           "cells": {str(c): dev[q * DRAWS:(q + 1) * DRAWS] for q, c in enumerate(CELLS)}}
    json.dump(out, open(os.path.join(here, "noise_fixture.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
