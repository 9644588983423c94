"""The phase-factor sin / cos of the diagonal runs (pyquafu_amd/csrc/sincos.h), compiled for the
host with g++ and checked against libm (numpy) over the angle range the runs produce."""

import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = r"""
#include <cstdio>
#include <cstdlib>
#include "sincos.h"
int main(int argc, char** argv) {
    FILE* f = fopen(argv[1], "rb");
    FILE* g = fopen(argv[2], "wb");
    double x;
    while (fread(&x, sizeof x, 1, f) == 1) {
        double s, c;
        float sf, cf;
        qfvm::sin_cos(x, &s, &c);
        qfvm::sin_cos((float)x, &sf, &cf);
        double out[4] = {s, c, (double)sf, (double)cf};
        fwrite(out, sizeof out, 1, g);
    }
    fclose(f);
    fclose(g);
    return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("sincos")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    exe = d / "h"
    r = subprocess.run(["g++", "-O2", "-ffp-contract=off", "-I", os.path.join(ROOT, "pyquafu_amd", "csrc"),
                        str(src), "-o", str(exe)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("g++ unavailable: " + r.stderr[:200])
    return d, exe


def run(harness, x):
    d, exe = harness
    inp, out = d / "x.bin", d / "y.bin"
    x.astype(np.float64).tofile(inp)
    subprocess.run([str(exe), str(inp), str(out)], check=True)
    return np.fromfile(out, dtype=np.float64).reshape(-1, 4)


def test_double_against_libm(harness):
    rng = np.random.Generator(np.random.PCG64(2001))
    x = np.concatenate([rng.uniform(-1000, 1000, 200000), rng.uniform(-4, 4, 100000),
                        np.arange(-64, 65) * (np.pi / 4), [0.0, -0.0, 1e-300, np.pi, -np.pi]])
    y = run(harness, x)
    assert np.abs(y[:, 0] - np.sin(x)).max() < 4e-16
    assert np.abs(y[:, 1] - np.cos(x)).max() < 4e-16
    # exact quadrant points
    assert y[-5, 0] == 0.0 and y[-5, 1] == 1.0


def test_float_against_libm(harness):
    rng = np.random.Generator(np.random.PCG64(2002))
    x = rng.uniform(-3000, 3000, 100000).astype(np.float32).astype(np.float64)
    y = run(harness, x)
    assert np.abs(y[:, 2] - np.sin(x)).max() < 6e-7
    assert np.abs(y[:, 3] - np.cos(x)).max() < 6e-7
