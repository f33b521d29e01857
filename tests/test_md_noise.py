"""The Brownian-loop noise generator (gdt_rng.h), compiled for the host with
g++, must give bit-identical normals and wrapped displacements to the numpy
restatement in gpudelaunaytriangulation_amd/inputs.py; the numpy side is
pinned by its moments and by the polar method's acceptance rate."""
import os
import subprocess

import numpy as np
import pytest

from gpudelaunaytriangulation_amd import inputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpu_emul", "rng.cpp")
CSRC = os.path.join(ROOT, "gpudelaunaytriangulation_amd", "csrc")


@pytest.fixture(scope="module")
def rng_exe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("rng") / "rng")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", f"-I{CSRC}", SRC, "-o", exe],
                   check=True)
    return exe


@pytest.mark.parametrize("seed,step,set_index", [(1, 0, 0), (12345, 999, 255), (2**63 + 5, 7, 3)])
def test_host_build_matches_numpy_bitwise(rng_exe, tmp_path, seed, step, set_index):
    n, L, sigma = 5000, 70.7, 0.01
    xy = inputs.uniform_points(n, L, L, seed=3)
    pin, pout = str(tmp_path / "p.bin"), str(tmp_path / "o.bin")
    xy.tofile(pin)
    subprocess.run([rng_exe, str(seed), str(step), str(set_index), str(n), repr(sigma), repr(L),
                    pin, pout], check=True)
    raw = np.fromfile(pout, dtype=np.float64)
    g_c, xy_c = raw[:2 * n].reshape(n, 2), raw[2 * n:].reshape(n, 2)
    g_np = inputs.md_noise(seed, step, set_index, n)
    assert np.array_equal(g_c.view(np.uint64), g_np.view(np.uint64))
    xy_np = inputs.md_step(xy, L, L, sigma, seed, step, set_index)
    assert np.array_equal(xy_c.view(np.uint64), xy_np.view(np.uint64))
    assert np.all((xy_np >= 0) & (xy_np < L))


def test_noise_moments_and_streams():
    g = inputs.md_noise(7, 3, 5, 1000000)
    # 2e6 samples: standard errors 7e-4 (mean), 5e-4 (std), 7e-3 (fourth
    # moment, which is 3 for a standard normal); tolerances about 6 sigma
    assert abs(g.mean()) < 0.005 and abs(g.std() - 1.0) < 0.003
    assert abs((g ** 4).mean() - 3.0) < 0.04
    # different steps and sets give independent streams
    h = inputs.md_noise(7, 4, 5, 1000000)
    k = inputs.md_noise(7, 3, 6, 1000000)
    assert abs(np.corrcoef(g[:, 0], h[:, 0])[0, 1]) < 0.01
    assert abs(np.corrcoef(g[:, 0], k[:, 0])[0, 1]) < 0.01
