"""Conserved noise of euler_mobility: the oracle's generator and noisy step.

The reference draws the noise from std::mt19937 (EulerMobilityCPU.cpp:17-18,
65-68) or per-cell curand states (EulerMobilityCUDA.cu:52-71); neither stream
is reproducible across implementations, so both the oracle and the HIP path
use Philox4x32-10 keyed by the seed with the counter (cell, step).  Pinned here:
  * Philox4x32-10 against the published known-answer vectors of Salmon et al.
    (Random123 kat_vectors, philox4x32 with 10 rounds);
  * the normals: Kolmogorov-Smirnov against N(0, 1), first two moments;
  * the step: from a uniform field (mu uniform, so only noise moves mass) the
    increment has variance 4 d M dt r^2 / dx^(d+2) (r = noise_rescale_factor)
    and neighbour correlation -1/(2d) along each axis (the two cells share one
    face flux with opposite signs); total mass is unchanged; a negative face
    mobility gives no noise (sqrt(max(M_face, 0)), EulerMobilityCPU.cpp:65);
    r = 0 reproduces the noiseless step bit for bit.
"""
import numpy as np
import pytest
from scipy import stats

import oracle as O


@pytest.mark.parametrize("ctr,key,expected", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_known_answers(ctr, key, expected):
    assert tuple(int(v) for v in O.philox4x32_10(ctr, key)) == expected


def test_cell_normals_are_standard_normal():
    z = np.array([O.cell_normals(987654321, c, 17)[:4] for c in range(25000)]).ravel()
    assert abs(z.mean()) < 0.02
    assert abs(z.var() - 1.0) < 0.03
    assert stats.kstest(z, "norm").pvalue > 1e-3
    # the step index and the seed both change the draw
    assert not np.array_equal(O.cell_normals(1, 5, 0), O.cell_normals(1, 5, 1))
    assert not np.array_equal(O.cell_normals(1, 5, 0), O.cell_normals(2, 5, 0))


@pytest.mark.parametrize("dims,N", [(1, 65536), (2, 256), (3, 32)])
def test_one_step_variance_and_neighbour_correlation(dims, N):
    M, dt, dx, r = 1.3, 0.01, 0.7, 1.7
    p = O.Params(dims=dims, N=N, dt=dt, dx=dx, k=1.0, mobility=O.MOB_CONSTANT, M=M)
    m = O.Model(O.LANDAU, epsilon=0.9)
    rho0 = np.full(N ** dims, 0.3)
    d = O.run_noise(p, m, rho0, 1, noise_rescale=r, seed=2024) - rho0
    expected = 4.0 * dims * M * dt * r * r / dx ** (dims + 2)
    assert abs(d.var() / expected - 1.0) < 0.05
    grid = d.reshape((N,) * dims)
    for ax in range(dims):
        corr = np.mean(grid * np.roll(grid, 1, axis=ax)) / np.mean(grid * grid)
        assert abs(corr + 1.0 / (2 * dims)) < 0.03
    assert abs(d.sum()) < 1e-12 * N ** dims * np.sqrt(expected)


def test_mass_conserved_and_rescale_zero_is_noiseless():
    p = O.Params(dims=2, N=32, dt=1e-4, dx=1.0, k=1.0, mobility=O.MOB_LINEAR, M=1.0, rho_ref=0.05)
    m = O.Model(O.SIMPLE_WERTHEIM, B2=1.0, delta=1e4, valence=4)
    rng = np.random.default_rng(3)
    rho0 = 0.05 * (1.0 + 0.05 * (2 * rng.random(32 * 32) - 1))
    noisy = O.run_noise(p, m, rho0, 50, noise_rescale=0.01, seed=11)
    assert abs(noisy.sum() - rho0.sum()) < 1e-12 * rho0.sum()
    assert np.abs(noisy - O.run(O.EULER_MOBILITY, p, m, rho0, 50)).max() > 1e-8
    assert np.array_equal(O.run_noise(p, m, rho0, 50, noise_rescale=0.0, seed=11),
                          O.run(O.EULER_MOBILITY, p, m, rho0, 50))


def test_negative_face_mobility_has_no_noise():
    # linear mobility M rho / rho_ref < 0 for rho < 0: sqrt(max(M_face, 0)) = 0
    p = O.Params(dims=2, N=16, dt=0.01, dx=1.0, k=1.0, mobility=O.MOB_LINEAR, M=1.0, rho_ref=1.0)
    m = O.Model(O.LANDAU, epsilon=0.9)
    rho0 = np.full(256, -0.2)
    assert np.array_equal(O.run_noise(p, m, rho0, 3, noise_rescale=1.0, seed=5), rho0)
    rho0 = np.full(256, 0.2)
    assert np.abs(O.run_noise(p, m, rho0, 1, noise_rescale=1.0, seed=5) - rho0).max() > 1e-3
