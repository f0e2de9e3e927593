"""Independent pins of the oracle's density-error statistic
(oracle/pbf_oracle.cpp density_impl; rho as in computeLambda,
fluids/Simulator_kernel.cuh:77-120, plus the wall term of DensityBoundary,
fluids/Simulator.cu:139-161).

Each check computes the expected per-particle density without the oracle's
grid or kernels: from lattice shell counts of the textbook poly6 kernel, from
an O(N^2) brute-force sum in numpy, and from the closed-form wall density.
A dropped self term fails the lattice test, a dropped or misplaced wall term
fails the wall test, and a density queried from the wrong position (a wrong
permutation, a neighbour's or a stale coordinate) fails the brute-force test;
every test also checks that mean and max aggregate the per-particle values.
CPU only."""
import math

import numpy as np
import pytest

from pbf import params as P

PI_REF = 3.14159265359  # fluids/helper.h:4


def poly6_ref(h, r2):
    """315/(64 pi h^9) (h^2 - r^2)^3 for r^2 < h^2, with the reference's pi
    (h is the parameter block's float, widened)."""
    h = float(np.float32(h))
    r2 = np.asarray(r2, np.float64)
    c = 315.0 / (64.0 * PI_REF * h**9)
    return np.where(r2 < h * h, c * (h * h - r2) ** 3, 0.0)


def wall_ref(h, d):
    """Wall density per face: 0 beyond h, 2 pi/3 at or behind the wall,
    (2 pi/3)(h - d)^2 (h + d) between."""
    h = float(np.float32(h))
    d = np.asarray(d, np.float64)
    mid = (2.0 * PI_REF / 3.0) * (h - d) ** 2 * (h + d)
    return np.where(d > h, 0.0, np.where(d <= 0.0, 2.0 * PI_REF / 3.0, mid))


def _check_aggregates(mean, mx, err):
    assert mean == pytest.approx(float(err.mean()), rel=1e-12, abs=1e-15)
    assert mx == pytest.approx(float(err.max()), rel=1e-12, abs=1e-15)


def test_unjittered_lattice_shell_counts(orc):
    """A 0.5h lattice in shuffled input order.  Every particle's density is
    the sum over lattice offsets d in {-1,0,1}^3 (self included) that stay
    inside the lattice of W(|d| 0.5h); offsets of length 2 and more lie at
    r >= h.  Interior particles give W(0)(1 + 6*0.421875 + 12*0.125 +
    8*0.015625) = 8078.20 at h = 0.1."""
    p = P.dam_params()
    h, s = p.h, 0.5 * p.h
    L = 10
    idx = np.stack(np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij"),
                   -1).reshape(-1, 3)
    rng = np.random.default_rng(11)
    perm = rng.permutation(idx.shape[0])
    idx = idx[perm]
    lo = np.array([0.3, 0.3, 0.3])
    x = (lo + (idx + 0.5) * s).astype(np.float32)
    ulim, llim = (1.6, 1.6, 1.6), (0.0, 0.0, 0.0)

    expected = np.zeros(idx.shape[0])
    for d in np.stack(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1], indexing="ij"),
                      -1).reshape(-1, 3):
        inside = np.all((idx + d >= 0) & (idx + d < L), axis=1)
        expected += inside * float(poly6_ref(h, (s * s) * float(d @ d)))
    interior = np.all((idx >= 1) & (idx <= L - 2), axis=1)
    w0 = float(poly6_ref(h, 0.0))
    assert w0 * (1 + 6 * 0.421875 + 12 * 0.125 + 8 * 0.015625) == pytest.approx(8078.20, abs=0.01)
    np.testing.assert_allclose(expected[interior], 8078.20, atol=0.01)
    want = np.abs(expected / p.pho0 - 1.0)

    # The FP32 coordinates sit up to half an ulp off the ideal lattice, which
    # moves rho by ~0.01; the smallest shell term, W(sqrt(3) h/2) = 24.5,
    # is over 500 times the tolerance, so a missing or extra neighbour fails.
    tol = 5e-6  # in |rho/pho0 - 1|, i.e. 0.03 in rho
    assert float(poly6_ref(h, 0.75 * h * h)) / p.pho0 > 500 * tol
    m64, x64, e64 = orc.density_error(p, ulim, llim, x, per_particle=True)
    np.testing.assert_allclose(e64, want, rtol=0, atol=tol)
    _check_aggregates(m64, x64, e64)
    m32, x32, e32 = orc.density_error_f32(p, ulim, llim, x)
    np.testing.assert_allclose(e32, want, rtol=0, atol=tol)
    _check_aggregates(m32, x32, e32)
    # interior: the closed-form value
    np.testing.assert_allclose(e64[interior], abs(8078.20 / p.pho0 - 1.0), atol=tol)


@pytest.mark.parametrize("k_b", [0.0, 0.5])
def test_brute_force_density_on_jittered_particles(orc, k_b):
    """3,000 particles: a jittered lattice block against two walls plus a
    scattered cloud, some outside the box.  Expected rho_i = sum_j W(|x_i -
    x_j|) over ALL j (self included) by brute force, plus k_b times the six
    faces' wall density at x_i."""
    p = P.dam_params().replace(k_boundaryDensity=k_b)
    h = p.h
    rng = np.random.default_rng(5)
    ulim, llim = (1.2, 0.9, 1.0), (0.0, 0.0, 0.0)
    g = np.stack(np.meshgrid(np.arange(14), np.arange(12), np.arange(14), indexing="ij"),
                 -1).reshape(-1, 3)
    block = (g + 0.5) * 0.05 + rng.uniform(-0.01, 0.01, g.shape)
    cloud = rng.uniform([-0.05, -0.05, -0.05], [1.25, 0.95, 1.05], (3000 - block.shape[0], 3))
    x = np.concatenate([block, cloud]).astype(np.float32)
    x = x[rng.permutation(x.shape[0])]
    X = x.astype(np.float64)
    d = X[:, None, :] - X[None, :, :]
    rho = poly6_ref(h, np.einsum("ijk,ijk->ij", d, d)).sum(1)
    u, l = np.array(ulim, np.float32).astype(np.float64), np.array(llim, np.float64)
    walls = (wall_ref(h, u[None, :] - X) + wall_ref(h, X - l[None, :])).sum(1)
    rho = rho + k_b * walls
    want = np.abs(rho / p.pho0 - 1.0)

    m64, x64, e64 = orc.density_error(p, ulim, llim, x, per_particle=True)
    np.testing.assert_allclose(e64, want, rtol=1e-9, atol=1e-12)
    _check_aggregates(m64, x64, e64)
    m32, x32, e32 = orc.density_error_f32(p, ulim, llim, x)
    np.testing.assert_allclose(e32, want, rtol=1e-4, atol=5e-6)
    _check_aggregates(m32, x32, e32)


def test_wall_term_at_known_distances(orc):
    """Isolated particles (more than h apart) at set distances from the
    floor and the x walls: rho = W(0) + k_b * (wall density of every face
    within h), so |rho/pho0 - 1| follows from the closed form alone."""
    p = P.dam_params().replace(k_boundaryDensity=0.5)
    h = p.h
    ulim, llim = (3.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    zs = [-0.02, 0.0, 0.01, 0.03, 0.05, 0.08, 0.099, 0.15, 0.5]
    pts = [(0.3 + 0.3 * k, 0.5, z) for k, z in enumerate(zs)]
    pts += [(0.04, 0.2, 0.5), (2.97, 0.2, 0.5), (0.0, 0.8, 0.5), (3.02, 0.8, 0.5),
            (0.06, 0.05, 0.5)]  # two walls at once
    x = np.array(pts, np.float32)
    X = x.astype(np.float64)
    gap = np.linalg.norm(X[:, None] - X[None, :], axis=-1) + 10 * np.eye(len(pts))
    assert gap.min() > 1.2 * h  # isolated: no pair term
    u, l = np.array(ulim, np.float32).astype(np.float64), np.array(llim, np.float64)
    walls = (wall_ref(h, u[None, :] - X) + wall_ref(h, X - l[None, :])).sum(1)
    assert walls[0] == pytest.approx(2 * PI_REF / 3) and walls[7] == 0.0
    assert walls[3] == pytest.approx((2 * PI_REF / 3) * 0.07**2 * 0.13, rel=1e-6)
    rho = float(poly6_ref(h, 0.0)) + p.k_boundaryDensity * walls
    want = np.abs(rho / p.pho0 - 1.0)
    m64, x64, e64 = orc.density_error(p, ulim, llim, x, per_particle=True)
    np.testing.assert_allclose(e64, want, rtol=1e-9)
    _check_aggregates(m64, x64, e64)
    m32, x32, e32 = orc.density_error_f32(p, ulim, llim, x)
    np.testing.assert_allclose(e32, want, rtol=1e-6)
    _check_aggregates(m32, x32, e32)
    # the wall term is what separates these particles: without it all equal
    assert np.ptp(e64) > 1e-6
    assert math.isclose(np.max(e64), abs(float(poly6_ref(h, 0.0)) / p.pho0 - 1.0), rel_tol=1e-9)


def test_empty_input(orc):
    p = P.dam_params()
    x = np.zeros((0, 3), np.float32)
    assert orc.density_error(p, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), x) == (0.0, 0.0)
