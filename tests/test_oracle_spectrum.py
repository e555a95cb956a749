"""Pins of the oracle's mass assignment (gridmodule.c:9-53, 146-270) and
power spectrum (spectrum_h.c:30-167, powermodule.c:4-108) on the CPU.

Mass assignment is compared with the textbook kernels evaluated at every
cell centre (NGP: the containing cell; CIC: prod(1 - |d|); TSC: 3/4 - d^2
or (3/2 - |d|)^2 / 2; sphere: the cells whose centre lies within R;
exponential: exp(-r^2 / 2R^2) within R R_times), with mass conservation
checked by hypothesis.  The power spectrum is compared with a single
Fourier mode, whose power, window (sinc powers; 3 j1(kR)/(kR) for the
sphere; exp(-k^2 R^2 / 2)), interlacing and multipole factors are known in
closed form, and with mode counts enumerated by numpy.
"""

import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st

from scipy.special import spherical_jn

from oracle import pipeline as O


def _kernel(d, window):
    a = np.abs(d)
    if window == 1:
        return ((d >= -0.5) & (d < 0.5)).astype(float)
    if window == 2:
        return np.clip(1.0 - a, 0.0, None)
    return np.where(a <= 0.5, 0.75 - d * d, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))


def _brute_grid(pos, nd, L, window, mass=None):
    """Sum of the kernel over every cell centre, with periodic distances
    in cell units."""
    Ld = L / nd
    g = np.zeros((nd, nd, nd))
    c = np.arange(nd) + 0.5
    for n, x in enumerate(pos):
        w = []
        for a in range(3):
            d = x[a] / Ld - c
            d = (d + nd / 2) % nd - nd / 2
            w.append(_kernel(d, window))
        m = 1.0 if mass is None else mass[n]
        g += m * w[0][:, None, None] * w[1][None, :, None] * w[2][None, None, :]
    return g


@pytest.mark.parametrize("window", [1, 2, 5])
def test_deposit_weights_match_kernels(window):
    nd, L = 12, 48.0
    rng = np.random.default_rng(window)
    pos = rng.uniform(0, L, (40, 3))
    pos[0] = [0.01, L - 0.02, 47.999]  # near the periodic edges
    mass = rng.uniform(0.5, 2.0, 40)
    g = O.grid_compute(pos, nd, L, window=window, mass=mass)[0, 0]
    np.testing.assert_allclose(g, _brute_grid(pos, nd, L, window, mass), atol=1e-12)


def _brute_sphere(pos, nd, L, R, R_times, expw, mass=None):
    """Every cell centre at periodic distance r from the particle: m if
    r <= R (sphere), m exp(-r^2 / 2R^2) if r <= R R_times (exponential)."""
    Ld = L / nd
    c = (np.indices((nd, nd, nd)) + 0.5) * Ld
    g = np.zeros((nd, nd, nd))
    M = 0.0
    for n, x in enumerate(pos):
        d = x[:, None, None, None] - c
        d = (d + L / 2) % L - L / 2
        r2 = (d * d).sum(axis=0)
        m = 1.0 if mass is None else mass[n]
        if expw:
            w = np.where(r2 <= (R * R_times) ** 2, np.exp(-0.5 * r2 / R**2), 0.0)
        else:
            w = (r2 <= R * R).astype(float)
        g += m * w
        M += m * w.sum()
    return g, M


@pytest.mark.parametrize("window,R,R_times", [(3, 5.0, 1.0), (3, 9.0, 1.0), (4, 3.0, 2.0),
                                              (4, 4.5, 1.5)])
def test_sphere_and_exponential_deposits(window, R, R_times):
    """Windows 3 and 4 (gridmodule.c:57-119) against every cell centre: with
    masses the grid is the weighted sum; without, it is normalised by the
    deposited weight (not the particle count) to rho / rho_bar - 1."""
    nd, L = 12, 48.0
    rng = np.random.default_rng(window + int(R))
    pos = rng.uniform(0, L, (25, 3))
    pos[0] = [0.01, L - 0.02, 47.999]  # near the periodic edges
    pos[1] = [2.0, 2.0, 2.0]           # a cell centre
    mass = rng.uniform(0.5, 2.0, 25)
    g = O.grid_compute(pos, nd, L, window=window, mass=mass, R=R, R_times=R_times)[0, 0]
    ref, _ = _brute_sphere(pos, nd, L, R, R_times, window == 4, mass)
    np.testing.assert_allclose(g, ref, atol=1e-12)
    g = O.grid_compute(pos, nd, L, window=window, R=R, R_times=R_times)[0, 0]
    ref, M = _brute_sphere(pos, nd, L, R, R_times, window == 4)
    np.testing.assert_allclose(g, ref / (M / nd**3) - 1.0, atol=1e-12)
    if window == 3:  # a particle at a cell centre fills the lattice ball
        one = O.grid_compute(pos[1:2], nd, L, window=3, mass=np.ones(1), R=R)[0, 0]
        i = np.arange(-4, 5)
        n_ball = int(((i[:, None, None] ** 2 + i[None, :, None] ** 2 + i[None, None, :] ** 2)
                      * (L / nd) ** 2 <= R * R).sum())
        assert one.sum() == n_ball


def test_single_particle_at_a_cell_centre():
    """All the weight goes to that cell; the unweighted grid is normalised
    to rho / rho_bar - 1 (gridmodule.c:197-202)."""
    nd, L = 8, 32.0
    Ld = L / nd
    pos = np.array([[(2 + 0.5) * Ld, (5 + 0.5) * Ld, (7 + 0.5) * Ld]])
    for window in (1, 2, 5):
        g = O.grid_compute(pos, nd, L, window=window)[0, 0]
        if window == 5:  # TSC spreads to the neighbours even at the centre
            continue
        expect = np.full((nd, nd, nd), -1.0)
        expect[2, 5, 7] = nd**3 - 1.0
        np.testing.assert_allclose(g, expect, atol=1e-12)


@pytest.mark.parametrize("window", [1, 2, 5])
def test_interlaced_grid_is_the_half_cell_shifted_deposit(window):
    """The second grid holds the particles moved by -Lb/2 on every axis
    (gridmodule.c:233-262)."""
    nd, L = 10, 40.0
    pos = np.random.default_rng(7).uniform(0, L, (30, 3))
    g = O.grid_compute(pos, nd, L, window=window, interlacing=True, mass=np.ones(30))
    np.testing.assert_allclose(g[0, 0], _brute_grid(pos, nd, L, window), atol=1e-12)
    shifted = np.mod(pos - 0.5 * L / nd, L)
    np.testing.assert_allclose(g[0, 1], _brute_grid(shifted, nd, L, window), atol=1e-12)


def test_redshift_space_shift():
    """direction d moves x_d by v_d (1+z)^3 / H(z), H = 100 sqrt(Om (1+z)^3
    + 1 - Om) (gridmodule.c:167-169)."""
    nd, L, z, Om = 10, 40.0, 0.5, 0.3
    rng = np.random.default_rng(8)
    pos = rng.uniform(0, L, (25, 3))
    vel = rng.normal(0, 200, (25, 3))
    H = 100 * np.sqrt(Om * (1 + z) ** 3 + 1 - Om)
    moved = pos.copy()
    moved[:, 1] = np.mod(moved[:, 1] + vel[:, 1] * (1 + z) ** 3 / H, L)
    g = O.grid_compute(pos, nd, L, window=2, mass=np.ones(25), direction=1, vel=vel, Om0=Om,
                       z=z)[0, 0]
    np.testing.assert_allclose(g, _brute_grid(moved, nd, L, 2), atol=1e-10)


@settings(max_examples=25, deadline=None)
@given(n=st.integers(1, 60), seed=st.integers(0, 2**31 - 1),
       window=st.sampled_from([1, 2, 5]))
def test_mass_is_conserved(n, seed, window):
    nd, L = 8, 16.0
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, L, (n, 3))
    mass = rng.uniform(0.1, 3.0, n)
    g = O.grid_compute(pos, nd, L, window=window, mass=mass)
    assert abs(g.sum() - mass.sum()) < 1e-10 * mass.sum()
    # unweighted: the normalised grid averages to zero
    g = O.grid_compute(pos, nd, L, window=window)
    assert abs(g.mean()) < 1e-12


# ------------------------------------------------------------------- P(k)
def _mode_table(nd, L, k_min, k_max, Nk):
    """All half-complex modes kept by the reference's mask
    (powermodule.c:50-51), with |k| and float64 bin index."""
    kn = 2 * np.pi / L
    i, j, k = np.meshgrid(np.arange(nd), np.arange(nd), np.arange(nd // 2 + 1), indexing="ij")
    plane = (k == 0) | (k == nd // 2)
    drop = plane & (((i > nd // 2) & (j > nd // 2)) | (i == 0) | (i == nd // 2) | (j == 0) |
                    (j == nd // 2) | ((i > nd // 2) & (j < nd // 2)))
    kx = np.where(2 * i < nd, i, i - nd) * kn
    ky = np.where(2 * j < nd, j, j - nd) * kn
    kz = np.where(2 * k < nd, k, k - nd) * kn
    kmod = np.sqrt(kx * kx + ky * ky + kz * kz)
    dk = (k_max - k_min) / Nk
    b = np.floor((kmod - k_min) / dk).astype(int)
    keep = ~drop & (b >= 0) & (b < Nk)
    return kmod[keep], b[keep]


@pytest.mark.parametrize("nd,Nk", [(16, 8), (16, 40), (12, 5)])
def test_mode_counts_and_mean_k(nd, Nk):
    L = 4.0 * nd
    kmin, kmax = 2 * np.pi / L, np.pi / L * nd
    kmod, b = _mode_table(nd, L, kmin, kmax, Nk)
    grid = np.random.default_rng(1).normal(size=(1, 1, nd, nd, nd))
    out = O.power_compute(grid, 1, nd, L, window=0, Nk=Nk, k_min=kmin, k_max=kmax)
    np.testing.assert_array_equal(out["Nk"], np.bincount(b, minlength=Nk))
    kmean = np.bincount(b, weights=kmod, minlength=Nk) / np.maximum(out["Nk"], 1)
    np.testing.assert_allclose(out["k64"], kmean, rtol=1e-13)


def test_total_mode_count_closed_form():
    """Over all |k|: nd^2 (nd/2 - 1) interior modes plus 2 (nd/2 - 1)
    (nd - 2) in the kz in {0, nd/2} planes survive the mask."""
    nd, L = 16, 64.0
    grid = np.zeros((1, 1, nd, nd, nd))
    out = O.power_compute(grid, 1, nd, L, window=0, Nk=1, k_min=0.0,
                          k_max=1.01 * np.sqrt(3) * np.pi * nd / L)
    assert out["Nk"][0] == nd * nd * (nd // 2 - 1) + 2 * (nd // 2 - 1) * (nd - 2)


@pytest.mark.parametrize("window", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("il", [False, True])
def test_single_z_mode_power(window, il):
    """grid = A cos(2 pi m z / nd): the only non-zero kept mode is (0, 0, m)
    with |delta_k|^2 = A^2 L^3 / 4 after Normk = (L/nd^2)^(3/2); it is
    divided by W^2 (W = sinc(pi m/nd)^p, p = 1, 2, 3 for NGP, CIC, TSC,
    spectrum_h.c:58-59), multiplied by cos^2(theta/2), theta = pi m / nd,
    when an identical interlaced copy is combined (spectrum_h.c:130-131),
    and averaged over the bin's modes; the l = 2 multipole with the z
    direction has P_2(1) = 1 and the factor 5.  Windows 3 and 4 divide by
    the sphere's 3 j1(kR)/(kR) and by exp(-k^2 R^2 / 2) (spectrum_h.c:60-67)."""
    nd, L, Nk, m, A, R = 16, 64.0, 8, 3, 0.6, 5.0
    kmin, kmax = 2 * np.pi / L, np.pi / L * nd
    z = np.indices((nd, nd, nd))[2]
    g = A * np.cos(2 * np.pi * m * z / nd)
    grid = np.stack([g, g])[None] if il else g[None, None]
    out = O.power_compute(grid, 1, nd, L, window=window, interlacing=il, Nk=Nk, k_min=kmin,
                          k_max=kmax, l_max=2, direction=2, R=R)
    kmod, b = _mode_table(nd, L, kmin, kmax, Nk)
    counts = np.bincount(b, minlength=Nk)
    bm = int(np.floor((2 * np.pi * m / L - kmin) / ((kmax - kmin) / Nk)))
    u = np.pi * m / nd
    kz = 2 * np.pi * m / L
    if window == 3:
        W = 3 * spherical_jn(1, kz * R) / (kz * R)
    elif window == 4:
        W = np.exp(-0.5 * (kz * R) ** 2)
    else:
        W = 1.0 if window == 0 else (np.sin(u) / u) ** {1: 1, 2: 2, 5: 3}[window]
    P = A**2 * L**3 / 4 / W**2 / counts[bm]
    if il:
        P *= np.cos(u / 2) ** 2
    expect0 = np.zeros(Nk)
    expect0[bm] = P
    np.testing.assert_allclose(out["Pk64"][0, 0], expect0, rtol=1e-10, atol=1e-12 * P)
    np.testing.assert_allclose(out["Pk64"][0, 1], 5 * expect0, rtol=1e-10, atol=1e-12 * P)
    # along x, mu = 0 and P_2(0) = -1/2
    out = O.power_compute(grid, 1, nd, L, window=window, interlacing=il, Nk=Nk, k_min=kmin,
                          k_max=kmax, l_max=2, direction=0, R=R)
    np.testing.assert_allclose(out["Pk64"][0, 1], -2.5 * expect0, rtol=1e-10, atol=1e-12 * P)


def test_cross_spectrum_of_two_tracers():
    """Pairs (l, m <= l) in order (0,0), (1,0), (1,1): the cross power of a
    grid with twice itself is twice the auto power (powermodule.c:66-72)."""
    nd, L = 16, 64.0
    g = np.random.default_rng(3).normal(size=(nd, nd, nd))
    grid = np.stack([g[None], 2 * g[None]])
    out = O.power_compute(grid, 2, nd, L, window=2, Nk=6)
    P = out["Pk64"][:, 0]
    np.testing.assert_allclose(P[1], 2 * P[0], rtol=1e-12)
    np.testing.assert_allclose(P[2], 4 * P[0], rtol=1e-12)


@pytest.mark.parametrize("precision", [4, 8])
@pytest.mark.parametrize("given_range", [False, True])
def test_abundance(precision, given_range):
    """abundance_compute / Measure_Abundance (spectrum.c:535-600,
    abundance.c:4-36): counts per logarithmic bin against np.histogram on the
    edges ln Mmin + i dlnM (dlnM a float), bin 0 left empty as the reference
    counts only 0 < bin < Nm; each bin's mean mass lies inside the bin;
    dn V dlnM is the integer count and dn_err the Poisson error."""
    R = np.float32 if precision == 4 else np.float64
    rng = np.random.default_rng(7)
    M = np.exp(rng.uniform(np.log(2e12), np.log(3e15), 5000)).astype(R)
    Nm, Lc, nd = 12, 4.0, 64
    if given_range:
        Mmin, Mmax = R(5e12), R(1e15)
    else:
        Mmin, Mmax = R(M.min() * R(0.9999)), R(M.max() * R(1.0001))
    out = O.abundance(M, Mmin if given_range else -1.0, Mmax if given_range else -1.0, Nm, Lc,
                      nd, nd, nd, precision=precision)
    dlnM = np.float32((np.log(np.float64(Mmax)) - np.log(np.float64(Mmin))) / Nm)
    edges = np.log(np.float64(Mmin)) + np.arange(Nm + 1) * np.float64(dlnM)
    counts, _ = np.histogram(np.log(M.astype(np.float64)), bins=edges)
    counts[0] = 0
    V = (Lc * nd) ** 3
    N = out["dn"].astype(np.float64) * V * np.float64(dlnM)
    np.testing.assert_allclose(N, counts, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(out["dn_err"].astype(np.float64) * V * np.float64(dlnM),
                               np.sqrt(counts), rtol=1e-5)
    full = counts > 0
    lnm = np.log(out["Mh"][full].astype(np.float64))
    lo = edges[:-1][full]
    assert np.all(lnm >= lo - 1e-6) and np.all(lnm < lo + dlnM + 1e-6)
    assert out["Mh"][0] == 0 and out["dn"][0] == 0


def test_bispectrum_configurations():
    """Configuration order of bimodule.c:49-317: bins non-increasing in each
    triple, no triple twice, every b >= floor(a/2) triple closes in bin space
    (c >= a - b - 1), and each a ends with (a, a, a)."""
    tri = O.bispectrum_triangles(12)
    assert len(set(tri)) == len(tri)
    assert all(a >= b >= c for a, b, c in tri)
    assert all(c >= a - b - 1 for a, b, c in tri if b >= a // 2 and (a, b, c) != (a, a, b))
    ends = [i for i, t in enumerate(tri) if t[0] == t[1] == t[2]]
    assert [tri[i][0] for i in ends] == list(range(12))
    assert ends[-1] == len(tri) - 1


@pytest.mark.parametrize("window", [0, 2])
def test_bispectrum_equals_fourier_triangle_sums(window):
    """Sum over cells of F_a F_b F_c, F_s the inverse transform of the modes of
    shell s, equals ng times the sum over closed triangles k1 + k2 + k3 = 0
    (k_i in shells a, b, c) of D(k1) D(k2) D(k3), D the deconvolved,
    normalised modes on the full grid; the I fields count closed triangles;
    P is Parseval's sum over the shell."""
    nd, L, Nk = 8, 64.0, 4
    rng = np.random.default_rng(5 + window)
    g = rng.normal(size=(nd, nd, nd))
    out = O.bispectrum(g[None, None], 1, nd, L, window=window, Nk=Nk, precision=8)
    kn = 2 * np.pi / L
    Lb = L / nd
    f = np.fft.fftfreq(nd, 1.0 / nd)
    kx, ky, kz = np.meshgrid(f * kn, f * kn, f * kn, indexing="ij")
    W = 1.0
    if window == 2:
        s = lambda x: np.where(x == 0, 1.0, np.sin(x) / np.where(x == 0, 1.0, x))  # noqa: E731
        W = (s(kx * Lb / 2) * s(ky * Lb / 2) * s(kz * Lb / 2)) ** 2
    D = np.fft.fftn(g) * (L / nd**2) ** 1.5 / W
    kmin, kmax = 2 * np.pi / L, np.pi / L * nd
    b = np.floor((np.sqrt(kx**2 + ky**2 + kz**2) - kmin) / ((kmax - kmin) / Nk)).astype(int)
    idx = np.indices((nd, nd, nd)).reshape(3, -1).T
    bf, Df = b.reshape(-1), D.reshape(-1)
    ng = nd**3
    for n, (a, bb, c) in enumerate(O.bispectrum_triangles(Nk)):
        i1, i2 = np.nonzero(bf == a)[0], np.nonzero(bf == bb)[0]
        k3 = (-(idx[i1][:, None, :] + idx[i2][None, :, :])) % nd
        j3 = (k3[..., 0] * nd + k3[..., 1]) * nd + k3[..., 2]
        closed = bf[j3] == c
        S = ng * np.sum(Df[i1][:, None] * Df[i2][None, :] * Df[j3] * closed)
        assert abs(S.imag) < 1e-8 * max(abs(S.real), 1.0)
        np.testing.assert_allclose(out["B64"][0, n], S.real, rtol=1e-9,
                                   atol=1e-9 * np.max(np.abs(out["B64"])))
        np.testing.assert_allclose(out["IB64"][n], ng * closed.sum(), rtol=1e-12)
    for a in range(Nk):
        m = bf == a
        P = ng * np.sum(np.abs(Df[m]) ** 2) / (ng * m.sum())
        np.testing.assert_allclose(out["Pk"][0, a], P, rtol=1e-9)
        np.testing.assert_allclose(out["Nk"][a], m.sum(), rtol=1e-12)


@pytest.mark.parametrize("prec", [4, 8])
def test_openmp_build_equals_the_parity_build(prec):
    """The CPU baseline's OpenMP build (threaded deposit with atomic adds,
    deconvolution and binning) gives the sequential build's grids and
    spectra up to the order of the float sums."""
    R = np.float32 if prec == 4 else np.float64
    rng = np.random.default_rng(12)
    nd, L = 24, 96.0
    pos = (rng.random((30000, 3)) * L).astype(R)
    ref_g = O.grid_compute(pos, nd, L, window=2, interlacing=True, precision=prec)
    ref_p = O.power_compute(ref_g, 1, nd, L, window=2, interlacing=True, Nk=10, precision=prec)
    O.use_openmp(True)
    try:
        assert O.lib().ora_omp_threads(4) == 4
        g = O.grid_compute(pos, nd, L, window=2, interlacing=True, precision=prec)
        p = O.power_compute(ref_g, 1, nd, L, window=2, interlacing=True, Nk=10, precision=prec)
    finally:
        O.use_openmp(False)
    tol = 2e-6 if prec == 4 else 1e-13
    np.testing.assert_allclose(g, ref_g, rtol=0, atol=tol * np.abs(ref_g).max())
    np.testing.assert_array_equal(p["Nk"], ref_p["Nk"])
    np.testing.assert_allclose(p["Pk64"], ref_p["Pk64"], rtol=1e-12)
