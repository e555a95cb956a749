"""Pins of the oracle's halo finder (find_halos.c:4-563) on the CPU.

Checks: the sphere-count table against OEIS A000605 and brute force, the
peak finder against a numpy mask, the Lomuto permutation against a hand
trace and an independent recursive implementation, growth against the
one-spike closed form and against invariants every catalogue must satisfy,
the taus generator against GSL's published known-answer value, the sigma
and mass-function tables against numpy/scipy integrals, and the box
parameters against their closed forms.
"""

import math
import sys

import numpy as np
import pytest
from scipy.integrate import cumulative_trapezoid

from exshalos_amd import synth
from oracle import pipeline as O

LC = 4.0


# ----------------------------------------------------------------- spheres
def test_cells_in_spheres_matches_oeis_a000605():
    """Number of lattice points with i^2 + j^2 + k^2 <= n over the
    representable n (OEIS A000605 with repeats removed)."""
    a000605 = [1, 7, 19, 27, 33, 57, 81, 81, 93, 123, 147, 171, 179, 203, 251, 251, 257,
               305, 341, 365, 389, 437, 461, 461, 485, 515, 587, 619, 619, 691, 739]
    distinct = [v for i, v in enumerate(a000605) if i == 0 or v != a000605[i - 1]]
    got = O.cells_in_spheres(200)
    np.testing.assert_array_equal(got[: len(distinct)], distinct)


def test_cells_in_spheres_brute_force():
    t = 12
    r = np.arange(-t, t + 1)
    s = (r[:, None, None] ** 2 + r[None, :, None] ** 2 + r[None, None, :] ** 2).ravel()
    vals = np.unique(s[s < t * t])
    expect = [(s <= v).sum() for v in vals]
    got = O.cells_in_spheres(t * t)
    np.testing.assert_array_equal(got[: len(expect)], expect)


# ------------------------------------------------------------------- peaks
@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("quant", [0, 16])
def test_peaks_match_numpy_mask(prec, quant):
    """Check_Peak: strictly above all six periodic neighbours, listed in
    lexicographic order (find_halos.c:96-156).  Quantised fields have ties,
    which are never peaks."""
    R = np.float32 if prec == 4 else np.float64
    d = np.random.default_rng(5).standard_normal((16, 12, 20))
    if quant:
        d = np.round(d * quant) / quant
    d = d.astype(R)
    m = np.ones(d.shape, bool)
    for ax in range(3):
        for s in (1, -1):
            m &= d > np.roll(d, s, axis=ax)
    px, pd = O.find_peaks(d, prec)
    np.testing.assert_array_equal(px, np.argwhere(m))
    np.testing.assert_array_equal(pd, d[m])


# ------------------------------------------------------------------ Lomuto
def _lomuto_recursive(a):
    """Textbook Lomuto quicksort (last element pivot, "<=" to the left),
    written recursively over (key, original index) pairs."""
    a = list(a)
    sys.setrecursionlimit(10000)

    def part(lo, hi):
        piv = a[hi][0]
        i = lo - 1
        for j in range(lo, hi):
            if a[j][0] <= piv:
                i += 1
                a[i], a[j] = a[j], a[i]
        a[i + 1], a[hi] = a[hi], a[i + 1]
        return i + 1

    def qs(lo, hi):
        while lo < hi:
            p = part(lo, hi)
            if p - lo < hi - p:
                qs(lo, p - 1)
                lo = p + 1
            else:
                qs(p + 1, hi)
                hi = p - 1

    qs(0, len(a) - 1)
    return a


def test_lomuto_hand_trace():
    """Keys [2, 1, 2, 1, 2]: traced by hand, the unstable Lomuto order of
    the original indices is [1, 3, 2, 0, 4]."""
    key, perm = O.lomuto_sort(np.array([2, 1, 2, 1, 2], dtype=np.float64))
    np.testing.assert_array_equal(key, [1, 1, 2, 2, 2])
    np.testing.assert_array_equal(perm, [1, 3, 2, 0, 4])


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("n,levels", [(3000, 7), (3000, 3000), (2500, 40)])
def test_lomuto_matches_recursive_implementation(prec, n, levels):
    R = np.float32 if prec == 4 else np.float64
    keys = np.random.default_rng(n + levels).integers(0, levels, n).astype(R)
    key, perm = O.lomuto_sort(keys, prec)
    ref = _lomuto_recursive(list(zip(keys.tolist(), range(n))))
    np.testing.assert_array_equal(perm, [v for _, v in ref])
    np.testing.assert_array_equal(key, np.sort(keys))


# ------------------------------------------------------------ growth KATs
def _two_spikes(nd, d0, d1=0.1):
    """Zero field with a spike d0 at (4, 5, 6) and a weak peak d1 far away
    (the least dense peak, index 0 after the sort, is never visited,
    find_halos.c:227)."""
    delta = np.zeros((nd, nd, nd))
    delta[4, 5, 6] = d0
    delta[nd - 4, nd - 3, nd - 2] = d1
    return delta


def _lattice_r2(t):
    r = np.arange(-t, t + 1)
    return (r[:, None, None] ** 2 + r[None, :, None] ** 2 + r[None, None, :] ** 2)


@pytest.mark.parametrize("mult", [1.0, 3.0, 10.0, 40.0, 100.0])
def test_spike_constant_barrier_closed_form(bbks, mult):
    """With the constant barrier dc, a spike d0 on a zero background grows
    to the largest sphere with d0 / N >= dc: N = max{N_s : N_s <= d0/dc}
    over the sphere counts N_s, cells with r^2 below the first failing
    shell.  Its mass lies between the masses of N and the next sphere."""
    k, P = bbks
    nd = 24
    p = O.params(nd, LC)
    d0 = mult * p.dc
    res = O.find_halos(_two_spikes(nd, d0), k, P, LC)
    sph = O.cells_in_spheres(400)
    N = max(s for s in sph if d0 / s >= p.dc)
    assert len(res["count"]) == 1
    assert res["count"][0] == N
    np.testing.assert_array_equal(res["x"][0], [4, 5, 6])
    # flagged cells: exactly the sphere r^2 < r2 around the centre
    r2 = int(res["r2"][0])
    c = _lattice_r2(6)
    assert (c < r2).sum() == N
    flag = res["flag"]
    ng = nd**3
    assert (flag == 0).sum() == N  # halo 0; its centre is -0 = 0
    assert ((flag == ng) | (flag == 0)).all()
    nxt = sph[list(sph).index(N) + 1]
    Mh = res["Mh"][0]
    assert N * p.Mcell * (1 - 1e-6) <= Mh <= nxt * p.Mcell * (1 + 1e-6)


def test_spike_below_barrier_gives_no_halo(bbks):
    k, P = bbks
    p = O.params(24, LC)
    res = O.find_halos(_two_spikes(24, 0.99 * p.dc), k, P, LC)
    assert len(res["count"]) == 0
    assert (res["flag"] == 24**3).all()


def test_spike_ellipsoidal_barrier_shrinks(bbks):
    """Ellipsoidal barrier: growth runs against the smallest barrier
    Bt[Ncells-1], then shrinks while the mean lies below Bt[count]
    (find_halos.c:270-340).  For a spike the result is the largest sphere
    N <= N_grow with d0 / N >= Bt[N]."""
    k, P = bbks
    nd = 24
    a, beta, alpha = 0.7, 0.4, 0.6
    p = O.params(nd, LC)
    tabs = O.halo_tables(p, k, P, a, beta, alpha)
    Bt = tabs["Bt"]
    Bmin = Bt[tabs["Ncells"] - 1]
    sph = O.cells_in_spheres(400)
    for mult in (2.0, 8.0, 30.0):
        d0 = mult * Bmin
        res = O.find_halos(_two_spikes(nd, d0), k, P, LC, a=a, beta=beta, alpha=alpha)
        n_grow = max(s for s in sph if d0 / s >= Bmin)
        ok = [s for s in sph if s <= n_grow and d0 / s >= Bt[s]]
        if not ok:
            assert len(res["count"]) == 0
            continue
        assert res["count"][0] == max(ok)


def test_nmin_rejects_and_flags_ng_plus_one(bbks):
    """A halo smaller than Nmin is rejected: its cells get ng + 1
    (find_halos.c:371-375) and it is not listed."""
    k, P = bbks
    nd = 24
    p = O.params(nd, LC)
    res = O.find_halos(_two_spikes(nd, 10.0 * p.dc), k, P, LC, Nmin=8)  # N = 7 < 8
    assert len(res["count"]) == 0
    flag = res["flag"]
    assert (flag == nd**3 + 1).sum() == 7


# -------------------------------------------------------------- invariants
def _catalogue_invariants(delta, res, p, Bt, nmin):
    nd = delta.shape
    ng = int(np.prod(nd))
    flag = res["flag"].reshape(-1)
    nh = len(res["count"])
    sph = set(O.cells_in_spheres(4000).tolist())
    # flags partition the cells
    assert ((flag >= 0) & (flag <= ng + 1) | (flag < 0) & (flag > -nh)).all()
    members = np.bincount(np.abs(flag[np.abs(flag) < max(nh, 1)]), minlength=nh)
    if nh:
        np.testing.assert_array_equal(members[:nh], res["count"])
    d = delta.reshape(-1).astype(np.float64)
    dens = []
    for h in range(nh):
        c = res["x"][h]
        cid = (c[0] * nd[1] + c[1]) * nd[2] + c[2]
        assert flag[cid] == -h
        n = res["count"][h]
        assert n in sph and n >= nmin
        # the accepted mean satisfies the barrier of its size
        mean = d[np.abs(flag) == h].sum() / n if h else d[flag == 0].sum() / n
        assert mean >= Bt[n] * (1 - 1e-9)
        dens.append(d[cid])
    # halos come in order of decreasing centre density
    assert all(dens[i] >= dens[i + 1] for i in range(len(dens) - 1))
    return members


@pytest.mark.parametrize("seed", [1, 2])
def test_catalogue_invariants_random_field(bbks, seed):
    k, P = bbks
    nd = 32
    noise = np.random.default_rng(seed).standard_normal((nd, nd, nd // 2 + 1, 2))
    delta, _, p = O.field_from_modes(k, P, nd, LC, noise)
    res = O.find_halos(delta, k, P, LC)
    assert len(res["count"]) > 20
    _catalogue_invariants(delta, res, p, res["tables"]["Bt"], 1)


def test_catalogue_is_maximal_for_constant_barrier(bbks):
    """With Nmin = 1 and the constant barrier nothing is rejected, so when
    halo h grew, exactly halos 0..h-1 were flagged.  Adding its next
    non-empty shell (flagged cells counting -Mtot, find_halos.c:301) must
    bring the mean below dc."""
    k, P = bbks
    nd = 32
    noise = np.random.default_rng(4).standard_normal((nd, nd, nd // 2 + 1, 2))
    delta, _, p = O.field_from_modes(k, P, nd, LC, noise)
    res = O.find_halos(delta, k, P, LC)
    flag = res["flag"]
    nh = len(res["count"])
    t = 8
    rr = _lattice_r2(t)
    offs = np.argwhere(rr >= 0) - t
    r2 = rr.ravel()
    for h in range(nh):
        c = res["x"][h]
        hr2 = int(res["r2"][h])
        nxt = r2[r2 >= hr2].min()
        sel = offs[r2 <= nxt]
        cells = tuple(((c[None, :] + sel) % nd).T)
        f = flag[cells]
        earlier = (np.abs(f) < h) | ((f == 0) & (h > 0))
        vals = np.where(earlier, -p.Mtot, delta[cells])
        assert vals.sum() / len(sel) < p.dc


# -------------------------------------------------------------------- taus
def test_taus_gsl_known_answer():
    """GSL's test suite checks gsl_rng_taus seeded with 1: its 10000th
    draw is 2733957125 (rng_test in GSL rng/test.c).  The uniforms are the
    draws divided by 2^32 (gsl_rng_uniform)."""
    u = O.taus_uniforms(1, 10000)
    assert int(round(u[-1] * 2.0**32)) == 2733957125
    assert (u >= 0).all() and (u < 1).all()
    # seed 0 is replaced by 1 (taus_set)
    np.testing.assert_array_equal(O.taus_uniforms(0, 50), O.taus_uniforms(1, 50))


# ------------------------------------------------------------------ tables
def test_box_parameters_closed_form():
    """set_cosmology / set_box (exshalos_h.c:43-106) at Lc = 4, Om0 = 0.31,
    z = 0: rho_m = 0.31 * 2.775e11, Mcell = rho_m Lc^3, dc = 1.686 Om^0.0055,
    Ncells = floor(1e16 / Mcell) = 1816."""
    p = O.params(64, LC)
    rho = 0.31 * 2.775e11
    assert math.isclose(p.rhomz, rho, rel_tol=1e-15)
    assert math.isclose(p.Mcell, rho * 64.0, rel_tol=1e-15)
    assert math.isclose(p.dc, 1.686 * 0.31**0.0055, rel_tol=1e-15)
    assert int(np.floor(1e16 / p.Mcell)) == 1816
    assert math.isclose(p.Mtot, rho * 256.0**3, rel_tol=1e-15)
    assert math.isclose(p.Normk * 64**3 * p.Normx, 1.0, rel_tol=1e-14)
    # a cell mass instead of a size
    q = O.params(64, 1e13)
    assert math.isclose(q.Lc, (1e13 / rho) ** (1 / 3), rel_tol=1e-14)
    with pytest.raises(ValueError):
        O.params(64, 500.0)


def _tophat(x):
    return 3.0 * (np.sin(x) - x * np.cos(x)) / x**3


def test_sigma_table_matches_direct_integral(bbks):
    """Sig_grid[i] = sigma^2 of the top-hat of mass i Mcell, with sigma^2(R)
    the trapezoid of P k^2 W(kR)^2 / (2 pi^2) over the input k
    (find_halos.c:4-44).  The table goes through a spline of sigma(M) on 500
    log-spaced radii; against the direct integral it agrees to 1e-5."""
    k, P = bbks
    p = O.params(64, LC)
    tabs = O.halo_tables(p, k, P)
    sig = tabs["Sig_grid"]
    assert sig[0] == 1e30 and tabs["Ncells"] == 1816
    for i in (1, 2, 7, 19, 123, 1000, 1815):
        R = (3.0 * i * p.Mcell / (4.0 * np.pi * p.rhomz)) ** (1.0 / 3.0)
        s2 = np.trapezoid(P * (k * _tophat(k * R)) ** 2, k) / (2 * np.pi**2)
        assert abs(sig[i] / s2 - 1.0) < 1e-5, i
    # sigma decreases with mass; constant barrier table
    assert np.all(np.diff(sig[1:]) < 0)
    np.testing.assert_array_equal(tabs["Bt"], np.full(len(tabs["Bt"]), p.dc))


def test_mass_function_integral_is_a_cumulative_trapezoid(bbks):
    """Int[i] accumulates -(ln sigma) trapezoids of f(sigma)/R^3
    (find_halos.c:75-93).  scipy's cumulative trapezoid of the same
    integrand on the same knots gives the same table."""
    k, P = bbks
    p = O.params(64, LC)
    tabs = O.halo_tables(p, k, P)
    M = tabs["Mknots"]
    R = (3.0 * M / (4.0 * np.pi * p.rhomz)) ** (1.0 / 3.0)
    sig = np.array([np.sqrt(np.trapezoid(P * (k * _tophat(k * r)) ** 2, k) / (2 * np.pi**2))
                    for r in R])
    f = 0.466 * ((sig / 0.99) ** -2.06 + sig**-0.48) * np.exp(-1.31 / sig**2)
    expect = -cumulative_trapezoid(f / R**3, np.log(sig), initial=0.0)
    np.testing.assert_allclose(tabs["Int"], expect, rtol=1e-5, atol=1e-12 * expect.max())
    assert np.all(np.diff(tabs["Int"]) > 0)


@pytest.mark.parametrize("prec", [4, 8])
def test_fp64_sigma_table_over_64_radii(bbks, prec):
    """The FP64 table option (sigma_radii = 64): knots at 64 radii
    log-spaced from the radius of 0.9 Mcell to that of M_max = 1e16, sigma^2
    by numpy's trapezoid in double over the input k, the Tinker integral a
    cumulative trapezoid in ln sigma -- both to 1e-12 whatever the grid's
    float type -- and sigma^2 per cell count within 1e-4 of the direct
    integral (the spline over 64 knots)."""
    k, P = bbks
    R = np.float32 if prec == 4 else np.float64
    kk = np.asarray(k, dtype=R).astype(np.float64)
    PP = np.asarray(P, dtype=R).astype(np.float64)
    p = O.params(64, LC, precision=prec)
    tabs = O.halo_tables(p, k, P, precision=prec, sigma_radii=64)
    M, Int = tabs["Mknots"], tabs["Int"]
    assert len(M) == len(Int) == 64
    rho = float(p.rhomz)
    r0 = (float(p.Mcell) * 0.9 * 3.0 / (4.0 * np.pi * rho)) ** (1.0 / 3.0)
    r1 = (1e16 * 3.0 / (4.0 * np.pi * rho)) ** (1.0 / 3.0)
    Rr = np.logspace(np.log10(r0), np.log10(r1), 64)
    np.testing.assert_allclose(M, 4.0 / 3.0 * np.pi * rho * Rr**3, rtol=1e-13)
    sig = np.array([np.sqrt(np.trapezoid(PP * (kk * _tophat(kk * r)) ** 2, kk) / (2 * np.pi**2))
                    for r in Rr])
    f = 0.466 * ((sig / 0.99) ** -2.06 + sig**-0.48) * np.exp(-1.31 / sig**2)
    expect = -cumulative_trapezoid(f / Rr**3, np.log(sig), initial=0.0)
    np.testing.assert_allclose(Int, expect, rtol=1e-12, atol=1e-14 * expect.max())
    sg = tabs["Sig_grid"]
    for i in (1, 2, 7, 19, 123, 1000, tabs["Ncells"] - 1):
        r = (3.0 * i * float(p.Mcell) / (4.0 * np.pi * rho)) ** (1.0 / 3.0)
        s2 = np.trapezoid(PP * (kk * _tophat(kk * r)) ** 2, kk) / (2 * np.pi**2)
        assert abs(sg[i] / s2 - 1.0) < 1e-4, (i, sg[i] / s2 - 1.0)


def test_fp64_sigma_table_effect_on_the_catalogue(bbks):
    """What the 64-radius FP64 table changes, on one FP64 field of 64^3
    under the ellipsoidal barrier: sigma^2 per cell count moves by < 1e-3
    and the barrier by < 1e-4, which leaves this catalogue (centres, counts,
    flags) unchanged; masses of halos below 50 cells move by < 1e-3, those
    of the largest halos by up to ~8% -- the natural spline of M(Int) over
    64 knots is poorly conditioned where Int(M) flattens.  The reference's
    500-radius table is converged: the FP64 table over 2000 radii gives its
    masses to 1e-5."""
    k, P = bbks
    nd = 64
    eb = (0.80, 0.29, 0.53)
    delta, _, _ = O.field_from_modes(k, P, nd, LC, synth.white_noise(nd, 9, np.float64))
    p = O.params(nd, LC)
    t0 = O.halo_tables(p, k, P, *eb)
    t64 = O.halo_tables(p, k, P, *eb, sigma_radii=64)
    n = t0["Ncells"]
    assert np.abs(t64["Sig_grid"][1:n] / t0["Sig_grid"][1:n] - 1).max() < 1e-3
    assert np.abs(t64["Bt"][1:n] / t0["Bt"][1:n] - 1).max() < 1e-4
    runs = {r: O.find_halos(delta, k, P, LC, 0.31, 0.0, -1.0, 1, *eb, sigma_radii=r)
            for r in (0, 64, 2000)}
    a, b = runs[0], runs[64]
    assert len(a["Mh"]) > 1000
    for f in ("x", "count", "flag"):
        np.testing.assert_array_equal(a[f], b[f])
    rel = np.abs(b["Mh"] / a["Mh"] - 1)
    small = a["count"] < 50
    assert rel[small].max() < 1e-3
    assert 1e-2 < rel.max() < 0.1
    np.testing.assert_allclose(runs[2000]["Mh"], a["Mh"], rtol=1e-5)


def test_fp64_sigma_table_mass_error_by_radii(bbks):
    """The mass drawn for each sphere count (u = 0, 1/2, 0.999) from the FP64
    table over 64, 128 and 256 radii against the reference's FP64 table (one
    radius per k: 500), for the nd = 512, Lc = 4 box (Ncells = 1816):
    64 radii leave the natural spline of M(Int) so coarse where Int(M)
    flattens that a draw near 600 cells is off by more than 50%; 128 radii
    bring every draw within 1e-2 and 256 within 1e-3.  The draw follows
    find_halos.c:453-476 with scipy's natural splines, the same functions
    as the oracle's (pinned by the mass-draw test below)."""
    from scipy.interpolate import CubicSpline

    k, P = bbks
    p = O.params(512, LC)
    sph = O.cells_in_spheres(20000)
    counts = [int(c) for c in sph[:-1] if c * p.Mcell <= 1e16]

    def draws(nr):
        t = O.halo_tables(p, k, P, sigma_radii=nr)
        I = CubicSpline(t["Mknots"], t["Int"], bc_type="natural")
        inv = CubicSpline(t["Int"], t["Mknots"], bc_type="natural")
        out = []
        for i, c in enumerate(counts):
            a, b = I(c * p.Mcell), I(sph[i + 1] * p.Mcell)
            out.append([inv(a + (b - a) * u) for u in (0.0, 0.5, 0.999)])
        return np.array(out)

    ref = draws(0)
    err = {nr: np.abs(draws(nr) / ref - 1).max(axis=1) for nr in (64, 128, 256)}
    cs = np.array(counts)
    assert err[64].max() > 0.5 and 300 < cs[err[64].argmax()] < 1000
    assert err[64][cs < 50].max() < 1e-3
    assert err[128].max() < 1e-2
    assert err[256].max() < 1e-3


# ------------------------------------------------------------- mass draw
@pytest.mark.parametrize("prec", [4, 8])
def test_mass_draw_uses_the_halo_index_uniform(bbks, prec):
    """Compute_Mass (find_halos.c:453-476): halo i (catalogue order, densest
    first) takes the i-th uniform of taus seeded with the barrier seed and
    Mh = InvI(I(count Mcell) + (I(next Mcell) - I(count Mcell)) u_i), with
    the products count*Mcell in the grid's float type and the interpolated
    integral rounded to it (den_tmp is an fft_real).  Recomputed here from
    the taus stream (pinned by GSL's known answer), scipy's natural cubic
    splines of the Tinker table (pinned above) and the sphere table (pinned
    by OEIS A000605), for four isolated spikes of different heights, so
    their counts differ and a permuted or shifted uniform fails."""
    from scipy.interpolate import CubicSpline

    k, P = bbks
    R = np.float32 if prec == 4 else np.float64
    nd = 48
    p = O.params(nd, LC, precision=prec)
    delta = np.zeros((nd, nd, nd))
    spots = [(6, 6, 6), (30, 8, 10), (8, 30, 34), (34, 34, 20), (20, 44, 44)]
    heights = [300.0, 120.0, 40.0, 12.0, 0.05]  # the last is the unvisited peak 0
    for s, hgt in zip(spots, heights):
        delta[s] = hgt * p.dc
    seed = 4321
    res = O.find_halos(delta.astype(R), k, P, LC, seed=seed, precision=prec, p=p)
    cnt = res["count"]
    assert len(cnt) == 4 and len(set(cnt.tolist())) == 4
    np.testing.assert_array_equal(res["x"], np.array(spots[:4], dtype=np.int32))
    tabs = res["tables"]
    I = CubicSpline(tabs["Mknots"], tabs["Int"], bc_type="natural")
    InvI = CubicSpline(tabs["Int"], tabs["Mknots"], bc_type="natural")
    sph = list(O.cells_in_spheres(2000))
    u = O.taus_uniforms(seed, len(cnt))

    def draw(uu):
        out = []
        for c, ui in zip(cnt, uu):
            nxt = sph[sph.index(int(c)) + 1]
            m0 = float(R(c) * R(p.Mcell))
            m1 = float(R(nxt) * R(p.Mcell))
            v = R(I(m0) + (I(m1) - I(m0)) * ui)
            out.append(R(InvI(float(v))))
        return np.array(out, dtype=np.float64)

    tol = 2e-6 if prec == 4 else 1e-9
    got = res["Mh"].astype(np.float64)
    np.testing.assert_allclose(got, draw(u), rtol=tol, atol=0)
    # the bar separates the right draw from plausible wrong ones
    for wrong in (u[::-1], np.roll(u, 1), O.taus_uniforms(seed + 1, len(cnt))):
        assert np.max(np.abs(got / draw(wrong) - 1.0)) > 100 * tol
