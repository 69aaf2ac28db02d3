"""Pins the FP64 oracle (oracle/oracle.c) before it is trusted as the checker of the HIP library.

Which check pins which oracle function:
  orc_philox4x64        numpy.random.Philox (the same published Philox4x64-10), several counters/keys
  orc_mt64_*            the C++ standard's known answer (10000th draw of a default-seeded mt19937_64)
                        and g++'s std::mt19937_64 on other seeds (compiled here when g++ is present)
  orc_binomial_steps    mean/variance of Binomial(2n, 1/2)
  orc_lennard_jones,    the reference's own code-carsten LennardJones / potential_energy, compiled from
  orc_potential_*       its sources (oracle/_ref) and frozen in tests/golden/carsten_potential.json;
                        closed forms U(2^(1/6) sigma) = -eps, U(sigma) = 0, U(rc-) = 4(rc^-12 - rc^-6)
  orc_force_binning     all-images brute force (the reference's all_neighbour strategy), test_binning.cpp
                        tolerances, and the 5-point derivative of V (test_binning.cpp:108-168)
  orc_binning(_refalgo) the O(N*cells) restatement of binning.cpp, invariants, stability
  orc_virial_binning    two-particle closed form and -dU/dlambda under affine scaling
  orc_kick / orc_drift  time reversibility of the integrators and dH ~ dt^2 (LEAP, OMF2) / dt^4 (OMF4)
  orc_kick_lp / orc_drift_lp (lp_leapfrog, lp_omf4)  bit-exact reversibility (Levesque-Verlet) on the grid,
                        states on the grid, the plain scheme at n_prec = 16
  orc_rdf               brute-force numpy minimum-image histogram, ideal-gas g(r) -> 1, pair-count sum rule
  orc_heatbath          moments <p^2> = m/beta and <beta K> = 3N/2
  orc_hmc_run           Creutz equality <exp(-dH)> = 1 and acceptance bookkeeping
  orc_lattice_init      site/jitter bounds and the wrap into [0, L)
"""
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
RC = 2.5


def box(N, rho):
    return (N / rho) ** (1.0 / 3.0)


# ---- RNG ------------------------------------------------------------------------------------
def test_philox_matches_numpy(orc):
    for key, ctr in [((0, 0), (0, 0, 0, 0)), ((123, 456), (7, 8, 9, 10)), ((2**64 - 1, 5), (2**63, 1, 2, 2**64 - 2))]:
        g = np.random.Philox(key=np.array(key, dtype=np.uint64), counter=np.array(ctr, dtype=np.uint64))
        want = g.random_raw(4).astype(np.uint64)
        nxt = list(ctr)
        nxt[0] = (nxt[0] + 1) % 2**64   # numpy increments the counter before each block
        if nxt[0] == 0:
            nxt[1] = (nxt[1] + 1) % 2**64
        assert np.array_equal(orc.philox4x64(np.array(nxt, dtype=np.uint64), np.array(key, dtype=np.uint64)), want)


def test_mt19937_64_known_answer(orc):
    g = orc.MT19937_64(5489)
    for _ in range(9999):
        g.next()
    assert g.next() == 9981545732273789042


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ for std::mt19937_64")
def test_mt19937_64_vs_std(orc):
    src = r'''
#include <random>
#include <cstdio>
int main(){ for (unsigned long long s : {1ULL, 123ULL, 987654321ULL}) { std::mt19937_64 g(s);
  for (int i=0;i<5;i++) printf("%llu\n", (unsigned long long)g());
  double r = ((double)g() - g.min()) / (g.max() - g.min()); printf("%.17g\n", r);} }
'''
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "m.cc"), "w").write(src)
        subprocess.run(["g++", "-O1", "-o", os.path.join(d, "m"), os.path.join(d, "m.cc")], check=True)
        out = subprocess.run([os.path.join(d, "m")], capture_output=True, text=True, check=True).stdout.split()
    k = 0
    for s in (1, 123, 987654321):
        g = orc.MT19937_64(s)
        for _ in range(5):
            assert g.next() == int(out[k])
            k += 1
        assert g.uniform() == float(out[k])
        k += 1


def test_binomial_steps_moments(orc):
    g = orc.MT19937_64(42)
    s = np.array([g.binomial_steps(10) for _ in range(4000)])
    assert abs(s.mean() - 10) < 0.1 and abs(s.var() - 5) < 0.5


# ---- LJ arithmetic vs the reference's own code and closed forms --------------------------------
def test_lennard_jones_vs_reference_golden(orc):
    gold = json.load(open(os.path.join(HERE, "golden", "carsten_potential.json")))
    for e in gold["lennard_jones"]:
        assert orc.lennard_jones(e["r"], e["n"], e["a"], e["b"]) == pytest.approx(e["value"], rel=1e-15)
    for c in gold["cases"]:
        x = np.asarray(c["x"])
        assert orc.potential_energy_open(x) == pytest.approx(c["potential_energy"], rel=1e-13)
        # the periodic all-neighbour sum with sigma = 0.1, eps = 0.25 is the same function when no
        # image and no pair reaches the cutoff
        v = orc.potential_all(x, [1000.0] * 3, 100.0, sigma=0.1, eps=0.25)
        assert v == pytest.approx(c["potential_energy"], rel=1e-12)


def test_reference_lib_live(orc):
    ref = orc.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    rng = np.random.default_rng(9)
    x = np.ascontiguousarray(rng.uniform(0, 0.6, size=(30, 3)))
    assert orc.potential_energy_open(x) == pytest.approx(ref.ref_potential_energy(x, 30), rel=1e-13)


def test_pair_closed_forms(orc):
    L = [20.0] * 3
    for r, want in [(2 ** (1 / 6), -1.0), (1.0, 0.0), (2.4999999, 4 * (2.4999999 ** -12 - 2.4999999 ** -6))]:
        x = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]])
        assert orc.potential_binning(x, L, RC) == pytest.approx(want, abs=1e-12)
        assert orc.potential_all(x, L, RC) == pytest.approx(want, abs=1e-12)
    x = np.array([[5.0, 5.0, 5.0], [5.0 + 2 ** (1 / 6), 5.0, 5.0]])
    assert np.max(np.abs(orc.force_binning(x, L, RC))) < 1e-12
    # beyond the cutoff nothing interacts (truncated, unshifted: Q3)
    x = np.array([[5.0, 5.0, 5.0], [7.5000001, 5.0, 5.0]])
    assert orc.potential_binning(x, L, RC) == 0.0


def test_virial_two_particles(orc):
    L = [20.0] * 3
    r = 1.3
    x = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]])
    want = 48 * (r ** -12 - 0.5 * r ** -6)          # -r dU/dr of 4 (r^-12 - r^-6)
    assert orc.virial_binning(x, L, RC) == pytest.approx(want, rel=1e-13)


def test_virial_is_minus_dU_dlambda(orc):
    N, rho = 400, 0.7
    L = box(N, rho)
    x = orc.lattice_init(N, [L] * 3, 3, 0, 0.2)
    h = 1e-6
    up = orc.potential_all(x * (1 + h), [L * (1 + h)] * 3, RC)
    dn = orc.potential_all(x * (1 - h), [L * (1 - h)] * 3, RC)
    assert orc.virial_binning(x, [L] * 3, RC) == pytest.approx(-(up - dn) / (2 * h), rel=1e-6)


# ---- binning ---------------------------------------------------------------------------------
@pytest.mark.parametrize("N,rho,L", [(512, 0.8, None), (125, 0.58, None), (51, 0.8, None), (300, None, (8.0, 6.0, 4.5))])
def test_binning_refalgo_and_invariants(orc, N, rho, L):
    L = [box(N, rho)] * 3 if L is None else list(L)
    x = orc.lattice_init(N, L, 5, 0, 0.3)
    nbin, sizebin = orc.cutoff_binning(L, RC)
    perm, cnt, off = orc.binning(x, L, RC)
    p2, c2, o2 = orc.binning(x, L, RC, refalgo=True)
    assert np.array_equal(perm, p2) and np.array_equal(cnt, c2) and np.array_equal(off, o2)
    assert cnt.sum() == N and np.array_equal(off, np.concatenate([[0], np.cumsum(cnt)[:-1]]))
    assert np.array_equal(np.sort(perm), np.arange(N))
    for c in range(len(cnt)):
        members = perm[off[c]:off[c] + cnt[c]]
        assert np.all(np.diff(members) > 0)                # stable sort of the identity
        b = np.minimum(np.floor(x[members] / sizebin).astype(int), np.array(nbin) - 1)
        assert np.all(b[:, 0] + nbin[0] * (b[:, 1] + nbin[1] * b[:, 2]) == c)
    # second binning from a shuffled permutation keeps that order within cells (reference Q8)
    rng = np.random.default_rng(1)
    prev = rng.permutation(N).astype(np.int32)
    p3, _, _ = orc.binning(x, L, RC, perm=prev)
    pos = np.argsort(prev)
    for c in range(len(cnt)):
        assert np.all(np.diff(pos[p3[off[c]:off[c] + cnt[c]]]) > 0)


def test_cutoff_binning_geometry(orc):
    nbin, sb = orc.cutoff_binning([33.8585] * 3, 2.5)
    assert nbin == [13] * 3 and sb[0] == pytest.approx(33.8585 / 13)
    with pytest.raises(ValueError):
        orc.cutoff_binning([2.5, 5, 5], 2.5)


# ---- forces ----------------------------------------------------------------------------------
@pytest.mark.parametrize("N,rho,L", [(512, 0.8, None), (125, 0.58, None), (51, 0.8, None), (300, None, (8.0, 6.0, 4.5))])
def test_force_vs_all_images(orc, N, rho, L):
    L = [box(N, rho)] * 3 if L is None else list(L)
    x = orc.lattice_init(N, L, 9, 0, 0.1)
    fb = orc.force_binning(x, L, RC)
    fa = orc.force_all(x, L, RC)
    # test_binning.cpp:14-42: relative 1e-7 where f^2 > 1e-6, else absolute 1e-7
    big = fa ** 2 > 1e-6
    assert np.all(np.abs(fb - fa)[big] <= 1e-7 * np.abs(fa)[big])
    assert np.all(np.abs(fb - fa)[~big] <= 1e-7)
    vb = orc.potential_binning(x, L, RC)
    assert abs(vb - orc.potential_all(x, L, RC)) <= 1e-6 * abs(vb)
    assert abs(np.sum(fb, axis=0)).max() < 1e-9 * np.abs(fb).max()        # Newton's third law


def test_force_five_point_derivative(orc):
    N, rho = 512, 0.8
    L = [box(N, rho)] * 3
    x = orc.lattice_init(N, L, 2, 0, 0.05)
    bins = orc.binning(x, L, RC)
    f = orc.force_binning(x, L, RC, bins=bins)
    h = 1e-6
    rng = np.random.default_rng(0)
    for i, d in zip(rng.integers(0, N, 12), rng.integers(0, 3, 12)):
        vals = []
        for s in (2, 1, -1, -2):
            xx = x.copy()
            xx[i, d] += s * h
            vals.append(orc.potential_binning(xx, L, RC, bins=bins))
        num = (-vals[0] + 8 * vals[1] - 8 * vals[2] + vals[3]) / (12 * h)
        assert abs(num - f[i, d]) <= 1e-3 * max(abs(num), 1e-6)       # f = +grad U (Q1)


# ---- integrators -----------------------------------------------------------------------------
def _sim(orc, N=216, rho=0.6, T=1.5, seed=4):
    L = box(N, rho)
    s = orc.Sim(N, [L] * 3, cutoff=RC, beta=1 / T)
    s.x[:] = orc.lattice_init(N, [L] * 3, seed, 0, 0.05)
    s.p[:] = orc.heatbath(N, 1.0, 1 / T, seed, 0, 1)
    return s, L


@pytest.mark.parametrize("scheme", [0, 6, 2])
def test_integrator_reversible(orc, scheme):
    s, L = _sim(orc)
    x0 = s.x.copy()
    s.integrate(scheme, 0.004, 8, 0.1932)
    s.p[:] *= -1
    s.integrate(scheme, 0.004, 8, 0.1932)
    d = s.x - x0
    d -= L * np.round(d / L)
    assert np.max(np.abs(d)) < 1e-9


def _blob(orc, n_prec, scheme_seed=5):
    """64 atoms (4 x 4 x 4 at spacing 1.1, jittered) in the middle of a 20-sided box: no atom reaches a box
    face in a short run; positions and momenta on the grid 10^-n_prec."""
    L = 20.0
    s = orc.Sim(64, [L] * 3, cutoff=RC, beta=1 / 1.5)
    s.set_lp_exponent(n_prec)
    g = 10.0 ** n_prec
    rng = np.random.default_rng(scheme_seed)
    site = np.stack(np.meshgrid(*[np.arange(4)] * 3, indexing="ij"), -1).reshape(64, 3)
    s.x[:] = np.round((8.35 + 1.1 * site + rng.uniform(-0.05, 0.05, (64, 3))) * g) / g
    s.p[:] = np.round(orc.heatbath(64, 1.0, 1 / 1.5, scheme_seed, 0, 1) * g) / g
    return s, g


def _on_grid(v, g):
    return bool(np.all(np.round(v * g) / g == v))


@pytest.mark.parametrize("scheme,plain", [(1, 0), (3, 2)])
def test_lp_integrators_bit_reversible(orc, scheme, plain):
    """lp_leapfrog / lp_omf4 (code-carsten/integrator.hh:51-79, 156-197) round every update to the grid
    10^-n: forward, flip p, forward again and flip back returns the start state bit for bit (Levesque-Verlet
    reversibility), and every state stays on the grid.  The plain scheme from the same start comes back only
    to rounding level, so a restatement that skips the rounding of either variable fails."""
    s, g = _blob(orc, 6)
    x0, p0 = s.x.copy(), s.p.copy()
    s.integrate(scheme, 0.004, 12, 0.1932)
    assert np.max(np.abs(s.x - x0)) > 1e-3
    assert _on_grid(s.x, g) and _on_grid(s.p, g)
    s.p[:] *= -1
    s.integrate(scheme, 0.004, 12, 0.1932)
    s.p[:] *= -1
    assert np.array_equal(s.x, x0) and np.array_equal(s.p, p0)
    t, _ = _blob(orc, 6)
    t.integrate(plain, 0.004, 12, 0.1932)
    t.p[:] *= -1
    t.integrate(plain, 0.004, 12, 0.1932)
    t.p[:] *= -1
    dx = np.max(np.abs(t.x - x0))
    assert 0 < dx < 1e-10 or not np.array_equal(t.p, p0)


@pytest.mark.parametrize("scheme,plain", [(1, 0), (3, 2)])
def test_lp_integrators_fine_grid_is_plain(orc, scheme, plain):
    """With n_prec = 16 (the reference's default) the rounding is below FP64 resolution of the state, so
    lp_* follows the plain scheme to rounding level; with a coarse grid (n_prec = 3) it visibly does not."""
    a, _ = _sim(orc)
    b, _ = _sim(orc)
    a.set_lp_exponent(16)
    a.integrate(scheme, 0.004, 8, 0.1932)
    b.integrate(plain, 0.004, 8, 0.1932)
    assert np.max(np.abs(a.x - b.x)) < 1e-12 and np.max(np.abs(a.p - b.p)) < 1e-11
    c, _ = _sim(orc)
    c.set_lp_exponent(3)
    c.integrate(scheme, 0.004, 8, 0.1932)
    assert _on_grid(c.x, 1e3) and _on_grid(c.p, 1e3)
    assert np.max(np.abs(c.x - b.x)) > 1e-5


@pytest.mark.parametrize("scheme,order", [(0, 2), (6, 2), (2, 4)])
def test_integrator_order(orc, scheme, order):
    # dH over a fixed trajectory length scales as dt^order.  A compact 2x2x2 cluster in a large box
    # keeps every pair inside the cutoff, so no energy jump of the truncated potential (Q3) enters.
    dhs = []
    for nsteps in (20, 40):
        L = 20.0
        s = orc.Sim(8, [L] * 3, cutoff=RC, beta=2.0)
        g = np.stack(np.meshgrid([0, 1], [0, 1], [0, 1], indexing="ij"), -1).reshape(8, 3)
        s.x[:] = 9.0 + 1.12 * g + 0.02 * np.arange(24).reshape(8, 3) / 24
        s.p[:] = orc.heatbath(8, 1.0, 2.0, 8, 0, 1)
        H0 = s.compute_potential() + orc.kinetic(s.p, 1.0)
        s.integrate(scheme, 0.08 / nsteps, nsteps, 0.1932)
        dhs.append(abs(s.compute_potential() + orc.kinetic(s.p, 1.0) - H0))
    ratio = dhs[0] / dhs[1]
    assert 2 ** order * 0.5 < ratio < 2 ** order * 2.0, ratio


def test_update_sequencing_counts(orc):
    s, _ = _sim(orc)
    s.integrate(0, 0.005, 20)
    assert s.force_evals == 21         # LEAP: steps + 1 (Q12)
    s.integrate(6, 0.005, 10, 0.1938)
    assert s.force_evals == 21 + 21    # OMF2: 2 steps + 1


# ---- heat bath, lattice ------------------------------------------------------------------------
def test_heatbath_moments(orc):
    N, m, beta = 20000, 2.0, 0.7
    p = orc.heatbath(N, m, beta, 5, 0, 3)
    assert abs(np.mean(p ** 2) - m / beta) < 0.02 * m / beta
    assert abs(np.mean(p)) < 0.03
    K = orc.kinetic(p, m)
    assert abs(beta * K - 1.5 * N) < 4 * np.sqrt(1.5 * N)
    assert not np.array_equal(p, orc.heatbath(N, m, beta, 5, 0, 4))       # fresh draw per trajectory
    assert not np.array_equal(p, orc.heatbath(N, m, beta, 5, 1, 3))       # independent chains


def test_lattice(orc):
    N, L = 1000, 11.0
    x = orc.lattice_init(N, [L] * 3, 1, 0, 0.05)
    a = L / 10
    site = (np.stack([np.arange(N) % 10, (np.arange(N) // 10) % 10, np.arange(N) // 100], 1) + 0.5) * a
    assert np.all(np.abs(x - site) <= 0.05) and np.all((x >= 0) & (x < L))
    x5 = orc.lattice_init(1030302, [107.4942] * 3, 1, 0, 0.05)[:3]       # 102^3 > N: truncated lattice
    assert x5.shape == (3, 3)


# ---- RDF ---------------------------------------------------------------------------------------
def test_rdf_brute_force_and_ideal_gas(orc):
    N, L, nb = 400, 9.0, 90
    rng = np.random.default_rng(2)
    x = rng.uniform(0, L, size=(N, 3))
    counts, g = orc.rdf(x, [L] * 3, nb, L / (2 * nb))
    d = x[None, :, :] - x[:, None, :]
    d -= L * np.round(d / L)
    r = np.sqrt((d ** 2).sum(-1))[~np.eye(N, dtype=bool)]
    b = (r / (L / (2 * nb))).astype(int)
    brute = np.bincount(b[b < nb], minlength=nb)
    assert np.abs(counts.astype(np.int64) - brute).sum() <= 4      # ties at rounding only
    assert counts.sum() == np.sum(r < L / 2) or abs(int(counts.sum()) - int(np.sum(r < L / 2))) <= 4
    assert abs(np.mean(g[nb // 4:]) - 1.0) < 0.03                   # ideal gas


# ---- HMC ---------------------------------------------------------------------------------------
def test_hmc_creutz_and_bookkeeping(orc):
    N, rho, T = 125, 0.6, 1.5
    L = box(N, rho)
    s = orc.Sim(N, [L] * 3, cutoff=RC, beta=1 / T)
    s.x[:] = orc.lattice_init(N, [L] * 3, 1, 0, 0.05)
    r = s.hmc(300, 20, 0, 0.005, 10, seed=3)
    assert r["nacc"] == int(r["acc"][19:].sum())               # i < therm accepted without a draw
    assert np.all(r["acc"][:19] == 1)
    assert abs(np.mean(np.exp(-r["dh"][50:])) - 1.0) < 0.05
    # rejected trajectories keep U
    acc = r["acc"]
    for t in range(1, len(acc)):
        if not acc[t]:
            assert r["u"][t] == r["u"][t - 1]
