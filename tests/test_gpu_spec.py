"""Speculative trajectories (chmc_set_speculation) against the sequential HMC loop of HMC.cpp:77-117.

K copies of one chain run its next K trajectories from the same state; the Metropolis tests take them in
order up to the first accepted one, whose end state every copy takes over.  The sequential loop starts
trajectory t + k from that same state exactly when t + 1 .. t + k - 1 were rejected, so the chain must be
the sequential one bit for bit: every record field (dH, U, K, W, decision), the acceptance count, and
the final positions and momenta.  Checked at an acceptance near one half (most batches end early and hand
an accepted copy's positions and lists to the others), through the accept-all thermalisation (every
batch keeps one trajectory) and at N = 32768, T* = 0.722 with the OMF2 10 x 0.01 trajectories of the
benchmark (acceptance of a few percent: most batches keep every trajectory), graph-replayed and direct,
with and without start rebuilds (chmc_set_list_start_rebuild: the lists rebuilt at the start of the
trajectory after an accepted one, in both loops).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RC = 2.5


def _pair(chmc, N, rho, T, K, lanes=8, check=2, skin=0.6, start_rebuild=False, bits=0):
    L = (N / rho) ** (1 / 3)
    ch = chmc.Chain(N=N, L=L, cutoff=RC, beta=1 / T)
    seq = chmc.Engine(ch)
    spec = chmc.Engine(ch, speculate=K)
    assert spec.speculation() == K and spec.nchains == K
    for e in (seq, spec):
        e.set_list_entry_bits(bits)
        e.set_skin(skin)
        e.set_list_lanes(lanes)
        e.set_list_check(check)
        e.set_list_start_rebuild(start_rebuild)
        assert e.list_start_rebuild() == start_rebuild
    return seq, spec


def _same_run(chmc, seq, spec, ntraj, therm, dt, steps, seed, use_graph):
    r1, n1 = seq.hmc(ntraj, therm, chmc.OMF2, dt, steps, seed=seed, use_graph=use_graph)
    r2, n2 = spec.hmc(ntraj, therm, chmc.OMF2, dt, steps, seed=seed, use_graph=use_graph)
    print("accepted", n1[0], "of", ntraj - max(therm - 1, 0), "dh[:8]", np.array2string(r1["dh"][:8, 0], precision=4))
    assert r2.shape == (ntraj, 1) and n2.shape == (1,)
    assert np.array_equal(r1[:, 0], r2[:, 0])            # every field of every record
    assert n1[0] == n2[0]
    assert np.array_equal(seq.positions(), spec.positions())
    assert np.array_equal(seq.momenta(), spec.momenta())
    return r1[:, 0], n1[0]


@pytest.mark.parametrize("use_graph,start_rebuild,bits", [(True, False, 0), (False, False, 0), (True, True, 0),
                                                          (False, True, 0), (True, True, 16), (False, False, 16)])
def test_speculative_equals_sequential_mixed_acceptance(chmc, use_graph, start_rebuild, bits):
    """bits 16: the 16-bit list entries, whose broadcast copies the rows as they are and moves the
    segment starts of every list cell instead."""
    seq, spec = _pair(chmc, 4000, 0.8442, 0.722, 4, start_rebuild=start_rebuild, bits=bits)
    seq.init_lattice(3, 0.05)
    x0 = seq.positions()
    spec.set_positions(x0)                 # chain 0 only: the run copies it into the other copies
    seq.hmc(30, 30, chmc.OMF2, 0.01, 10, seed=5, use_graph=use_graph)
    spec.hmc(30, 30, chmc.OMF2, 0.01, 10, seed=5, use_graph=use_graph)
    assert np.array_equal(seq.positions(), spec.positions())
    rec, nacc = _same_run(chmc, seq, spec, 80, 0, 0.01, 10, 7, use_graph)
    assert 8 <= nacc <= 72, nacc                      # both decisions occur often
    # a second call continues from the state (and the lists) both hold
    _same_run(chmc, seq, spec, 25, 0, 0.01, 10, 8, use_graph)
    seq.close()
    spec.close()


@pytest.mark.parametrize("start_rebuild", [False, True])
def test_speculative_equals_sequential_c2(chmc, start_rebuild):
    seq, spec = _pair(chmc, 32768, 0.8442, 0.722, 8, lanes=816, skin=0.7, start_rebuild=start_rebuild)
    seq.init_lattice(1, 0.05)
    spec.set_positions(seq.positions())
    # accept-all thermalisation (one trajectory per batch), then the benchmark's trajectories
    _same_run(chmc, seq, spec, 12, 12, 0.01, 10, 10, True)
    _same_run(chmc, seq, spec, 60, 0, 0.01, 10, 12, True)
    seq.close()
    spec.close()


def test_speculative_equals_sequential_large_cell_grid(chmc):
    """N = 131072: 19^3 list cells, more than one build workgroup scans, so rebuilds take the separate
    scan kernel and the broadcast's cell offsets are the ones the build reads; lp_leapfrog trajectories
    short enough to be accepted part of the time."""
    seq, spec = _pair(chmc, 131072, 0.8442, 0.722, 2, start_rebuild=True)
    for e in (seq, spec):
        e.set_lp_exponent(12)
    seq.init_lattice(2, 0.05)
    spec.set_positions(seq.positions())
    seq.hmc(4, 4, chmc.OMF2, 0.002, 10, seed=3)
    spec.hmc(4, 4, chmc.OMF2, 0.002, 10, seed=3)
    # dH at this size comes mostly from pairs crossing the unshifted cut-off, in proportion to the
    # trajectory's length: 4 x 0.0005 keeps it near one
    r1, n1 = seq.hmc(24, 0, chmc.LP_LEAPFROG, 0.0005, 4, seed=9)
    r2, n2 = spec.hmc(24, 0, chmc.LP_LEAPFROG, 0.0005, 4, seed=9)
    print("accepted", n1[0], "of 24", np.array2string(r1["dh"][:, 0], precision=3))
    assert 1 <= n1[0] <= 23, n1
    assert np.array_equal(r1[:, 0], r2[:, 0]) and n1[0] == n2[0]
    assert np.array_equal(seq.positions(), spec.positions())
    assert np.array_equal(seq.momenta(), spec.momenta())
    seq.close()
    spec.close()


def test_speculation_refusals(chmc):
    L = (512 / 0.8) ** (1 / 3)
    a, b = chmc.Chain(N=512, L=L, cutoff=RC, beta=1 / 1.5), chmc.Chain(N=512, L=L, cutoff=RC, beta=1 / 1.4)
    e = chmc.Engine([a, b])
    with pytest.raises(chmc.ChmcError):
        e.set_speculation(2)                   # not copies of one chain
    with pytest.raises(chmc.ChmcError):
        e.set_speculation(3)                   # not the context's chain count
    e.close()
    e = chmc.Engine(a, speculate=2)
    e.init_lattice(1, 0.05)
    with pytest.raises(chmc.ChmcError):
        e.hmc(4, 0, chmc.LEAPFROG, 0.005, 10, seed=1, randomize=True)
    e.set_chain_skins([0.5, 0.6])
    with pytest.raises(chmc.ChmcError):
        e.hmc(4, 0, chmc.LEAPFROG, 0.005, 10, seed=1)      # copies with different skins
    e.set_chain_skins([0.5, 0.5])
    e.set_speculation(1)                       # off: two independent chains again
    assert e.speculation() == 1
    rec, nacc = e.hmc(3, 0, chmc.LEAPFROG, 0.005, 10, seed=1)
    assert rec.shape == (3, 2)
    e.close()
    f = chmc.Engine([a, a], fp64_forces=True)
    with pytest.raises(chmc.ChmcError):
        f.set_speculation(2)                   # FP64 forces run on the cell path, not the lists
    f.close()
