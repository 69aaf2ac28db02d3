"""The reference's programs over the engine: "main -i infile" (HMC with xyz output, host and device
random state files, append/restart) and "RDF -i infile" (offline g(r) of every saved configuration).

  * xyz output: every saved configuration's potential equals the FP64 oracle's potential of the
    printed positions (relative 1e-9: positions carry 12 significant digits);
  * restart: 6 trajectories + an appended run of 4 write the same trajectories as one run of 10
    (trajectory numbers, heat-bath stream and the host mt19937_64 all continue): identical bytes up
    to the restart, and within the rounding of the 12 printed digits after it;
  * RDF: the program's g(r) of the last configuration equals the oracle's (printed with %g);
  * the reference's own input file (code/builds/test.yaml, kept in tests/golden) runs unchanged as the
    appended continuation of a short run that creates its start file and random states.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAIN = os.path.join(ROOT, "chemhmc_amd", "bin", "main")
RDF = os.path.join(ROOT, "chemhmc_amd", "bin", "RDF")
N = 216
L = (N / 0.8) ** (1 / 3)


def yaml_text(ntraj, start="cold", append=False, startfile=None):
    s = f"""StartCondition: {start}
{"start_configuration_file: " + startfile if startfile else ""}
output_file: out.xyz
append: {"true" if append else "false"}
seed: 7
rng_host_state: rng_host
rng_device_state: rng_device
Ntrajectories: {ntraj}
thermalization_steps: 2
save_every: 2
randomize_trajectory: 0

geometry:
  Lx: {L!r}
  Ly: {L!r}
  Lz: {L!r}

particles:
  name: identical_particles
  name_xyz: Ar
  N: {N}
  mass: 1
  beta: 0.6666666666666666
  cutoff: 2.5
  eps: 1
  sigma: 1
  algorithm: parallel_binning
  RDF:
    Lmax: 3.0
    size_bin: 0.03
    output_file: RDF.out

integrator:
  name: LEAP
  dt: 0.005
  steps: 10
"""
    return s


def run(prog, d, name):
    assert os.path.exists(prog), "build first: make -C chemhmc_amd all"
    r = subprocess.run([prog, "-i", name], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def read_xyz(path, n):
    lines = open(path).read().splitlines()
    assert len(lines) % (n + 2) == 0
    confs = []
    for c in range(len(lines) // (n + 2)):
        blk = lines[c * (n + 2):(c + 1) * (n + 2)]
        head = blk[1].split()
        traj, K, V = int(head[1]), float(head[3]), float(head[5])
        x = np.array([[float(t) for t in ln.split()[1:4]] for ln in blk[2:]])
        confs.append((traj, K, V, x))
    return confs


def test_main_xyz_restart_and_rdf(tmp_path, orc):
    one, two = tmp_path / "one", tmp_path / "two"
    one.mkdir()
    two.mkdir()
    (one / "in.yaml").write_text(yaml_text(10))
    run(MAIN, one, "in.yaml")
    (two / "in.yaml").write_text(yaml_text(6))
    run(MAIN, two, "in.yaml")
    (two / "more.yaml").write_text(yaml_text(4, start="read", append=True, startfile="out.xyz"))
    out = run(MAIN, two, "more.yaml")
    assert "trajectory 7 " in out and "trajectory 10 " in out
    confs = read_xyz(one / "out.xyz", N)
    confs2 = read_xyz(two / "out.xyz", N)
    assert [c[0] for c in confs] == [2, 4, 6, 8, 10] and [c[0] for c in confs2] == [2, 4, 6, 8, 10]
    # up to the restart the files are identical; after it the run starts from the 12 printed digits
    # of the positions (as the reference's restart does), so it stays within rounding of the other
    a, b = open(one / "out.xyz").read().splitlines(), open(two / "out.xyz").read().splitlines()
    assert a[:3 * (N + 2)] == b[:3 * (N + 2)]
    for (t1, K1, V1, x1), (t2, K2, V2, x2) in zip(confs[3:], confs2[3:]):
        assert np.max(np.abs(x1 - x2)) < 1e-6 and abs(V1 - V2) < 1e-6 * abs(V1) and abs(K1 - K2) < 1e-6 * K1
    for traj, K, V, x in confs:
        Vo = orc.potential_all(x, [L] * 3, 2.5)
        assert abs(V - Vo) <= 1e-9 * abs(Vo), (traj, V, Vo)
        assert 0.5 * 3 * N * 1.5 * 0.5 < K < 2.0 * 3 * N * 1.5     # K ~ 3N/(2 beta)

    run(RDF, one, "in.yaml")
    lines = open(one / "RDF.out").read().splitlines()
    nb = int(3.0 / 0.03)
    assert [float(t) for t in lines[:3]] == pytest.approx([L] * 3, rel=1e-5)
    assert int(lines[3]) == N and int(lines[11]) == 5
    body = lines[12:]
    assert len(body) == 5 * (nb + 1)
    last = body[4 * (nb + 1):]
    assert int(last[0]) == 4
    g = np.array([float(ln.split()[1]) for ln in last[1:]])
    go = orc.rdf(confs[-1][3], [L] * 3, nb, 0.03)[1]
    assert np.allclose(g, go, rtol=1e-5, atol=1e-6)


def test_reference_input_file_runs_as_a_continuation(tmp_path):
    d = tmp_path / "ref"
    d.mkdir()
    gold = open(os.path.join(ROOT, "tests", "golden", "reference_test.yaml")).read()
    # a first run with the file's physics writes out_xyz.txt and the random-state files it appends to
    first = (gold.replace("StartCondition: read", "StartCondition: hot")
                 .replace("append: true", "append: false")
                 .replace("Ntrajectories: 4", "Ntrajectories: 2"))
    (d / "first.yaml").write_text(first)
    run(MAIN, d, "first.yaml")
    shutil.copy(os.path.join(ROOT, "tests", "golden", "reference_test.yaml"), d / "test.yaml")
    out = run(MAIN, d, "test.yaml")
    assert "trajectory 3 " in out and "trajectory 6 " in out
    trajs = [int(ln.split()[1]) for ln in open(d / "out_xyz.txt").read().splitlines() if ln.startswith("trajectory=")]
    assert trajs == [1, 2, 3, 4, 5, 6]
    run(RDF, d, "test.yaml")
    assert os.path.exists(d / "RDF.out_xyz")
