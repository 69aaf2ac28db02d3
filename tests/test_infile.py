"""The YAML front end (include/chmc/read_infile.hh) on the reference's own input file
(code/builds/test.yaml, kept as tests/golden/reference_test.yaml) and on edge cases: values,
nesting, the missing-key error of check_and_assign_value (read_infile.cpp:18-25) and the int-through-
stod rule (read_infile.cpp:47).  Host-only: bin/infile_check makes no device call."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "chemhmc_amd", "bin", "infile_check")
GOLD = os.path.join(ROOT, "tests", "golden", "reference_test.yaml")


def run(path):
    assert os.path.exists(BIN), "build first: make -C chemhmc_amd all"
    r = subprocess.run([BIN, "-i", path], capture_output=True, text=True, timeout=30)
    vals = {}
    for line in r.stdout.splitlines():
        k, _, v = line.partition(" ")
        vals[k] = v
    return r, vals


def test_reference_input_file():
    r, v = run(GOLD)
    assert r.returncode == 0, r.stderr
    assert v["L"] == "3 3 3" and v["seed"] == "123" and v["StartCondition"] == "read"
    assert v["start_configuration_file"] == "out_xyz.txt" and v["output_file"] == "out_xyz.txt"
    assert v["append"] == "1" and v["name_RDF"] == "RDF.out_xyz"
    assert v["N"] == "10" and v["mass"] == "0.1" and v["beta"] == "0.5" and v["cutoff"] == "0.5"
    assert v["eps"] == "0.1" and v["sigma"] == "0.1" and v["name_xyz"] == "Ar" and v["algorithm"] == "all_neighbour"
    assert v["RDF"] == "1 1" and v["integrator"] == "OMF2 0.1 5"
    assert v["Ntrajectories"] == "4" and v["thermalization_steps"] == "0" and v["save_every"] == "1"
    assert v["randomize_trajectory"] == "absent"


def _variant(tmp_path, edit):
    text = open(GOLD).read()
    p = tmp_path / "in.yaml"
    p.write_text(edit(text))
    return run(str(p))


def test_int_through_stod_and_quotes(tmp_path):
    r, v = _variant(tmp_path, lambda t: t.replace("N: 10", "N: 1e1").replace("name_xyz: Ar", 'name_xyz: "Ar"'))
    assert r.returncode == 0, r.stderr
    assert v["N"] == "10" and v["name_xyz"] == "Ar"


def test_missing_key_is_an_error(tmp_path):
    r, _ = _variant(tmp_path, lambda t: t.replace("  Lz: 3\n", ""))
    assert r.returncode != 0
    assert "tag Lz is not present in the input file" in r.stdout
    r, _ = _variant(tmp_path, lambda t: t.replace("seed: 123", "seed: abc"))
    assert r.returncode != 0 and "impossible to read tag seed" in r.stdout


def test_append_rules(tmp_path):
    r, _ = _variant(tmp_path, lambda t: t.replace("StartCondition: read", "StartCondition: hot"))
    assert r.returncode != 0 and "start condition must be read" in r.stderr
    r, _ = _variant(tmp_path, lambda t: t.replace("rng_device_state: rng_device", "rng_device_state: rng_host"))
    assert r.returncode != 0 and "rng_device_state must be different" in r.stderr
