"""Host side of the library, no GPU needed: the C ABI exports what
include/gpuga.h declares, the set-up path (readers, neighbour list) agrees
bit for bit with the oracle, and malformed inputs fail with the reference's
messages (src/fitness.cu:163-314, src/ga.cu:214-316)."""
import os
import re
import subprocess

import numpy as np
import pytest

import gpuga_amd as G
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gpuga.h")).read()
    return sorted(set(re.findall(r"\b(gpuga_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", G.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (gpuga_[a-z_0-9]+)\b", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = G.lib()
    for n in names:
        assert hasattr(lib, n)


@pytest.mark.parametrize("name", ["si_diamond", "gaas", "si_type2", "nn2b", "hfo2"])
def test_neighbor_list_bit_exact_with_oracle(golden, oracle_case, name):
    NN, start, NL = G.inspect_neighbors(golden(name))
    a, b, c = oracle_case(name).neighbors()
    assert np.array_equal(NN, a) and np.array_equal(start, b) and np.array_equal(NL, c)


def test_synthetic_sets_round_trip_through_both_readers(tmp_path):
    from gpuga_amd import synth
    d = synth.tersoff_si(Nc=12, sizes="loguniform", seed=4)
    d = synth.label(d, O.Case(potential=dict(potential_type=1, cutoff=3.0,
                                             weights=(0.05, 0.15, 0.8), pmin=d["pmin"],
                                             pmax=d["pmax"]), arrays=d).find_force)
    synth.write_dir(d, str(tmp_path), ga={"maximum_generation": 5, "population_size": 20,
                                          "parent_number": 10, "mutation_rate": 0.2})
    case = O.Case(str(tmp_path))
    assert case.N == int(np.sum(d["Na"]))
    # float values survive the text round trip bit for bit
    assert np.array_equal(case.array("r", 3 * case.N), np.asarray(d["r"], np.float32))
    ncf = int(d["Nc_force"])  # force configurations carry no energy line
    assert np.array_equal(case.array("pe_ref", case.Nc)[ncf:], d["pe_ref"][ncf:])
    NN, start, NL = G.inspect_neighbors(str(tmp_path))
    a, b, c = case.neighbors()
    assert np.array_equal(NN, a) and np.array_equal(NL, c)
    ga = G.read_ga_in(str(tmp_path))
    assert ga["population_size"] == 20 and ga["seed"] == 12345678


def test_large_train_in_pieces(tmp_path):
    """A train.in of several MB is read in pieces on several threads; the
    list built from it equals the oracle's (which reads the file on its
    own), and a damaged token deep in the file still fails with the
    reference's message (the sequential reader reports it)."""
    from gpuga_amd import synth
    d = synth.tersoff_si(Nc=600, sizes="loguniform", seed=9)
    synth.write_dir(d, str(tmp_path), ga={"maximum_generation": 5, "population_size": 20,
                                          "parent_number": 10, "mutation_rate": 0.2})
    path = os.path.join(str(tmp_path), "train.in")
    assert os.path.getsize(path) > 2 << 20  # 32 pieces or more: several threads
    case = O.Case(str(tmp_path))
    NN, start, NL = G.inspect_neighbors(str(tmp_path))
    a, b, c = case.neighbors()
    assert np.array_equal(NN, a) and np.array_equal(start, b) and np.array_equal(NL, c)
    text = open(path).read()
    cut = text.rfind(" ", 0, len(text) * 3 // 4)
    with open(path, "w") as f:
        f.write(text[:cut] + " 1.0x" + text[cut:])
    with pytest.raises(G.GpugaError, match="reading error for train.in"):
        G.inspect_neighbors(str(tmp_path))


def write(path, name, text):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, name), "w") as f:
        f.write(text)


POT = "potential_type 1\ncutoff 3.0\nweight_force 0.05\nweight_energy 0.15\nweight_virial 0.8\n" + \
      "".join(f"p{k} 1 2\n" for k in range(9))


@pytest.mark.parametrize("train,msg", [
    ("1 1\n2\n", "Number of configurations should >= 2"),
    ("2 0\n2\n2\n", "Number of force configurations should >= 1"),
    ("2 2\n2\n2\n", "Number of potential configurations should >= 1"),
    ("2 1\n1\n2\n", "Number of atoms 1 should >= 2"),
    ("2 1\n2\n2\n10 0 0 0 10 0 0 0 10\n0 0 0 0 0 0 0\n", "reading error for train.in"),
])
def test_train_in_errors(tmp_path, train, msg):
    d = str(tmp_path)
    write(d, "potential.in", POT)
    write(d, "train.in", train)
    with pytest.raises(G.GpugaError, match=msg):
        G.inspect_neighbors(d)


@pytest.mark.parametrize("pot,msg", [
    ("potential_type 7\n", "unsupported potential type"),
    ("potential_type 1\ncutoff 3\nweight_force -1\nweight_energy 0\nweight_virial 0\n",
     "weight for force should >= 0"),
    ("potential_type 1\ncutoff 3\nweight_force 1\nweight_energy 1\nweight_virial 1\np 1\n",
     "reading error for potential.in"),
])
def test_potential_in_errors(tmp_path, pot, msg):
    d = str(tmp_path)
    write(d, "potential.in", pot)
    write(d, "train.in", "2 1\n2\n2\n")
    with pytest.raises(G.GpugaError, match=msg):
        G.inspect_neighbors(d)


@pytest.mark.parametrize("ga,msg", [
    ("generations 10\n", "'generations' is invalid keyword"),
    ("population_size 10\n", "population_size should >= 20"),
    ("population_size 30\nparent_number 11\n", "population_size - parent_number should be even"),
    ("mutation_rate 2\n", "mutation_rate should <= 1"),
    ("maximum_generation 0\n", "maximum_generation should be positive"),
    ("maximum_generation ten\n", "maximum_generation should be integer"),
    ("population_size 20 30\n", "population_size should have 1 parameter"),
])
def test_ga_in_errors(tmp_path, ga, msg):
    d = str(tmp_path)
    write(d, "ga.in", ga)
    with pytest.raises(G.GpugaError, match=msg):
        G.read_ga_in(d)


def test_ga_in_comments_defaults_and_seed(tmp_path):
    d = str(tmp_path)
    write(d, "ga.in", "# comment line\n\nmaximum_generation 7  # trailing\npopulation_size 256\n"
                      "parent_number 128\nseed 99\n")
    ga = G.read_ga_in(d)
    assert ga == {"maximum_generation": 7, "population_size": 256, "parent_number": 128,
                  "mutation_rate": pytest.approx(0.2), "seed": 99}
    og = O.lib()
    del og


@pytest.mark.parametrize("name", ["si_diamond", "gaas", "hfo2"])
def test_host_geometry_tables(golden, name):
    """The host geometry tables (the reference the device build is compared
    with bit for bit): bond vectors of reverse bonds are exact negatives
    (the minimum image is odd in IEEE arithmetic), lengths are the rounded
    norms, triplet cosines are symmetric bit for bit and equal the cosine of
    the bond vectors, and the gradient is orthogonal to r_ij (d cos / d r_ij
    . r_ij = 0); Tersoff records repeat the tables in their slots."""
    d = golden(name)
    bond, tri, rec = G.inspect_geometry(d)
    NN, start, NL = G.inspect_neighbors(d)
    N, B = len(NN), len(NL)
    bond = bond[:B].astype(np.float64)
    owner = np.repeat(np.arange(N), NN)
    # reverse bonds: keys i N + j are ascending over the list (atoms in
    # order, each list ascending), so the slot of (j, i) is a binary search
    keys = owner.astype(np.int64) * N + NL
    rev = np.searchsorted(keys, NL.astype(np.int64) * N + owner)
    assert np.array_equal(NL[rev], owner)
    assert np.array_equal(bond[rev, :3], -bond[:, :3])
    norm = np.sqrt((bond[:, :3] ** 2).sum(1))
    np.testing.assert_allclose(bond[:, 3], norm, rtol=2e-7)
    if name == "hfo2":
        assert len(rec) == 0
        return
    for i in range(0, N, max(1, N // 200)):
        nn, b0 = NN[i], start[i]
        off = int(np.sum(NN[:i].astype(np.int64) ** 2))
        T = tri[off:off + nn * nn].reshape(nn, nn, 4)
        u = bond[b0:b0 + nn, :3] / bond[b0:b0 + nn, 3:4]
        for s in range(nn):
            for t in range(nn):
                if s == t:
                    assert not T[s, t].any()
                    continue
                assert T[s, t, 0].view(np.uint32) == T[t, s, 0].view(np.uint32)
                assert T[s, t, 0] == pytest.approx(u[s] @ u[t], abs=1e-6)
                assert abs(T[s, t, 1:] @ bond[b0 + s, :3]) < 1e-5
        M = 4 if NN.max() <= 4 else (8 if NN.max() <= 8 else 16)
        RS = 1 + M + M * M + M // 2
        R = rec[i * RS:(i + 1) * RS]
        assert R[0].view(np.int32)[0] == nn and R[0].view(np.int32)[2] == b0
        assert np.array_equal(R[1:1 + nn], bond[b0:b0 + nn].astype(np.float32))
        assert np.array_equal(R[1 + M:1 + M + M * M].reshape(M, M, 4)[:nn, :nn], T)


def test_pruned_records_equal_records_of_the_shorter_list(tmp_path):
    """A 4.6 A cut-off (28 neighbours) with R2 at most 3.2 A: the population
    kernel's atom records hold only the bonds shorter than 3.2 A, and are
    then bit for bit
    the records of the same set under a 3.2 A cut-off (the same bonds in
    the same order, with their reverse-bond and first-bond indices)."""
    from gpuga_amd import synth
    d = synth.tersoff_si(Nc=6, reps=(3, 3, 3), seed=9)
    recs = {}
    for cutoff in (4.6, 3.2):
        d["cutoff"] = cutoff
        path = str(tmp_path / str(cutoff))
        synth.write_dir(d, path)
        NN, _, _ = G.inspect_neighbors(path)
        recs[cutoff] = (int(NN.max()), G.inspect_geometry(path)[2])
    assert recs[4.6][0] == 28 and recs[3.2][0] == 4
    a, b = recs[4.6][1], recs[3.2][1]
    assert len(a) == int(np.sum(d["Na"])) * (1 + 4 + 16 + 2)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
