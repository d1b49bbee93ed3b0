"""The oracle and the host side of the product against fixtures the reference's own
code wrote (tests/golden/make_golden.py, run on the reference's host headers).
CPU only."""
import hashlib
import json
import os

import numpy as np
import pytest

import ngia2_amd as ng
import oracle_lib as orc

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GOLD = json.load(open(os.path.join(G, "golden.json")))
LEIDEN = np.load(os.path.join(G, "leiden.npz"))


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def test_fixture_inputs_unchanged():
    assert sha(os.path.join(G, "edge_cases.fa")) == GOLD["edge_cases"]["fasta_sha256"]
    assert sha(os.path.join(G, "c1_small.fa")) == GOLD["c1_small"]["fasta_sha256"]


def test_oracle_makedb_edge_cases(tmp_path):
    out = str(tmp_path / "o.packed")
    assert orc.makedb(os.path.join(G, "edge_cases.fa"), out) == 0
    assert open(out, "rb").read() == open(os.path.join(G, "edge_cases.packed"), "rb").read()


def test_oracle_makedb_c1_small(tmp_path):
    out = str(tmp_path / "o.packed")
    assert orc.makedb(os.path.join(G, "c1_small.fa"), out) == 0
    assert os.path.getsize(out) == GOLD["c1_small"]["packed_size"]
    assert sha(out) == GOLD["c1_small"]["packed_sha256"]
    d = orc.read_db_numpy(out)
    np.testing.assert_array_equal(d["rep"], np.load(os.path.join(G, "c1_small_rep.npy")))


def test_oracle_sketch_bucket_reproduce_reference_rows():
    """Sketch + bucket of the reference DB's own packed records give its rep rows."""
    d = orc.read_db_numpy(os.path.join(G, "edge_cases.packed"))
    sig = orc.sketch_all(d["packed"], d["word_off"])
    np.testing.assert_array_equal(orc.bucket(sig), d["rep"])


def test_product_reader_on_reference_db():
    d = orc.read_db_numpy(os.path.join(G, "edge_cases.packed"))
    p = ng.read_db(os.path.join(G, "edge_cases.packed"))
    assert p.n == d["n"]
    np.testing.assert_array_equal(p.lengths, d["lengths"])
    np.testing.assert_array_equal(p.packed, d["packed"])
    np.testing.assert_array_equal(p.word_off, d["word_off"])
    np.testing.assert_array_equal(p.rep_rows, d["rep"])


@pytest.mark.parametrize("name", sorted(GOLD["leiden"]))
def test_leiden_oracle_and_host_match_reference(name):
    info = GOLD["leiden"][name]
    e = LEIDEN["edges_" + name]
    want = LEIDEN["part_" + name]
    assert orc.resolution(info["identity"]) == np.float32(info["resolution"])
    assert ng.resolution(info["identity"]) == np.float32(info["resolution"])
    np.testing.assert_array_equal(orc.leiden(e, info["n"], info["resolution"]), want)
    np.testing.assert_array_equal(ng.leiden(e, info["n"], info["resolution"]), want)


@pytest.mark.parametrize("name", ["c1_i30", "c1_i50", "c1_i70", "c1_i90"])
def test_result_file_matches_reference(tmp_path, name):
    info = GOLD["leiden"][name]
    db = str(tmp_path / "c1.packed")
    assert orc.makedb(os.path.join(G, "c1_small.fa"), db) == 0
    part = LEIDEN["part_" + name]
    a, b = str(tmp_path / "a.txt"), str(tmp_path / "b.txt")
    assert orc.save_result(part, db, a) == 0
    ng.save_result(part, db, b)
    assert sha(a) == info["result_sha256"]
    assert sha(b) == info["result_sha256"]


def test_graph_on_reference_rows_is_what_leiden_saw():
    """The edge lists in the fixtures are the oracle's graphs of the reference DB."""
    d = orc.read_db_numpy(os.path.join(G, "edge_cases.packed"))
    for identity in (30, 90):
        e, _ = orc.build_graph(d["packed"], d["word_off"], d["lengths"], d["rep"], 64, identity)
        assert np.all(e["u"] < e["v"])
        key = e["u"].astype(np.int64) * (1 << 32) + e["v"]
        assert np.all(np.diff(key) > 0)


@pytest.mark.skipif(not os.path.exists(orc.REF_SO),
                    reason="reference build absent (it is built only where the reference tree is)")
def test_live_reference_makedb_and_leiden(tmp_path):
    """Where the reference build is present: fresh random inputs, three-way agreement."""
    ref = orc.ref()
    rng = np.random.default_rng(99)
    for trial in range(3):
        fa = str(tmp_path / ("r%d.fa" % trial))
        p = ng.gen_config(1, scale=0.02 + 0.01 * trial, seed=100 + trial)
        ng.gen_fasta(p, fa)
        a, b = str(tmp_path / "a.packed"), str(tmp_path / "b.packed")
        assert ref.ref_makedb(fa.encode(), a.encode()) == 0
        assert orc.makedb(fa, b) == 0
        assert open(a, "rb").read() == open(b, "rb").read()
        n = int(rng.integers(50, 400))
        m = int(rng.integers(n, 6 * n))
        e = np.zeros(m, dtype=orc.EDGE_DTYPE)
        e["u"] = rng.integers(0, n, m)
        e["v"] = rng.integers(0, n, m)
        e["w"] = 1.0
        res = float(np.float32(rng.integers(1, 100)) / np.float32(100))
        want = orc.ref_leiden(e, n, res)
        np.testing.assert_array_equal(orc.leiden(e, n, res), want)
        np.testing.assert_array_equal(ng.leiden(e, n, res), want)
