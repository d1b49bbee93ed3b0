"""Host side of the product without a GPU: the C-ABI library loads and exports
its whole header, the generator, the packed-DB reader, the FASTA checks, the
host Leiden and result writer, and the command line's argument handling and
exit codes.  CPU only."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import ngia2_amd as ng
import oracle_lib as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NGIA2 = os.path.join(ROOT, "ngia2_amd", "bin", "ngia2")


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ngia2.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ngia_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import ctypes
    lib = ctypes.CDLL(ng.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
    from ngia2_amd import _lib
    assert set(names) == set(_lib.PROTOTYPES), set(names) ^ set(_lib.PROTOTYPES)


def test_generator_deterministic(tmp_path):
    p = ng.gen_config(1, scale=0.05, seed=9)
    a, b = str(tmp_path / "a.fa"), str(tmp_path / "b.fa")
    ng.gen_fasta(p, a)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    code = ("import sys; sys.path.insert(0, %r); import ngia2_amd as ng; "
            "ng.gen_fasta(ng.gen_config(1, scale=0.05, seed=9), %r)" % (ROOT, b))
    subprocess.run([sys.executable, "-c", code], check=True, env=env)
    assert open(a, "rb").read() == open(b, "rb").read()
    c = str(tmp_path / "c.fa")
    ng.gen_fasta(ng.gen_config(1, scale=0.05, seed=10), c)
    assert open(a, "rb").read() != open(c, "rb").read()


@pytest.mark.parametrize("config", [1, 2, 3, 4])
def test_gen_packed_equals_makedb_records(tmp_path, config):
    p = ng.gen_config(config, scale=0.002 if config != 1 else 0.05, seed=3)
    fa, db = str(tmp_path / "g.fa"), str(tmp_path / "g.packed")
    n = ng.gen_fasta(p, fa)
    assert orc.makedb(fa, db) == 0
    d = orc.read_db_numpy(db)
    g = ng.gen_packed(p)
    assert g.n == d["n"] == n
    np.testing.assert_array_equal(g.lengths, d["lengths"])
    np.testing.assert_array_equal(g.packed, d["packed"])
    np.testing.assert_array_equal(g.word_off, d["word_off"])
    np.testing.assert_array_equal(g.pack_off, d["pack_off"])
    np.testing.assert_array_equal(g.text_off, d["text_off"])


def test_config_shapes():
    p = ng.gen_config(2)
    assert p.n_families * p.members == 1_000_000 and p.len_a == 350 and p.len_max == 5000
    p = ng.gen_config(3)
    assert p.n_families * p.members == 2_000_000 and p.dna == 1 and p.len_normal == 1
    p = ng.gen_config(4)
    assert p.tail_alpha == 1.5 and p.tail_max == 35000
    g = ng.gen_packed(ng.gen_config(4, scale=0.01, seed=2))
    assert g.lengths.max() > 1000  # the Pareto tail is present
    d = ng.gen_packed(ng.gen_config(3, scale=0.001, seed=2))
    codes = set()
    for i in range(min(d.n, 50)):
        rec = d.packed[d.word_off[i]:]
        L = int(rec[0])
        for j in range(L):
            c = 0
            for e in range(4):
                c |= ((int(rec[1 + j // 32 * 4 + e]) >> (j % 32)) & 1) << e
            codes.add(c)
    assert codes <= {1, 2, 6, 15}


def test_read_db_matches_numpy_parser(tmp_path):
    fa, db = str(tmp_path / "r.fa"), str(tmp_path / "r.packed")
    ng.gen_fasta(ng.gen_config(1, scale=0.02, seed=4), fa)
    assert orc.makedb(fa, db) == 0
    d = orc.read_db_numpy(db, bands=16)
    p = ng.read_db(db, bands=16)
    np.testing.assert_array_equal(p.rep_rows, d["rep"])
    np.testing.assert_array_equal(p.packed, d["packed"])
    with pytest.raises(ng.NgiaError):
        ng.read_db(db, bands=65)
    trunc = str(tmp_path / "t.packed")
    open(trunc, "wb").write(open(db, "rb").read()[:1000])
    with pytest.raises(ng.NgiaError):
        ng.read_db(trunc)


def test_fasta_check(tmp_path):
    ok = tmp_path / "ok.fa"
    ok.write_bytes(b">a\nACD\n")
    assert ng.fasta_check(str(ok)) is None
    crlf = tmp_path / "crlf.fa"
    crlf.write_bytes(b">a\r\nACD\r\n")
    assert "\\r" in ng.fasta_check(str(crlf))
    nonl = tmp_path / "nonl.fa"
    nonl.write_bytes(b">a\nACD")
    assert "end with" in ng.fasta_check(str(nonl))
    assert "does not exist" in ng.fasta_check(str(tmp_path / "missing.fa"))
    empty = tmp_path / "empty.fa"
    empty.write_bytes(b"")
    assert ng.fasta_check(str(empty)) is None


def test_host_leiden_random_graphs_match_oracle():
    rng = np.random.default_rng(31)
    for trial in range(25):
        n = int(rng.integers(2, 600))
        m = int(rng.integers(0, 5 * n))
        e = np.zeros(m, dtype=orc.EDGE_DTYPE)
        e["u"] = rng.integers(-2, n + 2, m)  # out-of-range ids are dropped
        e["v"] = rng.integers(0, n, m)
        e["w"] = rng.random(m).astype(np.float32)  # weights are ignored
        res = orc.resolution(int(rng.integers(1, 100)))
        np.testing.assert_array_equal(ng.leiden(e, n, res), orc.leiden(e, n, res), str(trial))


def test_host_save_result_matches_oracle(tmp_path):
    fa, db = str(tmp_path / "s.fa"), str(tmp_path / "s.packed")
    ng.gen_fasta(ng.gen_config(1, scale=0.02, seed=8), fa)
    assert orc.makedb(fa, db) == 0
    n = orc.read_db_numpy(db)["n"]
    rng = np.random.default_rng(2)
    for cl in (np.arange(n), np.zeros(n), rng.integers(0, 17, n), rng.integers(0, 3 * n, n)):
        a, b = str(tmp_path / "a.txt"), str(tmp_path / "b.txt")
        orc.save_result(cl, db, a)
        ng.save_result(cl, db, b)
        assert open(a, "rb").read() == open(b, "rb").read()


def run_cli(args, name="ngia2", cwd=None):
    exe = os.path.join(os.path.dirname(NGIA2), name)
    return subprocess.run([exe] + args, capture_output=True, text=True, cwd=cwd, timeout=120)


def test_cli_usage_exit_codes(tmp_path):
    # usage and check failures exit 0 (main.cu:45, 56-60, 88-93, 159-161)
    r = run_cli([])
    assert r.returncode == 0 and "usage" in r.stdout
    r = run_cli(["-m", "bogus"])
    assert r.returncode == 0 and "Usage" in r.stdout
    r = run_cli(["-m", "makedb", "-f", str(tmp_path / "missing.fa"), "-p", str(tmp_path / "x")])
    assert r.returncode == 0 and "does not exist" in r.stdout
    crlf = tmp_path / "crlf.fa"
    crlf.write_bytes(b">a\r\nACD\n")
    r = run_cli(["-f", str(crlf), "-p", str(tmp_path / "x")], name="makeDB")
    assert r.returncode == 0 and "\\r" in r.stdout
    r = run_cli(["-m", "clustering", "-p", str(tmp_path / "none"), "-i", "150"])
    assert r.returncode == 0 and "identity should be 1-99" in r.stdout
    r = run_cli(["-i", "0", "-p", str(tmp_path / "none")], name="cluster")
    assert r.returncode == 0 and "identity should be 1-99" in r.stdout
    r = run_cli(["-m", "fast", "-p", str(tmp_path / "none"), "-b", "65"])
    assert r.returncode == 0 and "bands should be 1-64" in r.stdout
    r = run_cli(["-m", "clustering", "-p", str(tmp_path / "none"), "-i", "90"])
    assert r.returncode == 0 and "does not exists" in r.stdout
    empty = tmp_path / "zero.packed"
    empty.write_bytes(b"\x00\x00\x00\x00")
    r = run_cli(["-m", "clustering", "-p", str(empty)])
    assert r.returncode == 0 and "contains 0 sequences" in r.stdout


def test_cli_parser_quirks(tmp_path):
    # last occurrence wins; a trailing flag without a value is ignored
    r = run_cli(["-m", "clustering", "-i", "50", "-i", "150", "-p", str(tmp_path / "none")])
    assert "identity should be 1-99" in r.stdout
    r = run_cli(["-m", "clustering", "-i", "150", "-i", "50", "-p", str(tmp_path / "none"), "-i"])
    assert "identity should be 1-99" not in r.stdout and "does not exists" in r.stdout


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_plan_blocks():
    """The block cut of the out-of-HBM graph: consecutive records, a new block
    when the next record would pass the word or record limit, an oversized
    record alone, every record in exactly one block."""
    rng = np.random.default_rng(5)
    lengths = rng.integers(0, 3000, size=500).astype(np.uint32)
    words = 1 + 4 * ((lengths.astype(np.int64) + 31) // 32)
    for block_words, block_records in ((1, 0), (400, 0), (5000, 0), (5000, 7), (10 ** 9, 0),
                                       (10 ** 9, 100)):
        b = ng.plan_blocks(lengths, block_words, block_records)
        assert b[0] == 0 and b[-1] == len(lengths) and np.all(np.diff(b.astype(np.int64)) > 0)
        for k in range(len(b) - 1):
            w = words[b[k]:b[k + 1]]
            # within the limits, or a single record that alone exceeds them
            assert len(w) == 1 or w.sum() <= block_words
            assert block_records == 0 or len(w) <= block_records
            if k + 1 < len(b) - 1:  # greedy: the next record did not fit
                nxt = words[b[k + 1]]
                assert w.sum() + nxt > block_words or (block_records and len(w) == block_records)
    assert len(ng.plan_blocks(lengths, 1)) == len(lengths) + 1
    assert len(ng.plan_blocks(lengths, 10 ** 9)) == 2
    assert list(ng.plan_blocks(np.zeros(0, np.uint32), 10)) == [0]
    with pytest.raises(ng.NgiaError):
        ng.plan_blocks(lengths, 0)
