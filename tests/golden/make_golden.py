#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden from the reference's own code.

Runs where the reference tree is mounted: oracle/_ref/libngia2_ref.so is the
reference's host headers (makedb.hpp, leidenParallel.hpp, saveResult.hpp)
compiled as they lie by `make -C oracle ref`.  Outputs are data only: input
FASTA files, the packed DB bytes the reference writes for them, Leiden
partitions for fixed edge lists, and digests of the result files.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as orc  # noqa: E402
import ngia2_amd as ng  # noqa: E402


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def edge_case_fasta(path):
    """Awkward but valid input: lowercase, unknown letters, empty and short
    records, multi-line sequences, blank lines, a name that does not start
    with '>', a 65,534-residue read (kept) and a 65,535-residue read
    (dropped, makedb.hpp:118), and equal lengths (stable order)."""
    rng = np.random.default_rng(2024)
    aa = b"ACDEFGHIKLMNPQRSTVWY"
    rnd = lambda L: bytes(aa[i] for i in rng.integers(20, size=L))
    recs = [
        b"first line without a marker\nACDEFGHIKLMNPQRSTVWY\n",
        b">empty\n\n",
        b">no_residue_lines\n",
        b">short1\nA\n",
        b">short7\nACDEFGH\n",
        b">short8\nACDEFGHI\n",
        b">lower\nacdefghiklmnpqrstvwy\n",
        b">unknown\nXXBJOZ*-ACDE\n",
        b">multi line record\n" + rnd(40) + b"\n" + rnd(35) + b"\n\n" + rnd(50) + b"\n",
        b">dna\nACGTACGTTTGACCA\n",
        b">eq_a\n" + rnd(33) + b"\n",
        b">eq_b\n" + rnd(33) + b"\n",
        b">eq_c\n" + rnd(33) + b"\n",
        b">long_kept\n" + rnd(65534) + b"\n",
        b">long_dropped\n" + rnd(65535) + b"\n",
        b">tail\n" + rnd(64) + b"\n",
    ]
    for k in range(20):
        recs.append(b">fam_%d\n" % k + rnd(int(rng.integers(20, 400))) + b"\n")
    with open(path, "wb") as fh:
        fh.write(b"".join(recs))


def random_graph(rng, n, m):
    u = rng.integers(0, n, size=m)
    v = rng.integers(0, n, size=m)
    e = np.zeros(m, dtype=orc.EDGE_DTYPE)
    e["u"], e["v"], e["w"] = u, v, 1.0
    return e


def main():
    if orc.ref() is None:
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    ref = orc.ref()
    assert ref is not None, "reference build missing"
    out = {}

    # 1. packed DB bytes for the edge-case FASTA
    fa = os.path.join(HERE, "edge_cases.fa")
    edge_case_fasta(fa)
    db = os.path.join(HERE, "edge_cases.packed")
    assert ref.ref_makedb(fa.encode(), db.encode()) == 0
    out["edge_cases"] = {"fasta_sha256": sha(fa), "packed_sha256": sha(db),
                         "packed_size": os.path.getsize(db)}

    # 2. a config-1 style set (10 families x 50 members)
    fa2 = os.path.join(HERE, "c1_small.fa")
    ng.gen_fasta(ng.gen_config(1, scale=0.05, seed=5), fa2)
    db2 = os.path.join(HERE, "_c1_small.packed")
    assert ref.ref_makedb(fa2.encode(), db2.encode()) == 0
    out["c1_small"] = {"fasta_sha256": sha(fa2), "packed_sha256": sha(db2),
                       "packed_size": os.path.getsize(db2)}
    d = orc.read_db_numpy(db2)
    np.save(os.path.join(HERE, "c1_small_rep.npy"), d["rep"])

    # 3. Leiden partitions and result files of the reference on fixed edge lists
    parts = {}
    rng = np.random.default_rng(77)
    graphs = {}
    for identity in (30, 50, 70, 90):
        e, _ = orc.build_graph(d["packed"], d["word_off"], d["lengths"], d["rep"], 64, identity)
        graphs["c1_i%d" % identity] = (e, d["n"], identity)
    graphs["rand_sparse"] = (random_graph(rng, 2000, 3000), 2000, 60)
    graphs["rand_dense"] = (random_graph(rng, 300, 6000), 300, 50)
    graphs["rand_selfloops"] = (random_graph(rng, 50, 400), 50, 80)
    cliques = []
    for c in range(40):  # 40 cliques of 5 plus a few bridges
        base = 5 * c
        cliques += [(base + a, base + b) for a in range(5) for b in range(a + 1, 5)]
    cliques += [(5 * c + 4, 5 * c + 5) for c in range(0, 39, 3)]
    ce = np.zeros(len(cliques), dtype=orc.EDGE_DTYPE)
    ce["u"] = [a for a, _ in cliques]
    ce["v"] = [b for _, b in cliques]
    ce["w"] = 1.0
    graphs["cliques"] = (ce, 200, 90)
    arrays = {}
    for name, (e, n, identity) in graphs.items():
        res = orc.resolution(identity)
        part = orc.ref_leiden(e, n, res)
        arrays["edges_" + name] = e
        arrays["part_" + name] = part
        parts[name] = {"n": int(n), "identity": identity, "resolution": float(res),
                       "clusters": int(part.max()) + 1 if n else 0}
        if name.startswith("c1_"):
            rp = os.path.join(HERE, "_result.txt")
            ref.ref_save(part.ctypes.data_as(orc.C.c_void_p), n, db2.encode(), rp.encode())
            parts[name]["result_sha256"] = sha(rp)
            os.remove(rp)
    np.savez_compressed(os.path.join(HERE, "leiden.npz"), **arrays)
    out["leiden"] = parts
    os.remove(db2)
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1)[:2000])


if __name__ == "__main__":
    main()
