import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def ctx():
    import ngia2_amd
    c = ngia2_amd.Context(0)
    yield c
    c.close()


AMINO = b"ACDEFGHIKLMNPQRSTVWY"


def mutate(rng, seq: bytes, sub: float, indel: float, alpha=AMINO) -> bytes:
    out = bytearray()
    for ch in seq:
        if rng.random() < indel:
            if rng.random() < 0.5:
                continue
            out.append(alpha[rng.integers(len(alpha))])
        if rng.random() < sub:
            ch = alpha[rng.integers(len(alpha))]
        out.append(ch)
    return bytes(out)


def family_fasta(path, rng, lengths, members, sub, indel, alpha=AMINO):
    """Write families: per ancestor length, `members` mutated copies."""
    recs = []
    for f, L in enumerate(lengths):
        anc = bytes(alpha[i] for i in rng.integers(len(alpha), size=L))
        for m in range(members):
            recs.append((b">fam%d_m%d" % (f, m), mutate(rng, anc, sub, indel, alpha)))
    order = rng.permutation(len(recs))
    with open(path, "wb") as fh:
        for k in order:
            name, s = recs[k]
            fh.write(name + b"\n" + s + b"\n")
    return len(recs)
