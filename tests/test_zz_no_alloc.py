"""ngia_align_noalloc (the kernel_dynamic contract on caller-owned device
memory, row scratch in a static device array) against ngia_align on the same
jobs, through the C-ABI: two calls in a row, the second on a smaller input
with a max_len below its longest rep (those jobs take the full-row tier).
Device memory for the inputs and outputs comes from hipMalloc through ctypes;
the library allocates nothing in the call."""
import ctypes as C

import numpy as np
import pytest

import ngia2_amd as ng
from ngia2_amd import _lib
import oracle_lib as orc
from conftest import AMINO, mutate

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


class DevArray:
    """A device copy of a numpy array (hipMalloc / hipMemcpy through ctypes)."""

    hip = None

    def __init__(self, host):
        if DevArray.hip is None:
            DevArray.hip = C.CDLL("libamdhip64.so")
        self.host = np.ascontiguousarray(host)
        self.nbytes = max(self.host.nbytes, 4)
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0
        assert self.hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p),
                                  C.c_size_t(self.host.nbytes), 1) == 0  # host to device

    def get(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr,
                                  C.c_size_t(self.host.nbytes), 2) == 0  # device to host
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def records(rng):
    """Families of sequences of 20 to 9,000 residues, longest first."""
    seqs = []
    for L in (9000, 4100, 2600, 1500, 900, 500, 300, 120, 33, 20):
        anc = bytes(AMINO[i] for i in rng.integers(20, size=L))
        for _ in range(5):
            seqs.append(mutate(rng, anc, float(rng.uniform(0.0, 0.2)), 0.01))
    seqs.sort(key=lambda s: -len(s))
    packed, off = orc.pack_many(seqs)
    return packed, off, np.array([len(s) for s in seqs], dtype=np.uint32)


def jobs_of(rng, n, lengths, count, identity):
    """Random (earlier, later) pairs plus every pair within five neighbours."""
    a = rng.integers(0, n, size=count)
    b = rng.integers(0, n, size=count)
    near = [(i, j) for i in range(n) for j in range(i + 1, min(n, i + 6))]
    jobs = np.zeros((count + len(near), 3), dtype=np.uint32)
    jobs[:count, 0] = np.minimum(a, b)
    jobs[:count, 1] = np.maximum(a, b)
    jobs[count:, :2] = np.array(near, dtype=np.uint32)
    return jobs


def noalloc(ctx, packed, off, jobs, thr, max_len):
    L = _lib.lib()
    dp, do, dj = DevArray(packed), DevArray(off), DevArray(jobs)
    try:
        rc = L.ngia_align_noalloc(ctx.h, dp.ptr, do.ptr, dj.ptr, len(jobs), C.c_float(thr),
                                  max_len, None)
        assert rc == 0, L.ngia_last_error(ctx.h)
        ctx.sync()
        return dj.get()
    finally:
        for x in (dp, do, dj):
            x.free()


def sub_database(packed, off, lengths, keep):
    """The records `keep` (ascending) repacked as a database of their own."""
    recs = [packed[int(off[i]):int(off[i]) + 1 + 4 * ((int(lengths[i]) + 31) // 32)] for i in keep]
    o = np.zeros(len(recs), dtype=np.uint64)
    o[1:] = np.cumsum([len(r) for r in recs[:-1]])
    return np.concatenate(recs).astype(np.uint32), o, lengths[keep]


def test_noalloc_matches_align_twice(ctx):
    """Call 1: every record (20 to ~9,000 residues), max_len the longest.
    Call 2, smaller: the records of at most 900 residues with max_len 400, so
    that the longer reps run on the full-row tier."""
    rng = np.random.default_rng(41)
    packed, off, lengths = records(rng)
    p2, o2, l2 = sub_database(packed, off, lengths, np.nonzero(lengths <= 900)[0])
    assert int(l2.max()) > 400
    cases = [(packed, off, jobs_of(rng, len(lengths), lengths, 3000, 70), 70, int(lengths.max())),
             (p2, o2, jobs_of(rng, len(l2), l2, 500, 90), 90, 400)]
    for k, (pk, of, jobs, identity, max_len) in enumerate(cases):
        thr = float(np.float32(identity) / np.float32(100))
        say("noalloc call", k + 1, "jobs", len(jobs), "identity", identity, "max_len", max_len)
        want = ctx.align(pk, of, jobs, thr)
        got = noalloc(ctx, pk, of, jobs, thr, max_len)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(got[:, 2], orc.align_jobs(pk, of, jobs, thr)[:, 2])
        assert (got[:, 2] > 0).sum() > 20


def test_noalloc_rejects_oversized_max_len(ctx):
    L = _lib.lib()
    rc = L.ngia_align_noalloc(ctx.h, None, None, None, 0, C.c_float(0.9), 70000, None)
    assert rc == ng.NGIA_ERR_ARG
    assert b"65534" in L.ngia_last_error(ctx.h)
