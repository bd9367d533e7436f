"""HIP kernels through the C ABI against the reference's known answers and the oracle."""
import numpy as np
import pytest

from helpers import load_jsonl
from hero_amd import gpu, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("aligner", [1, 0], ids=["global", "affine"])
def test_band_align_batch_known_answers(aligner):
    cases = [c for c in load_jsonl("dp_cases.jsonl") if c["aligner"] == aligner]
    with gpu.Context(aligner=aligner) as ctx:
        res = ctx.band_align([(c["S"], c["Q"], c["band"], c["flag"]) for c in cases])
    bad = [(c, r) for c, r in zip(cases, res) if r != (c["cigar"], c["out_band"])]
    assert not bad, bad[:3]


def test_hash_kmers_known_answers():
    cases = load_jsonl("hash_cases.jsonl")
    with gpu.Context() as ctx:
        by_k = {}
        for c in cases:
            by_k.setdefault((len(c["kmer"]), c["P"]), []).append(c)
        for (k, P), cs in by_k.items():
            out = ctx.hash_kmers([c["kmer"] for c in cs], P)
            assert [int(x) for x in out] == [c["hash"] for c in cs]


@pytest.mark.parametrize("k", [21, 31, 32, 39, 41])
def test_lookup_buckets_match_oracle(k):
    rs = synth.generate(seed=k, genome_len=5000, len_min=100, coverage=10.0, sub_5p=0.01)
    reads = [rs.read(i) for i in range(rs.n)]
    with gpu.Context(k=k, min_overlap=40) as ctx:
        ctx.set_queries(rs)
        ctx.build_index()
        wins = [reads[i][p:p + k] for i in range(0, rs.n, 7) for p in (0, 5, len(reads[i]) - k)]
        got = ctx.lookup(wins)
    cfg = O.make_cfg(k=k, call_consensus=1, min_overlap=40)
    for w, g in zip(wins, got):
        # bucket = every good query's key occurrence equal to w, in (innerId, pos) order
        exp = []
        for q, s in enumerate(reads):
            if len(s) <= 40:
                continue
            L = len(s)
            pos = list(range(0, L - k + 1, k))
            if pos[-1] + k < L:
                pos.append(L - k)
            exp += [(p << 40) | q for p in pos if s[p:p + k] == w]
        assert [int(x) for x in g] == exp
