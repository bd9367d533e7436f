"""The oracle (oracle/hero_oracle.c) against the reference's own outputs (tests/golden/)."""
import numpy as np
import pytest

from helpers import fasta_text, flags_to_opts, load_e2e, load_jsonl, load_mapping, read_digests, sha256
from hero_amd import synth
from oracle import oracle as O


def test_dp_known_answers():
    bad = []
    for c in load_jsonl("dp_cases.jsonl"):
        cfg = O.make_cfg(aligner=c["aligner"])
        cig, band, _ = O.band_align(cfg, c["S"], c["Q"], c["band"], c["flag"])
        if cig != c["cigar"] or band != c["out_band"]:
            bad.append((c, cig, band))
    assert not bad, bad[:3]


def test_hash_known_answers():
    for c in load_jsonl("hash_cases.jsonl"):
        assert O.hash_kmer(c["kmer"], c["P"]) == c["hash"], c


def test_hash_closed_form():
    # k <= 32: the 2-bit string after a leading 1 (A0 C1 T2 G3), modulo P; k = 32 loses the 1.
    assert O.hash_kmer("A", 1_000_003) == 4
    assert O.hash_kmer("G", 1_000_003) == 7
    assert O.hash_kmer("ACTG", 1 << 40) == (1 << 8) | 0b00011011
    assert O.hash_kmer("C" * 32, (1 << 63) + 11) == int("01" * 32, 2)


def test_table_size_is_prime_above_factor_n():
    for n in (1, 10, 29703, 1666667):
        p = O.table_size(n, 32)
        assert p > 32 * n
        assert all(p % d for d in range(2, min(int(p ** 0.5) + 1, 20000)))


def test_record_scoring_known_answers():
    for c in load_jsonl("record_cases.jsonl"):
        cfg = O.make_cfg(mismatch_rate=c["mis_rate"], indel_rate=c["indel_rate"])
        out, merged = O.record_score(cfg, [tuple(x) for x in c["ops"]], c["cap_mis"], c["cap_indel"])
        assert merged == c["merged"], c
        assert (out[0], out[1], out[2]) == (c["nmis"], c["nindel"], c["length"]), c
        assert (int(out[3]), int(out[4])) == (c["good_caps"], c["good_rates"]), c
        assert out[5] == c["rate_sum"], c


def test_consensus_known_answers():
    for c in load_jsonl("consensus_cases.jsonl"):
        cfg = O.make_cfg(consensus_freq=c["thr"])
        assert O.consensus_read(cfg, c["seq"], np.array(c["tally"], dtype=np.uint32)) == c["out"], c


@pytest.mark.parametrize("case", load_e2e(), ids=lambda c: c["name"])
def test_end_to_end_matches_reference(case):
    rs = synth.generate(**case["synth"])
    assert rs.n == case["n_reads"]
    opts = flags_to_opts(case["flags"])
    r = O.run(O.make_cfg(call_consensus=1, **opts), rs, threads=4)
    assert read_digests(r["corrected"], rs) == case["read_digests"]
    assert sha256(fasta_text(r["corrected"], rs)) == case["fasta_sha256"]
    a = O.run(O.make_cfg(call_consensus=0, **opts), rs, threads=4, want_aln=True)
    assert sha256(a["aln"].encode()) == case["aln_sha256"]


@pytest.mark.parametrize("case", load_mapping(), ids=lambda c: c["name"])
def test_mapping_histogram_matches_reference(case):
    """--mapping: per subject the best record covering it (chooseQueryWithBestAlignment, including
    the keep-all case of exactly one covering record), then the coverage histogram; each of the
    selection rules changes these outputs when altered (checked by planting them in the oracle)."""
    rs = synth.generate(**case["synth"])
    assert rs.n == case["n_reads"]
    opts = flags_to_opts(case["flags"])
    r = O.run(O.make_cfg(call_consensus=1, mapping=1, **opts), rs, threads=4, want_tally=True)
    assert O.coverage_text(rs, r["tally"]) == case["coverage_text"]


def test_seed_search_brute_force():
    # seeds of one subject strand == brute-force string search of every query key
    rs = synth.generate(seed=7, genome_len=3000, len_min=60, coverage=8.0, sub_5p=0.01)
    k = 21
    cfg = O.make_cfg(k=k, call_consensus=1, min_overlap=30)
    reads = [rs.read(i) for i in range(rs.n)]
    comp = str.maketrans("ACGT", "TGCA")
    for s in range(0, rs.n, 37):
        for rev in (0, 1):
            subj = reads[s][::-1].translate(comp) if rev else reads[s]
            got = O.seeds(cfg, rs, reads[s], s, rev)
            exp = []
            for q, qs in enumerate(reads):
                if q == s or not (len(qs) > 30):
                    continue
                L = len(qs)
                pos = list(range(0, L - k + 1, k))
                if pos[-1] + k < L:
                    pos.append(L - k)
                for pq in pos:
                    key = qs[pq:pq + k]
                    for ps in range(len(subj) - k + 1):
                        if subj[ps:ps + k] == key:
                            exp.append((q, ps, pq))
            exp.sort(key=lambda x: (x[0], x[2], x[1]))
            assert [tuple(map(int, x)) for x in got] == exp


@pytest.mark.parametrize("name", ["c1_full_global_i2", "c1_full_global_i3"])
def test_oracle_at_c1_full_size_matches_reference(name):
    """The oracle on the whole 29,703-read C1 set reproduces the reference's corrected FASTA."""
    from helpers import block_digests, load_scale
    case = load_scale()[name]
    rs = synth.generate(**case["synth"])
    assert rs.n == case["n_reads"]
    r = O.run(O.make_cfg(call_consensus=1, **flags_to_opts(case["flags"])), rs, threads=8)
    assert block_digests(r["corrected"], rs, case["block"]) == case["block_digests"]
    assert sha256(fasta_text(r["corrected"], rs)) == case["fasta_sha256"]
