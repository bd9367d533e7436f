"""The device logic (hero_amd/csrc/*_core.h) run on the CPU against the oracle."""
import numpy as np
import pytest

from helpers import load_jsonl
from hero_amd import synth
from oracle import oracle as O
import hostsim as H


def test_dp_core_known_answers():
    bad = []
    for c in load_jsonl("dp_cases.jsonl"):
        cfg = O.make_cfg(aligner=c["aligner"])
        cig, band = H.band_align(cfg, c["S"], c["Q"], c["band"], c["flag"])
        if (cig, band) != (c["cigar"], c["out_band"]):
            bad.append((c, cig, band))
    assert not bad, bad[:3]


CASES = [
    dict(rs=dict(seed=11, genome_len=8000, len_min=101, coverage=25.0, sub_5p=0.01),
         cfg=dict(k=31, aligner=1, mismatch_rate=0.04, indel_rate=0.02)),
    dict(rs=dict(seed=12, genome_len=8000, len_min=150, coverage=30.0, sub_5p=0.005, sub_3p=0.02, indel_rate=0.002),
         cfg=dict(k=39, aligner=0, mismatch_rate=0.04, indel_rate=0.0267)),
    dict(rs=dict(seed=13, genome_len=8000, len_min=75, len_max=250, coverage=30.0, sub_5p=0.01, indel_rate=0.002),
         cfg=dict(k=25, aligner=1, mismatch_rate=0.04, indel_rate=0.025)),
    dict(rs=dict(seed=14, genome_len=8000, len_min=75, len_max=250, coverage=30.0, sub_5p=0.01, indel_rate=0.002),
         cfg=dict(k=41, aligner=0, mismatch_rate=0.05, indel_rate=0.03, dx=-7, dy=-7)),
    dict(rs=dict(seed=15, genome_len=8000, len_min=101, coverage=25.0, sub_5p=0.01),
         cfg=dict(k=31, aligner=1, mismatch_rate=0.04, indel_rate=0.02, band_override=5, max_candidates=8)),
]


@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_pipeline_core_matches_oracle(case):
    rs = synth.generate(**case["rs"])
    cfg = O.make_cfg(call_consensus=1, min_overlap=40, **case["cfg"])
    ro = O.run(cfg, rs, threads=4, want_tally=True)
    rh = H.run(cfg, rs)
    assert np.array_equal(ro["tally"], rh["tally"])
    assert np.array_equal(ro["corrected"], rh["corrected"])
    for k in ("dp_calls", "dp_cells", "records_built", "records_kept", "seeds"):
        assert ro["stats"][k] == rh["stats"][k], k
    assert rh["stats"]["oob"] == 0


def test_band_cells_closed_form_equals_row_sum():
    """The closed form of the per-call cell count equals the row-by-row sum it replaces."""
    import ctypes

    f = H.lib().hostsim_band_cells
    f.restype = ctypes.c_uint64
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    for ls in range(0, 40):
        for lq in range(0, 40):
            for b in range(0, 45):
                want = sum(min(lq, i + b) - max(0, i - b) + 1 for i in range(ls + 1)) - 1
                assert f(ls, lq, b) == want % (1 << 64), (ls, lq, b)
    for ls, lq, b in [(500, 480, 20), (16383, 16000, 383), (150, 149, 1), (1, 300, 299)]:
        want = sum(min(lq, i + b) - max(0, i - b) + 1 for i in range(ls + 1)) - 1
        assert f(ls, lq, b) == want


@pytest.mark.parametrize("bmax", [4, 8, 12])
@pytest.mark.parametrize("aligner", [0, 1], ids=["affine", "global"])
def test_packed_two_task_dp_equals_one_task_dp(aligner, bmax):
    """The packed 16-bit two-tasks-per-lane DP writes the same pointer rows as the 32-bit DP."""
    costs = dict(dx=-7, dy=-7, ex=-1, miss=-4, hit=1) if aligner == 0 else {}
    cfg = O.make_cfg(aligner=aligner, **costs)
    for max_len in (3, 20, 140, 500):
        assert H.pk16_check(cfg, 400, 12345 + bmax + max_len, bmax, max_len) == 0, (aligner, bmax, max_len)


@pytest.mark.parametrize("bmax", [2, 4, 6, 8, 10, 12])
@pytest.mark.parametrize(
    "costs",
    [
        dict(dx=-4, dy=-4, ex=-1, sx=0, miss=-4, hit=1),
        dict(dx=-7, dy=-7, ex=-2, sx=-1, miss=-3, hit=2),
        dict(dx=-3, dy=-3, ex=-3, sx=0, miss=-1, hit=0),
        dict(dx=-5, dy=-5, ex=-1, sx=-2, miss=-6, hit=3),
    ],
    ids=["default", "hit2", "hit0", "hit3"],
)
def test_tagged_affine_dp_equals_plain_dp(costs, bmax):
    """The tagged affine DP (pointer in the low bits of scaled, offset scores) traces the same
    paths as the plain 32-bit DP, with each task's own row variants and with every row forced
    through the cut / all-edge variants (as a lane among other tasks is)."""
    cfg = O.make_cfg(aligner=0, **costs)
    for max_len in (3, 20, 140, 500):
        for mode_or in (0, 4, 7):
            seed = 777 + bmax + max_len + 13 * mode_or
            assert H.tagged_check(cfg, 300, seed, bmax, max_len, mode_or) == 0, (costs, bmax, max_len, mode_or)


@pytest.mark.parametrize("bmax", [2, 4, 6, 8, 10, 12])
@pytest.mark.parametrize(
    "costs",
    [
        dict(),
        dict(dx=-2, dy=-3, sx=0, miss=-2, hit=1),
        dict(dx=-3, dy=-1, sx=-1, miss=-1, hit=0),
        dict(dx=-4, dy=-4, sx=-2, miss=-5, hit=3),
    ],
    ids=["default", "dx_ne_dy", "hit0", "hit3"],
)
def test_tagged_global_dp_equals_plain_dp(costs, bmax):
    """The tagged global DP (value x 8, offset 4 hit (i + j), move tags 3 match / 2 mismatch / 1 left /
    0 up) traces the same paths as the plain 32-bit global DP, with each task's own row variants and
    with every row forced through the cut / all-edge variants."""
    cfg = O.make_cfg(aligner=1, **costs)
    for max_len in (3, 20, 140, 500):
        for mode_or in (0, 4, 7):
            seed = 919 + bmax + max_len + 13 * mode_or
            assert H.tagged_check(cfg, 300, seed, bmax, max_len, mode_or) == 0, (costs, bmax, max_len, mode_or)


@pytest.mark.parametrize("bmax", [4, 6, 8, 10, 12, 16, 24, 32])
def test_split_pointer_store_reads_back_every_word(bmax):
    """The DP kernels' split pointer store (the first rows in LDS, a side word per 16 rows for a
    last word of one cell, the other rows in scratch): seven lanes put random rows into one area,
    then every word reads back as put, nothing lands past a lane's near words, and rows_for gives
    the most rows that fit (near budgets from none to more than every row)."""
    for max_len in (0, 5, 20, 70, 140):
        for near_words in (0, 1, 2, 3, 17, 48, 52, 300):
            assert H.split_ptrs_roundtrip(20, 3 + bmax + max_len + near_words, bmax, max_len, near_words) == 0, (
                max_len, near_words)


@pytest.mark.parametrize("bmax", [4, 6, 8, 10, 12, 16, 24, 32])
@pytest.mark.parametrize(
    "aligner,costs",
    [(0, dict(dx=-7, dy=-7, ex=-1, sx=0, miss=-4, hit=1)), (0, dict(dx=-3, dy=-5, ex=-1, sx=0, miss=-4, hit=1)),
     (1, dict())],
    ids=["tagged_affine", "affine", "global"],
)
def test_split_pointer_store_gives_the_same_tracebacks(aligner, costs, bmax):
    """Forward passes of five lanes into one split store, then their tracebacks: each equals the
    traceback over a plain row store (the kernels' forms: tagged affine, general affine, tagged
    global rows; near budgets with the boundary inside and outside the tasks)."""
    cfg = O.make_cfg(aligner=aligner, **costs)
    for max_len in (3, 20, 70, 140):
        for near_words in (0, 1, 3, 17, 48, 52, 300):
            assert H.split_ptrs_check(cfg, 40, 99 + 7 * bmax + max_len + near_words, bmax, max_len, near_words) == 0, (
                max_len, near_words)


@pytest.mark.parametrize("bmax", [4, 6, 8, 10, 12])
@pytest.mark.parametrize(
    "costs",
    [
        dict(dx=-4, dy=-4, ex=-1, sx=0, miss=-4, hit=1),
        dict(dx=-7, dy=-7, ex=-1, sx=0, miss=-4, hit=1),
        dict(dx=-7, dy=-7, ex=-2, sx=-1, miss=-3, hit=2),
        dict(dx=-3, dy=-3, ex=-3, sx=0, miss=-1, hit=0),
    ],
    ids=["default", "c2", "hit2", "hit0"],
)
def test_tagged_pk16_dp_equals_plain_dp(costs, bmax):
    """The tagged rows in packed 16-bit halves, two tasks per lane, trace the same paths as the
    plain 32-bit DP for every task of the pair, up to the longest calls the 16-bit range admits,
    with each pair's own row variants and with every row forced through the cut / all-edge ones."""
    cfg = O.make_cfg(aligner=0, **costs)
    for max_len in (3, 20, 140, 400):
        for mode_or in (0, 4, 7):
            seed = 313 + bmax + max_len + 13 * mode_or
            r = H.pk16t_check(cfg, 300, seed, bmax, max_len, mode_or)
            assert r == 0, (costs, bmax, max_len, mode_or, r)


def _byte_reads(reads):
    """ReadSet from a list of bytes objects (any byte values)."""
    lens = np.array([len(r) for r in reads], dtype=np.int32)
    offs = np.zeros(len(reads), dtype=np.int64)
    offs[1:] = np.cumsum(lens[:-1].astype(np.int64))
    seq = np.frombuffer(b"".join(reads), dtype=np.uint8).copy()
    return synth.ReadSet(seq, offs, lens, np.arange(len(reads), dtype=np.int64))


def _predicate_reads(seed=5):
    rng = np.random.default_rng(seed)
    reads = []
    for n in list(range(0, 48)) + [254, 255, 256, 510, 511, 900, 2047]:
        for _ in range(3):
            reads.append(rng.choice(list(b"ACGT"), n).astype(np.uint8).tobytes())
        if n:
            # one foreign byte anywhere (N, lower case, DEL, high bytes)
            r = bytearray(rng.choice(list(b"ACGT"), n).astype(np.uint8).tobytes())
            r[int(rng.integers(n))] = int(rng.choice([ord("N"), ord("a"), ord("t"), 0x7F, 0x80, 0xFF, 0]))
            reads.append(bytes(r))
            # one letter right at, just under and just over the 80% threshold
            thr = int(n * 0.8)
            for cnt in (thr - 1, thr, thr + 1):
                if 0 <= cnt <= n:
                    other = [c for c in b"ACGT" if c != b"ACGT"[n % 4]]
                    body = [b"ACGT"[n % 4]] * cnt + list(rng.choice(other, n - cnt))
                    reads.append(bytes(rng.permutation(np.array(body, dtype=np.uint8))))
    return _byte_reads(reads)


def test_oracle_read_predicates_known_answers():
    # QueryDataset::qualityFilter (src/QueryDataset.cpp:60-75) fixes these by hand: letters other
    # than upper-case A/C/G/T fail; a letter at floor(0.8 n) or more fails; the length tests
    # compare with minOverlap cast to 16 bits (src/QueryDataset.cpp:226, SubjectDataset.cpp:157-162).
    reads = [b"ACGT" * 10, b"AAAAAAAACG", b"AAAAAAACGT", b"ACGTN" * 8, b"acgt" * 10, b"", b"ACGT" * 9 + b"ACG"]
    rs = _byte_reads(reads)
    q, s = O.read_flags(O.make_cfg(min_overlap=39), rs)
    assert q.tolist() == [1, 0, 0, 0, 0, 0, 0]
    assert s.tolist() == [1, 1, 1, 0, 0, 1, 1]
    q, s = O.read_flags(O.make_cfg(min_overlap=0), rs)
    assert q.tolist() == [1, 0, 1, 0, 0, 0, 1]
    assert s.tolist() == [1, 0, 1, 0, 0, 0, 1]
    q, s = O.read_flags(O.make_cfg(min_overlap=65536 + 9), rs)  # compared as 9
    assert q.tolist() == [1, 0, 1, 0, 0, 0, 1]
    assert s.tolist() == [1, 0, 1, 0, 0, 1, 1]


@pytest.mark.parametrize("min_overlap,k", [(40, 39), (0, 1), (12, 30), (65536 + 20, 25)])
def test_read_flags_match_oracle_predicates(min_overlap, k):
    rs = _predicate_reads()
    q, s = O.read_flags(O.make_cfg(min_overlap=min_overlap), rs)
    good, active = H.read_flags(rs, min_overlap, k)
    n = rs.length.astype(np.int64)
    assert np.array_equal(good, q)
    # subjects too short for the overlap test or for one k-mer need no device work
    assert np.array_equal(active, s & (n >= min_overlap) & (n >= k))


def test_pack_matches_plain_packing():
    rs = _predicate_reads(seed=6)
    words, woff = H.pack(rs)
    codes = ((rs.seq >> 1) & 3).astype(np.uint64)
    expect = np.zeros_like(words)
    w = 0
    for i in range(rs.n):
        o, L = int(rs.offset[i]), int(rs.length[i])
        assert int(woff[i]) == w
        for j in range(0, L, 32):
            c = codes[o + j:o + min(L, j + 32)]
            expect[w] = np.bitwise_or.reduce(c << (62 - 2 * np.arange(len(c), dtype=np.uint64)))
            w += 1
    assert w + 2 == len(words)
    assert np.array_equal(words, expect)
