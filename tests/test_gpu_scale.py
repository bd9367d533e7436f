"""Parity at the workloads' stated sizes: the full C1 read set (29,703 reads) and the C2 read set
(1,666,667 query reads) against a 49,999-read subject sample, through the HIP library.

Expected outputs come from tests/golden/scale.json: the reference program's corrected FASTA at its
own flags, and the oracle's vote tally at the C2 knobs (open -6 / extend -1, band 8, 64 candidates
per subject strand)."""
import hashlib

import numpy as np
import pytest

from helpers import block_digests, fasta_text, first_differing_block, flags_to_opts, load_scale, sha256
from hero_amd import gpu, synth

pytestmark = pytest.mark.gpu

SCALE = load_scale()


def _check_fasta(out, rs, case):
    got = block_digests(out, rs, case["block"])
    assert first_differing_block(got, case["block_digests"]) is None, \
        f"first differing block of {case['block']} reads: {first_differing_block(got, case['block_digests'])}"
    assert sha256(fasta_text(out, rs)) == case["fasta_sha256"]


@pytest.mark.parametrize("name", ["c1_full_global_i2", "c1_full_global_i3"])
def test_c1_full_size_matches_reference(name):
    case = SCALE[name]
    rs = synth.generate(**case["synth"])
    assert rs.n == case["n_reads"] == 29703
    with gpu.Context(call_consensus=1, **flags_to_opts(case["flags"])) as ctx:
        ctx.set_queries(rs)
        ctx.build_index()
        ctx.align_self()
        out = ctx.consensus()
        st = ctx.stats()
    assert st["vote_oob"] == 0 and st["dp_skipped"] == 0
    assert st["reads_changed"] == case["reads_changed"]
    _check_fasta(out, rs, case)


@pytest.fixture(scope="module")
def c2_reads():
    rs = synth.generate(**SCALE["c2_sample_ref"]["synth"])
    assert rs.n == 1666667
    return rs


def test_c2_sample_at_reference_flags(c2_reads):
    case = SCALE["c2_sample_ref"]
    rs = c2_reads
    subj = rs.subset(np.arange(case["n_subjects"]))
    with gpu.Context(call_consensus=1, **flags_to_opts(case["flags"])) as ctx:
        ctx.set_queries(rs)
        ctx.build_index()
        ctx.align_subjects(subj)
        out = ctx.consensus()
        st = ctx.stats()
    assert st["vote_oob"] == 0 and st["dp_skipped"] == 0
    assert st["reads_changed"] == case["reads_changed"]
    _check_fasta(out, rs, case)


def test_c2_sample_at_workload_knobs(c2_reads):
    case = SCALE["c2_sample_knobs"]
    rs = c2_reads
    subj = rs.subset(np.arange(case["n_subjects"]))
    with gpu.Context(**case["options"]) as ctx:
        ctx.set_queries(rs)
        ctx.build_index()
        ctx.align_subjects(subj)
        tally = ctx.tally()
        out = ctx.consensus()
        st = ctx.stats()
    assert st["vote_oob"] == 0 and st["dp_skipped"] == 0
    assert st["dp_cells_exec"] == case["dp_cells"] and st["dp_calls_exec"] == case["dp_calls"]
    assert int(tally.astype(np.uint64).sum()) == case["tally_sum"]
    assert hashlib.sha256(tally.astype("<u4").tobytes()).hexdigest() == case["tally_sha256"]
    assert st["reads_changed"] == case["reads_changed"]
    got = block_digests(out, rs, case["block"])
    assert first_differing_block(got, case["block_digests"]) is None
