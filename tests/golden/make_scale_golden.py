"""Writes tests/golden/scale.json: parity fixtures at the workloads' stated sizes.

Run in the container that holds the reference sources, after `make -C oracle ref oracle`:

  c1_full          the 100 kb / 101 bp / 30x read set (29,703 reads), k=31, global (-a g), -i 2 and -i 3,
                   through the reference program (oracle/_ref/hero_ref): corrected FASTA sha256,
                   digests of 1,024-read blocks, changed-read count
  c2_sample_ref    the 5 Mb / 150 bp / 50x read set (1,666,667 reads) as queries, its first 49,999 reads
                   as subjects, the reference's own costs (-3/-1), no candidate cap, -i 2.67: the same
                   digests from the reference program
  c2_sample_knobs  the same queries and subjects at the workload's knobs (affine open -6 / extend -1,
                   mismatch -4, band 8, 64 candidates per subject strand) through the oracle: sha256
                   of the vote tally (u32 little-endian, 4 per base), its sum, and the corrected-read
                   digests

The read sets are regenerated from their seeds by hero_amd.synth on any machine; only parameters,
flags and digests are stored.
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
HERO_REF = os.path.join(ROOT, "oracle", "_ref", "hero_ref")
BLOCK = 1024

C1 = dict(seed=101, genome_len=100_000, len_min=101, coverage=30.0, sub_5p=0.01)
C2 = dict(seed=202, genome_len=5_000_000, len_min=150, coverage=50.0, sub_5p=0.005, sub_3p=0.02, indel_rate=0.001)
C2_SUBJECTS = 49_999
# the workload's knobs in the C ABI's option names (hero_options / oracle Cfg)
C2_KNOBS = dict(call_consensus=1, k=39, aligner=0, mismatch_rate=0.04, indel_rate=0.0267, min_overlap=40,
                dx=-7, dy=-7, ex=-1, miss=-4, hit=1, band_override=8, max_candidates=64, consensus_freq=0.6)


def block_digests(rs, seq_bytes):
    """blake2b-64 of each block of BLOCK reads' corrected bases (reads joined by newlines)."""
    out = []
    for b0 in range(0, rs.n, BLOCK):
        h = hashlib.blake2b(digest_size=8)
        for i in range(b0, min(rs.n, b0 + BLOCK)):
            o = int(rs.offset[i])
            h.update(seq_bytes[o:o + int(rs.length[i])].tobytes())
            h.update(b"\n")
        out.append(h.hexdigest())
    return out


def corrected_from_fasta(rs, text):
    """The corrected bases of a --singleReads FASTA written by the reference, in read order."""
    import numpy as np
    lines = text.split(b"\n")
    seq = np.empty_like(rs.seq)
    for i in range(rs.n):
        s = lines[2 * i + 1]
        o = int(rs.offset[i])
        assert len(s) == int(rs.length[i])
        seq[o:o + len(s)] = np.frombuffer(s, dtype=np.uint8)
    return seq


def changed(rs, seq_bytes):
    n = 0
    for i in range(rs.n):
        o = int(rs.offset[i])
        L = int(rs.length[i])
        if (seq_bytes[o:o + L] != rs.seq[o:o + L]).any():
            n += 1
    return n


def run_reference(rs, n_subjects, flags, tag):
    d = tempfile.mkdtemp(prefix="golden_scale_", dir="/tmp")
    q = os.path.join(d, "q.fasta")
    rs.write_fasta(q)
    s = q
    if n_subjects is not None:
        s = os.path.join(d, "s.fasta")
        rs.write_fasta(s, 0, n_subjects)
    prefix = os.path.join(d, "out")
    cmd = [HERO_REF, "-q", q, "-s", s, "-o", prefix, "-t", "8"] + flags + ["-c", "--singleReads"]
    t0 = time.time()
    subprocess.run(cmd, capture_output=True, check=True)
    print(f"{tag}: reference ran in {time.time() - t0:.1f} s", flush=True)
    text = open(prefix + ".fasta", "rb").read()
    return text


def ref_case(name, synth_args, flags, n_subjects=None):
    from hero_amd import synth
    rs = synth.generate(**synth_args)
    assert (n_subjects or rs.n) % 400 != 0, "the reference's subject streamer aborts on multiples of 400 reads"
    text = run_reference(rs, n_subjects, flags, name)
    seq = corrected_from_fasta(rs, text)
    out = dict(name=name, synth=synth_args, flags=flags, n_reads=rs.n, n_subjects=n_subjects or rs.n,
               fasta_sha256=hashlib.sha256(text).hexdigest(), block=BLOCK, block_digests=block_digests(rs, seq),
               reads_changed=changed(rs, seq))
    print(name, rs.n, out["reads_changed"], flush=True)
    return out


def knob_case():
    import numpy as np
    from hero_amd import synth
    from oracle import oracle as O
    rs = synth.generate(**C2)
    subj = rs.subset(np.arange(C2_SUBJECTS))
    t0 = time.time()
    res = O.run(O.make_cfg(**C2_KNOBS), rs, subj, threads=8, want_tally=True)
    print(f"c2_sample_knobs: oracle ran in {time.time() - t0:.1f} s", flush=True)
    tally = res["tally"]
    seq = res["corrected"]
    out = dict(name="c2_sample_knobs", synth=C2, options=C2_KNOBS, n_reads=rs.n, n_subjects=C2_SUBJECTS,
               tally_sha256=hashlib.sha256(tally.astype("<u4").tobytes()).hexdigest(),
               tally_sum=int(tally.astype(np.uint64).sum()), block=BLOCK, block_digests=block_digests(rs, seq),
               reads_changed=changed(rs, seq), dp_cells=int(res["stats"]["dp_cells"]),
               dp_calls=int(res["stats"]["dp_calls"]))
    print("c2_sample_knobs", out["reads_changed"], out["tally_sum"], flush=True)
    return out


def main():
    cases = [
        ref_case("c1_full_global_i2", C1, ["-k", "31", "-m", "4", "-i", "2", "-l", "40", "-a", "g"]),
        ref_case("c1_full_global_i3", C1, ["-k", "31", "-m", "4", "-i", "3", "-l", "40", "-a", "g"]),
        ref_case("c2_sample_ref", C2, ["-k", "39", "-m", "4", "-i", "2.67", "-l", "40", "-a", "a"], C2_SUBJECTS),
        knob_case(),
    ]
    with open(os.path.join(HERE, "scale.json"), "w") as f:
        json.dump(cases, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
