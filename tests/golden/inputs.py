"""Deterministic input files for the file-format fixtures (formats.json): the same files are written
by make_golden.py (for the reference program) and by the tests (for the host classes).

Each writer takes a directory and returns the paths it wrote. Data only: read sets from
hero_amd.synth (or a seeded numpy generator), written in the layouts the reference's loaders take.
"""
import os

import numpy as np

from hero_amd import synth

COMP = str.maketrans("ACGT", "TGCA")


def _reads(rs):
    return [rs.read(i) for i in range(rs.n)]


def _fastq(path, names, seqs):
    with open(path, "w") as f:
        for n, s in zip(names, seqs):
            f.write(f"@{n}\n{s}\n+\n{'I' * len(s)}\n")


def _fasta_wrapped(path, names, seqs, width):
    with open(path, "w") as f:
        for n, s in zip(names, seqs):
            f.write(f">{n}\n")
            for i in range(0, len(s), width):
                f.write(s[i:i + width] + "\n")


def _lower_some(seqs, every):
    return [s.lower() if i % every == 3 else s for i, s in enumerate(seqs)]


def parse_edges(d):
    """Query files whose records exercise the loader: FASTA wrapped at 7 and 60 columns, a header with
    spaces and a non-numeric name, lower-case bases, N, an empty sequence in the middle, and no
    newline at the end; read by -q a,b (two files)."""
    a = os.path.join(d, "a.fasta")
    b = os.path.join(d, "b.fasta")
    rng = np.random.default_rng(9)
    seqs = ["".join(rng.choice(list("ACGT"), int(rng.integers(20, 90)))) for _ in range(12)]
    with open(a, "w") as f:
        f.write(">0\n" + seqs[0][:7] + "\n" + seqs[0][7:] + "\n")
        f.write(">1 some description\n" + seqs[1].lower() + "\n")
        f.write(">x2\n" + seqs[2][:10] + "N" + seqs[2][11:] + "\n")
        f.write(">3\n\n")  # empty sequence in the middle of the file
        f.write(">4\n" + "\n".join(seqs[4][i:i + 60] for i in range(0, len(seqs[4]), 60)) + "\n")
    with open(b, "w") as f:
        for i in range(5, 12):
            f.write(f">{i}\n{seqs[i]}\n")
        f.write(">12\n" + seqs[0][::-1])  # no newline at the end
    return [a, b]


def paired_fastq(d):
    """C1-like reads (20 kb genome, 101 bp, 30x) as two FASTQ query files (every fourth read lower
    case) and three subject FASTA files wrapped at 60 columns; an even read count."""
    rs = synth.generate(seed=121, genome_len=20000, len_min=101, coverage=30.0, sub_5p=0.01)
    seqs = _reads(rs)
    if len(seqs) % 2:
        seqs = seqs[:-1]
    names = [str(i) for i in range(len(seqs))]
    half = len(seqs) // 2 + 1
    q1, q2 = os.path.join(d, "q1.fastq"), os.path.join(d, "q2.fastq")
    low = _lower_some(seqs, 4)
    _fastq(q1, names[:half], low[:half])
    _fastq(q2, names[half:], low[half:])
    thirds = [0, len(seqs) // 3, 2 * len(seqs) // 3 + 1, len(seqs)]
    subs = []
    for t in range(3):
        p = os.path.join(d, f"s{t}.fasta")
        _fasta_wrapped(p, names[thirds[t]:thirds[t + 1]], seqs[thirds[t]:thirds[t + 1]], 60)
        subs.append(p)
    return [q1, q2], subs


def mates(d, seed=31):
    """Paired reads: fragments of 150-190 bp from a 15 kb genome at 40x; read 2i is a fragment's first
    100 bases, read 2i+1 the reverse complement of its last 100, 0.5 % substitutions. One FASTA file
    used as queries and subjects."""
    rng = np.random.default_rng(seed)
    g = "".join(rng.choice(list("ACGT"), 15000))
    n_frag = int(15000 * 40 / 200) - 1  # 5,998 reads: the reference aborts on multiples of its 400-read chunk
    seqs = []
    for _ in range(n_frag):
        flen = int(rng.integers(150, 191))
        st = int(rng.integers(0, len(g) - flen))
        frag = g[st:st + flen]
        left, right = frag[:100], frag[-100:].translate(COMP)[::-1]
        for s in (left, right):
            s = list(s)
            for p in range(len(s)):
                if rng.random() < 0.005:
                    s[p] = str(rng.choice([c for c in "ACGT" if c != s[p]]))
            seqs.append("".join(s))
    p = os.path.join(d, "mates.fasta")
    with open(p, "w") as f:
        for i, s in enumerate(seqs):
            f.write(f">{i}\n{s}\n")
    return p


def named(d):
    """Reads with names that are not numbers (read_<i>/1), as FASTQ: for -v, which keeps the names."""
    rs = synth.generate(seed=141, genome_len=12000, len_min=90, len_max=140, coverage=15.0, sub_5p=0.01,
                        indel_rate=0.002)
    seqs = _reads(rs)
    p = os.path.join(d, "named.fastq")
    _fastq(p, [f"read_{i}/1" for i in range(len(seqs))], seqs)
    return p


# name -> (writer producing (query files, subject files), reference flags, output suffixes)
def case_files(name, d):
    if name == "parse_edges":
        q = parse_edges(d)
        return q, q
    if name == "paired_fastq_covinfo":
        return paired_fastq(d)
    if name in ("merge_mates", "merge_mates_covinfo"):
        p = mates(d)
        return [p], [p]
    if name == "long_format":
        p = named(d)
        return [p], [p]
    raise KeyError(name)


CASES = {
    # every read written as parsed (consensus threshold above 1 changes no base)
    "parse_edges": dict(flags=["-k", "11", "-m", "4", "-i", "2", "-l", "10", "-a", "g", "-c", "1.1",
                               "--singleReads"], outputs=[".fasta"]),
    "paired_fastq_covinfo": dict(flags=["-k", "31", "-m", "4", "-i", "2", "-l", "40", "-a", "g", "-c", "--covinfo"],
                                 outputs=[".fasta", ".info"]),
    "merge_mates": dict(flags=["-k", "25", "-m", "4", "-i", "2", "-l", "40", "-a", "a", "-c", "--merge"],
                        outputs=["_merged.fasta", "_unmerged.fasta"]),
    "merge_mates_covinfo": dict(flags=["-k", "25", "-m", "4", "-i", "2", "-l", "40", "-a", "g", "-c", "--merge",
                                       "--covinfo", "-z", "150"],
                                outputs=["_merged.fasta", "_unmerged.fasta", ".info"]),
    "long_format": dict(flags=["-k", "31", "-m", "4", "-i", "3", "-l", "40", "-a", "a", "-v"], outputs=[".aln"]),
}
