"""Shared helpers of the test suite."""
import hashlib
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_jsonl(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return [json.loads(line) for line in f if line.strip()]


def load_e2e():
    with open(os.path.join(GOLDEN, "e2e.json")) as f:
        return json.load(f)


def load_mapping():
    """Reference --mapping runs (-c): synthetic read-set parameters, flags, coverage histogram text."""
    with open(os.path.join(GOLDEN, "mapping.json")) as f:
        return json.load(f)


def flags_to_opts(flags):
    """Reference CLI flags (-k -m -i -l -a) -> option dict (rates as fractions, like Config.cpp:165-170)."""
    o = {}
    it = iter(flags)
    for f in it:
        v = next(it)
        if f == "-k":
            o["k"] = int(v)
        elif f == "-m":
            o["mismatch_rate"] = float(v) / 100.0
        elif f == "-i":
            o["indel_rate"] = float(v) / 100.0
        elif f == "-l":
            o["min_overlap"] = int(v)
        elif f == "-a":
            o["aligner"] = 0 if v == "a" else 1
    return o


def read_digests(seq_bytes, rs):
    out = []
    for i in range(rs.n):
        o = int(rs.offset[i])
        s = seq_bytes[o:o + int(rs.length[i])].tobytes()
        out.append(hashlib.blake2b(s, digest_size=8).hexdigest())
    return out


def fasta_text(seq_bytes, rs):
    parts = []
    for i in range(rs.n):
        o = int(rs.offset[i])
        parts.append(b">" + str(int(rs.unique_id[i])).encode() + b"\n" + seq_bytes[o:o + int(rs.length[i])].tobytes() + b"\n")
    return b"".join(parts)


def sha256(b):
    return hashlib.sha256(b).hexdigest()


def cigar_ops(cig):
    ops, num = [], ""
    for ch in cig:
        if ch.isdigit():
            num += ch
        else:
            ops.append((int(num), ch))
            num = ""
    return ops


def load_scale():
    """Fixtures at the workloads' stated sizes (tests/golden/make_scale_golden.py)."""
    with open(os.path.join(GOLDEN, "scale.json")) as f:
        return {c["name"]: c for c in json.load(f)}


def block_digests(seq_bytes, rs, block):
    """blake2b-64 of each block of `block` reads' bases, reads joined by newlines."""
    out = []
    for b0 in range(0, rs.n, block):
        h = hashlib.blake2b(digest_size=8)
        for i in range(b0, min(rs.n, b0 + block)):
            o = int(rs.offset[i])
            h.update(seq_bytes[o:o + int(rs.length[i])].tobytes())
            h.update(b"\n")
        out.append(h.hexdigest())
    return out


def first_differing_block(got, want):
    return next((i for i, (a, b) in enumerate(zip(got, want)) if a != b), None)
