"""The C++ host classes with the reference's shapes (hero_amd/host/hero_host.h), driven like the
reference's main.cpp: corrected FASTA and .aln bytes equal the reference's, BandAligner calls equal
the reference's CIGARs and widened bands."""
import os
import subprocess

import pytest

from helpers import load_e2e, load_jsonl, load_mapping, sha256
from hero_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "native", "host_shim_main")

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not os.path.exists(DRIVER), reason="host_shim_main is not built")]


def _flag(flags, name):
    return flags[flags.index(name) + 1]


def _run(mode, case, tmp_path, chunk=0):
    rs = synth.generate(**case["synth"])
    fa = str(tmp_path / "reads.fasta")
    rs.write_fasta(fa)
    out = str(tmp_path / "out")
    f = case["flags"]
    cmd = [DRIVER, mode, fa, _flag(f, "-k"), _flag(f, "-m"), _flag(f, "-i"), _flag(f, "-l"), _flag(f, "-a"), out]
    if chunk:
        cmd.append(str(chunk))
    subprocess.run(cmd, check=True, timeout=120)
    return open(out, "rb").read()


@pytest.mark.parametrize("case", load_e2e()[:3], ids=lambda c: c["name"])
def test_corrected_fasta_through_host_classes(case, tmp_path):
    assert sha256(_run("correct", case, tmp_path)) == case["fasta_sha256"]


@pytest.mark.parametrize("case", load_mapping(), ids=lambda c: c["name"])
def test_mapping_histogram_through_host_classes(case, tmp_path):
    """-c --mapping through the host classes: the histogram file equals the reference's."""
    assert _run("mapping", case, tmp_path).decode() == case["coverage_text"]


def test_alignment_records_through_host_classes_in_subject_chunks(tmp_path):
    case = load_e2e()[2]
    out = _run("aln", case, tmp_path, chunk=400)  # SubjectDataset chunks of the reference's default size
    assert len(out) == case["aln_bytes"] and sha256(out) == case["aln_sha256"]


@pytest.mark.parametrize("aligner", [1, 0], ids=["global", "affine"])
def test_band_aligner_protocol(aligner, tmp_path):
    cases = [c for c in load_jsonl("dp_cases.jsonl") if c["aligner"] == aligner][:400]
    inp = tmp_path / "cases.txt"
    inp.write_text("".join(f"{c['S'] or '-'} {c['Q'] or '-'} {c['band']} {c['flag']}\n" for c in cases))
    out = tmp_path / "out.txt"
    subprocess.run([DRIVER, "band", "a" if aligner == 0 else "g", str(inp), str(out)], check=True, timeout=120)
    got = out.read_text().splitlines()
    assert len(got) == len(cases)
    for c, line in zip(cases, got):
        band, cigar = (line.split(" ") + [""])[:2]
        assert (int(band), cigar) == (c["out_band"], c["cigar"]), c


def _formats():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "formats.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("case", _formats(), ids=lambda c: c["name"])
def test_file_formats_through_host_classes(case, tmp_path):
    """main.cpp's flow over the host classes on FASTQ / wrapped FASTA / multi-file inputs: paired
    and single-read FASTA, .info (--covinfo), --merge's _merged / _unmerged files and the -v
    long-format .aln equal the reference's bytes."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import inputs
    q, s = inputs.case_files(case["name"], str(tmp_path))
    prefix = str(tmp_path / "out")
    cmd = [DRIVER, "run", "-q", ",".join(q), "-s", ",".join(s), "-o", prefix, "-t", "8"] + case["flags"]
    subprocess.run(cmd, check=True, timeout=300)
    for suf, want in case["files"].items():
        data = open(prefix + suf, "rb").read()
        assert (len(data), sha256(data)) == (want["bytes"], want["sha256"]), suf


def _tables():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "tables.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("case", _tables(), ids=lambda c: c["name"])
def test_hash_table_method_seam_matches_reference(case, tmp_path):
    """HashTableLongKey through its HashTableMethod base: the table size from the reference's prime
    list, the reference's collision counts (its own printout), iBitOfReadId, and getListOfReads
    returning the table's own list (the same pointer twice) holding read 0's occurrence."""
    rs = synth.generate(**case["synth"])
    fa = str(tmp_path / "reads.fasta")
    rs.write_fasta(fa)
    mapping = "--mapping" in case["flags"]
    factor = case["flags"][case["flags"].index("-f") + 1] if "-f" in case["flags"] else "32"
    cmd = [DRIVER, "table", fa, str(case["k"]), factor] + (["mapping"] if mapping else [])
    out = subprocess.run(cmd, check=True, timeout=120, capture_output=True, text=True).stdout.splitlines()
    head = out[0].split()
    got = dict(zip(head[0::2], (int(x) for x in head[1::2])))
    assert (got["size"], got["collisions"], got["max_single"]) == (case["size"], case["collisions"], case["max_single"])
    bits = 35 if mapping else 40
    assert got["bits"] == bits
    k = case["k"]
    for line in out[1:]:
        pos = int(line.split()[1].rstrip(":"))
        entries = [int(x) for x in line.split(":")[1].split()[:-2]]
        assert line.endswith("same 1")
        assert ((pos << bits) | 0) in entries, line
        assert all(rs.read(e & ((1 << bits) - 1))[e >> bits:(e >> bits) + k] == rs.read(0)[pos:pos + k] for e in entries)
