"""The host classes' FASTA / FASTQ loader (hero_amd/host/hero_host.cpp ReadFiles) against the
reference's loader, without a device: on inputs with wrapped and empty sequences, non-numeric
names, lower case, N and a missing final newline, read as two query files, every read comes out as
the reference wrote it (a run whose consensus threshold above 1 changes no base)."""
import json
import os
import subprocess
import sys

import pytest

from helpers import sha256

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "native", "host_shim_main")
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="host_shim_main is not built")
def test_loader_matches_reference_parse(tmp_path):
    import inputs
    case = {c["name"]: c for c in json.load(open(os.path.join(ROOT, "tests", "golden", "formats.json")))}["parse_edges"]
    q, _ = inputs.case_files("parse_edges", str(tmp_path))
    out = str(tmp_path / "parsed.fasta")
    subprocess.run([DRIVER, "parse", out, ",".join(q)], check=True, timeout=60)
    data = open(out, "rb").read()
    want = case["files"][".fasta"]
    assert (len(data), sha256(data)) == (want["bytes"], want["sha256"]), data.decode()[:400]
