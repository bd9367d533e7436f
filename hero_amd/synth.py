"""Deterministic synthetic read sets (ctypes over libhero_synth.so).

Workload shapes for the correction path: genome size, read length range, coverage, the
substitution ramp along the read, 1-bp indel rate and optional repeat families. Read names are
their indices, so a read's uniqueId equals its index.
"""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

_LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib")


class _Params(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("genome_len", ctypes.c_uint64),
        ("len_min", ctypes.c_uint32),
        ("len_max", ctypes.c_uint32),
        ("coverage", ctypes.c_double),
        ("n_reads", ctypes.c_uint64),
        ("sub_5p", ctypes.c_double),
        ("sub_3p", ctypes.c_double),
        ("indel_rate", ctypes.c_double),
        ("repeat_frac", ctypes.c_double),
        ("repeat_seg_len", ctypes.c_uint32),
        ("repeat_copies_per_family", ctypes.c_uint32),
        ("repeat_div", ctypes.c_double),
    ]


_lib = None


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(_LIB_DIR, "libhero_synth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        lib = ctypes.CDLL(path)
        lib.hero_synth_create.restype = ctypes.c_void_p
        lib.hero_synth_create.argtypes = [ctypes.POINTER(_Params)]
        lib.hero_synth_num_reads.restype = ctypes.c_int64
        lib.hero_synth_num_reads.argtypes = [ctypes.c_void_p]
        lib.hero_synth_num_bases.restype = ctypes.c_int64
        lib.hero_synth_num_bases.argtypes = [ctypes.c_void_p]
        lib.hero_synth_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.hero_synth_genome.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
        lib.hero_synth_destroy.argtypes = [ctypes.c_void_p]
        lib.hero_write_fasta.restype = ctypes.c_int
        lib.hero_write_fasta.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64]
        _lib = lib
    return _lib


@dataclass
class ReadSet:
    """Reads as the datasets hold them after loading: upper-case ASCII, concatenated."""

    seq: np.ndarray        # uint8, concatenated bases
    offset: np.ndarray     # int64
    length: np.ndarray     # int32
    unique_id: np.ndarray  # int64, atoi(read name)

    @property
    def n(self):
        return int(self.length.shape[0])

    def read(self, i):
        o = int(self.offset[i])
        return self.seq[o:o + int(self.length[i])].tobytes().decode()

    def subset(self, idx):
        idx = np.asarray(idx, dtype=np.int64)
        lens = self.length[idx]
        offs = np.zeros(len(idx), dtype=np.int64)
        if len(idx):
            offs[1:] = np.cumsum(lens[:-1].astype(np.int64))
        total = int(lens.astype(np.int64).sum())
        seq = np.empty(total, dtype=np.uint8)
        for j, i in enumerate(idx):
            o = int(self.offset[i])
            seq[offs[j]:offs[j] + lens[j]] = self.seq[o:o + lens[j]]
        return ReadSet(seq, offs, lens.astype(np.int32), self.unique_id[idx].copy())

    @staticmethod
    def from_strings(seqs, unique_ids=None):
        lens = np.array([len(s) for s in seqs], dtype=np.int32)
        offs = np.zeros(len(seqs), dtype=np.int64)
        if len(seqs):
            offs[1:] = np.cumsum(lens[:-1].astype(np.int64))
        seq = np.frombuffer("".join(seqs).encode(), dtype=np.uint8).copy()
        uid = np.arange(len(seqs), dtype=np.int64) if unique_ids is None else np.asarray(unique_ids, dtype=np.int64)
        return ReadSet(seq, offs, lens, uid)

    def write_fasta(self, path, r0=0, r1=None):
        """Two-line FASTA of reads [r0, r1), named by unique id."""
        r1 = self.n if r1 is None else r1
        rc = _load().hero_write_fasta(path.encode(), self.seq.ctypes.data, self.offset.ctypes.data,
                                      self.length.ctypes.data, self.unique_id.ctypes.data, r0, r1)
        if rc != 0:
            raise OSError(f"cannot write {path}")


def generate(seed, genome_len, len_min, len_max=None, coverage=0.0, n_reads=0, sub_5p=0.0, sub_3p=None,
             indel_rate=0.0, repeat_frac=0.0, repeat_seg_len=0, repeat_copies_per_family=0, repeat_div=0.0,
             return_genome=False):
    lib = _load()
    p = _Params(seed, genome_len, len_min, len_max if len_max is not None else len_min, coverage, n_reads,
                sub_5p, sub_5p if sub_3p is None else sub_3p, indel_rate, repeat_frac, repeat_seg_len,
                repeat_copies_per_family, repeat_div)
    h = lib.hero_synth_create(ctypes.byref(p))
    if not h:
        raise ValueError("invalid synthetic parameters")
    try:
        n = lib.hero_synth_num_reads(h)
        nb = lib.hero_synth_num_bases(h)
        seq = np.empty(nb, dtype=np.uint8)
        offs = np.empty(n, dtype=np.int64)
        lens = np.empty(n, dtype=np.int32)
        lib.hero_synth_copy(h, seq.ctypes.data, offs.ctypes.data, lens.ctypes.data)
        rs = ReadSet(seq, offs, lens, np.arange(n, dtype=np.int64))
        if return_genome:
            g = np.empty(genome_len, dtype=np.uint8)
            lib.hero_synth_genome(h, g.ctypes.data, genome_len)
            return rs, g
        return rs
    finally:
        lib.hero_synth_destroy(h)


# The five workload shapes (read sets only; k, band, aligner and rates are run options).
WORKLOADS = {
    "c1": dict(seed=101, genome_len=100_000, len_min=101, coverage=30.0, sub_5p=0.01),
    "c2": dict(seed=202, genome_len=5_000_000, len_min=150, coverage=50.0, sub_5p=0.005, sub_3p=0.02,
               indel_rate=0.001),
    "c3": dict(seed=303, genome_len=10_000_000, len_min=75, len_max=250, coverage=40.0, sub_5p=0.01),
    "c4": dict(seed=404, genome_len=20_000_000, len_min=150, coverage=60.0, sub_5p=0.01),
    "c5": dict(seed=505, genome_len=50_000_000, len_min=150, coverage=100.0, sub_5p=0.02, indel_rate=0.002,
               repeat_frac=0.10, repeat_seg_len=5000, repeat_copies_per_family=20, repeat_div=0.01),
}
