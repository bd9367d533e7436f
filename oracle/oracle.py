"""ctypes binding of the CPU oracle (oracle/_build/libhero_oracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libhero_oracle.so")


class Cfg(ctypes.Structure):
    _fields_ = [
        ("k", ctypes.c_int32), ("aligner", ctypes.c_int32), ("call_consensus", ctypes.c_int32),
        ("hash_factor", ctypes.c_int32), ("min_overlap", ctypes.c_int64), ("mismatch_rate", ctypes.c_double),
        ("indel_rate", ctypes.c_double), ("consensus_freq", ctypes.c_double), ("band_override", ctypes.c_int32),
        ("max_candidates", ctypes.c_int32), ("dx", ctypes.c_int32), ("dy", ctypes.c_int32), ("ex", ctypes.c_int32),
        ("sx", ctypes.c_int32), ("miss", ctypes.c_int32), ("hit", ctypes.c_int32), ("unobserved", ctypes.c_int32),
        ("mapping", ctypes.c_int32),
    ]


class Reads(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("seq", ctypes.c_void_p), ("offset", ctypes.c_void_p),
                ("length", ctypes.c_void_p), ("unique_id", ctypes.c_void_p)]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "table_size", "distinct_keys", "total_keys", "max_probe", "lookups", "seeds", "dp_calls", "dp_cells",
        "records_built", "records_kept", "reads_changed")] + [
        ("t_index", ctypes.c_double), ("t_align", ctypes.c_double), ("t_consensus", ctypes.c_double)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_default_cfg.argtypes = [ctypes.POINTER(Cfg)]
        L.oracle_hash_kmer.restype = ctypes.c_uint64
        L.oracle_hash_kmer.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_uint64]
        L.oracle_table_size.restype = ctypes.c_uint64
        L.oracle_table_size.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_band_align.restype = ctypes.c_int
        L.oracle_band_align.argtypes = [ctypes.POINTER(Cfg), ctypes.c_char_p, ctypes.c_int64, ctypes.c_int64,
                                        ctypes.c_char_p, ctypes.c_int64, ctypes.c_int64,
                                        ctypes.POINTER(ctypes.c_int64), ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_record_score.restype = ctypes.c_int
        L.oracle_record_score.argtypes = [ctypes.POINTER(Cfg), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p]
        L.oracle_consensus_read.restype = ctypes.c_int
        L.oracle_consensus_read.argtypes = [ctypes.POINTER(Cfg), ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]
        L.oracle_run.restype = ctypes.c_int
        L.oracle_run.argtypes = [ctypes.POINTER(Cfg), ctypes.POINTER(Reads), ctypes.POINTER(Reads),
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                 ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.POINTER(Stats)]
        L.oracle_seeds.restype = ctypes.c_int64
        L.oracle_seeds.argtypes = [ctypes.POINTER(Cfg), ctypes.POINTER(Reads), ctypes.c_char_p, ctypes.c_int32,
                                   ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64]
        L.oracle_read_flags.argtypes = [ctypes.POINTER(Cfg), ctypes.POINTER(Reads), ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_free.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def make_cfg(**kw):
    c = Cfg()
    lib().oracle_default_cfg(ctypes.byref(c))
    for key, v in kw.items():
        if not hasattr(c, key):
            raise KeyError(key)
        setattr(c, key, v)
    return c


def _reads(rs):
    r = Reads(rs.n, rs.seq.ctypes.data, rs.offset.ctypes.data, rs.length.ctypes.data, rs.unique_id.ctypes.data)
    r._keep = rs
    return r


def read_flags(cfg, rs):
    """Per-read (query predicate, subject predicate) as two uint8 arrays."""
    q = np.zeros(rs.n, dtype=np.uint8)
    s = np.zeros(rs.n, dtype=np.uint8)
    lib().oracle_read_flags(ctypes.byref(cfg), _reads(rs), q.ctypes.data, s.ctypes.data)
    return q, s


def hash_kmer(kmer, P):
    return lib().oracle_hash_kmer(kmer.encode(), len(kmer), P)


def table_size(n, factor=32):
    return lib().oracle_table_size(n, factor)


def band_align(cfg, S, Q, band, flag=0, s0=0, s1=None, q0=0, q1=None):
    """One BandAligner call; returns (cigar string, widened band, cells)."""
    s1 = len(S) if s1 is None else s1
    q1 = len(Q) if q1 is None else q1
    cap = len(S) + len(Q) + 4
    cnt = np.zeros(cap, dtype=np.int32)
    typ = np.zeros(cap, dtype=np.uint8)
    b = ctypes.c_int64(band)
    cells = ctypes.c_uint64(0)
    n = lib().oracle_band_align(ctypes.byref(cfg), S.encode(), s0, s1, Q.encode(), q0, q1, ctypes.byref(b), flag,
                                cnt.ctypes.data, typ.ctypes.data, cap, ctypes.byref(cells))
    if n < 0:
        raise ValueError(f"oracle_band_align failed: {n}")
    cig = "".join(f"{cnt[i]}{chr(typ[i])}" for i in range(n))
    return cig, b.value, cells.value


def record_score(cfg, ops, cap_mis, cap_indel):
    cnt = np.array([c for c, _ in ops], dtype=np.int32)
    typ = np.frombuffer("".join(t for _, t in ops).encode(), dtype=np.uint8).copy()
    out = np.zeros(6, dtype=np.float64)
    m = lib().oracle_record_score(ctypes.byref(cfg), cnt.ctypes.data, typ.ctypes.data, len(ops), cap_mis, cap_indel,
                                  out.ctypes.data)
    merged = "".join(f"{cnt[i]}{chr(typ[i])}" for i in range(m))
    return out, merged


def consensus_read(cfg, seq, tally):
    buf = np.frombuffer(seq.encode(), dtype=np.uint8).copy()
    t = np.ascontiguousarray(tally, dtype=np.uint32).copy()
    lib().oracle_consensus_read(ctypes.byref(cfg), buf.ctypes.data, len(seq), t.ctypes.data)
    return buf.tobytes().decode()


def seeds(cfg, queries, subject, subject_uid, reverse):
    cap = 1 << 20
    out = np.zeros(3 * cap, dtype=np.int64)
    n = lib().oracle_seeds(ctypes.byref(cfg), _reads(queries), subject.encode(), len(subject), subject_uid,
                           int(reverse), out.ctypes.data, cap)
    if n < 0:
        raise ValueError(f"oracle_seeds failed: {n}")
    return out[:3 * n].reshape(n, 3)


def coverage_text(queries, tally):
    """--mapping output text (COV / FRE histogram) from a vote tally of run(..., want_tally=True)."""
    f = lib().oracle_coverage_text
    f.restype = ctypes.c_void_p
    f.argtypes = [ctypes.POINTER(Reads), ctypes.c_void_p]
    t = np.ascontiguousarray(tally, dtype=np.uint32)
    r = _reads(queries)
    p = f(ctypes.byref(r), t.ctypes.data)
    s = ctypes.string_at(p).decode()
    lib().oracle_free(ctypes.c_void_p(p))
    return s


def run(cfg, queries, subjects=None, threads=0, want_tally=False, want_aln=False, want_records=False):
    """Full path. Returns dict(corrected=ReadSet seq bytes or None, tally, aln, records, stats)."""
    subjects = queries if subjects is None else subjects
    corrected = np.empty_like(queries.seq) if cfg.call_consensus else None
    tally = np.zeros(4 * len(queries.seq), dtype=np.uint32) if want_tally else None
    aln_p = ctypes.c_void_p(0)
    rec_p = ctypes.c_void_p(0)
    st = Stats()
    rc = lib().oracle_run(ctypes.byref(cfg), _reads(queries), _reads(subjects),
                          corrected.ctypes.data if corrected is not None else None,
                          tally.ctypes.data if tally is not None else None,
                          ctypes.byref(aln_p) if want_aln else None, ctypes.byref(rec_p) if want_records else None,
                          threads, ctypes.byref(st))
    if rc != 0:
        raise RuntimeError(f"oracle_run failed: {rc}")
    out = dict(corrected=corrected, tally=tally, stats=st.as_dict(), aln=None, records=None)
    if want_aln and aln_p.value:
        out["aln"] = ctypes.string_at(aln_p.value).decode()
        lib().oracle_free(aln_p)
    if want_records and rec_p.value:
        out["records"] = ctypes.string_at(rec_p.value).decode()
        lib().oracle_free(rec_p)
    return out
