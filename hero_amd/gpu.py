"""ctypes binding of libhero_gpu.so (include/hero_gpu.h).

This is the product path: it raises if the HIP library is missing and never falls back to a CPU
implementation.
"""
import ctypes
import os

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libhero_gpu.so")

HERO_OK = 0
ERRORS = {-1: "HERO_ERR_ARG", -2: "HERO_ERR_HIP", -3: "HERO_ERR_NOMEM", -4: "HERO_ERR_STATE", -5: "HERO_ERR_LIMIT",
          -6: "HERO_ERR_REF_ABORT"}


class Options(ctypes.Structure):
    _fields_ = [
        ("k", ctypes.c_int32), ("aligner", ctypes.c_int32), ("call_consensus", ctypes.c_int32),
        ("hash_factor", ctypes.c_int32), ("min_overlap", ctypes.c_int64), ("mismatch_rate", ctypes.c_double),
        ("indel_rate", ctypes.c_double), ("consensus_freq", ctypes.c_double),
        ("dx", ctypes.c_int32), ("dy", ctypes.c_int32), ("ex", ctypes.c_int32), ("sx", ctypes.c_int32),
        ("miss", ctypes.c_int32), ("hit", ctypes.c_int32), ("unobserved", ctypes.c_int32),
        ("band_override", ctypes.c_int32), ("max_candidates", ctypes.c_int32),
        ("device", ctypes.c_int32), ("dp_form", ctypes.c_int32), ("chunk_subjects", ctypes.c_int64),
        ("dp_waves", ctypes.c_int32), ("records", ctypes.c_int32), ("dp_per_lane", ctypes.c_int32),
        ("dp_eager", ctypes.c_int32), ("chunk_workers", ctypes.c_int32), ("dp_untagged", ctypes.c_int32),
        ("mapping", ctypes.c_int32), ("band_align_form", ctypes.c_int32), ("long_format", ctypes.c_int32),
        ("merge", ctypes.c_int32), ("file_order", ctypes.c_int32), ("chunk_seeds_max", ctypes.c_int64),
    ]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "table_size", "distinct_keys", "total_keys", "max_probe", "lookups", "seeds", "dp_tasks", "dp_calls_exec",
        "dp_cells_exec", "dp_cells_all", "records_built", "records_kept", "dp_skipped", "vote_oob", "subjects",
        "queries", "reads_changed")] + [(n, ctypes.c_double) for n in (
            "ms_index", "ms_seed", "ms_plan", "ms_dp", "ms_score_vote", "ms_consensus", "ms_total",
            "ms_seed_probe")] + [(n, ctypes.c_uint64) for n in ("dp_graph_builds", "dp_graph_launches")]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class BandPair(ctypes.Structure):
    _fields_ = [("s_off", ctypes.c_int64), ("q_off", ctypes.c_int64), ("s_len", ctypes.c_int32),
                ("q_len", ctypes.c_int32), ("band", ctypes.c_int32), ("flag", ctypes.c_int32)]


class TuneCandidate(ctypes.Structure):
    _fields_ = [("per_lane", ctypes.c_int32), ("waves_per_cu", ctypes.c_int32), ("ms", ctypes.c_float)]


class TuneClass(ctypes.Structure):
    _fields_ = [("band", ctypes.c_int32), ("per_lane", ctypes.c_int32), ("waves_per_cu", ctypes.c_int32),
                ("ncand", ctypes.c_int32), ("tasks", ctypes.c_int64), ("cells", ctypes.c_int64), ("ms", ctypes.c_float),
                ("ms_default", ctypes.c_float), ("cand", TuneCandidate * 16)]


class TuneReport(ctypes.Structure):
    _fields_ = [("classes", ctypes.c_int32), ("_pad", ctypes.c_int32), ("subjects", ctypes.c_int64),
                ("cls", TuneClass * 12)]  # HERO_TUNE_MAX_CLASSES

    def as_list(self):
        out = []
        for i in range(self.classes):
            c = self.cls[i]
            if c.tasks == 0:
                continue
            out.append(dict(band=c.band, tasks=c.tasks, cells=c.cells, per_lane=c.per_lane, waves_per_cu=c.waves_per_cu,
                            ms=round(c.ms, 4), ms_default=round(c.ms_default, 4),
                            gcups=round(c.cells / (c.ms * 1e6), 2) if c.ms > 0 else None,
                            candidates=[dict(per_lane=c.cand[j].per_lane, waves_per_cu=c.cand[j].waves_per_cu,
                                             ms=round(c.cand[j].ms, 4)) for j in range(c.ncand)]))
        return out


class ServerOptions(ctypes.Structure):
    _fields_ = [("workers", ctypes.c_int32), ("max_step_reads", ctypes.c_int32), ("hbm_budget", ctypes.c_int64),
                ("reserved", ctypes.c_int32 * 8)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("step_reads", ctypes.c_int32), ("step_batches", ctypes.c_int32),
                ("_pad", ctypes.c_int32), ("step_id", ctypes.c_int64), ("t_submit", ctypes.c_double),
                ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_ if f[0] != "_pad"}


# exported symbols and their signatures (kept in step with include/hero_gpu.h)
_P = ctypes.c_void_p
SIGNATURES = {
    "hero_default_options": (None, [ctypes.POINTER(Options)]),
    "hero_create": (ctypes.c_int, [ctypes.POINTER(Options), ctypes.POINTER(ctypes.c_void_p)]),
    "hero_destroy": (None, [_P]),
    "hero_last_error": (ctypes.c_char_p, [_P]),
    "hero_set_queries": (ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, _P]),
    "hero_build_index": (ctypes.c_int, [_P]),
    "hero_align_subjects": (ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, _P]),
    "hero_align_self": (ctypes.c_int, [_P]),
    "hero_call_consensus": (ctypes.c_int, [_P, _P]),
    "hero_copy_corrected": (ctypes.c_int, [_P, _P]),
    "hero_take_alignments": (ctypes.c_int, [_P, _P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]),
    "hero_get_tally": (ctypes.c_int, [_P, _P]),
    "hero_set_tally": (ctypes.c_int, [_P, _P]),
    "hero_coverage_histogram": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "hero_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(Stats)]),
    "hero_reset_stats": (ctypes.c_int, [_P]),
    "hero_band_align_batch": (ctypes.c_int, [_P, ctypes.c_int32, _P, _P, _P, _P, _P, _P, _P, ctypes.c_int64]),
    "hero_band_align_batch_noalloc": (ctypes.c_int, [_P, ctypes.c_int32, _P, _P, ctypes.c_int32, ctypes.c_int32, _P, _P,
                                                     _P, _P, _P, _P]),
    "hero_prime_larger_than": (ctypes.c_uint64, [ctypes.c_uint64]),
    "hero_get_extension_tally": (ctypes.c_int, [_P, _P]),
    "hero_take_records": (ctypes.c_int, [_P, _P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), _P, ctypes.c_int64,
                                         ctypes.POINTER(ctypes.c_int64)]),
    "hero_hash_kmers": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, _P, ctypes.c_uint64, _P]),
    "hero_lookup": (ctypes.c_int, [_P, ctypes.c_int32, _P, _P, _P, _P, ctypes.c_int64]),
    "hero_autotune": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.POINTER(TuneReport)]),
    "hero_set_subjects": (ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, _P]),
    "hero_build_subject_index": (ctypes.c_int, [_P]),
    "hero_correct_batch": (ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, _P, _P]),
    "hero_server_start": (ctypes.c_int, [_P, ctypes.POINTER(ServerOptions), ctypes.POINTER(ctypes.c_void_p)]),
    "hero_server_submit": (ctypes.c_int, [_P, ctypes.c_int64, _P, _P, _P, _P, _P, ctypes.POINTER(ctypes.c_int64)]),
    "hero_server_wait": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.POINTER(RequestInfo)]),
    "hero_server_stop": (ctypes.c_int, [_P, ctypes.POINTER(Stats)]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} is missing: the HIP library must be built (__graft_entry__.build())")
        L = ctypes.CDLL(_LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def default_options(**kw):
    o = Options()
    lib().hero_default_options(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


class HeroError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Context:
    """One device context (hero_ctx). Reads are ReadSet-like objects (seq/offset/length/unique_id)."""

    def __init__(self, **opts):
        self.opts = default_options(**opts)
        h = ctypes.c_void_p()
        rc = lib().hero_create(ctypes.byref(self.opts), ctypes.byref(h))
        if rc != HERO_OK:
            raise HeroError(rc, "hero_create failed (no usable GPU or invalid options)")
        self.h = h
        self._keep = []

    def _check(self, rc):
        if rc != HERO_OK:
            raise HeroError(rc, lib().hero_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().hero_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()

    def set_queries(self, rs):
        self._queries = rs
        self._check(lib().hero_set_queries(self.h, rs.n, rs.seq.ctypes.data, rs.offset.ctypes.data,
                                           rs.length.ctypes.data, rs.unique_id.ctypes.data))

    def build_index(self):
        self._check(lib().hero_build_index(self.h))

    def align_subjects(self, rs):
        self._check(lib().hero_align_subjects(self.h, rs.n, rs.seq.ctypes.data, rs.offset.ctypes.data,
                                              rs.length.ctypes.data, rs.unique_id.ctypes.data))

    def align_self(self):
        self._check(lib().hero_align_self(self.h))

    def autotune(self, max_subjects=0):
        """Times the DP schedule candidates per band class on the first subjects; returns the report."""
        r = TuneReport()
        self._check(lib().hero_autotune(self.h, max_subjects, ctypes.byref(r)))
        return r.as_list()

    def set_subjects(self, rs):
        """Resident subject set of the serving form."""
        self._subjects = rs
        self._check(lib().hero_set_subjects(self.h, rs.n, rs.seq.ctypes.data, rs.offset.ctypes.data,
                                            rs.length.ctypes.data, rs.unique_id.ctypes.data))

    def build_subject_index(self):
        self._check(lib().hero_build_subject_index(self.h))

    def correct_batch(self, rs):
        """One query batch corrected against the resident subjects (bytes in the layout of rs.seq)."""
        out = np.empty_like(rs.seq)
        self._check(lib().hero_correct_batch(self.h, rs.n, rs.seq.ctypes.data, rs.offset.ctypes.data,
                                             rs.length.ctypes.data, rs.unique_id.ctypes.data, out.ctypes.data))
        return out

    def consensus(self, device_only=False, out=None):
        """Corrected reads (bytes in the layout of the query seq, into `out` when given);
        device_only keeps them on the GPU."""
        if device_only:
            self._check(lib().hero_call_consensus(self.h, None))
            return None
        if out is None:
            out = np.empty_like(self._queries.seq)
        assert out.dtype == np.uint8 and out.nbytes >= self._queries.seq.nbytes and out.flags.c_contiguous
        self._check(lib().hero_call_consensus(self.h, out.ctypes.data))
        return out

    def copy_corrected(self, out=None):
        if out is None:
            out = np.empty_like(self._queries.seq)
        assert out.dtype == np.uint8 and out.nbytes >= self._queries.seq.nbytes and out.flags.c_contiguous
        self._check(lib().hero_copy_corrected(self.h, out.ctypes.data))
        return out

    def take_alignments(self):
        """The .aln text of the subjects aligned since the last call (record mode)."""
        n = ctypes.c_int64(0)
        self._check(lib().hero_take_alignments(self.h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(max(1, n.value))
        self._check(lib().hero_take_alignments(self.h, buf, n.value, ctypes.byref(n)))
        return buf.raw[:n.value]

    def tally(self):
        t = np.zeros(4 * len(self._queries.seq), dtype=np.uint32)
        self._check(lib().hero_get_tally(self.h, t.ctypes.data))
        return t

    def extension_tally(self):
        """--merge: 20 x 4 u32 per query read (QueryRead::aExtraCoverageMatrix) via hero_get_extension_tally."""
        t = np.zeros(20 * 4 * len(self._queries.length), dtype=np.uint32)
        self._check(lib().hero_get_extension_tally(self.h, t.ctypes.data))
        return t

    def set_tally(self, t):
        """Replaces the vote matrices (4 u32 per base, A,C,T,G) via hero_set_tally."""
        t = np.ascontiguousarray(t, dtype=np.uint32)
        assert t.size == 4 * len(self._queries.seq)
        self._check(lib().hero_set_tally(self.h, t.ctypes.data))

    def coverage_histogram(self, cap=1 << 16):
        """--mapping output: bases by coverage (counts[v], v = 0 .. max) via hero_coverage_histogram."""
        counts = np.zeros(cap, dtype=np.uint64)
        m = ctypes.c_int32(0)
        self._check(lib().hero_coverage_histogram(self.h, counts.ctypes.data, cap, ctypes.byref(m)))
        return counts[:m.value + 1]

    def coverage_text(self, cap=1 << 16):
        """The text the reference's --mapping run writes (QueryDataset::getCoverageAndWriteResults)."""
        h = self.coverage_histogram(cap)
        return "COV\tFRE\n" + "".join(f"{v}\t{int(n)}\n" for v, n in enumerate(h) if n)

    def stats(self):
        s = Stats()
        self._check(lib().hero_get_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        self._check(lib().hero_reset_stats(self.h))

    def band_align(self, pairs):
        """pairs: list of (S, Q, band, flag) strings. Returns list of (cigar, band)."""
        seq = "".join(s + q for s, q, _, _ in pairs).encode()
        arr = (BandPair * len(pairs))()
        off = 0
        for i, (s, q, b, f) in enumerate(pairs):
            arr[i] = BandPair(off, off + len(s), len(s), len(q), b, f)
            off += len(s) + len(q)
        n = len(pairs)
        cap = sum(len(s) + len(q) for s, q, _, _ in pairs) + 4
        band = np.zeros(n, dtype=np.int32)
        coff = np.zeros(n, dtype=np.int64)
        cn = np.zeros(n, dtype=np.int32)
        cc = np.zeros(cap, dtype=np.int32)
        ct = np.zeros(cap, dtype=np.uint8)
        buf = ctypes.create_string_buffer(seq, len(seq) + 1)
        self._check(lib().hero_band_align_batch(self.h, n, ctypes.addressof(arr), buf, band.ctypes.data,
                                                coff.ctypes.data, cn.ctypes.data, cc.ctypes.data, ct.ctypes.data, cap))
        out = []
        for i in range(n):
            o = int(coff[i])
            out.append(("".join(f"{cc[o + t]}{chr(ct[o + t])}" for t in range(cn[i])), int(band[i])))
        return out

    def hash_kmers(self, kmers, table_size):
        k = len(kmers[0])
        buf = "".join(kmers).encode()
        out = np.zeros(len(kmers), dtype=np.uint64)
        self._check(lib().hero_hash_kmers(self.h, len(kmers), k, buf, table_size, out.ctypes.data))
        return out

    def lookup(self, windows):
        n = len(windows)
        buf = "".join(windows).encode()
        cnt = np.zeros(n, dtype=np.int64)
        off = np.zeros(n, dtype=np.int64)
        cap = 1 << 20
        ent = np.zeros(cap, dtype=np.uint64)
        self._check(lib().hero_lookup(self.h, n, buf, cnt.ctypes.data, off.ctypes.data, ent.ctypes.data, cap))
        return [ent[off[i]:off[i] + cnt[i]].copy() for i in range(n)]


def correct(rs, subjects=None, **opts):
    """Self-correction (or queries vs a subject set) in one call; returns (corrected bytes, stats)."""
    with Context(**opts) as ctx:
        ctx.set_queries(rs)
        ctx.build_index()
        if subjects is None:
            ctx.align_self()
        else:
            ctx.align_subjects(subjects)
        out = ctx.consensus() if ctx.opts.call_consensus else None
        return out, ctx.stats()


class Server:
    """The native request loop (hero_server_*) over a context with resident subjects."""

    def __init__(self, ctx, workers=0, max_step_reads=0, hbm_budget=0):
        self.ctx = ctx
        o = ServerOptions(workers=workers, max_step_reads=max_step_reads, hbm_budget=hbm_budget)
        h = ctypes.c_void_p()
        ctx._check(lib().hero_server_start(ctx.h, ctypes.byref(o), ctypes.byref(h)))
        self.h = h
        self._live = {}

    def submit(self, rs):
        out = np.empty_like(rs.seq)
        t = ctypes.c_int64(-1)
        rc = lib().hero_server_submit(self.h, rs.n, rs.seq.ctypes.data, rs.offset.ctypes.data, rs.length.ctypes.data,
                                      rs.unique_id.ctypes.data, out.ctypes.data, ctypes.byref(t))
        if rc != HERO_OK:
            raise HeroError(rc, "hero_server_submit refused the batch")
        self._live[t.value] = (rs, out)
        return t.value

    def wait(self, ticket):
        info = RequestInfo()
        rc = lib().hero_server_wait(self.h, ticket, ctypes.byref(info))
        rs, out = self._live.pop(ticket)
        if rc != HERO_OK:
            raise HeroError(rc, lib().hero_last_error(self.ctx.h).decode())
        return out, info.as_dict()

    def stop(self):
        st = Stats()
        if self.h:
            rc = lib().hero_server_stop(self.h, ctypes.byref(st))
            self.h = None
            if rc != HERO_OK:
                raise HeroError(rc, lib().hero_last_error(self.ctx.h).decode())
        return st.as_dict()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.stop()
