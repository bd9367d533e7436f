"""ctypes binding of tests/native/libhero_hostsim.so (CPU run of the device logic)."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhero_hostsim.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        from oracle.oracle import Cfg, Reads  # same layouts
        L = ctypes.CDLL(LIB_PATH)
        L.hostsim_run.restype = ctypes.c_int
        L.hostsim_run.argtypes = [ctypes.POINTER(Cfg), ctypes.POINTER(Reads), ctypes.POINTER(Reads), ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p]
        L.hostsim_band_align.restype = ctypes.c_int
        L.hostsim_band_align.argtypes = [ctypes.POINTER(Cfg), ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int)]
        L.hostsim_read_flags.argtypes = [ctypes.POINTER(Reads), ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_void_p]
        L.hostsim_pack.restype = ctypes.c_int64
        L.hostsim_pack.argtypes = [ctypes.POINTER(Reads), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        _lib = L
    return _lib


def read_flags(reads, min_overlap, k):
    """The product's (query good, active as subject) flags per read."""
    from oracle.oracle import _reads
    good = np.zeros(reads.n, dtype=np.uint8)
    active = np.zeros(reads.n, dtype=np.uint8)
    lib().hostsim_read_flags(_reads(reads), min_overlap, k, good.ctypes.data, active.ctypes.data)
    return good, active


def pack(reads):
    """The product's 2-bit packing: (words, per-read word offsets)."""
    from oracle.oracle import _reads
    cap = int(((reads.length.astype(np.int64) + 31) // 32).sum()) + 2
    words = np.zeros(cap, dtype=np.uint64)
    woff = np.zeros(reads.n, dtype=np.uint64)
    got = lib().hostsim_pack(_reads(reads), words.ctypes.data, cap, woff.ctypes.data)
    if got != cap:
        raise RuntimeError(f"hostsim_pack wrote {got} words, expected {cap}")
    return words, woff


def run(cfg, queries, subjects=None):
    from oracle.oracle import _reads
    subjects = queries if subjects is None else subjects
    corrected = np.empty_like(queries.seq)
    tally = np.zeros(4 * len(queries.seq), dtype=np.uint32)
    stats = np.zeros(8, dtype=np.uint64)
    rc = lib().hostsim_run(ctypes.byref(cfg), _reads(queries), _reads(subjects), corrected.ctypes.data,
                           tally.ctypes.data, stats.ctypes.data)
    if rc < 0:
        raise RuntimeError(f"hostsim_run failed: {rc}")
    keys = ["seeds", "tasks", "dp_calls", "dp_cells", "records_built", "records_kept", "cells_all", "oob"]
    return dict(corrected=corrected, tally=tally, stats=dict(zip(keys, map(int, stats))))


def band_align(cfg, S, Q, band, flag=0):
    buf = ctypes.create_string_buffer(4 * (len(S) + len(Q)) + 16)
    b = ctypes.c_int(0)
    rc = lib().hostsim_band_align(ctypes.byref(cfg), S.encode(), len(S), Q.encode(), len(Q), band, flag, buf,
                                  len(buf), ctypes.byref(b))
    if rc != 0:
        raise RuntimeError(f"hostsim_band_align failed: {rc}")
    return buf.value.decode(), b.value


def pk16_check(cfg, n, seed, bmax, max_len):
    """Task pairs whose packed two-task DP pointers differ from the one-task DP's (0 = all equal)."""
    f = lib().hostsim_pk16_check
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
    return f(ctypes.addressof(cfg), n, seed, bmax, max_len)


def tagged_check(cfg, n, seed, bmax, max_len, mode_or=0):
    """Tasks whose tagged affine DP traceback differs from the plain one-task DP's (0 = all
    equal; -1 = options outside the tagged form). mode_or forces row variants on every row."""
    f = lib().hostsim_tagged_check
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    return f(ctypes.addressof(cfg), n, seed, bmax, max_len, mode_or)


def pk16t_check(cfg, n, seed, bmax, max_len, mode_or=0):
    """Tasks whose tagged packed two-task DP traceback differs from the one-task DP's (0 = all
    equal; -1 = options or lengths outside the tagged 16-bit form). mode_or forces row variants."""
    f = lib().hostsim_pk16t_check
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    return f(ctypes.addressof(cfg), n, seed, bmax, max_len, mode_or)


def split_ptrs_check(cfg, n, seed, bmax, max_len, near_words):
    """Tasks whose traceback over the split pointer store (near rows in near_words words per lane,
    side words for a last word of one cell, the rest far) differs from the plain row store's
    (0 = all equal); five lanes share the near and far areas."""
    f = lib().hostsim_split_ptrs_check
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    return f(ctypes.addressof(cfg), n, seed, bmax, max_len, near_words)


def split_ptrs_roundtrip(n, seed, bmax, max_len, near_words):
    """Words read back wrong from the split pointer store after seven lanes put random rows into
    one near and far area (0 = every word as put)."""
    f = lib().hostsim_split_ptrs_roundtrip
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    return f(n, seed, bmax, max_len, near_words)
