"""hero_band_align_batch_noalloc (device arrays, fixed device scratch) against hero_band_align_batch
on the same calls: two calls in a row on one context, the second smaller."""
import ctypes

import numpy as np
import pytest

from helpers import load_jsonl
from hero_amd import gpu

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    """The HIP runtime libhero_gpu.so links (the caller's device memory comes from it)."""
    global _hip
    if _hip is None:
        gpu.lib()
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


class DevArray:
    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), max(1, self.host.nbytes)) == 0
        assert hip().hipMemcpy(self.p, self.host.ctypes.data, self.host.nbytes, 1) == 0  # host to device

    def get(self):
        out = np.empty_like(self.host)
        assert hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        hip().hipFree(self.p)


def noalloc_align(ctx, pairs, max_len, max_band):
    seq = "".join(s + q for s, q, _, _ in pairs).encode()
    arr = (gpu.BandPair * len(pairs))()
    off, cap = 0, []
    for i, (s, q, b, f) in enumerate(pairs):
        arr[i] = gpu.BandPair(off, off + len(s), len(s), len(q), b, f)
        off += len(s) + len(q)
        cap.append(len(s) + len(q))
    n = len(pairs)
    coff = np.zeros(n, dtype=np.int64)
    coff[1:] = np.cumsum(cap)[:-1]
    total = int(sum(cap))
    d_pairs = DevArray(np.frombuffer(bytes(arr), dtype=np.uint8))
    d_seq = DevArray(np.frombuffer(seq, dtype=np.uint8))
    d_band = DevArray(np.full(n, -99, dtype=np.int32))
    d_off = DevArray(coff)
    d_n = DevArray(np.full(n, -99, dtype=np.int32))
    d_cnt = DevArray(np.zeros(total, dtype=np.int32))
    d_typ = DevArray(np.zeros(total, dtype=np.uint8))
    try:
        rc = gpu.lib().hero_band_align_batch_noalloc(ctx.h, n, d_pairs.p, d_seq.p, max_len, max_band, d_band.p, d_off.p,
                                                     d_n.p, d_cnt.p, d_typ.p, None)
        assert rc == 0, gpu.lib().hero_last_error(ctx.h)
        assert hip().hipDeviceSynchronize() == 0
        band, cn, cnt, typ = d_band.get(), d_n.get(), d_cnt.get(), d_typ.get()
    finally:
        for d in (d_pairs, d_seq, d_band, d_off, d_n, d_cnt, d_typ):
            d.free()
    out = []
    for i in range(n):
        assert cn[i] >= 0, (i, cn[i])
        o = int(coff[i])
        out.append(("".join(f"{cnt[o + t]}{chr(typ[o + t])}" for t in range(cn[i])), int(band[i])))
    return out


def _long_pairs(rng, n):
    pairs = []
    for t in range(n):
        L = int(rng.integers(600, 2000))
        a = "".join(rng.choice(list("ACGT"), L))
        b = list(a)
        for _ in range(L // 30):
            b[int(rng.integers(0, L))] = str(rng.choice(list("ACGT")))
        del b[int(rng.integers(0, L - 5)):][:3]
        pairs.append((a, "".join(b), int(rng.choice([8, 90, 300])), int(rng.integers(0, 3))))
    return pairs


@pytest.mark.parametrize("aligner", [1, 0], ids=["global", "affine"])
def test_noalloc_equals_blocking_entry_twice(aligner):
    cases = [c for c in load_jsonl("dp_cases.jsonl") if c["aligner"] == aligner]
    rng = np.random.default_rng(11 + aligner)
    first = [(c["S"], c["Q"], c["band"], c["flag"]) for c in cases] + _long_pairs(rng, 12)
    second = first[::7][:200]  # a smaller input on the same context
    with gpu.Context(aligner=aligner) as ctx:
        for pairs in (first, second):
            want = ctx.band_align(pairs)
            max_len = max(max(len(s), len(q)) for s, q, _, _ in pairs)
            max_band = max(max(b, abs(len(s) - len(q))) for s, q, b, _ in pairs)
            got = noalloc_align(ctx, pairs, max_len, max_band)
            bad = [(i, g, w) for i, (g, w) in enumerate(zip(got, want)) if g != w]
            assert not bad, bad[:3]


def test_noalloc_rejects_limits_and_reports_per_call_status():
    with gpu.Context() as ctx:
        assert gpu.lib().hero_band_align_batch_noalloc(ctx.h, 1, None, None, 100, 2000, None, None, None, None, None,
                                                       None) == -1  # null arrays
        z = DevArray(np.zeros(64, dtype=np.uint8))
        try:
            assert gpu.lib().hero_band_align_batch_noalloc(ctx.h, 1, z.p, z.p, 100, 2000, z.p, z.p, z.p, z.p, z.p,
                                                           None) == -5  # max_band above 1023
            assert gpu.lib().hero_band_align_batch_noalloc(ctx.h, 1, z.p, z.p, 20000, 100, z.p, z.p, z.p, z.p, z.p,
                                                           None) == -5  # max_len above 16383
        finally:
            z.free()
        # a call longer than max_len is reported in its cigar_n, the others run
        pairs = [("ACGTACGTAC", "ACGTTCGTAC", 2, 0), ("A" * 50, "A" * 50, 2, 0), ("ACGT", "ACGA", 1, 0)]
        seq = "".join(s + q for s, q, _, _ in pairs).encode()
        arr = (gpu.BandPair * 3)()
        off = 0
        for i, (s, q, b, f) in enumerate(pairs):
            arr[i] = gpu.BandPair(off, off + len(s), len(s), len(q), b, f)
            off += len(s) + len(q)
        d = [DevArray(np.frombuffer(bytes(arr), dtype=np.uint8)), DevArray(np.frombuffer(seq, dtype=np.uint8)),
             DevArray(np.zeros(3, dtype=np.int32)), DevArray(np.array([0, 20, 120], dtype=np.int64)),
             DevArray(np.zeros(3, dtype=np.int32)), DevArray(np.zeros(128, dtype=np.int32)),
             DevArray(np.zeros(128, dtype=np.uint8))]
        try:
            assert gpu.lib().hero_band_align_batch_noalloc(ctx.h, 3, d[0].p, d[1].p, 10, 4, d[2].p, d[3].p, d[4].p,
                                                           d[5].p, d[6].p, None) == 0
            assert hip().hipDeviceSynchronize() == 0
            cn = d[4].get()
        finally:
            for x in d:
                x.free()
        assert cn[1] == -5 and cn[0] > 0 and cn[2] > 0
