"""Benchmark of the read-correction hot path on one MI355X.

Workload (step = one pass of the hot path over the whole read set, self-correction):
  5 Mb uniform-random genome, 150 bp reads at 50x (1,666,667 reads), substitution rate rising
  linearly from 0.5% (5' end) to 2% (3' end), 0.1% 1-bp indels; k = 39 seeds; affine banded DP
  (match +1, mismatch -4, gap open -6, extend -1), band half-width 8, at most 64 candidate query
  reads per subject strand; consensus threshold 0.6.
A step = index build + seed search / chaining / DP / records / votes over all subjects +
consensus, with the packed reads already resident in HBM; the corrected bases end in HBM. The
copy of the corrected reads to host memory is timed once after the steps (pcie_inclusive).

Prints one JSON line (see the keys below). Extra measurements:
  gcups            executed in-band DP cells / DP kernel time, batched form vs plain form, on this
                   workload and on a mixed-length workload (75-250 bp, band 5% of length)
  roofline         DP kernel: integer VALU ops (18 per affine cell) per second vs the VALU peak
  roofline_probe   seed-probe kernel: algorithmic HBM bytes per second vs 8 TB/s
  mapping          the same reads with -c --mapping (best covering record per subject, coverage
                   histogram): reads/s of index + subjects + histogram
  cpu_baseline     the reference program (oracle/_ref/hero_ref) or the CPU oracle on host cores
"""
import argparse
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from hero_amd import gpu, synth  # noqa: E402

METRIC = "banded-DP GCUPS + corrected reads/s at 1 GPU vs plain; p50/p99 batch latency"
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # 256 CUs x 64 lanes/clk x 2.4 GHz int32 ops
HBM_PEAK_GBS = 8000.0
AFFINE_OPS_PER_CELL = 18
GLOBAL_OPS_PER_CELL = 9

C2_READS = synth.WORKLOADS["c2"]
C2_OPTS = dict(call_consensus=1, k=39, aligner=0, mismatch_rate=0.04, indel_rate=0.0267, min_overlap=40,
               dx=-7, dy=-7, ex=-1, miss=-4, hit=1, band_override=8, max_candidates=64, consensus_freq=0.6)
# serving: 20 Mb genome, 8,000,000 resident 150 bp reads (60x) + a pool of fresh reads from the same genome
C4_RESIDENT = 8_000_000
C4_POOL = 262_144
C4_READS = dict(seed=404, genome_len=20_000_000, len_min=150, n_reads=C4_RESIDENT + C4_POOL, sub_5p=0.005,
                sub_3p=0.02, indel_rate=0.001)
# the same reads at the reference program's own flags (-c -k 39 -m 4 -i 2.67 -l 40 -a a: open -3 / extend -1,
# band 2*round(0.0267*overlap), no candidate cap): the work its CPU leg times
C2_REF_OPTS = dict(call_consensus=1, k=39, aligner=0, mismatch_rate=0.04, indel_rate=0.0267, min_overlap=40)
C4_OPTS = dict(C2_OPTS, max_candidates=0)  # a per-subject candidate cap would depend on the batch mix
C4_RATES = (5, 20, 50, 100, 200)
# mixed lengths (the autotuner's workload): 10 Mb genome, 75-250 bp at 40x, band 2*round(0.025*overlap)
C3_READS = synth.WORKLOADS["c3"]
C3_LEGS = [("k25_global", dict(call_consensus=1, k=25, aligner=1, mismatch_rate=0.04, indel_rate=0.025, min_overlap=40)),
           ("k41_affine", dict(call_consensus=1, k=41, aligner=0, mismatch_rate=0.04, indel_rate=0.025, min_overlap=40))]
# repeat-heavy share: 50 Mb genome with 10 % in 5 kb repeat families, 150 bp at 100x (33.3 M reads),
# 2 % substitutions, 0.2 % indels, affine band 10, no candidate cap
C5_READS = synth.WORKLOADS["c5"]
C5_OPTS = dict(call_consensus=1, k=39, aligner=0, mismatch_rate=0.04, indel_rate=0.0333, min_overlap=40, band_override=10)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def check_clean(st, what):
    """Every DP call ran and every vote landed in its read's tally."""
    if st["vote_oob"] or st["dp_skipped"]:
        raise SystemExit(f"{what}: vote_oob={st['vote_oob']} dp_skipped={st['dp_skipped']}")


def run_steps(ctx, steps):
    """Timed steps: index build, every subject, consensus; the corrected reads stay in HBM."""
    ctx.reset_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.build_index()
        ctx.align_self()
        ctx.consensus(device_only=True)
    t1 = time.perf_counter()
    st = ctx.stats()
    check_clean(st, "timed steps")
    return (t1 - t0) * 1000.0 / steps, st


def host_to_host(rs, opts, steps):
    """Reads/s from ASCII reads in host memory to corrected ASCII bases in host memory: each step
    packs and uploads the reads (hero_set_queries), builds the index, aligns every subject and
    copies the consensus back (hero_call_consensus into a host buffer). After one untimed step."""
    out = np.empty_like(rs.seq)
    phases = []
    with gpu.Context(**opts) as ctx:
        for i in range(steps + 1):
            t = [time.perf_counter()]
            ctx.set_queries(rs)
            t.append(time.perf_counter())
            ctx.build_index()
            t.append(time.perf_counter())
            ctx.align_self()
            t.append(time.perf_counter())
            ctx.consensus(out=out)
            t.append(time.perf_counter())
            if i:
                phases.append(np.diff(np.array(t)) * 1000.0)
        check_clean(ctx.stats(), "host-to-host steps")
    ph = np.mean(np.array(phases), axis=0)
    ms = float(ph.sum())
    ms_apart = ms - float(ph[1])
    return dict(reads_per_s=round(rs.n / (ms / 1000.0), 1), ms_per_step=round(ms, 3), steps=steps,
                phases_ms=dict(pack_upload=round(ph[0], 3), index=round(ph[1], 3), align=round(ph[2], 3),
                               consensus_copy_back=round(ph[3], 3)),
                reads_per_s_index_apart=round(rs.n / (ms_apart / 1000.0), 1),
                note="ASCII reads in host memory to corrected ASCII in host memory, wall clock per step; "
                     "reads_per_s_index_apart leaves the index build out (reported on its own), the "
                     "ASCII packing stays in")


def sample_parity(rs):
    """The timed workload's knobs on its C2 sample (all reads as queries, the first 49,999 as
    subjects) against the committed oracle tally (tests/golden/scale.json)."""
    path = os.path.join(ROOT, "tests", "golden", "scale.json")
    case = {c["name"]: c for c in json.load(open(path))}["c2_sample_knobs"]
    with gpu.Context(**case["options"]) as ctx:
        ctx.set_queries(rs)
        ctx.build_index()
        ctx.align_subjects(rs.subset(np.arange(case["n_subjects"])))
        tally = ctx.tally()
        st = ctx.stats()
    sha = hashlib.sha256(tally.astype("<u4").tobytes()).hexdigest()
    ok = sha == case["tally_sha256"] and st["dp_cells_exec"] == case["dp_cells"]
    return dict(sample=f"all {rs.n} reads indexed, first {case['n_subjects']} subjects, workload knobs",
                tally_sha256_equal=sha == case["tally_sha256"], dp_cells_equal=st["dp_cells_exec"] == case["dp_cells"],
                result="bit-exact" if ok else "MISMATCH")


def mapping_run(rs, steps=2):
    """--mapping on the same reads: index build, every subject (only the best record covering each
    subject votes), coverage histogram; after one untimed step."""
    with gpu.Context(mapping=1, **C2_OPTS) as ctx:
        ctx.set_queries(rs)
        ctx.build_index()
        ctx.align_self()
        ctx.coverage_histogram()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.build_index()
            ctx.align_self()
            h = ctx.coverage_histogram()
        ms = (time.perf_counter() - t0) * 1000.0 / steps
    return dict(workload="c2 reads with -c --mapping: per subject the best record covering it votes; output the "
                         "coverage histogram (QueryDataset::getCoverageAndWriteResults)",
                reads_per_s=round(rs.n / (ms / 1000.0), 1), ms_per_step=round(ms, 3), steps=steps,
                max_coverage=int(len(h) - 1), bases_covered=int(h[1:].sum()))


def gcups_of(st):
    return st["dp_cells_exec"] / (st["ms_dp"] / 1000.0) / 1e9 if st["ms_dp"] > 0 else 0.0


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, from the committed counter summary
    (profiles/c2_hbm_traffic.json: FETCH_SIZE / WRITE_SIZE runs scaled by a 1 GiB copy), or None."""
    path = os.path.join(ROOT, "profiles", "c2_hbm_traffic.json")
    try:
        k = json.load(open(path))["kernels"][kernel]
    except (OSError, KeyError, ValueError):
        return None
    return int((k["fetch_bytes"] + k["write_bytes"]) / max(k["launches"], 1))


def one_form(rs, opts, dp_form, tune=None, autotune=True):
    """One step with one subject chunk in flight at a time, so each stage's event time is its own
    (with two chunk workers the stages of different chunks overlap)."""
    with gpu.Context(dp_form=dp_form, chunk_workers=1, **opts) as ctx:
        ctx.set_queries(rs)
        if dp_form == 0 and autotune:  # the batched form runs the schedule its autotuner picks
            ctx.build_index()
            rep = ctx.autotune()
            if tune is not None:
                tune.extend(rep)
        run_steps(ctx, 1)  # untimed: buffers grow and DP graphs are captured once
        ms, st = run_steps(ctx, 1)
    return st


def cpu_threads():
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    return min(n, int(env)) if env and env.isdigit() else n


def cpu_model():
    """The host CPU as /proc/cpuinfo names it, with its family / model / stepping numbers (a
    virtualized host may report only the brand, e.g. "AMD EPYC")."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = (x.strip() for x in line.split(":", 1))
                if k in ("model name", "cpu family", "model", "stepping", "vendor_id") and k not in info:
                    info[k] = v
            if not line.strip() and info:
                break
    except OSError:
        pass
    name = info.get("model name", "unknown")
    nums = ", ".join(f"{k} {info[k]}" for k in ("cpu family", "model", "stepping") if k in info)
    return f"{name} ({info.get('vendor_id', '?')}, {nums})" if nums else name


def reference_baseline(rs, sample, threads):
    """The reference program on host cores: all reads indexed, `sample` subjects streamed."""
    exe = os.path.join(ROOT, "oracle", "_ref", "hero_ref")
    if not os.path.exists(exe):
        return None
    if sample % 400 == 0:  # the reference's subject streamer aborts on multiples of its 400-read chunk
        sample -= 1
    d = tempfile.mkdtemp(prefix="hero_cpu_")
    q = os.path.join(d, "reads.fasta")
    s = os.path.join(d, "sample.fasta")
    rs.write_fasta(q)
    rs.write_fasta(s, 0, sample)
    cmd = [exe, "-q", q, "-s", s, "-o", os.path.join(d, "out"), "-t", str(threads), "-c", "-k", "39", "-m", "4",
           "-i", "2.67", "-l", "40", "--singleReads", "-a", "a"]
    log("cpu baseline (reference):", " ".join(cmd[1:]))
    t0 = time.perf_counter()
    p = subprocess.run(cmd, capture_output=True, text=True)
    wall = time.perf_counter() - t0
    if p.returncode != 0:
        log(p.stdout[-2000:], p.stderr[-2000:])
        return None
    out = p.stdout

    def grab(pat):
        m = re.search(pat, out)
        return float(m.group(1)) if m else float("nan")

    t_index = grab(r"Function insertQueryDataset\(\) finished in ([0-9.e+-]+) Seconds")
    t_start = grab(r"Function start\(\) finished in ([0-9.e+-]+) Seconds")
    t_load = grab(r"([0-9.e+-]+) seconds \(Time for loading\)")
    t_cons = grab(r"([0-9.e+-]+) seconds \(Time for calling consensus\)")
    t_pass = t_start - t_load
    n = rs.n
    est = t_index + t_pass * n / sample + t_cons
    return dict(value=round(n / est, 1), unit="corrected reads/s", cores=threads, kind="reference",
                sample=(f"reference -c -k 39 -m 4 -i 2.67 -l 40 -a a (its own costs: open -3/extend -1, band "
                        f"2*round(0.0267*overlap), no candidate cap) on {cpu_model()}: all {n} reads indexed, first "
                        f"{sample} subjects streamed; reads/s = N / (index {t_index:.2f} s + subject pass "
                        f"{t_pass:.2f} s x N/{sample} + consensus {t_cons:.2f} s), from the program's own timers; "
                        f"wall {wall:.1f} s"))


def port_baseline(rs, sample, threads):
    """The CPU oracle at the exact workload configuration on the same sample."""
    try:
        from oracle import oracle as O
    except Exception:
        return None
    sub = rs.subset(np.arange(sample))
    cfg = O.make_cfg(**{k: v for k, v in C2_OPTS.items()})
    t0 = time.perf_counter()
    r = O.run(cfg, rs, sub, threads=threads)
    wall = time.perf_counter() - t0
    st = r["stats"]
    n = rs.n
    est = st["t_index"] + st["t_align"] * n / sample + st["t_consensus"]
    cells_rate = st["dp_cells"] / st["t_align"] / 1e9 if st["t_align"] > 0 else 0
    return dict(value=round(n / est, 1), unit="corrected reads/s", cores=threads, kind="port",
                sample=(f"CPU oracle (C, OpenMP) at the exact workload options on {cpu_model()}: all {n} reads "
                        f"indexed, first {sample} subjects aligned; reads/s = N / (index {st['t_index']:.2f} s + "
                        f"align {st['t_align']:.2f} s x N/{sample} + consensus {st['t_consensus']:.2f} s); "
                        f"wall {wall:.1f} s"),
                gcups_incl_seeding=round(cells_rate, 3))


def _percentile(v, q):
    return float(np.percentile(np.asarray(v), q)) if len(v) else None


def _drive(srv, batches, arrivals):
    """Submits batches at the given arrival times (s) from this thread, then waits for all of them.
    Returns per-batch latencies (s) and the reads/s between the first submit and the last completion."""
    tickets = []
    t0 = time.perf_counter()
    for b, at in zip(batches, arrivals):
        delay = at - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        tickets.append(srv.submit(b))
    infos = [srv.wait(t)[1] for t in tickets]
    lat = [i["t_done"] - i["t_submit"] for i in infos]
    span = max(i["t_done"] for i in infos) - min(i["t_submit"] for i in infos)
    reads = sum(b.n for b in batches)
    merged = sum(1 for i in infos if i["step_batches"] > 1)
    return lat, reads / span if span > 0 else 0.0, merged


def serving_latency(seed_duration=2.0, fixed_duration=4.0, workers=2, max_step_reads=32768, sweep_rates=C4_RATES,
                    prepared=None):
    """Serving form on the resident-subject workload: offline throughput of the same path, a fixed
    4096-read batch stream at 82% of that offline rate, and Poisson sweeps with log-uniform batch
    sizes in [1024, 65536]."""
    if prepared is None:
        prepared = prepare_serving()
    rs, pool_idx, ctx, index_s = prepared
    rng = np.random.default_rng(4040)

    def batch(n, start):
        i = (start + np.arange(n)) % pool_idx.size
        return rs.subset(pool_idx[i])

    # offline: the same path over the pool in 65536-read batches, back to back
    big = [batch(65536, i * 65536) for i in range(pool_idx.size // 65536)]
    ctx.correct_batch(big[0])
    t0 = time.perf_counter()
    for b in big:
        ctx.correct_batch(b)
    offline = sum(b.n for b in big) / (time.perf_counter() - t0)
    log(f"offline (65536-read batches): {offline:.0f} reads/s")
    out = dict(workload=("c4: 20 Mb genome, 8,000,000 resident 150 bp reads (60x) indexed at every k-mer position; "
                         "fresh query reads from the same genome; k=39, affine +1/-4/-6/-1, band 8, no candidate cap"),
               index_build_s=round(index_s, 3), offline_reads_per_s=round(offline, 1), workers=workers,
               max_step_reads=max_step_reads)
    with gpu.Server(ctx, workers=workers, max_step_reads=max_step_reads) as srv:
        # fixed 4096-read batches at 82% of the offline rate
        rate = 0.82 * offline / 4096
        nb = max(20, int(rate * fixed_duration))
        arrivals = np.cumsum(rng.exponential(1.0 / rate, nb))
        batches = [batch(4096, j * 4096) for j in range(nb)]
        # untimed: full-size steps on every worker grow their buffers once, then 2 s of the same
        # stream: the first seconds of a server still grow buffers (each move re-captures a worker's
        # DP graph), so a queue builds up then; the timed stream measures the steady state
        warm = [srv.submit(batches[j % nb]) for j in range(4 * workers * max(1, max_step_reads // 4096))]
        for t in warm:
            srv.wait(t)
        nw = max(20, int(rate * 2.0))
        warm_lat, _, _ = _drive(srv, [batch(4096, j * 4096) for j in range(nw)], np.cumsum(rng.exponential(1.0 / rate, nw)))
        lat, rps, merged = _drive(srv, batches, arrivals)
        out["fixed_4k"] = dict(rate_batches_per_s=round(rate, 2), batches=nb, p50_ms=round(1e3 * _percentile(lat, 50), 3),
                               p99_ms=round(1e3 * _percentile(lat, 99), 3), reads_per_s=round(rps, 1),
                               vs_offline=round(rps / offline, 4), merged_steps=merged,
                               warmup_stream=dict(batches=nw, p99_ms=round(1e3 * _percentile(warm_lat, 99), 3)))
        log("serving fixed 4k:", json.dumps(out["fixed_4k"]))
        # the largest Poisson rate of 4096-read batches whose p99 stays under 50 ms: rates rise by 12 %
        # from half the offline rate until a point fails; each point runs >= 200 batches (>= 1 s)
        r = 0.5 * offline / 4096
        points, best = [], None
        while len(points) < 16:
            nb = max(200, int(r * 1.0))
            arrivals = np.cumsum(rng.exponential(1.0 / r, nb))
            batches = [batch(4096, j * 4096) for j in range(nb)]
            lat, rps, merged = _drive(srv, batches, arrivals)
            pt = dict(rate_batches_per_s=round(r, 1), batches=nb, p50_ms=round(1e3 * _percentile(lat, 50), 3),
                      p99_ms=round(1e3 * _percentile(lat, 99), 3), reads_per_s=round(rps, 1),
                      vs_offline=round(rps / offline, 4), merged_steps=merged)
            points.append(pt)
            log("serving 4k search:", json.dumps(pt))
            if pt["p99_ms"] >= 50.0:
                break
            best = pt
            r *= 1.12
        out["max_rate_4k_p99_under_50ms"] = dict(best=best, points=points)
        sweep = []
        for r in sweep_rates:
            nb = max(200, int(r * seed_duration))
            arrivals = np.cumsum(rng.exponential(1.0 / r, nb))
            sizes = np.exp(rng.uniform(np.log(1024), np.log(65536), nb)).astype(int)
            starts = np.cumsum(np.concatenate([[0], sizes[:-1]]))
            batches = [batch(int(n), int(st)) for n, st in zip(sizes, starts)]
            lat, rps, merged = _drive(srv, batches, arrivals)
            sweep.append(dict(rate_batches_per_s=r, batches=nb, p50_ms=round(1e3 * _percentile(lat, 50), 3),
                              p99_ms=round(1e3 * _percentile(lat, 99), 3), reads_per_s=round(rps, 1),
                              merged_steps=merged))
            log("serving poisson:", json.dumps(sweep[-1]))
        out["poisson_log_uniform_1k_64k"] = sweep
    return out


def prepare_serving():
    """The resident subject set on the device with its all-position index, and the query pool."""
    t0 = time.perf_counter()
    rs = synth.generate(**C4_READS)
    resident = rs.subset(np.arange(C4_RESIDENT))
    pool_idx = np.arange(C4_RESIDENT, rs.n)
    log(f"serving workload: {resident.n} resident + {pool_idx.size} pool reads generated in "
        f"{time.perf_counter() - t0:.1f} s")
    ctx = gpu.Context(**C4_OPTS)
    t0 = time.perf_counter()
    ctx.set_subjects(resident)
    t1 = time.perf_counter()
    ctx.build_subject_index()
    t2 = time.perf_counter()
    log(f"resident subjects uploaded in {t1 - t0:.1f} s, all-position index built in {t2 - t1:.1f} s")
    return rs, pool_idx, ctx, t2 - t1


def profile_pass(mode):
    """One pass of a single form, so that a profiler's per-kernel averages cover launches of one kind."""
    if mode == "serve":
        rs, pool_idx, ctx, _ = prepare_serving()
        with gpu.Server(ctx, workers=2, max_step_reads=32768) as srv:
            batches = [rs.subset(pool_idx[(j * 4096 + np.arange(4096)) % pool_idx.size]) for j in range(400)]
            t0 = time.perf_counter()
            for t in [srv.submit(b) for b in batches]:
                srv.wait(t)
            dt = time.perf_counter() - t0
        ctx.close()
        log(f"profile serve: {len(batches)} batches of 4096 reads in {dt * 1e3:.1f} ms "
            f"({len(batches) * 4096 / dt:.0f} reads/s)")
        return
    rs = synth.generate(**C2_READS)
    if mode == "serial":
        st = one_form(rs, C2_OPTS, 0, autotune=False)
        log(f"profile serial (two identical passes, the second timed): index {st['ms_index']:.1f}, seeds "
            f"{st['ms_seed']:.1f}, plan {st['ms_plan']:.1f}, dp {st['ms_dp']:.1f}, score+votes "
            f"{st['ms_score_vote']:.1f}, consensus {st['ms_consensus']:.1f} ms; per pass: dp_tasks {st['dp_tasks']}, "
            f"dp_cells_all {st['dp_cells_all']}, dp_cells_exec {st['dp_cells_exec']}, lookups {st['lookups']}, "
            f"seeds {st['seeds']}")
        return
    with gpu.Context(**C2_OPTS) as ctx:
        ctx.set_queries(rs)
        run_steps(ctx, 1)
        ms, st = run_steps(ctx, 1)
    log(f"profile step (a warmup and a timed step): {ms:.1f} ms, {rs.n / (ms / 1e3):.0f} reads/s")


def copy_calibration(nbytes=1 << 30):
    """A plain device copy of nbytes (reads nbytes, writes nbytes): under a counter profiler its
    FETCH_SIZE / WRITE_SIZE give the scale between counter units and bytes."""
    import torch

    x = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    log(f"copy calibration: {nbytes} bytes copied")
    del x, y


def hbm_used():
    """Device bytes in use (total - free), from the HIP runtime the library links."""
    import ctypes
    gpu.lib()
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total))
    return total.value - free.value


def c5_run(steps):
    """The repeat-heavy share at its real shape: every read indexed and streamed as a subject, one
    untimed step, then `steps` timed steps; reads/s, GCUPS, the stage split and peak HBM."""
    t0 = time.perf_counter()
    rs = synth.generate(**C5_READS)
    log(f"c5: {rs.n} reads generated in {time.perf_counter() - t0:.1f} s")
    base = hbm_used()
    with gpu.Context(**C5_OPTS) as ctx:
        t0 = time.perf_counter()
        ctx.set_queries(rs)
        log(f"c5: reads packed and resident in {time.perf_counter() - t0:.1f} s")
        run_steps(ctx, 1)
        peak = hbm_used() - base
        log(f"c5: untimed step done, {peak / 2**30:.1f} GiB of HBM in use")
        ms, st = run_steps(ctx, steps)
    line = dict(metric=METRIC, workload="c5: 50 Mb genome with 10% in 5 kb repeat families at 1% divergence, 150 bp, "
                                        "100x, 2% substitutions, 0.2% indels, k=39, affine -3/-1, band 10, no cap",
                reads=rs.n, steps=steps, ms_per_step=round(ms, 3), reads_per_s=round(rs.n / (ms / 1000.0), 1),
                gcups=round(gcups_of(st), 2) if steps else None,
                stages_ms={k: round(st[k] / steps, 3) for k in ("ms_index", "ms_seed", "ms_plan", "ms_dp",
                                                               "ms_score_vote", "ms_consensus")},
                counts={k: st[k] // steps for k in ("seeds", "dp_tasks", "dp_calls_exec", "dp_cells_exec",
                                                    "records_kept", "reads_changed")},
                hbm_gib=round(peak / 2**30, 2), note="stage times summed over the chunks of a step; two chunks run at once")
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=50_000, help="subjects streamed by the CPU baselines")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the plain-form and mixed-length runs")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving (latency) runs")
    ap.add_argument("--no-autotune", action="store_true",
                    help="default DP schedules instead of the autotuner's (profiles of exactly the step's kernels)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device copy (the reference point for HBM byte counters)")
    ap.add_argument("--workload", default="c2", choices=["c2", "c5"],
                    help="c2: the headline step (default); c5: the repeat-heavy share at its real shape")
    ap.add_argument("--profile", choices=["step", "serial", "serve"],
                    help="one pass for a profiler, no JSON line: step = a warmup and a timed C2 step (two chunk "
                         "workers); serial = the same with one chunk in flight; serve = 4,096-read batches "
                         "through the request loop")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path; a subject set "
              "shards across GPUs by subject chunks with a vote-tally reduction, which is not built.",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "c5":
        c5_run(args.steps)
        return
    if args.copy_calibration:
        copy_calibration()
    if args.profile:
        profile_pass(args.profile)
        return
    t0 = time.perf_counter()
    rs = synth.generate(**C2_READS)
    log(f"workload: {rs.n} reads x 150 bp generated in {time.perf_counter() - t0:.1f} s")
    ctx = gpu.Context(**C2_OPTS)
    ctx.set_queries(rs)  # packed reads resident in HBM before any timed step
    ctx.build_index()
    tune_c2 = []
    if not args.no_autotune:
        tune_c2 = ctx.autotune()  # DP schedule per band class, timed on the first 65,536 subjects
        log("autotune:", json.dumps([{k: v for k, v in t.items() if k != "candidates"} for t in tune_c2]))
    for _ in range(args.warmup):
        run_steps(ctx, 1)
    ms, st = run_steps(ctx, args.steps)
    t0 = time.perf_counter()
    ctx.copy_corrected()  # the corrected reads to host memory over PCIe (outside the timed steps)
    copy_ms = (time.perf_counter() - t0) * 1000.0
    ctx.close()
    value = rs.n / (ms / 1000.0)
    steps = args.steps
    # stage times, GCUPS and the rooflines from the same step with one chunk in flight at a time
    # (the timed steps overlap two chunks, whose stage times add up to more than the step)
    s1 = one_form(rs, C2_OPTS, 0, autotune=not args.no_autotune)
    dp_ms = s1["ms_dp"]
    cells_all = s1["dp_cells_all"]
    gcups_batched = gcups_of(s1)
    log(f"step {ms:.1f} ms, {value:.0f} reads/s; one chunk at a time: index {s1['ms_index']:.1f}, seeds "
        f"{s1['ms_seed']:.1f}, plan {s1['ms_plan']:.1f}, dp {dp_ms:.1f}, score+votes {s1['ms_score_vote']:.1f}, "
        f"consensus {s1['ms_consensus']:.1f} ms; {gcups_batched:.1f} GCUPS")

    cells_exec = s1["dp_cells_exec"]
    achieved = cells_exec * AFFINE_OPS_PER_CELL / (dp_ms / 1000.0) / 1e12
    roofline = dict(bound="valu", kernel="k_dp_batched (all band classes)", achieved=round(achieved, 3),
                    peak=round(VALU_PEAK_TOPS, 2), unit="Tops/s (int32 VALU)", frac=round(achieved / VALU_PEAK_TOPS, 4),
                    traffic=measured_traffic("hero::k_dp_batched<8, true, true>"), traffic_unit="HBM bytes per launch",
                    ops_per_cell=AFFINE_OPS_PER_CELL, cells_per_step=int(cells_exec),
                    cells_computed_per_step=int(cells_all), dp_ms_per_step=round(dp_ms, 3),
                    note="cells of the DP calls the reference's control flow executes")
    # seed probes: packed subject bases + one 32 B slot per canonical probe (a forward window and
    # its reverse-strand partner share one; "lookups" counts the reference's per-strand lookups)
    # + 8 B per bucket entry read
    probe_ms = s1["ms_seed_probe"]
    bases = int(rs.length.astype(np.int64).sum())
    probe_bytes = bases / 4 + s1["lookups"] / 2 * 32 + s1["seeds"] * 8
    probe_gbs = probe_bytes / (probe_ms / 1000.0) / 1e9 if probe_ms > 0 else 0.0
    roofline_probe = dict(bound="hbm", kernel="k_seed_count", achieved=round(probe_gbs, 1), peak=HBM_PEAK_GBS,
                          unit="GB/s", frac=round(probe_gbs / HBM_PEAK_GBS, 4),
                          traffic=measured_traffic("hero::k_seed_count"), traffic_unit="HBM bytes per launch",
                          bytes_per_step=int(probe_bytes), ms_per_step=round(probe_ms, 3))

    gcups = dict(batched=round(gcups_batched, 2))
    if not args.no_compare:
        stp = one_form(rs, C2_OPTS, 1)
        gcups["plain"] = round(gcups_of(stp), 2)
        gcups["batched_over_plain"] = round(gcups_batched / gcups["plain"], 2) if gcups["plain"] else None
        # the plain form's own roofline: one wavefront per call, a lane per pair of band diagonals, so
        # at band 8 about 9 of 64 lanes hold a cell on each anti-diagonal step
        plain_tops = gcups["plain"] * AFFINE_OPS_PER_CELL / 1e3
        lane_bound = (C2_OPTS["band_override"] + 1) / 64
        gcups["plain_roofline"] = dict(bound="valu", achieved=round(plain_tops, 3), peak=round(VALU_PEAK_TOPS, 2),
                                       unit="Tops/s (int32 VALU)", frac=round(plain_tops / VALU_PEAK_TOPS, 4),
                                       lane_bound=round(lane_bound, 4),
                                       frac_of_lane_bound=round(plain_tops / VALU_PEAK_TOPS / lane_bound, 4))
        rs3 = synth.generate(**C3_READS)
        mixed = dict(workload=f"10 Mb genome, 75-250 bp reads at 40x ({rs3.n} reads), band 2*round(0.025*overlap)")
        for leg, opts in C3_LEGS:
            tune_c3 = []
            s3b, s3p = one_form(rs3, opts, 0, tune_c3), one_form(rs3, opts, 1)
            gb, gp = gcups_of(s3b), gcups_of(s3p)
            mixed[leg] = dict(batched=round(gb, 2), plain=round(gp, 2), batched_over_plain=round(gb / max(gp, 1e-9), 2),
                              dp_cells=s3b["dp_cells_exec"], ms_dp_batched=round(s3b["ms_dp"], 3),
                              ms_dp_plain=round(s3p["ms_dp"], 3),
                              autotune=[{k: v for k, v in t.items() if k != "candidates"} for t in tune_c3])
            log(f"mixed lengths {leg}:", json.dumps(mixed[leg]))
        del rs3
        gcups["mixed_length"] = mixed
        log("gcups:", json.dumps(gcups))

    mapping = None
    if not args.no_compare:
        mapping = mapping_run(rs)
        log("mapping:", json.dumps(mapping))

    h2h = host_to_host(rs, C2_OPTS, 2)
    log("host to host:", json.dumps(h2h))
    with gpu.Context(**C2_REF_OPTS) as rctx:
        rctx.set_queries(rs)
        run_steps(rctx, 1)
        ref_ms, ref_st = run_steps(rctx, 2)
    ref_flags = dict(reads_per_s=round(rs.n / (ref_ms / 1000.0), 1), ms_per_step=round(ref_ms, 3),
                     dp_cells_per_step=ref_st["dp_cells_exec"] // 2,
                     flags="-c -k 39 -m 4 -i 2.67 -l 40 -a a (open -3 / extend -1, no candidate cap)")
    log("reference flags:", json.dumps(ref_flags))
    parity = sample_parity(rs)
    log("parity:", json.dumps(parity))

    latency = None
    if not args.no_serve:
        prepared = prepare_serving()
        latency = serving_latency(prepared=prepared)
        prepared[2].close()

    cpu = None
    port = None
    if not args.no_cpu:
        threads = cpu_threads()
        sample = min(args.cpu_sample, rs.n - 1)
        cpu = reference_baseline(rs, sample, threads)
        port = port_baseline(rs, sample, threads)
        if cpu is None:
            cpu, port = port, None

    line = dict(metric=METRIC, value=round(value, 1), unit="corrected reads/s", n_gpus=1, steps=steps,
                warmup=args.warmup, ms_per_step=round(ms, 3), higher_is_better=True, scaling="weak",
                vs_baseline=None, dtype="int32", data="synthetic",
                config=dict(workload=("c2: 5 Mb genome, 150 bp, 50x, 0.5->2% substitutions, 0.1% indels, k=39, affine "
                                      "+1/-4/-6/-1, band 8, cap 64 candidates, self-correction"),
                            reads=rs.n, bases=bases, parallelism="1 GPU"),
                value_note="device-resident: packed reads in HBM at the start of each step, corrected reads left in HBM",
                host_to_host=h2h, reference_flags=ref_flags, parity=parity,
                pcie_inclusive=dict(copy_ms=round(copy_ms, 3), reads_per_s=round(rs.n / ((ms + copy_ms) / 1000.0), 1)),
                gcups=gcups, autotune=tune_c2, latency=latency, mapping=mapping, roofline=roofline,
                roofline_probe=roofline_probe, cpu_baseline=cpu,
                cpu_baseline_port=port,
                stages_ms=dict(index=round(s1["ms_index"], 3), seed=round(s1["ms_seed"], 3),
                               plan=round(s1["ms_plan"], 3), dp=round(dp_ms, 3),
                               score_vote=round(s1["ms_score_vote"], 3), consensus=round(s1["ms_consensus"], 3),
                               note="one chunk in flight at a time (chunk_workers=1)"),
                counts=dict(dp_tasks=st["dp_tasks"] // steps, dp_calls_exec=st["dp_calls_exec"] // steps,
                            dp_cells_exec=st["dp_cells_exec"] // steps, records_kept=st["records_kept"] // steps,
                            reads_changed=st["reads_changed"] // steps))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
