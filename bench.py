"""Prover-core benchmark on one MI355X.

Workload (one step = one pass of the prover core over a resident trace):
  Goldilocks trace of 2^20 rows x 128 columns, uniform in [0, p) from seed
  0x5EED0002 (SplitMix64, values >= p rejected), resident in HBM;
  iNTT 2^20 -> coset LDE x8 (shift 7) -> SHA3 row Merkle tree over the 2^23
  LDE rows -> f0 = sum_c (c+1) LDE_c -> STIR deep_fri_prove(f0) with fold
  schedule [8,8,8,8,8,4] (2^23 -> 2^6), 80 queries, bound to the row root ->
  80 row openings. The output (proof bytes) is checked each step for
  determinism; parity with the oracle is covered by tests/ at reduced sizes.

Prints one JSON line. `value` = prover-core proofs/s (whole job, one GPU).
`roofline` reports the dominant kernel (most GPU time) with its per-launch
algorithmic bytes over its average launch time, measured with HIP events on
the library's stream during the timed steps. `cpu_baseline` times the C
oracle on a bounded sample of the same pipeline on this host.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--calib-copy]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

P = 2**64 - 2**32 + 1
HBM_PEAK_GBS = 8000.0
METRIC = "proofs/s on 1 MI355X (batched vs plain, bit-exact); LDE GB/s; p50/p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def splitmix_field(seed, n):
    """n uniform field elements: SplitMix64 stream, values >= p rejected."""
    out = np.empty(0, dtype=np.uint64)
    state = np.uint64(seed)
    with np.errstate(over="ignore"):
        while out.size < n:
            m = int((n - out.size) * 1.01) + 64
            idx = np.arange(1, m + 1, dtype=np.uint64)
            z = state + idx * np.uint64(0x9E3779B97F4A7C15)
            state = z[-1]
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
            out = np.concatenate([out, z[z < np.uint64(P)]])
    return out[:n]


CPU_SHARE = 16  # host cores one GPU's share of the box gives a job


def cpu_threads():
    """threads the CPU legs use: the host's cores, at most one GPU's share"""
    return max(1, min(os.cpu_count() or 1, CPU_SHARE))


def cpu_baseline(w_full, lg_full, blowup, schedule_full, s0):
    """Time the C oracle on a bounded sample of the prover core in the
    reference's own threading (one column at a time in the LDE -- ark-poly
    without its parallel feature -- and the rayon loops on cpu_threads()
    threads: Merkle leaves and levels at >= 256, FRI folds and coset
    extraction at >= 2^12), and scale to the full workload by rows x columns
    (which leaves out the LDE's extra log factor, so it favours the CPU)."""
    from tests import _oracle as O
    T = cpu_threads()
    lg, w = 15, 128
    n = 1 << lg
    trace = splitmix_field(0x5EED0002, w * n).reshape(w, n)
    lg0 = lg + blowup.bit_length() - 1
    sched = [8] * (lg0 // 3) + ([1 << (lg0 % 3)] if lg0 % 3 else [])
    params = O.fri_params(sched, r=s0, s0=s0)
    rr = ctypes.create_string_buffer(32)
    pp, pl = O.u8p(), O.sz()
    with O.threads(T):
        t = time.perf_counter()
        rc = O.lib().ref_prover_core(O.ptr(trace), w, n, blowup, 7, params, rr, None, pp, pl)
        dt = time.perf_counter() - t
    assert rc == 0
    O.take_bytes(pp, pl)
    scale = (w_full << lg_full) / (w << lg)
    full_s = dt * scale
    with O.threads(T):
        small = cpu_small_proofs(O)
    return {"value": 1.0 / full_s, "unit": "proofs/s", "cores": T, "nproc": os.cpu_count(), "kind": "port",
            "threading": "reference threading: LDE one column at a time (ark-poly 0.4.2 without 'parallel'); "
                         "Merkle leaves / levels (>= 256, merkle lib.rs:345), FRI folds and coset extraction "
                         "(>= 2^12, fri.rs:89) on %d threads" % T,
            "sample": "oracle ref_prover_core on 2^%d x %d, blowup %d, schedule %s, %d queries: %.2f s on %d "
                      "threads; scaled x%.0f by rows*columns to the 2^%d x %d workload" %
                      (lg, w, blowup, sched, s0, dt, T, scale, lg_full, w_full),
            "cpu": _cpu_model(), "small_proofs": small}


def oracle_sha_proof(O, alg, msg, fsb, height):
    """prove_ds_ksk_binding's pipeline for one C4 job with oracle pieces:
    trace padded to `height` by repeating the last useful row -> LDE x32 ->
    merge -> STIR proof with 54 queries bound to fsb"""
    tr, nb, _ = (O.sha256_trace if alg == 256 else O.sha512_trace)(msg)
    useful = 128 * nb
    if height > tr.shape[1]:
        tr = np.concatenate([tr, np.repeat(tr[:, useful - 1:useful], height - tr.shape[1], axis=1)], axis=1)
    lde = O.lde_columns(tr, 32)
    c = (O.merge_sha256 if alg == 256 else O.merge_sha512)(lde, height, 32, nb, 1)
    del lde
    lg = (32 * height).bit_length() - 1
    sched = [8] * (lg // 3) + ([1 << (lg % 3)] if lg % 3 else [])
    return O.deep_fri_prove(c, O.fri_params(sched, r=54, s0=54, pi_hash=fsb))


def cpu_c4(jobs, budget_s=15.0):
    """The C4 CPU leg: the oracle (reference threading, cpu_threads()) proves
    the jobs in order of size, smallest first, until budget_s is spent; the
    whole batch's time is the sampled time scaled by the LDE work
    sum(columns x n0 x log2 n0) of all jobs over that of the sampled ones."""
    from tests import _oracle as O
    T = cpu_threads()

    def work(a, h):
        n0 = 32 * h
        return (756 if a == 256 else 1510) * n0 * (n0.bit_length() - 1)
    order = sorted(range(len(jobs)), key=lambda i: (jobs[i][2], jobs[i][0]))
    done, spent, w_s = 0, 0.0, 0
    with O.threads(T):
        for i in order:
            alg, msg, h, fsb = jobs[i]
            t = time.perf_counter()
            oracle_sha_proof(O, alg, msg, fsb, h)
            spent += time.perf_counter() - t
            w_s += work(alg, h)
            done += 1
            if spent >= budget_s:
                break
    w_all = sum(work(a, h) for a, _, h, _ in jobs)
    est = spent * w_all / w_s
    return {"proofs_per_s": round(len(jobs) / est, 4), "cores": T, "kind": "port",
            "sample": "%d of the %d proofs (the smallest, 2^%d..2^%d rows) in %.1f s on %d threads; scaled by "
                      "LDE work sum(columns x n0 x log2 n0) x%.1f" %
                      (done, len(jobs), min(jobs[i][2] for i in order[:done]).bit_length() - 1,
                       max(jobs[i][2] for i in order[:done]).bit_length() - 1, spent, T, w_all / w_s)}


def cpu_ed25519(blowup, queries, cols=512, pts=16):
    """Seconds the C oracle needs for one Ed25519 proof (k_scalar 256, a
    2-block message) in the reference threading, from a bounded sample: the
    trace fill in full, the LDE of `cols` of the 41 519 columns scaled to all
    of them (the LDE runs one column at a time), the 70 606 constraints at
    `pts` points scaled to every LDE point and divided over the threads (the
    merge's into_par_iter), and one STIR proof over the LDE domain."""
    from stark_swarm_amd.zone import synthetic_ed25519_signatures
    from tests import _oracle as O
    T = cpu_threads()
    pk, sig, msg = synthetic_ed25519_signatures(1, 0x5EED0003)[0]
    t = time.perf_counter()
    e = O.Ed25519(pk, sig, msg, 256)
    tr = e.trace()
    fill = time.perf_counter() - t
    h = tr.shape[1]
    n0 = h * blowup
    t = time.perf_counter()
    O.lde_columns(tr[:cols], blowup)
    lde_s = (time.perf_counter() - t) * tr.shape[0] / cols
    cur, nxt = np.ascontiguousarray(tr[:, 0]), np.ascontiguousarray(tr[:, 1])
    t = time.perf_counter()
    for i in range(pts):
        e.constraints(cur, nxt, i % (h - 1))
    merge_s = (time.perf_counter() - t) * n0 / pts / T
    lg = n0.bit_length() - 1
    sched = [8] * (lg // 3) + ([1 << (lg % 3)] if lg % 3 else [])
    c = splitmix_field(0x5EED000C, n0)
    with O.threads(T):
        t = time.perf_counter()
        O.deep_fri_prove(c, O.fri_params(sched, r=queries, s0=queries, pi_hash=bytes(32)))
        fri_s = time.perf_counter() - t
    total = fill + lde_s + merge_s + fri_s
    return {"proofs_per_s": round(1.0 / total, 5), "cores": T, "kind": "port",
            "sample": "trace fill (%.2f s) + LDE of %d of %d columns x%.1f (%.1f s) + %d of %d constraint rows "
                      "scaled to %d points over %d threads (%.1f s) + STIR over %d points (%.2f s)" %
                      (fill, cols, tr.shape[0], tr.shape[0] / cols, lde_s, pts, n0, n0, T, merge_s, n0, fri_s)}


def cpu_small_proofs(O):
    """One whole proof of each small-proof line with the C oracle (single
    thread), as proofs/s: the CPU side of the plain-vs-batched lines."""
    import hashlib
    import random
    out = {}
    key = splitmix_field(0x5EED0001, 8).astype("<u8").tobytes()
    t = time.perf_counter()
    O.prove_ds_ksk_binding(key, hashlib.sha256(key).digest(), b"\xAA" * 32)
    out["ds_ksk"] = round(1.0 / (time.perf_counter() - t), 3)
    rng = random.Random(0x5EED0006)
    m = rng.getrandbits(2048) | 1 | (1 << 2047)
    sig = rng.randrange(m)
    t = time.perf_counter()
    O.prove_rsa_exp(m, sig, pow(sig, 65537, m), bytes(32))
    out["rsa"] = round(1.0 / (time.perf_counter() - t), 3)
    leaves = O.zone_leaf_hashes(0, 1024, b"swarm-test-zone1")
    t = time.perf_counter()
    O.prove_inner_shard(b"swarm-test-zone1", leaves, bytes(32))
    out["inner_shard"] = round(1.0 / (time.perf_counter() - t), 3)
    out["p256"] = round(1.0 / cpu_p256_seconds(O), 5)
    out["unit"] = ("proofs/s in the reference threading on %d threads, one proof each (64-byte DS-KSK key, one "
                   "RSA-2048 record, 1024-record shard; one ECDSA-P256 record estimated from a sample, see "
                   "p256_sample)" % O.lib().ref_get_threads())
    out["p256_sample"] = ("the full 256-row trace, the LDE of 2048 of the 151 334 columns (x73.9, one column at a "
                          "time), the 155 774 constraints at 64 of the 8192 points (x128, divided over the "
                          "threads) and one STIR proof over 8192 points, summed; the oracle's whole proof takes "
                          "about 140 s on one core")
    return out


def cpu_p256_seconds(O):
    """Seconds the C oracle (one thread) needs for one ECDSA-P256 record proof,
    from a bounded sample: trace fill in full, the LDE of 2048 columns scaled
    to 151 334, the constraint evaluation at 64 points scaled to 8192, and one
    STIR proof over 8192 points."""
    from stark_swarm_amd.zone import synthetic_ecdsa_records
    r = synthetic_ecdsa_records(1, 0x5EED000B)[0]
    t = time.perf_counter()
    tr, _ = O.p256_ecdsa_trace(r.qx, r.qy, r.r, r.s, r.digest)
    fill = time.perf_counter() - t
    cols = 2048
    t = time.perf_counter()
    lde = O.lde_columns(tr[:cols], 32)
    lde_s = (time.perf_counter() - t) * tr.shape[0] / cols
    del lde
    pts = 64
    rows = np.ascontiguousarray(tr[:, :2].T)
    t = time.perf_counter()
    for i in range(pts):
        O.p256_constraints(rows[0], rows[1], i % 255, 256)
    merge_s = (time.perf_counter() - t) * 8192 / pts / O.lib().ref_get_threads()  # the merge's par_iter
    sched = [8, 8, 8, 8, 2]
    c = splitmix_field(0x5EED000C, 8192)
    t = time.perf_counter()
    O.deep_fri_prove(c, O.fri_params(sched, r=54, s0=54, pi_hash=bytes(32)))
    fri_s = time.perf_counter() - t
    return fill + lde_s + merge_s + fri_s


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _profile_line(tag, kernel):
    path = os.path.join(ROOT, "profiles", "hbm_traffic__prover_core_2e20x128.md")
    try:
        for line in open(path):
            f = line.split()
            if len(f) >= 3 and f[0] == tag and f[1] == kernel:
                return [float(x) for x in f[2:]]
    except OSError:
        pass
    return None


def profiled_traffic(kernel):
    """HBM bytes per launch (read + write) of `kernel` on the 2^20 x 128 workload
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary, or None."""
    v = _profile_line("TRAFFIC", kernel)
    return v[0] + v[1] if v else None


def profiled_valu_load(kernel):
    """VALU issue load of `kernel` (SQ_INSTS_VALU x 4 cycles over the SIMD
    cycles of its launch) from the same committed summary, or None."""
    v = _profile_line("VALU_LOAD", kernel)
    return v[0] if v else None


def calib_copy(nbytes=1 << 30):
    """A plain device copy of known size, for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch
    a = torch.empty(nbytes // 8, dtype=torch.int64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log("calib copy: 3 x %d bytes" % nbytes)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log-rows", type=int, default=20)
    ap.add_argument("--cols", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", action="store_true")
    ap.add_argument("--streams", type=int, default=6,
                    help="proofs in flight for the batched measurement (2 / 3 / 4 / 6 / 8 measured 27.9 / 27.5 / "
                         "28.2 / 29.1 / 29.0 proofs/s on MI355X)")
    ap.add_argument("--shards", type=int, default=128, help="inner-shard proofs for the plain-vs-batched line (0 = skip)")
    ap.add_argument("--dsksk", type=int, default=64, help="DS-KSK (SHA-256 AIR) proofs for the plain-vs-batched line (0 = skip)")
    ap.add_argument("--rsa", type=int, default=64, help="RSA-2048 record proofs for the plain-vs-batched line (0 = skip)")
    ap.add_argument("--p256", type=int, default=16, help="ECDSA-P256 record proofs for the plain-vs-batched line (0 = skip)")
    ap.add_argument("--ed25519", type=int, default=64,
                    help="Ed25519 signatures for the C3 plain-vs-batched lines (blowup 8 with 90 queries, and the "
                         "reference's blowup 32 with 54) (0 = skip)")
    ap.add_argument("--serve", default="100,200,300,400",
                    help="serving: Poisson arrival rates (requests/s), '' = skip")
    ap.add_argument("--serve-requests", type=int, default=300, help="serving: requests per rate")
    ap.add_argument("--sha-batch", default="256,10,17",
                    help="batched uneven SHA-256/512 proofs: count,log2 min rows,log2 max rows ('' = skip)")
    ap.add_argument("--only", choices=["shards", "ds_ksk", "rsa", "p256", "sha_batch", "serve", "ed25519"],
                    help="run only this plain-vs-batched line (for profiling it) and print its JSON")
    ap.add_argument("--shards-only", action="store_true",
                    help="run only the inner-shard plain-vs-batched measurement (for profiling that path)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: contexts time the pass splits of each transform shape on first use (in the warm-ups) "
                         "and keep the fastest; 0: default splits")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the prover core runs on a single MI355X and the library has no "
              "multi-GPU path (independent proofs shard as one process per GPU). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if args.calib_copy:
        calib_copy()
        return

    import stark_swarm_amd as S
    S.set_autotune_default(bool(args.autotune))
    if args.shards_only:
        args.only = "shards"
    if args.only:
        if args.only == "shards":
            line = shard_batch(S, args.streams, max(1, args.shards))
        elif args.only == "ds_ksk":
            line = ds_ksk_batch(S, max(1, args.dsksk))
        elif args.only == "rsa":
            line = rsa_batch(S, max(1, args.rsa))
        elif args.only == "p256":
            line = p256_batch(S, max(1, args.p256))
        elif args.only == "ed25519":
            line = {"ed25519": ed25519_batch(S, max(1, args.ed25519), 8, 90)}
        elif args.only == "sha_batch":
            cnt, lo, hi = (int(x) for x in args.sha_batch.split(","))
            line = c4_batch(S, cnt, lo, hi)
        else:
            line = serve(S, [int(r) for r in args.serve.split(",")], args.serve_requests)
        print(json.dumps(line))
        return
    ctx = S.default_context()
    w, lg, blowup = args.cols, args.log_rows, 8
    n = 1 << lg
    n0 = n * blowup
    schedule = [8, 8, 8, 8, 8, 4] if lg == 20 else S.make_schedule(n0)
    s0 = 80
    params = S.DeepFriParams(schedule, s0, S.SEED_Z, True, 1, True, s0, None)
    log("bench: generating %d x 2^%d trace" % (w, lg))
    trace = splitmix_field(0x5EED0002, w * n).reshape(w, n)
    d_tr = ctx.alloc(trace.nbytes).upload(trace)
    d_lde = ctx.alloc(trace.nbytes * blowup)
    del trace

    ref_blob = None
    for i in range(args.warmup):
        t = time.perf_counter()
        root, blob = S.prover_core(d_tr, w, n, blowup, 7, params, d_lde)
        ref_blob = blob
        log("warmup %d: %.1f ms, %d proof bytes" % (i, 1e3 * (time.perf_counter() - t), len(blob)))

    ctx.profile_reset()
    ctx.profile(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        root, blob = S.prover_core(d_tr, w, n, blowup, 7, params, d_lde)
        if ref_blob is not None and blob != ref_blob:
            raise SystemExit("non-deterministic proof bytes at step %d" % i)
        ref_blob = blob
        log("step %d done" % i)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ctx.profile(False)
    stats = ctx.profile_stats()
    # the plans the warm-up tuned for this workload's transforms
    plans = {"%s 2^%d x%d, 2^%d cols" % (k[0], k[1], 1 << k[2], k[3]):
             {"split": [best[0], k[1] - best[0]] if best[0] < k[1] else [best[0]], "tile_log": best[1],
              "ms": {"%d/tile%d" % p: round(v, 4) for p, v in sorted(ms.items())}}
             for k, (best, ms) in sorted(S.autotune_report().items())}

    ms_step = 1e3 * dt / args.steps
    # stage accounting
    gi = 1 << 30
    merkle_ms = sum(stats.get(k, {}).get("ms", 0.0) for k in ["merkle_leaves", "merkle_nodes"]) / args.steps
    host_stats = {k: v for k, v in stats.items() if k.startswith("host:")}
    stats = {k: v for k, v in stats.items() if not k.startswith("host:")}
    # The dominant stage is the coset LDE: pass A (coefficients -> the B
    # pre-twiddled column DFTs, scratch) and pass B (row DFTs -> the LDE),
    # one launch each per proof. Its algorithmic bytes are the coefficients
    # read once and the LDE written once; the scratch round trip is traffic,
    # not work.
    def stage(names, alg_bytes):
        ms = sum(stats.get(k, {}).get("ms", 0.0) for k in names) / args.steps
        return {"kernels": names, "ms": round(ms, 3), "algorithmic_bytes": alg_bytes,
                "GBps": round(alg_bytes / (ms * 1e-3) / 1e9, 1) if ms > 0 else None,
                "frac": round(alg_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if ms > 0 else None}
    stages = {"intt": stage(["intt_passA", "intt_passB", "intt_small"], 16.0 * w * n),
              "lde": stage(["lde_passA", "lde_passB", "lde_small"], 8.0 * w * n + 8.0 * w * n0),
              "merkle": stage(["merkle_leaves", "merkle_nodes"], 8.0 * w * n0 + 32.0 * n0)}
    lde_launch_ms = stages["lde"]["ms"]
    lde_alg = stages["lde"]["algorithmic_bytes"]
    achieved = lde_alg / (lde_launch_ms * 1e-3) / 1e9 if lde_launch_ms > 0 else 0.0
    full = (lg, w) == (20, 128)
    traffic = None
    if full:
        ta, tb = profiled_traffic("lde_passA"), profiled_traffic("lde_passB")
        traffic = ta + tb if ta is not None and tb is not None else None
    kernels = {k: {"ms_per_step": round(v["ms"] / args.steps, 3), "launches_per_step": v["launches"] / args.steps,
                   "GBps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None}
               for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"])}
    res = {
        "metric": METRIC,
        "value": round(1e3 / ms_step, 4),
        "unit": "proofs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (Goldilocks)",
        "data": "synthetic",
        "config": {"workload": "prover-core microbench: 2^%d x %d Goldilocks trace, iNTT, coset LDE x%d (shift 7), "
                               "SHA3 row Merkle over 2^%d rows, STIR fold %s, %d queries" %
                               (lg, w, blowup, lg + 3, schedule, s0),
                   "rows": n, "cols": w, "blowup": blowup, "schedule": schedule, "queries": s0},
        "lde_GBps": stages["lde"]["GBps"],
        "lde_ms": stages["lde"]["ms"],
        "merkle_ms": round(merkle_ms, 3),
        "proof_bytes": len(ref_blob),
        "roofline": {"bound": "valu", "kernel": "coset LDE stage: lde_passA + lde_passB (ntt_tile, one launch "
                                                "each per proof)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": "profiles/hbm_traffic__prover_core_2e20x128.md: FETCH_SIZE x2 + WRITE_SIZE "
                                       "x1 (calibrated on a 1 GiB copy) of lde_passA + lde_passB",
                     "algorithmic_bytes_per_launch": lde_alg, "avg_launch_ms": round(lde_launch_ms, 4),
                     "valu_issue_load": {k: profiled_valu_load(k) for k in ("lde_passA", "lde_passB")} if full
                     else None,
                     "note": "algorithmic bytes = the 2^20 x 128 coefficients read once + the 2^23 x 128 LDE written "
                             "once, over the two passes' time; bound from the counters: integer VALU work (64-bit "
                             "modular products and lazily reduced butterflies on 32-bit lanes, 274 and 218 "
                             "instructions per LDE element) issued at the valu_issue_load above (about one "
                             "instruction per 4 SIMD cycles: both passes are issue-bound); their HBM traffic (the "
                             "scratch round trip included) stays far below peak"},
        "stages": stages,
        "kernels": kernels,
        "autotune": {"enabled": bool(args.autotune), "plans": plans,
                     "note": "split = log2 sizes of the two passes (one entry: one pass with the cosets as "
                             "repetitions), tile_log = log2 of the LDS tile (13: 512 lanes, 14: 1024 lanes); "
                             "ms = one timed run per candidate (split/tile) in the warm-up"},
    }
    if args.streams > 1:
        bres = batched(S, args, w, n, blowup, params, ref_blob, ms_step)
        res.update(bres)
        # headline value: batched throughput (K proofs in flight); the plain
        # one-proof-at-a-time rate stays in batched.plain_proofs_per_s
        b = bres["batched"]
        res["value"] = b["proofs_per_s"]
        res["ms_per_step"] = round(1e3 * b["seconds"] / args.steps, 3)
        res["config"]["proofs_in_flight"] = args.streams
        res["config"]["step"] = "%d concurrent prover-core proofs (one per HIP stream)" % args.streams
    if args.shards:
        log("inner shards: plain vs batched")
        res.update(shard_batch(S, args.streams, args.shards))
    if args.dsksk:
        log("ds-ksk (SHA-256 AIR): plain vs batched")
        res.update(ds_ksk_batch(S, args.dsksk))
    if args.rsa:
        log("rsa-2048 records: plain vs batched")
        res.update(rsa_batch(S, args.rsa))
    if args.p256:
        log("ecdsa-p256 records: plain vs batched")
        res.update(p256_batch(S, args.p256))
    if args.ed25519:
        log("ed25519 (C3): plain vs batched, blowup 8 / 90 queries and blowup 32 / 54 queries")
        res["ed25519"] = ed25519_batch(S, args.ed25519, 8, 90)
        res["ed25519_reference_params"] = ed25519_batch(S, args.ed25519, 32, 54)
    if args.sha_batch:
        cnt, lo, hi = (int(x) for x in args.sha_batch.split(","))
        log("sha batch: %d uneven SHA-256/512 proofs, 2^%d..2^%d rows" % (cnt, lo, hi))
        res.update(c4_batch(S, cnt, lo, hi))
    if args.serve:
        log("serving: Poisson arrivals")
        res.update(serve(S, [int(r) for r in args.serve.split(",")], args.serve_requests))
    if not args.no_cpu:
        log("cpu baseline: timing the oracle on a bounded sample (%d threads)" % cpu_threads())
        res["cpu_baseline"] = cpu_baseline(w, lg, blowup, schedule, s0)
        if args.ed25519:
            log("cpu: ed25519 proof estimate")
            res["ed25519"]["cpu"] = cpu_ed25519(8, 90)
            res["ed25519_reference_params"]["cpu"] = cpu_ed25519(32, 54)
        if args.sha_batch:
            log("cpu: sha batch sample")
            a_, m_, h_, f_ = c4_jobs(*(int(x) for x in args.sha_batch.split(",")))
            res["sha_batch"]["cpu"] = cpu_c4(list(zip(a_, m_, h_, f_)))
    print(json.dumps(res))


def shard_batch(S, K, nproofs, records_per_shard=1024):
    """prove_inner_shard (HashRollup, n0 = 2^17 at 1024 records) on distinct
    synthetic shards: one proof per call (plain) vs all of them as one
    lockstep device batch. Every batched proof must equal the plain proof of
    the same shard."""
    from stark_swarm_amd.zone import build_synthetic_zone
    salt = b"swarm-test-zone1"
    zone = build_synthetic_zone(records_per_shard * 4)
    jobs = []
    for i in range(nproofs):
        off = (i * 97) % (len(zone) - records_per_shard)
        fsb = S.shard_fs_binding(b"\x11" * 32, b"\x22" * 32, i + 1, i.to_bytes(32, "little"))
        jobs.append((zone[off:off + records_per_shard], fsb))
    ctx0 = S.default_context()
    # records are marshaled into C arrays before the clock starts, for both legs
    singles = [S.PackedShards([recs]) for recs, _ in jobs]
    S.prove_inner_shard_batch(salt, singles[0], [jobs[0][1]], ctx=ctx0)  # warm
    t0 = time.perf_counter()
    lat = []
    plain = []
    for i, (recs, fsb) in enumerate(jobs):
        t = time.perf_counter()
        plain.append(S.prove_inner_shard_batch(salt, singles[i], [fsb], ctx=ctx0)[0].proof_blob)
        lat.append(time.perf_counter() - t)
    plain_dt = time.perf_counter() - t0
    # lockstep batch: all shards in one stark_prove_inner_shard_batch call
    packed = S.PackedShards([j[0] for j in jobs])
    fsbs = [j[1] for j in jobs]
    S.prove_inner_shard_batch(salt, packed, fsbs, ctx=ctx0)  # warm: the batch's buffers, once
    t0 = time.perf_counter()
    outs = S.prove_inner_shard_batch(salt, packed, fsbs, ctx=ctx0)
    ldt = time.perf_counter() - t0
    exact = all(outs[i].proof_blob == plain[i] for i in range(nproofs))
    if not exact:
        raise SystemExit("lockstep-batched inner-shard proofs differ from the plain proofs")
    lat.sort()
    return {"inner_shard": {"records": records_per_shard, "n0": 1 << 17, "proofs": nproofs,
                            "plain_proofs_per_s": round(nproofs / plain_dt, 2),
                            "batched_proofs_per_s": round(nproofs / ldt, 2),
                            "speedup": round(plain_dt / ldt, 3),
                            "batch": "one lockstep device batch (stark_prove_inner_shard_batch)",
                            "plain_latency_ms_p50": round(1e3 * lat[len(lat) // 2], 3),
                            "plain_latency_ms_p99": round(1e3 * lat[min(len(lat) - 1, (99 * len(lat)) // 100)], 3),
                            "bit_exact_vs_plain": exact}}


def ds_ksk_batch(S, nproofs):
    """prove_ds_ksk_binding at C1: 64-byte messages (2 SHA-256 blocks, n_trace
    256, n0 = 8192, 756 columns, 766 constraints) from SplitMix64 seed
    0x5EED0001, fs_binding = [0xAA; 32]: one proof per call vs one batch call.
    Every batched proof must equal the single-call proof of the same key."""
    import hashlib
    raw = splitmix_field(0x5EED0001, 8 * nproofs).astype("<u8").tobytes()
    keys = [raw[64 * i:64 * (i + 1)] for i in range(nproofs)]
    parents = [hashlib.sha256(k).digest() for k in keys]
    fsb = [b"\xAA" * 32] * nproofs
    S.prove_ds_ksk_binding(keys[0], parents[0], fsb[0])  # warm
    lat, plain = [], []
    t0 = time.perf_counter()
    for k, pa, f in zip(keys, parents, fsb):
        t = time.perf_counter()
        plain.append(S.prove_ds_ksk_binding(k, pa, f).proof_blob)
        lat.append(time.perf_counter() - t)
    plain_dt = time.perf_counter() - t0
    S.prove_ds_ksk_binding_batch(keys, parents, fsb)  # warm: the batch's buffers, once
    t0 = time.perf_counter()
    outs = S.prove_ds_ksk_binding_batch(keys, parents, fsb)
    bdt = time.perf_counter() - t0
    exact = all(o.proof_blob == p for o, p in zip(outs, plain)) and all(o.asserted_digest == pa for o, pa in
                                                                        zip(outs, parents))
    if not exact:
        raise SystemExit("batched DS-KSK proofs differ from the single-call proofs")
    lat.sort()
    return {"ds_ksk": {"message_bytes": 64, "n_blocks": outs[0].n_blocks, "n_trace": outs[0].n_trace,
                       "n0": outs[0].n_trace * 32, "proofs": nproofs,
                       "plain_proofs_per_s": round(nproofs / plain_dt, 2),
                       "batched_proofs_per_s": round(nproofs / bdt, 2), "speedup": round(plain_dt / bdt, 3),
                       "plain_latency_ms_p50": round(1e3 * lat[len(lat) // 2], 3),
                       "plain_latency_ms_p99": round(1e3 * lat[min(len(lat) - 1, (99 * len(lat)) // 100)], 3),
                       "bit_exact_vs_plain": exact}}


def rsa_records(S, n, seed):
    """n RSA-2048 records: odd 2048-bit moduli with the top bit set,
    signatures s < n and em = s^65537 mod n, from a seeded generator (the
    proving cost does not depend on n being a product of two primes)."""
    import random
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        m = rng.getrandbits(2048) | 1 | (1 << 2047)
        sig = rng.randrange(m)
        out.append(S.RsaStackedRecord(m, sig, pow(sig, 65537, m)))
    return out


def rsa_batch(S, nproofs):
    """RSA-2048 signature records (the per-record prover of the
    mixed-algorithm zone example: 11 041 columns, 11 280 constraints,
    n_trace 32, n0 = 1024): one proof per call vs one batch call, every
    batched proof equal to its single-call proof."""
    import hashlib
    recs = rsa_records(S, nproofs, 0x5EED0006)
    pis = [hashlib.sha3_256(b"rsa-pis-%d" % i).digest() for i in range(nproofs)]
    S.prove_rsa_exp_batch(recs[:1], pis[:1])  # warm
    lat, plain = [], []
    t0 = time.perf_counter()
    for r, p in zip(recs, pis):
        t = time.perf_counter()
        plain.append(S.prove_rsa_exp_batch([r], [p])[0].proof_blob)
        lat.append(time.perf_counter() - t)
    plain_dt = time.perf_counter() - t0
    S.prove_rsa_exp_batch(recs, pis)  # warm: the batch's buffers, once
    t0 = time.perf_counter()
    outs = S.prove_rsa_exp_batch(recs, pis)
    bdt = time.perf_counter() - t0
    exact = all(o.proof_blob == p for o, p in zip(outs, plain))
    if not exact:
        raise SystemExit("batched RSA proofs differ from the single-call proofs")
    lat.sort()
    return {"rsa": {"records_per_proof": 1, "n_trace": 32, "n0": 1024, "columns": 11041, "constraints": 11280,
                    "proofs": nproofs, "plain_proofs_per_s": round(nproofs / plain_dt, 2),
                    "batched_proofs_per_s": round(nproofs / bdt, 2), "speedup": round(plain_dt / bdt, 3),
                    "plain_latency_ms_p50": round(1e3 * lat[len(lat) // 2], 3),
                    "plain_latency_ms_p99": round(1e3 * lat[min(len(lat) - 1, (99 * len(lat)) // 100)], 3),
                    "bit_exact_vs_plain": exact}}


def p256_batch(S, nproofs):
    """ECDSA-P256 signature records (the per-record prover of the
    mixed-algorithm zone example: two 256-step scalar-multiplication chains,
    151 334 columns, 155 774 constraints, n_trace 256, n0 = 8192): one proof
    per call vs one batch call, every batched proof equal to its single-call
    proof."""
    from stark_swarm_amd.zone import synthetic_ecdsa_records
    recs = synthetic_ecdsa_records(nproofs, 0x5EED0008)
    S.prove_record_ecdsa_batch(recs[:1])  # warm
    lat, plain = [], []
    t0 = time.perf_counter()
    for r in recs:
        t = time.perf_counter()
        plain.append(S.prove_record_ecdsa_batch([r])[0].proof_blob)
        lat.append(time.perf_counter() - t)
    plain_dt = time.perf_counter() - t0
    S.prove_record_ecdsa_batch(recs)  # warm: the batch's buffers, once
    t0 = time.perf_counter()
    outs = S.prove_record_ecdsa_batch(recs)
    bdt = time.perf_counter() - t0
    exact = all(o.proof_blob == p for o, p in zip(outs, plain))
    if not exact:
        raise SystemExit("batched ECDSA-P256 proofs differ from the single-call proofs")
    lat.sort()
    lde_bytes = 8.0 * 151334 * (256 + 8192)
    return {"p256": {"n_trace": 256, "n0": 8192, "columns": 151334, "constraints": 155774,
                     "note": "reference_lde_bytes_per_proof is the blowup-32 LDE the reference reads in its merge; "
                             "this library composes the gadget constraints on a blowup-2 LDE and interpolates",
                     "proofs": nproofs, "plain_proofs_per_s": round(nproofs / plain_dt, 2),
                     "batched_proofs_per_s": round(nproofs / bdt, 2), "speedup": round(plain_dt / bdt, 3),
                     "plain_latency_ms_p50": round(1e3 * lat[len(lat) // 2], 3),
                     "plain_latency_ms_p99": round(1e3 * lat[min(len(lat) - 1, (99 * len(lat)) // 100)], 3),
                     "reference_lde_bytes_per_proof": lde_bytes, "bit_exact_vs_plain": exact}}


def ed25519_batch(S, nproofs, blowup, queries):
    """prove_zsk_ksk_binding_v2 (crates/swarm-dns/src/prover.rs:856-951) on
    nproofs Ed25519 signatures from seeded keys over seeded random 32-256
    byte messages: the v16 verify AIR (k_scalar 256, 41 519 columns, 70 606
    constraints, 1024 or 2048 rows), one proof per call vs one batch call
    (lockstep per trace shape), every batched proof equal to its single-call
    proof."""
    from stark_swarm_amd.zone import synthetic_ed25519_signatures
    sigs = synthetic_ed25519_signatures(nproofs, 0x5EED0003)
    reqs = [(pk, sig, msg, bytes([i & 0xFF]) * 32, bytes(32)) for i, (pk, sig, msg) in enumerate(sigs)]
    S.prove_zsk_ksk_binding_v2_batch(reqs[:1], blowup=blowup, queries=queries)  # warm
    lat, plain = [], []
    t0 = time.perf_counter()
    for r in reqs:
        t = time.perf_counter()
        plain.append(S.prove_zsk_ksk_binding_v2_batch([r], blowup=blowup, queries=queries)[0].proof_blob)
        lat.append(time.perf_counter() - t)
    plain_dt = time.perf_counter() - t0
    S.prove_zsk_ksk_binding_v2_batch(reqs, blowup=blowup, queries=queries)  # warm: the batch's buffers, once
    t0 = time.perf_counter()
    outs = S.prove_zsk_ksk_binding_v2_batch(reqs, blowup=blowup, queries=queries)
    bdt = time.perf_counter() - t0
    exact = all(o.proof_blob == p for o, p in zip(outs, plain))
    if not exact:
        raise SystemExit("batched Ed25519 proofs differ from the single-call proofs")
    lat.sort()
    heights = sorted({o.n_trace for o in outs})
    return {"proofs": nproofs, "k_scalar": 256, "blowup": blowup, "queries": queries, "trace_rows": heights,
            "columns": 41519, "constraints": 70606, "message_bytes": [32, 256],
            "plain_proofs_per_s": round(nproofs / plain_dt, 2), "batched_proofs_per_s": round(nproofs / bdt, 2),
            "speedup": round(plain_dt / bdt, 3), "plain_latency_ms_p50": round(1e3 * lat[len(lat) // 2], 3),
            "plain_latency_ms_p99": round(1e3 * lat[min(len(lat) - 1, (99 * len(lat)) // 100)], 3),
            "reference_lde_bytes_per_proof": 8.0 * 41519 * 1024 * blowup, "bit_exact_vs_plain": exact}


def c4_jobs(nproofs, log_lo, log_hi):
    """The C4 jobs: alternating SHA-256 / SHA-512, seeded 64-4096 byte
    messages, heights 2^e for e uniform in [log_lo, log_hi] raised to each
    message's minimum, seeded 32-byte bindings."""
    rng = np.random.default_rng(0x5EED0004)
    algs, msgs, hs = [], [], []
    for i in range(nproofs):
        alg = 256 if i % 2 == 0 else 512
        m = rng.bytes(int(rng.integers(64, 4097)))
        blocks = (len(m) + 9 + 63) // 64 if alg == 256 else (len(m) + 17 + 127) // 128
        hmin = 1 << max(7, (128 * blocks - 1).bit_length())
        h = max(hmin, 1 << int(rng.integers(log_lo, log_hi + 1)))
        algs.append(alg)
        msgs.append(m)
        hs.append(h)
    fsb = [rng.bytes(32) for _ in range(nproofs)]
    return algs, msgs, hs, fsb


def c4_batch(S, nproofs, log_lo, log_hi):
    """Batched uneven SHA-256/512 proofs: nproofs alternating SHA-256 and
    SHA-512 AIR proofs (the prove_ds_ksk_binding pipeline: LDE x32, merge,
    STIR with 54 queries), messages of seeded uniform 64-4096 bytes, trace
    heights log-uniform in [2^log_lo, 2^log_hi] raised to each message's
    minimum (padding rows repeat the last useful row). One proof per call
    (plain) vs one stark_prove_sha_batch call (lockstep per alg and height,
    HBM-sized waves); every batched proof must equal its single-call proof."""
    algs, msgs, hs, fsb = c4_jobs(nproofs, log_lo, log_hi)
    lde_bytes = sum(8 * (756 if a == 256 else 1510) * 32 * h for a, h in zip(algs, hs))
    S.prove_sha_batch([256, 512], [b"w", b"w"], fsb[:2])  # warm
    # one untimed single proof per (alg, height): the first call of a shape
    # tunes its transform plans when autotune is on
    shapes = {}
    for i in range(nproofs):
        shapes.setdefault((algs[i], hs[i]), i)
    for i in shapes.values():
        S.prove_sha_batch(algs[i], [msgs[i]], [fsb[i]], heights=[hs[i]])
    t0 = time.perf_counter()
    plain = []
    for i in range(nproofs):
        plain.append(S.prove_sha_batch(algs[i], [msgs[i]], [fsb[i]], heights=[hs[i]])[0].proof_blob)
        if (i + 1) % 16 == 0:
            log("sha batch: %d/%d plain proofs, %.1f s" % (i + 1, nproofs, time.perf_counter() - t0))
    plain_dt = time.perf_counter() - t0
    S.prove_sha_batch(algs, msgs, fsb, heights=hs)  # warm: the batch's buffers, once
    t0 = time.perf_counter()
    outs = S.prove_sha_batch(algs, msgs, fsb, heights=hs)
    bdt = time.perf_counter() - t0
    exact = all(o.proof_blob == p for o, p in zip(outs, plain))
    if not exact:
        raise SystemExit("batched SHA proofs differ from the single-call proofs")
    return {"sha_batch": {"proofs": nproofs, "algs": "alternating SHA-256 / SHA-512",
                          "trace_rows_log2": [log_lo, log_hi], "message_bytes": [64, 4096],
                          "reference_lde_GB": round(lde_bytes / 1e9, 2),
                          "plain_proofs_per_s": round(nproofs / plain_dt, 2),
                          "batched_proofs_per_s": round(nproofs / bdt, 2), "speedup": round(plain_dt / bdt, 3),
                          "bit_exact_vs_plain": exact,
                          "note": "reference_lde_GB is the blowup-32 LDE the reference writes for these proofs; "
                                  "this library composes the gate-group polynomials on a blowup-2 LDE and "
                                  "interpolates them, so it never builds that LDE"}}


def serve(S, rates, per_rate, workers=4, max_batch=32):
    """Serving: a synthetic signed zone of 10 000 RRsets (owner names of 2-4
    labels of 3-12 lowercase letters; 60 % A / 25 % AAAA / 15 % TXT with
    16-255 byte TXT rdata; signatures 50 % Ed25519 / 30 % ECDSA-P256 / 20 %
    RSA-2048 from seeded keys). Each request is the proof of one RRset's
    signature -- the per-record proof of the mixed-algorithm zone example
    (Ed25519 verify AIR v16 at k_scalar 256, the P-256 ECDSA double-multirow
    AIR, the RSA-2048 exponentiation AIR; blowup 32, 54 STIR queries, the
    record's pk_hash_pre as public-input hash). Requests arrive as a Poisson
    process at each rate (requests/s) and go through the proving server
    (HBM-budget admission, continuous batching over `workers` streams).
    Latency = completion - arrival. A sample of results is checked
    byte-equal to the same request proven alone."""
    import threading

    from stark_swarm_amd.server import ProvingServer
    from stark_swarm_amd.zone import ZoneSigner, synthetic_rrset_zone
    rng = np.random.default_rng(0x5EED0005)
    log("serve: zone of 10000 RRsets, seeded keys")
    zone = synthetic_rrset_zone(10000, 0x5EED000C)
    signer = ZoneSigner(0x5EED000D)
    root0 = bytes(32)  # the record proof has no merkle root; it enters only the binding's own pi_hash
    out = []
    # the workers take the plans tuned so far but time none themselves:
    # tuning a new batch shape mid-stream would stall the requests behind it
    S.set_autotune_default(False)

    def submit(srv, q):
        if q[0] == "ed25519":
            return srv.submit_ed25519(q[1], q[2], q[3], q[4], root0)
        if q[0] == "p256":
            return srv.submit_p256(q[1])
        return srv.submit_rsa(q[1], q[2])

    def alone(q):
        if q[0] == "ed25519":
            return S.prove_zsk_ksk_binding_v2(q[1], q[2], q[3], q[4], root0)
        if q[0] == "p256":
            return S.prove_record_ecdsa_batch([q[1]])[0]
        return S.prove_rsa_exp_batch([q[1]], [q[2]])[0]

    with ProvingServer(workers=workers, hbm_budget=64 << 30, max_batch=max_batch) as srv:
        # warm every worker's context once per kind and Ed25519 trace shape (SHA-512 blocks of R || A || M)
        warm = {}
        for r in zone[:3000]:
            shape = (r.alg, (64 + len(r.message()) + 17 + 127) // 128 if r.alg == 15 else 0)
            if shape not in warm:
                warm[shape] = signer.request(r)
        for q in warm.values():
            for rid in [submit(srv, q) for _ in range(workers)]:
                srv.wait(rid)
        for rate in rates:
            n = per_rate
            picks = rng.integers(0, len(zone), size=n)
            log("serve rate %d/s: signing %d RRsets" % (rate, n))
            reqs = [signer.request(zone[int(i)]) for i in picks]
            gaps = rng.exponential(1.0 / rate, size=n)
            ids = [None] * n
            res = [None] * n

            def collect(lo, hi):
                for j in range(lo, hi):
                    while ids[j] is None:
                        time.sleep(0.0005)
                    res[j] = srv.wait(ids[j], timeout=600)

            waiters = [threading.Thread(target=collect, args=(q * n // 4, (q + 1) * n // 4)) for q in range(4)]
            for t in waiters:
                t.start()
            t0 = time.perf_counter()
            due = t0
            for j, q in enumerate(reqs):
                due += gaps[j]
                d = due - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids[j] = submit(srv, q)
            for t in waiters:
                t.join()
            wall = time.perf_counter() - t0

            def pct(v, f):
                v = sorted(v)
                return round(1e3 * v[min(len(v) - 1, int(f * len(v)))], 2)
            lat = [r.latency for r in res]
            line = {"rate_per_s": rate, "requests": n, "proofs_per_s": round(n / wall, 1),
                    "latency_ms_p50": pct(lat, 0.5), "latency_ms_p99": pct(lat, 0.99),
                    "mean_batch": round(sum(r.batch_size for r in res) / n, 2)}
            for kind in ("ed25519", "p256", "rsa"):
                lk = [r.latency for q, r in zip(reqs, res) if q[0] == kind]
                if lk:
                    line[kind] = {"requests": len(lk), "latency_ms_p50": pct(lk, 0.5), "latency_ms_p99": pct(lk, 0.99)}
            out.append(line)
            log("serve rate %d/s: %s" % (rate, line))
            # byte-equality of a sample against the request proven alone
            for j in range(0, n, max(1, n // 9)):
                a = alone(reqs[j])
                if a.proof_blob != res[j].proof_blob or a.pi_hash != res[j].pi_hash:
                    raise SystemExit("served proof %d (%s) differs from the proof made alone" % (j, reqs[j][0]))
        st = srv.stats()
    return {"serve": {"workers": workers, "max_batch": max_batch, "hbm_budget_GiB": 64,
                      "peak_admitted_GiB": round(st["peak_bytes_admitted"] / 2**30, 3),
                      "zone": "10000 RRsets: 60% A / 25% AAAA / 15% TXT (16-255 B), owner names of 2-4 labels of "
                              "3-12 lowercase letters",
                      "mix": "one RRset signature proof per request: 50% Ed25519 (verify AIR v16, k_scalar 256), "
                             "30% ECDSA-P256, 20% RSA-2048; blowup 32, 54 queries, seeded keys",
                      "rates": out, "sample_bit_exact_vs_alone": True}}


def batched(S, args, w, n, blowup, params, ref_blob, plain_ms):
    """K independent prover-core proofs in flight at once: one context (HIP
    stream) and one host thread per proof slot, each over its own resident
    trace (the same seeded trace, so every proof must equal the plain one)."""
    import threading
    K = args.streams
    log("batched: %d streams" % K)
    trace = splitmix_field(0x5EED0002, w * n).reshape(w, n)
    slots = []
    for k in range(K):
        ctx = S.Context(0)
        slots.append((ctx, ctx.alloc(trace.nbytes).upload(trace), ctx.alloc(trace.nbytes * blowup)))
    del trace
    per = max(1, args.steps)
    errors = []

    def worker(k, count):
        ctx, d_tr, d_lde = slots[k]
        try:
            for _ in range(count):
                _, blob = S.prover_core(d_tr, w, n, blowup, 7, params, d_lde, ctx=ctx)
                if blob != ref_blob:
                    errors.append("slot %d: proof differs from the plain proof" % k)
        except Exception as e:  # noqa: BLE001
            errors.append("slot %d: %r" % (k, e))

    # warm each slot once (allocator, code objects)
    ths = [threading.Thread(target=worker, args=(k, 1)) for k in range(K)]
    [t.start() for t in ths]
    [t.join() for t in ths]
    if errors:
        raise SystemExit("; ".join(errors))
    t0 = time.perf_counter()
    ths = [threading.Thread(target=worker, args=(k, per)) for k in range(K)]
    [t.start() for t in ths]
    [t.join() for t in ths]
    dt = time.perf_counter() - t0
    if errors:
        raise SystemExit("; ".join(errors))
    for ctx, a, b in slots:
        a.free()
        b.free()
        ctx.close()
    total = K * per
    log("batched: %d proofs in %.3f s" % (total, dt))
    return {"batched": {"streams": K, "proofs": total, "seconds": round(dt, 4),
                        "proofs_per_s": round(total / dt, 4), "plain_proofs_per_s": round(1e3 / plain_ms, 4),
                        "speedup_vs_plain": round((total / dt) / (1e3 / plain_ms), 3),
                        "bit_exact_vs_plain": True}}


if __name__ == "__main__":
    main()
