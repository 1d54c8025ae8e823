"""Benchmark: Fibonacci Circle-STARK proof at log_n = 24 on one MI355X.

One step = one complete proof of the seeded Fibonacci statement (a0 = 1,
a1 = SplitMix64(42) >> 33): device trace generation, interpolate, LDE, trace /
quotient / FRI Blake2s commitments, device Fiat-Shamir, 80 queries,
decommitment and the download of the finished KRAK v1 proof bytes. The
prover's twiddle tables and buffers are built once before timing (the
reference's ProverCache, src/prover.rs:172-218).

Prints one JSON line (the driver's contract). `value` is milliseconds per
proof (lower is better). `roofline` covers the dominant kernel, the trace-tree
Blake2s subtree hash, which is VALU-bound; its duration comes from HIP events
recorded around that kernel on the prover's stream. `cpu_baseline` times the
C oracle (a restatement of the same algorithm producing the same bytes) on
this host: config A (log_n 16) median of 5 proofs single-threaded and with
OpenMP, and one full log_n = 24 proof with OpenMP (its `value`), each proof's
digest compared with the GPU's. Sub-objects time config A on the GPU, config
B with PoW 26 (KRAK v2) and config C (log_n 30, blowup 1, PoW 26).
"""
import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Hardware queues of this process, read by the HIP runtime when it starts
# (before the first HIP call below): the serving leg runs one stream per
# slot, and with HIP's default of 4 queues streams beyond the fourth share a
# queue and serialise. 16 queues for the 16 default slots measured 1.08x the
# proofs/s of 4 queues; the single-proof legs use one stream and do not care.
SERVING_HW_QUEUES = 16
if int(os.environ.get("GPU_MAX_HW_QUEUES") or 0) < SERVING_HW_QUEUES:
    os.environ["GPU_MAX_HW_QUEUES"] = str(SERVING_HW_QUEUES)

# Reference number for this metric: 133 ms on an RTX 5090 (README.md:37-43).
REFERENCE_MS = 133.0
# MI355X: 256 CUs x 64 int32 lanes x 2.4 GHz
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
HBM_PEAK_GBS = 8000.0
# VALU instructions of one Blake2s compression: 10 rounds x 8 G x
# (2 add3 + 2 add + 4 xor + 4 rotate)
OPS_PER_COMPRESSION = 10 * 8 * 12
# VALU instructions of one canonical M31 butterfly: product (64-bit multiply,
# fold, conditional subtract: 6), add (3) and subtract (3)
OPS_PER_BUTTERFLY = 12
MASK64 = (1 << 64) - 1
P = 0x7FFFFFFF


def fib_a1(seed):
    s = seed & MASK64
    while True:
        s = (s + 0x9E3779B97F4A7C15) & MASK64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        r = (z ^ (z >> 31)) >> 33
        if r < P:
            return r


PROFILE_SUMMARY = os.path.join(ROOT, "profiles", "fib_log24_prove_kernels.md")


def profiled_traffic():
    """HBM bytes per launch of the trace-tree kernel at log_n = 24, blowup 4,
    from the committed rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE
    runs of `bench.py --profile`, each scaled by what the counter reads for a
    1 GiB device copy); None if the summary is absent."""
    import re

    try:
        with open(PROFILE_SUMMARY) as f:
            m = re.search(r"traffic_bytes_per_launch:\s*(\d+)", f.read())
    except OSError:
        return None
    return int(m.group(1)) if m else None


def golden_digest(log_n, blowup, pow_bits):
    try:
        with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
            g = json.load(f)
    except OSError:
        return None
    for p in g.get("proofs_c", []):
        if (p["log_n"], p["blowup_bits"], p["pow_bits"]) == (log_n, blowup, pow_bits):
            return p["sha256"]
    return None


def host_info():
    """Host CPU and the oracle's compiler flags (oracle/Makefile)."""
    import platform
    import re
    import subprocess

    model = None
    try:
        with open("/proc/cpuinfo") as f:
            m = re.search(r"model name\s*:\s*(.+)", f.read())
            model = m.group(1).strip() if m else None
    except OSError:
        pass
    try:
        cc = subprocess.run(["gcc", "--version"], capture_output=True, text=True, timeout=10).stdout.splitlines()[0]
    except (OSError, subprocess.SubprocessError, IndexError):
        cc = "gcc"
    return {"nproc": os.cpu_count(), "cpu_model": model or platform.processor(),
            "compiler": f"{cc}: -O3 -march=x86-64-v2 (liboracle.so); the same plus -fopenmp (liboracle_omp.so)"}


def omp_threads():
    return int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)


def cpu_baseline(gpu_a, gpu_b_sha, runs=5):
    """The C oracle (the same algorithm and bytes as the GPU path) on this
    host: config A (log_n 16, blowup 4, 80 queries; PoW 0 and 10) as the
    median of `runs` proofs after one warm-up, single-threaded and with
    OpenMP; config B (log_n 24, PoW 0) once with OpenMP. `value` is config
    B's time, the same workload as the headline. Every CPU proof's SHA-256 is
    compared with the GPU proof of the same statement."""
    import statistics

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    a1 = fib_a1(42)
    threads = omp_threads()

    def timed(log_n, pw, omp):
        t0 = time.perf_counter()
        data = pyoracle.prove(1, a1, log_n, 2, 80, pw, omp=omp)
        return (time.perf_counter() - t0) * 1000.0, hashlib.sha256(data).hexdigest()

    config_a = {}
    for pw in (0, 10):
        row = {"runs": runs}
        for omp in (False, True):
            timed(16, pw, omp)
            ts, shas = zip(*(timed(16, pw, omp) for _ in range(runs)))
            row["omp_ms_median" if omp else "single_thread_ms_median"] = round(statistics.median(ts), 2)
            row["omp_sha256_matches_gpu" if omp else "single_thread_sha256_matches_gpu"] = all(
                h == gpu_a[pw]["proof_sha256"] for h in shas)
        row["gpu_ms"] = gpu_a[pw]["ms_per_proof"]
        row["speedup_gpu_vs_omp"] = round(row["omp_ms_median"] / row["gpu_ms"], 1)
        config_a[f"pow{pw}"] = row
    b_ms, b_sha = timed(24, 0, True)
    return {
        "value": round(b_ms, 1),
        "unit": "ms",
        "cores": threads,
        "kind": "port",
        "sample": f"one full config-B proof (log_n 24, blowup 4, 80 queries) by the C oracle (oracle/liboracle_omp.so) "
                  f"on {threads} OpenMP threads of this host, bytes compared with the GPU proof",
        "config_b_sha256_matches_gpu": b_sha == gpu_b_sha,
        "config_a": config_a,
        "host": host_info(),
    }


def prove_leg(vx, log_n, bb, pw, steps, warmup):
    """ms per proof (graph replays, host clock over `steps` proofs), one timed
    proof's stage breakdown, the digest against the committed oracle digest
    when there is one, and the host verifier's verdict."""
    a0, a1 = 1, fib_a1(42)
    prover = vx.Prover(log_n, bb, vx.N_QUERIES, pw)
    try:
        for _ in range(warmup):
            prover.prove(a0, a1)
        vx._lib.call("vx_device_sync")
        t0 = time.perf_counter()
        for _ in range(steps):
            proof = prover.prove(a0, a1)
        vx._lib.call("vx_device_sync")
        ms = (time.perf_counter() - t0) * 1000.0 / steps
        prover.enqueue(a0, a1, timed=True)
        prover.wait()
        stages = prover.stage_times()
        dev_bytes = prover.device_bytes
    finally:
        prover.close()
    digest = hashlib.sha256(proof).hexdigest()
    want = golden_digest(log_n, bb, pw)
    ok, why = vx.verify(proof, bb)
    out = {"log_n": log_n, "blowup_bits": bb, "pow_bits": pw, "n_queries": vx.N_QUERIES,
           "proof_format": "KRAK v2" if pw else "KRAK v1", "steps": steps, "ms_per_proof": round(ms, 3),
           "stages_ms": {k: round(v, 4) for k, v in stages.items()}, "proof_bytes": len(proof),
           "proof_sha256": digest, "bit_exact_vs_oracle": (digest == want) if want else None, "verified": ok,
           "device_bytes": dev_bytes}
    if not ok:
        out["verify_error"] = why
    return out


# VALU instructions per second measured on the box by a chain-of-instructions
# probe (v_alignbit_b32 + v_xor_b32, 8 independent chains per lane): the
# engine clock runs above the nominal 2.4 GHz under integer load, so
# fractions are also reported against this rate.
VALU_MEASURED_TOPS = 41.2
# Lane instructions the trace-tree subtree kernel issues per compression
# (SQ_INSTS_VALU x 64 / compressions, profiles/fib_log24_valu_counters.md):
# 960 of the compression function plus loads, addressing and stores.
ISSUED_PER_COMPRESSION = 997.1
# VALU lane instructions a proof of the serving mix issues on average
# (rocprofv3 SQ_INSTS_VALU x 64 over `bench.py --profile-serving-mix`, every
# class of the mix proved eagerly, per-class table setup left out): the
# issue-rate ceiling of the mix in proofs/s is the VALU rate over it.
MIX_ISSUED_PER_PROOF = 2.493e10


def proof_valu_ops(air, log_n, blowup_bits, pow_bits, n_cols=34, n_cons=35):
    """Algorithmic VALU instructions of one proof: Blake2s compressions of
    every tree (a leaf of <= 16 words is one compression, a node one) and of
    the expected PoW search (2^pow_bits nonces), times 960; the NTT
    butterflies (interpolate on 2^log_n, LDE of log_n layers on 2^le points),
    times 12; for the multi-column AIR the quotient's 35 constraints at about
    14 instructions per constraint and row (one M31 product, two QM31 x M31
    lazy multiply-adds). Folds, decommitment and transcript are left out
    (under 1 %)."""
    n, N = 1 << log_n, 1 << (log_n + blowup_bits)
    le = log_n + blowup_bits
    trees = (2 * N - 1) * (2 if air == 0 else 1 + -(-n_cols // 16))  # trace tree(s) + quotient tree
    fri = sum(2 * (1 << (le - 1 - k)) - 1 for k in range(le - 3))
    comp = trees + fri + ((1 << pow_bits) if pow_bits else 0)
    w = 1 if air == 0 else n_cols
    butterflies = w * log_n * (n // 2 + N // 2)
    quot = 0 if air == 0 else N * n_cons * 14
    return comp * OPS_PER_COMPRESSION + butterflies * OPS_PER_BUTTERFLY + quot


def serving_floor():
    """The serving mix's VALU floor in proofs/s: the peak VALU rate over the
    mean algorithmic instructions per proof of the mix (half Fibonacci with
    log_n uniform in 14..22, half the 34-column AIR with log_n uniform in
    14..20; blowup 4, PoW 20)."""
    fib = [proof_valu_ops(0, L, 2, 20) for L in range(14, 23)]
    d2 = [proof_valu_ops(1, L, 2, 20) for L in range(14, 21)]
    mean = 0.5 * sum(fib) / len(fib) + 0.5 * sum(d2) / len(d2)
    return {"ops_per_proof_mean": round(mean), "floor_proofs_per_s": round(VALU_PEAK_TOPS * 1e12 / mean, 1),
            "floor_proofs_per_s_measured_rate": round(VALU_MEASURED_TOPS * 1e12 / mean, 1)}


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[k]


SERVING_MIX = ("50% Fibonacci (log_n uniform 14..22), 50% synthetic 34-column AIR with 35 degree-2 "
               "constraints (log_n uniform 14..20); blowup 4, 80 queries, PoW 20")


def serving_requests(vx, n, seed):
    """The serving mix: half Fibonacci proofs (log_n uniform in 14..22,
    inputs (1, random)), half 34-column synthetic-AIR proofs (log_n uniform
    in 14..20, random trace seed, 35 constraints); blowup 4, 80 queries,
    PoW 20. Drawn from Python's seeded Mersenne Twister."""
    import random

    rng = random.Random(seed)
    reqs = []
    for _ in range(n):
        if rng.random() < 0.5:
            reqs.append(vx.request(rng.randint(14, 22), a=1, b=rng.randrange(P), blowup_bits=2, pow_bits=20))
        else:
            reqs.append(vx.request(rng.randint(14, 20), air=vx.AIR_D2, seed=rng.getrandbits(64), blowup_bits=2,
                                   pow_bits=20))
    return reqs


def run_server(vx, srv, reqs, arrivals):
    """Replay the trace; returns latencies (s), throughput, proof digests."""
    proofs, times = srv.run_trace(reqs, arrivals, keep=True)
    digests = [hashlib.sha256(p).digest() for p in proofs]
    sample = proofs[:: max(1, len(proofs) // 16)]
    verified = all(vx.verify(p, 2)[0] for p in sample)
    by_air = {}
    for r, t, a in zip(reqs, times, arrivals):
        by_air.setdefault("fibonacci" if r.air == vx.AIR_FIBONACCI else "d2_34col", []).append(t[2] - a)
    del proofs
    lat = [t[2] - a for t, a in zip(times, arrivals)]
    span = max(t[2] for t in times) - min(arrivals)
    return {
        "proofs_per_s": round(len(reqs) / span, 1),
        "p50_ms": round(percentile(lat, 50) * 1e3, 3),
        "p99_ms": round(percentile(lat, 99) * 1e3, 3),
        "max_ms": round(max(lat) * 1e3, 3),
        "span_s": round(span, 3),
        "p99_ms_by_air": {k: round(percentile(v, 99) * 1e3, 3) for k, v in sorted(by_air.items())},
        "verified_sample": verified,
    }, digests


def serving(vx, n_requests, slots, seed=42, load=0.8):
    """Continuous batching vs the one-at-a-time loop on one Poisson stream.

    1. Saturation: all requests queued at t = 0, once per server: the plain
       loop's capacity C_plain and the batching server's capacity.
    2. Open loop: Poisson arrivals at lambda = load * C_plain; both servers
       replay the same arrival trace; p50 / p99 of (completion - arrival).
    Every proof of the batching server must equal, byte for byte, the proof
    of the same request made alone by the plain loop."""
    import random

    reqs = serving_requests(vx, n_requests, seed)
    plain = vx.Server(n_slots=1)
    # about 3 GB of resident provers per slot for this mix (48 GB at 16 slots)
    batch = vx.Server(n_slots=slots, mem_budget=max(64 << 30, slots * (4 << 30)))
    classes = sorted({(r.air, r.log_n, r.blowup_bits, r.pow_bits) for r in reqs})
    warm = [vx.request(ln, air=air, blowup_bits=bb, pow_bits=pw) for (air, ln, bb, pw) in classes]
    for w in warm:
        plain.prepare(w)
        batch.prepare(w)
    # warm every class once on every slot (first replays, page-in of pinned buffers)
    plain.run_trace(warm, [0.0] * len(warm), keep=False)
    batch.run_trace(warm * slots, [0.0] * (len(warm) * slots), keep=False)

    zero = [0.0] * n_requests
    sat_plain, d_plain = run_server(vx, plain, reqs, zero)
    sat_batch, d_batch = run_server(vx, batch, reqs, zero)
    lam = load * sat_plain["proofs_per_s"]
    rng = random.Random(seed + 1)
    t, arrivals = 0.0, []
    for _ in reqs:
        t += rng.expovariate(lam)
        arrivals.append(t)
    open_plain, o_plain = run_server(vx, plain, reqs, arrivals)
    open_batch, o_batch = run_server(vx, batch, reqs, arrivals)
    identical = d_plain == d_batch == o_plain == o_batch
    roof = serving_floor()
    roof["bound"] = "valu"
    roof["achieved_proofs_per_s"] = sat_batch["proofs_per_s"]
    roof["frac"] = round(sat_batch["proofs_per_s"] / roof["floor_proofs_per_s"], 4)
    roof["plain_loop_frac"] = round(sat_plain["proofs_per_s"] / roof["floor_proofs_per_s"], 4)
    ceiling = VALU_MEASURED_TOPS * 1e12 / MIX_ISSUED_PER_PROOF
    roof["issued_instructions_per_proof"] = MIX_ISSUED_PER_PROOF
    roof["issue_ceiling_proofs_per_s"] = round(ceiling, 1)
    roof["frac_of_issue_ceiling"] = round(sat_batch["proofs_per_s"] / ceiling, 4)
    roof["plain_loop_frac_of_issue_ceiling"] = round(sat_plain["proofs_per_s"] / ceiling, 4)
    # a capacity ratio of 2 needs the batched server at twice the loop
    roof["ratio_2_needs_frac_of_issue_ceiling"] = round(2 * sat_plain["proofs_per_s"] / ceiling, 4)
    resident = batch.resident_bytes
    plain.close()
    batch.close()
    return {
        "requests": n_requests,
        "mix": SERVING_MIX,
        "slots": slots,
        "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
        "lambda_per_s": round(lam, 1),
        "plain_loop": {"saturation": sat_plain, "poisson": open_plain},
        "continuous_batching": {"saturation": sat_batch, "poisson": open_batch},
        "capacity_ratio": round(sat_batch["proofs_per_s"] / sat_plain["proofs_per_s"], 3),
        "p99_ratio": round(open_batch["p99_ms"] / open_plain["p99_ms"], 3),
        "identical_to_alone": identical,
        "resident_bytes": resident,
        "roofline": roof,
    }


def serving_mix_profile(vx):
    """Every class of the serving mix proved eagerly (no graph, so each
    dispatch is visible to a kernel trace), built-in tree shapes: 7 proofs per
    Fibonacci class (log_n 14..22) and 9 per synthetic class (log_n 14..20),
    the mix's 1:1 weighting of the two AIRs over 9 and 7 classes; PoW 20."""
    vx.tuner(False)
    t0 = time.perf_counter()
    n = 0
    for ln in range(14, 23):
        p = vx.Prover(ln, pow_bits=20, graphs=False)
        for k in range(7):
            p.prove(1, 1000 + k)
        n += 7
        p.close()
        print(f"fibonacci log_n={ln} done", file=sys.stderr, flush=True)
    for ln in range(14, 21):
        p = vx.D2Prover(ln, pow_bits=20, graphs=False)
        for k in range(9):
            p.prove(77 + k)
        n += 9
        p.close()
        print(f"synthetic log_n={ln} done", file=sys.stderr, flush=True)
    return {"proofs": n, "wall_s": round(time.perf_counter() - t0, 3)}


def batched(vx, steps, n_cols=96, seed=42, blowup_bits=2):
    """The batched uneven-commit workload: n_cols columns with log sizes drawn
    from {14..22} by the seed, values from the synthetic generator; one step =
    work-list LDE at blowup 4 + batched bit reverse + lifted Blake2s
    commitment of all columns into one tree (inputs resident in HBM)."""
    logs = vx.batch_log_sizes(n_cols, seed)
    b = vx.Batch(logs, blowup_bits, True)
    b.fill_random(seed)
    root = b.run()
    vx._lib.call("vx_device_sync")
    t0 = time.perf_counter()
    for _ in range(steps):
        b.enqueue()
    root = b.wait()
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    b.run(timed=True)
    st = b.stage_times()
    S = logs[-1] + blowup_bits
    n_in = sum(1 << L for L in logs)
    n_ev = sum(1 << (L + blowup_bits) for L in logs)
    blocks = (n_cols + 15) // 16
    compressions = blocks * (1 << S) + (1 << S) - 1
    lde_bytes = 4 * n_in + 4 * n_in + 4 * n_ev   # read inputs, write coefficients, write evaluations
    # interpolate: L layers of 2^(L-1) butterflies; evaluate: the L layers
    # that are run (the blowup layers only replicate) of 2^(L+blowup-1)
    lde_bfly = sum(L * (1 << (L - 1)) + L * (1 << (L + blowup_bits - 1)) for L in logs)
    br_bytes = 8 * n_ev
    want = None
    try:
        with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
            g = json.load(f)
        want = next((x["root"] for x in g.get("batch_c", []) if x["n_cols"] == n_cols and x["seed"] == seed), None)
    except OSError:
        pass
    # the whole batched-columns proof (LDE, bit reverse, lifted tree, RLC
    # quotient, FRI, queries, decommitment), KRAK v4
    p = vx.BatchProver(logs, blowup_bits, vx.N_QUERIES, 0)
    proof = p.prove(seed)
    vx._lib.call("vx_device_sync")
    t0 = time.perf_counter()
    for _ in range(steps):
        proof = p.prove(seed)
    proof_ms = (time.perf_counter() - t0) * 1000.0 / steps
    p.enqueue(seed, timed=True)
    p.wait()
    proof_stages = {k: round(v, 4) for k, v in p.stage_times().items() if k != "trace_subtree_kernel"}
    ok, why = vx.verify(proof, blowup_bits)
    digest = hashlib.sha256(proof).hexdigest()
    want_proof = None
    try:
        with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
            g = json.load(f)
        want_proof = next((x["sha256"] for x in g.get("proofs_batch_c", [])
                           if x["n_cols"] == n_cols and x["seed"] == seed and x["pow_bits"] == 0), None)
    except OSError:
        pass
    p.close()
    out = {
        "workload": f"{n_cols} columns, log sizes 14..22 (seed {seed}), blowup 4, bit-reversed, one lifted tree",
        "ms_per_batch": round(ms, 3),
        "stages_ms": {k: round(v, 4) for k, v in st.items()},
        "columns_per_s": round(n_cols / (ms * 1e-3), 1),
        "input_elements": n_in,
        "lde_elements": n_ev,
        "lde_gbs": round(lde_bytes / (st["lde"] * 1e-3) / 1e9, 1),
        "lde_hbm_frac": round(lde_bytes / (st["lde"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "lde_valu_frac": round(lde_bfly * OPS_PER_BUTTERFLY / (st["lde"] * 1e-3) / 1e12 / VALU_PEAK_TOPS, 4),
        "bit_reverse_gbs": round(br_bytes / (st["bit_reverse"] * 1e-3) / 1e9, 1),
        "bit_reverse_hbm_frac": round(br_bytes / (st["bit_reverse"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "commit_tops": round(compressions * OPS_PER_COMPRESSION / (st["commit"] * 1e-3) / 1e12, 2),
        "commit_valu_frac": round(compressions * OPS_PER_COMPRESSION / (st["commit"] * 1e-3) / 1e12 / VALU_PEAK_TOPS, 4),
        "lifted_root_matches_oracle": (root == want) if want else None,
        "device_bytes": b.device_bytes,
        "full_proof": {"ms_per_proof": round(proof_ms, 3), "stages_ms": proof_stages, "proof_bytes": len(proof),
                       "format": "KRAK v4 (RLC quotient + FRI, 80 queries)", "verified": ok,
                       "bit_exact_vs_oracle": (digest == want_proof) if want_proof else None},
    }
    b.close()
    return out


def cairo_commit(vx, steps, log_eval=24, seed=0xCA1F0, step=4):
    """The Cairo-shaped commitment at trace log_n = log_eval - 2 (blowup 4):
    34 trace columns as lo / hi / dict trees (16, 15, 3 columns) and the
    Cairo VM AIR quotient (40 weighted constraints over the 34 trace and 32
    interaction columns, divided by Z_H) committed as a QM31 tree; inputs from
    the synthetic generator, resident in HBM. The quotient is timed alone
    (vx_cairo_quotient, then the tree of its 4 columns) and fused with its
    tree (vx_cairo_quotient_commit, one subtree launch)."""
    import ctypes as C

    call = vx._lib.call
    n = 1 << log_eval
    bufs = []

    def alloc(words):
        p = C.c_void_p()
        call("vx_malloc", C.byref(p), 4 * words)
        bufs.append(p)
        return p.value

    def levels_words(L, lo):
        return sum(8 << (L - l) for l in range(lo, L + 1))

    try:
        cols = [alloc(n) for _ in range(70)]
        for c, ptr in enumerate(cols):
            call("vx_m31_fill_random", ptr, log_eval, seed, c, None)
        alpha, chal = alloc(256), alloc(32)
        call("vx_m31_fill_random", alpha, 8, seed, 68, None)
        call("vx_m31_fill_random", chal, 5, seed, 69, None)
        outs = [alloc(n) for _ in range(4)]
        lv, root = alloc(levels_words(log_eval, 10)), alloc(8)
        dcols = alloc(2 * 34)
        arr = (C.c_uint64 * 34)(*cols[:34])
        call("vx_memcpy", dcols, C.cast(arr, C.c_void_p), 8 * 34, 1)
        hint = (C.c_void_p * 32)(*cols[34:66])
        trees = [(0, 16), (16, 31), (31, 34)]

        def trace_trees():
            for lo, hi in trees:
                hc = (C.c_void_p * (hi - lo))(*cols[lo:hi])
                call("vx_merkle_commit_tiled", hc, hi - lo, log_eval, 10, lv, root, None)

        def quotient_alone():
            call("vx_cairo_quotient", dcols, *cols[34:66], *outs, alpha, cols[66], cols[67], chal, n, step, None)

        def quotient_tree():
            hc = (C.c_void_p * 4)(*outs)
            call("vx_merkle_commit_tiled", hc, 4, log_eval, 10, lv, root, None)

        def fused():
            call("vx_cairo_quotient_commit", dcols, hint, *outs, alpha, cols[66], cols[67], chal, log_eval, step, 10, lv,
                 root, None)

        def timed(fn):
            fn()
            call("vx_device_sync")
            t0 = time.perf_counter()
            for _ in range(steps):
                fn()
            call("vx_device_sync")
            return (time.perf_counter() - t0) * 1000.0 / steps

        t_trees = timed(trace_trees)
        t_q = timed(quotient_alone)
        t_qt = timed(quotient_tree)
        t_fused = timed(fused)
    finally:
        for p in bufs:
            vx.load().vx_free(p)
    # unique bytes per row: 34 trace + 32 interaction + Z_H^-1 + transition
    # factor read, 4 quotient words written (the next-row reads hit cache)
    q_bytes = n * 4 * (34 + 32 + 2 + 4)
    comps = 2 * n - 1
    return {
        "workload": f"Cairo-shaped commitment, trace log_n {log_eval - 2}, blowup 4 ({n} rows), synthetic columns",
        "ms_trace_trees": round(t_trees, 3),
        "ms_quotient": round(t_q, 3),
        "ms_quotient_tree": round(t_qt, 3),
        "ms_quotient_fused_commit": round(t_fused, 3),
        "ms_separate_quotient_and_tree": round(t_q + t_qt, 3),
        "quotient_hbm_gbs": round(q_bytes / (t_q * 1e-3) / 1e9, 1),
        "quotient_hbm_frac": round(q_bytes / (t_q * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "fused_tree_valu_frac": round(comps * OPS_PER_COMPRESSION / (t_fused * 1e-3) / 1e12 / VALU_PEAK_TOPS, 4),
        "steps": steps,
        "synchronous": True,
    }


def bytecode_program(seed, n_ops, n_cols, n_constraints):
    """A Cairo-sized constraint program in the stwo backend's register
    bytecode: loads of every column at row offsets 0 and +1, mostly M31
    products and sums with some QM31 work, n_constraints constraints."""
    import random

    rng = random.Random(seed)
    words, live, wide = [], [], []
    reg = 0

    def new():
        nonlocal reg
        reg += 1
        return reg - 1

    for c in range(n_cols):
        for off in (0, 1):
            d = new()
            words += [(0x03 << 24) | d, c, off]
            live.append(d)
    per = max(1, n_ops // n_constraints)
    for i in range(n_ops):
        d = new()
        a = rng.choice(live[-12:])
        b = rng.choice(live[:2 * n_cols]) if rng.random() < 0.3 else rng.choice(live[-12:])
        kind = rng.random()
        if kind < 0.45:
            words += [(0x12 << 24) | d, (a << 16) | b]       # mul
            live.append(d)
        elif kind < 0.75:
            words += [(rng.choice((0x10, 0x11)) << 24) | d, (a << 16) | b]  # add / sub
            live.append(d)
        elif kind < 0.85:
            words += [(0x15 << 24) | d, a << 16, rng.randrange(P)]  # mul by a constant
            live.append(d)
        elif kind < 0.95 or not wide:
            words += [(0x2B << 24) | d, a << 16] + [rng.randrange(P) for _ in range(4)]  # base * secure const
            wide.append(d)
        else:
            words += [(0x22 << 24) | d, (rng.choice(wide[-4:]) << 16) | rng.choice(wide[-4:])]  # QM31 product
            wide.append(d)
        if i % per == per - 1:
            words += [(0x40 << 24) | (wide[-1] if wide and rng.random() < 0.3 else live[-1])]
    return words


def bytecode_eval(vx, steps, log_eval=22, log_expand=2, n_ops=800, n_cols=34, n_constraints=40):
    """The bytecode constraint evaluator (vx_bytecode_constraint_eval) on a
    Cairo-sized program over 34 synthetic columns resident in HBM: one call
    per step (each reads the program back and checks it), timed with the
    compiled tier (the kernel built on the first call) and with the
    interpreter. Unique HBM bytes per row: each column once (the +1 row
    reads hit cache) and the four accumulator words read and written."""
    import ctypes as C

    import numpy as np

    call = vx._lib.call
    n = 1 << log_eval
    words = np.array(bytecode_program(7, n_ops, n_cols, n_constraints), np.uint32)
    m, q, k = C.c_uint32(), C.c_uint32(), C.c_uint32()
    call("vx_bytecode_slots", words.ctypes.data, words.size, n_cols, C.byref(m), C.byref(q), C.byref(k))
    bufs = []

    def alloc(nwords):
        p = C.c_void_p()
        call("vx_malloc", C.byref(p), 4 * nwords)
        bufs.append(p)
        return p.value

    try:
        cols = [alloc(n) for _ in range(n_cols)]
        for c, ptr in enumerate(cols):
            call("vx_m31_fill_random", ptr, log_eval, 0xB7C0DE, c, None)
        table = alloc(2 * n_cols)
        arr = (C.c_uint64 * n_cols)(*cols)
        call("vx_memcpy", table, C.cast(arr, C.c_void_p), 8 * n_cols, 1)
        prog = alloc(words.size)
        call("vx_memcpy", prog, C.c_void_p(words.ctypes.data), 4 * words.size, 1)
        coeff = alloc(4 * max(k.value, 1))
        call("vx_m31_fill_random", coeff, 8, 0xB7C0DE, 99, None)
        denom = alloc(1 << log_expand)
        call("vx_m31_fill_random", denom, log_expand, 0xB7C0DE, 98, None)
        acc = [alloc(n) for _ in range(4)]
        for a in acc:
            call("vx_memset", a, 0, 4 * n)

        def run():
            call("vx_bytecode_constraint_eval", prog, int(words.size), table, None, n_cols, n, n >> log_expand,
                 coeff, denom, log_expand, *acc, 0, None)

        def timed():
            run()  # the compiled tier builds its kernel here, untimed
            call("vx_device_sync")
            t0 = time.perf_counter()
            for _ in range(steps):
                run()
            call("vx_device_sync")
            return (time.perf_counter() - t0) * 1000.0 / steps

        ms = timed()
        call("vx_bytecode_set_jit", 0)
        ms_interp = timed()
    finally:
        call("vx_bytecode_set_jit", 1)
        for p in bufs:
            vx.load().vx_free(p)
    hbm = n * (4 * n_cols + 32)
    return {
        "workload": f"{n_ops}-op constraint program over {n_cols} columns (offsets 0, +1), {k.value} constraints, "
                    f"{n} rows (trace 2^{log_eval - log_expand}, blowup {1 << log_expand})",
        "tier": "compiled (hipRTC, gfx950, cached per program)",
        "ms": round(ms, 3),
        "ms_interpreted": round(ms_interp, 3),
        "interpreter_register_file_words_per_row": m.value + 4 * q.value,
        "rows_per_s": round(n / (ms * 1e-3), 1),
        "hbm_gbs": round(hbm / (ms * 1e-3) / 1e9, 1),
        "hbm_frac": round(hbm / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "steps": steps,
        "synchronous": True,
    }


def gkr(vx, steps, log_n=24):
    """GKR operations of the stwo plugin (vx_gkr_*) on secure columns of
    2^log_n QM31 elements resident in HBM, asynchronous calls back to back
    on one stream: fold of the first variable, the LogUp next layer, the
    round sums of a LogUp layer (n_terms = 2^(log_n - 2)) and the eq table
    (2^log_n entries) in one launch and as the plugin's log_n passes. HBM
    bytes are what each call must move: its inputs read once, its outputs
    written once."""
    import ctypes as C

    import numpy as np
    from vortexstark_amd.device import DeviceBuffer

    call = vx._lib.call
    n = 1 << log_n
    rng = np.random.default_rng(11)
    num = DeviceBuffer.from_host(rng.integers(0, P, 4 * n, dtype=np.uint32))
    den = DeviceBuffer.from_host(rng.integers(0, P, 4 * n, dtype=np.uint32))
    small = DeviceBuffer.from_host(rng.integers(0, P, 4 * 32, dtype=np.uint32))  # r, lambda, v, y
    out_n, out_d = DeviceBuffer.zeros(2 * n), DeviceBuffer.zeros(2 * n)
    eq = DeviceBuffer.zeros(4 * n)
    part = DeviceBuffer.zeros(8 * (n // 4 // 256))
    nb = C.c_uint32(0)
    vp = lambda b, off=0: C.c_void_p(b.as_ptr() + 4 * off)
    legs = {
        "fix_first_variable_secure": (lambda: call("vx_gkr_fix_first_variable_secure", vp(num), vp(out_n),
                                                   vp(small), n, None), 16 * n + 8 * n),
        "next_layer_logup_generic": (lambda: call("vx_gkr_next_layer_logup_generic", vp(num), vp(den), vp(out_n),
                                                  vp(out_d), n, None), 32 * n + 16 * n),
        "sum_poly_logup_generic": (lambda: call("vx_gkr_sum_poly_logup_generic", vp(eq), vp(num), vp(den),
                                                vp(small, 4), vp(part), n // 4, C.byref(nb), None),
                                   16 * (n // 4) + 32 * n),
        "gen_eq_evals": (lambda: call("vx_gkr_gen_eq_evals", vp(eq), vp(small, 12), log_n, vp(small, 8), None),
                         16 * n),
    }

    def eq_passes():
        call("vx_gkr_gen_eq_evals_init", vp(eq), vp(small, 8), None)
        for i in range(log_n):
            call("vx_gkr_gen_eq_evals_pass", vp(eq), vp(small, 12 + 4 * (log_n - 1 - i)), 1 << i, None)
    # a pass reads and writes 2^i elements and writes 2^i more: 48 * 2^i bytes
    legs["gen_eq_evals_plugin_passes"] = (eq_passes, 48 * (n - 1) + 16)
    res = {"workload": f"QM31 columns of 2^{log_n} elements in HBM (sum: 2^{log_n - 2} terms of a LogUp layer)"}
    for name, (fn, nbytes) in legs.items():
        for _ in range(2):
            fn()
        call("vx_device_sync")
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        call("vx_device_sync")
        ms = (time.perf_counter() - t0) * 1000.0 / steps
        gbs = nbytes / (ms * 1e-3) / 1e9
        res[name] = {"ms": round(ms, 4), "hbm_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
    res["steps"] = steps
    return res


def poseidon(vx, steps):
    """Poseidon252 operations of the stwo plugin (the Stark-field Hades
    permutation, one state per lane): a parent layer of 2^20 nodes, lifted
    leaves over 32 columns of 2^16 rows (4 packed elements, 3 permutations
    per leaf) and a grind batch of 2^22 nonces that no nonce passes (PoW
    128), so every nonce is hashed. Reported as permutations per second."""
    import ctypes as C

    import numpy as np
    from vortexstark_amd.device import DeviceBuffer

    call = vx._lib.call
    rng = np.random.default_rng(12)
    # canonical field elements below 2^250 < p are valid Montgomery words
    def felts(n):
        w = rng.integers(0, 1 << 32, 8 * n, dtype=np.uint64).astype(np.uint32)
        w[7::8] &= 0x03FFFFFF
        return w
    n_par, n_leaf, n_cols, n_grind = 1 << 20, 1 << 16, 32, 1 << 22
    prev, par = DeviceBuffer.from_host(felts(2 * n_par)), DeviceBuffer.zeros(8 * n_par)
    cols = [DeviceBuffer.from_host(rng.integers(0, P, n_leaf, dtype=np.uint32)) for _ in range(n_cols)]
    ptrs = DeviceBuffer.from_host(np.array([c.as_ptr() for c in cols], np.uint64).view(np.uint32))
    logs = DeviceBuffer.from_host(np.full(n_cols, 16, np.uint32))
    leaves = DeviceBuffer.zeros(8 * n_leaf)
    prefix = DeviceBuffer.from_host(felts(1))
    best = DeviceBuffer.from_host(np.full(2, 0xFFFFFFFF, np.uint32))
    vp = lambda b: C.c_void_p(b.as_ptr())
    legs = {
        "next_layer": (lambda: call("vx_build_next_layer_poseidon252", vp(prev), vp(par), n_par, None), n_par),
        "leaves_32_cols": (lambda: call("vx_build_leaves_poseidon252", vp(ptrs), vp(logs), n_cols, 16, vp(leaves),
                                        n_leaf, None), 3 * n_leaf),
        "grind_batch": (lambda: call("vx_grind_pow_poseidon", vp(prefix), vp(best), 128, 0, n_grind, None), n_grind),
    }
    res = {"workload": "Poseidon252 (Stark field) permutations, one state per lane, inputs in HBM"}
    for name, (fn, perms) in legs.items():
        fn()
        call("vx_device_sync")
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        call("vx_device_sync")
        ms = (time.perf_counter() - t0) * 1000.0 / steps
        res[name] = {"ms": round(ms, 4), "permutations_per_s": round(perms / (ms * 1e-3), 1)}
    res["steps"] = steps
    return res


def stwo_ntt(vx, steps, log_n=24, batch_log=20, batch_cols=16, cpu=True):
    """stwo-basis circle NTT (the stwo plugin's PolyOps): evaluate and
    interpolate of one 2^log_n column resident in HBM, asynchronous calls
    back to back on one stream, and the synchronous batched call over
    batch_cols separately allocated 2^batch_log columns. The HBM figure counts
    what the passes must move: each pass reads and writes the column once
    (two passes at 2^24: a strided 10-bit and a contiguous 14-bit one)."""
    import ctypes as C

    import numpy as np
    from vortexstark_amd import stwo
    from vortexstark_amd.device import DeviceBuffer

    call = vx._lib.call
    n = 1 << log_n
    cache = stwo.StwoNttCache(log_n)
    x = np.random.default_rng(7).integers(0, P, n, dtype=np.uint32)
    d = DeviceBuffer.from_host(x)
    ptr, tw, itw = C.c_void_p(d.as_ptr()), C.c_void_p(cache.d_fwd_twiddles.as_ptr()), \
        C.c_void_p(cache.d_inv_twiddles.as_ptr())
    res = {}
    for name, fn, t in (("evaluate", "vx_stwo_ntt_evaluate_async", tw),
                        ("interpolate", "vx_stwo_ntt_interpolate_async", itw)):
        for _ in range(2):
            call(fn, ptr, t, n, 1, 0, None)
        call("vx_device_sync")
        t0 = time.perf_counter()
        for _ in range(steps):
            call(fn, ptr, t, n, 1, 0, None)
        call("vx_device_sync")
        res[name] = (time.perf_counter() - t0) * 1000.0 / steps
    # evaluate and interpolate ran `steps + 2` times each: the column is back
    roundtrip_ok = bool(np.array_equal(d.to_host(), x))
    # PolyOps::evaluate of 2^(log_n - 2) coefficients on the 2^log_n domain:
    # fused (replication layers skipped) against the plugin's extend
    # (zero-fill + copy) followed by the full transform
    cn = n >> 2
    d_co = DeviceBuffer.from_host(x[:cn])

    def lde_fused():
        call("vx_stwo_ntt_evaluate_lde", ptr, C.c_void_p(d_co.as_ptr()), cn, tw, n, None)

    def lde_plugin():
        call("vx_memset", ptr, 0, 4 * n)
        call("vx_memcpy", ptr, C.c_void_p(d_co.as_ptr()), 4 * cn, 3)
        call("vx_stwo_ntt_evaluate_async", ptr, tw, n, 1, 0, None)

    lde_ms = {}
    for name, fn in (("fused", lde_fused), ("extend_then_evaluate", lde_plugin)):
        fn()
        call("vx_device_sync")
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        call("vx_device_sync")
        lde_ms[name] = round((time.perf_counter() - t0) * 1000.0 / steps, 4)
    d_co.free()
    passes = 1 if log_n <= 14 else 2
    bytes_moved = passes * 8 * n
    m = 1 << batch_log
    cols = [DeviceBuffer.from_host(np.random.default_rng(c).integers(0, P, m, dtype=np.uint32))
            for c in range(batch_cols)]
    cache_b = stwo.StwoNttCache(batch_log)
    stwo.evaluate_stwo_batch(cols, cache_b)
    t0 = time.perf_counter()
    for _ in range(steps):
        stwo.evaluate_stwo_batch(cols, cache_b)
    batch_ms = (time.perf_counter() - t0) * 1000.0 / steps
    # PolyOps::eval_at_point of the same column (now values, any 2^24 words
    # serve) and barycentric_eval_at_point over 2^22 values with the plugin's
    # 256 blocks; synchronous calls, as the plugin makes them
    zx, zy = (3, 1, 4, 1), (5, 9, 2, 6)
    stwo.eval_at_point(d, n, zx, zy)
    t0 = time.perf_counter()
    for _ in range(steps):
        stwo.eval_at_point(d, n, zx, zy)
    eval_pt_ms = (time.perf_counter() - t0) * 1000.0 / steps
    nb = 1 << 22
    e = DeviceBuffer.from_host(np.random.default_rng(8).integers(0, P, nb, dtype=np.uint32))
    w = DeviceBuffer.from_host(np.random.default_rng(9).integers(0, P, 4 * nb, dtype=np.uint32))
    stwo.barycentric_eval(e, w, nb)
    t0 = time.perf_counter()
    for _ in range(steps):
        stwo.barycentric_eval(e, w, nb)
    bary_ms = (time.perf_counter() - t0) * 1000.0 / steps
    # QuotientOps at 2^22 rows: accumulate_numerators over 16 columns, then
    # compute_quotients_combine of 2 accumulators (asynchronous calls)
    qr, qc = 1 << 22, 16
    qcols = [DeviceBuffer.from_host(np.random.default_rng(40 + i).integers(0, P, qr, dtype=np.uint32))
             for i in range(qc)]

    def dptrs(bufs):
        return DeviceBuffer.from_host(np.array([b.as_ptr() for b in bufs], np.uint64).view(np.uint32))

    d_cols = dptrs(qcols)
    d_idx = DeviceBuffer.from_host(np.arange(qc, dtype=np.uint32))
    d_b = DeviceBuffer.from_host(np.random.default_rng(50).integers(0, P, 4 * qc, dtype=np.uint32))
    d_c = DeviceBuffer.from_host(np.random.default_rng(51).integers(0, P, 4 * qc, dtype=np.uint32))
    accs = [[DeviceBuffer(qr) for _ in range(4)] for _ in range(2)]
    vp = C.c_void_p

    def acc_call(o):
        call("vx_accumulate_numerators", vp(d_cols.as_ptr()), vp(d_idx.as_ptr()), vp(d_b.as_ptr()),
             vp(d_c.as_ptr()), qc, qr, *[vp(x.as_ptr()) for x in o], None)

    for o in accs:
        acc_call(o)
    call("vx_device_sync")
    t0 = time.perf_counter()
    for _ in range(steps):
        acc_call(accs[0])
    call("vx_device_sync")
    accnum_ms = (time.perf_counter() - t0) * 1000.0 / steps
    rng = np.random.default_rng(52)
    d_sp = [DeviceBuffer.from_host(rng.integers(0, P, 8, dtype=np.uint32)) for _ in range(3)]
    d_nu = [dptrs([accs[j][k] for j in range(2)]) for k in range(4)]
    d_logs = DeviceBuffer.from_host(np.array([22, 22], np.uint32))
    d_xs, d_ys = (DeviceBuffer.from_host(rng.integers(0, P, qr, dtype=np.uint32)) for _ in range(2))
    q_out = [DeviceBuffer(qr) for _ in range(4)]

    def comb_call():
        call("vx_compute_quotients_combine", *[vp(x.as_ptr()) for x in d_sp], *[vp(x.as_ptr()) for x in d_nu],
             vp(d_logs.as_ptr()), 2, vp(d_xs.as_ptr()), vp(d_ys.as_ptr()), 22, qr,
             *[vp(x.as_ptr()) for x in q_out], None)

    comb_call()
    call("vx_device_sync")
    t0 = time.perf_counter()
    for _ in range(steps):
        comb_call()
    call("vx_device_sync")
    comb_ms = (time.perf_counter() - t0) * 1000.0 / steps
    # OODS-style: 64 columns of 2^20 coefficients at one point, one batched
    # call against 64 single calls (each synchronous, as the plugin's
    # per-column eval_at_point)
    oc, ol = 64, 20
    ocols = [DeviceBuffer.from_host(np.random.default_rng(90 + i).integers(0, P, 1 << ol, dtype=np.uint32))
             for i in range(oc)]
    d_f = DeviceBuffer.from_host(stwo.folding_factors(zx, zy, ol))
    d_res = DeviceBuffer(4 * oc)
    o_ptrs = (C.c_void_p * oc)(*[b.as_ptr() for b in ocols])
    o_logs = (C.c_uint32 * oc)(*([ol] * oc))

    def batch_call():
        call("vx_eval_at_point_batch", o_ptrs, o_logs, oc, vp(d_f.as_ptr()), ol, vp(d_res.as_ptr()), None)

    batch_call()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch_call()
    oods_batch_ms = (time.perf_counter() - t0) * 1000.0 / steps
    batch_vals = d_res.to_host().reshape(oc, 4)
    t0 = time.perf_counter()
    singles = [stwo.eval_at_point(b, 1 << ol, zx, zy) for b in ocols]
    oods_single_ms = (time.perf_counter() - t0) * 1000.0
    oods_same = all(tuple(int(v) for v in batch_vals[i]) == singles[i] for i in range(oc))
    # the OODS step over the same 64 columns (trace 2^20): draw z, both
    # points' values for every column, mix them into the channel
    d_chan = DeviceBuffer.from_host(np.arange(9, dtype=np.uint32))
    d_opts, d_ovals = DeviceBuffer(16), DeviceBuffer(8 * oc)

    def oods_call():
        call("vx_oods_sample", vp(d_chan.as_ptr()), o_ptrs, o_logs, oc, ol, vp(d_opts.as_ptr()),
             vp(d_ovals.as_ptr()), None)

    oods_call()
    t0 = time.perf_counter()
    for _ in range(steps):
        oods_call()
    oods_step_ms = (time.perf_counter() - t0) * 1000.0 / steps
    # half-coset natural -> canonic bit-reversed order of one 2^log_n column
    # (asynchronous calls back to back): reads and writes the column once
    d_perm = DeviceBuffer(n)

    def perm_call():
        call("vx_permute_hc_to_canonic_brt", ptr, vp(d_perm.as_ptr()), n, log_n, None)

    perm_call()
    call("vx_device_sync")
    t0 = time.perf_counter()
    for _ in range(steps):
        perm_call()
    call("vx_device_sync")
    perm_ms = (time.perf_counter() - t0) * 1000.0 / steps
    out = {
        "workload": f"stwo-basis evaluate / interpolate, one 2^{log_n} column in HBM",
        "evaluate_ms": round(res["evaluate"], 4),
        "interpolate_ms": round(res["interpolate"], 4),
        "passes": passes,
        "hbm_bytes_per_transform": bytes_moved,
        "evaluate_hbm_frac": round(bytes_moved / (res["evaluate"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "evaluate_valu_frac": round(log_n * (n // 2) * OPS_PER_BUTTERFLY / (res["evaluate"] * 1e-3) / 1e12
                                    / VALU_PEAK_TOPS, 4),
        "interpolate_hbm_frac": round(bytes_moved / (res["interpolate"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "roundtrip_ok": roundtrip_ok,
        "evaluate_lde_ms": {"coeffs": cn, "domain": n, **lde_ms},
        "batch": {"columns": batch_cols, "log_n": batch_log, "ms_per_call": round(batch_ms, 4),
                  "columns_per_s": round(batch_cols / (batch_ms * 1e-3), 1), "synchronous": True},
        "eval_at_point": {"log_n": log_n, "ms_per_call": round(eval_pt_ms, 4),
                          "read_gbs": round(4 * n / (eval_pt_ms * 1e-3) / 1e9, 1), "synchronous": True},
        "barycentric": {"n": nb, "blocks": 256, "ms_per_call": round(bary_ms, 4),
                        "read_gbs": round(20 * nb / (bary_ms * 1e-3) / 1e9, 1), "synchronous": True},
        "accumulate_numerators": {"rows": qr, "columns": qc, "ms_per_call": round(accnum_ms, 4),
                                  "hbm_gbs": round(4 * qr * (qc + 4) / (accnum_ms * 1e-3) / 1e9, 1)},
        "eval_at_point_batch": {"columns": oc, "log_n": ol, "ms_per_call": round(oods_batch_ms, 4),
                                "ms_as_single_calls": round(oods_single_ms, 3),
                                "read_gbs": round(4 * oc * (1 << ol) / (oods_batch_ms * 1e-3) / 1e9, 1),
                                "same_as_single_calls": oods_same},
        "quotients_combine": {"rows": qr, "accumulators": 2, "ms_per_call": round(comb_ms, 4),
                              "hbm_gbs": round(4 * qr * (2 + 8 + 4) / (comb_ms * 1e-3) / 1e9, 1)},
        "oods_sample": {"columns": oc, "log_n": ol, "ms_per_call": round(oods_step_ms, 4),
                        "read_gbs": round(2 * 4 * oc * (1 << ol) / (oods_step_ms * 1e-3) / 1e9, 1),
                        "synchronous": True},
        "permute_hc_to_canonic_brt": {"log_n": log_n, "ms_per_call": round(perm_ms, 4),
                                      "hbm_gbs": round(8 * n / (perm_ms * 1e-3) / 1e9, 1),
                                      "hbm_frac": round(8 * n / (perm_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
    }
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po

        fwd, inv = po.stwo_twiddles(log_n)
        t0 = time.perf_counter()
        ev = po.stwo_evaluate(x, fwd)
        te = time.perf_counter() - t0
        t0 = time.perf_counter()
        back = po.stwo_interpolate(ev, inv)
        ti = time.perf_counter() - t0
        out["cpu_baseline"] = {"evaluate_ms": round(te * 1000.0, 1), "interpolate_ms": round(ti * 1000.0, 1),
                               "cores": 1, "kind": "port", "roundtrip_ok": bool(np.array_equal(back, x)),
                               "sample": f"one full 2^{log_n} evaluate and interpolate, C oracle, 1 thread"}
    for b in cols + [e, w, d] + qcols + [x for a in accs for x in a] + q_out + ocols:
        b.free()
    return out


def copy_calibration(vx, mib=1024):
    """A plain device-to-device copy of known size (1 GiB, four times the
    256 MB MALL, so the source streams from HBM): its FETCH_SIZE / WRITE_SIZE
    counters scale the per-kernel HBM traffic reported from rocprofv3."""
    import ctypes as C

    src, dst = C.c_void_p(), C.c_void_p()
    nbytes = mib << 20
    vx._lib.call("vx_malloc", C.byref(src), nbytes)
    vx._lib.call("vx_malloc", C.byref(dst), nbytes)
    vx._lib.call("vx_memset", src, 1, nbytes)
    vx._lib.call("vx_memcpy", dst, src, nbytes, 3)
    vx._lib.call("vx_device_sync")
    vx._lib.call("vx_free", src)
    vx._lib.call("vx_free", dst)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log-n", type=int, default=24)
    ap.add_argument("--blowup-bits", type=int, default=2)
    ap.add_argument("--pow-bits", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the config A (log_n 16), config B PoW 26 and config C (log_n 30) legs")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--slots", type=int, default=16)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-only", action="store_true")
    ap.add_argument("--no-stwo", action="store_true")
    ap.add_argument("--no-cairo", action="store_true")
    ap.add_argument("--cairo-only", action="store_true")
    ap.add_argument("--no-bytecode", action="store_true")
    ap.add_argument("--bytecode-only", action="store_true")
    ap.add_argument("--stwo-only", action="store_true")
    ap.add_argument("--no-gkr", action="store_true")
    ap.add_argument("--gkr-only", action="store_true")
    ap.add_argument("--no-poseidon", action="store_true")
    ap.add_argument("--poseidon-only", action="store_true")
    ap.add_argument("--profile-serving-mix", action="store_true",
                    help="eager proofs of every serving class (7 per Fibonacci class, 9 per synthetic class) "
                         "for a kernel-trace profile of the mix")
    ap.add_argument("--tune-cache", default=None,
                    help="autotuner cache file: choices are loaded from it and new ones appended, so that "
                         "profiled runs use the shapes an unprofiled run chose")
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3: one 1 GiB device copy (counter calibration), "
                         "1 warmup + 2 timed proofs, no CPU baseline")
    args = ap.parse_args()
    if args.profile:
        args.steps, args.warmup, args.no_cpu_baseline, args.no_serve, args.no_batch = 2, 1, True, True, True
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import vortexstark_amd as vx

    vx.load()
    if args.tune_cache:
        vx.tuner(True, args.tune_cache)
    if args.profile_serving_mix:
        print(json.dumps({"serving_mix_profile": serving_mix_profile(vx)}), flush=True)
        return
    if args.serve_only:
        print(json.dumps({"serving": serving(vx, args.serve_requests, args.slots)}), flush=True)
        return
    if args.batch_only:
        print(json.dumps({"batched_commit": batched(vx, args.steps)}), flush=True)
        return
    if args.bytecode_only:
        print(json.dumps({"bytecode_eval": bytecode_eval(vx, args.steps)}), flush=True)
        return
    if args.cairo_only:
        print(json.dumps({"cairo_commit": cairo_commit(vx, args.steps)}), flush=True)
        return
    if args.gkr_only:
        print(json.dumps({"gkr": gkr(vx, args.steps)}), flush=True)
        return
    if args.poseidon_only:
        print(json.dumps({"poseidon252": poseidon(vx, args.steps)}), flush=True)
        return
    if args.stwo_only:
        print(json.dumps({"stwo_ntt": stwo_ntt(vx, args.steps, cpu=not args.no_cpu_baseline)}), flush=True)
        return
    if args.profile:
        copy_calibration(vx)
    log_n, bb, pw = args.log_n, args.blowup_bits, args.pow_bits
    a0, a1 = 1, fib_a1(42)
    prover = vx.Prover(log_n, bb, vx.N_QUERIES, pw)
    for _ in range(args.warmup):
        prover.prove(a0, a1)
    vx._lib.call("vx_device_sync")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        proof = prover.prove(a0, a1)
    vx._lib.call("vx_device_sync")
    elapsed = time.perf_counter() - t0
    ms = elapsed * 1000.0 / args.steps

    # one timed proof for the stage breakdown and the dominant kernel
    prover.enqueue(a0, a1, timed=True)
    timed_proof = prover.wait()
    stages = prover.stage_times()
    digest = hashlib.sha256(proof).hexdigest()
    want = golden_digest(log_n, bb, pw)
    tv = time.perf_counter()
    ok, why = vx.verify(proof, bb)
    verify_ms = (time.perf_counter() - tv) * 1000.0

    le = log_n + bb
    N = 1 << le
    tile = min(le, 12)
    compressions = N + (N - (N >> tile))  # leaves + nodes inside the 4096-leaf tiles
    k_ms = stages["trace_subtree_kernel"]
    achieved = compressions * OPS_PER_COMPRESSION / (k_ms * 1e-3) / 1e12
    lde_bytes = 4 * (1 << log_n) + 4 * N  # read coefficients once, write the evaluation once
    lde_gbs = lde_bytes / (stages["lde"] * 1e-3) / 1e9
    # the log_n layers that are run (the blowup layers only replicate), N/2 butterflies each
    lde_tops = log_n * (N // 2) * OPS_PER_BUTTERFLY / (stages["lde"] * 1e-3) / 1e12

    out = {
        "metric": "Fib log_n=24 prove ms" if log_n == 24 else f"Fib log_n={log_n} prove ms",
        "value": round(ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": round(ms / REFERENCE_MS, 4) if (log_n, bb, pw) == (24, 2, 0) else None,
        "dtype": "u32 (M31/QM31 exact)",
        "data": "synthetic (seeded Fibonacci statement a0=1, a1=SplitMix64(42))",
        "config": {"workload": "fibonacci_stark_prove", "log_n": log_n, "blowup_bits": bb, "lde_log": le,
                   "n_queries": vx.N_QUERIES, "pow_bits": pw, "proof_format": "KRAK v2" if pw else "KRAK v1",
                   "parallelism": "single GPU"},
        "roofline": {
            "kernel": "k_subtree<LEAF_COLS,1,4> (trace-tree Blake2s, %d leaves)" % N,
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": round(VALU_PEAK_TOPS, 2),
            "unit": "TOP/s",
            "frac": round(achieved / VALU_PEAK_TOPS, 4),
            "frac_vs_measured_rate": round(achieved / VALU_MEASURED_TOPS, 4),
            "issued_frac": round(compressions * ISSUED_PER_COMPRESSION / (k_ms * 1e-3) / 1e12 / VALU_PEAK_TOPS, 4),
            "traffic": profiled_traffic() if (log_n, bb) == (24, 2) else None,
            "kernel_ms": round(k_ms, 4),
            "work": f"{compressions} Blake2s compressions x {OPS_PER_COMPRESSION} int ops",
        },
        "roofline_lde": {"bound": "hbm", "achieved": round(lde_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(lde_gbs / HBM_PEAK_GBS, 4), "stage_ms": round(stages["lde"], 4),
                         "algorithmic_bytes": lde_bytes,
                         "valu_tops": round(lde_tops, 2), "valu_frac": round(lde_tops / VALU_PEAK_TOPS, 4),
                         "note": "VALU-bound: 12 int ops per butterfly; the HBM fraction is not the limit"},
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "proof_bytes": len(proof),
        "proof_sha256": digest,
        "bit_exact_vs_oracle": (digest == want) if want else None,
        "timed_proof_identical": timed_proof == proof,
        "verified": ok,
        "verify_ms": round(verify_ms, 2),
        "device_bytes": prover.device_bytes,
        "ntt_plans": [{k: e[k] for k in ("forward", "log_n", "bits", "n_cols", "split", "order")}
                      for e in vx.tuner_ntt_entries()],
    }
    if not ok:
        out["verify_error"] = why
    headline = (log_n, bb, pw) == (24, 2, 0) and not args.profile
    gpu_a = None
    if headline and not (args.no_configs and args.no_cpu_baseline):
        # config A on the GPU (the CPU baseline's workload), PoW 0 and 10
        gpu_a = {p: prove_leg(vx, 16, 2, p, 20, 3) for p in (0, 10)}
        out["config_a_gpu"] = {f"pow{p}": v for p, v in gpu_a.items()}
    if headline and not args.no_configs:
        # config B as specified (PoW 26, KRAK v2) and config C (log_n 30,
        # blowup 1: the whole 2^31-point circle group, everything resident)
        out["config_b_pow26"] = prove_leg(vx, 24, 2, 26, 5, 1)
        out["config_c"] = prove_leg(vx, 30, 1, 26, 3, 1)
    if not args.no_serve and not args.profile:
        out["serving"] = serving(vx, args.serve_requests, args.slots)
    if not args.no_batch and not args.profile:
        out["batched_commit"] = batched(vx, args.steps)
    if not args.no_cairo and not args.profile:
        out["cairo_commit"] = cairo_commit(vx, args.steps)
    if not args.no_bytecode and not args.profile:
        out["bytecode_eval"] = bytecode_eval(vx, args.steps)
    if not args.no_stwo and not args.profile:
        out["stwo_ntt"] = stwo_ntt(vx, args.steps, cpu=not args.no_cpu_baseline)
    if not args.no_gkr and not args.profile:
        out["gkr"] = gkr(vx, args.steps)
    if not args.no_poseidon and not args.profile:
        out["poseidon252"] = poseidon(vx, args.steps)
    if not args.no_cpu_baseline and gpu_a is not None:
        out["cpu_baseline"] = cpu_baseline(gpu_a, digest)
    prover.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
