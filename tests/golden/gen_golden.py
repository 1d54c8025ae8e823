"""Writes tests/golden/golden.json: known-answer vectors for the prove path.

Sources, per section:
  blake2s      Python hashlib.blake2s (standard Blake2s-256)
  twiddles,    tests/golden/pyref.py, an independent big-int restatement that
  ntt, folds   derives twiddles from closed-form point powers
  proofs_py    pyref.prove (independent of the C oracle), SHA-256 of the bytes
  proofs_c     the C oracle (oracle/liboracle_omp.so) at sizes pyref cannot
               reach; pinned transitively by proofs_py agreement at small sizes
  pow          brute-force smallest nonce with hashlib
  proofs_d2_py pyref.prove_d2 (synthetic 34-column AIR, KRAK v3), independent
               of the C oracle
  proofs_d2_c  the C oracle's multi-column AIR proofs at larger sizes
  batch_c      the C oracle's lifted root of the batched uneven-column
               workload (per-column LDE, bit reverse, lifted tree)
  proofs_batch_py / proofs_batch_c
               the batched-columns proof (KRAK v4): pyref.prove_batch at tiny
               sizes, the C oracle at the workload's sizes

  cairo_commit_c  the Cairo-shaped commitment at trace log_n 20 (lo / hi / dict
               trace trees and the quotient tree), C oracle
  serving_e    one request of every class of the serving mix (Fibonacci
               log_n 14..22, 34-column AIR log_n 14..20, PoW 20), C oracle

Run from the repo root:  python tests/golden/gen_golden.py [--big | --serving | --cairo]
(--big adds the log_n = 20, 24 and 25 (blowup 1) oracle digests, the
log_n = 24 PoW-26 digest and the log_n = 20 multi-column AIR digest, which
take minutes; --serving / --cairo recompute only serving_e / cairo_commit_c
and keep the rest.)
"""
import hashlib
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

import pyref  # noqa: E402

P = pyref.P
MASK64 = (1 << 64) - 1


def splitmix64(seed):
    s = seed & MASK64
    while True:
        s = (s + 0x9E3779B97F4A7C15) & MASK64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        yield z ^ (z >> 31)


def fib_a1(seed):
    """a1 of the seeded trace recipe: r = next() >> 33, rejected while r >= P."""
    for z in splitmix64(seed):
        r = z >> 33
        if r < P:
            return r


def serving_requests_e(seed=20261203):
    """One request of every class of the serving mix: Fibonacci (a0 = 1,
    a1 from SplitMix64) at log_n 14..22 and the 34-column synthetic AIR
    (35 constraints, air seed 0xD2A1, trace seed from SplitMix64) at log_n
    14..20; blowup 4, 80 queries, PoW 20."""
    out = []
    s = seed
    for log_n in range(14, 23):
        s += 1
        out.append({"air": 0, "log_n": log_n, "a": 1, "b": fib_a1(s), "blowup_bits": 2, "n_queries": 80,
                    "pow_bits": 20})
    for log_n in range(14, 21):
        s += 1
        out.append({"air": 1, "log_n": log_n, "seed": next(splitmix64(s)), "n_cols": 34, "n_cons": 35,
                    "air_seed": 0xD2A1, "blowup_bits": 2, "n_queries": 80, "pow_bits": 20})
    return out


def serving_digests():
    import pyoracle
    rows = []
    for r in serving_requests_e():
        if r["air"] == 0:
            data = pyoracle.prove(r["a"], r["b"], r["log_n"], blowup_bits=2, n_queries=80, pow_bits=20, omp=True)
        else:
            data = pyoracle.d2_prove(r["seed"], r["log_n"], 34, 35, 0xD2A1, 2, 80, 20, omp=True)
        rows.append(dict(r, len=len(data), sha256=hashlib.sha256(data).hexdigest()))
        print("serving digest", r["air"], r["log_n"], len(data), flush=True)
    return rows


CAIRO_SEED, CAIRO_LOG_EVAL, CAIRO_STEP = 0xCA1F0, 22, 4


def cairo_commit_digest():
    """The Cairo-shaped commitment (trace log_n 20, blowup 4: 2^22 rows):
    column c < 70 of the synthetic generator with seed CAIRO_SEED (34 trace
    columns, 32 interaction columns, Z_H^-1, transition factors), alphas
    and challenges from columns 68 and 69; roots of the lo / hi / dict trace
    trees (columns 0-15, 16-30, 31-33) and of the quotient's QM31 tree."""
    import pyoracle
    L = CAIRO_LOG_EVAL
    cols = [pyoracle.d2_column(CAIRO_SEED, c, L) for c in range(68)]
    alpha = pyoracle.d2_column(CAIRO_SEED, 68, 8)[:160]
    chal = pyoracle.d2_column(CAIRO_SEED, 69, 5)[:24]
    roots = [pyoracle.merkle_root(cols[lo:hi]) for lo, hi in ((0, 16), (16, 31), (31, 34))]
    q = pyoracle.cairo_quotient(cols[:34], cols[34:66], alpha, cols[66], cols[67], chal, CAIRO_STEP)
    roots.append(pyoracle.merkle_root(q))
    print("cairo commitment digest", flush=True)
    return {"seed": CAIRO_SEED, "log_eval": L, "blowup_step": CAIRO_STEP,
            "roots": [[int(x) for x in r] for r in roots]}


def main():
    big = "--big" in sys.argv
    out = os.path.join(HERE, "golden.json")
    if "--cairo" in sys.argv:
        g = json.load(open(out))
        g["cairo_commit_c"] = cairo_commit_digest()
        with open(out, "w") as f:
            json.dump(g, f, separators=(",", ":"))
        print("wrote", out, os.path.getsize(out), "bytes")
        return
    if "--serving" in sys.argv:
        # only (re)compute the serving-mix digests, keep everything else
        g = json.load(open(out))
        g["serving_e"] = serving_digests()
        with open(out, "w") as f:
            json.dump(g, f, separators=(",", ":"))
        print("wrote", out, os.path.getsize(out), "bytes")
        return
    g = {}
    rng = random.Random(20261202)

    g["blake2s"] = []
    for n in [0, 1, 4, 16, 37, 40, 52, 63, 64, 65, 127, 128, 129, 160, 300]:
        data = bytes(rng.randrange(256) for _ in range(n))
        g["blake2s"].append({"in": data.hex(), "out": hashlib.blake2s(data).hexdigest()})

    g["twiddles"] = []
    for L in [2, 3, 4, 5]:
        for name, c in [("half_coset", pyref.half_coset(L)), ("half_odds", pyref.half_odds(L)),
                        ("subgroup", pyref.subgroup(L))]:
            line, circle = pyref.twiddle_tables(c)
            g["twiddles"].append({"kind": name, "log": L, "line": line, "circle": circle})

    g["ntt_basis"] = []
    for L in [2, 3, 4]:
        c = pyref.half_coset(L)
        n = 1 << L
        ev = [pyref.evaluate([int(i == k) for i in range(n)], c) for k in range(n)]
        g["ntt_basis"].append({"log": L, "evaluate_unit": ev})

    g["ntt_random"] = []
    for L in [5, 6, 8]:
        n = 1 << L
        v = [rng.randrange(P) for _ in range(n)]
        c = pyref.half_coset(L)
        g["ntt_random"].append({"log": L, "in": v, "evaluate": pyref.evaluate(v, c),
                                "interpolate": pyref.interpolate(v, c)})

    g["fold_twiddles"] = []
    for L in [3, 4, 5, 6]:
        for name, c, use_y in [("half_coset_y", pyref.half_coset(L), True),
                               ("half_odds_x", pyref.half_odds(L), False)]:
            g["fold_twiddles"].append({"kind": name, "log": L,
                                       "tw": [pyref.fold_tw(c, i, use_y) for i in range(1 << (L - 1))]})

    g["channel"] = []
    ch = pyref.Channel()
    ops = []
    digest = [rng.randrange(1 << 32) for _ in range(8)]
    ch.mix_digest(digest)
    ops.append({"op": "mix_digest", "digest": digest, "state": ch.state.hex()})
    for _ in range(3):
        ops.append({"op": "draw_felt", "out": list(ch.draw_felt())})
    felts = [tuple(rng.randrange(P) for _ in range(4)) for _ in range(8)]
    ch.mix_felts(felts)
    ops.append({"op": "mix_felts", "felts": [list(f) for f in felts], "state": ch.state.hex()})
    ops.append({"op": "draw_number", "bound": 1 << 26, "out": ch.draw_number(1 << 26)})
    ch.mix_u64(0x123456789ABCDEF)
    ops.append({"op": "mix_u64", "v": 0x123456789ABCDEF, "state": ch.state.hex()})
    g["channel"] = ops

    g["pow"] = []
    for bits in [4, 8, 12]:
        c2 = pyref.Channel()
        c2.mix_digest([bits] * 8)
        nonce = 0
        while not c2.pow_ok(bits, nonce):
            nonce += 1
        g["pow"].append({"digest": [bits] * 8, "bits": bits, "nonce": nonce})

    g["proofs_py"] = []
    for a, b, log_n, bb, pw in [(1, 1, 4, 2, 0), (1, 1, 5, 2, 0), (1, 1, 6, 2, 0), (2, 3, 6, 2, 0),
                                (42, 99, 6, 2, 0), (1000, 2000, 6, 2, 0), (1, 1, 8, 2, 0),
                                (1, 1, 6, 1, 0), (1, 1, 5, 3, 0), (5, 8, 5, 2, 6)]:
        data = pyref.prove(a, b, log_n, blowup_bits=bb, pow_bits=pw)
        g["proofs_py"].append({"a": a, "b": b, "log_n": log_n, "blowup_bits": bb, "pow_bits": pw,
                               "len": len(data), "sha256": hashlib.sha256(data).hexdigest()})

    import pyoracle
    seed = 42
    a1 = fib_a1(seed)
    g["workload_seed"] = {"seed": seed, "a0": 1, "a1": a1}
    sizes = [(10, 2, 0), (12, 2, 0), (16, 2, 0), (16, 2, 10), (14, 1, 0)]
    if big:
        sizes += [(20, 2, 0), (24, 2, 0), (25, 1, 0), (24, 2, 26)]
    g["proofs_c"] = []
    for log_n, bb, pw in sizes:
        data = pyoracle.prove(1, a1, log_n, blowup_bits=bb, pow_bits=pw, omp=True)
        g["proofs_c"].append({"a": 1, "b": a1, "log_n": log_n, "blowup_bits": bb, "pow_bits": pw,
                              "len": len(data), "sha256": hashlib.sha256(data).hexdigest()})
        print("oracle digest", log_n, bb, pw, len(data), flush=True)

    # synthetic degree-2 multi-column AIR (KRAK v3): the Python restatement
    # at small sizes, the C oracle at larger ones
    g["proofs_d2_py"] = []
    for sd, log_n, w, k, air, bb, nq, pw in [(7, 4, 34, 35, 99, 2, 80, 0), (7, 4, 3, 2, 99, 1, 5, 0),
                                             (11, 5, 17, 4, 5, 2, 10, 6), (3, 5, 34, 35, 0xD2A1, 2, 80, 4)]:
        data = pyref.prove_d2(sd, log_n, w, k, air, bb, nq, pw)
        g["proofs_d2_py"].append({"seed": sd, "log_n": log_n, "n_cols": w, "n_cons": k, "air_seed": air,
                                  "blowup_bits": bb, "n_queries": nq, "pow_bits": pw, "len": len(data),
                                  "sha256": hashlib.sha256(data).hexdigest()})
    g["proofs_d2_c"] = []
    d2_sizes = [(10, 0), (14, 20), (16, 0)] + ([(20, 20)] if big else [])
    for log_n, pw in d2_sizes:
        data = pyoracle.d2_prove(seed, log_n, 34, 35, 0xD2A1, 2, 80, pw, omp=True)
        g["proofs_d2_c"].append({"seed": seed, "log_n": log_n, "n_cols": 34, "n_cons": 35, "air_seed": 0xD2A1,
                                 "blowup_bits": 2, "n_queries": 80, "pow_bits": pw, "len": len(data),
                                 "sha256": hashlib.sha256(data).hexdigest()})
        print("oracle d2 digest", log_n, pw, len(data), flush=True)

    # batched uneven commitment: 96 columns, sizes 2^14..2^22 from the seed,
    # blowup 4, bit reversed, lifted tree (C oracle)
    import vortexstark_amd as vx_sizes  # only the size schedule (pure Python)
    g["batch_c"] = []
    for n_cols, sd in [(12, 7)] + ([(96, seed)] if big else []):
        logs = vx_sizes.batch_log_sizes(n_cols, sd)
        root = pyoracle.batch_root(logs, sd, 2, True)
        g["batch_c"].append({"n_cols": n_cols, "seed": sd, "log_sizes": logs, "blowup_bits": 2,
                             "bit_reverse": True, "root": [int(x) for x in root]})
        print("oracle batch root", n_cols, flush=True)

    # batched-columns proof (KRAK v4): Python restatement at tiny sizes, the C
    # oracle at the workload's sizes
    g["proofs_batch_py"] = []
    for sd, logs, bb, nq, pw in [(5, [2, 3, 3, 4], 2, 8, 0), (9, [3], 1, 5, 0),
                                 (1, [2, 2, 2, 3, 3] + [4] * 14, 2, 80, 0), (3, [3, 5], 2, 16, 6)]:
        data = pyref.prove_batch(sd, logs, bb, nq, pw)
        g["proofs_batch_py"].append({"seed": sd, "log_sizes": logs, "blowup_bits": bb, "n_queries": nq,
                                     "pow_bits": pw, "len": len(data), "sha256": hashlib.sha256(data).hexdigest()})
    g["proofs_batch_c"] = []
    for n_cols, sd, pw in [(12, 7, 0)] + ([(96, seed, 0), (96, seed, 20)] if big else []):
        logs = vx_sizes.batch_log_sizes(n_cols, sd)
        data = pyoracle.batch_prove(sd, logs, 2, 80, pw)
        g["proofs_batch_c"].append({"n_cols": n_cols, "seed": sd, "log_sizes": logs, "blowup_bits": 2,
                                    "n_queries": 80, "pow_bits": pw, "len": len(data),
                                    "sha256": hashlib.sha256(data).hexdigest()})
        print("oracle batch proof", n_cols, pw, len(data), flush=True)

    if os.path.exists(out):
        g["serving_e"] = json.load(open(out)).get("serving_e", [])
        g["cairo_commit_c"] = json.load(open(out)).get("cairo_commit_c")
    if big:
        g["cairo_commit_c"] = cairo_commit_digest()
    if not big and os.path.exists(out):
        old = json.load(open(out))
        keep = [p for p in old.get("proofs_c", []) if p["log_n"] >= 20]
        g["proofs_c"] += keep
        g["proofs_d2_c"] += [p for p in old.get("proofs_d2_c", []) if p["log_n"] >= 20]
        g["batch_c"] += [p for p in old.get("batch_c", []) if p["n_cols"] >= 96]
        g["proofs_batch_c"] += [p for p in old.get("proofs_batch_c", []) if p["n_cols"] >= 96]
    with open(out, "w") as f:
        json.dump(g, f, separators=(",", ":"))
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
