"""vortexstark_amd: MI355X (gfx950) Circle-STARK prover hot path.

Python mirror of the reference's prove/verify API (src/prover.rs:221-262,
src/verifier.rs:39) over the C-ABI in include/vortex_hip.h. Proofs are KRAK
byte strings (src/bin/stark_cli.rs:1092-1120); v1 when pow_bits == 0.

    p = Prover(log_n=24)          # ProverCache::new: tables + buffers, once
    proof = p.prove(1, 1)         # prove_cached(a, b)
    ok, why = verify(proof)       # verifier::verify
"""
import ctypes as C
import os

from . import _lib
from ._lib import VortexError, load

BLOWUP_BITS = 2   # src/prover.rs:106-109
N_QUERIES = 80    # src/prover.rs:114-117
P = 0x7FFFFFFF

STAGES = ("trace", "interpolate", "lde", "trace_commit", "quotient_commit", "fri", "queries", "decommit",
          "download", "total", "trace_subtree_kernel")


def proof_size(log_n, blowup_bits=BLOWUP_BITS, n_queries=N_QUERIES, pow_bits=0):
    return int(load().vx_proof_size(log_n, blowup_bits, n_queries, pow_bits))


class Prover:
    """Cached prover for one size class (ProverCache + prove_cached).

    graphs=True replays a HIP graph of the whole proof (captured on the first
    untimed proof); False launches kernel by kernel. The bytes are the same."""

    def __init__(self, log_n, blowup_bits=BLOWUP_BITS, n_queries=N_QUERIES, pow_bits=0, stream=None, graphs=True):
        self.log_n, self.blowup_bits, self.n_queries, self.pow_bits = log_n, blowup_bits, n_queries, pow_bits
        self._h = C.c_void_p()
        _lib.call("vx_prover_create", log_n, blowup_bits, n_queries, pow_bits, stream, C.byref(self._h))
        self.size = proof_size(log_n, blowup_bits, n_queries, pow_bits)
        _lib.call("vx_prover_set_graphs", self._h, 1 if graphs else 0)

    def prove(self, a, b):
        self.enqueue(a, b)
        return self.wait()

    def enqueue(self, a, b, timed=False):
        _lib.call("vx_prover_enqueue", self._h, a, b, 1 if timed else 0)

    def wait(self):
        """Wait for the last enqueued proof; one copy out of pinned memory."""
        n = C.c_size_t()
        _lib.call("vx_prover_wait", self._h, None, 0, C.byref(n))
        ptr, _ = self.buffer(101)
        return C.string_at(ptr, n.value)

    def stage_times(self):
        out = (C.c_float * 11)()
        _lib.call("vx_prover_stage_times", self._h, out)
        return dict(zip(STAGES, list(out)))

    @property
    def device_bytes(self):
        return int(load().vx_prover_device_bytes(self._h))

    def buffer(self, which):
        """(device pointer, words) of an internal buffer of the last proof."""
        ptr, words = C.c_void_p(), C.c_uint64()
        _lib.call("vx_prover_buffer", self._h, which, C.byref(ptr), C.byref(words))
        return ptr.value, words.value

    def close(self):
        if self._h:
            _lib.call("vx_prover_destroy", self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


D2_COLS, D2_CONS, D2_AIR_SEED = 34, 35, 0xD2A1


def d2_proof_size(log_n, n_cols=D2_COLS, blowup_bits=BLOWUP_BITS, n_queries=N_QUERIES):
    return int(load().vx_d2_proof_size(log_n, n_cols, blowup_bits, n_queries))


class D2Prover(Prover):
    """Cached prover of the synthetic multi-column AIR (KRAK v3): n_cols
    seeded random M31 columns, n_cons degree-2 transition constraints."""

    def __init__(self, log_n, n_cols=D2_COLS, n_cons=D2_CONS, air_seed=D2_AIR_SEED, blowup_bits=BLOWUP_BITS,
                 n_queries=N_QUERIES, pow_bits=0, stream=None, graphs=True):
        self.log_n, self.blowup_bits, self.n_queries, self.pow_bits = log_n, blowup_bits, n_queries, pow_bits
        self.n_cols, self.n_cons, self.air_seed = n_cols, n_cons, air_seed
        self._h = C.c_void_p()
        _lib.call("vx_d2_prover_create", log_n, n_cols, n_cons, air_seed, blowup_bits, n_queries, pow_bits, stream,
                  C.byref(self._h))
        self.size = d2_proof_size(log_n, n_cols, blowup_bits, n_queries)
        _lib.call("vx_prover_set_graphs", self._h, 1 if graphs else 0)

    def prove(self, seed):
        self.enqueue(seed)
        return self.wait()

    def enqueue(self, seed, timed=False):
        _lib.call("vx_prover_enqueue", self._h, seed & 0xFFFFFFFF, seed >> 32, 1 if timed else 0)


def _u32_array(xs):
    return (C.c_uint32 * len(xs))(*xs)


def _ptr_array(ptrs):
    return (C.c_void_p * len(ptrs))(*ptrs)


class Batch:
    """Resident batch of M31 columns of uneven log sizes (sorted ascending,
    14..24): one run = work-list LDE at 2^blowup_bits, bit reverse, lifted
    Blake2s commitment (vx_batch_*)."""

    def __init__(self, log_sizes, blowup_bits=BLOWUP_BITS, bit_reverse=True, stream=None):
        self.log_sizes = list(log_sizes)
        self._h = C.c_void_p()
        _lib.call("vx_batch_create", _u32_array(self.log_sizes), len(self.log_sizes), blowup_bits,
                  1 if bit_reverse else 0, stream, C.byref(self._h))

    def column(self, c, which):
        """(device pointer, words): which 0 input, 1 coefficients, 2 evaluation."""
        ptr, words = C.c_void_p(), C.c_uint64()
        _lib.call("vx_batch_column", self._h, c, which, C.byref(ptr), C.byref(words))
        return ptr.value, words.value

    def fill_random(self, seed):
        """Input column c = the synthetic generator's column c for `seed`."""
        for c, L in enumerate(self.log_sizes):
            _lib.call("vx_m31_fill_random", self.column(c, 0)[0], L, seed, c, None)

    def run(self, timed=False):
        self.enqueue(timed)
        return self.wait()

    def enqueue(self, timed=False):
        _lib.call("vx_batch_enqueue", self._h, 1 if timed else 0)

    def wait(self):
        root = (C.c_uint32 * 8)()
        _lib.call("vx_batch_wait", self._h, root)
        return list(root)

    def stage_times(self):
        out = (C.c_float * 4)()
        _lib.call("vx_batch_stage_times", self._h, out)
        return dict(zip(("lde", "bit_reverse", "commit", "total"), list(out)))

    @property
    def device_bytes(self):
        return int(load().vx_batch_device_bytes(self._h))

    def close(self):
        if self._h:
            _lib.call("vx_batch_destroy", self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_proof_size(log_sizes, blowup_bits=BLOWUP_BITS, n_queries=N_QUERIES):
    return int(load().vx_batch_proof_size(_u32_array(list(log_sizes)), len(log_sizes), blowup_bits, n_queries))


class BatchProver(Prover):
    """Cached prover of the batched-columns proof (KRAK v4): columns of the
    given log sizes (sorted ascending, 14..24) from the synthetic generator,
    bit-reversed LDE, lifted tree, random-linear-combination quotient, FRI."""

    def __init__(self, log_sizes, blowup_bits=BLOWUP_BITS, n_queries=N_QUERIES, pow_bits=0, stream=None,
                 graphs=True):
        self.log_sizes = list(log_sizes)
        self.log_n, self.blowup_bits, self.n_queries, self.pow_bits = self.log_sizes[-1], blowup_bits, n_queries, pow_bits
        self._h = C.c_void_p()
        _lib.call("vx_batch_prover_create", _u32_array(self.log_sizes), len(self.log_sizes), blowup_bits, n_queries,
                  pow_bits, stream, C.byref(self._h))
        self.size = batch_proof_size(self.log_sizes, blowup_bits, n_queries)
        _lib.call("vx_prover_set_graphs", self._h, 1 if graphs else 0)

    def prove(self, seed):
        self.enqueue(seed)
        return self.wait()

    def enqueue(self, seed, timed=False):
        _lib.call("vx_prover_enqueue", self._h, seed & 0xFFFFFFFF, seed >> 32, 1 if timed else 0)


def batch_log_sizes(n_cols, seed, lo=14, hi=22):
    """The batched-commit workload's column sizes: SplitMix64(seed) outputs
    mod (hi - lo + 1) + lo, sorted ascending."""
    M = (1 << 64) - 1
    s, out = seed, []
    for _ in range(n_cols):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        out.append(lo + z % (hi - lo + 1))
    return sorted(out)


def prove(a, b, log_n, blowup_bits=BLOWUP_BITS, n_queries=N_QUERIES, pow_bits=0):
    """prover::prove (src/prover.rs:221): one-shot proof bytes."""
    size = proof_size(log_n, blowup_bits, n_queries, pow_bits)
    buf = C.create_string_buffer(size)
    n = C.c_size_t()
    _lib.call("vx_prove", a, b, log_n, blowup_bits, n_queries, pow_bits, buf, size, C.byref(n))
    return buf.raw[:n.value]


def verify(proof, blowup_bits=BLOWUP_BITS, n_queries=N_QUERIES):
    """verifier::verify (src/verifier.rs:39): (accepted, reason). Host code."""
    err = C.create_string_buffer(512)
    rc = load().vx_verify(proof, len(proof), blowup_bits, n_queries, err, 512)
    if rc not in (0, 4):
        raise VortexError("vx_verify", rc, err.value.decode())
    return rc == 0, err.value.decode()


AIR_FIBONACCI, AIR_D2 = 0, 1


def request(log_n, a=1, b=1, blowup_bits=BLOWUP_BITS, n_queries=N_QUERIES, pow_bits=0, air=AIR_FIBONACCI, seed=0,
            n_cols=D2_COLS, n_cons=D2_CONS, air_seed=D2_AIR_SEED):
    """A proof request (vx_request): Fibonacci (a, b), or with air=AIR_D2 the
    synthetic multi-column AIR over the trace of `seed`."""
    d2 = air == AIR_D2
    return _lib.Request(air, log_n, blowup_bits, n_queries, pow_bits, a if not d2 else 0, b if not d2 else 0,
                        n_cols if d2 else 0, n_cons if d2 else 0, 0, seed if d2 else 0, air_seed if d2 else 0)


TUNER_FIELDS = ("device", "leaf_kind", "n_cols", "n_cons", "lift_n", "log_n", "store_lo", "tile_log", "lane_log")


def tuner(enabled=True, cache_file=None):
    """Configure the Merkle-shape autotuner (process-wide): when enabled, the
    first commitment of each (leaf kind, columns, size) is timed with every
    valid subtree shape and the fastest is kept; cache_file persists the
    choices across processes. Proof bytes never depend on it."""
    _lib.call("vx_tuner_configure", 1 if enabled else 0,
              None if cache_file is None else os.fsencode(cache_file))


def tuner_entries():
    """The cached tuning choices, one dict per key."""
    n = C.c_uint32()
    _lib.call("vx_tuner_entries", None, 0, C.byref(n))
    buf = (C.c_uint32 * (9 * max(n.value, 1)))()
    _lib.call("vx_tuner_entries", buf, n.value, C.byref(n))
    return [dict(zip(TUNER_FIELDS, buf[9 * i:9 * i + 9])) for i in range(n.value)]


NTT_TUNER_FIELDS = ("device", "forward", "log_n", "bits", "n_cols", "stwo", "split", "order")


def tuner_ntt_entries():
    """The circle NTT's tuned plans, one dict per transform shape."""
    n = C.c_uint32()
    _lib.call("vx_tuner_ntt_entries", None, 0, C.byref(n))
    buf = (C.c_uint32 * (8 * max(n.value, 1)))()
    _lib.call("vx_tuner_ntt_entries", buf, n.value, C.byref(n))
    return [dict(zip(NTT_TUNER_FIELDS, buf[8 * i:8 * i + 8])) for i in range(n.value)]


def request_size(req):
    if req.air == AIR_D2:
        return d2_proof_size(req.log_n, req.n_cols, req.blowup_bits, req.n_queries)
    return proof_size(req.log_n, req.blowup_bits, req.n_queries, req.pow_bits)


class Server:
    """Continuous-batching proof server (vx_server_*): n_slots HIP streams,
    resident provers per size class admitted against mem_budget bytes of
    device memory. n_slots=1 is the one-proof-at-a-time loop of the
    reference's worker (src/bin/dashboard.rs:66-96)."""

    def __init__(self, n_slots=4, mem_budget=64 << 30, graphs=True):
        self._h = C.c_void_p()
        _lib.call("vx_server_create", n_slots, mem_budget, 1 if graphs else 0, C.byref(self._h))
        self.n_slots = n_slots
        self._bufs = {}

    def prepare(self, req):
        _lib.call("vx_server_prepare", self._h, C.byref(req))

    def submit(self, req):
        size = request_size(req)
        buf = C.create_string_buffer(size)
        t = C.c_uint64()
        _lib.call("vx_server_submit", self._h, C.byref(req), buf, size, C.byref(t))
        self._bufs[t.value] = buf
        return t.value

    def done(self, ticket):
        rc = load().vx_server_done(self._h, ticket)
        if rc < 0:
            _lib.check("vx_server_done", rc)
        return rc == 1

    def wait(self, ticket):
        """(proof bytes, (submit, start, done) seconds on the server clock)."""
        n = C.c_size_t()
        times = (C.c_double * 3)()
        buf = self._bufs.pop(ticket)
        _lib.call("vx_server_wait", self._h, ticket, C.byref(n), times)
        return buf.raw[:n.value], tuple(times)

    def run_trace(self, reqs, arrivals, keep=True):
        """Replay an arrival trace (seconds, sorted). Returns (proofs or
        None, times[n][3] = submit/start/done seconds from the trace start)."""
        n = len(reqs)
        arr = (_lib.Request * n)(*reqs)
        at = (C.c_double * n)(*arrivals)
        sizes = [request_size(r) for r in reqs]
        bufs = [C.create_string_buffer(sz) for sz in sizes]
        outs = (C.c_void_p * n)(*[C.addressof(b) for b in bufs])
        caps = (C.c_size_t * n)(*sizes)
        lens = (C.c_size_t * n)()
        times = (C.c_double * (3 * n))()
        _lib.call("vx_server_run_trace", self._h, arr, at, n, outs, caps, lens, times)
        t = [tuple(times[3 * i:3 * i + 3]) for i in range(n)]
        proofs = [bufs[i].raw[:lens[i]] for i in range(n)] if keep else None
        return proofs, t

    @property
    def resident_bytes(self):
        return int(load().vx_server_resident_bytes(self._h))

    @property
    def stats(self):
        """{'created', 'evicted', 'deferred'}: provers built, provers evicted,
        admissions postponed until a running proof or a copy-out freed memory."""
        out = (C.c_uint64 * 3)()
        _lib.call("vx_server_stats", self._h, out)
        return {"created": out[0], "evicted": out[1], "deferred": out[2]}

    def close(self):
        if self._h:
            _lib.call("vx_server_destroy", self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
