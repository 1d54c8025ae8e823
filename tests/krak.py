"""KRAK proof (src/bin/stark_cli.rs:1051-1230) parser for tests: returns the
fields together with their word offsets, so tests can tamper with one field
and re-serialise."""
import struct


class Krak:
    def __init__(self, data: bytes):
        self.words = list(struct.unpack(f"<{len(data) // 4}I", data))
        w = self.words
        p = 0

        def take(n=1):
            nonlocal p
            start = p
            p += n
            return start

        self.magic, self.version = w[0], w[1]
        take(2)
        self.log_n, self.a, self.b = w[2], w[3], w[4]
        take(3)
        if self.version == 2:
            self.blowup_bits, self.pow_bits = w[5], w[6]
            self.nonce = w[7] | (w[8] << 32)
            take(4)
        self.off_trace_root = take(8)
        self.off_quotient_root = take(8)
        n_fri = w[take()]
        self.off_fri_roots = [take(8) for _ in range(n_fri)]
        n_last = w[take()]
        self.off_last = take(4 * n_last)
        nq = w[take()]
        self.off_queries = take(nq)
        self.queries = w[self.off_queries:self.off_queries + nq]

        def decom(width):
            n = w[take()]
            d = {"n": n, "width": width, "values": take(n * width), "siblings": take(n * width),
                 "paths": [], "sibling_paths": []}
            for key in ("paths", "sibling_paths"):
                for _ in range(n):
                    depth = w[take()]
                    d[key].append((p, depth))
                    take(8 * depth)
            return d

        self.trace = decom(1)
        self.quotient = decom(4)
        n_fd = w[take()]
        self.off_n_fri_decom = p - 1
        self.fri = [decom(4) for _ in range(n_fd)]
        assert p == len(w), "trailing bytes in proof"

    def to_bytes(self):
        return struct.pack(f"<{len(self.words)}I", *self.words)

    def root(self, off):
        return self.words[off:off + 8]
