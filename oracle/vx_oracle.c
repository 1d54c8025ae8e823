/*
 * vx_oracle.c -- CPU restatement of the VortexSTARK Fibonacci prove path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the correctness oracle and the timed
 * CPU baseline. It is loaded by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, always as the checker and never as the thing
 * measured on the GPU. The product library (vortexstark_amd) does not link it.
 *
 * The reference is Rust + CUDA and has no CPU path for this prover; this file
 * restates its algorithm from the source text. Each function names the
 * reference file:line it follows. Arithmetic is exact over GF(2^31 - 1), so
 * evaluation order never changes a result.
 *
 * Built by oracle/Makefile into oracle/liboracle.so (single-threaded) and
 * oracle/liboracle_omp.so (OpenMP over the data-parallel loops; same bytes).
 */
#include "vx_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define P OR_P

/* ======================================================================
 * Field: src/field/m31.rs:24-35,86-113; cm31.rs:76-86; qm31.rs:113-132
 * ====================================================================== */
static inline uint32_t m31_add(uint32_t a, uint32_t b) {
    uint32_t r = a + b;
    return r >= P ? r - P : r;
}
static inline uint32_t m31_sub(uint32_t a, uint32_t b) {
    return a >= b ? a - b : a + P - b;
}
static inline uint32_t m31_reduce(uint64_t x) {
    uint32_t lo = (uint32_t)(x & P);
    uint32_t hi = (uint32_t)(x >> 31);
    uint32_t r = lo + hi;
    return r >= P ? r - P : r;
}
static inline uint32_t m31_mul(uint32_t a, uint32_t b) {
    return m31_reduce((uint64_t)a * (uint64_t)b);
}
static inline uint32_t m31_neg(uint32_t a) { return a == 0 ? 0 : P - a; }
static uint32_t m31_pow(uint32_t base, uint32_t e) {
    uint32_t r = 1;
    while (e) {
        if (e & 1) r = m31_mul(r, base);
        base = m31_mul(base, base);
        e >>= 1;
    }
    return r;
}
/* m31.rs:32-35: Fermat inverse, so inv(0) = 0 */
static inline uint32_t m31_inv(uint32_t a) { return m31_pow(a, P - 2); }

uint32_t or_m31_add(uint32_t a, uint32_t b) { return m31_add(a, b); }
uint32_t or_m31_sub(uint32_t a, uint32_t b) { return m31_sub(a, b); }
uint32_t or_m31_mul(uint32_t a, uint32_t b) { return m31_mul(a, b); }
uint32_t or_m31_inv(uint32_t a) { return m31_inv(a); }

typedef struct { uint32_t a, b; } cm31;
typedef struct { uint32_t v[4]; } qm31; /* [a.re, a.im, b.re, b.im] */

static inline cm31 cm_add(cm31 x, cm31 y) { return (cm31){m31_add(x.a, y.a), m31_add(x.b, y.b)}; }
static inline cm31 cm_sub(cm31 x, cm31 y) { return (cm31){m31_sub(x.a, y.a), m31_sub(x.b, y.b)}; }
/* cm31.rs:76-86: (a+bi)(c+di) = (ac-bd) + (ad+bc)i */
static inline cm31 cm_mul(cm31 x, cm31 y) {
    return (cm31){m31_sub(m31_mul(x.a, y.a), m31_mul(x.b, y.b)),
                  m31_add(m31_mul(x.a, y.b), m31_mul(x.b, y.a))};
}
static inline cm31 cm_scale(cm31 x, uint32_t s) { return (cm31){m31_mul(x.a, s), m31_mul(x.b, s)}; }
/* cm31.rs:38-51 */
static cm31 cm_inv(cm31 x) {
    uint32_t inv_norm = m31_inv(m31_add(m31_mul(x.a, x.a), m31_mul(x.b, x.b)));
    return (cm31){m31_mul(x.a, inv_norm), m31_mul(m31_neg(x.b), inv_norm)};
}
/* (p+qi)(2+i) = (2p-q) + (p+2q)i ; qm31.rs:120-124 */
static inline cm31 cm_mul_u2(cm31 x) {
    return (cm31){m31_sub(m31_add(x.a, x.a), x.b), m31_add(x.a, m31_add(x.b, x.b))};
}
static inline cm31 q_a(qm31 x) { return (cm31){x.v[0], x.v[1]}; }
static inline cm31 q_b(qm31 x) { return (cm31){x.v[2], x.v[3]}; }
static inline qm31 q_make(cm31 a, cm31 b) { return (qm31){{a.a, a.b, b.a, b.b}}; }
static inline qm31 q_add(qm31 x, qm31 y) {
    qm31 r;
    for (int i = 0; i < 4; i++) r.v[i] = m31_add(x.v[i], y.v[i]);
    return r;
}
static inline qm31 q_sub(qm31 x, qm31 y) {
    qm31 r;
    for (int i = 0; i < 4; i++) r.v[i] = m31_sub(x.v[i], y.v[i]);
    return r;
}
/* qm31.rs:113-132: (a+bu)(c+du) = (ac + bd(2+i)) + (ad+bc)u */
static inline qm31 q_mul(qm31 x, qm31 y) {
    cm31 ac = cm_mul(q_a(x), q_a(y)), bd = cm_mul(q_b(x), q_b(y));
    cm31 ad = cm_mul(q_a(x), q_b(y)), bc = cm_mul(q_b(x), q_a(y));
    return q_make(cm_add(ac, cm_mul_u2(bd)), cm_add(ad, bc));
}
static inline qm31 q_scale(qm31 x, uint32_t s) {
    qm31 r;
    for (int i = 0; i < 4; i++) r.v[i] = m31_mul(x.v[i], s);
    return r;
}
/* qm31.rs:67-91: conj / norm */
static qm31 q_inv(qm31 x) {
    cm31 a2 = cm_mul(q_a(x), q_a(x)), b2 = cm_mul(q_b(x), q_b(x));
    cm31 inv_norm = cm_inv(cm_sub(a2, cm_mul_u2(b2)));
    cm31 nb = (cm31){m31_neg(x.v[2]), m31_neg(x.v[3])};
    return q_make(cm_mul(q_a(x), inv_norm), cm_mul(nb, inv_norm));
}
static inline qm31 q_load(const uint32_t *p) { return (qm31){{p[0], p[1], p[2], p[3]}}; }

void or_cm31_mul(const uint32_t a[2], const uint32_t b[2], uint32_t out[2]) {
    cm31 r = cm_mul((cm31){a[0], a[1]}, (cm31){b[0], b[1]});
    out[0] = r.a;
    out[1] = r.b;
}
void or_qm31_mul(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]) {
    qm31 r = q_mul(q_load(a), q_load(b));
    memcpy(out, r.v, 16);
}
void or_qm31_inv(const uint32_t a[4], uint32_t out[4]) {
    qm31 r = q_inv(q_load(a));
    memcpy(out, r.v, 16);
}

/* ======================================================================
 * Circle group: src/circle.rs:12-185
 * ====================================================================== */
typedef struct { uint32_t x, y; } cpt;
static const cpt C_ID = {1, 0};
static const cpt C_GEN = {2, 1268011823u}; /* circle.rs:26-29 */

/* circle.rs:33-38 */
static inline cpt c_mul(cpt a, cpt b) {
    return (cpt){m31_sub(m31_mul(a.x, b.x), m31_mul(a.y, b.y)),
                 m31_add(m31_mul(a.x, b.y), m31_mul(a.y, b.x))};
}
static cpt c_rdouble(cpt p, uint32_t n) {
    for (uint32_t i = 0; i < n; i++) p = c_mul(p, p);
    return p;
}
/* circle.rs:65-76 (u32 scalar in the reference; u64 here, identical below 2^32) */
static cpt c_pow(cpt base, uint64_t k) {
    cpt r = C_ID;
    while (k) {
        if (k & 1) r = c_mul(r, base);
        base = c_mul(base, base);
        k >>= 1;
    }
    return r;
}

typedef struct { cpt init, step; uint32_t log_size; } coset_t;

static coset_t coset_unpack(const uint32_t c[5]) {
    coset_t r = {{c[0], c[1]}, {c[2], c[3]}, c[4]};
    return r;
}
static void coset_pack(coset_t c, uint32_t out[5]) {
    out[0] = c.init.x; out[1] = c.init.y; out[2] = c.step.x; out[3] = c.step.y; out[4] = c.log_size;
}
/* circle.rs:101-143 */
static coset_t coset_make(int kind, uint32_t L) {
    coset_t c;
    c.log_size = L;
    c.step = c_rdouble(C_GEN, 31 - L);
    if (kind == OR_SUBGROUP) c.init = C_ID;
    else if (kind == OR_HALF_COSET) c.init = c_rdouble(C_GEN, 30 - L);
    else c.init = c_rdouble(C_GEN, L >= 29 ? 0 : 29 - L); /* half_odds: saturating_sub */
    return c;
}
/* prover.rs:655-659: the eval domain is subgroup(31) when log_eval=31 */
static coset_t eval_domain(uint32_t L) {
    return L <= 30 ? coset_make(OR_HALF_COSET, L) : coset_make(OR_SUBGROUP, L);
}
/* circle.rs:258-267 */
static cpt coset_at(coset_t c, uint64_t i) { return c_mul(c.init, c_pow(c.step, i)); }

void or_coset(int kind, uint32_t log_size, uint32_t out[5]) { coset_pack(coset_make(kind, log_size), out); }
void or_eval_domain(uint32_t log_size, uint32_t out[5]) { coset_pack(eval_domain(log_size), out); }
void or_coset_at(const uint32_t c[5], uint64_t i, uint32_t out[2]) {
    cpt p = coset_at(coset_unpack(c), i);
    out[0] = p.x;
    out[1] = p.y;
}
void or_circle_mul(const uint32_t a[2], const uint32_t b[2], uint32_t out[2]) {
    cpt p = c_mul((cpt){a[0], a[1]}, (cpt){b[0], b[1]});
    out[0] = p.x;
    out[1] = p.y;
}

/* ======================================================================
 * Blake2s-256 (RFC 7693), as src/channel.rs:216-298 and cuda/include/
 * blake2s.cuh:53-136: h0 ^= 0x01010020, t0 = byte length, final flag.
 * ====================================================================== */
static const uint32_t B2_IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                  0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static const uint8_t B2_SIGMA[10][16] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15},
    {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
    {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4},
    {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
    {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13},
    {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
    {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11},
    {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
    {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5},
    {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0},
};
static inline uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static void b2_compress(uint32_t h[8], const uint32_t m[16], uint64_t t, int last) {
    uint32_t v[16];
    for (int i = 0; i < 8; i++) { v[i] = h[i]; v[i + 8] = B2_IV[i]; }
    v[12] ^= (uint32_t)t;
    v[13] ^= (uint32_t)(t >> 32);
    if (last) v[14] ^= 0xFFFFFFFFu;
#define G2(a, b, c, d, x, y)                 \
    do {                                     \
        v[a] = v[a] + v[b] + (x);            \
        v[d] = rotr32(v[d] ^ v[a], 16);      \
        v[c] = v[c] + v[d];                  \
        v[b] = rotr32(v[b] ^ v[c], 12);      \
        v[a] = v[a] + v[b] + (y);            \
        v[d] = rotr32(v[d] ^ v[a], 8);       \
        v[c] = v[c] + v[d];                  \
        v[b] = rotr32(v[b] ^ v[c], 7);       \
    } while (0)
    for (int r = 0; r < 10; r++) {
        const uint8_t *s = B2_SIGMA[r];
        G2(0, 4, 8, 12, m[s[0]], m[s[1]]);
        G2(1, 5, 9, 13, m[s[2]], m[s[3]]);
        G2(2, 6, 10, 14, m[s[4]], m[s[5]]);
        G2(3, 7, 11, 15, m[s[6]], m[s[7]]);
        G2(0, 5, 10, 15, m[s[8]], m[s[9]]);
        G2(1, 6, 11, 12, m[s[10]], m[s[11]]);
        G2(2, 7, 8, 13, m[s[12]], m[s[13]]);
        G2(3, 4, 9, 14, m[s[14]], m[s[15]]);
    }
#undef G2
    for (int i = 0; i < 8; i++) h[i] ^= v[i] ^ v[i + 8];
}
static inline uint32_t le32(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
static inline void put32(uint8_t *p, uint32_t v) {
    p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); p[2] = (uint8_t)(v >> 16); p[3] = (uint8_t)(v >> 24);
}

/* Standard unkeyed Blake2s-256 over any length (the `blake2` crate path used
 * by Channel::mix_felts, src/channel.rs:46-59). */
void or_blake2s(const uint8_t *in, size_t len, uint8_t out[32]) {
    uint32_t h[8], m[16];
    memcpy(h, B2_IV, sizeof h);
    h[0] ^= 0x01010020u;
    uint64_t t = 0;
    size_t off = 0;
    while (len - off > 64) {
        for (int i = 0; i < 16; i++) m[i] = le32(in + off + 4 * i);
        t += 64;
        b2_compress(h, m, t, 0);
        off += 64;
    }
    uint8_t blk[64] = {0};
    memcpy(blk, in + off, len - off);
    t += len - off;
    for (int i = 0; i < 16; i++) m[i] = le32(blk + 4 * i);
    b2_compress(h, m, t, 1);
    for (int i = 0; i < 8; i++) put32(out + 4 * i, h[i]);
}
/* src/channel.rs:254-257: the single-block helper hashes min(len, 64) bytes. */
void or_blake2s_single(const uint8_t *in, size_t len, uint8_t out[32]) {
    or_blake2s(in, len < 64 ? len : 64, out);
}

/* Word-level helpers matching cuda/blake2s.cu:13-87 (leaf of n<=16 words,
 * node of two 8-word children). */
static void hash_words(const uint32_t *w, uint32_t nw, uint32_t out[8]) {
    uint32_t h[8], m[16] = {0};
    memcpy(h, B2_IV, sizeof h);
    h[0] ^= 0x01010020u;
    if (nw > 16) nw = 16;
    for (uint32_t i = 0; i < nw; i++) m[i] = w[i];
    b2_compress(h, m, 4u * nw, 1);
    memcpy(out, h, 32);
}
static void hash_node(const uint32_t *l, const uint32_t *r, uint32_t out[8]) {
    uint32_t m[16];
    memcpy(m, l, 32);
    memcpy(m + 8, r, 32);
    hash_words(m, 16, out);
}

/* ======================================================================
 * Twiddles: src/circle.rs:284-347 (compute_twiddles), cuda/twiddles.cu:
 * 80-142 (coset points, extract_and_squash), cuda/batch_inverse.cu:8-41.
 * ====================================================================== */
void or_compute_twiddles(const uint32_t c5[5], uint32_t *line_tw, uint32_t *circle_tw,
                         uint32_t *offsets, uint32_t *sizes) {
    coset_t c = coset_unpack(c5);
    uint64_t n = 1ull << c.log_size;
    uint32_t *xs = (uint32_t *)malloc(n * 4);
    /* coset points p_i = init * step^i (twiddles.cu:80-105); sequential
     * multiplication gives identical exact values. */
    cpt p = c.init;
    for (uint64_t i = 0; i < n; i++) {
        xs[i] = p.x;
        if (i < n / 2) circle_tw[i] = p.y;
        p = c_mul(p, c.step);
    }
    uint32_t n_line_layers = c.log_size > 0 ? c.log_size - 1 : 0;
    uint64_t cur_n = n, off = 0;
    for (uint32_t l = 0; l < n_line_layers; l++) {
        uint64_t half = cur_n / 2;
        offsets[l] = (uint32_t)off;
        sizes[l] = (uint32_t)half;
        /* extract_and_squash: tw[i] = x[2i]; next[i] = 2 x[2i]^2 - 1 */
        for (uint64_t i = 0; i < half; i++) {
            uint32_t x = xs[2 * i];
            line_tw[off + i] = x;
            uint32_t x2 = m31_mul(x, x);
            xs[i] = m31_sub(m31_add(x2, x2), 1);
        }
        off += half;
        cur_n = half;
    }
    free(xs);
}

/* batch_inverse.cu:8-41: inverse with inverse(0) := 0 (chunking does not
 * change exact results). */
void or_batch_inverse(const uint32_t *in, uint32_t *out, size_t n) {
    for (size_t i = 0; i < n; i++) out[i] = in[i] == 0 ? 0 : m31_inv(in[i]);
}

static inline uint64_t bitrev(uint64_t x, uint32_t bits) {
    uint64_t r = 0;
    for (uint32_t i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

/* twiddles.cu:23-52 + batch inverse (src/fri.rs:112-139):
 * out[i] = coord(D.at(bitrev_L(2i)))^-1 for i < |D|/2 */
void or_fold_twiddles(const uint32_t c5[5], int extract_y, uint32_t *out) {
    coset_t c = coset_unpack(c5);
    uint64_t half = (1ull << c.log_size) / 2;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)half; i++) {
        cpt p = coset_at(c, bitrev((uint64_t)i << 1, c.log_size));
        uint32_t v = extract_y ? p.y : p.x;
        out[i] = v == 0 ? 0 : m31_inv(v);
    }
}

/* ======================================================================
 * Twisted circle NTT: cuda/circle_ntt.cu:74-117 (butterflies and layer
 * indexing), :157-165 (scale), :248-312 (layer order).
 * ====================================================================== */
void or_ntt_layer(uint32_t *d, const uint32_t *tw, uint32_t li, uint64_t n, int forward) {
    uint64_t half = n / 2, stride = 1ull << li;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < (int64_t)half; t++) {
        uint64_t h = (uint64_t)t >> li, l = (uint64_t)t & (stride - 1);
        uint64_t i0 = (h << (li + 1)) + l, i1 = i0 + stride;
        uint32_t v0 = d[i0], v1 = d[i1], w = tw[h];
        if (forward) {
            uint32_t tmp = m31_mul(v1, w);
            d[i0] = m31_add(v0, tmp);
            d[i1] = m31_sub(v0, tmp);
        } else {
            d[i0] = m31_add(v0, v1);
            d[i1] = m31_mul(m31_sub(v0, v1), w);
        }
    }
}
void or_ntt_evaluate(uint32_t *d, const uint32_t *line_tw, const uint32_t *circle_tw,
                     const uint32_t *offsets, uint32_t n_line_layers, uint64_t n) {
    for (int l = (int)n_line_layers - 1; l >= 0; l--)
        or_ntt_layer(d, line_tw + offsets[l], (uint32_t)l + 1, n, 1);
    or_ntt_layer(d, circle_tw, 0, n, 1);
}
void or_ntt_interpolate(uint32_t *d, const uint32_t *line_itw, const uint32_t *circle_itw,
                        const uint32_t *offsets, uint32_t n_line_layers, uint64_t n) {
    or_ntt_layer(d, circle_itw, 0, n, 0);
    for (uint32_t l = 0; l < n_line_layers; l++)
        or_ntt_layer(d, line_itw + offsets[l], l + 1, n, 0);
    uint32_t log_n = 0;
    for (uint64_t t = n; t > 1; t >>= 1) log_n++;
    uint32_t e = (30u * log_n) % 31u;
    uint32_t inv_n = e == 0 ? 1u : (1u << e);
    for (uint64_t i = 0; i < n; i++) d[i] = m31_mul(d[i], inv_n);
}
/* Single outputs of the forward transform above (or_ntt_evaluate on the
 * domain of coset dom, 2^dom_log points) applied to the zero-padded
 * coefficients c[0 .. 2^log_n), each in O(2^log_n) instead of the whole
 * O(N log N) transform.
 *
 * Unrolling the butterflies of cuda/circle_ntt.cu:74-117 from output i back
 * to the input: layer j pairs positions that differ in bit j; the output's
 * bits above j fix the twiddle index h = i >> (j + 1) and bit j of i picks
 * + or -, so input k reaches output i with the factor
 *     prod over bits j set in k of  s_j(i) = (-1)^{i_j} w_j(i),
 * w_0(i) = y(p_{i >> 1}) (circle layer), w_j(i) = x(pi^{j-1}(p_{2^j (i >> (j+1))}))
 * for j >= 1 (line layer j - 1), p_t = init * step^t and pi(x) = 2x^2 - 1:
 * the closed form of the tables or_compute_twiddles builds (src/circle.rs:
 * 284-347, circle_tw[h] = y(p_h), line layer l entry h = x of p_{2^{l+1} h}
 * squashed l times). Bits j >= log_n of k are zero (zero padding), so only
 * the first log_n layers contribute (the others replicate).
 *
 * Outputs that share all bits above bit 1 share s_j for j >= 2, so they are
 * evaluated together: for each such group the 4 partial sums over the
 * coefficients with each value of bits 0..1 are accumulated once (2^log_n
 * multiply-adds per group), then combined per output with its own s_0, s_1. */
static uint32_t ntt_twiddle_at(coset_t c, uint32_t j, uint64_t i) {
    if (j == 0) return coset_at(c, i >> 1).y;
    uint32_t x = coset_at(c, (i >> (j + 1)) << j).x;
    for (uint32_t r = 1; r < j; r++) {
        uint32_t x2 = m31_mul(x, x);
        x = m31_sub(m31_add(x2, x2), 1);
    }
    return x;
}
static uint32_t ntt_sign_twiddle(coset_t c, uint32_t j, uint64_t i) {
    uint32_t w = ntt_twiddle_at(c, j, i);
    return ((i >> j) & 1) ? m31_neg(w) : w;
}
/* table[t] = prod over bits b of t set of f[b], t < 2^nb */
static void product_table(const uint32_t *f, uint32_t nb, uint32_t *table) {
    table[0] = 1;
    for (uint32_t b = 0; b < nb; b++)
        for (uint64_t t = 0; t < (1ull << b); t++) table[t + (1ull << b)] = m31_mul(table[t], f[b]);
}
static inline uint64_t fold62(uint64_t v) { return (v & P) + (v >> 31); }
static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y);
}
void or_ntt_evaluate_at(const uint32_t *c, uint32_t log_n, const uint32_t dom5[5], const uint64_t *idx,
                        uint32_t n_idx, uint32_t *out) {
    if (n_idx == 0) return;
    coset_t d = coset_unpack(dom5);
    const uint32_t m = log_n < 2 ? log_n : 2;       /* bits combined per output */
    const uint32_t B = log_n < 16 ? log_n : 16;     /* bits per coefficient block */
    const uint32_t mid_bits = B - m, hi_bits = log_n - B;
    const uint64_t n_low = 1ull << m, n_mid = 1ull << mid_bits, n_hi = 1ull << hi_bits;
    /* distinct groups (output index >> m) */
    uint64_t *grp = (uint64_t *)malloc(8 * (size_t)n_idx);
    for (uint32_t k = 0; k < n_idx; k++) grp[k] = idx[k] >> m;
    qsort(grp, n_idx, 8, cmp_u64);
    uint32_t ng = 0;
    for (uint32_t k = 0; k < n_idx; k++)
        if (k == 0 || grp[k] != grp[ng - 1]) grp[ng++] = grp[k];
    uint32_t *wmid = (uint32_t *)malloc(4 * (size_t)ng * n_mid);
    uint32_t *whi = (uint32_t *)malloc(4 * (size_t)ng * n_hi);
    uint64_t *acc = (uint64_t *)calloc((size_t)ng * n_low, 8);
    for (uint32_t g = 0; g < ng; g++) {
        uint32_t f[32];
        const uint64_t base = grp[g] << m;
        for (uint32_t j = m; j < B; j++) f[j - m] = ntt_sign_twiddle(d, j, base);
        product_table(f, mid_bits, wmid + (size_t)g * n_mid);
        for (uint32_t j = B; j < log_n; j++) f[j - B] = ntt_sign_twiddle(d, j, base);
        product_table(f, hi_bits, whi + (size_t)g * n_hi);
    }
#pragma omp parallel
    {
        uint64_t *part = (uint64_t *)calloc((size_t)ng * n_low, 8);
#pragma omp for schedule(static)
        for (int64_t h = 0; h < (int64_t)n_hi; h++) {
            const uint32_t *blk = c + ((uint64_t)h << B);
            for (uint32_t g = 0; g < ng; g++) {
                const uint32_t *wm = wmid + (size_t)g * n_mid;
                uint64_t a[4] = {0, 0, 0, 0};
                for (uint64_t t = 0; t < n_mid; t++)
                    for (uint64_t lo = 0; lo < n_low; lo++)
                        a[lo] += fold62((uint64_t)blk[(t << m) | lo] * wm[t]);
                const uint32_t wh = whi[(size_t)g * n_hi + (uint64_t)h];
                for (uint64_t lo = 0; lo < n_low; lo++)
                    part[(size_t)g * n_low + lo] += fold62((uint64_t)(uint32_t)(a[lo] % P) * wh);
            }
        }
#pragma omp critical
        for (size_t k = 0; k < (size_t)ng * n_low; k++) acc[k] = (acc[k] + part[k] % P) % P;
        free(part);
    }
    for (uint32_t k = 0; k < n_idx; k++) {
        uint64_t lo_g = idx[k] >> m;
        uint32_t g = 0, hi = ng;
        while (hi - g > 1) { /* binary search in the sorted distinct groups */
            uint32_t mid = (g + hi) / 2;
            if (grp[mid] <= lo_g) g = mid; else hi = mid;
        }
        uint32_t f[2], tab[4];
        for (uint32_t j = 0; j < m; j++) f[j] = ntt_sign_twiddle(d, j, idx[k]);
        product_table(f, m, tab);
        uint32_t v = 0;
        for (uint64_t lo = 0; lo < n_low; lo++)
            v = m31_add(v, m31_mul((uint32_t)acc[(size_t)g * n_low + lo], tab[lo]));
        out[k] = v;
    }
    free(grp);
    free(wmid);
    free(whi);
    free(acc);
}

/* circle_ntt.cu:234-244 */
void or_bit_reverse_m31(uint32_t *d, uint32_t log_n) {
    uint64_t n = 1ull << log_n;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t j = bitrev(i, log_n);
        if (i < j) { uint32_t t = d[i]; d[i] = d[j]; d[j] = t; }
    }
}

/* half-coset natural order -> canonic bit-reversed order and back
 * (src/circle.rs:192-220, cuda/permute.cu:30-57): natural index k of
 * half_coset(log_n) is canonic index cn = k / 2 (k even) or n - 1 - k / 2
 * (k odd), stored at bit-reverse(cn). */
static uint64_t hc_to_brt(uint64_t k, uint32_t log_n) {
    uint64_t n = 1ull << log_n;
    uint64_t cn = (k & 1) ? n - 1 - k / 2 : k / 2;
    return bitrev(cn, log_n);
}
void or_permute_hc_to_canonic_brt(const uint32_t *src, uint32_t *dst, uint32_t log_n) {
    uint64_t n = 1ull << log_n;
    for (uint64_t k = 0; k < n; k++) dst[hc_to_brt(k, log_n)] = src[k];
}
void or_permute_canonic_brt_to_hc_natural(const uint32_t *src, uint32_t *dst, uint32_t log_n) {
    uint64_t n = 1ull << log_n;
    for (uint64_t k = 0; k < n; k++) dst[k] = src[hc_to_brt(k, log_n)];
}

/* ======================================================================
 * AIR: src/air.rs:30-45 ; constraint: cuda/constraint.cu:10-36, 81-90
 * ====================================================================== */
void or_fibonacci_trace(uint32_t a, uint32_t b, uint32_t log_n, uint32_t *out) {
    uint64_t n = 1ull << log_n;
    out[0] = a;
    out[1] = b;
    for (uint64_t i = 2; i < n; i++) out[i] = m31_add(out[i - 1], out[i - 2]);
}
void or_zero_pad(const uint32_t *src, uint32_t *dst, uint64_t src_n, uint64_t dst_n) {
    for (uint64_t i = 0; i < dst_n; i++) dst[i] = i < src_n ? src[i] : 0;
}
void or_fibonacci_quotient(const uint32_t *t, uint32_t *o0, uint32_t *o1, uint32_t *o2,
                           uint32_t *o3, const uint32_t alpha[4], uint64_t n) {
    qm31 al = q_load(alpha);
#pragma omp parallel for schedule(static)
    for (int64_t si = 0; si < (int64_t)n; si++) {
        uint64_t i = (uint64_t)si;
        uint32_t c = m31_sub(t[(i + 2) % n], m31_add(t[(i + 1) % n], t[i]));
        qm31 r = q_scale(al, c);
        o0[i] = r.v[0]; o1[i] = r.v[1]; o2[i] = r.v[2]; o3[i] = r.v[3];
    }
}

/* ======================================================================
 * FRI folds: cuda/fri.cu:58-87 (fold_line), :92-126 (fold_circle_into_line)
 * ====================================================================== */
static inline qm31 fold_pair(qm31 f0, qm31 f1, qm31 alpha, uint32_t tw) {
    return q_add(q_add(f0, f1), q_mul(alpha, q_scale(q_sub(f0, f1), tw)));
}
void or_fold_line(const uint32_t *const in[4], const uint32_t *tw, uint32_t *const out[4],
                  const uint32_t alpha[4], uint64_t half_n) {
    qm31 al = q_load(alpha);
#pragma omp parallel for schedule(static)
    for (int64_t si = 0; si < (int64_t)half_n; si++) {
        uint64_t i = (uint64_t)si;
        qm31 f0 = {{in[0][2 * i], in[1][2 * i], in[2][2 * i], in[3][2 * i]}};
        qm31 f1 = {{in[0][2 * i + 1], in[1][2 * i + 1], in[2][2 * i + 1], in[3][2 * i + 1]}};
        qm31 r = fold_pair(f0, f1, al, tw[i]);
        for (int k = 0; k < 4; k++) out[k][i] = r.v[k];
    }
}
void or_fold_circle_into_line(uint32_t *const dst[4], const uint32_t *const src[4],
                              const uint32_t *tw, const uint32_t alpha[4],
                              const uint32_t alpha_sq[4], uint64_t half_n) {
    qm31 al = q_load(alpha), al2 = q_load(alpha_sq);
#pragma omp parallel for schedule(static)
    for (int64_t si = 0; si < (int64_t)half_n; si++) {
        uint64_t i = (uint64_t)si;
        qm31 f0 = {{src[0][2 * i], src[1][2 * i], src[2][2 * i], src[3][2 * i]}};
        qm31 f1 = {{src[0][2 * i + 1], src[1][2 * i + 1], src[2][2 * i + 1], src[3][2 * i + 1]}};
        qm31 prev = {{dst[0][i], dst[1][i], dst[2][i], dst[3][i]}};
        qm31 r = q_add(q_mul(prev, al2), fold_pair(f0, f1, al, tw[i]));
        for (int k = 0; k < 4; k++) dst[k][i] = r.v[k];
    }
}

/* ======================================================================
 * Merkle: cuda/blake2s.cu:13-87, src/merkle.rs:520-593, 841-884
 * ====================================================================== */
void or_merkle_leaf_hashes(const uint32_t *const *cols, uint32_t n_cols, uint64_t n_leaves,
                           uint32_t *hashes) {
    uint32_t nc = n_cols > 16 ? 16 : n_cols;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n_leaves; i++) {
        uint32_t w[16];
        for (uint32_t c = 0; c < nc; c++) w[c] = cols[c][i];
        hash_words(w, nc, hashes + 8 * (uint64_t)i);
    }
}
void or_merkle_hash_nodes(const uint32_t *children, uint32_t *parents, uint64_t n_parents) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n_parents; i++)
        hash_node(children + 16 * (uint64_t)i, children + 16 * (uint64_t)i + 8, parents + 8 * (uint64_t)i);
}

/* A full tree: layers[0] = leaf hashes (n), layers[k] = n >> k nodes. */
typedef struct {
    uint32_t log_n;
    uint32_t *nodes;      /* all layers concatenated, 8 words per node */
    uint64_t off[40];     /* word offset of each layer */
} tree_t;

static void tree_build_from_leaves(tree_t *t, uint32_t log_n, uint32_t *leaf_hashes_owned) {
    uint64_t n = 1ull << log_n;
    t->log_n = log_n;
    t->nodes = (uint32_t *)realloc(leaf_hashes_owned, (2 * n - 1) * 32);
    uint64_t o = 0, cnt = n;
    for (uint32_t k = 0; k <= log_n; k++) {
        t->off[k] = o;
        o += cnt * 8;
        cnt >>= 1;
    }
    for (uint32_t k = 1; k <= log_n; k++)
        or_merkle_hash_nodes(t->nodes + t->off[k - 1], t->nodes + t->off[k], n >> k);
}
static void tree_commit(tree_t *t, const uint32_t *const *cols, uint32_t n_cols, uint32_t log_n) {
    uint64_t n = 1ull << log_n;
    uint32_t *leaves = (uint32_t *)malloc(n * 32);
    or_merkle_leaf_hashes(cols, n_cols, n, leaves);
    tree_build_from_leaves(t, log_n, leaves);
}
static const uint32_t *tree_root(const tree_t *t) { return t->nodes + t->off[t->log_n]; }
static void tree_free(tree_t *t) { free(t->nodes); t->nodes = NULL; }

void or_merkle_root(const uint32_t *const *cols, uint32_t n_cols, uint32_t log_n, uint32_t root[8]) {
    tree_t t;
    tree_commit(&t, cols, n_cols, log_n);
    memcpy(root, tree_root(&t), 32);
    tree_free(&t);
}

/* src/merkle.rs:520-566 */
int or_verify_auth_path(const uint32_t root[8], const uint32_t leaf[8], uint64_t index,
                        const uint32_t *path, uint32_t depth) {
    uint32_t cur[8];
    memcpy(cur, leaf, 32);
    for (uint32_t k = 0; k < depth; k++) {
        const uint32_t *sib = path + 8 * (uint64_t)k;
        if (index % 2 == 0) hash_node(cur, sib, cur);
        else hash_node(sib, cur, cur);
        index /= 2;
    }
    return memcmp(cur, root, 32) == 0;
}

/* ======================================================================
 * Channel: src/channel.rs:15-185
 * ====================================================================== */
void or_channel_init(or_channel *ch) { memset(ch, 0, sizeof *ch); }

/* channel.rs:30-38 */
void or_channel_mix_digest(or_channel *ch, const uint32_t d[8]) {
    uint8_t in[64];
    memcpy(in, ch->state, 32);
    for (int i = 0; i < 8; i++) put32(in + 32 + 4 * i, d[i]);
    or_blake2s_single(in, 64, ch->state);
    ch->n_draws = 0;
}
/* channel.rs:46-59 */
void or_channel_mix_felts(or_channel *ch, const uint32_t *felts, size_t n_felts) {
    if (n_felts == 0) return;
    size_t len = 32 + 16 * n_felts;
    uint8_t *in = (uint8_t *)malloc(len);
    memcpy(in, ch->state, 32);
    for (size_t i = 0; i < 4 * n_felts; i++) put32(in + 32 + 4 * i, felts[i]);
    or_blake2s(in, len, ch->state);
    free(in);
    ch->n_draws = 0;
}
/* channel.rs:177-184 */
static void ch_squeeze(or_channel *ch, uint8_t out[32]) {
    uint8_t in[37];
    memcpy(in, ch->state, 32);
    put32(in + 32, ch->n_draws);
    in[36] = 0;
    ch->n_draws++;
    or_blake2s_single(in, 37, out);
}
/* channel.rs:66-83 */
void or_channel_draw_felt(or_channel *ch, uint32_t out[4]) {
    for (;;) {
        uint8_t b[32];
        ch_squeeze(ch, b);
        int ok = 1;
        for (int i = 0; i < 8; i++)
            if (le32(b + 4 * i) >= 2 * P) ok = 0;
        if (ok) {
            for (int i = 0; i < 4; i++) out[i] = le32(b + 4 * i) % P;
            return;
        }
    }
}
/* channel.rs:91-95 */
uint64_t or_channel_draw_number(or_channel *ch, uint64_t bound) {
    uint8_t b[32];
    ch_squeeze(ch, b);
    uint64_t raw = (uint64_t)le32(b) | ((uint64_t)le32(b + 4) << 32);
    return raw % bound;
}
/* channel.rs:120-129 */
void or_channel_mix_u64(or_channel *ch, uint64_t v) {
    uint8_t in[40];
    memcpy(in, ch->state, 32);
    put32(in + 32, (uint32_t)v);
    put32(in + 36, (uint32_t)(v >> 32));
    or_blake2s_single(in, 40, ch->state);
    ch->n_draws = 0;
}
/* channel.rs:140-168 */
static void pow_prefix(const or_channel *ch, uint32_t pow_bits, uint8_t prefix[32]) {
    uint8_t in[52] = {0};
    put32(in, 0x12345678u);
    memcpy(in + 16, ch->state, 32);
    put32(in + 48, pow_bits);
    or_blake2s_single(in, 52, prefix);
}
static uint32_t pow_tz(const uint8_t prefix[32], uint64_t nonce) {
    uint8_t in[40], r[32];
    memcpy(in, prefix, 32);
    put32(in + 32, (uint32_t)nonce);
    put32(in + 36, (uint32_t)(nonce >> 32));
    or_blake2s_single(in, 40, r);
    uint64_t lo = (uint64_t)le32(r) | ((uint64_t)le32(r + 4) << 32);
    uint64_t hi = (uint64_t)le32(r + 8) | ((uint64_t)le32(r + 12) << 32);
    if (lo != 0) return (uint32_t)__builtin_ctzll(lo);
    return 64 + (hi ? (uint32_t)__builtin_ctzll(hi) : 64);
}
int or_channel_verify_pow_nonce(const or_channel *ch, uint32_t pow_bits, uint64_t nonce) {
    uint8_t prefix[32];
    pow_prefix(ch, pow_bits, prefix);
    return pow_tz(prefix, nonce) >= pow_bits;
}
/* Deterministic grind: the smallest nonce in 0,1,2,... that passes
 * verify_pow_nonce (the search order the GPU grind reproduces). */
uint64_t or_grind_pow(const or_channel *ch, uint32_t pow_bits) {
    uint8_t prefix[32];
    pow_prefix(ch, pow_bits, prefix);
    for (uint64_t n = 0;; n++)
        if (pow_tz(prefix, n) >= pow_bits) return n;
}

/* The stwo plugin's grind (crates/cuda-backend/src/grind_ops.rs:71-128):
 * nonces (hi << 32) | low, low < 2^20, for hi = 0, 1, ...; the smallest
 * valid low of the first hi that has one. */
uint64_t or_grind_pow_stwo(const or_channel *ch, uint32_t pow_bits) {
    uint8_t prefix[32];
    pow_prefix(ch, pow_bits, prefix);
    for (uint64_t hi = 0;; hi++)
        for (uint64_t low = 0; low < (1u << 20); low++)
            if (pow_tz(prefix, (hi << 32) | low) >= pow_bits) return (hi << 32) | low;
}

/* The grind test over a caller-given 8-word prefixed digest. m31_out
 * reduces every output word mod P before the trailing-zero count, as stwo's
 * Blake2sChannelGeneric<true> does (grind_ops.rs:154-180; the device form
 * cuda/grind_m31_output.cu:44-63). The reduction here is the plain x % P. */
uint32_t or_pow_tz_words(const uint32_t prefix[8], uint64_t nonce, int m31_out) {
    uint8_t in[40], r[32];
    for (int i = 0; i < 8; i++) put32(in + 4 * i, prefix[i]);
    put32(in + 32, (uint32_t)nonce);
    put32(in + 36, (uint32_t)(nonce >> 32));
    or_blake2s_single(in, 40, r);
    for (int i = 0; i < 4; i++) {
        uint32_t w = le32(r + 4 * i);
        if (m31_out) w %= 0x7FFFFFFFu;
        if (w) return 32u * (uint32_t)i + (uint32_t)__builtin_ctz(w);
    }
    return 128;
}
/* The plugin's search order (grind_ops.rs:71-128, :190-240) over a given
 * prefixed digest. */
uint64_t or_grind_pow_stwo_words(const uint32_t prefix[8], uint32_t pow_bits, int m31_out) {
    for (uint64_t hi = 0;; hi++)
        for (uint64_t low = 0; low < (1u << 20); low++)
            if (or_pow_tz_words(prefix, (hi << 32) | low, m31_out) >= pow_bits) return (hi << 32) | low;
}

/* ======================================================================
 * Proof bytes: KRAK v1 (src/bin/stark_cli.rs:1051-1110). pow_bits > 0 selects
 * the KRAK v2 extension: header gains blowup_bits, pow_bits, nonce(lo,hi);
 * everything after the header is unchanged.
 * ====================================================================== */
#define KRAK_MAGIC 0x4B52414Bu

typedef struct { uint8_t *p; size_t len, cap; } wbuf;
static void wb_u32(wbuf *w, uint32_t v) {
    if (w->len + 4 > w->cap) {
        w->cap = w->cap ? 2 * w->cap : 1 << 16;
        w->p = (uint8_t *)realloc(w->p, w->cap);
    }
    put32(w->p + w->len, v);
    w->len += 4;
}
static void wb_words(wbuf *w, const uint32_t *v, size_t n) {
    for (size_t i = 0; i < n; i++) wb_u32(w, v[i]);
}
/* auth path of leaf `idx` from a full tree (src/prover.rs:2262-2273) */
static void wb_path(wbuf *w, const tree_t *t, uint64_t idx) {
    wb_u32(w, t->log_n);
    for (uint32_t k = 0; k < t->log_n; k++) {
        wb_words(w, t->nodes + t->off[k] + 8 * (idx ^ 1), 8);
        idx >>= 1;
    }
}
/* write_decommitment_{u32,4} (stark_cli.rs:1061-1089) for data held as
 * `width` SoA columns. */
static void wb_decommit(wbuf *w, const tree_t *t, const uint32_t *const *cols, uint32_t width,
                        const uint64_t *idx, uint32_t nq) {
    wb_u32(w, nq);
    for (uint32_t q = 0; q < nq; q++)
        for (uint32_t c = 0; c < width; c++) wb_u32(w, cols[c][idx[q]]);
    for (uint32_t q = 0; q < nq; q++)
        for (uint32_t c = 0; c < width; c++) wb_u32(w, cols[c][idx[q] ^ 1]);
    for (uint32_t q = 0; q < nq; q++) wb_path(w, t, idx[q]);
    for (uint32_t q = 0; q < nq; q++) wb_path(w, t, idx[q] ^ 1);
}

size_t or_proof_size(uint32_t log_n, uint32_t blowup_bits, uint32_t nq, uint32_t pow_bits) {
    uint32_t le = log_n + blowup_bits;
    uint32_t n_fri = le - 3;
    size_t words = (pow_bits ? 9 : 5) + 16 + 1 + 8 * (size_t)n_fri + 1 + 32 + 1 + nq;
    /* one decommitment: n, values, siblings, 2*nq paths of (1 + 8*depth) */
    words += 1 + 2 * (size_t)nq + 2 * (size_t)nq * (1 + 8 * (size_t)le);           /* trace */
    words += 1 + 8 * (size_t)nq + 2 * (size_t)nq * (1 + 8 * (size_t)le);           /* quotient */
    words += 1;
    for (uint32_t k = 0; k < n_fri; k++) {
        uint32_t L = le - 1 - k;
        words += 1 + 8 * (size_t)nq + 2 * (size_t)nq * (1 + 8 * (size_t)L);
    }
    return 4 * words;
}

/* FRI commit phase and query draws (src/prover.rs:380-557), shared by the
 * Fibonacci prover and the synthetic multi-column AIR below: circle fold of
 * the quotient q (4 SoA columns on the eval domain of size 2^le), line folds
 * down to 8 values, mix_felts(last layer), optional PoW, nq query draws. */
typedef struct {
    uint32_t le, n_fri, nq;
    uint32_t *(*fl)[4];
    tree_t *ftree;
    uint32_t last[32];
    uint64_t nonce;
    uint64_t *qi;
} fri_run_t;

static void fri_commit_and_query(or_channel *ch, uint32_t *const q[4], uint32_t le, uint32_t nq,
                                 uint32_t pow_bits, fri_run_t *f) {
    uint64_t N = 1ull << le;
    coset_t edom = eval_domain(le);
    f->le = le;
    f->nq = nq;
    f->n_fri = le - 3;
    uint32_t n_fri = f->n_fri;
    f->fl = (uint32_t * (*)[4]) calloc(n_fri, sizeof *f->fl);
    f->ftree = (tree_t *)calloc(n_fri, sizeof(tree_t));
    uint32_t *(*fl)[4] = f->fl;
    tree_t *ftree = f->ftree;
    uint32_t *tw = (uint32_t *)malloc(N / 2 * 4);
    {
        uint32_t fa[4], fa2[4], c5[5];
        or_channel_draw_felt(ch, fa);
        or_qm31_mul(fa, fa, fa2);
        coset_pack(edom, c5);
        or_fold_twiddles(c5, 1, tw);
        for (int k = 0; k < 4; k++) fl[0][k] = (uint32_t *)calloc(N / 2, 4);
        or_fold_circle_into_line(fl[0], (const uint32_t *const *)q, tw, fa, fa2, N / 2);
        tree_commit(&ftree[0], (const uint32_t *const *)fl[0], 4, le - 1);
        or_channel_mix_digest(ch, tree_root(&ftree[0]));
    }
    uint32_t cur = le - 1;
    for (uint32_t k = 1; k < n_fri; k++) {
        uint32_t fa[4], c5[5];
        or_channel_draw_felt(ch, fa);
        coset_pack(coset_make(OR_HALF_ODDS, cur), c5);
        or_fold_twiddles(c5, 0, tw);
        uint64_t half = 1ull << (cur - 1);
        for (int c = 0; c < 4; c++) fl[k][c] = (uint32_t *)malloc(half * 4);
        or_fold_line((const uint32_t *const *)fl[k - 1], tw, fl[k], fa, half);
        cur--;
        tree_commit(&ftree[k], (const uint32_t *const *)fl[k], 4, cur);
        or_channel_mix_digest(ch, tree_root(&ftree[k]));
    }
    free(tw);
    for (int i = 0; i < 8; i++)
        for (int c = 0; c < 4; c++) f->last[4 * i + c] = fl[n_fri - 1][c][i];
    or_channel_mix_felts(ch, f->last, 8);
    f->nonce = 0;
    if (pow_bits) {
        f->nonce = or_grind_pow(ch, pow_bits);
        or_channel_mix_u64(ch, f->nonce);
    }
    f->qi = (uint64_t *)malloc((nq + 1) * 8);
    for (uint32_t i = 0; i < nq; i++) f->qi[i] = or_channel_draw_number(ch, N);
}
/* n_fri, FRI roots, last layer, queries (stark_cli.rs:1100-1106) */
static void fri_write_roots_and_queries(wbuf *w, const fri_run_t *f) {
    wb_u32(w, f->n_fri);
    for (uint32_t k = 0; k < f->n_fri; k++) wb_words(w, tree_root(&f->ftree[k]), 8);
    wb_u32(w, 8);
    wb_words(w, f->last, 32);
    wb_u32(w, f->nq);
    for (uint32_t i = 0; i < f->nq; i++) wb_u32(w, (uint32_t)f->qi[i]);
}
/* n_fri, then each FRI layer's decommitment at query >> (k + 1) */
static void fri_write_decommits(wbuf *w, const fri_run_t *f) {
    uint64_t *fi = (uint64_t *)malloc((f->nq + 1) * 8);
    wb_u32(w, f->n_fri);
    for (uint32_t i = 0; i < f->nq; i++) fi[i] = f->qi[i] / 2;
    for (uint32_t k = 0; k < f->n_fri; k++) {
        wb_decommit(w, &f->ftree[k], (const uint32_t *const *)f->fl[k], 4, fi, f->nq);
        for (uint32_t i = 0; i < f->nq; i++) fi[i] /= 2;
    }
    free(fi);
}
static void fri_free(fri_run_t *f) {
    for (uint32_t k = 0; k < f->n_fri; k++) {
        tree_free(&f->ftree[k]);
        for (int c = 0; c < 4; c++) free(f->fl[k][c]);
    }
    free(f->ftree);
    free(f->fl);
    free(f->qi);
}

/* In place: the first 2^log_n words of `col` (trace values on
 * half_coset(log_n)) become the 2^le-point evaluation on the eval domain of
 * the interpolated polynomial (src/prover.rs:280-318, 655-659). */
static void lde_column(uint32_t *col, uint32_t log_n, uint32_t le) {
    uint64_t n = 1ull << log_n, N = 1ull << le;
    {
        uint32_t c5[5];
        coset_pack(coset_make(OR_HALF_COSET, log_n), c5);
        uint32_t *ltw = (uint32_t *)malloc(n * 4), *ctw = (uint32_t *)malloc(n / 2 * 4);
        uint32_t offs[32], sizes[32];
        or_compute_twiddles(c5, ltw, ctw, offs, sizes);
        or_batch_inverse(ltw, ltw, n - 2);
        or_batch_inverse(ctw, ctw, n / 2);
        or_ntt_interpolate(col, ltw, ctw, offs, log_n - 1, n);
        free(ltw);
        free(ctw);
    }
    for (uint64_t i = n; i < N; i++) col[i] = 0; /* zero_pad */
    {
        uint32_t c5[5];
        coset_pack(eval_domain(le), c5);
        uint32_t *ltw = (uint32_t *)malloc(N * 4), *ctw = (uint32_t *)malloc(N / 2 * 4);
        uint32_t offs[32], sizes[32];
        or_compute_twiddles(c5, ltw, ctw, offs, sizes);
        or_ntt_evaluate(col, ltw, ctw, offs, le - 1, N);
        free(ltw);
        free(ctw);
    }
}

/* src/prover.rs:264-619 (prove_lean_inner); prove_lean_fused and
 * prove_with_cache produce the same bytes. */
int or_prove(uint32_t a, uint32_t b, uint32_t log_n, uint32_t blowup_bits, uint32_t nq,
             uint32_t pow_bits, uint8_t **out, size_t *out_len) {
    if (log_n < 4 || blowup_bits < 1 || log_n + blowup_bits > 31 || a >= P || b >= P || nq == 0)
        return -1;
    if (log_n + blowup_bits > 31) return -1;
    uint32_t le = log_n + blowup_bits;
    uint64_t n = 1ull << log_n, N = 1ull << le;

    /* trace, interpolate on half_coset(log_n), zero-pad, evaluate on the
     * eval domain (prover.rs:280-318, 655-659) */
    uint32_t *ev = (uint32_t *)malloc(N * 4);
    or_fibonacci_trace(a, b, log_n, ev);
    lde_column(ev, log_n, le);
    (void)n;
    /* trace commitment (prover.rs:321-324) */
    tree_t ttree;
    const uint32_t *evc[1] = {ev};
    tree_commit(&ttree, evc, 1, le);

    /* quotient (prover.rs:330-352) */
    or_channel ch;
    or_channel_init(&ch);
    or_channel_mix_digest(&ch, tree_root(&ttree));
    uint32_t alpha[4];
    or_channel_draw_felt(&ch, alpha);
    uint32_t *q[4];
    for (int k = 0; k < 4; k++) q[k] = (uint32_t *)malloc(N * 4);
    or_fibonacci_quotient(ev, q[0], q[1], q[2], q[3], alpha, N);
    tree_t qtree;
    tree_commit(&qtree, (const uint32_t *const *)q, 4, le);
    or_channel_mix_digest(&ch, tree_root(&qtree));

    /* FRI + queries (prover.rs:380-557) */
    fri_run_t fr;
    fri_commit_and_query(&ch, q, le, nq, pow_bits, &fr);
    uint64_t nonce = fr.nonce, *qi = fr.qi;

    /* serialize (stark_cli.rs:1092-1120) */
    wbuf w = {0};
    wb_u32(&w, KRAK_MAGIC);
    wb_u32(&w, pow_bits ? 2 : 1);
    wb_u32(&w, log_n);
    wb_u32(&w, a);
    wb_u32(&w, b);
    if (pow_bits) {
        wb_u32(&w, blowup_bits);
        wb_u32(&w, pow_bits);
        wb_u32(&w, (uint32_t)nonce);
        wb_u32(&w, (uint32_t)(nonce >> 32));
    }
    wb_words(&w, tree_root(&ttree), 8);
    wb_words(&w, tree_root(&qtree), 8);
    fri_write_roots_and_queries(&w, &fr);
    wb_decommit(&w, &ttree, evc, 1, qi, nq);
    wb_decommit(&w, &qtree, (const uint32_t *const *)q, 4, qi, nq);
    fri_write_decommits(&w, &fr);

    tree_free(&ttree);
    tree_free(&qtree);
    fri_free(&fr);
    for (int k = 0; k < 4; k++) free(q[k]);
    free(ev);
    *out = w.p;
    *out_len = w.len;
    return 0;
}

void or_free(void *p) { free(p); }

/* ======================================================================
 * Verifier: src/verifier.rs:39-310 over KRAK bytes (stark_cli.rs:1189-1230)
 * ====================================================================== */
typedef struct { const uint8_t *p; size_t len, pos; int bad; } rbuf;
static uint32_t rb_u32(rbuf *r) {
    if (r->pos + 4 > r->len) { r->bad = 1; return 0; }
    uint32_t v = le32(r->p + r->pos);
    r->pos += 4;
    return v;
}
typedef struct {
    uint32_t n, width;
    uint32_t *vals, *sibs;        /* n * width */
    uint32_t *depth, *sdepth;     /* n */
    uint32_t **paths, **spaths;   /* n pointers to depth*8 words */
} decom_t;
static void decom_free(decom_t *d) {
    if (d->paths)
        for (uint32_t i = 0; i < d->n; i++) { free(d->paths[i]); free(d->spaths[i]); }
    free(d->vals); free(d->sibs); free(d->depth); free(d->sdepth); free(d->paths); free(d->spaths);
    memset(d, 0, sizeof *d);
}
static int rb_decom(rbuf *r, decom_t *d, uint32_t width) {
    memset(d, 0, sizeof *d);
    d->width = width;
    d->n = rb_u32(r);
    if (r->bad || d->n > (1u << 20) || (size_t)d->n * width * 8 > r->len) return 0;
    uint32_t n = d->n;
    d->vals = (uint32_t *)calloc((size_t)n * width + 1, 4);
    d->sibs = (uint32_t *)calloc((size_t)n * width + 1, 4);
    d->depth = (uint32_t *)calloc(n + 1, 4);
    d->sdepth = (uint32_t *)calloc(n + 1, 4);
    d->paths = (uint32_t **)calloc(n + 1, sizeof(uint32_t *));
    d->spaths = (uint32_t **)calloc(n + 1, sizeof(uint32_t *));
    for (size_t i = 0; i < (size_t)n * width; i++) d->vals[i] = rb_u32(r);
    for (size_t i = 0; i < (size_t)n * width; i++) d->sibs[i] = rb_u32(r);
    for (int s = 0; s < 2; s++)
        for (uint32_t i = 0; i < n; i++) {
            uint32_t dep = rb_u32(r);
            if (r->bad || dep > 64) return 0;
            uint32_t *pth = (uint32_t *)calloc((size_t)dep * 8 + 1, 4);
            for (uint32_t k = 0; k < 8 * dep; k++) pth[k] = rb_u32(r);
            if (s == 0) { d->depth[i] = dep; d->paths[i] = pth; }
            else { d->sdepth[i] = dep; d->spaths[i] = pth; }
        }
    return !r->bad;
}
/* verifier.rs:274-310 */
static int check_decom(const decom_t *d, const uint32_t root[8], const uint64_t *idx, uint32_t nq) {
    if (d->n != nq) return 0;
    for (uint32_t q = 0; q < nq; q++) {
        uint32_t h[8];
        hash_words(d->vals + (size_t)q * d->width, d->width, h);
        if (!or_verify_auth_path(root, h, idx[q], d->paths[q], d->depth[q])) return 0;
        hash_words(d->sibs + (size_t)q * d->width, d->width, h);
        if (!or_verify_auth_path(root, h, idx[q] ^ 1, d->spaths[q], d->sdepth[q])) return 0;
    }
    return 1;
}
/* verifier.rs:30-36 */
static uint32_t fold_twiddle_at(coset_t dom, uint64_t folded, int circle) {
    cpt p = coset_at(dom, bitrev(folded << 1, dom.log_size));
    return m31_inv(circle ? p.y : p.x);
}
static qm31 pair_fold(const decom_t *d, uint32_t q, uint64_t idx, qm31 alpha, uint32_t tw) {
    qm31 v = q_load(d->vals + 4 * (size_t)q), s = q_load(d->sibs + 4 * (size_t)q);
    return idx % 2 == 0 ? fold_pair(v, s, alpha, tw) : fold_pair(s, v, alpha, tw);
}

#define VFAIL(...) do { snprintf(err, err_cap, __VA_ARGS__); goto done; } while (0)

int or_verify(const uint8_t *bytes, size_t len, uint32_t blowup_bits, uint32_t nq,
              char *err, size_t err_cap) {
    int ok = 0;
    rbuf r = {bytes, len, 0, 0};
    decom_t td = {0}, qd = {0}, *fd = NULL;
    uint32_t n_fd = 0, n_fri = 0, n_last = 0, n_q = 0;
    uint32_t (*froots)[8] = NULL, *lastw = NULL;
    uint64_t *qidx = NULL, *fidx = NULL;
    qm31 *alphas = NULL;
    char dummy[8];
    if (!err || err_cap == 0) { err = dummy; err_cap = sizeof dummy; }
    err[0] = 0;

    if (rb_u32(&r) != KRAK_MAGIC) VFAIL("bad magic");
    uint32_t version = rb_u32(&r);
    if (version != 1 && version != 2) VFAIL("unsupported proof version %u", version);
    uint32_t log_n = rb_u32(&r), a = rb_u32(&r), b = rb_u32(&r);
    uint32_t pow_bits = 0;
    uint64_t nonce = 0;
    if (version == 2) {
        uint32_t bb = rb_u32(&r);
        pow_bits = rb_u32(&r);
        nonce = rb_u32(&r);
        nonce |= (uint64_t)rb_u32(&r) << 32;
        if (bb != blowup_bits) VFAIL("blowup mismatch");
        if (pow_bits > 128) VFAIL("bad pow bits");
    }
    uint32_t troot[8], qroot[8];
    for (int i = 0; i < 8; i++) troot[i] = rb_u32(&r);
    for (int i = 0; i < 8; i++) qroot[i] = rb_u32(&r);
    n_fri = rb_u32(&r);
    if (r.bad || n_fri > 64) VFAIL("truncated proof");
    froots = (uint32_t(*)[8])calloc(n_fri + 1, 32);
    for (uint32_t k = 0; k < n_fri; k++)
        for (int i = 0; i < 8; i++) froots[k][i] = rb_u32(&r);
    n_last = rb_u32(&r);
    if (r.bad || n_last > 4096) VFAIL("truncated proof");
    lastw = (uint32_t *)calloc(4 * (size_t)n_last + 1, 4);
    for (size_t i = 0; i < 4 * (size_t)n_last; i++) lastw[i] = rb_u32(&r);
    n_q = rb_u32(&r);
    if (r.bad || n_q > (1u << 20)) VFAIL("truncated proof");
    qidx = (uint64_t *)calloc(n_q + 1, 8);
    for (uint32_t i = 0; i < n_q; i++) qidx[i] = rb_u32(&r);
    if (!rb_decom(&r, &td, 1)) VFAIL("truncated proof");
    if (!rb_decom(&r, &qd, 4)) VFAIL("truncated proof");
    n_fd = rb_u32(&r);
    if (r.bad || n_fd > 64) VFAIL("truncated proof");
    fd = (decom_t *)calloc(n_fd + 1, sizeof(decom_t));
    for (uint32_t k = 0; k < n_fd; k++)
        if (!rb_decom(&r, &fd[k], 4)) VFAIL("truncated proof");

    uint32_t le = log_n + blowup_bits;
    if (log_n < 1 || le > 31) VFAIL("bad log size");
    uint64_t N = 1ull << le;
    /* verifier.rs:46-53 */
    if (a >= P || b >= P) VFAIL("Public inputs out of M31 range");
    if (n_q != nq) VFAIL("Expected %u queries, got %u", nq, n_q);
    /* verifier.rs:56-70 */
    or_channel ch;
    or_channel_init(&ch);
    or_channel_mix_digest(&ch, troot);
    uint32_t tmp[4];
    or_channel_draw_felt(&ch, tmp);
    or_channel_mix_digest(&ch, qroot);
    alphas = (qm31 *)calloc(n_fri + 2, sizeof(qm31));
    or_channel_draw_felt(&ch, alphas[0].v);
    for (uint32_t k = 0; k < n_fri; k++) {
        or_channel_mix_digest(&ch, froots[k]);
        or_channel_draw_felt(&ch, alphas[k + 1].v);
    }
    /* verifier.rs:73-87 */
    if (n_fri != (le >= 3 ? le - 3 : 0)) VFAIL("Expected %u FRI layers, got %u", le - 3, n_fri);
    if (n_last != 8) VFAIL("Expected 8 FRI last layer values, got %u", n_last);
    /* verifier.rs:90-96 (+ KRAK v2 PoW check before the query draws) */
    or_channel_mix_felts(&ch, lastw, n_last);
    if (version == 2) {
        if (!or_channel_verify_pow_nonce(&ch, pow_bits, nonce)) VFAIL("PoW nonce invalid");
        or_channel_mix_u64(&ch, nonce);
    }
    for (uint32_t i = 0; i < nq; i++)
        if (or_channel_draw_number(&ch, N) != qidx[i])
            VFAIL("Query indices don't match Fiat-Shamir derivation");
    /* verifier.rs:99-104 */
    {
        int tz = 1, qz = 1;
        for (int i = 0; i < 8; i++) { tz &= troot[i] == 0; qz &= qroot[i] == 0; }
        if (tz) VFAIL("Trace commitment is zero");
        if (qz) VFAIL("Quotient commitment is zero");
    }
    /* verifier.rs:107-144 */
    if (!check_decom(&td, troot, qidx, nq)) VFAIL("trace auth path invalid");
    if (!check_decom(&qd, qroot, qidx, nq)) VFAIL("quotient auth path invalid");
    fidx = (uint64_t *)calloc(nq + 1, 8);
    for (uint32_t i = 0; i < nq; i++) fidx[i] = qidx[i] / 2;
    for (uint32_t k = 0; k < n_fd && k < n_fri; k++) {
        if (!check_decom(&fd[k], froots[k], fidx, nq)) VFAIL("FRI layer %u auth path invalid", k);
        for (uint32_t i = 0; i < nq; i++) fidx[i] /= 2;
    }
    /* verifier.rs:153-233: fold equations */
    if (n_fd == 0) VFAIL("missing FRI decommitments");
    for (uint32_t k = 0; k < n_fd; k++)
        if (fd[k].n != nq) VFAIL("FRI layer %u decommitment size mismatch", k);
    for (uint32_t q = 0; q < nq; q++) {
        uint64_t ci = qidx[q];
        uint32_t cl = le;
        coset_t dom = eval_domain(cl);
        qm31 f0 = q_load(qd.vals + 4 * (size_t)q), f1 = q_load(qd.sibs + 4 * (size_t)q);
        if (ci % 2) { qm31 t = f0; f0 = f1; f1 = t; }
        qm31 e = fold_pair(f0, f1, alphas[0], fold_twiddle_at(dom, ci / 2, 1));
        if (memcmp(e.v, fd[0].vals + 4 * (size_t)q, 16)) VFAIL("Circle fold mismatch at query %u", q);
        ci /= 2;
        cl--;
        for (uint32_t k = 0; k + 1 < n_fd; k++) {
            coset_t ld = coset_make(OR_HALF_ODDS, cl);
            qm31 ex = pair_fold(&fd[k], q, ci, alphas[k + 1], fold_twiddle_at(ld, ci / 2, 0));
            if (memcmp(ex.v, fd[k + 1].vals + 4 * (size_t)q, 16))
                VFAIL("Line fold mismatch at query %u, layer %u", q, k);
            ci /= 2;
            cl--;
        }
        if (ci >= n_last) VFAIL("FRI last layer index out of bounds");
        if (memcmp(fd[n_fd - 1].vals + 4 * (size_t)q, lastw + 4 * ci, 16))
            VFAIL("FRI last layer mismatch at query %u", q);
    }
    ok = 1;
done:
    decom_free(&td);
    decom_free(&qd);
    if (fd)
        for (uint32_t k = 0; k < n_fd; k++) decom_free(&fd[k]);
    free(fd);
    free(froots);
    free(lastw);
    free(qidx);
    free(fidx);
    free(alphas);
    return ok;
}

/* ======================================================================
 * Synthetic degree-2 multi-column AIR ("d2"): the serving workload's
 * 34-column request. No reference counterpart; its multi-column commitment
 * follows the reference's Cairo trace (columns in Merkle trees of <= 16
 * columns, src/cairo_air/prover.rs:341-350, because the leaf hash takes at
 * most 16 words, cuda/blake2s.cu:48-53) and its verifier checks the
 * constraints at the query rows from values at q and q + 1, as the Cairo
 * verifier does (trace_values_at_queries[_next], prover.rs:415-419).
 *
 *   trace     t_c[i] = mix64(seed + G * ((c << 32 | i) + 1)) mod P,
 *             G = 0x9E3779B97F4A7C15, mix64 = SplitMix64 finaliser
 *   AIR       constraint k < K:  C_k(i) = c_k t_a[i] t_b[i+1] + d_k t_e[i],
 *             a = k mod W, b = (5k + 1) mod W, e = (11k + 7) mod W,
 *             (c_0, d_0, c_1, d_1, ...) = SplitMix64(air_seed) outputs mod P
 *   quotient  Q(i) = sum_k alpha^(k+1) C_k(i) on the LDE (row i + 1 taken
 *             mod N in eval order, as the Fibonacci quotient does)
 *   transcript  mix the G = ceil(W / 16) trace roots, draw alpha, commit Q,
 *             then exactly the Fibonacci FRI / PoW / query phase
 *   bytes     KRAK v3: magic, 3, log_n, W, K, blowup_bits, pow_bits,
 *             nonce lo/hi, seed lo/hi, air_seed lo/hi, G, G roots, Q root,
 *             FRI roots, last layer, queries, per group: n, values at q,
 *             values at q + 1, paths at q, paths at q + 1; then the
 *             quotient and FRI decommitments as in KRAK v1.
 * ====================================================================== */
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
#define D2_GOLDEN 0x9E3779B97F4A7C15ull

uint32_t or_d2_value(uint64_t seed, uint32_t col, uint64_t row) {
    uint64_t key = ((uint64_t)col << 32) | row;
    return (uint32_t)(mix64(seed + D2_GOLDEN * (key + 1)) % P);
}
void or_d2_trace(uint64_t seed, uint32_t n_cols, uint32_t log_n, uint32_t *out) {
    uint64_t n = 1ull << log_n;
    for (uint32_t c = 0; c < n_cols; c++)
        for (uint64_t i = 0; i < n; i++) out[c * n + i] = or_d2_value(seed, c, i);
}
void or_d2_coeffs(uint64_t air_seed, uint32_t n_cons, uint32_t *cd) {
    uint64_t s = air_seed;
    for (uint32_t k = 0; k < 2 * n_cons; k++) {
        s += D2_GOLDEN;
        cd[k] = (uint32_t)(mix64(s) % P);
    }
}
static inline uint32_t d2_a(uint32_t k, uint32_t w) { return k % w; }
static inline uint32_t d2_b(uint32_t k, uint32_t w) { return (5 * k + 1) % w; }
static inline uint32_t d2_e(uint32_t k, uint32_t w) { return (11 * k + 7) % w; }

/* Q at row i of the evaluation (cols: W columns of N values). */
static qm31 d2_q_at(const uint32_t *const *cols, uint32_t w, uint32_t k_cons, const uint32_t *cd,
                    const qm31 *beta, uint64_t i, uint64_t mask) {
    qm31 acc = {{0, 0, 0, 0}};
    uint64_t j = (i + 1) & mask;
    for (uint32_t k = 0; k < k_cons; k++) {
        uint32_t m = m31_add(m31_mul(cd[2 * k], m31_mul(cols[d2_a(k, w)][i], cols[d2_b(k, w)][j])),
                             m31_mul(cd[2 * k + 1], cols[d2_e(k, w)][i]));
        acc = q_add(acc, q_scale(beta[k], m));
    }
    return acc;
}
/* Same from the values of one row and of the next (verifier side). */
static qm31 d2_q_rows(const uint32_t *row, const uint32_t *next, uint32_t w, uint32_t k_cons,
                      const uint32_t *cd, const qm31 *beta) {
    qm31 acc = {{0, 0, 0, 0}};
    for (uint32_t k = 0; k < k_cons; k++) {
        uint32_t m = m31_add(m31_mul(cd[2 * k], m31_mul(row[d2_a(k, w)], next[d2_b(k, w)])),
                             m31_mul(cd[2 * k + 1], row[d2_e(k, w)]));
        acc = q_add(acc, q_scale(beta[k], m));
    }
    return acc;
}
static void d2_powers(const uint32_t alpha[4], uint32_t k_cons, qm31 *beta) {
    qm31 a = q_load(alpha), p = a;
    for (uint32_t k = 0; k < k_cons; k++) {
        beta[k] = p;
        p = q_mul(p, a);
    }
}
void or_d2_quotient(const uint32_t *eval, uint32_t n_cols, uint32_t le, uint32_t n_cons, uint64_t air_seed,
                    const uint32_t alpha[4], uint32_t *q0, uint32_t *q1, uint32_t *q2, uint32_t *q3) {
    uint64_t N = 1ull << le;
    const uint32_t **cols = (const uint32_t **)malloc(n_cols * sizeof(uint32_t *));
    for (uint32_t c = 0; c < n_cols; c++) cols[c] = eval + c * N;
    uint32_t *cd = (uint32_t *)malloc(8 * n_cons + 8);
    or_d2_coeffs(air_seed, n_cons, cd);
    qm31 *beta = (qm31 *)malloc(sizeof(qm31) * (n_cons + 1));
    d2_powers(alpha, n_cons, beta);
    uint32_t *o[4] = {q0, q1, q2, q3};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)N; i++) {
        qm31 v = d2_q_at(cols, n_cols, n_cons, cd, beta, (uint64_t)i, N - 1);
        for (int c = 0; c < 4; c++) o[c][i] = v.v[c];
    }
    free(cols);
    free(cd);
    free(beta);
}

static uint32_t d2_groups(uint32_t w) { return (w + 15) / 16; }
static uint32_t d2_group_cols(uint32_t w, uint32_t g) { return w - 16 * g < 16 ? w - 16 * g : 16; }
#define KRAK_D2_HEADER 13

size_t or_d2_proof_size(uint32_t log_n, uint32_t n_cols, uint32_t blowup_bits, uint32_t nq) {
    uint32_t le = log_n + blowup_bits, n_fri = le - 3, G = d2_groups(n_cols);
    size_t words = KRAK_D2_HEADER + 1 + 8 * (size_t)G + 8 + 1 + 8 * (size_t)n_fri + 1 + 32 + 1 + nq;
    for (uint32_t g = 0; g < G; g++)
        words += 1 + 2 * (size_t)nq * d2_group_cols(n_cols, g) + 2 * (size_t)nq * (1 + 8 * (size_t)le);
    words += 1 + 8 * (size_t)nq + 2 * (size_t)nq * (1 + 8 * (size_t)le);
    words += 1;
    for (uint32_t k = 0; k < n_fri; k++) {
        uint32_t L = le - 1 - k;
        words += 1 + 8 * (size_t)nq + 2 * (size_t)nq * (1 + 8 * (size_t)L);
    }
    return 4 * words;
}

int or_d2_prove(uint64_t seed, uint32_t log_n, uint32_t n_cols, uint32_t n_cons, uint64_t air_seed,
                uint32_t blowup_bits, uint32_t nq, uint32_t pow_bits, uint8_t **out, size_t *out_len) {
    if (log_n < 4 || blowup_bits < 1 || log_n + blowup_bits > 31 || nq == 0 || n_cols < 1 || n_cols > 256 ||
        n_cons < 1 || n_cons > 1024)
        return -1;
    uint32_t le = log_n + blowup_bits, G = d2_groups(n_cols);
    uint64_t n = 1ull << log_n, N = 1ull << le;
    uint32_t *ev = (uint32_t *)malloc((size_t)n_cols * N * 4);
    for (uint32_t c = 0; c < n_cols; c++) {
        uint32_t *col = ev + c * N;
        for (uint64_t i = 0; i < n; i++) col[i] = or_d2_value(seed, c, i);
        lde_column(col, log_n, le);
    }
    /* trace trees, one per group of <= 16 columns */
    tree_t *tt = (tree_t *)calloc(G, sizeof(tree_t));
    const uint32_t **cols = (const uint32_t **)malloc(n_cols * sizeof(uint32_t *));
    for (uint32_t c = 0; c < n_cols; c++) cols[c] = ev + c * N;
    or_channel ch;
    or_channel_init(&ch);
    for (uint32_t g = 0; g < G; g++) {
        tree_commit(&tt[g], cols + 16 * g, d2_group_cols(n_cols, g), le);
        or_channel_mix_digest(&ch, tree_root(&tt[g]));
    }
    uint32_t alpha[4];
    or_channel_draw_felt(&ch, alpha);
    uint32_t *q[4];
    for (int k = 0; k < 4; k++) q[k] = (uint32_t *)malloc(N * 4);
    or_d2_quotient(ev, n_cols, le, n_cons, air_seed, alpha, q[0], q[1], q[2], q[3]);
    tree_t qtree;
    tree_commit(&qtree, (const uint32_t *const *)q, 4, le);
    or_channel_mix_digest(&ch, tree_root(&qtree));
    fri_run_t fr;
    fri_commit_and_query(&ch, q, le, nq, pow_bits, &fr);

    wbuf w = {0};
    wb_u32(&w, KRAK_MAGIC);
    wb_u32(&w, 3);
    wb_u32(&w, log_n);
    wb_u32(&w, n_cols);
    wb_u32(&w, n_cons);
    wb_u32(&w, blowup_bits);
    wb_u32(&w, pow_bits);
    wb_u32(&w, (uint32_t)fr.nonce);
    wb_u32(&w, (uint32_t)(fr.nonce >> 32));
    wb_u32(&w, (uint32_t)seed);
    wb_u32(&w, (uint32_t)(seed >> 32));
    wb_u32(&w, (uint32_t)air_seed);
    wb_u32(&w, (uint32_t)(air_seed >> 32));
    wb_u32(&w, G);
    for (uint32_t g = 0; g < G; g++) wb_words(&w, tree_root(&tt[g]), 8);
    wb_words(&w, tree_root(&qtree), 8);
    fri_write_roots_and_queries(&w, &fr);
    uint64_t *nx = (uint64_t *)malloc((nq + 1) * 8);
    for (uint32_t i = 0; i < nq; i++) nx[i] = (fr.qi[i] + 1) & (N - 1);
    for (uint32_t g = 0; g < G; g++) {
        uint32_t wg = d2_group_cols(n_cols, g);
        const uint32_t *const *gc = cols + 16 * g;
        wb_u32(&w, nq);
        for (uint32_t i = 0; i < nq; i++)
            for (uint32_t c = 0; c < wg; c++) wb_u32(&w, gc[c][fr.qi[i]]);
        for (uint32_t i = 0; i < nq; i++)
            for (uint32_t c = 0; c < wg; c++) wb_u32(&w, gc[c][nx[i]]);
        for (uint32_t i = 0; i < nq; i++) wb_path(&w, &tt[g], fr.qi[i]);
        for (uint32_t i = 0; i < nq; i++) wb_path(&w, &tt[g], nx[i]);
    }
    wb_decommit(&w, &qtree, (const uint32_t *const *)q, 4, fr.qi, nq);
    fri_write_decommits(&w, &fr);

    for (uint32_t g = 0; g < G; g++) tree_free(&tt[g]);
    free(tt);
    tree_free(&qtree);
    fri_free(&fr);
    for (int k = 0; k < 4; k++) free(q[k]);
    free(cols);
    free(ev);
    free(nx);
    *out = w.p;
    *out_len = w.len;
    return 0;
}

/* Trace decommitment of one group: values at q and at q + 1, auth paths of both. */
typedef struct {
    uint32_t n, width;
    uint32_t *vals, *next;
    uint32_t **paths, **npaths;
    uint32_t *depth, *ndepth;
} d2dec_t;
static void d2dec_free(d2dec_t *d) {
    if (d->paths)
        for (uint32_t i = 0; i < d->n; i++) { free(d->paths[i]); free(d->npaths[i]); }
    free(d->vals); free(d->next); free(d->paths); free(d->npaths); free(d->depth); free(d->ndepth);
    memset(d, 0, sizeof *d);
}
static int rb_d2dec(rbuf *r, d2dec_t *d, uint32_t width) {
    memset(d, 0, sizeof *d);
    d->width = width;
    d->n = rb_u32(r);
    if (r->bad || d->n > (1u << 20) || (size_t)d->n * width * 8 > r->len) return 0;
    uint32_t n = d->n;
    d->vals = (uint32_t *)calloc((size_t)n * width + 1, 4);
    d->next = (uint32_t *)calloc((size_t)n * width + 1, 4);
    d->paths = (uint32_t **)calloc(n + 1, sizeof(uint32_t *));
    d->npaths = (uint32_t **)calloc(n + 1, sizeof(uint32_t *));
    d->depth = (uint32_t *)calloc(n + 1, 4);
    d->ndepth = (uint32_t *)calloc(n + 1, 4);
    for (size_t i = 0; i < (size_t)n * width; i++) d->vals[i] = rb_u32(r);
    for (size_t i = 0; i < (size_t)n * width; i++) d->next[i] = rb_u32(r);
    for (int s = 0; s < 2; s++)
        for (uint32_t i = 0; i < n; i++) {
            uint32_t dep = rb_u32(r);
            if (r->bad || dep > 64) return 0;
            uint32_t *pth = (uint32_t *)calloc((size_t)dep * 8 + 1, 4);
            for (uint32_t k = 0; k < 8 * dep; k++) pth[k] = rb_u32(r);
            if (s == 0) { d->depth[i] = dep; d->paths[i] = pth; }
            else { d->ndepth[i] = dep; d->npaths[i] = pth; }
        }
    return !r->bad;
}

int or_d2_verify(const uint8_t *bytes, size_t len, uint32_t blowup_bits, uint32_t nq, char *err, size_t err_cap) {
    int ok = 0;
    rbuf r = {bytes, len, 0, 0};
    decom_t qd = {0}, *fd = NULL;
    d2dec_t *td = NULL;
    uint32_t G = 0, n_fd = 0, n_fri = 0, n_last = 0, n_q = 0;
    uint32_t (*troots)[8] = NULL, (*froots)[8] = NULL, *lastw = NULL, *cd = NULL, *row = NULL, *nrow = NULL;
    uint64_t *qidx = NULL, *fidx = NULL;
    qm31 *alphas = NULL, *beta = NULL;
    char dummy[8];
    if (!err || err_cap == 0) { err = dummy; err_cap = sizeof dummy; }
    err[0] = 0;

    if (rb_u32(&r) != KRAK_MAGIC) VFAIL("bad magic");
    if (rb_u32(&r) != 3) VFAIL("unsupported proof version");
    uint32_t log_n = rb_u32(&r), W = rb_u32(&r), K = rb_u32(&r), bb = rb_u32(&r), pow_bits = rb_u32(&r);
    uint64_t nonce = rb_u32(&r);
    nonce |= (uint64_t)rb_u32(&r) << 32;
    uint64_t seed = rb_u32(&r);
    seed |= (uint64_t)rb_u32(&r) << 32;
    uint64_t air_seed = rb_u32(&r);
    air_seed |= (uint64_t)rb_u32(&r) << 32;
    (void)seed;
    if (r.bad) VFAIL("truncated proof");
    if (bb != blowup_bits) VFAIL("blowup mismatch");
    if (W < 1 || W > 256 || K < 1 || K > 1024) VFAIL("bad AIR shape");
    if (pow_bits > 128) VFAIL("bad pow bits");
    G = rb_u32(&r);
    if (r.bad || G != d2_groups(W)) VFAIL("bad trace group count");
    troots = (uint32_t(*)[8])calloc(G, 32);
    for (uint32_t g = 0; g < G; g++)
        for (int i = 0; i < 8; i++) troots[g][i] = rb_u32(&r);
    uint32_t qroot[8];
    for (int i = 0; i < 8; i++) qroot[i] = rb_u32(&r);
    n_fri = rb_u32(&r);
    if (r.bad || n_fri > 64) VFAIL("truncated proof");
    froots = (uint32_t(*)[8])calloc(n_fri + 1, 32);
    for (uint32_t k = 0; k < n_fri; k++)
        for (int i = 0; i < 8; i++) froots[k][i] = rb_u32(&r);
    n_last = rb_u32(&r);
    if (r.bad || n_last > 4096) VFAIL("truncated proof");
    lastw = (uint32_t *)calloc(4 * (size_t)n_last + 1, 4);
    for (size_t i = 0; i < 4 * (size_t)n_last; i++) lastw[i] = rb_u32(&r);
    n_q = rb_u32(&r);
    if (r.bad || n_q > (1u << 20)) VFAIL("truncated proof");
    qidx = (uint64_t *)calloc(n_q + 1, 8);
    for (uint32_t i = 0; i < n_q; i++) qidx[i] = rb_u32(&r);
    td = (d2dec_t *)calloc(G, sizeof(d2dec_t));
    for (uint32_t g = 0; g < G; g++)
        if (!rb_d2dec(&r, &td[g], d2_group_cols(W, g))) VFAIL("truncated proof");
    if (!rb_decom(&r, &qd, 4)) VFAIL("truncated proof");
    n_fd = rb_u32(&r);
    if (r.bad || n_fd > 64) VFAIL("truncated proof");
    fd = (decom_t *)calloc(n_fd + 1, sizeof(decom_t));
    for (uint32_t k = 0; k < n_fd; k++)
        if (!rb_decom(&r, &fd[k], 4)) VFAIL("truncated proof");
    if (r.pos != r.len) VFAIL("trailing bytes");

    uint32_t le = log_n + blowup_bits;
    if (log_n < 1 || le > 31) VFAIL("bad log size");
    uint64_t N = 1ull << le;
    if (n_q != nq) VFAIL("Expected %u queries, got %u", nq, n_q);
    or_channel ch;
    or_channel_init(&ch);
    for (uint32_t g = 0; g < G; g++) or_channel_mix_digest(&ch, troots[g]);
    uint32_t alpha[4];
    or_channel_draw_felt(&ch, alpha);
    or_channel_mix_digest(&ch, qroot);
    alphas = (qm31 *)calloc(n_fri + 2, sizeof(qm31));
    or_channel_draw_felt(&ch, alphas[0].v);
    for (uint32_t k = 0; k < n_fri; k++) {
        or_channel_mix_digest(&ch, froots[k]);
        or_channel_draw_felt(&ch, alphas[k + 1].v);
    }
    if (n_fri != (le >= 3 ? le - 3 : 0)) VFAIL("Expected %u FRI layers, got %u", le - 3, n_fri);
    if (n_last != 8) VFAIL("Expected 8 FRI last layer values, got %u", n_last);
    or_channel_mix_felts(&ch, lastw, n_last);
    if (pow_bits) {
        if (!or_channel_verify_pow_nonce(&ch, pow_bits, nonce)) VFAIL("PoW nonce invalid");
        or_channel_mix_u64(&ch, nonce);
    } else if (nonce != 0) {
        VFAIL("nonce without PoW");
    }
    for (uint32_t i = 0; i < nq; i++)
        if (or_channel_draw_number(&ch, N) != qidx[i]) VFAIL("Query indices don't match Fiat-Shamir derivation");
    /* trace groups: auth paths at q and q + 1 */
    for (uint32_t g = 0; g < G; g++) {
        const d2dec_t *d = &td[g];
        if (d->n != nq) VFAIL("trace group %u decommitment size mismatch", g);
        for (uint32_t q = 0; q < nq; q++) {
            uint32_t h[8];
            uint64_t nx = (qidx[q] + 1) & (N - 1);
            hash_words(d->vals + (size_t)q * d->width, d->width, h);
            if (d->depth[q] != le || !or_verify_auth_path(troots[g], h, qidx[q], d->paths[q], d->depth[q]))
                VFAIL("trace group %u auth path invalid", g);
            hash_words(d->next + (size_t)q * d->width, d->width, h);
            if (d->ndepth[q] != le || !or_verify_auth_path(troots[g], h, nx, d->npaths[q], d->ndepth[q]))
                VFAIL("trace group %u next-row auth path invalid", g);
        }
    }
    if (!check_decom(&qd, qroot, qidx, nq)) VFAIL("quotient auth path invalid");
    /* the AIR: Q(q) from the committed rows q and q + 1 */
    cd = (uint32_t *)malloc(8 * (size_t)K + 8);
    or_d2_coeffs(air_seed, K, cd);
    beta = (qm31 *)malloc(sizeof(qm31) * (K + 1));
    d2_powers(alpha, K, beta);
    row = (uint32_t *)malloc(4 * (size_t)W);
    nrow = (uint32_t *)malloc(4 * (size_t)W);
    for (uint32_t q = 0; q < nq; q++) {
        for (uint32_t g = 0; g < G; g++)
            for (uint32_t c = 0; c < td[g].width; c++) {
                row[16 * g + c] = td[g].vals[(size_t)q * td[g].width + c];
                nrow[16 * g + c] = td[g].next[(size_t)q * td[g].width + c];
            }
        qm31 e = d2_q_rows(row, nrow, W, K, cd, beta);
        if (memcmp(e.v, qd.vals + 4 * (size_t)q, 16)) VFAIL("Constraint quotient mismatch at query %u", q);
    }
    fidx = (uint64_t *)calloc(nq + 1, 8);
    for (uint32_t i = 0; i < nq; i++) fidx[i] = qidx[i] / 2;
    for (uint32_t k = 0; k < n_fd && k < n_fri; k++) {
        if (!check_decom(&fd[k], froots[k], fidx, nq)) VFAIL("FRI layer %u auth path invalid", k);
        for (uint32_t i = 0; i < nq; i++) fidx[i] /= 2;
    }
    if (n_fd != n_fri || n_fd == 0) VFAIL("FRI decommitment count mismatch");
    for (uint32_t q = 0; q < nq; q++) {
        uint64_t ci = qidx[q];
        uint32_t cl = le;
        qm31 f0 = q_load(qd.vals + 4 * (size_t)q), f1 = q_load(qd.sibs + 4 * (size_t)q);
        if (ci % 2) { qm31 t = f0; f0 = f1; f1 = t; }
        qm31 e = fold_pair(f0, f1, alphas[0], fold_twiddle_at(eval_domain(cl), ci / 2, 1));
        if (memcmp(e.v, fd[0].vals + 4 * (size_t)q, 16)) VFAIL("Circle fold mismatch at query %u", q);
        ci /= 2;
        cl--;
        for (uint32_t k = 0; k + 1 < n_fd; k++) {
            qm31 ex = pair_fold(&fd[k], q, ci, alphas[k + 1], fold_twiddle_at(coset_make(OR_HALF_ODDS, cl), ci / 2, 0));
            if (memcmp(ex.v, fd[k + 1].vals + 4 * (size_t)q, 16)) VFAIL("Line fold mismatch at query %u, layer %u", q, k);
            ci /= 2;
            cl--;
        }
        if (ci >= n_last) VFAIL("FRI last layer index out of bounds");
        if (memcmp(fd[n_fd - 1].vals + 4 * (size_t)q, lastw + 4 * ci, 16)) VFAIL("FRI last layer mismatch at query %u", q);
    }
    ok = 1;
done:
    if (td)
        for (uint32_t g = 0; g < G; g++) d2dec_free(&td[g]);
    free(td);
    decom_free(&qd);
    if (fd)
        for (uint32_t k = 0; k < n_fd; k++) decom_free(&fd[k]);
    free(fd);
    free(troots);
    free(froots);
    free(lastw);
    free(qidx);
    free(fidx);
    free(alphas);
    free(beta);
    free(cd);
    free(row);
    free(nrow);
    return ok;
}

/* ======================================================================
 * Batched uneven columns: per-column LDE (the Fibonacci prover's
 * interpolate + zero-pad + evaluate, src/prover.rs:280-318) and the lifted
 * Merkle tree of the reference's stwo backend (cuda/merkle_leaves_lifted.cu:
 * 26-98, columns sorted by ascending size, src/../merkle_ops.rs schedule).
 * ====================================================================== */
void or_lde(const uint32_t *in, uint32_t log_n, uint32_t blowup_bits, uint32_t *out) {
    uint64_t n = 1ull << log_n;
    memcpy(out, in, n * 4);
    lde_column(out, log_n, log_n + blowup_bits);
}

static uint64_t lifted_row(uint64_t idx, uint32_t S, uint32_t L) {
    return L == S ? idx : (((idx >> (S - L + 1)) << 1) | (idx & 1));
}
/* leaf i = Blake2s(all columns' words at their lifted rows), multi-block */
void or_build_leaves_lifted(const uint32_t *const *cols, const uint32_t *log_sizes, uint32_t n_cols,
                            uint32_t lifting_log, uint32_t *hashes) {
    uint64_t n = 1ull << lifting_log;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        uint8_t *buf = (uint8_t *)malloc(4 * (size_t)n_cols);
        for (uint32_t c = 0; c < n_cols; c++)
            put32(buf + 4 * c, cols[c][lifted_row((uint64_t)i, lifting_log, log_sizes[c])]);
        uint8_t h[32];
        or_blake2s(buf, 4 * (size_t)n_cols, h);
        for (int k = 0; k < 8; k++) hashes[8 * (uint64_t)i + k] = le32(h + 4 * k);
        free(buf);
    }
}
void or_merkle_root_lifted(const uint32_t *const *cols, const uint32_t *log_sizes, uint32_t n_cols,
                           uint32_t root[8]) {
    uint32_t S = log_sizes[n_cols - 1];
    uint64_t n = 1ull << S;
    uint32_t *leaves = (uint32_t *)malloc(n * 32);
    or_build_leaves_lifted(cols, log_sizes, n_cols, S, leaves);
    tree_t t;
    tree_build_from_leaves(&t, S, leaves);
    memcpy(root, tree_root(&t), 32);
    tree_free(&t);
}
/* column `col` of the synthetic generator (or_d2_value) */
void or_d2_column(uint64_t seed, uint32_t col, uint32_t log_n, uint32_t *out) {
    uint64_t n = 1ull << log_n;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) out[i] = or_d2_value(seed, col, (uint64_t)i);
}

/* ======================================================================
 * Batched-columns proof (KRAK v4): the batched uneven-commit workload as a
 * full proof. No reference counterpart as a whole; its pieces are the
 * reference's: per-column LDE (src/prover.rs:280-318), bit reverse
 * (cuda/circle_ntt.cu:234-244), the lifted tree (cuda/merkle_leaves_lifted.cu),
 * and the Fibonacci quotient-commitment / FRI / query phase.
 *   columns   c < C, log sizes L_c sorted ascending, t_c[i] = or_d2_value(seed, c, i)
 *   trace     E_c = L_c + blowup, bit-reversed LDE; lifted tree of 2^S leaves, S = max E_c
 *   quotient  Q(i) = sum_c alpha^(c+1) col_c[lifted_row(i, S, E_c)], alpha drawn
 *             after mixing Blake2s(log sizes) and the lifted root; then
 *             exactly the Fibonacci FRI
 *   bytes     magic, 4, C, blowup, pow, nonce lo/hi, seed lo/hi, C log sizes,
 *             lifted root, Q root, FRI roots, last layer, queries, trace
 *             decommitment (n, C values at q, C values at q ^ 1, paths at q,
 *             paths at q ^ 1), quotient and FRI decommitments as in KRAK v1.
 * ====================================================================== */
#define KRAK_BATCH_HEADER 9

size_t or_batch_proof_size(const uint32_t *log_sizes, uint32_t n_cols, uint32_t blowup_bits, uint32_t nq) {
    uint32_t S = log_sizes[n_cols - 1] + blowup_bits, n_fri = S - 3;
    size_t words = KRAK_BATCH_HEADER + n_cols + 8 + 8 + 1 + 8 * (size_t)n_fri + 1 + 32 + 1 + nq;
    words += 1 + 2 * (size_t)nq * n_cols + 2 * (size_t)nq * (1 + 8 * (size_t)S);
    words += 1 + 8 * (size_t)nq + 2 * (size_t)nq * (1 + 8 * (size_t)S);
    words += 1;
    for (uint32_t k = 0; k < n_fri; k++) {
        uint32_t L = S - 1 - k;
        words += 1 + 8 * (size_t)nq + 2 * (size_t)nq * (1 + 8 * (size_t)L);
    }
    return 4 * words;
}

/* the column sizes enter the transcript first: Blake2s of their LE words */
static void batch_sizes_digest(const uint32_t *log_sizes, uint32_t n_cols, uint32_t out[8]) {
    uint8_t *buf = (uint8_t *)malloc(4 * (size_t)n_cols);
    for (uint32_t c = 0; c < n_cols; c++) put32(buf + 4 * c, log_sizes[c]);
    uint8_t h[32];
    or_blake2s(buf, 4 * (size_t)n_cols, h);
    for (int k = 0; k < 8; k++) out[k] = le32(h + 4 * k);
    free(buf);
}

int or_batch_prove(uint64_t seed, const uint32_t *log_sizes, uint32_t n_cols, uint32_t blowup_bits, uint32_t nq,
                   uint32_t pow_bits, uint8_t **out, size_t *out_len) {
    if (n_cols < 1 || n_cols > 4096 || blowup_bits < 1 || nq == 0) return -1;
    for (uint32_t c = 0; c < n_cols; c++)
        if (log_sizes[c] < 2 || log_sizes[c] + blowup_bits > 31 || (c && log_sizes[c] < log_sizes[c - 1]))
            return -1;
    const uint32_t S = log_sizes[n_cols - 1] + blowup_bits;
    if (S < 4) return -1;
    const uint64_t N = 1ull << S;
    uint32_t **cols = (uint32_t **)calloc(n_cols, sizeof(uint32_t *));
    uint32_t *elog = (uint32_t *)malloc(4 * (size_t)n_cols);
    for (uint32_t c = 0; c < n_cols; c++) {
        uint32_t L = log_sizes[c];
        elog[c] = L + blowup_bits;
        cols[c] = (uint32_t *)malloc(4ull << elog[c]);
        or_d2_column(seed, c, L, cols[c]);
        lde_column(cols[c], L, elog[c]);
        or_bit_reverse_m31(cols[c], elog[c]);
    }
    /* lifted tree */
    uint32_t *leaves = (uint32_t *)malloc(N * 32);
    or_build_leaves_lifted((const uint32_t *const *)cols, elog, n_cols, S, leaves);
    tree_t tt;
    tree_build_from_leaves(&tt, S, leaves);
    or_channel ch;
    or_channel_init(&ch);
    uint32_t sd[8];
    batch_sizes_digest(log_sizes, n_cols, sd);
    or_channel_mix_digest(&ch, sd);
    or_channel_mix_digest(&ch, tree_root(&tt));
    uint32_t alpha[4];
    or_channel_draw_felt(&ch, alpha);
    /* RLC quotient */
    qm31 *beta = (qm31 *)malloc(sizeof(qm31) * n_cols);
    d2_powers(alpha, n_cols, beta);
    uint32_t *q[4];
    for (int k = 0; k < 4; k++) q[k] = (uint32_t *)malloc(N * 4);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)N; i++) {
        qm31 acc = {{0, 0, 0, 0}};
        for (uint32_t c = 0; c < n_cols; c++)
            acc = q_add(acc, q_scale(beta[c], cols[c][lifted_row((uint64_t)i, S, elog[c])]));
        for (int k = 0; k < 4; k++) q[k][i] = acc.v[k];
    }
    tree_t qtree;
    tree_commit(&qtree, (const uint32_t *const *)q, 4, S);
    or_channel_mix_digest(&ch, tree_root(&qtree));
    fri_run_t fr;
    fri_commit_and_query(&ch, q, S, nq, pow_bits, &fr);

    wbuf w = {0};
    wb_u32(&w, KRAK_MAGIC);
    wb_u32(&w, 4);
    wb_u32(&w, n_cols);
    wb_u32(&w, blowup_bits);
    wb_u32(&w, pow_bits);
    wb_u32(&w, (uint32_t)fr.nonce);
    wb_u32(&w, (uint32_t)(fr.nonce >> 32));
    wb_u32(&w, (uint32_t)seed);
    wb_u32(&w, (uint32_t)(seed >> 32));
    wb_words(&w, log_sizes, n_cols);
    wb_words(&w, tree_root(&tt), 8);
    wb_words(&w, tree_root(&qtree), 8);
    fri_write_roots_and_queries(&w, &fr);
    wb_u32(&w, nq);
    for (int side = 0; side < 2; side++)
        for (uint32_t i = 0; i < nq; i++) {
            uint64_t idx = fr.qi[i] ^ (uint64_t)side;
            for (uint32_t c = 0; c < n_cols; c++) wb_u32(&w, cols[c][lifted_row(idx, S, elog[c])]);
        }
    for (uint32_t i = 0; i < nq; i++) wb_path(&w, &tt, fr.qi[i]);
    for (uint32_t i = 0; i < nq; i++) wb_path(&w, &tt, fr.qi[i] ^ 1);
    wb_decommit(&w, &qtree, (const uint32_t *const *)q, 4, fr.qi, nq);
    fri_write_decommits(&w, &fr);

    tree_free(&tt);
    tree_free(&qtree);
    fri_free(&fr);
    for (int k = 0; k < 4; k++) free(q[k]);
    for (uint32_t c = 0; c < n_cols; c++) free(cols[c]);
    free(cols);
    free(elog);
    free(beta);
    *out = w.p;
    *out_len = w.len;
    return 0;
}

static void lifted_hash_words(const uint32_t *vals, uint32_t n, uint32_t out[8]) {
    uint8_t *buf = (uint8_t *)malloc(4 * (size_t)n + 4);
    for (uint32_t c = 0; c < n; c++) put32(buf + 4 * c, vals[c]);
    uint8_t h[32];
    or_blake2s(buf, 4 * (size_t)n, h);
    for (int k = 0; k < 8; k++) out[k] = le32(h + 4 * k);
    free(buf);
}

int or_batch_verify(const uint8_t *bytes, size_t len, uint32_t blowup_bits, uint32_t nq, char *err, size_t err_cap) {
    int ok = 0;
    rbuf r = {bytes, len, 0, 0};
    decom_t qd = {0}, td = {0}, *fd = NULL;
    uint32_t n_fd = 0, n_fri = 0, n_last = 0, n_q = 0, C = 0;
    uint32_t *logs = NULL, (*froots)[8] = NULL, *lastw = NULL;
    uint64_t *qidx = NULL, *fidx = NULL;
    qm31 *alphas = NULL, *beta = NULL;
    char dummy[8];
    if (!err || err_cap == 0) { err = dummy; err_cap = sizeof dummy; }
    err[0] = 0;

    if (rb_u32(&r) != KRAK_MAGIC) VFAIL("bad magic");
    if (rb_u32(&r) != 4) VFAIL("unsupported proof version");
    C = rb_u32(&r);
    uint32_t bb = rb_u32(&r), pow_bits = rb_u32(&r);
    uint64_t nonce = rb_u32(&r);
    nonce |= (uint64_t)rb_u32(&r) << 32;
    (void)rb_u32(&r);
    (void)rb_u32(&r);
    if (r.bad || C < 1 || C > 4096) VFAIL("bad column count");
    if (bb != blowup_bits) VFAIL("blowup mismatch");
    if (pow_bits > 128) VFAIL("bad pow bits");
    logs = (uint32_t *)calloc(C, 4);
    for (uint32_t c = 0; c < C; c++) {
        logs[c] = rb_u32(&r);
        if (r.bad || logs[c] < 1 || logs[c] + bb > 31 || (c && logs[c] < logs[c - 1])) VFAIL("bad column sizes");
    }
    uint32_t S = logs[C - 1] + bb;
    uint32_t troot[8], qroot[8];
    for (int i = 0; i < 8; i++) troot[i] = rb_u32(&r);
    for (int i = 0; i < 8; i++) qroot[i] = rb_u32(&r);
    n_fri = rb_u32(&r);
    if (r.bad || n_fri > 64) VFAIL("truncated proof");
    froots = (uint32_t(*)[8])calloc(n_fri + 1, 32);
    for (uint32_t k = 0; k < n_fri; k++)
        for (int i = 0; i < 8; i++) froots[k][i] = rb_u32(&r);
    n_last = rb_u32(&r);
    if (r.bad || n_last > 4096) VFAIL("truncated proof");
    lastw = (uint32_t *)calloc(4 * (size_t)n_last + 1, 4);
    for (size_t i = 0; i < 4 * (size_t)n_last; i++) lastw[i] = rb_u32(&r);
    n_q = rb_u32(&r);
    if (r.bad || n_q > (1u << 20)) VFAIL("truncated proof");
    qidx = (uint64_t *)calloc(n_q + 1, 8);
    for (uint32_t i = 0; i < n_q; i++) qidx[i] = rb_u32(&r);
    if (!rb_decom(&r, &td, C)) VFAIL("truncated proof");
    if (!rb_decom(&r, &qd, 4)) VFAIL("truncated proof");
    n_fd = rb_u32(&r);
    if (r.bad || n_fd > 64) VFAIL("truncated proof");
    fd = (decom_t *)calloc(n_fd + 1, sizeof(decom_t));
    for (uint32_t k = 0; k < n_fd; k++)
        if (!rb_decom(&r, &fd[k], 4)) VFAIL("truncated proof");
    if (r.pos != r.len) VFAIL("trailing bytes");

    uint64_t N = 1ull << S;
    if (n_q != nq) VFAIL("Expected %u queries, got %u", nq, n_q);
    or_channel ch;
    or_channel_init(&ch);
    uint32_t sd[8];
    batch_sizes_digest(logs, C, sd);
    or_channel_mix_digest(&ch, sd);
    or_channel_mix_digest(&ch, troot);
    uint32_t alpha[4];
    or_channel_draw_felt(&ch, alpha);
    or_channel_mix_digest(&ch, qroot);
    alphas = (qm31 *)calloc(n_fri + 2, sizeof(qm31));
    or_channel_draw_felt(&ch, alphas[0].v);
    for (uint32_t k = 0; k < n_fri; k++) {
        or_channel_mix_digest(&ch, froots[k]);
        or_channel_draw_felt(&ch, alphas[k + 1].v);
    }
    if (n_fri != S - 3) VFAIL("Expected %u FRI layers, got %u", S - 3, n_fri);
    if (n_last != 8) VFAIL("Expected 8 FRI last layer values, got %u", n_last);
    or_channel_mix_felts(&ch, lastw, n_last);
    if (pow_bits) {
        if (!or_channel_verify_pow_nonce(&ch, pow_bits, nonce)) VFAIL("PoW nonce invalid");
        or_channel_mix_u64(&ch, nonce);
    } else if (nonce != 0) {
        VFAIL("nonce without PoW");
    }
    for (uint32_t i = 0; i < nq; i++)
        if (or_channel_draw_number(&ch, N) != qidx[i]) VFAIL("Query indices don't match Fiat-Shamir derivation");
    /* lifted trace openings (leaf = Blake2s of all C values) */
    if (td.n != nq) VFAIL("trace decommitment size mismatch");
    for (uint32_t q = 0; q < nq; q++) {
        uint32_t h[8];
        lifted_hash_words(td.vals + (size_t)q * C, C, h);
        if (td.depth[q] != S || !or_verify_auth_path(troot, h, qidx[q], td.paths[q], td.depth[q]))
            VFAIL("trace auth path invalid");
        lifted_hash_words(td.sibs + (size_t)q * C, C, h);
        if (td.sdepth[q] != S || !or_verify_auth_path(troot, h, qidx[q] ^ 1, td.spaths[q], td.sdepth[q]))
            VFAIL("trace auth path invalid");
    }
    if (!check_decom(&qd, qroot, qidx, nq)) VFAIL("quotient auth path invalid");
    /* the random linear combination at q and q ^ 1 */
    beta = (qm31 *)malloc(sizeof(qm31) * C);
    d2_powers(alpha, C, beta);
    for (uint32_t q = 0; q < nq; q++)
        for (int side = 0; side < 2; side++) {
            const uint32_t *v = (side ? td.sibs : td.vals) + (size_t)q * C;
            qm31 acc = {{0, 0, 0, 0}};
            for (uint32_t c = 0; c < C; c++) acc = q_add(acc, q_scale(beta[c], v[c]));
            if (memcmp(acc.v, (side ? qd.sibs : qd.vals) + 4 * (size_t)q, 16))
                VFAIL("Quotient combination mismatch at query %u", q);
        }
    fidx = (uint64_t *)calloc(nq + 1, 8);
    for (uint32_t i = 0; i < nq; i++) fidx[i] = qidx[i] / 2;
    for (uint32_t k = 0; k < n_fd && k < n_fri; k++) {
        if (!check_decom(&fd[k], froots[k], fidx, nq)) VFAIL("FRI layer %u auth path invalid", k);
        for (uint32_t i = 0; i < nq; i++) fidx[i] /= 2;
    }
    if (n_fd != n_fri || n_fd == 0) VFAIL("FRI decommitment count mismatch");
    for (uint32_t q = 0; q < nq; q++) {
        uint64_t ci = qidx[q];
        uint32_t cl = S;
        qm31 f0 = q_load(qd.vals + 4 * (size_t)q), f1 = q_load(qd.sibs + 4 * (size_t)q);
        if (ci % 2) { qm31 t = f0; f0 = f1; f1 = t; }
        qm31 e = fold_pair(f0, f1, alphas[0], fold_twiddle_at(eval_domain(cl), ci / 2, 1));
        if (memcmp(e.v, fd[0].vals + 4 * (size_t)q, 16)) VFAIL("Circle fold mismatch at query %u", q);
        ci /= 2;
        cl--;
        for (uint32_t k = 0; k + 1 < n_fd; k++) {
            qm31 ex = pair_fold(&fd[k], q, ci, alphas[k + 1], fold_twiddle_at(coset_make(OR_HALF_ODDS, cl), ci / 2, 0));
            if (memcmp(ex.v, fd[k + 1].vals + 4 * (size_t)q, 16)) VFAIL("Line fold mismatch at query %u, layer %u", q, k);
            ci /= 2;
            cl--;
        }
        if (ci >= n_last) VFAIL("FRI last layer index out of bounds");
        if (memcmp(fd[n_fd - 1].vals + 4 * (size_t)q, lastw + 4 * ci, 16)) VFAIL("FRI last layer mismatch at query %u", q);
    }
    ok = 1;
done:
    decom_free(&td);
    decom_free(&qd);
    if (fd)
        for (uint32_t k = 0; k < n_fd; k++) decom_free(&fd[k]);
    free(fd);
    free(logs);
    free(froots);
    free(lastw);
    free(qidx);
    free(fidx);
    free(alphas);
    free(beta);
    return ok;
}

/* ======================================================================
 * stwo-basis circle NTT (the stwo Backend plugin's PolyOps).
 * Twiddles: src/ntt.rs:296-354 (StwoNttCache::new, which restates stwo's
 * slow_precompute_twiddles, garrick247/stwo-fork rev 5c41bf4a, for the
 * half_odds(log_n - 1) coset). Transform: cuda/circle_ntt_stwo.cu:30-56
 * (line layer), :202-248 (circle layer), :268-402 (layer order and the
 * 1/n scale). The fused shared-memory and radix-4 kernels of that file
 * (:71-200) apply the same butterflies with the same twiddle index per
 * butterfly, so one layer at a time gives identical values.
 * Basis and evaluation order: src/oods.rs:133-161 (eval_at_oods_from_coeffs)
 * and src/ntt.rs:517-548 (value k is f at CircleDomain point bitrev(k)).
 * ====================================================================== */
void or_stwo_twiddles(uint32_t log_n, uint32_t *fwd, uint32_t *inv) {
    uint64_t tc = 1ull << (log_n - 1), pos = 0;
    coset_t c = coset_make(OR_HALF_ODDS, log_n - 1);
    uint32_t *xs = (uint32_t *)malloc((tc > 1 ? tc / 2 : 1) * 4);
    while (c.log_size > 0) {
        /* x of the first half of the coset, stored at bit-reversed positions */
        uint32_t layer_log = c.log_size - 1;
        uint64_t half = 1ull << layer_log;
        cpt p = c.init;
        for (uint64_t j = 0; j < half; j++) {
            xs[j] = p.x;
            p = c_mul(p, c.step);
        }
        for (uint64_t j = 0; j < half; j++) fwd[pos + bitrev(j, layer_log)] = xs[j];
        pos += half;
        c.init = c_mul(c.init, c.init);
        c.step = c_mul(c.step, c.step);
        c.log_size--;
    }
    fwd[pos] = 1; /* pos == tc - 1: the pad */
    free(xs);
    /* every entry is the x of a point of order >= 8, never 0 */
    for (uint64_t i = 0; i < tc; i++) inv[i] = m31_inv(fwd[i]);
}

/* circle layer twiddle of butterfly t: (y, -y, -x, x)[t & 3] with
 * (x, y) = (tw[2(t >> 2)], tw[2(t >> 2) + 1]) (circle_ntt_stwo.cu:214-234) */
static uint32_t *stwo_circle_table(const uint32_t *tw, uint64_t half_n) {
    uint32_t *ct = (uint32_t *)malloc(half_n * 4);
    for (uint64_t t = 0; t < half_n; t++) {
        uint32_t x = tw[2 * (t >> 2)], y = tw[2 * (t >> 2) + 1];
        uint32_t r = t & 3;
        ct[t] = r == 0 ? y : r == 1 ? m31_neg(y) : r == 2 ? m31_neg(x) : x;
    }
    return ct;
}

void or_stwo_ntt_evaluate(uint32_t *d, const uint32_t *tw, uint32_t log_n) {
    uint64_t n = 1ull << log_n, half_n = n / 2;
    uint32_t nl = log_n - 1;
    for (int k = (int)nl - 1; k >= 0; k--)
        or_ntt_layer(d, tw + (half_n - (1ull << (nl - (uint32_t)k))), (uint32_t)k + 1, n, 1);
    uint32_t *ct = stwo_circle_table(tw, half_n);
    or_ntt_layer(d, ct, 0, n, 1);
    free(ct);
}

void or_stwo_ntt_interpolate(uint32_t *d, const uint32_t *itw, uint32_t log_n) {
    uint64_t n = 1ull << log_n, half_n = n / 2;
    uint32_t nl = log_n - 1;
    uint32_t *ct = stwo_circle_table(itw, half_n);
    or_ntt_layer(d, ct, 0, n, 0);
    free(ct);
    for (uint32_t k = 0; k < nl; k++) or_ntt_layer(d, itw + (half_n - (1ull << (nl - k))), k + 1, n, 0);
    uint32_t e = (30u * log_n) % 31u;
    uint32_t inv_n = e == 0 ? 1u : (1u << e);
    for (uint64_t i = 0; i < n; i++) d[i] = m31_mul(d[i], inv_n);
}

/* src/oods.rs:133-161, 258-290: f(z) = sum_i c_i prod_{bit b of i} m_b with
 * m_0 = z.y, m_1 = z.x, m_b = 2 m_{b-1}^2 - 1 (src/oods.rs:96-99); folded
 * from the lowest bit up (exact, so the order does not matter). */
void or_stwo_eval_at_point(const uint32_t *coeffs, uint32_t log_n, const uint32_t zx[4], const uint32_t zy[4],
                           uint32_t out[4]) {
    uint64_t n = 1ull << log_n;
    qm31 *v = (qm31 *)malloc(n * sizeof(qm31));
    for (uint64_t i = 0; i < n; i++) v[i] = (qm31){{coeffs[i], 0, 0, 0}};
    qm31 x = q_load(zx), m = q_load(zy);
    const qm31 one = {{1, 0, 0, 0}};
    for (uint32_t b = 0; b < log_n; b++) {
        if (b == 1) m = x;
        if (b >= 2) {
            qm31 m2 = q_mul(m, m);
            m = q_sub(q_add(m2, m2), one);
        }
        uint64_t half = n >> (b + 1);
        for (uint64_t i = 0; i < half; i++) v[i] = q_add(v[2 * i], q_mul(m, v[2 * i + 1]));
    }
    memcpy(out, v[0].v, 16);
    free(v);
}

/* cuda/circle_ntt.cu:186-231, 429-468 (cuda_eval_at_point): fold the
 * coefficient vector in halves, level l combining v[t] + factors[l] * v[t + half]. */
void or_eval_at_point_factors(const uint32_t *coeffs, const uint32_t *factors, uint32_t log_n, uint32_t out[4]) {
    uint64_t n = 1ull << log_n;
    if (log_n == 0) {
        out[0] = coeffs[0]; out[1] = out[2] = out[3] = 0;
        return;
    }
    uint64_t half = n / 2;
    qm31 *v = (qm31 *)malloc(half * sizeof(qm31));
    qm31 f = q_load(factors);
    for (uint64_t t = 0; t < half; t++) {
        qm31 bf = q_scale(f, coeffs[t + half]);
        v[t] = bf;
        v[t].v[0] = m31_add(coeffs[t], bf.v[0]);
    }
    for (uint32_t level = 1; level < log_n; level++) {
        half /= 2;
        f = q_load(factors + 4 * level);
        for (uint64_t t = 0; t < half; t++) v[t] = q_add(v[t], q_mul(v[t + half], f));
    }
    memcpy(out, v[0].v, 16);
    free(v);
}

/* cuda/barycentric_eval.cu:41-110: out[b] = sum of evals[i] * weights[i]
 * over i = b*256 + t + k*n_blocks*256, t < 256. */
void or_barycentric_partials(const uint32_t *evals, const uint32_t *weights, uint64_t n, uint32_t n_blocks,
                             uint32_t *out) {
    for (uint32_t b = 0; b < n_blocks; b++) {
        qm31 acc = {{0, 0, 0, 0}};
        for (uint64_t base = (uint64_t)b * 256; base < n; base += (uint64_t)n_blocks * 256)
            for (uint64_t i = base; i < base + 256 && i < n; i++)
                acc = q_add(acc, q_scale(q_load(weights + 4 * i), evals[i]));
        memcpy(out + 4 * (size_t)b, acc.v, 16);
    }
}

/* ======================================================================
 * stwo plugin QuotientOps: cuda/quotient.cu:13-50 (accumulate_numerators),
 * :56-131 (the fused z / z_next form), :153-233 (compute_quotients_combine).
 * Column and numerator pointer arrays are given as host arrays here.
 * ====================================================================== */
void or_accumulate_numerators(const uint32_t *const *cols, const uint32_t *col_indices, const uint32_t *b,
                              const uint32_t *c, uint32_t n_batch_cols, uint64_t n_rows, uint32_t *o0,
                              uint32_t *o1, uint32_t *o2, uint32_t *o3) {
    for (uint64_t row = 0; row < n_rows; row++) {
        qm31 acc = {{0, 0, 0, 0}};
        for (uint32_t i = 0; i < n_batch_cols; i++) {
            uint32_t f = cols[col_indices[i]][row];
            acc = q_add(acc, q_sub(q_scale(q_load(c + 4 * i), f), q_load(b + 4 * i)));
        }
        o0[row] = acc.v[0]; o1[row] = acc.v[1]; o2[row] = acc.v[2]; o3[row] = acc.v[3];
    }
}

void or_quotients_combine(const uint32_t *sp_x, const uint32_t *sp_y, const uint32_t *first_linear_acc,
                          const uint32_t *const *numer0, const uint32_t *const *numer1,
                          const uint32_t *const *numer2, const uint32_t *const *numer3,
                          const uint32_t *acc_log_sizes, uint32_t n_accs, const uint32_t *domain_xs,
                          const uint32_t *domain_ys, uint32_t lifting_log_size, uint64_t n_rows, uint32_t *o0,
                          uint32_t *o1, uint32_t *o2, uint32_t *o3) {
    for (uint64_t row = 0; row < n_rows; row++) {
        uint32_t dx = domain_xs[row], dy = domain_ys[row];
        qm31 quot = {{0, 0, 0, 0}};
        for (uint32_t j = 0; j < n_accs; j++) {
            cm31 prx = {sp_x[4 * j], sp_x[4 * j + 1]}, pix = {sp_x[4 * j + 2], sp_x[4 * j + 3]};
            cm31 pry = {sp_y[4 * j], sp_y[4 * j + 1]}, piy = {sp_y[4 * j + 2], sp_y[4 * j + 3]};
            cm31 denom = cm_sub(cm_mul(cm_sub(prx, (cm31){dx, 0}), piy), cm_mul(cm_sub(pry, (cm31){dy, 0}), pix));
            cm31 den_inv = cm_inv(denom);
            uint32_t lr = lifting_log_size - acc_log_sizes[j];
            uint64_t li = ((row >> (lr + 1)) << 1) + (row & 1);
            qm31 pn = {{numer0[j][li], numer1[j][li], numer2[j][li], numer3[j][li]}};
            qm31 full = q_sub(pn, q_scale(q_load(first_linear_acc + 4 * j), dy));
            cm31 ra = cm_mul(q_a(full), den_inv), rb = cm_mul(q_b(full), den_inv);
            quot = q_add(quot, q_make(ra, rb));
        }
        o0[row] = quot.v[0]; o1[row] = quot.v[1]; o2[row] = quot.v[2]; o3[row] = quot.v[3];
    }
}

/* ---- stwo plugin column ops: cuda/accumulate_lift.cu:14-40 and the
 *      plugin's lift_and_accumulate loop (accumulation_ops.rs:44-85) ---- */
void or_accumulate_lift(uint32_t *col, const uint32_t *curr, uint64_t col_n, uint32_t log_ratio) {
    for (uint64_t i = 0; i < col_n; i++)
        col[i] = m31_add(col[i], curr[((i >> (log_ratio + 1)) << 1) | (i & 1)]);
}
/* one channel: curr = zeros(2); for each column (ascending):
 * col += lift(curr); curr = col. Columns are updated in place. */
void or_lift_and_accumulate(uint32_t *const *cols, const uint32_t *log_sizes, uint32_t n_cols) {
    uint32_t zeros[2] = {0, 0};
    const uint32_t *curr = zeros;
    uint32_t curr_log = 1;
    for (uint32_t k = 0; k < n_cols; k++) {
        or_accumulate_lift(cols[k], curr, 1ull << log_sizes[k], log_sizes[k] - curr_log);
        curr = cols[k];
        curr_log = log_sizes[k];
    }
}

/* The OODS step (src/oods.rs): t = draw_felt, z = ((1 - t^2)/(1 + t^2),
 * 2t/(1 + t^2)) (:44-52), z_next = z * step, step = generator doubled
 * 31 - trace_log times (:74-81, circle.rs:26-29), every column evaluated at
 * z and z_next (or_stwo_eval_at_point, the basis sum), then the values
 * mixed into the channel, all f_k(z) first (:395-406, channel.rs:46-59).
 * points: z.x, z.y, z_next.x, z_next.y; values: 8 n_cols words. */
void or_oods_sample(or_channel *ch, const uint32_t *const *cols, const uint32_t *log_sizes, uint32_t n_cols,
                    uint32_t trace_log, uint32_t points[16], uint32_t *values) {
    uint32_t tw[4];
    or_channel_draw_felt(ch, tw);
    const qm31 t = q_load(tw), one = {{1, 0, 0, 0}};
    const qm31 t2 = q_mul(t, t);
    const qm31 inv = q_inv(q_add(one, t2));
    const qm31 zx = q_mul(q_sub(one, t2), inv), zy = q_mul(q_add(t, t), inv);
    uint32_t st[2] = {2u, 1268011823u};
    for (uint32_t i = 0; i < 31 - trace_log; i++) {
        uint32_t sq[2];
        or_circle_mul(st, st, sq);
        st[0] = sq[0];
        st[1] = sq[1];
    }
    const qm31 nx = q_sub(q_scale(zx, st[0]), q_scale(zy, st[1]));
    const qm31 ny = q_add(q_scale(zx, st[1]), q_scale(zy, st[0]));
    const qm31 pts[4] = {zx, zy, nx, ny};
    for (int p = 0; p < 4; p++)
        for (int k = 0; k < 4; k++) points[4 * p + k] = pts[p].v[k];
    for (uint32_t c = 0; c < n_cols; c++) {
        or_stwo_eval_at_point(cols[c], log_sizes[c], zx.v, zy.v, values + 4 * c);
        or_stwo_eval_at_point(cols[c], log_sizes[c], nx.v, ny.v, values + 4 * (n_cols + c));
    }
    or_channel_mix_felts(ch, values, 2 * (size_t)n_cols);
}


/* ======================================================================
 * Cairo VM AIR constraint quotient: cuda/cairo_constraint.cu.
 * 34 trace columns (pc, ap, fp, inst_lo, inst_hi, 15 flags, dst_addr, dst,
 * op0_addr, op0, op1_addr, op1, res, off0..2, dst_inv, dict key/new/active,
 * :8-53); interaction columns as QM31 SoA (4 words each); challenges
 * z_mem, alpha_mem, alpha_mem^2, z_rc, z_dict, alpha_dict (:179-180).
 * ====================================================================== */
enum {
    CA_PC = 0, CA_AP = 1, CA_FP = 2, CA_INST_LO = 3, CA_INST_HI = 4, CA_FLAGS = 5, CA_DST_ADDR = 20,
    CA_DST = 21, CA_OP0_ADDR = 22, CA_OP0 = 23, CA_OP1_ADDR = 24, CA_OP1 = 25, CA_RES = 26, CA_OFF0 = 27,
    CA_OFF1 = 28, CA_OFF2 = 29, CA_DST_INV = 30, CA_DICT_KEY = 31, CA_DICT_NEW = 32, CA_DICT_ACTIVE = 33
};
#define CA_BIAS 0x8000u

/* Z_H(x) = f_{log_n}(x) + 1 with f_0 = x, f_{k+1} = 2 f_k^2 - 1 (:64-75) */
static uint32_t cairo_vanishing(uint32_t x, uint32_t log_n) {
    uint32_t v = x;
    for (uint32_t k = 0; k < log_n; k++) {
        uint32_t s = m31_mul(v, v);
        v = m31_sub(m31_add(s, s), 1);
    }
    return m31_add(v, 1);
}
/* :77-118, :990-1019: Z_H (or its inverse, 0 for 0) at eval position i,
 * the point init * step^{bitrev(i)} */
void or_cairo_vanishing(const uint32_t init[2], const uint32_t step[2], uint32_t *out, uint32_t log_eval,
                        uint32_t log_n, int inverse) {
    cpt in = {init[0], init[1]}, st = {step[0], step[1]};
    uint64_t n = 1ull << log_eval;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        cpt p = c_mul(in, c_pow(st, bitrev((uint64_t)i, log_eval)));
        uint32_t z = cairo_vanishing(p.x, log_n);
        out[i] = inverse ? (z == 0 ? 0 : m31_inv(z)) : z;
    }
}
/* :1037-1046 */
void or_axpy_m31(uint32_t s, const uint32_t *x, uint32_t *y, uint64_t n) {
    for (uint64_t i = 0; i < n; i++) y[i] = m31_add(y[i], m31_mul(s, x[i]));
}

static qm31 q_from(uint32_t a) { return (qm31){{a, 0, 0, 0}}; }

/* Constraints 0..30 of one row: the M31 part shared by both kernels
 * (:243-462 and the chunk kernel's copy). Each c is weighted by the next
 * alpha (qm31_mul_m31). */
static qm31 cairo_m31_part(const uint32_t *const cols[34], uint64_t i, uint64_t ni, uint32_t tf,
                           const uint32_t *alpha, int *ci) {
    uint32_t f[15];
    for (int j = 0; j < 15; j++) f[j] = cols[CA_FLAGS + j][i];
    const uint32_t pc = cols[CA_PC][i], ap = cols[CA_AP][i], fp = cols[CA_FP][i];
    const uint32_t dst = cols[CA_DST][i], op0 = cols[CA_OP0][i], op1 = cols[CA_OP1][i], res = cols[CA_RES][i];
    const uint32_t npc = cols[CA_PC][ni], nap = cols[CA_AP][ni], nfp = cols[CA_FP][ni];
    const uint32_t off0 = cols[CA_OFF0][i], off1 = cols[CA_OFF1][i], off2 = cols[CA_OFF2][i];
    const uint32_t dinv = cols[CA_DST_INV][i];
    const uint32_t op1_imm = f[2], res_add = f[5], res_mul = f[6], jabs = f[7], jrel = f[8], jnz = f[9];
    const uint32_t ap_add = f[10], ap_add1 = f[11], call = f[12], ret = f[13], asrt = f[14];
    uint32_t c[31];
    int k = 0;
    for (int j = 0; j < 15; j++) c[k++] = m31_mul(f[j], m31_sub(1, f[j]));
    {   /* 15 result */
        uint32_t exp = m31_add(m31_add(m31_mul(m31_sub(m31_sub(1, res_add), res_mul), op1),
                                       m31_mul(res_add, m31_add(op0, op1))),
                               m31_mul(res_mul, m31_mul(op0, op1)));
        c[k++] = m31_mul(m31_sub(1, jnz), m31_sub(res, exp));
    }
    {   /* 16 pc update */
        uint32_t notj = m31_sub(m31_sub(m31_sub(1, jabs), jrel), jnz);
        uint32_t tgt = m31_add(m31_add(m31_mul(notj, m31_add(pc, m31_add(1, op1_imm))), m31_mul(jabs, res)),
                               m31_mul(jrel, m31_add(pc, res)));
        uint32_t nonjnz = m31_mul(m31_sub(1, jnz), m31_sub(npc, tgt));
        uint32_t jz = m31_mul(jnz, m31_mul(dst, m31_sub(npc, m31_add(pc, op1))));
        c[k++] = m31_mul(m31_add(nonjnz, jz), tf);
    }
    /* 17 ap update, 18 fp update, 19 assert_eq */
    c[k++] = m31_mul(m31_sub(nap, m31_add(m31_add(m31_add(ap, m31_mul(ap_add, res)), ap_add1), m31_mul(call, 2))), tf);
    c[k++] = m31_mul(m31_sub(nfp, m31_add(m31_add(m31_mul(m31_sub(m31_sub(1, call), ret), fp),
                                                 m31_mul(call, m31_add(ap, 2))), m31_mul(ret, dst))), tf);
    c[k++] = m31_mul(asrt, m31_sub(dst, res));
    /* 20..22 operand addresses, offsets biased by 2^15 */
    c[k++] = m31_sub(cols[CA_DST_ADDR][i],
                     m31_sub(m31_add(m31_add(m31_mul(m31_sub(1, f[0]), ap), m31_mul(f[0], fp)), off0), CA_BIAS));
    c[k++] = m31_sub(cols[CA_OP0_ADDR][i],
                     m31_sub(m31_add(m31_add(m31_mul(m31_sub(1, f[1]), ap), m31_mul(f[1], fp)), off1), CA_BIAS));
    {
        uint32_t dflt = m31_sub(m31_sub(m31_sub(1, op1_imm), f[3]), f[4]);
        uint32_t base = m31_add(m31_add(m31_add(m31_mul(op1_imm, pc), m31_mul(f[3], fp)), m31_mul(f[4], ap)),
                                m31_mul(dflt, op0));
        c[k++] = m31_sub(cols[CA_OP1_ADDR][i], m31_sub(m31_add(base, off2), CA_BIAS));
    }
    /* 23 jnz fall-through, 24 jnz inverse */
    c[k++] = m31_mul(m31_mul(jnz, m31_mul(m31_sub(1, m31_mul(dst, dinv)), m31_sub(npc, m31_add(pc, m31_add(1, op1_imm))))),
                     tf);
    c[k++] = m31_mul(jnz, m31_mul(dst, m31_sub(1, m31_mul(dst, dinv))));
    /* 25..29 exclusivity */
    c[k++] = m31_mul(op1_imm, f[3]);
    c[k++] = m31_mul(op1_imm, f[4]);
    c[k++] = m31_mul(f[3], f[4]);
    c[k++] = m31_add(m31_add(m31_mul(jabs, jrel), m31_mul(jabs, jnz)), m31_mul(jrel, jnz));
    c[k++] = m31_add(m31_add(m31_mul(call, ret), m31_mul(call, asrt)), m31_mul(ret, asrt));
    {   /* 30 instruction word: 2^31 = 1, 2^32 = 2, 2^(48+j) = 2^(17+j), 2^62 = 1 */
        uint32_t rhs = m31_add(m31_add(off0, m31_mul(off1, 1u << 16)), m31_mul(off2, 2));
        for (int j = 0; j < 14; j++) rhs = m31_add(rhs, m31_mul(f[j], 1u << (17 + j)));
        rhs = m31_add(rhs, f[14]);
        c[k++] = m31_sub(m31_add(cols[CA_INST_LO][i], cols[CA_INST_HI][i]), rhs);
    }
    qm31 acc = {{0, 0, 0, 0}};
    for (int j = 0; j < k; j++) acc = q_add(acc, q_scale(q_load(alpha + 4 * (*ci)++), c[j]));
    return acc;
}

/* QM31 column value at row i of a SoA group of 4 columns */
static qm31 q_at(const uint32_t *const g[4], uint64_t i) { return (qm31){{g[0][i], g[1][i], g[2][i], g[3][i]}}; }

/* cuda_cairo_quotient / _slab (:157-580, wrappers :860-960): 40 weighted
 * sub-constraints; LogUp (31a-d), range check (32a-c) and dict (34) in
 * polynomial form over the interaction columns. inter: 32 column pointers in
 * FFI order (S_logup, T1, T2, T3, S_rc, U1, U2, S_dict; 4 words each).
 * Full mode: rows 0..n of n, next row (i + blowup_step) mod n. Slab mode:
 * the columns are slabs of n + blowup_step rows starting at global row
 * chunk_offset, next row i + blowup_step; vh_inv, trans and out are global. */
void or_cairo_quotient(const uint32_t *const cols[34], const uint32_t *const inter[32], uint32_t *const out[4],
                       const uint32_t alpha[160], const uint32_t *vh_inv, const uint32_t *trans,
                       const uint32_t chal[24], uint64_t n, uint32_t blowup_step, uint64_t chunk_offset,
                       int slab_mode) {
    const qm31 z_mem = q_load(chal), a_mem = q_load(chal + 4), a_mem2 = q_load(chal + 8);
    const qm31 z_rc = q_load(chal + 12), z_dict = q_load(chal + 16), a_dict = q_load(chal + 20);
    const qm31 one = {{1, 0, 0, 0}};
#pragma omp parallel for schedule(static)
    for (int64_t si = 0; si < (int64_t)n; si++) {
        const uint64_t i = (uint64_t)si;
        const uint64_t g = slab_mode ? chunk_offset + i : i;
        const uint64_t ni = slab_mode ? i + blowup_step : (i + blowup_step) % n;
        const uint32_t tf = trans[g];
        int ci = 0;
        qm31 acc = cairo_m31_part(cols, i, ni, tf, alpha, &ci);
        const uint32_t pc = cols[CA_PC][i], dst = cols[CA_DST][i], op0 = cols[CA_OP0][i], op1 = cols[CA_OP1][i];
        const qm31 s = q_at(inter + 0, i), sn = q_at(inter + 0, ni);
        const qm31 t1 = q_at(inter + 4, i), t2 = q_at(inter + 8, i), t3 = q_at(inter + 12, i);
        const qm31 d0 = q_sub(z_mem, q_add(q_from(pc), q_add(q_scale(a_mem, cols[CA_INST_LO][i]),
                                                              q_scale(a_mem2, cols[CA_INST_HI][i]))));
        const qm31 d1 = q_sub(z_mem, q_add(q_from(cols[CA_DST_ADDR][i]), q_scale(a_mem, dst)));
        const qm31 d2 = q_sub(z_mem, q_add(q_from(cols[CA_OP0_ADDR][i]), q_scale(a_mem, op0)));
        const qm31 d3 = q_sub(z_mem, q_add(q_from(cols[CA_OP1_ADDR][i]), q_scale(a_mem, op1)));
        qm31 cq[8];
        cq[0] = q_sub(q_mul(q_sub(t1, s), d0), one);
        cq[1] = q_sub(q_mul(q_sub(t2, t1), d1), one);
        cq[2] = q_sub(q_mul(q_sub(t3, t2), d2), one);
        cq[3] = q_scale(q_sub(q_mul(q_sub(sn, t3), d3), one), tf);
        const qm31 r = q_at(inter + 16, i), rn = q_at(inter + 16, ni);
        const qm31 u1 = q_at(inter + 20, i), u2 = q_at(inter + 24, i);
        cq[4] = q_sub(q_mul(q_sub(u1, r), q_sub(z_rc, q_from(cols[CA_OFF0][i]))), one);
        cq[5] = q_sub(q_mul(q_sub(u2, u1), q_sub(z_rc, q_from(cols[CA_OFF1][i]))), one);
        cq[6] = q_scale(q_sub(q_mul(q_sub(rn, u2), q_sub(z_rc, q_from(cols[CA_OFF2][i]))), one), tf);
        for (int k = 0; k < 7; k++) acc = q_add(acc, q_mul(q_load(alpha + 4 * ci++), cq[k]));
        const uint32_t act = cols[CA_DICT_ACTIVE][i];
        acc = q_add(acc, q_scale(q_load(alpha + 4 * ci++), m31_mul(act, m31_sub(1, act))));
        const qm31 ds = q_at(inter + 28, i), dsn = q_at(inter + 28, ni);
        const qm31 entry = q_add(q_from(cols[CA_DICT_KEY][i]), q_scale(a_dict, cols[CA_DICT_NEW][i]));
        const qm31 c34 = q_scale(q_sub(q_mul(q_sub(dsn, ds), q_sub(z_dict, entry)), q_from(act)), tf);
        acc = q_add(acc, q_mul(q_load(alpha + 4 * ci++), c34));
        const qm31 q = q_scale(acc, vh_inv[g]);
        for (int k = 0; k < 4; k++) out[k][g] = q.v[k];
    }
}

/* cuda_cairo_quotient_chunk (:582-850): rows offset .. offset + chunk_n of
 * global_n, 35 weighted constraints; LogUp (31) and range check (32) as one
 * step constraint each with the inverse-sum delta (:120-153); outputs at the
 * chunk-local row. inter: 12 pointers (S_logup, S_rc, S_dict; 4 words each). */
void or_cairo_quotient_chunk(const uint32_t *const cols[34], const uint32_t *const inter[12],
                             uint32_t *const out[4], const uint32_t alpha[140], const uint32_t *vh_inv,
                             const uint32_t *trans, const uint32_t chal[24], uint64_t offset, uint64_t chunk_n,
                             uint64_t global_n, uint32_t blowup_step) {
    const qm31 z_mem = q_load(chal), a_mem = q_load(chal + 4), a_mem2 = q_load(chal + 8);
    const qm31 z_rc = q_load(chal + 12), z_dict = q_load(chal + 16), a_dict = q_load(chal + 20);
#pragma omp parallel for schedule(static)
    for (int64_t sl = 0; sl < (int64_t)chunk_n; sl++) {
        const uint64_t i = offset + (uint64_t)sl;
        const uint64_t ni = (i + blowup_step) % global_n;
        const uint32_t tf = trans[i];
        int ci = 0;
        qm31 acc = cairo_m31_part(cols, i, ni, tf, alpha, &ci);
        const uint32_t pc = cols[CA_PC][i], dst = cols[CA_DST][i], op0 = cols[CA_OP0][i], op1 = cols[CA_OP1][i];
        const qm31 d0 = q_sub(z_mem, q_add(q_from(pc), q_add(q_scale(a_mem, cols[CA_INST_LO][i]),
                                                              q_scale(a_mem2, cols[CA_INST_HI][i]))));
        const qm31 d1 = q_sub(z_mem, q_add(q_from(cols[CA_DST_ADDR][i]), q_scale(a_mem, dst)));
        const qm31 d2 = q_sub(z_mem, q_add(q_from(cols[CA_OP0_ADDR][i]), q_scale(a_mem, op0)));
        const qm31 d3 = q_sub(z_mem, q_add(q_from(cols[CA_OP1_ADDR][i]), q_scale(a_mem, op1)));
        const qm31 dl = q_add(q_add(q_add(q_inv(d0), q_inv(d1)), q_inv(d2)), q_inv(d3));
        const qm31 c31 = q_scale(q_sub(q_sub(q_at(inter + 0, ni), q_at(inter + 0, i)), dl), tf);
        acc = q_add(acc, q_mul(q_load(alpha + 4 * ci++), c31));
        const qm31 dr = q_add(q_add(q_inv(q_sub(z_rc, q_from(cols[CA_OFF0][i]))),
                                    q_inv(q_sub(z_rc, q_from(cols[CA_OFF1][i])))),
                              q_inv(q_sub(z_rc, q_from(cols[CA_OFF2][i]))));
        const qm31 c32 = q_scale(q_sub(q_sub(q_at(inter + 4, ni), q_at(inter + 4, i)), dr), tf);
        acc = q_add(acc, q_mul(q_load(alpha + 4 * ci++), c32));
        const uint32_t act = cols[CA_DICT_ACTIVE][i];
        acc = q_add(acc, q_scale(q_load(alpha + 4 * ci++), m31_mul(act, m31_sub(1, act))));
        const qm31 entry = q_add(q_from(cols[CA_DICT_KEY][i]), q_scale(a_dict, cols[CA_DICT_NEW][i]));
        const qm31 c34 = q_scale(q_sub(q_mul(q_sub(q_at(inter + 8, ni), q_at(inter + 8, i)), q_sub(z_dict, entry)),
                                       q_from(act)),
                                 tf);
        acc = q_add(acc, q_mul(q_load(alpha + 4 * ci++), c34));
        const qm31 q = q_scale(acc, vh_inv[i]);
        for (int k = 0; k < 4; k++) out[k][sl] = q.v[k];
    }
}

/* ======================================================================
 * Generic constraint evaluation (cuda/constraint_eval.cu:97-407): the
 * reference's register bytecode interpreted row by row.
 * ====================================================================== */

static uint32_t bitrev32(uint32_t v, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t k = 0; k < bits; k++) r |= ((v >> k) & 1u) << (bits - 1 - k);
    return r;
}

/* stwo's offset_bit_reversed_circle_domain_index (cuda/constraint_eval.cu:
 * 63-94), with the step in 64 bits */
uint32_t or_bc_offset_row(uint32_t i, uint32_t domain_log, uint32_t eval_log, int32_t offset) {
    uint32_t prev = bitrev32(i, eval_log);
    const int64_t half = (int64_t)1 << (eval_log - 1);
    const int64_t step = (int64_t)offset * ((int64_t)1 << (eval_log - domain_log - 1));
    int64_t idx;
    if ((int64_t)prev < half) {
        idx = ((int64_t)prev + step) % half;
        if (idx < 0) idx += half;
    } else {
        idx = ((int64_t)prev - half - step) % half;
        if (idx < 0) idx += half;
        idx += half;
    }
    return bitrev32((uint32_t)idx, eval_log);
}

static uint32_t ilog2_u32(uint32_t v) {
    uint32_t l = 0;
    while ((1u << l) < v) l++;
    return l;
}

int or_bytecode_constraint_eval(const uint32_t *bc, uint32_t n_words, const uint32_t *const *cols,
                                uint32_t n_cols, uint32_t n_rows, uint32_t trace_n_rows, const uint32_t *coeff,
                                const uint32_t *denom_inv, uint32_t *acc0, uint32_t *acc1, uint32_t *acc2,
                                uint32_t *acc3) {
    qm31 *regs = (qm31 *)calloc(1u << 16, sizeof(qm31));
    if (!regs) return -1;
    const uint32_t eval_log = ilog2_u32(n_rows), trace_log = ilog2_u32(trace_n_rows);
    int rc = 0;
    for (uint32_t row = 0; row < n_rows && rc == 0; row++) {
        qm31 accq = {{0, 0, 0, 0}};
        uint32_t k = 0, pc = 0;
#define BC_NEXT(x)                          \
    do {                                    \
        if (pc >= n_words) { rc = -1; break; } \
        (x) = bc[pc++];                     \
    } while (0)
        while (pc < n_words && rc == 0) {
            const uint32_t w = bc[pc++], op = w >> 24, dst = w & 0xFFFFu;
            uint32_t w2 = 0, w3 = 0, c[4] = {0, 0, 0, 0};
            qm31 r = {{0, 0, 0, 0}};
            int writes = 1;
            switch (op) {
            case 0x01: BC_NEXT(w2); r.v[0] = w2; break;                         /* LoadConst */
            case 0x02: for (int j = 0; j < 4; j++) BC_NEXT(c[j]);               /* LoadSecureConst */
                       r = q_load(c); break;
            case 0x03: {                                                        /* LoadTrace */
                BC_NEXT(w2); BC_NEXT(w3);
                if (rc) break;
                if (w2 >= n_cols) { rc = -1; break; }
                const int32_t off = (w3 >> 31) ? -(int32_t)(w3 & 0x7FFFFFFFu) : (int32_t)(w3 & 0x7FFFFFFFu);
                const uint32_t rr = off == 0 ? row : or_bc_offset_row(row, trace_log, eval_log, off);
                r.v[0] = cols[w2] ? cols[w2][rr] : 0;
                break;
            }
            case 0x10: BC_NEXT(w2); r.v[0] = m31_add(regs[w2 >> 16].v[0], regs[w2 & 0xFFFFu].v[0]); break;
            case 0x11: BC_NEXT(w2); r.v[0] = m31_sub(regs[w2 >> 16].v[0], regs[w2 & 0xFFFFu].v[0]); break;
            case 0x12: BC_NEXT(w2); r.v[0] = m31_mul(regs[w2 >> 16].v[0], regs[w2 & 0xFFFFu].v[0]); break;
            case 0x13: BC_NEXT(w2); r.v[0] = m31_neg(regs[w2 >> 16].v[0]); break;
            case 0x14: BC_NEXT(w2); BC_NEXT(w3); r.v[0] = m31_add(regs[w2 >> 16].v[0], w3); break;
            case 0x15: BC_NEXT(w2); BC_NEXT(w3); r.v[0] = m31_mul(regs[w2 >> 16].v[0], w3); break;
            case 0x20: BC_NEXT(w2); r = q_add(regs[w2 >> 16], regs[w2 & 0xFFFFu]); break;
            case 0x21: BC_NEXT(w2); r = q_sub(regs[w2 >> 16], regs[w2 & 0xFFFFu]); break;
            case 0x22: BC_NEXT(w2); r = q_mul(regs[w2 >> 16], regs[w2 & 0xFFFFu]); break;
            case 0x23: BC_NEXT(w2); r = q_sub((qm31){{0, 0, 0, 0}}, regs[w2 >> 16]); break;
            case 0x24: BC_NEXT(w2); for (int j = 0; j < 4; j++) BC_NEXT(c[j]);
                       r = q_add(regs[w2 >> 16], q_load(c)); break;
            case 0x25: BC_NEXT(w2); for (int j = 0; j < 4; j++) BC_NEXT(c[j]);
                       r = q_mul(regs[w2 >> 16], q_load(c)); break;
            case 0x28: BC_NEXT(w2); r = regs[w2 >> 16];                         /* WideAddBase */
                       r.v[0] = m31_add(r.v[0], regs[w2 & 0xFFFFu].v[0]); break;
            case 0x29: BC_NEXT(w2); r = q_scale(regs[w2 >> 16], regs[w2 & 0xFFFFu].v[0]); break;
            case 0x2A: BC_NEXT(w2); for (int j = 0; j < 4; j++) BC_NEXT(c[j]);
                       r = q_add((qm31){{regs[w2 >> 16].v[0], 0, 0, 0}}, q_load(c)); break;
            case 0x2B: BC_NEXT(w2); for (int j = 0; j < 4; j++) BC_NEXT(c[j]);
                       r = q_scale(q_load(c), regs[w2 >> 16].v[0]); break;
            case 0x30: BC_NEXT(w2); r.v[0] = regs[w2 >> 16].v[0]; break;       /* Widen */
            case 0x31: BC_NEXT(w2); BC_NEXT(w3);                                /* CombineEF */
                       r = (qm31){{regs[w2 >> 16].v[0], regs[w2 & 0xFFFFu].v[0], regs[w3 >> 16].v[0],
                                   regs[w3 & 0xFFFFu].v[0]}};
                       break;
            case 0x40:                                                          /* AddConstraint */
                accq = q_add(accq, q_mul(regs[dst], q_load(coeff + 4 * k)));
                k++;
                writes = 0;
                break;
            default: rc = -1;
            }
            if (rc) break;
            if (writes) regs[dst] = r;
        }
#undef BC_NEXT
        if (rc) break;
        const qm31 res = q_scale(accq, denom_inv[row >> trace_log]);
        acc0[row] = m31_add(acc0[row], res.v[0]);
        acc1[row] = m31_add(acc1[row], res.v[1]);
        acc2[row] = m31_add(acc2[row], res.v[2]);
        acc3[row] = m31_add(acc3[row], res.v[3]);
    }
    free(regs);
    return rc;
}

/* ======================================================================
 * GKR sum-check operations of the stwo plugin: cuda/gkr.cu,
 * crates/cuda-backend/src/gkr_ops.rs. QM31 element i of a secure column is
 * words 4i..4i+3; a base column is one word per element.
 * ====================================================================== */
static inline void q_put(uint32_t *p, qm31 x) {
    p[0] = x.v[0];
    p[1] = x.v[1];
    p[2] = x.v[2];
    p[3] = x.v[3];
}

/* out[i] = r (in[i + n/2] - in[i]) + in[i], i < n/2 (cuda/gkr.cu:20-45; the
 * fold_mle_evals of the MLE's first variable). */
void or_gkr_fix_first_variable_base(const uint32_t *in, uint32_t *out, const uint32_t r[4], uint64_t n) {
    const uint64_t h = n / 2;
    const qm31 rq = q_load(r);
    for (uint64_t i = 0; i < h; i++) {
        qm31 o = q_scale(rq, m31_sub(in[i + h], in[i]));
        o.v[0] = m31_add(o.v[0], in[i]);
        q_put(out + 4 * i, o);
    }
}

/* The same over a secure column (cuda/gkr.cu:47-74). */
void or_gkr_fix_first_variable_secure(const uint32_t *in, uint32_t *out, const uint32_t r[4], uint64_t n) {
    const uint64_t h = n / 2;
    const qm31 rq = q_load(r);
    for (uint64_t i = 0; i < h; i++) {
        const qm31 a = q_load(in + 4 * i), b = q_load(in + 4 * (i + h));
        q_put(out + 4 * i, q_add(q_mul(rq, q_sub(b, a)), a));
    }
}

/* One doubling pass of gen_eq_evals (cuda/gkr.cu:84-118): for j < cur,
 * t = buf[j] y; buf[cur + j] = t; buf[j] -= t. */
void or_gkr_gen_eq_evals_pass(uint32_t *buf, const uint32_t y[4], uint64_t cur) {
    const qm31 yq = q_load(y);
    for (uint64_t j = 0; j < cur; j++) {
        const qm31 v = q_load(buf + 4 * j), t = q_mul(v, yq);
        q_put(buf + 4 * (cur + j), t);
        q_put(buf + 4 * j, q_sub(v, t));
    }
}

/* gen_eq_evals(y[0..k), v) (gkr_ops.rs:129-161): buf[0] = v, then passes
 * with y[k-1], ..., y[0] over 1, 2, ..., 2^(k-1) entries; out has 2^k. */
void or_gkr_gen_eq_evals(const uint32_t *y, uint32_t k, const uint32_t v[4], uint32_t *out) {
    memcpy(out, v, 16);
    for (uint32_t i = 0; i < k; i++) or_gkr_gen_eq_evals_pass(out, y + 4 * (k - 1 - i), 1ull << i);
}

/* next_layer (cuda/gkr.cu:120-212): kind 0 GrandProduct out[i] = in[2i] in[2i+1]
 * (in_num is the layer); 1 LogUpGeneric, 2 LogUpMultiplicities (M31
 * numerators, one word each), 3 LogUpSingles (numerators 1, in_num unused):
 * (n0 d1 + n1 d0, d0 d1) of the pair (2i, 2i+1). n = input length. */
void or_gkr_next_layer(int kind, const uint32_t *in_num, const uint32_t *in_den, uint32_t *out_num,
                       uint32_t *out_den, uint64_t n) {
    const uint64_t h = n / 2;
    for (uint64_t i = 0; i < h; i++) {
        if (kind == 0) {
            q_put(out_num + 4 * i, q_mul(q_load(in_num + 8 * i), q_load(in_num + 8 * i + 4)));
            continue;
        }
        const qm31 d0 = q_load(in_den + 8 * i), d1 = q_load(in_den + 8 * i + 4);
        qm31 rn;
        if (kind == 1) rn = q_add(q_mul(q_load(in_num + 8 * i), d1), q_mul(q_load(in_num + 8 * i + 4), d0));
        else if (kind == 2) rn = q_add(q_scale(d1, in_num[2 * i]), q_scale(d0, in_num[2 * i + 1]));
        else rn = q_add(d0, d1);
        q_put(out_num + 4 * i, rn);
        q_put(out_den + 4 * i, q_mul(d0, d1));
    }
}

/* sum_as_poly_in_first_variable's device sums (cuda/gkr.cu:260-540): for
 * i < n_terms, the pair at t = 0 is (2i, 2i+1) of the layer and at t = 1
 * (2(n_terms + i), 2(n_terms + i) + 1); t = 2 is 2 (t = 1) - (t = 0) per
 * element; F(pair) is the product (kind 0) or the pair's fraction sum
 * num + lambda den (kinds 1..3, as next_layer); out[0..4) = sum_i eq[i]
 * F(pair at 0), out[4..8) = sum_i eq[i] F(pair at 2). */
void or_gkr_sum_poly(int kind, const uint32_t *eq, const uint32_t *in_num, const uint32_t *in_den,
                     const uint32_t lambda[4], uint64_t n_terms, uint32_t out[8]) {
    const qm31 lam = kind == 0 ? (qm31){{0, 0, 0, 0}} : q_load(lambda);
    qm31 s0 = {{0, 0, 0, 0}}, s2 = {{0, 0, 0, 0}};
    for (uint64_t i = 0; i < n_terms; i++) {
        const uint64_t a = 2 * i, b = 2 * (n_terms + i);
        qm31 f[2];
        for (int t = 0; t < 2; t++) {  /* t = 0 and t = 2 */
            qm31 num[2], den[2];
            for (int e = 0; e < 2; e++) {
                qm31 x0, x1, y0 = {{0, 0, 0, 0}}, y1 = {{0, 0, 0, 0}};
                if (kind == 2) {
                    x0 = (qm31){{in_num[a + e], 0, 0, 0}};
                    x1 = (qm31){{in_num[b + e], 0, 0, 0}};
                } else if (kind == 3) {
                    x0 = x1 = (qm31){{1, 0, 0, 0}};
                } else {
                    x0 = q_load(in_num + 4 * (a + e));
                    x1 = q_load(in_num + 4 * (b + e));
                }
                if (kind != 0) {
                    y0 = q_load(in_den + 4 * (a + e));
                    y1 = q_load(in_den + 4 * (b + e));
                }
                num[e] = t == 0 ? x0 : q_sub(q_add(x1, x1), x0);
                den[e] = t == 0 ? y0 : q_sub(q_add(y1, y1), y0);
            }
            if (kind == 0) {
                f[t] = q_mul(num[0], num[1]);
            } else {
                const qm31 fn = q_add(q_mul(num[0], den[1]), q_mul(num[1], den[0]));
                f[t] = q_add(fn, q_mul(lam, q_mul(den[0], den[1])));
            }
        }
        const qm31 e = q_load(eq + 4 * i);
        s0 = q_add(s0, q_mul(e, f[0]));
        s2 = q_add(s2, q_mul(e, f[1]));
    }
    q_put(out, s0);
    q_put(out + 4, s2);
}
