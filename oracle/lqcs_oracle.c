/*
 * lqcs_oracle.c -- CPU restatement of the reference statevector path.
 * TEST INFRASTRUCTURE ONLY (see lqcs_oracle.h for who may load it and how
 * each part is pinned). Citations are reference paths relative to its root.
 *
 * Arithmetic mirrors std::complex<double> as compiled by g++ without FMA:
 * (a+ib)(c+id) = (ac-bd) + i(ad+bc); complex*double scales both parts;
 * norm = re*re + im*im. Build with -ffp-contract=off.
 */
#include "lqcs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PAR_THRESHOLD (1ull << 14) /* src/backend/kernels/gate_kernels.cpp:16 */

/* ------------------------------------------------------------------ */
/* complex helpers                                                     */
/* ------------------------------------------------------------------ */
static inline orc_c cc(double re, double im) { orc_c z = {re, im}; return z; }
static inline orc_c cmul(orc_c a, orc_c b) {
    return cc(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline orc_c cadd(orc_c a, orc_c b) { return cc(a.re + b.re, a.im + b.im); }
static inline orc_c csub(orc_c a, orc_c b) { return cc(a.re - b.re, a.im - b.im); }
static inline orc_c cscale(orc_c a, double s) { return cc(a.re * s, a.im * s); }
static inline orc_c cconj(orc_c a) { return cc(a.re, -a.im); }
static inline double cnorm(orc_c a) { return a.re * a.re + a.im * a.im; }
static inline double cabs_(orc_c a) { return hypot(a.re, a.im); }
static inline orc_c expi(double t) { return cc(cos(t), sin(t)); }

/* include/lqcs/core/bit_utils.hpp:9-16 */
static inline uint64_t insert_zero_bit(uint64_t v, unsigned pos) {
    const uint64_t low = (1ull << pos) - 1;
    return ((v & ~low) << 1) | (v & low);
}
static inline int test_bit(uint64_t v, unsigned pos) { return (int)((v >> pos) & 1ull); }

/* ------------------------------------------------------------------ */
/* RNG: include/lqcs/core/random.hpp:10-23 = std::mt19937_64 feeding    */
/* uniform_real_distribution<double>(0,1); libstdc++ 11 maps one 64-bit */
/* word w to double(w) / 2^64, clamped below 1 (bits/random.tcc:3348). */
/* mt19937_64 parameters are the ones [rand.predef] fixes.              */
/* ------------------------------------------------------------------ */
void orc_rng_seed(orc_rng* r, uint64_t seed) {
    r->mt[0] = seed;
    for (uint32_t i = 1; i < 312; ++i) {
        r->mt[i] = 6364136223846793005ull * (r->mt[i - 1] ^ (r->mt[i - 1] >> 62)) + i;
    }
    r->idx = 312;
}

uint64_t orc_rng_next(orc_rng* r) {
    const uint64_t upper = 0xFFFFFFFF80000000ull, lower = 0x7FFFFFFFull;
    if (r->idx >= 312) {
        for (uint32_t k = 0; k < 312; ++k) {
            const uint64_t y = (r->mt[k] & upper) | (r->mt[(k + 1) % 312] & lower);
            uint64_t next = r->mt[(k + 156) % 312] ^ (y >> 1);
            if (y & 1ull) next ^= 0xB5026F5AA96619E9ull;
            r->mt[k] = next;
        }
        r->idx = 0;
    }
    uint64_t z = r->mt[r->idx++];
    z ^= (z >> 29) & 0x5555555555555555ull;
    z ^= (z << 17) & 0x71D67FFFEDA60000ull;
    z ^= (z << 37) & 0xFFF7EEE000000000ull;
    z ^= z >> 43;
    return z;
}

double orc_rng_uniform(orc_rng* r) {
    const double v = (double)orc_rng_next(r) / 18446744073709551616.0;
    return v >= 1.0 ? nextafter(1.0, 0.0) : v;
}

void orc_uniform_stream(uint64_t seed, uint64_t n, double* out) {
    orc_rng r;
    orc_rng_seed(&r, seed);
    for (uint64_t i = 0; i < n; ++i) out[i] = orc_rng_uniform(&r);
}

/* ------------------------------------------------------------------ */
/* gate model: src/circuit/gate.cpp                                    */
/* ------------------------------------------------------------------ */
static uint32_t ctz64(uint64_t v) { return v ? (uint32_t)__builtin_ctzll(v) : 64u; }

/* gate.cpp:27-49 */
uint32_t orc_base_qubits(const orc_inst* g) {
    switch (g->type) {
        case ORC_SWAP: case ORC_ISWAP: case ORC_RXX: case ORC_RYY: case ORC_RZZ:
            return 2;
        case ORC_UNITARY: {
            const uint64_t dim = (uint64_t)sqrt((double)g->mat_len);
            return ctz64(dim);
        }
        case ORC_PERMUTATION: return ctz64(g->perm_len);
        case ORC_PAULI_ROTATION: return g->pauli_len;
        case ORC_BARRIER: return 0;
        default: return 1;
    }
}

static int is_unitary(const orc_inst* g) {  /* gate.cpp:70-73 */
    return g->type != ORC_MEASURE && g->type != ORC_RESET && g->type != ORC_BARRIER;
}

typedef struct { uint64_t x, zy; uint32_t n_y; } pauli_masks;
/* gate.cpp:51-68: local bit j <-> pauli[k-1-j] */
static pauli_masks masks_of(const char* p, uint32_t k) {
    pauli_masks m = {0, 0, 0};
    for (uint32_t j = 0; j < k; ++j) {
        const char ch = p[k - 1 - j];
        const uint64_t bit = 1ull << j;
        if (ch == 'X') m.x |= bit;
        else if (ch == 'Y') { m.x |= bit; m.zy |= bit; m.n_y++; }
        else if (ch == 'Z') m.zy |= bit;
    }
    return m;
}

/* gate.cpp:75-92 */
int orc_is_diagonal(const orc_inst* g) {
    switch (g->type) {
        case ORC_I: case ORC_Z: case ORC_S: case ORC_SDG: case ORC_T: case ORC_TDG:
        case ORC_RZ: case ORC_P: case ORC_RZZ:
            return 1;
        case ORC_PAULI_ROTATION: return masks_of(g->pauli, g->pauli_len).x == 0;
        default: return 0;
    }
}

static const orc_c kIPow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};

/* gate.cpp:94-201. The entries are written with the same expressions
 * (e.g. -i*sin(h) as (-i)*sin(h)) so the doubles agree bit for bit. */
uint64_t orc_base_matrix(const orc_inst* g, orc_c* o) {
    const orc_c I = cc(0.0, 1.0), mI = cc(-0.0, -1.0), one = cc(1, 0), zero = cc(0, 0);
    const double r2 = 1.4142135623730951 / 2.0; /* std::numbers::sqrt2 / 2 */
    const double pi = 3.141592653589793;
    switch (g->type) {
        case ORC_I: o[0] = one; o[1] = zero; o[2] = zero; o[3] = one; return 2;
        case ORC_X: o[0] = zero; o[1] = one; o[2] = one; o[3] = zero; return 2;
        case ORC_Y: o[0] = zero; o[1] = mI; o[2] = I; o[3] = zero; return 2;
        case ORC_Z: o[0] = one; o[1] = zero; o[2] = zero; o[3] = cc(-1, 0); return 2;
        case ORC_H: o[0] = cc(r2, 0); o[1] = cc(r2, 0); o[2] = cc(r2, 0); o[3] = cc(-r2, 0); return 2;
        case ORC_S: o[0] = one; o[1] = zero; o[2] = zero; o[3] = I; return 2;
        case ORC_SDG: o[0] = one; o[1] = zero; o[2] = zero; o[3] = mI; return 2;
        case ORC_T: o[0] = one; o[1] = zero; o[2] = zero; o[3] = expi(pi / 4); return 2;
        case ORC_TDG: o[0] = one; o[1] = zero; o[2] = zero; o[3] = expi(-pi / 4); return 2;
        case ORC_SX:
            o[0] = cc(0.5, 0.5); o[1] = cc(0.5, -0.5); o[2] = cc(0.5, -0.5); o[3] = cc(0.5, 0.5);
            return 2;
        case ORC_RX: {
            const double h = g->params[0] / 2;
            o[0] = cc(cos(h), 0); o[1] = cscale(mI, sin(h)); o[2] = cscale(mI, sin(h));
            o[3] = cc(cos(h), 0);
            return 2;
        }
        case ORC_RY: {
            const double h = g->params[0] / 2;
            o[0] = cc(cos(h), 0); o[1] = cc(-sin(h), 0); o[2] = cc(sin(h), 0); o[3] = cc(cos(h), 0);
            return 2;
        }
        case ORC_RZ: {
            const double h = g->params[0] / 2;
            o[0] = expi(-h); o[1] = zero; o[2] = zero; o[3] = expi(h);
            return 2;
        }
        case ORC_P: o[0] = one; o[1] = zero; o[2] = zero; o[3] = expi(g->params[0]); return 2;
        case ORC_U: {
            const double h = g->params[0] / 2, phi = g->params[1], lam = g->params[2];
            const orc_c el = expi(lam);
            o[0] = cc(cos(h), 0);
            o[1] = cscale(cc(-el.re, -el.im), sin(h));
            o[2] = cscale(expi(phi), sin(h));
            o[3] = cscale(expi(phi + lam), cos(h));
            return 2;
        }
        case ORC_SWAP: case ORC_ISWAP: {
            for (int i = 0; i < 16; ++i) o[i] = zero;
            const orc_c off = g->type == ORC_SWAP ? one : I;
            o[0] = one; o[6] = off; o[9] = off; o[15] = one;
            return 4;
        }
        case ORC_RXX: case ORC_RYY: {
            const double h = g->params[0] / 2;
            const orc_c c = cc(cos(h), 0), is = cscale(I, sin(h)), mis = cc(-is.re, -is.im);
            for (int i = 0; i < 16; ++i) o[i] = zero;
            o[0] = c; o[5] = c; o[10] = c; o[15] = c;
            o[6] = mis; o[9] = mis;
            o[3] = g->type == ORC_RXX ? mis : is;
            o[12] = g->type == ORC_RXX ? mis : is;
            return 4;
        }
        case ORC_RZZ: {
            const double h = g->params[0] / 2;
            for (int i = 0; i < 16; ++i) o[i] = zero;
            o[0] = expi(-h); o[5] = expi(h); o[10] = expi(h); o[15] = expi(-h);
            return 4;
        }
        case ORC_PAULI_ROTATION: {  /* gate.cpp:166-188 */
            const uint32_t k = g->pauli_len;
            const uint64_t dim = 1ull << k;
            const pauli_masks m = masks_of(g->pauli, k);
            const double h = g->params[0] / 2;
            const orc_c c = cc(cos(h), 0), mis = cc(0.0, -sin(h));
            const orc_c yp = kIPow[m.n_y % 4];
            for (uint64_t i = 0; i < dim * dim; ++i) o[i] = zero;
            for (uint64_t col = 0; col < dim; ++col) {
                o[col * dim + col] = cadd(o[col * dim + col], c);
                const uint64_t row = col ^ m.x;
                const double sign = (__builtin_popcountll(col & m.zy) % 2 == 0) ? 1.0 : -1.0;
                o[row * dim + col] = cadd(o[row * dim + col], cscale(cmul(mis, yp), sign));
            }
            return dim;
        }
        case ORC_UNITARY: {
            const uint64_t dim = 1ull << orc_base_qubits(g);
            memcpy(o, g->mat, sizeof(orc_c) * dim * dim);
            return dim;
        }
        case ORC_PERMUTATION: {
            const uint64_t dim = g->perm_len;
            for (uint64_t i = 0; i < dim * dim; ++i) o[i] = zero;
            for (uint64_t l = 0; l < dim; ++l) o[g->perm[l] * dim + l] = one;
            return dim;
        }
        default: return 0;
    }
}

/* ------------------------------------------------------------------ */
/* kernels: src/backend/kernels/gate_kernels.cpp                        */
/* ------------------------------------------------------------------ */
/* :26-39 */
void orc_apply_single_qubit(orc_c* s, uint64_t n, uint32_t t, const orc_c m[4]) {
    const orc_c m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3];
    const int64_t groups = (int64_t)(n / 2);
#pragma omp parallel for schedule(static) if (n >= PAR_THRESHOLD)
    for (int64_t g = 0; g < groups; ++g) {
        const uint64_t i0 = insert_zero_bit((uint64_t)g, t), i1 = i0 | (1ull << t);
        const orc_c a0 = s[i0], a1 = s[i1];
        s[i0] = cadd(cmul(m0, a0), cmul(m1, a1));
        s[i1] = cadd(cmul(m2, a0), cmul(m3, a1));
    }
}

/* :41-47 */
void orc_apply_diagonal_single_qubit(orc_c* s, uint64_t n, uint32_t t, orc_c d0, orc_c d1) {
#pragma omp parallel for schedule(static) if (n >= PAR_THRESHOLD)
    for (int64_t i = 0; i < (int64_t)n; ++i) s[i] = cmul(s[i], test_bit((uint64_t)i, t) ? d1 : d0);
}

/* :49-69 */
void orc_apply_controlled_single_qubit(orc_c* s, uint64_t n, uint32_t c, uint32_t t,
                                       const orc_c m[4]) {
    const uint32_t lo = c < t ? c : t, hi = c < t ? t : c;
    const orc_c m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3];
    const int64_t groups = (int64_t)(n / 4);
#pragma omp parallel for schedule(static) if (n >= PAR_THRESHOLD)
    for (int64_t g = 0; g < groups; ++g) {
        const uint64_t base = insert_zero_bit(insert_zero_bit((uint64_t)g, lo), hi);
        const uint64_t i0 = base | (1ull << c), i1 = i0 | (1ull << t);
        const orc_c a0 = s[i0], a1 = s[i1];
        s[i0] = cadd(cmul(m0, a0), cmul(m1, a1));
        s[i1] = cadd(cmul(m2, a0), cmul(m3, a1));
    }
}

/* :71-84 */
void orc_apply_swap(orc_c* s, uint64_t n, uint32_t a, uint32_t b) {
    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    const int64_t groups = (int64_t)(n / 4);
#pragma omp parallel for schedule(static) if (n >= PAR_THRESHOLD)
    for (int64_t g = 0; g < groups; ++g) {
        const uint64_t base = insert_zero_bit(insert_zero_bit((uint64_t)g, lo), hi);
        const orc_c tmp = s[base | (1ull << a)];
        s[base | (1ull << a)] = s[base | (1ull << b)];
        s[base | (1ull << b)] = tmp;
    }
}

/* :86-106 */
void orc_apply_two_qubit(orc_c* s, uint64_t n, uint32_t q0, uint32_t q1, const orc_c m[16]) {
    const uint32_t lo = q0 < q1 ? q0 : q1, hi = q0 < q1 ? q1 : q0;
    const int64_t groups = (int64_t)(n / 4);
#pragma omp parallel for schedule(static) if (n >= PAR_THRESHOLD)
    for (int64_t g = 0; g < groups; ++g) {
        const uint64_t base = insert_zero_bit(insert_zero_bit((uint64_t)g, lo), hi);
        const uint64_t idx[4] = {base, base | (1ull << q0), base | (1ull << q1),
                                 base | (1ull << q0) | (1ull << q1)};
        const orc_c a[4] = {s[idx[0]], s[idx[1]], s[idx[2]], s[idx[3]]};
        for (int r = 0; r < 4; ++r) {
            orc_c acc = cmul(m[r * 4], a[0]);
            acc = cadd(acc, cmul(m[r * 4 + 1], a[1]));
            acc = cadd(acc, cmul(m[r * 4 + 2], a[2]));
            acc = cadd(acc, cmul(m[r * 4 + 3], a[3]));
            s[idx[r]] = acc;
        }
    }
}

/* :118-139 -- fixed positions (sorted), control mask, target offsets */
typedef struct {
    uint32_t fixed[64];
    uint32_t n_fixed;
    uint64_t c_mask;
    uint64_t* offset;
    uint64_t groups;
} layout;

static void make_layout(layout* L, uint64_t n, const uint32_t* ctl, uint32_t nc,
                        const uint32_t* tgt, uint32_t nt) {
    L->n_fixed = 0;
    for (uint32_t i = 0; i < nc; ++i) L->fixed[L->n_fixed++] = ctl[i];
    for (uint32_t i = 0; i < nt; ++i) L->fixed[L->n_fixed++] = tgt[i];
    for (uint32_t i = 1; i < L->n_fixed; ++i) {  /* insertion sort ascending */
        const uint32_t v = L->fixed[i];
        uint32_t j = i;
        while (j > 0 && L->fixed[j - 1] > v) { L->fixed[j] = L->fixed[j - 1]; --j; }
        L->fixed[j] = v;
    }
    L->c_mask = 0;
    for (uint32_t i = 0; i < nc; ++i) L->c_mask |= 1ull << ctl[i];
    const uint64_t dim = 1ull << nt;
    L->offset = (uint64_t*)calloc(dim, sizeof(uint64_t));
    for (uint64_t l = 0; l < dim; ++l)
        for (uint32_t j = 0; j < nt; ++j)
            if (test_bit(l, j)) L->offset[l] |= 1ull << tgt[j];
    L->groups = n >> L->n_fixed;
}

static inline uint64_t expand(uint64_t g, const layout* L) {  /* :19-22 */
    for (uint32_t i = 0; i < L->n_fixed; ++i) g = insert_zero_bit(g, L->fixed[i]);
    return g;
}

/* :143-167 */
void orc_apply_controlled_unitary(orc_c* s, uint64_t n, const uint32_t* ctl, uint32_t nc,
                                  const uint32_t* tgt, uint32_t nt, const orc_c* m) {
    layout L;
    make_layout(&L, n, ctl, nc, tgt, nt);
    const uint64_t dim = 1ull << nt;
#pragma omp parallel if (n >= PAR_THRESHOLD)
    {
        orc_c* a = (orc_c*)malloc(sizeof(orc_c) * dim);
        uint64_t* idx = (uint64_t*)malloc(sizeof(uint64_t) * dim);
#pragma omp for schedule(static)
        for (int64_t g = 0; g < (int64_t)L.groups; ++g) {
            const uint64_t base = expand((uint64_t)g, &L) | L.c_mask;
            for (uint64_t l = 0; l < dim; ++l) { idx[l] = base | L.offset[l]; a[l] = s[idx[l]]; }
            for (uint64_t r = 0; r < dim; ++r) {
                orc_c sum = cc(0, 0);
                for (uint64_t c = 0; c < dim; ++c) sum = cadd(sum, cmul(m[r * dim + c], a[c]));
                s[idx[r]] = sum;
            }
        }
        free(a);
        free(idx);
    }
    free(L.offset);
}

/* :169-189 */
void orc_apply_controlled_permutation(orc_c* s, uint64_t n, const uint32_t* ctl, uint32_t nc,
                                      const uint32_t* tgt, uint32_t nt, const uint64_t* perm) {
    layout L;
    make_layout(&L, n, ctl, nc, tgt, nt);
    const uint64_t dim = 1ull << nt;
#pragma omp parallel if (n >= PAR_THRESHOLD)
    {
        orc_c* a = (orc_c*)malloc(sizeof(orc_c) * dim);
#pragma omp for schedule(static)
        for (int64_t g = 0; g < (int64_t)L.groups; ++g) {
            const uint64_t base = expand((uint64_t)g, &L) | L.c_mask;
            for (uint64_t l = 0; l < dim; ++l) a[l] = s[base | L.offset[l]];
            for (uint64_t l = 0; l < dim; ++l) s[base | L.offset[perm[l]]] = a[l];
        }
        free(a);
    }
    free(L.offset);
}

/* :191-225 */
void orc_apply_pauli_rotation(orc_c* s, uint64_t n, uint64_t x, uint64_t zy, uint32_t n_y,
                              double theta) {
    const double h = theta / 2;
    const orc_c c = cc(cos(h), 0), mis = cc(0.0, -sin(h));
    const orc_c yphase = cmul(mis, kIPow[n_y % 4]);
    if (x == 0) {
#pragma omp parallel for schedule(static) if (n >= PAR_THRESHOLD)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            const double sign = (__builtin_popcountll((uint64_t)i & zy) % 2 == 0) ? 1.0 : -1.0;
            s[i] = cmul(s[i], cadd(c, cscale(yphase, sign)));
        }
        return;
    }
    const uint64_t lsb = x & (~x + 1);
#pragma omp parallel for schedule(static) if (n >= PAR_THRESHOLD)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint64_t i = (uint64_t)ii;
        if (i & lsb) continue;
        const uint64_t j = i ^ x;
        const double si = (__builtin_popcountll(i & zy) % 2 == 0) ? 1.0 : -1.0;
        const double sj = (__builtin_popcountll(j & zy) % 2 == 0) ? 1.0 : -1.0;
        const orc_c ai = s[i], aj = s[j];
        s[i] = cadd(cmul(c, ai), cmul(cscale(yphase, sj), aj));
        s[j] = cadd(cmul(cscale(yphase, si), ai), cmul(c, aj));
    }
}

/* :227-247 (reduction in index order = the reference at one thread) */
int orc_collapse_qubit(orc_c* s, uint64_t n, uint32_t q, double u) {
    double p1 = 0.0;
    for (uint64_t i = 0; i < n; ++i)
        if (test_bit(i, q)) p1 += cnorm(s[i]);
    const int outcome = u < p1;
    const double keep = outcome ? p1 : 1.0 - p1;
    const double scale = 1.0 / sqrt(keep);
#pragma omp parallel for schedule(static) if (n >= PAR_THRESHOLD)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        if (test_bit((uint64_t)i, q) == outcome) s[i] = cscale(s[i], scale);
        else s[i] = cc(0, 0);
    }
    return outcome;
}

/* :249-260 */
void orc_reset_qubit(orc_c* s, uint64_t n, uint32_t q, double u) {
    if (!orc_collapse_qubit(s, n, q, u)) return;
    const uint64_t mask = 1ull << q;
    for (uint64_t g = 0; g < n / 2; ++g) {
        const uint64_t i0 = insert_zero_bit(g, q);
        s[i0] = s[i0 | mask];
        s[i0 | mask] = cc(0, 0);
    }
}

/* :262-277 */
double orc_kraus_probability(const orc_c* s, uint64_t n, uint32_t t, const orc_c k[4]) {
    const orc_c k0 = k[0], k1 = k[1], k2 = k[2], k3 = k[3];
    double p = 0.0;
    for (uint64_t g = 0; g < n / 2; ++g) {
        const uint64_t i0 = insert_zero_bit(g, t), i1 = i0 | (1ull << t);
        const orc_c a0 = s[i0], a1 = s[i1];
        p += cnorm(cadd(cmul(k0, a0), cmul(k1, a1))) + cnorm(cadd(cmul(k2, a0), cmul(k3, a1)));
    }
    return p;
}

/* ------------------------------------------------------------------ */
/* simulator: src/backend/statevector_simulator.cpp                      */
/* ------------------------------------------------------------------ */
/* :22-85 -- dispatch order: Permutation, PauliRotation (nc=0), nc=0 k=1
 * diagonal/dense, SWAP, k=2 dense, nc=1 k=1, generic controlled unitary. */
int orc_apply_instruction(orc_c* s, uint64_t n, const orc_inst* g) {
    const uint32_t nc = g->n_controls, k = orc_base_qubits(g);
    const uint32_t* ctl = g->qubits;
    const uint32_t* tgt = g->qubits + nc;
    if (!is_unitary(g)) return -1;
    if (g->type == ORC_PERMUTATION) {
        orc_apply_controlled_permutation(s, n, ctl, nc, tgt, k, g->perm);
        return 0;
    }
    if (g->type == ORC_PAULI_ROTATION && nc == 0) {
        const pauli_masks pm = masks_of(g->pauli, g->pauli_len);
        uint64_t xg = 0, zg = 0;
        for (uint32_t j = 0; j < k; ++j) {
            if (pm.x & (1ull << j)) xg |= 1ull << tgt[j];
            if (pm.zy & (1ull << j)) zg |= 1ull << tgt[j];
        }
        orc_apply_pauli_rotation(s, n, xg, zg, pm.n_y, g->params[0]);
        return 0;
    }
    const uint64_t dim = 1ull << k;
    orc_c stack_m[16];
    orc_c* m = dim <= 4 ? stack_m : (orc_c*)malloc(sizeof(orc_c) * dim * dim);
    orc_base_matrix(g, m);
    if (nc == 0 && k == 1) {
        if (orc_is_diagonal(g)) orc_apply_diagonal_single_qubit(s, n, tgt[0], m[0], m[3]);
        else orc_apply_single_qubit(s, n, tgt[0], m);
    } else if (nc == 0 && g->type == ORC_SWAP) {
        orc_apply_swap(s, n, tgt[0], tgt[1]);
    } else if (nc == 0 && k == 2) {
        orc_apply_two_qubit(s, n, tgt[0], tgt[1], m);
    } else if (nc == 1 && k == 1) {
        orc_apply_controlled_single_qubit(s, n, ctl[0], tgt[0], m);
    } else {
        orc_apply_controlled_unitary(s, n, ctl, nc, tgt, k, m);
    }
    if (m != stack_m) free(m);
    return 0;
}

/* ---- greedy single-qubit fusion: src/transpiler/gate_fusion.cpp:32-68 */
struct orc_fused {
    orc_circuit circ;
    orc_inst* inst;
    orc_c* mats;       /* 4 per fused unitary */
    uint32_t* qubits;  /* one per fused unitary */
};

static void matmul2(const orc_c* a, const orc_c* b, orc_c* out) {  /* :14-17 */
    orc_c r[4];
    r[0] = cadd(cmul(a[0], b[0]), cmul(a[1], b[2]));
    r[1] = cadd(cmul(a[0], b[1]), cmul(a[1], b[3]));
    r[2] = cadd(cmul(a[2], b[0]), cmul(a[3], b[2]));
    r[3] = cadd(cmul(a[2], b[1]), cmul(a[3], b[3]));
    memcpy(out, r, sizeof r);
}

static int is_identity2(const orc_c* m) {  /* :19-23 */
    const double tol = 1e-12;
    return cabs_(csub(m[0], cc(1, 0))) < tol && cabs_(m[1]) < tol && cabs_(m[2]) < tol &&
           cabs_(csub(m[3], cc(1, 0))) < tol;
}

static int is_fusable(const orc_inst* g) {  /* :25-28 */
    return is_unitary(g) && g->n_controls == 0 && orc_base_qubits(g) == 1;
}

orc_fused* orc_fuse(const orc_circuit* c) {
    orc_fused* f = (orc_fused*)calloc(1, sizeof(orc_fused));
    const uint32_t nq = c->n_qubits;
    f->inst = (orc_inst*)calloc(c->n_inst + nq + 1, sizeof(orc_inst));
    f->mats = (orc_c*)calloc(4 * (c->n_inst + 1), sizeof(orc_c));
    f->qubits = (uint32_t*)calloc(c->n_inst + 1, sizeof(uint32_t));
    orc_c* pend = (orc_c*)calloc(4 * (size_t)nq, sizeof(orc_c));
    int* has = (int*)calloc(nq, sizeof(int));
    uint64_t out = 0, fused = 0;

#define FLUSH(q)                                                                  \
    do {                                                                          \
        const uint32_t fq = (q);                                                  \
        if (has[fq]) {                                                            \
            if (!is_identity2(pend + 4 * fq)) {                                   \
                orc_c* dst = f->mats + 4 * fused;                                 \
                memcpy(dst, pend + 4 * fq, 4 * sizeof(orc_c));                    \
                f->qubits[fused] = fq;                                            \
                orc_inst u;                                                       \
                memset(&u, 0, sizeof u);                                          \
                u.type = ORC_UNITARY; u.n_qubits = 1; u.qubits = f->qubits + fused; \
                u.mat_len = 4; u.mat = dst;                                       \
                f->inst[out++] = u;                                               \
                ++fused;                                                          \
            }                                                                     \
            has[fq] = 0;                                                          \
        }                                                                         \
    } while (0)

    for (uint64_t i = 0; i < c->n_inst; ++i) {
        const orc_inst* g = c->inst + i;
        if (is_fusable(g)) {
            const uint32_t q = g->qubits[0];
            orc_c m[4];
            orc_base_matrix(g, m);
            if (has[q]) matmul2(m, pend + 4 * q, pend + 4 * q);  /* later gate left-multiplies */
            else { memcpy(pend + 4 * q, m, sizeof m); has[q] = 1; }
            continue;
        }
        if (g->type == ORC_BARRIER) {
            for (uint32_t q = 0; q < nq; ++q) FLUSH(q);
        } else {
            for (uint32_t j = 0; j < g->n_qubits; ++j) FLUSH(g->qubits[j]);
        }
        f->inst[out++] = *g;
    }
    for (uint32_t q = 0; q < nq; ++q) FLUSH(q);
#undef FLUSH
    free(pend);
    free(has);
    f->circ.n_qubits = c->n_qubits;
    f->circ.n_clbits = c->n_clbits;
    f->circ.n_inst = out;
    f->circ.inst = f->inst;
    return f;
}

const orc_circuit* orc_fused_circuit(const orc_fused* f) { return &f->circ; }

void orc_fused_free(orc_fused* f) {
    if (!f) return;
    free(f->inst);
    free(f->mats);
    free(f->qubits);
    free(f);
}

static void init_zero_state(orc_c* s, uint64_t n) {  /* statevector.cpp:11-25 */
    memset(s, 0, sizeof(orc_c) * n);
    s[0] = cc(1, 0);
}

/* :197-207 evolve (Measure/Barrier skipped, Reset rejected) */
static int evolve(const orc_circuit* c, orc_c* s) {
    const uint64_t n = 1ull << c->n_qubits;
    init_zero_state(s, n);
    for (uint64_t i = 0; i < c->n_inst; ++i) {
        if (c->inst[i].type == ORC_RESET) return -1;
        if (is_unitary(c->inst + i)) orc_apply_instruction(s, n, c->inst + i);
    }
    return 0;
}

/* :222-230 */
int orc_run_statevector(const orc_circuit* c, int fuse, orc_c* out) {
    if (!fuse) return evolve(c, out);
    orc_fused* f = orc_fuse(c);
    const int rc = evolve(&f->circ, out);
    orc_fused_free(f);
    return rc;
}

/* :241-257 path selection; returns 1 per-shot, 0 sampling, -2 no measure */
static int needs_per_shot(const orc_circuit* c, int noisy) {
    int seen = 0, per_shot = noisy;
    for (uint64_t i = 0; i < c->n_inst; ++i) {
        const orc_inst* g = c->inst + i;
        if (g->type == ORC_MEASURE) seen = 1;
        else if (g->type == ORC_RESET) per_shot = 1;
        else if (seen && is_unitary(g)) per_shot = 1;
    }
    if (!seen) return -2;
    return per_shot;
}

/* :88-119 -- serial inclusive prefix sum, r = u * total, upper_bound */
int orc_run_sampling(const orc_circuit* c, int fuse, uint64_t seed, uint64_t shots,
                     uint64_t* idx_out, double* cum_out) {
    orc_fused* f = fuse ? orc_fuse(c) : NULL;
    const orc_circuit* opt = f ? &f->circ : c;
    const int path = needs_per_shot(opt, 0);
    if (path != 0) { orc_fused_free(f); return path == -2 ? -2 : -1; }
    const uint64_t n = 1ull << c->n_qubits;
    orc_c* s = (orc_c*)malloc(sizeof(orc_c) * n);
    evolve(opt, s);
    orc_fused_free(f);
    double* cum = cum_out ? cum_out : (double*)malloc(sizeof(double) * n);
    for (uint64_t i = 0; i < n; ++i) cum[i] = cnorm(s[i]);
    free(s);
    for (uint64_t i = 1; i < n; ++i) cum[i] = cum[i - 1] + cum[i];
    const double total = cum[n - 1];
    orc_rng r;
    orc_rng_seed(&r, seed);
    for (uint64_t sh = 0; sh < shots; ++sh) {
        const double x = orc_rng_uniform(&r) * total;
        uint64_t lo = 0, hi = n;  /* first index with cum[i] > x */
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (x < cum[mid]) hi = mid; else lo = mid + 1;
        }
        idx_out[sh] = lo;
    }
    if (!cum_out) free(cum);
    return 0;
}

uint64_t orc_draws_per_shot(const orc_circuit* c, const orc_noise* noise) {
    const uint32_t nch = noise ? noise->n_channels : 0;
    uint64_t d = 0;
    for (uint64_t i = 0; i < c->n_inst; ++i) {
        const orc_inst* g = c->inst + i;
        if (g->type == ORC_MEASURE || g->type == ORC_RESET) d += 1;
        else if (g->type == ORC_BARRIER) continue;
        else d += (uint64_t)nch * g->n_qubits;
    }
    return d;
}

/* :122-144 */
static void sample_kraus_branch(orc_c* s, uint64_t n, const orc_channel* ch, uint32_t q,
                                orc_rng* r) {
    const double u = orc_rng_uniform(r);
    double cum = 0.0, chosen_p = 0.0;
    const orc_c* chosen = NULL;
    for (uint32_t k = 0; k < ch->n_ops; ++k) {
        const orc_c* K = ch->ops + 4 * k;
        const double p = orc_kraus_probability(s, n, q, K);
        if (p <= 0.0) continue;
        chosen = K;
        chosen_p = p;
        cum += p;
        if (u < cum) break;
    }
    if (!chosen) return;
    const double inv = 1.0 / sqrt(chosen_p);
    const orc_c m[4] = {cscale(chosen[0], inv), cscale(chosen[1], inv),
                        cscale(chosen[2], inv), cscale(chosen[3], inv)};
    orc_apply_single_qubit(s, n, q, m);
}

/* :156-190 */
int orc_run_per_shot(const orc_circuit* c, const orc_noise* noise, uint64_t seed,
                     uint64_t shots, uint8_t* clbits_out) {
    const int noisy = noise && noise->n_channels > 0;
    if (needs_per_shot(c, noisy) == -2) return -2;
    const uint64_t n = 1ull << c->n_qubits;
    orc_c* s = (orc_c*)malloc(sizeof(orc_c) * n);
    orc_rng r;
    orc_rng_seed(&r, seed);
    for (uint64_t sh = 0; sh < shots; ++sh) {
        uint8_t* creg = clbits_out + sh * c->n_clbits;
        memset(creg, 0, c->n_clbits);
        init_zero_state(s, n);
        for (uint64_t i = 0; i < c->n_inst; ++i) {
            const orc_inst* g = c->inst + i;
            if (g->type == ORC_MEASURE) {
                creg[g->clbit] = (uint8_t)orc_collapse_qubit(s, n, g->qubits[0], orc_rng_uniform(&r));
            } else if (g->type == ORC_RESET) {
                orc_reset_qubit(s, n, g->qubits[0], orc_rng_uniform(&r));
            } else if (g->type == ORC_BARRIER) {
                continue;
            } else {
                orc_apply_instruction(s, n, g);
                if (noisy)
                    for (uint32_t k = 0; k < noise->n_channels; ++k)
                        for (uint32_t j = 0; j < g->n_qubits; ++j)
                            sample_kraus_branch(s, n, noise->channels + k, g->qubits[j], &r);
            }
        }
    }
    free(s);
    return 0;
}

/* ------------------------------------------------------------------ */
/* observables: src/backend/statevector.cpp, src/algorithms/vqe.cpp      */
/* ------------------------------------------------------------------ */
double orc_norm(const orc_c* s, uint64_t n) {  /* statevector.cpp:27-31 */
    double sum = 0.0;
    for (uint64_t i = 0; i < n; ++i) sum += cnorm(s[i]);
    return sqrt(sum);
}

/* statevector.cpp:39-78: leftmost char <-> qubit n-1 */
double orc_expectation_value(const orc_c* s, uint32_t nq, const char* pauli) {
    uint64_t x = 0, zy = 0;
    uint32_t ny = 0;
    for (uint32_t j = 0; j < nq; ++j) {
        const uint64_t bit = 1ull << (nq - 1 - j);
        switch (pauli[j]) {
            case 'X': x |= bit; break;
            case 'Z': zy |= bit; break;
            case 'Y': x |= bit; zy |= bit; ++ny; break;
            default: break;
        }
    }
    const orc_c yp = kIPow[ny % 4];
    const uint64_t n = 1ull << nq;
    double sre = 0.0, sim = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        const double sign = (__builtin_popcountll(i & zy) % 2 == 0) ? 1.0 : -1.0;
        const orc_c t = cmul(cscale(s[i], sign), cconj(s[i ^ x]));
        sre += t.re;
        sim += t.im;
    }
    return cmul(yp, cc(sre, sim)).re;
}

double orc_expectation(const orc_c* s, uint32_t nq, const orc_term* h, uint64_t nt) {
    double e = 0.0;  /* vqe.cpp:11-17 */
    for (uint64_t t = 0; t < nt; ++t) e += h[t].coeff * orc_expectation_value(s, nq, h[t].pauli);
    return e;
}

/* ParametricCircuit::bind (parametric_circuit.cpp:73-96) writes values into
 * params[0] of the parameter instructions of a template copy. */
static double bound_energy(const orc_circuit* tmpl, const uint64_t* param_inst, uint64_t np,
                           const double* theta, const orc_term* h, uint64_t nt, orc_c* s) {
    orc_inst* inst = (orc_inst*)malloc(sizeof(orc_inst) * (tmpl->n_inst ? tmpl->n_inst : 1));
    double* pv = (double*)malloc(sizeof(double) * (np ? np : 1));
    memcpy(inst, tmpl->inst, sizeof(orc_inst) * tmpl->n_inst);
    for (uint64_t p = 0; p < np; ++p) {
        pv[p] = theta[p];
        inst[param_inst[p]].params = pv + p;
        inst[param_inst[p]].n_params = 1;
    }
    orc_circuit bound = {tmpl->n_qubits, tmpl->n_clbits, tmpl->n_inst, inst};
    orc_run_statevector(&bound, 1, s);  /* default Options: fuse_gates = true */
    const double e = orc_expectation(s, tmpl->n_qubits, h, nt);
    free(inst);
    free(pv);
    return e;
}

int orc_bound_energy(const orc_circuit* tmpl, const uint64_t* param_inst, uint64_t np,
                     const double* theta, const orc_term* h, uint64_t nt, double* energy) {
    orc_c* s = (orc_c*)malloc(sizeof(orc_c) << tmpl->n_qubits);
    *energy = bound_energy(tmpl, param_inst, np, theta, h, nt, s);
    free(s);
    return 0;
}

/* vqe.cpp:68-94 */
int orc_parameter_shift_gradient(const orc_circuit* tmpl, const uint64_t* param_inst,
                                 uint64_t np, const double* theta, const orc_term* h,
                                 uint64_t nt, double* grad) {
    const double half_pi = 3.141592653589793 / 2;
    double* th = (double*)malloc(sizeof(double) * (np ? np : 1));
    memcpy(th, theta, sizeof(double) * np);
    orc_c* s = (orc_c*)malloc(sizeof(orc_c) << tmpl->n_qubits);
    for (uint64_t k = 0; k < np; ++k) {
        const double saved = th[k];
        th[k] = saved + half_pi;
        const double ep = bound_energy(tmpl, param_inst, np, th, h, nt, s);
        th[k] = saved - half_pi;
        const double em = bound_energy(tmpl, param_inst, np, th, h, nt, s);
        th[k] = saved;
        grad[k] = 0.5 * (ep - em);
    }
    free(s);
    free(th);
    return 0;
}

/* ------------------------------------------------------------------ */
/* density matrix: src/backend/density_matrix.cpp:34-84,                */
/* density_matrix_simulator.cpp:13-28                                   */
/* ------------------------------------------------------------------ */
static void dm_apply_instruction(orc_c* rho, uint32_t nq, const orc_inst* g) {
    const uint64_t n = 1ull << (2 * nq);
    const uint32_t nc = g->n_controls, k = orc_base_qubits(g);
    uint32_t ket[64], bra[64];
    for (uint32_t j = 0; j < g->n_qubits; ++j) { ket[j] = g->qubits[j]; bra[j] = g->qubits[j] + nq; }
    if (g->type == ORC_PERMUTATION) {
        orc_apply_controlled_permutation(rho, n, ket, nc, ket + nc, k, g->perm);
        orc_apply_controlled_permutation(rho, n, bra, nc, bra + nc, k, g->perm);
        return;
    }
    const uint64_t dim = 1ull << k;
    orc_c* m = (orc_c*)malloc(sizeof(orc_c) * dim * dim);
    orc_c* mc = (orc_c*)malloc(sizeof(orc_c) * dim * dim);
    orc_base_matrix(g, m);
    for (uint64_t i = 0; i < dim * dim; ++i) mc[i] = cconj(m[i]);
    orc_apply_controlled_unitary(rho, n, ket, nc, ket + nc, k, m);
    orc_apply_controlled_unitary(rho, n, bra, nc, bra + nc, k, mc);
    free(m);
    free(mc);
}

static void dm_apply_channel(orc_c* rho, uint32_t nq, const orc_channel* ch, uint32_t q,
                             orc_c* acc, orc_c* work) {
    const uint64_t n = 1ull << (2 * nq);
    memset(acc, 0, sizeof(orc_c) * n);
    for (uint32_t k = 0; k < ch->n_ops; ++k) {
        const orc_c* K = ch->ops + 4 * k;
        const orc_c kc[4] = {cconj(K[0]), cconj(K[1]), cconj(K[2]), cconj(K[3])};
        memcpy(work, rho, sizeof(orc_c) * n);
        orc_apply_single_qubit(work, n, q, K);
        orc_apply_single_qubit(work, n, q + nq, kc);
        for (uint64_t i = 0; i < n; ++i) acc[i] = cadd(acc[i], work[i]);
    }
    memcpy(rho, acc, sizeof(orc_c) * n);
}

int orc_run_density_matrix(const orc_circuit* c, const orc_noise* noise, orc_c* rho) {
    const uint32_t nq = c->n_qubits;
    const uint64_t n = 1ull << (2 * nq);
    memset(rho, 0, sizeof(orc_c) * n);
    rho[0] = cc(1, 0);
    orc_c* acc = (orc_c*)malloc(sizeof(orc_c) * n);
    orc_c* work = (orc_c*)malloc(sizeof(orc_c) * n);
    int rc = 0;
    for (uint64_t i = 0; i < c->n_inst; ++i) {
        const orc_inst* g = c->inst + i;
        if (g->type == ORC_RESET) { rc = -1; break; }
        if (!is_unitary(g)) continue;
        dm_apply_instruction(rho, nq, g);
        if (noise)
            for (uint32_t k = 0; k < noise->n_channels; ++k)
                for (uint32_t j = 0; j < g->n_qubits; ++j)
                    dm_apply_channel(rho, nq, noise->channels + k, g->qubits[j], acc, work);
    }
    free(acc);
    free(work);
    return rc;
}

/* :232-265 -- StatevectorSimulator::run. Noise disables fusion (:239);
 * clbits_out receives shots * n_clbits bytes. Returns 0 when the terminal
 * sampling path ran, 1 for the per-shot path, -2 without Measure, -3 for
 * shots == 0. */
int orc_run(const orc_circuit* c, const orc_noise* noise, int fuse, uint64_t seed,
            uint64_t shots, uint8_t* clbits_out) {
    if (shots == 0) return -3;
    const int noisy = noise && noise->n_channels > 0;
    orc_fused* f = (!noisy && fuse) ? orc_fuse(c) : NULL;
    const orc_circuit* opt = f ? &f->circ : c;
    const int path = needs_per_shot(opt, noisy);
    int rc;
    if (path == -2) {
        rc = -2;
    } else if (path == 1) {
        orc_run_per_shot(opt, noisy ? noise : NULL, seed, shots, clbits_out);
        rc = 1;
    } else {
        uint64_t* idx = (uint64_t*)malloc(sizeof(uint64_t) * shots);
        orc_run_sampling(opt, 0, seed, shots, idx, NULL);
        int32_t* src = (int32_t*)malloc(sizeof(int32_t) * (c->n_clbits + 1));
        for (uint32_t k = 0; k < c->n_clbits; ++k) src[k] = -1;
        for (uint64_t i = 0; i < opt->n_inst; ++i)
            if (opt->inst[i].type == ORC_MEASURE) src[opt->inst[i].clbit] = (int32_t)opt->inst[i].qubits[0];
        for (uint64_t sh = 0; sh < shots; ++sh)
            for (uint32_t k = 0; k < c->n_clbits; ++k)
                clbits_out[sh * c->n_clbits + k] =
                    (uint8_t)(src[k] >= 0 && test_bit(idx[sh], (unsigned)src[k]));
        free(src);
        free(idx);
        rc = 0;
    }
    orc_fused_free(f);
    return rc;
}
