/*
 * CPU ORACLE — test infrastructure only.
 *
 * A plain-C restatement of ncsim's dense `StateVector<C64>` kernels
 * (src/state/vector.rs), used to check the HIP library and as the timed CPU
 * baseline.  Only the test suite, the smoke check and the benchmark's CPU
 * baseline leg load this file's library; the product never does.
 *
 * Each function names the reference lines it follows.  The arithmetic is the
 * reference's: complex products as (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re),
 * exact-zero skips, the fiber engine with BLOCK = 4096, the pair engine, and
 * the pairwise summation tree of `par_sum`.  Build with -ffp-contract=off so
 * no product is fused into an FMA (Rust does not contract either).
 *
 * Parallelism: OpenMP over the same independent slices rayon uses; results
 * do not depend on the thread count (the summation tree is fixed by the
 * index range, as in the reference).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    double re, im;
} cplx;

static inline cplx cmul(cplx a, cplx b) {
    cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline cplx cadd(cplx a, cplx b) {
    cplx r = {a.re + b.re, a.im + b.im};
    return r;
}
static inline cplx cconj(cplx a) {
    cplx r = {a.re, -a.im};
    return r;
}
static inline int czero(cplx a) { return a.re == 0.0 && a.im == 0.0; }

/* S::i_pow(k) for C64 (scalar.rs:54-61): one, i, one.negated(), i.negated(). */
static const cplx IPOW[4] = {{1.0, 0.0}, {0.0, 1.0}, {-1.0, -0.0}, {-0.0, -1.0}};

#define BLOCK 4096u /* vector.rs:192 */

/* ------------------------------------------------------------------------ */
/* fiber engine: vector.rs:194-301                                          */
/* ------------------------------------------------------------------------ */

enum { F_MATRIX = 0, F_DIAG = 1, F_PERM = 2 };

typedef struct {
    int mode;
    int k;
    /* F_MATRIX: the nonzero entries of each row (vector.rs:72-79), CSR */
    const int *nz_off, *nz_col;
    const cplx *nz_val;
    /* F_DIAG: the local states with a non-unit phase (vector.rs:115-120) */
    int nph;
    const int *ph_x;
    const cplx *ph_c;
    /* F_PERM: images y[x], phases c[x], unit flags (vector.rs:106-121) */
    const int *img_y;
    const cplx *img_c;
    const unsigned char *unit;
} fiber_op;

/* The closures handed to for_each_fiber by apply_matrix (vector.rs:80-91)
 * and apply_monomial (123-136). */
static void fiber_apply(const fiber_op *op, cplx *v, cplx *w) {
    const int gd = 1 << op->k;
    switch (op->mode) {
    case F_MATRIX:
        for (int a = 0; a < gd; a++) {
            cplx acc = {0.0, 0.0};
            for (int e = op->nz_off[a]; e < op->nz_off[a + 1]; e++) {
                int b = op->nz_col[e];
                if (!czero(v[b])) acc = cadd(acc, cmul(op->nz_val[e], v[b]));
            }
            w[a] = acc;
        }
        memcpy(v, w, sizeof(cplx) * gd);
        break;
    case F_DIAG:
        for (int e = 0; e < op->nph; e++) {
            int x = op->ph_x[e];
            if (!czero(v[x])) v[x] = cmul(op->ph_c[e], v[x]);
        }
        break;
    default:
        for (int x = 0; x < gd; x++)
            w[op->img_y[x]] = op->unit[x] ? v[x] : cmul(op->img_c[x], v[x]);
        memcpy(v, w, sizeof(cplx) * gd);
        break;
    }
}

/* for_each_fiber / split_fibers / fiber_leaf (vector.rs:201-301).  The
 * recursive split over the "high" gate qubits (those with 2<<q > BLOCK)
 * yields leaves made of 2^{#high} sub-slices of length S; each leaf walks its
 * sub-slices in BLOCK-sized chunks with the low gate qubits as patterns.  The
 * leaves are independent, as the rayon tasks are. */
static void for_each_fiber(cplx *amp, int n, const int *qubits, int k, const fiber_op *op) {
    int hq[64], hj[64], lq[64], lj[64];
    int nh = 0, nl = 0;
    /* order: qubits by decreasing index (vector.rs:206-208) */
    int idx[64];
    for (int j = 0; j < k; j++) idx[j] = j;
    for (int a = 0; a < k; a++)
        for (int b = a + 1; b < k; b++)
            if (qubits[idx[b]] > qubits[idx[a]]) {
                int t = idx[a];
                idx[a] = idx[b];
                idx[b] = t;
            }
    for (int e = 0; e < k; e++) {
        int q = qubits[idx[e]];
        if ((2ull << q) > BLOCK && nl == 0) {
            hq[nh] = q;
            hj[nh] = idx[e];
            nh++;
        } else {
            lq[nl] = q;
            lj[nl] = idx[e];
            nl++;
        }
    }
    const uint64_t len = 1ull << n;
    const uint64_t sub = nh ? (1ull << hq[nh - 1]) : len;     /* sub-slice length */
    const uint64_t S = sub < BLOCK ? sub : BLOCK;              /* leaf chunk */
    uint64_t himask = 0;
    for (int e = 0; e < nh; e++) himask |= 1ull << hq[e];
    /* leaves: every index with bits >= log2(S) outside the high qubits */
    int lgS = 0;
    while ((1ull << lgS) < S) lgS++;
    const int free_bits = n - lgS - nh;
    const uint64_t nleaves = 1ull << free_bits;
    uint64_t lowmask = 0;
    for (int e = 0; e < nl; e++) lowmask |= 1ull << lq[e];
    const int npat = 1 << nl, nsl = 1 << nh;

#pragma omp parallel
    {
        cplx *v = (cplx *)malloc(sizeof(cplx) << k);
        cplx *w = (cplx *)malloc(sizeof(cplx) << k);
        uint64_t *poff = (uint64_t *)malloc(sizeof(uint64_t) * npat);
        int *ploc = (int *)malloc(sizeof(int) * npat);
        uint64_t *soff = (uint64_t *)malloc(sizeof(uint64_t) * nsl);
        int *sloc = (int *)malloc(sizeof(int) * nsl);
        /* offsets and local bits of the low patterns (vector.rs:271-283) */
        for (int p = 0; p < npat; p++) {
            uint64_t off = 0;
            int loc = 0;
            for (int i = 0; i < nl; i++)
                if ((p >> i) & 1) {
                    off |= 1ull << lq[i];
                    loc |= 1 << lj[i];
                }
            poff[p] = off;
            ploc[p] = loc;
        }
        /* the sub-slices of a leaf: lower / upper halves at each high qubit
         * (vector.rs:238-248) */
        for (int s = 0; s < nsl; s++) {
            uint64_t off = 0;
            int loc = 0;
            for (int i = 0; i < nh; i++)
                if ((s >> i) & 1) {
                    off |= 1ull << hq[i];
                    loc |= 1 << hj[i];
                }
            soff[s] = off;
            sloc[s] = loc;
        }
#pragma omp for schedule(static)
        for (int64_t L = 0; L < (int64_t)nleaves; L++) {
            /* deposit the leaf number into the bits >= lgS that are not high */
            uint64_t base = 0, rest = (uint64_t)L;
            for (int b = lgS; b < n && rest; b++) {
                if (himask >> b & 1) continue;
                base |= (rest & 1) << b;
                rest >>= 1;
            }
            uint64_t i = 0;
            while (i < S) {
                for (int s = 0; s < nsl; s++)
                    for (int p = 0; p < npat; p++) v[sloc[s] | ploc[p]] = amp[base + soff[s] + (i | poff[p])];
                fiber_apply(op, v, w);
                for (int s = 0; s < nsl; s++)
                    for (int p = 0; p < npat; p++) amp[base + soff[s] + (i | poff[p])] = v[sloc[s] | ploc[p]];
                i = ((i | lowmask) + 1) & ~lowmask; /* vector.rs:299 */
            }
        }
        free(v);
        free(w);
        free(poff);
        free(ploc);
        free(soff);
        free(sloc);
    }
}

/* apply_matrix (vector.rs:67-93): m is 2^k x 2^k row-major, interleaved. */
void orc_apply_matrix(double *amp_, int n, const double *m_, const int *qubits, int k) {
    cplx *amp = (cplx *)amp_;
    const cplx *m = (const cplx *)m_;
    const int gd = 1 << k;
    int *off = (int *)malloc(sizeof(int) * (gd + 1));
    int *col = (int *)malloc(sizeof(int) * gd * gd);
    cplx *val = (cplx *)malloc(sizeof(cplx) * gd * gd);
    int e = 0;
    for (int a = 0; a < gd; a++) {
        off[a] = e;
        for (int b = 0; b < gd; b++)
            if (!czero(m[a * gd + b])) {
                col[e] = b;
                val[e] = m[a * gd + b];
                e++;
            }
    }
    off[gd] = e;
    fiber_op op = {F_MATRIX, k, off, col, val, 0, NULL, NULL, NULL, NULL, NULL};
    for_each_fiber(amp, n, qubits, k, &op);
    free(off);
    free(col);
    free(val);
}

/* apply_monomial (vector.rs:96-139) given the images of the local basis
 * states: U|x> = c[x] |y[x]>.  `c` is interleaved; the diagonal case
 * touches only the non-unit phases. */
void orc_apply_monomial(double *amp_, int n, const int *qubits, int k, const int *y, const double *c_) {
    cplx *amp = (cplx *)amp_;
    const cplx *c = (const cplx *)c_;
    const int gd = 1 << k;
    int diagonal = 1;
    for (int x = 0; x < gd; x++)
        if (y[x] != x) diagonal = 0;
    unsigned char *unit = (unsigned char *)malloc(gd);
    int *phx = (int *)malloc(sizeof(int) * gd);
    cplx *phc = (cplx *)malloc(sizeof(cplx) * gd);
    int nph = 0;
    for (int x = 0; x < gd; x++) {
        unit[x] = (c[x].re == 1.0 && c[x].im == 0.0);
        if (!unit[x]) {
            phx[nph] = x;
            phc[nph] = c[x];
            nph++;
        }
    }
    fiber_op op = {diagonal ? F_DIAG : F_PERM, k, NULL, NULL, NULL, nph, phx, phc, y, c, unit};
    for_each_fiber(amp, n, qubits, k, &op);
    free(unit);
    free(phx);
    free(phc);
}

/* ------------------------------------------------------------------------ */
/* pair engine: vector.rs:303-331                                           */
/* ------------------------------------------------------------------------ */

static inline int k_of(uint64_t c, uint64_t zm, int ph) {
    return (ph + 2 * __builtin_popcountll(zm & c)) & 3; /* vector.rs:147 */
}

enum { P_PAULI = 0, P_PROJECT = 1 };

static void for_each_pair(cplx *amp, int n, uint64_t x, uint64_t zm, int ph, int mode, const cplx *sgn) {
    const int h = 63 - __builtin_clzll(x);
    const uint64_t half = 1ull << h;
    const uint64_t xl = x & (half - 1);
    uint64_t span = 1;
    if (xl) {
        span = 1;
        while (span < xl + 1) span <<= 1;
    }
    uint64_t b = BLOCK < half ? BLOCK : half;
    if (span > b) b = span;
    const uint64_t nchunks = (1ull << n) / (2 * half);
    const uint64_t nblk = half / b;
    const cplx hf = {0.5, 0.0};
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t ci = 0; ci < (int64_t)nchunks; ci++)
        for (int64_t bi = 0; bi < (int64_t)nblk; bi++) {
            const uint64_t base = (uint64_t)ci * 2 * half + (uint64_t)bi * b;
            cplx *lb = amp + base;
            cplx *hb = amp + base + half;
            for (uint64_t i = 0; i < b; i++) {
                const uint64_t c = base + i;
                cplx *pa = &lb[i], *pb = &hb[i ^ xl];
                cplx a = *pa, bb = *pb;
                if (mode == P_PAULI) {
                    /* vector.rs:156-162 */
                    cplx na = cmul(IPOW[k_of(c ^ x, zm, ph)], bb);
                    cplx nb = cmul(IPOW[k_of(c, zm, ph)], a);
                    *pa = na;
                    *pb = nb;
                } else {
                    /* vector.rs:470-476 */
                    const uint64_t d = c ^ x;
                    cplx na = cmul(cadd(a, cmul(sgn[k_of(d, zm, ph)], bb)), hf);
                    cplx nb = cmul(cadd(bb, cmul(sgn[k_of(c, zm, ph)], a)), hf);
                    *pa = na;
                    *pb = nb;
                }
            }
        }
}

/* apply_pauli (vector.rs:143-164): psi <- i^ph X^x Z^z psi. */
void orc_apply_pauli(double *amp_, int n, uint64_t x, uint64_t z, int ph) {
    cplx *amp = (cplx *)amp_;
    if (x == 0) {
        const int64_t len = (int64_t)1 << n;
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < len; c++) {
            int kk = k_of((uint64_t)c, z, ph);
            if (kk != 0 && !czero(amp[c])) amp[c] = cmul(IPOW[kk], amp[c]);
        }
    } else {
        for_each_pair(amp, n, x, z, ph, P_PAULI, NULL);
    }
}

/* ------------------------------------------------------------------------ */
/* pairwise sums: par_sum (vector.rs:334-350)                               */
/* ------------------------------------------------------------------------ */

enum { S_NORM = 0, S_EXPECT = 1, S_INNER = 2 };

typedef struct {
    int mode;
    const cplx *a, *b;
    uint64_t xm, zm;
    int ph;
} sum_term;

static inline cplx term(const sum_term *t, uint64_t i) {
    const cplx z = {0.0, 0.0};
    switch (t->mode) {
    case S_NORM: /* norm_sq (vector.rs:397-406), abs_sq = a * conj(a) (scalar.rs:71-73) */
        return czero(t->a[i]) ? z : cmul(t->a[i], cconj(t->a[i]));
    case S_EXPECT: { /* expectation (vector.rs:412-429) */
        const cplx ac = t->a[i], ax = t->a[i ^ t->xm];
        if (czero(ac) || czero(ax)) return z;
        int kk = k_of(i, t->zm, t->ph);
        return cmul(cconj(ax), cmul(IPOW[kk], ac));
    }
    default: /* inner (vector.rs:431-443) */
        if (czero(t->a[i]) || czero(t->b[i])) return z;
        return cmul(cconj(t->a[i]), t->b[i]);
    }
}

static cplx pair_sum_seq(const sum_term *t, uint64_t lo, uint64_t hi) {
    const uint64_t len = hi - lo;
    if (len <= 8) {
        cplx acc = {0.0, 0.0};
        for (uint64_t i = lo; i < hi; i++) acc = cadd(acc, term(t, i));
        return acc;
    }
    const uint64_t mid = lo + len / 2;
    cplx a = pair_sum_seq(t, lo, mid);
    cplx b = pair_sum_seq(t, mid, hi);
    return cadd(a, b);
}

/* The same tree as par_sum: for a power-of-two range the top levels split
 * it into 2^d equal blocks, summed in parallel and then combined in the tree's
 * order, so the association is the reference's whatever the thread count. */
static cplx par_sum(const sum_term *t, uint64_t len) {
    int d = 0;
    while (d < 10 && (len >> (d + 1)) >= (1u << 14)) d++;
    if (d == 0 || (len & (len - 1))) return pair_sum_seq(t, 0, len);
    const int nb = 1 << d;
    const uint64_t bl = len >> d;
    cplx part[1024];
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < nb; i++) part[i] = pair_sum_seq(t, (uint64_t)i * bl, (uint64_t)(i + 1) * bl);
    for (int w = 1; w < nb; w <<= 1)
        for (int i = 0; i + w < nb; i += 2 * w) part[i] = cadd(part[i], part[i + w]);
    return part[0];
}

void orc_norm_sq(const double *amp, int n, double *out) {
    sum_term t = {S_NORM, (const cplx *)amp, NULL, 0, 0, 0};
    cplx r = par_sum(&t, 1ull << n);
    out[0] = r.re;
    out[1] = r.im;
}

void orc_expectation(const double *amp, int n, uint64_t x, uint64_t z, int ph, double *out) {
    sum_term t = {S_EXPECT, (const cplx *)amp, NULL, x, z, ph & 3};
    cplx r = par_sum(&t, 1ull << n);
    out[0] = r.re;
    out[1] = r.im;
}

void orc_inner(const double *a, const double *b, int n, double *out) {
    sum_term t = {S_INNER, (const cplx *)a, (const cplx *)b, 0, 0, 0};
    cplx r = par_sum(&t, 1ull << n);
    out[0] = r.re;
    out[1] = r.im;
}

/* project (vector.rs:445-479) for a Hermitian Pauli (checked by the caller);
 * returns the weight norm_sq() in out. */
void orc_project(double *amp_, int n, uint64_t x, uint64_t z, int ph, int outcome, double *out) {
    cplx *amp = (cplx *)amp_;
    cplx sgn[4];
    for (int kk = 0; kk < 4; kk++) {
        cplx c = IPOW[kk];
        if (outcome) { /* negated() */
            c.re = -c.re;
            c.im = -c.im;
        }
        sgn[kk] = c;
    }
    const cplx hf = {0.5, 0.0};
    if (x == 0) {
        const int64_t len = (int64_t)1 << n;
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < len; c++) {
            cplx a = amp[c];
            if (!czero(a)) amp[c] = cmul(cadd(a, cmul(sgn[k_of((uint64_t)c, z, ph)], a)), hf);
        }
    } else {
        for_each_pair(amp, n, x, z, ph, P_PROJECT, sgn);
    }
    orc_norm_sq(amp_, n, out);
}

/* scale (vector.rs:481-488). */
void orc_scale(double *amp_, int n, double cre, double cim) {
    cplx *amp = (cplx *)amp_;
    const cplx c = {cre, cim};
    const int64_t len = (int64_t)1 << n;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < len; i++)
        if (!czero(amp[i])) amp[i] = cmul(amp[i], c);
}

/* kron (vector.rs:168-188): out[xa + 2^na xb] = a[xa] b[xb]; out zeroed here. */
void orc_kron(const double *a_, int na, const double *b_, int nb, double *out_) {
    const cplx *a = (const cplx *)a_, *b = (const cplx *)b_;
    cplx *out = (cplx *)out_;
    const int64_t la = (int64_t)1 << na, lb = (int64_t)1 << nb;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < lb; j++) {
        cplx *ch = out + j * la;
        for (int64_t i = 0; i < la; i++) {
            ch[i].re = 0.0;
            ch[i].im = 0.0;
        }
        if (czero(b[j])) continue;
        for (int64_t i = 0; i < la; i++)
            if (!czero(a[i])) ch[i] = cmul(a[i], b[j]);
    }
}

static inline double abs_sq_f64(cplx a) { return a.re * a.re + a.im * a.im; }

typedef struct {
    double u;
    int64_t k;
} shot_key;

static int cmp_shot(const void *p, const void *q) {
    const shot_key *a = (const shot_key *)p, *b = (const shot_key *)q;
    if (a->u < b->u) return -1;
    if (a->u > b->u) return 1;
    return (a->k > b->k) - (a->k < b->k);
}

/* sample (vector.rs:494-540): `uniforms[k]` are the shots' draws in [0, 1)
 * (rng.gen::<f64>() in the reference, in order).  Returns 0, or 6 (ZeroState)
 * when the total weight is not positive. */
int orc_sample(const double *amp_, int n, const double *uniforms, int64_t shots, uint64_t *out) {
    const cplx *a = (const cplx *)amp_;
    const uint64_t len = 1ull << n;
    const uint64_t bl = BLOCK < len ? BLOCK : len;
    const int64_t nblk = (int64_t)(len / bl);
    double *wts = (double *)malloc(sizeof(double) * nblk);
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nblk; b++) {
        double s = 0.0;
        for (uint64_t i = 0; i < bl; i++) s += abs_sq_f64(a[(uint64_t)b * bl + i]);
        wts[b] = s;
    }
    double total = 0.0;
    for (int64_t b = 0; b < nblk; b++) total += wts[b];
    if (!(total > 0.0)) {
        free(wts);
        return 6;
    }
    shot_key *us = (shot_key *)malloc(sizeof(shot_key) * (shots ? shots : 1));
    for (int64_t k = 0; k < shots; k++) {
        us[k].u = uniforms[k] * total;
        us[k].k = k;
    }
    qsort(us, shots, sizeof(shot_key), cmp_shot);
    int64_t next = 0;
    double acc = 0.0;
    for (int64_t b = 0; b < nblk; b++) {
        if (next == shots) break;
        const double hi = acc + wts[b];
        const int last_block = (b + 1 == nblk);
        if (us[next].u < hi || last_block) {
            double run = acc;
            const uint64_t start = (uint64_t)b * bl;
            const uint64_t end = start + bl < len ? start + bl : len;
            uint64_t idx = start;
            while (next < shots && (us[next].u < hi || last_block)) {
                const double u = us[next].u;
                while (idx + 1 < end && run + abs_sq_f64(a[idx]) <= u) {
                    run += abs_sq_f64(a[idx]);
                    idx++;
                }
                uint64_t pick = idx;
                while (czero(a[pick]) && pick > start) pick--;
                out[us[next].k] = pick;
                next++;
            }
        }
        acc = hi;
    }
    free(us);
    free(wts);
    return 0;
}

/* Thread count of the OpenMP loops (the CPU baseline times 1 thread and all). */
void orc_set_num_threads(int t) {
#ifdef _OPENMP
    omp_set_num_threads(t > 0 ? t : 1);
#else
    (void)t;
#endif
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
