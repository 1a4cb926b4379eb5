/* Plain-C restatement of the reference hafnian -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library,
 * as the checker.  The product never links it.
 *
 * It follows MPS_cpu.py:117-187 (hafnian dispatch, recursive_hafnian, solve) one matrix at
 * a time, with NumPy's complex128 arithmetic written out explicitly so that the result bits
 * do not depend on the host's NumPy build:
 *   complex multiply  re = fma(ar, br, -(ai*bi)),  im = fma(ar, bi, ai*br)
 *   complex add       componentwise IEEE add
 * (NumPy 2.2 on x86-64 evaluates complex128 products in exactly this FMA form; the golden
 * vectors in tests/golden pin it.)  Build with -ffp-contract=off so that no other
 * expression is fused.
 *
 * The recursion keeps every truncated-polynomial coefficient, including the ones that are
 * structurally zero, exactly as the reference loops do.
 *
 * loop_hafnian_* extend the same recursion with loop weights (a vertex may be matched to
 * itself with weight loops[i]).  The reference computes plain hafnians only (its pure state
 * has zero mean), so this extension has no reference counterpart: it is pinned by brute-force
 * sums over matchings with loops, by closed forms (involution counts, rank-one matrices) and by
 * reducing bit for bit to the plain recursion when every loop weight is zero.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cplx;

static inline cplx cmul(cplx a, cplx b) {
    cplx r;
    r.re = fma(a.re, b.re, -(a.im * b.im));
    r.im = fma(a.re, b.im, a.im * b.re);
    return r;
}
static inline cplx cadd(cplx a, cplx b) {
    cplx r;
    r.re = a.re + b.re;
    r.im = a.im + b.im;
    return r;
}

/* solve(b, s, w, g, n), MPS_cpu.py:163-187.  b: [s(s-1)/2][n+1], g: [n+1]. */
static cplx solve(const cplx *b, int s, int w, const cplx *g, int n) {
    if (s == 0) {
        cplx wc = {(double)w, 0.0};
        return cmul(wc, g[n]); /* w * g[:, n] with w promoted to complex */
    }
    const int np1 = n + 1;
    const int nc = (s - 2) * (s - 3) / 2;
    cplx *c = (cplx *)calloc((size_t)(nc > 0 ? nc : 1) * np1, sizeof(cplx));
    cplx *e = (cplx *)malloc((size_t)np1 * sizeof(cplx));
    int i = 0;
    for (int j = 1; j < s - 2; ++j)
        for (int k = 0; k < j; ++k) {
            memcpy(&c[(size_t)i * np1], &b[(size_t)((j + 1) * (j + 2) / 2 + k + 2) * np1],
                   (size_t)np1 * sizeof(cplx));
            ++i;
        }
    cplx h = solve(c, s - 2, -w, g, n);
    memcpy(e, g, (size_t)np1 * sizeof(cplx));
    for (int u = 0; u < n; ++u)
        for (int v = 0; v < n - u; ++v)
            e[u + v + 1] = cadd(e[u + v + 1], cmul(g[u], b[0 * np1 + v]));
    for (int j = 1; j < s - 2; ++j)
        for (int k = 0; k < j; ++k)
            for (int u = 0; u < n; ++u)
                for (int v = 0; v < n - u; ++v) {
                    const cplx *bj0 = &b[(size_t)((j + 1) * (j + 2) / 2) * np1];
                    const cplx *bk1 = &b[(size_t)((k + 1) * (k + 2) / 2 + 1) * np1];
                    const cplx *bk0 = &b[(size_t)((k + 1) * (k + 2) / 2) * np1];
                    const cplx *bj1 = &b[(size_t)((j + 1) * (j + 2) / 2 + 1) * np1];
                    cplx t1 = cmul(bj0[u], bk1[v]);
                    cplx t2 = cmul(bk0[u], bj1[v]);
                    cplx t3 = cadd(t1, t2);
                    cplx *dst = &c[(size_t)(j * (j - 1) / 2 + k) * np1 + u + v + 1];
                    *dst = cadd(*dst, t3);
                }
    cplx r = solve(c, s - 2, w, e, n);
    free(c);
    free(e);
    return cadd(h, r);
}

/* hafnian of A = Sigma[t][:, t] (MPS_cpu.py:117-140 dispatch + :143-160). */
static cplx hafnian_one(const cplx *sigma, int ld, const int *t, int nsel) {
    cplx zero = {0.0, 0.0}, one = {1.0, 0.0};
    if (nsel == 0) return one;
    if (nsel % 2) return zero;
#define A(r, c) sigma[(size_t)t[r] * ld + t[c]]
    if (nsel == 2) return A(0, 1);
    if (nsel == 4)
        return cadd(cadd(cmul(A(0, 1), A(2, 3)), cmul(A(0, 2), A(1, 3))), cmul(A(0, 3), A(1, 2)));
    const int n = nsel / 2, np1 = n + 1;
    const int nz = n * (2 * n - 1);
    cplx *z = (cplx *)calloc((size_t)nz * np1, sizeof(cplx));
    for (int j = 1; j < 2 * n; ++j) {
        int ind = j * (j - 1) / 2;
        for (int k = 0; k < j; ++k) z[(size_t)(ind + k) * np1] = A(j, k);
    }
#undef A
    cplx *g = (cplx *)calloc((size_t)np1, sizeof(cplx));
    g[0] = one;
    cplx r = solve(z, 2 * n, 1, g, n);
    free(z);
    free(g);
    return r;
}

/* out[i] = hafnian(Sigma[target_i][:, target_i]) for a batch of targets [n_batch][n_select].
 * sigma, out: interleaved complex128. */
void oracle_hafnian_targets(const double *sigma, int ld, const int *targets, long n_batch,
                            int n_select, double *out) {
    const cplx *S = (const cplx *)sigma;
    for (long i = 0; i < n_batch; ++i) {
        cplx h = hafnian_one(S, ld, targets + (size_t)i * n_select, n_select);
        out[2 * i] = h.re;
        out[2 * i + 1] = h.im;
    }
}

/* hafnian of explicit matrices A [n_batch][n][n] (interleaved complex128). */
void oracle_hafnian_dense(const double *A, long n_batch, int n, double *out) {
    int *t = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) t[i] = i;
    for (long b = 0; b < n_batch; ++b) {
        cplx h = hafnian_one((const cplx *)A + (size_t)b * n * n, n, t, n);
        out[2 * b] = h.re;
        out[2 * b + 1] = h.im;
    }
    free(t);
}

/* Loop-weight extension of solve().  l: [s][n+1], the loop polynomial of every remaining row:
 * the weight of the chains of matching and pairing edges that start at a looped vertex of an
 * already contracted pair and end at this row.  Contracting the pair (0, 1)
 *   closes a cycle (b(1,0)) or a chain looped at both ends (l(0) l(1)) at the pair:
 *       e = g + x g (b(1,0) + l(0) l(1)),
 *   extends chains through it as in solve(), and extends loop chains through it:
 *       l'(j) = l(j) + x (l(0) b(j,1) + l(1) b(j,0)).
 * With l = 0 every added term is an exact zero, so the value equals solve()'s. */
static cplx solve_loop(const cplx *b, const cplx *l, int s, int w, const cplx *g, int n) {
    if (s == 0) {
        cplx wc = {(double)w, 0.0};
        return cmul(wc, g[n]);
    }
    const int np1 = n + 1;
    const int nc = (s - 2) * (s - 3) / 2;
    cplx *c = (cplx *)calloc((size_t)(nc > 0 ? nc : 1) * np1, sizeof(cplx));
    cplx *lc = (cplx *)calloc((size_t)(s > 2 ? s - 2 : 1) * np1, sizeof(cplx));
    cplx *e = (cplx *)malloc((size_t)np1 * sizeof(cplx));
    cplx *q = (cplx *)malloc((size_t)np1 * sizeof(cplx));
    int i = 0;
    for (int j = 1; j < s - 2; ++j)
        for (int k = 0; k < j; ++k) {
            memcpy(&c[(size_t)i * np1], &b[(size_t)((j + 1) * (j + 2) / 2 + k + 2) * np1],
                   (size_t)np1 * sizeof(cplx));
            ++i;
        }
    for (int j = 0; j < s - 2; ++j) memcpy(&lc[(size_t)j * np1], &l[(size_t)(j + 2) * np1], (size_t)np1 * sizeof(cplx));
    cplx h = solve_loop(c, lc, s - 2, -w, g, n);
    /* q[v] = b(1,0)[v] + sum_p l(0)[p] l(1)[v-p], the p-sum in ascending p */
    for (int v = 0; v < n; ++v) {
        cplx acc = cmul(l[0 * np1 + 0], l[1 * np1 + v]);
        for (int p = 1; p <= v; ++p) acc = cadd(acc, cmul(l[0 * np1 + p], l[1 * np1 + v - p]));
        q[v] = cadd(b[0 * np1 + v], acc);
    }
    memcpy(e, g, (size_t)np1 * sizeof(cplx));
    for (int u = 0; u < n; ++u)
        for (int v = 0; v < n - u; ++v) e[u + v + 1] = cadd(e[u + v + 1], cmul(g[u], q[v]));
    for (int j = 1; j < s - 2; ++j)
        for (int k = 0; k < j; ++k)
            for (int u = 0; u < n; ++u)
                for (int v = 0; v < n - u; ++v) {
                    const cplx *bj0 = &b[(size_t)((j + 1) * (j + 2) / 2) * np1];
                    const cplx *bk1 = &b[(size_t)((k + 1) * (k + 2) / 2 + 1) * np1];
                    const cplx *bk0 = &b[(size_t)((k + 1) * (k + 2) / 2) * np1];
                    const cplx *bj1 = &b[(size_t)((j + 1) * (j + 2) / 2 + 1) * np1];
                    cplx t3 = cadd(cmul(bj0[u], bk1[v]), cmul(bk0[u], bj1[v]));
                    cplx *dst = &c[(size_t)(j * (j - 1) / 2 + k) * np1 + u + v + 1];
                    *dst = cadd(*dst, t3);
                }
    /* row j + 2 of b: b(j+2, 0) and b(j+2, 1) */
    for (int j = 0; j < s - 2; ++j)
        for (int u = 0; u < n; ++u)
            for (int v = 0; v < n - u; ++v) {
                const cplx *bj0 = &b[(size_t)((j + 1) * (j + 2) / 2) * np1];
                const cplx *bj1 = &b[(size_t)((j + 1) * (j + 2) / 2 + 1) * np1];
                cplx t3 = cadd(cmul(l[0 * np1 + u], bj1[v]), cmul(l[1 * np1 + u], bj0[v]));
                cplx *dst = &lc[(size_t)j * np1 + u + v + 1];
                *dst = cadd(*dst, t3);
            }
    cplx r = solve_loop(c, lc, s - 2, w, e, n);
    free(c);
    free(lc);
    free(e);
    free(q);
    return cadd(h, r);
}

/* Loop hafnian of the n x n matrix at(r, c) with loop weights lw(r).  An odd n is padded with
 * one vertex of loop weight 1 and no edges (which must be looped, so the value is unchanged). */
static cplx loop_hafnian_one(const cplx *A, const int *t, int ldA, const cplx *lw, int nsel) {
    cplx one = {1.0, 0.0}, zero = {0.0, 0.0};
    if (nsel == 0) return one;
    const int m = (nsel + 1) / 2, np1 = m + 1, ne = 2 * m;
    cplx *z = (cplx *)calloc((size_t)(ne * (ne - 1) / 2) * np1, sizeof(cplx));
    cplx *l = (cplx *)calloc((size_t)ne * np1, sizeof(cplx));
    for (int j = 1; j < ne; ++j)
        for (int k = 0; k < j; ++k)
            z[(size_t)(j * (j - 1) / 2 + k) * np1] = j < nsel ? A[(size_t)t[j] * ldA + t[k]] : zero;
    for (int j = 0; j < ne; ++j) l[(size_t)j * np1] = j < nsel ? lw[t[j]] : one;
    cplx *g = (cplx *)calloc((size_t)np1, sizeof(cplx));
    g[0] = one;
    cplx r = solve_loop(z, l, ne, 1, g, m);
    free(z);
    free(l);
    free(g);
    return r;
}

/* out[i] = lhaf(Sigma[t_i][:, t_i]) with loop weights loops[t_i] (loops [ld], complex128). */
void oracle_loop_hafnian_targets(const double *sigma, int ld, const double *loops, const int *targets,
                                 long n_batch, int n_select, double *out) {
    for (long i = 0; i < n_batch; ++i) {
        cplx h = loop_hafnian_one((const cplx *)sigma, targets + (size_t)i * n_select, ld, (const cplx *)loops,
                                  n_select);
        out[2 * i] = h.re;
        out[2 * i + 1] = h.im;
    }
}

/* out[b] = lhaf(A[b]) with loop weights loops[b] (A [n_batch][n][n], loops [n_batch][n]). */
void oracle_loop_hafnian_dense(const double *A, const double *loops, long n_batch, int n, double *out) {
    int *t = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) t[i] = i;
    for (long b = 0; b < n_batch; ++b) {
        cplx h = loop_hafnian_one((const cplx *)A + (size_t)b * n * n, t, n, (const cplx *)loops + (size_t)b * n, n);
        out[2 * b] = h.re;
        out[2 * b + 1] = h.im;
    }
    free(t);
}
