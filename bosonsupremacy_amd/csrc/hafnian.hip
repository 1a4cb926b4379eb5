// Batched hafnians for the MPS site tensors, gfx950.
//
// The arithmetic follows the reference recursion (MPS_cpu.py:117-187: hafnian dispatch,
// recursive_hafnian, solve) operation for operation, with NumPy's complex128 semantics, so
// results are bit-identical to the reference's complex128 evaluation.  Terms that are
// structurally zero (truncated-polynomial coefficients above the degree a node can have,
// and leaves whose polynomial cannot reach degree n/2) are skipped: adding them only ever
// adds +-0, which leaves every non-zero value unchanged.
//
// Three launch shapes:
//   * lane path   (n <= 10): one lane per matrix; the recursion is unrolled at compile time
//                 and the polynomial stack lives in VGPRs.  The gathered matrix is staged
//                 in LDS, laid out [entry][lane] so each wave reads it conflict-free.
//   * group path  (n >= 6):  G = 16/32/64 lanes per matrix follow a precomputed depth-first
//                 schedule of the recursion tree; each node's child matrix is built with one
//                 lane per entry, the polynomial stack lives in LDS.
//   * split path  (n = 12..34): the upper recursion tree breadth-first in HBM, one launch
//                 per level and one lane per (matrix, child entry); then one lane per
//                 (matrix, subtree) for every subtree that fits one lane's registers; then a
//                 tree-order combine.
// The lane and group paths gather each matrix from Sigma (L2-resident) through its target
// indices; the split path gathers it once into its scratch.
//
// Coefficients that cannot reach the result (live_degree below) are never computed.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <type_traits>
#include <mutex>
#include <vector>

#include "common.h"

namespace {

constexpr int kWave = 64;
constexpr int cmin(int a, int b) { return a < b ? a : b; }
__host__ __device__ constexpr int tri(int J, int K) { return J * (J - 1) / 2 + K; }
// Row J of packed index i = tri(J, K), 0 <= K < J (folds to a constant for a constant i).
__host__ __device__ constexpr int tri_row(int i) {
    int J = 1;
    while ((J + 1) * J / 2 <= i) ++J;
    return J;
}

// Highest polynomial coefficient of matrix entry (J, K), J > K, that can still reach the
// leaf value g[m] of a 2m x 2m hafnian.  A coefficient of degree t reaches g only by raising
// its degree: by 1 when the entry is a pivot pair (2p+1, 2p) feeding g at level p, by at
// least 2 for any other entry (it first feeds some child entry as a pivot).  Coefficients
// above this bound never contribute to the result, so they are not computed; every
// computed value is produced by the same operations in the same order as before.
__host__ __device__ constexpr int live_degree(int m, int J, int K) {
    return (J == K + 1 && (K & 1) == 0) ? m - 1 : m - 2;
}

// Workgroup barrier that orders LDS accesses only.  __syncthreads() also waits for the
// workgroup's outstanding global stores, which the cooperative kernel issues between its
// phases (frontier data) and never reads back in the same launch.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---------------------------------------------------------------------------------------
// Item sources: where an item's matrix comes from and where its result goes.
// fetch() fills the target indices t[0..n-1] and returns the base/leading dimension of the
// matrix the targets index; store() applies the epilogue and writes the result.
// ---------------------------------------------------------------------------------------

struct TargetsSource {  // A_elem: explicit targets, optional denominator
    const double* sigma;
    int ld;
    const int32_t* targets;
    int n_select;
    const double* den;
    double* out;
    int64_t n_items;
    __device__ void fetch(int64_t item, int k, int& tk) const {
        tk = targets[item * n_select + k];
    }
    __device__ const double* base(int64_t) const { return sigma; }
    __device__ int ldim() const { return ld; }
    __device__ void store(int64_t item, cd h) const {
        if (den) h = cscale(h, 1.0 / den[item]);
        st_cd(out, item, h);
    }
};

struct DenseSource {  // hafnian of explicit matrices
    const double* A;
    int n;
    double* out;
    int64_t n_items;
    __device__ void fetch(int64_t, int k, int& tk) const { tk = k; }
    __device__ const double* base(int64_t item) const { return A + 2 * item * (int64_t)n * n; }
    __device__ int ldim() const { return n; }
    __device__ void store(int64_t item, cd h) const { st_cd(out, item, h); }
};

// Tables of one site tensor, in device memory: Sigma, the per-label targets, denominators and
// photon-class permutations of both sides, Lambda, and where Gamma goes.
struct GammaSite {
    const double* sigma;
    const double* sqrt_fact;
    const int32_t* lt;
    const double* lden;
    const double* lam;
    const int32_t* lperm;
    const int32_t* loff;  // class offsets into lperm
    const int32_t* rt;
    const double* rden;
    const int32_t* rperm;
    const int32_t* roff;
    const int32_t* rcnt;  // class counts (right)
    double* gamma;
    double z, inv_z;
    int ld, d, chi, wl, wr, vacuum_real_div;
};

// A rectangle class_L[p] x class_R[q] x {j} of one site's (j, a, b) items of one photon total;
// items are numbered consecutively over the rectangles of a launch (start = first item).
struct GammaRect {
    int64_t start;
    int j, p, q, site;
};

struct GammaSource {  // site tensor entries Gamma[b, a, j] of one photon total n, of one or many sites
    const GammaSite* sites;
    const GammaRect* rects;  // sorted by start
    int n_rects;
    int64_t n_items;

    struct Decoded {
        int j, p, q, a, b;
        const GammaSite* s;
    };
    __device__ Decoded decode(int64_t item) const {
        int lo = 0, hi = n_rects - 1;
        while (lo < hi) {  // last rect with start <= item
            int mid = (lo + hi + 1) >> 1;
            if (rects[mid].start <= item) lo = mid;
            else hi = mid - 1;
        }
        const GammaRect r = rects[lo];
        const GammaSite* st = sites + r.site;
        const int64_t off = item - r.start;
        const int cr = st->rcnt[r.q];
        const int64_t ia = off / cr;
        const int ib = (int)(off - ia * cr);
        Decoded dc;
        dc.j = r.j;
        dc.p = r.p;
        dc.q = r.q;
        dc.a = st->lperm[st->loff[r.p] + ia];
        dc.b = st->rperm[st->roff[r.q] + ib];
        dc.s = st;
        return dc;
    }
    __device__ void fetch_dec(const Decoded& dc, int k, int& tk) const {
        if (k < dc.j) tk = 0;
        else if (k < dc.j + dc.p) tk = dc.s->lt[(int64_t)dc.a * dc.s->wl + (k - dc.j)];
        else tk = dc.s->rt[(int64_t)dc.b * dc.s->wr + (k - dc.j - dc.p)];
    }
    __device__ static const double* base_dec(const Decoded& dc) { return dc.s->sigma; }
    __device__ static int ldim_dec(const Decoded& dc) { return dc.s->ld; }
    __device__ void store_dec(const Decoded& dc, cd h) const {
        const GammaSite* st = dc.s;
        cd v;
        if (st->vacuum_real_div && dc.j + dc.p + dc.q == 0) {
            v.re = (1.0 / st->z) / st->lam[dc.a];
            v.im = 0.0;
        } else {
            const double den = (st->sqrt_fact[dc.j] * st->lden[dc.a]) * st->rden[dc.b];
            v = cscale(cscale(cscale(h, 1.0 / den), st->inv_z), 1.0 / st->lam[dc.a]);
        }
        st_cd(st->gamma, ((int64_t)dc.b * st->chi + dc.a) * st->d + dc.j, v);
    }
};

// Uniform adaptor so the kernels see one interface.
template <class S>
struct Src {
    const S& s;
    __device__ explicit Src(const S& src) : s(src) {}
};

// ---------------------------------------------------------------------------------------
// Lane path: compile-time recursion.  Level L node of a 2M x 2M hafnian holds the entries
// (J, K), J > K >= 2L, of its matrix as truncated polynomials; coefficient 0 is always
// the gathered A(J, K) (the recursion never updates it), coefficients 1..D live in P.
// ---------------------------------------------------------------------------------------

template <int M>
struct LaneDims {
    static constexpr int NE = M * (2 * M - 1) > 0 ? M * (2 * M - 1) : 1;
    static constexpr int NC = M > 0 ? M : 1;
};

template <int M, int L, int D, int G>
struct Node {
    static constexpr int NE = LaneDims<M>::NE;
    static constexpr int NC = LaneDims<M>::NC;
    static constexpr int D2 = cmin(M, 2 * D + 1);
    static constexpr int G2 = cmin(M, G + D + 1);

    static constexpr bool nonzero() {
        if constexpr (L == M) return G >= M;
        else return Node<M, L + 1, D, G>::nonzero() || Node<M, L + 1, D2, G2>::nonzero();
    }
    static constexpr int cap(int J, int K) { return live_degree(M, J, K); }

    template <class AF>
    __device__ __forceinline__ static cd coef(const AF& A, const cd (&P)[NE][NC], int J, int K, int u) {
        return u == 0 ? A(J, K) : P[tri(J, K)][u - 1];
    }

    template <class AF>
    __device__ __forceinline__ static cd run(const AF& A, const cd (&P)[NE][NC], const cd (&g)[M + 1],
                                             double w) {
        if constexpr (L == M) {
            return cmul(cd{w, 0.0}, g[M]);  // solve(): w * g[:, n]
        } else {
            constexpr bool lnz = Node<M, L + 1, D, G>::nonzero();
            constexpr bool rnz = Node<M, L + 1, D2, G2>::nonzero();
            cd left = czero(), right = czero();
            if constexpr (lnz) left = Node<M, L + 1, D, G>::run(A, P, g, -w);
            if constexpr (rnz) {
                constexpr int b0 = 2 * L, b1 = 2 * L + 1;
                cd Q[NE][NC];
                cd e[M + 1];
#pragma unroll
                for (int J = 2 * L + 3; J < 2 * M; ++J) {
#pragma unroll
                    for (int K = 2 * L + 2; K < J; ++K) {
#pragma unroll
                        for (int t = 1; t <= D2; ++t) {
                            if (t > cap(J, K)) continue;
                            cd acc = (t <= D) ? P[tri(J, K)][t - 1] : czero();
#pragma unroll
                            for (int u = 0; u < t; ++u) {
                                const int v = t - 1 - u;
                                if (u <= D && v <= D) {
                                    const cd t1 = cmul(coef(A, P, J, b0, u), coef(A, P, K, b1, v));
                                    const cd t2 = cmul(coef(A, P, K, b0, u), coef(A, P, J, b1, v));
                                    acc = cadd(acc, cadd(t1, t2));
                                }
                            }
                            Q[tri(J, K)][t - 1] = acc;
                        }
                    }
                }
                e[0] = g[0];
#pragma unroll
                for (int t = 1; t <= M; ++t) {
                    cd acc = (t <= G) ? g[t] : czero();
#pragma unroll
                    for (int u = 0; u < t; ++u) {
                        const int v = t - 1 - u;
                        if (u <= G && v <= D) acc = cadd(acc, cmul(g[u], coef(A, P, b1, b0, v)));
                    }
                    e[t] = acc;
                }
                right = Node<M, L + 1, D2, G2>::run(A, Q, e, w);
            }
            if constexpr (lnz && rnz) return cadd(left, right);
            else if constexpr (lnz) return left;
            else return right;
        }
    }
};

// A accessor reading a lane's gathered matrix from LDS ([entry][lane] layout).
struct LdsA {
    const cd* base;  // &lds[0 * stride + lane]
    int stride;
    __device__ __forceinline__ cd operator()(int J, int K) const { return base[tri(J, K) * stride]; }
};

template <int M>
__device__ __forceinline__ cd lane_hafnian(const int (&t)[2 * M > 0 ? 2 * M : 1], const double* S,
                                           int ld, cd* lds_lane, int stride) {
    if constexpr (M == 0) {
        return cd{1.0, 0.0};
    } else if constexpr (M == 1) {
        return ld_cd(S, (int64_t)t[0] * ld + t[1]);  // A[:, 0, 1]
    } else if constexpr (M == 2) {
        // A01*A23 + A02*A13 + A03*A12, evaluated left to right (MPS_cpu.py:137-138)
        auto A = [&](int r, int c) { return ld_cd(S, (int64_t)t[r] * ld + t[c]); };
        return cadd(cadd(cmul(A(0, 1), A(2, 3)), cmul(A(0, 2), A(1, 3))), cmul(A(0, 3), A(1, 2)));
    } else {
        // stage the strictly-lower triangle A[J][K] = Sigma[t_J][t_K] (recursive_hafnian :154-157)
#pragma unroll
        for (int J = 1; J < 2 * M; ++J)
#pragma unroll
            for (int K = 0; K < J; ++K)
                lds_lane[tri(J, K) * stride] = ld_cd(S, (int64_t)t[J] * ld + t[K]);
        LdsA A{lds_lane, stride};
        cd P0[LaneDims<M>::NE][LaneDims<M>::NC];
        cd g[M + 1];
        g[0] = cd{1.0, 0.0};
#pragma unroll
        for (int i = 1; i <= M; ++i) g[i] = czero();
        return Node<M, 0, 0, 0>::run(A, P0, g, 1.0);
    }
}

template <int M, class S>
__global__ __launch_bounds__(kWave) void haf_lane_kernel(S src, int64_t item0) {
    __shared__ cd lds[LaneDims<M>::NE * kWave];
    const int lane = threadIdx.x;
    const int64_t item = item0 + (int64_t)blockIdx.x * kWave + lane;
    if (item >= src.n_items) return;
    int t[2 * M > 0 ? 2 * M : 1];
    if constexpr (std::is_same<S, GammaSource>::value) {
        const auto dc = src.decode(item);
#pragma unroll
        for (int k = 0; k < 2 * M; ++k) src.fetch_dec(dc, k, t[k]);
        const cd h = lane_hafnian<M>(t, GammaSource::base_dec(dc), GammaSource::ldim_dec(dc), lds + lane, kWave);
        src.store_dec(dc, h);
    } else {
#pragma unroll
        for (int k = 0; k < 2 * M; ++k) src.fetch(item, k, t[k]);
        const cd h = lane_hafnian<M>(t, src.base(item), src.ldim(), lds + lane, kWave);
        src.store(item, h);
    }
}

// ---------------------------------------------------------------------------------------
// Group path: schedule-driven depth-first traversal with the stack in LDS.
// ---------------------------------------------------------------------------------------

// Schedule op, packed in one int4:
//   x = kind | level << 8 | owner << 16        (kind: 0 RIGHT, 1 LEAF, 2 ADD, 3 END)
//   y = D | D2 << 8 | G << 16 | G2 << 24
//   z = g_src | g_dst << 8 | (w > 0) << 16
struct Layout {  // offsets in complex elements within one matrix's LDS region
    int m;
    int off_A;
    int off_S[16];  // level storage, 1..m-1
    int cap[16];    // coefficients stored per entry at that level
    int off_g;      // (m+1) levels x (m+1) coefficients
    int off_val;    // value stack (m+2)
    int off_tgt;    // 2m ints (packed 4 per complex slot)
    int total;      // complex elements
};

Layout make_layout(int m) {
    Layout L{};
    L.m = m;
    int off = 0;
    L.off_A = off;
    off += m * (2 * m - 1);
    for (int l = 1; l < m; ++l) {
        const int s = 2 * m - 2 * l;
        const int n_ent = s * (s - 1) / 2;
        const int cap = std::min(m - 1, (1 << std::min(l, 20)) - 1);  // live coefficients only
        L.off_S[l] = off;
        L.cap[l] = cap;
        off += n_ent * cap;
    }
    L.off_g = off;
    off += (m + 1) * (m + 1);
    L.off_val = off;
    off += m + 2;
    L.off_tgt = off;
    off += (2 * m + 3) / 4 + 1;
    L.total = off;
    return L;
}

struct ScheduleBuilder {
    int m;
    std::vector<int4> ops;
    std::map<long, bool> memo;
    bool nonzero(int level, int D, int G) {
        if (level == m) return G >= m;
        const long key = ((long)level << 20) | (D << 10) | G;
        auto it = memo.find(key);
        if (it != memo.end()) return it->second;
        const int D2 = std::min(m, 2 * D + 1), G2 = std::min(m, G + D + 1);
        const bool r = nonzero(level + 1, D, G) || nonzero(level + 1, D2, G2);
        memo[key] = r;
        return r;
    }
    void gen(int level, int owner, int D, int G, int gl, int w) {
        if (level == m) {
            ops.push_back(make_int4(1 | (level << 8), 0, gl | ((w > 0) << 16), 0));
            return;
        }
        const int D2 = std::min(m, 2 * D + 1), G2 = std::min(m, G + D + 1);
        const bool lnz = nonzero(level + 1, D, G), rnz = nonzero(level + 1, D2, G2);
        if (lnz) gen(level + 1, owner, D, G, gl, -w);
        if (rnz) {
            ops.push_back(make_int4(0 | (level << 8) | (owner << 16), D | (D2 << 8) | (G << 16) | (G2 << 24),
                                    gl | ((level + 1) << 8) | ((w > 0) << 16), 0));
            gen(level + 1, level + 1, D2, G2, level + 1, w);
        }
        if (lnz && rnz) ops.push_back(make_int4(2, 0, 0, 0));
    }
};

struct GroupPlan {
    int m;
    Layout lay;
    int4* d_ops;  // device copy
    int n_ops;
};

std::mutex g_plan_mu;
std::map<std::pair<int, int>, GroupPlan> g_plans;  // (device, m)

int get_plan(int device, int m, GroupPlan** out) {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto key = std::make_pair(device, m);
    auto it = g_plans.find(key);
    if (it == g_plans.end()) {
        ScheduleBuilder sb;
        sb.m = m;
        sb.gen(0, 0, 0, 0, 0, 1);
        sb.ops.push_back(make_int4(3, 0, 0, 0));
        sb.ops.push_back(make_int4(3, 0, 0, 0));  // read ahead by the kernel
        GroupPlan p;
        p.m = m;
        p.lay = make_layout(m);
        p.n_ops = (int)sb.ops.size();
        BSMPS_HIP(hipMalloc(&p.d_ops, sizeof(int4) * sb.ops.size()));
        BSMPS_HIP(hipMemcpy(p.d_ops, sb.ops.data(), sizeof(int4) * sb.ops.size(), hipMemcpyHostToDevice));
        it = g_plans.emplace(key, p).first;
    }
    *out = &it->second;
    return BSMPS_OK;
}

__device__ __forceinline__ void decode_pair(int i, int& J, int& K) {
    // i = J(J-1)/2 + K with J > K >= 0
    int j = (int)((1.0f + __fsqrt_rn(1.0f + 8.0f * (float)i)) * 0.5f);
    while (j * (j - 1) / 2 > i) --j;
    while ((j + 1) * j / 2 <= i) ++j;
    J = j;
    K = i - j * (j - 1) / 2;
}

// Right-branch child-matrix entry (J, K): its live coefficients t = 1..min(D2, live degree) of
//   c(J,K)[t] = b(J,K)[t] + sum_u (b(J,b0)[u] b(K,b1)[v] + b(K,b0)[u] b(J,b1)[v]),  v = t-1-u,
// accumulated in ascending u as solve() does (MPS_cpu.py:179-186).  The node's four pivot
// polynomials are loaded once into registers; D (the node's degree) and D2 are compile-time,
// which is possible because a node's degree is always 2^r - 1 or m.
template <int D, int D2>
__device__ __forceinline__ void child_entry(const cd* Ar, const cd* So, int obase, int ocap, cd* dst, int J,
                                            int K, int b0, int b1, int tmax) {
    auto co = [&](int r, int c, int u) -> cd {
        return u == 0 ? Ar[tri(r, c)] : So[tri(r - obase, c - obase) * ocap + (u - 1)];
    };
    constexpr int DU = D < D2 - 1 ? D : D2 - 1;  // highest pivot degree any t <= D2 uses
    cd X[DU + 1], Y[DU + 1], Z[DU + 1], W[DU + 1];
#pragma unroll
    for (int u = 0; u <= DU; ++u) {
        const bool need = u < tmax;
        X[u] = need ? co(J, b0, u) : czero();
        Y[u] = need ? co(K, b1, u) : czero();
        Z[u] = need ? co(K, b0, u) : czero();
        W[u] = need ? co(J, b1, u) : czero();
    }
#pragma unroll
    for (int t = 1; t <= D2; ++t) {
        if (t <= tmax) {
            cd acc = (t <= D) ? co(J, K, t) : czero();
#pragma unroll
            for (int u = 0; u <= DU; ++u) {
                const int v = t - 1 - u;
                if (v >= 0 && v <= D && u <= D) acc = cadd(acc, cadd(cmul(X[u], Y[v]), cmul(Z[u], W[v])));
            }
            dst[t - 1] = acc;
        }
    }
}

template <int M>
__device__ __forceinline__ void child_entry_dispatch(int D, const cd* Ar, const cd* So, int obase, int ocap, cd* dst,
                                                     int J, int K, int b0, int b1, int tmax) {
    if (D == M) {
        child_entry<M, M>(Ar, So, obase, ocap, dst, J, K, b0, b1, tmax);
    } else if (D == 0) {
        child_entry<0, cmin(M, 1)>(Ar, So, obase, ocap, dst, J, K, b0, b1, tmax);
    } else if (D == 1) {
        if constexpr (M > 1) child_entry<1, cmin(M, 3)>(Ar, So, obase, ocap, dst, J, K, b0, b1, tmax);
    } else if (D == 3) {
        if constexpr (M > 3) child_entry<3, cmin(M, 7)>(Ar, So, obase, ocap, dst, J, K, b0, b1, tmax);
    } else if (D == 7) {
        if constexpr (M > 7) child_entry<7, cmin(M, 15)>(Ar, So, obase, ocap, dst, J, K, b0, b1, tmax);
    }
}

template <int M, class S>
__global__ __launch_bounds__(kWave) void haf_group_kernel(S src, int64_t item0, Layout lay,
                                                          const int4* __restrict__ ops, int G) {
    extern __shared__ cd smem[];
    const int PER = kWave / G;  // matrices per workgroup
    const int grp = threadIdx.x / G;
    const int gl = threadIdx.x % G;
    constexpr int m = M;
    const int n = 2 * m;
    cd* R = smem + grp * lay.total;
    int* tg = reinterpret_cast<int*>(R + lay.off_tgt);
    const int64_t item = item0 + (int64_t)blockIdx.x * PER + grp;
    const bool active = item < src.n_items;
    const int64_t eff = active ? item : src.n_items - 1;

    typename std::conditional<std::is_same<S, GammaSource>::value, GammaSource::Decoded, int>::type dc{};
    if constexpr (std::is_same<S, GammaSource>::value) dc = src.decode(eff);
    for (int k = gl; k < n; k += G) {
        int tk;
        if constexpr (std::is_same<S, GammaSource>::value) src.fetch_dec(dc, k, tk);
        else src.fetch(eff, k, tk);
        tg[k] = tk;
    }
    lds_barrier();
    const double* Sg;
    int ld;
    if constexpr (std::is_same<S, GammaSource>::value) {
        Sg = GammaSource::base_dec(dc);
        ld = GammaSource::ldim_dec(dc);
    } else {
        Sg = src.base(eff);
        ld = src.ldim();
    }
    for (int i = gl; i < m * (2 * m - 1); i += G) {
        int J, K;
        decode_pair(i, J, K);
        R[lay.off_A + i] = ld_cd(Sg, (int64_t)tg[J] * ld + tg[K]);
    }
    if (gl <= m) R[lay.off_g + gl] = gl == 0 ? cd{1.0, 0.0} : czero();  // g at level 0
    lds_barrier();
    int sp = 0;  // value-stack depth (uniform)
    int4 op = ops[0];
    for (int pc = 0;; ++pc) {
        const int4 next = ops[pc + 1];  // the schedule ends with two END ops; fetched ahead
        const int kind = op.x & 0xff;
        if (kind == 3) break;
        if (kind == 1) {  // LEAF: push w * g[m]
            if (gl == 0) {
                const int gs = op.z & 0xff;
                const double w = (op.z >> 16) & 1 ? 1.0 : -1.0;
                R[lay.off_val + sp] = cmul(cd{w, 0.0}, R[lay.off_g + gs * (m + 1) + m]);
            }
            ++sp;
        } else if (kind == 2) {  // ADD: left + right
            if (gl == 0) R[lay.off_val + sp - 2] = cadd(R[lay.off_val + sp - 2], R[lay.off_val + sp - 1]);
            --sp;
        } else {  // RIGHT: build the right child of a level-`level` node
            const int level = (op.x >> 8) & 0xff;
            const int owner = (op.x >> 16) & 0xff;
            const int D = op.y & 0xff, D2 = (op.y >> 8) & 0xff;
            const int Gd = (op.y >> 16) & 0xff, G2 = (op.y >> 24) & 0xff;
            const int gsrc = op.z & 0xff, gdst = (op.z >> 8) & 0xff;
            const int s = n - 2 * level;
            const int n_child = (s - 2) * (s - 3) / 2;
            const int b0 = 2 * level, b1 = 2 * level + 1;
            const int obase = 2 * owner;
            const cd* So = R + lay.off_S[owner > 0 ? owner : 1];
            const int ocap = owner > 0 ? lay.cap[owner] : 0;
            cd* Sd = R + lay.off_S[level + 1 < m ? level + 1 : 1];
            const int dcap = level + 1 < m ? lay.cap[level + 1] : 0;
            // coefficient u of entry (J, K) of the current node's matrix
            auto coef = [&](int J, int K, int u) -> cd {
                if (u == 0) return R[lay.off_A + tri(J, K)];
                return So[tri(J - obase, K - obase) * ocap + (u - 1)];
            };
            const cd* gs = R + lay.off_g + gsrc * (m + 1);
            cd* gd = R + lay.off_g + gdst * (m + 1);
            // One work item per child entry (all its live coefficients), then one per g update e[t].
            for (int it = gl; it < n_child + m; it += G) {
                if (it < n_child) {
                    int Jl, Kl;
                    decode_pair(it, Jl, Kl);
                    const int J = Jl + b0 + 2, K = Kl + b0 + 2;
                    const int tmax = std::min(D2, live_degree(m, J, K));
                    child_entry_dispatch<M>(D, R + lay.off_A, So, obase, ocap, Sd + tri(Jl, Kl) * dcap, J, K, b0, b1,
                                            tmax);
                } else {
                    const int t = it - n_child + 1;  // e[t], t = 1..m
                    cd acc = (t <= Gd) ? gs[t] : czero();
                    const int ulo = t - 1 - D > 0 ? t - 1 - D : 0;
                    const int uhi = t - 1 < Gd ? t - 1 : Gd;
                    for (int u = ulo; u <= uhi; ++u) acc = cadd(acc, cmul(gs[u], coef(b1, b0, t - 1 - u)));
                    gd[t] = acc;
                }
            }
            if (gl == 0) gd[0] = gs[0];
            (void)G2;
            (void)D2;
            lds_barrier();
        }
        op = next;
    }
    lds_barrier();
    if (active && gl == 0) {
        const cd h = R[lay.off_val];
        if constexpr (std::is_same<S, GammaSource>::value) src.store_dec(dc, h);
        else src.store(item, h);
    }
}

// ---------------------------------------------------------------------------------------
// Frontier split (sizes 12..34).  The recursion tree is cut at "frontier" nodes whose subtree
// fits one lane's registers.  The part above the frontier runs breadth-first, one launch per
// tree level, with one lane per (hafnian, child-matrix entry): every node that owns new
// polynomials (the root and each right child) keeps them in HBM, laid out
// data[slot * stride + h] with the hafnian index h fastest, so all accesses are coalesced and
// no lane idles.  Then one lane per (hafnian, frontier subtree) evaluates the subtree with the
// compile-time recursion, and a last kernel adds the subtree values in tree order and applies
// the item's epilogue.  Every value is produced by the same operations in the same order as
// the depth-first recursion, so results are bit-identical to it.
// ---------------------------------------------------------------------------------------

// One RIGHT step of the upper tree: the node (level, degree D, g degree G) reads its matrix
// from its owner's block (rows >= 2 plo, pcap coefficients per entry; po < 0: the root, whose
// matrix is A alone) and its g from gsrc; the right child's block goes to pdst (dcap
// coefficients per entry) and its g to gdst.
struct BfsOp {
    int level, D, D2, G;
    int po, plo, pcap, gsrc;
    int pdst, dcap, gdst, pad;
};

struct BfsData {
    double* data;     // complex slots, slot-major
    int64_t stride;   // hafnians per slot (chunk capacity)
    int64_t n_h;      // hafnians in this chunk
    int a_off;        // A(J, K) at slot a_off + tri(J, K)
};

__device__ __forceinline__ cd bfs_ld(const BfsData& bd, int slot, int64_t h) {
    return ld_cd(bd.data, (int64_t)slot * bd.stride + h);
}

// XCD-aware work mapping of the split path's (work item, block of 64 hafnians) grids.  The
// hardware hands workgroups to the 8 XCDs round-robin by linear workgroup id, and each XCD has
// its own L2; all work items of one hafnian block read that block's shared data (the pivot
// polynomials of a RIGHT step, the gathered matrix of a subtree), so the W = n_items *
// n_hblocks work units, ordered hafnian-block-major, are cut into 8 contiguous ranges of
// Q = ceil(W / 8), one per XCD: linear id L runs on XCD c = L % 8 and takes unit c Q + L / 8.
// A hafnian block is then split between at most two XCDs; the grid is 8 Q workgroups, of which
// at most 7 have no unit.
constexpr int kXcds = 8;
__device__ __forceinline__ bool xcd_work(int n_items, int n_hblocks, int& item, int& hblock) {
    const int64_t W = (int64_t)n_items * n_hblocks;
    const int64_t Q = (W + kXcds - 1) / kXcds;
    const int64_t L = blockIdx.x;
    const int64_t w = (L % kXcds) * Q + L / kXcds;
    if (w >= W) return false;
    hblock = (int)(w / n_items);
    item = (int)(w % n_items);
    return true;
}
inline int64_t xcd_grid(int n_items, int n_hblocks) {
    return ((int64_t)n_items * n_hblocks + kXcds - 1) / kXcds * kXcds;
}

// Gather A and the root's g of every hafnian of the chunk (one lane per hafnian).
template <int M, class S>
__global__ __launch_bounds__(kWave) void haf_bfs_prep_kernel(S src, int64_t item0, BfsData bd, int rootg) {
    const int64_t h = (int64_t)blockIdx.x * kWave + threadIdx.x;
    if (h >= bd.n_h) return;
    const int64_t item = item0 + h;
    int t[2 * M];
    const double* Sg;
    int ld;
    if constexpr (std::is_same<S, GammaSource>::value) {
        const auto dc = src.decode(item);
#pragma unroll
        for (int k = 0; k < 2 * M; ++k) src.fetch_dec(dc, k, t[k]);
        Sg = GammaSource::base_dec(dc);
        ld = GammaSource::ldim_dec(dc);
    } else {
#pragma unroll
        for (int k = 0; k < 2 * M; ++k) src.fetch(item, k, t[k]);
        Sg = src.base(item);
        ld = src.ldim();
    }
    // the gathered entries go out in batches of 8, each batch loaded before any of it is stored
    // (a load placed after a store to possibly the same memory waits for nothing but cannot be
    // moved above it, so one-at-a-time gathers would each wait a full latency)
    constexpr int NA = M * (2 * M - 1);
    constexpr int kB = 8;
#pragma unroll
    for (int i0 = 0; i0 < NA; i0 += kB) {
        cd v[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            if (i0 + u < NA) {
                const int J = tri_row(i0 + u), K = i0 + u - tri(J, 0);
                v[u] = ld_cd(Sg, (int64_t)t[J] * ld + t[K]);
            }
        }
#pragma unroll
        for (int u = 0; u < kB; ++u)
            if (i0 + u < NA) st_cd(bd.data, (int64_t)(bd.a_off + i0 + u) * bd.stride + h, v[u]);
    }
#pragma unroll
    for (int i = 0; i <= M; ++i)
        st_cd(bd.data, (int64_t)(rootg + i) * bd.stride + h, i == 0 ? cd{1.0, 0.0} : czero());
}

// Child entry (J, K) of a RIGHT step (all its live coefficients), or the step's g update when
// e < 0.  D and D2 are compile-time (a node's degree is 2^r - 1 or m).
template <int M, int D, int D2>
__device__ __forceinline__ void bfs_item(const BfsData& bd, const BfsOp& op, int e, int64_t h) {
    const int b0 = 2 * op.level, b1 = b0 + 1;
    const int ob = 2 * op.plo;
    auto co = [&](int r, int c, int u) -> cd {
        return u == 0 ? bfs_ld(bd, bd.a_off + tri(r, c), h) : bfs_ld(bd, op.po + tri(r - ob, c - ob) * op.pcap + (u - 1), h);
    };
    if (e >= 0) {
        int Jl, Kl;
        decode_pair(e, Jl, Kl);
        const int J = Jl + b0 + 2, K = Kl + b0 + 2;
        const int tmax = cmin(D2, live_degree(M, J, K));
        constexpr int DU = D < D2 - 1 ? D : D2 - 1;
        cd X[DU + 1], Y[DU + 1], Z[DU + 1], W[DU + 1];
#pragma unroll
        for (int u = 0; u <= DU; ++u) {
            const bool need = u < tmax;
            X[u] = need ? co(J, b0, u) : czero();
            Y[u] = need ? co(K, b1, u) : czero();
            Z[u] = need ? co(K, b0, u) : czero();
            W[u] = need ? co(J, b1, u) : czero();
        }
        // The entry's own coefficients are all loaded before the first store: loads and stores
        // share one buffer, so a load placed after a store cannot be moved above it, and each
        // coefficient would then wait a full memory latency of its own.  They are read by this
        // item only, and the child coefficients below are written once, so both stream past the
        // L2 (non-temporal), which keeps the pivot polynomials every item of the node re-reads.
        cd acc[D2 + 1];
#pragma unroll
        for (int t = 1; t <= D2; ++t)
            acc[t] = (t <= tmax && t <= D)
                         ? ld_cd_nt(bd.data, (int64_t)(op.po + tri(J - ob, K - ob) * op.pcap + (t - 1)) * bd.stride + h)
                         : czero();
        const int64_t dst = (int64_t)(op.pdst + e * op.dcap) * bd.stride + h;
#pragma unroll
        for (int t = 1; t <= D2; ++t) {
            if (t <= tmax) {
                // c(J,K)[t] = b(J,K)[t] + sum_u (b(J,b0)[u] b(K,b1)[v] + b(K,b0)[u] b(J,b1)[v]),
                // v = t-1-u, ascending u (MPS_cpu.py:179-186)
#pragma unroll
                for (int u = 0; u <= DU; ++u) {
                    const int v = t - 1 - u;
                    if (v >= 0 && v <= D && u <= D) acc[t] = cadd(acc[t], cadd(cmul(X[u], Y[v]), cmul(Z[u], W[v])));
                }
                st_cd_nt(bd.data, dst + (int64_t)(t - 1) * bd.stride, acc[t]);
            }
        }
    } else {
        // e = g + sum_{u,v} g[u] b(b1,b0)[v] into e[u+v+1] (MPS_cpu.py:175-178)
        cd g[M + 1], p[D + 1];
#pragma unroll
        for (int i = 0; i <= M; ++i) g[i] = i <= op.G ? bfs_ld(bd, op.gsrc + i, h) : czero();
#pragma unroll
        for (int v = 0; v <= D; ++v) p[v] = v < M ? co(b1, b0, v) : czero();
        st_cd(bd.data, (int64_t)op.gdst * bd.stride + h, g[0]);
#pragma unroll
        for (int t = 1; t <= M; ++t) {
            cd acc = (t <= op.G) ? g[t] : czero();
#pragma unroll
            for (int u = 0; u < t; ++u) {
                const int v = t - 1 - u;
                if (u <= op.G && v <= D) acc = cadd(acc, cmul(g[u], p[v]));
            }
            st_cd(bd.data, (int64_t)(op.gdst + t) * bd.stride + h, acc);
        }
    }
}

// One tree level of the upper phase, for the RIGHT steps of one node degree D: workgroup =
// (item (op, entry) of this level and degree, block of 64 hafnians), placed by xcd_work.  Each
// degree has its own kernel, so a launch of low-degree items is not held to the registers of the
// degree-M items (202 VGPRs at M = 11, two waves per SIMD).
template <int M, int D>
__global__ __launch_bounds__(kWave) void haf_bfs_level_kernel(BfsData bd, const BfsOp* __restrict__ ops,
                                                              const int2* __restrict__ items, int n_items,
                                                              int n_hblocks) {
    int item, hb;
    if (!xcd_work(n_items, n_hblocks, item, hb)) return;
    const int64_t h = (int64_t)hb * kWave + threadIdx.x;
    if (h >= bd.n_h) return;
    const int2 it = items[item];
    const BfsOp op = ops[it.x];
    bfs_item<M, D, (D == M ? M : cmin(M, 2 * D + 1))>(bd, op, it.y, h);
}

template <int OFF>
struct LdsAOff {  // A(J, K) of a subtree whose rows start at OFF, [entry][lane] in LDS
    const cd* base;
    __device__ __forceinline__ cd operator()(int J, int K) const { return base[tri(J - OFF, K - OFF) * kWave]; }
};

// Frontier node: {k | D << 8 | G << 16 | (w > 0) << 24, po, plo | pcap << 8, g slot}.
// Workgroup = (frontier node of this shape, block of 64 hafnians), placed by xcd_work.
template <int M, int K0, int D>
__global__ __launch_bounds__(kWave) void haf_subtree_kernel(BfsData bd, const int4* __restrict__ front,
                                                            const int* __restrict__ tlist, double* __restrict__ vals,
                                                            int n_front, int n_hblocks) {
    constexpr int S0 = 2 * (M - K0);
    constexpr int NA = S0 * (S0 - 1) / 2;
    __shared__ cd lds[NA * kWave];
    const int lane = threadIdx.x;
    int fi, hb;
    if (!xcd_work(n_front, n_hblocks, fi, hb)) return;
    const int64_t h = (int64_t)hb * kWave + lane;
    const int f = tlist[fi];
    const int4 fr = front[f];
    if (h >= bd.n_h) return;
#pragma unroll
    for (int J = 2 * K0 + 1; J < 2 * M; ++J)
#pragma unroll
        for (int K = 2 * K0; K < J; ++K) lds[tri(J - 2 * K0, K - 2 * K0) * kWave + lane] = bfs_ld(bd, bd.a_off + tri(J, K), h);
    LdsAOff<2 * K0> A{lds + lane};
    cd P[LaneDims<M>::NE][LaneDims<M>::NC];
    const int po = fr.y, ob = 2 * (fr.z & 0xff), pcap = fr.z >> 8;
#pragma unroll
    for (int J = 2 * K0 + 1; J < 2 * M; ++J)
#pragma unroll
        for (int K = 2 * K0; K < J; ++K)
#pragma unroll
            for (int t = 0; t < D; ++t)
                if (t < live_degree(M, J, K))  // read once per launch: streamed past the L2
                    P[tri(J, K)][t] = ld_cd_nt(bd.data, (int64_t)(po + tri(J - ob, K - ob) * pcap + t) * bd.stride + h);
    cd g[M + 1];
#pragma unroll
    for (int t = 0; t <= M; ++t) g[t] = t <= D ? bfs_ld(bd, fr.w + t, h) : czero();
    const double w = (fr.x >> 24) & 1 ? 1.0 : -1.0;
    st_cd(vals, (int64_t)f * bd.stride + h, Node<M, K0, D, D>::run(A, P, g, w));
}

// Frontier split, last phase: add the subtree values in the order of the recursion tree and
// apply the item's epilogue.  The combine program is postfix: c >= 0 pushes the value in
// slot c, -1 replaces the top two by their sum (lower + upper, the recursion's left + right).
// A long program is first cut into independent parts -- complete subtrees of the add tree,
// each a contiguous stretch of the program -- that run in parallel (haf_combine_part_kernel,
// one workgroup per (part, block of 64 hafnians)) and leave their sums in the slot of their
// first value; the final kernel then runs the program of the tree above the parts.  Every sum
// is formed from the same operands in the same order as in one pass over the whole program.
constexpr int kCombineStack = 18;
constexpr int kCombineBatch = 16;

// Evaluates prog[0..n) for hafnian h; the values pushed by the next kCombineBatch steps are
// loaded before any is used, so a lane has that many loads in flight instead of one (the
// program is uniform across lanes).
__device__ __forceinline__ cd combine_eval(const double* vals, int64_t stride, int64_t h,
                                           const int* __restrict__ prog, int n, cd* stack, int lane) {
    int sp = 0;
    for (int i0 = 0; i0 < n; i0 += kCombineBatch) {
        cd v[kCombineBatch];
#pragma unroll
        for (int u = 0; u < kCombineBatch; ++u) {
            const int c = i0 + u < n ? prog[i0 + u] : -1;
            v[u] = c >= 0 ? ld_cd(vals, (int64_t)c * stride + h) : czero();
        }
#pragma unroll
        for (int u = 0; u < kCombineBatch; ++u) {
            if (i0 + u < n) {
                if (prog[i0 + u] >= 0) {
                    stack[sp * kWave + lane] = v[u];
                    ++sp;
                } else {
                    stack[(sp - 2) * kWave + lane] =
                        cadd(stack[(sp - 2) * kWave + lane], stack[(sp - 1) * kWave + lane]);
                    --sp;
                }
            }
        }
    }
    return stack[lane];
}

// One part of the combine program: parts[p] = {first step, steps, slot of the sum}.
// Workgroup = (part, block of 64 hafnians), placed by xcd_work.
__global__ __launch_bounds__(kWave) void haf_combine_part_kernel(int64_t n_h, int64_t stride, double* vals,
                                                                 const int* __restrict__ comb,
                                                                 const int4* __restrict__ parts, int n_parts,
                                                                 int n_hblocks) {
    __shared__ cd stack[kCombineStack * kWave];
    const int lane = threadIdx.x;
    int pi, hb;
    if (!xcd_work(n_parts, n_hblocks, pi, hb)) return;
    const int64_t h = (int64_t)hb * kWave + lane;
    if (h >= n_h) return;
    const int4 part = parts[pi];
    const cd v = combine_eval(vals, stride, h, comb + part.x, part.y, stack, lane);
    st_cd(vals, (int64_t)part.z * stride + h, v);
}

template <class S>
__global__ __launch_bounds__(kWave) void haf_combine_kernel(S src, int64_t item0, int64_t n_h, int64_t stride,
                                                            const double* __restrict__ vals,
                                                            const int* __restrict__ comb, int n_comb) {
    __shared__ cd stack[kCombineStack * kWave];
    const int lane = threadIdx.x;
    const int64_t h = (int64_t)blockIdx.x * kWave + lane;
    if (h >= n_h) return;
    const cd v = combine_eval(vals, stride, h, comb, n_comb, stack, lane);
    const int64_t item = item0 + h;
    if constexpr (std::is_same<S, GammaSource>::value) src.store_dec(src.decode(item), v);
    else src.store(item, v);
}

// ---------------------------------------------------------------------------------------
// Loop hafnians: a vertex may also be matched to itself, with weight loops[i].  The reference
// computes plain hafnians only (its pure state has zero mean); this is the same truncated-
// polynomial recursion with one more polynomial per row, l(J): the weight of the chains of
// matching and pairing edges that start at a looped vertex of an already contracted pair and
// end at row J (l(J) = loops[J] at the root).  Contracting the pivot pair (0, 1) closes a
// cycle or a chain looped at both ends there, e = g + x g (b(1,0) + l(0) l(1)), and extends
// the loop chains through it, l'(J) = l(J) + x (l(0) b(J,1) + l(1) b(J,0)); b is updated as in
// the plain recursion.  An odd size is padded with one vertex of loop weight 1 and no edges.
//
// One lane per matrix walks the recursion tree depth first.  Every node of level L has the
// rows 2L..2M-1, so one node per level is stored: its entries, loop polynomials and g (all
// M + 1 coefficients, no term skipped) and the value of its first child, slot-major
// [slot][matrix] in scratch so the lanes of a wave access consecutive matrices.  The terms are
// added in the same order as the C restatement in the oracle, so the results are identical to
// it; with every loop weight zero the added terms are exact zeros and the value is the plain
// recursion's.
// ---------------------------------------------------------------------------------------

constexpr int kLoopMaxM = BSMPS_MAX_LOOP_HAFNIAN_SIZE / 2;

template <int M>
struct LoopLayout {
    static constexpr int NP = M + 1;
    __host__ __device__ static constexpr int rows(int L) { return 2 * M - 2 * L; }
    __host__ __device__ static constexpr int entries(int L) { return rows(L) * (rows(L) - 1) / 2; }
    __host__ __device__ static constexpr int level_slots(int L) { return (entries(L) + rows(L) + 1) * NP + 1; }
    __host__ __device__ static constexpr int off(int L) {
        int o = 0;
        for (int i = 0; i < L; ++i) o += level_slots(i);
        return o;
    }
    __host__ __device__ static constexpr int total() { return off(M + 1); }
    // slots relative to off(L)
    __host__ __device__ static constexpr int b(int J, int K, int t) { return tri(J, K) * NP + t; }
    __host__ __device__ static constexpr int l(int L, int J, int t) { return (entries(L) + J) * NP + t; }
    __host__ __device__ static constexpr int g(int L, int t) { return (entries(L) + rows(L)) * NP + t; }
    __host__ __device__ static constexpr int first(int L) { return (entries(L) + rows(L) + 1) * NP; }
};

struct LoopData {
    double* data;    // complex slots, slot-major
    int64_t stride;  // matrices per slot (chunk capacity)
    int64_t n_h;     // matrices in this chunk
};

template <int M>
struct LoopNode {
    using Lay = LoopLayout<M>;
    static constexpr int NP = M + 1;
    double* D;
    int64_t st, h;
    __device__ cd ld(int o, int s) const { return ld_cd(D, (int64_t)(o + s) * st + h); }
    __device__ void put(int o, int s, cd v) const { st_cd(D, (int64_t)(o + s) * st + h, v); }

    // left child of the level-L node: rows 2.. of it, same g
    __device__ void left(int L) const {
        const int o = Lay::off(L), o1 = Lay::off(L + 1), s1 = Lay::rows(L) - 2;
        for (int j = 1; j < s1; ++j)
            for (int k = 0; k < j; ++k)
#pragma unroll
                for (int t = 0; t < NP; ++t) put(o1, Lay::b(j, k, t), ld(o, Lay::b(j + 2, k + 2, t)));
        for (int j = 0; j < s1; ++j)
#pragma unroll
            for (int t = 0; t < NP; ++t) put(o1, Lay::l(L + 1, j, t), ld(o, Lay::l(L, j + 2, t)));
#pragma unroll
        for (int t = 0; t < NP; ++t) put(o1, Lay::g(L + 1, t), ld(o, Lay::g(L, t)));
    }

    // right child of the level-L node: pair (0, 1) contracted
    __device__ void right(int L) const {
        const int o = Lay::off(L), o1 = Lay::off(L + 1), s1 = Lay::rows(L) - 2;
        cd l0[M], l1[M];
#pragma unroll
        for (int u = 0; u < M; ++u) {
            l0[u] = ld(o, Lay::l(L, 0, u));
            l1[u] = ld(o, Lay::l(L, 1, u));
        }
        {
            // q[v] = b(1,0)[v] + sum_p l(0)[p] l(1)[v-p] (ascending p); e = g + x g q
            cd q[M], g[NP];
#pragma unroll
            for (int v = 0; v < M; ++v) {
                cd acc = cmul(l0[0], l1[v]);
#pragma unroll
                for (int p = 1; p <= v; ++p) acc = cadd(acc, cmul(l0[p], l1[v - p]));
                q[v] = cadd(ld(o, Lay::b(1, 0, v)), acc);
            }
#pragma unroll
            for (int t = 0; t < NP; ++t) g[t] = ld(o, Lay::g(L, t));
#pragma unroll
            for (int t = 0; t < NP; ++t) {
                cd e = g[t];
#pragma unroll
                for (int u = 0; u < t; ++u) e = cadd(e, cmul(g[u], q[t - 1 - u]));
                put(o1, Lay::g(L + 1, t), e);
            }
        }
        for (int j = 1; j < s1; ++j) {
            cd X[M], W[M];
#pragma unroll
            for (int u = 0; u < M; ++u) {
                X[u] = ld(o, Lay::b(j + 2, 0, u));
                W[u] = ld(o, Lay::b(j + 2, 1, u));
            }
            for (int k = 0; k < j; ++k) {
                cd Y[M], Z[M];
#pragma unroll
                for (int u = 0; u < M; ++u) {
                    Y[u] = ld(o, Lay::b(k + 2, 1, u));
                    Z[u] = ld(o, Lay::b(k + 2, 0, u));
                }
#pragma unroll
                for (int t = 0; t < NP; ++t) {
                    cd acc = ld(o, Lay::b(j + 2, k + 2, t));
#pragma unroll
                    for (int u = 0; u < t; ++u) {
                        const int v = t - 1 - u;
                        acc = cadd(acc, cadd(cmul(X[u], Y[v]), cmul(Z[u], W[v])));
                    }
                    put(o1, Lay::b(j, k, t), acc);
                }
            }
        }
        for (int j = 0; j < s1; ++j) {
            cd B0[M], B1[M];
#pragma unroll
            for (int u = 0; u < M; ++u) {
                B0[u] = ld(o, Lay::b(j + 2, 0, u));
                B1[u] = ld(o, Lay::b(j + 2, 1, u));
            }
#pragma unroll
            for (int t = 0; t < NP; ++t) {
                cd acc = ld(o, Lay::l(L, j + 2, t));
#pragma unroll
                for (int u = 0; u < t; ++u) {
                    const int v = t - 1 - u;
                    acc = cadd(acc, cadd(cmul(l0[u], B1[v]), cmul(l1[u], B0[v])));
                }
                put(o1, Lay::l(L + 1, j, t), acc);
            }
        }
    }
};

// loops: per matrix [n_batch][n_sel] when loop_stride = n_sel (explicit matrices), else a
// vector over the rows of the matrix the targets index (loop weight of row J = loops[t[J]]).
template <int M, class S>
__global__ __launch_bounds__(kWave) void haf_loop_kernel(S src, const double* __restrict__ loops, int loop_stride,
                                                         int n_sel, int64_t item0, LoopData lp) {
    using Lay = LoopLayout<M>;
    constexpr int NP = M + 1, NE = 2 * M;
    const int64_t h = (int64_t)blockIdx.x * kWave + threadIdx.x;
    if (h >= lp.n_h) return;
    const int64_t item = item0 + h;
    const LoopNode<M> nd{lp.data, lp.stride, h};
    const double* A = src.base(item);
    const int lda = src.ldim();
    for (int J = 0; J < NE; ++J) {  // rows J >= n_sel pad an odd size
        int tj = 0;
        if (J < n_sel) src.fetch(item, J, tj);
        for (int K = 0; K < J; ++K) {
            int tk = 0;
            src.fetch(item, K, tk);
            nd.put(0, Lay::b(J, K, 0), J < n_sel ? ld_cd(A, (int64_t)tj * lda + tk) : czero());
#pragma unroll
            for (int c = 1; c < NP; ++c) nd.put(0, Lay::b(J, K, c), czero());
        }
        const cd lw = J < n_sel ? ld_cd(loops, loop_stride ? item * loop_stride + J : (int64_t)tj) : cd{1.0, 0.0};
        nd.put(0, Lay::l(0, J, 0), lw);
#pragma unroll
        for (int c = 1; c < NP; ++c) nd.put(0, Lay::l(0, J, c), czero());
    }
#pragma unroll
    for (int c = 0; c < NP; ++c) nd.put(0, Lay::g(0, c), c == 0 ? cd{1.0, 0.0} : czero());
    // depth-first walk; state of level L in bits 2L..2L+1: 0 fresh, 1 left child taken,
    // 2 right child taken.  A child returns its value in rv; w is the current node's sign.
    uint32_t states = 0;
    int L = 0;
    double w = 1.0;
    bool ret = false;
    cd rv = czero();
    for (;;) {
        const uint32_t stl = (states >> (2 * L)) & 3u;
        if (ret) {
            if (stl == 1u) {  // left child done: keep its value, take the right child
                nd.put(Lay::off(L), Lay::first(L), rv);
                w = -w;
                ret = false;
            } else {  // right child done: h + r, returned to the parent
                rv = cadd(nd.ld(Lay::off(L), Lay::first(L)), rv);
                if (L == 0) break;
                --L;
                continue;
            }
        }
        if (L == M) {  // leaf: w g[M]
            rv = cmul(cd{w, 0.0}, nd.ld(Lay::off(M), Lay::g(M, M)));
            ret = true;
            --L;
            continue;
        }
        if (stl == 0u) {
            nd.left(L);
            states |= 1u << (2 * L);
            ++L;
            states &= ~(3u << (2 * L));
            w = -w;
        } else {
            nd.right(L);
            states = (states & ~(3u << (2 * L))) | (2u << (2 * L));
            ++L;
            states &= ~(3u << (2 * L));
        }
    }
    src.store(item, rv);
}

// ---------------------------------------------------------------------------------------
// Host-side dispatch
// ---------------------------------------------------------------------------------------

constexpr int64_t kMaxGrid = 1 << 30;

template <class S>
int launch_lane(bsmps_ctx* ctx, const S& src, int m) {
    if (src.n_items == 0) return BSMPS_OK;
    const int64_t blocks = (src.n_items + kWave - 1) / kWave;
    for (int64_t b0 = 0; b0 < blocks; b0 += kMaxGrid) {
        const int64_t nb = std::min<int64_t>(kMaxGrid, blocks - b0);
        const int64_t item0 = b0 * kWave;
        dim3 grid((unsigned)nb), block(kWave);
        switch (m) {
            case 0: hipLaunchKernelGGL((haf_lane_kernel<0, S>), grid, block, 0, ctx->stream, src, item0); break;
            case 1: hipLaunchKernelGGL((haf_lane_kernel<1, S>), grid, block, 0, ctx->stream, src, item0); break;
            case 2: hipLaunchKernelGGL((haf_lane_kernel<2, S>), grid, block, 0, ctx->stream, src, item0); break;
            case 3: hipLaunchKernelGGL((haf_lane_kernel<3, S>), grid, block, 0, ctx->stream, src, item0); break;
            case 4: hipLaunchKernelGGL((haf_lane_kernel<4, S>), grid, block, 0, ctx->stream, src, item0); break;
            case 5: hipLaunchKernelGGL((haf_lane_kernel<5, S>), grid, block, 0, ctx->stream, src, item0); break;
            default:
                bsmps_set_error("lane hafnian path supports n <= 10, got n = %d", 2 * m);
                return BSMPS_ERR_INVALID;
        }
        BSMPS_HIP(hipGetLastError());
    }
    return BSMPS_OK;
}

template <class S>
int launch_group(bsmps_ctx* ctx, const S& src, int m) {
    if (src.n_items == 0) return BSMPS_OK;
    GroupPlan* plan = nullptr;
    BSMPS_TRY(get_plan(ctx->device, m, &plan));
    const int G = ctx->group_lanes_by_m[m];
    const int per = kWave / G;
    const size_t lds = sizeof(cd) * (size_t)plan->lay.total * per;
    if (lds > (size_t)ctx->max_lds) {
        bsmps_set_error("hafnian of size %d needs %zu bytes of LDS per workgroup with %d lanes per matrix "
                        "(device limit %d)", 2 * m, lds, G, ctx->max_lds);
        return BSMPS_ERR_UNSUPPORTED;
    }
    const int64_t blocks = (src.n_items + per - 1) / per;
    switch (m) {
#define BSMPS_GROUP_CASE(MM)                                                                                  \
    case MM: {                                                                                                \
        if (lds > 64 * 1024)                                                                                  \
            BSMPS_HIP(hipFuncSetAttribute((const void*)&haf_group_kernel<MM, S>,                       \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));             \
        for (int64_t b0 = 0; b0 < blocks; b0 += kMaxGrid) {                                                   \
            const int64_t nb = std::min<int64_t>(kMaxGrid, blocks - b0);                                      \
            hipLaunchKernelGGL((haf_group_kernel<MM, S>), dim3((unsigned)nb), dim3(kWave), lds, ctx->stream, \
                               src, b0 * per, plan->lay, plan->d_ops, G);                                     \
            BSMPS_HIP(hipGetLastError());                                                                     \
        }                                                                                                     \
        break;                                                                                                \
    }
        BSMPS_GROUP_CASE(3)
        BSMPS_GROUP_CASE(4)
        BSMPS_GROUP_CASE(5)
        BSMPS_GROUP_CASE(6)
        BSMPS_GROUP_CASE(7)
        BSMPS_GROUP_CASE(8)
        BSMPS_GROUP_CASE(9)
        BSMPS_GROUP_CASE(10)
        BSMPS_GROUP_CASE(11)
        BSMPS_GROUP_CASE(12)
#undef BSMPS_GROUP_CASE
        default:
            bsmps_set_error("group hafnian path supports sizes 6..24, got %d", 2 * m);
            return BSMPS_ERR_UNSUPPORTED;
    }
    return BSMPS_OK;
}

// Frontier-split plan of one size: storage slots of every owner node, the RIGHT steps grouped
// by tree level with their work items, the frontier nodes grouped by subtree shape, and the
// combine program.  A node is a frontier node when its subtree fits one lane: depth <= 4 with
// degree <= 1, depth <= 3 with degree <= 3, or depth <= 2.
struct BfsBuilder {
    int m;
    ScheduleBuilder nzb;
    int slots = 0;
    std::vector<BfsOp> ops;   // depth-first order
    std::vector<int4> front;  // {k | D << 8 | G << 16 | (w > 0) << 24, po, plo | pcap << 8, g slot}
    std::vector<int> comb;
    static bool is_frontier(int m, int level, int D) {
        const int depth = m - level;
        return (depth <= 4 && D <= 1) || (depth <= 3 && D <= 3) || depth <= 2;
    }
    void gen(int level, int po, int plo, int pcap, int gslot, int D, int G, int w) {
        if (is_frontier(m, level, D)) {
            comb.push_back((int)front.size());
            front.push_back(make_int4(level | (D << 8) | (G << 16) | ((w > 0) << 24), po, plo | (pcap << 8), gslot));
            return;
        }
        const int D2 = std::min(m, 2 * D + 1), G2 = std::min(m, G + D + 1);
        const bool lnz = nzb.nonzero(level + 1, D, G), rnz = nzb.nonzero(level + 1, D2, G2);
        if (lnz) gen(level + 1, po, plo, pcap, gslot, D, G, -w);
        if (rnz) {
            const int s0 = 2 * m - 2 * (level + 1);
            const int dcap = std::min(m - 1, D2);
            BfsOp op{level, D, D2, G, po, plo, pcap, gslot, slots, dcap, slots + s0 * (s0 - 1) / 2 * dcap, 0};
            slots += s0 * (s0 - 1) / 2 * dcap + m + 1;
            ops.push_back(op);
            gen(level + 1, op.pdst, level + 1, dcap, op.gdst, D2, G2, w);
        }
        if (lnz && rnz) comb.push_back(-1);
    }
};

// Subtree shapes with a compiled one-lane kernel: (K0, D) = (M-4, 0), (M-4, 1), (M-3, 3),
// (M-2, 7), (M-2, 15) and (M-2, M), for M = 6..17 (a node's degree is 2^r - 1 after r RIGHT
// steps, capped at M, so up to M = 17 every degree is 0, 1, 3, 7, 15 or M).
constexpr int kSplitMinM = 6, kSplitMaxM = BSMPS_MAX_HAFNIAN_SIZE / 2;
bool subtree_compiled(int M, int K0, int D) {
    if (M < kSplitMinM || M > kSplitMaxM) return false;
    return (K0 == M - 4 && (D == 0 || D == 1)) || (K0 == M - 3 && D == 3) ||
           (K0 == M - 2 && (D == M || (D == 7 && M > 7) || (D == 15 && M > 15)));
}

struct FrontierPlan {
    bool ok;
    int slots;                       // complex slots per hafnian (A, root g, owner blocks)
    int a_off, rootg;
    BfsOp* d_ops;
    int2* d_items;                   // (op, entry or -1), grouped by level
    std::vector<int> level_first, level_count, level_degree;  // item groups of one (level, degree)
    int4* d_front;
    int n_front;
    int* d_tlist;
    std::vector<int2> types;         // (K0, D)
    std::vector<int> type_first, type_count;
    int* d_comb;
    int n_comb;
    int4* d_parts;                   // combine parts {first step, steps, sum slot}
    int n_parts;
    int* d_top;                      // combine program above the parts (d_comb when none)
    int n_top;
};

// Parts of a combine program: complete subtrees of its add tree of at most kCombinePartSteps
// steps, found by descending from the root; "top" is the program of the tree above them, which
// pushes each part's sum from the slot of the part's first value (a subtree's program starts
// with its leftmost value).  Programs of at most 2 kCombinePartSteps steps are not cut.
constexpr int kCombinePartSteps = 48;
struct CombineCut {
    struct Node {
        int start, end, left, right;  // program range [start, end); children of an add
    };
    const std::vector<int>& prog;
    std::vector<Node> nodes;
    std::vector<int4> parts;
    std::vector<int> top;
    explicit CombineCut(const std::vector<int>& c) : prog(c) {
        std::vector<int> st;
        for (int i = 0; i < (int)prog.size(); ++i) {
            if (prog[i] >= 0) {
                nodes.push_back({i, i + 1, -1, -1});
            } else {
                const int r = st.back();
                st.pop_back();
                const int l = st.back();
                st.pop_back();
                nodes.push_back({nodes[l].start, i + 1, l, r});
            }
            st.push_back((int)nodes.size() - 1);
        }
        if (prog.size() > 2 * (size_t)kCombinePartSteps && st.size() == 1) cut(st.back());
    }
    void cut(int x) {
        const Node c = nodes[x];
        if (c.end - c.start <= kCombinePartSteps || c.left < 0) {
            if (c.end - c.start > 1) parts.push_back(make_int4(c.start, c.end - c.start, prog[c.start], 0));
            top.push_back(prog[c.start]);
            return;
        }
        cut(c.left);
        cut(c.right);
        top.push_back(-1);
    }
};

std::map<std::pair<int, int>, FrontierPlan> g_fplans;  // (device, m)

int get_fplan(int device, int m, FrontierPlan** out) {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto key = std::make_pair(device, m);
    auto it = g_fplans.find(key);
    if (it == g_fplans.end()) {
        BfsBuilder fb;
        fb.m = m;
        fb.nzb.m = m;
        FrontierPlan p{};
        p.ok = true;
        p.a_off = 0;
        p.rootg = m * (2 * m - 1);
        fb.slots = p.rootg + m + 1;
        fb.gen(0, -1, 0, 0, p.rootg, 0, 0, 1);
        p.slots = fb.slots;
        // work items of each level: every child entry of every RIGHT step, then its g update
        std::vector<int2> items;
        // (a level's items are independent, so its degree groups may run one after another)
        const int degrees[6] = {m, 0, 1, 3, 7, 15};
        for (int l = 0; l < m; ++l) {
            for (int di = 0; di < 6; ++di) {
                const int D = degrees[di];
                if (di > 0 && D == m) continue;  // M coincides with 7 or 15
                const int first = (int)items.size();
                for (size_t o = 0; o < fb.ops.size(); ++o) {
                    if (fb.ops[o].level != l || fb.ops[o].D != D) continue;
                    const int s0 = 2 * m - 2 * (l + 1);
                    for (int e = 0; e < s0 * (s0 - 1) / 2; ++e) items.push_back(make_int2((int)o, e));
                    items.push_back(make_int2((int)o, -1));
                }
                if ((int)items.size() > first) {
                    p.level_first.push_back(first);
                    p.level_count.push_back((int)items.size() - first);
                    p.level_degree.push_back(D);
                }
            }
        }
        for (const BfsOp& op : fb.ops)
            if (!(op.D == 0 || op.D == 1 || op.D == 3 || op.D == 7 || op.D == 15 || op.D == m)) p.ok = false;
        size_t n_level_items = 0;
        for (int c : p.level_count) n_level_items += (size_t)c;
        if (n_level_items != items.size()) p.ok = false;
        p.n_front = (int)fb.front.size();
        std::map<std::pair<int, int>, std::vector<int>> by_type;
        for (int f = 0; f < p.n_front; ++f) {
            const int4 fr = fb.front[f];
            const int k = fr.x & 0xff, D = (fr.x >> 8) & 0xff, G = (fr.x >> 16) & 0xff;
            if (G != D || !subtree_compiled(m, k, D)) p.ok = false;
            by_type[std::make_pair(k, D)].push_back(f);
        }
        std::vector<int> tlist;
        for (auto& kv : by_type) {
            p.types.push_back(make_int2(kv.first.first, kv.first.second));
            p.type_first.push_back((int)tlist.size());
            p.type_count.push_back((int)kv.second.size());
            tlist.insert(tlist.end(), kv.second.begin(), kv.second.end());
        }
        p.n_comb = (int)fb.comb.size();
        CombineCut cc(fb.comb);
        p.n_parts = (int)cc.parts.size();
        p.n_top = p.n_parts ? (int)cc.top.size() : p.n_comb;
        for (const std::vector<int>* prog : {&fb.comb, &cc.top}) {
            int dep = 0, mx = 0;
            for (int c : *prog) {
                dep += c >= 0 ? 1 : -1;
                mx = std::max(mx, dep);
            }
            if (!prog->empty() && (mx > kCombineStack || dep != 1)) p.ok = false;
        }
        int depth = 0, max_depth = 0;
        for (int c : fb.comb) {
            depth += c >= 0 ? 1 : -1;
            max_depth = std::max(max_depth, depth);
        }
        if (max_depth > kCombineStack || depth != 1) p.ok = false;
        auto upload = [](auto** dst, const auto& v) -> int {
            const size_t bytes = sizeof(v[0]) * std::max<size_t>(v.size(), 1);
            BSMPS_HIP(hipMalloc(reinterpret_cast<void**>(dst), bytes));
            if (!v.empty()) BSMPS_HIP(hipMemcpy(*dst, v.data(), sizeof(v[0]) * v.size(), hipMemcpyHostToDevice));
            return BSMPS_OK;
        };
        BSMPS_TRY(upload(&p.d_ops, fb.ops));
        BSMPS_TRY(upload(&p.d_items, items));
        BSMPS_TRY(upload(&p.d_front, fb.front));
        BSMPS_TRY(upload(&p.d_tlist, tlist));
        BSMPS_TRY(upload(&p.d_comb, fb.comb));
        BSMPS_TRY(upload(&p.d_parts, cc.parts));
        if (p.n_parts) BSMPS_TRY(upload(&p.d_top, cc.top));
        else p.d_top = p.d_comb;
        it = g_fplans.emplace(key, p).first;
    }
    *out = &it->second;
    return BSMPS_OK;
}

template <int M>
int launch_level_m(hipStream_t st, int D, int n_items, int n_hblocks, const BfsData& bd, const BfsOp* ops,
                   const int2* items) {
    const dim3 grid((unsigned)xcd_grid(n_items, n_hblocks));
#define BSMPS_LEVEL(DD) \
    hipLaunchKernelGGL((haf_bfs_level_kernel<M, DD>), grid, dim3(kWave), 0, st, bd, ops, items, n_items, n_hblocks)
    if (D == M) BSMPS_LEVEL(M);
    else if (D == 0) BSMPS_LEVEL(0);
    else if (D == 1) BSMPS_LEVEL(1);
    else if (D == 3 && M > 3) BSMPS_LEVEL((M > 3 ? 3 : M));
    else if (D == 7 && M > 7) BSMPS_LEVEL((M > 7 ? 7 : M));
    else if (D == 15 && M > 15) BSMPS_LEVEL((M > 15 ? 15 : M));
    else {
        bsmps_set_error("no level kernel for size %d, degree %d", 2 * M, D);
        return BSMPS_ERR_UNSUPPORTED;
    }
#undef BSMPS_LEVEL
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

template <int M>
int launch_subtrees_m(hipStream_t st, int K0, int D, int n_front, int n_hblocks, const BfsData& bd,
                      const int4* front, const int* tlist, double* vals) {
    const dim3 grid((unsigned)xcd_grid(n_front, n_hblocks));
#define BSMPS_SUBTREE(KK, DD)                                                                              \
    hipLaunchKernelGGL((haf_subtree_kernel<M, KK, DD>), grid, dim3(kWave), 0, st, bd, front, tlist, vals, \
                       n_front, n_hblocks)
    if (K0 == M - 4 && D == 0) BSMPS_SUBTREE(M - 4, 0);
    else if (K0 == M - 4 && D == 1) BSMPS_SUBTREE(M - 4, 1);
    else if (K0 == M - 3 && D == 3) BSMPS_SUBTREE(M - 3, 3);
    else if (K0 == M - 2 && D == M) BSMPS_SUBTREE(M - 2, M);
    else if (K0 == M - 2 && D == 7 && M > 7) BSMPS_SUBTREE(M - 2, (M > 7 ? 7 : M));
    else if (K0 == M - 2 && D == 15 && M > 15) BSMPS_SUBTREE(M - 2, (M > 15 ? 15 : M));
    else {
        bsmps_set_error("no subtree kernel for size %d, level %d, degree %d", 2 * M, K0, D);
        return BSMPS_ERR_UNSUPPORTED;
    }
#undef BSMPS_SUBTREE
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

// Frontier-split launch: per chunk of hafnians (sized so its HBM storage fits the budget),
// the gather, one launch per upper-tree level, one per subtree shape, then the combine.
template <class S>
int launch_split(bsmps_ctx* ctx, const S& src, int m, FrontierPlan* fp) {
    const size_t per_h = sizeof(cd) * (size_t)(fp->slots + fp->n_front);
    int64_t H = (int64_t)(ctx->split_chunk_bytes / per_h);
    H = std::max<int64_t>(kWave, H / kWave * kWave);
    H = std::min<int64_t>(H, (src.n_items + kWave - 1) / kWave * kWave);
    // one launch covers (items of a level or subtree shape) x (hafnian blocks) workgroups of
    // 64 lanes, which must stay below 2^32 threads
    int max_items = 1;
    for (int c : fp->level_count) max_items = std::max(max_items, c);
    for (int c : fp->type_count) max_items = std::max(max_items, c);
    max_items = std::max(max_items, fp->n_parts);
    const int64_t max_hblocks = (((int64_t)1 << 32) / kWave - kXcds) / max_items;
    if (max_hblocks < 1) {
        bsmps_set_error("frontier-split hafnian of size %d: %d work items per launch exceed the grid limit", 2 * m,
                        max_items);
        return BSMPS_ERR_UNSUPPORTED;
    }
    H = std::min<int64_t>(H, max_hblocks * kWave);
    void* scratch = nullptr;
    BSMPS_TRY(bsmps_ctx_scratch(ctx, per_h * (size_t)H, &scratch));
    double* data = reinterpret_cast<double*>(scratch);
    double* vals = data + 2 * (size_t)fp->slots * H;
    for (int64_t c0 = 0; c0 < src.n_items; c0 += H) {
        const int64_t nh = std::min<int64_t>(H, src.n_items - c0);
        const unsigned hblocks = (unsigned)((nh + kWave - 1) / kWave);
        BfsData bd{data, H, nh, fp->a_off};
        switch (m) {
#define BSMPS_SPLIT_CASE(MM)                                                                                  \
    case MM:                                                                                                  \
        hipLaunchKernelGGL((haf_bfs_prep_kernel<MM, S>), dim3(hblocks), dim3(kWave), 0, ctx->stream, src, c0, bd, \
                           fp->rootg);                                                                        \
        BSMPS_HIP(hipGetLastError());                                                                         \
        for (size_t l = 0; l < fp->level_first.size(); ++l)                                                   \
            BSMPS_TRY(launch_level_m<MM>(ctx->stream, fp->level_degree[l], fp->level_count[l], (int)hblocks, bd, \
                                         fp->d_ops, fp->d_items + fp->level_first[l]));                       \
        for (size_t t = 0; t < fp->types.size(); ++t)                                                         \
            BSMPS_TRY(launch_subtrees_m<MM>(ctx->stream, fp->types[t].x, fp->types[t].y, fp->type_count[t],   \
                                            (int)hblocks, bd, fp->d_front, fp->d_tlist + fp->type_first[t],   \
                                            vals));                                                           \
        break;
            BSMPS_SPLIT_CASE(6)
            BSMPS_SPLIT_CASE(7)
            BSMPS_SPLIT_CASE(8)
            BSMPS_SPLIT_CASE(9)
            BSMPS_SPLIT_CASE(10)
            BSMPS_SPLIT_CASE(11)
            BSMPS_SPLIT_CASE(12)
            BSMPS_SPLIT_CASE(13)
            BSMPS_SPLIT_CASE(14)
            BSMPS_SPLIT_CASE(15)
            BSMPS_SPLIT_CASE(16)
            BSMPS_SPLIT_CASE(17)
#undef BSMPS_SPLIT_CASE
            default:
                bsmps_set_error("frontier-split hafnian path supports sizes 12..%d, got %d", BSMPS_MAX_HAFNIAN_SIZE, 2 * m);
                return BSMPS_ERR_INVALID;
        }
        if (fp->n_parts) {
            hipLaunchKernelGGL(haf_combine_part_kernel, dim3((unsigned)xcd_grid(fp->n_parts, (int)hblocks)),
                               dim3(kWave), 0, ctx->stream, nh, H, vals, fp->d_comb, fp->d_parts, fp->n_parts,
                               (int)hblocks);
            BSMPS_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL((haf_combine_kernel<S>), dim3(hblocks), dim3(kWave), 0, ctx->stream, src, c0, nh, H, vals,
                           fp->d_top, fp->n_top);
        BSMPS_HIP(hipGetLastError());
    }
    return BSMPS_OK;
}

template <class S>
int launch_hafnians(bsmps_ctx* ctx, const S& src, int n) {
    if (n % 2) return BSMPS_OK;  // odd sizes are zero; callers zero-fill
    const int m = n / 2;
    if (m <= ctx->lane_max_m || m <= 2) return launch_lane(ctx, src, m);
    if (m >= kSplitMinM && m <= kSplitMaxM && ctx->split_by_m[m] && src.n_items > 0) {
        FrontierPlan* fp = nullptr;
        BSMPS_TRY(get_fplan(ctx->device, m, &fp));
        if (fp->ok) return launch_split(ctx, src, m, fp);
    }
    return launch_group(ctx, src, m);
}

__global__ void sigma_select_kernel(const double* sigma, int ld, const int32_t* targets, int64_t n_batch,
                                    int n_select, double* out) {
    const int64_t per = (int64_t)n_select * n_select;
    const int64_t total = n_batch * per;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t l = i / per;
        const int rc = (int)(i - l * per);
        const int r = rc / n_select, c = rc - r * n_select;
        const int tr = targets[l * n_select + r], tc = targets[l * n_select + c];
        reinterpret_cast<double2*>(out)[i] = reinterpret_cast<const double2*>(sigma)[(int64_t)tr * ld + tc];
    }
}

__global__ void fill_kernel(double* out, int64_t n, double re, double im) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        reinterpret_cast<double2*>(out)[i] = make_double2(re, im);
}

int fill_complex(bsmps_ctx* ctx, double* out, int64_t n, double re, double im) {
    if (n == 0) return BSMPS_OK;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, out, n, re, im);
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

template <int M, class S>
int launch_loop_m(bsmps_ctx* ctx, const S& src, const double* loops, int loop_stride, int n_sel) {
    using Lay = LoopLayout<M>;
    const size_t per_h = sizeof(cd) * (size_t)Lay::total();
    int64_t H = (int64_t)(ctx->split_chunk_bytes / per_h);
    H = std::max<int64_t>(kWave, H / kWave * kWave);
    H = std::min<int64_t>(H, (src.n_items + kWave - 1) / kWave * kWave);
    void* scratch = nullptr;
    BSMPS_TRY(bsmps_ctx_scratch(ctx, per_h * (size_t)H, &scratch));
    for (int64_t c0 = 0; c0 < src.n_items; c0 += H) {
        const int64_t nh = std::min<int64_t>(H, src.n_items - c0);
        LoopData lp{reinterpret_cast<double*>(scratch), H, nh};
        hipLaunchKernelGGL((haf_loop_kernel<M, S>), dim3((unsigned)((nh + kWave - 1) / kWave)), dim3(kWave), 0,
                           ctx->stream, src, loops, loop_stride, n_sel, c0, lp);
        BSMPS_HIP(hipGetLastError());
    }
    return BSMPS_OK;
}

template <class S>
int launch_loop(bsmps_ctx* ctx, const S& src, const double* loops, int loop_stride, int n_sel) {
    switch ((n_sel + 1) / 2) {
#define BSMPS_LOOP_CASE(MM) \
    case MM:                \
        return launch_loop_m<MM, S>(ctx, src, loops, loop_stride, n_sel);
        BSMPS_LOOP_CASE(1)
        BSMPS_LOOP_CASE(2)
        BSMPS_LOOP_CASE(3)
        BSMPS_LOOP_CASE(4)
        BSMPS_LOOP_CASE(5)
        BSMPS_LOOP_CASE(6)
        BSMPS_LOOP_CASE(7)
        BSMPS_LOOP_CASE(8)
        BSMPS_LOOP_CASE(9)
        BSMPS_LOOP_CASE(10)
        BSMPS_LOOP_CASE(11)
        BSMPS_LOOP_CASE(12)
#undef BSMPS_LOOP_CASE
        default:
            bsmps_set_error("loop hafnian of size %d: sizes 1..%d are supported", n_sel, BSMPS_MAX_LOOP_HAFNIAN_SIZE);
            return BSMPS_ERR_INVALID;
    }
}

}  // namespace

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------

extern "C" int bsmps_sigma_select(bsmps_ctx* ctx, const double* sigma, int ld, const int32_t* targets,
                                  int64_t n_batch, int n_select, double* out) {
    if (!ctx || n_batch < 0 || n_select < 0 || ld <= 0) {
        bsmps_set_error("bsmps_sigma_select: invalid arguments");
        return BSMPS_ERR_INVALID;
    }
    if (n_batch == 0 || n_select == 0) return BSMPS_OK;
    if (!sigma || !targets || !out) {
        bsmps_set_error("bsmps_sigma_select: null pointer");
        return BSMPS_ERR_INVALID;
    }
    const int64_t total = n_batch * (int64_t)n_select * n_select;
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(sigma_select_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, sigma, ld,
                       targets, n_batch, n_select, out);
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

extern "C" int bsmps_hafnian_dense(bsmps_ctx* ctx, const double* A, int64_t n_batch, int n, double* out) {
    if (!ctx || n_batch < 0 || n < 0 || n > BSMPS_MAX_HAFNIAN_SIZE) {
        bsmps_set_error("bsmps_hafnian_dense: invalid arguments (n = %d, max %d)", n, BSMPS_MAX_HAFNIAN_SIZE);
        return BSMPS_ERR_INVALID;
    }
    if (n_batch == 0) return BSMPS_OK;
    if (!out || (n > 0 && !A)) {
        bsmps_set_error("bsmps_hafnian_dense: null pointer");
        return BSMPS_ERR_INVALID;
    }
    if (n == 0) return fill_complex(ctx, out, n_batch, 1.0, 0.0);
    if (n % 2) return fill_complex(ctx, out, n_batch, 0.0, 0.0);
    DenseSource s{A, n, out, n_batch};
    return launch_hafnians(ctx, s, n);
}

extern "C" int bsmps_hafnian_targets(bsmps_ctx* ctx, const double* sigma, int ld, const int32_t* targets,
                                     int64_t n_batch, int n_select, const double* den, double* out) {
    if (!ctx || n_batch < 0 || n_select < 0 || n_select > BSMPS_MAX_HAFNIAN_SIZE || ld <= 0) {
        bsmps_set_error("bsmps_hafnian_targets: invalid arguments (n_select = %d, max %d)", n_select,
                        BSMPS_MAX_HAFNIAN_SIZE);
        return BSMPS_ERR_INVALID;
    }
    if (n_batch == 0) return BSMPS_OK;
    if (!out || !sigma || (n_select > 0 && !targets)) {
        bsmps_set_error("bsmps_hafnian_targets: null pointer");
        return BSMPS_ERR_INVALID;
    }
    if (n_select % 2) return fill_complex(ctx, out, n_batch, 0.0, 0.0);
    TargetsSource s{sigma, ld, targets, n_select, den, out, n_batch};
    return launch_hafnians(ctx, s, n_select);
}

extern "C" int bsmps_loop_hafnian_dense(bsmps_ctx* ctx, const double* A, const double* loops, int64_t n_batch,
                                        int n, double* out) {
    if (!loops) return bsmps_hafnian_dense(ctx, A, n_batch, n, out);
    if (!ctx || n_batch < 0 || n < 0 || n > BSMPS_MAX_LOOP_HAFNIAN_SIZE) {
        bsmps_set_error("bsmps_loop_hafnian_dense: invalid arguments (n = %d, max %d)", n,
                        BSMPS_MAX_LOOP_HAFNIAN_SIZE);
        return BSMPS_ERR_INVALID;
    }
    if (n_batch == 0) return BSMPS_OK;
    if (!out || (n > 0 && !A)) {
        bsmps_set_error("bsmps_loop_hafnian_dense: null pointer");
        return BSMPS_ERR_INVALID;
    }
    if (n == 0) return fill_complex(ctx, out, n_batch, 1.0, 0.0);
    DenseSource s{A, n, out, n_batch};
    return launch_loop(ctx, s, loops, n, n);
}

extern "C" int bsmps_loop_hafnian_targets(bsmps_ctx* ctx, const double* sigma, int ld, const double* loops,
                                          const int32_t* targets, int64_t n_batch, int n_select, const double* den,
                                          double* out) {
    if (!loops) return bsmps_hafnian_targets(ctx, sigma, ld, targets, n_batch, n_select, den, out);
    if (!ctx || n_batch < 0 || n_select < 0 || n_select > BSMPS_MAX_LOOP_HAFNIAN_SIZE || ld <= 0) {
        bsmps_set_error("bsmps_loop_hafnian_targets: invalid arguments (n_select = %d, max %d)", n_select,
                        BSMPS_MAX_LOOP_HAFNIAN_SIZE);
        return BSMPS_ERR_INVALID;
    }
    if (n_batch == 0) return BSMPS_OK;
    if (!out || !sigma || (n_select > 0 && !targets)) {
        bsmps_set_error("bsmps_loop_hafnian_targets: null pointer");
        return BSMPS_ERR_INVALID;
    }
    // the empty matrix has no loops: its value (1, then the epilogue) is the plain hafnian's
    if (n_select == 0) return bsmps_hafnian_targets(ctx, sigma, ld, targets, n_batch, n_select, den, out);
    TargetsSource s{sigma, ld, targets, n_select, den, out, n_batch};
    return launch_loop(ctx, s, loops, 0, n_select);
}

namespace {

// Checks one site's arguments and derives its class offsets and counts.
int gamma_site_tables(const bsmps_gamma_site_args* a, const char* who, std::vector<int32_t>& loff,
                      std::vector<int32_t>& roff, std::vector<int32_t>& rcnt, int& max_n) {
    if (!a) {
        bsmps_set_error("%s: null argument", who);
        return BSMPS_ERR_INVALID;
    }
    if (a->kind < 0 || a->kind > 2 || a->d <= 0 || a->chi <= 0 || a->ld <= 0 || a->n_left < 0 || a->n_right < 0) {
        bsmps_set_error("%s: invalid sizes (kind %d, d %d, chi %d, ld %d, n_left %d, n_right %d)", who,
                        a->kind, a->d, a->chi, a->ld, a->n_left, a->n_right);
        return BSMPS_ERR_INVALID;
    }
    if (a->n_left > a->chi) {
        bsmps_set_error("%s: %d left labels exceed chi = %d", who, a->n_left, a->chi);
        return BSMPS_ERR_INVALID;
    }
    if (a->n_left_class <= 0 || a->n_right_class <= 0) {
        bsmps_set_error("%s: empty photon classes (%d left, %d right)", who, a->n_left_class, a->n_right_class);
        return BSMPS_ERR_INVALID;
    }
    if (!a->sigma || !a->sqrt_fact || !a->left_den || !a->lambda || !a->left_perm || !a->right_den ||
        !a->right_perm || !a->left_class_count_host || !a->right_class_count_host) {
        bsmps_set_error("%s: null pointer (sigma %p sqrt_fact %p left_den %p lambda %p left_perm %p "
                        "right_den %p right_perm %p)", who, (const void*)a->sigma, (const void*)a->sqrt_fact,
                        (const void*)a->left_den, (const void*)a->lambda, (const void*)a->left_perm,
                        (const void*)a->right_den, (const void*)a->right_perm);
        return BSMPS_ERR_INVALID;
    }
    const int nlc = a->n_left_class, nrc = a->n_right_class;
    if (nlc - 1 > a->w_left || nrc - 1 > a->w_right) {
        bsmps_set_error("%s: photon classes exceed target widths", who);
        return BSMPS_ERR_INVALID;
    }
    loff.assign(nlc + 1, 0);
    roff.assign(nrc + 1, 0);
    rcnt.assign(nrc, 0);
    int64_t lsum = 0, rsum = 0;
    for (int p = 0; p < nlc; ++p) {
        if (a->left_class_count_host[p] < 0) {
            bsmps_set_error("%s: negative class count", who);
            return BSMPS_ERR_INVALID;
        }
        loff[p + 1] = loff[p] + a->left_class_count_host[p];
        lsum += a->left_class_count_host[p];
    }
    for (int q = 0; q < nrc; ++q) {
        if (a->right_class_count_host[q] < 0) {
            bsmps_set_error("%s: negative class count", who);
            return BSMPS_ERR_INVALID;
        }
        roff[q + 1] = roff[q] + a->right_class_count_host[q];
        rcnt[q] = a->right_class_count_host[q];
        rsum += a->right_class_count_host[q];
    }
    if (lsum != a->n_left || rsum != a->n_right) {
        bsmps_set_error("%s: class counts do not add up to the label counts", who);
        return BSMPS_ERR_INVALID;
    }
    max_n = (a->d - 1) + (nlc - 1) + (nrc - 1);
    if (max_n > BSMPS_MAX_HAFNIAN_SIZE + 1) {
        bsmps_set_error("%s: hafnians up to size %d exceed the supported %d", who, max_n, BSMPS_MAX_HAFNIAN_SIZE);
        return BSMPS_ERR_UNSUPPORTED;
    }
    return BSMPS_OK;
}

// Site tensors of any number of sites: for each even photon total n (largest first) one hafnian
// launch covers the (j, p, q) rectangles of every site, numbered site after site.  The tables go
// to the context workspace as [GammaSite x n_sites][GammaRect...][int32 class tables].
int gamma_sites(bsmps_ctx* ctx, int n_sites, const bsmps_gamma_site_args* args, double* const* gammas,
                const char* who) {
    if (!ctx || n_sites < 0 || (n_sites > 0 && (!args || !gammas))) {
        bsmps_set_error("%s: null argument", who);
        return BSMPS_ERR_INVALID;
    }
    std::vector<std::vector<int32_t>> loff(n_sites), roff(n_sites), rcnt(n_sites);
    int max_n = -1;
    for (int i = 0; i < n_sites; ++i) {
        int mn = 0;
        BSMPS_TRY(gamma_site_tables(&args[i], who, loff[i], roff[i], rcnt[i], mn));
        if (!gammas[i]) {
            bsmps_set_error("%s: null output of site %d", who, i);
            return BSMPS_ERR_INVALID;
        }
        max_n = std::max(max_n, mn);
    }
    if (max_n < 0) return BSMPS_OK;
    // rectangles grouped by even total n, each group numbered site after site
    std::vector<std::vector<GammaRect>> by_n(max_n + 1);
    std::vector<int64_t> n_items(max_n + 1, 0);
    for (int n = 0; n <= max_n; n += 2) {
        int64_t start = 0;
        for (int i = 0; i < n_sites; ++i) {
            const bsmps_gamma_site_args* a = &args[i];
            for (int j = 0; j < a->d && j <= n; ++j)
                for (int p = 0; p < a->n_left_class && j + p <= n; ++p) {
                    const int q = n - j - p;
                    if (q >= a->n_right_class) continue;
                    const int64_t cnt = (int64_t)a->left_class_count_host[p] * a->right_class_count_host[q];
                    if (cnt == 0) continue;
                    by_n[n].push_back(GammaRect{start, j, p, q, i});
                    start += cnt;
                }
        }
        n_items[n] = start;
    }
    size_t n_rects = 0, n_tab = 0;
    for (const auto& v : by_n) n_rects += v.size();
    for (int i = 0; i < n_sites; ++i) n_tab += loff[i].size() + roff[i].size() + rcnt[i].size();
    const size_t bytes_sites = sizeof(GammaSite) * (size_t)n_sites;
    const size_t bytes_rects = sizeof(GammaRect) * std::max<size_t>(n_rects, 1);
    const size_t bytes_tab = sizeof(int32_t) * n_tab;
    void* ws = nullptr;
    BSMPS_TRY(bsmps_ctx_workspace(ctx, bytes_sites + bytes_rects + bytes_tab + 256, &ws));
    char* dbase = reinterpret_cast<char*>(ws);
    std::vector<char> host(bytes_sites + bytes_rects + bytes_tab);
    GammaSite* hs = reinterpret_cast<GammaSite*>(host.data());
    int32_t* htab = reinterpret_cast<int32_t*>(host.data() + bytes_sites + bytes_rects);
    const int32_t* dtab = reinterpret_cast<const int32_t*>(dbase + bytes_sites + bytes_rects);
    size_t t = 0;
    for (int i = 0; i < n_sites; ++i) {
        const bsmps_gamma_site_args* a = &args[i];
        GammaSite& g = hs[i];
        g.sigma = a->sigma;
        g.sqrt_fact = a->sqrt_fact;
        g.lt = a->left_targets;
        g.lden = a->left_den;
        g.lam = a->lambda;
        g.lperm = a->left_perm;
        g.rt = a->right_targets;
        g.rden = a->right_den;
        g.rperm = a->right_perm;
        g.gamma = gammas[i];
        g.z = a->z;
        g.inv_z = a->inv_z;
        g.ld = a->ld;
        g.d = a->d;
        g.chi = a->chi;
        g.wl = a->w_left;
        g.wr = a->w_right;
        g.vacuum_real_div = a->kind != BSMPS_SITE_FIRST;
        const std::vector<int32_t>* tabs[3] = {&loff[i], &roff[i], &rcnt[i]};
        const int32_t** dst[3] = {&g.loff, &g.roff, &g.rcnt};
        for (int k = 0; k < 3; ++k) {
            *dst[k] = dtab + t;
            memcpy(htab + t, tabs[k]->data(), sizeof(int32_t) * tabs[k]->size());
            t += tabs[k]->size();
        }
    }
    GammaRect* hr = reinterpret_cast<GammaRect*>(host.data() + bytes_sites);
    const GammaRect* dr = reinterpret_cast<const GammaRect*>(dbase + bytes_sites);
    std::vector<size_t> n_first(max_n + 1, 0);
    size_t r = 0;
    for (int n = 0; n <= max_n; ++n) {
        n_first[n] = r;
        for (const GammaRect& x : by_n[n]) hr[r++] = x;
    }
    BSMPS_TRY(bsmps_ctx_upload(ctx, host.data(), host.size(), ws));
    const GammaSite* ds = reinterpret_cast<const GammaSite*>(dbase);
    // largest sizes first: they are the longest-running launches
    for (int n = max_n; n >= 0; --n) {
        if (n % 2 || n_items[n] == 0) continue;
        GammaSource src{ds, dr + n_first[n], (int)by_n[n].size(), n_items[n]};
        BSMPS_TRY(launch_hafnians(ctx, src, n));
    }
    return BSMPS_OK;
}

}  // namespace

extern "C" int bsmps_gamma_site(bsmps_ctx* ctx, const bsmps_gamma_site_args* a, double* gamma) {
    if (!ctx || !a || !gamma) {
        bsmps_set_error("bsmps_gamma_site: null argument");
        return BSMPS_ERR_INVALID;
    }
    return gamma_sites(ctx, 1, a, &gamma, "bsmps_gamma_site");
}

extern "C" int bsmps_gamma_sites(bsmps_ctx* ctx, int n_sites, const bsmps_gamma_site_args* args,
                                 double* const* gammas) {
    return gamma_sites(ctx, n_sites, args, gammas, "bsmps_gamma_sites");
}
