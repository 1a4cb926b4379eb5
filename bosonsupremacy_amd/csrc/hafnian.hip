// Batched hafnians for the MPS site tensors, gfx950.
//
// The arithmetic follows the reference recursion (MPS_cpu.py:117-187: hafnian dispatch,
// recursive_hafnian, solve) operation for operation, with NumPy's complex128 semantics, so
// results are bit-identical to the reference's complex128 evaluation.  Terms that are
// structurally zero (truncated-polynomial coefficients above the degree a node can have,
// and leaves whose polynomial cannot reach degree n/2) are skipped: adding them only ever
// adds +-0, which leaves every non-zero value unchanged.
//
// Two launch shapes:
//   * lane path   (n <= 10): one lane per matrix; the recursion is unrolled at compile time
//                 and the polynomial stack lives in VGPRs.  The gathered matrix is staged
//                 in LDS, laid out [entry][lane] so each wave reads it conflict-free.
//   * group path  (n >= 6):  G = 16/32/64 lanes per matrix follow a precomputed depth-first
//                 schedule of the recursion tree; each node's child matrix is built with one
//                 lane per entry, the polynomial stack lives in LDS.
// The matrices never exist in HBM: each lane gathers its entries from Sigma (L2-resident)
// through its target indices.
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

struct GammaSource {  // site tensor entries Gamma[b, a, j] of one photon total n
    const double* sigma;
    int ld;
    int d, chi;
    int vacuum_real_div;
    double z, inv_z;
    const double* sqrt_fact;
    const int32_t* lt;
    int wl;
    const double* lden;
    const double* lam;
    const int32_t* lperm;
    const int32_t* loff;  // class offsets into lperm
    const int32_t* rt;
    int wr;
    const double* rden;
    const int32_t* rperm;
    const int32_t* roff;
    const int32_t* rcnt;  // class counts (right)
    const int4* rects;    // {j, p, q, start} for this launch, sorted by start
    int n_rects;
    double* gamma;
    int64_t n_items;

    struct Decoded {
        int j, p, q, a, b;
    };
    __device__ Decoded decode(int64_t item) const {
        int lo = 0, hi = n_rects - 1;
        while (lo < hi) {  // last rect with start <= item
            int mid = (lo + hi + 1) >> 1;
            if (rects[mid].w <= item) lo = mid;
            else hi = mid - 1;
        }
        const int4 r = rects[lo];
        const int64_t off = item - r.w;
        const int cr = rcnt[r.z];
        const int64_t ia = off / cr;
        const int ib = (int)(off - ia * cr);
        Decoded dc;
        dc.j = r.x;
        dc.p = r.y;
        dc.q = r.z;
        dc.a = lperm[loff[r.y] + ia];
        dc.b = rperm[roff[r.z] + ib];
        return dc;
    }
    __device__ void fetch_dec(const Decoded& dc, int k, int& tk) const {
        if (k < dc.j) tk = 0;
        else if (k < dc.j + dc.p) tk = lt[(int64_t)dc.a * wl + (k - dc.j)];
        else tk = rt[(int64_t)dc.b * wr + (k - dc.j - dc.p)];
    }
    __device__ const double* base(int64_t) const { return sigma; }
    __device__ int ldim() const { return ld; }
    __device__ void store_dec(const Decoded& dc, cd h) const {
        cd v;
        if (vacuum_real_div && dc.j + dc.p + dc.q == 0) {
            v.re = (1.0 / z) / lam[dc.a];
            v.im = 0.0;
        } else {
            const double den = (sqrt_fact[dc.j] * lden[dc.a]) * rden[dc.b];
            v = cscale(cscale(cscale(h, 1.0 / den), inv_z), 1.0 / lam[dc.a]);
        }
        st_cd(gamma, ((int64_t)dc.b * chi + dc.a) * d + dc.j, v);
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
        const cd h = lane_hafnian<M>(t, src.base(item), src.ldim(), lds + lane, kWave);
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
        const int cap = std::min(m, (1 << std::min(l, 20)) - 1);
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

// Right-branch child-matrix entry (J, K): all coefficients t = 1..D2 of
//   c(J,K)[t] = b(J,K)[t] + sum_u (b(J,b0)[u] b(K,b1)[v] + b(K,b0)[u] b(J,b1)[v]),  v = t-1-u,
// accumulated in ascending u as solve() does (MPS_cpu.py:179-186).  The node's four pivot
// polynomials are loaded once into registers; D (the node's degree) and D2 are compile-time,
// which is possible because a node's degree is always 2^r - 1 or m.
template <int D, int D2>
__device__ __forceinline__ void child_entry(const cd* Ar, const cd* So, int obase, int ocap, cd* dst, int J,
                                            int K, int b0, int b1) {
    auto co = [&](int r, int c, int u) -> cd {
        return u == 0 ? Ar[tri(r, c)] : So[tri(r - obase, c - obase) * ocap + (u - 1)];
    };
    cd X[D + 1], Y[D + 1], Z[D + 1], W[D + 1];
#pragma unroll
    for (int u = 0; u <= D; ++u) {
        X[u] = co(J, b0, u);
        Y[u] = co(K, b1, u);
        Z[u] = co(K, b0, u);
        W[u] = co(J, b1, u);
    }
#pragma unroll
    for (int t = 1; t <= D2; ++t) {
        cd acc = (t <= D) ? co(J, K, t) : czero();
#pragma unroll
        for (int u = 0; u <= D; ++u) {
            const int v = t - 1 - u;
            if (v >= 0 && v <= D) acc = cadd(acc, cadd(cmul(X[u], Y[v]), cmul(Z[u], W[v])));
        }
        dst[t - 1] = acc;
    }
}

template <int M>
__device__ __forceinline__ void child_entry_dispatch(int D, const cd* Ar, const cd* So, int obase, int ocap, cd* dst,
                                                     int J, int K, int b0, int b1) {
    if (D == M) {
        child_entry<M, M>(Ar, So, obase, ocap, dst, J, K, b0, b1);
    } else if (D == 0) {
        child_entry<0, cmin(M, 1)>(Ar, So, obase, ocap, dst, J, K, b0, b1);
    } else if (D == 1) {
        if constexpr (M > 1) child_entry<1, cmin(M, 3)>(Ar, So, obase, ocap, dst, J, K, b0, b1);
    } else if (D == 3) {
        if constexpr (M > 3) child_entry<3, cmin(M, 7)>(Ar, So, obase, ocap, dst, J, K, b0, b1);
    } else if (D == 7) {
        if constexpr (M > 7) child_entry<7, cmin(M, 15)>(Ar, So, obase, ocap, dst, J, K, b0, b1);
    }
}

// Where the split path's upper phase writes the frontier subtrees' inputs: complex slot `i` of
// hafnian `h` (0 <= h < n_h, relative to the chunk's first item) lives at fdata[i * stride + h],
// hafnian index fastest, so the one-lane-per-subtree kernels read it coalesced.
struct SplitOut {
    double* fdata;
    int64_t stride;
    int64_t n_h;
    const int4* front;  // per frontier node: {k | D << 8 | G << 16, owner | g_level << 8, offset, w > 0}
    int abot_k;         // the bottom block holds A(J, K), J > K >= 2 abot_k, at slots abot_off..
    int abot_off;
};

template <int M, class S, bool SPLIT>
__global__ __launch_bounds__(kWave) void haf_group_kernel(S src, int64_t item0, Layout lay,
                                                          const int4* __restrict__ ops, int G, SplitOut so) {
    extern __shared__ cd smem[];
    const int PER = kWave / G;  // matrices per workgroup
    const int grp = threadIdx.x / G;
    const int gl = threadIdx.x % G;
    constexpr int m = M;
    const int n = 2 * m;
    cd* R = smem + grp * lay.total;
    int* tg = reinterpret_cast<int*>(R + lay.off_tgt);
    const int64_t hl = (int64_t)blockIdx.x * PER + grp;  // position within this launch
    const int64_t item = item0 + hl;
    const bool active = SPLIT ? hl < so.n_h : item < src.n_items;
    const int64_t eff = active ? item : (SPLIT ? item0 + so.n_h - 1 : src.n_items - 1);

    typename std::conditional<std::is_same<S, GammaSource>::value, GammaSource::Decoded, int>::type dc{};
    if constexpr (std::is_same<S, GammaSource>::value) dc = src.decode(eff);
    for (int k = gl; k < n; k += G) {
        int tk;
        if constexpr (std::is_same<S, GammaSource>::value) src.fetch_dec(dc, k, tk);
        else src.fetch(eff, k, tk);
        tg[k] = tk;
    }
    __syncthreads();
    const double* Sg = src.base(eff);
    const int ld = src.ldim();
    for (int i = gl; i < m * (2 * m - 1); i += G) {
        int J, K;
        decode_pair(i, J, K);
        R[lay.off_A + i] = ld_cd(Sg, (int64_t)tg[J] * ld + tg[K]);
    }
    if (gl <= m) R[lay.off_g + gl] = gl == 0 ? cd{1.0, 0.0} : czero();  // g at level 0
    __syncthreads();
    if constexpr (SPLIT) {  // bottom block of the gathered matrix, read by every subtree
        const int s0 = n - 2 * so.abot_k;
        if (active)
            for (int i = gl; i < s0 * (s0 - 1) / 2; i += G) {
                int Jl, Kl;
                decode_pair(i, Jl, Kl);
                st_cd(so.fdata, (so.abot_off + i) * so.stride + hl,
                      R[lay.off_A + tri(Jl + 2 * so.abot_k, Kl + 2 * so.abot_k)]);
            }
    }

    int sp = 0;  // value-stack depth (uniform)
    for (int pc = 0;; ++pc) {
        const int4 op = ops[pc];
        const int kind = op.x & 0xff;
        if (kind == 3) break;
        if (SPLIT && kind == 4) {  // EMIT: copy a frontier node's polynomials and g out
            const int4 fr = so.front[op.y];
            const int k = fr.x & 0xff, D = (fr.x >> 8) & 0xff, Gd = (fr.x >> 16) & 0xff;
            const int owner = fr.y & 0xff, glev = (fr.y >> 8) & 0xff;
            const int s0 = n - 2 * k;
            const int n_p = s0 * (s0 - 1) / 2 * D;
            const cd* So = R + lay.off_S[owner > 0 ? owner : 1];
            const int ocap = owner > 0 ? lay.cap[owner] : 0;
            if (active)
                for (int i = gl; i < n_p + Gd + 1; i += G) {
                    cd v;
                    if (i < n_p) {
                        const int e = i / D, t = i - e * D;  // coefficient t + 1 of entry e
                        int Jl, Kl;
                        decode_pair(e, Jl, Kl);
                        v = So[tri(Jl + 2 * (k - owner), Kl + 2 * (k - owner)) * ocap + t];
                    } else {
                        v = R[lay.off_g + glev * (m + 1) + (i - n_p)];
                    }
                    st_cd(so.fdata, (fr.z + i) * so.stride + hl, v);
                }
            __syncthreads();
            continue;
        }
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
            for (int it = gl; it < n_child + m; it += G) {
                if (it < n_child) {
                    int Jl, Kl;
                    decode_pair(it, Jl, Kl);
                    const int J = Jl + b0 + 2, K = Kl + b0 + 2;
                    child_entry_dispatch<M>(D, R + lay.off_A, So, obase, ocap, Sd + tri(Jl, Kl) * dcap, J, K, b0, b1);
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
            __syncthreads();
        }
    }
    __syncthreads();
    if (!SPLIT && active && gl == 0) {
        const cd h = R[lay.off_val];
        if constexpr (std::is_same<S, GammaSource>::value) src.store_dec(dc, h);
        else src.store(item, h);
    }
}

// ---------------------------------------------------------------------------------------
// Frontier split, lower phase: one lane per (hafnian, frontier subtree).  A subtree rooted at
// level K0 with node degree D (and g degree D: every frontier node the plan makes has G == D)
// is evaluated by the same compile-time recursion as the lane path, from the polynomials and
// g the upper phase wrote out; its value goes to vals[f * stride + h].
// ---------------------------------------------------------------------------------------

template <int OFF>
struct LdsAOff {  // A(J, K) of a subtree whose rows start at OFF, [entry][lane] in LDS
    const cd* base;
    __device__ __forceinline__ cd operator()(int J, int K) const { return base[tri(J - OFF, K - OFF) * kWave]; }
};

template <int M, int K0, int D>
__global__ __launch_bounds__(kWave) void haf_subtree_kernel(const double* __restrict__ fdata, int64_t stride,
                                                            int64_t n_h, const int4* __restrict__ front,
                                                            const int* __restrict__ tlist, int abot_k,
                                                            int abot_off, double* __restrict__ vals) {
    constexpr int S0 = 2 * (M - K0);
    constexpr int NA = S0 * (S0 - 1) / 2;
    __shared__ cd lds[NA * kWave];
    const int lane = threadIdx.x;
    const int64_t h = (int64_t)blockIdx.x * kWave + lane;
    const int f = tlist[blockIdx.y];
    const int4 fr = front[f];
    if (h >= n_h) return;
    const int ak = 2 * abot_k;
#pragma unroll
    for (int J = 2 * K0 + 1; J < 2 * M; ++J)
#pragma unroll
        for (int K = 2 * K0; K < J; ++K)
            lds[tri(J - 2 * K0, K - 2 * K0) * kWave + lane] = ld_cd(fdata, (abot_off + tri(J - ak, K - ak)) * stride + h);
    LdsAOff<2 * K0> A{lds + lane};
    cd P[LaneDims<M>::NE][LaneDims<M>::NC];
    const int64_t base = fr.z;
#pragma unroll
    for (int J = 2 * K0 + 1; J < 2 * M; ++J)
#pragma unroll
        for (int K = 2 * K0; K < J; ++K)
#pragma unroll
            for (int t = 0; t < D; ++t)
                P[tri(J, K)][t] = ld_cd(fdata, (base + tri(J - 2 * K0, K - 2 * K0) * D + t) * stride + h);
    cd g[M + 1];
#pragma unroll
    for (int t = 0; t <= M; ++t) g[t] = t <= D ? ld_cd(fdata, (base + NA * D + t) * stride + h) : czero();
    const double w = fr.w ? 1.0 : -1.0;
    st_cd(vals, (int64_t)f * stride + h, Node<M, K0, D, D>::run(A, P, g, w));
}

// Frontier split, last phase: add the subtree values in the order of the recursion tree
// (comb: f >= 0 pushes subtree f, -1 adds the top two) and apply the item's epilogue.
constexpr int kCombineStack = 18;
template <class S>
__global__ __launch_bounds__(kWave) void haf_combine_kernel(S src, int64_t item0, int64_t n_h, int64_t stride,
                                                            const double* __restrict__ vals,
                                                            const int* __restrict__ comb, int n_comb) {
    __shared__ cd stack[kCombineStack * kWave];
    const int lane = threadIdx.x;
    const int64_t h = (int64_t)blockIdx.x * kWave + lane;
    if (h >= n_h) return;
    int sp = 0;
    for (int i = 0; i < n_comb; ++i) {
        const int c = comb[i];
        if (c >= 0) {
            stack[sp * kWave + lane] = ld_cd(vals, (int64_t)c * stride + h);
            ++sp;
        } else {
            stack[(sp - 2) * kWave + lane] = cadd(stack[(sp - 2) * kWave + lane], stack[(sp - 1) * kWave + lane]);
            --sp;
        }
    }
    const cd v = stack[lane];
    const int64_t item = item0 + h;
    if constexpr (std::is_same<S, GammaSource>::value) src.store_dec(src.decode(item), v);
    else src.store(item, v);
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
            BSMPS_HIP(hipFuncSetAttribute((const void*)&haf_group_kernel<MM, S, false>,                       \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));             \
        for (int64_t b0 = 0; b0 < blocks; b0 += kMaxGrid) {                                                   \
            const int64_t nb = std::min<int64_t>(kMaxGrid, blocks - b0);                                      \
            hipLaunchKernelGGL((haf_group_kernel<MM, S, false>), dim3((unsigned)nb), dim3(kWave), lds, ctx->stream, \
                               src, b0 * per, plan->lay, plan->d_ops, G, SplitOut{});                         \
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
            return BSMPS_ERR_INVALID;
    }
    return BSMPS_OK;
}

// Frontier-split plan of one size: the upper schedule (RIGHT / EMIT / END ops), the frontier
// nodes with their data offsets, the combine program and the frontier nodes grouped by type.
// A node is a frontier node when its subtree fits one lane: depth <= 4 with degree <= 1,
// depth <= 3 with degree <= 3, or depth <= 2.
struct FrontierBuilder {
    int m;
    ScheduleBuilder nzb;
    std::vector<int4> ops;
    std::vector<int4> front;  // {k | D << 8 | G << 16, owner | g_level << 8, offset, w > 0}
    std::vector<int> comb;
    int slots = 0;
    bool is_frontier(int level, int D) const {
        const int depth = m - level;
        return (depth <= 4 && D <= 1) || (depth <= 3 && D <= 3) || depth <= 2;
    }
    void gen(int level, int owner, int D, int G, int gl, int w) {
        if (is_frontier(level, D)) {
            const int s0 = 2 * m - 2 * level;
            comb.push_back((int)front.size());
            ops.push_back(make_int4(4, (int)front.size(), 0, 0));
            front.push_back(make_int4(level | (D << 8) | (G << 16), owner | (gl << 8), slots, w > 0));
            slots += s0 * (s0 - 1) / 2 * D + G + 1;
            return;
        }
        const int D2 = std::min(m, 2 * D + 1), G2 = std::min(m, G + D + 1);
        const bool lnz = nzb.nonzero(level + 1, D, G), rnz = nzb.nonzero(level + 1, D2, G2);
        if (lnz) gen(level + 1, owner, D, G, gl, -w);
        if (rnz) {
            ops.push_back(make_int4(0 | (level << 8) | (owner << 16), D | (D2 << 8) | (G << 16) | (G2 << 24),
                                    gl | ((level + 1) << 8) | ((w > 0) << 16), 0));
            gen(level + 1, level + 1, D2, G2, level + 1, w);
        }
        if (lnz && rnz) comb.push_back(-1);
    }
};

// Subtree shapes with a compiled one-lane kernel: (K0, D) = (M-4, 0), (M-4, 1), (M-3, 3),
// (M-2, 7) and (M-2, M), for M = 6..10.
constexpr int kSplitMinM = 6, kSplitMaxM = 10;
bool subtree_compiled(int M, int K0, int D) {
    if (M < kSplitMinM || M > kSplitMaxM) return false;
    return (K0 == M - 4 && (D == 0 || D == 1)) || (K0 == M - 3 && D == 3) ||
           (K0 == M - 2 && (D == M || (D == 7 && M > 7)));
}

struct FrontierPlan {
    bool ok;
    Layout lay;
    int4* d_ops;
    int4* d_front;
    int* d_comb;
    int n_comb;
    int* d_tlist;
    std::vector<int2> types;  // (K0, D)
    std::vector<int> type_first, type_count;
    int n_front;
    int data_slots;  // frontier inputs + bottom block, per hafnian
    int abot_k, abot_off;
};

std::map<std::pair<int, int>, FrontierPlan> g_fplans;  // (device, m)

int get_fplan(int device, int m, FrontierPlan** out) {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto key = std::make_pair(device, m);
    auto it = g_fplans.find(key);
    if (it == g_fplans.end()) {
        FrontierBuilder fb;
        fb.m = m;
        fb.nzb.m = m;
        fb.gen(0, 0, 0, 0, 0, 1);
        fb.ops.push_back(make_int4(3, 0, 0, 0));
        FrontierPlan p{};
        p.ok = true;
        p.lay = make_layout(m);
        p.n_front = (int)fb.front.size();
        p.abot_k = m;
        std::map<std::pair<int, int>, std::vector<int>> by_type;
        for (int f = 0; f < p.n_front; ++f) {
            const int4 fr = fb.front[f];
            const int k = fr.x & 0xff, D = (fr.x >> 8) & 0xff, G = (fr.x >> 16) & 0xff;
            if (G != D || !subtree_compiled(m, k, D)) p.ok = false;
            p.abot_k = std::min(p.abot_k, k);
            by_type[std::make_pair(k, D)].push_back(f);
        }
        const int sb = 2 * m - 2 * p.abot_k;
        p.abot_off = fb.slots;
        p.data_slots = fb.slots + sb * (sb - 1) / 2;
        std::vector<int> tlist;
        for (auto& kv : by_type) {
            p.types.push_back(make_int2(kv.first.first, kv.first.second));
            p.type_first.push_back((int)tlist.size());
            p.type_count.push_back((int)kv.second.size());
            tlist.insert(tlist.end(), kv.second.begin(), kv.second.end());
        }
        p.n_comb = (int)fb.comb.size();
        int depth = 0, max_depth = 0;
        for (int c : fb.comb) {
            depth += c >= 0 ? 1 : -1;
            max_depth = std::max(max_depth, depth);
        }
        if (max_depth > kCombineStack || depth != 1) p.ok = false;
        BSMPS_HIP(hipMalloc(&p.d_ops, sizeof(int4) * fb.ops.size()));
        BSMPS_HIP(hipMemcpy(p.d_ops, fb.ops.data(), sizeof(int4) * fb.ops.size(), hipMemcpyHostToDevice));
        BSMPS_HIP(hipMalloc(&p.d_front, sizeof(int4) * fb.front.size()));
        BSMPS_HIP(hipMemcpy(p.d_front, fb.front.data(), sizeof(int4) * fb.front.size(), hipMemcpyHostToDevice));
        BSMPS_HIP(hipMalloc(&p.d_comb, sizeof(int) * fb.comb.size()));
        BSMPS_HIP(hipMemcpy(p.d_comb, fb.comb.data(), sizeof(int) * fb.comb.size(), hipMemcpyHostToDevice));
        BSMPS_HIP(hipMalloc(&p.d_tlist, sizeof(int) * tlist.size()));
        BSMPS_HIP(hipMemcpy(p.d_tlist, tlist.data(), sizeof(int) * tlist.size(), hipMemcpyHostToDevice));
        it = g_fplans.emplace(key, p).first;
    }
    *out = &it->second;
    return BSMPS_OK;
}

template <int M>
int launch_subtrees_m(hipStream_t st, int K0, int D, dim3 grid, const double* fdata, int64_t stride, int64_t n_h,
                      const int4* front, const int* tlist, int abot_k, int abot_off, double* vals) {
#define BSMPS_SUBTREE(KK, DD)                                                                                 \
    hipLaunchKernelGGL((haf_subtree_kernel<M, KK, DD>), grid, dim3(kWave), 0, st, fdata, stride, n_h, front,  \
                       tlist, abot_k, abot_off, vals)
    if (K0 == M - 4 && D == 0) BSMPS_SUBTREE(M - 4, 0);
    else if (K0 == M - 4 && D == 1) BSMPS_SUBTREE(M - 4, 1);
    else if (K0 == M - 3 && D == 3) BSMPS_SUBTREE(M - 3, 3);
    else if (K0 == M - 2 && D == M) BSMPS_SUBTREE(M - 2, M);
    else if (K0 == M - 2 && D == 7 && M > 7) BSMPS_SUBTREE(M - 2, (M > 7 ? 7 : M));
    else {
        bsmps_set_error("no subtree kernel for size %d, level %d, degree %d", 2 * M, K0, D);
        return BSMPS_ERR_UNSUPPORTED;
    }
#undef BSMPS_SUBTREE
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

// Frontier-split launch: per chunk of hafnians, the upper phase (cooperative lanes), one launch
// per subtree shape, then the combine; the chunk is sized so its scratch fits the budget.
template <class S>
int launch_split(bsmps_ctx* ctx, const S& src, int m, FrontierPlan* fp) {
    const int G = ctx->group_lanes_by_m[m];
    const int per = kWave / G;
    const size_t lds = sizeof(cd) * (size_t)fp->lay.total * per;
    if (lds > (size_t)ctx->max_lds) {
        bsmps_set_error("hafnian of size %d needs %zu bytes of LDS per workgroup with %d lanes per matrix "
                        "(device limit %d)", 2 * m, lds, G, ctx->max_lds);
        return BSMPS_ERR_UNSUPPORTED;
    }
    const size_t per_h = sizeof(cd) * (size_t)(fp->data_slots + fp->n_front);
    int64_t H = (int64_t)(ctx->split_chunk_bytes / per_h);
    H = std::max<int64_t>(kWave, H / kWave * kWave);
    H = std::min<int64_t>(H, (src.n_items + kWave - 1) / kWave * kWave);
    void* scratch = nullptr;
    BSMPS_TRY(bsmps_ctx_scratch(ctx, per_h * (size_t)H, &scratch));
    double* fdata = reinterpret_cast<double*>(scratch);
    double* vals = fdata + 2 * (size_t)fp->data_slots * H;
    for (int64_t c0 = 0; c0 < src.n_items; c0 += H) {
        const int64_t nh = std::min<int64_t>(H, src.n_items - c0);
        SplitOut so{fdata, H, nh, fp->d_front, fp->abot_k, fp->abot_off};
        const unsigned blocks = (unsigned)((nh + per - 1) / per);
        switch (m) {
#define BSMPS_SPLIT_CASE(MM)                                                                                  \
    case MM:                                                                                                  \
        if (lds > 64 * 1024)                                                                                  \
            BSMPS_HIP(hipFuncSetAttribute((const void*)&haf_group_kernel<MM, S, true>,                        \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));             \
        hipLaunchKernelGGL((haf_group_kernel<MM, S, true>), dim3(blocks), dim3(kWave), lds, ctx->stream, src, c0, \
                           fp->lay, fp->d_ops, G, so);                                                        \
        BSMPS_HIP(hipGetLastError());                                                                         \
        for (size_t t = 0; t < fp->types.size(); ++t)                                                         \
            BSMPS_TRY(launch_subtrees_m<MM>(ctx->stream, fp->types[t].x, fp->types[t].y,                      \
                                            dim3((unsigned)((nh + kWave - 1) / kWave), fp->type_count[t]),    \
                                            fdata, H, nh, fp->d_front, fp->d_tlist + fp->type_first[t],       \
                                            fp->abot_k, fp->abot_off, vals));                                 \
        break;
            BSMPS_SPLIT_CASE(6)
            BSMPS_SPLIT_CASE(7)
            BSMPS_SPLIT_CASE(8)
            BSMPS_SPLIT_CASE(9)
            BSMPS_SPLIT_CASE(10)
#undef BSMPS_SPLIT_CASE
            default:
                bsmps_set_error("frontier-split hafnian path supports sizes 12..20, got %d", 2 * m);
                return BSMPS_ERR_INVALID;
        }
        hipLaunchKernelGGL((haf_combine_kernel<S>), dim3((unsigned)((nh + kWave - 1) / kWave)), dim3(kWave), 0,
                           ctx->stream, src, c0, nh, H, vals, fp->d_comb, fp->n_comb);
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

extern "C" int bsmps_gamma_site(bsmps_ctx* ctx, const bsmps_gamma_site_args* a, double* gamma) {
    if (!ctx || !a || !gamma) {
        bsmps_set_error("bsmps_gamma_site: null argument");
        return BSMPS_ERR_INVALID;
    }
    if (a->kind < 0 || a->kind > 2 || a->d <= 0 || a->chi <= 0 || a->ld <= 0 || a->n_left < 0 ||
        a->n_right < 0 || a->n_left > a->chi || a->n_right > a->chi || a->n_left_class <= 0 ||
        a->n_right_class <= 0 || !a->sigma || !a->sqrt_fact || !a->left_den || !a->lambda || !a->left_perm ||
        !a->right_den || !a->right_perm || !a->left_class_count_host || !a->right_class_count_host) {
        bsmps_set_error("bsmps_gamma_site: invalid arguments");
        return BSMPS_ERR_INVALID;
    }
    const int nlc = a->n_left_class, nrc = a->n_right_class;
    if (nlc - 1 > a->w_left || nrc - 1 > a->w_right) {
        bsmps_set_error("bsmps_gamma_site: photon classes exceed target widths");
        return BSMPS_ERR_INVALID;
    }
    // class offsets and counts
    std::vector<int32_t> loff(nlc + 1, 0), roff(nrc + 1, 0), rcnt(nrc, 0);
    int64_t lsum = 0, rsum = 0;
    for (int p = 0; p < nlc; ++p) {
        if (a->left_class_count_host[p] < 0) { bsmps_set_error("bsmps_gamma_site: negative class count"); return BSMPS_ERR_INVALID; }
        loff[p + 1] = loff[p] + a->left_class_count_host[p];
        lsum += a->left_class_count_host[p];
    }
    for (int q = 0; q < nrc; ++q) {
        if (a->right_class_count_host[q] < 0) { bsmps_set_error("bsmps_gamma_site: negative class count"); return BSMPS_ERR_INVALID; }
        roff[q + 1] = roff[q] + a->right_class_count_host[q];
        rcnt[q] = a->right_class_count_host[q];
        rsum += a->right_class_count_host[q];
    }
    if (lsum != a->n_left || rsum != a->n_right) {
        bsmps_set_error("bsmps_gamma_site: class counts do not add up to the label counts");
        return BSMPS_ERR_INVALID;
    }
    const int max_n = (a->d - 1) + (nlc - 1) + (nrc - 1);
    if (max_n > BSMPS_MAX_HAFNIAN_SIZE + 1) {
        bsmps_set_error("bsmps_gamma_site: hafnians up to size %d exceed the supported %d", max_n,
                        BSMPS_MAX_HAFNIAN_SIZE);
        return BSMPS_ERR_UNSUPPORTED;
    }
    // rectangles (j, p, q) grouped by even total n = j + p + q, items numbered per n
    std::vector<int4> rects;
    std::vector<int> n_first(max_n + 2, 0);
    std::vector<int64_t> n_items(max_n + 1, 0);
    for (int n = 0; n <= max_n; ++n) {
        n_first[n] = (int)rects.size();
        if (n % 2) continue;
        int64_t start = 0;
        for (int j = 0; j < a->d && j <= n; ++j)
            for (int p = 0; p < nlc && j + p <= n; ++p) {
                const int q = n - j - p;
                if (q >= nrc) continue;
                const int64_t cnt = (int64_t)a->left_class_count_host[p] * a->right_class_count_host[q];
                if (cnt == 0) continue;
                if (start > INT32_MAX) { bsmps_set_error("bsmps_gamma_site: too many entries of one size"); return BSMPS_ERR_UNSUPPORTED; }
                rects.push_back(make_int4(j, p, q, (int)start));
                start += cnt;
            }
        n_items[n] = start;
    }
    n_first[max_n + 1] = (int)rects.size();
    // upload tables into the context workspace: rects | loff | roff | rcnt
    const size_t bytes_rects = sizeof(int4) * std::max<size_t>(rects.size(), 1);
    const size_t bytes_tab = sizeof(int32_t) * (loff.size() + roff.size() + rcnt.size());
    void* ws = nullptr;
    BSMPS_TRY(bsmps_ctx_workspace(ctx, bytes_rects + bytes_tab + 256, &ws));
    std::vector<char> host(bytes_rects + bytes_tab);
    memcpy(host.data(), rects.data(), sizeof(int4) * rects.size());
    char* p = host.data() + bytes_rects;
    memcpy(p, loff.data(), sizeof(int32_t) * loff.size());
    p += sizeof(int32_t) * loff.size();
    memcpy(p, roff.data(), sizeof(int32_t) * roff.size());
    p += sizeof(int32_t) * roff.size();
    memcpy(p, rcnt.data(), sizeof(int32_t) * rcnt.size());
    BSMPS_HIP(hipMemcpyAsync(ws, host.data(), host.size(), hipMemcpyHostToDevice, ctx->stream));
    BSMPS_HIP(hipStreamSynchronize(ctx->stream));  // `host` goes out of scope
    const int4* d_rects = reinterpret_cast<const int4*>(ws);
    const int32_t* d_loff = reinterpret_cast<const int32_t*>(reinterpret_cast<char*>(ws) + bytes_rects);
    const int32_t* d_roff = d_loff + loff.size();
    const int32_t* d_rcnt = d_roff + roff.size();

    GammaSource base{};
    base.sigma = a->sigma;
    base.ld = a->ld;
    base.d = a->d;
    base.chi = a->chi;
    base.vacuum_real_div = a->kind != BSMPS_SITE_FIRST;
    base.z = a->z;
    base.inv_z = a->inv_z;
    base.sqrt_fact = a->sqrt_fact;
    base.lt = a->left_targets;
    base.wl = a->w_left;
    base.lden = a->left_den;
    base.lam = a->lambda;
    base.lperm = a->left_perm;
    base.loff = d_loff;
    base.rt = a->right_targets;
    base.wr = a->w_right;
    base.rden = a->right_den;
    base.rperm = a->right_perm;
    base.roff = d_roff;
    base.rcnt = d_rcnt;
    base.gamma = gamma;
    // largest sizes first: they are the longest-running launches
    for (int n = max_n; n >= 0; --n) {
        if (n % 2 || n_items[n] == 0) continue;
        GammaSource s = base;
        s.rects = d_rects + n_first[n];
        s.n_rects = n_first[n + 1] - n_first[n];
        s.n_items = n_items[n];
        BSMPS_TRY(launch_hafnians(ctx, s, n));
    }
    return BSMPS_OK;
}
