// Sampling path (L4): random displacements, site contraction, conditional probabilities
// and the draw (sampling_cpu.py:30-159; GPU path distributed_sampling.py:62-175).
//
//   bsmps_mu_to_alpha      alpha = (sqrtW @ z) -> complex amplitudes            (:37-41)
//   bsmps_batch_displaces  expm(alpha a^dag - alpha^* a), FP64 spectral form    (:133-148)
//   bsmps_site_contract    T = (pre * Lambda_prev) @ Gamma, FP64 MFMA GEMM      (:75-80)
//   bsmps_sample_site      D(alpha) T, |.|^2 Lambda^2, inverse CDF, select      (:81-113)
//   bsmps_philox_*         counter-based RNG replacing np.random                (:37, 52, 92)
#include <math.h>

#include <algorithm>
#include <type_traits>

#include "common.h"

namespace {

// ---------------------------------------------------------------------------------------
// alpha from normals: mu[i] = sum_k sqrtW[i][k] z[k][s]; alpha = (mu[i] + 1j mu[M+i]) / 2
// ---------------------------------------------------------------------------------------
__global__ void mu_to_alpha_kernel(const double* __restrict__ sqrtW, int M, const double* __restrict__ z,
                                   int64_t n, double* __restrict__ alpha) {
    const int64_t total = n * M;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = idx / M;
        const int i = (int)(idx - s * M);
        double re = 0.0, im = 0.0;
        const double* wr = sqrtW + (int64_t)i * 2 * M;
        const double* wi = sqrtW + (int64_t)(M + i) * 2 * M;
        for (int k = 0; k < 2 * M; ++k) {
            const double zk = z[(int64_t)k * n + s];
            re = fma(wr[k], zk, re);
            im = fma(wi[k], zk, im);
        }
        reinterpret_cast<double2*>(alpha)[s * M + i] = make_double2(re * 0.5, im * 0.5);
    }
}

// ---------------------------------------------------------------------------------------
// Displacement operator on the dd-dim truncated space.  With alpha = r e^{i phi},
// alpha a^dag - alpha^* a = R (r (a^dag - a)) R^-1, R = diag(e^{i n phi}), and
// a^dag - a = P (-i Q) P^-1 with P = diag(i^n), Q = a + a^dag = W diag(x) W^T (real symmetric
// tridiagonal).  Hence D[m][k] = e^{i (m - k) theta} sum_q W[m][q] W[k][q] e^{-i r x_q},
// theta = phi + pi/2 -- the exact exponential of the truncated generator.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void displace_entry_setup(cd alpha, int dd, const double* x, double* cr, double* sr,
                                                     double& theta) {
    const double r = hypot(alpha.re, alpha.im);
    theta = atan2(alpha.im, alpha.re) + 1.5707963267948966;
    for (int q = 0; q < dd; ++q) {
        double s, c;
        sincos(r * x[q], &s, &c);
        cr[q] = c;
        sr[q] = -s;  // e^{-i r x_q}
    }
}

__global__ void batch_displaces_kernel(int dd, const double* __restrict__ alpha, int64_t n, int64_t astride,
                                       const double* __restrict__ W, const double* __restrict__ x,
                                       double* __restrict__ out) {
    // one sample per block; threads over (m, k)
    extern __shared__ double sh[];
    double* cr = sh;
    double* sr = sh + dd;
    __shared__ double s_theta;
    for (int64_t s = blockIdx.x; s < n; s += gridDim.x) {
        if (threadIdx.x == 0) {
            const cd a = ld_cd(alpha, s * astride);
            double th;
            displace_entry_setup(a, dd, x, cr, sr, th);
            s_theta = th;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < dd * dd; e += blockDim.x) {
            const int m = e / dd, k = e - m * dd;
            double re = 0.0, im = 0.0;
            for (int q = 0; q < dd; ++q) {
                const double w = W[m * dd + q] * W[k * dd + q];
                re = fma(w, cr[q], re);
                im = fma(w, sr[q], im);
            }
            double sp, cp;
            sincos((double)(m - k) * s_theta, &sp, &cp);
            st_cd(out, s * dd * dd + e, cd{re * cp - im * sp, re * sp + im * cp});
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// Site contraction as a real GEMM on MFMA:
//   C[s][2c + {0,1}] = sum_{b} A[s][2b + {0,1}] * Bx[2b + {0,1}][2c + {0,1}]
// with A[s][2b] = Re(pre) lambda, A[s][2b+1] = Im(pre) lambda and Bx the 2x2 real block
// [[Re g, Im g], [-Im g, Re g]] of each complex Gamma entry g -- so C is the complex product
// stored interleaved, exactly T's layout.  M = n, N = 2 chi d, K = 2 chi_prev; 8 n chi^2 d
// real flops.  Workgroup tile 128 x 128 x TKB (256 threads, four waves of 64 x 64 = 4 x 4 MFMA
// 16x16 blocks); the next K tile is fetched into registers while the current one computes.
// ---------------------------------------------------------------------------------------
constexpr int TM = 128, TN = 128;

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

template <typename E>
struct MfmaTraits;
template <>
struct MfmaTraits<double> {
    typedef double4_t acc_t;
    typedef double2 pair_t;
    __device__ static acc_t mfma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // register r of lane l holds D[4 r + l/16][l % 16]
    __device__ static int row(int lane, int r) { return 4 * r + lane / 16; }
};
template <>
struct MfmaTraits<float> {
    typedef float4_t acc_t;
    typedef float2 pair_t;
    __device__ static acc_t mfma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // register r of lane l holds D[4 (l/16) + r][l % 16]
    __device__ static int row(int lane, int r) { return 4 * (lane / 16) + r; }
};

// Where the contraction kernels put output column c (complex) of sample row s: T[s][c] in a row of
// ncols when MAPPED is false, else T[s * ld + map[c]] (rows of ld complex values; an entry outside
// [0, ld) is not written).  The photon-number parity split contracts each parity block of a site
// separately and scatters its columns this way.
struct ColMap {
    int64_t ld;
    const int* map;
};

template <typename E, bool MAPPED = false>
__global__ __launch_bounds__(256, 2) void site_gemm_kernel(int64_t n, int chi_prev, int ncols /* complex cols = chi*d */,
                                                        const E* __restrict__ pre, const double* __restrict__ lam,
                                                        const E* __restrict__ gamma, E* __restrict__ T, ColMap om) {
    typedef MfmaTraits<E> Tr;
    typedef typename Tr::acc_t acc_t;
    typedef typename Tr::pair_t E2;
    // real k per LDS stage: FP64 takes 32, so a stage's MFMA work (~128 per wave) covers the
    // next stage's prefetch (0.685 -> 0.726 of the FP64 MFMA peak at n = 10^5, chi = 512, d = 6);
    // FP32 MFMAs are twice as fast and its registers prefer 16 (32 measured 0.72 -> 0.67)
    constexpr int TKB = sizeof(E) == 8 ? 32 : 16;
    __shared__ E As[TKB][TM + 2];  // As[k][row]
    __shared__ E Bs[TKB][TN + 2];  // Bs[k][col]
    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int64_t row0 = (int64_t)blockIdx.y * TM;
    const int64_t col0 = (int64_t)blockIdx.x * TN;  // real column
    const int64_t Nreal = 2 * (int64_t)ncols;
    const int Kreal = 2 * chi_prev;
    const int wm = (wave / 2) * 64, wn = (wave % 2) * 64;

    acc_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    // per thread: QA complex of the A tile (128 rows x TKB/2 b) and QB of the B tile (TKB/2 b x 64 c)
    constexpr int QA = TM * (TKB / 2) / 256, QB = (TKB / 2) * (TN / 2) / 256;
    E2 ra[QA], rb[QB];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int e = tid + 256 * q;
            const int r = e / (TKB / 2), bb = e % (TKB / 2);
            const int64_t srow = row0 + r;
            const int b = k0 / 2 + bb;
            E2 v;
            v.x = 0;
            v.y = 0;
            if (srow < n && b < chi_prev) {
                v = reinterpret_cast<const E2*>(pre)[srow * chi_prev + b];
                const double l = lam[b];
                v.x = (E)(v.x * l);
                v.y = (E)(v.y * l);
            }
            ra[q] = v;
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int e = tid + 256 * q;
            const int bb = e / (TN / 2), cc = e % (TN / 2);
            const int b = k0 / 2 + bb;
            const int64_t c = col0 / 2 + cc;
            E2 g;
            g.x = 0;
            g.y = 0;
            if (b < chi_prev && c < ncols) g = reinterpret_cast<const E2*>(gamma)[(int64_t)b * ncols + c];
            rb[q] = g;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int e = tid + 256 * q;
            const int r = e / (TKB / 2), bb = e % (TKB / 2);
            As[2 * bb][r] = ra[q].x;
            As[2 * bb + 1][r] = ra[q].y;
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int e = tid + 256 * q;
            const int bb = e / (TN / 2), cc = e % (TN / 2);
            Bs[2 * bb][2 * cc] = rb[q].x;
            Bs[2 * bb][2 * cc + 1] = rb[q].y;
            Bs[2 * bb + 1][2 * cc] = -rb[q].y;
            Bs[2 * bb + 1][2 * cc + 1] = rb[q].x;
        }
    };

    fetch(0);
    for (int k0 = 0; k0 < Kreal; k0 += TKB) {
        stash();
        __syncthreads();
        if (k0 + TKB < Kreal) fetch(k0 + TKB);
#pragma unroll
        for (int kk = 0; kk < TKB; kk += 4) {
            E a[4], b[4];
            // operand layout: lane holds A[lane%16][lane/16] and B[lane/16][lane%16]
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk + lane / 16][wm + i * 16 + lane % 16];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk + lane / 16][wn + j * 16 + lane % 16];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = Tr::mfma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t srow = row0 + wm + i * 16 + Tr::row(lane, r);
                const int64_t c = col0 + wn + j * 16 + lane % 16;
                if (srow < n && c < Nreal) {
                    if constexpr (MAPPED) {
                        const int64_t oc = om.map[c >> 1];
                        if (oc >= 0 && oc < om.ld) T[(srow * om.ld + oc) * 2 + (c & 1)] = acc[i][j][r];
                    } else {
                        T[srow * Nreal + c] = acc[i][j][r];
                    }
                }
            }
}

// ---------------------------------------------------------------------------------------
// Site contraction, Gauss form: T[s][c] = sum_b (pre[s][b] lambda[b]) Gamma[b][c], complex,
// stored interleaved (T's layout).  The complex product runs as three real products
// (Gauss's method: P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi); Re = P1 - P2,
// Im = P3 - P1 - P2), 6 n chi_prev chi d real MFMA flops instead of the 8 of the complex
// product; the extra rounding is a few ulps of |A||B| (the FP64 form stays within 1e-13 of
// the exact product).  Four waves of 64 rows x 16 JB complex columns (3 x 4 x JB MFMA 16x16
// blocks each), WR of them along the rows: workgroup tile 64 WR rows x 64 JB (4 / WR) complex
// columns x KC complex k; the next K tile is fetched into registers while the current one
// computes.  FP32: JB = 2, WR = 1 (64 x 128; 128 x 64 was 10% slower), KC = 8.  FP64: JB = 1
// (the 24 FP64 accumulators of JB = 2 took 256 VGPRs and AGPRs besides, one wave per SIMD),
// WR = 1 (a square tile: 128 x 32 re-read the carried vectors once per 32 columns) and KC = 16,
// so a stage's MFMA work still covers the next stage's fetch.
// Selected per precision by bsmps_set_contract_form (the autotuner times both forms).
// ---------------------------------------------------------------------------------------
template <typename E>
struct GaussShape;
template <>
struct GaussShape<double> {
    static constexpr int JB = 1, WR = 1, KC = 16;
    static constexpr int ROWS = 64 * WR, COLS = 16 * JB * (4 / WR);  // workgroup tile
};
template <>
struct GaussShape<float> {
    static constexpr int JB = 2, WR = 1, KC = 8;
    static constexpr int ROWS = 64 * WR, COLS = 16 * JB * (4 / WR);
};

template <typename E, bool MAPPED = false>
__global__ __launch_bounds__(256, 2) void site_gemm3m_kernel(int64_t n, int chi_prev, int ncols /* complex cols = chi*d */,
                                                           const E* __restrict__ pre, const double* __restrict__ lam,
                                                           const E* __restrict__ gamma, E* __restrict__ T, ColMap om) {
    typedef MfmaTraits<E> Tr;
    typedef typename Tr::acc_t acc_t;
    typedef typename Tr::pair_t E2;
    constexpr int JB = GaussShape<E>::JB, WC = 4 / GaussShape<E>::WR;  // waves along the columns
    constexpr int KC = GaussShape<E>::KC;    // complex k per tile
    constexpr int MR = GaussShape<E>::ROWS;  // rows per tile
    constexpr int NC = GaussShape<E>::COLS;  // complex columns per tile
    constexpr int QA = KC * MR / 256, QB = KC * NC / 256;  // complex per thread of the A and B tiles
    static_assert(QA * 256 == KC * MR && QB * 256 == KC * NC, "tiles split evenly over 256 threads");
    __shared__ E Ar[KC][MR + 2], Ai[KC][MR + 2];  // [k][row]
    __shared__ E Br[KC][NC + 2], Bi[KC][NC + 2];  // [k][complex col]
    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int64_t row0 = (int64_t)blockIdx.y * MR;
    const int64_t ccol0 = (int64_t)blockIdx.x * NC;  // complex column
    const int wm = (wave / WC) * 64, wn = (wave % WC) * 16 * JB;

    // three real products per 16x16 block (Gauss): P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi)
    acc_t p1[4][JB], p2[4][JB], p3[4][JB];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) p1[i][j] = p2[i][j] = p3[i][j] = acc_t{0, 0, 0, 0};

    E2 ra[QA], rb[QB];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int e = tid + 256 * q;
            const int r = e / KC, bb = e % KC;
            const int64_t srow = row0 + r;
            const int b = k0 + bb;
            E2 v;
            v.x = 0;
            v.y = 0;
            if (srow < n && b < chi_prev) {
                v = reinterpret_cast<const E2*>(pre)[srow * chi_prev + b];
                const double l = lam[b];
                v.x = (E)(v.x * l);
                v.y = (E)(v.y * l);
            }
            ra[q] = v;
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int e = tid + 256 * q;
            const int bb = e / NC, cc = e % NC;
            const int b = k0 + bb;
            const int64_t c = ccol0 + cc;
            E2 g;
            g.x = 0;
            g.y = 0;
            if (b < chi_prev && c < ncols) g = reinterpret_cast<const E2*>(gamma)[(int64_t)b * ncols + c];
            rb[q] = g;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int e = tid + 256 * q;
            const int r = e / KC, bb = e % KC;
            Ar[bb][r] = ra[q].x;
            Ai[bb][r] = ra[q].y;
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int e = tid + 256 * q;
            const int bb = e / NC, cc = e % NC;
            Br[bb][cc] = rb[q].x;
            Bi[bb][cc] = rb[q].y;
        }
    };

    fetch(0);
    for (int k0 = 0; k0 < chi_prev; k0 += KC) {
        stash();
        __syncthreads();
        if (k0 + KC < chi_prev) fetch(k0 + KC);
#pragma unroll
        for (int kk = 0; kk < KC; kk += 4) {
            // operand layout: lane holds A[lane%16][lane/16] and B[lane/16][lane%16]
            E ar[4], ai[4], as[4], br[JB], bi[JB], bs[JB];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                ar[i] = Ar[kk + lane / 16][wm + i * 16 + lane % 16];
                ai[i] = Ai[kk + lane / 16][wm + i * 16 + lane % 16];
                as[i] = ar[i] + ai[i];
            }
#pragma unroll
            for (int j = 0; j < JB; ++j) {
                br[j] = Br[kk + lane / 16][wn + j * 16 + lane % 16];
                bi[j] = Bi[kk + lane / 16][wn + j * 16 + lane % 16];
                bs[j] = br[j] + bi[j];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < JB; ++j) {
                    p1[i][j] = Tr::mfma(ar[i], br[j], p1[i][j]);
                    p2[i][j] = Tr::mfma(ai[i], bi[j], p2[i][j]);
                    p3[i][j] = Tr::mfma(as[i], bs[j], p3[i][j]);
                }
        }
        __syncthreads();
    }
    // re = P1 - P2, im = (P3 - P1) - P2
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t srow = row0 + wm + i * 16 + Tr::row(lane, r);
                const int64_t c = ccol0 + wn + j * 16 + lane % 16;
                if (srow < n && c < ncols) {
                    E2 v;
                    v.x = p1[i][j][r] - p2[i][j][r];
                    v.y = (p3[i][j][r] - p1[i][j][r]) - p2[i][j][r];
                    if constexpr (MAPPED) {
                        const int64_t oc = om.map[c];
                        if (oc >= 0 && oc < om.ld) reinterpret_cast<E2*>(T)[srow * om.ld + oc] = v;
                    } else {
                        reinterpret_cast<E2*>(T)[srow * ncols + c] = v;
                    }
                }
            }
}

// ---------------------------------------------------------------------------------------
// BF16x3 split-precision form of the contraction (v_mfma_f32_16x16x32_bf16, FP32
// accumulation).  Every FP32 operand x is split as x = hi + lo with hi = bf16(x) and
// lo = bf16(x - hi); a product is taken as hi*hi + hi*lo + lo*hi (the lo*lo term, ~2^-16
// relative, is dropped).  Gamma is split once per site (bsmps_split_bf16); the carried
// vectors are split as they are staged.  Tile 128 x 128 x 32 (real), LDS rows k-contiguous
// so each lane's 8 consecutive k values are one 16-byte read.
// ---------------------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int BK = 32;             // real K per tile (16 complex b)
constexpr int BKP = BK + 8;        // padded row length (80 bytes)

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

__global__ void split_bf16_kernel(const float* __restrict__ g, int64_t count, __bf16* __restrict__ hi,
                                  __bf16* __restrict__ lo) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        split_bf16(g[i], hi[i], lo[i]);
}

__global__ __launch_bounds__(256) void site_gemm_bf16x3_kernel(int64_t n, int chi_prev, int ncols,
                                                               const float* __restrict__ pre,
                                                               const double* __restrict__ lam,
                                                               const __bf16* __restrict__ ghi,
                                                               const __bf16* __restrict__ glo, float* __restrict__ T) {
    __shared__ __attribute__((aligned(16))) __bf16 Ah[TM][BKP];
    __shared__ __attribute__((aligned(16))) __bf16 Al[TM][BKP];
    __shared__ __attribute__((aligned(16))) __bf16 Bh[TN][BKP];
    __shared__ __attribute__((aligned(16))) __bf16 Bl[TN][BKP];
    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int64_t row0 = (int64_t)blockIdx.y * TM;
    const int64_t col0 = (int64_t)blockIdx.x * TN;
    const int64_t Nreal = 2 * (int64_t)ncols;
    const int Kreal = 2 * chi_prev;
    const int wm = (wave / 2) * 64, wn = (wave % 2) * 64;
    float4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};

    // per thread: 8 complex of A (128 rows x 16 b) and 4 of B (16 b x 64 c)
    float2 ra[8];
    uint32_t rbh[4], rbl[4];  // (re, im) bf16 pairs
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + 256 * q;
            const int r = e / (BK / 2), bb = e % (BK / 2);
            const int64_t srow = row0 + r;
            const int b = k0 / 2 + bb;
            float2 v = make_float2(0.f, 0.f);
            if (srow < n && b < chi_prev) {
                v = reinterpret_cast<const float2*>(pre)[srow * chi_prev + b];
                const double l = lam[b];
                v.x = (float)(v.x * l);
                v.y = (float)(v.y * l);
            }
            ra[q] = v;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + 256 * q;
            const int bb = e / (TN / 2), cc = e % (TN / 2);
            const int b = k0 / 2 + bb;
            const int64_t c = col0 / 2 + cc;
            uint32_t h = 0, l = 0;
            if (b < chi_prev && c < ncols) {
                h = reinterpret_cast<const uint32_t*>(ghi)[(int64_t)b * ncols + c];
                l = reinterpret_cast<const uint32_t*>(glo)[(int64_t)b * ncols + c];
            }
            rbh[q] = h;
            rbl[q] = l;
        }
    };
    auto neg = [](uint32_t pair_hi16) -> __bf16 {  // -(bf16 stored in the high half)
        const unsigned short u = (unsigned short)(pair_hi16 >> 16) ^ 0x8000u;
        return __builtin_bit_cast(__bf16, u);
    };
    auto lo16 = [](uint32_t p) -> __bf16 { return __builtin_bit_cast(__bf16, (unsigned short)(p & 0xffffu)); };
    auto hi16 = [](uint32_t p) -> __bf16 { return __builtin_bit_cast(__bf16, (unsigned short)(p >> 16)); };
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + 256 * q;
            const int r = e / (BK / 2), bb = e % (BK / 2);
            __bf16 h, l;
            split_bf16(ra[q].x, h, l);
            Ah[r][2 * bb] = h;
            Al[r][2 * bb] = l;
            split_bf16(ra[q].y, h, l);
            Ah[r][2 * bb + 1] = h;
            Al[r][2 * bb + 1] = l;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + 256 * q;
            const int bb = e / (TN / 2), cc = e % (TN / 2);
            // B_exp rows (2bb, 2bb+1), cols (2cc, 2cc+1): [[gr, gi], [-gi, gr]], stored [col][k]
            Bh[2 * cc][2 * bb] = lo16(rbh[q]);
            Bh[2 * cc + 1][2 * bb] = hi16(rbh[q]);
            Bh[2 * cc][2 * bb + 1] = neg(rbh[q]);
            Bh[2 * cc + 1][2 * bb + 1] = lo16(rbh[q]);
            Bl[2 * cc][2 * bb] = lo16(rbl[q]);
            Bl[2 * cc + 1][2 * bb] = hi16(rbl[q]);
            Bl[2 * cc][2 * bb + 1] = neg(rbl[q]);
            Bl[2 * cc + 1][2 * bb + 1] = lo16(rbl[q]);
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < Kreal; k0 += BK) {
        stash();
        __syncthreads();
        if (k0 + BK < Kreal) fetch(k0 + BK);
        bf16x8 ah[4], al[4], bh[4], bl[4];
        const int ko = 8 * (lane / 16);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            ah[i] = *reinterpret_cast<const bf16x8*>(&Ah[wm + i * 16 + lane % 16][ko]);
            al[i] = *reinterpret_cast<const bf16x8*>(&Al[wm + i * 16 + lane % 16][ko]);
            bh[i] = *reinterpret_cast<const bf16x8*>(&Bh[wn + i * 16 + lane % 16][ko]);
            bl[i] = *reinterpret_cast<const bf16x8*>(&Bl[wn + i * 16 + lane % 16][ko]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
            }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t srow = row0 + wm + i * 16 + 4 * (lane / 16) + r;
                const int64_t c = col0 + wn + j * 16 + lane % 16;
                if (srow < n && c < Nreal) T[srow * Nreal + c] = acc[i][j][r];
            }
}

// ---------------------------------------------------------------------------------------
// BF16x6 form of the contraction: FP32-accurate products on BF16 MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation).  Every FP32 value x is split exactly into
// three BF16 parts x = h + m + l (h = bf16(x), m = bf16(x - h), l = bf16(x - h - m); both
// differences are exact in FP32, so h + m + l carries all 24 significand bits), and a product
// keeps the six terms of relative order >= 2^-16: hh + hm + mh + hl + mm + lh (the dropped
// ones are ~2^-24, the size of an FP32 rounding).  The complex product runs on planes:
// Re = Ar Br + (-Ai) Bi, Im = Ar Bi + Ai Br.  Both operands arrive split, as six bf16 planes
// (parts h, m, l of the real part, then of the imaginary part) with k contiguous and padded
// to a multiple of 32: lambda_prev * Gamma once per site (bsmps_split_bf16x6, [part][col][k]),
// the carried vectors once per step (bsmps_split_rows_bf16x6, [part][row][k]), so the GEMM
// itself does no conversion.  Workgroup tile 128 rows x 64 complex columns x 32 complex k,
// staged in LDS with an XOR swizzle of the 16-byte k groups (conflict-free fragment reads);
// four waves of 64 rows x 32 columns = 4 x 2 blocks of 16 x 16, each with a real and an
// imaginary accumulator.
// ---------------------------------------------------------------------------------------
constexpr int X6_TM = 128, X6_TN = 64, X6_KC = 32;
// Split-form contractions with more than kSplitKMin values of k run in k segments of kSplitKSeg
// (see site_gemm_split_kernel)
constexpr int64_t kSplitKMin = 4096, kSplitKSeg = 2048;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
    h = (__bf16)x;
    const float r1 = x - (float)h;
    m = (__bf16)r1;
    l = (__bf16)(r1 - (float)m);
}

// 16-byte group g (of 4) of LDS row r is stored in slot g ^ f(r)
__device__ __forceinline__ int x6_slot(int r, int g) { return g ^ (((r >> 2) ^ (r >> 1)) & 3); }

// Plane layout of the split forms, for a row (carried vector) or column (Gamma) r: element k
// of part p at ((r kpad / KC + k / KC) NPL + p) KC + k % KC.  One KC-step of all NPL parts of a
// row is NPL * 64 contiguous bytes, so a GEMM step reads whole 128-byte lines however long the
// rows are (with the parts in separate planes each step read half a line per part, and at
// kpad = 10^4 the other half was evicted before the next step).
__device__ __forceinline__ int64_t x6_at(int64_t r, int k, int kpad, int npl, int p) {
    return ((r * (kpad / X6_KC) + k / X6_KC) * npl + p) * X6_KC + k % X6_KC;
}

// Operand formats of the split-precision contraction.  An FP32 value is split into NP parts
// of a 16-bit type; prod() adds the retained part products of one real product to acc.
struct Bf16x6Fmt {  // three bf16 parts (exact), products of order >= 2^-16: six MFMAs
    typedef __bf16 E;
    typedef bf16x8 V;
    static constexpr int NP = 3;
    static constexpr bool SCALED = false;
    static constexpr bool DMA = true;  // staged by LDS DMA (site_gemm_split_dma_kernel)
    __device__ static float4_t mfma(V a, V b, float4_t c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
    __device__ static float4_t prod(const V (&a)[NP], const V* b, float4_t acc) {
        acc = mfma(a[2], b[0], acc);
        acc = mfma(a[1], b[1], acc);
        acc = mfma(a[0], b[2], acc);
        acc = mfma(a[1], b[0], acc);
        acc = mfma(a[0], b[1], acc);
        return mfma(a[0], b[0], acc);
    }
};
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
struct Fp16x3Fmt {  // two fp16 parts (22 bits) of values scaled to max 1: three MFMAs
    typedef _Float16 E;
    typedef half8 V;
    static constexpr int NP = 2;
    static constexpr bool SCALED = true;
    static constexpr bool DMA = false;  // staged through registers (site_gemm_split_kernel)
    __device__ static float4_t mfma(V a, V b, float4_t c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
    __device__ static float4_t prod(const V (&a)[NP], const V* b, float4_t acc) {
        acc = mfma(a[1], b[0], acc);
        acc = mfma(a[0], b[1], acc);
        return mfma(a[0], b[0], acc);
    }
};

template <class V>
__device__ __forceinline__ V neg16x8(V v) {  // sign flip of 8 16-bit floats (bf16 or fp16)
    return __builtin_bit_cast(V, __builtin_bit_cast(u32x4, v) ^ u32x4{0x80008000u, 0x80008000u, 0x80008000u,
                                                                      0x80008000u});
}

__device__ __forceinline__ void split2_fp16(float x, _Float16& h, _Float16& l) {
    h = (_Float16)x;
    l = (_Float16)(x - (float)h);
}

// The NP parts of one value in the format's 16-bit type (exact three-way bf16 split, or the
// two-way fp16 split of a value already divided by its column scale).
template <class F>
__device__ __forceinline__ void split_value(float x, typename F::E (&part)[F::NP]) {
    if constexpr (F::NP == 3) split3(x, part[0], part[1], part[2]);
    else split2_fp16(x, part[0], part[1]);
}

// Gamma's planes for the split-form contraction: parts p of lambda[k] gamma[k][c] (divided by
// colscale[c] in the scaled format) at x6_at(c, k, kpad, NPL, p) for c < ncols_pad, k < kpad, zero
// past chi_prev and ncols.  One tile of X6_KC k x 64 columns per iteration: the tile's rows are
// read 64 consecutive columns at a time (512 contiguous bytes per row), scaled as in an
// element-wise form, transposed through LDS, and each column's X6_KC k of every plane written as
// one contiguous NPL * 64-byte block, four threads per column with 16 bytes per plane each.
// (Thread per element with k fastest, the reads of one wave were 64 rows apart: 0.85 TB/s.)
template <class F>
__global__ __launch_bounds__(256) void split_gamma_kernel(const float* __restrict__ gamma,
                                                          const double* __restrict__ lam,
                                                          const float* __restrict__ colscale, int chi_prev, int ncols,
                                                          int kpad, int ncols_pad, typename F::E* __restrict__ planes) {
    typedef typename F::V V;
    constexpr int NP = F::NP, NPL = 2 * F::NP;
    __shared__ float2 tile[X6_KC][64 + 1];
    const int tid = threadIdx.x;
    const int nct = ncols_pad / 64;
    const int64_t ntiles = (int64_t)nct * (kpad / X6_KC);
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int c0 = (int)(t % nct) * 64, k0 = (int)(t / nct) * X6_KC;
        __syncthreads();  // the previous tile is no longer read
#pragma unroll
        for (int i = 0; i < X6_KC * 64 / 256; ++i) {
            const int e = tid + 256 * i, kk = e / 64, cc = e % 64, k = k0 + kk, c = c0 + cc;
            float2 g = make_float2(0.f, 0.f);
            if (c < ncols && k < chi_prev) {
                g = reinterpret_cast<const float2*>(gamma)[(int64_t)k * ncols + c];
                const double l = lam[k];
                g.x = (float)(g.x * l);
                g.y = (float)(g.y * l);
                if constexpr (F::SCALED) {
                    const float sc = colscale[c];
                    g.x = g.x / sc;
                    g.y = g.y / sc;
                }
            }
            tile[kk][cc] = g;
        }
        __syncthreads();
        const int cc = tid / 4, kq = (tid % 4) * 8;
        V out[NPL];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float2 g = tile[kq + j][cc];
            typename F::E re[NP], im[NP];
            split_value<F>(g.x, re);
            split_value<F>(g.y, im);
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                out[q][j] = re[q];
                out[NP + q][j] = im[q];
            }
        }
        V* dst = reinterpret_cast<V*>(planes + x6_at(c0 + cc, k0 + kq, kpad, NPL, 0));
#pragma unroll
        for (int q = 0; q < NPL; ++q) dst[q * (X6_KC / 8)] = out[q];
    }
}

// parts p of pre[r][k] (complex64 [n][chi_prev]) at x6_at(r, k, kpad, 6, p); k < kpad (zero padded)
__global__ void split_rows_bf16x6_kernel(const float* __restrict__ pre, int64_t n, int chi_prev, int kpad,
                                         __bf16* __restrict__ planes) {
    const int64_t total = n * kpad;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / kpad;
        const int k = (int)(i % kpad);
        const float2 v = k < chi_prev ? reinterpret_cast<const float2*>(pre)[r * chi_prev + k] : make_float2(0.f, 0.f);
        split3(v.x, planes[x6_at(r, k, kpad, 6, 0)], planes[x6_at(r, k, kpad, 6, 1)], planes[x6_at(r, k, kpad, 6, 2)]);
        split3(v.y, planes[x6_at(r, k, kpad, 6, 3)], planes[x6_at(r, k, kpad, 6, 4)], planes[x6_at(r, k, kpad, 6, 5)]);
    }
}

// Tile order of the GEMM grid.  Workgroups reach the 8 XCDs round-robin by linear id and each
// XCD has its own L2, so XCD x takes the column blocks [x gx / 8, (x + 1) gx / 8) and walks its
// (row block, column block) tiles in groups of 8 x 8: the ~64 workgroups an XCD runs at once then
// share 8 row blocks of A and 8 column blocks of B in its L2, whatever the aspect ratio of the
// product (n much smaller than chi d at chi = 10^4, or the reverse).  The grid is 8 x the
// largest per-XCD tile count; workgroups past an XCD's count return at once.
constexpr int kGemmXcds = 8, kTileGroup = 8;
__device__ __forceinline__ bool gemm_tile(int gx, int gy, int& rb, int& cb) {
    const int x = (int)(blockIdx.x % kGemmXcds);
    const int64_t s = blockIdx.x / kGemmXcds;
    const int c_lo = (int)((int64_t)x * gx / kGemmXcds), c_hi = (int)((int64_t)(x + 1) * gx / kGemmXcds);
    const int nc = c_hi - c_lo;
    if (nc <= 0 || s >= (int64_t)nc * gy) return false;
    const int64_t rg = s / ((int64_t)kTileGroup * nc);        // group of 8 row blocks
    const int R = min(kTileGroup, gy - (int)rg * kTileGroup);
    const int64_t off = s - rg * kTileGroup * nc;              // < R * nc
    const int cg = (int)(off / (R * kTileGroup));              // group of 8 column blocks
    const int off2 = (int)(off - (int64_t)cg * R * kTileGroup);  // < R * min(8, nc - 8 cg)
    rb = (int)rg * kTileGroup + off2 % R;
    cb = c_lo + cg * kTileGroup + off2 / R;
    return true;
}
inline int64_t gemm_tile_grid(int64_t gx, int64_t gy) {
    return kGemmXcds * ((gx + kGemmXcds - 1) / kGemmXcds) * gy;
}

// Sum of the k-segment partials of the split-form contraction: T[i] = (float) sum_s P_s[i], the
// partials added in FP64 in segment order (the same bits however the grid is shaped).
__global__ __launch_bounds__(256) void split_k_combine_kernel(const float* __restrict__ P, int64_t count, int segs,
                                                              float* __restrict__ T) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256) {
        double acc = 0.0;
        for (int s = 0; s < segs; ++s) acc += (double)P[(int64_t)s * count + i];
        T[i] = (float)acc;
    }
}

// The same sum written through a column map (ColMap): partial element i is part i % 2 of complex
// column (i / 2) % ncols of row (i / 2) / ncols.
__global__ __launch_bounds__(256) void split_k_combine_cols_kernel(const float* __restrict__ P, int64_t n, int ncols,
                                                                   int segs, float* __restrict__ T, ColMap om) {
    const int64_t count = n * ncols * 2;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256) {
        double acc = 0.0;
        for (int s = 0; s < segs; ++s) acc += (double)P[(int64_t)s * count + i];
        const int64_t e = i >> 1, row = e / ncols, oc = om.map[e - row * ncols];
        if (oc >= 0 && oc < om.ld) T[(row * om.ld + oc) * 2 + (i & 1)] = (float)acc;
    }
}

// Split-form contraction kernels: F::NP parts per value, 2 F::NP planes per operand (parts of
// the real part, then of the imaginary part).  With F::SCALED the planes hold values divided by a
// per-row (A) and per-column (B) scale, which the epilogue multiplies back.
//
// The k range may be cut into segments (blockIdx.y = segment, kseg values of k each, a multiple of
// X6_KC): segment s accumulates k in [s kseg, min((s + 1) kseg, kpad)) and writes its partial T to
// T + s * (n * ncols * 2) floats; split_k_combine_kernel then adds the partials in FP64.  A shorter
// FP32 accumulation chain per output is the point (at k = 10^4 the ~2000 FP32 roundings of one
// chain, not the products, set the error).  kseg = kpad is the unsegmented product.
//
// Two staging forms share the tile (128 rows x 64 complex columns x 32 complex k, four waves of
// 64 x 32), the LDS layout, the MFMA sequence of every accumulator (so both give the same bits)
// and the epilogue:
//   * registers (site_gemm_split_kernel, FP16x3): one LDS stage, the next k step fetched into
//     registers while this one is multiplied, two barriers per step, two workgroups per CU;
//   * LDS DMA (site_gemm_split_dma_kernel, BF16x6): two LDS stages filled straight from global
//     memory (global_load_lds_dwordx4), one barrier per step, one workgroup per CU.
// Each step has a fixed cost of about 800 cycles (barrier, the wait for the stage, the first
// fragment reads) that the second workgroup of the register form hides and the DMA form does
// not: the DMA form is 1.3x faster for BF16x6 (192 MFMAs per wave and step, and no spills in
// its register budget) and 10% slower for FP16x3 (96 MFMAs per wave and step; also with three
// stages, two steps in flight).

// Fragments of one k step from a stage in LDS ([plane][row][slot] of 16 bytes, slot x6_slot) into
// the wave's accumulators: lane holds k 8 (lane / 16) .. + 8 of row / column lane % 16.
template <class F>
__device__ __forceinline__ void split_multiply(const typename F::E* As, const typename F::E* Bs, int wm, int wn,
                                               int lane, float4_t (&acc_re)[4][2], float4_t (&acc_im)[4][2]) {
    typedef typename F::V V;
    constexpr int NP = F::NP, NPL = 2 * F::NP;
    const int g = lane / 16;
    V bfr[2][NPL];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int c = wn + 16 * j + lane % 16;
#pragma unroll
        for (int pl = 0; pl < NPL; ++pl)
            bfr[j][pl] = *reinterpret_cast<const V*>(Bs + ((pl * X6_TN + c) * 4 + x6_slot(c, g)) * 8);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = wm + 16 * i + lane % 16;
        V ar[NP], ai[NP], nai[NP];
#pragma unroll
        for (int t = 0; t < NP; ++t) {
            ar[t] = *reinterpret_cast<const V*>(As + ((t * X6_TM + r) * 4 + x6_slot(r, g)) * 8);
            ai[t] = *reinterpret_cast<const V*>(As + (((t + NP) * X6_TM + r) * 4 + x6_slot(r, g)) * 8);
            nai[t] = neg16x8(ai[t]);
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            acc_re[i][j] = F::prod(ar, &bfr[j][0], acc_re[i][j]);
            acc_re[i][j] = F::prod(nai, &bfr[j][NP], acc_re[i][j]);
            acc_im[i][j] = F::prod(ar, &bfr[j][NP], acc_im[i][j]);
            acc_im[i][j] = F::prod(ai, &bfr[j][0], acc_im[i][j]);
        }
    }
}

// Epilogue: output register r of lane l holds row 4 (l / 16) + r, column l % 16 of each block.
template <class F, bool MAPPED>
__device__ __forceinline__ void split_store(const float4_t (&acc_re)[4][2], const float4_t (&acc_im)[4][2],
                                            int64_t n, int ncols, int64_t row0, int col0, int wm, int wn, int lane,
                                            const float* __restrict__ rowscale, const float* __restrict__ colscale,
                                            float* __restrict__ T, ColMap om) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t srow = row0 + wm + 16 * i + 4 * (lane / 16) + r;
                const int c = col0 + wn + 16 * j + lane % 16;
                if (srow < n && c < ncols) {
                    float re = acc_re[i][j][r], im = acc_im[i][j][r];
                    if constexpr (F::SCALED) {
                        const float sc = rowscale[srow] * colscale[c];
                        re *= sc;
                        im *= sc;
                    }
                    if constexpr (MAPPED) {
                        const int64_t oc = om.map[c];
                        if (oc >= 0 && oc < om.ld) reinterpret_cast<float2*>(T)[srow * om.ld + oc] = make_float2(re, im);
                    } else {
                        reinterpret_cast<float2*>(T)[srow * ncols + c] = make_float2(re, im);
                    }
                }
            }
}

// Register staging.  Pieces of 16 bytes: piece w = 4 pl + g of row r is the NPL * 64-byte block
// of the step (x6_at).  A: 128 rows x 4 NPL pieces (2 NPL per thread, fetched into registers one
// step ahead), B: 64 columns x 4 NPL pieces (NPL per thread, fetched one step ahead as well);
// consecutive threads read consecutive pieces.  Rows past n read row n - 1 (their results are
// not stored).
template <class F, bool MAPPED>
__device__ __forceinline__ void split_gemm_registers(int64_t n, int ncols, int kpad, int ncols_pad,
                                                     const typename F::E* __restrict__ apl,
                                                     const float* __restrict__ rowscale,
                                                     const typename F::E* __restrict__ bpl,
                                                     const float* __restrict__ colscale, float* __restrict__ T,
                                                     int kseg, ColMap om) {
    typedef typename F::E E;
    constexpr int NPL = 2 * F::NP;
    __shared__ __attribute__((aligned(16))) E As[NPL * X6_TM * X6_KC];
    __shared__ __attribute__((aligned(16))) E Bs[NPL * X6_TN * X6_KC];
    int rb, cb;
    if (!gemm_tile(ncols_pad / X6_TN, (int)((n + X6_TM - 1) / X6_TM), rb, cb)) return;
    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int64_t row0 = (int64_t)rb * X6_TM;
    const int col0 = cb * X6_TN;  // complex column
    const int wm = (wave / 2) * 64, wn = (wave % 2) * 32;
    const int64_t ksteps = kpad / X6_KC;
    const int kbeg = (int)blockIdx.y * kseg;
    const int kend = kbeg + kseg < kpad ? kbeg + kseg : kpad;
    T += (int64_t)blockIdx.y * n * ncols * 2;

    float4_t acc_re[4][2], acc_im[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc_re[i][j] = acc_im[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};

    constexpr int PW = 4 * NPL;  // pieces per row and step
    u32x4 ra[2 * NPL], rbp[NPL];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 2 * NPL; ++u) {
            const int q = tid + 256 * u, r = q / PW, w = q % PW;
            const int64_t srow = row0 + r < n ? row0 + r : n - 1;
            ra[u] = *reinterpret_cast<const u32x4*>(apl + (srow * ksteps + k0 / X6_KC) * (NPL * X6_KC) + 8 * w);
        }
#pragma unroll
        for (int u = 0; u < NPL; ++u) {
            const int q = tid + 256 * u, c = q / PW, w = q % PW;
            rbp[u] = *reinterpret_cast<const u32x4*>(bpl + ((int64_t)(col0 + c) * ksteps + k0 / X6_KC) * (NPL * X6_KC) +
                                                     8 * w);
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int u = 0; u < 2 * NPL; ++u) {
            const int q = tid + 256 * u, r = q / PW, w = q % PW, pl = w / 4, g = w % 4;
            *reinterpret_cast<u32x4*>(As + ((pl * X6_TM + r) * 4 + x6_slot(r, g)) * 8) = ra[u];
        }
#pragma unroll
        for (int u = 0; u < NPL; ++u) {
            const int q = tid + 256 * u, c = q / PW, w = q % PW, pl = w / 4, g = w % 4;
            *reinterpret_cast<u32x4*>(Bs + ((pl * X6_TN + c) * 4 + x6_slot(c, g)) * 8) = rbp[u];
        }
    };

    fetch(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += X6_KC) {
        stash();
        __syncthreads();
        if (k0 + X6_KC < kend) fetch(k0 + X6_KC);
        split_multiply<F>(As, Bs, wm, wn, lane, acc_re, acc_im);
        __syncthreads();
    }
    split_store<F, MAPPED>(acc_re, acc_im, n, ncols, row0, col0, wm, wn, lane, rowscale, colscale, T, om);
}

// The register form as its own kernels, plain and with a column map (separate signatures keep
// the plain kernel's code independent of the map).
template <class F>
__global__ __launch_bounds__(256, 2) void site_gemm_split_kernel(int64_t n, int ncols, int kpad, int ncols_pad,
                                                                 const typename F::E* __restrict__ apl,
                                                                 const float* __restrict__ rowscale,
                                                                 const typename F::E* __restrict__ bpl,
                                                                 const float* __restrict__ colscale,
                                                                 float* __restrict__ T, int kseg) {
    split_gemm_registers<F, false>(n, ncols, kpad, ncols_pad, apl, rowscale, bpl, colscale, T, kseg, ColMap{0, nullptr});
}
template <class F>
__global__ __launch_bounds__(256, 2) void site_gemm_split_cols_kernel(int64_t n, int ncols, int kpad, int ncols_pad,
                                                                      const typename F::E* __restrict__ apl,
                                                                      const float* __restrict__ rowscale,
                                                                      const typename F::E* __restrict__ bpl,
                                                                      const float* __restrict__ colscale,
                                                                      float* __restrict__ T, int kseg, ColMap om) {
    split_gemm_registers<F, true>(n, ncols, kpad, ncols_pad, apl, rowscale, bpl, colscale, T, kseg, om);
}

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) const void global_void_t;

// 16 bytes per lane from global memory straight into LDS (global_load_lds_dwordx4): lane l of
// the wave writes dst + 16 l, so one instruction fills 1 KB of LDS; dst must be wave-uniform.
__device__ __forceinline__ void lds_dma16(const void* src, void* dst) {
    __builtin_amdgcn_global_load_lds((global_void_t*)src, (lds_void_t*)dst, 16, 0, 0);
}
// Waits for this wave's outstanding vector-memory loads, LDS DMA included: s_waitcnt vmcnt(0),
// with expcnt and lgkmcnt left at their maxima.
__device__ __forceinline__ void wait_vm_loads() { __builtin_amdgcn_s_waitcnt(0x0f70); }

// Stage of the split-form contraction in LDS: A (the carried vectors' tile) and B (Gamma's).
template <class F>
struct SplitStage {
    __attribute__((aligned(16))) typename F::E a[2 * F::NP * X6_TM * X6_KC];
    __attribute__((aligned(16))) typename F::E b[2 * F::NP * X6_TN * X6_KC];
};

// LDS DMA staging.  The two stages are separate LDS objects, so the compiler sees that the reads
// of one do not depend on the DMA in flight into the other and inserts no wait between them.
template <class F, bool MAPPED = false>
__global__ __launch_bounds__(256, 1) void site_gemm_split_dma_kernel(int64_t n, int ncols, int kpad, int ncols_pad,
                                                                     const typename F::E* __restrict__ apl,
                                                                     const float* __restrict__ rowscale,
                                                                     const typename F::E* __restrict__ bpl,
                                                                     const float* __restrict__ colscale,
                                                                     float* __restrict__ T, int kseg, ColMap om) {
    constexpr int NPL = 2 * F::NP;
    __shared__ SplitStage<F> st0, st1;
    int rb, cb;
    if (!gemm_tile(ncols_pad / X6_TN, (int)((n + X6_TM - 1) / X6_TM), rb, cb)) return;
    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int64_t row0 = (int64_t)rb * X6_TM;
    const int col0 = cb * X6_TN;  // complex column
    const int wm = (wave / 2) * 64, wn = (wave % 2) * 32;
    const int64_t ksteps = kpad / X6_KC;
    const int kbeg = (int)blockIdx.y * kseg;
    const int kend = kbeg + kseg < kpad ? kbeg + kseg : kpad;
    T += (int64_t)blockIdx.y * n * ncols * 2;

    float4_t acc_re[4][2], acc_im[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc_re[i][j] = acc_im[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};

    // Wave w's DMA instruction u fills the 64 consecutive stage pieces starting at 64 (4 u + w):
    // plane u / 2, rows 16 (4 (u % 2) + w) to + 15 for A (plane u, columns 16 w to + 15 for B),
    // four slots each; each lane loads the global piece that belongs in its position -- k group g
    // of that plane of its row, 16 contiguous bytes of the step's NPL * 64-byte block of the row
    // (x6_at).  A lane thus needs two row offsets and one column offset; the planes add
    // constants.  Rows past n read row n - 1 (their results are not stored).
    constexpr int AU = NPL * X6_TM * 4 / 256, BU = NPL * X6_TN * 4 / 256;  // DMA instructions per wave
    static_assert(AU == 2 * NPL && BU == NPL, "A: two row groups per plane, B: one column group per plane");
    int64_t aoff[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int r = (4 * h + wave) * 16 + lane / 4;
        const int64_t srow = row0 + r < n ? row0 + r : n - 1;
        aoff[h] = srow * ksteps * (NPL * X6_KC) + x6_slot(r, lane % 4) * 8;  // x6_slot is an involution in g
    }
    const int cdma = wave * 16 + lane / 4;
    const int64_t boff = (int64_t)(col0 + cdma) * ksteps * (NPL * X6_KC) + x6_slot(cdma, lane % 4) * 8;
    auto issue = [&](int k0, SplitStage<F>& stg) {
        const int64_t kb = (int64_t)(k0 / X6_KC) * (NPL * X6_KC);
#pragma unroll
        for (int u = 0; u < AU; ++u)
            lds_dma16(apl + aoff[u % 2] + kb + (u / 2) * X6_KC, stg.a + (4 * u + wave) * 64 * 8);
#pragma unroll
        for (int u = 0; u < BU; ++u) lds_dma16(bpl + boff + kb + u * X6_KC, stg.b + (4 * u + wave) * 64 * 8);
    };
    // One barrier per k step: after it, the stage of this step is complete (every wave waited for
    // its own DMA into it) and every wave has finished reading the stage of the step before, which
    // the DMA issued next then overwrites while this step is multiplied.
    issue(kbeg, st0);
    for (int k0 = kbeg; k0 < kend; k0 += 2 * X6_KC) {
        wait_vm_loads();
        __syncthreads();
        if (k0 + X6_KC < kend) issue(k0 + X6_KC, st1);
        split_multiply<F>(st0.a, st0.b, wm, wn, lane, acc_re, acc_im);
        if (k0 + X6_KC >= kend) break;
        wait_vm_loads();
        __syncthreads();
        if (k0 + 2 * X6_KC < kend) issue(k0 + 2 * X6_KC, st0);
        split_multiply<F>(st1.a, st1.b, wm, wn, lane, acc_re, acc_im);
    }
    split_store<F, MAPPED>(acc_re, acc_im, n, ncols, row0, col0, wm, wn, lane, rowscale, colscale, T, om);
}

// FP16x3 operand preparation.  Gamma: colscale[c] = max_k max(|Re|, |Im|) of lambda[k] gamma[k][c]
// (1 for an all-zero column), then the two fp16 parts of that entry divided by colscale[c] at
// x6_at(c, k, kpad, 4, p) (split_gamma_kernel).  Carried vectors: one wave per row, rowscale[r] =
// the row's largest component (1 for a zero row), the parts of pre[r][k] / rowscale[r] at
// x6_at(r, k, kpad, 4, p).
// colscale (as float bits) = max over k of max(|Re|, |Im|) of lambda[k] gamma[k][c]: blocks of
// 64 columns x a slice of kslice k, four k lanes per column; the block's maximum goes to the
// column by atomicMax on the bits (non-negative floats order as their bit patterns, and a
// maximum does not depend on the order it is taken in).  The caller zeroes the bits first;
// colscale_finish_kernel then turns them into the scale, 1 for an all-zero column.
__global__ __launch_bounds__(256) void colmax_fp16x3_kernel(const float* __restrict__ gamma,
                                                            const double* __restrict__ lam, int chi_prev, int ncols,
                                                            int kslice, unsigned* __restrict__ bits) {
    __shared__ float red[4][64];
    const int cc = threadIdx.x % 64, ks = threadIdx.x / 64;
    const int c = blockIdx.x * 64 + cc;
    const int kb = (int)blockIdx.y * kslice, ke = kb + kslice < chi_prev ? kb + kslice : chi_prev;
    float m = 0.f;
    if (c < ncols)
        for (int k = kb + ks; k < ke; k += 4) {
            const float2 g = reinterpret_cast<const float2*>(gamma)[(int64_t)k * ncols + c];
            const double l = lam[k];
            m = fmaxf(m, fmaxf(fabsf((float)(g.x * l)), fabsf((float)(g.y * l))));
        }
    red[ks][cc] = m;
    __syncthreads();
    if (ks == 0 && c < ncols) {
        m = fmaxf(fmaxf(red[0][cc], red[1][cc]), fmaxf(red[2][cc], red[3][cc]));
        if (m > 0.f) atomicMax(bits + c, __float_as_uint(m));
    }
}

__global__ void zero_u32_kernel(unsigned* __restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void colscale_finish_kernel(float* __restrict__ colscale, int ncols_pad) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncols_pad; c += gridDim.x * blockDim.x) {
        const float m = colscale[c];
        colscale[c] = m > 0.f ? m : 1.f;
    }
}

__global__ __launch_bounds__(256) void split_rows_fp16x3_kernel(const float* __restrict__ pre, int64_t n, int chi_prev,
                                                                int kpad, _Float16* __restrict__ planes,
                                                                float* __restrict__ rowscale) {
    const int lane = threadIdx.x % 64;
    for (int64_t r = (int64_t)blockIdx.x * 4 + threadIdx.x / 64; r < n; r += (int64_t)gridDim.x * 4) {
        const float2* row = reinterpret_cast<const float2*>(pre) + r * chi_prev;
        float m = 0.f;
        for (int k = lane; k < chi_prev; k += 64) m = fmaxf(m, fmaxf(fabsf(row[k].x), fabsf(row[k].y)));
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
        const float sc = m > 0.f ? m : 1.f;
        if (lane == 0) rowscale[r] = sc;
        for (int k = lane; k < kpad; k += 64) {
            const float2 v = k < chi_prev ? row[k] : make_float2(0.f, 0.f);
            split2_fp16(v.x / sc, planes[x6_at(r, k, kpad, 4, 0)], planes[x6_at(r, k, kpad, 4, 1)]);
            split2_fp16(v.y / sc, planes[x6_at(r, k, kpad, 4, 2)], planes[x6_at(r, k, kpad, 4, 3)]);
        }
    }
}

// ---------------------------------------------------------------------------------------
// Displacement + probabilities + draw, one workgroup per sample.
// ---------------------------------------------------------------------------------------
template <typename TT>
__device__ __forceinline__ cd ld_c(const TT* p, int64_t i);
template <>
__device__ __forceinline__ cd ld_c<double>(const double* p, int64_t i) {
    return ld_cd(p, i);
}
template <>
__device__ __forceinline__ cd ld_c<float>(const float* p, int64_t i) {
    const float2 v = reinterpret_cast<const float2*>(p)[i];
    return cd{(double)v.x, (double)v.y};
}
__device__ __forceinline__ void st_c(double* p, int64_t i, cd v) { st_cd(p, i, v); }
__device__ __forceinline__ void st_c(float* p, int64_t i, cd v) {
    reinterpret_cast<float2*>(p)[i] = make_float2((float)v.re, (float)v.im);
}

constexpr int kMaxDD = 64;
constexpr int kMaxd = 32;

// Orders LDS writes of one wave before its later LDS reads (the waves of a workgroup work on
// separate samples and separate LDS slices, so no workgroup barrier is needed).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// D(alpha_s) of every sample of one site, Dout[s][k][j] (k < dd, j < d, complex128): one wave per
// sample, cos/sin of |alpha| x_q in LDS (the spectral form of bsmps_batch_displaces).  Kept out
// of the draw kernel, whose occupancy the FP64 sincos registers would halve.
template <int DDMAX>
__global__ __launch_bounds__(256) void site_displace_kernel(int64_t n, int d, int dd, const double* __restrict__ alpha,
                                                            int64_t astride, const double* __restrict__ W,
                                                            const double* __restrict__ x, double* __restrict__ Dout) {
    __shared__ double s_cr[4][DDMAX], s_sr[4][DDMAX];
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    double* cr = s_cr[wave];
    double* sr = s_sr[wave];
    const int64_t n_rounds = (n + 3) / 4;
    for (int64_t rnd = blockIdx.x; rnd < n_rounds; rnd += gridDim.x) {
        const int64_t s = rnd * 4 + wave;
        if (s >= n) continue;  // uniform per wave
        const cd a = ld_cd(alpha, s * astride);
        const double r = hypot(a.re, a.im);
        const double theta = atan2(a.im, a.re) + 1.5707963267948966;
        if (lane < dd) {
            double sn, cs;
            sincos(r * x[lane], &sn, &cs);
            cr[lane] = cs;
            sr[lane] = -sn;  // e^{-i r x_q}
        }
        wave_lds_sync();
        for (int e = lane; e < dd * d; e += 64) {
            const int k = e / d, j = e - k * d;
            double re = 0.0, im = 0.0;
            for (int q = 0; q < dd; ++q) {
                const double w = W[k * dd + q] * W[j * dd + q];
                re = fma(w, cr[q], re);
                im = fma(w, sr[q], im);
            }
            double sp, cp;
            sincos((double)(k - j) * theta, &sp, &cp);
            st_cd(Dout, s * dd * d + e, cd{re * cp - im * sp, re * sp + im * cp});
        }
        wave_lds_sync();
    }
}

// DMAX / DDMAX bound d and dd at compile time so the per-thread arrays stay in VGPRs (the
// tightest of a few bounds is launched: fewer registers, more waves per SIMD).
// One wave per sample, four samples per workgroup, each wave independent; D(alpha_s) comes from
// site_displace_kernel.  TT: storage type of T and pre (double: FP64 form; float:
// reduced-precision forms, where the carried vector is rescaled after every site so that a long
// chain cannot underflow -- by 1/sqrt(p_pick), which makes its lam-weighted norm 1; the
// reference's GPU path rescales by the largest component, distributed_sampling.py:147, and the
// draws only see ratios of probabilities either way).  The probabilities and the inverse CDF are
// evaluated in FP64 in every form.  Used for d > 10 or dd > 20 (sample_site_gram_kernel below
// covers the smaller cutoffs).
template <int DMAX, int DDMAX, typename TT, bool RENORM>
__global__ __launch_bounds__(256) void sample_site_kernel(int64_t n, int chi, int d, int dd,
                                                          const TT* __restrict__ T, int64_t tstride,
                                                          const double* __restrict__ Dg,
                                                          const double* __restrict__ lam,
                                                          const double* __restrict__ uni, int32_t* __restrict__ samples,
                                                          int64_t sstride, TT* __restrict__ pre,
                                                          double* __restrict__ probs) {
    __shared__ cd s_D[4][DDMAX * DMAX];  // D[k][j], j < d
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    cd* Dm = s_D[wave];
    const int64_t n_rounds = (n + 3) / 4;
    for (int64_t rnd = blockIdx.x; rnd < n_rounds; rnd += gridDim.x) {
        const int64_t s = rnd * 4 + wave;
        const bool live = s < n;   // uniform per wave
        const int64_t se = live ? s : n - 1;
        for (int e = lane; e < dd * d; e += 64) {
            const int k = e / d, j = e - k * d;
            Dm[k * DMAX + j] = ld_cd(Dg, se * dd * d + e);
        }
        wave_lds_sync();
        const TT* Ts = T + 2 * se * tstride;
        // p[k] = sum_a |sum_j D[k][j] T[a][j]|^2 lam[a]^2   (last site: a = 0 only, no lambda)
        double pk[DDMAX];
#pragma unroll
        for (int k = 0; k < DDMAX; ++k) pk[k] = 0.0;
        const int a_end = lam ? chi : 1;
        for (int aa = lane; aa < a_end; aa += 64) {
            cd t[DMAX];
#pragma unroll
            for (int j = 0; j < DMAX; ++j) t[j] = j < d ? ld_c(Ts, (int64_t)aa * d + j) : czero();
            const double l2 = lam ? lam[aa] * lam[aa] : 1.0;
#pragma unroll
            for (int k = 0; k < DDMAX; ++k) {
                if (k < dd) {
                    double ur = 0.0, ui = 0.0;
#pragma unroll
                    for (int j = 0; j < DMAX; ++j) {
                        if (j < d) {
                            const cd dk = Dm[k * DMAX + j];
                            ur = fma(dk.re, t[j].re, fma(-dk.im, t[j].im, ur));
                            ui = fma(dk.re, t[j].im, fma(dk.im, t[j].re, ui));
                        }
                    }
                    pk[k] = fma(fma(ur, ur, ui * ui), l2, pk[k]);
                }
            }
        }
        // wave reduction; lane 0 ends with the totals
#pragma unroll
        for (int k = 0; k < DDMAX; ++k) {
            if (k < dd) {
                double v = pk[k];
                for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
                pk[k] = v;
            }
        }
        int pick = 0;
        double scale = 1.0;
        if (lane == 0) {
            double tot = 0.0;
#pragma unroll
            for (int k = 0; k < DDMAX; ++k)
                if (k < dd) tot += pk[k];
            const double u = uni[se];
            double cdf = 0.0;
            int cntk = 0;
#pragma unroll
            for (int k = 0; k < DDMAX; ++k) {
                if (k < dd) {
                    const double pnorm = pk[k] / tot;
                    cdf += pnorm;
                    if (u > cdf) ++cntk;
                    if (probs && live) probs[s * dd + k] = pnorm;
                }
            }
            pick = cntk < dd - 1 ? cntk : dd - 1;
            if (live) samples[s * sstride] = pick;
            double pp = 0.0;
#pragma unroll
            for (int k = 0; k < DDMAX; ++k)
                if (k == pick) pp = pk[k];
            scale = pp > 0.0 ? 1.0 / sqrt(pp) : 1.0;
        }
        pick = __shfl(pick, 0, 64);
        if constexpr (RENORM) scale = __shfl(scale, 0, 64);
        if (pre && live) {
            auto sel = [&](int aa) -> cd {
                double ur = 0.0, ui = 0.0;
                for (int j = 0; j < d; ++j) {
                    const cd dk = Dm[pick * DMAX + j];
                    const cd t = ld_c(Ts, (int64_t)aa * d + j);
                    ur = fma(dk.re, t.re, fma(-dk.im, t.im, ur));
                    ui = fma(dk.re, t.im, fma(dk.im, t.re, ui));
                }
                return cd{ur, ui};
            };
            for (int aa = lane; aa < chi; aa += 64) {
                cd u = sel(aa);
                if constexpr (RENORM) u = cscale(u, scale);
                st_c(pre, s * chi + aa, u);
            }
        }
        wave_lds_sync();
    }
}

// One DPP step of a 16-lane butterfly on a double (two 32-bit moves).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sum over the 16 lanes of a DPP row; every lane of the row ends with the same total (each
// step adds a pair of equal partial sums in either order).
__device__ __forceinline__ double row16_sum(double v) {
    v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);  // row_half_mirror
    v += dpp_f64<0x140>(v);  // row_mirror
    return v;
}

// Probabilities + draw + carried vector through the per-sample Gram matrix, for d <= 10 and
// dd <= 20: one workgroup of 256 threads per sample.  The sample's rows of T are staged through
// LDS in chunks of R rows with coalesced loads (consecutive threads read consecutive elements;
// row a's element j lands at a * RS + j, RS = d rounded up to odd against bank conflicts), and
// when all chi rows fit (R >= chi) the carried-vector pass reads them from LDS again instead of
// from HBM.  Pass 1 accumulates G = sum_a lam_a^2 t_a t_a^H (d(d+1)/2 complex entries in VGPRs,
// one row per thread per step); a DPP butterfly sums each 16-lane row, and the 16 row partials
// are added in a fixed order through LDS.  Thread k then forms p_k = D_k G D_k^H -- the same
// conditional probability as sum_a lam_a^2 |D_k t_a|^2 at (d+1)/(2 dd) of the FP64 work -- and
// thread 0 draws (sequential total and cdf, as in sample_site_kernel).  Pass 2 writes the carried
// vector pre[a] = D_pick t_a with the same operation order as sample_site_kernel, so the FP64
// form's carried vectors are the same bits.  RENORM scales the carried vector by 1/sqrt(p_pick),
// which makes its lam-weighted norm 1 at every site (the draws only see ratios of probabilities).
template <typename TT>
struct GramStage;
template <>
struct GramStage<double> {
    static constexpr int kBytes = 64 * 1024;  // two workgroups per CU (160 KB of LDS)
    using V = double2;
};
template <>
struct GramStage<float> {
    static constexpr int kBytes = 40 * 1024;  // three workgroups per CU with the other arrays
    using V = float2;
};

// SPLIT = 2 (d <= 10): the workgroup's two halves -- waves 0-1 and 2-3 -- accumulate the first
// and the second half of the G entries over the same rows, so each thread holds half of them
// (55 complex entries at d = 10 would not fit one thread's registers); each half's 8 row
// partials of an entry are then added in a fixed order.
template <int DMAX, int DDMAX, typename TT, bool RENORM, int SPLIT = 1>
__global__ __launch_bounds__(256) void sample_site_gram_kernel(int64_t n, int chi, int d, int dd,
                                                               const TT* __restrict__ T, int64_t tstride,
                                                               const double* __restrict__ Dg,
                                                               const double* __restrict__ lam,
                                                               const double* __restrict__ uni,
                                                               int32_t* __restrict__ samples, int64_t sstride,
                                                               TT* __restrict__ pre, double* __restrict__ probs,
                                                               int R, int RS, unsigned dmagic) {
    static_assert(DDMAX <= 64 && DMAX <= DDMAX, "one candidate photon number per thread");
    static_assert(SPLIT == 1 || SPLIT == 2, "the G entries are held by one or two halves of the workgroup");
    using V = typename GramStage<TT>::V;
    constexpr int NP = DMAX * (DMAX + 1) / 2;
    constexpr int NPH = (NP + SPLIT - 1) / SPLIT;  // G entries per thread
    constexpr int RSTEP = 256 / SPLIT;             // rows walked in parallel by one half
    constexpr int WPH = 16 / SPLIT;                // 16-lane rows per half
    __shared__ V s_T[GramStage<TT>::kBytes / sizeof(V)];
    __shared__ double s_part[16][2 * NP];  // per 16-lane row: G partial (re, im per entry)
    __shared__ double s_G[2 * NP];
    __shared__ double s_p[DDMAX];
    __shared__ double s_scale;
    __shared__ int s_pick;
    const int tid = threadIdx.x, r16 = tid % 16, row = tid / 16;
    const int half = tid / RSTEP, rtid = tid % RSTEP;  // wave-uniform half
    // rows [r0, r0 + nr) of the sample into s_T, four loads in flight per thread (more loads
    // per batch were measured slower); the caller synchronises around it
    auto stage = [&](const V* Tv, int r0, int nr) {
        const int ne = nr * d;
        const V* src = Tv + (int64_t)r0 * d;
        int e = tid;
        for (; e + 3 * 256 < ne; e += 4 * 256) {
            V v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = src[e + u * 256];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int ee = e + u * 256;
                const int a = (int)__umulhi((unsigned)ee, dmagic);
                s_T[a * RS + (ee - a * d)] = v[u];
            }
        }
        for (; e < ne; e += 256) {
            const int a = (int)__umulhi((unsigned)e, dmagic);
            s_T[a * RS + (e - a * d)] = src[e];
        }
    };
    auto row_of = [&](int a, cd* t) {
#pragma unroll
        for (int j = 0; j < DMAX; ++j) {
            if (j < d) {
                const V v = s_T[a * RS + j];
                t[j] = cd{(double)v.x, (double)v.y};
            } else {
                t[j] = czero();
            }
        }
    };
    for (int64_t s = blockIdx.x; s < n; s += gridDim.x) {
        const V* Tv = reinterpret_cast<const V*>(T) + s * tstride;
        double gr[NPH], gi[NPH];
#pragma unroll
        for (int p = 0; p < NPH; ++p) gr[p] = gi[p] = 0.0;
        const int a_end = lam ? chi : 1;
        for (int r0 = 0; r0 < a_end; r0 += R) {
            const int nr = a_end - r0 < R ? a_end - r0 : R;
            __syncthreads();  // the previous chunk / sample is no longer read
            stage(Tv, r0, nr);
            __syncthreads();
            for (int aa = rtid; aa < nr; aa += RSTEP) {
                const double l = lam ? lam[r0 + aa] : 1.0;
                cd t[DMAX];
                row_of(aa, t);
                int p = 0;
#pragma unroll
                for (int j = 0; j < DMAX; ++j) {
#pragma unroll
                    for (int q = j; q < DMAX; ++q, ++p) {
                        if (q < d && p / NPH == half) {
                            const cd x = cscale(t[j], l), y = cscale(t[q], l);
                            const int i = p % NPH;
                            if (q == j) {
                                gr[i] = fma(x.re, x.re, fma(x.im, x.im, gr[i]));
                            } else {  // t_j conj(t_q)
                                gr[i] = fma(x.re, y.re, fma(x.im, y.im, gr[i]));
                                gi[i] = fma(x.im, y.re, fma(-x.re, y.im, gi[i]));
                            }
                        }
                    }
                }
            }
        }
        {
            int p = 0;
#pragma unroll
            for (int j = 0; j < DMAX; ++j) {
#pragma unroll
                for (int q = j; q < DMAX; ++q, ++p) {
                    if (q < d && p / NPH == half) {
                        const double vr = row16_sum(gr[p % NPH]);
                        const double vi = q != j ? row16_sum(gi[p % NPH]) : 0.0;
                        if (r16 == 0) {
                            s_part[row][2 * p] = vr;
                            s_part[row][2 * p + 1] = vi;
                        }
                    }
                }
            }
        }
        __syncthreads();
        if (tid < 2 * NP) {  // entry tid / 2 was accumulated by the 16-lane rows of its half
            const int w0 = (tid / 2) / NPH * WPH;
            double acc = 0.0;
#pragma unroll
            for (int w = 0; w < WPH; ++w) acc += s_part[w0 + w][tid];
            s_G[tid] = acc;
        }
        __syncthreads();
        // p_k = sum_j |D_kj|^2 G_jj + 2 Re sum_{j<q} D_kj G_jq conj(D_kq), thread k
        if (tid < dd) {
            cd dk[DMAX];
#pragma unroll
            for (int j = 0; j < DMAX; ++j) dk[j] = j < d ? ld_cd(Dg, (s * dd + tid) * d + j) : czero();
            double diag = 0.0, off = 0.0;
            int p = 0;
#pragma unroll
            for (int j = 0; j < DMAX; ++j) {
#pragma unroll
                for (int q = j; q < DMAX; ++q, ++p) {
                    if (q < d) {
                        const double g_re = s_G[2 * p], g_im = s_G[2 * p + 1];
                        if (q == j) {
                            diag = fma(fma(dk[j].re, dk[j].re, dk[j].im * dk[j].im), g_re, diag);
                        } else {
                            const double zr = dk[j].re * g_re - dk[j].im * g_im;
                            const double zi = dk[j].re * g_im + dk[j].im * g_re;
                            off = fma(zr, dk[q].re, fma(zi, dk[q].im, off));
                        }
                    }
                }
            }
            s_p[tid] = fmax(fma(2.0, off, diag), 0.0);
        }
        __syncthreads();
        if (tid == 0) {
            double tot = 0.0;
            for (int k = 0; k < dd; ++k) tot += s_p[k];
            const double u = uni[s];
            double cdf = 0.0;
            int cntk = 0;
            for (int k = 0; k < dd; ++k) {
                const double pnorm = s_p[k] / tot;
                cdf += pnorm;
                if (u > cdf) ++cntk;
                if (probs) probs[s * dd + k] = pnorm;
            }
            const int pick = cntk < dd - 1 ? cntk : dd - 1;
            samples[s * sstride] = pick;
            const double pp = s_p[pick];
            s_pick = pick;
            s_scale = pp > 0.0 ? 1.0 / sqrt(pp) : 1.0;
        }
        __syncthreads();
        if (pre) {
            const int pick = s_pick;
            const double scale = s_scale;
            cd dp[DMAX];
#pragma unroll
            for (int j = 0; j < DMAX; ++j) dp[j] = j < d ? ld_cd(Dg, (s * dd + pick) * d + j) : czero();
            const bool resident = a_end == chi && R >= chi;  // pass 1 left every row in s_T
            for (int r0 = 0; r0 < chi; r0 += R) {
                const int nr = chi - r0 < R ? chi - r0 : R;
                if (!resident) {
                    __syncthreads();
                    stage(Tv, r0, nr);
                    __syncthreads();
                }
                for (int aa = tid; aa < nr; aa += 256) {
                    cd t[DMAX];
                    row_of(aa, t);
                    double ur = 0.0, ui = 0.0;
#pragma unroll
                    for (int j = 0; j < DMAX; ++j) {
                        if (j < d) {
                            ur = fma(dp[j].re, t[j].re, fma(-dp[j].im, t[j].im, ur));
                            ui = fma(dp[j].re, t[j].im, fma(dp[j].im, t[j].re, ui));
                        }
                    }
                    cd v{ur, ui};
                    if constexpr (RENORM) v = cscale(v, scale);
                    st_c(pre, s * chi + r0 + aa, v);
                }
            }
        }
        // s_T is next written behind the next sample's first barrier; s_part / s_G / s_p behind
        // its later barriers, s_pick / s_scale behind its third after the staging
    }
}

// Rows per staging chunk and the odd row stride of sample_site_gram_kernel.
template <typename TT>
void gram_stage_shape(int chi, int d, int& R, int& RS, unsigned& dmagic) {
    RS = d | 1;
    const int cap = GramStage<TT>::kBytes / (int)(RS * 2 * sizeof(TT));
    R = chi < cap ? chi : cap;
    dmagic = (unsigned)((((uint64_t)1 << 32) + d - 1) / d);
}

// ---------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
// ---------------------------------------------------------------------------------------
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ inline double u53(uint32_t lo, uint32_t hi) {
    const uint64_t v = ((uint64_t)hi << 32) | lo;
    return (double)(v >> 11) * (1.0 / 9007199254740992.0);
}

__global__ void philox_kernel(uint64_t seed, uint32_t stream0, int n_streams, int64_t first, int64_t n, int normal,
                              double* out) {
    const int64_t total = n * n_streams;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = idx / n;
        const int64_t e = first + (idx - r * n);
        uint32_t c[4] = {(uint32_t)e, (uint32_t)((uint64_t)e >> 32), stream0 + (uint32_t)r, 0u};
        philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        const double u1 = u53(c[0], c[1]);
        if (!normal) {
            out[idx] = u1;
        } else {
            const double u2 = u53(c[2], c[3]);
            out[idx] = sqrt(-2.0 * log(1.0 - u1)) * cos(6.283185307179586 * u2);
        }
    }
}

// Grid of the Gauss-form contraction: (column tiles, row tiles) of its workgroup tile.
template <typename E>
dim3 gauss_grid(int64_t n, int64_t ncols) {
    return dim3((unsigned)((ncols + GaussShape<E>::COLS - 1) / GaussShape<E>::COLS),
                (unsigned)((n + GaussShape<E>::ROWS - 1) / GaussShape<E>::ROWS));
}

int grid_for(int64_t work, int per_block) {
    const int64_t b = (work + per_block - 1) / per_block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 1 << 16));
}

}  // namespace

extern "C" int bsmps_mu_to_alpha(bsmps_ctx* ctx, const double* sqrtW, int M, const double* normals, int64_t n,
                                 double* alpha) {
    if (!ctx || M <= 0 || n < 0) {
        bsmps_set_error("bsmps_mu_to_alpha: invalid arguments");
        return BSMPS_ERR_INVALID;
    }
    if (n == 0) return BSMPS_OK;
    if (!sqrtW || !normals || !alpha) {
        bsmps_set_error("bsmps_mu_to_alpha: null pointer");
        return BSMPS_ERR_INVALID;
    }
    hipLaunchKernelGGL(mu_to_alpha_kernel, dim3(grid_for(n * M, 256)), dim3(256), 0, ctx->stream, sqrtW, M,
                       normals, n, alpha);
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

extern "C" int bsmps_batch_displaces(bsmps_ctx* ctx, int dd, const double* alpha, int64_t n, int64_t alpha_stride,
                                     const double* eigvec, const double* eigval, double* out) {
    if (!ctx || dd <= 0 || dd > kMaxDD || n < 0 || alpha_stride < 0) {
        bsmps_set_error("bsmps_batch_displaces: invalid arguments (dd must be 1..%d)", kMaxDD);
        return BSMPS_ERR_INVALID;
    }
    if (n == 0) return BSMPS_OK;
    if (!alpha || !eigvec || !eigval || !out) {
        bsmps_set_error("bsmps_batch_displaces: null pointer");
        return BSMPS_ERR_INVALID;
    }
    hipLaunchKernelGGL(batch_displaces_kernel, dim3(grid_for(n, 1)), dim3(128), sizeof(double) * 2 * dd, ctx->stream,
                       dd, alpha, n, alpha_stride, eigvec, eigval, out);
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

// FP64 (E = double) and FP32 (E = float) contraction in the form the context selects (Gauss or
// blocks), stored plainly (om null) or through a column map.
template <typename E>
static int site_contract_mfma(bsmps_ctx* ctx, const char* name, int64_t n, int chi_prev, int chi, int d, const E* pre,
                              const double* lambda_prev, const E* gamma, E* T, const ColMap* om) {
    constexpr int form = sizeof(E) == 8 ? 0 : 1;
    if (!ctx || n < 0 || chi_prev <= 0 || chi <= 0 || d <= 0 || (om && om->ld <= 0)) {
        bsmps_set_error("%s: invalid arguments", name);
        return BSMPS_ERR_INVALID;
    }
    if (n == 0) return BSMPS_OK;
    if (!pre || !lambda_prev || !gamma || !T || (om && !om->map)) {
        bsmps_set_error("%s: null pointer", name);
        return BSMPS_ERR_INVALID;
    }
    const int64_t ncols = (int64_t)chi * d;
    const int64_t gx = (2 * ncols + TN - 1) / TN, gy = (n + TM - 1) / TM;
    const int64_t gy_gauss = (n + GaussShape<E>::ROWS - 1) / GaussShape<E>::ROWS;
    if (gx > 0x7fffffff || gy > 65535 || (ctx->contract_gauss[form] && gy_gauss > 65535)) {
        bsmps_set_error("%s: problem too large for one launch", name);
        return BSMPS_ERR_UNSUPPORTED;
    }
    const ColMap none{0, nullptr};
    const dim3 bgrid((unsigned)gx, (unsigned)gy);
    if (ctx->contract_gauss[form]) {
        if (om)
            hipLaunchKernelGGL((site_gemm3m_kernel<E, true>), gauss_grid<E>(n, ncols), dim3(256), 0, ctx->stream, n,
                               chi_prev, (int)ncols, pre, lambda_prev, gamma, T, *om);
        else
            hipLaunchKernelGGL(site_gemm3m_kernel<E>, gauss_grid<E>(n, ncols), dim3(256), 0, ctx->stream, n, chi_prev,
                               (int)ncols, pre, lambda_prev, gamma, T, none);
    } else {
        if (om)
            hipLaunchKernelGGL((site_gemm_kernel<E, true>), bgrid, dim3(256), 0, ctx->stream, n, chi_prev, (int)ncols,
                               pre, lambda_prev, gamma, T, *om);
        else
            hipLaunchKernelGGL(site_gemm_kernel<E>, bgrid, dim3(256), 0, ctx->stream, n, chi_prev, (int)ncols, pre,
                               lambda_prev, gamma, T, none);
    }
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

extern "C" int bsmps_site_contract(bsmps_ctx* ctx, int64_t n, int chi_prev, int chi, int d, const double* pre,
                                   const double* lambda_prev, const double* gamma, double* T) {
    return site_contract_mfma<double>(ctx, "bsmps_site_contract", n, chi_prev, chi, d, pre, lambda_prev, gamma, T,
                                      nullptr);
}

extern "C" int bsmps_site_contract_cols(bsmps_ctx* ctx, int64_t n, int chi_prev, int chi, int d, const double* pre,
                                        const double* lambda_prev, const double* gamma, double* T, int64_t ldT,
                                        const int32_t* cmap) {
    const ColMap om{ldT, cmap};
    return site_contract_mfma<double>(ctx, "bsmps_site_contract_cols", n, chi_prev, chi, d, pre, lambda_prev, gamma,
                                      T, &om);
}

extern "C" int bsmps_sample_site(bsmps_ctx* ctx, int64_t n, int chi, int d, int dd, const double* T,
                                 int64_t T_sample_stride, const double* alpha, int64_t alpha_stride,
                                 const double* eigvec, const double* eigval, double* dwork, const double* lambda,
                                 const double* uniforms, int32_t* samples, int64_t samples_stride, double* pre,
                                 double* probs) {
    if (!ctx || n < 0 || chi <= 0 || d <= 0 || d > kMaxd || dd < d || dd > kMaxDD || T_sample_stride < 0 ||
        alpha_stride < 0 || samples_stride < 0) {
        bsmps_set_error("bsmps_sample_site: invalid arguments (d <= %d, d <= dd <= %d)", kMaxd, kMaxDD);
        return BSMPS_ERR_INVALID;
    }
    if (n == 0) return BSMPS_OK;
    if (!T || !alpha || !eigvec || !eigval || !dwork || !uniforms || !samples) {
        bsmps_set_error("bsmps_sample_site: null pointer");
        return BSMPS_ERR_INVALID;
    }
    const dim3 grid(grid_for((n + 3) / 4, 1)), block(256);
    if (dd <= 16)
        hipLaunchKernelGGL(site_displace_kernel<16>, grid, block, 0, ctx->stream, n, d, dd, alpha, alpha_stride, eigvec,
                           eigval, dwork);
    else
        hipLaunchKernelGGL(site_displace_kernel<kMaxDD>, grid, block, 0, ctx->stream, n, d, dd, alpha, alpha_stride,
                           eigvec, eigval, dwork);
    BSMPS_HIP(hipGetLastError());
#define BSMPS_SAMPLE_LAUNCH(DM, DDM)                                                                           \
    hipLaunchKernelGGL((sample_site_kernel<DM, DDM, double, false>), grid, block, 0, ctx->stream, n, chi, d, dd, T, \
                       T_sample_stride, dwork, lambda, uniforms, samples, samples_stride, pre, probs)
#define BSMPS_GRAM_LAUNCH(DM, DDM, SPL)                                                                         \
    hipLaunchKernelGGL((sample_site_gram_kernel<DM, DDM, double, false, SPL>), grid16, block, 0, ctx->stream, n, chi, \
                       d, dd, T, T_sample_stride, dwork, lambda, uniforms, samples, samples_stride, pre, probs, R, RS,  \
                       dmagic)
    const dim3 grid16(grid_for(n, 1));
    int R = 0, RS = 0;
    unsigned dmagic = 0;
    gram_stage_shape<double>(chi, d, R, RS, dmagic);
    if (d <= 4 && dd <= 8) BSMPS_GRAM_LAUNCH(4, 8, 1);
    else if (d <= 6 && dd <= 12) BSMPS_GRAM_LAUNCH(6, 12, 1);
    else if (d <= 8 && dd <= 16) BSMPS_GRAM_LAUNCH(8, 16, 1);
    else if (d <= 10 && dd <= 20) BSMPS_GRAM_LAUNCH(10, 20, 2);
    else if (d <= 16 && dd <= 32) BSMPS_SAMPLE_LAUNCH(16, 32);
    else BSMPS_SAMPLE_LAUNCH(32, 64);
#undef BSMPS_GRAM_LAUNCH
#undef BSMPS_SAMPLE_LAUNCH
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

extern "C" int bsmps_philox(bsmps_ctx* ctx, uint64_t seed, uint32_t stream0, int n_streams, int64_t first, int64_t n,
                            int normal, double* out) {
    if (!ctx || n < 0 || n_streams < 0 || first < 0) {
        bsmps_set_error("bsmps_philox: invalid arguments");
        return BSMPS_ERR_INVALID;
    }
    if (n == 0 || n_streams == 0) return BSMPS_OK;
    if (!out) {
        bsmps_set_error("bsmps_philox: null pointer");
        return BSMPS_ERR_INVALID;
    }
    hipLaunchKernelGGL(philox_kernel, dim3(grid_for(n * n_streams, 256)), dim3(256), 0, ctx->stream, seed, stream0,
                       n_streams, first, n, normal, out);
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

extern "C" int bsmps_site_contract_c64(bsmps_ctx* ctx, int64_t n, int chi_prev, int chi, int d, const float* pre,
                                       const double* lambda_prev, const float* gamma, float* T) {
    return site_contract_mfma<float>(ctx, "bsmps_site_contract_c64", n, chi_prev, chi, d, pre, lambda_prev, gamma, T,
                                     nullptr);
}

extern "C" int bsmps_site_contract_c64_cols(bsmps_ctx* ctx, int64_t n, int chi_prev, int chi, int d, const float* pre,
                                            const double* lambda_prev, const float* gamma, float* T, int64_t ldT,
                                            const int32_t* cmap) {
    const ColMap om{ldT, cmap};
    return site_contract_mfma<float>(ctx, "bsmps_site_contract_c64_cols", n, chi_prev, chi, d, pre, lambda_prev,
                                     gamma, T, &om);
}

extern "C" int bsmps_sample_site_c64(bsmps_ctx* ctx, int64_t n, int chi, int d, int dd, const float* T,
                                     int64_t T_sample_stride, const double* alpha, int64_t alpha_stride,
                                     const double* eigvec, const double* eigval, double* dwork, const double* lambda,
                                     const double* uniforms, int32_t* samples, int64_t samples_stride, float* pre,
                                     double* probs) {
    if (!ctx || n < 0 || chi <= 0 || d <= 0 || d > kMaxd || dd < d || dd > kMaxDD || T_sample_stride < 0 ||
        alpha_stride < 0 || samples_stride < 0) {
        bsmps_set_error("bsmps_sample_site_c64: invalid arguments (d <= %d, d <= dd <= %d)", kMaxd, kMaxDD);
        return BSMPS_ERR_INVALID;
    }
    if (n == 0) return BSMPS_OK;
    if (!T || !alpha || !eigvec || !eigval || !dwork || !uniforms || !samples) {
        bsmps_set_error("bsmps_sample_site_c64: null pointer");
        return BSMPS_ERR_INVALID;
    }
    const dim3 grid(grid_for((n + 3) / 4, 1)), block(256);
    if (dd <= 16)
        hipLaunchKernelGGL(site_displace_kernel<16>, grid, block, 0, ctx->stream, n, d, dd, alpha, alpha_stride, eigvec,
                           eigval, dwork);
    else
        hipLaunchKernelGGL(site_displace_kernel<kMaxDD>, grid, block, 0, ctx->stream, n, d, dd, alpha, alpha_stride,
                           eigvec, eigval, dwork);
    BSMPS_HIP(hipGetLastError());
#define BSMPS_SAMPLE_LAUNCH(DM, DDM)                                                                           \
    hipLaunchKernelGGL((sample_site_kernel<DM, DDM, float, true>), grid, block, 0, ctx->stream, n, chi, d, dd, T, \
                       T_sample_stride, dwork, lambda, uniforms, samples, samples_stride, pre, probs)
#define BSMPS_GRAM_LAUNCH(DM, DDM, SPL)                                                                         \
    hipLaunchKernelGGL((sample_site_gram_kernel<DM, DDM, float, true, SPL>), grid16, block, 0, ctx->stream, n, chi, \
                       d, dd, T, T_sample_stride, dwork, lambda, uniforms, samples, samples_stride, pre, probs, R, RS, \
                       dmagic)
    const dim3 grid16(grid_for(n, 1));
    int R = 0, RS = 0;
    unsigned dmagic = 0;
    gram_stage_shape<float>(chi, d, R, RS, dmagic);
    if (d <= 4 && dd <= 8) BSMPS_GRAM_LAUNCH(4, 8, 1);
    else if (d <= 6 && dd <= 12) BSMPS_GRAM_LAUNCH(6, 12, 1);
    else if (d <= 8 && dd <= 16) BSMPS_GRAM_LAUNCH(8, 16, 1);
    else if (d <= 10 && dd <= 20) BSMPS_GRAM_LAUNCH(10, 20, 2);
    else if (d <= 16 && dd <= 32) BSMPS_SAMPLE_LAUNCH(16, 32);
    else BSMPS_SAMPLE_LAUNCH(32, 64);
#undef BSMPS_GRAM_LAUNCH
#undef BSMPS_SAMPLE_LAUNCH
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

extern "C" int bsmps_split_bf16(bsmps_ctx* ctx, const float* x, int64_t count, void* hi, void* lo) {
    if (!ctx || count < 0) {
        bsmps_set_error("bsmps_split_bf16: invalid arguments");
        return BSMPS_ERR_INVALID;
    }
    if (count == 0) return BSMPS_OK;
    if (!x || !hi || !lo) {
        bsmps_set_error("bsmps_split_bf16: null pointer");
        return BSMPS_ERR_INVALID;
    }
    hipLaunchKernelGGL(split_bf16_kernel, dim3(grid_for(count, 256)), dim3(256), 0, ctx->stream, x, count,
                       reinterpret_cast<__bf16*>(hi), reinterpret_cast<__bf16*>(lo));
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

extern "C" int bsmps_site_contract_bf16x3(bsmps_ctx* ctx, int64_t n, int chi_prev, int chi, int d, const float* pre,
                                          const double* lambda_prev, const void* gamma_hi, const void* gamma_lo,
                                          float* T) {
    if (!ctx || n < 0 || chi_prev <= 0 || chi <= 0 || d <= 0) {
        bsmps_set_error("bsmps_site_contract_bf16x3: invalid arguments");
        return BSMPS_ERR_INVALID;
    }
    if (n == 0) return BSMPS_OK;
    if (!pre || !lambda_prev || !gamma_hi || !gamma_lo || !T) {
        bsmps_set_error("bsmps_site_contract_bf16x3: null pointer");
        return BSMPS_ERR_INVALID;
    }
    const int64_t ncols = (int64_t)chi * d;
    const int64_t gx = (2 * ncols + TN - 1) / TN, gy = (n + TM - 1) / TM;
    if (gx > 0x7fffffff || gy > 65535) {
        bsmps_set_error("bsmps_site_contract_bf16x3: problem too large for one launch");
        return BSMPS_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL(site_gemm_bf16x3_kernel, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, ctx->stream, n,
                       chi_prev, (int)ncols, pre, lambda_prev, reinterpret_cast<const __bf16*>(gamma_hi),
                       reinterpret_cast<const __bf16*>(gamma_lo), T);
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

// padded plane extents of the BF16x6 form
static int64_t x6_kpad(int chi_prev) { return ((int64_t)chi_prev + X6_KC - 1) / X6_KC * X6_KC; }
static int64_t x6_cols(int64_t ncols) { return (ncols + X6_TN - 1) / X6_TN * X6_TN; }

extern "C" int64_t bsmps_bf16x6_plane_bytes(int chi_prev, int chi, int d) {
    if (chi_prev <= 0 || chi <= 0 || d <= 0) return -1;
    return 6 * x6_kpad(chi_prev) * x6_cols((int64_t)chi * d) * (int64_t)sizeof(__bf16);
}

extern "C" int64_t bsmps_bf16x6_row_plane_bytes(int64_t n, int chi_prev) {
    if (n < 0 || chi_prev <= 0) return -1;
    return 6 * n * x6_kpad(chi_prev) * (int64_t)sizeof(__bf16);
}

extern "C" int bsmps_split_bf16x6(bsmps_ctx* ctx, int chi_prev, int chi, int d, const float* gamma,
                                  const double* lambda_prev, void* planes) {
    if (!ctx || chi_prev <= 0 || chi <= 0 || d <= 0) {
        bsmps_set_error("bsmps_split_bf16x6: invalid arguments");
        return BSMPS_ERR_INVALID;
    }
    if (!gamma || !lambda_prev || !planes) {
        bsmps_set_error("bsmps_split_bf16x6: null pointer");
        return BSMPS_ERR_INVALID;
    }
    const int64_t ncols = (int64_t)chi * d, kpad = x6_kpad(chi_prev), ncols_pad = x6_cols(ncols);
    if (ncols_pad > 0x7fffffff) {
        bsmps_set_error("bsmps_split_bf16x6: Gamma too large");
        return BSMPS_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL(split_gamma_kernel<Bf16x6Fmt>, dim3(grid_for(kpad / X6_KC * (ncols_pad / 64), 1)), dim3(256), 0,
                       ctx->stream, gamma, lambda_prev, nullptr, chi_prev, (int)ncols, (int)kpad, (int)ncols_pad,
                       reinterpret_cast<__bf16*>(planes));
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

extern "C" int bsmps_split_rows_bf16x6(bsmps_ctx* ctx, int64_t n, int chi_prev, const float* pre, void* planes) {
    if (!ctx || n < 0 || chi_prev <= 0) {
        bsmps_set_error("bsmps_split_rows_bf16x6: invalid arguments");
        return BSMPS_ERR_INVALID;
    }
    if (n == 0) return BSMPS_OK;
    if (!pre || !planes) {
        bsmps_set_error("bsmps_split_rows_bf16x6: null pointer");
        return BSMPS_ERR_INVALID;
    }
    const int64_t kpad = x6_kpad(chi_prev);
    hipLaunchKernelGGL(split_rows_bf16x6_kernel, dim3(grid_for(n * kpad, 256)), dim3(256), 0, ctx->stream, pre, n,
                       chi_prev, (int)kpad, reinterpret_cast<__bf16*>(planes));
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

// The split-form contraction of either format.  Long k runs in segments of kSplitKSeg values
// with the partials in the context workspace and an FP64 combine in segment order.  With a column
// map (om non-null) the unsegmented product is stored through it and a segmented one is combined
// through it; the partials keep their own contiguous rows.
template <class F>
int launch_split_contract(bsmps_ctx* ctx, int64_t n, int64_t ncols, int64_t kpad, int64_t ncols_pad, int64_t gx,
                          int64_t gy, const typename F::E* apl, const float* rowscale, const typename F::E* bpl,
                          const float* colscale, float* T, const ColMap* om) {
    const int segs = kpad > kSplitKMin ? (int)((kpad + kSplitKSeg - 1) / kSplitKSeg) : 1;
    float* out = T;
    const int64_t count = n * ncols * 2;
    if (segs > 1) {
        void* ws = nullptr;
        BSMPS_TRY(bsmps_ctx_workspace(ctx, sizeof(float) * (size_t)count * segs, &ws));
        out = static_cast<float*>(ws);
    }
    const dim3 grid((unsigned)gemm_tile_grid(gx, gy), (unsigned)segs);
    const int kseg = segs > 1 ? (int)kSplitKSeg : (int)kpad;
    const ColMap none{0, nullptr};
    if (om && segs == 1) {
        if constexpr (F::DMA)
            hipLaunchKernelGGL((site_gemm_split_dma_kernel<F, true>), grid, dim3(256), 0, ctx->stream, n, (int)ncols,
                               (int)kpad, (int)ncols_pad, apl, rowscale, bpl, colscale, out, kseg, *om);
        else
            hipLaunchKernelGGL(site_gemm_split_cols_kernel<F>, grid, dim3(256), 0, ctx->stream, n, (int)ncols,
                               (int)kpad, (int)ncols_pad, apl, rowscale, bpl, colscale, out, kseg, *om);
    } else {
        if constexpr (F::DMA)
            hipLaunchKernelGGL(site_gemm_split_dma_kernel<F>, grid, dim3(256), 0, ctx->stream, n, (int)ncols, (int)kpad,
                               (int)ncols_pad, apl, rowscale, bpl, colscale, out, kseg, none);
        else
            hipLaunchKernelGGL(site_gemm_split_kernel<F>, grid, dim3(256), 0, ctx->stream, n, (int)ncols, (int)kpad,
                               (int)ncols_pad, apl, rowscale, bpl, colscale, out, kseg);
    }
    BSMPS_HIP(hipGetLastError());
    if (segs > 1) {
        const dim3 cgrid((unsigned)std::min<int64_t>((count + 255) / 256, 65536));
        if (om)
            hipLaunchKernelGGL(split_k_combine_cols_kernel, cgrid, dim3(256), 0, ctx->stream, out, n, (int)ncols, segs,
                               T, *om);
        else
            hipLaunchKernelGGL(split_k_combine_kernel, cgrid, dim3(256), 0, ctx->stream, out, count, segs, T);
        BSMPS_HIP(hipGetLastError());
    }
    return BSMPS_OK;
}

static int site_contract_bf16x6(bsmps_ctx* ctx, const char* name, int64_t n, int chi_prev, int chi, int d,
                                const void* pre_planes, const void* gamma_planes, float* T, const ColMap* om) {
    if (!ctx || n < 0 || chi_prev <= 0 || chi <= 0 || d <= 0 || (om && om->ld <= 0)) {
        bsmps_set_error("%s: invalid arguments", name);
        return BSMPS_ERR_INVALID;
    }
    if (n == 0) return BSMPS_OK;
    if (!pre_planes || !gamma_planes || !T || (om && !om->map)) {
        bsmps_set_error("%s: null pointer", name);
        return BSMPS_ERR_INVALID;
    }
    const int64_t ncols = (int64_t)chi * d, kpad = x6_kpad(chi_prev), ncols_pad = x6_cols(ncols);
    const int64_t gx = ncols_pad / X6_TN, gy = (n + X6_TM - 1) / X6_TM;
    if (gy > 0x7fffffff || ncols > 0x7fffffff || gemm_tile_grid(gx, gy) > ((int64_t)1 << 32) / 256 - 1) {
        bsmps_set_error("%s: problem too large for one launch", name);
        return BSMPS_ERR_UNSUPPORTED;
    }
    return launch_split_contract<Bf16x6Fmt>(ctx, n, ncols, kpad, ncols_pad, gx, gy,
                                            reinterpret_cast<const __bf16*>(pre_planes), nullptr,
                                            reinterpret_cast<const __bf16*>(gamma_planes), nullptr, T, om);
}

extern "C" int bsmps_site_contract_bf16x6(bsmps_ctx* ctx, int64_t n, int chi_prev, int chi, int d,
                                          const void* pre_planes, const void* gamma_planes, float* T) {
    return site_contract_bf16x6(ctx, "bsmps_site_contract_bf16x6", n, chi_prev, chi, d, pre_planes, gamma_planes, T,
                                nullptr);
}

extern "C" int bsmps_site_contract_bf16x6_cols(bsmps_ctx* ctx, int64_t n, int chi_prev, int chi, int d,
                                               const void* pre_planes, const void* gamma_planes, float* T,
                                               int64_t ldT, const int32_t* cmap) {
    const ColMap om{ldT, cmap};
    return site_contract_bf16x6(ctx, "bsmps_site_contract_bf16x6_cols", n, chi_prev, chi, d, pre_planes, gamma_planes,
                                T, &om);
}

extern "C" int64_t bsmps_fp16x3_plane_bytes(int chi_prev, int chi, int d) {
    if (chi_prev <= 0 || chi <= 0 || d <= 0) return -1;
    return 4 * x6_kpad(chi_prev) * x6_cols((int64_t)chi * d) * (int64_t)sizeof(_Float16);
}

extern "C" int64_t bsmps_fp16x3_row_plane_bytes(int64_t n, int chi_prev) {
    if (n < 0 || chi_prev <= 0) return -1;
    return 4 * n * x6_kpad(chi_prev) * (int64_t)sizeof(_Float16);
}

extern "C" int bsmps_split_fp16x3(bsmps_ctx* ctx, int chi_prev, int chi, int d, const float* gamma,
                                  const double* lambda_prev, void* planes, float* colscale) {
    if (!ctx || chi_prev <= 0 || chi <= 0 || d <= 0) {
        bsmps_set_error("bsmps_split_fp16x3: invalid arguments");
        return BSMPS_ERR_INVALID;
    }
    if (!gamma || !lambda_prev || !planes || !colscale) {
        bsmps_set_error("bsmps_split_fp16x3: null pointer");
        return BSMPS_ERR_INVALID;
    }
    const int64_t ncols = (int64_t)chi * d, kpad = x6_kpad(chi_prev), ncols_pad = x6_cols(ncols);
    if (ncols_pad > 0x7fffffff) {
        bsmps_set_error("bsmps_split_fp16x3: Gamma too large");
        return BSMPS_ERR_UNSUPPORTED;
    }
    // column maxima over k slices, enough blocks to fill the GPU (a fill kernel, not a memset, zeroes
    // the bits, so the call can be captured in a graph)
    unsigned* bits = reinterpret_cast<unsigned*>(colscale);
    const int64_t cblocks = (ncols + 63) / 64;
    const int64_t slices = std::max<int64_t>(1, std::min<int64_t>((chi_prev + 63) / 64, (4096 + cblocks - 1) / cblocks));
    const int kslice = (int)((chi_prev + slices - 1) / slices);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(grid_for(ncols_pad, 256)), dim3(256), 0, ctx->stream, bits, ncols_pad);
    BSMPS_HIP(hipGetLastError());
    hipLaunchKernelGGL(colmax_fp16x3_kernel, dim3((unsigned)cblocks, (unsigned)((chi_prev + kslice - 1) / kslice)),
                       dim3(256), 0, ctx->stream, gamma, lambda_prev, chi_prev, (int)ncols, kslice, bits);
    BSMPS_HIP(hipGetLastError());
    hipLaunchKernelGGL(colscale_finish_kernel, dim3(grid_for(ncols_pad, 256)), dim3(256), 0, ctx->stream, colscale,
                       (int)ncols_pad);
    BSMPS_HIP(hipGetLastError());
    hipLaunchKernelGGL(split_gamma_kernel<Fp16x3Fmt>, dim3(grid_for(kpad / X6_KC * (ncols_pad / 64), 1)), dim3(256), 0,
                       ctx->stream, gamma, lambda_prev, colscale, chi_prev, (int)ncols, (int)kpad, (int)ncols_pad,
                       reinterpret_cast<_Float16*>(planes));
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

extern "C" int bsmps_split_rows_fp16x3(bsmps_ctx* ctx, int64_t n, int chi_prev, const float* pre, void* planes,
                                       float* rowscale) {
    if (!ctx || n < 0 || chi_prev <= 0) {
        bsmps_set_error("bsmps_split_rows_fp16x3: invalid arguments");
        return BSMPS_ERR_INVALID;
    }
    if (n == 0) return BSMPS_OK;
    if (!pre || !planes || !rowscale) {
        bsmps_set_error("bsmps_split_rows_fp16x3: null pointer");
        return BSMPS_ERR_INVALID;
    }
    hipLaunchKernelGGL(split_rows_fp16x3_kernel, dim3(grid_for((n + 3) / 4, 1)), dim3(256), 0, ctx->stream, pre, n,
                       chi_prev, (int)x6_kpad(chi_prev), reinterpret_cast<_Float16*>(planes), rowscale);
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

static int site_contract_fp16x3(bsmps_ctx* ctx, const char* name, int64_t n, int chi_prev, int chi, int d,
                                const void* pre_planes, const float* rowscale, const void* gamma_planes,
                                const float* colscale, float* T, const ColMap* om) {
    if (!ctx || n < 0 || chi_prev <= 0 || chi <= 0 || d <= 0 || (om && om->ld <= 0)) {
        bsmps_set_error("%s: invalid arguments", name);
        return BSMPS_ERR_INVALID;
    }
    if (n == 0) return BSMPS_OK;
    if (!pre_planes || !rowscale || !gamma_planes || !colscale || !T || (om && !om->map)) {
        bsmps_set_error("%s: null pointer", name);
        return BSMPS_ERR_INVALID;
    }
    const int64_t ncols = (int64_t)chi * d, kpad = x6_kpad(chi_prev), ncols_pad = x6_cols(ncols);
    const int64_t gx = ncols_pad / X6_TN, gy = (n + X6_TM - 1) / X6_TM;
    if (gy > 0x7fffffff || ncols > 0x7fffffff || gemm_tile_grid(gx, gy) > ((int64_t)1 << 32) / 256 - 1) {
        bsmps_set_error("%s: problem too large for one launch", name);
        return BSMPS_ERR_UNSUPPORTED;
    }
    return launch_split_contract<Fp16x3Fmt>(ctx, n, ncols, kpad, ncols_pad, gx, gy,
                                            reinterpret_cast<const _Float16*>(pre_planes), rowscale,
                                            reinterpret_cast<const _Float16*>(gamma_planes), colscale, T, om);
}

extern "C" int bsmps_site_contract_fp16x3(bsmps_ctx* ctx, int64_t n, int chi_prev, int chi, int d,
                                          const void* pre_planes, const float* rowscale, const void* gamma_planes,
                                          const float* colscale, float* T) {
    return site_contract_fp16x3(ctx, "bsmps_site_contract_fp16x3", n, chi_prev, chi, d, pre_planes, rowscale,
                                gamma_planes, colscale, T, nullptr);
}

extern "C" int bsmps_site_contract_fp16x3_cols(bsmps_ctx* ctx, int64_t n, int chi_prev, int chi, int d,
                                               const void* pre_planes, const float* rowscale,
                                               const void* gamma_planes, const float* colscale, float* T,
                                               int64_t ldT, const int32_t* cmap) {
    const ColMap om{ldT, cmap};
    return site_contract_fp16x3(ctx, "bsmps_site_contract_fp16x3_cols", n, chi_prev, chi, d, pre_planes, rowscale,
                                gamma_planes, colscale, T, &om);
}
