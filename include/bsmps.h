/* bsmps -- MI355X (gfx950) MPS construction and sampling for lossy Gaussian boson sampling.
 *
 * C ABI.  Plain pointers and sizes; no framework types.  Conventions:
 *   - every entry point returns BSMPS_OK (0) or an error code; bsmps_last_error() gives a
 *     thread-local message for the last failure;
 *   - array arguments are caller-owned DEVICE memory unless the name ends in _host;
 *   - complex arrays are complex128, interleaved (re, im), row-major, like NumPy's;
 *   - work is enqueued on the context's stream and is asynchronous unless stated;
 *     bsmps_synchronize() waits for it;
 *   - a context is used from one host thread at a time; contexts are independent.
 * Sizes are checked on the host before any launch; a bad size returns BSMPS_ERR_INVALID
 * and launches nothing.
 *
 * Reference interfaces each entry point replaces are cited as file:line of the reference
 * repository (sss441803/BosonSupremacy).
 */
#ifndef BSMPS_H
#define BSMPS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    BSMPS_OK = 0,
    BSMPS_ERR_INVALID = 1,     /* bad argument (size, null pointer, unsupported shape) */
    BSMPS_ERR_HIP = 2,         /* a HIP runtime call failed */
    BSMPS_ERR_UNSUPPORTED = 3, /* valid request outside what this build implements */
};

/* Largest hafnian matrix size (rows) the hafnian entry points accept. */
#define BSMPS_MAX_HAFNIAN_SIZE 34
/* Largest matrix size (rows) the loop-hafnian entry points accept with a loop-weight vector. */
#define BSMPS_MAX_LOOP_HAFNIAN_SIZE 24

typedef struct bsmps_ctx bsmps_ctx;

/* ---- context --------------------------------------------------------------------------- */

/* Create a context on HIP device `device` that enqueues all work on `stream` (a hipStream_t;
 * NULL is HIP's null stream, which is PyTorch's default stream).  Callers must enqueue the
 * producers and consumers of the buffers they pass on the same stream, or synchronise.
 * (The reference has no context: it runs on CuPy's current device and default stream,
 * MPS_utils.py:15-19.) */
int bsmps_create(int device, void *stream, bsmps_ctx **out);
int bsmps_destroy(bsmps_ctx *ctx);
int bsmps_synchronize(bsmps_ctx *ctx);
const char *bsmps_last_error(void);
int bsmps_version(void);
/* Launch tuning (no reference counterpart; set by bosonsupremacy_amd/autotune.py):
 * hafnians of size <= 2*lane_max_m use one lane per matrix (lane_max_m in 0..5; sizes <= 4
 * always do); larger ones use `group_lanes` (16, 32 or 64) cooperating lanes per matrix,
 * for every size (bsmps_set_tuning) or for size n only (bsmps_set_group_lanes). */
int bsmps_set_tuning(bsmps_ctx *ctx, int lane_max_m, int group_lanes);
int bsmps_set_group_lanes(bsmps_ctx *ctx, int n, int group_lanes);
/* Hafnians of size n in 12..34 take the frontier-split path when `enable` is 1 (default): the
 * upper part of the recursion tree runs breadth-first, one launch per tree level, with its
 * polynomials in device scratch, and every subtree small enough for one lane's registers is
 * evaluated by a one-lane-per-subtree kernel; results are identical to the unsplit path (no
 * reference counterpart). */
int bsmps_set_split(bsmps_ctx *ctx, int n, int enable);
/* Device scratch the split path may hold (default 16 GiB, at least 1 MiB): batches of
 * hafnians are processed in chunks whose frontier data fits it, with at least 64 hafnians
 * per chunk (one wavefront), which alone needs 2.2 GB at size 30, 4.8 GB at 32 and 10.3 GB
 * at 34. */
int bsmps_set_split_budget(bsmps_ctx *ctx, size_t bytes);
/* Form of the site contraction for `precision` 0 (bsmps_site_contract, FP64) or 1
 * (bsmps_site_contract_c64, FP32): gauss = 0 runs the complex product as one real GEMM with
 * 2x2 real blocks (8 n chi_prev chi d flops, default), gauss = 1 as three real products
 * (Gauss: 6 n chi_prev chi d flops, more registers per tile, a few ulps more rounding).
 * Launch tuning only; no reference counterpart. */
int bsmps_set_contract_form(bsmps_ctx *ctx, int precision, int gauss);

/* ---- hafnians (L3) --------------------------------------------------------------------- */

/* Sigma_select, MPS_cpu.py:19-37 / direct_mps_kernels.cu:22-39 (sigma_select_cdouble):
 * out[l, r, c] = sigma[targets[l, r] * ld + targets[l, c]].
 * sigma [ld][ld], targets int32 [n_batch][n_select], out [n_batch][n_select][n_select]. */
int bsmps_sigma_select(bsmps_ctx *ctx, const double *sigma, int ld, const int32_t *targets,
                       int64_t n_batch, int n_select, double *out);

/* hafnian, MPS_cpu.py:117-187 (dispatch + recursive_hafnian + solve): out[b] = haf(A[b]) for
 * A [n_batch][n][n].  Bit-exact to the reference's complex128 evaluation order.  Odd n gives
 * zeros, n = 0 gives ones (MPS_cpu.py:122-126). */
int bsmps_hafnian_dense(bsmps_ctx *ctx, const double *A, int64_t n_batch, int n, double *out);

/* A_elem, MPS_cpu.py:334-355 (MPS_utils.py:273-297): out[b] = haf(Sigma[t_b][:, t_b]) / den[b]
 * with the gather fused (the n x n matrices are never materialised).  den may be NULL
 * (no division).  Division is NumPy's complex/real: both parts times 1/den. */
int bsmps_hafnian_targets(bsmps_ctx *ctx, const double *sigma, int ld, const int32_t *targets,
                          int64_t n_batch, int n_select, const double *den, double *out);

/* The hafnian entry points with an optional loop-weight vector (a vertex may be matched to
 * itself with weight loops[i]).  The reference computes plain hafnians only (MPS_cpu.py:117-140:
 * its pure state has zero mean), so loops = NULL is the parity case and calls the plain entry
 * point (bit-exact to the reference).  With loops, the reference recursion extended by one loop
 * polynomial per row (sizes 1..BSMPS_MAX_LOOP_HAFNIAN_SIZE; an odd size is padded with a vertex
 * of loop weight 1 and no edges); an all-zero vector gives the plain recursion's value.
 * Dense: loops [n_batch][n], one vector per matrix.  Targets: loops [ld], a weight per row of
 * sigma, gathered through the targets like sigma (row r of a matrix has weight loops[t_r]). */
int bsmps_loop_hafnian_dense(bsmps_ctx *ctx, const double *A, const double *loops, int64_t n_batch, int n,
                             double *out);
int bsmps_loop_hafnian_targets(bsmps_ctx *ctx, const double *sigma, int ld, const double *loops,
                               const int32_t *targets, int64_t n_batch, int n_select, const double *den,
                               double *out);

/* One site tensor Gamma_k, the j / photon-total / chunk loops of MPS_cpu.py:402-520
 * (first site :402-428, last site :430-459, middle sites :461-520; GPU path
 * distributed_MPS.py:160-258).  Every entry Gamma[b, a, j] with j + |a| + |b| even is
 * computed from its labels on the device:
 *   target = [0]*j ++ left_targets[a] ++ right_targets[b]      (Sigma indices)
 *   Gamma  = ((haf * (1/den)) * (1/Z)) * (1/lambda[a]),  den = (sqrt(j!) * left_den[a]) * right_den[b]
 * and, for kind != FIRST, the vacuum entry (j = |a| = |b| = 0) is (1/Z) / lambda[a]
 * (MPS_cpu.py:449-451, 494-496).  gamma must be zeroed by the caller; odd-size entries are
 * left untouched (the reference's hafnian is 0 there). */
enum { BSMPS_SITE_FIRST = 0, BSMPS_SITE_MIDDLE = 1, BSMPS_SITE_LAST = 2 };
typedef struct {
    int kind;                     /* BSMPS_SITE_* */
    int d;                        /* local cutoff: j = 0..d-1 */
    int chi;                      /* Gamma is [n_right][chi][d] (the reference: [chi][chi][d]); n_left <= chi */
    int ld;                       /* Sigma is [ld][ld], ld = 2 M' */
    const double *sigma;          /* device complex128 [ld][ld] */
    double z;                     /* Z = sqrt(prod cosh(sq)), MPS_cpu.py:483 */
    double inv_z;                 /* 1.0 / Z */
    const double *sqrt_fact;      /* device [d]: sqrt(j!) */
    /* left labels a (Schmidt vectors of cut k; a single empty label for the last site) */
    int n_left, w_left;           /* count, row width of left_targets */
    const int32_t *left_targets;  /* device [n_left][w_left], Sigma indices in 1..M'-1 */
    const double *left_den;       /* device [n_left]: sqrt(prod num_a!) (1.0 for last site) */
    const double *lambda;         /* device [n_left]: Lambda_k[a] (1.0 for last site) */
    const int32_t *left_perm;     /* device [n_left]: labels ordered by photon count */
    const int32_t *left_class_count_host; /* host [n_left_class]: labels with p photons */
    int n_left_class;
    /* right labels b (Schmidt vectors of cut k-1; a single empty label for the first site) */
    int n_right, w_right;
    const int32_t *right_targets; /* device [n_right][w_right], Sigma indices in M'..2M'-1 */
    const double *right_den;      /* device [n_right] (1.0 for the first site) */
    const int32_t *right_perm;    /* device [n_right] */
    const int32_t *right_class_count_host; /* host [n_right_class] */
    int n_right_class;
} bsmps_gamma_site_args;

int bsmps_gamma_site(bsmps_ctx *ctx, const bsmps_gamma_site_args *args, double *gamma);

/* Many site tensors at once, of one instance or of several: args[i] describes site i and
 * gammas[i] is its output (zeroed by the caller), as for bsmps_gamma_site.  Per even photon total
 * one hafnian launch covers the (j, a, b) items of every listed site, so sites and instances with
 * uneven cutoffs, ranks and hafnian sizes share launches.  The values are those of
 * bsmps_gamma_site, bit for bit.  Replaces the per-site loop of MPS_cpu.py:384-525 and the
 * site-per-rank offload of distributed_MPS.py:188-246 for batches of independent instances. */
int bsmps_gamma_sites(bsmps_ctx *ctx, int n_sites, const bsmps_gamma_site_args *args, double *const *gammas);

/* ---- Schmidt spectra (L2) -------------------------------------------------------------- */

/* The pruned Kronecker loop of get_cumsum_kron, kron_cpu.py:106-125 (GPU path
 * distributed_kron.py:105-137), for n_cuts independent cuts in one launch.  Cut c has
 * n_modes_host[c] >= 1 modes with per-mode thermal distributions tp[c][i][0..d-1]
 * (thermal_photons, kron_cpu.py:92-93, evaluated on the host).  Per step the candidates
 * res[a]*tp[i][n] above err_tol are reduced to the max_dim largest; the chi largest of the
 * last step are returned in descending order with their occupation labels.
 *   tp   device double [n_cuts][max_modes][d]
 *   res  device double [n_cuts][chi]            (descending; unused tail zero)
 *   num  device int8   [n_cuts][chi][max_modes] (label of each kept weight)
 *   len  device int32  [n_cuts]                 (number of kept weights, <= chi)
 * Exact ties are ordered by (mode-i photon number, then earlier rank); the reference's
 * argsort leaves their order unspecified. */
int bsmps_kron_cuts(bsmps_ctx *ctx, int n_cuts, const int32_t *n_modes_host, int max_modes, int d,
                    const double *tp, int64_t max_dim, int chi, double err_tol, double *res,
                    int8_t *num, int32_t *len);

/* Allocation-free variant of bsmps_kron_cuts.  It has no counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, results bit-identical to
 * bsmps_kron_cuts.  Its scratch is a fixed device array of bsmps_kron_noalloc_scratch_bytes()
 * bytes (256 MiB) per GPU, held until the process ends; it calls no hipMalloc/hipFree, copies
 * nothing from the host and does not synchronise (cut descriptors travel in the kernel
 * arguments, cuts run in launches of at most 64 that fit the scratch).  A cut needing more
 * scratch than that -- max_dim * (32 + 4 * (n_modes - 1)) bytes, about 1300 modes at
 * max_dim = 1e5 -- is refused with BSMPS_ERR_UNSUPPORTED before anything is launched.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU (other streams, other
 * contexts, other threads) race. */
int bsmps_kron_cuts_noalloc(bsmps_ctx *ctx, int n_cuts, const int32_t *n_modes_host, int max_modes,
                            int d, const double *tp, int64_t max_dim, int chi, double err_tol,
                            double *res, int8_t *num, int32_t *len);
int64_t bsmps_kron_noalloc_scratch_bytes(void);

/* ---- sampling (L4) --------------------------------------------------------------------- */

/* alpha[s][i] = (mu[s][i] + 1j mu[s][M+i]) / 2 with mu = sqrtW @ normals[:, s]
 * (sampling_cpu.py:37-41 and batch_mu_to_alpha :156-159, hbar = 2).
 * sqrtW device [2M][2M], normals device [2M][n], alpha device complex128 [n][M]. */
int bsmps_mu_to_alpha(bsmps_ctx *ctx, const double *sqrtW, int M, const double *normals,
                      int64_t n, double *alpha);

/* batch_displaces, sampling_cpu.py:133-148: D[s] = expm(alpha_s a^dag - conj(alpha_s) a) of the
 * dd x dd truncated ladder operators, in FP64 via the exact spectral form
 * D[m][k] = e^{i(m-k)theta} sum_q W[m][q] W[k][q] e^{-i|alpha| x_q}, theta = arg(alpha)+pi/2,
 * where W, x are the eigenvectors/eigenvalues of a + a^dag (dd x dd, computed once on the
 * host).  alpha complex128 with element stride alpha_stride (in complex elements);
 * eigvec [dd][dd] (row m, column q), eigval [dd]; out complex128 [n][dd][dd]. */
int bsmps_batch_displaces(bsmps_ctx *ctx, int dd, const double *alpha, int64_t n,
                          int64_t alpha_stride, const double *eigvec, const double *eigval,
                          double *out);

/* Site contraction, sampling_cpu.py:75-80 (distributed_sampling.py:116-117):
 * T[s][a][j] = sum_b (pre[s][b] * lambda_prev[b]) * gamma[b][a][j],  j < d,
 * as an FP64 MFMA GEMM.  pre [n][chi_prev], gamma [chi_prev][chi][d], T [n][chi][d]. */
int bsmps_site_contract(bsmps_ctx *ctx, int64_t n, int chi_prev, int chi, int d,
                        const double *pre, const double *lambda_prev, const double *gamma,
                        double *T);

/* Displacement, conditional probabilities and draw of one site, sampling_cpu.py:81-113
 * (first site :55-72): per sample s,
 *   U[a][k]  = sum_{j<d} D_s[k][j] T[s][a][j]         (D_s from alpha_s as above)
 *   p[k]     = sum_a |U[a][k]|^2 lambda[a]^2, or |U[0][k]|^2 when lambda is NULL (last site)
 *   n_s      = #{k : u_s > cumsum(p / sum p)[k]}, clamped to dd-1
 *   pre[s][a] = U[a][n_s]
 * T is read at T + s*T_sample_stride (complex elements; stride 0 broadcasts one [chi][d]
 * tensor, as the first site does).  probs (may be NULL) receives p / sum p [n][dd].
 * dwork: caller-owned scratch for the n displacements, complex128 [n][dd][d] (16*n*dd*d
 * bytes); D_s is computed into it by a first launch and read by the draw. */
int bsmps_sample_site(bsmps_ctx *ctx, int64_t n, int chi, int d, int dd, const double *T,
                      int64_t T_sample_stride, const double *alpha, int64_t alpha_stride,
                      const double *eigvec, const double *eigval, double *dwork, const double *lambda,
                      const double *uniforms, int32_t *samples, int64_t samples_stride,
                      double *pre, double *probs);

/* Reduced-precision form of the two calls above (no reference counterpart; the reference's
 * only precision knob is complex64 throughout, MPS_cpu.py:17): the contraction runs on FP32
 * MFMA (v_mfma_f32_16x16x4_f32) with FP32 accumulation over complex64 pre / gamma / T, and the
 * draw reads complex64 T, computes D(alpha) and the probabilities in FP64, and writes the
 * carried vector as complex64 rescaled by its largest component (as the reference's GPU path
 * does, distributed_sampling.py:147), which leaves the normalised probabilities unchanged. */
int bsmps_site_contract_c64(bsmps_ctx *ctx, int64_t n, int chi_prev, int chi, int d, const float *pre,
                            const double *lambda_prev, const float *gamma, float *T);
int bsmps_sample_site_c64(bsmps_ctx *ctx, int64_t n, int chi, int d, int dd, const float *T,
                          int64_t T_sample_stride, const double *alpha, int64_t alpha_stride,
                          const double *eigvec, const double *eigval, double *dwork, const double *lambda,
                          const double *uniforms, int32_t *samples, int64_t samples_stride, float *pre,
                          double *probs);

/* BF16x3 split-precision form of the contraction (no reference counterpart): each FP32
 * operand x is split into bf16 hi = bf16(x) and lo = bf16(x - hi); products are
 * hi*hi + hi*lo + lo*hi on bf16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation.
 * gamma_hi / gamma_lo come from bsmps_split_bf16 applied to gamma's complex64 storage
 * (2*chi_prev*chi*d floats); pre and T are complex64 as in bsmps_site_contract_c64. */
int bsmps_split_bf16(bsmps_ctx *ctx, const float *x, int64_t count, void *hi, void *lo);
int bsmps_site_contract_bf16x3(bsmps_ctx *ctx, int64_t n, int chi_prev, int chi, int d, const float *pre,
                               const double *lambda_prev, const void *gamma_hi, const void *gamma_lo, float *T);

/* BF16x6 form of the contraction, FP32-accurate on bf16 MFMA (no reference counterpart):
 * each FP32 value is split exactly into three bf16 parts h + m + l and a product keeps the
 * six terms hh + hm + mh + hl + mm + lh (dropped terms ~2^-24 relative), FP32 accumulation.
 * Both operands are passed split into six bf16 parts (h, m, l of the real part, then of the
 * imaginary part), k padded with zeros to kpad = round_up(chi_prev, 32), laid out per row in
 * 32-k steps: part p of element k of row r at [r][k / 32][p][k % 32] (one step of all parts of
 * a row is 384 contiguous bytes):
 *   bsmps_split_bf16x6: lambda_prev[k] * gamma[k][c] (gamma complex64 [chi_prev][chi*d]), rows
 *     c < round_up(chi*d, 64), once per site; bsmps_bf16x6_plane_bytes gives their size;
 *   bsmps_split_rows_bf16x6: the carried vectors pre (complex64 [n][chi_prev]), rows r < n,
 *     once per step; bsmps_bf16x6_row_plane_bytes gives their size.
 * The size functions return -1 for invalid extents.  T is complex64 [n][chi*d] as in
 * bsmps_site_contract_c64. */
int64_t bsmps_bf16x6_plane_bytes(int chi_prev, int chi, int d);
int64_t bsmps_bf16x6_row_plane_bytes(int64_t n, int chi_prev);
int bsmps_split_bf16x6(bsmps_ctx *ctx, int chi_prev, int chi, int d, const float *gamma,
                       const double *lambda_prev, void *planes);
int bsmps_split_rows_bf16x6(bsmps_ctx *ctx, int64_t n, int chi_prev, const float *pre, void *planes);
int bsmps_site_contract_bf16x6(bsmps_ctx *ctx, int64_t n, int chi_prev, int chi, int d, const void *pre_planes,
                               const void *gamma_planes, float *T);

/* FP16x3 form of the contraction (no reference counterpart): values are scaled to a largest
 * component of 1 -- per column of lambda_prev * gamma, per row of pre -- and split into two fp16
 * parts h = fp16(x), l = fp16(x - h) (22 significand bits); a product keeps hh + hl + lh on fp16
 * MFMA (v_mfma_f32_16x16x32_f16) with FP32 accumulation, and the epilogue multiplies the scales
 * back (relative error ~2^-21 per product, three MFMAs per real product instead of six).
 * Layout as in the BF16x6 form with four parts (h, l of the real part, then of the imaginary
 * part): [round_up(chi*d, 64)][kpad / 32][4][32] (once per site, with colscale float
 * [round_up(chi*d, 64)]) and [n][kpad / 32][4][32] (once per step, with rowscale float [n]). */
int64_t bsmps_fp16x3_plane_bytes(int chi_prev, int chi, int d);
int64_t bsmps_fp16x3_row_plane_bytes(int64_t n, int chi_prev);
int bsmps_split_fp16x3(bsmps_ctx *ctx, int chi_prev, int chi, int d, const float *gamma,
                       const double *lambda_prev, void *planes, float *colscale);
int bsmps_split_rows_fp16x3(bsmps_ctx *ctx, int64_t n, int chi_prev, const float *pre, void *planes,
                            float *rowscale);
int bsmps_site_contract_fp16x3(bsmps_ctx *ctx, int64_t n, int chi_prev, int chi, int d, const void *pre_planes,
                               const float *rowscale, const void *gamma_planes, const float *colscale, float *T);

/* The four contractions above with the output stored through a column map (no reference
 * counterpart; the photon-number parity split of the contraction of sampling_cpu.py:75-80):
 * output column c < chi*d of sample row s goes to T[s * ldT + cmap[c]] (complex elements, rows of
 * ldT), so the product of one parity block of a site -- its labels' carried vectors against its
 * columns -- lands in the site's full T without a separate scatter.  cmap is device int32
 * [chi*d]; entries outside [0, ldT) are not written, and T's other columns are left as they are.
 * Everything else (operands, forms, segments, results bit for bit) is as in the plain call. */
int bsmps_site_contract_cols(bsmps_ctx *ctx, int64_t n, int chi_prev, int chi, int d, const double *pre,
                             const double *lambda_prev, const double *gamma, double *T, int64_t ldT,
                             const int32_t *cmap);
int bsmps_site_contract_c64_cols(bsmps_ctx *ctx, int64_t n, int chi_prev, int chi, int d, const float *pre,
                                 const double *lambda_prev, const float *gamma, float *T, int64_t ldT,
                                 const int32_t *cmap);
int bsmps_site_contract_bf16x6_cols(bsmps_ctx *ctx, int64_t n, int chi_prev, int chi, int d,
                                    const void *pre_planes, const void *gamma_planes, float *T, int64_t ldT,
                                    const int32_t *cmap);
int bsmps_site_contract_fp16x3_cols(bsmps_ctx *ctx, int64_t n, int chi_prev, int chi, int d,
                                    const void *pre_planes, const float *rowscale, const void *gamma_planes,
                                    const float *colscale, float *T, int64_t ldT, const int32_t *cmap);

/* Counter-based Philox4x32-10 streams (Salmon et al., SC'11), identical on host and device.
 * out[r][i] (r < n_streams, i < n) is element first+i of stream stream0+r under `seed`:
 * counter (e_lo, e_hi, stream, 0) with e = first+i, key (seed_lo, seed_hi).  A uniform is
 * (x >> 11) * 2^-53 of the first two output words (x = w1 << 32 | w0); a normal (normal != 0)
 * is sqrt(-2 ln(1 - u1)) cos(2 pi u2) with u2 from words 2, 3 (Box-Muller, cosine branch).
 * Replaces np.random.normal / np.random.rand, sampling_cpu.py:37, 52, 92.
 * out device double [n_streams][n]. */
int bsmps_philox(bsmps_ctx *ctx, uint64_t seed, uint32_t stream0, int n_streams, int64_t first, int64_t n,
                 int normal, double *out);

/* ---- learned-cutoff form: a small MLP trained and evaluated on the GPU (no reference
 * counterpart; the reference has no learned component) ------------------------------------ */

/* Forward pass of the in -> width -> width -> out ReLU network:
 *   H1 = relu(X W1^T + b1), H2 = relu(H1 W2^T + b2), Y = H2 W3^T + b3.
 * X float [n][d_in]; W1 bf16 [width][d_in], W2 bf16 [width][width], W3 bf16 [d_out][width];
 * biases float; H1, H2 float [n][width]; Y float [n][d_out].  Products on BF16 MFMA with FP32
 * accumulation; FP32 activations are rounded to BF16 (nearest-even) as operands. */
int bsmps_mlp_forward(bsmps_ctx *ctx, int n, int d_in, int width, int d_out, const float *X, const void *W1,
                      const float *b1, const void *W2, const float *b2, const void *W3, const float *b3,
                      float *H1, float *H2, float *Y);
/* Backward pass for the loss gradient dY float [n][d_out] (H1, H2 from the forward pass):
 * weight gradients float [fan_out][fan_in] and bias gradients float, summed over the n rows.
 * scratch: float [2 n width]. */
int bsmps_mlp_backward(bsmps_ctx *ctx, int n, int d_in, int width, int d_out, const float *X, const void *W2,
                       const void *W3, const float *H1, const float *H2, const float *dY, float *scratch,
                       float *dW1, float *db1, float *dW2, float *db2, float *dW3, float *db3);
/* dst[i] = bf16(src[i]), round to nearest even. */
int bsmps_f32_to_bf16(bsmps_ctx *ctx, int64_t count, const float *src, void *dst);
/* One Adam step (Kingma & Ba, 2015) with bias correction for step >= 1:
 * m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2; p -= lr (m / (1-b1^step)) / (sqrt(v / (1-b2^step)) + eps). */
int bsmps_adam_step(bsmps_ctx *ctx, int64_t count, float *param, const float *grad, float *m, float *v, float lr,
                    float beta1, float beta2, float eps, int step);
/* Mean squared error gradient over n x d_out: dY = 2 (Y - T) / (n d_out); err (optional) = (Y - T)^2. */
int bsmps_mse_grad(bsmps_ctx *ctx, int n, int d_out, const float *Y, const float *T, float *dY, float *err);

#ifdef __cplusplus
}
#endif

#endif /* BSMPS_H */
