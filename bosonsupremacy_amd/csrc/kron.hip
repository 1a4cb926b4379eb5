// Schmidt spectra of the MPS cuts: the pruned Kronecker loop of get_cumsum_kron
// (kron_cpu.py:106-125), one workgroup per cut, all cuts in one launch.
//
// Step i of a cut multiplies the kept weights res[a] (sorted descending) by the thermal
// distribution tp[i][n] of mode i.  For a fixed n the candidates res[a] * tp[i][n] are
// already sorted (rounding is monotone), so the reference's threshold + argsort + top
// max_dim (kron_cpu.py:110-114) becomes: a binary search per n for the candidates above
// err_tol, then a d-way merge of d sorted lists truncated to max_dim, done as d-1
// merge-path passes inside the workgroup.  No sort, no materialised outer product.
// Labels are not copied per step (kron_cpu.py:118-119 copies K x i bytes per step): each
// step stores one back-pointer (parent rank, photon number) per kept weight, and only the
// chi survivors walk their back-pointers at the end.
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "common.h"

namespace {

constexpr int kThreads = 1024;
constexpr int kMaxD = 32;  // photon numbers fit in the low 5 bits of a back-pointer

struct CutDesc {
    int n_modes;
    int pad;
    int64_t bp_off;  // int32 offset of this cut's back-pointer block in the workspace
};

// Position in an implicit list Y (values cur[a] * f) -- the same single multiply as
// np.outer(res, tp) in kron_cpu.py:110.
__device__ __forceinline__ double yval(const double* cur, double f, int64_t a) { return cur[a] * f; }

// Cut c of a launch (workgroup c), described by cd.
__device__ __forceinline__ void kron_cut(const CutDesc cd, int c, int max_modes, int d,
                                         const double* __restrict__ tp_all, int64_t max_dim, int chi,
                                         double err_tol, double* __restrict__ res_out, int8_t* __restrict__ num_out,
                                         int32_t* __restrict__ len_out, double* __restrict__ vbuf,
                                         int32_t* __restrict__ pbuf, int32_t* __restrict__ bp_all) {
    const int n_modes = cd.n_modes;
    const double* tp = tp_all + (int64_t)c * max_modes * d;
    double* V[3] = {vbuf + (int64_t)c * 3 * max_dim, vbuf + (int64_t)c * 3 * max_dim + max_dim,
                    vbuf + (int64_t)c * 3 * max_dim + 2 * max_dim};
    int32_t* P[2] = {pbuf + (int64_t)c * 2 * max_dim, pbuf + (int64_t)c * 2 * max_dim + max_dim};
    int32_t* bp = bp_all + cd.bp_off;  // step s (1..n-1) at bp[(s-1)*max_dim]; step 0 perm in perm0

    __shared__ int perm0[kMaxD];
    __shared__ int64_t cnt[kMaxD];
    __shared__ int64_t s_len;
    const int tid = threadIdx.x;

    // step 0: the first mode's distribution, ordered descending (ties: lower n first)
    if (tid == 0) {
        for (int n = 0; n < d; ++n) perm0[n] = n;
        for (int i = 1; i < d; ++i) {  // insertion sort, d <= 32
            int x = perm0[i];
            int j = i - 1;
            while (j >= 0 && tp[perm0[j]] < tp[x]) {
                perm0[j + 1] = perm0[j];
                --j;
            }
            perm0[j + 1] = x;
        }
        s_len = d < max_dim ? d : max_dim;
    }
    __syncthreads();
    int cur = 0;
    for (int r = tid; r < s_len; r += kThreads) V[0][r] = tp[perm0[r]];
    __syncthreads();

    for (int i = 1; i < n_modes; ++i) {
        const double* tpi = tp + (int64_t)i * d;
        const int64_t K = s_len;
        const double* X0 = V[cur];
        // candidates above err_tol: a prefix of each list n
        if (tid < d) {
            const double f = tpi[tid];
            int64_t lo = 0, hi = K;
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (yval(X0, f, mid) > err_tol) lo = mid + 1;
                else hi = mid;
            }
            cnt[tid] = lo;
        }
        __syncthreads();
        // list 0 materialised, then merged with lists 1..d-1
        int vo = (cur + 1) % 3, vo2 = (cur + 2) % 3;
        int po = 0;
        int64_t nx = cnt[0] < max_dim ? cnt[0] : max_dim;
        {
            const double f = tpi[0];
            int32_t* Pd = (d == 1) ? bp + (int64_t)(i - 1) * max_dim : P[po];
            for (int64_t r = tid; r < nx; r += kThreads) {
                V[vo][r] = yval(X0, f, r);
                Pd[r] = (int32_t)(r << 5);
            }
        }
        __syncthreads();
        for (int n = 1; n < d; ++n) {
            const double f = tpi[n];
            const int64_t ny = cnt[n];
            const int64_t no = (nx + ny) < max_dim ? (nx + ny) : max_dim;
            const double* Xv = V[vo];
            const int32_t* Xp = P[po];
            double* Ov = V[vo2];
            int32_t* Op = (n == d - 1) ? bp + (int64_t)(i - 1) * max_dim : P[1 - po];
            const int64_t chunk = (no + kThreads - 1) / kThreads;
            const int64_t p0 = (int64_t)tid * chunk;
            if (p0 < no) {
                const int64_t p1 = p0 + chunk < no ? p0 + chunk : no;
                // merge-path split: ix elements of X precede output position p0
                int64_t lo = p0 - ny > 0 ? p0 - ny : 0, hi = p0 < nx ? p0 : nx;
                while (lo < hi) {
                    const int64_t mid = (lo + hi) >> 1;
                    if (Xv[mid] >= yval(X0, f, p0 - 1 - mid)) lo = mid + 1;
                    else hi = mid;
                }
                int64_t ix = lo, iy = p0 - lo;
                for (int64_t p = p0; p < p1; ++p) {
                    const bool takeX = iy >= ny || (ix < nx && Xv[ix] >= yval(X0, f, iy));
                    if (takeX) {
                        Ov[p] = Xv[ix];
                        Op[p] = Xp[ix];
                        ++ix;
                    } else {
                        Ov[p] = yval(X0, f, iy);
                        Op[p] = (int32_t)((iy << 5) | n);
                        ++iy;
                    }
                }
            }
            __syncthreads();
            nx = no;
            const int t = vo;
            vo = vo2;
            vo2 = t;
            po = 1 - po;
        }
        // the merged list is in V[vo]; it becomes the current list
        cur = vo;
        if (tid == 0) s_len = nx;
        __syncthreads();
    }

    // output: the chi largest, descending, with labels from the back-pointers
    const int64_t K = s_len;
    const int len = (int)(K < chi ? K : chi);
    if (tid == 0) len_out[c] = len;
    double* res = res_out + (int64_t)c * chi;
    int8_t* num = num_out + (int64_t)c * chi * max_modes;
    for (int r = tid; r < chi; r += kThreads) {
        int8_t* lab = num + (int64_t)r * max_modes;
        if (r >= len) {
            res[r] = 0.0;
            for (int k = 0; k < max_modes; ++k) lab[k] = 0;
            continue;
        }
        res[r] = V[cur][r];
        int64_t idx = r;
        for (int s = n_modes - 1; s >= 1; --s) {
            const int32_t pay = bp[(int64_t)(s - 1) * max_dim + idx];
            lab[s] = (int8_t)(pay & 31);
            idx = pay >> 5;
        }
        lab[0] = (int8_t)perm0[idx];
        for (int k = n_modes; k < max_modes; ++k) lab[k] = 0;
    }
}

// Cut descriptors read from device memory (the allocating entry point uploads them).
__global__ __launch_bounds__(kThreads) void kron_cut_kernel(const CutDesc* __restrict__ cuts, int max_modes, int d,
                                                            const double* __restrict__ tp_all, int64_t max_dim,
                                                            int chi, double err_tol, double* __restrict__ res_out,
                                                            int8_t* __restrict__ num_out, int32_t* __restrict__ len_out,
                                                            double* __restrict__ vbuf, int32_t* __restrict__ pbuf,
                                                            int32_t* __restrict__ bp_all) {
    kron_cut(cuts[blockIdx.x], blockIdx.x, max_modes, d, tp_all, max_dim, chi, err_tol, res_out, num_out, len_out,
             vbuf, pbuf, bp_all);
}

// Cut descriptors passed by value in the kernel arguments (the allocation-free entry point:
// no upload, no host staging buffer, nothing to synchronise on).
constexpr int kArgCuts = 64;  // 64 x 16 bytes of kernel arguments
struct CutDescArgs {
    CutDesc c[kArgCuts];
};
__global__ __launch_bounds__(kThreads) void kron_cut_args_kernel(CutDescArgs cuts, int max_modes, int d,
                                                                 const double* __restrict__ tp_all, int64_t max_dim,
                                                                 int chi, double err_tol, double* __restrict__ res_out,
                                                                 int8_t* __restrict__ num_out,
                                                                 int32_t* __restrict__ len_out,
                                                                 double* __restrict__ vbuf, int32_t* __restrict__ pbuf,
                                                                 int32_t* __restrict__ bp_all) {
    kron_cut(cuts.c[blockIdx.x], blockIdx.x, max_modes, d, tp_all, max_dim, chi, err_tol, res_out, num_out, len_out,
             vbuf, pbuf, bp_all);
}

// Scratch of the allocation-free entry point: a file-scope device array (one copy per GPU,
// held until the process ends).  Host code never passes its name to a kernel -- on the host
// that name is a stand-in, not a device address -- but the per-device address that
// hipGetSymbolAddress returns, looked up once per device when the device's first context is
// created (bsmps_kron_noalloc_init, called by bsmps_create).
constexpr size_t kNoAllocScratchBytes = (size_t)256 << 20;
__device__ __align__(256) char kron_noalloc_scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Bytes of scratch cut c0..c1-1 take: value and payload buffers plus the back-pointer block.
int64_t cut_fixed_bytes(int64_t max_dim) { return max_dim * (3 * 8 + 2 * 4); }
int64_t cut_bp_words(int nm, int64_t max_dim) { return (int64_t)std::max(nm - 1, 1) * max_dim; }

}  // namespace

int bsmps_kron_noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) {
        bsmps_set_error("bsmps_create: device %d beyond the %d the allocation-free scratch table holds", device,
                        kMaxDevices);
        return BSMPS_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return BSMPS_OK;
    void* p = nullptr;
    BSMPS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(kron_noalloc_scratch)));  // the current device's copy
    g_scratch_ptr[device] = p;
    return BSMPS_OK;
}

extern "C" int64_t bsmps_kron_noalloc_scratch_bytes(void) { return (int64_t)kNoAllocScratchBytes; }

extern "C" int bsmps_kron_cuts_noalloc(bsmps_ctx* ctx, int n_cuts, const int32_t* n_modes_host, int max_modes,
                                       int d, const double* tp, int64_t max_dim, int chi, double err_tol,
                                       double* res, int8_t* num, int32_t* len) {
    if (!ctx || n_cuts < 0 || !n_modes_host || max_modes <= 0 || d <= 0 || d > kMaxD || max_dim <= 0 ||
        max_dim > (int64_t(1) << 26) || chi <= 0) {
        bsmps_set_error("bsmps_kron_cuts_noalloc: invalid arguments (d must be 1..%d, max_dim 1..2^26)", kMaxD);
        return BSMPS_ERR_INVALID;
    }
    if (n_cuts == 0) return BSMPS_OK;
    if (!tp || !res || !num || !len) {
        bsmps_set_error("bsmps_kron_cuts_noalloc: null pointer");
        return BSMPS_ERR_INVALID;
    }
    char* base = ctx->device >= 0 && ctx->device < kMaxDevices ? static_cast<char*>(g_scratch_ptr[ctx->device])
                                                               : nullptr;
    if (!base) {
        bsmps_set_error("bsmps_kron_cuts_noalloc: no scratch address for device %d", ctx->device);
        return BSMPS_ERR_UNSUPPORTED;
    }
    const int64_t cap = (int64_t)kNoAllocScratchBytes;
    // every cut must fit the scratch on its own; checked before anything is launched
    for (int c = 0; c < n_cuts; ++c) {
        const int nm = n_modes_host[c];
        if (nm < 1 || nm > max_modes) {
            bsmps_set_error("bsmps_kron_cuts_noalloc: cut %d has %d modes (max_modes %d)", c, nm, max_modes);
            return BSMPS_ERR_INVALID;
        }
        const int64_t need = cut_fixed_bytes(max_dim) + cut_bp_words(nm, max_dim) * 4 + 1024;
        if (need > cap) {
            bsmps_set_error("bsmps_kron_cuts_noalloc: cut %d (%d modes, max_dim %lld) needs %lld bytes of scratch, "
                            "more than the %lld this entry point has", c, nm, (long long)max_dim, (long long)need,
                            (long long)cap);
            return BSMPS_ERR_UNSUPPORTED;
        }
    }
    // chunks of at most kArgCuts cuts whose buffers fit the scratch, one launch each, in order
    // on the context's stream (the next chunk reuses the scratch after the previous has run)
    int c0 = 0;
    while (c0 < n_cuts) {
        CutDescArgs args;
        int64_t bp_words = 0;
        int c1 = c0;
        while (c1 < n_cuts && c1 - c0 < kArgCuts) {
            const int64_t words = cut_bp_words(n_modes_host[c1], max_dim);
            const int64_t nc = c1 - c0 + 1;
            const int64_t bytes = ((nc * 3 * max_dim * 8 + 255) / 256) * 256 + ((nc * 2 * max_dim * 4 + 255) / 256) * 256 +
                                  (bp_words + words) * 4;
            if (c1 > c0 && bytes > cap) break;
            args.c[c1 - c0] = CutDesc{n_modes_host[c1], 0, bp_words};
            bp_words += words;
            ++c1;
        }
        const int nc = c1 - c0;
        const size_t bytes_v = ((size_t)nc * 3 * max_dim * sizeof(double) + 255) / 256 * 256;
        const size_t bytes_p = ((size_t)nc * 2 * max_dim * sizeof(int32_t) + 255) / 256 * 256;
        hipLaunchKernelGGL(kron_cut_args_kernel, dim3(nc), dim3(kThreads), 0, ctx->stream, args, max_modes, d,
                           tp + (int64_t)c0 * max_modes * d, max_dim, chi, err_tol, res + (int64_t)c0 * chi,
                           num + (int64_t)c0 * chi * max_modes, len + c0, reinterpret_cast<double*>(base),
                           reinterpret_cast<int32_t*>(base + bytes_v), reinterpret_cast<int32_t*>(base + bytes_v + bytes_p));
        BSMPS_HIP(hipGetLastError());
        c0 = c1;
    }
    return BSMPS_OK;
}

extern "C" int bsmps_kron_cuts(bsmps_ctx* ctx, int n_cuts, const int32_t* n_modes_host, int max_modes, int d,
                               const double* tp, int64_t max_dim, int chi, double err_tol, double* res, int8_t* num,
                               int32_t* len) {
    if (!ctx || n_cuts < 0 || !n_modes_host || max_modes <= 0 || d <= 0 || d > kMaxD || max_dim <= 0 ||
        max_dim > (int64_t(1) << 26) || chi <= 0) {
        bsmps_set_error("bsmps_kron_cuts: invalid arguments (d must be 1..%d, max_dim 1..2^26)", kMaxD);
        return BSMPS_ERR_INVALID;
    }
    if (n_cuts == 0) return BSMPS_OK;
    if (!tp || !res || !num || !len) {
        bsmps_set_error("bsmps_kron_cuts: null pointer");
        return BSMPS_ERR_INVALID;
    }
    // Cuts are processed in chunks that keep the back-pointer workspace bounded.
    const int64_t ws_limit = int64_t(2) << 30;
    int c0 = 0;
    while (c0 < n_cuts) {
        std::vector<CutDesc> desc;
        int64_t bp_words = 0;
        int c1 = c0;
        while (c1 < n_cuts) {
            const int nm = n_modes_host[c1];
            if (nm < 1 || nm > max_modes) {
                bsmps_set_error("bsmps_kron_cuts: cut %d has %d modes (max_modes %d)", c1, nm, max_modes);
                return BSMPS_ERR_INVALID;
            }
            const int64_t words = (int64_t)std::max(nm - 1, 1) * max_dim;
            const int64_t fixed = (int64_t)(c1 - c0 + 1) * max_dim * (3 * 8 + 2 * 4);
            if (c1 > c0 && (bp_words + words) * 4 + fixed > ws_limit) break;
            desc.push_back(CutDesc{nm, 0, bp_words});
            bp_words += words;
            ++c1;
        }
        const int nc = c1 - c0;
        const size_t bytes_desc = ((sizeof(CutDesc) * nc + 255) / 256) * 256;
        const size_t bytes_v = ((size_t)nc * 3 * max_dim * sizeof(double) + 255) / 256 * 256;
        const size_t bytes_p = ((size_t)nc * 2 * max_dim * sizeof(int32_t) + 255) / 256 * 256;
        const size_t bytes_bp = (size_t)bp_words * sizeof(int32_t);
        void* ws = nullptr;
        BSMPS_TRY(bsmps_ctx_workspace(ctx, bytes_desc + bytes_v + bytes_p + bytes_bp, &ws));
        char* base = reinterpret_cast<char*>(ws);
        BSMPS_HIP(hipMemcpyAsync(base, desc.data(), sizeof(CutDesc) * nc, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(kron_cut_kernel, dim3(nc), dim3(kThreads), 0, ctx->stream,
                           reinterpret_cast<const CutDesc*>(base), max_modes, d, tp + (int64_t)c0 * max_modes * d,
                           max_dim, chi, err_tol, res + (int64_t)c0 * chi, num + (int64_t)c0 * chi * max_modes,
                           len + c0, reinterpret_cast<double*>(base + bytes_desc),
                           reinterpret_cast<int32_t*>(base + bytes_desc + bytes_v),
                           reinterpret_cast<int32_t*>(base + bytes_desc + bytes_v + bytes_p));
        BSMPS_HIP(hipGetLastError());
        BSMPS_HIP(hipStreamSynchronize(ctx->stream));  // `desc` is a host staging buffer
        c0 = c1;
    }
    return BSMPS_OK;
}
