// Shared device/host helpers for the bsmps HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bsmps.h"

// complex128 as two doubles (interleaved re, im in memory, like NumPy's complex128).
struct cd {
    double re, im;
};

// NumPy's complex128 multiply on x86-64 is evaluated in this FMA form; the hafnian and
// the Gamma epilogue reproduce it bit for bit.  This file is compiled with
// -ffp-contract=off, so nothing else is fused.
__device__ __forceinline__ cd cmul(cd a, cd b) {
    cd r;
    r.re = __builtin_fma(a.re, b.re, -(a.im * b.im));
    r.im = __builtin_fma(a.re, b.im, a.im * b.re);
    return r;
}
__device__ __forceinline__ cd cadd(cd a, cd b) {
    cd r;
    r.re = a.re + b.re;
    r.im = a.im + b.im;
    return r;
}
// complex / real as NumPy evaluates it: multiply both parts by the rounded reciprocal.
__device__ __forceinline__ cd cscale(cd a, double s) {
    cd r;
    r.re = a.re * s;
    r.im = a.im * s;
    return r;
}
__device__ __forceinline__ cd czero() { return cd{0.0, 0.0}; }

__device__ __forceinline__ cd ld_cd(const double* p, long i) {
    const double2 v = reinterpret_cast<const double2*>(p)[i];
    return cd{v.x, v.y};
}
__device__ __forceinline__ void st_cd(double* p, long i, cd v) {
    reinterpret_cast<double2*>(p)[i] = make_double2(v.re, v.im);
}
// Streaming forms (non-temporal cache policy) for data a kernel reads or writes once, so that
// it does not displace the data other work items of the same launch read again from the L2.
typedef double dbl2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ cd ld_cd_nt(const double* p, long i) {
    const dbl2_t v = __builtin_nontemporal_load(reinterpret_cast<const dbl2_t*>(p) + i);
    return cd{v.x, v.y};
}
__device__ __forceinline__ void st_cd_nt(double* p, long i, cd v) {
    __builtin_nontemporal_store(dbl2_t{v.re, v.im}, reinterpret_cast<dbl2_t*>(p) + i);
}

// Host-side error plumbing shared by every translation unit.
void bsmps_set_error(const char* fmt, ...);
int bsmps_check_hip(hipError_t e, const char* what);

struct bsmps_ctx {
    int device;
    hipStream_t stream;
    int own_stream;
    // device workspace, grown on demand (never on the per-item hot path)
    void* ws;
    size_t ws_bytes;
    // autotuned launch parameters
    int lane_max_m;   // hafnians of size <= 2*lane_max_m take the one-lane-per-hafnian path
    int group_lanes_by_m[BSMPS_MAX_HAFNIAN_SIZE / 2 + 1];  // lanes per hafnian in the cooperative path (16, 32 or 64), by n/2
    int max_lds;      // bytes of LDS one workgroup may use on this device
    int split_by_m[BSMPS_MAX_HAFNIAN_SIZE / 2 + 1];  // 1: size 2m takes the frontier-split path (n = 12..34)
    // device scratch of the frontier-split hafnian path (separate from `ws`, which holds the
    // per-call tables that are live while the hafnian launches run)
    void* hws;
    size_t hws_bytes;
    size_t split_chunk_bytes;  // upper bound of `hws` per chunk of hafnians
    int contract_gauss[2];     // site contraction form per precision (0 FP64, 1 FP32): 1 = three real products
    // a ring of pinned host staging buffers for small per-call tables; an event marks when the
    // copy out of each has completed, and a buffer is reused only after that (so the host can
    // run up to kPinSlots uploads ahead of the stream)
    static constexpr int kPinSlots = 16;
    void* pin[kPinSlots];
    size_t pin_bytes[kPinSlots];
    hipEvent_t pin_ev[kPinSlots];
    int pin_next;
};

int bsmps_ctx_workspace(bsmps_ctx* ctx, size_t bytes, void** out);
// Looks up (once per device, with the device current) the address of the allocation-free
// entry point's scratch array (kron.hip).
int bsmps_kron_noalloc_init(int device);
int bsmps_ctx_scratch(bsmps_ctx* ctx, size_t bytes, void** out);
// Asynchronous host -> device copy of a small table through pinned staging: returns once the
// table is staged, without waiting for the stream.
int bsmps_ctx_upload(bsmps_ctx* ctx, const void* src, size_t bytes, void* dst);

#define BSMPS_TRY(expr)                      \
    do {                                     \
        int _st = (expr);                    \
        if (_st != BSMPS_OK) return _st;     \
    } while (0)
#define BSMPS_HIP(expr) BSMPS_TRY(bsmps_check_hip((expr), #expr))
