// Context, workspace and error reporting for the bsmps C ABI.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "common.h"

static thread_local char g_last_error[512] = "";

void bsmps_set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

int bsmps_check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return BSMPS_OK;
    bsmps_set_error("%s failed: %s", what, hipGetErrorString(e));
    return BSMPS_ERR_HIP;
}

int bsmps_ctx_workspace(bsmps_ctx* ctx, size_t bytes, void** out) {
    if (bytes > ctx->ws_bytes) {
        if (ctx->ws) {
            BSMPS_HIP(hipStreamSynchronize(ctx->stream));
            BSMPS_HIP(hipFree(ctx->ws));
            ctx->ws = nullptr;
            ctx->ws_bytes = 0;
        }
        const size_t want = bytes + bytes / 4;
        BSMPS_HIP(hipMalloc(&ctx->ws, want));
        ctx->ws_bytes = want;
    }
    *out = ctx->ws;
    return BSMPS_OK;
}

int bsmps_ctx_scratch(bsmps_ctx* ctx, size_t bytes, void** out) {
    if (bytes > ctx->hws_bytes) {
        if (ctx->hws) {
            BSMPS_HIP(hipStreamSynchronize(ctx->stream));
            BSMPS_HIP(hipFree(ctx->hws));
            ctx->hws = nullptr;
            ctx->hws_bytes = 0;
        }
        BSMPS_HIP(hipMalloc(&ctx->hws, bytes));
        ctx->hws_bytes = bytes;
    }
    *out = ctx->hws;
    return BSMPS_OK;
}

int bsmps_ctx_upload(bsmps_ctx* ctx, const void* src, size_t bytes, void* dst) {
    if (bytes == 0) return BSMPS_OK;
    const int i = ctx->pin_next;
    ctx->pin_next = (ctx->pin_next + 1) % bsmps_ctx::kPinSlots;
    if (ctx->pin_ev[i]) BSMPS_HIP(hipEventSynchronize(ctx->pin_ev[i]));  // the copy kPinSlots uploads ago
    else BSMPS_HIP(hipEventCreateWithFlags(&ctx->pin_ev[i], hipEventDisableTiming));
    if (bytes > ctx->pin_bytes[i]) {
        if (ctx->pin[i]) BSMPS_HIP(hipHostFree(ctx->pin[i]));
        ctx->pin[i] = nullptr;
        ctx->pin_bytes[i] = 0;
        const size_t want = bytes + bytes / 4 + 4096;
        BSMPS_HIP(hipHostMalloc(&ctx->pin[i], want, hipHostMallocDefault));
        ctx->pin_bytes[i] = want;
    }
    memcpy(ctx->pin[i], src, bytes);
    BSMPS_HIP(hipMemcpyAsync(dst, ctx->pin[i], bytes, hipMemcpyHostToDevice, ctx->stream));
    BSMPS_HIP(hipEventRecord(ctx->pin_ev[i], ctx->stream));
    return BSMPS_OK;
}

extern "C" const char* bsmps_last_error(void) { return g_last_error; }

extern "C" int bsmps_version(void) { return 1; }

extern "C" int bsmps_create(int device, void* stream, bsmps_ctx** out) {
    if (!out) {
        bsmps_set_error("bsmps_create: null output pointer");
        return BSMPS_ERR_INVALID;
    }
    *out = nullptr;
    int n = 0;
    BSMPS_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        bsmps_set_error("bsmps_create: device %d out of range (%d devices)", device, n);
        return BSMPS_ERR_INVALID;
    }
    BSMPS_HIP(hipSetDevice(device));
    BSMPS_TRY(bsmps_kron_noalloc_init(device));  // the device's allocation-free scratch address
    bsmps_ctx* c = new bsmps_ctx();
    c->device = device;
    c->ws = nullptr;
    c->ws_bytes = 0;
    c->lane_max_m = 4;
    for (int m = 0; m <= BSMPS_MAX_HAFNIAN_SIZE / 2; ++m) c->group_lanes_by_m[m] = 64;
    for (int m = 0; m <= BSMPS_MAX_HAFNIAN_SIZE / 2; ++m) c->split_by_m[m] = m >= 6 ? 1 : 0;
    c->hws = nullptr;
    c->hws_bytes = 0;
    c->split_chunk_bytes = (size_t)16 << 30;
    c->contract_gauss[0] = 0;
    c->contract_gauss[1] = 0;
    for (int i = 0; i < bsmps_ctx::kPinSlots; ++i) {
        c->pin[i] = nullptr;
        c->pin_bytes[i] = 0;
        c->pin_ev[i] = nullptr;
    }
    c->pin_next = 0;
    c->max_lds = 64 * 1024;
    {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0)
            c->max_lds = v;
    }
    // NULL is HIP's null stream, which is also PyTorch's default stream: work enqueued here
    // is ordered with the caller's allocations, fills and copies on that stream.
    c->stream = reinterpret_cast<hipStream_t>(stream);
    c->own_stream = 0;
    *out = c;
    return BSMPS_OK;
}

extern "C" int bsmps_destroy(bsmps_ctx* ctx) {
    if (!ctx) return BSMPS_OK;
    int st = BSMPS_OK;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) st = BSMPS_ERR_HIP;
    if (ctx->ws) hipFree(ctx->ws);
    if (ctx->hws) hipFree(ctx->hws);
    for (int i = 0; i < bsmps_ctx::kPinSlots; ++i) {
        if (ctx->pin[i]) hipHostFree(ctx->pin[i]);
        if (ctx->pin_ev[i]) hipEventDestroy(ctx->pin_ev[i]);
    }
    if (ctx->own_stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return st;
}

extern "C" int bsmps_synchronize(bsmps_ctx* ctx) {
    if (!ctx) {
        bsmps_set_error("bsmps_synchronize: null context");
        return BSMPS_ERR_INVALID;
    }
    BSMPS_HIP(hipStreamSynchronize(ctx->stream));
    return BSMPS_OK;
}

extern "C" int bsmps_set_tuning(bsmps_ctx* ctx, int lane_max_m, int group_lanes) {
    if (!ctx || lane_max_m < 0 || lane_max_m > 5 ||
        (group_lanes != 16 && group_lanes != 32 && group_lanes != 64)) {
        bsmps_set_error("bsmps_set_tuning: lane_max_m must be 0..5 and group_lanes 16, 32 or 64");
        return BSMPS_ERR_INVALID;
    }
    ctx->lane_max_m = lane_max_m;
    for (int m = 0; m <= BSMPS_MAX_HAFNIAN_SIZE / 2; ++m) ctx->group_lanes_by_m[m] = group_lanes;
    return BSMPS_OK;
}

extern "C" int bsmps_set_group_lanes(bsmps_ctx* ctx, int n, int group_lanes) {
    if (!ctx || n < 0 || n > BSMPS_MAX_HAFNIAN_SIZE || n % 2 ||
        (group_lanes != 16 && group_lanes != 32 && group_lanes != 64)) {
        bsmps_set_error("bsmps_set_group_lanes: n must be even and <= %d, group_lanes 16, 32 or 64",
                        BSMPS_MAX_HAFNIAN_SIZE);
        return BSMPS_ERR_INVALID;
    }
    ctx->group_lanes_by_m[n / 2] = group_lanes;
    return BSMPS_OK;
}

extern "C" int bsmps_set_split(bsmps_ctx* ctx, int n, int enable) {
    if (!ctx || n < 12 || n > BSMPS_MAX_HAFNIAN_SIZE || n % 2 || (enable != 0 && enable != 1)) {
        bsmps_set_error("bsmps_set_split: n must be even in 12..%d and enable 0 or 1", BSMPS_MAX_HAFNIAN_SIZE);
        return BSMPS_ERR_INVALID;
    }
    ctx->split_by_m[n / 2] = enable;
    return BSMPS_OK;
}

extern "C" int bsmps_set_contract_form(bsmps_ctx* ctx, int precision, int gauss) {
    if (!ctx || precision < 0 || precision > 1 || (gauss != 0 && gauss != 1)) {
        bsmps_set_error("bsmps_set_contract_form: precision must be 0 (FP64) or 1 (FP32), gauss 0 or 1");
        return BSMPS_ERR_INVALID;
    }
    ctx->contract_gauss[precision] = gauss;
    return BSMPS_OK;
}

extern "C" int bsmps_set_split_budget(bsmps_ctx* ctx, size_t bytes) {
    if (!ctx || bytes < ((size_t)1 << 20)) {
        bsmps_set_error("bsmps_set_split_budget: the budget must be at least 1 MiB");
        return BSMPS_ERR_INVALID;
    }
    ctx->split_chunk_bytes = bytes;
    return BSMPS_OK;
}
