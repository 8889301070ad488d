// Internal to libcwfield: the field handle and the helpers that the solve
// (cwfield.hip) and the coupling/holds/step side (cw_coupling.hip) share.
#pragma once

#include "../../include/cwfield.h"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

namespace cwi {

// Records the message for cw_last_error (per thread) and returns code.
cw_status fail(cw_status code, const char* fmt, ...);
// One-time device check (a gfx950 HIP device is present).
cw_status device_init();
// hipMalloc plus an asynchronous byte fill on stream s.
cw_status alloc_fill(void** p, size_t bytes, hipStream_t s, int fill);

constexpr uint32_t kObstacle = 0xFFFFFFFFu;  // cpm.rs:137 OBSTACLE = u32::MAX

}  // namespace cwi

#define CW_HIP(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return cwi::fail(CW_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr,              \
                             hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)

// Grid-stride reductions run at most kRedBlocks workgroups; each writes its
// partial result to its own slot and one workgroup folds the slots.
constexpr int kScanInts = 64;
constexpr int kHostSmallInts = 64;  // cw_field::host_small
constexpr int kRedBlocks = 2048;
constexpr int kRedSlot = 8;

// Device mirror of ScalarField (crates/engine/src/pde.rs:30-53). Every
// per-voxel array is pitched: row pitch P = ceil(W/64)*64 elements, plane
// stride P*H, x fastest. The device copies of held/uptake/production are the
// field's state; the solve's coefficient words are derived from them.
struct cw_field {
    uint32_t W = 0, H = 0, D = 0;
    int device = 0;            // HIP device the field's memory lives on
    int P = 0;
    size_t n = 0, npad = 0;
    double diff = 0, decay = 0, source = 0;
    hipStream_t stream = nullptr;
    double* c[2] = {nullptr, nullptr};  // values / next ping-pong
    int cur = 0;
    uint8_t* held = nullptr;   // 1 = held; padding 1
    double* up = nullptr;      // uptake (clamped at zero), padding 0
    double* pr = nullptr;      // production (clamped at zero), padding 0
    double* k = nullptr;       // folded coefficient, NaN for held and padding
    double* k2 = nullptr;      // uptake, general form only
    bool general = false;
    float* e[2] = {nullptr, nullptr};
    float* r32 = nullptr;
    float* lb = nullptr;
    int* scan = nullptr;       // kScanInts of small results, then kRedBlocks slots of kRedSlot
                               // ints: per-workgroup partials of the reductions
    mutable void* stage = nullptr;  // flat device staging of host copies
    mutable size_t stage_bytes = 0;
    int* host_small = nullptr;  // pinned host ints for small read-backs (spans, boxes, flags)
    uint8_t* mask8[2] = {nullptr, nullptr};  // u8 scratch of npad (the holds' masks,
                                             // the flood's open set and reach), kept
                                             // so a hold allocates once per field
    bool has_span = false;
    int span[6] = {0, 0, 0, 0, 0, 0};
    bool has_sink = false;
    bool coef_dirty = true;
    // Launch plans of this field's single-field solves, kept between calls so
    // a warm-started relax allocates nothing (cwfield.hip owns the type).
    void* plans = nullptr;
};

namespace cwi {

// Derive k (and k2), the sink flag and the free-span box from the device
// state (cw_coupling.hip). Runs before a solve whenever the state changed.
// Without may_alloc it fails (CW_ERR_ARG) instead of allocating the
// general-form coefficient buffer a field needs the first time it keeps
// production and uptake apart.
cw_status build_coefficients(cw_field* f, bool may_alloc = true);
// Flat (W*H*D) host copies of the device state.
cw_status download_held(const cw_field* f, uint8_t* out);
cw_status download_f64(const cw_field* f, const double* dev, double* out);
// Host flat (W*H*D, x fastest) <-> device pitched copies of elem-byte
// elements (1, 4 or 8): one contiguous PCIe copy through a flat device
// staging buffer and a pack/unpack kernel, instead of one copy per row.
cw_status upload_flat(const cw_field* f, void* dst, const void* src, int elem);
cw_status download_flat(const cw_field* f, const void* src, void* dst, int elem);

}  // namespace cwi
