// cwfield: MI355X (gfx950) diffusion field solve behind a C ABI.
//
// Host side of the library: field handles (device mirrors of
// crates/engine/src/pde.rs:30-53), launch plans, the sweep loop of
// ScalarField::relax (pde.rs:245-288), FP32 smoothing with FP64 defect
// correction, and the batched solve. Kernels live in cw_kernels.hpp.
#include "cw_kernels.hpp"
#include "cw_internal.hpp"
#include "cw_shapes.hpp"
#include "cw_campaign.hpp"

#include <algorithm>
#include <memory>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

using namespace cw;

namespace cwi {

namespace {
thread_local std::string g_err;
std::once_flag g_init_once;
cw_status g_init_status = CW_ERR_DEVICE;
std::string g_device_name;
std::string g_version;
}  // namespace

cw_status fail(cw_status code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

cw_status device_init() {
    std::call_once(g_init_once, [] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
            g_init_status = CW_ERR_DEVICE;
            return;
        }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
            g_init_status = CW_ERR_DEVICE;
            return;
        }
        g_device_name = prop.gcnArchName;
        g_init_status = CW_OK;
    });
    if (g_init_status != CW_OK) return fail(CW_ERR_DEVICE, "no usable HIP device (gfx950 expected)");
    return CW_OK;
}

cw_status alloc_fill(void** p, size_t bytes, hipStream_t s, int fill) {
    CW_HIP(hipMalloc(p, bytes));
    CW_HIP(hipMemsetAsync(*p, fill, bytes, s));
    return CW_OK;
}

const char* last_error_text() { return g_err.c_str(); }
const std::string& device_name() { return g_device_name; }
std::string& version_text() { return g_version; }

}  // namespace cwi

using namespace cwi;

namespace {

cw_status alloc_zero(void** p, size_t bytes, hipStream_t s, int fill) {
    return alloc_fill(p, bytes, s, fill);
}

constexpr double kPi = 3.14159265358979323846;

// The plane tile shapes (kPlainShapes, kSmoothShapes) are in cw_shapes.hpp.
// 3D shapes: kind 0 = in-place red and black half sweeps (half3d, two
// launches per sweep); kind 1 = z-streaming relax3d_stream with NW waves of
// NR rows and zc planes per chunk (one launch per sweep, ping-pong); kind 2 =
// the pipelined FP32 smoother relax3d_pipe, two sweeps per launch (below).
struct Shape3 { int kind, nw, nr, zc; };
constexpr Shape3 k3DShapes[] = {{0, 4, 1, 1},  {1, 8, 4, 16}, {1, 4, 8, 16}, {1, 8, 4, 32},
                                 {1, 8, 4, 8},  {1, 8, 4, 24}, {1, 4, 8, 32}, {1, 8, 4, 48},
                                 {1, 4, 8, 24}, {1, 8, 4, 4},  {1, 4, 8, 8},  {1, 4, 8, 4},
                                 {2, 16, 2, 16}, {2, 16, 2, 32}, {2, 16, 2, 8}, {1, 8, 4, 2},
                                 {1, 4, 8, 2},  {1, 8, 4, 1},  {2, 16, 2, 0}};
// kind 2: two sweeps per launch (relax3d_pipe<2>, FP32 smoothing only), 16
// waves of 2 rows; its remainder launches use the 8 x 4 one-sweep kernel,
// which has the same 32-row region. zc 0: the chunk length that fits the
// plan's workgroups on the CUs in the fewest rounds (pipe_chunks).
constexpr bool two_sweep(int shape3) { return k3DShapes[shape3].kind == 2; }
constexpr int kVolumePipeAuto = 18;
constexpr int kNum3DShapes = (int)(sizeof(k3DShapes) / sizeof(k3DShapes[0]));
constexpr int kDefault3DShape = 1;

// Planes whose largest problem has at least this many sweep tiles fold each
// checked launch's tile partials once per problem (fold_sums) instead of in
// every workgroup of the next launch. Below it the extra launch costs more
// than the workgroups' reads (at the default shapes a 1024^2 plane has 170
// to 220 tiles, 4096^2 2,700 to 3,200). CW_FOLD_TILES overrides it; both ways
// give the same bits.
constexpr int kFoldTiles = 512;
int fold_threshold() {
    const char* e = getenv("CW_FOLD_TILES");
    if (!e || !*e) return kFoldTiles;
    const long v = strtol(e, nullptr, 10);
    return v < 1 ? 1 : (int)std::min<long>(v, 1L << 30);
}

// Volumes: the z-streaming kernels are latency-bound when workgroups are
// few, so the halo planes a shorter chunk re-reads matter less than the
// workgroups it adds, down to a point:
//   FP64: the longest chunk of 16 or 8 planes that gives the batch 220
//     workgroups, else the longest of 4, 2 or 1 plane that gives 150;
//   FP32: the longest chunk of 16, 8, 4 or 2 planes that gives 256, else 2
//     planes if they give 150, else 1 plane if it gives 150;
// failing both, spans under 2^18 voxels take the in-place red and black half
// sweeps (an elementwise grid, two launches a sweep) and larger ones 1-plane
// chunks. Whole solves of spheroids to 1e-8, ms (in-place / 1 / 2 / 4 / 8
// planes; the tiles of the chosen chunk):
//   FP64  64^3: 3.61 / 3.30 / 3.85 / 4.84 / -     (186)
//         80^3: 5.10 / 4.31 / 4.97 / 6.26 / -     (234)
//         96^3: 7.39 / 7.27 / 6.15 / 7.73 / 10.7  (188)
//        120^3: 11.2 / 12.1 / 11.8 / 10.8 / 15.0  (150)
//        160^3: 36.7 / -    / 28.8 / 26.0 / 23.1  (240)
//   FP32  64^3: 2.23 / 1.88 / 2.24 / 2.96 / -     (186)
//         80^3: 3.14 / 2.28 / 2.76 / 3.74 / -     (234)
//         96^3: 4.92 / 3.56 / 3.39 / 4.55 / 6.82  (188)
//        120^3: 8.29 / 6.17 / 5.81 / 6.10 / 9.19  (295)
//        160^3: 33.3 / -    / 14.1 / 11.8 / 13.6  (480)
// and per sweep on tissue fields (in-place / 4 / 8 / 16 planes, us): FP64
// 151^3 61 / 46 / 39 / 64, FP32 151^3 63 / 27 / 29 / 47; 16 planes are best
// at 254^3 (480 workgroups).
constexpr int kVolumeInPlace = 0, kVolumeChunk16 = 1, kVolumeChunk8 = 4, kVolumeChunk4 = 9,
              kVolumeChunk2 = 15, kVolumeChunk1 = 17;
long volume_tiles(cw_field* const* F, uint32_t nf, int si, uint64_t* voxels) {
    const Shape3 sh = k3DShapes[si];
    const int ty = sh.nw * sh.nr - 4;
    long tiles = 0;
    *voxels = 0;
    for (uint32_t i = 0; i < nf; i++) {
        const cw_field* f = F[i];
        if (!f->has_span) continue;
        const int sx = f->span[1] - f->span[0] + 1, sy = f->span[3] - f->span[2] + 1;
        const int sz = f->span[5] - f->span[4] + 1;
        tiles += (long)((sx + 123) / 124) * ((sy + ty - 1) / ty) * ((sz + sh.zc - 1) / sh.zc);
        *voxels += (uint64_t)sx * sy * sz;
    }
    return tiles;
}
// relax3d_pipe addresses a field's FP32 arrays with 32-bit byte offsets
bool pipe_addressable(const cw_field* f) {
    return (uint64_t)f->P * f->H * f->D * sizeof(float) < (1ull << 32);
}

// FP32 volumes of 2^18 voxels or more take the pipelined two-sweep smoother
// with its automatic chunks (kVolumePipeAuto): FP32+refine to 1e-8 on the
// spheroids, ms (one-sweep rule above / two-sweep, automatic chunks):
//   96^3 3.46 / 2.41   128^3 8.53 / 4.63   160^3 11.9 / 7.21
//   192^3 21.2 / 10.4   256^3 50.8 / 30.1 (the same bits, sweep for sweep)
int default_volume_shape(cw_field* const* F, uint32_t nf, bool fp64) {
    uint64_t voxels = 0;
    if (!fp64) {
        bool fits = true;
        for (uint32_t i = 0; i < nf; i++) fits = fits && pipe_addressable(F[i]);
        volume_tiles(F, nf, kVolumeChunk16, &voxels);
        if (fits && voxels >= (1u << 18)) return kVolumePipeAuto;
    }
    if (fp64) {
        for (int c : {kVolumeChunk16, kVolumeChunk8})
            if (volume_tiles(F, nf, c, &voxels) >= 220) return c;
        for (int c : {kVolumeChunk4, kVolumeChunk2, kVolumeChunk1})
            if (volume_tiles(F, nf, c, &voxels) >= 150) return c;
    } else {
        for (int c : {kVolumeChunk16, kVolumeChunk8, kVolumeChunk4, kVolumeChunk2})
            if (volume_tiles(F, nf, c, &voxels) >= 256) return c;
        for (int c : {kVolumeChunk2, kVolumeChunk1})
            if (volume_tiles(F, nf, c, &voxels) >= 150) return c;
    }
    return voxels < (1u << 18) ? kVolumeInPlace : kVolumeChunk1;
}
int default_smooth_shape(int s) { return s <= 1 ? 0 : s == 2 ? 16 : s <= 4 ? 14 : 15; }

// FP32 planes with no sweep count asked for, from the tiles of the whole
// batch (peers included):
//   - 128 x 64 regions at S = 12 (shape 23, 512-thread workgroups, two a CU)
//     when they fit on the CUs at once: more, shorter workgroups;
//   - else 128 x 128 regions at S = 16 (shape 24) when they fit at once: one
//     round of the deepest tile, whose per-sweep latency is least;
//   - else, for fields launched one sequence at a time (a single field, or a
//     batch of more than four), the shape with the least time a sweep when
//     each launch takes whole rounds of the CUs: rounds x (RY / 128) x
//     (L + S w) / S, with a tile load L of 20 sweeps w, over the 128-row
//     shapes at S = 8, 10, 12, 16 and the 160-row ones at S = 8, 10, 12;
//   - else (two to four fields solved concurrently, their launches
//     overlapping) the least tile-work a sweep, tiles x (L + S w) / S: over
//     the 128-row shapes while the batch has fewer than three rounds of
//     128 x 128 tiles at S = 8 (a launch of 160-row regions that spills just
//     past one round holds its stream back a whole workgroup), with the
//     160-row ones too, weighted by RY / 128, beyond that.
// FP32+refine to 1e-8, ms (shapes 23 / 24 / 15 / 19 / 20, the rule's choice
// in brackets):
//   128^2 0.74 / 0.78 / 0.96 / 0.94 / 0.87 (23)   256^2 1.10 / 1.37 / 1.82 / 1.59 / 1.39 (23)
//   512^2 2.14 / 2.52 / 3.33 / 3.00 / 2.77 (23)   768^2 3.36 / 3.47 / 4.35 / 3.91 / 3.62 (23)
//   1024^2 7.04 / 4.29 / 5.87 / 5.07 / 4.73 (24)
//   three 256^2 1.68 / 1.73 / 2.17 / 2.04 / 1.93 (23)   two 512^2 2.60 / 2.81 / 3.70 / 3.28 / 3.06 (23)
//   three 512^2 3.05 / 2.88 / 3.80 / 3.40 / 3.11 (24)
// and (shapes 15 / 19 / 20 / 25 / 26 / 27, the 160-row ones at S = 8, 10, 12):
//   1024^2 5.18 / 4.55 / 4.26 / 5.99 / 5.29 / 4.86 (24: 3.90)   2048^2 21.5 / 24.8 / 22.8 / 23.6 / 21.6 / 19.7 (27)
//   three 768^2 4.80 / 4.69 / 4.54 / 5.48 / 5.23 / 4.59 (15)   three 1024^2 7.81 / 8.20 / 8.77 / 8.08 / 7.93 / 8.53 (15)
//   three 1536^2 23.5 / 25.0 / 27.6 / 21.6 / 22.3 / 23.9 (25)  three 2048^2 54.8 / 55.7 / 61.1 / 46.9 / 50.8 / 53.8 (25)
constexpr int kPlaneDeepShapes[] = {15, 19, 20, 24};  // S = 8, 10, 12, 16 on 128 x 128 regions
constexpr int kPlaneTallShapes[] = {25, 26, 27};      // S = 8, 10, 12 on 128 x 160 regions
constexpr int kPlaneShortShape = 23;                  // S = 12 on 128 x 64 regions
constexpr int kPlaneDeepest = 24;
constexpr double kTileLoadSweeps = 20.0;
long plane_tiles(cw_field* const* F, uint32_t nf, const Shape& sh) {
    const int tx = 128 - 4 * sh.s, ty = sh.ry - 4 * sh.s;
    long tiles = 0;
    for (uint32_t i = 0; i < nf; i++) {
        const cw_field* f = F[i];
        if (!f->has_span) continue;
        const int sx = f->span[1] - f->span[0] + 1, sy = f->span[3] - f->span[2] + 1;
        tiles += (long)((sx + tx - 1) / tx) * ((sy + ty - 1) / ty);
    }
    return tiles;
}
bool run_concurrently(cw_field* const* F, uint32_t nf);
int default_smooth2d_shape(cw_field* const* F, uint32_t nf) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    if (plane_tiles(F, nf, kSmoothShapes[kPlaneShortShape]) <= 2L * cus) return kPlaneShortShape;
    if (plane_tiles(F, nf, kSmoothShapes[kPlaneDeepest]) <= cus) return kPlaneDeepest;
    const bool overlap = run_concurrently(F, nf);
    const bool tall = !overlap || plane_tiles(F, nf, kSmoothShapes[kPlaneDeepShapes[0]]) >= 3L * cus;
    int best = kPlaneDeepShapes[0];
    double best_cost = 0.0;
    auto consider = [&](int si) {
        const Shape sh = kSmoothShapes[si];
        const long tiles = plane_tiles(F, nf, sh);
        const double units = overlap ? (double)tiles : (double)((tiles + cus - 1) / cus);
        const double rows = tall ? sh.ry / 128.0 : 1.0;
        const double cost = units * rows * (kTileLoadSweeps + sh.s) / sh.s;
        if (si == kPlaneDeepShapes[0] || cost < best_cost) {
            best = si;
            best_cost = cost;
        }
    };
    for (int si : kPlaneDeepShapes) consider(si);
    if (tall)
        for (int si : kPlaneTallShapes) consider(si);
    return best;
}

// FP64 planes: s sweeps per launch as asked (1, 2, 4 or 8), else the
// library's choice: 8 sweeps on 128 x 64 regions (shape 9, one 512-thread
// workgroup a CU at 136 VGPRs) while the batch's tiles, peers included, fit
// on the CUs at once, and 4 sweeps on 128 x 64 regions (shape 8, two
// workgroups a CU) beyond. FP64 to 1e-8, ms (4 sweeps on 128 x 32 / 4 on
// 128 x 64 / 8 on 128 x 64; tiles of shape 9):
//   200^2 max change 9.35 / 11.8 / 9.40 (21)    256^2   3.45 / 4.31 / 3.26 (24)
//   512^2   6.71 / 8.60 / 6.57 (96)             768^2  10.6 / 10.8 / 8.36 (192)
//   896^2  12.5 / 12.8 / 17.5 (280)             1024^2 17.6 / 15.1 / 20.5 (352)
//   three 1024^2 fields 35.9 / 29.8 / 48.4 (1056)
// Two sweeps a launch on 128 x 32 regions took 14.1 ms at 200^2 and 35.9 ms
// for the three 1024^2 fields.
constexpr int kPlainDeep32 = 7, kPlainDeep64 = 8, kPlainDeep8 = 9;
int default_plain_shape(int s, cw_field* const* F, uint32_t nf) {
    if (s == 1) return 2;
    if (s == 2) return 3;
    if (s == 8) return kPlainDeep8;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    if (s == 4)
        return plane_tiles(F, nf, kPlainShapes[kPlainDeep32]) <= cus ? kPlainDeep32 : kPlainDeep64;
    return plane_tiles(F, nf, kPlainShapes[kPlainDeep8]) <= cus ? kPlainDeep8 : kPlainDeep64;
}

}  // namespace

namespace {

cw_status ensure_fp32(cw_field* f) {
    if (f->r32) return CW_OK;
    cw_status st;
    if ((st = alloc_zero((void**)&f->e[0], f->npad * sizeof(float), f->stream, 0))) return st;
    if ((st = alloc_zero((void**)&f->e[1], f->npad * sizeof(float), f->stream, 0))) return st;
    if ((st = alloc_zero((void**)&f->lb, f->npad * sizeof(float), f->stream, 0))) return st;
    if ((st = alloc_zero((void**)&f->r32, f->npad * sizeof(float), f->stream, 0xFF))) return st;
    CW_HIP(hipStreamSynchronize(f->stream));
    return CW_OK;
}

double reference_omega(const cw_field* f) {
    if (f->D > 1) {
        int longest = 0;
        for (int a = 0; a < 3; a++) longest = std::max(longest, f->span[2 * a + 1] - f->span[2 * a] + 1);
        longest = std::max(longest, 2);
        return 2.0 / (1.0 + std::sin(kPi / (double)longest));
    }
    const double n = (double)std::max(f->W, f->H);
    return std::min(2.0 / (1.0 + std::sin(kPi / n)), 1.95);
}

// The model-problem optimum 2 / (1 + sin(pi / L)), L the longest edge of the
// free-span box; the reference's volume rule, applied to planes too.
double optimal_omega(const cw_field* f) {
    int longest = 2;
    for (int a = 0; a < 3; a++) longest = std::max(longest, f->span[2 * a + 1] - f->span[2 * a] + 1);
    return 2.0 / (1.0 + std::sin(kPi / (double)longest));
}

double pick_omega(double asked, const cw_field* f) {
    if (asked > 0.0) return asked;
    return asked < 0.0 ? optimal_omega(f) : reference_omega(f);
}

// ---------------------------------------------------------------------------
// Launch plan: problems, control blocks, tile tables, reduction scratch.
// ---------------------------------------------------------------------------
enum Tiling { TL_MAIN = 0, TL_ONE = 1, TL_EL = 2 };

struct Plan {
    hipStream_t stream = nullptr;
    int nprob = 0;
    int s_main = 1;
    int shape = 0;       // kShapes index of the main sweep launch
    bool fp64 = true;
    bool plain = true;   // FP64: floored field (false: the unfloored adjoint)
    bool general = false; // some problem keeps production and uptake apart
    bool three_d = false;
    bool lower_bound = true;  // FP32 plane smoothing: some problem has a lower bound
    int shape3 = kDefault3DShape;   // k3DShapes index for volumes
    int pipe_k = 1;                 // z chunks of each volume in the automatic two-sweep shape
    bool pipe_fits = true;          // every FP32 array of the plan under 4 GiB (relax3d_pipe)
    bool in_place() const { return three_d && k3DShapes[shape3].kind == 0; }
    std::vector<Prob> hp;
    std::vector<Ctl> hc;
    std::vector<int> tp[3];
    Prob* dp = nullptr;
    Ctl* dc = nullptr;
    int* dtp[3] = {nullptr, nullptr, nullptr};
    double* partials = nullptr;
    int part_stride = 0;
    // Per-problem residual sums of a checked launch (fold_sums), used on
    // sweep tilings whose largest problem has at least fold_tiles tiles.
    double* sums = nullptr;
    int max_tiles[3] = {0, 0, 0};
    int fold_tiles = kFoldTiles;
    bool fold(int tl) const { return !three_d && max_tiles[tl] >= fold_tiles; }
    size_t cap_prob = 0, cap_ctl = 0, cap_tp[3] = {0, 0, 0}, cap_part = 0, cap_sums = 0;
    // host mirror of the pass sequence of the current run
    int pass = 0, sweeps_before = 0, prev_s = 0, prev_tl = 0, prev_check = 0;
    // captured chunk of passes (hipGraph), replayed with a device-resident base
    bool use_graph = false;
    int* run = nullptr;
    int run_host[2] = {0, 0};
    hipGraphExec_t gexec = nullptr;
    int g_k = 0, g_crit = -1, g_ce = 0, g_max = 0;
    double g_tol = 0.0;
    // What a captured graph's launches bake in: buffers, grids, shapes. A
    // graph is replayed only while the plan still has this layout, so a plan
    // kept between solves reuses its graph for as long as the field does not
    // change shape.
    struct Layout {
        const void* buf[8];
        int n[8];
        bool operator==(const Layout& o) const {
            return memcmp(this, &o, sizeof *this) == 0;
        }
    };
    Layout layout() const {
        Layout L;
        memset(&L, 0, sizeof L);
        L.buf[0] = dp; L.buf[1] = dc; L.buf[2] = dtp[0]; L.buf[3] = dtp[1];
        L.buf[4] = dtp[2]; L.buf[5] = partials; L.buf[6] = run; L.buf[7] = sums;
        L.n[0] = (int)tp[0].size(); L.n[1] = (int)tp[1].size(); L.n[2] = (int)tp[2].size();
        L.n[3] = nprob; L.n[4] = part_stride; L.n[5] = s_main;
        L.n[6] = shape | (shape3 << 8) | (three_d << 16) | (fp64 << 17) | (plain << 18) |
                 (general << 19) | (lower_bound << 20);
        L.n[7] = fold(TL_MAIN) | (fold(TL_ONE) << 1);
        return L;
    }
    Layout g_layout{};
    // run_passes lays the plan out again over its unstopped problems
    bool compact = false;
    std::vector<char> act;
    int nact = 0;
    // Arena mode: the device buffers are carved out of a region the plan does
    // not own (the allocation-free entry point's scratch) instead of being
    // allocated; plan_upload fails if the layout does not fit.
    unsigned char* arena = nullptr;
    size_t arena_bytes = 0;
    // Pinned host staging for the plan's uploads and control-block traffic:
    // copies from pageable memory are staged by the runtime, one blocking
    // bounce (and a blit kernel) each, which cost a warm-started relax more
    // than its sweeps. Uploads are packed into `stage` and copied from there;
    // the control blocks go both ways through `ctl_pin`. An event after each
    // batch of copies guards the buffer until the copies have read it. The
    // allocation-free mode (arena) allocates nothing and copies as before.
    unsigned char* stage = nullptr;
    size_t stage_cap = 0, stage_off = 0;
    hipEvent_t stage_ev = nullptr;
    bool stage_pending = false;
    Ctl* ctl_pin = nullptr;
    int* run_pin = nullptr;  // two ints after the control blocks
    size_t ctl_cap = 0;
    int run_dev[2] = {0, 0};  // the graph's pass base as the device holds it
    bool run_dev_ok = false;
    hipEvent_t ctl_ev = nullptr;
    bool ctl_pending = false;
    // what the device's tile tables and problems hold (not in arena mode):
    // an upload of the same content to the same buffer is skipped
    std::vector<int> dev_tp[3];
    const int* dev_tp_at[3] = {nullptr, nullptr, nullptr};

    // Back to a fresh plan, keeping the device buffers and their capacities
    // (and dropping a captured graph, whose launch grids and arguments belong
    // to the previous layout).
    void reuse() {
        // a captured graph stays: capture_chunk replays it only if the new
        // solve has the same layout and pattern
        stream = nullptr;
        nprob = 0; s_main = 1; shape = 0;
        fp64 = true; plain = true; general = false; three_d = false; lower_bound = true;
        shape3 = kDefault3DShape;
        hp.clear(); hc.clear();
        for (int i = 0; i < 3; i++) tp[i].clear();
        part_stride = 0;
        pass = 0; sweeps_before = 0; prev_s = 0; prev_tl = 0; prev_check = 0;
        use_graph = false;
        run_host[0] = run_host[1] = 0;
        compact = false; act.clear(); nact = 0;
        pipe_k = 1; pipe_fits = true;
        run_dev_ok = false;
    }

    ~Plan() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (stage_ev) (void)hipEventSynchronize(stage_ev);
        if (ctl_ev) (void)hipEventSynchronize(ctl_ev);
        if (stage) (void)hipHostFree(stage);
        if (ctl_pin) (void)hipHostFree(ctl_pin);
        if (stage_ev) (void)hipEventDestroy(stage_ev);
        if (ctl_ev) (void)hipEventDestroy(ctl_ev);
        if (arena) return;  // nothing else here was allocated by the plan
        (void)hipFree(run);
        (void)hipFree(dp); (void)hipFree(dc); (void)hipFree(partials); (void)hipFree(sums);
        for (int i = 0; i < 3; i++) (void)hipFree(dtp[i]);
    }
};

template <typename T>
cw_status grow(T** p, size_t* cap, size_t need) {
    if (need <= *cap) return CW_OK;
    (void)hipFree(*p);
    *p = nullptr;
    size_t c = std::max<size_t>(need, 2 * *cap);
    CW_HIP(hipMalloc((void**)p, c * sizeof(T)));
    *cap = c;
    return CW_OK;
}

// A field's own plans for single-field solves: `main` for the FP64 sweep or
// the refinement's outer (FP64) plan, `aux` for the refinement's FP32 inner
// plan. Batches of several fields in one plan allocate their own.
struct PlanCache {
    Plan main, aux;
};

Plan& field_plan(cw_field* f, int which) {
    if (!f->plans) f->plans = new PlanCache();
    PlanCache* pc = (PlanCache*)f->plans;
    Plan& p = which ? pc->aux : pc->main;
    p.reuse();
    return p;
}

// z chunks per volume for the automatic two-sweep shape. A relax3d_pipe
// workgroup (16 waves, 112 VGPRs at most: one a CU) streams its chunk of zc
// planes in zc + 7 steps of about the same time, so a launch takes
// rounds x (zc + 7) steps, rounds = ceil(workgroups / CUs): the chunk count k
// that minimises ceil(columns k / CUs) (ceil(planes / k) + 7) over the plan's
// active volumes (256^3: 33 columns, k = 7, 37-plane chunks, one round).
int pipe_chunks(const Plan& pl, const std::vector<char>* active) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    long cols = 0;
    int depth = 1;
    for (int p = 0; p < pl.nprob; p++) {
        if (active && !(*active)[p]) continue;
        const Prob& pr = pl.hp[p];
        const int sx = pr.x1 - pr.xs + 1, sy = pr.y1 - (pr.ys & ~1) + 1;
        cols += (long)((sx + 119) / 120) * ((sy + 23) / 24);
        depth = std::max(depth, pr.z1 - pr.zs + 1);
    }
    if (cols <= 0) return 1;
    int best = 1;
    long best_cost = -1;
    for (int k = 1; k <= depth; k++) {
        const long rounds = (cols * k + cus - 1) / cus;
        const long cost = rounds * ((depth + k - 1) / k + 7);
        if (best_cost < 0 || cost < best_cost) {
            best = k;
            best_cost = cost;
        }
    }
    return best;
}

// Makes the plan's pinned upload buffer hold `bytes` once the copies that
// still read it are done (no staging in arena mode).
cw_status stage_begin(Plan& pl, size_t bytes) {
    pl.stage_off = 0;
    if (pl.arena) return CW_OK;
    if (pl.stage_pending) {
        CW_HIP(hipEventSynchronize(pl.stage_ev));
        pl.stage_pending = false;
    }
    if (!pl.stage_ev) CW_HIP(hipEventCreateWithFlags(&pl.stage_ev, hipEventDisableTiming));
    if (bytes > pl.stage_cap) {
        if (pl.stage) (void)hipHostFree(pl.stage);
        pl.stage = nullptr;
        pl.stage_cap = 0;
        const size_t cap = std::max<size_t>(bytes + bytes / 2, 4096);
        CW_HIP(hipHostMalloc((void**)&pl.stage, cap, hipHostMallocDefault));
        pl.stage_cap = cap;
    }
    return CW_OK;
}

// Host-to-device copy of `bytes` from src through the upload buffer.
cw_status stage_h2d(Plan& pl, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return CW_OK;
    if (pl.arena || !pl.stage) {
        CW_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, pl.stream));
        return CW_OK;
    }
    if (pl.stage_off + bytes > pl.stage_cap) return fail(CW_ERR_DEVICE, "upload staging overrun");
    unsigned char* at = pl.stage + pl.stage_off;
    memcpy(at, src, bytes);
    CW_HIP(hipMemcpyAsync(dst, at, bytes, hipMemcpyHostToDevice, pl.stream));
    pl.stage_off += (bytes + 255) / 256 * 256;
    return CW_OK;
}

cw_status stage_end(Plan& pl) {
    if (pl.arena || !pl.stage || pl.stage_off == 0) return CW_OK;
    CW_HIP(hipEventRecord(pl.stage_ev, pl.stream));
    pl.stage_pending = true;
    return CW_OK;
}

// The control blocks' pinned buffer, free of pending copies.
cw_status ctl_buffer(Plan& pl) {
    if (pl.arena) return CW_OK;
    if (pl.ctl_pending) {
        CW_HIP(hipEventSynchronize(pl.ctl_ev));
        pl.ctl_pending = false;
    }
    if (!pl.ctl_ev) CW_HIP(hipEventCreateWithFlags(&pl.ctl_ev, hipEventDisableTiming));
    const size_t need = (size_t)std::max(1, pl.nprob);
    if (need > pl.ctl_cap) {
        if (pl.ctl_pin) (void)hipHostFree(pl.ctl_pin);
        pl.ctl_pin = nullptr;
        pl.run_pin = nullptr;
        pl.ctl_cap = 0;
        CW_HIP(hipHostMalloc((void**)&pl.ctl_pin, need * sizeof(Ctl) + 64, hipHostMallocDefault));
        pl.ctl_cap = need;
        pl.run_pin = (int*)((unsigned char*)pl.ctl_pin + need * sizeof(Ctl));
    }
    return CW_OK;
}

// Lays out the tilings of every problem (s-sweep tile, one-sweep tile,
// elementwise tile) and uploads problems, control blocks and tables.
// With `active`, problems marked 0 get no tiles: launches skip them without
// dispatching a workgroup. (A captured graph whose grids followed the old tile
// lists no longer matches the plan's layout and is captured again.)
cw_status plan_upload(Plan& pl, const std::vector<char>* active = nullptr) {
    const int np = pl.nprob;
    for (int tl = 0; tl < 3; tl++) pl.tp[tl].clear();
    if (pl.three_d && two_sweep(pl.shape3) && k3DShapes[pl.shape3].zc == 0)
        pl.pipe_k = pipe_chunks(pl, active);
    for (int p = 0; p < np; p++) {
        Prob& pr = pl.hp[p];
        const int sx = pr.x1 - pr.xs + 1, sy = pr.y1 - pr.ys + 1, sz = pr.z1 - pr.zs + 1;
        const bool on = !active || (*active)[p];
        for (int tl = 0; tl < 3; tl++) {
            int nx, ny, nz;
            if (!on) {
                nx = ny = nz = 0;
            } else if (tl == TL_EL) {
                nx = (sx + 127) / 128; ny = (sy + 3) / 4; nz = sz;
            } else if (pl.three_d) {
                const Shape3 sh = k3DShapes[pl.shape3];
                if (sh.kind == 0) {
                    nx = ny = nz = 0;
                } else {
                    // two sweeps need a halo of 4 (one sweep 2) in x and y;
                    // the two-sweep kernel's y tiling starts on an even row
                    const bool two = sh.kind == 2 && tl == TL_MAIN;
                    const int tx = two ? 120 : 124;
                    const int ty = sh.nw * sh.nr - (two ? 8 : 4);
                    const int syk = two ? pr.y1 - (pr.ys & ~1) + 1 : sy;
                    // one chunk length for both tilings (the kernels share pr.zc)
                    const int zc = sh.zc == 0 ? (sz + pl.pipe_k - 1) / pl.pipe_k : sh.zc;
                    nx = (sx + tx - 1) / tx; ny = (syk + ty - 1) / ty; nz = (sz + zc - 1) / zc;
                    pr.zc = zc;
                }
            } else {
                const int si = tl == TL_MAIN ? pl.shape : 0;
                const Shape sh = (pl.fp64 && pl.plain) ? kPlainShapes[si] : kSmoothShapes[si];
                const int tx = 128 - 4 * sh.s, ty = sh.ry - 4 * sh.s;
                nx = (sx + tx - 1) / tx; ny = (sy + ty - 1) / ty; nz = 1;
            }
            const int base = (int)pl.tp[tl].size(), count = nx * ny * nz;
            if (tl == TL_EL) {
                pr.el_base = base; pr.el_count = count; pr.elx = nx; pr.ely = ny;
            } else {
                pr.sw_base[tl] = base; pr.sw_count[tl] = count; pr.ntx[tl] = nx; pr.nty[tl] = ny;
            }
            pl.tp[tl].insert(pl.tp[tl].end(), (size_t)count, p);
        }
    }
    cw_status st;
    size_t maxt = 1;
    for (int tl = 0; tl < 3; tl++) maxt = std::max(maxt, pl.tp[tl].size());
    pl.part_stride = (int)(2 * maxt);
    for (int tl = 0; tl < 3; tl++) {
        pl.max_tiles[tl] = 0;
        if (tl == TL_EL) continue;
        for (int p = 0; p < np; p++) pl.max_tiles[tl] = std::max(pl.max_tiles[tl], pl.hp[p].sw_count[tl]);
    }
    pl.fold_tiles = fold_threshold();
    if (pl.arena) {
        // problems, control blocks, the three tile tables and the partials,
        // each on a 256-byte boundary
        size_t off = 0;
        auto take = [&](size_t bytes) {
            unsigned char* p = pl.arena + off;
            off += (bytes + 255) / 256 * 256;
            return p;
        };
        pl.dp = (Prob*)take(std::max(1, np) * sizeof(Prob));
        pl.dc = (Ctl*)take(std::max(1, np) * sizeof(Ctl));
        for (int tl = 0; tl < 3; tl++) pl.dtp[tl] = (int*)take(std::max<size_t>(1, pl.tp[tl].size()) * sizeof(int));
        pl.partials = (double*)take(2 * (size_t)pl.part_stride * sizeof(double));
        pl.sums = (double*)take(2 * (size_t)std::max(1, np) * sizeof(double));
        if (off > pl.arena_bytes)
            return fail(CW_ERR_ARG,
                        "the solve needs %zu bytes of launch-plan scratch; the allocation-free "
                        "entry point holds %zu",
                        off, pl.arena_bytes);
    } else {
        if ((st = grow(&pl.dp, &pl.cap_prob, (size_t)std::max(1, np)))) return st;
        if ((st = grow(&pl.dc, &pl.cap_ctl, (size_t)std::max(1, np)))) return st;
        for (int tl = 0; tl < 3; tl++) {
            const size_t cap0 = pl.cap_tp[tl];
            if (!pl.tp[tl].empty() && (st = grow(&pl.dtp[tl], &pl.cap_tp[tl], pl.tp[tl].size())))
                return st;
            // a new allocation may reuse the old address: its content is new
            if (pl.cap_tp[tl] != cap0) pl.dev_tp_at[tl] = nullptr;
        }
        if ((st = grow(&pl.partials, &pl.cap_part, 2 * (size_t)pl.part_stride))) return st;
        if ((st = grow(&pl.sums, &pl.cap_sums, 2 * (size_t)std::max(1, np)))) return st;
    }
    size_t bytes = 0;
    for (int tl = 0; tl < 3; tl++) bytes += (pl.tp[tl].size() * sizeof(int) + 255) / 256 * 256;
    bytes += (np * sizeof(Prob) + 255) / 256 * 256 + (np * sizeof(Ctl) + 255) / 256 * 256;
    if ((st = stage_begin(pl, bytes))) return st;
    for (int tl = 0; tl < 3; tl++) {
        if (pl.tp[tl].empty()) continue;
        // a warm-started relax of an unchanged field lays out the same tiles
        if (!pl.arena && pl.dev_tp_at[tl] == pl.dtp[tl] && pl.dev_tp[tl] == pl.tp[tl]) continue;
        if ((st = stage_h2d(pl, pl.dtp[tl], pl.tp[tl].data(), pl.tp[tl].size() * sizeof(int))))
            return st;
        if (!pl.arena) {
            pl.dev_tp[tl] = pl.tp[tl];
            pl.dev_tp_at[tl] = pl.dtp[tl];
        }
    }
    if ((st = stage_h2d(pl, pl.dp, pl.hp.data(), np * sizeof(Prob)))) return st;
    if ((st = stage_h2d(pl, pl.dc, pl.hc.data(), np * sizeof(Ctl)))) return st;
    return stage_end(pl);
}

PassArgs make_args(Plan& pl, int tl, int mode, int crit, double tol, int max_sweeps) {
    PassArgs a{};
    a.probs = pl.dp;
    a.ctl = pl.dc;
    a.tile_prob = pl.dtp[tl];
    a.partials = pl.partials;
    a.part_stride = pl.part_stride;
    a.nprob = pl.nprob;
    a.pass = pl.pass;
    a.sweeps_before = pl.sweeps_before;
    a.prev_s = pl.prev_s;
    a.prev_tl = pl.prev_tl;
    a.prev_check = pl.prev_check;
    a.tl = tl == TL_EL ? 0 : tl;
    a.crit = crit;
    a.mode = mode;
    a.max_sweeps = max_sweeps;
    a.tol = tol;
    a.project = 1;
    return a;
}

cw_status launch_sweep(Plan& pl, const PassArgs& a_in) {
    const int nt = (int)pl.tp[a_in.tl].size();
    if (nt == 0) return CW_OK;
    PassArgs a = a_in;
    a.sums = a.prev_check && pl.fold(a.prev_tl) ? pl.sums : nullptr;
    const int shape = a.tl == TL_MAIN ? pl.shape : 0;
    const dim3 grid(nt), b256(256), b512(512), b1024(1024);
    hipStream_t st = pl.stream;
#define CW_L(kern, blk) hipLaunchKernelGGL(kern, grid, blk, 0, st, a)
// the register smoother with or without the lower bound of fields with sinks
#define CW_LB(S_, NW_, NR_, blk)                                                   \
    do {                                                                           \
        if (pl.lower_bound) CW_L((relax2d_reg<float, S_, NW_, NR_, false, true>), blk); \
        else CW_L((relax2d_reg<float, S_, NW_, NR_, false, false>), blk);          \
    } while (0)
    if (pl.fp64 && pl.plain) {
        if (pl.general) {
            switch (shape) {
            case 0: CW_L((sweep2d<1, 32, 4, true>), b256); break;
            case 1: CW_L((sweep2d<2, 32, 4, true>), b256); break;
            case 2: CW_L((relax2d_reg<double, 1, 4, 8, true, true>), b256); break;
            case 3: CW_L((relax2d_reg<double, 2, 4, 8, true, true>), b256); break;
            case 4: CW_L((relax2d_reg<double, 2, 8, 8, true, true>), b512); break;
            case 5: CW_L((relax2d_reg<double, 2, 16, 4, true, true>), b1024); break;
            case 7: CW_L((relax2d_reg<double, 4, 4, 8, true, true>), b256); break;
            case 8: CW_L((relax2d_reg<double, 4, 8, 8, true, true>), b512); break;
            case 9: CW_L((relax2d_reg<double, 8, 8, 8, true, true>), b512); break;
            default: CW_L((relax2d_reg<double, 1, 8, 4, true, true>), b512); break;
            }
        } else {
            switch (shape) {
            case 0: CW_L((sweep2d<1, 32, 4, false>), b256); break;
            case 1: CW_L((sweep2d<2, 32, 4, false>), b256); break;
            case 2: CW_L((relax2d_reg<double, 1, 4, 8, true, false>), b256); break;
            case 3: CW_L((relax2d_reg<double, 2, 4, 8, true, false>), b256); break;
            case 4: CW_L((relax2d_reg<double, 2, 8, 8, true, false>), b512); break;
            case 5: CW_L((relax2d_reg<double, 2, 16, 4, true, false>), b1024); break;
            case 7: CW_L((relax2d_reg<double, 4, 4, 8, true, false>), b256); break;
            case 8: CW_L((relax2d_reg<double, 4, 8, 8, true, false>), b512); break;
            case 9: CW_L((relax2d_reg<double, 8, 8, 8, true, false>), b512); break;
            default: CW_L((relax2d_reg<double, 1, 8, 4, true, false>), b512); break;
            }
        }
    } else if (pl.fp64) {
        switch (shape) {
        case 0: CW_L((smooth2d<double, 1, 32, 4>), b256); break;
        case 1: CW_L((smooth2d<double, 2, 32, 4>), b256); break;
        default: CW_L((relax2d_reg<double, 2, 4, 8, false, false>), b256); break;
        }
    } else {
        switch (shape) {
        case 0: CW_L((smooth2d<float, 1, 32, 4>), b256); break;
        case 1: CW_L((smooth2d<float, 2, 32, 4>), b256); break;
        case 2: CW_L((smooth2d<float, 4, 64, 8>), b512); break;
        case 3: CW_L((smooth2d<float, 4, 64, 4>), b256); break;
        case 4: CW_L((smooth2d<float, 8, 128, 8>), b512); break;
        case 5: CW_L((smooth2d<float, 8, 128, 16>), b1024); break;
        case 6: CW_L((smooth2d_reg<float, 2, 4, 8>), b256); break;
        case 7: CW_L((smooth2d_reg<float, 4, 16, 4>), b1024); break;
        case 8: CW_L((smooth2d_reg<float, 4, 8, 8>), b512); break;
        case 9: CW_L((smooth2d_reg<float, 8, 16, 8>), b1024); break;
        case 10: CW_L((smooth2d_reg<float, 2, 8, 8>), b512); break;
        case 11: CW_L((smooth2d_reg<float, 4, 16, 8>), b1024); break;
        case 12: CW_LB(2, 4, 8, b256); break;
        case 13: CW_LB(4, 8, 8, b512); break;
        case 14: CW_LB(4, 16, 4, b1024); break;
        case 15: CW_LB(8, 16, 8, b1024); break;
        case 16: CW_LB(2, 8, 8, b512); break;
        case 17: CW_LB(4, 16, 8, b1024); break;
        case 18: CW_LB(6, 16, 8, b1024); break;
        case 19: CW_LB(10, 16, 8, b1024); break;
        case 21: CW_LB(6, 8, 8, b512); break;
        case 22: CW_LB(8, 8, 8, b512); break;
        case 23: CW_LB(12, 8, 8, b512); break;
        case 24: CW_LB(16, 16, 8, b1024); break;
        case 25: CW_LB(8, 16, 10, b1024); break;
        case 26: CW_LB(10, 16, 10, b1024); break;
        case 27: CW_LB(12, 16, 10, b1024); break;
        default: CW_LB(12, 16, 8, b1024); break;
        }
    }
#undef CW_LB
#undef CW_L
    CW_HIP(hipGetLastError());
    if (a.check && pl.fold(a.tl)) {
        a.sums = pl.sums;
        hipLaunchKernelGGL(fold_sums, dim3((pl.nprob + 3) / 4), dim3(256), 0, st, a);
        CW_HIP(hipGetLastError());
    }
    return CW_OK;
}

cw_status launch_stream3d(Plan& pl, const PassArgs& a) {
    const int nt = (int)pl.tp[a.tl].size();
    if (nt == 0) return CW_OK;
    const dim3 grid(nt), b256(256), b512(512), b1024(1024);
    hipStream_t st = pl.stream;
    const bool w4 = k3DShapes[pl.shape3].nw == 4;  // 4 waves of 8 rows, else 8 of 4
#define CW_L(kern, blk) hipLaunchKernelGGL(kern, grid, blk, 0, st, a)
    if (two_sweep(pl.shape3) && a.tl == TL_MAIN) {
        if (pl.fp64) return fail(CW_ERR_ARG, "two-sweep volume shapes are for FP32 smoothing only");
        if (!pl.pipe_fits)
            return fail(CW_ERR_ARG, "two-sweep volume shapes address arrays under 4 GiB only");
        if (pl.lower_bound) CW_L((relax3d_pipe<2, 16, 2, true>), b1024);
        else CW_L((relax3d_pipe<2, 16, 2, false>), b1024);
    } else if (pl.fp64 && pl.plain) {
        if (pl.general) {
            if (w4) CW_L((relax3d_stream<double, 4, 8, true, true>), b256);
            else CW_L((relax3d_stream<double, 8, 4, true, true>), b512);
        } else {
            if (w4) CW_L((relax3d_stream<double, 4, 8, true, false>), b256);
            else CW_L((relax3d_stream<double, 8, 4, true, false>), b512);
        }
    } else if (pl.fp64) {
        if (w4) CW_L((relax3d_stream<double, 4, 8, false, false>), b256);
        else CW_L((relax3d_stream<double, 8, 4, false, false>), b512);
    } else {
        if (w4) CW_L((relax3d_stream<float, 4, 8, false, false>), b256);
        else CW_L((relax3d_stream<float, 8, 4, false, false>), b512);
    }
#undef CW_L
    CW_HIP(hipGetLastError());
    return CW_OK;
}

cw_status launch_half3d(Plan& pl, PassArgs a, int colour) {
    const int nt = (int)pl.tp[TL_EL].size();
    if (nt == 0) return CW_OK;
    a.tile_prob = pl.dtp[TL_EL];
    if (pl.fp64 && pl.plain) hipLaunchKernelGGL((half3d<double, true>), dim3(nt), dim3(256), 0, pl.stream, a, colour);
    else if (pl.fp64) hipLaunchKernelGGL((half3d<double, false>), dim3(nt), dim3(256), 0, pl.stream, a, colour);
    else hipLaunchKernelGGL((half3d<float, false>), dim3(nt), dim3(256), 0, pl.stream, a, colour);
    CW_HIP(hipGetLastError());
    return CW_OK;
}

cw_status launch_decide(Plan& pl, int mode, int crit, double tol, int max_sweeps) {
    PassArgs a = make_args(pl, TL_MAIN, mode, crit, tol, max_sweeps);
    hipLaunchKernelGGL(decide, dim3(1), dim3(kDecideThreads), 0, pl.stream, a);
    CW_HIP(hipGetLastError());
    return CW_OK;
}

cw_status read_ctl(Plan& pl) {
    cw_status st;
    if ((st = ctl_buffer(pl))) return st;
    Ctl* to = pl.arena ? pl.hc.data() : pl.ctl_pin;
    CW_HIP(hipMemcpyAsync(to, pl.dc, pl.nprob * sizeof(Ctl), hipMemcpyDeviceToHost, pl.stream));
    CW_HIP(hipStreamSynchronize(pl.stream));
    if (to != pl.hc.data()) memcpy(pl.hc.data(), to, pl.nprob * sizeof(Ctl));
    return CW_OK;
}

cw_status write_ctl(Plan& pl) {
    cw_status st;
    if ((st = ctl_buffer(pl))) return st;
    if (pl.arena) {
        CW_HIP(hipMemcpyAsync(pl.dc, pl.hc.data(), pl.nprob * sizeof(Ctl), hipMemcpyHostToDevice,
                              pl.stream));
        return CW_OK;
    }
    memcpy(pl.ctl_pin, pl.hc.data(), pl.nprob * sizeof(Ctl));
    CW_HIP(hipMemcpyAsync(pl.dc, pl.ctl_pin, pl.nprob * sizeof(Ctl), hipMemcpyHostToDevice,
                          pl.stream));
    CW_HIP(hipEventRecord(pl.ctl_ev, pl.stream));
    pl.ctl_pending = true;
    return CW_OK;
}

bool any_status(const Plan& pl, int st) {
    for (int p = 0; p < pl.nprob; p++)
        if (pl.hc[p].status == st) return true;
    return false;
}

void run_reset(Plan& pl) {
    pl.pass = 0; pl.sweeps_before = 0; pl.prev_s = 0; pl.prev_tl = 0; pl.prev_check = 0;
}

// One sweep launch (2D: s sweeps; 3D: a red and a black half sweep).
cw_status one_pass(Plan& pl, int s, int crit, int check, double tol, int max_sweeps) {
    cw_status st;
    if (pl.three_d) {
        if (pl.in_place()) {
            PassArgs a = make_args(pl, TL_EL, MODE_SWEEP, crit, tol, max_sweeps);
            a.s = 1;
            if ((st = launch_half3d(pl, a, 0))) return st;
            if ((st = launch_half3d(pl, a, 1))) return st;
            s = 1;
        } else {
            // s sweeps: the main tiling at the plan's depth, the one-sweep
            // tiling for a remainder
            const int tl = (s == pl.s_main) ? TL_MAIN : TL_ONE;
            PassArgs a = make_args(pl, tl, MODE_SWEEP, crit, tol, max_sweeps);
            a.s = s;
            if ((st = launch_stream3d(pl, a))) return st;
            pl.prev_tl = tl;
        }
        check = 0;
    } else {
        const int tl = (s == pl.s_main) ? TL_MAIN : TL_ONE;
        PassArgs a = make_args(pl, tl, MODE_SWEEP, crit, tol, max_sweeps);
        a.s = s;
        a.check = check;
        if ((st = launch_sweep(pl, a))) return st;
        pl.prev_tl = tl;
    }
    pl.pass += 1;
    pl.sweeps_before += s;
    pl.prev_s = s;
    pl.prev_check = check;
    return CW_OK;
}

struct LoopResult { long passes = 0; long graph_replays = 0; };

// Captures K passes (launch j carrying pass offset j), the end-of-chunk
// verdict and the advance of the device base into one hipGraph; cached per
// plan for the same pattern.
cw_status capture_chunk(Plan& pl, int K, int crit, int ce, double tol, int maxs) {
    const int S = pl.s_main;
    if (!pl.run) CW_HIP(hipMalloc((void**)&pl.run, 2 * sizeof(int)));
    if (pl.gexec && pl.g_k == K && pl.g_crit == crit && pl.g_ce == ce && pl.g_tol == tol &&
        pl.g_max == maxs && pl.g_layout == pl.layout())
        return CW_OK;
    if (pl.gexec) {
        (void)hipGraphExecDestroy(pl.gexec);
        pl.gexec = nullptr;
    }
    CW_HIP(hipStreamBeginCapture(pl.stream, hipStreamCaptureModeThreadLocal));
    cw_status st = CW_OK;
    auto pattern_check = [&](int j) { return crit == CRIT_RESID && (j % ce == 0) ? 1 : 0; };
    for (int j = 0; j < K && st == CW_OK; j++) {
        const int tl = pl.three_d && pl.in_place() ? TL_EL : TL_MAIN;
        PassArgs a = make_args(pl, tl, MODE_SWEEP, crit, tol, maxs);
        a.pass = j;
        a.sweeps_before = j * S;
        a.s = S;
        a.prev_s = S;
        a.prev_tl = TL_MAIN;
        a.prev_check = pl.three_d ? 0 : pattern_check((j + K - 1) % K);
        a.check = pl.three_d ? 0 : pattern_check(j);
        a.run = pl.run;
        if (pl.three_d && pl.in_place()) {
            st = launch_half3d(pl, a, 0);
            if (st == CW_OK) st = launch_half3d(pl, a, 1);
        } else if (pl.three_d) {
            st = launch_stream3d(pl, a);
        } else {
            st = launch_sweep(pl, a);
        }
    }
    if (st == CW_OK) {
        PassArgs a = make_args(pl, TL_MAIN, DECIDE_PASS, crit, tol, maxs);
        a.pass = K;
        a.sweeps_before = K * S;
        a.prev_s = S;
        a.prev_tl = TL_MAIN;
        a.prev_check = pl.three_d ? 0 : pattern_check(K - 1);
        a.run = pl.run;
        hipLaunchKernelGGL(decide, dim3(1), dim3(kDecideThreads), 0, pl.stream, a);
        hipLaunchKernelGGL(advance_run, dim3(1), dim3(64), 0, pl.stream, pl.run, K, K * S);
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(pl.stream, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (ec != hipSuccess) return fail(CW_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&pl.gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess)
        return fail(CW_ERR_DEVICE, "graph instantiate failed: %s", hipGetErrorString(ei));
    pl.g_k = K; pl.g_crit = crit; pl.g_ce = ce; pl.g_tol = tol; pl.g_max = maxs;
    pl.g_layout = pl.layout();
    return CW_OK;
}

// Lays a batch plan out again over its unstopped problems once a quarter of
// them have stopped, so the tail of a batch does not dispatch workgroups that
// only find their problem stopped. Called right after a chunk's verdicts were
// read: the chunk-end decide has already judged the last launch exactly as
// the next launch would, so that launch is told not to judge it again from
// partials laid out by the old tiling (prev_check cleared). The change slots
// are per problem and unaffected.
cw_status compact_plan(Plan& pl) {
    if (!pl.compact || pl.nprob < 2) return CW_OK;
    if (pl.act.empty()) {
        pl.act.assign(pl.nprob, 1);
        pl.nact = pl.nprob;
    }
    int nrun = 0;
    for (int p = 0; p < pl.nprob; p++)
        nrun += pl.hc[p].status == ST_RUNNING || pl.hc[p].status == ST_REPLAY;
    if (nrun * 4 > pl.nact * 3) return CW_OK;
    for (int p = 0; p < pl.nprob; p++)
        pl.act[p] = pl.hc[p].status == ST_RUNNING || pl.hc[p].status == ST_REPLAY;
    pl.nact = nrun;
    pl.prev_check = 0;
    return plan_upload(pl, &pl.act);
}

// Runs passes until no problem is RUNNING or `budget` sweeps have run; the
// device stops a problem once the run has done `dev_max` sweeps. Problems
// that stopped inside a fused pass are replayed one sweep at a time from the
// pass's input.
cw_status run_passes(Plan& pl, int crit, double tol, uint64_t budget, uint64_t dev_max,
                     int check_every, LoopResult* out) {
    cw_status st;
    const int maxs = (int)std::min<uint64_t>(dev_max, 0x7fffffff);
    const uint64_t start = (uint64_t)pl.sweeps_before;
    long launched = 0;
    const int S = pl.s_main;
    // Without a stopping rule the run ends at the budget: enqueue all of it
    // and read the verdicts once. With one, start small and double, so a
    // warm start that settles in a few sweeps does not queue many no-op
    // launches.
    // A launch runs S sweeps, or one (the one-sweep tiling) for the last
    // budget % S.
    int chunk = crit == CRIT_NONE
                    ? (int)std::min<uint64_t>(budget / (uint64_t)S + budget % (uint64_t)S, 1u << 20)
                    : 4;
    constexpr int kGraphPasses = 16;
    while (true) {
        // steady state: replay a captured chunk of passes
        const uint64_t done0 = (uint64_t)pl.sweeps_before - start;
        const bool steady =
            pl.pass == 0 ||
            (pl.prev_s == S && pl.prev_tl == TL_MAIN &&
             pl.prev_check == ((crit == CRIT_RESID && !pl.three_d && ((pl.pass - 1) % check_every == 0)) ? 1 : 0));
        if (pl.use_graph && steady && kGraphPasses % check_every == 0 &&
            pl.pass % check_every == 0 && done0 + (uint64_t)(kGraphPasses * S) <= budget &&
            !(pl.three_d && crit == CRIT_RESID)) {
            if ((st = capture_chunk(pl, kGraphPasses, crit, check_every, tol, maxs))) return st;
            // the graph reads its pass base from the device; the chunk's
            // last node advances it, so it is uploaded only when eager
            // passes (or a new run) moved the host's count away from it
            if (!pl.run_dev_ok || pl.run_dev[0] != pl.pass || pl.run_dev[1] != pl.sweeps_before) {
                if ((st = ctl_buffer(pl))) return st;
                int* from = pl.arena ? pl.run_host : pl.run_pin;
                from[0] = pl.run_host[0] = pl.pass;
                from[1] = pl.run_host[1] = pl.sweeps_before;
                CW_HIP(hipMemcpyAsync(pl.run, from, 2 * sizeof(int), hipMemcpyHostToDevice,
                                      pl.stream));
                if (!pl.arena) {
                    CW_HIP(hipEventRecord(pl.ctl_ev, pl.stream));
                    pl.ctl_pending = true;
                }
            }
            CW_HIP(hipGraphLaunch(pl.gexec, pl.stream));
            pl.run_dev[0] = pl.pass + kGraphPasses;
            pl.run_dev[1] = pl.sweeps_before + kGraphPasses * S;
            pl.run_dev_ok = true;
            pl.pass += kGraphPasses;
            pl.sweeps_before += kGraphPasses * S;
            pl.prev_s = S;
            pl.prev_tl = TL_MAIN;
            pl.prev_check = (crit == CRIT_RESID && !pl.three_d && ((kGraphPasses - 1) % check_every == 0)) ? 1 : 0;
            launched += kGraphPasses;
            if (out) out->graph_replays++;
            // Without a stopping rule nothing is decided between replays (a
            // field whose budget is spent finds so in each launch's
            // prologue): the next replay is queued behind this one at once,
            // and the verdicts are read once, after the remainder.
            if (crit == CRIT_NONE) continue;
            if ((st = read_ctl(pl))) return st;
            if (!any_status(pl, ST_RUNNING)) break;
            if ((st = compact_plan(pl))) return st;
            continue;
        }
        for (int j = 0; j < chunk; j++) {
            const uint64_t done = (uint64_t)pl.sweeps_before - start;
            if (done >= budget) break;
            const int s = budget - done >= (uint64_t)pl.s_main ? pl.s_main : 1;
            const int check = crit == CRIT_RESID && (pl.pass % check_every == 0);
            if ((st = one_pass(pl, s, crit, check, tol, maxs))) return st;
            launched++;
        }
        if ((st = launch_decide(pl, DECIDE_PASS, crit, tol, maxs))) return st;
        if ((st = read_ctl(pl))) return st;
        if (!any_status(pl, ST_RUNNING)) break;
        if ((uint64_t)pl.sweeps_before - start >= budget) break;
        if ((st = compact_plan(pl))) return st;
        chunk = std::min(chunk * 2, 64);
    }
    if (any_status(pl, ST_REPLAY)) {
        int most = 0;
        for (int p = 0; p < pl.nprob; p++)
            if (pl.hc[p].status == ST_REPLAY) most = std::max(most, pl.hc[p].replay);
        for (int r = 0; r < most; r++) {
            PassArgs a = make_args(pl, TL_ONE, MODE_REPLAY, crit, tol, maxs);
            a.s = 1;
            a.replay_step = r;
            if ((st = launch_sweep(pl, a))) return st;
        }
        if ((st = launch_decide(pl, DECIDE_REPLAY_END, crit, tol, maxs))) return st;
        if ((st = read_ctl(pl))) return st;
    }
    if (out) out->passes += launched;
    return CW_OK;
}

// FP64 residual of every running problem (optionally writing the FP32
// defect inputs), then r2/b2 recorded and converged problems stopped.
cw_status run_resid(Plan& pl, int crit, double tol, int write32, int project = 1) {
    const int nt = (int)pl.tp[TL_EL].size();
    if (nt == 0) return CW_OK;
    PassArgs a = make_args(pl, TL_EL, DECIDE_RESID, crit, tol, 0);
    a.project = project;
    hipLaunchKernelGGL(resid64, dim3(nt), dim3(256), 0, pl.stream, a, write32);
    CW_HIP(hipGetLastError());
    cw_status st = launch_decide(pl, DECIDE_RESID, crit, tol, 0);
    if (st) return st;
    return read_ctl(pl);
}

void fill_prob_common(Prob& pr, const cw_field* f, double omega) {
    memset(&pr, 0, sizeof pr);
    pr.W = (int)f->W; pr.H = (int)f->H; pr.D = (int)f->D; pr.P = f->P;
    pr.xs = f->span[0] & ~1; pr.ys = f->span[2]; pr.zs = f->span[4];
    pr.x1 = f->span[1]; pr.y1 = f->span[3]; pr.z1 = f->span[5];
    pr.d = f->diff; pr.decay = f->decay; pr.source = f->source;
    pr.omega = omega; pr.omo = 1.0 - omega;
}

void init_ctl(Ctl& c, int status) {
    memset(&c, 0, sizeof c);
    c.status = status;
}

// Starts the timed windows of fields timed concurrently together, so each
// stream's launches overlap the others' as they do in a solve. A thread that
// never arrives (an error before its timing) cannot hold the others for
// long: they wait at most a second.
struct Gate {
    std::mutex m;
    std::condition_variable cv;
    int expected = 0, arrived = 0;
    void arrive_and_wait() {
        std::unique_lock<std::mutex> lk(m);
        if (++arrived >= expected) {
            cv.notify_all();
            return;
        }
        cv.wait_for(lk, std::chrono::seconds(1), [&] { return arrived >= expected; });
    }
};

struct Opts {
    double omega;
    int crit;
    double tol;
    uint64_t max_sweeps;
    int precision;
    int s;
    int check_every;
    double inner_tol;
    uint64_t max_outer;
    int shape = -1;            // tile shape index forced by the caller (-1: from s)
    bool use_graph = false;    // replay steady chunks of passes from a captured hipGraph
    long time_passes = 0;      // > 0: time this many sweep launches instead of solving
    double* time_ms = nullptr; // average duration of one sweep launch (HIP events)
    Gate* time_gate = nullptr; // fields timed concurrently start their windows together
    // the fields solved concurrently with this one (all of them share the
    // GPU, so the plane tile shape is chosen for their tiles together)
    cw_field* const* peers = nullptr;
    uint32_t npeers = 0;
    // allocation-free solve: the launch plan lives in this region (no graphs)
    unsigned char* arena = nullptr;
    size_t arena_bytes = 0;
};

// Times n sweep launches of the plan on its stream with HIP events (the
// kernels' own durations plus the gaps between back-to-back launches).
cw_status time_launches(Plan& pl, long n, double* ms, Gate* gate = nullptr) {
    if (gate) {
        CW_HIP(hipStreamSynchronize(pl.stream));  // set-up done before the window opens
        gate->arrive_and_wait();
    }
    hipEvent_t e0, e1;
    CW_HIP(hipEventCreate(&e0));
    CW_HIP(hipEventCreate(&e1));
    cw_status st = CW_OK;
    run_reset(pl);
    CW_HIP(hipEventRecord(e0, pl.stream));
    for (long i = 0; i < n && st == CW_OK; i++)
        st = one_pass(pl, pl.s_main, CRIT_NONE, 0, 0.0, 0x7fffffff);
    CW_HIP(hipEventRecord(e1, pl.stream));
    CW_HIP(hipEventSynchronize(e1));
    float t = 0.0f;
    CW_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) return st;
    *ms = (double)t / (double)n;
    return CW_OK;
}

cw_status check_opts(const cw_relax_opts* o, Opts* out) {
    if (!o) return fail(CW_ERR_ARG, "null options");
    if (o->criterion != CW_CRIT_MAX_CHANGE && o->criterion != CW_CRIT_REL_RESIDUAL)
        return fail(CW_ERR_ARG, "unknown criterion %d", o->criterion);
    if (o->precision != CW_PREC_FP64 && o->precision != CW_PREC_FP32_REFINE)
        return fail(CW_ERR_ARG, "unknown precision %d", o->precision);
    if (o->precision == CW_PREC_FP32_REFINE && o->criterion != CW_CRIT_REL_RESIDUAL)
        return fail(CW_ERR_ARG, "FP32 refinement stops on the relative residual only");
    if (!(o->tolerance == o->tolerance)) return fail(CW_ERR_ARG, "tolerance is NaN");
    if (o->omega > 0.0 && !(o->omega < 2.0))
        return fail(CW_ERR_ARG, "omega %g outside (0, 2)", o->omega);
    if (o->sweeps_per_pass < 0 || o->sweeps_per_pass > 8 ||
        (o->sweeps_per_pass & (o->sweeps_per_pass - 1)) != 0)
        return fail(CW_ERR_ARG, "sweeps_per_pass must be 0, 1, 2, 4 or 8");
    {
        // planes index the 2D table of the precision, volumes the 3D table;
        // the fields' dimensionality is checked where the plan is built
        const bool plain = o->precision == CW_PREC_FP64;
        const int n2 = plain ? kNumPlainShapes : kNumSmoothShapes;
        if (o->tile_shape >= std::max(n2, kNum3DShapes))
            return fail(CW_ERR_ARG, "unknown tile shape %d", o->tile_shape);
        const int ts = (o->tile_shape < 0 || o->tile_shape >= n2) ? 0
                       : plain ? kPlainShapes[o->tile_shape].s
                               : kSmoothShapes[o->tile_shape].s;
        if (o->tile_shape >= 0 && o->sweeps_per_pass && ts != o->sweeps_per_pass)
            return fail(CW_ERR_ARG, "tile shape %d does not do %d sweeps per pass", o->tile_shape,
                        o->sweeps_per_pass);
    }
    out->omega = o->omega;
    out->crit = o->criterion == CW_CRIT_MAX_CHANGE ? CRIT_MAXCHANGE : CRIT_RESID;
    out->tol = o->tolerance;
    out->max_sweeps = o->max_sweeps;
    out->precision = o->precision;
    out->s = o->sweeps_per_pass;
    out->check_every = o->check_every > 0 ? o->check_every : 0;
    // The FP32 smoother at omega near 2 stalls after a 1e-3..1e-4 reduction of
    // the defect; asking each segment for 2e-2 keeps every segment at the full
    // SOR rate (fewest total sweeps on planes, volumes and campaigns alike).
    out->inner_tol = o->inner_tolerance > 0.0 ? o->inner_tolerance : 2e-2;
    out->max_outer = o->max_outer > 0 ? o->max_outer : 1000;
    out->shape = o->tile_shape;
    out->use_graph = o->use_graph != 0;
    return CW_OK;
}

uint64_t span_voxels(const cw_field* f) {
    return (uint64_t)(f->span[1] - f->span[0] + 1) * (f->span[3] - f->span[2] + 1) *
           (f->span[5] - f->span[4] + 1);
}

// Plain FP64 solve of a set of fields (each sweeps its own span box).
cw_status solve_fp64(cw_field* const* F, uint32_t nf, const Opts& o, cw_relax_stats* stats) {
    cw_status st;
    hipStream_t stream = F[0]->stream;
    const bool three_d = F[0]->D > 1;
    for (uint32_t i = 0; i < nf; i++)
        if ((F[i]->D > 1) != three_d)
            return fail(CW_ERR_ARG, "a batch mixes planes and volumes");
    int s = o.s ? std::min(o.s, kMaxS) : 0;  // 0: the library's choice
    if (three_d) s = 1;
    Plan local;
    Plan& pl = (nf == 1 && !o.arena) ? field_plan(F[0], 0) : local;
    pl.arena = o.arena;
    pl.arena_bytes = o.arena_bytes;
    pl.stream = stream;
    pl.fp64 = true;
    pl.three_d = three_d;
    pl.s_main = s;
    pl.shape = o.shape >= 0 ? o.shape
               : o.peers ? default_plain_shape(s, o.peers, o.npeers)
                         : default_plain_shape(s, F, nf);
    if (!three_d && pl.shape >= kNumPlainShapes)
        return fail(CW_ERR_ARG, "unknown tile shape %d", pl.shape);
    if (three_d && o.shape >= kNum3DShapes)
        return fail(CW_ERR_ARG, "unknown volume tile shape %d", o.shape);
    pl.s_main = three_d ? 1 : kPlainShapes[pl.shape].s;
    pl.use_graph = o.use_graph;
    pl.compact = nf > 1 && o.time_passes == 0;
    if (three_d) {
        // volumes: tile_shape picks a 3D shape, one sweep per launch
        pl.shape3 = (o.shape >= 0 && o.shape < kNum3DShapes) ? o.shape
                                                               : default_volume_shape(F, nf, true);
        if (two_sweep(pl.shape3))
            return fail(CW_ERR_ARG, "volume shape %d sweeps twice per launch: FP32 smoothing only",
                        pl.shape3);
        pl.shape = 0;
        pl.s_main = 1;
    }
    s = pl.s_main;
    std::vector<int> idx;  // plan problem -> field
    std::vector<double> omegas(nf);
    for (uint32_t i = 0; i < nf; i++) {
        cw_field* f = F[i];
        memset(&stats[i], 0, sizeof(cw_relax_stats));
        stats[i].rel_residual = std::nan("");
        stats[i].sweeps_per_pass = s;
        stats[i].tile_shape = three_d ? pl.shape3 : pl.shape;
        if (!f->has_span) {
            // Nothing to balance: Ok(1) (pde.rs:256-259).
            stats[i].sweeps = o.crit == CRIT_MAXCHANGE ? 1 : 0;
            continue;
        }
        omegas[i] = pick_omega(o.omega, f);
        stats[i].omega = omegas[i];
        stats[i].span_voxels = span_voxels(f);
        if (o.max_sweeps == 0 && o.time_passes == 0) {
            stats[i].status = CW_ERR_NOT_CONVERGED;
            continue;
        }
        // Outside the swept box both buffers must agree (pde.rs:249-252, :278).
        CW_HIP(hipMemcpyAsync(f->c[f->cur ^ 1], f->c[f->cur], f->npad * sizeof(double),
                              hipMemcpyDeviceToDevice, f->stream));
        Prob pr;
        fill_prob_common(pr, f, omegas[i]);
        pr.x[0] = f->c[f->cur];
        pr.x[1] = (three_d && k3DShapes[pl.shape3].kind == 0) ? f->c[f->cur] : f->c[f->cur ^ 1];
        pr.k = f->k;
        pr.k2 = f->general ? f->k2 : nullptr;
        if (f->general) pl.general = true;
        pr.c64 = f->c[f->cur];
        pl.hp.push_back(pr);
        Ctl c;
        init_ctl(c, ST_RUNNING);
        pl.hc.push_back(c);
        idx.push_back((int)i);
    }
    for (uint32_t i = 0; i < nf; i++)
        if (F[i]->stream != stream) CW_HIP(hipStreamSynchronize(F[i]->stream));
    pl.nprob = (int)idx.size();
    if (pl.nprob == 0) return CW_OK;
    if ((st = plan_upload(pl))) return st;
    if (o.time_passes > 0) {
        if ((st = time_launches(pl, o.time_passes, o.time_ms, o.time_gate))) return st;
        for (int p = 0; p < pl.nprob; p++)
            if (!pl.in_place() && (pl.pass & 1)) F[idx[p]]->cur ^= 1;
        return CW_OK;
    }
    const int check_every = o.check_every ? o.check_every : (three_d ? 8 : 4);
    LoopResult lr;
    run_reset(pl);
    if (o.crit == CRIT_RESID) {
        if ((st = run_resid(pl, CRIT_RESID, o.tol, 0))) return st;
    }
    if (three_d && o.crit == CRIT_RESID) {
        // Volumes test the residual in its own pass every check_every sweeps.
        while (any_status(pl, ST_RUNNING) && (uint64_t)pl.sweeps_before < o.max_sweeps) {
            const uint64_t n =
                std::min<uint64_t>(check_every, o.max_sweeps - (uint64_t)pl.sweeps_before);
            if ((st = run_passes(pl, CRIT_NONE, o.tol, n, o.max_sweeps, 1, &lr))) return st;
            for (int p = 0; p < pl.nprob; p++) pl.hp[p].c64 = (double*)pl.hp[p].x[pl.pass & 1];
            CW_HIP(hipMemcpyAsync(pl.dp, pl.hp.data(), pl.nprob * sizeof(Prob),
                                  hipMemcpyHostToDevice, pl.stream));
            if ((st = run_resid(pl, CRIT_RESID, o.tol, 0))) return st;
        }
        if ((st = launch_decide(pl, DECIDE_PASS, CRIT_NONE, o.tol,
                                (int)std::min<uint64_t>(o.max_sweeps, 0x7fffffff))))
            return st;
        if ((st = read_ctl(pl))) return st;
    } else {
        if ((st = run_passes(pl, o.crit, o.tol, o.max_sweeps, o.max_sweeps, check_every, &lr)))
            return st;
    }
    for (int p = 0; p < pl.nprob; p++) {
        cw_field* f = F[idx[p]];
        const Ctl& c = pl.hc[p];
        if (!pl.in_place() && (c.parity & 1)) f->cur ^= 1;
        cw_status fs = CW_OK;
        if (c.nonfinite) fs = CW_ERR_NONFINITE;
        else if (c.status != ST_DONE) fs = CW_ERR_NOT_CONVERGED;
        cw_relax_stats& S = stats[idx[p]];
        S.sweeps = (uint64_t)c.sweeps;
        S.passes = (uint64_t)lr.passes;
        S.status = fs;
        if (o.crit == CRIT_RESID && c.b2 > 0.0) S.rel_residual = std::sqrt(c.r2 / c.b2);
    }
    return CW_OK;
}

// Sweep budget of the next FP32 segment of one problem. From the FP64
// residuals at the last two corrections, the observed rate gives the sweeps
// that reach the target, but one segment is asked for no more than
// `per_segment` (FP32 cannot carry a correction much further); 10% margin,
// clamped to [lo, hi]. While the residual has not yet started to fall (the
// SOR transient) the segment doubles.
int next_segment(const std::vector<double>& res, const std::vector<int>& at, double target,
                 double per_segment, int first, int lo, int hi, int s, double model_rate = 0.0) {
    int k = first;
    const size_t n = res.size();
    if (n >= 2) {
        const int last = at[n - 1] - at[n - 2];
        const double goal = std::max(target / res[n - 1], per_segment);
        if (res[n - 1] > 0.0 && res[n - 2] > res[n - 1] && last > 0) {
            const double rate = std::log(res[n - 1] / res[n - 2]) / (double)last;
            const double need = std::log(goal) / rate;
            k = need > 0.0 ? (int)std::ceil(1.1 * need) + s : lo;
        } else {
            k = 2 * std::max(last, first);
        }
        // A rate measured across the SOR transient can be far slower than the
        // asymptotic one; the model problem's rate bounds the budget.
        if (model_rate < 0.0 && goal < 1.0)
            k = std::min(k, (int)std::ceil(1.5 * std::log(goal) / model_rate) + first / 2);
    }
    k = std::max(lo, std::min(hi, k));
    return (k + s - 1) / s * s;
}

// log of the asymptotic residual reduction per sweep of red-black SOR on the
// field's free-span box with all faces held (the model problem): with mu the
// Jacobi radius, rho = omega - 1 above the optimum and
// ((omega mu + sqrt(omega^2 mu^2 - 4 (omega - 1))) / 2)^2 below it (Young's
// theory for consistently ordered matrices). Decay shrinks mu by A_d / (A_d +
// decay). Fields held on more than the box converge at least this fast.
double model_log_rate(const cw_field* f, double omega) {
    const int dims = f->D > 1 ? 3 : 2;
    double mu = 0.0;
    for (int a = 0; a < dims; a++)
        mu += std::cos(kPi / (double)(f->span[2 * a + 1] - f->span[2 * a] + 2));
    mu /= (double)dims;
    const double ad = 2.0 * dims * f->diff;
    if (ad + f->decay > 0.0) mu *= ad / (ad + f->decay);
    const double disc = omega * omega * mu * mu - 4.0 * (omega - 1.0);
    double rho = omega - 1.0;
    if (disc > 0.0) {
        const double r = 0.5 * (omega * mu + std::sqrt(disc));
        rho = r * r;
    }
    rho = std::fabs(rho);
    return (rho > 0.0 && rho < 1.0) ? std::log(rho) : 0.0;
}

// FP32 smoothing of the defect equation with FP64 residual and correction.
//   r = b - A c (FP64), A e = r smoothed in FP32 from e = 0 with e >= -c,
//   c <- max(c + e, 0) (FP64), until ||rt|| / ||b_eff|| <= tol.
// SOR is stationary, so restarting e from zero on the fresh FP64 residual
// keeps the error the iteration works on: segments cost only their FP64
// passes and keep the correction small enough for FP32.
cw_status solve_refine(cw_field* const* F, uint32_t nf, const Opts& o, cw_relax_stats* stats) {
    cw_status st;
    hipStream_t stream = F[0]->stream;
    const bool three_d = F[0]->D > 1;
    for (uint32_t i = 0; i < nf; i++)
        if ((F[i]->D > 1) != three_d)
            return fail(CW_ERR_ARG, "a batch mixes planes and volumes");
    int shape = o.shape >= 0 ? o.shape
                : o.s      ? default_smooth_shape(o.s)
                : three_d  ? 0
                : o.peers  ? default_smooth2d_shape(o.peers, o.npeers)
                           : default_smooth2d_shape(F, nf);
    if (!three_d && shape >= kNumSmoothShapes)
        return fail(CW_ERR_ARG, "unknown tile shape %d", shape);
    if (three_d && o.shape >= kNum3DShapes)
        return fail(CW_ERR_ARG, "unknown volume tile shape %d", o.shape);
    if (three_d) shape = 0;
    // Segment lengths are multiples of s. Volumes round them to two sweeps
    // whatever the shape, so the one- and two-sweep smoothers run the same
    // segments (and give the same bits), and a batch's two-sweep launches
    // end every field's segment on a launch boundary.
    const int s = three_d ? 2 : kSmoothShapes[shape].s;
    Plan local_outer, local_inner;
    Plan& outer = nf == 1 ? field_plan(F[0], 0) : local_outer;
    Plan& inner = nf == 1 ? field_plan(F[0], 1) : local_inner;
    outer.stream = inner.stream = stream;
    outer.fp64 = true;
    inner.fp64 = false;
    outer.three_d = inner.three_d = three_d;
    outer.s_main = 1;
    inner.s_main = s;
    inner.shape = shape;
    inner.use_graph = o.use_graph;
    if (three_d) {
        inner.shape3 = (o.shape >= 0 && o.shape < kNum3DShapes) ? o.shape
                                                                 : default_volume_shape(F, nf, false);
        inner.s_main = two_sweep(inner.shape3) ? 2 : 1;
        for (uint32_t i = 0; i < nf; i++)
            if (!pipe_addressable(F[i])) inner.pipe_fits = false;
    }
    const int s_used = inner.s_main;
    inner.lower_bound = false;
    std::vector<int> idx;
    std::vector<double> omegas(nf);
    int longest = 2;
    for (uint32_t i = 0; i < nf; i++) {
        cw_field* f = F[i];
        memset(&stats[i], 0, sizeof(cw_relax_stats));
        stats[i].rel_residual = std::nan("");
        stats[i].sweeps_per_pass = s_used;
        stats[i].tile_shape = three_d ? inner.shape3 : inner.shape;
        if (!f->has_span) continue;
        if ((st = ensure_fp32(f))) return st;
        omegas[i] = pick_omega(o.omega, f);
        stats[i].omega = omegas[i];
        stats[i].span_voxels = span_voxels(f);
        for (int a = 0; a < 3; a++) longest = std::max(longest, f->span[2 * a + 1] - f->span[2 * a] + 1);
        Prob pr;
        fill_prob_common(pr, f, omegas[i]);
        pr.k = f->k;
        pr.k2 = f->general ? f->k2 : nullptr;
        pr.c64 = f->c[f->cur];
        pr.e0 = f->e[0];
        // correct64 reads the inner answer from e[parity]; an in-place inner
        // plan keeps it in e[0] whatever the parity
        pr.e1 = (three_d && k3DShapes[inner.shape3].kind == 0) ? f->e[0] : f->e[1];
        pr.r32 = f->r32;
        pr.lb32 = f->has_sink ? f->lb : nullptr;
        pr.inner_idx = (int)idx.size();
        outer.hp.push_back(pr);
        Prob ip = pr;
        ip.x[0] = f->e[0];
        ip.x[1] = (three_d && k3DShapes[inner.shape3].kind == 0) ? f->e[0] : f->e[1];
        ip.k = f->r32;
        ip.k2 = nullptr;
        ip.lb = pr.lb32;
        if (ip.lb) inner.lower_bound = true;
        inner.hp.push_back(ip);
        Ctl c;
        init_ctl(c, ST_RUNNING);
        outer.hc.push_back(c);
        init_ctl(c, ST_IDLE);
        inner.hc.push_back(c);
        idx.push_back((int)i);
    }
    const int np = (int)idx.size();
    outer.nprob = inner.nprob = np;
    if (np == 0) return CW_OK;
    if ((st = plan_upload(outer))) return st;
    if ((st = plan_upload(inner))) return st;
    run_reset(outer);
    if ((st = run_resid(outer, CRIT_RESID, o.tol, 1))) return st;
    if (o.time_passes > 0) {
        for (int p = 0; p < np; p++) init_ctl(inner.hc[p], ST_RUNNING);
        if ((st = write_ctl(inner))) return st;
        return time_launches(inner, o.time_passes, o.time_ms, o.time_gate);
    }
    std::vector<double> model_rate(np);
    for (int p = 0; p < np; p++) model_rate[p] = model_log_rate(F[idx[p]], outer.hp[p].omega);
    std::vector<std::vector<double>> hist(np);
    std::vector<std::vector<int>> hist_at(np);
    std::vector<int> total(np, 0), outer_it(np, 0);
    for (int p = 0; p < np; p++) {
        const Ctl& c = outer.hc[p];
        hist[p].push_back(c.b2 > 0.0 ? std::sqrt(c.r2 / c.b2) : std::sqrt(c.r2));
        hist_at[p].push_back(0);
    }
    // First segment: a fraction of the box's longest side, the scale of the
    // SOR transient; later ones follow the measured rate.
    const int first = std::max(32, longest / 4);
    const int hi = std::max(256, 4 * longest);
    long passes = 0;
    // Fields that have settled keep no tiles: once a quarter of the plan's
    // fields have stopped, both plans are laid out again over the rest, so a
    // batch's tail does not dispatch workgroups that only find their field
    // stopped.
    std::vector<char> act(np, 1);
    int nact = np;
    // Sweeps field p may still run under max_sweeps. In a batch the launches
    // run as many sweeps as the longest segment needs (s at a time), and a
    // field's own budget is judged only between launches, so its allowance
    // is whole launches; alone, a remainder runs one sweep a launch.
    auto room = [&](int p) -> uint64_t {
        const uint64_t left = o.max_sweeps > (uint64_t)total[p] ? o.max_sweeps - (uint64_t)total[p] : 0;
        return np > 1 ? left / (uint64_t)s * (uint64_t)s : left;
    };
    for (uint64_t it = 0; it < o.max_outer && any_status(outer, ST_RUNNING); it++) {
        int nrun = 0;
        for (int p = 0; p < np; p++) nrun += outer.hc[p].status == ST_RUNNING;
        if (np > 1 && nrun * 4 <= nact * 3) {
            for (int p = 0; p < np; p++) act[p] = outer.hc[p].status == ST_RUNNING;
            if ((st = plan_upload(outer, &act))) return st;
            if ((st = plan_upload(inner, &act))) return st;
            nact = nrun;
        }
        int most = 0;
        for (int p = 0; p < np; p++) {
            const bool run = outer.hc[p].status == ST_RUNNING && room(p) > 0;
            init_ctl(inner.hc[p], run ? ST_RUNNING : ST_IDLE);
            if (!run) continue;
            int k = next_segment(hist[p], hist_at[p], o.tol, o.inner_tol, first, 16, hi, s,
                                 model_rate[p]);
            if ((uint64_t)k > room(p)) k = (int)room(p);
            inner.hc[p].budget = std::max(k, 1);
            most = std::max(most, inner.hc[p].budget);
            outer_it[p]++;
        }
        if (most <= 0) break;
        if ((st = write_ctl(inner))) return st;
        run_reset(inner);
        LoopResult lr;
        if ((st = run_passes(inner, CRIT_NONE, 0.0, (uint64_t)most, 0x7fffffff, 1, &lr))) return st;
        passes += lr.passes;
        for (int p = 0; p < np; p++)
            if (outer.hc[p].status == ST_RUNNING) total[p] += inner.hc[p].sweeps;
        {
            const int nt = (int)outer.tp[TL_EL].size();
            PassArgs a = make_args(outer, TL_EL, MODE_SWEEP, CRIT_NONE, 0.0, 0);
            a.inner_ctl = inner.dc;
            hipLaunchKernelGGL(correct64, dim3(nt), dim3(256), 0, stream, a);
            CW_HIP(hipGetLastError());
        }
        if ((st = run_resid(outer, CRIT_RESID, o.tol, 1))) return st;
        bool any_left = false;
        for (int p = 0; p < np; p++) {
            const Ctl& c = outer.hc[p];
            if (inner.hc[p].status == ST_IDLE) continue;
            hist[p].push_back(c.b2 > 0.0 ? std::sqrt(c.r2 / c.b2) : std::sqrt(c.r2));
            hist_at[p].push_back(total[p]);
            if (c.status == ST_RUNNING && room(p) > 0) any_left = true;
        }
        if (!any_left) break;
    }
    for (int p = 0; p < np; p++) {
        const Ctl& c = outer.hc[p];
        cw_status fs = CW_OK;
        if (c.nonfinite) fs = CW_ERR_NONFINITE;
        else if (c.status != ST_DONE) fs = CW_ERR_NOT_CONVERGED;
        cw_relax_stats& S = stats[idx[p]];
        S.sweeps = (uint64_t)total[p];
        S.passes = (uint64_t)passes;
        S.outer_iters = (uint64_t)outer_it[p];
        S.status = fs;
        if (c.b2 > 0.0) S.rel_residual = std::sqrt(c.r2 / c.b2);
    }
    return CW_OK;
}

// A batch of two to four fields runs as concurrent solves: one host thread
// and the field's own stream each. The fields' launches then overlap on the
// GPU, so one field's tile loads run under another's sweeps instead of every
// CU loading its tile at the start of each launch together (three 1024^2
// fields to 1e-8: FP32+refine 14.2 ms against 18.3 ms as one batched launch
// sequence, FP64 38.2 ms against 57.2 ms). The batch never mixes problems, so
// at one tile shape each field gets the same bits either way (the default
// shape is picked from the tiles of the whole batch). Larger batches keep the single
// launch sequence: they fill the GPU on their own. In timing mode each field
// times its own launches and the result is the longest average interval.
constexpr uint32_t kConcurrentMax = 4;  // hardware queues of a process

bool run_concurrently(cw_field* const* F, uint32_t nf) {
    if (nf < 2 || nf > kConcurrentMax) return false;
    for (uint32_t i = 0; i < nf; i++)
        for (uint32_t j = i + 1; j < nf; j++)
            if (F[i]->stream == F[j]->stream) return false;
    return true;
}

cw_status solve_one(cw_field* f, const Opts& o, cw_relax_stats* st) {
    cw_field* one[1] = {f};
    return o.precision == CW_PREC_FP64 ? solve_fp64(one, 1, o, st) : solve_refine(one, 1, o, st);
}

cw_status solve_concurrently(cw_field* const* F, uint32_t nf, const Opts& o,
                             cw_relax_stats* stats) {
    int dev = 0;
    CW_HIP(hipGetDevice(&dev));
    std::vector<cw_status> sts(nf, CW_OK);
    std::vector<std::string> msgs(nf);
    std::vector<double> ms(nf, 0.0);
    Gate gate;
    gate.expected = (int)nf;
    auto work = [&](uint32_t i) {
        Opts oi = o;
        oi.time_ms = o.time_passes > 0 ? &ms[i] : nullptr;
        oi.time_gate = o.time_passes > 0 ? &gate : nullptr;
        oi.peers = F;
        oi.npeers = nf;
        cw_status st;
        try {
            st = hipSetDevice(dev) == hipSuccess
                     ? solve_one(F[i], oi, &stats[i])
                     : fail(CW_ERR_DEVICE, "hipSetDevice(%d) failed in a solve thread", dev);
        } catch (const std::exception& e) {
            st = fail(CW_ERR_DEVICE, "host error in a solve thread: %s", e.what());
        }
        if (st == CW_OK && hipStreamSynchronize(F[i]->stream) != hipSuccess)
            st = fail(CW_ERR_DEVICE, "stream synchronisation failed in a solve thread");
        sts[i] = st;
        if (st) msgs[i] = last_error_text();
    };
    std::vector<std::thread> th;
    th.reserve(nf);
    uint32_t started = 0;
    try {
        for (; started < nf; started++) th.emplace_back(work, started);
    } catch (...) {
        // no thread to spare: the rest run on this one
    }
    for (uint32_t i = started; i < nf; i++) work(i);
    for (auto& t : th) t.join();
    for (uint32_t i = 0; i < nf; i++)
        if (sts[i]) return fail(sts[i], "%s", msgs[i].c_str());
    if (o.time_passes > 0 && o.time_ms) *o.time_ms = *std::max_element(ms.begin(), ms.end());
    return CW_OK;
}

cw_status solve_impl(cw_field* const* F, uint32_t nf, const cw_relax_opts* opts,
                     cw_relax_stats* stats, long time_passes, double* time_ms,
                     unsigned char* arena, size_t arena_bytes);

// Every solve entry point comes through here: no C++ exception crosses the
// C ABI (host allocations for plans and threads can throw).
cw_status solve(cw_field* const* F, uint32_t nf, const cw_relax_opts* opts, cw_relax_stats* stats,
                long time_passes = 0, double* time_ms = nullptr, unsigned char* arena = nullptr,
                size_t arena_bytes = 0) {
    try {
        return solve_impl(F, nf, opts, stats, time_passes, time_ms, arena, arena_bytes);
    } catch (const std::bad_alloc&) {
        return fail(CW_ERR_DEVICE, "out of host memory");
    } catch (const std::exception& e) {
        return fail(CW_ERR_DEVICE, "host error: %s", e.what());
    }
}

cw_status solve_impl(cw_field* const* F, uint32_t nf, const cw_relax_opts* opts,
                     cw_relax_stats* stats, long time_passes, double* time_ms,
                     unsigned char* arena, size_t arena_bytes) {
    auto t0 = std::chrono::steady_clock::now();
    if (!F || nf == 0) return fail(CW_ERR_ARG, "no fields");
    Opts o;
    cw_status st = check_opts(opts, &o);
    if (st) return st;
    o.time_passes = time_passes;
    o.time_ms = time_ms;
    o.arena = arena;
    o.arena_bytes = arena_bytes;
    if (arena && (nf != 1 || o.precision != CW_PREC_FP64 || o.use_graph || time_passes))
        return fail(CW_ERR_ARG, "the allocation-free solve is one FP64 field without graphs");
    for (uint32_t i = 0; i < nf; i++) {
        if (!F[i]) return fail(CW_ERR_ARG, "null field %u", i);
        if (F[i]->coef_dirty && (st = build_coefficients(F[i], arena == nullptr))) return st;
    }
    // A batch swept on the first field's stream reads the others' freshly
    // built coefficients: their streams finish first. (Alone, or as
    // concurrent solves on their own streams, each field's work is ordered
    // on its stream.)
    if (!run_concurrently(F, nf))
        for (uint32_t i = 1; i < nf; i++)
            if (F[i]->stream != F[0]->stream) CW_HIP(hipStreamSynchronize(F[i]->stream));
    std::vector<cw_relax_stats> local;
    if (!stats) {
        local.resize(nf);
        stats = local.data();
    }
    for (uint32_t i = 1; i < nf; i++)
        if ((F[i]->D > 1) != (F[0]->D > 1)) return fail(CW_ERR_ARG, "a batch mixes planes and volumes");
    if (run_concurrently(F, nf)) {
        if ((st = solve_concurrently(F, nf, o, stats))) return st;
    } else {
        if (o.precision == CW_PREC_FP64) st = solve_fp64(F, nf, o, stats);
        else st = solve_refine(F, nf, o, stats);
        if (st) return st;
        CW_HIP(hipStreamSynchronize(F[0]->stream));
    }
    if (time_passes > 0) return CW_OK;
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    cw_status worst = CW_OK;
    for (uint32_t i = 0; i < nf; i++) {
        stats[i].seconds = secs;
        if (stats[i].status != CW_OK && worst == CW_OK) worst = (cw_status)stats[i].status;
    }
    if (worst == CW_ERR_NOT_CONVERGED)
        return fail(CW_ERR_NOT_CONVERGED,
                    "diffusion did not settle within %llu sweeps at tolerance %g",
                    (unsigned long long)o.max_sweeps, o.tol);
    if (worst == CW_ERR_NONFINITE) return fail(CW_ERR_NONFINITE, "the residual is not finite");
    return CW_OK;
}

// ---------------------------------------------------------------------------
// Device-driven batch of planes (cw_campaign_relax; cw_campaign.hpp): each
// group of fields (a size class) is one launch plan swept with its own tile
// shape; every field's state lives on the device, and one captured graph
// holds every group's chain of steps as parallel branches. The host launches
// the graph as often as the model problem predicts, waits once and reads the
// fields' states; only a field still running then costs another launch.
// ---------------------------------------------------------------------------
constexpr int kCampSteps = 64;  // rounds (steps of every group) in the captured graph

struct CampGroup {
    Plan pl;                      // sweep view of every problem, tiles, partials, control
    std::vector<Prob> hp64;       // FP32 refinement: the FP64 view
    std::vector<CampField> hst;   // per problem, host copy
    std::vector<int> idx;         // plan problem -> field
    unsigned char* dev = nullptr; // CampField[np] | Prob[np] | count | list[ntiles]
    CampField* dst = nullptr;
    Prob* dp64 = nullptr;
    int* run = nullptr;           // the step base every group shares
    int* count = nullptr;
    int* list = nullptr;
    int grid = 0, grid_tail = 0, shape = 0, s = 1;
    std::vector<int> pred;        // per problem: the last step the model problem predicts
    bool gen = false;
    hipStream_t stream = nullptr;
    ~CampGroup() { (void)hipFree(dev); }
};

CampArgs camp_args(CampGroup& g, int step, const Opts& o, int check_every, bool fp64) {
    CampArgs ca{};
    ca.pa = make_args(g.pl, TL_MAIN, MODE_SWEEP, CRIT_RESID, o.tol, 0);
    ca.pa.pass = step;
    ca.pa.s = g.s;
    ca.pa.run = g.run;
    ca.probs64 = g.dp64;
    ca.st = g.dst;
    ca.list = g.list;
    ca.count = g.count;
    ca.fp64 = fp64 ? 1 : 0;
    ca.check_every = check_every;
    ca.lo = 16;
    ca.max_sweeps = (int)std::min<uint64_t>(o.max_sweeps, 0x7fffffff);
    ca.max_outer = (int)std::min<uint64_t>(o.max_outer, 0x7fffffff);
    ca.tol = o.tol;
    ca.inner_tol = o.inner_tol;
    return ca;
}

cw_status campaign_impl(cw_field* const* F, uint32_t nf, const uint32_t* group,
                        const int32_t* shapes, uint32_t ngroups, const cw_relax_opts* opts,
                        cw_relax_stats* stats) {
    auto t0 = std::chrono::steady_clock::now();
    if (!F || nf == 0 || !group || ngroups == 0 || !stats) return fail(CW_ERR_ARG, "bad argument");
    Opts o;
    cw_status st = check_opts(opts, &o);
    if (st) return st;
    if (o.crit != CRIT_RESID)
        return fail(CW_ERR_ARG, "the device-driven batch stops on the relative residual");
    const bool fp64 = o.precision == CW_PREC_FP64;
    const int check_every = o.check_every ? o.check_every : 4;
    for (uint32_t i = 0; i < nf; i++) {
        if (!F[i]) return fail(CW_ERR_ARG, "null field %u", i);
        if (F[i]->D > 1) return fail(CW_ERR_ARG, "the device-driven batch solves planes only");
        if (group[i] >= ngroups) return fail(CW_ERR_ARG, "field %u: group %u of %u", i, group[i], ngroups);
        for (uint32_t j = 0; j < i; j++)
            if (F[j] == F[i]) return fail(CW_ERR_ARG, "field %u appears twice", i);
        if (F[i]->coef_dirty && (st = build_coefficients(F[i]))) return st;
        if (!fp64 && (st = ensure_fp32(F[i]))) return st;
        memset(&stats[i], 0, sizeof(cw_relax_stats));
        stats[i].rel_residual = std::nan("");
    }
    // every field's own stream is idle before the graph reads its memory
    for (uint32_t i = 0; i < nf; i++) CW_HIP(hipStreamSynchronize(F[i]->stream));
    // Each group's shape; groups of one shape then form one device group (one
    // tile list and one step kernel a round), since a persistent step kernel
    // takes every CU and two of them would only run one after the other.
    std::vector<int> shape_of(ngroups, -1);
    for (uint32_t g = 0; g < ngroups; g++) {
        std::vector<cw_field*> members;
        for (uint32_t i = 0; i < nf; i++)
            if (group[i] == g) members.push_back(F[i]);
        if (members.empty()) continue;
        const int nm = (int)members.size();
        const int shape = (shapes && shapes[g] >= 0) ? shapes[g]
                          : fp64 ? default_plain_shape(0, members.data(), nm)
                                 : default_smooth2d_shape(members.data(), nm);
        if (shape >= (fp64 ? kNumPlainShapes : kNumSmoothShapes) || !camp_shape_supported(shape, fp64))
            return fail(CW_ERR_ARG,
                        "tile shape %d is not one the device-driven batch runs (FP64: 8 or 9; "
                        "FP32: 15 or 18 to 24)", shape);
        shape_of[g] = shape;
    }
    std::vector<int> distinct;
    for (uint32_t g = 0; g < ngroups; g++)
        if (shape_of[g] >= 0 && std::find(distinct.begin(), distinct.end(), shape_of[g]) == distinct.end())
            distinct.push_back(shape_of[g]);
    std::vector<std::unique_ptr<CampGroup>> gs;
    int longest_steps = 1;
    for (int shape : distinct) {
        std::vector<cw_field*> members;
        std::vector<uint32_t> which;
        for (uint32_t g = 0; g < ngroups; g++) {
            if (shape_of[g] != shape) continue;
            for (uint32_t i = 0; i < nf; i++)
                if (group[i] == g) { members.push_back(F[i]); which.push_back(i); }
        }
        auto G = std::make_unique<CampGroup>();
        Plan& pl = G->pl;
        G->stream = members[0]->stream;
        pl.stream = G->stream;
        pl.fp64 = fp64;
        pl.plain = true;
        pl.three_d = false;
        pl.lower_bound = false;
        const int nm = (int)members.size();
        G->shape = pl.shape = shape;
        G->s = pl.s_main = fp64 ? kPlainShapes[shape].s : kSmoothShapes[shape].s;
        for (int m = 0; m < nm; m++) {
            cw_field* f = members[m];
            cw_relax_stats& S = stats[which[m]];
            S.sweeps_per_pass = G->s;
            S.tile_shape = shape;
            if (!f->has_span) continue;  // nothing free: 0 sweeps, as relax reports it
            const double omega = pick_omega(o.omega, f);
            S.omega = omega;
            S.span_voxels = span_voxels(f);
            Prob pr;
            fill_prob_common(pr, f, omega);
            if (fp64) {
                CW_HIP(hipMemcpyAsync(f->c[f->cur ^ 1], f->c[f->cur], f->npad * sizeof(double),
                                      hipMemcpyDeviceToDevice, G->stream));
                pr.x[0] = f->c[f->cur];
                pr.x[1] = f->c[f->cur ^ 1];
                pr.k = f->k;
                pr.k2 = f->general ? f->k2 : nullptr;
                pr.c64 = f->c[f->cur];
                if (f->general) pl.general = true;
                pl.hp.push_back(pr);
            } else {
                Prob p64 = pr;
                p64.k = f->k;
                p64.k2 = f->general ? f->k2 : nullptr;
                p64.c64 = f->c[f->cur];
                p64.e0 = f->e[0];
                p64.e1 = f->e[1];
                p64.r32 = f->r32;
                p64.lb32 = f->has_sink ? f->lb : nullptr;
                Prob ps = pr;
                ps.x[0] = f->e[0];
                ps.x[1] = f->e[1];
                ps.k = f->r32;
                ps.k2 = nullptr;
                ps.lb = p64.lb32;
                if (ps.lb) pl.lower_bound = true;
                pl.hp.push_back(ps);
                G->hp64.push_back(p64);
            }
            Ctl c;
            init_ctl(c, ST_RUNNING);
            pl.hc.push_back(c);
            CampField cf;
            memset(&cf, 0, sizeof cf);
            cf.phase = PH_INIT;
            int longest = 2;
            for (int a = 0; a < 2; a++) longest = std::max(longest, f->span[2 * a + 1] - f->span[2 * a] + 1);
            cf.first = std::max(32, longest / 4);
            cf.hi = std::max(256, 4 * longest);
            cf.model_rate = model_log_rate(f, omega);
            G->hst.push_back(cf);
            G->idx.push_back((int)which[m]);
            // steps the model problem predicts: its sweeps to tol with a 30%
            // margin, plus the refinement's residual and correction steps
            const double need = cf.model_rate < 0.0 ? 1.3 * std::log(o.tol) / cf.model_rate : 1e4;
            const int steps = (int)std::min(1e7, std::ceil(need / G->s)) + (fp64 ? 2 : 40);
            longest_steps = std::max(longest_steps, steps);
            G->pred.push_back(steps);
        }
        const int np = (int)G->idx.size();
        pl.nprob = np;
        if (np == 0) continue;
        if ((st = plan_upload(pl))) return st;
        if (!fp64)
            for (int p = 0; p < np; p++) {
                Prob& q = G->hp64[p];
                const Prob& r = pl.hp[p];
                memcpy(q.sw_base, r.sw_base, sizeof q.sw_base);
                memcpy(q.sw_count, r.sw_count, sizeof q.sw_count);
                memcpy(q.ntx, r.ntx, sizeof q.ntx);
                memcpy(q.nty, r.nty, sizeof q.nty);
                q.el_base = r.el_base; q.el_count = r.el_count; q.elx = r.elx; q.ely = r.ely;
            }
        const size_t ntiles = pl.tp[TL_MAIN].size();
        auto up = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t b_st = up(np * sizeof(CampField)), b_p64 = up(np * sizeof(Prob)),
                     b_run = up(kCampCounterStride * 9 * sizeof(int)), b_list = up(std::max<size_t>(ntiles, 1) * sizeof(int));
        CW_HIP(hipMalloc((void**)&G->dev, b_st + b_p64 + b_run + b_list));
        G->dst = (CampField*)G->dev;
        G->dp64 = (Prob*)(G->dev + b_st);
        G->count = (int*)(G->dev + b_st + b_p64);
        G->list = (int*)(G->dev + b_st + b_p64 + b_run);
        CW_HIP(hipMemcpyAsync(G->dst, G->hst.data(), np * sizeof(CampField), hipMemcpyHostToDevice,
                              G->stream));
        if (!fp64)
            CW_HIP(hipMemcpyAsync(G->dp64, G->hp64.data(), np * sizeof(Prob), hipMemcpyHostToDevice,
                                  G->stream));
        CW_HIP(hipMemsetAsync(G->count, 0, b_run, G->stream));
        G->gen = fp64 ? pl.general : pl.lower_bound;
        if ((st = camp_grid(shape, fp64, G->gen, false, (long)ntiles, &G->grid))) return st;
        if ((st = camp_grid(shape, fp64, G->gen, true, (long)ntiles, &G->grid_tail))) return st;
        if (fp64) {
            // the verdict on the input, as the batched FP64 solve takes it:
            // resid64 and decide fix r2 and b2 (the tolerance's scale) and stop
            // a field that is already converged
            const int nt = (int)pl.tp[TL_EL].size();
            PassArgs a = make_args(pl, TL_EL, DECIDE_RESID, CRIT_RESID, o.tol, 0);
            hipLaunchKernelGGL(resid64, dim3(nt), dim3(256), 0, pl.stream, a, 0);
            CW_HIP(hipGetLastError());
            if ((st = launch_decide(pl, DECIDE_RESID, CRIT_RESID, o.tol, 0))) return st;
        }
        gs.push_back(std::move(G));
    }
    long steps_run = 0;
    // the groups as the transition kernels read them, and the shared step base
    struct Shared {
        unsigned char* dev = nullptr;
        ~Shared() { (void)hipFree(dev); }
    } shared;
    if (!gs.empty()) {
        hipStream_t s0 = gs[0]->stream;
        const size_t b_groups = (gs.size() * sizeof(CampGroupDev) + 255) / 256 * 256;
        CW_HIP(hipMalloc((void**)&shared.dev, b_groups + 256));
        CampGroupDev* dgroups = (CampGroupDev*)shared.dev;
        int* run = (int*)(shared.dev + b_groups);
        std::vector<CampGroupDev> hg(gs.size());
        int nfields_all = 0;
        for (size_t g = 0; g < gs.size(); g++) {
            gs[g]->run = run;
            hg[g].ca = camp_args(*gs[g], 0, o, check_every, fp64);
            hg[g].field_base = nfields_all;
            hg[g].nfields = (int)gs[g]->idx.size();
            nfields_all += hg[g].nfields;
        }
        CW_HIP(hipMemcpyAsync(dgroups, hg.data(), gs.size() * sizeof(CampGroupDev),
                              hipMemcpyHostToDevice, s0));
        CW_HIP(hipMemsetAsync(run, 0, 256, s0));
        // The set-up each group enqueued on its own stream comes first.
        std::vector<hipEvent_t> setup(gs.size(), nullptr);
        for (size_t g = 1; g < gs.size(); g++) {
            CW_HIP(hipEventCreateWithFlags(&setup[g], hipEventDisableTiming));
            const hipError_t e1 = hipEventRecord(setup[g], gs[g]->stream);
            const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(s0, setup[g], 0) : e1;
            (void)hipEventDestroy(setup[g]);
            if (e2 != hipSuccess) return fail(CW_ERR_DEVICE, "set-up join failed");
        }
        // One graph of kCampSteps rounds: each round the transition for every
        // group on the first group's stream, then the step launches forked onto
        // lanes of their own and joined back. A full-width FP64 step is split
        // over kCampSplit lanes (one queue cannot dispatch a workgroup per tile
        // fast enough); the FP64 tail form and FP32 take one lane per group.
        auto capture = [&](bool tail, hipGraphExec_t* out) -> cw_status {
            struct Lane {
                int g, part;
                hipStream_t stream;
                bool own;
            };
            std::vector<Lane> lanes;
            for (size_t g = 0; g < gs.size(); g++) {
                const int parts = (fp64 && !tail) ? kCampSplit : 1;
                for (int q = 0; q < parts; q++) lanes.push_back({(int)g, q, nullptr, false});
            }
            std::vector<hipEvent_t> ev(lanes.size(), nullptr);
            auto drop = [&]() {
                for (hipEvent_t e : ev)
                    if (e) (void)hipEventDestroy(e);
                for (Lane& ln : lanes)
                    if (ln.own && ln.stream) (void)hipStreamDestroy(ln.stream);
            };
            for (size_t i = 0; i < lanes.size(); i++) {
                Lane& ln = lanes[i];
                if (i == 0) ln.stream = s0;
                else if (ln.part == 0) ln.stream = gs[ln.g]->stream;
                else if (hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking) == hipSuccess) ln.own = true;
                else { ln.stream = nullptr; drop(); return fail(CW_ERR_DEVICE, "hipStreamCreate failed"); }
            }
            for (auto& e : ev)
                if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                    drop();
                    return fail(CW_ERR_DEVICE, "hipEventCreate failed");
                }
            hipGraph_t graph = nullptr;
            if (hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                drop();
                return fail(CW_ERR_DEVICE, "graph capture failed to begin");
            }
            cw_status cst = CW_OK;
            for (int j = 0; j < kCampSteps && cst == CW_OK; j++) {
                cst = camp_launch_transition(dgroups, (int)gs.size(), nfields_all, j, run, s0);
                if (cst == CW_OK && lanes.size() > 1 && hipEventRecord(ev[0], s0) != hipSuccess)
                    cst = fail(CW_ERR_DEVICE, "fork record failed");
                for (size_t i = 1; i < lanes.size() && cst == CW_OK; i++)
                    if (hipStreamWaitEvent(lanes[i].stream, ev[0], 0) != hipSuccess)
                        cst = fail(CW_ERR_DEVICE, "fork wait failed");
                for (size_t i = 0; i < lanes.size() && cst == CW_OK; i++) {
                    CampGroup& G = *gs[lanes[i].g];
                    CampArgs ca = camp_args(G, j, o, check_every, fp64);
                    const int grid = tail ? G.grid_tail : G.grid;
                    ca.slot_base = lanes[i].part * grid;
                    cst = camp_launch_step(G.shape, fp64, G.gen, tail, ca, grid, lanes[i].stream);
                }
                for (size_t i = 1; i < lanes.size() && cst == CW_OK; i++)
                    if (hipEventRecord(ev[i], lanes[i].stream) != hipSuccess ||
                        hipStreamWaitEvent(s0, ev[i], 0) != hipSuccess)
                        cst = fail(CW_ERR_DEVICE, "join failed");
            }
            if (cst == CW_OK) {
                hipLaunchKernelGGL(advance_run, dim3(1), dim3(64), 0, s0, run, kCampSteps, 0);
                if (hipGetLastError() != hipSuccess) cst = fail(CW_ERR_DEVICE, "advance_run failed");
            }
            const hipError_t ec = hipStreamEndCapture(s0, &graph);
            if (cst == CW_OK && ec != hipSuccess)
                cst = fail(CW_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(ec));
            if (cst == CW_OK && hipGraphInstantiate(out, graph, nullptr, nullptr, 0) != hipSuccess)
                cst = fail(CW_ERR_DEVICE, "graph instantiate failed");
            if (graph) (void)hipGraphDestroy(graph);
            drop();
            return cst;
        };
        // FP64: full-width rounds while the model problem predicts more tiles
        // running than twice the tail grid, tail rounds after that; FP32: one form.
        long head_steps = 0;
        if (fp64) {
            std::vector<std::pair<int, long>> ends;  // predicted last step, tiles
            long tail_grid = 0;
            for (auto& G : gs) {
                tail_grid += G->grid_tail;
                for (size_t p = 0; p < G->idx.size(); p++)
                    ends.push_back({G->pred[p], (long)G->pl.hp[p].sw_count[TL_MAIN]});
            }
            std::sort(ends.begin(), ends.end());
            long running = 0;
            for (auto& e : ends) running += e.second;
            for (auto& e : ends) {
                if (running < 2 * tail_grid) break;
                head_steps = e.first;
                running -= e.second;
            }
        }
        hipGraphExec_t full = nullptr, tailx = nullptr;
        struct Execs {
            hipGraphExec_t* a;
            hipGraphExec_t* b;
            ~Execs() {
                if (*a) (void)hipGraphExecDestroy(*a);
                if (*b) (void)hipGraphExecDestroy(*b);
            }
        } execs{&full, &tailx};
        const long head_chains = fp64 ? (head_steps + kCampSteps - 1) / kCampSteps : 0;
        if (head_chains > 0 && (st = capture(false, &full))) return st;
        if ((st = capture(fp64, &tailx))) return st;
        // launch the predicted number of chains, wait once; a field still
        // running after that costs one more round trip each time
        const long limit_steps =
            (long)(std::min<uint64_t>(o.max_sweeps, 0x7fffffff) + 2 * std::min<uint64_t>(o.max_outer, 1 << 20)) + 8;
        long chains = std::max(1L, (longest_steps + kCampSteps - 1) / kCampSteps - head_chains), more = 1;
        for (long c = 0; c < head_chains && st == CW_OK; c++)
            if (hipGraphLaunch(full, s0) != hipSuccess) st = fail(CW_ERR_DEVICE, "graph launch failed");
        steps_run += head_chains * kCampSteps;
        while (st == CW_OK) {
            for (long c = 0; c < chains && st == CW_OK; c++)
                if (hipGraphLaunch(tailx, s0) != hipSuccess) st = fail(CW_ERR_DEVICE, "graph launch failed");
            steps_run += chains * kCampSteps;
            if (st) break;
            for (auto& G : gs)
                if (hipMemcpyAsync(G->hst.data(), G->dst, G->hst.size() * sizeof(CampField),
                                   hipMemcpyDeviceToHost, s0) != hipSuccess)
                    st = fail(CW_ERR_DEVICE, "state read failed");
            if (st == CW_OK && hipStreamSynchronize(s0) != hipSuccess)
                st = fail(CW_ERR_DEVICE, "graph run failed: %s", hipGetErrorString(hipGetLastError()));
            if (st) break;
            bool running = false;
            for (auto& G : gs)
                for (const CampField& cf : G->hst) running |= camp_active(cf.phase) || cf.phase == PH_INIT;
            if (!running || steps_run >= limit_steps) break;
            chains = more;
            more = std::min<long>(2 * more, 16);
        }
        if (st) return st;
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    cw_status worst = CW_OK;
    for (auto& G : gs)
        for (size_t p = 0; p < G->idx.size(); p++) {
            const CampField& cf = G->hst[p];
            cw_field* f = F[G->idx[p]];
            cw_relax_stats& S = stats[G->idx[p]];
            S.sweeps = (uint64_t)cf.total;
            S.outer_iters = (uint64_t)cf.outer;
            S.passes = (uint64_t)steps_run;
            if (cf.b2 > 0.0) S.rel_residual = std::sqrt(cf.r2 / cf.b2);
            else if (cf.phase == PH_DONE) S.rel_residual = std::sqrt(cf.r2);
            S.status = cf.phase == PH_DONE ? CW_OK : cf.nonfinite ? CW_ERR_NONFINITE : CW_ERR_NOT_CONVERGED;
            if (fp64 && cf.phase == PH_DONE && (cf.parity & 1)) f->cur ^= 1;
            if (S.status != CW_OK && worst == CW_OK) worst = (cw_status)S.status;
        }
    for (uint32_t i = 0; i < nf; i++) stats[i].seconds = secs;
    if (worst == CW_ERR_NOT_CONVERGED)
        return fail(CW_ERR_NOT_CONVERGED, "diffusion did not settle within %llu sweeps at tolerance %g",
                    (unsigned long long)o.max_sweeps, o.tol);
    if (worst == CW_ERR_NONFINITE) return fail(CW_ERR_NONFINITE, "the residual is not finite");
    return CW_OK;
}

cw_status campaign_solve(cw_field* const* F, uint32_t nf, const uint32_t* group,
                         const int32_t* shapes, uint32_t ngroups, const cw_relax_opts* opts,
                         cw_relax_stats* stats) {
    try {
        return campaign_impl(F, nf, group, shapes, ngroups, opts, stats);
    } catch (const std::bad_alloc&) {
        return fail(CW_ERR_DEVICE, "out of host memory");
    } catch (const std::exception& e) {
        return fail(CW_ERR_DEVICE, "host error: %s", e.what());
    }
}

// Adjoint of the steady state (new; the reference fits parameters by hand).
// Loss L = 0.5 * sum over all voxels of (c - target)^2. With A c = b on the
// free voxels and lambda = A^-1 (c - target) (A symmetric, lambda = 0 on held
// voxels), the gradients are
//   dL/dD   = sum_i lambda_i sum_{real nbrs j} (c_j - c_i)   (D in px^2 per s)
//   dL/du_k = -sum of lambda over the voxels taking up at cell k's rate
//   dL/dg   = D sum_i lambda_i #held_nbrs(i) + sum_held (c_h - target_h)
// valid while no free voxel is floored. lambda is solved in FP64 by the
// pre-scaled red-black smoother, tested on its unprojected relative residual.
cw_status field_adjoint(cw_field* f, const double* target, const uint32_t* grid,
                        uint32_t ncells, const cw_relax_opts* opts, double* dl_du,
                        cw_adjoint_result* res) {
    auto t0 = std::chrono::steady_clock::now();
    if (!f || !target || !res || !opts) return fail(CW_ERR_ARG, "null argument");
    if (ncells && dl_du && !grid) return fail(CW_ERR_ARG, "per-cell gradients need the lattice");
    Opts o;
    cw_status st = check_opts(opts, &o);
    if (st) return st;
    if (f->coef_dirty && (st = build_coefficients(f))) return st;
    memset(res, 0, sizeof *res);
    res->rel_residual = std::nan("");
    if (dl_du)
        for (uint32_t k = 0; k < ncells; k++) dl_du[k] = 0.0;
    // forward state on the host: loss, the held voxels' own term, the floor
    std::vector<double> c(f->n);
    if ((st = download_f64(f, f->c[f->cur], c.data()))) return st;
    std::vector<uint8_t> held(f->n);
    std::vector<double> uptake;
    if ((st = download_held(f, held.data()))) return st;
    if (dl_du && ncells) {
        uptake.resize(f->n);
        if ((st = download_f64(f, f->up, uptake.data()))) return st;
    }
    double loss = 0.0, dg_held = 0.0;
    size_t floored = 0;
    for (size_t i = 0; i < f->n; i++) {
        const double r = c[i] - target[i];
        loss += r * r;
        if (held[i]) dg_held += r;
        else if (!(c[i] > 0.0)) floored++;
    }
    res->loss = 0.5 * loss;
    res->dl_dg = dg_held;
    if (floored)
        return fail(CW_ERR_ARG,
                    "the floor is active at %zu free voxels: the linear adjoint does not apply",
                    floored);
    if (!f->has_span) {
        res->status = CW_OK;
        return CW_OK;
    }
    const size_t bytes = f->npad * sizeof(double);
    double *tgt = nullptr, *rhs = nullptr, *beta = nullptr, *lam[2] = {nullptr, nullptr};
    auto cleanup = [&]() {
        (void)hipFree(tgt); (void)hipFree(rhs); (void)hipFree(beta);
        (void)hipFree(lam[0]); (void)hipFree(lam[1]);
    };
    auto bail = [&](cw_status s2) { cleanup(); return s2; };
    if (hipMalloc((void**)&tgt, bytes) != hipSuccess || hipMalloc((void**)&rhs, bytes) != hipSuccess ||
        hipMalloc((void**)&beta, bytes) != hipSuccess ||
        hipMalloc((void**)&lam[0], bytes) != hipSuccess ||
        hipMalloc((void**)&lam[1], bytes) != hipSuccess)
        return bail(fail(CW_ERR_DEVICE, "adjoint buffers: out of device memory"));
    // Outside the swept box every voxel is held: NaN coefficients, zero lambda.
    if (hipMemsetAsync(rhs, 0xFF, bytes, f->stream) != hipSuccess ||
        hipMemsetAsync(beta, 0xFF, bytes, f->stream) != hipSuccess ||
        hipMemsetAsync(lam[0], 0, bytes, f->stream) != hipSuccess ||
        hipMemsetAsync(lam[1], 0, bytes, f->stream) != hipSuccess ||
        hipMemsetAsync(tgt, 0, bytes, f->stream) != hipSuccess ||
        upload_flat(f, tgt, target, 8) != CW_OK)
        return bail(fail(CW_ERR_DEVICE, "adjoint target upload failed"));
    const bool three_d = f->D > 1;
    const double omega = pick_omega(o.omega, f);
    // smoother plan (lambda, pre-scaled beta) and residual plan (lambda, rhs)
    Plan sm, rs;
    sm.stream = rs.stream = f->stream;
    sm.fp64 = rs.fp64 = true;
    sm.plain = false;
    sm.three_d = rs.three_d = three_d;
    sm.s_main = three_d ? 1 : 2;
    sm.shape = three_d ? 0 : 1;
    rs.s_main = 1;
    Prob pr;
    fill_prob_common(pr, f, omega);
    pr.k = f->k;
    pr.c64 = f->c[f->cur];
    Ctl c0;
    init_ctl(c0, ST_RUNNING);
    sm.hp.push_back(pr); sm.hc.push_back(c0); sm.nprob = 1;
    rs.hp.push_back(pr); rs.hc.push_back(c0); rs.nprob = 1;
    if ((st = plan_upload(sm)) || (st = plan_upload(rs))) return bail(st);
    {
        const int nt = (int)sm.tp[TL_EL].size();
        PassArgs a = make_args(sm, TL_EL, MODE_SWEEP, CRIT_NONE, 0.0, 0);
        hipLaunchKernelGGL(adjoint_setup, dim3(nt), dim3(256), 0, f->stream, a,
                           (const double*)f->c[f->cur], (const double*)tgt, rhs, beta, lam[0],
                           lam[1]);
        if (hipGetLastError() != hipSuccess) return bail(fail(CW_ERR_DEVICE, "adjoint_setup"));
    }
    int cur = 0;
    auto set_state = [&]() -> cw_status {
        sm.hp[0].x[0] = lam[cur];
        sm.hp[0].x[1] = sm.in_place() ? lam[cur] : lam[cur ^ 1];
        sm.hp[0].k = beta;
        rs.hp[0].k = rhs;
        rs.hp[0].c64 = lam[cur];
        CW_HIP(hipMemcpyAsync(sm.dp, sm.hp.data(), sizeof(Prob), hipMemcpyHostToDevice, f->stream));
        CW_HIP(hipMemcpyAsync(rs.dp, rs.hp.data(), sizeof(Prob), hipMemcpyHostToDevice, f->stream));
        return CW_OK;
    };
    if ((st = set_state())) return bail(st);
    int longest = 2;
    for (int ax = 0; ax < 3; ax++) longest = std::max(longest, f->span[2 * ax + 1] - f->span[2 * ax] + 1);
    std::vector<double> hist;
    std::vector<int> at;
    uint64_t total = 0;
    bool converged = false;
    while (true) {
        run_reset(rs);
        init_ctl(rs.hc[0], ST_RUNNING);
        if ((st = write_ctl(rs)) || (st = run_resid(rs, CRIT_RESID, o.tol, 0, 0))) return bail(st);
        const Ctl& rc = rs.hc[0];
        const double rel = rc.b2 > 0.0 ? std::sqrt(rc.r2 / rc.b2) : std::sqrt(rc.r2);
        res->rel_residual = rel;
        if (rc.nonfinite) return bail(fail(CW_ERR_NONFINITE, "the adjoint residual is not finite"));
        if (rc.status == ST_DONE) { converged = true; break; }
        if (total >= o.max_sweeps) break;
        hist.push_back(rel);
        at.push_back((int)total);
        int k = next_segment(hist, at, o.tol, 1e-300, std::max(32, longest / 2), 16,
                             std::max(256, 8 * longest), sm.s_main);
        k = (int)std::min<uint64_t>((uint64_t)k, o.max_sweeps - total);
        init_ctl(sm.hc[0], ST_RUNNING);
        sm.hc[0].budget = k;
        if ((st = write_ctl(sm))) return bail(st);
        run_reset(sm);
        LoopResult lr;
        if ((st = run_passes(sm, CRIT_NONE, 0.0, (uint64_t)k, 0x7fffffff, 1, &lr))) return bail(st);
        total += (uint64_t)sm.hc[0].sweeps;
        if (!sm.in_place() && (sm.hc[0].parity & 1)) cur ^= 1;
        if ((st = set_state())) return bail(st);
    }
    res->sweeps = total;
    // gradient partials over the swept box
    {
        const int nt = (int)rs.tp[TL_EL].size();
        PassArgs a = make_args(rs, TL_EL, MODE_SWEEP, CRIT_NONE, 0.0, 0);
        hipLaunchKernelGGL(adjoint_grads, dim3(nt), dim3(256), 0, f->stream, a,
                           (const double*)f->c[f->cur], (const double*)lam[cur],
                           (const double*)f->k);
        if (hipGetLastError() != hipSuccess) return bail(fail(CW_ERR_DEVICE, "adjoint_grads"));
        std::vector<double> part(2 * (size_t)nt);
        CW_HIP(hipMemcpyAsync(part.data(), rs.partials, part.size() * sizeof(double),
                              hipMemcpyDeviceToHost, f->stream));
        CW_HIP(hipStreamSynchronize(f->stream));
        double gd = 0.0, gg = 0.0;
        for (int t = 0; t < nt; t++) { gd += part[2 * t]; gg += part[2 * t + 1]; }
        res->dl_dd = gd;
        res->dl_dg += gg;
    }
    if (dl_du && ncells) {
        // CSR of cell -> voxels taking up at that cell's rate (pitched indices)
        std::vector<long long> off(ncells + 1, 0);
        const size_t W = f->W, H = f->H;
        for (size_t i = 0; i < f->n; i++) {
            const uint32_t id = grid[i];
            if (held[i] || !(uptake[i] > 0.0) || id == 0 || id == kObstacle || id >= ncells)
                continue;
            off[id + 1]++;
        }
        for (uint32_t k = 0; k < ncells; k++) off[k + 1] += off[k];
        std::vector<long long> idx((size_t)off[ncells]);
        std::vector<long long> fillp(off.begin(), off.end() - 1);
        for (size_t i = 0; i < f->n; i++) {
            const uint32_t id = grid[i];
            if (held[i] || !(uptake[i] > 0.0) || id == 0 || id == kObstacle || id >= ncells)
                continue;
            const size_t x = i % W, y = (i / W) % H, z = i / (W * H);
            idx[(size_t)fillp[id]++] = (long long)((z * H + y) * (size_t)f->P + x);
        }
        long long *doff = nullptr, *didx = nullptr;
        double* dout = nullptr;
        bool ok = hipMalloc((void**)&doff, off.size() * sizeof(long long)) == hipSuccess &&
                  hipMalloc((void**)&didx, std::max<size_t>(1, idx.size()) * sizeof(long long)) == hipSuccess &&
                  hipMalloc((void**)&dout, ncells * sizeof(double)) == hipSuccess;
        if (ok) {
            ok = hipMemcpyAsync(doff, off.data(), off.size() * sizeof(long long),
                                hipMemcpyHostToDevice, f->stream) == hipSuccess &&
                 (idx.empty() || hipMemcpyAsync(didx, idx.data(), idx.size() * sizeof(long long),
                                                hipMemcpyHostToDevice, f->stream) == hipSuccess);
            if (ok) {
                hipLaunchKernelGGL(cell_sums, dim3((ncells + 255) / 256), dim3(256), 0, f->stream,
                                   (const double*)lam[cur], (const long long*)doff,
                                   (const long long*)didx, (int)ncells, dout);
                ok = hipGetLastError() == hipSuccess &&
                     hipMemcpyAsync(dl_du, dout, ncells * sizeof(double), hipMemcpyDeviceToHost,
                                    f->stream) == hipSuccess &&
                     hipStreamSynchronize(f->stream) == hipSuccess;
            }
        }
        (void)hipFree(doff); (void)hipFree(didx); (void)hipFree(dout);
        if (!ok) return bail(fail(CW_ERR_DEVICE, "per-cell gradient sums failed"));
    }
    cleanup();
    res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    res->status = converged ? CW_OK : CW_ERR_NOT_CONVERGED;
    if (!converged)
        return fail(CW_ERR_NOT_CONVERGED, "the adjoint did not reach %g within %llu sweeps",
                    o.tol, (unsigned long long)o.max_sweeps);
    return CW_OK;
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free relax (cw_field_relax_noalloc): the launch
// plan of one field (problems, control blocks, tile tables, residual
// partials) lives here instead of in allocated device memory. Every GPU the
// library runs on holds one copy until the process ends, and every call on
// that GPU shares it: two such calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = (size_t)4 << 20;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_addr[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// The device address of `scratch` on device dev, looked up once per device
// with that device current. Host code uses only this address: the array's
// own name is a host-side stand-in, not device memory.
cw_status scratch_on(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(CW_ERR_DEVICE, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        int cur = 0;
        CW_HIP(hipGetDevice(&cur));
        if (cur != dev) CW_HIP(hipSetDevice(dev));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(CW_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                        hipGetErrorString(e));
        g_scratch_addr[dev] = (unsigned char*)p;
    }
    *out = g_scratch_addr[dev];
    return CW_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

void cw_relax_opts_default(cw_relax_opts* o) {
    if (!o) return;
    memset(o, 0, sizeof *o);
    o->omega = 0.0;
    o->criterion = CW_CRIT_MAX_CHANGE;
    o->tolerance = 1e-8;
    o->max_sweeps = 2000000;
    o->precision = CW_PREC_FP64;
    o->sweeps_per_pass = 0;
    o->check_every = 0;
    o->inner_tolerance = 0.0;
    o->max_outer = 0;
    o->tile_shape = -1;
    o->use_graph = 0;
}

const char* cw_last_error(void) { return last_error_text(); }

const char* cw_version(void) {
    std::string& v = version_text();
    if (v.empty()) {
        v = "cwfield 0.2.0 gfx950";
        if (device_init() == CW_OK) v += " on " + device_name();
    }
    return v.c_str();
}

cw_status cw_field_create(uint32_t w, uint32_t h, uint32_t d, double diff, double decay,
                          double initial, cw_field** out) {
    if (!out) return fail(CW_ERR_ARG, "null output handle");
    *out = nullptr;
    if (w == 0 || h == 0 || d == 0)
        return fail(CW_ERR_CONFIG, "field dimensions must be positive, got %ux%ux%u", w, h, d);
    if (!std::isfinite(diff) || !std::isfinite(decay) || diff < 0.0 || decay < 0.0)
        return fail(CW_ERR_CONFIG,
                    "diffusion and decay must be finite and not negative, got D = %g, decay = %g",
                    diff, decay);
    if ((uint64_t)w * h * d > (1ull << 34)) return fail(CW_ERR_CONFIG, "field too large");
    cw_status st = device_init();
    if (st) return st;
    cw_field* f = new cw_field();
    if (hipGetDevice(&f->device) != hipSuccess) {
        delete f;
        return fail(CW_ERR_DEVICE, "hipGetDevice failed");
    }
    f->W = w; f->H = h; f->D = d;
    f->P = (int)((w + 63) / 64 * 64);
    f->n = (size_t)w * h * d;
    f->npad = (size_t)f->P * h * d;
    f->diff = diff; f->decay = decay;
    auto bail = [&](cw_status s) { cw_field_destroy(f); return s; };
    if (hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(CW_ERR_DEVICE, "hipStreamCreate failed"));
    if (hipHostMalloc((void**)&f->host_small, kHostSmallInts * sizeof(int), hipHostMallocDefault) !=
        hipSuccess) {
        f->host_small = nullptr;
        return bail(fail(CW_ERR_DEVICE, "hipHostMalloc of %d ints failed", kHostSmallInts));
    }
    const size_t b8 = f->npad * sizeof(double);
    if ((st = alloc_zero((void**)&f->c[0], b8, f->stream, 0))) return bail(st);
    if ((st = alloc_zero((void**)&f->c[1], b8, f->stream, 0))) return bail(st);
    if ((st = alloc_zero((void**)&f->k, b8, f->stream, 0xFF))) return bail(st);
    if ((st = alloc_zero((void**)&f->up, b8, f->stream, 0))) return bail(st);
    if ((st = alloc_zero((void**)&f->pr, b8, f->stream, 0))) return bail(st);
    if ((st = alloc_zero((void**)&f->held, f->npad, f->stream, 1))) return bail(st);
    if ((st = alloc_zero((void**)&f->scan, (kScanInts + kRedBlocks * kRedSlot) * sizeof(int),
                         f->stream, 0)))
        return bail(st);
    // nothing held inside the box; the padding stays held
    CW_HIP(hipMemset2DAsync(f->held, f->P, 0, f->W, (size_t)f->H * f->D, f->stream));
    {
        std::vector<double> init(f->n, initial);
        if ((st = upload_flat(f, f->c[0], init.data(), 8))) return bail(st);
    }
    f->coef_dirty = true;
    if ((st = build_coefficients(f))) return bail(st);
    *out = f;
    return CW_OK;
}

cw_status cw_field_destroy(cw_field* f) {
    if (!f) return CW_OK;
    if (f->stream) hipStreamSynchronize(f->stream);
    delete (PlanCache*)f->plans;
    hipFree(f->c[0]); hipFree(f->c[1]); hipFree(f->k); hipFree(f->k2);
    hipFree(f->up); hipFree(f->pr); hipFree(f->held); hipFree(f->scan); hipFree(f->stage);
    hipFree(f->e[0]); hipFree(f->e[1]); hipFree(f->r32); hipFree(f->lb);
    hipFree(f->mask8[0]); hipFree(f->mask8[1]);
    if (f->host_small) hipHostFree(f->host_small);
    if (f->stream) hipStreamDestroy(f->stream);
    delete f;
    return CW_OK;
}

cw_status cw_field_set_source(cw_field* f, double source) {
    if (!f) return fail(CW_ERR_ARG, "null field");
    f->source = source;
    f->coef_dirty = true;
    return CW_OK;
}

cw_status cw_field_upload(cw_field* f, const double* values, const uint8_t* held,
                          const double* uptake, const double* production) {
    if (!f) return fail(CW_ERR_ARG, "null field");
    cw_status st;
    if (held) {
        std::vector<uint8_t> h(f->n);
        for (size_t i = 0; i < f->n; i++) h[i] = held[i] ? 1 : 0;
        if ((st = upload_flat(f, f->held, h.data(), 1))) return st;
        f->coef_dirty = true;
    }
    // The setters clamp at zero (pde.rs:146-182); NaN clamps to zero as
    // f64::max(NaN, 0.0) does.
    std::vector<double> tmp;
    for (int which = 0; which < 2; which++) {
        const double* src = which ? production : uptake;
        if (!src) continue;
        tmp.resize(f->n);
        for (size_t i = 0; i < f->n; i++) tmp[i] = src[i] > 0.0 ? src[i] : 0.0;
        if ((st = upload_flat(f, which ? f->pr : f->up, tmp.data(), 8))) return st;
        f->coef_dirty = true;
    }
    if (values && (st = upload_flat(f, f->c[f->cur], values, 8))) return st;
    if (f->coef_dirty) return build_coefficients(f);
    return CW_OK;
}

cw_status cw_field_download(const cw_field* f, double* values) {
    if (!f || !values) return fail(CW_ERR_ARG, "null argument");
    return download_flat(f, f->c[f->cur], values, 8);
}

cw_status cw_field_relax(cw_field* f, double tolerance, uint64_t max_sweeps, uint64_t* sweeps) {
    if (!f) return fail(CW_ERR_ARG, "null field");
    cw_relax_opts o;
    cw_relax_opts_default(&o);
    o.tolerance = tolerance;
    o.max_sweeps = max_sweeps;
    // Steady chunks replay from a graph the field keeps between calls: the
    // warm-started relax of every lattice step launches once per 16 sweeps
    // instead of once per sweep. Same sweeps and bits as eager launches.
    o.use_graph = 1;
    cw_relax_stats s{};
    cw_status st = solve(&f, 1, &o, &s);
    if (sweeps) *sweeps = s.sweeps;
    return st;
}

cw_status cw_field_relax_noalloc(cw_field* f, double tolerance, uint64_t max_sweeps,
                                 uint64_t* sweeps) {
    if (!f) return fail(CW_ERR_ARG, "null field");
    unsigned char* sc = nullptr;
    cw_status st = scratch_on(f->device, &sc);
    if (st) return st;
    // cw_field_relax's options, with eager launches: instantiating a graph
    // allocates
    cw_relax_opts o;
    cw_relax_opts_default(&o);
    o.tolerance = tolerance;
    o.max_sweeps = max_sweeps;
    o.use_graph = 0;
    cw_relax_stats s{};
    st = solve(&f, 1, &o, &s, 0, nullptr, sc, kNoAllocScratchBytes);
    if (sweeps) *sweeps = s.sweeps;
    return st;
}

cw_status cw_field_relax_ex(cw_field* f, const cw_relax_opts* opts, cw_relax_stats* stats) {
    if (!f) return fail(CW_ERR_ARG, "null field");
    return solve(&f, 1, opts, stats);
}

cw_status cw_batch_relax(cw_field* const* fields, uint32_t count, const cw_relax_opts* opts,
                         cw_relax_stats* stats) {
    return solve(fields, count, opts, stats);
}

cw_status cw_campaign_relax(cw_field* const* fields, uint32_t count, const uint32_t* group,
                            const int32_t* shapes, uint32_t ngroups, const cw_relax_opts* opts,
                            cw_relax_stats* stats) {
    return campaign_solve(fields, count, group, shapes, ngroups, opts, stats);
}

cw_status cw_batch_time_passes(cw_field* const* fields, uint32_t count, const cw_relax_opts* opts,
                               uint32_t passes, double* ms_per_pass) {
    if (!fields || count == 0 || !opts || !ms_per_pass || passes == 0)
        return fail(CW_ERR_ARG, "bad argument");
    return solve(fields, count, opts, nullptr, (long)passes, ms_per_pass);
}

cw_status cw_autotune(cw_field* const* fields, uint32_t count, const cw_relax_opts* opts,
                      uint32_t launches, int32_t* shapes, double* us_per_sweep, uint32_t cap,
                      uint32_t* n_out) {
    if (!fields || count == 0 || !opts || !shapes || !us_per_sweep || !n_out || launches == 0)
        return fail(CW_ERR_ARG, "bad argument");
    const bool three_d = fields[0]->D > 1;
    const bool plain = opts->precision == CW_PREC_FP64;
    const int n = three_d ? kNum3DShapes : plain ? kNumPlainShapes : kNumSmoothShapes;
    std::vector<std::pair<double, int>> res;
    for (int sh = 0; sh < n; sh++) {
        if (three_d && plain && two_sweep(sh)) continue;  // FP32 smoothing only
        cw_relax_opts o = *opts;
        o.tile_shape = sh;
        o.sweeps_per_pass = 0;
        o.use_graph = 0;
        double ms = 0.0;
        cw_status st = solve(fields, count, &o, nullptr, (long)launches, &ms);
        if (st) return st;
        const int S = three_d ? (two_sweep(sh) ? 2 : 1) : plain ? kPlainShapes[sh].s : kSmoothShapes[sh].s;
        res.push_back({1e3 * ms / S, sh});
    }
    std::sort(res.begin(), res.end());
    uint32_t k = 0;
    for (; k < res.size() && k < cap; k++) {
        shapes[k] = res[k].second;
        us_per_sweep[k] = res[k].first;
    }
    *n_out = k;
    return CW_OK;
}

cw_status cw_field_fill(cw_field* f, double value) {
    if (!f) return fail(CW_ERR_ARG, "null field");
    const size_t rows = (size_t)f->H * f->D;
    const int blocks = (int)std::min<size_t>((rows * f->W + 255) / 256, 65536);
    hipLaunchKernelGGL(fill_cells, dim3(blocks), dim3(256), 0, f->stream, f->c[f->cur],
                       (int)f->W, (long)rows, f->P, value);
    CW_HIP(hipGetLastError());
    CW_HIP(hipStreamSynchronize(f->stream));
    return CW_OK;
}

cw_status cw_field_adjoint(cw_field* f, const double* target, const uint32_t* grid,
                           uint32_t ncells, const cw_relax_opts* opts, double* dl_du,
                           cw_adjoint_result* result) {
    return field_adjoint(f, target, grid, ncells, opts, dl_du, result);
}

cw_status cw_field_residual(cw_field* f, double* rel) {
    if (!f || !rel) return fail(CW_ERR_ARG, "null argument");
    cw_status st;
    if (f->coef_dirty && (st = build_coefficients(f))) return st;
    if (!f->has_span) { *rel = 0.0; return CW_OK; }
    Plan pl;
    pl.stream = f->stream;
    pl.fp64 = true;
    pl.three_d = f->D > 1;
    pl.s_main = 1;
    Prob pr;
    fill_prob_common(pr, f, 1.0);
    pr.k = f->k;
    pr.k2 = f->general ? f->k2 : nullptr;
    pr.c64 = f->c[f->cur];
    pl.hp.push_back(pr);
    Ctl c;
    init_ctl(c, ST_RUNNING);
    pl.hc.push_back(c);
    pl.nprob = 1;
    if ((st = plan_upload(pl))) return st;
    run_reset(pl);
    if ((st = run_resid(pl, CRIT_NONE, 0.0, 0))) return st;
    const Ctl& r = pl.hc[0];
    *rel = r.b2 > 0.0 ? std::sqrt(r.r2 / r.b2) : std::sqrt(r.r2);
    return CW_OK;
}

}  // extern "C"
