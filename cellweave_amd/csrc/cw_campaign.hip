// Kernels of the device-driven batch of planes (cw_campaign.hpp): the
// transition between steps and the persistent step kernel, compiled for the
// plane shapes a campaign's size classes use.
#include "cw_campaign.hpp"
#include "cw_shapes.hpp"

#include <algorithm>

using namespace cw;

namespace cw {

// Length of a field's next FP32 smoothing segment, from its residual history:
// the sweeps the observed rate needs to reach the target, asked for at most a
// `per_segment` reduction (FP32 cannot carry a correction much further), 10%
// margin, capped by the model problem's rate; doubled while the residual has
// not started to fall; within [lo, hi] and a multiple of s. The same rule as
// the host-driven refinement's next_segment.
__device__ inline int camp_segment(const CampField& f, double tol, double per_segment, int lo,
                                   int s) {
    int k = f.first;
    if (f.nhist >= 2) {
        const double r0 = f.hist[0], r1 = f.hist[1];
        const int last = f.hist_at[1] - f.hist_at[0];
        const double goal = fmax(tol / r1, per_segment);
        if (r1 > 0.0 && r0 > r1 && last > 0) {
            const double rate = log(r1 / r0) / (double)last;
            const double need = fmin(log(goal) / rate, 1e9);
            k = need > 0.0 ? (int)ceil(1.1 * need) + s : lo;
        } else {
            k = 2 * (last > f.first ? last : f.first);
        }
        if (f.model_rate < 0.0 && goal < 1.0) {
            const double cap = fmin(ceil(1.5 * log(goal) / f.model_rate), 1e9);
            k = min(k, (int)cap + f.first / 2);
        }
    }
    k = max(lo, min(f.hi, k));
    return (k + s - 1) / s * s;
}

// Phase of field p in step L from its phase in step L - 1 and, where that
// step left residual sums, from them. One thread per field.
__device__ inline void camp_next(const CampArgs& ca, const PassArgs& a, CampField& f, int p,
                                 int L, bool have, double r2, double b2) {
    const int s = a.s;
    const int prev = f.phase;
    if (prev == PH_DONE || prev == PH_MAXED) return;
    if (ca.fp64) {
        if (prev == PH_INIT) {
            // the verdict resid64 + decide took on the input
            const Ctl& c = a.ctl[p];
            f.r2 = c.r2;
            f.b2 = c.b2;
            f.total = 0;
            f.parity = 0;
            if (c.status == ST_DONE) { f.phase = PH_DONE; return; }
            if (c.status != ST_RUNNING || ca.max_sweeps <= 0) {
                f.nonfinite = c.nonfinite;
                f.phase = PH_MAXED;
                return;
            }
            f.phase = PH_SWEEP;
            return;
        }
        if (have) {
            // the test of step L - 1's input: that input is the answer
            if (!(r2 == r2) || r2 > 1.7e308) {
                f.r2 = r2; f.nonfinite = 1; f.parity = (L - 1) & 1; f.total = (L - 1) * s;
                f.phase = PH_MAXED;
                return;
            }
            if (r2 <= ca.tol * ca.tol * (f.b2 > 0.0 ? f.b2 : 1.0)) {
                f.r2 = r2; f.parity = (L - 1) & 1; f.total = (L - 1) * s;
                f.phase = PH_DONE;
                return;
            }
        }
        // a step that would run past max_sweeps is not started
        if ((long long)(L + 1) * s > ca.max_sweeps) {
            f.parity = L & 1; f.total = L * s;
            f.phase = PH_MAXED;
            return;
        }
        f.phase = PH_SWEEP;
        return;
    }
    switch (prev) {
    case PH_INIT:
    case PH_CORRECT:
        f.phase = PH_RESID;
        return;
    case PH_SMOOTH:
        f.seg_done += s;
        f.total += s;
        f.phase = f.seg_done >= f.budget ? PH_CORRECT : PH_SMOOTH;
        return;
    case PH_RESID: {
        f.r2 = r2;
        f.b2 = b2;
        if (!(r2 == r2) || r2 > 1.7e308) { f.nonfinite = 1; f.phase = PH_MAXED; return; }
        if (r2 <= ca.tol * ca.tol * (b2 > 0.0 ? b2 : 1.0)) { f.phase = PH_DONE; return; }
        if (f.total >= ca.max_sweeps || f.outer >= ca.max_outer) { f.phase = PH_MAXED; return; }
        const double rel = b2 > 0.0 ? sqrt(r2 / b2) : sqrt(r2);
        if (f.nhist == 2) {
            f.hist[0] = f.hist[1];
            f.hist_at[0] = f.hist_at[1];
        }
        const int h = f.nhist < 2 ? f.nhist : 1;
        f.hist[h] = rel;
        f.hist_at[h] = f.total;
        if (f.nhist < 2) f.nhist++;
        int k = camp_segment(f, ca.tol, ca.inner_tol, ca.lo, s);
        // a segment never runs past max_sweeps: it is cut to whole steps
        const int rem = ca.max_sweeps - f.total;
        if (k > rem) k = rem / s * s;
        if (k < s) { f.phase = PH_MAXED; return; }
        f.budget = k;
        f.seg_done = 0;
        f.seg_start = L;
        f.outer++;
        f.phase = PH_SMOOTH;
        return;
    }
    default:
        return;
    }
}

constexpr int kCampThreads = 1024;

// Between steps L - 1 and L, part 1: every field's phase for step L, one wave
// per field over every group (a field's tile partials summed in the fixed
// order of problem_sums, so FP64 stops where the host-driven batch stops).
__global__ __launch_bounds__(kCampThreads) void camp_verdicts(const CampGroupDev* groups,
                                                              int ngroups, int step,
                                                              const int* run) {
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * (kCampThreads / 64) + (threadIdx.x >> 6);
    int g = 0;
    while (g < ngroups && w >= groups[g].field_base + groups[g].nfields) g++;
    if (g >= ngroups) return;
    const CampArgs& ca = groups[g].ca;
    PassArgs a = ca.pa;
    a.pass = step + run[0];
    const int L = a.pass;
    const int p = w - groups[g].field_base;
    CampField& f = ca.st[p];
    const Prob& pr = a.probs[p];
    const int prev = f.phase;
    const bool have = ca.fp64 ? (prev == PH_SWEEP && L > 0 && (L - 1) % ca.check_every == 0)
                              : prev == PH_RESID;
    double r2 = 0.0, b2 = 0.0;
    if (have)
        problem_sums(a.partials + ((L - 1) & 1) * a.part_stride, pr.sw_base[0], pr.sw_count[0],
                     &r2, &b2);
    if (lane == 0) camp_next(ca, a, f, p, L, have, r2, b2);
}

// Part 2: the list of the tiles that work in step L, one workgroup per group,
// field by field.
__global__ __launch_bounds__(kCampThreads) void camp_lists(const CampGroupDev* groups) {
    const CampArgs& ca = groups[blockIdx.x].ca;
    const PassArgs& a = ca.pa;
    constexpr int NWV = kCampThreads / 64;
    __shared__ int s_off[kCampThreads], s_cnt[kCampThreads], s_base[kCampThreads];
    __shared__ int s_wave[NWV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int np = a.nprob;
    const int per = (np + blockDim.x - 1) / blockDim.x;
    const int p0 = min(np, (int)threadIdx.x * per), p1 = min(np, p0 + per);
    // each thread's working tiles (its problems' loads issued together)
    int mine = 0, base1 = 0;
    for (int p = p0; p < p1; p++) {
        const int ph = ca.st[p].phase;
        const int cnt = a.probs[p].sw_count[0], base = a.probs[p].sw_base[0];
        if (camp_active(ph)) mine += cnt;
        base1 = base;
    }
    // exclusive offsets: a shuffle scan in each wave, then over the waves
    int x = mine;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    if (wave == 0) {
        int w = lane < NWV ? s_wave[lane] : 0;
#pragma unroll
        for (int d = 1; d < NWV; d <<= 1) {
            const int y = __shfl_up(w, d, 64);
            if (lane >= d) w += y;
        }
        if (lane < NWV) s_wave[lane] = w;
    }
    __syncthreads();
    const int off = (wave > 0 ? s_wave[wave - 1] : 0) + x - mine;
    if (per == 1) {
        // one problem a thread: each wave writes whole problems' runs of
        // tiles, 64 consecutive entries a store
        s_off[threadIdx.x] = off;
        s_cnt[threadIdx.x] = mine;
        s_base[threadIdx.x] = base1;
        __syncthreads();
        for (int p = wave; p < np; p += NWV) {
            const int cnt = s_cnt[p], o = s_off[p], b = s_base[p];
            for (int j = lane; j < cnt; j += 64) ca.list[o + j] = b + j;
        }
    } else {
        int o = off;
        for (int p = p0; p < p1; p++) {
            if (!camp_active(ca.st[p].phase)) continue;
            const int base = a.probs[p].sw_base[0], cnt = a.probs[p].sw_count[0];
            for (int j = 0; j < cnt; j++) ca.list[o + j] = base + j;
            o += cnt;
        }
    }
    if (threadIdx.x == blockDim.x - 1) *ca.count = off + mine;
    // the FP32 step's per-XCD tile counters start the step at zero
    if (threadIdx.x < 8) ca.count[kCampCounterStride * (threadIdx.x + 1)] = 0;
}

// The interior of sweep tile t (the cells it stores), clipped to the span.
struct Interior {
    int x0, y0, w, n;
};

template <int S, int RY>
__device__ inline Interior interior_of(const Prob& pr, int t) {
    constexpr int TX = 128 - 4 * S, TY = RY - 4 * S;
    const int lt = t - pr.sw_base[0];
    const int ti = lt % pr.ntx[0], tj = lt / pr.ntx[0];
    Interior in;
    in.x0 = pr.xs + ti * TX;
    in.y0 = pr.ys + tj * TY;
    const int x1 = min(min(in.x0 + TX - 1, pr.x1), pr.W - 1);
    const int y1 = min(in.y0 + TY - 1, pr.y1);
    in.w = x1 - in.x0 + 1;
    in.n = (in.w > 0 && y1 >= in.y0) ? in.w * (y1 - in.y0 + 1) : 0;
    return in;
}

// FP64 residual of c on the interior of tile t (the projected natural
// residual and b_eff, as resid64 computes them), the FP32 smoother's
// right-hand side omega r / A, the correction's lower bound -c rounded toward
// zero, e = 0 for the coming segment, and the tile's partial sums into the
// ring slot of this step.
template <int S, int RY>
__device__ void camp_resid_tile(const PassArgs& a, const Prob& pr, int t) {
    const Interior in = interior_of<S, RY>(pr, t);
    const int P = pr.P, W = pr.W, Hh = pr.H;
    const double* c = pr.c64;
    const double* kk = (const double*)pr.k;
    const double* kk2 = (const double*)pr.k2;
    const double d = pr.d, decay = pr.decay;
    double r2 = 0.0, b2 = 0.0;
    for (int idx = threadIdx.x; idx < in.n; idx += blockDim.x) {
        const int x = in.x0 + idx % in.w, y = in.y0 + idx / in.w;
        const size_t i = (size_t)y * P + x;
        const double kq = kk[i];
        if (kq != kq) {
            pr.r32[i] = __int_as_float(0x7fc00000);
            pr.e0[i] = 0.0f;
            continue;
        }
        const double ci = c[i];
        double sum = 0.0, hsum = 0.0;
        int real = 0;
#define CW_NB(cond, j)                 \
    if (cond) {                        \
        const double v = c[j];         \
        sum += v;                      \
        real += 1;                     \
        if (kk[j] != kk[j]) hsum += v; \
    }
        CW_NB(x > 0, i - 1)
        CW_NB(x + 1 < W, i + 1)
        CW_NB(y > 0, i - P)
        CW_NB(y + 1 < Hh, i + P)
#undef CW_NB
        const double A = (double)real * d + decay;
        const double b = kk2 ? ((pr.source + kq) - kk2[i]) : kq;
        const double r = (b + d * sum) - A * ci;
        double rt = r;
        if (A > 0.0 && !(ci + r / A > 0.0)) rt = -A * ci;
        const double be = b + d * hsum;
        r2 += rt * rt;
        b2 += be * be;
        pr.r32[i] = (float)(pr.omega * r / A);
        pr.e0[i] = 0.0f;
        if (pr.lb32) pr.lb32[i] = __double2float_rz(-ci);
    }
    block_partial(a.partials + (a.pass & 1) * a.part_stride, t, r2, b2);
}

// c <- max(c + e, 0) on the free cells of tile t's interior, e from buffer par.
template <int S, int RY>
__device__ void camp_correct_tile(const Prob& pr, int t, int par) {
    const Interior in = interior_of<S, RY>(pr, t);
    const float* e = par ? pr.e1 : pr.e0;
    const double* kk = (const double*)pr.k;
    for (int idx = threadIdx.x; idx < in.n; idx += blockDim.x) {
        const size_t i = (size_t)(in.y0 + idx / in.w) * pr.P + (in.x0 + idx % in.w);
        if (kk[i] != kk[i]) continue;
        const double v = pr.c64[i] + (double)e[i];
        pr.c64[i] = v > 0.0 ? v : 0.0;
    }
}

// FP32 refinement step: a fixed grid loops over the step's tile list; each
// XCD owns a contiguous run of it (neighbouring tiles share one L2), its
// workgroups take the run's tiles one at a time from the XCD's counter, and
// once the run is taken they help the other XCDs with theirs. Tiles cost
// what their field's phase costs (s smoothing sweeps, or one residual or
// correction pass), and fields in different phases sit side by side in the
// list, so tiles are handed out as workgroups free up rather than dealt out
// in advance. Each tile does its field's phase.
template <int S, int NW, int NR, bool GEN>
__global__ __launch_bounds__(NW * 64) void camp_step(CampArgs ca) {
    PassArgs a = resolve(ca.pa);
    const int L = a.pass;
    const int n = *ca.count;
    const int x = blockIdx.x & 7;
    const int chunk = (n + 7) >> 3;
    int* ctr = ca.count + kCampCounterStride;
    __shared__ int s_tile;
    int tries = 0;  // thread 0: XCD runs found empty so far (own run first)
    a.mode = MODE_SWEEP;
    a.tl = 0;
    a.check = 0;
    while (true) {
        __syncthreads();  // the previous tile's reads of LDS (and of s_tile) are done
        if (threadIdx.x == 0) {
            int t = -1;
            while (tries < 8) {
                const int y = (x + tries) & 7;
                const int lo = y * chunk, hi = min(n, lo + chunk);
                if (lo < hi) {
                    const int j = atomicAdd(&ctr[kCampCounterStride * y], 1);
                    if (lo + j < hi) {
                        t = ca.list[lo + j];
                        break;
                    }
                }
                tries++;
            }
            s_tile = t;
        }
        __syncthreads();
        // uniform, so the tile's problem record is read with scalar loads
        const int t = __builtin_amdgcn_readfirstlane(s_tile);
        if (t < 0) break;
        const int p = a.tile_prob[t];
        const CampField& f = ca.st[p];
        const int phase = f.phase;
        const int par = (L - f.seg_start) & 1;
        if (phase == PH_SMOOTH) relax2d_reg_tile<float, S, NW, NR, false, GEN>(a, t, par);
        else if (phase == PH_CORRECT) camp_correct_tile<S, NW * NR>(ca.probs64[p], t, par);
        else camp_resid_tile<S, NW * NR>(a, ca.probs64[p], t);
    }
}

// FP64 step, full width: one workgroup per slot of the group's tile list (the
// launches of a step together cover every tile of the group; slots past the
// list return at once), each XCD taking a contiguous run of a launch's slots.
// The sweep body is expanded in this kernel itself, so it compiles to the
// standalone sweep's registers (two 512-thread workgroups a CU at four
// sweeps). A step is split over several launches on several queues: one queue
// dispatches about 20 workgroups a microsecond, about what the CUs retire.
template <int S, int NW, int NR, bool GEN>
__global__ __launch_bounds__(NW * 64) void camp_sweep64(CampArgs ca) {
    using T = double;
    constexpr bool PLAIN = true;
    const int slot = ca.slot_base + xcd_remap(blockIdx.x, gridDim.x);
    if (slot >= *ca.count) return;
    PassArgs a = resolve(ca.pa);
    a.mode = MODE_SWEEP;
    a.tl = 0;
    a.check = a.pass % ca.check_every == 0;
#define CW_TILE_INDEX ca.list[slot]
#define CW_INPUT_PARITY (a.pass & 1)
#define CW_TILE_PROLOGUE 0
#include "cw_relax2d_tile.inc"
#undef CW_TILE_INDEX
#undef CW_INPUT_PARITY
#undef CW_TILE_PROLOGUE
}

// FP64 step for the tail of a solve, when few tiles are left: a fixed grid of
// resident workgroups loops over the list (each XCD a contiguous run), so a
// step dispatches only those instead of one workgroup per tile of the group.
// The loop keeps more registers live (one workgroup a CU), which matters
// little with few tiles.
template <int S, int NW, int NR, bool GEN>
__global__ __launch_bounds__(NW * 64) void camp_sweep64_tail(CampArgs ca) {
    using T = double;
    constexpr bool PLAIN = true;
    PassArgs a = resolve(ca.pa);
    a.mode = MODE_SWEEP;
    a.tl = 0;
    a.check = a.pass % ca.check_every == 0;
    const int n = *ca.count;
    const int x = blockIdx.x & 7, k = blockIdx.x >> 3, per_xcd = gridDim.x >> 3;
    const int chunk = (n + 7) >> 3;
    const int hi = min(n, (x + 1) * chunk);
    for (int slot = x * chunk + k; slot < hi; slot += per_xcd) {
        __syncthreads();  // the previous tile's reads of LDS are done
#define CW_TILE_INDEX ca.list[slot]
#define CW_INPUT_PARITY (a.pass & 1)
#define CW_TILE_PROLOGUE 0
#include "cw_relax2d_tile.inc"
#undef CW_TILE_INDEX
#undef CW_INPUT_PARITY
#undef CW_TILE_PROLOGUE
    }
}

}  // namespace cw

namespace cwi {

namespace {

template <int S, int NW, int NR, bool FP64, bool GEN>
void* step_kernel(bool tail) {
    if constexpr (FP64)
        return tail ? (void*)&camp_sweep64_tail<S, NW, NR, GEN> : (void*)&camp_sweep64<S, NW, NR, GEN>;
    else return (void*)&camp_step<S, NW, NR, GEN>;
}

// The step kernel of a shape (FP64: full width or tail), or null when it is
// not compiled.
void* pick(int shape, bool fp64, bool gen, bool tail, int* threads) {
#define CW_K(S_, NW_, NR_)                                                              \
    do {                                                                                \
        *threads = NW_ * 64;                                                            \
        return gen ? step_kernel<S_, NW_, NR_, FP64_, true>(tail)                       \
                   : step_kernel<S_, NW_, NR_, FP64_, false>(tail);                     \
    } while (0)
    if (fp64) {
        constexpr bool FP64_ = true;
        switch (shape) {
        case 8: CW_K(4, 8, 8);
        case 9: CW_K(8, 8, 8);
        default: return nullptr;
        }
    }
    constexpr bool FP64_ = false;
    switch (shape) {
    case 15: CW_K(8, 16, 8);
    case 18: CW_K(6, 16, 8);
    case 19: CW_K(10, 16, 8);
    case 20: CW_K(12, 16, 8);
    case 21: CW_K(6, 8, 8);
    case 22: CW_K(8, 8, 8);
    case 23: CW_K(12, 8, 8);
    case 24: CW_K(16, 16, 8);
    case 25: CW_K(8, 16, 10);
    case 26: CW_K(10, 16, 10);
    case 27: CW_K(12, 16, 10);
    default: return nullptr;
    }
#undef CW_K
}

}  // namespace

bool camp_shape_supported(int shape, bool fp64) {
    int th = 0;
    return pick(shape, fp64, false, false, &th) != nullptr;
}

cw_status camp_grid(int shape, bool fp64, bool gen, bool tail, long ntiles, int* grid) {
    int th = 0;
    void* k = pick(shape, fp64, gen, tail, &th);
    if (!k) return fail(CW_ERR_ARG, "plane shape %d is not compiled for the device-driven batch", shape);
    if (fp64 && !tail) {
        // one workgroup per tile of the group, over kCampSplit launches
        const long per = (ntiles + kCampSplit - 1) / kCampSplit;
        if (per > 0x7fffffff) return fail(CW_ERR_ARG, "too many tiles in one group");
        *grid = (int)std::max<long>(1, per);
        return CW_OK;
    }
    int dev = 0, per = 0;
    hipDeviceProp_t prop;
    CW_HIP(hipGetDevice(&dev));
    CW_HIP(hipGetDeviceProperties(&prop, dev));
    CW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, th, 0));
    const int g = std::max(1, per) * prop.multiProcessorCount;
    *grid = std::max(8, g / 8 * 8);
    return CW_OK;
}

cw_status camp_launch_transition(const CampGroupDev* groups, int ngroups, int nfields, int step,
                                 const int* run, hipStream_t stream) {
    const int waves = kCampThreads / 64;
    hipLaunchKernelGGL(camp_verdicts, dim3((nfields + waves - 1) / waves), dim3(kCampThreads), 0,
                       stream, groups, ngroups, step, run);
    hipLaunchKernelGGL(camp_lists, dim3(ngroups), dim3(kCampThreads), 0, stream, groups);
    CW_HIP(hipGetLastError());
    return CW_OK;
}

cw_status camp_launch_step(int shape, bool fp64, bool gen, bool tail, const CampArgs& ca,
                           int grid, hipStream_t stream) {
    int th = 0;
    void* k = pick(shape, fp64, gen, tail, &th);
    if (!k) return fail(CW_ERR_ARG, "plane shape %d is not compiled for the device-driven batch", shape);
    CampArgs arg = ca;
    void* args[] = {&arg};
    CW_HIP(hipLaunchKernel(k, dim3(grid), dim3(th), args, 0, stream));
    return CW_OK;
}

}  // namespace cwi
