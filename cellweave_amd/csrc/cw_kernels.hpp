// HIP kernels of the diffusion field solve for gfx950 (CDNA4, wave64).
//
// The per-voxel update restates crates/engine/src/pde.rs:353-398:
//   sum over real neighbours in the order x-1, x+1, y-1, y+1, z-1, z+1,
//   balanced = (d*sum + source + production - uptake) / (real*d + decay),
//   after    = max((1 - omega)*before + omega*balanced, 0).
// FP64 kernels are compiled with -ffp-contract=off and keep that order, so
// each update rounds exactly as the reference's does. A neighbour outside the
// box is a zero in LDS: adding +0.0 to a sum that starts at +0.0 and only
// gathers non-negative terms leaves it unchanged to the bit, so the sum needs
// no branch, while `real` is counted from coordinates.
//
// Every launch sweeps a plan: a set of fields ("problems") whose tiles are
// numbered consecutively; blockIdx -> (problem, tile) via a table, remapped so
// that each XCD gets a contiguous run of tiles (neighbouring tiles share halo
// lines in that XCD's L2).
//
// Stopping without fences: launch L of a run finds out by itself whether a
// problem stopped after launch L-1. Every workgroup of the problem reads what
// launch L-1 left in memory (the per-sweep largest changes, reduced with
// atomicMax, and the residual partials of the input) and reaches the same
// verdict; the problem's first tile records it in the problem's control
// block. Kernel boundaries make launch L-1's writes visible, so no workgroup
// ever waits for or fences against another. A one-workgroup `decide` kernel
// records the verdict on the last launch of a chunk before the host reads it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace cw {

constexpr int kMaxS = 8;
constexpr int kWave = 64;

enum Status : int { ST_RUNNING = 0, ST_DONE = 1, ST_MAXED = 2, ST_REPLAY = 3, ST_IDLE = 4 };
enum Crit : int { CRIT_NONE = 0, CRIT_MAXCHANGE = 1, CRIT_RESID = 2 };
enum Mode : int {
    MODE_SWEEP = 0,       // a pass of s full sweeps over running problems
    MODE_REPLAY = 1,      // single sweeps for problems rolled back by an exact stop
    DECIDE_PASS = 2,      // decide: verdict on the previous pass
    DECIDE_RESID = 3,     // decide: residual sums after resid64
    DECIDE_REPLAY_END = 4 // decide: replayed problems are done
};

// One field in a launch plan. Arrays are pitched: row stride P elements
// (P even, P >= W), plane stride P*H. Cells in the padding hold 0 in state
// arrays and NaN in coefficient arrays.
struct Prob {
    int W, H, D, P;
    int xs, ys, zs;          // tile-grid origin (xs even): first swept column/row/plane
    int x1, y1, z1;          // last swept column/row/plane (inclusive)
    int sw_base[2], sw_count[2], ntx[2], nty[2];  // sweep tilings: [0] s sweeps, [1] one
    int el_base, el_count, elx, ely;              // elementwise tiling (128 x 4 per plane)
    int inner_idx;           // this problem's index in the inner plan (refinement)
    int zc;                  // 3D streaming: planes per z chunk
    void* x[2];              // state ping-pong (T); equal for in-place 3D sweeps
    const void* k;           // coefficient (T): NaN = held. FP64: production minus
                             // uptake when at most one is nonzero; else production
    const void* k2;          // FP64 general form: uptake (else null)
    const void* lb;          // FP32 inner: lower bound of the correction (or null)
    double d, decay, omega, omo, source;  // omo = 1 - omega as the host rounds it
    // FP64 refinement arrays
    double* c64;             // the FP64 state being corrected / tested
    float* e0;               // correction buffers of the inner plan
    float* e1;
    float* r32;              // FP32 defect right-hand side (NaN = held)
    float* lb32;             // FP32 lower bound of the correction (or null)
};

struct Ctl {
    int status;    // Status
    int parity;    // when stopped: x[parity] holds the answer (replay: the input)
    int sweeps;    // when stopped: sweeps run
    int replay;    // REPLAY: single sweeps still to run from x[parity]
    int budget;    // > 0: stop DONE once this many sweeps have run (inner segments)
    int nonfinite;
    int pad0, pad1;
    unsigned long long moved[3][kMaxS];  // largest change per sweep, ring by pass % 3
    double r2, b2;                       // last residual sums (squared norms)
};

struct PassArgs {
    const Prob* probs;
    Ctl* ctl;
    const int* tile_prob;
    double* partials;      // 2 ring slots of part_stride doubles (r2, b2 per tile)
    int part_stride;
    int nprob;
    int pass;              // launch index L: running problems have done L passes
    int sweeps_before;     // sweeps running problems have done before this launch
    int s, tl;             // this launch: sweeps, sweep tiling
    int prev_s, prev_tl, prev_check;
    int check;             // this launch computes the residual of its input
    int crit, max_sweeps, mode, replay_step;
    double tol;
    const Ctl* inner_ctl;  // correct64: the inner plan's control blocks
    int project;           // resid64: project onto the floor (0 for the adjoint's linear system)
    int pad_;
    const int* run;        // graph replays: device base (pass, sweeps) added to pass, sweeps_before
    const double* sums;    // non-null: the previous (checked) launch's residual sums per problem,
                           // folded by fold_sums; null: each workgroup sums the tile partials
};

// A launch replayed from a captured graph carries pass and sweeps_before as
// offsets from a device-resident base that a small kernel advances after each
// chunk; every kernel resolves them first.
__device__ __forceinline__ PassArgs resolve(const PassArgs& a0) {
    PassArgs a = a0;
    if (a.run) {
        a.pass += a.run[0];
        a.sweeps_before += a.run[1];
    }
    return a;
}

static __global__ void advance_run(int* run, int passes, int sweeps) {
    if (threadIdx.x == 0) {
        run[0] += passes;
        run[1] += sweeps;
    }
}

template <typename T> struct V2;
template <> struct V2<double> { using type = double2; };
template <> struct V2<float> { using type = float2; };

__device__ __forceinline__ int xcd_remap(int b, int n) {
    // blockIdx b runs on XCD b % 8; give XCD x the contiguous logical range
    // [x*q + min(x, r), ...) so halo neighbours meet in one L2.
    const int q = n >> 3, r = n & 7;
    const int x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ __forceinline__ unsigned long long wave_umax(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long w = __shfl_xor(v, o, kWave);
        v = w > v ? w : v;
    }
    return v;
}

// Largest change of a sweep as the bit pattern of a non-negative double:
// integer order equals numeric order, so atomicMax is exact and
// order-independent. NaN changes are dropped, as f64::max drops them.
__device__ __forceinline__ unsigned long long change_bits(double m) {
    return (m == m) ? (unsigned long long)__double_as_longlong(m) : 0ull;
}

// Sum of one problem's tile partials of one ring slot, by one whole wave:
// lane l adds tiles l, l + 64, l + 128, ... in that order, then a fixed
// shuffle tree adds the lanes, so every wave that computes it gets the same
// bits. Each (r2, b2) pair is one 16-byte load, eight of them in flight a
// lane: a problem's partials come from other XCDs' writes, and a wave that
// waited on one load at a time would spend about a microsecond per 64 tiles.
__device__ __forceinline__ void problem_sums(const double* part, int base, int cnt, double* r2,
                                             double* b2) {
    const int lane = threadIdx.x & 63;
    const double2* p2 = reinterpret_cast<const double2*>(part) + base;
    double r = 0.0, b = 0.0;
    int t = lane;
    for (; t + 7 * 64 < cnt; t += 8 * 64) {
        double2 v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = p2[t + 64 * k];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            r += v[k].x;
            b += v[k].y;
        }
    }
    for (; t < cnt; t += 64) {
        const double2 v = p2[t];
        r += v.x;
        b += v.y;
    }
    *r2 = wave_sum(r);
    *b2 = wave_sum(b);
}

// Second level of the stopping verdict's reduction, launched right after a
// checked launch on a plan whose problems have many tiles: one wave per
// problem sums that launch's tile partials once (problem_sums: the order, and
// so the bits, every workgroup would get) and writes (r2, b2) to sums[2p].
// The next launch's workgroups then read two doubles each instead of all of
// their problem's partials, which would be O(tiles^2) reads a launch.
static __global__ __launch_bounds__(256) void fold_sums(PassArgs a_in) {
    const PassArgs a = resolve(a_in);
    const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= a.nprob) return;  // wave-uniform
    const Prob& pr = a.probs[p];
    double r2, b2;
    problem_sums(a.partials + (a.pass & 1) * a.part_stride, pr.sw_base[a.tl], pr.sw_count[a.tl],
                 &r2, &b2);
    if ((threadIdx.x & 63) == 0) {
        double* out = const_cast<double*>(a.sums);
        out[2 * p] = r2;
        out[2 * p + 1] = b2;
    }
}

struct Verdict {
    int status, parity, sweeps, replay, nonfinite;
};

// Fate of a running problem after launch a.pass - 1.
__device__ inline Verdict assess(const PassArgs& a, const Ctl& c, bool have_r2, double r2) {
    Verdict v{ST_RUNNING, a.pass & 1, a.sweeps_before, 0, 0};
    if (a.pass == 0) return v;
    const int lp = a.pass - 1;
    const int sb = a.sweeps_before - a.prev_s;
    if (have_r2) {
        if (!(r2 == r2) || r2 > 1.7e308) {
            v.status = ST_MAXED; v.nonfinite = 1; v.parity = lp & 1; v.sweeps = sb;
            return v;
        }
        if (r2 <= a.tol * a.tol * (c.b2 > 0.0 ? c.b2 : 1.0)) {
            // converged on the input of launch L-1: that input is the answer
            v.status = ST_DONE; v.parity = lp & 1; v.sweeps = sb;
            return v;
        }
    }
    if (a.crit == CRIT_MAXCHANGE) {
        int found = -1;
        for (int j = 0; j < a.prev_s; j++)
            if (__longlong_as_double((long long)c.moved[lp % 3][j]) < a.tol) { found = j; break; }
        if (found == a.prev_s - 1) { v.status = ST_DONE; return v; }
        if (found >= 0) {
            v.status = ST_REPLAY; v.parity = lp & 1; v.sweeps = sb; v.replay = found + 1;
            return v;
        }
    }
    if (c.budget > 0 && a.sweeps_before >= c.budget) { v.status = ST_DONE; return v; }
    if (a.sweeps_before >= a.max_sweeps) v.status = ST_MAXED;
    return v;
}

__device__ __forceinline__ void record(Ctl& c, const Verdict& v) {
    c.parity = v.parity;
    c.sweeps = v.sweeps;
    c.replay = v.replay;
    c.nonfinite = v.nonfinite;
    c.status = v.status;
}

// Prologue of every sweep launch: returns whether this workgroup's problem
// runs in this launch and which buffer holds its input. Thread 0 decides; the
// problem's first tile records a stop and clears the change slot the next
// launch will use.
__device__ inline bool sweep_prologue(const PassArgs& a, const Prob& pr, int p, int lt,
                                      int* src_parity) {
    __shared__ int s_run, s_par;
    Ctl& c = a.ctl[p];
    if (a.mode == MODE_REPLAY) {
        if (threadIdx.x == 0) {
            const bool run = c.status == ST_REPLAY && a.replay_step < c.replay;
            s_run = run;
            s_par = (c.parity + a.replay_step) & 1;
        }
        __syncthreads();
        *src_parity = s_par;
        return s_run;
    }
    const bool have_r2 = a.prev_check && a.crit == CRIT_RESID && a.pass > 0;
    double r2 = 0.0, b2 = 0.0;
    if (have_r2 && threadIdx.x < 64) {
        if (a.sums) {
            r2 = a.sums[2 * p];
            b2 = a.sums[2 * p + 1];
        } else {
            const int tl = a.prev_tl;
            problem_sums(a.partials + ((a.pass - 1) & 1) * a.part_stride, pr.sw_base[tl],
                         pr.sw_count[tl], &r2, &b2);
        }
    }
    if (threadIdx.x == 0) {
        int run = 0;
        if (c.status == ST_RUNNING) {
            const Verdict v = assess(a, c, have_r2, r2);
            run = v.status == ST_RUNNING;
            if (!run && lt == 0) {
                if (have_r2) c.r2 = r2;
                record(c, v);
            }
        }
        if (lt == 0) {
#pragma unroll
            for (int j = 0; j < kMaxS; j++) c.moved[(a.pass + 1) % 3][j] = 0ull;
        }
        s_run = run;
    }
    __syncthreads();
    *src_parity = a.pass & 1;
    return s_run;
}

__device__ __forceinline__ void block_partial(double* slot, int tile, double r2, double b2) {
    __shared__ double s_r[16], s_b[16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    r2 = wave_sum(r2);
    b2 = wave_sum(b2);
    if (lane == 0) { s_r[wv] = r2; s_b[wv] = b2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tr = 0.0, tb = 0.0;
        for (int w = 0; w < nw; w++) { tr += s_r[w]; tb += s_b[w]; }
        slot[2 * tile] = tr;
        slot[2 * tile + 1] = tb;
    }
}

template <int NW, int S>
__device__ __forceinline__ void block_moved(Ctl& c, int pass, const unsigned long long (&mv)[S]) {
    __shared__ unsigned long long s_m[NW][S];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < S; j++) {
        const unsigned long long m = wave_umax(mv[j]);
        if (lane == 0) s_m[wv][j] = m;
    }
    __syncthreads();
    if (threadIdx.x < S) {
        unsigned long long m = 0ull;
        for (int w = 0; w < NW; w++) m = s_m[w][threadIdx.x] > m ? s_m[w][threadIdx.x] : m;
        // Thousands of workgroups' atomics on one word serialise. The slot
        // only grows during a launch, so a workgroup whose largest change is
        // not above the slot's current value (a coherent load) changes
        // nothing and skips its atomic: the result is the same maximum.
        unsigned long long* slot = &c.moved[pass % 3][threadIdx.x];
        if (m && m > __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(slot, m);
    }
}

__device__ __forceinline__ int real_count(int gx, int gy, int gz, int W, int H, int D) {
    return (gx > 0) + (gx < W - 1) + (gy > 0) + (gy < H - 1) + (gz > 0) + (gz < D - 1);
}

// Typed fused multiply-add (the untyped builtin would widen floats to double)
// and the reciprocal used by the smoothers' coefficients.
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float rcp_t(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float max_t(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double max_t(double a, double b) { return __builtin_fmax(a, b); }
template <typename T> __device__ __forceinline__ T neg_inf();
template <> __device__ __forceinline__ float neg_inf<float>() { return -__builtin_huge_valf(); }
template <> __device__ __forceinline__ double neg_inf<double>() { return -__builtin_huge_val(); }
__device__ __forceinline__ double rcp_t(double x) { return 1.0 / x; }

template <typename T> __device__ __forceinline__ T nan_of();
template <> __device__ __forceinline__ double nan_of<double>() {
    return __longlong_as_double(0x7ff8000000000000ll);
}
template <> __device__ __forceinline__ float nan_of<float>() { return __int_as_float(0x7fc00000); }

// ---------------------------------------------------------------------------
// 2D LDS tile sweeps: s full sweeps (2s half sweeps) per launch.
//
// A workgroup loads a region 128 wide and RY tall (halo H = 2s around a
// (128-2H) x (RY-2H) interior) from src into LDS, split by the parity of the
// local x so every neighbour read is unit-stride across a wave. Wave w owns
// rows w, w+NW, ... (NW even, so all its rows share a parity); lane l owns the
// pair of columns (2l, 2l+1) and keeps the pair's coefficients in registers,
// stored by colour so every register index is a compile-time constant. Half
// sweep h updates the cells of colour h&1 at distance > h from the region's
// edge (the rest is stale halo), so after 2s half sweeps the interior is exact
// and is stored to dst. The state loads are issued before the stop verdict
// (the input parity of a sweep launch is known from the pass index), so their
// latency overlaps it.
// ---------------------------------------------------------------------------
template <int RY, int NW, int S>
struct Region {
    static constexpr int H = 2 * S;
    static constexpr int RX = 128;
    static constexpr int TX = RX - 2 * H;
    static constexpr int TY = RY - 2 * H;
    static constexpr int NR = RY / NW;
    static_assert(RY % NW == 0 && NW % 2 == 0, "rows split evenly over an even wave count");
    static_assert(TY > 0, "interior rows");
};

// FP64 plain sweep: the reference's arithmetic and floor, the per-sweep
// largest change for the reference stopping rule, and optionally the
// residual of the input for the relative-residual rule.
template <int S, int RY, int NW, bool GEN>
__global__ __launch_bounds__(NW * 64) void sweep2d(PassArgs a_in) {
    const PassArgs a = resolve(a_in);
    using R = Region<RY, NW, S>;
    constexpr int H = R::H, RX = R::RX, TX = R::TX, TY = R::TY, NR = R::NR;
    static_assert(S <= kMaxS, "change slots");
    __shared__ double cs[RY][2][64];

    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int p = a.tile_prob[t];
    const Prob& pr = a.probs[p];
    const int tl = a.tl;
    const int lt = t - pr.sw_base[tl];
    const int ti = lt % pr.ntx[tl], tj = lt / pr.ntx[tl];
    const int gxo = pr.xs + ti * TX - H;
    const int gyo = pr.ys + tj * TY - H;
    const int W = pr.W, Hh = pr.H, P = pr.P, bx1 = pr.x1, by1 = pr.y1;
    const int gx = gxo + 2 * lane;
    const bool colin = gx >= 0 && gx < W;
    const double* kk = (const double*)pr.k;
    const double* kk2 = (const double*)pr.k2;
    constexpr bool general = GEN;  // separate production and uptake (source != 0 or both set)
    // real neighbours: x part per cell of the pair, y part per row (2D: D = 1)
    const int rxe = (gx > 0) + (gx < W - 1), rxo = (gx + 1 > 0) + (gx + 1 < W - 1);
    const int base = (gxo + gyo) & 1;
    const int cb = (base + wv) & 1;  // colour of the even column on this wave's rows
    const int rx0 = cb ? rxo : rxe, rx1 = cb ? rxe : rxo;  // by colour

    double k0[NR], k1[NR], u0[GEN ? NR : 1], u1[GEN ? NR : 1], v0[NR], v1[NR];
    const bool replay = a.mode == MODE_REPLAY;
#pragma unroll
    for (int i = 0; i < NR; i++) {
        const int gy = gyo + wv + NW * i;
        double ce = nan_of<double>(), co = nan_of<double>(), ue = 0.0, uo = 0.0;
        v0[i] = 0.0; v1[i] = 0.0;
        if (colin && gy >= 0 && gy < Hh) {
            const size_t off = (size_t)gy * P + gx;
            const double2 c = *(const double2*)(kk + off);
            ce = c.x; co = c.y;
            if (general) { const double2 u = *(const double2*)(kk2 + off); ue = u.x; uo = u.y; }
            if (!replay) {
                const double2 v = *(const double2*)((const double*)pr.x[a.pass & 1] + off);
                v0[i] = v.x; v1[i] = v.y;
            }
        }
        k0[i] = cb ? co : ce; k1[i] = cb ? ce : co;
        if (GEN) { u0[GEN ? i : 0] = cb ? uo : ue; u1[GEN ? i : 0] = cb ? ue : uo; }
    }

    int par;
    if (!sweep_prologue(a, pr, p, lt, &par)) return;
    const double* src = (const double*)pr.x[par];
    double* dst = (double*)pr.x[par ^ 1];

#pragma unroll
    for (int i = 0; i < NR; i++) {
        const int y = wv + NW * i;
        const int gy = gyo + y;
        if (replay && colin && gy >= 0 && gy < Hh) {
            const double2 v = *(const double2*)(src + (size_t)gy * P + gx);
            v0[i] = v.x; v1[i] = v.y;
        }
        cs[y][0][lane] = v0[i];
        cs[y][1][lane] = v1[i];
    }
    __syncthreads();

    const double d = pr.d, decay = pr.decay, omega = pr.omega, omo = pr.omo, source = pr.source;

    // Residual of the input state on the interior (projected natural residual).
    double r2 = 0.0;
    if (a.check) {
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const int y = wv + NW * i;
            if (y < H || y >= H + TY) continue;
            const int gy = gyo + y;
            const int ry = (gy > 0) + (gy < Hh - 1);
#pragma unroll
            for (int col = 0; col < 2; col++) {
                const int q = col ^ cb;
                const int x = 2 * lane + q;
                if (x < H || x >= H + TX) continue;
                const double kq = col ? k1[i] : k0[i];
                if (kq != kq) continue;
                const double c = cs[y][q][lane];
                const double w = q ? cs[y][0][lane] : cs[y][1][lane - 1];
                const double e = q ? cs[y][0][lane + 1] : cs[y][1][lane];
                double sum = 0.0;
                sum += w; sum += e; sum += cs[y - 1][q][lane]; sum += cs[y + 1][q][lane];
                const double A = (double)((col ? rx1 : rx0) + ry) * d + decay;
                const double b = general ? ((source + kq) - (col ? u1[GEN ? i : 0] : u0[GEN ? i : 0])) : kq;
                double r = (b + d * sum) - A * c;
                if (A > 0.0 && !(c + r / A > 0.0)) r = -A * c;
                r2 += r * r;
            }
        }
    }

    unsigned long long mv[S];
#pragma unroll
    for (int j = 0; j < S; j++) mv[j] = 0ull;
#pragma unroll
    for (int h = 0; h < 2 * S; h++) {
        const int colour = h & 1;
        const int lo = h + 1, hiY = RY - 2 - h, hiX = RX - 2 - h;
        const int q = colour ^ cb;
        const int x = 2 * lane + q;
        const bool xin = x >= lo && x <= hiX;
        const int rx = colour ? rx1 : rx0;
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const int y = wv + NW * i;
            if (y < lo || y > hiY) continue;
            const double kq = colour ? k1[i] : k0[i];
            if (!xin || kq != kq) continue;
            const int gy = gyo + y;
            const int ry = (gy > 0) + (gy < Hh - 1);
            const double before = cs[y][q][lane];
            const double w = q ? cs[y][0][lane] : cs[y][1][lane - 1];
            const double e = q ? cs[y][0][lane + 1] : cs[y][1][lane];
            double sum = 0.0;
            sum += w; sum += e; sum += cs[y - 1][q][lane]; sum += cs[y + 1][q][lane];
            double num = d * sum;
            if (general) {
                num = num + source;
                num = num + kq;
                num = num - (colour ? u1[GEN ? i : 0] : u0[GEN ? i : 0]);
            } else {
                num = num + kq;
            }
            const double bal = num / ((double)(rx + ry) * d + decay);
            double after = omo * before + omega * bal;
            after = after > 0.0 ? after : 0.0;
            cs[y][q][lane] = after;
            if (y >= H && y < H + TY && x >= H && x < H + TX) {
                const unsigned long long m = change_bits(__builtin_fabs(after - before));
                mv[h >> 1] = m > mv[h >> 1] ? m : mv[h >> 1];
            }
        }
        __syncthreads();
    }

    // Store the interior (only rows and pairs that reach the swept box).
    if (lane >= H / 2 && lane < (H + TX) / 2 && gx <= bx1) {
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const int y = wv + NW * i;
            if (y < H || y >= H + TY) continue;
            const int gy = gyo + y;
            if (gy > by1) continue;
            double2 v;
            v.x = cs[y][0][lane];
            v.y = cs[y][1][lane];
            *(double2*)(dst + (size_t)gy * P + gx) = v;
        }
    }

    if (a.check) block_partial(a.partials + (a.pass & 1) * a.part_stride, t, r2, 0.0);
    if (a.crit == CRIT_MAXCHANGE && a.mode == MODE_SWEEP) block_moved<NW, S>(a.ctl[p], a.pass, mv);
}

// Linear smoother in pre-scaled form: e <- max(lb, (1-omega) e + alpha sum + beta)
// with alpha = omega d / A and beta = omega r / A (A = real d + decay), which
// is the SOR update of A e = r written as two fused multiply-adds. beta is
// the stored coefficient (NaN = held); lb, when present, bounds the FP32
// correction from below so c + e stays >= 0 (the floor of the field).
template <typename T, int S, int RY, int NW>
__global__ __launch_bounds__(NW * 64) void smooth2d(PassArgs a_in) {
    const PassArgs a = resolve(a_in);
    using R = Region<RY, NW, S>;
    constexpr int H = R::H, RX = R::RX, TX = R::TX, TY = R::TY, NR = R::NR;
    using T2 = typename V2<T>::type;
    __shared__ T cs[RY][2][64];

    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int p = a.tile_prob[t];
    const Prob& pr = a.probs[p];
    const int tl = a.tl;
    const int lt = t - pr.sw_base[tl];
    const int ti = lt % pr.ntx[tl], tj = lt / pr.ntx[tl];
    const int gxo = pr.xs + ti * TX - H;
    const int gyo = pr.ys + tj * TY - H;
    const int W = pr.W, Hh = pr.H, P = pr.P, bx1 = pr.x1, by1 = pr.y1;
    const int gx = gxo + 2 * lane;
    const bool colin = gx >= 0 && gx < W;
    const T* kk = (const T*)pr.k;
    const T* lbp = (const T*)pr.lb;
    const bool has_lb = lbp != nullptr;
    const T* srcp = (const T*)pr.x[a.pass & 1];
    const int rxe = (gx > 0) + (gx < W - 1), rxo = (gx + 1 > 0) + (gx + 1 < W - 1);
    const int base = (gxo + gyo) & 1;
    const int cb = (base + wv) & 1;
    const int rx0 = cb ? rxo : rxe, rx1 = cb ? rxe : rxo;
    const T d = (T)pr.d, decay = (T)pr.decay, omega = (T)pr.omega, omo = (T)pr.omo;

    T b0[NR], b1[NR], l0[NR], l1[NR], a0[NR], a1[NR], v0[NR], v1[NR];
#pragma unroll
    for (int i = 0; i < NR; i++) {
        const int gy = gyo + wv + NW * i;
        T ce = nan_of<T>(), co = nan_of<T>(), le = T(0), lo = T(0);
        v0[i] = T(0); v1[i] = T(0);
        if (colin && gy >= 0 && gy < Hh) {
            const size_t off = (size_t)gy * P + gx;
            const T2 c = *(const T2*)(kk + off);
            ce = c.x; co = c.y;
            if (has_lb) { const T2 l = *(const T2*)(lbp + off); le = l.x; lo = l.y; }
            const T2 v = *(const T2*)(srcp + off);
            v0[i] = v.x; v1[i] = v.y;
        }
        b0[i] = cb ? co : ce; b1[i] = cb ? ce : co;
        l0[i] = cb ? lo : le; l1[i] = cb ? le : lo;
        const int ry = (gy > 0) + (gy < Hh - 1);
        a0[i] = omega * d * rcp_t((T)(rx0 + ry) * d + decay);
        a1[i] = omega * d * rcp_t((T)(rx1 + ry) * d + decay);
    }

    int par;
    if (!sweep_prologue(a, pr, p, lt, &par)) return;
    T* dst = (T*)pr.x[par ^ 1];

#pragma unroll
    for (int i = 0; i < NR; i++) {
        const int y = wv + NW * i;
        cs[y][0][lane] = v0[i];
        cs[y][1][lane] = v1[i];
    }
    __syncthreads();

#pragma unroll
    for (int h = 0; h < 2 * S; h++) {
        const int colour = h & 1;
        const int lo = h + 1, hiY = RY - 2 - h, hiX = RX - 2 - h;
        const int q = colour ^ cb;
        const int x = 2 * lane + q;
        const bool xin = x >= lo && x <= hiX;
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const int y = wv + NW * i;
            if (y < lo || y > hiY) continue;
            const T bq = colour ? b1[i] : b0[i];
            if (!xin || bq != bq) continue;
            const T before = cs[y][q][lane];
            const T w = q ? cs[y][0][lane] : cs[y][1][lane - 1];
            const T e = q ? cs[y][0][lane + 1] : cs[y][1][lane];
            const T sum = (w + e) + (cs[y - 1][q][lane] + cs[y + 1][q][lane]);
            T after = fma_t(omo, before, fma_t(colour ? a1[i] : a0[i], sum, bq));
            if (has_lb) {
                const T l = colour ? l1[i] : l0[i];
                after = after > l ? after : l;
            }
            cs[y][q][lane] = after;
        }
        __syncthreads();
    }

    if (lane >= H / 2 && lane < (H + TX) / 2 && gx <= bx1) {
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const int y = wv + NW * i;
            if (y < H || y >= H + TY) continue;
            const int gy = gyo + y;
            if (gy > by1) continue;
            T2 v;
            v.x = cs[y][0][lane];
            v.y = cs[y][1][lane];
            *(T2*)(dst + (size_t)gy * P + gx) = v;
        }
    }
}

// Neighbour values across lanes without LDS: DPP wave shifts (GFX9 family).
// from_left gives lane l the value of lane l-1, from_right that of lane l+1;
// the wave's end lanes get 0 (their cells sit on the region's edge and are
// never updated). The FP32 forms ask for the zero through the bound control,
// so a shift the compiler cannot fold into its add needs no zeroed
// destination first.
__device__ __forceinline__ float from_left(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float from_right(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));
}
__device__ __forceinline__ double from_left(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), 0x138, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double from_right(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), 0x130, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Register-blocked linear smoother, the same update as smooth2d. Wave w owns
// the contiguous rows [w*NR, (w+1)*NR) of a 128 x (NW*NR) region and lane l
// the columns (2l, 2l+1) of each: the lane's 2*NR cells live in registers for
// the whole launch. Vertical neighbours are the lane's own registers, the
// left or right neighbour across lanes comes by a DPP wave shift, and only a
// wave's first and last rows go through LDS, to the waves above and below
// (double-buffered by half-sweep parity: one barrier per half sweep).
template <typename T, int S, int NW, int NR>
__global__ __launch_bounds__(NW * 64) void smooth2d_reg(PassArgs a_in) {
    const PassArgs a = resolve(a_in);
    constexpr int H = 2 * S;
    constexpr int RX = 128;
    constexpr int RY = NW * NR;
    constexpr int TX = RX - 2 * H;
    constexpr int TY = RY - 2 * H;
    static_assert(TY > 0 && NR >= 2, "region shape");
    using T2 = typename V2<T>::type;
    __shared__ T edge[2][NW][2][2][64];  // [slot][wave][first/last row][column parity][lane]

    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int p = a.tile_prob[t];
    const Prob& pr = a.probs[p];
    const int tl = a.tl;
    const int lt = t - pr.sw_base[tl];
    const int ti = lt % pr.ntx[tl], tj = lt / pr.ntx[tl];
    const int gxo = pr.xs + ti * TX - H;
    const int gyo = pr.ys + tj * TY - H;
    const int W = pr.W, Hh = pr.H, P = pr.P, bx1 = pr.x1, by1 = pr.y1;
    const int gx = gxo + 2 * lane;
    const bool colin = gx >= 0 && gx < W;
    const T* kk = (const T*)pr.k;
    const T* lbp = (const T*)pr.lb;
    const bool has_lb = lbp != nullptr;
    const T* srcp = (const T*)pr.x[a.pass & 1];
    const T d = (T)pr.d, decay = (T)pr.decay, omega = (T)pr.omega, omo = (T)pr.omo;
    const int y0 = wv * NR;  // first region row of this wave

    // alpha = omega d / (real d + decay) by the real neighbour count in x of
    // each column (rx, per lane) and in y of the row (ry = 0, 1, 2, uniform).
    const int rxe = (gx > 0) + (gx < W - 1), rxo = (gx + 1 > 0) + (gx + 1 < W - 1);
    const T od = omega * d;
    const T ae0 = od * rcp_t((T)rxe * d + decay), ae1 = od * rcp_t((T)(rxe + 1) * d + decay),
            ae2 = od * rcp_t((T)(rxe + 2) * d + decay);
    const T ao0 = od * rcp_t((T)rxo * d + decay), ao1 = od * rcp_t((T)(rxo + 1) * d + decay),
            ao2 = od * rcp_t((T)(rxo + 2) * d + decay);

    T ve[NR], vo[NR], be[NR], bo[NR], le[NR], lo[NR];
#pragma unroll
    for (int i = 0; i < NR; i++) {
        const int gy = gyo + y0 + i;
        ve[i] = T(0); vo[i] = T(0); be[i] = nan_of<T>(); bo[i] = nan_of<T>();
        le[i] = T(0); lo[i] = T(0);
        if (colin && gy >= 0 && gy < Hh) {
            const size_t off = (size_t)gy * P + gx;
            const T2 c = *(const T2*)(kk + off);
            be[i] = c.x; bo[i] = c.y;
            if (has_lb) { const T2 l = *(const T2*)(lbp + off); le[i] = l.x; lo[i] = l.y; }
            const T2 v = *(const T2*)(srcp + off);
            ve[i] = v.x; vo[i] = v.y;
        }
    }

    int par;
    if (!sweep_prologue(a, pr, p, lt, &par)) return;
    T* dst = (T*)pr.x[par ^ 1];

    // publish the edge rows for half sweep 0
    edge[0][wv][0][0][lane] = ve[0];
    edge[0][wv][0][1][lane] = vo[0];
    edge[0][wv][1][0][lane] = ve[NR - 1];
    edge[0][wv][1][1][lane] = vo[NR - 1];
    __syncthreads();

    const int base = (gxo + gyo + y0) & 1;  // colour of column 2l on the wave's first row
#pragma unroll
    for (int h = 0; h < 2 * S; h++) {
        const int slot = h & 1;
        const int colour = h & 1;
        const int lo_ = h + 1, hiY = RY - 2 - h, hiX = RX - 2 - h;
        // q0: column parity updated on the wave's first row; row i updates q0 ^ (i & 1)
        const int q0 = (colour + base) & 1;
        const bool ein = 2 * lane >= lo_ && 2 * lane <= hiX;
        const bool oin = 2 * lane + 1 >= lo_ && 2 * lane + 1 <= hiX;
        // halo rows above and below the wave's block (other colour, current values)
        T above = T(0), below = T(0);
        {
            const int qa = q0;                 // column updated on row 0
            const int qb = q0 ^ ((NR - 1) & 1); // column updated on row NR-1
            if (wv > 0) above = edge[slot][wv - 1][1][qa][lane];
            if (wv < NW - 1) below = edge[slot][wv + 1][0][qb][lane];
        }
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const int r = y0 + i;
            if (r < lo_ || r > hiY) continue;       // uniform per wave
            const int gy = gyo + r;
            const int ry = (gy > 0) + (gy < Hh - 1);
            const int q = q0 ^ (i & 1);             // uniform per wave
            // neighbours of column q on row i
            const T up = i > 0 ? (q ? vo[i - 1] : ve[i - 1]) : above;
            const T dn = i < NR - 1 ? (q ? vo[i + 1] : ve[i + 1]) : below;
            if (q == 0) {
                const T left = from_left(vo[i]);
                const T sum = (left + vo[i]) + (up + dn);
                const T al = ry == 2 ? ae2 : (ry == 1 ? ae1 : ae0);
                T after = fma_t(omo, ve[i], fma_t(al, sum, be[i]));
                if (has_lb) after = after > le[i] ? after : le[i];
                const bool upd = ein && be[i] == be[i];
                ve[i] = upd ? after : ve[i];
            } else {
                const T right = from_right(ve[i]);
                const T sum = (ve[i] + right) + (up + dn);
                const T al = ry == 2 ? ao2 : (ry == 1 ? ao1 : ao0);
                T after = fma_t(omo, vo[i], fma_t(al, sum, bo[i]));
                if (has_lb) after = after > lo[i] ? after : lo[i];
                const bool upd = oin && bo[i] == bo[i];
                vo[i] = upd ? after : vo[i];
            }
        }
        if (h + 1 < 2 * S) {
            const int ns = slot ^ 1;
            edge[ns][wv][0][0][lane] = ve[0];
            edge[ns][wv][0][1][lane] = vo[0];
            edge[ns][wv][1][0][lane] = ve[NR - 1];
            edge[ns][wv][1][1][lane] = vo[NR - 1];
            __syncthreads();
        }
    }

    if (lane >= H / 2 && lane < (H + TX) / 2 && gx <= bx1) {
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const int r = y0 + i;
            if (r < H || r >= H + TY) continue;
            const int gy = gyo + r;
            if (gy > by1) continue;
            T2 v;
            v.x = ve[i];
            v.y = vo[i];
            *(T2*)(dst + (size_t)gy * P + gx) = v;
        }
    }
}

// ---------------------------------------------------------------------------
// Register-blocked red-black sweep (2D), for both forms:
//   PLAIN: the reference's FP64 update (pde.rs:353-398) bit for bit, with the
//          floor, the per-sweep largest change and the optional residual of
//          the input; GEN keeps production and uptake apart.
//   else:  the pre-scaled linear smoother e <- max(lb, (1-w) e + a sum + b).
// Wave w owns the contiguous rows [w*NR, (w+1)*NR) of a 128 x (NW*NR)
// region; lane l owns columns (2l, 2l+1) of each row. The lane keeps its
// cells in registers by colour: st[c][i] is the colour-c cell of row i. For a
// colour-c cell, the vertical neighbours are st[1-c][i-1] and st[1-c][i+1]
// (same column, other colour), one horizontal neighbour is st[1-c][i] (the
// lane's other cell) and the other is the neighbouring lane's st[1-c][i],
// fetched by a DPP wave shift. The x parity of colour c on row i is
// P0 ^ (i & 1) with P0 wave-uniform: one branch per half sweep picks an
// instantiation in which every row's parity is a compile-time constant. Only
// a wave's first and last rows go through LDS (double-buffered).
// ---------------------------------------------------------------------------
template <typename T, int NR>
struct RegCells {
    T st[2][NR];   // state by colour
    T kc[2][NR];   // PLAIN: coefficient (NaN = held or outside); smoother: beta
    T k2[2][NR];   // PLAIN+GEN: uptake; smoother: lower bound
    T al[2][NR];   // smoother: alpha = omega d / A (0 on held or outside cells)
    T ga[2][NR];   // FP64 smoother: gamma = -omega (0 on held or outside cells); the
                   // FP32 smoother uses -omega everywhere (below)
};

// A workgroup barrier that orders LDS only: __syncthreads() also waits for
// every global store of the wave to complete (its release fence), and on this
// target a wait on stores is a wait on the vector-memory counter, which the
// loads in flight share. Kernels that prefetch across a barrier use this one.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// A workgroup barrier for LDS that the waves write and read with ds
// instructions: each wave's ds accesses complete before it arrives, and the
// compiler moves no memory access across it. Nothing else is waited for.
__device__ __forceinline__ void ds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// The FP32 smoother keeps gamma = -omega on every cell: a held or outside
// cell has alpha = beta = 0 and a correction of exactly zero (the correction
// buffers are zero wherever no free cell's value was stored: zeroed when
// made, outside the free span whenever the coefficients are rebuilt, and on
// every held cell of the span by the residual pass), so
// 0 + 0*sum + fma(-omega, 0, 0) keeps it at +0 with one register array fewer.
template <typename T, bool PLAIN>
constexpr bool kConstGamma = !PLAIN && std::is_same<T, float>::value;

template <typename T, bool PLAIN, bool GEN, int NR, int C, int P0, int S>
__device__ __forceinline__ void reg_half(RegCells<T, NR>& R, T above, T below, int r0, int lo,
                                         int hiY, int hiX, int lane, int gyo, int Hh,
                                         const int (&rx)[2], T d, T decay, T omega, T omo,
                                         T source, const T (&al)[2][3], bool has_lb, int H,
                                         int TX, int TY, unsigned long long* mv) {
#pragma unroll
    for (int k = 0; k < NR; k++) {
        // FP32: the wave's first and last rows, which wait for the edge rows
        // read from LDS after the barrier, go last, so the interior rows'
        // updates cover that read (a colour's cells are independent).
        const int i = !kConstGamma<T, PLAIN> ? k : k < NR - 2 ? k + 1 : k == NR - 2 ? 0 : NR - 1;
        const int q = P0 ^ (i & 1);  // compile-time after unrolling
        const int r = r0 + i;
        if (PLAIN && (r < lo || r > hiY)) continue;
        if constexpr (kConstGamma<T, PLAIN>) {
            // FP32 smoother, cells by column parity: this colour's cell of
            // row i is at parity q, the other colour's at q ^ 1, and the
            // cells above and below (the other colour) at parity q too.
            const T o = R.st[q ^ 1][i];
            const T across = q ? from_right(o) : from_left(o);
            const T up = i > 0 ? R.st[q][i - 1] : above;
            const T dn = i < NR - 1 ? R.st[q][i + 1] : below;
            const T before = R.st[q][i];
            const T sum = (across + o) + (up + dn);
            const T after = before + fma_t(R.al[q][i], sum, fma_t(-omega, before, R.kc[q][i]));
            R.st[q][i] = GEN ? max_t(after, R.k2[q][i]) : after;
            continue;
        }
        const int x = 2 * lane + q;
        const T o = R.st[C ^ 1][i];
        const T across = q ? from_right(o) : from_left(o);
        const T up = i > 0 ? R.st[C ^ 1][i - 1] : above;
        const T dn = i < NR - 1 ? R.st[C ^ 1][i + 1] : below;
        const T before = R.st[C][i];
        T after;
        if (PLAIN) {
            const T kq = R.kc[C][i];
            const bool upd = x >= lo && x <= hiX && kq == kq;
            const int gy = gyo + r;
            const int ry = (gy > 0) + (gy < Hh - 1);
            // x-1, x+1, y-1, y+1, in the reference's order
            T sum = T(0);
            if (q == 0) { sum += across; sum += o; } else { sum += o; sum += across; }
            sum += up;
            sum += dn;
            T num = d * sum;
            if (GEN) {
                num = num + source;
                num = num + kq;
                num = num - R.k2[C][i];
            } else {
                num = num + kq;
            }
            const T bal = num / ((T)(rx[q] + ry) * d + decay);
            after = omo * before + omega * bal;
            after = after > T(0) ? after : T(0);
            if (upd && r >= H && r < H + TY && x >= H && x < H + TX && S > 0) {
                const unsigned long long m = change_bits((double)__builtin_fabs(after - before));
                *mv = m > *mv ? m : *mv;
            }
            R.st[C][i] = upd ? after : before;
        } else {
            // FP64 smoother: before + alpha*sum + (beta - omega*before); a
            // held or outside cell has alpha = beta = gamma = 0 and keeps its
            // value exactly. Cells of the stale halo are updated too: they
            // are invalid at this depth whatever they hold, and never reach
            // a valid cell.
            const T sum = (across + o) + (up + dn);
            after = before + fma_t(R.al[C][i], sum, fma_t(R.ga[C][i], before, R.kc[C][i]));
            // lower bound (GEN: the launch has one; -inf for a field without)
            R.st[C][i] = GEN ? max_t(after, R.k2[C][i]) : after;
        }
    }
}

// The register-blocked sweep's workgroup body lives in cw_relax2d_tile.inc:
// the kernel below and the device-driven batch's tile function (the tile and
// its input buffer given, no launch prologue) each expand it in place.
template <typename T, int S, int NW, int NR, bool PLAIN, bool GEN>
__global__ __launch_bounds__(NW * 64) void relax2d_reg(PassArgs a_in) {
    const PassArgs a = resolve(a_in);
#define CW_TILE_INDEX xcd_remap(blockIdx.x, gridDim.x)
#define CW_INPUT_PARITY (a.pass & 1)
#define CW_TILE_PROLOGUE 1
#include "cw_relax2d_tile.inc"
#undef CW_TILE_INDEX
#undef CW_INPUT_PARITY
#undef CW_TILE_PROLOGUE
}

// One tile of the device-driven batch: tile t, input in buffer in_par.
template <typename T, int S, int NW, int NR, bool PLAIN, bool GEN>
__device__ __forceinline__ void relax2d_reg_tile(const PassArgs& a, const int t_in, const int in_par) {
#define CW_TILE_INDEX t_in
#define CW_INPUT_PARITY in_par
#define CW_TILE_PROLOGUE 0
#include "cw_relax2d_tile.inc"
#undef CW_TILE_INDEX
#undef CW_INPUT_PARITY
#undef CW_TILE_PROLOGUE
}

// ---------------------------------------------------------------------------
// 3D red-black sweep streaming in z (one full sweep per launch, ping-pong).
//
// A workgroup owns a 128 x (NW*NR) tile in (x, y) (halo 2, interior
// 124 x (NW*NR - 4)) and a chunk of zc planes. Lanes hold column pairs, waves
// contiguous rows, as in relax2d_reg, and each plane's cells sit in registers
// by colour. Colour flips with z, so the z neighbours of st[c][i] in plane z
// are st[1-c][i] of planes z-1 and z+1: the same register slot. Step z loads
// plane z+1, updates red in plane z (old black around it), then black in plane
// z-1 (new red around it) and stores plane z-1: every plane is read and
// written once per sweep. The chunk recomputes red one plane before its first
// plane; in-plane halos are recomputed as in the 2D tiles.
// ---------------------------------------------------------------------------
template <typename T, int NR>
struct Plane {
    T st[2][NR];
};
template <typename T, int NR>
struct PlaneK {
    T k[2][NR];   // coefficient (NaN = held or outside); PLAIN: folded k or production
    T u[2][NR];   // PLAIN+GEN: uptake; smoother: lower bound (-inf without one)
};

template <typename T, bool PLAIN, bool GEN, int NR, int C, int P0>
__device__ __forceinline__ void plane_update(Plane<T, NR>& cur, const Plane<T, NR>& below_z,
                                             const Plane<T, NR>& above_z, const PlaneK<T, NR>& K,
                                             T above, T below, int lane, const int (&rx)[2],
                                             const int (&ry)[NR], int rz, T d, T decay, T omega,
                                             T omo, T source, bool track, int r0, int TYlo,
                                             int TYhi, unsigned long long* mv) {
#pragma unroll
    for (int i = 0; i < NR; i++) {
        const int q = P0 ^ (i & 1);
        const T o = cur.st[C ^ 1][i];
        const T across = q ? from_right(o) : from_left(o);
        const T up = i > 0 ? cur.st[C ^ 1][i - 1] : above;
        const T dn = i < NR - 1 ? cur.st[C ^ 1][i + 1] : below;
        const T zm = below_z.st[C ^ 1][i];
        const T zp = above_z.st[C ^ 1][i];
        const T kq = K.k[C][i];
        const T before = cur.st[C][i];
        const int real = rx[q] + ry[i] + rz;
        T after;
        if (PLAIN) {
            // x-1, x+1, y-1, y+1, z-1, z+1, in the reference's order
            T sum = T(0);
            if (q == 0) { sum += across; sum += o; } else { sum += o; sum += across; }
            sum += up;
            sum += dn;
            sum += zm;
            sum += zp;
            T num = d * sum;
            if (GEN) {
                num = num + source;
                num = num + kq;
                num = num - K.u[C][i];
            } else {
                num = num + kq;
            }
            const T bal = num / ((T)real * d + decay);
            after = omo * before + omega * bal;
            after = after > T(0) ? after : T(0);
        } else {
            const T sum = ((across + o) + (up + dn)) + (zm + zp);
            const T a = omega * d * rcp_t((T)real * d + decay);
            after = fma_t(omo, before, fma_t(a, sum, kq));
            after = max_t(after, K.u[C][i]);
        }
        const bool upd = kq == kq;
        if (PLAIN && track && upd) {
            const int x = 2 * lane + q;
            const int r = r0 + i;
            if (x >= 2 && x < 126 && r >= TYlo && r < TYhi) {
                const unsigned long long m = change_bits((double)__builtin_fabs(after - before));
                *mv = m > *mv ? m : *mv;
            }
        }
        cur.st[C][i] = upd ? after : before;
    }
}

template <typename T, int NW, int NR, bool PLAIN, bool GEN>
__global__ __launch_bounds__(NW * 64) void relax3d_stream(PassArgs a_in) {
    const PassArgs a = resolve(a_in);
    constexpr int RY = NW * NR;
    constexpr int TX = 124, TY = RY - 4;
    static_assert(NR >= 2 && TY > 0, "tile shape");
    using T2 = typename V2<T>::type;
    __shared__ T edge[2][NW][2][64];  // [slot][wave][first/last row][lane]: one colour

    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int p = a.tile_prob[t];
    const Prob& pr = a.probs[p];
    const int tl = a.tl;
    const int lt = t - pr.sw_base[tl];
    const int nxy = pr.ntx[tl] * pr.nty[tl];
    const int ch = lt / nxy, txy = lt % nxy;
    const int ti = txy % pr.ntx[tl], tj = txy / pr.ntx[tl];
    const int gxo = pr.xs + ti * TX - 2;
    const int gyo = pr.ys + tj * TY - 2;
    const int za = pr.zs + ch * pr.zc;
    const int zb = za + pr.zc - 1 < pr.z1 ? za + pr.zc - 1 : pr.z1;
    const int W = pr.W, Hh = pr.H, D = pr.D, P = pr.P;
    const int bx1 = pr.x1, by1 = pr.y1, bz1 = pr.z1;
    const size_t plane = (size_t)P * Hh;
    const int gx = gxo + 2 * lane;
    const bool colin = gx >= 0 && gx < W;
    const T* kk = (const T*)pr.k;
    const T* k2p = (const T*)(PLAIN ? pr.k2 : pr.lb);
    const bool has_k2 = PLAIN ? GEN : (k2p != nullptr);
    const T d = (T)pr.d, decay = (T)pr.decay, omega = (T)pr.omega, omo = (T)pr.omo;
    const T source = (T)pr.source;
    const int r0 = wv * NR;
    const int rx[2] = {(gx > 0) + (gx < W - 1), (gx + 1 > 0) + (gx + 1 < W - 1)};
    int ry[NR];
    bool rowin[NR];
#pragma unroll
    for (int i = 0; i < NR; i++) {
        const int gy = gyo + r0 + i;
        ry[i] = (gy > 0) + (gy < Hh - 1);
        rowin[i] = colin && gy >= 0 && gy < Hh;
    }
    const int base0 = (gxo + gyo + r0) & 1;  // colour of column 2l, row 0, plane 0
    const T* src = (const T*)pr.x[a.pass & 1];

    auto load_state = [&](Plane<T, NR>& pl, int z) {
        const int sw0 = (base0 + z) & 1;
#pragma unroll
        for (int i = 0; i < NR; i++) {
            T ve = T(0), vo = T(0);
            if (rowin[i] && z >= 0 && z < D) {
                const T2 v = *(const T2*)(src + (size_t)z * plane + (size_t)(gyo + r0 + i) * P + gx);
                ve = v.x; vo = v.y;
            }
            const bool sw = (sw0 ^ (i & 1)) != 0;
            pl.st[0][i] = sw ? vo : ve;
            pl.st[1][i] = sw ? ve : vo;
        }
    };
    auto load_coef = [&](PlaneK<T, NR>& K, int z) {
        const int sw0 = (base0 + z) & 1;
#pragma unroll
        for (int i = 0; i < NR; i++) {
            T ce = nan_of<T>(), co = nan_of<T>(), ue = PLAIN ? T(0) : neg_inf<T>(), uo = ue;
            if (rowin[i] && z >= 0 && z < D) {
                const size_t off = (size_t)z * plane + (size_t)(gyo + r0 + i) * P + gx;
                const T2 c = *(const T2*)(kk + off);
                ce = c.x; co = c.y;
                if (has_k2) { const T2 u = *(const T2*)(k2p + off); ue = u.x; uo = u.y; }
            }
            const bool sw = (sw0 ^ (i & 1)) != 0;
            K.k[0][i] = sw ? co : ce; K.k[1][i] = sw ? ce : co;
            K.u[0][i] = sw ? uo : ue; K.u[1][i] = sw ? ue : uo;
        }
    };

    Plane<T, NR> m2, m1, c0, p1;
    PlaneK<T, NR> km1, k0, kp1;
    // pipeline fill: planes za-2, za-1, za; coefficients za-1, za
    load_state(m2, za - 2);
    load_state(m1, za - 1);
    load_state(c0, za);
    load_coef(km1, za - 1);
    load_coef(k0, za);

    int par;
    if (!sweep_prologue(a, pr, p, lt, &par)) return;
    T* dst = (T*)pr.x[par ^ 1];

    unsigned long long mv = 0ull;
    const bool track = PLAIN && a.crit == CRIT_MAXCHANGE;
    int slot = 0;
    // red in one plane: publish the plane's other-colour edge rows, then update
    auto red_phase = [&](Plane<T, NR>& cur, const Plane<T, NR>& bz, const Plane<T, NR>& az,
                         const PlaneK<T, NR>& K, int z, bool mtrack) {
        edge[slot][wv][0][lane] = cur.st[1][0];
        edge[slot][wv][1][lane] = cur.st[1][NR - 1];
        __syncthreads();
        const T above = wv > 0 ? edge[slot][wv - 1][1][lane] : T(0);
        const T below = wv < NW - 1 ? edge[slot][wv + 1][0][lane] : T(0);
        slot ^= 1;
        const int rz = (z > 0) + (z < D - 1);
        const int p0 = (base0 + z) & 1;  // x parity of colour 0 on row 0 of this plane
        if (p0) plane_update<T, PLAIN, GEN, NR, 0, 1>(cur, bz, az, K, above, below, lane, rx, ry, rz, d, decay, omega, omo, source, mtrack, r0, 2, 2 + TY, &mv);
        else plane_update<T, PLAIN, GEN, NR, 0, 0>(cur, bz, az, K, above, below, lane, rx, ry, rz, d, decay, omega, omo, source, mtrack, r0, 2, 2 + TY, &mv);
    };
    auto black_phase = [&](Plane<T, NR>& cur, const Plane<T, NR>& bz, const Plane<T, NR>& az,
                           const PlaneK<T, NR>& K, int z, bool mtrack) {
        edge[slot][wv][0][lane] = cur.st[0][0];
        edge[slot][wv][1][lane] = cur.st[0][NR - 1];
        __syncthreads();
        const T above = wv > 0 ? edge[slot][wv - 1][1][lane] : T(0);
        const T below = wv < NW - 1 ? edge[slot][wv + 1][0][lane] : T(0);
        slot ^= 1;
        const int rz = (z > 0) + (z < D - 1);
        const int p1_ = ((base0 + z) & 1) ^ 1;  // x parity of colour 1 on row 0
        if (p1_) plane_update<T, PLAIN, GEN, NR, 1, 1>(cur, bz, az, K, above, below, lane, rx, ry, rz, d, decay, omega, omo, source, mtrack, r0, 2, 2 + TY, &mv);
        else plane_update<T, PLAIN, GEN, NR, 1, 0>(cur, bz, az, K, above, below, lane, rx, ry, rz, d, decay, omega, omo, source, mtrack, r0, 2, 2 + TY, &mv);
    };

    // red of plane za-1 (needed by black of za), only inside the swept box
    if (za - 1 >= pr.zs) red_phase(m1, m2, c0, km1, za - 1, false);
    for (int z = za; z <= zb + 1; z++) {
        load_state(p1, z + 1);
        load_coef(kp1, z + 1);
        if (z <= bz1) red_phase(c0, m1, p1, k0, z, track && z <= zb);
        if (z - 1 >= za) {
            black_phase(m1, m2, c0, km1, z - 1, track);
            // store plane z-1: interior rows and pairs inside the swept box
            if (lane >= 1 && lane < 63 && gx <= bx1) {
                const int zz = z - 1;
                const int sw0 = (base0 + zz) & 1;
#pragma unroll
                for (int i = 0; i < NR; i++) {
                    const int r = r0 + i;
                    if (r < 2 || r >= 2 + TY) continue;
                    const int gy = gyo + r;
                    if (gy > by1) continue;
                    const bool sw = (sw0 ^ (i & 1)) != 0;
                    T2 v;
                    v.x = sw ? m1.st[1][i] : m1.st[0][i];
                    v.y = sw ? m1.st[0][i] : m1.st[1][i];
                    *(T2*)(dst + (size_t)zz * plane + (size_t)gy * P + gx) = v;
                }
            }
        }
        m2 = m1; m1 = c0; c0 = p1;
        km1 = k0; k0 = kp1;
    }
    if (track) {
        unsigned long long mvv[1] = {mv};
        block_moved<NW, 1>(a.ctl[p], a.pass, mvv);
    }
}

// S red-black sweeps per pass, for the FP32 volume smoother, as a pipeline
// of 2S half-sweep phases in one z stream: at loop step z, phase h (colour
// h & 1 of sweep h / 2) updates plane z - h, in the order h = 0, 1, ..., so
// every phase reads its plane's z neighbours at the level the sequential
// sweeps give them (plane z-h+1 was just updated by phase h-1, plane z-h-1
// not yet by phase h+1). Plane z-2S+1 then holds its final values and is
// stored. The halo is 2S cells in x and y; phase h covers the chunk plus
// 2S-1-h planes either side, so the stored (128-4S) x (NW*NR-4S) x zc block
// is exact, and red-black SOR fixes every cell's operands: the bits are the
// one-sweep smoother's.
//
// One barrier a step. A wave's rows at the edge of its block need the
// neighbouring waves' edge rows, and every phase takes them from the step
// before: phase h of step z needs the other colour of plane z-h as phase h-1
// left it, and phase h-1 updated plane z-h at step z-1 (phase 0 needs plane
// z's original values, published at step z-1 when that plane is the next
// one). So each wave publishes its edge rows after every phase into the slot
// of its step, reads those of the previous step, and one barrier at the end
// of the step orders the two (double-buffered by step parity). The next
// plane's loads are issued a step ahead.
// Slot of plane z+2-j in a ring of L planes at a step with z mod L == R.
template <int L, int R, int J>
__device__ __forceinline__ constexpr int ring_slot() { return ((R + 2 - J) % L + 2 * L) % L; }

template <int I, int N, typename F>
__device__ __attribute__((always_inline)) void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// The FP32 volume smoother's update of one colour of a plane whose cells are
// kept by column parity (st[q][i]: row i, column 2l + q), the colour's parity
// on row 0 being P0, with alpha = omega d / (real d + decay) given per parity
// and row: the same arithmetic, in the same order, as plane_update<float,
// false, false>, so the same bits. The other colour's cells in the row are
// the lane's other parity; the cells above, below and in the planes either
// side are at the same parity.
template <int NR, int P0>
__device__ __forceinline__ void pipe_update(Plane<float, NR>& cur, const Plane<float, NR>& below_z,
                                            const Plane<float, NR>& above_z,
                                            const PlaneK<float, NR>& K, float above, float below,
                                            const float (&al)[2][NR], float omo) {
#pragma unroll
    for (int i = 0; i < NR; i++) {
        const int q = P0 ^ (i & 1);
        const float o = cur.st[q ^ 1][i];
        const float across = q ? from_right(o) : from_left(o);
        const float up = i > 0 ? cur.st[q][i - 1] : above;
        const float dn = i < NR - 1 ? cur.st[q][i + 1] : below;
        const float zm = below_z.st[q][i];
        const float zp = above_z.st[q][i];
        const float kq = K.k[q][i];
        const float before = cur.st[q][i];
        const float sum = ((across + o) + (up + dn)) + (zm + zp);
        float after = fma_t(omo, before, fma_t(al[q][i], sum, kq));
        after = max_t(after, K.u[q][i]);
        cur.st[q][i] = kq == kq ? after : before;
    }
}

template <int S, int NW, int NR, bool LB>
__global__ __launch_bounds__(NW * 64) void relax3d_pipe(PassArgs a_in) {
    using T = float;
    const PassArgs a = resolve(a_in);
    constexpr int RY = NW * NR;
    constexpr int HALO = 2 * S;
    constexpr int TX = 128 - 2 * HALO, TY = RY - 2 * HALO;
    constexpr int NPH = 2 * S;        // phases a step
    constexpr int L = NPH + 4;        // ring length (even): planes z+2 .. z-2S live at step z
    static_assert(NR >= 2 && NR % 2 == 0 && TY > 0 && S >= 1, "tile shape");
    __shared__ T edge[2][NPH][NW][2][64];  // [step parity][phase][wave][first/last row][lane]

    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int p = a.tile_prob[t];
    const Prob& pr = a.probs[p];
    const int tl = a.tl;
    const int lt = t - pr.sw_base[tl];
    const int nxy = pr.ntx[tl] * pr.nty[tl];
    const int ch = lt / nxy, txy = lt % nxy;
    const int ti = txy % pr.ntx[tl], tj = txy / pr.ntx[tl];
    // Regions start on even columns and rows (pr.xs is even; the y tiling
    // of this kernel starts at the even row at or before the span's first),
    // so column 2l of row 0 of plane z has colour z & 1 in every region.
    const int gxo = pr.xs + ti * TX - HALO;
    const int gyo = (pr.ys & ~1) + tj * TY - HALO;
    const int za = pr.zs + ch * pr.zc;
    const int zb = za + pr.zc - 1 < pr.z1 ? za + pr.zc - 1 : pr.z1;
    const int W = pr.W, Hh = pr.H, D = pr.D, P = pr.P;
    // byte offsets fit 32 bits (the host takes this kernel only for arrays
    // under 4 GiB), so every load is a base plus a 32-bit lane offset
    const uint32_t plane_bytes = (uint32_t)P * (uint32_t)Hh * 4u;
    const int gx = gxo + 2 * lane;
    const bool colin = gx >= 0 && gx < W;
    typedef float F2 __attribute__((ext_vector_type(2)));
    using GB = __attribute__((address_space(1))) const unsigned char;
    using GT2 = __attribute__((address_space(1))) const F2;
    typedef __attribute__((address_space(1))) F2 GF2;
    // Pointers in the global address space: their loads count against the
    // vector-memory counter only, so LDS waits do not wait for the prefetch.
    GB* kk = (GB*)pr.k;
    GB* lbp = (GB*)pr.lb;  // LB: the launch's fields have lower bounds
    const bool has_lb = LB && lbp != nullptr;
    // the problem's bounds in registers: the stores below would otherwise
    // make the compiler reload them (and wait for every load in flight)
    const int zlo = pr.zs, zhi = pr.z1, xhi = pr.x1, yhi = pr.y1;
    const T d = (T)pr.d, decay = (T)pr.decay, omega = (T)pr.omega, omo = (T)pr.omo;
    const int r0 = wv * NR;
    const int rx[2] = {(gx > 0) + (gx < W - 1), (gx + 1 > 0) + (gx + 1 < W - 1)};
    int ry[NR];
    bool rowin[NR];
    uint32_t rowoff[NR];  // byte offset of the lane's pair in its row of a plane
#pragma unroll
    for (int i = 0; i < NR; i++) {
        const int gy = gyo + r0 + i;
        ry[i] = (gy > 0) + (gy < Hh - 1);
        rowin[i] = colin && gy >= 0 && gy < Hh;
        rowoff[i] = rowin[i] ? ((uint32_t)gy * (uint32_t)P + (uint32_t)gx) * 4u : 0u;
    }
    // alpha of interior planes (z neighbours on both sides) per parity and row
    T a2[2][NR];
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
        for (int i = 0; i < NR; i++) a2[q][i] = omega * d * rcp_t((T)(rx[q] + ry[i] + 2) * d + decay);
    GB* src = (GB*)pr.x[a.pass & 1];

    // Raw loads of a plane's pairs as stored (a row outside the field reads
    // element 0); the masks are applied when the plane is sorted, a step
    // later, so nothing waits for a load before its plane is needed.
    auto load_state = [&](F2 (&v)[NR], int z) __attribute__((always_inline)) {
        const bool zin = z >= 0 && z < D;
        const uint32_t zo = zin ? (uint32_t)z * plane_bytes : 0u;
#pragma unroll
        for (int i = 0; i < NR; i++) v[i] = *(GT2*)(src + (zin ? zo + rowoff[i] : 0u));
    };
    auto load_coef = [&](F2 (&c)[NR], F2 (&u)[NR], int z) __attribute__((always_inline)) {
        const bool zin = z >= 0 && z < D;
        const uint32_t zo = zin ? (uint32_t)z * plane_bytes : 0u;
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const uint32_t off = zin ? zo + rowoff[i] : 0u;
            c[i] = *(GT2*)(kk + off);
            if (has_lb) u[i] = *(GT2*)(lbp + off);
        }
    };
    // Cells stay by column parity as loaded (st[q][i]: row i, column 2l + q);
    // sorting a plane is only masking what lies outside the field.
    auto sort_state = [&](Plane<T, NR>& pl, const F2 (&v)[NR], int z) __attribute__((always_inline)) {
        const bool zin = z >= 0 && z < D;
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const bool in = rowin[i] && zin;
            pl.st[0][i] = in ? v[i].x : T(0);
            pl.st[1][i] = in ? v[i].y : T(0);
        }
    };
    auto sort_coef = [&](PlaneK<T, NR>& K, const F2 (&c)[NR], const F2 (&u)[NR], int z) __attribute__((always_inline)) {
        const bool zin = z >= 0 && z < D;
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const bool in = rowin[i] && zin;
            K.k[0][i] = in ? c[i].x : nan_of<T>();
            K.k[1][i] = in ? c[i].y : nan_of<T>();
            K.u[0][i] = has_lb && in ? u[i].x : neg_inf<T>();
            K.u[1][i] = has_lb && in ? u[i].y : neg_inf<T>();
        }
    };

    // A ring of L planes: plane p sits in slot (p - zr) mod L, so with the
    // step's (z - zr) mod L a compile-time constant (the loop below is
    // unrolled L times) every slot index is one and no register moves. zr is
    // the even plane at or before the first step z0; the loop starts there
    // and runs whole blocks of L steps, the steps outside [z0, zend] loading,
    // updating and storing nothing. Every step sorts its planes whatever it
    // does, so no slot keeps a stale plane alive across the loop. At step z
    // the ring holds the states of planes z+2 (sorted at the step's start,
    // loaded a step before) .. z-2S and the coefficients of planes z ..
    // z-2S+1 (plane z's sorted at the step's start, loaded a step before).
    // Phase h starts at step z0 + h.
    const int z0 = za - (NPH - 1);
    const int zend = zb + (NPH - 1);
    const int zr = z0 & ~1;
    Plane<T, NR> q[L];
    PlaneK<T, NR> K[L];
    F2 nv[NR], nc[NR], nu[NR] = {};
    {
        F2 v[NR];
        load_state(v, zr + 1); sort_state(q[ring_slot<L, 0, 1>()], v, zr + 1);
        load_state(v, zr); sort_state(q[ring_slot<L, 0, 2>()], v, zr);
        load_state(v, zr - 1); sort_state(q[ring_slot<L, 0, 3>()], v, zr - 1);
        // plane zr's original colour 1 (parity 1 on row 0: zr is even) for
        // phase 0 of step zr
        const Plane<T, NR>& first = q[ring_slot<L, 0, 2>()];
        edge[1][0][wv][0][lane] = first.st[1][0];
        edge[1][0][wv][1][lane] = first.st[0][NR - 1];
    }
    load_state(nv, zr + 2);
    load_coef(nc, nu, zr);

    int par;
    if (!sweep_prologue(a, pr, p, lt, &par)) return;
    __attribute__((address_space(1))) unsigned char* dst =
        (__attribute__((address_space(1))) unsigned char*)pr.x[par ^ 1];
    lds_barrier();

    // Phase h of a step: colour h & 1 of plane z, whose cells of that colour
    // sit at parity P0 on row 0.
    auto phase = [&](auto p0c, int h, Plane<T, NR>& cur, const Plane<T, NR>& bz,
                     const Plane<T, NR>& az, const PlaneK<T, NR>& Kc, int z, int rd, int wr)
        __attribute__((always_inline)) {
        constexpr int P0 = decltype(p0c)::value;
        const T above = edge[rd][h][wv > 0 ? wv - 1 : wv][1][lane];
        const T below = edge[rd][h][wv < NW - 1 ? wv + 1 : wv][0][lane];
        const T ab = wv > 0 ? above : T(0);
        const T bl = wv < NW - 1 ? below : T(0);
        const int rz = (z > 0) + (z < D - 1);
        if (rz == 2) {
            pipe_update<NR, P0>(cur, bz, az, Kc, ab, bl, a2, omo);
        } else {
            // a plane on a z face of the field: fewer real neighbours
            T af[2][NR];
#pragma unroll
            for (int qq = 0; qq < 2; qq++)
#pragma unroll
                for (int i = 0; i < NR; i++)
                    af[qq][i] = omega * d * rcp_t((T)(rx[qq] + ry[i] + rz) * d + decay);
            pipe_update<NR, P0>(cur, bz, az, Kc, ab, bl, af, omo);
        }
        // this colour of the plane as the next phase reads it at the next step
        if (h + 1 < NPH) {
            edge[wr][h + 1][wv][0][lane] = cur.st[P0][0];
            edge[wr][h + 1][wv][1][lane] = cur.st[P0 ^ 1][NR - 1];
        }
    };

    // Step z, with (z - zr) mod L == RC: every ring slot is a compile-time
    // constant, and so is every colour's parity, since zr and L are even:
    // plane z-h's colour h & 1 sits at parity (z - h + h) & 1 == RC & 1 on
    // row 0 for every h.
    auto step = [&](auto rc, int z) __attribute__((always_inline)) {
        constexpr int RC = decltype(rc)::value;
        constexpr int P0 = RC & 1;
        const int rd = (z - 1) & 1, wr = z & 1;
        const bool live = z >= z0 && z <= zend;
        // the state loaded a step ago becomes plane z+2, the coefficients
        // plane z's; the next ones go in flight while some step still needs
        // them (plane p's state from step p-1, its coefficients from step p)
        sort_state(q[ring_slot<L, RC, 0>()], nv, z + 2);
        sort_coef(K[ring_slot<L, RC, 2>()], nc, nu, z);
        if (z + 3 <= zend + 1) load_state(nv, z + 3);
        if (z + 1 <= zend) load_coef(nc, nu, z + 1);
        // plane z+1's original colour 1 (parity P0 on row 0), for phase 0 of
        // the next step
        {
            const Plane<T, NR>& nx = q[ring_slot<L, RC, 1>()];
            edge[wr][0][wv][0][lane] = nx.st[P0][0];
            edge[wr][0][wv][1][lane] = nx.st[P0 ^ 1][NR - 1];
        }
        static_for<0, NPH>([&](auto hc) __attribute__((always_inline)) {
            constexpr int h = decltype(hc)::value;
            const int zh = z - h;
            if (live && zh >= zlo && zh <= zhi && zh >= za - (NPH - 1 - h))
                phase(std::integral_constant<int, P0>{}, h, q[ring_slot<L, RC, h + 2>()],
                      q[ring_slot<L, RC, h + 3>()], q[ring_slot<L, RC, h + 1>()],
                      K[ring_slot<L, RC, h + 2>()], zh, rd, wr);
        });
        // plane z-2S+1 is final: interior rows and pairs inside the swept box
        const int zs = z - (NPH - 1);
        if (zs >= za && zs <= zb && lane >= S && lane < 64 - S && gx <= xhi) {
            const Plane<T, NR>& fin = q[ring_slot<L, RC, NPH + 1>()];
#pragma unroll
            for (int i = 0; i < NR; i++) {
                const int r = r0 + i;
                if (r < HALO || r >= HALO + TY) continue;
                const int gy = gyo + r;
                if (gy > yhi) continue;
                F2 v;
                v.x = fin.st[0][i];
                v.y = fin.st[1][i];
                *(GF2*)(dst + ((uint32_t)zs * plane_bytes + rowoff[i])) = v;
            }
        }
        lds_barrier();
    };

    for (int zz = zr; zz <= zend; zz += L)
        static_for<0, L>([&](auto rc) __attribute__((always_inline)) {
            step(rc, zz + decltype(rc)::value);
        });
}

// ---------------------------------------------------------------------------
// Elementwise tiling: 128 cells in x (64 pairs) by 4 rows of one plane.
// ---------------------------------------------------------------------------
struct ElTile {
    int gx, gy, gz;   // pair start column, row, plane
    bool ok;
};

__device__ __forceinline__ ElTile el_tile(const Prob& pr, int t) {
    const int lt = t - pr.el_base;
    const int tx = lt % pr.elx;
    const int rest = lt / pr.elx;
    const int ty = rest % pr.ely;
    const int tz = rest / pr.ely;
    ElTile e;
    e.gx = pr.xs + tx * 128 + 2 * (threadIdx.x & 63);
    e.gy = pr.ys + ty * 4 + (threadIdx.x >> 6);
    e.gz = pr.zs + tz;
    e.ok = e.gx <= pr.x1 && e.gy <= pr.y1 && e.gz <= pr.z1;
    return e;
}

// 3D red or black half sweep in place (x[0] == x[1]): cells of one colour read
// only the other colour, which this launch never writes, so the result equals
// the reference's read-old/write-new half sweep (pde.rs:322-328). Red and
// black launches of one pass carry the same pass index.
template <typename T, bool PLAIN>
__global__ __launch_bounds__(256) void half3d(PassArgs a_in, int colour) {
    const PassArgs a = resolve(a_in);
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int p = a.tile_prob[t];
    const Prob& pr = a.probs[p];
    const int lt = t - pr.el_base;
    int par;
    if (!sweep_prologue(a, pr, p, lt, &par)) return;
    unsigned long long mv[1] = {0ull};
    const ElTile et = el_tile(pr, t);
    if (et.ok) {
        const int W = pr.W, Hh = pr.H, D = pr.D, P = pr.P;
        const size_t plane = (size_t)P * Hh;
        T* xs = (T*)pr.x[0];
        const T* kk = (const T*)pr.k;
        const T* kk2 = (const T*)pr.k2;
        const T* lbp = (const T*)pr.lb;
        const bool general = PLAIN && kk2 != nullptr;
        const bool has_lb = !PLAIN && lbp != nullptr;
        const T d = (T)pr.d, decay = (T)pr.decay, omega = (T)pr.omega, omo = (T)pr.omo;
        const int q = (colour + et.gx + et.gy + et.gz) & 1;
        const int x = et.gx + q;
        if (x < W && x <= pr.x1) {
            const size_t i = (size_t)et.gz * plane + (size_t)et.gy * P + x;
            const T kq = kk[i];
            if (kq == kq) {
                const T before = xs[i];
                T sum = T(0);
                if (x > 0) sum += xs[i - 1];
                if (x + 1 < W) sum += xs[i + 1];
                if (et.gy > 0) sum += xs[i - P];
                if (et.gy + 1 < Hh) sum += xs[i + P];
                if (et.gz > 0) sum += xs[i - plane];
                if (et.gz + 1 < D) sum += xs[i + plane];
                const int real = real_count(x, et.gy, et.gz, W, Hh, D);
                T after;
                if (PLAIN) {
                    // reference order (FP64, no contraction)
                    T num = d * sum;
                    if (general) {
                        num = num + (T)pr.source;
                        num = num + kq;
                        num = num - kk2[i];
                    } else {
                        num = num + kq;
                    }
                    const T bal = num / ((T)real * d + decay);
                    after = omo * before + omega * bal;
                    after = after > T(0) ? after : T(0);
                } else {
                    // pre-scaled linear smoother: kq = omega r / A
                    const T alpha = omega * d * rcp_t((T)real * d + decay);
                    after = fma_t(omo, before, fma_t(alpha, sum, kq));
                    if (has_lb) { const T l = lbp[i]; after = after > l ? after : l; }
                }
                xs[i] = after;
                if (PLAIN) mv[0] = change_bits((double)__builtin_fabs(after - before));
            }
        }
    }
    if (PLAIN && a.crit == CRIT_MAXCHANGE) block_moved<4, 1>(a.ctl[p], a.pass, mv);
}

// FP64 residual of the current state (2D and 3D), elementwise tiling, into
// ring slot 0 of the partials. Per free cell: b = (source + production) -
// uptake (or the folded k), r = (b + d*sum) - A*c, projected to the natural
// residual; b_eff adds the held neighbours. With write32 it also writes the
// FP32 smoother's right-hand side omega*r/A, the correction's lower bound -c
// (rounded toward zero) and zeroes e[0] for the FP32 inner solve.
static __global__ __launch_bounds__(256) void resid64(PassArgs a, int write32) {
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int p = a.tile_prob[t];
    const Prob& pr = a.probs[p];
    double r2 = 0.0, b2 = 0.0;
    if (a.ctl[p].status != ST_RUNNING) return;
    const ElTile et = el_tile(pr, t);
    if (et.ok) {
        // Everything the stores need is read first: a store may alias the
        // problem record as far as the compiler knows, so a field read after
        // one would be reloaded behind it.
        const int W = pr.W, Hh = pr.H, D = pr.D, P = pr.P;
        const size_t plane = (size_t)P * Hh;
        const double* c = pr.c64;
        const double* kk = (const double*)pr.k;
        const double* kk2 = (const double*)pr.k2;
        float* r32 = pr.r32;
        float* e0 = pr.e0;
        float* lb32 = pr.lb32;
        const double d = pr.d, decay = pr.decay, omega = pr.omega, source = pr.source;
        const bool project = a.project;
        size_t at[2] = {0, 0};
        bool live[2] = {false, false}, held[2] = {false, false};
        float rs[2] = {0.0f, 0.0f}, ls[2] = {0.0f, 0.0f};
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int x = et.gx + q;
            if (x >= W) continue;
            const size_t i = (size_t)et.gz * plane + (size_t)et.gy * P + x;
            at[q] = i;
            live[q] = true;
            const double kq = kk[i];
            if (kq != kq) {
                held[q] = true;
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
            CW_NB(et.gy > 0, i - P)
            CW_NB(et.gy + 1 < Hh, i + P)
            CW_NB(et.gz > 0, i - plane)
            CW_NB(et.gz + 1 < D, i + plane)
#undef CW_NB
            const double A = (double)real * d + decay;
            const double b = kk2 ? ((source + kq) - kk2[i]) : kq;
            const double r = (b + d * sum) - A * ci;
            double rt = r;
            if (project && A > 0.0 && !(ci + r / A > 0.0)) rt = -A * ci;
            const double be = b + d * hsum;
            r2 += rt * rt;
            b2 += be * be;
            if (write32) {
                // the FP32 smoother's pre-scaled right-hand side omega r / A
                rs[q] = (float)(omega * r / A);
                ls[q] = __double2float_rz(-ci);
            }
        }
        if (write32) {
#pragma unroll
            for (int q = 0; q < 2; q++) {
                if (!live[q]) continue;
                const size_t i = at[q];
                r32[i] = held[q] ? __int_as_float(0x7fc00000) : rs[q];
                e0[i] = 0.0f;
                if (lb32 && !held[q]) lb32[i] = ls[q];
            }
        }
    }
    block_partial(a.partials, t, r2, b2);
}

// c <- max(c + e, 0) on the free cells, e from the buffer the inner plan ended in.
static __global__ __launch_bounds__(256) void correct64(PassArgs a) {
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int p = a.tile_prob[t];
    const Prob& pr = a.probs[p];
    if (a.ctl[p].status != ST_RUNNING) return;
    const ElTile et = el_tile(pr, t);
    if (!et.ok) return;
    const int W = pr.W, P = pr.P;
    const size_t plane = (size_t)P * pr.H;
    const Ctl& ic = a.inner_ctl[pr.inner_idx];
    const float* e = (ic.parity & 1) ? pr.e1 : pr.e0;
    const double* kk = (const double*)pr.k;
    double* c64 = pr.c64;
    // both cells are read before either is written (see resid64)
    size_t at[2] = {0, 0};
    bool live[2] = {false, false};
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int x = et.gx + q;
        if (x >= W) continue;
        const size_t i = (size_t)et.gz * plane + (size_t)et.gy * P + x;
        if (kk[i] != kk[i]) continue;
        at[q] = i;
        live[q] = true;
        v[q] = c64[i] + (double)e[i];
    }
#pragma unroll
    for (int q = 0; q < 2; q++)
        if (live[q]) c64[at[q]] = v[q] > 0.0 ? v[q] : 0.0;
}

// One workgroup: records verdicts for the host. DECIDE_PASS assesses the last
// launch of a chunk (as the next sweep launch would); DECIDE_RESID turns the
// resid64 partials into r2, b2 and stops converged problems; DECIDE_REPLAY_END
// closes replayed problems.
// Verdict of one running problem from its residual sums (DECIDE_RESID) or
// from launch L-1's slots (DECIDE_PASS); one thread calls it.
__device__ inline void decide_one(const PassArgs& a, Ctl& c, bool have_r2, double r2, double b2) {
    if (a.mode == DECIDE_RESID) {
        c.r2 = r2;
        c.b2 = b2;
        if (!(r2 == r2) || r2 > 1.7e308) {
            c.nonfinite = 1;
            c.status = ST_MAXED;
            c.sweeps = a.sweeps_before;
        } else if (a.crit == CRIT_RESID && r2 <= a.tol * a.tol * (b2 > 0.0 ? b2 : 1.0)) {
            c.status = ST_DONE;
            c.sweeps = a.sweeps_before;
            c.parity = a.pass & 1;
        }
        return;
    }
    const Verdict v = assess(a, c, have_r2, r2);
    if (v.status != ST_RUNNING) {
        if (have_r2) c.r2 = r2;
        record(c, v);
    }
}

// Partials a problem's verdict sums, or cnt = 0 when it needs none.
__device__ inline void decide_partials(const PassArgs& a, const Prob& pr, bool* have_r2,
                                       const double** part, int* base, int* cnt) {
    *part = a.partials;
    *base = 0;
    *cnt = 0;
    *have_r2 = false;
    if (a.mode == DECIDE_RESID) {
        *base = pr.el_base;
        *cnt = pr.el_count;
        *have_r2 = true;
    } else if (a.prev_check && a.crit == CRIT_RESID && a.pass > 0) {
        *part = a.partials + ((a.pass - 1) & 1) * a.part_stride;
        *base = pr.sw_base[a.prev_tl];
        *cnt = pr.sw_count[a.prev_tl];
        *have_r2 = true;
    }
}

constexpr int kDecideThreads = 1024;
constexpr int kDecideShared = 4;  // up to this many problems: the whole block sums each one

// One workgroup of kDecideThreads: records verdicts for the host.
// DECIDE_PASS assesses the last launch of a chunk (as the next sweep launch
// would); DECIDE_RESID turns the resid64 partials into r2, b2 and stops
// converged problems; DECIDE_REPLAY_END closes replayed problems. A few
// problems are summed by the whole block in turn (a large field has tens of
// thousands of partials); many are summed one per wave. Either way each sum
// has a fixed order.
static __global__ __launch_bounds__(kDecideThreads) void decide(PassArgs a_in) {
    const PassArgs a = resolve(a_in);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (a.mode != DECIDE_REPLAY_END && a.nprob <= kDecideShared) {
        __shared__ double sr[kDecideThreads / 64], sb[kDecideThreads / 64];
        for (int p = 0; p < a.nprob; p++) {
            const Prob& pr = a.probs[p];
            Ctl& c = a.ctl[p];
            const bool running = c.status == ST_RUNNING;
            bool have_r2;
            const double* part;
            int base, cnt;
            decide_partials(a, pr, &have_r2, &part, &base, &cnt);
            double r = 0.0, b = 0.0;
            if (running)
                for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
                    r += part[2 * (base + t)];
                    b += part[2 * (base + t) + 1];
                }
            r = wave_sum(r);
            b = wave_sum(b);
            if (lane == 0) {
                sr[wv] = r;
                sb[wv] = b;
            }
            __syncthreads();
            if (threadIdx.x == 0 && running) {
                double r2 = 0.0, b2 = 0.0;
                for (int w = 0; w < nw; w++) {
                    r2 += sr[w];
                    b2 += sb[w];
                }
                decide_one(a, c, have_r2, r2, b2);
            }
            __syncthreads();
        }
        return;
    }
    for (int p = wv; p < a.nprob; p += nw) {
        const Prob& pr = a.probs[p];
        Ctl& c = a.ctl[p];
        if (a.mode == DECIDE_REPLAY_END) {
            if (lane == 0 && c.status == ST_REPLAY) {
                c.sweeps += c.replay;
                c.parity = (c.parity + c.replay) & 1;
                c.replay = 0;
                c.status = ST_DONE;
            }
            continue;
        }
        if (c.status != ST_RUNNING) continue;
        bool have_r2;
        const double* part;
        int base, cnt;
        decide_partials(a, pr, &have_r2, &part, &base, &cnt);
        double r2 = 0.0, b2 = 0.0;
        if (cnt) problem_sums(part, base, cnt, &r2, &b2);
        if (lane == 0) decide_one(a, c, have_r2, r2, b2);
    }
}

// Adjoint set-up over the swept box: rhs = c - target and the smoother's
// pre-scaled beta = omega * rhs / A on free cells (NaN on held ones), and
// lambda = 0 in both buffers.
static __global__ __launch_bounds__(256) void adjoint_setup(PassArgs a, const double* c,
                                                     const double* target, double* rhs,
                                                     double* beta, double* lam0, double* lam1) {
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int p = a.tile_prob[t];
    const Prob& pr = a.probs[p];
    const ElTile et = el_tile(pr, t);
    if (!et.ok) return;
    const int W = pr.W, Hh = pr.H, D = pr.D, P = pr.P;
    const size_t plane = (size_t)P * Hh;
    const double* kk = (const double*)pr.k;
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int x = et.gx + q;
        if (x >= W) continue;
        const size_t i = (size_t)et.gz * plane + (size_t)et.gy * P + x;
        lam0[i] = 0.0;
        lam1[i] = 0.0;
        if (kk[i] != kk[i]) {
            rhs[i] = nan_of<double>();
            beta[i] = nan_of<double>();
            continue;
        }
        const int real = real_count(x, et.gy, et.gz, W, Hh, D);
        const double A = (double)real * pr.d + pr.decay;
        const double r = c[i] - target[i];
        rhs[i] = r;
        beta[i] = pr.omega * r / A;
    }
}

// Adjoint gradient partials per elementwise tile (r2 slot: dL/dD, b2 slot:
// dL/dg from the free cells), for the closed forms
//   dL/dD = sum_i lambda_i sum_{real nbrs j} (c_j - c_i),
//   dL/dg = D sum_i lambda_i #held_nbrs(i)   (+ the held cells' own term).
static __global__ __launch_bounds__(256) void adjoint_grads(PassArgs a, const double* c,
                                                     const double* lam, const double* kk) {
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int p = a.tile_prob[t];
    const Prob& pr = a.probs[p];
    double gd = 0.0, gg = 0.0;
    const ElTile et = el_tile(pr, t);
    if (et.ok) {
        const int W = pr.W, Hh = pr.H, D = pr.D, P = pr.P;
        const size_t plane = (size_t)P * Hh;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int x = et.gx + q;
            if (x >= W) continue;
            const size_t i = (size_t)et.gz * plane + (size_t)et.gy * P + x;
            if (kk[i] != kk[i]) continue;
            const double ci = c[i];
            double ds = 0.0, nh = 0.0;
#define CW_NB(cond, j)                       \
    if (cond) {                              \
        ds += c[j] - ci;                     \
        if (kk[j] != kk[j]) nh += 1.0;       \
    }
            CW_NB(x > 0, i - 1)
            CW_NB(x + 1 < W, i + 1)
            CW_NB(et.gy > 0, i - P)
            CW_NB(et.gy + 1 < Hh, i + P)
            CW_NB(et.gz > 0, i - plane)
            CW_NB(et.gz + 1 < D, i + plane)
#undef CW_NB
            gd += lam[i] * ds;
            gg += pr.d * lam[i] * nh;
        }
    }
    block_partial(a.partials, t, gd, gg);
}

// dL/du_k = -sum of lambda over the voxels whose uptake is cell k's rate,
// one thread per cell over a CSR list, in a fixed order.
static __global__ __launch_bounds__(256) void cell_sums(const double* lam, const long long* off,
                                                 const long long* idx, int ncells,
                                                 double* out) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= ncells) return;
    double s = 0.0;
    for (long long j = off[k]; j < off[k + 1]; j++) s += lam[idx[j]];
    out[k] = -s;
}

// Every cell of the box (not the row padding) set to v: ScalarField::fill
// (pde.rs:212-215) on the device.
static __global__ __launch_bounds__(256) void fill_cells(double* c, int W, long rows, int P, double v) {
    const long n = rows * (long)W;
    for (long i = blockIdx.x * 256l + threadIdx.x; i < n; i += (long)gridDim.x * 256l) {
        const long r = i / W, x = i - r * W;
        c[r * P + x] = v;
    }
}

}  // namespace cw
