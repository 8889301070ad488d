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
// lines in that XCD's L2). Each problem has a device control block (Ctl) that
// the last workgroup of a launch advances: pass parity, sweep count, status.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cw {

constexpr int kMaxS = 4;
constexpr int kWave = 64;

enum Status : int { ST_RUNNING = 0, ST_DONE = 1, ST_MAXED = 2, ST_REPLAY = 3, ST_IDLE = 4 };
enum Crit : int { CRIT_NONE = 0, CRIT_MAXCHANGE = 1, CRIT_RESID = 2, CRIT_INNER = 3 };
enum Mode : int {
    MODE_SWEEP = 0,     // a pass of s full sweeps; finalize per criterion
    MODE_REPLAY = 1,    // one sweep for problems rolled back by an exact stop
    MODE_HALF_RED = 2,  // 3D red half sweep: no finalize
    MODE_HALF_BLACK = 3,// 3D black half sweep: finalize one sweep
    MODE_RESID = 4,     // residual pass: r2, b2 per problem
    MODE_PLAIN = 5      // no finalize (elementwise helpers)
};

// One field in a launch plan. Arrays are pitched: row stride P elements
// (P even, P >= W), plane stride P*H. Cells in the padding hold 0 in state
// arrays and NaN in coefficient arrays.
struct Prob {
    int W, H, D, P;
    int xs, ys, zs;        // tile-grid origin (xs even): first swept column/row/plane
    int x1, y1, z1;        // last swept column/row/plane (inclusive)
    int ntx, nty, ntz;     // sweep tiles (2D: ntz = 1)
    int sw_base, sw_count; // sweep tiling: first tile, tile count
    int el_base, el_count; // elementwise tiling (128 x 4 cells per plane)
    int elx, ely;          // elementwise tiles along x, y
    int in_place;          // 3D half-sweep kernels: x[0] == x[1]
    void* x[2];            // state ping-pong (T)
    const void* k;         // coefficient (T): NaN = held. FP64 plain: production
                           // minus uptake when at most one is nonzero; else production
    const void* k2;        // FP64 general form: uptake (else null)
    const void* lb;        // FP32 inner: lower bound of the correction (or null)
    double d, decay, omega, omo, source;  // omo = 1 - omega as the host rounds it
    float inv_den[8];      // FP32: 1/(real*d + decay), real = 0..6
    // FP64 refinement arrays (FP32_REFINE only)
    double* c64;           // the FP64 state being corrected
    float* e0;             // correction buffers of the inner plan (e[0], e[1])
    float* e1;
    float* r32;            // FP32 defect right-hand side (NaN = held)
    float* lb32;           // FP32 lower bound of the correction (or null)
    int inner_idx;         // this problem's index in the inner plan's Ctl
    int pad_;
};

struct Ctl {
    int status, pass, sweeps, replay;
    unsigned long long moved[kMaxS];
    double r2, b2, r2_first;
    int checks, nonfinite;
};

struct PassArgs {
    const Prob* probs;
    Ctl* ctl;
    const int* tile_prob;
    double* partials;   // 2 doubles per tile
    unsigned* counter;
    int nprob, ntiles;
    int crit, check, mode, max_sweeps, s;
    int tiling;         // 0 sweep tiling, 1 elementwise tiling
    double tol;
    const Ctl* inner_ctl;  // correction kernel: pass parity of the inner plan
};

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

__device__ __forceinline__ double bits_to_double(unsigned long long b) {
    return __longlong_as_double((long long)b);
}

// Largest change of a sweep as the bit pattern of a non-negative double:
// integer order equals numeric order, so atomicMax is exact and
// order-independent. NaN changes are dropped, as f64::max drops them.
__device__ __forceinline__ unsigned long long change_bits(double m) {
    return (m == m) ? (unsigned long long)__double_as_longlong(m) : 0ull;
}

// ---------------------------------------------------------------------------
// Finalize: run by every thread of the last workgroup of a launch.
// ---------------------------------------------------------------------------
__device__ inline void finalize(const PassArgs& a) {
    for (int p = threadIdx.x; p < a.nprob; p += blockDim.x) {
        const Prob& pr = a.probs[p];
        Ctl& c = a.ctl[p];
        const int want = (a.mode == MODE_REPLAY) ? ST_REPLAY : ST_RUNNING;
        if (c.status != want) continue;
        const int base = a.tiling ? pr.el_base : pr.sw_base;
        const int cnt = a.tiling ? pr.el_count : pr.sw_count;
        double r2 = 0.0, b2 = 0.0;
        if (a.check || a.mode == MODE_RESID) {
            for (int t = 0; t < cnt; t++) {
                r2 += __hip_atomic_load(&a.partials[2 * (base + t)], __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
                b2 += __hip_atomic_load(&a.partials[2 * (base + t) + 1], __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
            }
            if (!(r2 == r2) || r2 > 1.7e308) c.nonfinite = 1;
        }
        if (a.mode == MODE_RESID) {
            c.r2 = r2;
            c.b2 = b2;
            const double lim = a.tol * a.tol * (b2 > 0.0 ? b2 : 1.0);
            if (a.crit == CRIT_RESID && r2 <= lim) c.status = ST_DONE;
            if (c.nonfinite) c.status = ST_MAXED;
            continue;
        }
        if (a.mode == MODE_REPLAY) {
            c.sweeps += 1;
            c.pass += 1;
            c.replay -= 1;
            if (c.replay <= 0) c.status = ST_DONE;
            c.moved[0] = 0ull;
            continue;
        }
        bool advance = true;
        if (a.crit == CRIT_MAXCHANGE) {
            int found = -1;
            for (int j = 0; j < a.s; j++) {
                const unsigned long long mb =
                    __hip_atomic_load(&c.moved[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (bits_to_double(mb) < a.tol) { found = j; break; }
            }
            if (found == a.s - 1) {
                c.sweeps += a.s;
                c.pass += 1;
                c.status = ST_DONE;
                advance = false;
            } else if (found >= 0) {
                // Stopped inside the pass: keep the input (src) and replay
                // found+1 single sweeps from it.
                c.status = ST_REPLAY;
                c.replay = found + 1;
                advance = false;
            }
        } else if (a.crit == CRIT_RESID && a.check) {
            c.r2 = r2;
            const double lim = a.tol * a.tol * (c.b2 > 0.0 ? c.b2 : 1.0);
            if (r2 <= lim) { c.status = ST_DONE; advance = false; }
        } else if (a.crit == CRIT_INNER && a.check) {
            if (c.checks == 0) c.r2_first = r2;
            c.checks += 1;
            c.r2 = r2;
            if (r2 <= a.tol * a.tol * c.r2_first) { c.status = ST_DONE; advance = false; }
        }
        if (advance) {
            c.sweeps += a.s;
            c.pass += 1;
            if (c.status == ST_RUNNING && c.sweeps >= a.max_sweeps) c.status = ST_MAXED;
        }
        if (c.nonfinite && c.status == ST_RUNNING) c.status = ST_MAXED;
        for (int j = 0; j < kMaxS; j++) c.moved[j] = 0ull;
    }
}

// Arrival counter: the last workgroup to arrive runs finalize and re-arms.
__device__ inline void arrive_and_finalize(const PassArgs& a, bool do_finalize) {
    __shared__ int s_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        unsigned prev = atomicAdd(a.counter, 1u);
        s_last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (s_last) {
        __threadfence();
        if (do_finalize) finalize(a);
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            *a.counter = 0u;
        }
    }
}

__device__ __forceinline__ void block_partials(const PassArgs& a, int tile, double r2, double b2) {
    __shared__ double s_r[16], s_b[16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    r2 = wave_sum(r2);
    b2 = wave_sum(b2);
    if (lane == 0) { s_r[wv] = r2; s_b[wv] = b2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tr = 0.0, tb = 0.0;
        for (int w = 0; w < nw; w++) { tr += s_r[w]; tb += s_b[w]; }
        a.partials[2 * tile] = tr;
        a.partials[2 * tile + 1] = tb;
    }
}

__device__ __forceinline__ int real_count(int gx, int gy, int gz, int W, int H, int D) {
    return (gx > 0) + (gx < W - 1) + (gy > 0) + (gy < H - 1) + (gz > 0) + (gz < D - 1);
}

// ---------------------------------------------------------------------------
// 2D LDS tile sweep: s full sweeps (2s half sweeps) per launch.
//
// A workgroup loads a region 128 wide and RY tall (halo H = 2s around a
// (128-2H) x (RY-2H) interior) from src into LDS, split by the parity of the
// local x so every neighbour read is unit-stride across a wave. Wave w owns
// rows w, w+NW, ...; lane l owns the pair of columns (2l, 2l+1) and keeps its
// two coefficients per row in registers. Half sweep h updates the cells of
// colour h&1 at distance > h from the region's edge (the rest is stale halo),
// so after 2s half sweeps the interior is exact and is stored to dst.
// ---------------------------------------------------------------------------
template <typename T> struct Arith;
template <> struct Arith<double> {
    // Reference order, no contraction (file compiled -ffp-contract=off).
    static __device__ __forceinline__ double balanced(double d, double sum, double k, double k2,
                                                       double source, bool general, int real,
                                                       double decay, const float*) {
        double num = d * sum;
        if (general) {
            num = num + source;
            num = num + k;
            num = num - k2;
        } else {
            num = num + k;
        }
        return num / ((double)real * d + decay);
    }
    static __device__ __forceinline__ double relax(double omo, double omega, double before,
                                                    double bal) {
        return omo * before + omega * bal;
    }
};
template <> struct Arith<float> {
    // A smoother only: the fast reciprocal (1 ulp) replaces the division.
    static __device__ __forceinline__ float balanced(float d, float sum, float k, float, float,
                                                     bool, int real, float decay,
                                                     const float*) {
        return __builtin_fmaf(d, sum, k) * __builtin_amdgcn_rcpf(__builtin_fmaf((float)real, d, decay));
    }
    static __device__ __forceinline__ float relax(float omo, float omega, float before,
                                                  float bal) {
        return __builtin_fmaf(omo, before, omega * bal);
    }
};

template <typename T> __device__ __forceinline__ T nan_of();
template <> __device__ __forceinline__ double nan_of<double>() { return __longlong_as_double(0x7ff8000000000000ll); }
template <> __device__ __forceinline__ float nan_of<float>() { return __int_as_float(0x7fc00000); }

template <typename T, int S, int RY, int NW, bool PLAIN>
__global__ __launch_bounds__(NW * 64) void sweep2d(PassArgs a) {
    constexpr int H = 2 * S;
    constexpr int RX = 128;
    constexpr int TX = RX - 2 * H;
    constexpr int TY = RY - 2 * H;
    constexpr int NR = RY / NW;
    static_assert(RY % NW == 0, "rows split evenly over waves");
    static_assert(TY > 0, "interior rows");
    using T2 = typename V2<T>::type;
    __shared__ T cs[RY][2][64];

    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int p = a.tile_prob[t];
    const Prob& pr = a.probs[p];
    const Ctl& ctl = a.ctl[p];
    const int want = (a.mode == MODE_REPLAY) ? ST_REPLAY : ST_RUNNING;
    double r2 = 0.0;
    unsigned long long mv[S];
#pragma unroll
    for (int j = 0; j < S; j++) mv[j] = 0ull;

    if (ctl.status == want) {
        const int lt = t - pr.sw_base;
        const int ti = lt % pr.ntx, tj = lt / pr.ntx;
        const int gxo = pr.xs + ti * TX - H;
        const int gyo = pr.ys + tj * TY - H;
        const int W = pr.W, Hh = pr.H, D = pr.D, P = pr.P;
        const int gz = 0;
        const int base = (gxo + gyo + gz) & 1;
        const int par = ctl.pass & 1;
        const T* src = (const T*)pr.x[par];
        T* dst = (T*)pr.x[par ^ 1];
        const T* kk = (const T*)pr.k;
        const T* kk2 = (const T*)pr.k2;
        const T* lbp = (const T*)pr.lb;
        const bool general = PLAIN && kk2 != nullptr;
        const bool has_lb = !PLAIN && lbp != nullptr;
        const T d = (T)pr.d, decay = (T)pr.decay, omega = (T)pr.omega, omo = (T)pr.omo;
        const T source = (T)pr.source;
        const int gx = gxo + 2 * lane;
        const bool colin = gx >= 0 && gx < W;

        T kr[NR][2], k2r[NR][2], lbr[NR][2];
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const int y = wv + NW * i;
            const int gy = gyo + y;
            T v0 = T(0), v1 = T(0);
            T c0 = nan_of<T>(), c1 = nan_of<T>();
            T u0 = T(0), u1 = T(0), l0 = T(0), l1 = T(0);
            if (colin && gy >= 0 && gy < Hh) {
                const size_t off = (size_t)gy * P + gx;
                const T2 v = *(const T2*)(src + off);
                const T2 c = *(const T2*)(kk + off);
                v0 = v.x; v1 = v.y; c0 = c.x; c1 = c.y;
                if (general) { const T2 u = *(const T2*)(kk2 + off); u0 = u.x; u1 = u.y; }
                if (has_lb) { const T2 l = *(const T2*)(lbp + off); l0 = l.x; l1 = l.y; }
            }
            cs[y][0][lane] = v0;
            cs[y][1][lane] = v1;
            kr[i][0] = c0; kr[i][1] = c1;
            k2r[i][0] = u0; k2r[i][1] = u1;
            lbr[i][0] = l0; lbr[i][1] = l1;
        }
        __syncthreads();

        const float* inv_den = pr.inv_den;
        // Residual of the input state on the interior (projected natural residual).
        if (a.check) {
#pragma unroll
            for (int i = 0; i < NR; i++) {
                const int y = wv + NW * i;
                if (y < H || y >= H + TY) continue;
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    const int x = 2 * lane + q;
                    if (x < H || x >= H + TX) continue;
                    const T kq = kr[i][q];
                    if (kq != kq) continue;
                    const T c = cs[y][q][lane];
                    const T w = q ? cs[y][0][lane] : cs[y][1][lane - 1];
                    const T e = q ? cs[y][0][lane + 1] : cs[y][1][lane];
                    T sum = T(0);
                    sum += w; sum += e; sum += cs[y - 1][q][lane]; sum += cs[y + 1][q][lane];
                    const int real = real_count(gxo + x, gyo + y, gz, W, Hh, D);
                    const double A = (double)real * (double)d + (double)decay;
                    double b;
                    if (general) b = ((double)source + (double)kq) - (double)k2r[i][q];
                    else b = (double)kq;
                    double r = (b + (double)d * (double)sum) - A * (double)c;
                    if (PLAIN) {
                        if (A > 0.0 && !((double)c + r / A > 0.0)) r = -A * (double)c;
                    } else if (has_lb) {
                        const double lo = (double)lbr[i][q];
                        if (A > 0.0 && !((double)c + r / A > lo)) r = A * (lo - (double)c);
                    }
                    r2 += r * r;
                }
            }
        }

#pragma unroll
        for (int h = 0; h < 2 * S; h++) {
            const int colour = h & 1;
            const int lo = h + 1, hiY = RY - 2 - h, hiX = RX - 2 - h;
#pragma unroll
            for (int i = 0; i < NR; i++) {
                const int y = wv + NW * i;
                if (y < lo || y > hiY) continue;
                const int q = (colour + base + y) & 1;
                const int x = 2 * lane + q;
                if (x < lo || x > hiX) continue;
                const T kq = q ? kr[i][1] : kr[i][0];
                if (kq != kq) continue;
                const T before = cs[y][q][lane];
                const T w = q ? cs[y][0][lane] : cs[y][1][lane - 1];
                const T e = q ? cs[y][0][lane + 1] : cs[y][1][lane];
                T sum = T(0);
                sum += w; sum += e; sum += cs[y - 1][q][lane]; sum += cs[y + 1][q][lane];
                const int real = real_count(gxo + x, gyo + y, gz, W, Hh, D);
                const T k2q = q ? k2r[i][1] : k2r[i][0];
                const T bal = Arith<T>::balanced(d, sum, kq, k2q, source, general, real, decay, inv_den);
                T after = Arith<T>::relax(omo, omega, before, bal);
                if (PLAIN) {
                    after = after > T(0) ? after : T(0);
                } else if (has_lb) {
                    const T l = q ? lbr[i][1] : lbr[i][0];
                    after = after > l ? after : l;
                }
                cs[y][q][lane] = after;
                if (PLAIN && y >= H && y < H + TY && x >= H && x < H + TX) {
                    const unsigned long long m = change_bits((double)__builtin_fabs(after - before));
                    mv[h >> 1] = m > mv[h >> 1] ? m : mv[h >> 1];
                }
            }
            __syncthreads();
        }

        // Store the interior (only rows and pairs that reach the swept box).
        if (lane >= H / 2 && lane < (H + TX) / 2 && gx <= pr.x1) {
#pragma unroll
            for (int i = 0; i < NR; i++) {
                const int y = wv + NW * i;
                if (y < H || y >= H + TY) continue;
                const int gy = gyo + y;
                if (gy > pr.y1) continue;
                T2 v;
                v.x = cs[y][0][lane];
                v.y = cs[y][1][lane];
                *(T2*)(dst + (size_t)gy * P + gx) = v;
            }
        }
    }

    if (a.check) block_partials(a, t, r2, 0.0);
    if (PLAIN) {
        __shared__ unsigned long long s_m[NW][S];
#pragma unroll
        for (int j = 0; j < S; j++) {
            const unsigned long long m = wave_umax(mv[j]);
            if (lane == 0) s_m[wv][j] = m;
        }
        __syncthreads();
        if (threadIdx.x < S) {
            unsigned long long m = 0ull;
            for (int w = 0; w < NW; w++) m = s_m[w][threadIdx.x] > m ? s_m[w][threadIdx.x] : m;
            if (m) atomicMax(&a.ctl[p].moved[threadIdx.x], m);
        }
    }
    arrive_and_finalize(a, true);
}

// ---------------------------------------------------------------------------
// Elementwise tiling: 128 cells in x (64 pairs) by NW rows of one plane.
// ---------------------------------------------------------------------------
struct ElTile {
    int gx, gy, gz;   // pair start column, row, plane
    bool ok;
};

__device__ __forceinline__ ElTile el_tile(const Prob& pr, int t) {
    const int lt = t - pr.el_base;
    const int nw = blockDim.x >> 6;
    const int tx = lt % pr.elx;
    const int rest = lt / pr.elx;
    const int ty = rest % pr.ely;
    const int tz = rest / pr.ely;
    ElTile e;
    e.gx = pr.xs + tx * 128 + 2 * (threadIdx.x & 63);
    e.gy = pr.ys + ty * nw + (threadIdx.x >> 6);
    e.gz = pr.zs + tz;
    e.ok = e.gx <= pr.x1 && e.gy <= pr.y1 && e.gz <= pr.z1;
    return e;
}

// 3D red or black half sweep in place (x[0] == x[1]): cells of one colour read
// only the other colour, which this launch never writes, so the result equals
// the reference's read-old/write-new half sweep (pde.rs:322-328).
template <typename T, bool PLAIN>
__global__ __launch_bounds__(256) void half3d(PassArgs a, int colour) {
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int p = a.tile_prob[t];
    const Prob& pr = a.probs[p];
    const Ctl& ctl = a.ctl[p];
    const int want = (a.mode == MODE_REPLAY) ? ST_REPLAY : ST_RUNNING;
    unsigned long long mv = 0ull;
    if (ctl.status == want) {
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
                    const T k2q = general ? kk2[i] : T(0);
                    const T bal = Arith<T>::balanced(d, sum, kq, k2q, (T)pr.source, general, real,
                                                     decay, pr.inv_den);
                    T after = Arith<T>::relax(omo, omega, before, bal);
                    if (PLAIN) after = after > T(0) ? after : T(0);
                    else if (has_lb) { const T l = lbp[i]; after = after > l ? after : l; }
                    xs[i] = after;
                    if (PLAIN) mv = change_bits((double)__builtin_fabs(after - before));
                }
            }
        }
    }
    if (PLAIN) {
        __shared__ unsigned long long s_m[4];
        const unsigned long long m = wave_umax(mv);
        if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long mm = 0ull;
            for (int w = 0; w < (int)(blockDim.x >> 6); w++) mm = s_m[w] > mm ? s_m[w] : mm;
            if (mm) atomicMax(&a.ctl[p].moved[0], mm);
        }
    }
    arrive_and_finalize(a, a.mode == MODE_HALF_BLACK || a.mode == MODE_REPLAY);
}

// FP64 residual of the current state (2D and 3D), elementwise tiling.
// Per free cell: b = (source + production) - uptake (or the folded k),
// r = (b + d*sum) - A*c, projected to the natural residual; b_eff adds the
// held neighbours. With write32 it also writes the FP32 defect right-hand
// side r, the correction's lower bound -c (rounded toward zero) and zeroes
// e[0] for the FP32 inner solve.
__global__ __launch_bounds__(256) void resid64(PassArgs a, int write32) {
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int p = a.tile_prob[t];
    const Prob& pr = a.probs[p];
    const Ctl& ctl = a.ctl[p];
    double r2 = 0.0, b2 = 0.0;
    if (ctl.status == ST_RUNNING) {
        const ElTile et = el_tile(pr, t);
        if (et.ok) {
            const int W = pr.W, Hh = pr.H, D = pr.D, P = pr.P;
            const size_t plane = (size_t)P * Hh;
            const double* c = pr.c64;
            const double* kk = (const double*)pr.k;
            const double* kk2 = (const double*)pr.k2;
            const double d = pr.d, decay = pr.decay;
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int x = et.gx + q;
                if (x >= W) continue;
                const size_t i = (size_t)et.gz * plane + (size_t)et.gy * P + x;
                const double kq = kk[i];
                if (kq != kq) {
                    if (write32) { pr.r32[i] = __int_as_float(0x7fc00000); pr.e0[i] = 0.0f; }
                    continue;
                }
                const double ci = c[i];
                double sum = 0.0, hsum = 0.0;
                int real = 0;
#define CW_NB(cond, j)                                   \
    if (cond) {                                          \
        const double v = c[j];                           \
        sum += v;                                        \
        real += 1;                                       \
        if (kk[j] != kk[j]) hsum += v;                   \
    }
                CW_NB(x > 0, i - 1)
                CW_NB(x + 1 < W, i + 1)
                CW_NB(et.gy > 0, i - P)
                CW_NB(et.gy + 1 < Hh, i + P)
                CW_NB(et.gz > 0, i - plane)
                CW_NB(et.gz + 1 < D, i + plane)
#undef CW_NB
                const double A = (double)real * d + decay;
                const double b = kk2 ? ((pr.source + kq) - kk2[i]) : kq;
                const double r = (b + d * sum) - A * ci;
                double rt = r;
                if (A > 0.0 && !(ci + r / A > 0.0)) rt = -A * ci;
                const double be = b + d * hsum;
                r2 += rt * rt;
                b2 += be * be;
                if (write32) {
                    pr.r32[i] = (float)r;
                    pr.e0[i] = 0.0f;
                    if (pr.lb32) pr.lb32[i] = __double2float_rz(-ci);
                }
            }
        }
    }
    block_partials(a, t, r2, b2);
    arrive_and_finalize(a, true);
}

// c <- max(c + e, 0) on the free cells, e from the buffer the inner plan ended in.
__global__ __launch_bounds__(256) void correct64(PassArgs a) {
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int p = a.tile_prob[t];
    const Prob& pr = a.probs[p];
    const Ctl& ctl = a.ctl[p];
    if (ctl.status == ST_RUNNING) {
        const ElTile et = el_tile(pr, t);
        if (et.ok) {
            const int W = pr.W, P = pr.P;
            const size_t plane = (size_t)P * pr.H;
            const Ctl& ic = a.inner_ctl[pr.inner_idx];
            const float* e = (ic.pass & 1) ? pr.e1 : pr.e0;
            const double* kk = (const double*)pr.k;
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int x = et.gx + q;
                if (x >= W) continue;
                const size_t i = (size_t)et.gz * plane + (size_t)et.gy * P + x;
                if (kk[i] != kk[i]) continue;
                const double v = pr.c64[i] + (double)e[i];
                pr.c64[i] = v > 0.0 ? v : 0.0;
            }
        }
    }
    arrive_and_finalize(a, false);
}

}  // namespace cw
