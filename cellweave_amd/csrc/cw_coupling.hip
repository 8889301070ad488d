// cwfield, the field's state on the device: the coefficient words the solve
// reads, the gathers from the Cellular Potts cell-id lattice
// (crates/engine/src/coupling.rs), the holds and the bath flood
// (pde.rs:465-631), the inflow and free count (pde.rs:415-461), the explicit
// step (pde.rs:651-728) and the per-cell means (coupling.rs:164-182).
//
// Every per-voxel array is pitched like the field (row pitch P, plane stride
// P*H); the lattice's cell ids use the same layout, so each gather is one
// coalesced pass that reads 4 B of id per voxel plus an L2-resident per-cell
// table. The operations are synchronous, as the reference's are.
#include "cw_internal.hpp"

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

using namespace cwi;

// Device copy of the lattice's cell ids (CpmLattice::grid, cpm.rs:198), plus
// the staging of per-cell tables and the cell -> voxel lists that the
// per-cell means use (rebuilt after each upload).
struct cw_lattice {
    uint32_t W = 0, H = 0, D = 0;
    int P = 0;
    size_t n = 0, npad = 0;
    hipStream_t stream = nullptr;
    uint32_t* ids = nullptr;        // pitched, padding 0
    uint8_t* flags = nullptr;       // staged per-cell u8 table
    double* table = nullptr;        // staged per-cell f64 table
    size_t table_cap = 0;
    // cell -> voxel lists (voxels in increasing flat index), valid for csr_ncells
    uint32_t csr_ncells = 0;
    bool csr_valid = false;
    uint32_t *keys[2] = {nullptr, nullptr}, *vals[2] = {nullptr, nullptr};
    uint32_t *first = nullptr, *last = nullptr;
    size_t range_cap = 0;
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    void* stage = nullptr;
    size_t stage_bytes = 0;
};

namespace {

constexpr int kBlock = 256;

struct Geo {
    int P, W, H, D;
    size_t npad;
};

Geo geo_of(const cw_field* f) { return Geo{f->P, (int)f->W, (int)f->H, (int)f->D, f->npad}; }

unsigned grid_of(size_t npad) { return (unsigned)((npad + kBlock - 1) / kBlock); }

// (x, y, z) of pitched element i; x >= W marks padding.
__device__ inline void coords(const Geo& g, size_t i, int& x, int& y, int& z) {
    const size_t row = i / (size_t)g.P;
    x = (int)(i - row * (size_t)g.P);
    y = (int)(row % (size_t)g.H);
    z = (int)(row / (size_t)g.H);
}

__device__ inline bool valid_id(uint32_t id, uint32_t ncells) {
    return id != 0 && id != kObstacle && id < ncells;
}

__device__ inline double clamp0(double v) { return v > 0.0 ? v : 0.0; }

__device__ inline int wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline int wave_or(int v) {
    for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
    return v;
}

// ---- coefficient scan and build -------------------------------------------
// scan[10] flag bits (1 general, 2 folded sink, 4 general sink, 8 any free),
// scan[4..9] the free-span box (min x, max x, min y, max y, min z, max z).
__global__ void k_scan_init(int* scan) {
    if (threadIdx.x < 16) {
        const int t = threadIdx.x;
        scan[t] = (t >= 4 && t < 10 && (t & 1) == 0) ? INT_MAX : (t >= 4 && t < 10 ? -1 : 0);
    }
}

// Reductions walk whole rows: block b takes rows b, b + grid, ... and writes
// its partial to slot b of the field's reduction slots; one block folds them.
// (Thousands of blocks updating one cache line with atomics serialise at the
// L2: 2048 blocks took ~140 us that way.)
unsigned red_grid(const cw_field* f) {
    return (unsigned)std::min<size_t>(kRedBlocks, (size_t)f->H * f->D);
}

int* red_slots(const cw_field* f) { return f->scan + kScanInts; }

// Block-wide combine of (flags, box); thread 0 gets the block's result.
__device__ inline void box_reduce(int& flags, int& mnx, int& mxx, int& mny, int& mxy, int& mnz,
                                  int& mxz) {
    __shared__ int sh[kBlock / 64][7];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    flags = wave_or(flags);
    mnx = wave_min(mnx); mxx = wave_max(mxx);
    mny = wave_min(mny); mxy = wave_max(mxy);
    mnz = wave_min(mnz); mxz = wave_max(mxz);
    if (lane == 0) {
        sh[wv][0] = flags; sh[wv][1] = mnx; sh[wv][2] = mxx; sh[wv][3] = mny;
        sh[wv][4] = mxy; sh[wv][5] = mnz; sh[wv][6] = mxz;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; w++) {
            flags |= sh[w][0];
            mnx = min(mnx, sh[w][1]); mxx = max(mxx, sh[w][2]);
            mny = min(mny, sh[w][3]); mxy = max(mxy, sh[w][4]);
            mnz = min(mnz, sh[w][5]); mxz = max(mxz, sh[w][6]);
        }
    }
}

__device__ inline void box_partial(int flags, int mnx, int mxx, int mny, int mxy, int mnz, int mxz,
                                   int* slots) {
    box_reduce(flags, mnx, mxx, mny, mxy, mnz, mxz);
    if (threadIdx.x == 0) {
        int* o = slots + (size_t)blockIdx.x * kRedSlot;
        o[0] = flags; o[1] = mnx; o[2] = mxx; o[3] = mny; o[4] = mxy; o[5] = mnz; o[6] = mxz;
    }
}

// One block: folds `nblocks` (flags, box) slots into out_flags and out_box,
// which hold the identity (0, INT_MAX / -1 pairs) or earlier results.
__global__ __launch_bounds__(kBlock) void k_box_fold(const int* slots, int nblocks,
                                                     int* out_flags, int* out_box) {
    int flags = 0;
    int mnx = INT_MAX, mxx = -1, mny = INT_MAX, mxy = -1, mnz = INT_MAX, mxz = -1;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) {
        const int* o = slots + (size_t)b * kRedSlot;
        flags |= o[0];
        mnx = min(mnx, o[1]); mxx = max(mxx, o[2]);
        mny = min(mny, o[3]); mxy = max(mxy, o[4]);
        mnz = min(mnz, o[5]); mxz = max(mxz, o[6]);
    }
    box_reduce(flags, mnx, mxx, mny, mxy, mnz, mxz);
    if (threadIdx.x == 0 && flags) {
        *out_flags |= flags;
        if (mxx >= 0) {
            out_box[0] = min(out_box[0], mnx); out_box[1] = max(out_box[1], mxx);
            out_box[2] = min(out_box[2], mny); out_box[3] = max(out_box[3], mxy);
            out_box[4] = min(out_box[4], mnz); out_box[5] = max(out_box[5], mxz);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_scan(Geo g, const uint8_t* held, const double* up,
                                                 const double* pr, double source, int* slots) {
    int flags = 0;
    int mnx = INT_MAX, mxx = -1, mny = INT_MAX, mxy = -1, mnz = INT_MAX, mxz = -1;
    const size_t rows = (size_t)g.H * g.D;
    for (size_t row = blockIdx.x; row < rows; row += gridDim.x)
    for (int x = threadIdx.x; x < g.W; x += kBlock) {
        const size_t i = row * (size_t)g.P + x;
        if (held[i]) continue;
        const int y = (int)(row % (size_t)g.H), z = (int)(row / (size_t)g.H);
        const double p = pr[i], u = up[i];
        flags |= (p != 0.0 && u != 0.0) | (((p != 0.0 ? p : -u) < 0.0) << 1) |
                 ((((source + p) - u) < 0.0) << 2) | 8;
        mnx = min(mnx, x); mxx = max(mxx, x);
        mny = min(mny, y); mxy = max(mxy, y);
        mnz = min(mnz, z); mxz = max(mxz, z);
    }
    box_partial(flags, mnx, mxx, mny, mxy, mnz, mxz, slots);
}

// k = p or -u (folded, exact when no voxel has both and source == 0) or
// k = p, k2 = u (general); held voxels and padding carry NaN.
__global__ __launch_bounds__(kBlock) void k_build(Geo g, const uint8_t* held, const double* up,
                                                  const double* pr, int general, double* k,
                                                  double* k2) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.npad) return;
    int x, y, z;
    coords(g, i, x, y, z);
    const double qnan = __longlong_as_double(0x7ff8000000000000ll);
    if (x >= g.W || held[i]) {
        k[i] = qnan;
        if (general) k2[i] = 0.0;
        return;
    }
    const double p = pr[i], u = up[i];
    if (general) {
        k[i] = p;
        k2[i] = u;
    } else {
        k[i] = p != 0.0 ? p : -u;
    }
}

// ---- gathers from the lattice (coupling.rs:26-73) -------------------------
// per_cell == nullptr: every consuming cell takes up `rate` (apply_uptake);
// otherwise its own per_cell[id] where that is positive.
__global__ __launch_bounds__(kBlock) void k_uptake(Geo g, const uint32_t* ids,
                                                   const uint8_t* consumes, const double* per_cell,
                                                   uint32_t ncells, double rate, double* up) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.npad) return;
    int x, y, z;
    coords(g, i, x, y, z);
    if (x >= g.W) return;
    const uint32_t id = ids[i];
    double v = 0.0;
    if (valid_id(id, ncells) && (!consumes || consumes[id])) {
        if (per_cell) {
            const double r = per_cell[id];
            v = r > 0.0 ? r : 0.0;
        } else {
            v = clamp0(rate);
        }
    }
    up[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_background(Geo g, const uint32_t* ids, double rate,
                                                       double* up) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.npad) return;
    int x, y, z;
    coords(g, i, x, y, z);
    if (x < g.W && ids[i] == 0) up[i] = clamp0(rate);
}

__global__ __launch_bounds__(kBlock) void k_production(Geo g, const uint32_t* ids,
                                                       const double* per_cell, uint32_t ncells,
                                                       double* pr) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.npad) return;
    int x, y, z;
    coords(g, i, x, y, z);
    if (x >= g.W) return;
    const uint32_t id = ids[i];
    double v = 0.0;
    if (valid_id(id, ncells) && per_cell[id] > 0.0) v = per_cell[id];
    pr[i] = v;
}

// ---- holds (pde.rs:468-498, coupling.rs:80-82) ----------------------------
// mode 0: hold where mask != 0; 1: where the lattice has an obstacle;
// 2: everywhere; 3: release every hold.
__global__ __launch_bounds__(kBlock) void k_hold(Geo g, int mode, const uint8_t* mask,
                                                 const uint32_t* ids, double value,
                                                 uint8_t* held, double* c) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.npad) return;
    int x, y, z;
    coords(g, i, x, y, z);
    if (x >= g.W) return;
    if (mode == 3) {
        held[i] = 0;
        return;
    }
    const bool m = mode == 2 || (mode == 0 ? mask[i] != 0 : ids[i] == kObstacle);
    if (m) {
        held[i] = 1;
        c[i] = value;
    }
}

// ---- bath flood (pde.rs:550-631) --------------------------------------------
struct Box {
    int x0, x1, y0, y1, z0, z1;
};

__device__ inline bool in_box(const Box& b, int x, int y, int z) {
    return x >= b.x0 && x <= b.x1 && y >= b.y0 && y <= b.y1 && z >= b.z0 && z <= b.z1;
}

__global__ __launch_bounds__(kBlock) void k_open_from_ids(Geo g, const uint32_t* ids,
                                                          uint8_t* open) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.npad) return;
    int x, y, z;
    coords(g, i, x, y, z);
    open[i] = (x < g.W && ids[i] == 0) ? 1 : 0;
}

// Box around the closed voxels: scan[0] any, [1..6] min/max x, y, z.
__global__ void k_box_init(int* scan) {
    if (threadIdx.x < 8) {
        const int t = threadIdx.x;
        scan[t] = (t == 0 || t == 7) ? 0 : ((t & 1) ? INT_MAX : -1);
    }
}

__global__ __launch_bounds__(kBlock) void k_closed_box(Geo g, const uint8_t* open, int* slots) {
    int flags = 0;
    int mnx = INT_MAX, mxx = -1, mny = INT_MAX, mxy = -1, mnz = INT_MAX, mxz = -1;
    const size_t rows = (size_t)g.H * g.D;
    for (size_t row = blockIdx.x; row < rows; row += gridDim.x)
    for (int x = threadIdx.x; x < g.W; x += kBlock) {
        const size_t i = row * (size_t)g.P + x;
        if (open[i]) continue;
        const int y = (int)(row % (size_t)g.H), z = (int)(row / (size_t)g.H);
        flags = 1;
        mnx = min(mnx, x); mxx = max(mxx, x);
        mny = min(mny, y); mxy = max(mxy, y);
        mnz = min(mnz, z); mxz = max(mxz, z);
    }
    box_partial(flags, mnx, mxx, mny, mxy, mnz, mxz, slots);
}

// Seeds: open, not yet held voxels on the outer layer of the box (its first
// and last slice count only in a volume).
__global__ __launch_bounds__(kBlock) void k_flood_seed(Geo g, Box b, const uint8_t* open,
                                                       const uint8_t* held, uint8_t* reach) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.npad) return;
    int x, y, z;
    coords(g, i, x, y, z);
    uint8_t r = 0;
    if (x < g.W && in_box(b, x, y, z) && open[i] && !held[i]) {
        const bool outer = x == b.x0 || x == b.x1 || y == b.y0 || y == b.y1 ||
                           (g.D > 1 && (z == b.z0 || z == b.z1));
        r = outer ? 1 : 0;
    }
    reach[i] = r;
}

// One pass of face-connected propagation over tiles of the box: each tile
// iterates in LDS to its own fixed point, with its one-voxel halo as the last
// launch (or a neighbour tile of this launch) left it. Reach only grows, so a
// stale halo read delays propagation to the next launch and never loses it;
// *changed says whether another launch is needed.
template <int TX, int TY, int TZ>
__global__ __launch_bounds__(kBlock) void k_flood(Geo g, Box b, const uint8_t* open,
                                                  const uint8_t* held, uint8_t* reach,
                                                  int* changed) {
    __shared__ uint8_t t[TZ + 2][TY + 2][TX + 2];
    __shared__ uint8_t el[TZ][TY][TX];
    const int bw = b.x1 - b.x0 + 1, bh = b.y1 - b.y0 + 1;
    const int ntx = (bw + TX - 1) / TX, nty = (bh + TY - 1) / TY;
    const int tile = blockIdx.x;
    const int ox = b.x0 + (tile % ntx) * TX;
    const int oy = b.y0 + ((tile / ntx) % nty) * TY;
    const int oz = b.z0 + (tile / (ntx * nty)) * TZ;
    const size_t plane = (size_t)g.P * g.H;
    constexpr int NH = (TZ + 2) * (TY + 2) * (TX + 2);
    for (int s = threadIdx.x; s < NH; s += kBlock) {
        const int lx = s % (TX + 2), ly = (s / (TX + 2)) % (TY + 2), lz = s / ((TX + 2) * (TY + 2));
        const int x = ox + lx - 1, y = oy + ly - 1, z = oz + lz - 1;
        uint8_t v = 0;
        if (in_box(b, x, y, z)) v = reach[(size_t)z * plane + (size_t)y * g.P + x];
        t[lz][ly][lx] = v;
    }
    constexpr int NO = TZ * TY * TX;
    for (int s = threadIdx.x; s < NO; s += kBlock) {
        const int lx = s % TX, ly = (s / TX) % TY, lz = s / (TX * TY);
        const int x = ox + lx, y = oy + ly, z = oz + lz;
        uint8_t e = 0;
        if (in_box(b, x, y, z)) {
            const size_t i = (size_t)z * plane + (size_t)y * g.P + x;
            e = open[i] && !held[i];
        }
        el[lz][ly][lx] = e;
    }
    __syncthreads();
    int grew = 0;
    for (int it = 0; it < 4 * (TX + TY + TZ); it++) {
        int ch = 0;
        for (int s = threadIdx.x; s < NO; s += kBlock) {
            const int lx = s % TX, ly = (s / TX) % TY, lz = s / (TX * TY);
            if (!el[lz][ly][lx] || t[lz + 1][ly + 1][lx + 1]) continue;
            const bool nb = t[lz + 1][ly + 1][lx] | t[lz + 1][ly + 1][lx + 2] |
                            t[lz + 1][ly][lx + 1] | t[lz + 1][ly + 2][lx + 1] |
                            t[lz][ly + 1][lx + 1] | t[lz + 2][ly + 1][lx + 1];
            if (nb) {
                t[lz + 1][ly + 1][lx + 1] = 1;
                ch = 1;
            }
        }
        ch = __syncthreads_or(ch);
        if (!ch) break;
        grew = 1;
    }
    if (grew) {
        for (int s = threadIdx.x; s < NO; s += kBlock) {
            const int lx = s % TX, ly = (s / TX) % TY, lz = s / (TX * TY);
            const int x = ox + lx, y = oy + ly, z = oz + lz;
            if (in_box(b, x, y, z) && t[lz + 1][ly + 1][lx + 1])
                reach[(size_t)z * plane + (size_t)y * g.P + x] = 1;
        }
        if (threadIdx.x == 0) atomicOr(changed, 1);
    }
}

// Everything outside the box is held without being visited; everything the
// flood reached is held; then every held voxel takes the bath's value.
__global__ __launch_bounds__(kBlock) void k_flood_finish(Geo g, Box b, int has_box,
                                                         const uint8_t* reach, double value,
                                                         uint8_t* held, double* c) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.npad) return;
    int x, y, z;
    coords(g, i, x, y, z);
    if (x >= g.W) return;
    bool h = held[i] != 0;
    if (!has_box || !in_box(b, x, y, z) || reach[i]) h = true;
    if (h) {
        held[i] = 1;
        c[i] = value;
    }
}

// ---- reductions: free count and inflow (pde.rs:415-461) ----------------------
__global__ __launch_bounds__(kBlock) void k_count_free(Geo g, const uint8_t* held,
                                                       unsigned long long* slots) {
    __shared__ unsigned long long sh[kBlock / 64];
    unsigned long long v = 0;
    const size_t rows = (size_t)g.H * g.D;
    for (size_t row = blockIdx.x; row < rows; row += gridDim.x)
        for (int x = threadIdx.x; x < g.W; x += kBlock) v += held[row * (size_t)g.P + x] ? 0 : 1;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; w++) v += sh[w];
        slots[blockIdx.x] = v;
    }
}

__global__ __launch_bounds__(kBlock) void k_sum_u64(const unsigned long long* slots, int nblocks,
                                                    unsigned long long* out) {
    __shared__ unsigned long long sh[kBlock / 64];
    unsigned long long v = 0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) v += slots[b];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; w++) v += sh[w];
        *out = v;
    }
}

// Changed voxels of the lattice: entry j sets voxel v[j] (flat index,
// x fastest) to id[j]. The host has already kept only the last entry for
// each voxel, so no two entries write the same word.
__global__ __launch_bounds__(kBlock) void k_lattice_set(const uint64_t* v, const uint32_t* id,
                                                        size_t n, int W, int P, uint32_t* ids) {
    const size_t j = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const uint64_t row = v[j] / (uint64_t)W;
    ids[row * (uint64_t)P + (v[j] - row * (uint64_t)W)] = id[j];
}

// ---- host <-> device copies through a flat staging buffer ---------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_unflatten(const T* flat, T* pitched, int W,
                                                      size_t rows, int P) {
    const size_t j = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= (size_t)W * rows) return;
    const size_t row = j / (size_t)W, x = j - row * (size_t)W;
    pitched[row * (size_t)P + x] = flat[j];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_flatten(const T* pitched, T* flat, int W, size_t rows,
                                                    int P) {
    const size_t j = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= (size_t)W * rows) return;
    const size_t row = j / (size_t)W, x = j - row * (size_t)W;
    flat[j] = pitched[row * (size_t)P + x];
}

cw_status stage_ensure(void** stage, size_t* cap, size_t bytes) {
    if (*cap >= bytes) return CW_OK;
    (void)hipFree(*stage);
    *stage = nullptr;
    *cap = 0;
    CW_HIP(hipMalloc(stage, bytes));
    *cap = bytes;
    return CW_OK;
}

cw_status copy_in(hipStream_t s, int W, size_t rows, int P, void* dst, const void* src, int elem,
                  void** stage, size_t* cap) {
    const size_t n = (size_t)W * rows;
    cw_status st = stage_ensure(stage, cap, n * elem);
    if (st) return st;
    CW_HIP(hipMemcpyAsync(*stage, src, n * elem, hipMemcpyHostToDevice, s));
    const unsigned gb = (unsigned)((n + kBlock - 1) / kBlock);
    if (elem == 1)
        hipLaunchKernelGGL(k_unflatten<uint8_t>, dim3(gb), dim3(kBlock), 0, s,
                           (const uint8_t*)*stage, (uint8_t*)dst, W, rows, P);
    else if (elem == 4)
        hipLaunchKernelGGL(k_unflatten<uint32_t>, dim3(gb), dim3(kBlock), 0, s,
                           (const uint32_t*)*stage, (uint32_t*)dst, W, rows, P);
    else
        hipLaunchKernelGGL(k_unflatten<uint64_t>, dim3(gb), dim3(kBlock), 0, s,
                           (const uint64_t*)*stage, (uint64_t*)dst, W, rows, P);
    CW_HIP(hipGetLastError());
    CW_HIP(hipStreamSynchronize(s));
    return CW_OK;
}

cw_status copy_out(hipStream_t s, int W, size_t rows, int P, const void* src, void* dst, int elem,
                   void** stage, size_t* cap) {
    const size_t n = (size_t)W * rows;
    cw_status st = stage_ensure(stage, cap, n * elem);
    if (st) return st;
    const unsigned gb = (unsigned)((n + kBlock - 1) / kBlock);
    if (elem == 1)
        hipLaunchKernelGGL(k_flatten<uint8_t>, dim3(gb), dim3(kBlock), 0, s, (const uint8_t*)src,
                           (uint8_t*)*stage, W, rows, P);
    else if (elem == 4)
        hipLaunchKernelGGL(k_flatten<uint32_t>, dim3(gb), dim3(kBlock), 0, s,
                           (const uint32_t*)src, (uint32_t*)*stage, W, rows, P);
    else
        hipLaunchKernelGGL(k_flatten<uint64_t>, dim3(gb), dim3(kBlock), 0, s,
                           (const uint64_t*)src, (uint64_t*)*stage, W, rows, P);
    CW_HIP(hipGetLastError());
    CW_HIP(hipMemcpyAsync(dst, *stage, n * elem, hipMemcpyDeviceToHost, s));
    CW_HIP(hipStreamSynchronize(s));
    return CW_OK;
}

constexpr int kInflowPer = 4;  // pitched elements per thread

// The flux d * (c_j - c_i) across every face between a free voxel i and a
// held neighbour j, in the neighbour order x-1, x+1, y-1, y+1, z-1, z+1;
// summed per thread over consecutive elements, then over the block in a
// fixed tree. The sum does not depend on scheduling.
__global__ __launch_bounds__(kBlock) void k_inflow(Geo g, const double* c, const uint8_t* held,
                                                   double d, double* partials) {
    __shared__ double s[kBlock];
    const size_t base = ((size_t)blockIdx.x * kBlock + threadIdx.x) * kInflowPer;
    const size_t plane = (size_t)g.P * g.H;
    double acc = 0.0;
    for (int q = 0; q < kInflowPer; q++) {
        const size_t i = base + q;
        if (i >= g.npad) break;
        int x, y, z;
        coords(g, i, x, y, z);
        if (x >= g.W || held[i]) continue;
        const double ci = c[i];
        if (x > 0 && held[i - 1]) acc += d * (c[i - 1] - ci);
        if (x + 1 < g.W && held[i + 1]) acc += d * (c[i + 1] - ci);
        if (y > 0 && held[i - g.P]) acc += d * (c[i - g.P] - ci);
        if (y + 1 < g.H && held[i + g.P]) acc += d * (c[i + g.P] - ci);
        if (z > 0 && held[i - plane]) acc += d * (c[i - plane] - ci);
        if (z + 1 < g.D && held[i + plane]) acc += d * (c[i + plane] - ci);
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(kBlock) void k_sum_partials(const double* partials, int n,
                                                         double* out) {
    __shared__ double s[kBlock];
    double acc = 0.0;
    for (int j = threadIdx.x; j < n; j += kBlock) acc += partials[j];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// ---- explicit step (pde.rs:651-667, 705-728) --------------------------------
// The reference's expression order: lap = (((W + E) + N) + S) - 4c, plus
// ((B + T) - 2c) in a volume, with a missing neighbour replaced by c; change
// = (((d*lap - decay*c) + source) + p) - u; next = max(c + dt*change, 0).
__global__ __launch_bounds__(kBlock) void k_step(Geo g, const double* c, double* next,
                                                 const uint8_t* held, const double* up,
                                                 const double* pr, double d, double decay,
                                                 double source, double dt) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.npad) return;
    int x, y, z;
    coords(g, i, x, y, z);
    if (x >= g.W) return;
    const double ci = c[i];
    if (held[i]) {
        next[i] = ci;
        return;
    }
    const size_t plane = (size_t)g.P * g.H;
    const double xm = x > 0 ? c[i - 1] : ci;
    const double xp = x + 1 < g.W ? c[i + 1] : ci;
    const double ym = y > 0 ? c[i - g.P] : ci;
    const double yp = y + 1 < g.H ? c[i + g.P] : ci;
    double lap = xm + xp + ym + yp - 4.0 * ci;
    if (g.D > 1) {
        const double zm = z > 0 ? c[i - plane] : ci;
        const double zp = z + 1 < g.D ? c[i + plane] : ci;
        lap += zm + zp - 2.0 * ci;
    }
    const double change = d * lap - decay * ci + source + pr[i] - up[i];
    const double v = ci + dt * change;
    next[i] = v > 0.0 ? v : 0.0;
}

// ---- per-cell means (coupling.rs:164-182) -------------------------------------
// Keys in flat (x, y, z) order: the cell id, or ncells for voxels that count
// for no cell; values: the pitched index. A stable sort keeps each cell's
// voxels in increasing flat index, the order the reference adds them in.
__global__ __launch_bounds__(kBlock) void k_cell_keys(Geo g, const uint32_t* ids,
                                                      uint32_t ncells, uint32_t* keys,
                                                      uint32_t* vals) {
    const size_t j = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t n = (size_t)g.W * g.H * g.D;
    if (j >= n) return;
    const size_t x = j % (size_t)g.W, row = j / (size_t)g.W;
    const size_t i = row * (size_t)g.P + x;
    const uint32_t id = ids[i];
    keys[j] = valid_id(id, ncells) ? id : ncells;
    vals[j] = (uint32_t)i;
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* a, size_t n, uint32_t v) {
    const size_t j = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (j < n) a[j] = v;
}

__global__ __launch_bounds__(kBlock) void k_cell_ranges(const uint32_t* keys, size_t n,
                                                        uint32_t ncells, uint32_t* first,
                                                        uint32_t* last) {
    const size_t j = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const uint32_t k = keys[j];
    if (k >= ncells) return;
    if (j == 0 || keys[j - 1] != k) first[k] = (uint32_t)j;
    if (j + 1 == n || keys[j + 1] != k) last[k] = (uint32_t)(j + 1);
}

// One wave per cell. The lanes gather 64 of the cell's values at a time and
// the sum takes them in lane order, which is ascending voxel index (the CSR
// lists come from a stable sort): the reference's sequential order, so the
// same bits, without one lane walking a cell's voxels load after load.
__global__ __launch_bounds__(kBlock) void k_cell_means(const double* c, const uint32_t* vals,
                                                       const uint32_t* first, const uint32_t* last,
                                                       uint32_t ncells, double* out) {
    const uint32_t k = (uint32_t)(((size_t)blockIdx.x * kBlock + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= ncells) return;  // whole waves
    const uint32_t a = first[k], b = last[k];
    double s = 0.0;
    for (uint32_t j0 = a; j0 < b; j0 += 64) {
        const uint32_t j = j0 + (uint32_t)lane;
        const double v = j < b ? c[vals[j]] : 0.0;
        const int m = (int)min(64u, b - j0);
        for (int l = 0; l < m; l++) s += __shfl(v, l, 64);
    }
    if (lane == 0) out[k] = b > a ? s / (double)(b - a) : 0.0;
}

// ---- helpers ------------------------------------------------------------------
cw_status check_pair(const cw_field* f, const cw_lattice* l) {
    if (!f || !l) return fail(CW_ERR_ARG, "null field or lattice");
    if (f->W != l->W || f->H != l->H || f->D != l->D)
        return fail(CW_ERR_ARG, "lattice %ux%ux%u does not match the field %ux%ux%u", l->W, l->H,
                    l->D, f->W, f->H, f->D);
    return CW_OK;
}

cw_status stage_tables(cw_lattice* l, hipStream_t s, const uint8_t* flags, const double* table,
                       uint32_t ncells) {
    if (ncells > l->table_cap) {
        (void)hipFree(l->flags);
        (void)hipFree(l->table);
        l->flags = nullptr;
        l->table = nullptr;
        l->table_cap = 0;
        CW_HIP(hipMalloc((void**)&l->flags, std::max<size_t>(ncells, 1)));
        CW_HIP(hipMalloc((void**)&l->table, std::max<size_t>(ncells, 1) * sizeof(double)));
        l->table_cap = ncells;
    }
    if (flags && ncells) CW_HIP(hipMemcpyAsync(l->flags, flags, ncells, hipMemcpyHostToDevice, s));
    if (table && ncells)
        CW_HIP(hipMemcpyAsync(l->table, table, ncells * sizeof(double), hipMemcpyHostToDevice, s));
    return CW_OK;
}

// The field's two u8 scratch arrays of its pitched size (open mask, reach),
// allocated on first use and kept: a bath hold every lattice step then
// allocates and frees nothing (hipFree waits for the device).
struct Scratch8 {
    uint8_t* a = nullptr;
    uint8_t* b = nullptr;
};

cw_status scratch8(cw_field* f, Scratch8* s) {
    for (int i = 0; i < 2; i++)
        if (!f->mask8[i]) CW_HIP(hipMalloc((void**)&f->mask8[i], f->npad));
    s->a = f->mask8[0];
    s->b = f->mask8[1];
    return CW_OK;
}

__global__ void k_zero_ints(int* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

cw_status finish(cw_field* f) {
    CW_HIP(hipGetLastError());
    CW_HIP(hipStreamSynchronize(f->stream));
    return CW_OK;
}

cw_status upload_mask(const cw_field* f, const uint8_t* host, uint8_t* dev) {
    std::vector<uint8_t> m(f->n);
    for (size_t i = 0; i < f->n; i++) m[i] = host[i] ? 1 : 0;
    CW_HIP(hipMemsetAsync(dev, 0, f->npad, f->stream));
    return upload_flat(f, dev, m.data(), 1);
}

// pde.rs:550-631 with `open` already on the device.
cw_status flood_hold(cw_field* f, const uint8_t* open, uint8_t* reach, double value) {
    const Geo g = geo_of(f);
    const unsigned gb = grid_of(f->npad);
    hipLaunchKernelGGL(k_box_init, dim3(1), dim3(64), 0, f->stream, f->scan);
    hipLaunchKernelGGL(k_closed_box, dim3(red_grid(f)), dim3(kBlock), 0, f->stream, g, open,
                       red_slots(f));
    hipLaunchKernelGGL(k_box_fold, dim3(1), dim3(kBlock), 0, f->stream, (const int*)red_slots(f),
                       (int)red_grid(f), f->scan, f->scan + 1);
    int* box = f->host_small;  // 8 ints, pinned
    CW_HIP(hipMemcpyAsync(box, f->scan, 8 * sizeof(int), hipMemcpyDeviceToHost, f->stream));
    CW_HIP(hipStreamSynchronize(f->stream));
    Box b{0, -1, 0, -1, 0, -1};
    const int has_box = box[0];
    if (has_box) {
        const int lim[3] = {(int)f->W, (int)f->H, (int)f->D};
        int lo[3] = {box[1], box[3], box[5]}, hi[3] = {box[2], box[4], box[6]};
        for (int a = 0; a < 3; a++) {
            lo[a] = std::max(lo[a] - 1, 0);
            hi[a] = std::min(hi[a] + 1, lim[a] - 1);
        }
        b = Box{lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]};
        hipLaunchKernelGGL(k_flood_seed, dim3(gb), dim3(kBlock), 0, f->stream, g, b, open,
                           (const uint8_t*)f->held, reach);
        const bool vol = f->D > 1;
        const int TX = vol ? 32 : 64, TY = vol ? 8 : 16, TZ = vol ? 4 : 1;
        const long ntiles = (long)((b.x1 - b.x0 + TX) / TX) * ((b.y1 - b.y0 + TY) / TY) *
                            ((b.z1 - b.z0 + TZ) / TZ);
        // Launches go in batches of kFloodBatch, each with its own change
        // flag, and the host reads the flags once a batch: the reach set
        // only grows towards its one fixed point, so once a launch changes
        // nothing the later ones change nothing either, and the batch's
        // last flag says whether another batch is needed. (A lattice step's
        // flood settles in about four launches; one host round trip each
        // cost more than the launches.)
        constexpr int kFloodBatch = 4;
        int* changed = f->scan + 16;
        const long cap = 4 * ntiles + 64;
        for (long it = 0;; it += kFloodBatch) {
            if (it >= cap) return fail(CW_ERR_DEVICE, "bath flood did not settle");
            hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, f->stream, changed, kFloodBatch);
            for (int j = 0; j < kFloodBatch; j++) {
                if (vol)
                    hipLaunchKernelGGL((k_flood<32, 8, 4>), dim3((unsigned)ntiles), dim3(kBlock),
                                       0, f->stream, g, b, open, (const uint8_t*)f->held, reach,
                                       changed + j);
                else
                    hipLaunchKernelGGL((k_flood<64, 16, 1>), dim3((unsigned)ntiles), dim3(kBlock),
                                       0, f->stream, g, b, open, (const uint8_t*)f->held, reach,
                                       changed + j);
            }
            CW_HIP(hipGetLastError());
            int* ch = f->host_small + 8;  // pinned, after the box
            CW_HIP(hipMemcpyAsync(ch, changed, kFloodBatch * sizeof(int), hipMemcpyDeviceToHost,
                                  f->stream));
            CW_HIP(hipStreamSynchronize(f->stream));
            if (!ch[kFloodBatch - 1]) break;
        }
    }
    hipLaunchKernelGGL(k_flood_finish, dim3(gb), dim3(kBlock), 0, f->stream, g, b, has_box,
                       (const uint8_t*)reach, value, f->held, f->c[f->cur]);
    f->coef_dirty = true;
    return finish(f);
}

cw_status hold(cw_field* f, int mode, const uint8_t* mask_dev, const uint32_t* ids, double value) {
    hipLaunchKernelGGL(k_hold, dim3(grid_of(f->npad)), dim3(kBlock), 0, f->stream, geo_of(f),
                       mode, mask_dev, ids, value, f->held, f->c[f->cur]);
    f->coef_dirty = true;
    return finish(f);
}

cw_status build_csr(cw_lattice* l, uint32_t ncells) {
    if (l->csr_valid && l->csr_ncells == ncells) return CW_OK;
    if (l->n > 0xFFFFFFFFull || l->npad > 0xFFFFFFFFull)
        return fail(CW_ERR_CONFIG, "per-cell means need fewer than 2^32 voxels");
    const size_t n = l->n;
    if (!l->keys[0]) {
        for (int q = 0; q < 2; q++) {
            CW_HIP(hipMalloc((void**)&l->keys[q], n * sizeof(uint32_t)));
            CW_HIP(hipMalloc((void**)&l->vals[q], n * sizeof(uint32_t)));
        }
    }
    if (ncells > l->range_cap) {
        (void)hipFree(l->first);
        (void)hipFree(l->last);
        l->first = l->last = nullptr;
        l->range_cap = 0;
        CW_HIP(hipMalloc((void**)&l->first, std::max<size_t>(ncells, 1) * sizeof(uint32_t)));
        CW_HIP(hipMalloc((void**)&l->last, std::max<size_t>(ncells, 1) * sizeof(uint32_t)));
        l->range_cap = ncells;
    }
    const Geo g{l->P, (int)l->W, (int)l->H, (int)l->D, l->npad};
    const unsigned gn = grid_of(n);
    hipLaunchKernelGGL(k_cell_keys, dim3(gn), dim3(kBlock), 0, l->stream, g,
                       (const uint32_t*)l->ids, ncells, l->keys[0], l->vals[0]);
    unsigned end_bit = 1;
    while (end_bit < 32 && ((uint64_t)1 << end_bit) <= ncells) end_bit++;
    size_t need = 0;
    CW_HIP(rocprim::radix_sort_pairs(nullptr, need, l->keys[0], l->keys[1], l->vals[0],
                                     l->vals[1], n, 0, end_bit, l->stream));
    if (need > l->sort_tmp_bytes) {
        (void)hipFree(l->sort_tmp);
        l->sort_tmp = nullptr;
        l->sort_tmp_bytes = 0;
        CW_HIP(hipMalloc(&l->sort_tmp, need));
        l->sort_tmp_bytes = need;
    }
    CW_HIP(rocprim::radix_sort_pairs(l->sort_tmp, need, l->keys[0], l->keys[1], l->vals[0],
                                     l->vals[1], n, 0, end_bit, l->stream));
    const unsigned gc = grid_of(std::max<size_t>(ncells, 1));
    hipLaunchKernelGGL(k_fill_u32, dim3(gc), dim3(kBlock), 0, l->stream, l->first, (size_t)ncells, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(gc), dim3(kBlock), 0, l->stream, l->last, (size_t)ncells, 0u);
    hipLaunchKernelGGL(k_cell_ranges, dim3(gn), dim3(kBlock), 0, l->stream,
                       (const uint32_t*)l->keys[1], n, ncells, l->first, l->last);
    CW_HIP(hipGetLastError());
    CW_HIP(hipStreamSynchronize(l->stream));
    l->csr_ncells = ncells;
    l->csr_valid = true;
    return CW_OK;
}

}  // namespace

// ---- shared with the solve ------------------------------------------------------
namespace cwi {

cw_status build_coefficients(cw_field* f, bool may_alloc) {
    const Geo g = geo_of(f);
    const unsigned gb = grid_of(f->npad);
    hipLaunchKernelGGL(k_scan_init, dim3(1), dim3(64), 0, f->stream, f->scan);
    hipLaunchKernelGGL(k_scan, dim3(red_grid(f)), dim3(kBlock), 0, f->stream, g,
                       (const uint8_t*)f->held,
                       (const double*)f->up, (const double*)f->pr, f->source, red_slots(f));
    hipLaunchKernelGGL(k_box_fold, dim3(1), dim3(kBlock), 0, f->stream, (const int*)red_slots(f),
                       (int)red_grid(f), f->scan + 10, f->scan + 4);
    int* raw = f->host_small;  // 11 ints, pinned
    CW_HIP(hipMemcpyAsync(raw, f->scan, 11 * sizeof(int), hipMemcpyDeviceToHost, f->stream));
    CW_HIP(hipStreamSynchronize(f->stream));
    // scan[10] holds the flag bits: 1 general, 2 folded sink, 4 general sink, 8 any free
    int sc[10];
    for (int a = 0; a < 10; a++) sc[a] = raw[a];
    for (int b = 0; b < 4; b++) sc[b] = (raw[10] >> b) & 1;
    const bool general = f->source != 0.0 || sc[0];
    if (general && !f->k2) {
        if (!may_alloc)
            return fail(CW_ERR_ARG,
                        "the field keeps production and uptake apart and has no buffer for it "
                        "yet; an allocating call (cw_field_relax) creates it");
        cw_status st = alloc_fill((void**)&f->k2, f->npad * sizeof(double), f->stream, 0);
        if (st) return st;
    }
    hipLaunchKernelGGL(k_build, dim3(gb), dim3(kBlock), 0, f->stream, g, (const uint8_t*)f->held,
                       (const double*)f->up, (const double*)f->pr, (int)general, f->k, f->k2);
    CW_HIP(hipGetLastError());
    if (f->r32) CW_HIP(hipMemsetAsync(f->r32, 0xFF, f->npad * sizeof(float), f->stream));
    // The FP32 smoother needs a zero correction on every held cell; outside
    // the new free span no pass writes it, so clear what an earlier, larger
    // span may have left there.
    for (int b = 0; b < 2; b++)
        if (f->e[b]) CW_HIP(hipMemsetAsync(f->e[b], 0, f->npad * sizeof(float), f->stream));
    // no wait here: whatever reads the coefficients runs on the field's stream
    f->general = general;
    f->has_sink = general ? sc[2] != 0 : sc[1] != 0;
    f->has_span = sc[3] != 0;
    for (int a = 0; a < 6; a++) f->span[a] = f->has_span ? sc[4 + a] : 0;
    f->coef_dirty = false;
    return CW_OK;
}

cw_status upload_flat(const cw_field* f, void* dst, const void* src, int elem) {
    return copy_in(f->stream, (int)f->W, (size_t)f->H * f->D, f->P, dst, src, elem, &f->stage,
                   &f->stage_bytes);
}

cw_status download_flat(const cw_field* f, const void* src, void* dst, int elem) {
    return copy_out(f->stream, (int)f->W, (size_t)f->H * f->D, f->P, src, dst, elem, &f->stage,
                    &f->stage_bytes);
}

cw_status download_held(const cw_field* f, uint8_t* out) {
    return download_flat(f, f->held, out, 1);
}

cw_status download_f64(const cw_field* f, const double* dev, double* out) {
    return download_flat(f, dev, out, 8);
}

}  // namespace cwi

// ---- C ABI ------------------------------------------------------------------------
extern "C" {

cw_status cw_lattice_create(uint32_t w, uint32_t h, uint32_t d, cw_lattice** out) {
    if (!out) return fail(CW_ERR_ARG, "null output handle");
    *out = nullptr;
    if (w == 0 || h == 0 || d == 0)
        return fail(CW_ERR_CONFIG, "lattice dimensions must be positive, got %ux%ux%u", w, h, d);
    if ((uint64_t)w * h * d > (1ull << 34)) return fail(CW_ERR_CONFIG, "lattice too large");
    cw_status st = device_init();
    if (st) return st;
    cw_lattice* l = new cw_lattice();
    l->W = w; l->H = h; l->D = d;
    l->P = (int)((w + 63) / 64 * 64);
    l->n = (size_t)w * h * d;
    l->npad = (size_t)l->P * h * d;
    auto bail = [&](cw_status s) { cw_lattice_destroy(l); return s; };
    if (hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(CW_ERR_DEVICE, "hipStreamCreate failed"));
    if ((st = alloc_fill((void**)&l->ids, l->npad * sizeof(uint32_t), l->stream, 0))) return bail(st);
    if (hipStreamSynchronize(l->stream) != hipSuccess)
        return bail(fail(CW_ERR_DEVICE, "lattice set-up failed"));
    *out = l;
    return CW_OK;
}

cw_status cw_lattice_destroy(cw_lattice* l) {
    if (!l) return CW_OK;
    if (l->stream) (void)hipStreamSynchronize(l->stream);
    (void)hipFree(l->ids); (void)hipFree(l->flags); (void)hipFree(l->table);
    for (int q = 0; q < 2; q++) { (void)hipFree(l->keys[q]); (void)hipFree(l->vals[q]); }
    (void)hipFree(l->first); (void)hipFree(l->last); (void)hipFree(l->sort_tmp);
    (void)hipFree(l->stage);
    if (l->stream) (void)hipStreamDestroy(l->stream);
    delete l;
    return CW_OK;
}

cw_status cw_lattice_upload(cw_lattice* l, const uint32_t* ids) {
    if (!l || !ids) return fail(CW_ERR_ARG, "null argument");
    cw_status st = copy_in(l->stream, (int)l->W, (size_t)l->H * l->D, l->P, l->ids, ids, 4,
                           &l->stage, &l->stage_bytes);
    if (st) return st;
    l->csr_valid = false;
    return CW_OK;
}

cw_status cw_lattice_set(cw_lattice* l, const uint64_t* voxels, const uint32_t* ids,
                         uint64_t count) {
    if (!l) return fail(CW_ERR_ARG, "null lattice");
    if (count == 0) return CW_OK;
    if (!voxels || !ids) return fail(CW_ERR_ARG, "null voxel or id list");
    // Keep the last entry for each voxel, as applying the list in order would.
    std::vector<std::pair<uint64_t, uint64_t>> order((size_t)count);  // (voxel, position)
    for (uint64_t j = 0; j < count; j++) {
        if (voxels[j] >= l->n)
            return fail(CW_ERR_ARG, "voxel %llu outside the %ux%ux%u lattice",
                        (unsigned long long)voxels[j], l->W, l->H, l->D);
        order[(size_t)j] = {voxels[j], j};
    }
    std::sort(order.begin(), order.end());
    std::vector<uint64_t> v;
    std::vector<uint32_t> id;
    v.reserve((size_t)count);
    id.reserve((size_t)count);
    for (size_t k = 0; k < order.size(); k++) {
        if (k + 1 < order.size() && order[k + 1].first == order[k].first) continue;
        v.push_back(order[k].first);
        id.push_back(ids[order[k].second]);
    }
    const size_t n = v.size();
    cw_status st = stage_ensure(&l->stage, &l->stage_bytes, n * (sizeof(uint64_t) + sizeof(uint32_t)));
    if (st) return st;
    uint64_t* dv = (uint64_t*)l->stage;
    uint32_t* did = (uint32_t*)(dv + n);
    CW_HIP(hipMemcpyAsync(dv, v.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice, l->stream));
    CW_HIP(hipMemcpyAsync(did, id.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, l->stream));
    hipLaunchKernelGGL(k_lattice_set, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       l->stream, (const uint64_t*)dv, (const uint32_t*)did, n, (int)l->W, l->P,
                       l->ids);
    CW_HIP(hipGetLastError());
    CW_HIP(hipStreamSynchronize(l->stream));
    l->csr_valid = false;
    return CW_OK;
}

cw_status cw_field_apply_uptake(cw_field* f, cw_lattice* l, const uint8_t* consumes,
                                uint32_t ncells, double rate) {
    cw_status st = check_pair(f, l);
    if (st) return st;
    if (ncells && !consumes) return fail(CW_ERR_ARG, "null consumes table");
    if ((st = stage_tables(l, f->stream, consumes, nullptr, ncells))) return st;
    hipLaunchKernelGGL(k_uptake, dim3(grid_of(f->npad)), dim3(kBlock), 0, f->stream, geo_of(f),
                       (const uint32_t*)l->ids, (const uint8_t*)l->flags, (const double*)nullptr,
                       ncells, rate, f->up);
    f->coef_dirty = true;
    return finish(f);
}

cw_status cw_field_apply_uptake_per_cell(cw_field* f, cw_lattice* l, const uint8_t* consumes,
                                         const double* rates, uint32_t ncells) {
    cw_status st = check_pair(f, l);
    if (st) return st;
    if (ncells && !rates) return fail(CW_ERR_ARG, "null rate table");
    if ((st = stage_tables(l, f->stream, consumes, rates, ncells))) return st;
    hipLaunchKernelGGL(k_uptake, dim3(grid_of(f->npad)), dim3(kBlock), 0, f->stream, geo_of(f),
                       (const uint32_t*)l->ids,
                       consumes ? (const uint8_t*)l->flags : (const uint8_t*)nullptr,
                       (const double*)l->table, ncells, 0.0, f->up);
    f->coef_dirty = true;
    return finish(f);
}

cw_status cw_field_apply_background_uptake(cw_field* f, cw_lattice* l, double rate) {
    cw_status st = check_pair(f, l);
    if (st) return st;
    hipLaunchKernelGGL(k_background, dim3(grid_of(f->npad)), dim3(kBlock), 0, f->stream,
                       geo_of(f), (const uint32_t*)l->ids, rate, f->up);
    f->coef_dirty = true;
    return finish(f);
}

cw_status cw_field_apply_production(cw_field* f, cw_lattice* l, const double* per_cell,
                                    uint32_t ncells) {
    cw_status st = check_pair(f, l);
    if (st) return st;
    if (ncells && !per_cell) return fail(CW_ERR_ARG, "null production table");
    if ((st = stage_tables(l, f->stream, nullptr, per_cell, ncells))) return st;
    hipLaunchKernelGGL(k_production, dim3(grid_of(f->npad)), dim3(kBlock), 0, f->stream,
                       geo_of(f), (const uint32_t*)l->ids, (const double*)l->table, ncells, f->pr);
    f->coef_dirty = true;
    return finish(f);
}

cw_status cw_field_hold_obstacles_at(cw_field* f, cw_lattice* l, double value) {
    cw_status st = check_pair(f, l);
    if (st) return st;
    return hold(f, 1, nullptr, l->ids, value);
}

cw_status cw_field_hold_where(cw_field* f, const uint8_t* mask, double value) {
    if (!f || !mask) return fail(CW_ERR_ARG, "null argument");
    Scratch8 s;
    cw_status st = scratch8(f, &s);
    if (st) return st;
    if ((st = upload_mask(f, mask, s.a))) return st;
    return hold(f, 0, s.a, nullptr, value);
}

cw_status cw_field_release_all(cw_field* f) {
    if (!f) return fail(CW_ERR_ARG, "null field");
    return hold(f, 3, nullptr, nullptr, 0.0);
}

cw_status cw_field_hold_bath_at(cw_field* f, const uint8_t* open, double value) {
    if (!f || !open) return fail(CW_ERR_ARG, "null argument");
    Scratch8 s;
    cw_status st = scratch8(f, &s);
    if (st) return st;
    if ((st = upload_mask(f, open, s.a))) return st;
    return flood_hold(f, s.a, s.b, value);
}

cw_status cw_field_hold_medium_bath(cw_field* f, cw_lattice* l, double value) {
    cw_status st = check_pair(f, l);
    if (st) return st;
    Scratch8 s;
    if ((st = scratch8(f, &s))) return st;
    hipLaunchKernelGGL(k_open_from_ids, dim3(grid_of(f->npad)), dim3(kBlock), 0, f->stream,
                       geo_of(f), (const uint32_t*)l->ids, s.a);
    CW_HIP(hipGetLastError());
    return flood_hold(f, s.a, s.b, value);
}

cw_status cw_field_free_count(cw_field* f, uint64_t* out) {
    if (!f || !out) return fail(CW_ERR_ARG, "null argument");
    unsigned long long* dev = (unsigned long long*)(f->scan + 32);
    unsigned long long* slots = (unsigned long long*)red_slots(f);
    hipLaunchKernelGGL(k_count_free, dim3(red_grid(f)), dim3(kBlock), 0, f->stream,
                       geo_of(f), (const uint8_t*)f->held, slots);
    hipLaunchKernelGGL(k_sum_u64, dim3(1), dim3(kBlock), 0, f->stream,
                       (const unsigned long long*)slots, (int)red_grid(f), dev);
    CW_HIP(hipGetLastError());
    unsigned long long v = 0;
    CW_HIP(hipMemcpyAsync(&v, dev, sizeof v, hipMemcpyDeviceToHost, f->stream));
    CW_HIP(hipStreamSynchronize(f->stream));
    *out = (uint64_t)v;
    return CW_OK;
}

cw_status cw_field_inflow(cw_field* f, double* out) {
    if (!f || !out) return fail(CW_ERR_ARG, "null argument");
    const size_t per_block = (size_t)kBlock * kInflowPer;
    const unsigned nb = (unsigned)((f->npad + per_block - 1) / per_block);
    double* part = nullptr;
    CW_HIP(hipMalloc((void**)&part, (nb + 1) * sizeof(double)));
    hipLaunchKernelGGL(k_inflow, dim3(nb), dim3(kBlock), 0, f->stream, geo_of(f),
                       (const double*)f->c[f->cur], (const uint8_t*)f->held, f->diff, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, f->stream, (const double*)part,
                       (int)nb, part + nb);
    double v = 0.0;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&v, part + nb, sizeof v, hipMemcpyDeviceToHost, f->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(f->stream);
    (void)hipFree(part);
    if (e != hipSuccess) return fail(CW_ERR_DEVICE, "inflow failed: %s", hipGetErrorString(e));
    *out = v;
    return CW_OK;
}

cw_status cw_field_step(cw_field* f, double dt) {
    if (!f) return fail(CW_ERR_ARG, "null field");
    // check_step, pde.rs:673-701
    if (std::isnan(dt) || dt <= 0.0)
        return fail(CW_ERR_CONFIG, "time step must be positive, got %g", dt);
    const double limit = f->D > 1 ? 0.25 * 2.0 / 3.0 : 0.25;
    if (f->diff * dt > limit)
        return fail(CW_ERR_CONFIG, "unstable time step: D*dt = %.4f is above %g (D = %g, dt = %g)",
                    f->diff * dt, limit, f->diff, dt);
    if (f->decay * dt > 1.0)
        return fail(CW_ERR_CONFIG,
                    "unstable time step: decay*dt = %.4f is above 1 (decay = %g, dt = %g)",
                    f->decay * dt, f->decay, dt);
    hipLaunchKernelGGL(k_step, dim3(grid_of(f->npad)), dim3(kBlock), 0, f->stream, geo_of(f),
                       (const double*)f->c[f->cur], f->c[f->cur ^ 1], (const uint8_t*)f->held,
                       (const double*)f->up, (const double*)f->pr, f->diff, f->decay, f->source,
                       dt);
    cw_status st = finish(f);
    if (st) return st;
    f->cur ^= 1;
    return CW_OK;
}

cw_status cw_field_mean_per_cell(cw_field* f, cw_lattice* l, uint32_t ncells, double* out) {
    cw_status st = check_pair(f, l);
    if (st) return st;
    if (!out && ncells) return fail(CW_ERR_ARG, "null output");
    if (ncells == 0) return CW_OK;
    if ((st = build_csr(l, ncells))) return st;
    if ((st = stage_tables(l, f->stream, nullptr, nullptr, ncells))) return st;
    hipLaunchKernelGGL(k_cell_means, dim3(grid_of((size_t)ncells * 64)), dim3(kBlock), 0, f->stream,
                       (const double*)f->c[f->cur], (const uint32_t*)l->vals[1],
                       (const uint32_t*)l->first, (const uint32_t*)l->last, ncells, l->table);
    CW_HIP(hipGetLastError());
    CW_HIP(hipMemcpyAsync(out, l->table, ncells * sizeof(double), hipMemcpyDeviceToHost, f->stream));
    CW_HIP(hipStreamSynchronize(f->stream));
    return CW_OK;
}

cw_status cw_field_download_state(cw_field* f, uint8_t* held, double* uptake,
                                  double* production) {
    if (!f) return fail(CW_ERR_ARG, "null field");
    cw_status st;
    if (held && (st = download_held(f, held))) return st;
    if (uptake && (st = download_f64(f, f->up, uptake))) return st;
    if (production && (st = download_f64(f, f->pr, production))) return st;
    return CW_OK;
}

}  // extern "C"
