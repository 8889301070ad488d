// Device-driven batch of planes (the campaign of src/campaign.rs:196-204 run
// together on the GPU). Every field carries its own state in device memory;
// between two steps a one-workgroup kernel moves each field to its next phase
// and lists the tiles that work in the coming step, and a persistent step
// kernel (a fixed grid that loops over that list) does each tile's phase. A
// captured chain of steps therefore runs a whole solve with no host decision:
//   FP64: plain sweeps of s sweeps a step, the relative residual of the input
//         tested every `check_every` steps (the same kernels, cadence and bits
//         as cw_batch_relax);
//   FP32 refinement: FP64 residual (writes the smoother's right-hand side),
//         then a smoothing segment of FP32 sweeps whose length the transition
//         picks from the field's residual history, then the FP64 correction.
#pragma once

#include "cw_internal.hpp"
#include "cw_kernels.hpp"

namespace cw {

enum CampPhase : int {
    PH_INIT = 0,     // before the first step
    PH_SWEEP = 1,    // FP64: s plain sweeps
    PH_RESID = 2,    // FP32 refinement: FP64 residual of c
    PH_SMOOTH = 3,   // FP32 refinement: s smoother sweeps of A e = r
    PH_CORRECT = 4,  // FP32 refinement: c <- max(c + e, 0)
    PH_DONE = 5,
    PH_MAXED = 6     // stopped without converging (budget, or a non-finite residual)
};

__host__ __device__ inline bool camp_active(int phase) {
    return phase == PH_SWEEP || phase == PH_RESID || phase == PH_SMOOTH || phase == PH_CORRECT;
}

struct CampField {
    int phase;       // what the current step does
    int seg_start;   // FP32: step at which the current smoothing segment began
    int budget;      // FP32: sweeps asked of the current segment
    int seg_done;    // FP32: sweeps it has run
    int total;       // FP64: sweeps before the answer; FP32: smoother sweeps summed
    int outer;       // FP32: segments begun
    int parity;      // FP64: the state buffer holding the answer once stopped
    int nonfinite;
    int first, hi;   // FP32: first and longest segment (from the span box)
    int nhist;       // FP32: residual history entries kept (at most 2)
    int hist_at[2];  // FP32: sweeps at which hist[] was measured (hist[1] newest)
    int pad_;
    double hist[2];
    double r2, b2;   // last residual sums
    double model_rate;  // log of the model problem's SOR rate per sweep (< 0)
};

// The FP32 step's eight per-XCD tile counters sit after a group's tile count,
// at count[kCampCounterStride * (x + 1)] (a cache line apart).
constexpr int kCampCounterStride = 32;

struct CampArgs {
    PassArgs pa;           // sweep view: problems, tile table (main tiling), partials,
                           // control blocks (FP64: the verdict on the input), s, run base
    const Prob* probs64;   // FP32 refinement: the FP64 view of every problem
    CampField* st;
    int* list;             // the tiles that work in the current step
    int* count;            // the list's length, then the FP32 step's per-XCD counters
    int fp64;
    int check_every;
    int lo;                // shortest segment
    int max_sweeps, max_outer;
    int slot_base;         // FP64 step: first list slot of this launch (a step is split
                           // over several launches on several queues)
    double tol, inner_tol;
};

// A group (size class) as the transition kernels read it: its arguments, with
// the step index supplied by the launch, and its fields' place in the one
// numbering of every group's fields.
struct CampGroupDev {
    CampArgs ca;
    int field_base;
    int nfields;
    int pad_[2];
};

}  // namespace cw

namespace cwi {

// Plane shapes the step kernel is compiled for (indices into kPlainShapes for
// FP64, kSmoothShapes for FP32).
bool camp_shape_supported(int shape, bool fp64);
// Launches a full-width FP64 step is split into (each on its own queue).
constexpr int kCampSplit = 4;
// Grid of one launch of a group's step kernel: a full-width FP64 step, the
// group's tiles over kCampSplit launches; FP32 and the FP64 tail (persistent
// loops over the list), the workgroups resident on the device at once (a
// multiple of 8).
cw_status camp_grid(int shape, bool fp64, bool gen, bool tail, long ntiles, int* grid);
// Between two steps: every group's fields move to their next phase, then
// every group's tile list is rebuilt (groups: nfields fields in all, device
// memory; the step index is `step` plus run[0]).
cw_status camp_launch_transition(const cw::CampGroupDev* groups, int ngroups, int nfields,
                                 int step, const int* run, hipStream_t stream);
// One launch of a group's step kernel (FP64: full width, or the tail form).
cw_status camp_launch_step(int shape, bool fp64, bool gen, bool tail, const cw::CampArgs& ca,
                           int grid, hipStream_t stream);

}  // namespace cwi
