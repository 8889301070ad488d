// cellweave host side in C++: ScalarField and the lattice coupling with the
// reference's interface (crates/engine/src/pde.rs, crates/engine/src/coupling.rs,
// crates/core/src/{error,types,traits}.rs), solved on the GPU through the C ABI
// of include/cwfield.h.
//
// Names, argument meanings and error behaviour follow the Rust functions; a
// Result's error becomes an exception (ConfigError for CellweaveError::Config,
// SimulationError for CellweaveError::Simulation). The field keeps host copies
// of its arrays, as ScalarField does, and knows for each whether the host or
// the device holds the current copy: per-voxel setters work on the host, the
// solve, the step and the gathers from a Lattice run on the device, and each
// side is brought up to date in one bulk copy only when the other needs it.
#pragma once

#include "../../include/cwfield.h"

#include <cstddef>
#include <cstdint>
#include <functional>
#include <optional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

namespace cellweave {

// crates/core/src/error.rs:7-14
class CellweaveError : public std::runtime_error {
  public:
    using std::runtime_error::runtime_error;
};
class ConfigError : public CellweaveError {
  public:
    explicit ConfigError(const std::string& m) : CellweaveError("configuration error: " + m) {}
};
class SimulationError : public CellweaveError {
  public:
    explicit SimulationError(const std::string& m) : CellweaveError("simulation error: " + m) {}
};

// crates/core/src/types.rs:29
struct TimeStep {
    double value;
    explicit TimeStep(double v) : value(v) {}
};

// The box around the free pixels, inclusive (pde.rs:305-320).
struct Span {
    uint32_t x0, x1, y0, y1, z0, z1;
    bool operator==(const Span& o) const {
        return x0 == o.x0 && x1 == o.x1 && y0 == o.y0 && y1 == o.y1 && z0 == o.z0 && z1 == o.z1;
    }
};

// The cell-id lattice (cpm.rs:194-198) as the coupling reads it: ids per
// pixel (0 the medium, OBSTACLE an obstacle) and, per id, whether the cell's
// kind consumes (coupling.rs:15-20). A device copy is uploaded when the ids
// change.
class Lattice {
  public:
    static constexpr uint32_t OBSTACLE = 0xFFFFFFFFu;  // cpm.rs:137

    Lattice(uint32_t width, uint32_t height, uint32_t depth = 1);
    ~Lattice();
    Lattice(const Lattice&) = delete;
    Lattice& operator=(const Lattice&) = delete;

    // Replace the ids (width*height*depth, x fastest).
    void set_grid(std::vector<uint32_t> ids);
    // Set the listed voxels (flat indices) to the listed ids, in order: what
    // a lattice step changed. A device copy already in use takes only these
    // voxels (cw_lattice_set) instead of the whole grid.
    void set_ids(const std::vector<uint64_t>& voxels, const std::vector<uint32_t>& ids);
    // consumes[id] for ids 0..cell_count(); index 0 is the medium.
    void set_cells(std::vector<uint8_t> consumes);
    uint32_t occupant_at(size_t i) const { return grid_.at(i); }
    size_t cell_count() const { return consumes_.empty() ? 0 : consumes_.size() - 1; }
    const std::vector<uint32_t>& grid() const { return grid_; }
    const std::vector<uint8_t>& consumes() const { return consumes_; }
    std::tuple<uint32_t, uint32_t, uint32_t> dims3() const { return {w_, h_, d_}; }
    cw_lattice* device();  // uploads the ids when they changed

  private:
    uint32_t w_, h_, d_;
    std::vector<uint32_t> grid_;
    std::vector<uint8_t> consumes_;
    cw_lattice* dev_ = nullptr;
    bool dirty_ = true;
};

class ScalarField {
  public:
    // pde.rs:82-131. ConfigError for a zero dimension or a negative or
    // non-finite diffusion coefficient or decay rate.
    static ScalarField create(uint32_t width, uint32_t height, double d, double decay,
                              double initial);
    static ScalarField with_depth(uint32_t width, uint32_t height, uint32_t depth, double d,
                                  double decay, double initial);
    ScalarField(ScalarField&& o) noexcept;
    ScalarField& operator=(ScalarField&& o) noexcept;
    ScalarField(const ScalarField&) = delete;
    ~ScalarField();

    ScalarField with_source(double source) &&;  // pde.rs:135-138

    // pde.rs:146-220: setters clamp at zero and ignore positions outside the grid.
    void set_uptake(uint32_t x, uint32_t y, double rate);
    void set_production(uint32_t x, uint32_t y, double rate);
    void set_uptake_at(size_t i, double rate);
    void set_production_at(size_t i, double rate);
    double value_at(size_t i) const;  // zero past the end
    size_t len() const { return n_; }
    bool is_empty() const { return n_ == 0; }
    std::tuple<uint32_t, uint32_t, uint32_t> dims3() const { return {w_, h_, d_}; }
    void clear_production();
    void fill(double value);
    void clear_uptake();

    // pde.rs:245-288: sweeps until no pixel moves by `tolerance`; the sweep
    // count, or SimulationError after max_sweeps.
    size_t relax(TimeStep dt, double tolerance, size_t max_sweeps);
    // relax_ex: every knob of cw_relax_opts (criterion, precision, omega, ...).
    cw_relax_stats relax_with(const cw_relax_opts& opts);

    double inflow() const;            // pde.rs:415-456
    size_t free_count() const;        // pde.rs:460-463
    void release_all();               // pde.rs:468-470
    std::optional<Span> free_span() const;  // pde.rs:305-320

    void hold_pixels(double value, const std::function<bool(uint32_t, uint32_t)>& is_source);
    void hold_where(double value, const std::function<bool(size_t)>& is_source);
    void hold_pixels_by(const std::function<std::optional<double>(uint32_t, uint32_t)>& level);
    void hold_bath(double value, const std::function<bool(uint32_t, uint32_t)>& is_open);
    void hold_bath_at(double value, const std::function<bool(size_t)>& is_open);

    // DiffusionField (crates/core/src/traits.rs:47-60, pde.rs:704-740)
    void step(TimeStep dt);
    double concentration_at(uint32_t x, uint32_t y) const;
    std::pair<uint32_t, uint32_t> dims() const { return {w_, h_}; }

    // host views (brought up to date from the device first)
    const std::vector<double>& values() const;
    const std::vector<double>& uptake() const;
    const std::vector<double>& production() const;
    const std::vector<uint8_t>& held() const;

    cw_field* device();  // the device state, brought up to date from the host

  private:
    friend void apply_uptake(ScalarField&, Lattice&, double);
    friend void apply_background_uptake(ScalarField&, Lattice&, double);
    friend void apply_production(ScalarField&, Lattice&, const std::vector<double>&);
    friend void hold_obstacles_at(ScalarField&, Lattice&, double);
    friend void hold_vessels_at(ScalarField&, Lattice&, double,
                                const std::function<bool(uint32_t, uint32_t)>&);
    friend void hold_medium_bath(ScalarField&, Lattice&, double);
    friend std::vector<double> mean_per_cell(ScalarField&, Lattice&);

    enum class Side { Host, Device, Both };
    enum Array { kValues = 0, kHeld, kUptake, kProduction, kArrays };

    ScalarField() = default;
    void pull(Array a) const;  // host copy up to date
    void touched_on_device(std::initializer_list<Array> as);
    void touched_on_host(Array a);
    size_t idx(uint32_t x, uint32_t y) const { return (size_t)y * w_ + x; }

    uint32_t w_ = 0, h_ = 0, d_ = 0;
    size_t n_ = 0;
    double source_ = 0.0;
    cw_field* dev_ = nullptr;
    mutable std::vector<double> values_, uptake_, production_;
    mutable std::vector<uint8_t> held_;
    mutable Side at_[kArrays] = {Side::Both, Side::Both, Side::Both, Side::Both};
};

// crates/engine/src/coupling.rs, on the device
void apply_uptake(ScalarField& field, Lattice& lattice, double rate);             // :26-39
void apply_background_uptake(ScalarField& field, Lattice& lattice, double rate);  // :47-53
void apply_production(ScalarField& field, Lattice& lattice,
                      const std::vector<double>& per_cell);                      // :61-73
void hold_obstacles_at(ScalarField& field, Lattice& lattice, double value);       // :80-82
// :89-99: every obstacle voxel whose (x, y) `carries` says carries blood is held at value
void hold_vessels_at(ScalarField& field, Lattice& lattice, double value,
                     const std::function<bool(uint32_t, uint32_t)>& carries);
void hold_medium_bath(ScalarField& field, Lattice& lattice, double value);        // :156-158
std::vector<double> mean_per_cell(ScalarField& field, Lattice& lattice);          // :164-182
// :187-201: the mean over the voxels whose occupant is `id` (any id, the
// medium and obstacles included), added in increasing voxel order; 0 for none
double mean_in(const ScalarField& field, const Lattice& lattice, uint32_t id);

// crates/engine/src/simulation.rs:2114-2125 (settle_bath_level): the level a
// still medium of diffusion coefficient `diffusion`, held far away at the
// source's level 1 and taking up nothing, sits at around the tissue: a ball of
// the free voxels' volume that takes up what the held voxels hand the field,
// 1 - inflow / (4 pi diffusion R), floored at 0; 1 when nothing is free.
double still_medium_level(const ScalarField& oxygen, double diffusion);

}  // namespace cellweave
