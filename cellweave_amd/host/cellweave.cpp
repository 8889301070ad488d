// cellweave host side in C++ over the C ABI (see cellweave.hpp).
#include "cellweave.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace cellweave {

namespace {

[[noreturn]] void raise(cw_status st) {
    const std::string m = cw_last_error();
    switch (st) {
        case CW_ERR_CONFIG: throw ConfigError(m);
        case CW_ERR_NOT_CONVERGED:
        case CW_ERR_NONFINITE: throw SimulationError(m);
        case CW_ERR_ARG: throw std::invalid_argument(m);
        default: throw std::runtime_error("device error: " + m);
    }
}

void check(cw_status st) {
    if (st != CW_OK) raise(st);
}

std::string fmt(const char* f, double a, double b) {
    char buf[256];
    std::snprintf(buf, sizeof buf, f, a, b);
    return buf;
}

}  // namespace

// ---- Lattice -----------------------------------------------------------------

Lattice::Lattice(uint32_t width, uint32_t height, uint32_t depth)
    : w_(width), h_(height), d_(depth) {
    if (width == 0 || height == 0 || depth == 0)
        throw ConfigError("lattice dimensions must be positive");
    grid_.assign((size_t)width * height * depth, 0u);
}

Lattice::~Lattice() { (void)cw_lattice_destroy(dev_); }

void Lattice::set_grid(std::vector<uint32_t> ids) {
    if (ids.size() != grid_.size()) throw std::invalid_argument("lattice size mismatch");
    grid_ = std::move(ids);
    dirty_ = true;
}

void Lattice::set_ids(const std::vector<uint64_t>& voxels, const std::vector<uint32_t>& ids) {
    if (voxels.size() != ids.size()) throw std::invalid_argument("voxel and id lists differ");
    for (size_t j = 0; j < voxels.size(); j++) {
        if (voxels[j] >= grid_.size()) throw std::invalid_argument("voxel outside the lattice");
        grid_[voxels[j]] = ids[j];
    }
    if (dev_ && !dirty_ && !voxels.empty())
        check(cw_lattice_set(dev_, voxels.data(), ids.data(), voxels.size()));
}

void Lattice::set_cells(std::vector<uint8_t> consumes) { consumes_ = std::move(consumes); }

cw_lattice* Lattice::device() {
    if (!dev_) check(cw_lattice_create(w_, h_, d_, &dev_));
    if (dirty_) {
        check(cw_lattice_upload(dev_, grid_.data()));
        dirty_ = false;
    }
    return dev_;
}

// ---- ScalarField ---------------------------------------------------------------

ScalarField ScalarField::create(uint32_t width, uint32_t height, double d, double decay,
                                double initial) {
    return with_depth(width, height, 1, d, decay, initial);
}

ScalarField ScalarField::with_depth(uint32_t width, uint32_t height, uint32_t depth, double d,
                                    double decay, double initial) {
    // pde.rs:106-116, checked before anything touches the device
    if (width == 0 || height == 0 || depth == 0) {
        char buf[128];
        std::snprintf(buf, sizeof buf, "field dimensions must be positive, got %ux%ux%u", width,
                      height, depth);
        throw ConfigError(buf);
    }
    if (!std::isfinite(d) || !std::isfinite(decay) || d < 0.0 || decay < 0.0)
        throw ConfigError(fmt("diffusion and decay must be finite and not negative, got D = %g, "
                              "decay = %g", d, decay));
    ScalarField f;
    f.w_ = width;
    f.h_ = height;
    f.d_ = depth;
    f.n_ = (size_t)width * height * depth;
    f.values_.assign(f.n_, initial);
    f.held_.assign(f.n_, 0);
    f.uptake_.assign(f.n_, 0.0);
    f.production_.assign(f.n_, 0.0);
    check(cw_field_create(width, height, depth, d, decay, initial, &f.dev_));
    return f;
}

ScalarField::ScalarField(ScalarField&& o) noexcept { *this = std::move(o); }

ScalarField& ScalarField::operator=(ScalarField&& o) noexcept {
    if (this == &o) return *this;
    (void)cw_field_destroy(dev_);
    w_ = o.w_; h_ = o.h_; d_ = o.d_; n_ = o.n_; source_ = o.source_;
    dev_ = o.dev_;
    o.dev_ = nullptr;
    values_ = std::move(o.values_);
    uptake_ = std::move(o.uptake_);
    production_ = std::move(o.production_);
    held_ = std::move(o.held_);
    for (int a = 0; a < kArrays; a++) at_[a] = o.at_[a];
    return *this;
}

ScalarField::~ScalarField() { (void)cw_field_destroy(dev_); }

ScalarField ScalarField::with_source(double source) && {
    source_ = source;
    check(cw_field_set_source(dev_, source));
    return std::move(*this);
}

void ScalarField::pull(Array a) const {
    if (at_[a] != Side::Device) return;
    if (a == kValues)
        check(cw_field_download(dev_, values_.data()));
    else
        check(cw_field_download_state(dev_, a == kHeld ? held_.data() : nullptr,
                                      a == kUptake ? uptake_.data() : nullptr,
                                      a == kProduction ? production_.data() : nullptr));
    at_[a] = Side::Both;
}

cw_field* ScalarField::device() {
    const bool v = at_[kValues] == Side::Host, h = at_[kHeld] == Side::Host;
    const bool u = at_[kUptake] == Side::Host, p = at_[kProduction] == Side::Host;
    if (v || h || u || p) {
        check(cw_field_upload(dev_, v ? values_.data() : nullptr, h ? held_.data() : nullptr,
                              u ? uptake_.data() : nullptr, p ? production_.data() : nullptr));
        for (int a = 0; a < kArrays; a++)
            if (at_[a] == Side::Host) at_[a] = Side::Both;
    }
    return dev_;
}

void ScalarField::touched_on_device(std::initializer_list<Array> as) {
    for (Array a : as) at_[a] = Side::Device;
}

void ScalarField::touched_on_host(Array a) { at_[a] = Side::Host; }

void ScalarField::set_uptake(uint32_t x, uint32_t y, double rate) {
    if (x >= w_ || y >= h_) return;
    set_uptake_at(idx(x, y), rate);
}

void ScalarField::set_production(uint32_t x, uint32_t y, double rate) {
    if (x >= w_ || y >= h_) return;
    set_production_at(idx(x, y), rate);
}

// rate.max(0.0): a NaN rate clamps to zero as f64::max does
void ScalarField::set_uptake_at(size_t i, double rate) {
    if (i >= n_) return;
    pull(kUptake);
    uptake_[i] = rate > 0.0 ? rate : 0.0;
    touched_on_host(kUptake);
}

void ScalarField::set_production_at(size_t i, double rate) {
    if (i >= n_) return;
    pull(kProduction);
    production_[i] = rate > 0.0 ? rate : 0.0;
    touched_on_host(kProduction);
}

double ScalarField::value_at(size_t i) const {
    if (i >= n_) return 0.0;
    pull(kValues);
    return values_[i];
}

void ScalarField::clear_production() {
    std::fill(production_.begin(), production_.end(), 0.0);
    touched_on_host(kProduction);
}

void ScalarField::clear_uptake() {
    std::fill(uptake_.begin(), uptake_.end(), 0.0);
    touched_on_host(kUptake);
}

void ScalarField::fill(double value) {
    std::fill(values_.begin(), values_.end(), value);
    check(cw_field_fill(dev_, value));
    at_[kValues] = Side::Both;
}

size_t ScalarField::relax(TimeStep, double tolerance, size_t max_sweeps) {
    uint64_t sweeps = 0;
    const cw_status st = cw_field_relax(device(), tolerance, (uint64_t)max_sweeps, &sweeps);
    touched_on_device({kValues});  // the sweeps ran even when they did not settle
    check(st);
    return (size_t)sweeps;
}

cw_relax_stats ScalarField::relax_with(const cw_relax_opts& opts) {
    cw_relax_stats s{};
    const cw_status st = cw_field_relax_ex(device(), &opts, &s);
    touched_on_device({kValues});
    check(st);
    return s;
}

double ScalarField::inflow() const {
    double v = 0.0;
    check(cw_field_inflow(const_cast<ScalarField*>(this)->device(), &v));
    return v;
}

size_t ScalarField::free_count() const {
    uint64_t v = 0;
    check(cw_field_free_count(const_cast<ScalarField*>(this)->device(), &v));
    return (size_t)v;
}

void ScalarField::release_all() {
    std::fill(held_.begin(), held_.end(), 0);
    touched_on_host(kHeld);
}

std::optional<Span> ScalarField::free_span() const {
    pull(kHeld);
    bool found = false;
    Span s{0, 0, 0, 0, 0, 0};
    for (size_t i = 0; i < n_; i++) {
        if (held_[i]) continue;
        const uint32_t x = (uint32_t)(i % w_), y = (uint32_t)((i / w_) % h_);
        const uint32_t z = (uint32_t)(i / ((size_t)w_ * h_));
        if (!found) {
            s = Span{x, x, y, y, z, z};
            found = true;
        }
        s.x0 = std::min(s.x0, x); s.x1 = std::max(s.x1, x);
        s.y0 = std::min(s.y0, y); s.y1 = std::max(s.y1, y);
        s.z1 = z;
    }
    if (!found) return std::nullopt;
    return s;
}

void ScalarField::hold_pixels(double value, const std::function<bool(uint32_t, uint32_t)>& is_source) {
    // pde.rs:477-488: the pixels of a flat field, by (x, y)
    std::vector<uint8_t> mask(n_, 0);
    for (uint32_t y = 0; y < h_; y++)
        for (uint32_t x = 0; x < w_; x++) mask[idx(x, y)] = is_source(x, y) ? 1 : 0;
    check(cw_field_hold_where(device(), mask.data(), value));
    touched_on_device({kValues, kHeld});
}

void ScalarField::hold_where(double value, const std::function<bool(size_t)>& is_source) {
    std::vector<uint8_t> mask(n_, 0);
    for (size_t i = 0; i < n_; i++) mask[i] = is_source(i) ? 1 : 0;
    check(cw_field_hold_where(device(), mask.data(), value));
    touched_on_device({kValues, kHeld});
}

void ScalarField::hold_pixels_by(
    const std::function<std::optional<double>(uint32_t, uint32_t)>& level) {
    // each pixel its own value: on the host
    pull(kValues);
    pull(kHeld);
    for (uint32_t y = 0; y < h_; y++)
        for (uint32_t x = 0; x < w_; x++)
            if (const auto v = level(x, y)) {
                held_[idx(x, y)] = 1;
                values_[idx(x, y)] = *v;
            }
    touched_on_host(kValues);
    touched_on_host(kHeld);
}

void ScalarField::hold_bath(double value, const std::function<bool(uint32_t, uint32_t)>& is_open) {
    // pde.rs:535-538
    const uint32_t w = w_;
    hold_bath_at(value, [&](size_t i) { return is_open((uint32_t)(i % w), (uint32_t)(i / w)); });
}

void ScalarField::hold_bath_at(double value, const std::function<bool(size_t)>& is_open) {
    std::vector<uint8_t> open(n_, 0);
    for (size_t i = 0; i < n_; i++) open[i] = is_open(i) ? 1 : 0;
    check(cw_field_hold_bath_at(device(), open.data(), value));
    touched_on_device({kValues, kHeld});
}

void ScalarField::step(TimeStep dt) {
    const cw_status st = cw_field_step(device(), dt.value);
    check(st);
    touched_on_device({kValues});
}

double ScalarField::concentration_at(uint32_t x, uint32_t y) const {
    if (x >= w_ || y >= h_) return 0.0;
    pull(kValues);
    return values_[idx(x, y)];
}

const std::vector<double>& ScalarField::values() const { pull(kValues); return values_; }
const std::vector<double>& ScalarField::uptake() const { pull(kUptake); return uptake_; }
const std::vector<double>& ScalarField::production() const { pull(kProduction); return production_; }
const std::vector<uint8_t>& ScalarField::held() const { pull(kHeld); return held_; }

// ---- coupling.rs ----------------------------------------------------------------

namespace {
void same_shape(const ScalarField& f, const Lattice& l) {
    if (f.dims3() != l.dims3()) throw std::invalid_argument("lattice and field differ in size");
}
}  // namespace

void apply_uptake(ScalarField& field, Lattice& lattice, double rate) {
    same_shape(field, lattice);
    const auto& c = lattice.consumes();
    check(cw_field_apply_uptake(field.device(), lattice.device(), c.empty() ? nullptr : c.data(),
                                (uint32_t)c.size(), rate));
    field.touched_on_device({ScalarField::kUptake});
}

void apply_background_uptake(ScalarField& field, Lattice& lattice, double rate) {
    same_shape(field, lattice);
    check(cw_field_apply_background_uptake(field.device(), lattice.device(), rate));
    field.touched_on_device({ScalarField::kUptake});
}

void apply_production(ScalarField& field, Lattice& lattice, const std::vector<double>& per_cell) {
    same_shape(field, lattice);
    check(cw_field_apply_production(field.device(), lattice.device(),
                                    per_cell.empty() ? nullptr : per_cell.data(),
                                    (uint32_t)per_cell.size()));
    field.touched_on_device({ScalarField::kProduction});
}

void hold_obstacles_at(ScalarField& field, Lattice& lattice, double value) {
    same_shape(field, lattice);
    check(cw_field_hold_obstacles_at(field.device(), lattice.device(), value));
    field.touched_on_device({ScalarField::kValues, ScalarField::kHeld});
}

void hold_vessels_at(ScalarField& field, Lattice& lattice, double value,
                     const std::function<bool(uint32_t, uint32_t)>& carries) {
    same_shape(field, lattice);
    const auto [w, h, d] = lattice.dims3();
    (void)d;
    const auto& grid = lattice.grid();
    std::vector<uint8_t> mask(grid.size(), 0);
    for (size_t i = 0; i < grid.size(); i++) {
        // position(i), cpm.rs:960-963
        const uint32_t x = (uint32_t)(i % w), y = (uint32_t)((i / w) % h);
        mask[i] = grid[i] == Lattice::OBSTACLE && carries(x, y) ? 1 : 0;
    }
    check(cw_field_hold_where(field.device(), mask.data(), value));
    field.touched_on_device({ScalarField::kValues, ScalarField::kHeld});
}

void hold_medium_bath(ScalarField& field, Lattice& lattice, double value) {
    same_shape(field, lattice);
    check(cw_field_hold_medium_bath(field.device(), lattice.device(), value));
    field.touched_on_device({ScalarField::kValues, ScalarField::kHeld});
}

std::vector<double> mean_per_cell(ScalarField& field, Lattice& lattice) {
    same_shape(field, lattice);
    std::vector<double> out(lattice.cell_count() + 1, 0.0);
    check(cw_field_mean_per_cell(field.device(), lattice.device(), (uint32_t)out.size(),
                                 out.data()));
    return out;
}

double mean_in(const ScalarField& field, const Lattice& lattice, uint32_t id) {
    same_shape(field, lattice);
    const auto& grid = lattice.grid();
    const auto& v = field.values();
    double sum = 0.0;
    uint32_t count = 0;
    for (size_t i = 0; i < grid.size(); i++)
        if (grid[i] == id) {
            sum = sum + v[i];
            count += 1;
        }
    return count == 0 ? 0.0 : sum / (double)count;
}

double still_medium_level(const ScalarField& oxygen, double diffusion) {
    const double pi = 3.14159265358979323846;
    const double taken = oxygen.inflow();
    const double radius = std::cbrt(3.0 * (double)oxygen.free_count() / (4.0 * pi));
    if (!(radius > 0.0)) return 1.0;
    const double level = 1.0 - taken / (4.0 * pi * diffusion * radius);
    return level > 0.0 ? level : 0.0;
}

}  // namespace cellweave
