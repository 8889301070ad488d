// The reference's ScalarField and coupling tests (crates/engine/src/pde.rs
// tests module, crates/engine/src/coupling.rs tests module), restated against
// the C++ host side and run on the GPU. Each test keeps the reference's name;
// the assertions check what the reference's do.
//
//   build/test_reference            run every test
//   build/test_reference NAME ...   run the named ones
#include "cellweave.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

using namespace cellweave;

namespace {

struct Failure {
    std::string what;
};

std::vector<std::pair<const char*, void (*)()>>& registry() {
    static std::vector<std::pair<const char*, void (*)()>> r;
    return r;
}

struct Reg {
    Reg(const char* n, void (*f)()) { registry().push_back({n, f}); }
};

#define TEST(name)                          \
    static void name();                     \
    static Reg reg_##name(#name, name);     \
    static void name()

#define CHECK(cond)                                                                   \
    do {                                                                              \
        if (!(cond)) {                                                                \
            std::ostringstream os_;                                                   \
            os_ << __FILE__ << ":" << __LINE__ << ": " << #cond;                      \
            throw Failure{os_.str()};                                                 \
        }                                                                             \
    } while (0)

#define CHECK_THROWS(expr, type)                                                      \
    do {                                                                              \
        bool thrown_ = false;                                                         \
        try {                                                                         \
            expr;                                                                     \
        } catch (const type&) {                                                       \
            thrown_ = true;                                                           \
        }                                                                             \
        if (!thrown_) throw Failure{std::string("expected " #type " from " #expr)};   \
    } while (0)

const double D = 0.2;  // pde.rs:758

// pde.rs:748-755
std::function<bool(uint32_t, uint32_t)> outside_disc(uint32_t size, uint32_t radius) {
    const double c = (double)(size / 2), r2 = (double)radius * radius;
    return [=](uint32_t x, uint32_t y) {
        const double dx = (double)x - c, dy = (double)y - c;
        return dx * dx + dy * dy > r2;
    };
}

// pde.rs:762-775
ScalarField bathed_disc(uint32_t size, uint32_t radius, double uptake) {
    auto outside = outside_disc(size, radius);
    auto field = ScalarField::create(size, size, D, 0.0, 1.0);
    for (uint32_t y = 0; y < size; y++)
        for (uint32_t x = 0; x < size; x++)
            if (!outside(x, y)) field.set_uptake(x, y, uptake);
    field.hold_bath(1.0, outside);
    field.relax(TimeStep(1.0), 1e-11, 50000);
    return field;
}

double disc_centre(uint32_t size, uint32_t radius, double uptake) {
    return bathed_disc(size, radius, uptake).concentration_at(size / 2, size / 2);
}

}  // namespace

// ---- pde.rs ---------------------------------------------------------------------

TEST(the_size_of_the_box_does_not_change_the_tissue) {  // pde.rs:783-807
    const double small = disc_centre(30, 6, 0.004), large = disc_centre(60, 6, 0.004);
    CHECK(std::fabs(small - large) < 1e-9);
    CHECK(small < 0.9);
    CHECK(small > 0.1);
}

TEST(a_bathed_disc_matches_the_textbook_profile) {  // pde.rs:810-830
    const uint32_t R = 8;
    const double q = 0.003;
    const double measured = disc_centre(32, R, q);
    const double depth = q * (double)(R * R) / (4.0 * D);
    CHECK(std::fabs(measured - (1.0 - depth)) / depth < 0.12);
}

TEST(a_pocket_walled_in_by_tissue_is_never_reached) {  // pde.rs:833-845
    auto f = ScalarField::create(11, 11, D, 0.0, 0.0);
    f.hold_bath(1.0, [](uint32_t x, uint32_t y) {
        return std::max(std::abs((int)x - 5), std::abs((int)y - 5)) != 3;
    });
    CHECK(f.concentration_at(5, 5) == 0.0);
    CHECK(f.concentration_at(0, 0) == 1.0);
}

TEST(one_gap_in_the_wall_lets_the_bath_in) {  // pde.rs:848-858
    auto f = ScalarField::create(11, 11, D, 0.0, 0.0);
    f.hold_bath(1.0, [](uint32_t x, uint32_t y) {
        const bool wall = std::max(std::abs((int)x - 5), std::abs((int)y - 5)) == 3 &&
                          !(x == 5 && y == 2);
        return !wall;
    });
    CHECK(f.concentration_at(5, 5) == 1.0);
}

TEST(a_wall_that_only_touches_at_corners_still_seals) {  // pde.rs:861-880
    auto f = ScalarField::create(11, 11, D, 0.0, 0.0);
    f.hold_bath(1.0, [](uint32_t x, uint32_t y) {
        return std::abs((int)x - 5) + std::abs((int)y - 5) != 3;
    });
    CHECK(f.concentration_at(5, 5) == 0.0);
    CHECK(f.concentration_at(0, 0) == 1.0);
}

TEST(a_closed_box_keeps_what_it_holds) {  // pde.rs:883-905
    auto f = ScalarField::create(20, 20, D, 0.0, 0.0);
    f.hold_pixels_by([](uint32_t x, uint32_t y) -> std::optional<double> {
        if (x >= 2 && x < 6 && y >= 2 && y < 6) return 1.0;
        return std::nullopt;
    });
    f.release_all();  // the square keeps its value and moves freely
    double before = 0.0;
    for (double v : f.values()) before += v;
    for (int k = 0; k < 500; k++) f.step(TimeStep(1.0));
    double after = 0.0;
    for (double v : f.values()) after += v;
    CHECK(std::fabs(after - before) < 1e-9);
}

TEST(an_unstable_time_step_is_refused) {  // pde.rs:908-914
    auto f = ScalarField::create(20, 20, 0.3, 0.0, 1.0);
    CHECK_THROWS(f.step(TimeStep(1.0)), ConfigError);
}

TEST(a_decay_that_removes_more_than_everything_is_refused) {  // pde.rs:917-923
    auto f = ScalarField::create(20, 20, 0.1, 2.0, 1.0);
    CHECK_THROWS(f.step(TimeStep(1.0)), ConfigError);
}

TEST(a_time_step_going_nowhere_is_refused) {  // pde.rs:926-931
    auto f = ScalarField::create(20, 20, 0.1, 0.0, 1.0);
    CHECK_THROWS(f.step(TimeStep(0.0)), ConfigError);
    CHECK_THROWS(f.step(TimeStep(-1.0)), ConfigError);
}

TEST(a_field_with_no_pixels_is_refused) {  // pde.rs:934-938
    CHECK_THROWS(ScalarField::create(0, 10, D, 0.0, 1.0), ConfigError);
    CHECK_THROWS(ScalarField::create(10, 0, D, 0.0, 1.0), ConfigError);
}

TEST(a_field_that_is_not_a_number_is_refused) {  // pde.rs:941-949
    CHECK_THROWS(ScalarField::create(10, 10, NAN, 0.0, 1.0), ConfigError);
    CHECK_THROWS(ScalarField::create(10, 10, D, NAN, 1.0), ConfigError);
    CHECK_THROWS(ScalarField::create(10, 10, INFINITY, 0.0, 1.0), ConfigError);
    auto f = ScalarField::create(10, 10, D, 0.0, 1.0);
    CHECK_THROWS(f.step(TimeStep(NAN)), ConfigError);
}

TEST(field_decays_toward_zero_without_source) {  // pde.rs:952-961
    auto f = ScalarField::create(20, 20, 0.25, 0.1, 1.0);
    for (int k = 0; k < 200; k++) f.step(TimeStep(0.1));
    CHECK(f.concentration_at(10, 10) < 0.5);
}

TEST(concentration_never_negative) {  // pde.rs:964-979
    auto f = ScalarField::create(30, 30, 0.5, 2.0, 1.0);
    for (int k = 0; k < 500; k++) f.step(TimeStep(0.05));
    for (double v : f.values()) CHECK(v >= 0.0);
}

TEST(the_number_of_threads_does_not_change_a_single_bit) {  // pde.rs:982-1000
    // On the device the analogue is that two solves of the same inputs give
    // the same bits, whatever the launch order of their tiles.
    auto a = bathed_disc(61, 20, 0.0005);
    auto b = bathed_disc(61, 20, 0.0005);
    CHECK(std::memcmp(a.values().data(), b.values().data(), a.len() * sizeof(double)) == 0);
}

TEST(what_the_bath_hands_a_settled_tissue_is_what_it_takes_up) {  // pde.rs:1003-1012
    auto f = bathed_disc(30, 6, 0.004);
    const double taken = 0.004 * (double)f.free_count();
    CHECK(f.free_count() > 100);
    CHECK(std::fabs(f.inflow() - taken) < 1e-6 * taken);
}

TEST(relaxing_settles_the_field_and_says_how_long_it_took) {  // pde.rs:1015-1027
    auto f = bathed_disc(30, 6, 0.004);
    const std::vector<double> settled = f.values();
    CHECK(f.relax(TimeStep(1.0), 1e-11, 50000) == 1);
    for (size_t i = 0; i < f.len(); i++) CHECK(std::fabs(f.value_at(i) - settled[i]) < 1e-6);
}

TEST(relaxing_refuses_to_pretend_when_it_cannot_settle_in_time) {  // pde.rs:1030-1045
    auto outside = outside_disc(30, 6);
    auto f = ScalarField::create(30, 30, D, 0.0, 1.0);
    for (uint32_t y = 0; y < 30; y++)
        for (uint32_t x = 0; x < 30; x++)
            if (!outside(x, y)) f.set_uptake(x, y, 0.004);
    f.hold_bath(1.0, outside);
    CHECK_THROWS(f.relax(TimeStep(1.0), 1e-9, 3), SimulationError);
}

TEST(a_held_pixel_is_a_boundary_condition_and_does_not_move) {  // pde.rs:1048-1067
    auto f = bathed_disc(30, 6, 0.05);
    for (int k = 0; k < 200; k++) f.step(TimeStep(1.0));
    CHECK(f.concentration_at(0, 0) == 1.0);
    CHECK(f.concentration_at(15, 3) == 1.0);
    CHECK(f.concentration_at(15, 15) < 0.5);
}

TEST(clearing_the_uptake_lets_the_field_refill) {  // pde.rs:1070-1081
    auto f = bathed_disc(30, 6, 0.01);
    CHECK(f.concentration_at(15, 15) < 0.9);
    f.clear_uptake();
    f.relax(TimeStep(1.0), 1e-7, 50000);
    CHECK(std::fabs(f.concentration_at(15, 15) - 1.0) < 1e-5);
}

TEST(a_held_vessel_supplies_from_the_inside_out) {  // pde.rs:1084-1113
    auto f = ScalarField::create(60, 60, D, 0.0, 0.0);
    auto vessel = [](uint32_t x, uint32_t y) {
        const int dx = (int)x - 30, dy = (int)y - 30;
        return dx * dx + dy * dy <= 9;
    };
    for (uint32_t y = 0; y < 60; y++)
        for (uint32_t x = 0; x < 60; x++)
            if (!vessel(x, y)) f.set_uptake(x, y, 0.002);
    f.hold_pixels(1.0, vessel);
    f.relax(TimeStep(1.0), 1e-7, 50000);
    auto at = [&](uint32_t r) { return f.concentration_at(30 + r, 30); };
    CHECK(at(0) == 1.0);
    CHECK(at(5) > at(10));
    CHECK(at(10) > at(15));
    CHECK(f.concentration_at(1, 1) < 0.01);
}

TEST(releasing_lets_a_former_bath_pixel_move_again) {  // pde.rs:1116-1138
    auto f = ScalarField::create(20, 20, D, 0.0, 1.0);
    f.hold_bath(1.0, [](uint32_t, uint32_t) { return true; });
    for (uint32_t y = 0; y < 20; y++)
        for (uint32_t x = 0; x < 20; x++) f.set_uptake(x, y, 0.01);
    f.step(TimeStep(1.0));
    CHECK(f.concentration_at(10, 10) == 1.0);
    f.release_all();
    f.step(TimeStep(1.0));
    CHECK(f.concentration_at(10, 10) < 1.0);
}

TEST(a_bath_and_a_vessel_can_hold_at_once) {  // pde.rs:1141-1156
    auto f = ScalarField::create(30, 30, D, 0.0, 0.0);
    auto tissue = [](uint32_t x, uint32_t y) { return x >= 5 && x < 25 && y >= 5 && y < 25; };
    f.hold_bath(1.0, [&](uint32_t x, uint32_t y) { return !tissue(x, y); });
    f.hold_pixels(0.5, [](uint32_t x, uint32_t y) { return x == 15 && y == 15; });
    CHECK(f.concentration_at(0, 0) == 1.0);
    CHECK(f.concentration_at(15, 15) == 0.5);
    CHECK(f.concentration_at(10, 10) == 0.0);
}

TEST(a_sweep_is_held_to_the_box_around_what_is_free) {  // pde.rs:1159-1175
    auto f = ScalarField::create(20, 20, D, 0.0, 0.0);
    auto tissue = [](uint32_t x, uint32_t y) { return x >= 4 && x < 9 && y >= 11 && y < 14; };
    f.hold_bath(1.0, [&](uint32_t x, uint32_t y) { return !tissue(x, y); });
    const auto span = f.free_span();
    CHECK(span.has_value());
    CHECK((*span == Span{4, 8, 11, 13, 0, 0}));
    f.hold_pixels(1.0, [](uint32_t, uint32_t) { return true; });
    CHECK(!f.free_span().has_value());
    CHECK(f.relax(TimeStep(1.0), 1e-9, 10) == 1);
}

TEST(the_box_swept_leaves_what_lies_outside_it_as_it_was_held) {  // pde.rs:1178-1200
    auto f = ScalarField::create(40, 40, D, 0.0, 1.0);
    auto disc = [](uint32_t x, uint32_t y) {
        const int dx = (int)x - 20, dy = (int)y - 20;
        return dx * dx + dy * dy <= 64;
    };
    for (uint32_t y = 0; y < 40; y++)
        for (uint32_t x = 0; x < 40; x++)
            if (disc(x, y)) f.set_uptake(x, y, 0.001);
    f.hold_bath(1.0, [&](uint32_t x, uint32_t y) { return !disc(x, y); });
    f.relax(TimeStep(1.0), 1e-9, 100000);
    for (uint32_t y = 0; y < 40; y++)
        for (uint32_t x = 0; x < 40; x++)
            if (!disc(x, y)) CHECK(f.concentration_at(x, y) == 1.0);
    CHECK(f.concentration_at(20, 20) < 1.0);
}

TEST(a_starving_ball_settles_where_the_sphere_formula_puts_it) {  // pde.rs:1203-1229
    const uint32_t side = 31;
    const double r = 10.0, u = 0.006, c = (double)(side / 2);
    auto field = ScalarField::with_depth(side, side, side, D, 0.0, 1.0);
    auto inside = [&](size_t i) {
        const size_t s = side;
        const double x = (double)(i % s), y = (double)((i / s) % s), z = (double)(i / (s * s));
        return (x - c) * (x - c) + (y - c) * (y - c) + (z - c) * (z - c) <= r * r;
    };
    for (size_t i = 0; i < field.len(); i++)
        if (inside(i)) field.set_uptake_at(i, u);
    field.hold_bath_at(1.0, [&](size_t i) { return !inside(i); });
    field.relax(TimeStep(1.0), 1e-10, 1000000);
    const size_t s = side;
    const double centre = field.value_at((s / 2) * (s * s + s + 1));
    CHECK(std::fabs(centre - (1.0 - u * r * r / (6.0 * D))) < 0.05);
}

TEST(a_bath_reaches_a_volume_through_its_six_faces) {  // pde.rs:1232-1248
    auto column = [](size_t i) { return i % 64 != 4 * 8 + 4; };  // open everywhere but one column
    auto vol = ScalarField::with_depth(8, 8, 8, D, 0.0, 0.0);
    vol.hold_bath_at(1.0, column);
    CHECK(vol.value_at(3 * 64 + 4 * 8 + 3) == 1.0);
    CHECK(vol.value_at(3 * 64 + 4 * 8 + 4) == 0.0);
    auto flat = ScalarField::create(8, 8, D, 0.0, 0.0);
    flat.hold_bath_at(1.0, column);
    CHECK(flat.value_at(4 * 8 + 4) == 0.0);
}

TEST(setters_clamp_at_zero_and_ignore_out_of_range) {  // pde.rs:146-182
    auto f = ScalarField::create(5, 5, D, 0.0, 0.0);
    f.set_uptake(1, 1, -3.0);
    f.set_production(2, 2, NAN);
    f.set_uptake(9, 9, 1.0);
    f.set_production_at(10000, 1.0);
    double su = 0, sp = 0;
    for (double v : f.uptake()) su += v;
    for (double v : f.production()) sp += v;
    CHECK(su == 0.0 && sp == 0.0);
    CHECK(f.value_at(1000) == 0.0);
    CHECK(f.concentration_at(7, 1) == 0.0);
}

// ---- coupling.rs ------------------------------------------------------------------

namespace {

// coupling.rs:211-223 in spirit: one cell in the middle of an 80x80 lattice
// ringed by eight others, discs of radius 6 (ids 1..9).
void cluster(Lattice& l) {
    std::vector<uint32_t> g(80 * 80, 0);
    const int cx[9] = {40, 53, 49, 40, 31, 27, 31, 40, 49};
    const int cy[9] = {40, 40, 49, 53, 49, 40, 31, 27, 31};
    for (int c = 0; c < 9; c++)
        for (int y = 0; y < 80; y++)
            for (int x = 0; x < 80; x++)
                if ((x - cx[c]) * (x - cx[c]) + (y - cy[c]) * (y - cy[c]) <= 36)
                    g[(size_t)y * 80 + x] = (uint32_t)(c + 1);
    l.set_grid(g);
    std::vector<uint8_t> living(10, 1);
    living[0] = 0;
    l.set_cells(living);
}

std::vector<double> settled_means(Lattice& l) {
    auto f = ScalarField::create(80, 80, D, 0.0, 1.0);
    apply_uptake(f, l, 0.003);
    f.release_all();
    hold_medium_bath(f, l, 1.0);
    f.relax(TimeStep(1.0), 1e-7, 50000);
    return mean_per_cell(f, l);
}

}  // namespace

TEST(a_cell_in_the_middle_of_a_cluster_reads_less_oxygen_than_one_on_the_rim) {  // :234-254
    Lattice l(80, 80);
    cluster(l);
    const auto o2 = settled_means(l);
    double outer = 0.0;
    for (size_t k = 2; k < o2.size(); k++) outer = std::max(outer, o2[k]);
    CHECK(o2[1] < outer);
    CHECK(o2[1] < 0.9);
    CHECK(outer > o2[1] + 0.05);
}

TEST(a_necrotic_cell_stops_consuming) {  // coupling.rs:257-269
    Lattice l(80, 80);
    cluster(l);
    const double starved = settled_means(l)[1];
    std::vector<uint8_t> living = l.consumes();
    living[1] = 0;  // the middle cell turns necrotic
    l.set_cells(living);
    const double relieved = settled_means(l)[1];
    CHECK(relieved > starved);
}

TEST(the_medium_reads_zero_and_every_cell_reads_a_mean) {  // coupling.rs:272-279
    Lattice l(80, 80);
    cluster(l);
    const auto o2 = settled_means(l);
    CHECK(o2.size() == 10);
    CHECK(o2[0] == 0.0);
    for (size_t k = 1; k < o2.size(); k++) CHECK(o2[k] > 0.0 && o2[k] <= 1.0);
}

TEST(gathers_skip_medium_obstacles_and_unknown_ids) {  // coupling.rs:26-73
    Lattice l(6, 1);
    l.set_grid({0, 1, 2, Lattice::OBSTACLE, 7, 2});
    l.set_cells({0, 1, 0});
    auto f = ScalarField::create(6, 1, D, 0.0, 0.0);
    apply_uptake(f, l, 0.5);
    CHECK((f.uptake() == std::vector<double>{0, 0.5, 0, 0, 0, 0}));
    apply_background_uptake(f, l, 0.25);
    CHECK((f.uptake() == std::vector<double>{0.25, 0.5, 0, 0, 0, 0}));
    apply_production(f, l, {9.0, -1.0, 3.0});
    CHECK((f.production() == std::vector<double>{0, 0, 3.0, 0, 0, 3.0}));
    hold_obstacles_at(f, l, 0.7);
    CHECK(f.value_at(3) == 0.7 && f.held()[3] == 1);
}

TEST(a_vessel_feeds_the_cell_beside_it_and_not_the_one_far_away) {  // coupling.rs:281-302
    // an obstacle disc of radius 4 at (40, 40), cells of radius 5 at (50, 40)
    // (id 1) and (70, 40) (id 2)
    std::vector<uint32_t> g(80 * 80, 0);
    auto disc = [](int x, int y, int cx, int cy, int r) {
        return (x - cx) * (x - cx) + (y - cy) * (y - cy) <= r * r;
    };
    for (int y = 0; y < 80; y++)
        for (int x = 0; x < 80; x++) {
            uint32_t& v = g[(size_t)y * 80 + x];
            if (disc(x, y, 40, 40, 4)) v = Lattice::OBSTACLE;
            else if (disc(x, y, 50, 40, 5)) v = 1;
            else if (disc(x, y, 70, 40, 5)) v = 2;
        }
    Lattice l(80, 80);
    l.set_grid(g);
    l.set_cells({0, 1, 1});
    auto field = ScalarField::create(80, 80, D, 0.0, 0.0);
    apply_uptake(field, l, 0.003);
    hold_obstacles_at(field, l, 1.0);
    field.relax(TimeStep(1.0), 1e-7, 50000);
    const auto o2 = mean_per_cell(field, l);
    CHECK(o2[1] > o2[2] + 0.1);
    CHECK(o2.size() == l.cell_count() + 1);
    // the same through hold_vessels_at with every vessel carrying blood
    // (coupling.rs:84-88: then it is hold_obstacles_at)
    auto again = ScalarField::create(80, 80, D, 0.0, 0.0);
    apply_uptake(again, l, 0.003);
    hold_vessels_at(again, l, 1.0, [](uint32_t, uint32_t) { return true; });
    again.relax(TimeStep(1.0), 1e-7, 50000);
    CHECK(mean_per_cell(again, l) == o2);
    CHECK(mean_in(again, l, 1) == o2[1] && mean_in(again, l, 2) == o2[2]);
}

TEST(hold_vessels_at_holds_only_the_obstacles_that_carry) {  // coupling.rs:89-99
    Lattice l(5, 2);
    const uint32_t O = Lattice::OBSTACLE;
    l.set_grid({O, O, 0, 1, O, 0, O, O, 2, 0});
    l.set_cells({0, 1, 1});
    auto f = ScalarField::create(5, 2, D, 0.0, 0.0);
    hold_vessels_at(f, l, 0.8, [](uint32_t x, uint32_t y) { return x < 2 || y == 1; });
    const std::vector<uint8_t> want = {1, 1, 0, 0, 0, 0, 1, 1, 0, 0};
    CHECK(f.held() == want);
    for (size_t i = 0; i < 10; i++) CHECK(f.value_at(i) == (want[i] ? 0.8 : 0.0));
}

TEST(mean_in_is_the_mean_over_one_id_in_voxel_order) {  // coupling.rs:187-201
    Lattice l(4, 2);
    const uint32_t O = Lattice::OBSTACLE;
    l.set_grid({1, 0, 1, O, 2, 1, 0, 1});
    l.set_cells({0, 1, 1});
    auto f = ScalarField::create(4, 2, D, 0.0, 0.0);
    const double v[8] = {0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8};
    f.hold_where(0.0, [](size_t) { return false; });
    for (size_t i = 0; i < 8; i++) f.hold_where(v[i], [i](size_t j) { return j == i; });
    CHECK(mean_in(f, l, 1) == (((0.1 + 0.3) + 0.6) + 0.8) / 4.0);
    CHECK(mean_in(f, l, 2) == 0.5);
    CHECK(mean_in(f, l, 0) == (0.2 + 0.7) / 2.0);  // the medium too, unlike mean_per_cell
    CHECK(mean_in(f, l, O) == 0.4);
    CHECK(mean_in(f, l, 7) == 0.0);  // an id with no voxel
}

TEST(a_still_medium_sits_below_the_source_by_what_the_tissue_takes_up) {  // simulation.rs:2657-2689
    // The first cell of the volume scenario (simulation.rs:2619-2637): a ball
    // of 123 voxels at the centre of a 40^3 box, each taking up 0.01, in a
    // medium held as a bath at 1; settled as Simulation::start settles it
    // (warm-up tolerance 1e-8, simulation.rs:95-98).
    Lattice l(40, 40, 40);
    std::vector<uint32_t> g(40 * 40 * 40, 0);
    size_t ball = 0;
    for (int z = 0; z < 40; z++)
        for (int y = 0; y < 40; y++)
            for (int x = 0; x < 40; x++)
                if ((x - 20) * (x - 20) + (y - 20) * (y - 20) + (z - 20) * (z - 20) <= 9) {
                    g[((size_t)z * 40 + y) * 40 + x] = 1;
                    ball++;
                }
    CHECK(ball == 123);
    l.set_grid(g);
    l.set_cells({0, 1});
    auto o2 = ScalarField::with_depth(40, 40, 40, D, 0.0, 1.0);
    apply_uptake(o2, l, 0.01);
    o2.release_all();
    hold_medium_bath(o2, l, 1.0);
    o2.relax(TimeStep(1.0), 1e-8, 2000000);
    const double pi = 3.14159265358979323846;
    const double radius = std::cbrt(3.0 * 123.0 / (4.0 * pi));
    const double expected = 1.0 - 0.01 * 123.0 / (4.0 * pi * 0.2 * radius);
    const double level = still_medium_level(o2, 0.2);
    CHECK(std::fabs(level - expected) < 1e-6);
    // a medium oxygen goes through faster is depleted less
    const double faster = still_medium_level(o2, 2.0);
    CHECK(faster > level && faster < 1.0);
    // nothing free: the level is the source's
    auto held = ScalarField::with_depth(4, 4, 4, D, 0.0, 1.0);
    held.hold_where(1.0, [](size_t) { return true; });
    CHECK(still_medium_level(held, 0.2) == 1.0);
}

TEST(host_and_device_copies_stay_in_step) {
    // a setter after a device gather, a gather after a setter, a hold after a step
    Lattice l(16, 16);
    std::vector<uint32_t> g(256, 0);
    for (size_t i = 0; i < 256; i++) g[i] = (i % 3 == 0) ? 1 : 0;
    l.set_grid(g);
    l.set_cells({0, 1});
    auto f = ScalarField::create(16, 16, D, 0.0, 1.0);
    apply_uptake(f, l, 0.01);
    f.set_uptake(1, 0, 0.5);  // voxel 1 is medium: a host change on top of the gather
    CHECK(f.uptake()[0] == 0.01 && f.uptake()[1] == 0.5 && f.uptake()[2] == 0.0);
    f.step(TimeStep(1.0));
    const double after_step = f.value_at(1);
    CHECK(after_step < 1.0);
    f.hold_where(0.25, [](size_t i) { return i == 5; });
    CHECK(f.value_at(5) == 0.25 && f.value_at(1) == after_step);
    apply_background_uptake(f, l, 0.02);  // overwrites the medium voxel's host value
    CHECK(f.uptake()[1] == 0.02);
}

int main(int argc, char** argv) {
    int failed = 0, ran = 0;
    for (const auto& [name, fn] : registry()) {
        if (argc > 1) {
            bool want = false;
            for (int a = 1; a < argc; a++) want |= std::strcmp(argv[a], name) == 0;
            if (!want) continue;
        }
        ran++;
        try {
            fn();
            std::printf("ok   %s\n", name);
        } catch (const Failure& f) {
            failed++;
            std::printf("FAIL %s: %s\n", name, f.what.c_str());
        } catch (const std::exception& e) {
            failed++;
            std::printf("FAIL %s: exception %s\n", name, e.what());
        }
        std::fflush(stdout);
    }
    std::printf("%d of %d passed\n", ran - failed, ran);
    return failed ? 1 : 0;
}
