// Plane sweep-tile shapes shared by the solve (cwfield.hip) and the
// device-driven batch (cw_campaign.hip); a shape is an index into one table.
#pragma once

namespace cwi {

// Sweep-tile shapes: s sweeps per launch over regions 128 x RY cells with NW
// waves; the interior is (128 - 4s) x (RY - 4s). kind 0: LDS tile kernel;
// 1: register-blocked by column parity; 2: register-blocked by colour
// (relax2d_reg, NR = RY / NW rows per wave). Index 0 of each table is the
// one-sweep shape used for remainders and exact-stop replays.
struct Shape { int s, ry, nw, kind; };
constexpr Shape kPlainShapes[] = {{1, 32, 4, 0}, {2, 32, 4, 0},  {1, 32, 4, 2}, {2, 32, 4, 2},
                                  {2, 64, 8, 2}, {2, 64, 16, 2}, {1, 32, 8, 2}, {4, 32, 4, 2},
                                  {4, 64, 8, 2}, {8, 64, 8, 2}};
constexpr Shape kSmoothShapes[] = {{1, 32, 4, 0},   {2, 32, 4, 0},   {4, 64, 8, 0},   {4, 64, 4, 0},
                                   {8, 128, 8, 0},  {8, 128, 16, 0}, {2, 32, 4, 1},   {4, 64, 16, 1},
                                   {4, 64, 8, 1},   {8, 128, 16, 1}, {2, 64, 8, 1},   {4, 128, 16, 1},
                                   {2, 32, 4, 2},   {4, 64, 8, 2},   {4, 64, 16, 2},  {8, 128, 16, 2},
                                   {2, 64, 8, 2},   {4, 128, 16, 2}, {6, 128, 16, 2},
                                   {10, 128, 16, 2}, {12, 128, 16, 2}, {6, 64, 8, 2},
                                   {8, 64, 8, 2},   {12, 64, 8, 2},  {16, 128, 16, 2},
                                   {8, 160, 16, 2}, {10, 160, 16, 2}, {12, 160, 16, 2}};
constexpr int kNumPlainShapes = (int)(sizeof(kPlainShapes) / sizeof(kPlainShapes[0]));
constexpr int kNumSmoothShapes = (int)(sizeof(kSmoothShapes) / sizeof(kSmoothShapes[0]));

}  // namespace cwi
