// Dormand-Prince 5(4) tableau, Shampine's dense-output midpoint weights and the step-control
// parameters shared by the adaptive orbit kernel and the adaptive response kernel.
#pragma once

namespace ssc {
namespace dp5 {

// Dormand & Prince (1980) tableau.
constexpr double A21 = 1.0 / 5;
constexpr double A31 = 3.0 / 40, A32 = 9.0 / 40;
constexpr double A41 = 44.0 / 45, A42 = -56.0 / 15, A43 = 32.0 / 9;
constexpr double A51 = 19372.0 / 6561, A52 = -25360.0 / 2187, A53 = 64448.0 / 6561,
                 A54 = -212.0 / 729;
constexpr double A61 = 9017.0 / 3168, A62 = -355.0 / 33, A63 = 46732.0 / 5247, A64 = 49.0 / 176,
                 A65 = -5103.0 / 18656;
constexpr double C2 = 1.0 / 5, C3 = 3.0 / 10, C4 = 4.0 / 5, C5 = 8.0 / 9;
constexpr double B1 = 35.0 / 384, B3 = 500.0 / 1113, B4 = 125.0 / 192, B5 = -2187.0 / 6784,
                 B6 = 11.0 / 84;
// b - bhat
constexpr double E1 = 71.0 / 57600, E3 = -71.0 / 16695, E4 = 71.0 / 1920,
                 E5 = -17253.0 / 339200, E6 = 22.0 / 525, E7 = -1.0 / 40;
// Shampine's midpoint weights (y at s = 1/2)
constexpr double M1 = 6025192743.0 / 30085553152.0 / 2, M3 = 51252292925.0 / 65400821598.0 / 2,
                 M4 = -2691868925.0 / 45128329728.0 / 2,
                 M5 = 187940372067.0 / 1594534317056.0 / 2,
                 M6 = -1776094331.0 / 19743644256.0 / 2, M7 = 11237099.0 / 235043384.0 / 2;

struct Dp5Ctl {
  double rtol, atol, dtmin, dtmax, dt0;  // dtmin / dtmax / dt0 <= 0: none / none / heuristic
  int max_steps;
  int n_stops;
};

}  // namespace dp5
}  // namespace ssc
