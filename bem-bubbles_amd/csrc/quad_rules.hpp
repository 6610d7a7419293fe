// Quadrature rules of the collocation path.
//
// The reference selects `quadrature_7` for the far field and `gauss_7` for the
// singular self-panel rule (Simulation.hpp:44-45; tables at quadrature.hpp:30-38
// and :88-96).  Both are published rules, so the constants are necessarily the
// same numbers; here they are written from their closed forms:
//   * 7-point Gauss-Legendre on (0,1): nodes (1 +- t)/2 with t the roots of P_7,
//     weights w/2 (sum 1);
//   * Strang-Fix / Dunavant degree-5 7-point triangle rule on the unit triangle
//     (0,0),(1,0),(0,1): centroid weight 9/80, orbits
//     ((9 +- 2 sqrt 15)/21, (6 -+ sqrt 15)/21) with weights (155 -+ sqrt 15)/2400
//     (sum 1/2, the triangle's area).
#pragma once

namespace bemhip {

// roots of P_7 (positive half) and Gauss weights on (-1, 1)
constexpr double GL7_T1 = 0.40584515137739716690660641;
constexpr double GL7_T2 = 0.74153118559939443986386477;
constexpr double GL7_T3 = 0.94910791234275852452618968;
constexpr double GL7_W0 = 0.41795918367346938775510204;
constexpr double GL7_W1 = 0.38183005050511894495036978;
constexpr double GL7_W2 = 0.27970539148927666790146778;
constexpr double GL7_W3 = 0.12948496616886969327061143;

// nodes / weights on (0,1), ascending
constexpr double GQ_X[7] = {0.5 * (1.0 - GL7_T3), 0.5 * (1.0 - GL7_T2), 0.5 * (1.0 - GL7_T1), 0.5,
                            0.5 * (1.0 + GL7_T1), 0.5 * (1.0 + GL7_T2), 0.5 * (1.0 + GL7_T3)};
constexpr double GQ_W[7] = {0.5 * GL7_W3, 0.5 * GL7_W2, 0.5 * GL7_W1, 0.5 * GL7_W0,
                            0.5 * GL7_W1, 0.5 * GL7_W2, 0.5 * GL7_W3};

// (9 + 2 sqrt15)/21, (6 - sqrt15)/21, (9 - 2 sqrt15)/21, (6 + sqrt15)/21
constexpr double TQ_A1 = 0.79742698535308732239802527;
constexpr double TQ_B1 = 0.10128650732345633880098737;
constexpr double TQ_A2 = 0.05971587178976982045911758;
constexpr double TQ_B2 = 0.47014206410511508977044121;
constexpr double TQ_W0 = 0.1125;                         // 9/80
constexpr double TQ_W1 = 0.06296959027241357629784197;   // (155 - sqrt15)/2400
constexpr double TQ_W2 = 0.06619707639425309036881862;   // (155 + sqrt15)/2400

constexpr double TQ_X[7] = {1.0 / 3.0, TQ_A1, TQ_B1, TQ_B1, TQ_A2, TQ_B2, TQ_B2};
constexpr double TQ_Y[7] = {1.0 / 3.0, TQ_B1, TQ_A1, TQ_B1, TQ_B2, TQ_A2, TQ_B2};
constexpr double TQ_W[7] = {TQ_W0, TQ_W1, TQ_W1, TQ_W1, TQ_W2, TQ_W2, TQ_W2};

// Degree-9 19-point triangle rule (Dunavant 1985) on the same unit triangle,
// as the reference lists it for the exterior potential (quadrature_19,
// quadrature.hpp:126-146; weights there are written w * 0.5): the centroid,
// four orbits (a, b), (b, a), (b, b) and one six-point orbit, in the
// reference's order so that sums run in the same sequence.
constexpr double Q19_C = 0.33333333333333331;
constexpr double Q19_A1 = 2.06349616025259287e-2, Q19_B1 = 0.48968251919873701;
constexpr double Q19_A2 = 0.12582081701412900, Q19_B2 = 0.43708959149293553;
constexpr double Q19_A3 = 0.62359292876193562, Q19_B3 = 0.18820353561903219;
constexpr double Q19_A4 = 0.91054097321109406, Q19_B4 = 4.47295133944529688e-2;
constexpr double Q19_E1 = 0.74119859878449801, Q19_E2 = 3.68384120547362581e-2, Q19_E3 = 0.22196298916076573;
constexpr double Q19_W0 = 9.71357962827961025e-2 * 0.5, Q19_W1 = 3.13347002271398278e-2 * 0.5,
                 Q19_W2 = 7.78275410047754301e-2 * 0.5, Q19_W3 = 7.96477389272090969e-2 * 0.5,
                 Q19_W4 = 2.55776756586981006e-2 * 0.5, Q19_W5 = 4.32835393772893970e-2 * 0.5;
constexpr double Q19_X[19] = {Q19_C,  Q19_A1, Q19_B1, Q19_B1, Q19_A2, Q19_B2, Q19_B2, Q19_A3, Q19_B3, Q19_B3,
                              Q19_A4, Q19_B4, Q19_B4, Q19_E1, Q19_E1, Q19_E2, Q19_E2, Q19_E3, Q19_E3};
constexpr double Q19_Y[19] = {Q19_C,  Q19_B1, Q19_A1, Q19_B1, Q19_B2, Q19_A2, Q19_B2, Q19_B3, Q19_A3, Q19_B3,
                              Q19_B4, Q19_A4, Q19_B4, Q19_E2, Q19_E3, Q19_E1, Q19_E3, Q19_E1, Q19_E2};
constexpr double Q19_W[19] = {Q19_W0, Q19_W1, Q19_W1, Q19_W1, Q19_W2, Q19_W2, Q19_W2, Q19_W3, Q19_W3, Q19_W3,
                              Q19_W4, Q19_W4, Q19_W4, Q19_W5, Q19_W5, Q19_W5, Q19_W5, Q19_W5, Q19_W5};

constexpr double PI = 3.14159265358979323846264338;

}  // namespace bemhip
