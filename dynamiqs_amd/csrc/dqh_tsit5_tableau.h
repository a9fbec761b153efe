// Tableaus of the 5(4) pairs and step clipping. Tsit5: Ch. Tsitouras, Comput. Math.
// Appl. 62 (2011) 770-775. Published constants of the methods; each translation
// unit gets its own __constant__ copy.
#pragma once

#include <hip/hip_runtime.h>

#include "dqh.h"

namespace dqh {

// 7-stage FSAL pairs of order 5(4): stage times C, stage coefficients A (the last row holds the
// solution weights b, b_7 = 0), error weights E = b - b_hat and the continuous
// extension b_i(theta) = sum_p R[i][p] theta^(p+1).
struct RKTab {
  double C[7];  // stage times t + C_s h
  double A[7][7];
  double E[7];
  double R[7][4];
};

// [0] Tsit5. [1] Dopri5: J. R. Dormand, P. J. Prince, J. Comput. Appl. Math. 6 (1980)
// 19-26. Its error weights are b - b_hat with Shampine's fourth-order b_hat
// (1951/21600, 0, 22642/50085, 451/720, -12231/42400, 649/6300, 1/60; Math. Comp.
// 46 (1986) 135-150), not the 1980 paper's (5179/57600, ..., 1/40). The dense
// output is the quartic through y0, y1, f0, f1 and Shampine's midpoint value, in
// the same b_i(theta) form. Its theta^4 column equals the dense-output constants
// of the DOPRI5 code.
static __constant__ RKTab kRK[2] = {
    {{0, 0.161, 0.327, 0.9, 0.9800255409045097, 1, 1},
     {{0, 0, 0, 0, 0, 0, 0},
      {0.161, 0, 0, 0, 0, 0, 0},
      {-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0, 0},
      {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0, 0},
      {5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525, 0, 0, 0},
      {5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401, -0.028269050394068383, 0, 0},
      {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081, 2.324710524099774, 0}},
     {-0.00178001105222577714, -0.0008164344596567469, 0.007880878010261995, -0.1447110071732629,
      0.5823571654525552, -0.45808210592918697, 0.015151515151515152},
     {{1.0, -2.763706197274826, 2.9132554618219126, -1.0530884977290216},
      {0.0, 0.13169999999999998, -0.2234, 0.1017},
      {0.0, 3.9302962368947516, -5.941033872131505, 2.490627285651253},
      {0.0, -12.411077166933676, 30.33818863028232, -16.548102889244902},
      {0.0, 37.50931341651104, -88.1789048947664, 47.37952196281928},
      {0.0, -27.896526289197286, 65.09189467479366, -34.87065786149661},
      {0.0, 1.5, -4.0, 2.5}}},
    {{0, 0.2, 0.3, 0.8, 0.8888888888888888, 1, 1},
     {{0, 0, 0, 0, 0, 0, 0},
      {0.2, 0, 0, 0, 0, 0, 0},
      {0.075, 0.225, 0, 0, 0, 0, 0},
      {0.9777777777777777, -3.7333333333333334, 3.5555555555555554, 0, 0, 0, 0},
      {2.9525986892242035, -11.595793324188385, 9.822892851699436, -0.2908093278463649, 0, 0, 0},
      {2.8462752525252526, -10.757575757575758, 8.906422717743473, 0.2784090909090909, -0.2735313036020583, 0, 0},
      {0.09114583333333333, 0, 0.44923629829290207, 0.6510416666666666, -0.322376179245283, 0.13095238095238096, 0}},
     {0.0008217592592592593, 0, -0.0028351801936707596, 0.024652777777777777, -0.03390919811320755,
      0.027936507936507936, -0.016666666666666666},
     {{1.0, -2.8535800653862835, 3.0717434641059005, -1.1270175653862835},
      {0, 0, 0, 0},
      {0, 4.023133379230305, -6.249321565289, 2.675424484351598},
      {0, -3.7324019615885042, 10.068970589843675, -5.685526961588504},
      {0, 2.5548038301849423, -6.399112377351017, 3.5219323679207912},
      {0, -1.3744241142186024, 3.272657752246729, -1.7672812570757455},
      {0, 1.3824689317781436, -3.764937863556287, 2.382468931778144}}},
};

// Initial step h1 = (0.01 / max(d1, d2))^kInitExp (Hairer-Norsett-Wanner II.4). The
// exponent is one over the order of the embedded solution (4), while the
// controller uses one over the order of the pair (5). This is the exponent with
// which both pairs take exactly the step counts the reference prints for its
// examples (README.md:75, apis/mesolve.py:225, apis/sesolve.py:164,
// apis/floquet.py:139, docs/documentation/basics/workflow.md:88). With 1/5,
// every Tsit5 example takes one step fewer.
constexpr double kInitExp = 0.25;

// the pair of a descriptor: Dopri5 with DQH_FLAG_DOPRI5, Tsit5 otherwise
__device__ __forceinline__ const RKTab &rk_tab(int64_t flags) { return kRK[(flags & DQH_FLAG_DOPRI5) ? 1 : 0]; }

// A step that would stop within kSnap * max(1, |target|) short of its target (the
// end time or the next discontinuity) ends on the target instead: otherwise a
// step landing a few ulps before a breakpoint leaves a sliver step to it, far
// below any meaningful step size (and below the vanishing-step guard).
constexpr double kSnap = 1e-12;
__device__ __forceinline__ bool snaps_to(double t, double h, double target) {
  return t + h >= target - kSnap * fmax(1.0, fabs(target));
}

// Increment form of the stages (generator constant over the step): stages 1..5
// hold g_s = f(Delta_s), Delta_s = h sum_j a_sj k_j, and k_s = k_0 + g_s, so the
// weights w_0..w_6 of the true stages k_0..k_6 become weights of
// (k_0, g_1..g_5, k_6) with w_0 <- w_0 + ... + w_5.
__device__ __forceinline__ void fold_k0(double (&w)[7]) {
  double a = 0.0;
#pragma unroll
  for (int j = 0; j < 6; ++j) a += w[j];
  w[0] = a;
}

}  // namespace dqh
