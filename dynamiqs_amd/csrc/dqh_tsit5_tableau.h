// Tsit5 tableau and step clipping. Tableau (Ch. Tsitouras, Comput. Math. Appl. 62 (2011) 770-775): stage
// coefficients kA (the last row holds the solution weights b, FSAL), error
// weights kE = b - b_hat and the continuous extension kR. Published constants
// of the method; each translation unit gets its own __constant__ copy.
#pragma once

#include <hip/hip_runtime.h>

namespace dqh {

static __constant__ double kA[7][7] = {
    {0, 0, 0, 0, 0, 0, 0},
    {0.161, 0, 0, 0, 0, 0, 0},
    {-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0, 0},
    {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0, 0},
    {5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525, 0, 0, 0},
    {5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401, -0.028269050394068383, 0, 0},
    {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081, 2.324710524099774, 0},
};
static __constant__ double kE[7] = {-0.00178001105222577714, -0.0008164344596567469, 0.007880878010261995,
                             -0.1447110071732629,     0.5823571654525552,     -0.45808210592918697,
                             0.015151515151515152};
// continuous extension: b_i(theta) = sum_p kR[i][p] theta^(p+1)
static __constant__ double kR[7][4] = {
    {1.0, -2.763706197274826, 2.9132554618219126, -1.0530884977290216},
    {0.0, 0.13169999999999998, -0.2234, 0.1017},
    {0.0, 3.9302962368947516, -5.941033872131505, 2.490627285651253},
    {0.0, -12.411077166933676, 30.33818863028232, -16.548102889244902},
    {0.0, 37.50931341651104, -88.1789048947664, 47.37952196281928},
    {0.0, -27.896526289197286, 65.09189467479366, -34.87065786149661},
    {0.0, 1.5, -4.0, 2.5},
};

// A step that would stop within kSnap * max(1, |target|) short of its target (the
// end time or the next discontinuity) ends on the target instead: otherwise a
// step landing a few ulps before a breakpoint leaves a sliver step to it, far
// below any meaningful step size (and below the vanishing-step guard).
constexpr double kSnap = 1e-12;
__device__ __forceinline__ bool snaps_to(double t, double h, double target) {
  return t + h >= target - kSnap * fmax(1.0, fabs(target));
}

}  // namespace dqh
