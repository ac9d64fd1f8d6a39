// Face tiles of the viscous and Spalart-Allmaras kernels and the face
// gradient rule they share.
#pragma once
#include "fv3d_kernels.h"

namespace fv3d {

// ------------------------------------------------------------ viscous flux
// A tile of TI x TJ x TK cells stages u v w T of its cells and a one-cell halo
// in LDS, evaluates every face of the tile (lower faces of its cells and the
// upper faces on its far sides) once, keeps the face fluxes in LDS and sums
// them per cell.
template <int TI, int TJ, int TK>
struct VTile {
    static constexpr int HI = TI + 2, HJ = TJ + 2, HK = TK + 2, NH = HI * HJ * HK;
    static constexpr int NFI = (TI + 1) * TJ * TK, NFJ = TI * (TJ + 1) * TK, NFK = TI * TJ * (TK + 1);
    static constexpr int NF = NFI + NFJ + NFK;
    __device__ static int h(int a, int b, int c) { return (a * HJ + b) * HK + c; }
};

// Gradients of u v w T at the face of direction D whose higher-index cell is
// (i, j, k), halo position l (source/Spectral_curl_grad.cpp:85-493).  Along D
// the difference of the two cells; along each other direction t the average of
// the cells x, x+t, x+t-D, x-D minus that of x, x-f, x-f-s, x-s (f, s the
// lower and higher of D and t), summed in that order.  A cell outside the
// block along both D and t (a never-filled ghost-box corner) is left out and
// the other three averaged; at the J face j == JM, k == 1 the reference also
// scales the lower three-cell sum by 0.25 (line 331).
template <int D, class TL, int NV = 4>
__device__ inline void face_dv(const double *sw, int l, int i, int j, int k, const BlockDev &B, double dv[3][NV])
{
    constexpr int NH = TL::NH;
    const int st[3] = {TL::HJ * TL::HK, TL::HK, 1};
    const int pos[3] = {i, j, k};
    const int lim[3] = {B.ni, B.nj, B.nk};
    const int lD = l - st[D];
#pragma unroll
    for (int q = 0; q < NV; ++q) dv[D][q] = sw[q * NH + l] - sw[q * NH + lD];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        if (t == D) continue;
        const int f = D < t ? D : t, s = D < t ? t : D;
        // offsets (along D, along t) of the four cells of each average
        const int pD[4] = {0, 0, -1, -1}, pT[4] = {0, 1, 1, 0};
        const int mf[4] = {0, -1, -1, 0}, ms[4] = {0, 0, -1, -1};
        double avg[2][NV];
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            double sum[NV];
#pragma unroll
            for (int q = 0; q < NV; ++q) sum[q] = 0.0;
            int n = 0;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                int oD, oT;
                if (hh == 0) {
                    oD = pD[m];
                    oT = pT[m];
                } else {
                    oD = f == D ? mf[m] : ms[m];
                    oT = f == D ? ms[m] : mf[m];
                }
                const int yD = pos[D] + oD, yT = pos[t] + oT;
                const bool outD = yD < 1 || yD > lim[D], outT = yT < 1 || yT > lim[t];
                const bool keep = !(outD && outT);
                const int ly = l + oD * st[D] + oT * st[t];
#pragma unroll
                for (int q = 0; q < NV; ++q) {
                    const double v = sw[q * NH + ly];
                    sum[q] = keep ? (m == 0 ? v : sum[q] + v) : sum[q];
                }
                n += keep ? 1 : 0;
            }
            const bool quirk = hh == 1 && D == 1 && t == 2 && j == B.JM && k == 1;
#pragma unroll
            for (int q = 0; q < NV; ++q)
                avg[hh][q] = n == 4 ? sum[q] / F32C(4.0)
                                    : (quirk ? F32C(0.25) * sum[q] / F32C(3.0) : sum[q] / F32C(3.0));
        }
#pragma unroll
        for (int q = 0; q < NV; ++q) dv[t][q] = avg[0][q] - avg[1][q];
    }
}


}  // namespace fv3d
