// Streaming convective residual: the production form of the upwind face loop
// of CSpaceDiscr::Upwind_Convective (source/Upwind_Convective.cpp:5-360), the
// WallFlux override (source/WallFlux.cpp:70-175) and the cell assembly of
// CMesh::ConvectiveFlux (source/ComputeAllBlockVars.cpp:73-119).
//
// A workgroup owns a TJ x TK column tile of one block's (j, k) plane and
// marches along i over a chunk of ci cells.  The tile's cells of planes
// i-1 .. i+2 sit in an LDS ring (the I-face stencils; each plane is read once
// as it enters), and plane i's two-cell halo in j and k beside them (the J and
// K stencils).  Per cell the state is read about 1.5 times (halo and chunk
// overlap) instead of once per face of the four-cell stencils, and every face
// vector, area and node once per face.  Every face is evaluated once, by face_core (shared with the
// per-tile kernel k_convective), and every cell sums its six faces in the same
// order, so the result is bit for bit that of k_convective.
//
// Workgroups are mapped to tiles XCD by XCD: the hardware hands workgroup w to
// XCD w mod 8, and XCD x takes the x-th eighth of the tile list in list order
// (block, i chunk, j, k), so the halos of neighbouring tiles are read through
// the same L2.
#include "fv3d_kernels.h"

namespace fv3d {

// Component c of the stencil variables of cell (i, j, k) of block B, at most one
// index outside the block's cells (then the Dummy layer 1 or 2 of that side).
// MODE 0: primitive rho u v w p (W); 1: conservative (Q); 2: characteristic
// reconstruction, Q (c < 5) then the primitive rho u v w p (c >= 5).
template <int MODE>
__device__ inline double stencil_val(const BlockDev &B, int i, int j, int k, int c)
{
    const bool fromW = MODE == 0 || (MODE == 2 && c >= 5);
    const int cc = (MODE == 2 && c >= 5) ? c - 5 : c;
    if (i >= 1 && i <= B.ni && j >= 1 && j <= B.nj && k >= 1 && k <= B.nk) {
        const long x = cidx(B, i, j, k);
        return fromW ? B.W[cc * B.box + x] : B.Q[cc * B.box + x];
    }
    int s, layer, a, b;
    if (i < 1 || i > B.ni) {
        s = i < 1 ? 0 : 1; layer = i < 1 ? 1 - i : i - B.ni; a = j; b = k;
    } else if (j < 1 || j > B.nj) {
        s = j < 1 ? 2 : 3; layer = j < 1 ? 1 - j : j - B.nj; a = i; b = k;
    } else {
        s = k < 1 ? 4 : 5; layer = k < 1 ? 1 - k : k - B.nk; a = i; b = j;
    }
    return *dummy_ptr(B, s, layer, (fromW ? 5 : 0) + cc, a, b);
}

// Face (i, j, k) of direction d from its stencil (NV variables of the four
// cells), with the wall override on wall faces of the block boundary.
template <int MODE, int NV>
__device__ inline void stream_face(const BlockDev &B, const Gas &g, int d, int i, int j, int k, const double v0[NV],
                                   const double v1[NV], const double v2[NV], const double v3[NV], double out[5])
{
    const long box = B.box;
    const long xf = cidx(B, i, j, k);
    const int N = d == 0 ? B.IM : d == 1 ? B.JM : B.KM;
    const int p = d == 0 ? i : d == 1 ? j : k;
    if (p == 1 || p == N) {
        const int s = 2 * d + (p == N);
        const int ta = d == 0 ? j : i, tb = d == 2 ? j : k;
        const long tix = (long)ta * B.Db[s] + tb;
        if (B.wmask[s] != nullptr && B.wmask[s][tix]) {
            const long xb = xf - (d == 0 ? B.si : d == 1 ? B.sj : 1);
            double wa[5], wb[5];
            for (int c = 0; c < 5; ++c) { wa[c] = B.W[c * box + xf]; wb[c] = B.W[c * box + xb]; }
            wall_face_core(g.gamma, wa, wb, B.fS[d][xf], B.fS[d][box + xf], B.fS[d][2 * box + xf], B.fc[d][xf],
                           B.fc[d][box + xf], B.wrot[s][tix], out);
            return;
        }
    }
    double st[4][5], wF[6] = {0, 0, 0, 0, 0, 0}, wB[6] = {0, 0, 0, 0, 0, 0};
    for (int c = 0; c < 5; ++c) { st[0][c] = v0[c]; st[1][c] = v1[c]; st[2][c] = v2[c]; st[3][c] = v3[c]; }
    if (MODE == 2)
        for (int c = 0; c < 5; ++c) { wF[c] = v1[5 + c]; wB[c] = v2[5 + c]; }
    const long xd = d == 0 ? cidx(B, i, j + 1, k + 1) : d == 1 ? cidx(B, i + 1, j, k + 1) : cidx(B, i + 1, j + 1, k);
    double fcx = 0.0, fcy = 0.0;
    if (B.w_rot != 0.0) { fcx = B.fc[d][xf]; fcy = B.fc[d][box + xf]; }
    face_core(g, st, wF, wB, B.fS[d][xf], B.fS[d][box + xf], B.fS[d][2 * box + xf], B.fs[d][xf], fcx, fcy, B.w_rot,
              B.node[xd] - B.node[xf], B.node[box + xd] - B.node[box + xf], B.node[2 * box + xd] - B.node[2 * box + xf],
              out);
}

template <int TJ, int TK, int MODE>
__global__ void __launch_bounds__(TJ *TK) k_conv_stream(const BlockDev *__restrict__ blocks, const int4 *__restrict__ tiles,
                                                        int ntiles, int ci, Gas g)
{
    constexpr int NV = MODE == 2 ? 10 : 5;
    constexpr int NT = TJ * TK;
    __shared__ double ring[4][NV][TJ][TK];   // tile cells of planes p = i-1 .. i+2 (slot p & 3)
    __shared__ double hJ[NV][4][TK];         // plane i: rows j0-1, j0, j0+TJ+1, j0+TJ+2 over the tile's k
    __shared__ double hK[NV][TJ][4];         // plane i: columns k0-1, k0, k0+TK+1, k0+TK+2 over the tile's j
    __shared__ double FJ[5][TJ + 1][TK];     // J-face fluxes of plane i
    __shared__ double FK[5][TJ][TK + 1];     // K-face fluxes of plane i
    const int nper = (ntiles + 7) >> 3;
    const int tix = (int)(blockIdx.x & 7) * nper + (int)(blockIdx.x >> 3);
    if (tix >= ntiles) return;
    const int4 tl = tiles[tix];
    const BlockDev &B = blocks[tl.x];
    const int t = threadIdx.x, tk = t % TK, tj = t / TK;
    const int j = tl.z + tj + 1, k = tl.w + tk + 1;
    const bool own = j <= B.nj && k <= B.nk;
    // a Dummy cell of the J+ or K+ side inside a partial tile (stencil of the last faces)
    const bool dum = !own && ((j <= B.nj + 2 && k <= B.nk) || (k <= B.nk + 2 && j <= B.nj));
    const int ia = tl.y + 1, ib = min(tl.y + ci, B.ni);
    const long box = B.box;
    const double w_rot = B.w_rot;

    // plane p of the own column into its ring slot (own cells for p in
    // [-1, ni+2], the Dummy cells of a partial tile for p in [1, ni])
    auto load_plane = [&](int p) {
        if (own || (dum && p >= 1 && p <= B.ni)) {
#pragma unroll
            for (int c = 0; c < NV; ++c) ring[p & 3][c][tj][tk] = stencil_val<MODE>(B, p, j, k, c);
        }
    };
    // stencil variable c of plane i at LDS cross position (pj, pk) = offset + 2
    auto cross = [&](int slot, int c, int pj, int pk) -> double {
        if (pj < 2) return hJ[c][pj][pk - 2];
        if (pj >= TJ + 2) return hJ[c][pj - TJ][pk - 2];
        if (pk < 2) return hK[c][pj - 2][pk];
        if (pk >= TK + 2) return hK[c][pj - 2][pk - TK];
        return ring[slot][c][pj - 2][pk - 2];
    };

    double flo[5];
    for (int p = ia - 2; p <= ia + 1; ++p) load_plane(p);
    if (own) {
        double v0[NV], v1[NV], v2[NV], v3[NV];
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            v0[c] = ring[(ia - 2) & 3][c][tj][tk]; v1[c] = ring[(ia - 1) & 3][c][tj][tk];
            v2[c] = ring[ia & 3][c][tj][tk]; v3[c] = ring[(ia + 1) & 3][c][tj][tk];
        }
        stream_face<MODE, NV>(B, g, 0, ia, j, k, v0, v1, v2, v3, flo);
    }
    for (int i = ia; i <= ib; ++i) {
        const int sl = i & 3;
        load_plane(i + 2);
        // halo of plane i: rows first, then columns
        for (int h = t; h < 4 * (TJ + TK); h += NT) {
            int r, jj, kk;
            if (h < 4 * TK) {
                r = h / TK;
                jj = tl.z + (r < 2 ? r : TJ + r) - 1;
                kk = tl.w + h % TK + 1;
            } else {
                const int u = h - 4 * TK;
                r = u & 3;
                jj = tl.z + (u >> 2) + 1;
                kk = tl.w + (r < 2 ? r : TK + r) - 1;
            }
            const bool jin = jj >= 1 && jj <= B.nj, kin = kk >= 1 && kk <= B.nk;
            if (jj >= -1 && jj <= B.nj + 2 && kk >= -1 && kk <= B.nk + 2 && (jin || kin)) {
                if (h < 4 * TK) {
#pragma unroll
                    for (int c = 0; c < NV; ++c) hJ[c][r][h % TK] = stencil_val<MODE>(B, i, jj, kk, c);
                } else {
#pragma unroll
                    for (int c = 0; c < NV; ++c) hK[c][(h - 4 * TK) >> 2][r] = stencil_val<MODE>(B, i, jj, kk, c);
                }
            }
        }
        __syncthreads();
        // J and K faces of plane i, one flat list over the workgroup
        constexpr int nJ = (TJ + 1) * TK, nK = TJ * (TK + 1);
        for (int f = t; f < nJ + nK; f += NT) {
            double fl[5], v0[NV], v1[NV], v2[NV], v3[NV];
            if (f < nJ) {
                const int fj = f / TK, fk = f % TK;
                const int jf = tl.z + fj + 1, kk = tl.w + fk + 1;
                if (jf > B.JM || kk > B.nk) continue;
#pragma unroll
                for (int c = 0; c < NV; ++c) {
                    v0[c] = cross(sl, c, fj, fk + 2); v1[c] = cross(sl, c, fj + 1, fk + 2);
                    v2[c] = cross(sl, c, fj + 2, fk + 2); v3[c] = cross(sl, c, fj + 3, fk + 2);
                }
                stream_face<MODE, NV>(B, g, 1, i, jf, kk, v0, v1, v2, v3, fl);
#pragma unroll
                for (int c = 0; c < 5; ++c) FJ[c][fj][fk] = fl[c];
            } else {
                const int u = f - nJ, fj = u / (TK + 1), fk = u % (TK + 1);
                const int jj = tl.z + fj + 1, kf = tl.w + fk + 1;
                if (jj > B.nj || kf > B.KM) continue;
#pragma unroll
                for (int c = 0; c < NV; ++c) {
                    v0[c] = cross(sl, c, fj + 2, fk); v1[c] = cross(sl, c, fj + 2, fk + 1);
                    v2[c] = cross(sl, c, fj + 2, fk + 2); v3[c] = cross(sl, c, fj + 2, fk + 3);
                }
                stream_face<MODE, NV>(B, g, 2, i, jj, kf, v0, v1, v2, v3, fl);
#pragma unroll
                for (int c = 0; c < 5; ++c) FK[c][fj][fk] = fl[c];
            }
        }
        __syncthreads();
        if (own) {
            double fhi[5];
            {
                double v0[NV], v1[NV], v2[NV], v3[NV];
#pragma unroll
                for (int c = 0; c < NV; ++c) {
                    v0[c] = ring[(i - 1) & 3][c][tj][tk]; v1[c] = ring[sl][c][tj][tk];
                    v2[c] = ring[(i + 1) & 3][c][tj][tk]; v3[c] = ring[(i + 2) & 3][c][tj][tk];
                }
                stream_face<MODE, NV>(B, g, 0, i + 1, j, k, v0, v1, v2, v3, fhi);
            }
            const long x = cidx(B, i, j, k);
            const double vol = B.vol[x];
            double dens, uvel, vvel;
            if (MODE == 0) { dens = ring[sl][0][tj][tk]; uvel = ring[sl][1][tj][tk]; vvel = ring[sl][2][tj][tk]; }
            else if (MODE == 2) { dens = ring[sl][5][tj][tk]; uvel = ring[sl][6][tj][tk]; vvel = ring[sl][7][tj][tk]; }
            else { dens = B.W[x]; uvel = B.W[box + x]; vvel = B.W[2 * box + x]; }
            double Qc[5] = {0.0, dens * (w_rot * vvel) * vol, F32C(-1.0) * dens * (w_rot * uvel) * vol, 0.0, 0.0};
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                double s = fhi[c] - flo[c];
                s = s + FJ[c][tj + 1][tk];
                s = s - FJ[c][tj][tk];
                s = s + FK[c][tj][tk + 1];
                s = s - FK[c][tj][tk];
                s = s - Qc[c];
                B.R[c * box + x] = (s - 0.0) - 0.0;
                flo[c] = fhi[c];
            }
        }
    }
}

#define FV3D_STREAM(S, TJ, TK)                                                                            \
    template __global__ void k_conv_stream<TJ, TK, 0>(const BlockDev *, const int4 *, int, int, Gas);     \
    template __global__ void k_conv_stream<TJ, TK, 1>(const BlockDev *, const int4 *, int, int, Gas);     \
    template __global__ void k_conv_stream<TJ, TK, 2>(const BlockDev *, const int4 *, int, int, Gas);
FV3D_FOR_STREAM_TILES(FV3D_STREAM)
#undef FV3D_STREAM

}  // namespace fv3d
