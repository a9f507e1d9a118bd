// Interior leapfrog updates of the velocity-stress scheme on gfx950.
//
// One launch per half step, as in the reference time loop
// (opesci/templates/staggered3d_tmpl.py:40-58): the stress kernel writes the
// six stresses at t1 from the velocities at t0 (opesci/staggeredgrid.py:
// 728-737), the velocity kernel writes U, V, W at t1 from the stresses at t1
// (:739-748).  Interior = [m, dim-m) on every axis, the same bounds the
// reference loops use for staggered and unstaggered fields.
//
// Stencil (coefficients folded on the host, see opesci/plan.py):
//   T_ii += sum_j m_ij sum_k cn[i][j][k] (V_j[+k] - V_j[-1-k])          along j
//   T_ij += m_s (sum_k cs[s][0][k] (V_i[+1+k] - V_i[-k]) along j + ... along i)
//   V_i  += m_i sum_j sum_k cv[i][j][k] (T[a+k] - T[a-1-k])             along j
// The velocity sum is one fma chain over j = x, y, z (the order that keeps
// FP32 closest to FP64 on the 512^3 8th-order sweep; y, z, x measured 20-40 %
// farther): the fused step keeps the y and z differences of an x plane while
// its stresses are staged and sums the chain M planes later.
// with m = 1 for constant media, and lambda(+2mu) / mu_ij / beta_i arrays in
// the read path.  Arithmetic in the compute type R (float or double); fields
// stored as S (float, double, _Float16, bf16).
//
// Three implementations (kernel_variant):
//  * 0, pair sweep (default for FP32 arithmetic, any storage): the sweep below
//    with two z points per lane and 128(z) x TY(y) tiles; FP64 uses 2.
//  * 2, sweep: a block owns a 64(z) x TY(y) column tile and sweeps x.
//    Fields differentiated along x live in per-thread register queues, the
//    current x plane of fields differentiated along y or z is staged with its
//    halo in LDS, so every field is read from HBM once per sweep.  Blocks are
//    remapped so that the tiles of one XCD form a contiguous slab and share
//    halo lines in that XCD's L2.  The cell's own previous value and every
//    store are non-temporal (streamed once); only the stencil operands, whose
//    halo lines neighbouring tiles re-read, are ordinary cached loads.
//  * 1, gather: one thread per (y,z) column, every stencil operand loaded from
//    global memory (caches only); kept as the simple reference form.
// All three keep one accumulation order, so their results agree bit for bit.
#include "common.h"

#include <algorithm>

namespace opesci {

template <typename R>
__device__ __forceinline__ R fmac(R a, R b, R c) { return fma(a, b, c); }

// XCD-aware block remap: hardware dispatches linear block b to XCD b % 8.
// Give XCD k the logical tiles [k*per, (k+1)*per) so one XCD's blocks are
// spatial neighbours.  Returns -1 for padding blocks.
__device__ __forceinline__ int logical_block(int nblocks) {
    const int b = blockIdx.x;
    const int per = (nblocks + 7) / 8;
    const int l = (b % 8) * per + b / 8;
    return l < nblocks ? l : -1;
}

// Origin of logical block l of a box list: z tile start, first row, row
// end of its box, x range.  The box is picked with constant indices (the
// list stays in kernel arguments).
struct TileOrigin {
    int z0, y0, yend, xb, xe;
};
template <int TY>
__device__ __forceinline__ TileOrigin tile_of(const Boxes& B, int l, int tz, int M) {
    int first = B.first[0], x0 = B.x0[0], x1 = B.x1[0], y0 = B.y0[0], y1 = B.y1[0], tyn = B.tyn[0], xc = B.xc[0];
#pragma unroll
    for (int i = 1; i < 4; ++i)
        if (i < B.n && l >= B.first[i]) {
            first = B.first[i];
            x0 = B.x0[i];
            x1 = B.x1[i];
            y0 = B.y0[i];
            y1 = B.y1[i];
            tyn = B.tyn[i];
            xc = B.xc[i];
        }
    const int loc = l - first;
    const int bz = loc % tz, by = (loc / tz) % tyn;
    const int bx = loc / (tz * tyn);
    TileOrigin o;
    o.z0 = M + bz * 128;
    o.y0 = y0 + by * TY;
    o.yend = y1;
    o.xb = x0 + bx * xc;
    o.xe = min(o.xb + xc, x1);
    return o;
}

// ------------------------------------------------------------------ gather
template <typename R, typename S, int M, bool HET>
__global__ __launch_bounds__(256) void stress_gather(Geom g, FieldPtrs<S> F, MediaPtrs<R> md, Coefs<R> c,
                                                     int t0, int t1, int tz, int ty, int xchunk, int nblocks) {
    const int l = logical_block(nblocks);
    if (l < 0) return;
    const int bz = l % tz, by = (l / tz) % ty, bx = l / (tz * ty);
    const int z = M + bz * 64 + (int)(threadIdx.x & 63);
    const int y = M + by * 4 + (int)(threadIdx.x >> 6);
    if (z >= g.nz - M || y >= g.ny - M) return;
    const int xb = M + bx * xchunk;
    const int xe = min(xb + xchunk, g.nx - M);
    using C = Conv<S, R>;
    const int64_t st[3] = {g.sx, g.sy, 1};
    for (int x = xb; x < xe; ++x) {
        const int64_t q = g.at(x, y, z);
        R lam = 0, l2m = 0;
        if (HET) {
            lam = md.m[OPESCI_M_LAMBDA][q];
            l2m = lam + (R)2 * md.m[OPESCI_M_MU][q];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            R acc = C::ld(F.f[3 + i][t0] + q);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const S* v = F.f[j][t0] + q;
                R s = 0;
#pragma unroll
                for (int k = 0; k < M; ++k)
                    s = fmac(c.cn[i][j][k], C::ld(v + k * st[j]) - C::ld(v - (k + 1) * st[j]), s);
                acc = HET ? fmac(j == i ? l2m : lam, s, acc) : acc + s;
            }
            C::st(F.f[3 + i][t1] + q, acc);
        }
        const int pr[3][2] = {{0, 1}, {1, 2}, {0, 2}};
        const int mid[3] = {OPESCI_M_MU12, OPESCI_M_MU23, OPESCI_M_MU13};
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const int ia = pr[s][0], ja = pr[s][1];
            R acc = C::ld(F.f[6 + s][t0] + q);
            R sum = 0;
#pragma unroll
            for (int gq = 0; gq < 2; ++gq) {
                const int src = gq == 0 ? ia : ja, ax = gq == 0 ? ja : ia;
                const S* v = F.f[src][t0] + q;
#pragma unroll
                for (int k = 0; k < M; ++k)
                    sum = fmac(c.cs[s][gq][k], C::ld(v + (1 + k) * st[ax]) - C::ld(v - k * st[ax]), sum);
            }
            acc = HET ? fmac(md.m[mid[s]][q], sum, acc) : acc + sum;
            C::st(F.f[6 + s][t1] + q, acc);
        }
    }
}

template <typename R, typename S, int M, bool HET>
__global__ __launch_bounds__(256) void velocity_gather(Geom g, FieldPtrs<S> F, MediaPtrs<R> md, Coefs<R> c,
                                                       int t0, int t1, int tz, int ty, int xchunk, int nblocks) {
    const int l = logical_block(nblocks);
    if (l < 0) return;
    const int bz = l % tz, by = (l / tz) % ty, bx = l / (tz * ty);
    const int z = M + bz * 64 + (int)(threadIdx.x & 63);
    const int y = M + by * 4 + (int)(threadIdx.x >> 6);
    if (z >= g.nz - M || y >= g.ny - M) return;
    const int xb = M + bx * xchunk;
    const int xe = min(xb + xchunk, g.nx - M);
    using C = Conv<S, R>;
    const int64_t st[3] = {g.sx, g.sy, 1};
    // source stress of velocity i along axis j
    const int src[3][3] = {{3, 6, 8}, {6, 4, 7}, {8, 7, 5}};
    for (int x = xb; x < xe; ++x) {
        const int64_t q = g.at(x, y, z);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            R acc = C::ld(F.f[i][t0] + q);
            R sum = 0;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int a = (i == j) ? 1 : 0;
                const S* t = F.f[src[i][j]][t1] + q;
#pragma unroll
                for (int k = 0; k < M; ++k)
                    sum = fmac(c.cv[i][j][k], C::ld(t + (a + k) * st[j]) - C::ld(t + (a - 1 - k) * st[j]), sum);
            }
            acc = HET ? fmac(md.m[OPESCI_M_BETA1 + i][q], sum, acc) : acc + sum;
            C::st(F.f[i][t1] + q, acc);
        }
    }
}

// ------------------------------------------------------------------ sweep
// LDS plane of a 64 x TY tile with an M-cell halo.  Halo cells: the y-halo
// (M rows above and below, tile columns only) and the z-halo (M columns left
// and right, tile rows only); corners are never read.
template <int M, int TY>
struct Plane {
    static constexpr int W = 64 + 2 * M;
    static constexpr int H = TY + 2 * M;
    static constexpr int NT = 64 * TY;
    static constexpr int NY = 2 * M * 64;   // y-halo cells
    static constexpr int NZ = TY * 2 * M;   // z-halo cells
    static constexpr int JY = (NY + NT - 1) / NT;
    static_assert(NZ <= NT, "z halo larger than the block");
    __device__ __forceinline__ static void yhalo(int h, int& row, int& col) {
        const int r = h / 64;
        row = r < M ? r : r + TY;
        col = M + h % 64;
    }
    __device__ __forceinline__ static void zhalo(int h, int& row, int& col) {
        row = M + h / (2 * M);
        const int c = h % (2 * M);
        col = c < M ? c : 64 + c;
    }
};

// Per-thread halo slots: LDS offset and in-plane global offset of every halo
// cell this thread stages (-1 = none).
template <int M, int TY>
struct HaloSlots {
    using P = Plane<M, TY>;
    int ly[P::JY], gy[P::JY];
    int lz, gz;
    __device__ void init(const Geom& g, int y0, int z0) {
        const int t = threadIdx.x;
#pragma unroll
        for (int j = 0; j < P::JY; ++j) {
            const int h = t + j * P::NT;
            if (h < P::NY) {
                int row, cl;
                P::yhalo(h, row, cl);
                const int yy = min(max(y0 - M + row, 0), g.ny - 1), zz = min(max(z0 - M + cl, 0), g.nz - 1);
                ly[j] = row * P::W + cl;
                gy[j] = yy * (int)g.sy + zz;
            } else {
                ly[j] = -1;
                gy[j] = 0;
            }
        }
        if (t < P::NZ) {
            int row, cl;
            P::zhalo(t, row, cl);
            const int yy = min(max(y0 - M + row, 0), g.ny - 1), zz = min(max(z0 - M + cl, 0), g.nz - 1);
            lz = row * P::W + cl;
            gz = yy * (int)g.sy + zz;
        } else {
            lz = -1;
            gz = 0;
        }
    }
};

// Values staged for one plane: y-halo and z-halo registers.
template <typename R, int M, int TY, bool HY, bool HZ>
struct Staged {
    using P = Plane<M, TY>;
    R y[HY ? P::JY : 1];
    R z;
    template <typename S>
    __device__ __forceinline__ void fetch(const S* base, int64_t px, const HaloSlots<M, TY>& hs) {
        using C = Conv<S, R>;
        if (HY) {
#pragma unroll
            for (int j = 0; j < P::JY; ++j)
                if (hs.ly[j] >= 0) y[j] = C::ld(base + px + hs.gy[j]);
        }
        if (HZ && hs.lz >= 0) z = C::ld(base + px + hs.gz);
    }
    __device__ __forceinline__ void put(R* lds, const HaloSlots<M, TY>& hs) const {
        if (HY) {
#pragma unroll
            for (int j = 0; j < P::JY; ++j)
                if (hs.ly[j] >= 0) lds[hs.ly[j]] = y[j];
        }
        if (HZ && hs.lz >= 0) lds[hs.lz] = z;
    }
};

template <typename R, typename S, int M, int TY, bool HET>
__global__ __launch_bounds__(64 * TY) void stress_sweep(Geom g, FieldPtrs<S> F, MediaPtrs<R> md, Coefs<R> c,
                                                        int t0, int t1, int tz, int ty, int xchunk, int nblocks) {
    using P = Plane<M, TY>;
    using C = Conv<S, R>;
    __shared__ R sU[2][P::H * P::W], sV[2][P::H * P::W], sW[2][P::H * P::W];
    const int l = logical_block(nblocks);
    if (l < 0) return;
    const int bz = l % tz, by = (l / tz) % ty, bx = l / (tz * ty);
    const int lz = threadIdx.x & 63, ly = threadIdx.x >> 6;
    const int z0 = M + bz * 64, y0 = M + by * TY;
    const int z = z0 + lz, y = y0 + ly;
    const bool active = z < g.nz - M && y < g.ny - M;
    const int64_t col = (int64_t)min(y, g.ny - 1) * g.sy + min(z, g.nz - 1);
    const int xb = M + bx * xchunk, xe = min(xb + xchunk, g.nx - M);
    const S* U0 = F.f[0][t0];
    const S* V0 = F.f[1][t0];
    const S* W0 = F.f[2][t0];
    HaloSlots<M, TY> hs;
    hs.init(g, y0, z0);
    // register queues along x: qU[i] = U(x+i-M), qV[i] = V(x+i-M+1), qW[i] = W(x+i-M+1)
    R qU[2 * M], qV[2 * M], qW[2 * M];
#pragma unroll
    for (int i = 1; i < 2 * M; ++i) {
        qU[i] = C::ld(U0 + (int64_t)(xb - M - 1 + i) * g.sx + col);
        qV[i] = C::ld(V0 + (int64_t)(xb - M + i) * g.sx + col);
        qW[i] = C::ld(W0 + (int64_t)(xb - M + i) * g.sx + col);
    }
    // values of the next plane, fetched one iteration ahead
    R nU, nV, nW, nT[6];
    Staged<R, M, TY, true, true> nhU, nhV, nhW;
    auto fetch = [&](int x) {
        const int64_t px = (int64_t)x * g.sx;
        nU = C::ld(U0 + px + (int64_t)(M - 1) * g.sx + col);
        nV = C::ld(V0 + px + (int64_t)M * g.sx + col);
        nW = C::ld(W0 + px + (int64_t)M * g.sx + col);
#pragma unroll
        for (int f = 0; f < 6; ++f) nT[f] = C::ldnt(F.f[3 + f][t0] + px + col);
        nhU.fetch(U0, px, hs);
        nhV.fetch(V0, px, hs);
        nhW.fetch(W0, px, hs);
    };
    fetch(xb);
    for (int x = xb; x < xe; ++x) {
        const int b = (x - xb) & 1;
#pragma unroll
        for (int i = 0; i < 2 * M - 1; ++i) {
            qU[i] = qU[i + 1];
            qV[i] = qV[i + 1];
            qW[i] = qW[i + 1];
        }
        qU[2 * M - 1] = nU;
        qV[2 * M - 1] = nV;
        qW[2 * M - 1] = nW;
        R T0[6];
#pragma unroll
        for (int f = 0; f < 6; ++f) T0[f] = nT[f];
        R* LU = sU[b];
        R* LV = sV[b];
        R* LW = sW[b];
        const int me = (ly + M) * P::W + lz + M;
        LU[me] = qU[M];
        LV[me] = qV[M - 1];
        LW[me] = qW[M - 1];
        nhU.put(LU, hs);
        nhV.put(LV, hs);
        nhW.put(LW, hs);
        if (x + 1 < xe) fetch(x + 1);
        __syncthreads();
        if (active) {
            const int64_t q = (int64_t)x * g.sx + col;
            R dx[M], dy[M], dz[M];
#pragma unroll
            for (int k = 0; k < M; ++k) {
                dx[k] = qU[M + k] - qU[M - 1 - k];
                dy[k] = LV[me + k * P::W] - LV[me - (1 + k) * P::W];
                dz[k] = LW[me + k] - LW[me - 1 - k];
            }
            R lam = 0, l2m = 0;
            if (HET) {
                lam = __builtin_nontemporal_load(md.m[OPESCI_M_LAMBDA] + q);
                l2m = lam + (R)2 * __builtin_nontemporal_load(md.m[OPESCI_M_MU] + q);
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                R acc = T0[i];
                R s = 0;
#pragma unroll
                for (int k = 0; k < M; ++k) s = fmac(c.cn[i][0][k], dx[k], s);
                acc = HET ? fmac(i == 0 ? l2m : lam, s, acc) : acc + s;
                s = 0;
#pragma unroll
                for (int k = 0; k < M; ++k) s = fmac(c.cn[i][1][k], dy[k], s);
                acc = HET ? fmac(i == 1 ? l2m : lam, s, acc) : acc + s;
                s = 0;
#pragma unroll
                for (int k = 0; k < M; ++k) s = fmac(c.cn[i][2][k], dz[k], s);
                acc = HET ? fmac(i == 2 ? l2m : lam, s, acc) : acc + s;
                C::stnt(F.f[3 + i][t1] + q, acc);
            }
            // Txy: dU/dy then dV/dx;  Tyz: dV/dz then dW/dy;  Txz: dU/dz then dW/dx
            R sxy = 0, syz = 0, sxz = 0;
#pragma unroll
            for (int k = 0; k < M; ++k) sxy = fmac(c.cs[0][0][k], LU[me + (1 + k) * P::W] - LU[me - k * P::W], sxy);
#pragma unroll
            for (int k = 0; k < M; ++k) sxy = fmac(c.cs[0][1][k], qV[M + k] - qV[M - 1 - k], sxy);
#pragma unroll
            for (int k = 0; k < M; ++k) syz = fmac(c.cs[1][0][k], LV[me + 1 + k] - LV[me - k], syz);
#pragma unroll
            for (int k = 0; k < M; ++k) syz = fmac(c.cs[1][1][k], LW[me + (1 + k) * P::W] - LW[me - k * P::W], syz);
#pragma unroll
            for (int k = 0; k < M; ++k) sxz = fmac(c.cs[2][0][k], LU[me + 1 + k] - LU[me - k], sxz);
#pragma unroll
            for (int k = 0; k < M; ++k) sxz = fmac(c.cs[2][1][k], qW[M + k] - qW[M - 1 - k], sxz);
            if (HET) {
                C::stnt(F.f[6][t1] + q, fmac(__builtin_nontemporal_load(md.m[OPESCI_M_MU12] + q), sxy, T0[3]));
                C::stnt(F.f[7][t1] + q, fmac(__builtin_nontemporal_load(md.m[OPESCI_M_MU23] + q), syz, T0[4]));
                C::stnt(F.f[8][t1] + q, fmac(__builtin_nontemporal_load(md.m[OPESCI_M_MU13] + q), sxz, T0[5]));
            } else {
                C::stnt(F.f[6][t1] + q, T0[3] + sxy);
                C::stnt(F.f[7][t1] + q, T0[4] + syz);
                C::stnt(F.f[8][t1] + q, T0[5] + sxz);
            }
        }
    }
}

template <typename R, typename S, int M, int TY, bool HET>
__global__ __launch_bounds__(64 * TY) void velocity_sweep(Geom g, FieldPtrs<S> F, MediaPtrs<R> md, Coefs<R> c,
                                                          int t0, int t1, int tz, int ty, int xchunk, int nblocks) {
    using P = Plane<M, TY>;
    using C = Conv<S, R>;
    __shared__ R sYY[2][P::H * P::W], sZZ[2][P::H * P::W], sXY[2][P::H * P::W], sYZ[2][P::H * P::W],
        sXZ[2][P::H * P::W];
    const int l = logical_block(nblocks);
    if (l < 0) return;
    const int bz = l % tz, by = (l / tz) % ty, bx = l / (tz * ty);
    const int lz = threadIdx.x & 63, ly = threadIdx.x >> 6;
    const int z0 = M + bz * 64, y0 = M + by * TY;
    const int z = z0 + lz, y = y0 + ly;
    const bool active = z < g.nz - M && y < g.ny - M;
    const int64_t col = (int64_t)min(y, g.ny - 1) * g.sy + min(z, g.nz - 1);
    const int xb = M + bx * xchunk, xe = min(xb + xchunk, g.nx - M);
    const S* TXX = F.f[3][t1];
    const S* TYY = F.f[4][t1];
    const S* TZZ = F.f[5][t1];
    const S* TXY = F.f[6][t1];
    const S* TYZ = F.f[7][t1];
    const S* TXZ = F.f[8][t1];
    HaloSlots<M, TY> hs;
    hs.init(g, y0, z0);
    // qXX[i] = Txx(x+i-M+1) (a = 1), qXY[i] = Txy(x+i-M), qXZ[i] = Txz(x+i-M) (a = 0)
    R qXX[2 * M], qXY[2 * M], qXZ[2 * M];
#pragma unroll
    for (int i = 1; i < 2 * M; ++i) {
        qXX[i] = C::ldnt(TXX + (int64_t)(xb - M + i) * g.sx + col);
        qXY[i] = C::ld(TXY + (int64_t)(xb - M - 1 + i) * g.sx + col);
        qXZ[i] = C::ld(TXZ + (int64_t)(xb - M - 1 + i) * g.sx + col);
    }
    R nXX, nXY, nXZ, nYY, nZZ, nYZ, nU, nV, nW;
    Staged<R, M, TY, true, false> hYY, hXY;
    Staged<R, M, TY, false, true> hZZ, hXZ;
    Staged<R, M, TY, true, true> hYZ;
    auto fetch = [&](int x) {
        const int64_t px = (int64_t)x * g.sx;
        nXX = C::ldnt(TXX + px + (int64_t)M * g.sx + col);   // read along x only: streamed once
        nXY = C::ld(TXY + px + (int64_t)(M - 1) * g.sx + col);
        nXZ = C::ld(TXZ + px + (int64_t)(M - 1) * g.sx + col);
        nYY = C::ld(TYY + px + col);
        nZZ = C::ld(TZZ + px + col);
        nYZ = C::ld(TYZ + px + col);
        nU = C::ldnt(F.f[0][t0] + px + col);
        nV = C::ldnt(F.f[1][t0] + px + col);
        nW = C::ldnt(F.f[2][t0] + px + col);
        hYY.fetch(TYY, px, hs);
        hXY.fetch(TXY, px, hs);
        hZZ.fetch(TZZ, px, hs);
        hXZ.fetch(TXZ, px, hs);
        hYZ.fetch(TYZ, px, hs);
    };
    fetch(xb);
    for (int x = xb; x < xe; ++x) {
        const int b = (x - xb) & 1;
#pragma unroll
        for (int i = 0; i < 2 * M - 1; ++i) {
            qXX[i] = qXX[i + 1];
            qXY[i] = qXY[i + 1];
            qXZ[i] = qXZ[i + 1];
        }
        qXX[2 * M - 1] = nXX;
        qXY[2 * M - 1] = nXY;
        qXZ[2 * M - 1] = nXZ;
        const R u0 = nU, v0 = nV, w0 = nW;
        R* LYY = sYY[b];
        R* LZZ = sZZ[b];
        R* LXY = sXY[b];
        R* LYZ = sYZ[b];
        R* LXZ = sXZ[b];
        const int me = (ly + M) * P::W + lz + M;
        LYY[me] = nYY;
        LZZ[me] = nZZ;
        LYZ[me] = nYZ;
        LXY[me] = qXY[M];
        LXZ[me] = qXZ[M];
        hYY.put(LYY, hs);
        hXY.put(LXY, hs);
        hZZ.put(LZZ, hs);
        hXZ.put(LXZ, hs);
        hYZ.put(LYZ, hs);
        if (x + 1 < xe) fetch(x + 1);
        __syncthreads();
        if (active) {
            const int64_t q = (int64_t)x * g.sx + col;
            R su = 0, sv = 0, sw = 0;
#pragma unroll
            for (int k = 0; k < M; ++k) su = fmac(c.cv[0][0][k], qXX[M + k] - qXX[M - 1 - k], su);
#pragma unroll
            for (int k = 0; k < M; ++k) su = fmac(c.cv[0][1][k], LXY[me + k * P::W] - LXY[me - (1 + k) * P::W], su);
#pragma unroll
            for (int k = 0; k < M; ++k) su = fmac(c.cv[0][2][k], LXZ[me + k] - LXZ[me - 1 - k], su);
#pragma unroll
            for (int k = 0; k < M; ++k) sv = fmac(c.cv[1][0][k], qXY[M + k] - qXY[M - 1 - k], sv);
#pragma unroll
            for (int k = 0; k < M; ++k) sv = fmac(c.cv[1][1][k], LYY[me + (1 + k) * P::W] - LYY[me - k * P::W], sv);
#pragma unroll
            for (int k = 0; k < M; ++k) sv = fmac(c.cv[1][2][k], LYZ[me + k] - LYZ[me - 1 - k], sv);
#pragma unroll
            for (int k = 0; k < M; ++k) sw = fmac(c.cv[2][0][k], qXZ[M + k] - qXZ[M - 1 - k], sw);
#pragma unroll
            for (int k = 0; k < M; ++k) sw = fmac(c.cv[2][1][k], LYZ[me + k * P::W] - LYZ[me - (1 + k) * P::W], sw);
#pragma unroll
            for (int k = 0; k < M; ++k) sw = fmac(c.cv[2][2][k], LZZ[me + 1 + k] - LZZ[me - k], sw);
            if (HET) {
                C::stnt(F.f[0][t1] + q, fmac(__builtin_nontemporal_load(md.m[OPESCI_M_BETA1] + q), su, u0));
                C::stnt(F.f[1][t1] + q, fmac(__builtin_nontemporal_load(md.m[OPESCI_M_BETA2] + q), sv, v0));
                C::stnt(F.f[2][t1] + q, fmac(__builtin_nontemporal_load(md.m[OPESCI_M_BETA3] + q), sw, w0));
            } else {
                C::stnt(F.f[0][t1] + q, u0 + su);
                C::stnt(F.f[1][t1] + q, v0 + sv);
                C::stnt(F.f[2][t1] + q, w0 + sw);
            }
        }
    }
}

// ------------------------------------------------------------------ pair sweeps (FP32 arithmetic)
// The same sweeps with two z points per lane: a tile is 128 (z) x TY (y),
// lane lz owns z0 + 2 lz and z0 + 2 lz + 1, and every own-column value and
// y-halo row is moved as one pair (32 bits for FP16/BF16 storage, 64 bits
// for FP32; tiles start at z = m, which the layout aligns to 128 bytes, so
// pairs are aligned; the row padding holds the partner of the last point).
// Twice the bytes per load instruction and half the blocks' z halos: at
// 256^3 the FP32 read path runs 11 % faster than with one point per lane.
// Per point the arithmetic is the single-point sweeps' and the gather
// kernels'.
template <int M, int TY>
struct Plane2 {
    static constexpr int W = 128 + 2 * M;
    static constexpr int H = TY + 2 * M;
    static constexpr int NT = 64 * TY;
    static constexpr int NY = 2 * M * 64;   // y-halo pairs
    static constexpr int NZ = TY * 2 * M;   // z-halo cells
    static constexpr int JY = (NY + NT - 1) / NT;
    static_assert(NZ <= NT, "z halo larger than the block");
};

// first z of the aligned pair holding z (z - m even), kept inside the row
__device__ __forceinline__ int pair_start(const Geom& g, int z) {
    z = min(max(z, 0), g.nz - 1);
    return z - ((z - g.m) & 1);
}

template <int M, int TY>
struct HaloSlots2 {
    using P = Plane2<M, TY>;
    int ly[P::JY], gy[P::JY];
    int lz, gz;
    __device__ void init(const Geom& g, int y0, int z0) {
        const int t = threadIdx.x;
#pragma unroll
        for (int j = 0; j < P::JY; ++j) {
            const int h = t + j * P::NT;
            if (h < P::NY) {
                const int r = h / 64, row = r < M ? r : r + TY, cl = M + 2 * (h % 64);
                const int yy = min(max(y0 - M + row, 0), g.ny - 1), zz = pair_start(g, z0 - M + cl);
                ly[j] = row * P::W + cl;
                gy[j] = yy * (int)g.sy + zz;
            } else {
                ly[j] = -1;
                gy[j] = 0;
            }
        }
        if (t < P::NZ) {
            const int row = M + t / (2 * M), c = t % (2 * M), cl = c < M ? c : 128 + c;
            const int yy = min(max(y0 - M + row, 0), g.ny - 1), zz = min(max(z0 - M + cl, 0), g.nz - 1);
            lz = row * P::W + cl;
            gz = yy * (int)g.sy + zz;
        } else {
            lz = -1;
            gz = 0;
        }
    }
};

template <int M, int TY, bool HY, bool HZ>
struct Staged2 {
    using P = Plane2<M, TY>;
    float y0[HY ? P::JY : 1], y1[HY ? P::JY : 1];
    float z;
    template <typename S>
    __device__ __forceinline__ void fetch(const S* base, int64_t px, const HaloSlots2<M, TY>& hs) {
        if (HY) {
#pragma unroll
            for (int j = 0; j < P::JY; ++j)
                if (hs.ly[j] >= 0) ld2(base + px + hs.gy[j], y0[j], y1[j]);
        }
        if (HZ && hs.lz >= 0) z = Conv<S, float>::ld(base + px + hs.gz);
    }
    __device__ __forceinline__ void put(float* lds, const HaloSlots2<M, TY>& hs) const {
        if (HY) {
#pragma unroll
            for (int j = 0; j < P::JY; ++j)
                if (hs.ly[j] >= 0) {
                    lds[hs.ly[j]] = y0[j];
                    lds[hs.ly[j] + 1] = y1[j];
                }
        }
        if (HZ && hs.lz >= 0) lds[hs.lz] = z;
    }
};

// FP32 media of a lane's two points as one 8-byte load (the media use the
// fields' element layout, so z - m even is 8-byte aligned for them)
typedef float mpair_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ mpair_t media2(const float* p) {
    return __builtin_nontemporal_load(reinterpret_cast<const mpair_t*>(p));
}

// one or two of a lane's points, as one pair or a single value (non-temporal:
// ordinary stores, which leave the fields in the memory-side cache for the
// next launch, measured slower for both sweeps at 256^3)
template <typename S>
__device__ __forceinline__ void store_points(S* p, float a, float b, bool act0, bool act1) {
    if (act1)
        stnt2(p, a, b);
    else if (act0)
        Conv<S, float>::stnt(p, a);
}

// v as lane `lane` (wave-uniform) holds it, rounded to the storage type as a
// boundary statement reading it back from memory sees it
template <typename S>
__device__ __forceinline__ float lane_read(float v, int lane) {
    const float r = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
    return Conv<S, float>::from(Conv<S, float>::to(r));
}

template <typename S, int M, int TY, bool HET, bool ZF>
__global__ __launch_bounds__(64 * TY) void stress_sweep2(Geom g, FieldPtrs<S> F, MediaPtrs<float> md,
                                                         Coefs<float> c, int t0, int t1, int tz, int nblocks,
                                                         Boxes B, ZFaceArgs zf) {
    using R = float;
    using P = Plane2<M, TY>;
    __shared__ R sU[2][P::H * P::W], sV[2][P::H * P::W], sW[2][P::H * P::W];
    const int l = logical_block(nblocks);
    if (l < 0) return;
    const TileOrigin o = tile_of<TY>(B, l, tz, M);
    const int lz = threadIdx.x & 63, ly = threadIdx.x >> 6;
    const int z0 = o.z0, y0 = o.y0;
    const int z = z0 + 2 * lz, y = y0 + ly;
    const bool act0 = z < g.nz - M && y < o.yend, act1 = z + 1 < g.nz - M && y < o.yend;
    const int64_t col = (int64_t)min(y, g.ny - 1) * g.sy + pair_start(g, z);
    const int xb = o.xb, xe = o.xe;
    const S* U0 = F.f[0][t0];
    const S* V0 = F.f[1][t0];
    const S* W0 = F.f[2][t0];
    HaloSlots2<M, TY> hs;
    hs.init(g, y0, z0);
    static_assert(!(ZF && HET), "in-sweep z faces are built for constant media");
    // z-face statements: coefficient block staged in LDS once per block (kept
    // out of the kernel arguments, whose scalar registers it would exhaust)
    __shared__ ZCoef zl;
    const int zrh = g.nz - M - 1 - z0;                 // high face relative to the tile
    const bool zlo = z0 == M, zhi = zrh >= 0 && zrh < 128;
    const int zlane = zlo && (threadIdx.x & 63) == 0 ? 0 : (zhi && (int)(threadIdx.x & 63) == (zrh >> 1) ? 1 : -1);
    if constexpr (ZF) {
        for (int w = threadIdx.x; w < ZCOEF_WORDS; w += blockDim.x)
            reinterpret_cast<int*>(&zl)[w] = reinterpret_cast<const int*>(zf.coef)[w];
        __syncthreads();
    }
    // per point v: qU[v][i] = U(x+i-M), qV[v][i] = V(x+i-M+1), qW[v][i] = W(x+i-M+1)
    R qU[2][2 * M], qV[2][2 * M], qW[2][2 * M];
#pragma unroll
    for (int i = 1; i < 2 * M; ++i) {
        ld2(U0 + (int64_t)(xb - M - 1 + i) * g.sx + col, qU[0][i], qU[1][i]);
        ld2(V0 + (int64_t)(xb - M + i) * g.sx + col, qV[0][i], qV[1][i]);
        ld2(W0 + (int64_t)(xb - M + i) * g.sx + col, qW[0][i], qW[1][i]);
    }
    R nU[2], nV[2], nW[2], nT[6][2];
    Staged2<M, TY, true, true> nhU, nhV, nhW;
    auto fetch = [&](int x) {
        const int64_t px = (int64_t)x * g.sx;
        ld2(U0 + px + (int64_t)(M - 1) * g.sx + col, nU[0], nU[1]);
        ld2(V0 + px + (int64_t)M * g.sx + col, nV[0], nV[1]);
        ld2(W0 + px + (int64_t)M * g.sx + col, nW[0], nW[1]);
#pragma unroll
        for (int f = 0; f < 6; ++f) ldnt2(F.f[3 + f][t0] + px + col, nT[f][0], nT[f][1]);
        nhU.fetch(U0, px, hs);
        nhV.fetch(V0, px, hs);
        nhW.fetch(W0, px, hs);
    };
    fetch(xb);
    for (int x = xb; x < xe; ++x) {
        const int b = (x - xb) & 1;
#pragma unroll
        for (int v = 0; v < 2; ++v) {
#pragma unroll
            for (int i = 0; i < 2 * M - 1; ++i) {
                qU[v][i] = qU[v][i + 1];
                qV[v][i] = qV[v][i + 1];
                qW[v][i] = qW[v][i + 1];
            }
            qU[v][2 * M - 1] = nU[v];
            qV[v][2 * M - 1] = nV[v];
            qW[v][2 * M - 1] = nW[v];
        }
        R T0[6][2];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            T0[f][0] = nT[f][0];
            T0[f][1] = nT[f][1];
        }
        R* LU = sU[b];
        R* LV = sV[b];
        R* LW = sW[b];
        const int me0 = (ly + M) * P::W + 2 * lz + M;
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            LU[me0 + v] = qU[v][M];
            LV[me0 + v] = qV[v][M - 1];
            LW[me0 + v] = qW[v][M - 1];
        }
        nhU.put(LU, hs);
        nhV.put(LV, hs);
        nhW.put(LW, hs);
        if (x + 1 < xe) fetch(x + 1);
        __syncthreads();
        if (act0) {
            const int64_t q = (int64_t)x * g.sx + col;
            R out[6][2];
            mpair_t mlam, mmu, m12, m23, m13;
            if (HET) {
                mlam = media2(md.m[OPESCI_M_LAMBDA] + q);
                mmu = media2(md.m[OPESCI_M_MU] + q);
                m12 = media2(md.m[OPESCI_M_MU12] + q);
                m23 = media2(md.m[OPESCI_M_MU23] + q);
                m13 = media2(md.m[OPESCI_M_MU13] + q);
            }
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                const int me = me0 + v;
                R dx[M], dy[M], dz[M];
#pragma unroll
                for (int k = 0; k < M; ++k) {
                    dx[k] = qU[v][M + k] - qU[v][M - 1 - k];
                    dy[k] = LV[me + k * P::W] - LV[me - (1 + k) * P::W];
                    dz[k] = LW[me + k] - LW[me - 1 - k];
                }
                R lam = 0, l2m = 0;
                if (HET) {
                    lam = mlam[v];
                    l2m = lam + (R)2 * mmu[v];
                }
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    R acc = T0[i][v];
                    R s = 0;
#pragma unroll
                    for (int k = 0; k < M; ++k) s = fmac(c.cn[i][0][k], dx[k], s);
                    acc = HET ? fmac(i == 0 ? l2m : lam, s, acc) : acc + s;
                    s = 0;
#pragma unroll
                    for (int k = 0; k < M; ++k) s = fmac(c.cn[i][1][k], dy[k], s);
                    acc = HET ? fmac(i == 1 ? l2m : lam, s, acc) : acc + s;
                    s = 0;
#pragma unroll
                    for (int k = 0; k < M; ++k) s = fmac(c.cn[i][2][k], dz[k], s);
                    acc = HET ? fmac(i == 2 ? l2m : lam, s, acc) : acc + s;
                    out[i][v] = acc;
                }
                R sxy = 0, syz = 0, sxz = 0;
#pragma unroll
                for (int k = 0; k < M; ++k) sxy = fmac(c.cs[0][0][k], LU[me + (1 + k) * P::W] - LU[me - k * P::W], sxy);
#pragma unroll
                for (int k = 0; k < M; ++k) sxy = fmac(c.cs[0][1][k], qV[v][M + k] - qV[v][M - 1 - k], sxy);
#pragma unroll
                for (int k = 0; k < M; ++k) syz = fmac(c.cs[1][0][k], LV[me + 1 + k] - LV[me - k], syz);
#pragma unroll
                for (int k = 0; k < M; ++k) syz = fmac(c.cs[1][1][k], LW[me + (1 + k) * P::W] - LW[me - k * P::W], syz);
#pragma unroll
                for (int k = 0; k < M; ++k) sxz = fmac(c.cs[2][0][k], LU[me + 1 + k] - LU[me - k], sxz);
#pragma unroll
                for (int k = 0; k < M; ++k) sxz = fmac(c.cs[2][1][k], qW[v][M + k] - qW[v][M - 1 - k], sxz);
                if (HET) {
                    out[3][v] = fmac(m12[v], sxy, T0[3][v]);
                    out[4][v] = fmac(m23[v], syz, T0[4][v]);
                    out[5][v] = fmac(m13[v], sxz, T0[5][v]);
                } else {
                    out[3][v] = T0[3][v] + sxy;
                    out[4][v] = T0[4][v] + syz;
                    out[5][v] = T0[5][v] + sxz;
                }
            }
            if constexpr (ZF) {
                // free-surface statements of the z faces on the deep rows (the
                // wave is one row: x and y are wave-uniform)
                if (x >= zf.F && x < g.nx - zf.F && y >= zf.F && y < g.ny - zf.F) {
                    const int64_t row = (int64_t)x * g.sx + (int64_t)y * g.sy;
                    const int rh = zrh;
                    // image sources [face][Tzz, Tyz, Txz][k] (interior values,
                    // before any replacement), read from the lanes holding
                    // them while every interior lane is active
                    float src[2][3][M];
#pragma unroll
                    for (int k = 0; k < M; ++k) {
                        if (zlo) {
                            src[0][0][k] = k >= 1 ? lane_read<S>(out[2][k & 1], k >> 1) : 0.0f;
                            src[0][1][k] = lane_read<S>(out[4][k & 1], k >> 1);
                            src[0][2][k] = lane_read<S>(out[5][k & 1], k >> 1);
                        }
                        if (zhi) {
                            const int r0 = rh - k, r1 = rh - 1 - k;
                            src[1][0][k] = k >= 1 ? lane_read<S>((r0 & 1) ? out[2][1] : out[2][0], r0 >> 1) : 0.0f;
                            src[1][1][k] = lane_read<S>((r1 & 1) ? out[4][1] : out[4][0], r1 >> 1);
                            src[1][2][k] = lane_read<S>((r1 & 1) ? out[5][1] : out[5][0], r1 >> 1);
                        }
                    }
                    // the lane holding a surface point: its face's Levander
                    // recompute of Txx and Tyy (one pass over the nine terms,
                    // point element e), surface values and ghost cells
                    if (zlane >= 0) {
                        const int face = zlane, e = face ? (rh & 1) : 0;
                        float rxx = 0.0f, ryy = 0.0f;
                        if constexpr (M == 2) {
                            if (zf.rec) {
#pragma unroll
                                for (int k = 0; k < 8; ++k) {
                                    const float v = k < 4 ? (e ? qU[1][k] : qU[0][k]) : LV[me0 + e + (k - 6) * P::W];
                                    rxx = fmac(zl.rc[face][0][k], v, rxx);
                                    ryy = fmac(zl.rc[face][1][k], v, ryy);
                                }
                                rxx = fmac(zl.rc[face][0][8], e ? T0[0][1] : T0[0][0], rxx);
                                ryy = fmac(zl.rc[face][1][8], e ? T0[1][1] : T0[1][0], ryy);
                            }
                        }
                        float sv[3][M];
#pragma unroll
                        for (int f = 0; f < 3; ++f)
#pragma unroll
                            for (int k = 0; k < M; ++k) sv[f][k] = face ? src[1][f][k] : src[0][f][k];
                        // ghost cells: ordinary stores, so the few cells of
                        // each face line merge in L2 before they reach HBM
                        using CS = Conv<S, float>;
                        if (face == 0) {
#pragma unroll
                            for (int k = 0; k < M; ++k) {
                                if (k >= 1) CS::st(F.f[5][t1] + row + M - k, fmac(zl.zc[0][0][k], sv[0][k], 0.0f));
                                CS::st(F.f[7][t1] + row + M - 1 - k, fmac(zl.zc[0][1][k], sv[1][k], 0.0f));
                                CS::st(F.f[8][t1] + row + M - 1 - k, fmac(zl.zc[0][2][k], sv[2][k], 0.0f));
                            }
                            if (zf.rec) {
                                out[0][0] = rxx;
                                out[1][0] = ryy;
                            }
                            out[2][0] = 0.0f;
                        } else {
                            const int bhi = g.nz - M - 1;
#pragma unroll
                            for (int k = 1; k < M; ++k) {
                                CS::st(F.f[5][t1] + row + bhi + k, fmac(zl.zc[1][0][k], sv[0][k], 0.0f));
                                CS::st(F.f[7][t1] + row + bhi + k, fmac(zl.zc[1][1][k], sv[1][k], 0.0f));
                                CS::st(F.f[8][t1] + row + bhi + k, fmac(zl.zc[1][2][k], sv[2][k], 0.0f));
                            }
                            const float tyz = fmac(zl.zc[1][1][0], sv[1][0], 0.0f);
                            const float txz = fmac(zl.zc[1][2][0], sv[2][0], 0.0f);
#pragma unroll
                            for (int v = 0; v < 2; ++v)
                                if (v == e) {
                                    if (zf.rec) {
                                        out[0][v] = rxx;
                                        out[1][v] = ryy;
                                    }
                                    out[2][v] = 0.0f;
                                    out[4][v] = tyz;
                                    out[5][v] = txz;
                                }
                        }
                    }
                }
            }
#pragma unroll
            for (int f = 0; f < 6; ++f)
                store_points(F.f[3 + f][t1] + q, out[f][0], out[f][1], act0, act1);
        }
    }
}

template <typename S, int M, int TY, bool HET>
__global__ __launch_bounds__(64 * TY) void velocity_sweep2(Geom g, FieldPtrs<S> F, MediaPtrs<float> md,
                                                           Coefs<float> c, int t0, int t1, int tz, int nblocks,
                                                           Boxes B) {
    using R = float;
    using P = Plane2<M, TY>;
    __shared__ R sYY[2][P::H * P::W], sZZ[2][P::H * P::W], sXY[2][P::H * P::W], sYZ[2][P::H * P::W],
        sXZ[2][P::H * P::W];
    const int l = logical_block(nblocks);
    if (l < 0) return;
    const TileOrigin o = tile_of<TY>(B, l, tz, M);
    const int lz = threadIdx.x & 63, ly = threadIdx.x >> 6;
    const int z0 = o.z0, y0 = o.y0;
    const int z = z0 + 2 * lz, y = y0 + ly;
    const bool act0 = z < g.nz - M && y < o.yend, act1 = z + 1 < g.nz - M && y < o.yend;
    const int64_t col = (int64_t)min(y, g.ny - 1) * g.sy + pair_start(g, z);
    const int xb = o.xb, xe = o.xe;
    const S* TXX = F.f[3][t1];
    const S* TYY = F.f[4][t1];
    const S* TZZ = F.f[5][t1];
    const S* TXY = F.f[6][t1];
    const S* TYZ = F.f[7][t1];
    const S* TXZ = F.f[8][t1];
    HaloSlots2<M, TY> hs;
    hs.init(g, y0, z0);
    // per point v: qXX[v][i] = Txx(x+i-M+1), qXY[v][i] = Txy(x+i-M), qXZ[v][i] = Txz(x+i-M)
    R qXX[2][2 * M], qXY[2][2 * M], qXZ[2][2 * M];
#pragma unroll
    for (int i = 1; i < 2 * M; ++i) {
        ldnt2(TXX + (int64_t)(xb - M + i) * g.sx + col, qXX[0][i], qXX[1][i]);
        ld2(TXY + (int64_t)(xb - M - 1 + i) * g.sx + col, qXY[0][i], qXY[1][i]);
        ld2(TXZ + (int64_t)(xb - M - 1 + i) * g.sx + col, qXZ[0][i], qXZ[1][i]);
    }
    R nXX[2], nXY[2], nXZ[2], nYY[2], nZZ[2], nYZ[2], nU[2], nV[2], nW[2];
    Staged2<M, TY, true, false> hYY, hXY;
    Staged2<M, TY, false, true> hZZ, hXZ;
    Staged2<M, TY, true, true> hYZ;
    auto fetch = [&](int x) {
        const int64_t px = (int64_t)x * g.sx;
        ldnt2(TXX + px + (int64_t)M * g.sx + col, nXX[0], nXX[1]);
        ld2(TXY + px + (int64_t)(M - 1) * g.sx + col, nXY[0], nXY[1]);
        ld2(TXZ + px + (int64_t)(M - 1) * g.sx + col, nXZ[0], nXZ[1]);
        ld2(TYY + px + col, nYY[0], nYY[1]);
        ld2(TZZ + px + col, nZZ[0], nZZ[1]);
        ld2(TYZ + px + col, nYZ[0], nYZ[1]);
        ldnt2(F.f[0][t0] + px + col, nU[0], nU[1]);
        ldnt2(F.f[1][t0] + px + col, nV[0], nV[1]);
        ldnt2(F.f[2][t0] + px + col, nW[0], nW[1]);
        hYY.fetch(TYY, px, hs);
        hXY.fetch(TXY, px, hs);
        hZZ.fetch(TZZ, px, hs);
        hXZ.fetch(TXZ, px, hs);
        hYZ.fetch(TYZ, px, hs);
    };
    fetch(xb);
    for (int x = xb; x < xe; ++x) {
        const int b = (x - xb) & 1;
#pragma unroll
        for (int v = 0; v < 2; ++v) {
#pragma unroll
            for (int i = 0; i < 2 * M - 1; ++i) {
                qXX[v][i] = qXX[v][i + 1];
                qXY[v][i] = qXY[v][i + 1];
                qXZ[v][i] = qXZ[v][i + 1];
            }
            qXX[v][2 * M - 1] = nXX[v];
            qXY[v][2 * M - 1] = nXY[v];
            qXZ[v][2 * M - 1] = nXZ[v];
        }
        const R u0[2] = {nU[0], nU[1]}, v0[2] = {nV[0], nV[1]}, w0[2] = {nW[0], nW[1]};
        R* LYY = sYY[b];
        R* LZZ = sZZ[b];
        R* LXY = sXY[b];
        R* LYZ = sYZ[b];
        R* LXZ = sXZ[b];
        const int me0 = (ly + M) * P::W + 2 * lz + M;
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            LYY[me0 + v] = nYY[v];
            LZZ[me0 + v] = nZZ[v];
            LYZ[me0 + v] = nYZ[v];
            LXY[me0 + v] = qXY[v][M];
            LXZ[me0 + v] = qXZ[v][M];
        }
        hYY.put(LYY, hs);
        hXY.put(LXY, hs);
        hZZ.put(LZZ, hs);
        hXZ.put(LXZ, hs);
        hYZ.put(LYZ, hs);
        if (x + 1 < xe) fetch(x + 1);
        __syncthreads();
        if (act0) {
            const int64_t q = (int64_t)x * g.sx + col;
            R ou[2], ov[2], ow[2];
            mpair_t b1, b2, b3;
            if (HET) {
                b1 = media2(md.m[OPESCI_M_BETA1] + q);
                b2 = media2(md.m[OPESCI_M_BETA2] + q);
                b3 = media2(md.m[OPESCI_M_BETA3] + q);
            }
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                const int me = me0 + v;
                R su = 0, sv = 0, sw = 0;
#pragma unroll
                for (int k = 0; k < M; ++k) su = fmac(c.cv[0][0][k], qXX[v][M + k] - qXX[v][M - 1 - k], su);
#pragma unroll
                for (int k = 0; k < M; ++k) su = fmac(c.cv[0][1][k], LXY[me + k * P::W] - LXY[me - (1 + k) * P::W], su);
#pragma unroll
                for (int k = 0; k < M; ++k) su = fmac(c.cv[0][2][k], LXZ[me + k] - LXZ[me - 1 - k], su);
#pragma unroll
                for (int k = 0; k < M; ++k) sv = fmac(c.cv[1][0][k], qXY[v][M + k] - qXY[v][M - 1 - k], sv);
#pragma unroll
                for (int k = 0; k < M; ++k) sv = fmac(c.cv[1][1][k], LYY[me + (1 + k) * P::W] - LYY[me - k * P::W], sv);
#pragma unroll
                for (int k = 0; k < M; ++k) sv = fmac(c.cv[1][2][k], LYZ[me + k] - LYZ[me - 1 - k], sv);
#pragma unroll
                for (int k = 0; k < M; ++k) sw = fmac(c.cv[2][0][k], qXZ[v][M + k] - qXZ[v][M - 1 - k], sw);
#pragma unroll
                for (int k = 0; k < M; ++k) sw = fmac(c.cv[2][1][k], LYZ[me + k * P::W] - LYZ[me - (1 + k) * P::W], sw);
#pragma unroll
                for (int k = 0; k < M; ++k) sw = fmac(c.cv[2][2][k], LZZ[me + 1 + k] - LZZ[me - k], sw);
                if (HET) {
                    ou[v] = fmac(b1[v], su, u0[v]);
                    ov[v] = fmac(b2[v], sv, v0[v]);
                    ow[v] = fmac(b3[v], sw, w0[v]);
                } else {
                    ou[v] = u0[v] + su;
                    ov[v] = v0[v] + sv;
                    ow[v] = w0[v] + sw;
                }
            }
            store_points(F.f[0][t1] + q, ou[0], ou[1], act0, act1);
            store_points(F.f[1][t1] + q, ov[0], ov[1], act0, act1);
            store_points(F.f[2][t1] + q, ow[0], ow[1], act0, act1);
        }
    }
}

// ------------------------------------------------------------------ fused step
// Both half steps of one leapfrog step in one x sweep, for FP32 arithmetic
// on constant media with the deep z-face statements of the stress phase in
// the sweep.  It covers a box of rows far enough from the x and y faces
// that no stress boundary statement writes a cell its velocity update reads
// (the host checks this against the boundary program); the frame around the
// box keeps the separate stress and velocity launches, and the velocity
// boundary statements run after both as before.
//
// A block owns a 128 (z) x TY (y) tile of an x chunk [xb, xe).  Its waves
// hold rows y0 - M .. y0 + TY + M - 1: TY own rows between M halo rows on
// either side.  At step xs of the sweep (xs = xb - M .. xe + M - 1) every
// wave computes the stresses of its row at plane xs exactly as the pair
// stress sweep does (velocity plane xs staged in LDS with 2M halo rows and
// 2M z columns either side, x stencils in register queues), lanes 0 and 63
// of the own waves add the z-halo pairs below and above the tile (Txz, Tyz,
// Tzz) where a neighbouring tile, not a face, lies there, and the z-face
// statements replace the surface values and produce the ghost cells.  The
// stresses go to an LDS plane (rounded to the storage type, as the separate
// velocity launch would read them back) and, for own rows and xs in
// [xb, xe), to memory.  The own waves then keep the y and z stencil
// differences of the velocity sums of plane xs from that plane for M
// planes, until the stresses up to xs + M are known: at step xs they form
// the velocities of plane xs - M, x part from rings of Txx, Txy, Txz first,
// then the kept differences, the separate velocity sweep's fma chain.  So
// every value is the separate launches' bit for bit, and a step reads U, V,
// W and the six stresses once and writes them once (18 streams against
// 15 + 12).
template <int M, int TY>
struct FusedTile {
    static constexpr int NW = TY + 2 * M;   // waves
    static constexpr int NT = 64 * NW;
    static constexpr int VW = 128 + 4 * M;  // velocity plane: rows y0-2M.., z from z0-2M
    static constexpr int VH = TY + 4 * M;
    static constexpr int SW = 128 + 2 * M;  // stress plane: rows y0-M.., z from z0-M
    static constexpr int SH = TY + 2 * M;
    static constexpr int NO = 3 * 2 * M * 64;  // velocity pairs of the 2M outer rows (3 fields)
    static constexpr int JO = (NO + NT - 1) / NT;
    static constexpr int NZF = SH * 2 * M;             // velocity pairs of the z columns per field (M either side)
    static constexpr int NZP = (NZF + 63) / 64 * 64;   // padded: one field per wave
    static constexpr int NZ = 3 * NZP;
    static constexpr int JZ = (NZ + NT - 1) / NT;
};

// staging slots of the velocity plane beyond the waves' own pairs
template <int M, int TY>
struct FusedSlots {
    using T = FusedTile<M, TY>;
    int lo[T::JO], go[T::JO], fo[T::JO];  // outer rows: LDS slot, offset, field (-1: none)
    int lz[T::JZ], gz[T::JZ], fz[T::JZ];  // z columns
    __device__ void init(const Geom& g, int y0, int z0) {
        const int t = threadIdx.x;
#pragma unroll
        for (int j = 0; j < T::JO; ++j) {
            const int h = t + j * T::NT;
            fo[j] = -1;
            lo[j] = 0;
            go[j] = 0;
            if (h < T::NO) {
                const int f = h / (2 * M * 64), r = (h / 64) % (2 * M), p = h % 64;
                const int row = r < M ? r : r + TY + 2 * M;
                const int yy = min(max(y0 - 2 * M + row, 0), g.ny - 1);
                fo[j] = f;
                lo[j] = row * T::VW + 2 * M + 2 * p;
                go[j] = yy * (int)g.sy + pair_start(g, z0 + 2 * p);
            }
        }
#pragma unroll
        for (int j = 0; j < T::JZ; ++j) {
            const int h = t + j * T::NT;
            fz[j] = -1;
            lz[j] = 0;
            gz[j] = 0;
            if (h < T::NZ && h % T::NZP < T::NZF) {
                const int f = h / T::NZP, rem = h % T::NZP;
                const int row = rem / (2 * M) + M, c = rem % (2 * M);
                const int zz = c < M ? z0 - 2 * M + 2 * c : z0 + 128 + 2 * (c - M);
                const int yy = min(max(y0 - 2 * M + row, 0), g.ny - 1);
                fz[j] = f;
                lz[j] = row * T::VW + (zz - (z0 - 2 * M));
                gz[j] = yy * (int)g.sy + pair_start(g, zz);
            }
        }
    }
};

// field pointers of one fused step, levels resolved on the host
template <typename S>
struct StepPtrs {
    const S* v0[3];   // U, V, W at t0
    const S* s0[6];   // Txx, Tyy, Tzz, Txy, Tyz, Txz at t0
    S* s1[6];         // the stresses at t1
    S* v1[3];         // U, V, W at t1
};

template <typename S, int M, int TY>
__global__ __launch_bounds__(64 * (TY + 2 * M)) void step_sweep2(Geom g, StepPtrs<S> P, Coefs<float> c, int tz,
                                                                 int nblocks, Boxes B, ZFaceArgs zf) {
    using R = float;
    using T = FusedTile<M, TY>;
    using CS = Conv<S, float>;
    constexpr int L = 2 * M + 1;    // planes kept in the LDS rings
    constexpr int NOWN = 64 * TY;   // own-row lanes
    __shared__ R vU[T::VH * T::VW], vV[T::VH * T::VW], vW[T::VH * T::VW];
    __shared__ R sXY[T::SH * T::SW], sYY[T::SH * T::SW], sYZ[T::SH * T::SW], sXZ[T::SH * T::SW],
        sZZ[T::SH * T::SW];
    // own lanes: Txx, Txy, Txz of the last planes, ring slot p % L for plane
    // p, [slot][element][lane]; the y and z differences of the velocity
    // sums of the last M + 1 planes, [p % (M + 1)][element][difference][lane]
    // (U: Txy along y, Txz along z; V: Tyy, Tyz; W: Tyz, Tzz; M each)
    constexpr int LD = M + 1, ND = 6 * M;
    __shared__ R rXX[L * 2 * NOWN], rXY[L * 2 * NOWN], rXZ[L * 2 * NOWN], rD[LD * 2 * ND * NOWN];
    // z-halo pairs: U and W of the last planes, [row][side][slot][field][element]
    __shared__ R rI[TY * 2 * L * 2 * 2];
    __shared__ ZCoef zl;
    const int l = logical_block(nblocks);
    if (l < 0) return;
    const TileOrigin o = tile_of<TY>(B, l, tz, M);
    const int w = threadIdx.x >> 6, lz = threadIdx.x & 63;
    const bool own = w >= M && w < M + TY;
    const int tw = (w - M) * 64 + lz;   // own-lane index (own waves)
    const int z0 = o.z0, y0 = o.y0, xb = o.xb, xe = o.xe;
    const int y = y0 - M + w, z = z0 + 2 * lz;
    const bool rowc = y < o.yend + M, rowo = own && y < o.yend;
    const bool zin0 = z < g.nz - M, zin1 = z + 1 < g.nz - M;
    const bool act0 = rowc && zin0, act1 = rowc && zin1;
    const int64_t col = (int64_t)min(y, g.ny - 1) * g.sy + pair_start(g, z);
    const int zrh = g.nz - M - 1 - z0;
    const bool zlo = z0 == M, zhi = zrh >= 0 && zrh < 128;
    const int zlane = zlo && lz == 0 ? 0 : (zhi && lz == (zrh >> 1) ? 1 : -1);
    const bool item = own && ((lz == 0 && !zlo) || (lz == 63 && !zhi && z0 + 128 < g.nz - M));
    const int zi = lz == 0 ? z0 - 2 : z0 + 128;
    const int64_t coli = (int64_t)min(y, g.ny - 1) * g.sy + pair_start(g, zi);
    R* ring_i = rI + (((own ? w - M : 0) * 2 + (lz == 63 ? 1 : 0)) * L) * 4;
    FusedSlots<M, TY> hs;
    hs.init(g, y0, z0);
    for (int i = threadIdx.x; i < T::SH * T::SW; i += T::NT) {
        sXY[i] = 0.0f;
        sYY[i] = 0.0f;
        sYZ[i] = 0.0f;
        sXZ[i] = 0.0f;
        sZZ[i] = 0.0f;
    }
    for (int i = threadIdx.x; i < ZCOEF_WORDS; i += T::NT)
        reinterpret_cast<int*>(&zl)[i] = reinterpret_cast<const int*>(zf.coef)[i];
    const int xs0 = xb - M, xs1 = xe + M;
    // stress side: qU[v][i] = U(x-M+i), qV[v][i] = V(x-M+i), qW[v][i] = W(x-M+i)
    R qU[2][2 * M], qV[2][2 * M + 1], qW[2][2 * M + 1];
#pragma unroll
    for (int i = 1; i < 2 * M; ++i) ld2(P.v0[0] + (int64_t)(xs0 - M - 1 + i) * g.sx + col, qU[0][i], qU[1][i]);
#pragma unroll
    for (int i = 1; i < 2 * M + 1; ++i) {
        ld2(P.v0[1] + (int64_t)(xs0 - M - 1 + i) * g.sx + col, qV[0][i], qV[1][i]);
        ld2(P.v0[2] + (int64_t)(xs0 - M - 1 + i) * g.sx + col, qW[0][i], qW[1][i]);
    }
    // z-halo pairs: U(x-M .. x+M-2) and W(x-M+1 .. x+M-1) into ring slots,
    // the newest plane arriving with each step
    if (item) {
        for (int p = xs0 - M; p < xs0 + M; ++p) {
            const int sl = ((p % L) + L) % L;
            R a0, a1;
            if (p < xs0 + M - 1) {
                ld2(P.v0[0] + (int64_t)p * g.sx + coli, a0, a1);
                ring_i[(sl * 2 + 0) * 2 + 0] = a0;
                ring_i[(sl * 2 + 0) * 2 + 1] = a1;
            }
            if (p > xs0 - M) {
                ld2(P.v0[2] + (int64_t)p * g.sx + coli, a0, a1);
                ring_i[(sl * 2 + 1) * 2 + 0] = a0;
                ring_i[(sl * 2 + 1) * 2 + 1] = a1;
            }
        }
    }
    R nU[2], nV[2], nW[2], nT[6][2], nUi[2] = {0.0f, 0.0f}, nWi[2] = {0.0f, 0.0f}, nTi[3][2] = {}, so[T::JO][2],
        sz[T::JZ][2];
    auto fetch = [&](int x) {
        const int64_t px = (int64_t)x * g.sx;
        ld2(P.v0[0] + px + (int64_t)(M - 1) * g.sx + col, nU[0], nU[1]);
        ld2(P.v0[1] + px + (int64_t)M * g.sx + col, nV[0], nV[1]);
        ld2(P.v0[2] + px + (int64_t)M * g.sx + col, nW[0], nW[1]);
#pragma unroll
        for (int f = 0; f < 6; ++f) ldnt2(P.s0[f] + px + col, nT[f][0], nT[f][1]);
        if (item) {
            ld2(P.v0[0] + px + (int64_t)(M - 1) * g.sx + coli, nUi[0], nUi[1]);
            ld2(P.v0[2] + px + (int64_t)M * g.sx + coli, nWi[0], nWi[1]);
            ld2(P.s0[2] + px + coli, nTi[0][0], nTi[0][1]);
            ld2(P.s0[4] + px + coli, nTi[1][0], nTi[1][1]);
            ld2(P.s0[5] + px + coli, nTi[2][0], nTi[2][1]);
        }
#pragma unroll
        for (int j = 0; j < T::JO; ++j)
            if (hs.fo[j] >= 0) ld2(P.v0[hs.fo[j]] + px + hs.go[j], so[j][0], so[j][1]);
#pragma unroll
        for (int j = 0; j < T::JZ; ++j)
            if (hs.fz[j] >= 0) ld2(P.v0[hs.fz[j]] + px + hs.gz[j], sz[j][0], sz[j][1]);
    };
    fetch(xs0);
    const int vme0 = (w + M) * T::VW + 2 * M + 2 * lz;
    const int sme0 = w * T::SW + M + 2 * lz;
    const int vmi = (w + M) * T::VW + (zi - (z0 - 2 * M));
    const int smi = w * T::SW + (zi - (z0 - M));
    auto rnd = [](float v) { return CS::from(CS::to(v)); };
    int slot = ((xs0 % L) + L) % L;     // ring slot of plane x
    int dslot = ((xs0 % LD) + LD) % LD;  // difference ring slot of plane x
    auto back = [&](int d) { return slot >= d ? slot - d : slot - d + L; };   // slot of plane x - d
    for (int x = xs0; x < xs1; ++x, slot = slot + 1 == L ? 0 : slot + 1, dslot = dslot + 1 == LD ? 0 : dslot + 1) {
#pragma unroll
        for (int v = 0; v < 2; ++v) {
#pragma unroll
            for (int i = 0; i < 2 * M - 1; ++i) qU[v][i] = qU[v][i + 1];
#pragma unroll
            for (int i = 0; i < 2 * M; ++i) {
                qV[v][i] = qV[v][i + 1];
                qW[v][i] = qW[v][i + 1];
            }
            qU[v][2 * M - 1] = nU[v];
            qV[v][2 * M] = nV[v];
            qW[v][2 * M] = nW[v];
        }
        R T0[6][2], T0i[3][2];
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            T0[f][0] = nT[f][0];
            T0[f][1] = nT[f][1];
        }
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            T0i[f][0] = nTi[f][0];
            T0i[f][1] = nTi[f][1];
        }
        if (item) {
            // U(x+M-1) and W(x+M) of the z-halo pair
            const int su = back(L - (M - 1)), sw = back(L - M);
            ring_i[(su * 2 + 0) * 2 + 0] = nUi[0];
            ring_i[(su * 2 + 0) * 2 + 1] = nUi[1];
            ring_i[(sw * 2 + 1) * 2 + 0] = nWi[0];
            ring_i[(sw * 2 + 1) * 2 + 1] = nWi[1];
        }
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            vU[vme0 + v] = qU[v][M];
            vV[vme0 + v] = qV[v][M];
            vW[vme0 + v] = qW[v][M];
        }
#pragma unroll
        for (int j = 0; j < T::JO; ++j)
            if (hs.fo[j] >= 0) {
                R* Lp = hs.fo[j] == 0 ? vU : (hs.fo[j] == 1 ? vV : vW);
                Lp[hs.lo[j]] = so[j][0];
                Lp[hs.lo[j] + 1] = so[j][1];
            }
#pragma unroll
        for (int j = 0; j < T::JZ; ++j)
            if (hs.fz[j] >= 0) {
                R* Lp = hs.fz[j] == 0 ? vU : (hs.fz[j] == 1 ? vV : vW);
                Lp[hs.lz[j]] = sz[j][0];
                Lp[hs.lz[j] + 1] = sz[j][1];
            }
        if (x + 1 < xs1) fetch(x + 1);
        __syncthreads();
        const bool xo = x >= xb && x < xe;
        const int64_t q = (int64_t)x * g.sx + col;
        R out[6][2] = {};
        if (act0) {
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                const int me = vme0 + v;
                R dx[M], dy[M], dz[M];
#pragma unroll
                for (int k = 0; k < M; ++k) {
                    dx[k] = qU[v][M + k] - qU[v][M - 1 - k];
                    dy[k] = vV[me + k * T::VW] - vV[me - (1 + k) * T::VW];
                    dz[k] = vW[me + k] - vW[me - 1 - k];
                }
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    R acc = T0[i][v];
                    R s = 0;
#pragma unroll
                    for (int k = 0; k < M; ++k) s = fmac(c.cn[i][0][k], dx[k], s);
                    acc = acc + s;
                    s = 0;
#pragma unroll
                    for (int k = 0; k < M; ++k) s = fmac(c.cn[i][1][k], dy[k], s);
                    acc = acc + s;
                    s = 0;
#pragma unroll
                    for (int k = 0; k < M; ++k) s = fmac(c.cn[i][2][k], dz[k], s);
                    acc = acc + s;
                    out[i][v] = acc;
                }
                R sxy = 0, syz = 0, sxz = 0;
#pragma unroll
                for (int k = 0; k < M; ++k)
                    sxy = fmac(c.cs[0][0][k], vU[me + (1 + k) * T::VW] - vU[me - k * T::VW], sxy);
#pragma unroll
                for (int k = 0; k < M; ++k) sxy = fmac(c.cs[0][1][k], qV[v][M + k + 1] - qV[v][M - k], sxy);
#pragma unroll
                for (int k = 0; k < M; ++k) syz = fmac(c.cs[1][0][k], vV[me + 1 + k] - vV[me - k], syz);
#pragma unroll
                for (int k = 0; k < M; ++k)
                    syz = fmac(c.cs[1][1][k], vW[me + (1 + k) * T::VW] - vW[me - k * T::VW], syz);
#pragma unroll
                for (int k = 0; k < M; ++k) sxz = fmac(c.cs[2][0][k], vU[me + 1 + k] - vU[me - k], sxz);
#pragma unroll
                for (int k = 0; k < M; ++k) sxz = fmac(c.cs[2][1][k], qW[v][M + k + 1] - qW[v][M - k], sxz);
                out[3][v] = T0[3][v] + sxy;
                out[4][v] = T0[4][v] + syz;
                out[5][v] = T0[5][v] + sxz;
            }
        }
        // z-face statements (box rows are deep rows: x and y in [F, dim - F))
        float ghost[3][M];   // [Tzz, Tyz, Txz][k] below the low face or above the high face
        const bool zrow = act0 && x >= zf.F && x < g.nx - zf.F && y >= zf.F && y < g.ny - zf.F;
        // image sources [face][Tzz, Tyz, Txz][k], read from the lanes holding
        // them while every interior lane (each source among them) is active
        float src[2][3][M];
        if (zrow) {
            const int rh = zrh;
#pragma unroll
            for (int k = 0; k < M; ++k) {
                if (zlo) {
                    src[0][0][k] = k >= 1 ? lane_read<S>(out[2][k & 1], k >> 1) : 0.0f;
                    src[0][1][k] = lane_read<S>(out[4][k & 1], k >> 1);
                    src[0][2][k] = lane_read<S>(out[5][k & 1], k >> 1);
                }
                if (zhi) {
                    const int r0 = rh - k, r1 = rh - 1 - k;
                    src[1][0][k] = k >= 1 ? lane_read<S>((r0 & 1) ? out[2][1] : out[2][0], r0 >> 1) : 0.0f;
                    src[1][1][k] = lane_read<S>((r1 & 1) ? out[4][1] : out[4][0], r1 >> 1);
                    src[1][2][k] = lane_read<S>((r1 & 1) ? out[5][1] : out[5][0], r1 >> 1);
                }
            }
        }
        if (zrow && zlane >= 0) {
            const int face = zlane;
            const int eh = face ? (zrh & 1) : 0;
            float rxx = 0.0f, ryy = 0.0f;
            if constexpr (M == 2) {
                if (zf.rec) {
                    // Levander recompute of Txx, Tyy on the surface point
#pragma unroll
                    for (int k = 0; k < 9; ++k) {
                        const float v = k < 4 ? (eh ? qU[1][k] : qU[0][k])
                                              : (k < 8 ? vV[vme0 + eh + (k - 6) * T::VW] : 0.0f);
                        if (k < 8) {
                            rxx = fmac(zl.rc[face][0][k], v, rxx);
                            ryy = fmac(zl.rc[face][1][k], v, ryy);
                        } else {
                            rxx = fmac(zl.rc[face][0][k], eh ? T0[0][1] : T0[0][0], rxx);
                            ryy = fmac(zl.rc[face][1][k], eh ? T0[1][1] : T0[1][0], ryy);
                        }
                    }
                }
            }
            if (face == 0) {
#pragma unroll
                for (int k = 0; k < M; ++k) {
                    ghost[0][k] = k >= 1 ? fmac(zl.zc[0][0][k], src[0][0][k], 0.0f) : 0.0f;
                    ghost[1][k] = fmac(zl.zc[0][1][k], src[0][1][k], 0.0f);
                    ghost[2][k] = fmac(zl.zc[0][2][k], src[0][2][k], 0.0f);
                }
                if (zf.rec) {
                    out[0][0] = rxx;
                    out[1][0] = ryy;
                }
                out[2][0] = 0.0f;
            } else {
                const float tyz = fmac(zl.zc[1][1][0], src[1][1][0], 0.0f);
                const float txz = fmac(zl.zc[1][2][0], src[1][2][0], 0.0f);
#pragma unroll
                for (int k = 0; k < M; ++k) {
                    ghost[0][k] = k >= 1 ? fmac(zl.zc[1][0][k], src[1][0][k], 0.0f) : 0.0f;
                    ghost[1][k] = k >= 1 ? fmac(zl.zc[1][1][k], src[1][1][k], 0.0f) : 0.0f;
                    ghost[2][k] = k >= 1 ? fmac(zl.zc[1][2][k], src[1][2][k], 0.0f) : 0.0f;
                }
#pragma unroll
                for (int v = 0; v < 2; ++v)
                    if (v == eh) {
                        if (zf.rec) {
                            out[0][v] = rxx;
                            out[1][v] = ryy;
                        }
                        out[2][v] = 0.0f;
                        out[4][v] = tyz;
                        out[5][v] = txz;
                    }
            }
        }
        // z-halo pair of the neighbouring tile: Tzz, Tyz, Txz as its own lanes compute them
        R outi[3][2] = {};
        if (item && rowc) {
            // U(x-M+i) for i < 2M, W(x-M+1+i) for i < 2M from the ring
            R ui[2 * M][2], wi[2 * M][2];
#pragma unroll
            for (int i = 0; i < 2 * M; ++i) {
                const int su = back((M - i + L) % L), sw = back((M - 1 - i + L) % L);
                ui[i][0] = ring_i[(su * 2 + 0) * 2 + 0];
                ui[i][1] = ring_i[(su * 2 + 0) * 2 + 1];
                wi[i][0] = ring_i[(sw * 2 + 1) * 2 + 0];
                wi[i][1] = ring_i[(sw * 2 + 1) * 2 + 1];
            }
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                const int me = vmi + v;
                R s = 0, acc = T0i[0][v];
#pragma unroll
                for (int k = 0; k < M; ++k) s = fmac(c.cn[2][0][k], ui[M + k][v] - ui[M - 1 - k][v], s);
                acc = acc + s;
                s = 0;
#pragma unroll
                for (int k = 0; k < M; ++k) s = fmac(c.cn[2][1][k], vV[me + k * T::VW] - vV[me - (1 + k) * T::VW], s);
                acc = acc + s;
                s = 0;
#pragma unroll
                for (int k = 0; k < M; ++k) s = fmac(c.cn[2][2][k], vW[me + k] - vW[me - 1 - k], s);
                outi[0][v] = acc + s;
                R syz = 0, sxz = 0;
#pragma unroll
                for (int k = 0; k < M; ++k) syz = fmac(c.cs[1][0][k], vV[me + 1 + k] - vV[me - k], syz);
#pragma unroll
                for (int k = 0; k < M; ++k)
                    syz = fmac(c.cs[1][1][k], vW[me + (1 + k) * T::VW] - vW[me - k * T::VW], syz);
#pragma unroll
                for (int k = 0; k < M; ++k) sxz = fmac(c.cs[2][0][k], vU[me + 1 + k] - vU[me - k], sxz);
#pragma unroll
                for (int k = 0; k < M; ++k) sxz = fmac(c.cs[2][1][k], wi[M + k][v] - wi[M - 1 - k][v], sxz);
                outi[1][v] = T0i[1][v] + syz;
                outi[2][v] = T0i[2][v] + sxz;
            }
        }
        // stress plane x, rounded as the separate velocity launch reads it back
        if (act0) {
#pragma unroll
            for (int v = 0; v < 2; ++v)
                if (v == 0 || act1) {
                    const int sm = sme0 + v;
                    sXY[sm] = rnd(out[3][v]);
                    sYY[sm] = rnd(out[1][v]);
                    sYZ[sm] = rnd(out[4][v]);
                    if (own) {
                        sXZ[sm] = rnd(out[5][v]);
                        sZZ[sm] = rnd(out[2][v]);
                    }
                }
        }
        if (item && rowc) {
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                const int cz = zi + v - (z0 - M);
                if (cz >= 0 && cz < T::SW) {
                    sZZ[smi + v] = rnd(outi[0][v]);
                    sYZ[smi + v] = rnd(outi[1][v]);
                    sXZ[smi + v] = rnd(outi[2][v]);
                }
            }
        }
        if (zrow && zlane >= 0) {
            // ghost cells: below the low face (z0 = m) columns M - k (Tzz) and
            // M - 1 - k; above the high face columns ch + k
            const int ch = zrh + M;
#pragma unroll
            for (int k = 0; k < M; ++k) {
                if (zlane == 0) {
                    if (k >= 1) sZZ[w * T::SW + M - k] = rnd(ghost[0][k]);
                    sYZ[w * T::SW + M - 1 - k] = rnd(ghost[1][k]);
                    sXZ[w * T::SW + M - 1 - k] = rnd(ghost[2][k]);
                } else if (k >= 1) {
                    sZZ[w * T::SW + ch + k] = rnd(ghost[0][k]);
                    sYZ[w * T::SW + ch + k] = rnd(ghost[1][k]);
                    sXZ[w * T::SW + ch + k] = rnd(ghost[2][k]);
                }
            }
        }
        if (rowo && xo) {
            const int64_t row = (int64_t)x * g.sx + (int64_t)y * g.sy;
            if (zrow && zlane >= 0) {
                const int b = zlane == 0 ? M : g.nz - M - 1;
#pragma unroll
                for (int k = 0; k < M; ++k) {
                    if (zlane == 0) {
                        if (k >= 1) CS::st(P.s1[2] + row + b - k, ghost[0][k]);
                        CS::st(P.s1[4] + row + b - 1 - k, ghost[1][k]);
                        CS::st(P.s1[5] + row + b - 1 - k, ghost[2][k]);
                    } else if (k >= 1) {
                        CS::st(P.s1[2] + row + b + k, ghost[0][k]);
                        CS::st(P.s1[4] + row + b + k, ghost[1][k]);
                        CS::st(P.s1[5] + row + b + k, ghost[2][k]);
                    }
                }
            }
            if (act0) {
#pragma unroll
                for (int f = 0; f < 6; ++f) store_points(P.s1[f] + q, out[f][0], out[f][1], act0, act1);
            }
        }
        // own rows: Txx, Txy, Txz of plane x into the rings
        if (own && rowc) {
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                const int r = (slot * 2 + v) * NOWN + tw;
                rXX[r] = rnd(out[0][v]);
                rXY[r] = rnd(out[3][v]);
                rXZ[r] = rnd(out[5][v]);
            }
        }
        __syncthreads();
        // velocity: y and z differences of plane x kept; the sums of plane
        // x - M formed in the x, y, z order of the separate velocity sweep
        if (own && rowc) {
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                const int sm = sme0 + v;
                R* d = rD + (dslot * 2 + v) * ND * NOWN + tw;
#pragma unroll
                for (int k = 0; k < M; ++k) {
                    d[(0 * M + k) * NOWN] = sXY[sm + k * T::SW] - sXY[sm - (1 + k) * T::SW];
                    d[(1 * M + k) * NOWN] = sXZ[sm + k] - sXZ[sm - 1 - k];
                    d[(2 * M + k) * NOWN] = sYY[sm + (1 + k) * T::SW] - sYY[sm - k * T::SW];
                    d[(3 * M + k) * NOWN] = sYZ[sm + k] - sYZ[sm - 1 - k];
                    d[(4 * M + k) * NOWN] = sYZ[sm + k * T::SW] - sYZ[sm - (1 + k) * T::SW];
                    d[(5 * M + k) * NOWN] = sZZ[sm + 1 + k] - sZZ[sm - k];
                }
            }
            const int xv = x - M;
            if (rowo && xv >= xb && xv < xe) {
                R ou[2], ov[2], ow[2];
                const int ds = dslot >= M ? dslot - M : dslot - M + LD;   // plane x - M
#pragma unroll
                for (int v = 0; v < 2; ++v) {
                    const R* d = rD + (ds * 2 + v) * ND * NOWN + tw;
                    R su = 0, sv = 0, sw = 0;
                    // Txx(xv-M+1 .. xv+M) = planes x-2M+1 .. x; Txy, Txz(xv-M .. xv+M-1)
#pragma unroll
                    for (int k = 0; k < M; ++k) {
                        const R a = rXX[(back(M - 1 - k) * 2 + v) * NOWN + tw];
                        const R b = rXX[(back(M + k) * 2 + v) * NOWN + tw];
                        su = fmac(c.cv[0][0][k], a - b, su);
                    }
#pragma unroll
                    for (int k = 0; k < M; ++k) su = fmac(c.cv[0][1][k], d[(0 * M + k) * NOWN], su);
#pragma unroll
                    for (int k = 0; k < M; ++k) su = fmac(c.cv[0][2][k], d[(1 * M + k) * NOWN], su);
#pragma unroll
                    for (int k = 0; k < M; ++k) {
                        const R a = rXY[(back(M - k) * 2 + v) * NOWN + tw];
                        const R b = rXY[(back(M + 1 + k) * 2 + v) * NOWN + tw];
                        sv = fmac(c.cv[1][0][k], a - b, sv);
                    }
#pragma unroll
                    for (int k = 0; k < M; ++k) sv = fmac(c.cv[1][1][k], d[(2 * M + k) * NOWN], sv);
#pragma unroll
                    for (int k = 0; k < M; ++k) sv = fmac(c.cv[1][2][k], d[(3 * M + k) * NOWN], sv);
#pragma unroll
                    for (int k = 0; k < M; ++k) {
                        const R a = rXZ[(back(M - k) * 2 + v) * NOWN + tw];
                        const R b = rXZ[(back(M + 1 + k) * 2 + v) * NOWN + tw];
                        sw = fmac(c.cv[2][0][k], a - b, sw);
                    }
#pragma unroll
                    for (int k = 0; k < M; ++k) sw = fmac(c.cv[2][1][k], d[(4 * M + k) * NOWN], sw);
#pragma unroll
                    for (int k = 0; k < M; ++k) sw = fmac(c.cv[2][2][k], d[(5 * M + k) * NOWN], sw);
                    ou[v] = qU[v][0] + su;
                    ov[v] = qV[v][0] + sv;
                    ow[v] = qW[v][0] + sw;
                }
                const int64_t qv = (int64_t)xv * g.sx + col;
                store_points(P.v1[0] + qv, ou[0], ou[1], zin0, zin1);
                store_points(P.v1[1] + qv, ov[0], ov[1], zin0, zin1);
                store_points(P.v1[2] + qv, ow[0], ow[1], zin0, zin1);
            }
        }
    }
}

// ------------------------------------------------------------------ dispatch
struct Tiling {
    int ty, tz, tyn, tx, xchunk, nb, grid;
};

template <int M>
static Tiling tiling(const Geom& g, const InteriorLaunch& cfg, int ty, int tzw = 64) {
    Tiling t;
    const int ni[3] = {g.nx - 2 * M, g.ny - 2 * M, g.nz - 2 * M};
    t.ty = ty;
    t.tz = (ni[2] + tzw - 1) / tzw;
    t.tyn = (ni[1] + ty - 1) / ty;
    int xc = cfg.x_chunk;
    if (xc <= 0) {
        // about 512 blocks of 512 threads (256 of 1024): two per CU.  Long x
        // sweeps load fewer x halos and measured fastest at 256^3 (x chunks
        // 8..128 timed); 8..64 planes
        const int cols = t.tz * t.tyn, target = ty >= 16 ? 256 : 512;
        xc = (ni[0] * cols + target - 1) / target;
        xc = xc < 8 ? 8 : (xc > 64 ? 64 : xc);
    }
    t.xchunk = xc;
    t.tx = (ni[0] + xc - 1) / xc;
    t.nb = t.tz * t.tyn * t.tx;
    t.grid = ((t.nb + 7) / 8) * 8;
    return t;
}

// Tile counts of a box list for tiles of ty rows and 128 z points; x chunk
// xc0 (0: about 512 blocks over all boxes, 8..64 planes, as tiling()).
template <int M>
static Boxes fill_boxes(const Geom& g, Boxes B, int ty, int xc0, int& tz) {
    tz = (g.nz - 2 * M + 127) / 128;
    int cols = 0, len = 0;
    for (int i = 0; i < B.n; ++i) {
        B.tyn[i] = std::max(0, (B.y1[i] - B.y0[i] + ty - 1) / ty);
        cols += tz * B.tyn[i];
        len = std::max(len, B.x1[i] - B.x0[i]);
    }
    int xc = xc0;
    if (xc <= 0) {
        long long work = 0;
        for (int i = 0; i < B.n; ++i) work += (long long)tz * B.tyn[i] * std::max(0, B.x1[i] - B.x0[i]);
        xc = (int)((work + 511) / 512);
        xc = xc < 8 ? 8 : (xc > 64 ? 64 : xc);
    }
    B.first[0] = 0;
    for (int i = 0; i < B.n; ++i) {
        const int xl = std::max(0, B.x1[i] - B.x0[i]);
        B.xc[i] = std::max(1, std::min(xc, xl));
        const int tx = xl > 0 ? (xl + B.xc[i] - 1) / B.xc[i] : 0;
        B.first[i + 1] = B.first[i] + tz * B.tyn[i] * tx;
    }
    return B;
}

// Rows per tile of the single-point sweeps when the caller leaves it open: 16
// for FP32 on constant media (fewer y-halo rows; measured 2-3 % faster at
// 256^3), else 8 (the media streams and the 16-bit conversions favour the
// smaller tile).
template <typename R, typename S, bool HET>
static int default_tile_y() {
    return (!HET && sizeof(R) == 4 && sizeof(S) == 4) ? 16 : 8;
}

template <typename R, typename S, int M, bool HET, bool VEL>
static hipError_t run_interior(const Geom& g, void* fields, const void* media, const void* coefs, int t0, int t1,
                               const InteriorLaunch& cfg, hipStream_t s, const ZFaceArgs* zf, const void* pool,
                               const Boxes* boxes) {
    const FieldPtrs<S>& F = *static_cast<const FieldPtrs<S>*>(fields);
    const MediaPtrs<R>& md = *static_cast<const MediaPtrs<R>*>(media);
    const Coefs<R>& c = *static_cast<const Coefs<R>*>(coefs);
    // box lists run on the pair sweeps only
    if (boxes && (sizeof(R) != 4 || cfg.variant != 0)) return hipErrorInvalidValue;
    if (cfg.variant == 1) {
        const Tiling t = tiling<M>(g, cfg, 4);
        if (VEL)
            hipLaunchKernelGGL((velocity_gather<R, S, M, HET>), dim3(t.grid), dim3(256), 0, s, g, F, md, c, t0, t1,
                               t.tz, t.tyn, t.xchunk, t.nb);
        else
            hipLaunchKernelGGL((stress_gather<R, S, M, HET>), dim3(t.grid), dim3(256), 0, s, g, F, md, c, t0, t1,
                               t.tz, t.tyn, t.xchunk, t.nb);
        return hipGetLastError();
    }
    if constexpr (sizeof(R) == 4) {
        if (cfg.variant == 0) {
            // FP32 arithmetic: the pair sweeps (two z points per lane, 128-wide
            // z tiles).  Stress tiles have 4 or 8 rows (default 8), as
            // 1024-thread blocks would spill the doubled register queues; the
            // velocity kernel, with fewer of them, takes 16 rows at orders 2
            // and 4 (the default for FP32 storage: fewer y-halo rows, 2-5 %
            // faster at 256^3)
            int ty = cfg.tile_y == 4 ? 4 : 8;
            if constexpr (VEL && M <= 2)
                if (cfg.tile_y == 16 || (cfg.tile_y == 0 && sizeof(S) == 4)) ty = 16;
            Boxes B;
            int tz = 0;
            if (boxes) {
                B = fill_boxes<M>(g, *boxes, ty, cfg.x_chunk, tz);
            } else {
                const Tiling t = tiling<M>(g, cfg, ty, 128);
                B.n = 1;
                B.x0[0] = M;
                B.x1[0] = g.nx - M;
                B.y0[0] = M;
                B.y1[0] = g.ny - M;
                B.tyn[0] = t.tyn;
                B.xc[0] = t.xchunk;
                B.first[1] = t.nb;
                tz = t.tz;
            }
            const int nb = B.first[B.n];
            if (nb == 0) return hipSuccess;
            const int grid = ((nb + 7) / 8) * 8;
#define OPESCI_PAIR(TYV)                                                                                        \
    if (ty == TYV) {                                                                                            \
        if (VEL)                                                                                                \
            hipLaunchKernelGGL((velocity_sweep2<S, M, TYV, HET>), dim3(grid), dim3(64 * TYV), 0, s, g, F, md, c,   \
                               t0, t1, tz, nb, B);                                                              \
        else if (zf && zf->F > 0)                                                                               \
            hipLaunchKernelGGL((stress_sweep2<S, M, TYV, false, true>), dim3(grid), dim3(64 * TYV), 0, s, g, F,     \
                               md, c, t0, t1, tz, nb, B, *zf);                                                  \
        else                                                                                                    \
            hipLaunchKernelGGL((stress_sweep2<S, M, TYV, HET, false>), dim3(grid), dim3(64 * TYV), 0, s, g, F, md, \
                               c, t0, t1, tz, nb, B, ZFaceArgs());                                              \
        return hipGetLastError();                                                                               \
    }
            OPESCI_PAIR(4)
            OPESCI_PAIR(8)
#undef OPESCI_PAIR
            if constexpr (VEL && M <= 2) {
                if (ty == 16) {
                    hipLaunchKernelGGL((velocity_sweep2<S, M, 16, HET>), dim3(grid), dim3(1024), 0, s, g, F, md, c,
                                       t0, t1, tz, nb, B);
                    return hipGetLastError();
                }
            }
        }
    }
    // one point per lane: kernel_variant 2, and every FP64 sweep
    const int tile_y = cfg.tile_y ? cfg.tile_y : default_tile_y<R, S, HET>();
    if (tile_y == 16) {
        const Tiling t = tiling<M>(g, cfg, 16);
        if (VEL)
            hipLaunchKernelGGL((velocity_sweep<R, S, M, 16, HET>), dim3(t.grid), dim3(1024), 0, s, g, F, md, c, t0, t1,
                               t.tz, t.tyn, t.xchunk, t.nb);
        else
            hipLaunchKernelGGL((stress_sweep<R, S, M, 16, HET>), dim3(t.grid), dim3(1024), 0, s, g, F, md, c, t0, t1,
                               t.tz, t.tyn, t.xchunk, t.nb);
        return hipGetLastError();
    }
    if (tile_y == 4) {
        const Tiling t = tiling<M>(g, cfg, 4);
        if (VEL)
            hipLaunchKernelGGL((velocity_sweep<R, S, M, 4, HET>), dim3(t.grid), dim3(256), 0, s, g, F, md, c, t0, t1,
                               t.tz, t.tyn, t.xchunk, t.nb);
        else
            hipLaunchKernelGGL((stress_sweep<R, S, M, 4, HET>), dim3(t.grid), dim3(256), 0, s, g, F, md, c, t0, t1,
                               t.tz, t.tyn, t.xchunk, t.nb);
        return hipGetLastError();
    }
    const Tiling t = tiling<M>(g, cfg, 8);
    if (VEL)
        hipLaunchKernelGGL((velocity_sweep<R, S, M, 8, HET>), dim3(t.grid), dim3(512), 0, s, g, F, md, c, t0, t1,
                           t.tz, t.tyn, t.xchunk, t.nb);
    else
        hipLaunchKernelGGL((stress_sweep<R, S, M, 8, HET>), dim3(t.grid), dim3(512), 0, s, g, F, md, c, t0, t1,
                           t.tz, t.tyn, t.xchunk, t.nb);
    return hipGetLastError();
}

template <typename R, typename S, bool VEL>
static hipError_t dispatch_m(const Geom& g, bool het, void* fields, const void* media, const void* coefs, int t0,
                             int t1, const InteriorLaunch& cfg, hipStream_t s, const ZFaceArgs* zf,
                             const void* pool, const Boxes* boxes) {
#define OPESCI_CASE(MM)                                                                                          \
    case MM:                                                                                                     \
        return het ? run_interior<R, S, MM, true, VEL>(g, fields, media, coefs, t0, t1, cfg, s, zf, pool, boxes) \
                   : run_interior<R, S, MM, false, VEL>(g, fields, media, coefs, t0, t1, cfg, s, zf, pool, boxes);
    switch (g.m) {
        OPESCI_CASE(1)
        OPESCI_CASE(2)
        OPESCI_CASE(3)
        OPESCI_CASE(4)
        default:
            return hipErrorInvalidValue;
    }
#undef OPESCI_CASE
}

template <bool VEL>
static hipError_t dispatch(const Geom& g, int real_bytes, int storage, bool het, void* fields, const void* media,
                           const void* coefs, int t0, int t1, const InteriorLaunch& cfg, hipStream_t s,
                           const ZFaceArgs* zf, const void* pool, const Boxes* boxes) {
    if (real_bytes == 8)
        return dispatch_m<double, double, VEL>(g, het, fields, media, coefs, t0, t1, cfg, s, nullptr, nullptr, boxes);
    if (storage == OPESCI_STORE_F16)
        return dispatch_m<float, _Float16, VEL>(g, het, fields, media, coefs, t0, t1, cfg, s, zf, pool, boxes);
    if (storage == OPESCI_STORE_BF16)
        return dispatch_m<float, bf16_t, VEL>(g, het, fields, media, coefs, t0, t1, cfg, s, zf, pool, boxes);
    return dispatch_m<float, float, VEL>(g, het, fields, media, coefs, t0, t1, cfg, s, zf, pool, boxes);
}

// ------------------------------------------------------------------ fused step launch
#ifndef OPESCI_FUSED_TY
#define OPESCI_FUSED_TY 4
#endif
bool fused_step_builds(int m) { return m == 1 || m == 2; }

template <typename S, int M>
static hipError_t run_fused(const Geom& g, void* fields, const void* coefs, int t0, int t1, const ZFaceArgs& zf,
                            const Boxes& box, int x_chunk, hipStream_t s) {
    constexpr int TY = OPESCI_FUSED_TY;
    const FieldPtrs<S>& F = *static_cast<const FieldPtrs<S>*>(fields);
    const Coefs<float>& c = *static_cast<const Coefs<float>*>(coefs);
    StepPtrs<S> P;
    for (int f = 0; f < 3; ++f) {
        P.v0[f] = F.f[f][t0];
        P.v1[f] = F.f[f][t1];
    }
    for (int f = 0; f < 6; ++f) {
        P.s0[f] = F.f[3 + f][t0];
        P.s1[f] = F.f[3 + f][t1];
    }
    if (box.n != 1 || box.x0[0] < 2 * M || box.y0[0] < 2 * M || box.x1[0] > g.nx - 2 * M ||
        box.y1[0] > g.ny - 2 * M || !zf.coef || zf.F <= 0)
        return hipErrorInvalidValue;
    int xc = x_chunk;
    if (xc <= 0) {
        // one block per CU (the 12-wave blocks hold a CU's registers): split
        // x so that the block count is close to 256
        const int tz = (g.nz - 2 * M + 127) / 128, tyn = (box.y1[0] - box.y0[0] + TY - 1) / TY;
        const int cols = std::max(1, tz * tyn), xl = box.x1[0] - box.x0[0];
        const int tx = std::max(1, std::min(xl, (256 + cols / 2) / cols));
        xc = (xl + tx - 1) / tx;
    }
    int tz = 0;
    const Boxes B = fill_boxes<M>(g, box, TY, xc, tz);
    const int nb = B.first[B.n];
    if (nb == 0) return hipSuccess;
    const int grid = ((nb + 7) / 8) * 8;
    hipLaunchKernelGGL((step_sweep2<S, M, TY>), dim3(grid), dim3(64 * (TY + 2 * M)), 0, s, g, P, c, tz, nb, B, zf);
    return hipGetLastError();
}

template <typename S>
static hipError_t fused_m(const Geom& g, void* fields, const void* coefs, int t0, int t1, const ZFaceArgs& zf,
                          const Boxes& box, int x_chunk, hipStream_t s) {
    switch (g.m) {
        case 1:
            return run_fused<S, 1>(g, fields, coefs, t0, t1, zf, box, x_chunk, s);
        case 2:
            return run_fused<S, 2>(g, fields, coefs, t0, t1, zf, box, x_chunk, s);
        default:
            return hipErrorInvalidValue;
    }
}

hipError_t launch_step_fused(const Geom& g, int storage, void* fields, const void* coefs, int t0, int t1,
                             const ZFaceArgs& zf, const Boxes& box, int x_chunk, hipStream_t s) {
    if (storage == OPESCI_STORE_F16) return fused_m<_Float16>(g, fields, coefs, t0, t1, zf, box, x_chunk, s);
    if (storage == OPESCI_STORE_BF16) return fused_m<bf16_t>(g, fields, coefs, t0, t1, zf, box, x_chunk, s);
    return fused_m<float>(g, fields, coefs, t0, t1, zf, box, x_chunk, s);
}

bool stress_applies_zfaces(int real_bytes, bool het, const InteriorLaunch& cfg) {
    // the pair sweeps (FP32 arithmetic, variant 0) with their 4- or 8-row
    // stress tiles, constant media (on the read path the surface recompute
    // takes per-point coefficients, whose prefetch costs the sweep more
    // occupancy than the separate launch costs time)
    return real_bytes == 4 && !het && cfg.variant == 0;
}

hipError_t launch_stress(const Geom& g, int real_bytes, int storage, bool het, void* fields, const void* media,
                         const void* coefs, int t0, int t1, const InteriorLaunch& cfg, hipStream_t s,
                         const ZFaceArgs* zf, const void* pool, const Boxes* boxes) {
    if (zf && zf->F > 0 && !stress_applies_zfaces(real_bytes, het, cfg)) return hipErrorInvalidValue;
    return dispatch<false>(g, real_bytes, storage, het, fields, media, coefs, t0, t1, cfg, s, zf, pool, boxes);
}

hipError_t launch_velocity(const Geom& g, int real_bytes, int storage, bool het, void* fields, const void* media,
                           const void* coefs, int t0, int t1, const InteriorLaunch& cfg, hipStream_t s,
                           const Boxes* boxes) {
    return dispatch<true>(g, real_bytes, storage, het, fields, media, coefs, t0, t1, cfg, s, nullptr, nullptr, boxes);
}

}  // namespace opesci
