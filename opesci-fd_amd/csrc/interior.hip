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
// with m = 1 for constant media, and lambda(+2mu) / mu_ij / beta_i arrays in
// the read path.  Arithmetic in the compute type R (float or double); fields
// stored as S (float, double, _Float16, bf16).
//
// Two implementations:
//  * sweep (default): a block owns a 64(z) x TY(y) column tile and sweeps x.
//    Fields differentiated along x live in per-thread register queues, the
//    current x plane of fields differentiated along y or z is staged with its
//    halo in LDS, so every field is read from HBM once per sweep.  Blocks are
//    remapped so that the tiles of one XCD form a contiguous slab and share
//    halo lines in that XCD's L2.
//  * gather: one thread per (y,z) column, every stencil operand loaded from
//    global memory (caches only); kept as the simple reference form.
#include "common.h"

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

// ------------------------------------------------------------------ dispatch
template <typename R, typename S, int M, bool HET>
static hipError_t run_stress(const Geom& g, void* fields, const void* media, const void* coefs, int t0, int t1,
                             const InteriorLaunch& cfg, hipStream_t s) {
    const FieldPtrs<S>& F = *static_cast<const FieldPtrs<S>*>(fields);
    const MediaPtrs<R>& md = *static_cast<const MediaPtrs<R>*>(media);
    const Coefs<R>& c = *static_cast<const Coefs<R>*>(coefs);
    const int ni[3] = {g.nx - 2 * M, g.ny - 2 * M, g.nz - 2 * M};
    const int xchunk = cfg.x_chunk > 0 ? cfg.x_chunk : 16;
    const int tz = (ni[2] + 63) / 64, ty = (ni[1] + 3) / 4, tx = (ni[0] + xchunk - 1) / xchunk;
    const int nb = tz * ty * tx;
    const int grid = ((nb + 7) / 8) * 8;
    hipLaunchKernelGGL((stress_gather<R, S, M, HET>), dim3(grid), dim3(256), 0, s, g, F, md, c, t0, t1, tz, ty,
                       xchunk, nb);
    return hipGetLastError();
}

template <typename R, typename S, int M, bool HET>
static hipError_t run_velocity(const Geom& g, void* fields, const void* media, const void* coefs, int t0, int t1,
                               const InteriorLaunch& cfg, hipStream_t s) {
    const FieldPtrs<S>& F = *static_cast<const FieldPtrs<S>*>(fields);
    const MediaPtrs<R>& md = *static_cast<const MediaPtrs<R>*>(media);
    const Coefs<R>& c = *static_cast<const Coefs<R>*>(coefs);
    const int ni[3] = {g.nx - 2 * M, g.ny - 2 * M, g.nz - 2 * M};
    const int xchunk = cfg.x_chunk > 0 ? cfg.x_chunk : 16;
    const int tz = (ni[2] + 63) / 64, ty = (ni[1] + 3) / 4, tx = (ni[0] + xchunk - 1) / xchunk;
    const int nb = tz * ty * tx;
    const int grid = ((nb + 7) / 8) * 8;
    hipLaunchKernelGGL((velocity_gather<R, S, M, HET>), dim3(grid), dim3(256), 0, s, g, F, md, c, t0, t1, tz, ty,
                       xchunk, nb);
    return hipGetLastError();
}

template <typename R, typename S, bool VEL>
static hipError_t dispatch_m(const Geom& g, bool het, void* fields, const void* media, const void* coefs, int t0,
                             int t1, const InteriorLaunch& cfg, hipStream_t s) {
#define OPESCI_CASE(MM)                                                                            \
    case MM:                                                                                       \
        if (VEL)                                                                                   \
            return het ? run_velocity<R, S, MM, true>(g, fields, media, coefs, t0, t1, cfg, s)     \
                       : run_velocity<R, S, MM, false>(g, fields, media, coefs, t0, t1, cfg, s);   \
        return het ? run_stress<R, S, MM, true>(g, fields, media, coefs, t0, t1, cfg, s)           \
                   : run_stress<R, S, MM, false>(g, fields, media, coefs, t0, t1, cfg, s);
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
                           const void* coefs, int t0, int t1, const InteriorLaunch& cfg, hipStream_t s) {
    if (real_bytes == 8) return dispatch_m<double, double, VEL>(g, het, fields, media, coefs, t0, t1, cfg, s);
    if (storage == OPESCI_STORE_F16) return dispatch_m<float, _Float16, VEL>(g, het, fields, media, coefs, t0, t1, cfg, s);
    if (storage == OPESCI_STORE_BF16) return dispatch_m<float, bf16_t, VEL>(g, het, fields, media, coefs, t0, t1, cfg, s);
    return dispatch_m<float, float, VEL>(g, het, fields, media, coefs, t0, t1, cfg, s);
}

hipError_t launch_stress(const Geom& g, int real_bytes, int storage, bool het, void* fields, const void* media,
                         const void* coefs, int t0, int t1, const InteriorLaunch& cfg, hipStream_t s) {
    return dispatch<false>(g, real_bytes, storage, het, fields, media, coefs, t0, t1, cfg, s);
}

hipError_t launch_velocity(const Geom& g, int real_bytes, int storage, bool het, void* fields, const void* media,
                           const void* coefs, int t0, int t1, const InteriorLaunch& cfg, hipStream_t s) {
    return dispatch<true>(g, real_bytes, storage, het, fields, media, coefs, t0, t1, cfg, s);
}

}  // namespace opesci
