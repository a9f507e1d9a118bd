// Discrete adjoint of the velocity-stress leapfrog step (read path, native
// storage), for the gradient of a receiver misfit with respect to the media.
//
// One forward step (runtime.hip step_once) is: stress interior update, stress
// boundary launches, velocity interior update, velocity boundary launches,
// receivers, source.  Its adjoint runs the transposed pieces in reverse:
// receiver residual, velocity boundary levels (last first), velocity interior,
// stress boundary levels, stress interior.  Adjoint fields `bar` mirror the
// forward fields (both time levels, same padded layout); a forward write of a
// cell zeroes its adjoint after propagating it, so cells a step never writes
// carry their adjoint through unchanged.
//
// Interior transposes are gathers (no atomics): the adjoint of
//   T_ii(p) += m_ij(p) sum_k cn[i][j][k] (V_j(p + k e_j) - V_j(p - (1+k) e_j))
// collects, at each velocity cell q, the terms of every interior p whose
// stencil reads q; likewise for the shear and velocity updates (stencils in
// interior.hip).  Boundary statements dst = sum c_k src_k are transposed per
// point: a = bar[dst]; bar[dst] = 0; bar[src_k] += c_k a.  Points of a level
// may share a source cell (a level's sources are never its destinations), so
// the additions are gathered: each point saves a in a buffer, then each
// source cell sums its contributions in a fixed order (no atomics, results
// independent of scheduling).  Their forward operands are the state before
// the level, which an undo log restores level by level.
//
// Media gradients: lambda, mu (normal stresses), mu12/mu23/mu13 (shear),
// beta1..3 (velocities) from the interior updates; media-dependent boundary
// coefficients accumulate d misfit / d coefficient per point in `gpool`, mapped
// at the end to the media cells through the generator's derivative programs.
#include "bcdev.h"

namespace opesci {

template <typename R>
struct AdjPtrs {
    R* f[NF][2];
};

template <typename R>
struct GradPtrs {
    R* m[OPESCI_M_COUNT];   // media gradients, same layout as the media
};

__device__ __forceinline__ bool inside(const Geom& g, int x, int y, int z) {
    const int M = g.m;
    return x >= M && x < g.nx - M && y >= M && y < g.ny - M && z >= M && z < g.nz - M;
}

__device__ __forceinline__ void cell_of(const Geom& g, int64_t p, int& x, int& y, int& z) {
    z = (int)(p % g.nz);
    y = (int)((p / g.nz) % g.ny);
    x = (int)(p / ((int64_t)g.nz * g.ny));
}

// ------------------------------------------------------------------ stress interior
// bar V_j(t0) += transposed stencils of the stress update (reads bar T(t1)).
template <typename R, int M>
__global__ void adj_stress_to_v(Geom g, AdjPtrs<R> B, MediaPtrs<R> md, Coefs<R> c, int t0, int t1) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const R* lam = md.m[OPESCI_M_LAMBDA];
    const R* mu = md.m[OPESCI_M_MU];
    const R* ms[3] = {md.m[OPESCI_M_MU12], md.m[OPESCI_M_MU23], md.m[OPESCI_M_MU13]};
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < 3 * n; e += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(e / n);
        int x, y, z;
        cell_of(g, e % n, x, y, z);
        const int q[3] = {x, y, z};
        R acc = 0;
        // normal stresses i read V_j along j at +k and -(1+k)
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < M; ++k) {
                for (int sgn = 0; sgn < 2; ++sgn) {
                    int p[3] = {q[0], q[1], q[2]};
                    p[j] += sgn == 0 ? -k : 1 + k;
                    if (!inside(g, p[0], p[1], p[2])) continue;
                    const int64_t a = g.at(p[0], p[1], p[2]);
                    const R m = i == j ? lam[a] + (R)2 * mu[a] : lam[a];
                    const R v = m * c.cn[i][j][k] * B.f[3 + i][t1][a];
                    acc += sgn == 0 ? v : -v;
                }
            }
        // shear: (shear index, axis, coefficient set) pairs that read V_j,
        // each at +(1+k) and -k along its axis
        int sh[2], ax[2], cc[2];
        if (j == 0) {
            sh[0] = 0; ax[0] = 1; cc[0] = 0;   // Txy, dU/dy
            sh[1] = 2; ax[1] = 2; cc[1] = 0;   // Txz, dU/dz
        } else if (j == 1) {
            sh[0] = 0; ax[0] = 0; cc[0] = 1;   // Txy, dV/dx
            sh[1] = 1; ax[1] = 2; cc[1] = 0;   // Tyz, dV/dz
        } else {
            sh[0] = 1; ax[0] = 1; cc[0] = 1;   // Tyz, dW/dy
            sh[1] = 2; ax[1] = 0; cc[1] = 1;   // Txz, dW/dx
        }
        for (int u = 0; u < 2; ++u)
            for (int k = 0; k < M; ++k)
                for (int sgn = 0; sgn < 2; ++sgn) {
                    int p[3] = {q[0], q[1], q[2]};
                    p[ax[u]] += sgn == 0 ? -(1 + k) : k;
                    if (!inside(g, p[0], p[1], p[2])) continue;
                    const int64_t a = g.at(p[0], p[1], p[2]);
                    const R v = ms[sh[u]][a] * c.cs[sh[u]][cc[u]][k] * B.f[6 + sh[u]][t1][a];
                    acc += sgn == 0 ? v : -v;
                }
        B.f[j][t0][g.at(x, y, z)] += acc;
    }
}

// bar T(t0) += bar T(t1), bar T(t1) = 0 on the interior; media gradients of
// the stress update from the forward velocities at t0.
template <typename R, int M>
__global__ void adj_stress_point(Geom g, FieldPtrs<R> F, AdjPtrs<R> B, GradPtrs<R> gm, Coefs<R> c, int t0, int t1) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int64_t st[3] = {g.sx, g.sy, 1};
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int x, y, z;
        cell_of(g, e, x, y, z);
        if (!inside(g, x, y, z)) continue;
        const int64_t a = g.at(x, y, z);
        R bt[6];
        for (int f = 0; f < 6; ++f) {
            bt[f] = B.f[3 + f][t1][a];
            B.f[3 + f][t0][a] += bt[f];
            B.f[3 + f][t1][a] = 0;
        }
        // normal: T_ii += (i == j ? lambda + 2 mu : lambda) Sn[i][j]
        R glam = 0, gmu = 0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                R s = 0;
                for (int k = 0; k < M; ++k)
                    s += c.cn[i][j][k] * (F.f[j][t0][a + k * st[j]] - F.f[j][t0][a - (1 + k) * st[j]]);
                glam += bt[i] * s;
                if (i == j) gmu += (R)2 * bt[i] * s;
            }
        gm.m[OPESCI_M_LAMBDA][a] += glam;
        gm.m[OPESCI_M_MU][a] += gmu;
        // shear s: mu_s (sum_k cs[s][0][k] (A(+(1+k)) - A(-k)) along ax0 + cs[s][1] B along ax1)
        const int fa[3] = {0, 1, 0}, fb[3] = {1, 2, 2};   // differentiated fields
        const int aa[3] = {1, 2, 2}, ab[3] = {0, 1, 0};   // their axes
        const int mid[3] = {OPESCI_M_MU12, OPESCI_M_MU23, OPESCI_M_MU13};
        for (int s = 0; s < 3; ++s) {
            R v = 0;
            for (int k = 0; k < M; ++k) {
                v += c.cs[s][0][k] * (F.f[fa[s]][t0][a + (1 + k) * st[aa[s]]] - F.f[fa[s]][t0][a - k * st[aa[s]]]);
                v += c.cs[s][1][k] * (F.f[fb[s]][t0][a + (1 + k) * st[ab[s]]] - F.f[fb[s]][t0][a - k * st[ab[s]]]);
            }
            gm.m[mid[s]][a] += bt[3 + s] * v;
        }
    }
}

// ------------------------------------------------------------------ velocity interior
// bar T(t1) += transposed stencils of the velocity update (reads bar V(t1)).
template <typename R, int M>
__global__ void adj_vel_to_t(Geom g, AdjPtrs<R> B, MediaPtrs<R> md, Coefs<R> c, int t1) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const R* beta[3] = {md.m[OPESCI_M_BETA1], md.m[OPESCI_M_BETA2], md.m[OPESCI_M_BETA3]};
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < 6 * n; e += (int64_t)gridDim.x * blockDim.x) {
        const int f = (int)(e / n);   // stress 0..5 = Txx, Tyy, Tzz, Txy, Tyz, Txz
        int x, y, z;
        cell_of(g, e % n, x, y, z);
        const int q[3] = {x, y, z};
        // (velocity i, axis j) pairs reading this stress; a = 1 for T_ii along i
        int vi[2], vj[2], cnt;
        if (f < 3) {
            vi[0] = f; vj[0] = f; cnt = 1;
        } else if (f == 3) {
            vi[0] = 0; vj[0] = 1; vi[1] = 1; vj[1] = 0; cnt = 2;
        } else if (f == 4) {
            vi[0] = 1; vj[0] = 2; vi[1] = 2; vj[1] = 1; cnt = 2;
        } else {
            vi[0] = 0; vj[0] = 2; vi[1] = 2; vj[1] = 0; cnt = 2;
        }
        R acc = 0;
        for (int u = 0; u < cnt; ++u) {
            const int i = vi[u], j = vj[u], off = (f < 3) ? 1 : 0;
            for (int k = 0; k < M; ++k)
                for (int sgn = 0; sgn < 2; ++sgn) {
                    // V_i(p) reads T(p + (off + k) e_j) (+) and T(p + (off - 1 - k) e_j) (-)
                    int p[3] = {q[0], q[1], q[2]};
                    p[j] -= sgn == 0 ? off + k : off - 1 - k;
                    if (!inside(g, p[0], p[1], p[2])) continue;
                    const int64_t a = g.at(p[0], p[1], p[2]);
                    const R v = beta[i][a] * c.cv[i][j][k] * B.f[i][t1][a];
                    acc += sgn == 0 ? v : -v;
                }
        }
        B.f[3 + f][t1][g.at(x, y, z)] += acc;
    }
}

// bar V(t0) += bar V(t1), bar V(t1) = 0 on the interior; beta gradients from
// the forward stresses at t1.
template <typename R, int M>
__global__ void adj_vel_point(Geom g, FieldPtrs<R> F, AdjPtrs<R> B, GradPtrs<R> gm, Coefs<R> c, int t0, int t1) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const int64_t st[3] = {g.sx, g.sy, 1};
    const int mid[3] = {OPESCI_M_BETA1, OPESCI_M_BETA2, OPESCI_M_BETA3};
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int x, y, z;
        cell_of(g, e, x, y, z);
        if (!inside(g, x, y, z)) continue;
        const int64_t a = g.at(x, y, z);
        for (int i = 0; i < 3; ++i) {
            const R b = B.f[i][t1][a];
            B.f[i][t0][a] += b;
            B.f[i][t1][a] = 0;
            R s = 0;
            for (int j = 0; j < 3; ++j) {
                const int fld = j == i ? 3 + i : 6 + ((i + j == 1) ? 0 : (i + j == 3 ? 1 : 2));
                const int off = j == i ? 1 : 0;
                for (int k = 0; k < M; ++k)
                    s += c.cv[i][j][k] *
                         (F.f[fld][t1][a + (off + k) * st[j]] - F.f[fld][t1][a + (off - 1 - k) * st[j]]);
            }
            gm.m[mid[i]][a] += b * s;
        }
    }
}

// ------------------------------------------------------------------ fast interior transposes
// The transposed interior updates have the forward stencils' shapes:
//   bar T(t1) += stress stencil of Q_i = beta_i bar V_i(t1)
//                (cn'[i][i] = -cv[i][i], cs'[s] = -cv of the two velocities
//                 that read T_s)
//   bar V(t0) += velocity stencil of P_j = lambda (bar Txx + bar Tyy + bar Tzz)
//                + 2 mu bar T_jj and W_s = mu_s bar T_s
//                (cv'[j][j] = -cn[.][j], cv'[i][j] = -cs of the stress term
//                 that reads V_i along j)
// The P form needs cn[i][j][k] independent of i, which the launcher checks.
// A point kernel forms Q (or P, W) in scratch arrays that are zero outside
// the interior, moves bar t1 into bar t0 and accumulates the media gradients;
// the forward sweep kernels (homogeneous form) then update interior targets
// in place, and a rim kernel gathers the targets in the M-thick face slabs
// (halo cells that interior stencils read along one axis).
template <typename R>
struct Scratch {
    R* a[6];
};

__device__ __forceinline__ bool interior_cell(const Geom& g, int& x, int& y, int& z) {
    const int M = g.m;
    z = M + (int)blockIdx.x * 64 + (int)(threadIdx.x & 63);
    y = M + (int)blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    x = M + (int)blockIdx.z;
    return z < g.nz - M && y < g.ny - M;
}

// Q_i = beta_i bar V_i(t1); beta gradients from the forward stresses at t1.
// The identity part of the update's transpose (bar V(t0) += bar V(t1),
// bar V(t1) = 0) is a swap of the two levels' arrays on the host plus
// adj_move_cells (see adj_swap_levels).
template <typename R, int M>
__global__ __launch_bounds__(256) void adj_vel_prep(Geom g, FieldPtrs<R> F, AdjPtrs<R> B, MediaPtrs<R> md,
                                                    GradPtrs<R> gm, Coefs<R> c, Scratch<R> S, int t0, int t1) {
    int x, y, z;
    if (!interior_cell(g, x, y, z)) return;
    const int64_t a = g.at(x, y, z);
    const int64_t st[3] = {g.sx, g.sy, 1};
    const int src[3][3] = {{3, 6, 8}, {6, 4, 7}, {8, 7, 5}};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        // own-cell streams are touched once: non-temporal (forward-field
        // stencil reads stay cached)
        const R b = __builtin_nontemporal_load(B.f[i][t1] + a);
        __builtin_nontemporal_store(__builtin_nontemporal_load(md.m[OPESCI_M_BETA1 + i] + a) * b, S.a[i] + a);
        R s = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int o = i == j ? 1 : 0;
            const R* t = F.f[src[i][j]][t1] + a;
#pragma unroll
            for (int k = 0; k < M; ++k) s = fma(c.cv[i][j][k], t[(o + k) * st[j]] - t[(o - 1 - k) * st[j]], s);
        }
        R* gp = gm.m[OPESCI_M_BETA1 + i] + a;
        __builtin_nontemporal_store(__builtin_nontemporal_load(gp) + b * s, gp);
    }
}

// P_j, W_s from bar T(t1); lambda, mu, mu12/mu23/mu13 gradients from the
// forward velocities at t0.  The identity part (bar T(t0) += bar T(t1),
// bar T(t1) = 0) is done by adj_swap_levels.
template <typename R, int M>
__global__ __launch_bounds__(256) void adj_stress_prep(Geom g, FieldPtrs<R> F, AdjPtrs<R> B, MediaPtrs<R> md,
                                                       GradPtrs<R> gm, Coefs<R> c, Scratch<R> S, int t0, int t1) {
    int x, y, z;
    if (!interior_cell(g, x, y, z)) return;
    const int64_t a = g.at(x, y, z);
    const int64_t st[3] = {g.sx, g.sy, 1};
    // own-cell streams are touched once: non-temporal (forward-field stencil
    // reads stay cached)
    R bt[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) bt[f] = __builtin_nontemporal_load(B.f[3 + f][t1] + a);
    const R lam = __builtin_nontemporal_load(md.m[OPESCI_M_LAMBDA] + a);
    const R mu = __builtin_nontemporal_load(md.m[OPESCI_M_MU] + a);
    const R tr = bt[0] + bt[1] + bt[2];
    const int mid[3] = {OPESCI_M_MU12, OPESCI_M_MU23, OPESCI_M_MU13};
#pragma unroll
    for (int j = 0; j < 3; ++j) __builtin_nontemporal_store(lam * tr + (R)2 * mu * bt[j], S.a[j] + a);
#pragma unroll
    for (int s = 0; s < 3; ++s)
        __builtin_nontemporal_store(__builtin_nontemporal_load(md.m[mid[s]] + a) * bt[3 + s], S.a[3 + s] + a);
    R glam = 0, gmu = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const R* v = F.f[j][t0] + a;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            R d = 0;
#pragma unroll
            for (int k = 0; k < M; ++k) d = fma(c.cn[i][j][k], v[k * st[j]] - v[-(1 + k) * st[j]], d);
            glam += bt[i] * d;
            if (i == j) gmu += (R)2 * bt[i] * d;
        }
    }
    R* gl = gm.m[OPESCI_M_LAMBDA] + a;
    R* gu = gm.m[OPESCI_M_MU] + a;
    __builtin_nontemporal_store(__builtin_nontemporal_load(gl) + glam, gl);
    __builtin_nontemporal_store(__builtin_nontemporal_load(gu) + gmu, gu);
    const int fa[3] = {0, 1, 0}, fb[3] = {1, 2, 2};   // differentiated velocities
    const int aa[3] = {1, 2, 2}, ab[3] = {0, 1, 0};   // their axes
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const R* va = F.f[fa[s]][t0] + a;
        const R* vb = F.f[fb[s]][t0] + a;
        R v = 0;
#pragma unroll
        for (int k = 0; k < M; ++k) {
            v = fma(c.cs[s][0][k], va[(1 + k) * st[aa[s]]] - va[-k * st[aa[s]]], v);
            v = fma(c.cs[s][1][k], vb[(1 + k) * st[ab[s]]] - vb[-k * st[ab[s]]], v);
        }
        R* gs = gm.m[mid[s]] + a;
        __builtin_nontemporal_store(__builtin_nontemporal_load(gs) + bt[3 + s] * v, gs);
    }
}

// Face-slab targets: slab blockIdx.y = 2 * axis + side, cells with that
// coordinate in the M halo planes of the side and the other two interior.
__device__ __forceinline__ bool rim_cell(const Geom& g, int& x, int& y, int& z) {
    const int M = g.m;
    const int sl = (int)blockIdx.y, ax = sl >> 1, side = sl & 1;
    const int n[3] = {g.nx, g.ny, g.nz};
    const int b = ax == 0 ? 1 : 0, cc = ax == 2 ? 1 : 2;
    const int nb = n[b] - 2 * M, nc = n[cc] - 2 * M;
    const int e = (int)blockIdx.x * (int)blockDim.x + (int)threadIdx.x;
    if (e >= M * nb * nc) return false;
    const int d = e / (nb * nc), r = e % (nb * nc);
    int q[3];
    q[ax] = side ? n[ax] - M + d : d;
    q[b] = M + r / nc;
    q[cc] = M + r % nc;
    x = q[0];
    y = q[1];
    z = q[2];
    return true;
}

// bar T(t1) at face-slab cells from Q (velocity-update transpose).
template <typename R, int M>
__global__ __launch_bounds__(256) void adj_rim_t(Geom g, AdjPtrs<R> B, Scratch<R> S, Coefs<R> c, int t1) {
    int x, y, z;
    if (!rim_cell(g, x, y, z)) return;
    const int q[3] = {x, y, z};
    const int64_t at = g.at(x, y, z);
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        // (velocity i, axis j) pairs reading stress f
        int vi[2], vj[2], cnt;
        if (f < 3) {
            vi[0] = f; vj[0] = f; cnt = 1;
        } else if (f == 3) {
            vi[0] = 0; vj[0] = 1; vi[1] = 1; vj[1] = 0; cnt = 2;
        } else if (f == 4) {
            vi[0] = 1; vj[0] = 2; vi[1] = 2; vj[1] = 1; cnt = 2;
        } else {
            vi[0] = 0; vj[0] = 2; vi[1] = 2; vj[1] = 0; cnt = 2;
        }
        const int off = f < 3 ? 1 : 0;
        R acc = 0;
        for (int u = 0; u < cnt; ++u) {
            const int i = vi[u], j = vj[u];
            for (int k = 0; k < M; ++k)
                for (int sgn = 0; sgn < 2; ++sgn) {
                    int p[3] = {q[0], q[1], q[2]};
                    p[j] -= sgn == 0 ? off + k : off - 1 - k;
                    if (!inside(g, p[0], p[1], p[2])) continue;
                    const R v = c.cv[i][j][k] * S.a[i][g.at(p[0], p[1], p[2])];
                    acc += sgn == 0 ? v : -v;
                }
        }
        B.f[3 + f][t1][at] += acc;
    }
}

// bar V(t0) at face-slab cells from P, W (stress-update transpose).
template <typename R, int M>
__global__ __launch_bounds__(256) void adj_rim_v(Geom g, AdjPtrs<R> B, Scratch<R> S, Coefs<R> c, int t0) {
    int x, y, z;
    if (!rim_cell(g, x, y, z)) return;
    const int q[3] = {x, y, z};
    const int64_t at = g.at(x, y, z);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        R acc = 0;
        for (int k = 0; k < M; ++k)
            for (int sgn = 0; sgn < 2; ++sgn) {
                int p[3] = {q[0], q[1], q[2]};
                p[j] += sgn == 0 ? -k : 1 + k;
                if (!inside(g, p[0], p[1], p[2])) continue;
                const R v = c.cn[0][j][k] * S.a[j][g.at(p[0], p[1], p[2])];
                acc += sgn == 0 ? v : -v;
            }
        // (shear slot, axis, coefficient set) of the stress terms reading V_j
        int sh[2], ax[2], cc[2];
        if (j == 0) {
            sh[0] = 0; ax[0] = 1; cc[0] = 0;
            sh[1] = 2; ax[1] = 2; cc[1] = 0;
        } else if (j == 1) {
            sh[0] = 0; ax[0] = 0; cc[0] = 1;
            sh[1] = 1; ax[1] = 2; cc[1] = 0;
        } else {
            sh[0] = 1; ax[0] = 1; cc[0] = 1;
            sh[1] = 2; ax[1] = 0; cc[1] = 1;
        }
        for (int u = 0; u < 2; ++u)
            for (int k = 0; k < M; ++k)
                for (int sgn = 0; sgn < 2; ++sgn) {
                    int p[3] = {q[0], q[1], q[2]};
                    p[ax[u]] += sgn == 0 ? -(1 + k) : k;
                    if (!inside(g, p[0], p[1], p[2])) continue;
                    const R v = c.cs[sh[u]][cc[u]][k] * S.a[3 + sh[u]][g.at(p[0], p[1], p[2])];
                    acc += sgn == 0 ? v : -v;
                }
        B.f[j][t0][at] += acc;
    }
}

// ------------------------------------------------------------------ boundary levels
// Forward statement with the destination's previous value logged
// (log[(descriptor - first) * maxpts + point]).
template <typename R>
__global__ __launch_bounds__(256) void bc_log_kernel(Geom g, FieldPtrs<R> F, const BcDesc* __restrict__ desc,
                                                     const R* __restrict__ pool, int first, int maxpts, R* log,
                                                     int t0, int t1) {
    const BcDesc* D = desc + first + blockIdx.y;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D->npts) return;
    const int n1 = D->n1, i0 = D->lo0 + p / n1, i1 = D->lo1 + p % n1;
    const int o0 = D->axis == 0 ? 1 : 0, o1 = D->axis == 2 ? 1 : 2;
    const int64_t so0 = o0 == 0 ? g.sx : g.sy, so1 = o1 == 1 ? g.sy : 1;
    const int64_t dst = (int64_t)i0 * so0 + (int64_t)i1 * so1 + D->disp;
    const int fl = D->fl;
    log[(int64_t)blockIdx.y * maxpts + p] = F.f[fl >> 1][(fl & 1) ? t1 : t0][dst];
    bc_point<R, R>(g, F, D, pool, p, i0, i1, t0, t1);
}

// Transposed statements of one level, first pass: each point's adjoint
// value a = bar[dst] into the level's buffer, bar[dst] = 0, and the
// coefficient gradients into gpool (one slot per point and term); then the
// forward value the level overwrote at dst is restored from the undo log
// (the pass reads forward values at the level's sources only, which no
// statement of the level writes).
template <typename R>
__global__ __launch_bounds__(256) void adj_bc_kernel(Geom g, FieldPtrs<R> F, AdjPtrs<R> B,
                                                     const BcDesc* __restrict__ desc, const R* __restrict__ pool,
                                                     R* gpool, int first, const int32_t* __restrict__ boff, R* buf,
                                                     const R* __restrict__ log, int maxpts, int t0, int t1) {
    const int d = first + blockIdx.y;
    const BcDesc* D = desc + d;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D->npts) return;
    const int n1 = D->n1, i0 = D->lo0 + p / n1, i1 = D->lo1 + p % n1;
    const int o0 = D->axis == 0 ? 1 : 0, o1 = D->axis == 2 ? 1 : 2;
    const int64_t so0 = o0 == 0 ? g.sx : g.sy, so1 = o1 == 1 ? g.sy : 1;
    const int64_t dst = (int64_t)i0 * so0 + (int64_t)i1 * so1 + D->disp;
    const int fl = D->fl;
    R* bd = B.f[fl >> 1][(fl & 1) ? t1 : t0] + dst;
    const R a = *bd;
    *bd = 0;
    buf[boff[d] + p] = a;
    if (a != (R)0)
        for (int k = 0; k < D->nterms; ++k)
            if (D->t[k].cslot >= 0) {
                const int sfl = D->t[k].fl;
                const int64_t ci = D->cpool + (int64_t)D->t[k].cslot * D->npts + p;
                gpool[ci] += a * F.f[sfl >> 1][(sfl & 1) ? t1 : t0][dst + D->t[k].disp];
            }
    F.f[fl >> 1][(fl & 1) ? t1 : t0][dst] = log[(int64_t)blockIdx.y * maxpts + p];
}

// Second pass: bar[src] += c_k a over the source cell's contributions, in
// (statement, point, term) order.
template <typename R>
__global__ __launch_bounds__(256) void adj_bc_gather_kernel(AdjPtrs<R> B, const BcDesc* __restrict__ desc,
                                                            const R* __restrict__ pool, AdjGather ga, int t0,
                                                            int t1) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ga.ntgt) return;
    const R* buf = static_cast<const R*>(ga.buf);
    const int fl = ga.key[t];
    R* q = B.f[fl >> 1][(fl & 1) ? t1 : t0] + ga.off[t];
    R acc = *q;
    for (int e = ga.start[t]; e < ga.start[t + 1]; ++e) {
        const int ci = ga.cinfo[e];
        const R coef = ci >= 0 ? pool[ci] : (R)desc[(-1 - ci) / BC_MAXT].t[(-1 - ci) % BC_MAXT].coef;
        acc += coef * buf[ga.src[e]];
    }
    *q = acc;
}

// gpool -> media gradients through one derivative program: term k of
// descriptor `d`, program q, at every point of the statement.  Distinct
// points reach distinct media cells (the cell is the point's anchor plus the
// program's fixed offset), so one launch has no conflicting updates; the
// launches run one after another in (statement, term, program) order, which
// fixes the order of every cell's additions.
template <typename R>
__global__ __launch_bounds__(256) void adj_coef_media_kernel(Geom g, MediaPtrs<R> md, GradPtrs<R> gm,
                                                             const BcDesc* __restrict__ desc,
                                                             const OpesciExprOp* __restrict__ code,
                                                             const OpesciBcDTerm* __restrict__ dterms,
                                                             const R* __restrict__ gpool, int d, int k, int q) {
    const BcDesc* D = desc + d;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D->npts) return;
    const R gc = gpool[D->cpool + (int64_t)D->t[k].cslot * D->npts + p];
    if (gc == (R)0) return;
    const int n1 = D->n1, axis = D->axis;
    const int i0 = D->lo0 + p / n1, i1 = D->lo1 + p % n1;
    const int ax = axis == 0 ? 0 : i0;
    const int ay = axis == 0 ? i0 : (axis == 1 ? 0 : i1);
    const int az = axis == 2 ? 0 : i1;
    const double c0[3] = {0, 0, 0};
    const OpesciBcDTerm& T = dterms[q];
    const double dv = (double)D->t[k].coef * vm_eval<R>(code + T.prog, T.prog_len, c0, 0.0, md, g, ax, ay, az);
    R* out = gm.m[T.media] + g.at(ax + T.off[0], ay + T.off[1], az + T.off[2]);
    *out += (R)(gc * (R)dv);
}

// Receiver misfit residual: bar U,V,W(t1) += field - observed, one thread
// per receiver cell summing its receivers in receiver order.
template <typename R>
__global__ void adj_receiver_kernel(FieldPtrs<R> F, AdjPtrs<R> B, int t1, int ngroups,
                                    const int32_t* __restrict__ start, const int64_t* __restrict__ cell,
                                    const int32_t* __restrict__ order, const float* __restrict__ obs) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= ngroups) return;
    const int64_t q = cell[gi];
    for (int c = 0; c < 3; ++c) {
        R acc = B.f[c][t1][q];
        const R v = F.f[c][t1][q];
        for (int e = start[gi]; e < start[gi + 1]; ++e) acc += v - (R)obs[3 * order[e] + c];
        B.f[c][t1][q] = acc;
    }
}

__global__ void set_step_kernel(int* step, int value) { *step = value; }

// Txx, Tyy, Tzz of one cell at level t: read into buf (mode 0) or written
// from it (mode 1); brackets the point source in the reverse sweep.
template <typename R>
__global__ void src_cells_kernel(FieldPtrs<R> F, int t, int64_t q, R* buf, int mode) {
    const int f = (int)threadIdx.x;
    if (f >= 3) return;
    if (mode == 0)
        buf[f] = F.f[3 + f][t][q];
    else
        F.f[3 + f][t][q] = buf[f];
}

// ------------------------------------------------------------------ launchers
static dim3 grid_for(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return dim3((unsigned)std::min<int64_t>(b, 65536));
}

// cn[i][j][k] independent of i: the P form of the stress transpose applies
template <typename R>
static bool normal_coefs_shared(const Coefs<R>& c, int M) {
    for (int j = 0; j < 3; ++j)
        for (int k = 0; k < M; ++k)
            if (c.cn[1][j][k] != c.cn[0][j][k] || c.cn[2][j][k] != c.cn[0][j][k]) return false;
    return true;
}

static dim3 interior_grid(const Geom& g) {
    const int M = g.m;
    return dim3((unsigned)((g.nz - 2 * M + 63) / 64), (unsigned)((g.ny - 2 * M + 3) / 4), (unsigned)(g.nx - 2 * M));
}

static dim3 rim_grid(const Geom& g) {
    const int M = g.m;
    const int ni[3] = {g.nx - 2 * M, g.ny - 2 * M, g.nz - 2 * M};
    const int slab = M * std::max(ni[1] * ni[2], std::max(ni[0] * ni[2], ni[0] * ni[1]));
    return dim3((unsigned)((slab + 255) / 256), 6);
}

// After the two levels' arrays swapped places: for a list of distinct cells,
// each entry (offset << 5) | exchange << 4 | field, either
// bar F(t0)[c] += bar F(t1)[c], bar F(t1)[c] = 0 (interior cells read at
// level 0) or an exchange of the two values (non-interior cells, which keep
// their values in their physical arrays).
template <typename R>
__global__ __launch_bounds__(256) void adj_move_cells(AdjPtrs<R> B, const int64_t* __restrict__ cells, int n, int t0,
                                                      int t1) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t e = cells[i];
    const int f = (int)(e & 15);
    const int64_t off = e >> 5;
    R* a = B.f[f][t0] + off;
    R* b = B.f[f][t1] + off;
    const R va = *a, vb = *b;
    if (e & 16) {
        *a = vb;
        *b = va;
    } else {
        *a = va + vb;
        *b = (R)0;
    }
}

// The identity part of an update's transpose, bar F(t0) += bar F(t1) over
// the interior and bar F(t1) = 0 there, for the fields [f0, f1), without
// touching the arrays: at this point bar F(t0) is zero in the interior
// except at the cells that boundary statements of the step read at level
// 0 (their transposes ran first).  So the two levels' arrays swap places in
// the host-side pointer image; those cells move, and the non-interior
// cells that boundary statements read take back their physical arrays'
// values (the remaining non-interior cells are never read as adjoint
// values: a written cell's is consumed and zeroed by its statement's
// transpose before this point, an unwritten cell's is never used).  Every
// later launch, of this step and of the steps before it, reads the swapped
// image.
template <typename R>
static hipError_t adj_swap_levels(AdjPtrs<R>& B, int f0, int f1, const AdjCells& cells, int t0, int t1,
                                  hipStream_t s) {
    for (int f = f0; f < f1; ++f) std::swap(B.f[f][t0], B.f[f][t1]);
    if (cells.n > 0)
        hipLaunchKernelGGL((adj_move_cells<R>), dim3((unsigned)((cells.n + 255) / 256)), dim3(256), 0, s, B,
                           cells.d, cells.n, t0, t1);
    return hipGetLastError();
}

template <typename R, int M>
static hipError_t adj_stress_t(const Geom& g, void* fields, void* bar, const void* media, void* grad,
                               const void* coefs, void* const* scratch, const AdjCells& cells,
                               const InteriorLaunch& cfg, int t0, int t1, hipStream_t s) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const FieldPtrs<R>& F = *static_cast<const FieldPtrs<R>*>(fields);
    AdjPtrs<R>& B = *static_cast<AdjPtrs<R>*>(bar);
    const MediaPtrs<R>& md = *static_cast<const MediaPtrs<R>*>(media);
    const GradPtrs<R>& gm = *static_cast<const GradPtrs<R>*>(grad);
    const Coefs<R>& c = *static_cast<const Coefs<R>*>(coefs);
    if (!scratch || cfg.variant == 1 || !normal_coefs_shared(c, M)) {
        hipLaunchKernelGGL((adj_stress_to_v<R, M>), grid_for(3 * n), dim3(256), 0, s, g, B, md, c, t0, t1);
        hipLaunchKernelGGL((adj_stress_point<R, M>), grid_for(n), dim3(256), 0, s, g, F, B, gm, c, t0, t1);
        return hipGetLastError();
    }
    Scratch<R> S;
    for (int i = 0; i < 6; ++i) S.a[i] = static_cast<R*>(scratch[i]);
    // the weighted stresses must be read from bar T(t1) before the rim and
    // sweep kernels touch bar V(t0)
    hipLaunchKernelGGL((adj_stress_prep<R, M>), interior_grid(g), dim3(256), 0, s, g, F, B, md, gm, c, S, t0, t1);
    if (hipError_t e = adj_swap_levels(B, 3, NF, cells, t0, t1, s)) return e;
    // velocity-shaped sweep: level 0 = level 1 = bar V(t0), stresses = P, W
    FieldPtrs<R> A;
    for (int i = 0; i < 3; ++i) A.f[i][0] = A.f[i][1] = B.f[i][t0];
    for (int f = 0; f < 6; ++f) A.f[3 + f][0] = A.f[3 + f][1] = S.a[f];
    Coefs<R> cb = {};
    const int pr[3][2] = {{0, 1}, {1, 2}, {0, 2}};
    for (int k = 0; k < M; ++k) {
        for (int j = 0; j < 3; ++j) cb.cv[j][j][k] = -c.cn[0][j][k];
        for (int sl = 0; sl < 3; ++sl) {
            const int ia = pr[sl][0], ja = pr[sl][1];
            cb.cv[ia][ja][k] = -c.cs[sl][0][k];   // T_s reads V_ia along ja
            cb.cv[ja][ia][k] = -c.cs[sl][1][k];   // and V_ja along ia
        }
    }
    if (hipError_t e = launch_velocity(g, sizeof(R), OPESCI_STORE_NATIVE, false, &A, media, &cb, 0, 1, cfg, s))
        return e;
    hipLaunchKernelGGL((adj_rim_v<R, M>), rim_grid(g), dim3(256), 0, s, g, B, S, c, t0);
    return hipGetLastError();
}

template <typename R, int M>
static hipError_t adj_vel_t(const Geom& g, void* fields, void* bar, const void* media, void* grad, const void* coefs,
                            void* const* scratch, const AdjCells& cells, const InteriorLaunch& cfg, int t0, int t1,
                            hipStream_t s) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    const FieldPtrs<R>& F = *static_cast<const FieldPtrs<R>*>(fields);
    AdjPtrs<R>& B = *static_cast<AdjPtrs<R>*>(bar);
    const MediaPtrs<R>& md = *static_cast<const MediaPtrs<R>*>(media);
    const GradPtrs<R>& gm = *static_cast<const GradPtrs<R>*>(grad);
    const Coefs<R>& c = *static_cast<const Coefs<R>*>(coefs);
    if (!scratch || cfg.variant == 1) {
        hipLaunchKernelGGL((adj_vel_to_t<R, M>), grid_for(6 * n), dim3(256), 0, s, g, B, md, c, t1);
        hipLaunchKernelGGL((adj_vel_point<R, M>), grid_for(n), dim3(256), 0, s, g, F, B, gm, c, t0, t1);
        return hipGetLastError();
    }
    Scratch<R> S;
    for (int i = 0; i < 6; ++i) S.a[i] = static_cast<R*>(scratch[i]);
    hipLaunchKernelGGL((adj_vel_prep<R, M>), interior_grid(g), dim3(256), 0, s, g, F, B, md, gm, c, S, t0, t1);
    if (hipError_t e = adj_swap_levels(B, 0, 3, cells, t0, t1, s)) return e;
    // stress-shaped sweep: velocities = Q, level 0 = level 1 = bar T(t1)
    FieldPtrs<R> A;
    for (int i = 0; i < 3; ++i) A.f[i][0] = A.f[i][1] = S.a[i];
    for (int f = 3; f < NF; ++f) A.f[f][0] = A.f[f][1] = B.f[f][t1];
    Coefs<R> ca = {};
    const int pr[3][2] = {{0, 1}, {1, 2}, {0, 2}};
    for (int k = 0; k < M; ++k) {
        for (int i = 0; i < 3; ++i) ca.cn[i][i][k] = -c.cv[i][i][k];
        for (int sl = 0; sl < 3; ++sl) {
            const int ia = pr[sl][0], ja = pr[sl][1];
            ca.cs[sl][0][k] = -c.cv[ia][ja][k];   // V_ia reads T_s along ja
            ca.cs[sl][1][k] = -c.cv[ja][ia][k];   // V_ja reads T_s along ia
        }
    }
    if (hipError_t e = launch_stress(g, sizeof(R), OPESCI_STORE_NATIVE, false, &A, media, &ca, 0, 1, cfg, s))
        return e;
    hipLaunchKernelGGL((adj_rim_t<R, M>), rim_grid(g), dim3(256), 0, s, g, B, S, c, t1);
    return hipGetLastError();
}

#define OPESCI_ADJ_M(FN, R)                                                                     \
    switch (g.m) {                                                                              \
        case 1: return FN<R, 1>(g, fields, bar, media, grad, coefs, scratch, cells, cfg, t0, t1, s); \
        case 2: return FN<R, 2>(g, fields, bar, media, grad, coefs, scratch, cells, cfg, t0, t1, s); \
        case 3: return FN<R, 3>(g, fields, bar, media, grad, coefs, scratch, cells, cfg, t0, t1, s); \
        case 4: return FN<R, 4>(g, fields, bar, media, grad, coefs, scratch, cells, cfg, t0, t1, s); \
        default: return hipErrorInvalidValue;                                                   \
    }

hipError_t launch_adj_stress(const Geom& g, int real_bytes, void* fields, void* bar, const void* media, void* grad,
                             const void* coefs, void* const* scratch, const AdjCells& cells, const InteriorLaunch& cfg,
                             int t0, int t1, hipStream_t s) {
    if (real_bytes == 8) {
        OPESCI_ADJ_M(adj_stress_t, double)
    }
    OPESCI_ADJ_M(adj_stress_t, float)
}

hipError_t launch_adj_velocity(const Geom& g, int real_bytes, void* fields, void* bar, const void* media, void* grad,
                               const void* coefs, void* const* scratch, const AdjCells& cells,
                               const InteriorLaunch& cfg, int t0, int t1, hipStream_t s) {
    if (real_bytes == 8) {
        OPESCI_ADJ_M(adj_vel_t, double)
    }
    OPESCI_ADJ_M(adj_vel_t, float)
}

template <typename R>
static hipError_t bc_log_t(const Geom& g, void* fields, const BcDesc* desc, const void* pool, const BcLevel& lv,
                           void* log, int t0, int t1, hipStream_t s) {
    if (lv.count <= 0 || lv.maxpts <= 0) return hipSuccess;
    dim3 grid((lv.maxpts + 255) / 256, lv.count);
    hipLaunchKernelGGL((bc_log_kernel<R>), grid, dim3(256), 0, s, g, *static_cast<const FieldPtrs<R>*>(fields), desc,
                       static_cast<const R*>(pool), lv.first, lv.maxpts, static_cast<R*>(log), t0, t1);
    return hipGetLastError();
}

hipError_t launch_bc_logged(const Geom& g, int real_bytes, void* fields, const BcDesc* desc, const void* pool,
                            const BcLevel& lv, void* log, int t0, int t1, hipStream_t s) {
    return real_bytes == 8 ? bc_log_t<double>(g, fields, desc, pool, lv, log, t0, t1, s)
                           : bc_log_t<float>(g, fields, desc, pool, lv, log, t0, t1, s);
}

template <typename R>
static hipError_t adj_bc_t(const Geom& g, void* fields, void* bar, const BcDesc* desc, const void* pool, void* gpool,
                           const BcLevel& lv, const AdjGather& ga, const void* log, int t0, int t1, hipStream_t s) {
    if (lv.count <= 0 || lv.maxpts <= 0) return hipSuccess;
    dim3 grid((lv.maxpts + 255) / 256, lv.count);
    const AdjPtrs<R>& B = *static_cast<const AdjPtrs<R>*>(bar);
    hipLaunchKernelGGL((adj_bc_kernel<R>), grid, dim3(256), 0, s, g, *static_cast<const FieldPtrs<R>*>(fields), B,
                       desc, static_cast<const R*>(pool), static_cast<R*>(gpool), lv.first, ga.boff,
                       static_cast<R*>(ga.buf), static_cast<const R*>(log), lv.maxpts, t0, t1);
    if (hipError_t e = hipGetLastError()) return e;
    if (ga.ntgt <= 0) return hipSuccess;
    hipLaunchKernelGGL((adj_bc_gather_kernel<R>), dim3((ga.ntgt + 255) / 256), dim3(256), 0, s, B, desc,
                       static_cast<const R*>(pool), ga, t0, t1);
    return hipGetLastError();
}

hipError_t launch_adj_bc(const Geom& g, int real_bytes, void* fields, void* bar, const BcDesc* desc, const void* pool,
                         void* gpool, const BcLevel& lv, const AdjGather& ga, const void* log, int t0, int t1,
                         hipStream_t s) {
    return real_bytes == 8 ? adj_bc_t<double>(g, fields, bar, desc, pool, gpool, lv, ga, log, t0, t1, s)
                           : adj_bc_t<float>(g, fields, bar, desc, pool, gpool, lv, ga, log, t0, t1, s);
}

template <typename R>
static hipError_t adj_coef_t(const Geom& g, const void* media, void* grad, const BcDesc* desc,
                             const OpesciExprOp* code, const OpesciBcDTerm* dterms, const void* gpool, int d,
                             int npts, int k, int q, hipStream_t s) {
    if (npts <= 0) return hipSuccess;
    hipLaunchKernelGGL((adj_coef_media_kernel<R>), dim3((npts + 255) / 256), dim3(256), 0, s, g,
                       *static_cast<const MediaPtrs<R>*>(media), *static_cast<const GradPtrs<R>*>(grad), desc, code,
                       dterms, static_cast<const R*>(gpool), d, k, q);
    return hipGetLastError();
}

hipError_t launch_adj_coef_media(const Geom& g, int real_bytes, const void* media, void* grad, const BcDesc* desc,
                                 const OpesciExprOp* code, const OpesciBcDTerm* dterms, const void* gpool, int d,
                                 int npts, int k, int q, hipStream_t s) {
    return real_bytes == 8 ? adj_coef_t<double>(g, media, grad, desc, code, dterms, gpool, d, npts, k, q, s)
                           : adj_coef_t<float>(g, media, grad, desc, code, dterms, gpool, d, npts, k, q, s);
}

hipError_t launch_adj_receivers(const Geom& g, int real_bytes, void* fields, void* bar, int t1, int ngroups,
                                const int32_t* start, const int64_t* cell, const int32_t* order, const float* obs,
                                hipStream_t s) {
    (void)g;
    if (ngroups <= 0) return hipSuccess;
    if (real_bytes == 8)
        hipLaunchKernelGGL((adj_receiver_kernel<double>), dim3((ngroups + 255) / 256), dim3(256), 0, s,
                           *static_cast<const FieldPtrs<double>*>(fields), *static_cast<const AdjPtrs<double>*>(bar),
                           t1, ngroups, start, cell, order, obs);
    else
        hipLaunchKernelGGL((adj_receiver_kernel<float>), dim3((ngroups + 255) / 256), dim3(256), 0, s,
                           *static_cast<const FieldPtrs<float>*>(fields), *static_cast<const AdjPtrs<float>*>(bar),
                           t1, ngroups, start, cell, order, obs);
    return hipGetLastError();
}

hipError_t launch_src_cells(const Geom& g, int real_bytes, void* fields, int t, int x, int y, int z, void* buf,
                            int mode, hipStream_t s) {
    const int64_t q = g.at(x, y, z);
    if (real_bytes == 8)
        hipLaunchKernelGGL((src_cells_kernel<double>), dim3(1), dim3(64), 0, s,
                           *static_cast<const FieldPtrs<double>*>(fields), t, q, static_cast<double*>(buf), mode);
    else
        hipLaunchKernelGGL((src_cells_kernel<float>), dim3(1), dim3(64), 0, s,
                           *static_cast<const FieldPtrs<float>*>(fields), t, q, static_cast<float*>(buf), mode);
    return hipGetLastError();
}

// D.f[field][td][offset] = S.f[field][ts][offset] for a list of distinct
// cells, each entry (offset << 4) | field.
template <typename R>
__global__ __launch_bounds__(256) void copy_cells_kernel(FieldPtrs<R> D, int td, FieldPtrs<R> S, int ts,
                                                         const int64_t* __restrict__ cells, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t e = cells[i];
    const int f = (int)(e & 15);
    const int64_t off = e >> 4;
    D.f[f][td][off] = S.f[f][ts][off];
}

hipError_t launch_copy_cells(int real_bytes, void* dst, int td, void* src, int ts, const int64_t* cells, int n,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (real_bytes == 8)
        hipLaunchKernelGGL((copy_cells_kernel<double>), grid, dim3(256), 0, s,
                           *static_cast<const FieldPtrs<double>*>(dst), td,
                           *static_cast<const FieldPtrs<double>*>(src), ts, cells, n);
    else
        hipLaunchKernelGGL((copy_cells_kernel<float>), grid, dim3(256), 0, s,
                           *static_cast<const FieldPtrs<float>*>(dst), td, *static_cast<const FieldPtrs<float>*>(src),
                           ts, cells, n);
    return hipGetLastError();
}

hipError_t launch_set_step(int* step, int value, hipStream_t s) {
    hipLaunchKernelGGL(set_step_kernel, dim3(1), dim3(1), 0, s, step, value);
    return hipGetLastError();
}

}  // namespace opesci
