// Boundary program, analytic initial state, L2 norms, effective media,
// layout conversion and point source / receivers.
#include "bcdev.h"

#include <mutex>

namespace opesci {

// real_t x = dx*(i - m [+ 0.5]) of the reference loops, rounded once to R
template <typename R>
__device__ __forceinline__ double coord_of(double dx, int i, int m, int stagger) {
    const R h = (R)dx;
    return (double)(R)((double)h * ((double)(i - m) + (stagger ? 0.5 : 0.0)));
}

// ------------------------------------------------------------------ boundary program
// One launch per hazard level, blockIdx.y = statement.  Statements of a
// level write disjoint cells and never read what another statement of the
// level writes.  Terms are loaded independently and summed in term order.
template <typename R, typename S>
__global__ __launch_bounds__(256) void bc_kernel(Geom g, FieldPtrs<S> F, const BcDesc* __restrict__ desc,
                                                 const R* __restrict__ pool, int first, int t0, int t1) {
    const BcDesc* D = desc + first + blockIdx.y;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D->npts) return;
    const int n1 = D->n1;
    bc_point<R, S>(g, F, D, pool, p, D->lo0 + p / n1, D->lo1 + p % n1, t0, t1);
}

// Per-point media-dependent coefficients of the statements [first,
// first+count): coef * program, rounded to R exactly as bc_point does.
template <typename R>
__global__ __launch_bounds__(256) void bc_coef_kernel(Geom g, MediaPtrs<R> md, const BcDesc* __restrict__ desc,
                                                      const OpesciExprOp* __restrict__ code, R* __restrict__ pool,
                                                      int first) {
    const BcDesc* D = desc + first + blockIdx.y;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D->npts) return;
    const int n1 = D->n1, axis = D->axis;
    const int i0 = D->lo0 + p / n1, i1 = D->lo1 + p % n1;
    const int ax = axis == 0 ? 0 : i0;
    const int ay = axis == 0 ? i0 : (axis == 1 ? 0 : i1);
    const int az = axis == 2 ? 0 : i1;
    const double c0[3] = {0, 0, 0};
    for (int k = 0; k < D->nterms; ++k)
        if (D->t[k].cslot >= 0)
            pool[D->cpool + (int64_t)D->t[k].cslot * D->npts + p] =
                (R)((double)D->t[k].coef * vm_eval<R>(code + D->t[k].prog, D->prog_len[k], c0, 0.0, md, g, ax, ay, az));
}

struct SolTable {
    double dx[3];
    double t[NF];
};

// The tables of the split analytic programs: slot blockIdx.y at every index
// i along its axis, coordinates as the point kernels compute them.
template <typename R>
__global__ __launch_bounds__(256) void sol_table_kernel(Geom g, MediaPtrs<R> md, SolProg sp, SolTable tab) {
    const SolSlot S = sp.slots[blockIdx.y];
    const int n = S.axis == 0 ? g.nx : (S.axis == 1 ? g.ny : g.nz);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double c[3] = {0.0, 0.0, 0.0};
    c[S.axis] = coord_of<R>(tab.dx[S.axis], i, g.m, stag(S.field, S.axis));
    sp.tab[S.tab + i] = vm_eval<R>(sp.code + S.prog, S.len, c, tab.t[S.field], md, g, 0, 0, 0);
}

// Field f's tables for the point kernels: copied into the block's LDS when
// they fit lds_cap doubles (the launch's dynamic LDS), else read in place.
constexpr int SOL_LDS_CAP = 4096;   // doubles (32 KiB)
__device__ __forceinline__ const double* stage_tables(const SolProg& sol, int f, int lds_cap) {
    extern __shared__ double stab[];
    const int n = sol.tab_len[f];
    if (n == 0 || n > lds_cap) return sol.tab + sol.tab_base[f];
    for (int i = threadIdx.x; i < n; i += blockDim.x) stab[i] = sol.tab[sol.tab_base[f] + i];
    __syncthreads();
    return stab;
}

// ------------------------------------------------------------------ initial state, L2 norms
// Field f's analytic values at z on K rows (x[k], y[k]) of a block: the K
// points of a shallow program in one pass over its ops, the others one
// point at a time.
constexpr int SOL_ROWS = 4;
template <typename R>
__device__ __forceinline__ void sol_points(const SolProg& sol, const OpesciExprOp* __restrict__ code, int f,
                                           const SolTable& tab, const MediaPtrs<R>& md, const Geom& g,
                                           const int x[SOL_ROWS], const int y[SOL_ROWS], int z,
                                           const double* tabs, double v[SOL_ROWS]) {
    const int m = g.m;
    double cx[SOL_ROWS], cy[SOL_ROWS];
#pragma unroll
    for (int k = 0; k < SOL_ROWS; ++k) {
        cx[k] = coord_of<R>(tab.dx[0], x[k], m, stag(f, 0));
        cy[k] = coord_of<R>(tab.dx[1], y[k], m, stag(f, 1));
    }
    const double cz = coord_of<R>(tab.dx[2], z, m, stag(f, 2));
    if (sol.shallow[f]) {
        vm_eval_k<R, SOL_ROWS>(code + sol.off[f], sol.len[f], cx, cy, cz, tab.t[f], md, g, x, y, z, tabs, v);
    } else {
        for (int k = 0; k < SOL_ROWS; ++k) {
            const double c[3] = {cx[k], cy[k], cz};
            v[k] = vm_eval_tos<R>(code + sol.off[f], sol.len[f], c, tab.t[f], md, g, x[k], y[k], z, tabs);
        }
    }
}

// rows r0 .. r0 + SOL_ROWS - 1 of field f's interior (x, y) rows; past the
// last row, repeats of it flagged not ok
__device__ __forceinline__ void sol_rows(int r0, int rows, int ny, int m, int x[SOL_ROWS], int y[SOL_ROWS],
                                         bool ok[SOL_ROWS]) {
#pragma unroll
    for (int k = 0; k < SOL_ROWS; ++k) {
        const int r = min(r0 + k, rows - 1);
        ok[k] = r0 + k < rows;
        x[k] = m + r / ny;
        y[k] = m + r % ny;
    }
}

// StaggeredGrid.initialise (opesci/staggeredgrid.py:612-659): level 0,
// ranges [m, dim-m) or [m, dim-m-1) when staggered, t = dt/2 for velocities.
template <typename R, typename S>
__global__ __launch_bounds__(256) void initial_kernel(Geom g, FieldPtrs<S> F, MediaPtrs<R> md, SolProg sol,
                                                      const OpesciExprOp* __restrict__ code, SolTable tab,
                                                      int lds_cap) {
    using C = Conv<S, R>;
    const int f = blockIdx.y;
    const int m = g.m;
    const int nx = g.nx - 2 * m - stag(f, 0), ny = g.ny - 2 * m - stag(f, 1), nz = g.nz - 2 * m - stag(f, 2);
    const int rows = nx * ny;
    const double* tabs = stage_tables(sol, f, lds_cap);
    // SOL_ROWS (x, y) rows per block iteration (block-uniform indices), z across threads
    for (int r0 = blockIdx.x * SOL_ROWS; r0 < rows; r0 += gridDim.x * SOL_ROWS) {
        int x[SOL_ROWS], y[SOL_ROWS];
        bool ok[SOL_ROWS];
        sol_rows(r0, rows, ny, m, x, y, ok);
        for (int z = m + (int)threadIdx.x; z < m + nz; z += blockDim.x) {
            double v[SOL_ROWS];
            sol_points<R>(sol, code, f, tab, md, g, x, y, z, tabs, v);
#pragma unroll
            for (int k = 0; k < SOL_ROWS; ++k)
                if (ok[k]) C::st(F.f[f][0] + g.at(x[k], y[k], z), (R)v[k]);
        }
    }
}

// sum over the interior of (F - F_exact(t_n))^2 in double
// (StaggeredGrid.converge_test, opesci/staggeredgrid.py:892-945); wavefront
// reduction by shuffles, one partial per block.
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename R, typename S>
__global__ __launch_bounds__(256) void l2_kernel(Geom g, FieldPtrs<S> F, MediaPtrs<R> md, SolProg sol,
                                                 const OpesciExprOp* __restrict__ code, SolTable tab, int level,
                                                 double* __restrict__ partial, int lds_cap) {
    using C = Conv<S, R>;
    __shared__ double red[4];
    const int f = blockIdx.y;
    const int m = g.m;
    const int nx = g.nx - 2 * m - stag(f, 0), ny = g.ny - 2 * m - stag(f, 1), nz = g.nz - 2 * m - stag(f, 2);
    const int rows = nx * ny;
    const double* tabs = stage_tables(sol, f, lds_cap);
    double acc = 0.0;
    for (int r0 = blockIdx.x * SOL_ROWS; r0 < rows; r0 += gridDim.x * SOL_ROWS) {
        int x[SOL_ROWS], y[SOL_ROWS];
        bool ok[SOL_ROWS];
        sol_rows(r0, rows, ny, m, x, y, ok);
        for (int z = m + (int)threadIdx.x; z < m + nz; z += blockDim.x) {
            double v[SOL_ROWS];
            sol_points<R>(sol, code, f, tab, md, g, x, y, z, tabs, v);
#pragma unroll
            for (int k = 0; k < SOL_ROWS; ++k)
                if (ok[k]) {
                    const double e = (double)C::ld(F.f[f][level] + g.at(x[k], y[k], z)) - v[k];
                    acc += e * e;
                }
        }
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[(size_t)f * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// ------------------------------------------------------------------ media
// StaggeredGrid.read_data (opesci/staggeredgrid.py:522-598): beta = 1/rho,
// lambda = rho (vp^2 - 2 vs^2), mu = rho vs^2 on every cell; beta1..3 =
// arithmetic means over the +x/+y/+z neighbour, mu12/mu13/mu23 = harmonic
// means over the 4-cell faces (the last index clamps to itself).
template <typename R>
__global__ void media_base_kernel(Geom g, const float* __restrict__ rho, const float* __restrict__ vp,
                                  const float* __restrict__ vs, MediaPtrs<R> out) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(p % g.nz), y = (int)((p / g.nz) % g.ny), x = (int)(p / ((int64_t)g.nz * g.ny));
        const double r = rho[p], a = vp[p], b = vs[p];
        const int64_t q = g.at(x, y, z);
        R* const* o = (R* const*)out.m;
        o[OPESCI_M_RHO][q] = (R)r;
        o[OPESCI_M_VP][q] = (R)a;
        o[OPESCI_M_VS][q] = (R)b;
        o[OPESCI_M_BETA][q] = (R)(1.0 / r);
        o[OPESCI_M_LAMBDA][q] = (R)(r * (a * a - 2.0 * b * b));
        o[OPESCI_M_MU][q] = (R)(r * (b * b));
    }
}

template <typename R>
__global__ void media_eff_kernel(Geom g, MediaPtrs<R> md) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    R* const* o = (R* const*)md.m;
    const R* beta = md.m[OPESCI_M_BETA];
    const R* mu = md.m[OPESCI_M_MU];
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(p % g.nz), y = (int)((p / g.nz) % g.ny), x = (int)(p / ((int64_t)g.nz * g.ny));
        const int x1 = min(x + 1, g.nx - 1), y1 = min(y + 1, g.ny - 1), z1 = min(z + 1, g.nz - 1);
        const int64_t q = g.at(x, y, z);
        const double b0 = beta[q];
        o[OPESCI_M_BETA1][q] = (R)((b0 + (double)beta[g.at(x1, y, z)]) / 2.0);
        o[OPESCI_M_BETA2][q] = (R)((b0 + (double)beta[g.at(x, y1, z)]) / 2.0);
        o[OPESCI_M_BETA3][q] = (R)((b0 + (double)beta[g.at(x, y, z1)]) / 2.0);
        const double i0 = 1.0 / (double)mu[q];
        const double ix = 1.0 / (double)mu[g.at(x1, y, z)], iy = 1.0 / (double)mu[g.at(x, y1, z)],
                     iz = 1.0 / (double)mu[g.at(x, y, z1)];
        o[OPESCI_M_MU12][q] = (R)(1.0 / (0.25 * (i0 + ix + iy + 1.0 / (double)mu[g.at(x1, y1, z)])));
        o[OPESCI_M_MU13][q] = (R)(1.0 / (0.25 * (i0 + ix + iz + 1.0 / (double)mu[g.at(x1, y, z1)])));
        o[OPESCI_M_MU23][q] = (R)(1.0 / (0.25 * (i0 + iy + iz + 1.0 / (double)mu[g.at(x, y1, z1)])));
    }
}

// ------------------------------------------------------------------ layout
template <typename R, typename S>
__global__ void pack_kernel(Geom g, const R* __restrict__ dense, S* __restrict__ field) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(p % g.nz), y = (int)((p / g.nz) % g.ny), x = (int)(p / ((int64_t)g.nz * g.ny));
        Conv<S, R>::st(field + g.at(x, y, z), dense[p]);
    }
}

template <typename R, typename S>
__global__ void unpack_kernel(Geom g, const S* __restrict__ field, R* __restrict__ dense) {
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(p % g.nz), y = (int)((p / g.nz) % g.ny), x = (int)(p / ((int64_t)g.nz * g.ny));
        dense[p] = Conv<S, R>::ld(field + g.at(x, y, z));
    }
}

__global__ void zero_kernel(uint4* __restrict__ p, size_t n16) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0, 0, 0, 0);
}

// ------------------------------------------------------------------ source / receivers
// tests/src/test_ref_iso_elastic.cpp:227-290: after the velocity update record
// U, V, W and (Txx+Tyy+Tzz)/3 at every receiver, then Tnn -= w[step]/3 at the
// source.  The step counter lives on the device so a captured graph can be
// replayed.
// Receivers (tests/src/test_ref_iso_elastic.cpp:227-272): U, V, W and
// (Txx+Tyy+Tzz)/3 at each receiver's grid point after the velocity update,
// one thread per receiver, trace[step][receiver][4] in the arithmetic type R
// (FP64 contexts keep FP64 traces, so misfits of FP64 runs are not rounded).
template <typename R> struct Trace4;
template <> struct Trace4<float> {
    static __device__ void put(void* t, size_t i, float a, float b, float c, float d) {
        reinterpret_cast<float4*>(t)[i] = make_float4(a, b, c, d);
    }
};
template <> struct Trace4<double> {
    static __device__ void put(void* t, size_t i, double a, double b, double c, double d) {
        double2* p = reinterpret_cast<double2*>(t) + 2 * i;
        p[0] = make_double2(a, b);
        p[1] = make_double2(c, d);
    }
};

template <typename R, typename S>
__global__ __launch_bounds__(256) void receiver_kernel(Geom g, FieldPtrs<S> F, int t1, int nrec,
                                                       const int32_t* __restrict__ rec, void* __restrict__ trace,
                                                       const int* __restrict__ step, int ntsteps) {
    using C = Conv<S, R>;
    const int s = *step;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ntsteps || r >= nrec) return;
    const int64_t q = g.at(rec[3 * r], rec[3 * r + 1], rec[3 * r + 2]);
    Trace4<R>::put(trace, (size_t)s * nrec + r, C::ld(F.f[0][t1] + q), C::ld(F.f[1][t1] + q), C::ld(F.f[2][t1] + q),
                   (C::ld(F.f[3][t1] + q) + C::ld(F.f[4][t1] + q) + C::ld(F.f[5][t1] + q)) / (R)3);
}

// Point source (test_ref_iso_elastic.cpp:274-290), after the receivers:
// Txx -= wx[t]/3, Tyy -= wy[t]/3, Tzz -= wz[t]/3 (one trace for all three
// when ntr == 1); then the device step counter advances.
template <typename R, typename S>
__global__ void source_kernel(Geom g, FieldPtrs<S> F, int t1, int ntr, int sx, int sy, int sz,
                              const float* __restrict__ w, int* step, int ntsteps) {
    using C = Conv<S, R>;
    const int s = *step;
    if (ntr > 0 && s < ntsteps) {
        const int64_t q = g.at(sx, sy, sz);
        for (int f = 0; f < 3; ++f) {
            const R v = (R)w[(ntr == 3 ? f : 0) * (size_t)ntsteps + s] / (R)3;
            C::st(F.f[3 + f][t1] + q, C::ld(F.f[3 + f][t1] + q) - v);
        }
    }
    *step = s + 1;
}

// ------------------------------------------------------------------ launchers
template <typename R, typename S>
static hipError_t bc_t(const Geom& g, void* fields, const BcDesc* desc, const void* pool, const BcLevel& lv, int t0,
                       int t1, hipStream_t s) {
    if (lv.count <= 0 || lv.maxpts <= 0) return hipSuccess;
    dim3 grid((lv.maxpts + 255) / 256, lv.count);
    hipLaunchKernelGGL((bc_kernel<R, S>), grid, dim3(256), 0, s, g, *static_cast<const FieldPtrs<S>*>(fields), desc,
                       static_cast<const R*>(pool), lv.first, t0, t1);
    return hipGetLastError();
}

template <typename R>
static hipError_t bc_coef_t(const Geom& g, const void* media, const BcDesc* desc, const OpesciExprOp* code,
                            void* pool, int first, int count, int maxpts, hipStream_t s) {
    if (count <= 0 || maxpts <= 0) return hipSuccess;
    dim3 grid((maxpts + 255) / 256, count);
    hipLaunchKernelGGL((bc_coef_kernel<R>), grid, dim3(256), 0, s, g, *static_cast<const MediaPtrs<R>*>(media), desc,
                       code, static_cast<R*>(pool), first);
    return hipGetLastError();
}

hipError_t launch_bc_coef(const Geom& g, int real_bytes, const void* media, const BcDesc* d_desc,
                          const OpesciExprOp* d_code, void* pool, int first, int count, int maxpts, hipStream_t s) {
    return real_bytes == 8 ? bc_coef_t<double>(g, media, d_desc, d_code, pool, first, count, maxpts, s)
                           : bc_coef_t<float>(g, media, d_desc, d_code, pool, first, count, maxpts, s);
}

#define OPESCI_TYPED(RB, ST, CALL_F, CALL_D, CALL_H, CALL_B) \
    ((RB) == 8 ? (CALL_D) : (ST) == OPESCI_STORE_F16 ? (CALL_H) : (ST) == OPESCI_STORE_BF16 ? (CALL_B) : (CALL_F))

hipError_t launch_bc(const Geom& g, int real_bytes, int storage, void* fields, const BcDesc* d_desc, const void* pool,
                     const BcLevel& lvl, int t0, int t1, hipStream_t s) {
    return OPESCI_TYPED(real_bytes, storage, (bc_t<float, float>(g, fields, d_desc, pool, lvl, t0, t1, s)),
                        (bc_t<double, double>(g, fields, d_desc, pool, lvl, t0, t1, s)),
                        (bc_t<float, _Float16>(g, fields, d_desc, pool, lvl, t0, t1, s)),
                        (bc_t<float, bf16_t>(g, fields, d_desc, pool, lvl, t0, t1, s)));
}

static SolTable make_tab(const double* dx) {
    SolTable t;
    for (int f = 0; f < NF; ++f) t.t[f] = 0.0;
    for (int a = 0; a < 3; ++a) t.dx[a] = dx[a];
    return t;
}

// doubles of dynamic LDS for the point kernels: the largest field's tables
// when within SOL_LDS_CAP (fields with larger tables read them in place)
static int sol_lds(const SolProg& sol) {
    int n = 0;
    for (int f = 0; f < NF; ++f)
        if (sol.tab_len[f] <= SOL_LDS_CAP) n = std::max(n, sol.tab_len[f]);
    return n;
}

// the tables for this launch's times (stream-ordered before the point kernel)
template <typename R>
static hipError_t sol_tables_t(const Geom& g, const void* media, const SolProg& sol, const SolTable& tab,
                               hipStream_t s) {
    if (sol.nslots <= 0) return hipSuccess;
    hipLaunchKernelGGL((sol_table_kernel<R>), dim3((sol.maxdim + 255) / 256, sol.nslots), dim3(256), 0, s, g,
                       *static_cast<const MediaPtrs<R>*>(media), sol, tab);
    return hipGetLastError();
}

template <typename R, typename S>
static hipError_t init_t(const Geom& g, void* fields, const void* media, const SolProg& sol, SolTable tab,
                         hipStream_t s) {
    if (hipError_t e = sol_tables_t<R>(g, media, sol, tab, s)) return e;
    const int cap = sol_lds(sol);
    hipLaunchKernelGGL((initial_kernel<R, S>), dim3(512, NF), dim3(256), cap * sizeof(double), s, g,
                       *static_cast<const FieldPtrs<S>*>(fields), *static_cast<const MediaPtrs<R>*>(media), sol,
                       sol.code, tab, cap);
    return hipGetLastError();
}

hipError_t launch_initial(const Geom& g, int real_bytes, int storage, void* fields, const void* media,
                          const SolProg& sol, const double* spacing, double dt, hipStream_t s) {
    SolTable tab = make_tab(spacing);
    for (int f = 0; f < NF; ++f) tab.t[f] = time_stag(f) ? dt / 2.0 : 0.0;
    return OPESCI_TYPED(real_bytes, storage, (init_t<float, float>(g, fields, media, sol, tab, s)),
                        (init_t<double, double>(g, fields, media, sol, tab, s)),
                        (init_t<float, _Float16>(g, fields, media, sol, tab, s)),
                        (init_t<float, bf16_t>(g, fields, media, sol, tab, s)));
}

template <typename R, typename S>
static hipError_t l2_t(const Geom& g, void* fields, const void* media, const SolProg& sol, SolTable tab,
                       int level, double* partial, int nblk, hipStream_t s) {
    if (hipError_t e = sol_tables_t<R>(g, media, sol, tab, s)) return e;
    const int cap = sol_lds(sol);
    hipLaunchKernelGGL((l2_kernel<R, S>), dim3(nblk, NF), dim3(256), cap * sizeof(double), s, g,
                       *static_cast<const FieldPtrs<S>*>(fields), *static_cast<const MediaPtrs<R>*>(media), sol,
                       sol.code, tab, level, partial, cap);
    return hipGetLastError();
}

hipError_t launch_l2(const Geom& g, int real_bytes, int storage, void* fields, const void* media,
                     const SolProg& sol, const double* spacing, int level, const double t_field[NF],
                     double* d_partial, int partial_cap, double* out_host, hipStream_t s) {
    SolTable tab = make_tab(spacing);
    for (int f = 0; f < NF; ++f) tab.t[f] = t_field[f];
    const int nblk = partial_cap / NF;
    hipError_t e = OPESCI_TYPED(real_bytes, storage,
                                (l2_t<float, float>(g, fields, media, sol, tab, level, d_partial, nblk, s)),
                                (l2_t<double, double>(g, fields, media, sol, tab, level, d_partial, nblk, s)),
                                (l2_t<float, _Float16>(g, fields, media, sol, tab, level, d_partial, nblk, s)),
                                (l2_t<float, bf16_t>(g, fields, media, sol, tab, level, d_partial, nblk, s)));
    if (e != hipSuccess) return e;
    std::vector<double> h((size_t)NF * nblk);
    e = hipMemcpyAsync(h.data(), d_partial, h.size() * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    const double vol = spacing[0] * spacing[1] * spacing[2];
    for (int f = 0; f < NF; ++f) {
        double a = 0.0;
        for (int b = 0; b < nblk; ++b) a += h[(size_t)f * nblk + b];
        out_host[f] = sqrt(a * vol);
    }
    return hipSuccess;
}

// ------------------------------------------------------------------ L2 without allocation
// The per-block partial sums of opesci_ctx_l2_noalloc live in one
// library-owned array per device: NF x L2_BLOCKS doubles (36 KiB), held
// until the process ends and shared by every call on that device (calls
// that overlap on one device race on it).  Every block of l2_kernel writes
// its partial, so nothing needs clearing between calls.  In host code the
// array's name stands for a host-side shadow, never for device memory: the
// device address is looked up once per device with hipGetSymbolAddress and
// only that address is handed to kernels.
constexpr int L2_BLOCKS = 512;
constexpr int L2_MAX_DEVICES = 64;
static __device__ __align__(256) double l2_scratch[NF * L2_BLOCKS];

// out[f] = sqrt(vol * sum_b partial[f][b]), summed in block order like the
// host reduction of launch_l2
__global__ void l2_finish_kernel(const double* __restrict__ partial, int nblk, double vol, double* __restrict__ out) {
    const int f = threadIdx.x;
    if (f >= NF) return;
    double a = 0.0;
    for (int b = 0; b < nblk; ++b) a += partial[(size_t)f * nblk + b];
    out[f] = sqrt(a * vol);
}

static hipError_t l2_scratch_of_current_device(double** out) {
    static std::mutex mu;
    static double* addr[L2_MAX_DEVICES] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= L2_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(l2_scratch));
        if (e != hipSuccess) return e;
        addr[dev] = static_cast<double*>(p);
    }
    *out = addr[dev];
    return hipSuccess;
}

int l2_noalloc_partials() { return L2_BLOCKS; }

hipError_t launch_l2_noalloc(const Geom& g, int real_bytes, int storage, void* fields, const void* media,
                             const SolProg& sol, const double* spacing, int level, const double t_field[NF],
                             double* d_out, hipStream_t s) {
    double* partial = nullptr;
    hipError_t e = l2_scratch_of_current_device(&partial);
    if (e != hipSuccess) return e;
    SolTable tab = make_tab(spacing);
    for (int f = 0; f < NF; ++f) tab.t[f] = t_field[f];
    e = OPESCI_TYPED(real_bytes, storage,
                     (l2_t<float, float>(g, fields, media, sol, tab, level, partial, L2_BLOCKS, s)),
                     (l2_t<double, double>(g, fields, media, sol, tab, level, partial, L2_BLOCKS, s)),
                     (l2_t<float, _Float16>(g, fields, media, sol, tab, level, partial, L2_BLOCKS, s)),
                     (l2_t<float, bf16_t>(g, fields, media, sol, tab, level, partial, L2_BLOCKS, s)));
    if (e != hipSuccess) return e;
    const double vol = spacing[0] * spacing[1] * spacing[2];
    hipLaunchKernelGGL(l2_finish_kernel, dim3(1), dim3(64), 0, s, (const double*)partial, L2_BLOCKS, vol, d_out);
    return hipGetLastError();
}

hipError_t launch_media(const Geom& g, int real_bytes, const float* d_rho, const float* d_vp, const float* d_vs,
                        void* const* d_media, hipStream_t s) {
    if (real_bytes == 8) {
        MediaPtrs<double> mp;
        for (int i = 0; i < OPESCI_M_COUNT; ++i) mp.m[i] = (const double*)d_media[i];
        hipLaunchKernelGGL((media_base_kernel<double>), dim3(1024), dim3(256), 0, s, g, d_rho, d_vp, d_vs, mp);
        hipLaunchKernelGGL((media_eff_kernel<double>), dim3(1024), dim3(256), 0, s, g, mp);
    } else {
        MediaPtrs<float> mp;
        for (int i = 0; i < OPESCI_M_COUNT; ++i) mp.m[i] = (const float*)d_media[i];
        hipLaunchKernelGGL((media_base_kernel<float>), dim3(1024), dim3(256), 0, s, g, d_rho, d_vp, d_vs, mp);
        hipLaunchKernelGGL((media_eff_kernel<float>), dim3(1024), dim3(256), 0, s, g, mp);
    }
    return hipGetLastError();
}

template <typename R, typename S>
static hipError_t pack_t(const Geom& g, const void* dense, void* field, hipStream_t s) {
    hipLaunchKernelGGL((pack_kernel<R, S>), dim3(1024), dim3(256), 0, s, g, (const R*)dense, (S*)field);
    return hipGetLastError();
}
template <typename R, typename S>
static hipError_t unpack_t(const Geom& g, const void* field, void* dense, hipStream_t s) {
    hipLaunchKernelGGL((unpack_kernel<R, S>), dim3(1024), dim3(256), 0, s, g, (const S*)field, (R*)dense);
    return hipGetLastError();
}

hipError_t launch_pack(const Geom& g, int real_bytes, int storage, const void* dense, void* field, hipStream_t s) {
    return OPESCI_TYPED(real_bytes, storage, (pack_t<float, float>(g, dense, field, s)),
                        (pack_t<double, double>(g, dense, field, s)), (pack_t<float, _Float16>(g, dense, field, s)),
                        (pack_t<float, bf16_t>(g, dense, field, s)));
}

hipError_t launch_unpack(const Geom& g, int real_bytes, int storage, const void* field, void* dense, hipStream_t s) {
    return OPESCI_TYPED(real_bytes, storage, (unpack_t<float, float>(g, field, dense, s)),
                        (unpack_t<double, double>(g, field, dense, s)),
                        (unpack_t<float, _Float16>(g, field, dense, s)), (unpack_t<float, bf16_t>(g, field, dense, s)));
}

hipError_t launch_pack_media(const Geom& g, int real_bytes, const void* dense, void* media, hipStream_t s) {
    return real_bytes == 8 ? pack_t<double, double>(g, dense, media, s) : pack_t<float, float>(g, dense, media, s);
}

hipError_t launch_zero(void* p, size_t bytes, hipStream_t s) {
    const size_t n16 = bytes / 16;
    if (n16) hipLaunchKernelGGL(zero_kernel, dim3(2048), dim3(256), 0, s, (uint4*)p, n16);
    return hipGetLastError();
}

template <typename R, typename S>
static hipError_t rec_t(const Geom& g, void* fields, int t1, int nrec, const int32_t* rec, void* trace, int ntr,
                        const int32_t* si, const float* w, int* step, int ntsteps, hipStream_t s) {
    const FieldPtrs<S>& F = *static_cast<const FieldPtrs<S>*>(fields);
    if (nrec > 0 && trace)
        hipLaunchKernelGGL((receiver_kernel<R, S>), dim3((nrec + 255) / 256), dim3(256), 0, s, g, F, t1, nrec, rec,
                           trace, step, ntsteps);
    hipLaunchKernelGGL((source_kernel<R, S>), dim3(1), dim3(1), 0, s, g, F, t1, ntr, si[0], si[1], si[2], w, step,
                       ntsteps);
    return hipGetLastError();
}

hipError_t launch_record_source(const Geom& g, int real_bytes, int storage, void* fields, int t1, int nrec,
                                const int32_t* d_rec, void* d_trace, int src_on, const int32_t* src_idx,
                                const float* d_wavelet, int* d_step, int ntsteps, hipStream_t s) {
    return OPESCI_TYPED(
        real_bytes, storage,
        (rec_t<float, float>(g, fields, t1, nrec, d_rec, d_trace, src_on, src_idx, d_wavelet, d_step, ntsteps, s)),
        (rec_t<double, double>(g, fields, t1, nrec, d_rec, d_trace, src_on, src_idx, d_wavelet, d_step, ntsteps, s)),
        (rec_t<float, _Float16>(g, fields, t1, nrec, d_rec, d_trace, src_on, src_idx, d_wavelet, d_step, ntsteps, s)),
        (rec_t<float, bf16_t>(g, fields, t1, nrec, d_rec, d_trace, src_on, src_idx, d_wavelet, d_step, ntsteps, s)));
}

}  // namespace opesci
