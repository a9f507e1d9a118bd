// Internal definitions shared by the libopesci_hip translation units.
//
// Device layout of one field time level: F[x][y][zoff + z], x slowest,
// z contiguous, rows padded to `pitch` elements so that the first interior
// z (= m) of every row starts on a 128-byte boundary for 4-byte storage.
// Pointers handed to kernels are pre-offset by zoff, so element (x,y,z) is
// p[(x*ny + y)*pitch + z] for z in [0, nz); the padding in front of z = 0
// keeps every stencil read (|offset| <= m) inside the allocation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "opesci_hip.h"

namespace opesci {

constexpr int NF = 9;
constexpr int MAXM = OPESCI_MAX_HALF_ORDER;

// field staggering (x, y, z) in the order U, V, W, Txx, Tyy, Tzz, Txy, Tyz, Txz
__host__ __device__ constexpr int stag(int f, int a) {
    return (f == 0 && a == 0) || (f == 1 && a == 1) || (f == 2 && a == 2) ||
                   (f == 6 && a != 2) || (f == 7 && a != 0) || (f == 8 && a != 1)
               ? 1
               : 0;
}
__host__ __device__ constexpr int time_stag(int f) { return f < 3 ? 1 : 0; }

// ---------------------------------------------------------------- storage
struct bf16_t {
    uint16_t bits;
};

template <typename S, typename R>
struct Conv;
// ld/st: ordinary accesses.  ldnt/stnt: non-temporal, for streams a sweep
// kernel touches once (its own cell's previous value, every store), so the
// caches keep the halo lines that neighbouring tiles re-read.
template <>
struct Conv<float, float> {
    __device__ __forceinline__ static float ld(const float* p) { return *p; }
    __device__ __forceinline__ static void st(float* p, float v) { *p = v; }
    __device__ __forceinline__ static float ldnt(const float* p) { return __builtin_nontemporal_load(p); }
    __device__ __forceinline__ static void stnt(float* p, float v) { __builtin_nontemporal_store(v, p); }
    __device__ __forceinline__ static float from(float v) { return v; }
    __device__ __forceinline__ static float to(float v) { return v; }
};
template <>
struct Conv<double, double> {
    __device__ __forceinline__ static double ld(const double* p) { return *p; }
    __device__ __forceinline__ static void st(double* p, double v) { *p = v; }
    __device__ __forceinline__ static double ldnt(const double* p) { return __builtin_nontemporal_load(p); }
    __device__ __forceinline__ static void stnt(double* p, double v) { __builtin_nontemporal_store(v, p); }
    __device__ __forceinline__ static double from(double v) { return v; }
    __device__ __forceinline__ static double to(double v) { return v; }
};
// FP32 -> FP16, round to nearest even, from a value already rounded to FP32.
// The empty register constraint keeps the compiler from fusing a producing
// fma with the conversion into one mixed-precision fma (v_fma_mix*), which
// rounds once instead of twice and so differs in the last bit now and then
// -- and does so in some kernels and not in others, while every kernel that
// stores a value must round it the same way.
__device__ __forceinline__ _Float16 to_f16(float v) {
    asm("" : "+v"(v));
    return (_Float16)v;
}
template <>
struct Conv<_Float16, float> {
    __device__ __forceinline__ static float from(_Float16 v) { return (float)v; }
    __device__ __forceinline__ static _Float16 to(float v) { return to_f16(v); }
    __device__ __forceinline__ static float ld(const _Float16* p) { return from(*p); }
    __device__ __forceinline__ static void st(_Float16* p, float v) { *p = to(v); }
    __device__ __forceinline__ static float ldnt(const _Float16* p) { return from(__builtin_nontemporal_load(p)); }
    __device__ __forceinline__ static void stnt(_Float16* p, float v) { __builtin_nontemporal_store(to(v), p); }
};
template <>
struct Conv<bf16_t, float> {
    __device__ __forceinline__ static float from(bf16_t v) { return __uint_as_float((uint32_t)v.bits << 16); }
    __device__ __forceinline__ static bf16_t to(float v) {
        uint32_t u = __float_as_uint(v);
        bf16_t r;
        if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) {
            r.bits = (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
        } else {
            u += 0x7fffu + ((u >> 16) & 1u);          // round to nearest even
            r.bits = (uint16_t)(u >> 16);
        }
        return r;
    }
    __device__ __forceinline__ static float ld(const bf16_t* p) { return from(*p); }
    __device__ __forceinline__ static void st(bf16_t* p, float v) { *p = to(v); }
    __device__ __forceinline__ static float ldnt(const bf16_t* p) {
        bf16_t r;
        r.bits = __builtin_nontemporal_load(&p->bits);
        return from(r);
    }
    __device__ __forceinline__ static void stnt(bf16_t* p, float v) { __builtin_nontemporal_store(to(v).bits, &p->bits); }
};

// Two adjacent 16-bit values as one 32-bit access (the address is 4-byte
// aligned): a 64-lane wave then moves 256 bytes per load or store, the rate
// at which the memory path streams, instead of 128.
template <typename S>
struct Pair;
template <>
struct Pair<_Float16> {
    __device__ __forceinline__ static void split(uint32_t w, float& a, float& b) {
        a = (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu));
        b = (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16));
    }
    __device__ __forceinline__ static uint32_t join(float a, float b) {
        return (uint32_t)__builtin_bit_cast(uint16_t, to_f16(a)) |
               ((uint32_t)__builtin_bit_cast(uint16_t, to_f16(b)) << 16);
    }
};
template <>
struct Pair<bf16_t> {
    __device__ __forceinline__ static void split(uint32_t w, float& a, float& b) {
        a = __uint_as_float(w << 16);
        b = __uint_as_float(w & 0xffff0000u);
    }
    __device__ __forceinline__ static uint32_t join(float a, float b) {
        using C = Conv<bf16_t, float>;
        return (uint32_t)C::to(a).bits | ((uint32_t)C::to(b).bits << 16);
    }
};
template <typename S>
__device__ __forceinline__ void ld2(const S* p, float& a, float& b) {
    Pair<S>::split(*reinterpret_cast<const uint32_t*>(p), a, b);
}
template <typename S>
__device__ __forceinline__ void ldnt2(const S* p, float& a, float& b) {
    Pair<S>::split(__builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(p)), a, b);
}
template <typename S>
__device__ __forceinline__ void stnt2(S* p, float a, float b) {
    __builtin_nontemporal_store(Pair<S>::join(a, b), reinterpret_cast<uint32_t*>(p));
}

// FP32 storage: a lane's two points as one 8-byte access (z - m even is
// 8-byte aligned: z = m starts a 128-byte line)
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void ld2(const float* p, float& a, float& b) {
    const f32x2_t v = *reinterpret_cast<const f32x2_t*>(p);
    a = v.x;
    b = v.y;
}
__device__ __forceinline__ void ldnt2(const float* p, float& a, float& b) {
    const f32x2_t v = __builtin_nontemporal_load(reinterpret_cast<const f32x2_t*>(p));
    a = v.x;
    b = v.y;
}
__device__ __forceinline__ void stnt2(float* p, float a, float b) {
    f32x2_t v;
    v.x = a;
    v.y = b;
    __builtin_nontemporal_store(v, reinterpret_cast<f32x2_t*>(p));
}

// ---------------------------------------------------------------- geometry
struct Geom {
    int nx, ny, nz;   // dims incl. ghosts
    int m;            // margin = half order
    int pitch;        // row pitch in elements
    int zoff;         // elements in front of z = 0
    int64_t sy, sx;   // strides (sy = pitch, sx = ny*pitch)
    __host__ __device__ int64_t at(int x, int y, int z) const { return (int64_t)x * sx + (int64_t)y * sy + z; }
};

// per-field, per-level typed pointers (already offset by zoff)
template <typename S>
struct FieldPtrs {
    S* f[NF][2];
};

template <typename R>
struct MediaPtrs {
    const R* m[OPESCI_M_COUNT];  // padded layout like the fields; null when constant
};

// coefficients, converted to the compute type
template <typename R>
struct Coefs {
    R cn[3][3][MAXM];
    R cs[3][2][MAXM];
    R cv[3][3][MAXM];
};

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define OPESCI_HIP_TRY(call)                                                                   \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ::opesci::fail(OPESCI_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------- launch wrappers
// interior kernels (interior.hip); variant 0 = LDS/register-queue sweep, 1 = plain gather
struct InteriorLaunch {
    int variant;
    int tile_y;
    int x_chunk;
};

// Up to four (x, y) boxes of interior rows, all of z, that one launch of the
// pair sweeps covers, blocks numbered box by box (the launch fills first,
// tyn and xc from the tile shape).  The whole interior is one box; the
// rows around the fused step's box are four.
struct Boxes {
    int n = 0;
    int first[5] = {};                     // first logical block of each box, first[n] = blocks
    int x0[4] = {}, x1[4] = {}, y0[4] = {}, y1[4] = {};
    int tyn[4] = {}, xc[4] = {};           // y tiles and x chunk of each box
};


// Free-surface statements of the stress phase on the z faces, applied by the
// FP32 pair stress sweep in place, for the deep rows x, y in [F, dim - F)
// (where they commute with every other statement of the phase, so they may
// follow the interior update directly).  Per face (b = m low, dim - m - 1
// high) the statements have three fixed forms, which the host checks
// against the lowered boundary program before enabling them:
//   zero      Tzz(b) = 0
//   image     Tzz(b -+ k) = c Tzz(b +- k), k = 1..M-1;
//             Tyz, Txz(b - 1 - k) = c T(b + k) (low),  T(b + k) = c T(b - 1 - k) (high), k = 0..M-1
//   recompute Txx, Tyy(b) = sum of 9 terms in the canonical order
//             U(x-2..x+1), V(y-2..y+1), T(t0) (Levander, half order 2),
//             literal coefficients (constant media).
// Values, term order and fma accumulation equal bc_point's, so the fields
// are those of the boundary launches bit for bit.
// coefficient block in device memory, staged into LDS once per block
struct ZCoef {
    float rc[2][2][9];     // recompute literals [low, high][Txx, Tyy][term]
    float zc[2][3][MAXM];  // image coefficients [low, high][Tzz, Tyz, Txz][k]
};
constexpr int ZCOEF_WORDS = (int)(sizeof(ZCoef) / 4);
static_assert(sizeof(ZCoef) % 4 == 0, "ZCoef is copied as 32-bit words");
// scalar part, passed by value
struct ZFaceArgs {
    int F = 0;        // 0: off
    int rec = 0;      // recompute statements present
    int n1 = 0, npts = 0;
    const ZCoef* coef = nullptr;
};

// `boxes` (x0..y1 and n set) restricts the update to those rows; it needs
// the pair sweeps (FP32 arithmetic, kernel variant 0)
hipError_t launch_stress(const Geom& g, int real_bytes, int storage, bool het, void* fields, const void* media,
                         const void* coefs, int t0, int t1, const InteriorLaunch& cfg, hipStream_t s,
                         const ZFaceArgs* zf = nullptr, const void* pool = nullptr, const Boxes* boxes = nullptr);
// whether launch_stress applies ZFaceArgs for this configuration
bool stress_applies_zfaces(int real_bytes, bool het, const InteriorLaunch& cfg);
hipError_t launch_velocity(const Geom& g, int real_bytes, int storage, bool het, void* fields, const void* media,
                           const void* coefs, int t0, int t1, const InteriorLaunch& cfg, hipStream_t s,
                           const Boxes* boxes = nullptr);
// Fused step (both half steps in one x sweep) over the rows of box (one box,
// x0..y1 set) for FP32 arithmetic on constant media with the deep z faces
// in the sweep; x_chunk 0 picks the chunk.  fused_step_builds(m): whether
// it is built for half order m.
hipError_t launch_step_fused(const Geom& g, int storage, void* fields, const void* coefs, int t0, int t1,
                             const ZFaceArgs& zf, const Boxes& box, int x_chunk, hipStream_t s);
bool fused_step_builds(int m);

// auxiliary kernels (aux.hip)
// Boundary program as the device runs it: one self-contained descriptor per
// statement (fixed stride, terms inlined), grouped by hazard level, so a
// thread does one descriptor load and then independent value loads.  Reads
// are the destination's linear index plus a displacement.  Coefficients are
// float literals (the generator rounds every folded constant to float), so a
// float holds them exactly in either precision.
constexpr int BC_MAXT = 16;
struct BcTermDev {
    int32_t fl;      // field*2 + level
    int32_t disp;    // source - destination linear offset
    float coef;
    int32_t prog;    // -1 or offset of a media program (length in prog_len)
    int32_t cslot;   // -1 or the term's slot in the statement's block of the coefficient pool
    int32_t dterm0;  // adjoint: the term's derivative programs [dterm0, dterm0 + ndterm)
    int32_t ndterm;
};
struct BcDesc {
    int32_t fl;      // destination field*2 + level
    int32_t disp;    // plane * stride(axis)
    int32_t nterms;
    int32_t axis;
    int32_t lo0, lo1, n1, npts;
    int32_t cpool;   // first pool element of this statement's per-point coefficients
    int32_t prog_len[BC_MAXT];
    BcTermDev t[BC_MAXT];
};
struct BcLevel {
    int first, count, maxpts;   // descriptors [first, first+count)
};
// Media-dependent coefficients are time-invariant: bc_coef evaluates every
// program once into the pool (pool[cpool + cslot*npts + point], compute
// type R), and the step launches read them from there.
hipError_t launch_bc(const Geom& g, int real_bytes, int storage, void* fields, const BcDesc* d_desc, const void* pool,
                     const BcLevel& lvl, int t0, int t1, hipStream_t s);
hipError_t launch_bc_coef(const Geom& g, int real_bytes, const void* media, const BcDesc* d_desc,
                          const OpesciExprOp* d_code, void* pool, int first, int count, int maxpts, hipStream_t s);
// adjoint (adjoint.hip); native storage, compute type given by real_bytes
// scratch: six field-sized arrays, zero outside the interior (null selects
// the simple gather kernels, as does cfg.variant == 1).  bar is the host-side
// image of the adjoint fields' level pointers: with the scratch path the
// launch swaps the two levels of the updated fields in it (the identity part
// of the transpose) and fixes up only `cells` (distinct entries
// (offset << 5) | exchange << 4 | field, built from the cells the step's
// boundary statements read).
struct AdjCells {
    const int64_t* d = nullptr;
    int n = 0;
};
hipError_t launch_adj_stress(const Geom& g, int real_bytes, void* fields, void* bar, const void* media, void* grad,
                             const void* coefs, void* const* scratch, const AdjCells& cells, const InteriorLaunch& cfg,
                             int t0, int t1, hipStream_t s);
hipError_t launch_adj_velocity(const Geom& g, int real_bytes, void* fields, void* bar, const void* media, void* grad,
                               const void* coefs, void* const* scratch, const AdjCells& cells,
                               const InteriorLaunch& cfg, int t0, int t1, hipStream_t s);
// one boundary level, each overwritten value first saved into `log`
// (descriptor-major, lv.maxpts slots per descriptor); launch_adj_bc restores them
hipError_t launch_bc_logged(const Geom& g, int real_bytes, void* fields, const BcDesc* desc, const void* pool,
                            const BcLevel& lv, void* log, int t0, int t1, hipStream_t s);
// Fixed-order accumulation of one level's transposes: targets in ascending
// cell order, each summing its contributions in (statement, point, term)
// order from a buffer the level's points fill first, so results do not
// depend on scheduling.  Device arrays; start[t]..start[t+1] are target t's
// entries.
struct AdjGather {
    int ntgt = 0;
    const int32_t* start = nullptr;   // ntgt + 1
    const int32_t* key = nullptr;     // target field*2 + level
    const int64_t* off = nullptr;     // target linear offset
    const int32_t* src = nullptr;     // per entry: slot in the buffer
    const int32_t* cinfo = nullptr;   // per entry: pool index, or -(1 + descriptor*BC_MAXT + term)
    const int32_t* boff = nullptr;    // per descriptor: its first buffer slot
    void* buf = nullptr;              // the level's slots (compute type)
};
// the level's transposes; its first pass also restores the forward values
// the level overwrote from `log` (the level's undo log, as bc_log wrote it)
hipError_t launch_adj_bc(const Geom& g, int real_bytes, void* fields, void* bar, const BcDesc* desc, const void* pool,
                         void* gpool, const BcLevel& lv, const AdjGather& ga, const void* log, int t0, int t1,
                         hipStream_t s);
// media gradients through derivative program q of term k of descriptor d
// (one launch per program, in a fixed order)
hipError_t launch_adj_coef_media(const Geom& g, int real_bytes, const void* media, void* grad, const BcDesc* desc,
                                 const OpesciExprOp* code, const OpesciBcDTerm* dterms, const void* gpool, int d,
                                 int npts, int k, int q, hipStream_t s);
// receiver residuals, the receivers of one cell (group g: order[start[g] ..
// start[g+1])) summed in receiver order
hipError_t launch_adj_receivers(const Geom& g, int real_bytes, void* fields, void* bar, int t1, int ngroups,
                                const int32_t* start, const int64_t* cell, const int32_t* order, const float* obs,
                                hipStream_t s);
hipError_t launch_set_step(int* step, int value, hipStream_t s);
// dst level td := src level ts at a list of distinct cells ((offset << 4) |
// field), both FieldPtrs images of the compute type
hipError_t launch_copy_cells(int real_bytes, void* dst, int td, void* src, int ts, const int64_t* cells, int n,
                             hipStream_t s);
// mode 0: buf[0..2] = Txx, Tyy, Tzz at (x, y, z) of level t; mode 1: the reverse
hipError_t launch_src_cells(const Geom& g, int real_bytes, void* fields, int t, int x, int y, int z, void* buf,
                            int mode, hipStream_t s);

// Analytic solution programs, split: every maximal subtree of a field's
// program that reads at most one coordinate and no media is evaluated once
// per index along that axis (slot: code [prog, prog + len) into tab[tab +
// i], i < dim(axis)) and the per-point program [off[f], off[f] + len[f])
// reads it back through OPX_TABLE.  The same double operations run in the
// same order, so the values equal per-point evaluation bit for bit.
constexpr int OPX_TABLE = 64;
struct SolSlot {
    int32_t field, axis, prog, len, tab;
};
struct SolProg {
    const OpesciExprOp* code = nullptr;  // per-point programs, then the slots' programs
    int off[NF] = {}, len[NF] = {};
    int tab_base[NF] = {}, tab_len[NF] = {};  // field f's tables: tab[tab_base[f] ..], OPX_TABLE offsets relative to it
    int shallow[NF] = {};                      // field f's point program never exceeds stack depth 3
    const SolSlot* slots = nullptr;
    int nslots = 0;
    int maxdim = 0;
    double* tab = nullptr;               // the tables (device, written by every launch below)
};
hipError_t launch_initial(const Geom& g, int real_bytes, int storage, void* fields, const void* media,
                          const SolProg& sol, const double* spacing, double dt, hipStream_t s);
hipError_t launch_l2(const Geom& g, int real_bytes, int storage, void* fields, const void* media,
                     const SolProg& sol, const double* spacing, int level, const double t_field[NF],
                     double* d_partial, int partial_cap, double* out_host, hipStream_t s);
// the same L2 with its partial sums in a fixed library-owned device array
// and the norms written to d_out (9 doubles, device memory), stream-ordered,
// no allocation and no host synchronisation; the block count equals
// launch_l2's, so the two agree
int l2_noalloc_partials();
hipError_t launch_l2_noalloc(const Geom& g, int real_bytes, int storage, void* fields, const void* media,
                             const SolProg& sol, const double* spacing, int level, const double t_field[NF],
                             double* d_out, hipStream_t s);
hipError_t launch_media(const Geom& g, int real_bytes, const float* d_rho, const float* d_vp, const float* d_vs,
                        void* const* d_media, hipStream_t s);
hipError_t launch_pack(const Geom& g, int real_bytes, int storage, const void* dense, void* field, hipStream_t s);
hipError_t launch_unpack(const Geom& g, int real_bytes, int storage, const void* field, void* dense, hipStream_t s);
hipError_t launch_pack_media(const Geom& g, int real_bytes, const void* dense, void* media, hipStream_t s);
hipError_t launch_zero(void* p, size_t bytes, hipStream_t s);
hipError_t launch_record_source(const Geom& g, int real_bytes, int storage, void* fields, int t1, int nrec,
                                const int32_t* d_rec, void* d_trace, int src_on, const int32_t* src_idx,
                                const float* d_wavelet, int* d_step, int ntsteps, hipStream_t s);

}  // namespace opesci
