// Device-side boundary statement evaluation: expression bytecode (media
// programs, analytic solutions) and the per-point statement of the boundary
// launches.
#pragma once

#include "common.h"

namespace opesci {

// ------------------------------------------------------------------ bytecode
// Stack machine over OpesciExprOp (opcodes in include/opesci_hip.h),
// evaluated in double.  MEDIA reads media a[0] at anchor + (a[1],a[2],a[3]).
// OPX_TABLE (library-internal, in programs the library rewrites) reads
// tabs[a[0] + anchor index along axis a[1]].
template <typename R>
__device__ double vm_eval(const OpesciExprOp* code, int n, const double coord[3], double t,
                          const MediaPtrs<R>& md, const Geom& g, int ax, int ay, int az,
                          const double* __restrict__ tabs = nullptr) {
    double st[16];
    int sp = 0;
    for (int i = 0; i < n; ++i) {
        const OpesciExprOp o = code[i];
        switch (o.op) {
            case OPESCI_OP_CONST: st[sp++] = o.val; break;
            case OPESCI_OP_COORD: st[sp++] = coord[o.a[0]]; break;
            case OPESCI_OP_TIME: st[sp++] = t; break;
            case OPESCI_OP_MEDIA: {
                const R* p = md.m[o.a[0]];
                st[sp++] = p ? (double)p[g.at(ax + o.a[1], ay + o.a[2], az + o.a[3])] : 0.0;
                break;
            }
            case OPESCI_OP_ADD: --sp; st[sp - 1] = st[sp - 1] + st[sp]; break;
            case OPESCI_OP_SUB: --sp; st[sp - 1] = st[sp - 1] - st[sp]; break;
            case OPESCI_OP_MUL: --sp; st[sp - 1] = st[sp - 1] * st[sp]; break;
            case OPESCI_OP_DIV: --sp; st[sp - 1] = st[sp - 1] / st[sp]; break;
            case OPESCI_OP_NEG: st[sp - 1] = -st[sp - 1]; break;
            case OPESCI_OP_POWC: st[sp - 1] = pow(st[sp - 1], o.val); break;
            case OPESCI_OP_SQRT: st[sp - 1] = sqrt(st[sp - 1]); break;
            case OPESCI_OP_SIN: st[sp - 1] = sin(st[sp - 1]); break;
            case OPESCI_OP_COS: st[sp - 1] = cos(st[sp - 1]); break;
            case OPESCI_OP_EXP: st[sp - 1] = exp(st[sp - 1]); break;
            case OPESCI_OP_LOG: st[sp - 1] = log(st[sp - 1]); break;
            case OPESCI_OP_ROUNDF: st[sp - 1] = (double)(float)st[sp - 1]; break;
            case OPX_TABLE: st[sp++] = tabs[o.a[0] + (o.a[1] == 0 ? ax : (o.a[1] == 1 ? ay : az))]; break;
            default: break;
        }
    }
    return sp > 0 ? st[0] : 0.0;
}

// The same machine with the top two stack entries in fixed registers and
// only deeper entries in the indexed array (the point kernels' fallback for
// programs deeper than vm_eval_k below takes): the same double operations
// in the same order as vm_eval.
template <typename R>
__device__ double vm_eval_tos(const OpesciExprOp* code, int n, const double coord[3], double t,
                              const MediaPtrs<R>& md, const Geom& g, int ax, int ay, int az,
                              const double* __restrict__ tabs) {
    double t0 = 0.0, t1 = 0.0, rest[14];
    int depth = 0;
    for (int i = 0; i < n; ++i) {
        const OpesciExprOp o = code[i];
        const int op = o.op;
        if (op <= OPESCI_OP_MEDIA || op == OPX_TABLE) {
            double v;
            if (op == OPX_TABLE)
                v = tabs[o.a[0] + (o.a[1] == 0 ? ax : (o.a[1] == 1 ? ay : az))];
            else if (op == OPESCI_OP_CONST)
                v = o.val;
            else if (op == OPESCI_OP_COORD)
                v = coord[o.a[0]];
            else if (op == OPESCI_OP_TIME)
                v = t;
            else {
                const R* p = md.m[o.a[0]];
                v = p ? (double)p[g.at(ax + o.a[1], ay + o.a[2], az + o.a[3])] : 0.0;
            }
            if (depth >= 2) rest[depth - 2] = t1;
            t1 = t0;
            t0 = v;
            ++depth;
        } else if (op <= OPESCI_OP_DIV) {
            t0 = op == OPESCI_OP_ADD ? t1 + t0 : (op == OPESCI_OP_SUB ? t1 - t0 : (op == OPESCI_OP_MUL ? t1 * t0 : t1 / t0));
            --depth;
            t1 = depth >= 2 ? rest[depth - 2] : 0.0;
        } else {
            switch (op) {
                case OPESCI_OP_NEG: t0 = -t0; break;
                case OPESCI_OP_POWC: t0 = pow(t0, o.val); break;
                case OPESCI_OP_SQRT: t0 = sqrt(t0); break;
                case OPESCI_OP_SIN: t0 = sin(t0); break;
                case OPESCI_OP_COS: t0 = cos(t0); break;
                case OPESCI_OP_EXP: t0 = exp(t0); break;
                case OPESCI_OP_LOG: t0 = log(t0); break;
                case OPESCI_OP_ROUNDF: t0 = (double)(float)t0; break;
                default: break;
            }
        }
    }
    return depth > 0 ? (depth == 1 ? t0 : rest[0]) : 0.0;
}

// K points of one program at once (the same ops, decoded once for all K),
// for programs whose stack never exceeds depth 3: t0/t1/t2 are the top
// three entries of each point's stack.  Per point, the same double
// operations in the same order as vm_eval.
template <typename R, int K>
__device__ __forceinline__ void vm_eval_k(const OpesciExprOp* __restrict__ code, int n, const double cx[K],
                                          const double cy[K], double cz, double t, const MediaPtrs<R>& md,
                                          const Geom& g, const int ax[K], const int ay[K], int az,
                                          const double* __restrict__ tabs, double out[K]) {
    double t0[K], t1[K], t2[K];
#pragma unroll
    for (int k = 0; k < K; ++k) t0[k] = t1[k] = t2[k] = 0.0;
    for (int i = 0; i < n; ++i) {
        const OpesciExprOp o = code[i];
        const int op = o.op;
        if (op <= OPESCI_OP_MEDIA || op == OPX_TABLE) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                double v;
                if (op == OPX_TABLE)
                    v = tabs[o.a[0] + (o.a[1] == 0 ? ax[k] : (o.a[1] == 1 ? ay[k] : az))];
                else if (op == OPESCI_OP_CONST)
                    v = o.val;
                else if (op == OPESCI_OP_COORD)
                    v = o.a[0] == 0 ? cx[k] : (o.a[0] == 1 ? cy[k] : cz);
                else if (op == OPESCI_OP_TIME)
                    v = t;
                else {
                    const R* p = md.m[o.a[0]];
                    v = p ? (double)p[g.at(ax[k] + o.a[1], ay[k] + o.a[2], az + o.a[3])] : 0.0;
                }
                t2[k] = t1[k];
                t1[k] = t0[k];
                t0[k] = v;
            }
        } else {
            switch (op) {
#define OPESCI_K_LOOP(expr) _Pragma("unroll") for (int k = 0; k < K; ++k) t0[k] = (expr); break
#define OPESCI_K_POP(expr) _Pragma("unroll") for (int k = 0; k < K; ++k) { t0[k] = (expr); t1[k] = t2[k]; } break
                case OPESCI_OP_ADD: OPESCI_K_POP(t1[k] + t0[k]);
                case OPESCI_OP_SUB: OPESCI_K_POP(t1[k] - t0[k]);
                case OPESCI_OP_MUL: OPESCI_K_POP(t1[k] * t0[k]);
                case OPESCI_OP_DIV: OPESCI_K_POP(t1[k] / t0[k]);
                case OPESCI_OP_NEG: OPESCI_K_LOOP(-t0[k]);
                case OPESCI_OP_POWC: OPESCI_K_LOOP(pow(t0[k], o.val));
                case OPESCI_OP_SQRT: OPESCI_K_LOOP(sqrt(t0[k]));
                case OPESCI_OP_SIN: OPESCI_K_LOOP(sin(t0[k]));
                case OPESCI_OP_COS: OPESCI_K_LOOP(cos(t0[k]));
                case OPESCI_OP_EXP: OPESCI_K_LOOP(exp(t0[k]));
                case OPESCI_OP_LOG: OPESCI_K_LOOP(log(t0[k]));
                case OPESCI_OP_ROUNDF: OPESCI_K_LOOP((double)(float)t0[k]);
#undef OPESCI_K_LOOP
#undef OPESCI_K_POP
                default: break;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = t0[k];
}

// One boundary statement at point p of its loop: dst = sum_k coef_k * src_k,
// terms loaded independently and accumulated with fma in term order.  A
// term's coefficient is its literal or, for media-dependent ones (cslot >=
// 0), the per-point value precomputed in the pool.
// (i0, i1) are the loop indices of the two axes other than D->axis, in
// increasing axis order.
template <typename R, typename S>
__device__ __forceinline__ void bc_point(const Geom& g, const FieldPtrs<S>& F, const BcDesc* __restrict__ D,
                                         const R* __restrict__ pool, int p, int i0, int i1, int t0, int t1) {
    using C = Conv<S, R>;
    const int axis = D->axis, nt = D->nterms;
    const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
    const int64_t so0 = o0 == 0 ? g.sx : g.sy, so1 = o1 == 1 ? g.sy : 1;
    const int64_t dst = (int64_t)i0 * so0 + (int64_t)i1 * so1 + D->disp;
    R v[BC_MAXT];
#pragma unroll
    for (int k = 0; k < BC_MAXT; ++k)
        if (k < nt) {
            const int fl = D->t[k].fl;
            v[k] = C::ld(F.f[fl >> 1][(fl & 1) ? t1 : t0] + dst + D->t[k].disp);
        }
    R acc = (R)0;
#pragma unroll
    for (int k = 0; k < BC_MAXT; ++k)
        if (k < nt) {
            const R coef = D->t[k].cslot >= 0 ? pool[D->cpool + (int64_t)D->t[k].cslot * D->npts + p]
                                              : (R)D->t[k].coef;
            acc = fma(coef, v[k], acc);
        }
    const int fl = D->fl;
    C::st(F.f[fl >> 1][(fl & 1) ? t1 : t0] + dst, acc);
}

}  // namespace opesci
