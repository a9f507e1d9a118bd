// Device-side boundary statement evaluation: expression bytecode (media
// programs, analytic solutions) and the per-point statement of the boundary
// launches.
#pragma once

#include "common.h"

namespace opesci {

// ------------------------------------------------------------------ bytecode
// Stack machine over OpesciExprOp (opcodes in include/opesci_hip.h),
// evaluated in double.  MEDIA reads media a[0] at anchor + (a[1],a[2],a[3]).
template <typename R>
__device__ double vm_eval(const OpesciExprOp* code, int n, const double coord[3], double t,
                          const MediaPtrs<R>& md, const Geom& g, int ax, int ay, int az) {
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
            default: break;
        }
    }
    return sp > 0 ? st[0] : 0.0;
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
