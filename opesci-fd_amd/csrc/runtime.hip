// Context management, time loop and the C ABI of libopesci_hip.so.
//
// A context owns one model on the current device: 9 fields x 2 time levels
// in the padded layout of common.h, media arrays, the boundary program and
// the analytic-solution bytecode.  The time loop follows the reference
// (opesci/templates/staggered3d_tmpl.py:40-58): t0 = step % 2, t1 = 1 - t0;
// stress kernel, stress boundary launches, velocity kernel, velocity boundary
// launches [, receivers + source].  Two steps are captured in one hipGraph
// and replayed.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <memory>
#include <new>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "common.h"

namespace opesci {

static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }
int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

}  // namespace opesci

using namespace opesci;

struct OpesciContext {
    OpesciModelDesc desc;
    Geom g;
    int real_bytes = 4, storage = 0, elem = 4;
    bool het = false;
    hipStream_t stream = nullptr;
    int device = 0;                 // the device current at creation
    void* base[NF][2] = {};
    void* fptr[NF][2] = {};         // FieldPtrs<S> image
    void* mbase[OPESCI_M_COUNT] = {};
    void* mptr[OPESCI_M_COUNT] = {};  // MediaPtrs<R> image
    Coefs<float> cf;
    Coefs<double> cd;
    std::vector<OpesciBcStmt> stmts;  // sorted by (phase, sched, axis, ranges)
    std::vector<BcLevel> levels[5];    // phases 0..3 + 4 (deep z-face rows of the stress phase, one launch)
    BcDesc* d_desc = nullptr;
    std::vector<BcDesc> h_desc;       // host copy of the descriptors (adjoint cell lists)
    void* d_pool = nullptr;           // precomputed media-dependent boundary coefficients (compute type)
    int64_t pool_elems = 0;
    OpesciBcDTerm* d_dterms = nullptr;  // derivative programs of those coefficients (adjoint)
    // adjoint state (allocated by the first opesci_ctx_gradient)
    void* abase[NF][2] = {};
    void* aptr[NF][2] = {};           // AdjPtrs<R> image
    void* gbase[OPESCI_M_COUNT] = {};
    void* gptr[OPESCI_M_COUNT] = {};  // MediaPtrs<R> image of the media gradients
    void* sbase[6] = {};              // adjoint scratch (weighted adjoint fields)
    void* sptr[6] = {};
    void* d_gpool = nullptr;
    int64_t* d_tcells = nullptr;      // adjoint level-swap cells of the stresses
    int64_t* d_vcells = nullptr;      // ... and of the velocities
    int64_t* d_scells = nullptr;      // cells boundary statements read at level t1 outside the interior
    int n_tcells = 0, n_vcells = 0, n_scells = 0;
    void* d_log = nullptr;            // undo log of one step's boundary levels
    std::vector<int64_t> log_off[5];  // per phase and level
    char* d_ckpt = nullptr;           // checkpoint states, 18 field levels each
    char* d_seg = nullptr;            // states of one checkpoint segment
    size_t ckpt_cap = 0, seg_cap = 0;  // bytes
    int64_t log_step = 0;             // undo-log elements of one step
    int64_t log_cap = 0;              // undo-log steps allocated
    int64_t srclog_cap = 0;           // source-cell log steps allocated
    void* d_srclog = nullptr;         // Txx/Tyy/Tzz at the source before and after injection, per step
    float* d_obs = nullptr;
    OpesciExprOp* d_code = nullptr;
    OpesciExprOp* d_sol = nullptr;    // the analytic programs, split (SolProg)
    SolSlot* d_solslots = nullptr;
    double* d_soltab = nullptr;
    SolProg sol;
    double* d_partial = nullptr;
    int partial_cap = 0;
    void* d_stage = nullptr;          // dense staging buffer (one field level, real_t)
    int32_t* d_rec = nullptr;
    void* d_trace = nullptr;          // [ntsteps][nrec][4] of the arithmetic type
    float* d_wavelet = nullptr;
    int* d_step = nullptr;
    int nrec = 0;
    int src_ntr = 0;                  // source traces: 0, 1 (all of Txx/Tyy/Tzz) or 3
    int32_t src_idx[3] = {0, 0, 0};
    hipGraphExec_t graph = nullptr;
    int steps = 0;
    bool initialised = false;
    size_t field_bytes = 0;
    InteriorLaunch cfg{0, 0, 0};
    ZFaceArgs zf;                   // deep z-face statements the pair stress sweep applies (F = 0: none)
    std::vector<OpesciBcTerm> h_terms;  // boundary terms as given (host)
    int fused_F = 0;                // fused step: box [F, dim - F) in x and y (0: off)
    Boxes fbox, frame;              // its box and the four boxes of rows around it
    int grad_every = 0;             // checkpoint interval of the last forward pass (0: none pending)
    bool grad_keep_t0 = false;
    ZCoef* d_zcoef = nullptr;       // their coefficient block (device)
    // adjoint: fixed-order gathers of the boundary transposes per level
    // (phases 2-4) and the receivers grouped by cell
    std::vector<OpesciBcDTerm> h_dterms;
    std::vector<int32_t> h_rec;
    std::vector<AdjGather> gat_bc[5];
    char* d_gather = nullptr;       // their index arrays, one allocation
    void* d_abuf = nullptr;         // the per-level slot buffer (compute type)
    int nrgroups = 0;
    const int32_t* rg_start = nullptr;
    const int64_t* rg_cell = nullptr;
    const int32_t* rg_order = nullptr;

    const void* coefs() const { return real_bytes == 8 ? (const void*)&cd : (const void*)&cf; }
};

namespace {

OpesciContext* g_default = nullptr;
OpesciModelDesc g_loaded;
bool g_have_loaded = false;
std::vector<char> g_loaded_blob;  // deep copy of the arrays the loaded desc points to

template <typename T>
int upload(T** dst, const T* src, size_t n, hipStream_t s) {
    if (n == 0) {
        *dst = nullptr;
        return 0;
    }
    OPESCI_HIP_TRY(hipMalloc((void**)dst, n * sizeof(T)));
    OPESCI_HIP_TRY(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
    OPESCI_HIP_TRY(hipStreamSynchronize(s));  // src may be a temporary
    return 0;
}

// Stack discipline of a bytecode program: every op sees the operands it
// pops, the stack never exceeds the device evaluator's 16 slots and ends with
// exactly one value.
bool program_ok(const OpesciExprOp* code, int n) {
    int depth = 0;
    for (int i = 0; i < n; ++i) {
        const int op = code[i].op;
        if (op == OPESCI_OP_CONST || op == OPESCI_OP_COORD || op == OPESCI_OP_TIME || op == OPESCI_OP_MEDIA) {
            if (++depth > 16) return false;
        } else if (op >= OPESCI_OP_ADD && op <= OPESCI_OP_DIV) {
            if (depth < 2) return false;
            --depth;
        } else if (op >= OPESCI_OP_NEG && op <= OPESCI_OP_ROUNDF) {
            if (depth < 1) return false;
        } else {
            return false;
        }
        if (op == OPESCI_OP_COORD && (code[i].a[0] < 0 || code[i].a[0] > 2)) return false;
    }
    return depth == 1;
}

int validate(const OpesciModelDesc* d) {
    if (!d) return fail(OPESCI_EINVAL, "null model description");
    if (d->abi_version != OPESCI_ABI_VERSION) return fail(OPESCI_EINVAL, "model description ABI version mismatch");
    if (d->space_order < 2 || d->space_order > 2 * MAXM || d->space_order % 2)
        return fail(OPESCI_EUNSUPPORTED, "space order must be 2, 4, 6 or 8");
    const int m = d->space_order / 2;
    for (int a = 0; a < 3; ++a)
        if (d->dims[a] < 2 * m + 2 || d->dims[a] > 1 << 15) return fail(OPESCI_EINVAL, "grid dims out of range");
    if (d->real_bytes != 4 && d->real_bytes != 8) return fail(OPESCI_EINVAL, "real_bytes must be 4 or 8");
    if (d->storage < 0 || d->storage > 2) return fail(OPESCI_EINVAL, "unknown storage type");
    if (d->storage != OPESCI_STORE_NATIVE && d->real_bytes != 4)
        return fail(OPESCI_EUNSUPPORTED, "16-bit storage needs FP32 compute");
    if (d->ntsteps < 0) return fail(OPESCI_EINVAL, "negative ntsteps");
    if (d->kernel_variant < 0 || d->kernel_variant > 2 || d->x_chunk < 0 ||
        (d->tile_y != 0 && d->tile_y != 4 && d->tile_y != 8 && d->tile_y != 16))
        return fail(OPESCI_EINVAL, "kernel_variant 0|1|2, tile_y 0|4|8|16, x_chunk >= 0");
    if (d->media_arrays && (!d->rho || !d->vp || !d->vs)) return fail(OPESCI_EINVAL, "media arrays missing");
    if (d->nstmts < 0 || d->nterms < 0 || d->ncode < 0 || d->nsol < 0)
        return fail(OPESCI_EINVAL, "negative table size");
    if ((d->nstmts && !d->stmts) || (d->nterms && !d->terms) || (d->ncode && !d->bc_code) || (d->nsol && !d->sol_code))
        return fail(OPESCI_EINVAL, "null table pointer");
    // every boundary access must stay inside the grid
    for (int i = 0; i < d->nstmts; ++i) {
        const OpesciBcStmt& s = d->stmts[i];
        if (s.field < 0 || s.field >= NF || s.axis < 0 || s.axis > 2 || s.level < 0 || s.level > 1 || s.phase < 0 ||
            s.phase > 4 || s.sched < 0 || s.term0 < 0 || s.nterms < 0 || s.term0 + s.nterms > d->nterms)
            return fail(OPESCI_EINVAL, "malformed boundary statement");
        const int o[2] = {s.axis == 0 ? 1 : 0, s.axis == 2 ? 1 : 2};
        if (s.plane < 0 || s.plane >= d->dims[s.axis]) return fail(OPESCI_EINVAL, "boundary plane outside the grid");
        for (int k = 0; k < 2; ++k)
            if (s.hi[k] > s.lo[k] && (s.lo[k] < 0 || s.hi[k] > d->dims[o[k]]))
                return fail(OPESCI_EINVAL, "boundary loop outside the grid");
        for (int t = s.term0; t < s.term0 + s.nterms; ++t) {
            const OpesciBcTerm& tm = d->terms[t];
            if (tm.field < 0 || tm.field >= NF || tm.level < 0 || tm.level > 1 || tm.plane < 0 ||
                tm.plane >= d->dims[s.axis])
                return fail(OPESCI_EINVAL, "malformed boundary term");
            for (int k = 0; k < 2; ++k)
                if (s.hi[k] > s.lo[k] && (s.lo[k] + tm.off[k] < 0 || s.hi[k] + tm.off[k] > d->dims[o[k]]))
                    return fail(OPESCI_EINVAL, "boundary term reads outside the grid");
            if (tm.prog >= 0) {
                if (tm.prog + tm.prog_len > d->ncode || !program_ok(d->bc_code + tm.prog, tm.prog_len))
                    return fail(OPESCI_EINVAL, "boundary program out of range or malformed");
                for (int c = tm.prog; c < tm.prog + tm.prog_len; ++c) {
                    const OpesciExprOp& op = d->bc_code[c];
                    if (op.op == OPESCI_OP_MEDIA) {
                        if (!d->media_arrays || op.a[0] < 0 || op.a[0] >= OPESCI_M_COUNT)
                            return fail(OPESCI_EINVAL, "media reference without media arrays");
                        for (int k = 0; k < 2; ++k)
                            if (s.hi[k] > s.lo[k] &&
                                (s.lo[k] + op.a[1 + o[k]] < 0 || s.hi[k] + op.a[1 + o[k]] > d->dims[o[k]]))
                                return fail(OPESCI_EINVAL, "media read outside the grid");
                        if (op.a[1 + s.axis] < 0 || op.a[1 + s.axis] >= d->dims[s.axis])
                            return fail(OPESCI_EINVAL, "media plane outside the grid");
                    }
                }
            }
        }
    }
    // adjoint derivative programs: ordered by term, valid media references
    // (their cells are among those the term's own program reads)
    if (d->ndterms < 0 || (d->ndterms && !d->dterms)) return fail(OPESCI_EINVAL, "bad derivative table");
    for (int i = 0; i < d->ndterms; ++i) {
        const OpesciBcDTerm& t = d->dterms[i];
        if (t.term < 0 || t.term >= d->nterms || (i > 0 && t.term < d->dterms[i - 1].term) || t.media < 0 ||
            t.media >= OPESCI_M_COUNT || !d->media_arrays || d->terms[t.term].prog < 0)
            return fail(OPESCI_EINVAL, "malformed derivative entry");
        if (t.prog < 0 || t.prog_len < 0 || t.prog + t.prog_len > d->ncode || !program_ok(d->bc_code + t.prog, t.prog_len))
            return fail(OPESCI_EINVAL, "derivative program out of range or malformed");
        bool found = false;
        const OpesciBcTerm& tm = d->terms[t.term];
        for (int c = tm.prog; c < tm.prog + tm.prog_len && !found; ++c) {
            const OpesciExprOp& op = d->bc_code[c];
            found = op.op == OPESCI_OP_MEDIA && op.a[0] == t.media && op.a[1] == t.off[0] && op.a[2] == t.off[1] &&
                    op.a[3] == t.off[2];
        }
        if (!found) return fail(OPESCI_EINVAL, "derivative of a media cell the coefficient does not read");
        for (int c = t.prog; c < t.prog + t.prog_len; ++c) {
            const OpesciExprOp& op = d->bc_code[c];
            if (op.op != OPESCI_OP_MEDIA) continue;
            bool known = false;
            for (int c2 = tm.prog; c2 < tm.prog + tm.prog_len && !known; ++c2) {
                const OpesciExprOp& o2 = d->bc_code[c2];
                known = o2.op == OPESCI_OP_MEDIA && o2.a[0] == op.a[0] && o2.a[1] == op.a[1] && o2.a[2] == op.a[2] &&
                        o2.a[3] == op.a[3];
            }
            if (!known) return fail(OPESCI_EINVAL, "derivative program reads a cell the coefficient does not");
        }
    }
    for (int f = 0; f < NF; ++f) {
        if (d->sol_off[f] < 0 || d->sol_len[f] < 0 || d->sol_off[f] + d->sol_len[f] > d->nsol)
            return fail(OPESCI_EINVAL, "analytic solution table out of range");
        if (d->sol_len[f] > 0 && !program_ok(d->sol_code + d->sol_off[f], d->sol_len[f]))
            return fail(OPESCI_EINVAL, "malformed analytic-solution program");
        for (int c = d->sol_off[f]; c < d->sol_off[f] + d->sol_len[f]; ++c) {
            const OpesciExprOp& op = d->sol_code[c];
            if (op.op == OPESCI_OP_MEDIA) {
                if (!d->media_arrays || op.a[0] < 0 || op.a[0] >= OPESCI_M_COUNT)
                    return fail(OPESCI_EINVAL, "media reference without media arrays");
                for (int k = 1; k < 4; ++k)
                    if (op.a[k] < -m || op.a[k] > m) return fail(OPESCI_EINVAL, "media offset out of range");
            }
        }
    }
    if (d->src_on != 0 && d->src_on != 1 && d->src_on != 3) return fail(OPESCI_EINVAL, "src_on must be 0, 1 or 3");
    if (d->src_on) {
        if (!d->src_wavelet) return fail(OPESCI_EINVAL, "source without wavelet");
        for (int a = 0; a < 3; ++a)
            if (d->src_idx[a] < 0 || d->src_idx[a] >= d->dims[a]) return fail(OPESCI_EINVAL, "source outside grid");
    }
    if (d->nrec < 0 || (d->nrec && !d->rec_idx)) return fail(OPESCI_EINVAL, "bad receivers");
    for (int r = 0; r < d->nrec; ++r)
        for (int a = 0; a < 3; ++a)
            if (d->rec_idx[3 * r + a] < 0 || d->rec_idx[3 * r + a] >= d->dims[a])
                return fail(OPESCI_EINVAL, "receiver outside grid");
    return 0;
}

// Sort statements by (phase, launch), build one self-contained descriptor
// per statement with the linear displacements the device uses (all indices
// were validated against the grid; the grid is checked to fit int32 offsets).
int build_program(OpesciContext* c, const OpesciBcTerm* terms, const OpesciBcDTerm* dterms, int ndterms,
                  int nterms) {
    // derivative programs per term (contiguous, ordered by term)
    std::vector<int> dfirst(std::max(nterms, 1), 0), dcount(std::max(nterms, 1), 0);
    for (int i = 0; i < ndterms; ++i) {
        if (dcount[dterms[i].term] == 0) dfirst[dterms[i].term] = i;
        ++dcount[dterms[i].term];
    }
    std::stable_sort(c->stmts.begin(), c->stmts.end(), [](const OpesciBcStmt& a, const OpesciBcStmt& b) {
        return a.phase != b.phase ? a.phase < b.phase : a.sched < b.sched;
    });
    const Geom& g = c->g;
    if ((int64_t)g.nx * g.sx >= ((int64_t)1 << 31)) return fail(OPESCI_EUNSUPPORTED, "grid too large for int32 offsets");
    const int64_t stride[3] = {g.sx, g.sy, 1};
    std::vector<BcDesc> descs;
    const int n = (int)c->stmts.size();
    for (int i = 0; i < n;) {
        const OpesciBcStmt& L = c->stmts[i];
        BcLevel lv{(int)descs.size(), 0, 0};
        int j = i;
        for (; j < n && c->stmts[j].phase == L.phase && c->stmts[j].sched == L.sched; ++j) {
            const OpesciBcStmt& a = c->stmts[j];
            if (a.nterms > BC_MAXT) return fail(OPESCI_EUNSUPPORTED, "boundary statement with more than 16 terms");
            const int n0 = std::max(0, a.hi[0] - a.lo[0]), n1 = std::max(0, a.hi[1] - a.lo[1]);
            if (n0 * n1 == 0) continue;
            BcDesc d;
            memset(&d, 0, sizeof(d));
            d.fl = a.field * 2 + a.level;
            d.disp = (int32_t)(a.plane * stride[a.axis]);
            d.nterms = a.nterms;
            d.axis = a.axis;
            d.lo0 = a.lo[0];
            d.lo1 = a.lo[1];
            d.n1 = n1;
            d.npts = n0 * n1;
            d.cpool = 0;
            int nslot = 0;
            const int o[2] = {a.axis == 0 ? 1 : 0, a.axis == 2 ? 1 : 2};
            for (int k = 0; k < a.nterms; ++k) {
                const OpesciBcTerm& tm = terms[a.term0 + k];
                d.t[k].fl = tm.field * 2 + tm.level;
                d.t[k].disp = (int32_t)((int64_t)(tm.plane - a.plane) * stride[a.axis] + (int64_t)tm.off[0] * stride[o[0]] +
                                        (int64_t)tm.off[1] * stride[o[1]]);
                d.t[k].coef = (float)tm.coef;
                if ((double)d.t[k].coef != tm.coef)
                    return fail(OPESCI_EINVAL, "boundary coefficient is not a float literal");
                d.t[k].prog = tm.prog;
                d.prog_len[k] = tm.prog_len;
                d.t[k].cslot = tm.prog >= 0 ? nslot++ : -1;
                d.t[k].dterm0 = dfirst[a.term0 + k];
                d.t[k].ndterm = dcount[a.term0 + k];
            }
            if (nslot > 0) {
                if (c->pool_elems + (int64_t)nslot * d.npts >= ((int64_t)1 << 31))
                    return fail(OPESCI_EUNSUPPORTED, "boundary coefficient pool beyond int32 offsets");
                d.cpool = (int32_t)c->pool_elems;
                c->pool_elems += (int64_t)nslot * d.npts;
            }
            descs.push_back(d);
            lv.maxpts = std::max(lv.maxpts, d.npts);
        }
        lv.count = (int)descs.size() - lv.first;
        if (lv.count > 65535) return fail(OPESCI_EUNSUPPORTED, "too many boundary statements in one launch");
        if (lv.count > 0) c->levels[L.phase].push_back(lv);
        i = j;
    }
    if (int e = upload(&c->d_desc, descs.data(), descs.size(), c->stream)) return e;
    c->h_desc = std::move(descs);
    return 0;
}

// Split the analytic programs (SolProg): parse each field's postfix program
// into subtrees (a subtree is a contiguous range of ops ending at its root),
// mark which coordinates and whether media each reads, and replace every
// maximal subtree that reads at most one coordinate and no media (and is not
// a lone constant) by a table read along that coordinate (axis 0 for none).
// OPESCI_FLAG_POINTWISE_SOLUTION keeps the programs whole.
int build_solution(OpesciContext* c, const OpesciModelDesc* d) {
    std::vector<OpesciExprOp> pc, tc;    // point programs, slot programs
    std::vector<SolSlot> slots;
    const int dims[3] = {d->dims[0], d->dims[1], d->dims[2]};
    int64_t ntab = 0;
    const bool split = !(d->flags & OPESCI_FLAG_POINTWISE_SOLUTION);
    for (int f = 0; f < NF; ++f) {
        const OpesciExprOp* ops = d->sol_code + d->sol_off[f];
        const int n = d->sol_len[f];
        std::vector<int> start(n), mask(n), stk;
        for (int i = 0; i < n; ++i) {
            const int op = ops[i].op;
            const int arity = op <= OPESCI_OP_MEDIA ? 0 : (op <= OPESCI_OP_DIV ? 2 : 1);
            int m = op == OPESCI_OP_COORD ? 1 << ops[i].a[0] : (op == OPESCI_OP_MEDIA ? 8 : 0), st = i;
            for (int k = 0; k < arity; ++k) {
                const int r = stk.back();
                stk.pop_back();
                m |= mask[r];
                st = std::min(st, start[r]);
            }
            start[i] = st;
            mask[i] = m;
            stk.push_back(i);
        }
        c->sol.off[f] = (int)pc.size();
        c->sol.tab_base[f] = (int)ntab;
        auto emit = [&](auto&& self, int i) -> void {
            const int op = ops[i].op, m = mask[i];
            const bool one_axis = !(m & 8) && (m & (m - 1)) == 0;
            if (split && one_axis && (i > start[i] || op == OPESCI_OP_COORD)) {
                const int axis = m == 0 ? 0 : (m == 1 ? 0 : (m == 2 ? 1 : 2));
                slots.push_back(SolSlot{f, axis, (int32_t)tc.size(), i - start[i] + 1, (int32_t)ntab});
                tc.insert(tc.end(), ops + start[i], ops + i + 1);
                OpesciExprOp t;
                memset(&t, 0, sizeof(t));
                t.op = OPX_TABLE;
                t.a[0] = (int32_t)(ntab - c->sol.tab_base[f]);
                t.a[1] = axis;
                pc.push_back(t);
                ntab += dims[axis];
                return;
            }
            const int arity = op <= OPESCI_OP_MEDIA ? 0 : (op <= OPESCI_OP_DIV ? 2 : 1);
            if (arity == 2) {
                self(self, start[i - 1] - 1);   // left operand ends just before the right one starts
                self(self, i - 1);
            } else if (arity == 1) {
                self(self, i - 1);
            }
            pc.push_back(ops[i]);
        };
        if (n > 0) emit(emit, n - 1);
        if (ntab >= ((int64_t)1 << 31)) return fail(OPESCI_EUNSUPPORTED, "analytic-solution tables too large");
        c->sol.len[f] = (int)pc.size() - c->sol.off[f];
        c->sol.tab_len[f] = (int)ntab - c->sol.tab_base[f];
        int depth = 0, maxd = 0;
        for (int i = c->sol.off[f]; i < (int)pc.size(); ++i) {
            const int op = pc[i].op;
            depth += (op <= OPESCI_OP_MEDIA || op == OPX_TABLE) ? 1 : (op <= OPESCI_OP_DIV ? -1 : 0);
            maxd = std::max(maxd, depth);
        }
        c->sol.shallow[f] = maxd <= 3;
    }
    for (SolSlot& s : slots) s.prog += (int32_t)pc.size();
    pc.insert(pc.end(), tc.begin(), tc.end());
    if (int e = upload(&c->d_sol, pc.data(), pc.size(), c->stream)) return e;
    if (int e = upload(&c->d_solslots, slots.data(), slots.size(), c->stream)) return e;
    if (ntab > 0) OPESCI_HIP_TRY(hipMalloc((void**)&c->d_soltab, (size_t)ntab * sizeof(double)));
    c->sol.code = c->d_sol;
    c->sol.slots = c->d_solslots;
    c->sol.nslots = (int)slots.size();
    c->sol.maxdim = std::max(dims[0], std::max(dims[1], dims[2]));
    c->sol.tab = c->d_soltab;
    return 0;
}

int alloc_zero(OpesciContext* c, void** p, size_t bytes) {
    OPESCI_HIP_TRY(hipMalloc(p, bytes));
    OPESCI_HIP_TRY(launch_zero(*p, bytes, c->stream));
    return 0;
}

// The deep z-face statements (phase 4, one launch level) when they have the
// fixed forms the pair stress sweep applies in place: zero and images of
// Tzz, images of Tyz/Txz, and the Levander recompute of Txx/Tyy with its
// terms in the canonical order.  Anything else leaves zf.F = 0 and the
// statements run as their own launch.
int build_zfaces(OpesciContext* c) {
    c->zf = ZFaceArgs();
    const Geom& g = c->g;
    const int M = g.m;
    if (c->levels[4].size() != 1) return 0;
    const int blo = M, bhi = g.nz - M - 1;
    // image sources within the wave that holds the face: 2M interior points
    // on either side, and the high face at least M points into its tile
    if (bhi - blo < 2 * M || (bhi - M) % 128 < M) return 0;
    ZFaceArgs z;
    ZCoef zc;
    memset(&zc, 0, sizeof(zc));
    bool zero[2] = {false, false}, img[2][3][MAXM] = {}, rec[2][2] = {};
    int F = -1;
    const BcLevel& lv = c->levels[4][0];
    for (int di = lv.first; di < lv.first + lv.count; ++di) {
        const BcDesc& d = c->h_desc[di];
        const int f = d.fl >> 1;
        if (d.axis != 2 || (d.fl & 1) != 1) return 0;
        const int f0 = d.lo0;
        if (d.lo1 != f0 || f0 < M || d.n1 != g.ny - 2 * f0 || d.npts != (g.nx - 2 * f0) * d.n1) return 0;
        if (F >= 0 && F != f0) return 0;
        F = f0;
        const int plane = d.disp;
        const int face = plane <= blo ? 0 : (plane >= bhi ? 1 : -1);
        if (face < 0) return 0;
        const int b = face ? bhi : blo;
        if (f == 5) {
            if (plane == b) {
                if (d.nterms != 0) return 0;
                zero[face] = true;
                continue;
            }
            const int k = face ? plane - b : b - plane;
            const int src = face ? b - k : b + k;
            if (k < 1 || k >= M || d.nterms != 1) return 0;
            if (d.t[0].fl != 5 * 2 + 1 || d.t[0].cslot >= 0 || d.t[0].disp != src - plane) return 0;
            zc.zc[face][0][k] = d.t[0].coef;
            img[face][0][k] = true;
        } else if (f == 7 || f == 8) {
            const int fi = f == 7 ? 1 : 2;
            const int k = face ? plane - b : b - 1 - plane;
            const int src = face ? b - 1 - k : b + k;
            if (k < 0 || k >= M || d.nterms != 1) return 0;
            if (d.t[0].fl != f * 2 + 1 || d.t[0].cslot >= 0 || d.t[0].disp != src - plane) return 0;
            zc.zc[face][fi][k] = d.t[0].coef;
            img[face][fi][k] = true;
        } else if (f == 3 || f == 4) {
            if (M != 2 || plane != b || d.nterms != 9) return 0;
            for (int k = 0; k < 9; ++k) {
                const BcTermDev& t = d.t[k];
                const int want_fl = k < 4 ? 0 : (k < 8 ? 2 : f * 2);
                const int64_t want_disp = k < 4 ? (k - 2) * g.sx : (k < 8 ? (k - 6) * g.sy : 0);
                // literal coefficients only (constant media)
                if (t.fl != want_fl || t.disp != want_disp || t.cslot >= 0) return 0;
                zc.rc[face][f - 3][k] = t.coef;
            }
            rec[face][f - 3] = true;
        } else {
            return 0;
        }
    }
    for (int face = 0; face < 2; ++face) {
        if (!zero[face]) return 0;
        for (int k = 0; k < M; ++k)
            if ((k >= 1 && !img[face][0][k]) || !img[face][1][k] || !img[face][2][k]) return 0;
    }
    const int nrec = rec[0][0] + rec[0][1] + rec[1][0] + rec[1][1];
    if (nrec != 0 && nrec != 4) return 0;
    z.rec = nrec == 4;
    z.F = F;
    z.n1 = g.ny - 2 * F;
    z.npts = (g.nx - 2 * F) * z.n1;
    if (int e = upload(&c->d_zcoef, &zc, 1, c->stream)) return e;
    z.coef = c->d_zcoef;
    c->zf = z;
    return 0;
}

// whether this step's stress launch applies the deep z-face statements
bool zfaces_in_sweep(const OpesciContext* c) {
    return c->zf.F > 0 && stress_applies_zfaces(c->real_bytes, c->het, c->cfg);
}

// The fused step's box: the largest [F, dim - F) in x and y such that no
// stress boundary statement left in phase 2 writes a cell the velocity
// update of a point in the box reads (x and y in [F - M, dim - F + M)), and
// no such statement reads a velocity at t1 (which the fused step would have
// written before it).  The box rows are then deep z-face rows at every halo
// plane and row (F - M >= the z faces' F), and the x and y stencils of the
// sweep stay inside the grid (F >= 2M).
int build_fused(OpesciContext* c) {
    c->fused_F = 0;
    c->fbox = Boxes();
    c->frame = Boxes();
    const Geom& g = c->g;
    const int M = g.m;
    if (!(c->desc.flags & OPESCI_FLAG_FUSED_STEP) || c->real_bytes != 4 || c->het || c->zf.F <= 0 ||
        !fused_step_builds(M))
        return 0;
    for (const OpesciBcStmt& a : c->stmts) {
        if (a.phase != 2) continue;
        if (a.field < 3) return 0;
        for (int k = 0; k < a.nterms; ++k) {
            const OpesciBcTerm& t = c->h_terms[a.term0 + k];
            if (t.field < 3 && t.level == 1) return 0;
        }
    }
    const int fmax = std::min(g.nx, g.ny) / 2 - 1;
    for (int F = std::max(c->zf.F + M, 2 * M); F <= fmax; ++F) {
        const int lo = F - M, hix = g.nx - F + M, hiy = g.ny - F + M;
        bool clear = true;
        for (const OpesciBcStmt& a : c->stmts) {
            if (a.phase != 2 || a.hi[0] <= a.lo[0] || a.hi[1] <= a.lo[1]) continue;
            const int o0 = a.axis == 0 ? 1 : 0, o1 = a.axis == 2 ? 1 : 2;
            int rx[2] = {0, 0}, ry[2] = {0, 0};
            auto range = [&](int axis, int* r) {
                if (axis == a.axis) {
                    r[0] = a.plane;
                    r[1] = a.plane + 1;
                } else {
                    const int j = axis == o0 ? 0 : 1;
                    r[0] = a.lo[j];
                    r[1] = a.hi[j];
                }
            };
            (void)o1;
            range(0, rx);
            range(1, ry);
            if (rx[0] < hix && lo < rx[1] && ry[0] < hiy && lo < ry[1]) {
                clear = false;
                break;
            }
        }
        if (!clear) continue;
        if (g.nx - 2 * F < 1 || g.ny - 2 * F < 1) return 0;
        c->fused_F = F;
        Boxes& b = c->fbox;
        b.n = 1;
        b.x0[0] = F;
        b.x1[0] = g.nx - F;
        b.y0[0] = F;
        b.y1[0] = g.ny - F;
        Boxes& f = c->frame;
        f.n = 4;
        const int X0[4] = {M, g.nx - F, F, F}, X1[4] = {F, g.nx - M, g.nx - F, g.nx - F};
        const int Y0[4] = {M, M, M, g.ny - F}, Y1[4] = {g.ny - M, g.ny - M, F, g.ny - M};
        for (int i = 0; i < 4; ++i) {
            f.x0[i] = X0[i];
            f.x1[i] = X1[i];
            f.y0[i] = Y0[i];
            f.y1[i] = Y1[i];
        }
        return 0;
    }
    return 0;
}

// whether steps run the fused kernel (FP32 pair sweeps with the z faces in place)
bool fused_step(const OpesciContext* c) {
    return c->fused_F > 0 && c->cfg.variant == 0 && zfaces_in_sweep(c);
}

int run_bc_phase(OpesciContext* c, int phase, int t0, int t1) {
    for (const BcLevel& lv : c->levels[phase])
        OPESCI_HIP_TRY(launch_bc(c->g, c->real_bytes, c->storage, c->fptr, c->d_desc, c->d_pool, lv, t0, t1, c->stream));
    return 0;
}

int step_once(OpesciContext* c, int t0) {
    const int t1 = 1 - t0;
    if (fused_step(c)) {
        OPESCI_HIP_TRY(launch_step_fused(c->g, c->storage, c->fptr, c->coefs(), t0, t1, c->zf, c->fbox, 0, c->stream));
        OPESCI_HIP_TRY(launch_stress(c->g, c->real_bytes, c->storage, c->het, c->fptr, c->mptr, c->coefs(), t0, t1,
                                     c->cfg, c->stream, &c->zf, c->d_pool, &c->frame));
        if (int e = run_bc_phase(c, 2, t0, t1)) return e;
        OPESCI_HIP_TRY(launch_velocity(c->g, c->real_bytes, c->storage, c->het, c->fptr, c->mptr, c->coefs(), t0, t1,
                                       c->cfg, c->stream, &c->frame));
        if (int e = run_bc_phase(c, 3, t0, t1)) return e;
        if (c->nrec > 0 || c->src_ntr)
            OPESCI_HIP_TRY(launch_record_source(c->g, c->real_bytes, c->storage, c->fptr, t1, c->nrec, c->d_rec,
                                                c->d_trace, c->src_ntr, c->src_idx, c->d_wavelet, c->d_step,
                                                c->desc.ntsteps, c->stream));
        return 0;
    }
    const bool zin = zfaces_in_sweep(c);
    OPESCI_HIP_TRY(launch_stress(c->g, c->real_bytes, c->storage, c->het, c->fptr, c->mptr, c->coefs(), t0, t1,
                                 c->cfg, c->stream, zin ? &c->zf : nullptr, c->d_pool));
    if (!zin)
        if (int e = run_bc_phase(c, 4, t0, t1)) return e;
    if (int e = run_bc_phase(c, 2, t0, t1)) return e;
    OPESCI_HIP_TRY(launch_velocity(c->g, c->real_bytes, c->storage, c->het, c->fptr, c->mptr, c->coefs(), t0, t1,
                                   c->cfg, c->stream));
    if (int e = run_bc_phase(c, 3, t0, t1)) return e;
    if (c->nrec > 0 || c->src_ntr)
        OPESCI_HIP_TRY(launch_record_source(c->g, c->real_bytes, c->storage, c->fptr, t1, c->nrec, c->d_rec,
                                            c->d_trace, c->src_ntr, c->src_idx, c->d_wavelet, c->d_step,
                                            c->desc.ntsteps, c->stream));
    return 0;
}

int build_graph(OpesciContext* c) {
    if (c->graph) return 0;
    hipGraph_t gr = nullptr;
    OPESCI_HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int e = step_once(c, 0);
    if (!e) e = step_once(c, 1);
    hipError_t he = hipStreamEndCapture(c->stream, &gr);
    if (e) {
        if (gr) (void)hipGraphDestroy(gr);
        return e;
    }
    if (he != hipSuccess) return fail(OPESCI_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(he));
    he = hipGraphInstantiate(&c->graph, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (he != hipSuccess) {
        c->graph = nullptr;
        return fail(OPESCI_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(he));
    }
    return 0;
}

void destroy(OpesciContext* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    for (int f = 0; f < NF; ++f)
        for (int l = 0; l < 2; ++l)
            if (c->base[f][l]) (void)hipFree(c->base[f][l]);
    for (int i = 0; i < OPESCI_M_COUNT; ++i)
        if (c->mbase[i]) (void)hipFree(c->mbase[i]);
    void* ptrs[] = {c->d_desc,  c->d_code,    c->d_sol,  c->d_solslots, c->d_soltab, c->d_partial, c->d_stage, c->d_rec,
                    c->d_trace, c->d_wavelet, c->d_step, c->d_pool,    c->d_dterms, c->d_gpool,
                    c->d_log,   c->d_ckpt,    c->d_seg,  c->d_obs,     c->d_srclog, c->d_tcells,
                    c->d_vcells, c->d_scells, c->d_zcoef, c->d_gather, c->d_abuf};
    for (int f = 0; f < NF; ++f)
        for (int l = 0; l < 2; ++l)
            if (c->abase[f][l]) (void)hipFree(c->abase[f][l]);
    for (int i = 0; i < OPESCI_M_COUNT; ++i)
        if (c->gbase[i]) (void)hipFree(c->gbase[i]);
    for (int i = 0; i < 6; ++i)
        if (c->sbase[i]) (void)hipFree(c->sbase[i]);
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int create(const OpesciModelDesc* d, OpesciContext** out) {
    if (int e = validate(d)) return e;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(OPESCI_EHIP, "no HIP device available");
    std::unique_ptr<OpesciContext, void (*)(OpesciContext*)> c(new OpesciContext(), destroy);
    c->desc = *d;
    c->real_bytes = d->real_bytes;
    c->storage = d->storage;
    c->elem = d->real_bytes == 8 ? 8 : (d->storage == OPESCI_STORE_NATIVE ? 4 : 2);
    c->het = d->media_arrays != 0;
    c->cfg.variant = d->kernel_variant;
    c->cfg.tile_y = d->tile_y;
    c->cfg.x_chunk = d->x_chunk;
    OPESCI_HIP_TRY(hipGetDevice(&c->device));
    OPESCI_HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    Geom& g = c->g;
    g.nx = d->dims[0];
    g.ny = d->dims[1];
    g.nz = d->dims[2];
    g.m = d->space_order / 2;
    const int align = 128 / c->elem;  // first interior z on a 128-byte boundary
    g.zoff = (align - g.m % align) % align;
    g.pitch = ((g.zoff + g.nz + g.m + align - 1) / align) * align;
    g.sy = g.pitch;
    g.sx = (int64_t)g.ny * g.pitch;
    c->field_bytes = (size_t)g.nx * g.sx * c->elem;
    for (int f = 0; f < NF; ++f)
        for (int l = 0; l < 2; ++l) {
            if (int e = alloc_zero(c.get(), &c->base[f][l], c->field_bytes)) return e;
            c->fptr[f][l] = (char*)c->base[f][l] + (size_t)g.zoff * c->elem;
        }
    const size_t rb = d->real_bytes;
    const size_t dense = (size_t)g.nx * g.ny * g.nz;
    OPESCI_HIP_TRY(hipMalloc(&c->d_stage, dense * std::max(rb, sizeof(float) * 3)));
    if (c->het) {
        const size_t mbytes = (size_t)g.nx * g.sx * rb;
        for (int i = 0; i < OPESCI_M_COUNT; ++i) {
            if (int e = alloc_zero(c.get(), &c->mbase[i], mbytes)) return e;
            c->mptr[i] = (char*)c->mbase[i] + (size_t)g.zoff * rb;
        }
        float* st = (float*)c->d_stage;
        OPESCI_HIP_TRY(hipMemcpyAsync(st, d->rho, dense * 4, hipMemcpyHostToDevice, c->stream));
        OPESCI_HIP_TRY(hipMemcpyAsync(st + dense, d->vp, dense * 4, hipMemcpyHostToDevice, c->stream));
        OPESCI_HIP_TRY(hipMemcpyAsync(st + 2 * dense, d->vs, dense * 4, hipMemcpyHostToDevice, c->stream));
        OPESCI_HIP_TRY(launch_media(g, d->real_bytes, st, st + dense, st + 2 * dense, c->mptr, c->stream));
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < MAXM; ++k) {
                c->cd.cn[i][j][k] = d->cn[i][j][k];
                c->cf.cn[i][j][k] = (float)d->cn[i][j][k];
                c->cd.cv[i][j][k] = d->cv[i][j][k];
                c->cf.cv[i][j][k] = (float)d->cv[i][j][k];
            }
    for (int s = 0; s < 3; ++s)
        for (int gq = 0; gq < 2; ++gq)
            for (int k = 0; k < MAXM; ++k) {
                c->cd.cs[s][gq][k] = d->cs[s][gq][k];
                c->cf.cs[s][gq][k] = (float)d->cs[s][gq][k];
            }
    c->stmts.assign(d->stmts, d->stmts + d->nstmts);
    c->h_terms.assign(d->terms, d->terms + d->nterms);
    if (int e = build_program(c.get(), d->terms, d->dterms, d->ndterms, d->nterms)) return e;
    if (int e = build_zfaces(c.get())) return e;
    if (int e = build_fused(c.get())) return e;
    if (int e = upload(&c->d_dterms, d->dterms, (size_t)d->ndterms, c->stream)) return e;
    c->h_dterms.assign(d->dterms, d->dterms + d->ndterms);
    if (int e = upload(&c->d_code, d->bc_code, (size_t)d->ncode, c->stream)) return e;
    // time-invariant media-dependent coefficients, once (the media are ready)
    if (c->pool_elems > 0) {
        OPESCI_HIP_TRY(hipMalloc(&c->d_pool, (size_t)c->pool_elems * c->real_bytes));
        for (int ph = 0; ph < 5; ++ph)
            for (const BcLevel& lv : c->levels[ph])
                OPESCI_HIP_TRY(launch_bc_coef(c->g, c->real_bytes, c->mptr, c->d_desc, c->d_code, c->d_pool, lv.first,
                                              lv.count, lv.maxpts, c->stream));
        OPESCI_HIP_TRY(hipStreamSynchronize(c->stream));
    }
    if (int e = build_solution(c.get(), d)) return e;
    c->partial_cap = NF * 512;
    OPESCI_HIP_TRY(hipMalloc((void**)&c->d_partial, c->partial_cap * sizeof(double)));
    c->nrec = d->nrec;
    c->src_ntr = d->src_on;
    if (c->nrec || c->src_ntr) {
        if (int e = alloc_zero(c.get(), (void**)&c->d_step, 256)) return e;
        if (c->nrec) {
            if (int e = upload(&c->d_rec, d->rec_idx, (size_t)3 * c->nrec, c->stream)) return e;
            c->h_rec.assign(d->rec_idx, d->rec_idx + (size_t)3 * c->nrec);
            const size_t tb = (size_t)std::max(d->ntsteps, 1) * c->nrec * 4 * d->real_bytes;
            if (int e = alloc_zero(c.get(), (void**)&c->d_trace, tb)) return e;
        }
        if (c->src_ntr) {
            if (int e = upload(&c->d_wavelet, d->src_wavelet, (size_t)std::max(d->ntsteps, 1) * c->src_ntr, c->stream))
                return e;
            for (int a = 0; a < 3; ++a) c->src_idx[a] = d->src_idx[a];
        }
    }
    // pointers into the caller's arrays are not kept
    c->desc.rho = c->desc.vp = c->desc.vs = nullptr;
    c->desc.stmts = nullptr;
    c->desc.terms = nullptr;
    c->desc.bc_code = nullptr;
    c->desc.sol_code = nullptr;
    c->desc.src_wavelet = nullptr;
    c->desc.rec_idx = nullptr;
    OPESCI_HIP_TRY(hipStreamSynchronize(c->stream));
    *out = c.release();
    return 0;
}

int ctx_init(OpesciContext* c) {
    c->grad_every = 0;
    for (int f = 0; f < NF; ++f)
        for (int l = 0; l < 2; ++l) OPESCI_HIP_TRY(launch_zero(c->base[f][l], c->field_bytes, c->stream));
    OPESCI_HIP_TRY(launch_initial(c->g, c->real_bytes, c->storage, c->fptr, c->mptr, c->sol, c->desc.spacing,
                                  c->desc.dt, c->stream));
    if (int e = run_bc_phase(c, 0, 0, 0)) return e;
    if (int e = run_bc_phase(c, 1, 0, 0)) return e;
    if (c->d_step) OPESCI_HIP_TRY(launch_zero(c->d_step, 256, c->stream));
    OPESCI_HIP_TRY(hipStreamSynchronize(c->stream));
    c->steps = 0;
    c->initialised = true;
    return 0;
}

int ctx_run(OpesciContext* c, int n, double* seconds) {
    if (!c->initialised) return fail(OPESCI_ESTATE, "opesci_ctx_init must run before the time loop");
    c->grad_every = 0;
    if (n < 0) return fail(OPESCI_EINVAL, "negative step count");
    if (c->desc.use_graph && n >= 2) {
        if (int e = build_graph(c)) return e;
    }
    hipEvent_t e0, e1;
    OPESCI_HIP_TRY(hipEventCreate(&e0));
    OPESCI_HIP_TRY(hipEventCreate(&e1));
    OPESCI_HIP_TRY(hipEventRecord(e0, c->stream));
    int left = n;
    int rc = 0;
    if (c->steps % 2 == 1 && left > 0) {
        rc = step_once(c, 1);
        --left;
        ++c->steps;
    }
    while (!rc && left >= 2 && c->graph) {
        hipError_t he = hipGraphLaunch(c->graph, c->stream);
        if (he != hipSuccess) rc = fail(OPESCI_EHIP, std::string("hipGraphLaunch: ") + hipGetErrorString(he));
        left -= 2;
        c->steps += 2;
    }
    while (!rc && left > 0) {
        rc = step_once(c, c->steps % 2);
        --left;
        ++c->steps;
    }
    OPESCI_HIP_TRY(hipEventRecord(e1, c->stream));
    OPESCI_HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    OPESCI_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    OPESCI_HIP_TRY(hipGetLastError());
    if (seconds) *seconds = ms * 1e-3;
    return 0;
}

int ctx_l2(OpesciContext* c, double out[NF]) {
    if (!c->initialised) return fail(OPESCI_ESTATE, "no state: call opesci_ctx_init first");
    double t[NF];
    for (int f = 0; f < NF; ++f) t[f] = c->desc.dt * c->steps + (time_stag(f) ? c->desc.dt / 2.0 : 0.0);
    OPESCI_HIP_TRY(launch_l2(c->g, c->real_bytes, c->storage, c->fptr, c->mptr, c->sol, c->desc.spacing,
                             c->steps % 2, t, c->d_partial, c->partial_cap, out, c->stream));
    return 0;
}

int ctx_download(OpesciContext* c, int f, int level, void* host) {
    if (f < 0 || f >= NF || level < -1 || level > 1 || !host) return fail(OPESCI_EINVAL, "bad download arguments");
    const int l = level < 0 ? c->steps % 2 : level;
    const size_t dense = (size_t)c->g.nx * c->g.ny * c->g.nz;
    OPESCI_HIP_TRY(launch_unpack(c->g, c->real_bytes, c->storage, c->fptr[f][l], c->d_stage, c->stream));
    OPESCI_HIP_TRY(hipMemcpyAsync(host, c->d_stage, dense * c->real_bytes, hipMemcpyDeviceToHost, c->stream));
    OPESCI_HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int ctx_upload(OpesciContext* c, int f, int level, const void* host) {
    if (f < 0 || f >= NF || level < 0 || level > 1 || !host) return fail(OPESCI_EINVAL, "bad upload arguments");
    c->grad_every = 0;
    const size_t dense = (size_t)c->g.nx * c->g.ny * c->g.nz;
    OPESCI_HIP_TRY(hipMemcpyAsync(c->d_stage, host, dense * c->real_bytes, hipMemcpyHostToDevice, c->stream));
    OPESCI_HIP_TRY(launch_pack(c->g, c->real_bytes, c->storage, c->d_stage, c->fptr[f][level], c->stream));
    if (!c->initialised) {
        // a state built by uploads starts at step 0: the source and receiver
        // step counter restarts with it (a gradient leaves it elsewhere)
        if (c->d_step) OPESCI_HIP_TRY(launch_zero(c->d_step, 256, c->stream));
        c->steps = 0;
    }
    OPESCI_HIP_TRY(hipStreamSynchronize(c->stream));
    c->initialised = true;
    return 0;
}

double flops_per_point(const OpesciContext* c) {
    // reference op model (RegularGrid.get_kernel_ai): one add + one multiply
    // per stencil point of each update: 3*3*2M (normal) + 3*2*2M (shear) + 3*3*2M (velocity)
    const int m = c->g.m;
    return 2.0 * (18.0 * m + 12.0 * m + 18.0 * m);
}

double interior_points(const OpesciContext* c) {
    const int m = c->g.m;
    return (double)(c->g.nx - 2 * m) * (c->g.ny - 2 * m) * (c->g.nz - 2 * m);
}

}  // namespace

// C++ exceptions (host allocation, thread start) never cross the C ABI:
// every entry point returns them as error codes.
template <typename F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(OPESCI_ENOMEM, "host allocation failed");
    } catch (const std::system_error& e) {
        return fail(OPESCI_ENOMEM, std::string("host resources: ") + e.what());
    } catch (const std::exception& e) {
        return fail(OPESCI_EUNSUPPORTED, e.what());
    }
}

extern "C" {

const char* opesci_last_error(void) { return g_error.c_str(); }
int opesci_abi_version(void) { return OPESCI_ABI_VERSION; }

int opesci_device_info(char* name, int32_t name_len, char* arch, int32_t arch_len) {
    return guarded([&]() -> int {
        if (name && name_len > 0) name[0] = 0;
        if (arch && arch_len > 0) arch[0] = 0;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(OPESCI_EHIP, "no HIP device available");
        hipDeviceProp_t p;
        OPESCI_HIP_TRY(hipGetDeviceProperties(&p, 0));
        if (name && name_len > 0) {
            strncpy(name, p.name, name_len - 1);
            name[name_len - 1] = 0;
        }
        if (arch && arch_len > 0) {
            strncpy(arch, p.gcnArchName, arch_len - 1);
            arch[arch_len - 1] = 0;
        }
        return 0;
    });
}

int opesci_ctx_create(const OpesciModelDesc* desc, OpesciContext** out) {
    return guarded([&]() -> int {
        if (!out) return fail(OPESCI_EINVAL, "null output handle");
        *out = nullptr;
        return create(desc, out);
    });
}

int opesci_ctx_destroy(OpesciContext* ctx) {
    return guarded([&]() -> int {
        if (ctx == g_default) g_default = nullptr;
        destroy(ctx);
        return 0;
    });
}

int opesci_ctx_init(OpesciContext* ctx) {
    return guarded([&]() -> int {
        if (!ctx) return fail(OPESCI_EINVAL, "null context");
        return ctx_init(ctx);
    });
}

int opesci_ctx_run(OpesciContext* ctx, int32_t nsteps, double* seconds) {
    return guarded([&]() -> int {
        if (!ctx) return fail(OPESCI_EINVAL, "null context");
        return ctx_run(ctx, nsteps, seconds);
    });
}

int opesci_ctx_l2(OpesciContext* ctx, double out[OPESCI_NFIELDS]) {
    return guarded([&]() -> int {
        if (!ctx || !out) return fail(OPESCI_EINVAL, "null argument");
        return ctx_l2(ctx, out);
    });
}

int opesci_ctx_l2_noalloc(OpesciContext* ctx, double* d_out, void* stream) {
    return guarded([&]() -> int {
        if (!ctx || !d_out) return fail(OPESCI_EINVAL, "null argument");
        if (!ctx->initialised) return fail(OPESCI_ESTATE, "no state: call opesci_ctx_init first");
        int dev = -1;
        OPESCI_HIP_TRY(hipGetDevice(&dev));
        if (dev != ctx->device) return fail(OPESCI_ESTATE, "the context's device must be the current device");
        double t[NF];
        for (int f = 0; f < NF; ++f) t[f] = ctx->desc.dt * ctx->steps + (time_stag(f) ? ctx->desc.dt / 2.0 : 0.0);
        const hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
        OPESCI_HIP_TRY(launch_l2_noalloc(ctx->g, ctx->real_bytes, ctx->storage, ctx->fptr, ctx->mptr, ctx->sol,
                                         ctx->desc.spacing, ctx->steps % 2, t, d_out, s));
        return 0;
    });
}

int opesci_ctx_download(OpesciContext* ctx, int32_t field, int32_t level, void* host) {
    return guarded([&]() -> int {
        if (!ctx) return fail(OPESCI_EINVAL, "null context");
        return ctx_download(ctx, field, level, host);
    });
}

int opesci_ctx_upload(OpesciContext* ctx, int32_t field, int32_t level, const void* host) {
    return guarded([&]() -> int {
        if (!ctx) return fail(OPESCI_EINVAL, "null context");
        return ctx_upload(ctx, field, level, host);
    });
}

int opesci_ctx_steps(OpesciContext* ctx, int32_t* steps) {
    return guarded([&]() -> int {
        if (!ctx || !steps) return fail(OPESCI_EINVAL, "null argument");
        *steps = ctx->steps;
        return 0;
    });
}

}  // extern "C"

// traces recorded so far, copied out and converted to T
template <typename T>
static int receivers_as(OpesciContext* ctx, T* host, int64_t capacity) {
    if (!ctx) return fail(OPESCI_EINVAL, "null context");
    if (ctx->nrec == 0) return 0;
    const int64_t n = (int64_t)std::min(ctx->steps, ctx->desc.ntsteps) * ctx->nrec * 4;
    if (!host || capacity < n) return fail(OPESCI_EINVAL, "receiver buffer too small");
    if ((size_t)ctx->real_bytes == sizeof(T)) {
        OPESCI_HIP_TRY(hipMemcpyAsync(host, ctx->d_trace, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
        OPESCI_HIP_TRY(hipStreamSynchronize(ctx->stream));
        return 0;
    }
    std::vector<char> tmp((size_t)n * ctx->real_bytes);
    OPESCI_HIP_TRY(hipMemcpyAsync(tmp.data(), ctx->d_trace, tmp.size(), hipMemcpyDeviceToHost, ctx->stream));
    OPESCI_HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < n; ++i)
        host[i] = ctx->real_bytes == 8 ? (T) reinterpret_cast<const double*>(tmp.data())[i]
                                       : (T) reinterpret_cast<const float*>(tmp.data())[i];
    return 0;
}

extern "C" {

int opesci_ctx_receivers(OpesciContext* ctx, float* host, int64_t capacity) {
    return guarded([&]() -> int {
        return receivers_as(ctx, host, capacity);
    });
}

int opesci_ctx_receivers_f64(OpesciContext* ctx, double* host, int64_t capacity) {
    return guarded([&]() -> int {
        return receivers_as(ctx, host, capacity);
    });
}

int opesci_ctx_kernel_times(OpesciContext* ctx, int32_t nsteps, double out_ms[3]) {
    return guarded([&]() -> int {
        if (!ctx || !out_ms || nsteps <= 0) return fail(OPESCI_EINVAL, "bad arguments");
        if (!ctx->initialised) return fail(OPESCI_ESTATE, "opesci_ctx_init must run first");
        std::vector<hipEvent_t> ev((size_t)nsteps * 5 + 1);
        for (auto& e : ev) OPESCI_HIP_TRY(hipEventCreate(&e));
        int rc = 0;
        const bool fused = fused_step(ctx);
        for (int s = 0; s < nsteps && !rc && fused; ++s) {
            // fused kernel | frame stress | stress BCs | frame velocity | velocity BCs, sources
            const int t0 = ctx->steps % 2, t1 = 1 - t0;
            hipEvent_t* e = &ev[(size_t)s * 5];
            OPESCI_HIP_TRY(hipEventRecord(e[0], ctx->stream));
            OPESCI_HIP_TRY(launch_step_fused(ctx->g, ctx->storage, ctx->fptr, ctx->coefs(), t0, t1, ctx->zf, ctx->fbox, 0,
                                             ctx->stream));
            OPESCI_HIP_TRY(hipEventRecord(e[1], ctx->stream));
            OPESCI_HIP_TRY(launch_stress(ctx->g, ctx->real_bytes, ctx->storage, ctx->het, ctx->fptr, ctx->mptr,
                                         ctx->coefs(), t0, t1, ctx->cfg, ctx->stream, &ctx->zf, ctx->d_pool, &ctx->frame));
            OPESCI_HIP_TRY(hipEventRecord(e[2], ctx->stream));
            rc = run_bc_phase(ctx, 2, t0, t1);
            OPESCI_HIP_TRY(hipEventRecord(e[3], ctx->stream));
            OPESCI_HIP_TRY(launch_velocity(ctx->g, ctx->real_bytes, ctx->storage, ctx->het, ctx->fptr, ctx->mptr,
                                           ctx->coefs(), t0, t1, ctx->cfg, ctx->stream, &ctx->frame));
            OPESCI_HIP_TRY(hipEventRecord(e[4], ctx->stream));
            if (!rc) rc = run_bc_phase(ctx, 3, t0, t1);
            if (!rc && (ctx->nrec > 0 || ctx->src_ntr))
                OPESCI_HIP_TRY(launch_record_source(ctx->g, ctx->real_bytes, ctx->storage, ctx->fptr, t1, ctx->nrec,
                                                    ctx->d_rec, ctx->d_trace, ctx->src_ntr, ctx->src_idx,
                                                    ctx->d_wavelet, ctx->d_step, ctx->desc.ntsteps, ctx->stream));
            OPESCI_HIP_TRY(hipEventRecord(e[5], ctx->stream));
            ++ctx->steps;
        }
        for (int s = 0; s < nsteps && !rc && !fused; ++s) {
            const int t0 = ctx->steps % 2, t1 = 1 - t0;
            hipEvent_t* e = &ev[(size_t)s * 5];
            OPESCI_HIP_TRY(hipEventRecord(e[0], ctx->stream));
            const bool zin = zfaces_in_sweep(ctx);
            OPESCI_HIP_TRY(launch_stress(ctx->g, ctx->real_bytes, ctx->storage, ctx->het, ctx->fptr, ctx->mptr,
                                         ctx->coefs(), t0, t1, ctx->cfg, ctx->stream, zin ? &ctx->zf : nullptr,
                                         ctx->d_pool));
            OPESCI_HIP_TRY(hipEventRecord(e[1], ctx->stream));
            if (!zin) rc = run_bc_phase(ctx, 4, t0, t1);
            if (!rc) rc = run_bc_phase(ctx, 2, t0, t1);
            OPESCI_HIP_TRY(hipEventRecord(e[2], ctx->stream));
            OPESCI_HIP_TRY(launch_velocity(ctx->g, ctx->real_bytes, ctx->storage, ctx->het, ctx->fptr, ctx->mptr,
                                           ctx->coefs(), t0, t1, ctx->cfg, ctx->stream));
            OPESCI_HIP_TRY(hipEventRecord(e[3], ctx->stream));
            if (!rc) rc = run_bc_phase(ctx, 3, t0, t1);
            if (!rc && (ctx->nrec > 0 || ctx->src_ntr))
                OPESCI_HIP_TRY(launch_record_source(ctx->g, ctx->real_bytes, ctx->storage, ctx->fptr, t1, ctx->nrec,
                                                    ctx->d_rec, ctx->d_trace, ctx->src_ntr, ctx->src_idx,
                                                    ctx->d_wavelet, ctx->d_step, ctx->desc.ntsteps, ctx->stream));
            OPESCI_HIP_TRY(hipEventRecord(e[4], ctx->stream));
            ++ctx->steps;
        }
        OPESCI_HIP_TRY(hipStreamSynchronize(ctx->stream));
        double acc[3] = {0, 0, 0};
        for (int s = 0; s < nsteps && fused; ++s) {
            hipEvent_t* e = &ev[(size_t)s * 5];
            float a = 0, b = 0, c2 = 0, d2 = 0, e2 = 0;
            OPESCI_HIP_TRY(hipEventElapsedTime(&a, e[0], e[1]));
            OPESCI_HIP_TRY(hipEventElapsedTime(&b, e[1], e[2]));
            OPESCI_HIP_TRY(hipEventElapsedTime(&c2, e[2], e[3]));
            OPESCI_HIP_TRY(hipEventElapsedTime(&d2, e[3], e[4]));
            OPESCI_HIP_TRY(hipEventElapsedTime(&e2, e[4], e[5]));
            acc[0] += a;
            acc[1] += b + d2;
            acc[2] += c2 + e2;
        }
        for (int s = 0; s < nsteps && !fused; ++s) {
            hipEvent_t* e = &ev[(size_t)s * 5];
            float a = 0, b = 0, c2 = 0, d2 = 0;
            OPESCI_HIP_TRY(hipEventElapsedTime(&a, e[0], e[1]));
            OPESCI_HIP_TRY(hipEventElapsedTime(&b, e[2], e[3]));
            OPESCI_HIP_TRY(hipEventElapsedTime(&c2, e[1], e[2]));
            OPESCI_HIP_TRY(hipEventElapsedTime(&d2, e[3], e[4]));
            acc[0] += a;
            acc[1] += b;
            acc[2] += c2 + d2;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        for (int i = 0; i < 3; ++i) out_ms[i] = acc[i] / nsteps;
        return rc;
    });
}

int opesci_ctx_stats(OpesciContext* ctx, int32_t out[5]) {
    return guarded([&]() -> int {
        if (!ctx || !out) return fail(OPESCI_EINVAL, "null argument");
        int n4 = 0;
        for (const BcLevel& lv : ctx->levels[4]) n4 += lv.count;
        const bool zin = zfaces_in_sweep(ctx);
        out[0] = (int)(ctx->levels[2].size() + ctx->levels[3].size() + (zin ? 0 : ctx->levels[4].size()));
        out[1] = n4;
        out[2] = zin ? 0 : (int)ctx->levels[4].size();
        out[3] = zin ? 1 : 0;
        out[4] = fused_step(ctx) ? ctx->fused_F : 0;
        return 0;
    });
}

int opesci_ctx_configure(OpesciContext* ctx, int32_t kernel_variant, int32_t tile_y, int32_t x_chunk) {
    return guarded([&]() -> int {
        if (!ctx) return fail(OPESCI_EINVAL, "null context");
        if (kernel_variant < 0 || kernel_variant > 2 || (tile_y != 0 && tile_y != 4 && tile_y != 8 && tile_y != 16) ||
            x_chunk < 0)
            return fail(OPESCI_EINVAL, "kernel_variant 0|1|2, tile_y 0|4|8|16, x_chunk >= 0");
        ctx->cfg = InteriorLaunch{kernel_variant, tile_y, x_chunk};
        ctx->desc.kernel_variant = kernel_variant;
        ctx->desc.tile_y = tile_y;
        ctx->desc.x_chunk = x_chunk;
        if (ctx->graph) {
            OPESCI_HIP_TRY(hipStreamSynchronize(ctx->stream));
            (void)hipGraphExecDestroy(ctx->graph);
            ctx->graph = nullptr;
        }
        return 0;
    });
}

// ---------------------------------------------------------------- adjoint
// Index arrays of the fixed-order gathers (AdjGather): per level, every
// scattered addition of the transposes (target: the source cell of a term,
// value: coefficient x the point's adjoint value), generated in (statement,
// point, term) order and stably sorted by target, so each target's sum runs
// in that order on every call.  Receivers sharing a cell are grouped the
// same way, in receiver order.
struct GatherEntry {
    int32_t key;
    int32_t src;
    int64_t off;
    int32_t cinfo;
};
struct GatherLevel {   // one level's gather, level-local indices
    std::vector<int32_t> start, key, src, cinfo;
    std::vector<int64_t> off;
    int64_t slots = 0;
    int err = 0;
};

// Sort one level's entries by (key, offset), ties in generation order: one
// 64-bit sort key per entry, (target rank << ibits) | entry index.
static void finish_level(std::vector<GatherEntry>& es, GatherLevel& out) {
    const size_t n = es.size();
    out.start.clear();
    if (n == 0) {
        out.start.push_back(0);
        return;
    }
    int64_t omin = es[0].off, omax = es[0].off;
    int32_t kmax = 0;
    for (const GatherEntry& e : es) {
        omin = std::min(omin, e.off);
        omax = std::max(omax, e.off);
        kmax = std::max(kmax, e.key);
    }
    int ibits = 1;
    while (((size_t)1 << ibits) < n) ++ibits;
    const uint64_t span = (uint64_t)(omax - omin) + 1;
    std::vector<uint64_t> sk(n);
    if ((unsigned __int128)((uint64_t)kmax + 1) * span < ((unsigned __int128)1 << (64 - ibits))) {
        // least-significant-digit radix sort of the packed keys, 11-bit
        // digits over the bits the keys use
        uint64_t kmaxv = 0;
        for (size_t i = 0; i < n; ++i) {
            sk[i] = (((uint64_t)es[i].key * span + (uint64_t)(es[i].off - omin)) << ibits) | i;
            kmaxv = std::max(kmaxv, sk[i]);
        }
        std::vector<uint64_t> tmp(n);
        std::vector<size_t> cnt(2048);
        for (int sh = 0; sh < 64 && (kmaxv >> sh) != 0; sh += 11) {
            std::fill(cnt.begin(), cnt.end(), 0);
            for (size_t i = 0; i < n; ++i) ++cnt[(sk[i] >> sh) & 2047];
            size_t acc = 0;
            for (size_t& c : cnt) {
                const size_t v = c;
                c = acc;
                acc += v;
            }
            for (size_t i = 0; i < n; ++i) tmp[cnt[(sk[i] >> sh) & 2047]++] = sk[i];
            sk.swap(tmp);
        }
        for (uint64_t& v : sk) v &= ((uint64_t)1 << ibits) - 1;
    } else {
        for (size_t i = 0; i < n; ++i) sk[i] = i;
        std::stable_sort(sk.begin(), sk.end(), [&](uint64_t a, uint64_t b) {
            return es[a].key != es[b].key ? es[a].key < es[b].key : es[a].off < es[b].off;
        });
    }
    out.src.resize(n);
    out.cinfo.resize(n);
    for (size_t i = 0; i < n; ++i) {
        const GatherEntry& e = es[sk[i]];
        if (i == 0 || e.key != es[sk[i - 1]].key || e.off != es[sk[i - 1]].off) {
            out.start.push_back((int32_t)i);
            out.key.push_back(e.key);
            out.off.push_back(e.off);
        }
        out.src[i] = e.src;
        out.cinfo[i] = e.cinfo;
    }
    out.start.push_back((int32_t)n);
}

// transposes of one level: target = source cell of term k, value = its
// coefficient (pool slot or literal) x the point's adjoint value
static void gather_bc_level(const OpesciContext* c, const BcLevel& lv, int32_t* boff, GatherLevel& out) {
    const Geom& g = c->g;
    std::vector<GatherEntry> es;
    int64_t slots = 0;
    for (int di = lv.first; di < lv.first + lv.count; ++di) {
        const BcDesc& d = c->h_desc[di];
        boff[di] = (int32_t)slots;
        const int o0 = d.axis == 0 ? 1 : 0, o1 = d.axis == 2 ? 1 : 2;
        const int64_t so0 = o0 == 0 ? g.sx : g.sy, so1 = o1 == 1 ? g.sy : 1;
        for (int p = 0; p < d.npts; ++p) {
            const int64_t dst = (int64_t)(d.lo0 + p / d.n1) * so0 + (int64_t)(d.lo1 + p % d.n1) * so1 + d.disp;
            for (int k = 0; k < d.nterms; ++k) {
                const int32_t ci = d.t[k].cslot >= 0 ? (int32_t)(d.cpool + (int64_t)d.t[k].cslot * d.npts + p)
                                                     : -(1 + di * BC_MAXT + k);
                es.push_back({d.t[k].fl, (int32_t)(slots + p), dst + d.t[k].disp, ci});
            }
        }
        slots += d.npts;
        if (slots > INT32_MAX || es.size() > (size_t)INT32_MAX) {
            out.err = 1;
            return;
        }
    }
    out.slots = slots;
    finish_level(es, out);
}

static int build_adj_gathers(OpesciContext* c) {
    const Geom& g = c->g;
    const size_t rb = c->real_bytes;
    const int nd = (int)c->h_desc.size();
    std::vector<int32_t> boff_bc(nd, 0);
    // one task per level of the stepping phases, run on a few host threads
    struct Task {
        int ph, L;
    };
    std::vector<Task> tasks;
    for (int ph = 2; ph < 5; ++ph)
        for (int L = 0; L < (int)c->levels[ph].size(); ++L) tasks.push_back({ph, L});
    std::vector<GatherLevel> res(tasks.size());
    {
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t t; (t = next.fetch_add(1)) < tasks.size();) {
                const Task& k = tasks[t];
                gather_bc_level(c, c->levels[k.ph][k.L], boff_bc.data(), res[t]);
            }
        };
        const int nt = (int)std::min<size_t>(tasks.size(), std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
        std::vector<std::thread> pool;
        for (int i = 1; i < nt; ++i) pool.emplace_back(work);
        work();
        for (std::thread& th : pool) th.join();
    }
    // concatenate in task order
    std::vector<int32_t> start, key, src, cinfo, rstart, rorder;
    std::vector<int64_t> off, rcell;
    struct Span {
        size_t e0, s0, t0;   // first entry, first start, first target
        int ntgt;
    };
    std::vector<Span> sp_bc[5];
    for (int ph = 0; ph < 2; ++ph) sp_bc[ph].assign(c->levels[ph].size(), Span{0, 0, 0, 0});
    int64_t maxslots = 0;
    for (size_t t = 0; t < tasks.size(); ++t) {
        GatherLevel& r = res[t];
        if (r.err) return fail(OPESCI_EUNSUPPORTED, "boundary level too large for the adjoint");
        sp_bc[tasks[t].ph].push_back(Span{src.size(), start.size(), key.size(), (int)r.key.size()});
        maxslots = std::max(maxslots, r.slots);
        start.insert(start.end(), r.start.begin(), r.start.end());
        key.insert(key.end(), r.key.begin(), r.key.end());
        off.insert(off.end(), r.off.begin(), r.off.end());
        src.insert(src.end(), r.src.begin(), r.src.end());
        cinfo.insert(cinfo.end(), r.cinfo.begin(), r.cinfo.end());
        r = GatherLevel();
    }
    // receivers by cell
    {
        std::vector<std::pair<int64_t, int32_t>> rc((size_t)c->nrec);
        for (int r = 0; r < c->nrec; ++r)
            rc[r] = {g.at(c->h_rec[3 * r], c->h_rec[3 * r + 1], c->h_rec[3 * r + 2]), r};
        std::sort(rc.begin(), rc.end());
        for (size_t i = 0; i < rc.size(); ++i) {
            if (i == 0 || rc[i].first != rc[i - 1].first) {
                rstart.push_back((int32_t)i);
                rcell.push_back(rc[i].first);
            }
            rorder.push_back(rc[i].second);
        }
        rstart.push_back((int32_t)rc.size());
        c->nrgroups = (int)rcell.size();
    }
    // one allocation: sections aligned to 256 bytes
    size_t bytes = 0;
    auto sec = [&](size_t n) {
        const size_t o = bytes;
        bytes += (n + 255) & ~size_t(255);
        return o;
    };
    const size_t o_start = sec(start.size() * 4), o_key = sec(key.size() * 4), o_off = sec(off.size() * 8),
                 o_src = sec(src.size() * 4), o_ci = sec(cinfo.size() * 4), o_bbc = sec((size_t)nd * 4),
                 o_rs = sec(rstart.size() * 4), o_rc = sec(rcell.size() * 8), o_ro = sec(rorder.size() * 4);
    OPESCI_HIP_TRY(hipMalloc((void**)&c->d_gather, std::max(bytes, (size_t)256)));
    std::vector<char> blob(bytes);
    auto put = [&](size_t o, const void* p, size_t n) {
        if (n) memcpy(blob.data() + o, p, n);
    };
    put(o_start, start.data(), start.size() * 4);
    put(o_key, key.data(), key.size() * 4);
    put(o_off, off.data(), off.size() * 8);
    put(o_src, src.data(), src.size() * 4);
    put(o_ci, cinfo.data(), cinfo.size() * 4);
    put(o_bbc, boff_bc.data(), (size_t)nd * 4);
    put(o_rs, rstart.data(), rstart.size() * 4);
    put(o_rc, rcell.data(), rcell.size() * 8);
    put(o_ro, rorder.data(), rorder.size() * 4);
    if (bytes) {
        OPESCI_HIP_TRY(hipMemcpyAsync(c->d_gather, blob.data(), bytes, hipMemcpyHostToDevice, c->stream));
        OPESCI_HIP_TRY(hipStreamSynchronize(c->stream));
    }
    if (int e = alloc_zero(c, &c->d_abuf, ((size_t)std::max<int64_t>(maxslots, 1) * rb + 255) & ~size_t(255))) return e;
    char* base = c->d_gather;
    auto views = [&](const std::vector<Span>& sp, size_t o_boff, std::vector<AdjGather>& out) {
        out.clear();
        for (const Span& s : sp) {
            AdjGather a;
            a.ntgt = s.ntgt;
            a.start = (const int32_t*)(base + o_start) + s.s0;
            a.key = (const int32_t*)(base + o_key) + s.t0;
            a.off = (const int64_t*)(base + o_off) + s.t0;
            a.src = (const int32_t*)(base + o_src) + s.e0;
            a.cinfo = (const int32_t*)(base + o_ci) + s.e0;
            a.boff = (const int32_t*)(base + o_boff);
            a.buf = c->d_abuf;
            out.push_back(a);
        }
    };
    for (int ph = 0; ph < 5; ++ph) views(sp_bc[ph], o_bbc, c->gat_bc[ph]);
    c->rg_start = (const int32_t*)(base + o_rs);
    c->rg_cell = (const int64_t*)(base + o_rc);
    c->rg_order = (const int32_t*)(base + o_ro);
    return 0;
}

static int adj_alloc(OpesciContext* c, int ckpt, size_t seg_bytes, int log_steps) {
    const Geom& g = c->g;
    const size_t rb = c->real_bytes;
    if (!c->abase[0][0]) {
        for (int f = 0; f < NF; ++f)
            for (int l = 0; l < 2; ++l) {
                if (int e = alloc_zero(c, &c->abase[f][l], c->field_bytes)) return e;
                c->aptr[f][l] = (char*)c->abase[f][l] + (size_t)g.zoff * rb;
            }
        const size_t mbytes = (size_t)g.nx * g.sx * rb;
        for (int i = 0; i < OPESCI_M_COUNT; ++i) {
            if (int e = alloc_zero(c, &c->gbase[i], mbytes)) return e;
            c->gptr[i] = (char*)c->gbase[i] + (size_t)g.zoff * rb;
        }
        // zero outside the interior for good: the adjoint kernels write interior cells only
        for (int i = 0; i < 6; ++i) {
            if (int e = alloc_zero(c, &c->sbase[i], c->field_bytes)) return e;
            c->sptr[i] = (char*)c->sbase[i] + (size_t)g.zoff * rb;
        }
        if (c->pool_elems > 0) {
            const size_t pb = ((size_t)c->pool_elems * rb + 15) & ~size_t(15);
            if (int e = alloc_zero(c, &c->d_gpool, pb)) return e;
        }
        // Cells where the level swap of an interior transpose must act
        // (entries (offset << 5) | exchange << 4 | field):
        //  * interior cells that the step's boundary statements read at
        //    level 0 hold adjoint values of level t0 at the swap: they move
        //    into the swapped-in level t0 (add, then zero);
        //  * non-interior cells that boundary statements read at either
        //    level keep their values in their physical arrays, as the point
        //    update leaves every non-interior cell (exchange).  A read of a
        //    level-t1 cell that a later statement of the step writes sees
        //    the value of two steps before, which lives in the same array.
        // Stresses: statements of every stepping phase (their transposes all
        // run before the stress swap).  Velocities: the velocity phase; the
        // stress phase may read velocities at level t0 only (its transposes
        // run after the velocity swap, into the swapped-in level t0).
        {
            std::vector<int64_t> tc, vc, sc;
            const int64_t so[3] = {g.sx, g.sy, 1};
            const int m = g.m;
            for (int ph = 2; ph < 5; ++ph)
                for (const BcLevel& lv : c->levels[ph])
                    for (int di = lv.first; di < lv.first + lv.count; ++di) {
                        const BcDesc& d = c->h_desc[di];
                        const int o0 = d.axis == 0 ? 1 : 0, o1 = d.axis == 2 ? 1 : 2;
                        for (int k = 0; k < d.nterms; ++k) {
                            const int f = d.t[k].fl >> 1, lvl = d.t[k].fl & 1;
                            if (f < 3 && ph != 3 && lvl == 1)
                                return fail(OPESCI_EUNSUPPORTED,
                                            "gradient: a stress boundary statement reads velocities at t1");
                            // velocities read by the stress phase (level t0) need
                            // no swap fix-up: their transposes run after the swap
                            std::vector<int64_t>* out = f >= 3 ? &tc : (ph == 3 ? &vc : nullptr);
                            for (int p = 0; p < d.npts; ++p) {
                                const int64_t cell = (int64_t)(d.lo0 + p / d.n1) * so[o0] +
                                                     (int64_t)(d.lo1 + p % d.n1) * so[o1] + d.disp + d.t[k].disp;
                                const int x = (int)(cell / g.sx), y = (int)((cell % g.sx) / g.sy),
                                          z = (int)(cell % g.sy);
                                const bool inner = x >= m && x < g.nx - m && y >= m && y < g.ny - m && z >= m &&
                                                   z < g.nz - m;
                                if (!inner && lvl == 1) sc.push_back(cell * 16 + f);
                                if (!out || (inner && lvl == 1)) continue;   // (inner: moved by the swap)
                                out->push_back(cell * 32 + (inner ? 0 : 16) + f);
                            }
                        }
                    }
            for (std::vector<int64_t>* v : {&tc, &vc, &sc}) {
                std::sort(v->begin(), v->end());
                v->erase(std::unique(v->begin(), v->end()), v->end());
            }
            if (tc.size() > (size_t)INT32_MAX || vc.size() > (size_t)INT32_MAX || sc.size() > (size_t)INT32_MAX)
                return fail(OPESCI_EUNSUPPORTED, "too many boundary cells for the adjoint");
            if (int e = upload(&c->d_tcells, tc.data(), tc.size(), c->stream)) return e;
            if (int e = upload(&c->d_vcells, vc.data(), vc.size(), c->stream)) return e;
            if (int e = upload(&c->d_scells, sc.data(), sc.size(), c->stream)) return e;
            c->n_tcells = (int)tc.size();
            c->n_vcells = (int)vc.size();
            c->n_scells = (int)sc.size();
        }
        if (int e = build_adj_gathers(c)) return e;
        int64_t off = 0;
        for (int ph = 2; ph < 5; ++ph) {
            c->log_off[ph].clear();
            for (const BcLevel& lv : c->levels[ph]) {
                c->log_off[ph].push_back(off);
                off += (int64_t)lv.count * lv.maxpts;
            }
        }
        c->log_step = off;
    }
    if (c->log_step > 0 && log_steps > c->log_cap) {
        if (c->d_log) OPESCI_HIP_TRY(hipFree(c->d_log));
        c->d_log = nullptr;
        OPESCI_HIP_TRY(hipMalloc(&c->d_log, (size_t)c->log_step * log_steps * rb));
        c->log_cap = log_steps;
    }
    if (log_steps > c->srclog_cap) {
        if (c->d_srclog) OPESCI_HIP_TRY(hipFree(c->d_srclog));
        c->d_srclog = nullptr;
        OPESCI_HIP_TRY(hipMalloc(&c->d_srclog, (size_t)6 * log_steps * rb));
        c->srclog_cap = log_steps;
    }
    const size_t state = (size_t)2 * NF * c->field_bytes;
    if (state * ckpt > c->ckpt_cap) {
        if (c->d_ckpt) OPESCI_HIP_TRY(hipFree(c->d_ckpt));
        c->d_ckpt = nullptr;
        OPESCI_HIP_TRY(hipMalloc((void**)&c->d_ckpt, state * ckpt));
        c->ckpt_cap = state * ckpt;
    }
    if (seg_bytes > c->seg_cap) {
        if (c->d_seg) OPESCI_HIP_TRY(hipFree(c->d_seg));
        c->d_seg = nullptr;
        OPESCI_HIP_TRY(hipMalloc((void**)&c->d_seg, seg_bytes));
        c->seg_cap = seg_bytes;
    }
    return 0;
}

// copy the 18 field levels to / from a state slot
static int state_copy(OpesciContext* c, char* slot, bool save) {
    for (int f = 0; f < NF; ++f)
        for (int l = 0; l < 2; ++l) {
            char* s = slot + (size_t)(2 * f + l) * c->field_bytes;
            OPESCI_HIP_TRY(hipMemcpyAsync(save ? s : (char*)c->base[f][l], save ? c->base[f][l] : s, c->field_bytes,
                                          hipMemcpyDeviceToDevice, c->stream));
        }
    return 0;
}

// one forward step up to its velocity boundary levels on the fields of the
// FieldPtrs image `fp`, with every boundary level's overwritten values
// logged into `log` (one step's log)
static int step_logged(OpesciContext* c, void* fp, int t0, char* log) {
    const int t1 = 1 - t0;
    auto levels = [&](int ph) -> int {
        for (size_t L = 0; L < c->levels[ph].size(); ++L)
            OPESCI_HIP_TRY(launch_bc_logged(c->g, c->real_bytes, fp, c->d_desc, c->d_pool, c->levels[ph][L],
                                            log + c->log_off[ph][L] * c->real_bytes, t0, t1, c->stream));
        return 0;
    };
    OPESCI_HIP_TRY(launch_stress(c->g, c->real_bytes, c->storage, c->het, fp, c->mptr, c->coefs(), t0, t1, c->cfg,
                                 c->stream));
    if (int e = levels(4)) return e;
    if (int e = levels(2)) return e;
    OPESCI_HIP_TRY(launch_velocity(c->g, c->real_bytes, c->storage, c->het, fp, c->mptr, c->coefs(), t0, t1, c->cfg,
                                   c->stream));
    // no receivers / source here: the adjoint reads the state the velocity
    // boundary left (the source, added after it, is a constant forcing with
    // an identity adjoint, and must not show in the stresses it reads)
    return levels(3);
}

// adjoint of step n; `fp` (a FieldPtrs image) holds that step's logged
// forward result, `log` its undo log
static int step_adjoint(OpesciContext* c, int n, void* fp, char* log) {
    const int t0 = n % 2, t1 = 1 - t0;
    auto levels = [&](int ph) -> int {
        for (int L = (int)c->levels[ph].size() - 1; L >= 0; --L) {
            const BcLevel& lv = c->levels[ph][L];
            // transposes, with the level's forward values restored from its log
            OPESCI_HIP_TRY(launch_adj_bc(c->g, c->real_bytes, fp, c->aptr, c->d_desc, c->d_pool, c->d_gpool, lv,
                                         c->gat_bc[ph][L], log + c->log_off[ph][L] * c->real_bytes, t0, t1,
                                         c->stream));
        }
        return 0;
    };
    OPESCI_HIP_TRY(launch_adj_receivers(c->g, c->real_bytes, fp, c->aptr, t1, c->nrgroups, c->rg_start, c->rg_cell,
                                        c->rg_order, c->d_obs + (size_t)n * c->nrec * 3, c->stream));
    if (int e = levels(3)) return e;
    OPESCI_HIP_TRY(launch_adj_velocity(c->g, c->real_bytes, fp, c->aptr, c->mptr, c->gptr, c->coefs(), c->sptr,
                                       AdjCells{c->d_vcells, c->n_vcells}, c->cfg, t0, t1, c->stream));
    if (int e = levels(2)) return e;
    if (int e = levels(4)) return e;
    OPESCI_HIP_TRY(launch_adj_stress(c->g, c->real_bytes, fp, c->aptr, c->mptr, c->gptr, c->coefs(), c->sptr,
                                     AdjCells{c->d_tcells, c->n_tcells}, c->cfg, t0, t1, c->stream));
    return 0;
}

// every stepping boundary statement writes level t1 (so a step leaves level t0 as it found it)
static bool steps_keep_t0(const OpesciContext* c) {
    for (const OpesciBcStmt& st : c->stmts)
        if (st.phase >= 2 && st.level == 0) return false;
    return true;
}

// Forward pass of the gradient: checkpoints every `every` steps and the
// misfit J of the recorded U, V, W traces.  Leaves the final state in the
// fields (a valid stepped state) and the checkpoints for
// opesci_ctx_gradient_reverse.
int opesci_ctx_misfit_forward(OpesciContext* c, const float* observed, int64_t count, int32_t every,
                              double* misfit) {
    return guarded([&]() -> int {
        if (!c || !observed || every <= 0) return fail(OPESCI_EINVAL, "bad arguments");
        if (!c->het || c->storage != OPESCI_STORE_NATIVE)
            return fail(OPESCI_EUNSUPPORTED, "gradients need the read path (media arrays) and native storage");
        if (c->nrec <= 0) return fail(OPESCI_EINVAL, "gradients need receivers");
        const int N = c->desc.ntsteps;
        if (count != (int64_t)N * c->nrec * 3) return fail(OPESCI_EINVAL, "observed must hold ntsteps * nrec * 3 values");
        c->grad_every = 0;
        // a failure part way leaves a partial state: stepping needs
        // opesci_ctx_init again
        struct InitOnFailure {
            OpesciContext* c;
            bool ok = false;
            ~InitOnFailure() {
                if (!ok) {
                    c->initialised = false;
                    c->steps = 0;
                }
            }
        } guard{c};
        const int nck = (N + every - 1) / every;
        const size_t state = (size_t)2 * NF * c->field_bytes;
        // Segment scheme.  keep_t0: the segment's forward pass logs every step and
        // writes its level t1 (nine arrays) into a slot of its own; the reverse
        // step runs on pointers into those slots (level t0 = the previous step's
        // slot, or the checkpoint).
        // Otherwise (and with the simple launch variant 1, which selects the
        // reference forms throughout): full states before each step, and the step
        // recomputed with its log before its adjoint.
        const bool keep_t0 = steps_keep_t0(c) && c->cfg.variant != 1;
        const size_t slot = keep_t0 ? (size_t)NF * c->field_bytes : state;
        if (int e = adj_alloc(c, std::max(nck, 1), slot * every, keep_t0 ? every : 1)) return e;
        if (!c->d_obs) OPESCI_HIP_TRY(hipMalloc((void**)&c->d_obs, (size_t)count * sizeof(float)));
        OPESCI_HIP_TRY(hipMemcpyAsync(c->d_obs, observed, (size_t)count * sizeof(float), hipMemcpyDefault, c->stream));
        std::vector<float> obs_host((size_t)count);
        OPESCI_HIP_TRY(hipMemcpyAsync(obs_host.data(), c->d_obs, (size_t)count * sizeof(float), hipMemcpyDeviceToHost,
                                      c->stream));
        // forward with checkpoints
        if (int e = ctx_init(c)) return e;
        for (int n = 0; n < N; ++n) {
            if (n % every == 0)
                if (int e = state_copy(c, c->d_ckpt + (size_t)(n / every) * state, true)) return e;
            if (int e = step_once(c, n % 2)) return e;
        }
        c->steps = N;
        // misfit from the recorded U, V, W traces
        std::vector<double> tr((size_t)N * c->nrec * 4);
        if (int e = receivers_as(c, tr.data(), (int64_t)tr.size())) return e;
        double J = 0;
        for (int64_t i = 0; i < (int64_t)N * c->nrec; ++i)
            for (int k = 0; k < 3; ++k) {
                const double r = tr[(size_t)i * 4 + k] - (double)obs_host[(size_t)i * 3 + k];
                J += 0.5 * r * r;
            }
        if (misfit) *misfit = J;
        c->grad_every = every;
        c->grad_keep_t0 = keep_t0;
        guard.ok = true;
        return 0;
    });
}

// Reverse sweep after opesci_ctx_misfit_forward: media gradients into the
// context's gradient arrays.  Leaves a recomputed early state in the fields.
int opesci_ctx_gradient_reverse(OpesciContext* c) {
    return guarded([&]() -> int {
        if (!c) return fail(OPESCI_EINVAL, "null context");
        if (c->grad_every <= 0) return fail(OPESCI_ESTATE, "no forward pass: call opesci_ctx_misfit_forward first");
        // the fields end up holding a recomputed state (or, after a failure, a
        // partial one): on every return, stepping needs opesci_ctx_init again,
        // and a new forward pass precedes the next reverse sweep
        struct NeedsInit {
            OpesciContext* c;
            ~NeedsInit() {
                c->initialised = false;
                c->steps = 0;
                c->grad_every = 0;
            }
        } needs_init{c};
        const int N = c->desc.ntsteps, every = c->grad_every;
        const int nck = (N + every - 1) / every;
        const size_t state = (size_t)2 * NF * c->field_bytes;
        const bool keep_t0 = c->grad_keep_t0;
        const size_t slot = keep_t0 ? (size_t)NF * c->field_bytes : state;
        // reverse sweep, segment by segment
        for (int f = 0; f < NF; ++f)
            for (int l = 0; l < 2; ++l) OPESCI_HIP_TRY(launch_zero(c->abase[f][l], c->field_bytes, c->stream));
        const size_t mbytes = (size_t)c->g.nx * c->g.sx * c->real_bytes;
        for (int i = 0; i < OPESCI_M_COUNT; ++i) OPESCI_HIP_TRY(launch_zero(c->gbase[i], mbytes, c->stream));
        if (c->d_gpool)
            OPESCI_HIP_TRY(launch_zero(c->d_gpool, ((size_t)c->pool_elems * c->real_bytes + 15) & ~size_t(15), c->stream));
        const size_t rb = c->real_bytes, zo = (size_t)c->g.zoff * rb;
        const int* si = c->src_idx;
        char* srclog = static_cast<char*>(c->d_srclog);
        // keep_t0: slot k of a segment holds level t1 of step k0 + k, written in
        // place by the logged forward (the step reads level t0 from the slot
        // before it, or from the checkpoint).  Cells no step writes hold a
        // constant per level, so a slot is filled once with the whole level of
        // its parity (slot_par; once per gradient when `every` is even).  Cells
        // a boundary statement reads at level t1 before a later statement of the
        // step writes them hold the value of two steps before, copied in from
        // the slot two back (or the checkpoint) ahead of the step.
        std::vector<int> slot_par(keep_t0 ? every : 0, -1);
        const size_t fb = c->field_bytes;
        for (int seg = nck - 1; seg >= 0; --seg) {
            const int k0 = seg * every, k1 = std::min(N, k0 + every);
            char* ck = c->d_ckpt + (size_t)seg * state;
            OPESCI_HIP_TRY(launch_set_step(c->d_step, k0, c->stream));
            if (!keep_t0) {
                if (int e = state_copy(c, ck, false)) return e;
                for (int n = k0; n < k1; ++n) {
                    if (int e = state_copy(c, c->d_seg + (size_t)(n - k0) * state, true)) return e;
                    if (int e = step_once(c, n % 2)) return e;
                }
                for (int n = k1 - 1; n >= k0; --n) {
                    if (int e = state_copy(c, c->d_seg + (size_t)(n - k0) * state, false)) return e;
                    if (int e = step_logged(c, c->fptr, n % 2, static_cast<char*>(c->d_log))) return e;
                    if (int e = step_adjoint(c, n, c->fptr, static_cast<char*>(c->d_log))) return e;
                }
                continue;
            }
            // level pointers of step n: t1 in slot n - k0, t0 in the slot before
            // (or the checkpoint)
            auto slot_image = [&](int n, void* fp[NF][2]) {
                const int k = n - k0, t0 = n % 2, t1 = 1 - t0;
                for (int f = 0; f < NF; ++f) {
                    fp[f][t1] = c->d_seg + (size_t)k * slot + (size_t)f * fb + zo;
                    fp[f][t0] = k > 0 ? c->d_seg + (size_t)(k - 1) * slot + (size_t)f * fb + zo
                                      : ck + (size_t)(2 * f + t0) * fb + zo;
                }
            };
            for (int n = k0; n < k1; ++n) {
                const int k = n - k0, t0 = n % 2, t1 = 1 - t0;
                char* sk = c->d_seg + (size_t)k * slot;
                if (slot_par[k] != t1) {
                    for (int f = 0; f < NF; ++f)
                        OPESCI_HIP_TRY(hipMemcpyAsync(sk + (size_t)f * fb, ck + (size_t)(2 * f + t1) * fb, fb,
                                                      hipMemcpyDeviceToDevice, c->stream));
                    slot_par[k] = t1;
                }
                void* fp[NF][2];
                slot_image(n, fp);
                void* fs[NF][2];
                for (int f = 0; f < NF; ++f)
                    fs[f][0] = fs[f][1] = k >= 2 ? c->d_seg + (size_t)(k - 2) * slot + (size_t)f * fb + zo
                                                 : ck + (size_t)(2 * f + t1) * fb + zo;
                OPESCI_HIP_TRY(launch_copy_cells(c->real_bytes, fp, t1, fs, t1, c->d_scells, c->n_scells, c->stream));
                char* log = static_cast<char*>(c->d_log) + (size_t)k * c->log_step * rb;
                char* sl = srclog + (size_t)6 * k * rb;
                if (int e = step_logged(c, fp, t0, log)) return e;
                if (c->src_ntr)
                    OPESCI_HIP_TRY(launch_src_cells(c->g, c->real_bytes, fp, t1, si[0], si[1], si[2], sl, 0, c->stream));
                // source and step counter only (the traces are already recorded)
                OPESCI_HIP_TRY(launch_record_source(c->g, c->real_bytes, c->storage, fp, t1, 0, c->d_rec, c->d_trace,
                                                    c->src_ntr, c->src_idx, c->d_wavelet, c->d_step, c->desc.ntsteps,
                                                    c->stream));
                if (c->src_ntr)
                    OPESCI_HIP_TRY(launch_src_cells(c->g, c->real_bytes, fp, t1, si[0], si[1], si[2], sl + 3 * rb, 0,
                                                    c->stream));
            }
            for (int n = k1 - 1; n >= k0; --n) {
                const int t1 = 1 - n % 2;
                void* fp[NF][2];
                slot_image(n, fp);
                // the slots hold the injected source (what the next step read);
                // the adjoint of step n reads the state before the injection
                if (c->src_ntr)
                    OPESCI_HIP_TRY(launch_src_cells(c->g, c->real_bytes, fp, t1, si[0], si[1], si[2],
                                                    srclog + (size_t)6 * (n - k0) * rb, 1, c->stream));
                char* log = static_cast<char*>(c->d_log) + (size_t)(n - k0) * c->log_step * rb;
                if (int e = step_adjoint(c, n, fp, log)) return e;
            }
        }
        // media-dependent boundary coefficients -> media cells, one launch per
        // (statement, term, derivative program) of the stepping phases (the
        // only ones whose coefficient gradients the transposes accumulate)
        for (int ph = 2; ph < 5; ++ph)
            for (const BcLevel& lv : c->levels[ph])
                for (int di = lv.first; di < lv.first + lv.count; ++di) {
                    const BcDesc& d = c->h_desc[di];
                    for (int k = 0; k < d.nterms; ++k)
                        if (d.t[k].cslot >= 0)
                            for (int q = d.t[k].dterm0; q < d.t[k].dterm0 + d.t[k].ndterm; ++q)
                                OPESCI_HIP_TRY(launch_adj_coef_media(c->g, c->real_bytes, c->mptr, c->gptr, c->d_desc,
                                                                     c->d_code, c->d_dterms, c->d_gpool, di, d.npts, k, q,
                                                                     c->stream));
                }
        OPESCI_HIP_TRY(hipStreamSynchronize(c->stream));
        OPESCI_HIP_TRY(hipGetLastError());
        return 0;
    });
}

int opesci_ctx_gradient(OpesciContext* c, const float* observed, int64_t count, int32_t every, double* misfit) {
    return guarded([&]() -> int {
        if (int e = opesci_ctx_misfit_forward(c, observed, count, every, misfit)) return e;
        return opesci_ctx_gradient_reverse(c);
    });
}

int opesci_ctx_download_gradient(OpesciContext* c, int32_t media, void* host) {
    return guarded([&]() -> int {
        if (!c || !host || media < 0 || media >= OPESCI_M_COUNT) return fail(OPESCI_EINVAL, "bad arguments");
        if (!c->gbase[media]) return fail(OPESCI_ESTATE, "no gradient: call opesci_ctx_gradient first");
        const size_t dense = (size_t)c->g.nx * c->g.ny * c->g.nz;
        OPESCI_HIP_TRY(launch_unpack(c->g, c->real_bytes, OPESCI_STORE_NATIVE, c->gptr[media], c->d_stage, c->stream));
        OPESCI_HIP_TRY(hipMemcpyAsync(host, c->d_stage, dense * c->real_bytes, hipMemcpyDefault, c->stream));
        OPESCI_HIP_TRY(hipStreamSynchronize(c->stream));
        return 0;
    });
}

int opesci_ctx_download_media(OpesciContext* c, int32_t media, void* host) {
    return guarded([&]() -> int {
        if (!c || !host || media < 0 || media >= OPESCI_M_COUNT) return fail(OPESCI_EINVAL, "bad arguments");
        if (!c->het || !c->mptr[media]) return fail(OPESCI_EUNSUPPORTED, "media arrays exist on the read path only");
        const size_t dense = (size_t)c->g.nx * c->g.ny * c->g.nz;
        OPESCI_HIP_TRY(launch_unpack(c->g, c->real_bytes, OPESCI_STORE_NATIVE, c->mptr[media], c->d_stage, c->stream));
        OPESCI_HIP_TRY(hipMemcpyAsync(host, c->d_stage, dense * c->real_bytes, hipMemcpyDefault, c->stream));
        OPESCI_HIP_TRY(hipStreamSynchronize(c->stream));
        return 0;
    });
}

int opesci_ctx_upload_media(OpesciContext* c, int32_t media, const void* host) {
    return guarded([&]() -> int {
        if (!c || !host || media < 0 || media >= OPESCI_M_COUNT) return fail(OPESCI_EINVAL, "bad arguments");
        if (!c->het || !c->mptr[media]) return fail(OPESCI_EUNSUPPORTED, "media arrays exist on the read path only");
        c->grad_every = 0;
        const size_t dense = (size_t)c->g.nx * c->g.ny * c->g.nz;
        OPESCI_HIP_TRY(hipMemcpyAsync(c->d_stage, host, dense * c->real_bytes, hipMemcpyDefault, c->stream));
        OPESCI_HIP_TRY(launch_pack_media(c->g, c->real_bytes, c->d_stage, c->mptr[media], c->stream));
        if (c->d_pool)
            for (int ph = 0; ph < 5; ++ph)
                for (const BcLevel& lv : c->levels[ph])
                    OPESCI_HIP_TRY(launch_bc_coef(c->g, c->real_bytes, c->mptr, c->d_desc, c->d_code, c->d_pool, lv.first,
                                                  lv.count, lv.maxpts, c->stream));
        OPESCI_HIP_TRY(hipStreamSynchronize(c->stream));
        return 0;
    });
}

// ---------------------------------------------------------------- batches
// Independent contexts stepped together: one graph holds two time steps of
// every context as parallel branches (fork from and join into the batch's
// own stream), so small grids of different shapes and orders fill the GPU
// side by side and one graph launch advances them all.
struct OpesciBatch {
    std::vector<OpesciContext*> ctx;
    hipStream_t origin = nullptr;
    std::vector<hipEvent_t> ev;  // one join event per context + the fork event
    hipGraphExec_t graph = nullptr;
};

static void batch_free(OpesciBatch* b) {
    if (!b) return;
    if (b->origin) (void)hipStreamSynchronize(b->origin);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (hipEvent_t e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->origin) (void)hipStreamDestroy(b->origin);
    delete b;
}

static int batch_graph(OpesciBatch* b) {
    if (b->graph) return 0;
    const int n = (int)b->ctx.size();
    hipGraph_t gr = nullptr;
    OPESCI_HIP_TRY(hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
    int e = 0;
    hipError_t he = hipEventRecord(b->ev[n], b->origin);
    for (int i = 0; i < n && !e && he == hipSuccess; ++i) {
        OpesciContext* c = b->ctx[i];
        he = hipStreamWaitEvent(c->stream, b->ev[n], 0);
        if (he == hipSuccess) e = step_once(c, 0);
        if (!e && he == hipSuccess) e = step_once(c, 1);
        if (!e && he == hipSuccess) he = hipEventRecord(b->ev[i], c->stream);
        if (!e && he == hipSuccess) he = hipStreamWaitEvent(b->origin, b->ev[i], 0);
    }
    hipError_t he2 = hipStreamEndCapture(b->origin, &gr);
    if (e || he != hipSuccess || he2 != hipSuccess) {
        if (gr) (void)hipGraphDestroy(gr);
        if (e) return e;
        return fail(OPESCI_EHIP, std::string("batch capture: ") + hipGetErrorString(he != hipSuccess ? he : he2));
    }
    he = hipGraphInstantiate(&b->graph, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (he != hipSuccess) {
        b->graph = nullptr;
        return fail(OPESCI_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(he));
    }
    return 0;
}

int opesci_batch_create(OpesciContext** ctxs, int32_t n, OpesciBatch** out) {
    return guarded([&]() -> int {
        if (!ctxs || !out || n <= 0) return fail(OPESCI_EINVAL, "batch needs at least one context");
        *out = nullptr;
        for (int i = 0; i < n; ++i) {
            if (!ctxs[i]) return fail(OPESCI_EINVAL, "null context in batch");
            for (int j = 0; j < i; ++j)
                if (ctxs[j] == ctxs[i]) return fail(OPESCI_EINVAL, "context listed twice in batch");
        }
        OpesciBatch* b = new OpesciBatch();
        b->ctx.assign(ctxs, ctxs + n);
        b->ev.assign(n + 1, nullptr);
        hipError_t he = hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking);
        for (int i = 0; i <= n && he == hipSuccess; ++i) he = hipEventCreateWithFlags(&b->ev[i], hipEventDisableTiming);
        if (he != hipSuccess) {
            batch_free(b);
            return fail(OPESCI_EHIP, std::string("batch setup: ") + hipGetErrorString(he));
        }
        *out = b;
        return 0;
    });
}

int opesci_batch_destroy(OpesciBatch* b) {
    return guarded([&]() -> int {
        batch_free(b);
        return 0;
    });
}

int opesci_batch_run(OpesciBatch* b, int32_t nsteps, double* seconds) {
    return guarded([&]() -> int {
        if (!b) return fail(OPESCI_EINVAL, "null batch");
        if (nsteps < 0) return fail(OPESCI_EINVAL, "negative step count");
        const int s0 = b->ctx[0]->steps;
        for (OpesciContext* c : b->ctx) {
            if (!c->initialised) return fail(OPESCI_ESTATE, "every context of the batch must be initialised");
            if (c->steps != s0) return fail(OPESCI_ESTATE, "contexts of a batch must be at the same step");
        }
        // the graph holds an even/odd step pair: bring an odd start to even first
        int left = nsteps;
        double extra = 0;
        if (s0 % 2 == 1 && left > 0) {
            for (OpesciContext* c : b->ctx) {
                double s = 0;
                if (int e = ctx_run(c, 1, &s)) return e;
                extra += s;
            }
            --left;
        }
        if (left >= 2)
            if (int e = batch_graph(b)) return e;
        hipEvent_t e0, e1;
        OPESCI_HIP_TRY(hipEventCreate(&e0));
        OPESCI_HIP_TRY(hipEventCreate(&e1));
        for (OpesciContext* c : b->ctx) OPESCI_HIP_TRY(hipStreamSynchronize(c->stream));
        OPESCI_HIP_TRY(hipEventRecord(e0, b->origin));
        int rc = 0;
        const int reps = left / 2;
        for (int r = 0; r < reps && !rc; ++r) {
            hipError_t he = hipGraphLaunch(b->graph, b->origin);
            if (he != hipSuccess) rc = fail(OPESCI_EHIP, std::string("hipGraphLaunch: ") + hipGetErrorString(he));
        }
        OPESCI_HIP_TRY(hipEventRecord(e1, b->origin));
        OPESCI_HIP_TRY(hipEventSynchronize(e1));
        float ms = 0;
        OPESCI_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) return rc;
        for (OpesciContext* c : b->ctx) c->steps += 2 * reps;
        if (left % 2 == 1)
            for (OpesciContext* c : b->ctx) {
                double s = 0;
                if (int e = ctx_run(c, 1, &s)) return e;
                extra += s;
            }
        if (seconds) *seconds = ms * 1e-3 + extra;
        return 0;
    });
}

// ---------------------------------------------------------------- reference-named ABI
int opesci_model_load(const OpesciModelDesc* desc) {
    return guarded([&]() -> int {
        if (int e = validate(desc)) return e;
        if (g_default) {
            destroy(g_default);
            g_default = nullptr;
        }
        // keep a deep copy: the generated TU owned its tables for the process lifetime
        const size_t dense = (size_t)desc->dims[0] * desc->dims[1] * desc->dims[2];
        const size_t sz_st = (size_t)desc->nstmts * sizeof(OpesciBcStmt), sz_te = (size_t)desc->nterms * sizeof(OpesciBcTerm);
        const size_t sz_co = (size_t)desc->ncode * sizeof(OpesciExprOp), sz_so = (size_t)desc->nsol * sizeof(OpesciExprOp);
        const size_t sz_me = desc->media_arrays ? 3 * dense * sizeof(float) : 0;
        const size_t sz_w = (size_t)std::max(desc->ntsteps, 1) * desc->src_on * sizeof(float);
        const size_t sz_r = (size_t)desc->nrec * 3 * sizeof(int32_t);
        const size_t sz_dt = (size_t)desc->ndterms * sizeof(OpesciBcDTerm);
        g_loaded_blob.assign(sz_st + sz_te + sz_co + sz_so + sz_me + sz_w + sz_r + sz_dt + 72, 0);
        char* p = g_loaded_blob.data();
        g_loaded = *desc;
        auto take = [&](const void* src, size_t n) -> void* {
            if (!n) return nullptr;
            void* dst = p;
            memcpy(dst, src, n);
            p += (n + 7) & ~size_t(7);
            return dst;
        };
        g_loaded.stmts = (const OpesciBcStmt*)take(desc->stmts, sz_st);
        g_loaded.terms = (const OpesciBcTerm*)take(desc->terms, sz_te);
        g_loaded.bc_code = (const OpesciExprOp*)take(desc->bc_code, sz_co);
        g_loaded.sol_code = (const OpesciExprOp*)take(desc->sol_code, sz_so);
        if (desc->media_arrays) {
            g_loaded.rho = (const float*)take(desc->rho, dense * sizeof(float));
            g_loaded.vp = (const float*)take(desc->vp, dense * sizeof(float));
            g_loaded.vs = (const float*)take(desc->vs, dense * sizeof(float));
        }
        g_loaded.src_wavelet = desc->src_on ? (const float*)take(desc->src_wavelet, sz_w) : nullptr;
        g_loaded.rec_idx = desc->nrec ? (const int32_t*)take(desc->rec_idx, sz_r) : nullptr;
        g_loaded.dterms = desc->ndterms ? (const OpesciBcDTerm*)take(desc->dterms, sz_dt) : nullptr;
        g_have_loaded = true;
        return 0;
    });
}

int opesci_execute(OpesciGrid* grid, OpesciProfiling* profiling) {
    return guarded([&]() -> int {
        if (!grid) return fail(OPESCI_EINVAL, "null grid");
        if (!g_have_loaded) return fail(OPESCI_ESTATE, "no model loaded: call opesci_model_load first");
        if (g_default) {
            destroy(g_default);
            g_default = nullptr;
        }
        if (int e = create(&g_loaded, &g_default)) return e;
        OpesciContext* c = g_default;
        if (int e = ctx_init(c)) return e;
        double secs = 0;
        if (int e = ctx_run(c, c->desc.ntsteps, &secs)) return e;
        const size_t dense = (size_t)c->g.nx * c->g.ny * c->g.nz;
        void** slots = (void**)grid;
        for (int f = 0; f < NF; ++f) {
            void* h = nullptr;
            if (posix_memalign(&h, 4096, 2 * dense * c->real_bytes) != 0 || !h)
                return fail(OPESCI_ENOMEM, "host allocation for the result fields failed");
            for (int l = 0; l < 2; ++l)
                if (int e = ctx_download(c, f, l, (char*)h + (size_t)l * dense * c->real_bytes)) {
                    free(h);
                    return e;
                }
            slots[f] = h;
        }
        if (profiling) {
            profiling->g_rtime = (float)secs;
            profiling->g_ptime = (float)secs;
            profiling->g_mflops =
                secs > 0 ? (float)(flops_per_point(c) * interior_points(c) * c->desc.ntsteps / secs * 1e-6) : 0.f;
        }
        return 0;
    });
}

int opesci_convergence(OpesciGrid* grid, void* conv) {
    return guarded([&]() -> int {
        if (!grid || !conv) return fail(OPESCI_EINVAL, "null argument");
        if (!g_default || !g_default->initialised)
            return fail(OPESCI_ESTATE, "convergence before execute: call opesci_execute first");
        double l2[NF];
        if (int e = ctx_l2(g_default, l2)) return e;
        for (int f = 0; f < NF; ++f) {
            if (g_default->real_bytes == 8)
                ((double*)conv)[f] = l2[f];
            else
                ((float*)conv)[f] = (float)l2[f];
        }
        return 0;
    });
}

int opesci_free(OpesciGrid* grid) {
    return guarded([&]() -> int {
        if (!grid) return fail(OPESCI_EINVAL, "null grid");
        void** slots = (void**)grid;
        for (int f = 0; f < NF; ++f) {
            free(slots[f]);
            slots[f] = nullptr;
        }
        return 0;
    });
}

}  // extern "C"
