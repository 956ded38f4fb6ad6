// engine.hip -- batched polar Newton-Raphson power flow on one MI355X (gfx950).
//
// The Newton iteration is the one of the reference's nr_klu (p3s/cpp/nr_klu.cpp:569-711):
// mismatch and Jacobian in polar variables, static-pivot LU on a frozen pattern with a
// pivot-ratio guard, Armijo backtracking (beta 0.5, at most 10 halvings, c = 1e-4, best trial
// taken when none passes). Every case of a batch iterates on its own.
//
// LAYOUT. A per-case quantity q of item i (bus, Jacobian entry, factor position, ...) is stored
// at [i * S + c], S = number of cases rounded up to 64: a wavefront covers 64 cases of one item,
// so every value load is one contiguous 512-byte line and every index (Ybus column, factor
// position, operand pair) is the same in all lanes.
//
// STATE. The voltage of bus i is kept both rectangular (Ve + j Vf, what the mismatch reads) and
// polar (Vm, Va, what the Newton step updates), with the base point (Vm0, Va0) of the step in
// progress. The mismatch is the current-injection form S = V conj(Y V); the Jacobian's magnitude
// columns are scaled by |V| (derivatives in ln |V|), so the step solves for dVm / Vm.
//
// CONTROL. One 32-bit word per case (ctl) holds its stage, line-search trial count, precision
// class and iteration count. The host runs passes; a pass is
//     [LU of the cases in STEP] -> k_begin / k_advance -> k_evaluate -> k_control -> counters
// and every kernel works only on the cases whose stage concerns it, so a case's backtracking
// trial, another's next Newton iteration and a newly loaded case's start point share one pass
// (serving slots). A whole-batch solve holds the cases in STEP while any case is in its line
// search, so that all cases take their k-th step in one LU launch.
//
// KERNELS:
//   k_begin       start voltage -> state of the cases in START
//   k_advance     V = base + alpha dx of the cases in STEP (alpha 1, base saved), TRIAL, COMMIT
//   k_evaluate    mismatch R = Sspec - V conj(I), the Jacobian (except for TRIAL cases) and the
//                 largest |R| of each (row tile, case); row tiles balanced by Ybus nonzeros,
//                 their CSR segment staged in LDS
//   k_control     per-case Newton bookkeeping (convergence, Armijo, iteration limit, precision
//                 vote) and the stage counters the host reads back
//   k_lu_blk, k_lu_stream   static-pivot LU + forward/backward substitution (see below)
//   k_resid, k_refine_*     FP64 residual and update of the mixed-precision refinement
//   k_islands, k_admittance_fill   N-1: islands, references, pins; per-case Ybus from magnitude/angle
//   k_emit        state -> complex V (n, B), NaN at unserved buses
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include <unistd.h>

#include "p3smi.h"
#include "plan.hpp"

namespace {

// ---- the per-case control word
//   bits  0..3   stage
//   bits  4..7   backtracking trials taken in the current step (0..10)
//   bits  8..9   precision class of the next factorization (mixed modes)
//   bit   10     refining: the case's FP32 step is still being refined (mixed mode)
//   bits 11..12  pivoting: frozen plan / needs a pivoting factorization / has its own pivot order /
//                pivots on the plan's diagonal (a pivoting factorization that chose every diagonal)
//   bit   13     rerun: the case is on its run from scratch with pivoting (no further rerun)
//   bit   14     pending: its step this iteration waits for the pivoting factorization
//   bits 16..31  Newton iterations so far (linear solves)
enum : unsigned {
    SG_IDLE = 0,    // free slot, or a padding column
    SG_START = 1,   // start voltage loaded, start point not evaluated yet
    SG_STEP = 2,    // evaluated and iterating: next a linear solve and the full step
    SG_TRIAL = 3,   // line search: move to base + alpha dx, mismatch only
    SG_COMMIT = 4,  // line search over: move to base + alpha dx, full evaluation
    SG_CONV = 5,    // converged
    SG_FAIL = 6,    // non-finite mismatch, or no usable factorization
    SG_STOP = 7     // iteration limit reached
};
// precision classes: FP64 this time; FP32; FP64 for good (after a slow or cut FP32 step or
// failed FP32 pivots); FP64 this once (an FP32 voter whose 64-case group takes an FP64 step)
enum : unsigned { PR_F64 = 0, PR_F32 = 1, PR_F64_ALWAYS = 2, PR_F64_ONCE = 3 };
constexpr unsigned CTL_REFINING = 1u << 10;
enum : unsigned { PV_NONE = 0, PV_NEED = 1, PV_HAVE = 2, PV_DIAG = 3 };
constexpr unsigned CTL_RERUN = 1u << 13, CTL_PENDING = 1u << 14;
constexpr int MAX_ITER_LIMIT = 65535;  // the iteration count has 16 bits

__host__ __device__ constexpr unsigned stage_of(unsigned w) { return w & 15u; }
__host__ __device__ constexpr unsigned trials_of(unsigned w) { return (w >> 4) & 15u; }
__host__ __device__ constexpr unsigned prec_of(unsigned w) { return (w >> 8) & 3u; }
__host__ __device__ constexpr int iters_of(unsigned w) { return (int)(w >> 16); }
__host__ __device__ constexpr unsigned set_stage(unsigned w, unsigned s) { return (w & ~15u) | s; }
__host__ __device__ constexpr unsigned set_trials(unsigned w, unsigned t) { return (w & ~(15u << 4)) | (t << 4); }
__host__ __device__ constexpr unsigned set_prec(unsigned w, unsigned p) { return (w & ~(3u << 8)) | (p << 8); }
__host__ __device__ constexpr unsigned set_iters(unsigned w, int it) { return (w & 0xFFFFu) | ((unsigned)it << 16); }
__host__ __device__ constexpr bool evaluated_stage(unsigned s) { return s >= SG_START && s <= SG_COMMIT; }
__host__ __device__ constexpr unsigned piv_of(unsigned w) { return (w >> 11) & 3u; }
__host__ __device__ constexpr unsigned set_piv(unsigned w, unsigned p) { return (w & ~(3u << 11)) | (p << 11); }
// in STEP on the plan's pivot order (frozen, or kept by a pivoting factorization that chose the
// diagonal everywhere), step not pending: what the static-pivot LU launches take
__host__ __device__ constexpr bool frozen_step(unsigned w) {
    return stage_of(w) == SG_STEP && (piv_of(w) == PV_NONE || piv_of(w) == PV_DIAG) && !(w & CTL_PENDING);
}
// in STEP and factored by the pivoting LU (a fresh factorization, or a refactorization on the
// case's own pivot order)
__host__ __device__ constexpr bool pivoting_step(unsigned w) {
    return stage_of(w) == SG_STEP && (piv_of(w) == PV_NEED || piv_of(w) == PV_HAVE);
}
// a case this pass advances and evaluates: START, TRIAL, COMMIT, and STEP unless its step waits
// for the pivoting factorization or the whole batch holds its STEP cases
__host__ __device__ constexpr bool moves(unsigned w, int hold) {
    return evaluated_stage(stage_of(w)) && !(stage_of(w) == SG_STEP && (hold || (w & CTL_PENDING)));
}
// The first run of a case ended without converging on the frozen plan: it starts again from its
// start voltage, factoring with pivoting from its first iteration (p3s/cpp/nr_klu.cpp:844-857).
__host__ __device__ constexpr unsigned rerun_word(unsigned w) {
    return set_iters(set_piv(set_prec(set_trials(set_stage(w & ~(CTL_PENDING | CTL_REFINING), SG_START), 0), PR_F64_ALWAYS),
                             PV_NEED),
                     0) |
           CTL_RERUN;
}

// Stage counters as the C ABI reports them (p3smi_batch_pass): active, converged, failed,
// backtrack, needj, stopped, idle, new.
constexpr int NSTAT = 8;
enum : int { CNT_ACTIVE = 0, CNT_CONV, CNT_FAIL, CNT_TRIAL, CNT_COMMIT, CNT_STOP, CNT_IDLE, CNT_START };
constexpr int CNT_PIVOT = NSTAT;   // internal: cases in STEP off the frozen plan (the pivoting solve)
constexpr int NCOUNT = NSTAT + 1;  // counters k_control writes
__host__ __device__ constexpr int counter_of(unsigned s) {
    return s == SG_START ? CNT_START : s == SG_STEP ? CNT_ACTIVE : s == SG_TRIAL ? CNT_TRIAL
         : s == SG_COMMIT ? CNT_COMMIT : s == SG_CONV ? CNT_CONV : s == SG_FAIL ? CNT_FAIL
         : s == SG_STOP ? CNT_STOP : CNT_IDLE;
}

// Pass plan of a batch whose passes run on the device (group solves, several passes per graph
// launch): what the next pass does, written by the last workgroup of k_control from the stage
// counters, read by every kernel of the pass. The host writes the first plan of each launch.
enum : int {
    PL_HOLD = 0,     // the cases in STEP wait (lock-step: some case is in its line search)
    PL_LU = 1,       // LU launch shape of the cases in STEP: 0 none, 1 the batch shape, 2 the few-case one
    PL_HALT = 2,     // nothing to do on the device: no live case, or a case waits for a pivoting solve
    PL_START = 3,    // cases in START
    PL_ADVANCE = 4,  // k_advance has work (a step, a trial or a commit)
    NPLAN = 8
};
// The plan for the pass after one whose counters are cnt (the decisions of the host's pass loop).
__host__ __device__ inline void plan_from_counts(const int* cnt, int few_max, int* plan) {
    const int live = cnt[CNT_START] + cnt[CNT_ACTIVE] + cnt[CNT_TRIAL] + cnt[CNT_COMMIT];
    const bool searching = cnt[CNT_TRIAL] + cnt[CNT_COMMIT] > 0;
    const bool hold = searching && cnt[CNT_ACTIVE] > 0;
    const int n_solve = hold ? 0 : cnt[CNT_ACTIVE];
    plan[PL_HOLD] = hold ? 1 : 0;
    plan[PL_LU] = n_solve == 0 ? 0 : (n_solve <= few_max ? 2 : 1);
    plan[PL_HALT] = (live == 0 || (!hold && cnt[CNT_PIVOT] > 0)) ? 1 : 0;
    plan[PL_START] = cnt[CNT_START] > 0 ? 1 : 0;
    plan[PL_ADVANCE] = (n_solve > 0 || searching) ? 1 : 0;
    for (int k = PL_ADVANCE + 1; k < NPLAN; ++k) plan[k] = 0;
}

// Workgroups of a launch over several batches ("segments"): segment s owns the workgroups
// [off[s], off[s + 1]), off[0] = 0, each range a multiple of 8 so that a segment's workgroups
// are dealt to the 8 XCDs as those of a launch of its own. Returns s (uniform in a workgroup).
__device__ __forceinline__ int seg_find(const int* __restrict__ off, int nseg, int b) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

constexpr double ARMIJO_C = 1e-4;
constexpr int LS_MAX_TRIALS = 10;
constexpr double REUSE_RCOND_RATIO = 1e-3;
// k_evaluate: workgroups of 4 wavefronts over one row tile; a tile's CSR segment (column,
// Jacobian positions) is staged in LDS when it has at most EV_NZ nonzeros
// (the host cuts tiles to fit, except around a single row with more).
constexpr int EV_WAVES = 4;
// compact one-case LU arenas per batch (LuCompact): the stragglers of a pass and single solves
constexpr int LU_COMPACT_SLOTS = 16;
constexpr int EV_NZ = 768;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(expr)                                                                                 \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw Error(P3SMI_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));            \
    } while (0)

thread_local std::string g_err;
std::atomic<unsigned long long> g_solve_ids{0};

// ===================================================================== device code
__device__ __forceinline__ size_t at(int item, int S, int c) { return (size_t)item * (size_t)S + (size_t)c; }

// largest of w and a, a NaN once seen kept
__device__ __forceinline__ double nan_max(double w, double a) { return (a > w || a != a) ? a : w; }

// Cases [0, B) start (START, no iterations, precision class p0, and `extra`: the pivoting bits
// when the frozen plan is not usable), padding columns [B, S) idle.
__global__ void k_ctl_reset(unsigned* __restrict__ ctl, int B, int S, unsigned p0, unsigned extra) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < S) ctl[c] = c < B ? (set_prec(SG_START, p0) | extra) : SG_IDLE;
}

// The zero and one rows of every compact one-case arena (LuCompact): slot w's rows start at
// w * frows.
__global__ void k_compact_rows(double* __restrict__ p, int cap, long long frows, long long zero) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= cap) return;
    p[(size_t)w * (size_t)frows + (size_t)zero] = 0.0;
    p[(size_t)w * (size_t)frows + (size_t)zero + 1] = 1.0;
}

__global__ void k_clear_int(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Start voltages of the cases in START: V0 (n) shared (v0_stride 0) or (n, v0_stride) complex,
// column c. The rectangular state is V0 itself; the polar one its modulus and argument. Sets the
// case's precision class p0 (one thread per case writes it). Workgroup (bx, by) of gy along the
// buses takes the 64 cases of group bx and the buses by * 4 + wave, stepping by gy * 4.
struct BeginArgs {
    const double* V0;
    int v0_stride;
    unsigned* ctl;
    unsigned p0;
    double *Vm, *Va, *Ve, *Vf;
    int n, S;
    int gx, gy;        // workgroups along the cases and along the buses
    const int* plan;   // passes driven on the device: skip unless the plan has cases in START
};

__device__ __forceinline__ void begin_body(const double* __restrict__ V0, int v0_stride, unsigned* __restrict__ ctl,
                                           unsigned p0, double* __restrict__ Vm, double* __restrict__ Va,
                                           double* __restrict__ Ve, double* __restrict__ Vf, int n, int S, int bx,
                                           int by, int gy) {
    const int c = bx * 64 + (threadIdx.x & 63);
    const unsigned w = ctl[c];
    if (stage_of(w) != SG_START) return;
    for (int bus = by * 4 + (int)(threadIdx.x >> 6); bus < n; bus += gy * 4) {
        const size_t src = 2 * (v0_stride ? (size_t)bus * v0_stride + c : (size_t)bus);
        const double re = V0[src], im = V0[src + 1];
        const size_t o = at(bus, S, c);
        Ve[o] = re;
        Vf[o] = im;
        Vm[o] = hypot(re, im);
        Va[o] = atan2(im, re);
        if (bus == 0) ctl[c] = set_prec(w, p0);  // the stage stays START: other threads may read it
    }
}

__global__ __launch_bounds__(256) void k_begin(BeginArgs a) {
    begin_body(a.V0, a.v0_stride, a.ctl, a.p0, a.Vm, a.Va, a.Ve, a.Vf, a.n, a.S, blockIdx.x, blockIdx.y, gridDim.y);
}

__global__ __launch_bounds__(256) void k_begin_multi(const BeginArgs* __restrict__ tab, const int* __restrict__ off,
                                                     int nseg) {
    const int s = seg_find(off, nseg, blockIdx.x);
    const BeginArgs& a = tab[s];
    const int b = blockIdx.x - off[s];
    if (b >= a.gx * a.gy || a.plan[PL_HALT] || !a.plan[PL_START]) return;
    begin_body(a.V0, a.v0_stride, a.ctl, a.p0, a.Vm, a.Va, a.Ve, a.Vf, a.n, a.S, b % a.gx, b / a.gx, a.gy);
}

// Per-case admittance G + jB at [k * S + c] of columns [0, nc): mode 0 copies the shared base
// values (G0, B0 indexed by k), mode 1 converts magnitude/angle (nnz, ld) host-layout arrays.
// One thread per (nonzero, case), cases along x.
__global__ __launch_bounds__(256) void k_admittance_fill(int mode, const double* __restrict__ p0,
                                                         const double* __restrict__ p1, int ld, int nc,
                                                         double* __restrict__ G, double* __restrict__ Bv, int nnz,
                                                         int S) {
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int k = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (k >= nnz || c >= S) return;
    const size_t o = at(k, S, c);
    if (mode == 0 || c >= nc) {
        G[o] = mode == 0 ? p0[k] : 0.0;
        Bv[o] = mode == 0 ? p1[k] : 0.0;
    } else {
        const double mag = p0[(size_t)k * ld + c], ang = p1[(size_t)k * ld + c];
        double s, co;
        sincos(ang, &s, &co);
        G[o] = mag * co;
        Bv[o] = mag * s;
    }
}

// Mismatch, Jacobian and per-tile infinity norm of every case that takes an evaluation this
// pass (stages START, STEP, TRIAL, COMMIT; TRIAL without the Jacobian).
//
// Bus i's current injection I_i = sum_k Y_ik V_k is accumulated over its Ybus row, and
//   S_i = V_i conj(I_i),   R = Sspec - S at the P row (pv, pq buses) and Q row (pq buses).
// With w_ik = V_i conj(Y_ik V_k) the derivatives (angle columns, and magnitude columns in
// ln |V|, i.e. multiplied by |V_k|) are, off the diagonal,
//   dP_i/dVa_k =  Im w_ik   dQ_i/dVa_k = -Re w_ik   dP_i/dln|V_k| = Re w_ik   dQ_i/dln|V_k| = Im w_ik
// and on it, with m2 = |V_i|^2 and Y_ii = g + jb,
//   dP_i/dVa_i = -Q_i - m2 b   dQ_i/dVa_i = P_i - m2 g   dP_i/dln|V_i| = P_i + m2 g
//   dQ_i/dln|V_i| = Q_i - m2 b
// (dS/dVa = j diag(V) conj(diag(I) - Y diag(V)), |V| dS/d|V| = diag(V) conj(Y diag(V)) +
// diag(conj(I)) diag(V) in matrix form; the polar derivatives of p3s/cpp/nr_klu.cpp:288-378
// times the column scale). A pinned bus (N-1 island reference, unserved bus) has identity rows
// and zero mismatch (p3s/cpp/nr_klu.cpp:398-423).
//
// A workgroup takes one row tile [tile_ptr[t], tile_ptr[t + 1]) and 64 cases; wavefront w walks
// the tile's rows w, w + 4, ...; the tile's column indices, Jacobian positions and (shared
// Ybus) admittances are read from LDS. part[t * S + c] = largest |R| of the tile's rows (NaN
// kept).
struct EvalArgs {
    const int* tile_ptr;
    const int* Yp;
    const int* Yj;
    const int* Ydiag;
    const int4* posJ;
    const double* G;   // (nnz) shared, or (nnz, S) per case
    const double* B;
    // per-case deltas on the shared admittance (YM 2): case c subtracts the values dval of its
    // entries [dptr[dc0 + c], dptr[dc0 + c + 1]) at the Ybus positions dpos, in list order
    const int* dptr;
    const int* dpos;
    const double* dval;  // complex
    int dc0;
    const double* Ve;
    const double* Vf;
    const double* Ssp;  // complex (n) shared (s_stride 0) or (n, s_stride)
    int s_stride;
    const int* row_p;
    const int* row_q;
    const unsigned char* pin;  // (n, S) or null
    const unsigned* ctl;
    int hold;                  // 1: the cases in STEP wait this pass (no LU ran for them)
    double* J;
    double* R;
    double* part;
    int n, S;
    int gx;                    // row tiles (workgroups along x)
    const int* plan;           // passes driven on the device: the hold comes from it
};

// YM: the admittance of the cases -- 0 shared, 1 per case ((nnz, S) arrays), 2 shared with
// per-case deltas (an outage case's stamps, subtracted as the evaluation reads the row: the
// same subtractions in the same order as a per-case copy would hold, without the copy).
enum : int { YM_SHARED = 0, YM_CASE = 1, YM_DELTA = 2 };

template <int YM, bool PIN>
__device__ __forceinline__ void evaluate_body(const EvalArgs& a, int hold, int bx, int by) {
    constexpr bool CASE_Y = YM == YM_CASE;
    // a shared admittance entry is read with a wave-uniform global load rather than staged in LDS:
    // the smaller block footprint keeps more waves resident for the voltage gathers (measured
    // 3-4% faster evaluations on the 2000- and 10000-bus workloads)
    __shared__ int s_col[EV_NZ];
    __shared__ int4 s_pos[EV_NZ];
    __shared__ double s_worst[EV_WAVES][64];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int S = a.S;
    const int c = by * 64 + lane;
    const int r0 = a.tile_ptr[bx], r1 = a.tile_ptr[bx + 1];
    const int k0 = a.Yp[r0], k1 = a.Yp[r1];
    const bool staged = k1 - k0 <= EV_NZ;
    if (staged) {
        for (int k = k0 + (int)threadIdx.x; k < k1; k += 64 * EV_WAVES) {
            s_col[k - k0] = a.Yj[k];
            s_pos[k - k0] = a.posJ[k];
        }
    }
    __syncthreads();
    const unsigned wc = a.ctl[c];
    const unsigned sg = stage_of(wc);
    const bool run = moves(wc, hold);
    const bool jac = run && sg != SG_TRIAL;
    double worst = 0.0;
    // YM_DELTA: the case's deltas, and whether any falls in this tile's nonzeros
    int d0 = 0, d1 = 0;
    bool dtile = false;
    if (YM == YM_DELTA && run) {
        d0 = a.dptr[a.dc0 + c];
        d1 = a.dptr[a.dc0 + c + 1];
        for (int d = d0; d < d1; ++d) dtile |= a.dpos[d] >= k0 && a.dpos[d] < k1;
    }
    if (run) {
        for (int row = r0 + wave; row < r1; row += EV_WAVES) {
            const size_t oi = at(row, S, c);
            const double ei = a.Ve[oi], fi = a.Vf[oi];
            const bool pinned = PIN && a.pin[oi];
            double ir = 0.0, ii = 0.0, gd = 0.0, bd = 0.0;
            const int q1 = a.Yp[row + 1] - k0;
            for (int q = a.Yp[row] - k0; q < q1; ++q) {
                const int col = staged ? s_col[q] : a.Yj[k0 + q];
                double g, b;
                if (CASE_Y) {
                    g = a.G[at(k0 + q, S, c)];
                    b = a.B[at(k0 + q, S, c)];
                } else {
                    g = a.G[k0 + q];
                    b = a.B[k0 + q];
                }
                if (YM == YM_DELTA && dtile)
                    for (int d = d0; d < d1; ++d)
                        if (a.dpos[d] == k0 + q) {
                            g -= a.dval[2 * (size_t)d];
                            b -= a.dval[2 * (size_t)d + 1];
                        }
                const size_t ok = at(col, S, c);
                const double ek = a.Ve[ok], fk = a.Vf[ok];
                const double tr = g * ek - b * fk, ti = g * fk + b * ek;  // Y_ik V_k
                ir += tr;
                ii += ti;
                if (col == row) {
                    gd = g;
                    bd = b;
                } else if (jac) {
                    const double wr = ei * tr + fi * ti, wi = fi * tr - ei * ti;  // V_i conj(Y_ik V_k)
                    const int4 pj = staged ? s_pos[q] : a.posJ[k0 + q];
                    if (pj.x >= 0) a.J[at(pj.x, S, c)] = pinned ? 0.0 : wi;
                    if (pj.y >= 0) a.J[at(pj.y, S, c)] = pinned ? 0.0 : wr;
                    if (pj.z >= 0) a.J[at(pj.z, S, c)] = pinned ? 0.0 : -wr;
                    if (pj.w >= 0) a.J[at(pj.w, S, c)] = pinned ? 0.0 : wi;
                }
            }
            const double P = ei * ir + fi * ii, Q = fi * ir - ei * ii;  // V_i conj(I_i)
            if (jac) {
                const int d = a.Ydiag[row];
                if (d >= 0) {
                    const int4 pj = (staged && d >= k0 && d < k1) ? s_pos[d - k0] : a.posJ[d];
                    const double m2 = ei * ei + fi * fi;
                    if (pj.x >= 0) a.J[at(pj.x, S, c)] = pinned ? 1.0 : -Q - m2 * bd;
                    if (pj.y >= 0) a.J[at(pj.y, S, c)] = pinned ? 0.0 : P + m2 * gd;
                    if (pj.z >= 0) a.J[at(pj.z, S, c)] = pinned ? 0.0 : P - m2 * gd;
                    if (pj.w >= 0) a.J[at(pj.w, S, c)] = pinned ? 1.0 : Q - m2 * bd;
                }
            }
            const size_t so = 2 * (a.s_stride ? (size_t)row * a.s_stride + c : (size_t)row);
            const int rp = a.row_p[row], rq = a.row_q[row];
            if (rp >= 0) {
                const double r = pinned ? 0.0 : a.Ssp[so] - P;
                a.R[at(rp, S, c)] = r;
                worst = nan_max(worst, fabs(r));
            }
            if (rq >= 0) {
                const double r = pinned ? 0.0 : a.Ssp[so + 1] - Q;
                a.R[at(rq, S, c)] = r;
                worst = nan_max(worst, fabs(r));
            }
        }
    }
    s_worst[wave][lane] = worst;
    __syncthreads();
    if (wave == 0 && run) {
        for (int v = 1; v < EV_WAVES; ++v) worst = nan_max(worst, s_worst[v][lane]);
        a.part[at(bx, S, c)] = worst;
    }
}

template <int YM, bool PIN>
__global__ __launch_bounds__(64 * EV_WAVES) void k_evaluate(EvalArgs a) {
    evaluate_body<YM, PIN>(a, a.hold, blockIdx.x, blockIdx.y);
}

template <int YM, bool PIN>
__global__ __launch_bounds__(64 * EV_WAVES) void k_evaluate_multi(const EvalArgs* __restrict__ tab,
                                                                  const int* __restrict__ off, int nseg) {
    const int s = seg_find(off, nseg, blockIdx.x);
    const EvalArgs& a = tab[s];
    const int b = blockIdx.x - off[s];
    if (b >= a.gx * (a.S / 64) || a.plan[PL_HALT]) return;
    evaluate_body<YM, PIN>(a, a.plan[PL_HOLD], b % a.gx, b / a.gx);
}

// Summation order shared by every LU kernel, so that a case's factors and step are the same
// bits whichever launch shape (cases per workgroup, task lists) solved it: one accumulator
// runs over the operand pairs of an entry in stream order from zero, s = fma(-a, b, s), and
// the start value (Jacobian entry, right-hand side) is added once at the end.
// (Each LU kernel below keeps this order; k_lu_blk's 2x2 blocks add a pair's own term last.)
//
// Static-pivot LU + substitution of the cases in STEP. A workgroup owns G consecutive cases and
// runs TL = 64 * NS / G task lists over them. Level l of the elimination tree:
//   sub-level A: pivots and U entries   acc = a - sum L(i,t) U(t,k); pivot -> 1/acc
//                forward rows            y_i = R[perm i] - sum L(i,j) y_j
//   barrier
//   sub-level B: L entries               acc = (a - sum L(i,t) U(t,k)) / U(k,k)
//   barrier
// then backward rows by depth: x_i = (y_i - sum U(i,j) x_j) / U(i,i), columns descending.
// FACTOR = true: factorize (real = double in the FP64 factor array, float in the FP32 one) and
// solve J x = rhs; FACTOR = false: solve only, with the factors of an earlier launch (the
// refinement solves). Substitutions accumulate in FP64 whatever the factor precision.

// Which cases an LU launch works on (besides being in STEP): every one (SEL_ALL, the FP64
// solve), the FP64 classes (SEL_F64), FP32 (SEL_F32), or FP32 and still refining (SEL_REFINE);
// SEL_NEED: the cases waiting for a fresh pivoting factorization, factored on the plan's
// diagonal as the first try of that factorization (k_pivot_verify decides whether it stands).
enum : int { SEL_ALL = 0, SEL_F64 = 1, SEL_F32 = 2, SEL_REFINE = 3, SEL_NEED = 4 };

__device__ __forceinline__ bool selected(int sel, unsigned w) {
    if (sel == SEL_NEED) return stage_of(w) == SG_STEP && piv_of(w) == PV_NEED;
    if (!frozen_step(w)) return false;
    const unsigned p = prec_of(w);
    switch (sel) {
        case SEL_F64: return p != PR_F32;
        case SEL_F32: return p == PR_F32;
        case SEL_REFINE: return p == PR_F32 && (w & CTL_REFINING) != 0u;
        default: return true;
    }
}

// Switching mixed mode, before its two LU launches. A wave factors 64 aligned cases, so it
// costs its full time in a launch as soon as one of them is selected: with FP32 and FP64
// classes mixed inside a wave, the wave would run in both launches. Here the unit is the wave:
// when any case of a 64-case group in STEP needs FP64, the group's FP32 cases take this step in
// FP64 as well (PR_F64_ONCE: a fresh vote after the step), so every group runs in one launch.
// The groups are fixed by case index, not by launch shape. Counts the FP32 and FP64
// factorizations into fcount[1] and fcount[2].
__global__ __launch_bounds__(256) void k_group_vote(unsigned* __restrict__ ctl, int* __restrict__ fcount, int S) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;  // blockDim 256: each wave is one group
    const unsigned w = c < S ? ctl[c] : SG_IDLE;
    const bool act = frozen_step(w);
    const bool f32 = act && prec_of(w) == PR_F32;
    const bool any64 = __ballot(act && !f32) != 0ull;
    if (f32 && any64) ctl[c] = set_prec(w, PR_F64_ONCE);
    const int n_act = __popcll(__ballot(act));
    const int n32 = any64 ? 0 : n_act;
    if ((threadIdx.x & 63) == 0 && n_act) {
        if (n32) atomicAdd(&fcount[1], n32);
        if (n_act - n32) atomicAdd(&fcount[2], n_act - n32);
    }
}

// Pivot check at the end of a factorization (p3s/cpp/lean_lu.hpp:146-177, p3s/cpp/nr_klu.cpp:
// 610-620), by the one thread that owns case c's statistics. FP32: a zero or non-finite pivot,
// or a pivot ratio under the guard, moves the case to FP64 for good; it is factored in FP64 in
// the same iteration (guard_trips[2] counts these). FP64: a zero or non-finite pivot or a ratio
// under the guard rejects the frozen plan for the rest of the case's run (guard_trips[0]): its
// step waits for the pivoting factorization (pending). A case on its kept diagonal order has no
// guard: a zero or non-finite pivot sends it back to a fresh pivoting factorization (the failed
// refactorization of p3s/cpp/nr_klu.cpp:629-640). A first try on the diagonal (SEL_NEED) leaves
// the control word to k_pivot_verify. Returns true when the case takes no step from these factors.
template <bool SINGLE>
__device__ __forceinline__ bool pivot_verdict(unsigned* ctl, int c, double lo, double hi, double sum,
                                              double rcond_min, int* guard_trips, int sel) {
    const bool broken = !(sum < DBL_MAX && lo > 0.0);
    if (sel == SEL_NEED) return broken;
    const unsigned w = ctl[c];
    if (!SINGLE && piv_of(w) == PV_DIAG) {
        if (broken) ctl[c] = set_piv(w, PV_NEED) | CTL_PENDING;
        return broken;
    }
    if (SINGLE) {
        if (broken || lo < rcond_min * hi) {
            ctl[c] = set_prec(w, PR_F64_ALWAYS);
            atomicAdd(guard_trips + 2, 1);
            return true;
        }
        return false;
    }
    if (broken || lo < rcond_min * hi) {
        ctl[c] = set_piv(w, PV_NEED) | CTL_PENDING;
        atomicAdd(guard_trips, 1);
        return true;
    }
    return false;
}

// ---- operand streams, software-pipelined ----------------------------------------------
// With 64 cases per workgroup a task list is one wavefront and its control flow is the
// same in every lane (with fewer cases per workgroup, each lane group walks its own list). The host lays the operand pairs of a task list out back to back in
// execution order (one stream per list and sub-level), so the wavefront can walk the
// stream in chunks of K pairs with the loads of the next chunk in flight while it
// multiplies the current one, across task boundaries: a task's end only stores a result.
// The start values of a task (its Jacobian entry, pivot inverse, right-hand side) are
// loaded one task ahead and added at its end. The memory-level parallelism of a wavefront
// is thus 2K operand pairs whatever the task lengths, where a task-at-a-time loop would
// wait for every task's first loads.
// Chunk length: 8 pairs when an operand is FP64, 16 when both are FP32 (the same registers
// and the same bytes in flight per wavefront).
template <typename TA, typename TB>
constexpr int lu_chunk() { return sizeof(TA) + sizeof(TB) <= 8 ? 16 : 8; }

template <int K, typename TA, typename TB>
__device__ __forceinline__ void load_chunk(const int2* __restrict__ st, int e, int eend, const TA* __restrict__ A,
                                           const TB* __restrict__ Bv, int S, int c, TA (&a)[K], TB (&b)[K]) {
#pragma unroll
    for (int u = 0; u < K; ++u) {
        const int2 p = st[min(e + u, eend - 1)];  // past the end: repeat the last pair (unused)
        a[u] = A[at(p.x, S, c)];
        b[u] = Bv[at(p.y, S, c)];
    }
}

// Tasks [q0, q1) of one list; task q = (w0, w1, pbeg, pend) with pend(q) == pbeg(q + 1).
// start(q, task) -> In issues the task's start loads; finish(task, In, sum) consumes
// sum = -sum_pairs a*b in type ACC.
template <int K, typename ACC, typename In, typename TA, typename TB, typename Start, typename Finish>
__device__ __forceinline__ void seg_stream(const int4* __restrict__ tasks, int q0, int q1,
                                           const int2* __restrict__ st, const TA* __restrict__ A,
                                           const TB* __restrict__ Bv, int S, int c, Start start, Finish finish) {
    if (q0 >= q1) return;
    int q = q0;
    int4 tk = tasks[q];
    In in = start(q, tk);
    int4 tkn = tk;
    In inn = in;
    if (q + 1 < q1) {
        tkn = tasks[q + 1];
        inn = start(q + 1, tkn);
    }
    const int eend = tasks[q1 - 1].w;
    ACC acc = 0;
    auto advance = [&]() {
        finish(tk, in, acc);
        acc = 0;
        ++q;
        if (q < q1) {
            tk = tkn;
            in = inn;
            if (q + 1 < q1) {
                tkn = tasks[q + 1];
                inn = start(q + 1, tkn);
            }
        }
    };
    auto consume = [&](int e, const TA (&a)[K], const TB (&b)[K]) {
#pragma unroll
        for (int u = 0; u < K; ++u) {
            const int pos = e + u;
            while (q < q1 && pos == tk.w) advance();
            if (pos < eend) acc = fma(-(ACC)a[u], (ACC)b[u], acc);
        }
    };
    // Two chunk buffers in turn, the loop unrolled by two: while a next chunk exists it is
    // loaded before the current one is consumed, so it is in flight meanwhile, and no register
    // is copied while a load into it may still be pending (which would make the wave wait for
    // every load in flight). No chunk past the end of the stream is loaded.
    int e = tk.z;
    if (e < eend) {
        TA a0[K], a1[K];
        TB b0[K], b1[K];
        load_chunk<K>(st, e, eend, A, Bv, S, c, a0, b0);
        bool last0 = true;  // the chunk in buffer 0 is the last one
        while (e + K < eend) {
            load_chunk<K>(st, e + K, eend, A, Bv, S, c, a1, b1);
            consume(e, a0, b0);
            e += K;
            if (e + K >= eend) {
                last0 = false;
                break;
            }
            load_chunk<K>(st, e + K, eend, A, Bv, S, c, a0, b0);
            consume(e, a1, b1);
            e += K;
        }
        if (last0) consume(e, a0, b0);
        else consume(e, a1, b1);
    }
    while (q < q1) advance();
}

struct In1 { double v; };
struct In2 { double v, d; };

// task marker of the one-entry schedule: a chunk of a cut sum (its sum goes, negated, to its
// partial slot; the next chunk's first operand pair (slot, one row) brings it back exactly)
constexpr int LU_PART = INT_MIN;

// The LU with one-entry tasks. A workgroup owns G consecutive cases (G = 64, 32, 16 or 1) and
// TL = 64 * NS / G task lists; lane t works on case t % G and list t / G. With G = 64 a wavefront is one list and every loop bound
// is uniform (scalar); with G = 32 or 16 the two or four lists of a wavefront walk their
// streams side by side (divergent only where their task boundaries differ), which gives a
// batch of few cases more lists per case and more workgroups for the compute units.
// The arguments of an LU launch over one batch (either kernel's schedule arrays), for launches
// that cover several batches (k_lu_stream_multi, k_lu_blk_multi: one table entry per batch).
struct LuArgs {
    const int* ptrA;
    const int4* taskA;
    const int* ptrB;
    const int4* taskB;
    const int* diagB;    // one-entry tasks
    const int2* pairs;   // one-entry tasks
    const int4* steps;   // 2x2 blocks
    const int* ptrF;
    const int4* fwd;
    const int2* fent;    // one-entry tasks
    const int4* fsteps;  // 2x2 blocks
    const int* ptrW;
    const int4* bwd;
    const int2* bent;    // one-entry tasks
    const int4* bsteps;  // 2x2 blocks
    const double* Jx;
    void* fac;
    const double* rhs;
    double* x;
    unsigned* ctl;
    double* pstat;
    int* guard;
    int sel;
    double rcond_min;
    int nlev, nblev, S;
    int gx;             // workgroups over the batch's cases (S / G)
    const int* plan;    // passes driven on the device: run when plan[PL_LU] == want
    int want;
};

// Compact arenas of one-case launches (G = 1, FP64 factor + solve): a workgroup that takes a
// slot (w < cap, from a counter the host clears before the launch) factors and substitutes in
// rows of stride 1 of its own -- fac rows [w * frows, (w + 1) * frows), x rows [w * xrows, ...)
// -- instead of the batch's rows of stride S, where each of its operands is alone in a
// 128-byte line: its factors (8 nnz_LU bytes) then stay in one L2 instead of spanning
// nnz_LU lines. The same arithmetic in the same order; the solution goes to the batch's x at
// the end. Nothing reads the factors of such a launch afterwards. cap 0: batch rows.
struct LuCompact {
    double* fac = nullptr;  // cap x frows, the zero and one rows of every slot set
    double* x = nullptr;    // cap x xrows
    int* ctr = nullptr;
    int cap = 0;
    long long frows = 0, xrows = 0;
    int m = 0;
};

template <int G, int NS, typename real, bool FACTOR>
__device__ __forceinline__ void lu_stream_body(
    const int* __restrict__ ptrA, const int4* __restrict__ taskA, const int* __restrict__ ptrB,
    const int4* __restrict__ taskB, const int* __restrict__ diagB, const int2* __restrict__ pairs,
    const int* __restrict__ ptrF, const int4* __restrict__ fwd, const int2* __restrict__ fent,
    const int* __restrict__ ptrW, const int4* __restrict__ bwd, const int2* __restrict__ bent,
    const double* __restrict__ Jx, real* __restrict__ fac_b, const double* __restrict__ rhs, double* __restrict__ x_b,
    unsigned* __restrict__ ctl, double* __restrict__ pstat, int* __restrict__ guard_trips, int sel, double rcond_min,
    int nlev, int nblev, int S, int bx, int gx, const LuCompact& cmp = LuCompact{}) {
    constexpr bool SINGLE = sizeof(real) == 4;
    constexpr int TL = 64 * NS / G;
    // fewer than 64 cases per workgroup: lists differ within a wavefront, so their indices and
    // addresses sit in vector registers -- shorter chunks keep 16 wavefronts within 128
    // registers: with one case per workgroup a quarter with FP64 factors, half with FP32 ones;
    // half for FP32 factors with 16 or 32 cases per workgroup. One case on 4 wavefronts (a
    // straggler, a single solve) has registers to spare for chunks twice the batch length:
    // the long Crout sums near the top of the elimination tree are serial chains whose time
    // is the load latency per chunk, so the deeper chunks shorten the critical path.
    constexpr int KDIV = G == 1 ? (NS == 4 ? 1 : (SINGLE ? 2 : 4)) : (G < 64 && SINGLE ? 2 : 1);
    constexpr int KMUL = G == 1 && NS == 4 ? 2 : 1;
    constexpr int K = lu_chunk<real, real>() * KMUL / KDIV;
    constexpr int KF = lu_chunk<real, double>() * KMUL / KDIV;
    __shared__ double smin[TL][G], smax[TL][G], ssum[TL][G];
    __shared__ int sbad[G];
    const int lane = threadIdx.x % G;
    const int slot = G == 64 ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x / G)) : (int)(threadIdx.x / G);
    // With one case per workgroup, a factor row's 128-byte line holds 16 cases (FP64), and
    // workgroups are dealt to the 8 XCDs in turn: XCD x takes the consecutive cases
    // [x * n / 8, (x + 1) * n / 8) (workgroups x, x + 8, ...), so the cases of a line share an L2.
    int wg = bx;
    if (G == 1 && (gx & 7) == 0) wg = (bx & 7) * (gx >> 3) + (bx >> 3);
    const int c = wg * G + lane;
    bool act = selected(sel, ctl[c]);
    if (!__syncthreads_or(act)) return;
    // factor and solution rows: the batch's (stride S, column c) or a compact slot (stride 1)
    real* __restrict__ fac = fac_b;
    double* __restrict__ x = x_b;
    int Sf = S, cf = c;
    if constexpr (G == 1 && FACTOR && !SINGLE) {
        if (cmp.cap > 0) {
            __shared__ int s_slot;
            if (threadIdx.x == 0) s_slot = atomicAdd(cmp.ctr, 1);
            __syncthreads();
            const int w = s_slot;
            if (w < cmp.cap) {
                fac = reinterpret_cast<real*>(cmp.fac + (size_t)w * (size_t)cmp.frows);
                x = cmp.x + (size_t)w * (size_t)cmp.xrows;
                Sf = 1;
                cf = 0;
            }
        }
    }

    double pmin = DBL_MAX, pmax = 0.0, psum = 0.0;
    // sub-level A: pivots (y = -src - 2) and U entries (y = Jacobian entry or -1 for fill); in
    // either sub-level y = LU_PART marks a chunk of a cut sum, which stores its sum negated
    auto startA = [&](int, int4 t) -> In1 {
        if (t.y == LU_PART) return {0.0};
        const int src = t.y < -1 ? -t.y - 2 : t.y;
        return {src >= 0 ? Jx[at(src, S, c)] : 0.0};
    };
    auto finishA = [&](int4 t, In1 in, real sum) {
        if (t.y == LU_PART) {
            fac[at(t.x, Sf, cf)] = -sum;
            return;
        }
        real v = (real)in.v + sum;
        if (t.y < -1) {
            const double ab = fabs((double)v);
            pmin = fmin(pmin, ab);
            pmax = fmax(pmax, ab);
            psum += ab;
            v = (real)1 / v;
        }
        fac[at(t.x, Sf, cf)] = v;
    };
    // sub-level B: L entries, scaled by the pivot inverse diagB[q]
    auto startB = [&](int q, int4 t) -> In2 {
        return {t.y >= 0 ? Jx[at(t.y, S, c)] : 0.0, (double)fac[at(diagB[q], Sf, cf)]};
    };
    auto finishB = [&](int4 t, In2 in, real sum) {
        fac[at(t.x, Sf, cf)] = t.y == LU_PART ? -sum : ((real)in.v + sum) * (real)in.d;
    };
    // forward rows (row, right-hand side index): y = rhs - sum L y; a chunk of a cut row
    // (y = LU_PART) stores its sum negated to its solution-vector slot (row index >= m)
    auto startF = [&](int, int4 t) -> In1 { return {t.y == LU_PART ? 0.0 : rhs[at(t.y, S, c)]}; };
    auto finishF = [&](int4 t, In1 in, double sum) { x[at(t.x, Sf, cf)] = t.y == LU_PART ? -sum : in.v + sum; };
    // backward rows (row, pivot position): x = (y - sum U x) / U(p,p); chunks as forward
    auto startW = [&](int, int4 t) -> In2 {
        if (t.y == LU_PART) return {0.0, 0.0};
        return {x[at(t.x, Sf, cf)], (double)fac[at(t.y, Sf, cf)]};
    };
    auto finishW = [&](int4 t, In2 in, double sum) {
        x[at(t.x, Sf, cf)] = t.y == LU_PART ? -sum : (in.v + sum) * in.d;
    };

    for (int l = 0; l < nlev; ++l) {
        const int i0 = l * TL + slot;
        if (act) {
            if (FACTOR)
                seg_stream<K, real, In1>(taskA, ptrA[i0], ptrA[i0 + 1], pairs, fac, fac, Sf, cf, startA, finishA);
            seg_stream<KF, double, In1>(fwd, ptrF[i0], ptrF[i0 + 1], fent, fac, x, Sf, cf, startF, finishF);
        }
        __syncthreads();
        if (FACTOR) {
            if (act) seg_stream<K, real, In2>(taskB, ptrB[i0], ptrB[i0 + 1], pairs, fac, fac, Sf, cf, startB, finishB);
            __syncthreads();
        }
    }
    if (FACTOR) {
        smin[slot][lane] = pmin;
        smax[slot][lane] = pmax;
        ssum[slot][lane] = psum;
        __syncthreads();
        if (slot == 0) {
            int bad = 0;
            if (act) {
                for (int s2 = 1; s2 < TL; ++s2) {
                    pmin = fmin(pmin, smin[s2][lane]);
                    pmax = fmax(pmax, smax[s2][lane]);
                    psum += ssum[s2][lane];
                }
                pstat[2 * (size_t)c] = pmin;
                pstat[2 * (size_t)c + 1] = pmax;
                bad = pivot_verdict<SINGLE>(ctl, c, pmin, pmax, psum, rcond_min, guard_trips, sel) ? 1 : 0;
            }
            sbad[lane] = bad;
        }
        __syncthreads();
        if (sbad[lane]) act = false;
    }
    for (int l = 0; l < nblev; ++l) {
        const int i0 = l * TL + slot;
        if (act) seg_stream<KF, double, In2>(bwd, ptrW[i0], ptrW[i0 + 1], bent, fac, x, Sf, cf, startW, finishW);
        __syncthreads();
    }
    if (x != x_b && act)  // a compact slot: the solution to the batch's rows
        for (int r = (int)threadIdx.x; r < cmp.m; r += (int)blockDim.x) x_b[at(r, S, c)] = x[r];
}

template <int G, int NS, typename real, bool FACTOR>
__global__ __launch_bounds__(64 * NS) void k_lu_stream(
    const int* __restrict__ ptrA, const int4* __restrict__ taskA, const int* __restrict__ ptrB,
    const int4* __restrict__ taskB, const int* __restrict__ diagB, const int2* __restrict__ pairs,
    const int* __restrict__ ptrF, const int4* __restrict__ fwd, const int2* __restrict__ fent,
    const int* __restrict__ ptrW, const int4* __restrict__ bwd, const int2* __restrict__ bent,
    const double* __restrict__ Jx, real* __restrict__ fac, const double* __restrict__ rhs, double* __restrict__ x,
    unsigned* __restrict__ ctl, double* __restrict__ pstat, int* __restrict__ guard_trips, int sel, double rcond_min,
    int nlev, int nblev, int S, LuCompact cmp) {
    lu_stream_body<G, NS, real, FACTOR>(ptrA, taskA, ptrB, taskB, diagB, pairs, ptrF, fwd, fent, ptrW, bwd, bent, Jx,
                                        fac, rhs, x, ctl, pstat, guard_trips, sel, rcond_min, nlev, nblev, S,
                                        blockIdx.x, gridDim.x, cmp);
}

template <int G, int NS, typename real, bool FACTOR>
__global__ __launch_bounds__(64 * NS) void k_lu_stream_multi(const LuArgs* __restrict__ tab,
                                                             const int* __restrict__ off, int nseg) {
    const int s = seg_find(off, nseg, blockIdx.x);
    const LuArgs& a = tab[s];
    const int b = blockIdx.x - off[s];
    if (b >= a.gx || a.plan[PL_HALT] || a.plan[PL_LU] != a.want) return;
    lu_stream_body<G, NS, real, FACTOR>(a.ptrA, a.taskA, a.ptrB, a.taskB, a.diagB, a.pairs, a.ptrF, a.fwd, a.fent,
                                        a.ptrW, a.bwd, a.bent, a.Jx, static_cast<real*>(a.fac), a.rhs, a.x, a.ctl,
                                        a.pstat, a.guard, a.sel, a.rcond_min, a.nlev, a.nblev, a.S, b, a.gx);
}

// ---- 2x2 supernode blocks, G = 64 ----------------------------------------------------
// A block task (3 int4: {kind | nr << 8 | nc << 16, -, sbeg, send}, {tgt0, tgt1, aux0,
// aux1}, {src00, src01, src10, src11}) updates up to four entries (rows R x columns C of
// a pivot group's diagonal, U or L block) from one step stream: a step (int4 of factor
// positions, the zero row where an entry is absent) loads L(R, t) and U(t, C) and makes four
// multiply-adds. A block without a second row (column) has -1 there, skips that load (the
// step is the same in every lane) and ignores the extra sums. The terms of t = p for the second row or
// column of a pivot pair are added in the finish, after the shared ones, so every entry
// sums in ascending pivot order and the factors equal those of k_lu bit for bit.
struct BlkJ { double j[4]; };   // Jacobian start values, loaded one task ahead
struct BlkX { double x[3]; };   // pivot-group values, loaded when the task begins

// Steps per chunk: four row loads per step. The task finish is inlined at every step of a
// chunk, so the chunk length sets the kernel's size: 4 steps keep it at ~22k instructions
// (as k_lu_stream) and ran faster than 8 (fewer instruction-cache misses, 0.57% of fetches
// at 8 steps and three consume sites), and 16 wavefronts fit in 128 VGPRs.
template <typename real>
constexpr int blk_chunk() { return sizeof(real) == 4 ? 8 : 4; }

template <int K, typename real>
__device__ __forceinline__ void load_steps(const int4* __restrict__ st, int e, int eend, const real* __restrict__ fac,
                                           int S, int c, real (&a0)[K], real (&a1)[K], real (&b0)[K], real (&b1)[K]) {
#pragma unroll
    for (int u = 0; u < K; ++u) {
        const int4 p = st[min(e + u, eend - 1)];  // past the end: repeat the last step (unused)
        a0[u] = fac[at(p.x, S, c)];
        b0[u] = fac[at(p.z, S, c)];
        if (p.y >= 0) a1[u] = fac[at(p.y, S, c)];  // the step's positions are uniform
        if (p.w >= 0) b1[u] = fac[at(p.w, S, c)];
    }
}

// Block tasks [q0, q1) of one list (steps back to back); pivot statistics of the diagonal
// blocks go to (pmin, pmax, psum).
template <int K, typename real>
__device__ __forceinline__ void blk_stream(const int4* __restrict__ tasks, int q0, int q1, const int4* __restrict__ st,
                                           const double* __restrict__ Jx, real* __restrict__ fac, int S, int c,
                                           double& pmin, double& pmax, double& psum) {
    if (q0 >= q1) return;
    using ACC = real;
    auto startJ = [&](int q) -> BlkJ {
        const int4 w2 = tasks[3 * q + 2];
        BlkJ in;
        in.j[0] = w2.x >= 0 ? Jx[at(w2.x, S, c)] : 0.0;
        in.j[1] = w2.y >= 0 ? Jx[at(w2.y, S, c)] : 0.0;
        in.j[2] = w2.z >= 0 ? Jx[at(w2.z, S, c)] : 0.0;
        in.j[3] = w2.w >= 0 ? Jx[at(w2.w, S, c)] : 0.0;
        return in;
    };
    auto startX = [&](int q) -> BlkX {
        const int4 w0 = tasks[3 * q], w1 = tasks[3 * q + 1];
        const int kind = w0.x & 255;
        BlkX in;
        in.x[0] = in.x[1] = in.x[2] = 0.0;
        if (kind >= p3smi::BLK_U2 && kind <= p3smi::BLK_L2) in.x[0] = (double)fac[at(w1.z, S, c)];  // L(p+1,p) or 1/U(p,p)
        if (kind == p3smi::BLK_L2) {
            in.x[1] = (double)fac[at(w1.z + 1, S, c)];  // U(p, p+1)
            in.x[2] = (double)fac[at(w1.w, S, c)];      // 1/U(p+1, p+1)
        }
        return in;
    };
    auto pivot = [&](real v) -> real {
        const double ab = fabs((double)v);
        pmin = fmin(pmin, ab);
        pmax = fmax(pmax, ab);
        psum += ab;
        return (real)1 / v;
    };
    auto finish = [&](int q, const BlkJ& in, const BlkX& ix, const ACC (&acc)[4]) {
        const int4 w0 = tasks[3 * q], w1 = tasks[3 * q + 1];
        const int kind = w0.x & 255, nr = (w0.x >> 8) & 255, nc = w0.x >> 16;
        switch (kind) {
            case p3smi::BLK_DIAG1:
                fac[at(w1.x, S, c)] = pivot((real)in.j[0] + acc[0]);
                break;
            case p3smi::BLK_DIAG2: {
                const real inv0 = pivot((real)in.j[0] + acc[0]);
                const real u01 = (real)in.j[1] + acc[1];
                const real l10 = ((real)in.j[2] + acc[2]) * inv0;
                fac[at(w1.x, S, c)] = inv0;
                fac[at(w1.z, S, c)] = u01;
                fac[at(w1.w, S, c)] = l10;
                fac[at(w1.y, S, c)] = pivot((real)in.j[3] + fma(-(ACC)l10, (ACC)u01, acc[3]));
                break;
            }
            case p3smi::BLK_U1:
                fac[at(w1.x, S, c)] = (real)in.j[0] + acc[0];
                if (nc == 2) fac[at(w1.x + 1, S, c)] = (real)in.j[1] + acc[1];
                break;
            case p3smi::BLK_U2: {
                const real l10 = (real)ix.x[0];
                const real u00 = (real)in.j[0] + acc[0];
                fac[at(w1.x, S, c)] = u00;
                fac[at(w1.y, S, c)] = (real)in.j[2] + fma(-(ACC)l10, (ACC)u00, acc[2]);
                if (nc == 2) {
                    const real u01 = (real)in.j[1] + acc[1];
                    fac[at(w1.x + 1, S, c)] = u01;
                    fac[at(w1.y + 1, S, c)] = (real)in.j[3] + fma(-(ACC)l10, (ACC)u01, acc[3]);
                }
                break;
            }
            case p3smi::BLK_L1: {
                const real inv = (real)ix.x[0];
                fac[at(w1.x, S, c)] = ((real)in.j[0] + acc[0]) * inv;
                if (nr == 2) fac[at(w1.x + 1, S, c)] = ((real)in.j[2] + acc[2]) * inv;
                break;
            }
            case p3smi::BLK_PART:  // a chunk of a cut block: its sums, negated, for the next chunk
                fac[at(w1.x, S, c)] = -acc[0];
                if (nc == 2) fac[at(w1.x + 1, S, c)] = -acc[1];
                if (nr == 2) fac[at(w1.x + 2, S, c)] = -acc[2];
                if (nr == 2 && nc == 2) fac[at(w1.x + 3, S, c)] = -acc[3];
                break;
            default: {  // BLK_L2
                const real inv0 = (real)ix.x[0], u01 = (real)ix.x[1], inv1 = (real)ix.x[2];
                const real l00 = ((real)in.j[0] + acc[0]) * inv0;
                fac[at(w1.x, S, c)] = l00;
                fac[at(w1.y, S, c)] = ((real)in.j[1] + fma(-(ACC)l00, (ACC)u01, acc[1])) * inv1;
                if (nr == 2) {
                    const real l10 = ((real)in.j[2] + acc[2]) * inv0;
                    fac[at(w1.x + 1, S, c)] = l10;
                    fac[at(w1.y + 1, S, c)] = ((real)in.j[3] + fma(-(ACC)l10, (ACC)u01, acc[3])) * inv1;
                }
                break;
            }
        }
    };
    int q = q0;
    int4 tk = tasks[3 * q];
    BlkJ in = startJ(q);
    BlkX ix = startX(q);
    BlkJ inn = in;
    if (q + 1 < q1) inn = startJ(q + 1);
    const int eend = tasks[3 * (q1 - 1)].w;
    ACC acc[4] = {0, 0, 0, 0};
    auto advance = [&]() {
        finish(q, in, ix, acc);
        acc[0] = acc[1] = acc[2] = acc[3] = 0;
        ++q;
        if (q < q1) {
            tk = tasks[3 * q];
            in = inn;
            ix = startX(q);
            if (q + 1 < q1) inn = startJ(q + 1);
        }
    };
    auto consume = [&](int e, const real (&a0)[K], const real (&a1)[K], const real (&b0)[K], const real (&b1)[K]) {
#pragma unroll
        for (int u = 0; u < K; ++u) {
            const int pos = e + u;
            while (q < q1 && pos == tk.w) advance();
            if (pos < eend) {
                acc[0] = fma(-(ACC)a0[u], (ACC)b0[u], acc[0]);
                acc[1] = fma(-(ACC)a0[u], (ACC)b1[u], acc[1]);
                acc[2] = fma(-(ACC)a1[u], (ACC)b0[u], acc[2]);
                acc[3] = fma(-(ACC)a1[u], (ACC)b1[u], acc[3]);
            }
        }
    };
    // Two chunk buffers in turn, as in seg_stream, but with two consume sites instead of three
    // (the last chunk is consumed inside the loop; a chunk past the end is not loaded): every
    // site inlines the task finish K times, and the kernel's size decides its instruction-cache
    // misses.
    int e = tk.z;
    if (e < eend) {
        real a0[K], a1[K], b0[K], b1[K], c0[K], c1[K], d0[K], d1[K];
#pragma unroll
        for (int u = 0; u < K; ++u) a1[u] = b1[u] = c1[u] = d1[u] = 0;  // second rows absent: stale, unused
        load_steps<K>(st, e, eend, fac, S, c, a0, a1, b0, b1);
        while (true) {
            if (e + K < eend) load_steps<K>(st, e + K, eend, fac, S, c, c0, c1, d0, d1);
            consume(e, a0, a1, b0, b1);
            e += K;
            if (e >= eend) break;
            if (e + K < eend) load_steps<K>(st, e + K, eend, fac, S, c, a0, a1, b0, b1);
            consume(e, c0, c1, d0, d1);
            e += K;
            if (e >= eend) break;
        }
    }
    while (q < q1) advance();
}

// Forward rows of pivot groups (2 int4 per task: {row, nr, sbeg, send}, {rsrc0, rsrc1,
// lpos, pin}; steps int4 {L(row, j), L(row+1, j) or -1, j, -}): a step loads L(R, j) and y(j)
// once for the group's rows; the pair's term L(row+1, row) y(row) comes last. Chunks of
// K steps in two buffers, as blk_stream. A row cut into pieces: pin >= 0 is the solution-vector
// slot whose nr rows hold the previous piece's sums, the start of this piece's; a piece other
// than the last has nr + 4 in place of nr and its slot in place of row, and stores its sums
// there instead of finishing the row.
template <int K, typename real>
__device__ __forceinline__ void fblk_stream(const int4* __restrict__ tasks, int q0, int q1, const int4* __restrict__ st,
                                            const real* __restrict__ fac, double* __restrict__ x,
                                            const double* __restrict__ rhs, int S, int c) {
    if (q0 >= q1) return;
    struct FIn { double r0, r1, l10, p0, p1; };
    auto start = [&](int q) -> FIn {
        const int4 w0 = tasks[2 * q], w1 = tasks[2 * q + 1];
        const int nr = w0.y > 2 ? w0.y - 4 : w0.y;
        FIn in{0.0, 0.0, 0.0, 0.0, 0.0};
        if (w0.y <= 2) {
            in.r0 = rhs[at(w1.x, S, c)];
            in.r1 = nr == 2 ? rhs[at(w1.y, S, c)] : 0.0;
            in.l10 = nr == 2 ? (double)fac[at(w1.z, S, c)] : 0.0;
        }
        if (w1.w >= 0) {
            in.p0 = x[at(w1.w, S, c)];
            if (nr == 2) in.p1 = x[at(w1.w + 1, S, c)];
        }
        return in;
    };
    auto finish = [&](int q, const FIn& in, double a0, double a1) {
        const int4 w0 = tasks[2 * q];
        if (w0.y > 2) {  // a piece of a cut row: its sums for the next piece
            x[at(w0.x, S, c)] = a0;
            if (w0.y == 6) x[at(w0.x + 1, S, c)] = a1;
            return;
        }
        const double y0 = in.r0 + a0;
        x[at(w0.x, S, c)] = y0;
        if (w0.y == 2) x[at(w0.x + 1, S, c)] = in.r1 + fma(-in.l10, y0, a1);
    };
    int q = q0;
    int4 tk = tasks[2 * q];
    FIn in = start(q);
    FIn inn = in;
    if (q + 1 < q1) inn = start(q + 1);
    const int eend = tasks[2 * (q1 - 1)].w;
    double acc0 = in.p0, acc1 = in.p1;
    auto advance = [&]() {
        finish(q, in, acc0, acc1);
        ++q;
        if (q < q1) {
            tk = tasks[2 * q];
            in = inn;
            acc0 = in.p0;
            acc1 = in.p1;
            if (q + 1 < q1) inn = start(q + 1);
        }
    };
    auto load = [&](int e, real (&a0)[K], real (&a1)[K], double (&y)[K]) {
#pragma unroll
        for (int u = 0; u < K; ++u) {
            const int4 p = st[min(e + u, eend - 1)];
            a0[u] = fac[at(p.x, S, c)];
            if (p.y >= 0) a1[u] = fac[at(p.y, S, c)];
            y[u] = x[at(p.z, S, c)];
        }
    };
    auto consume = [&](int e, const real (&a0)[K], const real (&a1)[K], const double (&y)[K]) {
#pragma unroll
        for (int u = 0; u < K; ++u) {
            const int pos = e + u;
            while (q < q1 && pos == tk.w) advance();
            if (pos < eend) {
                acc0 = fma(-(double)a0[u], y[u], acc0);
                acc1 = fma(-(double)a1[u], y[u], acc1);
            }
        }
    };
    int e = tk.z;
    if (e < eend) {
        real a0[K], a1[K], b0[K], b1[K];
        double y0[K], y1[K];
#pragma unroll
        for (int u = 0; u < K; ++u) a1[u] = b1[u] = 0;  // second rows absent: stale, unused
        load(e, a0, a1, y0);
        while (true) {
            if (e + K < eend) load(e + K, b0, b1, y1);
            consume(e, a0, a1, y0);
            e += K;
            if (e >= eend) break;
            if (e + K < eend) load(e + K, a0, a1, y0);
            consume(e, b0, b1, y1);
            e += K;
            if (e >= eend) break;
        }
    }
    while (q < q1) advance();
}

// Backward rows of pivot groups (2 int4 per task: {row, nr, sbeg, send}, {1/U(row,row),
// 1/U(row+1,row+1), pin, -} positions; steps int4 {U(row, j), U(row+1, j) or -1, j, -} over
// the columns of the group's last row, descending): a step loads x(j) once for the group's
// rows; for a pair, x(row+1) first, then U(row, row+1) x(row+1) last in row's sum. Cut rows
// as fblk_stream (pin in the third field of the second int4).
template <int K, typename real>
__device__ __forceinline__ void bblk_stream(const int4* __restrict__ tasks, int q0, int q1, const int4* __restrict__ st,
                                            const real* __restrict__ fac, double* __restrict__ x, int S, int c) {
    if (q0 >= q1) return;
    struct BIn { double y0, y1, inv0, inv1, u01, p0, p1; };
    auto start = [&](int q) -> BIn {
        const int4 w0 = tasks[2 * q], w1 = tasks[2 * q + 1];
        const int nr = w0.y > 2 ? w0.y - 4 : w0.y;
        BIn in{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (w0.y <= 2) {
            in.y0 = x[at(w0.x, S, c)];
            in.inv0 = (double)fac[at(w1.x, S, c)];
            if (nr == 2) {
                in.y1 = x[at(w0.x + 1, S, c)];
                in.inv1 = (double)fac[at(w1.y, S, c)];
                in.u01 = (double)fac[at(w1.x + 1, S, c)];
            }
        }
        if (w1.z >= 0) {
            in.p0 = x[at(w1.z, S, c)];
            if (nr == 2) in.p1 = x[at(w1.z + 1, S, c)];
        }
        return in;
    };
    auto finish = [&](int q, const BIn& in, double a0, double a1) {
        const int4 w0 = tasks[2 * q];
        if (w0.y > 2) {  // a piece of a cut row: its sums for the next piece
            x[at(w0.x, S, c)] = a0;
            if (w0.y == 6) x[at(w0.x + 1, S, c)] = a1;
            return;
        }
        if (w0.y == 2) {
            const double x1 = (in.y1 + a1) * in.inv1;
            x[at(w0.x + 1, S, c)] = x1;
            x[at(w0.x, S, c)] = (in.y0 + fma(-in.u01, x1, a0)) * in.inv0;
        } else {
            x[at(w0.x, S, c)] = (in.y0 + a0) * in.inv0;
        }
    };
    int q = q0;
    int4 tk = tasks[2 * q];
    BIn in = start(q);
    BIn inn = in;
    if (q + 1 < q1) inn = start(q + 1);
    const int eend = tasks[2 * (q1 - 1)].w;
    double acc0 = in.p0, acc1 = in.p1;
    auto advance = [&]() {
        finish(q, in, acc0, acc1);
        ++q;
        if (q < q1) {
            tk = tasks[2 * q];
            in = inn;
            acc0 = in.p0;
            acc1 = in.p1;
            if (q + 1 < q1) inn = start(q + 1);
        }
    };
    auto load = [&](int e, real (&a0)[K], real (&a1)[K], double (&xv)[K]) {
#pragma unroll
        for (int u = 0; u < K; ++u) {
            const int4 p = st[min(e + u, eend - 1)];
            a0[u] = fac[at(p.x, S, c)];
            if (p.y >= 0) a1[u] = fac[at(p.y, S, c)];
            xv[u] = x[at(p.z, S, c)];
        }
    };
    auto consume = [&](int e, const real (&a0)[K], const real (&a1)[K], const double (&xv)[K]) {
#pragma unroll
        for (int u = 0; u < K; ++u) {
            const int pos = e + u;
            while (q < q1 && pos == tk.w) advance();
            if (pos < eend) {
                acc0 = fma(-(double)a0[u], xv[u], acc0);
                acc1 = fma(-(double)a1[u], xv[u], acc1);
            }
        }
    };
    int e = tk.z;
    if (e < eend) {
        real a0[K], a1[K], b0[K], b1[K];
        double x0[K], x1[K];
#pragma unroll
        for (int u = 0; u < K; ++u) a1[u] = b1[u] = 0;  // second rows absent: stale, unused
        load(e, a0, a1, x0);
        while (true) {
            if (e + K < eend) load(e + K, b0, b1, x1);
            consume(e, a0, a1, x0);
            e += K;
            if (e >= eend) break;
            if (e + K < eend) load(e + K, a0, a1, x0);
            consume(e, b0, b1, x1);
            e += K;
            if (e >= eend) break;
        }
    }
    while (q < q1) advance();
}

// k_lu_stream with the 2x2-block schedule (G = 64): same contract and, bit for bit, the
// same factors and solutions. Per level: sub-level A = diagonal blocks and U blocks of
// single pivots; sub-level B = the U blocks of pivot pairs, the L blocks and the forward
// rows by pivot group (which read L(p+1, p), written by the diagonal block of a pair in
// sub-level A).
template <int NS, typename real, bool FACTOR>
__device__ __forceinline__ void lu_blk_body(
    const int* __restrict__ ptrA, const int4* __restrict__ blkA, const int* __restrict__ ptrB,
    const int4* __restrict__ blkB, const int4* __restrict__ steps, const int* __restrict__ ptrF,
    const int4* __restrict__ fwd, const int4* __restrict__ fsteps, const int* __restrict__ ptrW,
    const int4* __restrict__ bwd, const int4* __restrict__ bsteps, const double* __restrict__ Jx,
    real* __restrict__ fac, const double* __restrict__ rhs, double* __restrict__ x, unsigned* __restrict__ ctl,
    double* __restrict__ pstat, int* __restrict__ guard_trips, int sel, double rcond_min, int nlev, int nblev, int S,
    int bx, int cpg) {
    constexpr bool SINGLE = sizeof(real) == 4;
    constexpr int G = 64, TL = NS;
    __shared__ double smin[TL][G], smax[TL][G], ssum[TL][G];
    __shared__ int sbad[G];
    const int lane = threadIdx.x % G;
    const int slot = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / G));
    // cases bx * cpg + [0, cpg): cpg of the 64 lanes carry one case each
    const int c = bx * cpg + lane;
    bool act = lane < cpg && c < S && selected(sel, ctl[c]);
    if (!__syncthreads_or(act)) return;

    // pivot statistics live in LDS (a diagonal block updates them; registers are scarce here)
    double& pmin = smin[slot][lane];
    double& pmax = smax[slot][lane];
    double& psum = ssum[slot][lane];
    pmin = DBL_MAX;
    pmax = 0.0;
    psum = 0.0;

    // The FP64 factorization also runs on the lanes of cases outside STEP, whose factors are
    // dead (a case refactors before it solves again; after a launch only cases in STEP have
    // their factors read), so that its row loads and stores cover the group's whole rows. With
    // part of the lanes off, a factorization ran 4.1 ms against 3.6 ms with all on (2,000-bus
    // grid, 5,702 of 8,760 cases selected). Their solutions and pivot statistics are not touched.
    const bool actf = act || (FACTOR && !SINGLE && (sel == SEL_ALL || sel == SEL_F64) && lane < cpg && c < S &&
                              stage_of(ctl[c]) != SG_STEP);
    for (int l = 0; l < nlev; ++l) {
        const int i0 = l * TL + slot;
        if (FACTOR) {
            if (actf) blk_stream<blk_chunk<real>(), real>(blkA, ptrA[i0], ptrA[i0 + 1], steps, Jx, fac, S, c, pmin, pmax, psum);
            __syncthreads();
        }
        if (FACTOR && actf)
            blk_stream<blk_chunk<real>(), real>(blkB, ptrB[i0], ptrB[i0 + 1], steps, Jx, fac, S, c, pmin, pmax, psum);
        if (act) fblk_stream<4, real>(fwd, ptrF[i0], ptrF[i0 + 1], fsteps, fac, x, rhs, S, c);
        __syncthreads();
    }
    if (FACTOR) {
        __syncthreads();
        if (slot == 0) {
            int bad = 0;
            if (act) {
                double lo = smin[0][lane], hi = smax[0][lane], sum = ssum[0][lane];
                for (int s2 = 1; s2 < TL; ++s2) {
                    lo = fmin(lo, smin[s2][lane]);
                    hi = fmax(hi, smax[s2][lane]);
                    sum += ssum[s2][lane];
                }
                pstat[2 * (size_t)c] = lo;
                pstat[2 * (size_t)c + 1] = hi;
                bad = pivot_verdict<SINGLE>(ctl, c, lo, hi, sum, rcond_min, guard_trips, sel) ? 1 : 0;
            }
            sbad[lane] = bad;
        }
        __syncthreads();
        if (sbad[lane]) act = false;
    }
    for (int l = 0; l < nblev; ++l) {
        const int i0 = l * TL + slot;
        if (act) bblk_stream<4, real>(bwd, ptrW[i0], ptrW[i0 + 1], bsteps, fac, x, S, c);
        __syncthreads();
    }
}

template <int NS, typename real, bool FACTOR>
__global__ __launch_bounds__(64 * NS) void k_lu_blk(
    const int* __restrict__ ptrA, const int4* __restrict__ blkA, const int* __restrict__ ptrB,
    const int4* __restrict__ blkB, const int4* __restrict__ steps, const int* __restrict__ ptrF,
    const int4* __restrict__ fwd, const int4* __restrict__ fsteps, const int* __restrict__ ptrW,
    const int4* __restrict__ bwd, const int4* __restrict__ bsteps, const double* __restrict__ Jx,
    real* __restrict__ fac, const double* __restrict__ rhs, double* __restrict__ x, unsigned* __restrict__ ctl,
    double* __restrict__ pstat, int* __restrict__ guard_trips, int sel, double rcond_min, int nlev, int nblev, int S,
    int cpg) {
    lu_blk_body<NS, real, FACTOR>(ptrA, blkA, ptrB, blkB, steps, ptrF, fwd, fsteps, ptrW, bwd, bsteps, Jx, fac, rhs, x,
                                  ctl, pstat, guard_trips, sel, rcond_min, nlev, nblev, S, blockIdx.x, cpg);
}

template <int NS, typename real, bool FACTOR>
__global__ __launch_bounds__(64 * NS) void k_lu_blk_multi(const LuArgs* __restrict__ tab, const int* __restrict__ off,
                                                          int nseg) {
    const int s = seg_find(off, nseg, blockIdx.x);
    const LuArgs& a = tab[s];
    const int b = blockIdx.x - off[s];
    if (b >= a.gx || a.plan[PL_HALT] || a.plan[PL_LU] != a.want) return;
    lu_blk_body<NS, real, FACTOR>(a.ptrA, a.taskA, a.ptrB, a.taskB, a.steps, a.ptrF, a.fwd, a.fsteps, a.ptrW, a.bwd,
                                  a.bsteps, a.Jx, static_cast<real*>(a.fac), a.rhs, a.x, a.ctl, a.pstat, a.guard,
                                  a.sel, a.rcond_min, a.nlev, a.nblev, a.S, b, 64);
}

// Refinement residual r = R - J x in FP64 (J in equation-row CSR order, x in pivot order:
// xcol[e] = pivot index of the column of entry e), one wavefront per row and 64 cases.
__global__ __launch_bounds__(256) void k_resid(const int* __restrict__ Jp, const int* __restrict__ xcol,
                                               const double* __restrict__ Jx, const double* __restrict__ R,
                                               const double* __restrict__ x, double* __restrict__ r,
                                               const unsigned* __restrict__ ctl, int m, int S) {
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int row = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (row >= m || !selected(SEL_REFINE, ctl[c])) return;
    double acc = R[at(row, S, c)], acc1 = 0.0;
    const int e1 = Jp[row + 1];
    int e = Jp[row];
    for (; e + 2 <= e1; e += 2) {
        acc = fma(-Jx[at(e, S, c)], x[at(xcol[e], S, c)], acc);
        acc1 = fma(-Jx[at(e + 1, S, c)], x[at(xcol[e + 1], S, c)], acc1);
    }
    if (e < e1) acc = fma(-Jx[at(e, S, c)], x[at(xcol[e], S, c)], acc);
    r[at(row, S, c)] = acc + acc1;
}

// x += d for the refining cases; per workgroup (64 pivots x 64 cases) the largest |d| and
// |x| of each case go to part[2 * (blockIdx.x * S + c) + {0, 1}].
__global__ __launch_bounds__(256) void k_refine_update(double* __restrict__ x, const double* __restrict__ d,
                                                       double* __restrict__ part, const unsigned* __restrict__ ctl,
                                                       int m, int S) {
    __shared__ double sd[4][64], sx[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = blockIdx.y * 64 + lane;
    const bool on = selected(SEL_REFINE, ctl[c]);
    double dm = 0.0, xm = 0.0;
    if (on) {
        const int p1 = min(m, (int)(blockIdx.x + 1) * 64);
        for (int p = blockIdx.x * 64 + wave; p < p1; p += 4) {
            const size_t o = at(p, S, c);
            const double dv = d[o], xv = x[o] + dv;
            x[o] = xv;
            dm = nan_max(dm, fabs(dv));
            xm = nan_max(xm, fabs(xv));
        }
    }
    sd[wave][lane] = dm;
    sx[wave][lane] = xm;
    __syncthreads();
    if (wave == 0 && on) {
        for (int v = 1; v < 4; ++v) {
            dm = nan_max(dm, sd[v][lane]);
            xm = nan_max(xm, sx[v][lane]);
        }
        part[2 * at(blockIdx.x, S, c)] = dm;
        part[2 * at(blockIdx.x, S, c) + 1] = xm;
    }
}

// Per refining case: stop when the correction |d| <= rtol |x|; fall back to FP64 for good when a
// step did not at least halve |d| (refinement stalls: cond(J) times the FP32 rounding is too big)
// or when the last allowed step still misses rtol. counts[0] += cases still refining, counts[1]
// += fallbacks. Every step is refined to rtol, also far from the solution: on ill-conditioned
// cases whose line search cuts steps, an inexact early step shifts the converged voltages by up
// to |J^-1| tol (measured: 4e-9 with an inexact-Newton forcing term of 1e-4 ||F||).
__global__ void k_refine_decide(const double* __restrict__ part, int nparts, double* __restrict__ prevd,
                                unsigned* __restrict__ ctl, int* __restrict__ counts, double rtol, int last, int S) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned w = c < S ? ctl[c] : SG_IDLE;
    if (selected(SEL_REFINE, w)) {
        double dm = 0.0, xm = 0.0;
        for (int b = 0; b < nparts; ++b) {
            dm = nan_max(dm, part[2 * at(b, S, c)]);
            xm = nan_max(xm, part[2 * at(b, S, c) + 1]);
        }
        if (dm <= rtol * xm) {
            ctl[c] = w & ~CTL_REFINING;
        } else if (!(dm <= 0.5 * prevd[c]) || last) {
            ctl[c] = set_prec(w & ~CTL_REFINING, PR_F64_ALWAYS);
            atomicAdd(&counts[1], 1);
        } else {
            prevd[c] = dm;
            atomicAdd(&cnt, 1);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && cnt) atomicAdd(&counts[0], cnt);
}

// Start of a mixed-precision iteration: the FP32 cases in STEP refine, no previous correction;
// counts[2] += cases in STEP of an FP64 class (factored in FP64 this iteration whatever happens).
__global__ __launch_bounds__(256) void k_refine_init(unsigned* __restrict__ ctl, double* __restrict__ prevd,
                                                     int* __restrict__ counts, int S) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned w = c < S ? ctl[c] : SG_IDLE;
    const bool step = frozen_step(w);
    const bool on = step && prec_of(w) == PR_F32;
    const int n64 = __popcll(__ballot(step && !on));
    if ((threadIdx.x & 63) == 0 && n64) atomicAdd(&counts[2], n64);
    if (c >= S) return;
    ctl[c] = on ? (w | CTL_REFINING) : (w & ~CTL_REFINING);
    prevd[c] = DBL_MAX;
}

template <typename T>
__global__ void k_zero(T* p, size_t n) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = (T)0;
}

template <typename T>
__global__ void k_one(T* p, size_t n) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = (T)1;
}

// One case's column of an interleaved array and a dense vector: put (dense -> column c of the
// [item * S + c] layout) or take (column -> dense), n items.
__global__ void k_column(const double* __restrict__ src, double* __restrict__ dst, int n, int S, int c, int take) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (take) dst[i] = src[(size_t)i * S + c];
    else dst[(size_t)i * S + c] = src[i];
}

// Newton move of the cases this pass advances (p3s/cpp/nr_klu.cpp:665-700): STEP takes the full
// step (alpha 1) from the current point and saves that point as the step's base; TRIAL and
// COMMIT move to base + alpha dx. The solve's x is in pivot order: x[xva[bus]] the angle step,
// x[xvm[bus]] the relative magnitude step dVm / Vm (magnitude columns are in ln |V|), so
// Vm = Vm0 (1 + alpha x). One thread per (bus, case); the slack has neither and stays. Workgroup
// (bx, by) of gy along the buses: the 64 cases of group bx, buses by * 4 + wave stepping by gy * 4.
struct AdvanceArgs {
    const double* x;
    const int *xva, *xvm;
    const unsigned* ctl;
    const double* alpha;
    double *Vm, *Va, *Ve, *Vf, *Vm0, *Va0;
    int hold, n, S;
    int gx, gy;
    const int* plan;  // passes driven on the device: the hold and whether to run come from it
};

__device__ __forceinline__ void advance_body(const double* __restrict__ x, const int* __restrict__ xva,
                                             const int* __restrict__ xvm, const unsigned* __restrict__ ctl,
                                             const double* __restrict__ alpha, double* __restrict__ Vm,
                                             double* __restrict__ Va, double* __restrict__ Ve,
                                             double* __restrict__ Vf, double* __restrict__ Vm0,
                                             double* __restrict__ Va0, int hold, int n, int S, int bx, int by,
                                             int gy) {
    const int c = bx * 64 + (threadIdx.x & 63);
    const unsigned w = ctl[c];
    const unsigned sg = stage_of(w);
    if (sg == SG_START || !moves(w, hold)) return;
    const double al = sg == SG_STEP ? 1.0 : alpha[c];
    for (int bus = by * 4 + (int)(threadIdx.x >> 6); bus < n; bus += gy * 4) {
        const int iva = xva[bus], ivm = xvm[bus];
        if (iva < 0 && ivm < 0) continue;
        const size_t o = at(bus, S, c);
        double va, vm, step;
        if (sg == SG_STEP) {
            va = Va[o];
            vm = Vm[o];
            Va0[o] = va;
            Vm0[o] = vm;
            step = 1.0;
        } else {
            va = Va0[o];
            vm = Vm0[o];
            step = al;
        }
        if (iva >= 0) va += step * x[at(iva, S, c)];
        if (ivm >= 0) vm = fma(vm, step * x[at(ivm, S, c)], vm);
        double s, co;
        sincos(va, &s, &co);
        Va[o] = va;
        Vm[o] = vm;
        Ve[o] = vm * co;
        Vf[o] = vm * s;
    }
}

__global__ __launch_bounds__(256) void k_advance(AdvanceArgs a) {
    advance_body(a.x, a.xva, a.xvm, a.ctl, a.alpha, a.Vm, a.Va, a.Ve, a.Vf, a.Vm0, a.Va0, a.hold, a.n, a.S, blockIdx.x,
                 blockIdx.y, gridDim.y);
}

__global__ __launch_bounds__(256) void k_advance_multi(const AdvanceArgs* __restrict__ tab,
                                                       const int* __restrict__ off, int nseg) {
    const int s = seg_find(off, nseg, blockIdx.x);
    const AdvanceArgs& a = tab[s];
    const int b = blockIdx.x - off[s];
    if (b >= a.gx * a.gy || a.plan[PL_HALT] || !a.plan[PL_ADVANCE]) return;
    advance_body(a.x, a.xva, a.xvm, a.ctl, a.alpha, a.Vm, a.Va, a.Ve, a.Vf, a.Vm0, a.Va0, a.plan[PL_HOLD], a.n, a.S,
                 b % a.gx, b / a.gx, a.gy);
}

// Precision vote of the switching mixed mode after an accepted step (the reference GPU
// solver's rule, p3s/cuda/nr_polar_kernels.cu:209-275): FP32 while the mismatch is above `sw`,
// FP64 below it, FP64 for good once an FP32 step failed to halve the mismatch. (A cut step
// sets FP64 for good where the line search starts.)
__device__ __forceinline__ unsigned switch_vote(unsigned p, double before, double after, double sw) {
    if (p == PR_F64_ALWAYS) return p;
    if (p == PR_F32 && !(after <= 0.5 * before)) return PR_F64_ALWAYS;
    return after > sw ? PR_F32 : PR_F64;
}

// Newton bookkeeping after k_evaluate, one thread per case, by the stage the case was evaluated
// in (p3s/cpp/nr_klu.cpp:590-709):
//   START   the start point: converged if ||R|| < tol, failed if non-finite, else iterating
//   STEP    the full step of iteration it + 1: accepted if ||R|| <= (1 - c) ||R_0|| (always
//           without line search), else the line search starts at alpha 1/2
//   TRIAL   accepted if ||R|| <= (1 - c alpha) ||R_0||, else alpha halves; after 10 trials the
//           best one is taken
//   COMMIT  the Jacobian rebuilt at the accepted point: converged, failed or iterating
// nr_klu leaves its loop after the step of iteration max_iter without a further convergence
// test, so such a case ends STOP. A non-finite norm counts as DBL_MAX, so no comparison
// against it passes. A case that ends STOP or FAIL on its first run starts again, factoring with
// pivoting (rerun_word); the result of that run is the case's. The stage counters of the launch are summed over its workgroups; the last
// workgroup to finish copies them to counts_out (what the host reads) and clears the sums for
// the next launch, so no clearing kernel or memset is needed.
struct CtlArgs {
    const double* part;  // (nparts, S) largest |R| per row tile
    int nparts;
    unsigned* ctl;
    double* nrm;         // mismatch norm at the accepted point
    double* alpha;
    double* best_alpha;
    double* best_nrm;
    int* tally;          // NCOUNT sums + 1 finish counter, zero between launches
    int* counts_out;     // NCOUNT
    int* reruns;         // += cases restarted with pivoting
    double tol, sw;      // sw > 0: switching mixed mode
    int line_search, max_iter, S;
    int hold;            // 1: the cases in STEP waited this pass
    int gx;              // workgroups (S / 256 rounded up)
    int few_max;         // LU: the few-case shape for at most this many cases in STEP
    int* plan;           // passes driven on the device: the hold comes from it, and the last
                         // workgroup writes the plan of the next pass there
};

__device__ __forceinline__ void control_body(const CtlArgs& a, int hold, int bx, int gx) {
    __shared__ int s_cnt[NCOUNT];
    __shared__ int s_last;
    if (threadIdx.x < NCOUNT) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const int c = bx * blockDim.x + threadIdx.x;
    if (c < a.S) {
        unsigned w = a.ctl[c];
        unsigned sg = stage_of(w);
        if (moves(w, hold)) {
            double nw = 0.0;
            for (int t = 0; t < a.nparts; ++t) nw = nan_max(nw, a.part[at(t, a.S, c)]);
            if (!(nw < DBL_MAX)) nw = DBL_MAX;
            int it = iters_of(w);
            unsigned prec = prec_of(w);
            const double before = a.nrm[c];
            auto settle = [&](double v, bool stepped) -> unsigned {
                if (stepped && it >= a.max_iter) return SG_STOP;
                if (v < a.tol) return SG_CONV;
                return v >= DBL_MAX ? SG_FAIL : SG_STEP;
            };
            if (sg == SG_START) {
                it = 0;
                a.nrm[c] = nw;
                sg = settle(nw, false);
                if (a.sw > 0.0) prec = nw > a.sw ? PR_F32 : PR_F64;
            } else if (sg == SG_STEP) {
                it += 1;
                if (!a.line_search || nw <= (1.0 - ARMIJO_C) * before) {
                    if (a.sw > 0.0) prec = switch_vote(prec, before, nw, a.sw);
                    a.nrm[c] = nw;
                    sg = settle(nw, true);
                } else {
                    a.best_alpha[c] = 1.0;
                    a.best_nrm[c] = nw;
                    a.alpha[c] = 0.5;
                    w = set_trials(w, 1);
                    sg = SG_TRIAL;
                    if (a.sw > 0.0) prec = PR_F64_ALWAYS;
                }
            } else if (sg == SG_TRIAL) {
                const double al = a.alpha[c];
                if (nw < a.best_nrm[c]) {
                    a.best_nrm[c] = nw;
                    a.best_alpha[c] = al;
                }
                const unsigned tr = trials_of(w);
                if (nw <= (1.0 - ARMIJO_C * al) * before) {
                    sg = SG_COMMIT;
                } else if ((int)tr >= LS_MAX_TRIALS) {
                    a.alpha[c] = a.best_alpha[c];
                    sg = SG_COMMIT;
                } else {
                    w = set_trials(w, tr + 1);
                    a.alpha[c] = 0.5 * al;
                }
            } else {  // SG_COMMIT
                a.nrm[c] = nw;
                sg = settle(nw, true);
                w = set_trials(w, 0);
            }
            w = set_iters(set_prec(set_stage(w, sg), prec), it);
            if ((sg == SG_STOP || sg == SG_FAIL) && !(w & CTL_RERUN)) {
                w = rerun_word(w);
                sg = SG_START;
                atomicAdd(a.reruns, 1);
            }
            a.ctl[c] = w;
        }
        atomicAdd(&s_cnt[counter_of(sg)], 1);
        if (pivoting_step(w)) atomicAdd(&s_cnt[CNT_PIVOT], 1);
    }
    __syncthreads();
    if (threadIdx.x < NCOUNT && s_cnt[threadIdx.x]) atomicAdd(&a.tally[threadIdx.x], s_cnt[threadIdx.x]);
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) s_last = atomicAdd(&a.tally[NCOUNT], 1) == gx - 1;
    __syncthreads();
    if (s_last) {
        __threadfence();
        if (threadIdx.x < NCOUNT) {
            const int v = atomicExch(&a.tally[threadIdx.x], 0);
            a.counts_out[threadIdx.x] = v;
            s_cnt[threadIdx.x] = v;
        }
        if (threadIdx.x == 0) atomicExch(&a.tally[NCOUNT], 0);
        __syncthreads();
        if (a.plan && threadIdx.x == 0) {
            int pl[NPLAN];
            plan_from_counts(s_cnt, a.few_max, pl);
            for (int k = 0; k < NPLAN; ++k) a.plan[k] = pl[k];
        }
    }
}

__global__ __launch_bounds__(256) void k_control(CtlArgs a) { control_body(a, a.hold, blockIdx.x, gridDim.x); }

__global__ __launch_bounds__(256) void k_control_multi(const CtlArgs* __restrict__ tab, const int* __restrict__ off,
                                                       int nseg) {
    const int s = seg_find(off, nseg, blockIdx.x);
    const CtlArgs& a = tab[s];
    const int b = blockIdx.x - off[s];
    if (b >= a.gx || a.plan[PL_HALT]) return;
    control_body(a, a.plan[PL_HOLD], b, a.gx);
}

// Complex V (n, B) of the cases' state, the layout of the numpy result; NaN at buses a served
// mask marks unserved (p3s/contingency/solver_cpp.py:147-150).
__global__ __launch_bounds__(256) void k_emit(const double* __restrict__ Ve, const double* __restrict__ Vf,
                                              const unsigned char* __restrict__ served, double* __restrict__ V, int n,
                                              int S, int B) {
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int bus = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (bus >= n || c >= B) return;
    const size_t o = at(bus, S, c);
    double re = Ve[o], im = Vf[o];
    if (served && !served[o]) re = im = __longlong_as_double(0x7ff8000000000000ll);
    V[2 * ((size_t)bus * B + c)] = re;
    V[2 * ((size_t)bus * B + c) + 1] = im;
}


// ===================================================================== pivoting factorization
// The rescue of the frozen plan: a left-looking sparse LU with threshold partial pivoting
// (Gilbert-Peierls: per column a depth-first reach through the columns of L gives the nonzero
// rows in topological order, a sparse triangular solve, then the pivot), standing in for KLU's
// klu_factor / klu_refactor / klu_solve at p3s/cpp/nr_klu.cpp:609-652. Rows are scaled by their
// largest |a_ij| (KLU's default); the column order is the plan's minimum-degree order; a column's
// pivot is its own row when that row is still free and holds at least PIVOT_TOL of the column's
// largest candidate, else the largest (the first in reach order among equals). A refactorization
// reuses the pivot order, row scales and pattern of the case's last factorization.
//
// One wavefront per case, each in a persistent slot of a pool (factors, pattern, work vectors,
// slot-major). Lane 0 walks the depth-first search; the 64 lanes split every column update, the
// pivot search and the store of an L column. The work vector w is zero between columns.
constexpr double PIVOT_TOL = 1e-3;

struct PivArgs {
    const int* Cp;  // J by column (reference order): rows Ci, CSR entries Ce
    const int* Ci;
    const int* Ce;
    const int* Jp;  // J rows (row scales)
    const int* Q;   // column order
    int m, S;
    const double* J;  // [e * S + c]
    const double* R;  // right-hand side, Jacobian row order [row * S + c]
    double* x;        // solution, pivot order [k * S + c]
    unsigned* ctl;
    const int* list;  // the cases of this launch
    const int* slot_of;
    int* res;         // per list entry: 0 solved, 1 singular, 2 factor capacity exceeded
    int* guard;       // [3] reruns, [4] pivoting factorizations
    long long cap;    // L and U entries per slot
    int *Lp, *Li, *Up, *Ui, *P, *pinv, *xi, *stk, *pstk, *flag;
    double *Lx, *Ux, *Ud, *rs, *w, *y;
};

struct PivSlot {
    int *Lp, *Li, *Up, *Ui, *P, *pinv, *xi, *stk, *pstk, *flag;
    double *Lx, *Ux, *Ud, *rs, *w, *y;
};

__device__ __forceinline__ PivSlot piv_slot(const PivArgs& a, int s) {
    const size_t m = (size_t)a.m, m1 = m + 1, cp = (size_t)a.cap;
    return PivSlot{a.Lp + s * m1, a.Li + s * cp, a.Up + s * m1, a.Ui + s * cp, a.P + s * m,
                   a.pinv + s * m,  a.xi + s * m, a.stk + s * m, a.pstk + s * m, a.flag + s * m,
                   a.Lx + s * cp,  a.Ux + s * cp, a.Ud + s * m,  a.rs + s * m,   a.w + s * m,
                   a.y + s * m};
}

__device__ __forceinline__ void wave_sync() {
    __threadfence_block();
    __syncthreads();
}

// rows reachable from column col of A through the columns of L (lane 0): xi[top, m) in
// topological order; flag[i] == k marks them
__device__ int piv_reach(const PivArgs& a, const PivSlot& f, int col, int k) {
    int top = a.m;
    for (int p = a.Cp[col]; p < a.Cp[col + 1]; ++p) {
        const int i0 = a.Ci[p];
        if (f.flag[i0] == k) continue;
        int head = 0;
        f.stk[0] = i0;
        while (head >= 0) {
            const int i = f.stk[head];
            const int j = f.pinv[i];
            if (f.flag[i] != k) {
                f.flag[i] = k;
                f.pstk[head] = j < 0 ? 0 : f.Lp[j];
            }
            bool done = true;
            if (j >= 0) {
                const int pend = f.Lp[j + 1];
                for (int q = f.pstk[head]; q < pend; ++q) {
                    const int r = f.Li[q];
                    if (f.flag[r] == k) continue;
                    f.pstk[head] = q + 1;
                    f.stk[++head] = r;
                    done = false;
                    break;
                }
            }
            if (done) {
                --head;
                f.xi[--top] = i;
            }
        }
    }
    return top;
}

// w[i] -= L(i, j) xj over column j of L, split over the lanes
__device__ __forceinline__ void piv_axpy(const PivSlot& f, int j, double xj, int lane) {
    for (int q = f.Lp[j] + lane; q < f.Lp[j + 1]; q += 64) f.w[f.Li[q]] -= f.Lx[q] * xj;
}

// A(:, col) / row scale into w
__device__ __forceinline__ void piv_scatter(const PivArgs& a, const PivSlot& f, int c, int col, int lane) {
    for (int p = a.Cp[col] + lane; p < a.Cp[col + 1]; p += 64) {
        const int i = a.Ci[p];
        f.w[i] += a.J[at(a.Ce[p], a.S, c)] / f.rs[i];
    }
}

// Fresh factorization. 0 = done, 1 = structurally or numerically singular, 2 = capacity.
__device__ int piv_factor(const PivArgs& a, const PivSlot& f, int c, int lane) {
    __shared__ int s_top;
    const int m = a.m;
    for (int i = lane; i < m; i += 64) {
        double r = 0.0;
        for (int e = a.Jp[i]; e < a.Jp[i + 1]; ++e) r = fmax(r, fabs(a.J[at(e, a.S, c)]));
        f.rs[i] = (r > 0.0 && r <= DBL_MAX) ? r : 1.0;
        f.pinv[i] = -1;
        f.flag[i] = -1;
        f.w[i] = 0.0;
    }
    if (lane == 0) f.Lp[0] = f.Up[0] = 0;
    wave_sync();
    long long lnz = 0, unz = 0;
    for (int k = 0; k < m; ++k) {
        const int col = a.Q[k];
        if (lane == 0) s_top = piv_reach(a, f, col, k);
        wave_sync();
        const int top = s_top;
        if (unz + (m - top) > a.cap || lnz + (m - top) > a.cap) return 2;
        piv_scatter(a, f, c, col, lane);
        wave_sync();
        for (int q = top; q < m; ++q) {
            const int i = f.xi[q];
            const int j = f.pinv[i];
            if (j < 0) continue;
            const double xj = f.w[i];
            if (lane == 0) {
                f.Ui[unz] = j;
                f.Ux[unz] = xj;
            }
            ++unz;
            piv_axpy(f, j, xj, lane);
            wave_sync();
        }
        if (lane == 0) f.Up[k + 1] = (int)unz;
        // pivot search over the free rows of the reach: the largest |w|, the first in reach order
        double amax = -1.0;
        int piv = -1, pq = INT_MAX;
        for (int q = top + lane; q < m; q += 64) {
            const int i = f.xi[q];
            if (f.pinv[i] >= 0) continue;
            const double v = fabs(f.w[i]);
            if (v > amax) {
                amax = v;
                piv = i;
                pq = q;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double oa = __shfl_xor(amax, off);
            const int op = __shfl_xor(piv, off), oq = __shfl_xor(pq, off);
            if (oa > amax || (oa == amax && oq < pq)) {
                amax = oa;
                piv = op;
                pq = oq;
            }
        }
        if (f.pinv[col] < 0 && f.flag[col] == k && fabs(f.w[col]) >= PIVOT_TOL * amax) piv = col;
        if (piv < 0 || !(amax > 0.0) || !(amax <= DBL_MAX)) return 1;
        const double d = f.w[piv];
        wave_sync();
        if (lane == 0) {
            f.Ud[k] = d;
            f.P[k] = piv;
            f.pinv[piv] = k;
        }
        wave_sync();
        // column k of L: the free rows of the reach in reach order, divided by the pivot
        for (int base = top; base < m; base += 64) {
            const int q = base + lane;
            bool take = false;
            int i = 0;
            double v = 0.0;
            if (q < m) {
                i = f.xi[q];
                if (f.pinv[i] < 0) {
                    take = true;
                    v = f.w[i] / d;
                }
                f.w[i] = 0.0;
            }
            const unsigned long long bal = __ballot(take);
            const int before = __popcll(bal & ((1ull << lane) - 1ull));
            if (take) {
                f.Li[lnz + before] = i;
                f.Lx[lnz + before] = v;
            }
            lnz += __popcll(bal);
        }
        if (lane == 0) f.Lp[k + 1] = (int)lnz;
        wave_sync();
    }
    return 0;
}

// Numeric refactorization on the case's pivot order and pattern. false: a zero or non-finite
// pivot (w is left dirty; a fresh factorization clears it).
__device__ bool piv_refactor(const PivArgs& a, const PivSlot& f, int c, int lane) {
    const int m = a.m;
    for (int i = lane; i < m; i += 64) f.w[i] = 0.0;
    wave_sync();
    for (int k = 0; k < m; ++k) {
        piv_scatter(a, f, c, a.Q[k], lane);
        wave_sync();
        for (int q = f.Up[k]; q < f.Up[k + 1]; ++q) {
            const int j = f.Ui[q];
            const int i = f.P[j];
            const double xj = __shfl(f.w[i], 0);
            if (lane == 0) {
                f.Ux[q] = xj;
                f.w[i] = 0.0;
            }
            piv_axpy(f, j, xj, lane);
            wave_sync();
        }
        const int pk = f.P[k];
        const double d = __shfl(f.w[pk], 0);
        if (!(fabs(d) > 0.0) || !(fabs(d) <= DBL_MAX)) return false;
        if (lane == 0) {
            f.w[pk] = 0.0;
            f.Ud[k] = d;
        }
        for (int q = f.Lp[k] + lane; q < f.Lp[k + 1]; q += 64) {
            const int i = f.Li[q];
            f.Lx[q] = f.w[i] / d;
            f.w[i] = 0.0;
        }
        wave_sync();
    }
    return true;
}

// x = A^-1 R with the case's factors (klu_solve): y = (R / rs) permuted by P, L y, U z, then
// z in column order, which is the plan's pivot order.
__device__ void piv_solve(const PivArgs& a, const PivSlot& f, int c, int lane) {
    const int m = a.m;
    for (int k = lane; k < m; k += 64) {
        const int i = f.P[k];
        f.y[k] = a.R[at(i, a.S, c)] / f.rs[i];
    }
    wave_sync();
    for (int k = 0; k < m; ++k) {
        const double yk = f.y[k];
        if (yk != 0.0)
            for (int q = f.Lp[k] + lane; q < f.Lp[k + 1]; q += 64) f.y[f.pinv[f.Li[q]]] -= f.Lx[q] * yk;
        wave_sync();
    }
    for (int k = m - 1; k >= 0; --k) {
        const double zk = __shfl(f.y[k], 0) / f.Ud[k];
        if (lane == 0) f.y[k] = zk;
        if (zk != 0.0)
            for (int q = f.Up[k] + lane; q < f.Up[k + 1]; q += 64) f.y[f.Ui[q]] -= f.Ux[q] * zk;
        wave_sync();
    }
    for (int k = lane; k < m; k += 64) a.x[at(k, a.S, c)] = f.y[k];
}

// The pivoting solve of every case in the list (cases in STEP off the frozen plan): a fresh
// factorization for a case that needs one, else a refactorization on its order -- a fresh one
// after all when that meets a zero pivot (p3s/cpp/nr_klu.cpp:629-651) -- then the solve. A case
// whose fresh factorization is singular ends FAIL with this iteration counted, or, on its first
// run, starts again with pivoting (as k_control restarts it). res 2: the slot's capacity was too
// small for a fresh factorization; the case is left as it was, for a larger pool.
__global__ __launch_bounds__(64) void k_pivot_solve(PivArgs a) {
    const int lane = threadIdx.x;
    const int c = a.list[blockIdx.x];
    const unsigned w0 = a.ctl[c];
    if (piv_of(w0) == PV_DIAG) {  // k_pivot_verify took the static factorization on the diagonal
        if (lane == 0) a.res[blockIdx.x] = 4;
        return;
    }
    const int slot = a.slot_of[c];
    if (slot < 0) {  // no slot assigned: the host reports it
        if (lane == 0) a.res[blockIdx.x] = 3;
        return;
    }
    const PivSlot f = piv_slot(a, slot);
    bool fresh = piv_of(w0) != PV_HAVE;
    bool ok = !fresh && piv_refactor(a, f, c, lane);
    if (!ok) {
        fresh = true;
        const int r = piv_factor(a, f, c, lane);
        if (r == 2) {
            if (lane == 0) {
                a.ctl[c] = set_piv(w0, PV_NEED) | CTL_PENDING;  // still waits for its step
                a.res[blockIdx.x] = 2;
            }
            return;
        }
        if (lane == 0) atomicAdd(&a.guard[4], 1);
        ok = r == 0;
    }
    if (!ok) {
        if (lane == 0) {
            if (w0 & CTL_RERUN) {
                a.ctl[c] = set_iters(set_stage(w0 & ~CTL_PENDING, SG_FAIL), iters_of(w0) + 1);
            } else {
                a.ctl[c] = rerun_word(w0);
                atomicAdd(&a.guard[3], 1);
            }
            a.res[blockIdx.x] = 1;
        }
        return;
    }
    piv_solve(a, f, c, lane);
    if (lane == 0) {
        a.ctl[c] = set_piv(w0 & ~CTL_PENDING, PV_HAVE);
        a.res[blockIdx.x] = 0;
    }
}

// First try of a fresh pivoting factorization. The threshold-pivoting LU above, taking the
// diagonal at every column, computes the static-pivot factors of the plan's order; an SEL_NEED
// launch has just computed those (and the step from them) for the cases waiting for a fresh
// factorization. With rows scaled by their largest |J| (d), column p's candidates are the free
// rows of its reach -- the rows i of L(:, p) -- with values U(p,p) L(i,p) / d_i, and its diagonal
// is U(p,p) / d_p, so the diagonal is the pivot of every column exactly when
//     |L(i, p)| d_p / d_i <= 1 / PIVOT_TOL   for every L entry,
// with every pivot finite and nonzero. Such a case keeps the plan's order (PV_DIAG: its later
// refactorizations are the static-pivot launches, without the guard, as a refactorization on a
// kept order has none) and takes its step from that launch; any other keeps PV_NEED for the
// Gilbert-Peierls factorization (the off-diagonal pivots it needs). Measured on the 94 N-1
// subsets of the Poisson serving mix (8k-43k-bus grids), the pivoting factorizations the
// oracle's restatement of KLU ran, reruns included, chose the diagonal at every column.
// One workgroup per listed case; the row scales go to the case's slot.
struct VerifyArgs {
    const int* Jp;    // J rows (row scales)
    const int* perm;  // pivot -> Jacobian row / column
    const int* Lp;    // L patterns in pivot space: rows Li[Lp[p], Lp[p + 1]) of column p, ascending
    const int* Li;
    const int* base;  // factor position of 1 / U(p,p); U row, then L column, follow
    const double* J;
    const double* fac;
    unsigned* ctl;
    const int* list;
    const int* slot_of;
    double* rs;       // slot-major row scales (m per slot)
    int* guard;       // [4] pivoting factorizations, [5] of them on the diagonal
    int m, S;
};

__global__ __launch_bounds__(256) void k_pivot_verify(VerifyArgs a) {
    const int c = a.list[blockIdx.x];
    const unsigned w = a.ctl[c];
    const int slot = a.slot_of[c];
    if (piv_of(w) != PV_NEED || slot < 0) return;  // the same in the whole workgroup
    double* rs = a.rs + (size_t)slot * (size_t)a.m;
    for (int i = threadIdx.x; i < a.m; i += blockDim.x) {
        double r = 0.0;
        for (int e = a.Jp[i]; e < a.Jp[i + 1]; ++e) r = fmax(r, fabs(a.J[at(e, a.S, c)]));
        rs[i] = (r > 0.0 && r <= DBL_MAX) ? r : 1.0;
    }
    __syncthreads();
    bool ok = true;
    for (int p = threadIdx.x; p < a.m && ok; p += blockDim.x) {
        const double inv = fabs(a.fac[at(a.base[p], a.S, c)]);
        if (!(inv > 0.0 && inv <= DBL_MAX)) {
            ok = false;
            break;
        }
        const double dp = rs[a.perm[p]];
        const int l0 = a.Lp[p], l1 = a.Lp[p + 1];
        const int lpos = a.base[p] + 1 + (l1 - l0);
        for (int r = l0; r < l1; ++r) {
            const double v = fabs(a.fac[at(lpos + (r - l0), a.S, c)]) * dp / rs[a.perm[a.Li[r]]];
            if (!(v <= 1.0 / PIVOT_TOL)) {
                ok = false;
                break;
            }
        }
    }
    ok = __syncthreads_and(ok);
    if (threadIdx.x == 0 && ok) {
        a.ctl[c] = set_prec(set_piv(w & ~(CTL_PENDING | CTL_REFINING), PV_DIAG), PR_F64_ALWAYS);
        atomicAdd(&a.guard[4], 1);
        atomicAdd(&a.guard[5], 1);
    }
}

// The cases in STEP factored by the pivoting LU, appended to list (count[0] = how many).
__global__ void k_pivot_list(const unsigned* __restrict__ ctl, int S, int* __restrict__ list, int* __restrict__ count) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= S) return;
    const unsigned w = ctl[c];
    if (pivoting_step(w)) list[atomicAdd(count, 1)] = c;
}

// Time-series injections on the device (p3s/timeseries.py:36-141): for bus i and case t,
// S[i, t] = s0[i] + sum_k coef[k] * (X[row[k], t] - base[k]) over the contributions k of
// bus i, [ptr[i], ptr[i + 1]), in their order, added to the imaginary part where imag[k].
// X stacks the element series (rows, T). One thread per (bus, case): the case index runs
// along the thread x dimension, so X rows and S rows are read and written in contiguous
// lines. Products and sums are rounded one by one (no contraction), as the reference's
// loop rounds them.
__global__ void k_sbus(const double* __restrict__ s0, const int* __restrict__ ptr, const int* __restrict__ row,
                       const double* __restrict__ coef, const double* __restrict__ base,
                       const unsigned char* __restrict__ imag, const double* __restrict__ X, double* __restrict__ out,
                       int n, int T) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    for (int i = blockIdx.y; i < n; i += gridDim.y) {
        double re = s0[2 * i], im = s0[2 * i + 1];
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
            const double d = X[(size_t)row[k] * T + t] - base[k];
            const double v = coef[k] * d;
            if (imag[k]) im = im + v;
            else re = re + v;
        }
        out[2 * ((size_t)i * T + t)] = re;
        out[2 * ((size_t)i * T + t) + 1] = im;
    }
}

// Voltage band of every case: min and max |V| over its buses, skipping buses a served mask
// marks unserved. NewtonPowerflowCpp.calculate accepts a converged solve only if every |V| lies
// in [0.5, 1.5] p.u. (p3s/NewtonPowerflowCpp.py:553-584); |V| is that of the returned voltage.
// One thread per case, reads coalesced across cases.
__global__ void k_vrange(const double* __restrict__ Ve, const double* __restrict__ Vf,
                         const unsigned char* __restrict__ served, double* __restrict__ vmin,
                         double* __restrict__ vmax, int n, int S, int B) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= B) return;
    double lo = DBL_MAX, hi = -DBL_MAX;
    for (int i = 0; i < n; ++i) {
        const size_t o = at(i, S, c);
        if (served && !served[o]) continue;
        const double v = hypot(Ve[o], Vf[o]);
        lo = fmin(lo, v);
        hi = fmax(hi, v);
    }
    vmin[c] = lo;
    vmax[c] = hi;
}

// ===================================================================== N-1 post-processing
// Branch results of every outage case from the resident solve state, as
// compute_branch_loading does them on the (n, L) voltage table
// (p3s/contingency/find_overloads.py:128-262): with V = Vm cos + j Vm sin at each end and
// the intact branch stamps, S_from = conj(yff V_f + yft V_t) V_f sn, S_to = conj(ytf V_f +
// ytt V_t) V_t sn; I = |S| / (vn |V| sqrt 3); loading = max(I_from, I_to) / rating * 100
// for a line, max(|S_from|-equivalent, |S_to|-equivalent) / rating * 100 for a
// transformer. A branch touching an unserved bus is NaN (its voltages are), a branch of the
// case's own outage group is 0 (it carries nothing), a non-converged case is NaN (reactive
// powers 0, as the reference's complex NaN assignment leaves them): in that order of
// precedence, as the reference applies them. Products and sums are rounded one by one in
// the reference's order. Thread per (case, branch), cases along x.
__device__ __forceinline__ void cmul(double ar, double ai, double br, double bi, double& cr, double& ci) {
#pragma clang fp contract(off)
    cr = ar * br - ai * bi;
    ci = ar * bi + ai * br;
}

__global__ void k_branch_flows(const int* __restrict__ fb, const int* __restrict__ tb, const double* __restrict__ yval,
                               const int* __restrict__ gptr, const int* __restrict__ gbr, const double* __restrict__ Ve,
                               const double* __restrict__ Vf, const unsigned char* __restrict__ served,
                               const unsigned* __restrict__ ctl,
                               const double* __restrict__ vnf, const double* __restrict__ vnt,
                               const double* __restrict__ rating, const unsigned char* __restrict__ trafo, double sn_mva,
                               int nbr, int L, int S, double* __restrict__ pf, double* __restrict__ qf,
                               double* __restrict__ pt, double* __restrict__ qt, double* __restrict__ i_from,
                               double* __restrict__ i_to, double* __restrict__ i_ka, double* __restrict__ loading) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= L) return;
    const double nan = __longlong_as_double(0x7ff8000000000000ll);
    const double sqrt3 = 1.7320508075688772;  // the double nearest sqrt(3)
    for (int k = blockIdx.y; k < nbr; k += gridDim.y) {
        const int f = fb[k], t = tb[k];
        const size_t of = at(f, S, c), ot = at(t, S, c);
        double vfr = Ve[of], vfi = Vf[of];
        double vtr = Ve[ot], vti = Vf[ot];
        if (!served[of]) vfr = vfi = nan;
        if (!served[ot]) vtr = vti = nan;
        const double* y = yval + 8 * (size_t)k;  // ff, ft, tf, tt (re, im)
        // Yf V and Yt V, terms in ascending bus order (the CSR order of the sparse product)
        double a1r, a1i, a2r, a2i, b1r, b1i, b2r, b2i;
        cmul(y[0], y[1], vfr, vfi, a1r, a1i);  // yff V_f
        cmul(y[2], y[3], vtr, vti, a2r, a2i);  // yft V_t
        cmul(y[4], y[5], vfr, vfi, b1r, b1i);  // ytf V_f
        cmul(y[6], y[7], vtr, vti, b2r, b2i);  // ytt V_t
        double ifr, ifi, itr, iti;
        if (f < t) {
            ifr = a1r + a2r; ifi = a1i + a2i;
            itr = b1r + b2r; iti = b1i + b2i;
        } else {
            ifr = a2r + a1r; ifi = a2i + a1i;
            itr = b2r + b1r; iti = b2i + b1i;
        }
        double sfr, sfi, str_, sti;
        cmul(ifr, -ifi, vfr, vfi, sfr, sfi);  // conj(I_f) V_f
        cmul(itr, -iti, vtr, vti, str_, sti);
        sfr = sfr * sn_mva; sfi = sfi * sn_mva;
        str_ = str_ * sn_mva; sti = sti * sn_mva;
        const double vmf = hypot(vfr, vfi), vmt = hypot(vtr, vti);
        double cif = hypot(sfr, sfi) / (vnf[k] * vmf * sqrt3);
        double cit = hypot(str_, sti) / (vnt[k] * vmt * sqrt3);
        double imax = (cif != cif || cit != cit) ? nan : fmax(cif, cit);
        double ld;
        if (trafo[k]) {
            const double a = cif * vnf[k] * sqrt3, b = cit * vnt[k] * sqrt3;
            ld = (a != a || b != b) ? nan : fmax(a, b);
            ld = ld / rating[k] * 100.0;
        } else {
            ld = imax / rating[k] * 100.0;
        }
        bool out = false;
        for (int q = gptr[c]; q < gptr[c + 1]; ++q) out = out || gbr[q] == k;
        if (out) sfr = sfi = str_ = sti = cif = cit = imax = ld = 0.0;
        if (stage_of(ctl[c]) != SG_CONV) {
            // the reference assigns NaN into its complex power arrays, which stores NaN + 0j:
            // a non-converged case reports NaN active and 0 reactive power
            // (p3s/contingency/find_overloads.py:249-252)
            sfr = str_ = cif = cit = imax = ld = nan;
            sfi = sti = 0.0;
        }
        const size_t o = (size_t)k * L + c;
        if (pf) { pf[o] = sfr; qf[o] = sfi; pt[o] = str_; qt[o] = sti; }
        if (i_from) { i_from[o] = cif; i_to[o] = cit; }
        i_ka[o] = imax;
        loading[o] = ld;
    }
}

// Per branch: the largest loading over the cases and the first case reaching it (NaN and
// -1 when every case is NaN), as nanmax / nanargmax do it; and every (branch, case) above
// the threshold into a list (atomic append, order arbitrary: the host sorts it).
__global__ __launch_bounds__(256) void k_branch_worst(const double* __restrict__ loading, int L,
                                                      double* __restrict__ best, int* __restrict__ arg) {
    __shared__ double sv[256];
    __shared__ int si[256];
    const int k = blockIdx.x;
    double v = -DBL_MAX;
    int idx = INT_MAX;
    for (int c = threadIdx.x; c < L; c += blockDim.x) {
        const double x = loading[(size_t)k * L + c];
        if (x == x && (x > v || (x == v && c < idx))) { v = x; idx = c; }
    }
    sv[threadIdx.x] = v;
    si[threadIdx.x] = idx;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double ov = sv[threadIdx.x + w];
            const int oi = si[threadIdx.x + w];
            if (ov > sv[threadIdx.x] || (ov == sv[threadIdx.x] && oi < si[threadIdx.x])) {
                sv[threadIdx.x] = ov;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const bool none = si[0] == INT_MAX;
        best[k] = none ? __longlong_as_double(0x7ff8000000000000ll) : sv[0];
        arg[k] = none ? -1 : si[0];
    }
}

__global__ void k_over_list(const double* __restrict__ loading, const double* __restrict__ i_ka, int nbr, int L,
                            double threshold, int cap, int* __restrict__ count, int* __restrict__ ob,
                            int* __restrict__ oc, double* __restrict__ ol, double* __restrict__ oi) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= L) return;
    for (int k = blockIdx.y; k < nbr; k += gridDim.y) {
        const size_t o = (size_t)k * L + c;
        const double x = loading[o];
        if (x > threshold) {  // NaN never compares true
            const int j = atomicAdd(count, 1);
            if (j < cap) { ob[j] = k; oc[j] = c; ol[j] = x; oi[j] = i_ka[o]; }
        }
    }
}

// Bus voltage-band violations (p3s/contingency/voltage_band.py:103-182): |V| below
// vm_min - tol (kind 0) or above vm_max + tol (kind 1) at served, in-service buses of
// converged cases; a NaN bound never flags.
__global__ void k_vviol(const double* __restrict__ Ve, const double* __restrict__ Vf,
                        const unsigned char* __restrict__ served, const unsigned* __restrict__ ctl,
                        const double* __restrict__ vmin, const double* __restrict__ vmax,
                        const unsigned char* __restrict__ in_service, double tol, int n, int L, int S, int cap,
                        int* __restrict__ count, int* __restrict__ ob, int* __restrict__ oc, double* __restrict__ ov,
                        unsigned char* __restrict__ okind) {
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= L || stage_of(ctl[c]) != SG_CONV) return;
    for (int i = blockIdx.y; i < n; i += gridDim.y) {
        const size_t o = at(i, S, c);
        if (!served[o] || !in_service[i]) continue;
        const double vm = hypot(Ve[o], Vf[o]);
        const bool low = vm < vmin[i] - tol, high = vm > vmax[i] + tol;
        if (low || high) {
            const int j = atomicAdd(count, 1);
            if (j < cap) { ob[j] = i; oc[j] = c; ov[j] = vm; okind[j] = high ? 1 : 0; }
        }
    }
}

// ===================================================================== slots (serving)
// Copy k staged cases (S and V0, (n, k) complex) into slot columns and mark them START.
__global__ __launch_bounds__(256) void k_load_slots(const double* __restrict__ Sst, const double* __restrict__ Vst,
                                                    const int* __restrict__ slots, double* __restrict__ Ssp,
                                                    double* __restrict__ V0, unsigned* __restrict__ ctl, unsigned extra,
                                                    int n, int k, int S) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int bus = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (j >= k || bus >= n) return;
    const int c = slots[j];
    const size_t src = 2 * ((size_t)bus * k + j), dst = 2 * ((size_t)bus * S + c);
    Ssp[dst] = Sst[src];
    Ssp[dst + 1] = Sst[src + 1];
    V0[dst] = Vst[src];
    V0[dst + 1] = Vst[src + 1];
    if (bus == 0) ctl[c] = SG_START | extra;
}

// Results of k slots into (n, k) complex V, iterations, flags, final mismatch.
__global__ __launch_bounds__(256) void k_fetch_slots(const int* __restrict__ slots, const double* __restrict__ Ve,
                                                     const double* __restrict__ Vf, const unsigned* __restrict__ ctl,
                                                     const double* __restrict__ nrm, double* __restrict__ V,
                                                     int* __restrict__ it_out, unsigned char* __restrict__ cv_out,
                                                     double* __restrict__ mis_out, int n, int k, int S) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int bus = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (j >= k || bus >= n) return;
    const int c = slots[j];
    const size_t o = at(bus, S, c);
    V[2 * ((size_t)bus * k + j)] = Ve[o];
    V[2 * ((size_t)bus * k + j) + 1] = Vf[o];
    if (bus == 0) {
        const unsigned w = ctl[c];
        it_out[j] = iters_of(w);
        cv_out[j] = stage_of(w) == SG_CONV;
        mis_out[j] = nrm[c];
    }
}

__global__ void k_release_slots(const int* __restrict__ slots, unsigned* __restrict__ ctl, int k) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < k) ctl[slots[j]] = SG_IDLE;
}

// ===================================================================== N-1 islands
// Union-find over the live, non-outaged branches of one case per workgroup. A root is
// hooked under the smaller of the two roots, so every non-root points to a smaller bus
// index and the final root of an island is its smallest bus. The parent array lives in
// global memory; every access to it is an agent-scope atomic (CAS hooking and the reads
// that race with it go to L2, no stale L1 line is ever read), and every find walks a
// strictly decreasing chain of bus indices, so it ends.
__device__ __forceinline__ int uf_ld(int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void uf_st(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int uf_find(int* parent, int x) {
    for (;;) {
        const int p = uf_ld(&parent[x]);
        if (p == x) return x;
        const int gp = uf_ld(&parent[p]);
        if (gp != p) uf_st(&parent[x], gp);  // path halving: gp < p < x
        x = gp;
    }
}

__device__ __forceinline__ void uf_unite(int* parent, int a, int b) {
    for (;;) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a < b) { const int t = a; a = b; b = t; }
        if (atomicCAS(&parent[a], a, b) == a) return;  // else a was hooked meanwhile: retry
    }
}

constexpr int MAX_GROUP = 256;   // branches one outage case may remove
constexpr unsigned long long INF_BITS = 0x7ff0000000000000ull;

// Islands and solver inputs of cases [c0, c0 + gridDim.x) (p3s/contingency/case_generator.py:
// 355-430, p3s/contingency/solver_cpp.py:119-134). Column cl = blockIdx.x of the outputs
// (stride S): served, ref bus (or -1), pin (re-slack references and unserved buses) and the
// start voltage (|V| + 0j at a re-slack reference, 1 + 0j at unserved buses).
// LDS: the union-find parents and island references of the case (8 n bytes) live in the
// workgroup's dynamic LDS, so the unions are LDS atomics; else in ufbuf (2 n ints per case).
// Only the workgroup's own threads share these arrays (and its |Z| scratch), so a barrier
// between phases is all the ordering they need; a device-scope fence adds nothing but its
// cost to the LDS form.
template <bool LDS>
__device__ __forceinline__ void phase_end() {
    if (!LDS) __threadfence();
    __syncthreads();
}

template <bool LDS>
__global__ __launch_bounds__(256) void k_islands(
    int n, int nbr, const int* __restrict__ f, const int* __restrict__ t, const unsigned char* __restrict__ live,
    const int* __restrict__ gptr, const int* __restrict__ gbr, int nref, const int* __restrict__ ref, int ngen,
    const int* __restrict__ gbus, const double* __restrict__ gz, int reslack, int c0, int* __restrict__ ufbuf,
    unsigned long long* __restrict__ zscratch, const double* __restrict__ Vstart, unsigned char* __restrict__ served,
    int* __restrict__ refbus, unsigned char* __restrict__ pin, double* __restrict__ V0, int* __restrict__ npinned,
    int S) {
    __shared__ int gm[MAX_GROUP];
    extern __shared__ int s_uf[];
    const int cl = blockIdx.x;
    const int cs = c0 + cl;
    const int g0 = gptr[cs], gsz = gptr[cs + 1] - g0;
    for (int q = threadIdx.x; q < gsz; q += blockDim.x) gm[q] = gbr[g0 + q];
    int* parent = LDS ? s_uf : ufbuf + (size_t)cl * 2 * n;
    int* cref = parent + n;
    for (int v = threadIdx.x; v < n; v += blockDim.x) {
        uf_st(&parent[v], v);
        uf_st(&cref[v], INT_MAX);
    }
    phase_end<LDS>();
    for (int e = threadIdx.x; e < nbr; e += blockDim.x) {
        if (!live[e]) continue;
        bool out = false;
        for (int q = 0; q < gsz; ++q) out |= gm[q] == e;
        if (!out) uf_unite(parent, f[e], t[e]);
    }
    phase_end<LDS>();
    for (int v = threadIdx.x; v < n; v += blockDim.x) uf_st(&parent[v], uf_find(parent, v));
    phase_end<LDS>();
    // the lowest-index slack bus of each island
    for (int i = threadIdx.x; i < nref; i += blockDim.x) atomicMin(&cref[uf_ld(&parent[ref[i]])], ref[i]);
    phase_end<LDS>();
    unsigned long long* zb = zscratch ? zscratch + (size_t)cl * n : nullptr;
    if (reslack && ngen > 0) {
        // islands without a slack: the lowest generator |Z|, then the lowest bus index among
        // the generators within np.isclose of it (rtol 1e-5, atol 1e-8)
        for (int v = threadIdx.x; v < n; v += blockDim.x) zb[v] = INF_BITS;
        phase_end<LDS>();
        for (int i = threadIdx.x; i < ngen; i += blockDim.x) {
            const int r = uf_ld(&parent[gbus[i]]);
            if (uf_ld(&cref[r]) == INT_MAX) atomicMin(&zb[r], (unsigned long long)__double_as_longlong(gz[i]));
        }
        phase_end<LDS>();
        for (int i = threadIdx.x; i < ngen; i += blockDim.x) {
            const int r = uf_ld(&parent[gbus[i]]);
            const unsigned long long zr = __hip_atomic_load(&zb[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (zr == INF_BITS) continue;
            const double best = __longlong_as_double((long long)zr);
            if (fabs(gz[i] - best) <= 1e-8 + 1e-5 * fabs(best)) atomicMin(&cref[r], gbus[i]);
        }
        phase_end<LDS>();
    }
    for (int v = threadIdx.x; v < n; v += blockDim.x) {
        const int r = uf_ld(&parent[v]);
        const int cr = uf_ld(&cref[r]);
        const bool srv = cr != INT_MAX;
        const bool reslacked = srv && zb != nullptr && reslack &&
                               __hip_atomic_load(&zb[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != INF_BITS;
        const size_t o = (size_t)v * S + cl;
        served[o] = srv;
        if (refbus) refbus[o] = srv ? cr : -1;
        const bool pinned_ref = reslacked && cr == v;
        if (pinned_ref) atomicAdd(&npinned[cl], 1);
        if (pin) pin[o] = (!srv || pinned_ref) ? 1 : 0;
        if (V0) {
            double re = Vstart[2 * v], im = Vstart[2 * v + 1];
            if (!srv) { re = 1.0; im = 0.0; }
            else if (pinned_ref) { re = hypot(re, im); im = 0.0; }
            V0[2 * o] = re;
            V0[2 * o + 1] = im;
        }
    }
}

// ===================================================================== host side
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        if (count == 0) count = 1;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            throw Error(P3SMI_E_MEMORY, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        alloc(count);
        if (count) HIPCHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

// ---- host copies of the callers' arrays
// Copies between pageable host memory and the device run at about half of what the link
// does from pinned memory (the runtime stages them itself, on one thread), and a download
// into pages the caller has not touched yet pays their faults on that one thread too. Large
// transfers of the one-call entry points and of resident batches therefore go through a ring
// of pinned slots: a few host threads pack (upload) or scatter (download) one slice of rows
// while the DMA of the neighbouring slice runs.

// Workers that run one function on every worker and return when all are done. One pool per
// process, created on first use (again in a forked child) and never destroyed; callers from
// several threads take turns.
class HostPool {
  public:
    explicit HostPool(int n) : n_(n), pid_(getpid()) {
        for (int i = 1; i < n; ++i) std::thread([this, i] { loop(i); }).detach();
    }
    int size() const { return n_; }
    pid_t pid() const { return pid_; }
    void run(const std::function<void(int)>& f) {
        std::lock_guard<std::mutex> turn(run_mu_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &f;
            pending_ = n_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        f(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return pending_ == 0; });
        job_ = nullptr;
    }

  private:
    void loop(int i) {
        unsigned long long seen = 0;
        for (;;) {
            const std::function<void(int)>* f;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                f = job_;
            }
            (*f)(i);
            std::lock_guard<std::mutex> lk(mu_);
            if (--pending_ == 0) done_.notify_one();
        }
    }
    const int n_;
    const pid_t pid_;
    std::mutex run_mu_, mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)>* job_ = nullptr;
    int pending_ = 0;
    unsigned long long gen_ = 0;
};

HostPool& host_pool() {
    static std::mutex mu;
    static HostPool* pool = nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (!pool || pool->pid() != getpid()) {
        const unsigned hc = std::thread::hardware_concurrency();
        pool = new HostPool((int)std::max(1u, std::min(8u, hc / 2)));
    }
    return *pool;
}

// Pinned slots of one batch, allocated on its first large transfer.
struct Staging {
    static constexpr int NSLOT = 4;
    static constexpr size_t SLOT = (size_t)8 << 20;
    static constexpr size_t MIN_BYTES = (size_t)16 << 20;  // smaller transfers copy directly
    char* buf[NSLOT] = {};
    hipEvent_t ev[NSLOT] = {};
    DevBuf<char> xfer;  // device side of narrow-row transfers (h2d_rows, d2h_rows)
    Staging() = default;
    Staging(const Staging&) = delete;
    Staging& operator=(const Staging&) = delete;
    ~Staging() {
        for (int s = 0; s < NSLOT; ++s) {
            if (ev[s]) (void)hipEventSynchronize(ev[s]);
            if (ev[s]) (void)hipEventDestroy(ev[s]);
            if (buf[s]) (void)hipHostFree(buf[s]);
        }
    }
    void ensure() {
        for (int s = 0; s < NSLOT; ++s) {
            if (!buf[s] && hipHostMalloc((void**)&buf[s], SLOT, hipHostMallocDefault) != hipSuccess) {
                buf[s] = nullptr;
                throw Error(P3SMI_E_MEMORY, "pinned staging allocation failed");
            }
            if (!ev[s]) HIPCHECK(hipEventCreateWithFlags(&ev[s], hipEventDisableTiming));
        }
    }
};

// Worker w of W copies bytes [split64(nb, w, W), split64(nb, w + 1, W)) of a slice: 64-byte
// aligned cuts, the last worker ends at nb.
size_t split64(size_t nb, size_t w, size_t W) { return w >= W ? nb : (nb * w / W) & ~(size_t)63; }

// Bytes [b0, b1) of a slice packed row after row (row = rowb bytes) <-> rows of a strided array
// (row r at base + r * ldb).
void pack_range(char* packed, const char* base, size_t ldb, size_t rowb, size_t b0, size_t b1) {
    while (b0 < b1) {
        const size_t r = b0 / rowb, off = b0 % rowb, len = std::min(rowb - off, b1 - b0);
        std::memcpy(packed + b0, base + r * ldb + off, len);
        b0 += len;
    }
}
void unpack_range(char* base, const char* packed, size_t ldb, size_t rowb, size_t b0, size_t b1) {
    while (b0 < b1) {
        const size_t r = b0 / rowb, off = b0 % rowb, len = std::min(rowb - off, b1 - b0);
        std::memcpy(base + r * ldb + off, packed + b0, len);
        b0 += len;
    }
}

// Rows of `width` bytes between pageable host memory and the device (host pitch spitch or
// dpitch). A 2D copy from or to pageable memory moves row by row, so rows narrower than
// WIDE_ROW (a few cases' worth) are packed on the host, copied in one piece, and spread or
// gathered on the device by a device-to-device 2D copy. Both return once the host side is
// done with.
constexpr size_t WIDE_ROW = 4096;

void h2d_rows(Staging& sg, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t rows,
              hipStream_t st) {
    if (rows == 0 || width == 0) return;
    if (width >= WIDE_ROW) {
        HIPCHECK(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, hipMemcpyHostToDevice, st));
        return;
    }
    const size_t nb = width * rows;
    std::vector<char> packed(nb);
    pack_range(packed.data(), (const char*)src, spitch, width, 0, nb);
    if (dpitch == width) {
        HIPCHECK(hipMemcpyAsync(dst, packed.data(), nb, hipMemcpyHostToDevice, st));
    } else {
        if (sg.xfer.n < nb) sg.xfer.alloc(nb);
        HIPCHECK(hipMemcpyAsync(sg.xfer.p, packed.data(), nb, hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpy2DAsync(dst, dpitch, sg.xfer.p, width, width, rows, hipMemcpyDeviceToDevice, st));
    }
    HIPCHECK(hipStreamSynchronize(st));
}

void d2h_rows(Staging& sg, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t rows,
              hipStream_t st) {
    if (rows == 0 || width == 0) return;
    if (width >= WIDE_ROW) {
        HIPCHECK(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        return;
    }
    const size_t nb = width * rows;
    std::vector<char> packed(nb);
    if (spitch == width) {
        HIPCHECK(hipMemcpyAsync(packed.data(), src, nb, hipMemcpyDeviceToHost, st));
    } else {
        if (sg.xfer.n < nb) sg.xfer.alloc(nb);
        HIPCHECK(hipMemcpy2DAsync(sg.xfer.p, width, src, spitch, width, rows, hipMemcpyDeviceToDevice, st));
        HIPCHECK(hipMemcpyAsync(packed.data(), sg.xfer.p, nb, hipMemcpyDeviceToHost, st));
    }
    HIPCHECK(hipStreamSynchronize(st));
    unpack_range((char*)dst, packed.data(), dpitch, width, 0, nb);
}

// Column block [c0, c0 + cols) of a (rows, ld) host array of `elem` bytes -> device (rows, cols),
// queued on `st`.
void upload_cols(Staging& sg, void* dst, const void* src, size_t rows, size_t ld, size_t c0, size_t cols,
                 size_t elem, hipStream_t st) {
    const size_t rowb = cols * elem, ldb = ld * elem;
    const char* base = (const char*)src + c0 * elem;
    if (rows * rowb < Staging::MIN_BYTES || rowb > Staging::SLOT) {
        if (rowb == ldb) HIPCHECK(hipMemcpyAsync(dst, base, rows * rowb, hipMemcpyHostToDevice, st));
        else h2d_rows(sg, dst, rowb, base, ldb, rowb, rows, st);
        return;
    }
    sg.ensure();
    HostPool& pool = host_pool();
    const size_t prow = Staging::SLOT / rowb;  // rows per slice
    for (size_t r0 = 0, i = 0; r0 < rows; r0 += prow, ++i) {
        const int s = (int)(i % Staging::NSLOT);
        HIPCHECK(hipEventSynchronize(sg.ev[s]));  // the slot's previous DMA is done
        const size_t nb = std::min(prow, rows - r0) * rowb;
        char* buf = sg.buf[s];
        const char* sp = base + r0 * ldb;
        const size_t W = (size_t)pool.size();
        pool.run([&](int w) { pack_range(buf, sp, ldb, rowb, split64(nb, w, W), split64(nb, w + 1, W)); });
        HIPCHECK(hipMemcpyAsync((char*)dst + r0 * rowb, buf, nb, hipMemcpyHostToDevice, st));
        HIPCHECK(hipEventRecord(sg.ev[s], st));
    }
}

// Device (rows, cols) -> column block [c0, c0 + cols) of a (rows, ld) host array. Waits for the
// work queued on `st` before it.
void download_cols(Staging& sg, void* dst, const void* src, size_t rows, size_t ld, size_t c0, size_t cols,
                   size_t elem, hipStream_t st) {
    const size_t rowb = cols * elem, ldb = ld * elem;
    char* base = (char*)dst + c0 * elem;
    if (rows * rowb < Staging::MIN_BYTES || rowb > Staging::SLOT) {
        if (rowb == ldb) {
            HIPCHECK(hipMemcpyAsync(base, src, rows * rowb, hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
        } else {
            d2h_rows(sg, base, ldb, src, rowb, rowb, rows, st);
        }
        return;
    }
    sg.ensure();
    HostPool& pool = host_pool();
    const size_t prow = Staging::SLOT / rowb;
    const size_t nslice = (rows + prow - 1) / prow;
    auto issue = [&](size_t i) {
        const int s = (int)(i % Staging::NSLOT);
        const size_t r0 = i * prow, nb = std::min(prow, rows - r0) * rowb;
        HIPCHECK(hipMemcpyAsync(sg.buf[s], (const char*)src + r0 * rowb, nb, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipEventRecord(sg.ev[s], st));
    };
    for (int s = 0; s < Staging::NSLOT; ++s) HIPCHECK(hipEventSynchronize(sg.ev[s]));  // earlier uses
    for (size_t i = 0; i < std::min(nslice, (size_t)Staging::NSLOT); ++i) issue(i);
    for (size_t i = 0; i < nslice; ++i) {
        const int s = (int)(i % Staging::NSLOT);
        HIPCHECK(hipEventSynchronize(sg.ev[s]));
        const size_t r0 = i * prow, nb = std::min(prow, rows - r0) * rowb;
        const char* buf = sg.buf[s];
        char* dp = base + r0 * ldb;
        const size_t W = (size_t)pool.size();
        pool.run([&](int w) { unpack_range(dp, buf, ldb, rowb, split64(nb, w, W), split64(nb, w + 1, W)); });
        if (i + Staging::NSLOT < nslice) issue(i + Staging::NSLOT);
    }
}

struct DevSched {  // one level schedule dealt to TL task lists
    int TL = 0, nlev = 0, nblev = 0;
    bool blk = false;  // 2x2-block schedule: taskA / taskB hold 3 int4 per block
    DevBuf<int> ptrA, ptrB, ptrF, ptrW, diagB;
    DevBuf<int4> taskA, taskB, fwd, bwd;
    DevBuf<int4> steps;   // 2x2-block schedule: the block steps
    DevBuf<int4> fsteps;  // 2x2-block schedule: the forward-row steps (fwd: 2 int4 per group)
    DevBuf<int4> bsteps;  // 2x2-block schedule: the backward-row steps (bwd: 2 int4 per group)
    DevBuf<int2> pairs, fent, bent;  // operand streams in task-list order
};

struct DevTopo {
    DevBuf<int> Yp, Yj, Ydiag, row_p, row_q, posJ, xva, xvm, Jp, xcol;
    DevBuf<int> Cp, Ci, Ce, perm;  // the Jacobian by column and the column order (pivoting LU)
    DevBuf<int> Lp, Li, base;      // the plan's L patterns and factor positions (k_pivot_verify)
    DevBuf<double> G, B;
    DevBuf<int> tiles;    // k_evaluate row tiles: first row of each, then n
    int ntiles = 0, tile_rows = 0;
    std::map<int, std::unique_ptr<DevSched>> sched;
    bool ready = false;
};

// LU launch shape: G cases per workgroup, NS wavefronts; grp = 0: 2x2 supernode blocks
// when G = 64 (k_lu_blk), single entries otherwise; grp > 0: runs of up to grp entries of
// one pivot dealt together (1 = single entries).
struct LuConfig {
    int G, NS, grp;
    int tl() const { return 64 * NS / G; }
    bool blk() const { return G == 64 && grp == 0; }
    int key() const { return (tl() * 8192 + grp) * 2 + (blk() ? 1 : 0); }
};

}  // namespace

struct p3smi_topology {
    p3smi::Topology host;
    DevTopo dev;
    LuConfig lu_batch{64, 16, 0};  // LU launch shape while many cases are active
    LuConfig lu_few{1, 16, 0};     // ... and for the last few
    int lu_few_max = 256;         // "few": at most this many active cases
    bool graphs = false;          // one-call solves replay captured Newton iterations
    double mixed_rtol = 1e-12;    // refinement stops once |dx correction| <= rtol |dx|
    double mixed_switch = 1e-3;   // switching mode: FP32 factors while the mismatch is above this
    int mixed_kmax = 6;           // refinement steps before the FP64 fallback
    int max_chunk = 0;            // one-call entry points: cases per chunk at most (0: sized to free HBM)
    int eval_rows = 32;           // k_evaluate: rows per tile at most (the autotuner may change it)
    double guard_ratio = REUSE_RCOND_RATIO;  // frozen plan rejected below guard_ratio x rcond0
    bool pivot_diagonal_first = true;  // a fresh pivoting factorization tries the plan's diagonal first
    int n_cu = 256;                    // compute units of the current device (set with the stream)
    hipStream_t stream = nullptr;
    std::unique_ptr<p3smi_batch> cached;  // workspace of the one-call entry points
    std::unique_ptr<p3smi_batch> dc;      // one-case workspace of the DC start
    std::mutex mu;
    ~p3smi_topology();
};

// Slots of the pivoting factorization (see k_pivot_solve): per slot the L and U factors (cap
// entries each), pattern, pivot order, row scales and work vectors, slot-major. A case keeps its
// slot while it runs off the frozen plan (its factors persist between iterations).
struct PivotPool {
    int slots = 0, m = 0;
    long long cap = 0;
    DevBuf<int> Lp, Li, Up, Ui, P, pinv, xi, stk, pstk, flag;
    DevBuf<double> Lx, Ux, Ud, rs, w, y;
    DevBuf<int> slot_of, list, res, count;
    std::vector<int> h_slot_of;  // host mirror of slot_of
    std::vector<int> free;       // free slots
};

struct p3smi_batch {
    p3smi_topology* topo = nullptr;
    int B = 0, S = 0;
    DevBuf<double> Vm, Va, Ve, Vf, Vm0, Va0, Jx, fac, R, x;  // polar, rectangular and base state
    DevBuf<float> fac32;                // FP32 factors (mixed precision)
    DevBuf<double> cmp_fac, cmp_x;      // compact one-case arenas (LuCompact), cmp_cap slots
    DevBuf<int> cmp_ctr;
    int cmp_cap = 0;
    DevBuf<double> rr, dvec, prevd, rpart;  // refinement residual, correction, norms
    int nparts_r = 0;
    bool mixed_ready = false;
    DevBuf<double> nrm, alpha, best_alpha, best_nrm, part, pstat;
    DevBuf<unsigned> ctl;               // per-case control word
    DevBuf<int> tally, counts, guard, rcount;  // k_control sums + finish counter, its result; guards; refining
    DevBuf<double> Ssp, V0, Vout;       // inputs / output in the caller's layouts
    DevBuf<double> Yg, Yb;              // per-case Ybus (N-1)
    DevBuf<unsigned char> pin;          // per-case pin mask (N-1)
    DevBuf<unsigned char> served;       // per-case served mask (N-1 outage sets)
    int s_stride = 0, v0_stride = 0;
    bool per_case_y = false, has_pin = false, uploaded = false, mask_unserved = false;
    // per-case admittance deltas on the shared Ybus (outage sets, stamps): null = none
    const int* ydptr = nullptr;
    const int* ydpos = nullptr;
    const double* ydval = nullptr;
    int ydc0 = 0;
    int ymode() const { return per_case_y ? 1 : (ydptr ? 2 : 0); }
    int* h_counts = nullptr;            // pinned host mirror of counts (stage counters, C ABI order)
    int* h_rcount = nullptr;            // pinned host mirror of rcount
    bool pending_start = false;         // cases in START since the last pass
    DevBuf<double> st_S, st_V, st_out;  // staging of slot loads / fetches
    DevBuf<int> st_slots, st_it;
    DevBuf<unsigned char> st_cv;
    DevBuf<double> st_mis;
    bool timing = false;
    bool slot_mode = false;             // driven by slots_init / load / pass (continuous batching)
    int precision = 0;                  // precision of the solve in progress
    unsigned long long solve_id = 0;    // stamp of the last run_newton on this batch
    bool sbus_built = false;            // Ssp holds injections built on the device (k_sbus)
    Staging stage;                      // pinned slots of large host transfers
    DevBuf<double> sbX;                 // element series of the last device Sbus build
    bool graphs = false;                // replay captured Newton iterations (FP64, timers off)
    std::map<std::string, std::pair<hipGraph_t, hipGraphExec_t>> gcache;
    PivotPool piv;                      // pivoting factorizations (allocated on first use)
    double guard_override = -1.0;       // >= 0: the frozen plan's guard ratio for this batch's LU
    p3smi_stats stats{};
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
    size_t ev_used = 0;
    void drop_graphs() {
        for (auto& g : gcache) {
            (void)hipGraphExecDestroy(g.second.second);
            (void)hipGraphDestroy(g.second.first);
        }
        gcache.clear();
    }
    ~p3smi_batch() {
        drop_graphs();
        if (h_counts) (void)hipHostFree(h_counts);
        if (h_rcount) (void)hipHostFree(h_rcount);
        for (auto& e : ev) {
            (void)hipEventDestroy(e.second.first);
            (void)hipEventDestroy(e.second.second);
        }
    }
};

struct p3smi_outages {
    p3smi_topology* topo = nullptr;
    int nbr = 0, ncases = 0, nref = 0, ngen = 0, reslack = 0;
    DevBuf<int> f, t, gptr, gbr, ref, gbus, ypos;
    DevBuf<unsigned char> live;
    DevBuf<double> gz, yval;
    DevBuf<int> dptr, dpos;  // per case, the Ybus deltas of its outaged branches in group order
    DevBuf<double> dval;
    DevBuf<int> scratch, npinned;
    DevBuf<unsigned long long> zscratch;
    // the batch and solve stamp holding this set's last results on the device (one chunk)
    const p3smi_batch* res_batch = nullptr;
    unsigned long long res_id = 0;
};

p3smi_topology::~p3smi_topology() {
    cached.reset();
    dc.reset();
    if (stream) (void)hipStreamDestroy(stream);
}

namespace {

void ensure_device(p3smi_topology* t) {
    if (t->dev.ready) return;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        throw Error(P3SMI_E_DEVICE, "no HIP device available (the MI355X engine needs a GPU)");
    if (!t->stream) HIPCHECK(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
    {
        int dev = 0, ncu = 0;
        HIPCHECK(hipGetDevice(&dev));
        HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        if (ncu > 0) t->n_cu = ncu;
    }
    const p3smi::Topology& h = t->host;
    hipStream_t st = t->stream;
    DevTopo& d = t->dev;
    d.Yp.upload(h.Yp.data(), h.Yp.size(), st);
    d.Yj.upload(h.Yj.data(), h.Yj.size(), st);
    d.Ydiag.upload(h.Ydiag.data(), h.Ydiag.size(), st);
    d.G.upload(h.G.data(), h.G.size(), st);
    d.B.upload(h.B.data(), h.B.size(), st);
    d.row_p.upload(h.row_p.data(), h.row_p.size(), st);
    d.row_q.upload(h.row_q.data(), h.row_q.size(), st);
    d.posJ.upload(h.posJ.data(), h.posJ.size(), st);
    d.xva.upload(h.xva.data(), h.xva.size(), st);
    d.xvm.upload(h.xvm.data(), h.xvm.size(), st);
    d.Jp.upload(h.Jp.data(), h.Jp.size(), st);
    std::vector<int> xcol(h.nnzJ);
    for (int e = 0; e < h.nnzJ; ++e) xcol[e] = h.pinv[h.Jj[e]];
    d.xcol.upload(xcol.data(), xcol.size(), st);
    d.Cp.upload(h.Cp.data(), h.Cp.size(), st);
    d.Ci.upload(h.Ci.data(), h.Ci.size(), st);
    d.Ce.upload(h.Ce.data(), h.Ce.size(), st);
    d.perm.upload(h.perm.data(), h.perm.size(), st);
    d.Lp.upload(h.Lp.data(), h.Lp.size(), st);
    d.Li.upload(h.Li.data(), h.Li.size(), st);
    d.base.upload(h.base.data(), h.base.size(), st);
    HIPCHECK(hipStreamSynchronize(st));
    d.ready = true;
}

void ensure_tiles(p3smi_topology* t);

// Level schedule for a launch shape (TL task lists, dealing group), built and uploaded on
// first use.
// The 2x2-block schedule on the device. Every position a block task or step names is
// checked against the factor and Jacobian sizes here, before any kernel reads it.
std::unique_ptr<DevSched> upload_block_sched(p3smi_topology* t, int TL) {
    const p3smi::Topology& h = t->host;
    const p3smi::BlockSchedule s = p3smi::make_block_schedule(h, TL);
    const int64_t nlu = h.nnz_lu, nj = h.nnzJ;
    auto bad = [](const char* what) { throw Error(P3SMI_E_INTERNAL, std::string("block schedule: ") + what); };
    // steps: positions in the factor arena [0, fac_rows) (nnz_lu = the zero row, nnz_lu + 1 the one
    // row, then the partial slots); the second row / column is -1 exactly when the block has one
    // row / column
    const int64_t nsteps = (int64_t)s.steps.size() / 4;
    if (nlu + 2 + 4 * (int64_t)s.nparts > h.fac_rows) bad("partial slots beyond the factor arena");
    for (int64_t k = 0; k < nsteps; ++k)
        for (int w = 0; w < 4; ++w) {
            const int v = s.steps[4 * k + w];
            if (v >= h.fac_rows || v < ((w & 1) ? -1 : 0)) bad("step position out of range");
        }
    auto pack = [&](const std::vector<p3smi::BlockTask>& ts, const std::vector<int>& ptr) {
        std::vector<int4> out(3 * ts.size());
        for (size_t i = 0; i < ts.size(); ++i) {
            const p3smi::BlockTask& k = ts[i];
            if (k.kind < p3smi::BLK_DIAG1 || k.kind > p3smi::BLK_PART || k.nr < 1 || k.nr > 2 || k.nc < 1 || k.nc > 2)
                bad("task shape");
            if (k.sbeg < 0 || k.send < k.sbeg || k.send > nsteps) bad("step range");
            for (int e = k.sbeg; e < k.send; ++e)
                if ((s.steps[4 * e + 1] < 0) != (k.nr == 1) || (s.steps[4 * e + 3] < 0) != (k.nc == 1)) bad("step shape");
            if (k.kind == p3smi::BLK_PART) {  // writes its partial slot, four rows past the one row
                if (k.tgt0 < nlu + 2 || k.tgt0 + 3 >= h.fac_rows) bad("partial slot");
                for (int j = 0; j < 4; ++j)
                    if (k.src[j] != -1) bad("partial chunk with a Jacobian source");
                out[3 * i] = make_int4(k.kind | (k.nr << 8) | (k.nc << 16), 0, k.sbeg, k.send);
                out[3 * i + 1] = make_int4(k.tgt0, 0, 0, 0);
                out[3 * i + 2] = make_int4(-1, -1, -1, -1);
                continue;
            }
            const bool two = k.kind == p3smi::BLK_DIAG2 || k.kind == p3smi::BLK_U2 || k.kind == p3smi::BLK_L2;
            const int ext = (k.kind == p3smi::BLK_U1 || k.kind == p3smi::BLK_U2) ? k.nc - 1
                            : (k.kind == p3smi::BLK_L1 || k.kind == p3smi::BLK_L2) ? k.nr - 1 : 0;
            if (k.tgt0 < 0 || k.tgt0 + ext >= nlu) bad("target position");
            if (two && (k.tgt1 < 0 || k.tgt1 + ext >= nlu)) bad("target position");
            if (k.kind == p3smi::BLK_DIAG2 && (k.aux0 < 0 || k.aux0 >= nlu || k.aux1 < 0 || k.aux1 >= nlu)) bad("aux");
            if (k.kind >= p3smi::BLK_U2 && (k.aux0 < 0 || k.aux0 + 1 >= nlu)) bad("aux");
            if (k.kind == p3smi::BLK_L2 && (k.aux1 < 0 || k.aux1 >= nlu)) bad("aux");
            for (int j = 0; j < 4; ++j)
                if (k.src[j] >= nj) bad("Jacobian source");
            out[3 * i] = make_int4(k.kind | (k.nr << 8) | (k.nc << 16), 0, k.sbeg, k.send);
            out[3 * i + 1] = make_int4(k.tgt0, k.tgt1, k.aux0, k.aux1);
            out[3 * i + 2] = make_int4(k.src[0], k.src[1], k.src[2], k.src[3]);
        }
        // a list's tasks must have their steps back to back
        for (size_t r = 0; r + 1 < ptr.size(); ++r)
            for (int q = ptr[r]; q + 1 < ptr[r + 1]; ++q)
                if (ts[q].send != ts[q + 1].sbeg) bad("steps not contiguous");
        if (out.empty()) out.push_back(make_int4(0, 0, 0, 0));
        return out;
    };
    const std::vector<int4> ta = pack(s.taskA, s.ptrA), tb = pack(s.taskB, s.ptrB);
    std::unique_ptr<DevSched> d(new DevSched());
    hipStream_t st = t->stream;
    d->TL = TL;
    d->nlev = s.nlev;
    d->nblev = s.nblev;
    d->blk = true;
    d->ptrA.upload(s.ptrA.data(), s.ptrA.size(), st);
    d->ptrB.upload(s.ptrB.data(), s.ptrB.size(), st);
    d->ptrF.upload(s.ptrF.data(), s.ptrF.size(), st);
    d->ptrW.upload(s.ptrW.data(), s.ptrW.size(), st);
    d->diagB.alloc(1);
    d->taskA.upload(ta.data(), ta.size(), st);
    d->taskB.upload(tb.data(), tb.size(), st);
    std::vector<int4> st4(std::max<int64_t>(nsteps, 1), make_int4(0, -1, 0, -1));
    for (int64_t k = 0; k < nsteps; ++k)
        st4[k] = make_int4(s.steps[4 * k], s.steps[4 * k + 1], s.steps[4 * k + 2], s.steps[4 * k + 3]);
    // forward rows by pivot group: {row, nr, sbeg, send}, {rsrc0, rsrc1, lpos, pin}; a piece of
    // a cut row other than the last {slot, nr + 4, sbeg, send}; steps {L(row, j), L(row+1, j)
    // or -1, j, 0}
    const int64_t nfs = (int64_t)s.fsteps.size() / 3;
    const int64_t xr = h.x_rows;  // cut rows' slots: rows [m, x_rows) of the solution vector
    auto slot_ok = [&](int v, int nr) { return v >= h.m && v + nr <= xr; };
    std::vector<int4> fw(2 * std::max<size_t>(s.fwd.size(), 1), make_int4(0, 1, 0, -1));
    std::vector<int4> fs(std::max<int64_t>(nfs, 1), make_int4(0, -1, 0, 0));
    for (int64_t k = 0; k < nfs; ++k) {
        const int la = s.fsteps[3 * k], lb = s.fsteps[3 * k + 1], j = s.fsteps[3 * k + 2];
        if (la < 0 || la > nlu || lb < -1 || lb > nlu || j < 0 || j >= h.m) bad("forward step out of range");
        fs[k] = make_int4(la, lb, j, 0);
    }
    for (size_t i = 0; i < s.fwd.size(); ++i) {
        const p3smi::FwdBlock& f = s.fwd[i];
        if (f.nr < 1 || f.nr > 2 || f.row < 0 || f.row + f.nr > h.m || f.rsrc0 < 0 || f.rsrc0 >= h.m) bad("forward row");
        if (f.nr == 2 && (f.rsrc1 < 0 || f.rsrc1 >= h.m || f.lpos < 0 || f.lpos >= nlu)) bad("forward pair");
        if (f.sbeg < 0 || f.send < f.sbeg || f.send > nfs) bad("forward step range");
        for (int e = f.sbeg; e < f.send; ++e)
            if ((s.fsteps[3 * e + 1] < 0) != (f.nr == 1)) bad("forward step shape");
        if (f.pin != -1 && !slot_ok(f.pin, f.nr)) bad("forward slot");
        if (f.pout != -1 && !slot_ok(f.pout, f.nr)) bad("forward slot");
        fw[2 * i] = f.pout >= 0 ? make_int4(f.pout, f.nr + 4, f.sbeg, f.send) : make_int4(f.row, f.nr, f.sbeg, f.send);
        fw[2 * i + 1] = make_int4(f.rsrc0, f.nr == 2 ? f.rsrc1 : 0, f.nr == 2 ? f.lpos : 0, f.pin);
    }
    for (size_t r = 0; r + 1 < s.ptrF.size(); ++r)
        for (int q = s.ptrF[r]; q + 1 < s.ptrF[r + 1]; ++q)
            if (s.fwd[q].send != s.fwd[q + 1].sbeg) bad("forward steps not contiguous");
    // backward rows by pivot group: {row, nr, sbeg, send}, {base0, base1, pin, 0}; pieces of cut
    // rows as the forward ones; steps {U(row, j), U(row+1, j) or -1, j, 0}
    const int64_t nbs = (int64_t)s.bsteps.size() / 3;
    std::vector<int4> bw(2 * std::max<size_t>(s.bwd.size(), 1), make_int4(0, 1, -1, 0));
    std::vector<int4> bs4(std::max<int64_t>(nbs, 1), make_int4(0, -1, 0, 0));
    for (int64_t k = 0; k < nbs; ++k) {
        const int ua = s.bsteps[3 * k], ub = s.bsteps[3 * k + 1], j = s.bsteps[3 * k + 2];
        if (ua < 0 || ua >= nlu || ub < -1 || ub >= nlu || j < 0 || j >= h.m) bad("backward step out of range");
        bs4[k] = make_int4(ua, ub, j, 0);
    }
    for (size_t i = 0; i < s.bwd.size(); ++i) {
        const p3smi::BwdBlock& w = s.bwd[i];
        if (w.nr < 1 || w.nr > 2 || w.row < 0 || w.row + w.nr > h.m || w.base0 < 0 || w.base0 + 1 >= nlu + 1)
            bad("backward row");
        if (w.nr == 2 && (w.base1 < 0 || w.base1 >= nlu || w.base0 + 1 >= nlu)) bad("backward pair");
        if (w.sbeg < 0 || w.send < w.sbeg || w.send > nbs) bad("backward step range");
        for (int e = w.sbeg; e < w.send; ++e)
            if ((s.bsteps[3 * e + 1] < 0) != (w.nr == 1)) bad("backward step shape");
        if (w.pin != -1 && !slot_ok(w.pin, w.nr)) bad("backward slot");
        if (w.pout != -1 && !slot_ok(w.pout, w.nr)) bad("backward slot");
        bw[2 * i] = w.pout >= 0 ? make_int4(w.pout, w.nr + 4, w.sbeg, w.send) : make_int4(w.row, w.nr, w.sbeg, w.send);
        bw[2 * i + 1] = make_int4(w.base0, w.nr == 2 ? w.base1 : 0, w.pin, 0);
    }
    for (size_t r = 0; r + 1 < s.ptrW.size(); ++r)
        for (int q = s.ptrW[r]; q + 1 < s.ptrW[r + 1]; ++q)
            if (s.bwd[q].send != s.bwd[q + 1].sbeg) bad("backward steps not contiguous");
    d->pairs.alloc(1);
    d->steps.upload(st4.data(), st4.size(), st);
    d->fsteps.upload(fs.data(), fs.size(), st);
    d->bsteps.upload(bs4.data(), bs4.size(), st);
    d->fent.alloc(1);
    d->bent.alloc(1);
    d->fwd.upload(fw.data(), fw.size(), st);
    d->bwd.upload(bw.data(), bw.size(), st);
    HIPCHECK(hipStreamSynchronize(st));
    return d;
}

const DevSched& get_sched(p3smi_topology* t, LuConfig cfg) {
    const int TL = cfg.tl();
    auto it = t->dev.sched.find(cfg.key());
    if (it != t->dev.sched.end()) return *it->second;
    const p3smi::Topology& h = t->host;
    if (cfg.blk()) return *(t->dev.sched[cfg.key()] = upload_block_sched(t, TL));
    const p3smi::Schedule s = p3smi::make_schedule(h, TL, cfg.grp);
    std::unique_ptr<DevSched> d(new DevSched());
    hipStream_t st = t->stream;
    d->TL = TL;
    d->nlev = s.nlev;
    d->nblev = s.nblev;
    d->ptrA.upload(s.ptrA.data(), s.ptrA.size(), st);
    d->ptrB.upload(s.ptrB.data(), s.ptrB.size(), st);
    d->ptrF.upload(s.ptrF.data(), s.ptrF.size(), st);
    d->ptrW.upload(s.ptrW.data(), s.ptrW.size(), st);
    d->diagB.upload(s.diagB.data(), s.diagB.size(), st);
    // Operand streams: the entries every task reads, laid out in execution order (tasks of
    // one list and sub-level back to back), so that a task's range ends where the next begins.
    std::vector<int4> ta(s.taskA.size()), tb(s.taskB.size()), fw(s.fwd.size()), bw(s.bwd.size());
    std::vector<int2> pr, fe, be;
    pr.reserve((size_t)h.n_pairs);
    const int64_t one = h.nnz_lu + 1;
    auto put_pairs = [&](const p3smi::Task& k) {
        const int b0 = (int)pr.size();
        if (k.pin >= 0) {  // the previous chunk's sum, negated, times the one row
            if (k.pin < one + 1 || k.pin >= h.fac_rows) throw Error(P3SMI_E_INTERNAL, "schedule: partial slot");
            pr.push_back(make_int2(k.pin, (int)one));
        }
        for (int e = k.pbeg; e < k.pend; ++e) pr.push_back(make_int2(h.pa[e], h.pb[e]));
        return std::make_pair(b0, (int)pr.size());
    };
    auto dst_of = [&](const p3smi::Task& k) {
        if (k.part ? (k.dst < one + 1 || k.dst >= h.fac_rows) : (k.dst < 0 || k.dst >= h.nnz_lu))
            throw Error(P3SMI_E_INTERNAL, "schedule: target position");
        return k.dst;
    };
    for (size_t i = 0; i < ta.size(); ++i) {
        const auto& k = s.taskA[i];
        int src = k.asrc;
        if (k.part) {
            src = LU_PART;
        } else if (s.isdiagA[i]) {
            if (src < 0) throw Error(P3SMI_E_INTERNAL, "pivot without a Jacobian entry");
            src = -src - 2;
        }
        const auto r = put_pairs(k);
        ta[i] = make_int4(dst_of(k), src, r.first, r.second);
    }
    for (size_t i = 0; i < tb.size(); ++i) {
        const auto& k = s.taskB[i];
        const auto r = put_pairs(k);
        tb[i] = make_int4(dst_of(k), k.part ? LU_PART : k.asrc, r.first, r.second);
    }
    // a cut row's chunks pass their sums on, negated, through solution-vector slots [m, x_rows):
    // (one row, slot) brings the previous chunk's back exactly
    auto xslot = [&](int v) {
        if (v < h.m || v >= h.x_rows) throw Error(P3SMI_E_INTERNAL, "schedule: solution-vector slot");
        return v;
    };
    for (size_t i = 0; i < fw.size(); ++i) {
        const auto& f = s.fwd[i];
        const int b0 = (int)fe.size();
        if (f.pin >= 0) fe.push_back(make_int2((int)one, xslot(f.pin)));
        for (int e = f.ebeg; e < f.eend; ++e) fe.push_back(make_int2(h.Fpos[e], h.Fcol[e]));
        fw[i] = f.dst >= 0 ? make_int4(xslot(f.dst), LU_PART, b0, (int)fe.size())
                           : make_int4(f.row, f.rsrc, b0, (int)fe.size());
    }
    for (size_t i = 0; i < bw.size(); ++i) {
        const auto& w = s.bwd[i];
        const int b0 = (int)be.size();
        // columns descending: every backward row sums in this order (the 2x2-block kernel adds a
        // pair's own term U(p, p+1) x(p+1) last); a chunk walks its part of the row
        if (w.pin >= 0) be.push_back(make_int2((int)one, xslot(w.pin)));
        const int l0 = h.Lp[w.row];
        for (int e = w.lend - 1; e >= w.lbeg; --e) be.push_back(make_int2(w.base + 1 + (e - l0), h.Li[e]));
        bw[i] = w.dst >= 0 ? make_int4(xslot(w.dst), LU_PART, b0, (int)be.size())
                           : make_int4(w.row, w.base, b0, (int)be.size());
    }
    if (pr.empty()) pr.push_back(make_int2(0, 0));  // keep every stream a valid pointer
    if (fe.empty()) fe.push_back(make_int2(0, 0));
    if (be.empty()) be.push_back(make_int2(0, 0));
    d->pairs.upload(pr.data(), pr.size(), st);
    d->fent.upload(fe.data(), fe.size(), st);
    d->bent.upload(be.data(), be.size(), st);
    d->taskA.upload(ta.data(), ta.size(), st);
    d->taskB.upload(tb.data(), tb.size(), st);
    d->fwd.upload(fw.data(), fw.size(), st);
    d->bwd.upload(bw.data(), bw.size(), st);
    HIPCHECK(hipStreamSynchronize(st));
    return *(t->dev.sched[cfg.key()] = std::move(d));
}

// Clear n elements with a kernel (no memset node if a caller captures the stream).
template <typename T>
void zero_row(T* p, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
    HIPCHECK(hipGetLastError());
}

// Set n elements to 1 (the factor arena's one row).
template <typename T>
void one_row(T* p, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_one<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
    HIPCHECK(hipGetLastError());
}

// Row tiles of k_evaluate: consecutive rows, at most `rows` of them and at most EV_NZ Ybus
// nonzeros, so that a tile's CSR segment fits the kernel's LDS (a single row with more
// nonzeros is a tile of its own and read from global memory).
std::vector<int> eval_tiles(const p3smi::Topology& h, int rows) {
    std::vector<int> ptr{0};
    int r0 = 0;
    for (int r = 0; r < h.n; ++r)
        if (r > r0 && (r + 1 - r0 > rows || h.Yp[r + 1] - h.Yp[r0] > EV_NZ)) {
            ptr.push_back(r);
            r0 = r;
        }
    ptr.push_back(h.n);
    return ptr;
}

void ensure_tiles(p3smi_topology* t) {
    DevTopo& d = t->dev;
    if (d.ntiles > 0 && d.tile_rows == t->eval_rows) return;
    const std::vector<int> ptr = eval_tiles(t->host, t->eval_rows);
    d.tiles.upload(ptr.data(), ptr.size(), t->stream);
    HIPCHECK(hipStreamSynchronize(t->stream));
    d.ntiles = (int)ptr.size() - 1;
    d.tile_rows = t->eval_rows;
}

// rows per tile the evaluation may use at most (its partial maxima are sized for the most tiles)
int max_tiles(const p3smi::Topology& h) { return (int)eval_tiles(h, 1).size() - 1; }

void batch_alloc(p3smi_batch* b, p3smi_topology* t, int cases) {
    ensure_device(t);
    ensure_tiles(t);
    const p3smi::Topology& h = t->host;
    b->topo = t;
    b->B = cases;
    b->S = (cases + 63) / 64 * 64;
    const size_t S = b->S;
    for (auto* v : {&b->Vm, &b->Va, &b->Ve, &b->Vf, &b->Vm0, &b->Va0}) v->alloc((size_t)h.n * S);
    b->Jx.alloc((size_t)h.nnzJ * S);
    // the factor arena: + the zero row and the one row the 2x2-block steps read, + partial slots
    b->fac.alloc((size_t)h.fac_rows * S);
    zero_row(b->fac.p + (size_t)h.nnz_lu * S, S, t->stream);
    one_row(b->fac.p + ((size_t)h.nnz_lu + 1) * S, S, t->stream);
    b->R.alloc((size_t)h.m * S);
    b->x.alloc((size_t)h.x_rows * S);  // + the slots of cut substitution rows
    // compact arenas of one-case FP64 launches (LuCompact): one per case, at most
    // LU_COMPACT_SLOTS (a launch with more cases runs the others in the batch's rows)
    b->cmp_cap = std::min(cases, LU_COMPACT_SLOTS);
    b->cmp_fac.alloc((size_t)h.fac_rows * (size_t)b->cmp_cap);
    b->cmp_x.alloc((size_t)h.x_rows * (size_t)b->cmp_cap);
    b->cmp_ctr.alloc(1);
    hipLaunchKernelGGL(k_compact_rows, dim3((unsigned)((b->cmp_cap + 63) / 64)), dim3(64), 0, t->stream, b->cmp_fac.p,
                       b->cmp_cap, (long long)h.fac_rows, (long long)h.nnz_lu);
    for (auto* v : {&b->nrm, &b->alpha, &b->best_alpha, &b->best_nrm}) v->alloc(S);
    b->part.alloc((size_t)t->dev.ntiles * S);
    b->pstat.alloc(2 * S);
    b->ctl.alloc(S);
    b->tally.alloc(NCOUNT + 1);
    b->counts.alloc(NCOUNT);
    // guard trips, FP32 and FP64 factorizations of the switching mode, reruns, pivoting factorizations
    b->guard.alloc(6);
    b->rcount.alloc(4);  // refining, fallbacks, FP64 classes, FP32 pivot failures before
    hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, t->stream, b->tally.p, NCOUNT + 1);
    hipLaunchKernelGGL(k_ctl_reset, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, t->stream, b->ctl.p, 0,
                       (int)S, 0u, 0u);
    HIPCHECK(hipGetLastError());
    b->mixed_ready = false;
    if (!b->h_counts) HIPCHECK(hipHostMalloc((void**)&b->h_counts, NCOUNT * sizeof(int), hipHostMallocDefault));
    if (!b->h_rcount) HIPCHECK(hipHostMalloc((void**)&b->h_rcount, 4 * sizeof(int), hipHostMallocDefault));
    std::memset(b->h_counts, 0, NCOUNT * sizeof(int));
    b->h_counts[CNT_IDLE] = (int)S;
    b->piv.h_slot_of.assign(S, -1);
    b->piv.free.clear();
    for (int k = b->piv.slots - 1; k >= 0; --k) b->piv.free.push_back(k);
    HIPCHECK(hipStreamSynchronize(t->stream));
}

// kernel timing (optional): kinds 0 eval, 1 lu, 2 step, 3 decide, 4 other, 5 refine, 6 lu of the
// batch shape in FP64
struct KTimer {
    p3smi_batch* b;
    int kind;
    hipEvent_t stop = nullptr;
    KTimer(p3smi_batch* bb, int k) : b(bb), kind(k) {
        if (!b->timing) return;
        if (b->ev_used == b->ev.size()) {
            hipEvent_t e0, e1;
            HIPCHECK(hipEventCreate(&e0));
            HIPCHECK(hipEventCreate(&e1));
            b->ev.push_back({k, {e0, e1}});
        }
        auto& slot = b->ev[b->ev_used++];
        slot.first = k;
        HIPCHECK(hipEventRecord(slot.second.first, b->topo->stream));
        stop = slot.second.second;
    }
    ~KTimer() {
        if (stop) (void)hipEventRecord(stop, b->topo->stream);
    }
};

void collect_timing(p3smi_batch* b) {
    if (!b->timing) return;
    HIPCHECK(hipStreamSynchronize(b->topo->stream));
    for (size_t i = 0; i < b->ev_used; ++i) {
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, b->ev[i].second.first, b->ev[i].second.second));
        switch (b->ev[i].first) {
            case 0: b->stats.ms_eval += ms; b->stats.n_eval++; break;
            case 1: b->stats.ms_lu += ms; b->stats.n_lu++; break;
            case 6:
                b->stats.ms_lu += ms;
                b->stats.n_lu++;
                b->stats.ms_lu_batch += ms;
                b->stats.n_lu_batch++;
                break;
            case 2: b->stats.ms_step += ms; b->stats.n_step++; break;
            case 3: b->stats.ms_decide += ms; b->stats.n_decide++; break;
            case 5: b->stats.ms_refine += ms; b->stats.n_refine++; break;
            default: b->stats.ms_other += ms; break;
        }
    }
    b->ev_used = 0;
}

dim3 grid_items(int items, int S) { return dim3((unsigned)(S / 64), (unsigned)((items + 3) / 4)); }

EvalArgs eval_args(p3smi_batch* b, int hold) {
    p3smi_topology* t = b->topo;
    const p3smi::Topology& h = t->host;
    const DevTopo& d = t->dev;
    EvalArgs a;
    a.tile_ptr = d.tiles.p;
    a.Yp = d.Yp.p;
    a.Yj = d.Yj.p;
    a.Ydiag = d.Ydiag.p;
    a.posJ = reinterpret_cast<const int4*>(d.posJ.p);
    a.G = b->per_case_y ? b->Yg.p : d.G.p;
    a.B = b->per_case_y ? b->Yb.p : d.B.p;
    a.dptr = b->ydptr;
    a.dpos = b->ydpos;
    a.dval = b->ydval;
    a.dc0 = b->ydc0;
    a.Ve = b->Ve.p;
    a.Vf = b->Vf.p;
    a.Ssp = b->Ssp.p;
    a.s_stride = b->s_stride;
    a.row_p = d.row_p.p;
    a.row_q = d.row_q.p;
    a.pin = b->has_pin ? b->pin.p : nullptr;
    a.ctl = b->ctl.p;
    a.hold = hold;
    a.J = b->Jx.p;
    a.R = b->R.p;
    a.part = b->part.p;
    a.n = h.n;
    a.S = b->S;
    a.gx = d.ntiles;
    a.plan = nullptr;
    return a;
}

void launch_evaluate(p3smi_batch* b, int hold) {
    p3smi_topology* t = b->topo;
    const DevTopo& d = t->dev;
    if (b->part.n < (size_t)d.ntiles * b->S) b->part.alloc((size_t)d.ntiles * b->S);
    KTimer tm(b, 0);
    const EvalArgs a = eval_args(b, hold);
    const dim3 grid((unsigned)d.ntiles, (unsigned)(b->S / 64));
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, grid, dim3(64 * EV_WAVES), 0, t->stream, a); };
    switch (2 * b->ymode() + (int)b->has_pin) {
        case 0: go(k_evaluate<YM_SHARED, false>); break;
        case 1: go(k_evaluate<YM_SHARED, true>); break;
        case 2: go(k_evaluate<YM_CASE, false>); break;
        case 3: go(k_evaluate<YM_CASE, true>); break;
        case 4: go(k_evaluate<YM_DELTA, false>); break;
        default: go(k_evaluate<YM_DELTA, true>); break;
    }
    HIPCHECK(hipGetLastError());
}

// Cases per workgroup of a 2x2-block LU launch: 48 FP64 cases (rows of three 128-byte lines)
// while the batch's groups of 48 fit one per compute unit, else 64. A workgroup's time is set
// by its chain of dependent loads and hardly depends on the width of its rows; narrower rows
// over more compute units ease the load on the memory system (2,000-bus grid, 8,760 cases:
// 183 workgroups, 3.84 -> 3.76 ms per launch). 32 cases on that batch need 274 workgroups,
// more than the 256 compute units: 6.3 ms. FP32 rows keep 64 cases (48 would start every
// other row mid-line).
template <typename real>
int blk_cases(const p3smi_batch* b) {
    if (!std::is_same<real, double>::value) return 64;
    return (b->S + 47) / 48 <= b->topo->n_cu ? 48 : 64;
}

template <int G, int NS, typename real, bool FACTOR>
void launch_lu_cfg(p3smi_batch* b, const DevSched& ds, real* fac, const double* rhs, double* xout, int sel) {
    const p3smi::Topology& h = b->topo->host;
    const double rcond_min = (b->guard_override >= 0.0 ? b->guard_override : b->topo->guard_ratio) * h.rcond0;
    if constexpr (G == 64) {
        if (ds.blk) {
            const int cpg = blk_cases<real>(b);
            hipLaunchKernelGGL((k_lu_blk<NS, real, FACTOR>), dim3((unsigned)((b->S + cpg - 1) / cpg)), dim3(64 * NS),
                               0, b->topo->stream, ds.ptrA.p, ds.taskA.p, ds.ptrB.p, ds.taskB.p, ds.steps.p,
                               ds.ptrF.p, ds.fwd.p, ds.fsteps.p, ds.ptrW.p, ds.bwd.p, ds.bsteps.p, b->Jx.p, fac, rhs,
                               xout, b->ctl.p, b->pstat.p, b->guard.p, sel, rcond_min, ds.nlev, ds.nblev, b->S, cpg);
            return;
        }
    }
    if (ds.blk) throw Error(P3SMI_E_INTERNAL, "2x2-block schedule launched with G != 64");
    LuCompact cmp;  // one-case FP64 factor + solve whose factors nothing reads later: compact arenas
    if constexpr (G == 1 && FACTOR && std::is_same<real, double>::value) {
        if ((sel == SEL_ALL || sel == SEL_F64) && b->cmp_cap > 0 && xout == b->x.p) {
            hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, b->topo->stream, b->cmp_ctr.p, 1);
            cmp.fac = b->cmp_fac.p;
            cmp.x = b->cmp_x.p;
            cmp.ctr = b->cmp_ctr.p;
            cmp.cap = b->cmp_cap;
            cmp.frows = h.fac_rows;
            cmp.xrows = h.x_rows;
            cmp.m = h.m;
        }
    }
    hipLaunchKernelGGL((k_lu_stream<G, NS, real, FACTOR>), dim3((unsigned)(b->S / G)), dim3(64 * NS), 0,
                       b->topo->stream, ds.ptrA.p, ds.taskA.p, ds.ptrB.p, ds.taskB.p, ds.diagB.p, ds.pairs.p,
                       ds.ptrF.p, ds.fwd.p, ds.fent.p, ds.ptrW.p, ds.bwd.p, ds.bent.p, b->Jx.p, fac, rhs, xout,
                       b->ctl.p, b->pstat.p, b->guard.p, sel, rcond_min, ds.nlev, ds.nblev, b->S, cmp);
}

// The arguments of an LU launch of shape cfg over batch b (factor + solve of the selected cases).
LuArgs lu_args(p3smi_batch* b, LuConfig cfg, const DevSched& ds, void* fac, const double* rhs, double* xout, int sel) {
    const p3smi::Topology& h = b->topo->host;
    LuArgs a{};
    a.ptrA = ds.ptrA.p;
    a.taskA = ds.taskA.p;
    a.ptrB = ds.ptrB.p;
    a.taskB = ds.taskB.p;
    a.diagB = ds.diagB.p;
    a.pairs = ds.pairs.p;
    a.steps = ds.steps.p;
    a.ptrF = ds.ptrF.p;
    a.fwd = ds.fwd.p;
    a.fent = ds.fent.p;
    a.fsteps = ds.fsteps.p;
    a.ptrW = ds.ptrW.p;
    a.bwd = ds.bwd.p;
    a.bent = ds.bent.p;
    a.bsteps = ds.bsteps.p;
    a.Jx = b->Jx.p;
    a.fac = fac;
    a.rhs = rhs;
    a.x = xout;
    a.ctl = b->ctl.p;
    a.pstat = b->pstat.p;
    a.guard = b->guard.p;
    a.sel = sel;
    a.rcond_min = (b->guard_override >= 0.0 ? b->guard_override : b->topo->guard_ratio) * h.rcond0;
    a.nlev = ds.nlev;
    a.nblev = ds.nblev;
    a.S = b->S;
    a.gx = b->S / cfg.G;
    a.plan = nullptr;
    a.want = 0;
    return a;
}

bool valid_lu_config(LuConfig c) {
    const bool g = c.G == 64 || c.G == 32 || c.G == 16 || c.G == 1;
    const bool n = c.NS == 8 || c.NS == 16 || (c.NS == 4 && c.G == 1);
    return g && n;
}

LuConfig lu_config_for(const p3smi_topology* t, int n_active) {
    return n_active <= t->lu_few_max ? t->lu_few : t->lu_batch;
}

// One LU launch over the selected cases; n_active picks the batch or the few-cases shape.
template <typename real, bool FACTOR>
void launch_lu(p3smi_batch* b, int n_active, real* fac, const double* rhs, double* xout, int sel) {
    p3smi_topology* t = b->topo;
    const LuConfig cfg = lu_config_for(t, n_active);
    const DevSched& ds = get_sched(t, cfg);
    // FP64 factor + solve launches of the batch shape over every stepping case are timed apart
    // (kind 6): the dominant kernel of a whole-batch FP64 solve, with their cases counted
    const bool batch64 = FACTOR && std::is_same<real, double>::value && sel == SEL_ALL && n_active > t->lu_few_max;
    if (batch64 && b->timing) b->stats.lu_batch_cases += n_active;
    KTimer tm(b, batch64 ? 6 : 1);
#define P3SMI_LU_CASE(GG, NN)                                                                \
    if (cfg.G == GG && cfg.NS == NN) {                                                       \
        launch_lu_cfg<GG, NN, real, FACTOR>(b, ds, fac, rhs, xout, sel);                     \
        HIPCHECK(hipGetLastError());                                                         \
        return;                                                                              \
    }
#define P3SMI_LU_G(GG) P3SMI_LU_CASE(GG, 8) P3SMI_LU_CASE(GG, 16)
    P3SMI_LU_G(64) P3SMI_LU_G(32) P3SMI_LU_G(16) P3SMI_LU_G(1) P3SMI_LU_CASE(1, 4)
#undef P3SMI_LU_G
#undef P3SMI_LU_CASE
    throw Error(P3SMI_E_ARG, "unsupported LU launch shape");
}

void ensure_mixed(p3smi_batch* b) {
    if (b->mixed_ready) return;
    const p3smi::Topology& h = b->topo->host;
    const size_t S = b->S;
    b->fac32.alloc((size_t)h.fac_rows * S);  // + the zero and one rows, partial slots
    zero_row(b->fac32.p + (size_t)h.nnz_lu * S, S, b->topo->stream);
    one_row(b->fac32.p + ((size_t)h.nnz_lu + 1) * S, S, b->topo->stream);
    b->rr.alloc((size_t)h.m * S);
    b->dvec.alloc((size_t)h.x_rows * S);  // an LU solution too: + the slots of cut rows
    b->prevd.alloc(S);
    b->nparts_r = (h.m + 63) / 64;
    b->rpart.alloc(2 * (size_t)b->nparts_r * S);
    b->mixed_ready = true;
}

// Newton step of the cases in STEP. FP64: one factor+solve launch. Mixed: FP32 factors and
// solve, then refinement x += LU32^-1 (R - J x) with the residual in FP64 until the correction
// is below rtol |x|; cases whose FP32 pivots fail or whose refinement stalls are factored in
// FP64 in the same iteration.
void linear_solve(p3smi_batch* b, int n_active, int precision) {
    p3smi_topology* t = b->topo;
    const p3smi::Topology& h = t->host;
    hipStream_t st = t->stream;
    if (precision == P3SMI_PREC_DOUBLE) {
        b->stats.double_factorizations += n_active;
        launch_lu<double, true>(b, n_active, b->fac.p, b->R.p, b->x.p, SEL_ALL);
        return;
    }
    if (precision == P3SMI_PREC_MIXED_SWITCH) {
        // FP32 factors and solve for the 64-case groups of FP32 cases alone, no refinement; FP64
        // for the others and for the cases whose FP32 pivots failed, in the same iteration. The
        // factorization counts are kept on the device (guard[1], guard[2]).
        ensure_mixed(b);
        {
            KTimer tm(b, 4);
            hipLaunchKernelGGL(k_group_vote, dim3((unsigned)((b->S + 255) / 256)), dim3(256), 0, st, b->ctl.p,
                               b->guard.p, b->S);
            HIPCHECK(hipGetLastError());
        }
        launch_lu<float, true>(b, n_active, b->fac32.p, b->R.p, b->x.p, SEL_F32);
        launch_lu<double, true>(b, n_active, b->fac.p, b->R.p, b->x.p, SEL_F64);
        return;
    }
    ensure_mixed(b);
    const unsigned gS = (unsigned)((b->S + 255) / 256);
    {
        KTimer tm(b, 4);
        hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, st, b->rcount.p, 4);
        hipLaunchKernelGGL(k_refine_init, dim3(gS), dim3(256), 0, st, b->ctl.p, b->prevd.p, b->rcount.p, b->S);
        HIPCHECK(hipGetLastError());
    }
    // rcount[3] = FP32 pivot failures so far (guard[2]), to tell this iteration's apart
    HIPCHECK(hipMemcpyAsync(b->rcount.p + 3, b->guard.p + 2, sizeof(int), hipMemcpyDeviceToDevice, st));
    launch_lu<float, true>(b, n_active, b->fac32.p, b->R.p, b->x.p, SEL_F32);
    for (int k = 1; k <= t->mixed_kmax; ++k) {
        {
            KTimer tm(b, 5);
            hipLaunchKernelGGL(k_resid, dim3((unsigned)(b->S / 64), (unsigned)((h.m + 3) / 4)), dim3(256), 0, st,
                               t->dev.Jp.p, t->dev.xcol.p, b->Jx.p, b->R.p, b->x.p, b->rr.p, b->ctl.p, h.m, b->S);
        }
        launch_lu<float, false>(b, n_active, b->fac32.p, b->rr.p, b->dvec.p, SEL_REFINE);
        {
            KTimer tm(b, 5);
            hipLaunchKernelGGL(k_refine_update, dim3((unsigned)b->nparts_r, (unsigned)(b->S / 64)), dim3(256), 0, st,
                               b->x.p, b->dvec.p, b->rpart.p, b->ctl.p, h.m, b->S);
            hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, st, b->rcount.p, 1);  // [0] only
            hipLaunchKernelGGL(k_refine_decide, dim3(gS), dim3(256), 0, st, b->rpart.p, b->nparts_r, b->prevd.p,
                               b->ctl.p, b->rcount.p, t->mixed_rtol, (int)(k == t->mixed_kmax), b->S);
            HIPCHECK(hipGetLastError());
        }
        b->stats.refinement_steps += 1;
        HIPCHECK(hipMemcpyAsync(b->h_rcount, b->rcount.p, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        if (b->h_rcount[0] == 0) break;
    }
    int g2[2];
    HIPCHECK(hipMemcpyAsync(&g2[0], b->rcount.p + 3, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&g2[1], b->guard.p + 2, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(b->h_rcount, b->rcount.p, 3 * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    const int fp64_class = b->h_rcount[2], fallbacks = b->h_rcount[1];
    // FP64 factorizations of this iteration: the FP64 classes, stalled refinements, failed FP32 pivots
    const int to_fp64 = fp64_class + fallbacks + (g2[1] - g2[0]);
    b->stats.single_factorizations += n_active - fp64_class;
    b->stats.double_factorizations += fp64_class + fallbacks;  // failed FP32 pivots: guard[2], at the end
    // in the launch shape for their number
    if (to_fp64 > 0) launch_lu<double, true>(b, to_fp64, b->fac.p, b->R.p, b->x.p, SEL_F64);
}

void check_solve_args(int max_iter, double tol, int precision) {
    if (max_iter <= 0) throw Error(P3SMI_E_ARG, "max_iter must be > 0");
    if (max_iter > MAX_ITER_LIMIT) throw Error(P3SMI_E_ARG, "max_iter must be <= 65535");
    if (!(tol > 0.0)) throw Error(P3SMI_E_ARG, "tol must be > 0");
    if (precision != P3SMI_PREC_DOUBLE && precision != P3SMI_PREC_MIXED && precision != P3SMI_PREC_MIXED_SWITCH)
        throw Error(P3SMI_E_ARG, "precision must be 0 (double), 1 (mixed) or 2 (mixed_switch)");
}

// ---- pivoting factorizations ------------------------------------------------------------
// Control bits a case starts with: none on a usable frozen plan; otherwise it factors with
// pivoting from its first iteration and has no rerun (nothing frozen was reused,
// p3s/cpp/nr_klu.cpp:844-857).
unsigned start_extra(const p3smi_topology* t) { return t->host.lean_valid ? 0u : (set_piv(0u, PV_NEED) | CTL_RERUN); }

// rows x width array -> rows' x width' keeping every row's leading min(width, width') entries
template <typename T>
void regrow(DevBuf<T>& buf, size_t rows, size_t width, size_t rows2, size_t width2, hipStream_t st) {
    DevBuf<T> nb;
    nb.alloc(rows2 * width2);
    if (buf.p && rows && width)
        HIPCHECK(hipMemcpy2DAsync(nb.p, width2 * sizeof(T), buf.p, width * sizeof(T), std::min(width, width2) * sizeof(T),
                                  std::min(rows, rows2), hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipStreamSynchronize(st));
    std::swap(buf.p, nb.p);
    std::swap(buf.n, nb.n);
}

// At least `slots` slots of `cap` factor entries each; existing slots keep their factors.
void pivot_pool_ensure(p3smi_batch* b, int slots, long long cap) {
    PivotPool& P = b->piv;
    const size_t m = (size_t)b->topo->host.m;
    hipStream_t st = b->topo->stream;
    if (P.slots > 0 && P.m != (int)m) throw Error(P3SMI_E_INTERNAL, "pivot pool: dimension changed");
    if (slots <= P.slots && cap <= P.cap) return;
    const size_t s0 = P.slots, s1 = std::max(slots, P.slots);
    const size_t c0 = (size_t)P.cap, c1 = (size_t)std::max(cap, P.cap);
    for (auto* v : {&P.Lp, &P.Up}) regrow(*v, s0, m + 1, s1, m + 1, st);
    for (auto* v : {&P.P, &P.pinv, &P.xi, &P.stk, &P.pstk, &P.flag}) regrow(*v, s0, m, s1, m, st);
    for (auto* v : {&P.Ud, &P.rs, &P.w, &P.y}) regrow(*v, s0, m, s1, m, st);
    for (auto* v : {&P.Li, &P.Ui}) regrow(*v, s0, c0, s1, c1, st);
    for (auto* v : {&P.Lx, &P.Ux}) regrow(*v, s0, c0, s1, c1, st);
    for (size_t k = s1; k-- > s0;) P.free.push_back((int)k);
    P.slots = (int)s1;
    P.cap = (long long)c1;
    P.m = (int)m;
}

void pivot_release(p3smi_batch* b, int c) {
    PivotPool& P = b->piv;
    if (c >= 0 && c < (int)P.h_slot_of.size() && P.h_slot_of[c] >= 0) {
        P.free.push_back(P.h_slot_of[c]);
        P.h_slot_of[c] = -1;
    }
}

void pivot_release_all(p3smi_batch* b) {
    PivotPool& P = b->piv;
    P.h_slot_of.assign(b->S, -1);
    P.free.clear();
    for (int k = P.slots - 1; k >= 0; --k) P.free.push_back(k);
}

// The pivoting solves of this pass: every case in STEP factored by the pivoting LU gets its step
// here (slots assigned here; a pool too small for a fresh factorization grows and the case is
// factored again). A case that needs a fresh factorization first tries the plan's diagonal
// (an SEL_NEED static-pivot launch checked by k_pivot_verify); the others, and those the check
// rejects, go to k_pivot_solve. Returns the number of cases solved.
int pivot_solve(p3smi_batch* b) {
    p3smi_topology* t = b->topo;
    const p3smi::Topology& h = t->host;
    hipStream_t st = t->stream;
    PivotPool& P = b->piv;
    if (P.list.n < (size_t)b->S) {
        P.list.alloc(b->S);
        P.res.alloc(b->S);
        P.slot_of.alloc(b->S);
    }
    P.count.alloc(1);
    if (P.h_slot_of.size() != (size_t)b->S) pivot_release_all(b);
    hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, st, P.count.p, 1);
    hipLaunchKernelGGL(k_pivot_list, dim3((unsigned)((b->S + 255) / 256)), dim3(256), 0, st, b->ctl.p, b->S, P.list.p,
                       P.count.p);
    HIPCHECK(hipGetLastError());
    int n = 0;
    HIPCHECK(hipMemcpyAsync(&n, P.count.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (n <= 0) return 0;
    if (n > b->S) throw Error(P3SMI_E_INTERNAL, "pivot list overflow");
    std::vector<int> list(n);
    HIPCHECK(hipMemcpyAsync(list.data(), P.list.p, n * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    int lacking = 0;
    for (int c : list) lacking += P.h_slot_of[c] < 0;
    if (lacking > (int)P.free.size()) {
        const int want = P.slots + lacking - (int)P.free.size();
        pivot_pool_ensure(b, (want + 15) / 16 * 16, std::max<long long>(P.cap, std::max<long long>(h.nnz_lu, h.m)));
    }
    for (int c : list)
        if (P.h_slot_of[c] < 0) {
            P.h_slot_of[c] = P.free.back();
            P.free.pop_back();
        }
    HIPCHECK(hipMemcpyAsync(P.slot_of.p, P.h_slot_of.data(), b->S * sizeof(int), hipMemcpyHostToDevice, st));
    if (t->pivot_diagonal_first) {
        launch_lu<double, true>(b, n, b->fac.p, b->R.p, b->x.p, SEL_NEED);
        KTimer tm(b, 1);
        VerifyArgs v;
        v.Jp = t->dev.Jp.p;
        v.perm = t->dev.perm.p;
        v.Lp = t->dev.Lp.p;
        v.Li = t->dev.Li.p;
        v.base = t->dev.base.p;
        v.J = b->Jx.p;
        v.fac = b->fac.p;
        v.ctl = b->ctl.p;
        v.list = P.list.p;
        v.slot_of = P.slot_of.p;
        v.rs = P.rs.p;
        v.guard = b->guard.p;
        v.m = h.m;
        v.S = b->S;
        hipLaunchKernelGGL(k_pivot_verify, dim3((unsigned)n), dim3(256), 0, st, v);
        HIPCHECK(hipGetLastError());
    }
    KTimer tm(b, 1);
    std::vector<int> res(n), diagonal;
    for (;;) {
        PivArgs a;
        a.Cp = t->dev.Cp.p;
        a.Ci = t->dev.Ci.p;
        a.Ce = t->dev.Ce.p;
        a.Jp = t->dev.Jp.p;
        a.Q = t->dev.perm.p;
        a.m = h.m;
        a.S = b->S;
        a.J = b->Jx.p;
        a.R = b->R.p;
        a.x = b->x.p;
        a.ctl = b->ctl.p;
        a.list = P.list.p;
        a.slot_of = P.slot_of.p;
        a.res = P.res.p;
        a.guard = b->guard.p;
        a.cap = P.cap;
        a.Lp = P.Lp.p; a.Li = P.Li.p; a.Up = P.Up.p; a.Ui = P.Ui.p; a.P = P.P.p; a.pinv = P.pinv.p;
        a.xi = P.xi.p; a.stk = P.stk.p; a.pstk = P.pstk.p; a.flag = P.flag.p;
        a.Lx = P.Lx.p; a.Ux = P.Ux.p; a.Ud = P.Ud.p; a.rs = P.rs.p; a.w = P.w.p; a.y = P.y.p;
        hipLaunchKernelGGL(k_pivot_solve, dim3((unsigned)n), dim3(64), 0, st, a);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(res.data(), P.res.p, n * sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        std::vector<int> again;
        for (int q = 0; q < n; ++q) {
            if (res[q] == 3) throw Error(P3SMI_E_INTERNAL, "pivoting factorization: a case without a slot");
            if (res[q] == 2) again.push_back(list[q]);
            if (res[q] == 4) diagonal.push_back(list[q]);
        }
        if (again.empty()) break;
        // fill beyond the pool's capacity: a pool with twice the capacity, the same slots
        if (P.cap > (long long)h.m * (long long)h.m) throw Error(P3SMI_E_INTERNAL, "pivoting factorization: fill beyond m^2");
        pivot_pool_ensure(b, P.slots, 2 * P.cap);
        list = again;
        n = (int)list.size();
        HIPCHECK(hipMemcpyAsync(P.list.p, list.data(), n * sizeof(int), hipMemcpyHostToDevice, st));
    }
    for (int c : diagonal) pivot_release(b, c);  // on the plan's order: no slot needed any more
    return (int)list.size();
}

// precision class a case starts in
unsigned start_class(int precision) { return precision == P3SMI_PREC_MIXED ? PR_F32 : PR_F64; }

BeginArgs begin_args(p3smi_batch* b, int precision, dim3 grid) {
    BeginArgs a;
    a.V0 = b->V0.p;
    a.v0_stride = b->v0_stride;
    a.ctl = b->ctl.p;
    a.p0 = start_class(precision);
    a.Vm = b->Vm.p;
    a.Va = b->Va.p;
    a.Ve = b->Ve.p;
    a.Vf = b->Vf.p;
    a.n = b->topo->host.n;
    a.S = b->S;
    a.gx = (int)grid.x;
    a.gy = (int)grid.y;
    a.plan = nullptr;
    return a;
}

AdvanceArgs advance_args(p3smi_batch* b, int hold, dim3 grid) {
    const DevTopo& d = b->topo->dev;
    AdvanceArgs a;
    a.x = b->x.p;
    a.xva = d.xva.p;
    a.xvm = d.xvm.p;
    a.ctl = b->ctl.p;
    a.alpha = b->alpha.p;
    a.Vm = b->Vm.p;
    a.Va = b->Va.p;
    a.Ve = b->Ve.p;
    a.Vf = b->Vf.p;
    a.Vm0 = b->Vm0.p;
    a.Va0 = b->Va0.p;
    a.hold = hold;
    a.n = b->topo->host.n;
    a.S = b->S;
    a.gx = (int)grid.x;
    a.gy = (int)grid.y;
    a.plan = nullptr;
    return a;
}

struct PassParams {
    int max_iter, line_search, precision;
    double tol;
    bool lockstep;  // hold the cases in STEP while any case is in its line search
};

CtlArgs ctl_args(p3smi_batch* b, const PassParams& p, int hold) {
    p3smi_topology* t = b->topo;
    CtlArgs a;
    a.part = b->part.p;
    a.nparts = t->dev.ntiles;
    a.ctl = b->ctl.p;
    a.nrm = b->nrm.p;
    a.alpha = b->alpha.p;
    a.best_alpha = b->best_alpha.p;
    a.best_nrm = b->best_nrm.p;
    a.tally = b->tally.p;
    a.counts_out = b->counts.p;
    a.tol = p.tol;
    a.sw = p.precision == P3SMI_PREC_MIXED_SWITCH ? t->mixed_switch : 0.0;
    a.line_search = p.line_search;
    a.max_iter = p.max_iter;
    a.S = b->S;
    a.hold = hold;
    a.reruns = b->guard.p + 3;
    a.gx = (b->S + 255) / 256;
    a.few_max = t->lu_few_max;
    a.plan = nullptr;
    return a;
}

// The kernels of one pass (see the file header); the LU only when cases are in STEP, k_begin
// only when cases are in START.
void pass_kernels(p3smi_batch* b, const PassParams& p, int n_solve, bool starting, int hold, bool pivoting) {
    p3smi_topology* t = b->topo;
    const p3smi::Topology& h = t->host;
    hipStream_t st = t->stream;
    if (n_solve > 0) linear_solve(b, n_solve, p.precision);
    if (pivoting) pivot_solve(b);
    if (starting) {
        KTimer tm(b, 4);
        const dim3 grid = grid_items(h.n, b->S);
        hipLaunchKernelGGL(k_begin, grid, dim3(256), 0, st, begin_args(b, p.precision, grid));
        HIPCHECK(hipGetLastError());
    }
    if ((n_solve > 0 && !hold) || b->h_counts[CNT_TRIAL] > 0 || b->h_counts[CNT_COMMIT] > 0) {
        KTimer tm(b, 2);
        const dim3 grid = grid_items(h.n, b->S);
        hipLaunchKernelGGL(k_advance, grid, dim3(256), 0, st, advance_args(b, hold, grid));
        HIPCHECK(hipGetLastError());
    }
    launch_evaluate(b, hold);
    {
        KTimer tm(b, 3);
        const CtlArgs a = ctl_args(b, p, hold);
        hipLaunchKernelGGL(k_control, dim3((unsigned)a.gx), dim3(256), 0, st, a);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipMemcpyAsync(b->h_counts, b->counts.p, NCOUNT * sizeof(int), hipMemcpyDeviceToHost, st));
}

// ---- hipGraph replay of passes ---------------------------------------------------------
// A pass depends only on device state (the control words steer it), so it is recorded once per
// (LU launch shape, whether it factors and whether it starts cases, solve parameters, buffers)
// and replayed with one graph launch and one synchronisation. The stage counters reset
// themselves in k_control (no memset node is recorded); schedules are uploaded before recording.
// Kernel timers are not recorded: graphs run with timing off. FP64 solves only (the mixed
// modes read refinement counters back between their launches).
template <typename F>
void replay(p3smi_batch* b, const std::string& key, F&& record) {
    hipStream_t st = b->topo->stream;
    auto it = b->gcache.find(key);
    if (it == b->gcache.end()) {
        if (b->gcache.size() >= 32) b->drop_graphs();  // inputs re-allocated many times: start over
        HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            record();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
        HIPCHECK(hipStreamEndCapture(st, &g));
        if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(g);
            throw Error(P3SMI_E_DEVICE, "hipGraphInstantiate failed");
        }
        it = b->gcache.emplace(key, std::make_pair(g, x)).first;
    }
    HIPCHECK(hipGraphLaunch(it->second.second, st));
}

int live_cases(const int* cnt) { return cnt[CNT_START] + cnt[CNT_ACTIVE] + cnt[CNT_TRIAL] + cnt[CNT_COMMIT]; }

// One pass; returns the stage counters after it (C ABI order).
const int* run_pass(p3smi_batch* b, const PassParams& p) {
    p3smi_topology* t = b->topo;
    hipStream_t st = t->stream;
    const bool searching = b->h_counts[CNT_TRIAL] + b->h_counts[CNT_COMMIT] > 0;
    const int hold = (p.lockstep && searching && b->h_counts[CNT_ACTIVE] > 0) ? 1 : 0;
    const int n_solve = hold ? 0 : b->h_counts[CNT_ACTIVE];
    const bool starting = b->pending_start || b->h_counts[CNT_START] > 0;
    // cases off the frozen plan take their step from the pivoting factorization (host round trips:
    // this pass runs kernel by kernel)
    const bool pivoting = !hold && b->h_counts[CNT_PIVOT] > 0;
    if (n_solve > 0) b->stats.newton_passes++;
    if (b->h_counts[CNT_TRIAL] > 0) b->stats.line_search_passes++;
    if (b->part.n < (size_t)t->dev.ntiles * b->S) b->part.alloc((size_t)t->dev.ntiles * b->S);  // never under capture
    if (b->graphs && !b->timing && p.precision == P3SMI_PREC_DOUBLE && !pivoting) {
        const LuConfig cfg = lu_config_for(t, n_solve);
        if (n_solve > 0) {
            (void)get_sched(t, cfg);  // upload before recording
            b->stats.double_factorizations += n_solve;
        }
        char key[360];
        std::snprintf(key, sizeof key, "%d/%d/%d/%d/%d/%a/%d/%d/%p/%d/%p/%p/%p/%p/%d/%d/%d/%p", n_solve > 0 ? cfg.key() : -1,
                      cfg.G, cfg.NS, (int)starting + 2 * hold, (int)searching, p.tol,
                      p.max_iter, p.line_search, (const void*)b->Ssp.p, b->s_stride, (const void*)b->Yg.p,
                      (const void*)b->Yb.p, (const void*)b->pin.p, (const void*)b->V0.p, b->v0_stride,
                      b->ymode() * 2 + (int)b->has_pin, t->dev.tile_rows, (const void*)b->part.p);
        replay(b, key, [&] {
            PassParams q = p;
            const int64_t before = b->stats.double_factorizations;
            pass_kernels(b, q, n_solve, starting, hold, false);
            b->stats.double_factorizations = before;  // counted above, once per replay
        });
    } else {
        pass_kernels(b, p, n_solve, starting, hold, pivoting);
    }
    HIPCHECK(hipStreamSynchronize(st));
    b->pending_start = false;
    return b->h_counts;
}

void reset_cases(p3smi_batch* b, int precision) {
    hipStream_t st = b->topo->stream;
    hipLaunchKernelGGL(k_ctl_reset, dim3((unsigned)((b->S + 255) / 256)), dim3(256), 0, st, b->ctl.p, b->B, b->S,
                       start_class(precision), start_extra(b->topo));
    hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, st, b->guard.p, 6);
    HIPCHECK(hipGetLastError());
    std::memset(b->h_counts, 0, NCOUNT * sizeof(int));
    b->h_counts[CNT_START] = b->B;
    b->h_counts[CNT_IDLE] = b->S - b->B;
    b->pending_start = true;
    pivot_release_all(b);
}

void run_newton(p3smi_batch* b, int max_iter, double tol, int line_search, int precision) {
    if (!b->uploaded) throw Error(P3SMI_E_ARG, "batch has no inputs (upload first)");
    check_solve_args(max_iter, tol, precision);
    hipStream_t st = b->topo->stream;
    ensure_tiles(b->topo);
    std::memset(&b->stats, 0, sizeof b->stats);
    b->precision = precision;
    b->solve_id = ++g_solve_ids;
    b->stats.cases = b->B;
    {
        KTimer tm(b, 4);
        reset_cases(b, precision);
    }
    const PassParams p{max_iter, line_search ? 1 : 0, precision, tol, true};
    // a run of a case needs at most 1 + max_iter * (1 + 1 + LS_MAX_TRIALS + 1) passes (start; per
    // iteration a pending pass after a guard trip, the step, the trials and the commit), and a case
    // has at most two runs
    const long long cap = 2 * (1 + (long long)max_iter * (LS_MAX_TRIALS + 3)) + 2;
    const int* cnt = b->h_counts;
    long long passes = 0;
    while (live_cases(cnt) > 0) {
        if (++passes > cap) throw Error(P3SMI_E_INTERNAL, "Newton passes did not finish every case");
        cnt = run_pass(b, p);
    }
    b->stats.failed_factorizations = cnt[CNT_FAIL];
    int g[6];
    HIPCHECK(hipMemcpyAsync(g, b->guard.p, sizeof g, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    b->stats.guard_trips = g[0];
    b->stats.reruns = g[3];
    b->stats.pivot_factorizations = g[4];
    b->stats.diagonal_factorizations = g[5];
    if (precision == P3SMI_PREC_MIXED_SWITCH) {
        b->stats.single_factorizations = g[1];
        b->stats.double_factorizations = g[2];
    } else if (precision == P3SMI_PREC_MIXED) {
        b->stats.double_factorizations += g[2];  // FP32 pivot failures refactored in FP64
    }
    collect_timing(b);
}

// V goes to columns [c0, c0 + B) of the caller's (n, ld) array.
void download(p3smi_batch* b, double* V_out, size_t ld, size_t c0, int32_t* iterations, uint8_t* converged,
              double* mismatch) {
    const p3smi::Topology& h = b->topo->host;
    hipStream_t st = b->topo->stream;
    b->Vout.alloc((size_t)2 * h.n * b->B);
    hipLaunchKernelGGL(k_emit, grid_items(h.n, b->S), dim3(256), 0, st, b->Ve.p, b->Vf.p,
                       b->mask_unserved ? b->served.p : nullptr, b->Vout.p, h.n, b->S, b->B);
    HIPCHECK(hipGetLastError());
    std::vector<unsigned> ctl(b->B);
    HIPCHECK(hipMemcpyAsync(ctl.data(), b->ctl.p, b->B * sizeof(unsigned), hipMemcpyDeviceToHost, st));
    if (mismatch) HIPCHECK(hipMemcpyAsync(mismatch, b->nrm.p, b->B * sizeof(double), hipMemcpyDeviceToHost, st));
    if (V_out) download_cols(b->stage, V_out, b->Vout.p, h.n, ld, c0, b->B, 16, st);
    HIPCHECK(hipStreamSynchronize(st));
    for (int c = 0; c < b->B; ++c) {
        if (iterations) iterations[c] = iters_of(ctl[c]);
        if (converged) converged[c] = stage_of(ctl[c]) == SG_CONV;
    }
}

void check_v0(const double* V0, size_t rows, size_t cols, const char* who) {
    for (size_t i = 0; i < rows; ++i)
        for (size_t c = 0; c < cols; ++c) {
            const double re = V0[2 * (i * cols + c)], im = V0[2 * (i * cols + c) + 1];
            if (!std::isfinite(re) || !std::isfinite(im))
                throw Error(P3SMI_E_ARG, std::string(who) + "invalid initial voltage at bus " + std::to_string(i) +
                                             ", step " + std::to_string(c));
            const double vm = std::hypot(re, im);
            if (vm <= 0.0 || vm > 10.0)
                throw Error(P3SMI_E_ARG, std::string(who) + "invalid voltage magnitude " + std::to_string(vm) +
                                             " at bus " + std::to_string(i));
        }
}

size_t bytes_per_case(const p3smi::Topology& h, bool per_case_y, bool pin, bool per_case_s) {
    size_t b = (size_t)h.n * 6 * 8 + (size_t)h.nnzJ * 8 + (size_t)h.fac_rows * 8 + (size_t)h.m * 16;
    b += (size_t)max_tiles(h) * 8 + 64;
    b += (size_t)h.n * 16 * 2;  // V0 / V out
    if (per_case_s) b += (size_t)h.n * 16;
    if (per_case_y) b += (size_t)h.nnzY * 16 * 2;  // G, B and the upload staging
    if (pin) b += h.n;
    return b;
}

int chunk_cases(p3smi_topology* t, int total, bool per_case_y, bool pin, bool per_case_s) {
    size_t freeb = 0, totb = 0;
    HIPCHECK(hipMemGetInfo(&freeb, &totb));
    const size_t per = bytes_per_case(t->host, per_case_y, pin, per_case_s);
    size_t cap = (size_t)(0.6 * (double)freeb) / per;
    if (t->cached) cap += (size_t)t->cached->S;  // its memory is reusable
    if (cap < 64) cap = 64;
    if (t->max_chunk > 0 && (size_t)t->max_chunk < cap) cap = (size_t)t->max_chunk;
    if ((size_t)total <= cap) return total;
    // whole waves of cases, and chunks of (nearly) equal size: a small last chunk cannot fill
    // the GPU yet costs a full chunk's iterations (as p3s/cuda/nr_polar_solver.py:630-637)
    cap = cap / 64 * 64;
    const size_t n_chunks = ((size_t)total + cap - 1) / cap;
    const size_t even = ((size_t)total + n_chunks - 1) / n_chunks;
    return (int)std::min(cap, (even + 63) / 64 * 64);
}

p3smi_batch* cached_batch(p3smi_topology* t, int cases) {
    if (!t->cached || t->cached->B != cases) {
        t->cached.reset(new p3smi_batch());
        batch_alloc(t->cached.get(), t, cases);
    }
    if (t->cached->graphs && !t->graphs) t->cached->drop_graphs();
    t->cached->graphs = t->graphs;
    return t->cached.get();
}

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return P3SMI_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return P3SMI_E_MEMORY;
    } catch (const std::exception& e) {
        g_err = e.what();
        return P3SMI_E_ARG;
    }
}

}  // namespace

// ===================================================================== C ABI
extern "C" {

const char* p3smi_last_error(void) { return g_err.c_str(); }
int p3smi_version(void) { return 1; }

int p3smi_device_count(int32_t* count) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return P3SMI_OK;
}

int p3smi_device_memory(int64_t* free_bytes, int64_t* total_bytes) {
    return guarded([&] {
        if (!free_bytes || !total_bytes) throw Error(P3SMI_E_ARG, "device_memory: null argument");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            throw Error(P3SMI_E_DEVICE, "device_memory: no HIP device available");
        size_t f = 0, t = 0;
        HIPCHECK(hipMemGetInfo(&f, &t));
        *free_bytes = (int64_t)f;
        *total_bytes = (int64_t)t;
    });
}

// HBM reference patterns with known byte counts, for scaling the FETCH_SIZE / WRITE_SIZE
// counters and for the bandwidth a pattern can reach. Rows are 512 bytes: one factor entry
// (or one Jacobian entry, one voltage component) of a 64-case group, the unit the LU and the
// evaluation move. Both kernels read every source byte once and write every destination byte
// once.
__global__ void __launch_bounds__(256) k_probe_fill(double2* p, size_t n2) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_double2((double)(i & 1023), 1.0);
}

__global__ void __launch_bounds__(256) k_probe_copy(const double2* __restrict__ src, double2* __restrict__ dst,
                                                    size_t n2) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// One wavefront per row: row r of dst takes row (r * A) mod R of src and lands at row
// (r * B) mod R; A and B are odd and R a power of two, so both maps are permutations and
// consecutive rows fall far apart.
__global__ void __launch_bounds__(256) k_probe_rows(const double* __restrict__ src, double* __restrict__ dst,
                                                    uint64_t rows) {
    const uint64_t mask = rows - 1;
    const int lane = threadIdx.x & 63;
    for (uint64_t r = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; r < rows;
         r += ((uint64_t)gridDim.x * blockDim.x) >> 6) {
        const uint64_t a = (r * 0x9E3779B1ull) & mask, b = (r * 0x85EBCA77ull) & mask;
        dst[b * 64 + lane] = src[a * 64 + lane];
    }
}

int p3smi_hbm_probe(int32_t pattern, int64_t bytes, int32_t repeats, double* ms_per_launch) {
    return guarded([&] {
        if (!ms_per_launch) throw Error(P3SMI_E_ARG, "hbm_probe: null argument");
        if (pattern != 0 && pattern != 1) throw Error(P3SMI_E_ARG, "hbm_probe: pattern must be 0 (copy) or 1 (rows)");
        if (bytes < (1 << 20) || bytes > (int64_t(1) << 36) || (bytes & (bytes - 1)))
            throw Error(P3SMI_E_ARG, "hbm_probe: bytes must be a power of two in [1 MiB, 64 GiB]");
        if (repeats < 1 || repeats > 1000) throw Error(P3SMI_E_ARG, "hbm_probe: repeats must be in [1, 1000]");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            throw Error(P3SMI_E_DEVICE, "hbm_probe: no HIP device available");
        DevBuf<double> src, dst;
        const size_t nd = (size_t)bytes / sizeof(double);
        src.alloc(nd);
        dst.alloc(nd);
        hipStream_t st = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        std::unique_ptr<void, void (*)(void*)> st_guard(st, [](void* s) { (void)hipStreamDestroy((hipStream_t)s); });
        HIPCHECK(hipEventCreate(&e0));
        std::unique_ptr<void, void (*)(void*)> e0_guard(e0, [](void* e) { (void)hipEventDestroy((hipEvent_t)e); });
        HIPCHECK(hipEventCreate(&e1));
        std::unique_ptr<void, void (*)(void*)> e1_guard(e1, [](void* e) { (void)hipEventDestroy((hipEvent_t)e); });
        const size_t n2 = nd / 2;
        const uint64_t rows = (uint64_t)bytes / 512;
        k_probe_fill<<<4096, 256, 0, st>>>((double2*)src.p, n2);
        HIPCHECK(hipGetLastError());
        auto launch = [&] {
            if (pattern == 0)
                k_probe_copy<<<8192, 256, 0, st>>>((const double2*)src.p, (double2*)dst.p, n2);
            else
                k_probe_rows<<<8192, 256, 0, st>>>(src.p, dst.p, rows);
        };
        launch();  // warm-up
        HIPCHECK(hipEventRecord(e0, st));
        for (int i = 0; i < repeats; ++i) launch();
        HIPCHECK(hipEventRecord(e1, st));
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
        *ms_per_launch = (double)ms / repeats;
    });
}

int p3smi_topology_create(const int32_t* Yp, const int32_t* Yj, const double* Yx, int32_t n, int32_t nnz,
                          const int32_t* pv, int32_t npv, const int32_t* pq, int32_t npq, int32_t ordering,
                          int32_t n_slots, p3smi_topology** out) {
    return guarded([&] {
        if (!out) throw Error(P3SMI_E_ARG, "out is NULL");
        *out = nullptr;
        if (ordering != 0 && ordering != 2) throw Error(P3SMI_E_ARG, "ordering must be 0 (fill-reducing) or 2 (natural)");
        if (n_slots == 0) n_slots = 8;
        if (n_slots != 1 && n_slots != 2 && n_slots != 4 && n_slots != 8 && n_slots != 16)
            throw Error(P3SMI_E_ARG, "n_slots must be 1, 2, 4, 8 or 16");
        std::unique_ptr<p3smi_topology> t(new p3smi_topology());
        p3smi::build_topology(t->host, Yp, Yj, Yx, n, nnz, pv, npv, pq, npq, ordering, n_slots);
        *out = t.release();
    });
}

void p3smi_topology_free(p3smi_topology* t) { delete t; }

int p3smi_topology_info(const p3smi_topology* t, p3smi_info* o) {
    return guarded([&] {
        const p3smi::Topology& h = t->host;
        o->n = h.n; o->m = h.m; o->npv = h.npv; o->npq = h.npq; o->nnz_y = h.nnzY; o->nnz_j = h.nnzJ;
        o->nnz_lu = h.nnz_lu; o->n_pairs = h.n_pairs; o->n_levels = h.sched.nlev; o->n_blevels = h.sched.nblev;
        o->n_slots = h.sched.nslots; o->ordering = h.ordering; o->rcond0 = h.rcond0;
        o->lu_g = t->lu_batch.G; o->lu_ns = t->lu_batch.NS; o->lu_few_g = t->lu_few.G; o->lu_few_ns = t->lu_few.NS;
        o->lu_few_max = t->lu_few_max; o->lu_group = t->lu_batch.grp; o->eval_rows = t->eval_rows;
    });
}

int p3smi_topology_set_mixed(p3smi_topology* t, double rtol, int32_t kmax) {
    return guarded([&] {
        if (!(rtol > 0.0) || kmax < 1) throw Error(P3SMI_E_ARG, "mixed precision: rtol must be > 0 and kmax >= 1");
        std::lock_guard<std::mutex> lk(t->mu);
        t->mixed_rtol = rtol;
        t->mixed_kmax = kmax;
    });
}

int p3smi_topology_set_mixed_switch(p3smi_topology* t, double mismatch) {
    return guarded([&] {
        if (!(mismatch > 0.0) || !std::isfinite(mismatch))
            throw Error(P3SMI_E_ARG, "mixed_switch: the switching mismatch must be finite and > 0");
        std::lock_guard<std::mutex> lk(t->mu);
        t->mixed_switch = mismatch;
    });
}

int p3smi_topology_set_max_chunk(p3smi_topology* t, int32_t cases) {
    return guarded([&] {
        if (!t) throw Error(P3SMI_E_ARG, "set_max_chunk: null topology");
        if (cases != 0 && (cases < 64 || cases % 64))
            throw Error(P3SMI_E_ARG, "set_max_chunk: cases must be 0 or a positive multiple of 64");
        std::lock_guard<std::mutex> lk(t->mu);
        t->max_chunk = cases;
    });
}

int p3smi_topology_set_lu_shape(p3smi_topology* t, int32_t g, int32_t ns, int32_t few_g, int32_t few_ns,
                                int32_t few_max, int32_t group) {
    return guarded([&] {
        if (group < 0 || group > 4096) throw Error(P3SMI_E_ARG, "LU launch shape: group must be in [0, 4096]");
        const LuConfig a{g, ns, group}, f{few_g, few_ns, group};
        if (!valid_lu_config(a) || !valid_lu_config(f))
            throw Error(P3SMI_E_ARG, "LU launch shape: g must be 1, 16, 32 or 64 and ns 8 or 16");
        if (few_max < 0) throw Error(P3SMI_E_ARG, "LU launch shape: few_max must be >= 0");
        std::lock_guard<std::mutex> lk(t->mu);
        t->lu_batch = a;
        t->lu_few = f;
        t->lu_few_max = few_max;
    });
}

int p3smi_topology_set_guard(p3smi_topology* t, double ratio) {
    return guarded([&] {
        if (!t) throw Error(P3SMI_E_ARG, "set_guard: null topology");
        if (!(ratio >= 0.0) || !std::isfinite(ratio)) throw Error(P3SMI_E_ARG, "set_guard: ratio must be finite and >= 0");
        std::lock_guard<std::mutex> lk(t->mu);
        t->guard_ratio = ratio;
    });
}

int p3smi_topology_set_pivoting(p3smi_topology* t, int32_t diagonal_first) {
    return guarded([&] {
        if (!t) throw Error(P3SMI_E_ARG, "set_pivoting: null topology");
        if (diagonal_first != 0 && diagonal_first != 1) throw Error(P3SMI_E_ARG, "set_pivoting: diagonal_first must be 0 or 1");
        std::lock_guard<std::mutex> lk(t->mu);
        t->pivot_diagonal_first = diagonal_first != 0;
    });
}

int p3smi_topology_set_eval_tile(p3smi_topology* t, int32_t rows) {
    return guarded([&] {
        if (!t) throw Error(P3SMI_E_ARG, "set_eval_tile: null topology");
        if (rows < 1 || rows > 1024) throw Error(P3SMI_E_ARG, "set_eval_tile: rows must be in [1, 1024]");
        std::lock_guard<std::mutex> lk(t->mu);
        t->eval_rows = rows;
    });
}

int p3smi_dc_start(p3smi_topology* t, const double* Bdc, const double* P, const double* V_flat, int32_t nref,
                   const int32_t* ref, double* V_out) {
    return guarded([&] {
        if (!t || !Bdc || !P || !V_flat || !V_out || (nref > 0 && !ref) || nref < 0)
            throw Error(P3SMI_E_ARG, "dc_start: null argument");
        std::lock_guard<std::mutex> lk(t->mu);
        const p3smi::Topology& h = t->host;
        const int n = h.n;
        std::vector<double> theta(n, 0.0);
        std::vector<char> is_ref(n, 0);
        for (int q = 0; q < nref; ++q) {
            if (ref[q] < 0 || ref[q] >= n) throw Error(P3SMI_E_ARG, "dc_start: reference bus out of range");
            is_ref[ref[q]] = 1;
            theta[ref[q]] = std::atan2(V_flat[2 * (size_t)ref[q] + 1], V_flat[2 * (size_t)ref[q]]);
        }
        // the matrix [B'(pvpq, pvpq) 0; 0 I] on the Jacobian's pattern and the right-hand side
        // P(pvpq) - B'(pvpq, ref) theta_ref (p3s/NewtonPowerflowCpp.py:422-440)
        auto bus_of = [&](int r) { return r < h.npvpq ? h.pvpq[r] : h.pq[r - h.npvpq]; };
        std::vector<double> Jv(h.nnzJ), R(h.m, 0.0);
        for (int r = 0; r < h.m; ++r)
            for (int e = h.Jp[r]; e < h.Jp[r + 1]; ++e) {
                const int blk = h.Jb[e];
                Jv[e] = blk == 0 ? Bdc[h.Jk[e]] : (blk == 3 && bus_of(r) == bus_of(h.Jj[e])) ? 1.0 : 0.0;
            }
        for (int r = 0; r < h.npvpq; ++r) {
            const int i = h.pvpq[r];
            double rhs = P[i];
            for (int k = h.Yp[i]; k < h.Yp[i + 1]; ++k)
                if (is_ref[h.Yj[k]]) rhs -= Bdc[k] * theta[h.Yj[k]];
            R[r] = rhs;
        }
        // one static-pivot factorization and solve on a one-case batch (B' is symmetric positive
        // definite on a grid whose islands all hold a reference, so no pivot ratio guard applies)
        // (the workspace is kept for later calls; the dense vectors go up and down in one copy
        // each and are put into / taken from column 0 on the device)
        if (!t->dc) {
            std::unique_ptr<p3smi_batch> nb(new p3smi_batch());
            batch_alloc(nb.get(), t, 1);
            t->dc = std::move(nb);
        }
        p3smi_batch& b = *t->dc;
        b.guard_override = 0.0;
        hipStream_t st = t->stream;
        const int S = (int)b.S;
        DevBuf<double> dense;
        dense.alloc((size_t)std::max(h.nnzJ, h.m));
        auto put = [&](const std::vector<double>& v, double* col) {
            HIPCHECK(hipMemcpyAsync(dense.p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_column, dim3((unsigned)((v.size() + 255) / 256)), dim3(256), 0, st, dense.p, col,
                               (int)v.size(), S, 0, 0);
            HIPCHECK(hipGetLastError());
        };
        put(Jv, b.Jx.p);
        put(R, b.R.p);
        const unsigned w = SG_STEP;
        HIPCHECK(hipMemcpyAsync(b.ctl.p, &w, sizeof w, hipMemcpyHostToDevice, st));
        launch_lu<double, true>(&b, 1, b.fac.p, b.R.p, b.x.p, SEL_ALL);
        std::vector<double> x(h.m);
        unsigned w1 = 0;
        hipLaunchKernelGGL(k_column, dim3((unsigned)((h.m + 255) / 256)), dim3(256), 0, st, b.x.p, dense.p, h.m, S, 0, 1);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(x.data(), dense.p, (size_t)h.m * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(&w1, b.ctl.p, sizeof w1, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        const bool singular = (w1 & CTL_PENDING) != 0;  // a zero or non-finite pivot
        // angles at pv and pq buses (a singular solve gives NaN, taken as 0 as numpy's
        // nan_to_num does, p3s/timeseries.py:200-244), magnitudes of the start kept
        for (int i = 0; i < 2 * n; ++i) V_out[i] = V_flat[i];
        bool finite = true;
        for (int r = 0; r < h.npvpq; ++r) {
            const int i = h.pvpq[r];
            double th = singular ? 0.0 : x[h.pinv[r]];
            if (std::isnan(th)) th = 0.0;
            else if (std::isinf(th)) th = th > 0 ? DBL_MAX : -DBL_MAX;
            const double vm = std::hypot(V_flat[2 * (size_t)i], V_flat[2 * (size_t)i + 1]);
            V_out[2 * (size_t)i] = vm * std::cos(th);
            V_out[2 * (size_t)i + 1] = vm * std::sin(th);
            finite = finite && std::isfinite(V_out[2 * (size_t)i]) && std::isfinite(V_out[2 * (size_t)i + 1]);
        }
        if (!finite)
            for (int i = 0; i < 2 * n; ++i) V_out[i] = V_flat[i];
    });
}

int p3smi_update_y(p3smi_topology* t, const double* Yx, int32_t nnz) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(t->mu);
        if (nnz != t->host.nnzY) throw Error(P3SMI_E_ARG, "update_Y: nnz mismatch (topology changed; rebuild Solver)");
        p3smi::update_values(t->host, Yx);
        if (t->dev.ready) {
            t->dev.G.upload(t->host.G.data(), t->host.G.size(), t->stream);
            t->dev.B.upload(t->host.B.data(), t->host.B.size(), t->stream);
            HIPCHECK(hipStreamSynchronize(t->stream));
        }
    });
}

int p3smi_host_jacobian(const p3smi_topology* t, const double* S, const double* V, int32_t* Jp, int32_t* Jj,
                        double* Jx, double* R) {
    return guarded([&] {
        const p3smi::Topology& h = t->host;
        std::vector<double> Vm(h.n), Va(h.n);
        for (int i = 0; i < h.n; ++i) { Vm[i] = std::hypot(V[2 * i], V[2 * i + 1]); Va[i] = std::atan2(V[2 * i + 1], V[2 * i]); }
        p3smi::host_eval(h, Vm.data(), Va.data(), nullptr, nullptr, S, nullptr, R, Jx);
        std::copy(h.Jp.begin(), h.Jp.end(), Jp);
        std::copy(h.Jj.begin(), h.Jj.end(), Jj);
    });
}

int p3smi_host_factor_solve(const p3smi_topology* t, const double* Jx, const double* R, double* x, double* pivots) {
    return guarded([&] {
        const p3smi::Topology& h = t->host;
        std::vector<double> fac(h.nnz_lu), y(h.m);
        p3smi::host_factor(h, Jx, fac.data(), &pivots[0], &pivots[1], &pivots[2]);
        p3smi::host_solve(h, fac.data(), R, y.data());
        for (int p = 0; p < h.m; ++p) x[h.perm[p]] = y[p];
    });
}

int p3smi_host_lu_emulate(const p3smi_topology* t, int32_t blocks, int32_t nslots, int32_t single, const double* Jx,
                          const double* R, double* x, double* pivots, int64_t* checks) {
    return guarded([&] {
        if (!t || !Jx || !R || !x || !pivots || !checks) throw Error(P3SMI_E_ARG, "host_lu_emulate: null argument");
        if (nslots < 1 || nslots > 4096) throw Error(P3SMI_E_ARG, "host_lu_emulate: nslots must be in [1, 4096]");
        const p3smi::Topology& h = t->host;
        std::vector<double> y(h.m);
        p3smi::host_lu_emulate(h, blocks != 0, nslots, single != 0, Jx, R, y.data(), &pivots[0], &pivots[1], checks);
        for (int p = 0; p < h.m; ++p) x[h.perm[p]] = y[p];
    });
}

int p3smi_host_maps(const p3smi_topology* t, int32_t* row_p, int32_t* row_q, int32_t* perm) {
    return guarded([&] {
        const p3smi::Topology& h = t->host;
        if (row_p) std::copy(h.row_p.begin(), h.row_p.end(), row_p);
        if (row_q) std::copy(h.row_q.begin(), h.row_q.end(), row_q);
        if (perm) std::copy(h.perm.begin(), h.perm.end(), perm);
    });
}

int p3smi_batch_create(p3smi_topology* t, int32_t cases, p3smi_batch** out) {
    return guarded([&] {
        if (cases <= 0) throw Error(P3SMI_E_ARG, "cases must be > 0");
        std::unique_ptr<p3smi_batch> b(new p3smi_batch());
        batch_alloc(b.get(), t, cases);
        *out = b.release();
    });
}

void p3smi_batch_free(p3smi_batch* b) { delete b; }

int p3smi_batch_upload(p3smi_batch* b, const double* S, const double* V0, int32_t v0_per_case) {
    return guarded([&] {
        const p3smi::Topology& h = b->topo->host;
        hipStream_t st = b->topo->stream;
        check_v0(V0, h.n, v0_per_case ? b->B : 1, "solve_batch: ");
        if (S) {
            b->Ssp.alloc((size_t)2 * h.n * b->B);
            upload_cols(b->stage, b->Ssp.p, S, h.n, b->B, 0, b->B, 16, st);
        }
        else if (!b->sbus_built) throw Error(P3SMI_E_ARG, "batch_upload: S is NULL and no Sbus was built on the device");
        b->sbus_built = false;
        b->s_stride = b->B;
        if (v0_per_case) {
            b->V0.alloc((size_t)2 * h.n * b->B);
            upload_cols(b->stage, b->V0.p, V0, h.n, b->B, 0, b->B, 16, st);
        } else {
            b->V0.upload(V0, (size_t)2 * h.n, st);
        }
        b->v0_stride = v0_per_case ? b->B : 0;
        b->per_case_y = false;
        b->ydptr = nullptr;
        b->has_pin = false;
        b->mask_unserved = false;
        HIPCHECK(hipStreamSynchronize(st));
        b->uploaded = true;
        b->slot_mode = false;
    });
}

int p3smi_batch_solve(p3smi_batch* b, int32_t max_iter, double tol, int32_t line_search, int32_t precision) {
    return guarded([&] { run_newton(b, max_iter, tol, line_search, precision); });
}

int p3smi_batch_download(p3smi_batch* b, double* V_out, int32_t* iterations, uint8_t* converged, double* mismatch) {
    return guarded([&] { download(b, V_out, (size_t)b->B, 0, iterations, converged, mismatch); });
}

static void build_sbus(p3smi_batch* b, const double* s0, const int32_t* ptr, int32_t nk, const int32_t* row,
                       const double* coef, const double* base, const uint8_t* imag, int32_t nblk,
                       const int32_t* blk_rows, const double* const* blk) {
    if (!b || !s0 || !ptr || (nk > 0 && (!row || !coef || !base || !imag)) || nblk < 0 || (nblk > 0 && !blk_rows))
        throw Error(P3SMI_E_ARG, "build_sbus: null argument");
    int64_t nrows = 0;
    for (int k = 0; k < nblk; ++k) {
        if (blk_rows[k] < 0) throw Error(P3SMI_E_ARG, "build_sbus: negative block row count");
        if (blk_rows[k] > 0 && (!blk || !blk[k])) throw Error(P3SMI_E_ARG, "build_sbus: null argument");
        nrows += blk_rows[k];
    }
    if (nrows > INT_MAX) throw Error(P3SMI_E_ARG, "build_sbus: too many series rows");
    const p3smi::Topology& h = b->topo->host;
    const int n = h.n, T = b->B;
    if (ptr[0] != 0 || ptr[n] != nk) throw Error(P3SMI_E_ARG, "build_sbus: ptr must start at 0 and end at nk");
    for (int i = 0; i < n; ++i)
        if (ptr[i + 1] < ptr[i]) throw Error(P3SMI_E_ARG, "build_sbus: ptr is not monotone");
    for (int k = 0; k < nk; ++k)
        if (row[k] < 0 || row[k] >= nrows) throw Error(P3SMI_E_ARG, "build_sbus: row index out of range");
    hipStream_t st = b->topo->stream;
    DevBuf<double> ds0, dcoef, dbase;
    DevBuf<int> dptr, drow;
    DevBuf<unsigned char> dimag;
    ds0.upload(s0, 2 * (size_t)n, st);
    dptr.upload(ptr, (size_t)n + 1, st);
    drow.upload(row, (size_t)nk, st);
    dcoef.upload(coef, (size_t)nk, st);
    dbase.upload(base, (size_t)nk, st);
    dimag.upload(imag, (size_t)nk, st);
    // the series, block after block, into the batch's (nrows, cases) array
    if (b->sbX.n < (size_t)std::max<int64_t>(nrows, 1) * T) b->sbX.alloc((size_t)std::max<int64_t>(nrows, 1) * T);
    for (int k = 0, r0 = 0; k < nblk; r0 += blk_rows[k], ++k)
        if (blk_rows[k] > 0)
            upload_cols(b->stage, b->sbX.p + (size_t)r0 * T, blk[k], (size_t)blk_rows[k], (size_t)T, 0, (size_t)T, 8,
                        st);
    b->Ssp.alloc((size_t)2 * n * T);
    const dim3 grid((unsigned)((T + 255) / 256), (unsigned)std::min(n, 65535));
    hipLaunchKernelGGL(k_sbus, grid, dim3(256), 0, st, ds0.p, dptr.p, drow.p, dcoef.p, dbase.p, dimag.p, b->sbX.p,
                       b->Ssp.p, n, T);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
    b->sbus_built = true;
}

int p3smi_batch_build_sbus(p3smi_batch* b, const double* s0, const int32_t* ptr, int32_t nk, const int32_t* row,
                           const double* coef, const double* base, const uint8_t* imag, int32_t nrows, const double* X) {
    return guarded([&] {
        if (nrows < 0) throw Error(P3SMI_E_ARG, "build_sbus: nrows must be >= 0");
        if (nrows > 0 && !X) throw Error(P3SMI_E_ARG, "build_sbus: null argument");
        build_sbus(b, s0, ptr, nk, row, coef, base, imag, 1, &nrows, &X);
    });
}

int p3smi_batch_build_sbus_blocks(p3smi_batch* b, const double* s0, const int32_t* ptr, int32_t nk,
                                  const int32_t* row, const double* coef, const double* base, const uint8_t* imag,
                                  int32_t cols, int32_t nblk, const int32_t* blk_rows, const double* const* blk) {
    return guarded([&] {
        if (!b) throw Error(P3SMI_E_ARG, "build_sbus: null argument");
        // every block is read as (rows, cols) with row stride cols: the width must be the batch's
        if (cols != b->B)
            throw Error(P3SMI_E_ARG, "build_sbus: the series have " + std::to_string(cols) + " columns, the batch " +
                                         std::to_string(b->B) + " cases");
        build_sbus(b, s0, ptr, nk, row, coef, base, imag, nblk, blk_rows, blk);
    });
}

int p3smi_batch_download_sbus(p3smi_batch* b, double* S_out) {
    return guarded([&] {
        const p3smi::Topology& h = b->topo->host;
        if (!b->sbus_built && b->s_stride != b->B) throw Error(P3SMI_E_ARG, "download_sbus: the batch has no (n, B) Sbus");
        download_cols(b->stage, S_out, b->Ssp.p, h.n, b->B, 0, b->B, 16, b->topo->stream);
        HIPCHECK(hipStreamSynchronize(b->topo->stream));
    });
}

int p3smi_topology_release_cache(p3smi_topology* t) {
    return guarded([&] {
        if (!t) throw Error(P3SMI_E_ARG, "topology_release_cache: null topology");
        std::lock_guard<std::mutex> lk(t->mu);
        t->cached.reset();
    });
}

int p3smi_topology_last_batch(p3smi_topology* t, p3smi_batch** out) {
    if (!t || !out) {
        g_err = "topology_last_batch: null argument";
        return P3SMI_E_ARG;
    }
    *out = t->cached.get();
    if (!*out) {
        g_err = "topology_last_batch: no one-call solve has run on this topology";
        return P3SMI_E_ARG;
    }
    return P3SMI_OK;
}

int p3smi_batch_vrange(p3smi_batch* b, double* vmin, double* vmax) {
    return guarded([&] {
        if (!b || !vmin || !vmax) throw Error(P3SMI_E_ARG, "batch_vrange: null argument");
        const p3smi::Topology& h = b->topo->host;
        hipStream_t st = b->topo->stream;
        DevBuf<double> lo, hi;
        lo.alloc(b->B);
        hi.alloc(b->B);
        hipLaunchKernelGGL(k_vrange, dim3((unsigned)((b->B + 255) / 256)), dim3(256), 0, st, b->Ve.p, b->Vf.p,
                           b->mask_unserved ? b->served.p : nullptr, lo.p, hi.p, h.n, b->S, b->B);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(vmin, lo.p, b->B * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(vmax, hi.p, b->B * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    });
}

int p3smi_batch_stats(const p3smi_batch* b, p3smi_stats* out) {
    return guarded([&] {
        *out = b->stats;
        if (b->slot_mode) {
            // passes of a slot batch do not read the device counters back: read them here
            // (guard trips; FP32 and FP64 factorizations of the switching mode) since slots_init
            int g[6];
            HIPCHECK(hipMemcpyAsync(g, b->guard.p, sizeof g, hipMemcpyDeviceToHost, b->topo->stream));
            HIPCHECK(hipStreamSynchronize(b->topo->stream));
            out->guard_trips = g[0];
            out->reruns = g[3];
            out->pivot_factorizations = g[4];
            out->diagonal_factorizations = g[5];
            if (b->precision == P3SMI_PREC_MIXED_SWITCH) {
                out->single_factorizations = g[1];
                out->double_factorizations = g[2];
            }
        }
    });
}

int p3smi_topology_set_graphs(p3smi_topology* t, int32_t on) {
    return guarded([&] {
        if (!t) throw Error(P3SMI_E_ARG, "topology_set_graphs: null topology");
        std::lock_guard<std::mutex> lk(t->mu);
        t->graphs = on != 0;
    });
}

int p3smi_batch_set_graphs(p3smi_batch* b, int32_t on) {
    if (!b) {
        g_err = "batch_set_graphs: null batch";
        return P3SMI_E_ARG;
    }
    b->graphs = on != 0;
    if (!b->graphs) b->drop_graphs();
    return P3SMI_OK;
}

int p3smi_batch_set_timing(p3smi_batch* b, int32_t on) {
    b->timing = on != 0;
    return P3SMI_OK;
}

// Algorithmic bytes per case and pass: every array a kernel must touch, touched once.
int p3smi_batch_bytes_per_case(const p3smi_batch* b, double* out4) {
    const p3smi::Topology& h = b->topo->host;
    const double n = h.n, m = h.m, nnzJ = h.nnzJ, lu = (double)h.nnz_lu;
    // evaluation: Ve, Vf (16 n) + S (16 n) + J write (8 nnzJ) + R write (8 m) [+ per-case Y 16 nnzY]
    out4[0] = 16 * n + (b->s_stride ? 16 * n : 0) + 8 * nnzJ + 8 * m + (b->per_case_y ? 16.0 * h.nnzY : 0.0);
    // lu: J read (8 nnzJ) + factor write (8 nnz_lu) + the substitutions' read of every factor
    // entry (8 nnz_lu: a 64-case group's factors, nnz_lu * 512 B, do not stay on chip between
    // the factor and solve phases) + R read (8 m) + x write/read (16 m)
    out4[1] = 8 * nnzJ + 16 * lu + 8 * m + 16 * m;
    // step: x read (8 m) + Vm, Va read/write + cs/sn write + base write (8 n * 8)
    out4[2] = 8 * m + 64 * n;
    out4[3] = 64;
    return P3SMI_OK;
}

// Returns the number of chunks the cases were solved in.
static void add_stats(p3smi_stats& a, const p3smi_stats& s) {
    a.cases += s.cases;
    a.newton_passes += s.newton_passes;
    a.line_search_passes += s.line_search_passes;
    a.failed_factorizations += s.failed_factorizations;
    a.guard_trips += s.guard_trips;
    a.single_factorizations += s.single_factorizations;
    a.double_factorizations += s.double_factorizations;
    a.refinement_steps += s.refinement_steps;
    a.ms_eval += s.ms_eval;
    a.ms_lu += s.ms_lu;
    a.ms_step += s.ms_step;
    a.ms_decide += s.ms_decide;
    a.ms_other += s.ms_other;
    a.ms_refine += s.ms_refine;
    a.reruns += s.reruns;
    a.pivot_factorizations += s.pivot_factorizations;
    a.diagonal_factorizations += s.diagonal_factorizations;
    a.n_eval += s.n_eval;
    a.n_lu += s.n_lu;
    a.n_step += s.n_step;
    a.n_decide += s.n_decide;
    a.n_refine += s.n_refine;
    a.ms_lu_batch += s.ms_lu_batch;
    a.n_lu_batch += s.n_lu_batch;
    a.lu_batch_cases += s.lu_batch_cases;
}

static int solve_chunked(p3smi_topology* t, int T, bool per_case_y, bool pin_given, bool per_case_s,
                         const std::function<void(p3smi_batch*, int, int)>& upload, int max_iter, double tol,
                         int line_search, int precision, double* V_out, int32_t* iters, uint8_t* conv,
                         double* mismatch) {
    const p3smi::Topology& h = t->host;
    ensure_device(t);
    const int cs = chunk_cases(t, T, per_case_y, pin_given, per_case_s);
    p3smi_stats total{};  // the counters of every chunk, left in the last chunk's batch
    p3smi_batch* b = nullptr;
    for (int c0 = 0; c0 < T; c0 += cs) {
        const int nc = std::min(cs, T - c0);
        b = cached_batch(t, nc);
        upload(b, c0, nc);
        b->uploaded = true;
        b->slot_mode = false;
        run_newton(b, max_iter, tol, line_search, precision);
        add_stats(total, b->stats);
        download(b, V_out, (size_t)T, (size_t)c0, iters ? iters + c0 : nullptr, conv ? conv + c0 : nullptr,
                 mismatch ? mismatch + c0 : nullptr);
    }
    b->stats = total;
    return (T + cs - 1) / cs;
}

int p3smi_solve_batch(p3smi_topology* t, const double* S, const double* V0, int32_t v0_per_case, int32_t T,
                      int32_t max_iter, double tol, int32_t line_search, int32_t precision, double* V_out,
                      int32_t* iterations, uint8_t* converged, double* mismatch) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(t->mu);
        const p3smi::Topology& h = t->host;
        if (T <= 0) throw Error(P3SMI_E_ARG, "solve_batch: T (time steps) must be > 0");
        if (max_iter <= 0) throw Error(P3SMI_E_ARG, "solve_batch: max_iter must be > 0");
        if (!(tol > 0.0)) throw Error(P3SMI_E_ARG, "solve_batch: tol must be > 0");
        check_v0(V0, h.n, v0_per_case ? T : 1, "solve_batch: ");
        auto up = [&](p3smi_batch* b, int c0, int nc) {
            hipStream_t st = t->stream;
            b->Ssp.alloc((size_t)2 * h.n * nc);
            upload_cols(b->stage, b->Ssp.p, S, h.n, T, c0, nc, 16, st);
            b->s_stride = nc;
            if (v0_per_case) {
                b->V0.alloc((size_t)2 * h.n * nc);
                upload_cols(b->stage, b->V0.p, V0, h.n, T, c0, nc, 16, st);
                b->v0_stride = nc;
            } else {
                b->V0.upload(V0, (size_t)2 * h.n, st);
                b->v0_stride = 0;
            }
            b->per_case_y = false;
            b->ydptr = nullptr;
            b->has_pin = false;
            b->mask_unserved = false;
        };
        solve_chunked(t, T, false, false, true, up, max_iter, tol, line_search, precision, V_out, iterations, converged,
                      mismatch);
    });
}

// upload_y null: the cases' admittances are the shared Ybus minus per-case deltas (dptr, dpos,
// dval, device arrays indexed by the case), subtracted as the evaluation reads them
static void contingency_common(p3smi_topology* t, const double* S, const double* V0, const uint8_t* pin, int L,
                               const std::function<void(p3smi_batch*, int, int)>& upload_y, int max_iter, double tol,
                               int line_search, int precision, double* V_out, int32_t* iters, uint8_t* conv,
                               double* mismatch, const int* dptr = nullptr, const int* dpos = nullptr,
                               const double* dval = nullptr) {
    const p3smi::Topology& h = t->host;
    if (L <= 0) throw Error(P3SMI_E_ARG, "solve_batch_contingency: L (cases) must be > 0");
    if (max_iter <= 0) throw Error(P3SMI_E_ARG, "solve_batch_contingency: max_iter must be > 0");
    if (!(tol > 0.0)) throw Error(P3SMI_E_ARG, "solve_batch_contingency: tol must be > 0");
    auto up = [&](p3smi_batch* b, int c0, int nc) {
        hipStream_t st = t->stream;
        b->Ssp.upload(S, (size_t)2 * h.n, st);
        b->s_stride = 0;
        b->V0.alloc((size_t)2 * h.n * nc);
        upload_cols(b->stage, b->V0.p, V0, h.n, L, c0, nc, 16, st);
        b->v0_stride = nc;
        if (upload_y) {
            b->Yg.alloc((size_t)h.nnzY * b->S);
            b->Yb.alloc((size_t)h.nnzY * b->S);
            upload_y(b, c0, nc);
            b->per_case_y = true;
            b->ydptr = nullptr;
        } else {
            b->per_case_y = false;
            b->ydptr = dptr;
            b->ydpos = dpos;
            b->ydval = dval;
            b->ydc0 = c0;
        }
        b->has_pin = pin != nullptr;
        b->mask_unserved = false;
        if (pin) {
            // pin (n, L) -> device (n, S): zero the padding columns
            b->pin.alloc((size_t)h.n * b->S);
            HIPCHECK(hipMemsetAsync(b->pin.p, 0, (size_t)h.n * b->S, st));
            h2d_rows(b->stage, b->pin.p, b->S, pin + c0, L, nc, h.n, st);
        }
    };
    solve_chunked(t, L, (bool)upload_y, pin != nullptr, false, up, max_iter, tol, line_search, precision, V_out, iters,
                  conv, mismatch);
}

int p3smi_solve_batch_contingency(p3smi_topology* t, const double* Yx_mag, const double* Yx_ang, const double* S,
                                  const double* V0, const uint8_t* pin, int32_t L, int32_t max_iter, double tol,
                                  int32_t line_search, int32_t precision, double* V_out, int32_t* iterations,
                                  uint8_t* converged, double* mismatch) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(t->mu);
        const p3smi::Topology& h = t->host;
        auto upy = [&](p3smi_batch* b, int c0, int nc) {
            hipStream_t st = t->stream;
            DevBuf<double> mag, ang;
            mag.alloc((size_t)h.nnzY * nc);
            ang.alloc((size_t)h.nnzY * nc);
            upload_cols(b->stage, mag.p, Yx_mag, h.nnzY, L, c0, nc, 8, st);
            upload_cols(b->stage, ang.p, Yx_ang, h.nnzY, L, c0, nc, 8, st);
            hipLaunchKernelGGL(k_admittance_fill, grid_items(h.nnzY, b->S), dim3(256), 0, st, 1, mag.p, ang.p, nc, nc,
                               b->Yg.p, b->Yb.p, h.nnzY, b->S);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipStreamSynchronize(st));
        };
        contingency_common(t, S, V0, pin, L, upy, max_iter, tol, line_search, precision, V_out, iterations, converged,
                           mismatch);
    });
}

int p3smi_solve_batch_contingency_stamp(p3smi_topology* t, const int32_t* stamp_k, const int32_t* stamp_c,
                                        const double* stamp_val, int64_t n_stamp, const double* S, const double* V0,
                                        const uint8_t* pin, int32_t L, int32_t max_iter, double tol,
                                        int32_t line_search, int32_t precision, double* V_out, int32_t* iterations,
                                        uint8_t* converged, double* mismatch) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(t->mu);
        const p3smi::Topology& h = t->host;
        for (int64_t e = 0; e < n_stamp; ++e)
            if (stamp_k[e] < 0 || stamp_k[e] >= h.nnzY || stamp_c[e] < 0 || stamp_c[e] >= L)
                throw Error(P3SMI_E_ARG, "solve_batch_contingency_stamp: stamp index out of range");
        if (n_stamp > INT_MAX) throw Error(P3SMI_E_ARG, "solve_batch_contingency_stamp: too many stamp entries");
        // the stamp as per-case delta lists (entries of a case in input order)
        std::vector<int> dptr((size_t)L + 1, 0), dpos((size_t)std::max<int64_t>(n_stamp, 1));
        std::vector<double> dval(2 * (size_t)std::max<int64_t>(n_stamp, 1));
        for (int64_t e = 0; e < n_stamp; ++e) dptr[(size_t)stamp_c[e] + 1]++;
        for (int c = 0; c < L; ++c) dptr[c + 1] += dptr[c];
        {
            std::vector<int> fill(dptr.begin(), dptr.end() - 1);
            for (int64_t e = 0; e < n_stamp; ++e) {
                const int q = fill[stamp_c[e]]++;
                dpos[q] = stamp_k[e];
                dval[2 * (size_t)q] = stamp_val[2 * e];
                dval[2 * (size_t)q + 1] = stamp_val[2 * e + 1];
            }
        }
        DevBuf<int> ddp, ddq;
        DevBuf<double> ddv;
        ensure_device(t);
        ddp.upload(dptr.data(), dptr.size(), t->stream);
        ddq.upload(dpos.data(), dpos.size(), t->stream);
        ddv.upload(dval.data(), dval.size(), t->stream);
        // the cases' admittances: the shared Ybus minus their stamps, subtracted in the evaluation
        contingency_common(t, S, V0, pin, L, nullptr, max_iter, tol, line_search, precision, V_out, iterations,
                           converged, mismatch, ddp.p, ddq.p, ddv.p);
    });
}

}  // extern "C"

// ---- N-1 outage sets ---------------------------------------------------------------
namespace {

constexpr int ISLAND_CHUNK = 2048;  // cases per islands launch (bounds the scratch)

bool islands_need_z(const p3smi_outages* o) { return o->reslack && o->ngen; }

// Islands + solver inputs of cases [c0, c0 + nc) into column-strided (stride S) outputs, in
// launches of at most `chunk` cases. uf holds chunk * 2n ints (union-find parents and island
// references), z (with re-slacking) chunk * n words of generator |Z|.
// Whether a case's union-find (8 n bytes) fits the LDS of one workgroup of the current device
// next to the kernel's static 1 KiB; the answer is asked of the device once.
bool islands_in_lds(int n) {
    static std::mutex mu;
    static std::map<int, int> max_lds;  // device -> LDS bytes per workgroup
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    int lim = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = max_lds.find(dev);
        if (it == max_lds.end()) {
            if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) lim = 0;
            max_lds[dev] = lim;
        } else {
            lim = it->second;
        }
    }
    return (size_t)8 * n + sizeof(int) * MAX_GROUP <= (size_t)lim;
}

void launch_islands(p3smi_outages* o, int c0, int nc, int S, int chunk, int* uf, unsigned long long* z,
                    const double* dVstart, unsigned char* served, int* refbus, unsigned char* pin, double* V0,
                    int* npinned) {
    p3smi_topology* t = o->topo;
    const int n = t->host.n;
    hipStream_t st = t->stream;
    const bool lds = islands_in_lds(n);
    const size_t lds_bytes = lds ? (size_t)8 * n : 0;
    if (lds && lds_bytes > (size_t)64 * 1024)
        HIPCHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_islands<true>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
    if (!lds && !uf) throw Error(P3SMI_E_INTERNAL, "islands: no union-find scratch");
    for (int a = 0; a < nc; a += chunk) {
        const int k = std::min(chunk, nc - a);
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3((unsigned)k), dim3(256), lds_bytes, st, n, o->nbr, o->f.p, o->t.p,
                               o->live.p, o->gptr.p, o->gbr.p, o->nref, o->ref.p, o->ngen, o->gbus.p, o->gz.p,
                               o->reslack, c0 + a, uf, islands_need_z(o) ? z : nullptr, dVstart, served + a,
                               refbus ? refbus + a : nullptr, pin ? pin + a : nullptr,
                               V0 ? V0 + 2 * (size_t)a : nullptr, npinned + a, S);
        };
        if (lds) go(k_islands<true>);
        else go(k_islands<false>);
        HIPCHECK(hipGetLastError());
    }
}

void run_islands(p3smi_outages* o, int c0, int nc, int S, const double* dVstart, unsigned char* served,
                 int* refbus, unsigned char* pin, double* V0, int* npinned) {
    const int n = o->topo->host.n;
    if (!islands_in_lds(n)) o->scratch.alloc((size_t)ISLAND_CHUNK * 2 * n);
    if (islands_need_z(o)) o->zscratch.alloc((size_t)ISLAND_CHUNK * n);
    launch_islands(o, c0, nc, S, ISLAND_CHUNK, o->scratch.p, o->zscratch.p, dVstart, served, refbus, pin, V0,
                   npinned);
}

// ---- islands without allocation ----------------------------------------------------
// The union-find scratch of p3smi_outages_islands_noalloc lives in this array, one copy per
// GPU for the life of the process. Its device address comes from hipGetSymbolAddress (once
// per device, with that device current); host code never passes the array's name itself.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;  // held for a whole no-allocation call: calls in this process never overlap
void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of the current device's copy of `scratch` (caller holds g_noalloc_mu).
unsigned char* noalloc_scratch() {
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) throw Error(P3SMI_E_DEVICE, "noalloc: device index out of range");
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(P3SMI_E_DEVICE, std::string("noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_noalloc_ptr[dev] = p;
    }
    return (unsigned char*)g_noalloc_ptr[dev];
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of scratch `k` cases of n buses need: parents and references, then (re-slacking) the
// |Z| words at a 256-byte boundary.
size_t noalloc_island_bytes(size_t k, size_t n, bool z) { return align256(k * 2 * n * sizeof(int)) + (z ? k * n * 8 : 0); }

}  // namespace

extern "C" {

int p3smi_outages_create(p3smi_topology* t, int32_t nbr, const int32_t* from, const int32_t* to, const int32_t* ypos,
                         const double* yval, int32_t n_cases, const int32_t* gptr, const int32_t* gbr, int32_t nref,
                         const int32_t* ref, const double* gen_z, int32_t reslack, p3smi_outages** out) {
    return guarded([&] {
        const p3smi::Topology& h = t->host;
        if (nbr < 0 || n_cases < 0 || nref < 0) throw Error(P3SMI_E_ARG, "outages: negative size");
        for (int k = 0; k < nbr; ++k) {
            if (from[k] < 0 || from[k] >= h.n || to[k] < 0 || to[k] >= h.n)
                throw Error(P3SMI_E_ARG, "outages: branch end out of range");
            for (int j = 0; j < 4; ++j)
                if (ypos[4 * k + j] < 0 || ypos[4 * k + j] >= h.nnzY)
                    throw Error(P3SMI_E_ARG, "outages: stamp position out of range");
        }
        if (gptr[0] != 0) throw Error(P3SMI_E_ARG, "outages: gptr[0] must be 0");
        for (int c = 0; c < n_cases; ++c) {
            if (gptr[c + 1] < gptr[c]) throw Error(P3SMI_E_ARG, "outages: gptr not monotone");
            if (gptr[c + 1] - gptr[c] > MAX_GROUP)
                throw Error(P3SMI_E_ARG, "outages: an outage case removes more than 256 branches");
        }
        for (int q = 0; q < (n_cases ? gptr[n_cases] : 0); ++q)
            if (gbr[q] < 0 || gbr[q] >= nbr) throw Error(P3SMI_E_ARG, "outages: branch index out of range");
        for (int i = 0; i < nref; ++i)
            if (ref[i] < 0 || ref[i] >= h.n) throw Error(P3SMI_E_ARG, "outages: slack bus out of range");
        std::unique_ptr<p3smi_outages> o(new p3smi_outages());
        o->topo = t;
        o->nbr = nbr; o->ncases = n_cases; o->nref = nref; o->reslack = reslack ? 1 : 0;
        ensure_device(t);
        hipStream_t st = t->stream;
        std::vector<unsigned char> live(nbr);
        for (int k = 0; k < nbr; ++k)
            live[k] = (yval[2 * (4 * k + 1)] != 0.0 || yval[2 * (4 * k + 1) + 1] != 0.0 ||
                       yval[2 * (4 * k + 2)] != 0.0 || yval[2 * (4 * k + 2) + 1] != 0.0);
        std::vector<int> gb;
        std::vector<double> gzv;
        if (gen_z)
            for (int i = 0; i < h.n; ++i)
                if (std::isfinite(gen_z[i])) {
                    if (gen_z[i] < 0) throw Error(P3SMI_E_ARG, "outages: generator |Z| must be >= 0");
                    gb.push_back(i);
                    gzv.push_back(gen_z[i]);
                }
        o->ngen = (int)gb.size();
        o->f.upload(from, nbr, st);
        o->t.upload(to, nbr, st);
        o->live.upload(live.data(), live.size(), st);
        o->ypos.upload(ypos, 4 * (size_t)nbr, st);
        o->yval.upload(yval, 8 * (size_t)nbr, st);
        o->gptr.upload(gptr, (size_t)n_cases + 1, st);
        o->gbr.upload(gbr, (size_t)(n_cases ? gptr[n_cases] : 0), st);
        o->ref.upload(ref, nref, st);
        o->gbus.upload(gb.data(), gb.size(), st);
        o->gz.upload(gzv.data(), gzv.size(), st);
        // each case's Ybus deltas: the four stamp entries of every outaged branch, in group order
        {
            const int ng = n_cases ? gptr[n_cases] : 0;
            std::vector<int> dptr((size_t)n_cases + 1), dpos(std::max(4 * ng, 1));
            std::vector<double> dval(2 * (size_t)std::max(4 * ng, 1));
            for (int c = 0; c <= n_cases; ++c) dptr[c] = 4 * gptr[c];
            for (int q = 0; q < ng; ++q)
                for (int j = 0; j < 4; ++j) {
                    const int k = gbr[q];
                    dpos[4 * q + j] = ypos[4 * k + j];
                    dval[2 * (4 * (size_t)q + j)] = yval[2 * (4 * (size_t)k + j)];
                    dval[2 * (4 * (size_t)q + j) + 1] = yval[2 * (4 * (size_t)k + j) + 1];
                }
            o->dptr.upload(dptr.data(), dptr.size(), st);
            o->dpos.upload(dpos.data(), dpos.size(), st);
            o->dval.upload(dval.data(), dval.size(), st);
        }
        HIPCHECK(hipStreamSynchronize(st));
        *out = o.release();
    });
}

void p3smi_outages_free(p3smi_outages* o) { delete o; }

int p3smi_outages_islands(p3smi_outages* o, uint8_t* served, int32_t* ref_bus, int32_t* pinned) {
    return guarded([&] {
        p3smi_topology* t = o->topo;
        std::lock_guard<std::mutex> lk(t->mu);
        const int n = t->host.n, L = o->ncases;
        if (L == 0) return;
        hipStream_t st = t->stream;
        DevBuf<unsigned char> dsrv;
        DevBuf<int> dref, dnp;
        DevBuf<double> dv;
        dsrv.alloc((size_t)n * L);
        dref.alloc((size_t)n * L);
        dnp.alloc(L);
        dv.alloc(2 * (size_t)n);
        HIPCHECK(hipMemsetAsync(dnp.p, 0, sizeof(int) * L, st));
        HIPCHECK(hipMemsetAsync(dv.p, 0, sizeof(double) * 2 * n, st));
        run_islands(o, 0, L, L, dv.p, dsrv.p, dref.p, nullptr, nullptr, dnp.p);
        if (served) HIPCHECK(hipMemcpyAsync(served, dsrv.p, (size_t)n * L, hipMemcpyDeviceToHost, st));
        if (ref_bus) HIPCHECK(hipMemcpyAsync(ref_bus, dref.p, (size_t)n * L * sizeof(int), hipMemcpyDeviceToHost, st));
        if (pinned) HIPCHECK(hipMemcpyAsync(pinned, dnp.p, (size_t)L * sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    });
}

int p3smi_outages_islands_noalloc(p3smi_outages* o, uint8_t* served, int32_t* ref_bus, int32_t* pinned) {
    return guarded([&] {
        if (!o || !served || !pinned) throw Error(P3SMI_E_ARG, "islands_noalloc: null argument");
        p3smi_topology* t = o->topo;
        std::lock_guard<std::mutex> lk(t->mu);
        std::lock_guard<std::mutex> lk2(g_noalloc_mu);
        if (!t->dev.ready || !t->stream) throw Error(P3SMI_E_ARG, "islands_noalloc: the topology is not resident");
        const size_t n = (size_t)t->host.n;
        const int L = o->ncases;
        if (L == 0) return;
        const bool z = islands_need_z(o);
        if (noalloc_island_bytes(1, n, z) > NOALLOC_SCRATCH_BYTES)
            throw Error(P3SMI_E_ARG, "islands_noalloc: a grid of " + std::to_string(n) + " buses needs " +
                                         std::to_string(noalloc_island_bytes(1, n, z)) +
                                         " bytes of scratch per case, above the " +
                                         std::to_string(NOALLOC_SCRATCH_BYTES) + "-byte no-allocation buffer");
        size_t chunk = NOALLOC_SCRATCH_BYTES / (n * (z ? 16 : 8));
        while (chunk > 1 && noalloc_island_bytes(chunk, n, z) > NOALLOC_SCRATCH_BYTES) --chunk;
        chunk = std::min(chunk, (size_t)L);
        unsigned char* base = noalloc_scratch();
        int* uf = (int*)base;
        unsigned long long* zb = z ? (unsigned long long*)(base + align256(chunk * 2 * n * sizeof(int))) : nullptr;
        hipStream_t st = t->stream;
        // the re-slack counters are summed with atomics: cleared by a kernel, not a memset
        hipLaunchKernelGGL(k_clear_int, dim3((unsigned)((L + 255) / 256)), dim3(256), 0, st, pinned, L);
        HIPCHECK(hipGetLastError());
        launch_islands(o, 0, L, L, (int)chunk, uf, zb, nullptr, served, ref_bus, nullptr, nullptr, pinned);
        HIPCHECK(hipStreamSynchronize(st));
    });
}

int p3smi_outages_solve(p3smi_outages* o, const double* S, const double* V_start, int32_t max_iter, double tol,
                        int32_t line_search, int32_t precision, double* V_out, int32_t* iterations,
                        uint8_t* converged, double* mismatch, uint8_t* served) {
    return guarded([&] {
        p3smi_topology* t = o->topo;
        std::lock_guard<std::mutex> lk(t->mu);
        const p3smi::Topology& h = t->host;
        const int L = o->ncases;
        if (L <= 0) throw Error(P3SMI_E_ARG, "outages: no cases");
        if (max_iter <= 0) throw Error(P3SMI_E_ARG, "solve_batch_contingency: max_iter must be > 0");
        if (!(tol > 0.0)) throw Error(P3SMI_E_ARG, "solve_batch_contingency: tol must be > 0");
        check_v0(V_start, h.n, 1, "solve_batch_contingency: ");
        DevBuf<double> dvs;
        dvs.upload(V_start, 2 * (size_t)h.n, t->stream);
        auto up = [&](p3smi_batch* b, int c0, int nc) {
            hipStream_t st = t->stream;
            b->Ssp.upload(S, 2 * (size_t)h.n, st);
            b->s_stride = 0;
            b->V0.alloc(2 * (size_t)h.n * b->S);
            b->v0_stride = b->S;
            b->pin.alloc((size_t)h.n * b->S);
            b->served.alloc((size_t)h.n * b->S);
            HIPCHECK(hipMemsetAsync(b->pin.p, 0, (size_t)h.n * b->S, st));
            o->npinned.alloc(b->S);
            HIPCHECK(hipMemsetAsync(o->npinned.p, 0, sizeof(int) * b->S, st));
            run_islands(o, c0, nc, b->S, dvs.p, b->served.p, nullptr, b->pin.p, b->V0.p, o->npinned.p);
            // each case's Ybus is the base minus its outaged stamps (case_generator.py:319-325),
            // subtracted in group order as the evaluation reads the row
            b->per_case_y = false;
            b->ydptr = o->dptr.p;
            b->ydpos = o->dpos.p;
            b->ydval = o->dval.p;
            b->ydc0 = c0;
            b->has_pin = true;
            b->mask_unserved = true;
            if (served) {
                d2h_rows(b->stage, served + c0, L, b->served.p, b->S, nc, h.n, st);
            }
        };
        o->res_batch = nullptr;
        const int chunks = solve_chunked(t, L, false, true, false, up, max_iter, tol, line_search, precision, V_out,
                                         iterations, converged, mismatch);
        if (chunks == 1) {  // the whole set's results stay on the device for post-processing
            o->res_batch = t->cached.get();
            o->res_id = t->cached->solve_id;
        }
    });
}

const p3smi_batch* resident_results(const p3smi_outages* o) {
    const p3smi_batch* b = o->res_batch;
    if (!b || b != o->topo->cached.get() || b->solve_id != o->res_id)
        throw Error(P3SMI_E_ARG, "outages: no resident results (solve the set in one chunk first; a later one-call "
                                 "solve on the topology replaces them)");
    return b;
}

int p3smi_outages_branch_flows(p3smi_outages* o, const double* vn_from_kv, const double* vn_to_kv, const double* rating,
                               const uint8_t* trafo, double sn_mva, double threshold, p3smi_branch_table* table,
                               double* max_loading, int32_t* worst_case, int32_t cap, int32_t* ov_branch,
                               int32_t* ov_case, double* ov_loading, double* ov_i_ka, int32_t* n_over) {
    return guarded([&] {
        if (!o || !vn_from_kv || !vn_to_kv || !rating || !trafo) throw Error(P3SMI_E_ARG, "branch_flows: null argument");
        if (cap < 0 || (cap > 0 && (!ov_branch || !ov_case || !ov_loading || !ov_i_ka)))
            throw Error(P3SMI_E_ARG, "branch_flows: violation list needs its four outputs");
        std::lock_guard<std::mutex> lk(o->topo->mu);
        const p3smi_batch* b = resident_results(o);
        hipStream_t st = o->topo->stream;
        const int nbr = o->nbr, L = o->ncases;
        const size_t NL = (size_t)nbr * L;
        DevBuf<double> vnf, vnt, rat, ld, ik, pfb, qfb, ptb, qtb, ifb, itb, best;
        DevBuf<unsigned char> tr;
        DevBuf<int> arg, cnt, ob, oc;
        DevBuf<double> ol, oi;
        vnf.upload(vn_from_kv, nbr, st);
        vnt.upload(vn_to_kv, nbr, st);
        rat.upload(rating, nbr, st);
        tr.upload(trafo, nbr, st);
        ld.alloc(NL);
        ik.alloc(NL);
        const bool full = table && table->p_from && table->q_from && table->p_to && table->q_to;
        const bool cur = table && table->i_from && table->i_to;
        if (full) { pfb.alloc(NL); qfb.alloc(NL); ptb.alloc(NL); qtb.alloc(NL); }
        if (cur) { ifb.alloc(NL); itb.alloc(NL); }
        const dim3 grid((unsigned)((L + 255) / 256), (unsigned)std::max(1, std::min(nbr, 65535)));
        if (nbr > 0)
            hipLaunchKernelGGL(k_branch_flows, grid, dim3(256), 0, st, o->f.p, o->t.p, o->yval.p, o->gptr.p, o->gbr.p,
                               b->Ve.p, b->Vf.p, b->served.p, b->ctl.p, vnf.p, vnt.p, rat.p, tr.p, sn_mva,
                               nbr, L, b->S, full ? pfb.p : nullptr, full ? qfb.p : nullptr, full ? ptb.p : nullptr,
                               full ? qtb.p : nullptr, cur ? ifb.p : nullptr, cur ? itb.p : nullptr, ik.p, ld.p);
        HIPCHECK(hipGetLastError());
        if (max_loading || worst_case) {
            best.alloc(nbr);
            arg.alloc(nbr);
            if (nbr > 0)
                hipLaunchKernelGGL(k_branch_worst, dim3((unsigned)nbr), dim3(256), 0, st, ld.p, L, best.p, arg.p);
            HIPCHECK(hipGetLastError());
            if (max_loading) HIPCHECK(hipMemcpyAsync(max_loading, best.p, nbr * sizeof(double), hipMemcpyDeviceToHost, st));
            if (worst_case) HIPCHECK(hipMemcpyAsync(worst_case, arg.p, nbr * sizeof(int), hipMemcpyDeviceToHost, st));
        }
        if (n_over) {
            cnt.alloc(1);
            HIPCHECK(hipMemsetAsync(cnt.p, 0, sizeof(int), st));
            ob.alloc(std::max(cap, 1)); oc.alloc(std::max(cap, 1)); ol.alloc(std::max(cap, 1)); oi.alloc(std::max(cap, 1));
            if (nbr > 0)
                hipLaunchKernelGGL(k_over_list, grid, dim3(256), 0, st, ld.p, ik.p, nbr, L, threshold, cap, cnt.p, ob.p,
                                   oc.p, ol.p, oi.p);
            HIPCHECK(hipGetLastError());
            int total = 0;
            HIPCHECK(hipMemcpyAsync(&total, cnt.p, sizeof(int), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            const int k = std::min(total, cap);
            if (k > 0) {
                HIPCHECK(hipMemcpyAsync(ov_branch, ob.p, k * sizeof(int), hipMemcpyDeviceToHost, st));
                HIPCHECK(hipMemcpyAsync(ov_case, oc.p, k * sizeof(int), hipMemcpyDeviceToHost, st));
                HIPCHECK(hipMemcpyAsync(ov_loading, ol.p, k * sizeof(double), hipMemcpyDeviceToHost, st));
                HIPCHECK(hipMemcpyAsync(ov_i_ka, oi.p, k * sizeof(double), hipMemcpyDeviceToHost, st));
            }
            *n_over = total;
        }
        if (table) {
            auto down = [&](double* dst, const DevBuf<double>& src) {
                if (dst && src.p) HIPCHECK(hipMemcpyAsync(dst, src.p, NL * sizeof(double), hipMemcpyDeviceToHost, st));
            };
            down(table->loading, ld);
            down(table->i_ka, ik);
            if (full) { down(table->p_from, pfb); down(table->q_from, qfb); down(table->p_to, ptb); down(table->q_to, qtb); }
            if (cur) { down(table->i_from, ifb); down(table->i_to, itb); }
        }
        HIPCHECK(hipStreamSynchronize(st));
    });
}

int p3smi_outages_voltage_violations(p3smi_outages* o, const double* vm_min, const double* vm_max,
                                     const uint8_t* in_service, double tol, int32_t cap, int32_t* bus, int32_t* case_,
                                     double* vm, uint8_t* kind, int32_t* n_viol) {
    return guarded([&] {
        if (!o || !vm_min || !vm_max || !in_service || !n_viol) throw Error(P3SMI_E_ARG, "voltage_violations: null argument");
        if (!(tol >= 0.0) || !std::isfinite(tol)) throw Error(P3SMI_E_ARG, "voltage_violations: tol must be finite and >= 0");
        if (cap < 0 || (cap > 0 && (!bus || !case_ || !vm || !kind)))
            throw Error(P3SMI_E_ARG, "voltage_violations: the list needs its four outputs");
        std::lock_guard<std::mutex> lk(o->topo->mu);
        const p3smi::Topology& h = o->topo->host;
        for (int i = 0; i < h.n; ++i)
            if (vm_min[i] > vm_max[i]) throw Error(P3SMI_E_ARG, "voltage_violations: vm_min_pu > vm_max_pu at bus " + std::to_string(i));
        const p3smi_batch* b = resident_results(o);
        hipStream_t st = o->topo->stream;
        const int L = o->ncases;
        DevBuf<double> lo, hi, ov;
        DevBuf<unsigned char> ins, ok;
        DevBuf<int> cnt, ob, oc;
        lo.upload(vm_min, h.n, st);
        hi.upload(vm_max, h.n, st);
        ins.upload(in_service, h.n, st);
        cnt.alloc(1);
        HIPCHECK(hipMemsetAsync(cnt.p, 0, sizeof(int), st));
        ob.alloc(std::max(cap, 1)); oc.alloc(std::max(cap, 1)); ov.alloc(std::max(cap, 1)); ok.alloc(std::max(cap, 1));
        const dim3 grid((unsigned)((L + 255) / 256), (unsigned)std::min(h.n, 65535));
        hipLaunchKernelGGL(k_vviol, grid, dim3(256), 0, st, b->Ve.p, b->Vf.p, b->served.p, b->ctl.p, lo.p,
                           hi.p, ins.p, tol, h.n, L, b->S, cap, cnt.p, ob.p, oc.p, ov.p, ok.p);
        HIPCHECK(hipGetLastError());
        int total = 0;
        HIPCHECK(hipMemcpyAsync(&total, cnt.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        const int k = std::min(total, cap);
        if (k > 0) {
            HIPCHECK(hipMemcpyAsync(bus, ob.p, k * sizeof(int), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipMemcpyAsync(case_, oc.p, k * sizeof(int), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipMemcpyAsync(vm, ov.p, k * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipMemcpyAsync(kind, ok.p, k, hipMemcpyDeviceToHost, st));
        }
        HIPCHECK(hipStreamSynchronize(st));
        *n_viol = total;
    });
}

}  // extern "C"

// ---- continuous batching: a resident batch used as a pool of slots ------------------
extern "C" {

int p3smi_batch_slots_init(p3smi_batch* b) {
    return guarded([&] {
        const p3smi::Topology& h = b->topo->host;
        hipStream_t st = b->topo->stream;
        b->Ssp.alloc(2 * (size_t)h.n * b->S);
        b->V0.alloc(2 * (size_t)h.n * b->S);
        b->s_stride = b->S;
        b->v0_stride = b->S;
        b->per_case_y = false;
        b->ydptr = nullptr;
        b->has_pin = false;
        b->mask_unserved = false;
        b->slot_mode = true;
        ensure_tiles(b->topo);
        hipLaunchKernelGGL(k_ctl_reset, dim3((unsigned)((b->S + 255) / 256)), dim3(256), 0, st, b->ctl.p, 0, b->S, 0u,
                           0u);
        hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, st, b->guard.p, 6);
        pivot_release_all(b);
        HIPCHECK(hipGetLastError());
        std::memset(b->h_counts, 0, NSTAT * sizeof(int));
        b->h_counts[CNT_IDLE] = b->S;
        HIPCHECK(hipStreamSynchronize(st));
        b->uploaded = true;
        b->pending_start = false;
    });
}

int p3smi_batch_load(p3smi_batch* b, int32_t k, const int32_t* slots, const double* S, const double* V0) {
    return guarded([&] {
        const p3smi::Topology& h = b->topo->host;
        hipStream_t st = b->topo->stream;
        if (k <= 0) return;
        for (int j = 0; j < k; ++j)
            if (slots[j] < 0 || slots[j] >= b->B) throw Error(P3SMI_E_ARG, "load: slot out of range");
        check_v0(V0, h.n, k, "solve_batch: ");
        b->st_S.upload(S, 2 * (size_t)h.n * k, st);
        b->st_V.upload(V0, 2 * (size_t)h.n * k, st);
        b->st_slots.upload(slots, k, st);
        hipLaunchKernelGGL(k_load_slots, dim3((unsigned)((k + 63) / 64), (unsigned)((h.n + 3) / 4)), dim3(256), 0, st,
                           b->st_S.p, b->st_V.p, b->st_slots.p, b->Ssp.p, b->V0.p, b->ctl.p, start_extra(b->topo), h.n,
                           k, b->S);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(st));
        b->pending_start = true;
        b->h_counts[CNT_START] += k;
        b->h_counts[CNT_IDLE] -= k;
    });
}

int p3smi_batch_pass(p3smi_batch* b, int32_t max_iter, double tol, int32_t line_search, int32_t precision,
                     int32_t* counts) {
    return guarded([&] {
        if (!b->uploaded) throw Error(P3SMI_E_ARG, "pass: batch slots not initialised");
        check_solve_args(max_iter, tol, precision);
        b->precision = precision;
        const int* cnt = b->h_counts;
        if (live_cases(cnt) > 0 || b->pending_start)
            cnt = run_pass(b, PassParams{max_iter, line_search ? 1 : 0, precision, tol, false});
        if (counts) std::copy(cnt, cnt + NSTAT, counts);
        collect_timing(b);
    });
}

int p3smi_batch_finished(p3smi_batch* b, int32_t* slots, int32_t* n) {
    return guarded([&] {
        std::vector<unsigned> ctl(b->B);
        HIPCHECK(hipMemcpyAsync(ctl.data(), b->ctl.p, b->B * sizeof(unsigned), hipMemcpyDeviceToHost, b->topo->stream));
        HIPCHECK(hipStreamSynchronize(b->topo->stream));
        int k = 0;
        for (int c = 0; c < b->B; ++c) {
            const unsigned sg = stage_of(ctl[c]);
            if (sg == SG_CONV || sg == SG_FAIL || sg == SG_STOP) slots[k++] = c;
        }
        *n = k;
    });
}

int p3smi_batch_fetch(p3smi_batch* b, int32_t k, const int32_t* slots, double* V, int32_t* iterations,
                      uint8_t* converged, double* mismatch) {
    return guarded([&] {
        const p3smi::Topology& h = b->topo->host;
        hipStream_t st = b->topo->stream;
        if (k <= 0) return;
        for (int j = 0; j < k; ++j)
            if (slots[j] < 0 || slots[j] >= b->B) throw Error(P3SMI_E_ARG, "fetch: slot out of range");
        b->st_slots.upload(slots, k, st);
        b->st_out.alloc(2 * (size_t)h.n * k);
        b->st_it.alloc(k);
        b->st_cv.alloc(k);
        b->st_mis.alloc(k);
        hipLaunchKernelGGL(k_fetch_slots, dim3((unsigned)((k + 63) / 64), (unsigned)((h.n + 3) / 4)), dim3(256), 0, st,
                           b->st_slots.p, b->Ve.p, b->Vf.p, b->ctl.p, b->nrm.p, b->st_out.p, b->st_it.p, b->st_cv.p,
                           b->st_mis.p, h.n, k, b->S);
        hipLaunchKernelGGL(k_release_slots, dim3((unsigned)((k + 255) / 256)), dim3(256), 0, st, b->st_slots.p,
                           b->ctl.p, k);
        for (int j = 0; j < k; ++j) pivot_release(b, slots[j]);
        HIPCHECK(hipGetLastError());
        if (V) HIPCHECK(hipMemcpyAsync(V, b->st_out.p, 2 * (size_t)h.n * k * sizeof(double), hipMemcpyDeviceToHost, st));
        if (iterations) HIPCHECK(hipMemcpyAsync(iterations, b->st_it.p, k * sizeof(int), hipMemcpyDeviceToHost, st));
        std::vector<unsigned char> cv(k);
        HIPCHECK(hipMemcpyAsync(cv.data(), b->st_cv.p, k, hipMemcpyDeviceToHost, st));
        if (mismatch) HIPCHECK(hipMemcpyAsync(mismatch, b->st_mis.p, k * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        if (converged) std::copy(cv.begin(), cv.end(), converged);
    });
}

}  // extern "C"

// ===================================================================== batch groups
// Batches of different topologies and sizes solved together (uneven batching): every pass is one
// launch per kernel -- per LU launch shape and per evaluation variant -- over all the batches,
// each launch's workgroups dealt to the batches by a segment table (seg_find). The passes run on
// the device: each batch's k_control writes the plan of its next pass (plan_from_counts, the
// decisions of the host's pass loop), every kernel reads it, and a batch with nothing to do
// halts. K passes are captured into one graph, so the host synchronises once per K passes.
// A batch with a case waiting for a pivoting factorization halts; the host runs that pass for it
// alone (run_pass) and the group goes on. FP64 solves of whole batches (lock-step passes).
struct GroupLaunch {
    int kind = 0;  // LU: 0 one-entry tasks (k_lu_stream_multi), 1 2x2 blocks (k_lu_blk_multi);
                   // evaluation: 2 * admittance mode (YM_*) + pin mask
    int G = 0, NS = 0;
    DevBuf<char> tab;
    DevBuf<int> off;
    int nseg = 0, blocks = 0;
};

struct p3smi_group {
    std::vector<p3smi_batch*> mem;
    hipStream_t stream = nullptr;
    DevBuf<int> plans;   // NPLAN per member
    DevBuf<int> counts;  // NCOUNT per member (k_control's counters)
    int* h_plans = nullptr;
    int* h_counts = nullptr;
    std::vector<std::unique_ptr<GroupLaunch>> lu, eval;
    GroupLaunch begin, adv, ctl;
    std::string built;   // solve parameters and buffers the tables were built for
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_passes = 0;
    long long device_passes = 0, eager_passes = 0, graph_launches = 0;
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_passes = 0;
    }
    ~p3smi_group() {
        drop_graph();
        if (h_plans) (void)hipHostFree(h_plans);
        if (h_counts) (void)hipHostFree(h_counts);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Segment table of one launch: args[i] with blocks[i] workgroups, ranges padded to 8.
template <class A>
void group_table(GroupLaunch& L, const std::vector<A>& args, const std::vector<int>& blocks, hipStream_t st) {
    std::vector<int> off(args.size() + 1, 0);
    for (size_t i = 0; i < args.size(); ++i) {
        if (blocks[i] <= 0) throw Error(P3SMI_E_INTERNAL, "group: a segment without workgroups");
        off[i + 1] = off[i] + (blocks[i] + 7) / 8 * 8;
    }
    L.tab.upload(reinterpret_cast<const char*>(args.data()), args.size() * sizeof(A), st);
    L.off.upload(off.data(), off.size(), st);
    L.nseg = (int)args.size();
    L.blocks = off.back();
    HIPCHECK(hipStreamSynchronize(st));  // the host vectors go out of scope
}

// workgroups along the buses of k_begin_multi / k_advance_multi (each thread steps through the rest)
int group_bus_blocks(int n) { return std::min((n + 3) / 4, 64); }

void group_build(p3smi_group* g, const PassParams& p) {
    hipStream_t st = g->stream;
    const int nm = (int)g->mem.size();
    g->lu.clear();
    g->eval.clear();
    g->drop_graph();
    std::map<std::tuple<int, int, int>, std::pair<std::vector<LuArgs>, std::vector<int>>> lu;
    std::map<int, std::pair<std::vector<EvalArgs>, std::vector<int>>> ev;
    std::vector<BeginArgs> bg;
    std::vector<AdvanceArgs> ad;
    std::vector<CtlArgs> ct;
    std::vector<int> bgb, adb, ctb;
    for (int i = 0; i < nm; ++i) {
        p3smi_batch* b = g->mem[i];
        p3smi_topology* t = b->topo;
        const p3smi::Topology& h = t->host;
        const int* plan = g->plans.p + (size_t)i * NPLAN;
        const LuConfig cfgs[2] = {t->lu_batch, t->lu_few};
        for (int w = 0; w < 2; ++w) {
            const LuConfig cfg = cfgs[w];
            if (!valid_lu_config(cfg)) throw Error(P3SMI_E_ARG, "group: unsupported LU launch shape");
            LuArgs a = lu_args(b, cfg, get_sched(t, cfg), b->fac.p, b->R.p, b->x.p, SEL_ALL);
            a.plan = plan;
            a.want = w + 1;
            auto& e = lu[std::make_tuple(cfg.blk() ? 1 : 0, cfg.G, cfg.NS)];
            e.first.push_back(a);
            e.second.push_back(a.gx);
        }
        const dim3 gi((unsigned)(b->S / 64), (unsigned)group_bus_blocks(h.n));
        BeginArgs ba = begin_args(b, p.precision, gi);
        ba.plan = plan;
        bg.push_back(ba);
        bgb.push_back(ba.gx * ba.gy);
        AdvanceArgs aa = advance_args(b, 0, gi);
        aa.plan = plan;
        ad.push_back(aa);
        adb.push_back(aa.gx * aa.gy);
        EvalArgs ea = eval_args(b, 0);
        ea.plan = plan;
        auto& e = ev[2 * b->ymode() + (int)b->has_pin];
        e.first.push_back(ea);
        e.second.push_back(t->dev.ntiles * (b->S / 64));
        CtlArgs ca = ctl_args(b, p, 0);
        ca.counts_out = g->counts.p + (size_t)i * NCOUNT;
        ca.plan = g->plans.p + (size_t)i * NPLAN;
        ct.push_back(ca);
        ctb.push_back(ca.gx);
    }
    for (auto& kv : lu) {
        std::unique_ptr<GroupLaunch> L(new GroupLaunch());
        L->kind = std::get<0>(kv.first);
        L->G = std::get<1>(kv.first);
        L->NS = std::get<2>(kv.first);
        group_table(*L, kv.second.first, kv.second.second, st);
        g->lu.push_back(std::move(L));
    }
    for (auto& kv : ev) {
        std::unique_ptr<GroupLaunch> L(new GroupLaunch());
        L->kind = kv.first;
        group_table(*L, kv.second.first, kv.second.second, st);
        g->eval.push_back(std::move(L));
    }
    group_table(g->begin, bg, bgb, st);
    group_table(g->adv, ad, adb, st);
    group_table(g->ctl, ct, ctb, st);
}

template <int G, int NS>
void group_lu_launch(const GroupLaunch& L, hipStream_t st) {
    const LuArgs* tab = reinterpret_cast<const LuArgs*>(L.tab.p);
    if constexpr (G == 64) {
        if (L.kind == 1) {
            hipLaunchKernelGGL((k_lu_blk_multi<NS, double, true>), dim3((unsigned)L.blocks), dim3(64 * NS), 0, st, tab,
                               L.off.p, L.nseg);
            return;
        }
    }
    hipLaunchKernelGGL((k_lu_stream_multi<G, NS, double, true>), dim3((unsigned)L.blocks), dim3(64 * NS), 0, st, tab,
                       L.off.p, L.nseg);
}

// One pass over every member (the kernels a batch's plan does not ask for return at once).
void group_pass(p3smi_group* g) {
    hipStream_t st = g->stream;
    for (const auto& L : g->lu) {
#define P3SMI_GLU(GG, NN)                          \
    if (L->G == GG && L->NS == NN) {               \
        group_lu_launch<GG, NN>(*L, st);           \
        continue;                                  \
    }
        P3SMI_GLU(64, 8) P3SMI_GLU(64, 16) P3SMI_GLU(32, 8) P3SMI_GLU(32, 16) P3SMI_GLU(16, 8) P3SMI_GLU(16, 16)
        P3SMI_GLU(1, 4) P3SMI_GLU(1, 8) P3SMI_GLU(1, 16)
#undef P3SMI_GLU
        throw Error(P3SMI_E_ARG, "group: unsupported LU launch shape");
    }
    hipLaunchKernelGGL(k_begin_multi, dim3((unsigned)g->begin.blocks), dim3(256), 0, st,
                       reinterpret_cast<const BeginArgs*>(g->begin.tab.p), g->begin.off.p, g->begin.nseg);
    hipLaunchKernelGGL(k_advance_multi, dim3((unsigned)g->adv.blocks), dim3(256), 0, st,
                       reinterpret_cast<const AdvanceArgs*>(g->adv.tab.p), g->adv.off.p, g->adv.nseg);
    for (const auto& L : g->eval) {
        const EvalArgs* tab = reinterpret_cast<const EvalArgs*>(L->tab.p);
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3((unsigned)L->blocks), dim3(64 * EV_WAVES), 0, st, tab, L->off.p, L->nseg);
        };
        switch (L->kind) {
            case 0: go(k_evaluate_multi<YM_SHARED, false>); break;
            case 1: go(k_evaluate_multi<YM_SHARED, true>); break;
            case 2: go(k_evaluate_multi<YM_CASE, false>); break;
            case 3: go(k_evaluate_multi<YM_CASE, true>); break;
            case 4: go(k_evaluate_multi<YM_DELTA, false>); break;
            default: go(k_evaluate_multi<YM_DELTA, true>); break;
        }
    }
    hipLaunchKernelGGL(k_control_multi, dim3((unsigned)g->ctl.blocks), dim3(256), 0, st,
                       reinterpret_cast<const CtlArgs*>(g->ctl.tab.p), g->ctl.off.p, g->ctl.nseg);
    HIPCHECK(hipGetLastError());
}

void group_solve(p3smi_group* g, int max_iter, double tol, int line_search, int precision, int K) {
    check_solve_args(max_iter, tol, precision);
    if (precision != P3SMI_PREC_DOUBLE) throw Error(P3SMI_E_ARG, "group solve: precision must be 0 (double)");
    if (K < 1 || K > 64) throw Error(P3SMI_E_ARG, "group solve: passes per graph must be in [1, 64]");
    const int nm = (int)g->mem.size();
    const PassParams p{max_iter, line_search ? 1 : 0, precision, tol, true};
    std::string key = std::to_string(max_iter) + "/" + std::to_string(tol) + "/" + std::to_string(line_search);
    for (p3smi_batch* b : g->mem) {
        p3smi_topology* t = b->topo;
        if (!b->uploaded) throw Error(P3SMI_E_ARG, "group solve: a batch has no inputs (upload first)");
        if (b->slot_mode) throw Error(P3SMI_E_ARG, "group solve: a batch is in slot mode");
        ensure_tiles(t);
        if (b->part.n < (size_t)t->dev.ntiles * b->S) b->part.alloc((size_t)t->dev.ntiles * b->S);
        (void)get_sched(t, t->lu_batch);
        (void)get_sched(t, t->lu_few);
        std::memset(&b->stats, 0, sizeof b->stats);
        b->precision = precision;
        b->solve_id = ++g_solve_ids;
        b->stats.cases = b->B;
        reset_cases(b, precision);
        HIPCHECK(hipStreamSynchronize(t->stream));
        char buf[160];
        std::snprintf(buf, sizeof buf, "|%p/%p/%p/%p/%d/%d/%d/%d", (void*)b, (void*)b->part.p, (void*)b->Ssp.p,
                      (void*)b->V0.p, t->lu_batch.key(), t->lu_few.key(), t->lu_few_max, t->dev.tile_rows);
        key += buf;
    }
    if (g->built != key) {
        group_build(g, p);
        g->built = key;
    }
    if (g->graph_passes != K) {
        g->drop_graph();
        HIPCHECK(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
        try {
            for (int k = 0; k < K; ++k) group_pass(g);
        } catch (...) {
            hipGraph_t gr = nullptr;
            (void)hipStreamEndCapture(g->stream, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            throw;
        }
        HIPCHECK(hipStreamEndCapture(g->stream, &g->graph));
        if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) {
            g->drop_graph();
            throw Error(P3SMI_E_DEVICE, "group solve: hipGraphInstantiate failed");
        }
        g->graph_passes = K;
    }
    // every run of a case takes at most 1 + max_iter * (LS_MAX_TRIALS + 3) passes, two runs per case
    const long long cap = 2 * (1 + (long long)max_iter * (LS_MAX_TRIALS + 3)) + 2;
    long long rounds = 0;
    g->device_passes = g->eager_passes = g->graph_launches = 0;
    for (;;) {
        if (++rounds > cap) throw Error(P3SMI_E_INTERNAL, "group solve: passes did not finish every case");
        bool any_live = false, any_device = false;
        for (int i = 0; i < nm; ++i) {
            p3smi_batch* b = g->mem[i];
            int* pl = g->h_plans + (size_t)i * NPLAN;
            plan_from_counts(b->h_counts, b->topo->lu_few_max, pl);
            if (live_cases(b->h_counts) == 0) continue;
            any_live = true;
            if (pl[PL_HALT]) {  // a case waits for its pivoting factorization: this pass on the host
                run_pass(b, p);
                g->eager_passes++;
                plan_from_counts(b->h_counts, b->topo->lu_few_max, pl);
            }
            if (!pl[PL_HALT]) any_device = true;
        }
        if (!any_live) break;
        if (!any_device) continue;
        HIPCHECK(hipMemcpyAsync(g->plans.p, g->h_plans, (size_t)nm * NPLAN * sizeof(int), hipMemcpyHostToDevice,
                                g->stream));
        HIPCHECK(hipGraphLaunch(g->exec, g->stream));
        HIPCHECK(hipMemcpyAsync(g->h_counts, g->counts.p, (size_t)nm * NCOUNT * sizeof(int), hipMemcpyDeviceToHost,
                                g->stream));
        HIPCHECK(hipStreamSynchronize(g->stream));
        g->graph_launches++;
        g->device_passes += K;
        for (int i = 0; i < nm; ++i) {
            p3smi_batch* b = g->mem[i];
            if (g->h_plans[(size_t)i * NPLAN + PL_HALT]) continue;  // the device did not touch it
            std::copy(g->h_counts + (size_t)i * NCOUNT, g->h_counts + (size_t)(i + 1) * NCOUNT, b->h_counts);
            b->pending_start = false;
        }
    }
    for (p3smi_batch* b : g->mem) {
        int gd[6];
        HIPCHECK(hipMemcpyAsync(gd, b->guard.p, sizeof gd, hipMemcpyDeviceToHost, g->stream));
        HIPCHECK(hipStreamSynchronize(g->stream));
        b->stats.failed_factorizations = b->h_counts[CNT_FAIL];
        b->stats.guard_trips = gd[0];
        b->stats.reruns = gd[3];
        b->stats.pivot_factorizations = gd[4];
        b->stats.diagonal_factorizations = gd[5];
    }
}

}  // namespace

extern "C" {

int p3smi_group_create(p3smi_batch* const* members, int32_t count, p3smi_group** out) {
    return guarded([&] {
        if (!out) throw Error(P3SMI_E_ARG, "group: null output");
        *out = nullptr;
        if (count <= 0 || !members) throw Error(P3SMI_E_ARG, "group: needs at least one batch");
        std::unique_ptr<p3smi_group> g(new p3smi_group());
        for (int i = 0; i < count; ++i) {
            if (!members[i]) throw Error(P3SMI_E_ARG, "group: null batch");
            for (p3smi_batch* b : g->mem)
                if (b == members[i]) throw Error(P3SMI_E_ARG, "group: a batch appears twice");
            g->mem.push_back(members[i]);
        }
        HIPCHECK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
        g->plans.alloc((size_t)count * NPLAN);
        g->counts.alloc((size_t)count * NCOUNT);
        HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&g->h_plans), (size_t)count * NPLAN * sizeof(int)));
        HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&g->h_counts), (size_t)count * NCOUNT * sizeof(int)));
        *out = g.release();
    });
}

void p3smi_group_free(p3smi_group* g) { delete g; }

int p3smi_group_solve(p3smi_group* g, int32_t max_iter, double tol, int32_t line_search, int32_t precision,
                      int32_t passes_per_graph) {
    return guarded([&] {
        if (!g) throw Error(P3SMI_E_ARG, "group solve: null group");
        std::vector<std::unique_lock<std::mutex>> locks;
        std::vector<p3smi_topology*> topos;
        for (p3smi_batch* b : g->mem)
            if (std::find(topos.begin(), topos.end(), b->topo) == topos.end()) topos.push_back(b->topo);
        std::sort(topos.begin(), topos.end());
        for (p3smi_topology* t : topos) locks.emplace_back(t->mu);
        group_solve(g, max_iter, tol, line_search, precision, passes_per_graph);
    });
}

int p3smi_group_stats(const p3smi_group* g, int64_t* out3) {
    return guarded([&] {
        if (!g || !out3) throw Error(P3SMI_E_ARG, "group stats: null argument");
        out3[0] = g->device_passes;
        out3[1] = g->eager_passes;
        out3[2] = g->graph_launches;
    });
}

}  // extern "C"
