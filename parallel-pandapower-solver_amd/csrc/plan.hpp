// plan.hpp -- host-side, once-per-topology setup of the batched polar Newton engine.
//
// Everything here depends only on the Ybus sparsity pattern and the pv/pq split, and is
// built once per topology on the host (the reference does the same setup in
// p3s/cpp/nr_klu.cpp:732-759 and p3s/cuda/polar_topology.py, lean_lu_plan.py):
//
//   * the polar topology: Ybus CSR, diagonal index, the P-row / Q-row of every bus;
//   * the Jacobian pattern (CSR by equation) and the map from each Ybus nonzero's four
//     derivative blocks to Jacobian entries (block ids 0 dP/dVa, 1 dP/dVm, 2 dQ/dVa,
//     3 dQ/dVm, as p3s/cpp/nr_klu.cpp:104-108);
//   * a symmetric fill-reducing order (minimum fill on the buses, P before Q), the
//     elimination tree and the static-pivot factor pattern;
//   * the level schedule the GPU executes: every factor entry is one task (a dot product
//     over precomputed operand pairs), tasks of one elimination-tree height run
//     concurrently, dealt to `nslots` task lists by longest-processing-time first.
//
// Factor layout (positions, per pivot p in elimination order, S_p = strict column
// pattern of L(:,p), rows ascending):
//     base[p]                      1 / U(p,p)
//     base[p] + 1 + r              U(p, S_p[r])
//     base[p] + 1 + |S_p| + r      L(S_p[r], p)      (unit lower triangle, scaled)
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace p3smi {

struct Task {      // one factor entry (or a chunk of its sum, see ONE_SPLIT_MIN)
    int dst;       // factor position written (a chunk other than the last: its partial slot)
    int asrc;      // Jacobian entry it starts from, -1 = fill (starts from 0)
    int pbeg, pend;  // operand pairs [pbeg, pend)
    int pin = -1;  // chunk: partial slot of the previous chunk (its sum negated), loaded first
    int part = 0;  // chunk other than the last: stores its sum, negated, to dst
};

// Long sums of the one-entry schedule (one accumulator per entry) are cut as the 2x2 blocks'
// are (BLK_SPLIT_MIN below): chunks of consecutive operand pairs in earlier sub-levels, the
// partial sum passed on negated through a slot of the factor arena and brought back by the
// pair (slot, one row). These slots and the block schedule's are never used by one launch,
// so they share the arena rows past the one row.
constexpr int ONE_SPLIT_MIN = 32;  // sums over more operand pairs than this are cut
constexpr int ONE_CHUNK = 16;
constexpr int ONE_TAIL = 8;

struct FwdRow {    // forward substitution row i: y_i = R[rsrc] - sum L(i,j) y_j
    int row, rsrc, ebeg, eend;
    int pin = -1;  // chunk: solution-vector slot (row index >= m) holding the previous chunk's sum, negated
    int dst = -1;  // chunk other than the last: the slot its sum goes to, negated
};

struct BwdRow {    // backward substitution row i: x_i = (y_i - sum U(i,j) x_j) / U(i,i), j descending
    int row, base, lbeg, lend;  // a chunk: columns [lbeg, lend) of row's S_row (still descending)
    int pin = -1;  // chunk: solution-vector slot of the previous chunk's sum, negated
    int dst = -1;  // chunk other than the last: the slot its sum goes to, negated
};

struct Schedule {
    int nslots = 0;
    int nlev = 0;   // factor / forward levels (elimination-tree heights)
    int nblev = 0;  // backward levels (elimination-tree depths)
    // per (level, slot): [ptr[l*nslots+s], ptr[l*nslots+s+1])
    std::vector<int> ptrA, ptrB, ptrF, ptrW;
    std::vector<Task> taskA;  // diagonal and U entries (diagonal marked by dst == base)
    std::vector<Task> taskB;  // L entries, scaled by 1/U(p,p)
    std::vector<int> diagB;   // factor position of the inverted pivot of each taskB
    std::vector<unsigned char> isdiagA;  // taskA entry is a pivot
    std::vector<FwdRow> fwd;
    std::vector<BwdRow> bwd;
    int nparts = 0;  // partial slots (one factor-arena row each, from nnz_lu + 2)
    int ncut = 0;    // entries cut into chunks
    int nxparts = 0; // solution-vector slots of cut forward / backward rows (rows m, m + 1, ...)
};

// ---- 2x2 supernode blocks
// Pivots p, p+1 form a pair when p+1 is p's elimination-tree parent and S_p = {p+1} + S_{p+1}
// (the P and Q unknowns of one PQ bus whenever the order puts them side by side, and the
// chain of single pivots at the top of the tree). A pivot group P (one pivot or a pair) owns
// the diagonal block P x P, a U block P x C and an L block C x P for every group C (a pair,
// or a single pivot) of its column pattern. A block (R, C) steps through the pivots t < P
// that contribute to any of its entries, in ascending order; a step loads L(R, t) and
// U(t, C) once -- |R| + |C| rows -- for |R| * |C| updates. Where one row or column of a pair
// has no entry at t, the step reads the zero row (position nnz_lu of the factor arena): the
// term is an exact zero, so every entry sums the same nonzero terms in the same ascending
// order as the one-entry-per-task schedule (identical bits while the factors are finite).
// The contribution of t = p to the second row or column of a pair comes last, in the finish.
enum BlockKind { BLK_DIAG1 = 0, BLK_DIAG2, BLK_U1, BLK_U2, BLK_L1, BLK_L2, BLK_PART };

// ---- chunks of long block sums
// A block whose sum runs over many pivots t would keep one task list busy while the others wait
// at the sub-level's barrier (the single long dot products at the top of the elimination tree).
// Such a sum is cut into chunks of consecutive t, in ascending order, placed in consecutive
// earlier sub-levels (each chunk where all its operands are already final). A chunk other than
// the last is a BLK_PART task: its finish stores its four accumulators, negated, to a partial
// slot (four rows of the factor arena past the zero row). The next chunk starts with one or two
// load steps that bring them back exactly: with -p in the L operands and the one row (1.0) or the
// zero row in the U operands, fma(p, 1, 0) = p and fma(p', 0, acc) = acc (an accumulator that
// starts at +0 and adds products with fma never becomes -0). Every entry therefore sums the same
// terms in the same order as an uncut block, so the factors keep their bits. Partial slots are
// reused once read; the cuts do not depend on the number of task lists.
// Forward and backward rows of a pivot group are cut the same way, over their steps in order:
// a chunk at the level before the next piece stores its one or two running sums to
// solution-vector rows past m, and the next piece starts its sums from them (the same
// additions in the same order as the uncut row).
constexpr int BLK_SPLIT_MIN = 32;  // sums over more pivots than this are cut
constexpr int BLK_CHUNK = 16;      // pivots per chunk (a leftover under BLK_TAIL joins its neighbour)
constexpr int BLK_TAIL = 8;
// In a thin sub-level (fewer than BLK_REF_LISTS x BLK_FINE_MIN pivots summed by all its tasks)
// a sum that sets the sub-level's length -- more pivots than its tasks hold on average over
// BLK_REF_LISTS lists -- is cut finer, from BLK_FINE_MIN pivots on into chunks of
// BLK_FINE_CHUNK. Such sums form the chain at the top of the elimination tree, where a level
// holds one or two pivots and the other lists wait on them. Wider sub-levels keep the coarse
// cut: their time follows the steps of all their tasks, and each cut adds steps.
constexpr int BLK_REF_LISTS = 16;
constexpr int BLK_FINE_MIN = 16;
constexpr int BLK_FINE_CHUNK = 8;
constexpr int BLK_FINE_TAIL = 4;

struct BlockTask {
    int kind;        // BlockKind
    int nr, nc;      // rows, columns of the block (1 or 2)
    int tgt0, tgt1;  // DIAG: 1/U(p,p), 1/U(p+1,p+1); U: U(p,C0), U(p+1,C0); L: L(R0,p), L(R0,p+1)
    int aux0, aux1;  // DIAG2: U(p,p+1), L(p+1,p); U2: L(p+1,p); L1: 1/U(p,p); L2: 1/U(p,p), 1/U(p+1,p+1)
    int src[4];      // Jacobian entries of (0,0) (0,1) (1,0) (1,1), -1 = fill or absent
    int sbeg, send;  // steps [sbeg, send)
};

// Steps of a block: positions of L(R0, t), L(R1, t), U(t, C0), U(t, C1); nnz_lu = the zero
// row; -1 for the second row (column) of a block that has one row (column).

// Forward rows of a pivot group: y(row..row+nr-1) = R - sum_j L(., j) y(j) over the columns j
// of either row (the zero row where one row has no entry); for a pair, the term
// L(row+1, row) y(row) comes last, in the finish.
struct FwdBlock {
    int row, nr;           // first row (pivot index), 1 or 2 rows
    int rsrc0, rsrc1;      // right-hand side entries (Jacobian row order)
    int lpos;              // pair: position of L(row+1, row)
    int sbeg, send;        // forward steps [sbeg, send)
    // a row cut into chunks (see BLK_SPLIT_MIN): pin = solution-vector slot (row >= m) whose
    // nr rows hold the sums of the previous chunk to start from, -1 none; pout = the slot a
    // chunk other than the last stores its sums to, -1 for the row's last piece
    int pin = -1, pout = -1;
};

// Backward rows of a pivot group, columns j in descending order (the order of every
// backward row): x(row+1) over S_{row+1}, and x(row) over the same columns followed, last,
// by U(row, row+1) x(row+1).
struct BwdBlock {
    int row, nr;           // first row (pivot index), 1 or 2 rows
    int base0, base1;      // 1/U(row,row) and (pair) 1/U(row+1,row+1) positions
    int sbeg, send;        // backward steps [sbeg, send)
    int pin = -1, pout = -1;  // cut row: solution-vector slots, as FwdBlock
};

struct BlockSchedule {
    int nslots = 0, nlev = 0, nblev = 0;
    int npairs = 0;  // pivot pairs
    int nparts = 0;  // partial slots (4 factor-arena rows each, from nnz_lu + 2)
    int ncut = 0;    // blocks cut into chunks
    int nxparts = 0; // solution-vector rows past m holding the cut substitution rows' sums
    int nfcut = 0, nbcut = 0;  // forward, backward rows cut into chunks
    // per (level, slot): sub-level A = diagonal blocks and U blocks of single pivots;
    // sub-level B = the other U and L blocks and the forward rows.
    std::vector<int> ptrA, ptrB, ptrF, ptrW;
    std::vector<BlockTask> taskA, taskB;  // steps of the tasks back to back in execution order
    std::vector<int> steps;               // 4 ints per step
    std::vector<FwdBlock> fwd;            // forward rows by pivot group
    std::vector<int> fsteps;              // 3 ints per step: L(row, j), L(row+1, j) (-1: one row), j
    std::vector<BwdBlock> bwd;            // backward rows by pivot group
    std::vector<int> bsteps;              // 3 ints per step: U(row, j), U(row+1, j) (-1: one row), j
};


struct Topology {
    // ---- polar topology
    int n = 0, m = 0, npv = 0, npq = 0, npvpq = 0, nnzY = 0, nnzJ = 0;
    std::vector<int> Yp, Yj, Ydiag;
    std::vector<double> G, B;             // Ybus = G + jB (rectangular)
    std::vector<int> pv, pq, pvpq;
    std::vector<int> row_p, row_q;        // bus -> Jacobian row of its P / Q equation, -1 = none
    // ---- Jacobian pattern, CSR by equation, columns ascending
    std::vector<int> Jp, Jj, Jk, Jb;      // entry -> (column, Ybus nonzero, block)
    // ---- the same pattern by column, in the order of p3s/cpp/nr_klu.cpp:230-285 (within a
    // column the P rows, then the Q rows, each in ascending bus order): Cp (m + 1), Ci the row and
    // Ce the CSR entry of every column slot (the pivoting factorization reads columns this way)
    std::vector<int> Cp, Ci, Ce;
    std::vector<int> posJ;                // (4 * nnzY) Jacobian entry of block b of nonzero k, -1 = none
    // ---- static-pivot LU plan
    std::vector<int> perm, pinv;          // perm[p] = Jacobian row/column eliminated p-th
    std::vector<int> parent, height, depth;
    std::vector<int> Lp, Li;              // S_p in pivot space, rows ascending
    std::vector<int> base;                // (m + 1)
    std::vector<int> apos;                // Jacobian entry -> factor position
    std::vector<int> xva, xvm;            // bus -> pivot index of its Va / Vm variable, -1 = none
    std::vector<int> pa, pb;              // operand pairs: L(a,t) * U(t,b)
    std::vector<int> prange;              // (nnz_lu + 1) pairs of factor position d: [prange[d], prange[d+1])
    std::vector<int> Fp, Fpos, Fcol;      // row i of L: entries [Fp[i], Fp[i+1]) = (position, column)
    int64_t nnz_lu = 0, n_pairs = 0;
    int64_t fac_rows = 0;                 // factor arena rows: nnz_lu entries, the zero row (nnz_lu), the
                                          // one row (nnz_lu + 1), the schedules' partial slots
    int64_t x_rows = 0;                   // solution-vector rows: m, then the slots of cut substitution rows
    Schedule sched;
    double rcond0 = 0.0;                  // min/max |pivot| of the intact grid at flat start
    bool lean_valid = true;               // false: the flat-start Jacobian has a zero or non-finite
                                          // static pivot; every solve then factors with pivoting
    int ordering = 0;
};

// Build everything. ordering: 0 = fill-reducing (minimum fill on the buses), 2 = natural.
// Throws std::runtime_error.
void build_topology(Topology& t, const int* Yp, const int* Yj, const double* Yx, int n, int nnzY, const int* pv,
                    int npv, const int* pq, int npq, int ordering, int nslots);

// Refresh the admittance values (same pattern) and the reference pivot ratio.
void update_values(Topology& t, const double* Yx);

// (Re)deal the level schedule to `nslots` task lists (t.sched), or build another one.
void build_schedule(Topology& t, int nslots);
// group > 0: deal runs of up to `group` entries of one pivot's U row / L column together.
Schedule make_schedule(const Topology& t, int nslots, int group);
// The 2x2-block schedule dealt to `nslots` task lists (reads position nnz_lu as zero).
BlockSchedule make_block_schedule(const Topology& t, int nslots);

// Host execution of the plan for ONE system (checks and rcond0): Jacobian values at
// (Vm, Va) with admittance (G, B) [nullptr = the topology's], pin mask [nullptr = none],
// residual R = Sspec - Scalc in Jacobian row order, J in entry order.
void host_eval(const Topology& t, const double* Vm, const double* Va, const double* G, const double* B,
               const double* Ssp /* complex (n) */, const unsigned char* pin, double* R, double* J);
// Factor J (entry order) into fac (nnz_lu) following the schedule; returns min|pivot|
// and max|pivot| (pmin > 0 and finite on success).
void host_factor(const Topology& t, const double* J, double* fac, double* pmin, double* pmax, double* psum);
// Solve with the factors: x (pivot space, m) from R (Jacobian row order, m).
void host_solve(const Topology& t, const double* fac, const double* R, double* x);

// Host emulation of the device LU arithmetic (sums from zero with fused multiply-adds in
// stream order, start value added at the end, pivots inverted, backward columns descending)
// on the one-entry schedule (blocks = false) or the 2x2-block schedule, dealt to nslots
// lists, in FP64 or FP32 (single). Within a sub-level the lists run last to first, so an
// entry read before the sub-level that writes it is caught. x in pivot space. checks:
// [0] violations (a factor position written twice or never, read before written; the zero
// row written; a substitution reading an unsolved value), [1] pivot pairs, [2] steps.
void host_lu_emulate(const Topology& t, bool blocks, int nslots, bool single, const double* J, const double* R,
                     double* x, double* pmin, double* pmax, int64_t* checks);

}  // namespace p3smi
