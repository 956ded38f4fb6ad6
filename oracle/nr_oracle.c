/*
 * nr_oracle.c -- CPU FP64 restatement of the reference's nr_klu polar Newton path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the correctness oracle and the host CPU
 * baseline. It is loaded by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, and by nothing else: the product (the HIP library) never links,
 * loads or calls it.
 *
 * What it restates (reference file:line, relative to the reference checkout):
 *   - Topology setup: |Y|, angle(Y), diagonal index, pv/pq position maps
 *       p3s/cpp/nr_klu.cpp:86-125, 210-216, 732-759
 *   - CSC Jacobian pattern and its (src_k, src_block) gather map
 *       p3s/cpp/nr_klu.cpp:230-285
 *   - fused mismatch + Jacobian, one sincos per Ybus nonzero, pin rows
 *       p3s/cpp/nr_klu.cpp:311-424
 *   - NaN-safe infinity norm and the mismatch-only trial norm
 *       p3s/cpp/nr_klu.cpp:434-488
 *   - Armijo damped Newton (beta 0.5, 10 trials, c 1e-4, best-trial commit)
 *       p3s/cpp/nr_klu.cpp:561-711
 *   - frozen-pivot ("lean") LU plan, left-looking refactor with pivot-ratio result,
 *     permuted triangular solves
 *       p3s/cpp/lean_lu.hpp:55-197, p3s/cpp/nr_klu.cpp:718-729
 *   - run_column retry on a pivoting factorization
 *       p3s/cpp/nr_klu.cpp:844-857
 *   - solve_batch / solve_batch_contingency validation, OpenMP loop, outputs
 *       p3s/cpp/nr_klu.cpp:776-793, 1049-1281
 *   - exponent-bit finiteness test
 *       p3s/cpp/finite.hpp:21-25
 *
 * The reference delegates ordering and pivoting to SuiteSparse KLU 7.14.1
 * (p3s/cpp/ci/build_klu.py:49-50), which is not available here. Its role is restated:
 *   - ordering: a fill-reducing order (KLU default: AMD): minimum fill, taken on the buses
 *     (the quotient graph of the Jacobian pattern), each bus's P equation before its Q
 *     equation -- the order the device plan factors in, so that the frozen plan's pivot
 *     ratios, and with them the reuse guard's decisions, are the same on both;
 *   - klu_factor: left-looking Gilbert-Peierls LU with threshold partial pivoting
 *     (tolerance 0.001, diagonal preferred, rows scaled by their max |a_ij|, as KLU's
 *     defaults do), KLU's pivot growth estimate rcond = min|u_kk| / max|u_kk|;
 *   - klu_refactor: the same numeric sweep on the pivot order of the last klu_factor.
 * Each batch column runs on a fresh solve state, so the KLU fallback's pivot order never
 * carries over between columns (the reference drops a carried order whenever its lean
 * plan is valid, p3s/cpp/nr_klu.cpp:616-619, 847-856).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ONORM_NONFINITE DBL_MAX
#define LS_BETA 0.5
#define LS_MAX_TRIALS 10
#define ARMIJO_C 1e-4
#define REUSE_RCOND_RATIO_DEFAULT 1e-3
/* pivot-ratio guard of the frozen plan (tests may change it, as the reference's tests
 * monkeypatch _RCOND_RATIO, tests/cuda/test_lean_lu.py:224-244) */
static double REUSE_RCOND_RATIO = REUSE_RCOND_RATIO_DEFAULT;
void oracle_set_lean_guard(double ratio) { REUSE_RCOND_RATIO = ratio; }
#define PIVOT_TOL 1e-3

static int o_isfinite(double x) {
    uint64_t b;
    memcpy(&b, &x, sizeof b);
    return (b & 0x7ff0000000000000ULL) != 0x7ff0000000000000ULL;
}

static void* xcalloc(size_t n, size_t sz) {
    void* p = calloc(n ? n : 1, sz ? sz : 1);
    if (!p) { fprintf(stderr, "nr_oracle: out of memory\n"); abort(); }
    return p;
}
#define NEWV(T, n) ((T*)xcalloc((size_t)(n), sizeof(T)))

/* ------------------------------------------------------------------ int vector */
typedef struct { int* a; int n, cap; } ivec;
static void iv_push(ivec* v, int x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 8;
        v->a = (int*)realloc(v->a, (size_t)v->cap * sizeof(int));
        if (!v->a) abort();
    }
    v->a[v->n++] = x;
}

/* ================================================================== ordering
 * Minimum degree on the symmetric pattern of A + A^T (diagonal ignored), on the explicit
 * elimination graph: eliminating v joins its remaining neighbours into a clique. A node's
 * degree is its number of live neighbours, or with weights w (may be NULL) the sum of
 * theirs. Ties go to the lower index, so the order is deterministic. perm[k] = column
 * eliminated k-th. */
typedef struct { int deg, node; } heapent;
static int he_less(heapent a, heapent b) { return a.deg < b.deg || (a.deg == b.deg && a.node < b.node); }
static void heap_push(heapent* h, int* hn, heapent e) {
    int i = (*hn)++;
    h[i] = e;
    while (i > 0) {
        int p = (i - 1) / 2;
        if (!he_less(h[i], h[p])) break;
        heapent t = h[i]; h[i] = h[p]; h[p] = t; i = p;
    }
}
static heapent heap_pop(heapent* h, int* hn) {
    heapent top = h[0];
    h[0] = h[--(*hn)];
    int i = 0;
    for (;;) {
        int l = 2 * i + 1, r = l + 1, s = i;
        if (l < *hn && he_less(h[l], h[s])) s = l;
        if (r < *hn && he_less(h[r], h[s])) s = r;
        if (s == i) break;
        heapent t = h[i]; h[i] = h[s]; h[s] = t; i = s;
    }
    return top;
}

static int o_degree(const ivec* adj, int v, const int* wt) {
    if (!wt) return adj[v].n;
    int d = 0;
    for (int q = 0; q < adj[v].n; ++q) d += wt[adj[v].a[q]];
    return d;
}

static void o_min_degree_w(int m, const int* Ap, const int* Ai, const int* wt, int* perm) {
    ivec* adj = NEWV(ivec, m);
    int* mark = NEWV(int, m);
    char* gone = NEWV(char, m);
    for (int i = 0; i < m; ++i) mark[i] = -1;
    for (int j = 0; j < m; ++j)
        for (int p = Ap[j]; p < Ap[j + 1]; ++p) {
            int i = Ai[p];
            if (i == j) continue;
            iv_push(&adj[i], j);
            iv_push(&adj[j], i);
        }
    for (int i = 0; i < m; ++i) { /* dedupe */
        int w = 0;
        for (int q = 0; q < adj[i].n; ++q) {
            int u = adj[i].a[q];
            if (mark[u] != i) { mark[u] = i; adj[i].a[w++] = u; }
        }
        adj[i].n = w;
    }
    for (int i = 0; i < m; ++i) mark[i] = -1;
    size_t hcap = (size_t)8 * m + 16;
    heapent* h = (heapent*)malloc(hcap * sizeof(heapent));
    int hn = 0;
    for (int i = 0; i < m; ++i) heap_push(h, &hn, (heapent){o_degree(adj, i, wt), i});
    int stamp = 0;
    for (int k = 0; k < m; ++k) {
        heapent e;
        for (;;) {
            e = heap_pop(h, &hn);
            if (!gone[e.node] && e.deg == o_degree(adj, e.node, wt)) break;
        }
        int v = e.node;
        perm[k] = v;
        gone[v] = 1;
        ivec nb = adj[v];
        for (int q = 0; q < nb.n; ++q) {
            int u = nb.a[q];
            ++stamp;
            int w = 0;
            for (int r = 0; r < adj[u].n; ++r) { /* drop v and eliminated nodes */
                int x = adj[u].a[r];
                if (gone[x]) continue;
                adj[u].a[w++] = x;
                mark[x] = stamp;
            }
            adj[u].n = w;
            mark[u] = stamp;
            for (int r = 0; r < nb.n; ++r) {
                int x = nb.a[r];
                if (mark[x] != stamp) { mark[x] = stamp; iv_push(&adj[u], x); }
            }
            if ((size_t)hn + 1 >= hcap) {
                hcap *= 2;
                h = (heapent*)realloc(h, hcap * sizeof(heapent));
            }
            heap_push(h, &hn, (heapent){o_degree(adj, u, wt), u});
        }
        /* rebuild the heap when stale entries dominate */
        if ((size_t)hn > (size_t)4 * (m - k) + 64) {
            int w = 0;
            for (int q = 0; q < hn; ++q)
                if (!gone[h[q].node] && h[q].deg == o_degree(adj, h[q].node, wt)) h[w++] = h[q];
            int hn2 = 0;
            for (int q = 0; q < w; ++q) heap_push(h, &hn2, h[q]);
            hn = hn2;
        }
    }
    for (int i = 0; i < m; ++i) free(adj[i].a);
    free(adj); free(mark); free(gone); free(h);
}

static void o_min_degree(int m, const int* Ap, const int* Ai, int* perm) { o_min_degree_w(m, Ap, Ai, NULL, perm); }

/* Minimum fill on a weighted symmetric pattern (diagonal ignored), on the explicit
 * elimination graph: eliminate the node whose clique adds the least weight (sum of wt_a wt_b
 * over its non-adjacent live neighbour pairs), ties by weighted degree, then lower index.
 * Keys of the clique and its neighbours are recomputed after each elimination (a binary heap
 * with versions; stale entries are skipped). perm[k] = node eliminated k-th. */
typedef struct { int64_t fill, deg; int node, ver; } fillent;
static int fe_less(fillent a, fillent b) {
    if (a.fill != b.fill) return a.fill < b.fill;
    if (a.deg != b.deg) return a.deg < b.deg;
    return a.node < b.node;
}
static void fheap_push(fillent** h, int* hn, int* cap, fillent e) {
    if (*hn >= *cap) { *cap *= 2; *h = (fillent*)realloc(*h, (size_t)*cap * sizeof(fillent)); if (!*h) abort(); }
    int i = (*hn)++;
    (*h)[i] = e;
    while (i > 0) {
        int p = (i - 1) / 2;
        if (!fe_less((*h)[i], (*h)[p])) break;
        fillent t = (*h)[i]; (*h)[i] = (*h)[p]; (*h)[p] = t; i = p;
    }
}
static fillent fheap_pop(fillent* h, int* hn) {
    fillent top = h[0];
    h[0] = h[--(*hn)];
    int i = 0;
    for (;;) {
        int l = 2 * i + 1, r = l + 1, s = i;
        if (l < *hn && fe_less(h[l], h[s])) s = l;
        if (r < *hn && fe_less(h[r], h[s])) s = r;
        if (s == i) break;
        fillent t = h[i]; h[i] = h[s]; h[s] = t; i = s;
    }
    return top;
}
static void o_min_fill_w(int m, const int* Ap, const int* Ai, const int* wt, int* perm) {
    ivec* adj = NEWV(ivec, m);
    int* mark = NEWV(int, m);
    int* stamp = NEWV(int, m);
    int* ver = NEWV(int, m);
    char* gone = NEWV(char, m);
    for (int i = 0; i < m; ++i) mark[i] = stamp[i] = -1;
    for (int j = 0; j < m; ++j)
        for (int p = Ap[j]; p < Ap[j + 1]; ++p) {
            int i = Ai[p];
            if (i == j) continue;
            iv_push(&adj[i], j);
            iv_push(&adj[j], i);
        }
    for (int i = 0; i < m; ++i) { /* dedupe */
        int w = 0;
        for (int q = 0; q < adj[i].n; ++q) {
            int u = adj[i].a[q];
            if (mark[u] != i) { mark[u] = i; adj[i].a[w++] = u; }
        }
        adj[i].n = w;
    }
    for (int i = 0; i < m; ++i) mark[i] = -1;
    int tick = 0, st = 0;
    #define O_FILL_KEY(v, F, D) do { \
        (F) = 0; (D) = 0; \
        for (int q_ = 0; q_ < adj[v].n; ++q_) (D) += wt[adj[v].a[q_]]; \
        for (int i_ = 0; i_ < adj[v].n; ++i_) { \
            int a_ = adj[v].a[i_]; ++tick; \
            for (int r_ = 0; r_ < adj[a_].n; ++r_) mark[adj[a_].a[r_]] = tick; \
            for (int j_ = i_ + 1; j_ < adj[v].n; ++j_) \
                if (mark[adj[v].a[j_]] != tick) (F) += (int64_t)wt[a_] * wt[adj[v].a[j_]]; \
        } } while (0)
    int cap = 2 * m + 16, hn = 0;
    fillent* h = (fillent*)malloc((size_t)cap * sizeof(fillent));
    for (int v = 0; v < m; ++v) {
        int64_t f, d;
        O_FILL_KEY(v, f, d);
        fheap_push(&h, &hn, &cap, (fillent){f, d, v, 0});
    }
    ivec touched = {0, 0, 0};
    for (int k = 0; k < m; ++k) {
        fillent e;
        for (;;) {
            e = fheap_pop(h, &hn);
            if (!gone[e.node] && e.ver == ver[e.node]) break;
        }
        int v = e.node;
        perm[k] = v;
        gone[v] = 1;
        ivec nb = adj[v];
        adj[v].a = NULL; adj[v].n = adj[v].cap = 0;
        for (int q = 0; q < nb.n; ++q) {
            int u = nb.a[q];
            if (gone[u]) continue;
            ++st;
            int w = 0;
            for (int r = 0; r < adj[u].n; ++r) { /* drop v and eliminated nodes */
                int x = adj[u].a[r];
                if (gone[x]) continue;
                adj[u].a[w++] = x;
                stamp[x] = st;
            }
            adj[u].n = w;
            stamp[u] = st;
            for (int r = 0; r < nb.n; ++r) {
                int x = nb.a[r];
                if (!gone[x] && stamp[x] != st) { stamp[x] = st; iv_push(&adj[u], x); }
            }
        }
        /* keys change for the clique and its neighbours */
        ++st;
        touched.n = 0;
        for (int q = 0; q < nb.n; ++q) {
            int u = nb.a[q];
            if (gone[u]) continue;
            if (stamp[u] != st) { stamp[u] = st; iv_push(&touched, u); }
            for (int r = 0; r < adj[u].n; ++r) {
                int x = adj[u].a[r];
                if (stamp[x] != st) { stamp[x] = st; iv_push(&touched, x); }
            }
        }
        for (int q = 0; q < touched.n; ++q) {
            int u = touched.a[q];
            int64_t f, d;
            O_FILL_KEY(u, f, d);
            fheap_push(&h, &hn, &cap, (fillent){f, d, u, ++ver[u]});
        }
        free(nb.a);
    }
    #undef O_FILL_KEY
    free(touched.a);
    for (int i = 0; i < m; ++i) free(adj[i].a);
    free(adj); free(mark); free(stamp); free(ver); free(gone); free(h);
}

/* Minimum fill on the buses: bus k = pvpq[k] (its P equation is row k), adjacent when any of
 * their equations couple in the Jacobian pattern, a bus weighing its equations (2 for a PQ
 * bus, 1 for a PV bus); then each bus's P equation before its Q equation (row npvpq + its
 * position in pq). Q[k] = column eliminated k-th. */
static void o_bus_min_degree(int m, int npvpq, const int* Jp, const int* Ji, const int* pvpq, const int* pq_pos,
                             const int* pvpq_pos, const int* pq, int* Q) {
    int* kb = NEWV(int, m); /* equation -> bus index */
    for (int v = 0; v < m; ++v) kb[v] = v < npvpq ? v : pvpq_pos[pq[v - npvpq]];
    ivec* col = NEWV(ivec, npvpq);
    for (int j = 0; j < m; ++j)
        for (int p = Jp[j]; p < Jp[j + 1]; ++p)
            if (kb[Ji[p]] != kb[j]) iv_push(&col[kb[j]], kb[Ji[p]]);
    int* Bp = NEWV(int, npvpq + 1);
    for (int b = 0; b < npvpq; ++b) Bp[b + 1] = Bp[b] + col[b].n;
    int* Bi = NEWV(int, Bp[npvpq] + 1);
    for (int b = 0; b < npvpq; ++b) {
        if (col[b].n) memcpy(Bi + Bp[b], col[b].a, (size_t)col[b].n * sizeof(int));
        free(col[b].a);
    }
    int* wt = NEWV(int, npvpq);
    for (int b = 0; b < npvpq; ++b) wt[b] = pq_pos[pvpq[b]] >= 0 ? 2 : 1;
    int* order = NEWV(int, npvpq);
    o_min_fill_w(npvpq, Bp, Bi, wt, order);
    int k = 0;
    for (int q = 0; q < npvpq; ++q) {
        const int b = order[q];
        Q[k++] = b;
        if (pq_pos[pvpq[b]] >= 0) Q[k++] = npvpq + pq_pos[pvpq[b]];
    }
    free(kb); free(col); free(Bp); free(Bi); free(wt); free(order);
}

/* ================================================================== pivoting LU
 * Left-looking Gilbert-Peierls LU of the row-scaled matrix R^-1 A Q with threshold
 * partial pivoting, standing in for klu_factor / klu_refactor / klu_solve. */
typedef struct {
    int m;
    int *Lp, *Li;      /* L (unit, diagonal not stored), CSC by step, original row indices */
    double* Lx;
    int *Up, *Ui;      /* U strictly upper part, CSC by step, step indices, topological order */
    double *Ux, *Ud;   /* Ud: pivots */
    int *P, *pinv;     /* P[k] = original row of pivot k */
    double* rs;        /* row scale factors (max |a_ij| of the row) */
    int lcap, ucap;
    double rcond;      /* min |u_kk| / max |u_kk| */
} gplu;

static void gp_free(gplu* f) {
    if (!f) return;
    free(f->Lp); free(f->Li); free(f->Lx); free(f->Up); free(f->Ui); free(f->Ux); free(f->Ud);
    free(f->P); free(f->pinv); free(f->rs);
    memset(f, 0, sizeof *f);
}

/* work buffers of one factorization sweep */
typedef struct { double* x; int *xi, *stack, *pstack, *flag, *list; } gpwork;
static void gpw_alloc(gpwork* w, int m) {
    w->x = NEWV(double, m); w->xi = NEWV(int, m); w->stack = NEWV(int, m);
    w->pstack = NEWV(int, m); w->flag = NEWV(int, m); w->list = NEWV(int, m);
    for (int i = 0; i < m; ++i) w->flag[i] = -1;
}
static void gpw_free(gpwork* w) { free(w->x); free(w->xi); free(w->stack); free(w->pstack); free(w->flag); free(w->list); }

/* Depth-first reach of the rows of A(:,col) through the columns of L (rows already
 * pivotal). Returns the count; xi[0..cnt) holds the reach in topological order. */
static int gp_reach(const gplu* f, const int* Ap, const int* Ai, int col, int k, gpwork* w) {
    int top = f->m;
    for (int p = Ap[col]; p < Ap[col + 1]; ++p) {
        int i0 = Ai[p];
        if (w->flag[i0] == k) continue;
        int head = 0;
        w->stack[0] = i0;
        while (head >= 0) {
            int i = w->stack[head];
            int j = f->pinv[i];
            if (w->flag[i] != k) {
                w->flag[i] = k;
                w->pstack[head] = (j < 0) ? 0 : f->Lp[j];
            }
            int done = 1;
            if (j >= 0) {
                int pend = f->Lp[j + 1];
                for (int q = w->pstack[head]; q < pend; ++q) {
                    int r = f->Li[q];
                    if (w->flag[r] == k) continue;
                    w->pstack[head] = q + 1;
                    w->stack[++head] = r;
                    done = 0;
                    break;
                }
            }
            if (done) { --head; w->xi[--top] = i; }
        }
    }
    int cnt = f->m - top;
    memmove(w->xi, w->xi + top, (size_t)cnt * sizeof(int));
    return cnt;
}

static void gp_grow(int** ia, double** xa, int* cap, int need) {
    if (need <= *cap) return;
    int nc = *cap ? *cap : 64;
    while (nc < need) nc *= 2;
    *ia = (int*)realloc(*ia, (size_t)nc * sizeof(int));
    *xa = (double*)realloc(*xa, (size_t)nc * sizeof(double));
    *cap = nc;
}

/* diagnostics: fresh pivoting factorizations, and the pivots they took off the diagonal */
static long long g_gp_factors = 0, g_gp_offdiag = 0;

/* klu_factor: returns 1 on success, 0 if structurally or numerically singular. */
static int gp_factor(gplu* f, int m, const int* Ap, const int* Ai, const double* Ax, const int* Q) {
    gp_free(f);
    f->m = m;
    f->Lp = NEWV(int, m + 1); f->Up = NEWV(int, m + 1);
    f->Ud = NEWV(double, m); f->P = NEWV(int, m); f->pinv = NEWV(int, m); f->rs = NEWV(double, m);
    for (int i = 0; i < m; ++i) f->pinv[i] = -1;
    for (int j = 0; j < m; ++j)
        for (int p = Ap[j]; p < Ap[j + 1]; ++p) {
            double a = fabs(Ax[p]);
            if (a > f->rs[Ai[p]]) f->rs[Ai[p]] = a;
        }
    for (int i = 0; i < m; ++i) if (!(f->rs[i] > 0.0) || !o_isfinite(f->rs[i])) f->rs[i] = 1.0;
    gpwork w;
    gpw_alloc(&w, m);
    int lnz = 0, unz = 0, ok = 1;
    double dmin = INFINITY, dmax = 0.0;
    for (int k = 0; k < m && ok; ++k) {
        int col = Q[k];
        int cnt = gp_reach(f, Ap, Ai, col, k, &w);
        for (int p = Ap[col]; p < Ap[col + 1]; ++p) w.x[Ai[p]] = Ax[p] / f->rs[Ai[p]];
        gp_grow(&f->Ui, &f->Ux, &f->ucap, unz + cnt);
        for (int q = 0; q < cnt; ++q) {
            int i = w.xi[q];
            int j = f->pinv[i];
            if (j < 0) continue;
            double xj = w.x[i];
            f->Ui[unz] = j; f->Ux[unz++] = xj;
            for (int s = f->Lp[j]; s < f->Lp[j + 1]; ++s) w.x[f->Li[s]] -= f->Lx[s] * xj;
        }
        f->Up[k + 1] = unz;
        /* pivot search over the non-pivotal rows of the reach */
        int piv = -1;
        double amax = -1.0;
        for (int q = 0; q < cnt; ++q) {
            int i = w.xi[q];
            if (f->pinv[i] >= 0) continue;
            double a = fabs(w.x[i]);
            if (a > amax) { amax = a; piv = i; }
        }
        if (f->pinv[col] < 0 && w.flag[col] == k && fabs(w.x[col]) >= PIVOT_TOL * amax) piv = col;
        if (piv < 0 || !(amax > 0.0) || !o_isfinite(amax)) { ok = 0; break; }
        double d = w.x[piv];
        f->Ud[k] = d;
        f->P[k] = piv;
        f->pinv[piv] = k;
        double ad = fabs(d);
        if (ad < dmin) dmin = ad;
        if (ad > dmax) dmax = ad;
        gp_grow(&f->Li, &f->Lx, &f->lcap, lnz + cnt);
        for (int q = 0; q < cnt; ++q) {
            int i = w.xi[q];
            if (f->pinv[i] >= 0) { w.x[i] = 0.0; continue; }
            f->Li[lnz] = i; f->Lx[lnz++] = w.x[i] / d;
            w.x[i] = 0.0;
        }
        w.x[piv] = 0.0;
        f->Lp[k + 1] = lnz;
    }
    gpw_free(&w);
    {
        long long off = 0;
        for (int k = 0; k < m && ok; ++k) off += f->P[k] != Q[k];
#pragma omp atomic
        g_gp_factors++;
#pragma omp atomic
        g_gp_offdiag += off;
    }
    if (!ok) { gp_free(f); return 0; }
    f->rcond = dmin / dmax;
    return 1;
}

/* klu_refactor: numeric sweep on the existing pivot order and pattern. 0 = singular. */
static int gp_refactor(gplu* f, const int* Ap, const int* Ai, const double* Ax, const int* Q, double* x) {
    int m = f->m;
    double dmin = INFINITY, dmax = 0.0;
    for (int k = 0; k < m; ++k) {
        int col = Q[k];
        for (int p = Ap[col]; p < Ap[col + 1]; ++p) x[Ai[p]] += Ax[p] / f->rs[Ai[p]];
        for (int q = f->Up[k]; q < f->Up[k + 1]; ++q) {
            int j = f->Ui[q];
            int i = f->P[j];
            double xj = x[i];
            f->Ux[q] = xj;
            x[i] = 0.0;
            for (int s = f->Lp[j]; s < f->Lp[j + 1]; ++s) x[f->Li[s]] -= f->Lx[s] * xj;
        }
        double d = x[f->P[k]];
        x[f->P[k]] = 0.0;
        f->Ud[k] = d;
        double ad = fabs(d);
        if (!(ad > 0.0) || !o_isfinite(ad)) {
            memset(x, 0, (size_t)m * sizeof(double));
            return 0;
        }
        if (ad < dmin) dmin = ad;
        if (ad > dmax) dmax = ad;
        for (int s = f->Lp[k]; s < f->Lp[k + 1]; ++s) { f->Lx[s] = x[f->Li[s]] / d; x[f->Li[s]] = 0.0; }
    }
    f->rcond = dmin / dmax;
    return 1;
}

/* klu_solve: b overwritten by A^-1 b. y is scratch (m). */
static void gp_solve(const gplu* f, const int* Q, double* b, double* y) {
    int m = f->m;
    for (int k = 0; k < m; ++k) { int i = f->P[k]; y[k] = b[i] / f->rs[i]; }
    for (int k = 0; k < m; ++k) {
        double yk = y[k];
        if (yk == 0.0) continue;
        for (int s = f->Lp[k]; s < f->Lp[k + 1]; ++s) y[f->pinv[f->Li[s]]] -= f->Lx[s] * yk;
    }
    for (int k = m - 1; k >= 0; --k) {
        double zk = y[k] / f->Ud[k];
        y[k] = zk;
        if (zk == 0.0) continue;
        for (int q = f->Up[k]; q < f->Up[k + 1]; ++q) y[f->Ui[q]] -= f->Ux[q] * zk;
    }
    for (int k = 0; k < m; ++k) b[Q[k]] = y[k];
}

/* ================================================================== topology */
typedef struct {
    int valid;
    int n;
    int *P, *Q;
    int *cp, *dpos, *row;
    int *aK, *aSrc, *aRow;
    int *dK, *dU, *dLb, *dLe;
    int nnz;
    double rcond0;
} leanplan;

typedef struct {
    int n, m, npvpq, npq, nnzY, nnzJ;
    int *Yp, *Yj, *Ydiag;
    double *Ym, *Ya;
    int *pvpq, *pq, *pvpq_pos, *pq_pos;
    int *Jp, *Ji, *src_k;
    signed char* src_block;
    int* Q;       /* fill-reducing column order */
    leanplan lean;
} otopo;

typedef struct {
    double *Jx, *Vm, *Va, *Vm_ls, *Va_ls, *Pspec, *Qspec, *F, *rhs;
    const double *Ym, *Ya;           /* per-case Ybus override, else the topology's */
    const unsigned char* pin;        /* per-case pin mask, or NULL */
    double *LU, *lu_x, *lu_y;
    double *buf, *Pcalc, *Qcalc;
    double *gx, *gy;                 /* pivoting-LU scratch */
    gplu fac;
    int factored, use_lean;
} ostate;

static void lean_free(leanplan* p) {
    free(p->P); free(p->Q); free(p->cp); free(p->dpos); free(p->row);
    free(p->aK); free(p->aSrc); free(p->aRow); free(p->dK); free(p->dU); free(p->dLb); free(p->dLe);
    memset(p, 0, sizeof *p);
}

/* CSC Jacobian pattern: p3s/cpp/nr_klu.cpp:230-285. Column c < npvpq is dVa at pvpq[c],
 * else dVm at pq[c - npvpq]; within a column the P rows (pvpq buses) come first in Ybus
 * column order, then the Q rows (pq buses). */
static void build_J_pattern(otopo* t) {
    int n = t->n, npvpq = t->npvpq, npq = t->npq;
    t->m = npvpq + npq;
    /* transpose of the Ybus pattern: for bus column j, its (row, k) entries by row */
    int* cnt = NEWV(int, n + 1);
    for (int k = 0; k < t->nnzY; ++k) cnt[t->Yj[k] + 1]++;
    for (int j = 0; j < n; ++j) cnt[j + 1] += cnt[j];
    int* crow = NEWV(int, t->nnzY);
    int* ck = NEWV(int, t->nnzY);
    int* fill = NEWV(int, n);
    for (int r = 0; r < n; ++r)
        for (int k = t->Yp[r]; k < t->Yp[r + 1]; ++k) {
            int j = t->Yj[k];
            int q = cnt[j] + fill[j]++;
            crow[q] = r; ck[q] = k;
        }
    int cap = 4 * t->nnzY + 4;
    t->Jp = NEWV(int, t->m + 1);
    t->Ji = NEWV(int, cap);
    t->src_k = NEWV(int, cap);
    t->src_block = NEWV(signed char, cap);
    int nz = 0, col = 0;
    for (int pass = 0; pass < 2; ++pass) {
        int ncols = pass == 0 ? npvpq : npq;
        for (int c = 0; c < ncols; ++c) {
            int bus = pass == 0 ? t->pvpq[c] : t->pq[c];
            for (int q = cnt[bus]; q < cnt[bus + 1]; ++q) {
                int rp = t->pvpq_pos[crow[q]];
                if (rp != -1) { t->Ji[nz] = rp; t->src_k[nz] = ck[q]; t->src_block[nz++] = pass ? 1 : 0; }
            }
            for (int q = cnt[bus]; q < cnt[bus + 1]; ++q) {
                int rq = t->pq_pos[crow[q]];
                if (rq != -1) { t->Ji[nz] = npvpq + rq; t->src_k[nz] = ck[q]; t->src_block[nz++] = pass ? 3 : 2; }
            }
            t->Jp[++col] = nz;
        }
    }
    t->nnzJ = nz;
    free(cnt); free(crow); free(ck); free(fill);
}

static void state_init(ostate* s, const otopo* t) {
    memset(s, 0, sizeof *s);
    s->Jx = NEWV(double, t->nnzJ);
    s->Vm = NEWV(double, t->n); s->Va = NEWV(double, t->n);
    s->Vm_ls = NEWV(double, t->n); s->Va_ls = NEWV(double, t->n);
    s->Pspec = NEWV(double, t->n); s->Qspec = NEWV(double, t->n);
    s->F = NEWV(double, t->m); s->rhs = NEWV(double, t->m);
    s->buf = NEWV(double, 4 * (size_t)t->nnzY);
    s->Pcalc = NEWV(double, t->n); s->Qcalc = NEWV(double, t->n);
    s->gx = NEWV(double, t->m); s->gy = NEWV(double, t->m);
    s->Ym = t->Ym; s->Ya = t->Ya;
    if (t->lean.valid) {
        s->LU = NEWV(double, t->lean.nnz);
        s->lu_x = NEWV(double, t->m);
        s->lu_y = NEWV(double, t->m);
    }
    s->use_lean = 1;
}

static void state_free(ostate* s) {
    free(s->Jx); free(s->Vm); free(s->Va); free(s->Vm_ls); free(s->Va_ls); free(s->Pspec); free(s->Qspec);
    free(s->F); free(s->rhs); free(s->buf); free(s->Pcalc); free(s->Qcalc); free(s->gx); free(s->gy);
    free(s->LU); free(s->lu_x); free(s->lu_y);
    gp_free(&s->fac);
}

/* ================================================================== mismatch + J */
/* p3s/cpp/nr_klu.cpp:311-424 */
static void eval_F_and_J(const otopo* w, ostate* st, const double* Vm, const double* Va, double* F) {
    const int nnzY = w->nnzY;
    const double* Ym = st->Ym;
    const double* Ya = st->Ya;
    double* dP_dVa = st->buf;
    double* dQ_dVa = st->buf + nnzY;
    double* dP_dVm = st->buf + 2 * (size_t)nnzY;
    double* dQ_dVm = st->buf + 3 * (size_t)nnzY;
    for (int row = 0; row < w->n; ++row) {
        int dix = w->Ydiag[row];
        if (dix < 0) continue;
        dP_dVa[dix] = 0.0; dQ_dVa[dix] = 0.0; dP_dVm[dix] = 0.0; dQ_dVm[dix] = 0.0;
    }
    for (int row = 0; row < w->n; ++row) {
        const int dix = w->Ydiag[row];
        double prow = 0.0, qrow = 0.0;
        for (int k = w->Yp[row]; k < w->Yp[row + 1]; ++k) {
            const int col = w->Yj[k];
            const double delta = Va[row] - Ya[k] - Va[col];
            const double s = sin(delta), c = cos(delta);
            const double VmYm = Vm[row] * Ym[k];
            const double vij = VmYm * Vm[col];
            prow += vij * c;
            qrow += vij * s;
            if (row != col) {
                const double dpva = vij * s;
                const double dqva = -vij * c;
                dP_dVa[k] = dpva;
                dQ_dVa[k] = dqva;
                dP_dVm[k] = VmYm * c;
                dQ_dVm[k] = VmYm * s;
                if (dix >= 0) {
                    dP_dVa[dix] -= dpva;
                    dQ_dVa[dix] -= dqva;
                    dP_dVm[dix] += Ym[k] * Vm[col] * c;
                    dQ_dVm[dix] += Ym[k] * Vm[col] * s;
                }
            } else {
                dP_dVm[dix] += 2.0 * Vm[row] * Ym[k] * c;
                dQ_dVm[dix] += 2.0 * Vm[row] * Ym[k] * s;
            }
        }
        st->Pcalc[row] = prow;
        st->Qcalc[row] = qrow;
    }
    const int npvpq = w->npvpq;
    for (int ri = 0; ri < npvpq; ++ri) { int row = w->pvpq[ri]; F[ri] = st->Pcalc[row] - st->Pspec[row]; }
    for (int ri = 0; ri < w->npq; ++ri) { int row = w->pq[ri]; F[npvpq + ri] = st->Qcalc[row] - st->Qspec[row]; }
    const double* blocks[4] = {dP_dVa, dP_dVm, dQ_dVa, dQ_dVm};
    for (int p = 0; p < w->nnzJ; ++p) st->Jx[p] = blocks[(int)w->src_block[p]][w->src_k[p]];
    if (st->pin) {
        const unsigned char* pin = st->pin;
        for (int col = 0; col < w->m; ++col)
            for (int p = w->Jp[col]; p < w->Jp[col + 1]; ++p) {
                int r = w->Ji[p];
                int rbus = r < npvpq ? w->pvpq[r] : w->pq[r - npvpq];
                if (pin[rbus]) st->Jx[p] = (r == col) ? 1.0 : 0.0;
            }
        for (int r = 0; r < npvpq; ++r) if (pin[w->pvpq[r]]) F[r] = 0.0;
        for (int r = 0; r < w->npq; ++r) if (pin[w->pq[r]]) F[npvpq + r] = 0.0;
    }
}

/* p3s/cpp/nr_klu.cpp:434-444 */
static double mismatch_norm(const double* F, int m) {
    double nrm = 0.0, sum = 0.0;
    for (int i = 0; i < m; ++i) {
        double a = fabs(F[i]);
        sum += a;
        if (a > nrm) nrm = a;
    }
    return o_isfinite(sum) ? nrm : ONORM_NONFINITE;
}

/* p3s/cpp/nr_klu.cpp:456-488 */
static double eval_F_norm(const otopo* w, const ostate* st, const double* Vm, const double* Va) {
    const double* Ym = st->Ym;
    const double* Ya = st->Ya;
    const unsigned char* pin = st->pin;
    double nrm = 0.0, sum = 0.0;
    for (int ri = 0; ri < w->npvpq; ++ri) {
        int row = w->pvpq[ri];
        if (pin && pin[row]) continue;
        double p = 0.0;
        for (int k = w->Yp[row]; k < w->Yp[row + 1]; ++k) {
            double delta = Va[row] - Ya[k] - Va[w->Yj[k]];
            p += Vm[row] * Ym[k] * Vm[w->Yj[k]] * cos(delta);
        }
        double a = fabs(p - st->Pspec[row]);
        sum += a;
        if (a > nrm) nrm = a;
    }
    for (int ri = 0; ri < w->npq; ++ri) {
        int row = w->pq[ri];
        if (pin && pin[row]) continue;
        double q = 0.0;
        for (int k = w->Yp[row]; k < w->Yp[row + 1]; ++k) {
            double delta = Va[row] - Ya[k] - Va[w->Yj[k]];
            q += Vm[row] * Ym[k] * Vm[w->Yj[k]] * sin(delta);
        }
        double a = fabs(q - st->Qspec[row]);
        sum += a;
        if (a > nrm) nrm = a;
    }
    return o_isfinite(sum) ? nrm : ONORM_NONFINITE;
}

/* ================================================================== lean plan
 * p3s/cpp/lean_lu.hpp:55-140: freeze the pivot order and L/U pattern of one pivoting
 * factorization. Per column k: U rows ascending, the pivot at dpos[k], then the L rows. */
static int lean_build(leanplan* pl, int m, const int* Ap, const int* Ai, const gplu* f, const int* Q) {
    lean_free(pl);
    pl->n = m;
    pl->P = NEWV(int, m); pl->Q = NEWV(int, m);
    memcpy(pl->P, f->P, (size_t)m * sizeof(int));
    memcpy(pl->Q, Q, (size_t)m * sizeof(int));
    double pmin = INFINITY, pmax = 0.0;
    for (int k = 0; k < m; ++k) {
        double a = fabs(f->rs[f->P[k]] * f->Ud[k]);
        if (a < pmin) pmin = a;
        if (a > pmax) pmax = a;
    }
    if (!o_isfinite(pmax) || !o_isfinite(pmin) || !(pmax > 0.0) || !(pmin > 0.0)) return 0;
    pl->rcond0 = pmin / pmax;
    pl->cp = NEWV(int, m + 1);
    pl->dpos = NEWV(int, m);
    for (int k = 0; k < m; ++k) {
        int nu = f->Up[k + 1] - f->Up[k], nl = f->Lp[k + 1] - f->Lp[k];
        pl->cp[k + 1] = pl->cp[k] + nu + 1 + nl;
    }
    pl->nnz = pl->cp[m];
    pl->row = NEWV(int, pl->nnz);
    int* Lbeg = NEWV(int, m);
    int* Lend = NEWV(int, m);
    int n_udep = 0;
    for (int k = 0; k < m; ++k) {
        int q = pl->cp[k];
        int q0 = q;
        for (int p = f->Up[k]; p < f->Up[k + 1]; ++p) pl->row[q++] = f->Ui[p];
        /* insertion sort: U rows ascending (the elimination order) */
        for (int a = q0 + 1; a < q; ++a) {
            int v = pl->row[a], b = a - 1;
            while (b >= q0 && pl->row[b] > v) { pl->row[b + 1] = pl->row[b]; --b; }
            pl->row[b + 1] = v;
        }
        n_udep += q - q0;
        pl->dpos[k] = q;
        pl->row[q++] = k;
        Lbeg[k] = q;
        for (int p = f->Lp[k]; p < f->Lp[k + 1]; ++p) pl->row[q++] = f->pinv[f->Li[p]];
        Lend[k] = q;
    }
    char* inpat = NEWV(char, m);
    pl->aK = NEWV(int, m + 1);
    pl->aSrc = NEWV(int, Ap[m] + 1);
    pl->aRow = NEWV(int, Ap[m] + 1);
    pl->dK = NEWV(int, m + 1);
    pl->dU = NEWV(int, n_udep + 1);
    pl->dLb = NEWV(int, n_udep + 1);
    pl->dLe = NEWV(int, n_udep + 1);
    int na = 0, nd = 0, ok = 1;
    for (int k = 0; k < m && ok; ++k) {
        for (int q = pl->cp[k]; q < pl->cp[k + 1]; ++q) inpat[pl->row[q]] = 1;
        int ac = Q[k];
        for (int p = Ap[ac]; p < Ap[ac + 1]; ++p) {
            int r = f->pinv[Ai[p]];
            if (!inpat[r]) { ok = 0; break; }
            pl->aSrc[na] = p; pl->aRow[na++] = r;
        }
        pl->aK[k + 1] = na;
        for (int q = pl->cp[k]; q < pl->dpos[k]; ++q) {
            int j = pl->row[q];
            pl->dU[nd] = q; pl->dLb[nd] = Lbeg[j]; pl->dLe[nd++] = Lend[j];
        }
        pl->dK[k + 1] = nd;
        for (int q = pl->cp[k]; q < pl->cp[k + 1]; ++q) inpat[pl->row[q]] = 0;
    }
    free(Lbeg); free(Lend); free(inpat);
    pl->valid = ok;
    return ok;
}

/* p3s/cpp/lean_lu.hpp:146-177 */
static double lean_refactor(const leanplan* pl, const double* Ax, double* LU, double* x) {
    const int* row = pl->row;
    double pmin = INFINITY, pmax = 0.0, asum = 0.0;
    int finite = 1;
    for (int k = 0; k < pl->n; ++k) {
        for (int e = pl->aK[k]; e < pl->aK[k + 1]; ++e) x[pl->aRow[e]] += Ax[pl->aSrc[e]];
        for (int d = pl->dK[k]; d < pl->dK[k + 1]; ++d) {
            double u = x[row[pl->dU[d]]];
            if (u == 0.0) continue;
            for (int s = pl->dLb[d]; s < pl->dLe[d]; ++s) x[row[s]] -= u * LU[s];
        }
        int dp = pl->dpos[k];
        for (int q = pl->cp[k]; q < dp; ++q) { int r = row[q]; LU[q] = x[r]; x[r] = 0.0; }
        double piv = x[k];
        x[k] = 0.0;
        LU[dp] = piv;
        double a = fabs(piv);
        if (!(a > 0.0)) finite = 0;
        asum += a;
        if (a < pmin) pmin = a;
        if (a > pmax) pmax = a;
        double inv = 1.0 / piv;
        for (int q = dp + 1; q < pl->cp[k + 1]; ++q) { int r = row[q]; LU[q] = x[r] * inv; x[r] = 0.0; }
    }
    if (!finite || !o_isfinite(asum) || pmax == 0.0) return 0.0;
    return pmin / pmax;
}

/* p3s/cpp/lean_lu.hpp:181-197 */
static void lean_solve(const leanplan* pl, const double* LU, double* b, double* y) {
    int n = pl->n;
    const int* row = pl->row;
    for (int k = 0; k < n; ++k) y[k] = b[pl->P[k]];
    for (int k = 0; k < n; ++k) {
        double yk = y[k];
        if (yk == 0.0) continue;
        for (int q = pl->dpos[k] + 1; q < pl->cp[k + 1]; ++q) y[row[q]] -= LU[q] * yk;
    }
    for (int k = n - 1; k >= 0; --k) {
        double zk = y[k] / LU[pl->dpos[k]];
        y[k] = zk;
        if (zk == 0.0) continue;
        for (int q = pl->cp[k]; q < pl->dpos[k]; ++q) y[row[q]] -= LU[q] * zk;
    }
    for (int k = 0; k < n; ++k) b[pl->Q[k]] = y[k];
}

/* ================================================================== Newton */
typedef struct { int iterations, converged; } nres;

/* diagnostics: how often the frozen-pivot plan was rejected / a column was re-run */
static long long g_lean_rejections = 0, g_column_reruns = 0, g_backtracks = 0;
long long oracle_lean_rejections(void) { return g_lean_rejections; }
long long oracle_pivoting_factors(void) { return g_gp_factors; }
long long oracle_offdiagonal_pivots(void) { return g_gp_offdiag; }
long long oracle_column_reruns(void) { return g_column_reruns; }
long long oracle_backtracks(void) { return g_backtracks; }
void oracle_reset_counters(void) { g_lean_rejections = 0; g_column_reruns = 0; g_backtracks = 0; g_gp_factors = 0; g_gp_offdiag = 0; }

/* p3s/cpp/nr_klu.cpp:569-711 */
static nres run_newton(const otopo* topo, ostate* st, int max_iter, double tol, int line_search) {
    const int npvpq = topo->npvpq, npq = topo->npq, m = topo->m;
    double *Vm = st->Vm, *Va = st->Va, *F = st->F, *rhs = st->rhs, *Vm_ls = st->Vm_ls, *Va_ls = st->Va_ls;
    eval_F_and_J(topo, st, Vm, Va, F);
    double nrm = mismatch_norm(F, m);
    int it = 0, converged = 0;
    while (it < max_iter) {
        if (nrm < tol) { converged = 1; break; }
        if (nrm >= ONORM_NONFINITE) break;
        ++it;
        int lean_done = 0;
        if (st->use_lean && topo->lean.valid) {
            double rc = lean_refactor(&topo->lean, st->Jx, st->LU, st->lu_x);
            if (rc >= REUSE_RCOND_RATIO * topo->lean.rcond0) {
                for (int i = 0; i < m; ++i) rhs[i] = -F[i];
                lean_solve(&topo->lean, st->LU, rhs, st->lu_y);
                lean_done = 1;
            } else {
#pragma omp atomic
                g_lean_rejections++;
                st->use_lean = 0;
                gp_free(&st->fac);
                st->factored = 0;
            }
        }
        if (!lean_done) {
            int need_factor = !st->factored;
            if (!need_factor) {
                if (!gp_refactor(&st->fac, topo->Jp, topo->Ji, st->Jx, topo->Q, st->gx)) {
                    gp_free(&st->fac);
                    need_factor = 1;
                }
            }
            if (need_factor) {
                st->factored = gp_factor(&st->fac, m, topo->Jp, topo->Ji, st->Jx, topo->Q);
                if (!st->factored) return (nres){it, 0};
            }
            for (int i = 0; i < m; ++i) rhs[i] = -F[i];
            gp_solve(&st->fac, topo->Q, rhs, st->gy);
        }
        /* full step scored with eval_F_and_J itself, then Armijo backtracking */
#define FORM_TRIAL(a)                                                                 \
    do {                                                                              \
        memcpy(Vm_ls, Vm, (size_t)topo->n * sizeof(double));                         \
        memcpy(Va_ls, Va, (size_t)topo->n * sizeof(double));                         \
        for (int i_ = 0; i_ < npvpq; ++i_) Va_ls[topo->pvpq[i_]] += (a) * rhs[i_];   \
        for (int i_ = 0; i_ < npq; ++i_) Vm_ls[topo->pq[i_]] += (a) * rhs[npvpq + i_]; \
    } while (0)
#define COMMIT_TRIAL()                                                                \
    do {                                                                              \
        for (int i_ = 0; i_ < npvpq; ++i_) Va[topo->pvpq[i_]] = Va_ls[topo->pvpq[i_]]; \
        for (int i_ = 0; i_ < npq; ++i_) Vm[topo->pq[i_]] = Vm_ls[topo->pq[i_]];    \
    } while (0)
        FORM_TRIAL(1.0);
        eval_F_and_J(topo, st, Vm_ls, Va_ls, F);
        const double nrm_full = mismatch_norm(F, m);
        if (!line_search || nrm_full <= (1.0 - ARMIJO_C) * nrm) {
            COMMIT_TRIAL();
            nrm = nrm_full;
            continue;
        }
#pragma omp atomic
        g_backtracks++;
        double alpha = 1.0, best_alpha = 1.0, best_nrm = nrm_full;
        for (int trial = 1; trial <= LS_MAX_TRIALS; ++trial) {
            alpha *= LS_BETA;
            FORM_TRIAL(alpha);
            double nrm_new = eval_F_norm(topo, st, Vm_ls, Va_ls);
            if (nrm_new < best_nrm) { best_nrm = nrm_new; best_alpha = alpha; }
            if (nrm_new <= (1.0 - ARMIJO_C * alpha) * nrm) break;
            if (trial == LS_MAX_TRIALS) {
                if (best_alpha != alpha) FORM_TRIAL(best_alpha);
                break;
            }
        }
        COMMIT_TRIAL();
        eval_F_and_J(topo, st, Vm, Va, F);
        nrm = mismatch_norm(F, m);
#undef FORM_TRIAL
#undef COMMIT_TRIAL
    }
    return (nres){it, converged};
}

/* p3s/cpp/nr_klu.cpp:844-857, with a fresh state per column */
typedef int (*init_fn)(ostate* s, const otopo* t, const void* ctx, int col, char* err, int errlen);

static nres run_column(const otopo* topo, ostate* s, init_fn init, const void* ctx, int col, int max_iter,
                       double tol, int line_search, char* err, int errlen, int* failed) {
    const int reused = topo->lean.valid;
    *failed = init(s, topo, ctx, col, err, errlen);
    if (*failed) return (nres){0, 0};
    s->use_lean = 1;
    s->factored = 0;
    nres r = run_newton(topo, s, max_iter, tol, line_search);
    if (r.converged || !reused) return r;
#pragma omp atomic
    g_column_reruns++;
    gp_free(&s->fac);
    s->factored = 0;
    s->use_lean = 0;
    *failed = init(s, topo, ctx, col, err, errlen);
    if (*failed) return (nres){0, 0};
    return run_newton(topo, s, max_iter, tol, line_search);
}

/* ================================================================== API */
static void set_err(char* err, int errlen, const char* msg) {
    if (err && errlen > 0) { strncpy(err, msg, (size_t)errlen - 1); err[errlen - 1] = 0; }
}

static void topo_free_fields(otopo* t) {
    free(t->Yp); free(t->Yj); free(t->Ydiag); free(t->Ym); free(t->Ya);
    free(t->pvpq); free(t->pq); free(t->pvpq_pos); free(t->pq_pos);
    free(t->Jp); free(t->Ji); free(t->src_k); free(t->src_block); free(t->Q);
    lean_free(&t->lean);
}

/* p3s/cpp/nr_klu.cpp:718-729: factor J at |V| = 1, Va = 0 and freeze its pivot order. */
static void build_lean_plan(otopo* t) {
    ostate s;
    memset(&t->lean, 0, sizeof t->lean); /* no lean buffers in the probe state */
    state_init(&s, t);
    for (int i = 0; i < t->n; ++i) s.Vm[i] = 1.0;
    eval_F_and_J(t, &s, s.Vm, s.Va, s.F);
    gplu f;
    memset(&f, 0, sizeof f);
    if (gp_factor(&f, t->m, t->Jp, t->Ji, s.Jx, t->Q)) lean_build(&t->lean, t->m, t->Jp, t->Ji, &f, t->Q);
    gp_free(&f);
    state_free(&s);
}

/* Yx given as interleaved (re, im) pairs. ordering: 0 = minimum degree on the buses (P
 * before Q of each bus), 2 = natural. */
void* oracle_topo_create(const int* Yp, const int* Yj, const double* Yx, int n, int nnzY, const int* pv, int npv,
                         const int* pq, int npq, int ordering, char* err, int errlen) {
    if (n <= 0 || nnzY <= 0 || Yp[n] != nnzY) { set_err(err, errlen, "invalid Ybus CSR"); return NULL; }
    otopo* t = NEWV(otopo, 1);
    t->n = n; t->nnzY = nnzY; t->npvpq = npv + npq; t->npq = npq;
    t->Yp = NEWV(int, n + 1); memcpy(t->Yp, Yp, (size_t)(n + 1) * sizeof(int));
    t->Yj = NEWV(int, nnzY); memcpy(t->Yj, Yj, (size_t)nnzY * sizeof(int));
    t->Ym = NEWV(double, nnzY); t->Ya = NEWV(double, nnzY);
    for (int k = 0; k < nnzY; ++k) { t->Ym[k] = hypot(Yx[2 * k], Yx[2 * k + 1]); t->Ya[k] = atan2(Yx[2 * k + 1], Yx[2 * k]); }
    t->Ydiag = NEWV(int, n);
    for (int r = 0; r < n; ++r) {
        t->Ydiag[r] = -1;
        for (int k = Yp[r]; k < Yp[r + 1]; ++k) if (Yj[k] == r) { t->Ydiag[r] = k; break; }
    }
    t->pvpq = NEWV(int, t->npvpq);
    t->pq = NEWV(int, npq);
    for (int i = 0; i < npv; ++i) t->pvpq[i] = pv[i];
    for (int i = 0; i < npq; ++i) { t->pvpq[npv + i] = pq[i]; t->pq[i] = pq[i]; }
    t->pvpq_pos = NEWV(int, n); t->pq_pos = NEWV(int, n);
    for (int i = 0; i < n; ++i) { t->pvpq_pos[i] = -1; t->pq_pos[i] = -1; }
    for (int i = 0; i < t->npvpq; ++i) t->pvpq_pos[t->pvpq[i]] = i;
    for (int i = 0; i < npq; ++i) t->pq_pos[t->pq[i]] = i;
    build_J_pattern(t);
    t->Q = NEWV(int, t->m);
    if (ordering == 2) for (int k = 0; k < t->m; ++k) t->Q[k] = k;
    else o_bus_min_degree(t->m, t->npvpq, t->Jp, t->Ji, t->pvpq, t->pq_pos, t->pvpq_pos, t->pq, t->Q);
    build_lean_plan(t);
    return t;
}

void oracle_topo_free(void* h) {
    if (!h) return;
    topo_free_fields((otopo*)h);
    free(h);
}

int oracle_lean_valid(void* h) { return ((otopo*)h)->lean.valid; }
int oracle_dims(void* h, int* out) {
    otopo* t = (otopo*)h;
    out[0] = t->n; out[1] = t->m; out[2] = t->nnzY; out[3] = t->nnzJ; out[4] = t->lean.valid ? t->lean.nnz : 0;
    return 0;
}
double oracle_rcond0(void* h) { return ((otopo*)h)->lean.rcond0; }

int oracle_update_Y(void* h, const double* Yx, int nnz, char* err, int errlen) {
    otopo* t = (otopo*)h;
    if (nnz != t->nnzY) { set_err(err, errlen, "update_Y: nnz mismatch (topology changed; rebuild Solver)"); return 1; }
    for (int k = 0; k < nnz; ++k) { t->Ym[k] = hypot(Yx[2 * k], Yx[2 * k + 1]); t->Ya[k] = atan2(Yx[2 * k + 1], Yx[2 * k]); }
    return 0;
}

/* ---------------------------------------------------------------- time-series batch */
typedef struct {
    const double *S, *V0;
    int T, v0_per_col;
} ts_ctx;

static int ts_init(ostate* s, const otopo* t, const void* vctx, int col, char* err, int errlen) {
    const ts_ctx* c = (const ts_ctx*)vctx;
    for (int i = 0; i < t->n; ++i) {
        size_t si = ((size_t)i * c->T + col) * 2;
        size_t vi = (c->v0_per_col ? ((size_t)i * c->T + col) : (size_t)i) * 2;
        double vr = c->V0[vi], vim = c->V0[vi + 1];
        if (!o_isfinite(vr) || !o_isfinite(vim)) {
            char msg[160];
            snprintf(msg, sizeof msg, "invalid initial voltage at bus %d, step %d", i, col);
            set_err(err, errlen, msg);
            return 1;
        }
        s->Pspec[i] = c->S[si]; s->Qspec[i] = c->S[si + 1];
        s->Vm[i] = hypot(vr, vim); s->Va[i] = atan2(vim, vr);
        if (s->Vm[i] <= 0.0 || s->Vm[i] > 10.0) {
            char msg[160];
            snprintf(msg, sizeof msg, "invalid voltage magnitude %f at bus %d", s->Vm[i], i);
            set_err(err, errlen, msg);
            return 1;
        }
    }
    return 0;
}

static int resolve_threads(int n_threads, int T) {
#ifdef _OPENMP
    int nt = n_threads <= 0 ? omp_get_max_threads() : n_threads;
    if (nt > omp_get_num_procs()) nt = omp_get_num_procs();
    if (nt > T) nt = T;
    return nt < 1 ? 1 : nt;
#else
    (void)n_threads; (void)T;
    return 1;
#endif
}

/* p3s/cpp/nr_klu.cpp:1049-1166. S: (n, T) complex row-major; V0: (n,) or (n, T).
 * Outputs V (n, T) complex row-major, iterations (T), converged (T). Returns 0, or 1 with
 * err set (the caller prefixes "solve_batch: "). */
int oracle_solve_batch(void* h, const double* S, const double* V0, int v0_per_col, int T, int max_iter, double tol,
                       int n_threads, int line_search, double* Vout, int* iters, unsigned char* conv, char* err,
                       int errlen) {
    otopo* t = (otopo*)h;
    if (T <= 0) { set_err(err, errlen, "T (time steps) must be > 0"); return 1; }
    if (max_iter <= 0) { set_err(err, errlen, "max_iter must be > 0"); return 1; }
    if (!(tol > 0.0)) { set_err(err, errlen, "tol must be > 0"); return 1; }
    if (n_threads < 0) { set_err(err, errlen, "n_threads must be >= 0"); return 1; }
    ts_ctx ctx = {S, V0, T, v0_per_col};
    int nt = resolve_threads(n_threads, T);
    int bad = 0;
    char first_err[256] = {0};
#pragma omp parallel num_threads(nt)
    {
        ostate s;
        state_init(&s, t);
        char lerr[256];
#pragma omp for schedule(dynamic)
        for (int col = 0; col < T; ++col) {
            int failed = 0;
            nres r = run_column(t, &s, ts_init, &ctx, col, max_iter, tol, line_search, lerr, sizeof lerr, &failed);
            if (failed) {
#pragma omp critical
                { if (!bad) { bad = 1; memcpy(first_err, lerr, sizeof lerr); } }
                continue;
            }
            for (int i = 0; i < t->n; ++i) {
                size_t o = ((size_t)i * T + col) * 2;
                Vout[o] = s.Vm[i] * cos(s.Va[i]);
                Vout[o + 1] = s.Vm[i] * sin(s.Va[i]);
            }
            iters[col] = r.iterations;
            conv[col] = (unsigned char)r.converged;
        }
        state_free(&s);
    }
    if (bad) { set_err(err, errlen, first_err); return 1; }
    return 0;
}

/* ---------------------------------------------------------------- contingency batch */
typedef struct {
    const double *Ymag, *Yang, *S, *V0;
    const unsigned char* pin;
    int L, nnz;
} n1_ctx;

typedef struct { double *Ym, *Ya; unsigned char* pin; } n1_slot;

static int n1_init_impl(ostate* s, const otopo* t, const n1_ctx* c, n1_slot* slot, int col) {
    for (int k = 0; k < t->nnzY; ++k) {
        slot->Ym[k] = c->Ymag[(size_t)k * c->L + col];
        slot->Ya[k] = c->Yang[(size_t)k * c->L + col];
    }
    for (int i = 0; i < t->n; ++i) slot->pin[i] = c->pin[(size_t)i * c->L + col];
    s->Ym = slot->Ym; s->Ya = slot->Ya; s->pin = slot->pin;
    for (int i = 0; i < t->n; ++i) {
        size_t vi = ((size_t)i * c->L + col) * 2;
        s->Pspec[i] = c->S[2 * (size_t)i]; s->Qspec[i] = c->S[2 * (size_t)i + 1];
        s->Vm[i] = hypot(c->V0[vi], c->V0[vi + 1]); s->Va[i] = atan2(c->V0[vi + 1], c->V0[vi]);
    }
    return 0;
}

typedef struct { const n1_ctx* c; n1_slot* slot; } n1_bound;
static int n1_init(ostate* s, const otopo* t, const void* v, int col, char* err, int errlen) {
    (void)err; (void)errlen;
    const n1_bound* b = (const n1_bound*)v;
    return n1_init_impl(s, t, b->c, b->slot, col);
}

/* p3s/cpp/nr_klu.cpp:1184-1281. Ymag/Yang: (nnz, L); S: (n,) complex; V0: (n, L) complex;
 * pin: (n, L) uint8. */
int oracle_solve_batch_contingency(void* h, const double* Ymag, const double* Yang, const double* S, const double* V0,
                                   const unsigned char* pin, int L, int max_iter, double tol, int n_threads,
                                   int line_search, double* Vout, int* iters, unsigned char* conv, char* err,
                                   int errlen) {
    otopo* t = (otopo*)h;
    if (L <= 0) { set_err(err, errlen, "L (cases) must be > 0"); return 1; }
    if (max_iter <= 0) { set_err(err, errlen, "max_iter must be > 0"); return 1; }
    if (!(tol > 0.0)) { set_err(err, errlen, "tol must be > 0"); return 1; }
    n1_ctx ctx = {Ymag, Yang, S, V0, pin, L, t->nnzY};
    int nt = resolve_threads(n_threads, L);
#pragma omp parallel num_threads(nt)
    {
        ostate s;
        state_init(&s, t);
        n1_slot slot = {NEWV(double, t->nnzY), NEWV(double, t->nnzY), NEWV(unsigned char, t->n)};
        n1_bound b = {&ctx, &slot};
        char lerr[64];
#pragma omp for schedule(dynamic)
        for (int col = 0; col < L; ++col) {
            int failed = 0;
            nres r = run_column(t, &s, n1_init, &b, col, max_iter, tol, line_search, lerr, sizeof lerr, &failed);
            for (int i = 0; i < t->n; ++i) {
                size_t o = ((size_t)i * L + col) * 2;
                Vout[o] = s.Vm[i] * cos(s.Va[i]);
                Vout[o + 1] = s.Vm[i] * sin(s.Va[i]);
            }
            iters[col] = r.iterations;
            conv[col] = (unsigned char)r.converged;
        }
        free(slot.Ym); free(slot.Ya); free(slot.pin);
        state_free(&s);
    }
    return 0;
}

/* Jacobian (CSC, the oracle's own column order) and mismatch F at V0: the reference's
 * debug_J (p3s/cpp/nr_klu.cpp:1284-1304). Jp (m+1), Ji/Jx (nnzJ), F (m). */
int oracle_debug_J(void* h, const double* S, const double* V0, int* Jp, int* Ji, double* Jx, double* F) {
    otopo* t = (otopo*)h;
    ostate s;
    state_init(&s, t);
    for (int i = 0; i < t->n; ++i) {
        s.Pspec[i] = S[2 * i]; s.Qspec[i] = S[2 * i + 1];
        s.Vm[i] = hypot(V0[2 * i], V0[2 * i + 1]); s.Va[i] = atan2(V0[2 * i + 1], V0[2 * i]);
    }
    eval_F_and_J(t, &s, s.Vm, s.Va, s.F);
    memcpy(Jp, t->Jp, (size_t)(t->m + 1) * sizeof(int));
    memcpy(Ji, t->Ji, (size_t)t->nnzJ * sizeof(int));
    memcpy(Jx, s.Jx, (size_t)t->nnzJ * sizeof(double));
    memcpy(F, s.F, (size_t)t->m * sizeof(double));
    state_free(&s);
    return 0;
}

/* One linear solve A x = b with the oracle's pivoting LU on a CSC matrix (tests pin the
 * stand-in for KLU against a dense solve). Returns 0 on success, 1 if singular. */
int oracle_gp_solve(int m, const int* Ap, const int* Ai, const double* Ax, int ordering, double* b) {
    int* Q = NEWV(int, m);
    if (ordering == 2) for (int k = 0; k < m; ++k) Q[k] = k;
    else o_min_degree(m, Ap, Ai, Q);
    gplu f;
    memset(&f, 0, sizeof f);
    int ok = gp_factor(&f, m, Ap, Ai, Ax, Q);
    if (ok) {
        double* y = NEWV(double, m);
        gp_solve(&f, Q, b, y);
        free(y);
    }
    gp_free(&f);
    free(Q);
    return ok ? 0 : 1;
}

/* Minimum-degree order of a CSC pattern (tests compare its fill with natural order). */
int oracle_min_degree(int m, const int* Ap, const int* Ai, int* perm) {
    o_min_degree(m, Ap, Ai, perm);
    return 0;
}

int oracle_openmp(void) {
#ifdef _OPENMP
    return 1;
#else
    return 0;
#endif
}
