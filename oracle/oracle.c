/*
 * CPU oracle for the AC Newton-Raphson N-1 path — TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * FP64 throughout. Own sparse LU (Gilbert-Peierls left-looking with threshold partial pivoting (KLU's default
 * tolerance 1e-3: the diagonal is kept unless 1000x smaller than the column maximum), after a
 * bus-level minimum degree order), own fill, own Newton loop with the reference's iteration accounting
 * and statuses, own local reactive-limit outer loop, own island detection. Nothing here is shared with the
 * GPU library: the GPU path builds its own pattern, ordering and factorization.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------------------ */
/* Branch formulas.                                                                                   */
/* The pi-branch flows in OLF's polar form, restated from the reference's generated header            */
/* native/cuda/equations/generated/closed_branch.cuh:31-229 (OLF ClosedBranchSide{1,2}*FlowEquationTerm) */
/* with R2 = 1, A2 = 0; derivatives derived by hand and pinned by finite differences and by the        */
/* textbook S = V conj(Y V) form in tests/test_oracle_formulas.py.                                    */
/* ------------------------------------------------------------------------------------------------ */
void oracle_closed_branch(const double *p8, double v1, double ph1, double v2, double ph2, double *o)
{
    const double y = p8[0], g1 = p8[1], g2 = p8[2], b1 = p8[3], b2 = p8[4], ksi = p8[5], r1 = p8[6], a1 = p8[7];
    const double t1 = ksi - a1 - ph1 + ph2; /* theta seen from side 1 */
    const double t2 = ksi + a1 + ph1 - ph2; /* theta seen from side 2 */
    const double sk = sin(ksi), ck = cos(ksi);
    const double s1 = sin(t1), c1 = cos(t1), s2 = sin(t2), c2 = cos(t2);
    const double yv1v2 = y * r1 * v1 * v2;
    /* side 1 active */
    o[0] = r1 * v1 * (g1 * r1 * v1 + y * r1 * v1 * sk - y * v2 * s1);
    o[1] = r1 * (2.0 * g1 * r1 * v1 + 2.0 * y * r1 * v1 * sk - y * v2 * s1);
    o[2] = -y * r1 * v1 * s1;
    o[3] = yv1v2 * c1;
    o[4] = -yv1v2 * c1;
    o[5] = yv1v2 * c1;
    o[6] = v1 * (2.0 * r1 * v1 * (g1 + y * sk) - y * v2 * s1);
    /* side 1 reactive */
    o[7] = r1 * v1 * (-b1 * r1 * v1 + y * r1 * v1 * ck - y * v2 * c1);
    o[8] = r1 * (-2.0 * b1 * r1 * v1 + 2.0 * y * r1 * v1 * ck - y * v2 * c1);
    o[9] = -y * r1 * v1 * c1;
    o[10] = -yv1v2 * s1;
    o[11] = yv1v2 * s1;
    o[12] = -yv1v2 * s1;
    o[13] = v1 * (2.0 * r1 * v1 * (-b1 + y * ck) - y * v2 * c1);
    /* side 2 active */
    o[14] = v2 * (g2 * v2 - y * r1 * v1 * s2 + y * v2 * sk);
    o[15] = -y * r1 * v2 * s2;
    o[16] = 2.0 * g2 * v2 - y * r1 * v1 * s2 + 2.0 * y * v2 * sk;
    o[17] = -yv1v2 * c2;
    o[18] = yv1v2 * c2;
    o[19] = -yv1v2 * c2;
    o[20] = -y * v1 * v2 * s2;
    /* side 2 reactive */
    o[21] = v2 * (-b2 * v2 - y * r1 * v1 * c2 + y * v2 * ck);
    o[22] = -y * r1 * v2 * c2;
    o[23] = -2.0 * b2 * v2 - y * r1 * v1 * c2 + 2.0 * y * v2 * ck;
    o[24] = yv1v2 * s2;
    o[25] = -yv1v2 * s2;
    o[26] = yv1v2 * s2;
    o[27] = -y * v1 * v2 * c2;
}

/* ------------------------------------------------------------------------------------------------ */
/* System: stable rows and the superset CSR pattern.                                                 */
/* ------------------------------------------------------------------------------------------------ */
struct oracle_sys {
    int n_bus, n_branch, n, nnz;
    double *p8;
    int *bus1, *bus2;
    int *adjptr, *adj;   /* bus adjacency, sorted, no duplicates, no self */
    int *rowptr, *colidx;
    int *diag;           /* CSR position of (r, r) */
    int *pos;            /* 16 CSR positions per branch: rows Pa Qa Pb Qb x cols pha va phb vb */
    int *colorder;       /* scalar column order for the LU (bus min degree, phi then V) */
    double *shunt_g, *shunt_b;   /* per bus, pu (NULL: no shunts) */
};

static int find_col(const int *colidx, int lo, int hi, int c)
{
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (colidx[mid] < c) lo = mid + 1; else hi = mid;
    }
    return lo;
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

oracle_sys *oracle_sys_create(int n_bus, int n_branch, const double *p8, const int *idx12)
{
    oracle_sys *s = (oracle_sys *)calloc(1, sizeof(oracle_sys));
    s->n_bus = n_bus;
    s->n_branch = n_branch;
    s->n = 2 * n_bus;
    s->p8 = (double *)malloc(sizeof(double) * 8 * (size_t)(n_branch > 0 ? n_branch : 1));
    memcpy(s->p8, p8, sizeof(double) * 8 * (size_t)n_branch);
    s->bus1 = (int *)malloc(sizeof(int) * (size_t)(n_branch + 1));
    s->bus2 = (int *)malloc(sizeof(int) * (size_t)(n_branch + 1));
    for (int e = 0; e < n_branch; e++) {
        s->bus1[e] = idx12[12 * e + 0] / 2;
        s->bus2[e] = idx12[12 * e + 2] / 2;
    }
    /* adjacency with duplicates, then sort + unique per bus */
    int *deg = (int *)calloc((size_t)n_bus + 1, sizeof(int));
    for (int e = 0; e < n_branch; e++) {
        deg[s->bus1[e]]++;
        deg[s->bus2[e]]++;
    }
    int *ptr = (int *)malloc(sizeof(int) * ((size_t)n_bus + 1));
    ptr[0] = 0;
    for (int i = 0; i < n_bus; i++) ptr[i + 1] = ptr[i] + deg[i];
    int *tmp = (int *)malloc(sizeof(int) * (size_t)(ptr[n_bus] + 1));
    int *fillp = (int *)malloc(sizeof(int) * ((size_t)n_bus + 1));
    memcpy(fillp, ptr, sizeof(int) * (size_t)n_bus);
    for (int e = 0; e < n_branch; e++) {
        tmp[fillp[s->bus1[e]]++] = s->bus2[e];
        tmp[fillp[s->bus2[e]]++] = s->bus1[e];
    }
    s->adjptr = (int *)malloc(sizeof(int) * ((size_t)n_bus + 1));
    s->adj = (int *)malloc(sizeof(int) * (size_t)(ptr[n_bus] + 1));
    s->adjptr[0] = 0;
    for (int i = 0; i < n_bus; i++) {
        int lo = ptr[i], hi = ptr[i + 1];
        qsort(tmp + lo, (size_t)(hi - lo), sizeof(int), cmp_int);
        int w = s->adjptr[i];
        for (int k = lo; k < hi; k++) {
            if (tmp[k] == i) continue;
            if (w > s->adjptr[i] && s->adj[w - 1] == tmp[k]) continue;
            s->adj[w++] = tmp[k];
        }
        s->adjptr[i + 1] = w;
    }
    free(deg); free(ptr); free(tmp); free(fillp);

    /* scalar CSR: rows 2i, 2i+1 hold cols {2j, 2j+1 : j in {i} U adj(i)} ascending */
    const int n = s->n;
    s->rowptr = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    s->rowptr[0] = 0;
    for (int i = 0; i < n_bus; i++) {
        int cnt = 2 * (1 + s->adjptr[i + 1] - s->adjptr[i]);
        s->rowptr[2 * i + 1] = s->rowptr[2 * i] + cnt;
        s->rowptr[2 * i + 2] = s->rowptr[2 * i + 1] + cnt;
    }
    s->nnz = s->rowptr[n];
    s->colidx = (int *)malloc(sizeof(int) * (size_t)s->nnz);
    for (int i = 0; i < n_bus; i++) {
        for (int rr = 0; rr < 2; rr++) {
            int w = s->rowptr[2 * i + rr];
            int inserted = 0;
            for (int k = s->adjptr[i]; k <= s->adjptr[i + 1]; k++) {
                int j = (k < s->adjptr[i + 1]) ? s->adj[k] : n_bus + 1;
                if (!inserted && i < j) {
                    s->colidx[w++] = 2 * i;
                    s->colidx[w++] = 2 * i + 1;
                    inserted = 1;
                }
                if (k < s->adjptr[i + 1]) {
                    s->colidx[w++] = 2 * j;
                    s->colidx[w++] = 2 * j + 1;
                }
            }
        }
    }
    s->diag = (int *)malloc(sizeof(int) * (size_t)n);
    for (int r = 0; r < n; r++) s->diag[r] = find_col(s->colidx, s->rowptr[r], s->rowptr[r + 1], r);
    s->pos = (int *)malloc(sizeof(int) * 16 * (size_t)(n_branch + 1));
    for (int e = 0; e < n_branch; e++) {
        int a = s->bus1[e], b = s->bus2[e];
        int rows[4] = {2 * a, 2 * a + 1, 2 * b, 2 * b + 1};
        int cols[4] = {2 * a, 2 * a + 1, 2 * b, 2 * b + 1};
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++)
                s->pos[16 * e + 4 * r + c] = find_col(s->colidx, s->rowptr[rows[r]], s->rowptr[rows[r] + 1], cols[c]);
    }
    /* LU column order */
    int *perm = (int *)malloc(sizeof(int) * (size_t)(n_bus + 1));
    oracle_min_degree(n_bus, s->adjptr, s->adj, perm);
    s->colorder = (int *)malloc(sizeof(int) * (size_t)n);
    for (int k = 0; k < n_bus; k++) {
        s->colorder[2 * k] = 2 * perm[k];
        s->colorder[2 * k + 1] = 2 * perm[k] + 1;
    }
    free(perm);
    return s;
}

void oracle_sys_free(oracle_sys *s)
{
    if (!s) return;
    free(s->p8); free(s->bus1); free(s->bus2); free(s->adjptr); free(s->adj);
    free(s->rowptr); free(s->colidx); free(s->diag); free(s->pos); free(s->colorder);
    free(s->shunt_g); free(s->shunt_b);
    free(s);
}

int oracle_sys_nnz(const oracle_sys *s) { return s->nnz; }

void oracle_sys_pattern(const oracle_sys *s, int *rowptr, int *colidx)
{
    memcpy(rowptr, s->rowptr, sizeof(int) * ((size_t)s->n + 1));
    memcpy(colidx, s->colidx, sizeof(int) * (size_t)s->nnz);
}

static int is_out(int e, const int *out_br, int n_out)
{
    for (int k = 0; k < n_out; k++)
        if (out_br[k] == e) return 1;
    return 0;
}

/* reference equations/fill_launch.cu:82-126 (F = -target, family sums, identity rows). */
void oracle_sys_set_shunts(oracle_sys *s, const double *g, const double *b)
{
    free(s->shunt_g);
    free(s->shunt_b);
    s->shunt_g = s->shunt_b = NULL;
    if (!g || !b) return;
    s->shunt_g = (double *)malloc(sizeof(double) * (size_t)s->n_bus);
    s->shunt_b = (double *)malloc(sizeof(double) * (size_t)s->n_bus);
    memcpy(s->shunt_g, g, sizeof(double) * (size_t)s->n_bus);
    memcpy(s->shunt_b, b, sizeof(double) * (size_t)s->n_bus);
}

void oracle_fill(const oracle_sys *s, const double *x, const int *mode, const double *target,
                 const int *out_br, int n_out, double *F, double *qinj, double *J)
{
    const int n = s->n;
    memset(F, 0, sizeof(double) * (size_t)n);
    if (qinj) memset(qinj, 0, sizeof(double) * (size_t)s->n_bus);
    if (J) memset(J, 0, sizeof(double) * (size_t)s->nnz);
    double o[28];
    for (int e = 0; e < s->n_branch; e++) {
        if (n_out > 0 && is_out(e, out_br, n_out)) continue;
        const int a = s->bus1[e], b = s->bus2[e];
        oracle_closed_branch(s->p8 + 8 * e, x[2 * a + 1], x[2 * a], x[2 * b + 1], x[2 * b], o);
        F[2 * a] += o[0];
        F[2 * a + 1] += o[7];
        F[2 * b] += o[14];
        F[2 * b + 1] += o[21];
        if (qinj) {
            qinj[a] += o[7];
            qinj[b] += o[21];
        }
        if (J) {
            const int *p = s->pos + 16 * e;
            /* row Pa: cols pha va phb vb */
            J[p[0]] += o[3]; J[p[1]] += o[1]; J[p[2]] += o[4]; J[p[3]] += o[2];
            /* row Qa */
            J[p[4]] += o[10]; J[p[5]] += o[8]; J[p[6]] += o[11]; J[p[7]] += o[9];
            /* row Pb */
            J[p[8]] += o[17]; J[p[9]] += o[15]; J[p[10]] += o[18]; J[p[11]] += o[16];
            /* row Qb */
            J[p[12]] += o[24]; J[p[13]] += o[22]; J[p[14]] += o[25]; J[p[15]] += o[23];
        }
    }
    if (s->shunt_g) {
        /* bus shunt (reference generated/shunt.cuh:47-60): p = g v^2, q = -b v^2 */
        for (int a = 0; a < s->n_bus; a++) {
            const double v = x[2 * a + 1], g = s->shunt_g[a], bb = s->shunt_b[a];
            if (g == 0.0 && bb == 0.0) continue;
            F[2 * a] += g * v * v;
            F[2 * a + 1] -= bb * v * v;
            if (qinj) qinj[a] -= bb * v * v;
            if (J) {
                J[s->diag[2 * a] + 1] += 2.0 * g * v;      /* row P_a, column V_a (right after column phi_a) */
                J[s->diag[2 * a + 1]] -= 2.0 * bb * v;     /* row Q_a, column V_a */
            }
        }
    }
    for (int r = 0; r < n; r++) {
        if (mode[r] == ORC_FIXED_VARIABLE) {
            F[r] = x[r] - target[r];
            if (J) {
                for (int k = s->rowptr[r]; k < s->rowptr[r + 1]; k++) J[k] = 0.0;
                J[s->diag[r]] = 1.0;
            }
        } else {
            F[r] -= target[r];
        }
    }
}

/* ------------------------------------------------------------------------------------------------ */
/* Minimum degree on an explicit elimination graph (sorted neighbour sets), lowest index on ties.     */
/* ------------------------------------------------------------------------------------------------ */
typedef struct { int *v; int n, cap; } ivec;

static void iv_push(ivec *a, int x)
{
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 8;
        a->v = (int *)realloc(a->v, sizeof(int) * (size_t)a->cap);
    }
    a->v[a->n++] = x;
}

void oracle_min_degree(int n, const int *adjptr, const int *adj, int *perm)
{
    ivec *g = (ivec *)calloc((size_t)n + 1, sizeof(ivec));
    for (int i = 0; i < n; i++)
        for (int k = adjptr[i]; k < adjptr[i + 1]; k++)
            if (adj[k] != i) iv_push(&g[i], adj[k]);
    for (int i = 0; i < n; i++) qsort(g[i].v, (size_t)g[i].n, sizeof(int), cmp_int);
    char *dead = (char *)calloc((size_t)n + 1, 1);
    int *merged = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    for (int step = 0; step < n; step++) {
        int best = -1, bd = 0x7fffffff;
        for (int i = 0; i < n; i++)
            if (!dead[i] && g[i].n < bd) { bd = g[i].n; best = i; }
        perm[step] = best;
        dead[best] = 1;
        ivec nb = g[best];
        for (int t = 0; t < nb.n; t++) {
            int u = nb.v[t];
            /* merged = (g[u] U nb) \ {u, best}, sorted */
            int i1 = 0, i2 = 0, w = 0;
            while (i1 < g[u].n || i2 < nb.n) {
                int c;
                if (i2 >= nb.n || (i1 < g[u].n && g[u].v[i1] < nb.v[i2])) c = g[u].v[i1++];
                else if (i1 >= g[u].n || nb.v[i2] < g[u].v[i1]) c = nb.v[i2++];
                else { c = g[u].v[i1++]; i2++; }
                if (c == u || c == best) continue;
                merged[w++] = c;
            }
            g[u].n = 0;
            for (int k = 0; k < w; k++) iv_push(&g[u], merged[k]);
        }
        free(g[best].v);
        g[best].v = NULL;
        g[best].n = 0;
    }
    for (int i = 0; i < n; i++) free(g[i].v);
    free(g); free(dead); free(merged);
}

/* ------------------------------------------------------------------------------------------------ */
/* Sparse LU: Gilbert-Peierls left-looking, threshold partial pivoting (diagonal kept if within 1e-3 of */
/* the column maximum), columns in a given order.                                                     */
/* ------------------------------------------------------------------------------------------------ */
typedef struct {
    int n;
    int *Lp, *Li; double *Lx; size_t Lcap;
    int *Up, *Ui; double *Ux; size_t Ucap;
    double *Ud;
    int *pinv, *prow;
} olu;

static void grow(int **i, double **x, size_t *cap, size_t need)
{
    if (need <= *cap) return;
    size_t c = *cap ? *cap : 1024;
    while (c < need) c *= 2;
    *i = (int *)realloc(*i, sizeof(int) * c);
    *x = (double *)realloc(*x, sizeof(double) * c);
    *cap = c;
}

static void olu_free(olu *f)
{
    free(f->Lp); free(f->Li); free(f->Lx); free(f->Up); free(f->Ui); free(f->Ux);
    free(f->Ud); free(f->pinv); free(f->prow);
}

/* A given in CSC (Ap, Ai, Ax). */
static int olu_factor(int n, const int *Ap, const int *Ai, const double *Ax, const int *q, olu *f)
{
    memset(f, 0, sizeof(*f));
    f->n = n;
    f->Lp = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    f->Up = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    f->Ud = (double *)malloc(sizeof(double) * (size_t)n);
    f->pinv = (int *)malloc(sizeof(int) * (size_t)n);
    f->prow = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; i++) f->pinv[i] = -1;
    double *x = (double *)calloc((size_t)n, sizeof(double));
    int *stamp = (int *)calloc((size_t)n, sizeof(int));     /* row in pattern of step j: stamp == j+1 */
    int *mark = (int *)calloc((size_t)n, sizeof(int));      /* L column visited in step j: mark == j+1 */
    int *xi = (int *)malloc(sizeof(int) * (size_t)n);       /* topological order of reached columns */
    int *stk = (int *)malloc(sizeof(int) * (size_t)n);
    int *pstk = (int *)malloc(sizeof(int) * (size_t)n);
    int *patt = (int *)malloc(sizeof(int) * (size_t)n);
    size_t lnz = 0, unz = 0;
    int rc = 0;
    for (int j = 0; j < n; j++) {
        const int col = q[j];
        const int tag = j + 1;
        f->Lp[j] = (int)lnz;
        f->Up[j] = (int)unz;
        /* reach of A(:,col) through L */
        int top = n;
        for (int p = Ap[col]; p < Ap[col + 1]; p++) {
            int k0 = f->pinv[Ai[p]];
            if (k0 < 0 || mark[k0] == tag) continue;
            int sp = 0;
            stk[0] = k0;
            pstk[0] = f->Lp[k0];
            mark[k0] = tag;
            while (sp >= 0) {
                int k = stk[sp];
                int end = f->Lp[k + 1];
                int pushed = 0;
                for (int pp = pstk[sp]; pp < end; pp++) {
                    int kk = f->pinv[f->Li[pp]];
                    if (kk < 0 || mark[kk] == tag) continue;
                    pstk[sp] = pp + 1;
                    mark[kk] = tag;
                    sp++;
                    stk[sp] = kk;
                    pstk[sp] = f->Lp[kk];
                    pushed = 1;
                    break;
                }
                if (!pushed) {
                    xi[--top] = k;
                    sp--;
                }
            }
        }
        /* pattern = rows of A(:,col) U rows of reached L columns U their pivot rows */
        int np = 0;
        for (int p = Ap[col]; p < Ap[col + 1]; p++) {
            int r = Ai[p];
            if (stamp[r] != tag) { stamp[r] = tag; patt[np++] = r; x[r] = 0.0; }
        }
        for (int t = top; t < n; t++) {
            int k = xi[t];
            int pr = f->prow[k];
            if (stamp[pr] != tag) { stamp[pr] = tag; patt[np++] = pr; x[pr] = 0.0; }
            for (int pp = f->Lp[k]; pp < f->Lp[k + 1]; pp++) {
                int r = f->Li[pp];
                if (stamp[r] != tag) { stamp[r] = tag; patt[np++] = r; x[r] = 0.0; }
            }
        }
        for (int p = Ap[col]; p < Ap[col + 1]; p++) x[Ai[p]] += Ax[p];
        /* sparse triangular solve, topological order */
        grow(&f->Ui, &f->Ux, &f->Ucap, unz + (size_t)(n - top) + 1);
        for (int t = top; t < n; t++) {
            int k = xi[t];
            double xk = x[f->prow[k]];
            f->Ui[unz] = k;
            f->Ux[unz] = xk;
            unz++;
            if (xk == 0.0) continue;
            for (int pp = f->Lp[k]; pp < f->Lp[k + 1]; pp++) x[f->Li[pp]] -= f->Lx[pp] * xk;
        }
        /* pivot */
        int piv = -1;
        double amax = 0.0;
        for (int t = 0; t < np; t++) {
            int r = patt[t];
            if (f->pinv[r] >= 0) continue;
            double a = fabs(x[r]);
            if (a > amax) { amax = a; piv = r; }
        }
        if (piv < 0 || amax == 0.0 || !isfinite(amax)) { rc = -1; f->Lp[j + 1] = (int)lnz; f->Up[j + 1] = (int)unz; break; }
        if (stamp[col] == tag && f->pinv[col] < 0 && fabs(x[col]) >= 1e-3 * amax) piv = col;
        double d = x[piv];
        f->Ud[j] = d;
        f->pinv[piv] = j;
        f->prow[j] = piv;
        grow(&f->Li, &f->Lx, &f->Lcap, lnz + (size_t)np + 1);
        for (int t = 0; t < np; t++) {
            int r = patt[t];
            if (f->pinv[r] >= 0) continue;
            if (x[r] != 0.0) {
                f->Li[lnz] = r;
                f->Lx[lnz] = x[r] / d;
                lnz++;
            }
        }
        f->Lp[j + 1] = (int)lnz;
        f->Up[j + 1] = (int)unz;
    }
    free(x); free(stamp); free(mark); free(xi); free(stk); free(pstk); free(patt);
    return rc;
}

static void olu_solve(const olu *f, const int *q, const double *b, double *out, double *w)
{
    const int n = f->n;
    memcpy(w, b, sizeof(double) * (size_t)n);  /* indexed by original rows */
    double *y = out;                            /* temporarily indexed by step */
    for (int j = 0; j < n; j++) {
        double yj = w[f->prow[j]];
        y[j] = yj;
        if (yj == 0.0) continue;
        for (int p = f->Lp[j]; p < f->Lp[j + 1]; p++) w[f->Li[p]] -= f->Lx[p] * yj;
    }
    for (int j = n - 1; j >= 0; j--) {
        double z = y[j] / f->Ud[j];
        y[j] = z;
        if (z == 0.0) continue;
        for (int p = f->Up[j]; p < f->Up[j + 1]; p++) y[f->Ui[p]] -= f->Ux[p] * z;
    }
    memcpy(w, y, sizeof(double) * (size_t)n);
    for (int j = 0; j < n; j++) out[q[j]] = w[j];
}

static void csr_to_csc(int n, const int *rowptr, const int *colidx, const double *vals, int *Ap, int *Ai, double *Ax)
{
    int nnz = rowptr[n];
    memset(Ap, 0, sizeof(int) * ((size_t)n + 1));
    for (int k = 0; k < nnz; k++) Ap[colidx[k] + 1]++;
    for (int c = 0; c < n; c++) Ap[c + 1] += Ap[c];
    int *w = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    memcpy(w, Ap, sizeof(int) * (size_t)n);
    for (int r = 0; r < n; r++)
        for (int k = rowptr[r]; k < rowptr[r + 1]; k++) {
            int d = w[colidx[k]]++;
            Ai[d] = r;
            Ax[d] = vals[k];
        }
    free(w);
}

static int lu_solve_ordered(int n, const int *rowptr, const int *colidx, const double *vals, const int *q,
                            const double *b, double *x)
{
    int nnz = rowptr[n];
    int *Ap = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    int *Ai = (int *)malloc(sizeof(int) * ((size_t)nnz + 1));
    double *Ax = (double *)malloc(sizeof(double) * ((size_t)nnz + 1));
    csr_to_csc(n, rowptr, colidx, vals, Ap, Ai, Ax);
    olu f;
    int rc = olu_factor(n, Ap, Ai, Ax, q, &f);
    if (rc == 0) {
        double *w = (double *)malloc(sizeof(double) * (size_t)n);
        olu_solve(&f, q, b, x, w);
        free(w);
    }
    olu_free(&f);
    free(Ap); free(Ai); free(Ax);
    return rc;
}

int oracle_lu_solve(int n, const int *rowptr, const int *colidx, const double *vals, const double *b, double *x)
{
    /* symmetric-pattern minimum degree on the pattern of A + A^T */
    int nnz = rowptr[n];
    int *deg = (int *)calloc((size_t)n + 1, sizeof(int));
    for (int r = 0; r < n; r++)
        for (int k = rowptr[r]; k < rowptr[r + 1]; k++)
            if (colidx[k] != r) { deg[r]++; deg[colidx[k]]++; }
    int *ptr = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    ptr[0] = 0;
    for (int i = 0; i < n; i++) ptr[i + 1] = ptr[i] + deg[i];
    int *adj = (int *)malloc(sizeof(int) * ((size_t)2 * nnz + 1));
    int *fp = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    memcpy(fp, ptr, sizeof(int) * (size_t)n);
    for (int r = 0; r < n; r++)
        for (int k = rowptr[r]; k < rowptr[r + 1]; k++)
            if (colidx[k] != r) { adj[fp[r]++] = colidx[k]; adj[fp[colidx[k]]++] = r; }
    int *q = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    oracle_min_degree(n, ptr, adj, q);
    int rc = lu_solve_ordered(n, rowptr, colidx, vals, q, b, x);
    free(deg); free(ptr); free(adj); free(fp); free(q);
    return rc;
}

/* ------------------------------------------------------------------------------------------------ */
/* Newton-Raphson with the local reactive-limit outer loop.                                           */
/* ------------------------------------------------------------------------------------------------ */
#define LIN_SOLVE_FAIL_RATIO 1e-2 /* reference solver/iterative_refinement.cuh:23 */

static double inf_norm(const double *v, int n, int *finite)
{
    double m = 0.0;
    int fin = 1;
    for (int i = 0; i < n; i++) {
        if (!isfinite(v[i])) fin = 0;
        double a = fabs(v[i]);
        if (a > m) m = a;
    }
    if (finite) *finite = fin;
    return m;
}

/* reference outerloop/reactive_limits.cu:321-371 (local controllers) plus the forced-PV election
 * (reactive_limits.cu:417-, ReactiveLimitsOuterLoop.switchPvPq): returns whether anything switched. */
static int rl_check(const oracle_rl *rl, const oracle_nr_params *prm, const unsigned char *islanded,
                    const double *x, const double *qinj, int *mode, double *target, signed char *lt,
                    signed char *sc, signed char *pass)
{
    int remaining = 0;
    for (int i = 0; i < rl->n; i++) {
        pass[i] = 0;
        const int bus = rl->bus[i];
        if (islanded && islanded[bus]) continue;
        const int vr = rl->voltage_row[i];
        if (mode[vr] == ORC_FIXED_VARIABLE && lt[i] == ORC_QLIM_NONE) {
            double q = qinj[bus] + rl->load_q[i];
            if (q < rl->min_q[i] - prm->rl_tol) {
                mode[vr] = ORC_EQUATION;
                target[vr] = rl->min_q[i] - rl->load_q[i];
                lt[i] = ORC_QLIM_MIN;
                sc[i]++;
                pass[i] = 1;
            } else if (q > rl->max_q[i] + prm->rl_tol) {
                mode[vr] = ORC_EQUATION;
                target[vr] = rl->max_q[i] - rl->load_q[i];
                lt[i] = ORC_QLIM_MAX;
                sc[i]++;
                pass[i] = 1;
            } else {
                remaining++;
            }
        } else if (lt[i] != ORC_QLIM_NONE) {
            double v = x[vr], tv = rl->target_v[i];
            int wants = (lt[i] == ORC_QLIM_MAX && v > tv) || (lt[i] == ORC_QLIM_MIN && v < tv);
            if (wants && sc[i] < prm->max_switch) {
                mode[vr] = ORC_FIXED_VARIABLE;
                target[vr] = tv;
                lt[i] = ORC_QLIM_NONE;
                pass[i] = 2;
            }
        }
    }
    if (remaining == 0) {
        /* forced PV: the strongest bus switched PV -> PQ in this check stays PV (all nominal voltages are
         * equal here: highest P target, then lowest entry) */
        int best = -1;
        double bp = -1e300;
        for (int i = 0; i < rl->n; i++) {
            if (pass[i] != 1) continue;
            double p = target[2 * rl->bus[i]];
            if (p > bp) { bp = p; best = i; }
        }
        if (best >= 0) {
            const int vr = rl->voltage_row[best];
            mode[vr] = ORC_FIXED_VARIABLE;
            target[vr] = rl->target_v[best];
            lt[best] = ORC_QLIM_NONE;
            sc[best]--;
            pass[best] = 0;
        }
    }
    for (int i = 0; i < rl->n; i++)
        if (pass[i]) return 1;
    return 0;
}

double oracle_bus_p(const oracle_sys *s, const double *x, const int *out_br, int n_out, int bus)
{
    double p = 0.0, o[28];
    for (int e = 0; e < s->n_branch; e++) {
        const int a = s->bus1[e], b = s->bus2[e];
        if (a != bus && b != bus) continue;
        if (n_out > 0 && is_out(e, out_br, n_out)) continue;
        oracle_closed_branch(s->p8 + 8 * e, x[2 * a + 1], x[2 * a], x[2 * b + 1], x[2 * b], o);
        p += a == bus ? o[0] : o[14];
    }
    if (s->shunt_g) p += s->shunt_g[bus] * x[2 * bus + 1] * x[2 * bus + 1];
    return p;
}

#define P_RESIDUE_EPS 1e-5 /* OLF ActivePowerDistribution.P_RESIDUE_EPS (outer_loop_driver.cu:204) */

/* One DistributedSlackOuterLoop check, the reference's aggregate (per-bus) path: the slack's retained share
 * (distributed_slack.cu:131-195, computeSlackRetain), the gate on the participants' share (:215-305), the reset of
 * every participant to its anchor and the water-fill of the whole total from there, factors renormalised over the
 * unsaturated participants each pass (:395-420, 530-566), and OLF's outcome: UNSTABLE when the targets (and the
 * slack's applied share) moved by more than 0.9 P_RESIDUE_EPS (:568-600). A participant whose P row is not an
 * equation (islanded) takes no part. *retain_applied is NaN before the first distribution. Returns 1 when unstable;
 * *failed when more than P_RESIDUE_EPS stayed unplaced. */
static int ds_check(const oracle_ds *ds, double p_slack, const int *mode, double *target, double *retain_applied,
                    double *prev, int *failed)
{
    const double gate = fmax(ds->tol, P_RESIDUE_EPS);
    const double m = p_slack - ds->slack_target_p;
    double fs_all = 0.0, fs_active = 0.0, absorbed = 0.0;
    for (int i = 0; i < ds->n; i++) {
        fs_all += ds->factor[i];
        const int row = 2 * ds->bus[i];
        if (mode[row] != ORC_EQUATION) continue;
        fs_active += ds->factor[i];
        absorbed += target[row] - ds->base_target_p[i];
    }
    const double rho = 1.0 - fs_all;
    const double rho_eff = fs_active != fs_all ? (fs_active + rho > 0 ? rho / (fs_active + rho) : 0.0) : rho;
    const double retain = rho_eff * (m + absorbed);
    double rem = m - retain;
    const double fs_global = fs_active == fs_all ? fs_active : (fs_active + rho > 0 ? fs_active / (fs_active + rho) : 0.0);
    const int no_capacity = !(fs_active > 0.0) && rho < 1e-12;
    *failed = 0;
    if (fs_global <= 0.0) {
        if (no_capacity) {
            if (fabs(m) > ds->tol && fabs(m) > P_RESIDUE_EPS) *failed = 1;
            return 0;
        }
        /* only the slack participates: it takes the whole mismatch, and the loop re-solves when that changed */
        const double before = isnan(*retain_applied) ? 0.0 : *retain_applied;
        if (fabs(m - before) > gate) {
            *retain_applied = m;
            return 1;
        }
        return 0;
    }
    if (fabs(rem) <= fs_global * gate) return 0;
    const double prev_retain = isnan(*retain_applied) ? 0.0 : *retain_applied;
    *retain_applied = retain;
    for (int i = 0; i < ds->n; i++) {
        const int row = 2 * ds->bus[i];
        if (mode[row] != ORC_EQUATION) continue;
        prev[i] = target[row];
        target[row] = ds->base_target_p[i];
    }
    rem = fs_global * (m + absorbed);
    for (int iter = 0; iter < ds->n && fabs(rem) > P_RESIDUE_EPS; iter++) {
        double fsum = 0.0;
        for (int i = 0; i < ds->n; i++) {
            const int row = 2 * ds->bus[i];
            if (mode[row] != ORC_EQUATION) continue;
            const double cum = target[row] - ds->base_target_p[i];
            const int sat = rem > 0 ? cum >= ds->max_delta[i] : cum <= ds->min_delta[i];
            if (!sat) fsum += ds->factor[i];
        }
        if (fsum <= 0.0) break;
        double done = 0.0;
        for (int i = 0; i < ds->n; i++) {
            const int row = 2 * ds->bus[i];
            if (mode[row] != ORC_EQUATION) continue;
            const double cum = target[row] - ds->base_target_p[i];
            const int sat = rem > 0 ? cum >= ds->max_delta[i] : cum <= ds->min_delta[i];
            if (sat) continue;
            const double next = fmin(fmax(cum + rem * (ds->factor[i] / fsum), ds->min_delta[i]), ds->max_delta[i]);
            if (next != cum) {
                target[row] = ds->base_target_p[i] + next;
                done += next - cum;
            }
        }
        rem -= done;
    }
    double moved = fabs(*retain_applied - prev_retain);
    for (int i = 0; i < ds->n; i++) {
        const int row = 2 * ds->bus[i];
        if (mode[row] != ORC_EQUATION) continue;
        moved += fabs(target[row] - prev[i]);
    }
    if (fabs(rem) > P_RESIDUE_EPS) *failed = 1;
    return moved > 0.9 * P_RESIDUE_EPS;
}

void oracle_nr_solve(const oracle_sys *s, const oracle_nr_params *prm, const oracle_rl *rl, const oracle_ds *ds,
                     const int *out_br, int n_out, const unsigned char *islanded,
                     double *x, int *mode, double *target, signed char *lt, signed char *sc, oracle_result *res)
{
    const int n = s->n;
    double *F = (double *)malloc(sizeof(double) * (size_t)n);
    double *b = (double *)malloc(sizeof(double) * (size_t)n);
    double *dx = (double *)malloc(sizeof(double) * (size_t)n);
    double *r = (double *)malloc(sizeof(double) * (size_t)n);
    double *J = (double *)malloc(sizeof(double) * (size_t)s->nnz);
    double *qinj = (double *)malloc(sizeof(double) * (size_t)s->n_bus);
    signed char *pass = (signed char *)malloc((size_t)(rl && rl->n > 0 ? rl->n : 1));
    double *prev = (double *)malloc(sizeof(double) * (size_t)(ds && ds->n > 0 ? ds->n : 1));
    double *xprev = (double *)malloc(sizeof(double) * (size_t)n);
    const int ls = prm->scaling_mode == 1;
    const int ls_max = prm->ls_max_iteration > 0 ? prm->ls_max_iteration : 10;
    const double ls_fold = prm->ls_step_fold > 1.0 ? prm->ls_step_fold : 4.0 / 3.0;
    memset(res, 0, sizeof(*res));
    const int ds_on = ds != NULL;
    const int rl_on = rl && rl->n > 0 && prm->reactive_limits;
    double retain_applied = NAN;
    const double fail_ratio = prm->lin_fail_ratio > 0.0 ? prm->lin_fail_ratio : LIN_SOLVE_FAIL_RATIO;
    const int fixed_iter = !(prm->tol > 0.0) && !(prm->conv_eps_per_eq > 0.0) && !(prm->row_thr_fixed && prm->row_thr_eq);
    int outer = 0;
    int loop = 0;             /* 0 distributed slack, 1 reactive limits, 2 end of the sweep */
    int sweep_changed = 0;
    for (;;) {
        int it = 0;
        double last_ratio = 0.0;
        int ended = 0, converged = 0;
        double norm = 0.0;
        int has_last = 0;          /* LINE_SEARCH: the 2-norm of the last tested step, per solve */
        double last_ssq = 0.0;
        for (;;) {
            oracle_fill(s, x, mode, target, out_br, n_out, F, qinj, J);
            int finite;
            norm = inf_norm(F, n, &finite);
            if (ls && it > 0) {
                /* LINE_SEARCH (newton_iteration.cu:497-530): a step that did not lower ||F||2 (or went non-finite)
                 * is walked back to xprev + mu dx, mu = fold^-b, b = 1.. ls_max, refilling each time */
                double ssq = 0.0;
                for (int i = 0; i < n; i++) ssq += F[i] * F[i];
                if (has_last && (!finite || ssq >= last_ssq)) {
                    for (int bt = 1; bt <= ls_max; bt++) {
                        const double mu = 1.0 / pow(ls_fold, bt);
                        for (int i = 0; i < n; i++) x[i] = xprev[i] + mu * dx[i];
                        oracle_fill(s, x, mode, target, out_br, n_out, F, qinj, J);
                        norm = inf_norm(F, n, &finite);
                        ssq = 0.0;
                        for (int i = 0; i < n; i++) ssq += F[i] * F[i];
                        if (finite && ssq < last_ssq) break;
                    }
                }
                last_ssq = ssq;
                has_last = 1;
            }
            res->norm = norm;
            res->iterations = it;
            if (fixed_iter) {
                /* exactly max_iter steps, no test; the status from the final mismatch */
                if (it == prm->max_iter) {
                    res->status = finite ? ORC_MAX_ITERATION_REACHED : ORC_DIVERGED;
                    ended = 1;
                    break;
                }
            } else if (!finite) { res->status = ORC_DIVERGED; ended = 1; break; }
            if (!fixed_iter && it > 0) {
                int below = norm < prm->tol;
                if (prm->row_thr_fixed && prm->row_thr_eq) {
                    below = 1;
                    for (int i = 0; i < n && below; i++)
                        below = fabs(F[i]) < (mode[i] == ORC_FIXED_VARIABLE ? prm->row_thr_fixed[i] : prm->row_thr_eq[i]);
                } else if (prm->conv_eps_per_eq > 0.0) {
                    double ssq = 0.0;
                    for (int i = 0; i < n; i++) ssq += F[i] * F[i];
                    const double na = prm->n_active > 0 ? (double)prm->n_active : (double)n;
                    below = sqrt(ssq) < sqrt(prm->conv_eps_per_eq * prm->conv_eps_per_eq * na);
                }
                if (below && it < prm->max_iter) { converged = 1; break; }
                if (it == prm->max_iter) {
                    int bad = !(last_ratio <= fail_ratio);
                    res->status = bad ? (prm->batch_of_one ? ORC_LINEAR_SOLVE_FAILED : ORC_REJECTED)
                                      : ORC_MAX_ITERATION_REACHED;
                    ended = 1;
                    break;
                }
            }
            for (int i = 0; i < n; i++) b[i] = -F[i];
            if (fixed_iter && !finite) {
                /* the fixed-iteration mode does not isolate a non-finite scenario: its steps stay non-finite */
                for (int i = 0; i < n; i++) dx[i] = NAN;
            } else if (lu_solve_ordered(n, s->rowptr, s->colidx, J, s->colorder, b, dx) != 0) {
                res->status = ORC_LINEAR_SOLVE_FAILED;
                ended = 1;
                break;
            }
            /* linear residual ratio ||J dx - b|| / ||F|| (reference newton_iteration.cu:612-652) */
            for (int i = 0; i < n; i++) {
                double acc = -b[i];
                for (int k = s->rowptr[i]; k < s->rowptr[i + 1]; k++) acc += J[k] * dx[s->colidx[k]];
                r[i] = acc;
            }
            double ratio = norm > 0.0 ? inf_norm(r, n, NULL) / norm : 0.0;
            if (!fixed_iter && res->solves == 0 && it == 1 && !(ratio <= fail_ratio)) {
                res->status = ORC_LINEAR_SOLVE_FAILED;
                ended = 1;
                break;
            }
            last_ratio = ratio;
            double f = 1.0;
            if (prm->scaling_mode == 2) {
                double mdv = 0.0, mdp = 0.0;
                for (int i = 0; i < n; i++) {
                    double a = fabs(dx[i]);
                    if (i & 1) { if (a > mdv) mdv = a; } else { if (a > mdp) mdp = a; }
                }
                if (mdv > prm->max_dv) f = fmin(f, prm->max_dv / mdv);
                if (mdp > prm->max_dphi) f = fmin(f, prm->max_dphi / mdp);
            }
            if (ls) memcpy(xprev, x, sizeof(double) * (size_t)n);
            for (int i = 0; i < n; i++) x[i] += f * dx[i];
            it++;
            res->total_iterations++;
        }
        if (ended) break;
        res->solves++;
        (void)converged;
        /* the outer loops, in order, on the converged state */
        int resolve = 0, rejected = 0;
        while (outer < prm->max_outer) {
            if (loop == 0) {
                if (ds_on) {
                    int failed = 0;
                    const double ps = oracle_bus_p(s, x, out_br, n_out, ds->slack_bus);
                    const int unstable = ds_check(ds, ps, mode, target, &retain_applied, prev, &failed);
                    res->slack_distribution_failed = failed;
                    if (failed && ds->reject_failure) { rejected = 1; break; }
                    if (unstable) { resolve = 1; break; }
                }
                loop = 1;
            } else if (loop == 1) {
                if (rl_on && rl_check(rl, prm, islanded, x, qinj, mode, target, lt, sc, pass)) {
                    resolve = 1;
                    break;
                }
                loop = 2;
            } else {
                if (!sweep_changed || !ds_on) break;
                sweep_changed = 0;
                loop = 0;
            }
        }
        if (rejected) { res->status = ORC_REJECTED; break; }
        if (resolve) {
            outer++;
            sweep_changed = 1;
            res->outer_switches = outer;
            continue;
        }
        res->status = ORC_CONVERGED;
        break;
    }
    free(F); free(b); free(dx); free(r); free(J); free(qinj); free(pass); free(prev); free(xprev);
}

int oracle_islands(const oracle_sys *s, int slack_bus, const int *out_br, int n_out, unsigned char *islanded)
{
    const int nb = s->n_bus;
    /* incident branch lists */
    int *deg = (int *)calloc((size_t)nb + 1, sizeof(int));
    for (int e = 0; e < s->n_branch; e++) { deg[s->bus1[e] + 1]++; deg[s->bus2[e] + 1]++; }
    for (int i = 0; i < nb; i++) deg[i + 1] += deg[i];
    int *inc = (int *)malloc(sizeof(int) * (size_t)(2 * s->n_branch + 1));
    int *fp = (int *)malloc(sizeof(int) * ((size_t)nb + 1));
    memcpy(fp, deg, sizeof(int) * (size_t)nb);
    for (int e = 0; e < s->n_branch; e++) { inc[fp[s->bus1[e]]++] = e; inc[fp[s->bus2[e]]++] = e; }
    for (int i = 0; i < nb; i++) islanded[i] = 1;
    int *queue = (int *)malloc(sizeof(int) * ((size_t)nb + 1));
    int head = 0, tail = 0;
    queue[tail++] = slack_bus;
    islanded[slack_bus] = 0;
    while (head < tail) {
        int u = queue[head++];
        for (int k = deg[u]; k < deg[u + 1]; k++) {
            int e = inc[k];
            if (n_out > 0 && is_out(e, out_br, n_out)) continue;
            int w = s->bus1[e] == u ? s->bus2[e] : s->bus1[e];
            if (islanded[w]) { islanded[w] = 0; queue[tail++] = w; }
        }
    }
    free(deg); free(inc); free(fp); free(queue);
    return nb - tail;
}

int oracle_n1(const oracle_sys *s, const oracle_nr_params *prm, const oracle_rl *rl, const oracle_ds *ds,
              int slack_bus, const double *x_base, const int *mode_base, const double *target_base,
              const signed char *limit_type_base, int n_ctg, int max_out, const int *ctg,
              int n_threads, double *x_out, int *status, int *iterations, int *total_iterations, double *norm,
              double *target_out, const double *x0_per)
{
    const int n = s->n;
    const int nrl = rl ? rl->n : 0;
    int used = 1;
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
    {
#pragma omp single
        used = omp_get_num_threads();
#else
    {
#endif
        int *mode = (int *)malloc(sizeof(int) * (size_t)n);
        double *target = (double *)malloc(sizeof(double) * (size_t)n);
        signed char *lt = (signed char *)malloc((size_t)nrl + 1);
        signed char *sc = (signed char *)malloc((size_t)nrl + 1);
        unsigned char *isl = (unsigned char *)malloc((size_t)s->n_bus + 1);
        int *outs = (int *)malloc(sizeof(int) * ((size_t)max_out + 1));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int c = 0; c < n_ctg; c++) {
            int n_out = 0;
            for (int k = 0; k < max_out; k++)
                if (ctg[(size_t)c * max_out + k] >= 0) outs[n_out++] = ctg[(size_t)c * max_out + k];
            double *x = x_out + (size_t)c * n;
            memcpy(x, x0_per ? x0_per + (size_t)c * n : x_base, sizeof(double) * (size_t)n);
            memcpy(mode, mode_base, sizeof(int) * (size_t)n);
            memcpy(target, target_base, sizeof(double) * (size_t)n);
            if (nrl) {
                memcpy(lt, limit_type_base, (size_t)nrl);
                memset(sc, 0, (size_t)nrl);
            }
            int ni = oracle_islands(s, slack_bus, outs, n_out, isl);
            if (ni > 0) {
                for (int i = 0; i < s->n_bus; i++)
                    if (isl[i]) {
                        mode[2 * i] = ORC_FIXED_VARIABLE;
                        target[2 * i] = 0.0;
                        mode[2 * i + 1] = ORC_FIXED_VARIABLE;
                        target[2 * i + 1] = 1.0;
                    }
            }
            oracle_result res;
            oracle_nr_params p = *prm;
            p.batch_of_one = 0;
            p.n_active = n - 2 * ni;
            oracle_nr_solve(s, &p, rl, ds, outs, n_out, ni > 0 ? isl : NULL, x, mode, target, lt, sc, &res);
            if (target_out) memcpy(target_out + (size_t)c * n, target, sizeof(double) * (size_t)n);
            status[c] = res.status;
            iterations[c] = res.iterations;
            total_iterations[c] = res.total_iterations;
            norm[c] = res.norm;
        }
        free(mode); free(target); free(lt); free(sc); free(isl); free(outs);
    }
    return used;
}

/* ------------------------------------------------------------------------------------------------ */
/* Limit violations: OLF's detection scan (sa/gpu/GpuLimitScanner.java:95-131,394-410).             */
/* ------------------------------------------------------------------------------------------------ */
int oracle_violations(const oracle_sys *s, const double *x, const int *out_br, int n_out,
                      const unsigned char *islanded, const oracle_limits *lim, oracle_violation *out, int max_out)
{
    int count = 0;
    double o[28];
    for (int b = 0; b < lim->n_branches; b++) {
        const int g0 = lim->group_start[b], g1 = lim->group_start[b + 1];
        if (g0 == g1) continue;
        const int e = lim->branch[b];
        if (n_out > 0 && is_out(e, out_br, n_out)) continue;
        const int a = s->bus1[e], c = s->bus2[e];
        if (islanded && (islanded[a] || islanded[c])) continue;
        oracle_closed_branch(s->p8 + 8 * e, x[2 * a + 1], x[2 * a], x[2 * c + 1], x[2 * c], o);
        /* outputs 0, 7, 14, 21: p1, q1, p2, q2 */
        for (int g = g0; g < g1; g++) {
            const int one = lim->group_side[g] == 1;
            const double p = one ? o[0] : o[14], q = one ? o[7] : o[21], v = one ? x[2 * a + 1] : x[2 * c + 1];
            double value;
            switch (lim->group_quantity[g]) {
                case 0: value = hypot(p, q) / v; break;
                case 1: value = p; break;
                default: value = sqrt(p * p + q * q); break;
            }
            if (lim->group_quantity[g] == 2 && isnan(value)) continue;
            const double compared = lim->group_quantity[g] == 1 ? fabs(value) : value;
            for (int l = lim->limit_start[g]; l < lim->limit_start[g + 1]; l++) {
                if (compared > lim->threshold[l]) {
                    if (count < max_out) {
                        oracle_violation v0 = {0, b, lim->group_side[g], lim->group_quantity[g], l, value,
                                               lim->threshold[l]};
                        out[count] = v0;
                    }
                    count++;
                    break;     /* the most serious one only */
                }
            }
        }
    }
    for (int i = 0; i < lim->n_buses; i++) {
        if (islanded && islanded[i]) continue;
        const double v = x[2 * i + 1];
        if (v > lim->bus_high[i]) {
            if (count < max_out) { oracle_violation v0 = {1, i, 0, 0, -1, v, lim->bus_high[i]}; out[count] = v0; }
            count++;
        }
        if (v < lim->bus_low[i]) {
            if (count < max_out) { oracle_violation v0 = {2, i, 0, 0, -1, v, lim->bus_low[i]}; out[count] = v0; }
            count++;
        }
    }
    for (int a = 0; a < lim->n_angle; a++) {
        const int f = lim->angle_from[a], t = lim->angle_to[a];
        if (islanded && (islanded[f] || islanded[t])) continue;
        const double d = x[2 * t] - x[2 * f];
        if (!isnan(lim->angle_high[a]) && d > lim->angle_high[a]) {
            if (count < max_out) { oracle_violation v0 = {3, a, 0, 0, -1, d, lim->angle_high[a]}; out[count] = v0; }
            count++;
        }
        if (!isnan(lim->angle_low[a]) && d < lim->angle_low[a]) {
            if (count < max_out) { oracle_violation v0 = {4, a, 0, 0, -1, d, lim->angle_low[a]}; out[count] = v0; }
            count++;
        }
    }
    return count;
}
