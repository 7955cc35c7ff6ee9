/* mfs_oracle.c -- CPU FP64 restatement of fastMBCUI's preconditioned MFS solve.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Built with -ffp-contract=off so that every sum of products is evaluated in
 * the order written, as the reference's host build (-O3, no -mfma) does.
 */
#include "mfs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------------
 * Squared distance in the reference's order: ((dx*dx + dy*dy) + dz*dz).
 * src/maxmin.h:370-374 (dist2_3d); the same order results from Eigen's
 * squaredNorm on a 3-vector and from nanoflann's L2_Adaptor::evalMetric
 * (src/nanoflann.hpp:383-405), so one helper serves ordering and pattern.
 * ------------------------------------------------------------------------- */
static inline double d2_3(const double *a, const double *b)
{
    const double dx = a[0] - b[0];
    const double dy = a[1] - b[1];
    const double dz = a[2] - b[2];
    return dx * dx + dy * dy + dz * dz;
}

/* ---------------------------------------------------------------------------
 * Max-heap over point ids, stored by position.  Same shape and rules as
 * src/maxmin.h:184-293: position k has children 2k+1, 2k+2; initial keys are
 * INFINITY for positions with fewer than two children and 10000 otherwise
 * (maxmin.h:277-291); sift-down prefers the right child unless left > right
 * strictly, and swaps only if parent < child strictly (maxmin.h:194-263).
 * ------------------------------------------------------------------------- */
typedef struct {
    int32_t n;
    double *key;    /* by position */
    int32_t *id;    /* point id at position */
    int32_t *pos;   /* position of point id */
} oheap;

static void oheap_swap(oheap *h, int32_t a, int32_t b)
{
    double tk = h->key[a]; h->key[a] = h->key[b]; h->key[b] = tk;
    int32_t ti = h->id[a]; h->id[a] = h->id[b]; h->id[b] = ti;
    h->pos[h->id[a]] = a;
    h->pos[h->id[b]] = b;
}

/* maxmin.h:266-271 update(): set the key, then only sift down. */
static void oheap_update(oheap *h, int32_t pid, double newkey)
{
    int32_t k = h->pos[pid];
    h->key[k] = newkey;
    for (;;) {
        const int32_t l = 2 * k + 1;
        if (l >= h->n) return;
        const int32_t r = l + 1;
        int32_t c;
        if (r >= h->n) {
            c = l;
        } else {
            c = (h->key[l] > h->key[r]) ? l : r;
        }
        if (h->key[k] < h->key[c]) {
            oheap_swap(h, k, c);
            k = c;
        } else {
            return;
        }
    }
}

/* Child lists per pivot (maxmin.h:295-336 ijlookup). */
typedef struct {
    int32_t *head;   /* head[p]..head[p+1] bounds the list of pivot p */
    int32_t *item;
    int64_t size, cap;
    int32_t cur;     /* index of the list being appended to */
} olists;

static int olists_push(olists *L, int32_t v)
{
    if (L->size == L->cap) {
        int64_t nc = L->cap * 2;
        int32_t *ni = (int32_t *)realloc(L->item, (size_t)nc * sizeof(int32_t));
        if (!ni) return MO_ERR_ALLOC;
        L->item = ni;
        L->cap = nc;
    }
    L->item[L->size++] = v;
    L->head[L->cur + 1] += 1;
    return MO_OK;
}

static void olists_open(olists *L)
{
    L->cur += 1;
    L->head[L->cur + 1] = L->head[L->cur];
}

/* maxmin.h:588-665 create_ordering_3d with no boundary tree (dist_bd = 1e16). */
static int ordering_3d(int32_t n, const double *x, int32_t root, int32_t *seq,
                       double *dist)
{
    int rc = MO_OK;
    oheap h;
    olists L;
    int32_t *parent = NULL;
    h.n = n;
    h.key = (double *)malloc((size_t)n * sizeof(double));
    h.id = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    h.pos = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    L.head = (int32_t *)calloc((size_t)n + 2, sizeof(int32_t));
    L.cap = n > 16 ? n : 16;
    L.item = (int32_t *)malloc((size_t)L.cap * sizeof(int32_t));
    L.size = 0;
    L.cur = 0;
    parent = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    if (!h.key || !h.id || !h.pos || !L.head || !L.item || !parent) {
        rc = MO_ERR_ALLOC;
        goto done;
    }
    for (int32_t k = 0; k < n; ++k) {
        const int two_children = (2 * (int64_t)k + 2) < n;
        h.key[k] = two_children ? 10000.0 : INFINITY;
        h.id[k] = k;
        h.pos[k] = k;
    }
    const double dist_bd = 1e16;

    /* Root pass, maxmin.h:637-646. */
    dist[0] = 0.0;
    for (int32_t k = 0; k < n; ++k) {
        if ((rc = olists_push(&L, k)) != MO_OK) goto done;
        const double dk = sqrt(d2_3(x + 3 * (int64_t)root, x + 3 * (int64_t)k));
        if (dk > dist[0]) dist[0] = dk;
        oheap_update(&h, k, dk < dist_bd ? dk : dist_bd);
        parent[k] = 0;
    }

    /* Pivot loop, maxmin.h:651-657 with determineChildren_3d, :440-464. */
    for (int32_t k = 1; k < n; ++k) {
        const int32_t pid = h.id[0];
        const double pdist = h.key[0];
        dist[k] = pdist;
        seq[k] = pid;
        const int32_t kmin = L.head[parent[pid]];
        const int32_t kmax = L.head[parent[pid] + 1];
        olists_open(&L);
        for (int32_t t = kmin; t < kmax; ++t) {
            const int32_t c = L.item[t];
            const double t2 = d2_3(x + 3 * (int64_t)pid, x + 3 * (int64_t)c);
            if (t2 < pdist * pdist) {
                double jd = h.key[h.pos[c]];
                if (t2 < jd * jd) {
                    const double s = sqrt(t2);
                    const double nk = s < dist_bd ? s : dist_bd;
                    oheap_update(&h, c, nk);
                    jd = nk;
                }
                if ((rc = olists_push(&L, c)) != MO_OK) goto done;
                if (sqrt(t2) + jd < pdist) parent[c] = k;
            }
        }
    }
done:
    free(h.key); free(h.id); free(h.pos);
    free(L.head); free(L.item); free(parent);
    return rc;
}

int mo_fast_fps(int32_t n, const double *xyz, int32_t *P, int32_t *Pinv,
                double *scale, int32_t *seq_out, double *dist_out)
{
    if (n <= 0 || !xyz || !P || !Pinv || !scale) return MO_ERR_ARG;
    /* preprocess.h:785, 815-823: corner = column-wise min; root = closest to
     * the corner, strict < from 1e10 so the lowest index wins ties. */
    double corner[3] = {xyz[0], xyz[1], xyz[2]};
    for (int32_t i = 1; i < n; ++i)
        for (int d = 0; d < 3; ++d)
            if (xyz[3 * (int64_t)i + d] < corner[d]) corner[d] = xyz[3 * (int64_t)i + d];
    int32_t root = -1;
    double best = 1e10;
    for (int32_t i = 0; i < n; ++i) {
        const double d = d2_3(xyz + 3 * (int64_t)i, corner);
        if (d < best) { best = d; root = i; }
    }
    if (root < 0) return MO_ERR_ARG;

    int32_t *seq = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    double *dist = (double *)malloc((size_t)n * sizeof(double));
    if (!seq || !dist) { free(seq); free(dist); return MO_ERR_ALLOC; }
    seq[0] = root;
    int rc = ordering_3d(n, xyz, root, seq, dist);
    if (rc == MO_OK) {
        /* preprocess.h:822-823: the sequence must be a permutation. */
        for (int32_t i = 0; i < n; ++i) P[i] = -1;
        for (int32_t k = 0; k < n; ++k) {
            if (seq[k] < 0 || seq[k] >= n || P[seq[k]] != -1) { rc = MO_ERR_DUPLICATE; break; }
            P[seq[k]] = n - 1 - k;   /* preprocess.h:825-828 */
        }
    }
    if (rc == MO_OK) {
        for (int32_t k = 0; k < n; ++k) {
            Pinv[n - 1 - k] = seq[k];
            scale[n - 1 - k] = dist[k];  /* preprocess.h:830 reverseInPlace */
        }
        /* preprocess.h:74-77 debug(): ascending scale; a zero scale means a
         * point coincides with a coarser one. */
        for (int32_t i = 0; i + 1 < n; ++i)
            if (!(scale[i] <= scale[i + 1])) { rc = MO_ERR_DUPLICATE; break; }
        if (rc == MO_OK && n > 1 && !(scale[0] > 0.0)) rc = MO_ERR_DUPLICATE;
    }
    if (seq_out) memcpy(seq_out, seq, (size_t)n * sizeof(int32_t));
    if (dist_out) memcpy(dist_out, dist, (size_t)n * sizeof(double));
    free(seq);
    free(dist);
    return rc;
}

int mo_maximin_brute(int32_t n, const double *xyz, int32_t root, int32_t *seq,
                     double *dist)
{
    if (n <= 0 || root < 0 || root >= n) return MO_ERR_ARG;
    double *d = (double *)malloc((size_t)n * sizeof(double));
    unsigned char *used = (unsigned char *)calloc((size_t)n, 1);
    if (!d || !used) { free(d); free(used); return MO_ERR_ALLOC; }
    dist[0] = 0.0;
    for (int32_t i = 0; i < n; ++i) {
        d[i] = sqrt(d2_3(xyz + 3 * (int64_t)root, xyz + 3 * (int64_t)i));
        if (d[i] > dist[0]) dist[0] = d[i];
    }
    seq[0] = root;
    used[root] = 1;
    for (int32_t k = 1; k < n; ++k) {
        int32_t best = -1;
        double bd = -1.0;
        for (int32_t i = 0; i < n; ++i)
            if (!used[i] && d[i] > bd) { bd = d[i]; best = i; }
        seq[k] = best;
        dist[k] = bd;
        used[best] = 1;
        for (int32_t i = 0; i < n; ++i) {
            if (used[i]) continue;
            const double s = sqrt(d2_3(xyz + 3 * (int64_t)best, xyz + 3 * (int64_t)i));
            if (s < d[i]) d[i] = s;
        }
    }
    free(d);
    free(used);
    return MO_OK;
}

/* ---------------------------------------------------------------------------
 * Pattern: preprocess.h:235-316.  Both of the reference's branches (kd-tree
 * radius search for early columns, linear scan for late ones) keep exactly
 * the rows that satisfy the safeguard predicate at :270 / :292, which is
 * what is evaluated here for every i >= j.
 * ------------------------------------------------------------------------- */
int mo_pattern(int32_t n, const double *xyz_new, const double *scale, double rho,
               int32_t *ptr, int32_t *ind)
{
    if (n <= 0 || !xyz_new || !scale || !ptr) return MO_ERR_ARG;
    if (!ind) {
        ptr[0] = 0;
        for (int32_t j = 0; j < n; ++j) {
            const double lj = rho * scale[j];
            int32_t cnt = 0;
            for (int32_t i = j; i < n; ++i) {
                const double li = rho * scale[i];
                const double lim = (li * li < lj * lj) ? li * li : lj * lj;
                if (d2_3(xyz_new + 3 * (int64_t)i, xyz_new + 3 * (int64_t)j) < lim) ++cnt;
            }
            ptr[j + 1] = ptr[j] + cnt;
        }
        return MO_OK;
    }
    for (int32_t j = 0; j < n; ++j) {
        const double lj = rho * scale[j];
        int32_t w = ptr[j];
        for (int32_t i = j; i < n; ++i) {
            const double li = rho * scale[i];
            const double lim = (li * li < lj * lj) ? li * li : lj * lj;
            if (d2_3(xyz_new + 3 * (int64_t)i, xyz_new + 3 * (int64_t)j) < lim) ind[w++] = i;
        }
        if (w != ptr[j + 1]) return MO_ERR_ARG;
    }
    return MO_OK;
}

/* preprocess.h:541-602 aggregate() with group {0, n}, dim 1. */
int mo_aggregate(int32_t n, const int32_t *ptr, const int32_t *ind,
                 const double *scale, double lambda, int32_t max_super,
                 int32_t *sup_ptr, int32_t *sup_ind, int32_t *sup_parent,
                 int32_t *n_super)
{
    if (n <= 0) return MO_ERR_ARG;
    unsigned char *vis = (unsigned char *)calloc((size_t)n, 1);
    if (!vis) return MO_ERR_ALLOC;
    int32_t ns = 0, w = 0;
    sup_ptr[0] = 0;
    for (int32_t j = 0; j < n; ++j) {
        if (vis[j]) continue;
        const int32_t start = w;
        for (int32_t e = ptr[j]; e < ptr[j + 1]; ++e) {
            const int32_t i = ind[e];
            if (!vis[i] && scale[i] <= lambda * scale[j]) {
                vis[i] = 1;
                sup_ind[w++] = i;
                sup_parent[i] = ns;
                if (w - start >= max_super) break;
            }
        }
        ++ns;
        sup_ptr[ns] = w;
    }
    free(vis);
    *n_super = ns;
    /* preprocess.h:593: every point belongs to exactly one supernode. */
    return (w == n) ? MO_OK : MO_ERR_DUPLICATE;
}

/* preprocess.h:604-671: idx_set[s] = sorted union of the rows of the columns
 * of s; column j keeps the part of idx_set[sup(j)] that is >= j. */
int mo_super_pattern(int32_t n, const int32_t *ptr, const int32_t *ind,
                     const int32_t *sup_parent, int32_t n_super,
                     int32_t *sptr, int32_t *sind)
{
    if (n <= 0 || n_super <= 0) return MO_ERR_ARG;
    /* Row-major walk (the reference walks the transpose column by column). */
    int32_t *rcnt = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    int64_t nnz = ptr[n];
    int32_t *rcol = (int32_t *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
    int32_t *last = (int32_t *)malloc((size_t)n_super * sizeof(int32_t));
    int32_t *setcnt = (int32_t *)calloc((size_t)n_super + 1, sizeof(int32_t));
    int32_t *setval = NULL;
    int rc = MO_OK;
    if (!rcnt || !rcol || !last || !setcnt) { rc = MO_ERR_ALLOC; goto done; }
    for (int64_t e = 0; e < nnz; ++e) rcnt[ind[e] + 1]++;
    for (int32_t i = 0; i < n; ++i) rcnt[i + 1] += rcnt[i];
    {
        int32_t *fill = (int32_t *)malloc((size_t)n * sizeof(int32_t));
        if (!fill) { rc = MO_ERR_ALLOC; goto done; }
        memcpy(fill, rcnt, (size_t)n * sizeof(int32_t));
        for (int32_t j = 0; j < n; ++j)
            for (int32_t e = ptr[j]; e < ptr[j + 1]; ++e) rcol[fill[ind[e]]++] = j;
        free(fill);
    }
    /* Pass 1: set sizes.  Rows arrive in ascending order, duplicates are
     * adjacent, so "unique" is a comparison with the last row pushed. */
    for (int32_t s = 0; s < n_super; ++s) last[s] = -1;
    for (int32_t i = 0; i < n; ++i)
        for (int32_t e = rcnt[i]; e < rcnt[i + 1]; ++e) {
            const int32_t s = sup_parent[rcol[e]];
            if (last[s] != i) { last[s] = i; setcnt[s + 1]++; }
        }
    for (int32_t s = 0; s < n_super; ++s) setcnt[s + 1] += setcnt[s];
    setval = (int32_t *)malloc((size_t)(setcnt[n_super] > 0 ? setcnt[n_super] : 1) * sizeof(int32_t));
    if (!setval) { rc = MO_ERR_ALLOC; goto done; }
    {
        int32_t *fill = (int32_t *)malloc((size_t)n_super * sizeof(int32_t));
        if (!fill) { rc = MO_ERR_ALLOC; goto done; }
        memcpy(fill, setcnt, (size_t)n_super * sizeof(int32_t));
        for (int32_t s = 0; s < n_super; ++s) last[s] = -1;
        for (int32_t i = 0; i < n; ++i)
            for (int32_t e = rcnt[i]; e < rcnt[i + 1]; ++e) {
                const int32_t s = sup_parent[rcol[e]];
                if (last[s] != i) { last[s] = i; setval[fill[s]++] = i; }
            }
        free(fill);
    }
    /* Column j: suffix of its supernode's set starting at the first row >= j. */
    sptr[0] = 0;
    for (int32_t j = 0; j < n; ++j) {
        const int32_t s = sup_parent[j];
        int32_t lo = setcnt[s], hi = setcnt[s + 1];
        while (lo < hi) {   /* lower_bound */
            const int32_t mid = lo + (hi - lo) / 2;
            if (setval[mid] < j) lo = mid + 1; else hi = mid;
        }
        const int32_t len = setcnt[s + 1] - lo;
        sptr[j + 1] = sptr[j] + len;
        if (sind) memcpy(sind + sptr[j], setval + lo, (size_t)len * sizeof(int32_t));
    }
done:
    free(rcnt); free(rcol); free(last); free(setcnt); free(setval);
    return rc;
}

/* ---------------------------------------------------------------------------
 * Factor.  For supernode s with local rows I (= rows of its first column),
 * TH[rev_p + rev_q*ns] = <Mp[:,I_p], Mp[:,I_q]>, rev = ns-1-local
 * (kl_chol.cu:208-221).  Upper Cholesky TH = U^T U, column-major, left-looking
 * (kl_chol.cu:16-36).  Column j with nb rows: x = U[0:nb,0:nb]^{-1} e_{nb-1},
 * reversed into ascending row order (kl_chol.cu:38-53, 381-386).
 * ------------------------------------------------------------------------- */
int mo_factor(int64_t m, int32_t n, const double *Mp,
              const int32_t *sup_ptr, const int32_t *sup_ind, int32_t n_super,
              const int32_t *sptr, const int32_t *sind, double *val,
              int32_t *bad_super)
{
    if (m <= 0 || n <= 0 || n_super <= 0) return MO_ERR_ARG;
    if (bad_super) *bad_super = -1;
    int rc = MO_OK;
    int32_t first_bad = -1;

    #pragma omp parallel for schedule(dynamic, 1)
    for (int32_t s = 0; s < n_super; ++s) {
        const int32_t first = sup_ind[sup_ptr[s]];
        const int32_t ns = sptr[first + 1] - sptr[first];
        const int32_t *I = sind + sptr[first];
        double *A = (double *)malloc((size_t)ns * (size_t)ns * sizeof(double));
        double *b = (double *)malloc((size_t)ns * sizeof(double));
        if (!A || !b) {
            free(A); free(b);
            #pragma omp critical
            rc = MO_ERR_ALLOC;
            continue;
        }
        for (int32_t p = 0; p < ns; ++p) {
            const double *cp = Mp + (int64_t)I[p] * m;
            for (int32_t q = p; q < ns; ++q) {
                const double *cq = Mp + (int64_t)I[q] * m;
                double g = 0.0;
                for (int64_t f = 0; f < m; ++f) g += cp[f] * cq[f];
                const int32_t rp = ns - 1 - p, rq = ns - 1 - q;
                A[rp + (int64_t)rq * ns] = g;
                A[rq + (int64_t)rp * ns] = g;
            }
        }
        /* A = U^T U in place (upper triangle holds U). */
        int ok = 1;
        for (int32_t j = 0; j < ns && ok; ++j) {
            int32_t i;
            for (i = 0; i < j; ++i) {
                double dot = 0.0;
                for (int32_t k = 0; k < i; ++k) dot += A[k + (int64_t)j * ns] * A[k + (int64_t)i * ns];
                A[i + (int64_t)j * ns] = (A[i + (int64_t)j * ns] - dot) / A[i + (int64_t)i * ns];
            }
            double dot = 0.0;
            for (int32_t k = 0; k < j; ++k) dot += A[k + (int64_t)j * ns] * A[k + (int64_t)j * ns];
            const double piv = A[j + (int64_t)j * ns] - dot;
            if (!(piv > 0.0)) { ok = 0; break; }
            A[j + (int64_t)j * ns] = sqrt(piv);
        }
        if (!ok) {
            #pragma omp critical
            {
                rc = MO_ERR_NOT_SPD;
                if (first_bad < 0 || s < first_bad) first_bad = s;
            }
            free(A); free(b);
            continue;
        }
        /* Member columns of s. */
        for (int32_t t = sup_ptr[s]; t < sup_ptr[s + 1]; ++t) {
            const int32_t j = sup_ind[t];
            const int32_t nb = sptr[j + 1] - sptr[j];
            const int32_t end = nb - 1;
            for (int32_t i = 0; i < nb; ++i) b[i] = 0.0;
            b[end] = 1.0;
            for (int32_t i = 0; i < nb; ++i) {
                for (int32_t k = 0; k < i; ++k)
                    b[end - i] -= b[end - k] * A[(end - i) + (int64_t)(end - k) * ns];
                b[end - i] /= A[(end - i) + (int64_t)(end - i) * ns];
            }
            double *out = val + sptr[j];
            for (int32_t i = 0; i < nb; ++i) out[i] = b[end - i];
        }
        free(A);
        free(b);
    }
    if (bad_super) *bad_super = first_bad;
    return rc;
}

void mo_spmv_lt(int32_t n, const int32_t *sptr, const int32_t *sind,
                const double *val, const double *r, double *t)
{
    for (int32_t j = 0; j < n; ++j) {
        double s = 0.0;
        for (int32_t e = sptr[j]; e < sptr[j + 1]; ++e) s += val[e] * r[sind[e]];
        t[j] = s;
    }
}

void mo_spmv_l(int32_t n, const int32_t *sptr, const int32_t *sind,
               const double *val, const double *t, double *z)
{
    for (int32_t i = 0; i < n; ++i) z[i] = 0.0;
    for (int32_t j = 0; j < n; ++j)
        for (int32_t e = sptr[j]; e < sptr[j + 1]; ++e) z[sind[e]] += val[e] * t[j];
}

static double dotn(int32_t n, const double *a, const double *b)
{
    double s = 0.0;
    for (int32_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* q = Mp^T (Mp p): two GEMV passes as at kl_chol.cu:992-1002. */
static void normal_matvec(int64_t m, int32_t n, const double *Mp, const double *p,
                          double *y, double *q)
{
    /* y = Mp p: row blocks in parallel, columns in order within each block. */
    const int64_t blk = 256;
    #pragma omp parallel for schedule(static)
    for (int64_t f0 = 0; f0 < m; f0 += blk) {
        const int64_t f1 = (f0 + blk < m) ? f0 + blk : m;
        for (int64_t f = f0; f < f1; ++f) y[f] = 0.0;
        for (int32_t c = 0; c < n; ++c) {
            const double pc = p[c];
            const double *col = Mp + (int64_t)c * m;
            for (int64_t f = f0; f < f1; ++f) y[f] += col[f] * pc;
        }
    }
    #pragma omp parallel for schedule(static)
    for (int32_t c = 0; c < n; ++c) {
        const double *col = Mp + (int64_t)c * m;
        double s = 0.0;
        for (int64_t f = 0; f < m; ++f) s += col[f] * y[f];
        q[c] = s;
    }
}

int mo_pcg(int64_t m, int32_t n, const double *Mp, const double *rhs,
           int rhs_space, const int32_t *sptr, const int32_t *sind,
           const double *val, int preconditioned, int32_t maxit, double tol,
           double *x, int32_t *iters, double *relres)
{
    if (m <= 0 || n <= 0) return MO_ERR_ARG;
    double *r = (double *)malloc((size_t)n * sizeof(double));
    double *p = (double *)malloc((size_t)n * sizeof(double));
    double *q = (double *)malloc((size_t)n * sizeof(double));
    double *z = (double *)malloc((size_t)n * sizeof(double));
    double *t = (double *)malloc((size_t)n * sizeof(double));
    double *y = (double *)malloc((size_t)m * sizeof(double));
    if (!r || !p || !q || !z || !t || !y) {
        free(r); free(p); free(q); free(z); free(t); free(y);
        return MO_ERR_ALLOC;
    }
    /* r = M^T b (kl_chol.cu:928-939) or r = g (kl_chol.cu:1118). */
    if (rhs_space == 0) {
        #pragma omp parallel for schedule(static)
        for (int32_t c = 0; c < n; ++c) {
            const double *col = Mp + (int64_t)c * m;
            double s = 0.0;
            for (int64_t f = 0; f < m; ++f) s += col[f] * rhs[f];
            r[c] = s;
        }
    } else {
        memcpy(r, rhs, (size_t)n * sizeof(double));
    }
    for (int32_t i = 0; i < n; ++i) x[i] = 0.0;
    const double rhs2 = dotn(n, r, r);
    *iters = 0;
    *relres = 0.0;
    if (rhs2 < 1e-16) {   /* kl_chol.cu:950-953 */
        free(r); free(p); free(q); free(z); free(t); free(y);
        return MO_OK;
    }
    const double threshold = tol * tol * rhs2;
    if (preconditioned) {
        mo_spmv_lt(n, sptr, sind, val, r, t);
        mo_spmv_l(n, sptr, sind, val, t, p);
    } else {
        memcpy(p, r, (size_t)n * sizeof(double));
    }
    double absnew = dotn(n, r, p);
    int32_t i = 0;
    double resd2 = 0.0;
    while (i < maxit) {
        ++i;
        normal_matvec(m, n, Mp, p, y, q);
        const double alpha = absnew / dotn(n, p, q);
        for (int32_t c = 0; c < n; ++c) x[c] += alpha * p[c];
        for (int32_t c = 0; c < n; ++c) r[c] += (-alpha) * q[c];
        resd2 = dotn(n, r, r);
        if (resd2 < threshold) break;
        if (preconditioned) {
            mo_spmv_lt(n, sptr, sind, val, r, t);
            mo_spmv_l(n, sptr, sind, val, t, z);
        } else {
            memcpy(z, r, (size_t)n * sizeof(double));
        }
        const double absold = absnew;
        absnew = dotn(n, r, z);
        const double beta = absnew / absold;
        for (int32_t c = 0; c < n; ++c) p[c] = beta * p[c] + z[c];
    }
    *iters = i;
    *relres = sqrt(resd2 / rhs2);
    free(r); free(p); free(q); free(z); free(t); free(y);
    return MO_OK;
}
