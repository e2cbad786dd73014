/*
 * ptg_oracle.c -- CPU restatement of pytreegrav's tree-gravity hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is compared against and the
 * CPU baseline timed beside it.  Nothing in the product path (pytreegrav_amd/) may link, load or
 * call it; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Every function restates one function of the reference (file:line cited at each).  The
 * arithmetic order of the build and moment passes is kept literally, and the file is compiled
 * with -ffp-contract=off and without -ffast-math, so Morton keys, topology and FP64 moments are
 * reproduced bit-for-bit; the walk and brute-force sums follow the reference formulas (the
 * reference compiles those with fastmath, so they agree with it to ulps, not bits).
 *
 * Parity pinning (what checks each part, run in tests/test_oracle_*.py):
 *   kernels        closed forms at q=0, TidalKernel(1,1)=3, continuity, T=-(1/r)dK/dr
 *   keys           bit-interleave against a per-bit loop; cell-membership invariants
 *   radix argsort  numpy.argsort(kind="stable")
 *   build          brute-force enumeration of dyadic cells (>=2 points), link-walk invariants,
 *                  re-key separation and coincident-point buckets (reference test inputs)
 *   moments        per-node COM / mass / hmax recomputed directly; root quadrupole analytic
 *   walk           theta->0 equals brute force; analytic Plummer; permutation invariance;
 *                  published accuracy table (README.md:120-125) on examples/error_benchmark.py's
 *                  Plummer sampler, seed 42
 *   brute force    two-body closed forms, momentum conservation, trace = -4 pi G rho
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MORTON_BITS 21
#define MORTON_SCALE 2097152.0 /* 2**21 */
#define MORTON_MAXQ ((1LL << MORTON_BITS) - 1)
#define MORTON_MAX_ROUNDS 3

/* octree.py:31-42 */
static const double octant_offsets[8][3] = {
    {-0.25, -0.25, -0.25}, {0.25, -0.25, -0.25}, {-0.25, 0.25, -0.25}, {0.25, 0.25, -0.25},
    {-0.25, -0.25, 0.25},  {0.25, -0.25, 0.25},  {-0.25, 0.25, 0.25},  {0.25, 0.25, 0.25},
};

/* ------------------------------------------------------------------------------------------ */
/* Softening kernels: kernel.py:4-75                                                          */
/* ------------------------------------------------------------------------------------------ */

/* kernel.py:4-25 */
double ptgo_force_kernel(double r, double h)
{
    if (r > h) return 1.0 / (r * r * r);
    double hinv = 1.0 / h;
    double q = r * hinv;
    if (q <= 0.5) return (10.666666666666666666 + q * q * (-38.4 + 32.0 * q)) * hinv * hinv * hinv;
    return (21.333333333333 - 48.0 * q + 38.4 * q * q - 10.666666666667 * q * q * q -
            0.066666666667 / (q * q * q)) *
           hinv * hinv * hinv;
}

/* kernel.py:28-50 */
double ptgo_tidal_kernel(double r, double h)
{
    if (r > h) {
        double r2 = r * r;
        return 3.0 / (r2 * r2 * r);
    }
    double hinv = 1.0 / h;
    double hinv2 = hinv * hinv;
    double hinv5 = hinv2 * hinv2 * hinv;
    double q = r * hinv;
    if (q <= 0.5) return (76.8 - 96.0 * q) * hinv5;
    double qinv = 1.0 / q;
    double qinv3 = qinv * qinv * qinv;
    return (48.0 * qinv - 76.8 + 32.0 * q - 0.2 * qinv3 * qinv * qinv) * hinv5;
}

/* kernel.py:53-75 */
double ptgo_potential_kernel(double r, double h)
{
    if (h == 0.0) return -1.0 / r;
    double hinv = 1.0 / h;
    double q = r * hinv;
    if (q <= 0.5) return (-2.8 + q * q * (5.33333333333333333 + q * q * (6.4 * q - 9.6))) * hinv;
    if (q <= 1)
        return (-3.2 + 0.066666666666666666666 / q +
                q * q * (10.666666666666666666666 + q * (-16.0 + q * (9.6 - 2.1333333333333333333333 * q)))) *
               hinv;
    return -1.0 / r;
}

/* ------------------------------------------------------------------------------------------ */
/* Keys and sort: octree.py:54-132, 539-574                                                  */
/* ------------------------------------------------------------------------------------------ */

/* octree.py:54-66 */
int64_t ptgo_spread_bits(int64_t x)
{
    x &= 0x1FFFFF;
    x = (x | (x << 32)) & 0x1F00000000FFFFLL;
    x = (x | (x << 16)) & 0x1F0000FF0000FFLL;
    x = (x | (x << 8)) & 0x100F00F00F00F00FLL;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3LL;
    x = (x | (x << 2)) & 0x1249249249249249LL;
    return x;
}

/* octree.py:69-98.  Python's int() truncates toward zero, as does the C cast. */
void ptgo_morton_keys(const double *pts, int64_t n, const double *center, double size, int64_t *keys)
{
    double inv = size > 0 ? 1.0 / size : 0.0;
    double cmin0 = center[0] - 0.5 * size;
    double cmin1 = center[1] - 0.5 * size;
    double cmin2 = center[2] - 0.5 * size;
    for (int64_t i = 0; i < n; i++) {
        int64_t qx = (int64_t)((pts[3 * i + 0] - cmin0) * inv * MORTON_SCALE);
        int64_t qy = (int64_t)((pts[3 * i + 1] - cmin1) * inv * MORTON_SCALE);
        int64_t qz = (int64_t)((pts[3 * i + 2] - cmin2) * inv * MORTON_SCALE);
        if (qx < 0) qx = 0; else if (qx > MORTON_MAXQ) qx = MORTON_MAXQ;
        if (qy < 0) qy = 0; else if (qy > MORTON_MAXQ) qy = MORTON_MAXQ;
        if (qz < 0) qz = 0; else if (qz > MORTON_MAXQ) qz = MORTON_MAXQ;
        keys[i] = ptgo_spread_bits(qx) | (ptgo_spread_bits(qy) << 1) | (ptgo_spread_bits(qz) << 2);
    }
}

/* octree.py:101-132: stable LSD radix argsort, 11-bit digits, keys carried alongside. */
void ptgo_radix_argsort(const int64_t *keys, int64_t n, int64_t *order)
{
    int64_t *idx = order;
    int64_t *idx_t = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    int64_t *k_cur = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    int64_t *k_t = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    int64_t count[2049];
    for (int64_t i = 0; i < n; i++) {
        idx[i] = i;
        k_cur[i] = keys[i];
    }
    const int RADIX_BITS = 11;
    const int64_t mask = (1 << RADIX_BITS) - 1;
    int passes = 0;
    for (int shift = 0; shift < 63; shift += RADIX_BITS, passes++) {
        memset(count, 0, sizeof(count));
        for (int64_t i = 0; i < n; i++) count[((k_cur[i] >> shift) & mask) + 1] += 1;
        for (int b = 0; b < 2048; b++) count[b + 1] += count[b];
        for (int64_t i = 0; i < n; i++) {
            int64_t d = (k_cur[i] >> shift) & mask;
            int64_t p = count[d]++;
            idx_t[p] = idx[i];
            k_t[p] = k_cur[i];
        }
        int64_t *s = idx; idx = idx_t; idx_t = s;
        s = k_cur; k_cur = k_t; k_t = s;
    }
    /* six passes: the result sits in the buffer `order` started in after an even swap count */
    if (idx != order) {
        memcpy(order, idx, sizeof(int64_t) * n);
        idx_t = idx;
    }
    free(idx_t);
    free(k_cur);
    free(k_t);
    (void)passes;
}

/* octree.py:539-574 */
void ptgo_bounding_cube(const double *p, int64_t n, double *center, double *size)
{
    double lo0 = p[0], hi0 = p[0], lo1 = p[1], hi1 = p[1], lo2 = p[2], hi2 = p[2];
    for (int64_t i = 1; i < n; i++) {
        double x = p[3 * i], y = p[3 * i + 1], z = p[3 * i + 2];
        if (x < lo0) lo0 = x; else if (x > hi0) hi0 = x;
        if (y < lo1) lo1 = y; else if (y > hi1) hi1 = y;
        if (z < lo2) lo2 = z; else if (z > hi2) hi2 = z;
    }
    center[0] = 0.5 * (hi0 + lo0);
    center[1] = 0.5 * (hi1 + lo1);
    center[2] = 0.5 * (hi2 + lo2);
    double s = hi0 - lo0;
    if (hi1 - lo1 > s) s = hi1 - lo1;
    if (hi2 - lo2 > s) s = hi2 - lo2;
    *size = s;
}

/* grouped_treewalk.py:269-275 */
void ptgo_morton_order(const double *pts, int64_t n, int64_t *order)
{
    if (n == 0) return;
    double c[3], s;
    ptgo_bounding_cube(pts, n, c, &s);
    int64_t *keys = (int64_t *)malloc(sizeof(int64_t) * n);
    ptgo_morton_keys(pts, n, c, s, keys);
    ptgo_radix_argsort(keys, n, order);
    free(keys);
}

/* ------------------------------------------------------------------------------------------ */
/* Tree: octree.py:664-723 (Octree), 817-1206 (BuildTreeRadix, RekeySlice, BuildBucket,      */
/* Initialize), 199-292 (RollupMoments, RollupQuadrupoles), 1280-1298 (growth)              */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    int64_t NumParticles;
    int64_t NumNodes; /* allocated element slots (particles + node capacity) */
    int64_t node_hi;  /* first unused node index after the build */
    int HasQuads, HasCoincidentPoints;
    double *Coordinates, *Masses, *Softenings, *Sizes, *Deltas, *Quadrupoles;
    int64_t *FirstSubnode, *NextBranch, *TreewalkIndices;
    /* build-local moment accumulators, indexed by node - NumParticles (octree.py:873-880) */
    double *acc;
    int64_t *parent, *pparent;
} ptgo_tree;

static void *xrealloc_tail(void *p, size_t elem, int64_t old_n, int64_t new_n, int fill_neg1)
{
    p = realloc(p, elem * (size_t)new_n);
    if (fill_neg1)
        memset((char *)p + elem * old_n, 0xff, elem * (size_t)(new_n - old_n));
    else
        memset((char *)p + elem * old_n, 0, elem * (size_t)(new_n - old_n));
    return p;
}

/* octree.py:1280-1298 (+ _grow_moment_arrays :142-152) */
static void increase_tree_size(ptgo_tree *t)
{
    int64_t old = t->NumNodes;
    int64_t inc = (int64_t)(old * 1.2 + 1) - old;
    if (inc < 1) inc = 1;
    int64_t nn = old + inc;
    t->Sizes = xrealloc_tail(t->Sizes, 8, old, nn, 0);
    t->Deltas = xrealloc_tail(t->Deltas, 8, old, nn, 0);
    t->Masses = xrealloc_tail(t->Masses, 8, old, nn, 0);
    t->Softenings = xrealloc_tail(t->Softenings, 8, old, nn, 0);
    t->NextBranch = xrealloc_tail(t->NextBranch, 8, old, nn, 1);
    t->FirstSubnode = xrealloc_tail(t->FirstSubnode, 8, old, nn, 1);
    t->Coordinates = xrealloc_tail(t->Coordinates, 24, old, nn, 0);
    if (t->HasQuads) t->Quadrupoles = xrealloc_tail(t->Quadrupoles, 72, old, nn, 0);
    int64_t np = t->NumParticles;
    t->acc = xrealloc_tail(t->acc, 24, old - np, nn - np, 0);
    t->parent = xrealloc_tail(t->parent, 8, old - np, nn - np, 1);
    t->NumNodes = nn;
}

/* octree.py:1188-1206 */
static void initialize(ptgo_tree *t, int64_t n)
{
    t->NumParticles = n;
    t->NumNodes = (int64_t)(1.5 * n + 1);
    int64_t nn = t->NumNodes;
    t->Sizes = calloc(nn, 8);
    t->Deltas = calloc(nn, 8);
    t->Masses = calloc(nn, 8);
    t->Quadrupoles = t->HasQuads ? calloc(nn, 72) : NULL;
    t->Softenings = calloc(nn, 8);
    t->Coordinates = calloc(nn, 24);
    t->NextBranch = malloc(nn * 8);
    t->FirstSubnode = malloc(nn * 8);
    memset(t->NextBranch, 0xff, nn * 8);
    memset(t->FirstSubnode, 0xff, nn * 8);
    t->acc = calloc(nn - n, 24);
    t->parent = malloc((nn - n) * 8);
    memset(t->parent, 0xff, (nn - n) * 8);
}

/* octree.py:170-196 (the linear-scan branch; the binary-search branch finds the same runs) */
static int octant_runs(const int64_t *keys, int64_t lo, int64_t hi, int shift, int64_t *bs, int64_t *be,
                       int64_t *bd)
{
    int nb = 0;
    int64_t start = lo;
    while (start < hi) {
        int64_t d = (keys[start] >> shift) & 7;
        int64_t end = start + 1;
        while (end < hi && ((keys[end] >> shift) & 7) == d) end++;
        bs[nb] = start;
        be[nb] = end;
        bd[nb] = d;
        nb++;
        start = end;
    }
    return nb;
}

/* octree.py:1093-1126 */
static void rekey_slice(ptgo_tree *t, int64_t *keys_sorted, int64_t node, int64_t lo, int64_t hi)
{
    int64_t g = hi - lo;
    int64_t *subkeys = malloc(8 * g), *perm = malloc(8 * g);
    double center[3] = {t->Coordinates[3 * node], t->Coordinates[3 * node + 1], t->Coordinates[3 * node + 2]};
    ptgo_morton_keys(t->Coordinates + 3 * lo, g, center, t->Sizes[node], subkeys);
    ptgo_radix_argsort(subkeys, g, perm);
    double *nc = malloc(24 * g), *nm = malloc(8 * g), *ns = malloc(8 * g);
    int64_t *ni = malloc(8 * g), *nk = malloc(8 * g);
    for (int64_t a = 0; a < g; a++) {
        int64_t p = perm[a];
        for (int k = 0; k < 3; k++) nc[3 * a + k] = t->Coordinates[3 * (lo + p) + k];
        nm[a] = t->Masses[lo + p];
        ns[a] = t->Softenings[lo + p];
        ni[a] = t->TreewalkIndices[lo + p];
        nk[a] = subkeys[p];
    }
    for (int64_t a = 0; a < g; a++) {
        for (int k = 0; k < 3; k++) t->Coordinates[3 * (lo + a) + k] = nc[3 * a + k];
        t->Masses[lo + a] = nm[a];
        t->Softenings[lo + a] = ns[a];
        t->TreewalkIndices[lo + a] = ni[a];
        keys_sorted[lo + a] = nk[a];
    }
    free(subkeys); free(perm); free(nc); free(nm); free(ns); free(ni); free(nk);
}

/* octree.py:1128-1175 */
static int64_t build_bucket(ptgo_tree *t, int64_t node, int64_t lo, int64_t hi, int64_t new_node_idx)
{
    int64_t np = t->NumParticles;
    t->HasCoincidentPoints = 1;
    int64_t count = hi - lo, cur = node, last = -1;
    int slot = 0;
    for (int64_t jj = 0; jj < count; jj++) {
        int64_t remaining = count - 1 - jj;
        if (slot == 7 && remaining > 0) {
            int64_t nn = new_node_idx++;
            for (int k = 0; k < 3; k++)
                t->Coordinates[3 * nn + k] = t->Coordinates[3 * cur + k] + t->Sizes[cur] * octant_offsets[7][k];
            t->Sizes[nn] = t->Sizes[cur] * 0.5;
            t->parent[nn - np] = cur;
            if (last < 0) t->FirstSubnode[cur] = nn;
            else t->NextBranch[last] = nn;
            t->NextBranch[nn] = t->NextBranch[cur];
            cur = nn;
            slot = 0;
            last = -1;
        }
        int64_t c = lo + jj;
        double mi = t->Masses[c];
        t->Masses[cur] += mi;
        int64_t ac = cur - np;
        for (int k = 0; k < 3; k++) t->acc[3 * ac + k] += mi * t->Coordinates[3 * c + k];
        if (!(t->Softenings[cur] >= t->Softenings[c])) t->Softenings[cur] = t->Softenings[c];
        if (t->HasQuads) t->pparent[c] = cur;
        if (last < 0) t->FirstSubnode[cur] = c;
        else t->NextBranch[last] = c;
        last = c;
        slot++;
    }
    if (last >= 0) t->NextBranch[last] = t->NextBranch[cur];
    return new_node_idx;
}

typedef struct {
    int64_t node, lo, hi, shift, rnd;
} frame;

/* octree.py:817-1091, serial path (the frontier-parallel path builds the same tree; the
 * reference's own test_parallel_split_matches_serial pins that equivalence). */
static void build_tree_radix(ptgo_tree *t, const double *pts, const double *m, const double *h, int64_t N)
{
    double center[3], size;
    ptgo_bounding_cube(pts, N, center, &size);
    int64_t *keys = malloc(8 * N);
    ptgo_morton_keys(pts, N, center, size, keys);
    t->TreewalkIndices = malloc(8 * N);
    ptgo_radix_argsort(keys, N, t->TreewalkIndices);
    initialize(t, N);
    int64_t *order = t->TreewalkIndices;
    int64_t *keys_sorted = malloc(8 * N);
    /* octree.py:577-603 */
    for (int64_t i = 0; i < N; i++) {
        int64_t oi = order[i];
        for (int k = 0; k < 3; k++) t->Coordinates[3 * i + k] = pts[3 * oi + k];
        t->Masses[i] = m[oi];
        t->Softenings[i] = h[oi];
        keys_sorted[i] = keys[oi];
    }
    free(keys);
    int64_t root = N;
    t->Sizes[root] = size;
    for (int k = 0; k < 3; k++) t->Coordinates[3 * root + k] = center[k];
    int64_t new_node_idx = N + 1;
    t->pparent = malloc(8 * (t->HasQuads ? N : 1));
    memset(t->pparent, 0xff, 8 * (t->HasQuads ? N : 1));

    int64_t cap = 64, top = 0;
    frame *st = malloc(sizeof(frame) * cap);
    st[top++] = (frame){root, 0, N, 3 * (MORTON_BITS - 1), 0};
    int64_t bs[8], be[8], bd[8];
    while (top > 0) {
        frame f = st[--top];
        int64_t node = f.node, lo = f.lo, hi = f.hi, shift = f.shift, rnd = f.rnd;
        if (shift < 0) {
            if (rnd >= MORTON_MAX_ROUNDS) {
                while (new_node_idx + (hi - lo) / 7 + 2 > t->NumNodes) increase_tree_size(t);
                new_node_idx = build_bucket(t, node, lo, hi, new_node_idx);
                continue;
            }
            rekey_slice(t, keys_sorted, node, lo, hi);
            shift = 3 * (MORTON_BITS - 1);
            rnd += 1;
        }
        int64_t last_child = -1;
        int nb = octant_runs(keys_sorted, lo, hi, (int)shift, bs, be, bd);
        for (int b = 0; b < nb; b++) {
            int64_t i = bs[b], j = be[b], d = bd[b], c;
            if (j - i == 1) {
                c = i;
                double mi = t->Masses[i];
                t->Masses[node] += mi;
                int64_t an = node - N;
                for (int k = 0; k < 3; k++) t->acc[3 * an + k] += mi * t->Coordinates[3 * i + k];
                if (!(t->Softenings[node] >= t->Softenings[i])) t->Softenings[node] = t->Softenings[i];
                if (t->HasQuads) t->pparent[i] = node;
            } else {
                while (new_node_idx + 1 > t->NumNodes) increase_tree_size(t);
                int64_t cnode = new_node_idx++;
                for (int k = 0; k < 3; k++)
                    t->Coordinates[3 * cnode + k] =
                        t->Coordinates[3 * node + k] + t->Sizes[node] * octant_offsets[d][k];
                t->Sizes[cnode] = t->Sizes[node] * 0.5;
                t->parent[cnode - N] = node;
                c = cnode;
                if (top >= cap) {
                    cap *= 2;
                    st = realloc(st, sizeof(frame) * cap);
                }
                st[top++] = (frame){cnode, i, j, shift - 3, rnd};
            }
            if (last_child < 0) t->FirstSubnode[node] = c;
            else t->NextBranch[last_child] = c;
            last_child = c;
        }
        if (last_child >= 0) t->NextBranch[last_child] = t->NextBranch[node];
    }
    free(st);
    free(keys_sorted);
    t->node_hi = new_node_idx;
}

/* octree.py:199-231 */
static void rollup_moments(ptgo_tree *t)
{
    int64_t np = t->NumParticles;
    for (int64_t node = t->node_hi - 1; node > np - 1; node--) {
        int64_t a = node - np;
        double mm = t->Masses[node];
        if (mm == 0.0) continue;
        double inv = 1.0 / mm;
        double d2 = 0.0;
        int64_t p = t->parent[a];
        for (int k = 0; k < 3; k++) {
            double com = t->acc[3 * a + k] * inv;
            double dx = com - t->Coordinates[3 * node + k];
            d2 += dx * dx;
            t->Coordinates[3 * node + k] = com;
        }
        t->Deltas[node] = sqrt(d2);
        if (p >= np) {
            int64_t pa = p - np;
            for (int k = 0; k < 3; k++) t->acc[3 * pa + k] += t->acc[3 * a + k];
            t->Masses[p] += mm;
            if (!(t->Softenings[p] >= t->Softenings[node])) t->Softenings[p] = t->Softenings[node];
        }
    }
}

/* octree.py:234-249 */
static void shift_quad_into(double *Q, int64_t p, double m, double r0, double r1, double r2)
{
    double rsq = r0 * r0 + r1 * r1 + r2 * r2;
    double *q = Q + 9 * p;
    q[0] += m * (3.0 * r0 * r0 - rsq);
    q[1] += m * 3.0 * r0 * r1;
    q[2] += m * 3.0 * r0 * r2;
    q[3] += m * 3.0 * r1 * r0;
    q[4] += m * (3.0 * r1 * r1 - rsq);
    q[5] += m * 3.0 * r1 * r2;
    q[6] += m * 3.0 * r2 * r0;
    q[7] += m * 3.0 * r2 * r1;
    q[8] += m * (3.0 * r2 * r2 - rsq);
}

/* octree.py:252-292 */
static void rollup_quadrupoles(ptgo_tree *t)
{
    int64_t np = t->NumParticles;
    double *Q = t->Quadrupoles, *C = t->Coordinates;
    for (int64_t i = 0; i < np; i++) {
        int64_t p = t->pparent[i];
        if (p < np) continue;
        shift_quad_into(Q, p, t->Masses[i], C[3 * i] - C[3 * p], C[3 * i + 1] - C[3 * p + 1],
                        C[3 * i + 2] - C[3 * p + 2]);
    }
    for (int64_t node = t->node_hi - 1; node > np - 1; node--) {
        int64_t p = t->parent[node - np];
        if (p < np) continue;
        double mm = t->Masses[node];
        if (mm == 0.0) continue;
        double r0 = C[3 * node] - C[3 * p], r1 = C[3 * node + 1] - C[3 * p + 1], r2 = C[3 * node + 2] - C[3 * p + 2];
        for (int k = 0; k < 9; k++) Q[9 * p + k] += Q[9 * node + k];
        shift_quad_into(Q, p, mm, r0, r1, r2);
    }
}

/* Octree.__init__ radix path (octree.py:668-723) */
void *ptgo_build(const double *pos, const double *m, const double *h, int64_t n, int quadrupole, int compute_moments)
{
    ptgo_tree *t = calloc(1, sizeof(ptgo_tree));
    t->HasQuads = quadrupole;
    build_tree_radix(t, pos, m, h, n);
    if (compute_moments) {
        rollup_moments(t);
        if (quadrupole) rollup_quadrupoles(t);
    }
    return t;
}

void ptgo_tree_free(void *vt)
{
    ptgo_tree *t = vt;
    if (!t) return;
    free(t->Coordinates); free(t->Masses); free(t->Softenings); free(t->Sizes); free(t->Deltas);
    free(t->Quadrupoles); free(t->FirstSubnode); free(t->NextBranch); free(t->TreewalkIndices);
    free(t->acc); free(t->parent); free(t->pparent);
    free(t);
}

/* info[0]=NumParticles, [1]=NumNodes (allocated), [2]=node_hi, [3]=HasQuads, [4]=HasCoincidentPoints */
void ptgo_tree_info(void *vt, int64_t *info)
{
    ptgo_tree *t = vt;
    info[0] = t->NumParticles;
    info[1] = t->NumNodes;
    info[2] = t->node_hi;
    info[3] = t->HasQuads;
    info[4] = t->HasCoincidentPoints;
}

/* Copies the first `count` elements of one attribute array into dst (caller-sized).
 * which: 0 Coordinates(3) 1 Masses 2 Softenings 3 Sizes 4 Deltas 5 Quadrupoles(9)
 *        6 FirstSubnode 7 NextBranch 8 TreewalkIndices(NumParticles)                        */
int ptgo_tree_get(void *vt, int which, int64_t count, void *dst)
{
    ptgo_tree *t = vt;
    switch (which) {
    case 0: memcpy(dst, t->Coordinates, 24 * count); return 0;
    case 1: memcpy(dst, t->Masses, 8 * count); return 0;
    case 2: memcpy(dst, t->Softenings, 8 * count); return 0;
    case 3: memcpy(dst, t->Sizes, 8 * count); return 0;
    case 4: memcpy(dst, t->Deltas, 8 * count); return 0;
    case 5: if (!t->Quadrupoles) return -1; memcpy(dst, t->Quadrupoles, 72 * count); return 0;
    case 6: memcpy(dst, t->FirstSubnode, 8 * count); return 0;
    case 7: memcpy(dst, t->NextBranch, 8 * count); return 0;
    case 8: memcpy(dst, t->TreewalkIndices, 8 * count); return 0;
    }
    return -2;
}

/* ------------------------------------------------------------------------------------------ */
/* Grouped walk: grouped_treewalk.py:45-241 (+ treewalk.acceptance_criterion :22-25)          */
/* ------------------------------------------------------------------------------------------ */

/* treewalk.py:22-25 */
static inline int acceptance_criterion(double r, double h, double size, double delta, double theta)
{
    double a = size / theta + delta, b = h + size * 0.6 + delta;
    return r > (a > b ? a : b);
}

/* grouped_treewalk.py:55-153, all flags as runtime arguments */
static inline void field_kernel(const ptgo_tree *t, int64_t no, int64_t a, int64_t b, const double *pos,
                                const double *soft, double *acc, int W, int want_pot, int want_accel,
                                int want_tidal, int quadrupole)
{
    double cx = t->Coordinates[3 * no], cy = t->Coordinates[3 * no + 1], cz = t->Coordinates[3 * no + 2];
    double M = t->Masses[no], hs = t->Softenings[no];
    int is_node = quadrupole ? (no >= t->NumParticles) : 0;
    double qxx = 0, qxy = 0, qxz = 0, qyx = 0, qyy = 0, qyz = 0, qzx = 0, qzy = 0, qzz = 0;
    if (quadrupole && is_node) {
        const double *q = t->Quadrupoles + 9 * no;
        qxx = q[0]; qxy = q[1]; qxz = q[2]; qyx = q[3]; qyy = q[4]; qyz = q[5]; qzx = q[6]; qzy = q[7]; qzz = q[8];
    }
    for (int64_t tt = a; tt < b; tt++) {
        double dx = cx - pos[3 * tt], dy = cy - pos[3 * tt + 1], dz = cz - pos[3 * tt + 2];
        double r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 > 0)) continue;
        double r = sqrt(r2);
        double ht = hs > soft[tt] ? hs : soft[tt];
        int softened = r < ht;
        double phi = 0, K = 0, Q = 0;
        if (want_pot) phi = softened ? M * ptgo_potential_kernel(r, ht) : -M / r;
        if (want_accel || want_tidal) K = softened ? M * ptgo_force_kernel(r, ht) : M / (r * r2);
        if (want_tidal) Q = softened ? M * ptgo_tidal_kernel(r, ht) : 3.0 * K / r2;
        double gx = 0, gy = 0, gz = 0;
        if (want_accel) { gx = K * dx; gy = K * dy; gz = K * dz; }
        double txx = 0, tyy = 0, tzz = 0, txy = 0, txz = 0, tyz = 0;
        if (want_tidal) {
            txx = Q * dx * dx - K; tyy = Q * dy * dy - K; tzz = Q * dz * dz - K;
            txy = Q * dx * dy; txz = Q * dx * dz; tyz = Q * dy * dz;
        }
        if (quadrupole && is_node) {
            double r5inv = 1.0 / (r2 * r2 * r);
            double r7inv = r5inv / r2;
            double qdx = qxx * dx + qxy * dy + qxz * dz;
            double qdy = qyx * dx + qyy * dy + qyz * dz;
            double qdz = qzx * dx + qzy * dy + qzz * dz;
            double A = dx * qdx + dy * qdy + dz * qdz;
            if (want_pot) phi += -0.5 * A * r5inv;
            if (want_accel) {
                double quad_fac = A * r7inv;
                gx += 2.5 * quad_fac * dx - qdx * r5inv;
                gy += 2.5 * quad_fac * dy - qdy * r5inv;
                gz += 2.5 * quad_fac * dz - qdz * r5inv;
            }
            if (want_tidal) {
                double diag = -2.5 * A * r7inv;
                double fac = 17.5 * A * r7inv / r2;
                txx += qxx * r5inv - 10.0 * dx * qdx * r7inv + diag + fac * dx * dx;
                tyy += qyy * r5inv - 10.0 * dy * qdy * r7inv + diag + fac * dy * dy;
                tzz += qzz * r5inv - 10.0 * dz * qdz * r7inv + diag + fac * dz * dz;
                txy += qxy * r5inv - 5.0 * (dx * qdy + dy * qdx) * r7inv + fac * dx * dy;
                txz += qxz * r5inv - 5.0 * (dx * qdz + dz * qdx) * r7inv + fac * dx * dz;
                tyz += qyz * r5inv - 5.0 * (dy * qdz + dz * qdy) * r7inv + fac * dy * dz;
            }
        }
        double *o = acc + (tt - a) * W;
        int k = 0;
        if (want_pot) o[k++] += phi;
        if (want_accel) { o[k] += gx; o[k + 1] += gy; o[k + 2] += gz; k += 3; }
        if (want_tidal) {
            o[k] += txx; o[k + 1] += txy; o[k + 2] += txz;
            o[k + 3] += txy; o[k + 4] += tyy; o[k + 5] += tyz;
            o[k + 6] += txz; o[k + 7] += tyz; o[k + 8] += tzz;
        }
    }
}

/* grouped_treewalk.py:169-239 (core) and :350-397 (FieldsTarget_grouped packing).
 * out is (n, W) with W = want_pot + 3*want_accel + 9*want_tidal, scaled by G.
 * counters (nullable): [0] = sum over groups of (elements summed x targets in group),
 *                      [1] = sum over groups of nodes/elements visited by the traversal.     */
void ptgo_fields_grouped(void *vt, const double *pos, const double *soft, int64_t n, int want_pot, int want_accel,
                         int want_tidal, int quadrupole, int64_t group_size, double G, double theta, int nthreads,
                         double *out, int64_t *counters)
{
    const ptgo_tree *t = vt;
    int W = want_pot + 3 * want_accel + 9 * want_tidal;
    int64_t ngroups = (n + group_size - 1) / group_size;
    int64_t tot_int = 0, tot_vis = 0;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) reduction(+ : tot_int, tot_vis)
#endif
    for (int64_t gi = 0; gi < ngroups; gi++) {
        int64_t a = gi * group_size, b = a + group_size < n ? a + group_size : n, m = b - a;
        double bmin0 = pos[3 * a], bmin1 = pos[3 * a + 1], bmin2 = pos[3 * a + 2];
        double bmax0 = bmin0, bmax1 = bmin1, bmax2 = bmin2, hmax = soft[a];
        for (int64_t tt = a + 1; tt < b; tt++) {
            const double *p = pos + 3 * tt;
            if (p[0] < bmin0) bmin0 = p[0]; else if (p[0] > bmax0) bmax0 = p[0];
            if (p[1] < bmin1) bmin1 = p[1]; else if (p[1] > bmax1) bmax1 = p[1];
            if (p[2] < bmin2) bmin2 = p[2]; else if (p[2] > bmax2) bmax2 = p[2];
            if (soft[tt] > hmax) hmax = soft[tt];
        }
        double *acc = calloc(m * W, sizeof(double));
        int64_t no = t->NumParticles;
        while (no > -1) {
            tot_vis++;
            const double *c = t->Coordinates + 3 * no;
            double dxm = 0, dym = 0, dzm = 0;
            if (c[0] < bmin0) dxm = bmin0 - c[0]; else if (c[0] > bmax0) dxm = c[0] - bmax0;
            if (c[1] < bmin1) dym = bmin1 - c[1]; else if (c[1] > bmax1) dym = c[1] - bmax1;
            if (c[2] < bmin2) dzm = bmin2 - c[2]; else if (c[2] > bmax2) dzm = c[2] - bmax2;
            double r_min = sqrt(dxm * dxm + dym * dym + dzm * dzm);
            double hh = t->Softenings[no] > hmax ? t->Softenings[no] : hmax;
            if (no < t->NumParticles || acceptance_criterion(r_min, hh, t->Sizes[no], t->Deltas[no], theta)) {
                field_kernel(t, no, a, b, pos, soft, acc, W, want_pot, want_accel, want_tidal, quadrupole);
                tot_int += m;
                no = t->NextBranch[no];
            } else {
                no = t->FirstSubnode[no];
            }
        }
        for (int64_t tt = a; tt < b; tt++)
            for (int k = 0; k < W; k++) out[tt * W + k] = G * acc[(tt - a) * W + k];
        free(acc);
    }
    if (counters) {
        counters[0] = tot_int;
        counters[1] = tot_vis;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Brute force: bruteforce.py:7-386 (target form and the parallel self forms)                 */
/* ------------------------------------------------------------------------------------------ */

/* Target form, any subset of fields packed like the walk (pot, accel, tidal).
 * Potential follows bruteforce.py:7-42 (dx = target - source, r == 0 skipped);
 * acceleration bruteforce.py:209-248 (dx = source - target, r2 == 0 skipped);
 * tidal bruteforce.py:353-386 with _add_tidal_pair :259-282.                                   */
void ptgo_fields_target_bruteforce(const double *xt, const double *ht, int64_t nt, const double *xs, const double *ms,
                                   const double *hs, int64_t ns, int want_pot, int want_accel, int want_tidal,
                                   double G, int nthreads, double *out)
{
    int W = want_pot + 3 * want_accel + 9 * want_tidal;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < nt; i++) {
        double phi = 0, ax = 0, ay = 0, az = 0, T[9] = {0};
        for (int64_t j = 0; j < ns; j++) {
            double h = hs[j] > ht[i] ? hs[j] : ht[i];
            if (want_pot) {
                double d0 = xt[3 * i] - xs[3 * j], d1 = xt[3 * i + 1] - xs[3 * j + 1], d2 = xt[3 * i + 2] - xs[3 * j + 2];
                double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                if (r != 0) {
                    if (r < h) phi += ms[j] * ptgo_potential_kernel(r, h);
                    else phi -= ms[j] / r;
                }
            }
            if (want_accel || want_tidal) {
                double d0 = xs[3 * j] - xt[3 * i], d1 = xs[3 * j + 1] - xt[3 * i + 1], d2 = xs[3 * j + 2] - xt[3 * i + 2];
                double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                if (r2 == 0) continue;
                double r = sqrt(r2);
                if (want_accel) {
                    if (r < h) {
                        double kern = ptgo_force_kernel(r, h);
                        ax += kern * ms[j] * d0; ay += kern * ms[j] * d1; az += kern * ms[j] * d2;
                    } else {
                        double fac = 1 / (r2 * r);
                        ax += ms[j] * fac * d0; ay += ms[j] * fac * d1; az += ms[j] * fac * d2;
                    }
                }
                if (want_tidal) {
                    double K, Q;
                    if (r < h) { K = ms[j] * ptgo_force_kernel(r, h); Q = ms[j] * ptgo_tidal_kernel(r, h); }
                    else { K = ms[j] / (r * r2); Q = 3.0 * K / r2; }
                    T[0] += Q * d0 * d0 - K; T[4] += Q * d1 * d1 - K; T[8] += Q * d2 * d2 - K;
                    double txy = Q * d0 * d1, txz = Q * d0 * d2, tyz = Q * d1 * d2;
                    T[1] += txy; T[3] += txy; T[2] += txz; T[6] += txz; T[5] += tyz; T[7] += tyz;
                }
            }
        }
        double *o = out + i * W;
        int k = 0;
        if (want_pot) o[k++] = G * phi;
        if (want_accel) { o[k] = G * ax; o[k + 1] = G * ay; o[k + 2] = G * az; k += 3; }
        if (want_tidal) for (int q = 0; q < 9; q++) o[k + q] = G * T[q];
    }
}

/* Self form: bruteforce.py:86-116 (potential: i == j excluded, coincident distinct particles
 * inside h keep the softened term), :164-206 (accel: r2 == 0 skipped), :319-350 (tidal).      */
void ptgo_fields_self_bruteforce(const double *x, const double *m, const double *h, int64_t n, int want_pot,
                                 int want_accel, int want_tidal, double G, int nthreads, double *out)
{
    int W = want_pot + 3 * want_accel + 9 * want_tidal;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < n; i++) {
        double phi = 0, ax = 0, ay = 0, az = 0, T[9] = {0};
        for (int64_t j = 0; j < n; j++) {
            if (i == j) continue;
            double hh = h[i] > h[j] ? h[i] : h[j];
            if (want_pot) {
                double d0 = x[3 * i] - x[3 * j], d1 = x[3 * i + 1] - x[3 * j + 1], d2 = x[3 * i + 2] - x[3 * j + 2];
                double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                if (r < hh) phi += m[j] * ptgo_potential_kernel(r, hh);
                else if (r > 0) phi -= m[j] / r;
            }
            if (want_accel || want_tidal) {
                double d0 = x[3 * j] - x[3 * i], d1 = x[3 * j + 1] - x[3 * i + 1], d2 = x[3 * j + 2] - x[3 * i + 2];
                double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                if (r2 == 0) continue;
                double r = sqrt(r2);
                if (want_accel) {
                    if (r < hh) {
                        double kern = ptgo_force_kernel(r, hh);
                        ax += kern * m[j] * d0; ay += kern * m[j] * d1; az += kern * m[j] * d2;
                    } else {
                        double fac = 1 / (r2 * r);
                        ax += m[j] * fac * d0; ay += m[j] * fac * d1; az += m[j] * fac * d2;
                    }
                }
                if (want_tidal) {
                    double K, Q;
                    if (r < hh) { K = m[j] * ptgo_force_kernel(r, hh); Q = m[j] * ptgo_tidal_kernel(r, hh); }
                    else { K = m[j] / (r * r2); Q = 3.0 * K / r2; }
                    T[0] += Q * d0 * d0 - K; T[4] += Q * d1 * d1 - K; T[8] += Q * d2 * d2 - K;
                    double txy = Q * d0 * d1, txz = Q * d0 * d2, tyz = Q * d1 * d2;
                    T[1] += txy; T[3] += txy; T[2] += txz; T[6] += txz; T[5] += tyz; T[7] += tyz;
                }
            }
        }
        double *o = out + i * W;
        int k = 0;
        if (want_pot) o[k++] = G * phi;
        if (want_accel) { o[k] = G * ax; o[k + 1] = G * ay; o[k + 2] = G * az; k += 3; }
        if (want_tidal) for (int q = 0; q < 9; q++) o[k + q] = G * T[q];
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Column density (treewalk.py:821-1276): particles are uniform spheres of radius Softenings  */
/* ------------------------------------------------------------------------------------------ */

#define PTGO_PI 3.14159265358979323846

/* ColumnDensityWalk_singleray (treewalk.py:904-973), one target and one unit ray; the ray-traced
 * walk is exact (a node opens whenever its inflated bounding sphere could meet the ray). */
static double column_singleray(const ptgo_tree *t, const double *pos, const double *ray)
{
    const double fac_density = 3.0 / (4.0 * PTGO_PI);
    double column = 0.0;
    int64_t no = t->NumParticles;
    const int64_t stop = t->NextBranch[no];
    while (no > -1 && no != stop) {
        double dx0 = t->Coordinates[3 * no] - pos[0];
        double dx1 = t->Coordinates[3 * no + 1] - pos[1];
        double dx2 = t->Coordinates[3 * no + 2] - pos[2];
        double r2 = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
        double r = sqrt(r2);
        double z = ray[0] * dx0 + ray[1] * dx1 + ray[2] * dx2;
        if (r2 - z * z < 0) {
            no = t->NextBranch[no];
            continue;
        }
        double h_no = t->Softenings[no];
        if (no < t->NumParticles) {
            if (h_no > 0) {
                double hinv = 1.0 / h_no;
                double fac = fac_density * t->Masses[no] * hinv * hinv;
                double r_proj = sqrt(r2 - z * z);
                double q = r_proj * hinv;
                if (r_proj < h_no) {
                    if (r > h_no) {
                        if (z > 0) column += fac * 2 * sqrt(1 - q * q);
                    } else {
                        column += fac * (z * hinv + sqrt(1 - q * q));
                    }
                }
            }
            no = t->NextBranch[no];
        } else {
            double R_eff = t->Sizes[no] * 0.8660254037844386 + t->Deltas[no];
            int hit = r < h_no + R_eff;
            if (!hit && z > 0) hit = (r2 - z * z) < (h_no + R_eff) * (h_no + R_eff);
            no = hit ? t->FirstSubnode[no] : t->NextBranch[no];
        }
    }
    return column;
}

/* ColumnDensity_tree's ray-traced branches (treewalk.py:1031-1084): out[t][i] for nt targets and
 * nr rays, shared (per_target = 0, rays (nr,3)) or per target (per_target = 1, rays (nt,nr,3)). */
void ptgo_column_rays(void *vt, const double *pos, int64_t nt, const double *rays, int64_t nr, int per_target,
                      double *out, int nthreads)
{
    const ptgo_tree *t = vt;
    if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
    for (int64_t j = 0; j < nt; j++)
        for (int64_t i = 0; i < nr; i++)
            out[j * nr + i] = column_singleray(t, pos + 3 * j, per_target ? rays + 3 * (j * nr + i) : rays + 3 * i);
}

/* angular_bin_scalar (treewalk.py:38-45) */
static int angular_bin(double d0, double d1, double d2)
{
    if (fabs(d0) > fabs(d1) && fabs(d0) > fabs(d2)) return d0 > 0 ? 0 : 1;
    if (fabs(d1) > fabs(d2)) return d1 > 0 ? 2 : 3;
    return d2 > 0 ? 4 : 5;
}

/* ColumnDensityBinned_grouped (treewalk.py:1197-1276): groups of gs consecutive targets share one
 * traversal opened by the node's distance to the group box (gs = 1: ColumnDensityWalk_binned,
 * treewalk.py:976-1028); out (nt, 6) mean column per equal-area sky bin. */
void ptgo_column_binned(void *vt, const double *pos, int64_t nt, int64_t gs, double theta, double *out, int nthreads)
{
    const ptgo_tree *t = vt;
    const double bin_size = (4.0 * PTGO_PI) / 6.0;
    if (nthreads <= 0) nthreads = 1;
    memset(out, 0, sizeof(double) * (size_t)nt * 6);
    const int64_t ngroups = (nt + gs - 1) / gs;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (int64_t gi = 0; gi < ngroups; gi++) {
        const int64_t a = gi * gs, b = a + gs < nt ? a + gs : nt;
        double bmin[3], bmax[3];
        for (int k = 0; k < 3; k++) bmin[k] = bmax[k] = pos[3 * a + k];
        for (int64_t j = a + 1; j < b; j++)
            for (int k = 0; k < 3; k++) {
                if (pos[3 * j + k] < bmin[k]) bmin[k] = pos[3 * j + k];
                else if (pos[3 * j + k] > bmax[k]) bmax[k] = pos[3 * j + k];
            }
        int64_t no = t->NumParticles;
        const int64_t stop = t->NextBranch[no];
        while (no > -1 && no != stop) {
            const double c0 = t->Coordinates[3 * no], c1 = t->Coordinates[3 * no + 1], c2 = t->Coordinates[3 * no + 2];
            const double h_no = t->Softenings[no], M = t->Masses[no];
            if (no < t->NumParticles) {
                const double hh = h_no * h_no;
                for (int64_t j = a; j < b; j++) {
                    double d0 = c0 - pos[3 * j], d1 = c1 - pos[3 * j + 1], d2 = c2 - pos[3 * j + 2];
                    double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                    int bin = angular_bin(d0, d1, d2);
                    double *o = out + 6 * j;
                    if (r2 > hh) {
                        o[bin] += M / r2 / bin_size;
                    } else if (h_no > 0) {
                        double col0 = M * (3 / (4 * PTGO_PI * hh));
                        double fac = sqrt(r2) / h_no;
                        double col_iso = (1 - fac) * col0;
                        for (int k = 0; k < 6; k++) o[k] += col_iso;
                        o[bin] += col0 * fac * 2;
                    }
                }
                no = t->NextBranch[no];
            } else {
                double dm[3];
                const double c[3] = {c0, c1, c2};
                for (int k = 0; k < 3; k++) dm[k] = c[k] < bmin[k] ? bmin[k] - c[k] : (c[k] > bmax[k] ? c[k] - bmax[k] : 0.0);
                double r_min = sqrt(dm[0] * dm[0] + dm[1] * dm[1] + dm[2] * dm[2]);
                double size = t->Sizes[no], delta = t->Deltas[no];
                double c_a = size / theta + delta, c_b = h_no + size * 0.6 + delta;
                if (r_min > (c_a > c_b ? c_a : c_b)) {
                    for (int64_t j = a; j < b; j++) {
                        double d0 = c0 - pos[3 * j], d1 = c1 - pos[3 * j + 1], d2 = c2 - pos[3 * j + 2];
                        double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                        out[6 * j + angular_bin(d0, d1, d2)] += M / r2 / bin_size;
                    }
                    no = t->NextBranch[no];
                } else {
                    no = t->FirstSubnode[no];
                }
            }
        }
    }
}
