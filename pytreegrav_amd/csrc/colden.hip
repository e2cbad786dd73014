// Column density through the octree (treewalk.py:821-1276): particles are uniform spheres of radius
// equal to their softening.  Both walks follow the tree's own depth-first links (FirstSubnode /
// NextBranch of the reference layout), so they need no stack and accumulate every sum in the
// reference's order.
//
// Ray-traced (ptg_column_rays): one thread per (target, ray), ray-major so that a wavefront walks 64
// Morton-adjacent targets along one direction and its lanes open nearly the same nodes.  The walk is
// exact: a node is opened whenever its bounding sphere, inflated by the largest particle radius in
// it, could meet the ray; only particles contribute, by the closed-form chord integral.
//
// Angular-binned (ptg_column_binned), the reference's grouped estimator: a node whose distance to
// the box of group_size consecutive targets passes the opening test puts its mass in one of six
// equal-area sky bins per target; particles are summed per target.  For group_size <= 64 one
// wavefront walks a group over the FP64 walk pack (64 candidates per step, opened nodes pushed on an
// LDS stack, accepted nodes and particles listed in LDS); every list entry is then evaluated for
// the group's targets by 64 / group_size lane slices at once, whose partial bins are summed at the
// end.  Larger groups take one thread per group along the depth-first links.
#include <algorithm>

#include "ptg_common.hpp"

#define CD_THREADS 256

// one 32-byte record of the FP64 walk pack
struct alignas(32) D4 {
    double x, y, z, w;
};

// a corrupt link (index out of range, or a chain longer than the element count) stops the walk and
// sets err[0]; the depth-first chain visits each element at most once
__device__ __forceinline__ bool cd_bad(int64_t no, int64_t ne, int64_t &visits)
{
    return no >= ne || ++visits > ne;
}

// adds one wave's counts to ctr[k] with one atomic (every lane of the wave takes part)
__device__ __forceinline__ void cd_count(unsigned long long *ctr, int k, int64_t v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if (ptg_lane() == 0 && v) atomicAdd(ctr + k, (unsigned long long)v);
}

// column along ray i from target j; visits counts the elements the walk touched
__device__ __forceinline__ double ray_column(const ptg_tree_arrays &t, int64_t n, int64_t ne, const double *tp,
                                             const double *rays, int64_t nr, int per_target, int64_t i, int64_t j,
                                             int32_t *err, int64_t &visits)
{
    const double *ray = per_target ? rays + 3 * (j * nr + i) : rays + 3 * i;
    const double rx = ray[0], ry = ray[1], rz = ray[2];
    const double px = tp[3 * j], py = tp[3 * j + 1], pz = tp[3 * j + 2];
    const double fac_density = 3.0 / (4.0 * 3.14159265358979323846);
    double column = 0.0;
    int64_t no = n;
    const int64_t stop = t.next_branch[n];
    while (no > -1 && no != stop) {
        if (cd_bad(no, ne, visits)) {
            atomicOr(err, 1);
            break;
        }
        const double dx = t.coordinates[3 * no] - px, dy = t.coordinates[3 * no + 1] - py;
        const double dz = t.coordinates[3 * no + 2] - pz;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double r = sqrt(r2);
        const double z = rx * dx + ry * dy + rz * dz;
        if (r2 - z * z < 0) {  // rounding below zero: skipped, as ColumnDensityWalk_singleray does
            no = t.next_branch[no];
            continue;
        }
        const double h_no = t.softenings[no];
        if (no < n) {
            if (h_no > 0) {  // a zero radius has no cross-section
                const double hinv = 1.0 / h_no;
                const double fac = fac_density * t.masses[no] * hinv * hinv;
                const double r_proj = sqrt(r2 - z * z);
                const double q = r_proj * hinv;
                if (r_proj < h_no) {
                    if (r > h_no) {  // target outside the sphere: the whole chord if it lies ahead
                        if (z > 0) column += fac * 2 * sqrt(1 - q * q);
                    } else {  // target inside: the forward half-chord
                        column += fac * (z * hinv + sqrt(1 - q * q));
                    }
                }
            }
            no = t.next_branch[no];
        } else {
            const double reach = h_no + t.sizes[no] * 0.8660254037844386 + t.deltas[no];
            bool open = r < reach;  // the node holds the ray's origin
            if (!open && z > 0) open = (r2 - z * z) < reach * reach;
            no = open ? t.first_subnode[no] : t.next_branch[no];
        }
    }
    return column;
}

__global__ __launch_bounds__(CD_THREADS) void k_column_rays(const ptg_tree_arrays t, int64_t n, int64_t ne,
                                                            const double *tp, int64_t nt, const double *rays,
                                                            int64_t nr, int per_target, double *out, int32_t *err,
                                                            unsigned long long *ctr)
{
    const int64_t id = (int64_t)blockIdx.x * CD_THREADS + threadIdx.x;
    int64_t visits = 0;
    if (id < nt * nr) {
        const int64_t i = id / nt, j = id - (id / nt) * nt;  // ray i, target j
        out[j * nr + i] = ray_column(t, n, ne, tp, rays, nr, per_target, i, j, err, visits);
    }
    if (ctr) cd_count(ctr, 0, visits);
}

// ---- ray-traced walk, one wavefront per 64 targets along one ray index ----
// The wavefront walks the union of its lanes' walks in depth-first preorder over the FP64 walk pack:
// a stack of (first slot, count, lane mask) sibling ranges, one element per iteration, its record
// read once for all lanes.  Each lane applies exactly the tests of the per-thread walk to the
// elements its own walk would reach (those whose ancestors it opened), in the same order, so every
// column gets the same terms summed in the same sequence.
#define CR_WAVES 4
#define CR_DEPTH 256  // sibling ranges on the stack (one per tree level on the current path)

struct CrLds {
    int s[CR_DEPTH], c[CR_DEPTH];
    uint64_t m[CR_DEPTH];
};

__global__ __launch_bounds__(64 * CR_WAVES) void k_column_rays_wave(const D4 *xyzm, const D4 *aux, const int2 *kids,
                                                                    int64_t nslots, const double *tp, int64_t nt,
                                                                    const double *rays, int64_t nr, int per_target,
                                                                    const int64_t *order, double *out, int32_t *err,
                                                                    unsigned long long *ctr)
{
    __shared__ CrLds lds[CR_WAVES];
    const int wave = threadIdx.x >> 6, lane = ptg_lane();
    CrLds &S = lds[wave];
    const int64_t ntb = (nt + 63) / 64;
    const int64_t task = (int64_t)blockIdx.x * CR_WAVES + wave;
    if (task >= ntb * nr) return;  // whole wavefronts only: task is uniform
    const int64_t i = task / ntb, jr = (task % ntb) * 64 + lane;  // ray i, rank jr of the target
    const bool live = jr < nt;
    // the target order of ray i (if given) puts targets whose rays run side by side in one wavefront
    const int64_t j = live && order ? order[i * nt + jr] : jr;
    double rx = 0, ry = 0, rz = 0, px = 0, py = 0, pz = 0;
    if (live) {
        const double *ray = per_target ? rays + 3 * (j * nr + i) : rays + 3 * i;
        rx = ray[0];
        ry = ray[1];
        rz = ray[2];
        px = tp[3 * j];
        py = tp[3 * j + 1];
        pz = tp[3 * j + 2];
    }
    const double fac_density = 3.0 / (4.0 * 3.14159265358979323846);
    const uint64_t me = 1ull << lane;
    double column = 0.0;
    int64_t visits = 0;
    int top = 0;
    const uint64_t all = ptg_ballot(live);  // outside the lane-0 branch: a ballot sees the active lanes only
    if (lane == 0) {
        S.s[0] = 0;  // the root's walk slot
        S.c[0] = 1;
        S.m[0] = all;
    }
    __builtin_amdgcn_wave_barrier();
    while (top >= 0) {
        const int sl = S.s[top], cnt = S.c[top];
        const uint64_t m = S.m[top];
        __builtin_amdgcn_wave_barrier();
        if (cnt > 1) {
            if (lane == 0) {
                S.s[top] = sl + 1;
                S.c[top] = cnt - 1;
            }
        } else {
            top--;
        }
        if (sl < 0 || sl >= nslots) {  // a corrupt range is reported, never followed
            if (lane == 0) atomicOr(err, 1);
            __builtin_amdgcn_wave_barrier();
            continue;
        }
        const D4 X = xyzm[sl], X2 = aux[sl];
        const int2 K = kids[sl];
        bool open = false;
        if (m & me) {
            visits++;
            const double dx = X.x - px, dy = X.y - py, dz = X.z - pz;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double r = sqrt(r2);
            const double z = rx * dx + ry * dy + rz * dz;
            if (!(r2 - z * z < 0)) {  // rounding below zero: skipped, as ColumnDensityWalk_singleray does
                const double h_no = X2.x;
                if (K.y == 0) {
                    if (h_no > 0) {  // a zero radius has no cross-section
                        const double hinv = 1.0 / h_no;
                        const double fac = fac_density * X.w * hinv * hinv;
                        const double r_proj = sqrt(r2 - z * z);
                        const double q = r_proj * hinv;
                        if (r_proj < h_no) {
                            if (r > h_no) {
                                if (z > 0) column += fac * 2 * sqrt(1 - q * q);
                            } else {
                                column += fac * (z * hinv + sqrt(1 - q * q));
                            }
                        }
                    }
                } else {
                    const double reach = h_no + X2.y * 0.8660254037844386 + X2.z;
                    open = r < reach;
                    if (!open && z > 0) open = (r2 - z * z) < reach * reach;
                }
            }
        }
        const uint64_t mo = ptg_ballot(open);
        if (mo && K.y > 0) {
            if (top + 1 >= CR_DEPTH) {  // deeper than the stack: the caller reruns the stackless walk
                if (lane == 0) atomicOr(err, 2);
                break;
            }
            top++;
            if (lane == 0) {
                S.s[top] = K.x;
                S.c[top] = std::min(K.y, 8);
                S.m[top] = mo;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (live) out[j * nr + i] = column;
    if (ctr) cd_count(ctr, 0, visits);
}

// angular_bin_scalar (treewalk.py:38-45)
__device__ __forceinline__ int cd_bin(double d0, double d1, double d2)
{
    if (fabs(d0) > fabs(d1) && fabs(d0) > fabs(d2)) return d0 > 0 ? 0 : 1;
    if (fabs(d1) > fabs(d2)) return d1 > 0 ? 2 : 3;
    return d2 > 0 ? 4 : 5;
}

__global__ __launch_bounds__(CD_THREADS) void k_column_binned(const ptg_tree_arrays t, int64_t n, int64_t ne,
                                                              const double *tp, int64_t nt, int64_t gs, double theta,
                                                              double *out, int32_t *err, unsigned long long *ctr)
{
    const int64_t gi = (int64_t)blockIdx.x * CD_THREADS + threadIdx.x;
    const int64_t a = gi * gs;
    if (a >= nt) return;
    const int64_t b = std::min(a + gs, nt);
    const double bin_size = (4.0 * 3.14159265358979323846) / 6.0;
    double bmin0 = tp[3 * a], bmin1 = tp[3 * a + 1], bmin2 = tp[3 * a + 2];
    double bmax0 = bmin0, bmax1 = bmin1, bmax2 = bmin2;
    for (int64_t j = a + 1; j < b; j++) {
        bmin0 = fmin(bmin0, tp[3 * j]);
        bmax0 = fmax(bmax0, tp[3 * j]);
        bmin1 = fmin(bmin1, tp[3 * j + 1]);
        bmax1 = fmax(bmax1, tp[3 * j + 1]);
        bmin2 = fmin(bmin2, tp[3 * j + 2]);
        bmax2 = fmax(bmax2, tp[3 * j + 2]);
    }
    for (int64_t j = a; j < b; j++)
        for (int k = 0; k < 6; k++) out[6 * j + k] = 0.0;
    int64_t no = n, visits = 0, pairs = 0;
    const int64_t stop = t.next_branch[n];
    while (no > -1 && no != stop) {
        if (cd_bad(no, ne, visits)) {
            atomicOr(err, 1);
            break;
        }
        const double cx = t.coordinates[3 * no], cy = t.coordinates[3 * no + 1], cz = t.coordinates[3 * no + 2];
        const double h_no = t.softenings[no], M = t.masses[no];
        if (no < n) {  // a particle contributes to every target of the group, no criterion
            pairs += b - a;
            const double hh = h_no * h_no;
            for (int64_t j = a; j < b; j++) {
                const double d0 = cx - tp[3 * j], d1 = cy - tp[3 * j + 1], d2 = cz - tp[3 * j + 2];
                const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                const int bin = cd_bin(d0, d1, d2);
                double *o = out + 6 * j;
                if (r2 > hh) {  // no overlap: a point mass at distance sqrt(r2)
                    o[bin] += M / r2 / bin_size;
                } else if (h_no > 0) {  // overlap: blend of the isotropic and directed limits
                    const double col0 = M * (3 / (4 * 3.14159265358979323846 * hh));
                    const double fac = sqrt(r2) / h_no;
                    const double col_iso = (1 - fac) * col0;
                    for (int k = 0; k < 6; k++) o[k] += col_iso;
                    o[bin] += col0 * fac * 2;
                }
            }
            no = t.next_branch[no];
        } else {
            // the reference's opening test (treewalk.py:22-25) at the node's distance to the box
            const double dxm = cx < bmin0 ? bmin0 - cx : (cx > bmax0 ? cx - bmax0 : 0.0);
            const double dym = cy < bmin1 ? bmin1 - cy : (cy > bmax1 ? cy - bmax1 : 0.0);
            const double dzm = cz < bmin2 ? bmin2 - cz : (cz > bmax2 ? cz - bmax2 : 0.0);
            const double r_min = sqrt(dxm * dxm + dym * dym + dzm * dzm);
            const double size = t.sizes[no], delta = t.deltas[no];
            const double c1 = size / theta + delta, c2 = h_no + size * 0.6 + delta;
            if (r_min > (c1 > c2 ? c1 : c2)) {
                pairs += b - a;
                for (int64_t j = a; j < b; j++) {
                    const double d0 = cx - tp[3 * j], d1 = cy - tp[3 * j + 1], d2 = cz - tp[3 * j + 2];
                    const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                    out[6 * j + cd_bin(d0, d1, d2)] += M / r2 / bin_size;
                }
                no = t.next_branch[no];
            } else {
                no = t.first_subnode[no];
            }
        }
    }
    if (ctr) {
        atomicAdd(ctr, (unsigned long long)pairs);
        atomicAdd(ctr + 1, (unsigned long long)visits);
    }
}


// ---- wavefront-per-group binned walk ----
#define CB_WAVES 4
#define CB_SCAP 1024   // LDS stack entries per wave (>= CB_SPILL + the 512 children one step can push)
#define CB_SPILL 512
#define CB_GCAP 16384  // global spill entries per wave
#define CB_ICAP 128
#define CB_BUDGET 32768  // main-pass visits after which a group is queued for the heavy pass
#define CB_HCAP 1024     // queued groups at most (further heavy groups finish in the main pass)
#define CB_W 64          // shares of a queued group
#define CB_D 4           // depth of the nodes whose children are dealt out to the shares

struct CbLds {
    int stack[CB_SCAP];
    D4 nodes[CB_ICAP];
    D4 parts[CB_ICAP];
    double ph[CB_ICAP];
};

__device__ __forceinline__ void cb_add(double (&acc)[6], int bin, double v)
{
#pragma unroll
    for (int k = 0; k < 6; k++) acc[k] += (k == bin) ? v : 0.0;
}

// list entries [0, cnt) on this lane's target, slice sl of `slices`
// far terms M / r^2 go to far[], divided by the bin's solid angle once at the end of the task;
// overlapping particles' terms go to near[]
__device__ __forceinline__ void cb_flush(const CbLds &L, int nn, int np, int sl, int slices, bool mine, double tx,
                                         double ty, double tz, double (&far)[6], double (&near)[6], int64_t &pairs)
{
    if (!mine) return;
    pairs += (nn > sl ? (nn - sl + slices - 1) / slices : 0) + (np > sl ? (np - sl + slices - 1) / slices : 0);
    for (int k = sl; k < nn; k += slices) {
        const D4 c = L.nodes[k];
        const double d0 = c.x - tx, d1 = c.y - ty, d2 = c.z - tz;
        const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
        cb_add(far, cd_bin(d0, d1, d2), c.w / r2);
    }
    for (int k = sl; k < np; k += slices) {
        const D4 c = L.parts[k];
        const double h = L.ph[k];
        const double d0 = c.x - tx, d1 = c.y - ty, d2 = c.z - tz;
        const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
        const int bin = cd_bin(d0, d1, d2);
        const double hh = h * h;
        if (r2 > hh) {
            cb_add(far, bin, c.w / r2);
        } else if (h > 0) {
            const double col0 = c.w * (3 / (4 * 3.14159265358979323846 * hh));
            const double fac = sqrt(r2) / h;
            const double iso = (1 - fac) * col0;
#pragma unroll
            for (int q = 0; q < 6; q++) near[q] += iso;
            cb_add(near, bin, col0 * fac * 2);
        }
    }
}

// One task: the traversal of group gi's targets, whole (main pass) or the share w of CB_W (heavy
// pass).  Stack entries carry the walk slot and the depth (slot * 8 + min(depth, 7)).  In the heavy
// pass every wave walks the levels above CB_D alike, wave 0 alone lists what they accept, and below
// an opened node of depth CB_D - 1 wave w descends only into the children whose slot is w modulo
// CB_W, so the CB_W shares cover the group's traversal exactly once.  Returns 0 done, 1 deferred (a
// main-pass group past CB_BUDGET visits, queued for the heavy pass), 2 stack overflow.
template <bool HEAVY>
__device__ int cb_task(CbLds &L, int *spill, const D4 *xyzm, const D4 *aux, const int2 *kids, int64_t nslots,
                       const double *tp, int64_t nt, int gs, double theta, int64_t gi, int w, int32_t *ctl,
                       int32_t *heavy, double (&acc)[6], int64_t &gpairs, int64_t &gvis)
{
    const int lane = ptg_lane();
    const int slices = 64 / gs;
    const int tl = lane % gs, sl = lane / gs;
    const int64_t a = gi * gs;
    const int mt = (int)std::min<int64_t>(gs, nt - a);
    // the group's box (lanes 0..mt-1 hold its targets)
    double bx0 = INFINITY, by0 = INFINITY, bz0 = INFINITY, bx1 = -INFINITY, by1 = -INFINITY, bz1 = -INFINITY;
    if (lane < mt) {
        bx0 = bx1 = tp[3 * (a + lane)];
        by0 = by1 = tp[3 * (a + lane) + 1];
        bz0 = bz1 = tp[3 * (a + lane) + 2];
    }
    bx0 = ptg_wave_min(bx0);
    by0 = ptg_wave_min(by0);
    bz0 = ptg_wave_min(bz0);
    bx1 = ptg_wave_max(bx1);
    by1 = ptg_wave_max(by1);
    bz1 = ptg_wave_max(bz1);
    const bool mine = sl < slices && tl < mt;
    const double tx = mine ? tp[3 * (a + tl)] : 0.0, ty = mine ? tp[3 * (a + tl) + 1] : 0.0;
    const double tz = mine ? tp[3 * (a + tl) + 2] : 0.0;
    double far[6], near[6];
    for (int q = 0; q < 6; q++) far[q] = near[q] = 0.0;
    gpairs = gvis = 0;
    bool budget = !HEAVY;
    int sp = 1, gsp = 0, nn = 0, np = 0;
    if (lane == 0) L.stack[0] = 0;  // the root: slot 0, depth 0
    __builtin_amdgcn_wave_barrier();
    for (;;) {
        if (sp == 0) {
            if (gsp == 0) break;
            const int k = std::min(gsp, CB_SPILL);
            for (int q = lane; q < k; q += 64) L.stack[q] = spill[gsp - k + q];
            __builtin_amdgcn_wave_barrier();
            gsp -= k;
            sp = k;
        }
        const int k = std::min(sp, 64);
        const int base = sp - k;
        const int enc = lane < k ? L.stack[base + lane] : 0;
        int e = enc >> 3;
        const int d = enc & 7;
        const bool ve = lane < k && enc >= 0 && e < nslots;
        if (!ve) e = 0;
        __builtin_amdgcn_wave_barrier();
        sp = base;
        // above the split depth only wave 0 lists and counts; the other waves only open
        const bool lists = !HEAVY || w == 0 || d >= CB_D;
        gvis += (int)__popcll(ptg_ballot(ve && lists));
        const D4 X = xyzm[e], X2 = aux[e];
        const int2 K = kids[e];
        const bool ispart = K.y == 0;
        const double dxm = X.x < bx0 ? bx0 - X.x : (X.x > bx1 ? X.x - bx1 : 0.0);
        const double dym = X.y < by0 ? by0 - X.y : (X.y > by1 ? X.y - by1 : 0.0);
        const double dzm = X.z < bz0 ? bz0 - X.z : (X.z > bz1 ? X.z - bz1 : 0.0);
        const double r_min = sqrt(dxm * dxm + dym * dym + dzm * dzm);
        const double c1 = X2.y / theta + X2.z, c2 = X2.x + X2.y * 0.6 + X2.z;
        const bool acc_node = ve && !ispart && r_min > (c1 > c2 ? c1 : c2);
        const bool open = ve && !ispart && !acc_node;
        const bool list_node = acc_node && lists, list_part = ve && ispart && lists;
        const uint64_t bn = ptg_ballot(list_node), bp = ptg_ballot(list_part);
        if (list_node) L.nodes[nn + (int)__popcll(bn & ptg_lanemask_lt())] = X;
        if (list_part) {
            const int at = np + (int)__popcll(bp & ptg_lanemask_lt());
            L.parts[at] = X;
            L.ph[at] = X2.x;
        }
        nn += (int)__popcll(bn);
        np += (int)__popcll(bp);
        const int nk = open ? std::min(std::max(K.y, 0), 8) : 0;
        // at depth CB_D - 1 (heavy pass) the one child, if any, whose slot is w modulo CB_W
        const bool split = HEAVY && d == CB_D - 1;
        const int cw = split ? (int)(((int64_t)w - K.x) % CB_W + CB_W) % CB_W : 0;
        const int cnt = split ? (cw < nk ? 1 : 0) : nk;
        const int incl = ptg_wave_incl_scan(cnt);
        const int tot = __shfl(incl, 63, 64);
        if (sp + tot > CB_SCAP) {
            if (gsp + CB_SPILL > CB_GCAP) return 2;
            for (int q = lane; q < CB_SPILL; q += 64) spill[gsp + q] = L.stack[q];
            __builtin_amdgcn_wave_barrier();
            for (int q = 0; q < sp - CB_SPILL; q += 64) {
                const int v = (q + lane < sp - CB_SPILL) ? L.stack[CB_SPILL + q + lane] : 0;
                __builtin_amdgcn_wave_barrier();
                if (q + lane < sp - CB_SPILL) L.stack[q + lane] = v;
                __builtin_amdgcn_wave_barrier();
            }
            gsp += CB_SPILL;
            sp -= CB_SPILL;
        }
        {
            const int at = sp + incl - cnt;
            const int cd = d < 7 ? d + 1 : 7;
            if (split) {
                if (cnt) L.stack[at] = (K.x + cw) * 8 + cd;
            } else {
                for (int c = 0; c < cnt; c++) L.stack[at + c] = (K.x + c) * 8 + cd;
            }
        }
        sp += tot;
        __builtin_amdgcn_wave_barrier();
        if (nn >= 64 || np >= 64) {
            cb_flush(L, nn, np, sl, slices, mine, tx, ty, tz, far, near, gpairs);
            __builtin_amdgcn_wave_barrier();
            nn = np = 0;
        }
        if (budget && gvis > CB_BUDGET) {
            int h = 0;
            if (lane == 0) h = atomicAdd(ctl + 1, 1);
            h = __builtin_amdgcn_readfirstlane(h);
            if (h < CB_HCAP) {
                if (lane == 0) heavy[h] = (int32_t)gi;
                return 1;
            }
            budget = false;  // the heavy queue is full: finish the group here
        }
    }
    cb_flush(L, nn, np, sl, slices, mine, tx, ty, tz, far, near, gpairs);
    __builtin_amdgcn_wave_barrier();
    const double bin_size = (4.0 * 3.14159265358979323846) / 6.0;
    for (int q = 0; q < 6; q++) acc[q] = far[q] / bin_size + near[q];
    // sum the slices' partial bins into lanes 0 .. gs-1
#pragma unroll
    for (int q = 0; q < 6; q++) {
        double v = acc[q];
        for (int s2 = 1; s2 < slices; s2++) {
            const double o = __shfl(acc[q], tl + s2 * gs, 64);
            if (sl == 0) v += o;
        }
        acc[q] = v;
    }
    return 0;
}

// Main pass (HEAVY = false): one task per group, results into out; groups past the visit budget are
// queued.  Heavy pass: CB_W tasks per queued group, partial bins into partial[task][target][6].
template <bool HEAVY>
__global__ __launch_bounds__(64 * CB_WAVES) void k_column_binned_wave(const D4 *xyzm, const D4 *aux, const int2 *kids,
                                                                      int64_t nslots, const double *tp, int64_t nt,
                                                                      int gs, double theta, double *out, int *spill_all,
                                                                      int32_t *ctl, int32_t *heavy, double *partial,
                                                                      int32_t *err, unsigned long long *counts)
{
    __shared__ CbLds lds[CB_WAVES];
    const int wave = threadIdx.x >> 6, lane = ptg_lane();
    CbLds &L = lds[wave];
    int *spill = spill_all + ((int64_t)blockIdx.x * CB_WAVES + wave) * CB_GCAP;
    const int64_t ntasks = HEAVY ? (int64_t)std::min(ctl[1], CB_HCAP) * CB_W : (nt + gs - 1) / gs;
    const int tl = lane % gs, sl = lane / gs;
    int64_t pairs = 0, vis = 0;
    for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(ctl + (HEAVY ? 2 : 0), 1);
        t = __builtin_amdgcn_readfirstlane(t);
        if (t >= ntasks) break;
        const int64_t gi = HEAVY ? (int64_t)heavy[t / CB_W] : (int64_t)t;
        const int w = HEAVY ? t % CB_W : 0;
        double acc[6];
        int64_t gpairs, gvis;
        const int r = cb_task<HEAVY>(L, spill, xyzm, aux, kids, nslots, tp, nt, gs, theta, gi, w, ctl, heavy, acc,
                                     gpairs, gvis);
        if (r == 2) {
            if (lane == 0) atomicOr(err, 2);
            continue;
        }
        if (r == 1) continue;
        pairs += gpairs;
        vis += gvis;
        const int64_t a = gi * gs;
        if (sl == 0 && a + tl < nt) {
            double *o = HEAVY ? partial + ((int64_t)t * gs + tl) * 6 : out + 6 * (a + tl);
            for (int q = 0; q < 6; q++) o[q] = acc[q];
        }
    }
    if (counts) {
        cd_count(counts, 0, pairs);
        cd_count(counts, 1, lane == 0 ? vis : 0);
    }
}

// out rows of the queued groups: their CB_W partial rows summed in share order
__global__ __launch_bounds__(256) void k_column_heavy_sum(const int32_t *ctl, const int32_t *heavy,
                                                          const double *partial, int64_t nt, int gs, double *out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t h = i / gs;
    const int tl = (int)(i % gs);
    if (h >= std::min(ctl[1], CB_HCAP)) return;
    const int64_t a = (int64_t)heavy[h] * gs;
    if (a + tl >= nt) return;
    for (int q = 0; q < 6; q++) {
        double v = 0.0;
        for (int w = 0; w < CB_W; w++) v += partial[((h * CB_W + w) * gs + tl) * 6 + q];
        out[6 * (a + tl) + q] = v;
    }
}

static int cb_cus()
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
    return cus;
}

// workspace: control words, the heavy-group queue, per-wave spill stacks, the heavy partial rows
struct CbWs {
    int32_t *ctl, *heavy;
    int *spill;
    double *partial;
};

static size_t cb_carve(void *ws, int64_t gs, CbWs *w)
{
    const int cus = cb_cus();
    char *p = (char *)ws;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *q = p ? p + off : nullptr;
        off += ptg_align256(bytes);
        return q;
    };
    w->ctl = (int32_t *)take(sizeof(int32_t) * 64);
    w->heavy = (int32_t *)take(sizeof(int32_t) * CB_HCAP);
    w->spill = (int *)take((size_t)cus * 2 * CB_WAVES * CB_GCAP * sizeof(int));
    w->partial = (double *)take((size_t)CB_HCAP * CB_W * std::max<int64_t>(gs, 1) * 6 * sizeof(double));
    return off;
}

extern "C" size_t ptg_column_workspace_bytes(int64_t group_size)
{
    CbWs w;
    return cb_carve(nullptr, group_size, &w);
}

__global__ void k_column_err_init(int32_t *err) { *err = 0; }
__global__ void k_column_ctl_init(int32_t *ctl)
{
    if (threadIdx.x < 64) ctl[threadIdx.x] = 0;
}

static ptg_status column_common(const ptg_tree_arrays *tree, int64_t n, int64_t nnodes, const double *targets,
                                int64_t nt, const char *who)
{
    if (!tree || !tree->coordinates || !tree->masses || !tree->softenings || !tree->sizes || !tree->deltas ||
        !tree->first_subnode || !tree->next_branch || n < 1 || nnodes < 1 || nt < 0 || (nt > 0 && !targets)) {
        ptg_set_error("%s: bad arguments", who);
        return PTG_ERR_ARG;
    }
    return PTG_OK;
}

extern "C" ptg_status ptg_column_rays(const ptg_tree_arrays *tree, int64_t n, int64_t nnodes,
                                      const ptg_walk_tree *wt, const double *targets, int64_t nt, const double *rays,
                                      int64_t nrays, int per_target, const int64_t *order, double *out, int32_t *err,
                                      int64_t *counters, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    ptg_status st = column_common(tree, n, nnodes, targets, nt, "ptg_column_rays");
    if (st) return st;
    if (nrays < 1 || !rays || !out || !err) {
        ptg_set_error("ptg_column_rays: needs nrays >= 1, rays, out and err");
        return PTG_ERR_ARG;
    }
    k_column_err_init<<<1, 1, 0, s>>>(err);
    PTG_LAUNCH_CHECK();
    const int64_t work = nt * nrays;
    if (work == 0) return PTG_OK;
    if (wt) {
        if (wt->prec != PTG_F64 || !wt->xyzm || !wt->aux || !wt->kids || wt->nslots != n + nnodes) {
            ptg_set_error("ptg_column_rays: wt must be the FP64 walk pack of the tree (or NULL)");
            return PTG_ERR_ARG;
        }
        const int64_t tasks = ptg_cdiv(nt, 64) * nrays;
        k_column_rays_wave<<<(unsigned)ptg_cdiv(tasks, CR_WAVES), 64 * CR_WAVES, 0, s>>>(
            (const D4 *)wt->xyzm, (const D4 *)wt->aux, (const int2 *)wt->kids, wt->nslots, targets, nt, rays, nrays,
            per_target != 0, order, out, err, (unsigned long long *)counters);
        PTG_LAUNCH_CHECK();
        return PTG_OK;
    }
    k_column_rays<<<(unsigned)ptg_cdiv(work, CD_THREADS), CD_THREADS, 0, s>>>(*tree, n, n + nnodes, targets, nt, rays,
                                                                              nrays, per_target != 0, out, err,
                                                                              (unsigned long long *)counters);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

/* Randomized ray grids (treewalk.py:1061-1078 with misc.py:6-54): target k sees the grid rotated by
 * the rotation drawn from a Mersenne-Twister stream seeded with k.  Four doubles are drawn, i.e. the
 * first eight 32-bit outputs of MT19937 after its integer seeding, so a thread needs only state
 * words 0..8 and 397..404 of the seeded table: the first twist combines word i, i+1 and i+397.
 * Each double is the 53-bit (a >> 5, b >> 6) combination of two outputs, as numpy's legacy
 * generator and numba's draw it; uniform(lo, hi) = lo + (hi - lo) * u. */
#define MT_M 397
#define MT_NEED 9 // words 0..8 of the table, and 397..404

__device__ __forceinline__ uint32_t mt_temper(uint32_t y)
{
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

__global__ void __launch_bounds__(256) k_random_rotations(int64_t first_seed, int64_t nt, const double *__restrict__ rays,
                                                          int nrays, double *__restrict__ out)
{
#pragma clang fp contract(off) // the host's rounding: no fused multiply-adds
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nt) return;
    uint32_t lo[MT_NEED], hi[MT_NEED - 1];
    uint32_t s = (uint32_t)(first_seed + k);
    // integer seeding: word i = 1812433253 * (word(i-1) ^ word(i-1) >> 30) + i
    for (int i = 0; i < MT_M + MT_NEED - 1; ++i) {
        if (i < MT_NEED) lo[i] = s;
        if (i >= MT_M) hi[i - MT_M] = s;
        s = 1812433253u * (s ^ (s >> 30)) + (uint32_t)(i + 1);
    }
    double u[4];
    for (int d = 0; d < 4; ++d) {
        uint32_t w[2];
        for (int q = 0; q < 2; ++q) {
            const int i = 2 * d + q;
            const uint32_t y = (lo[i] & 0x80000000u) | (lo[i + 1] & 0x7fffffffu);
            w[q] = mt_temper(hi[i] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u));
        }
        u[d] = ((double)(w[0] >> 5) * 67108864.0 + (double)(w[1] >> 6)) / 9007199254740992.0;
    }
    double R[3][3];
    for (int row = 0; row < 2; ++row) {
        const double c = -1.0 + 2.0 * u[2 * row];
        const double sn = sqrt(fmax(1.0 - c * c, 0.0));
        const double phi = 2.0 * M_PI * (0.0 + 1.0 * u[2 * row + 1]);
        R[row][0] = sn * cos(phi);
        R[row][1] = sn * sin(phi);
        R[row][2] = c;
    }
    double dot = 0.0;
    for (int c = 0; c < 3; ++c) dot += R[0][c] * R[1][c];
    for (int c = 0; c < 3; ++c) R[1][c] -= dot * R[0][c];
    double nrm = 0.0;
    for (int c = 0; c < 3; ++c) nrm += R[1][c] * R[1][c];
    nrm = sqrt(nrm);
    for (int c = 0; c < 3; ++c) R[1][c] /= nrm;
    for (int c = 0; c < 3; ++c) {
        const int a = (c + 1) % 3, b = (c + 2) % 3;
        R[2][c] = R[0][a] * R[1][b] - R[1][a] * R[0][b];
    }
    double *o = out + k * nrays * 3;
    for (int j = 0; j < nrays; ++j)
        for (int c = 0; c < 3; ++c)
            o[j * 3 + c] = rays[j * 3] * R[0][c] + rays[j * 3 + 1] * R[1][c] + rays[j * 3 + 2] * R[2][c];
}

extern "C" ptg_status ptg_random_rotations(int64_t first_seed, int64_t nt, const double *rays, int64_t nrays,
                                           double *out, void *stream)
{
    if (nt < 0 || nrays < 1 || nrays > (1 << 20) || !rays || (nt > 0 && !out) || first_seed < 0 ||
        first_seed + nt > (1ll << 32)) {
        ptg_set_error("ptg_random_rotations: needs 0 <= first_seed, first_seed + nt <= 2^32, 1 <= nrays <= 2^20, "
                      "rays and out");
        return PTG_ERR_ARG;
    }
    if (nt == 0) return PTG_OK;
    k_random_rotations<<<(unsigned)ptg_cdiv(nt, 256), 256, 0, (hipStream_t)stream>>>(first_seed, nt, rays, (int)nrays,
                                                                                     out);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

extern "C" ptg_status ptg_column_binned(const ptg_tree_arrays *tree, int64_t n, int64_t nnodes,
                                        const ptg_walk_tree *wt, const double *targets, int64_t nt,
                                        int64_t group_size, double theta, double *out, int32_t *err, void *ws,
                                        size_t ws_bytes, int64_t *counters, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    ptg_status st = column_common(tree, n, nnodes, targets, nt, "ptg_column_binned");
    if (st) return st;
    if (group_size < 1 || !(theta > 0) || !out || !err) {
        ptg_set_error("ptg_column_binned: needs group_size >= 1, theta > 0, out and err");
        return PTG_ERR_ARG;
    }
    k_column_err_init<<<1, 1, 0, s>>>(err);
    PTG_LAUNCH_CHECK();
    if (nt == 0) return PTG_OK;
    if (group_size <= 64) {
        if (!wt || wt->prec != PTG_F64 || !wt->xyzm || !wt->aux || !wt->kids || wt->nslots != n + nnodes || !ws ||
            ws_bytes < ptg_column_workspace_bytes(group_size) || n + nnodes >= (1ll << 28)) {
            ptg_set_error("ptg_column_binned: group_size <= 64 needs the FP64 walk pack of the tree (fewer than "
                          "2^28 elements) and ptg_column_workspace_bytes(group_size) of workspace");
            return PTG_ERR_ARG;
        }
        CbWs w;
        cb_carve(ws, group_size, &w);
        const int cus = cb_cus();
        k_column_ctl_init<<<1, 64, 0, s>>>(w.ctl);
        PTG_LAUNCH_CHECK();
        const int64_t ngroups = ptg_cdiv(nt, group_size);
        const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((int64_t)cus * 2, ptg_cdiv(ngroups, CB_WAVES)));
        const D4 *X = (const D4 *)wt->xyzm, *X2 = (const D4 *)wt->aux;
        const int2 *K = (const int2 *)wt->kids;
        k_column_binned_wave<false><<<(unsigned)blocks, 64 * CB_WAVES, 0, s>>>(
            X, X2, K, wt->nslots, targets, nt, (int)group_size, theta, out, w.spill, w.ctl, w.heavy, w.partial, err,
            (unsigned long long *)counters);
        k_column_binned_wave<true><<<(unsigned)cus * 2, 64 * CB_WAVES, 0, s>>>(
            X, X2, K, wt->nslots, targets, nt, (int)group_size, theta, out, w.spill, w.ctl, w.heavy, w.partial, err,
            (unsigned long long *)counters);
        k_column_heavy_sum<<<(unsigned)ptg_cdiv((int64_t)CB_HCAP * group_size, 256), 256, 0, s>>>(
            w.ctl, w.heavy, w.partial, nt, (int)group_size, out);
        PTG_LAUNCH_CHECK();
        return PTG_OK;
    }
    const int64_t ngroups = ptg_cdiv(nt, group_size);
    k_column_binned<<<(unsigned)ptg_cdiv(ngroups, CD_THREADS), CD_THREADS, 0, s>>>(*tree, n, n + nnodes, targets, nt,
                                                                                   group_size, theta, out, err,
                                                                                   (unsigned long long *)counters);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}
