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
// Angular-binned (ptg_column_binned): one thread per group of group_size consecutive targets, the
// reference's grouped estimator: a node whose distance to the group's box passes the opening test
// puts its mass in one of six equal-area sky bins per target; particles are summed per target.
#include <algorithm>

#include "ptg_common.hpp"

#define CD_THREADS 256

// a corrupt link (index out of range, or a chain longer than the element count) stops the walk and
// sets err[0]; the depth-first chain visits each element at most once
__device__ __forceinline__ bool cd_bad(int64_t no, int64_t ne, int64_t &visits)
{
    return no >= ne || ++visits > ne;
}

__global__ __launch_bounds__(CD_THREADS) void k_column_rays(const ptg_tree_arrays t, int64_t n, int64_t ne,
                                                            const double *tp, int64_t nt, const double *rays,
                                                            int64_t nr, int per_target, double *out, int32_t *err)
{
    const int64_t id = (int64_t)blockIdx.x * CD_THREADS + threadIdx.x;
    if (id >= nt * nr) return;
    const int64_t i = id / nt, j = id - (id / nt) * nt;  // ray i, target j
    const double *ray = per_target ? rays + 3 * (j * nr + i) : rays + 3 * i;
    const double rx = ray[0], ry = ray[1], rz = ray[2];
    const double px = tp[3 * j], py = tp[3 * j + 1], pz = tp[3 * j + 2];
    const double fac_density = 3.0 / (4.0 * 3.14159265358979323846);
    double column = 0.0;
    int64_t no = n, visits = 0;
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
    out[j * nr + i] = column;
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
                                                              double *out, int32_t *err)
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
    int64_t no = n, visits = 0;
    const int64_t stop = t.next_branch[n];
    while (no > -1 && no != stop) {
        if (cd_bad(no, ne, visits)) {
            atomicOr(err, 1);
            break;
        }
        const double cx = t.coordinates[3 * no], cy = t.coordinates[3 * no + 1], cz = t.coordinates[3 * no + 2];
        const double h_no = t.softenings[no], M = t.masses[no];
        if (no < n) {  // a particle contributes to every target of the group, no criterion
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
}

__global__ void k_column_err_init(int32_t *err) { *err = 0; }

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

extern "C" ptg_status ptg_column_rays(const ptg_tree_arrays *tree, int64_t n, int64_t nnodes, const double *targets,
                                      int64_t nt, const double *rays, int64_t nrays, int per_target, double *out,
                                      int32_t *err, void *stream)
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
    k_column_rays<<<(unsigned)ptg_cdiv(work, CD_THREADS), CD_THREADS, 0, s>>>(*tree, n, n + nnodes, targets, nt, rays,
                                                                              nrays, per_target != 0, out, err);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

extern "C" ptg_status ptg_column_binned(const ptg_tree_arrays *tree, int64_t n, int64_t nnodes, const double *targets,
                                        int64_t nt, int64_t group_size, double theta, double *out, int32_t *err,
                                        void *stream)
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
    const int64_t ngroups = ptg_cdiv(nt, group_size);
    k_column_binned<<<(unsigned)ptg_cdiv(ngroups, CD_THREADS), CD_THREADS, 0, s>>>(*tree, n, n + nnodes, targets, nt,
                                                                                   group_size, theta, out, err);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}
