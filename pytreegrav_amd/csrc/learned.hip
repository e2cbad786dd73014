// Learned opening criterion for gfx950.
//
// The geometric criterion (treewalk.py:22-25) opens a node unless r > size/theta + delta; the
// error-bounded criterion opens it unless 3.5 |Q| / r^4 < eps.  Both are per-node opening radii.
// Here a 10-32-32-1 MLP predicts each node's opening radius for a given absolute acceleration
// tolerance eps from scale-free features of the node (quadrupole norm, COM offset, children's
// mass distribution, softening, and the tolerance itself).  It is trained on per-node labels made
// by direct summation: for every sampled node, the worst monopole error over 32 directions at a
// ladder of distances gives the smallest distance beyond which the node's monopole is within eps.
//
// Kernels: features, error curves (labels, FP64 direct sums over the node's particles), training
// samples, minibatch forward + backward (deterministic block partials), Adam, batched inference,
// and the fused radius pass that writes the predicted radius of every node into the FP32 walk
// pack (aux[3]), where the walk's PTG_CRIT_RADIUS test reads it.
#include "ptg_common.hpp"

#define MLP_IN 10
#define MLP_H 32
#define OFF_W1 0                          // W1[in][h]
#define OFF_B1 (OFF_W1 + MLP_IN * MLP_H)  // b1[h]
#define OFF_W2 (OFF_B1 + MLP_H)           // W2[h_in][h_out]
#define OFF_B2 (OFF_W2 + MLP_H * MLP_H)   // b2[h]
#define OFF_W3 (OFF_B2 + MLP_H)           // W3[h]
#define OFF_B3 (OFF_W3 + MLP_H)           // b3
#define MLP_NP (OFF_B3 + 1)

static_assert(MLP_NP == PTG_MLP_NPARAM, "parameter layout");

static dim3 lgrid(int64_t n, int bs = 256) { return dim3((unsigned)std::max<int64_t>(1, ptg_cdiv(n, bs))); }

// ============================================================================================
// node features (scale free; the same function serves training and inference)
// ============================================================================================
__device__ __forceinline__ double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

// children of node e through the reference links (first_subnode / next_branch), at most 8
__device__ __forceinline__ int node_kids(const ptg_tree_arrays &t, int64_t e, int64_t ne, int32_t *kids)
{
    int32_t c = t.first_subnode[e], stop = t.next_branch[e];
    int nk = 0;
    while (c != stop && c >= 0 && c < ne && nk < 8) {
        kids[nk++] = c;
        c = t.next_branch[c];
    }
    return nk;
}

// Features of node e from its quadrupole norm qn, child count nk and child statistics: the largest
// child mass mmax, sum m d^2, sum m d^3 and the largest d over its children (d = child centre to the
// node's centre of mass).
__device__ void features_from(const ptg_tree_arrays &t, int64_t e, double eps, double s, double M, double qn, int nk,
                              double mmax, double m2, double m3, double dmax, float *f)
{
    // log10(r_Q / s) = log10(3.5 |Q| / (eps s^4)) / 4, r_Q = (3.5 |Q| / eps)^(1/4); the ratios are
    // formed in FP64 and clamped before the FP32 logarithms
    const double s2 = s * s;
    const float lq = log10f((float)clampd(3.5 * qn / (eps * s2 * s2), 1e-12, 1e12));
    f[0] = fminf(fmaxf(0.25f * lq, -2.f), 3.f) * 0.5f;
    f[1] = (float)clampd(t.deltas[e] / s, 0.0, 2.0);
    f[2] = (float)clampd(qn / (M * s * s), 0.0, 2.0);
    f[3] = (float)clampd(t.softenings[e] / s, 0.0, 4.0);
    f[4] = (float)nk * 0.125f;
    f[5] = (float)(mmax / M);
    f[6] = (float)clampd(m2 / (M * s * s), 0.0, 2.0);
    f[7] = (float)clampd(m3 / (M * s * s * s), 0.0, 2.0);
    f[8] = (float)clampd(dmax / s, 0.0, 2.0);
    f[9] = log10f((float)clampd(M / (eps * s2), 1e-4, 1e12)) * 0.25f;
}

// Returns false for a massless or zero-size node (its radius is then 0: the geometric test decides).
__device__ bool node_features(const ptg_tree_arrays &t, int64_t ne, int64_t e, double eps, float *f)
{
    const double s = t.sizes[e], M = t.masses[e];
    if (!(s > 0.0) || !(M > 0.0)) {
        for (int k = 0; k < MLP_IN; k++) f[k] = 0.f;
        return false;
    }
    const double cx = t.coordinates[3 * e], cy = t.coordinates[3 * e + 1], cz = t.coordinates[3 * e + 2];
    double qn = 0.0;
    if (t.quadrupoles) {
        const double *Q = t.quadrupoles + 9 * e;
        double s2 = 0.0;
        for (int k = 0; k < 9; k++) s2 += Q[k] * Q[k];
        qn = sqrt(s2);
    }
    int32_t kids[8];
    const int nk = node_kids(t, e, ne, kids);
    double mmax = 0.0, m2 = 0.0, m3 = 0.0, dmax = 0.0;
    for (int q = 0; q < nk; q++) {
        const int64_t c = kids[q];
        const double mc = t.masses[c];
        const double dx = t.coordinates[3 * c] - cx, dy = t.coordinates[3 * c + 1] - cy,
                     dz = t.coordinates[3 * c + 2] - cz;
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        mmax = fmax(mmax, mc);
        m2 += mc * d * d;
        m3 += mc * d * d * d;
        dmax = fmax(dmax, d);
    }
    features_from(t, e, eps, s, M, qn, nk, mmax, m2, m3, dmax, f);
    return true;
}

// The same features from the statistics the build stored per node (no walk over the links).
__device__ bool node_features_stats(const ptg_tree_arrays &t, int64_t n, int64_t e, double eps, float *f)
{
    const double s = t.sizes[e], M = t.masses[e];
    if (!(s > 0.0) || !(M > 0.0)) {
        for (int k = 0; k < MLP_IN; k++) f[k] = 0.f;
        return false;
    }
    const double *st = t.node_stats + 5 * (e - n);
    features_from(t, e, eps, s, M, st[0], t.nchild[e], st[1], st[2], st[3], st[4], f);
    return true;
}

// ============================================================================================
// MLP forward.  The hidden activation is tanh(z) = 1 - 2 / (exp(2z) + 1) with the hardware exp2 and
// reciprocal (absolute error ~1e-7; saturates correctly for large |z|); training and inference use
// the same function.
// ============================================================================================
__device__ __forceinline__ float act(float z)
{
    const float e = __builtin_amdgcn_exp2f(z * 2.8853900817779268f);  // exp(2z) = 2^(2z log2 e)
    return 1.f - 2.f * __builtin_amdgcn_rcpf(e + 1.f);
}

// FP32 forward with the parameters read straight from global memory: every lane reads the same
// address, so the compiler turns the fully unrolled, constant-offset loads into scalar loads and
// the FMAs take their weight from an SGPR.  Loops run input-major over the [in][out] rows.
__device__ __forceinline__ float mlp_forward_f32(const float *__restrict__ P, const float *x)
{
    float z[MLP_H], h[MLP_H];
#pragma unroll
    for (int j = 0; j < MLP_H; j++) z[j] = P[OFF_B1 + j];
#pragma unroll
    for (int k = 0; k < MLP_IN; k++) {
        const float xk = x[k];
#pragma unroll
        for (int j = 0; j < MLP_H; j++) z[j] = fmaf(xk, P[OFF_W1 + k * MLP_H + j], z[j]);
    }
#pragma unroll
    for (int j = 0; j < MLP_H; j++) {
        h[j] = act(z[j]);
        z[j] = P[OFF_B2 + j];
    }
#pragma unroll
    for (int i = 0; i < MLP_H; i++) {
        const float hi = h[i];
#pragma unroll
        for (int j = 0; j < MLP_H; j++) z[j] = fmaf(hi, P[OFF_W2 + i * MLP_H + j], z[j]);
    }
    float y = P[OFF_B3];
#pragma unroll
    for (int j = 0; j < MLP_H; j++) y = fmaf(act(z[j]), P[OFF_W3 + j], y);
    return y;
}

// The same FP32 forward for the 64 nodes of a wavefront (f: this lane's node's features), layers 1
// and 2 as 16x16x4 FP32 matrix products on the matrix cores: rows are the wave's nodes (four row
// tiles), columns the hidden units (two column tiles).  In a 16x16x4 product lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16] and receives D[4 (l / 16) + q][l % 16] in its q-th
// register.  Layer 2 hands lane group g = l / 16 the inputs 8g .. 8g+7 over its eight steps (the sum
// over the 32 inputs is the same), so each lane reads its A values as two contiguous runs.  The
// activations go through LDS (xs: 12 floats per node, hs: MLP_HS per node); the activation and
// layer 3 are those of mlp_forward_f32, in its order.  xs may alias hs.  Every lane of the wave must
// call it.
#define MLP_HS 33
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ float mlp_forward_f32_wave(const float *__restrict__ P, const float *f, float *xs, float *hs)
{
    const int lane = ptg_lane(), g = lane >> 4, c16 = lane & 15;
#pragma unroll
    for (int k = 0; k < 12; k++) xs[lane * 12 + k] = k < MLP_IN ? f[k] : 0.f;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    floatx4 acc[4][2];
    // layer 1: z = b1 + x W1 over the 10 inputs (padded to 12, three steps of 4)
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const float b = P[OFF_B1 + 16 * c + c16];
#pragma unroll
        for (int r = 0; r < 4; r++) acc[r][c] = floatx4{b, b, b, b};
    }
#pragma unroll
    for (int st = 0; st < 3; st++) {
        const int k = 4 * st + g;
        const float w0 = k < MLP_IN ? P[OFF_W1 + k * MLP_H + c16] : 0.f;
        const float w1 = k < MLP_IN ? P[OFF_W1 + k * MLP_H + 16 + c16] : 0.f;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float a = xs[(16 * r + c16) * 12 + k];
            acc[r][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, w0, acc[r][0], 0, 0, 0);
            acc[r][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, w1, acc[r][1], 0, 0, 0);
        }
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) hs[(16 * r + 4 * g + q) * MLP_HS + 16 * c + c16] = act(acc[r][c][q]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // layer 2: z = b2 + h W2 over the 32 hidden inputs
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const float b = P[OFF_B2 + 16 * c + c16];
#pragma unroll
        for (int r = 0; r < 4; r++) acc[r][c] = floatx4{b, b, b, b};
    }
#pragma unroll
    for (int st = 0; st < 8; st++) {
        const int k = 8 * g + st;
        const float w0 = P[OFF_W2 + k * MLP_H + c16], w1 = P[OFF_W2 + k * MLP_H + 16 + c16];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float a = hs[(16 * r + c16) * MLP_HS + k];
            acc[r][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, w0, acc[r][0], 0, 0, 0);
            acc[r][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, w1, acc[r][1], 0, 0, 0);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();  // every lane has read layer 1's activations
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) hs[(16 * r + 4 * g + q) * MLP_HS + 16 * c + c16] = act(acc[r][c][q]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // layer 3 for this lane's node, in mlp_forward_f32's order
    float y = P[OFF_B3];
#pragma unroll
    for (int j = 0; j < MLP_H; j++) y = fmaf(hs[lane * MLP_HS + j], P[OFF_W3 + j], y);
    return y;
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// BF16 weights, transposed so that each output unit's inputs are pairs: W1T[h][in/2], W2T[h][h/2];
// layer sums accumulate in FP32 through v_dot2_f32_bf16
struct Bf16Params {
    bf16x2 w1t[MLP_H][MLP_IN / 2];
    bf16x2 w2t[MLP_H][MLP_H / 2];
    float b1[MLP_H], b2[MLP_H], w3[MLP_H], b3;
};

__device__ __forceinline__ float mlp_forward_bf16(const Bf16Params &Q, const float *x)
{
    bf16x2 xp[MLP_IN / 2];
#pragma unroll
    for (int k = 0; k < MLP_IN / 2; k++) xp[k] = bf16x2{(__bf16)x[2 * k], (__bf16)x[2 * k + 1]};
    bf16x2 hp[MLP_H / 2];
#pragma unroll
    for (int j = 0; j < MLP_H; j += 2) {
        float z0 = Q.b1[j], z1 = Q.b1[j + 1];
#pragma unroll
        for (int k = 0; k < MLP_IN / 2; k++) {
            z0 = __builtin_amdgcn_fdot2_f32_bf16(xp[k], Q.w1t[j][k], z0, false);
            z1 = __builtin_amdgcn_fdot2_f32_bf16(xp[k], Q.w1t[j + 1][k], z1, false);
        }
        hp[j / 2] = bf16x2{(__bf16)act(z0), (__bf16)act(z1)};
    }
    float y = Q.b3;
    for (int j = 0; j < MLP_H; j++) {
        float z = Q.b2[j];
#pragma unroll
        for (int i = 0; i < MLP_H / 2; i++) z = __builtin_amdgcn_fdot2_f32_bf16(hp[i], Q.w2t[j][i], z, false);
        y = fmaf(act(z), Q.w3[j], y);
    }
    return y;
}

__device__ void load_params_f32(const float *P, float *sP)
{
    for (int i = threadIdx.x; i < MLP_NP; i += blockDim.x) sP[i] = P[i];
}

__device__ void load_params_bf16(const float *P, Bf16Params &Q)
{
    for (int i = threadIdx.x; i < MLP_H * MLP_IN; i += blockDim.x) {
        const int j = i / MLP_IN, k = i % MLP_IN;  // unit j, input k
        __bf16 v = (__bf16)P[OFF_W1 + k * MLP_H + j];
        if (k & 1) Q.w1t[j][k / 2].y = v;
        else Q.w1t[j][k / 2].x = v;
    }
    for (int i = threadIdx.x; i < MLP_H * MLP_H; i += blockDim.x) {
        const int j = i / MLP_H, k = i % MLP_H;  // output unit j, input unit k
        __bf16 v = (__bf16)P[OFF_W2 + k * MLP_H + j];
        if (k & 1) Q.w2t[j][k / 2].y = v;
        else Q.w2t[j][k / 2].x = v;
    }
    for (int j = threadIdx.x; j < MLP_H; j += blockDim.x) {
        Q.b1[j] = P[OFF_B1 + j];
        Q.b2[j] = P[OFF_B2 + j];
        Q.w3[j] = P[OFF_W3 + j];
    }
    if (threadIdx.x == 0) Q.b3 = P[OFF_B3];
}

// ============================================================================================
// labels: monopole error curves by direct summation over each node's particles
// ============================================================================================
#define EC_DIRS 32
#define EC_THREADS 256
#define EC_MAXR 64
#define EC_TPT ((EC_DIRS * EC_MAXR + EC_THREADS - 1) / EC_THREADS)  // targets per thread

// Particles of node e: Morton slots [lo, hi).  lo: the leftmost leaf below e; hi: the first leaf
// after e's subtree (next_branch, then leftmost leaf), or the end of e's tree.
__device__ bool node_particle_range(const ptg_tree_arrays &t, int64_t n, int64_t ne, int64_t e, int64_t tree_end,
                                    int64_t *lo, int64_t *hi)
{
    int64_t c = e;
    for (int s = 0; s < 512 && c >= n; s++) {
        const int32_t f = t.first_subnode[c];
        if (f < 0 || f >= ne) return false;
        c = f;
    }
    if (c >= n) return false;
    *lo = c;
    int64_t b = t.next_branch[e];
    if (b >= ne) return false;
    if (b < 0) {
        *hi = tree_end;
        return true;
    }
    for (int s = 0; s < 512 && b >= n; s++) {
        const int32_t f = t.first_subnode[b];
        if (f < 0 || f >= ne) return false;
        b = f;
    }
    if (b >= n) return false;
    *hi = b;
    return *hi > *lo;
}

__device__ __forceinline__ void fib_dir(int d, double *u)
{
    const double z = 1.0 - (2.0 * d + 1.0) / EC_DIRS;
    const double r = sqrt(fmax(0.0, 1.0 - z * z));
    const double phi = 2.399963229728653 * d;  // golden angle
    u[0] = r * cos(phi);
    u[1] = r * sin(phi);
    u[2] = z;
}

struct EcForest {
    int64_t nsys;
    const int64_t *soff;
    const int32_t *roots;
};

// One block per node.  err[row * nr + j] = max over EC_DIRS directions of |a_exact - a_monopole| at
// distance rho_j * size from the COM, in units of M / size^2 (G = 1, no softening: the walk's guard
// keeps accepted nodes beyond every softening length); -1 for rows that are skipped (more than cap
// particles, massless or malformed nodes).
__global__ __launch_bounds__(EC_THREADS) void k_errcurve(const ptg_tree_arrays t, int64_t n, int64_t ne, EcForest fo,
                                                         const int32_t *nodes, int64_t count, int64_t cap, int nr,
                                                         double rho_lo, double rho_hi, int stat, double *err)
{
    __shared__ double px[EC_THREADS], py[EC_THREADS], pz[EC_THREADS], pm[EC_THREADS];
    __shared__ double emax[EC_DIRS * EC_MAXR];
    __shared__ int64_t range[2];
    __shared__ int ok;
    const int64_t row = blockIdx.x;  // one block per row: the grid is sized to count
    if (row >= count) return;
    const int64_t e = nodes[row];
    if (threadIdx.x == 0) {
        ok = 0;
        if (e >= n && e < ne) {
            int64_t tree_end = n;
            if (fo.nsys > 1) {
                int64_t lo = 0, hi = fo.nsys;
                while (hi - lo > 1) {
                    const int64_t mid = (lo + hi) >> 1;
                    if (fo.roots[mid] <= e) lo = mid;
                    else hi = mid;
                }
                tree_end = fo.soff[lo + 1];
            }
            int64_t a, b;
            if (node_particle_range(t, n, ne, e, tree_end, &a, &b) && b - a <= cap && t.masses[e] > 0.0 &&
                t.sizes[e] > 0.0) {
                range[0] = a;
                range[1] = b;
                ok = 1;
            }
        }
    }
    __syncthreads();
    if (!ok) {
        for (int j = threadIdx.x; j < nr; j += blockDim.x) err[row * nr + j] = -1.0;
        return;
    }
    const int64_t a = range[0], b = range[1];
    const double s = t.sizes[e], M = t.masses[e];
    const double c[3] = {t.coordinates[3 * e], t.coordinates[3 * e + 1], t.coordinates[3 * e + 2]};
    const int T = EC_DIRS * nr;
    double tx[EC_TPT], ty[EC_TPT], tz[EC_TPT], ax[EC_TPT], ay[EC_TPT], az[EC_TPT];
    const double lr = nr > 1 ? log(rho_hi / rho_lo) / (nr - 1) : 0.0;
#pragma unroll
    for (int k = 0; k < EC_TPT; k++) {
        const int q = threadIdx.x + k * EC_THREADS;
        double u[3] = {0, 0, 0};
        double r = 0;
        if (q < T) {
            fib_dir(q % EC_DIRS, u);
            r = rho_lo * exp(lr * (q / EC_DIRS)) * s;
        }
        tx[k] = c[0] + r * u[0];
        ty[k] = c[1] + r * u[1];
        tz[k] = c[2] + r * u[2];
        ax[k] = ay[k] = az[k] = 0.0;
    }
    for (int64_t p0 = a; p0 < b; p0 += EC_THREADS) {
        const int64_t p = p0 + threadIdx.x;
        __syncthreads();
        if (p < b) {
            px[threadIdx.x] = t.coordinates[3 * p];
            py[threadIdx.x] = t.coordinates[3 * p + 1];
            pz[threadIdx.x] = t.coordinates[3 * p + 2];
            pm[threadIdx.x] = t.masses[p];
        }
        __syncthreads();
        const int cnt = (int)(b - p0 < EC_THREADS ? b - p0 : EC_THREADS);
        for (int i = 0; i < cnt; i++) {
            const double xi = px[i], yi = py[i], zi = pz[i], mi = pm[i];
#pragma unroll
            for (int k = 0; k < EC_TPT; k++) {
                const double dx = xi - tx[k], dy = yi - ty[k], dz = zi - tz[k];
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > 0.0) {
                    const double ir = 1.0 / sqrt(r2);
                    const double w = mi * ir * ir * ir;
                    ax[k] += w * dx;
                    ay[k] += w * dy;
                    az[k] += w * dz;
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < EC_TPT; k++) {
        const int q = threadIdx.x + k * EC_THREADS;
        if (q >= T) continue;
        const double dx = c[0] - tx[k], dy = c[1] - ty[k], dz = c[2] - tz[k];
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double ir = 1.0 / sqrt(r2);
        const double w = M * ir * ir * ir;
        const double ex = ax[k] - w * dx, ey = ay[k] - w * dy, ez = az[k] - w * dz;
        emax[q] = sqrt(ex * ex + ey * ey + ez * ez) * s * s / M;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < nr; j += blockDim.x) {
        double m = 0.0;
        if (stat == 0) {
            for (int d = 0; d < EC_DIRS; d++) m = fmax(m, emax[j * EC_DIRS + d]);
        } else {
            for (int d = 0; d < EC_DIRS; d++) m += emax[j * EC_DIRS + d] * emax[j * EC_DIRS + d];
            m = sqrt(m / EC_DIRS);
        }
        err[row * nr + j] = m;
    }
}

// Training samples: sample i pairs node row[i] of the curves with tolerance eps[i].  Writes the
// features x[i][10] and the label y[i] = log10(rho*), rho* the smallest ladder distance (in node
// sizes, log-interpolated) beyond which every ladder point's error is within eps; valid[i] = 0 for
// skipped rows, 2 when the error is still above eps at the last ladder point (label clipped).
__global__ __launch_bounds__(256) void k_samples(const ptg_tree_arrays t, int64_t ne, const int32_t *nodes,
                                                 const int32_t *row, const double *eps, int64_t count,
                                                 const double *err, int nr, double rho_lo, double rho_hi, float *x,
                                                 float *y, int32_t *valid)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int64_t r = row[i];
    const int64_t e = nodes[r];
    float f[MLP_IN] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const bool good = e >= 0 && e < ne && node_features(t, ne, e, eps[i], f);
    for (int k = 0; k < MLP_IN; k++) x[i * MLP_IN + k] = f[k];
    const double *ec = err + r * nr;
    if (!good || ec[0] < 0.0) {
        y[i] = 0.f;
        valid[i] = 0;
        return;
    }
    const double s = t.sizes[e], M = t.masses[e];
    const double unit = M / (s * s);
    const double lr = nr > 1 ? log(rho_hi / rho_lo) / (nr - 1) : 0.0;
    // last ladder index whose error exceeds eps
    int last = -1;
    for (int j = 0; j < nr; j++)
        if (ec[j] * unit > eps[i]) last = j;
    double lrho;
    int v = 1;
    if (last < 0) {
        lrho = log(rho_lo);
    } else if (last == nr - 1) {
        lrho = log(rho_hi);
        v = 2;
    } else {
        // interpolate log(err) linearly in log(rho) between ladder points last and last + 1
        const double e0 = log(ec[last] * unit), e1 = log(fmax(ec[last + 1] * unit, 1e-300));
        const double le = log(eps[i]);
        const double fr = e0 > e1 ? clampd((e0 - le) / (e0 - e1), 0.0, 1.0) : 1.0;
        lrho = log(rho_lo) + lr * (last + fr);
    }
    y[i] = (float)(lrho / log(10.0));
    valid[i] = v;
}

// ============================================================================================
// training: minibatch forward + backward with deterministic per-block partial gradients
// ============================================================================================
#define TR_B 64
struct TrainLds {
    float x[TR_B][MLP_IN + 1];
    float h1[TR_B][MLP_H + 1];
    float h2[TR_B][MLP_H + 1];
    float dz1[TR_B][MLP_H + 1];
    float dz2[TR_B][MLP_H + 1];
    float g[TR_B];
    float P[MLP_NP];
};

__device__ __forceinline__ uint64_t splitmix(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Loss over a batch of `batch` samples: (1/batch) sum w_i (yhat_i - y_i)^2 with w_i = under for
// yhat_i < y_i (a predicted radius below the true one accepts nodes too early), else 1.  Sample
// b of the batch is index[b] when index is given, else splitmix(seed, b) mod nsamp.
// part[block][MLP_NP + 1]: the block's gradient and loss contributions.
__global__ __launch_bounds__(TR_B) void k_mlp_grad(const float *P, const float *X, const float *Y, int64_t nsamp,
                                                   const int32_t *index, uint64_t seed, int64_t batch, float under,
                                                   float *part)
{
    __shared__ TrainLds S;
    load_params_f32(P, S.P);
    __syncthreads();
    const int b = threadIdx.x;
    const int64_t gb = (int64_t)blockIdx.x * TR_B + b;
    const bool live = gb < batch;
    int64_t si = 0;
    if (live) si = index ? (int64_t)index[gb] : (int64_t)(splitmix(seed * 0x100000001b3ull + (uint64_t)gb) % (uint64_t)nsamp);
    float x[MLP_IN];
    for (int k = 0; k < MLP_IN; k++) {
        x[k] = live ? X[si * MLP_IN + k] : 0.f;
        S.x[b][k] = x[k];
    }
    float h1[MLP_H];
    for (int j = 0; j < MLP_H; j++) {
        float z = S.P[OFF_B1 + j];
        for (int k = 0; k < MLP_IN; k++) z = fmaf(x[k], S.P[OFF_W1 + k * MLP_H + j], z);
        h1[j] = act(z);
        S.h1[b][j] = h1[j];
    }
    float yhat = S.P[OFF_B3];
    for (int j = 0; j < MLP_H; j++) {
        float z = S.P[OFF_B2 + j];
        for (int i = 0; i < MLP_H; i++) z = fmaf(h1[i], S.P[OFF_W2 + i * MLP_H + j], z);
        const float h2 = act(z);
        S.h2[b][j] = h2;
        yhat = fmaf(h2, S.P[OFF_W3 + j], yhat);
    }
    const float res = live ? yhat - Y[si] : 0.f;
    const float w = res < 0.f ? under : 1.f;
    const float loss = w * res * res / (float)batch;
    const float g = 2.f * w * res / (float)batch;  // dL/dyhat
    S.g[b] = g;
    for (int j = 0; j < MLP_H; j++) {
        const float h2 = S.h2[b][j];
        S.dz2[b][j] = g * S.P[OFF_W3 + j] * (1.f - h2 * h2);
    }
    for (int i = 0; i < MLP_H; i++) {
        float d = 0.f;
        for (int j = 0; j < MLP_H; j++) d = fmaf(S.P[OFF_W2 + i * MLP_H + j], S.dz2[b][j], d);
        S.dz1[b][i] = d * (1.f - h1[i] * h1[i]);
    }
    __syncthreads();
    // block sums over samples, one parameter per thread per pass, fixed order
    float *out = part + (int64_t)blockIdx.x * (MLP_NP + 1);
    for (int p = threadIdx.x; p < MLP_NP + 1; p += TR_B) {
        float acc = 0.f;
        if (p < OFF_B1) {
            const int k = p / MLP_H, j = p % MLP_H;
            for (int q = 0; q < TR_B; q++) acc = fmaf(S.x[q][k], S.dz1[q][j], acc);
        } else if (p < OFF_W2) {
            const int j = p - OFF_B1;
            for (int q = 0; q < TR_B; q++) acc += S.dz1[q][j];
        } else if (p < OFF_B2) {
            const int i = (p - OFF_W2) / MLP_H, j = (p - OFF_W2) % MLP_H;
            for (int q = 0; q < TR_B; q++) acc = fmaf(S.h1[q][i], S.dz2[q][j], acc);
        } else if (p < OFF_W3) {
            const int j = p - OFF_B2;
            for (int q = 0; q < TR_B; q++) acc += S.dz2[q][j];
        } else if (p < OFF_B3) {
            const int j = p - OFF_W3;
            for (int q = 0; q < TR_B; q++) acc = fmaf(S.h2[q][j], S.g[q], acc);
        } else if (p == OFF_B3) {
            for (int q = 0; q < TR_B; q++) acc += S.g[q];
        }
        out[p] = acc;
    }
    // loss: reuse the g column as scratch after everyone has read it
    __syncthreads();
    S.g[b] = loss;
    __syncthreads();
    if (threadIdx.x == 0) {
        float l = 0.f;
        for (int q = 0; q < TR_B; q++) l += S.g[q];
        out[MLP_NP] = l;
    }
}

// grad[p] = sum over blocks of part[block][p] (p = MLP_NP is the loss), in block order
__global__ __launch_bounds__(256) void k_reduce_parts(const float *part, int64_t nblocks, float *grad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > MLP_NP) return;
    float acc = 0.f;
    for (int64_t q = 0; q < nblocks; q++) acc += part[q * (MLP_NP + 1) + p];
    grad[p] = acc;
}

__global__ __launch_bounds__(256) void k_adam(float *P, float *m, float *v, const float *grad, int64_t step, float lr,
                                              float b1, float b2, float eps)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NP) return;
    const float g = grad[p];
    const float mp = b1 * m[p] + (1.f - b1) * g;
    const float vp = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mp;
    v[p] = vp;
    const float mh = mp / (1.f - powf(b1, (float)step));
    const float vh = vp / (1.f - powf(b2, (float)step));
    P[p] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ P, const float *X, int64_t count,
                                                     int bf16, float *Y)
{
    __shared__ Bf16Params Q;
    if (bf16) {
        load_params_bf16(P, Q);
        __syncthreads();
    }
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float x[MLP_IN];
    for (int k = 0; k < MLP_IN; k++) x[k] = X[i * MLP_IN + k];
    Y[i] = bf16 ? mlp_forward_bf16(Q, x) : mlp_forward_f32(P, x);
}

__global__ __launch_bounds__(256) void k_node_features(const ptg_tree_arrays t, int64_t ne, const int32_t *nodes,
                                                       int64_t count, double eps, float *X)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float f[MLP_IN] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int64_t e = nodes[i];
    if (e >= 0 && e < ne) node_features(t, ne, e, eps, f);
    for (int k = 0; k < MLP_IN; k++) X[i * MLP_IN + k] = f[k];
}

// Fused radius pass: features + MLP for every node, radius = size * 10^(yhat + margin) into the
// FP32 pack's aux[3] at the node's walk slot (0 for massless nodes and for particles).
__global__ __launch_bounds__(256) void k_open_radii(const ptg_tree_arrays t, int64_t n, int64_t nnodes, float *aux,
                                                    int64_t nslots, const float *__restrict__ P, float margin,
                                                    double eps, const double *scale, EcForest fo, int bf16,
                                                    float rho_max)
{
    __shared__ Bf16Params Q;
    // per wavefront: the features (64 x 12) then, over them, the activations (64 x MLP_HS); a wave's
    // layer-1 products have read every feature before the first activation is stored
    __shared__ float hs[4][64 * MLP_HS];
    if (bf16) {
        load_params_bf16(P, Q);
        __syncthreads();
    }
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // every lane reaches the FP32 forward (a wave-wide product); lanes without a node carry zeros
    bool valid = a < ptg_node_count(nnodes, t.node_count);  // a deferred tree may hold fewer nodes
    const int64_t e = n + a;
    const int64_t sl = valid ? t.slot[e] : -1;
    valid = valid && sl >= 0 && sl < nslots;
    float f[MLP_IN];
    bool has = false;
    if (valid) {
        if (scale) {  // one scale, or one per system of a forest
            int64_t sy = 0;
            if (fo.nsys > 1) {
                int64_t lo = 0, hi = fo.nsys;
                while (hi - lo > 1) {
                    const int64_t mid = (lo + hi) >> 1;
                    if (fo.roots[mid] <= e) lo = mid;
                    else hi = mid;
                }
                sy = lo;
            }
            eps *= scale[sy];
        }
        has = t.node_stats ? node_features_stats(t, n, e, eps, f) : node_features(t, n + nnodes, e, eps, f);
    } else {
#pragma unroll
        for (int k = 0; k < MLP_IN; k++) f[k] = 0.f;
    }
    float yhat;
    if (bf16) {
        yhat = mlp_forward_bf16(Q, f);
    } else {
        const int w = threadIdx.x >> 6;
        yhat = mlp_forward_f32_wave(P, f, hs[w], hs[w]);
    }
    if (!valid) return;
    float radius = 0.f;
    if (has) {
        const float rho = fminf(exp10f(yhat + margin), rho_max);
        radius = (float)(t.sizes[e] * (double)rho);
    }
    aux[4 * sl + 3] = radius;
}


// root mean square over rows of the vector norm of columns [col, col + ncols) of an (rows, stride)
// array; block b reduces segment b of rows / gridDim.x consecutive rows; fixed summation order
__global__ __launch_bounds__(256) void k_rms_rows(const double *x, int64_t rows, int stride, int col, int ncols,
                                                  double *out)
{
    __shared__ double sh[256];
    x += (int64_t)blockIdx.x * rows * stride;
    out += blockIdx.x;
    double acc = 0.0;
    for (int64_t r = threadIdx.x; r < rows; r += blockDim.x)
        for (int c = 0; c < ncols; c++) {
            const double v = x[r * stride + col + c];
            acc += v * v;
        }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = rows > 0 ? sqrt(sh[0] / (double)rows) : 0.0;
}

// ============================================================================================
// C ABI
// ============================================================================================
extern "C" ptg_status ptg_rms_rows(const double *x, int64_t rows, int stride, int col, int ncols, int64_t nseg,
                                   double *out, void *stream)
{
    if (!x || !out || rows < 0 || stride < 1 || col < 0 || ncols < 1 || col + ncols > stride || nseg < 1 ||
        nseg > 0x7fffffff || rows % nseg) {
        ptg_set_error("ptg_rms_rows: bad arguments (rows must be a multiple of nseg)");
        return PTG_ERR_ARG;
    }
    k_rms_rows<<<(unsigned)nseg, 256, 0, (hipStream_t)stream>>>(x, rows / nseg, stride, col, ncols, out);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

extern "C" ptg_status ptg_node_features(const ptg_tree_arrays *tree, int64_t n, int64_t nnodes, const int32_t *nodes,
                                        int64_t count, double eps, float *features, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (!tree || n < 1 || nnodes < 1 || count < 0 || (count > 0 && (!nodes || !features)) || !(eps > 0)) {
        ptg_set_error("ptg_node_features: bad arguments");
        return PTG_ERR_ARG;
    }
    if (count == 0) return PTG_OK;
    k_node_features<<<lgrid(count), 256, 0, s>>>(*tree, n + nnodes, nodes, count, eps, features);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

extern "C" ptg_status ptg_node_errcurve(const ptg_tree_arrays *tree, int64_t n, int64_t nnodes,
                                        const ptg_forest *forest, const int32_t *nodes, int64_t count, int64_t cap,
                                        int nradii, double rho_lo, double rho_hi, int stat, double *err, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (!tree || n < 1 || nnodes < 1 || count < 0 || (count > 0 && (!nodes || !err)) || nradii < 2 ||
        nradii > EC_MAXR || !(rho_lo > 0) || !(rho_hi > rho_lo) || cap < 1 || (stat != 0 && stat != 1)) {
        ptg_set_error("ptg_node_errcurve: bad arguments (2 <= nradii <= %d, 0 < rho_lo < rho_hi, cap >= 1, "
                      "stat 0 or 1)", EC_MAXR);
        return PTG_ERR_ARG;
    }
    if (forest && (forest->nsys < 1 || !forest->offsets || !forest->roots)) {
        ptg_set_error("ptg_node_errcurve: bad forest");
        return PTG_ERR_ARG;
    }
    if (count == 0) return PTG_OK;
    EcForest fo{forest ? forest->nsys : 1, forest ? forest->offsets : nullptr, forest ? forest->roots : nullptr};
    for (int64_t c0 = 0; c0 < count; c0 += 65535) {
        const int64_t c = std::min<int64_t>(65535, count - c0);
        k_errcurve<<<(unsigned)c, EC_THREADS, 0, s>>>(*tree, n, n + nnodes, fo, nodes + c0, c, cap, nradii, rho_lo,
                                                      rho_hi, stat, err + c0 * nradii);
        PTG_LAUNCH_CHECK();
    }
    return PTG_OK;
}

extern "C" ptg_status ptg_mlp_samples(const ptg_tree_arrays *tree, int64_t n, int64_t nnodes, const int32_t *nodes,
                                      const int32_t *rows, const double *eps, int64_t count, const double *err,
                                      int nradii, double rho_lo, double rho_hi, float *x, float *y, int32_t *valid,
                                      void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (!tree || n < 1 || nnodes < 1 || count < 0 ||
        (count > 0 && (!nodes || !rows || !eps || !err || !x || !y || !valid)) || nradii < 2 || nradii > EC_MAXR) {
        ptg_set_error("ptg_mlp_samples: bad arguments");
        return PTG_ERR_ARG;
    }
    if (count == 0) return PTG_OK;
    k_samples<<<lgrid(count), 256, 0, s>>>(*tree, n + nnodes, nodes, rows, eps, count, err, nradii, rho_lo, rho_hi, x,
                                          y, valid);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

extern "C" size_t ptg_mlp_grad_workspace_bytes(int64_t batch)
{
    return (size_t)std::max<int64_t>(1, ptg_cdiv(batch, TR_B)) * (MLP_NP + 1) * sizeof(float);
}

extern "C" ptg_status ptg_mlp_grad(const float *params, const float *x, const float *y, int64_t nsamples,
                                   const int32_t *index, uint64_t seed, int64_t batch, float under_weight,
                                   float *grad, void *ws, size_t ws_bytes, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (!params || !x || !y || !grad || batch < 1 || nsamples < 1 || !(under_weight > 0)) {
        ptg_set_error("ptg_mlp_grad: bad arguments");
        return PTG_ERR_ARG;
    }
    if (!ws || ws_bytes < ptg_mlp_grad_workspace_bytes(batch)) {
        ptg_set_error("ptg_mlp_grad: workspace too small (%zu < %zu)", ws_bytes, ptg_mlp_grad_workspace_bytes(batch));
        return PTG_ERR_WORKSPACE;
    }
    const int64_t nb = ptg_cdiv(batch, TR_B);
    k_mlp_grad<<<(unsigned)nb, TR_B, 0, s>>>(params, x, y, nsamples, index, seed, batch, under_weight, (float *)ws);
    k_reduce_parts<<<lgrid(MLP_NP + 1), 256, 0, s>>>((const float *)ws, nb, grad);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

extern "C" ptg_status ptg_mlp_adam(float *params, float *m, float *v, const float *grad, int64_t step, float lr,
                                   float beta1, float beta2, float eps, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (!params || !m || !v || !grad || step < 1 || !(lr > 0)) {
        ptg_set_error("ptg_mlp_adam: bad arguments");
        return PTG_ERR_ARG;
    }
    k_adam<<<lgrid(MLP_NP), 256, 0, s>>>(params, m, v, grad, step, lr, beta1, beta2, eps);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

extern "C" ptg_status ptg_mlp_forward(const float *params, const float *x, int64_t count, int bf16, float *y,
                                      void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (!params || count < 0 || (count > 0 && (!x || !y))) {
        ptg_set_error("ptg_mlp_forward: bad arguments");
        return PTG_ERR_ARG;
    }
    if (count == 0) return PTG_OK;
    k_mlp_forward<<<lgrid(count), 256, 0, s>>>(params, x, count, bf16 != 0, y);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

extern "C" ptg_status ptg_open_radii(const ptg_tree_arrays *tree, int64_t n, int64_t nnodes, const ptg_forest *forest,
                                     const ptg_walk_tree *wt, const float *params, float margin, double eps,
                                     const double *eps_scale, int bf16, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (!tree || !wt || !params || n < 1 || nnodes < 1 || wt->nslots != n + nnodes || !wt->aux || !(eps > 0) ||
        !tree->slot) {
        ptg_set_error("ptg_open_radii: bad arguments");
        return PTG_ERR_ARG;
    }
    if (wt->prec != PTG_F32) {
        ptg_set_error("ptg_open_radii: the learned criterion runs on the FP32 walk (prec=PTG_F32 pack)");
        return PTG_ERR_ARG;
    }
    if (forest && (forest->nsys < 1 || (forest->nsys > 1 && !forest->roots))) {
        ptg_set_error("ptg_open_radii: bad forest");
        return PTG_ERR_ARG;
    }
    EcForest fo{forest ? forest->nsys : 1, forest ? forest->offsets : nullptr, forest ? forest->roots : nullptr};
    k_open_radii<<<lgrid(nnodes), 256, 0, s>>>(*tree, n, nnodes, (float *)wt->aux, wt->nslots, params, margin, eps,
                                               eps_scale, fo, bf16 != 0, 1e4f);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}
