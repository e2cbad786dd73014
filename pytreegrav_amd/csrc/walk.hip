// Grouped Barnes-Hut walk for gfx950: one wavefront walks the tree for one group of Morton-
// consecutive targets (grouped_treewalk.py:163-241).  Traversal is breadth-ordered over a stack in
// LDS: each step pops up to 64 elements (one per lane), evaluates the opening test against the
// group's bounding box in parallel, appends accepted elements to LDS interaction lists with
// ballot/popcount compaction and pushes the children of opened nodes with a wave prefix sum.
// When a list holds 64 entries the wave evaluates them against its targets, every lane reading the
// same LDS entry (broadcast).  Persistent waves take groups from per-XCD counters.
//
// Group expansion (FP32 monopole walks, no reference counterpart): an accepted element far enough
// from the group's centre that a third-order Taylor expansion of its field about that centre is
// within the tolerance is not listed; the lane that accepted it adds its derivative tensors to the
// lane's local-expansion coefficients, the wave sums them once per group and every target
// evaluates the expansion at its offset from the centre.

#include <algorithm>
#include <mutex>

#include "ptg_common.hpp"

#define WALK_WAVES 4
#define SCAP 768       // LDS traversal stack entries per wave (>= 512, the most children one step pushes)
#define SPILL 256      // entries refilled from global memory at a time (<= SCAP - 512)
#define GCAP 16384     // global spill entries per wave
// a list is evaluated in multiples of LF entries once it holds LF (a power of 2: 64, or 32 for the
// six-wave variant); a list holds LF - 1 + one step's 64 entries (ICAP)
#define NXCD 8
#define CTR_STRIDE 32  // int32 words between per-XCD counters (128 B)

template <typename T> struct alignas(4 * sizeof(T)) V4 {
    T x, y, z, w;
};

template <typename T, bool QUAD, int LF> struct WaveLds {
    static constexpr int ICAP = LF + 64;
    int stack[SCAP];
    V4<T> nodes[ICAP];  // accepted nodes: x y z m
    T nh[sizeof(T) == 8 ? ICAP : 1];  // node softening (FP64 exact path only)
    T nq[QUAD ? 6 * ICAP : 1];
    V4<T> parts[ICAP];  // particles: x y z m
    T ph[ICAP];
};

struct WalkCtl {
    int32_t ctr[NXCD * CTR_STRIDE];   // per-XCD group counters
    int32_t err;
    int32_t hwm;  // deepest global spill stack (entries) any wave reached in the last walk
    int32_t pad[30];
};

template <typename T> struct WalkArgs {
    const V4<T> *xyzm;
    const V4<T> *aux;
    const int2 *kids;
    const T *quad;  // 8 per slot
    const double *pos;
    const double *soft;
    const int64_t *perm;
    int sorted;  // pos/soft are indexed by the sorted target position (perm maps the outputs only)
    int64_t nslots;
    int64_t nt;
    int64_t ngroups;
    const int32_t *goff;    // optional group offsets (ngroups + 1), else fixed groups of gs
    const int32_t *gcount;  // device group count when goff is given
    int gs;
    T theta;
    T inv_theta;  // FP32 path: 1 / theta
    int crit;  // 0: geometric (treewalk.py:22-25); 1: quadrupole error bound; 2: learned
    T eps;     // criterion 1/2: per-node acceleration error tolerance (already divided by 3.5)
    T eps_exp;                 // group expansion: per-element tolerance of the expansion test
    T exp_rel;                 // group expansion: cap of the truncation bound relative to M / D^2 (0: none)
    const double *eps_scale;   // nullable: eps_exp * eps_scale[system] (one scale per system)
    double G;
    double ox, oy, oz;
    // forest: group g walks the tree of the system s holding its first target (root at slot s),
    // with coordinates relative to that system's cube centre (FP32) or absolute (FP64)
    int64_t nsys;          // 0: one tree
    const int64_t *soff;   // (nsys + 1) system target ranges
    const double *cubes;   // (nsys, 4)
    double *out;
    int64_t *counters;
    WalkCtl *ctl;
    int *spill;
    int gcap;  // global spill entries per wave (GCAP, or what the static scratch of ptg_walk_noalloc gives)
    const int32_t *start = nullptr;  // subtree walks: device slot of the element every walk starts at
};

// ---- softening kernels (kernel.py:4-75), generic precision ----
template <typename T> __device__ __forceinline__ T force_kernel(T r, T h)
{
    if (r > h) return T(1) / (r * r * r);
    T hinv = T(1) / h;
    T q = r * hinv;
    if (q <= T(0.5)) return (T(10.666666666666666666) + q * q * (T(-38.4) + T(32.0) * q)) * hinv * hinv * hinv;
    return (T(21.333333333333) - T(48.0) * q + T(38.4) * q * q - T(10.666666666667) * q * q * q -
            T(0.066666666667) / (q * q * q)) *
           hinv * hinv * hinv;
}

template <typename T> __device__ __forceinline__ T tidal_kernel(T r, T h)
{
    if (r > h) {
        T r2 = r * r;
        return T(3) / (r2 * r2 * r);
    }
    T hinv = T(1) / h;
    T hinv2 = hinv * hinv;
    T hinv5 = hinv2 * hinv2 * hinv;
    T q = r * hinv;
    if (q <= T(0.5)) return (T(76.8) - T(96.0) * q) * hinv5;
    T qinv = T(1) / q;
    T qinv3 = qinv * qinv * qinv;
    return (T(48.0) * qinv - T(76.8) + T(32.0) * q - T(0.2) * qinv3 * qinv * qinv) * hinv5;
}

template <typename T> __device__ __forceinline__ T potential_kernel(T r, T h)
{
    if (h == T(0)) return T(-1) / r;
    T hinv = T(1) / h;
    T q = r * hinv;
    if (q <= T(0.5)) return (T(-2.8) + q * q * (T(5.33333333333333333) + q * q * (T(6.4) * q - T(9.6)))) * hinv;
    if (q <= T(1))
        return (T(-3.2) + T(0.066666666666666666666) / q +
                q * q * (T(10.666666666666666666666) + q * (T(-16.0) + q * (T(9.6) - T(2.1333333333333333333333) * q)))) *
               hinv;
    return T(-1) / r;
}

__device__ __forceinline__ float fast_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double fast_rsqrt(double x) { return 1.0 / sqrt(x); }

template <int FIELDS> struct NW {
    static constexpr int POT = (FIELDS & PTG_POT) ? 1 : 0;
    static constexpr int ACC = (FIELDS & PTG_ACC) ? 1 : 0;
    static constexpr int TID = (FIELDS & PTG_TIDAL) ? 1 : 0;
    static constexpr int W = POT + 3 * ACC + 9 * TID;
};

// per-target accumulators: phi, g(3), T(6 independent)
template <typename T> struct Acc {
    T phi, gx, gy, gz, txx, tyy, tzz, txy, txz, tyz;
};

// One source element on one target, with the reference's branches (grouped_treewalk.py:75-153):
// r2 > 0 self skip, softened iff r < max(h_s, h_t), quadrupole terms for nodes only.
template <typename T, int FIELDS, bool QUADTERM>
__device__ __forceinline__ void interact_exact(Acc<T> &a, T sx, T sy, T sz, T M, T hs, const T *q, T tx, T ty, T tz,
                                               T th)
{
    constexpr bool P = NW<FIELDS>::POT, G = NW<FIELDS>::ACC, D = NW<FIELDS>::TID;
    T dx = sx - tx, dy = sy - ty, dz = sz - tz;
    T r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 > T(0))) return;
    T r = sqrt(r2);
    T ht = hs > th ? hs : th;
    bool softened = r < ht;
    T phi = 0, K = 0, Q = 0;
    if (P) phi = softened ? M * potential_kernel(r, ht) : -M / r;
    if (G || D) K = softened ? M * force_kernel(r, ht) : M / (r * r2);
    if (D) Q = softened ? M * tidal_kernel(r, ht) : T(3) * K / r2;
    T gx = 0, gy = 0, gz = 0;
    if (G) {
        gx = K * dx;
        gy = K * dy;
        gz = K * dz;
    }
    T txx = 0, tyy = 0, tzz = 0, txy = 0, txz = 0, tyz = 0;
    if (D) {
        txx = Q * dx * dx - K;
        tyy = Q * dy * dy - K;
        tzz = Q * dz * dz - K;
        txy = Q * dx * dy;
        txz = Q * dx * dz;
        tyz = Q * dy * dz;
    }
    if (QUADTERM) {
        T qxx = q[0], qxy = q[1], qxz = q[2], qyy = q[3], qyz = q[4], qzz = q[5];
        T r5inv = T(1) / (r2 * r2 * r);
        T r7inv = r5inv / r2;
        T qdx = qxx * dx + qxy * dy + qxz * dz;
        T qdy = qxy * dx + qyy * dy + qyz * dz;
        T qdz = qxz * dx + qyz * dy + qzz * dz;
        T A = dx * qdx + dy * qdy + dz * qdz;
        if (P) phi += T(-0.5) * A * r5inv;
        if (G) {
            T quad_fac = A * r7inv;
            gx += T(2.5) * quad_fac * dx - qdx * r5inv;
            gy += T(2.5) * quad_fac * dy - qdy * r5inv;
            gz += T(2.5) * quad_fac * dz - qdz * r5inv;
        }
        if (D) {
            T diag = T(-2.5) * A * r7inv;
            T fac = T(17.5) * A * r7inv / r2;
            txx += qxx * r5inv - T(10.0) * dx * qdx * r7inv + diag + fac * dx * dx;
            tyy += qyy * r5inv - T(10.0) * dy * qdy * r7inv + diag + fac * dy * dy;
            tzz += qzz * r5inv - T(10.0) * dz * qdz * r7inv + diag + fac * dz * dz;
            txy += qxy * r5inv - T(5.0) * (dx * qdy + dy * qdx) * r7inv + fac * dx * dy;
            txz += qxz * r5inv - T(5.0) * (dx * qdz + dz * qdx) * r7inv + fac * dx * dz;
            tyz += qyz * r5inv - T(5.0) * (dy * qdz + dz * qdy) * r7inv + fac * dy * dz;
        }
    }
    a.phi += phi;
    a.gx += gx;
    a.gy += gy;
    a.gz += gz;
    a.txx += txx;
    a.tyy += tyy;
    a.tzz += tzz;
    a.txy += txy;
    a.txz += txz;
    a.tyz += tyz;
}

// Accepted node on one target, fast path: the acceptance test guarantees r > h_t, h_s and r > 0.
template <typename T, int FIELDS, bool QUAD>
__device__ __forceinline__ void interact_node_fast(Acc<T> &a, V4<T> s, const T *q, T tx, T ty, T tz)
{
    constexpr bool P = NW<FIELDS>::POT, G = NW<FIELDS>::ACC, D = NW<FIELDS>::TID;
    T dx = s.x - tx, dy = s.y - ty, dz = s.z - tz;
    T r2 = dx * dx + dy * dy + dz * dz;
    T rinv = fast_rsqrt(r2);
    T rinv2 = rinv * rinv;
    T mr = s.w * rinv;
    if (P) a.phi -= mr;
    T K = mr * rinv2;
    if (G) {
        a.gx += K * dx;
        a.gy += K * dy;
        a.gz += K * dz;
    }
    if (D) {
        T Q = T(3) * K * rinv2;
        a.txx += Q * dx * dx - K;
        a.tyy += Q * dy * dy - K;
        a.tzz += Q * dz * dz - K;
        a.txy += Q * dx * dy;
        a.txz += Q * dx * dz;
        a.tyz += Q * dy * dz;
    }
    if (QUAD) {
        T qxx = q[0], qxy = q[1], qxz = q[2], qyy = q[3], qyz = q[4], qzz = q[5];
        T r5inv = rinv2 * rinv2 * rinv;
        T r7inv = r5inv * rinv2;
        T qdx = qxx * dx + qxy * dy + qxz * dz;
        T qdy = qxy * dx + qyy * dy + qyz * dz;
        T qdz = qxz * dx + qyz * dy + qzz * dz;
        T A = dx * qdx + dy * qdy + dz * qdz;
        if (P) a.phi += T(-0.5) * A * r5inv;
        if (G) {
            T f = T(2.5) * A * r7inv;
            a.gx += f * dx - qdx * r5inv;
            a.gy += f * dy - qdy * r5inv;
            a.gz += f * dz - qdz * r5inv;
        }
        if (D) {
            T diag = T(-2.5) * A * r7inv;
            T fac = T(17.5) * A * r7inv * rinv2;
            a.txx += qxx * r5inv - T(10.0) * dx * qdx * r7inv + diag + fac * dx * dx;
            a.tyy += qyy * r5inv - T(10.0) * dy * qdy * r7inv + diag + fac * dy * dy;
            a.tzz += qzz * r5inv - T(10.0) * dz * qdz * r7inv + diag + fac * dz * dz;
            a.txy += qxy * r5inv - T(5.0) * (dx * qdy + dy * qdx) * r7inv + fac * dx * dy;
            a.txz += qxz * r5inv - T(5.0) * (dx * qdz + dz * qdx) * r7inv + fac * dx * dz;
            a.tyz += qyz * r5inv - T(5.0) * (dy * qdz + dz * qdy) * r7inv + fac * dy * dz;
        }
    }
}

// The M4 spline terms for r < h (kernel.py:4-75) with reciprocals instead of IEEE divisions, for
// the FP32 fast path: hinv = 1/h, qinv = h/r = h * rinv.
template <typename T> struct SplineFast {
    T K, phi, Q;
};
template <typename T, bool P, bool D>
__device__ __forceinline__ SplineFast<T> spline_fast(T r, T rinv, T h)
{
    T hinv = __builtin_amdgcn_rcpf(h);
    T q = r * hinv;
    T hinv2 = hinv * hinv;
    T hinv3 = hinv2 * hinv;
    SplineFast<T> o;
    o.phi = 0;
    o.Q = 0;
    if (q <= T(0.5)) {
        T q2 = q * q;
        o.K = (T(10.666666666666666666) + q2 * (T(-38.4) + T(32.0) * q)) * hinv3;
        if (P) o.phi = (T(-2.8) + q2 * (T(5.33333333333333333) + q2 * (T(6.4) * q - T(9.6)))) * hinv;
        if (D) o.Q = (T(76.8) - T(96.0) * q) * hinv3 * hinv2;
    } else {
        T qinv = h * rinv;
        T qinv3 = qinv * qinv * qinv;
        o.K = (T(21.333333333333) - T(48.0) * q + T(38.4) * q * q - T(10.666666666667) * q * q * q -
               T(0.066666666667) * qinv3) * hinv3;
        if (P)
            o.phi = (T(-3.2) + T(0.066666666666666666666) * qinv +
                     q * q * (T(10.666666666666666666666) + q * (T(-16.0) + q * (T(9.6) - T(2.1333333333333333333333) * q)))) *
                    hinv;
        if (D) o.Q = (T(48.0) * qinv - T(76.8) + T(32.0) * q - T(0.2) * qinv3 * qinv * qinv) * hinv3 * hinv2;
    }
    return o;
}

// Particle on one target, fast path: point mass unless any lane is inside the softening radius,
// in which case those lanes evaluate the spline (wave-uniform branch); r2 == 0 contributes zero.
template <typename T, int FIELDS>
__device__ __forceinline__ void interact_part_fast(Acc<T> &a, V4<T> s, T hs, T tx, T ty, T tz, T th)
{
    constexpr bool P = NW<FIELDS>::POT, G = NW<FIELDS>::ACC, D = NW<FIELDS>::TID;
    T dx = s.x - tx, dy = s.y - ty, dz = s.z - tz;
    T r2 = dx * dx + dy * dy + dz * dz;
    T ht = fmaxf(hs, th);
    bool nonzero = r2 > T(0);
    bool soft = nonzero && r2 < ht * ht;
    T rinv = nonzero ? fast_rsqrt(r2) : T(0);
    T rinv2 = rinv * rinv;
    T mr = s.w * rinv;
    T phi = -mr;
    T K = mr * rinv2;
    T Q = T(3) * K * rinv2;
    if (ptg_ballot(soft)) {
        if (soft) {
            SplineFast<T> sp = spline_fast<T, P, D>(r2 * rinv, rinv, ht);
            if (P) phi = s.w * sp.phi;
            K = s.w * sp.K;
            if (D) Q = s.w * sp.Q;
        }
    }
    if (P) a.phi += phi;
    if (G) {
        a.gx += K * dx;
        a.gy += K * dy;
        a.gz += K * dz;
    }
    if (D) {
        a.txx += Q * dx * dx - K;
        a.tyy += Q * dy * dy - K;
        a.tzz += Q * dz * dz - K;
        a.txy += Q * dx * dy;
        a.txz += Q * dx * dz;
        a.tyz += Q * dy * dz;
    }
}

// ---- group expansion ----
// A point mass M at D = y - c from the group centre c acts on a target at c + u through
// f(D - u) = 1/|D - u| = f - u.f1 + (1/2) u f2 u - (1/6) f3 uuu + O(|u|^4 / |D|^5) with the
// derivative tensors of 1/r at D: f1 = -D/D^3, f2 = 3 D D / D^5 - I / D^3,
// f3 = -15 D D D / D^7 + 3 (D I + I D + (I D)^T) / D^5.  Summed over the folded elements:
// a0 = sum M f, g = -sum M f1, a2 = sum M f2, a3 = sum M f3; then at u
//   phi = -a0 - u.g - (1/2) u a2 u + (1/6) a3 uuu,    acc = g + a2 u - (1/2) a3 uu.
// a2 and a3 are trace-free (1/r is harmonic), so five and seven components are kept:
// a2zz = -(xx + yy), a3xzz = -(xxx + xyy), a3yzz = -(xxy + yyy), a3zzz = -(xxz + yyz).
// Error bound: the dropped terms of 1/|D - u| are |u|^l P_l(cos g) / D^(l+1), l >= 4, and the
// gradient of the solid harmonic r^l P_l(cos) is at most l r^(l-1) long, so for |u| <= R < D the
// acceleration misses by at most (M / D^2) sum_{l>=4} l x^(l-1) = (M / D^2) x^3 (4 - 3x) / (1 - x)^2
// with x = R / D (the potential by (M / D) x^4 / (1 - x)); an element is folded when that is below
// the tolerance.
template <typename T> struct GroupExp {
    T a0, gx, gy, gz;
    T xx, yy, xy, xz, yz;
    T xxx, xxy, xxz, xyy, xyz, yyy, yyz;
};

template <typename T>
__device__ __forceinline__ void exp_add(GroupExp<T> &E, T M, T Dx, T Dy, T Dz, T rinv)
{
    const T rinv2 = rinv * rinv;
    const T mr = M * rinv;
    const T m3 = mr * rinv2;
    E.a0 += mr;
    E.gx += m3 * Dx;
    E.gy += m3 * Dy;
    E.gz += m3 * Dz;
    const T p = T(3) * m3 * rinv2;
    const T pDx = p * Dx, pDy = p * Dy;
    E.xx += pDx * Dx - m3;
    E.yy += pDy * Dy - m3;
    E.xy += pDx * Dy;
    E.xz += pDx * Dz;
    E.yz += pDy * Dz;
    const T q = T(-5) * p * rinv2;
    const T qxx = q * Dx * Dx, qyy = q * Dy * Dy;
    E.xxx += (qxx + T(3) * p) * Dx;
    E.xxy += (qxx + p) * Dy;
    E.xxz += (qxx + p) * Dz;
    E.xyy += (qyy + p) * Dx;
    E.xyz += q * Dx * Dy * Dz;
    E.yyy += (qyy + T(3) * p) * Dy;
    E.yyz += (qyy + p) * Dz;
}

// the fold test: truncation bound of the expansion about a centre at offset -D, radius rg, below
// the absolute tolerance eps and, when rel > 0, below rel times the element's own field M / D^2
// (bound / (M / D^2) = x^3 (4 - 3x) / (1 - x)^2 depends on x alone, so rel caps x)
template <typename T>
__device__ __forceinline__ bool fold_test(T M, T Dx, T Dy, T Dz, T rg, T eps, T rel, T &rinv)
{
    const T D2 = Dx * Dx + Dy * Dy + Dz * Dz;
    rinv = fast_rsqrt(D2);  // D2 == 0 gives x = inf or nan: no fold
    const T x = rg * rinv, omx = T(1) - x;
    const T lhs = x * x * x * (T(4) - T(3) * x);
    const T rhs = omx * omx;
    return x < T(1) && M * lhs < eps * D2 * rhs && (rel <= T(0) || lhs < rel * rhs);
}

__device__ __forceinline__ float wave_fsum(float v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

__device__ __forceinline__ void exp_wave_sum(GroupExp<float> &E)
{
    E.a0 = wave_fsum(E.a0);
    E.gx = wave_fsum(E.gx);
    E.gy = wave_fsum(E.gy);
    E.gz = wave_fsum(E.gz);
    E.xx = wave_fsum(E.xx);
    E.yy = wave_fsum(E.yy);
    E.xy = wave_fsum(E.xy);
    E.xz = wave_fsum(E.xz);
    E.yz = wave_fsum(E.yz);
    E.xxx = wave_fsum(E.xxx);
    E.xxy = wave_fsum(E.xxy);
    E.xxz = wave_fsum(E.xxz);
    E.xyy = wave_fsum(E.xyy);
    E.xyz = wave_fsum(E.xyz);
    E.yyy = wave_fsum(E.yyy);
    E.yyz = wave_fsum(E.yyz);
}

template <typename T, int FIELDS>
__device__ __forceinline__ void exp_eval(Acc<T> &a, const GroupExp<T> &E, T ux, T uy, T uz)
{
    const T zz = -(E.xx + E.yy);
    const T xzz = -(E.xxx + E.xyy), yzz = -(E.xxy + E.yyy), zzz = -(E.xxz + E.yyz);
    const T a2x = E.xx * ux + E.xy * uy + E.xz * uz;
    const T a2y = E.xy * ux + E.yy * uy + E.yz * uz;
    const T a2z = E.xz * ux + E.yz * uy + zz * uz;
    const T uxx = ux * ux, uyy = uy * uy, uzz = uz * uz;
    const T uxy = T(2) * ux * uy, uxz = T(2) * ux * uz, uyz = T(2) * uy * uz;
    const T a3x = E.xxx * uxx + E.xyy * uyy + xzz * uzz + E.xxy * uxy + E.xxz * uxz + E.xyz * uyz;
    const T a3y = E.xxy * uxx + E.yyy * uyy + yzz * uzz + E.xyy * uxy + E.xyz * uxz + E.yyz * uyz;
    const T a3z = E.xxz * uxx + E.yyz * uyy + zzz * uzz + E.xyz * uxy + xzz * uxz + yzz * uyz;
    if (NW<FIELDS>::POT)
        a.phi += -E.a0 - (ux * E.gx + uy * E.gy + uz * E.gz) - T(0.5) * (ux * a2x + uy * a2y + uz * a2z) +
                 T(1.0 / 6.0) * (ux * a3x + uy * a3y + uz * a3z);
    if (NW<FIELDS>::ACC) {
        a.gx += E.gx + a2x - T(0.5) * a3x;
        a.gy += E.gy + a2y - T(0.5) * a3y;
        a.gz += E.gz + a2z - T(0.5) * a3z;
    }
}

// gap between x and the interval [lo, hi] (0 inside): bmin - x and x - bmax are exact differences,
// so this equals the two-sided branch form bit for bit
__device__ __forceinline__ float box_gap(float x, float lo, float hi) { return fmaxf(fmaxf(lo - x, x - hi), 0.f); }
__device__ __forceinline__ double box_gap(double x, double lo, double hi) { return fmax(fmax(lo - x, x - hi), 0.0); }

// Stack spill: when sp entries plus the tot children about to be pushed exceed SCAP, the bottom
// ks entries (the excess rounded up to 64, at most sp; tot <= 512 <= SCAP so ks >= the excess)
// go to the wave's global spill stack and the rest move down.  Any order of visits gives the same
// accepted sets, so the global stack is refilled from its top when the LDS stack runs empty.
static_assert(SCAP >= 512 && SPILL >= 64 && SPILL <= SCAP - 512, "stack geometry");
__device__ __forceinline__ int spill_count(int sp, int tot)
{
    const int ex = sp + tot - SCAP;
    return std::min(sp, (ex + 63) & ~63);
}
__device__ __forceinline__ void spill_bottom(int *stack, int *dst, int sp, int ks, int lane)
{
    for (int q = lane; q < ks; q += 64) dst[q] = stack[q];
    __builtin_amdgcn_wave_barrier();
    // read [ks + q, ks + q + 64) before writing [q, q + 64); later reads lie above earlier writes
    for (int q = 0; q < sp - ks; q += 64) {
        const int v = (q + lane < sp - ks) ? stack[ks + q + lane] : 0;
        __builtin_amdgcn_wave_barrier();
        if (q + lane < sp - ks) stack[q + lane] = v;
        __builtin_amdgcn_wave_barrier();
    }
}

// lanes below this one whose bit is set in m
__device__ __forceinline__ int lanes_below(uint64_t m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// exclusive prefix sum across the wave of v in [0, 15] (children counts), one ballot per bit;
// tot receives the wave total
__device__ __forceinline__ int small_excl_scan(int v, int &tot)
{
    int below = 0, t = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const uint64_t m = ptg_ballot((v >> b) & 1);
        below += lanes_below(m) << b;
        t += (int)__popcll(m) << b;
    }
    tot = t;
    return below;
}

template <typename T, int TPL, int FIELDS, bool QUAD, bool EXACT, int LF>
__device__ __forceinline__ void flush_nodes(const WaveLds<T, QUAD, LF> &L, int cnt, Acc<T> (&acc)[TPL],
                                            const T (&tx)[TPL], const T (&ty)[TPL], const T (&tz)[TPL],
                                            const T (&th)[TPL])
{
    for (int k = 0; k < cnt; k++) {
        V4<T> s = L.nodes[k];
        T q[6];
        if (QUAD) {
#pragma unroll
            for (int c = 0; c < 6; c++) q[c] = L.nq[6 * k + c];
        }
#pragma unroll
        for (int j = 0; j < TPL; j++) {
            if constexpr (EXACT)
                interact_exact<T, FIELDS, QUAD>(acc[j], s.x, s.y, s.z, s.w, L.nh[k], q, tx[j], ty[j], tz[j], th[j]);
            else
                interact_node_fast<T, FIELDS, QUAD>(acc[j], s, q, tx[j], ty[j], tz[j]);
        }
    }
}

template <typename T, int TPL, int FIELDS, bool QUAD, bool EXACT, int LF>
__device__ __forceinline__ void flush_parts(const WaveLds<T, QUAD, LF> &L, int cnt, Acc<T> (&acc)[TPL],
                                            const T (&tx)[TPL], const T (&ty)[TPL], const T (&tz)[TPL],
                                            const T (&th)[TPL])
{
    for (int k = 0; k < cnt; k++) {
        V4<T> s = L.parts[k];
        T hs = L.ph[k];
#pragma unroll
        for (int j = 0; j < TPL; j++) {
            if constexpr (EXACT)
                interact_exact<T, FIELDS, false>(acc[j], s.x, s.y, s.z, s.w, hs, nullptr, tx[j], ty[j], tz[j], th[j]);
            else
                interact_part_fast<T, FIELDS>(acc[j], s, hs, tx[j], ty[j], tz[j], th[j]);
        }
    }
}

// Kernel variants of the FP32 one-target-per-lane walk (ptg_criterion.variant; the autotuner picks
// one per problem).  The traversal is latency-bound (3 -> 4 waves per SIMD: 4.5 -> 3.7 ms on the
// 1e7-particle walk), so registers are capped to fit more waves:
//   0: 5 waves per SIMD, lists evaluated at 64 entries (7.7 KB of LDS per wave, five 4-wave blocks per CU)
//   1: 4 waves per SIMD, lists at 64 (more registers per lane)
//   2: 6 waves per SIMD, lists at 32 (6.4 KB of LDS per wave, six blocks per CU)
// Every variant evaluates the same terms; variants 0 and 1 in the same order (bit-identical results),
// variant 2's shorter lists interleave the node and particle list evaluations differently, so its
// FP32 sums differ at the rounding level.
#define WALK_NVARIANTS 3
template <typename T, int TPL, int VAR> struct WalkOcc {
    static constexpr bool fast = sizeof(T) == 4 && TPL == 1;
    static constexpr int waves = !fast ? 1 : VAR == 1 ? 4 : VAR == 2 ? 6 : 5;
    static constexpr int lf = fast && VAR == 2 ? 32 : 64;
    static constexpr int blocks_per_cu = fast && VAR == 2 ? 6 : 5;
};

template <typename T, int TPL, int FIELDS, bool QUAD, bool EXACT, bool EXP, int VAR>
__global__ __launch_bounds__(64 * WALK_WAVES) __attribute__((amdgpu_waves_per_eu(WalkOcc<T, TPL, VAR>::waves)))
void k_walk(WalkArgs<T> A)
{
    constexpr int LF = WalkOcc<T, TPL, VAR>::lf;
    static_assert(!EXP || (sizeof(T) == 4 && !QUAD && !EXACT && !NW<FIELDS>::TID),
                  "the group expansion is an FP32 monopole potential/acceleration path");
    __shared__ WaveLds<T, QUAD, LF> lds[WALK_WAVES];
    const int wave = threadIdx.x >> 6, lane = ptg_lane();
    WaveLds<T, QUAD, LF> &L = lds[wave];
    const int64_t gwave = (int64_t)blockIdx.x * WALK_WAVES + wave;
    int *spill = A.spill + gwave * (int64_t)A.gcap;
    const int xcd = blockIdx.x % NXCD;
    int64_t n_int = 0, n_vis = 0, n_part = 0, n_exp = 0;
    int ghw = 0;  // this wave's deepest spill
    const int64_t ngroups = A.goff ? (int64_t)__builtin_amdgcn_readfirstlane(*A.gcount) : A.ngroups;

    for (int attempt = 0; attempt < NXCD; attempt++) {
        const int x = (xcd + attempt) % NXCD;
        const int64_t g_lo = ngroups * x / NXCD, g_hi = ngroups * (x + 1) / NXCD;
        for (;;) {
            int gi = 0;
            if (lane == 0) gi = atomicAdd(&A.ctl->ctr[x * CTR_STRIDE], 1);
            gi = __builtin_amdgcn_readfirstlane(gi);
            const int64_t g = g_lo + gi;
            if (g >= g_hi) break;
            const int64_t a0 = A.goff ? (int64_t)A.goff[g] : g * A.gs;
            const int64_t e0 = A.goff ? (int64_t)A.goff[g + 1] : A.nt;
            const int64_t b0 = e0 < a0 + A.gs ? e0 : a0 + A.gs;
            const int mt = (int)(b0 - a0);
            if (A.goff && e0 > b0 && lane == 0) atomicOr(&A.ctl->err, 2);  // a supplied group larger than group_size
            int root = 0;
            double ox = A.ox, oy = A.oy, oz = A.oz;
            if (A.nsys > 0) {
                int64_t lo = 0, hi = A.nsys;  // last system with soff[s] <= a0
                while (hi - lo > 1) {
                    const int64_t mid = (lo + hi) >> 1;
                    if (A.soff[mid] <= a0) lo = mid;
                    else hi = mid;
                }
                root = (int)lo;
                if (sizeof(T) == 4) {
                    ox = A.cubes[4 * lo];
                    oy = A.cubes[4 * lo + 1];
                    oz = A.cubes[4 * lo + 2];
                } else {
                    ox = oy = oz = 0.0;
                }
            }

            // ---- targets: lane holds a0 + lane + 64 j ----
            T tx[TPL], ty[TPL], tz[TPL], th[TPL];
            int64_t tidx[TPL];
            bool tval[TPL];
            T bmin0 = INFINITY, bmin1 = INFINITY, bmin2 = INFINITY;
            T bmax0 = -INFINITY, bmax1 = -INFINITY, bmax2 = -INFINITY, hmax = -INFINITY;
#pragma unroll
            for (int j = 0; j < TPL; j++) {
                const int o = lane + 64 * j;
                tval[j] = o < mt;
                tidx[j] = tval[j] ? A.perm[a0 + o] : 0;
                if (tval[j]) {
                    const int64_t src = A.sorted ? a0 + o : tidx[j];
                    tx[j] = (T)(A.pos[3 * src] - ox);
                    ty[j] = (T)(A.pos[3 * src + 1] - oy);
                    tz[j] = (T)(A.pos[3 * src + 2] - oz);
                    th[j] = (T)A.soft[src];
                    bmin0 = std::min(bmin0, tx[j]);
                    bmin1 = std::min(bmin1, ty[j]);
                    bmin2 = std::min(bmin2, tz[j]);
                    bmax0 = std::max(bmax0, tx[j]);
                    bmax1 = std::max(bmax1, ty[j]);
                    bmax2 = std::max(bmax2, tz[j]);
                    hmax = std::max(hmax, th[j]);
                } else {
                    tx[j] = ty[j] = tz[j] = th[j] = T(0);
                }
            }
            bmin0 = ptg_wave_min(bmin0);
            bmin1 = ptg_wave_min(bmin1);
            bmin2 = ptg_wave_min(bmin2);
            bmax0 = ptg_wave_max(bmax0);
            bmax1 = ptg_wave_max(bmax1);
            bmax2 = ptg_wave_max(bmax2);
            hmax = ptg_wave_max(hmax);
            // invalid lanes sit at the box corner so they never divide by zero
#pragma unroll
            for (int j = 0; j < TPL; j++)
                if (!tval[j]) {
                    tx[j] = bmin0 - T(1);
                    ty[j] = bmin1 - T(1);
                    tz[j] = bmin2 - T(1);
                }

            Acc<T> acc[TPL];
#pragma unroll
            for (int j = 0; j < TPL; j++) acc[j] = Acc<T>{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

            // group expansion: centre and radius of the group's box, this system's tolerance
            T cx = 0, cy = 0, cz = 0, rg = 0, eps_x = 0;
            GroupExp<T> E{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            if constexpr (EXP) {
                cx = T(0.5) * (bmin0 + bmax0);
                cy = T(0.5) * (bmin1 + bmax1);
                cz = T(0.5) * (bmin2 + bmax2);
                const T hx = T(0.5) * (bmax0 - bmin0), hy = T(0.5) * (bmax1 - bmin1), hz = T(0.5) * (bmax2 - bmin2);
                // rounded up by a relative 1e-6 so every target's offset from the centre is within rg
                rg = sqrt(hx * hx + hy * hy + hz * hz) * T(1.000001);
                eps_x = A.eps_exp * (A.eps_scale ? (T)A.eps_scale[root] : T(1));
            }

            // the root of system s occupies walk slot s; a subtree walk (the FP64 reference-branch
            // walk only) starts at its element's slot instead (wave-uniform; an out-of-range slot is
            // reported, and dropped when popped like any corrupt link, so it walks nothing)
            int first = root;
            if constexpr (EXACT) {
                if (A.start) {
                    first = __builtin_amdgcn_readfirstlane(*A.start);
                    if ((first < 0 || first >= A.nslots) && lane == 0) atomicOr(&A.ctl->err, 4);
                }
            }
            int sp = 1, gsp = 0, nn = 0, np = 0;
            bool overflow = false;
            if (lane == 0) L.stack[0] = first;
            __builtin_amdgcn_wave_barrier();

            for (;;) {
                // pop up to 64 stack entries, one per lane, refilling the LDS stack from the spill
                // region when it is empty
                if (sp == 0) {
                    if (gsp > 0) {
                        const int k = std::min(gsp, SPILL);
                        for (int q = lane; q < k; q += 64) L.stack[q] = spill[gsp - k + q];
                        __builtin_amdgcn_wave_barrier();
                        gsp -= k;
                        sp = k;
                    } else {
                        break;
                    }
                }
                const int k = std::min(sp, 64);
                const int base = sp - k;
                int e = lane < k ? L.stack[base + lane] : 0;
                const bool ve = lane < k && e >= 0 && e < A.nslots;  // a corrupt link is dropped, not followed
                if (!ve) e = 0;
                __builtin_amdgcn_wave_barrier();
                sp = base;
                n_vis += k;
                const V4<T> X = A.xyzm[e];
                const V4<T> X2 = A.aux[e];
                const int2 K = A.kids[e];
                const bool ispart = K.y == 0;
                // distance to the box per axis: max(bmin - x, x - bmax, 0), the same value as the
                // two-sided test (one max3 per axis)
                const T dxm = box_gap(X.x, bmin0, bmax0), dym = box_gap(X.y, bmin1, bmax1), dzm = box_gap(X.z, bmin2, bmax2);
                const T r2m = dxm * dxm + dym * dym + dzm * dzm;
                const T hh = X2.x > hmax ? X2.x : hmax;
                bool acc_node, far_part = false;
                if constexpr (EXACT) {
                    // the reference's test as written: r_min > max(size/theta + delta, h + 0.6 size + delta)
                    const T r_min = sqrt(r2m);
                    const T c1 = X2.y / A.theta + X2.z;
                    const T c2 = hh + X2.y * T(0.6) + X2.z;
                    acc_node = ve && !ispart && r_min > (c1 > c2 ? c1 : c2);
                } else {
                    // the same test on squared distances (every threshold is >= 0), with 1/theta;
                    // evaluated without branches (bitwise &) so the slot's aux record is one load
                    // issued with the others, not a second round trip inside a branch
                    const T c1 = X2.y * A.inv_theta + X2.z;
                    const T c2 = hh + X2.y * T(0.6) + X2.z;
                    const T c = c1 > c2 ? c1 : c2;
                    acc_node = ve & !ispart & (r2m > c * c);
                    // a particle farther from the group's box than every softening radius (and not
                    // inside it) needs neither the softening nor the self test -> node list
                    far_part = ve & ispart & (r2m > hh * hh);
                    if (A.crit == 1) {
                        // error-bounded: also require the quadrupole term the monopole drops,
                        // 3.5 |Q| / r^4, to stay below eps (theta is then only a loose cap)
                        acc_node = acc_node & (X2.w < A.eps * r2m * r2m);
                    } else if (A.crit == 2) {
                        // learned: r_min beyond the node's predicted opening radius
                        acc_node = acc_node & (r2m > X2.w * X2.w);
                    }
                }
                bool acc_part = ve && ispart && !far_part;
                acc_node = acc_node || far_part;
                bool open = ve && !ispart && !acc_node;
                if constexpr (EXP) {
                    if (acc_node) {
                        // accepted elements are outside every softening radius and r_min > 0,
                        // so D >= r_min > 0; fold if the truncation bound is within the tolerance
                        const T Dx = X.x - cx, Dy = X.y - cy, Dz = X.z - cz;
                        T rinv;
                        if (fold_test(X.w, Dx, Dy, Dz, rg, eps_x, A.exp_rel, rinv)) {
                            exp_add(E, X.w, Dx, Dy, Dz, rinv);
                            acc_node = false;
                            n_exp++;
                        }
                    }
                }

                const uint64_t bp = ptg_ballot(acc_part), bn = ptg_ballot(acc_node);
                if (acc_part) {
                    const int at = np + lanes_below(bp);
                    L.parts[at] = X;
                    L.ph[at] = X2.x;
                }
                if (acc_node) {
                    const int at = nn + lanes_below(bn);
                    L.nodes[at] = X;
                    if (EXACT) L.nh[at] = X2.x;
                    if (QUAD) {
                        const T *q = A.quad + 8 * (int64_t)e;
#pragma unroll
                        for (int c = 0; c < 6; c++) L.nq[6 * at + c] = q[c];
                    }
                }
                const int addp = (int)__popcll(bp), addn = (int)__popcll(bn);
                np += addp;
                nn += addn;
                n_int += (int64_t)(addp + addn) * mt;
                n_part += (int64_t)addp * mt;

                // push the children of opened nodes
                const int cnt = open ? std::min(std::max(K.y, 0), 8) : 0;
                int tot;
                const int below = small_excl_scan(cnt, tot);
                if (sp + tot > SCAP) {
                    // move the bottom ks entries of the LDS stack to this wave's global region
                    const int ks = spill_count(sp, tot);
                    if (gsp + ks > A.gcap) {
                        overflow = true;
                        break;
                    }
                    spill_bottom(L.stack, spill + gsp, sp, ks, lane);
                    gsp += ks;
                    sp -= ks;
                    ghw = std::max(ghw, gsp);
                }
                {
                    // unrolled: eight masked LDS stores with immediate offsets, no per-lane loop
                    int *dst = &L.stack[sp + below];
#pragma unroll
                    for (int c = 0; c < 8; c++)
                        if (c < cnt) dst[c] = K.x + c;
                }
                sp += tot;
                __builtin_amdgcn_wave_barrier();

                if (nn >= LF) {
                    const int c = nn & ~(LF - 1);  // c >= LF > nn - c
                    flush_nodes<T, TPL, FIELDS, QUAD, EXACT, LF>(L, c, acc, tx, ty, tz, th);
                    __builtin_amdgcn_wave_barrier();
                    if (lane < nn - c) {  // reads [c, nn) and writes [0, nn - c): disjoint
                        L.nodes[lane] = L.nodes[c + lane];
                        if (EXACT) L.nh[lane] = L.nh[c + lane];
                        if (QUAD)
                            for (int q = 0; q < 6; q++) L.nq[6 * lane + q] = L.nq[6 * (c + lane) + q];
                    }
                    __builtin_amdgcn_wave_barrier();
                    nn -= c;
                }
                if (np >= LF) {
                    const int c = np & ~(LF - 1);
                    flush_parts<T, TPL, FIELDS, QUAD, EXACT, LF>(L, c, acc, tx, ty, tz, th);
                    __builtin_amdgcn_wave_barrier();
                    if (lane < np - c) {  // disjoint ranges, as above
                        L.parts[lane] = L.parts[c + lane];
                        L.ph[lane] = L.ph[c + lane];
                    }
                    __builtin_amdgcn_wave_barrier();
                    np -= c;
                }
            }
            if (overflow) {
                if (lane == 0) atomicOr(&A.ctl->err, 1);
                continue;
            }
            flush_nodes<T, TPL, FIELDS, QUAD, EXACT, LF>(L, nn, acc, tx, ty, tz, th);
            flush_parts<T, TPL, FIELDS, QUAD, EXACT, LF>(L, np, acc, tx, ty, tz, th);
            __builtin_amdgcn_wave_barrier();
            if constexpr (EXP) {
                exp_wave_sum(E);
#pragma unroll
                for (int j = 0; j < TPL; j++) exp_eval<T, FIELDS>(acc[j], E, tx[j] - cx, ty[j] - cy, tz[j] - cz);
            }

            // ---- write G * fields in the caller's order ----
            constexpr int W = NW<FIELDS>::W;
#pragma unroll
            for (int j = 0; j < TPL; j++) {
                if (!tval[j]) continue;
                double *o = A.out + tidx[j] * W;
                int c = 0;
                if (NW<FIELDS>::POT) o[c++] = A.G * (double)acc[j].phi;
                if (NW<FIELDS>::ACC) {
                    o[c++] = A.G * (double)acc[j].gx;
                    o[c++] = A.G * (double)acc[j].gy;
                    o[c++] = A.G * (double)acc[j].gz;
                }
                if (NW<FIELDS>::TID) {
                    double xx = A.G * (double)acc[j].txx, yy = A.G * (double)acc[j].tyy, zz = A.G * (double)acc[j].tzz;
                    double xy = A.G * (double)acc[j].txy, xz = A.G * (double)acc[j].txz, yz = A.G * (double)acc[j].tyz;
                    o[c++] = xx; o[c++] = xy; o[c++] = xz;
                    o[c++] = xy; o[c++] = yy; o[c++] = yz;
                    o[c++] = xz; o[c++] = yz; o[c++] = zz;
                }
            }
        }
    }
    if (lane == 0 && ghw > 0) atomicMax(&A.ctl->hwm, ghw);
    if (lane == 0 && A.counters) {
        atomicAdd((unsigned long long *)&A.counters[0], (unsigned long long)n_int);
        atomicAdd((unsigned long long *)&A.counters[1], (unsigned long long)n_vis);
        atomicAdd((unsigned long long *)&A.counters[2], (unsigned long long)n_part);
    }
    if (EXP && A.counters) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) n_exp += __shfl_xor(n_exp, d, 64);  // lanes fold separately
        if (lane == 0) atomicAdd((unsigned long long *)&A.counters[3], (unsigned long long)n_exp);
    }
}


__global__ void k_walk_init(WalkCtl *ctl)
{
    for (int i = threadIdx.x; i < (int)(sizeof(WalkCtl) / 4); i += blockDim.x) ((int32_t *)ctl)[i] = 0;
}

// ============================================================================================
// host side
// ============================================================================================
static int g_num_cus = 0;

static int num_cus()
{
    if (g_num_cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) == hipSuccess) g_num_cus = p.multiProcessorCount;
        if (g_num_cus <= 0) g_num_cus = 256;
    }
    return g_num_cus;
}

#define WALK_MAX_BLOCKS_PER_CU 6  // the most any variant runs (variant 2)

// persistent grid: resident blocks per CU from the occupancy API, capped by the variant's plan and
// by the spill allocation
static int64_t walk_grid_blocks(int variant = -1)
{
    const int per = variant == 2 ? 6 : variant < 0 ? WALK_MAX_BLOCKS_PER_CU : 5;
    return (int64_t)num_cus() * per;
}

extern "C" size_t ptg_walk_workspace_bytes(int64_t nt, int group_size)
{
    (void)nt;
    (void)group_size;
    size_t waves = (size_t)walk_grid_blocks() * WALK_WAVES;
    return ptg_align256(sizeof(WalkCtl)) + ptg_align256(waves * GCAP * sizeof(int));
}

// ---- static scratch of ptg_walk_noalloc ----
// Callers that cannot allocate (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc) walk with this file-scope array as the workspace: the walk control block (per-XCD group
// counters, overflow flag) and the per-wave global spill stacks.  Every GPU that runs the library
// holds its own copy until the process ends; every call on a device shares it, so calls that
// overlap on one device (different streams) race.  Host code never passes the array's name to a
// kernel (that is a host-side stand-in): it uses the address hipGetSymbolAddress returns, looked up
// once per device with that device current and kept per device.
#define NOALLOC_BYTES ((80u << 20) + 4096u)  // 4096 spill entries per wave for 5120 resident waves
#define NOALLOC_MIN_GCAP 2048  // spill entries per wave the static scratch must leave
#define NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) char scratch[NOALLOC_BYTES];
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];

static ptg_status noalloc_scratch(void **p)
{
    int dev = 0;
    PTG_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        ptg_set_error("ptg_walk_noalloc: device %d beyond the %d the library tracks", dev, NOALLOC_MAX_DEVICES);
        return PTG_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            ptg_set_error("ptg_walk_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                          hipGetErrorString(e));
            return PTG_ERR_HIP;
        }
        g_scratch_ptr[dev] = ptr;
    }
    *p = g_scratch_ptr[dev];
    return PTG_OK;
}

// spill entries per wave the static scratch leaves for `blocks` resident blocks (a multiple of 64,
// at most GCAP)
static int noalloc_gcap(int64_t blocks)
{
    const int64_t waves = blocks * WALK_WAVES;
    int64_t per = (int64_t)(NOALLOC_BYTES - ptg_align256(sizeof(WalkCtl))) / (waves * (int64_t)sizeof(int));
    per = std::min<int64_t>(per, GCAP) & ~(int64_t)63;
    return (int)per;
}

__global__ void k_walk_err_out(const WalkCtl *ctl, int32_t *err)
{
    if (threadIdx.x == 0) *err = ctl->err;
}

template <typename T, int TPL, int FIELDS, bool QUAD, bool EXACT, bool EXP, int VAR = 0>
static ptg_status launch_t(const WalkArgs<T> &a, int64_t max_blocks, hipStream_t s)
{
    static int per_cu = 0;
    if (per_cu == 0) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_walk<T, TPL, FIELDS, QUAD, EXACT, EXP, VAR>,
                                                         64 * WALK_WAVES, 0) != hipSuccess ||
            nb < 1)
            nb = 1;
        per_cu = std::min(nb, WalkOcc<T, TPL, VAR>::blocks_per_cu);
    }
    int64_t blocks = std::min<int64_t>(max_blocks, (int64_t)num_cus() * per_cu);
    k_walk<T, TPL, FIELDS, QUAD, EXACT, EXP, VAR><<<(unsigned)blocks, 64 * WALK_WAVES, 0, s>>>(a);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

// the monopole potential/acceleration kernels of the one-target-per-lane FP32 walk come in
// WALK_NVARIANTS variants; every other kernel in variant 0
template <typename T, int TPL, int FIELDS, bool EXP>
static ptg_status launch_var(const WalkArgs<T> &a, int variant, int64_t blocks, hipStream_t s)
{
    if constexpr (sizeof(T) == 4 && TPL == 1) {
        if (variant == 1) return launch_t<T, TPL, FIELDS, false, false, EXP, 1>(a, blocks, s);
        if (variant == 2) return launch_t<T, TPL, FIELDS, false, false, EXP, 2>(a, blocks, s);
    }
    return launch_t<T, TPL, FIELDS, false, false, EXP, 0>(a, blocks, s);
}

template <typename T, int TPL, bool EXACT>
static ptg_status launch_fq(const WalkArgs<T> &a, int fields, bool quad, bool expand, int variant, int64_t blocks,
                            hipStream_t s)
{
    if constexpr (sizeof(T) == 4 && !EXACT) {
        if (expand) {  // checked by the caller: monopole, potential and/or acceleration
            if (fields == PTG_POT) return launch_var<T, TPL, PTG_POT, true>(a, variant, blocks, s);
            if (fields == PTG_ACC) return launch_var<T, TPL, PTG_ACC, true>(a, variant, blocks, s);
            return launch_var<T, TPL, PTG_POT | PTG_ACC, true>(a, variant, blocks, s);
        }
        if (!quad) {
            if (fields == PTG_POT) return launch_var<T, TPL, PTG_POT, false>(a, variant, blocks, s);
            if (fields == PTG_ACC) return launch_var<T, TPL, PTG_ACC, false>(a, variant, blocks, s);
            if (fields == (PTG_POT | PTG_ACC)) return launch_var<T, TPL, PTG_POT | PTG_ACC, false>(a, variant, blocks, s);
        }
    }
#define PTG_F(F)                                                              \
    case F:                                                                   \
        return quad ? launch_t<T, TPL, F, true, EXACT, false>(a, blocks, s)   \
                    : launch_t<T, TPL, F, false, EXACT, false>(a, blocks, s);
    switch (fields) {
        PTG_F(1) PTG_F(2) PTG_F(3) PTG_F(4) PTG_F(5) PTG_F(6) PTG_F(7)
    }
#undef PTG_F
    ptg_set_error("ptg_walk: fields must be a nonzero combination of PTG_POT|PTG_ACC|PTG_TIDAL");
    return PTG_ERR_ARG;
}

template <typename T, bool EXACT>
static ptg_status launch_walk(const WalkArgs<T> &a, int fields, bool quad, bool expand, int variant, int64_t blocks,
                              hipStream_t s)
{
    if (a.gs <= 64) return launch_fq<T, 1, EXACT>(a, fields, quad, expand, variant, blocks, s);
    if (a.gs <= 128) return launch_fq<T, 2, EXACT>(a, fields, quad, expand, variant, blocks, s);
    return launch_fq<T, 4, EXACT>(a, fields, quad, expand, variant, blocks, s);
}

// ptg_walk and ptg_walk_noalloc: ws/ws_bytes is the caller's workspace, or nullptr with noalloc set
// (the static scratch below; err_out then receives the overflow flag)
static ptg_status walk_impl(const ptg_walk_tree *wt, const ptg_forest *forest, const double *pos, const double *soft,
                            const int64_t *perm, int targets_sorted, int64_t nt, int fields, int quadrupole,
                            int group_size, const int32_t *group_offsets, const int32_t *group_count, double theta,
                            const ptg_criterion *criterion, double G, int exact, double *out, int64_t *counters,
                            void *ws, size_t ws_bytes, bool noalloc, int32_t *err_out, void *stream,
                            const int32_t *start_slot = nullptr)
{
    hipStream_t s = (hipStream_t)stream;
    if (!wt || nt < 0 || (nt > 0 && (!pos || !soft || !perm || !out))) {
        ptg_set_error("ptg_walk: bad arguments");
        return PTG_ERR_ARG;
    }
    if (forest && (forest->nsys < 1 || !forest->offsets || !forest->cubes)) {
        ptg_set_error("ptg_walk: a forest needs nsys >= 1, offsets and cubes");
        return PTG_ERR_ARG;
    }
    if (forest && forest->nsys > 1 && !group_offsets) {
        ptg_set_error("ptg_walk: a forest walk needs group offsets that respect system boundaries "
                      "(ptg_target_groups with the forest)");
        return PTG_ERR_ARG;
    }
    if (group_size < 1 || group_size > 256) {
        ptg_set_error("ptg_walk: group_size must be in [1, 256], got %d", group_size);
        return PTG_ERR_ARG;
    }
    if (!(theta > 0)) {
        ptg_set_error("ptg_walk: theta must be positive");
        return PTG_ERR_ARG;
    }
    if (quadrupole && (!wt->has_quad || !wt->quad)) {
        ptg_set_error("ptg_walk: quadrupole=True requires a walk tree packed with quadrupoles");
        return PTG_ERR_ARG;
    }
    if (fields < 1 || fields > 7) {
        ptg_set_error("ptg_walk: fields must be a nonzero combination of PTG_POT|PTG_ACC|PTG_TIDAL");
        return PTG_ERR_ARG;
    }
    if ((group_offsets == nullptr) != (group_count == nullptr)) {
        ptg_set_error("ptg_walk: group_offsets and group_count must be given together");
        return PTG_ERR_ARG;
    }
    const int crit = criterion ? criterion->criterion : PTG_CRIT_THETA;
    if (crit != PTG_CRIT_THETA && crit != PTG_CRIT_QERR && crit != PTG_CRIT_RADIUS) {
        ptg_set_error("ptg_walk: unknown criterion %d", crit);
        return PTG_ERR_ARG;
    }
    if (crit == PTG_CRIT_QERR && (exact || !(criterion->eps > 0))) {
        ptg_set_error("ptg_walk: the error-bounded criterion needs eps > 0 and the FP32 walk");
        return PTG_ERR_ARG;
    }
    if (crit == PTG_CRIT_RADIUS && (exact || wt->prec != PTG_F32)) {
        ptg_set_error("ptg_walk: the radius criterion runs on the FP32 walk with radii from ptg_open_radii");
        return PTG_ERR_ARG;
    }
    const double eps = crit == PTG_CRIT_QERR ? criterion->eps / 3.5 : 0.0;
    const bool expand = criterion && criterion->group_expansion != 0;
    if (expand) {
        if (exact || wt->prec != PTG_F32 || quadrupole || (fields & PTG_TIDAL)) {
            ptg_set_error("ptg_walk: the group expansion runs on the FP32 walk with monopole interactions, "
                          "for the potential and/or the acceleration");
            return PTG_ERR_ARG;
        }
        if (!(criterion->expansion_eps > 0)) {
            ptg_set_error("ptg_walk: the group expansion needs expansion_eps > 0");
            return PTG_ERR_ARG;
        }
        if (!(criterion->expansion_rel >= 0.0f)) {
            ptg_set_error("ptg_walk: expansion_rel must be >= 0 (0 = no relative cap)");
            return PTG_ERR_ARG;
        }
    }
    const double eps_exp = expand ? criterion->expansion_eps : 0.0;
    const double *eps_scale = expand ? criterion->eps_scale : nullptr;
    if (expand && criterion->group_expansion != 1) {
        ptg_set_error("ptg_walk: group_expansion must be 0 (off) or 1");
        return PTG_ERR_ARG;
    }
    const int variant = criterion ? criterion->variant : 0;
    if (variant < 0 || variant >= WALK_NVARIANTS) {
        ptg_set_error("ptg_walk: variant must be in [0, %d), got %d", WALK_NVARIANTS, variant);
        return PTG_ERR_ARG;
    }
    int64_t blocks = walk_grid_blocks(variant);
    int gcap = GCAP;
    if (noalloc) {
        if (!err_out) {
            ptg_set_error("ptg_walk_noalloc: err (a device int32 that receives the overflow flag) is required");
            return PTG_ERR_ARG;
        }
        ptg_status st = noalloc_scratch(&ws);
        if (st) return st;
        ws_bytes = NOALLOC_BYTES;
        gcap = noalloc_gcap(blocks);
        if (gcap < NOALLOC_MIN_GCAP) {
            ptg_set_error("ptg_walk_noalloc: %lld resident waves leave %d spill entries each in the %u-byte static "
                          "scratch (at least %d needed); use ptg_walk with a workspace",
                          (long long)(blocks * WALK_WAVES), gcap, NOALLOC_BYTES, NOALLOC_MIN_GCAP);
            return PTG_ERR_ARG;
        }
    }
    if (nt == 0) return PTG_OK;
    const size_t base_bytes = noalloc ? ws_bytes : ptg_walk_workspace_bytes(nt, group_size);
    const size_t need = base_bytes;
    if (ws_bytes < need || !ws) {
        ptg_set_error("ptg_walk: workspace too small (%zu < %zu)", ws_bytes, need);
        return PTG_ERR_WORKSPACE;
    }
    WalkCtl *ctl = (WalkCtl *)ws;
    int *spill = (int *)((char *)ws + ptg_align256(sizeof(WalkCtl)));
    const int64_t fn = forest ? forest->nsys : 0;
    const int64_t *fo = forest ? forest->offsets : nullptr;
    const double *fc = forest ? forest->cubes : nullptr;
    k_walk_init<<<1, 64, 0, s>>>(ctl);
    PTG_LAUNCH_CHECK();
    ptg_status st;
    if (wt->prec == PTG_F32) {
        WalkArgs<float> a{(const V4<float> *)wt->xyzm, (const V4<float> *)wt->aux, (const int2 *)wt->kids,
                          (const float *)wt->quad, pos, soft, perm, targets_sorted != 0, wt->nslots, nt, ptg_cdiv(nt, group_size), group_offsets, group_count,
                          group_size, (float)theta, (float)(1.0 / theta), crit, (float)eps, (float)eps_exp,
                          expand ? criterion->expansion_rel : 0.0f, eps_scale, G, wt->origin[0], wt->origin[1],
                          wt->origin[2],
                          fn, fo, fc, out, counters, ctl, spill, gcap, start_slot};
        if (exact) {
            ptg_set_error("ptg_walk: the exact branch set is evaluated in FP64; use prec=PTG_F64");
            return PTG_ERR_ARG;
        }
        st = launch_walk<float, false>(a, fields, quadrupole != 0, expand, variant, blocks, s);
    } else if (wt->prec == PTG_F64) {
        WalkArgs<double> a{(const V4<double> *)wt->xyzm, (const V4<double> *)wt->aux, (const int2 *)wt->kids,
                           (const double *)wt->quad, pos, soft, perm, targets_sorted != 0, wt->nslots, nt, ptg_cdiv(nt, group_size), group_offsets, group_count,
                           group_size, theta, 1.0 / theta, crit, eps, 0.0, 0.0, nullptr, G,
                           wt->origin[0], wt->origin[1], wt->origin[2],
                           fn, fo, fc, out, counters, ctl, spill, gcap, start_slot};
        st = launch_walk<double, true>(a, fields, quadrupole != 0, false, 0, blocks, s);
    } else {
        ptg_set_error("ptg_walk: prec must be PTG_F32 or PTG_F64");
        return PTG_ERR_ARG;
    }
    if (st == PTG_OK && noalloc) {
        k_walk_err_out<<<1, 64, 0, s>>>(ctl, err_out);
        PTG_LAUNCH_CHECK();
    }
    return st;
}

extern "C" ptg_status ptg_walk(const ptg_walk_tree *wt, const ptg_forest *forest, const double *pos,
                               const double *soft, const int64_t *perm, int targets_sorted, int64_t nt, int fields,
                               int quadrupole, int group_size, const int32_t *group_offsets,
                               const int32_t *group_count, double theta, const ptg_criterion *criterion, double G,
                               int exact, double *out, int64_t *counters, void *ws, size_t ws_bytes, void *stream)
{
    return walk_impl(wt, forest, pos, soft, perm, targets_sorted, nt, fields, quadrupole, group_size, group_offsets,
                     group_count, theta, criterion, G, exact, out, counters, ws, ws_bytes, false, nullptr, stream);
}

extern "C" ptg_status ptg_walk_noalloc(const ptg_walk_tree *wt, const ptg_forest *forest, const double *pos,
                                       const double *soft, const int64_t *perm, int targets_sorted, int64_t nt,
                                       int fields, int quadrupole, int group_size, const int32_t *group_offsets,
                                       const int32_t *group_count, double theta, const ptg_criterion *criterion,
                                       double G, int exact, double *out, int64_t *counters, int32_t *err,
                                       void *stream)
{
    return walk_impl(wt, forest, pos, soft, perm, targets_sorted, nt, fields, quadrupole, group_size, group_offsets,
                     group_count, theta, criterion, G, exact, out, counters, nullptr, 0, true, err, stream);
}

// every target walks the subtree of one element (treewalk.py:78-277 with no >= 0, stopping at
// subtree_end :28-34): the FP64 walk with the reference's branches, one target per group
extern "C" ptg_status ptg_walk_subtree(const ptg_walk_tree *wt, const int32_t *start_slot, const double *pos,
                                       const double *soft, const int64_t *perm, int64_t nt, int fields,
                                       int quadrupole, double theta, double G, double *out, int64_t *counters,
                                       void *ws, size_t ws_bytes, void *stream)
{
    if (!wt || !start_slot) {
        ptg_set_error("ptg_walk_subtree: bad arguments (walk tree and start slot are required)");
        return PTG_ERR_ARG;
    }
    if (wt->prec != PTG_F64) {
        ptg_set_error("ptg_walk_subtree: the per-target walks run on the FP64 walk tree (prec=PTG_F64)");
        return PTG_ERR_ARG;
    }
    return walk_impl(wt, nullptr, pos, soft, perm, 0, nt, fields, quadrupole, 1, nullptr, nullptr, theta, nullptr, G,
                     1, out, counters, ws, ws_bytes, false, nullptr, stream, start_slot);
}

// overflow flag of the last walk on this workspace (0 = fine); synchronises the stream
extern "C" ptg_status ptg_walk_check(void *ws, void *stream)
{
    int32_t err = 0;
    hipStream_t s = (hipStream_t)stream;
    PTG_HIP_CHECK(hipMemcpyAsync(&err, &((WalkCtl *)ws)->err, 4, hipMemcpyDeviceToHost, s));
    PTG_HIP_CHECK(hipStreamSynchronize(s));
    if (err & 2) {
        ptg_set_error("ptg_walk: a supplied target group holds more than group_size targets (its rows past "
                      "group_size were not evaluated)");
        return PTG_ERR_ARG;
    }
    if (err & 4) {
        ptg_set_error("ptg_walk_subtree: the start element's slot is outside the walk tree (nothing was walked)");
        return PTG_ERR_ARG;
    }
    if (err) {
        ptg_set_error("ptg_walk: a traversal stack exceeded its spill capacity");
        return PTG_ERR_OVERFLOW;
    }
    return PTG_OK;
}
