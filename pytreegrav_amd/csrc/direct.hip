// Direct summation (bruteforce.py): one lane per target, the sources split into slices over the waves
// of a block and read at wave-uniform addresses (every lane of a wave reads the same source,
// broadcast through the scalar cache).  FP64 is the oracle-grade path used to measure tree errors;
// FP32 is the fast brute-force method.
#include <algorithm>

#include "ptg_common.hpp"

#define DS_SLICES 16  // waves per block, each summing one contiguous slice of the sources

template <typename T> __device__ __forceinline__ T d_force_kernel(T r, T h)
{
    if (r > h) return T(1) / (r * r * r);
    T hinv = T(1) / h;
    T q = r * hinv;
    if (q <= T(0.5)) return (T(10.666666666666666666) + q * q * (T(-38.4) + T(32.0) * q)) * hinv * hinv * hinv;
    return (T(21.333333333333) - T(48.0) * q + T(38.4) * q * q - T(10.666666666667) * q * q * q -
            T(0.066666666667) / (q * q * q)) *
           hinv * hinv * hinv;
}

template <typename T> __device__ __forceinline__ T d_tidal_kernel(T r, T h)
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

template <typename T> __device__ __forceinline__ T d_potential_kernel(T r, T h)
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

// One block per 64 targets (one per lane); its DS_SLICES waves split the sources into contiguous
// slices, so even a few thousand targets keep every CU busy.  A wave's lanes read the same source at
// each step (a wave-uniform address, served by the scalar cache).  The slices' partial sums are
// added in slice order through LDS, so results do not depend on scheduling.
template <typename T, int FIELDS, bool SELF>
__global__ __launch_bounds__(64 * DS_SLICES) void k_direct(const double *pt, const double *ht, int64_t nt,
                                                           const double *ps, const double *ms, const double *hs,
                                                           int64_t ns, double G, double *out)
{
    constexpr bool P = FIELDS & PTG_POT, A = FIELDS & PTG_ACC, D = FIELDS & PTG_TIDAL;
    constexpr int W = (P ? 1 : 0) + (A ? 3 : 0) + (D ? 6 : 0);  // accumulators (tidal: 6 independent)
    __shared__ T part[DS_SLICES - 1][W][64];
    // the wave index (and so every source index below) is wave-uniform: scalar loads
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = ptg_lane();
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    const bool vt = i < nt;
    const T tx = vt ? (T)pt[3 * i] : T(0), ty = vt ? (T)pt[3 * i + 1] : T(0), tz = vt ? (T)pt[3 * i + 2] : T(0);
    const T th = vt ? (T)ht[i] : T(0);
    T phi = 0, ax = 0, ay = 0, az = 0, txx = 0, tyy = 0, tzz = 0, txy = 0, txz = 0, tyz = 0;
    const int64_t j0 = ns * wave / DS_SLICES, j1 = ns * (wave + 1) / DS_SLICES;
#pragma unroll 4
    for (int64_t j = j0; j < j1; j++) {
        if (SELF && j == i) continue;
        // dx = source - target (the potential is even in dx)
        const T dx = (T)ps[3 * j] - tx, dy = (T)ps[3 * j + 1] - ty, dz = (T)ps[3 * j + 2] - tz;
        const T r2 = dx * dx + dy * dy + dz * dz;
        const T hsj = (T)hs[j];
        const T h = hsj > th ? hsj : th;
        const T m = (T)ms[j];
        const T r = sqrt(r2);
        if (P) {
            if (SELF) {
                // bruteforce.py:110-115: coincident distinct particles keep the softened term
                if (r < h) phi += m * d_potential_kernel(r, h);
                else if (r > T(0)) phi -= m / r;
            } else if (r != T(0)) {
                if (r < h) phi += m * d_potential_kernel(r, h);
                else phi -= m / r;
            }
        }
        if ((A || D) && r2 != T(0)) {
            T K, Q = 0;
            if (r < h) {
                K = m * d_force_kernel(r, h);
                if (D) Q = m * d_tidal_kernel(r, h);
            } else {
                K = m / (r2 * r);
                if (D) Q = T(3) * K / r2;
            }
            if (A) {
                ax += K * dx;
                ay += K * dy;
                az += K * dz;
            }
            if (D) {
                txx += Q * dx * dx - K;
                tyy += Q * dy * dy - K;
                tzz += Q * dz * dz - K;
                txy += Q * dx * dy;
                txz += Q * dx * dz;
                tyz += Q * dy * dz;
            }
        }
    }
    T acc[W];
    {
        int c = 0;
        if (P) acc[c++] = phi;
        if (A) {
            acc[c++] = ax;
            acc[c++] = ay;
            acc[c++] = az;
        }
        if (D) {
            acc[c++] = txx;
            acc[c++] = tyy;
            acc[c++] = tzz;
            acc[c++] = txy;
            acc[c++] = txz;
            acc[c++] = tyz;
        }
    }
    if (wave > 0) {
#pragma unroll
        for (int c = 0; c < W; c++) part[wave - 1][c][lane] = acc[c];
    }
    __syncthreads();
    if (wave > 0 || !vt) return;
#pragma unroll
    for (int w = 0; w < DS_SLICES - 1; w++) {
#pragma unroll
        for (int c = 0; c < W; c++) acc[c] += part[w][c][lane];
    }
    constexpr int WO = (P ? 1 : 0) + (A ? 3 : 0) + (D ? 9 : 0);
    double *o = out + i * WO;
    int c = 0, k = 0;
    if (P) o[c++] = G * (double)acc[k++];
    if (A) {
        o[c++] = G * (double)acc[k++];
        o[c++] = G * (double)acc[k++];
        o[c++] = G * (double)acc[k++];
    }
    if (D) {
        double xx = G * (double)acc[k], yy = G * (double)acc[k + 1], zz = G * (double)acc[k + 2];
        double xy = G * (double)acc[k + 3], xz = G * (double)acc[k + 4], yz = G * (double)acc[k + 5];
        o[c++] = xx; o[c++] = xy; o[c++] = xz;
        o[c++] = xy; o[c++] = yy; o[c++] = yz;
        o[c++] = xz; o[c++] = yz; o[c++] = zz;
    }
}

template <typename T, bool SELF>
static ptg_status launch_direct(int fields, const double *pt, const double *ht, int64_t nt, const double *ps,
                                const double *ms, const double *hs, int64_t ns, double G, double *out, hipStream_t s)
{
    dim3 grid((unsigned)std::max<int64_t>(1, ptg_cdiv(nt, 64)));
#define PTG_D(F)                                                                                           \
    case F:                                                                                                \
        k_direct<T, F, SELF><<<grid, 64 * DS_SLICES, 0, s>>>(pt, ht, nt, ps, ms, hs, ns, G, out);          \
        break;
    switch (fields) {
        PTG_D(1) PTG_D(2) PTG_D(3) PTG_D(4) PTG_D(5) PTG_D(6) PTG_D(7)
    default:
        ptg_set_error("ptg_direct: fields must be a nonzero combination of PTG_POT|PTG_ACC|PTG_TIDAL");
        return PTG_ERR_ARG;
    }
#undef PTG_D
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

extern "C" ptg_status ptg_direct(const double *pos_t, const double *soft_t, int64_t nt, const double *pos_s,
                                 const double *m_s, const double *h_s, int64_t ns, int prec, int fields, int self_mode,
                                 double G, double *out, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (nt < 0 || ns < 0 || (nt > 0 && (!pos_t || !soft_t || !out)) || (ns > 0 && (!pos_s || !m_s || !h_s))) {
        ptg_set_error("ptg_direct: bad arguments");
        return PTG_ERR_ARG;
    }
    if (self_mode && nt != ns) {
        ptg_set_error("ptg_direct: self_mode needs nt == ns");
        return PTG_ERR_ARG;
    }
    if (nt == 0) return PTG_OK;
    if (prec == PTG_F64)
        return self_mode ? launch_direct<double, true>(fields, pos_t, soft_t, nt, pos_s, m_s, h_s, ns, G, out, s)
                         : launch_direct<double, false>(fields, pos_t, soft_t, nt, pos_s, m_s, h_s, ns, G, out, s);
    if (prec == PTG_F32)
        return self_mode ? launch_direct<float, true>(fields, pos_t, soft_t, nt, pos_s, m_s, h_s, ns, G, out, s)
                         : launch_direct<float, false>(fields, pos_t, soft_t, nt, pos_s, m_s, h_s, ns, G, out, s);
    ptg_set_error("ptg_direct: prec must be PTG_F32 or PTG_F64");
    return PTG_ERR_ARG;
}
