// Tree build for gfx950: bounding cube, Morton keys, LSD radix sort, collision re-keying,
// Karras-style node emission (octree levels from common key prefixes), links, FP64 moments and
// the sibling-contiguous walk pack.
//
// Compiled with -ffp-contract=off: keys, node geometry and FP64 moments reproduce the reference's
// IEEE arithmetic bit for bit (octree.py builds these without fastmath).
#pragma clang fp contract(off)

#include <stdarg.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "ptg_common.hpp"

// ============================================================================================
// error reporting
// ============================================================================================
static thread_local char g_err[1024] = "";

void ptg_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char *ptg_last_error(void) { return g_err; }

// ---- optional phase timing (no reference counterpart) ----
// Events a caller registers here are recorded on the build's stream at the phase boundaries of the
// next ptg_build_sort / ptg_build_tree calls of this thread (see PTG_PHASE_* in ptg.h).
static thread_local void *g_phase_ev[PTG_PHASE_COUNT];

extern "C" ptg_status ptg_build_phase_events(void *const *events, int nevents)
{
    if (nevents != 0 && (nevents != PTG_PHASE_COUNT || !events)) {
        ptg_set_error("ptg_build_phase_events: pass %d events, or 0 to stop recording", PTG_PHASE_COUNT);
        return PTG_ERR_ARG;
    }
    for (int i = 0; i < PTG_PHASE_COUNT; i++) g_phase_ev[i] = nevents ? events[i] : nullptr;
    return PTG_OK;
}

static void phase_mark(int i, hipStream_t s)
{
    if (g_phase_ev[i]) (void)hipEventRecord((hipEvent_t)g_phase_ev[i], s);
}
extern "C" int ptg_version(void) { return 10000; }

// ============================================================================================
// generic exclusive scan of int32 values produced by a loader functor
// ============================================================================================
// exclusive scans of n int32 values given by a load functor: per-tile sums, one workgroup scanning
// the tile sums, and a pass that rescans each tile from its offset.  Tiles of SC_TILE values are
// read striped (value base + q * SC_THREADS + thread, coalesced per load); the rescan moves them
// through LDS to contiguous runs of SC_ITEMS per thread and back.
#define SC_THREADS 256
#define SC_ITEMS 16
#define SC_TILE (SC_THREADS * SC_ITEMS)
#define SC_PAD(j) ((j) + ((j) >> 4))  // one pad word per 16: a thread's run of 16 starts on its own bank

template <typename T> __device__ __forceinline__ T block_excl_scan(T v, T *total, T *sh)
{
    // sh: at least 4 entries
    int lane = ptg_lane(), wave = threadIdx.x >> 6;
    T incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    if (lane == 63) sh[wave] = incl;
    __syncthreads();
    T woff = 0, tot = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
        T s = sh[w];
        if (w < wave) woff += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return woff + incl - v;
}

template <typename Load>
__global__ __launch_bounds__(SC_THREADS) void k_scan_reduce(Load ld, int64_t n, int32_t *bsum, const int32_t *gate)
{
    if (gate && *gate == 0) return;
    __shared__ int32_t sh[8];
    const int64_t base = (int64_t)blockIdx.x * SC_TILE + threadIdx.x;
    int32_t s = 0;
#pragma unroll
    for (int q = 0; q < SC_ITEMS; q++) {
        const int64_t i = base + (int64_t)q * SC_THREADS;
        if (i < n) s += ld(i);
    }
    int32_t tot;
    block_excl_scan<int32_t>(s, &tot, sh);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// single workgroup: exclusive scan of nb tile sums in place (each thread a contiguous slice);
// grand total to *total
__global__ __launch_bounds__(SC_THREADS) void k_scan_top(int32_t *bsum, int64_t nb, int32_t *total,
                                                      const int32_t *gate)
{
    if (gate && *gate == 0) return;
    __shared__ int32_t sh[8];
    const int64_t per = (nb + SC_THREADS - 1) / SC_THREADS;
    const int64_t a = (int64_t)threadIdx.x * per, b = std::min<int64_t>(a + per, nb);
    int32_t s = 0;
    for (int64_t i = a; i < b; i++) s += bsum[i];
    int32_t tot;
    int32_t run = block_excl_scan<int32_t>(s, &tot, sh);
    for (int64_t i = a; i < b; i++) {
        const int32_t v = bsum[i];
        bsum[i] = run;
        run += v;
    }
    if (threadIdx.x == 0 && total) *total = tot;
}

template <typename Load>
__global__ __launch_bounds__(SC_THREADS) void k_scan_down(Load ld, int64_t n, const int32_t *bsum, int32_t *out,
                                                       const int32_t *gate)
{
    if (gate && *gate == 0) return;
    __shared__ int32_t sh[8];
    __shared__ int32_t tile[SC_PAD(SC_TILE)];
    const int64_t base = (int64_t)blockIdx.x * SC_TILE;
#pragma unroll
    for (int q = 0; q < SC_ITEMS; q++) {
        const int j = q * SC_THREADS + threadIdx.x;
        tile[SC_PAD(j)] = base + j < n ? ld(base + j) : 0;
    }
    __syncthreads();
    const int j0 = threadIdx.x * SC_ITEMS;
    int32_t v[SC_ITEMS];
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        v[k] = tile[SC_PAD(j0 + k)];
        s += v[k];
    }
    int32_t tot;
    int32_t ex = block_excl_scan<int32_t>(s, &tot, sh) + bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        tile[SC_PAD(j0 + k)] = ex;
        ex += v[k];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SC_ITEMS; q++) {
        const int j = q * SC_THREADS + threadIdx.x;
        if (base + j < n) out[base + j] = tile[SC_PAD(j)];
    }
}

static int64_t scan_blocks(int64_t n) { return std::max<int64_t>(1, ptg_cdiv(n, SC_TILE)); }

// gate (nullable, device): when it reads 0 the scan does nothing (out and *total keep their contents)
template <typename Load>
static ptg_status excl_scan(Load ld, int64_t n, int32_t *out, int32_t *bsum, int32_t *total, hipStream_t s,
                            const int32_t *gate = nullptr)
{
    int64_t nb = scan_blocks(n);
    k_scan_reduce<<<nb, SC_THREADS, 0, s>>>(ld, n, bsum, gate);
    k_scan_top<<<1, SC_THREADS, 0, s>>>(bsum, nb, total, gate);
    k_scan_down<<<nb, SC_THREADS, 0, s>>>(ld, n, bsum, out, gate);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

struct LoadArr {
    const int32_t *a;
    __device__ int32_t operator()(int64_t i) const { return a[i]; }
};

// ============================================================================================
// LSD radix sort of (u64 key, u32 value) pairs, 8-bit digits.  Each 256-thread block owns a tile of
// RS_TILE consecutive elements.  Per pass: a tile histogram, the offset of every (digit, tile) (one
// block per digit scans that digit's tile counts and writes the digit's total; each scatter block
// turns the 256 totals into digit starts itself), and a scatter that ranks the tile's keys stably
// (ballot-based digit matching inside each wave, per-digit prefix over the waves), reorders the
// tile in LDS and writes each digit's run
// contiguously, so global writes come in runs rather than one scattered element at a time.  (A
// single-pass decoupled look-back variant measured slower here, 116 against about 100 us per pass
// at 1e7 keys: the inclusive prefixes propagate slowly through the many resident tiles, and every
// look-back round pays a device-coherent load.)
// ============================================================================================
#define RS_BITS 8
#define RS_BINS 256
#define RS_ITEMS 8
#define RS_WAVES 4
#define RS_THREADS (64 * RS_WAVES)
#define RS_TILE (RS_THREADS * RS_ITEMS)

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t *keys, int64_t n, int shift, int64_t ntiles,
                                                        int32_t *hist)
{
    __shared__ uint32_t cnt[RS_BINS];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int r = 0; r < RS_ITEMS; r++) {
        int64_t i = base + r * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (RS_BINS - 1)], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = (int32_t)cnt[threadIdx.x];
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint64_t *kin, const uint32_t *vin, int64_t n,
                                                           int shift, int64_t ntiles, const int32_t *offs,
                                                           const int32_t *dtot, uint64_t *kout, uint32_t *vout)
{
    __shared__ uint64_t sk[RS_TILE];
    __shared__ uint32_t sv[RS_TILE];
    __shared__ uint32_t run[RS_WAVES][RS_BINS];
    __shared__ int64_t gbase[RS_BINS];
    __shared__ int32_t sh[8], sh2[8];
    const int wave = threadIdx.x >> 6, lane = ptg_lane();
    const uint64_t lt = ptg_lanemask_lt();
    const int64_t tile0 = (int64_t)blockIdx.x * RS_TILE;
    const int64_t base = tile0 + (int64_t)wave * (64 * RS_ITEMS);
#pragma unroll
    for (int q = 0; q < RS_BINS / 64; q++) run[wave][lane * (RS_BINS / 64) + q] = 0;
    __builtin_amdgcn_wave_barrier();
    uint64_t K[RS_ITEMS];
    uint32_t V[RS_ITEMS], DR[RS_ITEMS];  // digit << 16 | rank inside the wave's quarter
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const int64_t i = base + r * 64 + lane;
        const bool valid = i < n;
        const uint64_t k = valid ? kin[i] : 0;
        const uint32_t d = (uint32_t)(k >> shift) & (RS_BINS - 1);
        uint64_t peers = ptg_ballot(valid);
#pragma unroll
        for (int b = 0; b < RS_BITS; b++) {
            const bool bit = (d >> b) & 1;
            const uint64_t bal = ptg_ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t rank = run[wave][d] + (uint32_t)__popcll(peers & lt);
        __builtin_amdgcn_wave_barrier();
        if (valid && (peers & lt) == 0) run[wave][d] = rank + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        K[r] = k;
        V[r] = valid ? vin[i] : 0;
        DR[r] = valid ? ((d << 16) | rank) : 0xffffffffu;
    }
    __syncthreads();
    // per digit: tile-local start, offset of each wave's quarter, global base of the digit's run
    {
        const int d = threadIdx.x;
        uint32_t c[RS_WAVES], tot = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; w++) {
            c[w] = run[w][d];
            tot += c[w];
        }
        int32_t all, keys_all;
        const int32_t start = block_excl_scan<int32_t>((int32_t)tot, &all, sh);
        // the digit's start in the output: exclusive scan of the digit totals (every block redoes
        // this 256-entry scan instead of a separate launch)
        const int32_t dstart = block_excl_scan<int32_t>(dtot[d], &keys_all, sh2);
        gbase[d] = (int64_t)dstart + offs[(int64_t)d * ntiles + blockIdx.x] - start;
        uint32_t o = (uint32_t)start;
#pragma unroll
        for (int w = 0; w < RS_WAVES; w++) {
            run[w][d] = o;
            o += c[w];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        if (DR[r] == 0xffffffffu) continue;
        const uint32_t p = run[wave][DR[r] >> 16] + (DR[r] & 0xffffu);
        sk[p] = K[r];
        sv[p] = V[r];
    }
    __syncthreads();
    const int64_t cnt = n - tile0 < RS_TILE ? n - tile0 : RS_TILE;
    for (int64_t p = threadIdx.x; p < cnt; p += RS_THREADS) {
        const uint64_t k = sk[p];
        const int64_t g = gbase[(k >> shift) & (RS_BINS - 1)] + p;
        kout[g] = k;
        vout[g] = sv[p];
    }
}

// number of sort tiles (sizes the histogram array: RS_BINS entries per tile)
static int64_t rs_nwt(int64_t n) { return std::max<int64_t>(1, ptg_cdiv(n, RS_TILE)); }

struct RsWS {
    int32_t *hist;   // (RS_BINS, ntiles) tile counts, scanned in place per digit
    int32_t *dtot;   // (RS_BINS) digit totals
    int64_t cap;     // elements the buffers were carved for
};

static void carve_rs(PtgArena &A, int64_t n, RsWS &w)
{
    w.hist = A.take<int32_t>((int64_t)RS_BINS * rs_nwt(n));
    w.dtot = A.take<int32_t>(RS_BINS);
    w.cap = n;
}

// digit d = blockIdx.x: exclusive scan of hist[d][0..ntiles) in place, total into dtot[d]
__global__ __launch_bounds__(256) void k_rs_tile_scan(int32_t *hist, int64_t ntiles, int32_t *dtot)
{
    __shared__ int32_t sh[8];
    int32_t *row = hist + (int64_t)blockIdx.x * ntiles;
    const int64_t per = (ntiles + 255) / 256;
    const int64_t a = (int64_t)threadIdx.x * per, b = std::min<int64_t>(a + per, ntiles);
    int32_t sum = 0;
    for (int64_t i = a; i < b; i++) sum += row[i];
    int32_t tot;
    int32_t run = block_excl_scan<int32_t>(sum, &tot, sh);
    for (int64_t i = a; i < b; i++) {
        const int32_t v = row[i];
        row[i] = run;
        run += v;
    }
    if (threadIdx.x == 0) dtot[blockIdx.x] = tot;
}

// sorts (k0,v0) by the low `nbits` bits; the result ends in (k0,v0) or (k1,v1): *in_first tells which
static ptg_status radix_sort(uint64_t *k0, uint32_t *v0, uint64_t *k1, uint32_t *v1, int64_t n, int nbits,
                             const RsWS &w, bool *in_first, hipStream_t s)
{
    if (n > w.cap) {
        ptg_set_error("radix_sort: %lld keys exceed the workspace's %lld", (long long)n, (long long)w.cap);
        return PTG_ERR_WORKSPACE;
    }
    const int64_t ntiles = rs_nwt(n);
    bool first = true;
    for (int shift = 0; shift < nbits; shift += RS_BITS) {
        uint64_t *ki = first ? k0 : k1, *ko = first ? k1 : k0;
        uint32_t *vi = first ? v0 : v1, *vo = first ? v1 : v0;
        k_rs_hist<<<ntiles, RS_THREADS, 0, s>>>(ki, n, shift, ntiles, w.hist);
        k_rs_tile_scan<<<RS_BINS, 256, 0, s>>>(w.hist, ntiles, w.dtot);
        k_rs_scatter<<<ntiles, RS_THREADS, 0, s>>>(ki, vi, n, shift, ntiles, w.hist, w.dtot, ko, vo);
        PTG_LAUNCH_CHECK();
        first = !first;
    }
    *in_first = first;
    return PTG_OK;
}

// ============================================================================================
// bounding cube (octree.py:539-574) and Morton keys (octree.py:69-98)
// ============================================================================================
#define BB_BLOCKS 1024
struct Geom {
    double center[3], size, cmin[3], inv;
};

__global__ __launch_bounds__(256) void k_bbox_partial(const double *pos, int64_t n, double *part)
{
    __shared__ double sh[6][4];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double x = pos[3 * i + k];
            lo[k] = fmin(lo[k], x);
            hi[k] = fmax(hi[k], x);
        }
    }
    int wave = threadIdx.x >> 6, lane = ptg_lane();
#pragma unroll
    for (int k = 0; k < 3; k++) {
        lo[k] = ptg_wave_min(lo[k]);
        hi[k] = ptg_wave_max(hi[k]);
    }
    if (lane == 0)
        for (int k = 0; k < 3; k++) {
            sh[k][wave] = lo[k];
            sh[3 + k][wave] = hi[k];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < 3; k++) {
            double a = sh[k][0], b = sh[3 + k][0];
            for (int w = 1; w < 4; w++) {
                a = fmin(a, sh[k][w]);
                b = fmax(b, sh[3 + k][w]);
            }
            part[6 * blockIdx.x + k] = a;
            part[6 * blockIdx.x + 3 + k] = b;
        }
    }
}

__global__ __launch_bounds__(256) void k_bbox_final(const double *part, int nb, Geom *g)
{
    __shared__ double sh[6][4];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int k = 0; k < 3; k++) {
            lo[k] = fmin(lo[k], part[6 * b + k]);
            hi[k] = fmax(hi[k], part[6 * b + 3 + k]);
        }
    int wave = threadIdx.x >> 6, lane = ptg_lane();
    for (int k = 0; k < 3; k++) {
        lo[k] = ptg_wave_min(lo[k]);
        hi[k] = ptg_wave_max(hi[k]);
    }
    if (lane == 0)
        for (int k = 0; k < 3; k++) {
            sh[k][wave] = lo[k];
            sh[3 + k][wave] = hi[k];
        }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 3; k++) {
        lo[k] = sh[k][0];
        hi[k] = sh[3 + k][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            lo[k] = fmin(lo[k], sh[k][w]);
            hi[k] = fmax(hi[k], sh[3 + k][w]);
        }
    }
    for (int k = 0; k < 3; k++) g->center[k] = 0.5 * (hi[k] + lo[k]);
    double sz = hi[0] - lo[0];
    if (hi[1] - lo[1] > sz) sz = hi[1] - lo[1];
    if (hi[2] - lo[2] > sz) sz = hi[2] - lo[2];
    g->size = sz;
    g->inv = sz > 0 ? 1.0 / sz : 0.0;
    for (int k = 0; k < 3; k++) g->cmin[k] = g->center[k] - 0.5 * sz;
}

__device__ __forceinline__ uint64_t morton_key(double x, double y, double z, const double *cmin, double inv)
{
    const double SCALE = 2097152.0;
    const int64_t MAXQ = (1 << 21) - 1;
    int64_t q[3];
    double p[3] = {x, y, z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double t = ((p[k] - cmin[k]) * inv) * SCALE;
        int64_t v = (int64_t)t;
        v = v < 0 ? 0 : (v > MAXQ ? MAXQ : v);
        q[k] = v;
    }
    return ptg_spread_bits(q[0]) | (ptg_spread_bits(q[1]) << 1) | (ptg_spread_bits(q[2]) << 2);
}

struct Counters {
    int32_t ncollide;     // adjacent pairs with equal 63-bit keys
    int32_t nnodes;       // total internal nodes
    int32_t nsub;         // particles in collision runs
    int32_t coincident;   // bit-identical extended keys exist
    int32_t max_level;
    int32_t nonfinite;    // some position, mass or softening is NaN or infinite
    int32_t badforest;    // forest offsets do not describe systems of >= 2 particles covering [0, n)
    int32_t eff_sub;      // deferred build: particles in collision runs, 0 if over its capacity
    int32_t overflow;     // deferred build: PTG_OVERFLOW_* bits
    int32_t pad[7];
};

// one particle's inputs in one 48-byte record, so that the Morton-order gather reads about one
// cache line per particle instead of one from each of the three input arrays
struct alignas(16) PartRec {
    double x, y, z, m, h, pad;
};

__device__ __forceinline__ void put_rec(PartRec *rec, int64_t i, double x, double y, double z, double m, double h)
{
    double2 *r = (double2 *)(rec + i);
    r[0] = make_double2(x, y);
    r[1] = make_double2(z, m);
    r[2] = make_double2(h, 0.0);
}

// Morton keys in input order; with m and h (and c), also the finiteness flag of every input and,
// with rec, the particle records
__global__ __launch_bounds__(256) void k_keys(const double *pos, const double *m, const double *h, int64_t n,
                                              const Geom *g, uint64_t *keys, uint32_t *vals, Counters *c,
                                              PartRec *rec)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (i < n) {
        const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        keys[i] = morton_key(x, y, z, g->cmin, g->inv);
        vals[i] = (uint32_t)i;
        if (c) {
            const double mi = m[i], hi = h[i];
            bad = !isfinite(x) || !isfinite(y) || !isfinite(z) || !isfinite(mi) || !isfinite(hi);
            if (rec) put_rec(rec, i, x, y, z, mi, hi);
        }
    }
    if (c && ptg_ballot(bad) && ptg_lane() == 0) c->nonfinite = 1;
}

// ---- forest: several independent systems built in one pass ----
__global__ __launch_bounds__(256) void k_sys_of(const int64_t *offsets, int64_t nsys, int64_t n, uint32_t *sys_in)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t lo = 0, hi = nsys;  // last s with offsets[s] <= i
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        if (offsets[mid] <= i) lo = mid;
        else hi = mid;
    }
    sys_in[i] = (uint32_t)lo;
}

// one block per system: its bounding cube (octree.py:539-574, per system)
__global__ __launch_bounds__(256) void k_bbox_sys(const double *pos, const int64_t *offsets, Geom *geoms)
{
    __shared__ double sh[6][4];
    const int64_t a = offsets[blockIdx.x], b = offsets[blockIdx.x + 1];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = a + threadIdx.x; i < b; i += blockDim.x)
        for (int k = 0; k < 3; k++) {
            double x = pos[3 * i + k];
            lo[k] = fmin(lo[k], x);
            hi[k] = fmax(hi[k], x);
        }
    int wave = threadIdx.x >> 6, lane = ptg_lane();
    for (int k = 0; k < 3; k++) {
        lo[k] = ptg_wave_min(lo[k]);
        hi[k] = ptg_wave_max(hi[k]);
    }
    if (lane == 0)
        for (int k = 0; k < 3; k++) {
            sh[k][wave] = lo[k];
            sh[3 + k][wave] = hi[k];
        }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 3; k++) {
        lo[k] = sh[k][0];
        hi[k] = sh[3 + k][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            lo[k] = fmin(lo[k], sh[k][w]);
            hi[k] = fmax(hi[k], sh[3 + k][w]);
        }
    }
    Geom *g = geoms + blockIdx.x;
    for (int k = 0; k < 3; k++) g->center[k] = 0.5 * (hi[k] + lo[k]);
    double sz = hi[0] - lo[0];
    if (hi[1] - lo[1] > sz) sz = hi[1] - lo[1];
    if (hi[2] - lo[2] > sz) sz = hi[2] - lo[2];
    g->size = sz;
    g->inv = sz > 0 ? 1.0 / sz : 0.0;
    for (int k = 0; k < 3; k++) g->cmin[k] = g->center[k] - 0.5 * sz;
}

__global__ __launch_bounds__(256) void k_keys_sys(const double *pos, const double *m, const double *h, int64_t n,
                                                  const Geom *geoms, const uint32_t *sys, uint64_t *keys,
                                                  uint64_t *kU, uint32_t *vals, PartRec *rec)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Geom *g = geoms + sys[i];
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    uint64_t k = morton_key(x, y, z, g->cmin, g->inv);
    keys[i] = k;
    kU[i] = k;
    vals[i] = (uint32_t)i;
    if (rec) put_rec(rec, i, x, y, z, m[i], h[i]);
}

__global__ __launch_bounds__(256) void k_gather_sys_key(const uint32_t *sys, const uint32_t *vals, int64_t n,
                                                        uint64_t *out)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) out[j] = vals[j] < n ? sys[vals[j]] : 0;  // vals[j] < n unless a sort gave up (reported)
}

__global__ __launch_bounds__(256) void k_gather_keys(const uint64_t *kU, const uint32_t *vals, int64_t n,
                                                     uint64_t *keys)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) keys[j] = vals[j] < n ? kU[vals[j]] : 0;  // vals[j] < n unless a sort gave up (reported)
}

// root cube of every system, as (centre x, y, z, side)
__global__ __launch_bounds__(256) void k_forest_cubes(const Geom *geoms, int64_t nsys, double *cubes)
{
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    for (int k = 0; k < 3; k++) cubes[4 * s + k] = geoms[s].center[k];
    cubes[4 * s + 3] = geoms[s].size;
}

// ============================================================================================
// build workspace
// ============================================================================================

__global__ __launch_bounds__(256) void k_check_finite(const double *pos, const double *m, const double *h, int64_t n,
                                                      Counters *c)
{
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        bad |= !isfinite(pos[3 * i]) || !isfinite(pos[3 * i + 1]) || !isfinite(pos[3 * i + 2]);
        if (m) bad |= !isfinite(m[i]);
        if (h) bad |= !isfinite(h[i]);
    }
    if (ptg_ballot(bad) && ptg_lane() == 0) c->nonfinite = 1;
}

struct BuildWS {
    double *part;
    Geom *geom;
    Counters *ctr;
    uint64_t *kA, *kB;
    uint32_t *vA, *vB;
    RsWS rs;
    int32_t *bsum;
    int32_t *lev, *off;
    // collision re-keying (extended keys K1..K3 per Morton slot, and the compact subset)
    uint64_t *ext;      // 3 * n
    int32_t *flag, *subpos;
    uint64_t *sk0, *sk1, *sk2, *sk3;
    uint32_t *sslot, *sval;
    uint64_t *tkA, *tkB;
    uint32_t *perm, *permB;
    // forest of several independent systems: per-system cubes, the system of each particle (system s
    // owns input indices [offsets[s], offsets[s+1]) and, after the per-system sort, the same range of
    // Morton slots, so one array serves both), keys before sorting
    Geom *geoms;
    uint32_t *sys;
    uint64_t *kU;
    const int64_t *soff;
    int64_t nsys;
    PartRec *rec;  // particle records in input order (written by the key pass when m and h are given)
};

static bool carve_build(PtgArena &A, int64_t n, BuildWS &w, int64_t nsys = 1)
{
    w.part = A.take<double>(6 * BB_BLOCKS);
    w.geom = A.take<Geom>(1);
    w.ctr = A.take<Counters>(1);
    w.kA = A.take<uint64_t>(n);
    w.kB = A.take<uint64_t>(n);
    w.vA = A.take<uint32_t>(n);
    w.vB = A.take<uint32_t>(n);
    carve_rs(A, n, w.rs);
    w.bsum = A.take<int32_t>(scan_blocks(n) + 64);
    w.lev = A.take<int32_t>(n);
    w.off = A.take<int32_t>(n);
    w.ext = A.take<uint64_t>(3 * n);
    w.flag = A.take<int32_t>(n);
    w.subpos = A.take<int32_t>(n);
    w.sk0 = A.take<uint64_t>(n);
    w.sk1 = A.take<uint64_t>(n);
    w.sk2 = A.take<uint64_t>(n);
    w.sk3 = A.take<uint64_t>(n);
    w.sslot = A.take<uint32_t>(n);
    w.sval = A.take<uint32_t>(n);
    w.tkA = A.take<uint64_t>(n);
    w.tkB = A.take<uint64_t>(n);
    w.perm = A.take<uint32_t>(n);
    w.permB = A.take<uint32_t>(n);
    w.nsys = nsys;
    w.soff = nullptr;
    w.geoms = A.take<Geom>(nsys);
    w.rec = A.take<PartRec>(n);
    if (nsys > 1) {
        w.sys = A.take<uint32_t>(n);
        w.kU = A.take<uint64_t>(n);
    } else {
        w.sys = nullptr;
        w.kU = nullptr;
    }
    return A.ok();
}

extern "C" size_t ptg_build_workspace_bytes(int64_t n, int64_t nsys)
{
    PtgArena A(nullptr, 0);
    BuildWS w;
    carve_build(A, std::max<int64_t>(n, 1), w, std::max<int64_t>(nsys, 1));
    return A.used;
}

// ============================================================================================
// pair levels, collision re-keying
// ============================================================================================
__global__ void k_zero_counters(Counters *c)
{
    int i = threadIdx.x;
    if (i < (int)(sizeof(Counters) / 4)) ((int32_t *)c)[i] = 0;
}

// lev[i] = whole octree levels shared by Morton slots i and i+1; -1 where slots i and i+1 belong to
// different systems of a forest (the pair then starts no node and the next pair starts a root)
// one block's sum and max of per-thread values, then one atomic of each per block
__device__ __forceinline__ void block_count_max(int cnt, int vmax, int32_t *sum_to, int32_t *max_to)
{
    __shared__ int ws[4], wm[4];
    cnt = ptg_wave_sum(cnt);
    vmax = ptg_wave_max(vmax);
    if (ptg_lane() == 0) {
        ws[threadIdx.x >> 6] = cnt;
        wm[threadIdx.x >> 6] = vmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int s = ws[0] + ws[1] + ws[2] + ws[3];
        const int m = max(max(wm[0], wm[1]), max(wm[2], wm[3]));
        if (sum_to && s) atomicAdd(sum_to, s);
        if (max_to && m > 0) atomicMax(max_to, m);
    }
}

#define LEV_BLOCKS 2048  // grid of the per-pair level passes (grid-stride; one atomic per block)
static unsigned lev_grid(int64_t n)
{
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(LEV_BLOCKS, ptg_cdiv(n - 1, 256)));
}

// shared levels of every adjacent pair, the count of colliding pairs and the deepest level
__global__ __launch_bounds__(256) void k_levels(const uint64_t *keys, const uint32_t *sys, int64_t n, int32_t *lev,
                                                Counters *c)
{
    int cnt = 0, vmax = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n - 1; i += (int64_t)gridDim.x * blockDim.x) {
        if (sys && sys[i] != sys[i + 1]) {
            lev[i] = -1;
        } else {
            const int L = ptg_common_levels(keys[i], keys[i + 1]);
            lev[i] = L;
            cnt += keys[i] == keys[i + 1];
            vmax = max(vmax, L);
        }
    }
    block_count_max(cnt, vmax, &c->ncollide, &c->max_level);
}

// slots whose 63-bit key is shared with a neighbour of the same system
__global__ __launch_bounds__(256) void k_collision_flags(const uint64_t *keys, const uint32_t *sys, int64_t n,
                                                         int32_t *flag, const int32_t *gate)
{
    if (gate && *gate == 0) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        bool f = (i > 0 && keys[i] == keys[i - 1] && (!sys || sys[i] == sys[i - 1])) ||
                 (i < n - 1 && keys[i] == keys[i + 1] && (!sys || sys[i] == sys[i + 1]));
        flag[i] = f ? 1 : 0;
    }
}

// extended keys for one particle: re-keying relative to its level-21, -42, -63 cells
// (RekeySlice, octree.py:1093-1126, with cell centres from the split's recurrence :974-978)
__device__ void extended_keys(const double *p, uint64_t k0, const Geom *g, uint64_t *k123)
{
    double c[3] = {g->center[0], g->center[1], g->center[2]};
    double s = g->size;
    uint64_t key = k0;
    for (int round = 0; round < 3; round++) {
        for (int L = 1; L <= 21; L++) {
            int d = (int)((key >> (63 - 3 * L)) & 7);
#pragma unroll
            for (int k = 0; k < 3; k++) c[k] = c[k] + s * ptg_octant_offset(d, k);
            s = s * 0.5;
        }
        double inv = s > 0 ? 1.0 / s : 0.0;
        double cmin[3];
        for (int k = 0; k < 3; k++) cmin[k] = c[k] - 0.5 * s;
        key = morton_key(p[0], p[1], p[2], cmin, inv);
        k123[round] = key;
    }
}

__global__ __launch_bounds__(256) void k_collision_gather(const double *pos, const uint64_t *keys,
                                                          const uint32_t *order, const int32_t *flag,
                                                          const int32_t *subpos, int64_t n, const Geom *geoms,
                                                          const uint32_t *sys, uint64_t *sk0, uint64_t *sk1,
                                                          uint64_t *sk2, uint64_t *sk3, uint32_t *sslot,
                                                          uint32_t *sval, uint32_t *perm, const int32_t *gate)
{
    if (gate && *gate == 0) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!flag[i]) continue;
        int64_t j = subpos[i];
        uint32_t o = order[i];
        double p[3] = {pos[3 * (int64_t)o], pos[3 * (int64_t)o + 1], pos[3 * (int64_t)o + 2]};
        uint64_t k3[3];
        extended_keys(p, keys[i], geoms + (sys ? sys[i] : 0), k3);
        sk0[j] = keys[i];
        sk1[j] = k3[0];
        sk2[j] = k3[1];
        sk3[j] = k3[2];
        sslot[j] = (uint32_t)i;
        sval[j] = o;
        perm[j] = (uint32_t)j;
    }
}

// small collision subsets: stable sort by (system, K0, K1, K2, K3) through rank counting, rank of j =
// number of subset elements ordered before it (index breaks ties, so the order is stable).  The
// subset is gathered in slot order, so its system ids are already non-decreasing: only elements of
// j's own system are compared.
__global__ __launch_bounds__(256) void k_rank_sort4(const uint64_t *k0, const uint64_t *k1, const uint64_t *k2,
                                                    const uint64_t *k3, const uint32_t *sslot, const uint32_t *sys,
                                                    int64_t m, uint32_t *perm, const int32_t *dm)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (dm) m = std::min<int64_t>(m, *dm);  // deferred build: the subset size is on the device
    if (j >= m) return;
    const uint64_t a0 = k0[j], a1 = k1[j], a2 = k2[j], a3 = k3[j];
    int64_t lo = 0, hi = m;
    if (sys) {
        const uint32_t sj = sys[sslot[j]];
        while (lo < j && sys[sslot[lo]] != sj) lo++;  // subsets here are small (<= 8192)
        hi = j + 1;
        while (hi < m && sys[sslot[hi]] == sj) hi++;
    }
    int64_t rank = lo;
    for (int64_t i = lo; i < hi; i++) {
        uint64_t b0 = k0[i], b1 = k1[i], b2 = k2[i], b3 = k3[i];
        bool before = b0 != a0 ? b0 < a0 : b1 != a1 ? b1 < a1 : b2 != a2 ? b2 < a2 : b3 != a3 ? b3 < a3 : i < j;
        rank += before;
    }
    perm[rank] = (uint32_t)j;
}

__global__ __launch_bounds__(256) void k_gather_sys_sub(const uint32_t *sys, const uint32_t *sslot,
                                                        const uint32_t *perm, int64_t m, uint64_t *dst)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m) dst[j] = sys[sslot[perm[j]]];
}

__global__ __launch_bounds__(256) void k_gather_u64(const uint64_t *src, const uint32_t *perm, int64_t m,
                                                    uint64_t *dst)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m) dst[j] = src[perm[j]];
}

// place the re-sorted collision subset back into its Morton slots
__global__ __launch_bounds__(256) void k_collision_apply(const uint32_t *perm, int64_t m, const uint32_t *sslot,
                                                         const uint32_t *sval, const uint64_t *sk1,
                                                         const uint64_t *sk2, const uint64_t *sk3, uint32_t *order,
                                                         uint64_t *ext, int64_t n, const int32_t *dm)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (dm) m = std::min<int64_t>(m, *dm);
    if (j >= m) return;
    uint32_t src = perm[j];
    int64_t slot = sslot[j];
    order[slot] = sval[src];
    ext[slot] = sk1[src];
    ext[n + slot] = sk2[src];
    ext[2 * n + slot] = sk3[src];
}

__device__ __forceinline__ bool ext_equal(const uint64_t *keys, const uint64_t *ext, int64_t n, int64_t a, int64_t b)
{
    return keys[a] == keys[b] && ext[a] == ext[b] && ext[n + a] == ext[n + b] && ext[2 * n + a] == ext[2 * n + b];
}

__device__ __forceinline__ bool ext_less(const uint64_t *keys, const uint64_t *ext, int64_t n, int64_t a, int64_t b)
{
    if (keys[a] != keys[b]) return keys[a] < keys[b];
    for (int r = 0; r < 3; r++)
        if (ext[r * n + a] != ext[r * n + b]) return ext[r * n + a] < ext[r * n + b];
    return false;
}

// levels of pairs inside collision runs: re-keyed levels 22..84, then bucket chains of 7
// (BuildBucket, octree.py:1128-1175: 7 particles and one deeper node per chain level)
__device__ __forceinline__ int level_ext(const uint64_t *keys, const uint64_t *ext, const uint32_t *sys,
                                         const int64_t *soff, int64_t n, int64_t i, Counters *c)
{
    int L;
    if (ext[i] != ext[i + 1]) L = 21 + ptg_common_levels(ext[i], ext[i + 1]);
    else if (ext[n + i] != ext[n + i + 1]) L = 42 + ptg_common_levels(ext[n + i], ext[n + i + 1]);
    else if (ext[2 * n + i] != ext[2 * n + i + 1]) L = 63 + ptg_common_levels(ext[2 * n + i], ext[2 * n + i + 1]);
    else {
        // identical to double precision: find where this run of identical extended keys starts
        int64_t lo = sys ? soff[sys[i]] : 0, hi = i;  // first index of the system with key == key[i]
        while (lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            if (ext_less(keys, ext, n, mid, i)) lo = mid + 1;
            else hi = mid;
        }
        L = 84 + (int)((i - lo) / 7);
        c->coincident = 1;
    }
    return L;
}

// the levels of pairs inside collision runs (grid-stride; the deepest level joins the counters' max)
__global__ __launch_bounds__(256) void k_levels_ext(const uint64_t *keys, const uint64_t *ext, const uint32_t *sys,
                                                    const int64_t *soff, int64_t n, int32_t *lev, Counters *c)
{
    if (c->ncollide == 0) return;
    int vmax = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n - 1; i += (int64_t)gridDim.x * blockDim.x) {
        if (keys[i] != keys[i + 1] || (sys && sys[i] != sys[i + 1])) continue;
        const int L = level_ext(keys, ext, sys, soff, n, i, c);
        lev[i] = L;
        vmax = max(vmax, L);
    }
    block_count_max(0, vmax, nullptr, &c->max_level);
}

struct LoadNodeCount {
    const int32_t *lev;
    __device__ int32_t operator()(int64_t i) const
    {
        int32_t p = i == 0 ? -1 : lev[i - 1];
        int32_t d = lev[i] - p;
        return d > 0 ? d : 0;
    }
};

__global__ void k_fill_i32(int32_t *a, int64_t n, int32_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_iota_u32(uint32_t *a, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = (uint32_t)i;
}

__global__ void k_u32_to_i64(const uint32_t *a, int64_t n, int64_t *b)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

static dim3 grid1(int64_t n, int bs = 256) { return dim3((unsigned)std::max<int64_t>(1, ptg_cdiv(n, bs))); }

// bbox + keys + sort; result (sorted keys, order) in (w.kA, w.vA).  With m and h, also the
// finiteness flag.
static ptg_status sort_by_keys(const double *pos, const double *m, const double *h, int64_t n, BuildWS &w,
                               hipStream_t s)
{
    int nb = (int)std::min<int64_t>(BB_BLOCKS, std::max<int64_t>(1, ptg_cdiv(n, 1024)));
    k_bbox_partial<<<nb, 256, 0, s>>>(pos, n, w.part);
    k_bbox_final<<<1, 256, 0, s>>>(w.part, nb, w.geom);
    if (m) phase_mark(PTG_PHASE_BBOX, s);
    k_keys<<<grid1(n), 256, 0, s>>>(pos, m, h, n, w.geom, w.kA, w.vA, m && h ? w.ctr : nullptr,
                                    m && h ? w.rec : nullptr);
    PTG_LAUNCH_CHECK();
    if (m) phase_mark(PTG_PHASE_KEYS, s);
    bool in_first;
    ptg_status st = radix_sort(w.kA, w.vA, w.kB, w.vB, n, 64, w.rs, &in_first, s);
    if (st) return st;
    if (!in_first) {
        PTG_HIP_CHECK(hipMemcpyAsync(w.kA, w.kB, n * 8, hipMemcpyDeviceToDevice, s));
        PTG_HIP_CHECK(hipMemcpyAsync(w.vA, w.vB, n * 4, hipMemcpyDeviceToDevice, s));
    }
    return PTG_OK;
}

static int sys_bits(int64_t nsys)
{
    int nbits = 1;
    while ((1ll << nbits) < nsys) nbits++;
    return ((nbits + RS_BITS - 1) / RS_BITS) * RS_BITS;
}

// forest: per-system cubes and keys, stable sort by key, then a stable sort by system, which leaves
// every system's particles in its own input range sorted by its own keys; result in (w.kA, w.vA)
static ptg_status sort_by_keys_forest(const double *pos, const double *m, const double *h, int64_t n, BuildWS &w,
                                      hipStream_t s)
{
    k_sys_of<<<grid1(n), 256, 0, s>>>(w.soff, w.nsys, n, w.sys);
    k_bbox_sys<<<(unsigned)w.nsys, 256, 0, s>>>(pos, w.soff, w.geoms);
    k_keys_sys<<<grid1(n), 256, 0, s>>>(pos, m, h, n, w.geoms, w.sys, w.kA, w.kU, w.vA, m && h ? w.rec : nullptr);
    PTG_LAUNCH_CHECK();
    bool in_first;
    ptg_status st = radix_sort(w.kA, w.vA, w.kB, w.vB, n, 64, w.rs, &in_first, s);
    if (st) return st;
    uint32_t *v = in_first ? w.vA : w.vB, *vo = in_first ? w.vB : w.vA;
    k_gather_sys_key<<<grid1(n), 256, 0, s>>>(w.sys, v, n, w.tkA);
    PTG_LAUNCH_CHECK();
    st = radix_sort(w.tkA, v, w.tkB, vo, n, sys_bits(w.nsys), w.rs, &in_first, s);
    if (st) return st;
    const uint32_t *vs = in_first ? v : vo;
    if (vs != w.vA) PTG_HIP_CHECK(hipMemcpyAsync(w.vA, vs, n * 4, hipMemcpyDeviceToDevice, s));
    k_gather_keys<<<grid1(n), 256, 0, s>>>(w.kU, w.vA, n, w.kA);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

// validates a forest descriptor against n; nsys_out = 1 for a plain tree
static ptg_status forest_check(const ptg_forest *f, int64_t n, const char *who, int64_t *nsys_out)
{
    *nsys_out = 1;
    if (!f) return PTG_OK;
    if (f->nsys < 1 || f->nsys > n / 2 || !f->offsets || !f->cubes || !f->roots) {
        ptg_set_error("%s: a forest needs 1 <= nsys <= n/2 and non-null offsets, cubes and roots", who);
        return PTG_ERR_ARG;
    }
    *nsys_out = f->nsys;
    return PTG_OK;
}

extern "C" ptg_status ptg_morton_order(const double *pos, int64_t n, int64_t *order, void *ws, size_t ws_bytes,
                                       void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (n < 0 || (n > 0 && (!pos || !order))) {
        ptg_set_error("ptg_morton_order: bad arguments");
        return PTG_ERR_ARG;
    }
    if (n == 0) return PTG_OK;
    if (n > 0x7fffffffLL) {
        ptg_set_error("ptg_morton_order: n exceeds the int32 index range");
        return PTG_ERR_ARG;
    }
    PtgArena A(ws, ws_bytes);
    BuildWS w;
    if (!ws || !carve_build(A, n, w)) {
        ptg_set_error("ptg_morton_order: workspace too small (%zu < %zu)", ws_bytes, ptg_build_workspace_bytes(n, 1));
        return PTG_ERR_WORKSPACE;
    }
    ptg_status st = sort_by_keys(pos, nullptr, nullptr, n, w, s);
    if (st) return st;
    k_u32_to_i64<<<grid1(n), 256, 0, s>>>(w.vA, n, order);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

// forest offsets: 0 = offsets[0] < offsets[1] - 1 < ... with offsets[nsys] = n (every system holds at
// least two particles)
__global__ __launch_bounds__(256) void k_check_offsets(const int64_t *off, int64_t nsys, int64_t n, Counters *c)
{
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nsys) return;
    bool bad = false;
    if (s == 0) bad = off[0] != 0;
    else bad = off[s] - off[s - 1] < 2;
    if (s == nsys) bad |= off[s] != n;
    if (bad) c->badforest = 1;
}

extern "C" ptg_status ptg_build_sort(const double *pos, const double *m, const double *h, int64_t n,
                                     const ptg_forest *forest, void *ws, size_t ws_bytes, ptg_build_info *info,
                                     void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (n < 1 || !pos || !m || !h || !info) {
        ptg_set_error("ptg_build_sort: need n >= 1 particles and non-null pos, m, h and info");
        return PTG_ERR_ARG;
    }
    if (n > 0x3fffffffLL) {
        ptg_set_error("ptg_build_sort: n=%lld exceeds the int32 element range", (long long)n);
        return PTG_ERR_ARG;
    }
    int64_t nsys;
    ptg_status st = forest_check(forest, n, "ptg_build_sort", &nsys);
    if (st) return st;
    PtgArena A(ws, ws_bytes);
    BuildWS w;
    if (!ws || !carve_build(A, n, w, nsys)) {
        ptg_set_error("ptg_build_sort: workspace too small (%zu < %zu)", ws_bytes, ptg_build_workspace_bytes(n, nsys));
        return PTG_ERR_WORKSPACE;
    }
    const bool F = nsys > 1;
    if (F) w.soff = forest->offsets;
    k_zero_counters<<<1, 64, 0, s>>>(w.ctr);
    Counters hc;
    auto nonfinite = [&]() {
        info->n = n;
        info->nnodes = 0;
        info->nonfinite = 1;
        ptg_set_error("ptg_build_sort: non-finite position, mass or softening");
        return PTG_ERR_ARG;
    };
    if (forest) {
        // forests: the offsets index the per-system kernels, so they (and the inputs) are checked
        // before anything else runs; a single tree checks finiteness inside the key pass instead
        k_check_finite<<<(unsigned)std::min<int64_t>(1024, ptg_cdiv(n, 256)), 256, 0, s>>>(pos, m, h, n, w.ctr);
        if (forest) k_check_offsets<<<grid1(nsys + 1), 256, 0, s>>>(forest->offsets, nsys, n, w.ctr);
        PTG_HIP_CHECK(hipMemcpyAsync(&hc, w.ctr, sizeof(Counters), hipMemcpyDeviceToHost, s));
        PTG_HIP_CHECK(hipStreamSynchronize(s));
        if (hc.nonfinite) return nonfinite();
        if (hc.badforest) {
            ptg_set_error("ptg_build_sort: forest offsets must start at 0, end at n and give every system >= 2 "
                          "particles");
            return PTG_ERR_ARG;
        }
    }
    phase_mark(PTG_PHASE_START, s);
    st = F ? sort_by_keys_forest(pos, m, h, n, w, s) : sort_by_keys(pos, m, h, n, w, s);
    if (st) return st;
    phase_mark(PTG_PHASE_SORT, s);
    if (!F) PTG_HIP_CHECK(hipMemcpyAsync(w.geoms, w.geom, sizeof(Geom), hipMemcpyDeviceToDevice, s));
    if (forest) {
        k_forest_cubes<<<grid1(nsys), 256, 0, s>>>(w.geoms, nsys, forest->cubes);
        PTG_LAUNCH_CHECK();
    }
    if (n >= 2) {
        k_levels<<<lev_grid(n), 256, 0, s>>>(w.kA, w.sys, n, w.lev, w.ctr);
        PTG_LAUNCH_CHECK();
        PTG_HIP_CHECK(hipMemcpyAsync(&hc, w.ctr, sizeof(Counters), hipMemcpyDeviceToHost, s));
        PTG_HIP_CHECK(hipStreamSynchronize(s));
        if (hc.nonfinite) return nonfinite();
        if (hc.ncollide > 0) {
            // collision runs: compute extended keys and re-sort each run by them (stable)
            k_collision_flags<<<grid1(n), 256, 0, s>>>(w.kA, w.sys, n, w.flag, nullptr);
            st = excl_scan(LoadArr{w.flag}, n, w.subpos, w.bsum, &w.ctr->nsub, s);
            if (st) return st;
            k_collision_gather<<<grid1(n), 256, 0, s>>>(pos, w.kA, w.vA, w.flag, w.subpos, n, w.geoms, w.sys,
                                                        w.sk0, w.sk1, w.sk2, w.sk3, w.sslot, w.sval, w.perm, nullptr);
            PTG_LAUNCH_CHECK();
            PTG_HIP_CHECK(hipMemcpyAsync(&hc, w.ctr, sizeof(Counters), hipMemcpyDeviceToHost, s));
            PTG_HIP_CHECK(hipStreamSynchronize(s));
            int64_t m = hc.nsub;
            if (m <= 8192) {
                k_rank_sort4<<<grid1(m), 256, 0, s>>>(w.sk0, w.sk1, w.sk2, w.sk3, w.sslot, w.sys, m, w.perm, nullptr);
                PTG_LAUNCH_CHECK();
            } else {
                uint64_t *karr[4] = {w.sk3, w.sk2, w.sk1, w.sk0};
                for (int r = 0; r < 4; r++) {
                    k_gather_u64<<<grid1(m), 256, 0, s>>>(karr[r], w.perm, m, w.tkA);
                    bool in_first;
                    st = radix_sort(w.tkA, w.perm, w.tkB, w.permB, m, 64, w.rs, &in_first, s);
                    if (st) return st;
                    if (!in_first)
                        PTG_HIP_CHECK(hipMemcpyAsync(w.perm, w.permB, m * 4, hipMemcpyDeviceToDevice, s));
                }
                if (F) {  // most significant: the system
                    k_gather_sys_sub<<<grid1(m), 256, 0, s>>>(w.sys, w.sslot, w.perm, m, w.tkA);
                    bool in_first;
                    st = radix_sort(w.tkA, w.perm, w.tkB, w.permB, m, sys_bits(nsys), w.rs, &in_first, s);
                    if (st) return st;
                    if (!in_first)
                        PTG_HIP_CHECK(hipMemcpyAsync(w.perm, w.permB, m * 4, hipMemcpyDeviceToDevice, s));
                }
            }
            PTG_HIP_CHECK(hipMemsetAsync(w.ext, 0, 3 * n * sizeof(uint64_t), s));
            k_collision_apply<<<grid1(m), 256, 0, s>>>(w.perm, m, w.sslot, w.sval, w.sk1, w.sk2, w.sk3, w.vA, w.ext,
                                                       n, nullptr);
            k_levels_ext<<<lev_grid(n), 256, 0, s>>>(w.kA, w.ext, w.sys, w.soff, n, w.lev, w.ctr);
            PTG_LAUNCH_CHECK();
        }
        st = excl_scan(LoadNodeCount{w.lev}, n - 1, w.off, w.bsum, &w.ctr->nnodes, s);
        if (st) return st;
        PTG_LAUNCH_CHECK();
    }
    phase_mark(PTG_PHASE_COUNT_NODES, s);
    Geom hg;
    PTG_HIP_CHECK(hipMemcpyAsync(&hc, w.ctr, sizeof(Counters), hipMemcpyDeviceToHost, s));
    PTG_HIP_CHECK(hipMemcpyAsync(&hg, w.geoms, sizeof(Geom), hipMemcpyDeviceToHost, s));
    PTG_HIP_CHECK(hipStreamSynchronize(s));
    if (hc.nonfinite) return nonfinite();
    info->n = n;
    info->nnodes = n >= 2 ? hc.nnodes : 1;
    info->has_collisions = hc.ncollide > 0;
    info->has_coincident = hc.coincident;
    info->max_level = n >= 2 ? hc.max_level : 0;
    info->nonfinite = 0;
    for (int k = 0; k < 3; k++) info->center[k] = hg.center[k];
    info->size = hg.size;
    if (n + info->nnodes > 0x7fffffffLL) {
        ptg_set_error("ptg_build_sort: %lld elements exceed the int32 link range", (long long)(n + info->nnodes));
        return PTG_ERR_ARG;
    }
    return PTG_OK;
}

// ============================================================================================
// node emission: node (i, L) for every level L in (lev[i-1], lev[i]] started by pair i.
// Node index = n + off[i] + (L - lev[i-1] - 1), i.e. DFS preorder of the octree.
// ============================================================================================
struct KeyView {
    const uint64_t *keys, *ext;  // ext == nullptr when no collisions
    const uint32_t *sys;         // system of each slot (forest), nullptr for one tree
    int64_t n;
    __device__ int digit(int64_t i, int L) const  // digit selecting the level-L cell (L >= 1)
    {
        if (L <= 21) return (int)((keys[i] >> (63 - 3 * L)) & 7);
        if (L <= 84) {
            int r = (L - 1) / 21 - 1;  // 0..2 -> K1..K3
            int l = L - 21 * (r + 1);
            return (int)((ext[(int64_t)r * n + i] >> (63 - 3 * l)) & 7);
        }
        return 7;  // bucket chains continue in octant 7 (octree.py:1146)
    }
    // do Morton slots a and b lie in the same level-L cell?
    __device__ bool same_cell(int64_t a, int64_t b, int L) const
    {
        if (sys && sys[a] != sys[b]) return false;
        if (L == 0) return true;
        if (L < 21) return (keys[a] >> (63 - 3 * L)) == (keys[b] >> (63 - 3 * L));
        if (keys[a] != keys[b]) return false;
        if (L == 21) return true;
        for (int r = 0; r < 3; r++) {
            int l = L - 21 * (r + 1);
            uint64_t x = ext[(int64_t)r * n + a], y = ext[(int64_t)r * n + b];
            if (l < 21) return (x >> (63 - 3 * l)) == (y >> (63 - 3 * l));
            if (x != y) return false;
            if (l == 21) return true;
        }
        return true;  // bucket levels: identical to double precision
    }
};

// element that starts at Morton slot q (q >= 1): the outermost node opened by pair q, else the
// particle; -1 past the end of the tree (the end of the particles, or of the system in a forest)
__device__ __forceinline__ int32_t head_of(int64_t q, int64_t n, const int32_t *lev, const int32_t *off,
                                           const uint32_t *sys)
{
    if (q >= n) return -1;
    if (sys && sys[q] != sys[q - 1]) return -1;
    if (q <= n - 2 && lev[q] > lev[q - 1]) return (int32_t)(n + off[q]);
    return (int32_t)q;
}

// Nodes are numbered breadth-first: level by level (system by system in a forest), each level in
// depth-first preorder, so that the nodes of one level, which the moment rollup visits together,
// and their children are contiguous.  Pair i opens the preorder offsets off[i] .. off[i] + li - lp - 1
// (levels lp+1 .. li); pi maps a preorder offset to its node offset.

// pass 1, one thread per pair: level, opening pair and ordering key (system, level) of every node, by
// preorder offset
__global__ __launch_bounds__(256) void k_emit_levels(const int32_t *lev, const int32_t *off, int64_t n,
                                                     const uint32_t *sys, int lbits, uint64_t *nkey, int32_t *nlev,
                                                     int32_t *npair, int64_t nnodes, const int32_t *count)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int lp = i == 0 ? -1 : lev[i - 1];
    const int li = lev[i];
    if (li <= lp) return;
    const uint64_t hi = sys ? (uint64_t)sys[i] << lbits : 0;
    const int64_t nn = ptg_node_count(nnodes, count);
    for (int L = lp + 1; L <= li; L++) {
        const int64_t a = off[i] + (L - lp - 1);
        if (a >= nn) break;  // a deferred build over its node capacity writes nothing
        nkey[a] = hi | (uint64_t)L;
        nlev[a] = L;
        npair[a] = (int32_t)i;
    }
}

__global__ __launch_bounds__(256) void k_invert(const uint32_t *order, int64_t m, uint32_t *pi)
{
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < m) pi[order[q]] = (uint32_t)q;
}

// deferred build: the node sort keys past the node count sort after every real node
__global__ __launch_bounds__(256) void k_pad_keys(uint64_t *nkey, int64_t cap, const int32_t *count)
{
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < cap && q >= ptg_node_count(cap, count)) nkey[q] = ~0ull;
}

// an element as head_of returns it (particle, preorder node n + offset, or -1) in the final numbering
__device__ __forceinline__ int32_t bfs_elem(int32_t h, int64_t n, const uint32_t *pi)
{
    return h < n ? h : (int32_t)(n + pi[h - n]);
}

// pass 2, one thread per node in node order (coalesced writes): geometry by the reference's centre
// recurrence from the root down the opening pair's key digits, FirstSubnode, and NextBranch (the
// element after the last slot of the node's cell, found by galloping then bisecting on "same level-L
// cell as slot i"; the long searches of the large cells run side by side).  order[q] is the
// preorder offset of node offset q.
__global__ __launch_bounds__(256) void k_emit_nodes(KeyView kv, const int32_t *lev, const int32_t *off, int64_t n,
                                                    int64_t nnodes, const Geom *geoms, const uint64_t *skey,
                                                    const uint32_t *order,
                                                    const int32_t *nlev, const int32_t *npair, const uint32_t *pi,
                                                    double *coord, double *sizes, int32_t *first, int32_t *next,
                                                    const int32_t *count)
{
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ptg_node_count(nnodes, count)) return;
    const int64_t a = order[q], idx = n + q;
    const int64_t i = npair[a];
    const int L = nlev[a];
    if (i < 0 || i >= n - 1 || L < 0) {  // every node is opened by one pair: never taken
        first[idx] = -1;
        next[idx] = -1;
        return;
    }
    const Geom *g = geoms + (kv.sys ? kv.sys[i] : 0);
    double c[3] = {g->center[0], g->center[1], g->center[2]};
    double s = g->size;
    for (int l = 0; l < L; l++) {
        const int d = kv.digit(i, l + 1);
#pragma unroll
        for (int k = 0; k < 3; k++) c[k] = c[k] + s * ptg_octant_offset(d, k);
        s = s * 0.5;
    }
    coord[3 * idx] = c[0];
    coord[3 * idx + 1] = c[1];
    coord[3 * idx + 2] = c[2];
    sizes[idx] = s;
    first[idx] = (L + 1 <= lev[i]) ? (int32_t)(n + pi[a + 1]) : (int32_t)i;
    // the next node of the same level and system (sorted key equal) starts the next level-L cell in
    // Morton order, so this cell ends before its first slot; where the cells are dense it ends just
    // there, found with one probe instead of a gallop
    const int64_t nn = ptg_node_count(nnodes, count);
    int64_t bound = n;  // first slot known to lie outside the cell (or n)
    if (q + 1 < nn && skey[q + 1] == skey[q]) {
        const int64_t i2 = npair[order[q + 1]];
        if (i2 >= i + 2 && i2 < n) bound = i2;
    }
    int64_t lo = i + 1;  // slot i+1 shares the pair's level >= L with slot i
    int64_t step = 1;
    if (bound < n && kv.same_cell(i, bound - 1, L)) {
        lo = bound - 1;
    } else {
        while (lo + step < bound && kv.same_cell(i, lo + step, L)) {
            lo += step;
            step <<= 1;
        }
    }
    int64_t hi = std::min<int64_t>(lo + step, bound);  // first candidate known not to match (or n)
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        if (kv.same_cell(i, mid, L)) lo = mid;
        else hi = mid;
    }
    next[idx] = bfs_elem(head_of(lo + 1, n, lev, off, kv.sys), n, pi);
}

// NextBranch of every particle
__global__ __launch_bounds__(256) void k_particle_next(const int32_t *lev, const int32_t *off, const uint32_t *sys,
                                                       int64_t n, const uint32_t *pi, int32_t *next,
                                                       const int32_t *count)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (count && *count == 0) return;  // a deferred build over its capacity
    if (i < n) next[i] = n == 1 ? -1 : bfs_elem(head_of(i + 1, n, lev, off, sys), n, pi);
}

// levels of the nodes in their final (system, level) order, for the level-sorted list of a forest
__global__ __launch_bounds__(256) void k_key_levels(const uint64_t *skey, int64_t m, int lbits, uint64_t *lv,
                                                    uint32_t *ids)
{
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    lv[q] = skey[q] & ((1ull << lbits) - 1);
    ids[q] = (uint32_t)q;
}

// particle slots: gather the input records into Morton order (octree.py:577-603); their NextBranch
// follows once the nodes are numbered (k_particle_next)
__global__ __launch_bounds__(256) void k_particles(const PartRec *rec, const uint32_t *order, const uint64_t *keys,
                                                   int64_t n, const ptg_tree_arrays t)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t o = order[i];
    const double2 *r = (const double2 *)(rec + o);
    const double2 r0 = r[0], r1 = r[1], r2 = r[2];
    t.coordinates[3 * i] = r0.x;
    t.coordinates[3 * i + 1] = r0.y;
    t.coordinates[3 * i + 2] = r1.x;
    t.masses[i] = r1.y;
    t.softenings[i] = r2.x;
    t.treewalk_indices[i] = o;
    if (t.keys) t.keys[i] = keys[i];
    t.first_subnode[i] = -1;
    t.nchild[i] = 0;
    t.child_base[i] = -1;
    // a particle's size, delta and quadrupole rows are not written: all are zero by definition, the
    // pack writes zeros for particle slots and no other kernel reads them
}

__global__ void k_single_root(const Geom *g, const ptg_tree_arrays t)
{
    // N = 1: the root (element 1) holds the lone particle
    for (int k = 0; k < 3; k++) t.coordinates[3 + k] = g->center[k];
    t.sizes[1] = g->size;
    t.first_subnode[1] = 0;
    t.next_branch[1] = -1;
}

// ============================================================================================
// links and moments
// ============================================================================================
// a child's moments as its parent reads them: mass, sum of m*x (nodes), softening, centre of mass
// last index s in [0, cnt) with a[s] <= v (a ascending, a[0] <= v)
template <typename I> __device__ __forceinline__ int64_t upper_index(const I *a, int64_t cnt, int64_t v)
{
    int64_t lo = 0, hi = cnt;
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        if ((int64_t)a[mid] <= v) lo = mid;
        else hi = mid;
    }
    return lo;
}

struct NodeWS {
    double *acc;        // (nnodes, 3) sum of m*x of each node, by node offset (element - n)
    int32_t *err;       // [0] corrupt-chain flag
    int32_t *bsum;
    uint64_t *lvA, *lvB;   // node (system, level) keys, then levels (sort keys)
    uint32_t *idA, *idB;   // node offsets (sort values)
    RsWS rs;
    int32_t *lvstart;      // (max_level + 2) first position of each level in the sorted list
    int32_t *elem;         // (n + nnodes) element stored at each walk slot
    int32_t *npair;        // (nnodes) the Morton pair that opened each node, by preorder offset
    int32_t *nlev;         // (nnodes) its level, by preorder offset
    uint32_t *pi;          // (nnodes) node offset of each preorder offset
    int32_t *qscan;        // (nnodes) child-count scan in preorder (forest slot layout)
};

static bool carve_nodes(PtgArena &A, int64_t n, int64_t nnodes, int64_t max_level, NodeWS &w)
{
    w.acc = A.take<double>(3 * nnodes);
    w.err = A.take<int32_t>(64);
    w.bsum = A.take<int32_t>(scan_blocks(nnodes) + 64);
    w.lvA = A.take<uint64_t>(nnodes);
    w.lvB = A.take<uint64_t>(nnodes);
    w.idA = A.take<uint32_t>(nnodes);
    w.idB = A.take<uint32_t>(nnodes);
    carve_rs(A, nnodes, w.rs);
    w.lvstart = A.take<int32_t>(max_level + 2);
    w.elem = A.take<int32_t>(n + nnodes);
    w.npair = A.take<int32_t>(nnodes);
    w.nlev = A.take<int32_t>(nnodes);
    w.pi = A.take<uint32_t>(nnodes);
    w.qscan = A.take<int32_t>(nnodes);
    return A.ok();
}

extern "C" size_t ptg_node_workspace_bytes(int64_t n, int64_t nnodes, int max_level)
{
    PtgArena A(nullptr, 0);
    NodeWS w;
    carve_nodes(A, n, std::max<int64_t>(nnodes, 1), std::max(max_level, 0), w);
    return A.used;
}

// children of a node in link order; -1 if the chain is malformed (more than 8 links or an index
// outside [0, ne)), so a build bug cannot turn into an out-of-bounds access
__device__ __forceinline__ int node_children(const int32_t *first, const int32_t *next, int64_t idx, int64_t ne,
                                             int32_t *kids)
{
    int32_t c = first[idx], stop = next[idx];
    int nk = 0;
    while (c != stop && c >= 0) {
        if (nk == 8 || c >= ne) return -1;
        kids[nk++] = c;
        c = next[c];
    }
    return nk;
}

__global__ __launch_bounds__(256) void k_node_init(NodeWS w)
{
    if (threadIdx.x < 64) w.err[threadIdx.x] = 0;
}

// child counts (the moments pass, or the particle fold, writes every node moment afterwards)
__global__ __launch_bounds__(256) void k_link(int64_t n, int64_t nnodes, const ptg_tree_arrays t, NodeWS w)
{
    int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    nnodes = ptg_node_count(nnodes, t.node_count);
    if (a >= nnodes) return;
    int64_t idx = n + a;
    int32_t kids[8];
    int nk = node_children(t.first_subnode, t.next_branch, idx, n + nnodes, kids);
    if (nk < 0) {
        w.err[0] = 1;
        nk = 0;
    }
    t.nchild[idx] = nk;
}

// walk slots: the root of system s sits at slot s, child q of node idx at slot child_base[idx] + q;
// record the element -> slot map (for the pack) and its inverse (for the moments pass)
__global__ __launch_bounds__(256) void k_slots(int64_t n, int64_t nnodes, int64_t nsys, const int32_t *roots,
                                               const ptg_tree_arrays t, NodeWS w)
{
    int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    nnodes = ptg_node_count(nnodes, t.node_count);
    if (a >= nnodes) return;
    int64_t idx = n + a;
    if (a < nsys) {
        const int32_t r = roots ? roots[a] : (int32_t)n;
        t.slot[r] = (int32_t)a;
        w.elem[a] = r;
    }
    int32_t kids[8];
    int nk = node_children(t.first_subnode, t.next_branch, idx, n + nnodes, kids);
    if (nk < 0) {
        w.err[0] = 1;
        return;
    }
    int64_t base = t.child_base[idx];
    for (int q = 0; q < nk; q++) {
        int64_t sl = base + q;
        if (sl < nsys || sl >= n + nnodes) {
            w.err[0] = 1;
            continue;
        }
        t.slot[kids[q]] = (int32_t)sl;
        w.elem[sl] = kids[q];
    }
}

// Bottom-up moments, level by level from the deepest up: every node of a level is computed after
// all of its children (a kernel boundary, or a workgroup barrier for runs of small levels, orders
// the levels), sums run in a fixed order, so the result is deterministic and needs no device-scope
// fences or FP64 atomics.
__global__ __launch_bounds__(256) void k_level_starts(const uint64_t *lv, int64_t nnodes, int max_level,
                                                      int32_t *lvstart, const int32_t *count)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    nnodes = ptg_node_count(nnodes, count);
    if (i > nnodes) return;
    // lvstart[L] = first sorted position with level >= L
    int64_t a = i == 0 ? -1 : (int64_t)lv[i - 1];
    int64_t b = i == nnodes ? (int64_t)max_level + 1 : (int64_t)lv[i];
    for (int64_t L = a + 1; L <= b && L <= max_level + 1; L++) lvstart[L] = (int32_t)i;
}


// One node's moments, one node per lane, in the reference's summation order: particle children in
// digit order (folded during the split, octree.py:956-966), then node children in descending index
// order, i.e. descending digit order (RollupMoments' reverse scan, octree.py:212-231).  Zero-mass
// node children contribute nothing (:215-216).  Quadrupoles follow RollupQuadrupoles (:252-292).
// Every child record is loaded unconditionally (absent children read a valid stand-in element) so
// that all of a lane's gathers are in flight together; absent children are then masked out of the
// sums, which therefore see exactly the reference's sequence of operations.
template <bool QUAD>
__device__ __forceinline__ void lane_moments(int64_t idx, int64_t n, int64_t ne, const ptg_tree_arrays &t, NodeWS &w)
{
    int nk = t.nchild[idx];
    const int64_t base = t.child_base[idx];
    if (nk < 0 || nk > 8 || base < 1 || base + nk > ne) {
        w.err[0] = 1;
        nk = 0;
    }
    int32_t c[8];
    bool part[8], node[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        int64_t e = k < nk ? (int64_t)w.elem[base + k] : 0;
        if (e < 0 || e >= ne) {  // a malformed slot map is reported, never followed
            w.err[0] = 1;
            e = 0;
        }
        c[k] = (int32_t)e;
        part[k] = k < nk && e < n;
        node[k] = k < nk && e >= n;
    }
    // per child: mass, softening and one 3-vector (a particle's position, a node's sum of m*x)
    double km[8], kh[8], kv0[8], kv1[8], kv2[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int64_t e = c[k];
        const double *v = node[k] ? w.acc + 3 * (e - n) : t.coordinates + 3 * e;
        km[k] = t.masses[e];
        kh[k] = t.softenings[e];
        kv0[k] = v[0];
        kv1[k] = v[1];
        kv2[k] = v[2];
    }
    double com0 = t.coordinates[3 * idx], com1 = t.coordinates[3 * idx + 1], com2 = t.coordinates[3 * idx + 2];
    double m = 0.0, a0 = 0.0, a1 = 0.0, a2 = 0.0, hs = 0.0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        if (!part[k]) continue;
        const double mi = km[k];
        m += mi;
        a0 += mi * kv0[k];
        a1 += mi * kv1[k];
        a2 += mi * kv2[k];
        if (!(hs >= kh[k])) hs = kh[k];
    }
#pragma unroll
    for (int k = 7; k >= 0; k--) {
        if (!node[k] || km[k] == 0.0) continue;
        a0 += kv0[k];
        a1 += kv1[k];
        a2 += kv2[k];
        m += km[k];
        if (!(hs >= kh[k])) hs = kh[k];
    }
    t.masses[idx] = m;
    t.softenings[idx] = hs;
    if (m != 0.0) {
        double inv = 1.0 / m;
        double c0 = a0 * inv, c1 = a1 * inv, c2 = a2 * inv;
        double dx0 = c0 - com0, dx1 = c1 - com1, dx2 = c2 - com2;
        double d2 = 0.0;
        d2 += dx0 * dx0;
        d2 += dx1 * dx1;
        d2 += dx2 * dx2;
        t.deltas[idx] = sqrt(d2);
        com0 = c0;
        com1 = c1;
        com2 = c2;
        t.coordinates[3 * idx] = c0;
        t.coordinates[3 * idx + 1] = c1;
        t.coordinates[3 * idx + 2] = c2;
    } else {
        t.deltas[idx] = 0.0;  // a massless node keeps its geometric centre
    }
    const int64_t ia = idx - n;
    w.acc[3 * ia] = a0;
    w.acc[3 * ia + 1] = a1;
    w.acc[3 * ia + 2] = a2;
    if constexpr (!QUAD) return;
    double Q[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    auto shift = [&](double mm, double r0, double r1, double r2) {
        double rsq = r0 * r0 + r1 * r1 + r2 * r2;
        Q[0] += mm * (3.0 * r0 * r0 - rsq);
        Q[1] += mm * 3.0 * r0 * r1;
        Q[2] += mm * 3.0 * r0 * r2;
        Q[3] += mm * 3.0 * r1 * r0;
        Q[4] += mm * (3.0 * r1 * r1 - rsq);
        Q[5] += mm * 3.0 * r1 * r2;
        Q[6] += mm * 3.0 * r2 * r0;
        Q[7] += mm * 3.0 * r2 * r1;
        Q[8] += mm * (3.0 * r2 * r2 - rsq);
    };
    // child statistics of the learned criterion's features, gathered on the way: largest child mass,
    // sums of m d^2 and m d^3, largest d (d = child centre to the centre of mass)
    const bool stats = t.node_stats != nullptr;
    double mmax = 0.0, m2 = 0.0, m3 = 0.0, dmax = 0.0;
    auto stat = [&](double mc, double r0, double r1, double r2) {
        const double d = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
        mmax = fmax(mmax, mc);
        m2 += mc * d * d;
        m3 += mc * d * d * d;
        dmax = fmax(dmax, d);
    };
#pragma unroll
    for (int k = 0; k < 8; k++) {
        if (!part[k]) continue;
        const double r0 = kv0[k] - com0, r1 = kv1[k] - com1, r2 = kv2[k] - com2;
        shift(km[k], r0, r1, r2);
        if (stats) stat(km[k], r0, r1, r2);
    }
#pragma unroll
    for (int k = 7; k >= 0; k--) {
        if (!node[k]) continue;
        const int64_t e = c[k];
        const double r0 = t.coordinates[3 * e] - com0, r1 = t.coordinates[3 * e + 1] - com1,
                     r2 = t.coordinates[3 * e + 2] - com2;
        if (km[k] == 0.0) {  // a massless child adds nothing but its distance
            if (stats) stat(0.0, r0, r1, r2);
            continue;
        }
        const double *KQ = t.quadrupoles + 9 * e;
#pragma unroll
        for (int j = 0; j < 9; j++) Q[j] += KQ[j];
        shift(km[k], r0, r1, r2);
        if (stats) stat(km[k], r0, r1, r2);
    }
#pragma unroll
    for (int j = 0; j < 9; j++) t.quadrupoles[9 * idx + j] = Q[j];
    if (stats) {
        double q2 = 0.0;
        for (int j = 0; j < 9; j++) q2 += Q[j] * Q[j];
        double *st = t.node_stats + 5 * (idx - n);
        st[0] = sqrt(q2);
        st[1] = mmax;
        st[2] = m2;
        st[3] = m3;
        st[4] = dmax;
    }
}

#define MOM_THREADS 256
#define MOM_DEFERRED_BLOCKS 2048  // grid of each level's launch in a deferred build (grid-stride over the level)

// one level, one node per lane
template <bool QUAD>
__global__ __launch_bounds__(MOM_THREADS) void k_moments_level(int64_t n, int64_t nnodes, int L,
                                                               const int32_t *lvstart, const uint32_t *byid,
                                                               const ptg_tree_arrays t, NodeWS w)
{
    const int64_t lo = lvstart[L], hi = lvstart[L + 1];
    const int64_t stride = (int64_t)gridDim.x * MOM_THREADS;
    for (int64_t qi = lo + (int64_t)blockIdx.x * MOM_THREADS + threadIdx.x; qi < hi; qi += stride)
        lane_moments<QUAD>(n + (byid ? (int64_t)byid[qi] : qi), n, n + nnodes, t, w);
}

// compute_moments=False: the split has still folded particle children into their parent's mass
// and softening (octree.py:956-966); nothing else is rolled up.
__global__ __launch_bounds__(256) void k_fold_particles(int64_t n, int64_t nnodes, const ptg_tree_arrays t)
{
    int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= ptg_node_count(nnodes, t.node_count)) return;
    int64_t idx = n + a;
    int32_t kids[8];
    int nk = node_children(t.first_subnode, t.next_branch, idx, n + nnodes, kids);
    double m = 0.0, hs = 0.0;
    for (int q = 0; q < nk; q++) {
        int32_t c = kids[q];
        if (c >= n) continue;
        m += t.masses[c];
        if (!(hs >= t.softenings[c])) hs = t.softenings[c];
    }
    t.masses[idx] = m;
    t.softenings[idx] = hs;
    t.deltas[idx] = 0.0;
    if (t.quadrupoles)
        for (int k = 0; k < 9; k++) t.quadrupoles[9 * idx + k] = 0.0;
}

__global__ __launch_bounds__(256) void k_child_base_fix(int64_t n, int64_t nnodes, int64_t nsys, int32_t *child_base,
                                                        const int32_t *count)
{
    int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a < ptg_node_count(nnodes, count)) child_base[n + a] += (int32_t)nsys;  // slots [0, nsys) hold the roots
}

// child counts in depth-first preorder (pi: preorder offset -> node offset)
__global__ __launch_bounds__(256) void k_nchild_preorder(int64_t n, int64_t nnodes, const uint32_t *pi,
                                                         const int32_t *nchild, int32_t *out, const int32_t *count)
{
    int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a < ptg_node_count(nnodes, count)) out[a] = nchild[n + pi[a]];
}

__global__ __launch_bounds__(256) void k_child_base_preorder(int64_t n, int64_t nnodes, int64_t nsys,
                                                             const uint32_t *pi, const int32_t *scan,
                                                             int32_t *child_base, const int32_t *count)
{
    int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a < ptg_node_count(nnodes, count)) child_base[n + pi[a]] = scan[a] + (int32_t)nsys;  // slots [0, nsys) hold the roots
}

// forest roots: pair offsets[s] (the first pair of system s) opens its root
__global__ __launch_bounds__(256) void k_forest_roots(const int64_t *soff, int64_t nsys, const int32_t *off, int64_t n,
                                                      const uint32_t *pi, int32_t *roots, const int32_t *count)
{
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (count && *count == 0) return;  // a deferred build over its capacity
    if (s < nsys) roots[s] = (int32_t)(n + pi[off[soff[s]]]);
}

static ptg_status pack_in_slot_order(const ptg_tree_arrays &t, int64_t n, int64_t nnodes, int64_t nsys,
                                     const ptg_forest *forest, const ptg_walk_tree *pack, const int32_t *elem,
                                     hipStream_t s);

// a deferred tree over its capacity has no nodes: its walk copy's root slots become massless leaves,
// so a walk of it (before the caller sees the overflow) finds nothing to follow
__global__ void k_empty_roots(const ptg_walk_tree wt, int64_t nsys, const int32_t *count)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (*count != 0 || r >= nsys) return;
    const int q = wt.prec == PTG_F64 ? 8 : 4;  // bytes per real
    char *x = (char *)wt.xyzm + r * 4 * q, *a = (char *)wt.aux + r * 4 * q;
    for (int k = 0; k < 4 * q; k++) x[k] = a[k] = 0;
    wt.kids[2 * r] = 0;
    wt.kids[2 * r + 1] = 0;
    if (wt.quad) {
        char *qq = (char *)wt.quad + r * 8 * q;
        for (int k = 0; k < 8 * q; k++) qq[k] = 0;
    }
}

// node emission, links, slots, moments and the walk copy: the second half of the build, shared by
// ptg_build_tree (nnodes and max_level exact) and ptg_build_deferred (capacities; t.node_count set)
static ptg_status tree_phase(const double *pos, const double *m, const double *h, int64_t n, int64_t nnodes,
                             int max_level, bool has_collisions, const ptg_forest *forest, int64_t nsys,
                             int quadrupole, int compute_moments, BuildWS &w, NodeWS &nw, const ptg_tree_arrays &t,
                             const ptg_walk_tree *pack, void *particles_ready, hipStream_t s)
{
    (void)pos;
    (void)m;
    (void)h;
    const bool F = nsys > 1;
    if (F) w.soff = forest->offsets;
    // a deferred build does not know whether keys collided: the extended keys are read (only ever
    // for colliding slots, which the collision pass wrote)
    KeyView kv{w.kA, has_collisions ? w.ext : nullptr, w.sys, n};
    const int32_t *cnt = t.node_count;  // device node count of a deferred build, else NULL
    ptg_status st;
    phase_mark(PTG_PHASE_TREE_START, s);
    k_particles<<<grid1(n), 256, 0, s>>>(w.rec, w.vA, w.kA, n, t);
    if (particles_ready) PTG_HIP_CHECK(hipEventRecord((hipEvent_t)particles_ready, s));
    phase_mark(PTG_PHASE_PERMUTE, s);
    // node order: (system, level), depth-first preorder within a level (stable LSD radix sort of the
    // preorder offsets)
    int lbits = 1;
    while ((1ll << lbits) <= (int64_t)max_level) lbits++;
    lbits = ((lbits + RS_BITS - 1) / RS_BITS) * RS_BITS;
    const int kbits = F ? ((lbits + sys_bits(nsys) + RS_BITS - 1) / RS_BITS) * RS_BITS : lbits;
    if (n == 1) {
        PTG_HIP_CHECK(hipMemsetAsync(nw.lvA, 0, sizeof(uint64_t), s));
    } else {
        k_emit_levels<<<grid1(n - 1), 256, 0, s>>>(w.lev, w.off, n, F ? w.sys : nullptr, lbits, nw.lvA, nw.nlev,
                                                   nw.npair, nnodes, cnt);
        if (cnt) k_pad_keys<<<grid1(nnodes), 256, 0, s>>>(nw.lvA, nnodes, cnt);
    }
    k_iota_u32<<<grid1(nnodes), 256, 0, s>>>(nw.idA, nnodes);
    bool in_first;
    st = radix_sort(nw.lvA, nw.idA, nw.lvB, nw.idB, nnodes, kbits, nw.rs, &in_first, s);
    if (st) return st;
    const uint64_t *skey = in_first ? nw.lvA : nw.lvB;
    k_invert<<<grid1(nnodes), 256, 0, s>>>(in_first ? nw.idA : nw.idB, nnodes, nw.pi);
    if (n == 1) {
        k_single_root<<<1, 1, 0, s>>>(w.geoms, t);
    } else {
        k_emit_nodes<<<grid1(nnodes), 256, 0, s>>>(kv, w.lev, w.off, n, nnodes, w.geoms, skey, in_first ? nw.idA : nw.idB,
                                                   nw.nlev, nw.npair, nw.pi, t.coordinates, t.sizes, t.first_subnode,
                                                   t.next_branch, cnt);
    }
    k_particle_next<<<grid1(n), 256, 0, s>>>(w.lev, w.off, w.sys, n, nw.pi, t.next_branch, cnt);
    if (F) k_forest_roots<<<grid1(nsys), 256, 0, s>>>(w.soff, nsys, w.off, n, nw.pi, forest->roots, cnt);
    else if (forest) k_fill_i32<<<1, 64, 0, s>>>(forest->roots, 1, (int32_t)n);
    PTG_LAUNCH_CHECK();
    // the level-sorted node list of the rollup: the node order itself for one tree; a forest sorts
    // its (system, level) order by level
    const uint32_t *byid = nullptr;
    const uint64_t *lv = skey;
    if (F) {
        uint64_t *lk = in_first ? nw.lvB : nw.lvA, *lk2 = in_first ? nw.lvA : nw.lvB;
        uint32_t *id = in_first ? nw.idB : nw.idA, *id2 = in_first ? nw.idA : nw.idB;
        k_key_levels<<<grid1(nnodes), 256, 0, s>>>(skey, nnodes, lbits, lk, id);
        bool first2;
        st = radix_sort(lk, id, lk2, id2, nnodes, lbits, nw.rs, &first2, s);
        if (st) return st;
        lv = first2 ? lk : lk2;
        byid = first2 ? id : id2;
    }
    phase_mark(PTG_PHASE_NODES, s);
    k_level_starts<<<grid1(nnodes + 1), 256, 0, s>>>(lv, nnodes, max_level, nw.lvstart, cnt);
    k_node_init<<<1, 64, 0, s>>>(nw);
    k_link<<<grid1(nnodes), 256, 0, s>>>(n, nnodes, t, nw);
    PTG_LAUNCH_CHECK();
    // sibling-contiguous walk slots: an exclusive scan of child counts in node order puts the
    // children of each level's nodes in one slot range; a forest scans in depth-first preorder
    // instead, which keeps each system's subtrees together (a faster batched walk)
    if (F) {
        // (a deferred tree's entries past its node count come after every real one: the exclusive
        // scan of the real entries does not see them)
        k_nchild_preorder<<<grid1(nnodes), 256, 0, s>>>(n, nnodes, nw.pi, t.nchild, nw.qscan, cnt);
        st = excl_scan(LoadArr{nw.qscan}, nnodes, nw.qscan, nw.bsum, nullptr, s);  // in place
        if (st) return st;
        k_child_base_preorder<<<grid1(nnodes), 256, 0, s>>>(n, nnodes, nsys, nw.pi, nw.qscan, t.child_base, cnt);
    } else {
        st = excl_scan(LoadArr{t.nchild + n}, nnodes, t.child_base + n, nw.bsum, nullptr, s);
        if (st) return st;
        k_child_base_fix<<<grid1(nnodes), 256, 0, s>>>(n, nnodes, nsys, t.child_base, cnt);
    }
    k_slots<<<grid1(nnodes), 256, 0, s>>>(n, nnodes, nsys, F ? forest->roots : nullptr, t, nw);
    PTG_LAUNCH_CHECK();
    phase_mark(PTG_PHASE_LINKS, s);
    if (compute_moments) {
        // one launch per level, deepest first, each reading its node range from the device (no host
        // round trip for the level sizes; small levels cost a few microseconds of dispatch and load
        // latency each: running the levels nearest the root in one workgroup with barriers measured
        // no faster, as each level's chain of dependent loads remains)
        for (int L = max_level; L >= 0; L--) {
            if (quadrupole)
                k_moments_level<true><<<MOM_DEFERRED_BLOCKS, MOM_THREADS, 0, s>>>(n, nnodes, L, nw.lvstart, byid, t, nw);
            else
                k_moments_level<false><<<MOM_DEFERRED_BLOCKS, MOM_THREADS, 0, s>>>(n, nnodes, L, nw.lvstart, byid, t, nw);
        }
    } else {
        k_fold_particles<<<grid1(nnodes), 256, 0, s>>>(n, nnodes, t);
    }
    PTG_LAUNCH_CHECK();
    phase_mark(PTG_PHASE_MOMENTS, s);
    if (pack) {  // the walk copy, in slot order through the build's slot map
        st = pack_in_slot_order(t, n, nnodes, nsys, forest, pack, nw.elem, s);
        if (st) return st;
        if (cnt) {
            k_empty_roots<<<grid1(nsys), 256, 0, s>>>(*pack, nsys, cnt);
            PTG_LAUNCH_CHECK();
        }
    }
    phase_mark(PTG_PHASE_PACK, s);
    return PTG_OK;
}

extern "C" ptg_status ptg_build_tree(const double *pos, const double *m, const double *h, const ptg_build_info *info,
                                     const ptg_forest *forest, int quadrupole, int compute_moments, void *ws,
                                     size_t ws_bytes, void *node_ws, size_t node_ws_bytes,
                                     const ptg_tree_arrays *tree, const ptg_walk_tree *pack, void *particles_ready,
                                     void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (!pos || !m || !h || !info || !tree || info->n < 1) {
        ptg_set_error("ptg_build_tree: bad arguments");
        return PTG_ERR_ARG;
    }
    int64_t nsys;
    ptg_status st = forest_check(forest, info->n, "ptg_build_tree", &nsys);
    if (st) return st;
    const ptg_tree_arrays t = *tree;
    if (!t.coordinates || !t.masses || !t.softenings || !t.sizes || !t.deltas || !t.first_subnode ||
        !t.next_branch || !t.treewalk_indices || !t.nchild || !t.child_base || !t.slot ||
        (quadrupole && !t.quadrupoles)) {
        ptg_set_error("ptg_build_tree: a required tree array is NULL");
        return PTG_ERR_ARG;
    }
    if (t.node_stats && !(quadrupole && compute_moments)) {
        ptg_set_error("ptg_build_tree: node_stats needs a build with quadrupoles and moments");
        return PTG_ERR_ARG;
    }
    if (pack && (!pack->xyzm || !pack->aux || !pack->kids || pack->nslots != info->n + info->nnodes ||
                 (pack->prec != PTG_F32 && pack->prec != PTG_F64) || (pack->quad && !quadrupole))) {
        ptg_set_error("ptg_build_tree: pack must describe a walk copy of this tree (see ptg_pack_walk)");
        return PTG_ERR_ARG;
    }
    int64_t n = info->n, nnodes = info->nnodes;
    PtgArena A(ws, ws_bytes);
    BuildWS w;
    if (!ws || !carve_build(A, n, w, nsys)) {
        ptg_set_error("ptg_build_tree: build workspace too small");
        return PTG_ERR_WORKSPACE;
    }
    PtgArena B(node_ws, node_ws_bytes);
    NodeWS nw;
    const int max_level = std::max(info->max_level, 0);
    if (!node_ws || !carve_nodes(B, n, nnodes, max_level, nw)) {
        ptg_set_error("ptg_build_tree: node workspace too small (%zu < %zu)", node_ws_bytes,
                      ptg_node_workspace_bytes(n, nnodes, max_level));
        return PTG_ERR_WORKSPACE;
    }
    if (t.node_count) {  // the count a later call reads back equals the exact node count
        k_fill_i32<<<1, 64, 0, s>>>(t.node_count, 1, (int32_t)nnodes);
        PTG_LAUNCH_CHECK();
    }
    return tree_phase(pos, m, h, n, nnodes, max_level, info->has_collisions != 0, forest, nsys, quadrupole,
                      compute_moments, w, nw, t, pack, particles_ready, s);
}

// ============================================================================================
// deferred build: both halves of the build without a host synchronisation
// ============================================================================================
// collision runs within the capacity of the rank sort; otherwise nothing is re-sorted (and the
// tree is reported incomplete)
__global__ void k_budget_sub(Counters *c, int32_t cap)
{
    if (threadIdx.x != 0) return;
    if (c->nsub > cap) {
        c->eff_sub = 0;
        c->overflow |= PTG_OVERFLOW_COLLISIONS;
    } else {
        c->eff_sub = c->nsub;
    }
}

// the node count the tree kernels use: the count, or 0 (they then do nothing) when the tree does not
// fit its capacities
__global__ void k_budget_nodes(Counters *c, int64_t nodes_cap, int level_cap, int64_t n, int32_t *node_count)
{
    if (threadIdx.x != 0) return;
    const int32_t nnodes = n >= 2 ? c->nnodes : 1;
    int32_t of = c->overflow;
    if (nnodes > nodes_cap) of |= PTG_OVERFLOW_NODES;
    if (c->max_level > level_cap) of |= PTG_OVERFLOW_LEVELS;
    c->overflow = of;
    *node_count = of ? 0 : nnodes;
}

__global__ void k_counts_out(const Counters *c, const Geom *g, int64_t n, ptg_build_counts *out)
{
    if (threadIdx.x != 0) return;
    out->nnodes = n >= 2 ? c->nnodes : 1;
    out->max_level = n >= 2 ? c->max_level : 0;
    out->has_collisions = c->ncollide > 0;
    out->has_coincident = c->coincident;
    out->nonfinite = c->nonfinite;
    out->overflow = c->overflow;
    out->pad[0] = out->pad[1] = 0;
    for (int k = 0; k < 3; k++) out->center[k] = g->center[k];
    out->size = g->size;
}

extern "C" ptg_status ptg_build_deferred(const double *pos, const double *m, const double *h, int64_t n,
                                         const ptg_forest *forest, int quadrupole, int compute_moments,
                                         int64_t nnodes_cap, int level_cap, void *ws, size_t ws_bytes, void *node_ws,
                                         size_t node_ws_bytes, const ptg_tree_arrays *tree, const ptg_walk_tree *pack,
                                         ptg_build_counts *counts, void *particles_ready, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (n < 2 || !pos || !m || !h || !tree || !counts || !tree->node_count) {
        ptg_set_error("ptg_build_deferred: need n >= 2, pos, m, h, counts and a tree with node_count");
        return PTG_ERR_ARG;
    }
    if (n > 0x3fffffffLL || nnodes_cap < 1 || nnodes_cap > 0x3fffffffLL - n || level_cap < 0 || level_cap > 255) {
        ptg_set_error("ptg_build_deferred: n=%lld, nnodes_cap=%lld, level_cap=%d out of range", (long long)n,
                      (long long)nnodes_cap, level_cap);
        return PTG_ERR_ARG;
    }
    int64_t nsys = 1;
    if (forest) {
        if (forest->nsys < 1 || !forest->offsets || !forest->roots || !forest->cubes) {
            ptg_set_error("ptg_build_deferred: a forest needs nsys >= 1, offsets, roots and cubes");
            return PTG_ERR_ARG;
        }
        nsys = forest->nsys;
    }
    const ptg_tree_arrays t = *tree;
    if (!t.coordinates || !t.masses || !t.softenings || !t.sizes || !t.deltas || !t.first_subnode ||
        !t.next_branch || !t.treewalk_indices || !t.nchild || !t.child_base || !t.slot ||
        (quadrupole && !t.quadrupoles) || (t.node_stats && !(quadrupole && compute_moments))) {
        ptg_set_error("ptg_build_deferred: a required tree array is NULL, or node_stats without quadrupole moments");
        return PTG_ERR_ARG;
    }
    if (pack && (!pack->xyzm || !pack->aux || !pack->kids || pack->nslots != n + nnodes_cap ||
                 (pack->prec != PTG_F32 && pack->prec != PTG_F64) || (pack->quad && !quadrupole))) {
        ptg_set_error("ptg_build_deferred: pack must describe a walk copy of n + nnodes_cap slots");
        return PTG_ERR_ARG;
    }
    PtgArena A(ws, ws_bytes);
    BuildWS w;
    if (!ws || !carve_build(A, n, w, nsys)) {
        ptg_set_error("ptg_build_deferred: workspace too small (%zu < %zu)", ws_bytes,
                      ptg_build_workspace_bytes(n, nsys));
        return PTG_ERR_WORKSPACE;
    }
    PtgArena B(node_ws, node_ws_bytes);
    NodeWS nw;
    if (!node_ws || !carve_nodes(B, n, nnodes_cap, level_cap, nw)) {
        ptg_set_error("ptg_build_deferred: node workspace too small (%zu < %zu)", node_ws_bytes,
                      ptg_node_workspace_bytes(n, nnodes_cap, level_cap));
        return PTG_ERR_WORKSPACE;
    }
    const bool F = nsys > 1;
    if (F) w.soff = forest->offsets;
    // ---- the first half of ptg_build_sort, every data-dependent step on the device ----
    k_zero_counters<<<1, 64, 0, s>>>(w.ctr);
    phase_mark(PTG_PHASE_START, s);
    ptg_status st = F ? sort_by_keys_forest(pos, m, h, n, w, s) : sort_by_keys(pos, m, h, n, w, s);
    if (st) return st;
    phase_mark(PTG_PHASE_SORT, s);
    if (!F) PTG_HIP_CHECK(hipMemcpyAsync(w.geoms, w.geom, sizeof(Geom), hipMemcpyDeviceToDevice, s));
    if (forest) {
        if (F) k_check_finite<<<(unsigned)std::min<int64_t>(1024, ptg_cdiv(n, 256)), 256, 0, s>>>(pos, m, h, n, w.ctr);
        k_forest_cubes<<<grid1(nsys), 256, 0, s>>>(w.geoms, nsys, forest->cubes);
    }
    k_levels<<<lev_grid(n), 256, 0, s>>>(w.kA, w.sys, n, w.lev, w.ctr);
    // collision runs: re-keyed here; the stage's kernels read k_levels' count of colliding pairs and
    // return at once when it is 0 (nsub then stays 0 from the zeroed counters)
    const int32_t *gate = &w.ctr->ncollide;
    k_collision_flags<<<lev_grid(n + 1), 256, 0, s>>>(w.kA, w.sys, n, w.flag, gate);
    st = excl_scan(LoadArr{w.flag}, n, w.subpos, w.bsum, &w.ctr->nsub, s, gate);
    if (st) return st;
    k_budget_sub<<<1, 64, 0, s>>>(w.ctr, PTG_DEFERRED_COLLISION_CAP);
    k_collision_gather<<<lev_grid(n + 1), 256, 0, s>>>(pos, w.kA, w.vA, w.flag, w.subpos, n, w.geoms, w.sys, w.sk0, w.sk1,
                                                w.sk2, w.sk3, w.sslot, w.sval, w.perm, gate);
    const int64_t mcap = std::min<int64_t>(n, PTG_DEFERRED_COLLISION_CAP);
    k_rank_sort4<<<grid1(mcap), 256, 0, s>>>(w.sk0, w.sk1, w.sk2, w.sk3, w.sslot, w.sys, mcap, w.perm,
                                             &w.ctr->eff_sub);
    // (the extended keys are read only for colliding slots, all of which this writes)
    k_collision_apply<<<grid1(mcap), 256, 0, s>>>(w.perm, mcap, w.sslot, w.sval, w.sk1, w.sk2, w.sk3, w.vA, w.ext, n,
                                                  &w.ctr->eff_sub);
    k_levels_ext<<<lev_grid(n), 256, 0, s>>>(w.kA, w.ext, w.sys, w.soff, n, w.lev, w.ctr);
    st = excl_scan(LoadNodeCount{w.lev}, n - 1, w.off, w.bsum, &w.ctr->nnodes, s);
    if (st) return st;
    k_budget_nodes<<<1, 64, 0, s>>>(w.ctr, nnodes_cap, level_cap, n, t.node_count);
    PTG_LAUNCH_CHECK();
    phase_mark(PTG_PHASE_COUNT_NODES, s);
    // ---- the second half, over the capacities ----
    st = tree_phase(pos, m, h, n, nnodes_cap, level_cap, true, forest, nsys, quadrupole, compute_moments, w, nw, t,
                    pack, particles_ready, s);
    if (st) return st;
    k_counts_out<<<1, 64, 0, s>>>(w.ctr, w.geoms, n, counts);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

// ============================================================================================
// walk pack
// ============================================================================================
template <typename T> struct alignas(4 * sizeof(T)) PackV4 {
    T x, y, z, w;
};

// one slot of the walk pack; xyzm, aux and kids go out as one vector store each (the slots of a
// thread block are scattered, so partial-line scalar stores would multiply the write traffic)
template <typename T>
__device__ void pack_slot(int64_t e, int64_t slot, int64_t n, const ptg_tree_arrays &t, const ptg_walk_tree &wt,
                          const double *origin)
{
    PackV4<T> X, X2;
    X.x = (T)(t.coordinates[3 * e] - origin[0]);
    X.y = (T)(t.coordinates[3 * e + 1] - origin[1]);
    X.z = (T)(t.coordinates[3 * e + 2] - origin[2]);
    X.w = (T)t.masses[e];
    X2.x = (T)t.softenings[e];
    X2.y = e >= n ? (T)t.sizes[e] : (T)0;  // particle rows of sizes and deltas are not written
    X2.z = e >= n ? (T)t.deltas[e] : (T)0;
    // Frobenius norm of the node's traceless quadrupole (the monopole truncation error scale used by
    // the error-bounded criterion); -1 when the tree carries no quadrupoles
    double qn = -1.0;
    if (e >= n && t.quadrupoles) {
        if (t.node_stats) {
            qn = t.node_stats[5 * (e - n)];  // the same sum, stored by the moment pass
        } else {
            const double *Q = t.quadrupoles + 9 * e;
            double s2 = 0.0;
            for (int k = 0; k < 9; k++) s2 += Q[k] * Q[k];
            qn = sqrt(s2);
        }
    } else if (e < n) {
        qn = 0.0;
    }
    X2.w = (T)qn;
    ((PackV4<T> *)wt.xyzm)[slot] = X;
    ((PackV4<T> *)wt.aux)[slot] = X2;
    bool node = e >= n;
    ((int2 *)wt.kids)[slot] = make_int2(node ? t.child_base[e] : (int32_t)e, node ? t.nchild[e] : 0);
    if (wt.quad) {
        T *q = (T *)wt.quad + 8 * slot;
        if (node && t.quadrupoles) {
            const double *Q = t.quadrupoles + 9 * e;
            q[0] = (T)Q[0]; q[1] = (T)Q[1]; q[2] = (T)Q[2];
            q[3] = (T)Q[4]; q[4] = (T)Q[5]; q[5] = (T)Q[8];
        } else {
            for (int k = 0; k < 6; k++) q[k] = (T)0;
        }
        q[6] = (T)0;
        q[7] = (T)0;
    }
}

// forest view for the pack: system s owns slots [soff[s], soff[s+1]) and nodes [roots[s], roots[s+1]);
// FP32 packs store coordinates relative to the system's cube centre, FP64 packs absolute ones
struct PackForest {
    int64_t nsys;
    const int64_t *soff;
    const int32_t *roots;
    const double *cubes;
};

template <typename T>
__global__ __launch_bounds__(256) void k_pack(int64_t n, int64_t nnodes, const ptg_tree_arrays t, const ptg_walk_tree wt,
                                              PackForest f)
{
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n + ptg_node_count(nnodes, t.node_count)) return;  // a deferred tree may hold fewer nodes
    int64_t sl = t.slot[e];
    if (sl < 0 || sl >= wt.nslots) return;
    double origin[3] = {wt.origin[0], wt.origin[1], wt.origin[2]};
    if (f.nsys > 0) {
        const int64_t sys = f.nsys == 1 ? 0 : e < n ? upper_index(f.soff, f.nsys, e) : upper_index(f.roots, f.nsys, e);
        const bool rel = sizeof(T) == 4;
        for (int k = 0; k < 3; k++) origin[k] = rel ? f.cubes[4 * sys + k] : 0.0;
    }
    pack_slot<T>(e, sl, n, t, wt, origin);
}

// The same copy written in slot order (coalesced stores), for the build: elem maps each slot to its
// element (the build's slot map), so the reads are the gathered side.
template <typename T>
__global__ __launch_bounds__(256) void k_pack_slots(int64_t n, int64_t ne, const ptg_tree_arrays t,
                                                    const ptg_walk_tree wt, PackForest f, const int32_t *elem)
{
    int64_t sl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= n + ptg_node_count(ne - n, t.node_count)) return;  // a deferred tree may hold fewer nodes
    const int64_t e = elem[sl];
    if (e < 0 || e >= ne) return;
    double origin[3] = {wt.origin[0], wt.origin[1], wt.origin[2]};
    if (f.nsys > 0) {
        const int64_t sys = f.nsys == 1 ? 0 : e < n ? upper_index(f.soff, f.nsys, e) : upper_index(f.roots, f.nsys, e);
        const bool rel = sizeof(T) == 4;
        for (int k = 0; k < 3; k++) origin[k] = rel ? f.cubes[4 * sys + k] : 0.0;
    }
    pack_slot<T>(e, sl, n, t, wt, origin);
}

static ptg_status pack_in_slot_order(const ptg_tree_arrays &t, int64_t n, int64_t nnodes, int64_t nsys,
                                     const ptg_forest *forest, const ptg_walk_tree *pack, const int32_t *elem,
                                     hipStream_t s)
{
    PackForest f{forest ? nsys : 0, forest ? forest->offsets : nullptr, forest ? forest->roots : nullptr,
                 forest ? forest->cubes : nullptr};
    const int64_t ne = n + nnodes;
    if (pack->prec == PTG_F32)
        k_pack_slots<float><<<grid1(ne), 256, 0, s>>>(n, ne, t, *pack, f, elem);
    else
        k_pack_slots<double><<<grid1(ne), 256, 0, s>>>(n, ne, t, *pack, f, elem);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

extern "C" ptg_status ptg_pack_walk(const ptg_tree_arrays *tree, int64_t n, int64_t nnodes, const ptg_forest *forest,
                                    const ptg_walk_tree *wt, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (!tree || !wt || !wt->xyzm || !wt->aux || !wt->kids || n < 1 || nnodes < 1 || wt->nslots != n + nnodes) {
        ptg_set_error("ptg_pack_walk: bad arguments");
        return PTG_ERR_ARG;
    }
    int64_t nsys;
    ptg_status st = forest_check(forest, n, "ptg_pack_walk", &nsys);
    if (st) return st;
    PackForest f{forest ? nsys : 0, forest ? forest->offsets : nullptr, forest ? forest->roots : nullptr,
                 forest ? forest->cubes : nullptr};
    if (wt->quad && !tree->quadrupoles) {
        ptg_set_error("ptg_pack_walk: a quadrupole pack needs a tree with quadrupoles");
        return PTG_ERR_ARG;
    }
    if (wt->prec == PTG_F32) k_pack<float><<<grid1(n + nnodes), 256, 0, s>>>(n, nnodes, *tree, *wt, f);
    else if (wt->prec == PTG_F64) k_pack<double><<<grid1(n + nnodes), 256, 0, s>>>(n, nnodes, *tree, *wt, f);
    else {
        ptg_set_error("ptg_pack_walk: prec must be PTG_F32 or PTG_F64");
        return PTG_ERR_ARG;
    }
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

// ============================================================================================
// cell-aligned target groups for the FP32 walk
//
// Fixed runs of group_size Morton-consecutive targets can straddle the boundary of a large cell:
// such a group's bounding box spans the cell and its walk opens most of the tree.  Here a group
// never crosses the boundary of an octree cell that holds more than `cell_cap` targets: targets
// are first split into maximal cells of at most cell_cap (= 4 x group_size) targets, and each such
// cell is cut into ceil(len / group_size) near-equal consecutive chunks.
// ============================================================================================
__global__ __launch_bounds__(256) void k_sorted_keys(const double *pos, const int64_t *perm, int64_t n,
                                                     const Geom *geoms, const int64_t *soff, int64_t nsys,
                                                     uint64_t *keys)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int64_t o = perm[t];
    const Geom *g = geoms + (nsys > 1 ? upper_index(soff, nsys, t) : 0);
    keys[t] = morton_key(pos[3 * o], pos[3 * o + 1], pos[3 * o + 2], g->cmin, g->inv);
}

// One workgroup per tile of GT_TILE consecutive sorted targets.  A capped cell starts at t when t
// starts a system or the smallest cell holding t-1 and t (the cell of the L levels their keys share)
// has more than cap = 4 x group_size targets.  With keys sorted within each system that cell is a
// contiguous run containing t-1 and t, so it holds more than cap targets exactly when one of the
// cap + 1-target windows [j, j + cap] with t - cap <= j <= t - 1 lies inside it, i.e. when the
// levels keys[j] and keys[j + cap] share (the window's own smallest cell; the window must not leave
// the system) are at least L.  Each tile therefore sets, for every level l = 0..21, one bit per
// window whose shared levels are >= l, and a cut flag is the OR of at most cap / 64 + 2 LDS words
// of level L's mask: no search.  Between two cuts there are at most cap targets (a longer stretch
// would contain a window inside the cell of its least-sharing adjacent pair), so the capped cell
// holding a target of the tile is found from the cut flags within margin >= cap of the tile.
#define GT_TILE 1024
#define GT_WORDS (GT_TILE / 64)
#define GT_MAX_MARGIN 1024  // group_size <= 256
#define GT_LEVELS 22

static size_t group_tiles_lds(int64_t margin, int64_t cap)
{
    const int64_t nkeys = GT_TILE + 3 * margin + cap, nwin = GT_TILE + 3 * margin;
    return (size_t)(nkeys + GT_LEVELS * (nwin / 64) + ptg_cdiv(nkeys, 64)) * sizeof(uint64_t);
}

// OR of bits [a, b] (a <= b) of the bitmask m
__device__ __forceinline__ uint64_t bits_any(const uint64_t *m, int64_t a, int64_t b)
{
    const int64_t wa = a >> 6, wb = b >> 6;
    uint64_t acc = 0;
    for (int64_t w = wa; w <= wb; ++w) {
        uint64_t x = m[w];
        if (w == wa) x &= ~0ull << (a & 63);
        if (w == wb) x &= ~0ull >> (63 - (b & 63));
        acc |= x;
    }
    return acc;
}

// gmask[tile]: one bit per target of the tile, set where a group starts; tcount[tile]: their number.
// cells == 0 is the fixed grouping (runs of gs restarted at every system start).  Dynamic LDS in
// cells mode: group_tiles_lds(margin, cap) bytes, holding the keys of positions
// [T0 - 2 margin, T0 + GT_TILE + margin + cap), the level masks of the windows starting in
// [T0 - 2 margin, T0 + GT_TILE + margin) and the system-start bits of the key positions.  Four
// wavefronts per tile, so that several tiles share a CU and hide each other's load latency.
#define GT_THREADS 256
__global__ __launch_bounds__(GT_THREADS) void k_group_tiles(const uint64_t *keys, int64_t n, const int64_t *soff,
                                                            int64_t nsys, int gs, int cells, int64_t margin,
                                                            uint64_t *gmask, int32_t *tcount)
{
    extern __shared__ uint64_t sk[];
    __shared__ uint64_t cutw[(GT_TILE + 2 * GT_MAX_MARGIN) / 64];
    __shared__ int32_t wcnt[GT_WORDS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t T0 = (int64_t)blockIdx.x * GT_TILE;
    const int64_t cap = 4 * (int64_t)gs;
    const int64_t K0 = T0 - 2 * margin;  // first key position held, also the first window
    if (cells) {
        const int64_t nwin = GT_TILE + 3 * margin;  // windows j in [K0, K0 + nwin), a multiple of 64
        const int64_t nkeys = nwin + cap;           // keys of positions [K0, K0 + nkeys)
        const int64_t nw = nwin >> 6;
        uint64_t *lvm = sk + nkeys;                 // level l's mask: lvm[l * nw + word]
        uint64_t *sst = lvm + GT_LEVELS * nw;       // system starts among the key positions
        // positions outside [0, n) load a clamped copy that is never read
#pragma unroll 4
        for (int64_t i = threadIdx.x; i < nkeys; i += GT_THREADS)
            sk[i] = keys[std::min<int64_t>(std::max<int64_t>(K0 + i, 0), n - 1)];
        if (nsys > 1) {
            // the system starts among the key positions: the first at or after K0 by one binary
            // search, then the following ones (few per window) one per thread
            __shared__ int64_t s0;
            const int64_t nsw = (nkeys + 63) >> 6;
            for (int64_t q = threadIdx.x; q < nsw; q += GT_THREADS) sst[q] = 0;
            if (threadIdx.x == 0) {
                int64_t lo = 0, hi = nsys;  // first system whose start is >= max(K0, 0)
                const int64_t v = K0 > 0 ? K0 : 0;
                while (lo < hi) {
                    const int64_t mid = (lo + hi) >> 1;
                    if (soff[mid] < v) lo = mid + 1;
                    else hi = mid;
                }
                s0 = lo;
            }
            __syncthreads();
            for (int64_t sy = s0 + threadIdx.x; sy < nsys && soff[sy] < K0 + nkeys; sy += GT_THREADS) {
                const int64_t i = soff[sy] - K0;
                atomicOr((unsigned long long *)&sst[i >> 6], 1ull << (i & 63));
            }
        }
        __syncthreads();
        // the levels each window [j, j + cap] shares, -1 where it leaves the targets or its system;
        // level l's bit is set where that is >= l: below the wave's smallest level every valid window
        // qualifies, above its largest none does, so only the levels in between take a ballot
        for (int64_t base = 64 * wave; base < nwin; base += GT_THREADS) {
            const int64_t j = K0 + base + lane;
            int lc = -1;
            if (j >= 0 && j + cap < n && (nsys == 1 || !bits_any(sst, j + 1 - K0, j + cap - K0)))
                lc = ptg_common_levels(sk[j - K0], sk[j + cap - K0]);
            int lo = lc < 0 ? GT_LEVELS : lc, hi = lc;
            for (int d = 32; d >= 1; d >>= 1) {
                lo = min(lo, __shfl_xor(lo, d));
                hi = max(hi, __shfl_xor(hi, d));
            }
            const uint64_t valid = __ballot(lc >= 0);
            uint64_t mine = lane <= lo ? valid : 0;
            for (int l = lo + 1; l <= hi; ++l) {
                const uint64_t b = __ballot(lc >= l);
                if (lane == l) mine = b;
            }
            if (lane < GT_LEVELS) lvm[lane * nw + (base >> 6)] = mine;
        }
        __syncthreads();
        // cut flags of positions [T0 - margin, T0 + GT_TILE + margin), one ballot word per 64; a
        // position at or past n counts as a cut (the end of the last cell), one before 0 never does
        // (position 0 always starts a cell)
        const int64_t E = GT_TILE + 2 * margin;
        for (int64_t base = 64 * wave; base < E; base += GT_THREADS) {
            const int64_t p = T0 - margin + base + lane;
            bool c = p >= n;
            if (p >= 0 && p < n) {
                const int64_t i = p - K0;
                if (p == 0 || (nsys > 1 && ((sst[i >> 6] >> (i & 63)) & 1))) {
                    c = true;
                } else {
                    const int L = ptg_common_levels(sk[i - 1], sk[i]);
                    c = bits_any(lvm + L * nw, i - cap, i - 1) != 0;
                }
            }
            const uint64_t b = __ballot(c);
            if (lane == 0) cutw[base >> 6] = b;
        }
        __syncthreads();
    }
    if (cells) {
        // group starts from every capped cell [ps, pe) starting in [T0 - margin, T0 + GT_TILE) (one
        // starting earlier ends before the tile, as cells hold at most cap <= margin targets): its
        // chunks of cs = ceil(len / ceil(len / gs)) targets start at ps, ps + cs, ...
        __shared__ uint32_t gm[GT_TILE / 32];
        for (int w = threadIdx.x; w < GT_TILE / 32; w += GT_THREADS) gm[w] = 0;
        __syncthreads();
        const int nwc = (int)((GT_TILE + 2 * margin) >> 6);
        for (int wi = threadIdx.x; wi < (int)((margin + GT_TILE) >> 6); wi += GT_THREADS) {
            uint64_t m = cutw[wi];
            while (m) {
                const int bit = __ffsll((long long)m) - 1;
                m &= m - 1;
                const int64_t ps = T0 - margin + 64 * wi + bit;
                if (ps >= n) break;
                // the next cut, within cap of ps, so inside the margin
                const int64_t i1 = ps - (T0 - margin) + 1;
                int w2 = (int)(i1 >> 6);
                uint64_t m2 = cutw[w2] & (~0ull << (i1 & 63));
                while (!m2 && w2 + 1 < nwc) m2 = cutw[++w2];
                if (!m2) break;  // not for keys sorted within each system
                const int64_t pe = T0 - margin + 64 * w2 + __ffsll((long long)m2) - 1;
                // len <= cap <= 1024: 32-bit arithmetic
                const uint32_t len = (uint32_t)(pe - ps), ugs = (uint32_t)gs;
                const uint32_t nch = (len + ugs - 1) / ugs;
                const uint32_t cs = (len + nch - 1) / nch;
                for (int64_t q = ps; q < pe; q += cs)
                    if (q >= T0 && q < T0 + GT_TILE) atomicOr(&gm[(q - T0) >> 5], 1u << ((q - T0) & 31));
            }
        }
        __syncthreads();
        for (int w = threadIdx.x; w < GT_WORDS; w += GT_THREADS) {
            const uint64_t b = (uint64_t)gm[2 * w] | ((uint64_t)gm[2 * w + 1] << 32);
            gmask[(int64_t)blockIdx.x * GT_WORDS + w] = b;
            wcnt[w] = __popcll(b);
        }
    } else {
        for (int k = 0; k < GT_TILE / GT_THREADS; ++k) {
            const int64_t t = T0 + k * GT_THREADS + threadIdx.x;
            bool g = false;
            if (t < n) {
                const int64_t start = nsys > 1 ? soff[upper_index(soff, nsys, t)] : 0;
                g = (t - start) % gs == 0;
            }
            const uint64_t b = __ballot(g);
            const int word = k * (GT_THREADS / 64) + wave;
            if (lane == 0) {
                gmask[(int64_t)blockIdx.x * GT_WORDS + word] = b;
                wcnt[word] = __popcll(b);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t c = 0;
        for (int w = 0; w < GT_WORDS; ++w) c += wcnt[w];
        tcount[blockIdx.x] = c;
    }
}

// offsets[tbase[tile] + rank] = position of each group start, one lane per 64-target word (a wave
// covers four tiles); the last word closes the list
__global__ __launch_bounds__(256) void k_group_write(const uint64_t *gmask, const int32_t *tcount,
                                                     const int32_t *tbase, int64_t n, int64_t ntiles,
                                                     int32_t *offsets, int32_t *ngroups)
{
    const int64_t word = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nwords = ntiles * GT_WORDS;
    const int sub = threadIdx.x & (GT_WORDS - 1);
    uint64_t w = word < nwords ? gmask[word] : 0;
    // exclusive prefix of the popcounts within each tile's GT_WORDS lanes
    int c = __popcll(w), incl = c;
    for (int d = 1; d < GT_WORDS; d <<= 1) {
        const int v = __shfl_up(incl, d, GT_WORDS);
        if (sub >= d) incl += v;
    }
    if (word >= nwords) return;
    const int64_t tile = word / GT_WORDS;
    int32_t r = tbase[tile] + incl - c;
    const int64_t p0 = word * 64;
    while (w) {
        offsets[r++] = (int32_t)(p0 + __ffsll((long long)w) - 1);
        w &= w - 1;
    }
    if (word == nwords - 1) {
        const int32_t total = tbase[tile] + tcount[tile];
        offsets[total] = (int32_t)n;
        *ngroups = total;
    }
}

struct GroupWS {
    double *part;
    Geom *geom;
    uint64_t *keys;
    uint64_t *gmask;
    int32_t *tcount, *tbase, *bsum;
};

static int64_t group_tiles(int64_t n) { return std::max<int64_t>(1, ptg_cdiv(n, GT_TILE)); }

static bool carve_groups(PtgArena &A, int64_t n, int64_t nsys, GroupWS &w)
{
    const int64_t nt = group_tiles(n);
    w.part = A.take<double>(6 * BB_BLOCKS);
    w.geom = A.take<Geom>(nsys);
    w.keys = A.take<uint64_t>(n);
    w.gmask = A.take<uint64_t>(nt * GT_WORDS);
    w.tcount = A.take<int32_t>(nt);
    w.tbase = A.take<int32_t>(nt);
    w.bsum = A.take<int32_t>(scan_blocks(nt) + 64);
    return A.ok();
}

extern "C" size_t ptg_groups_workspace_bytes(int64_t n, int64_t nsys)
{
    PtgArena A(nullptr, 0);
    GroupWS w;
    carve_groups(A, std::max<int64_t>(n, 1), std::max<int64_t>(nsys, 1), w);
    return A.used;
}

extern "C" ptg_status ptg_target_groups(const double *pos, const int64_t *perm, const uint64_t *keys, int64_t n,
                                        const ptg_forest *forest, int group_size, int mode, int32_t *offsets,
                                        int32_t *ngroups, void *ws, size_t ws_bytes, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (n < 1 || !pos || !perm || !offsets || !ngroups || group_size < 1 || group_size > 256) {
        ptg_set_error("ptg_target_groups: bad arguments");
        return PTG_ERR_ARG;
    }
    if (n > 0x7fffffffLL) {
        ptg_set_error("ptg_target_groups: n exceeds the int32 range");
        return PTG_ERR_ARG;
    }
    int64_t nsys;
    ptg_status st = forest_check(forest, n, "ptg_target_groups", &nsys);
    if (st) return st;
    PtgArena A(ws, ws_bytes);
    GroupWS w;
    if (!ws || !carve_groups(A, n, nsys, w)) {
        ptg_set_error("ptg_target_groups: workspace too small (%zu < %zu)", ws_bytes,
                      ptg_groups_workspace_bytes(n, nsys));
        return PTG_ERR_WORKSPACE;
    }
    const int64_t *soff = nsys > 1 ? forest->offsets : nullptr;
    if (mode != PTG_GROUPS_CELLS && mode != PTG_GROUPS_RUNS) {
        ptg_set_error("ptg_target_groups: mode must be PTG_GROUPS_RUNS or PTG_GROUPS_CELLS");
        return PTG_ERR_ARG;
    }
    const bool cells = mode == PTG_GROUPS_CELLS;
    if (cells && !keys) {
        if (nsys > 1) {
            // targets are the forest's own particles: system s owns input indices and sorted
            // positions [offsets[s], offsets[s+1])
            k_bbox_sys<<<(unsigned)nsys, 256, 0, s>>>(pos, soff, w.geom);
        } else {
            int nb = (int)std::min<int64_t>(BB_BLOCKS, std::max<int64_t>(1, ptg_cdiv(n, 1024)));
            k_bbox_partial<<<nb, 256, 0, s>>>(pos, n, w.part);
            k_bbox_final<<<1, 256, 0, s>>>(w.part, nb, w.geom);
        }
        k_sorted_keys<<<grid1(n), 256, 0, s>>>(pos, perm, n, w.geom, soff, nsys, w.keys);
        keys = w.keys;
    }
    const int64_t cap = 4 * (int64_t)group_size;
    const int64_t margin = cells ? ptg_cdiv(cap, 64) * 64 : 0;  // <= GT_MAX_MARGIN for group_size <= 256
    const size_t lds = cells ? group_tiles_lds(margin, cap) : 0;
    const int64_t nt = group_tiles(n);
    k_group_tiles<<<(unsigned)nt, GT_THREADS, lds, s>>>(keys, n, soff, nsys, group_size, cells ? 1 : 0, margin, w.gmask,
                                                     w.tcount);
    PTG_LAUNCH_CHECK();
    st = excl_scan(LoadArr{w.tcount}, nt, w.tbase, w.bsum, nullptr, s);
    if (st) return st;
    k_group_write<<<(unsigned)ptg_cdiv(nt * GT_WORDS, 256), 256, 0, s>>>(w.gmask, w.tcount, w.tbase, n, nt, offsets,
                                                                       ngroups);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}

// ============================================================================================
// field-scale sample targets: `samples` of each system's target groups, evenly spaced over the
// system's groups (group gfirst + k * gnum / samples for k < samples), each padded to gs targets by
// repeating its last target; row (s * samples + k) * gs + j of pos / soft.  The groups of system s
// are those whose start lies in [soff[s], soff[s+1]): a binary search over the group starts.
// ============================================================================================
__device__ __forceinline__ int64_t first_group_at(const int32_t *offsets, int64_t ng, int64_t v)
{
    int64_t lo = 0, hi = ng;  // first group whose start is >= v (ng if none)
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (offsets[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_field_sample(const int32_t *offsets, const int32_t *ngroups,
                                                      const int64_t *soff, int64_t n, int64_t nsys, int samples,
                                                      int gs, const double *coords, const double *soft, double *pos,
                                                      double *h)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nsys * samples * gs) return;
    const int64_t j = r % gs, sk = r / gs, k = sk % samples, s = sk / samples;
    const int64_t ng = *ngroups;
    int64_t g0 = 0, g1 = ng;
    if (nsys > 1) {
        g0 = first_group_at(offsets, ng, soff[s]);
        g1 = s + 1 < nsys ? first_group_at(offsets, ng, soff[s + 1]) : ng;
    }
    const int64_t g = g0 + (k * (g1 - g0)) / samples;
    const int64_t a = offsets[g], b = offsets[g + 1];
    const int64_t i = a + std::min<int64_t>(j, b - a - 1);
    pos[3 * r] = coords[3 * i];
    pos[3 * r + 1] = coords[3 * i + 1];
    pos[3 * r + 2] = coords[3 * i + 2];
    h[r] = soft[i];
}

extern "C" ptg_status ptg_field_sample(const int32_t *offsets, const int32_t *ngroups, const ptg_forest *forest,
                                       int64_t n, int samples, int group_size, const double *coordinates,
                                       const double *softenings, double *pos, double *soft, void *stream)
{
    if (n < 1 || !offsets || !ngroups || samples < 1 || group_size < 1 || !coordinates || !softenings || !pos ||
        !soft) {
        ptg_set_error("ptg_field_sample: bad arguments");
        return PTG_ERR_ARG;
    }
    int64_t nsys;
    ptg_status st = forest_check(forest, n, "ptg_field_sample", &nsys);
    if (st) return st;
    const int64_t rows = nsys * samples * (int64_t)group_size;
    k_field_sample<<<grid1(rows), 256, 0, (hipStream_t)stream>>>(offsets, ngroups, nsys > 1 ? forest->offsets : nullptr,
                                                                 n, nsys, samples, group_size, coordinates, softenings,
                                                                 pos, soft);
    PTG_LAUNCH_CHECK();
    return PTG_OK;
}
