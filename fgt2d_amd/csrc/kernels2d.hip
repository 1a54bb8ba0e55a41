// HIP kernels of the 2D plane-wave fast Gauss transform for gfx950.
//
// Layouts (all in HBM, tree order unless said otherwise):
//   points      double[2*n]        x,y interleaved (the reference's (2,n))
//   densities   double[nd*n]       (nd,n)
//   expansions  double[2*nexp*nd]  per slot, complex interleaved, index
//                                  a + (npw/2)*b (a: positive x nodes,
//                                  b: all y nodes), as pwexp(npw/2,npw,nd)
//   box arrays  indexed by box id of the host tree
//
// 64-wide wavefronts; 256-thread blocks (4 waves) throughout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "fgtpw_impl.hpp"

namespace fgtpw {

using u64 = unsigned long long;

static inline unsigned grid_for(long n, int bs)
{
    long g = (n + bs - 1) / bs;
    return (unsigned)(g < 1 ? 1 : g);
}

// ---------------------------------------------------------------------------
// bounding box of sources and targets (pts_tree_boxsize0, pts_tree2d4.f:1187)
// min/max are exact, so the reduction order does not matter.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bbox(const double* __restrict__ a, int na,
                                              const double* __restrict__ b, int nb,
                                              double* __restrict__ part)
{
    __shared__ double s[5][256];
    double xmn = HUGE_VAL, xmx = -HUGE_VAL, ymn = HUGE_VAL, ymx = -HUGE_VAL, bad = 0;
    long n = (long)na + nb;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double* p = i < na ? a + 2 * i : b + 2 * (i - na);
        double x = p[0], y = p[1];
        if (!isfinite(x) || !isfinite(y)) bad = 1;
        xmn = fmin(xmn, x);
        xmx = fmax(xmx, x);
        ymn = fmin(ymn, y);
        ymx = fmax(ymx, y);
    }
    int t = threadIdx.x;
    s[0][t] = xmn; s[1][t] = xmx; s[2][t] = ymn; s[3][t] = ymx; s[4][t] = bad;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            s[0][t] = fmin(s[0][t], s[0][t + w]);
            s[1][t] = fmax(s[1][t], s[1][t + w]);
            s[2][t] = fmin(s[2][t], s[2][t + w]);
            s[3][t] = fmax(s[3][t], s[3][t + w]);
            s[4][t] = fmax(s[4][t], s[4][t + w]);
        }
        __syncthreads();
    }
    if (t < 5) part[5 * blockIdx.x + t] = s[t][0];
}

__global__ __launch_bounds__(256) void k_bbox_final(const double* __restrict__ part, int nblk,
                                                    double* __restrict__ out)
{
    __shared__ double s[5][256];
    int t = threadIdx.x;
    double v[5] = {HUGE_VAL, -HUGE_VAL, HUGE_VAL, -HUGE_VAL, 0};
    for (int i = t; i < nblk; i += 256) {
        v[0] = fmin(v[0], part[5 * i]);
        v[1] = fmax(v[1], part[5 * i + 1]);
        v[2] = fmin(v[2], part[5 * i + 2]);
        v[3] = fmax(v[3], part[5 * i + 3]);
        v[4] = fmax(v[4], part[5 * i + 4]);
    }
    for (int k = 0; k < 5; ++k) s[k][t] = v[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            s[0][t] = fmin(s[0][t], s[0][t + w]);
            s[1][t] = fmax(s[1][t], s[1][t + w]);
            s[2][t] = fmin(s[2][t], s[2][t + w]);
            s[3][t] = fmax(s[3][t], s[3][t + w]);
            s[4][t] = fmax(s[4][t], s[4][t + w]);
        }
        __syncthreads();
    }
    if (t < 5) out[t] = s[t][0];
}

void launch_bbox(const double* a, int na, const double* b, int nb, double* part, int nblk,
                 double* out, hipStream_t s)
{
    k_bbox<<<nblk, 256, 0, s>>>(a, na, b, nb, part);
    k_bbox_final<<<1, 256, 0, s>>>(part, nblk, out);
}

// ---------------------------------------------------------------------------
// full-depth Morton keys.  The centre of a child is parent +- bs/2 in FP64
// and a point goes to the lower child when p - c < 0 (sort_pts_to_children,
// pts_tree2d4.f:594-689; tree_refine_boxes, tree_routs4.f:34-39): every
// comparison the reference makes along the point's path is made here with
// bit-identical centres, so the key prefix of any level names the point's
// box at that level.  Digit = xbit + 2*ybit, most significant = level 1.
// ---------------------------------------------------------------------------
template <class K>  // 32- or 64-bit keys
__global__ __launch_bounds__(256) void k_keys2d(const double* __restrict__ xy, int n, double cx0,
                                                double cy0, double bs0, int nlmax,
                                                K* __restrict__ key, unsigned* __restrict__ idx)
{
#pragma clang fp contract(off)
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double x = xy[2 * (long)i], y = xy[2 * (long)i + 1];
    double cx = cx0, cy = cy0, bs = bs0;
    K k = 0;
    for (int l = 0; l < nlmax; ++l) {
        bs = bs / 2;
        double h = bs / 2;
        int xb = !(x - cx < 0.0);
        int yb = !(y - cy < 0.0);
        k = (k << 2) | (K)(xb | (yb << 1));
        cx = cx + (xb ? h : -h);
        cy = cy + (yb ? h : -h);
    }
    key[i] = k;
    if (idx) idx[i] = (unsigned)i;
}

template <class K>
static void keys2d_launch(const double* xy, int n, double cx, double cy, double bs0, int nlmax,
                          K* key, unsigned* idx, hipStream_t s)
{
    if (n <= 0) return;
    k_keys2d<K><<<grid_for(n, 256), 256, 0, s>>>(xy, n, cx, cy, bs0, nlmax, key, idx);
}
void launch_keys2d(const double* xy, int n, double cx, double cy, double bs0, int nlmax, u64* key,
                   unsigned* idx, hipStream_t s)
{
    keys2d_launch(xy, n, cx, cy, bs0, nlmax, key, idx, s);
}
void launch_keys2d(const double* xy, int n, double cx, double cy, double bs0, int nlmax,
                   unsigned* key, unsigned* idx, hipStream_t s)
{
    keys2d_launch(xy, n, cx, cy, bs0, nlmax, key, idx, s);
}

// ---------------------------------------------------------------------------
// Leaf level of every source in the adaptive (pre level-restriction) tree.
// A box of level l is refined iff it holds more than ndiv(l) sources
// (ndiv0 below the cutoff level, ndiv at or beyond it) and l < nlmax
// (pts_tree_build, pts_tree2d4.f:507-523).  Counts shrink with depth and
// ndiv(l) grows with l, so along a point's path the refined levels are
// 0..L-1.  L is found from the finest level upwards: the point's run of equal
// level-l prefixes in the sorted keys is located by two binary searches in
// the window [i - lim, i + lim] (prefixes are sorted, so "same prefix" is a
// contiguous interval); a run that fills the window on either side holds more
// than lim points.  A point whose leaf prefix differs from its predecessor's
// starts a leaf; each block of 256 points also counts its starts.
// ---------------------------------------------------------------------------
template <class K>
__global__ __launch_bounds__(256) void k_leafdepth(const K* __restrict__ key, int n, int nlmax,
                                                   int levcut, int ndiv0, int ndiv, int bits,
                                                   int* __restrict__ level,
                                                   unsigned char* __restrict__ start,
                                                   unsigned* __restrict__ bcnt)
{
    // the block's keys and a halo of max(ndiv0, ndiv) on each side: every
    // window the searches visit, in LDS
    extern __shared__ __align__(8) unsigned char ld_s[];
    K* sk = (K*)ld_s;
    const int halo = max(ndiv0, ndiv);
    const int b0 = blockIdx.x * 256, lo0 = b0 - halo;
    for (int j = threadIdx.x; j < 256 + 2 * halo; j += 256) {
        const int g = lo0 + j;
        sk[j] = g >= 0 && g < n ? key[g] : K(0);
    }
    __syncthreads();
    const int i = b0 + threadIdx.x;
    int st = 0;
    if (i < n) {
        const K ki = sk[i - lo0];
        int L = 0;
        for (int l = nlmax - 1; l >= 0; --l) {
            const int sh = bits * (nlmax - l);  // < the key width (host guarantee)
            const K p = ki >> sh;
            const int lim = l < levcut ? ndiv0 : ndiv;
            // first index in [max(0, i - lim), i] with prefix p
            int a = max(0, i - lim), b = i;
            while (a < b) {
                int m = (a + b) >> 1;
                if ((sk[m - lo0] >> sh) == p) b = m; else a = m + 1;
            }
            const int lo = a;
            // last index in [i, min(n - 1, i + lim)] with prefix p
            a = i;
            b = min(n - 1, i + lim);
            while (a < b) {
                int m = (a + b + 1) >> 1;
                if ((sk[m - lo0] >> sh) == p) a = m; else b = m - 1;
            }
            if (a - lo + 1 > lim) {
                L = l + 1;
                break;
            }
        }
        level[i] = L;
        const int sh = bits * (nlmax - L);
        st = 1;
        if (i > 0) st = sh >= 8 * (int)sizeof(K) ? 0 : ((sk[i - 1 - lo0] >> sh) != (ki >> sh));
        start[i] = (unsigned char)st;
    }
    const int c = __syncthreads_count(st);  // every thread of the block reaches it
    if (threadIdx.x == 0) bcnt[blockIdx.x] = (unsigned)c;
}

template <class K>
static void leafdepth_launch(const K* key, int n, int nlmax, int levcut, int ndiv0, int ndiv,
                             int* level, unsigned char* start_flag, unsigned* bcnt, hipStream_t s,
                             int bits)
{
    const size_t lds = sizeof(K) * (256 + 2 * (size_t)std::max(ndiv0, ndiv));
    k_leafdepth<K><<<grid_for(n, 256), 256, lds, s>>>(key, n, nlmax, levcut, ndiv0, ndiv, bits,
                                                      level, start_flag, bcnt);
}
void launch_leafdepth(const u64* key, int n, int nlmax, int levcut, int ndiv0, int ndiv,
                      int* level, unsigned char* start_flag, unsigned* bcnt, hipStream_t s, int bits)
{
    leafdepth_launch(key, n, nlmax, levcut, ndiv0, ndiv, level, start_flag, bcnt, s, bits);
}
void launch_leafdepth(const unsigned* key, int n, int nlmax, int levcut, int ndiv0, int ndiv,
                      int* level, unsigned char* start_flag, unsigned* bcnt, hipStream_t s, int bits)
{
    leafdepth_launch(key, n, nlmax, levcut, ndiv0, ndiv, level, start_flag, bcnt, s, bits);
}

// exclusive scan in place of the nb block counts (one block of 1024
// threads); *total gets their sum.  Chunks of 8192 counts pass through LDS
// (coalesced loads and stores, all in flight at once): each thread scans 8
// consecutive counts of the chunk, the waves' shuffles and totals do the
// rest.
__global__ __launch_bounds__(1024) void k_scan_counts(unsigned* __restrict__ c, int nb,
                                                      int* __restrict__ total)
{
    __shared__ unsigned buf[8192];
    __shared__ unsigned wsum[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    unsigned carry = 0;
    for (int c0 = 0; c0 < nb; c0 += 8192) {
        const int m = min(8192, nb - c0);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int b = k * 1024 + t;
            buf[b] = b < m ? c[c0 + b] : 0u;
        }
        __syncthreads();
        unsigned v[8], sum = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            v[k] = buf[8 * t + k];
            sum += v[k];
        }
        unsigned x = sum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        unsigned run = carry + x - sum, all = 0;
        for (int ww = 0; ww < 16; ++ww) {
            if (ww < w) run += wsum[ww];
            all += wsum[ww];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            buf[8 * t + k] = run;
            run += v[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int b = k * 1024 + t;
            if (b < m) c[c0 + b] = buf[b];
        }
        carry += all;
        __syncthreads();  // buf and wsum are rewritten by the next chunk
    }
    if (t == 0) *total = (int)carry;
}

// the leaf records in Morton order: a leaf's place is the number of starts
// before it (its block's offset plus the starts of the earlier lanes of the
// block)
template <class K>
__global__ __launch_bounds__(256) void k_leaf_records(const unsigned char* __restrict__ start,
                                                      const unsigned* __restrict__ boff, int n,
                                                      const K* __restrict__ key,
                                                      const int* __restrict__ level, int nlmax,
                                                      int bits, LeafRun* __restrict__ out)
{
    __shared__ unsigned wc[4];
    const int i = blockIdx.x * 256 + threadIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool st = i < n && start[i];
    const unsigned long long m = __ballot(st);
    if (lane == 0) wc[w] = (unsigned)__popcll(m);
    __syncthreads();
    if (!st) return;
    unsigned r = boff[blockIdx.x] + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
    for (int ww = 0; ww < w; ++ww) r += wc[ww];
    const int L = level[i];
    LeafRun o;
    o.key = L == 0 ? 0ull : (u64)(key[i] >> (bits * (nlmax - L)));
    o.level = L;
    o.start = i;
    out[r] = o;
}

size_t leaf_records_temp(int n) { return 4 * ((size_t)n / 256 + 2); }

template <class K>
static void leaf_records_launch(const unsigned char* start_flag, unsigned* bcnt, int n,
                                const K* key, const int* level, int nlmax, int* nleaves,
                                LeafRun* out, hipStream_t s, int bits)
{
    const int nb = grid_for(n, 256);
    k_scan_counts<<<1, 1024, 0, s>>>(bcnt, nb, nleaves);
    k_leaf_records<K><<<nb, 256, 0, s>>>(start_flag, bcnt, n, key, level, nlmax, bits, out);
}
void launch_leaf_records(const unsigned char* start_flag, unsigned* bcnt, int n, const u64* key,
                         const int* level, int nlmax, int* nleaves, LeafRun* out, hipStream_t s,
                         int bits)
{
    leaf_records_launch(start_flag, bcnt, n, key, level, nlmax, nleaves, out, s, bits);
}
void launch_leaf_records(const unsigned char* start_flag, unsigned* bcnt, int n,
                         const unsigned* key, const int* level, int nlmax, int* nleaves,
                         LeafRun* out, hipStream_t s, int bits)
{
    leaf_records_launch(start_flag, bcnt, n, key, level, nlmax, nleaves, out, s, bits);
}

// ---------------------------------------------------------------------------
// sorting (rocPRIM)
// ---------------------------------------------------------------------------
// The scratch queries below ask rocPRIM for every configuration, and each
// query reads device attributes; the sizes depend on n only, so the last
// answer per thread is kept (a transform asks before its first launch).
size_t sort_pairs_u64_temp(int n)
{
    thread_local int last_n = -1;
    thread_local size_t last_b = 0;
    if (n == last_n) return last_b;
    size_t b = 0;
    HIPCHECK(rocprim::radix_sort_pairs(nullptr, b, (const u64*)nullptr, (u64*)nullptr,
                                       (const unsigned*)nullptr, (unsigned*)nullptr, n, 0, 64));
    last_n = n;
    last_b = b;
    return b;
}

void sort_pairs_u64(void* tmp, size_t tmpb, const u64* kin, u64* kout, const unsigned* vin,
                    unsigned* vout, int n, int endbit, hipStream_t s)
{
    HIPCHECK(rocprim::radix_sort_pairs(tmp, tmpb, kin, kout, vin, vout, n, 0, endbit, s));
}

// 32-bit keys: onesweep passes of 9, 11 or 10 bits so that the key widths
// of the trees (2 or 3 bits per level) take two passes up to 22 bits and
// three up to 30, instead of the 8-bit default's three and four (block
// shapes of the gfx950 default for 4-byte keys and values).
// The sort kernel's block shape is a tile shape of the tuning table: shape
// 0 is the gfx950 default for 4-byte keys and values (1024 threads x 16
// keys, about 120 blocks at 2M keys), 1..3 smaller blocks (512 x 8,
// 256 x 16, 256 x 8) for more blocks in flight.
template <unsigned BITS, unsigned BS, unsigned IPT>
using Onesweep32 = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 16>,
                                        rocprim::kernel_config<BS, IPT>, BITS,
                                        rocprim::block_radix_rank_algorithm::match>>;

template <class Cfg>
static size_t sort32_temp(int n, int endbit)
{
    size_t b = 0;
    HIPCHECK(rocprim::radix_sort_pairs<Cfg>(nullptr, b, (const unsigned*)nullptr, (unsigned*)nullptr,
                                            (const unsigned*)nullptr, (unsigned*)nullptr, n, 0,
                                            endbit));
    return b;
}

template <unsigned BS, unsigned IPT>
static size_t sort32_temp_all(int n)
{
    return std::max({sort32_temp<Onesweep32<9, BS, IPT>>(n, 18),
                     sort32_temp<Onesweep32<11, BS, IPT>>(n, 22),
                     sort32_temp<Onesweep32<10, BS, IPT>>(n, 30)});
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (32-bit key, 32-bit index) pairs for keys of at
// most 30 bits (tuning table key `sort` = 4 or 5): passes of at most 11 bits,
// each in three kernels over tiles of 256 R items (4 wavefronts x R rounds x
// 64 consecutive items, so that (wave, round, lane) is input order; R = 8,
// or 16 for digits of up to 10 bits under shape 5):
//   count:   the tile's digit histogram (LDS atomics), cnt[digit][tile];
//   offsets: per digit, the exclusive scan over tiles and the digit total;
//   scatter: an item's rank among its tile's items of its digit from
//            ballots (the lanes of its round with the same digit, then a
//            running count per wave and digit in LDS), plus the earlier
//            waves', tiles' and digits' counts.
// Equal keys keep their input order: the same permutation as any stable
// sort.  Without input values the first pass takes the indices 0..n-1.
// ---------------------------------------------------------------------------

// exclusive prefix of v over the 256 threads of a block (wavefront shuffles,
// then the four wave totals through LDS); *total gets the block's sum.
// wtot: 4 words of LDS, free again after the next barrier.
__device__ __forceinline__ unsigned rs_block_scan(unsigned v, unsigned* wtot, unsigned* total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) wtot[w] = x;
    __syncthreads();
    unsigned pre = 0;
    for (int ww = 0; ww < w; ++ww) pre += wtot[ww];
    *total = wtot[0] + wtot[1] + wtot[2] + wtot[3];
    return pre + x - v;
}

// exclusive scan, in place, of a[0..n) (n <= 256 * per) held in LDS; a
// barrier follows
__device__ __forceinline__ unsigned rs_lds_scan(unsigned* a, int n, unsigned* wtot)
{
    const int per = (n + 255) / 256, t = threadIdx.x;
    const int d0 = min(n, t * per), d1 = min(n, d0 + per);
    unsigned sum = 0;
    for (int d = d0; d < d1; ++d) sum += a[d];
    unsigned total;
    unsigned run = rs_block_scan(sum, wtot, &total);
    for (int d = d0; d < d1; ++d) {
        const unsigned c = a[d];
        a[d] = run;
        run += c;
    }
    __syncthreads();
    return total;
}

// the tile's digit histogram, cnt[digit][tile]; AGG: one LDS atomic per
// distinct digit of a wavefront (peers found by ballots), for passes whose
// digits repeat within a tile (high digits of clustered points); else one
// per key (low digits, nearly all distinct within a wavefront)
template <int R, bool AGG>
__global__ __launch_bounds__(256) void k_rs_count(const unsigned* __restrict__ key, int n, int shift,
                                                  int bits, unsigned* __restrict__ cnt, int ntile)
{
    constexpr int RS_TILE = 256 * R;
    extern __shared__ unsigned rs_h[];
    const int nd = 1 << bits;
    for (int d = threadIdx.x; d < nd; d += 256) rs_h[d] = 0;
    __syncthreads();
    const long base = (long)blockIdx.x * RS_TILE;
    const int lane = threadIdx.x & 63;
    unsigned kk[RS_TILE / 256];  // all loads in flight at once
#pragma unroll
    for (int r = 0; r < RS_TILE / 256; ++r) {
        const long i = base + threadIdx.x + 256 * r;
        kk[r] = i < n ? key[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < RS_TILE / 256; ++r) {
        const long i = base + threadIdx.x + 256 * r;
        const bool valid = i < n;
        const int d = valid ? (int)((kk[r] >> shift) & (unsigned)(nd - 1)) : 0;
        if (!AGG) {
            if (valid) atomicAdd(&rs_h[d], 1u);
            continue;
        }
        unsigned long long peers = __ballot(valid);
        for (int bit = 0; bit < bits; ++bit) {
            const unsigned long long m = __ballot((d >> bit) & 1);
            peers &= ((d >> bit) & 1) ? m : ~m;
        }
        if (valid && lane == __ffsll((long long)peers) - 1)
            atomicAdd(&rs_h[d], (unsigned)__popcll(peers));
    }
    __syncthreads();
    for (int d = threadIdx.x; d < nd; d += 256) cnt[(size_t)d * ntile + blockIdx.x] = rs_h[d];
}

// one block per digit: cnt[d][.] becomes its exclusive scan over the
// tiles, tot[d] its sum
__global__ __launch_bounds__(256) void k_rs_offsets(unsigned* __restrict__ cnt, int ntile,
                                                    unsigned* __restrict__ tot)
{
    __shared__ unsigned wtot[4];
    unsigned* c = cnt + (size_t)blockIdx.x * ntile;
    const int per = (ntile + 255) / 256, t = threadIdx.x;
    const int b0 = min(ntile, t * per), b1 = min(ntile, b0 + per);
    unsigned sum = 0;
    for (int b = b0; b < b1; ++b) sum += c[b];
    unsigned total;
    unsigned run = rs_block_scan(sum, wtot, &total);
    for (int b = b0; b < b1; ++b) {
        const unsigned v = c[b];
        c[b] = run;
        run += v;
    }
    if (t == 0) tot[blockIdx.x] = total;
}

template <int R>
__global__ __launch_bounds__(256) void k_rs_scatter(const unsigned* __restrict__ kin,
                                                    const unsigned* __restrict__ vin,
                                                    unsigned* __restrict__ kout,
                                                    unsigned* __restrict__ vout, int n, int shift,
                                                    int bits, const unsigned* __restrict__ pre,
                                                    const unsigned* __restrict__ tot, int ntile)
{
    constexpr int RS_ROUNDS = R, RS_TILE = 256 * R;
    extern __shared__ unsigned rs_s[];
    const int nd = 1 << bits, t = threadIdx.x, lane = t & 63, w = t >> 6;
    unsigned* wh = rs_s;               // 4 x nd: per-wave running counts, then wave offsets
    unsigned* dbase = rs_s + 4 * nd;   // nd: digits' first output position
    unsigned* lstart = rs_s + 5 * nd;  // nd: digits' first position in the tile
    unsigned* tk = rs_s + 6 * nd;      // RS_TILE keys in digit order
    unsigned* tv = tk + RS_TILE;       // RS_TILE values in digit order
    __shared__ unsigned wtot[4];
    const long i0 = (long)blockIdx.x * RS_TILE + (long)w * (RS_ROUNDS * 64);
    unsigned kk[RS_ROUNDS], vv[RS_ROUNDS], rk[RS_ROUNDS];
    int dg[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {  // all loads in flight at once
        const long i = i0 + r * 64 + lane;
        kk[r] = i < n ? kin[i] : 0u;
        vv[r] = i < n ? (vin ? vin[i] : (unsigned)i) : 0u;
    }
    for (int d = t; d < 4 * nd; d += 256) wh[d] = 0;
    for (int d = t; d < nd; d += 256) dbase[d] = tot[d];
    __syncthreads();
    rs_lds_scan(dbase, nd, wtot);  // the digit totals' exclusive scan
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned* my = wh + w * nd;
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const long i = i0 + r * 64 + lane;
        const bool valid = i < n;
        const int d = (int)((kk[r] >> shift) & (unsigned)(nd - 1));
        unsigned long long peers = __ballot(valid);
        for (int bit = 0; bit < bits; ++bit) {
            const unsigned long long m = __ballot((d >> bit) & 1);
            peers &= ((d >> bit) & 1) ? m : ~m;
        }
        const unsigned old = my[d];
        if (valid && lane == __ffsll((long long)peers) - 1) my[d] = old + (unsigned)__popcll(peers);
        rk[r] = old + (unsigned)__popcll(peers & below);
        dg[r] = valid ? d : -1;
    }
    __syncthreads();
    // the waves' counts per digit to exclusive offsets across the waves;
    // the tile's count per digit into lstart
    for (int d = t; d < nd; d += 256) {
        unsigned run = 0;
#pragma unroll
        for (int ww = 0; ww < 4; ++ww) {
            const unsigned c = wh[ww * nd + d];
            wh[ww * nd + d] = run;
            run += c;
        }
        lstart[d] = run;
    }
    __syncthreads();
    rs_lds_scan(lstart, nd, wtot);  // each digit's first place in the tile
    // the tile in digit order through LDS, then written out in that order:
    // consecutive threads write consecutive places of a digit's run
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int d = dg[r];
        if (d < 0) continue;
        const unsigned lp = lstart[d] + my[d] + rk[r];
        tk[lp] = kk[r];
        tv[lp] = vv[r];
    }
    __syncthreads();
    const int cnt = (int)min((long)RS_TILE, (long)n - (long)blockIdx.x * RS_TILE);
    for (int q = t; q < cnt; q += 256) {
        const unsigned k = tk[q];
        const int d = (int)((k >> shift) & (unsigned)(nd - 1));
        const unsigned pos = dbase[d] + pre[(size_t)d * ntile + blockIdx.x] + (q - lstart[d]);
        kout[pos] = k;
        vout[pos] = tv[q];
    }
}

static int rs_bits(int endbit, int* passes)
{
    const int np = (endbit + 10) / 11;
    *passes = np;
    return (endbit + np - 1) / np;
}

// scratch of the sort at the smaller tile (the larger one needs less)
static size_t rs_temp(int n)
{
    const size_t ntile = ((size_t)n + 2047) / 2048;
    return 256 + 2 * (((size_t)n * 4 + 255) & ~(size_t)255) + 4 * 2048 * (ntile + 1) + 4 * 2048;
}

template <int R>
static void rs_pass(const unsigned* ki, const unsigned* vi, unsigned* ko, unsigned* vo, int n,
                    int shift, int b, unsigned* cnt, unsigned* tot, hipStream_t s)
{
    const int ndb = 1 << b, ntile = (n + 256 * R - 1) / (256 * R);
    if (shift == 0)  // the lowest digit
        k_rs_count<R, false><<<ntile, 256, (size_t)ndb * 4, s>>>(ki, n, shift, b, cnt, ntile);
    else
        k_rs_count<R, true><<<ntile, 256, (size_t)ndb * 4, s>>>(ki, n, shift, b, cnt, ntile);
    k_rs_offsets<<<ndb, 256, 0, s>>>(cnt, ntile, tot);
    const size_t lds = ((size_t)6 * ndb + 2 * 256 * R) * 4;  // at most 64 KiB
    if (lds >= 65536) {
        static const bool set = [] {
            HIPCHECK(hipFuncSetAttribute((const void*)k_rs_scatter<R>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, 65536));
            return true;
        }();
        (void)set;
    }
    k_rs_scatter<R><<<ntile, 256, lds, s>>>(ki, vi, ko, vo, n, shift, b, cnt, tot, ntile);
}

static void rs_sort(void* tmp, size_t tmpb, const unsigned* kin, unsigned* kout, const unsigned* vin,
                    unsigned* vout, int n, int endbit, hipStream_t s, bool wide)
{
    if (n <= 0) return;
    if (tmpb < rs_temp(n)) fail(-2, "radix sort scratch too small");
    int np = 0;
    const int bits = rs_bits(std::max(endbit, 1), &np);
    const int ntile = (n + 2047) / 2048;
    const size_t pad = ((size_t)n * 4 + 255) & ~(size_t)255;
    char* p = (char*)tmp;
    unsigned* kb[2] = {(unsigned*)p, kout};
    unsigned* vb[2] = {(unsigned*)(p + pad), vout};
    unsigned* cnt = (unsigned*)(p + 2 * pad);
    unsigned* tot = cnt + (size_t)2048 * (ntile + 1);
    // ping-pong so that the last pass writes kout / vout
    const unsigned* ki = kin;
    const unsigned* vi = vin;  // null: the indices 0..n-1
    for (int pass = 0; pass < np; ++pass) {
        const int dst = (np - 1 - pass) % 2 == 0 ? 1 : 0;
        const int shift = pass * bits;
        const int b = std::min(bits, endbit - shift);
        if (wide && b <= 10)
            rs_pass<16>(ki, vi, kb[dst], vb[dst], n, shift, b, cnt, tot, s);
        else
            rs_pass<8>(ki, vi, kb[dst], vb[dst], n, shift, b, cnt, tot, s);
        ki = kb[dst];
        vi = vb[dst];
    }
}

size_t sort_pairs_u32_temp(int n)
{
    thread_local int last_n = -1;
    thread_local size_t last_b = 0;
    if (n == last_n) return last_b;
    last_b = std::max({sort32_temp_all<1024, 16>(n), sort32_temp_all<512, 8>(n),
                       sort32_temp_all<256, 16>(n), sort32_temp_all<256, 8>(n),
                       sort32_temp<rocprim::default_config>(n, 32), rs_temp(n)});
    last_n = n;
    return last_b;
}

template <unsigned BS, unsigned IPT>
static void sort32(void* tmp, size_t tmpb, const unsigned* kin, unsigned* kout,
                   const unsigned* vin, unsigned* vout, int n, int endbit, hipStream_t s)
{
    using C9 = Onesweep32<9, BS, IPT>;
    using C10 = Onesweep32<10, BS, IPT>;
    using C11 = Onesweep32<11, BS, IPT>;
    if (endbit <= 18)
        HIPCHECK(rocprim::radix_sort_pairs<C9>(tmp, tmpb, kin, kout, vin, vout, n, 0, endbit, s));
    else if (endbit <= 22)
        HIPCHECK(rocprim::radix_sort_pairs<C11>(tmp, tmpb, kin, kout, vin, vout, n, 0, endbit, s));
    else
        HIPCHECK(rocprim::radix_sort_pairs<C10>(tmp, tmpb, kin, kout, vin, vout, n, 0, endbit, s));
}

void sort_pairs_u32(void* tmp, size_t tmpb, const unsigned* kin, unsigned* kout,
                    const unsigned* vin, unsigned* vout, int n, int endbit, hipStream_t s,
                    int shape)
{
    if (endbit > 30) {
        HIPCHECK(rocprim::radix_sort_pairs(tmp, tmpb, kin, kout, vin, vout, n, 0, endbit, s));
        return;
    }
    switch (shape) {
    case 4: rs_sort(tmp, tmpb, kin, kout, vin, vout, n, endbit, s, false); break;
    case 5: rs_sort(tmp, tmpb, kin, kout, vin, vout, n, endbit, s, true); break;
    case 1: sort32<512, 8>(tmp, tmpb, kin, kout, vin, vout, n, endbit, s); break;
    case 2: sort32<256, 16>(tmp, tmpb, kin, kout, vin, vout, n, endbit, s); break;
    case 3: sort32<256, 8>(tmp, tmpb, kin, kout, vin, vout, n, endbit, s); break;
    default: sort32<1024, 16>(tmp, tmpb, kin, kout, vin, vout, n, endbit, s); break;
    }
}

// ---------------------------------------------------------------------------
// First point of every leaf of the final tree.  The points are sorted by
// full-depth Morton key and the leaves, in rank (Morton) order, tile the root
// box, so leaf r starts at the first point whose key is >= the leaf's
// smallest key, leafkey[r] (its prefix shifted to full depth).  One wavefront
// per leaf narrows [lo, hi] 64-fold per step: 64 pivots, one ballot.  Grid
// row 1, when present, does the same for a second point set (the targets).
// ---------------------------------------------------------------------------
template <class K>
__global__ __launch_bounds__(256) void k_leafstart_search(const K* __restrict__ key, int n,
                                                          const K* __restrict__ key2, int n2,
                                                          const K* __restrict__ leafkey,
                                                          int nleaves, int* __restrict__ ls,
                                                          int* __restrict__ ls2)
{
    if (blockIdx.y) {
        key = key2;
        n = n2;
        ls = ls2;
    }
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (r > nleaves) return;  // uniform over the wavefront
    if (r == nleaves) {
        if (lane == 0) ls[r] = n;
        return;
    }
    const K tgt = leafkey[r];
    // the answer, the first i with i == n or key[i] >= tgt, lies in [lo, hi];
    // a pivot at or past hi counts as >= tgt and is not read
    long lo = 0, hi = n;
    for (;;) {
        const long stride = hi - lo > 64 ? (hi - lo + 63) / 64 : 1;
        const long p = lo + lane * stride;
        const bool ge = p >= hi || key[p] >= tgt;
        const unsigned long long m = __ballot(ge);
        const int f = m ? __builtin_ctzll(m) : 64;
        if (stride == 1 || f == 0) {
            lo += (long)f * stride;  // stride 1: f <= hi - lo, as pivot hi - lo is >= tgt
            break;
        }
        const long nlo = lo + (long)(f - 1) * stride + 1;
        if (f < 64 && lo + (long)f * stride < hi) hi = lo + (long)f * stride;
        lo = nlo;
    }
    if (lane == 0) ls[r] = (int)lo;
}

template <class K>
static void leafstart_search_launch(const K* key, int n, const K* key2, int n2, const K* leafkey,
                                    int nleaves, int* ls, int* ls2, hipStream_t s)
{
    const dim3 grid((nleaves + 1 + 3) / 4, key2 ? 2 : 1);
    k_leafstart_search<K><<<grid, 256, 0, s>>>(key, n, key2, n2, leafkey, nleaves, ls, ls2);
}
void launch_leafstart_search(const u64* key, int n, const u64* key2, int n2, const u64* leafkey,
                             int nleaves, int* ls, int* ls2, hipStream_t s)
{
    leafstart_search_launch(key, n, key2, n2, leafkey, nleaves, ls, ls2, s);
}
void launch_leafstart_search(const unsigned* key, int n, const unsigned* key2, int n2,
                             const unsigned* leafkey, int nleaves, int* ls, int* ls2,
                             hipStream_t s)
{
    leafstart_search_launch(key, n, key2, n2, leafkey, nleaves, ls, ls2, s);
}

__global__ __launch_bounds__(256) void k_iota(unsigned* __restrict__ v, int n)
{
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = (unsigned)i;
}

void launch_iota(unsigned* v, int n, hipStream_t s)
{
    if (n <= 0) return;
    k_iota<<<grid_for(n, 256), 256, 0, s>>>(v, n);
}

__global__ __launch_bounds__(256) void k_box_ranges(const int* __restrict__ lo,
                                                    const int* __restrict__ hi, int nb,
                                                    const int* __restrict__ ls,
                                                    int2* __restrict__ se,
                                                    const int* __restrict__ ls2,
                                                    int2* __restrict__ se2)
{
    int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= nb) return;
    const int l = lo[b], h = hi[b];
    se[b] = make_int2(ls[l], ls[h]);
    if (se2) se2[b] = make_int2(ls2[l], ls2[h]);
}

void launch_box_ranges(const int* rank_lo, const int* rank_hi, int nb, const int* ls, int2* se,
                       const int* ls2, int2* se2, hipStream_t s)
{
    k_box_ranges<<<grid_for(nb, 256), 256, 0, s>>>(rank_lo, rank_hi, nb, ls, se, ls2, se2);
}

// V2: the points are 16-byte aligned and move as one 16-byte access each
template <bool V2>
__global__ __launch_bounds__(256) void k_gather(const double* __restrict__ xy,
                                                const double* __restrict__ q, int nd,
                                                const unsigned* __restrict__ perm, int n,
                                                double* __restrict__ xys, double* __restrict__ qs)
{
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    long j = perm[i];
    if (V2) {
        ((double2*)xys)[i] = ((const double2*)xy)[j];
    } else {
        xys[2 * (long)i] = xy[2 * j];
        xys[2 * (long)i + 1] = xy[2 * j + 1];
    }
    if (q) {
        if (nd == 1)
            qs[i] = q[j];
        else
            for (int d = 0; d < nd; ++d) qs[(long)nd * i + d] = q[(long)nd * j + d];
    }
}

// dipole weights in sorted order: w(c, ind, i) = rnormal(c, perm i) dipstr(ind, perm i)
// at ws[2 (nd i + ind) + c], the products the reference forms (g2drouts.f:239-240)
__global__ __launch_bounds__(256) void k_gather_dip(const double* __restrict__ rn,
                                                    const double* __restrict__ ds, int nd,
                                                    const unsigned* __restrict__ perm, int n,
                                                    double* __restrict__ ws)
{
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    long j = perm[i];
    const double n1 = rn[2 * j], n2 = rn[2 * j + 1];
    for (int d = 0; d < nd; ++d) {
        const double dd = ds[(long)nd * j + d];
        ws[2 * ((long)nd * i + d)] = n1 * dd;
        ws[2 * ((long)nd * i + d) + 1] = n2 * dd;
    }
}

void launch_gather_dip(const double* rn, const double* ds, int nd, const unsigned* perm, int n,
                       double* ws, hipStream_t s)
{
    if (n <= 0) return;
    k_gather_dip<<<grid_for(n, 256), 256, 0, s>>>(rn, ds, nd, perm, n, ws);
}

void launch_gather(const double* xy, const double* q, int nd, const unsigned* perm, int n,
                   double* xy_s, double* q_s, hipStream_t s)
{
    if (n <= 0) return;
    if (((uintptr_t)xy & 15) == 0 && ((uintptr_t)xy_s & 15) == 0)
        klaunch(KID_GATHER, k_gather<true>, grid_for(n, 256), 256, 0, s, xy, q, nd, perm, n, xy_s, q_s);
    else
        klaunch(KID_GATHER, k_gather<false>, grid_for(n, 256), 256, 0, s, xy, q, nd, perm, n, xy_s, q_s);
}

__global__ __launch_bounds__(256) void k_zero(double* __restrict__ p, size_t n)
{
    for (size_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0;
}

// up to three buffers cleared by one launch (null or empty ones skipped)
__global__ __launch_bounds__(256) void k_zero3(double* __restrict__ a, size_t na,
                                               double* __restrict__ b, size_t nb,
                                               double* __restrict__ c, size_t nc)
{
    const size_t n = na > nb ? (na > nc ? na : nc) : (nb > nc ? nb : nc);
    for (size_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        if (i < na) a[i] = 0;
        if (i < nb) b[i] = 0;
        if (i < nc) c[i] = 0;
    }
}

void launch_zero3(double* a, size_t na, double* b, size_t nb, double* c, size_t nc, hipStream_t s)
{
    if (!a) na = 0;
    if (!b) nb = 0;
    if (!c) nc = 0;
    const size_t n = std::max(na, std::max(nb, nc));
    if (n == 0) return;
    const size_t g = std::min<size_t>((n + 255) / 256, 1 << 20);
    k_zero3<<<(unsigned)g, 256, 0, s>>>(a, na, b, nb, c, nc);
}

void launch_zero(double* p, size_t n, hipStream_t s)
{
    if (n == 0) return;
    size_t g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    k_zero<<<(unsigned)g, 256, 0, s>>>(p, n);
}

// v[perm[i]*w + c] = v_s[i*w + c]  (dreorderi / dreorderiv, fmmcommon2d.f:29-90)
// v[perm[i]] = vs[i] (+ vs2[i]): sorted values back to input order, with
// a second sorted array added when vs2 is given; vs == nullptr reads zeros
__global__ __launch_bounds__(256) void k_scatter(const double* __restrict__ vs,
                                                 const double* __restrict__ vs2,
                                                 const unsigned* __restrict__ perm, int n, int w,
                                                 double* __restrict__ v)
{
    long t = blockIdx.x * 256L + threadIdx.x;
    if (t >= (long)n * w) return;
    long i = t / w;
    int c = (int)(t - i * w);
    double x = vs ? vs[t] : 0.0;
    if (vs2) x = vs2[t] + x;
    v[(long)perm[i] * w + c] = x;
}

// the same for the points of the listed boxes, with vs == nullptr: one
// block per box (the leaves that no evaluation covers)
__global__ __launch_bounds__(256) void k_scatter_boxes(const int* __restrict__ boxes,
                                                       const int2* __restrict__ se,
                                                       const double* __restrict__ nf,
                                                       const unsigned* __restrict__ perm, int w,
                                                       double* __restrict__ v)
{
    const int2 r = se[boxes[blockIdx.x]];
    const long t0 = (long)r.x * w, t1 = (long)r.y * w;
    for (long t = t0 + threadIdx.x; t < t1; t += 256) {
        const long i = t / w;
        const int c = (int)(t - i * w);
        v[(long)perm[i] * w + c] = nf[t] + 0.0;
    }
}

void launch_scatter_boxes(const int* boxes, int nbox, const int2* se, const double* nf,
                          const unsigned* perm, int w, double* v, hipStream_t s)
{
    if (nbox <= 0) return;
    klaunch(KID_SCATTER, k_scatter_boxes, nbox, 256, 0, s, boxes, se, nf, perm, w, v);
}

void launch_scatter(const double* v_s, const unsigned* perm, int n, int w, double* v,
                    hipStream_t s, const double* v_s2)
{
    if (n <= 0) return;
    klaunch(KID_SCATTER, k_scatter, grid_for((long)n * w, 256), 256, 0, s, v_s, v_s2, perm, n, w,
            v);
}

// ---------------------------------------------------------------------------
// Step 1, leaf formation (g2dformhc_vec g2drouts.f:100-168 followed by
// g2dh2pw_real_vec :1242-1329).  One block per leaf.
//   H(j,k)  = sum_i (q_i yh_i^k/sqrt(k!)) xh_i^j/sqrt(j!)      [LDS tiles]
//   T(a,k)  = sum_j h2x(j,a) H(j,k)          a over the positive x nodes
//   PW(a,b) = sum_k h2x(k,b) T(a,k)          b over all y nodes
// with h2x(j,a) = ws_a (-i ts_a)^j / sqrt(j!).
// ---------------------------------------------------------------------------
// With dipoles (w = d n per source and density) two more tables give
//   H(j,k) = sum_i X_j (q Y_k + w2 Y'_k) + X'_j w1 Y_k,
//   X'_j = dsq j xh^(j-1)/sqrt(j!), the source derivative of X_j
// (g2dformhd_vec / g2dformhcd_vec, g2drouts.f:173-330).
template <int MAXE, bool DIP>
__global__ __launch_bounds__(256) void k_form(FormArgs a, int chunk)
{
    extern __shared__ double sm[];
    int bi = blockIdx.x;
    int box = a.boxes[bi];
    int2 se = a.src_se[box];
    int n = se.y - se.x;
    if (n <= 0) return;
    const int m = a.p + 1, mm = m * m, t = threadIdx.x;
    const double cx = a.cen[2 * box], cy = a.cen[2 * box + 1];
    double* X = sm;              // chunk x m : xh^j / sqrt(j!)
    double* Y = sm + chunk * m;  // chunk x m : q yh^k / sqrt(k!) (+ w2 Y'_k)
    double* Xd = Y + chunk * m;  // chunk x m : X'_j           (dipoles)
    double* Yw = Xd + chunk * m; // chunk x m : w1 yh^k/sqrt(k!) (dipoles)
    double acc[MAXE];
#pragma unroll
    for (int r = 0; r < MAXE; ++r) acc[r] = 0;
    int jj[MAXE], kk[MAXE];
#pragma unroll
    for (int r = 0; r < MAXE; ++r) {
        int e = t + 256 * r;
        jj[r] = e % m;
        kk[r] = e / m;
    }
    for (int s0 = 0; s0 < n; s0 += chunk) {
        int cn = min(chunk, n - s0);
        for (int u = t; u < cn; u += 256) {
            long i = se.x + s0 + u;
            double x = (a.xs[2 * i] - cx) * a.dsq;
            double y = (a.xs[2 * i + 1] - cy) * a.dsq;
            const double q = a.qs[(long)a.nd * i + a.ind];
            if (!DIP) {
                double px = 1.0, py = q;
                X[u * m] = px;
                Y[u * m] = py;
                for (int j = 1; j < m; ++j) {
                    px = px * x * a.rsq[j];
                    py = py * y * a.rsq[j];
                    X[u * m + j] = px;
                    Y[u * m + j] = py;
                }
            } else {
                const double w1 = a.ws[2 * ((long)a.nd * i + a.ind)];
                const double w2 = a.ws[2 * ((long)a.nd * i + a.ind) + 1];
                double px = 1.0, py = 1.0;  // xh^j/sqrt(j!), yh^j/sqrt(j!)
                X[u * m] = 1.0;
                Y[u * m] = q;
                Xd[u * m] = 0.0;
                Yw[u * m] = w1;
                for (int j = 1; j < m; ++j) {
                    const double f = a.dsq * j * a.rsq[j];
                    Xd[u * m + j] = f * px;
                    const double dy = f * py;
                    px = px * x * a.rsq[j];
                    py = py * y * a.rsq[j];
                    X[u * m + j] = px;
                    Y[u * m + j] = q * py + w2 * dy;
                    Yw[u * m + j] = w1 * py;
                }
            }
        }
        __syncthreads();
        for (int s = 0; s < cn; ++s) {
#pragma unroll
            for (int r = 0; r < MAXE; ++r)
                if (t + 256 * r < mm) {
                    acc[r] += Y[s * m + kk[r]] * X[s * m + jj[r]];
                    if (DIP) acc[r] += Yw[s * m + kk[r]] * Xd[s * m + jj[r]];
                }
        }
        __syncthreads();
    }
    double* H = sm;                      // m x m, j fastest
    const int npw = a.npw, half = npw / 2;
    double* T = sm + mm;                 // complex (half, m): a fastest
#pragma unroll
    for (int r = 0; r < MAXE; ++r)
        if (t + 256 * r < mm) H[t + 256 * r] = acc[r];
    __syncthreads();
    for (int o = t; o < half * m; o += 256) {
        int aa = o % half, k = o / half;
        double re = 0, im = 0;
        for (int j = 0; j < m; ++j) {
            double h = H[j + m * k];
            re += a.h2x[2 * (j * npw + aa)] * h;
            im += a.h2x[2 * (j * npw + aa) + 1] * h;
        }
        T[2 * o] = re;
        T[2 * o + 1] = im;
    }
    __syncthreads();
    const int nexp = half * npw;
    double* out = a.mp + 2 * ((long)a.slot[box] * a.nd + a.ind) * nexp;
    for (int o = t; o < nexp; o += 256) {
        int aa = o % half, b = o / half;
        double re = 0, im = 0;
        for (int k = 0; k < m; ++k) {
            double wr = a.h2x[2 * (k * npw + b)], wi = a.h2x[2 * (k * npw + b) + 1];
            double tr = T[2 * (aa + half * k)], ti = T[2 * (aa + half * k) + 1];
            re += wr * tr - wi * ti;
            im += wr * ti + wi * tr;
        }
        out[2 * o] = re;
        out[2 * o + 1] = im;
    }
}

// ---------------------------------------------------------------------------
// Work decomposition.  Leaves hold from one to thousands of points, so the
// heavy steps run over fixed-size segments of each box's point range rather
// than over boxes: tstart (an exclusive scan of the per-box segment counts)
// numbers the segments, a persistent grid strides over them, and a block
// finds its box in segbox (filled here) or by binary search in tstart.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_seg_scan(SegLists L)
{
    // rounds of 1024 boxes: wavefront shuffles, then the 16 wave totals
    __shared__ int wsum[16];
    const SegList d = L.l[blockIdx.x];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int carry = 0;
    for (int base = 0; base <= d.n; base += 1024) {
        const int i = base + t;
        int c = 0;
        if (i < d.n) {
            int2 r = d.se[d.boxes[i]];
            int len = r.y - r.x;
            c = len > 0 ? (len + d.seg - 1) / d.seg : 0;
        }
        int x = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        int pre = 0, tot = 0;
#pragma unroll
        for (int ww = 0; ww < 16; ++ww) {
            const int v = wsum[ww];
            pre += ww < w ? v : 0;
            tot += v;
        }
        const int excl = carry + pre + x - c;
        if (i <= d.n) d.tstart[i] = excl;
        if (d.segbox)
            for (int j = 0; j < c; ++j) d.segbox[excl + j] = i;
        carry += tot;
        __syncthreads();  // wsum is rewritten by the next round
    }
}

void launch_seg_scan(const SegLists& L, hipStream_t s)
{
    if (L.count <= 0) return;
    k_seg_scan<<<L.count, 1024, 0, s>>>(L);
}

struct SegRange {
    int li, p0, p1;
};

// One point's evaluated values, for sorted point i at input position j =
// a.perm[i]: pot, the gradient's two and the Hessian's three entries, each
// added to the near field's sorted value as nf + (0 + v) and written to the
// caller's arrays at j (the sorted order's outputs are never formed).
// nfp is the near field's potential of the point (staged by the caller).
template <int PGH>
__device__ __forceinline__ void eval_out(const EvalArgs& a, long i, long j, double nfp, double v,
                                         double g0, double g1, double h0, double h1, double h2)
{
    const int nd = a.nd, c = a.ind;
    a.pot[j * nd + c] = nfp + (0.0 + v);
    if (PGH >= 2) {
        const long gi = i * 2 * nd + c, go = j * 2 * nd + c;
        a.grad[go] = a.ngrad[gi] + (0.0 + g0);
        a.grad[go + nd] = a.ngrad[gi + nd] + (0.0 + g1);
    }
    if (PGH >= 3) {
        const long hi = i * 3 * nd + c, ho = j * 3 * nd + c;
        a.hess[ho] = a.nhess[hi] + (0.0 + h0);
        a.hess[ho + nd] = a.nhess[hi + nd] + (0.0 + h1);
        a.hess[ho + 2 * nd] = a.nhess[hi + 2 * nd] + (0.0 + h2);
    }
}
// the same with the permutation and the near field read here; with no
// permutation (a.perm == nullptr: one box of points in input order, the
// large-delta branch) the values are added into the outputs at i
template <int PGH>
__device__ __forceinline__ void eval_out(const EvalArgs& a, long i, double v, double g0, double g1,
                                         double h0, double h1, double h2)
{
    if (a.perm) {
        eval_out<PGH>(a, i, (long)a.perm[i], a.npot[i * a.nd + a.ind], v, g0, g1, h0, h1, h2);
        return;
    }
    const int nd = a.nd, c = a.ind;
    a.pot[i * nd + c] += v;
    if (PGH >= 2) {
        const long g = i * 2 * nd + c;
        a.grad[g] += g0;
        a.grad[g + nd] += g1;
    }
    if (PGH >= 3) {
        const long h = i * 3 * nd + c;
        a.hess[h] += h0;
        a.hess[h + nd] += h1;
        a.hess[h + 2 * nd] += h2;
    }
}

// segment g of a list: its box index li and point range [p0, p1)
__device__ __forceinline__ SegRange seg_at(const int* tstart, int nbox, const int* boxes,
                                           const int2* se, int seg, int g,
                                           const int* segbox = nullptr)
{
    SegRange r;
    if (!tstart) {
        int2 x = se[boxes[g]];
        r.li = g;
        r.p0 = x.x;
        r.p1 = x.y;
        return r;
    }
    int lo = 0;
    if (segbox) {
        lo = segbox[g];
    } else {
        int hi = nbox - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (tstart[mid] <= g) lo = mid; else hi = mid - 1;
        }
    }
    int2 x = se[boxes[lo]];
    r.li = lo;
    r.p0 = x.x + (g - tstart[lo]) * seg;
    r.p1 = min(x.y, r.p0 + seg);
    return r;
}

__device__ __forceinline__ int seg_count(const int* tstart, int nbox)
{
    return tstart ? tstart[nbox] : nbox;
}

// ---------------------------------------------------------------------------
// FP64 matrix cores.  v_mfma_f64_16x16x4f64: D(16x16) += A(16x4) B(4x16) per
// wavefront; lane l supplies A[l%16][l/16] and B[l/16][l%16] and holds
// D[l/16 + 4r][l%16] in element r of the result (lane layout measured on
// gfx950).
// ---------------------------------------------------------------------------
typedef double d4_t __attribute__((ext_vector_type(4)));

// Kernels that take more than 64 KB of dynamic LDS (gfx950 has 160 KB per
// workgroup) declare it.  The attribute is a limit: it is raised only when a
// launch needs more than the kernel's largest earlier size, since the call
// costs host time on every launch path.
template <class F>
static void allow_lds(F* kernel, size_t bytes)
{
    if (bytes < 65536) return;
    static std::mutex mu;
    static std::vector<std::pair<const void*, size_t>> set;  // per kernel, the limit set
    const void* k = (const void*)kernel;
    std::lock_guard<std::mutex> lk(mu);
    auto it = std::find_if(set.begin(), set.end(), [k](const auto& e) { return e.first == k; });
    if (it != set.end() && it->second >= bytes) return;
    HIPCHECK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    if (it != set.end())
        it->second = bytes;
    else
        set.emplace_back(k, bytes);
}

__device__ __forceinline__ d4_t mfma16x16x4(double a, double b, d4_t c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// LDS writes of one wavefront made visible to its other lanes
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Step 1 on matrix cores, in two kernels.
// (a) per segment of a leaf's sources, the raw moments M = X diag(q) Y^T with
//     X(j,s) = xh_s^j and Y(k,s) = yh_s^k staged in LDS; each wave multiplies
//     4 sources per MFMA into TM x TM tiles of 16x16; the four waves' tiles are
//     summed in a fixed order and written as the segment's partial moments.
// (b) per leaf, the partial moments of its segments are summed in segment
//     order and converted, PW = h2xr^T M h2xr with h2xr(j,a) =
//     ws_a (-i ts_a)^j / j! (the 1/sqrt(j!) Hermite normalisation of
//     g2dformhc_vec twice), as g2dh2pw_real_vec.
template <int TM>
__global__ __launch_bounds__(256) void k_form_moments(FormArgs a)
{
    extern __shared__ double sm[];
    constexpr int MP = 16 * TM, RS = MP + 1;  // odd row stride: no LDS bank conflicts
    const int m = a.p + 1, t = threadIdx.x, lane = t & 63, w = t >> 6;
    double* X = sm;               // seg x RS, zero padded
    double* Y = sm + a.seg * RS;  // seg x RS, q folded in
    const int total = seg_count(a.tstart, a.nbox);
    for (int g = blockIdx.x; g < total; g += gridDim.x) {
        const SegRange r = seg_at(a.tstart, a.nbox, a.boxes, a.src_se, a.seg, g);
        const int box = a.boxes[r.li], cn = r.p1 - r.p0, cpad = (cn + 15) & ~15;
        const double cx = a.cen[2 * box], cy = a.cen[2 * box + 1];
        __syncthreads();
        for (int u = t; u < cpad; u += 256) {
            double* xr = X + u * RS;
            double* yr = Y + u * RS;
            int j0 = 0;
            if (u < cn) {
                long i = r.p0 + u;
                double x = (a.xs[2 * i] - cx) * a.dsq, y = (a.xs[2 * i + 1] - cy) * a.dsq;
                double px = 1.0, py = a.qs[(long)a.nd * i + a.ind];
                xr[0] = px;
                yr[0] = py;
                for (int j = 1; j < m; ++j) {
                    px *= x;
                    py *= y;
                    xr[j] = px;
                    yr[j] = py;
                }
                j0 = m;
            }
            for (int j = j0; j < MP; ++j) {
                xr[j] = 0;
                yr[j] = 0;
            }
        }
        __syncthreads();
        d4_t acc[TM][TM];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TM; ++j) acc[i][j] = (d4_t){0, 0, 0, 0};
        const int s_l = lane >> 4, i_l = lane & 15;
        for (int gg = w; gg < cpad / 4; gg += 4) {
            const int src = 4 * gg + s_l;
            double av[TM], bv[TM];
#pragma unroll
            for (int i = 0; i < TM; ++i) {
                av[i] = X[src * RS + i_l + 16 * i];
                bv[i] = Y[src * RS + i_l + 16 * i];
            }
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TM; ++j) acc[i][j] = mfma16x16x4(av[i], bv[j], acc[i][j]);
        }
        __syncthreads();
        double* Hf = sm;  // MP x MP, row = x order, column = y order (reuses X)
        for (int ww = 0; ww < 4; ++ww) {
            if (w == ww) {
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TM; ++j)
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            int row = (lane >> 4) + 4 * q + 16 * i, col = (lane & 15) + 16 * j;
                            double v = acc[i][j][q];
                            Hf[row * MP + col] = ww == 0 ? v : Hf[row * MP + col] + v;
                        }
            }
            __syncthreads();
        }
        double* out = a.partial + (long)g * a.pstride;  // m x m, x order fastest
        for (int o = t; o < m * m; o += 256) out[o] = Hf[(o / m) * MP + o % m];
    }
}

// Step 1a, one wavefront per segment (at most WSEG points), in chunks of 32
// sources: lane u < 32 writes the powers xh^e of source u of the chunk to
// the wave's slice of LDS and lane 32 + u the powers q yh^e, e < 16 TM,
// while the next chunk's points are already being loaded.  Every MFMA then
// takes 4 sources, lane l reading xh^e and q yh^e for e = l%16 (+16 per
// tile) of source l/16, so D(j,k) = sum_s xh_s^j q_s yh_s^k accumulates in
// registers.  Lane l holds x orders l/16 + 4r and y orders l%16 of each
// tile and writes the segment's partial moments, y order fastest.
constexpr int WSEG = 256;

//
// With dipoles (w = d n per source) the raw moments are
//   M(j,k) = sum_s xh^j (q yh^k + w2 dsq k yh^(k-1)) + (dsq j xh^(j-1)) w1 yh^k,
// the source derivatives of g2dformhd_vec / g2dformhcd_vec
// (g2drouts.f:173-330): two more tables and a second MFMA per step.  The
// block then holds blockDim.x / 64 waves (128 threads for TM >= 3, to fit
// the tables in LDS).
template <int TM, bool DIP>
__global__ __launch_bounds__(256) void k_form_wave(FormArgs a)
{
    extern __shared__ double sm[];
    constexpr int MP = 16 * TM, RS = MP + 1, CH = 32;  // odd row stride: no bank conflicts
    constexpr int NT = DIP ? 4 : 2;                     // tables per wave
    const int m = a.p + 1, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const int s_l = lane >> 4, e_l = lane & 15, u = lane & 31, side = lane >> 5;
    double* Px = sm + (size_t)w * NT * CH * RS;  // CH x RS: xh^e
    double* Py = Px + CH * RS;                    // CH x RS: q yh^e (+ w2 dsq e yh^(e-1))
    double* Pxd = Py + CH * RS;                   // CH x RS: dsq e xh^(e-1)   (dipoles)
    double* Pyw = Pxd + CH * RS;                  // CH x RS: w1 yh^e          (dipoles)
    double* mine = (side ? Py : Px) + u * RS;
    double* mine2 = (side ? Pyw : Pxd) + u * RS;
    const int total = seg_count(a.tstart, a.nbox);
    for (int g = blockIdx.x * nw + w; g < total; g += gridDim.x * nw) {
        const SegRange r = seg_at(a.tstart, a.nbox, a.boxes, a.src_se, a.seg, g, a.segbox);
        const int box = a.boxes[r.li], n = r.p1 - r.p0;
        const double c = a.cen[2 * box + side], dsq = a.dsq;
        d4_t acc[TM][TM];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TM; ++j) acc[i][j] = (d4_t){0, 0, 0, 0};
        // this lane's coordinate (x for side 0, y for side 1) and factors of
        // the first chunk: table 1 gets f v^e + g e v^(e-1), table 2
        // f2 v^e + g2 e v^(e-1)
        double v = 0, f = 0, g1 = 0, f2 = 0, g2 = 0;
        auto load = [&](int k, double& v_, double& f_, double& g_, double& f2_, double& g2_) {
            v_ = f_ = g_ = f2_ = g2_ = 0;
            if (k < n) {
                const long i = r.p0 + k;
                v_ = (a.xs[2 * i + side] - c) * dsq;
                f_ = side ? a.qs[(long)a.nd * i + a.ind] : 1.0;
                if (DIP) {
                    const double* wv = a.ws + 2 * ((long)a.nd * i + a.ind);
                    g_ = side ? dsq * wv[1] : 0.0;
                    f2_ = side ? wv[0] : 0.0;
                    g2_ = side ? 0.0 : dsq;
                }
            }
        };
        if (!DIP && a.seg <= 8 * CH) {
            // the whole segment's coordinates and charges requested at once
            // (one memory latency per segment instead of one per chunk)
            double V[8], F[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = u + CH * j;
                V[j] = F[j] = 0;
                if (k < n) {
                    const long i = r.p0 + k;
                    V[j] = a.xs[2 * i + side];
                    F[j] = side ? a.qs[(long)a.nd * i + a.ind] : 1.0;
                }
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int c0 = CH * j;
                if (c0 >= n) break;
                wave_lds_sync();
                const double vv = (V[j] - c) * dsq;
                double pw = F[j];
#pragma unroll
                for (int e = 0; e < MP; ++e) {
                    mine[e] = pw;
                    pw *= vv;
                }
                wave_lds_sync();
                const int nc = min(CH, n - c0);
                for (int gg = 0; 4 * gg < nc; ++gg) {
                    const int src = 4 * gg + s_l;
                    double av[TM], bv[TM];
#pragma unroll
                    for (int t = 0; t < TM; ++t) {
                        av[t] = Px[src * RS + e_l + 16 * t];
                        bv[t] = Py[src * RS + e_l + 16 * t];
                    }
#pragma unroll
                    for (int ii = 0; ii < TM; ++ii)
#pragma unroll
                        for (int jj = 0; jj < TM; ++jj)
                            acc[ii][jj] = mfma16x16x4(av[ii], bv[jj], acc[ii][jj]);
                }
            }
        } else {
        load(u, v, f, g1, f2, g2);
        for (int c0 = 0; c0 < n; c0 += CH) {
            double vn, fn, gn, f2n, g2n;  // next chunk, loaded ahead
            load(c0 + CH + u, vn, fn, gn, f2n, g2n);
            wave_lds_sync();  // the previous chunk's reads are done
            if (!DIP) {
                double pw = f;
#pragma unroll
                for (int e = 0; e < MP; ++e) {
                    mine[e] = pw;
                    pw *= v;
                }
            } else {
                double pw = 1.0, prev = 0.0;  // v^e, v^(e-1)
#pragma unroll
                for (int e = 0; e < MP; ++e) {
                    mine[e] = f * pw + g1 * e * prev;
                    mine2[e] = f2 * pw + g2 * e * prev;
                    prev = pw;
                    pw *= v;
                }
            }
            wave_lds_sync();
            const int nc = min(CH, n - c0);
            for (int gg = 0; 4 * gg < nc; ++gg) {
                const int src = 4 * gg + s_l;
                double av[TM], bv[TM];
#pragma unroll
                for (int t = 0; t < TM; ++t) {
                    av[t] = Px[src * RS + e_l + 16 * t];
                    bv[t] = Py[src * RS + e_l + 16 * t];
                }
#pragma unroll
                for (int ii = 0; ii < TM; ++ii)
#pragma unroll
                    for (int jj = 0; jj < TM; ++jj) acc[ii][jj] = mfma16x16x4(av[ii], bv[jj], acc[ii][jj]);
                if (DIP) {
#pragma unroll
                    for (int t = 0; t < TM; ++t) {
                        av[t] = Pxd[src * RS + e_l + 16 * t];
                        bv[t] = Pyw[src * RS + e_l + 16 * t];
                    }
#pragma unroll
                    for (int ii = 0; ii < TM; ++ii)
#pragma unroll
                        for (int jj = 0; jj < TM; ++jj)
                            acc[ii][jj] = mfma16x16x4(av[ii], bv[jj], acc[ii][jj]);
                }
            }
            v = vn;
            f = fn;
            g1 = gn;
            f2 = f2n;
            g2 = g2n;
        }
        }
        double* out = a.partial + (long)g * a.pstride;
#pragma unroll
        for (int ii = 0; ii < TM; ++ii)
#pragma unroll
            for (int jj = 0; jj < TM; ++jj)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int xo = s_l + 4 * q + 16 * ii, yo = e_l + 16 * jj;
                    if (xo < m && yo < m) out[yo + m * xo] = acc[ii][jj][q];
                }
    }
}

// ---------------------------------------------------------------------------
// FP32 matrix cores.  v_mfma_f32_16x16x4f32: the operands take the same lane
// layout as the FP64 instruction (lane l supplies A[l%16][l/16] and
// B[l/16][l%16]); lane l holds D[4 (l/16) + r][l%16] in element r of the
// result (rows of a lane contiguous, unlike the FP64 form).
// ---------------------------------------------------------------------------
typedef float f4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4_t mfma16x16x4f32(float a, float b, f4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Step 1a with FP32 operands (the mixed-precision route): k_form_wave with
// the offsets xh, yh formed in FP64 from the FP64 box centre, their powers
// (and charge and dipole factors) computed in FP64 and rounded once to FP32
// in the LDS tables, FP32 accumulation of one segment (at most WSEG
// sources) on the matrix cores, and the segment's partial moments stored in
// FP64 (summed across segments in FP64 by k_form_h2pw).
template <int TM, bool DIP>
__global__ __launch_bounds__(256) void k_form_wave_f32(FormArgs a)
{
    extern __shared__ float smf[];
    constexpr int MP = 16 * TM, RS = MP + 1, CH = 32;
    constexpr int NT = DIP ? 4 : 2;
    const int m = a.p + 1, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const int s_l = lane >> 4, e_l = lane & 15, u = lane & 31, side = lane >> 5;
    float* Px = smf + (size_t)w * NT * CH * RS;
    float* Py = Px + CH * RS;
    float* Pxd = Py + CH * RS;
    float* Pyw = Pxd + CH * RS;
    float* mine = (side ? Py : Px) + u * RS;
    float* mine2 = (side ? Pyw : Pxd) + u * RS;
    const int total = seg_count(a.tstart, a.nbox);
    for (int g = blockIdx.x * nw + w; g < total; g += gridDim.x * nw) {
        const SegRange r = seg_at(a.tstart, a.nbox, a.boxes, a.src_se, a.seg, g, a.segbox);
        const int box = a.boxes[r.li], n = r.p1 - r.p0;
        const double c = a.cen[2 * box + side], dsq = a.dsq;
        f4_t acc[TM][TM];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TM; ++j) acc[i][j] = (f4_t){0, 0, 0, 0};
        double v = 0, f = 0, g1 = 0, f2 = 0, g2 = 0;
        auto load = [&](int k, double& v_, double& f_, double& g_, double& f2_, double& g2_) {
            v_ = f_ = g_ = f2_ = g2_ = 0;
            if (k < n) {
                const long i = r.p0 + k;
                v_ = (a.xs[2 * i + side] - c) * dsq;
                f_ = side ? a.qs[(long)a.nd * i + a.ind] : 1.0;
                if (DIP) {
                    const double* wv = a.ws + 2 * ((long)a.nd * i + a.ind);
                    g_ = side ? dsq * wv[1] : 0.0;
                    f2_ = side ? wv[0] : 0.0;
                    g2_ = side ? 0.0 : dsq;
                }
            }
        };
        if (!DIP && a.seg <= 8 * CH) {
            // the whole segment's coordinates and charges requested at once
            double V[8], F[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = u + CH * j;
                V[j] = F[j] = 0;
                if (k < n) {
                    const long i = r.p0 + k;
                    V[j] = a.xs[2 * i + side];
                    F[j] = side ? a.qs[(long)a.nd * i + a.ind] : 1.0;
                }
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int c0 = CH * j;
                if (c0 >= n) break;
                wave_lds_sync();
                const double vv = (V[j] - c) * dsq;
                double pw = F[j];
#pragma unroll
                for (int e = 0; e < MP; ++e) {
                    mine[e] = (float)pw;
                    pw *= vv;
                }
                wave_lds_sync();
                const int nc = min(CH, n - c0);
                for (int gg = 0; 4 * gg < nc; ++gg) {
                    const int src = 4 * gg + s_l;
                    float av[TM], bv[TM];
#pragma unroll
                    for (int t = 0; t < TM; ++t) {
                        av[t] = Px[src * RS + e_l + 16 * t];
                        bv[t] = Py[src * RS + e_l + 16 * t];
                    }
#pragma unroll
                    for (int ii = 0; ii < TM; ++ii)
#pragma unroll
                        for (int jj = 0; jj < TM; ++jj)
                            acc[ii][jj] = mfma16x16x4f32(av[ii], bv[jj], acc[ii][jj]);
                }
            }
        } else {
        load(u, v, f, g1, f2, g2);
        for (int c0 = 0; c0 < n; c0 += CH) {
            double vn, fn, gn, f2n, g2n;
            load(c0 + CH + u, vn, fn, gn, f2n, g2n);
            wave_lds_sync();
            if (!DIP) {
                double pw = f;
#pragma unroll
                for (int e = 0; e < MP; ++e) {
                    mine[e] = (float)pw;
                    pw *= v;
                }
            } else {
                double pw = 1.0, prev = 0.0;
#pragma unroll
                for (int e = 0; e < MP; ++e) {
                    mine[e] = (float)(f * pw + g1 * e * prev);
                    mine2[e] = (float)(f2 * pw + g2 * e * prev);
                    prev = pw;
                    pw *= v;
                }
            }
            wave_lds_sync();
            const int nc = min(CH, n - c0);
            for (int gg = 0; 4 * gg < nc; ++gg) {
                const int src = 4 * gg + s_l;
                float av[TM], bv[TM];
#pragma unroll
                for (int t = 0; t < TM; ++t) {
                    av[t] = Px[src * RS + e_l + 16 * t];
                    bv[t] = Py[src * RS + e_l + 16 * t];
                }
#pragma unroll
                for (int ii = 0; ii < TM; ++ii)
#pragma unroll
                    for (int jj = 0; jj < TM; ++jj)
                        acc[ii][jj] = mfma16x16x4f32(av[ii], bv[jj], acc[ii][jj]);
                if (DIP) {
#pragma unroll
                    for (int t = 0; t < TM; ++t) {
                        av[t] = Pxd[src * RS + e_l + 16 * t];
                        bv[t] = Pyw[src * RS + e_l + 16 * t];
                    }
#pragma unroll
                    for (int ii = 0; ii < TM; ++ii)
#pragma unroll
                        for (int jj = 0; jj < TM; ++jj)
                            acc[ii][jj] = mfma16x16x4f32(av[ii], bv[jj], acc[ii][jj]);
                }
            }
            v = vn;
            f = fn;
            g1 = gn;
            f2 = f2n;
            g2 = g2n;
        }
        }
        double* out = a.partial + (long)g * a.pstride;
#pragma unroll
        for (int ii = 0; ii < TM; ++ii)
#pragma unroll
            for (int jj = 0; jj < TM; ++jj)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int xo = 4 * s_l + q + 16 * ii, yo = e_l + 16 * jj;
                    if (xo < m && yo < m) out[yo + m * xo] = (double)acc[ii][jj][q];
                }
    }
}

// E: element type of the plane-wave expansions written (double, or float
// on the mixed-precision route); the conversion itself runs in FP64.
template <class E>
__global__ __launch_bounds__(256) void k_form_h2pw(FormArgs a)
{
    extern __shared__ double sm[];
    const int li = blockIdx.x;
    const int box = a.boxes[li];
    const int g0 = a.tstart[li], g1 = a.tstart[li + 1];
    if (g1 <= g0) return;  // no sources: the expansion stays zero
    const int m = a.p + 1, t = threadIdx.x, npw = a.npw, half = npw / 2, nexp = half * npw;
    double* H = sm;                // m x m, x order fastest
    double* hx = H + m * m;        // complex (m, npw)
    double* T = hx + 2 * m * npw;  // complex (half, m)
    for (int o = t; o < m * m; o += 256) {  // partials y order fastest, H x order fastest
        double v = 0;
        for (int g = g0; g < g1; ++g) v += a.partial[(long)g * a.pstride + o];
        H[o / m + m * (o % m)] = v;
    }
    for (int o = t; o < 2 * m * npw; o += 256) hx[o] = a.h2xr[o];
    __syncthreads();
    for (int o = t; o < half * m; o += 256) {
        int aa = o % half, k = o / half;
        double re = 0, im = 0;
        for (int j = 0; j < m; ++j) {
            double h = H[j + m * k];
            re += hx[2 * (j * npw + aa)] * h;
            im += hx[2 * (j * npw + aa) + 1] * h;
        }
        T[2 * o] = re;
        T[2 * o + 1] = im;
    }
    __syncthreads();
    E* out = (E*)a.mp + 2 * ((long)a.slot[box] * a.nd + a.ind) * nexp;
    for (int o = t; o < nexp; o += 256) {
        int aa = o % half, b = o / half;
        double re = 0, im = 0;
        for (int k = 0; k < m; ++k) {
            double wr = hx[2 * (k * npw + b)], wi = hx[2 * (k * npw + b) + 1];
            double tr = T[2 * (aa + half * k)], ti = T[2 * (aa + half * k) + 1];
            re += wr * tr - wi * ti;
            im += wr * ti + wi * tr;
        }
        out[2 * o] = (E)re;
        out[2 * o + 1] = (E)im;
    }
}

// Step 1b on matrix cores, one wavefront per leaf (four leaves per block):
// the segment partials of the leaf summed in segment order into
// H(j,k) = sum_s xh_s^j q_s yh_s^k, then PW = h2xr^T H h2xr as
// g2dh2pw_real_vec (g2drouts.f:1242-1329) computes it.  With
// h2xr(j,a) = i^j r(j,a), r real (ws_a (-ts_a)^j / j!), the complex
// product splits by (j + k) mod 4:
//   PW(a,b) = sum_{j,k} i^(j+k) r(j,a) H(j,k) r(k,b)
//           = [R^T G_re R](a,b) + i [R^T G_im R](a,b),
//   G_re = H on j+k = 0 mod 4, -H on 2;  G_im = H on 1, -H on 3.
// Y = G R (rows j, columns b) runs first; its accumulators are the B
// operands of R^T Y as they stand (lane l holds Y(l/16 + 4 r + 16 t, .),
// which K step r + 4 t of the second product reads from lane l).
// TM: 16-row tiles of the order (m <= 16 TM); NB: 16-column tiles of b
// (npw <= 16 NB); HT: 16-row tiles of a (npw/2 <= 16 HT).
template <int TM, int NB, int HT, class E>
__global__ __launch_bounds__(256) void k_form_h2pw_wave(FormArgs a)
{
    constexpr int MP = 16 * TM;
    extern __shared__ double smh[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = blockIdx.x * 4 + w;
    if (li >= a.nbox) return;
    const int g0 = a.tstart[li], g1 = a.tstart[li + 1];
    if (g1 <= g0) return;  // no sources: the expansion stays zero
    const int box = a.boxes[li];
    const int m = a.p + 1, npw = a.npw, half = npw / 2;
    const int c16 = lane & 15, q4 = lane >> 4;
    double* Gre = smh + (size_t)w * 2 * MP * MP;  // G(j,k) at j + MP k
    double* Gim = Gre + MP * MP;
    // H(j,k): partials are y order fastest (k + m j); segments in order
    for (int o = lane; o < MP * MP; o += 64) {
        const int j = o % MP, k = o / MP;
        double h = 0;
        if (j < m && k < m)
            for (int g = g0; g < g1; ++g) h += a.partial[(long)g * a.pstride + k + m * j];
        const int q = (j + k) & 3;
        Gre[o] = q == 0 ? h : (q == 2 ? -h : 0.0);
        Gim[o] = q == 1 ? h : (q == 3 ? -h : 0.0);
    }
    wave_lds_sync();
    // r(j,a) from the complex table h2xr(j,a) = i^j r(j,a)
    auto rtab = [&](int j, int aa) -> double {
        const double* h = a.h2xr + 2 * ((long)j * npw + aa);
        switch (j & 3) {
        case 0: return h[0];
        case 1: return h[1];
        case 2: return -h[0];
        default: return -h[1];
        }
    };
    // Y(j,b) = sum_k G(j,k) r(k,b), re and im
    d4_t Y[TM][NB][2];
#pragma unroll
    for (int t = 0; t < TM; ++t)
#pragma unroll
        for (int bt = 0; bt < NB; ++bt) Y[t][bt][0] = Y[t][bt][1] = (d4_t){0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < 4 * TM; ++s) {
        const int k = 4 * s + q4;
        double bv[NB];
#pragma unroll
        for (int bt = 0; bt < NB; ++bt) {
            const int b = c16 + 16 * bt;
            bv[bt] = (k < m && b < npw) ? rtab(k, b) : 0.0;
        }
#pragma unroll
        for (int t = 0; t < TM; ++t) {
            const int j = c16 + 16 * t;
            const double ar = Gre[j + MP * k], ai = Gim[j + MP * k];
#pragma unroll
            for (int bt = 0; bt < NB; ++bt) {
                Y[t][bt][0] = mfma16x16x4(ar, bv[bt], Y[t][bt][0]);
                Y[t][bt][1] = mfma16x16x4(ai, bv[bt], Y[t][bt][1]);
            }
        }
    }
    // PW(a,b) = sum_j r(j,a) Y(j,b), a < half
    d4_t P[HT][NB][2];
#pragma unroll
    for (int ht = 0; ht < HT; ++ht)
#pragma unroll
        for (int bt = 0; bt < NB; ++bt) P[ht][bt][0] = P[ht][bt][1] = (d4_t){0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < TM; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int j = 16 * t + 4 * r + q4;  // K index of step r + 4 t in this lane
            double av[HT];
#pragma unroll
            for (int ht = 0; ht < HT; ++ht) {
                const int aa = c16 + 16 * ht;
                av[ht] = (j < m && aa < half) ? rtab(j, aa) : 0.0;
            }
#pragma unroll
            for (int ht = 0; ht < HT; ++ht)
#pragma unroll
                for (int bt = 0; bt < NB; ++bt) {
                    P[ht][bt][0] = mfma16x16x4(av[ht], Y[t][bt][0][r], P[ht][bt][0]);
                    P[ht][bt][1] = mfma16x16x4(av[ht], Y[t][bt][1][r], P[ht][bt][1]);
                }
        }
    E* out = (E*)a.mp + 2 * ((long)a.slot[box] * a.nd + a.ind) * (long)(half * npw);
#pragma unroll
    for (int ht = 0; ht < HT; ++ht)
#pragma unroll
        for (int bt = 0; bt < NB; ++bt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int aa = q4 + 4 * r + 16 * ht, b = c16 + 16 * bt;
                if (aa < half && b < npw) {
                    out[2 * (aa + half * b)] = (E)P[ht][bt][0][r];
                    out[2 * (aa + half * b) + 1] = (E)P[ht][bt][1][r];
                }
            }
}

template <int TM, int NB, int HT>
static void h2pw_wave_launch(const FormArgs& a, hipStream_t s)
{
    const unsigned lds = 4 * 2 * (16 * TM) * (16 * TM) * sizeof(double);
    const unsigned grid = (unsigned)((a.nbox + 3) / 4);
    if (a.prec) {
        allow_lds(k_form_h2pw_wave<TM, NB, HT, float>, lds);
        klaunch(KID_FORM_H2PW, k_form_h2pw_wave<TM, NB, HT, float>, grid, 256, lds, s, a);
    } else {
        allow_lds(k_form_h2pw_wave<TM, NB, HT, double>, lds);
        klaunch(KID_FORM_H2PW, k_form_h2pw_wave<TM, NB, HT, double>, grid, 256, lds, s, a);
    }
}

// the matrix-core Hermite -> PW for orders below 32 and npw <= 64; false
// when it does not apply
static bool launch_h2pw_wave(const FormArgs& a, hipStream_t s)
{
    const int m = a.p + 1, nb = (a.npw + 15) / 16, ht = (a.npw / 2 + 15) / 16;
    if (m > 32 || a.npw > 64 || !a.tstart) return false;
    const int tm = (m + 15) / 16;
#define FGT_H2PW(T_, B_, H_) \
    if (tm == T_ && nb <= B_ && ht <= H_) { h2pw_wave_launch<T_, B_, H_>(a, s); return true; }
    FGT_H2PW(1, 1, 1)
    FGT_H2PW(1, 2, 1)
    FGT_H2PW(1, 4, 2)
    FGT_H2PW(2, 1, 1)
    FGT_H2PW(2, 2, 1)
    FGT_H2PW(2, 4, 2)
#undef FGT_H2PW
    return false;
}

void launch_form(const FormArgs& a, hipStream_t s)
{
    if (a.nbox <= 0) return;
    int m = a.p + 1;
    if (m <= 64 && a.tstart) {
        const int TM = (m + 15) / 16, MP = 16 * TM;
        if (a.segbox && a.seg <= WSEG) {
            // per wave: two (four with dipoles) 32 x (MP + 1) power tables
            const bool dip = a.ws != nullptr;
            const int nwave = dip && TM >= 3 ? 2 : 4;
            size_t lds = (size_t)nwave * (dip ? 4 : 2) * 32 * (MP + 1) *
                         (a.prec ? sizeof(float) : sizeof(double));
            const int grid = (int)std::min<long>(2L * a.grid, (long)a.grid * 4 / nwave);
#define FGT_FORM_WAVE(T_, D_)                                                                  \
    if (a.prec) {                                                                              \
        allow_lds(k_form_wave_f32<T_, D_>, lds);                                               \
        klaunch(KID_FORM_MOMENTS, k_form_wave_f32<T_, D_>, grid, 64 * nwave, lds, s, a);                              \
    } else {                                                                                   \
        allow_lds(k_form_wave<T_, D_>, lds);                                                   \
        klaunch(KID_FORM_MOMENTS, k_form_wave<T_, D_>, grid, 64 * nwave, lds, s, a);                                  \
    }
            if (!dip) {
                switch (TM) {
                case 1: FGT_FORM_WAVE(1, false); break;
                case 2: FGT_FORM_WAVE(2, false); break;
                case 3: FGT_FORM_WAVE(3, false); break;
                default: FGT_FORM_WAVE(4, false); break;
                }
            } else {
                switch (TM) {
                case 1: FGT_FORM_WAVE(1, true); break;
                case 2: FGT_FORM_WAVE(2, true); break;
                case 3: FGT_FORM_WAVE(3, true); break;
                default: FGT_FORM_WAVE(4, true); break;
                }
            }
#undef FGT_FORM_WAVE
        } else {
            size_t lds = std::max((size_t)2 * a.seg * (MP + 1), (size_t)MP * MP) * sizeof(double);
            switch (TM) {
            case 1: allow_lds(k_form_moments<1>, lds); klaunch(KID_FORM_MOMENTS, k_form_moments<1>, a.grid, 256, lds, s, a); break;
            case 2: allow_lds(k_form_moments<2>, lds); klaunch(KID_FORM_MOMENTS, k_form_moments<2>, a.grid, 256, lds, s, a); break;
            case 3: allow_lds(k_form_moments<3>, lds); klaunch(KID_FORM_MOMENTS, k_form_moments<3>, a.grid, 256, lds, s, a); break;
            default: allow_lds(k_form_moments<4>, lds); klaunch(KID_FORM_MOMENTS, k_form_moments<4>, a.grid, 256, lds, s, a); break;
            }
        }
        size_t lds2 = ((size_t)m * m + 2 * (size_t)m * a.npw + (size_t)a.npw * m) * sizeof(double);
        if (launch_h2pw_wave(a, s)) {
            // matrix-core conversion, one wavefront per leaf
        } else if (a.prec) {
            allow_lds(k_form_h2pw<float>, lds2);
            klaunch(KID_FORM_H2PW, k_form_h2pw<float>, a.nbox, 256, lds2, s, a);
        } else {
            allow_lds(k_form_h2pw<double>, lds2);
            klaunch(KID_FORM_H2PW, k_form_h2pw<double>, a.nbox, 256, lds2, s, a);
        }
        return;
    }
    if (a.prec) fail(-3, "FP32 expansions need the segmented formation with Hermite order < 64");
    const bool dip = a.ws != nullptr;
    int chunk = (dip ? 2048 : 4096) / m;
    if (chunk > 256) chunk = 256;
    if (chunk < 16) chunk = 16;
    size_t lds = (size_t)(dip ? 4 : 2) * chunk * m * sizeof(double);
    size_t lds2 = ((size_t)m * m + (size_t)a.npw * m) * sizeof(double);
    if (lds2 > lds) lds = lds2;
    int ne = (m * m + 255) / 256;
#define FGT_FORM_VALU(E_)                                                                      \
    if (dip) {                                                                                 \
        allow_lds(k_form<E_, true>, lds);                                                      \
        klaunch(KID_FORM_VALU, k_form<E_, true>, a.nbox, 256, lds, s, a, chunk);                                   \
    } else {                                                                                   \
        allow_lds(k_form<E_, false>, lds);                                                     \
        klaunch(KID_FORM_VALU, k_form<E_, false>, a.nbox, 256, lds, s, a, chunk);                                  \
    }
    if (ne <= 1) {
        FGT_FORM_VALU(1)
    } else if (ne <= 2) {
        FGT_FORM_VALU(2)
    } else if (ne <= 4) {
        FGT_FORM_VALU(4)
    } else if (ne <= 8) {
        FGT_FORM_VALU(8)
    } else if (ne <= 16) {
        FGT_FORM_VALU(16)
    } else {
        fail(-3, "Hermite order too large for the formation kernel");
    }
#undef FGT_FORM_VALU
}

// ---------------------------------------------------------------------------
// Steps 2-4: diagonal plane-wave translations (g2dshiftpw_vec,
// g2drouts.f:2356-2402).  One thread per (box, coefficient); HBM-bound.
// Quadrant of a shift from its sign pattern as fgt2d.f:988-998:
// 0 = (+,+), 1 = (+,-), 2 = (-,+), 3 = (-,-).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int quadrant(double dx, double dy)
{
    return dx > 0 ? (dy > 0 ? 0 : 1) : (dy > 0 ? 2 : 3);
}

// E: element type of the expansions and factor tables (double; float on
// the mixed-precision route), complex interleaved.
template <class E>
__global__ __launch_bounds__(256) void k_merge(const int* __restrict__ parents, int np,
                                               const int4* __restrict__ child,
                                               const double* __restrict__ cen,
                                               const int2* __restrict__ src_se,
                                               const int* __restrict__ slot,
                                               const E* __restrict__ wms, int kbase, int nd,
                                               int nexp, E* __restrict__ mp)
{
    long t = blockIdx.x * 256L + threadIdx.x;
    int pi = (int)(t / nexp), e = (int)(t - (long)pi * nexp);
    if (pi >= np) return;
    int p = parents[pi];
    int4 c4 = child[p];
    int cs[4] = {c4.x, c4.y, c4.z, c4.w};
    for (int ind = 0; ind < nd; ++ind) {
        E* dst = mp + 2 * (((long)slot[p] * nd + ind) * nexp + e);
        E re = dst[0], im = dst[1];
        for (int j = 0; j < 4; ++j) {
            int c = cs[j];
            if (c < 0) continue;
            int2 r = src_se[c];
            if (r.y - r.x <= 0) continue;
            int k = quadrant(cen[2 * p] - cen[2 * c], cen[2 * p + 1] - cen[2 * c + 1]);
            const E* w = wms + 2 * ((long)(kbase + k) * nexp + e);
            const E* v = mp + 2 * (((long)slot[c] * nd + ind) * nexp + e);
            re += v[0] * w[0] - v[1] * w[1];
            im += v[0] * w[1] + v[1] * w[0];
        }
        dst[0] = re;
        dst[1] = im;
    }
}

void launch_merge(const int* parents, int np, const int4* child, const double* cen,
                  const int2* src_se, const int* slot, const double* wms, int kbase, int nd,
                  int nexp, double* mp, hipStream_t s, bool f32)
{
    if (np <= 0) return;
    const unsigned grid = grid_for((long)np * nexp, 256);
    if (f32)
        klaunch(KID_MERGE, k_merge<float>, grid, 256, 0, s, parents, np, child, cen, src_se, slot,
                                            (const float*)wms, kbase, nd, nexp, (float*)mp);
    else
        klaunch(KID_MERGE, k_merge<double>, grid, 256, 0, s, parents, np, child, cen, src_se, slot, wms, kbase,
                                             nd, nexp, mp);
}

// Step 3 (fgt2d.f:1026-1064): loc = sum over same-level neighbours j != self
// of mp_j shifted by (jx, jy) boxes, then + mp_self.
template <class E>
__global__ __launch_bounds__(256) void k_m2l(const int* __restrict__ boxes, int nb,
                                             const double* __restrict__ cen,
                                             const int* __restrict__ ncol,
                                             const int* __restrict__ col,
                                             const int2* __restrict__ src_se,
                                             const int2* __restrict__ trg_se, int use_src,
                                             int use_trg, const int* __restrict__ slot,
                                             const E* __restrict__ wm2l, double xmin,
                                             int do_list, int nd, int nexp,
                                             const E* __restrict__ mp,
                                             E* __restrict__ loc)
{
    long t = blockIdx.x * 256L + threadIdx.x;
    int bi = (int)(t / nexp), e = (int)(t - (long)bi * nexp);
    if (bi >= nb) return;
    int b = boxes[bi];
    int np = 0;
    if (use_trg) np += trg_se[b].y - trg_se[b].x;
    if (use_src) np += src_se[b].y - src_se[b].x;
    if (np <= 0) return;
    for (int ind = 0; ind < nd; ++ind) {
        E re = 0, im = 0;
        if (do_list)
            for (int i = 0; i < ncol[b]; ++i) {
                int j = col[9 * b + i];
                if (j == b) continue;
                int jx = (int)lround((cen[2 * b] - cen[2 * j]) / xmin);
                int jy = (int)lround((cen[2 * b + 1] - cen[2 * j + 1]) / xmin);
                jx = jx < -1 ? -1 : (jx > 1 ? 1 : jx);  // colleagues are +-1 box away
                jy = jy < -1 ? -1 : (jy > 1 ? 1 : jy);
                const E* w = wm2l + 2 * ((long)((jx + 1) + 3 * (jy + 1)) * nexp + e);
                const E* v = mp + 2 * (((long)slot[j] * nd + ind) * nexp + e);
                re += v[0] * w[0] - v[1] * w[1];
                im += v[0] * w[1] + v[1] * w[0];
            }
        const E* v = mp + 2 * (((long)slot[b] * nd + ind) * nexp + e);
        re += v[0];
        im += v[1];
        E* d = loc + 2 * (((long)slot[b] * nd + ind) * nexp + e);
        d[0] = re;
        d[1] = im;
    }
}

void launch_m2l(const int* boxes, int nb, const double* cen, const int* ncol, const int* col,
                const int2* src_se, const int2* trg_se, int use_src, int use_trg,
                const int* slot, const double* wm2l, double xmin, int do_list, int nd, int nexp,
                const double* mp, double* loc, hipStream_t s, bool f32)
{
    if (nb <= 0) return;
    const unsigned grid = grid_for((long)nb * nexp, 256);
    if (f32)
        klaunch(KID_M2L, k_m2l<float>, grid, 256, 0, s, boxes, nb, cen, ncol, col, src_se, trg_se, use_src,
                                          use_trg, slot, (const float*)wm2l, xmin, do_list, nd,
                                          nexp, (const float*)mp, (float*)loc);
    else
        klaunch(KID_M2L, k_m2l<double>, grid, 256, 0, s, boxes, nb, cen, ncol, col, src_se, trg_se, use_src,
                                           use_trg, slot, wm2l, xmin, do_list, nd, nexp, mp,
                                           loc);
}

// Step 4 (fgt2d.f:1082-1125): loc(child) = loc(parent) shifted to the child.
template <class E>
__global__ __launch_bounds__(256) void k_split(const int* __restrict__ children, int nc,
                                               const int* __restrict__ par,
                                               const double* __restrict__ cen,
                                               const int* __restrict__ slot,
                                               const E* __restrict__ wms, int kbase, int nd,
                                               int nexp, E* __restrict__ loc)
{
    long t = blockIdx.x * 256L + threadIdx.x;
    int ci = (int)(t / nexp), e = (int)(t - (long)ci * nexp);
    if (ci >= nc) return;
    int c = children[ci], p = par[c];
    int k = quadrant(cen[2 * c] - cen[2 * p], cen[2 * c + 1] - cen[2 * p + 1]);
    const E* w = wms + 2 * ((long)(kbase + k) * nexp + e);
    for (int ind = 0; ind < nd; ++ind) {
        const E* v = loc + 2 * (((long)slot[p] * nd + ind) * nexp + e);
        E* d = loc + 2 * (((long)slot[c] * nd + ind) * nexp + e);
        d[0] += v[0] * w[0] - v[1] * w[1];
        d[1] += v[0] * w[1] + v[1] * w[0];
    }
}

void launch_split(const int* children, int nc, const int* par, const double* cen,
                  const int* slot, const double* wms, int kbase, int nd, int nexp, double* loc,
                  hipStream_t s, bool f32)
{
    if (nc <= 0) return;
    const unsigned grid = grid_for((long)nc * nexp, 256);
    if (f32)
        klaunch(KID_SPLIT, k_split<float>, grid, 256, 0, s, children, nc, par, cen, slot, (const float*)wms,
                                            kbase, nd, nexp, (float*)loc);
    else
        klaunch(KID_SPLIT, k_split<double>, grid, 256, 0, s, children, nc, par, cen, slot, wms, kbase, nd, nexp,
                                             loc);
}

// ---------------------------------------------------------------------------
// Step 5a: plane wave -> Taylor (g2dpw2local_real_vec, g2drouts.f:1417-1540)
//   lx(k,b)  = i^k sum_a x2l(a,k) PW(a,b)          a positive x nodes
//   L(k,l)   = 2 Re[ i^l sum_{b<half} x2l(b,l) (lx(k,b) + (-1)^l lx(k,b+half)) ]
// with x2l(a,k) = ts_a^k / k!.  One block per leaf.  E: element type of
// the expansions read, TT of the Taylor coefficients written (float on the
// mixed-precision route); the arithmetic is FP64.
// ---------------------------------------------------------------------------
template <class E, class TT>
__global__ __launch_bounds__(256) void k_pw2local(EvalArgs a)
{
    extern __shared__ double sm[];
    int bi = blockIdx.x;
    int b = a.boxes[bi];
    const int npw = a.npw, half = npw / 2, nexp = half * npw, m = a.nl + 1, t = threadIdx.x;
    const E* pw = (const E*)a.loc + 2 * ((long)a.slot[b] * a.nd + a.ind) * nexp;
    double* P = sm;              // complex nexp
    double* LX = sm + 2 * nexp;  // complex (m, npw): k fastest
    for (int o = t; o < 2 * nexp; o += 256) P[o] = pw[o];
    __syncthreads();
    for (int o = t; o < m * npw; o += 256) {
        int k = o % m, bb = o / m;
        double cr = 0, ci = 0;
        for (int aa = 0; aa < half; ++aa) {
            double x = a.x2l[aa + npw * k];
            cr += x * P[2 * (aa + half * bb)];
            ci += x * P[2 * (aa + half * bb) + 1];
        }
        double r, im;
        switch (k & 3) {
        case 0: r = cr; im = ci; break;
        case 1: r = -ci; im = cr; break;
        case 2: r = -cr; im = -ci; break;
        default: r = ci; im = -cr; break;
        }
        LX[2 * o] = r;
        LX[2 * o + 1] = im;
    }
    __syncthreads();
    TT* L = (TT*)a.taylor + (long)bi * m * m;
    for (int o = t; o < m * m; o += 256) {
        int k = o % m, l = o / m;
        double sg = (l & 1) ? -1.0 : 1.0;
        double cr = 0, ci = 0;
        for (int bb = 0; bb < half; ++bb) {
            double x = a.x2l[bb + npw * l];
            cr += x * (LX[2 * (k + m * bb)] + sg * LX[2 * (k + m * (bb + half))]);
            ci += x * (LX[2 * (k + m * bb) + 1] + sg * LX[2 * (k + m * (bb + half)) + 1]);
        }
        double v;
        switch (l & 3) {
        case 0: v = 2 * cr; break;
        case 1: v = -2 * ci; break;
        case 2: v = -2 * cr; break;
        default: v = 2 * ci; break;
        }
        L[o] = (TT)v;
    }
}

static bool launch_pw2local_wave(const EvalArgs& a, hipStream_t s);

void launch_pw2local(const EvalArgs& a, hipStream_t s)
{
    if (a.nbox <= 0) return;
    if (launch_pw2local_wave(a, s)) return;  // matrix cores, one wavefront per leaf
    int half = a.npw / 2, m = a.nl + 1;
    size_t lds = (size_t)(2 * half * a.npw + 2 * m * a.npw) * sizeof(double);
    if (a.exp32 && a.tay32)
        klaunch(KID_PW2LOCAL, k_pw2local<float, float>, a.nbox, 256, lds, s, a);
    else if (a.exp32)
        klaunch(KID_PW2LOCAL, k_pw2local<float, double>, a.nbox, 256, lds, s, a);
    else
        klaunch(KID_PW2LOCAL, k_pw2local<double, double>, a.nbox, 256, lds, s, a);
}

// Step 5b: Taylor evaluation (g2dlevalp/g/h_vec, g2drouts.f:847-1080) by
// nested Horner schemes; derivatives carried along.  One block per leaf.
template <int PGH>
__global__ __launch_bounds__(256) void k_taylor_eval(EvalArgs a)
{
    extern __shared__ double sm[];
    const int nl = a.nl, m = nl + 1, t = threadIdx.x;
    const int total = seg_count(a.tstart, a.nbox);
    for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const SegRange sr = seg_at(a.tstart, a.nbox, a.boxes, a.se, a.seg, g);
    const int b = a.boxes[sr.li];
    const double* Lg = a.taylor + (long)sr.li * m * m;
    __syncthreads();
    for (int o = t; o < m * m; o += 256) sm[o] = Lg[o];
    __syncthreads();
    const double cx = a.cen[2 * b], cy = a.cen[2 * b + 1], dsq = a.dsq;
    for (int i = sr.p0 + t; i < sr.p1; i += 256) {
        double x = (a.pts[2 * (long)i] - cx) * dsq, y = (a.pts[2 * (long)i + 1] - cy) * dsq;
        // V = sum_k y^k P_k(x), U = sum_k y^k P_k'(x), Z = sum_k y^k P_k''(x)/2
        double V = 0, Vy = 0, Vyy = 0, U = 0, Uy = 0, Z = 0;
        for (int k = nl; k >= 0; --k) {
            const double* c = sm + m * k;
            double q = c[nl], q1 = 0, q2 = 0;
            for (int j = nl - 1; j >= 0; --j) {
                if (PGH >= 3) q2 = q2 * x + q1;
                if (PGH >= 2) q1 = q1 * x + q;
                q = q * x + c[j];
            }
            if (PGH >= 3) {
                Vyy = Vyy * y + Vy;
                Uy = Uy * y + U;
                Z = Z * y + q2;
            }
            if (PGH >= 2) {
                Vy = Vy * y + V;
                U = U * y + q1;
            }
            V = V * y + q;
        }
        eval_out<PGH>(a, i, V, U * dsq, Vy * dsq, 2 * Z * dsq * dsq, Uy * dsq * dsq,
                      2 * Vyy * dsq * dsq);
    }
    }
}

// Step 5b on matrix cores.  For a tile of 16 points of a leaf the MFMA
// computes D(k,t) = sum_j L(j,k) xh_t^j with A = L^T (LDS, x-derivative
// variants Lx, Lxx for gradients and Hessians) and B(j,t) = xh_t^j (running
// products in registers).  Lane l then holds y orders k = l/16 + 4r (+16 per
// column tile) of point t = l%16, so sum_k D(k,t) yh_t^k is a running
// product in registers plus a reduction over the 4 lanes of the point.
// Outputs are accumulated: the near field of a point may already be in
// them (each point is evaluated by exactly one leaf).
template <int TM, int PGH>
__global__ __launch_bounds__(256) void k_taylor_mfma(EvalArgs a)
{
    extern __shared__ double sm[];
    constexpr int MP = 16 * TM, NV = PGH >= 3 ? 3 : (PGH >= 2 ? 2 : 1);
    const int m = a.nl + 1, t = threadIdx.x, lane = t & 63, w = t >> 6;
    double* Lb = sm;  // NV x MP x MP, row j, column k
    const int total = seg_count(a.tstart, a.nbox);
    for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const SegRange sr = seg_at(a.tstart, a.nbox, a.boxes, a.se, a.seg, g);
    const int b = a.boxes[sr.li];
    const double* Lg = a.taylor + (long)sr.li * m * m;  // L(j,k) at j + m k
    __syncthreads();
    for (int o = t; o < MP * MP; o += 256) {
        int j = o / MP, k = o % MP;
        Lb[o] = (j < m && k < m) ? Lg[j + m * k] : 0.0;
        if (NV >= 2) Lb[MP * MP + o] = (j + 1 < m && k < m) ? (j + 1) * Lg[(j + 1) + m * k] : 0.0;
        if (NV >= 3)
            Lb[2 * MP * MP + o] =
                (j + 2 < m && k < m) ? (double)((j + 1) * (j + 2)) * Lg[(j + 2) + m * k] : 0.0;
    }
    __syncthreads();
    const double cx = a.cen[2 * b], cy = a.cen[2 * b + 1], dsq = a.dsq;
    const int col = lane & 15, kq = lane >> 4;
    const int npts = sr.p1 - sr.p0;
    for (int base = 16 * w; base < npts; base += 64) {
        const int ip = base + col;
        const bool act = ip < npts;
        const long i = sr.p0 + (act ? ip : 0);
        const double x = act ? (a.pts[2 * i] - cx) * dsq : 0.0;
        const double y = act ? (a.pts[2 * i + 1] - cy) * dsq : 0.0;
        double p = 1.0;
        for (int r = 0; r < kq; ++r) p *= x;  // x^(l/16)
        double x4 = x * x;
        x4 *= x4;
        d4_t acc[NV][TM];
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int j = 0; j < TM; ++j) acc[v][j] = (d4_t){0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < MP / 4; ++ks) {
            const int row = 4 * ks + kq;
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int j = 0; j < TM; ++j)
                    acc[v][j] = mfma16x16x4(Lb[v * MP * MP + row * MP + col + 16 * j], p, acc[v][j]);
            p *= x4;
        }
        // this lane: orders k = kq + 4r + 16j of point col
        double yk = 1.0;  // y^kq
        for (int r = 0; r < kq; ++r) yk *= y;
        double y4 = y * y;
        y4 *= y4;
        double V = 0, Vy = 0, Vyy = 0, U = 0, Uy = 0, Z = 0;
#pragma unroll
        for (int j = 0; j < TM; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                V += acc[0][j][r] * yk;
                if (PGH >= 2) U += acc[1][j][r] * yk;
                if (PGH >= 3) Z += acc[2][j][r] * yk;
                yk *= y4;
            }
        if (PGH >= 2) {
            // y derivatives: k y^(k-1) and k (k-1) y^(k-2) along this lane's
            // orders; e1, e2 start at the first order where they are nonzero
            const int f1 = kq >= 1 ? 0 : 1, f2 = kq >= 2 ? 0 : 1;
            const int p1 = kq >= 1 ? kq - 1 : 3, p2 = kq >= 2 ? kq - 2 : kq + 2;
            double e1 = 1.0, e2 = 1.0;
            for (int r = 0; r < p1; ++r) e1 *= y;
            for (int r = 0; r < p2; ++r) e2 *= y;
            int st = 0;
#pragma unroll
            for (int j = 0; j < TM; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r, ++st) {
                    const double k = kq + 4 * r + 16 * j;
                    if (st >= f1) {
                        Vy += k * acc[0][j][r] * e1;
                        if (PGH >= 3) Uy += k * acc[1][j][r] * e1;
                        e1 *= y4;
                    }
                    if (PGH >= 3 && st >= f2) {
                        Vyy += k * (k - 1) * acc[0][j][r] * e2;
                        e2 *= y4;
                    }
                }
        }
        // sum over the four lanes that hold this point
        V += __shfl_xor(V, 16);
        V += __shfl_xor(V, 32);
        if (PGH >= 2) {
            U += __shfl_xor(U, 16);
            U += __shfl_xor(U, 32);
            Vy += __shfl_xor(Vy, 16);
            Vy += __shfl_xor(Vy, 32);
        }
        if (PGH >= 3) {
            Z += __shfl_xor(Z, 16);
            Z += __shfl_xor(Z, 32);
            Uy += __shfl_xor(Uy, 16);
            Uy += __shfl_xor(Uy, 32);
            Vyy += __shfl_xor(Vyy, 16);
            Vyy += __shfl_xor(Vyy, 32);
        }
        if (kq == 0 && act)
            eval_out<PGH>(a, i, V, U * dsq, Vy * dsq, Z * dsq * dsq, Uy * dsq * dsq,
                          Vyy * dsq * dsq);
    }
    }
}

// Step 5a inside the evaluation wave: the Taylor coefficients of one leaf
// from its plane-wave expansion on the wave's FP64 matrix cores (the
// mathematics of g2dpw2local_real_vec, g2drouts.f:1417-1540, as k_pw2local):
//   c(k,b)  = sum_{a<half} x2l(a,k) PW(a,b),   lx(k,b) = i^k c(k,b)
//   L(k,l)  = 2 Re[ i^l sum_{b<half} x2l(b,l) (lx(k,b) + (-1)^l lx(k,b+half)) ]
// With Er(k,b) = Re lx(k,b) + Re lx(k,b+half) and Oi(k,b) = Im lx(k,b) -
// Im lx(k,b+half), L(k,l) = s sum_b x2l(b,l) Er(k,b) for even l (s = 2 at
// l = 0 mod 4, -2 at 2) and s sum_b x2l(b,l) Oi(k,b) for odd l (s = -2 at 1,
// 2 at 3).  The first product takes the columns b and b + half as two tiles
// of the same lanes, so Er and Oi form lane by lane.  Writes L(k,l) (x order
// k, y order l) to buf[k + MP l], zero beyond m; buf holds
// PW2L_BUF(TM, NBH) doubles.  Every lane of the wave calls it.
#define PW2L_BUF(TM, NBH) (2 * 16 * (TM) * 16 * (NBH) > 256 * (TM) * (TM) ? 2 * 16 * (TM) * 16 * (NBH) : 256 * (TM) * (TM))

template <int TM, int NBH, class E>
__device__ __forceinline__ void pw2local_wave(const EvalArgs& a, const E* __restrict__ pw,
                                              double* buf, int lane)
{
    constexpr int MP = 16 * TM, BW = 16 * NBH;
    const int m = a.nl + 1, npw = a.npw, half = npw / 2;
    const int c16 = lane & 15, q4 = lane >> 4;
    const double* __restrict__ x2l = a.x2l;  // x2l(a,k) at a + npw k
    d4_t C[TM][NBH][4];  // [it][jt][lo re, lo im, hi re, hi im]
#pragma unroll
    for (int it = 0; it < TM; ++it)
#pragma unroll
        for (int jt = 0; jt < NBH; ++jt)
#pragma unroll
            for (int u = 0; u < 4; ++u) C[it][jt][u] = (d4_t){0, 0, 0, 0};
    for (int k0 = 0; k0 < half; k0 += 4) {
        const int an = k0 + q4;  // the node of this lane's K index
        double av[TM];
#pragma unroll
        for (int it = 0; it < TM; ++it) {
            const int k = c16 + 16 * it;
            av[it] = (an < half && k < m) ? x2l[an + npw * k] : 0.0;
        }
#pragma unroll
        for (int jt = 0; jt < NBH; ++jt) {
            const int bc = c16 + 16 * jt;
            const bool ok = an < half && bc < half;
            const E* p0 = pw + 2 * (an + half * bc);           // PW(a, b)
            const E* p1 = pw + 2 * (an + half * (bc + half));  // PW(a, b + half)
            const double b0r = ok ? (double)p0[0] : 0.0, b0i = ok ? (double)p0[1] : 0.0;
            const double b1r = ok ? (double)p1[0] : 0.0, b1i = ok ? (double)p1[1] : 0.0;
#pragma unroll
            for (int it = 0; it < TM; ++it) {
                C[it][jt][0] = mfma16x16x4(av[it], b0r, C[it][jt][0]);
                C[it][jt][1] = mfma16x16x4(av[it], b0i, C[it][jt][1]);
                C[it][jt][2] = mfma16x16x4(av[it], b1r, C[it][jt][2]);
                C[it][jt][3] = mfma16x16x4(av[it], b1i, C[it][jt][3]);
            }
        }
    }
    double* Er = buf;            // Er(k,b) at k + MP b, b < BW
    double* Oi = buf + MP * BW;  // Oi(k,b) likewise
    wave_lds_sync();             // the previous segment's reads of buf are done
#pragma unroll
    for (int it = 0; it < TM; ++it)
#pragma unroll
        for (int jt = 0; jt < NBH; ++jt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = q4 + 4 * r + 16 * it, b = c16 + 16 * jt;
                const double lr = C[it][jt][0][r], li = C[it][jt][1][r];
                const double hr = C[it][jt][2][r], hi = C[it][jt][3][r];
                double er, oi;  // i^k times lo and hi
                switch (k & 3) {
                case 0: er = lr + hr; oi = li - hi; break;
                case 1: er = -li - hi; oi = lr - hr; break;
                case 2: er = -lr - hr; oi = -li + hi; break;
                default: er = li + hi; oi = -lr + hr; break;
                }
                Er[k + MP * b] = er;
                Oi[k + MP * b] = oi;
            }
    wave_lds_sync();
    d4_t P[2][TM][TM];
#pragma unroll
    for (int it = 0; it < TM; ++it)
#pragma unroll
        for (int lt = 0; lt < TM; ++lt) P[0][it][lt] = P[1][it][lt] = (d4_t){0, 0, 0, 0};
    for (int b0 = 0; b0 < half; b0 += 4) {
        const int bn = b0 + q4;
        double ae[TM], ao[TM], bv[TM];
#pragma unroll
        for (int it = 0; it < TM; ++it) {
            const int k = c16 + 16 * it;
            ae[it] = bn < half ? Er[k + MP * bn] : 0.0;
            ao[it] = bn < half ? Oi[k + MP * bn] : 0.0;
        }
#pragma unroll
        for (int lt = 0; lt < TM; ++lt) {
            const int l = c16 + 16 * lt;
            bv[lt] = (bn < half && l < m) ? x2l[bn + npw * l] : 0.0;
        }
#pragma unroll
        for (int it = 0; it < TM; ++it)
#pragma unroll
            for (int lt = 0; lt < TM; ++lt) {
                P[0][it][lt] = mfma16x16x4(ae[it], bv[lt], P[0][it][lt]);
                P[1][it][lt] = mfma16x16x4(ao[it], bv[lt], P[1][it][lt]);
            }
    }
    wave_lds_sync();  // Er, Oi read: buf takes L
#pragma unroll
    for (int it = 0; it < TM; ++it)
#pragma unroll
        for (int lt = 0; lt < TM; ++lt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = q4 + 4 * r + 16 * it, l = c16 + 16 * lt;
                double v;
                switch (l & 3) {
                case 0: v = 2 * P[0][it][lt][r]; break;
                case 1: v = -2 * P[1][it][lt][r]; break;
                case 2: v = -2 * P[0][it][lt][r]; break;
                default: v = 2 * P[1][it][lt][r]; break;
                }
                buf[k + MP * l] = (k < m && l < m) ? v : 0.0;
            }
    wave_lds_sync();
}

// Step 5a on matrix cores as its own kernel: one wavefront per leaf (four
// per block), pw2local_wave into the wave's LDS slice, then L(k,l) stored
// at k + m l as TT (FP32 on the mixed-precision route) for the Taylor
// evaluation kernels.
template <int TM, int NBH, class E, class TT>
__global__ __launch_bounds__(256) void k_pw2local_wave(EvalArgs a)
{
    constexpr int MP = 16 * TM;
    extern __shared__ double dsm[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int bi = blockIdx.x * 4 + w;
    if (bi >= a.nbox) return;  // wave-uniform: no block barrier below
    const int b = a.boxes[bi], m = a.nl + 1;
    double* buf = dsm + w * PW2L_BUF(TM, NBH);
    const long nexp = (long)(a.npw / 2) * a.npw;
    pw2local_wave<TM, NBH, E>(a, (const E*)a.loc + 2 * ((long)a.slot[b] * a.nd + a.ind) * nexp, buf,
                              lane);
    TT* L = (TT*)a.taylor + (long)bi * m * m;
    for (int o = lane; o < m * m; o += 64) L[o] = (TT)buf[o % m + MP * (o / m)];
}

template <int TM, int NBH>
static void pw2local_wave_launch(const EvalArgs& a, hipStream_t s)
{
    const unsigned lds = 4 * PW2L_BUF(TM, NBH) * sizeof(double);
    const unsigned grid = (unsigned)((a.nbox + 3) / 4);
#define FGT_P2L(E_, T_)                                                                        \
    {                                                                                          \
        allow_lds(k_pw2local_wave<TM, NBH, E_, T_>, lds);                                      \
        klaunch(KID_PW2LOCAL, k_pw2local_wave<TM, NBH, E_, T_>, grid, 256, lds, s, a);         \
    }
    if (a.exp32 && a.tay32)
        FGT_P2L(float, float)
    else if (a.exp32)
        FGT_P2L(float, double)
    else
        FGT_P2L(double, double)
#undef FGT_P2L
}

// orders up to 48 and npw <= 64; false when the VALU kernel has to run
static bool launch_pw2local_wave(const EvalArgs& a, hipStream_t s)
{
    const int m = a.nl + 1, nbh = (a.npw / 2 + 15) / 16;
    if (m > 48 || a.npw > 64) return false;
    const int tm = (m + 15) / 16;
    if (tm == 1 && nbh == 1) pw2local_wave_launch<1, 1>(a, s);
    else if (tm == 1) pw2local_wave_launch<1, 2>(a, s);
    else if (tm == 2 && nbh == 1) pw2local_wave_launch<2, 1>(a, s);
    else if (tm == 2) pw2local_wave_launch<2, 2>(a, s);
    else if (nbh == 1) pw2local_wave_launch<3, 1>(a, s);
    else pw2local_wave_launch<3, 2>(a, s);
    return true;
}

// Step 5b on matrix cores, one wavefront per segment (at most WSEG points):
// the lane's A operands L(4 ks + l/16, l%16 + 16 j) (and the x-derivative
// variants) are loaded once per segment into registers, the segment's
// scaled points are staged in the wave's slice of LDS in one coalesced
// pass, then every tile of 16 points runs as in k_taylor_mfma.
//
// FUSE = NBH > 0: the Taylor coefficients are not read from a.taylor but
// formed by the wave itself from the leaf's plane-wave expansion (element
// type E), pw2local_wave<TM, NBH, E>, into its slice of dynamic LDS.
template <int TM, int PGH, int FUSE = 0, class E = double>
__global__ __launch_bounds__(256) void k_taylor_wave(EvalArgs a)
{
    constexpr int MP = 16 * TM, KS = MP / 4, NV = PGH >= 3 ? 3 : (PGH >= 2 ? 2 : 1);
    __shared__ double stage[4][2][WSEG];
    __shared__ double stage_nf[4][WSEG];   // the near field's potentials
    __shared__ unsigned stage_j[4][WSEG];  // input positions
    extern __shared__ double dsm[];
    const int m = a.nl + 1, lane = threadIdx.x & 63;
    double* sx = stage[threadIdx.x >> 6][0];
    double* sy = stage[threadIdx.x >> 6][1];
    double* snf = stage_nf[threadIdx.x >> 6];
    unsigned* sj = stage_j[threadIdx.x >> 6];
    const int col = lane & 15, kq = lane >> 4;
    const int total = seg_count(a.tstart, a.nbox);
    // fused: each wave takes a run of a.spw consecutive segments (mostly of
    // one leaf) and forms a leaf's coefficients once per run; unfused: the
    // segments stride over the waves
    const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int g_lo = FUSE ? wid * a.spw : wid, g_hi = FUSE ? min(total, g_lo + a.spw) : total;
    const int g_step = FUSE ? 1 : gridDim.x * 4;
    int last_li = -1;
    for (int g = g_lo; g < g_hi; g += g_step) {
        const SegRange sr = seg_at(a.tstart, a.nbox, a.boxes, a.se, a.seg, g, a.segbox);
        const int b = a.boxes[sr.li];
        const double* Lg = a.taylor + (long)sr.li * m * m;  // L(j,k) at j + m k
        double* Lb = nullptr;                                // fused: L(j,k) at j + MP k
        if (FUSE) {
            Lb = dsm + (threadIdx.x >> 6) * PW2L_BUF(TM, FUSE);
            if (sr.li != last_li) {  // wave-uniform
                const long nexp = (long)(a.npw / 2) * a.npw;
                pw2local_wave<TM, (FUSE > 0 ? FUSE : 1), E>(
                    a, (const E*)a.loc + 2 * ((long)a.slot[b] * a.nd + a.ind) * nexp, Lb, lane);
                last_li = sr.li;
            }
        }
        auto Lv = [&](int j, int k) -> double { return FUSE ? Lb[j + MP * k] : Lg[j + m * k]; };
        double A[NV][KS][TM];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int jt = 0; jt < TM; ++jt) {
                const int j = 4 * ks + kq, k = col + 16 * jt;
                const bool kin = k < m;
                A[0][ks][jt] = (kin && j < m) ? Lv(j, k) : 0.0;
                if (NV >= 2) A[NV >= 2 ? 1 : 0][ks][jt] = (kin && j + 1 < m) ? (j + 1) * Lv(j + 1, k) : 0.0;
                if (NV >= 3)
                    A[NV >= 3 ? 2 : 0][ks][jt] =
                        (kin && j + 2 < m) ? (double)((j + 1) * (j + 2)) * Lv(j + 2, k) : 0.0;
            }
        const double cx = a.cen[2 * b], cy = a.cen[2 * b + 1], dsq = a.dsq;
        const int npts = sr.p1 - sr.p0;
        wave_lds_sync();  // the previous segment's reads are done
#pragma unroll
        for (int k = 0; k < WSEG / 64; ++k) {
            const int u = lane + 64 * k;
            if (u < npts) {
                const long i = sr.p0 + u;
                sx[u] = (a.pts[2 * i] - cx) * dsq;
                sy[u] = (a.pts[2 * i + 1] - cy) * dsq;
                snf[u] = a.npot[i * a.nd + a.ind];
                sj[u] = a.perm[i];
            }
        }
        wave_lds_sync();
        for (int base = 0; base < npts; base += 16) {
            const int ip = base + col;
            const bool act = ip < npts;
            const long i = sr.p0 + (act ? ip : 0);
            const double x = act ? sx[ip] : 0.0;
            const double y = act ? sy[ip] : 0.0;
            double p = 1.0;
            for (int r = 0; r < kq; ++r) p *= x;  // x^(l/16)
            double x4 = x * x;
            x4 *= x4;
            d4_t acc[NV][TM];
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int j = 0; j < TM; ++j) acc[v][j] = (d4_t){0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
                for (int v = 0; v < NV; ++v)
#pragma unroll
                    for (int j = 0; j < TM; ++j) acc[v][j] = mfma16x16x4(A[v][ks][j], p, acc[v][j]);
                p *= x4;
            }
            double yk = 1.0;  // y^kq
            for (int r = 0; r < kq; ++r) yk *= y;
            double y4 = y * y;
            y4 *= y4;
            double V = 0, Vy = 0, Vyy = 0, U = 0, Uy = 0, Z = 0;
#pragma unroll
            for (int j = 0; j < TM; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    V += acc[0][j][r] * yk;
                    if (PGH >= 2) U += acc[NV >= 2 ? 1 : 0][j][r] * yk;
                    if (PGH >= 3) Z += acc[NV >= 3 ? 2 : 0][j][r] * yk;
                    yk *= y4;
                }
            if (PGH >= 2) {
                const int f1 = kq >= 1 ? 0 : 1, f2 = kq >= 2 ? 0 : 1;
                const int p1 = kq >= 1 ? kq - 1 : 3, p2 = kq >= 2 ? kq - 2 : kq + 2;
                double e1 = 1.0, e2 = 1.0;
                for (int r = 0; r < p1; ++r) e1 *= y;
                for (int r = 0; r < p2; ++r) e2 *= y;
                int st = 0;
#pragma unroll
                for (int j = 0; j < TM; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r, ++st) {
                        const double k = kq + 4 * r + 16 * j;
                        if (st >= f1) {
                            Vy += k * acc[0][j][r] * e1;
                            if (PGH >= 3) Uy += k * acc[NV >= 2 ? 1 : 0][j][r] * e1;
                            e1 *= y4;
                        }
                        if (PGH >= 3 && st >= f2) {
                            Vyy += k * (k - 1) * acc[0][j][r] * e2;
                            e2 *= y4;
                        }
                    }
            }
            V += __shfl_xor(V, 16);
            V += __shfl_xor(V, 32);
            if (PGH >= 2) {
                U += __shfl_xor(U, 16);
                U += __shfl_xor(U, 32);
                Vy += __shfl_xor(Vy, 16);
                Vy += __shfl_xor(Vy, 32);
            }
            if (PGH >= 3) {
                Z += __shfl_xor(Z, 16);
                Z += __shfl_xor(Z, 32);
                Uy += __shfl_xor(Uy, 16);
                Uy += __shfl_xor(Uy, 32);
                Vyy += __shfl_xor(Vyy, 16);
                Vyy += __shfl_xor(Vyy, 32);
            }
            if (kq == 0 && act)
                eval_out<PGH>(a, i, sj[ip], snf[ip], V, U * dsq, Vy * dsq, Z * dsq * dsq,
                              Uy * dsq * dsq, Vyy * dsq * dsq);
        }
    }
}

// Step 5b with FP32 operands (the mixed-precision route): k_taylor_wave
// with FP32 Taylor coefficients (written by k_pw2local<., float>) and FP32
// powers of the offsets, which are formed in FP64 from the FP64 box centre;
// FP32 matrix-core products over the x orders, then the y weighting, the
// lane sums and the accumulation into the outputs in FP64.  A lane holds y
// orders 4 (l/16) + r + 16 j (the FP32 result layout).
template <int TM, int PGH, int FUSE = 0, class E = float>
__global__ __launch_bounds__(256) void k_taylor_wave_f32(EvalArgs a)
{
    constexpr int MP = 16 * TM, KS = MP / 4, NV = PGH >= 3 ? 3 : (PGH >= 2 ? 2 : 1);
    __shared__ float stage[4][2][WSEG];
    __shared__ double stage_nf[4][WSEG];   // the near field's potentials
    __shared__ unsigned stage_j[4][WSEG];  // input positions
    extern __shared__ double dsm[];
    const int m = a.nl + 1, lane = threadIdx.x & 63;
    float* sx = stage[threadIdx.x >> 6][0];
    float* sy = stage[threadIdx.x >> 6][1];
    double* snf = stage_nf[threadIdx.x >> 6];
    unsigned* sj = stage_j[threadIdx.x >> 6];
    const int col = lane & 15, kq = lane >> 4;
    const int total = seg_count(a.tstart, a.nbox);
    // fused: each wave takes a run of a.spw consecutive segments (mostly of
    // one leaf) and forms a leaf's coefficients once per run; unfused: the
    // segments stride over the waves
    const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int g_lo = FUSE ? wid * a.spw : wid, g_hi = FUSE ? min(total, g_lo + a.spw) : total;
    const int g_step = FUSE ? 1 : gridDim.x * 4;
    int last_li = -1;
    for (int g = g_lo; g < g_hi; g += g_step) {
        const SegRange sr = seg_at(a.tstart, a.nbox, a.boxes, a.se, a.seg, g, a.segbox);
        const int b = a.boxes[sr.li];
        const float* Lg = (const float*)a.taylor + (long)sr.li * m * m;  // L(j,k) at j + m k
        double* Lb = nullptr;  // fused: FP64 L(j,k) at j + MP k, rounded to FP32 below
        if (FUSE) {
            Lb = dsm + (threadIdx.x >> 6) * PW2L_BUF(TM, FUSE);
            if (sr.li != last_li) {  // wave-uniform
                const long nexp = (long)(a.npw / 2) * a.npw;
                pw2local_wave<TM, (FUSE > 0 ? FUSE : 1), E>(
                    a, (const E*)a.loc + 2 * ((long)a.slot[b] * a.nd + a.ind) * nexp, Lb, lane);
                last_li = sr.li;
            }
        }
        // (j+1) L and (j+1)(j+2) L as in the unfused form: FP32 products of
        // the FP32 coefficient
        auto Lv = [&](int j, int k) -> float { return FUSE ? (float)Lb[j + MP * k] : Lg[j + m * k]; };
        float A[NV][KS][TM];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int jt = 0; jt < TM; ++jt) {
                const int j = 4 * ks + kq, k = col + 16 * jt;
                const bool kin = k < m;
                A[0][ks][jt] = (kin && j < m) ? Lv(j, k) : 0.f;
                if (NV >= 2)
                    A[NV >= 2 ? 1 : 0][ks][jt] = (kin && j + 1 < m) ? (float)(j + 1) * Lv(j + 1, k) : 0.f;
                if (NV >= 3)
                    A[NV >= 3 ? 2 : 0][ks][jt] =
                        (kin && j + 2 < m) ? (float)((j + 1) * (j + 2)) * Lv(j + 2, k) : 0.f;
            }
        const double cx = a.cen[2 * b], cy = a.cen[2 * b + 1], dsq = a.dsq;
        const int npts = sr.p1 - sr.p0;
        wave_lds_sync();  // the previous segment's reads are done
#pragma unroll
        for (int k = 0; k < WSEG / 64; ++k) {
            const int u = lane + 64 * k;
            if (u < npts) {
                const long i = sr.p0 + u;
                sx[u] = (float)((a.pts[2 * i] - cx) * dsq);
                sy[u] = (float)((a.pts[2 * i + 1] - cy) * dsq);
                snf[u] = a.npot[i * a.nd + a.ind];
                sj[u] = a.perm[i];
            }
        }
        wave_lds_sync();
        for (int base = 0; base < npts; base += 16) {
            const int ip = base + col;
            const bool act = ip < npts;
            const long i = sr.p0 + (act ? ip : 0);
            const float x = act ? sx[ip] : 0.f;
            const double y = act ? (double)sy[ip] : 0.0;
            float p = 1.f;
            for (int r = 0; r < kq; ++r) p *= x;  // x^(l/16)
            float x4 = x * x;
            x4 *= x4;
            f4_t acc[NV][TM];
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int j = 0; j < TM; ++j) acc[v][j] = (f4_t){0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
                for (int v = 0; v < NV; ++v)
#pragma unroll
                    for (int j = 0; j < TM; ++j) acc[v][j] = mfma16x16x4f32(A[v][ks][j], p, acc[v][j]);
                p *= x4;
            }
            // y powers of this lane's rows: y^k, y^(k-1), y^(k-2) for
            // k = 4 kq + r + 16 j, advanced by y per r; at the first row of
            // tile j: t0 = y^(4 kq + 16 j) and y^(k-1), y^(k-2) from t1, t2
            // (tile 0; zero where the exponent is negative, whose factor k
            // or k (k-1) is zero) or u1, u2 (later tiles)
            const double y2 = y * y, y4 = y2 * y2, y8 = y4 * y4, y16 = y8 * y8;
            double tq = 1.0;  // y^(4 (kq - 1)) for kq >= 1
            for (int r = 1; r < kq; ++r) tq *= y4;
            double t0 = kq >= 1 ? tq * y4 : 1.0;
            const double t1 = kq >= 1 ? tq * y2 * y : 0.0, t2 = kq >= 1 ? tq * y2 : 0.0;
            double u1 = t0 * (y8 * y4 * y2 * y), u2 = t0 * (y8 * y4 * y2);
            double V = 0, Vy = 0, Vyy = 0, U = 0, Uy = 0, Z = 0;
#pragma unroll
            for (int j = 0; j < TM; ++j) {
                double p0 = t0, p1 = j == 0 ? t1 : u1, p2 = j == 0 ? t2 : u2;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double k = 4 * kq + r + 16 * j;
                    const double c0 = acc[0][j][r];
                    V += c0 * p0;
                    if (PGH >= 2) {
                        U += (double)acc[NV >= 2 ? 1 : 0][j][r] * p0;
                        Vy += k * c0 * p1;
                    }
                    if (PGH >= 3) {
                        Z += (double)acc[NV >= 3 ? 2 : 0][j][r] * p0;
                        Uy += k * (double)acc[NV >= 2 ? 1 : 0][j][r] * p1;
                        Vyy += k * (k - 1) * c0 * p2;
                    }
                    p2 = p1;
                    p1 = p0;
                    p0 *= y;
                }
                t0 *= y16;
                if (j > 0) {
                    u1 *= y16;
                    u2 *= y16;
                }
            }
            V += __shfl_xor(V, 16);
            V += __shfl_xor(V, 32);
            if (PGH >= 2) {
                U += __shfl_xor(U, 16);
                U += __shfl_xor(U, 32);
                Vy += __shfl_xor(Vy, 16);
                Vy += __shfl_xor(Vy, 32);
            }
            if (PGH >= 3) {
                Z += __shfl_xor(Z, 16);
                Z += __shfl_xor(Z, 32);
                Uy += __shfl_xor(Uy, 16);
                Uy += __shfl_xor(Uy, 32);
                Vyy += __shfl_xor(Vyy, 16);
                Vyy += __shfl_xor(Vyy, 32);
            }
            if (kq == 0 && act)
                eval_out<PGH>(a, i, sj[ip], snf[ip], V, U * dsq, Vy * dsq, Z * dsq * dsq,
                              Uy * dsq * dsq, Vyy * dsq * dsq);
        }
    }
}

// FUSE = NBH: PW -> Taylor inside the wave (a.fused), from FP64 or (a.exp32)
// FP32 expansions
// fused launches: a.spw segments per wave over at most a.segmax segments
static unsigned fused_grid(const EvalArgs& a)
{
    if (a.spw < 1 || a.segmax < 1) fail(-1, "fused evaluation without a segment bound");
    return (unsigned)std::max(1, (a.segmax + 4 * a.spw - 1) / (4 * a.spw));
}

template <int TM, int FUSE, class E>
static void taylor_wave_f32_launch(const EvalArgs& a, hipStream_t s)
{
    const unsigned lds = FUSE ? 4 * PW2L_BUF(TM, FUSE) * sizeof(double) : 0;
    const unsigned grid = FUSE ? fused_grid(a) : (unsigned)a.grid;
    if (a.ifpgh >= 3) {
        allow_lds(k_taylor_wave_f32<TM, 3, FUSE, E>, lds + sizeof(float) * 8 * WSEG);
        klaunch(KID_TAYLOR_EVAL, k_taylor_wave_f32<TM, 3, FUSE, E>, grid, 256, lds, s, a);
    } else if (a.ifpgh == 2) {
        allow_lds(k_taylor_wave_f32<TM, 2, FUSE, E>, lds + sizeof(float) * 8 * WSEG);
        klaunch(KID_TAYLOR_EVAL, k_taylor_wave_f32<TM, 2, FUSE, E>, grid, 256, lds, s, a);
    } else {
        allow_lds(k_taylor_wave_f32<TM, 1, FUSE, E>, lds + sizeof(float) * 8 * WSEG);
        klaunch(KID_TAYLOR_EVAL, k_taylor_wave_f32<TM, 1, FUSE, E>, grid, 256, lds, s, a);
    }
}

template <int TM, int FUSE, class E>
static void taylor_wave_launch(const EvalArgs& a, hipStream_t s)
{
    const unsigned lds = FUSE ? 4 * PW2L_BUF(TM, FUSE) * sizeof(double) : 0;
    const unsigned grid = FUSE ? fused_grid(a) : (unsigned)a.grid;
    if (a.ifpgh >= 3) {
        allow_lds(k_taylor_wave<TM, 3, FUSE, E>, lds + sizeof(double) * 8 * WSEG);
        klaunch(KID_TAYLOR_EVAL, k_taylor_wave<TM, 3, FUSE, E>, grid, 256, lds, s, a);
    } else if (a.ifpgh == 2) {
        allow_lds(k_taylor_wave<TM, 2, FUSE, E>, lds + sizeof(double) * 8 * WSEG);
        klaunch(KID_TAYLOR_EVAL, k_taylor_wave<TM, 2, FUSE, E>, grid, 256, lds, s, a);
    } else {
        allow_lds(k_taylor_wave<TM, 1, FUSE, E>, lds + sizeof(double) * 8 * WSEG);
        klaunch(KID_TAYLOR_EVAL, k_taylor_wave<TM, 1, FUSE, E>, grid, 256, lds, s, a);
    }
}

// the wave kernels of one Taylor order tile count: unfused (Taylor
// coefficients from a.taylor) or fused with NBH = 1 (npw <= 32) or 2
// (npw <= 64)
template <int TM>
static void taylor_wave_any(const EvalArgs& a, hipStream_t s)
{
    const int nbh = (a.npw / 2 + 15) / 16;
    if (!a.fused) {
        if (a.tay32)
            taylor_wave_f32_launch<TM, 0, float>(a, s);
        else
            taylor_wave_launch<TM, 0, double>(a, s);
    } else if (a.tay32) {
        if (nbh == 1)
            taylor_wave_f32_launch<TM, 1, float>(a, s);
        else
            taylor_wave_f32_launch<TM, 2, float>(a, s);
    } else if (a.exp32) {
        fail(-3, "FP32 expansions with FP64 Taylor evaluation are not fused");
    } else if (nbh == 1) {
        taylor_wave_launch<TM, 1, double>(a, s);
    } else {
        taylor_wave_launch<TM, 2, double>(a, s);
    }
}

template <int TM>
static void taylor_mfma_launch(const EvalArgs& a, hipStream_t s)
{
    constexpr int MP = 16 * TM;
    const int nv = a.ifpgh >= 3 ? 3 : (a.ifpgh == 2 ? 2 : 1);
    size_t lds = (size_t)nv * MP * MP * sizeof(double);
    const int grid = a.tstart ? a.grid : a.nbox;
    if (a.ifpgh >= 3) {
        allow_lds(k_taylor_mfma<TM, 3>, lds);
        klaunch(KID_TAYLOR_EVAL, k_taylor_mfma<TM, 3>, grid, 256, lds, s, a);
    } else if (a.ifpgh == 2) {
        allow_lds(k_taylor_mfma<TM, 2>, lds);
        klaunch(KID_TAYLOR_EVAL, k_taylor_mfma<TM, 2>, grid, 256, lds, s, a);
    } else {
        allow_lds(k_taylor_mfma<TM, 1>, lds);
        klaunch(KID_TAYLOR_EVAL, k_taylor_mfma<TM, 1>, grid, 256, lds, s, a);
    }
}

void launch_taylor_eval(const EvalArgs& a, hipStream_t s)
{
    if (a.nbox <= 0) return;
    int m = a.nl + 1;
    if (m <= 32 && a.tstart && a.segbox && a.seg <= WSEG) {
        if (a.fused && a.npw > 64) fail(-3, "fused PW -> Taylor needs npw <= 64");
        if (!a.perm || !a.npot) fail(-1, "the segmented evaluation writes permuted outputs");
        if (m <= 16)
            taylor_wave_any<1>(a, s);
        else
            taylor_wave_any<2>(a, s);
        return;
    }
    if (a.tay32 || a.fused)
        fail(-3, "FP32 or fused Taylor coefficients need the segmented evaluation with order < 32");
    if (m <= 64) {
        switch ((m + 15) / 16) {
        case 1: taylor_mfma_launch<1>(a, s); break;
        case 2: taylor_mfma_launch<2>(a, s); break;
        case 3: taylor_mfma_launch<3>(a, s); break;
        default: taylor_mfma_launch<4>(a, s); break;
        }
        return;
    }
    size_t lds = (size_t)m * m * sizeof(double);
    const int grid = a.tstart ? a.grid : a.nbox;
    if (a.ifpgh >= 3)
        klaunch(KID_TAYLOR_EVAL, k_taylor_eval<3>, grid, 256, lds, s, a);
    else if (a.ifpgh == 2)
        klaunch(KID_TAYLOR_EVAL, k_taylor_eval<2>, grid, 256, lds, s, a);
    else
        klaunch(KID_TAYLOR_EVAL, k_taylor_eval<1>, grid, 256, lds, s, a);
}

// Step 5, plane-wave route (g2dpwevalp/g/h_vec, g2drouts.f:1887-2200), on
// FP64 matrix cores: pot = 2 Re sum_b e^{i ts_b y} sum_{a<half} PW(a, b)
// e^{i ts_a x}.  A leaf's expansion PW(a, b) (x nodes
// ts_a, a < half; y nodes ts_b, b < npw, ts_{b+half} = -ts_b) is staged in
// LDS as real and imaginary planes, zero padded to KP x NP.  For 16 points
// of a wavefront, U(p, b) = sum_a e^{i ts_a x_p} PW(a, b) takes four real
// 16x16x4 products per tile (lane l supplies the phase of point l%16 and
// node l/16 + 4 ks); gradients add the products with ts_a e^{...}, Hessians
// with ts_a^2 e^{...}.  Each 16-column tile of U goes through LDS to four
// lanes per point, which add Re[U(p, b) E_b] with E_b = e^{i ts_b y_p} and
// the y derivatives (factors i ts_b).
constexpr int PWM_KS = 13;  // half <= 52

template <int PGH, class E>
__global__ __launch_bounds__(256) void k_pw_eval_mfma(EvalArgs a)
{
    extern __shared__ double sm[];
    const int npw = a.npw, half = npw / 2, nexp = half * npw;
    const int KS = (half + 3) / 4, NT = (npw + 15) / 16, KP = 4 * KS, NP = 16 * NT;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    double* Pr = sm;
    double* Pi = sm + KP * NP;
    double* Us = sm + 2 * KP * NP + w * (PGH * 512);  // [product][re, im][16 points][16 columns]
    const double* ts = a.ts;
    const double dsq = a.dsq;
    // the x nodes of this lane's A-operand rows, a0 + 4 ks (0 past half:
    // those rows of PW are zero)
    const int a0 = lane >> 4;
    double fa[PWM_KS];
#pragma unroll
    for (int ks = 0; ks < PWM_KS; ++ks) fa[ks] = a0 + 4 * ks < half ? ts[a0 + 4 * ks] : 0.0;
    const int pq = lane >> 2, cq = lane & 3;  // y sums: point pq, columns 4 cq .. 4 cq + 3
    const int total = seg_count(a.tstart, a.nbox);
    for (int g = blockIdx.x; g < total; g += gridDim.x) {
        const SegRange sr = seg_at(a.tstart, a.nbox, a.boxes, a.se, a.seg, g, a.segbox);
        const int b = a.boxes[sr.li];
        const E* pw = (const E*)a.loc + 2 * ((long)a.slot[b] * a.nd + a.ind) * nexp;
        __syncthreads();
        for (int o = t; o < 2 * KP * NP; o += 256) sm[o] = 0.0;
        __syncthreads();
        for (int e = t; e < nexp; e += 256) {
            const int aa = e % half, bb = e / half;
            Pr[aa * NP + bb] = (double)pw[2 * e];
            Pi[aa * NP + bb] = (double)pw[2 * e + 1];
        }
        __syncthreads();
        const double cx = a.cen[2 * b], cy = a.cen[2 * b + 1];
        for (int base = sr.p0 + 16 * w; base < sr.p1; base += 64) {
            // x phases of point base + lane%16 (clamped into the segment;
            // rows past its end are computed and not written)
            double wr[PWM_KS], wi[PWM_KS];
            {
                const int pa = min(base + (lane & 15), sr.p1 - 1);
                const double x = (a.pts[2 * (long)pa] - cx) * dsq;
                double s, c, s4, c4;
                sincos(fa[0] * x, &s, &c);
                sincos(8.0 * ts[0] * x, &s4, &c4);  // e^{i 4h x}, 4h = 8 ts_0
#pragma unroll
                for (int ks = 0; ks < PWM_KS; ++ks) {
                    wr[ks] = c;
                    wi[ks] = s;
                    const double tmp = c * c4 - s * s4;
                    s = c * s4 + s * c4;
                    c = tmp;
                }
            }
            // y phase data of point pq
            const int py = min(base + pq, sr.p1 - 1);
            const double y = (a.pts[2 * (long)py + 1] - cy) * dsq;
            double sh, ch;
            sincos(2.0 * ts[0] * y, &sh, &ch);  // e^{i h y}
            double zr = 0, gx = 0, gy = 0, hxx = 0, hxy = 0, hyy = 0;
            for (int nt = 0; nt < NT; ++nt) {
                d4_t u[PGH][2];
#pragma unroll
                for (int q = 0; q < PGH; ++q) u[q][0] = u[q][1] = d4_t{0, 0, 0, 0};
                const double* brow = Pr + a0 * NP + (lane & 15) + 16 * nt;
                const double* bimg = Pi + a0 * NP + (lane & 15) + 16 * nt;
#pragma unroll
                for (int ks = 0; ks < PWM_KS; ++ks) {
                    if (ks < KS) {
                        const double br = brow[4 * ks * NP], bi = bimg[4 * ks * NP];
                        double xr = wr[ks], xi = wi[ks];
#pragma unroll
                        for (int q = 0; q < PGH; ++q) {
                            u[q][0] = mfma16x16x4(xr, br, u[q][0]);
                            u[q][0] = mfma16x16x4(-xi, bi, u[q][0]);
                            u[q][1] = mfma16x16x4(xr, bi, u[q][1]);
                            u[q][1] = mfma16x16x4(xi, br, u[q][1]);
                            xr *= fa[ks];
                            xi *= fa[ks];
                        }
                    }
                }
                // the tile to LDS: lane holds rows lane/16 + 4 r of column lane%16
#pragma unroll
                for (int q = 0; q < PGH; ++q)
#pragma unroll
                    for (int ri = 0; ri < 2; ++ri)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            Us[((q * 2 + ri) * 16 + (lane >> 4) + 4 * r) * 16 + (lane & 15)] = u[q][ri][r];
                wave_lds_sync();
                double er = 0, ei = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int col = 4 * cq + j, bc = 16 * nt + col;
                    if (bc >= npw) break;
                    const double tb = ts[bc];
                    if (j == 0 || bc == half) {
                        sincos(tb * y, &ei, &er);
                    } else {
                        // next node: +h below half, -h above
                        const double si = bc < half ? sh : -sh;
                        const double tmp = er * ch - ei * si;
                        ei = er * si + ei * ch;
                        er = tmp;
                    }
                    const int o = pq * 16 + col;
                    const double ur = Us[o], ui = Us[256 + o];
                    const double re = ur * er - ui * ei, im = ur * ei + ui * er;
                    zr += re;
                    if (PGH >= 2) {
                        const double vr = Us[512 + o], vi = Us[768 + o];
                        const double rex = vr * er - vi * ei, imx = vr * ei + vi * er;
                        gx -= imx;
                        gy -= tb * im;
                        if (PGH >= 3) {
                            const double xr2 = Us[1024 + o], xi2 = Us[1280 + o];
                            hxx -= xr2 * er - xi2 * ei;
                            hxy -= tb * rex;
                            hyy -= tb * tb * re;
                        }
                    }
                }
                wave_lds_sync();
            }
            // the point's four lanes
#pragma unroll
            for (int m = 1; m <= 2; m <<= 1) {
                zr += __shfl_xor(zr, m);
                if (PGH >= 2) {
                    gx += __shfl_xor(gx, m);
                    gy += __shfl_xor(gy, m);
                }
                if (PGH >= 3) {
                    hxx += __shfl_xor(hxx, m);
                    hxy += __shfl_xor(hxy, m);
                    hyy += __shfl_xor(hyy, m);
                }
            }
            const int i = base + pq;
            if (cq == 0 && i < sr.p1)
                eval_out<PGH>(a, i, 2 * zr, 2 * gx * dsq, 2 * gy * dsq, 2 * hxx * dsq * dsq,
                              2 * hxy * dsq * dsq, 2 * hyy * dsq * dsq);
        }
    }
}

template <int PGH, class E>
static void pw_eval_mfma_launch(const EvalArgs& a, hipStream_t s)
{
    const int half = a.npw / 2, KP = 4 * ((half + 3) / 4), NP = 16 * ((a.npw + 15) / 16);
    const size_t lds = (2 * (size_t)KP * NP + 4 * PGH * 512) * sizeof(double);
    allow_lds(k_pw_eval_mfma<PGH, E>, lds);
    const int grid = a.tstart ? a.grid : a.nbox;
    klaunch(KID_PW_EVAL, k_pw_eval_mfma<PGH, E>, grid, 256, lds, s, a);
}

void launch_pw_eval(const EvalArgs& a, hipStream_t s)
{
    if (a.nbox <= 0) return;
    // the reference's evaluation holds at most 100 nodes (g2dpwevalp_vec,
    // g2drouts.f:1921-1922)
    if (a.npw > 100) fail(-3, "plane-wave length above 100");
    if (a.exp32) {
        if (a.ifpgh >= 3) pw_eval_mfma_launch<3, float>(a, s);
        else if (a.ifpgh == 2) pw_eval_mfma_launch<2, float>(a, s);
        else pw_eval_mfma_launch<1, float>(a, s);
    } else if (a.ifpgh >= 3) pw_eval_mfma_launch<3, double>(a, s);
    else if (a.ifpgh == 2) pw_eval_mfma_launch<2, double>(a, s);
    else pw_eval_mfma_launch<1, double>(a, s);
}

// ---------------------------------------------------------------------------
// One source's contribution at a target, r = t - s, e = exp(-|r|^2/delta):
// a charge q gives q e; a dipole w = d n gives d (n . grad_s) e =
// (2/delta) (w . r) e (g2d_directd*_vec, g2ddirect.f:217-420).  With
// u = -2 r / delta and A = q - w . u the sum is e A, its gradient
// e (u A - rfac w) and its Hessian e ((u_a u_b + rfac d_ab) A
// - rfac (u_a w_b + u_b w_a)), rfac = -2/delta.
// ---------------------------------------------------------------------------
template <int PGH, bool DIP>
__device__ __forceinline__ void pair_add(double dx, double dy, double q, double w1, double w2,
                                         double inv, double rfac, double dmax, double& p,
                                         double& gx, double& gy, double& hxx, double& hxy,
                                         double& hyy)
{
    const double rr = dx * dx + dy * dy;
    if (!(rr < dmax)) return;
    const double e = exp(-rr * inv);
    const double ux = rfac * dx, uy = rfac * dy;
    const double A = DIP ? q - (w1 * ux + w2 * uy) : q;
    p += e * A;
    if (PGH >= 2) {
        gx += e * (ux * A - (DIP ? rfac * w1 : 0.0));
        gy += e * (uy * A - (DIP ? rfac * w2 : 0.0));
        if (PGH >= 3) {
            hxx += e * ((rfac + ux * ux) * A - (DIP ? 2.0 * rfac * ux * w1 : 0.0));
            hxy += e * (ux * uy * A - (DIP ? rfac * (ux * w2 + uy * w1) : 0.0));
            hyy += e * ((rfac + uy * uy) * A - (DIP ? 2.0 * rfac * uy * w2 : 0.0));
        }
    }
}

// ---------------------------------------------------------------------------
// Step 6, near field (fgt2dpart_direct_vec fgt2d.f:1354 -> g2d_direct*_vec,
// g2ddirect.f:32-655), target-owned: one block per target unit, looping over
// its source boxes in LDS tiles; pairs with r^2 >= dmax are skipped.
// ---------------------------------------------------------------------------
template <int PGH, bool DIP>
__global__ __launch_bounds__(256) void k_direct(DirectArgs a)
{
    __shared__ double sx[256], sy[256], sq[256], sw1[DIP ? 256 : 1], sw2[DIP ? 256 : 1];
    const int u = a.units[blockIdx.x];
    const int2 tr = a.tgt_se[u];
    const int l0 = a.list_ptr[blockIdx.x], l1 = a.list_ptr[blockIdx.x + 1];
    const double inv = 1.0 / a.delta, rfac = -2.0 / a.delta;
    const int t = threadIdx.x;
    for (int ind = 0; ind < a.nd; ++ind)
        for (int t0 = tr.x; t0 < tr.y; t0 += 256) {
            int i = t0 + t;
            bool act = i < tr.y;
            double x = act ? a.xt[2 * (long)i] : 0, y = act ? a.xt[2 * (long)i + 1] : 0;
            double p = 0, gx = 0, gy = 0, hxx = 0, hxy = 0, hyy = 0;
            for (int li = l0; li < l1; ++li) {
                int2 sr = a.src_se[a.list[li]];
                for (int s0 = sr.x; s0 < sr.y; s0 += 256) {
                    int cn = min(256, sr.y - s0);
                    __syncthreads();
                    if (t < cn) {
                        long j = s0 + t;
                        sx[t] = a.xs[2 * j];
                        sy[t] = a.xs[2 * j + 1];
                        sq[t] = a.qs[j * a.nd + ind];
                        if (DIP) {
                            sw1[t] = a.ws[2 * (j * a.nd + ind)];
                            sw2[t] = a.ws[2 * (j * a.nd + ind) + 1];
                        }
                    }
                    __syncthreads();
                    for (int s = 0; s < cn; ++s)
                        pair_add<PGH, DIP>(x - sx[s], y - sy[s], sq[s], DIP ? sw1[s] : 0.0,
                                           DIP ? sw2[s] : 0.0, inv, rfac, a.dmax, p, gx, gy, hxx,
                                           hxy, hyy);
                }
            }
            if (act) {
                a.pot[(long)i * a.nd + ind] += p;
                if (PGH >= 2) {
                    long g = (long)i * 2 * a.nd + ind;
                    a.grad[g] += gx;
                    a.grad[g + a.nd] += gy;
                }
                if (PGH >= 3) {
                    long h = (long)i * 3 * a.nd + ind;
                    a.hess[h] += hxx;
                    a.hess[h + a.nd] += hxy;
                    a.hess[h + 2 * a.nd] += hyy;
                }
            }
        }
}

void launch_direct(const DirectArgs& a, hipStream_t s)
{
    if (a.nunits <= 0) return;
    const bool dip = a.ws != nullptr;
    if (a.ifpgh >= 3)
        dip ? klaunch(KID_DIRECT, k_direct<3, true>, a.nunits, 256, 0, s, a) : klaunch(KID_DIRECT, k_direct<3, false>, a.nunits, 256, 0, s, a);
    else if (a.ifpgh == 2)
        dip ? klaunch(KID_DIRECT, k_direct<2, true>, a.nunits, 256, 0, s, a) : klaunch(KID_DIRECT, k_direct<2, false>, a.nunits, 256, 0, s, a);
    else
        dip ? klaunch(KID_DIRECT, k_direct<1, true>, a.nunits, 256, 0, s, a) : klaunch(KID_DIRECT, k_direct<1, false>, a.nunits, 256, 0, s, a);
}

// ---------------------------------------------------------------------------
// Large-delta branch: Taylor coefficients of sum_i q_i exp(-|t - s_i|^2/delta)
// about one centre (g2dformlc_vec, g2drouts.f:621-680),
//   L(j,k) = sum_i q_i h_j(xh_i) h_k(yh_i),
//   h_0 = e^{-u^2}, h_1 = 2u e^{-u^2}, h_{j+1} = 2 (u h_j - h_{j-1})/(j+1).
// Each block sums a slice of the sources; a second kernel adds the slices.
// ---------------------------------------------------------------------------
//
// Dipoles (g2dformld_vec / g2dformlcd_vec, g2drouts.f:685-845) add
// dh_j = -dsq (j+1) h_{j+1}: L(j,k) += dh_j(xh) w1 h_k(yh) + h_j(xh) w2 dh_k(yh).
template <bool DIP>
__global__ __launch_bounds__(256) void k_formlc_part(const double* __restrict__ xy,
                                                     const double* __restrict__ q,
                                                     const double* __restrict__ ws, int nd,
                                                     int ind, int n, double cx, double cy,
                                                     double dsq, int nl, int nparts,
                                                     double* __restrict__ part)
{
    extern __shared__ double sm[];
    const int m = nl + 1, mm = m * m, t = threadIdx.x;
    const int chunk = 64;
    double* X = sm;
    double* Y = sm + chunk * m;
    double* Xd = Y + chunk * m;   // dipoles: dh_j(xh)
    double* Yw = Xd + chunk * m;  // dipoles: w1 h_k(yh)
    long per = (n + nparts - 1) / nparts;
    long i0 = blockIdx.x * per, i1 = min((long)n, i0 + per);
    double* out = part + (long)blockIdx.x * mm;
    for (int e = t; e < mm; e += 256) out[e] = 0;
    for (long s0 = i0; s0 < i1; s0 += chunk) {
        int cn = (int)min((long)chunk, i1 - s0);
        __syncthreads();
        if (t < cn) {
            long i = s0 + t;
            double x = (xy[2 * i] - cx) * dsq, y = (xy[2 * i + 1] - cy) * dsq;
            double fx = exp(-x * x), fy = exp(-y * y), qq = q ? q[(long)nd * i + ind] : 0.0;
            double w1 = DIP ? ws[2 * ((long)nd * i + ind)] : 0.0;
            double w2 = DIP ? ws[2 * ((long)nd * i + ind) + 1] : 0.0;
            // h_j for j = 0..m (one past the table for the derivatives)
            double a0 = fx, a1 = 2.0 * x * fx, b0 = fy, b1 = 2.0 * y * fy;
            for (int j = 0; j < m; ++j) {
                // here a0 = h_j(x), a1 = h_{j+1}(x) (same for b, y)
                X[t * m + j] = a0;
                if (DIP) {
                    const double dha = -(a1 * dsq * (j + 1.0)), dhb = -(b1 * dsq * (j + 1.0));
                    Xd[t * m + j] = dha;
                    Y[t * m + j] = qq * b0 + w2 * dhb;
                    Yw[t * m + j] = w1 * b0;
                } else {
                    Y[t * m + j] = qq * b0;
                }
                const double a2 = 2.0 * (x * a1 - a0) / (j + 2), b2 = 2.0 * (y * b1 - b0) / (j + 2);
                a0 = a1; a1 = a2; b0 = b1; b1 = b2;
            }
        }
        __syncthreads();
        for (int e = t; e < mm; e += 256) {
            int j = e % m, k = e / m;
            double acc = 0;
            for (int s = 0; s < cn; ++s) {
                acc += Y[s * m + k] * X[s * m + j];
                if (DIP) acc += Yw[s * m + k] * Xd[s * m + j];
            }
            out[e] += acc;
        }
    }
}

__global__ __launch_bounds__(256) void k_sum_parts(const double* __restrict__ part, int nparts,
                                                   int len, double* __restrict__ out)
{
    int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= len) return;
    double acc = 0;
    for (int p = 0; p < nparts; ++p) acc += part[(long)p * len + e];
    out[e] = acc;
}

void launch_formlc(const double* xy, const double* q, const double* ws, int nd, int ind, int n,
                   double cx, double cy, double dsq, int nl, double* part, int nparts,
                   double* out, hipStream_t s)
{
    int m = nl + 1;
    size_t lds = (size_t)(ws ? 4 : 2) * 64 * m * sizeof(double);
    if (ws) {
        allow_lds(k_formlc_part<true>, lds);
        k_formlc_part<true><<<nparts, 256, lds, s>>>(xy, q, ws, nd, ind, n, cx, cy, dsq, nl, nparts,
                                                     part);
    } else {
        allow_lds(k_formlc_part<false>, lds);
        k_formlc_part<false><<<nparts, 256, lds, s>>>(xy, q, ws, nd, ind, n, cx, cy, dsq, nl,
                                                      nparts, part);
    }
    k_sum_parts<<<grid_for(m * m, 256), 256, 0, s>>>(part, nparts, m * m, out);
}

// fgt2dpart_direct_vec (fgt2d.f:1354): every target against every source
// with the r^2 < dmax cutoff, accumulating.  Source tiles in LDS.
template <int PGH, bool DIP>
__global__ __launch_bounds__(256) void k_direct_all(const double* __restrict__ src,
                                                    const double* __restrict__ q,
                                                    const double* __restrict__ ws, int nd, int ns,
                                                    const double* __restrict__ tg, int nt,
                                                    double delta, double dmax,
                                                    double* __restrict__ pot,
                                                    double* __restrict__ grad,
                                                    double* __restrict__ hess)
{
    __shared__ double sx[256], sy[256], sq[256], sw1[DIP ? 256 : 1], sw2[DIP ? 256 : 1];
    const int t = threadIdx.x;
    const int i = blockIdx.x * 256 + t;
    const bool act = i < nt;
    const double inv = 1.0 / delta, rfac = -2.0 / delta;
    double x = act ? tg[2 * (long)i] : 0, y = act ? tg[2 * (long)i + 1] : 0;
    for (int ind = 0; ind < nd; ++ind) {
        double p = 0, gx = 0, gy = 0, hxx = 0, hxy = 0, hyy = 0;
        for (int s0 = 0; s0 < ns; s0 += 256) {
            int cn = min(256, ns - s0);
            __syncthreads();
            if (t < cn) {
                long j = s0 + t;
                sx[t] = src[2 * j];
                sy[t] = src[2 * j + 1];
                sq[t] = q ? q[j * nd + ind] : 0.0;
                if (DIP) {
                    sw1[t] = ws[2 * (j * nd + ind)];
                    sw2[t] = ws[2 * (j * nd + ind) + 1];
                }
            }
            __syncthreads();
            for (int s = 0; s < cn; ++s)
                pair_add<PGH, DIP>(x - sx[s], y - sy[s], sq[s], DIP ? sw1[s] : 0.0,
                                   DIP ? sw2[s] : 0.0, inv, rfac, dmax, p, gx, gy, hxx, hxy, hyy);
        }
        if (act) {
            pot[(long)i * nd + ind] += p;
            if (PGH >= 2) {
                grad[(long)i * 2 * nd + ind] += gx;
                grad[(long)i * 2 * nd + nd + ind] += gy;
            }
            if (PGH >= 3) {
                hess[(long)i * 3 * nd + ind] += hxx;
                hess[(long)i * 3 * nd + nd + ind] += hxy;
                hess[(long)i * 3 * nd + 2 * nd + ind] += hyy;
            }
        }
    }
}

void launch_direct_all(const double* src, const double* q, const double* ws, int nd, int ns,
                       const double* tg, int nt, double delta, double dmax, int ifpgh,
                       double* pot, double* grad, double* hess, hipStream_t s)
{
    if (nt <= 0 || ns <= 0) return;
    unsigned g = grid_for(nt, 256);
#define FGT_DALL(P_, D_)                                                                       \
    k_direct_all<P_, D_><<<g, 256, 0, s>>>(src, q, ws, nd, ns, tg, nt, delta, dmax, pot, grad, hess)
    if (ifpgh >= 3) {
        if (ws) FGT_DALL(3, true); else FGT_DALL(3, false);
    } else if (ifpgh == 2) {
        if (ws) FGT_DALL(2, true); else FGT_DALL(2, false);
    } else {
        if (ws) FGT_DALL(1, true); else FGT_DALL(1, false);
    }
#undef FGT_DALL
}

}  // namespace fgtpw
