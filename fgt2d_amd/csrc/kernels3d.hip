// HIP kernels of the 3D plane-wave fast Gauss transform for gfx950.
//
// Layouts (HBM, tree order):
//   points      double[3*n]          x,y,z interleaved (the reference's (3,n))
//   densities   double[nd*n]
//   expansions  double[2*nexp*nd]    per slot, complex interleaved, index
//                                    a + npw*(b + npw*c): a, b all nodes on
//                                    x, y; c the first npw/2 nodes on z, as
//                                    pwexp(npw,npw,npw/2,nd) (g3drouts.f)
//   nodes       ascending, ts_a = (a - npw/2 + 1/2) h (g3drouts.f:2669-2690);
//               the mirror of node a is npw-1-a
//
// 256-thread blocks (4 waves).  FP64 throughout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "fgtpw_impl.hpp"

namespace fgtpw {

using u64 = unsigned long long;

static inline unsigned gridn(long n, int bs)
{
    long g = (n + bs - 1) / bs;
    return (unsigned)(g < 1 ? 1 : g);
}

// FP64 matrix cores (v_mfma_f64_16x16x4f64): D(16x16) += A(16x4) B(4x16) per
// wavefront; lane l supplies A[l%16][l/16] and B[l/16][l%16] and holds
// D[l/16 + 4r][l%16] in element r of the result.
typedef double d4_t3 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4_t3 mfma3(double a, double b, d4_t3 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// i^n (re + i im), exact (swaps and sign flips)
__device__ __forceinline__ double2 rot_ij(double re, double im, int n)
{
    const double x = (n & 1) ? -im : re, y = (n & 1) ? re : im;
    return (n & 2) ? make_double2(-x, -y) : make_double2(x, y);
}

// Block barrier that orders LDS only: global loads in flight (prefetches)
// are not waited for, as __syncthreads' full fence would.
__device__ __forceinline__ void lds_barrier3()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// grid of a grid-stride kernel over n items: at most 2^20 blocks
static inline unsigned gridcap(long n, int bs)
{
    long g = (n + bs - 1) / bs;
    return (unsigned)(g < 1 ? 1 : (g > (1L << 20) ? (1L << 20) : g));
}

// ---------------------------------------------------------------------------
// bounding box (pts_tree_boxsize0, pts_tree3d.f:1597-1680): mins and maxes
// of x, y, z over sources and targets, plus a non-finite flag
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bbox3(const double* __restrict__ a, int na,
                                               const double* __restrict__ b, int nb,
                                               double* __restrict__ part)
{
    __shared__ double s[7][256];
    double v[7] = {HUGE_VAL, -HUGE_VAL, HUGE_VAL, -HUGE_VAL, HUGE_VAL, -HUGE_VAL, 0};
    long n = (long)na + nb;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double* p = i < na ? a + 3 * i : b + 3 * (i - na);
        for (int d = 0; d < 3; ++d) {
            double x = p[d];
            if (!isfinite(x)) v[6] = 1;
            v[2 * d] = fmin(v[2 * d], x);
            v[2 * d + 1] = fmax(v[2 * d + 1], x);
        }
    }
    int t = threadIdx.x;
    for (int k = 0; k < 7; ++k) s[k][t] = v[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < 7; ++k)
                s[k][t] = (k < 6 && !(k & 1)) ? fmin(s[k][t], s[k][t + w]) : fmax(s[k][t], s[k][t + w]);
        __syncthreads();
    }
    if (t < 7) part[7 * blockIdx.x + t] = s[t][0];
}

__global__ __launch_bounds__(256) void k_bbox3_final(const double* __restrict__ part, int nblk,
                                                     double* __restrict__ out)
{
    __shared__ double s[7][256];
    int t = threadIdx.x;
    double v[7] = {HUGE_VAL, -HUGE_VAL, HUGE_VAL, -HUGE_VAL, HUGE_VAL, -HUGE_VAL, 0};
    for (int i = t; i < nblk; i += 256)
        for (int k = 0; k < 7; ++k)
            v[k] = (k < 6 && !(k & 1)) ? fmin(v[k], part[7 * i + k]) : fmax(v[k], part[7 * i + k]);
    for (int k = 0; k < 7; ++k) s[k][t] = v[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < 7; ++k)
                s[k][t] = (k < 6 && !(k & 1)) ? fmin(s[k][t], s[k][t + w]) : fmax(s[k][t], s[k][t + w]);
        __syncthreads();
    }
    if (t < 7) out[t] = s[t][0];
}

void launch_bbox3(const double* a, int na, const double* b, int nb, double* part, int nblk,
                  double* out, hipStream_t s)
{
    k_bbox3<<<nblk, 256, 0, s>>>(a, na, b, nb, part);
    k_bbox3_final<<<1, 256, 0, s>>>(part, nblk, out);
}

// ---------------------------------------------------------------------------
// full-depth 3-bit Morton keys with the reference's FP64 centre recursion and
// "p - c < 0" tests (sort_pts_to_children, pts_tree3d.f:662-854;
// tree_refine_boxes, tree_routs3d.f:28-46).  Digit = x + 2y + 4z.
// ---------------------------------------------------------------------------
template <class K>  // 32- or 64-bit keys
__global__ __launch_bounds__(256) void k_keys3d(const double* __restrict__ p, int n, double cx0,
                                                double cy0, double cz0, double bs0, int nlmax,
                                                K* __restrict__ key, unsigned* __restrict__ idx)
{
#pragma clang fp contract(off)
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double x = p[3 * (long)i], y = p[3 * (long)i + 1], z = p[3 * (long)i + 2];
    double cx = cx0, cy = cy0, cz = cz0, bs = bs0;
    K k = 0;
    for (int l = 0; l < nlmax; ++l) {
        bs = bs / 2;
        double h = bs / 2;
        int xb = !(x - cx < 0.0), yb = !(y - cy < 0.0), zb = !(z - cz < 0.0);
        k = (k << 3) | (K)(xb | (yb << 1) | (zb << 2));
        cx = cx + (xb ? h : -h);
        cy = cy + (yb ? h : -h);
        cz = cz + (zb ? h : -h);
    }
    key[i] = k;
    if (idx) idx[i] = (unsigned)i;
}

void launch_keys3d(const double* p, int n, const double* c0, double bs0, int nlmax, u64* key,
                   unsigned* idx, hipStream_t s)
{
    if (n <= 0) return;
    k_keys3d<u64><<<gridn(n, 256), 256, 0, s>>>(p, n, c0[0], c0[1], c0[2], bs0, nlmax, key, idx);
}
void launch_keys3d(const double* p, int n, const double* c0, double bs0, int nlmax,
                   unsigned* key, unsigned* idx, hipStream_t s)
{
    if (n <= 0) return;
    k_keys3d<unsigned><<<gridn(n, 256), 256, 0, s>>>(p, n, c0[0], c0[1], c0[2], bs0, nlmax, key,
                                                     idx);
}

__global__ __launch_bounds__(256) void k_gather3(const double* __restrict__ p,
                                                 const double* __restrict__ q, int nd,
                                                 const unsigned* __restrict__ perm, int n,
                                                 double* __restrict__ ps, double* __restrict__ qs)
{
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    long j = perm[i];
    ps[3 * (long)i] = p[3 * j];
    ps[3 * (long)i + 1] = p[3 * j + 1];
    ps[3 * (long)i + 2] = p[3 * j + 2];
    if (q)
        for (int d = 0; d < nd; ++d) qs[(long)nd * i + d] = q[(long)nd * j + d];
}

// dipole weights in sorted order: w(c, ind, i) = rnormal(c, perm i) dipstr(ind, perm i)
// at ws[3 (nd i + ind) + c] (the products of g3drouts.f:273-275)
__global__ __launch_bounds__(256) void k_gather_dip3(const double* __restrict__ rn,
                                                     const double* __restrict__ ds, int nd,
                                                     const unsigned* __restrict__ perm, int n,
                                                     double* __restrict__ ws)
{
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    long j = perm[i];
    const double n1 = rn[3 * j], n2 = rn[3 * j + 1], n3 = rn[3 * j + 2];
    for (int d = 0; d < nd; ++d) {
        const double dd = ds[(long)nd * j + d];
        double* w = ws + 3 * ((long)nd * i + d);
        w[0] = n1 * dd;
        w[1] = n2 * dd;
        w[2] = n3 * dd;
    }
}

void launch_gather_dip3(const double* rn, const double* ds, int nd, const unsigned* perm, int n,
                        double* ws, hipStream_t s)
{
    if (n <= 0) return;
    k_gather_dip3<<<gridn(n, 256), 256, 0, s>>>(rn, ds, nd, perm, n, ws);
}

void launch_gather3(const double* p, const double* q, int nd, const unsigned* perm, int n,
                    double* ps, double* qs, hipStream_t s)
{
    if (n <= 0) return;
    kmark(KID_GATHER, false, s);
    k_gather3<<<gridn(n, 256), 256, 0, s>>>(p, q, nd, perm, n, ps, qs);
    kmark(KID_GATHER, true, s);
}

// ---------------------------------------------------------------------------
// Step 1, Hermite route (g3dformhc_vec g3drouts.f:109-195 followed by
// g3dh2pw_real_vec :1565-1690).  With raw moments M(j,k,l) = sum q x^j y^k
// z^l (x = (s - c)/sqrt(delta)) and C(j,a) = ws_a (-i ts_a)^j / j! -- the
// reference's ws (-ts)^j/sqrt(j!) times its i^j grouping, times the
// 1/sqrt(j!) of its Hermite normalisation --
//     PW(a,b,c) = sum_{j,k,l} M(j,k,l) C(j,a) C(k,b) C(l,c).
// One block per leaf (grid-strided), four phases through a per-block slice of
// global scratch (L2-resident): moments; N_c(j,k) = sum_l M C(l,c) for every
// z node c of the first half; then per c, P(a,k) = sum_j C(j,a) N_c(j,k)
// (LDS) and PW(a,b,c) = sum_k C(k,b) P(a,k).
// ---------------------------------------------------------------------------
constexpr int F3_CHUNK = 32;  // sources staged per pass

// the dynamic LDS limit of a kernel, raised only when a launch needs more
// than the kernel's largest earlier size (the call costs host time)
template <class F>
static void allow_lds3(F* kernel, size_t bytes)
{
    if (bytes <= 65536) return;
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

// F3_MAXM: register tile of the moment accumulators (Hermite orders up to
// F3_MAXM - 1); instantiated for 32 and 64.  With dipoles (w = d n per
// source, g3dformhd_vec / g3dformhcd_vec, g3drouts.f:197-409) the raw moments
// gain the source derivatives:
//   M(j,k,l) += x^j [q y^k z^l + dsq (w2 k y^(k-1) z^l + w3 l y^k z^(l-1))]
//             + dsq j x^(j-1) w1 y^k z^l.
template <int F3_MAXM, bool DIP>
__global__ __launch_bounds__(256) void k_form3_hermite(Form3Args a)
{
    extern __shared__ double sm[];
    const int m = a.p + 1, npw = a.npw, half = npw / 2, t = threadIdx.x;
    const long nexp = (long)npw * npw * half;
    double* Mg = a.scratch + (long)blockIdx.x * a.sstride;    // m^3
    double* Ng = Mg + (long)m * m * m;                        // complex (half, m, m)
    for (int li = blockIdx.x; li < a.nbox; li += gridDim.x) {
        const int box = a.boxes[li];
        const int2 r = a.src_se[box];
        if (r.y - r.x <= a.thresh) continue;  // empty, or on the direct plane-wave route
        const double cx = a.cen[3 * box], cy = a.cen[3 * box + 1], cz = a.cen[3 * box + 2];
        // ---- moments: thread per (k,l), all j in registers ----
        double* xp = sm;                          // [F3_CHUNK][m]
        double* yp = sm + F3_CHUNK * m;
        double* zq = sm + 2 * F3_CHUNK * m;       // q z^l (dipoles: z^l)
        double* xd = sm + 3 * F3_CHUNK * m;       // dipoles: j x^(j-1)
        double* yd = sm + 4 * F3_CHUNK * m;       //          k y^(k-1)
        double* zd = sm + 5 * F3_CHUNK * m;       //          l z^(l-1)
        double* sw = sm + 6 * F3_CHUNK * m;       //          q, dsq w1, dsq w2, dsq w3
        for (int p0 = 0; p0 < m * m; p0 += 256) {
            const int pr = p0 + t, k = pr % m, l = pr / m;
            const bool act = pr < m * m;
            double acc[F3_MAXM];
#pragma unroll
            for (int j = 0; j < F3_MAXM; ++j) acc[j] = 0;
            for (int s0 = r.x; s0 < r.y; s0 += F3_CHUNK) {
                const int cn = min(F3_CHUNK, r.y - s0);
                __syncthreads();
                if (t < cn) {
                    long i = s0 + t;
                    double x = (a.xs[3 * i] - cx) * a.dsq, y = (a.xs[3 * i + 1] - cy) * a.dsq,
                           z = (a.xs[3 * i + 2] - cz) * a.dsq;
                    const double q = a.qs[(long)a.nd * i + a.ind];
                    double px = 1, py = 1, pz = DIP ? 1.0 : q;
                    double qx = 0, qy = 0, qz = 0;  // previous powers
                    for (int j = 0; j < m; ++j) {
                        xp[t * m + j] = px;
                        yp[t * m + j] = py;
                        zq[t * m + j] = pz;
                        if (DIP) {
                            xd[t * m + j] = j * qx;
                            yd[t * m + j] = j * qy;
                            zd[t * m + j] = j * qz;
                            qx = px;
                            qy = py;
                            qz = pz;
                        }
                        px *= x;
                        py *= y;
                        pz *= z;
                    }
                    if (DIP) {
                        const double* wv = a.dipw + 3 * ((long)a.nd * i + a.ind);
                        sw[4 * t] = q;
                        sw[4 * t + 1] = a.dsq * wv[0];
                        sw[4 * t + 2] = a.dsq * wv[1];
                        sw[4 * t + 3] = a.dsq * wv[2];
                    }
                }
                __syncthreads();
                if (act)
                    for (int s = 0; s < cn; ++s) {
                        const double* xr = xp + s * m;
                        if (!DIP) {
                            const double w = yp[s * m + k] * zq[s * m + l];
#pragma unroll
                            for (int j = 0; j < F3_MAXM; ++j)
                                if (j < m) acc[j] += w * xr[j];
                        } else {
                            const double yk = yp[s * m + k], zl = zq[s * m + l];
                            const double F = sw[4 * s] * yk * zl + sw[4 * s + 2] * yd[s * m + k] * zl +
                                             sw[4 * s + 3] * yk * zd[s * m + l];
                            const double G = sw[4 * s + 1] * yk * zl;
                            const double* xdr = xd + s * m;
#pragma unroll
                            for (int j = 0; j < F3_MAXM; ++j)
                                if (j < m) acc[j] += F * xr[j] + G * xdr[j];
                        }
                    }
            }
            if (act)
#pragma unroll
                for (int j = 0; j < F3_MAXM; ++j)
                    if (j < m) Mg[j + m * (k + m * l)] = acc[j];
        }
        __syncthreads();
        // ---- C table in LDS: complex (npw, m), index a + npw*j ----
        double* Cs = sm;
        for (int o = t; o < npw * m; o += 256) {
            Cs[2 * o] = a.coef[2 * (o % npw + a.cstride * (o / npw))];
            Cs[2 * o + 1] = a.coef[2 * (o % npw + a.cstride * (o / npw)) + 1];
        }
        __syncthreads();
        // ---- N_c(j,k) = sum_l M(j,k,l) C(l,c), c < half ----
        for (int pr = t; pr < m * m; pr += 256) {
            for (int c0 = 0; c0 < half; c0 += 8) {
                double nr[8], ni[8];
#pragma unroll
                for (int cc = 0; cc < 8; ++cc) nr[cc] = ni[cc] = 0;
                for (int l = 0; l < m; ++l) {
                    const double mv = Mg[pr + (long)m * m * l];
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc) {
                        if (c0 + cc < half) {
                            nr[cc] += mv * Cs[2 * (c0 + cc + npw * l)];
                            ni[cc] += mv * Cs[2 * (c0 + cc + npw * l) + 1];
                        }
                    }
                }
#pragma unroll
                for (int cc = 0; cc < 8; ++cc)
                    if (c0 + cc < half) {
                        Ng[2 * (pr + (long)m * m * (c0 + cc))] = nr[cc];
                        Ng[2 * (pr + (long)m * m * (c0 + cc)) + 1] = ni[cc];
                    }
            }
        }
        __syncthreads();
        double* Ps = Cs + 2 * npw * m;  // complex (npw, m): P(a,k)
        double* out = a.mp + 2 * ((long)a.slot[box] * a.nd + a.ind) * nexp;
        for (int c = 0; c < half; ++c) {
            const double* Nc = Ng + 2L * m * m * c;
            for (int o = t; o < npw * m; o += 256) {
                const int aa = o % npw, k = o / npw;
                double pr_ = 0, pi_ = 0;
                for (int j = 0; j < m; ++j) {
                    const double cr = Cs[2 * (aa + npw * j)], ci = Cs[2 * (aa + npw * j) + 1];
                    const double nr = Nc[2 * (j + m * k)], ni = Nc[2 * (j + m * k) + 1];
                    pr_ += cr * nr - ci * ni;
                    pi_ += cr * ni + ci * nr;
                }
                Ps[2 * o] = pr_;
                Ps[2 * o + 1] = pi_;
            }
            __syncthreads();
            for (int o = t; o < npw * npw; o += 256) {
                const int aa = o % npw, b = o / npw;
                double vr = 0, vi = 0;
                for (int k = 0; k < m; ++k) {
                    const double cr = Cs[2 * (b + npw * k)], ci = Cs[2 * (b + npw * k) + 1];
                    const double pr_ = Ps[2 * (aa + npw * k)], pi_ = Ps[2 * (aa + npw * k) + 1];
                    vr += cr * pr_ - ci * pi_;
                    vi += cr * pi_ + ci * pr_;
                }
                double* d = out + 2 * (aa + (long)npw * (b + (long)npw * c));
                d[0] = vr;
                d[1] = vi;
            }
            __syncthreads();
        }
    }
}

// Hermite-route formation on matrix cores (m = p + 1 <= 32, npw <= 16 NB),
// one block of 8 wavefronts per leaf, in four contractions:
//   1. M(j, k, l) = sum_s x^j W_s(k, l), W = q y^k z^l (+ dipole terms): rows
//      j, columns (k, l) at k + m l (KLM = m^2 rounded up to 16 columns),
//      sources as K; to per-block scratch M[(k + m l) + KLM j];
//   2. N(j, k, c) = sum_l M(j, k, l) C(l, c), computed transposed (rows c,
//      columns (j, k) at k + m j, KLM of them) into scratch N[(k + m j) + KLM c];
//   3. per z node c: P_c(a, k) = sum_j C(j, a) N(j, k, c), into LDS;
//   4. PW(a, b, c) = sum_k C(k, b) P_c(a, k), computed transposed (rows b,
//      columns a) so that the store runs along a.
// C(j, a) = ws_a (-i ts_a)^j / j! = i^j r(a, j) with r real, so every
// contraction runs on the real r (two real products per complex operand, not
// four): phase 2 takes the even orders l into Re N and the odd ones into Im N,
// phase 2 stores N' = i^j N and phase 3 stores P' = i^k P_c, which phases 3
// and 4 then contract with r.  r sits in LDS at a + CW j for j < JR (JR a
// multiple of 8 >= m); KP = 16 MT is m rounded up to 16; all padding is zero.
template <int NB, bool DIP>
__global__ __launch_bounds__(512) void k_form3_mfma(Form3Args a)
{
    constexpr int CW = 16 * NB + 1;   // LDS row stride of the C table
    extern __shared__ double sm[];
    const int m = a.p + 1, npw = a.npw, half = npw / 2, t = threadIdx.x;
    const int w = t >> 6, lane = t & 63, i16 = lane & 15, k4 = lane >> 4;
    const int MT = (m + 15) / 16, KP = 16 * MT, KS = (m + 3) / 4, PK = KP + 1;
    const int KSE = ((m + 1) / 2 + 3) / 4, JR = 8 * KSE;  // steps over the even (odd) orders
    const int KLP = KP * m;           // scratch sizing (KLM <= KLP)
    const int KLM = 16 * ((m * m + 15) / 16);  // columns (k, l) of M, rows (j, k) of N, packed
    const int NC = (half + 15) / 16;
    const long nexp = (long)npw * npw * half;
    double* Mg = a.scratch + (long)blockIdx.x * a.sstride;  // KP rows j x KLP
    double* Ng = Mg + (long)KP * KLP;                        // complex KLP x half
    double* Rt = sm;                                         // real r(a, j), j < JR
    double* work = Rt + CW * JR;                             // sources, then P_c
    for (int o = t; o < CW * JR; o += 512) {
        const int aa = o % CW, j = o / CW;
        double v = 0.0;
        if (aa < npw && j < m) {
            // C = i^j r: r is the one non-zero part of C, signed
            const double* c = a.coef + 2 * (aa + a.cstride * j);
            v = (j & 1) ? c[1] : c[0];
            if ((j & 3) >= 2) v = -v;
        }
        Rt[o] = v;
    }
    for (int li = blockIdx.x; li < a.nbox; li += gridDim.x) {
        const int box = a.boxes[li];
        const int2 r = a.src_se[box];
        if (r.y - r.x <= a.thresh) continue;  // empty, or on the direct plane-wave route
        const double cx = a.cen[3 * box], cy = a.cen[3 * box + 1], cz = a.cen[3 * box + 2];
        // ---- 1. moments: tiles (j tile mt, column tile nk) over the waves ----
        {
            double* xp = work;                 // [32 sources][KP] x^j
            double* yp = xp + 32 * KP;         // y^k
            double* zq = yp + 32 * KP;         // q z^l (dipoles: z^l)
            double* xd = zq + 32 * KP;         // dipoles: j x^(j-1), k y^(k-1), l z^(l-1)
            double* yd = xd + 32 * KP;
            double* zd = yd + 32 * KP;
            double* sw = zd + 32 * KP;         // dipoles: q, dsq w1, dsq w2, dsq w3
            const int ntile = MT * (KLM / 16);
            constexpr int TPW = 4;  // tiles per wave and pass
            for (int tp = 0; tp < ntile; tp += 8 * TPW) {
                d4_t3 acc[TPW];
#pragma unroll
                for (int i = 0; i < TPW; ++i) acc[i] = (d4_t3){0, 0, 0, 0};
                for (int s0 = r.x; s0 < r.y; s0 += 32) {
                    const int cn = min(32, r.y - s0);
                    __syncthreads();
                    {
                        // 16 threads per source, each a strided subset of the orders
                        const int sl = t & 31, j0 = t >> 5;
                        const bool in = sl < cn;
                        const long i = s0 + sl;
                        double x = 0, y = 0, z = 0, q = 0;
                        if (in) {
                            x = (a.xs[3 * i] - cx) * a.dsq;
                            y = (a.xs[3 * i + 1] - cy) * a.dsq;
                            z = (a.xs[3 * i + 2] - cz) * a.dsq;
                            q = a.qs[(long)a.nd * i + a.ind];
                        }
                        for (int j = j0; j < KP; j += 16) {
                            const bool jm = in && j < m;
                            double px = 1, py = 1, pz = DIP ? 1.0 : q, qx = 0, qy = 0, qz = 0;
                            for (int e = 0; e < j; ++e) {
                                qx = px;
                                qy = py;
                                qz = pz;
                                px *= x;
                                py *= y;
                                pz *= z;
                            }
                            xp[sl * KP + j] = jm ? px : 0.0;
                            yp[sl * KP + j] = jm ? py : 0.0;
                            zq[sl * KP + j] = jm ? pz : 0.0;
                            if (DIP) {
                                xd[sl * KP + j] = jm ? j * qx : 0.0;
                                yd[sl * KP + j] = jm ? j * qy : 0.0;
                                zd[sl * KP + j] = jm ? j * qz : 0.0;
                            }
                        }
                        if (DIP && j0 == 0) {
                            double w1 = 0, w2 = 0, w3 = 0;
                            if (in) {
                                const double* wv = a.dipw + 3 * ((long)a.nd * i + a.ind);
                                w1 = wv[0], w2 = wv[1], w3 = wv[2];
                            }
                            sw[4 * sl] = q;
                            sw[4 * sl + 1] = a.dsq * w1;
                            sw[4 * sl + 2] = a.dsq * w2;
                            sw[4 * sl + 3] = a.dsq * w3;
                        }
                    }
                    __syncthreads();
                    const int nks = (cn + 3) / 4;
#pragma unroll
                    for (int i = 0; i < TPW; ++i) {
                        const int tile = tp + w + 8 * i;
                        if (tile >= ntile) break;
                        const int mt = tile % MT, nk = tile / MT;
                        // columns past m^2 read l >= m: zero rows of the tables
                        const int j = 16 * mt + i16, kl = 16 * nk + i16, k = kl % m, l = kl / m;
                        for (int ks = 0; ks < nks; ++ks) {
                            const int sa = 4 * ks + k4;  // this lane's source in the chunk
                            const double av = xp[sa * KP + j];
                            if (!DIP) {
                                const double bv = yp[sa * KP + k] * zq[sa * KP + l];
                                acc[i] = mfma3(av, bv, acc[i]);
                            } else {
                                const double yk = yp[sa * KP + k], zl = zq[sa * KP + l];
                                const double F = sw[4 * sa] * yk * zl +
                                                 sw[4 * sa + 2] * yd[sa * KP + k] * zl +
                                                 sw[4 * sa + 3] * yk * zd[sa * KP + l];
                                const double G = sw[4 * sa + 1] * yk * zl;
                                acc[i] = mfma3(av, F, acc[i]);
                                acc[i] = mfma3(xd[sa * KP + j], G, acc[i]);
                            }
                        }
                    }
                }
#pragma unroll
                for (int i = 0; i < TPW; ++i) {
                    const int tile = tp + w + 8 * i;
                    if (tile >= ntile) break;
                    const int mt = tile % MT, nk = tile / MT;
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr)
                        Mg[(16 * nk + i16) + (long)KLM * (16 * mt + k4 + 4 * rr)] = acc[i][rr];
                }
            }
        }
        __syncthreads();
        // ---- 2. N'^T(c, (j, k)) = i^j sum_l C(l, c) M(j, k, l):
        // Re N = sum over even l = 2e of (-1)^e r(c, l) M, Im N over odd
        // l = 2e + 1 of (-1)^e r(c, l) M ----
        {
            const int ntile = NC * (KLM / 16);
            for (int tile = w; tile < ntile; tile += 8) {
                const int nc = tile % NC, nr = tile / NC;
                const int rho = 16 * nr + i16, j = rho / m, k = rho % m;  // rho >= m^2: j >= m
                d4_t3 nr_ = {0, 0, 0, 0}, ni_ = {0, 0, 0, 0};
                for (int ks = 0; ks < KSE; ++ks) {
                    const int e = 4 * ks + k4, le = 2 * e, lo = 2 * e + 1;
                    const double sg = (e & 1) ? -1.0 : 1.0;
                    const double ce = sg * Rt[(16 * nc + i16) + CW * le];
                    const double co = sg * Rt[(16 * nc + i16) + CW * lo];
                    const double me = le < m && j < m ? Mg[(k + m * le) + (long)KLM * j] : 0.0;
                    const double mo = lo < m && j < m ? Mg[(k + m * lo) + (long)KLM * j] : 0.0;
                    nr_ = mfma3(ce, me, nr_);
                    ni_ = mfma3(co, mo, ni_);
                }
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int c = 16 * nc + k4 + 4 * rr;
                    if (c < half) {
                        double* d = Ng + 2 * ((16 * nr + i16) + (long)KLM * c);
                        const double2 v = rot_ij(nr_[rr], ni_[rr], j);
                        d[0] = v.x;
                        d[1] = v.y;
                    }
                }
            }
        }
        __syncthreads();
        // ---- 3, 4 per group of G z nodes; P_c double-buffered in LDS at
        // k + PK a, G buffers per parity ----
        double* out = a.mp + 2 * ((long)a.slot[box] * a.nd + a.ind) * nexp;
        constexpr int G = NB * NB >= 8 ? 1 : 8 / (NB * NB);
        for (int c0 = 0, par = 0; c0 < half; c0 += G, par ^= 1) {
            for (int tile = w; tile < G * NB * MT; tile += 8) {
                const int g = tile / (NB * MT), na = tile % NB, mt = (tile / NB) % MT;
                const int c = c0 + g;
                if (c >= half) continue;
                double* Pc = work + (par * G + g) * 2 * PK * 16 * NB;
                const double* Nc = Ng + 2L * KLM * c;
                d4_t3 pr = {0, 0, 0, 0}, pi = {0, 0, 0, 0};
                for (int ks = 0; ks < KS; ++ks) {
                    const int j = 4 * ks + k4;
                    const double cr = Rt[(16 * na + i16) + CW * j];
                    double nvr = 0, nvi = 0;
                    if (j < m && 16 * mt + i16 < m) {
                        const double* nv = Nc + 2 * ((16 * mt + i16) + m * j);
                        nvr = nv[0];
                        nvi = nv[1];
                    }
                    pr = mfma3(cr, nvr, pr);
                    pi = mfma3(cr, nvi, pi);
                }
                // P' = i^k P_c, k = 16 mt + i16
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int o = (16 * mt + i16) + PK * (16 * na + k4 + 4 * rr);
                    const double2 v = rot_ij(pr[rr], pi[rr], i16);
                    Pc[2 * o] = v.x;
                    Pc[2 * o + 1] = v.y;
                }
            }
            __syncthreads();
            for (int tile = w; tile < G * NB * NB; tile += 8) {
                const int g = tile / (NB * NB), nb = tile % NB, na = (tile / NB) % NB;
                const int c = c0 + g;
                if (c >= half) continue;
                const double* Pc = work + (par * G + g) * 2 * PK * 16 * NB;
                d4_t3 vr = {0, 0, 0, 0}, vi = {0, 0, 0, 0};
                for (int ks = 0; ks < KS; ++ks) {
                    const int k = 4 * ks + k4;
                    const int op = k + PK * (16 * na + i16);
                    const double cr = Rt[(16 * nb + i16) + CW * k];
                    const double xr = Pc[2 * op], xi = Pc[2 * op + 1];
                    vr = mfma3(cr, xr, vr);
                    vi = mfma3(cr, xi, vi);
                }
                const int aa = 16 * na + i16;
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int b = 16 * nb + k4 + 4 * rr;
                    if (aa < npw && b < npw) {
                        double* d = out + 2 * (aa + (long)npw * (b + (long)npw * c));
                        d[0] = vr[rr];
                        d[1] = vi[rr];
                    }
                }
            }
        }
        __syncthreads();  // the next leaf's staging reuses the P_c space
    }
}

// per-block scratch of the matrix-core formation, doubles
long form3_mfma_scratch(int m, int npw)
{
    const long KP = 16 * ((m + 15) / 16);
    return KP * KP * m + 2 * KP * m * (npw / 2);
}

template <int NB, bool DIP>
static void form3_mfma_launch(const Form3Args& a, hipStream_t s)
{
    const int m = a.p + 1, KP = 16 * ((m + 15) / 16), JR = 8 * (((m + 1) / 2 + 3) / 4);
    const size_t ct = (size_t)(16 * NB + 1) * JR;
    const size_t src = (size_t)(DIP ? 6 : 3) * 32 * KP + (DIP ? 4 * 32 : 0);
    constexpr int G = NB * NB >= 8 ? 1 : 8 / (NB * NB);
    const size_t pc = (size_t)2 * G * 2 * (KP + 1) * 16 * NB;
    const size_t lds = (ct + std::max(src, pc)) * sizeof(double);
    allow_lds3(k_form3_mfma<NB, DIP>, lds);
    k_form3_mfma<NB, DIP><<<a.grid, 512, lds, s>>>(a);
}

void launch_form3_hermite(const Form3Args& a, hipStream_t s)
{
    if (a.nbox <= 0) return;
    const int m = a.p + 1;
    if (m > 64) fail(-3, "3D Hermite order above 63 is not supported");
    const bool dip = a.dipw != nullptr;
    if (m <= 32 && a.npw <= 48) {
        if (a.sstride < form3_mfma_scratch(m, a.npw)) fail(-3, "formation scratch too small");
        kmark(KID_FORM_MOMENTS, false, s);
        if (a.npw <= 32) {
            if (dip) form3_mfma_launch<2, true>(a, s); else form3_mfma_launch<2, false>(a, s);
        } else {
            if (dip) form3_mfma_launch<3, true>(a, s); else form3_mfma_launch<3, false>(a, s);
        }
        kmark(KID_FORM_MOMENTS, true, s);
        return;
    }
    size_t lds = std::max((size_t)(dip ? 6 : 3) * F3_CHUNK * m + (dip ? 4 * F3_CHUNK : 0),
                          (size_t)4 * a.npw * m) * sizeof(double);
    kmark(KID_FORM_MOMENTS, false, s);
#define FGT_F3H(M_, D_)                                                                        \
    allow_lds3(k_form3_hermite<M_, D_>, lds);                                                  \
    k_form3_hermite<M_, D_><<<a.grid, 256, lds, s>>>(a)
    if (m <= 32) {
        if (dip) { FGT_F3H(32, true); } else { FGT_F3H(32, false); }
    } else {
        if (dip) { FGT_F3H(64, true); } else { FGT_F3H(64, false); }
    }
#undef FGT_F3H
    kmark(KID_FORM_MOMENTS, true, s);
}

// Step 1, direct plane-wave route (g3dformpwc_vec, g3drouts.f:1950-2040):
// PW(a,b,c) = sum_s q ws_a ws_b ws_c e^{-i(ts_a x + ts_b y + ts_c z)}.  One
// block per leaf; thread per (a,b), z nodes of the first half in chunks of 8.
// With dipoles (g3dformpwd_vec / g3dformpwcd_vec, g3drouts.f:2042-2266) the
// factor q becomes q - i dsq (w1 ts_a + w2 ts_b + w3 ts_c): ez carries
// q - i dsq w3 ts_c, ez2 the bare z factor, and each thread adds
// -i dsq (w1 ts_a + w2 ts_b) ex ey ez2.
constexpr int P3_CHUNK = 16;

template <bool DIP>
__global__ __launch_bounds__(256) void k_form3_pw(Form3Args a)
{
    extern __shared__ double sm[];
    const int npw = a.npw, half = npw / 2, t = threadIdx.x;
    const long nexp = (long)npw * npw * half;
    double* ex = sm;                       // [chunk][npw] complex
    double* ey = sm + 2 * P3_CHUNK * npw;
    double* ez = sm + 4 * P3_CHUNK * npw;  // [chunk][half] complex, q folded in
    double* ez2 = ez + 2 * P3_CHUNK * half;   // dipoles: [chunk][half] complex, bare
    double* sw = ez2 + 2 * P3_CHUNK * half;   // dipoles: dsq w1, dsq w2 per source
    for (int li = blockIdx.x; li < a.nbox; li += gridDim.x) {
        const int box = a.boxes[li];
        const int2 r = a.src_se[box];
        if (r.y <= r.x || r.y - r.x > a.thresh) continue;  // empty, or on the Hermite route
        const double c3[3] = {a.cen[3 * box], a.cen[3 * box + 1], a.cen[3 * box + 2]};
        double* out = a.mp + 2 * ((long)a.slot[box] * a.nd + a.ind) * nexp;
        for (int p0 = 0; p0 < npw * npw; p0 += 256) {
            const int pr = p0 + t, aa = pr % npw, b = pr / npw;
            const bool act = pr < npw * npw;
            for (int c0 = 0; c0 < half; c0 += 8) {
                double vr[8], vi[8];
#pragma unroll
                for (int cc = 0; cc < 8; ++cc) vr[cc] = vi[cc] = 0;
                for (int s0 = r.x; s0 < r.y; s0 += P3_CHUNK) {
                    const int cn = min(P3_CHUNK, r.y - s0);
                    __syncthreads();
                    for (int o = t; o < cn * npw; o += 256) {
                        const int s = o / npw, node = o % npw;
                        const long i = s0 + s;
                        const double q = a.qs[(long)a.nd * i + a.ind];
                        const double* wv = DIP ? a.dipw + 3 * ((long)a.nd * i + a.ind) : nullptr;
                        for (int d = 0; d < 3; ++d) {
                            if (d == 2 && node >= half) continue;
                            const double u = (a.xs[3 * i + d] - c3[d]) * a.dsq;
                            double sn, cs;
                            sincos(-a.ts[node] * u, &sn, &cs);
                            if (d < 2) {
                                double* e = (d == 0 ? ex : ey) + 2 * (s * npw + node);
                                e[0] = cs * a.ws[node];
                                e[1] = sn * a.ws[node];
                            } else if (!DIP) {
                                const double w = a.ws[node] * q;
                                ez[2 * (s * half + node)] = cs * w;
                                ez[2 * (s * half + node) + 1] = sn * w;
                            } else {
                                // (cs + i sn) ws (q - i g), g = dsq w3 ts_c
                                const double er = cs * a.ws[node], ei = sn * a.ws[node];
                                const double g = a.dsq * wv[2] * a.ts[node];
                                ez[2 * (s * half + node)] = er * q + ei * g;
                                ez[2 * (s * half + node) + 1] = ei * q - er * g;
                                ez2[2 * (s * half + node)] = er;
                                ez2[2 * (s * half + node) + 1] = ei;
                            }
                        }
                        if (DIP && node == 0) {
                            sw[2 * s] = a.dsq * wv[0];
                            sw[2 * s + 1] = a.dsq * wv[1];
                        }
                    }
                    __syncthreads();
                    if (act)
                        for (int s = 0; s < cn; ++s) {
                            const double xr = ex[2 * (s * npw + aa)], xi = ex[2 * (s * npw + aa) + 1];
                            const double yr = ey[2 * (s * npw + b)], yi = ey[2 * (s * npw + b) + 1];
                            const double tr = xr * yr - xi * yi, ti = xr * yi + xi * yr;
                            // dipoles: tg = t (-i gam), gam = dsq (w1 ts_a + w2 ts_b)
                            double gr = 0, gi = 0;
                            if (DIP) {
                                const double gam = sw[2 * s] * a.ts[aa] + sw[2 * s + 1] * a.ts[b];
                                gr = ti * gam;
                                gi = -tr * gam;
                            }
#pragma unroll
                            for (int cc = 0; cc < 8; ++cc)
                                if (c0 + cc < half) {
                                    const double zr = ez[2 * (s * half + c0 + cc)],
                                                 zi = ez[2 * (s * half + c0 + cc) + 1];
                                    vr[cc] += tr * zr - ti * zi;
                                    vi[cc] += tr * zi + ti * zr;
                                    if (DIP) {
                                        const double z2r = ez2[2 * (s * half + c0 + cc)],
                                                     z2i = ez2[2 * (s * half + c0 + cc) + 1];
                                        vr[cc] += gr * z2r - gi * z2i;
                                        vi[cc] += gr * z2i + gi * z2r;
                                    }
                                }
                        }
                }
                if (act)
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc)
                        if (c0 + cc < half) {
                            double* d = out + 2 * (aa + (long)npw * (b + (long)npw * (c0 + cc)));
                            d[0] = vr[cc];
                            d[1] = vi[cc];
                        }
            }
        }
    }
}

void launch_form3_pw(const Form3Args& a, hipStream_t s)
{
    if (a.nbox <= 0) return;
    const bool dip = a.dipw != nullptr;
    size_t lds = ((size_t)2 * P3_CHUNK * (2 * a.npw + (dip ? 2 : 1) * (a.npw / 2)) +
                  (dip ? 2 * P3_CHUNK : 0)) * sizeof(double);
    kmark(KID_FORM_VALU, false, s);
    if (dip) {
        allow_lds3(k_form3_pw<true>, lds);
        k_form3_pw<true><<<a.grid, 256, lds, s>>>(a);
    } else {
        allow_lds3(k_form3_pw<false>, lds);
        k_form3_pw<false><<<a.grid, 256, lds, s>>>(a);
    }
    kmark(KID_FORM_VALU, true, s);
}

// ---------------------------------------------------------------------------
// Steps 2-4: diagonal plane-wave translations (g3dshiftpw_vec,
// g3drouts.f:2844-2882).  Octant of a shift by its sign pattern, fgt3d.f:1120
// -1136: 0 (+,+,+), 1 (+,-,+), 2 (-,+,+), 3 (-,-,+), 4..7 the same with z<0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int octant(double dx, double dy, double dz)
{
    int k = dx > 0 ? (dy > 0 ? 0 : 1) : (dy > 0 ? 2 : 3);
    return dz > 0 ? k : k + 4;
}

// The shift kernels run one block per (box, chunk of SH3_E coefficients),
// chunk-major: the box's indices, octants and slots are block-uniform
// (scalar loads, once per block), each thread keeps SH3_V independent 16-byte
// loads per operand in flight, and consecutive blocks take neighbouring boxes
// at the same chunk, so expansions they share are read from cache.
constexpr int SH3_V = 4, SH3_E = 256 * SH3_V;

// a parent's expansion is the sum of its children's, shifted: written, not
// accumulated, so the slot needs no clearing beforehand
__global__ __launch_bounds__(256) void k_merge3(const int* __restrict__ parents, int np,
                                                const int* __restrict__ child,
                                                const double* __restrict__ cen,
                                                const int2* __restrict__ src_se,
                                                const int* __restrict__ slot,
                                                const double* __restrict__ wms, int kbase, int nd,
                                                long nexp, double* __restrict__ mp)
{
    const long nchunk = (nexp + SH3_E - 1) / SH3_E;
    const double2* w2 = reinterpret_cast<const double2*>(wms);
    double2* m2 = reinterpret_cast<double2*>(mp);
    for (long q = blockIdx.x; q < (long)np * nchunk; q += gridDim.x) {
        // parent-major: a parent's children share nothing with other
        // parents', so their expansions stream chunk after chunk
        const int p = parents[q / nchunk];
        const long e0 = (q % nchunk) * SH3_E + threadIdx.x;
        for (int ind = 0; ind < nd; ++ind) {
            double re[SH3_V], im[SH3_V];
#pragma unroll
            for (int v = 0; v < SH3_V; ++v) re[v] = im[v] = 0;
            for (int j = 0; j < 8; ++j) {
                const int c = child[8 * p + j];
                if (c < 0) continue;
                const int2 r = src_se[c];
                if (r.y - r.x <= 0) continue;
                const int k = octant(cen[3 * p] - cen[3 * c], cen[3 * p + 1] - cen[3 * c + 1],
                                     cen[3 * p + 2] - cen[3 * c + 2]);
                const double2* w = w2 + (long)(kbase + k) * nexp;
                const double2* x = m2 + ((long)slot[c] * nd + ind) * nexp;
                double2 xv[SH3_V], wv[SH3_V];
#pragma unroll
                for (int v = 0; v < SH3_V; ++v) {
                    const long e = e0 + 256 * v;
                    xv[v] = e < nexp ? x[e] : make_double2(0.0, 0.0);
                    wv[v] = e < nexp ? w[e] : make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int v = 0; v < SH3_V; ++v) {
                    re[v] += xv[v].x * wv[v].x - xv[v].y * wv[v].y;
                    im[v] += xv[v].x * wv[v].y + xv[v].y * wv[v].x;
                }
            }
            double2* dst = m2 + ((long)slot[p] * nd + ind) * nexp;
#pragma unroll
            for (int v = 0; v < SH3_V; ++v) {
                const long e = e0 + 256 * v;
                if (e < nexp) dst[e] = make_double2(re[v], im[v]);
            }
        }
    }
}

void launch_merge3(const int* parents, int np, const int* child8, const double* cen,
                   const int2* src_se, const int* slot, const double* wms, int kbase, int nd,
                   long nexp, double* mp, hipStream_t s)
{
    if (np <= 0) return;
    kmark(KID_MERGE, false, s);
    k_merge3<<<gridcap((long)np * ((nexp + SH3_E - 1) / SH3_E), 1), 256, 0, s>>>(
        parents, np, child8, cen, src_se, slot, wms, kbase, nd, nexp, mp);
    kmark(KID_MERGE, true, s);
}

// Step 3 (fgt3d.f:1162-1210): at the cutoff level, loc = sum over colleagues
// j != self that have children (gt3d_computelistpw, tree_routs3d.f:695) of
// mp_j shifted by (jx, jy, jz) boxes, then + mp_self; boxes without
// evaluation points get a zero expansion.  The shift factor of coefficient
// (a, b, c) is the product of the per-axis factors wm1(a, jx) wm1(b, jy)
// wm1(c, jz), formed from an LDS copy of wm1: only the colleagues'
// expansions are read from memory.  Multipole and local expansions
// have separate slot tables; a box without a multipole slot (slot_mp < 0)
// has a zero multipole expansion.
__global__ __launch_bounds__(256) void k_m2l3(const int* __restrict__ boxes, int nb,
                                              const double* __restrict__ cen,
                                              const int* __restrict__ ncol,
                                              const int* __restrict__ col,
                                              const int* __restrict__ nch,
                                              const int2* __restrict__ src_se,
                                              const int2* __restrict__ trg_se, int use_src,
                                              int use_trg, const int* __restrict__ slot_mp,
                                              const int* __restrict__ slot_loc,
                                              const double* __restrict__ wm1, int npw,
                                              double xmin, int do_list, int nd, long nexp,
                                              const double* __restrict__ mp,
                                              double* __restrict__ loc)
{
    extern __shared__ double2 f1[];  // complex (3, npw): offset -1, 0, +1 at 3 a + k + 1
    for (int o = threadIdx.x; o < 3 * npw; o += 256) f1[o] = reinterpret_cast<const double2*>(wm1)[o];
    __syncthreads();
    const long nchunk = (nexp + SH3_E - 1) / SH3_E;
    const double2* m2 = reinterpret_cast<const double2*>(mp);
    double2* l2 = reinterpret_cast<double2*>(loc);
    for (long q = blockIdx.x; q < (long)nb * nchunk; q += gridDim.x) {
        const int b = boxes[q % nb];
        const long e0 = (q / nb) * SH3_E + threadIdx.x;
        int npts = 0;
        if (use_trg) npts += trg_se[b].y - trg_se[b].x;
        if (use_src) npts += src_se[b].y - src_se[b].x;
        // (a, b, c) of this thread's coefficients, as LDS row offsets
        int ra[SH3_V], rb[SH3_V], rc[SH3_V];
#pragma unroll
        for (int v = 0; v < SH3_V; ++v) {
            const long e = e0 + 256 * v;
            const int ee = e < nexp ? (int)e : 0, bc = ee / npw;
            ra[v] = 3 * (ee - bc * npw) + 1;
            rb[v] = 3 * (bc % npw) + 1;
            rc[v] = 3 * (bc / npw) + 1;
        }
        for (int ind = 0; ind < nd; ++ind) {
            double re[SH3_V], im[SH3_V];
#pragma unroll
            for (int v = 0; v < SH3_V; ++v) re[v] = im[v] = 0;
            if (npts > 0 && do_list)
                for (int i = 0; i < ncol[b]; ++i) {
                    const int j = col[27 * b + i];
                    if (j == b || nch[j] <= 0 || slot_mp[j] < 0) continue;
                    const int jx = (int)rint((cen[3 * b] - cen[3 * j]) / xmin);
                    const int jy = (int)rint((cen[3 * b + 1] - cen[3 * j + 1]) / xmin);
                    const int jz = (int)rint((cen[3 * b + 2] - cen[3 * j + 2]) / xmin);
                    const double2* x = m2 + ((long)slot_mp[j] * nd + ind) * nexp;
                    double2 xv[SH3_V];
#pragma unroll
                    for (int v = 0; v < SH3_V; ++v) {
                        const long e = e0 + 256 * v;
                        xv[v] = e < nexp ? x[e] : make_double2(0.0, 0.0);
                    }
#pragma unroll
                    for (int v = 0; v < SH3_V; ++v) {
                        const double2 fa = f1[ra[v] + jx], fb = f1[rb[v] + jy], fc = f1[rc[v] + jz];
                        const double tr = fa.x * fb.x - fa.y * fb.y, ti = fa.x * fb.y + fa.y * fb.x;
                        const double wr = tr * fc.x - ti * fc.y, wi = tr * fc.y + ti * fc.x;
                        re[v] += xv[v].x * wr - xv[v].y * wi;
                        im[v] += xv[v].x * wi + xv[v].y * wr;
                    }
                }
            if (npts > 0 && slot_mp[b] >= 0) {
                const double2* x = m2 + ((long)slot_mp[b] * nd + ind) * nexp;
#pragma unroll
                for (int v = 0; v < SH3_V; ++v) {
                    const long e = e0 + 256 * v;
                    if (e < nexp) {
                        const double2 xv = x[e];
                        re[v] += xv.x;
                        im[v] += xv.y;
                    }
                }
            }
            double2* dst = l2 + ((long)slot_loc[b] * nd + ind) * nexp;
#pragma unroll
            for (int v = 0; v < SH3_V; ++v) {
                const long e = e0 + 256 * v;
                if (e < nexp) dst[e] = make_double2(re[v], im[v]);
            }
        }
    }
}

void launch_m2l3(const int* boxes, int nb, const double* cen, const int* ncol, const int* col27,
                 const int* nch, const int2* src_se, const int2* trg_se, int use_src, int use_trg,
                 const int* slot_mp, const int* slot_loc, const double* wm1, int npw, double xmin,
                 int do_list, int nd, long nexp, const double* mp, double* loc, hipStream_t s)
{
    if (nb <= 0) return;
    const size_t lds = (size_t)3 * npw * sizeof(double2);
    allow_lds3(k_m2l3, lds);
    kmark(KID_M2L, false, s);
    k_m2l3<<<gridcap((long)nb * ((nexp + SH3_E - 1) / SH3_E), 1), 256, lds, s>>>(
        boxes, nb, cen, ncol, col27, nch, src_se, trg_se, use_src, use_trg, slot_mp, slot_loc, wm1,
        npw, xmin, do_list, nd, nexp, mp, loc);
    kmark(KID_M2L, true, s);
}

// Step 4 (fgt3d.f:1215-1276): loc(child) = loc(parent) shifted to the child
// (the only contribution below the cutoff level, so written, not
// accumulated).  Children without evaluation points in their box -- and so
// in their subtree -- are skipped: nothing reads their expansion.
__global__ __launch_bounds__(256) void k_split3(const int* __restrict__ children, int nc,
                                                const int* __restrict__ par,
                                                const double* __restrict__ cen,
                                                const int* __restrict__ slot,
                                                const double* __restrict__ wms, int kbase, int nd,
                                                long nexp, double* __restrict__ loc,
                                                const int2* __restrict__ src_se,
                                                const int2* __restrict__ trg_se, int use_src,
                                                int use_trg)
{
    const long nchunk = (nexp + SH3_E - 1) / SH3_E;
    const double2* w2 = reinterpret_cast<const double2*>(wms);
    double2* l2 = reinterpret_cast<double2*>(loc);
    for (long q = blockIdx.x; q < (long)nc * nchunk; q += gridDim.x) {
        const int c = children[q % nc], p = par[c];
        const long e0 = (q / nc) * SH3_E + threadIdx.x;
        int npts = 0;
        if (use_trg) npts += trg_se[c].y - trg_se[c].x;
        if (use_src) npts += src_se[c].y - src_se[c].x;
        if (npts <= 0) continue;
        const int k = octant(cen[3 * c] - cen[3 * p], cen[3 * c + 1] - cen[3 * p + 1],
                             cen[3 * c + 2] - cen[3 * p + 2]);
        const double2* w = w2 + (long)(kbase + k) * nexp;
        for (int ind = 0; ind < nd; ++ind) {
            const double2* x = l2 + ((long)slot[p] * nd + ind) * nexp;
            double2* dst = l2 + ((long)slot[c] * nd + ind) * nexp;
            double2 xv[SH3_V], wv[SH3_V];
#pragma unroll
            for (int v = 0; v < SH3_V; ++v) {
                const long e = e0 + 256 * v;
                xv[v] = e < nexp ? x[e] : make_double2(0.0, 0.0);
                wv[v] = e < nexp ? w[e] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int v = 0; v < SH3_V; ++v) {
                const long e = e0 + 256 * v;
                if (e < nexp)
                    dst[e] = make_double2(xv[v].x * wv[v].x - xv[v].y * wv[v].y,
                                          xv[v].x * wv[v].y + xv[v].y * wv[v].x);
            }
        }
    }
}

void launch_split3(const int* children, int nc, const int* par, const double* cen,
                   const int* slot, const double* wms, int kbase, int nd, long nexp, double* loc,
                   const int2* src_se, const int2* trg_se, int use_src, int use_trg,
                   hipStream_t s)
{
    if (nc <= 0) return;
    kmark(KID_SPLIT, false, s);
    k_split3<<<gridcap((long)nc * ((nexp + SH3_E - 1) / SH3_E), 1), 256, 0, s>>>(
        children, nc, par, cen, slot, wms, kbase, nd, nexp, loc, src_se, trg_se, use_src, use_trg);
    kmark(KID_SPLIT, true, s);
}

// ---------------------------------------------------------------------------
// Step 5, Taylor route (g3dpw2local_real_vec g3drouts.f:1774-1945, then
// g3dlevalp/g/h_vec :1045-1365).  With D(j,a) = (i ts_a)^j / j!,
//   L(j3,j2,j1) = 2 Re sum_{c<half} D(j1,c) sum_b D(j2,b) sum_a D(j3,a) PW(a,b,c).
// One block per leaf; per z node c: S1(j3,b) in LDS, then each thread owns
// (j3,j2) pairs and keeps L(j3,j2,.) in registers across c.
// ---------------------------------------------------------------------------
constexpr int L3_MAXM = 64;  // largest Taylor order + 1 of the 3D kernels

// true when box b is evaluated from its plane waves directly (fewer than
// thresh sources + targets, fgt3d.f:1321-1324)
__device__ __forceinline__ bool pw_box(const Eval3Args& a, int b)
{
    if (a.thresh <= 0) return false;
    const int n = (a.rs[b].y - a.rs[b].x) + (a.rt[b].y - a.rt[b].x);
    return n < a.thresh;
}

// true when box b holds points where the potential is wanted
__device__ __forceinline__ bool has_eval(const Eval3Args& a, int b)
{
    int n = 0;
    if (a.use_src) n += a.rs[b].y - a.rs[b].x;
    if (a.use_trg) n += a.rt[b].y - a.rt[b].x;
    return n > 0;
}

template <int MM>
__global__ __launch_bounds__(256) void k_pw2local3(Eval3Args a)
{
    extern __shared__ double sm[];
    const int m = a.nl + 1, npw = a.npw, half = npw / 2, t = threadIdx.x;
    const long nexp = (long)npw * npw * half;
    double* Ds = sm;                  // complex (npw, m): a + npw*j
    double* S1 = sm + 2 * npw * m;    // complex (m, npw): j3 + m*b
    for (int o = t; o < npw * m; o += 256) {
        Ds[2 * o] = a.dcoef[2 * (o % npw + a.cstride * (o / npw))];
        Ds[2 * o + 1] = a.dcoef[2 * (o % npw + a.cstride * (o / npw)) + 1];
    }
    for (int li = blockIdx.x; li < a.nbox; li += gridDim.x) {
        const int box = a.boxes[li];
        if (pw_box(a, box) || !has_eval(a, box)) continue;
        const double* pw = a.loc + 2 * ((long)a.slot[box] * a.nd + a.ind) * nexp;
        double* L = a.taylor + (long)li * taylor3_stride(m, a.tlayout);
        const int MP = 16 * ((m + 15) / 16);
        if (a.tlayout)  // padding columns j2 = m..MP-1 of the transposed layout
            for (int o = t; o < (MP - m) * m * m; o += 256) {
                const int j2 = m + o % (MP - m), r = o / (MP - m);
                L[j2 + (long)MP * r] = 0;
            }
        for (int p0 = 0; p0 < m * m; p0 += 256) {
            const int pr = p0 + t, j3 = pr % m, j2 = pr / m;
            const bool act = pr < m * m;
            double acc[MM];
#pragma unroll
            for (int j = 0; j < MM; ++j) acc[j] = 0;
            for (int c = 0; c < half; ++c) {
                __syncthreads();
                for (int o = t; o < m * npw; o += 256) {
                    const int jj = o % m, b = o / m;
                    double sr = 0, si = 0;
                    const double* row = pw + 2 * ((long)npw * (b + (long)npw * c));
                    for (int aa = 0; aa < npw; ++aa) {
                        const double dr = Ds[2 * (aa + npw * jj)], di = Ds[2 * (aa + npw * jj) + 1];
                        const double pr_ = row[2 * aa], pi_ = row[2 * aa + 1];
                        sr += dr * pr_ - di * pi_;
                        si += dr * pi_ + di * pr_;
                    }
                    S1[2 * o] = sr;
                    S1[2 * o + 1] = si;
                }
                __syncthreads();
                if (act) {
                    double s2r = 0, s2i = 0;
                    for (int b = 0; b < npw; ++b) {
                        const double dr = Ds[2 * (b + npw * j2)], di = Ds[2 * (b + npw * j2) + 1];
                        const double xr = S1[2 * (j3 + m * b)], xi = S1[2 * (j3 + m * b) + 1];
                        s2r += dr * xr - di * xi;
                        s2i += dr * xi + di * xr;
                    }
                    // 2 Re(D(j1,c) S2)
#pragma unroll
                    for (int j1 = 0; j1 < MM; ++j1)
                        if (j1 < m) {
                            const double dr = Ds[2 * (c + npw * j1)], di = Ds[2 * (c + npw * j1) + 1];
                            acc[j1] += 2 * (dr * s2r - di * s2i);
                        }
                }
            }
            if (act)
#pragma unroll
                for (int j1 = 0; j1 < MM; ++j1)
                    if (j1 < m) {
                        if (a.tlayout)
                            L[j2 + (long)MP * (j3 + (long)m * j1)] = acc[j1];
                        else
                            L[j3 + m * (j2 + (long)m * j1)] = acc[j1];
                    }
        }
        __syncthreads();
    }
}


// PW -> Taylor on matrix cores for Taylor orders nl < 16 and npw <= 32, with
// D(a, j) = (i ts_a)^j / j! = i^j d(a, j), d real:
//   for each z node c < npw/2:
//     S1'(j3, b) = sum_a d(a, j3) PW(a, b, c)     MFMA, a padded to 32
//     S2'(j3, j2) = sum_b S1'(j3, b) d(b, j2)     MFMA
//     L(j3, j2, j1) += 2 d(c, j1) Re(i^(j1 + j2 + j3) S2'(j3, j2))  registers
// (S1 = i^j3 S1', S2 = i^(j2 + j3) S2'), two real products per complex
// operand instead of four.
// d(a, j) from the complex table D = i^j d: its one non-zero part, signed
__device__ __forceinline__ double dreal(const double* __restrict__ dc, int j)
{
    const double v = (j & 1) ? dc[1] : dc[0];
    return (j & 3) >= 2 ? -v : v;
}

// the factor of Re(i^(j + q) z) = f(j) * (j odd ? Im : Re)(i^q z) for order
// j: +1 for j = 0, 3 mod 4, -1 for j = 1, 2 mod 4
__device__ __forceinline__ constexpr double jsign(int j) { return ((j + 1) & 2) ? -1.0 : 1.0; }

constexpr int PW3_N = 32;  // padded npw
constexpr int PW3_S = 33;  // LDS stride of a slab column (bank-conflict-free lane stride)

// Two wavefronts per leaf: wave w stages half of the slab (times the octant
// shift when the split is fused) into LDS with padded strides, computes the
// S1 column tile nt = w into a shared LDS buffer; then both waves form S2
// (the same tile) and each keeps L for its half of j1 (8 orders) in
// registers, so that two leaves fit per SIMD.  The next z node's slab is
// loaded while the current one is computed: the block barriers order LDS
// only, so those loads stay in flight across them.  Lane l holds
// L(j3 = l/16 + 4r, j2 = l%16, j1 = 8w + j).  Writes the transposed layout
// Lt[j2 + 16 (j3 + m j1)]; columns j2 >= m hold the zeros computed for them.
__global__ __launch_bounds__(128, 2) void k_pw2local3_wave2(Eval3Args a)
{
    extern __shared__ double sm[];
    const int m = a.nl + 1, npw = a.npw, half = npw / 2, t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const long nexp = (long)npw * npw * half;
    double* P = sm;                        // complex (PW3_N a) x (PW3_N b) at a + PW3_S b
    double* S1 = sm + 2 * PW3_S * PW3_N;   // complex (16 j3) x (PW3_N b) at j3 + 17 b
    double* Dt = S1 + 2 * 17 * PW3_N;      // real d(a, j) at j + 16 a, a < PW3_N
    const int i16 = lane & 15, k4 = lane >> 4;
    const int q = (i16 + k4) & 3;          // (j2 + j3) mod 4 of this lane's S2' elements
    for (int o = t; o < PW3_N * 16; o += 128) {
        const int aa = o / 16, j = o % 16;
        Dt[o] = aa < npw && j < m ? dreal(a.dcoef + 2 * (aa + a.cstride * j), j) : 0.0;
    }
    for (int li = blockIdx.x; li < a.nbox; li += gridDim.x) {
        const int box = a.boxes[li];
        if (pw_box(a, box) || !has_eval(a, box)) continue;
        const int src = a.fused ? a.par[box] : box;
        const double* pw = a.loc + 2 * ((long)a.slot[src] * a.nd + a.ind) * nexp;
        const double* wk = nullptr;
        if (a.fused) {
            const int k = octant(a.cen[3 * box] - a.cen[3 * src], a.cen[3 * box + 1] - a.cen[3 * src + 1],
                                 a.cen[3 * box + 2] - a.cen[3 * src + 2]);
            wk = a.wms + 2 * ((long)(a.kbase + k) * nexp);
        }
        double L[4][8];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int j = 0; j < 8; ++j) L[r][j] = 0;
        // the slab of node c + 1 is loaded while node c is computed
        double2 v[8];
        auto load_slab = [&](int c) {
            const double* slab = pw + 2L * npw * npw * c;
#pragma unroll
            for (int it = 0; it < 8; ++it) {
                const int o = t + 128 * it, aa = o % PW3_N, b = o / PW3_N;
                v[it] = aa < npw && b < npw ? *(const double2*)(slab + 2 * (aa + npw * b))
                                            : make_double2(0.0, 0.0);
            }
        };
        if (half > 0) load_slab(0);
        for (int c = 0; c < half; ++c) {
            lds_barrier3();  // readers of P and S1 of the previous node are done
            {
                if (wk) {
                    const double* ws = wk + 2L * npw * npw * c;
#pragma unroll
                    for (int it = 0; it < 8; ++it) {
                        const int o = t + 128 * it, aa = o % PW3_N, b = o / PW3_N;
                        if (aa < npw && b < npw) {
                            const double2 wv = *(const double2*)(ws + 2 * (aa + npw * b));
                            v[it] = make_double2(v[it].x * wv.x - v[it].y * wv.y,
                                                 v[it].x * wv.y + v[it].y * wv.x);
                        }
                    }
                }
#pragma unroll
                for (int it = 0; it < 8; ++it) {
                    const int o = t + 128 * it, aa = o % PW3_N, b = o / PW3_N;
                    *(double2*)(P + 2 * (aa + PW3_S * b)) = v[it];
                }
                if (c + 1 < half) load_slab(c + 1);
            }
            lds_barrier3();
            {
                // S1 column tile nt = w
                d4_t3 s1r = {0, 0, 0, 0}, s1i = {0, 0, 0, 0};
#pragma unroll 2
                for (int ks = 0; ks < PW3_N / 4; ++ks) {
                    const int o = (4 * ks + k4) + PW3_S * (16 * w + i16);
                    const double pr = P[2 * o], pi = P[2 * o + 1];
                    const double d = Dt[i16 + 16 * (4 * ks + k4)];
                    s1r = mfma3(d, pr, s1r);
                    s1i = mfma3(d, pi, s1i);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int o = (k4 + 4 * r) + 17 * (16 * w + i16);
                    S1[2 * o] = s1r[r];
                    S1[2 * o + 1] = s1i[r];
                }
            }
            lds_barrier3();
            d4_t3 s2r = {0, 0, 0, 0}, s2i = {0, 0, 0, 0};
#pragma unroll 2
            for (int ks = 0; ks < PW3_N / 4; ++ks) {
                const int o = i16 + 17 * (4 * ks + k4);
                const double xr = S1[2 * o], xi = S1[2 * o + 1];
                const double d = Dt[i16 + 16 * (4 * ks + k4)];
                s2r = mfma3(xr, d, s2r);
                s2i = mfma3(xi, d, s2i);
            }
            double zr[4], zi[4];  // i^q S2'
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double2 z = rot_ij(s2r[r], s2i[r], q);
                zr[r] = z.x;
                zi[r] = z.y;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {  // j1 = 8 w + j, j1 = j mod 4
                const double f = 2.0 * jsign(j) * Dt[16 * c + 8 * w + j];
#pragma unroll
                for (int r = 0; r < 4; ++r) L[r][j] += f * ((j & 1) ? zi[r] : zr[r]);
            }
        }
        double* out = a.taylor + (long)li * taylor3_stride(m, 1);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int j3 = k4 + 4 * r;
            if (j3 < m)
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (8 * w + j < m) out[i16 + 16 * (j3 + (long)m * (8 * w + j))] = L[r][j];
        }
    }
}

// PW -> Taylor on matrix cores for 16 < m <= 32 (m = nl + 1) and npw <= 4 KA,
// one block of 8 wavefronts per leaf.  Per z node c:
//   the slab PW(., ., c) -> LDS P (a padded to 4 KA, b to 16 NB, zeros);
//   S1(j3, b) = sum_a D(a, j3) P(a, b): waves w < 2 NB, tile (j3: w % 2,
//     b: w / 2), into LDS;
//   S2(j3, j2) = sum_b S1(j3, b) D(b, j2): waves 4..7, tile (j3: w % 2,
//     j2: (w / 2) % 2), into LDS (times i^(j2 + j3), see below);
//   L(j3, j2, j1) += 2 Re(D(c, j1) S2(j3, j2)) for the wave's tile w % 4
//     and half of j1 (w / 4), in registers.
// The products run on the real d(a, j) of D = i^j d, as in
// k_pw2local3_wave2; d sits in LDS for every a < 4 KA, j < 32 (zero outside
// a < npw, j < m).  Writes the transposed layout Lt[j2 + 32 (j3 + m j1)].
template <int KA, int NB>
__global__ __launch_bounds__(512) void k_pw2local3_blk(Eval3Args a)
{
    constexpr int A4 = 4 * KA, BP = 16 * NB, PS = A4 + 1;
    extern __shared__ double sm[];
    double* P = sm;                     // complex (A4 a) x (BP b) at a + PS b
    double* S1 = P + 2 * PS * BP;       // complex (32 j3) x (BP b) at j3 + 33 b
    double* Dt = S1 + 2 * 33 * BP;      // real d(a, j) at j + 32 a, a < A4
    double* S2b = Dt + A4 * 32;         // 4 complex 16 x 16 tiles of i^q S2', lane order
    const int m = a.nl + 1, npw = a.npw, half = npw / 2, t = threadIdx.x;
    const int w = t >> 6, lane = t & 63, i16 = lane & 15, k4 = lane >> 4;
    const int mt = w & 1, jt = (w >> 1) & 1, jh = w >> 2;
    const int q = (i16 + k4) & 3;       // (j2 + j3) mod 4 of this lane's S2' elements
    const long nexp = (long)npw * npw * half;
    for (int o = t; o < A4 * 32; o += 512) {
        const int aa = o / 32, j = o % 32;
        Dt[o] = aa < npw && j < m ? dreal(a.dcoef + 2 * (aa + a.cstride * j), j) : 0.0;
    }
    for (int li = blockIdx.x; li < a.nbox; li += gridDim.x) {
        const int box = a.boxes[li];
        if (pw_box(a, box) || !has_eval(a, box)) continue;
        // the leaf's own expansion, or with a fused split its parent's,
        // multiplied by the octant shift while the slab is staged
        const int src = a.fused ? a.par[box] : box;
        const double* pw = a.loc + 2 * ((long)a.slot[src] * a.nd + a.ind) * nexp;
        const double* wk = nullptr;
        if (a.fused) {
            const int k = octant(a.cen[3 * box] - a.cen[3 * src], a.cen[3 * box + 1] - a.cen[3 * src + 1],
                                 a.cen[3 * box + 2] - a.cen[3 * src + 2]);
            wk = a.wms + 2 * ((long)(a.kbase + k) * nexp);
        }
        double L[4][16];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int j = 0; j < 16; ++j) L[r][j] = 0;
        for (int c = 0; c < half; ++c) {
            __syncthreads();  // readers of P, S1 of the previous node are done
            const double* slab = pw + 2L * npw * npw * c;
            {
                constexpr int NIT = (A4 * BP + 511) / 512;
                double2 v[NIT];
#pragma unroll
                for (int it = 0; it < NIT; ++it) {
                    const int o = t + 512 * it, aa = o % A4, b = o / A4;
                    v[it] = o < A4 * BP && aa < npw && b < npw
                                ? *(const double2*)(slab + 2 * (aa + npw * b))
                                : make_double2(0.0, 0.0);
                }
                if (wk) {
                    const double* ws = wk + 2L * npw * npw * c;
#pragma unroll
                    for (int it = 0; it < NIT; ++it) {
                        const int o = t + 512 * it, aa = o % A4, b = o / A4;
                        if (o < A4 * BP && aa < npw && b < npw) {
                            const double2 wv = *(const double2*)(ws + 2 * (aa + npw * b));
                            v[it] = make_double2(v[it].x * wv.x - v[it].y * wv.y,
                                                 v[it].x * wv.y + v[it].y * wv.x);
                        }
                    }
                }
#pragma unroll
                for (int it = 0; it < NIT; ++it) {
                    const int o = t + 512 * it, aa = o % A4, b = o / A4;
                    if (o < A4 * BP) *(double2*)(P + 2 * (aa + PS * b)) = v[it];
                }
            }
            __syncthreads();
            if (w < 2 * NB) {
                const int nb = w >> 1;
                d4_t3 sr = {0, 0, 0, 0}, si = {0, 0, 0, 0};
#pragma unroll 2
                for (int ks = 0; ks < KA; ++ks) {
                    const int o = (4 * ks + k4) + PS * (16 * nb + i16);
                    const double pr = P[2 * o], pi = P[2 * o + 1];
                    const double d = Dt[(16 * mt + i16) + 32 * (4 * ks + k4)];
                    sr = mfma3(d, pr, sr);
                    si = mfma3(d, pi, si);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int o = (16 * mt + k4 + 4 * r) + 33 * (16 * nb + i16);
                    S1[2 * o] = sr[r];
                    S1[2 * o + 1] = si[r];
                }
            }
            __syncthreads();
            double* S2t = S2b + 2 * 256 * (w & 3);  // this wave's tile (mt, jt)
            if (w >= 4) {
                d4_t3 s2r = {0, 0, 0, 0}, s2i = {0, 0, 0, 0};
#pragma unroll 2
                for (int ks = 0; ks < KA; ++ks) {
                    const int o = (16 * mt + i16) + 33 * (4 * ks + k4);
                    const double xr = S1[2 * o], xi = S1[2 * o + 1];
                    const double d = Dt[(16 * jt + i16) + 32 * (4 * ks + k4)];
                    s2r = mfma3(xr, d, s2r);
                    s2i = mfma3(xi, d, s2i);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    *(double2*)(S2t + 2 * (64 * r + lane)) = rot_ij(s2r[r], s2i[r], q);
            }
            __syncthreads();
            double zr[4], zi[4];  // i^q S2' of this lane's elements
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double2 z = *(const double2*)(S2t + 2 * (64 * r + lane));
                zr[r] = z.x;
                zi[r] = z.y;
            }
#pragma unroll
            for (int j = 0; j < 16; ++j) {  // j1 = 16 jh + j, j1 = j mod 4
                const double f = 2.0 * jsign(j) * Dt[32 * c + 16 * jh + j];
#pragma unroll
                for (int r = 0; r < 4; ++r) L[r][j] += f * ((j & 1) ? zi[r] : zr[r]);
                if ((j & 3) == 3) __builtin_amdgcn_sched_barrier(0);  // bounds the loads in flight
            }
        }
        double* out = a.taylor + (long)li * taylor3_stride(m, 1);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int j3 = 16 * mt + k4 + 4 * r, j2 = 16 * jt + i16;
            if (j3 < m)
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const int j1 = 16 * jh + j;
                    if (j1 < m) out[j2 + 32 * (j3 + (long)m * j1)] = L[r][j];
                }
        }
    }
}

template <int KA, int NB>
static void pw2local3_blk_launch(const Eval3Args& a, hipStream_t s)
{
    const size_t lds = ((size_t)2 * (4 * KA + 1) * 16 * NB + (size_t)2 * 33 * 16 * NB +
                        (size_t)4 * KA * 32 + (size_t)2 * 4 * 256) * sizeof(double);
    allow_lds3(k_pw2local3_blk<KA, NB>, lds);
    k_pw2local3_blk<KA, NB><<<std::min(a.nbox, 8192), 512, lds, s>>>(a);
}

// whether launch_pw2local3 runs a kernel that can fuse the parent -> leaf
// split (Eval3Args::fused) for Taylor order m and npw
bool pw2local3_fusable(int m, int npw)
{
    return (m <= 16 && npw <= PW3_N) || (m > 16 && m <= 32 && npw <= 48);
}

void launch_pw2local3(const Eval3Args& a, hipStream_t s)
{
    if (a.nbox <= 0) return;
    const int m = a.nl + 1;
    if (m > L3_MAXM) fail(-3, "3D Taylor order above 63 is not supported");
    if (a.fused && !(pw2local3_fusable(m, a.npw) && a.tlayout))
        fail(-3, "fused split requested for a PW -> Taylor kernel without it");
    if (m <= 16 && a.npw <= PW3_N && a.tlayout) {
        size_t lds = ((size_t)2 * PW3_S * PW3_N + (size_t)2 * 17 * PW3_N + (size_t)PW3_N * 16) *
                     sizeof(double);
        kmark(KID_PW2LOCAL, false, s);
        k_pw2local3_wave2<<<std::min(a.nbox, 65536), 128, lds, s>>>(a);
        kmark(KID_PW2LOCAL, true, s);
        return;
    }
    if (m > 16 && m <= 32 && a.npw <= 48 && a.tlayout) {
        kmark(KID_PW2LOCAL, false, s);
        if (a.npw <= 32)
            pw2local3_blk_launch<8, 2>(a, s);
        else if (a.npw <= 40)
            pw2local3_blk_launch<10, 3>(a, s);
        else
            pw2local3_blk_launch<12, 3>(a, s);
        kmark(KID_PW2LOCAL, true, s);
        return;
    }
    size_t lds = (size_t)4 * a.npw * m * sizeof(double);
    kmark(KID_PW2LOCAL, false, s);
    if (m <= 32) {
        allow_lds3(k_pw2local3<32>, lds);
        k_pw2local3<32><<<a.grid, 256, lds, s>>>(a);
    } else {
        allow_lds3(k_pw2local3<64>, lds);
        k_pw2local3<64><<<a.grid, 256, lds, s>>>(a);
    }
    kmark(KID_PW2LOCAL, true, s);
}

// Taylor evaluation over 256-point segments: sum L(j3,j2,j1) x^j3 y^j2 z^j1
// and its derivatives (Horner in x, explicit powers in y and z); outputs
// accumulated; Hessian order xx, yy, zz, xy, xz, yz.
template <int PGH>
__global__ __launch_bounds__(256) void k_taylor_eval3(Eval3Args a)
{
    const int nl = a.nl, m = nl + 1, t = threadIdx.x;
    const int total = a.tstart[a.nbox];
    for (int g = a.tstart[0] + blockIdx.x; g < total; g += gridDim.x) {
        int lo = 0, hi = a.nbox - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (a.tstart[mid] <= g) lo = mid; else hi = mid - 1;
        }
        const int b = a.boxes[lo];
        if (pw_box(a, b)) continue;
        const int2 rr = a.se[b];
        const int p0 = rr.x + (g - a.tstart[lo]) * a.seg, p1 = min(rr.y, p0 + a.seg);
        const double* L = a.taylor + (long)lo * m * m * m;
        const double cx = a.cen[3 * b], cy = a.cen[3 * b + 1], cz = a.cen[3 * b + 2], dsq = a.dsq;
        for (int i = p0 + t; i < p1; i += 256) {
            const double x = (a.pts[3 * (long)i] - cx) * dsq, y = (a.pts[3 * (long)i + 1] - cy) * dsq,
                         z = (a.pts[3 * (long)i + 2] - cz) * dsq;
            double pot = 0, gx = 0, gy = 0, gz = 0, hxx = 0, hyy = 0, hzz = 0, hxy = 0, hxz = 0,
                   hyz = 0;
            double zp = 1, zpm1 = 0, zpm2 = 0;  // z^j1, j1 z^(j1-1), j1(j1-1) z^(j1-2)
            for (int j1 = 0; j1 < m; ++j1) {
                double A = 0, Ay = 0, Ayy = 0, B = 0, By = 0, Cx = 0;
                double yp = 1, ypm1 = 0, ypm2 = 0;
                for (int j2 = 0; j2 < m; ++j2) {
                    const double* c = L + (long)m * (j2 + (long)m * j1);
                    double q = c[nl], q1 = 0, q2 = 0;
                    for (int j3 = nl - 1; j3 >= 0; --j3) {
                        if (PGH >= 3) q2 = q2 * x + q1;
                        if (PGH >= 2) q1 = q1 * x + q;
                        q = q * x + c[j3];
                    }
                    A += yp * q;
                    if (PGH >= 2) {
                        Ay += ypm1 * q;
                        B += yp * q1;
                    }
                    if (PGH >= 3) {
                        Ayy += ypm2 * q;
                        By += ypm1 * q1;
                        Cx += yp * 2 * q2;
                    }
                    ypm2 = (PGH >= 3) ? ypm2 * y + 2 * ypm1 : 0;
                    ypm1 = ypm1 * y + yp;
                    yp = yp * y;
                }
                pot += zp * A;
                if (PGH >= 2) {
                    gx += zp * B;
                    gy += zp * Ay;
                    gz += zpm1 * A;
                }
                if (PGH >= 3) {
                    hxx += zp * Cx;
                    hyy += zp * Ayy;
                    hzz += zpm2 * A;
                    hxy += zp * By;
                    hxz += zpm1 * B;
                    hyz += zpm1 * Ay;
                }
                zpm2 = (PGH >= 3) ? zpm2 * z + 2 * zpm1 : 0;
                zpm1 = zpm1 * z + zp;
                zp = zp * z;
            }
            const long o = (long)i * a.nd + a.ind;
            a.pot[o] += pot;
            if (PGH >= 2) {
                const long gi = (long)i * 3 * a.nd + a.ind;
                a.grad[gi] += gx * dsq;
                a.grad[gi + a.nd] += gy * dsq;
                a.grad[gi + 2 * a.nd] += gz * dsq;
            }
            if (PGH >= 3) {
                const long hi2 = (long)i * 6 * a.nd + a.ind;
                const double d2 = dsq * dsq;
                a.hess[hi2] += hxx * d2;
                a.hess[hi2 + a.nd] += hyy * d2;
                a.hess[hi2 + 2 * a.nd] += hzz * d2;
                a.hess[hi2 + 3 * a.nd] += hxy * d2;
                a.hess[hi2 + 4 * a.nd] += hxz * d2;
                a.hess[hi2 + 5 * a.nd] += hyz * d2;
            }
        }
    }
}

// Taylor evaluation on FP64 matrix cores (Taylor order nl < 32), the
// coefficients in the transposed layout Lt[j2 + MP (j3 + m j1)], MP = 16 NT,
// columns j2 >= m zero:
//   T(p, j2) = sum_j3 x_p^j3 Lt(j2, j3, j1)          MFMA, 16-point tiles
//   pot(p) += sum_j1 z_p^j1 sum_j2 y_p^j2 T(p, j2)     registers, 16-lane sum
// x derivatives come from the rows j3 x^(j3-1) and j3 (j3-1) x^(j3-2) (more
// MFMAs), y and z derivatives from the weights.  One wavefront per segment of
// at most 256 points of one leaf; lane l supplies the powers of point l%16
// of the tile and holds T(l/16 + 4r, 16 nt + l%16).
template <int NT, int PGH>
__global__ __launch_bounds__(256) void k_taylor3_mfma(Eval3Args a)
{
    constexpr int MP = 16 * NT, KS = 4 * NT;
    constexpr int NX = PGH >= 3 ? 3 : (PGH == 2 ? 2 : 1);
    constexpr int NACC = PGH >= 3 ? 10 : (PGH == 2 ? 4 : 1);
    const int m = a.nl + 1, lane = threadIdx.x & 63, i16 = lane & 15, k4 = lane >> 4;
    const long tst = taylor3_stride(m, 1);
    const int total = a.tstart[a.nbox];
    for (int g = a.tstart[0] + blockIdx.x * 4 + (threadIdx.x >> 6); g < total;
         g += gridDim.x * 4) {
        int lo = 0, hi = a.nbox - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (a.tstart[mid] <= g) lo = mid; else hi = mid - 1;
        }
        const int b = a.boxes[lo];
        if (pw_box(a, b)) continue;
        const int2 rr = a.se[b];
        const int p0 = rr.x + (g - a.tstart[lo]) * a.seg, p1 = min(rr.y, p0 + a.seg);
        const double* Lt = a.taylor + (long)lo * tst;
        const double cx = a.cen[3 * b], cy = a.cen[3 * b + 1], cz = a.cen[3 * b + 2], dsq = a.dsq;
        for (int t0 = p0; t0 < p1; t0 += 16) {
            // A operands: powers of x of point t0 + l%16, orders 4 ks + l/16
            double X[NX][KS];
            {
                const int pa = t0 + i16;
                const double x = pa < p1 ? (a.pts[3 * (long)pa] - cx) * dsq : 0.0;
                double q0 = 1, q1 = 0, q2 = 0;  // x^j, j x^(j-1), j (j-1) x^(j-2)
#pragma unroll
                for (int j = 0; j < 4 * KS; ++j) {
                    if ((j & 3) == k4) {
                        const bool in = j < m;
                        X[0][j >> 2] = in ? q0 : 0.0;
                        if (NX > 1) X[1 % NX][j >> 2] = in ? q1 : 0.0;
                        if (NX > 2) X[2 % NX][j >> 2] = in ? q2 : 0.0;
                    }
                    q2 = (j + 1) * q1;
                    q1 = (j + 1) * q0;
                    q0 *= x;
                }
            }
            // result rows: points t0 + l/16 + 4r; y weights of column 16 nt + l%16
            double zv[4], Y[3][4][NT];
            bool in_r[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int p = t0 + k4 + 4 * r;
                in_r[r] = p < p1;
                const double y = in_r[r] ? (a.pts[3 * (long)p + 1] - cy) * dsq : 0.0;
                zv[r] = in_r[r] ? (a.pts[3 * (long)p + 2] - cz) * dsq : 0.0;
                double q0 = 1, q1 = 0, q2 = 0;
#pragma unroll
                for (int j = 0; j < MP; ++j) {
                    if ((j & 15) == i16) {
                        const bool in = j < m;
                        Y[0][r][j >> 4] = in ? q0 : 0.0;
                        Y[1][r][j >> 4] = in ? q1 : 0.0;
                        Y[2][r][j >> 4] = in ? q2 : 0.0;
                    }
                    q2 = (j + 1) * q1;
                    q1 = (j + 1) * q0;
                    q0 *= y;
                }
            }
            // acc: pot; gx gy gz; hxx hyy hzz hxy hxz hyz
            double acc[NACC][4];
#pragma unroll
            for (int k = 0; k < NACC; ++k)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[k][r] = 0;
            double z0[4], z1[4], z2[4];  // z^j1, j1 z^(j1-1), j1 (j1-1) z^(j1-2)
#pragma unroll
            for (int r = 0; r < 4; ++r) z0[r] = 1, z1[r] = 0, z2[r] = 0;
            for (int j1 = 0; j1 < m; ++j1) {
                d4_t3 T[NX][NT];
#pragma unroll
                for (int ix = 0; ix < NX; ++ix)
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt) T[ix][nt] = (d4_t3){0, 0, 0, 0};
                const double* Lj = Lt + (long)MP * m * j1;
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) {
                    const int j3 = 4 * ks + k4;
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt) {
                        const double lb = j3 < m ? Lj[(16 * nt + i16) + (long)MP * j3] : 0.0;
#pragma unroll
                        for (int ix = 0; ix < NX; ++ix) T[ix][nt] = mfma3(X[ix][ks], lb, T[ix][nt]);
                    }
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    double s00 = 0, s01 = 0, s10 = 0, s02 = 0, s11 = 0, s20 = 0;
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt) {
                        s00 += T[0][nt][r] * Y[0][r][nt];
                        if (PGH >= 2) {
                            s01 += T[0][nt][r] * Y[1][r][nt];
                            s10 += T[1 % NX][nt][r] * Y[0][r][nt];
                        }
                        if (PGH >= 3) {
                            s02 += T[0][nt][r] * Y[2][r][nt];
                            s11 += T[1 % NX][nt][r] * Y[1][r][nt];
                            s20 += T[2 % NX][nt][r] * Y[0][r][nt];
                        }
                    }
                    acc[0][r] += z0[r] * s00;
                    if (PGH >= 2) {
                        acc[1 % NACC][r] += z0[r] * s10;
                        acc[2 % NACC][r] += z0[r] * s01;
                        acc[3 % NACC][r] += z1[r] * s00;
                    }
                    if (PGH >= 3) {
                        acc[4 % NACC][r] += z0[r] * s20;
                        acc[5 % NACC][r] += z0[r] * s02;
                        acc[6 % NACC][r] += z2[r] * s00;
                        acc[7 % NACC][r] += z0[r] * s11;
                        acc[8 % NACC][r] += z1[r] * s10;
                        acc[9 % NACC][r] += z1[r] * s01;
                    }
                    z2[r] = (j1 + 1) * z1[r];
                    z1[r] = (j1 + 1) * z0[r];
                    z0[r] *= zv[r];
                }
            }
            // sum over the 16 lanes that hold the same points
#pragma unroll
            for (int k = 0; k < NACC; ++k)
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int o = 1; o < 16; o <<= 1) acc[k][r] += __shfl_xor(acc[k][r], o, 64);
            if (i16 == 0)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (!in_r[r]) continue;
                    const long i = t0 + k4 + 4 * r;
                    a.pot[i * a.nd + a.ind] += acc[0][r];
                    if (PGH >= 2) {
                        const long gi = i * 3 * a.nd + a.ind;
                        a.grad[gi] += acc[1 % NACC][r] * dsq;
                        a.grad[gi + a.nd] += acc[2 % NACC][r] * dsq;
                        a.grad[gi + 2 * a.nd] += acc[3 % NACC][r] * dsq;
                    }
                    if (PGH >= 3) {
                        const long hi2 = i * 6 * a.nd + a.ind;
                        const double d2 = dsq * dsq;
                        a.hess[hi2] += acc[4 % NACC][r] * d2;
                        a.hess[hi2 + a.nd] += acc[5 % NACC][r] * d2;
                        a.hess[hi2 + 2 * a.nd] += acc[6 % NACC][r] * d2;
                        a.hess[hi2 + 3 * a.nd] += acc[7 % NACC][r] * d2;
                        a.hess[hi2 + 4 * a.nd] += acc[8 % NACC][r] * d2;
                        a.hess[hi2 + 5 * a.nd] += acc[9 % NACC][r] * d2;
                    }
                }
        }
    }
}

template <int NT>
static void taylor3_mfma_launch(const Eval3Args& a, hipStream_t s)
{
    const int grid = (a.grid + 3) / 4;
    if (a.ifpgh >= 3)
        k_taylor3_mfma<NT, 3><<<grid, 256, 0, s>>>(a);
    else if (a.ifpgh == 2)
        k_taylor3_mfma<NT, 2><<<grid, 256, 0, s>>>(a);
    else
        k_taylor3_mfma<NT, 1><<<grid, 256, 0, s>>>(a);
}

void launch_taylor_eval3(const Eval3Args& a, hipStream_t s)
{
    if (a.nbox <= 0) return;
    const int m = a.nl + 1;
    kmark(KID_TAYLOR_EVAL, false, s);
    if (a.tlayout) {
        if (m > 32) fail(-3, "transposed Taylor layout above order 31");
        if (m <= 16)
            taylor3_mfma_launch<1>(a, s);
        else
            taylor3_mfma_launch<2>(a, s);
    } else if (a.ifpgh >= 3)
        k_taylor_eval3<3><<<a.grid, 256, 0, s>>>(a);
    else if (a.ifpgh == 2)
        k_taylor_eval3<2><<<a.grid, 256, 0, s>>>(a);
    else
        k_taylor_eval3<1><<<a.grid, 256, 0, s>>>(a);
    kmark(KID_TAYLOR_EVAL, true, s);
}

// Step 5, plane-wave route (g3dpwevalp/g/h_vec, g3drouts.f:2268-2668):
// 2 Re sum_{a,b,c<half} PW(a,b,c) e^{i ts_a x} e^{i ts_b y} e^{i ts_c z} and
// derivatives (factors i ts dsq).  Phases by recurrence from the first node:
// e^{i ts_a u} = e^{i ts_0 u} (e^{i h u})^a.
template <int PGH>
__global__ __launch_bounds__(256) void k_pw_eval3(Eval3Args a)
{
    const int npw = a.npw, half = npw / 2, t = threadIdx.x;
    const long nexp = (long)npw * npw * half;
    const double h = a.ts[1] - a.ts[0];
    const int total = a.tstart[a.nbox];
    for (int g = a.tstart[0] + blockIdx.x; g < total; g += gridDim.x) {
        int lo = 0, hi = a.nbox - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (a.tstart[mid] <= g) lo = mid; else hi = mid - 1;
        }
        const int b = a.boxes[lo];
        if (!pw_box(a, b)) continue;
        const int2 rr = a.se[b];
        const int p0 = rr.x + (g - a.tstart[lo]) * a.seg, p1 = min(rr.y, p0 + a.seg);
        const double* pw = a.loc + 2 * ((long)a.slot[b] * a.nd + a.ind) * nexp;
        const double cx = a.cen[3 * b], cy = a.cen[3 * b + 1], cz = a.cen[3 * b + 2], dsq = a.dsq;
        for (int i = p0 + t; i < p1; i += 256) {
            const double u[3] = {(a.pts[3 * (long)i] - cx) * dsq, (a.pts[3 * (long)i + 1] - cy) * dsq,
                                 (a.pts[3 * (long)i + 2] - cz) * dsq};
            double w0r[3], w0i[3], str[3], sti[3];
            for (int d = 0; d < 3; ++d) {
                sincos(a.ts[0] * u[d], &w0i[d], &w0r[d]);
                sincos(h * u[d], &sti[d], &str[d]);
            }
            // accumulators (complex): pot, gx, gy, gz, hxx, hyy, hzz, hxy, hxz, hyz
            double R[10][2] = {};
            double wzr = w0r[2], wzi = w0i[2];
            for (int c = 0; c < half; ++c) {
                double d1[6][2] = {};  // d1, d1x, d1y, d1xx, d1yy, d1xy
                double wyr = w0r[1], wyi = w0i[1];
                for (int bb = 0; bb < npw; ++bb) {
                    double d2[3][2] = {};  // d2, d2x, d2xx
                    double wxr = w0r[0], wxi = w0i[0];
                    const double* row = pw + 2 * ((long)npw * (bb + (long)npw * c));
                    for (int aa = 0; aa < npw; ++aa) {
                        const double pr_ = row[2 * aa], pi_ = row[2 * aa + 1];
                        const double vr = pr_ * wxr - pi_ * wxi, vi = pr_ * wxi + pi_ * wxr;
                        d2[0][0] += vr;
                        d2[0][1] += vi;
                        if (PGH >= 2) {
                            const double f = a.ts[aa] * dsq;  // (i f) v
                            d2[1][0] += -f * vi;
                            d2[1][1] += f * vr;
                            if (PGH >= 3) {
                                d2[2][0] += -f * f * vr;
                                d2[2][1] += -f * f * vi;
                            }
                        }
                        const double tr = wxr * str[0] - wxi * sti[0];
                        wxi = wxr * sti[0] + wxi * str[0];
                        wxr = tr;
                    }
                    const double fy = a.ts[bb] * dsq;
                    auto madd = [](double* acc, const double* v, double wr, double wi) {
                        acc[0] += v[0] * wr - v[1] * wi;
                        acc[1] += v[0] * wi + v[1] * wr;
                    };
                    madd(d1[0], d2[0], wyr, wyi);
                    if (PGH >= 2) {
                        madd(d1[1], d2[1], wyr, wyi);
                        madd(d1[2], d2[0], -fy * wyi, fy * wyr);
                    }
                    if (PGH >= 3) {
                        madd(d1[3], d2[2], wyr, wyi);
                        madd(d1[4], d2[0], -fy * fy * wyr, -fy * fy * wyi);
                        madd(d1[5], d2[1], -fy * wyi, fy * wyr);
                    }
                    const double tr = wyr * str[1] - wyi * sti[1];
                    wyi = wyr * sti[1] + wyi * str[1];
                    wyr = tr;
                }
                const double fz = a.ts[c] * dsq;
                auto madd = [](double* acc, const double* v, double wr, double wi) {
                    acc[0] += v[0] * wr - v[1] * wi;
                    acc[1] += v[0] * wi + v[1] * wr;
                };
                madd(R[0], d1[0], wzr, wzi);
                if (PGH >= 2) {
                    madd(R[1], d1[1], wzr, wzi);
                    madd(R[2], d1[2], wzr, wzi);
                    madd(R[3], d1[0], -fz * wzi, fz * wzr);
                }
                if (PGH >= 3) {
                    madd(R[4], d1[3], wzr, wzi);
                    madd(R[5], d1[4], wzr, wzi);
                    madd(R[6], d1[0], -fz * fz * wzr, -fz * fz * wzi);
                    madd(R[7], d1[5], wzr, wzi);
                    madd(R[8], d1[1], -fz * wzi, fz * wzr);
                    madd(R[9], d1[2], -fz * wzi, fz * wzr);
                }
                const double tr = wzr * str[2] - wzi * sti[2];
                wzi = wzr * sti[2] + wzi * str[2];
                wzr = tr;
            }
            a.pot[(long)i * a.nd + a.ind] += 2 * R[0][0];
            if (PGH >= 2) {
                const long gi = (long)i * 3 * a.nd + a.ind;
                for (int k = 0; k < 3; ++k) a.grad[gi + k * a.nd] += 2 * R[1 + k][0];
            }
            if (PGH >= 3) {
                const long hi2 = (long)i * 6 * a.nd + a.ind;
                for (int k = 0; k < 6; ++k) a.hess[hi2 + k * a.nd] += 2 * R[4 + k][0];
            }
        }
    }
}

// Plane-wave evaluation of potentials on FP64 matrix cores (npw <= 16 NB):
//   T_c(p, b) = sum_a e^{i ts_a x_p} PW(a, b, c)       MFMA, 16-point tiles
//   pot(p) += 2 Re sum_c e^{i ts_c z_p} sum_b e^{i ts_b y_p} T_c(p, b)
// with the phases from sincos (no recurrence).  One wavefront per segment of
// at most 256 points of one leaf; lane l supplies the x phases of point l%16
// and holds T_c(l/16 + 4r, 16 nt + l%16); the b sum runs in registers per
// lane and over 16 lanes at the end.
template <int NB>
__global__ __launch_bounds__(256) void k_pw_eval3_mfma(Eval3Args a)
{
    constexpr int KA = 4 * NB;  // k-steps over a (npw <= 4 KA)
    const int npw = a.npw, half = npw / 2;
    const int lane = threadIdx.x & 63, i16 = lane & 15, k4 = lane >> 4;
    const long nexp = (long)npw * npw * half;
    const int total = a.tstart[a.nbox];
    for (int g = a.tstart[0] + blockIdx.x * 4 + (threadIdx.x >> 6); g < total;
         g += gridDim.x * 4) {
        int lo = 0, hi = a.nbox - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (a.tstart[mid] <= g) lo = mid; else hi = mid - 1;
        }
        const int b = a.boxes[lo];
        if (!pw_box(a, b)) continue;
        const int2 rr = a.se[b];
        const int p0 = rr.x + (g - a.tstart[lo]) * a.seg, p1 = min(rr.y, p0 + a.seg);
        const double* pw = a.loc + 2 * ((long)a.slot[b] * a.nd + a.ind) * nexp;
        const double cx = a.cen[3 * b], cy = a.cen[3 * b + 1], cz = a.cen[3 * b + 2], dsq = a.dsq;
        for (int t0 = p0; t0 < p1; t0 += 16) {
            // A operands: x phases of point t0 + l%16 at nodes 4 ks + l/16
            double exr[KA], exi[KA];
            {
                const int pa = t0 + i16;
                const double x = pa < p1 ? (a.pts[3 * (long)pa] - cx) * dsq : 0.0;
#pragma unroll
                for (int ks = 0; ks < KA; ++ks) {
                    const int aa = 4 * ks + k4;
                    double sn = 0, cs = 0;
                    if (aa < npw) sincos(a.ts[aa] * x, &sn, &cs);
                    exr[ks] = cs;
                    exi[ks] = sn;
                }
            }
            // rows: points t0 + l/16 + 4r; y phases of nodes 16 nt + l%16
            double eyr[4][NB], eyi[4][NB], zr[4];
            bool in_r[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int p = t0 + k4 + 4 * r;
                in_r[r] = p < p1;
                const double y = in_r[r] ? (a.pts[3 * (long)p + 1] - cy) * dsq : 0.0;
                zr[r] = in_r[r] ? (a.pts[3 * (long)p + 2] - cz) * dsq : 0.0;
#pragma unroll
                for (int nt = 0; nt < NB; ++nt) {
                    const int bb = 16 * nt + i16;
                    double sn = 0, cs = 0;
                    if (bb < npw) sincos(a.ts[bb] * y, &sn, &cs);
                    eyr[r][nt] = cs;
                    eyi[r][nt] = sn;
                }
            }
            double ur[4] = {0, 0, 0, 0}, ui[4] = {0, 0, 0, 0};
            for (int c = 0; c < half; ++c) {
                const double* pc = pw + 2L * npw * npw * c;
                d4_t3 tr[NB], ti[NB];
#pragma unroll
                for (int nt = 0; nt < NB; ++nt) {
                    tr[nt] = (d4_t3){0, 0, 0, 0};
                    ti[nt] = (d4_t3){0, 0, 0, 0};
                }
#pragma unroll
                for (int ks = 0; ks < KA; ++ks) {
                    const int aa = 4 * ks + k4;
#pragma unroll
                    for (int nt = 0; nt < NB; ++nt) {
                        const int bb = 16 * nt + i16;
                        double2 v = make_double2(0.0, 0.0);
                        if (aa < npw && bb < npw) v = *(const double2*)(pc + 2 * (aa + npw * bb));
                        tr[nt] = mfma3(exr[ks], v.x, tr[nt]);
                        tr[nt] = mfma3(-exi[ks], v.y, tr[nt]);
                        ti[nt] = mfma3(exr[ks], v.y, ti[nt]);
                        ti[nt] = mfma3(exi[ks], v.x, ti[nt]);
                    }
                }
                const double tc = a.ts[c];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    double sr = 0, si = 0;
#pragma unroll
                    for (int nt = 0; nt < NB; ++nt) {
                        sr += tr[nt][r] * eyr[r][nt] - ti[nt][r] * eyi[r][nt];
                        si += tr[nt][r] * eyi[r][nt] + ti[nt][r] * eyr[r][nt];
                    }
                    double szn, szc;
                    sincos(tc * zr[r], &szn, &szc);
                    ur[r] += szc * sr - szn * si;
                    ui[r] += szc * si + szn * sr;
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int o = 1; o < 16; o <<= 1) ur[r] += __shfl_xor(ur[r], o, 64);
            if (i16 == 0)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (in_r[r]) a.pot[(long)(t0 + k4 + 4 * r) * a.nd + a.ind] += 2 * ur[r];
        }
    }
}

void launch_pw_eval3(const Eval3Args& a, hipStream_t s)
{
    if (a.nbox <= 0) return;
    kmark(KID_PW_EVAL, false, s);
    if (a.ifpgh == 1 && a.npw <= 48) {
        const int grid = (a.grid + 3) / 4;
        if (a.npw <= 16)
            k_pw_eval3_mfma<1><<<grid, 256, 0, s>>>(a);
        else if (a.npw <= 32)
            k_pw_eval3_mfma<2><<<grid, 256, 0, s>>>(a);
        else
            k_pw_eval3_mfma<3><<<grid, 256, 0, s>>>(a);
    } else if (a.ifpgh >= 3)
        k_pw_eval3<3><<<a.grid, 256, 0, s>>>(a);
    else if (a.ifpgh == 2)
        k_pw_eval3<2><<<a.grid, 256, 0, s>>>(a);
    else
        k_pw_eval3<1><<<a.grid, 256, 0, s>>>(a);
    kmark(KID_PW_EVAL, true, s);
}

// ---------------------------------------------------------------------------
// Step 6, near field (g3d_directcp/cg/ch_vec g3ddirect.f:3-206 called from
// fgt3d.f:1460-1512): one block per target unit, its source boxes staged
// in LDS; pairs with r^2 < dmax; accumulated.
// ---------------------------------------------------------------------------
// One source's contribution at a target, r = t - s, e = exp(-|r|^2/delta),
// u = -2 r/delta, rfac = -2/delta: a charge q and a dipole w = d n
// (g3d_directd*_vec, g3ddirect.f:207-435) give A = q - w . u, the sum e A,
// its gradient e (u A - rfac w) and Hessian e ((u_a u_b + rfac d_ab) A -
// rfac (u_a w_b + u_b w_a)), ordered xx, yy, zz, xy, xz, yz.
template <int PGH, bool DIP>
__device__ __forceinline__ void pair_add3(double dx, double dy, double dz, double q,
                                          const double* w, double inv, double rfac, double dmax,
                                          double& p, double& gx, double& gy, double& gz,
                                          double* h)
{
    const double rr = dx * dx + dy * dy + dz * dz;
    if (!(rr < dmax)) return;
    const double e = exp(-rr * inv);
    const double ux = rfac * dx, uy = rfac * dy, uz = rfac * dz;
    const double w1 = DIP ? w[0] : 0.0, w2 = DIP ? w[1] : 0.0, w3 = DIP ? w[2] : 0.0;
    const double A = DIP ? q - (w1 * ux + w2 * uy + w3 * uz) : q;
    p += e * A;
    if (PGH >= 2) {
        gx += e * (ux * A - rfac * w1);
        gy += e * (uy * A - rfac * w2);
        gz += e * (uz * A - rfac * w3);
        if (PGH >= 3) {
            h[0] += e * ((rfac + ux * ux) * A - 2.0 * rfac * ux * w1);
            h[1] += e * ((rfac + uy * uy) * A - 2.0 * rfac * uy * w2);
            h[2] += e * ((rfac + uz * uz) * A - 2.0 * rfac * uz * w3);
            h[3] += e * (ux * uy * A - rfac * (ux * w2 + uy * w1));
            h[4] += e * (ux * uz * A - rfac * (ux * w3 + uz * w1));
            h[5] += e * (uy * uz * A - rfac * (uy * w3 + uz * w2));
        }
    }
}

template <int PGH, bool DIP>
__global__ __launch_bounds__(256) void k_direct3(DirectArgs a)
{
    __shared__ double sx[256], sy[256], sz[256], sq[256], sw[DIP ? 3 * 256 : 1];
    const int u = a.units[blockIdx.x];
    const int2 tr = a.tgt_se[u];
    const int l0 = a.list_ptr[blockIdx.x], l1 = a.list_ptr[blockIdx.x + 1];
    const double inv = 1.0 / a.delta, rfac = -2.0 / a.delta;
    const int t = threadIdx.x;
    for (int ind = 0; ind < a.nd; ++ind)
        for (int t0 = tr.x; t0 < tr.y; t0 += 256) {
            const int i = t0 + t;
            const bool act = i < tr.y;
            const double x = act ? a.xt[3 * (long)i] : 0, y = act ? a.xt[3 * (long)i + 1] : 0,
                         z = act ? a.xt[3 * (long)i + 2] : 0;
            double p = 0, gx = 0, gy = 0, gz = 0, h[6] = {0, 0, 0, 0, 0, 0};
            for (int li = l0; li < l1; ++li) {
                const int2 sr = a.src_se[a.list[li]];
                for (int s0 = sr.x; s0 < sr.y; s0 += 256) {
                    const int cn = min(256, sr.y - s0);
                    __syncthreads();
                    if (t < cn) {
                        const long j = s0 + t;
                        sx[t] = a.xs[3 * j];
                        sy[t] = a.xs[3 * j + 1];
                        sz[t] = a.xs[3 * j + 2];
                        sq[t] = a.qs[j * a.nd + ind];
                        if (DIP)
                            for (int c = 0; c < 3; ++c) sw[3 * t + c] = a.ws[3 * (j * a.nd + ind) + c];
                    }
                    __syncthreads();
                    for (int s = 0; s < cn; ++s)
                        pair_add3<PGH, DIP>(x - sx[s], y - sy[s], z - sz[s], sq[s],
                                            DIP ? sw + 3 * s : nullptr, inv, rfac, a.dmax, p, gx,
                                            gy, gz, h);
                }
            }
            if (act) {
                a.pot[(long)i * a.nd + ind] += p;
                if (PGH >= 2) {
                    const long g = (long)i * 3 * a.nd + ind;
                    a.grad[g] += gx;
                    a.grad[g + a.nd] += gy;
                    a.grad[g + 2 * a.nd] += gz;
                }
                if (PGH >= 3) {
                    const long hh = (long)i * 6 * a.nd + ind;
                    for (int k = 0; k < 6; ++k) a.hess[hh + k * a.nd] += h[k];
                }
            }
        }
}

void launch_direct3(const DirectArgs& a, hipStream_t s)
{
    if (a.nunits <= 0) return;
    kmark(KID_DIRECT, false, s);
    const bool dip = a.ws != nullptr;
    if (a.ifpgh >= 3)
        dip ? k_direct3<3, true><<<a.nunits, 256, 0, s>>>(a) : k_direct3<3, false><<<a.nunits, 256, 0, s>>>(a);
    else if (a.ifpgh == 2)
        dip ? k_direct3<2, true><<<a.nunits, 256, 0, s>>>(a) : k_direct3<2, false><<<a.nunits, 256, 0, s>>>(a);
    else
        dip ? k_direct3<1, true><<<a.nunits, 256, 0, s>>>(a) : k_direct3<1, false><<<a.nunits, 256, 0, s>>>(a);
    kmark(KID_DIRECT, true, s);
}

// all-pairs 3D direct sum with a cutoff (fgt3dpart_direct_vec,
// fgt3d.f:1533-1684), accumulating; thread per target, sources through LDS
template <int PGH, bool DIP>
__global__ __launch_bounds__(256) void k_direct_all3(const double* __restrict__ src,
                                                     const double* __restrict__ q,
                                                     const double* __restrict__ ws, int nd,
                                                     int ns, const double* __restrict__ tg, int nt,
                                                     double delta, double dmax,
                                                     double* __restrict__ pot,
                                                     double* __restrict__ grad,
                                                     double* __restrict__ hess)
{
    __shared__ double sx[256], sy[256], sz[256], sq[256], sw[DIP ? 3 * 256 : 1];
    const int t = threadIdx.x, i = blockIdx.x * 256 + t;
    const bool act = i < nt;
    const double x = act ? tg[3 * (long)i] : 0, y = act ? tg[3 * (long)i + 1] : 0,
                 z = act ? tg[3 * (long)i + 2] : 0;
    const double inv = 1.0 / delta, rfac = -2.0 / delta;
    for (int ind = 0; ind < nd; ++ind) {
        double p = 0, gx = 0, gy = 0, gz = 0, h[6] = {0, 0, 0, 0, 0, 0};
        for (int s0 = 0; s0 < ns; s0 += 256) {
            const int cn = min(256, ns - s0);
            __syncthreads();
            if (t < cn) {
                const long j = s0 + t;
                sx[t] = src[3 * j];
                sy[t] = src[3 * j + 1];
                sz[t] = src[3 * j + 2];
                sq[t] = q ? q[j * nd + ind] : 0.0;
                if (DIP)
                    for (int c = 0; c < 3; ++c) sw[3 * t + c] = ws[3 * (j * nd + ind) + c];
            }
            __syncthreads();
            for (int s = 0; s < cn; ++s)
                pair_add3<PGH, DIP>(x - sx[s], y - sy[s], z - sz[s], sq[s], DIP ? sw + 3 * s : nullptr,
                                    inv, rfac, dmax, p, gx, gy, gz, h);
        }
        if (act) {
            pot[(long)i * nd + ind] += p;
            if (PGH >= 2) {
                const long g = (long)i * 3 * nd + ind;
                grad[g] += gx;
                grad[g + nd] += gy;
                grad[g + 2 * nd] += gz;
            }
            if (PGH >= 3) {
                const long hh = (long)i * 6 * nd + ind;
                for (int k = 0; k < 6; ++k) hess[hh + k * nd] += h[k];
            }
        }
    }
}

void launch_direct_all3(const double* src, const double* q, const double* ws, int nd, int ns,
                        const double* tg, int nt, double delta, double dmax, int ifpgh,
                        double* pot, double* grad, double* hess, hipStream_t s)
{
    if (nt <= 0 || ns <= 0) return;
    const unsigned g = gridn(nt, 256);
#define FGT_DALL3(P_, D_)                                                                      \
    k_direct_all3<P_, D_><<<g, 256, 0, s>>>(src, q, ws, nd, ns, tg, nt, delta, dmax, pot, grad, hess)
    if (ifpgh >= 3) {
        if (ws) FGT_DALL3(3, true); else FGT_DALL3(3, false);
    } else if (ifpgh == 2) {
        if (ws) FGT_DALL3(2, true); else FGT_DALL3(2, false);
    } else {
        if (ws) FGT_DALL3(1, true); else FGT_DALL3(1, false);
    }
#undef FGT_DALL3
}

// ---------------------------------------------------------------------------
// large delta (fgt3d_large_delta, fgt3d.f:1783-1914): one Taylor expansion
// about the bounding-box centre, L(j3,j2,j1) = sum q h_j3(x) h_j2(y) h_j1(z)
// with the Hermite-function recurrence h_0 = e^{-u^2}, h_1 = 2u h_0,
// h_{j+1} = 2 (u h_j - h_{j-1})/(j+1) (g3dformlc_vec, g3drouts.f:784-856, with
// the z offset used for z).  Partial sums per block in a fixed order, then a
// fixed-order reduction: deterministic.
// ---------------------------------------------------------------------------
// Dipoles (the intended g3dformld_vec / g3dformlcd_vec, g3drouts.f:857-1040)
// add w . (dh_x h_y h_z, h_x dh_y h_z, h_x h_y dh_z), dh_j = -dsq (j+1) h_{j+1}.
template <int MM, bool DIP>
__global__ __launch_bounds__(256) void k_formlc3_part(const double* __restrict__ p,
                                                      const double* __restrict__ q,
                                                      const double* __restrict__ ws, int nd,
                                                      int ind, int n, double cx, double cy,
                                                      double cz, double dsq, int nl, int nparts,
                                                      double* __restrict__ part)
{
    extern __shared__ double sm[];
    const int m = nl + 1, t = threadIdx.x;
    const long per = (n + nparts - 1) / nparts;
    const long s0 = blockIdx.x * per, s1 = min((long)n, s0 + per);
    double* hx = sm;                 // [32][m]
    double* hy = sm + 32 * m;
    double* hz = sm + 64 * m;        // q folded in (without dipoles)
    double* dhx = sm + 96 * m;       // dipoles: dh_j per axis
    double* dhy = sm + 128 * m;
    double* dhz = sm + 160 * m;
    double* sw = sm + 192 * m;       // dipoles: q, w1, w2, w3 per source
    double* out = part + (long)blockIdx.x * m * m * m;
    for (int p0 = 0; p0 < m * m; p0 += 256) {
        const int pr = p0 + t, j2 = pr % m, j1 = pr / m;
        const bool act = pr < m * m;
        double acc[MM];
#pragma unroll
        for (int j = 0; j < MM; ++j) acc[j] = 0;
        for (long c0 = s0; c0 < s1; c0 += 32) {
            const int cn = (int)min(32L, s1 - c0);
            __syncthreads();
            if (t < cn) {
                const long i = c0 + t;
                const double c3[3] = {cx, cy, cz};
                double* hh[3] = {hx + t * m, hy + t * m, hz + t * m};
                double* dh[3] = {dhx + t * m, dhy + t * m, dhz + t * m};
                for (int d = 0; d < 3; ++d) {
                    const double u = (p[3 * i + d] - c3[d]) * dsq;
                    const double f = exp(-u * u);
                    double* H = hh[d];
                    // h_j for j < m; with dipoles one more for dh_{m-1}
                    double h0 = f, h1 = 2.0 * u * f;
                    for (int j = 0; j < m; ++j) {
                        H[j] = h0;
                        if (DIP) dh[d][j] = -(h1 * dsq * (j + 1.0));
                        const double h2 = 2.0 * (u * h1 - h0) / (j + 2);
                        h0 = h1;
                        h1 = h2;
                    }
                }
                const double qq = q ? q[(long)nd * i + ind] : 0.0;
                if (!DIP) {
                    for (int j = 0; j < m; ++j) hz[t * m + j] *= qq;
                } else {
                    sw[4 * t] = qq;
                    for (int c = 0; c < 3; ++c) sw[4 * t + 1 + c] = ws[3 * ((long)nd * i + ind) + c];
                }
            }
            __syncthreads();
            if (act)
                for (int s = 0; s < cn; ++s) {
                    if (!DIP) {
                        const double w = hz[s * m + j1] * hy[s * m + j2];
#pragma unroll
                        for (int j = 0; j < MM; ++j)
                            if (j < m) acc[j] += w * hx[s * m + j];
                    } else {
                        const double yv = hy[s * m + j2], zv = hz[s * m + j1];
                        const double F = sw[4 * s] * yv * zv + sw[4 * s + 2] * dhy[s * m + j2] * zv +
                                         sw[4 * s + 3] * yv * dhz[s * m + j1];
                        const double G = sw[4 * s + 1] * yv * zv;
#pragma unroll
                        for (int j = 0; j < MM; ++j)
                            if (j < m) acc[j] += F * hx[s * m + j] + G * dhx[s * m + j];
                    }
                }
        }
        if (act)
#pragma unroll
            for (int j = 0; j < MM; ++j)
                if (j < m) out[j + m * (j2 + (long)m * j1)] = acc[j];
    }
}

__global__ __launch_bounds__(256) void k_sum_parts3(const double* __restrict__ part, int nparts,
                                                    long len, double* __restrict__ out)
{
    long o = blockIdx.x * 256L + threadIdx.x;
    if (o >= len) return;
    double v = 0;
    for (int k = 0; k < nparts; ++k) v += part[(long)k * len + o];
    out[o] = v;
}

void launch_formlc3(const double* p, const double* q, const double* ws, int nd, int ind, int n,
                    const double* c, double dsq, int nl, double* part, int nparts, double* out,
                    hipStream_t s)
{
    const int m = nl + 1;
    if (m > L3_MAXM) fail(-3, "3D Taylor order above 63 is not supported");
    size_t lds = ((size_t)(ws ? 6 : 3) * 32 * m + (ws ? 4 * 32 : 0)) * sizeof(double);
#define FGT_LC3(M_, D_)                                                                        \
    allow_lds3(k_formlc3_part<M_, D_>, lds);                                                   \
    k_formlc3_part<M_, D_><<<nparts, 256, lds, s>>>(p, q, ws, nd, ind, n, c[0], c[1], c[2], dsq, \
                                                    nl, nparts, part)
    if (m <= 32) {
        if (ws) { FGT_LC3(32, true); } else { FGT_LC3(32, false); }
    } else {
        if (ws) { FGT_LC3(64, true); } else { FGT_LC3(64, false); }
    }
#undef FGT_LC3
    const long len = (long)m * m * m;
    k_sum_parts3<<<gridn(len, 256), 256, 0, s>>>(part, nparts, len, out);
}

}  // namespace fgtpw
