// Host side of the box trees (quadtree, D = 2; octree, D = 3): term
// counts, plane-wave nodes and the box-level tree built from the GPU's list
// of non-empty leaves.
//
// The GPU sorts the sources by their full-depth Morton key and finds, per
// point, the level at which the reference's refinement rule stops
// (kernels2d.hip, kernels3d.hip).  What comes back is one record per
// non-empty leaf of the adaptive tree, in Morton order.  From those records
// this file rebuilds the adaptive tree with the reference's box numbering
// (level by level, children of refined boxes in parent order; pts_tree_build,
// pts_tree2d4.f:363-590 and pts_tree3d.f:402-660), then applies the
// reference's level-restriction procedure (pts_tree_fix_lr,
// pts_tree2d4.f:693-996, pts_tree3d.f:855-1171) step for step, because the
// set of boxes it produces is part of the result the transform must match.
//
// Adjacency.  The reference decides "touching" and "colleague" with FP64
// centre distances against 1.05x the half-sum of the box sizes.  Centres of
// one level differ by whole multiples of the box size (up to rounding
// ~1e-16 relative), so those tests are exactly integer tests on the box
// coordinates of each level, which is what is used here:
//   same level:          |d| <= 1 on every axis
//   levels l and l+1:    the finer box's coordinate lies in [2c - 1, 2c + 2]
//                        on every axis.
// Child j of a box has bit (j >> a) & 1 on axis a (x, y, z), as the
// reference numbers children (tree_routs4.f:28-44, tree_routs3d.f:28-46).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <pthread.h>
#include <thread>

#include "fgtpw_impl.hpp"

namespace fgtpw {

namespace {

// Fork-join over index ranges for the tree builder: a few worker threads
// that spin briefly between forks (a build forks a few dozen times within a
// fraction of a millisecond) and sleep otherwise.  One caller at a time; a
// caller that finds the pool busy (concurrent transforms) runs its loop
// itself.  Only loops whose result does not depend on the order of their
// iterations are run here, so the tree is the same as the sequential one.
class TreePool {
public:
    // never destroyed: the workers live until the process ends (no join at
    // exit, and nothing to join in a forked child)
    static TreePool& get()
    {
        static TreePool* p = new TreePool;
        return *p;
    }
    // a forked child has none of the parent's workers: it runs sequentially
    bool acquire() { return nw_ > 0 && !forked_.load() && busy_.try_lock(); }
    void release() { busy_.unlock(); }
    // f(i0, i1) over [0, n) in chunks of `grain`; the caller holds the pool
    template <class F>
    void run(long n, long grain, F& f)
    {
        if (n <= grain) {
            f(0, n);
            return;
        }
        fn_ = [](void* c, long i0, long i1) { (*(F*)c)(i0, i1); };
        ctx_ = &f;
        total_ = n;
        grain_ = grain;
        next_.store(0, std::memory_order_relaxed);
        done_.store(0, std::memory_order_relaxed);
        gen_.fetch_add(1);  // sequentially consistent with the sleepers' check
        if (sleeping_.load() > 0) {
            std::lock_guard<std::mutex> lk(sleep_mu_);
            sleep_cv_.notify_all();
        }
        chunks();
        while (done_.load(std::memory_order_acquire) < nw_) __builtin_ia32_pause();
    }

private:
    TreePool()
    {
        const unsigned hw = std::thread::hardware_concurrency();
        nw_ = (int)std::min(7u, hw > 1 ? hw - 1 : 0u);
        for (int i = 0; i < nw_; ++i) std::thread([this] { worker(); }).detach();
        pthread_atfork(nullptr, nullptr, [] { forked_.store(true); });
    }
    static inline std::atomic<bool> forked_{false};
    void chunks()
    {
        for (;;) {
            const long i0 = next_.fetch_add(grain_, std::memory_order_relaxed);
            if (i0 >= total_) break;
            fn_(ctx_, i0, std::min(total_, i0 + grain_));
        }
    }
    void worker()
    {
        unsigned seen = 0;  // the generation at construction, before any fork
        for (;;) {
            int spin = 0;
            while (gen_.load(std::memory_order_acquire) == seen) {
                if (++spin < 100000) {  // ~1 ms of pause instructions
                    __builtin_ia32_pause();
                    continue;
                }
                std::unique_lock<std::mutex> lk(sleep_mu_);
                sleeping_.fetch_add(1);
                if (gen_.load() == seen)
                    sleep_cv_.wait_for(lk, std::chrono::milliseconds(5));
                sleeping_.fetch_sub(1);
                spin = 0;
            }
            seen = gen_.load(std::memory_order_acquire);
            chunks();
            done_.fetch_add(1, std::memory_order_acq_rel);
        }
    }
    int nw_ = 0;
    std::mutex busy_, sleep_mu_;
    std::condition_variable sleep_cv_;
    std::atomic<unsigned> gen_{0};
    std::atomic<long> next_{0};
    std::atomic<int> done_{0}, sleeping_{0};
    void (*fn_)(void*, long, long) = nullptr;
    void* ctx_ = nullptr;
    long total_ = 0, grain_ = 1;
};

// parallel loop when the builder holds the pool, sequential otherwise
struct Par {
    bool on = false;
    explicit Par(bool want) : on(want && TreePool::get().acquire()) {}
    ~Par()
    {
        if (on) TreePool::get().release();
    }
    template <class F>
    void range(long n, long grain, F f)
    {
        if (on && n > grain)
            TreePool::get().run(n, grain, f);
        else
            f(0, n);
    }
};
thread_local Par* t_par = nullptr;  // the builder's loop runner on this thread

template <class F>
void par_range(long n, long grain, F f)
{
    if (t_par)
        t_par->range(n, grain, f);
    else
        f(0, n);
}

}  // namespace

// x**n for integer n as Fortran lowers it: binary powering.
static double ipow(double a, int n)
{
    bool inv = n < 0;
    double r = 1;
    for (;;) {
        if (n & 1) r *= a;
        n /= 2;
        if (n == 0) break;
        a *= a;
    }
    return inv ? 1 / r : r;
}

// Hermite/Taylor order for boxes of side bsize, g2dhlterms (g2dterms.f:19-66;
// g3dhlterms, g3dterms.f:13-60, is the same routine): the smallest
// p >= int(e r^2)+1 with d*rk*p^(-1/4)*rp^p/(1-rp) < eps, r = bsize/sqrt(2
// delta), rp = r sqrt(e/p), rk = 1.09 (a single-precision constant in the
// reference) * (2 pi)^(-1/4).
int hermite_terms(int ndim, double bsize, double delta, double eps)
{
    const double e = std::exp(1.0);
    const double r = bsize / std::sqrt(2.0 * delta);
    const int p0 = (int)(e * r * r) + 1;
    const double rk = (double)1.09f * std::pow(8 * std::atan(1.0), -0.25);
    for (int p = p0; p <= 300; ++p) {
        double rp = r * std::sqrt(e / p);
        double bound = ndim * rk * std::pow((double)p, -0.25) * ipow(rp, p) / (1 - rp);
        if (bound < eps) return p;
    }
    return -1;
}

// Plane-wave length, g2dpwterms (g2dterms.f:71-133; g3dpwterms identical):
// Fourier truncation at pmax = 2 sqrt(ln 1/eps) and step
// h = 2 pi/(rmax + sqrt(ln 1/eps)), so that the aliased images of the
// Gaussian stay below eps; npw = 2 int(pmax/h).
void planewave_terms(double bsize, double delta, double eps, double* pmax, int* npw)
{
    const double pi = 4.0 * std::atan(1.0);
    double rmax = bsize / std::sqrt(delta);
    double d = std::sqrt(std::log(1.0 / eps));
    *pmax = 2.0 * d;
    double h = 2 * pi / (rmax + d);
    *npw = 2 * (int)(*pmax / h);
}

// Midpoint-rule nodes of exp(-u^2) = sum ws_k exp(i ts_k u), get_pwnodes
// (g2drouts.f:2205-2225): positive nodes first, then their negatives.
void planewave_nodes(double pmax, int npw, double* ws, double* ts)
{
    const double pi = 4.0 * std::atan(1.0);
    int half = npw / 2;
    double h = pmax / half;
    double w = h / (2.0 * std::sqrt(pi));
    for (int j = 0; j < half; ++j) {
        ts[j] = (j + 0.5) * h;
        ws[j] = w * std::exp(-ts[j] * ts[j] / 4);
        ts[j + half] = -ts[j];
        ws[j + half] = ws[j];
    }
}

// The 3D ordering of the same nodes (get_pwnodes, g3drouts.f:2669-2690):
// ascending, ts_j = (j + 1/2) h for j = -npw/2 .. npw/2-1.
void planewave_nodes3(double pmax, int npw, double* ws, double* ts)
{
    const double pi = 4.0 * std::atan(1.0);
    int half = npw / 2;
    double h = pmax / half;
    double w = h / (2.0 * std::sqrt(pi));
    for (int j = -half; j < half; ++j) {
        ts[j + half] = (j + 0.5) * h;
        ws[j + half] = w * std::exp(-ts[j + half] * ts[j + half] / 4);
    }
}

namespace {

template <int D>
struct Builder {
    static constexpr int CH = TreeD<D>::CH, NC = TreeD<D>::NC;
    TreeD<D>& T;
    // raw views of the per-box arrays, refreshed whenever they grow
    int *lev, *par, *nch, *ch, *ncol, *col, *flag;
    int* ic[D];
    double* cen;
    uint64_t* pre;
    explicit Builder(TreeD<D>& t) : T(t) { view(); }

    void view()
    {
        lev = T.lev.data(), par = T.par.data(), nch = T.nch.data(), ch = T.ch.data();
        ncol = T.ncol.data(), col = T.col.data(), flag = T.flag.data();
        cen = T.cen.data(), pre = T.pre.data();
        for (int a = 0; a < D; ++a) ic[a] = T.ic[a].data();
    }
    // append k boxes with no children, no colleagues and flag 0
    int grow(int k)
    {
        const size_t n = T.lev.size(), m = n + k;
        T.lev.resize(m);
        T.par.resize(m);
        T.nch.resize(m, 0);
        T.ch.resize(CH * m, -1);
        T.ncol.resize(m, 0);
        T.col.resize(NC * m, -1);
        T.flag.resize(m, 0);
        T.cen.resize(D * m);
        T.pre.resize(m);
        for (int a = 0; a < D; ++a) T.ic[a].resize(m);
        view();
        return (int)n;
    }

    // The CH children of box b at first..first+CH-1: child j sits at
    // c +- bs/2 on axis a by bit (j >> a) & 1, bs the child size
    // (tree_refine_boxes, tree_routs4.f:28-44, tree_routs3d.f:28-46).
    void set_children(int b, int first)
    {
        const int l = lev[b];
        const double bs = T.bsz[l + 1];
        for (int j = 0; j < CH; ++j) {
            const int c = first + j;
            lev[c] = l + 1;
            par[c] = b;
            for (int a = 0; a < D; ++a) {
                const int bit = (j >> a) & 1;
                cen[D * c + a] = cen[D * b + a] + (bit ? bs / 2 : -bs / 2);
                ic[a][c] = 2 * ic[a][b] + bit;
            }
            pre[c] = (pre[b] << D) | (uint64_t)j;
            ch[CH * b + j] = c;
        }
        nch[b] = CH;
    }

    // coarse box c (level l) and fine box f (level l+1) share a point
    bool near_finer(int c, int f) const
    {
        for (int a = 0; a < D; ++a) {
            int d = ic[a][f] - 2 * ic[a][c];
            if (d < -1 || d > 2) return false;
        }
        return true;
    }

    // Same-level neighbours (self included), found among the children of
    // the parent's neighbours in the reference's order (computecoll,
    // tree_routs4.f:100-211, tree_routs3d.f:108-225).  A child of the
    // parent's neighbour at offset o in {-1,0,1} on an axis, with bit c on
    // that axis, touches a box with bit t there iff |2o + c - t| <= 1: any c
    // for o = 0, c = 0 and t = 1 for o = 1, c = 1 and t = 0 for o = -1.
    // fam[o][t] is the CH-bit set of such children for the offset vector
    // o (base-3 digits, x first) and sibling t.
    struct Fam {
        unsigned short m[NC][CH];
        Fam()
        {
            static const unsigned char ok[3][2] = {{2, 0}, {3, 3}, {0, 1}};
            for (int o = 0; o < NC; ++o)
                for (int t = 0; t < CH; ++t) {
                    unsigned s = 0;
                    for (int j = 0; j < CH; ++j) {
                        bool in = true;
                        for (int a = 0, oo = o; a < D; ++a, oo /= 3)
                            in = in && ((ok[oo % 3][(t >> a) & 1] >> ((j >> a) & 1)) & 1);
                        if (in) s |= 1u << j;
                    }
                    m[o][t] = (unsigned short)s;
                }
        }
    };
    static const Fam& fam()
    {
        static const Fam f;
        return f;
    }

    // colleagues of the CH children of dad, in one pass over dad's list
    void family_colleagues(int dad)
    {
        const Fam& F = fam();
        const int* kids = ch + CH * dad;
        int n[CH] = {};
        for (int i = 0; i < ncol[dad]; ++i) {
            const int jb = col[NC * dad + i];
            if (nch[jb] == 0) continue;
            int o = 0;
            for (int a = D - 1; a >= 0; --a) o = 3 * o + (ic[a][jb] - ic[a][dad] + 1);
            const int* cj = ch + CH * jb;
            for (int t = 0; t < CH; ++t) {
                int* out = col + NC * kids[t];
                for (unsigned m = F.m[o][t]; m; m &= m - 1) out[n[t]++] = cj[__builtin_ctz(m)];
            }
        }
        for (int t = 0; t < CH; ++t) {
            const int b = kids[t];
            for (int i = n[t]; i < ncol[b]; ++i) col[NC * b + i] = -1;
            ncol[b] = n[t];
        }
    }

    // colleagues of one box (its siblings may not all need them)
    void colleagues(int b)
    {
        if (lev[b] == 0) {
            ncol[b] = 1;
            col[NC * b] = b;
            return;
        }
        const Fam& F = fam();
        const int dad = par[b];
        int t = 0;
        for (int a = 0; a < D; ++a) t |= (ic[a][b] & 1) << a;
        int n = 0;
        int* out = col + NC * b;
        for (int i = 0; i < ncol[dad]; ++i) {
            const int jb = col[NC * dad + i];
            if (nch[jb] == 0) continue;
            int o = 0;
            for (int a = D - 1; a >= 0; --a) o = 3 * o + (ic[a][jb] - ic[a][dad] + 1);
            const int* cj = ch + CH * jb;
            for (unsigned m = F.m[o][t]; m; m &= m - 1) out[n++] = cj[__builtin_ctz(m)];
        }
        for (int i = n; i < ncol[b]; ++i) out[i] = -1;
        ncol[b] = n;
    }

    // lists of the boxes of levels lo+1..hi from those of level lo (lo >= 0;
    // the root's list is set at lo = 0)
    void colleagues_between(int lo, int hi)
    {
        if (lo <= 0) {
            lo = 0;
            ncol[0] = 1;
            col[0] = 0;
        }
        for (int l = lo; l < std::min(T.nlevels, hi); ++l) {
            const int b0 = T.laddr[2 * l];
            par_range(T.laddr[2 * l + 1] - b0 + 1, 512, [&](long i0, long i1) {
                for (long i = i0; i < i1; ++i)
                    if (nch[b0 + i] > 0) family_colleagues(b0 + (int)i);
            });
        }
    }
    void all_colleagues(int maxlev) { colleagues_between(0, maxlev); }

    // Refine the boxes of [b0,b1] whose flag is positive; children of flag-1
    // boxes are marked 3 (still to be checked), of flag-2 boxes 0
    // (tree_refine_boxes_flag, tree_routs4.f:334-388, tree_routs3d.f:345-410).
    void refine_flagged(int b0, int b1)
    {
        int k = 0;
        for (int b = b0; b <= b1; ++b) k += flag[b] > 0;
        if (k == 0) return;
        int next = grow(CH * k);
        for (int b = b0; b <= b1; ++b) {
            if (flag[b] <= 0) continue;
            set_children(b, next);
            const int f = flag[b] == 1 ? 3 : 0;
            for (int j = 0; j < CH; ++j) flag[next + j] = f;
            next += CH;
        }
    }

    // The reference's final numbering (pts_tree_reorg, pts_tree2d4.f:1001-1150,
    // pts_tree3d.f:1172-1329, which runs after each restriction pass): every
    // level lists the boxes of the adaptive tree, then those the first pass
    // appended to it, then those of the second pass, each group in creation
    // order.  The passes run on creation indices; this one permutation then
    // renumbers every per-box array and every stored box index (parents,
    // children, colleague lists).
    // The permutation moves whole runs of boxes (a level's group of one
    // pass, in creation order) to contiguous places.
    struct Run {
        int from, to, n, l;  // first old index, first new index, count, level
    };
    template <class V>
    static void permute(V& v, int width, const std::vector<Run>& runs)
    {
        using E = typename V::value_type;
        thread_local std::vector<char> scratch;
        if (scratch.size() < v.size() * sizeof(E)) scratch.resize(v.size() * sizeof(E));
        E* old = (E*)scratch.data();
        std::memcpy(old, v.data(), v.size() * sizeof(E));
        for (const Run& r : runs)
            std::memcpy(v.data() + (size_t)width * r.to, old + (size_t)width * r.from,
                        sizeof(E) * width * (size_t)r.n);
    }
    // colleague lists of the levels up to `valid`, in the new order; the
    // deeper ones are stale (recomputed before anything reads them): cleared
    void permute_col(const std::vector<Run>& runs, int valid)
    {
        thread_local std::vector<int> scratch;
        if (scratch.size() < T.col.size()) scratch.resize(T.col.size());
        for (const Run& r : runs)
            if (r.l <= valid)
                std::memcpy(scratch.data() + (size_t)NC * r.from, T.col.data() + (size_t)NC * r.from,
                            sizeof(int) * NC * (size_t)r.n);
        for (const Run& r : runs) {
            int* dst = T.col.data() + (size_t)NC * r.to;
            if (r.l <= valid) {
                std::memcpy(dst, scratch.data() + (size_t)NC * r.from, sizeof(int) * NC * (size_t)r.n);
            } else {
                std::fill(dst, dst + (size_t)NC * r.n, -1);
                std::fill(T.ncol.data() + r.to, T.ncol.data() + r.to + r.n, 0);
            }
        }
    }
    // colleague lists are remapped for levels up to `valid` (the deeper ones
    // are stale and recomputed before anything reads them); the restriction
    // flags are not kept
    void renumber(const std::vector<int>& t1, const std::vector<int>& t2, int valid)
    {
        const int n = T.nboxes();
        thread_local std::vector<Run> runs;
        runs.clear();
        int cur = 0;
        bool moved = false;
        for (int l = 0; l <= T.nlevels; ++l) {
            const int f0 = T.laddr[2 * l], f1 = T.laddr[2 * l + 1];
            T.laddr[2 * l] = cur;
            const int grp[3][2] = {{f0, f1}, {t1[2 * l], t1[2 * l + 1]}, {t2[2 * l], t2[2 * l + 1]}};
            for (const auto& g : grp) {
                const int b0 = g[0], b1 = g[1];
                if (b1 < b0) continue;
                moved = moved || b0 != cur;
                runs.push_back(Run{b0, cur, b1 - b0 + 1, l});
                cur += b1 - b0 + 1;
            }
            T.laddr[2 * l + 1] = cur - 1;
        }
        if (!moved) return;  // the identity
        // new index of every old one, shifted by one so that -1 maps to -1
        std::vector<int>& map = T.map;
        map.resize(n + 1);
        map[0] = -1;
        for (const Run& r : runs)
            for (int i = 0; i < r.n; ++i) map[r.from + i + 1] = r.to + i;
        // the arrays are independent: one task each (large trees only; the
        // tasks may run on other threads, hence the run table by pointer)
        const std::vector<Run>* rp = &runs;
        auto each = [&, rp](long i0, long i1) {
            const std::vector<Run>& R = *rp;
            for (long k = i0; k < i1; ++k) switch (k) {
                    case 0: permute(T.lev, 1, R); break;
                    case 1: permute(T.nch, 1, R); break;
                    case 2: permute(T.cen, D, R); break;
                    case 3: permute(T.pre, 1, R); break;
                    case 4: permute(T.par, 1, R); break;
                    case 5: permute(T.ch, CH, R); break;
                    case 6:  // the lists and their counts together
                        permute(T.ncol, 1, R);
                        permute_col(R, valid);
                        break;
                    default: permute(T.ic[k - 7], 1, R); break;
                }
        };
        if (n >= 32768)
            par_range(7 + D, 1, each);
        else
            each(0, 7 + D);
        std::fill(T.flag.begin(), T.flag.end(), 0);
        view();
        // stored indices (-1 for none: children of leaves, unused list slots)
        const int* mp = map.data() + 1;
        const long nc = valid >= 0 ? T.laddr[2 * std::min(valid, T.nlevels) + 1] + 1 : 0;
        par_range(n, 4096, [&](long i0, long i1) {
            for (long b = i0; b < i1; ++b) par[b] = mp[par[b]];
            for (long k = CH * i0; k < CH * i1; ++k) ch[k] = mp[ch[k]];
            for (long k = NC * i0; k < NC * std::min(i1, nc); ++k) col[k] = mp[col[k]];
        });
    }

    // Colleague lists after the first restriction pass, as recomputing them
    // from the root would give them, for levels up to hi.  The pass only adds
    // children to leaves; a box's list depends on its parent's list and on
    // the children of the parent's colleagues, so the lists that change are
    // those of the children of boxes touching a refined box (colleague
    // relation is symmetric).  The boxes the pass adds to a list of level l
    // have no children, so the lists of level l+1 do not depend on that
    // change.  Level by level: a list is read only after its own level is
    // done.
    void refresh_after_first_pass(int hi)
    {
        thread_local std::vector<char> mark;
        thread_local std::vector<int> todo;
        mark.assign(T.nboxes(), 0);
        for (int l = 1; l < std::min(hi, T.nlevels - 1); ++l) {
            todo.clear();
            for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; ++b) {
                if (flag[b] <= 0) continue;  // refined by the pass
                for (int i = 0; i < ncol[b]; ++i) {
                    const int c = col[NC * b + i];
                    if (nch[c] > 0 && !mark[c]) {
                        mark[c] = 1;
                        todo.push_back(c);
                    }
                }
            }
            const int* tp = todo.data();  // this thread's list, read by the workers
            par_range((long)todo.size(), 256, [&, tp](long i0, long i1) {
                for (long i = i0; i < i1; ++i) family_colleagues(tp[i]);
            });
        }
    }

    // Colleague lists after the second restriction pass, for levels up to
    // hi, from the first level l0 that pass refined (lists of levels <= l0
    // are valid).  The pass ran on lists it did not update, so any list of
    // level l+1 may lack children of a box the pass refined at level l, or
    // of one the pass created; a box refined at level l by the pass has flag
    // 1 afterwards, and a created box with children was itself refined by
    // the pass.  A list of level l+1 is recomputed with its family, so the
    // families to recompute are those of the boxes (with children) in the
    // list of a box the pass refined at level l, colleagues being
    // symmetric.  Level by level, so that the lists of level l are final
    // when their families are read.
    void refresh_after_second_pass(int l0, int hi, const std::vector<int>& t1,
                                   const std::vector<int>& t2)
    {
        thread_local std::vector<char> mark;
        thread_local std::vector<int> todo;
        mark.assign(T.nboxes(), 0);
        for (int l = l0; l < hi; ++l) {
            todo.clear();
            const int grp[3][2] = {{T.laddr[2 * l], T.laddr[2 * l + 1]},
                                   {t1[2 * l], t1[2 * l + 1]},
                                   {t2[2 * l], t2[2 * l + 1]}};
            for (const auto& g : grp)
                for (int b = g[0]; b <= g[1]; ++b) {
                    if (flag[b] != 1) continue;
                    for (int i = 0; i < ncol[b]; ++i) {
                        const int c = col[NC * b + i];
                        if (nch[c] > 0 && !mark[c]) {
                            mark[c] = 1;
                            todo.push_back(c);
                        }
                    }
                }
            const int* tp = todo.data();  // this thread's list, read by the workers
            par_range((long)todo.size(), 256, [&, tp](long i0, long i1) {
                for (long i = i0; i < i1; ++i) family_colleagues(tp[i]);
            });
        }
    }

    // flag++ boxes with a touching grandchild-level box that has children are
    // marked for refinement (updateflags, tree_routs4.f:217-329,
    // tree_routs3d.f:226-344).
    void update_flags(int b0, int b1)
    {
        par_range(b1 - b0 + 1, 1024, [&](long i0, long i1) {
            for (long i = i0; i < i1; ++i) update_flag(b0 + (int)i);
        });
    }
    void update_flag(int b)
    {
        {
            if (flag[b] != 3) return;
            flag[b] = 0;
            for (int i = 0; i < ncol[b] && flag[b] == 0; ++i) {
                int jb = col[NC * b + i];
                if (nch[jb] == 0) continue;  // no children to look at
                for (int j = 0; j < CH; ++j) {
                    int kb = ch[CH * jb + j];
                    if (kb >= 0 && nch[kb] > 0 && near_finer(b, kb)) {
                        flag[b] = 1;
                        break;
                    }
                }
            }
        }
    }

    // Level restriction, pts_tree_fix_lr (pts_tree2d4.f:693-996,
    // pts_tree3d.f:855-1171), on colleague lists valid through level L-2 (the
    // deepest the procedure reads).  Returns the deepest level whose
    // colleague lists are valid afterwards: at least min(want, L-2).
    int fix_level_restriction(int want)
    {
        const int L = T.nlevels;
        std::fill(T.flag.begin(), T.flag.end(), 0);
        // leaves touching a box two levels finer, seen from its parent
        // (the test depends on the parent only: once per refined box of
        // levels 1..L-1)
        // (flags are set to the same value by whichever iteration gets
        // there, so the iterations of a level may run concurrently)
        std::atomic<int> nflag{0};
        for (int l = L - 1; l >= 1; --l) {
            const int d0 = T.laddr[2 * l];
            par_range(T.laddr[2 * l + 1] - d0 + 1, 1024, [&](long i0, long i1) {
                int nf = 0;
                for (long i = i0; i < i1; ++i) {
                    const int dad = d0 + (int)i;
                    if (nch[dad] == 0) continue;
                    int gd = par[dad];
                    for (int k = 0; k < ncol[gd]; ++k) {
                        int jb = col[NC * gd + k];
                        if (nch[jb] == 0 && __atomic_load_n(&flag[jb], __ATOMIC_RELAXED) == 0 &&
                            near_finer(jb, dad)) {
                            __atomic_store_n(&flag[jb], 1, __ATOMIC_RELAXED);
                            ++nf;
                        }
                    }
                }
                if (nf) nflag.fetch_add(nf);
            });
        }
        if (nflag.load() == 0) return L - 2;  // already level restricted: nothing below changes
        // coarser leaves touching a flagged box, propagated upwards
        for (int l = L; l >= 1; --l) {
            const int b0 = T.laddr[2 * l];
            par_range(T.laddr[2 * l + 1] - b0 + 1, 1024, [&](long i0, long i1) {
                for (long i = i0; i < i1; ++i) {
                    const int b = b0 + (int)i;
                    if (flag[b] != 1 && flag[b] != 2) continue;
                    int dad = par[b];
                    for (int k = 0; k < ncol[dad]; ++k) {
                        int jb = col[NC * dad + k];
                        if (nch[jb] == 0 && __atomic_load_n(&flag[jb], __ATOMIC_RELAXED) == 0 &&
                            near_finer(jb, b))
                            __atomic_store_n(&flag[jb], 2, __ATOMIC_RELAXED);
                    }
                }
            });
        }
        // boxes appended to each level by the first (t1) and second (t2) pass;
        // a level's boxes are its range in laddr, then t1, then t2
        std::vector<int>& t1 = T.tail;
        std::vector<int>& t2 = T.tail2;
        t1.assign(2 * (L + 1), 0);
        t2.assign(2 * (L + 1), 0);
        for (int l = 0; l <= L; ++l) t1[2 * l + 1] = t2[2 * l + 1] = -1;
        for (int l = 1; l <= L - 2; ++l) {
            t1[2 * (l + 1)] = T.nboxes();
            refine_flagged(T.laddr[2 * l], T.laddr[2 * l + 1]);
            t1[2 * (l + 1) + 1] = T.nboxes() - 1;
        }
        refresh_after_first_pass(L - 2);
        int n3 = 0;
        for (int b = 0, n = T.nboxes(); b < n; ++b) {
            if (flag[b] != 3) flag[b] = 0;
            else ++n3;
        }
        int valid = L - 2, first2 = -1;
        if (n3 > 0) {  // flag++ boxes: the second pass
            for (int l = 2; l <= L - 2; ++l) {
                update_flags(T.laddr[2 * l], T.laddr[2 * l + 1]);
                update_flags(t1[2 * l], t1[2 * l + 1]);
                update_flags(t2[2 * l], t2[2 * l + 1]);
                t2[2 * (l + 1)] = T.nboxes();
                refine_flagged(T.laddr[2 * l], T.laddr[2 * l + 1]);
                refine_flagged(t1[2 * l], t1[2 * l + 1]);
                refine_flagged(t2[2 * l], t2[2 * l + 1]);
                t2[2 * (l + 1) + 1] = T.nboxes() - 1;
                const int c0 = t2[2 * (l + 1)];
                // the new boxes get lists; the lists of the boxes of level l+1
                // that they touch are not updated (the reference does not
                // either, and the pass reads them in that state)
                if (T.nboxes() > c0 && first2 < 0) first2 = l;
                par_range(T.nboxes() - c0, 1024, [&](long i0, long i1) {
                    for (long i = i0; i < i1; ++i) colleagues(c0 + (int)i);
                });
            }
        }
        if (first2 >= 0) {
            const int hi = std::min(want, L - 2);
            if (hi > first2) refresh_after_second_pass(first2, hi, t1, t2);
            valid = std::max(hi, first2);
        }
        renumber(t1, t2, valid);
        return valid;
    }
};

}  // namespace

template <int D>
void complete_colleagues(TreeD<D>& T)
{
    Builder<D> B(T);
    for (int l = std::max(T.coll_levels, 0); l < T.nlevels; ++l)
        for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; ++b)
            if (B.nch[b] > 0) B.family_colleagues(b);
    T.coll_levels = T.nlevels;
}

template <int D>
void build_tree(TreeD<D>& T, const std::vector<LeafRun>& leaves, int nlmax, int npwlevel,
                double bs0, const double* cen0)
{
    constexpr int CH = TreeD<D>::CH, NC = TreeD<D>::NC;
    // storage is reused from call to call: clear, keep the capacity
    for (auto* v : {&T.laddr, &T.lev, &T.par, &T.nch, &T.ch, &T.ncol, &T.col, &T.flag,
                    &T.leaf_of_rank, &T.rank_of_box, &T.rank_lo, &T.rank_hi})
        v->clear();
    for (int a = 0; a < D; ++a) T.ic[a].clear();
    T.bsz.clear();
    T.cen.clear();
    T.pre.clear();
    T.coll_levels = -1;
    T.nlmax = nlmax;
    T.npwlevel = npwlevel;
    T.bs0 = bs0;
    for (int a = 0; a < D; ++a) T.cen0[a] = cen0[a];
    {
        // the adaptive tree has at most CH boxes per distinct proper prefix
        // of the leaves; the restriction pass adds a fraction more
        size_t cap = CH * leaves.size() + 64;
        cap += cap / 2;
        T.lev.reserve(cap); T.par.reserve(cap); T.nch.reserve(cap); T.ch.reserve(CH * cap);
        T.ncol.reserve(cap); T.col.reserve(NC * cap); T.flag.reserve(cap); T.cen.reserve(D * cap);
        T.pre.reserve(cap);
        for (int a = 0; a < D; ++a) T.ic[a].reserve(cap);
    }
    // the pool, for trees large enough to repay its forks and when no other
    // transform holds it
    Par par(leaves.size() >= 20000);
    struct Scope {
        Par* prev;
        explicit Scope(Par* p) : prev(t_par) { t_par = p; }
        ~Scope() { t_par = prev; }
    } scope(&par);
    Builder<D> B(T);
    // The refined boxes are the proper prefixes of the leaves.  In Morton
    // order a leaf shares its ancestors down to the level where its prefix
    // first differs from the previous leaf's and brings new refined boxes
    // below that; the most recent refined box of a level is the parent of
    // the next level's new ones.  Level l+1 holds the children of the
    // refined boxes of level l in their order, so one pass counts the
    // refined boxes per level (the level sizes) and a second creates every
    // box at its final index: the numbering of the level-by-level
    // construction.
    const int nleaf = (int)leaves.size();
    auto first_new = [&](int i) {  // the shallowest level with a box new for leaf i
        if (i == 0) return 0;
        const LeafRun& p = leaves[i - 1];
        const LeafRun& q = leaves[i];
        const int m = std::min(p.level, q.level);
        const uint64_t x = (p.key >> (D * (p.level - m))) ^ (q.key >> (D * (q.level - m)));
        if (x == 0) return m + 1;  // not two disjoint leaves: nothing shared below m
        return m - (63 - __builtin_clzll(x)) / D;
    };
    int maxlev = 0;
    for (const LeafRun& r : leaves) maxlev = std::max(maxlev, r.level);
    std::vector<int> cnt(maxlev + 1, 0), first(maxlev + 2, 0), cur(maxlev + 1, 0);
    for (int i = 0; i < nleaf; ++i)
        for (int l = first_new(i); l < leaves[i].level; ++l) ++cnt[l];
    int total = 1;
    for (int l = 0; l <= maxlev; ++l) {
        first[l + 1] = total;
        total += CH * cnt[l];
    }
    int nlev = 0;
    while (nlev < maxlev && cnt[nlev] > 0) ++nlev;
    T.bsz.push_back(bs0);
    for (int l = 0; l < nlev; ++l) T.bsz.push_back(T.bsz[l] / 2);
    B.grow(total);
    B.lev[0] = 0;
    B.par[0] = -1;
    for (int a = 0; a < D; ++a) {
        B.cen[a] = cen0[a];
        B.ic[a][0] = 0;
    }
    B.pre[0] = 0;
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int i = 0; i < nleaf; ++i) {
        const LeafRun& r = leaves[i];
        for (int l = first_new(i); l < r.level; ++l) {
            const int digit = (int)((r.key >> (D * (r.level - l))) & (CH - 1));
            const int b = l == 0 ? 0 : first[l] + CH * cur[l - 1] + digit;
            cur[l] = cnt[l]++;
            B.set_children(b, first[l + 1] + CH * cur[l]);
        }
    }
    T.laddr.assign(2 * (nlev + 1), 0);
    for (int l = 0; l <= nlev; ++l) {
        T.laddr[2 * l] = first[l];
        T.laddr[2 * l + 1] = first[l + 1] - 1;
    }
    T.nlevels = nlev;
    // Colleague lists: the level restriction reads them down to level
    // nlevels-2; the transform itself only at and above the cutoff level
    // (m2l, iflocal, near-field lists).  Deeper lists are filled on demand
    // by complete_colleagues().
    const int L = T.nlevels;
    B.all_colleagues(L - 2);
    int valid = std::max(L - 2, 0);
    const int keep = std::min(std::max(npwlevel, 0), L);
    if (L >= 2) valid = B.fix_level_restriction(keep);
    if (valid < keep) {
        B.colleagues_between(valid, keep);
        valid = keep;
    }
    T.coll_levels = valid;

    // leaves in Morton order (depth-first, children in digit order) and the
    // interval of leaf ranks of every box: the leaf counts bottom-up, then
    // each child's first rank from its parent's and its elder siblings'
    const int nbox = T.nboxes();
    T.rank_of_box.assign(nbox, -1);
    T.rank_lo.assign(nbox, 0);
    T.rank_hi.assign(nbox, 0);
    int* lo = T.rank_lo.data();
    int* hi = T.rank_hi.data();  // the leaf count first, the interval's end after
    const int* nc = T.nch.data();
    const int* chd = T.ch.data();
    for (int lv = T.nlevels; lv >= 0; --lv)
        for (int b = T.laddr[2 * lv]; b <= T.laddr[2 * lv + 1]; ++b) {
            int k = 1;
            if (nc[b] > 0) {
                k = 0;
                for (int j = 0; j < CH; ++j) k += hi[chd[CH * b + j]];
            }
            hi[b] = k;
        }
    T.leaf_of_rank.resize(hi[0]);
    lo[0] = 0;
    for (int lv = 0; lv <= T.nlevels; ++lv)
        for (int b = T.laddr[2 * lv]; b <= T.laddr[2 * lv + 1]; ++b) {
            const int r = lo[b];
            if (nc[b] > 0) {
                int next = r;
                for (int j = 0; j < CH; ++j) {
                    const int c = chd[CH * b + j];
                    lo[c] = next;
                    next += hi[c];
                }
            } else {
                T.rank_of_box[b] = r;
                T.leaf_of_rank[r] = b;
            }
            hi[b] += r;
        }
}

// memcpy of n bytes split over the tree builder's worker pool (in slices of
// at least 1 MiB); sequential when the pool is busy or absent
void parallel_copy(void* dst, const void* src, size_t n)
{
    constexpr size_t kSlice = (size_t)1 << 20;
    if (n < 2 * kSlice) {
        std::memcpy(dst, src, n);
        return;
    }
    Par par(true);
    const long slices = (long)((n + kSlice - 1) / kSlice);
    auto f = [&](long i0, long i1) {
        const size_t a = (size_t)i0 * kSlice, b = std::min(n, (size_t)i1 * kSlice);
        std::memcpy((char*)dst + a, (const char*)src + a, b - a);
    };
    par.range(slices, 1, f);
}

template void build_tree<2>(TreeD<2>&, const std::vector<LeafRun>&, int, int, double,
                            const double*);
template void build_tree<3>(TreeD<3>&, const std::vector<LeafRun>&, int, int, double,
                            const double*);
template void complete_colleagues<2>(TreeD<2>&);
template void complete_colleagues<3>(TreeD<3>&);

}  // namespace fgtpw
