// Host side of the quadtree: term counts, plane-wave nodes and the box-level
// tree built from the GPU's list of non-empty leaves.
//
// The GPU sorts the sources by their full-depth Morton key and finds, per
// point, the level at which the reference's refinement rule stops
// (kernels2d.hip).  What comes back is one record per non-empty leaf of the
// adaptive tree, in Morton order.  From those records this file rebuilds the
// adaptive tree with the reference's box numbering (level by level, children
// of refined boxes in parent order; pts_tree_build, pts_tree2d4.f:363-590),
// then applies the reference's level-restriction procedure
// (pts_tree_fix_lr, pts_tree2d4.f:693-996) step for step, because the set of
// boxes it produces is part of the result the transform must match.
//
// Adjacency.  The reference decides "touching" and "colleague" with FP64
// centre distances against 1.05x the half-sum of the box sizes.  Centres of
// one level differ by whole multiples of the box size (up to rounding
// ~1e-16 relative), so those tests are exactly integer tests on the box
// coordinates (ix, iy) of each level, which is what is used here:
//   same level:          |dix| <= 1 and |diy| <= 1
//   levels l and l+1:    the finer box's coordinate lies in
//                        [2 ic - 1, 2 ic + 2] on both axes.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "fgtpw_impl.hpp"

namespace fgtpw {

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

// Hermite/Taylor order for boxes of side bsize, g2dhlterms (g2dterms.f:19-66):
// the smallest p >= int(e r^2)+1 with d*rk*p^(-1/4)*rp^p/(1-rp) < eps,
// r = bsize/sqrt(2 delta), rp = r sqrt(e/p), rk = 1.09 (a single-precision
// constant in the reference) * (2 pi)^(-1/4).
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

// Plane-wave length, g2dpwterms (g2dterms.f:71-133): Fourier truncation at
// pmax = 2 sqrt(ln 1/eps) and step h = 2 pi/(rmax + sqrt(ln 1/eps)), so that
// the aliased images of the Gaussian stay below eps; npw = 2 int(pmax/h).
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

namespace {

struct Builder {
    Tree2D& T;
    // raw views of the per-box arrays, refreshed whenever they grow
    int *lev, *par, *nch, *ch, *ncol, *col, *flag, *ix, *iy;
    double* cen;
    uint64_t* pre;
    explicit Builder(Tree2D& t) : T(t) { view(); }

    void view()
    {
        lev = T.lev.data(), par = T.par.data(), nch = T.nch.data(), ch = T.ch.data();
        ncol = T.ncol.data(), col = T.col.data(), flag = T.flag.data(), ix = T.ix.data();
        iy = T.iy.data(), cen = T.cen.data(), pre = T.pre.data();
    }
    // append k boxes with no children, no colleagues and flag 0
    int grow(int k)
    {
        const size_t n = T.lev.size(), m = n + k;
        T.lev.resize(m);
        T.par.resize(m);
        T.nch.resize(m, 0);
        T.ch.resize(4 * m, -1);
        T.ncol.resize(m, 0);
        T.col.resize(9 * m, -1);
        T.flag.resize(m, 0);
        T.cen.resize(2 * m);
        T.pre.resize(m);
        T.ix.resize(m);
        T.iy.resize(m);
        view();
        return (int)n;
    }

    // Four children of box b at first..first+3; child j sits at
    // c + (-1)^(j+1) bs/2 in x and at -bs/2 (j < 2) or +bs/2 in y, bs the
    // child size (tree_refine_boxes, tree_routs4.f:28-44).
    void set_children(int b, int first)
    {
        const int l = lev[b];
        const double bs = T.bsz[l + 1];
        for (int j = 0; j < 4; ++j) {
            const int c = first + j;
            lev[c] = l + 1;
            par[c] = b;
            cen[2 * c] = cen[2 * b] + ((j & 1) ? bs / 2 : -bs / 2);
            cen[2 * c + 1] = cen[2 * b + 1] + ((j < 2) ? -bs / 2 : bs / 2);
            pre[c] = (pre[b] << 2) | (uint64_t)j;
            ix[c] = 2 * ix[b] + (j & 1);
            iy[c] = 2 * iy[b] + (j >> 1);
            ch[4 * b + j] = c;
        }
        nch[b] = 4;
    }

    // coarse box c (level l) and fine box f (level l+1) share a point
    bool near_finer(int c, int f) const
    {
        int dx = ix[f] - 2 * ix[c], dy = iy[f] - 2 * iy[c];
        return dx >= -1 && dx <= 2 && dy >= -1 && dy <= 2;
    }

    // Same-level neighbours (self included), searched among the children of
    // the parent's neighbours in the reference's order (computecoll,
    // tree_routs4.f:100-211).
    void colleagues(int b)
    {
        if (lev[b] == 0) {
            ncol[b] = 1;
            col[9 * b] = b;
            return;
        }
        // A child of the parent's neighbour jb, at offset o = (jb - dad) in
        // {-1,0,1} per axis, with bit c on that axis, is a neighbour of b
        // (bit t) iff |2o + c - t| <= 1: any c for o = 0, c = 0 and t = 1
        // for o = 1, c = 1 and t = 0 for o = -1.  ok[o+1][t] is the mask of
        // the allowed c (bit 0: c = 0, bit 1: c = 1).
        static const unsigned char ok[3][2] = {{2, 0}, {3, 3}, {0, 1}};
        const int dad = par[b], tx = ix[b] & 1, ty = iy[b] & 1;
        const int px = ix[dad], py = iy[dad];
        int n = 0;
        int* out = col + 9 * b;
        for (int i = 0; i < ncol[dad]; ++i) {
            const int jb = col[9 * dad + i];
            if (nch[jb] == 0) continue;
            const int mx = ok[ix[jb] - px + 1][tx], my = ok[iy[jb] - py + 1][ty];
            const int* cj = ch + 4 * jb;
            for (int j = 0; j < 4; ++j)
                if (((mx >> (j & 1)) & 1) && ((my >> (j >> 1)) & 1)) out[n++] = cj[j];
        }
        for (int i = n; i < ncol[b]; ++i) out[i] = -1;
        ncol[b] = n;
    }

    // colleagues() for the four children of dad at once: one pass over the
    // parent's neighbours; fam[o][t] is the 4-bit set of children of a
    // neighbour at offset o = 3 (oy+1) + (ox+1) that touch sibling t.
    void family_colleagues(int dad)
    {
        static unsigned char fam[9][4];
        static bool init = false;
        if (!init) {
            static const unsigned char ok[3][2] = {{2, 0}, {3, 3}, {0, 1}};
            for (int o = 0; o < 9; ++o)
                for (int t = 0; t < 4; ++t) {
                    int mx = ok[o % 3][t & 1], my = ok[o / 3][t >> 1], m = 0;
                    for (int j = 0; j < 4; ++j)
                        if (((mx >> (j & 1)) & 1) && ((my >> (j >> 1)) & 1)) m |= 1 << j;
                    fam[o][t] = (unsigned char)m;
                }
            init = true;
        }
        const int px = ix[dad], py = iy[dad];
        const int* kids = ch + 4 * dad;
        int n[4] = {0, 0, 0, 0};
        for (int i = 0; i < ncol[dad]; ++i) {
            const int jb = col[9 * dad + i];
            if (nch[jb] == 0) continue;
            const int o = 3 * (iy[jb] - py + 1) + (ix[jb] - px + 1);
            const int* cj = ch + 4 * jb;
            for (int t = 0; t < 4; ++t) {
                int* out = col + 9 * kids[t];
                for (unsigned m = fam[o][t]; m; m &= m - 1) out[n[t]++] = cj[__builtin_ctz(m)];
            }
        }
        for (int t = 0; t < 4; ++t) {
            const int b = kids[t];
            for (int i = n[t]; i < ncol[b]; ++i) col[9 * b + i] = -1;
            ncol[b] = n[t];
        }
    }

    void all_colleagues(int maxlev)
    {
        ncol[0] = 1;
        col[0] = 0;
        for (int l = 0; l < std::min(T.nlevels, maxlev); ++l)
            for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; ++b)
                if (nch[b] > 0) family_colleagues(b);
    }

    // Refine the boxes of [b0,b1] whose flag is positive; children of flag-1
    // boxes are marked 3 (still to be checked), of flag-2 boxes 0
    // (tree_refine_boxes_flag, tree_routs4.f:334-388).
    void refine_flagged(int b0, int b1)
    {
        int k = 0;
        for (int b = b0; b <= b1; ++b) k += flag[b] > 0;
        if (k == 0) return;
        int next = grow(4 * k);
        for (int b = b0; b <= b1; ++b) {
            if (flag[b] <= 0) continue;
            set_children(b, next);
            const int f = flag[b] == 1 ? 3 : 0;
            for (int j = 0; j < 4; ++j) flag[next + j] = f;
            next += 4;
        }
    }

    // Renumber so that each level lists its original boxes, then the boxes
    // appended to it (pts_tree_reorg, pts_tree2d4.f:1001-1150).
    template <class V>
    void permute(V& v, int width, const std::vector<int>& map)
    {
        using E = typename V::value_type;
        std::vector<char>& scratch = T.scratch;
        if (scratch.size() < v.size() * sizeof(E)) scratch.resize(v.size() * sizeof(E));
        E* old = (E*)scratch.data();
        std::memcpy(old, v.data(), v.size() * sizeof(E));
        E* dst = v.data();
        const int* mp = map.data();
        for (size_t b = 0; b < map.size(); ++b)
            for (int k = 0; k < width; ++k) dst[(size_t)width * mp[b] + k] = old[(size_t)width * b + k];
    }
    void reorg(const std::vector<int>& tail)
    {
        const int n = T.nboxes();
        std::vector<int>& map = T.map;
        map.resize(n);
        for (int l = 0; l <= 1 && l <= T.nlevels; ++l)
            for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; ++b) map[b] = b;
        int cur = T.laddr[4];
        for (int l = 2; l <= T.nlevels; ++l) {
            int f0 = T.laddr[2 * l], f1 = T.laddr[2 * l + 1];
            T.laddr[2 * l] = cur;
            for (int b = f0; b <= f1; ++b) map[b] = cur++;
            for (int b = tail[2 * l]; b <= tail[2 * l + 1]; ++b) map[b] = cur++;
            T.laddr[2 * l + 1] = cur - 1;
        }
        permute(T.lev, 1, map);
        permute(T.nch, 1, map);
        permute(T.cen, 2, map);
        permute(T.flag, 1, map);
        permute(T.pre, 1, map);
        permute(T.ix, 1, map);
        permute(T.iy, 1, map);
        permute(T.par, 1, map);
        permute(T.ch, 4, map);
        view();
        const int* mp = map.data();
        for (int b = 0; b < n; ++b)
            if (par[b] >= 0) par[b] = mp[par[b]];
        for (int i = 0; i < 4 * n; ++i)
            if (ch[i] >= 0) ch[i] = mp[ch[i]];
    }

    // flag++ boxes with a touching grandchild-level box that has children are
    // marked for refinement (updateflags, tree_routs4.f:217-329).
    void update_flags(int b0, int b1)
    {
        for (int b = b0; b <= b1; ++b) {
            if (flag[b] != 3) continue;
            flag[b] = 0;
            for (int i = 0; i < ncol[b] && flag[b] == 0; ++i) {
                int jb = col[9 * b + i];
                for (int j = 0; j < 4; ++j) {
                    int kb = ch[4 * jb + j];
                    if (kb >= 0 && nch[kb] > 0 && near_finer(b, kb)) {
                        flag[b] = 1;
                        break;
                    }
                }
            }
        }
    }

    // Level restriction, pts_tree_fix_lr (pts_tree2d4.f:693-996).  Returns
    // the deepest level whose colleague lists are still valid afterwards.
    int fix_level_restriction()
    {
        const int L = T.nlevels;
        std::fill(T.flag.begin(), T.flag.end(), 0);
        // leaves touching a box two levels finer, seen from its parent
        // (the test depends on the parent only: once per refined box of
        // levels 1..L-1)
        int nflag = 0;
        for (int l = L - 1; l >= 1; --l)
            for (int dad = T.laddr[2 * l]; dad <= T.laddr[2 * l + 1]; ++dad) {
                if (nch[dad] == 0) continue;
                int gd = par[dad];
                for (int i = 0; i < ncol[gd]; ++i) {
                    int jb = col[9 * gd + i];
                    if (nch[jb] == 0 && flag[jb] == 0 && near_finer(jb, dad)) {
                        flag[jb] = 1;
                        ++nflag;
                    }
                }
            }
        if (nflag == 0) return L - 1;  // already level restricted: nothing below changes
        // coarser leaves touching a flagged box, propagated upwards
        for (int l = L; l >= 1; --l)
            for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; ++b) {
                if (flag[b] != 1 && flag[b] != 2) continue;
                int dad = par[b];
                for (int i = 0; i < ncol[dad]; ++i) {
                    int jb = col[9 * dad + i];
                    if (nch[jb] == 0 && flag[jb] == 0 && near_finer(jb, b)) flag[jb] = 2;
                }
            }
        std::vector<int>& tail = T.tail;
        tail.assign(2 * (L + 1), 0);
        for (int l = 0; l <= L; ++l) tail[2 * l + 1] = -1;
        for (int l = 1; l <= L - 2; ++l) {
            tail[2 * (l + 1)] = T.nboxes();
            refine_flagged(T.laddr[2 * l], T.laddr[2 * l + 1]);
            tail[2 * (l + 1) + 1] = T.nboxes() - 1;
        }
        reorg(tail);
        all_colleagues(L - 1);
        int n3 = 0;
        for (int b = 0, n = T.nboxes(); b < n; ++b) {
            if (flag[b] != 3) flag[b] = 0;
            else ++n3;
        }
        if (n3 == 0) return L - 1;  // no flag++ box: the second pass adds nothing
        for (int l = 0; l <= L; ++l) { tail[2 * l] = 0; tail[2 * l + 1] = -1; }
        for (int l = 2; l <= L - 2; ++l) {
            update_flags(T.laddr[2 * l], T.laddr[2 * l + 1]);
            update_flags(tail[2 * l], tail[2 * l + 1]);
            tail[2 * (l + 1)] = T.nboxes();
            refine_flagged(T.laddr[2 * l], T.laddr[2 * l + 1]);
            refine_flagged(tail[2 * l], tail[2 * l + 1]);
            tail[2 * (l + 1) + 1] = T.nboxes() - 1;
            for (int b = tail[2 * (l + 1)]; b <= tail[2 * (l + 1) + 1]; ++b) colleagues(b);
        }
        reorg(tail);
        return -1;
    }
};

}  // namespace

void complete_colleagues(Tree2D& T)
{
    Builder B(T);
    for (int l = std::max(T.coll_levels, 0); l < T.nlevels; ++l)
        for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; ++b)
            if (B.nch[b] > 0) B.family_colleagues(b);
    T.coll_levels = T.nlevels;
}

void build_tree2d(Tree2D& T, const std::vector<LeafRun>& leaves, int nlmax, int npwlevel,
                  double bs0, const double* cen0)
{
    // storage is reused from call to call: clear, keep the capacity
    for (auto* v : {&T.laddr, &T.lev, &T.par, &T.nch, &T.ch, &T.ncol, &T.col, &T.flag, &T.ix,
                    &T.iy, &T.leaf_of_rank, &T.rank_of_box, &T.rank_lo, &T.rank_hi})
        v->clear();
    T.bsz.clear();
    T.cen.clear();
    T.pre.clear();
    T.coll_levels = -1;
    T.nlmax = nlmax;
    T.npwlevel = npwlevel;
    T.bs0 = bs0;
    T.cen0[0] = cen0[0];
    T.cen0[1] = cen0[1];
    {
        // the adaptive tree has at most 4 boxes per distinct proper prefix of
        // the leaves; the restriction pass adds a fraction more
        size_t cap = 4 * leaves.size() + 64;
        cap += cap / 2;
        T.lev.reserve(cap); T.par.reserve(cap); T.nch.reserve(cap); T.ch.reserve(4 * cap);
        T.ncol.reserve(cap); T.col.reserve(9 * cap); T.flag.reserve(cap); T.cen.reserve(2 * cap);
        T.pre.reserve(cap); T.ix.reserve(cap); T.iy.reserve(cap);
    }
    Builder B(T);
    T.bsz.push_back(bs0);
    B.grow(1);
    B.lev[0] = 0;
    B.par[0] = -1;
    B.cen[0] = cen0[0];
    B.cen[1] = cen0[1];
    B.pre[0] = 0;
    B.ix[0] = B.iy[0] = 0;
    T.laddr = {0, 0};
    // frontier: boxes of the current level with their range of leaf records
    std::vector<int> fbox{0}, fa{0}, fe{(int)leaves.size()}, nb, na, ne;
    int l = 0;
    for (;;) {
        int nref = 0;
        for (size_t i = 0; i < fbox.size(); ++i)
            if (fa[i] < fe[i] && leaves[fa[i]].level > l) ++nref;
        if (nref == 0) break;
        T.bsz.push_back(T.bsz[l] / 2);
        const int first = B.grow(4 * nref);
        int c0 = first;
        nb.clear();
        na.clear();
        ne.clear();
        for (size_t i = 0; i < fbox.size(); ++i) {
            int a = fa[i], e = fe[i];
            if (!(a < e && leaves[a].level > l)) continue;
            B.set_children(fbox[i], c0);
            int pos = a;
            for (int d = 0; d < 4; ++d) {
                int s = pos;
                while (pos < e) {
                    const LeafRun& r = leaves[pos];
                    int digit = (int)((r.key >> (2 * (r.level - (l + 1)))) & 3);
                    if (digit != d) break;
                    ++pos;
                }
                nb.push_back(c0 + d);
                na.push_back(s);
                ne.push_back(pos);
            }
            c0 += 4;
        }
        T.laddr.push_back(first);
        T.laddr.push_back(T.nboxes() - 1);
        fbox.swap(nb);
        fa.swap(na);
        fe.swap(ne);
        ++l;
    }
    T.nlevels = l;
    // Colleague lists: the level restriction reads them down to level
    // nlevels-1; the transform itself only at and above the cutoff level
    // (m2l, iflocal, near-field lists).  Deeper lists are filled on demand
    // by complete_colleagues().
    B.all_colleagues(T.nlevels - 1);
    int valid = T.nlevels - 1;
    if (T.nlevels >= 2) valid = B.fix_level_restriction();
    const int keep = std::min(std::max(npwlevel, 0), T.nlevels);
    if (valid < keep) {
        B.all_colleagues(keep);
        valid = keep;
    }
    T.coll_levels = valid;

    // leaves in Morton order: depth-first, children in digit order
    int nbox = T.nboxes();
    T.rank_of_box.assign(nbox, -1);
    T.rank_lo.assign(nbox, 0);
    T.rank_hi.assign(nbox, 0);
    {
        std::vector<int>& stack = T.map;
        stack.assign(1, 0);
        while (!stack.empty()) {
            int b = stack.back();
            stack.pop_back();
            if (T.nch[b] == 0) {
                T.rank_of_box[b] = (int)T.leaf_of_rank.size();
                T.leaf_of_rank.push_back(b);
            } else {
                for (int j = 3; j >= 0; --j) stack.push_back(T.ch[4 * b + j]);
            }
        }
    }
    for (int lv = T.nlevels; lv >= 0; --lv)
        for (int b = T.laddr[2 * lv]; b <= T.laddr[2 * lv + 1]; ++b) {
            if (T.nch[b] == 0) {
                T.rank_lo[b] = T.rank_of_box[b];
                T.rank_hi[b] = T.rank_of_box[b] + 1;
            } else {
                int lo = 1 << 30, hi = -1;
                for (int j = 0; j < 4; ++j) {
                    int c = T.ch[4 * b + j];
                    lo = std::min(lo, T.rank_lo[c]);
                    hi = std::max(hi, T.rank_hi[c]);
                }
                T.rank_lo[b] = lo;
                T.rank_hi[b] = hi;
            }
        }
}

}  // namespace fgtpw
