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
    explicit Builder(Tree2D& t) : T(t) {}

    int add_box(int level, int parent, double cx, double cy, uint64_t pre)
    {
        int id = T.nboxes();
        T.lev.push_back(level);
        T.par.push_back(parent);
        T.nch.push_back(0);
        for (int j = 0; j < 4; ++j) T.ch.push_back(-1);
        T.ncol.push_back(0);
        for (int j = 0; j < 9; ++j) T.col.push_back(-1);
        T.flag.push_back(0);
        T.cen.push_back(cx);
        T.cen.push_back(cy);
        T.pre.push_back(pre);
        return id;
    }

    // Four children of box b, numbered consecutively from the current end;
    // child j sits at c + (-1)^(j+1) bs/2 in x and at -bs/2 (j < 2) or +bs/2
    // in y, bs the child size (tree_refine_boxes, tree_routs4.f:28-44).
    int add_children(int b)
    {
        int l = T.lev[b];
        double bs = T.bsz[l + 1];
        int first = T.nboxes();
        for (int j = 0; j < 4; ++j) {
            double hx = (j & 1) ? bs / 2 : -bs / 2;
            double hy = (j < 2) ? -bs / 2 : bs / 2;
            int c = add_box(l + 1, b, T.cen[2 * b] + hx, T.cen[2 * b + 1] + hy, (T.pre[b] << 2) | j);
            T.ch[4 * b + j] = c;
        }
        T.nch[b] = 4;
        return first;
    }

    bool touch(int a, int b, double tol) const
    {
        return std::fabs(T.cen[2 * a] - T.cen[2 * b]) <= tol &&
               std::fabs(T.cen[2 * a + 1] - T.cen[2 * b + 1]) <= tol;
    }

    // Same-level neighbours (self included), searched among the children of
    // the parent's neighbours, |dc| <= 1.05 bs per axis, with the 1.05 a
    // single-precision constant as in computecoll (tree_routs4.f:100-211).
    void colleagues(int b)
    {
        int l = T.lev[b];
        T.ncol[b] = 0;
        if (l == 0) {
            T.ncol[b] = 1;
            T.col[9 * b] = b;
            return;
        }
        int dad = T.par[b];
        double tol = (double)1.05f * T.bsz[l];
        for (int i = 0; i < T.ncol[dad]; ++i) {
            int jb = T.col[9 * dad + i];
            for (int j = 0; j < 4; ++j) {
                int kb = T.ch[4 * jb + j];
                if (kb >= 0 && touch(kb, b, tol)) T.col[9 * b + T.ncol[b]++] = kb;
            }
        }
    }

    void all_colleagues()
    {
        for (int l = 0; l <= T.nlevels; ++l)
            for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; ++b) colleagues(b);
    }

    // Refine the boxes of [b0,b1] whose flag is positive; children of flag-1
    // boxes are marked 3 (still to be checked), of flag-2 boxes 0
    // (tree_refine_boxes_flag, tree_routs4.f:334-388).
    void refine_flagged(int b0, int b1)
    {
        for (int b = b0; b <= b1; ++b) {
            if (T.flag[b] <= 0) continue;
            int first = add_children(b);
            int f = T.flag[b] == 1 ? 3 : 0;
            for (int j = 0; j < 4; ++j) T.flag[first + j] = f;
        }
    }

    // Renumber so that each level lists its original boxes, then the boxes
    // appended to it (pts_tree_reorg, pts_tree2d4.f:1001-1150).
    void reorg(const std::vector<int>& tail)
    {
        int n = T.nboxes();
        Tree2D old = T;
        std::vector<int> map(n);
        for (int l = 0; l <= 1 && l <= T.nlevels; ++l)
            for (int b = old.laddr[2 * l]; b <= old.laddr[2 * l + 1]; ++b) map[b] = b;
        int cur = old.laddr[4];
        for (int l = 2; l <= T.nlevels; ++l) {
            T.laddr[2 * l] = cur;
            for (int pass = 0; pass < 2; ++pass) {
                int a = pass ? tail[2 * l] : old.laddr[2 * l];
                int e = pass ? tail[2 * l + 1] : old.laddr[2 * l + 1];
                for (int b = a; b <= e; ++b) {
                    T.lev[cur] = old.lev[b];
                    T.nch[cur] = old.nch[b];
                    T.cen[2 * cur] = old.cen[2 * b];
                    T.cen[2 * cur + 1] = old.cen[2 * b + 1];
                    T.flag[cur] = old.flag[b];
                    T.pre[cur] = old.pre[b];
                    map[b] = cur++;
                }
            }
            T.laddr[2 * l + 1] = cur - 1;
        }
        for (int b = 0; b < n; ++b) {
            T.par[map[b]] = old.par[b] < 0 ? -1 : map[old.par[b]];
            for (int j = 0; j < 4; ++j)
                T.ch[4 * map[b] + j] = old.ch[4 * b + j] < 0 ? -1 : map[old.ch[4 * b + j]];
        }
    }

    // flag++ boxes with a touching grandchild-level box that has children are
    // marked for refinement (updateflags, tree_routs4.f:217-329).
    void update_flags(int curlev, int b0, int b1)
    {
        double tol = 1.05 * (T.bsz[curlev] + T.bsz[curlev + 1]) / 2.0;
        for (int b = b0; b <= b1; ++b) {
            if (T.flag[b] != 3) continue;
            T.flag[b] = 0;
            for (int i = 0; i < T.ncol[b] && T.flag[b] == 0; ++i) {
                int jb = T.col[9 * b + i];
                for (int j = 0; j < 4; ++j) {
                    int kb = T.ch[4 * jb + j];
                    if (kb >= 0 && T.nch[kb] > 0 && touch(kb, b, tol)) {
                        T.flag[b] = 1;
                        break;
                    }
                }
            }
        }
    }

    // Level restriction, pts_tree_fix_lr (pts_tree2d4.f:693-996).
    void fix_level_restriction()
    {
        const int L = T.nlevels;
        std::fill(T.flag.begin(), T.flag.end(), 0);
        // leaves touching a box two levels finer, seen from its parent
        for (int l = L; l >= 2; --l) {
            double tol = 1.05 * (T.bsz[l - 1] + T.bsz[l - 2]) / 2.0;
            for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; ++b) {
                int dad = T.par[b], gd = T.par[dad];
                for (int i = 0; i < T.ncol[gd]; ++i) {
                    int jb = T.col[9 * gd + i];
                    if (T.nch[jb] == 0 && T.flag[jb] == 0 && touch(jb, dad, tol)) T.flag[jb] = 1;
                }
            }
        }
        // coarser leaves touching a flagged box, propagated upwards
        for (int l = L; l >= 1; --l) {
            double tol = 1.05 * (T.bsz[l] + T.bsz[l - 1]) / 2.0;
            for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; ++b) {
                if (T.flag[b] != 1 && T.flag[b] != 2) continue;
                int dad = T.par[b];
                for (int i = 0; i < T.ncol[dad]; ++i) {
                    int jb = T.col[9 * dad + i];
                    if (T.nch[jb] == 0 && T.flag[jb] == 0 && touch(jb, b, tol)) T.flag[jb] = 2;
                }
            }
        }
        std::vector<int> tail(2 * (L + 1));
        for (int l = 0; l <= L; ++l) { tail[2 * l] = 0; tail[2 * l + 1] = -1; }
        for (int l = 1; l <= L - 2; ++l) {
            tail[2 * (l + 1)] = T.nboxes();
            refine_flagged(T.laddr[2 * l], T.laddr[2 * l + 1]);
            tail[2 * (l + 1) + 1] = T.nboxes() - 1;
        }
        reorg(tail);
        all_colleagues();
        for (int& f : T.flag)
            if (f != 3) f = 0;
        for (int l = 0; l <= L; ++l) { tail[2 * l] = 0; tail[2 * l + 1] = -1; }
        for (int l = 2; l <= L - 2; ++l) {
            update_flags(l, T.laddr[2 * l], T.laddr[2 * l + 1]);
            update_flags(l, tail[2 * l], tail[2 * l + 1]);
            tail[2 * (l + 1)] = T.nboxes();
            refine_flagged(T.laddr[2 * l], T.laddr[2 * l + 1]);
            refine_flagged(tail[2 * l], tail[2 * l + 1]);
            tail[2 * (l + 1) + 1] = T.nboxes() - 1;
            for (int b = tail[2 * (l + 1)]; b <= tail[2 * (l + 1) + 1]; ++b) colleagues(b);
        }
        reorg(tail);
        all_colleagues();
    }
};

}  // namespace

void build_tree2d(Tree2D& T, const std::vector<LeafRun>& leaves, int nlmax, int npwlevel,
                  double bs0, const double* cen0)
{
    T = Tree2D();
    T.nlmax = nlmax;
    T.npwlevel = npwlevel;
    T.bs0 = bs0;
    T.cen0[0] = cen0[0];
    T.cen0[1] = cen0[1];
    Builder B(T);
    T.bsz.push_back(bs0);
    B.add_box(0, -1, cen0[0], cen0[1], 0);
    T.laddr = {0, 0};
    // frontier: boxes of the current level with their range of leaf records
    std::vector<int> fbox{0}, fa{0}, fe{(int)leaves.size()};
    int l = 0;
    for (;;) {
        bool any = false;
        for (size_t i = 0; i < fbox.size(); ++i)
            if (fa[i] < fe[i] && leaves[fa[i]].level > l) any = true;
        if (!any) break;
        T.bsz.push_back(T.bsz[l] / 2);
        int first = T.nboxes();
        std::vector<int> nb, na, ne;
        for (size_t i = 0; i < fbox.size(); ++i) {
            int a = fa[i], e = fe[i];
            if (!(a < e && leaves[a].level > l)) continue;
            int c0 = B.add_children(fbox[i]);
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
        }
        T.laddr.push_back(first);
        T.laddr.push_back(T.nboxes() - 1);
        fbox.swap(nb);
        fa.swap(na);
        fe.swap(ne);
        ++l;
    }
    T.nlevels = l;
    B.all_colleagues();
    if (T.nlevels >= 2) B.fix_level_restriction();

    // leaves in Morton order: their prefixes aligned to full depth are distinct
    int nb = T.nboxes();
    std::vector<std::pair<uint64_t, int>> lk;
    for (int b = 0; b < nb; ++b)
        if (T.nch[b] == 0) lk.push_back({T.pre[b] << (2 * (nlmax - T.lev[b])), b});
    std::sort(lk.begin(), lk.end());
    T.leaf_of_rank.resize(lk.size());
    T.rank_of_box.assign(nb, -1);
    T.rank_lo.assign(nb, 0);
    T.rank_hi.assign(nb, 0);
    for (size_t r = 0; r < lk.size(); ++r) {
        T.leaf_of_rank[r] = lk[r].second;
        T.rank_of_box[lk[r].second] = (int)r;
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
