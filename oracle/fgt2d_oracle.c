/*
 * fgt2d_oracle.c -- CPU FP64 restatement of the 2D plane-wave fast Gauss
 * transform, used ONLY as the test oracle and the CPU baseline.
 *
 * TEST INFRASTRUCTURE. Nothing in the product (fgt2d_amd/) links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py do, and only as the checker.
 *
 * It restates the algorithm of the reference (flatironinstitute/fgt2d,
 * 2d/planewave) step by step from its mathematics and loop structure, with
 * every function citing the reference routine it follows.  Paths are
 * relative to the reference root.  Indices are 0-based here; the reference
 * is 1-based Fortran.
 *
 * Pinning (what checks this restatement, see tests/):
 *   - every entry point against the reference itself, compiled from its own
 *     sources into oracle/_ref by oracle/Makefile (tree arrays bit-exact,
 *     potentials to ~1e-13 relative);
 *   - against the committed golden fixtures in tests/golden/ (written by the
 *     reference in this container);
 *   - against the FP64 direct sum within the requested eps;
 *   - closed forms: quadrature identity of the plane-wave nodes, term counts.
 *
 * Scope: charges and dipoles, alone or together, potentials and gradients
 * and Hessians at sources and/or targets (ifpgh, ifpghtarg in 0..3), any
 * nd >= 1.
 *
 * Threads.  Built with -fopenmp (liborc_omp.so, the all-core CPU baseline)
 * the per-box loops of the tree sort, formation, translations, evaluation,
 * near field and reorders run in parallel.  Every output keeps the order of
 * its additions (the near field runs per target box over its source boxes
 * in the sequential loop's order), so the threaded results are bit-identical
 * to the sequential ones.  Without -fopenmp (liborc.so) the pragmas vanish.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define ORC_MAXLEV 200

#ifdef _OPENMP
#include <omp.h>
#define OMP(x) _Pragma(#x)
#else
#define OMP(x)
#endif

/* threads of the OpenMP build (0: its default); returns the count in use,
 * 1 for the sequential build */
int orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* small helpers                                                       */
/* ------------------------------------------------------------------ */

/* real**integer as the Fortran compiler lowers it (binary powering,
 * compiler-rt __powidf2). */
static double orc_powi(double a, int b)
{
    int recip = b < 0;
    double r = 1;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1 / r : r;
}

static void *xcalloc(size_t n, size_t sz)
{
    void *p = calloc(n ? n : 1, sz);
    if (!p) abort();
    return p;
}

/* ------------------------------------------------------------------ */
/* term counts: 2d/planewave/src/g2dterms.f                            */
/* ------------------------------------------------------------------ */

/* g2dhlterms, g2dterms.f:19-66.  Smallest p >= int(e r^2)+1 with
 * ndim*rk*p^(-1/4)*rp^p/(1-rp) < eps, r = bsize/sqrt(2 delta),
 * rp = r*sqrt(e/p), rk = 1.09 (single-precision literal) * (2 pi)^(-1/4). */
int orc_hlterms(int ndim, double bsize, double delta, double eps)
{
    double e1 = exp(1.0);
    double r = bsize / sqrt(2.0 * delta);
    int p0 = (int)(e1 * r * r) + 1;
    double rk = (double)1.09f * pow(8 * atan(1.0), -0.25);
    for (int p = p0; p <= 300; p++) {
        double rp = r * sqrt(e1 / p);
        double err = ndim * rk * pow((double)p, -0.25) * orc_powi(rp, p) / (1 - rp);
        if (err < eps) return p;
    }
    return -1;
}

/* g2dpwterms, g2dterms.f:71-133.  pmax = 2 sqrt(ln 1/eps),
 * h = 2 pi/(rmax + sqrt(ln 1/eps)), npw = 2 int(pmax/h). */
void orc_pwterms(double bsize, double delta, double eps, double *pmax, int *npw)
{
    double pi = 4.0 * atan(1.0);
    double rmax = bsize / sqrt(delta);
    double d = sqrt(log(1.0 / eps));
    *pmax = 2.0 * d;
    double h = 2 * pi / (rmax + d);
    *npw = 2 * (int)(*pmax / h);
}

/* get_pwnodes, g2drouts.f:2205-2225: midpoint nodes ts_j = (j-1/2) h for the
 * first npw/2, negated for the second half; ws = h/(2 sqrt(pi)) e^{-ts^2/4}. */
void orc_pwnodes(double pmax, int npw, double *ws, double *ts)
{
    double pi = 4.0 * atan(1.0);
    int npw2 = npw / 2;
    double h = pmax / npw2;
    double w = h / (2.0 * sqrt(pi));
    for (int j = 1; j <= npw2; j++) {
        ts[j - 1] = (j - 0.5) * h;
        ws[j - 1] = w * exp(-ts[j - 1] * ts[j - 1] / 4);
        ts[j - 1 + npw2] = -ts[j - 1];
        ws[j - 1 + npw2] = ws[j - 1];
    }
}

/* pts_tree_boxsize0, common/pts_tree2d4.f:1187-1258. */
void orc_boxsize0(double delta, double eps, const double *src, int ns,
                  const double *targ, int nt, int *levcut, double *bs0, double *cen0)
{
    double xmin = src[0], xmax = src[0], ymin = src[1], ymax = src[1];
    for (int i = 0; i < ns; i++) {
        if (src[2 * i] < xmin) xmin = src[2 * i];
        if (src[2 * i] > xmax) xmax = src[2 * i];
        if (src[2 * i + 1] < ymin) ymin = src[2 * i + 1];
        if (src[2 * i + 1] > ymax) ymax = src[2 * i + 1];
    }
    for (int i = 0; i < nt; i++) {
        if (targ[2 * i] < xmin) xmin = targ[2 * i];
        if (targ[2 * i] > xmax) xmax = targ[2 * i];
        if (targ[2 * i + 1] < ymin) ymin = targ[2 * i + 1];
        if (targ[2 * i + 1] > ymax) ymax = targ[2 * i + 1];
    }
    double b = xmax - xmin, sizey = ymax - ymin;
    cen0[0] = (xmin + xmax) / 2;
    cen0[1] = (ymin + ymax) / 2;
    if (sizey > b) b = sizey;
    double dc = sqrt(delta * log(1.0 / eps));
    double dl = log(b / dc) / log(2.0);
    int lc = (int)dl;
    *bs0 = b;
    *levcut = lc;
    if (b < dc) return;
    *bs0 = (double)(1 << (lc + 1)) * dc;
    *levcut = lc + 1;
}

/* ------------------------------------------------------------------ */
/* quadtree: common/pts_tree2d4.f, common/tree_routs4.f                */
/* ------------------------------------------------------------------ */

typedef struct {
    int nlevels, nboxes, cap;
    int laddr[2 * (ORC_MAXLEV + 1)]; /* first,last box of each level */
    double bsz[ORC_MAXLEV + 1];
    int *lev, *par, *nch, *ch, *ncol, *col, *flag;
    double *cen;
} otree;

static void ot_grow(otree *T, int need)
{
    if (need <= T->cap) return;
    int nc = T->cap ? T->cap : 64;
    while (nc < need) nc *= 2;
    T->lev = realloc(T->lev, sizeof(int) * nc);
    T->par = realloc(T->par, sizeof(int) * nc);
    T->nch = realloc(T->nch, sizeof(int) * nc);
    T->ch = realloc(T->ch, sizeof(int) * 4 * nc);
    T->ncol = realloc(T->ncol, sizeof(int) * nc);
    T->col = realloc(T->col, sizeof(int) * 9 * nc);
    T->flag = realloc(T->flag, sizeof(int) * nc);
    T->cen = realloc(T->cen, sizeof(double) * 2 * nc);
    if (!T->lev || !T->par || !T->nch || !T->ch || !T->ncol || !T->col || !T->flag || !T->cen)
        abort();
    T->cap = nc;
}

static void ot_free(otree *T)
{
    free(T->lev); free(T->par); free(T->nch); free(T->ch);
    free(T->ncol); free(T->col); free(T->flag); free(T->cen);
    memset(T, 0, sizeof(*T));
}

/* child j (0..3) centre: tree_routs4.f:34-39, c + (-1)^(j+1) bs/2 in x,
 * and -bs/2 for j<2, +bs/2 otherwise in y (bs = child box size). */
static void ot_make_children(otree *T, int ibox, double bs, int clev, int first_new)
{
    T->nch[ibox] = 4;
    for (int j = 0; j < 4; j++) {
        int jb = first_new + j;
        double hx = (j & 1) ? bs / 2 : -bs / 2;
        double hy = (j < 2) ? -bs / 2 : bs / 2;
        T->cen[2 * jb] = T->cen[2 * ibox] + hx;
        T->cen[2 * jb + 1] = T->cen[2 * ibox + 1] + hy;
        T->par[jb] = ibox;
        T->nch[jb] = 0;
        for (int l = 0; l < 4; l++) T->ch[4 * jb + l] = -1;
        T->ch[4 * ibox + j] = jb;
        T->lev[jb] = clev;
        T->ncol[jb] = 0;
        T->flag[jb] = 0;
    }
}

/* sort_pts_to_children, pts_tree2d4.f:594-689: stable partition of the
 * box's index range into children 1..4 = (x<,y<),(x>=,y<),(x<,y>=),(x>=,y>=)
 * with the test "p - c < 0". se[2b] = first, se[2b+1] = one past last. */
static void ot_sort_children(const otree *T, int ibox, const double *xy,
                             int *perm, int *se, int *tmp)
{
    int s = se[2 * ibox], e = se[2 * ibox + 1];
    int cnt[4] = {0, 0, 0, 0};
    for (int i = s; i < e; i++) {
        int p = perm[i];
        int xb = !(xy[2 * p] - T->cen[2 * ibox] < 0);
        int yb = !(xy[2 * p + 1] - T->cen[2 * ibox + 1] < 0);
        cnt[xb + 2 * yb]++;
    }
    int off[4];
    off[0] = 0;
    for (int j = 1; j < 4; j++) off[j] = off[j - 1] + cnt[j - 1];
    int pos[4] = {off[0], off[1], off[2], off[3]};
    for (int i = s; i < e; i++) {
        int p = perm[i];
        int xb = !(xy[2 * p] - T->cen[2 * ibox] < 0);
        int yb = !(xy[2 * p + 1] - T->cen[2 * ibox + 1] < 0);
        tmp[s + pos[xb + 2 * yb]++] = p;  /* the box's own slice of tmp */
    }
    memcpy(perm + s, tmp + s, sizeof(int) * (e - s));
    for (int j = 0; j < 4; j++) {
        int jb = T->ch[4 * ibox + j];
        se[2 * jb] = s + off[j];
        se[2 * jb + 1] = s + off[j] + cnt[j];
    }
}

/* computecoll, tree_routs4.f:100-211: same-level boxes with
 * |dc| <= 1.05*bs per axis (1.05 is a single-precision literal there),
 * found among the children of the parent's colleagues. */
static void ot_colleagues(otree *T, int b0, int b1)
{
    OMP(omp parallel for schedule(dynamic, 64))
    for (int ib = b0; ib <= b1; ib++) {
        int lv = T->lev[ib];
        T->ncol[ib] = 0;
        if (lv == 0) { T->ncol[ib] = 1; T->col[9 * ib] = ib; continue; }
        int dad = T->par[ib];
        double tol = (double)1.05f * T->bsz[lv];
        for (int i = 0; i < T->ncol[dad]; i++) {
            int jb = T->col[9 * dad + i];
            for (int j = 0; j < 4; j++) {
                int kb = T->ch[4 * jb + j];
                if (kb < 0) continue;
                if (fabs(T->cen[2 * kb] - T->cen[2 * ib]) <= tol &&
                    fabs(T->cen[2 * kb + 1] - T->cen[2 * ib + 1]) <= tol)
                    T->col[9 * ib + T->ncol[ib]++] = kb;
            }
        }
    }
}

static void ot_colleagues_all(otree *T)
{
    for (int l = 0; l <= T->nlevels; l++)
        ot_colleagues(T, T->laddr[2 * l], T->laddr[2 * l + 1]);
}

/* tree_refine_boxes_flag, tree_routs4.f:334-388: boxes of [b0,b1] with flag>0
 * get 4 children appended at T->nboxes; children of flag-1 boxes become
 * flag 3 ("flag++"), children of flag-2 boxes flag 0. */
static void ot_refine_flagged(otree *T, int b0, int b1, int ilev)
{
    for (int ib = b0; ib <= b1; ib++) {
        if (T->flag[ib] <= 0) continue;
        ot_grow(T, T->nboxes + 4);
        int first = T->nboxes;
        T->nboxes += 4;
        ot_make_children(T, ib, T->bsz[ilev + 1], ilev + 1, first);
        int cf = T->flag[ib] == 1 ? 3 : 0;
        for (int j = 0; j < 4; j++) T->flag[first + j] = cf;
    }
}

/* pts_tree_reorg, pts_tree2d4.f:1001-1150: renumber so that each level holds
 * its original boxes followed by its appended ("tail") boxes. */
static void ot_reorg(otree *T, int *tail)
{
    int n = T->nboxes;
    int *map = xcalloc(n, sizeof(int));
    int *lev = xcalloc(n, sizeof(int)), *par = xcalloc(n, sizeof(int));
    int *nch = xcalloc(n, sizeof(int)), *ch = xcalloc(4 * (size_t)n, sizeof(int));
    int *flag = xcalloc(n, sizeof(int));
    double *cen = xcalloc(2 * (size_t)n, sizeof(double));
    memcpy(lev, T->lev, sizeof(int) * n);
    memcpy(par, T->par, sizeof(int) * n);
    memcpy(nch, T->nch, sizeof(int) * n);
    memcpy(ch, T->ch, sizeof(int) * 4 * n);
    memcpy(flag, T->flag, sizeof(int) * n);
    memcpy(cen, T->cen, sizeof(double) * 2 * n);
    for (int l = 0; l <= 1 && l <= T->nlevels; l++)
        for (int ib = T->laddr[2 * l]; ib <= T->laddr[2 * l + 1]; ib++) map[ib] = ib;
    int cur = T->laddr[4];
    for (int l = 2; l <= T->nlevels; l++) {
        int f0 = T->laddr[2 * l], f1 = T->laddr[2 * l + 1];
        T->laddr[2 * l] = cur;
        for (int pass = 0; pass < 2; pass++) {
            int a = pass ? tail[2 * l] : f0, b = pass ? tail[2 * l + 1] : f1;
            for (int ib = a; ib <= b; ib++) {
                T->lev[cur] = lev[ib];
                T->nch[cur] = nch[ib];
                T->cen[2 * cur] = cen[2 * ib];
                T->cen[2 * cur + 1] = cen[2 * ib + 1];
                T->flag[cur] = flag[ib];
                map[ib] = cur;
                cur++;
            }
        }
        T->laddr[2 * l + 1] = cur - 1;
    }
    for (int ib = 0; ib < n; ib++) {
        T->par[map[ib]] = par[ib] < 0 ? -1 : map[par[ib]];
        for (int j = 0; j < 4; j++)
            T->ch[4 * map[ib] + j] = ch[4 * ib + j] < 0 ? -1 : map[ch[4 * ib + j]];
    }
    free(map); free(lev); free(par); free(nch); free(ch); free(flag); free(cen);
}

/* updateflags, tree_routs4.f:217-329: a flag++ box is refined (flag 1) if a
 * child of one of its colleagues has children and touches it. */
static void ot_updateflags(otree *T, int curlev, int b0, int b1)
{
    double distest = 1.05 * (T->bsz[curlev] + T->bsz[curlev + 1]) / 2.0;
    for (int ib = b0; ib <= b1; ib++) {
        if (T->flag[ib] != 3) continue;
        T->flag[ib] = 0;
        for (int i = 0; i < T->ncol[ib] && T->flag[ib] == 0; i++) {
            int jb = T->col[9 * ib + i];
            for (int j = 0; j < 4; j++) {
                int kb = T->ch[4 * jb + j];
                if (kb < 0 || T->nch[kb] <= 0) continue;
                if (fabs(T->cen[2 * kb] - T->cen[2 * ib]) <= distest &&
                    fabs(T->cen[2 * kb + 1] - T->cen[2 * ib + 1]) <= distest) {
                    T->flag[ib] = 1;
                    break;
                }
            }
        }
    }
}

/* pts_tree_fix_lr, pts_tree2d4.f:693-996: make the tree level restricted. */
static void ot_fix_lr(otree *T)
{
    int L = T->nlevels;
    int tail[2 * (ORC_MAXLEV + 1)];
    for (int ib = 0; ib < T->nboxes; ib++) T->flag[ib] = 0;
    /* flag: leaf colleagues of the grandparent touching the parent */
    for (int l = L; l >= 2; l--) {
        double distest = 1.05 * (T->bsz[l - 1] + T->bsz[l - 2]) / 2.0;
        for (int ib = T->laddr[2 * l]; ib <= T->laddr[2 * l + 1]; ib++) {
            int dad = T->par[ib], gd = T->par[dad];
            for (int i = 0; i < T->ncol[gd]; i++) {
                int jb = T->col[9 * gd + i];
                if (T->nch[jb] == 0 && T->flag[jb] == 0) {
                    if (fabs(T->cen[2 * jb] - T->cen[2 * dad]) <= distest &&
                        fabs(T->cen[2 * jb + 1] - T->cen[2 * dad + 1]) <= distest)
                        T->flag[jb] = 1;
                }
            }
        }
    }
    /* flag+: leaf colleagues of the parent of a flagged box touching it */
    for (int l = L; l >= 1; l--) {
        double distest = 1.05 * (T->bsz[l] + T->bsz[l - 1]) / 2.0;
        for (int ib = T->laddr[2 * l]; ib <= T->laddr[2 * l + 1]; ib++) {
            if (T->flag[ib] != 1 && T->flag[ib] != 2) continue;
            int dad = T->par[ib];
            for (int i = 0; i < T->ncol[dad]; i++) {
                int jb = T->col[9 * dad + i];
                if (T->nch[jb] == 0 && T->flag[jb] == 0) {
                    if (fabs(T->cen[2 * jb] - T->cen[2 * ib]) <= distest &&
                        fabs(T->cen[2 * jb + 1] - T->cen[2 * ib + 1]) <= distest)
                        T->flag[jb] = 2;
                }
            }
        }
    }
    for (int l = 0; l <= L; l++) { tail[2 * l] = 0; tail[2 * l + 1] = -1; }
    for (int l = 1; l <= L - 2; l++) {
        tail[2 * (l + 1)] = T->nboxes;
        ot_refine_flagged(T, T->laddr[2 * l], T->laddr[2 * l + 1], l);
        tail[2 * (l + 1) + 1] = T->nboxes - 1;
    }
    ot_reorg(T, tail);
    ot_colleagues_all(T);
    for (int ib = 0; ib < T->nboxes; ib++)
        if (T->flag[ib] != 3) T->flag[ib] = 0;
    for (int l = 0; l <= L; l++) { tail[2 * l] = 0; tail[2 * l + 1] = -1; }
    for (int l = 2; l <= L - 2; l++) {
        ot_updateflags(T, l, T->laddr[2 * l], T->laddr[2 * l + 1]);
        ot_updateflags(T, l, tail[2 * l], tail[2 * l + 1]);
        tail[2 * (l + 1)] = T->nboxes;
        ot_refine_flagged(T, T->laddr[2 * l], T->laddr[2 * l + 1], l);
        ot_refine_flagged(T, tail[2 * l], tail[2 * l + 1], l);
        tail[2 * (l + 1) + 1] = T->nboxes - 1;
        ot_colleagues(T, tail[2 * (l + 1)], tail[2 * (l + 1) + 1]);
    }
    ot_reorg(T, tail);
    ot_colleagues_all(T);
}

/* pts_tree_mem + pts_tree_build, pts_tree2d4.f:34-356 and :363-590, with
 * idivflag = 0 (refine on sources only), nlmin = 0: a box at level l splits
 * when it holds more than ndiv0 sources (l < levcut) or ndiv (l >= levcut),
 * up to nlmax levels; then colleagues and the level-restriction fix. */
static void ot_build(otree *T, const double *src, int ns, int nlmax, int levcut,
                     int ndiv0, int ndiv, double bs0, const double *cen0)
{
    memset(T, 0, sizeof(*T));
    ot_grow(T, 64);
    T->bsz[0] = bs0;
    T->nboxes = 1;
    T->cen[0] = cen0[0];
    T->cen[1] = cen0[1];
    T->lev[0] = 0; T->par[0] = -1; T->nch[0] = 0; T->flag[0] = 0; T->ncol[0] = 0;
    for (int j = 0; j < 4; j++) T->ch[j] = -1;
    T->laddr[0] = 0; T->laddr[1] = 0;
    int *perm = xcalloc(ns, sizeof(int)), *tmp = xcalloc(ns, sizeof(int));
    for (int i = 0; i < ns; i++) perm[i] = i;
    int secap = 256;
    int *se = xcalloc(2 * secap, sizeof(int));
    se[0] = 0; se[1] = ns;
    int ilev;
    for (ilev = 0; ilev < nlmax; ilev++) {
        int f0 = T->laddr[2 * ilev], f1 = T->laddr[2 * ilev + 1];
        int lim = ilev < levcut ? ndiv0 : ndiv;
        int any = 0;
        for (int ib = f0; ib <= f1; ib++)
            if (se[2 * ib + 1] - se[2 * ib] > lim) any = 1;
        if (!any) break;
        T->bsz[ilev + 1] = T->bsz[ilev] / 2;
        T->laddr[2 * (ilev + 1)] = T->nboxes;
        /* children in parent order, then each parent's points partitioned
         * (disjoint index ranges) */
        for (int ib = f0; ib <= f1; ib++) {
            if (se[2 * ib + 1] - se[2 * ib] <= lim) continue;
            ot_grow(T, T->nboxes + 4);
            if (T->nboxes + 4 > secap) {
                while (secap < T->nboxes + 4) secap *= 2;
                se = realloc(se, sizeof(int) * 2 * secap);
            }
            int first = T->nboxes;
            T->nboxes += 4;
            ot_make_children(T, ib, T->bsz[ilev + 1], ilev + 1, first);
        }
        OMP(omp parallel for schedule(dynamic, 16))
        for (int ib = f0; ib <= f1; ib++)
            if (T->nch[ib] > 0) ot_sort_children(T, ib, src, perm, se, tmp);
        T->laddr[2 * (ilev + 1) + 1] = T->nboxes - 1;
    }
    T->nlevels = ilev;
    free(perm); free(tmp); free(se);
    ot_colleagues_all(T);
    if (T->nlevels >= 2) ot_fix_lr(T);
}

/* pts_tree_sort, pts_tree2d4.f:1157-1181: top-down stable partition of
 * points into the final tree. se has 2*nboxes entries. */
static void ot_sort_points(const otree *T, const double *xy, int n, int *perm, int *se)
{
    int *tmp = xcalloc(n, sizeof(int));
    for (int i = 0; i < n; i++) perm[i] = i;
    se[0] = 0; se[1] = n;
    for (int l = 0; l < T->nlevels; l++) {
        OMP(omp parallel for schedule(dynamic, 16))
        for (int ib = T->laddr[2 * l]; ib <= T->laddr[2 * l + 1]; ib++)
            if (T->nch[ib] > 0) ot_sort_children(T, ib, xy, perm, se, tmp);
    }
    free(tmp);
}

/* ------------------------------------------------------------------ */
/* tree export for the parity tests                                    */
/* ------------------------------------------------------------------ */

typedef struct {
    otree T;
    int npwlevel, nlmax;
    double bs0, cen0[2];
    int *isrc, *isrcse, *itarg, *itargse;
    int ns, nt;
} orc_tree_handle;

/* Builds the tree exactly as fgt2d does before its main routine
 * (fgt2d.f:124-221): boxsize0, nlmax from the Hermite term count,
 * ndiv0=10 / ndiv=80, tree, then source and target sorts. */
void *orc_tree2d_build(double delta, double eps, int ns, const double *src,
                       int nt, const double *targ, int *info)
{
    orc_tree_handle *H = xcalloc(1, sizeof(*H));
    orc_boxsize0(delta, eps, src, ns, targ, nt, &H->npwlevel, &H->bs0, H->cen0);
    int npwlevel = H->npwlevel;
    int nlevstart = npwlevel > 0 ? npwlevel : 0;
    int nlmax = npwlevel + 4;
    double bsize = H->bs0 / (double)(1 << nlevstart);
    for (int i = nlevstart; i <= 100; i++) {
        if (orc_hlterms(2, bsize, delta, eps) <= 18) { nlmax = i; break; }
        bsize = bsize / 2;
    }
    H->nlmax = nlmax;
    ot_build(&H->T, src, ns, nlmax, npwlevel, 10, 80, H->bs0, H->cen0);
    H->ns = ns; H->nt = nt;
    H->isrc = xcalloc(ns, sizeof(int));
    H->itarg = xcalloc(nt, sizeof(int));
    H->isrcse = xcalloc(2 * (size_t)H->T.nboxes, sizeof(int));
    H->itargse = xcalloc(2 * (size_t)H->T.nboxes, sizeof(int));
    ot_sort_points(&H->T, src, ns, H->isrc, H->isrcse);
    ot_sort_points(&H->T, targ, nt, H->itarg, H->itargse);
    info[0] = npwlevel; info[1] = nlmax; info[2] = H->T.nlevels; info[3] = H->T.nboxes;
    return H;
}

/* Copies the tree out (0-based). Arrays sized by the caller from info. */
void orc_tree2d_get(void *h, int *laddr, int *lev, int *par, int *nch, int *ch,
                    int *ncol, int *col, double *cen, double *bsz,
                    int *isrc, int *isrcse, int *itarg, int *itargse, double *bs0cen0)
{
    orc_tree_handle *H = h;
    const otree *T = &H->T;
    int n = T->nboxes;
    memcpy(laddr, T->laddr, sizeof(int) * 2 * (T->nlevels + 1));
    memcpy(lev, T->lev, sizeof(int) * n);
    memcpy(par, T->par, sizeof(int) * n);
    memcpy(nch, T->nch, sizeof(int) * n);
    memcpy(ch, T->ch, sizeof(int) * 4 * n);
    memcpy(ncol, T->ncol, sizeof(int) * n);
    memcpy(col, T->col, sizeof(int) * 9 * n);
    memcpy(cen, T->cen, sizeof(double) * 2 * n);
    memcpy(bsz, T->bsz, sizeof(double) * (T->nlevels + 1));
    memcpy(isrc, H->isrc, sizeof(int) * H->ns);
    memcpy(itarg, H->itarg, sizeof(int) * H->nt);
    memcpy(isrcse, H->isrcse, sizeof(int) * 2 * n);
    memcpy(itargse, H->itargse, sizeof(int) * 2 * n);
    bs0cen0[0] = H->bs0; bs0cen0[1] = H->cen0[0]; bs0cen0[2] = H->cen0[1];
}

void orc_tree2d_free(void *h)
{
    orc_tree_handle *H = h;
    ot_free(&H->T);
    free(H->isrc); free(H->itarg); free(H->isrcse); free(H->itargse);
    free(H);
}

/* ------------------------------------------------------------------ */
/* expansions: 2d/planewave/src/g2drouts.f                             */
/* complex numbers are stored as interleaved (re, im) doubles          */
/* ------------------------------------------------------------------ */

/* g2dformhc_vec / g2dformhd_vec / g2dformhcd_vec, g2drouts.f:100-330:
 * H(j,k) += q hx_j hy_k + d (n1 dhx_j hy_k + n2 hx_j dhy_k), with
 * hx_j = x^j/sqrt(j!), dhx_j = j dsq hx_{j-1}/sqrt(j), x = (s - c)/sqrt(delta),
 * in the reference's order of operations.  chg == NULL: no charges;
 * rn == NULL: no dipoles.  h is (nt+1)x(nt+1), j fastest. */
static void orc_formh(int nd, double delta, const double *src, int ns, const double *chg,
                      const double *rn, const double *ds, const double *c, int nt, double *h,
                      double *xp, double *yp, double *dxp, double *dyp)
{
    int m = nt + 1;
    memset(h, 0, sizeof(double) * m * m * nd);
    double dsq = 1.0 / sqrt(delta);
    for (int i = 0; i < ns; i++) {
        double x = (src[2 * i] - c[0]) * dsq, y = (src[2 * i + 1] - c[1]) * dsq;
        xp[0] = 1.0; yp[0] = 1.0;
        dxp[0] = 0.0; dyp[0] = 0.0;
        for (int k = 1; k <= nt; k++) {
            double t = sqrt((double)k);
            yp[k] = yp[k - 1] * y / t;
            xp[k] = xp[k - 1] * x / t;
            dxp[k] = k * dsq * xp[k - 1] / t;
            dyp[k] = k * dsq * yp[k - 1] / t;
        }
        for (int ind = 0; ind < nd; ind++) {
            double *hh = h + (size_t)ind * m * m;
            if (!rn) {
                double q = chg[(size_t)nd * i + ind];
                for (int k = 0; k <= nt; k++) {
                    double yt = q * yp[k];
                    for (int j = 0; j <= nt; j++) hh[j + m * k] += yt * xp[j];
                }
            } else {
                double dd = ds[(size_t)nd * i + ind];
                for (int k = 0; k <= nt; k++) {
                    double d1 = yp[k] * rn[2 * i] * dd, d2 = dyp[k] * rn[2 * i + 1] * dd;
                    if (chg) {
                        double yt = chg[(size_t)nd * i + ind] * yp[k];
                        for (int j = 0; j <= nt; j++)
                            hh[j + m * k] = hh[j + m * k] + dxp[j] * d1 + xp[j] * d2 + xp[j] * yt;
                    } else {
                        for (int j = 0; j <= nt; j++)
                            hh[j + m * k] = hh[j + m * k] + dxp[j] * d1 + xp[j] * d2;
                    }
                }
            }
        }
    }
}

/* g2drh2xmat, g2drouts.f:1645-1684: real part of the Hermite->PW 1-D matrix,
 * ws_i (-ts_i)^j / sqrt(j!); the i^j factor is applied by grouping j mod 4. */
static void orc_rh2xmat(int nt, int npw, const double *ws, const double *ts, double *m)
{
    double fac = 1.0;
    for (int j = 0; j <= nt; j++) {
        if (j > 0) fac = fac / sqrt((double)j);
        for (int i = 0; i < npw; i++) m[i + npw * j] = ws[i] * orc_powi(-ts[i], j) * fac;
    }
}

/* g2drx2lmat, g2drouts.f:1731-1769: ts_i^j / j!. */
static void orc_rx2lmat(int nl, int npw, const double *ts, double *m)
{
    double fac = 1.0;
    for (int j = 0; j <= nl; j++) {
        if (j > 0) fac = fac * j;
        for (int i = 0; i < npw; i++) m[i + npw * j] = orc_powi(ts[i], j) / fac;
    }
}

/* g2dh2pw_real_vec, g2drouts.f:1242-1329: PW(k2,k1) += sum_{j1,j2}
 * h2x(k2,j1) i^j1 h2x(k1,j2) i^j2 H(j1,j2); k2 runs over the positive-node
 * half (x), k1 over all nodes (y). pw is (npw/2) x npw complex, k2 fastest. */
static void orc_h2pw(int nd, int nt, int npw, const double *h2x, const double *h,
                     double *pw, double *pxhy)
{
    int m = nt + 1, npw2 = npw / 2;
    for (int ind = 0; ind < nd; ind++) {
        const double *hh = h + (size_t)ind * m * m;
        double *p = pw + (size_t)ind * 2 * npw2 * npw;
        for (int k2 = 0; k2 < npw2; k2++)
            for (int j2 = 0; j2 <= nt; j2++) {
                double c[4] = {0, 0, 0, 0};
                for (int r = 0; r < 4; r++)
                    for (int j1 = r; j1 <= nt; j1 += 4) c[r] += h2x[k2 + npw * j1] * hh[j1 + m * j2];
                pxhy[2 * (j2 + m * k2)] = c[0] - c[2];
                pxhy[2 * (j2 + m * k2) + 1] = c[1] - c[3];
            }
        for (int k1 = 0; k1 < npw; k1++)
            for (int k2 = 0; k2 < npw2; k2++) {
                double cr[4] = {0, 0, 0, 0}, ci[4] = {0, 0, 0, 0};
                for (int r = 0; r < 4; r++)
                    for (int j2 = r; j2 <= nt; j2 += 4) {
                        double a = h2x[k1 + npw * j2];
                        cr[r] += a * pxhy[2 * (j2 + m * k2)];
                        ci[r] += a * pxhy[2 * (j2 + m * k2) + 1];
                    }
                /* (c0 - c2) + i (c1 - c3) */
                p[2 * (k2 + npw2 * k1)] += (cr[0] - cr[2]) - (ci[1] - ci[3]);
                p[2 * (k2 + npw2 * k1) + 1] += (ci[0] - ci[2]) + (cr[1] - cr[3]);
            }
    }
}

/* g2dpw2local_real_vec, g2drouts.f:1417-1540: Taylor coefficients
 * L(k2,k1) of 2 Re sum PW(j2,j1) e^{i ts_j2 x} e^{i ts_j1 y}; overwrites L. */
static void orc_pw2local(int nd, int nl, int npw, const double *x2l, const double *pw,
                         double *loc, double *lxpy)
{
    int npw2 = npw / 2, m = nl + 1;
    for (int ind = 0; ind < nd; ind++) {
        const double *p = pw + (size_t)ind * 2 * npw2 * npw;
        double *L = loc + (size_t)ind * m * m;
        for (int k2 = 0; k2 <= nl; k2++)
            for (int j1 = 0; j1 < npw; j1++) {
                double cr = 0, ci = 0;
                for (int j2 = 0; j2 < npw2; j2++) {
                    cr += x2l[j2 + npw * k2] * p[2 * (j2 + npw2 * j1)];
                    ci += x2l[j2 + npw * k2] * p[2 * (j2 + npw2 * j1) + 1];
                }
                /* multiply by i^k2 */
                double r, im;
                switch (k2 & 3) {
                case 0: r = cr; im = ci; break;
                case 1: r = -ci; im = cr; break;
                case 2: r = -cr; im = -ci; break;
                default: r = ci; im = -cr; break;
                }
                lxpy[2 * (k2 + m * j1)] = r;
                lxpy[2 * (k2 + m * j1) + 1] = im;
            }
        for (int k1 = 0; k1 <= nl; k1++) {
            double sgn = (k1 & 1) ? -1.0 : 1.0;
            for (int k2 = 0; k2 <= nl; k2++) {
                double cr = 0, ci = 0;
                for (int j1 = 0; j1 < npw2; j1++) {
                    double a = x2l[j1 + npw * k1];
                    cr += a * (lxpy[2 * (k2 + m * j1)] + sgn * lxpy[2 * (k2 + m * (j1 + npw2))]);
                    ci += a * (lxpy[2 * (k2 + m * j1) + 1] + sgn * lxpy[2 * (k2 + m * (j1 + npw2)) + 1]);
                }
                double v;
                switch (k1 & 3) {
                case 0: v = cr * 2; break;
                case 1: v = -ci * 2; break;
                case 2: v = -cr * 2; break;
                default: v = ci * 2; break;
                }
                L[k2 + m * k1] = v;
            }
        }
    }
}

/* g2dlevalp_vec / g2dlevalg_vec / g2dlevalh_vec, g2drouts.f:847-1080:
 * Taylor evaluation sum L(j,k) x^j y^k; derivative tables carry the
 * 1/sqrt(delta) factors as there. */
static void orc_leval(int nd, double delta, const double *c, int nl, const double *loc,
                      const double *tg, int ntg, int ifpgh, double *pot, double *grad,
                      double *hess)
{
    int m = nl + 1;
    double dsq = 1.0 / sqrt(delta);
    double *xp = xcalloc(m + 2, sizeof(double)), *yp = xcalloc(m + 2, sizeof(double));
    double *dxp = xcalloc(m + 2, sizeof(double)), *dyp = xcalloc(m + 2, sizeof(double));
    double *ddxp = xcalloc(m + 2, sizeof(double)), *ddyp = xcalloc(m + 2, sizeof(double));
    for (int it = 0; it < ntg; it++) {
        double x = (tg[2 * it] - c[0]) * dsq, y = (tg[2 * it + 1] - c[1]) * dsq;
        xp[0] = 1.0; yp[0] = 1.0;
        for (int i = 1; i <= nl; i++) { xp[i] = x * xp[i - 1]; yp[i] = y * yp[i - 1]; }
        if (ifpgh >= 2) {
            dxp[0] = 0; dyp[0] = 0;
            for (int i = 1; i <= nl; i++) { dxp[i] = i * xp[i - 1] * dsq; dyp[i] = i * yp[i - 1] * dsq; }
        }
        if (ifpgh >= 3) {
            ddxp[0] = 0; ddyp[0] = 0; ddxp[1] = 0; ddyp[1] = 0;
            for (int i = 2; i <= nl; i++) {
                ddxp[i] = i * (i - 1) * xp[i - 2] * dsq * dsq;
                ddyp[i] = i * (i - 1) * yp[i - 2] * dsq * dsq;
            }
        }
        for (int ind = 0; ind < nd; ind++) {
            const double *L = loc + (size_t)ind * m * m;
            double dd = 0, g1 = 0, g2 = 0, h1 = 0, h2 = 0, h3 = 0;
            for (int k = 0; k <= nl; k++) {
                double d2 = 0, d2x = 0, d2xx = 0;
                for (int j = 0; j <= nl; j++) {
                    d2 += L[j + m * k] * xp[j];
                    if (ifpgh >= 2) d2x += L[j + m * k] * dxp[j];
                    if (ifpgh >= 3) d2xx += L[j + m * k] * ddxp[j];
                }
                dd += d2 * yp[k];
                if (ifpgh >= 2) { g1 += d2x * yp[k]; g2 += d2 * dyp[k]; }
                if (ifpgh >= 3) { h1 += d2xx * yp[k]; h2 += d2x * dyp[k]; h3 += d2 * ddyp[k]; }
            }
            size_t t = (size_t)it;
            pot[nd * t + ind] += dd;
            if (ifpgh >= 2) {
                grad[nd * 2 * t + ind] += g1;
                grad[nd * 2 * t + nd + ind] += g2;
            }
            if (ifpgh >= 3) {
                hess[nd * 3 * t + ind] += h1;
                hess[nd * 3 * t + nd + ind] += h2;
                hess[nd * 3 * t + 2 * nd + ind] += h3;
            }
        }
    }
    free(xp); free(yp); free(dxp); free(dyp); free(ddxp); free(ddyp);
}

/* g2dpwevalp_vec / g2dpwevalg_vec / g2dpwevalh_vec, g2drouts.f:1887-2200:
 * pot += 2 Re sum_{a<=npw/2} sum_b PW(a,b) e^{i ts_a x} e^{i ts_b y}; the
 * phases come from one complex exponential per axis and its odd powers,
 * derivatives multiply by i ts dsq per order. */
static inline void cmadd(double *zr, double *zi, double ar, double ai, double br, double bi)
{
    *zr += ar * br - ai * bi;
    *zi += ar * bi + ai * br;
}
static void orc_pweval(int nd, double delta, const double *c, int npw, const double *ts,
                       const double *pw, const double *tg, int ntg, int ifpgh,
                       double *pot, double *grad, double *hess)
{
    int npw2 = npw / 2;
    double dsq = 1.0 / sqrt(delta);
    /* w[0]: phase, w[1]: d/dx (or d/dy), w[2]: second derivative */
    double *w1 = xcalloc(6 * (size_t)npw, sizeof(double)), *w2 = xcalloc(6 * (size_t)npw, sizeof(double));
#define W(a, o, j) (a)[2 * ((size_t)(o) * npw + (j))]
#define WI(a, o, j) (a)[2 * ((size_t)(o) * npw + (j)) + 1]
    for (int it = 0; it < ntg; it++) {
        double x = (tg[2 * it] - c[0]) * dsq, y = (tg[2 * it + 1] - c[1]) * dsq;
        double qxr = cos(ts[0] * x), qxi = sin(ts[0] * x);
        double qyr = cos(ts[0] * y), qyi = sin(ts[0] * y);
        double q1r = qxr, q1i = qxi, q2r = qyr, q2i = qyi;
        double sxr = qxr * qxr - qxi * qxi, sxi = qxr * qxi + qxi * qxr;
        double syr = qyr * qyr - qyi * qyi, syi = qyr * qyi + qyi * qyr;
        for (int j = 0; j < npw2; j++) {
            double *a[2] = {w1, w2};
            double qr[2] = {q1r, q2r}, qi[2] = {q1i, q2i};
            for (int s = 0; s < 2; s++) {
                W(a[s], 0, j) = qr[s]; WI(a[s], 0, j) = qi[s];
                W(a[s], 0, j + npw2) = qr[s]; WI(a[s], 0, j + npw2) = -qi[s];
                for (int o = 1; o <= 2; o++) {
                    /* (i ts dsq) * previous order */
                    double f = ts[j] * dsq;
                    double pr = W(a[s], o - 1, j), pi = WI(a[s], o - 1, j);
                    W(a[s], o, j) = -f * pi; WI(a[s], o, j) = f * pr;
                    W(a[s], o, j + npw2) = W(a[s], o, j); WI(a[s], o, j + npw2) = -WI(a[s], o, j);
                }
            }
            double t;
            t = q1r * sxr - q1i * sxi; q1i = q1r * sxi + q1i * sxr; q1r = t;
            t = q2r * syr - q2i * syi; q2i = q2r * syi + q2i * syr; q2r = t;
        }
        for (int ind = 0; ind < nd; ind++) {
            const double *p = pw + (size_t)ind * 2 * npw2 * npw;
            double zr = 0, zi = 0, zxr = 0, zxi = 0, zyr = 0, zyi = 0;
            double zxxr = 0, zxxi = 0, zxyr = 0, zxyi = 0, zyyr = 0, zyyi = 0;
            for (int b = 0; b < npw; b++) {
                double cr = 0, ci = 0, cxr = 0, cxi = 0, cxxr = 0, cxxi = 0;
                for (int a = 0; a < npw2; a++) {
                    double pr = p[2 * (a + npw2 * b)], pi = p[2 * (a + npw2 * b) + 1];
                    cmadd(&cr, &ci, pr, pi, W(w1, 0, a), WI(w1, 0, a));
                    if (ifpgh >= 2) cmadd(&cxr, &cxi, pr, pi, W(w1, 1, a), WI(w1, 1, a));
                    if (ifpgh >= 3) cmadd(&cxxr, &cxxi, pr, pi, W(w1, 2, a), WI(w1, 2, a));
                }
                cmadd(&zr, &zi, cr, ci, W(w2, 0, b), WI(w2, 0, b));
                if (ifpgh >= 2) {
                    cmadd(&zxr, &zxi, cxr, cxi, W(w2, 0, b), WI(w2, 0, b));
                    cmadd(&zyr, &zyi, cr, ci, W(w2, 1, b), WI(w2, 1, b));
                }
                if (ifpgh >= 3) {
                    cmadd(&zxxr, &zxxi, cxxr, cxxi, W(w2, 0, b), WI(w2, 0, b));
                    cmadd(&zxyr, &zxyi, cxr, cxi, W(w2, 1, b), WI(w2, 1, b));
                    cmadd(&zyyr, &zyyi, cr, ci, W(w2, 2, b), WI(w2, 2, b));
                }
            }
            size_t t = (size_t)it;
            pot[nd * t + ind] += zr * 2;
            if (ifpgh >= 2) {
                grad[nd * 2 * t + ind] += zxr * 2;
                grad[nd * 2 * t + nd + ind] += zyr * 2;
            }
            if (ifpgh >= 3) {
                hess[nd * 3 * t + ind] += zxxr * 2;
                hess[nd * 3 * t + nd + ind] += zxyr * 2;
                hess[nd * 3 * t + 2 * nd + ind] += zyyr * 2;
            }
        }
    }
#undef W
#undef WI
    free(w1); free(w2);
}

/* g2d_directcp_vec / cg / ch, g2ddirect.f:32-210: for each target,
 * sum over sources with r^2 < ln(1/eps) delta of q e^{-r^2/delta}
 * (and derivatives: grad = -2/delta (t-s) e, hess as in g2ddirect.f:146).
 * With dipoles (rn != NULL) the six variants g2d_directdp/dg/dh and
 * g2d_directcdp/cdg/cdh, g2ddirect.f:217-655, in their order of
 * operations: the dipole kernel is d (n . grad_s) e^{-r^2/delta}. */
void orc_direct2d_cd(int nd, double delta, double eps, const double *src, int ns,
                     const double *chg, const double *rn, const double *ds, const double *tg,
                     int ntg, int ifpgh, double *pot, double *grad, double *hess)
{
    double dmax = log(1.0 / eps) * delta;
    for (int it = 0; it < ntg; it++) {
        size_t t = (size_t)it;
        double *P = pot + nd * t, *G = grad + nd * 2 * t, *H = hess + nd * 3 * t;
        for (int i = 0; i < ns; i++) {
            double xd = tg[2 * it] - src[2 * i], yd = tg[2 * it + 1] - src[2 * i + 1];
            double rr = xd * xd + yd * yd;
            if (!(rr < dmax)) continue;
            const double *q = chg ? chg + (size_t)nd * i : NULL;
            const double *d = rn ? ds + (size_t)nd * i : NULL;
            double n1 = rn ? rn[2 * i] : 0, n2 = rn ? rn[2 * i + 1] : 0;
            if (!rn) {
                double rfac = -2.0 / delta;
                double e = exp(-rr / delta);
                for (int ind = 0; ind < nd; ind++) {
                    P[ind] += e * q[ind];
                    if (ifpgh >= 2) {
                        double dx = rfac * xd, dy = rfac * yd;
                        G[ind] += e * q[ind] * dx;
                        G[nd + ind] += e * q[ind] * dy;
                        if (ifpgh >= 3) {
                            H[ind] += e * q[ind] * (dx * dx + rfac);
                            H[nd + ind] += e * q[ind] * dx * dy;
                            H[2 * nd + ind] += e * q[ind] * (dy * dy + rfac);
                        }
                    }
                }
            } else if (ifpgh == 1 && !q) { /* g2d_directdp_vec */
                double rfac = 2.0 / delta;
                double rtmp = exp(-rr / delta);
                double dx = rfac * xd, dy = rfac * yd;
                double rinc = (dx * n1 + dy * n2) * rtmp;
                for (int ind = 0; ind < nd; ind++) P[ind] = P[ind] + rinc * d[ind];
            } else if (ifpgh == 2 && !q) { /* g2d_directdg_vec */
                double rfac = 2.0 / delta;
                double rtmp = exp(-rr / delta);
                double dx = rfac * xd, dy = rfac * yd;
                double rinc = (dx * n1 + dy * n2);
                double rincx = n1 * rfac - rinc * dx, rincy = n2 * rfac - rinc * dy;
                for (int ind = 0; ind < nd; ind++) {
                    double dtmp = rtmp * d[ind];
                    P[ind] = P[ind] + rinc * dtmp;
                    G[ind] = G[ind] + rincx * dtmp;
                    G[nd + ind] = G[nd + ind] + rincy * dtmp;
                }
            } else if (!q) { /* g2d_directdh_vec */
                double rfac = 2.0 / delta;
                double dtmp = exp(-rr / delta);
                double dx = rfac * xd, dy = rfac * yd;
                double rinc = (dx * n1 + dy * n2);
                double rincx = n1 * rfac - rinc * dx, rincy = n2 * rfac - rinc * dy;
                double rincxx = -(rfac * (dx * n1 + rinc)) - rincx * dx;
                double rincyy = -(rfac * (dy * n2 + rinc)) - rincy * dy;
                double rincxy = -(rfac * dx * n2) - rincx * dy;
                for (int ind = 0; ind < nd; ind++) {
                    double rt = dtmp * d[ind];
                    P[ind] = P[ind] + rinc * rt;
                    G[ind] = G[ind] + rincx * rt;
                    G[nd + ind] = G[nd + ind] + rincy * rt;
                    H[ind] = H[ind] + rincxx * rt;
                    H[nd + ind] = H[nd + ind] + rincxy * rt;
                    H[2 * nd + ind] = H[2 * nd + ind] + rincyy * rt;
                }
            } else if (ifpgh == 1) { /* g2d_directcdp_vec */
                double rfac = 2.0 / delta;
                double dx = rfac * xd, dy = rfac * yd;
                double rtmp = exp(-rr / delta);
                for (int ind = 0; ind < nd; ind++) P[ind] = P[ind] + rtmp * q[ind];
                double rinc = (dx * n1 + dy * n2) * rtmp;
                for (int ind = 0; ind < nd; ind++) P[ind] = P[ind] + rinc * d[ind];
            } else if (ifpgh == 2) { /* g2d_directcdg_vec */
                double rfac = -2.0 / delta;
                double dtmp = exp(-rr / delta);
                double dx = rfac * xd, dy = rfac * yd;
                for (int ind = 0; ind < nd; ind++) {
                    double rinc = dtmp * q[ind];
                    P[ind] = P[ind] + rinc;
                    G[ind] = G[ind] + dx * rinc;
                    G[nd + ind] = G[nd + ind] + dy * rinc;
                }
                double rinc = -(dx * n1) - dy * n2;
                double rincx = rinc * dx - rfac * n1, rincy = rinc * dy - rfac * n2;
                for (int ind = 0; ind < nd; ind++) {
                    double rt = dtmp * d[ind];
                    P[ind] = P[ind] + rinc * rt;
                    G[ind] = G[ind] + rincx * rt;
                    G[nd + ind] = G[nd + ind] + rincy * rt;
                }
            } else { /* g2d_directcdh_vec */
                double rfac = -2.0 / delta;
                double dtmp = exp(-rr / delta);
                double dx = rfac * xd, dy = rfac * yd;
                double dxx = rfac + dx * dx, dyy = rfac + dy * dy, dxy = dx * dy;
                for (int ind = 0; ind < nd; ind++) {
                    double rinc = dtmp * q[ind];
                    P[ind] = P[ind] + rinc;
                    G[ind] = G[ind] + dx * rinc;
                    G[nd + ind] = G[nd + ind] + dy * rinc;
                    H[ind] = H[ind] + dxx * rinc;
                    H[nd + ind] = H[nd + ind] + dxy * rinc;
                    H[2 * nd + ind] = H[2 * nd + ind] + dyy * rinc;
                }
                double rinc = -(dx * n1) - dy * n2;
                double rincx = rinc * dx - n1 * rfac, rincy = rinc * dy - n2 * rfac;
                double rincxx = rfac * (rinc - dx * n1) + rincx * dx;
                double rincyy = rfac * (rinc - dy * n2) + rincy * dy;
                double rincxy = rincx * dy - rfac * dx * n2;
                for (int ind = 0; ind < nd; ind++) {
                    double rt = dtmp * d[ind];
                    P[ind] = P[ind] + rinc * rt;
                    G[ind] = G[ind] + rincx * rt;
                    G[nd + ind] = G[nd + ind] + rincy * rt;
                    H[ind] = H[ind] + rincxx * rt;
                    H[nd + ind] = H[nd + ind] + rincxy * rt;
                    H[2 * nd + ind] = H[2 * nd + ind] + rincyy * rt;
                }
            }
        }
    }
}

void orc_direct2d(int nd, double delta, double eps, const double *src, int ns,
                  const double *chg, const double *tg, int ntg, int ifpgh,
                  double *pot, double *grad, double *hess)
{
    orc_direct2d_cd(nd, delta, eps, src, ns, chg, NULL, NULL, tg, ntg, ifpgh, pot, grad, hess);
}

/* g2dformlc_vec / g2dformld_vec / g2dformlcd_vec, g2drouts.f:621-845:
 * Taylor coefficients of the sources' field about c, via the recurrence
 * h_{j+1} = 2 (x h_j - h_{j-1})/(j+1), h_0 = e^{-x^2}, h_1 = 2x e^{-x^2};
 * dipoles use dh_j = -h_{j+1} dsq (j+1), in the reference's order of
 * operations.  Accumulates into loc (the reference does not clear it). */
static void orc_formlc(int nd, double delta, const double *src, int ns, const double *chg,
                       const double *rn, const double *ds, const double *c, int nl, double *loc)
{
    int m = nl + 1;
    double dsq = 1.0 / sqrt(delta);
    double *hx = xcalloc(m + 1, sizeof(double)), *hy = xcalloc(m + 1, sizeof(double));
    double *dhx = xcalloc(m, sizeof(double)), *dhy = xcalloc(m, sizeof(double));
    for (int i = 0; i < ns; i++) {
        double x = (src[2 * i] - c[0]) * dsq, y = (src[2 * i + 1] - c[1]) * dsq;
        double fx = exp(-x * x), fy = exp(-y * y);
        hx[0] = fx; hy[0] = fy;
        hx[1] = 2.0 * x * fx; hy[1] = 2.0 * y * fy;
        int jtop = rn ? nl : nl - 1;
        for (int j = 1; j <= jtop; j++) {
            hx[j + 1] = 2.0 * (x * hx[j] - hx[j - 1]) / (j + 1);
            hy[j + 1] = 2.0 * (y * hy[j] - hy[j - 1]) / (j + 1);
        }
        if (rn)
            for (int j = 0; j <= nl; j++) {
                dhx[j] = -(hx[j + 1] * dsq * (j + 1.0));
                dhy[j] = -(hy[j + 1] * dsq * (j + 1.0));
            }
        for (int ind = 0; ind < nd; ind++) {
            double *L = loc + (size_t)ind * m * m;
            if (!rn) {
                double q = chg[(size_t)nd * i + ind];
                for (int j1 = 0; j1 <= nl; j1++) {
                    double yt = q * hy[j1];
                    for (int j2 = 0; j2 <= nl; j2++) L[j2 + m * j1] += yt * hx[j2];
                }
                continue;
            }
            double r1 = rn[2 * i] * ds[(size_t)nd * i + ind];
            double r2 = rn[2 * i + 1] * ds[(size_t)nd * i + ind];
            for (int j1 = 0; j1 <= nl; j1++) {
                double dy1 = hy[j1] * r1, dy2 = dhy[j1] * r2;
                if (chg) {
                    double yt = chg[(size_t)nd * i + ind] * hy[j1];
                    for (int j2 = 0; j2 <= nl; j2++)
                        L[j2 + m * j1] = L[j2 + m * j1] + dhx[j2] * dy1 + hx[j2] * dy2 + yt * hx[j2];
                } else {
                    for (int j2 = 0; j2 <= nl; j2++)
                        L[j2 + m * j1] = L[j2 + m * j1] + dhx[j2] * dy1 + hx[j2] * dy2;
                }
            }
        }
    }
    free(hx); free(hy); free(dhx); free(dhy);
}

/* ------------------------------------------------------------------ */
/* the transform: fgt2d (fgt2d.f:19-462) and fgt2dmain (:468-1348)     */
/* ------------------------------------------------------------------ */

typedef struct {
    int npwlevel, nlmax, nlevels, nboxes, npw, ntermmax;
    int nlocal[ORC_MAXLEV + 1];
    double pmax;
    double time_main;
} orc_stats;

static orc_stats g_last_stats;

void orc_last_stats(int *out)
{
    out[0] = g_last_stats.npwlevel;
    out[1] = g_last_stats.nlmax;
    out[2] = g_last_stats.nlevels;
    out[3] = g_last_stats.nboxes;
    out[4] = g_last_stats.npw;
    out[5] = g_last_stats.ntermmax;
    for (int i = 0; i <= g_last_stats.nlevels && i < 64; i++) out[6 + i] = g_last_stats.nlocal[i];
}

static void cmul_add(double *dst, const double *a, const double *w)
{
    double ar = a[0], ai = a[1], wr = w[0], wi = w[1];
    dst[0] += ar * wr - ai * wi;
    dst[1] += ar * wi + ai * wr;
}

/* merge_split_pw_matrices, g2drouts.f:2294-2350 (k = 1..nmax), and
 * pw_translation_matrices, g2drouts.f:2235-2289: repeated squaring of
 * e^{i ts xmin}. */
static void orc_shift_tables(int npw, const double *ts, double xmin_ms, int nmax_ms,
                             double xmin_m2l, double *wms, double *wm2l)
{
    int npw2 = npw / 2, nexp = npw * npw2;
    double *ww = xcalloc(2 * (size_t)npw * (nmax_ms + 1), sizeof(double));
    for (int j = 0; j < npw; j++) {
        double zr = cos(ts[j] * xmin_ms), zi = sin(ts[j] * xmin_ms);
        for (int k = 1; k <= nmax_ms; k++) {
            ww[2 * (j + npw * k)] = zr; ww[2 * (j + npw * k) + 1] = zi;
            double t = zr * zr - zi * zi; zi = 2 * zr * zi; zr = t;
        }
    }
    for (int k = 1; k <= nmax_ms; k++)
        for (int j1 = 0; j1 < npw; j1++)
            for (int j2 = 0; j2 < npw2; j2++) {
                double ar = ww[2 * (j2 + npw * k)], ai = ww[2 * (j2 + npw * k) + 1];
                double br = ww[2 * (j1 + npw * k)], bi = ww[2 * (j1 + npw * k) + 1];
                for (int q = 0; q < 4; q++) {
                    double xr = ar, xi = (q >= 2) ? -ai : ai;   /* pp,pm,mp,mm: x sign */
                    double yr = br, yi = (q & 1) ? -bi : bi;     /* y sign */
                    double *o = wms + 2 * ((size_t)(j2 + npw2 * j1) + (size_t)nexp * (q + 4 * (k - 1)));
                    o[0] = xr * yr - xi * yi;
                    o[1] = xr * yi + xi * yr;
                }
            }
    /* m2l: ww(j,k) for k = -1,0,1 */
    for (int kx = -1; kx <= 1; kx++)
        for (int ky = -1; ky <= 1; ky++)
            for (int j1 = 0; j1 < npw; j1++)
                for (int j2 = 0; j2 < npw2; j2++) {
                    double ar = 1, ai = 0, br = 1, bi = 0;
                    if (kx != 0) { ar = cos(ts[j2] * xmin_m2l); ai = kx * sin(ts[j2] * xmin_m2l); }
                    if (ky != 0) { br = cos(ts[j1] * xmin_m2l); bi = ky * sin(ts[j1] * xmin_m2l); }
                    double *o = wm2l + 2 * ((size_t)(j2 + npw2 * j1) + (size_t)nexp * ((kx + 1) + 3 * (ky + 1)));
                    o[0] = ar * br - ai * bi;
                    o[1] = ar * bi + ai * br;
                }
    free(ww);
}

static double wall_now(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* Whole transform with the reference's calling convention (0-based arrays,
 * column-major (2,n) points, (nd,n) densities). Returns 0, or -1 bad
 * argument, -3 unsupported (dipoles). Outputs are overwritten for the
 * requested quantities. */
int orc_fgt2d(int nd, double delta, double eps, int ns, const double *sources,
              int ifcharge, const double *charge, int ifdipole, const double *rnormal,
              const double *dipstr, int ifpgh, double *pot, double *grad, double *hess,
              int nt, const double *targ, int ifpghtarg, double *pottarg,
              double *gradtarg, double *hesstarg)
{
    if (nd < 1 || ns < 1 || nt < 0) return -1;
    if (ifcharge < 0 || ifcharge > 1 || ifdipole < 0 || ifdipole > 1) return -1;
    if (ifpgh < 0 || ifpgh > 3 || ifpghtarg < 0 || ifpghtarg > 3) return -1;
    if (!ifcharge) charge = NULL;
    if (!ifdipole) rnormal = dipstr = NULL;
    memset(&g_last_stats, 0, sizeof(g_last_stats));
    double t_start = wall_now();
    int npwlevel;
    double bs0, cen0[2];
    orc_boxsize0(delta, eps, sources, ns, targ, nt, &npwlevel, &bs0, cen0);
    g_last_stats.npwlevel = npwlevel;

    int npg = ifpgh > ifpghtarg ? ifpgh : ifpghtarg;
    if (npwlevel < -2) {
        /* fgt2d_large_delta, fgt2d.f:1616-1733: one Taylor expansion about the
         * bounding-box centre, evaluated everywhere.  As there, the values
         * are ADDED to the caller's output arrays (fgt2d.f:1700-1730). */
        int nl = orc_hlterms(2, bs0, delta, eps) + npg - 1;
        int m = nl + 1;
        double *loc = xcalloc((size_t)m * m * nd, sizeof(double));
        if (charge || rnormal)
            orc_formlc(nd, delta, sources, ns, charge, rnormal, dipstr, cen0, nl, loc);
        if (ifpghtarg > 0) {
            orc_leval(nd, delta, cen0, nl, loc, targ, nt, ifpghtarg, pottarg, gradtarg, hesstarg);
        }
        if (ifpgh > 0) {
            orc_leval(nd, delta, cen0, nl, loc, sources, ns, ifpgh, pot, grad, hess);
        }
        free(loc);
        g_last_stats.time_main = wall_now() - t_start;
        return 0;
    }

    /* nlmax: first level whose Hermite length is <= 18, fgt2d.f:175-186 */
    int nlevstart = npwlevel > 0 ? npwlevel : 0;
    int nlmax = npwlevel + 4;
    {
        double bsize = bs0 / (double)(1 << nlevstart);
        for (int i = nlevstart; i <= 100; i++) {
            if (orc_hlterms(2, bsize, delta, eps) <= 18) { nlmax = i; break; }
            bsize = bsize / 2;
        }
    }
    g_last_stats.nlmax = nlmax;
    otree T;
    ot_build(&T, sources, ns, nlmax, npwlevel, 10, 80, bs0, cen0);
    int nlevels = T.nlevels, nboxes = T.nboxes;
    g_last_stats.nlevels = nlevels;
    g_last_stats.nboxes = nboxes;
    int *isrc = xcalloc(ns, sizeof(int)), *itarg = xcalloc(nt ? nt : 1, sizeof(int));
    int *isrcse = xcalloc(2 * (size_t)nboxes, sizeof(int));
    int *itargse = xcalloc(2 * (size_t)nboxes, sizeof(int));
    ot_sort_points(&T, sources, ns, isrc, isrcse);
    ot_sort_points(&T, targ, nt, itarg, itargse);

    /* term counts per level, fgt2d.f:334-356 */
    int nlocal[ORC_MAXLEV + 1];
    memset(nlocal, 0, sizeof(nlocal));
    int nadd = ifdipole + npg;
    int ntermmax = 0;
    for (int i = nlevstart; i <= nlevels; i++) {
        nlocal[i] = orc_hlterms(2, T.bsz[i], delta, eps);
        if (ifdipole == 1 || ifpgh >= 2 || ifpghtarg >= 2) nlocal[i] += nadd;
        if (nlocal[i] > ntermmax) ntermmax = nlocal[i];
    }
    double pmax;
    int npw;
    {
        double bsize = 2 * bs0 / (double)(1 << nlevstart);
        double pweps = eps;
        if (nadd > 2) pweps = pweps / 100;
        orc_pwterms(bsize, delta, pweps, &pmax, &npw);
    }
    g_last_stats.npw = npw;
    g_last_stats.ntermmax = ntermmax;
    memcpy(g_last_stats.nlocal, nlocal, sizeof(nlocal));

    /* sorted copies, dreorderf fmmcommon2d.f:3-27 */
    double *ss = xcalloc(2 * (size_t)ns, sizeof(double)), *cs = xcalloc((size_t)nd * ns, sizeof(double));
    double *tsrt = xcalloc(2 * (size_t)(nt ? nt : 1), sizeof(double));
    double *rs = xcalloc(2 * (size_t)ns, sizeof(double)), *dss = xcalloc((size_t)nd * ns, sizeof(double));
    OMP(omp parallel for)
    for (int i = 0; i < ns; i++) {
        ss[2 * i] = sources[2 * (size_t)isrc[i]];
        ss[2 * i + 1] = sources[2 * (size_t)isrc[i] + 1];
        if (charge)
            for (int d = 0; d < nd; d++) cs[(size_t)nd * i + d] = charge[(size_t)nd * isrc[i] + d];
        if (rnormal) {
            rs[2 * i] = rnormal[2 * (size_t)isrc[i]];
            rs[2 * i + 1] = rnormal[2 * (size_t)isrc[i] + 1];
            for (int d = 0; d < nd; d++) dss[(size_t)nd * i + d] = dipstr[(size_t)nd * isrc[i] + d];
        }
    }
    const double *cq = charge ? cs : NULL, *crn = rnormal ? rs : NULL;
    OMP(omp parallel for)
    for (int i = 0; i < nt; i++) {
        tsrt[2 * i] = targ[2 * (size_t)itarg[i]];
        tsrt[2 * i + 1] = targ[2 * (size_t)itarg[i] + 1];
    }
    size_t nps = ifpgh > 0 ? (size_t)ns : 1, npt = ifpghtarg > 0 ? (size_t)(nt ? nt : 1) : 1;
    double *ps = xcalloc(nd * nps, sizeof(double)), *gs = xcalloc(2 * nd * nps, sizeof(double));
    double *hs = xcalloc(3 * nd * nps, sizeof(double));
    double *pt = xcalloc(nd * npt, sizeof(double)), *gt = xcalloc(2 * nd * npt, sizeof(double));
    double *ht = xcalloc(3 * nd * npt, sizeof(double));

    /* ---- fgt2dmain ---- */
    int nlevend = npwlevel < nlevels ? npwlevel : nlevels;
    int *ifhung = xcalloc(nboxes, sizeof(int));
    for (int l = 0; l <= nlevend; l++)
        for (int ib = T.laddr[2 * l]; ib <= T.laddr[2 * l + 1]; ib++)
            if (T.nch[ib] == 0 && isrcse[2 * ib + 1] > isrcse[2 * ib]) ifhung[ib] = 1;

    /* compute_modified_list1, tree_routs4.f:715-879 */
    int *nlist1 = xcalloc(nboxes, sizeof(int)), *list1 = xcalloc(13 * (size_t)nboxes, sizeof(int));
    if (T.nch[0] == 0) { nlist1[0] = 1; list1[0] = 0; }
    for (int l = 1; l <= nlevels; l++)
        for (int ib = T.laddr[2 * l]; ib <= T.laddr[2 * l + 1]; ib++) {
            if (T.nch[ib] != 0) continue;
            for (int i = 0; i < T.ncol[ib]; i++) {
                int jb = T.col[9 * ib + i];
                if (l >= npwlevel) {
                    list1[13 * ib + nlist1[ib]++] = jb;
                } else if (T.nch[jb] == 0) {
                    list1[13 * ib + nlist1[ib]++] = jb;
                } else {
                    double distest = 1.05 * (T.bsz[l] + T.bsz[l + 1]) / 2.0;
                    for (int j = 0; j < T.nch[jb]; j++) {
                        int kb = T.ch[4 * jb + j];
                        double xd = fabs(T.cen[2 * kb] - T.cen[2 * ib]);
                        double yd = fabs(T.cen[2 * kb + 1] - T.cen[2 * ib + 1]);
                        if (xd < distest && yd < distest) {
                            list1[13 * ib + nlist1[ib]++] = kb;
                            list1[13 * kb + nlist1[kb]++] = ib;
                        }
                    }
                }
            }
        }

    int npw2 = npw / 2, nexp = npw * npw2;
    double *mp = NULL, *lc = NULL;
    if (npwlevel < nlevels) {
        double ws[200], ts[200];
        orc_pwnodes(pmax, npw, ws, ts);
        /* expansion storage for levels nlevstart..nlevels (g2dmpalloc,
         * fgt2d.f:1518-1610); indexed directly by box number here */
        size_t per = (size_t)2 * nexp * nd;
        mp = xcalloc(per * nboxes, sizeof(double));
        lc = xcalloc(per * nboxes, sizeof(double));
        int nmax = nlevels - nlevstart;
        double *wms = xcalloc(2 * (size_t)nexp * 4 * (nmax + 1), sizeof(double));
        double *wm2l = xcalloc(2 * (size_t)nexp * 9, sizeof(double));
        double xmin_m2l = (npwlevel >= 0 ? T.bsz[npwlevel] : T.bsz[0]) / sqrt(delta);
        orc_shift_tables(npw, ts, T.bsz[nlevels] / sqrt(delta) / 2, nmax, xmin_m2l, wms, wm2l);
        double xmin = npwlevel >= 0 ? T.bsz[npwlevel] : T.bsz[0];
        double *rh2x = xcalloc((size_t)npw * (ntermmax + 1), sizeof(double));
        double *rx2l = xcalloc((size_t)npw * (ntermmax + 1), sizeof(double));
        orc_rh2xmat(ntermmax, npw, ws, ts, rh2x);
        orc_rx2lmat(ntermmax, npw, ts, rx2l);

        int *iflocal = xcalloc(nboxes, sizeof(int));
        if (npwlevel >= 0 && npwlevel <= nlevels)
            for (int ib = T.laddr[2 * npwlevel]; ib <= T.laddr[2 * npwlevel + 1]; ib++) {
                if (T.nch[ib] != 0) continue;
                for (int i = 0; i < T.ncol[ib]; i++)
                    if (T.nch[T.col[9 * ib + i]] > 0) { iflocal[ib] = 1; break; }
            }

        int mt = ntermmax + 1;
        /* step 1: leaf Hermite expansions -> PW, fgt2d.f:871-959 */
        int lstop = npwlevel + 1 > 0 ? npwlevel + 1 : 0;
        for (int l = nlevels; l >= lstop; l--)
            OMP(omp parallel)
            {
            double *hexp = xcalloc((size_t)mt * mt * nd, sizeof(double));
            double *pxhy = xcalloc(2 * (size_t)mt * npw, sizeof(double));
            double *xp = xcalloc(mt + 1, sizeof(double)), *yp = xcalloc(mt + 1, sizeof(double));
            double *dxp = xcalloc(mt + 1, sizeof(double)), *dyp = xcalloc(mt + 1, sizeof(double));
            OMP(omp for schedule(dynamic, 4))
            for (int ib = T.laddr[2 * l]; ib <= T.laddr[2 * l + 1]; ib++) {
                int s = isrcse[2 * ib], n = isrcse[2 * ib + 1] - s;
                if (T.nch[ib] != 0 || n <= 0) continue;
                if (!cq && !crn) continue; /* fgt2d.f:882-959: no branch forms anything */
                orc_formh(nd, delta, ss + 2 * (size_t)s, n, cq ? cs + (size_t)nd * s : NULL,
                          crn ? rs + 2 * (size_t)s : NULL, dss + (size_t)nd * s, T.cen + 2 * ib,
                          nlocal[l], hexp, xp, yp, dxp, dyp);
                orc_h2pw(nd, nlocal[l], npw, rh2x, hexp, mp + per * ib, pxhy);
            }
            free(hexp); free(pxhy); free(xp); free(yp); free(dxp); free(dyp);
            }
        /* step 2: merge children into parents, fgt2d.f:976-1009 */
        for (int l = nlevels - 1; l >= nlevstart; l--)
            OMP(omp parallel for schedule(dynamic, 16))
            for (int ib = T.laddr[2 * l]; ib <= T.laddr[2 * l + 1]; ib++)
                for (int i = 0; i < T.nch[ib]; i++) {
                    int jb = T.ch[4 * ib + i];
                    if (isrcse[2 * jb + 1] - isrcse[2 * jb] <= 0) continue;
                    double dx = T.cen[2 * ib] - T.cen[2 * jb], dy = T.cen[2 * ib + 1] - T.cen[2 * jb + 1];
                    int k = (dx > 0 && dy > 0) ? 0 : (dx > 0 && dy < 0) ? 1 : (dx < 0 && dy > 0) ? 2 : 3;
                    const double *w = wms + 2 * (size_t)nexp * (k + 4 * (nlevels - l - 1));
                    for (int d = 0; d < nd; d++)
                        for (int e = 0; e < nexp; e++)
                            cmul_add(mp + per * ib + 2 * ((size_t)d * nexp + e),
                                     mp + per * jb + 2 * ((size_t)d * nexp + e), w + 2 * e);
                }
        /* step 3: mp -> loc at the cutoff level, fgt2d.f:1026-1064 */
        OMP(omp parallel for schedule(dynamic, 16))
        for (int ib = T.laddr[2 * nlevstart]; ib <= T.laddr[2 * nlevstart + 1]; ib++) {
            int np = 0;
            if (ifpghtarg > 0) np += itargse[2 * ib + 1] - itargse[2 * ib];
            if (ifpgh > 0) np += isrcse[2 * ib + 1] - isrcse[2 * ib];
            if (np <= 0) continue;
            if (nlevstart == npwlevel)
                for (int i = 0; i < T.ncol[ib]; i++) {
                    int jb = T.col[9 * ib + i];
                    if (jb == ib) continue; /* gt2d_computelistpw_slow, tree_routs4.f:945-953 */
                    int jx = (int)lround((T.cen[2 * ib] - T.cen[2 * jb]) / xmin);
                    int jy = (int)lround((T.cen[2 * ib + 1] - T.cen[2 * jb + 1]) / xmin);
                    const double *w = wm2l + 2 * (size_t)nexp * ((jx + 1) + 3 * (jy + 1));
                    for (int d = 0; d < nd; d++)
                        for (int e = 0; e < nexp; e++)
                            cmul_add(lc + per * ib + 2 * ((size_t)d * nexp + e),
                                     mp + per * jb + 2 * ((size_t)d * nexp + e), w + 2 * e);
                }
            for (size_t e = 0; e < per; e++) lc[per * ib + e] += mp[per * ib + e];
        }
        /* step 4: split loc to children, fgt2d.f:1082-1125 */
        for (int l = nlevstart; l <= nlevels - 1; l++)
            OMP(omp parallel for schedule(dynamic, 16))
            for (int ib = T.laddr[2 * l]; ib <= T.laddr[2 * l + 1]; ib++) {
                int np = 0;
                if (ifpghtarg > 0) np += itargse[2 * ib + 1] - itargse[2 * ib];
                if (ifpgh > 0) np += isrcse[2 * ib + 1] - isrcse[2 * ib];
                if (np <= 0) continue;
                for (int i = 0; i < T.nch[ib]; i++) {
                    int jb = T.ch[4 * ib + i];
                    double dx = T.cen[2 * jb] - T.cen[2 * ib], dy = T.cen[2 * jb + 1] - T.cen[2 * ib + 1];
                    int k = (dx > 0 && dy > 0) ? 0 : (dx > 0 && dy < 0) ? 1 : (dx < 0 && dy > 0) ? 2 : 3;
                    const double *w = wms + 2 * (size_t)nexp * (k + 4 * (nlevels - l - 1));
                    for (int d = 0; d < nd; d++)
                        for (int e = 0; e < nexp; e++)
                            cmul_add(lc + per * jb + 2 * ((size_t)d * nexp + e),
                                     lc + per * ib + 2 * ((size_t)d * nexp + e), w + 2 * e);
                }
            }
        /* step 5: evaluate at leaves, fgt2d.f:1141-1271 */
        for (int l = nlevstart; l <= nlevels; l++)
            OMP(omp parallel)
            {
            double *loc = xcalloc((size_t)mt * mt * nd, sizeof(double));
            double *lxpy = xcalloc(2 * (size_t)mt * npw, sizeof(double));
            OMP(omp for schedule(dynamic, 4))
            for (int ib = T.laddr[2 * l]; ib <= T.laddr[2 * l + 1]; ib++) {
                if (l == npwlevel && !iflocal[ib]) continue;
                if (T.nch[ib] != 0) continue;
                int st = itargse[2 * ib], ntg = itargse[2 * ib + 1] - st;
                int s0 = isrcse[2 * ib], nsr = isrcse[2 * ib + 1] - s0;
                if (ntg + nsr <= 0) continue;
                if (nlocal[l] > 50) {
                    if (ntg > 0 && ifpghtarg > 0)
                        orc_pweval(nd, delta, T.cen + 2 * ib, npw, ts, lc + per * ib, tsrt + 2 * (size_t)st,
                                   ntg, ifpghtarg, pt + (size_t)nd * st, gt + (size_t)2 * nd * st,
                                   ht + (size_t)3 * nd * st);
                    if (nsr > 0 && ifpgh > 0)
                        orc_pweval(nd, delta, T.cen + 2 * ib, npw, ts, lc + per * ib, ss + 2 * (size_t)s0,
                                   nsr, ifpgh, ps + (size_t)nd * s0, gs + (size_t)2 * nd * s0,
                                   hs + (size_t)3 * nd * s0);
                } else {
                    orc_pw2local(nd, nlocal[l], npw, rx2l, lc + per * ib, loc, lxpy);
                    if (ntg > 0 && ifpghtarg > 0)
                        orc_leval(nd, delta, T.cen + 2 * ib, nlocal[l], loc, tsrt + 2 * (size_t)st, ntg,
                                  ifpghtarg, pt + (size_t)nd * st, gt + (size_t)2 * nd * st,
                                  ht + (size_t)3 * nd * st);
                    if (nsr > 0 && ifpgh > 0)
                        orc_leval(nd, delta, T.cen + 2 * ib, nlocal[l], loc, ss + 2 * (size_t)s0, nsr,
                                  ifpgh, ps + (size_t)nd * s0, gs + (size_t)2 * nd * s0,
                                  hs + (size_t)3 * nd * s0);
                }
            }
            free(loc); free(lxpy);
            }
        free(iflocal); free(wms); free(wm2l); free(rh2x); free(rx2l);
    }

    /* step 6: direct interactions of leaves at or above the cutoff,
     * fgt2d.f:1287-1333, cutoff eps/90.  The loop runs over target boxes
     * ib, each over its source boxes jb in the order of the reference's
     * loop over jb (so every target adds its source boxes in the same
     * order): a CSR of (ib <- jb) pairs, filled in that order. */
    int *nf_ptr = xcalloc((size_t)nboxes + 1, sizeof(int));
    for (int l = 0; l <= nlevend; l++)
        for (int jb = T.laddr[2 * l]; jb <= T.laddr[2 * l + 1]; jb++)
            if (ifhung[jb])
                for (int i = 0; i < nlist1[jb]; i++) nf_ptr[list1[13 * jb + i] + 1]++;
    for (int b = 0; b < nboxes; b++) nf_ptr[b + 1] += nf_ptr[b];
    int *nf_src = xcalloc((size_t)nf_ptr[nboxes] + 1, sizeof(int));
    int *nf_fill = xcalloc((size_t)nboxes + 1, sizeof(int));
    memcpy(nf_fill, nf_ptr, sizeof(int) * nboxes);
    for (int l = 0; l <= nlevend; l++)
        for (int jb = T.laddr[2 * l]; jb <= T.laddr[2 * l + 1]; jb++)
            if (ifhung[jb])
                for (int i = 0; i < nlist1[jb]; i++) nf_src[nf_fill[list1[13 * jb + i]]++] = jb;
    OMP(omp parallel for schedule(dynamic, 8))
    for (int ib = 0; ib < nboxes; ib++)
        for (int k = nf_ptr[ib]; k < nf_ptr[ib + 1]; k++) {
            int jb = nf_src[k];
            int js = isrcse[2 * jb], jn = isrcse[2 * jb + 1] - js;
            {
                int st = itargse[2 * ib], ntg = itargse[2 * ib + 1] - st;
                int s0 = isrcse[2 * ib], nsr = isrcse[2 * ib + 1] - s0;
                const double *q = cq ? cs + (size_t)nd * js : NULL;
                const double *r = crn ? rs + 2 * (size_t)js : NULL, *dd = dss + (size_t)nd * js;
                if (!q && !r) continue; /* fgt2dpart_direct_vec: no branch applies */
                if (ntg > 0 && ifpghtarg > 0)
                    orc_direct2d_cd(nd, delta, eps / 90, ss + 2 * (size_t)js, jn, q, r, dd,
                                    tsrt + 2 * (size_t)st, ntg, ifpghtarg, pt + (size_t)nd * st,
                                    gt + (size_t)2 * nd * st, ht + (size_t)3 * nd * st);
                if (nsr > 0 && ifpgh > 0)
                    orc_direct2d_cd(nd, delta, eps / 90, ss + 2 * (size_t)js, jn, q, r, dd,
                                    ss + 2 * (size_t)s0, nsr, ifpgh, ps + (size_t)nd * s0,
                                    gs + (size_t)2 * nd * s0, hs + (size_t)3 * nd * s0);
            }
        }
    free(nf_ptr); free(nf_src); free(nf_fill);

    /* back to input order, dreorderi / dreorderiv fmmcommon2d.f:29-90 */
    if (ifpgh > 0)
        OMP(omp parallel for)
        for (int i = 0; i < ns; i++) {
            for (int d = 0; d < nd; d++) pot[(size_t)nd * isrc[i] + d] = ps[(size_t)nd * i + d];
            if (ifpgh >= 2)
                for (int d = 0; d < 2 * nd; d++) grad[(size_t)2 * nd * isrc[i] + d] = gs[(size_t)2 * nd * i + d];
            if (ifpgh >= 3)
                for (int d = 0; d < 3 * nd; d++) hess[(size_t)3 * nd * isrc[i] + d] = hs[(size_t)3 * nd * i + d];
        }
    if (ifpghtarg > 0)
        OMP(omp parallel for)
        for (int i = 0; i < nt; i++) {
            for (int d = 0; d < nd; d++) pottarg[(size_t)nd * itarg[i] + d] = pt[(size_t)nd * i + d];
            if (ifpghtarg >= 2)
                for (int d = 0; d < 2 * nd; d++)
                    gradtarg[(size_t)2 * nd * itarg[i] + d] = gt[(size_t)2 * nd * i + d];
            if (ifpghtarg >= 3)
                for (int d = 0; d < 3 * nd; d++)
                    hesstarg[(size_t)3 * nd * itarg[i] + d] = ht[(size_t)3 * nd * i + d];
        }

    free(mp); free(lc); free(ifhung); free(nlist1); free(list1);
    free(ss); free(cs); free(rs); free(dss); free(tsrt); free(ps); free(gs); free(hs); free(pt); free(gt); free(ht);
    free(isrc); free(itarg); free(isrcse); free(itargse);
    ot_free(&T);
    g_last_stats.time_main = wall_now() - t_start;
    return 0;
}

double orc_last_time(void) { return g_last_stats.time_main; }

/* derr, 2d/planewave/test/test_fgt2d.f:204-225: relative l2 error, absolute
 * when the reference norm per entry is below 1e-10. */
double orc_derr(const double *vref, const double *vcomp, int n)
{
    double ra = 0, err = 0;
    for (int i = 0; i < n; i++) {
        ra += vref[i] * vref[i];
        err += (vcomp[i] - vref[i]) * (vcomp[i] - vref[i]);
    }
    if (sqrt(ra) / n < 1.0e-10) return sqrt(err) / n;
    return sqrt(err / ra);
}

/* DLARAN (common/dlaran.f:40-89) behind hkrand (common/hkrand.f): the 48-bit
 * multiplicative congruential generator seed <- seed * 33952834046453 mod 2^48
 * (multiplier 494,322,2508,2549 in base 4096), value seed / 2^48 (exact in
 * FP64).  hkrand's default seed is (0,0,0,1), i.e. 1.  Draws n values,
 * continuing from *state (pass 1 for a fresh hkrand stream). */
void orc_dlaran(uint64_t *state, int n, double *out)
{
    const uint64_t M = 33952834046453ull, MASK = (1ull << 48) - 1;
    uint64_t s = *state;
    for (int i = 0; i < n; i++) {
        /* 48x48-bit product mod 2^48 without overflow: split the multiplier */
        uint64_t lo = (s * (M & 0xFFFFFFull)) & MASK;
        uint64_t hi = ((s * (M >> 24)) & ((1ull << 24) - 1)) << 24;
        s = (lo + hi) & MASK;
        out[i] = (double)s / 281474976710656.0;
    }
    *state = s;
}
