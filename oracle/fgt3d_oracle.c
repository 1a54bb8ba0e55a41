/*
 * fgt3d_oracle.c -- CPU FP64 restatement of the 3D plane-wave fast Gauss
 * transform (flatironinstitute/fgt2d, 3d/planewave), used ONLY as the test
 * oracle.
 *
 * TEST INFRASTRUCTURE. Nothing in the product (fgt2d_amd/) links, loads or
 * calls this file; only tests/ and the checker legs do.
 *
 * Each function restates the reference routine it cites (paths relative to
 * the reference root, 3d/planewave/src/...), 0-based where the reference is
 * 1-based.  Pinning: the reference itself built by oracle/Makefile (ref3d),
 * golden fixtures, the FP64 direct sum, closed forms (tests/).
 *
 * One deliberate difference: g3dformlc_vec (g3drouts.f:784-856), the Taylor
 * expansion of the large-delta branch, assigns the scaled z offset to y and
 * leaves z undefined (g3drouts.f:803-804).  The restatement computes the
 * intended expansion (x, y, z offsets); that branch is pinned by the direct
 * sum, not by the reference.
 *
 * Scope: charges, ifpgh/ifpghtarg 0..3, any nd >= 1.  Dipoles: -3.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define O3_MAXLEV 200

int orc_hlterms(int ndim, double bsize, double delta, double eps);
void orc_pwterms(double bsize, double delta, double eps, double *pmax, int *npw);

static double o3_powi(double a, int b)
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

static void *xc(size_t n, size_t sz)
{
    void *p = calloc(n ? n : 1, sz);
    if (!p) abort();
    return p;
}

/* get_pwnodes, g3drouts.f:2669-2690: ts_j = (j + 1/2) h ascending,
 * j = -npw/2 .. npw/2-1, ws = h/(2 sqrt(pi)) e^{-ts^2/4}. */
void orc3_pwnodes(double pmax, int npw, double *ws, double *ts)
{
    double pi = 4.0 * atan(1.0);
    int npw2 = npw / 2;
    double h = pmax / npw2;
    double w = h / (2.0 * sqrt(pi));
    for (int j = -npw2; j < npw2; j++) {
        ts[j + npw2] = (j + 0.5) * h;
        ws[j + npw2] = w * exp(-ts[j + npw2] * ts[j + npw2] / 4);
    }
}

/* pts_tree_boxsize0, common/pts_tree3d.f:1597-1680. */
void orc3_boxsize0(double delta, double eps, const double *src, int ns, const double *targ,
                   int nt, int *levcut, double *bs0, double *cen0)
{
    double mn[3] = {src[0], src[1], src[2]}, mx[3] = {src[0], src[1], src[2]};
    for (int i = 0; i < ns; i++)
        for (int d = 0; d < 3; d++) {
            if (src[3 * i + d] < mn[d]) mn[d] = src[3 * i + d];
            if (src[3 * i + d] > mx[d]) mx[d] = src[3 * i + d];
        }
    for (int i = 0; i < nt; i++)
        for (int d = 0; d < 3; d++) {
            if (targ[3 * i + d] < mn[d]) mn[d] = targ[3 * i + d];
            if (targ[3 * i + d] > mx[d]) mx[d] = targ[3 * i + d];
        }
    double b = mx[0] - mn[0], sy = mx[1] - mn[1], sz = mx[2] - mn[2];
    if (sy > b) b = sy;
    if (sz > b) b = sz;
    for (int d = 0; d < 3; d++) cen0[d] = (mn[d] + mx[d]) / 2;
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
/* octree: common/pts_tree3d.f, common/tree_routs3d.f                  */
/* ------------------------------------------------------------------ */

typedef struct {
    int nlevels, nboxes, cap;
    int laddr[2 * (O3_MAXLEV + 1)];
    double bsz[O3_MAXLEV + 1];
    int *lev, *par, *nch, *ch, *ncol, *col, *flag;
    double *cen;
} o3tree;

static void t3_grow(o3tree *T, int need)
{
    if (need <= T->cap) return;
    int nc = T->cap ? T->cap : 64;
    while (nc < need) nc *= 2;
    T->lev = realloc(T->lev, sizeof(int) * nc);
    T->par = realloc(T->par, sizeof(int) * nc);
    T->nch = realloc(T->nch, sizeof(int) * nc);
    T->ch = realloc(T->ch, sizeof(int) * 8 * nc);
    T->ncol = realloc(T->ncol, sizeof(int) * nc);
    T->col = realloc(T->col, sizeof(int) * 27 * nc);
    T->flag = realloc(T->flag, sizeof(int) * nc);
    T->cen = realloc(T->cen, sizeof(double) * 3 * nc);
    if (!T->lev || !T->par || !T->nch || !T->ch || !T->ncol || !T->col || !T->flag || !T->cen)
        abort();
    T->cap = nc;
}

static void t3_free(o3tree *T)
{
    free(T->lev); free(T->par); free(T->nch); free(T->ch);
    free(T->ncol); free(T->col); free(T->flag); free(T->cen);
    memset(T, 0, sizeof(*T));
}

/* tree_refine_boxes, tree_routs3d.f:9-65: child j (0..7) at
 * c + (-1)^(j+1) bs/2 in x, -bs/2 for j in {0,1,4,5} else +bs/2 in y,
 * -bs/2 for j < 4 else +bs/2 in z (bs = child size). */
static void t3_make_children(o3tree *T, int ib, double bs, int clev, int first)
{
    T->nch[ib] = 8;
    for (int j = 0; j < 8; j++) {
        int jb = first + j;
        double hx = (j & 1) ? bs / 2 : -bs / 2;
        double hy = (j == 0 || j == 1 || j == 4 || j == 5) ? -bs / 2 : bs / 2;
        double hz = j < 4 ? -bs / 2 : bs / 2;
        T->cen[3 * jb] = T->cen[3 * ib] + hx;
        T->cen[3 * jb + 1] = T->cen[3 * ib + 1] + hy;
        T->cen[3 * jb + 2] = T->cen[3 * ib + 2] + hz;
        T->par[jb] = ib;
        T->nch[jb] = 0;
        for (int l = 0; l < 8; l++) T->ch[8 * jb + l] = -1;
        T->ch[8 * ib + j] = jb;
        T->lev[jb] = clev;
        T->ncol[jb] = 0;
        T->flag[jb] = 0;
    }
}

/* sort_pts_to_children, pts_tree3d.f:662-854: stable partitions on z, then
 * y, then x with the test "p - c < 0"; child = xbit + 2 ybit + 4 zbit. */
static void t3_sort_children(const o3tree *T, int ib, const double *p3, int *perm, int *se,
                             int *tmp)
{
    int s = se[2 * ib], e = se[2 * ib + 1];
    int cnt[8] = {0};
    for (int i = s; i < e; i++) {
        int p = perm[i], d = 0;
        for (int a = 0; a < 3; a++) d |= (!(p3[3 * p + a] - T->cen[3 * ib + a] < 0)) << a;
        cnt[d]++;
    }
    int off[8], pos[8];
    off[0] = 0;
    for (int j = 1; j < 8; j++) off[j] = off[j - 1] + cnt[j - 1];
    memcpy(pos, off, sizeof(off));
    for (int i = s; i < e; i++) {
        int p = perm[i], d = 0;
        for (int a = 0; a < 3; a++) d |= (!(p3[3 * p + a] - T->cen[3 * ib + a] < 0)) << a;
        tmp[pos[d]++] = p;
    }
    memcpy(perm + s, tmp, sizeof(int) * (e - s));
    for (int j = 0; j < 8; j++) {
        int jb = T->ch[8 * ib + j];
        se[2 * jb] = s + off[j];
        se[2 * jb + 1] = s + off[j] + cnt[j];
    }
}

static int t3_within(const o3tree *T, int a, int b, double tol, int strict)
{
    for (int d = 0; d < 3; d++) {
        double v = fabs(T->cen[3 * a + d] - T->cen[3 * b + d]);
        if (strict ? !(v < tol) : !(v <= tol)) return 0;
    }
    return 1;
}

/* computecoll, tree_routs3d.f:108-225: same-level boxes within
 * 1.05 (single-precision literal) * bs per axis among the children of the
 * parent's colleagues. */
static void t3_colleagues(o3tree *T, int b0, int b1)
{
    for (int ib = b0; ib <= b1; ib++) {
        int lv = T->lev[ib];
        T->ncol[ib] = 0;
        if (lv == 0) { T->ncol[ib] = 1; T->col[27 * ib] = ib; continue; }
        int dad = T->par[ib];
        double tol = (double)1.05f * T->bsz[lv];
        for (int i = 0; i < T->ncol[dad]; i++) {
            int jb = T->col[27 * dad + i];
            for (int j = 0; j < 8; j++) {
                int kb = T->ch[8 * jb + j];
                if (kb >= 0 && t3_within(T, kb, ib, tol, 0)) T->col[27 * ib + T->ncol[ib]++] = kb;
            }
        }
    }
}

static void t3_colleagues_all(o3tree *T)
{
    for (int l = 0; l <= T->nlevels; l++) t3_colleagues(T, T->laddr[2 * l], T->laddr[2 * l + 1]);
}

/* tree_refine_boxes_flag, tree_routs3d.f:345-410. */
static void t3_refine_flagged(o3tree *T, int b0, int b1, int ilev)
{
    for (int ib = b0; ib <= b1; ib++) {
        if (T->flag[ib] <= 0) continue;
        t3_grow(T, T->nboxes + 8);
        int first = T->nboxes;
        T->nboxes += 8;
        t3_make_children(T, ib, T->bsz[ilev + 1], ilev + 1, first);
        int cf = T->flag[ib] == 1 ? 3 : 0;
        for (int j = 0; j < 8; j++) T->flag[first + j] = cf;
    }
}

/* pts_tree_reorg, pts_tree3d.f:1172-1329. */
static void t3_reorg(o3tree *T, const int *tail)
{
    int n = T->nboxes;
    int *map = xc(n, sizeof(int)), *lev = xc(n, sizeof(int)), *par = xc(n, sizeof(int));
    int *nch = xc(n, sizeof(int)), *ch = xc(8 * (size_t)n, sizeof(int)), *flag = xc(n, sizeof(int));
    double *cen = xc(3 * (size_t)n, sizeof(double));
    memcpy(lev, T->lev, sizeof(int) * n);
    memcpy(par, T->par, sizeof(int) * n);
    memcpy(nch, T->nch, sizeof(int) * n);
    memcpy(ch, T->ch, sizeof(int) * 8 * n);
    memcpy(flag, T->flag, sizeof(int) * n);
    memcpy(cen, T->cen, sizeof(double) * 3 * n);
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
                for (int d = 0; d < 3; d++) T->cen[3 * cur + d] = cen[3 * ib + d];
                T->flag[cur] = flag[ib];
                map[ib] = cur++;
            }
        }
        T->laddr[2 * l + 1] = cur - 1;
    }
    for (int ib = 0; ib < n; ib++) {
        T->par[map[ib]] = par[ib] < 0 ? -1 : map[par[ib]];
        for (int j = 0; j < 8; j++)
            T->ch[8 * map[ib] + j] = ch[8 * ib + j] < 0 ? -1 : map[ch[8 * ib + j]];
    }
    free(map); free(lev); free(par); free(nch); free(ch); free(flag); free(cen);
}

/* updateflags, tree_routs3d.f:226-344. */
static void t3_updateflags(o3tree *T, int curlev, int b0, int b1)
{
    double distest = 1.05 * (T->bsz[curlev] + T->bsz[curlev + 1]) / 2.0;
    for (int ib = b0; ib <= b1; ib++) {
        if (T->flag[ib] != 3) continue;
        T->flag[ib] = 0;
        for (int i = 0; i < T->ncol[ib] && T->flag[ib] == 0; i++) {
            int jb = T->col[27 * ib + i];
            for (int j = 0; j < 8; j++) {
                int kb = T->ch[8 * jb + j];
                if (kb < 0 || T->nch[kb] <= 0) continue;
                if (t3_within(T, kb, ib, distest, 0)) { T->flag[ib] = 1; break; }
            }
        }
    }
}

/* pts_tree_fix_lr, pts_tree3d.f:855-1171. */
static void t3_fix_lr(o3tree *T)
{
    int L = T->nlevels;
    int tail[2 * (O3_MAXLEV + 1)];
    for (int ib = 0; ib < T->nboxes; ib++) T->flag[ib] = 0;
    for (int l = L; l >= 2; l--) {
        double distest = 1.05 * (T->bsz[l - 1] + T->bsz[l - 2]) / 2.0;
        for (int ib = T->laddr[2 * l]; ib <= T->laddr[2 * l + 1]; ib++) {
            int dad = T->par[ib], gd = T->par[dad];
            for (int i = 0; i < T->ncol[gd]; i++) {
                int jb = T->col[27 * gd + i];
                if (T->nch[jb] == 0 && T->flag[jb] == 0 && t3_within(T, jb, dad, distest, 0))
                    T->flag[jb] = 1;
            }
        }
    }
    for (int l = L; l >= 1; l--) {
        double distest = 1.05 * (T->bsz[l] + T->bsz[l - 1]) / 2.0;
        for (int ib = T->laddr[2 * l]; ib <= T->laddr[2 * l + 1]; ib++) {
            if (T->flag[ib] != 1 && T->flag[ib] != 2) continue;
            int dad = T->par[ib];
            for (int i = 0; i < T->ncol[dad]; i++) {
                int jb = T->col[27 * dad + i];
                if (T->nch[jb] == 0 && T->flag[jb] == 0 && t3_within(T, jb, ib, distest, 0))
                    T->flag[jb] = 2;
            }
        }
    }
    for (int l = 0; l <= L; l++) { tail[2 * l] = 0; tail[2 * l + 1] = -1; }
    for (int l = 1; l <= L - 2; l++) {
        tail[2 * (l + 1)] = T->nboxes;
        t3_refine_flagged(T, T->laddr[2 * l], T->laddr[2 * l + 1], l);
        tail[2 * (l + 1) + 1] = T->nboxes - 1;
    }
    t3_reorg(T, tail);
    t3_colleagues_all(T);
    for (int ib = 0; ib < T->nboxes; ib++)
        if (T->flag[ib] != 3) T->flag[ib] = 0;
    for (int l = 0; l <= L; l++) { tail[2 * l] = 0; tail[2 * l + 1] = -1; }
    for (int l = 2; l <= L - 2; l++) {
        t3_updateflags(T, l, T->laddr[2 * l], T->laddr[2 * l + 1]);
        t3_updateflags(T, l, tail[2 * l], tail[2 * l + 1]);
        tail[2 * (l + 1)] = T->nboxes;
        t3_refine_flagged(T, T->laddr[2 * l], T->laddr[2 * l + 1], l);
        t3_refine_flagged(T, tail[2 * l], tail[2 * l + 1], l);
        tail[2 * (l + 1) + 1] = T->nboxes - 1;
        t3_colleagues(T, tail[2 * (l + 1)], tail[2 * (l + 1) + 1]);
    }
    t3_reorg(T, tail);
    t3_colleagues_all(T);
}

/* pts_tree_mem + pts_tree_build, pts_tree3d.f:34-400 and :402-660, idivflag
 * 0, nlmin 0: a box splits when it holds more than ndiv0 (level < levcut) or
 * ndiv sources, up to nlmax; then colleagues and the level restriction. */
static void t3_build(o3tree *T, const double *src, int ns, int nlmax, int levcut, int ndiv0,
                     int ndiv, double bs0, const double *cen0)
{
    memset(T, 0, sizeof(*T));
    t3_grow(T, 64);
    T->bsz[0] = bs0;
    T->nboxes = 1;
    for (int d = 0; d < 3; d++) T->cen[d] = cen0[d];
    T->lev[0] = 0; T->par[0] = -1; T->nch[0] = 0; T->flag[0] = 0; T->ncol[0] = 0;
    for (int j = 0; j < 8; j++) T->ch[j] = -1;
    T->laddr[0] = 0; T->laddr[1] = 0;
    int *perm = xc(ns, sizeof(int)), *tmp = xc(ns, sizeof(int));
    for (int i = 0; i < ns; i++) perm[i] = i;
    int secap = 256;
    int *se = xc(2 * secap, sizeof(int));
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
        for (int ib = f0; ib <= f1; ib++) {
            if (se[2 * ib + 1] - se[2 * ib] <= lim) continue;
            t3_grow(T, T->nboxes + 8);
            if (T->nboxes + 8 > secap) {
                while (secap < T->nboxes + 8) secap *= 2;
                se = realloc(se, sizeof(int) * 2 * secap);
            }
            int first = T->nboxes;
            T->nboxes += 8;
            t3_make_children(T, ib, T->bsz[ilev + 1], ilev + 1, first);
            t3_sort_children(T, ib, src, perm, se, tmp);
        }
        T->laddr[2 * (ilev + 1) + 1] = T->nboxes - 1;
    }
    T->nlevels = ilev;
    free(perm); free(tmp); free(se);
    t3_colleagues_all(T);
    if (T->nlevels >= 2) t3_fix_lr(T);
}

/* pts_tree_sort, pts_tree3d.f:1330-1358. */
static void t3_sort_points(const o3tree *T, const double *p3, int n, int *perm, int *se)
{
    int *tmp = xc(n, sizeof(int));
    for (int i = 0; i < n; i++) perm[i] = i;
    se[0] = 0; se[1] = n;
    for (int l = 0; l < T->nlevels; l++)
        for (int ib = T->laddr[2 * l]; ib <= T->laddr[2 * l + 1]; ib++)
            if (T->nch[ib] > 0) t3_sort_children(T, ib, p3, perm, se, tmp);
    free(tmp);
}

/* the tree fgt3d builds (fgt3d.f:124-221): boxsize0, nlmax = first level
 * with nlocal <= 15, ndiv0 = 40, ndiv = 100 */
static int o3_nlmax(int npwlevel, double bs0, double delta, double eps)
{
    int nlevstart = npwlevel > 0 ? npwlevel : 0;
    int nlmax = npwlevel + 4;
    double bsize = bs0 / (double)(1 << nlevstart);
    for (int i = nlevstart; i <= 100; i++) {
        if (orc_hlterms(3, bsize, delta, eps) <= 15) { nlmax = i; break; }
        bsize = bsize / 2;
    }
    return nlmax;
}

typedef struct {
    o3tree T;
    int npwlevel, nlmax, ns, nt;
    double bs0, cen0[3];
    int *isrc, *isrcse, *itarg, *itargse;
} o3handle;

void *orc_tree3d_build(double delta, double eps, int ns, const double *src, int nt,
                       const double *targ, int *info)
{
    o3handle *H = xc(1, sizeof(*H));
    orc3_boxsize0(delta, eps, src, ns, targ, nt, &H->npwlevel, &H->bs0, H->cen0);
    H->nlmax = o3_nlmax(H->npwlevel, H->bs0, delta, eps);
    t3_build(&H->T, src, ns, H->nlmax, H->npwlevel, 40, 100, H->bs0, H->cen0);
    H->ns = ns; H->nt = nt;
    H->isrc = xc(ns, sizeof(int));
    H->itarg = xc(nt, sizeof(int));
    H->isrcse = xc(2 * (size_t)H->T.nboxes, sizeof(int));
    H->itargse = xc(2 * (size_t)H->T.nboxes, sizeof(int));
    t3_sort_points(&H->T, src, ns, H->isrc, H->isrcse);
    t3_sort_points(&H->T, targ, nt, H->itarg, H->itargse);
    info[0] = H->npwlevel; info[1] = H->nlmax; info[2] = H->T.nlevels; info[3] = H->T.nboxes;
    return H;
}

void orc_tree3d_get(void *h, int *laddr, int *lev, int *par, int *nch, int *ch, int *ncol,
                    int *col, double *cen, double *bsz, int *isrc, int *isrcse, int *itarg,
                    int *itargse, double *bs0cen0)
{
    o3handle *H = h;
    const o3tree *T = &H->T;
    int n = T->nboxes;
    memcpy(laddr, T->laddr, sizeof(int) * 2 * (T->nlevels + 1));
    memcpy(lev, T->lev, sizeof(int) * n);
    memcpy(par, T->par, sizeof(int) * n);
    memcpy(nch, T->nch, sizeof(int) * n);
    memcpy(ch, T->ch, sizeof(int) * 8 * n);
    memcpy(ncol, T->ncol, sizeof(int) * n);
    memcpy(col, T->col, sizeof(int) * 27 * n);
    memcpy(cen, T->cen, sizeof(double) * 3 * n);
    memcpy(bsz, T->bsz, sizeof(double) * (T->nlevels + 1));
    memcpy(isrc, H->isrc, sizeof(int) * H->ns);
    memcpy(itarg, H->itarg, sizeof(int) * H->nt);
    memcpy(isrcse, H->isrcse, sizeof(int) * 2 * n);
    memcpy(itargse, H->itargse, sizeof(int) * 2 * n);
    bs0cen0[0] = H->bs0;
    for (int d = 0; d < 3; d++) bs0cen0[1 + d] = H->cen0[d];
}

void orc_tree3d_free(void *h)
{
    o3handle *H = h;
    t3_free(&H->T);
    free(H->isrc); free(H->itarg); free(H->isrcse); free(H->itargse);
    free(H);
}

/* ------------------------------------------------------------------ */
/* expansions: 3d/planewave/src/g3drouts.f (complex interleaved)       */
/* PW layout (npw x, npw y, npw/2 z), x fastest; node index 0..npw-1   */
/* ascending, mirror of i is npw-1-i; the z half is the negative nodes */
/* ------------------------------------------------------------------ */

/* g3dformhc_vec, g3drouts.f:109-195: H(j,k,l) += q x^j/sqrt(j!) y^k/sqrt(k!)
 * z^l/sqrt(l!) (each power divided by sqrt(k) step by step). */
/* g3dformhc_vec / g3dformhd_vec / g3dformhcd_vec, g3drouts.f:109-409: with
 * dipoles (rn != NULL) H += dxhx_j hy_k hz_l w1 + hx_j (dyhy_k hz_l w2 +
 * hy_k (dzhz_l w3 + q hz_l)), dxhx_j = j dsq hx_{j-1}/sqrt(j), in the
 * reference's order of operations.  chg == NULL: no charges. */
static void o3_formhc(int nd, double delta, const double *src, int ns, const double *chg,
                      const double *rn, const double *ds, const double *c, int nt, double *h)
{
    int m = nt + 1;
    size_t m3 = (size_t)m * m * m;
    memset(h, 0, sizeof(double) * m3 * nd);
    double xp[301], yp[301], zp[301], sqk[301], dx[301], dy[301], dz[301];
    double dsq = 1.0 / sqrt(delta);
    for (int k = 1; k <= nt; k++) sqk[k] = sqrt((double)k);
    for (int i = 0; i < ns; i++) {
        double x = (src[3 * i] - c[0]) * dsq, y = (src[3 * i + 1] - c[1]) * dsq,
               z = (src[3 * i + 2] - c[2]) * dsq;
        xp[0] = yp[0] = zp[0] = 1.0;
        dx[0] = dy[0] = dz[0] = 0.0;
        for (int k = 1; k <= nt; k++) {
            xp[k] = xp[k - 1] * x / sqk[k];
            yp[k] = yp[k - 1] * y / sqk[k];
            zp[k] = zp[k - 1] * z / sqk[k];
            dx[k] = k * dsq * xp[k - 1] / sqk[k];
            dy[k] = k * dsq * yp[k - 1] / sqk[k];
            dz[k] = k * dsq * zp[k - 1] / sqk[k];
        }
        for (int ind = 0; ind < nd; ind++) {
            double *hh = h + (size_t)ind * m3;
            if (!rn) {
                double q = chg[(size_t)nd * i + ind];
                for (int l = 0; l <= nt; l++) {
                    double zt = q * zp[l];
                    for (int k = 0; k <= nt; k++) {
                        double rt = zt * yp[k];
                        for (int j = 0; j <= nt; j++) hh[j + m * (k + (size_t)m * l)] += rt * xp[j];
                    }
                }
                continue;
            }
            double dd = ds[(size_t)nd * i + ind];
            double r1 = rn[3 * i] * dd, r2 = rn[3 * i + 1] * dd, r3 = rn[3 * i + 2] * dd;
            for (int l = 0; l <= nt; l++) {
                double d1 = zp[l] * r1, d2 = zp[l] * r2, d3 = dz[l] * r3;
                double zt = chg ? chg[(size_t)nd * i + ind] * zp[l] : 0.0;
                for (int k = 0; k <= nt; k++) {
                    double c1 = yp[k] * d1;
                    double c2 = chg ? dy[k] * d2 + yp[k] * (d3 + zt) : dy[k] * d2 + yp[k] * d3;
                    double *hr = hh + m * (k + (size_t)m * l);
                    for (int j = 0; j <= nt; j++) hr[j] = hr[j] + dx[j] * c1 + xp[j] * c2;
                }
            }
        }
    }
}

/* g3drh2xmat, g3drouts.f:3019-3063: ws_i (-ts_i)^j / sqrt(j!), (0:nt, npw). */
void orc3_rh2xmat(int nt, int npw, const double *ws, const double *ts, double *m)
{
    double fac = 1.0;
    for (int j = 0; j <= nt; j++) {
        if (j > 0) fac = fac / sqrt((double)j);
        for (int i = 0; i < npw; i++) m[j + (nt + 1) * i] = ws[i] * o3_powi(-ts[i], j) * fac;
    }
}

/* g3drx2lmat, g3drouts.f:3106-3146: ts_i^j / j!, (npw, 0:nl). */
void orc3_rx2lmat(int nl, int npw, const double *ts, double *m)
{
    double fac = 1.0;
    for (int j = 0; j <= nl; j++) {
        if (j > 0) fac = fac * j;
        for (int i = 0; i < npw; i++) m[i + npw * j] = o3_powi(ts[i], j) / fac;
    }
}

/* sum_j a(j) i^j b(j) in four residue classes: returns c0-c2 (re part of
 * the even terms) and c1-c3 (coefficient of i); a real, b complex. */
#define I4SUM(nt, A, B, OUT_E, OUT_O)                                                  \
    do {                                                                               \
        double cr_[4] = {0, 0, 0, 0}, ci_[4] = {0, 0, 0, 0};                         \
        for (int r_ = 0; r_ < 4; r_++)                                                 \
            for (int j_ = r_; j_ <= (nt); j_ += 4) {                                   \
                double a_ = (A);                                                       \
                cr_[r_] += a_ * (B)[0];                                                \
                ci_[r_] += a_ * (B)[1];                                                \
            }                                                                          \
        (OUT_E)[0] = cr_[0] - cr_[2]; (OUT_E)[1] = ci_[0] - ci_[2];                    \
        (OUT_O)[0] = cr_[1] - cr_[3]; (OUT_O)[1] = ci_[1] - ci_[3];                    \
    } while (0)

/* g3dh2pw_real_vec, g3drouts.f:1565-1690: three 1-D transforms (x, then y,
 * then z over the first half), each sum_j h2x(j,k) i^j (.); pw incremented. */
static void o3_h2pw(int nd, int nt, int npw, const double *h2x, const double *h, double *pw)
{
    int m = nt + 1, npw2 = npw / 2;
    size_t m3 = (size_t)m * m * m;
    double *pxhyz = xc(2 * (size_t)npw * m * m, sizeof(double));
    double *pxyhz = xc(2 * (size_t)npw * npw * m, sizeof(double));
    for (int ind = 0; ind < nd; ind++) {
        const double *hh = h + (size_t)ind * m3;
        double *p = pw + (size_t)ind * 2 * npw * npw * npw2;
        for (int j1 = 0; j1 <= nt; j1++)
            for (int j2 = 0; j2 <= nt; j2++)
                for (int k3 = 0; k3 < npw2; k3++) {
                    double ce[2], co[2], b[2];
                    double c[4] = {0, 0, 0, 0};
                    for (int r = 0; r < 4; r++)
                        for (int j3 = r; j3 <= nt; j3 += 4)
                            c[r] += h2x[j3 + m * k3] * hh[j3 + m * (j2 + (size_t)m * j1)];
                    ce[0] = c[0] - c[2]; ce[1] = 0;
                    co[0] = 0; co[1] = c[1] - c[3];   /* eye*(c1-c3), real c */
                    (void)b;
                    double *o1 = pxhyz + 2 * (k3 + (size_t)npw * (j2 + (size_t)m * j1));
                    double *o2 = pxhyz + 2 * ((npw - 1 - k3) + (size_t)npw * (j2 + (size_t)m * j1));
                    o1[0] = ce[0] + co[0]; o1[1] = ce[1] + co[1];
                    o2[0] = ce[0] - co[0]; o2[1] = ce[1] - co[1];
                }
        for (int j1 = 0; j1 <= nt; j1++)
            for (int k2 = 0; k2 < npw2; k2++)
                for (int k3 = 0; k3 < npw; k3++) {
                    double e[2], o[2];
                    I4SUM(nt, h2x[j_ + m * k2], pxhyz + 2 * (k3 + (size_t)npw * (j_ + (size_t)m * j1)), e, o);
                    /* ce = e, co = eye*o = (-o.im, o.re) */
                    double *o1 = pxyhz + 2 * (k3 + (size_t)npw * (k2 + (size_t)npw * j1));
                    double *o2 = pxyhz + 2 * (k3 + (size_t)npw * ((npw - 1 - k2) + (size_t)npw * j1));
                    o1[0] = e[0] + (-o[1]); o1[1] = e[1] + o[0];
                    o2[0] = e[0] - (-o[1]); o2[1] = e[1] - o[0];
                }
        for (int k1 = 0; k1 < npw2; k1++)
            for (int k2 = 0; k2 < npw; k2++)
                for (int k3 = 0; k3 < npw; k3++) {
                    double e[2], o[2];
                    I4SUM(nt, h2x[j_ + m * k1], pxyhz + 2 * (k3 + (size_t)npw * (k2 + (size_t)npw * j_)), e, o);
                    double *d = p + 2 * (k3 + (size_t)npw * (k2 + (size_t)npw * k1));
                    /* pw += c0-c2 + eye*(c1-c3) */
                    double vr = e[0] + (-o[1]), vi = e[1] + o[0];
                    d[0] += vr;
                    d[1] += vi;
                }
    }
    free(pxhyz);
    free(pxyhz);
}

/* g3dformpwc_vec, g3drouts.f:1950-2040: PW(j1,j2,j3) += q ws ws ws
 * e^{-i ts x} e^{-i ts y} e^{-i ts z}, phases by recurrence from the first
 * positive node, negative nodes as conjugates. */
static inline void o3_cmul(double ar, double ai, double br, double bi, double *r, double *i)
{
    *r = ar * br - ai * bi;
    *i = ar * bi + ai * br;
}

/* with dipoles: g3dformpwd_vec / g3dformpwcd_vec, g3drouts.f:2042-2266,
 * PW += ww1x ry1 + ww1 ry2 with wwx = ww (-i ts dsq), ry1 = ww3 w1 ww2,
 * ry2 = ww3 w2 ww2y + (ww3z w3 + ww3 q) ww2, in the reference's order. */
static void o3_formpwc(int nd, double delta, const double *src, int ns, const double *chg,
                       const double *rn, const double *ds, const double *c, int npw,
                       const double *ws, const double *ts, double *pw)
{
    int npw2 = npw / 2;
    double dsq = 1.0 / sqrt(delta);
    double w[3][100][2], wd[3][100][2];
    for (int i = 0; i < ns; i++) {
        double u[3];
        for (int d = 0; d < 3; d++) u[d] = (src[3 * i + d] - c[d]) * dsq;
        for (int d = 0; d < 3; d++) {
            double qr = cos(-ts[npw2] * u[d]), qi = sin(-ts[npw2] * u[d]);
            double q1r = qr, q1i = qi;
            double q2r = qr * qr - qi * qi, q2i = qr * qi + qi * qr;
            for (int j = npw2; j < npw; j++) {
                w[d][j][0] = q1r * ws[j];
                w[d][j][1] = q1i * ws[j];
                double tr = q1r * q2r - q1i * q2i, ti = q1r * q2i + q1i * q2r;
                q1r = tr; q1i = ti;
                w[d][npw - 1 - j][0] = w[d][j][0];
                w[d][npw - 1 - j][1] = -w[d][j][1];
            }
            if (rn)
                for (int j = 0; j < npw; j++) {
                    /* ztmp = -i ts dsq; ww * ztmp with ztmp = (-0, -(ts dsq)) */
                    double zi = -(ts[j] * dsq);
                    wd[d][j][0] = -(w[d][j][1] * zi);
                    wd[d][j][1] = w[d][j][0] * zi;
                }
        }
        for (int ind = 0; ind < nd; ind++) {
            double *p = pw + (size_t)ind * 2 * npw * npw * npw2;
            if (!rn) {
                double q = chg[(size_t)nd * i + ind];
                for (int j3 = 0; j3 < npw2; j3++) {
                    double zr = q * w[2][j3][0], zi = q * w[2][j3][1];
                    for (int j2 = 0; j2 < npw; j2++) {
                        double yr = zr * w[1][j2][0] - zi * w[1][j2][1];
                        double yi = zr * w[1][j2][1] + zi * w[1][j2][0];
                        for (int j1 = 0; j1 < npw; j1++) {
                            double *d = p + 2 * (j1 + (size_t)npw * (j2 + (size_t)npw * j3));
                            d[0] += yr * w[0][j1][0] - yi * w[0][j1][1];
                            d[1] += yr * w[0][j1][1] + yi * w[0][j1][0];
                        }
                    }
                }
                continue;
            }
            double dd = ds[(size_t)nd * i + ind];
            double r1 = rn[3 * i] * dd, r2 = rn[3 * i + 1] * dd, r3 = rn[3 * i + 2] * dd;
            double q = chg ? chg[(size_t)nd * i + ind] : 0.0;
            for (int j3 = 0; j3 < npw2; j3++) {
                double z1r = w[2][j3][0] * r1, z1i = w[2][j3][1] * r1;
                double z2r = w[2][j3][0] * r2, z2i = w[2][j3][1] * r2;
                double z3r = wd[2][j3][0] * r3, z3i = wd[2][j3][1] * r3;
                if (chg) {  /* rz3 + ztmp, ztmp = ww3 q */
                    z3r = z3r + w[2][j3][0] * q;
                    z3i = z3i + w[2][j3][1] * q;
                }
                for (int j2 = 0; j2 < npw; j2++) {
                    double y1r, y1i, ar, ai, br, bi;
                    o3_cmul(z1r, z1i, w[1][j2][0], w[1][j2][1], &y1r, &y1i);
                    o3_cmul(z2r, z2i, wd[1][j2][0], wd[1][j2][1], &ar, &ai);
                    o3_cmul(z3r, z3i, w[1][j2][0], w[1][j2][1], &br, &bi);
                    double y2r = ar + br, y2i = ai + bi;
                    for (int j1 = 0; j1 < npw; j1++) {
                        double *d = p + 2 * (j1 + (size_t)npw * (j2 + (size_t)npw * j3));
                        double t1r, t1i, t2r, t2i;
                        o3_cmul(wd[0][j1][0], wd[0][j1][1], y1r, y1i, &t1r, &t1i);
                        o3_cmul(w[0][j1][0], w[0][j1][1], y2r, y2i, &t2r, &t2i);
                        d[0] = d[0] + t1r + t2r;
                        d[1] = d[1] + t1i + t2i;
                    }
                }
            }
        }
    }
}

/* g3dpw2local_real_vec, g3drouts.f:1774-1945: Taylor coefficients of
 * 2 Re sum PW e^{i ts x} e^{i ts y} e^{i ts z}; overwrites loc. */
static void o3_pw2local(int nd, int nl, int npw, const double *x2l, const double *pw,
                        double *loc)
{
    int npw2 = npw / 2, m = nl + 1;
    size_t m3 = (size_t)m * m * m;
    double *lx = xc(2 * (size_t)m * npw * npw2, sizeof(double));   /* (0:nl, npw, npw/2) */
    double *lxy = xc(2 * (size_t)m * m * npw2, sizeof(double));    /* (0:nl, 0:nl, npw/2) */
    for (int ind = 0; ind < nd; ind++) {
        const double *p = pw + (size_t)ind * 2 * npw * npw * npw2;
        double *L = loc + (size_t)ind * m3;
        /* x: pw(j3) +/- pw(mirror) against x2l over the first half */
        for (int j1 = 0; j1 < npw2; j1++)
            for (int j2 = 0; j2 < npw; j2++)
                for (int r = 0; r < 4; r++)
                    for (int k3 = r; k3 <= nl; k3 += 4) {
                        double cr = 0, ci = 0;
                        for (int j3 = 0; j3 < npw2; j3++) {
                            const double *a = p + 2 * (j3 + (size_t)npw * (j2 + (size_t)npw * j1));
                            const double *b = p + 2 * ((npw - 1 - j3) + (size_t)npw * (j2 + (size_t)npw * j1));
                            double sr = (r & 1) ? a[0] - b[0] : a[0] + b[0];
                            double si = (r & 1) ? a[1] - b[1] : a[1] + b[1];
                            cr += x2l[j3 + npw * k3] * sr;
                            ci += x2l[j3 + npw * k3] * si;
                        }
                        double *o = lx + 2 * (k3 + (size_t)m * (j2 + (size_t)npw * j1));
                        if (r == 0) { o[0] = cr; o[1] = ci; }
                        else if (r == 1) { o[0] = -ci; o[1] = cr; }     /* cd*eye */
                        else if (r == 2) { o[0] = -cr; o[1] = -ci; }
                        else { o[0] = ci; o[1] = -cr; }                 /* -eye*cd */
                    }
        /* y */
        for (int j1 = 0; j1 < npw2; j1++)
            for (int r = 0; r < 4; r++)
                for (int k2 = r; k2 <= nl; k2 += 4)
                    for (int k3 = 0; k3 <= nl; k3++) {
                        double cr = 0, ci = 0;
                        for (int j2 = 0; j2 < npw2; j2++) {
                            const double *a = lx + 2 * (k3 + (size_t)m * (j2 + (size_t)npw * j1));
                            const double *b = lx + 2 * (k3 + (size_t)m * ((npw - 1 - j2) + (size_t)npw * j1));
                            double sr = (r & 1) ? a[0] - b[0] : a[0] + b[0];
                            double si = (r & 1) ? a[1] - b[1] : a[1] + b[1];
                            cr += x2l[j2 + npw * k2] * sr;
                            ci += x2l[j2 + npw * k2] * si;
                        }
                        double *o = lxy + 2 * (k3 + (size_t)m * (k2 + (size_t)m * j1));
                        if (r == 0) { o[0] = cr; o[1] = ci; }
                        else if (r == 1) { o[0] = -ci; o[1] = cr; }
                        else if (r == 2) { o[0] = -cr; o[1] = -ci; }
                        else { o[0] = ci; o[1] = -cr; }
                    }
        /* z: first half only, 2 Re / 2 Im by residue */
        for (int r = 0; r < 4; r++)
            for (int k1 = r; k1 <= nl; k1 += 4)
                for (int k2 = 0; k2 <= nl; k2++)
                    for (int k3 = 0; k3 <= nl; k3++) {
                        double cr = 0, ci = 0;
                        for (int j1 = 0; j1 < npw2; j1++) {
                            const double *a = lxy + 2 * (k3 + (size_t)m * (k2 + (size_t)m * j1));
                            cr += x2l[j1 + npw * k1] * a[0];
                            ci += x2l[j1 + npw * k1] * a[1];
                        }
                        double v = r == 0 ? cr * 2 : r == 1 ? -ci * 2 : r == 2 ? -cr * 2 : ci * 2;
                        L[k3 + m * (k2 + (size_t)m * k1)] = v;
                    }
    }
    free(lx);
    free(lxy);
}

/* g3dlevalp/g/h_vec, g3drouts.f:1045-1365: sum L(j3,j2,j1) x^j3 y^j2 z^j1
 * and derivatives; pot/grad/hess incremented; hess order xx,yy,zz,xy,xz,yz. */
static void o3_leval(int nd, double delta, const double *c, int nl, const double *loc,
                     const double *tg, int nt, int ifpgh, double *pot, double *grad, double *hess)
{
    int m = nl + 1;
    size_t m3 = (size_t)m * m * m;
    double dsq = 1.0 / sqrt(delta);
    double xp[301], yp[301], zp[301], xpx[301], ypy[301], zpz[301], xpxx[301], ypyy[301], zpzz[301];
    for (int it = 0; it < nt; it++) {
        double x = (tg[3 * it] - c[0]) * dsq, y = (tg[3 * it + 1] - c[1]) * dsq,
               z = (tg[3 * it + 2] - c[2]) * dsq;
        xp[0] = yp[0] = zp[0] = 1.0;
        for (int i = 1; i <= nl; i++) {
            xp[i] = x * xp[i - 1];
            yp[i] = y * yp[i - 1];
            zp[i] = z * zp[i - 1];
        }
        xpx[0] = ypy[0] = zpz[0] = 0;
        for (int i = 1; i <= nl; i++) {
            xpx[i] = i * xp[i - 1] * dsq;
            ypy[i] = i * yp[i - 1] * dsq;
            zpz[i] = i * zp[i - 1] * dsq;
        }
        xpxx[0] = ypyy[0] = zpzz[0] = 0;
        if (nl >= 1) xpxx[1] = ypyy[1] = zpzz[1] = 0;
        for (int i = 2; i <= nl; i++) {
            xpxx[i] = i * (i - 1) * xp[i - 2] * dsq * dsq;
            ypyy[i] = i * (i - 1) * yp[i - 2] * dsq * dsq;
            zpzz[i] = i * (i - 1) * zp[i - 2] * dsq * dsq;
        }
        for (int ind = 0; ind < nd; ind++) {
            const double *L = loc + (size_t)ind * m3;
            double dd = 0, g1 = 0, g2 = 0, g3 = 0, h11 = 0, h22 = 0, h33 = 0, h12 = 0, h13 = 0,
                   h23 = 0;
            for (int j1 = 0; j1 <= nl; j1++) {
                double d1 = 0, d1x = 0, d1y = 0, d1xx = 0, d1yy = 0, d1xy = 0;
                for (int j2 = 0; j2 <= nl; j2++) {
                    double d2 = 0, d2x = 0, d2xx = 0;
                    const double *row = L + m * (j2 + (size_t)m * j1);
                    for (int j3 = 0; j3 <= nl; j3++) {
                        d2 = d2 + row[j3] * xp[j3];
                        if (ifpgh >= 2) d2x = d2x + row[j3] * xpx[j3];
                        if (ifpgh >= 3) d2xx = d2xx + row[j3] * xpxx[j3];
                    }
                    d1 = d1 + d2 * yp[j2];
                    if (ifpgh >= 2) {
                        d1x = d1x + d2x * yp[j2];
                        d1y = d1y + d2 * ypy[j2];
                    }
                    if (ifpgh >= 3) {
                        d1xx = d1xx + d2xx * yp[j2];
                        d1xy = d1xy + d2x * ypy[j2];
                        d1yy = d1yy + d2 * ypyy[j2];
                    }
                }
                dd = dd + d1 * zp[j1];
                if (ifpgh >= 2) {
                    g1 = g1 + d1x * zp[j1];
                    g2 = g2 + d1y * zp[j1];
                    g3 = g3 + d1 * zpz[j1];
                }
                if (ifpgh >= 3) {
                    h11 = h11 + d1xx * zp[j1];
                    h22 = h22 + d1yy * zp[j1];
                    h12 = h12 + d1xy * zp[j1];
                    h13 = h13 + d1x * zpz[j1];
                    h23 = h23 + d1y * zpz[j1];
                    h33 = h33 + d1 * zpzz[j1];
                }
            }
            pot[(size_t)nd * it + ind] += dd;
            if (ifpgh >= 2) {
                double *g = grad + (size_t)3 * nd * it;
                g[ind] += g1; g[nd + ind] += g2; g[2 * nd + ind] += g3;
            }
            if (ifpgh >= 3) {
                double *h = hess + (size_t)6 * nd * it;
                h[ind] += h11; h[nd + ind] += h22; h[2 * nd + ind] += h33;
                h[3 * nd + ind] += h12; h[4 * nd + ind] += h13; h[5 * nd + ind] += h23;
            }
        }
    }
}

static inline void cfma(double *z, const double *a, const double *b)
{
    z[0] += a[0] * b[0] - a[1] * b[1];
    z[1] += a[0] * b[1] + a[1] * b[0];
}

/* g3dpwevalp/g/h_vec, g3drouts.f:2268-2668: 2 Re sum PW w1 w2 w3 with
 * w = e^{i ts u} by recurrence, derivative factors i ts dsq. */
static void o3_pweval(int nd, double delta, const double *c, int npw, const double *ts,
                      const double *pw, const double *tg, int nt, int ifpgh, double *pot,
                      double *grad, double *hess)
{
    int npw2 = npw / 2;
    double dsq = 1.0 / sqrt(delta);
    double w[3][3][100][2];  /* [axis][order][node] */
    for (int it = 0; it < nt; it++) {
        for (int d = 0; d < 3; d++) {
            double u = (tg[3 * it + d] - c[d]) * dsq;
            double qr = cos(ts[npw2] * u), qi = sin(ts[npw2] * u);
            double q1r = qr, q1i = qi;
            double q2r = qr * qr - qi * qi, q2i = qr * qi + qi * qr;
            for (int j = npw2; j < npw; j++) {
                int jm = npw - 1 - j;
                w[d][0][j][0] = q1r; w[d][0][j][1] = q1i;
                w[d][0][jm][0] = q1r; w[d][0][jm][1] = -q1i;
                double zt = ts[j] * dsq;   /* eye*ts*dsq */
                w[d][1][j][0] = -zt * q1i; w[d][1][j][1] = zt * q1r;
                w[d][1][jm][0] = w[d][1][j][0]; w[d][1][jm][1] = -w[d][1][j][1];
                w[d][2][j][0] = -zt * w[d][1][j][1]; w[d][2][j][1] = zt * w[d][1][j][0];
                w[d][2][jm][0] = w[d][2][j][0]; w[d][2][jm][1] = -w[d][2][j][1];
                double tr = q1r * q2r - q1i * q2i, ti = q1r * q2i + q1i * q2r;
                q1r = tr; q1i = ti;
            }
        }
        for (int ind = 0; ind < nd; ind++) {
            const double *p = pw + (size_t)ind * 2 * npw * npw * npw2;
            double dd[2] = {0, 0}, g1[2] = {0, 0}, g2[2] = {0, 0}, g3[2] = {0, 0};
            double h11[2] = {0, 0}, h22[2] = {0, 0}, h33[2] = {0, 0}, h12[2] = {0, 0},
                   h13[2] = {0, 0}, h23[2] = {0, 0};
            for (int j3 = 0; j3 < npw2; j3++) {
                double d1[2] = {0, 0}, d1x[2] = {0, 0}, d1y[2] = {0, 0}, d1xx[2] = {0, 0},
                       d1yy[2] = {0, 0}, d1xy[2] = {0, 0};
                for (int j2 = 0; j2 < npw; j2++) {
                    double d2[2] = {0, 0}, d2x[2] = {0, 0}, d2xx[2] = {0, 0};
                    for (int j1 = 0; j1 < npw; j1++) {
                        const double *a = p + 2 * (j1 + (size_t)npw * (j2 + (size_t)npw * j3));
                        cfma(d2, a, w[0][0][j1]);
                        if (ifpgh >= 2) cfma(d2x, a, w[0][1][j1]);
                        if (ifpgh >= 3) cfma(d2xx, a, w[0][2][j1]);
                    }
                    cfma(d1, d2, w[1][0][j2]);
                    if (ifpgh >= 2) {
                        cfma(d1x, d2x, w[1][0][j2]);
                        cfma(d1y, d2, w[1][1][j2]);
                    }
                    if (ifpgh >= 3) {
                        cfma(d1xx, d2xx, w[1][0][j2]);
                        cfma(d1xy, d2x, w[1][1][j2]);
                        cfma(d1yy, d2, w[1][2][j2]);
                    }
                }
                cfma(dd, d1, w[2][0][j3]);
                if (ifpgh >= 2) {
                    cfma(g1, d1x, w[2][0][j3]);
                    cfma(g2, d1y, w[2][0][j3]);
                    cfma(g3, d1, w[2][1][j3]);
                }
                if (ifpgh >= 3) {
                    cfma(h11, d1xx, w[2][0][j3]);
                    cfma(h22, d1yy, w[2][0][j3]);
                    cfma(h12, d1xy, w[2][0][j3]);
                    cfma(h13, d1x, w[2][1][j3]);
                    cfma(h23, d1y, w[2][1][j3]);
                    cfma(h33, d1, w[2][2][j3]);
                }
            }
            pot[(size_t)nd * it + ind] += dd[0] * 2;
            if (ifpgh >= 2) {
                double *g = grad + (size_t)3 * nd * it;
                g[ind] += g1[0] * 2; g[nd + ind] += g2[0] * 2; g[2 * nd + ind] += g3[0] * 2;
            }
            if (ifpgh >= 3) {
                double *h = hess + (size_t)6 * nd * it;
                h[ind] += h11[0] * 2; h[nd + ind] += h22[0] * 2; h[2 * nd + ind] += h33[0] * 2;
                h[3 * nd + ind] += h12[0] * 2; h[4 * nd + ind] += h13[0] * 2;
                h[5 * nd + ind] += h23[0] * 2;
            }
        }
    }
}

/* g3d_directcp/cg/ch_vec, g3ddirect.f:3-206: pairs with r^2 < dmax,
 * pot += q e^{-r^2/delta}, grad = -2/delta d * (.), hess likewise.  With
 * dipoles (rn != NULL) the six variants g3d_directdp/dg/dh and
 * g3d_directcdp/cdg/cdh, g3ddirect.f:207-700, in their order of operations;
 * Hessians ordered xx, yy, zz, xy, xz, yz. */
static void o3_direct(int nd, double delta, double dmax, const double *src, int ns,
                      const double *chg, const double *rn, const double *ds, const double *tg,
                      int nt, int ifpgh, double *pot, double *grad, double *hess)
{
    for (int it = 0; it < nt; it++)
        for (int i = 0; i < ns; i++) {
            double xd = tg[3 * it] - src[3 * i], yd = tg[3 * it + 1] - src[3 * i + 1],
                   zd = tg[3 * it + 2] - src[3 * i + 2];
            double rr = xd * xd + yd * yd + zd * zd;
            if (!(rr < dmax)) continue;
            double *P = pot + (size_t)nd * it, *G = grad + (size_t)3 * nd * it,
                   *H = hess + (size_t)6 * nd * it;
            const double *q = chg ? chg + (size_t)nd * i : NULL;
            const double *d = rn ? ds + (size_t)nd * i : NULL;
            double n1 = rn ? rn[3 * i] : 0, n2 = rn ? rn[3 * i + 1] : 0, n3 = rn ? rn[3 * i + 2] : 0;
            if (!rn) {
                double rfac = -2.0 / delta;
                double rt = exp(-rr / delta);
                double dx = rfac * xd, dy = rfac * yd, dz = rfac * zd;
                for (int ii = 0; ii < nd; ii++) {
                    double rinc = rt * q[ii];
                    P[ii] += rinc;
                    if (ifpgh >= 2) { G[ii] += dx * rinc; G[nd + ii] += dy * rinc; G[2 * nd + ii] += dz * rinc; }
                    if (ifpgh >= 3) {
                        H[ii] += (rfac + dx * dx) * rinc;
                        H[nd + ii] += (rfac + dy * dy) * rinc;
                        H[2 * nd + ii] += (rfac + dz * dz) * rinc;
                        H[3 * nd + ii] += dx * dy * rinc;
                        H[4 * nd + ii] += dx * dz * rinc;
                        H[5 * nd + ii] += dy * dz * rinc;
                    }
                }
                continue;
            }
            if (!q && ifpgh == 1) { /* g3d_directdp_vec */
                double rfac = 2.0 / delta;
                double rtmp = exp(-rr / delta);
                double dx = rfac * xd, dy = rfac * yd, dz = rfac * zd;
                double rinc = (dx * n1 + dy * n2 + dz * n3) * rtmp;
                for (int ii = 0; ii < nd; ii++) P[ii] = P[ii] + rinc * d[ii];
            } else if (!q) { /* g3d_directdg_vec / g3d_directdh_vec */
                double rfac = 2.0 / delta;
                double dtmp = exp(-rr / delta);
                double dx = rfac * xd, dy = rfac * yd, dz = rfac * zd;
                double rinc = (dx * n1 + dy * n2 + dz * n3);
                double rx = n1 * rfac - rinc * dx, ry = n2 * rfac - rinc * dy, rz = n3 * rfac - rinc * dz;
                double rxx = -(rfac * (dx * n1 + rinc)) - rx * dx;
                double ryy = -(rfac * (dy * n2 + rinc)) - ry * dy;
                double rzz = -(rfac * (dz * n3 + rinc)) - rz * dz;
                double rxy = -(rfac * dx * n2) - rx * dy;
                double rxz = -(rfac * dx * n3) - rx * dz;
                double ryz = -(rfac * dy * n3) - ry * dz;
                for (int ii = 0; ii < nd; ii++) {
                    double rt = dtmp * d[ii];
                    P[ii] = P[ii] + rinc * rt;
                    G[ii] = G[ii] + rx * rt;
                    G[nd + ii] = G[nd + ii] + ry * rt;
                    G[2 * nd + ii] = G[2 * nd + ii] + rz * rt;
                    if (ifpgh >= 3) {
                        H[ii] = H[ii] + rxx * rt;
                        H[nd + ii] = H[nd + ii] + ryy * rt;
                        H[2 * nd + ii] = H[2 * nd + ii] + rzz * rt;
                        H[3 * nd + ii] = H[3 * nd + ii] + rxy * rt;
                        H[4 * nd + ii] = H[4 * nd + ii] + rxz * rt;
                        H[5 * nd + ii] = H[5 * nd + ii] + ryz * rt;
                    }
                }
            } else if (ifpgh == 1) { /* g3d_directcdp_vec */
                double rfac = 2.0 / delta;
                double rtmp = exp(-rr / delta);
                for (int ii = 0; ii < nd; ii++) P[ii] = P[ii] + rtmp * q[ii];
                double dx = rfac * xd, dy = rfac * yd, dz = rfac * zd;
                double rinc = (dx * n1 + dy * n2 + dz * n3) * rtmp;
                for (int ii = 0; ii < nd; ii++) P[ii] = P[ii] + rinc * d[ii];
            } else { /* g3d_directcdg_vec / g3d_directcdh_vec */
                double rfac = -2.0 / delta;
                double dtmp = exp(-rr / delta);
                double dx = rfac * xd, dy = rfac * yd, dz = rfac * zd;
                double dxx = rfac + dx * dx, dyy = rfac + dy * dy, dzz = rfac + dz * dz;
                double dxy = dx * dy, dxz = dx * dz, dyz = dy * dz;
                for (int ii = 0; ii < nd; ii++) {
                    double rinc = dtmp * q[ii];
                    P[ii] = P[ii] + rinc;
                    G[ii] = G[ii] + dx * rinc;
                    G[nd + ii] = G[nd + ii] + dy * rinc;
                    G[2 * nd + ii] = G[2 * nd + ii] + dz * rinc;
                    if (ifpgh >= 3) {
                        H[ii] = H[ii] + dxx * rinc;
                        H[nd + ii] = H[nd + ii] + dyy * rinc;
                        H[2 * nd + ii] = H[2 * nd + ii] + dzz * rinc;
                        H[3 * nd + ii] = H[3 * nd + ii] + dxy * rinc;
                        H[4 * nd + ii] = H[4 * nd + ii] + dxz * rinc;
                        H[5 * nd + ii] = H[5 * nd + ii] + dyz * rinc;
                    }
                }
                double rinc = -(dx * n1) - dy * n2 - dz * n3;
                double rx = rinc * dx - n1 * rfac, ry = rinc * dy - n2 * rfac, rz = rinc * dz - n3 * rfac;
                double rxx = rfac * (rinc - dx * n1) + rx * dx;
                double ryy = rfac * (rinc - dy * n2) + ry * dy;
                double rzz = rfac * (rinc - dz * n3) + rz * dz;
                double rxy = rx * dy - rfac * dx * n2;
                double rxz = rx * dz - rfac * dx * n3;
                double ryz = ry * dz - rfac * dy * n3;
                for (int ii = 0; ii < nd; ii++) {
                    double rt = dtmp * d[ii];
                    P[ii] = P[ii] + rinc * rt;
                    G[ii] = G[ii] + rx * rt;
                    G[nd + ii] = G[nd + ii] + ry * rt;
                    G[2 * nd + ii] = G[2 * nd + ii] + rz * rt;
                    if (ifpgh >= 3) {
                        H[ii] = H[ii] + rxx * rt;
                        H[nd + ii] = H[nd + ii] + ryy * rt;
                        H[2 * nd + ii] = H[2 * nd + ii] + rzz * rt;
                        H[3 * nd + ii] = H[3 * nd + ii] + rxy * rt;
                        H[4 * nd + ii] = H[4 * nd + ii] + rxz * rt;
                        H[5 * nd + ii] = H[5 * nd + ii] + ryz * rt;
                    }
                }
            }
        }
}

/* fgt3dpart_direct_vec (fgt3d.f:1533-1684) over whole arrays: increments. */
void orc_direct3d(int nd, double delta, double dmax, const double *src, int ns,
                  const double *chg, const double *tg, int nt, int ifpgh, double *pot,
                  double *grad, double *hess)
{
    o3_direct(nd, delta, dmax, src, ns, chg, NULL, NULL, tg, nt, ifpgh, pot, grad, hess);
}

/* the same with dipoles: chg == NULL no charges, rn == NULL no dipoles */
void orc_direct3d_cd(int nd, double delta, double dmax, const double *src, int ns,
                     const double *chg, const double *rn, const double *ds, const double *tg,
                     int nt, int ifpgh, double *pot, double *grad, double *hess)
{
    o3_direct(nd, delta, dmax, src, ns, chg, rn, ds, tg, nt, ifpgh, pot, grad, hess);
}

/* the intended g3dformlc_vec (g3drouts.f:784-856): Taylor coefficients of
 * sum q e^{-|t-s|^2/delta} about c from the Hermite-function recurrence
 * h_{j+1} = 2 (u h_j - h_{j-1})/(j+1), h_0 = e^{-u^2}, per axis.  Dipoles
 * (the intended g3dformld_vec / g3dformlcd_vec, g3drouts.f:857-1040, whose
 * x factors are indexed by the y order j2 where the x order j3 is meant)
 * add w . (dh_x h_y h_z, h_x dh_y h_z, h_x h_y dh_z), dh_j = -dsq (j+1)
 * h_{j+1}. */
static void o3_formlc(int nd, double delta, const double *src, int ns, const double *chg,
                      const double *rn, const double *ds, const double *c, int nl, double *loc)
{
    int m = nl + 1;
    size_t m3 = (size_t)m * m * m;
    double hx[302], hy[302], hz[302], dhx[301], dhy[301], dhz[301];
    double dsq = 1.0 / sqrt(delta);
    for (int i = 0; i < ns; i++) {
        double u[3];
        for (int d = 0; d < 3; d++) u[d] = (src[3 * i + d] - c[d]) * dsq;
        double *hh[3] = {hx, hy, hz}, *dh[3] = {dhx, dhy, dhz};
        int top = rn ? nl : nl - 1;
        for (int d = 0; d < 3; d++) {
            double f = exp(-u[d] * u[d]);
            hh[d][0] = 1.0 * f;
            hh[d][1] = 2.0 * u[d] * f;
            for (int j = 1; j <= top; j++) hh[d][j + 1] = 2.0 * (u[d] * hh[d][j] - hh[d][j - 1]) / (j + 1);
            if (rn)
                for (int j = 0; j <= nl; j++) dh[d][j] = -(hh[d][j + 1] * dsq * (j + 1.0));
        }
        for (int ind = 0; ind < nd; ind++) {
            double q = chg ? chg[(size_t)nd * i + ind] : 0.0;
            double *L = loc + (size_t)ind * m3;
            if (!rn) {
                for (int j1 = 0; j1 <= nl; j1++) {
                    double zt = q * hz[j1];
                    for (int j2 = 0; j2 <= nl; j2++) {
                        double yt = zt * hy[j2];
                        for (int j3 = 0; j3 <= nl; j3++) L[j3 + m * (j2 + (size_t)m * j1)] += yt * hx[j3];
                    }
                }
                continue;
            }
            double dd = ds[(size_t)nd * i + ind];
            double w1 = rn[3 * i] * dd, w2 = rn[3 * i + 1] * dd, w3 = rn[3 * i + 2] * dd;
            for (int j1 = 0; j1 <= nl; j1++)
                for (int j2 = 0; j2 <= nl; j2++) {
                    double a = hy[j2] * hz[j1];                                  /* x-derivative term */
                    double b = (q * hy[j2] + w2 * dhy[j2]) * hz[j1] + w3 * hy[j2] * dhz[j1];
                    for (int j3 = 0; j3 <= nl; j3++)
                        L[j3 + m * (j2 + (size_t)m * j1)] += w1 * a * dhx[j3] + b * hx[j3];
                }
        }
    }
}

/* merge_split_pw_matrices (g3drouts.f:2760-2842), table k (0..7) for shift
 * number kk: x factor conjugated for k in {2,3,6,7}, y for {1,3,5,7}, z for
 * k >= 4 (the pp p, pm p, mp p, mm p, pp m, ... order). */
static void o3_ms_tables(int npw, const double *ts, double xmin, int nmax, double *w)
{
    int npw2 = npw / 2;
    size_t nexp = (size_t)npw * npw * npw2;
    double (*ww)[2] = xc((size_t)npw * (nmax + 1), sizeof(double[2]));
    for (int j = 0; j < npw; j++) {
        double zr = cos(ts[j] * xmin), zi = sin(ts[j] * xmin);
        for (int k = 1; k <= nmax; k++) {
            ww[j * (nmax + 1) + k][0] = zr;
            ww[j * (nmax + 1) + k][1] = zi;
            double tr = zr * zr - zi * zi, ti = zr * zi + zi * zr;
            zr = tr; zi = ti;
        }
    }
    for (int k1 = 1; k1 <= nmax; k1++)
        for (int j1 = 0; j1 < npw2; j1++)
            for (int j2 = 0; j2 < npw; j2++)
                for (int j3 = 0; j3 < npw; j3++) {
                    size_t j = j3 + (size_t)npw * (j2 + (size_t)npw * j1);
                    const double *a = ww[j3 * (nmax + 1) + k1], *b = ww[j2 * (nmax + 1) + k1],
                                 *cz = ww[j1 * (nmax + 1) + k1];
                    for (int k = 0; k < 8; k++) {
                        double ar = a[0], ai = (k == 2 || k == 3 || k == 6 || k == 7) ? -a[1] : a[1];
                        double br = b[0], bi = (k & 1) ? -b[1] : b[1];
                        double cr = cz[0], ci = k >= 4 ? -cz[1] : cz[1];
                        double tr = ar * br - ai * bi, ti = ar * bi + ai * br;
                        double *o = w + 2 * ((size_t)(k + 8 * (k1 - 1)) * nexp + j);
                        o[0] = tr * cr - ti * ci;
                        o[1] = tr * ci + ti * cr;
                    }
                }
    free(ww);
}

/* pw_translation_matrices (g3drouts.f:2697-2757) with nmax = 1: table
 * (kx,ky,kz) in {-1,0,1}^3 at index (kx+1) + 3 (ky+1) + 9 (kz+1). */
static void o3_m2l_tables(int npw, const double *ts, double xmin, double *w)
{
    int npw2 = npw / 2;
    size_t nexp = (size_t)npw * npw * npw2;
    double (*ww)[3][2] = xc(npw, sizeof(double[3][2]));
    for (int j = 0; j < npw; j++) {
        double zr = cos(ts[j] * xmin), zi = sin(ts[j] * xmin);
        ww[j][1][0] = 1; ww[j][1][1] = 0;
        ww[j][2][0] = zr; ww[j][2][1] = zi;
        ww[j][0][0] = zr; ww[j][0][1] = -zi;
    }
    for (int kz = 0; kz < 3; kz++)
        for (int ky = 0; ky < 3; ky++)
            for (int kx = 0; kx < 3; kx++)
                for (int j1 = 0; j1 < npw2; j1++)
                    for (int j2 = 0; j2 < npw; j2++)
                        for (int j3 = 0; j3 < npw; j3++) {
                            size_t j = j3 + (size_t)npw * (j2 + (size_t)npw * j1);
                            const double *a = ww[j3][kx], *b = ww[j2][ky], *c = ww[j1][kz];
                            double tr = a[0] * b[0] - a[1] * b[1], ti = a[0] * b[1] + a[1] * b[0];
                            double *o = w + 2 * ((size_t)(kx + 3 * ky + 9 * kz) * nexp + j);
                            o[0] = tr * c[0] - ti * c[1];
                            o[1] = tr * c[1] + ti * c[0];
                        }
    free(ww);
}

static void o3_shift(int nd, size_t nexp, const double *src, double *dst, const double *w)
{
    for (int ind = 0; ind < nd; ind++)
        for (size_t j = 0; j < nexp; j++) {
            const double *a = src + 2 * ((size_t)ind * nexp + j);
            double *d = dst + 2 * ((size_t)ind * nexp + j);
            d[0] = d[0] + (a[0] * w[2 * j] - a[1] * w[2 * j + 1]);
            d[1] = d[1] + (a[0] * w[2 * j + 1] + a[1] * w[2 * j]);
        }
}

/* octant of a shift by its sign pattern, fgt3d.f:1120-1136: 0 (+,+,+),
 * 1 (+,-,+), 2 (-,+,+), 3 (-,-,+), then the same four with z < 0 */
static int o3_octant(double dx, double dy, double dz)
{
    int k = (dx > 0) ? ((dy > 0) ? 0 : 1) : ((dy > 0) ? 2 : 3);
    return dz > 0 ? k : k + 4;
}

static int g3_last[8 + O3_MAXLEV];

void orc3_last_stats(int *out)
{
    for (int i = 0; i < 8; i++) out[i] = g3_last[i];
}

/* fgt3d (fgt3d.f:19-480) and fgt3dmain (:483-1530) for charges. */
int orc_fgt3d(int nd, double delta, double eps, int ns, const double *sources,
              int ifcharge, const double *charge, int ifdipole, const double *rnormal,
              const double *dipstr, int ifpgh, double *pot, double *grad, double *hess, int nt,
              const double *targ, int ifpghtarg, double *pottarg, double *gradtarg,
              double *hesstarg)
{
    if (ifcharge < 0 || ifcharge > 1 || ifdipole < 0 || ifdipole > 1) return -1;
    if (!ifcharge) charge = NULL;
    if (!ifdipole) rnormal = dipstr = NULL;
    int npwlevel;
    double bs0, cen0[3];
    orc3_boxsize0(delta, eps, sources, ns, targ, nt, &npwlevel, &bs0, cen0);
    memset(g3_last, 0, sizeof(g3_last));
    g3_last[0] = npwlevel;
    if (npwlevel < -2) {
        /* fgt3d_large_delta, fgt3d.f:1783-1914 (outputs incremented) */
        int nl0 = orc_hlterms(3, bs0, delta, eps) + (ifpgh > ifpghtarg ? ifpgh : ifpghtarg) - 1;
        size_t m3 = (size_t)(nl0 + 1) * (nl0 + 1) * (nl0 + 1);
        double *loc = xc(m3 * nd, sizeof(double));
        if (charge || rnormal) o3_formlc(nd, delta, sources, ns, charge, rnormal, dipstr, cen0, nl0, loc);
        if (ifpghtarg >= 1) o3_leval(nd, delta, cen0, nl0, loc, targ, nt, ifpghtarg, pottarg, gradtarg, hesstarg);
        if (ifpgh >= 1) o3_leval(nd, delta, cen0, nl0, loc, sources, ns, ifpgh, pot, grad, hess);
        free(loc);
        g3_last[1] = nl0;
        return 0;
    }
    int nlmax = o3_nlmax(npwlevel, bs0, delta, eps);
    o3tree T;
    t3_build(&T, sources, ns, nlmax, npwlevel, 40, 100, bs0, cen0);
    int nb = T.nboxes, nlevels = T.nlevels;
    int *isrc = xc(ns, sizeof(int)), *itarg = xc(nt, sizeof(int));
    int *sse = xc(2 * (size_t)nb, sizeof(int)), *tse = xc(2 * (size_t)nb, sizeof(int));
    t3_sort_points(&T, sources, ns, isrc, sse);
    t3_sort_points(&T, targ, nt, itarg, tse);
    g3_last[1] = nlmax; g3_last[2] = nlevels; g3_last[3] = nb;

    /* sorted copies (dreorderf) and zeroed sorted outputs */
    double *ss = xc(3 * (size_t)ns, sizeof(double)), *ts3 = xc(3 * (size_t)nt + 3, sizeof(double));
    double *cs = xc((size_t)nd * ns, sizeof(double));
    double *rs = xc(3 * (size_t)ns, sizeof(double)), *dss = xc((size_t)nd * ns, sizeof(double));
    for (int i = 0; i < ns; i++) {
        for (int d = 0; d < 3; d++) ss[3 * i + d] = sources[3 * (size_t)isrc[i] + d];
        if (charge)
            for (int k = 0; k < nd; k++) cs[(size_t)nd * i + k] = charge[(size_t)nd * isrc[i] + k];
        if (rnormal) {
            for (int d = 0; d < 3; d++) rs[3 * i + d] = rnormal[3 * (size_t)isrc[i] + d];
            for (int k = 0; k < nd; k++) dss[(size_t)nd * i + k] = dipstr[(size_t)nd * isrc[i] + k];
        }
    }
    const double *cq = charge ? cs : NULL, *crn = rnormal ? rs : NULL;
    for (int i = 0; i < nt; i++)
        for (int d = 0; d < 3; d++) ts3[3 * i + d] = targ[3 * (size_t)itarg[i] + d];
    double *ps = xc((size_t)nd * ns + 1, sizeof(double)), *gs = xc((size_t)3 * nd * ns + 1, sizeof(double)),
           *hs = xc((size_t)6 * nd * ns + 1, sizeof(double));
    double *pt = xc((size_t)nd * nt + 1, sizeof(double)), *gt = xc((size_t)3 * nd * nt + 1, sizeof(double)),
           *ht = xc((size_t)6 * nd * nt + 1, sizeof(double));

    int nlevstart = npwlevel > 0 ? npwlevel : 0;
    int nlocal[O3_MAXLEV + 1];
    memset(nlocal, 0, sizeof(nlocal));
    int nadd = ifdipole + (ifpgh > ifpghtarg ? ifpgh : ifpghtarg); /* fgt3d.f:355-360 */
    for (int l = nlevstart; l <= nlevels; l++) {
        nlocal[l] = orc_hlterms(3, T.bsz[l], delta, eps);
        if (ifdipole == 1 || ifpgh >= 2 || ifpghtarg >= 2) nlocal[l] += nadd;
    }
    double pmax;
    int npw;
    {
        double bsize = 2 * bs0 / pow(2.0, nlevstart);
        double pweps = nadd > 2 ? eps / 100 : eps;
        orc_pwterms(bsize, delta, pweps, &pmax, &npw);
    }
    g3_last[4] = npw;
    int nlevend = npwlevel < nlevels ? npwlevel : nlevels;
    int npwh = npw / 2;
    size_t nexp = (size_t)npw * npw * npwh;

    /* ifhung, fgt3d.f:744-763 */
    char *ifhung = xc(nb, 1);
    for (int l = 0; l <= nlevend; l++)
        for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; b++)
            if (T.nch[b] == 0 && sse[2 * b + 1] > sse[2 * b]) ifhung[b] = 1;

    /* compute_modified_list1, tree_routs3d.f:511-620 */
    int mnl1 = 27 + 27 + 27 * 4 + 8;
    int *nl1 = xc(nb, sizeof(int)), *l1 = xc((size_t)mnl1 * nb, sizeof(int));
    if (T.nch[0] == 0) { nl1[0] = 1; l1[0] = 0; }
    for (int l = 1; l <= nlevels; l++)
        for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; b++) {
            if (T.nch[b] != 0) continue;
            int dad = T.par[b];
            for (int i = 0; i < T.ncol[b]; i++) {
                int jb = T.col[27 * b + i];
                if (l >= npwlevel || T.nch[jb] == 0) {
                    l1[(size_t)mnl1 * b + nl1[b]++] = jb;
                } else {
                    double dt = 1.05 * (T.bsz[l] + T.bsz[l + 1]) / 2.0;
                    for (int j = 0; j < 8; j++) {
                        int kb = T.ch[8 * jb + j];
                        if (kb >= 0 && t3_within(&T, kb, b, dt, 1)) l1[(size_t)mnl1 * b + nl1[b]++] = kb;
                    }
                }
            }
            for (int i = 0; i < T.ncol[dad]; i++) {
                int jb = T.col[27 * dad + i];
                if (T.nch[jb] == 0) {
                    double dt = 1.05 * (T.bsz[l] + T.bsz[l - 1]) / 2.0;
                    if (t3_within(&T, jb, b, dt, 1)) l1[(size_t)mnl1 * b + nl1[b]++] = jb;
                }
            }
        }

    if (npwlevel < nlevels) {
        double *ws = xc(npw, sizeof(double)), *tsn = xc(npw, sizeof(double));
        orc3_pwnodes(pmax, npw, ws, tsn);
        double xmin = T.bsz[nlevstart] / sqrt(delta);
        double *wm2l = xc(2 * 27 * nexp, sizeof(double));
        o3_m2l_tables(npw, tsn, xmin, wm2l);
        int nmax = nlevels - nlevstart;
        double *wms = xc(2 * 8 * nexp * (size_t)(nmax > 0 ? nmax : 1), sizeof(double));
        if (nmax > 0) o3_ms_tables(npw, tsn, T.bsz[nlevels] / sqrt(delta) / 2, nmax, wms);
        int ntermmax = 0;
        for (int l = 0; l <= nlevels; l++) if (nlocal[l] > ntermmax) ntermmax = nlocal[l];
        double *rh2x = xc((size_t)(ntermmax + 1) * npw, sizeof(double));
        orc3_rh2xmat(ntermmax, npw, ws, tsn, rh2x);
        double *rx2l = xc((size_t)(ntermmax + 1) * npw, sizeof(double));
        orc3_rx2lmat(ntermmax, npw, tsn, rx2l);

        /* iflocal, fgt3d.f:838-862 */
        char *iflocal = xc(nb, 1);
        if (npwlevel >= 0 && npwlevel <= nlevels)
            for (int b = T.laddr[2 * npwlevel]; b <= T.laddr[2 * npwlevel + 1]; b++)
                if (T.nch[b] == 0)
                    for (int i = 0; i < T.ncol[b]; i++)
                        if (T.nch[T.col[27 * b + i]] > 0) { iflocal[b] = 1; break; }

        /* expansion slots: boxes at levels >= nlevstart */
        long *slot = xc(nb, sizeof(long));
        long nsl = 0;
        for (int b = 0; b < nb; b++) slot[b] = T.lev[b] >= nlevstart ? nsl++ : -1;
        double *mp = xc(2 * nexp * nd * (size_t)(nsl ? nsl : 1), sizeof(double));
        double *lc = xc(2 * nexp * nd * (size_t)(nsl ? nsl : 1), sizeof(double));
#define MP(b) (mp + 2 * nexp * nd * (size_t)slot[b])
#define LC(b) (lc + 2 * nexp * nd * (size_t)slot[b])

        /* step 1: formation, fgt3d.f:940-1095 */
        for (int l = nlevels; l >= (npwlevel + 1 > 0 ? npwlevel + 1 : 0); l--) {
            int p = nlocal[l];
            double r1 = (p + 1.0) / npw, r3 = r1 * r1 * r1;
            int nptsswitch = r3 >= 3 ? 1000000 : 0;
            size_t m = p + 1;
            double *hexp = xc(m * m * m * nd, sizeof(double));
            double *h2xt = xc(m * npw, sizeof(double));
            for (int i = 0; i < npw; i++)
                for (int j = 0; j <= p; j++) h2xt[j + m * i] = rh2x[j + (size_t)(ntermmax + 1) * i];
            for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; b++) {
                int np = sse[2 * b + 1] - sse[2 * b];
                if (T.nch[b] != 0 || np <= 0) continue;
                if (!cq && !crn) continue; /* fgt3d.f:952-1090: no branch forms anything */
                const double *s = ss + 3 * (size_t)sse[2 * b];
                const double *q = cq ? cs + (size_t)nd * sse[2 * b] : NULL;
                const double *r = crn ? rs + 3 * (size_t)sse[2 * b] : NULL;
                const double *dd = dss + (size_t)nd * sse[2 * b];
                if (np > nptsswitch) {
                    o3_formhc(nd, delta, s, np, q, r, dd, T.cen + 3 * b, p, hexp);
                    o3_h2pw(nd, p, npw, h2xt, hexp, MP(b));
                } else {
                    o3_formpwc(nd, delta, s, np, q, r, dd, T.cen + 3 * b, npw, ws, tsn, MP(b));
                }
            }
            free(hexp);
            free(h2xt);
        }
        /* step 2: merge, fgt3d.f:1100-1160 */
        for (int l = nlevels - 1; l >= nlevstart; l--) {
            int klev = nlevels - l;
            for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; b++)
                for (int i = 0; i < T.nch[b]; i++) {
                    int c = T.ch[8 * b + i];
                    if (sse[2 * c + 1] - sse[2 * c] <= 0) continue;
                    int k = o3_octant(T.cen[3 * b] - T.cen[3 * c], T.cen[3 * b + 1] - T.cen[3 * c + 1],
                                      T.cen[3 * b + 2] - T.cen[3 * c + 2]);
                    o3_shift(nd, nexp, MP(c), MP(b), wms + 2 * (size_t)(k + 8 * (klev - 1)) * nexp);
                }
        }
        /* step 3: m2l at nlevstart, fgt3d.f:1162-1210 */
        {
            int l = nlevstart;
            double xm = T.bsz[nlevstart];
            for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; b++) {
                int np = 0;
                if (ifpghtarg > 0) np += tse[2 * b + 1] - tse[2 * b];
                if (ifpgh > 0) np += sse[2 * b + 1] - sse[2 * b];
                if (np <= 0) continue;
                if (l == npwlevel)
                    for (int i = 0; i < T.ncol[b]; i++) {
                        int jb = T.col[27 * b + i];
                        if (jb == b || T.nch[jb] <= 0) continue;   /* gt3d_computelistpw */
                        int jx = (int)lround((T.cen[3 * b] - T.cen[3 * jb]) / xm);
                        int jy = (int)lround((T.cen[3 * b + 1] - T.cen[3 * jb + 1]) / xm);
                        int jz = (int)lround((T.cen[3 * b + 2] - T.cen[3 * jb + 2]) / xm);
                        o3_shift(nd, nexp, MP(jb), LC(b),
                                 wm2l + 2 * (size_t)((jx + 1) + 3 * (jy + 1) + 9 * (jz + 1)) * nexp);
                    }
                for (int ind = 0; ind < nd; ind++)
                    for (size_t j = 0; j < 2 * nexp; j++)
                        LC(b)[2 * nexp * ind + j] += MP(b)[2 * nexp * ind + j];
            }
        }
        /* step 4: split, fgt3d.f:1215-1276 */
        for (int l = nlevstart; l <= nlevels - 1; l++)
            for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; b++) {
                int np = 0;
                if (ifpghtarg > 0) np += tse[2 * b + 1] - tse[2 * b];
                if (ifpgh > 0) np += sse[2 * b + 1] - sse[2 * b];
                if (np <= 0) continue;
                for (int i = 0; i < T.nch[b]; i++) {
                    int c = T.ch[8 * b + i];
                    int k = o3_octant(T.cen[3 * c] - T.cen[3 * b], T.cen[3 * c + 1] - T.cen[3 * b + 1],
                                      T.cen[3 * c + 2] - T.cen[3 * b + 2]);
                    o3_shift(nd, nexp, LC(b), LC(c), wms + 2 * (size_t)(k + 8 * (nlevels - l - 1)) * nexp);
                }
            }
        /* step 5: evaluation, fgt3d.f:1285-1455 */
        for (int l = nlevstart; l <= nlevels; l++) {
            int nl = nlocal[l];
            size_t m = nl + 1;
            double r1 = (nl + 1.0) / npw, r3 = r1 * r1 * r1;
            int nptsswitch = r3 >= 3 ? 1000000 : 0;
            double *loc = xc(m * m * m * nd, sizeof(double));
            double *x2lt = xc(m * npw, sizeof(double));
            for (int j = 0; j <= nl; j++)
                for (int i = 0; i < npw; i++) x2lt[i + npw * j] = rx2l[i + (size_t)npw * j];
            for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; b++) {
                if (l == npwlevel && !iflocal[b]) continue;
                if (T.nch[b] != 0) continue;
                int ntg = tse[2 * b + 1] - tse[2 * b], nsr = sse[2 * b + 1] - sse[2 * b];
                if (ntg + nsr <= 0) continue;
                if (ntg + nsr < nptsswitch) {
                    if (ntg > 0 && ifpghtarg > 0)
                        o3_pweval(nd, delta, T.cen + 3 * b, npw, tsn, LC(b), ts3 + 3 * (size_t)tse[2 * b], ntg,
                                  ifpghtarg, pt + (size_t)nd * tse[2 * b], gt + (size_t)3 * nd * tse[2 * b],
                                  ht + (size_t)6 * nd * tse[2 * b]);
                    if (nsr > 0 && ifpgh > 0)
                        o3_pweval(nd, delta, T.cen + 3 * b, npw, tsn, LC(b), ss + 3 * (size_t)sse[2 * b], nsr,
                                  ifpgh, ps + (size_t)nd * sse[2 * b], gs + (size_t)3 * nd * sse[2 * b],
                                  hs + (size_t)6 * nd * sse[2 * b]);
                } else {
                    o3_pw2local(nd, nl, npw, x2lt, LC(b), loc);
                    if (ntg > 0 && ifpghtarg > 0)
                        o3_leval(nd, delta, T.cen + 3 * b, nl, loc, ts3 + 3 * (size_t)tse[2 * b], ntg,
                                 ifpghtarg, pt + (size_t)nd * tse[2 * b], gt + (size_t)3 * nd * tse[2 * b],
                                 ht + (size_t)6 * nd * tse[2 * b]);
                    if (nsr > 0 && ifpgh > 0)
                        o3_leval(nd, delta, T.cen + 3 * b, nl, loc, ss + 3 * (size_t)sse[2 * b], nsr,
                                 ifpgh, ps + (size_t)nd * sse[2 * b], gs + (size_t)3 * nd * sse[2 * b],
                                 hs + (size_t)6 * nd * sse[2 * b]);
                }
            }
            free(loc);
            free(x2lt);
        }
#undef MP
#undef LC
        free(ws); free(tsn); free(wm2l); free(wms); free(rh2x); free(rx2l);
        free(iflocal); free(slot); free(mp); free(lc);
    }

    /* step 6: near field, fgt3d.f:1460-1512, dmax = ln(1/eps) delta */
    double dmax = log(1.0 / eps) * delta;
    for (int l = 0; l <= nlevend; l++)
        for (int jb = T.laddr[2 * l]; jb <= T.laddr[2 * l + 1]; jb++) {
            if (!ifhung[jb]) continue;
            int js = sse[2 * jb], je = sse[2 * jb + 1];
            for (int i = 0; i < nl1[jb]; i++) {
                int ib = l1[(size_t)mnl1 * jb + i];
                int t0 = tse[2 * ib], t1 = tse[2 * ib + 1], s0 = sse[2 * ib], s1 = sse[2 * ib + 1];
                const double *q = cq ? cs + (size_t)nd * js : NULL;
                const double *r = crn ? rs + 3 * (size_t)js : NULL, *dd = dss + (size_t)nd * js;
                if (!q && !r) continue; /* fgt3dpart_direct_vec: no branch applies */
                if (t1 > t0 && ifpghtarg > 0)
                    o3_direct(nd, delta, dmax, ss + 3 * (size_t)js, je - js, q, r, dd,
                              ts3 + 3 * (size_t)t0, t1 - t0, ifpghtarg, pt + (size_t)nd * t0,
                              gt + (size_t)3 * nd * t0, ht + (size_t)6 * nd * t0);
                if (s1 > s0 && ifpgh > 0)
                    o3_direct(nd, delta, dmax, ss + 3 * (size_t)js, je - js, q, r, dd,
                              ss + 3 * (size_t)s0, s1 - s0, ifpgh, ps + (size_t)nd * s0,
                              gs + (size_t)3 * nd * s0, hs + (size_t)6 * nd * s0);
            }
        }

    /* back to input order (dreorderi / dreorderiv): overwrite */
    for (int i = 0; i < ns && ifpgh > 0; i++) {
        size_t o = isrc[i];
        for (int k = 0; k < nd; k++) pot[nd * o + k] = ps[(size_t)nd * i + k];
        if (ifpgh >= 2) for (int k = 0; k < 3 * nd; k++) grad[3 * nd * o + k] = gs[(size_t)3 * nd * i + k];
        if (ifpgh >= 3) for (int k = 0; k < 6 * nd; k++) hess[6 * nd * o + k] = hs[(size_t)6 * nd * i + k];
    }
    for (int i = 0; i < nt && ifpghtarg > 0; i++) {
        size_t o = itarg[i];
        for (int k = 0; k < nd; k++) pottarg[nd * o + k] = pt[(size_t)nd * i + k];
        if (ifpghtarg >= 2) for (int k = 0; k < 3 * nd; k++) gradtarg[3 * nd * o + k] = gt[(size_t)3 * nd * i + k];
        if (ifpghtarg >= 3) for (int k = 0; k < 6 * nd; k++) hesstarg[6 * nd * o + k] = ht[(size_t)6 * nd * i + k];
    }
    for (int l = 0; l <= nlevels && 6 + l < 8 + O3_MAXLEV; l++) g3_last[6 + l] = nlocal[l];
    g3_last[5] = 0;
    for (int b = 0; b < nb; b++) g3_last[5] += T.nch[b] == 0;
    free(ifhung); free(nl1); free(l1);
    free(isrc); free(itarg); free(sse); free(tse); free(ss); free(ts3); free(cs); free(rs); free(dss);
    free(ps); free(gs); free(hs); free(pt); free(gt); free(ht);
    t3_free(&T);
    return 0;
}
