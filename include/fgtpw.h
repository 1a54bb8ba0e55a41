/*
 * fgtpw.h -- C-ABI of the MI355X plane-wave fast Gauss transform.
 *
 * Every entry point takes plain pointers and sizes and returns a status
 * (FGTPW_OK or a negative code); fgtpw_last_error() gives the message of the
 * last failure on the calling thread.  Array layouts are the reference's
 * Fortran layouts: points (2,n) column-major (x,y interleaved), densities
 * (nd,n), gradients (nd,2,n), Hessians (nd,3,n) ordered xx,xy,yy.
 *
 * Outputs are overwritten, not accumulated, for the quantities requested by
 * ifpgh / ifpghtarg (1 pot, 2 pot+grad, 3 pot+grad+hess); arrays of
 * quantities not requested are not touched and may be dummies.
 *
 * Supported in this build: charges (ifcharge = 1) and dipoles (ifdipole = 1,
 * rnormal (dim,n), dipstr (nd,n)), alone or together, any nd >= 1, in 2D
 * (fgt2d) and 3D (fgt3d).  With neither, the outputs are zero.
 */
#ifndef FGTPW_H
#define FGTPW_H

#ifdef __cplusplus
extern "C" {
#endif

#define FGTPW_OK 0
#define FGTPW_EINVAL (-1)      /* invalid argument (sizes, flags, non-finite points) */
#define FGTPW_ENOMEM (-2)      /* device or host allocation failed */
#define FGTPW_EUNSUPPORTED (-3) /* valid for the reference, not built here */
#define FGTPW_EDEVICE (-4)     /* HIP runtime error */

/* Text of the last error on this thread ("" if none). */
const char *fgtpw_last_error(void);

/* Optional: bind the library to a HIP device (default: the current device
 * at first use).  Returns FGTPW_OK. */
int fgtpw_init(int device);

/*
 * Replaces fgt2d (2d/planewave/src/fgt2d.f:19-22), same arguments and
 * meaning, host arrays.  iper is ignored, as in the reference (fgt2d.f:40).
 */
int fgtpw_fgt2d(int nd, double delta, double eps, int ns, const double *sources, int ifcharge,
                const double *charge, int ifdipole, const double *rnormal, const double *dipstr,
                int iper, int ifpgh, double *pot, double *grad, double *hess, int nt,
                const double *targ, int ifpghtarg, double *pottarg, double *gradtarg,
                double *hesstarg);

/*
 * Same transform with every array already in device memory (HBM); runs on
 * `stream` (a hipStream_t, NULL for the default stream) and returns when the
 * outputs are complete.  No reference counterpart: it is the same fgt2d
 * contract for callers whose data live on the GPU.
 */
int fgtpw_fgt2d_device(int nd, double delta, double eps, int ns, const double *sources,
                       int ifcharge, const double *charge, int ifdipole, const double *rnormal,
                       const double *dipstr, int iper, int ifpgh, double *pot, double *grad,
                       double *hess, int nt, const double *targ, int ifpghtarg, double *pottarg,
                       double *gradtarg, double *hesstarg, void *stream);

/*
 * fgtpw_fgt2d_device without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * inside a HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Every device buffer of the transform comes out of one static
 * scratch array of 256 MiB per device (held by the library until the
 * process ends); host staging is pageable; all work runs on `stream`.  It
 * calls no hipMalloc, hipFree, hipHostMalloc or allocating library routine
 * (the host tree build uses the host heap).  Returns FGTPW_EINVAL when the
 * point-proportional buffers of (ns, nt, nd) exceed the scratch (about
 * 2.0M sources + 2.0M targets for nd = 1, potentials only), and
 * FGTPW_ENOMEM when the tree and expansions of the input do not fit beside
 * them.  Every call on a device shares that device's array: calls are
 * serialised by the library lock and return after `stream` has drained;
 * two processes' calls do not share it (each holds its own copy); work on
 * one device that overlapped a call would race on the array.  Results are
 * bit-identical to fgtpw_fgt2d_device's.
 */
int fgtpw_fgt2d_device_noalloc(int nd, double delta, double eps, int ns, const double *sources,
                               int ifcharge, const double *charge, int ifdipole,
                               const double *rnormal, const double *dipstr, int iper, int ifpgh,
                               double *pot, double *grad, double *hess, int nt, const double *targ,
                               int ifpghtarg, double *pottarg, double *gradtarg, double *hesstarg,
                               void *stream);

/*
 * Replaces fgt2dpart_direct_vec (2d/planewave/src/fgt2d.f:1354-1512): adds to
 * pot/grad/hess at targets jstart..jend (1-based, inclusive) the direct
 * Gaussian sum over sources istart..iend, pairs with r^2 >= ln(1/eps) delta
 * skipped (g2ddirect.f:59-62).  Accumulates, as the reference does.  Host
 * arrays.
 */
int fgtpw_fgt2dpart_direct(int nd, double delta, double eps, int istart, int iend, int jstart,
                           int jend, const double *source, int ifcharge, const double *charge,
                           int ifdipole, const double *rnormal, const double *dipstr,
                           const double *targ, int ifpgh, double *pot, double *grad,
                           double *hess);

/*
 * Replaces fgt3d (3d/planewave/src/fgt3d.f:19-63), same arguments and
 * meaning, host arrays: points (3,n), gradients (nd,3,n), Hessians (nd,6,n)
 * ordered xx, yy, zz, xy, xz, yz.  iper is ignored, as in the reference.
 */
int fgtpw_fgt3d(int nd, double delta, double eps, int ns, const double *sources, int ifcharge,
                const double *charge, int ifdipole, const double *rnormal, const double *dipstr,
                int iper, int ifpgh, double *pot, double *grad, double *hess, int nt,
                const double *targ, int ifpghtarg, double *pottarg, double *gradtarg,
                double *hesstarg);

/* fgtpw_fgt3d with every array in device memory, on `stream` (a hipStream_t,
 * NULL for the default stream); returns when the outputs are complete.  No
 * reference counterpart. */
int fgtpw_fgt3d_device(int nd, double delta, double eps, int ns, const double *sources,
                       int ifcharge, const double *charge, int ifdipole, const double *rnormal,
                       const double *dipstr, int iper, int ifpgh, double *pot, double *grad,
                       double *hess, int nt, const double *targ, int ifpghtarg, double *pottarg,
                       double *gradtarg, double *hesstarg, void *stream);

/*
 * Replaces fgt3dpart_direct_vec (3d/planewave/src/fgt3d.f:1533-1684): adds
 * to pot/grad/hess at targets jstart..jend (1-based) the direct Gaussian sum
 * over sources istart..iend, pairs with r^2 >= dmax skipped (the 3D routine
 * takes dmax itself, g3ddirect.f:3-206).  Accumulates.  Host arrays.
 */
int fgtpw_fgt3dpart_direct(int nd, double delta, double dmax, int istart, int iend, int jstart,
                           int jend, const double *source, int ifcharge, const double *charge,
                           int ifdipole, const double *rnormal, const double *dipstr,
                           const double *targ, int ifpgh, double *pot, double *grad,
                           double *hess);

/*
 * Diagnostics of the last transform on this thread: the integers the
 * reference prints (fgt2d.f:127, :188, :196-197, :358-361) -- out[0]
 * npwlevel, out[1] nlmax, out[2] nlevels, out[3] nboxes, out[4] npw,
 * out[5] number of leaves, out[6..6+nlevels] nlocal per level.  n is the
 * length of out.  Returns the number of entries written.
 */
int fgtpw_last_stats(int *out, int n);

/*
 * Arithmetic of the next transforms (process-wide; returns the previous
 * mode, or FGTPW_EINVAL).  No reference counterpart: the reference computes
 * in FP64 only.
 *   0  FP64 everywhere (default): the reference's arithmetic.
 *   1  mixed precision for 2D transforms with eps >= 1e-6 that form
 *      expansions: the leaf moments and the Taylor evaluation run on FP32
 *      matrix cores (operands rounded to FP32 from offsets formed in FP64
 *      about FP64 box centres; FP32 accumulation within a segment of at most
 *      256 points, FP64 across segments and into the outputs), and the
 *      plane-wave expansions and their translation factors are FP32 through
 *      merge, m2l and split.  The Hermite->plane-wave and plane-wave->Taylor
 *      conversions, plane-wave evaluation phases, the tree and the near
 *      field stay FP64.  Other transforms (3D, eps < 1e-6) run in FP64.
 * fgtpw_last_precision() returns 1 when the last transform of the calling
 * thread took the mixed route, else 0.
 */
int fgtpw_set_precision(int mode);
int fgtpw_last_precision(void);

/*
 * The tile shapes the 2D kernels use, as "segf1=.. sege1=.. spw1=.. ...
 * sort=..;source=<where from>": points per formation and evaluation segment and
 * segments per wavefront of the fused evaluation for each order tile count
 * (orders up to 16, 32, 48, 64), and the radix sort's block shape.  They
 * come from defaults, from tuning.txt beside the library (written by
 * python -m fgt2d_amd.autotune) and from FGTPW_TUNE="key=value,...", read
 * once per process.  No reference counterpart.
 */
const char *fgtpw_tuning_table(void);

/*
 * Stage timing of the next transforms: when on (1), each call records HIP
 * events around its stages on its stream and fgtpw_stage_times() returns the
 * milliseconds of the last call per stage, in the order of
 * fgtpw_stage_name(i).  Returns the number of stages written.  on = 2 does
 * the same with the work of the side streams (near field, coarse levels)
 * queued on the call's stream instead, so that no two kernels overlap and
 * the per-kernel times of fgtpw_kernel_launches are each kernel's own.
 */
int fgtpw_set_timing(int on);
int fgtpw_stage_times(double *ms, int n);
const char *fgtpw_stage_name(int i);

/*
 * Per-kernel timing of the last call made with timing on: one entry per
 * launch of an instrumented kernel, in launch order -- kernel id (name from
 * fgtpw_kernel_name), tree level (-1 when not per level), milliseconds
 * between HIP events recorded on the launch's stream right before and after
 * it, and the units it processed (formation, evaluation: points of its
 * boxes; translations: boxes; near field, gather, scatter: points).  Fills
 * up to cap entries and returns the number of launches.  No reference
 * counterpart (the reference prints step times, fgt2d.f:1338).
 */
int fgtpw_kernel_launches(int *id, int *level, double *ms, long long *units, int cap);
const char *fgtpw_kernel_name(int id);

/*
 * Inspection of the quadtree of the last successful fgtpw_fgt2d* call on this
 * process (the tree fgt2d builds before its main routine, fgt2d.f:191-221),
 * 0-based: info[0] nlevels, info[1] nboxes, info[2] leaves, info[3] nlmax.
 * When laddr is non-NULL and cap >= nboxes the arrays are filled: laddr
 * (2,nlevels+1), lev, par, nch (nboxes), ch (4,nboxes), ncol, col
 * (9,nboxes), cen (2,nboxes), isrc (ns) and itarg (nt) permutations, and
 * isrcse/itargse (2,nboxes) as first index and one past the last.  The
 * layout is that of itree (pts_tree2d4.f:19-27) with 0-based box numbers.
 * No reference counterpart: the reference keeps its tree internal.
 */
int fgtpw_last_tree2d(int *info, int cap, int *laddr, int *lev, int *par, int *nch, int *ch,
                      int *ncol, int *col, double *cen, int *isrc, int *isrcse, int *itarg,
                      int *itargse);

/*
 * The octree of the last successful fgtpw_fgt3d* call (fgt3d.f:191-221),
 * 0-based, arrays as fgtpw_last_tree2d with ch (8,nboxes), col (27,nboxes)
 * and cen (3,nboxes) (pts_tree3d.f:34-400 itree layout).  No reference
 * counterpart.
 */
int fgtpw_last_tree3d(int *info, int cap, int *laddr, int *lev, int *par, int *nch, int *ch,
                      int *ncol, int *col, double *cen, int *isrc, int *isrcse, int *itarg,
                      int *itargse);

/*
 * Serving.  One transform request: the arguments of fgt2d / fgt3d (dim 2 or
 * 3), host arrays.  Replaces nothing in the reference, which has no request
 * interface; the fields mean what the fgt2d arguments of the same names mean
 * (2d/planewave/src/fgt2d.f:23-62; fgt3d.f:19-63 for dim 3).
 */
typedef struct fgtpw_call {
    int dim;
    int nd;
    double delta, eps;
    int ns;
    const double *sources; /* (dim, ns) */
    int ifcharge;
    const double *charge;  /* (nd, ns) */
    int ifdipole;
    const double *rnormal, *dipstr;
    int ifpgh;
    double *pot, *grad, *hess;
    int nt;
    const double *targ;    /* (dim, nt) */
    int ifpghtarg;
    double *pottarg, *gradtarg, *hesstarg;
} fgtpw_call;

/*
 * Starts the request server on the current device: `slots` worker slots
 * (1..16), each with its own stream and device buffers that stay allocated
 * between requests, and an HBM budget in GiB for the requests in flight
 * (0: 60% of the free device memory).  Requests are admitted in submission
 * order when a slot is free and their estimated footprint fits beside those
 * in flight (one request is always admitted when none runs).
 */
int fgtpw_serve_start(int slots, double hbm_budget_gib);

/*
 * Queues one transform and returns a ticket (> 0) or a negative status.
 * The input arrays are copied before it returns; the output arrays must stay
 * valid until fgtpw_wait for the ticket returns, and are written by then.
 * Thread-safe.
 */
long long fgtpw_submit(const fgtpw_call *call);

/*
 * Blocks until the request of `ticket` has finished and returns its status.
 * When ms is non-NULL: ms[0] queue time, ms[1] run time, ms[2] submit to
 * completion, in milliseconds.  Each ticket is waited for once.
 */
int fgtpw_wait(long long ticket, double *ms);

/*
 * Server counters: out[0] slots, out[1] budget bytes, out[2] queued,
 * out[3] running, out[4] bytes in flight, out[5] peak running, out[6] peak
 * bytes in flight.  Returns the number written (0 when no server runs).
 */
int fgtpw_serve_stats(long long *out, int n);

/* Finishes every queued request, then stops the workers and frees their
 * buffers.  Call fgtpw_wait for every ticket before. */
int fgtpw_serve_stop(void);

/*
 * Host-only: the box tree the runtime derives from a list of non-empty
 * leaves of the adaptive tree (Morton prefix keys[i] of level levels[i], in
 * Morton order) -- refinement, level restriction (pts_tree_fix_lr,
 * pts_tree2d4.f:693) and renumbering -- with the outputs of
 * fgtpw_last_tree2d.  No GPU needed; for tests of the host logic.
 */
int fgtpw_tree2d_from_leaves(int nlmax, int npwlevel, double bs0, const double *cen0,
                             int nleaves, const unsigned long long *keys, const int *levels,
                             int *info, int cap, int *laddr, int *lev, int *par, int *nch,
                             int *ch, int *ncol, int *col, double *cen);

#ifdef __cplusplus
}
#endif
#endif
