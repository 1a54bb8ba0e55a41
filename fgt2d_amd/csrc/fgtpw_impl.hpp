// Internal declarations shared by the host runtime (api.cpp, tree2d.cpp) and
// the HIP kernels (kernels2d.hip).  Not part of the public C-ABI, which is
// include/fgtpw.h.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <tuple>
#include <vector>

namespace fgtpw {

// ---------------------------------------------------------------------------
// error handling
// ---------------------------------------------------------------------------
struct Error {
    int code;
    std::string msg;
};
[[noreturn]] void fail(int code, const std::string& msg);
void check_hip(hipError_t e, const char* what);
#define HIPCHECK(x) ::fgtpw::check_hip((x), #x)

// ---------------------------------------------------------------------------
// per-kernel timing: with a clock installed on the calling thread, the
// launchers bracket the kernels named here with HIP events on their stream
// ---------------------------------------------------------------------------
enum KernelId {
    KID_FORM_MOMENTS, KID_FORM_H2PW, KID_FORM_VALU, KID_MERGE, KID_M2L, KID_SPLIT, KID_PW2LOCAL,
    KID_TAYLOR_EVAL, KID_PW_EVAL, KID_DIRECT, KID_GATHER, KID_SCATTER, KID_SORT, KID_N
};
struct KernelClock {
    // events recorded on the stream before (end = false) and after the
    // launches of kernel `id` that follow
    virtual void mark(int id, bool end, hipStream_t s) = 0;
    // a pair of events for one launch of kernel `id`, stamped by the
    // dispatch itself at the kernel's start and end (hipExtLaunchKernel)
    virtual void pair(int id, hipEvent_t* e0, hipEvent_t* e1) = 0;
    virtual ~KernelClock() = default;
};
extern thread_local KernelClock* t_kclock;
inline void kmark(int id, bool end, hipStream_t s)
{
    if (t_kclock) t_kclock->mark(id, end, s);
}

// Launch of an instrumented kernel: with a clock installed, the dispatch
// stamps a start and an end event of its own (the kernel's own duration,
// without the gap before it when the queue ran dry); otherwise a plain
// launch.
template <class... P, class... A>
void klaunch(int id, void (*f)(P...), dim3 grid, dim3 block, unsigned shm, hipStream_t s,
             A... a)
{
    if (!t_kclock) {
        f<<<grid, block, shm, s>>>(a...);
        return;
    }
    hipEvent_t e0, e1;
    t_kclock->pair(id, &e0, &e1);
    // the kernel's own parameter types, so that the argument buffer holds
    // exactly what the kernel reads
    auto args = std::tuple<P...>{a...};
    void* ptrs[sizeof...(P)];
    std::apply([&](auto&... v) {
        int k = 0;
        ((ptrs[k++] = (void*)&v), ...);
    }, args);
    HIPCHECK(hipExtLaunchKernel((const void*)f, grid, block, ptrs, shm, s, e0, e1, 0));
}

// ---------------------------------------------------------------------------
// term counts (g2dterms.f) and plane-wave nodes (g2drouts.f:2205)
// ---------------------------------------------------------------------------
int hermite_terms(int ndim, double bsize, double delta, double eps);
void planewave_terms(double bsize, double delta, double eps, double* pmax, int* npw);
void planewave_nodes(double pmax, int npw, double* ws, double* ts);

// ---------------------------------------------------------------------------
// host box tree (built from the GPU's list of pre-restriction leaves)
// ---------------------------------------------------------------------------
struct LeafRun {            // one non-empty leaf of the adaptive tree before
    uint64_t key;           // level restriction: its Morton prefix (level
    int level;              // digits, most significant first), its level
    int start;              // and its first position in the key-sorted source list
};

// Box tree in the reference's itree layout (D = 2: quadtree, D = 3: octree)
template <int D>
struct TreeD {
    static constexpr int CH = 1 << D;          // children per box
    static constexpr int NC = D == 2 ? 9 : 27;  // colleague slots per box
    int nlmax = 0, npwlevel = 0, nlevels = 0;
    double bs0 = 0, cen0[D] = {};
    std::vector<double> bsz;   // box size per level
    std::vector<int> laddr;    // first, last box of each level
    std::vector<int> lev, par, nch, ch, ncol, col, flag;
    std::vector<double> cen;   // D per box
    std::vector<uint64_t> pre;  // Morton prefix of every box
    std::vector<int> ic[D];     // integer box coordinates within the level, per axis
    int coll_levels = -1;       // colleague lists are valid for levels <= this
    int nboxes() const { return (int)lev.size(); }
    // leaves in Morton order and, per box, its interval of leaf ranks
    std::vector<int> leaf_of_rank, rank_of_box, rank_lo, rank_hi;
    // builder scratch, kept to reuse its storage
    std::vector<int> map, tail, tail2;
};
using Tree2D = TreeD<2>;
using Tree3D = TreeD<3>;

template <int D>
void build_tree(TreeD<D>& T, const std::vector<LeafRun>& leaves, int nlmax, int npwlevel,
                double bs0, const double* cen0);
template <int D>
void complete_colleagues(TreeD<D>& T);
inline void build_tree2d(Tree2D& T, const std::vector<LeafRun>& leaves, int nlmax, int npwlevel,
                         double bs0, const double* cen0)
{
    build_tree<2>(T, leaves, nlmax, npwlevel, bs0, cen0);
}
void planewave_nodes3(double pmax, int npw, double* ws, double* ts);
// host memcpy split over the worker threads of the tree builder
void parallel_copy(void* dst, const void* src, size_t n);

// ---------------------------------------------------------------------------
// device kernels (kernels2d.hip)
// ---------------------------------------------------------------------------
struct DevArena;  // fwd

void launch_bbox(const double* a, int na, const double* b, int nb, double* part, int nblk,
                 double* out, hipStream_t s);
// Morton keys of 64 bits, or of 32 bits when the key width (bits per
// level x nlmax) is at most 31
void launch_keys2d(const double* xy, int n, double cx, double cy, double bs0, int nlmax,
                   unsigned long long* key, unsigned* idx, hipStream_t s);
void launch_keys2d(const double* xy, int n, double cx, double cy, double bs0, int nlmax,
                   unsigned* key, unsigned* idx, hipStream_t s);
// bits: key bits per level (2 in 2D, 3 in 3D); bcnt: the number of leaf
// starts of every block of 256 points (leaf_records_temp(n) bytes)
void launch_leafdepth(const unsigned long long* key, int n, int nlmax, int levcut, int ndiv0,
                      int ndiv, int* level, unsigned char* start_flag, unsigned* bcnt,
                      hipStream_t s, int bits = 2);
void launch_leafdepth(const unsigned* key, int n, int nlmax, int levcut, int ndiv0, int ndiv,
                      int* level, unsigned char* start_flag, unsigned* bcnt, hipStream_t s,
                      int bits = 2);
size_t leaf_records_temp(int n);
// one record per leaf start, in Morton order, and their number (*nleaves);
// bcnt as launch_leafdepth left it (scanned in place)
void launch_leaf_records(const unsigned char* start_flag, unsigned* bcnt, int n,
                         const unsigned long long* key, const int* level, int nlmax, int* nleaves,
                         LeafRun* out, hipStream_t s, int bits = 2);
void launch_leaf_records(const unsigned char* start_flag, unsigned* bcnt, int n,
                         const unsigned* key, const int* level, int nlmax, int* nleaves,
                         LeafRun* out, hipStream_t s, int bits = 2);
size_t sort_pairs_u64_temp(int n);
void sort_pairs_u64(void* tmp, size_t tmpb, const unsigned long long* kin, unsigned long long* kout,
                    const unsigned* vin, unsigned* vout, int n, int endbit, hipStream_t s);
size_t sort_pairs_u32_temp(int n);
// shape (tuning table key `sort`): 0..3 the block shape of rocPRIM's onesweep
// kernel, 4 and 5 the library's own stable LSD sort (tiles of 2048 keys, or
// 4096 for digits of up to 10 bits)
void sort_pairs_u32(void* tmp, size_t tmpb, const unsigned* kin, unsigned* kout, const unsigned* vin,
                    unsigned* vout, int n, int endbit, hipStream_t s, int shape = 0);
// ls[r] = first of the n sorted keys >= leafkey[r], r < nleaves; ls[nleaves] = n;
// the same into ls2 for key2 (n2 keys) when key2 is not null
void launch_leafstart_search(const unsigned long long* key, int n, const unsigned long long* key2,
                             int n2, const unsigned long long* leafkey, int nleaves, int* ls,
                             int* ls2, hipStream_t s);
void launch_leafstart_search(const unsigned* key, int n, const unsigned* key2, int n2,
                             const unsigned* leafkey, int nleaves, int* ls, int* ls2,
                             hipStream_t s);
void launch_iota(unsigned* v, int n, hipStream_t s);
// se[b] = (ls[rank_lo[b]], ls[rank_hi[b]]), and se2 from ls2 when se2 is not null
void launch_box_ranges(const int* rank_lo, const int* rank_hi, int nb, const int* ls, int2* se,
                       const int* ls2, int2* se2, hipStream_t s);
void launch_gather(const double* xy, const double* q, int nd, const unsigned* perm, int n,
                   double* xy_s, double* q_s, hipStream_t s);
void launch_zero(double* p, size_t n, hipStream_t s);
void launch_zero3(double* a, size_t na, double* b, size_t nb, double* c, size_t nc, hipStream_t s);
void launch_scatter(const double* v_s, const unsigned* perm, int n, int w, double* v,
                    hipStream_t s, const double* v_s2 = nullptr);
// v[perm[i]*w + c] = nf[i*w + c] + 0 for the points of the listed boxes
// (ranges se): the points that no evaluation writes
void launch_scatter_boxes(const int* boxes, int nbox, const int2* se, const double* nf,
                          const unsigned* perm, int w, double* v, hipStream_t s);

struct FormArgs {
    const int* boxes;
    int nbox;
    const double* cen;      // double2 per box
    const int2* src_se;
    const int* slot;
    const double* xs;       // sorted source xy
    const double* qs;       // sorted charges (n, nd)
    int nd, ind;
    double dsq;
    int p;                  // Hermite order
    const double* h2x;      // complex (p+1, npw): ws (-i ts)^j / sqrt(j!)
    const double* h2xr;     // complex (p+1, npw): ws (-i ts)^j / j!
    const double* rsq;      // 1/sqrt(k)
    int npw;
    double* mp;             // complex per slot (nd, nexp)
    // source segments: tstart[i] = first segment of boxes[i] (exclusive scan
    // of ceil(count/seg)), partial moments of each segment (stride pstride)
    const int* tstart;
    int seg, grid;
    double* partial;
    int pstride;
    const int* segbox;      // list index of the box of every segment (or nullptr)
    const double* ws;       // dipole weights d n, (n, nd, 2) sorted, or nullptr
    // 1: FP32 moments on matrix cores (FP32 operands from FP64 offsets, FP64
    // sums across segments) and FP32 plane-wave output (mp holds floats)
    int prec = 0;
};
void launch_form(const FormArgs& a, hipStream_t s);

// Exclusive scans of per-box segment counts, one list per block (up to 32
// lists per launch): tstart[i] = sum_{j<i} ceil(len(se[boxes[j]]) / seg),
// tstart[n] = total.
struct SegList {
    const int* boxes;
    int n;
    const int2* se;
    int seg;
    int* tstart;
    int* segbox;  // optional: list index of the box of every segment
};
struct SegLists {
    SegList l[32];
    int count;
};
void launch_seg_scan(const SegLists& L, hipStream_t s);

// f32: the expansions (mp, loc) and factor tables (wms, wm2l) hold floats
// (complex interleaved) instead of doubles; the pointers are reinterpreted
void launch_merge(const int* parents, int np, const int4* child, const double* cen,
                  const int2* src_se, const int* slot, const double* wms, int kidx_base,
                  int nd, int nexp, double* mp, hipStream_t s, bool f32 = false);
void launch_m2l(const int* boxes, int nb, const double* cen, const int* ncol, const int* col,
                const int2* src_se, const int2* trg_se, int use_src, int use_trg,
                const int* slot, const double* wm2l, double xmin, int do_list, int nd, int nexp,
                const double* mp, double* loc, hipStream_t s, bool f32 = false);
void launch_split(const int* children, int nc, const int* par, const double* cen,
                  const int* slot, const double* wms, int kidx_base, int nd, int nexp,
                  double* loc, hipStream_t s, bool f32 = false);

struct EvalArgs {
    const int* boxes;
    int nbox;
    const double* cen;
    const int2* se;         // point ranges of the boxes
    const int* slot;
    const double* pts;      // sorted points
    int nd, ind;
    double dsq;
    int ifpgh;
    int nl;                 // Taylor order (Taylor route)
    const double* x2l;      // real (npw, nl+1): ts^j / j!
    const double* ts;       // nodes (PW route)
    int npw;
    const double* loc;      // PW local expansions per slot
    double* taylor;         // scratch: (nbox, (nl+1)^2)
    // outputs in input order, written at perm[i] for sorted point i as the
    // near field's sorted value plus the evaluated one
    double *pot, *grad, *hess;
    const unsigned* perm;
    const double *npot, *ngrad, *nhess;
    // point segments (as FormArgs); tstart == nullptr: one block per box
    const int* tstart;
    int seg, grid;
    const int* segbox;
    int exp32 = 0;          // 1: loc holds FP32 expansions
    int tay32 = 0;          // 1: FP32 Taylor coefficients, evaluated on FP32 matrix cores
    int fused = 0;          // 1: the evaluation forms the Taylor coefficients itself
                            //    (no k_pw2local launch, a.taylor unused)
    int spw = 2;            // fused: consecutive segments per wavefront
    int segmax = 0;         // an upper bound of the segment count (boxes + points / seg + 1)
};
void launch_pw2local(const EvalArgs& a, hipStream_t s);
void launch_taylor_eval(const EvalArgs& a, hipStream_t s);
void launch_pw_eval(const EvalArgs& a, hipStream_t s);

struct DirectArgs {
    const int* units;       // target units (boxes)
    int nunits;
    const int* list_ptr;    // CSR over units: source boxes
    const int* list;
    const int2* src_se;     // source ranges of boxes (sources sorted)
    const int2* tgt_se;     // target ranges of boxes
    const double* xs;
    const double* qs;
    const double* xt;
    int nd;
    double delta, dmax;
    int ifpgh;
    double *pot, *grad, *hess;
    const double* ws;       // dipole weights d n, (n, nd, dim) sorted, or nullptr
};
void launch_direct(const DirectArgs& a, hipStream_t s);

// Taylor expansion of all sources about one centre (large-delta branch)
void launch_formlc(const double* xy, const double* q, const double* ws, int nd, int ind, int n,
                   double cx, double cy, double dsq, int nl, double* part, int nparts,
                   double* out, hipStream_t s);
// dipole weights w(c, ind, i) = rnormal(c, perm i) dipstr(ind, perm i), sorted
void launch_gather_dip(const double* rn, const double* ds, int nd, const unsigned* perm, int n,
                       double* ws, hipStream_t s);
// all-pairs direct sum with a cutoff (fgt2dpart_direct_vec), accumulating;
// q == nullptr: no charges; ws: dipole weights (n, nd, 2) or nullptr
void launch_direct_all(const double* src, const double* q, const double* ws, int nd, int ns,
                       const double* tg, int nt, double delta, double dmax, int ifpgh,
                       double* pot, double* grad, double* hess, hipStream_t s);

// ---------------------------------------------------------------------------
// 3D kernels (kernels3d.hip).  Points (3,n), expansions complex
// (npw, npw, npw/2) per slot, ascending nodes.
// ---------------------------------------------------------------------------
void launch_bbox3(const double* a, int na, const double* b, int nb, double* part, int nblk,
                  double* out, hipStream_t s);  // out: xmin,xmax,ymin,ymax,zmin,zmax,bad
void launch_keys3d(const double* p, int n, const double* c0, double bs0, int nlmax,
                   unsigned long long* key, unsigned* idx, hipStream_t s);
void launch_keys3d(const double* p, int n, const double* c0, double bs0, int nlmax,
                   unsigned* key, unsigned* idx, hipStream_t s);
void launch_gather3(const double* p, const double* q, int nd, const unsigned* perm, int n,
                    double* ps, double* qs, hipStream_t s);

struct Form3Args {
    const int* boxes;       // leaves to form
    int nbox;
    const double* cen;      // 3 per box
    const int2* src_se;
    const int* slot;
    const double* xs;       // sorted sources (3,n)
    const double* qs;       // sorted charges (nd,n)
    int nd, ind;
    double dsq;
    int p;                  // Hermite order
    int npw;
    const double* coef;     // complex C(j,a) = ws_a (-i ts_a)^j / j! at a + cstride*j
    int cstride;
    const double *ws, *ts;  // nodes (ascending)
    double* mp;             // per slot (nd, nexp) complex
    double* scratch;        // per block: m^3 + 2 m^2 (npw/2) doubles
    long sstride;
    int grid;
    int thresh;             // Hermite route for boxes with more than thresh
                            // sources, direct plane-wave route for the others
    const double* dipw = nullptr;  // dipole weights d n, (n, nd, 3) sorted, or nullptr
};
void launch_form3_hermite(const Form3Args& a, hipStream_t s);
long form3_mfma_scratch(int m, int npw);  // per-block scratch of the matrix-core route
void launch_form3_pw(const Form3Args& a, hipStream_t s);

void launch_merge3(const int* parents, int np, const int* child8, const double* cen,
                   const int2* src_se, const int* slot, const double* wms, int kbase, int nd,
                   long nexp, double* mp, hipStream_t s);
void launch_m2l3(const int* boxes, int nb, const double* cen, const int* ncol, const int* col27,
                 const int* nch, const int2* src_se, const int2* trg_se, int use_src, int use_trg,
                 const int* slot_mp, const int* slot_loc, const double* wm1, int npw, double xmin,
                 int do_list, int nd, long nexp, const double* mp, double* loc, hipStream_t s);
void launch_split3(const int* children, int nc, const int* par, const double* cen,
                   const int* slot, const double* wms, int kbase, int nd, long nexp, double* loc,
                   const int2* src_se, const int2* trg_se, int use_src, int use_trg,
                   hipStream_t s);

struct Eval3Args {
    const int* boxes;
    int nbox;
    const double* cen;
    const int2* se;         // point ranges of the boxes
    const int* slot;
    const double* pts;      // sorted points (3,n)
    int nd, ind;
    double dsq;
    int ifpgh;
    int nl;                 // Taylor order
    const double* dcoef;    // complex D(j,a) = (i ts_a)^j / j! at a + cstride*j
    int cstride;
    const double* ts;       // nodes (ascending)
    int npw;
    const double* loc;      // PW local expansions per slot
    double* taylor;         // (nbox, (nl+1)^3)
    double *pot, *grad, *hess;
    // segments of the point ranges (evaluation kernels): box i of the list
    // owns segments tstart[i] .. tstart[i+1]-1; tstart[0] need not be 0, so a
    // sub-list of a scanned list is passed by offsetting boxes and tstart
    const int* tstart;
    int seg, grid;
    // route: boxes whose source + target count is below thresh are evaluated
    // from the plane waves directly, the others through Taylor
    const int2 *rs, *rt;
    int thresh;
    int use_src, use_trg;   // which points are evaluated (PW -> Taylor skips boxes with none)
    // Taylor coefficient layout: 0 L[j3 + m (j2 + m j1)]; 1 (m <= 32) the
    // transposed Lt[j2 + MP (j3 + m j1)], MP = m rounded up to 16, zero columns
    int tlayout;
    // fused split (PW -> Taylor only): the leaf's expansion is its parent's
    // (slot of par[box]) times the octant shift wms[kbase + octant]
    int fused;
    const int* par;
    const double* wms;
    int kbase;
};
bool pw2local3_fusable(int m, int npw);
// doubles per box of the Taylor buffer
__host__ __device__ inline long taylor3_stride(int m, int tlayout)
{
    return tlayout ? (long)(16 * ((m + 15) / 16)) * m * m : (long)m * m * m;
}
void launch_pw2local3(const Eval3Args& a, hipStream_t s);
void launch_taylor_eval3(const Eval3Args& a, hipStream_t s);
void launch_pw_eval3(const Eval3Args& a, hipStream_t s);
void launch_direct3(const DirectArgs& a, hipStream_t s);  // points (3,n)
void launch_direct_all3(const double* src, const double* q, const double* ws, int nd, int ns,
                        const double* tg, int nt, double delta, double dmax, int ifpgh,
                        double* pot, double* grad, double* hess, hipStream_t s);
void launch_formlc3(const double* p, const double* q, const double* ws, int nd, int ind, int n,
                    const double* c, double dsq, int nl, double* part, int nparts, double* out,
                    hipStream_t s);
// dipole weights w(c, ind, i) = rnormal(c, perm i) dipstr(ind, perm i), sorted, (n, nd, 3)
void launch_gather_dip3(const double* rn, const double* ds, int nd, const unsigned* perm, int n,
                        double* ws, hipStream_t s);

}  // namespace fgtpw
