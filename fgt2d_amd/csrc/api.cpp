// C-ABI and host runtime of the 2D plane-wave FGT.
//
// One call = one pass of fgt2d (2d/planewave/src/fgt2d.f:19-462) on the GPU:
//   bbox  -> term counts on the host (boxsize0, nlmax: fgt2d.f:124-186)
//   keys  -> radix sort -> leaf levels -> list of leaves  (device)
//   tree  -> level restriction, lists, tables, work lists  (host, O(boxes))
//   assign points to leaves, sort, gather                  (device)
//   form -> merge -> m2l -> split -> eval -> near field -> scatter (device)
// The host reads back three small things per call (bbox, leaf count, leaf
// list); everything proportional to the number of points stays in HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <climits>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>
#include <string>
#include <vector>

#include <dlfcn.h>
#include <unistd.h>

#include "fgtpw.h"
#include "fgtpw_impl.hpp"

namespace fgtpw {

namespace {
thread_local std::string t_err;
thread_local std::vector<int> t_stats;
thread_local std::vector<double> t_stage_ms;
thread_local bool t_mixed = false;  // the last 2D transform took the mixed-precision route
bool g_timing = false;
// timing mode 2, or FGTPW_SERIAL=1 for every call: the side streams' work
// runs on the call's stream (no two kernels of a call overlap)
const bool g_serial_env = [] {
    const char* e = std::getenv("FGTPW_SERIAL");
    return e && e[0] == '1';
}();
bool g_serial = g_serial_env;
// 0: FP64 everywhere; 1: the mixed-precision route for 2D transforms with
// eps >= kMixedEpsMin (FP32 matrix-core moments and Taylor evaluation, FP32
// plane-wave expansions and translation factors)
int g_precision = 0;
constexpr double kMixedEpsMin = 1e-6;
// 2D evaluation of Taylor orders below 32: PW -> Taylor as its own
// matrix-core kernel, one wavefront per leaf (default), or inside the
// evaluation waves (FGTPW_FUSE_PW2LOCAL=1; measured slower: the conversion's
// latency is not hidden at the evaluation kernel's occupancy)
const bool g_fuse_pw2local = [] {
    const char* e = std::getenv("FGTPW_FUSE_PW2LOCAL");
    return e && e[0] == '1';
}();
// Tile shapes of the 2D segmented kernels, per Taylor / Hermite order tile
// count TM (orders up to 16 TM): points per formation segment (segf),
// points per evaluation segment (sege, at most 256: the wave kernels' LDS
// stage) and consecutive segments per wavefront of the fused evaluation
// (spw).  Defaults below; a table written by the autotuner
// (python -m fgt2d_amd.autotune) in tuning.txt next to the library replaces
// them, and FGTPW_TUNE="key=value,..." overrides both (keys segf1..segf4,
// sege1..sege4, spw1..spw4, sort).
struct Tuning {
    int segf[4] = {256, 256, 128, 128};
    int sege[4] = {256, 256, 256, 256};
    int spw[4] = {2, 2, 2, 2};
    int sort = 4;  // 32-bit key sort: 4, 5 the library's LSD sort (tiles of 2048, 4096 keys),
                   // 0..3 a rocPRIM onesweep block shape
    std::string source = "defaults";
    bool set(const std::string& k, int v)
    {
        if (k == "sort") {
            if (v < 0 || v > 5) return false;
            sort = v;
            return true;
        }
        int* arr = k.rfind("segf", 0) == 0 ? segf : k.rfind("sege", 0) == 0 ? sege
                   : k.rfind("spw", 0) == 0 ? spw : nullptr;
        if (!arr || k.size() < 4) return false;
        const int t = k.back() - '1';
        if (t < 0 || t > 3) return false;
        const bool ok = arr == spw ? (v >= 1 && v <= 16)
                                   : ((v == 64 || v == 128 || v == 256) || (arr == segf && v == 512));
        if (ok) arr[t] = v;
        return ok;
    }
    void parse(const std::string& text, char sep)
    {
        size_t i = 0;
        while (i < text.size()) {
            size_t j = text.find(sep, i);
            if (j == std::string::npos) j = text.size();
            std::string line = text.substr(i, j - i);
            i = j + 1;
            const size_t h = line.find('#');
            if (h != std::string::npos) line.resize(h);
            const size_t eq = line.find('=');
            if (eq == std::string::npos) continue;
            auto trim = [](std::string x) {
                const size_t a = x.find_first_not_of(" \t\r"), b = x.find_last_not_of(" \t\r");
                return a == std::string::npos ? std::string() : x.substr(a, b - a + 1);
            };
            set(trim(line.substr(0, eq)), std::atoi(trim(line.substr(eq + 1)).c_str()));
        }
    }
};

const Tuning& tuning()
{
    static const Tuning t = [] {
        Tuning u;
        Dl_info info;
        if (dladdr((const void*)&tuning, &info) && info.dli_fname) {
            std::string path(info.dli_fname);
            const size_t slash = path.rfind('/');
            path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tuning.txt";
            if (FILE* f = std::fopen(path.c_str(), "r")) {
                std::string text;
                char buf[4096];
                size_t n;
                while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
                std::fclose(f);
                u.parse(text, '\n');
                u.source = path;
            }
        }
        if (const char* e = std::getenv("FGTPW_TUNE")) {
            u.parse(e, ',');
            u.source += "+FGTPW_TUNE";
        }
        return u;
    }();
    return t;
}
std::mutex g_mutex;  // one transform at a time per process (shared workspace)
int g_device = -1;

const char* const kStages[] = {"bbox",  "keys_sort", "leafdepth", "host_tree", "assign",
                               "ranges", "form",     "merge",     "m2l",       "split",
                               "eval",  "direct",    "scatter",   "host_build", "gather"};
// host_build: the box tree itself, part of host_tree (host clock); gather:
// the sources' points and charges into Morton order (side stream)
enum Stage { S_BBOX, S_KEYS, S_LEAF, S_HOST, S_ASSIGN, S_RANGES, S_FORM, S_MERGE, S_M2L, S_SPLIT,
             S_EVAL, S_DIRECT, S_SCATTER, S_HOST_BUILD, S_GATHER, S_N };
}  // namespace

thread_local KernelClock* t_kclock = nullptr;

void fail(int code, const std::string& msg) { throw Error{code, msg}; }

void check_hip(hipError_t e, const char* what)
{
    if (e == hipSuccess) return;
    if (e == hipErrorOutOfMemory) fail(FGTPW_ENOMEM, std::string(what) + ": out of device memory");
    fail(FGTPW_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

// ---------------------------------------------------------------------------
// workspace: named device buffers that only grow, plus pinned staging
// ---------------------------------------------------------------------------
struct Tables2D {
    bool valid = false;
    int npw = 0, ntermmax = 0, nmax = 0;
    double pmax = 0, xmin2 = 0, xm = 0;
    const double *ts = nullptr, *h2x = nullptr, *h2xr = nullptr, *x2l = nullptr, *rsq = nullptr,
                 *wms = nullptr, *wm2l = nullptr;
    const float *wms32 = nullptr, *wm2l32 = nullptr;  // the same factors rounded to FP32
};

struct Tables3D {
    bool valid = false;
    int npw = 0, ntermmax = 0, nmax = 0;
    double pmax = 0, xmin2 = 0, xm = 0;
    const double *ws = nullptr, *ts = nullptr, *coef = nullptr, *dcoef = nullptr, *wms = nullptr,
                 *wm1 = nullptr;  // wm1: complex (3, npw) m2l factors per axis
};

// One thread that runs a workspace's host-side job beside the launches of
// the same call (the near-field lists), started on first use.  wait()
// returns once the posted job has run and rethrows what it threw.
class HostHelper {
public:
    ~HostHelper()
    {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        if (th_ && pid_ == getpid()) {
            th_->join();
            delete th_;
        }
    }
    void post(std::function<void()> f)
    {
        // a forked child has none of its parent's threads: it starts its own
        // (the parent's std::thread object is left alone)
        if (!th_ || pid_ != getpid()) {
            th_ = new std::thread([this] { loop(); });
            pid_ = getpid();
        }
        {
            std::lock_guard<std::mutex> lk(m_);
            job_ = std::move(f);
            pending_ = true;
            err_ = nullptr;
        }
        cv_.notify_all();
    }
    // the posted job has ended (wait() then returns at once)
    bool done()
    {
        std::lock_guard<std::mutex> lk(m_);
        return !pending_;
    }
    void wait()
    {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return !pending_; });
        if (err_) {
            std::exception_ptr e = err_;
            err_ = nullptr;
            std::rethrow_exception(e);
        }
    }

private:
    void loop()
    {
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait(lk, [&] { return pending_ || stop_; });
            if (!pending_) return;  // stopping with nothing posted
            std::function<void()> f = std::move(job_);
            lk.unlock();
            std::exception_ptr e;
            try {
                f();
            } catch (...) {
                e = std::current_exception();
            }
            lk.lock();
            err_ = e;
            pending_ = false;
            cv_.notify_all();
        }
    }
    std::thread* th_ = nullptr;
    pid_t pid_ = 0;
    std::mutex m_;
    std::condition_variable cv_;
    std::function<void()> job_;
    bool pending_ = false, stop_ = false;
    std::exception_ptr err_;
};

// Growable host storage of a call's plan: pinned, so that its upload is one
// DMA from where it was packed, or pageable for the no-allocation entry
// point.  Grows only while a plan is packed, when the last call's upload
// from it has completed (a read-back on the same stream came since).
struct PlanStore {
    std::vector<char> pageable;
    char* pin = nullptr;
    size_t cap = 0;
    bool pinned = true;
    char* reserve(size_t n, size_t keep)
    {
        if (!pinned) {
            if (pageable.size() < n) pageable.resize(std::max(n, 2 * pageable.size()));
            return pageable.data();
        }
        if (cap < n) {
            const size_t want = std::max<size_t>(n + n / 2, 1 << 16);
            char* q = nullptr;
            HIPCHECK(hipHostMalloc((void**)&q, want, hipHostMallocDefault));
            if (keep) std::memcpy(q, pin, keep);
            if (pin) HIPCHECK(hipHostFree(pin));
            pin = q;
            cap = want;
        }
        return pin;
    }
    char* data() { return pinned ? pin : pageable.data(); }
    void release()
    {
        if (pin) (void)hipHostFree(pin);
        pin = nullptr;
        cap = 0;
    }
};

// Waits for a helper's job on every way out of a scope whose locals the job
// reads (h = nullptr once it has been waited for)
struct HelperJoin {
    HostHelper* h;
    ~HelperJoin()
    {
        if (h) try {
                h->wait();
            } catch (...) {
            }
    }
};

struct Workspace {
    struct Buf {
        void* p = nullptr;
        size_t cap = 0;
    };
    std::map<std::string, Buf> dev;
    void* pin = nullptr;
    size_t pin_cap = 0;
    void* pin2 = nullptr;
    size_t pin2_cap = 0;
    hipStream_t own = nullptr;
    // side streams (default priority): coarse levels of formation and
    // evaluation and the near field overlap the finest level's work
    hipStream_t aux[2] = {};
    // pinned staging ring of the host-array entry points: two halves, each
    // with the event of the copy that last used it
    char* ring = nullptr;
    hipEvent_t ring_ev[2] = {};
    std::vector<hipEvent_t> order_ev;  // stream-ordering events, reused from call to call
    size_t n_order = 0;
    std::vector<hipEvent_t> sev;  // stage timing events, grown on demand
    std::vector<hipEvent_t> kev;  // per-kernel timing events, grown on demand
    // host scratch of the plan, reused from call to call
    std::vector<int> h_units, h_uptr, h_ulist, h_nl1, h_l1, h_sp, h_sl, h_fill, h_slot, h_slot2;
    PlanStore h_plan, h_plan_nf;
    std::vector<unsigned long long> h_lkey64;  // smallest key of every leaf, rank order
    std::vector<unsigned> h_lkey32;
    size_t leaf_guess = 0;  // leaf records of the last call (size of the first read-back)
    std::unique_ptr<HostHelper> helper;  // host work beside a call's launches
    std::vector<std::vector<int>> h_form, h_merge, h_split, h_eval;  // per-level work lists
    std::vector<int> h_m2l;
    std::vector<int> h_resid;  // leaves that no evaluation covers
    std::vector<char> h_iflocal;
    // the finest level formed before the tree is built (2D): its boxes' leaf
    // records, centres and ranges, and the plan they travel in
    std::vector<int> h_erec, h_eiota;
    std::vector<double> h_ecen;
    std::vector<int2> h_ese;
    PlanStore h_plan_early;
    int last_nslot = 0;
    Tables2D tab;
    Tables3D tab3;
    // the tree of the last transform, kept for inspection
    Tree2D last;
    Tree3D last3;
    int last_dim = 2;
    int last_ns = 0, last_nt = 0;
    bool last_valid = false;

    // Arena mode (the no-allocation entry point): get() hands out pieces of
    // a fixed device region, reset at the start of every call, and the host
    // staging areas are pageable.  Nothing is allocated on the device.
    char* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    std::map<std::string, std::pair<size_t, size_t>> arena_map;  // name -> offset, bytes
    std::vector<char> stage, stage2;                              // pageable host staging
    void arena_reset()
    {
        arena_used = 0;
        arena_map.clear();
        tab.valid = tab3.valid = false;  // the tables live in the arena too
    }

    template <class T>
    T* get(const std::string& name, size_t count)
    {
        size_t bytes = std::max<size_t>(count * sizeof(T), 256);
        if (arena) {
            auto it = arena_map.find(name);
            if (it != arena_map.end() && it->second.second >= bytes)
                return (T*)(arena + it->second.first);
            const size_t off = (arena_used + 255) & ~size_t(255);
            if (off + bytes > arena_cap)
                fail(FGTPW_ENOMEM, "the tree and expansions of this transform need more than the " +
                                       std::to_string(arena_cap >> 20) +
                                       " MiB of the no-allocation scratch (buffer '" + name + "')");
            arena_used = off + bytes;
            arena_map[name] = {off, bytes};
            return (T*)(arena + off);
        }
        Buf& b = dev[name];
        if (b.cap < bytes) {
            if (b.p) HIPCHECK(hipFree(b.p));
            b.p = nullptr;
            b.cap = 0;
            size_t want = bytes + std::min<size_t>(bytes / 8, (size_t)256 << 20);
            HIPCHECK(hipMalloc(&b.p, want));
            b.cap = want;
        }
        return (T*)b.p;
    }
    size_t held(const std::string& name) const
    {
        auto it = dev.find(name);
        return it == dev.end() ? 0 : it->second.cap;
    }
    // an event after the work queued on `from` so far (a record costs the
    // host about 1 us; a wait on it, later, about 10 us)
    struct Mark {
        hipEvent_t e = nullptr;
        hipStream_t from = nullptr;
    };
    Mark mark(hipStream_t from)
    {
        if (n_order == order_ev.size()) {
            hipEvent_t e;
            HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            order_ev.push_back(e);
        }
        Mark m{order_ev[n_order++], from};
        HIPCHECK(hipEventRecord(m.e, from));
        return m;
    }
    // work queued on `to` after this call waits for the marked work
    void wait(hipStream_t to, const Mark& m)
    {
        if (m.e && m.from != to) HIPCHECK(hipStreamWaitEvent(to, m.e, 0));
    }
    // work queued on `to` after this call waits for the work queued on
    // `from` before it
    void order(hipStream_t from, hipStream_t to)
    {
        if (from == to) return;  // one stream: program order
        if (n_order == order_ev.size()) {
            hipEvent_t e;
            HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            order_ev.push_back(e);
        }
        hipEvent_t e = order_ev[n_order++];
        HIPCHECK(hipEventRecord(e, from));
        HIPCHECK(hipStreamWaitEvent(to, e, 0));
    }
    // frees every device and pinned buffer, the events and the streams
    void release()
    {
        for (auto& kv : dev)
            if (kv.second.p) (void)hipFree(kv.second.p);
        dev.clear();
        if (pin) (void)hipHostFree(pin);
        pin = nullptr;
        pin_cap = 0;
        h_plan.release();
        h_plan_nf.release();
        h_plan_early.release();
        if (pin2) (void)hipHostFree(pin2);
        pin2 = nullptr;
        pin2_cap = 0;
        if (ring) (void)hipHostFree(ring);
        ring = nullptr;
        for (auto& e : ring_ev)
            if (e) (void)hipEventDestroy(e);
        ring_ev[0] = ring_ev[1] = nullptr;
        for (auto& e : sev)
            if (e) (void)hipEventDestroy(e);
        sev.clear();
        for (auto& e : kev)
            if (e) (void)hipEventDestroy(e);
        kev.clear();
        for (auto& e : order_ev) (void)hipEventDestroy(e);
        order_ev.clear();
        for (auto& a : aux)
            if (a && a != own) (void)hipStreamDestroy(a);
        aux[0] = aux[1] = nullptr;
        if (own) (void)hipStreamDestroy(own);
        own = nullptr;
        tab.valid = tab3.valid = false;
    }
    void* pinned(size_t bytes)
    {
        if (arena) return grow_stage(stage, bytes);
        return grow_pinned(pin, pin_cap, bytes);
    }
    // a second staging area, for an upload queued while the first one's may
    // still be in flight
    void* pinned2(size_t bytes)
    {
        if (arena) return grow_stage(stage2, bytes);
        return grow_pinned(pin2, pin2_cap, bytes);
    }
    // pageable staging: a copy from pageable memory has read its source
    // when hipMemcpyAsync returns, and reads into it are followed by a
    // stream synchronisation
    static void* grow_stage(std::vector<char>& v, size_t bytes)
    {
        if (v.size() < bytes) v.resize(bytes);
        return v.data();
    }
    static void* grow_pinned(void*& p, size_t& cap, size_t bytes)
    {
        if (cap < bytes) {
            if (p) HIPCHECK(hipHostFree(p));
            p = nullptr;
            cap = 0;
            size_t want = std::max<size_t>(bytes + bytes / 4, 1 << 16);
            HIPCHECK(hipHostMalloc(&p, want, hipHostMallocDefault));
            cap = want;
        }
        return p;
    }
};

// side_streams = false: the workspace runs everything on its own stream
// (the server's slots: the device has few hardware queues, and a slot whose
// streams spread over several of them would put its long kernels ahead of
// the other slots' work)
Workspace* new_workspace(bool side_streams = true)
{
    Workspace* w = new Workspace();
    HIPCHECK(hipStreamCreateWithFlags(&w->own, hipStreamNonBlocking));
    for (auto& a : w->aux) {
        if (side_streams)
            HIPCHECK(hipStreamCreateWithFlags(&a, hipStreamNonBlocking));
        else
            a = w->own;
    }
    return w;
}

Workspace& workspace()
{
    static std::map<int, Workspace*> all;
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    Workspace*& w = all[dev];
    if (!w) w = new_workspace();
    return *w;
}

const char* const kKernels[KID_N] = {"form_moments", "form_h2pw", "form_valu", "merge",
                                     "m2l",          "split",     "pw2local",  "taylor_eval",
                                     "pw_eval",      "direct",    "gather",    "scatter",
                                     "sort"};
struct KLaunch {
    int id, level;
    double ms;
    long long units;
};
thread_local std::vector<KLaunch> t_klaunch;

// Events around each instrumented kernel launch of one call.  The unit count
// of a launch (points, boxes) is resolved after the call from the point
// ranges of its box list: side 0 sources, 1 targets, 2 the list length,
// -1 a count given directly.
struct LaunchClock : KernelClock {
    struct Rec {
        int id, level, side;
        const std::vector<int>* boxes;
        int r0, r1;
        long long units;
        hipEvent_t e0, e1;
    };
    std::vector<hipEvent_t>& pool;
    std::vector<Rec> recs;
    int level = -1, side = -1;
    const std::vector<int>* boxes = nullptr;
    int r0 = 0, r1 = -1;  // the launch covers (*boxes)[r0, r1), r1 < 0: all
    long long units = 0;
    explicit LaunchClock(std::vector<hipEvent_t>& p) : pool(p) {}
    void ctx(int l, const std::vector<int>* b, int sd, long long u = 0, int b0 = 0, int b1 = -1)
    {
        level = l, boxes = b, side = sd, units = u, r0 = b0, r1 = b1;
    }
    hipEvent_t ev(size_t i)
    {
        while (pool.size() <= i) {
            hipEvent_t e;
            HIPCHECK(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[i];
    }
    void mark(int id, bool end, hipStream_t st) override
    {
        if (!end) {
            Rec r{id, level, side, boxes, r0, r1, units, ev(2 * recs.size()), ev(2 * recs.size() + 1)};
            HIPCHECK(hipEventRecord(r.e0, st));
            recs.push_back(r);
        } else {
            HIPCHECK(hipEventRecord(recs.back().e1, st));
        }
    }
    void pair(int id, hipEvent_t* e0, hipEvent_t* e1) override
    {
        Rec r{id, level, side, boxes, r0, r1, units, ev(2 * recs.size()), ev(2 * recs.size() + 1)};
        recs.push_back(r);
        *e0 = r.e0;
        *e1 = r.e1;
    }
};

struct Timer {
    Workspace& W;
    hipStream_t s;
    bool on;
    double host_ms = 0, build_ms = 0;
    // stage spans: a stage may be entered several times per call (chunked
    // work); its time is the sum of its spans
    struct Span { int k; hipEvent_t e0, e1; };
    std::vector<Span> spans;
    int open[S_N];
    LaunchClock clock;
    const int2 *src_se = nullptr, *trg_se = nullptr;
    int nb = 0, nd = 1;
    Timer(Workspace& w, hipStream_t st) : W(w), s(st), on(g_timing), clock(w.kev)
    {
        t_klaunch.clear();
        for (int& o : open) o = -1;
        if (on) t_kclock = &clock;
    }
    hipEvent_t sev(size_t i)
    {
        while (W.sev.size() <= i) {
            hipEvent_t e;
            HIPCHECK(hipEventCreate(&e));
            W.sev.push_back(e);
        }
        return W.sev[i];
    }
    ~Timer() { t_kclock = nullptr; }
    void ctx(int l, const std::vector<int>* b, int sd, long long u = 0, int b0 = 0, int b1 = -1)
    {
        if (on) clock.ctx(l, b, sd, u, b0, b1);
    }
    // events around a library call's launches (the radix sorts), in context
    // (level, units)
    void bracket(int id, bool end, int l, long long u)
    {
        if (!on) return;
        if (!end) clock.ctx(l, nullptr, -1, u);
        clock.mark(id, end, s);
    }
    void begin(int k, hipStream_t st = nullptr)
    {
        if (!on) return;
        const size_t n = spans.size();
        spans.push_back({k, sev(2 * n), sev(2 * n + 1)});
        HIPCHECK(hipEventRecord(spans[n].e0, st ? st : s));
        open[k] = (int)n;
    }
    void end(int k, hipStream_t st = nullptr)
    {
        if (!on || open[k] < 0) return;
        HIPCHECK(hipEventRecord(spans[open[k]].e1, st ? st : s));
        open[k] = -1;
    }
    void finish()
    {
        t_stage_ms.assign(S_N, 0.0);
        if (!on) return;
        HIPCHECK(hipStreamSynchronize(s));
        for (const Span& sp : spans) {
            float ms = 0;
            HIPCHECK(hipEventElapsedTime(&ms, sp.e0, sp.e1));
            t_stage_ms[sp.k] += ms;
        }
        t_stage_ms[S_HOST] = host_ms;
        t_stage_ms[S_HOST_BUILD] = build_ms;
        std::vector<int2> hs, ht;
        if (nb > 0) {
            hs.resize(nb);
            ht.resize(nb);
            HIPCHECK(hipMemcpy(hs.data(), src_se, sizeof(int2) * nb, hipMemcpyDeviceToHost));
            HIPCHECK(hipMemcpy(ht.data(), trg_se, sizeof(int2) * nb, hipMemcpyDeviceToHost));
        }
        for (const auto& r : clock.recs) {
            float ms = 0;
            HIPCHECK(hipEventElapsedTime(&ms, r.e0, r.e1));
            long long u = r.units;
            const int b1 = r.boxes ? (r.r1 < 0 ? (int)r.boxes->size() : r.r1) : 0;
            if (r.boxes && r.side == 2) u = b1 - r.r0;
            if (r.boxes && (r.side == 0 || r.side == 1) && nb > 0) {
                const auto& se = r.side == 0 ? hs : ht;
                u = 0;
                for (int i = r.r0; i < b1; ++i) u += se[(*r.boxes)[i]].y - se[(*r.boxes)[i]].x;
            }
            if ((r.side == 3 || r.side == 4) && nb > 0) {
                // near field: source-target pairs of the unit lists, times nd
                // (side 3: the targets, 4: the sources as targets)
                const auto& te = r.side == 3 ? ht : hs;
                u = 0;
                for (size_t k = 0; k < W.h_units.size(); ++k) {
                    long long ns_ = 0;
                    for (int li = W.h_uptr[k]; li < W.h_uptr[k + 1]; ++li)
                        ns_ += hs[W.h_ulist[li]].y - hs[W.h_ulist[li]].x;
                    u += ns_ * (te[W.h_units[k]].y - te[W.h_units[k]].x);
                }
                u *= nd;
            }
            t_klaunch.push_back(KLaunch{r.id, r.level, ms, u});
        }
    }
};

// one transform, device arrays; chg == nullptr: no charges, rn == nullptr:
// no dipoles (rn (2 or 3, ns) normals, ds (nd, ns) strengths)
struct Call {
    int nd;
    double delta, eps;
    int ns;
    const double* src;
    const double* chg;
    int ifpgh;
    double *pot, *grad, *hess;
    int nt;
    const double* targ;
    int ifpghtarg;
    double *pottarg, *gradtarg, *hesstarg;
    const double* rn = nullptr;
    const double* ds = nullptr;
};

void copy_to_host(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// packed host->device upload of the per-call plan
// ---------------------------------------------------------------------------
struct Pack {
    PlanStore own;    // a pageable store of its own (the tables)
    PlanStore& st;    // storage, kept from call to call (no refill)
    size_t used = 0;  // bytes of this plan
    Pack() : st(own) { own.pinned = false; }
    Pack(PlanStore& storage, bool pinned) : st(storage) { st.pinned = pinned; }
    template <class T>
    size_t add(const T* p, size_t n)
    {
        const size_t off = (used + 255) & ~size_t(255);
        const size_t end = off + std::max<size_t>(n * sizeof(T), 1);
        char* b = st.reserve(end, used);
        if (n) std::memcpy(b + off, p, n * sizeof(T));
        used = end;
        return off;
    }
    const char* data() { return st.data(); }
    size_t size() const { return used; }
    template <class T>
    size_t add(const std::vector<T>& v)
    {
        return add(v.data(), v.size());
    }
};

// Smallest full-depth Morton key of every leaf of the final tree, in rank
// (Morton) order: the leaf's prefix shifted past the levels below it.  The
// GPU finds each leaf's first point by searching the sorted keys for it.
template <int D>
size_t pack_leaf_keys(Pack& pk, Workspace& W, const TreeD<D>& T, int nlmax, bool k32)
{
    const size_t nl = T.leaf_of_rank.size();
    if (k32) {
        std::vector<unsigned>& k = W.h_lkey32;
        k.resize(nl);
        for (size_t r = 0; r < nl; ++r) {
            const int b = T.leaf_of_rank[r];
            k[r] = (unsigned)(T.pre[b] << (D * (nlmax - T.lev[b])));
        }
        return pk.add(k);
    }
    std::vector<unsigned long long>& k = W.h_lkey64;
    k.resize(nl);
    for (size_t r = 0; r < nl; ++r) {
        const int b = T.leaf_of_rank[r];
        k[r] = (unsigned long long)T.pre[b] << (D * (nlmax - T.lev[b]));
    }
    return pk.add(k);
}

// ---------------------------------------------------------------------------
// plane-wave tables: nodes, Hermite->PW and PW->Taylor 1-D matrices, and the
// diagonal translation factors.  They depend only on (npw, pmax, orders,
// shift lengths), so they stay resident and are rebuilt only when those
// change.
// ---------------------------------------------------------------------------
const Tables2D& tables2d(Workspace& W, int npw, double pmax, int ntermmax, int nmax, double xmin2,
                         double xm, hipStream_t s)
{
    Tables2D& t = W.tab;
    if (t.valid && t.npw == npw && t.pmax == pmax && t.ntermmax == ntermmax && t.nmax == nmax &&
        t.xmin2 == xmin2 && t.xm == xm)
        return t;
    const int half = npw / 2, nexp = half * npw, mt = ntermmax + 1;
    std::vector<double> ws(npw), ts(npw);
    planewave_nodes(pmax, npw, ws.data(), ts.data());
    // h2x(j,a) = ws_a (-i ts_a)^j / sqrt(j!) (g2dh2xmat, g2drouts.f:1601-1640);
    // x2l(a,j) = ts_a^j / j! (g2drx2lmat, :1731-1769)
    std::vector<double> h2x(2 * (size_t)mt * npw), h2xr(2 * (size_t)mt * npw), x2l((size_t)npw * mt),
        rsq(mt + 1);
    rsq[0] = 1;
    for (int k = 1; k <= mt; ++k) rsq[k] = 1.0 / std::sqrt((double)k);
    double fh = 1.0, fl = 1.0;
    for (int j = 0; j < mt; ++j) {
        if (j > 0) { fh = fh / std::sqrt((double)j); fl = fl * j; }
        for (int a = 0; a < npw; ++a) {
            double r = ws[a] * std::pow(-ts[a], j) * fh;
            double re = 0, im = 0;  // times i^j
            switch (j & 3) {
            case 0: re = r; break;
            case 1: im = r; break;
            case 2: re = -r; break;
            default: im = -r; break;
            }
            h2x[2 * ((size_t)j * npw + a)] = re;
            h2x[2 * ((size_t)j * npw + a) + 1] = im;
            // the same with 1/j! in place of 1/sqrt(j!): for raw moments
            h2xr[2 * ((size_t)j * npw + a)] = re * fh;
            h2xr[2 * ((size_t)j * npw + a) + 1] = im * fh;
            x2l[a + (size_t)npw * j] = std::pow(ts[a], j) / fl;
        }
    }
    // merge/split factors e^{i ts xmin2 2^(k-1)} by repeated squaring
    // (merge_split_pw_matrices, g2drouts.f:2294-2350); 4 sign patterns
    std::vector<double> wms(2 * (size_t)nexp * 4 * std::max(nmax, 1));
    std::vector<double> wwr((size_t)npw * (nmax + 1)), wwi((size_t)npw * (nmax + 1));
    for (int j = 0; j < npw; ++j) {
        double zr = std::cos(ts[j] * xmin2), zi = std::sin(ts[j] * xmin2);
        for (int k = 1; k <= nmax; ++k) {
            wwr[j + (size_t)npw * k] = zr;
            wwi[j + (size_t)npw * k] = zi;
            double t2 = zr * zr - zi * zi;
            zi = zr * zi + zi * zr;
            zr = t2;
        }
    }
    for (int k = 1; k <= nmax; ++k)
        for (int qd = 0; qd < 4; ++qd)
            for (int b = 0; b < npw; ++b)
                for (int a = 0; a < half; ++a) {
                    double xr = wwr[a + (size_t)npw * k], xi = wwi[a + (size_t)npw * k];
                    double yr = wwr[b + (size_t)npw * k], yi = wwi[b + (size_t)npw * k];
                    if (qd >= 2) xi = -xi;  // x shift negative
                    if (qd & 1) yi = -yi;   // y shift negative
                    size_t o = 2 * ((size_t)(4 * (k - 1) + qd) * nexp + a + (size_t)half * b);
                    wms[o] = xr * yr - xi * yi;
                    wms[o + 1] = xr * yi + xi * yr;
                }
    // m2l factors for shifts of -1, 0, 1 boxes (pw_translation_matrices,
    // g2drouts.f:2235-2289)
    std::vector<double> wm2l(2 * (size_t)nexp * 9);
    for (int ky = -1; ky <= 1; ++ky)
        for (int kx = -1; kx <= 1; ++kx)
            for (int b = 0; b < npw; ++b)
                for (int a = 0; a < half; ++a) {
                    double xr = 1, xi = 0, yr = 1, yi = 0;
                    if (kx) { xr = std::cos(ts[a] * xm); xi = kx * std::sin(ts[a] * xm); }
                    if (ky) { yr = std::cos(ts[b] * xm); yi = ky * std::sin(ts[b] * xm); }
                    size_t o = 2 * ((size_t)((kx + 1) + 3 * (ky + 1)) * nexp + a + (size_t)half * b);
                    wm2l[o] = xr * yr - xi * yi;
                    wm2l[o + 1] = xr * yi + xi * yr;
                }
    std::vector<float> wms32(wms.begin(), wms.end()), wm2l32(wm2l.begin(), wm2l.end());
    Pack pk;
    size_t o_ts = pk.add(ts), o_h2x = pk.add(h2x), o_h2xr = pk.add(h2xr), o_x2l = pk.add(x2l),
           o_rsq = pk.add(rsq);
    size_t o_wms = pk.add(wms), o_wm2l = pk.add(wm2l);
    size_t o_wms32 = pk.add(wms32), o_wm2l32 = pk.add(wm2l32);
    char* d = W.get<char>("tables", pk.size());
    HIPCHECK(hipMemcpyAsync(d, pk.data(), pk.size(), hipMemcpyHostToDevice, s));
    HIPCHECK(hipStreamSynchronize(s));  // pk is freed on return
    t.ts = (const double*)(d + o_ts);
    t.h2x = (const double*)(d + o_h2x);
    t.h2xr = (const double*)(d + o_h2xr);
    t.x2l = (const double*)(d + o_x2l);
    t.rsq = (const double*)(d + o_rsq);
    t.wms = (const double*)(d + o_wms);
    t.wm2l = (const double*)(d + o_wm2l);
    t.wms32 = (const float*)(d + o_wms32);
    t.wm2l32 = (const float*)(d + o_wm2l32);
    t.npw = npw; t.pmax = pmax; t.ntermmax = ntermmax; t.nmax = nmax; t.xmin2 = xmin2; t.xm = xm;
    t.valid = true;
    return t;
}

// 3D plane-wave tables (ascending nodes, g3drouts.f:2669-3146): C(j,a) =
// ws_a (-i ts_a)^j / j! for raw moments (the reference's ws (-ts)^j/sqrt(j!)
// with its i^j grouping and the Hermite 1/sqrt(j!)), D(j,a) = (i ts_a)^j / j!
// (g3drx2lmat with the i^j grouping), merge/split factors for the 8 octants
// per shift length (merge_split_pw_matrices, :2760-2842) and the 27 m2l
// factors (pw_translation_matrices, :2697-2757).  Cached while unchanged.
const Tables3D& tables3d(Workspace& W, int npw, double pmax, int ntermmax, int nmax, double xmin2,
                         double xm, hipStream_t s)
{
    Tables3D& t = W.tab3;
    if (t.valid && t.npw == npw && t.pmax == pmax && t.ntermmax == ntermmax && t.nmax == nmax &&
        t.xmin2 == xmin2 && t.xm == xm)
        return t;
    const int half = npw / 2, mt = ntermmax + 1;
    const size_t nexp = (size_t)npw * npw * half;
    std::vector<double> ws(npw), ts(npw);
    planewave_nodes3(pmax, npw, ws.data(), ts.data());
    std::vector<double> coef(2 * (size_t)mt * npw), dcoef(2 * (size_t)mt * npw);
    double fh = 1.0, fl = 1.0;
    for (int j = 0; j < mt; ++j) {
        if (j > 0) { fh = fh / std::sqrt((double)j); fl = fl * j; }
        for (int a = 0; a < npw; ++a) {
            const double r = ws[a] * std::pow(-ts[a], j) * fh * fh;
            const double d = std::pow(ts[a], j) / fl;
            double* c = &coef[2 * (a + (size_t)npw * j)];
            double* e = &dcoef[2 * (a + (size_t)npw * j)];
            switch (j & 3) {  // times i^j
            case 0: c[0] = r; c[1] = 0; e[0] = d; e[1] = 0; break;
            case 1: c[0] = 0; c[1] = r; e[0] = 0; e[1] = d; break;
            case 2: c[0] = -r; c[1] = 0; e[0] = -d; e[1] = 0; break;
            default: c[0] = 0; c[1] = -r; e[0] = 0; e[1] = -d; break;
            }
        }
    }
    const int nm = std::max(nmax, 1);
    std::vector<double> wms(2 * nexp * 8 * nm), wm1(2 * 3 * (size_t)npw);
    {
        std::vector<double> wr((size_t)npw * (nm + 1)), wi((size_t)npw * (nm + 1));
        for (int j = 0; j < npw; ++j) {
            double zr = std::cos(ts[j] * xmin2), zi = std::sin(ts[j] * xmin2);
            for (int k = 1; k <= nm; ++k) {
                wr[j * (nm + 1) + k] = zr;
                wi[j * (nm + 1) + k] = zi;
                const double tr = zr * zr - zi * zi, ti = zr * zi + zi * zr;
                zr = tr;
                zi = ti;
            }
        }
        for (int k1 = 1; k1 <= nm; ++k1)
            for (int c = 0; c < half; ++c)
                for (int b = 0; b < npw; ++b)
                    for (int a = 0; a < npw; ++a) {
                        const size_t e = a + (size_t)npw * (b + (size_t)npw * c);
                        const double ar = wr[a * (nm + 1) + k1], ai0 = wi[a * (nm + 1) + k1];
                        const double br = wr[b * (nm + 1) + k1], bi0 = wi[b * (nm + 1) + k1];
                        const double cr = wr[c * (nm + 1) + k1], ci0 = wi[c * (nm + 1) + k1];
                        for (int k = 0; k < 8; ++k) {
                            // x conjugated for k in {2,3,6,7}, y for odd k, z for k >= 4
                            const double ai = (k == 2 || k == 3 || k == 6 || k == 7) ? -ai0 : ai0;
                            const double bi = (k & 1) ? -bi0 : bi0;
                            const double ci = k >= 4 ? -ci0 : ci0;
                            const double tr = ar * br - ai * bi, ti = ar * bi + ai * br;
                            double* o = &wms[2 * ((size_t)(k + 8 * (k1 - 1)) * nexp + e)];
                            o[0] = tr * cr - ti * ci;
                            o[1] = tr * ci + ti * cr;
                        }
                    }
    }
    // m2l factors e^{i ts_a k xm} for the offsets k = -1, 0, +1 at 3 a + k + 1;
    // the kernel multiplies the x, y and z factors of a shift
    for (int j = 0; j < npw; ++j) {
        const double zr = std::cos(ts[j] * xm), zi = std::sin(ts[j] * xm);
        wm1[2 * (3 * j + 0)] = zr; wm1[2 * (3 * j + 0) + 1] = -zi;  // offset -1
        wm1[2 * (3 * j + 1)] = 1;  wm1[2 * (3 * j + 1) + 1] = 0;    // offset 0
        wm1[2 * (3 * j + 2)] = zr; wm1[2 * (3 * j + 2) + 1] = zi;   // offset +1
    }
    Pack pk;
    size_t o_ws = pk.add(ws), o_ts = pk.add(ts), o_c = pk.add(coef), o_d = pk.add(dcoef);
    size_t o_ms = pk.add(wms), o_ml = pk.add(wm1);
    char* d = W.get<char>("tables3", pk.size());
    HIPCHECK(hipMemcpyAsync(d, pk.data(), pk.size(), hipMemcpyHostToDevice, s));
    HIPCHECK(hipStreamSynchronize(s));  // pk is freed on return
    t.ws = (const double*)(d + o_ws);
    t.ts = (const double*)(d + o_ts);
    t.coef = (const double*)(d + o_c);
    t.dcoef = (const double*)(d + o_d);
    t.wms = (const double*)(d + o_ms);
    t.wm1 = (const double*)(d + o_ml);
    t.valid = true;
    t.npw = npw;
    t.pmax = pmax;
    t.ntermmax = ntermmax;
    t.nmax = nmax;
    t.xmin2 = xmin2;
    t.xm = xm;
    return t;
}

// single Taylor expansion about the box centre for tiny boxes,
// fgt2d_large_delta (fgt2d.f:1616-1733)
void run_large_delta(Workspace& W, const Call& q, hipStream_t s, int npwlevel, double bs0,
                     const double* cen0, Timer& tm);

// outputs of a transform without charges or dipoles: zero (fgt2d.f:882-959
// and fgt2dpart_direct_vec form and add nothing)
void zero_outputs(const Call& q, int dim, hipStream_t s)
{
    const size_t nd = q.nd, ng = dim, nh = dim == 2 ? 3 : 6;
    if (q.ifpgh >= 1) launch_zero(q.pot, nd * q.ns, s);
    if (q.ifpgh >= 2) launch_zero(q.grad, ng * nd * q.ns, s);
    if (q.ifpgh >= 3) launch_zero(q.hess, nh * nd * q.ns, s);
    if (q.nt > 0) {
        if (q.ifpghtarg >= 1) launch_zero(q.pottarg, nd * q.nt, s);
        if (q.ifpghtarg >= 2) launch_zero(q.gradtarg, ng * nd * q.nt, s);
        if (q.ifpghtarg >= 3) launch_zero(q.hesstarg, nh * nd * q.nt, s);
    }
}

void run2d(Workspace& W, const Call& q, hipStream_t s)
{
    t_mixed = false;
    W.n_order = 0;
    Timer tm(W, s);
    const int ns = q.ns, nt = q.nt, nd = q.nd;
    if (!q.chg && !q.rn) {
        zero_outputs(q, 2, s);
        return;
    }
    const int ifdipole = q.rn ? 1 : 0;
    // ---------------- bounding box (pts_tree_boxsize0) ----------------
    tm.begin(S_BBOX);
    int nblk = (int)std::min<long>(1024, std::max<long>(1, ((long)ns + nt + 255) / 256));
    double* part = W.get<double>("bbox_part", 5 * (size_t)nblk);
    double* bbd = W.get<double>("bbox", 8);
    launch_bbox(q.src, ns, q.targ, nt, part, nblk, bbd, s);
    tm.end(S_BBOX);
    double* hb = (double*)W.pinned(64);
    copy_to_host(hb, bbd, 5 * sizeof(double), s);
    if (hb[4] != 0) fail(FGTPW_EINVAL, "non-finite source or target coordinates");
    double xmin = hb[0], xmax = hb[1], ymin = hb[2], ymax = hb[3];
    double bs0 = xmax - xmin, sizey = ymax - ymin;
    double cen0[2] = {(xmin + xmax) / 2, (ymin + ymax) / 2};
    if (sizey > bs0) bs0 = sizey;
    double dc = std::sqrt(q.delta * std::log(1.0 / q.eps));
    int npwlevel;
    if (bs0 > 0) {
        double dl = std::log(bs0 / dc) / std::log(2.0);
        if (dl > 29) fail(FGTPW_EUNSUPPORTED, "box too large relative to sqrt(delta): tree deeper than 30 levels");
        npwlevel = (int)dl;
        if (bs0 >= dc) {
            bs0 = (double)(1 << (npwlevel + 1)) * dc;
            npwlevel += 1;
        }
    } else {
        npwlevel = -1000;  // all points coincide: one expansion covers them
    }
    if (npwlevel < -2) {
        run_large_delta(W, q, s, npwlevel, bs0, cen0, tm);
        return;
    }
    // deepest level: first whose Hermite length is <= 18 (fgt2d.f:175-186)
    const int nlevstart = std::max(npwlevel, 0);
    int nlmax = npwlevel + 4;
    {
        double bsize = bs0 / (double)(1 << nlevstart);
        for (int i = nlevstart; i <= 100; ++i) {
            if (hermite_terms(2, bsize, q.delta, q.eps) <= 18) {
                nlmax = i;
                break;
            }
            bsize = bsize / 2;
        }
    }
    if (nlmax > 31) fail(FGTPW_EUNSUPPORTED, "tree deeper than 31 levels");

    // ---------------- keys, sort, leaves of the adaptive tree ----------------
    tm.begin(S_KEYS);
    const int nmx = std::max(ns, nt);
    using u64 = unsigned long long;
    // Points are kept in full-depth Morton order: every leaf of the final
    // tree is a Morton interval, so this order groups them by leaf (leaves in
    // depth-first order).  Equal keys keep their input order (stable sort).
    u64* ks = W.get<u64>("keys_s", ns);
    u64* ks2 = W.get<u64>("keys_s2", ns);
    unsigned* ix = W.get<unsigned>("idx_s", nmx);
    // the library's LSD sort numbers the points itself in its first pass:
    // no index array is written and read
    unsigned* isrc = W.get<unsigned>("isrc", ns);
    size_t tmpb = std::max(sort_pairs_u64_temp(nmx), sort_pairs_u32_temp(nmx));
    tmpb = std::max(tmpb, leaf_records_temp(ns));
    void* tmp = W.get<char>("sort_tmp", tmpb);
    const int keybits = std::max(2, 2 * nlmax);
    const bool k32 = keybits <= 31;  // keys fit 32 bits
    unsigned* ix32 = keybits <= 30 && (tuning().sort == 4 || tuning().sort == 5) ? nullptr : ix;
    if (k32)
        launch_keys2d(q.src, ns, cen0[0], cen0[1], bs0, nlmax, (unsigned*)ks, ix32, s);
    else
        launch_keys2d(q.src, ns, cen0[0], cen0[1], bs0, nlmax, ks, ix, s);
    tm.bracket(KID_SORT, false, -1, ns);
    if (k32)
        sort_pairs_u32(tmp, tmpb, (unsigned*)ks, (unsigned*)ks2, ix32, isrc, ns, keybits, s, tuning().sort);
    else
        sort_pairs_u64(tmp, tmpb, ks, ks2, ix, isrc, ns, keybits, s);
    tm.bracket(KID_SORT, true, -1, ns);
    tm.end(S_KEYS);
    tm.begin(S_LEAF);
    int* lvl = W.get<int>("leaf_level", ns);
    unsigned char* flg = W.get<unsigned char>("leaf_flag", ns);
    // the leaf count (first word) and the leaf records (from byte 256) in
    // one buffer: one read-back, sized by the last call's count
    char* rec = W.get<char>("leaf_runs", 256 + sizeof(LeafRun) * (size_t)ns);
    int* nsel = (int*)rec;
    LeafRun* runs = (LeafRun*)(rec + 256);
    unsigned* bcnt = (unsigned*)tmp;  // free after the sort
    if (k32)
        launch_leafdepth((const unsigned*)ks2, ns, nlmax, npwlevel, 10, 80, lvl, flg, bcnt, s);
    else
        launch_leafdepth(ks2, ns, nlmax, npwlevel, 10, 80, lvl, flg, bcnt, s);
    if (k32)
        launch_leaf_records(flg, bcnt, ns, (const unsigned*)ks2, lvl, nlmax, nsel, runs, s);
    else
        launch_leaf_records(flg, bcnt, ns, ks2, lvl, nlmax, nsel, runs, s);
    tm.end(S_LEAF);
    const size_t guess = std::min<size_t>(ns, std::max<size_t>(4096, W.leaf_guess + W.leaf_guess / 4));
    char* hl = (char*)W.pinned(256 + sizeof(LeafRun) * guess);
    copy_to_host(hl, rec, 256 + sizeof(LeafRun) * guess, s);
    const int nleaf0 = *(const int*)hl;
    W.leaf_guess = nleaf0;
    std::vector<LeafRun> leaves(nleaf0);
    std::memcpy(leaves.data(), hl + 256, sizeof(LeafRun) * std::min<size_t>(nleaf0, guess));
    if ((size_t)nleaf0 > guess) {
        const size_t rest = sizeof(LeafRun) * ((size_t)nleaf0 - guess);
        void* hp = W.pinned2(rest);
        copy_to_host(hp, runs + guess, rest, s);
        std::memcpy(leaves.data() + guess, hp, rest);
    }

    // points and charges into Morton order on a side stream: only the
    // expansion steps read them, so the gathers overlap the host's tree.
    // Queued once the leaves are back: beside a gather, the leaf searches
    // take twice as long, the one-block scan waits for a free CU and the
    // read-back's copy kernel takes 15 times as long.
    hipStream_t s_g = g_serial ? s : W.aux[0];
    W.order(s, s_g);
    tm.begin(S_GATHER, s_g);
    double* xs = W.get<double>("xs", 2 * (size_t)ns);
    double* qs = W.get<double>("qs", (size_t)nd * ns);
    tm.ctx(-1, nullptr, -1, ns);
    if (q.chg)
        launch_gather(q.src, q.chg, nd, isrc, ns, xs, qs, s_g);
    else {
        launch_gather(q.src, nullptr, nd, isrc, ns, xs, nullptr, s_g);
        launch_zero(qs, (size_t)nd * ns, s_g);
    }
    double* wsd = nullptr;  // dipole weights d n in sorted order
    if (q.rn) {
        wsd = W.get<double>("dip_w", 2 * (size_t)nd * ns);
        launch_gather_dip(q.rn, q.ds, nd, isrc, ns, wsd, s_g);
    }
    // the near field's sorted outputs, accumulated into by the near field
    // and cleared here; the evaluations add them to their own values and
    // write the sums to the caller's arrays in input order, and the points
    // that no evaluation covers get them alone (the sorted order's full
    // outputs are never formed)
    const bool use_src = q.ifpgh > 0, use_trg = q.ifpghtarg > 0 && nt > 0;
    const int ws_ = use_src ? ns : 0, wt_ = use_trg ? nt : 0;
    auto cleared = [&](const char* name, int k, int ifp, int n) {
        if (ifp < k) return (double*)nullptr;
        double* p = W.get<double>(name, (size_t)k * nd * n);
        launch_zero(p, (size_t)k * nd * n, s_g);
        return p;
    };
    double *nps = nullptr, *ngs = nullptr, *nhs = nullptr, *npt = nullptr, *ngt = nullptr,
           *nht = nullptr;
    if (use_src) {
        nps = cleared("nf_pot_s", 1, q.ifpgh, ws_);
        ngs = cleared("nf_grad_s", 2, q.ifpgh, ws_);
        nhs = cleared("nf_hess_s", 3, q.ifpgh, ws_);
    }
    if (use_trg) {
        npt = cleared("nf_pot_t", 1, q.ifpghtarg, wt_);
        ngt = cleared("nf_grad_t", 2, q.ifpghtarg, wt_);
        nht = cleared("nf_hess_t", 3, q.ifpghtarg, wt_);
    }
    tm.end(S_GATHER, s_g);

    // Work that needs no tree, queued now so that the GPU runs it while the
    // host builds the tree: targets to Morton order, their points gathered
    // into that order.
    tm.begin(S_ASSIGN);
    unsigned* itarg = nullptr;
    u64* kt2 = nullptr;
    double* xt = nullptr;
    if (nt > 0) {
        itarg = W.get<unsigned>("itarg", nt);
        xt = W.get<double>("xt", 2 * (size_t)nt);
        u64* kt = W.get<u64>("keys_t", nt);
        kt2 = W.get<u64>("keys_t2", nt);
        if (k32)
            launch_keys2d(q.targ, nt, cen0[0], cen0[1], bs0, nlmax, (unsigned*)kt, ix32, s);
        else
            launch_keys2d(q.targ, nt, cen0[0], cen0[1], bs0, nlmax, kt, ix, s);
        tm.bracket(KID_SORT, false, -2, nt);
        if (k32)
            sort_pairs_u32(tmp, tmpb, (unsigned*)kt, (unsigned*)kt2, ix32, itarg, nt, keybits, s, tuning().sort);
        else
            sort_pairs_u64(tmp, tmpb, kt, kt2, ix, itarg, nt, keybits, s);
        tm.bracket(KID_SORT, true, -2, nt);
        tm.ctx(-2, nullptr, -1, nt);  // level -2 tags the gather without densities
        launch_gather(q.targ, nullptr, 0, itarg, nt, xt, nullptr, s);
    }
    tm.end(S_ASSIGN);

    // ---------------- the finest level, formed beside the host tree ----------------
    // The level restriction never refines the two deepest levels, so the
    // finest level's boxes are those of the adaptive tree: the children of
    // the refined boxes one level up, in the order the leaf records first
    // reach them, with the records' point ranges.  Their centres follow the
    // builder's recursion (tree.cpp set_children), and the slot numbering puts
    // them first (slots 0..nL-1), so the helper thread forms them on the side
    // stream while this thread builds the tree.  After the tree is built its
    // boxes, centres and slot count are checked against what was used; on
    // any difference the level is formed again with the rest.  Not in the
    // timing modes (the per-kernel records follow the main thread) nor for
    // the no-allocation entry point (its scratch is sized without it).
    int L0 = 0;
    for (const LeafRun& r : leaves) L0 = std::max(L0, r.level);
    std::vector<double> bsz0(L0 + 1);
    bsz0[0] = bs0;
    for (int l = 0; l < L0; ++l) bsz0[l + 1] = bsz0[l] / 2;  // as the builder halves
    const bool early_ok = !g_serial && !tm.on && !W.arena && npwlevel < L0 && L0 >= 1 &&
                          L0 >= std::max(npwlevel + 1, 0);
    struct Early {
        bool on = false;
        int nL = 0, nl = 0, npw = 0;
        bool mixed = false;
        size_t cap = 0;  // slots the expansion store holds
    } early;
    if (early_ok) {
        early.on = true;
        if (!W.helper) W.helper.reset(new HostHelper);
        auto job = [&]() {
            const int L = L0;
            // leaf record of every finest-level box (4 per refined parent), or -1
            std::vector<int>& rec = W.h_erec;
            rec.clear();
            unsigned long long lastp = ~0ull;
            for (int r = 0; r < nleaf0; ++r) {
                if (leaves[r].level != L) continue;
                const unsigned long long par = leaves[r].key >> 2;
                if (par != lastp) {
                    rec.insert(rec.end(), 4, -1);
                    lastp = par;
                }
                rec[rec.size() - 4 + (size_t)(leaves[r].key & 3)] = r;
            }
            const int nL = (int)rec.size() / 4 * 4;
            early.nL = nL;
            std::vector<double>& ec = W.h_ecen;
            std::vector<int2>& se = W.h_ese;
            std::vector<int>& iota = W.h_eiota;
            ec.resize(2 * (size_t)nL);
            se.resize(nL);
            iota.resize(nL);
            for (int k = 0; k < nL; k += 4) {
                int r0 = -1;  // any record of the family gives the parent's prefix
                for (int j = 0; j < 4 && r0 < 0; ++j) r0 = rec[k + j];
                const unsigned long long par = leaves[r0].key >> 2;
                for (int j = 0; j < 4; ++j) {
                    const unsigned long long key = (par << 2) | (unsigned long long)j;
                    double c[2] = {cen0[0], cen0[1]};
                    for (int l = 0; l < L; ++l) {
                        const int dg = (int)((key >> (2 * (L - 1 - l))) & 3);
                        const double bs = bsz0[l + 1];
                        for (int a = 0; a < 2; ++a) c[a] = c[a] + (((dg >> a) & 1) ? bs / 2 : -bs / 2);
                    }
                    ec[2 * (size_t)(k + j)] = c[0];
                    ec[2 * (size_t)(k + j) + 1] = c[1];
                    const int r = rec[k + j];
                    se[k + j] = r < 0 ? int2{0, 0}
                                      : int2{leaves[r].start, r + 1 < nleaf0 ? leaves[r + 1].start : ns};
                    iota[k + j] = k + j;
                }
            }
            // terms, nodes and tables as the transform computes them after the tree
            const int npg0 = std::max(q.ifpgh, q.ifpghtarg), nadd0 = ifdipole + npg0;
            int ntm = 0, nlL = 0;
            for (int i = std::max(npwlevel, 0); i <= L; ++i) {
                int nl = hermite_terms(2, bsz0[i], q.delta, q.eps);
                if (ifdipole || q.ifpgh >= 2 || q.ifpghtarg >= 2) nl += nadd0;
                ntm = std::max(ntm, nl);
                if (i == L) nlL = nl;
            }
            double pmax0 = 0;
            int npw0 = 0;
            planewave_terms(2 * bs0 / std::pow(2.0, std::max(npwlevel, 0)), q.delta,
                            nadd0 > 2 ? q.eps / 100 : q.eps, &pmax0, &npw0);
            const bool mixed0 = g_precision == 1 && q.eps >= kMixedEpsMin && ntm < 63;
            const size_t esz0 = mixed0 ? sizeof(float) : sizeof(double);
            early.nl = nlL;
            early.npw = npw0;
            early.mixed = mixed0;
            const int nlevstart0 = std::max(npwlevel, 0);
            const Tables2D& tab0 =
                tables2d(W, npw0, pmax0, ntm, L - nlevstart0, bsz0[L] / std::sqrt(q.delta) / 2,
                         (npwlevel >= 0 ? bsz0[npwlevel] : bsz0[0]) / std::sqrt(q.delta), s_g);
            // the expansion store, sized for the slots of this call as the
            // last call's tree suggests (checked after the tree)
            const size_t per0 = 2 * (size_t)(npw0 / 2) * npw0 * nd;
            early.cap = std::max<size_t>(4 * (size_t)nL, (size_t)W.last_nslot + W.last_nslot / 4) + 1024;
            W.get<char>("mp", per0 * early.cap * esz0);
            Pack pe(W.h_plan_early, true);
            const size_t o_b = pe.add(iota), o_c = pe.add(ec), o_s = pe.add(se);
            char* d = W.get<char>("plan_early", pe.size());
            HIPCHECK(hipMemcpyAsync(d, pe.data(), pe.size(), hipMemcpyHostToDevice, s_g));
            const Tuning& tune = tuning();
            const int m = nlL + 1, seg = tune.segf[std::min(4, (nlL + 16) / 16) - 1];
            int* ts = W.get<int>("tstart_early", (size_t)nL + 1);
            int* sb = W.get<int>("segbox_early", (size_t)nL + ns / seg + 1);
            double* part = W.get<double>("form_partial_early", ((size_t)nL + ns / seg + 1) * m * m);
            SegLists sl{};
            sl.l[0] = SegList{(const int*)(d + o_b), nL, (const int2*)(d + o_s), seg, ts, sb};
            sl.count = 1;
            launch_seg_scan(sl, s_g);
            for (int ind = 0; ind < nd; ++ind) {
                FormArgs fa{(const int*)(d + o_b), nL, (const double*)(d + o_c), (const int2*)(d + o_s),
                            (const int*)(d + o_b), xs, qs, nd, ind, 1.0 / std::sqrt(q.delta), nlL,
                            tab0.h2x, tab0.h2xr, tab0.rsq, npw0,
                            (double*)W.get<char>("mp", per0 * early.cap * esz0), ts, seg,
                            (int)std::min<long>(4096, nL + ns / seg + 1), part, m * m, sb, wsd};
                fa.prec = mixed0 ? 1 : 0;
                launch_form(fa, s_g);
            }
        };
        W.helper->post(job);
    }
    HelperJoin early_join{early.on ? W.helper.get() : nullptr};

    // ---------------- host: tree, lists, tables ----------------
    auto th0 = std::chrono::steady_clock::now();
    W.last_valid = false;
    Tree2D& T = W.last;
    build_tree2d(T, leaves, nlmax, npwlevel, bs0, cen0);
    tm.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - th0).count();
    if (early.on) {
        early_join.h = nullptr;
        W.helper->wait();  // the early formation is queued (rethrows its errors)
    }
    W.last_ns = ns;
    W.last_nt = nt;
    W.last_dim = 2;
    const int nb = T.nboxes(), nlevels = T.nlevels, nleaves = (int)T.leaf_of_rank.size();
    const int npg = std::max(q.ifpgh, q.ifpghtarg);
    const int nadd = ifdipole + npg;  // fgt2d.f:340
    std::vector<int> nlocal(nlevels + 1, 0);
    int ntermmax = 0;
    for (int i = nlevstart; i <= nlevels; ++i) {
        nlocal[i] = hermite_terms(2, T.bsz[i], q.delta, q.eps);
        if (ifdipole || q.ifpgh >= 2 || q.ifpghtarg >= 2) nlocal[i] += nadd;
        ntermmax = std::max(ntermmax, nlocal[i]);
    }
    double pmax = 0;
    int npw = 0;
    {
        double bsize = 2 * bs0 / std::pow(2.0, nlevstart);
        double pweps = nadd > 2 ? q.eps / 100 : q.eps;
        planewave_terms(bsize, q.delta, pweps, &pmax, &npw);
    }
    const bool expansions = npwlevel < nlevels;
    const int half = npw / 2, nexp = half * npw;
    const double dsq = 1.0 / std::sqrt(q.delta);
    // the mixed-precision route: FP32 expansions, matrix-core moments and
    // Taylor evaluation where the order allows (formation orders below 64)
    const bool mixed = g_precision == 1 && expansions && q.eps >= kMixedEpsMin && ntermmax < 63;
    const size_t esz = mixed ? sizeof(float) : sizeof(double);

    // box arrays; expansion slots: the finest level first (its leaves are
    // formed first, possibly before the tree), then the other levels from
    // the cutoff level down
    std::vector<int>& slot = W.h_slot;
    slot.assign(nb, -1);
    int nslot = 0;
    const int bL0 = T.laddr[2 * nlevels], bL1 = T.laddr[2 * nlevels + 1], nL = bL1 - bL0 + 1;
    if (expansions) {
        for (int b = bL0; b <= bL1; ++b) slot[b] = nslot++;
        for (int b = 0; b < bL0; ++b)
            if (T.lev[b] >= nlevstart) slot[b] = nslot++;
    }
    // the early formation holds if it formed these boxes with these centres
    // and parameters and its store has room for every slot
    if (early.on)
        early.on = expansions && nlevels == L0 && nL == early.nL && (size_t)nslot <= early.cap &&
                   nlocal[nlevels] == early.nl && npw == early.npw && mixed == early.mixed &&
                   std::memcmp(T.cen.data() + 2 * (size_t)bL0, W.h_ecen.data(),
                               sizeof(double) * 2 * (size_t)nL) == 0;
    W.last_nslot = nslot;
    Pack pk(W.h_plan, !W.arena);
    size_t o_child = pk.add(T.ch), o_par = pk.add(T.par), o_cen = pk.add(T.cen);
    // colleague lists are read on the device by m2l only, for the boxes of
    // the top level (a contiguous range): only that slice travels
    const int colb0 = expansions ? T.laddr[2 * nlevstart] : 0;
    const int colb1 = expansions ? T.laddr[2 * nlevstart + 1] : -1;
    size_t o_ncol = pk.add(T.ncol);
    size_t o_col = pk.add(T.col.data() + 9 * (size_t)colb0, 9 * (size_t)(colb1 - colb0 + 1));
    size_t o_lkey = pack_leaf_keys(pk, W, T, nlmax, k32);
    size_t o_rlo = pk.add(T.rank_lo), o_rhi = pk.add(T.rank_hi), o_slot = pk.add(slot);

    // work lists of the expansion steps (storage kept from call to call)
    auto lists = [&](std::vector<std::vector<int>>& v) -> std::vector<std::vector<int>>& {
        if (v.size() < (size_t)nlevels + 1) v.resize(nlevels + 1);
        for (auto& x : v) x.clear();
        return v;
    };
    std::vector<std::vector<int>>& form_l = lists(W.h_form);
    std::vector<std::vector<int>>& merge_l = lists(W.h_merge);
    std::vector<std::vector<int>>& split_l = lists(W.h_split);
    std::vector<std::vector<int>>& eval_l = lists(W.h_eval);
    std::vector<int>& m2l_l = W.h_m2l;
    m2l_l.clear();
    std::vector<int>& resid = W.h_resid;
    resid.clear();
    std::vector<size_t> o_form(nlevels + 1), o_merge(nlevels + 1), o_split(nlevels + 1),
        o_eval(nlevels + 1);
    size_t o_m2l = 0;
    if (expansions) {
        // leaves of the cutoff level with a refined colleague (ifhung)
        std::vector<char>& iflocal = W.h_iflocal;
        iflocal.assign(nb, 0);
        if (npwlevel >= 0 && npwlevel <= nlevels)
            for (int b = T.laddr[2 * npwlevel]; b <= T.laddr[2 * npwlevel + 1]; ++b) {
                if (T.nch[b] != 0) continue;
                for (int i = 0; i < T.ncol[b]; ++i)
                    if (T.nch[T.col[9 * b + i]] > 0) { iflocal[b] = 1; break; }
            }
        const int* nch = T.nch.data();
        const int* ch = T.ch.data();
        for (int l = 0; l <= nlevels; ++l) {
            const int b0 = T.laddr[2 * l], b1 = T.laddr[2 * l + 1];
            const bool fl = l >= std::max(npwlevel + 1, 0), el = l >= nlevstart,
                       ml = l >= nlevstart && l <= nlevels - 1;
            std::vector<int>& f = form_l[l];
            std::vector<int>& m = merge_l[l];
            std::vector<int>& sp = split_l[l];
            std::vector<int>& e = eval_l[l];
            for (int b = b0; b <= b1; ++b) {
                if (nch[b] == 0) {
                    if (fl) f.push_back(b);
                    if (el && !(l == npwlevel && !iflocal[b]))
                        e.push_back(b);
                    else
                        resid.push_back(b);
                } else if (ml) {
                    m.push_back(b);
                    sp.insert(sp.end(), ch + 4 * b, ch + 4 * b + 4);
                }
            }
            if (l == nlevstart)
                for (int b = b0; b <= b1; ++b) m2l_l.push_back(b);
        }
        for (int l = 0; l <= nlevels; ++l) {
            o_form[l] = pk.add(form_l[l]);
            o_merge[l] = pk.add(merge_l[l]);
            o_split[l] = pk.add(split_l[l]);
            o_eval[l] = pk.add(eval_l[l]);
        }
        o_m2l = pk.add(m2l_l);
    }
    const size_t o_resid = pk.add(resid);
    const Tables2D& tab = expansions ? tables2d(W, npw, pmax, ntermmax, nlevels - nlevstart,
                                                T.bsz[nlevels] / std::sqrt(q.delta) / 2,
                                                (npwlevel >= 0 ? T.bsz[npwlevel] : T.bsz[0]) /
                                                    std::sqrt(q.delta),
                                                s)
                                     : W.tab;

    // near field: modified list1 (compute_modified_list1, tree_routs4.f:715-879)
    // inverted to target units; a unit is a leaf at or above the cutoff level
    // or a box of the cutoff level, and owns every target below it.  Built
    // after steps 1-3 are queued (only the side-stream near field reads
    // them), so this host work overlaps the GPU's.
    const int nlevend = std::min(npwlevel, nlevels);
    std::vector<int>& unit_ids = W.h_units;
    std::vector<int>& unit_ptr = W.h_uptr;
    std::vector<int>& unit_list = W.h_ulist;
    auto near_field_lists = [&]() {
    unit_ids.clear();
    unit_ptr.assign(1, 0);
    unit_list.clear();
    if (nlevend >= 0) {
        std::vector<int>& nl1 = W.h_nl1;
        std::vector<int>& l1 = W.h_l1;
        nl1.assign(nb, 0);
        if (l1.size() < 13 * (size_t)nb) l1.resize(13 * (size_t)nb);
        if (T.nch[0] == 0) { nl1[0] = 1; l1[0] = 0; }
        // only lists of boxes at or above the cutoff level are ever read
        for (int l = 1; l <= nlevend; ++l)
            for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; ++b) {
                if (T.nch[b] != 0) continue;
                for (int i = 0; i < T.ncol[b]; ++i) {
                    int jb = T.col[9 * b + i];
                    if (l >= npwlevel || T.nch[jb] == 0) {
                        l1[13 * (size_t)b + nl1[b]++] = jb;
                    } else {
                        double tol = 1.05 * (T.bsz[l] + T.bsz[l + 1]) / 2.0;
                        for (int j = 0; j < T.nch[jb]; ++j) {
                            int kb = T.ch[4 * jb + j];
                            if (std::fabs(T.cen[2 * kb] - T.cen[2 * b]) < tol &&
                                std::fabs(T.cen[2 * kb + 1] - T.cen[2 * b + 1]) < tol) {
                                l1[13 * (size_t)b + nl1[b]++] = kb;
                                l1[13 * (size_t)kb + nl1[kb]++] = b;
                            }
                        }
                    }
                }
            }
        // sources of every target box, in the reference's order of source
        // boxes (CSR: count, prefix sum, fill)
        std::vector<int>& sp = W.h_sp;
        std::vector<int>& sl = W.h_sl;
        sp.assign(nb + 1, 0);
        for (int l = 0; l <= nlevend; ++l)
            for (int sb = T.laddr[2 * l]; sb <= T.laddr[2 * l + 1]; ++sb) {
                if (T.nch[sb] != 0) continue;
                for (int i = 0; i < nl1[sb]; ++i) ++sp[l1[13 * (size_t)sb + i] + 1];
            }
        for (int b = 0; b < nb; ++b) sp[b + 1] += sp[b];
        sl.resize(sp[nb]);
        std::vector<int>& fill = W.h_fill;
        fill.assign(sp.begin(), sp.end() - 1);
        for (int l = 0; l <= nlevend; ++l)
            for (int sb = T.laddr[2 * l]; sb <= T.laddr[2 * l + 1]; ++sb) {
                if (T.nch[sb] != 0) continue;
                for (int i = 0; i < nl1[sb]; ++i) sl[fill[l1[13 * (size_t)sb + i]]++] = sb;
            }
        for (int b = 0; b < nb; ++b) {
            bool unit = (T.nch[b] == 0 && T.lev[b] <= npwlevel) || T.lev[b] == npwlevel;
            if (!unit) continue;
            size_t before = unit_list.size();
            for (int a = b; a >= 0; a = T.par[a])
                unit_list.insert(unit_list.end(), sl.begin() + sp[a], sl.begin() + sp[a + 1]);
            if (unit_list.size() == before) continue;
            unit_ids.push_back(b);
            unit_ptr.push_back((int)unit_list.size());
        }
    }
    };
    auto th1 = std::chrono::steady_clock::now();
    tm.host_ms = std::chrono::duration<double, std::milli>(th1 - th0).count();

    // upload the plan in one copy
    char* dplan = W.get<char>("plan", pk.size());
    {
        HIPCHECK(hipMemcpyAsync(dplan, pk.data(), pk.size(), hipMemcpyHostToDevice, s));
    }
    // the near field's lists on the helper thread, while this one queues
    // the expansion steps; the guard waits for them on every way out (the
    // job reads this call's locals)
    if (!W.helper) W.helper.reset(new HostHelper);
    W.helper->post(near_field_lists);
    HelperJoin nf_join{W.helper.get()};

    // the expansion stores are accumulated into: cleared on the side stream
    // while the point ranges are found
    const size_t per = 2 * (size_t)nexp * nd;  // elements (double or float) per slot
    double* mp = expansions ? (double*)W.get<char>("mp", per * nslot * esz) : nullptr;
    double* lc = expansions ? (double*)W.get<char>("loc", per * nslot * esz) : nullptr;
    const size_t nexp_d = per * nslot * esz / sizeof(double);  // per is even
    // the finest level's leaves store their expansions (no clearing); the
    // other boxes accumulate (merges)
    const size_t nfin_d = expansions ? per * nL * esz / sizeof(double) : 0;
    if (expansions) launch_zero3(nullptr, 0, mp + nfin_d, nexp_d - nfin_d, lc, nexp_d, s_g);
    auto D = [&](size_t off) { return (void*)(dplan + off); };
    const int4* d_child = (const int4*)D(o_child);
    const int* d_par = (const int*)D(o_par);
    const double* d_cen = (const double*)D(o_cen);
    const int* d_ncol = (const int*)D(o_ncol);
    const int* d_col = (const int*)D(o_col) - 9 * (ptrdiff_t)colb0;  // indexed by box
    const void* d_lkey = D(o_lkey);
    const int* d_rlo = (const int*)D(o_rlo);
    const int* d_rhi = (const int*)D(o_rhi);
    const int* d_slot = (const int*)D(o_slot);

    // ---------------- point ranges of the boxes ----------------
    // first point of every leaf (a search of the sorted keys), then the
    // range of every box
    tm.begin(S_RANGES);
    // sources and targets in one launch each
    int* ls = W.get<int>("leafstart", 2 * ((size_t)nleaves + 1));
    int* ls2 = ls + nleaves + 1;
    int2* src_se = W.get<int2>("src_se", nb);
    int2* trg_se = W.get<int2>("trg_se", nb);
    const bool tg = nt > 0;
    if (k32)
        launch_leafstart_search((const unsigned*)ks2, ns, tg ? (const unsigned*)kt2 : nullptr, nt,
                                (const unsigned*)d_lkey, nleaves, ls, ls2, s);
    else
        launch_leafstart_search(ks2, ns, tg ? kt2 : nullptr, nt, (const u64*)d_lkey, nleaves, ls,
                                ls2, s);
    launch_box_ranges(d_rlo, d_rhi, nb, ls, src_se, ls2, tg ? trg_se : nullptr, s);
    tm.src_se = src_se, tm.trg_se = trg_se, tm.nb = nb, tm.nd = nd;
    if (!tg) launch_zero((double*)trg_se, (size_t)nb, s);  // int2 pairs, 8 bytes each
    tm.end(S_RANGES);
    // everything after reads the side stream's work: the sorted points,
    // charges and dipole weights, and the cleared outputs and expansions
    W.order(s_g, s);

    // ---------------- step 6: near field, cutoff eps/90 (fgt2d.f:1316) ----------------
    // on a side stream, into outputs of its own (cleared with the others);
    // its lists are built on the host once every expansion step is queued
    hipStream_t s_nf = g_serial ? s : W.aux[1], s_lv = g_serial ? s : W.aux[0];
    W.order(s, s_nf);  // after the ranges, the sorted points and the cleared outputs
    Workspace::Mark nf_ready;
    bool nf_queued = false;
    auto near_field = [&]() {
    nf_queued = true;
    nf_join.h = nullptr;
    W.helper->wait();  // the lists
    tm.begin(S_DIRECT, s_nf);
    Pack pn(W.h_plan_nf, !W.arena);
    size_t o_units = pn.add(unit_ids), o_uptr = pn.add(unit_ptr), o_ulist = pn.add(unit_list);
    char* dnf = W.get<char>("plan_nf", pn.size());
    {
        HIPCHECK(hipMemcpyAsync(dnf, pn.data(), pn.size(), hipMemcpyHostToDevice, s_nf));
    }
    if (!unit_ids.empty()) {
        DirectArgs da{};
        da.units = (const int*)(dnf + o_units);
        da.nunits = (int)unit_ids.size();
        da.list_ptr = (const int*)(dnf + o_uptr);
        da.list = (const int*)(dnf + o_ulist);
        da.src_se = src_se;
        da.xs = xs;
        da.qs = qs;
        da.nd = nd;
        da.ws = wsd;
        da.delta = q.delta;
        double e90 = q.eps / 90;
        da.dmax = std::log(1.0 / e90) * q.delta;
        if (use_trg) {
            DirectArgs d2 = da;
            d2.tgt_se = trg_se;
            d2.xt = xt;
            d2.ifpgh = q.ifpghtarg;
            d2.pot = npt; d2.grad = ngt; d2.hess = nht;
            tm.ctx(npwlevel, nullptr, 3);  // units: the pairs, counted after the call
            launch_direct(d2, s_nf);
        }
        if (use_src) {
            DirectArgs d2 = da;
            d2.tgt_se = src_se;
            d2.xt = xs;
            d2.ifpgh = q.ifpgh;
            d2.pot = nps; d2.grad = ngs; d2.hess = nhs;
            tm.ctx(npwlevel, nullptr, 4);
            launch_direct(d2, s_nf);
        }
    }
    tm.end(S_DIRECT, s_nf);
    // the points that no evaluation writes: the near field alone, in input
    // order (with no expansions, every point)
    tm.begin(S_SCATTER, s_nf);
    auto residual = [&](const unsigned* perm, int n, const int2* se, double* v, const double* nf,
                        int w) {
        if (!v) return;
        if (!expansions)
            launch_scatter(nullptr, perm, n, w, v, s_nf, nf);
        else
            launch_scatter_boxes((const int*)D(o_resid), (int)resid.size(), se, nf, perm, w, v,
                                 s_nf);
    };
    if (use_src) {
        if (expansions)
            tm.ctx(-1, &resid, 0);  // units: the boxes' points, counted after the call
        else
            tm.ctx(-1, nullptr, -1, ns);
        residual(isrc, ns, src_se, q.pot, nps, nd);
        if (q.ifpgh >= 2) residual(isrc, ns, src_se, q.grad, ngs, 2 * nd);
        if (q.ifpgh >= 3) residual(isrc, ns, src_se, q.hess, nhs, 3 * nd);
    }
    if (use_trg) {
        if (expansions)
            tm.ctx(-1, &resid, 1);
        else
            tm.ctx(-1, nullptr, -1, nt);
        residual(itarg, nt, trg_se, q.pottarg, npt, nd);
        if (q.ifpghtarg >= 2) residual(itarg, nt, trg_se, q.gradtarg, ngt, 2 * nd);
        if (q.ifpghtarg >= 3) residual(itarg, nt, trg_se, q.hesstarg, nht, 3 * nd);
    }
    tm.end(S_SCATTER, s_nf);
    nf_ready = W.mark(s_nf);
    };
    // queued as soon as the helper has its lists (checked between steps)
    auto near_field_if_ready = [&]() {
        if (!nf_queued && W.helper->done()) near_field();
    };

    // ---------------- expansions: steps 1-5 ----------------
    if (expansions) {
        const double* d_h2x = tab.h2x;
        const double* d_h2xr = tab.h2xr;
        const double* d_x2l = tab.x2l;
        const double* d_rsq = tab.rsq;
        const double* d_ts = tab.ts;
        const double* d_wms = mixed ? (const double*)tab.wms32 : tab.wms;
        const double* d_wm2l = mixed ? (const double*)tab.wm2l32 : tab.wm2l;
        // Taylor coefficients of every evaluated box, one region per level
        // (levels are evaluated concurrently)
        std::vector<size_t> o_taylor(nlevels + 2, 0);
        for (int l = 0; l <= nlevels; ++l) {
            const bool fused = g_fuse_pw2local && nlocal[l] + 1 <= 32 && npw <= 64;
            o_taylor[l + 1] = o_taylor[l] + (nlocal[l] <= 50 && !fused ? eval_l[l].size() * (size_t)(nlocal[l] + 1) * (nlocal[l] + 1) : 0);
        }
        double* taylor = W.get<double>("taylor", o_taylor[nlevels + 1]);
        // segment numbering of every formation and evaluation list (one
        // launch): formation over sources in segments of SEGF, evaluation
        // over targets / sources in segments of sege(l)
        const Tuning& tune = tuning();
        auto tmof = [&](int l) { return std::min(4, (nlocal[l] + 16) / 16); };
        auto segf = [&](int l) { return tune.segf[tmof(l) - 1]; };
        auto sege = [&](int l) { return tune.sege[tmof(l) - 1]; };
        const int GRIDMAX = 4096;
        std::vector<int*> ts_form(nlevels + 1, nullptr), ts_trg(nlevels + 1, nullptr),
            ts_src(nlevels + 1, nullptr), sb_form(nlevels + 1, nullptr),
            sb_trg(nlevels + 1, nullptr), sb_src(nlevels + 1, nullptr);
        {
            // tstart: one entry per box and list plus the total; segbox: one
            // per segment, at most boxes + points / seg + 1 per list
            size_t ntot = 0, nseg = 0;
            for (int l = 0; l <= nlevels; ++l) {
                ntot += form_l[l].size() + 2 * eval_l[l].size() + 3;
                if (l >= std::max(npwlevel + 1, 0) && !form_l[l].empty())
                    nseg += form_l[l].size() + ns / segf(l) + 1;
                if (l >= nlevstart && !eval_l[l].empty())
                    nseg += 2 * eval_l[l].size() + (size_t)(nt + ns) / sege(l) + 2;
            }
            int* pool = W.get<int>("tstart", ntot);
            int* spool = W.get<int>("segbox", nseg);
            SegLists L{};
            auto add = [&](const std::vector<int>& v, size_t off, const int2* se, int seg,
                           int npts, int** sb) {
                int* t = pool;
                pool += v.size() + 1;
                *sb = spool;
                spool += v.size() + npts / seg + 1;
                if (L.count == 32) {
                    launch_seg_scan(L, s);
                    L.count = 0;
                }
                L.l[L.count++] = SegList{(const int*)D(off), (int)v.size(), se, seg, t, *sb};
                return t;
            };
            for (int l = nlevels - (early.on ? 1 : 0); l >= std::max(npwlevel + 1, 0); --l)
                if (!form_l[l].empty())
                    ts_form[l] = add(form_l[l], o_form[l], src_se, segf(l), ns, &sb_form[l]);
            for (int l = nlevstart; l <= nlevels; ++l) {
                if (eval_l[l].empty()) continue;
                if (use_trg) ts_trg[l] = add(eval_l[l], o_eval[l], trg_se, sege(l), nt, &sb_trg[l]);
                if (use_src) ts_src[l] = add(eval_l[l], o_eval[l], src_se, sege(l), ns, &sb_src[l]);
            }
            launch_seg_scan(L, s);
        }
        // partial moments of every segment, one region per level (the
        // finest level is formed on s, the others concurrently on s_lv)
        std::vector<size_t> o_part(nlevels + 2, 0);
        for (int l = 0; l <= nlevels; ++l) {
            size_t m = nlocal[l] + 1;
            o_part[l + 1] = o_part[l] + (l >= std::max(npwlevel + 1, 0) && !form_l[l].empty()
                                             ? (form_l[l].size() + ns / segf(l) + 1) * m * m
                                             : 0);
        }
        double* partial = W.get<double>("form_partial", o_part[nlevels + 1]);
        tm.begin(S_FORM);
        // the finest level on s (unless formed already), the others on s_lv;
        // the orders only when s_lv has a level to form (each costs host time)
        const int lfirst = nlevels - (early.on ? 1 : 0);
        bool coarse = false;
        for (int l = lfirst - 1; l >= std::max(npwlevel + 1, 0); --l) coarse = coarse || !form_l[l].empty();
        if (coarse) W.order(s, s_lv);
        for (int ind = 0; ind < nd; ++ind) {
            for (int l = lfirst; l >= std::max(npwlevel + 1, 0); --l) {
                if (form_l[l].empty()) continue;
                const int nb_l = (int)form_l[l].size(), m = nlocal[l] + 1;
                FormArgs fa{(const int*)D(o_form[l]), nb_l, d_cen, src_se, d_slot, xs, qs, nd, ind,
                            dsq, nlocal[l], d_h2x, d_h2xr, d_rsq, npw, mp, ts_form[l], segf(l),
                            (int)std::min<long>(GRIDMAX, nb_l + ns / segf(l) + 1),
                            partial + o_part[l], m * m, sb_form[l], wsd};
                fa.prec = mixed ? 1 : 0;
                tm.ctx(l, &form_l[l], 0);
                launch_form(fa, l == lfirst ? s : s_lv);
            }
        }
        if (coarse) W.order(s_lv, s);
        tm.end(S_FORM);
        near_field_if_ready();
        tm.begin(S_MERGE);
        for (int l = nlevels - 1; l >= nlevstart; --l)
            tm.ctx(l, &merge_l[l], 2), launch_merge((const int*)D(o_merge[l]), (int)merge_l[l].size(), d_child, d_cen, src_se,
                         d_slot, d_wms, 4 * (nlevels - l - 1), nd, nexp, mp, s, mixed);
        tm.end(S_MERGE);
        tm.begin(S_M2L);
        tm.ctx(nlevstart, &m2l_l, 2);
        launch_m2l((const int*)D(o_m2l), (int)m2l_l.size(), d_cen, d_ncol, d_col, src_se, trg_se,
                   use_src, use_trg, d_slot, d_wm2l, npwlevel >= 0 ? T.bsz[npwlevel] : T.bsz[0],
                   nlevstart == npwlevel, nd, nexp, mp, lc, s, mixed);
        tm.end(S_M2L);
        near_field_if_ready();
        auto eval_level = [&](int l, hipStream_t st) {
            for (int ind = 0; ind < nd; ++ind) {
                if (eval_l[l].empty()) continue;
                EvalArgs ea{};
                ea.boxes = (const int*)D(o_eval[l]);
                ea.nbox = (int)eval_l[l].size();
                ea.cen = d_cen;
                ea.slot = d_slot;
                ea.nd = nd;
                ea.ind = ind;
                ea.dsq = dsq;
                ea.nl = nlocal[l];
                ea.x2l = d_x2l;
                ea.ts = d_ts;
                ea.npw = npw;
                ea.loc = lc;
                ea.taylor = taylor + o_taylor[l];
                bool pwroute = nlocal[l] > 50;  // nlocalswitch, fgt2d.f:855
                ea.exp32 = mixed ? 1 : 0;
                ea.tay32 = mixed && !pwroute && nlocal[l] + 1 <= 32 ? 1 : 0;
                // orders below 32: PW -> Taylor inside the evaluation waves
                ea.fused = !pwroute && g_fuse_pw2local && nlocal[l] + 1 <= 32 && npw <= 64 ? 1 : 0;
                tm.ctx(l, &eval_l[l], 2);
                if (!pwroute && !ea.fused) launch_pw2local(ea, st);
                for (int side = 0; side < 2; ++side) {
                    if (side == 0 && !use_trg) continue;
                    if (side == 1 && !use_src) continue;
                    EvalArgs eb = ea;
                    eb.tstart = side == 0 ? ts_trg[l] : ts_src[l];
                    eb.segbox = side == 0 ? sb_trg[l] : sb_src[l];
                    eb.seg = sege(l);
                    eb.segmax = ea.nbox + (side == 0 ? nt : ns) / eb.seg + 1;
                    eb.grid = (int)std::min<long>(GRIDMAX, eb.segmax);
                    eb.spw = tune.spw[tmof(l) - 1];
                    eb.se = side == 0 ? trg_se : src_se;
                    eb.pts = side == 0 ? xt : xs;
                    eb.ifpgh = side == 0 ? q.ifpghtarg : q.ifpgh;
                    eb.pot = side == 0 ? q.pottarg : q.pot;
                    eb.grad = side == 0 ? q.gradtarg : q.grad;
                    eb.hess = side == 0 ? q.hesstarg : q.hess;
                    eb.perm = side == 0 ? itarg : isrc;
                    eb.npot = side == 0 ? npt : nps;
                    eb.ngrad = side == 0 ? ngt : ngs;
                    eb.nhess = side == 0 ? nht : nhs;
                    tm.ctx(l, &eval_l[l], side == 0 ? 1 : 0);
                    if (pwroute)
                        launch_pw_eval(eb, st);
                    else
                        launch_taylor_eval(eb, st);
                }
            }
        };
        // Every evaluation adds the near field's outputs, so it waits for
        // them.  The splits and the finest level's evaluation on s; each
        // coarse level's evaluation on s_nf after the near field, as soon as
        // the split before it has completed its local expansions (an event
        // on s), so that it overlaps the splits and the finest level.
        // Levels write disjoint points.
        if (!nf_queued) near_field();
        tm.begin(S_SPLIT);
        for (int l = nlevstart; l <= nlevels - 1; ++l) {
            if (!eval_l[l].empty()) {
                W.wait(s_nf, W.mark(s));
                eval_level(l, s_nf);
            }
            tm.ctx(l, &split_l[l], 2);
            launch_split((const int*)D(o_split[l]), (int)split_l[l].size(), d_par, d_cen, d_slot,
                         d_wms, 4 * (nlevels - l - 1), nd, nexp, lc, s, mixed);
        }
        tm.end(S_SPLIT);
        tm.begin(S_EVAL);
        W.wait(s, nf_ready);
        eval_level(nlevels, s);
        tm.end(S_EVAL);
    } else {
        near_field();
    }

    // s_lv's work (the gathers, the coarse levels' formation) is joined
    // already; each join costs a host-side wait of about 10 us
    W.order(s_nf, s);

    HIPCHECK(hipGetLastError());
    tm.finish();
    W.last_valid = true;

    t_mixed = mixed;
    t_stats.assign(6 + nlevels + 1, 0);
    t_stats[0] = npwlevel;
    t_stats[1] = nlmax;
    t_stats[2] = nlevels;
    t_stats[3] = nb;
    t_stats[4] = npw;
    t_stats[5] = nleaves;
    for (int l = 0; l <= nlevels; ++l) t_stats[6 + l] = nlocal[l];
}

}  // namespace

namespace {

// fgt2d_large_delta (fgt2d.f:1616-1733): the whole point set fits in a box
// much smaller than sqrt(delta); one Taylor expansion about the bounding-box
// centre (g2dformlc_vec, g2drouts.f:621) is evaluated at every point.  The
// outputs are overwritten here; the reference adds into the caller's arrays
// in this branch only.
void run_large_delta(Workspace& W, const Call& q, hipStream_t s, int npwlevel, double bs0,
                     const double* cen0, Timer& tm)
{
    const int nd = q.nd;
    int nl = hermite_terms(2, bs0, q.delta, q.eps) + std::max(q.ifpgh, q.ifpghtarg) - 1;
    if (nl < 0) fail(FGTPW_EUNSUPPORTED, "no Taylor order reaches eps for this box");
    const int m = nl + 1;
    const double dsq = 1.0 / std::sqrt(q.delta);
    int nparts = std::max(1, std::min(512, (q.ns + 1023) / 1024));
    double* part = W.get<double>("ld_part", (size_t)nparts * m * m);
    double* L = W.get<double>("ld_local", (size_t)m * m * nd);
    // one box (id 0) holding every point, evaluated in input order:
    // ints {box, pad, 0, ns, 0, nt} give the box list and two int2 ranges
    int hostarr[6] = {0, 0, 0, q.ns, 0, q.nt};
    int* dint = W.get<int>("ld_int", 8);
    double* dcen = W.get<double>("ld_cen", 2);
    char* hp = (char*)W.pinned(128);
    std::memcpy(hp, hostarr, sizeof(hostarr));
    std::memcpy(hp + 64, cen0, 2 * sizeof(double));
    HIPCHECK(hipMemcpyAsync(dint, hp, sizeof(hostarr), hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(dcen, hp + 64, 2 * sizeof(double), hipMemcpyHostToDevice, s));
    tm.begin(S_FORM);
    double* wsd = nullptr;  // dipole weights d n in input order
    if (q.rn) {
        unsigned* id = W.get<unsigned>("ld_iota", q.ns);
        launch_iota(id, q.ns, s);
        wsd = W.get<double>("ld_dipw", 2 * (size_t)nd * q.ns);
        launch_gather_dip(q.rn, q.ds, nd, id, q.ns, wsd, s);
    }
    for (int ind = 0; ind < nd; ++ind)
        launch_formlc(q.src, q.chg, wsd, nd, ind, q.ns, cen0[0], cen0[1], dsq, nl, part, nparts,
                      L + (size_t)ind * m * m, s);
    tm.end(S_FORM);
    tm.begin(S_EVAL);
    if (q.ifpgh > 0) {
        launch_zero(q.pot, (size_t)nd * q.ns, s);
        if (q.ifpgh >= 2) launch_zero(q.grad, 2 * (size_t)nd * q.ns, s);
        if (q.ifpgh >= 3) launch_zero(q.hess, 3 * (size_t)nd * q.ns, s);
    }
    if (q.ifpghtarg > 0 && q.nt > 0) {
        launch_zero(q.pottarg, (size_t)nd * q.nt, s);
        if (q.ifpghtarg >= 2) launch_zero(q.gradtarg, 2 * (size_t)nd * q.nt, s);
        if (q.ifpghtarg >= 3) launch_zero(q.hesstarg, 3 * (size_t)nd * q.nt, s);
    }
    for (int ind = 0; ind < nd; ++ind)
        for (int side = 0; side < 2; ++side) {
            EvalArgs e{};
            e.boxes = dint;
            e.nbox = 1;
            e.cen = dcen;
            e.nd = nd;
            e.ind = ind;
            e.dsq = dsq;
            e.nl = nl;
            e.taylor = L + (size_t)ind * m * m;
            if (side == 0) {
                if (q.ifpgh <= 0) continue;
                e.se = (const int2*)(dint + 2);
                e.pts = q.src;
                e.ifpgh = q.ifpgh;
                e.pot = q.pot; e.grad = q.grad; e.hess = q.hess;
            } else {
                if (q.ifpghtarg <= 0 || q.nt <= 0) continue;
                e.se = (const int2*)(dint + 4);
                e.pts = q.targ;
                e.ifpgh = q.ifpghtarg;
                e.pot = q.pottarg; e.grad = q.gradtarg; e.hess = q.hesstarg;
            }
            launch_taylor_eval(e, s);
        }
    tm.end(S_EVAL);
    HIPCHECK(hipGetLastError());
    tm.finish();
    t_stats.assign(7, 0);
    t_stats[0] = npwlevel;
    t_stats[6] = nl;
}

// fgt3d_large_delta (fgt3d.f:1783-1914): one Taylor expansion of all sources
// about the bounding-box centre, evaluated at every point (outputs
// overwritten).  The expansion uses the z offset for z; the reference's
// g3dformlc_vec loses it (g3drouts.f:803-804).
void run_large_delta3(Workspace& W, const Call& q, hipStream_t s, int npwlevel, double bs0,
                      const double* cen0, Timer& tm)
{
    const int nd = q.nd;
    int nl = hermite_terms(3, bs0, q.delta, q.eps) + std::max(q.ifpgh, q.ifpghtarg) - 1;
    if (nl < 0) fail(FGTPW_EUNSUPPORTED, "no Taylor order reaches eps for this box");
    const int m = nl + 1;
    const size_t m3 = (size_t)m * m * m;
    const double dsq = 1.0 / std::sqrt(q.delta);
    int nparts = std::max(1, std::min(256, (q.ns + 4095) / 4096));
    double* part = W.get<double>("ld3_part", (size_t)nparts * m3);
    double* L = W.get<double>("ld3_local", m3 * nd);
    // one box (id 0) holding every point in input order: ints {box, pad,
    // 0, ns, 0, nt, 0, 1} give the box list, two int2 ranges and a scan
    int hostarr[8] = {0, 0, 0, q.ns, 0, q.nt, 0, 0};
    int* dint = W.get<int>("ld3_int", 16);
    double* dcen = W.get<double>("ld3_cen", 4);
    char* hp = (char*)W.pinned(256);
    std::memcpy(hp, hostarr, sizeof(hostarr));
    std::memcpy(hp + 64, cen0, 3 * sizeof(double));
    HIPCHECK(hipMemcpyAsync(dint, hp, sizeof(hostarr), hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(dcen, hp + 64, 3 * sizeof(double), hipMemcpyHostToDevice, s));
    tm.begin(S_FORM);
    double* wsd = nullptr;  // dipole weights d n in input order
    if (q.rn) {
        unsigned* id = W.get<unsigned>("ld_iota", q.ns);
        launch_iota(id, q.ns, s);
        wsd = W.get<double>("ld_dipw", 3 * (size_t)nd * q.ns);
        launch_gather_dip3(q.rn, q.ds, nd, id, q.ns, wsd, s);
    }
    for (int ind = 0; ind < nd; ++ind)
        launch_formlc3(q.src, q.chg, wsd, nd, ind, q.ns, cen0, dsq, nl, part, nparts, L + ind * m3,
                       s);
    tm.end(S_FORM);
    tm.begin(S_EVAL);
    if (q.ifpgh > 0) {
        launch_zero(q.pot, (size_t)nd * q.ns, s);
        if (q.ifpgh >= 2) launch_zero(q.grad, 3 * (size_t)nd * q.ns, s);
        if (q.ifpgh >= 3) launch_zero(q.hess, 6 * (size_t)nd * q.ns, s);
    }
    if (q.ifpghtarg > 0 && q.nt > 0) {
        launch_zero(q.pottarg, (size_t)nd * q.nt, s);
        if (q.ifpghtarg >= 2) launch_zero(q.gradtarg, 3 * (size_t)nd * q.nt, s);
        if (q.ifpghtarg >= 3) launch_zero(q.hesstarg, 6 * (size_t)nd * q.nt, s);
    }
    const int SEG = 256;
    for (int side = 0; side < 2; ++side) {
        const int n = side == 0 ? q.ns : q.nt;
        const int ifp = side == 0 ? q.ifpgh : q.ifpghtarg;
        if (ifp <= 0 || n <= 0) continue;
        // segment scan of the one box: {0, ceil(n/SEG)}
        int* tst = W.get<int>(side == 0 ? "ld3_ts0" : "ld3_ts1", 2);
        int hs[2] = {0, (n + SEG - 1) / SEG};
        int* hp2 = (int*)((char*)W.pinned(256) + 128 + 16 * side);
        std::memcpy(hp2, hs, sizeof(hs));
        HIPCHECK(hipMemcpyAsync(tst, hp2, sizeof(hs), hipMemcpyHostToDevice, s));
        for (int ind = 0; ind < nd; ++ind) {
            Eval3Args e{};
            e.boxes = dint;
            e.nbox = 1;
            e.cen = dcen;
            e.se = (const int2*)(dint + (side == 0 ? 2 : 4));
            e.pts = side == 0 ? q.src : q.targ;
            e.nd = nd;
            e.ind = ind;
            e.dsq = dsq;
            e.ifpgh = ifp;
            e.nl = nl;
            e.taylor = L + ind * m3;
            e.pot = side == 0 ? q.pot : q.pottarg;
            e.grad = side == 0 ? q.grad : q.gradtarg;
            e.hess = side == 0 ? q.hess : q.hesstarg;
            e.tstart = tst;
            e.seg = SEG;
            e.grid = std::min(4096, hs[1]);
            launch_taylor_eval3(e, s);
        }
    }
    tm.end(S_EVAL);
    HIPCHECK(hipStreamSynchronize(s));  // the pinned staging above is reused by the next call
    HIPCHECK(hipGetLastError());
    tm.finish();
    t_stats.assign(7, 0);
    t_stats[0] = npwlevel;
    t_stats[6] = nl;
}

// The 3D transform (fgt3d, fgt3d.f:19-480, and fgt3dmain, :483-1530) on one
// stream: octree from Morton keys, host level restriction and lists, then
// the six steps.  Differences from run2d follow the reference's 3D code:
// ndiv0 = 40, ndiv = 100, nlmax at nlocal <= 15, formation and evaluation
// routes chosen by (p+1)/npw and the box's point count (fgt3d.f:971-983,
// :1297-1310), listpw without leaf colleagues, near field with
// dmax = ln(1/eps) delta and the 3D modified list1 (tree_routs3d.f:511-620).
void run3d(Workspace& W, const Call& q, hipStream_t s)
{
    t_mixed = false;
    if (!q.chg && !q.rn) {
        zero_outputs(q, 3, s);
        return;
    }
    const int ifdipole = q.rn ? 1 : 0;
    Timer tm(W, s);
    const int ns = q.ns, nt = q.nt, nd = q.nd;
    tm.begin(S_BBOX);
    int nblk = (int)std::min<long>(1024, std::max<long>(1, ((long)ns + nt + 255) / 256));
    double* part = W.get<double>("bbox_part", 7 * (size_t)nblk);
    double* bbd = W.get<double>("bbox", 8);
    launch_bbox3(q.src, ns, q.targ, nt, part, nblk, bbd, s);
    tm.end(S_BBOX);
    double* hb = (double*)W.pinned(64);
    copy_to_host(hb, bbd, 7 * sizeof(double), s);
    if (hb[6] != 0) fail(FGTPW_EINVAL, "non-finite source or target coordinates");
    double bs0 = hb[1] - hb[0];
    const double sy = hb[3] - hb[2], sz = hb[5] - hb[4];
    if (sy > bs0) bs0 = sy;
    if (sz > bs0) bs0 = sz;
    const double cen0[3] = {(hb[0] + hb[1]) / 2, (hb[2] + hb[3]) / 2, (hb[4] + hb[5]) / 2};
    const double dc = std::sqrt(q.delta * std::log(1.0 / q.eps));
    int npwlevel;
    if (bs0 > 0) {
        double dl = std::log(bs0 / dc) / std::log(2.0);
        if (dl > 20) fail(FGTPW_EUNSUPPORTED, "box too large relative to sqrt(delta): octree deeper than 21 levels");
        npwlevel = (int)dl;
        if (bs0 >= dc) {
            bs0 = (double)(1 << (npwlevel + 1)) * dc;
            npwlevel += 1;
        }
    } else {
        npwlevel = -1000;
    }
    if (npwlevel < -2) {
        W.last_valid = false;
        run_large_delta3(W, q, s, npwlevel, bs0, cen0, tm);
        return;
    }
    const int nlevstart = std::max(npwlevel, 0);
    int nlmax = npwlevel + 4;
    {
        double bsize = bs0 / (double)(1 << nlevstart);
        for (int i = nlevstart; i <= 100; ++i) {
            if (hermite_terms(3, bsize, q.delta, q.eps) <= 15) {
                nlmax = i;
                break;
            }
            bsize = bsize / 2;
        }
    }
    if (nlmax > 21) fail(FGTPW_EUNSUPPORTED, "octree deeper than 21 levels");

    // ---------------- keys, sort, leaves of the adaptive tree ----------------
    tm.begin(S_KEYS);
    using u64 = unsigned long long;
    const int nmx = std::max(ns, nt);
    u64* ks = W.get<u64>("keys_s", ns);
    u64* ks2 = W.get<u64>("keys_s2", ns);
    unsigned* ix = W.get<unsigned>("idx_s", nmx);
    // the library's LSD sort numbers the points itself in its first pass:
    // no index array is written and read
    unsigned* isrc = W.get<unsigned>("isrc", ns);
    size_t tmpb = std::max({sort_pairs_u64_temp(nmx), sort_pairs_u32_temp(nmx), leaf_records_temp(ns)});
    void* tmp = W.get<char>("sort_tmp", tmpb);
    const int keybits = std::max(3, 3 * nlmax);
    const bool k32 = keybits <= 31;  // keys fit 32 bits
    unsigned* ix32 = keybits <= 30 && (tuning().sort == 4 || tuning().sort == 5) ? nullptr : ix;
    if (k32)
        launch_keys3d(q.src, ns, cen0, bs0, nlmax, (unsigned*)ks, ix32, s);
    else
        launch_keys3d(q.src, ns, cen0, bs0, nlmax, ks, ix, s);
    tm.bracket(KID_SORT, false, -1, ns);
    if (k32)
        sort_pairs_u32(tmp, tmpb, (unsigned*)ks, (unsigned*)ks2, ix32, isrc, ns, keybits, s, tuning().sort);
    else
        sort_pairs_u64(tmp, tmpb, ks, ks2, ix, isrc, ns, keybits, s);
    tm.bracket(KID_SORT, true, -1, ns);
    tm.end(S_KEYS);
    tm.begin(S_LEAF);
    int* lvl = W.get<int>("leaf_level", ns);
    unsigned char* flg = W.get<unsigned char>("leaf_flag", ns);
    // the leaf count (first word) and the leaf records (from byte 256) in
    // one buffer: one read-back, sized by the last call's count
    char* rec = W.get<char>("leaf_runs", 256 + sizeof(LeafRun) * (size_t)ns);
    int* nsel = (int*)rec;
    LeafRun* runs = (LeafRun*)(rec + 256);
    unsigned* bcnt = (unsigned*)tmp;  // free after the sort
    if (k32)
        launch_leafdepth((const unsigned*)ks2, ns, nlmax, npwlevel, 40, 100, lvl, flg, bcnt, s, 3);
    else
        launch_leafdepth(ks2, ns, nlmax, npwlevel, 40, 100, lvl, flg, bcnt, s, 3);
    if (k32)
        launch_leaf_records(flg, bcnt, ns, (const unsigned*)ks2, lvl, nlmax, nsel, runs, s, 3);
    else
        launch_leaf_records(flg, bcnt, ns, ks2, lvl, nlmax, nsel, runs, s, 3);
    tm.end(S_LEAF);
    const size_t guess = std::min<size_t>(ns, std::max<size_t>(4096, W.leaf_guess + W.leaf_guess / 4));
    char* hl = (char*)W.pinned(256 + sizeof(LeafRun) * guess);
    copy_to_host(hl, rec, 256 + sizeof(LeafRun) * guess, s);
    const int nleaf0 = *(const int*)hl;
    W.leaf_guess = nleaf0;
    std::vector<LeafRun> leaves(nleaf0);
    std::memcpy(leaves.data(), hl + 256, sizeof(LeafRun) * std::min<size_t>(nleaf0, guess));
    if ((size_t)nleaf0 > guess) {
        const size_t rest = sizeof(LeafRun) * ((size_t)nleaf0 - guess);
        void* hp = W.pinned2(rest);
        copy_to_host(hp, runs + guess, rest, s);
        std::memcpy(leaves.data() + guess, hp, rest);
    }

    // points and charges into Morton order on a side stream, queued once the
    // leaves are back, overlapping the host's tree
    hipStream_t s_g = g_serial ? s : W.aux[0];
    W.order(s, s_g);
    tm.begin(S_GATHER, s_g);
    double* xs = W.get<double>("xs", 3 * (size_t)ns);
    double* qs = W.get<double>("qs", (size_t)nd * ns);
    tm.ctx(-1, nullptr, -1, ns);
    if (q.chg)
        launch_gather3(q.src, q.chg, nd, isrc, ns, xs, qs, s_g);
    else {
        launch_gather3(q.src, nullptr, nd, isrc, ns, xs, nullptr, s_g);
        launch_zero(qs, (size_t)nd * ns, s_g);
    }
    double* wsd = nullptr;  // dipole weights d n in sorted order
    if (q.rn) {
        wsd = W.get<double>("dip_w", 3 * (size_t)nd * ns);
        launch_gather_dip3(q.rn, q.ds, nd, isrc, ns, wsd, s_g);
    }
    tm.end(S_GATHER, s_g);
    tm.begin(S_ASSIGN);
    unsigned* itarg = nullptr;
    u64* kt2 = nullptr;
    double* xt = nullptr;
    if (nt > 0) {
        itarg = W.get<unsigned>("itarg", nt);
        xt = W.get<double>("xt", 3 * (size_t)nt);
        u64* kt = W.get<u64>("keys_t", nt);
        kt2 = W.get<u64>("keys_t2", nt);
        if (k32)
            launch_keys3d(q.targ, nt, cen0, bs0, nlmax, (unsigned*)kt, ix32, s);
        else
            launch_keys3d(q.targ, nt, cen0, bs0, nlmax, kt, ix, s);
        tm.bracket(KID_SORT, false, -2, nt);
        if (k32)
            sort_pairs_u32(tmp, tmpb, (unsigned*)kt, (unsigned*)kt2, ix32, itarg, nt, keybits, s, tuning().sort);
        else
            sort_pairs_u64(tmp, tmpb, kt, kt2, ix, itarg, nt, keybits, s);
        tm.bracket(KID_SORT, true, -2, nt);
        tm.ctx(-1, nullptr, -1, nt);
        launch_gather3(q.targ, nullptr, 0, itarg, nt, xt, nullptr, s);
    }
    tm.end(S_ASSIGN);

    // ---------------- host: tree, lists, tables ----------------
    auto th0 = std::chrono::steady_clock::now();
    W.last_valid = false;
    Tree3D& T = W.last3;
    build_tree<3>(T, leaves, nlmax, npwlevel, bs0, cen0);
    tm.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - th0).count();
    W.last_ns = ns;
    W.last_nt = nt;
    W.last_dim = 3;
    const int nb = T.nboxes(), nlevels = T.nlevels, nleaves = (int)T.leaf_of_rank.size();
    const bool use_src = q.ifpgh > 0, use_trg = q.ifpghtarg > 0 && nt > 0;
    const int nadd = ifdipole + std::max(q.ifpgh, q.ifpghtarg);  // fgt3d.f:355
    std::vector<int> nlocal(nlevels + 1, 0);
    int ntermmax = 0;
    for (int i = nlevstart; i <= nlevels; ++i) {
        nlocal[i] = hermite_terms(3, T.bsz[i], q.delta, q.eps);
        if (ifdipole || q.ifpgh >= 2 || q.ifpghtarg >= 2) nlocal[i] += nadd;
        ntermmax = std::max(ntermmax, nlocal[i]);
    }
    double pmax = 0;
    int npw = 0;
    {
        double bsize = 2 * bs0 / std::pow(2.0, nlevstart);
        double pweps = nadd > 2 ? q.eps / 100 : q.eps;
        planewave_terms(bsize, q.delta, pweps, &pmax, &npw);
    }
    const bool expansions = npwlevel < nlevels;
    const int half = npw / 2;
    const long nexp = (long)npw * npw * half;
    const double dsq = 1.0 / std::sqrt(q.delta);
    if (expansions && npw > 64) fail(FGTPW_EUNSUPPORTED, "3D plane-wave length above 64");
    // Route per level (fgt3d.f:971-983, :1297-1310): with (p+1)^3 >= 3 npw^3
    // the reference forms and evaluates the plane waves directly for boxes
    // under 10^6 points, else through Hermite / Taylor expansions.  thresh:
    // 0 = Hermite/Taylor for every box, 10^6 = the reference's switch,
    // INT_MAX = plane waves for every box (orders above 63, where the
    // Hermite/Taylor kernels stop; differs from the reference only for boxes
    // of 10^6 points or more).  The orders above 31 run on 64-wide register
    // tiles.
    std::vector<int> form_thresh(nlevels + 1, 0), eval_thresh(nlevels + 1, 0);
    if (expansions)
        for (int l = nlevstart; l <= nlevels; ++l) {
            const double r1 = (nlocal[l] + 1.0) / npw;
            const bool pw = r1 * r1 * r1 >= 3;
            const bool big = nlocal[l] + 1 > 64;
            if (big && !pw)
                fail(FGTPW_EUNSUPPORTED, "3D expansion order above 63 (eps too small for this box size)");
            form_thresh[l] = eval_thresh[l] = pw ? (big ? INT_MAX : 1000000) : 0;
        }

    // Expansion storage.  The reference keeps a multipole and a local
    // expansion for every box from the cutoff level down (fgt3d.f:388-400,
    // g3dmpalloc fgt3d.f:1685).  Here persistent slots hold only what lives
    // across levels -- multipole expansions of the merge parents (and of
    // leaves formed at the top level), local expansions of the top level and
    // of non-leaf boxes -- and one chunk region is reused by the leaves: on
    // the way up a chunk of leaves is formed there and merged into its
    // parents, on the way down a chunk is split into there and evaluated.
    // All of it lies in one buffer: the multipole table addresses it from
    // slot 0, the local table from slot n_mp, the chunk region starts at slot
    // n_mp + n_loc (multipole numbering) = n_loc (local numbering).
    std::vector<int>& slot_mp = W.h_slot;
    std::vector<int>& slot_loc = W.h_slot2;
    slot_mp.assign(nb, -1);
    slot_loc.assign(nb, -1);
    int n_mp = 0, n_loc = 0;
    std::vector<std::vector<int>> form_l(nlevels + 1), merge_l(nlevels + 1), split_l(nlevels + 1),
        eval_l(nlevels + 1);
    std::vector<int> m2l_l;
    // chunks: upward (level of the formed leaves, first and last parent in
    // merge_l[level - 1], first and last leaf in form_l[level]); downward
    // (level, first and last box in eval_l[level])
    struct UpChunk { int l, p0, p1, f0, f1; };
    struct DownChunk { int l, e0, e1; };
    std::vector<UpChunk> up;
    std::vector<DownChunk> down;
    // levels whose leaves all take the Taylor route through a kernel that
    // fuses the parent -> leaf split: no split launch, no chunk slots
    std::vector<char> fused(nlevels + 1, 0);
    const size_t per = 2 * (size_t)nexp * nd;  // doubles per slot
    size_t ntay = 1;                             // Taylor buffer, doubles
    int ncap = 0;                                // chunk region, slots
    if (expansions) {
        std::vector<char> iflocal(nb, 0);
        if (npwlevel >= 0 && npwlevel <= nlevels)
            for (int b = T.laddr[2 * npwlevel]; b <= T.laddr[2 * npwlevel + 1]; ++b) {
                if (T.nch[b] != 0) continue;
                for (int i = 0; i < T.ncol[b]; ++i)
                    if (T.nch[T.col[27 * b + i]] > 0) { iflocal[b] = 1; break; }
            }
        const int lform = std::max(npwlevel + 1, 0);
        for (int l = nlevstart; l <= nlevels; ++l)
            for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; ++b) {
                if (T.nch[b] > 0 && l <= nlevels - 1) {
                    merge_l[l].push_back(b);
                    slot_mp[b] = n_mp++;
                    for (int j = 0; j < 8; ++j) {
                        const int c = T.ch[8 * b + j];
                        if (c >= 0 && T.nch[c] > 0) split_l[l].push_back(c);
                    }
                }
                const bool evaluated = T.nch[b] == 0 && !(l == npwlevel && !iflocal[b]);
                if (l == nlevstart) {
                    // local expansions only where they are read: non-leaf
                    // boxes (split to their children) and evaluated leaves
                    if (T.nch[b] > 0 || evaluated) m2l_l.push_back(b);
                    if (T.nch[b] == 0 && l >= lform) {
                        form_l[l].push_back(b);
                        slot_mp[b] = n_mp++;
                    }
                }
                if ((l == nlevstart && evaluated) || T.nch[b] > 0) slot_loc[b] = n_loc++;
                if (evaluated) eval_l[l].push_back(b);
            }
        // leaves below the top level, grouped by parent
        for (int l = nlevstart + 1; l <= nlevels; ++l)
            for (int p : merge_l[l - 1])
                for (int j = 0; j < 8; ++j) {
                    const int c = T.ch[8 * p + j];
                    if (c >= 0 && T.nch[c] == 0) form_l[l].push_back(c);
                }
        // memory: persistent slots, then the chunk region and the Taylor
        // buffer as large as the largest level needs, or as the device allows
        size_t need_c = 0, need_tay = 1, m3max = 1;
        for (int l = nlevstart + 1; l <= nlevels; ++l) {
            fused[l] = eval_thresh[l] == 0 && pw2local3_fusable(nlocal[l] + 1, npw);
            need_c = std::max({need_c, form_l[l].size(), fused[l] ? (size_t)0 : eval_l[l].size()});
        }
        for (int l = nlevstart; l <= nlevels; ++l) {
            if (eval_thresh[l] == INT_MAX || eval_l[l].empty()) continue;
            const int m = nlocal[l] + 1;
            const size_t st = taylor3_stride(m, m <= 32);
            m3max = std::max(m3max, st);
            need_tay = std::max(need_tay, eval_l[l].size() * st);
        }
        size_t fr = 0, tot = 0;
        HIPCHECK(hipMemGetInfo(&fr, &tot));
        const double G = 1024.0 * 1024.0 * 1024.0;
        const double avail = (double)fr + W.held("exp3") + W.held("taylor3") - 1.0 * G - 0.02 * tot;
        const double pers = 8.0 * per * ((double)n_mp + n_loc);
        const double full = pers + 8.0 * per * need_c + 8.0 * need_tay;
        size_t cap_c = need_c, cap_tay = need_tay;
        if (full > avail) {
            const double rest = avail - pers;
            if (rest < 8.0 * per * std::min<size_t>(need_c, 8) + 8.0 * m3max) {
                char msg[200];
                std::snprintf(msg, sizeof msg,
                              "3D expansions need %.1f GiB of device memory at least, %.1f GiB free",
                              (pers + 8.0 * per * 8 + 8.0 * m3max) / G, avail / G);
                fail(FGTPW_ENOMEM, msg);
            }
            cap_tay = std::min(need_tay, std::max(m3max, (size_t)(0.1 * rest / 8)));
            cap_c = std::min(need_c, (size_t)((rest - 8.0 * cap_tay) / (8.0 * per)));
            cap_c = std::max(cap_c, std::min<size_t>(need_c, 8));
        }
        // FGTPW_MAX_CHUNK caps the boxes per chunk (at least 8 on the way up,
        // where chunks hold whole parents): less device memory for more launches
        size_t cap_env = 0;
        if (const char* e = std::getenv("FGTPW_MAX_CHUNK")) cap_env = std::max(0L, std::atol(e));
        if (cap_env > 0) cap_c = std::min(cap_c, std::max<size_t>(cap_env, std::min<size_t>(need_c, 8)));
        ncap = (int)cap_c;
        ntay = cap_tay;
        // upward chunks: whole parents, at most ncap leaves
        for (int l = nlevstart + 1; l <= nlevels; ++l) {
            const std::vector<int>& P = merge_l[l - 1];
            int p0 = 0, f0 = 0, f = 0;
            for (int i = 0; i < (int)P.size(); ++i) {
                int k = 0;
                for (int j = 0; j < 8; ++j) {
                    const int c = T.ch[8 * P[i] + j];
                    k += c >= 0 && T.nch[c] == 0;
                }
                if (f - f0 + k > ncap) {
                    up.push_back({l, p0, i, f0, f});
                    p0 = i;
                    f0 = f;
                }
                for (int j = 0; j < k; ++j, ++f) slot_mp[form_l[l][f]] = n_mp + n_loc + (f - f0);
            }
            if (!P.empty()) up.push_back({l, p0, (int)P.size(), f0, f});
        }
        // downward chunks: at most ncap leaves below the top level, and at
        // most the Taylor buffer's worth of boxes
        for (int l = nlevstart; l <= nlevels; ++l) {
            const size_t n = eval_l[l].size();
            if (n == 0) continue;
            const int m = nlocal[l] + 1;
            const size_t st = taylor3_stride(m, m <= 32);
            size_t cmax = eval_thresh[l] == INT_MAX ? n : std::max<size_t>(1, ntay / st);
            if (l > nlevstart && !fused[l]) cmax = std::min<size_t>(cmax, ncap);
            if (cap_env > 0) cmax = std::min(cmax, cap_env);
            for (size_t e0 = 0; e0 < n; e0 += cmax) {
                const size_t e1 = std::min(n, e0 + cmax);
                down.push_back({l, (int)e0, (int)e1});
                if (l > nlevstart && !fused[l])
                    for (size_t i = e0; i < e1; ++i) slot_loc[eval_l[l][i]] = n_loc + (int)(i - e0);
            }
        }
    }
    Pack pk(W.h_plan, !W.arena);
    size_t o_child = pk.add(T.ch), o_par = pk.add(T.par), o_cen = pk.add(T.cen);
    // colleague lists are read on the device by m2l only, for the boxes of
    // the top level (a contiguous range): only that slice travels
    const int colb0 = expansions ? T.laddr[2 * nlevstart] : 0;
    const int colb1 = expansions ? T.laddr[2 * nlevstart + 1] : -1;
    size_t o_ncol = pk.add(T.ncol);
    size_t o_col = pk.add(T.col.data() + 27 * (size_t)colb0, 27 * (size_t)(colb1 - colb0 + 1));
    size_t o_lkey = pack_leaf_keys(pk, W, T, nlmax, k32);
    size_t o_rlo = pk.add(T.rank_lo), o_rhi = pk.add(T.rank_hi);
    size_t o_smp = pk.add(slot_mp), o_sloc = pk.add(slot_loc);
    size_t o_nch = pk.add(T.nch);
    std::vector<size_t> o_form(nlevels + 1), o_merge(nlevels + 1), o_split(nlevels + 1),
        o_eval(nlevels + 1);
    size_t o_m2l = 0;
    if (expansions) {
        for (int l = 0; l <= nlevels; ++l) {
            o_form[l] = pk.add(form_l[l]);
            o_merge[l] = pk.add(merge_l[l]);
            o_split[l] = pk.add(split_l[l]);
            o_eval[l] = pk.add(eval_l[l]);
        }
        o_m2l = pk.add(m2l_l);
    }
    const Tables3D& tab = expansions
                              ? tables3d(W, npw, pmax, ntermmax, nlevels - nlevstart,
                                         T.bsz[nlevels] / std::sqrt(q.delta) / 2,
                                         T.bsz[nlevstart] / std::sqrt(q.delta), s)
                              : W.tab3;

    // near field: modified list1 of the source leaves at or above the cutoff
    // (compute_modified_list1, tree_routs3d.f:511-620, FP64 distance tests
    // as there), inverted to target units as in 2D.  Built once steps 1-5
    // are queued: the near field runs after them, so this host work overlaps
    // the GPU's.
    const int nlevend = std::min(npwlevel, nlevels);
    std::vector<int>& unit_ids = W.h_units;
    std::vector<int>& unit_ptr = W.h_uptr;
    std::vector<int>& unit_list = W.h_ulist;
    auto near_field_lists = [&]() {
    unit_ids.clear();
    unit_ptr.assign(1, 0);
    unit_list.clear();
    if (nlevend >= 0) {
        const int MNL = 27 + 27 * 8 + 27;
        std::vector<int>& nl1 = W.h_nl1;
        std::vector<int>& l1 = W.h_l1;
        nl1.assign(nb, 0);
        if (l1.size() < (size_t)MNL * nb) l1.resize((size_t)MNL * nb);
        if (T.nch[0] == 0) { nl1[0] = 1; l1[0] = 0; }
        auto touch = [&](int a, int b, double tol) {
            for (int d = 0; d < 3; ++d)
                if (!(std::fabs(T.cen[3 * a + d] - T.cen[3 * b + d]) < tol)) return false;
            return true;
        };
        for (int l = 1; l <= nlevend; ++l)
            for (int b = T.laddr[2 * l]; b <= T.laddr[2 * l + 1]; ++b) {
                if (T.nch[b] != 0) continue;
                int* out = &l1[(size_t)MNL * b];
                for (int i = 0; i < T.ncol[b]; ++i) {
                    int jb = T.col[27 * b + i];
                    if (l >= npwlevel || T.nch[jb] == 0) {
                        out[nl1[b]++] = jb;
                    } else {
                        const double tol = 1.05 * (T.bsz[l] + T.bsz[l + 1]) / 2.0;
                        for (int j = 0; j < 8; ++j) {
                            int kb = T.ch[8 * jb + j];
                            if (kb >= 0 && touch(kb, b, tol)) out[nl1[b]++] = kb;
                        }
                    }
                }
                const int dad = T.par[b];
                const double tol = 1.05 * (T.bsz[l] + T.bsz[l - 1]) / 2.0;
                for (int i = 0; i < T.ncol[dad]; ++i) {
                    int jb = T.col[27 * dad + i];
                    if (T.nch[jb] == 0 && touch(jb, b, tol)) out[nl1[b]++] = jb;
                }
            }
        std::vector<int>& sp = W.h_sp;
        std::vector<int>& sl = W.h_sl;
        sp.assign(nb + 1, 0);
        for (int l = 0; l <= nlevend; ++l)
            for (int sb = T.laddr[2 * l]; sb <= T.laddr[2 * l + 1]; ++sb) {
                if (T.nch[sb] != 0) continue;
                for (int i = 0; i < nl1[sb]; ++i) ++sp[l1[(size_t)MNL * sb + i] + 1];
            }
        for (int b = 0; b < nb; ++b) sp[b + 1] += sp[b];
        sl.resize(sp[nb]);
        std::vector<int>& fill = W.h_fill;
        fill.assign(sp.begin(), sp.end() - 1);
        for (int l = 0; l <= nlevend; ++l)
            for (int sb = T.laddr[2 * l]; sb <= T.laddr[2 * l + 1]; ++sb) {
                if (T.nch[sb] != 0) continue;
                for (int i = 0; i < nl1[sb]; ++i) sl[fill[l1[(size_t)MNL * sb + i]]++] = sb;
            }
        for (int b = 0; b < nb; ++b) {
            bool unit = (T.nch[b] == 0 && T.lev[b] <= npwlevel) || T.lev[b] == npwlevel;
            if (!unit) continue;
            size_t before = unit_list.size();
            for (int a = b; a >= 0; a = T.par[a])
                unit_list.insert(unit_list.end(), sl.begin() + sp[a], sl.begin() + sp[a + 1]);
            if (unit_list.size() == before) continue;
            unit_ids.push_back(b);
            unit_ptr.push_back((int)unit_list.size());
        }
    }
    };
    auto th1 = std::chrono::steady_clock::now();
    tm.host_ms = std::chrono::duration<double, std::milli>(th1 - th0).count();

    char* dplan = W.get<char>("plan", pk.size());
    {
        HIPCHECK(hipMemcpyAsync(dplan, pk.data(), pk.size(), hipMemcpyHostToDevice, s));
    }
    W.order(s_g, s);  // the sorted points, charges and dipole weights
    auto D = [&](size_t off) { return (void*)(dplan + off); };
    const int* d_child = (const int*)D(o_child);
    const int* d_par = (const int*)D(o_par);
    const double* d_cen = (const double*)D(o_cen);
    const int* d_ncol = (const int*)D(o_ncol);
    const int* d_col = (const int*)D(o_col) - 27 * (ptrdiff_t)colb0;  // indexed by box
    const void* d_lkey = D(o_lkey);
    const int* d_rlo = (const int*)D(o_rlo);
    const int* d_rhi = (const int*)D(o_rhi);
    const int* d_smp = (const int*)D(o_smp);
    const int* d_sloc = (const int*)D(o_sloc);
    const int* d_nch = (const int*)D(o_nch);

    // ---------------- point ranges of the boxes ----------------
    tm.begin(S_RANGES);
    // sources and targets in one launch each
    int* ls = W.get<int>("leafstart", 2 * ((size_t)nleaves + 1));
    int* ls2 = ls + nleaves + 1;
    int2* src_se = W.get<int2>("src_se", nb);
    int2* trg_se = W.get<int2>("trg_se", nb);
    const bool tg = nt > 0;
    if (k32)
        launch_leafstart_search((const unsigned*)ks2, ns, tg ? (const unsigned*)kt2 : nullptr, nt,
                                (const unsigned*)d_lkey, nleaves, ls, ls2, s);
    else
        launch_leafstart_search(ks2, ns, tg ? kt2 : nullptr, nt, (const u64*)d_lkey, nleaves, ls,
                                ls2, s);
    launch_box_ranges(d_rlo, d_rhi, nb, ls, src_se, ls2, tg ? trg_se : nullptr, s);
    tm.src_se = src_se, tm.trg_se = trg_se, tm.nb = nb, tm.nd = nd;
    if (!tg) launch_zero((double*)trg_se, (size_t)nb, s);  // int2 pairs, 8 bytes each
    tm.end(S_RANGES);

    const int ws_ = use_src ? ns : 0, wt_ = use_trg ? nt : 0;
    double *ps = nullptr, *gs = nullptr, *hs = nullptr, *pt = nullptr, *gt = nullptr, *ht = nullptr;
    if (use_src) {
        ps = W.get<double>("pot_s", (size_t)nd * ws_);
        launch_zero(ps, (size_t)nd * ws_, s);
        if (q.ifpgh >= 2) { gs = W.get<double>("grad_s", 3 * (size_t)nd * ws_); launch_zero(gs, 3 * (size_t)nd * ws_, s); }
        if (q.ifpgh >= 3) { hs = W.get<double>("hess_s", 6 * (size_t)nd * ws_); launch_zero(hs, 6 * (size_t)nd * ws_, s); }
    }
    if (use_trg) {
        pt = W.get<double>("pot_t", (size_t)nd * wt_);
        launch_zero(pt, (size_t)nd * wt_, s);
        if (q.ifpghtarg >= 2) { gt = W.get<double>("grad_t", 3 * (size_t)nd * wt_); launch_zero(gt, 3 * (size_t)nd * wt_, s); }
        if (q.ifpghtarg >= 3) { ht = W.get<double>("hess_t", 6 * (size_t)nd * wt_); launch_zero(ht, 6 * (size_t)nd * wt_, s); }
    }

    // ---------------- expansions: steps 1-5 ----------------
    if (expansions) {
        double* xb = W.get<double>("exp3", per * ((size_t)n_mp + n_loc + ncap));
        double* mp = xb;                       // multipole slots
        double* lc = xb + per * (size_t)n_mp;  // local slots (the chunk region follows)
        const int GRID3 = 1024, SEGE = 256, GRIDMAX = 4096;
        // segment scans of the evaluation lists
        std::vector<int*> ts_trg(nlevels + 1, nullptr), ts_src(nlevels + 1, nullptr);
        {
            size_t ntot = 0;
            for (int l = 0; l <= nlevels; ++l) ntot += 2 * eval_l[l].size() + 2;
            int* pool = W.get<int>("tstart", ntot);
            SegLists L{};
            auto add = [&](const std::vector<int>& v, size_t off, const int2* se) {
                int* t = pool;
                pool += v.size() + 1;
                if (L.count == 32) {
                    launch_seg_scan(L, s);
                    L.count = 0;
                }
                L.l[L.count++] = SegList{(const int*)D(off), (int)v.size(), se, SEGE, t, nullptr};
                return t;
            };
            for (int l = nlevstart; l <= nlevels; ++l) {
                if (eval_l[l].empty()) continue;
                if (use_trg) ts_trg[l] = add(eval_l[l], o_eval[l], trg_se);
                if (use_src) ts_src[l] = add(eval_l[l], o_eval[l], src_se);
            }
            if (L.count) launch_seg_scan(L, s);
        }
        // formation scratch per block: m^3 + 2 m^2 (npw/2), or the
        // matrix-core route's padded moments and z contractions
        long sstride = 0;
        for (int l = std::max(npwlevel + 1, 0); l <= nlevels; ++l) {
            if (form_thresh[l] == INT_MAX) continue;
            long m = nlocal[l] + 1;
            sstride = std::max(sstride, m * m * m + 2 * m * m * half);
            if (m <= 32 && npw <= 48) sstride = std::max(sstride, form3_mfma_scratch((int)m, npw));
        }
        double* fscr = W.get<double>("form3_scratch", (size_t)GRID3 * std::max(sstride, 1L));
        double* taylor = W.get<double>("taylor3", ntay);
        auto form_level = [&](int l, int f0, int f1) {
            for (int ind = 0; ind < nd; ++ind) {
                Form3Args fa{};
                fa.boxes = (const int*)D(o_form[l]) + f0;
                fa.nbox = f1 - f0;
                fa.cen = d_cen;
                fa.src_se = src_se;
                fa.slot = d_smp;
                fa.xs = xs;
                fa.qs = qs;
                fa.nd = nd;
                fa.ind = ind;
                fa.dsq = dsq;
                fa.p = nlocal[l];
                fa.npw = npw;
                fa.coef = tab.coef;
                fa.cstride = npw;
                fa.ws = tab.ws;
                fa.ts = tab.ts;
                fa.mp = mp;
                fa.scratch = fscr;
                fa.sstride = sstride;
                fa.grid = std::min(GRID3, fa.nbox);
                fa.thresh = form_thresh[l];
                fa.dipw = wsd;
                tm.ctx(l, &form_l[l], 0, 0, f0, f1);
                if (fa.nbox <= 0) continue;
                if (fa.thresh != INT_MAX) launch_form3_hermite(fa, s);
                if (fa.thresh > 0) launch_form3_pw(fa, s);
            }
        };
        // steps 1-2: leaves of the top level, then chunks of leaves formed
        // and merged into their parents, finest level first
        tm.begin(S_FORM);
        if (!form_l[nlevstart].empty()) form_level(nlevstart, 0, (int)form_l[nlevstart].size());
        tm.end(S_FORM);
        for (int ci = (int)up.size() - 1; ci >= 0;) {
            const int l = up[ci].l;
            int cj = ci;
            while (cj >= 0 && up[cj].l == l) --cj;
            // chunks of level l are up[cj+1 .. ci]
            for (int k = cj + 1; k <= ci; ++k) {
                const UpChunk& u = up[k];
                tm.begin(S_FORM);
                form_level(l, u.f0, u.f1);
                tm.end(S_FORM);
                tm.begin(S_MERGE);
                tm.ctx(l - 1, &merge_l[l - 1], 2, 0, u.p0, u.p1);
                launch_merge3((const int*)D(o_merge[l - 1]) + u.p0, u.p1 - u.p0, d_child, d_cen,
                              src_se, d_smp, tab.wms, 8 * (nlevels - l), nd, nexp, mp, s);
                tm.end(S_MERGE);
            }
            ci = cj;
        }
        tm.begin(S_M2L);
        tm.ctx(nlevstart, &m2l_l, 2);
        launch_m2l3((const int*)D(o_m2l), (int)m2l_l.size(), d_cen, d_ncol, d_col, d_nch, src_se,
                    trg_se, use_src, use_trg, d_smp, d_sloc, tab.wm1, npw, T.bsz[nlevstart],
                    nlevstart == npwlevel, nd, nexp, mp, lc, s);
        tm.end(S_M2L);
        // steps 4-5 by level: non-leaf children get their persistent local
        // expansions, then chunks of leaves are split into the chunk region
        // (below the top level) and evaluated
        size_t di = 0;
        for (int l = nlevstart; l <= nlevels; ++l) {
            if (l > nlevstart) {
                tm.begin(S_SPLIT);
                tm.ctx(l - 1, &split_l[l - 1], 2);
                launch_split3((const int*)D(o_split[l - 1]), (int)split_l[l - 1].size(), d_par,
                              d_cen, d_sloc, tab.wms, 8 * (nlevels - l), nd, nexp, lc, src_se,
                              trg_se, use_src, use_trg, s);
                tm.end(S_SPLIT);
            }
            for (; di < down.size() && down[di].l == l; ++di) {
                const DownChunk& dc = down[di];
                const int* boxes = (const int*)D(o_eval[l]) + dc.e0;
                const int nbox = dc.e1 - dc.e0;
                if (l > nlevstart && !fused[l]) {
                    tm.begin(S_SPLIT);
                    tm.ctx(l - 1, &eval_l[l], 2, 0, dc.e0, dc.e1);
                    launch_split3(boxes, nbox, d_par, d_cen, d_sloc, tab.wms, 8 * (nlevels - l), nd,
                                  nexp, lc, src_se, trg_se, use_src, use_trg, s);
                    tm.end(S_SPLIT);
                }
                tm.begin(S_EVAL);
                for (int ind = 0; ind < nd; ++ind) {
                    Eval3Args ea{};
                    ea.boxes = boxes;
                    ea.nbox = nbox;
                    ea.cen = d_cen;
                    ea.slot = d_sloc;
                    ea.nd = nd;
                    ea.ind = ind;
                    ea.dsq = dsq;
                    ea.nl = nlocal[l];
                    ea.dcoef = tab.dcoef;
                    ea.cstride = npw;
                    ea.ts = tab.ts;
                    ea.npw = npw;
                    ea.loc = lc;
                    ea.taylor = taylor;
                    ea.rs = src_se;
                    ea.rt = trg_se;
                    ea.use_src = use_src;
                    ea.use_trg = use_trg;
                    ea.thresh = eval_thresh[l];
                    ea.tlayout = nlocal[l] + 1 <= 32;
                    if (fused[l]) {
                        ea.fused = 1;
                        ea.par = d_par;
                        ea.wms = tab.wms;
                        ea.kbase = 8 * (nlevels - l);
                    }
                    ea.grid = std::min(2048, ea.nbox);
                    tm.ctx(l, &eval_l[l], 2, 0, dc.e0, dc.e1);
                    if (ea.thresh != INT_MAX) launch_pw2local3(ea, s);
                    for (int side = 0; side < 2; ++side) {
                        if (side == 0 && !use_trg) continue;
                        if (side == 1 && !use_src) continue;
                        Eval3Args eb = ea;
                        eb.tstart = (side == 0 ? ts_trg[l] : ts_src[l]) + dc.e0;
                        eb.seg = SEGE;
                        eb.grid = (int)std::min<long>(GRIDMAX, ea.nbox + (side == 0 ? nt : ns) / SEGE + 1);
                        eb.se = side == 0 ? trg_se : src_se;
                        eb.pts = side == 0 ? xt : xs;
                        eb.ifpgh = side == 0 ? q.ifpghtarg : q.ifpgh;
                        eb.pot = side == 0 ? pt : ps;
                        eb.grad = side == 0 ? gt : gs;
                        eb.hess = side == 0 ? ht : hs;
                        tm.ctx(l, &eval_l[l], side == 0 ? 1 : 0, 0, dc.e0, dc.e1);
                        if (eb.thresh != INT_MAX) launch_taylor_eval3(eb, s);
                        if (eb.thresh > 0) launch_pw_eval3(eb, s);
                    }
                }
                tm.end(S_EVAL);
            }
        }
    }

    // ---------------- step 6: near field, dmax = ln(1/eps) delta (fgt3d.f:1466) ----------------
    near_field_lists();
    Pack pn(W.h_plan_nf, !W.arena);
    size_t o_units = pn.add(unit_ids), o_uptr = pn.add(unit_ptr), o_ulist = pn.add(unit_list);
    char* dnf = W.get<char>("plan_nf", pn.size());
    {
        HIPCHECK(hipMemcpyAsync(dnf, pn.data(), pn.size(), hipMemcpyHostToDevice, s));
    }
    tm.begin(S_DIRECT);
    if (!unit_ids.empty()) {
        DirectArgs da{};
        da.units = (const int*)(dnf + o_units);
        da.nunits = (int)unit_ids.size();
        da.list_ptr = (const int*)(dnf + o_uptr);
        da.list = (const int*)(dnf + o_ulist);
        da.src_se = src_se;
        da.xs = xs;
        da.qs = qs;
        da.nd = nd;
        da.ws = wsd;
        da.delta = q.delta;
        da.dmax = std::log(1.0 / q.eps) * q.delta;
        if (use_trg) {
            DirectArgs d2 = da;
            d2.tgt_se = trg_se;
            d2.xt = xt;
            d2.ifpgh = q.ifpghtarg;
            d2.pot = pt; d2.grad = gt; d2.hess = ht;
            tm.ctx(npwlevel, nullptr, 3);  // units: the pairs, counted after the call
            launch_direct3(d2, s);
        }
        if (use_src) {
            DirectArgs d2 = da;
            d2.tgt_se = src_se;
            d2.xt = xs;
            d2.ifpgh = q.ifpgh;
            d2.pot = ps; d2.grad = gs; d2.hess = hs;
            tm.ctx(npwlevel, nullptr, 4);
            launch_direct3(d2, s);
        }
    }
    tm.end(S_DIRECT);

    tm.begin(S_SCATTER);
    if (use_src) {
        tm.ctx(-1, nullptr, -1, ns);
        launch_scatter(ps, isrc, ns, nd, q.pot, s);
        if (q.ifpgh >= 2) launch_scatter(gs, isrc, ns, 3 * nd, q.grad, s);
        if (q.ifpgh >= 3) launch_scatter(hs, isrc, ns, 6 * nd, q.hess, s);
    }
    if (use_trg) {
        tm.ctx(-1, nullptr, -1, nt);
        launch_scatter(pt, itarg, nt, nd, q.pottarg, s);
        if (q.ifpghtarg >= 2) launch_scatter(gt, itarg, nt, 3 * nd, q.gradtarg, s);
        if (q.ifpghtarg >= 3) launch_scatter(ht, itarg, nt, 6 * nd, q.hesstarg, s);
    }
    tm.end(S_SCATTER);
    HIPCHECK(hipGetLastError());
    tm.finish();
    W.last_valid = true;

    t_stats.assign(6 + nlevels + 1, 0);
    t_stats[0] = npwlevel;
    t_stats[1] = nlmax;
    t_stats[2] = nlevels;
    t_stats[3] = nb;
    t_stats[4] = npw;
    t_stats[5] = nleaves;
    for (int l = 0; l <= nlevels; ++l) t_stats[6 + l] = nlocal[l];
}

int run_guarded(const std::function<void()>& f)
{
    t_err.clear();
    try {
        f();
        return FGTPW_OK;
    } catch (const Error& e) {
        t_err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        t_err = "host allocation failed";
        return FGTPW_ENOMEM;
    } catch (const std::exception& e) {
        t_err = e.what();
        return FGTPW_EDEVICE;
    }
}

void validate(int nd, double delta, double eps, int ns, int ifcharge, int ifdipole, int ifpgh,
              int nt, int ifpghtarg)
{
    if (nd < 1) fail(FGTPW_EINVAL, "nd must be >= 1");
    if (ns < 1) fail(FGTPW_EINVAL, "ns must be >= 1");
    if (nt < 0) fail(FGTPW_EINVAL, "nt must be >= 0");
    if (!(delta > 0) || !std::isfinite(delta)) fail(FGTPW_EINVAL, "delta must be positive");
    if (!(eps > 0) || !(eps < 1)) fail(FGTPW_EINVAL, "eps must be in (0,1)");
    if (ifpgh < 0 || ifpgh > 3 || ifpghtarg < 0 || ifpghtarg > 3)
        fail(FGTPW_EINVAL, "ifpgh and ifpghtarg must be in 0..3");
    if (ifcharge != 0 && ifcharge != 1) fail(FGTPW_EINVAL, "ifcharge must be 0 or 1");
    if (ifdipole != 0 && ifdipole != 1) fail(FGTPW_EINVAL, "ifdipole must be 0 or 1");
}

void check_strengths(int ifcharge, const double* charge, int ifdipole, const double* rnormal,
                     const double* dipstr)
{
    if (ifcharge && !charge) fail(FGTPW_EINVAL, "charge is NULL");
    if (ifdipole && (!rnormal || !dipstr)) fail(FGTPW_EINVAL, "rnormal or dipstr is NULL");
}

void select_device()
{
    if (g_device >= 0) HIPCHECK(hipSetDevice(g_device));
}
}  // namespace
}  // namespace fgtpw

using namespace fgtpw;

extern "C" {

const char* fgtpw_last_error(void) { return t_err.c_str(); }

int fgtpw_init(int device)
{
    return run_guarded([&] {
        int n = 0;
        HIPCHECK(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) fail(FGTPW_EINVAL, "no such HIP device");
        g_device = device;
        HIPCHECK(hipSetDevice(device));
    });
}

int fgtpw_fgt2d_device(int nd, double delta, double eps, int ns, const double* sources,
                       int ifcharge, const double* charge, int ifdipole, const double* rnormal,
                       const double* dipstr, int iper, int ifpgh, double* pot, double* grad,
                       double* hess, int nt, const double* targ, int ifpghtarg, double* pottarg,
                       double* gradtarg, double* hesstarg, void* stream)
{
    (void)iper;
    return run_guarded([&] {
        validate(nd, delta, eps, ns, ifcharge, ifdipole, ifpgh, nt, ifpghtarg);
        check_strengths(ifcharge, charge, ifdipole, rnormal, dipstr);
        std::lock_guard<std::mutex> lk(g_mutex);
        select_device();
        Workspace& W = workspace();
        Call q{nd, delta, eps, ns, sources, ifcharge ? charge : nullptr, ifpgh, pot, grad, hess,
               nt, targ, ifpghtarg, pottarg, gradtarg, hesstarg};
        if (ifdipole) { q.rn = rnormal; q.ds = dipstr; }
        hipStream_t s = (hipStream_t)stream;
        run2d(W, q, s);
        HIPCHECK(hipStreamSynchronize(s));
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The no-allocation variant of fgtpw_fgt2d_device.  Every device buffer of
// the transform (keys, sort scratch, sorted points and outputs, the plan,
// expansions, partial moments, Taylor coefficients, tables) is carved out of
// `scratch`, one static array per device that the library holds until the
// process ends; host staging is pageable memory; all work runs on the
// caller's stream.  The array keeps the last call's contents: every buffer
// the transform reads before writing is cleared by a kernel of the call.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)256 << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

namespace fgtpw {
namespace {
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};   // device address of `scratch`, per device
Workspace* g_noalloc_ws[kMaxDevices] = {};

// Device address of the current device's copy of `scratch`, looked up once
// per device (with that device current).  Host code uses this address,
// never the array's name, which on the host is only a stand-in.
char* scratch_ptr(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) fail(FGTPW_EUNSUPPORTED, "device index above 63");
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail(FGTPW_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    return (char*)g_scratch_ptr[dev];
}

// Bytes of the buffers that scale with the points (keys, permutations, leaf
// records, sorted points, densities and outputs, sort scratch), each
// rounded to the arena's 256-byte alignment; the tree-sized rest is checked
// as it is carved out.
size_t noalloc_point_bytes(int nd, int ns, int nt, int ifdipole, int ifpgh, int ifpghtarg)
{
    auto r = [](size_t b) { return (std::max<size_t>(b, 256) + 255) & ~size_t(255); };
    const size_t NS = ns, NT = nt, ND = nd, nmx = std::max(ns, nt);
    auto outs = [&](int ifp, size_t n) {
        return (ifp >= 1 ? r(8 * ND * n) : 0) + (ifp >= 2 ? r(16 * ND * n) : 0) +
               (ifp >= 3 ? r(24 * ND * n) : 0);
    };
    size_t tmpb = std::max(sort_pairs_u64_temp((int)nmx), sort_pairs_u32_temp((int)nmx));
    tmpb = std::max(tmpb, leaf_records_temp(ns));
    // outputs once: the near field's sorted accumulators (the evaluations
    // write the caller's arrays)
    size_t b = 2 * r(8 * NS) + r(4 * nmx) + r(4 * NS) + r(tmpb) + r(4 * NS) + r(NS) +
               r(256 + sizeof(LeafRun) * NS) + r(16 * NS) + r(8 * ND * NS) +
               (ifdipole ? r(16 * ND * NS) : 0) + outs(ifpgh, NS);
    if (nt > 0) b += r(4 * NT) + r(16 * NT) + 2 * r(8 * NT) + outs(ifpghtarg, NT);
    return b;
}
}  // namespace
}  // namespace fgtpw

extern "C" {

int fgtpw_fgt2d_device_noalloc(int nd, double delta, double eps, int ns, const double* sources,
                               int ifcharge, const double* charge, int ifdipole,
                               const double* rnormal, const double* dipstr, int iper, int ifpgh,
                               double* pot, double* grad, double* hess, int nt, const double* targ,
                               int ifpghtarg, double* pottarg, double* gradtarg, double* hesstarg,
                               void* stream)
{
    (void)iper;
    return run_guarded([&] {
        validate(nd, delta, eps, ns, ifcharge, ifdipole, ifpgh, nt, ifpghtarg);
        check_strengths(ifcharge, charge, ifdipole, rnormal, dipstr);
        const size_t need = noalloc_point_bytes(nd, ns, nt, ifdipole, ifpgh, ifpghtarg);
        if (need > kScratchBytes)
            fail(FGTPW_EINVAL, "ns = " + std::to_string(ns) + ", nt = " + std::to_string(nt) +
                                   ", nd = " + std::to_string(nd) + " need " +
                                   std::to_string(need >> 20) + " MiB of point buffers; the "
                                   "no-allocation entry point has " +
                                   std::to_string(kScratchBytes >> 20) + " MiB of scratch");
        std::lock_guard<std::mutex> lk(g_mutex);
        select_device();
        int dev = 0;
        HIPCHECK(hipGetDevice(&dev));
        char* base = scratch_ptr(dev);
        Workspace*& wp = g_noalloc_ws[dev];
        if (!wp) wp = new Workspace();  // host object only: no streams, no device memory
        Workspace& W = *wp;
        W.arena = base;
        W.arena_cap = kScratchBytes;
        W.arena_reset();
        hipStream_t s = (hipStream_t)stream;
        W.own = W.aux[0] = W.aux[1] = s;  // one stream, the caller's
        Call q{nd, delta, eps, ns, sources, ifcharge ? charge : nullptr, ifpgh, pot, grad, hess,
               nt, targ, ifpghtarg, pottarg, gradtarg, hesstarg};
        if (ifdipole) { q.rn = rnormal; q.ds = dipstr; }
        run2d(W, q, s);
        HIPCHECK(hipStreamSynchronize(s));
    });
}

}  // extern "C"

namespace fgtpw {
namespace {
// Host-array transfers through the workspace's pinned ring: chunks of
// kRingHalf bytes alternate between the two halves, the host's copy of one
// chunk (split over the worker threads) running while the DMA of the other
// is in flight.
constexpr size_t kRingHalf = (size_t)8 << 20;
// FGTPW_STAGED_COPIES=0: plain pageable copies instead (for comparison)
const bool g_staged = [] {
    const char* e = std::getenv("FGTPW_STAGED_COPIES");
    return !(e && e[0] == '0');
}();

void ring_init(Workspace& W, hipStream_t s)
{
    if (W.ring) return;
    HIPCHECK(hipHostMalloc((void**)&W.ring, 2 * kRingHalf, hipHostMallocDefault));
    for (auto& e : W.ring_ev) {
        HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HIPCHECK(hipEventRecord(e, s));  // both halves start free
    }
}

void staged_h2d(Workspace& W, void* dst, const void* src, size_t n, hipStream_t s)
{
    if (!g_staged) {
        HIPCHECK(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, s));
        return;
    }
    for (size_t off = 0, i = 0; off < n; off += kRingHalf, ++i) {
        const int h = (int)(i & 1);
        const size_t len = std::min(kRingHalf, n - off);
        char* buf = W.ring + h * kRingHalf;
        HIPCHECK(hipEventSynchronize(W.ring_ev[h]));  // the half's last DMA is done
        parallel_copy(buf, (const char*)src + off, len);
        HIPCHECK(hipMemcpyAsync((char*)dst + off, buf, len, hipMemcpyHostToDevice, s));
        HIPCHECK(hipEventRecord(W.ring_ev[h], s));
    }
}

void staged_d2h(Workspace& W, void* dst, const void* src, size_t n, hipStream_t s)
{
    if (!g_staged) {
        HIPCHECK(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, s));
        return;
    }
    const size_t nch = (n + kRingHalf - 1) / kRingHalf;
    auto issue = [&](size_t i) {
        const int h = (int)(i & 1);
        const size_t off = i * kRingHalf, len = std::min(kRingHalf, n - off);
        HIPCHECK(hipEventSynchronize(W.ring_ev[h]));  // an upload may still read it
        HIPCHECK(hipMemcpyAsync(W.ring + h * kRingHalf, (const char*)src + off, len,
                                hipMemcpyDeviceToHost, s));
        HIPCHECK(hipEventRecord(W.ring_ev[h], s));
    };
    if (nch) issue(0);
    for (size_t i = 0; i < nch; ++i) {
        // the next chunk lands in the other half (emptied by the host in
        // the previous iteration) while this one is copied out
        if (i + 1 < nch) issue(i + 1);
        const int h = (int)(i & 1);
        const size_t off = i * kRingHalf, len = std::min(kRingHalf, n - off);
        HIPCHECK(hipEventSynchronize(W.ring_ev[h]));
        parallel_copy((char*)dst + off, W.ring + h * kRingHalf, len);
    }
}

// host-array entry (fgt2d / fgt3d): copies in, runs on the library's own
// stream, copies the requested outputs back
// charge == nullptr: no charges; rnormal == nullptr: no dipoles
void host_transform(Workspace& W, int dim, int nd, double delta, double eps, int ns,
                    const double* sources, const double* charge, const double* rnormal,
                    const double* dipstr, int ifpgh, double* pot, double* grad, double* hess,
                    int nt, const double* targ, int ifpghtarg, double* pottarg,
                    double* gradtarg, double* hesstarg)
{
    hipStream_t s = W.own;
    const size_t ng = dim, nh = dim == 2 ? 3 : 6;
    const size_t nps = dim * (size_t)ns, npt = dim * (size_t)nt;
    double* d_src = W.get<double>("in_src", nps);
    double* d_chg = charge ? W.get<double>("in_chg", (size_t)nd * ns) : nullptr;
    double* d_rn = rnormal ? W.get<double>("in_rn", nps) : nullptr;
    double* d_ds = rnormal ? W.get<double>("in_ds", (size_t)nd * ns) : nullptr;
    double* d_trg = W.get<double>("in_trg", std::max<size_t>(npt, 3));
    if (g_staged) ring_init(W, s);
    staged_h2d(W, d_src, sources, nps * 8, s);
    if (nt > 0) staged_h2d(W, d_trg, targ, npt * 8, s);
    if (charge) staged_h2d(W, d_chg, charge, (size_t)nd * ns * 8, s);
    if (rnormal) {
        staged_h2d(W, d_rn, rnormal, nps * 8, s);
        staged_h2d(W, d_ds, dipstr, (size_t)nd * ns * 8, s);
    }
    const size_t wps = ifpgh > 0 ? (size_t)ns : 0, wpt = ifpghtarg > 0 ? (size_t)nt : 0;
    double* o_p = W.get<double>("out_pot", std::max<size_t>(nd * wps, 1));
    double* o_g = W.get<double>("out_grad", std::max<size_t>(ng * nd * wps, 1));
    double* o_h = W.get<double>("out_hess", std::max<size_t>(nh * nd * wps, 1));
    double* o_pt = W.get<double>("out_pott", std::max<size_t>(nd * wpt, 1));
    double* o_gt = W.get<double>("out_gradt", std::max<size_t>(ng * nd * wpt, 1));
    double* o_ht = W.get<double>("out_hesst", std::max<size_t>(nh * nd * wpt, 1));
    Call q{nd, delta, eps, ns, d_src, d_chg, ifpgh, o_p, o_g, o_h,
           nt, d_trg, ifpghtarg, o_pt, o_gt, o_ht};
    q.rn = d_rn;
    q.ds = d_ds;
    if (dim == 2)
        run2d(W, q, s);
    else
        run3d(W, q, s);
    if (ifpgh >= 1) staged_d2h(W, pot, o_p, nd * wps * 8, s);
    if (ifpgh >= 2) staged_d2h(W, grad, o_g, ng * nd * wps * 8, s);
    if (ifpgh >= 3) staged_d2h(W, hess, o_h, nh * nd * wps * 8, s);
    if (ifpghtarg >= 1 && nt > 0) staged_d2h(W, pottarg, o_pt, nd * wpt * 8, s);
    if (ifpghtarg >= 2 && nt > 0) staged_d2h(W, gradtarg, o_gt, ng * nd * wpt * 8, s);
    if (ifpghtarg >= 3 && nt > 0) staged_d2h(W, hesstarg, o_ht, nh * nd * wpt * 8, s);
    HIPCHECK(hipStreamSynchronize(s));
}

int run_host(int dim, int nd, double delta, double eps, int ns, const double* sources,
             int ifcharge, const double* charge, int ifdipole, const double* rnormal,
             const double* dipstr, int ifpgh, double* pot, double* grad, double* hess, int nt,
             const double* targ, int ifpghtarg, double* pottarg, double* gradtarg,
             double* hesstarg)
{
    return run_guarded([&] {
        validate(nd, delta, eps, ns, ifcharge, ifdipole, ifpgh, nt, ifpghtarg);
        check_strengths(ifcharge, charge, ifdipole, rnormal, dipstr);
        std::lock_guard<std::mutex> lk(g_mutex);
        select_device();
        host_transform(workspace(), dim, nd, delta, eps, ns, sources, ifcharge ? charge : nullptr,
                       ifdipole ? rnormal : nullptr, dipstr, ifpgh, pot, grad, hess, nt, targ,
                       ifpghtarg, pottarg, gradtarg, hesstarg);
    });
}

// fgt2dpart_direct_vec / fgt3dpart_direct_vec: direct sums over ranges with
// a cutoff r^2 < dmax, accumulated into host arrays
int run_part_direct(int dim, int nd, double delta, double dmax, int istart, int iend, int jstart,
                    int jend, const double* source, int ifcharge, const double* charge,
                    int ifdipole, const double* rnormal, const double* dipstr,
                    const double* targ, int ifpgh, double* pot, double* grad, double* hess)
{
    return run_guarded([&] {
        if (nd < 1) fail(FGTPW_EINVAL, "nd must be >= 1");
        if (ifcharge != 0 && ifcharge != 1) fail(FGTPW_EINVAL, "ifcharge must be 0 or 1");
        if (ifdipole != 0 && ifdipole != 1) fail(FGTPW_EINVAL, "ifdipole must be 0 or 1");
        check_strengths(ifcharge, charge, ifdipole, rnormal, dipstr);
        if (!ifcharge && !ifdipole) return;  // fgt2dpart_direct_vec: no branch applies
        if (ifpgh < 1 || ifpgh > 3) return;  // the reference does nothing for other values
        if (!(delta > 0) || std::isnan(dmax)) fail(FGTPW_EINVAL, "delta must be positive");
        const int ns = iend - istart + 1, nt = jend - jstart + 1;
        if (ns <= 0 || nt <= 0) return;
        if (istart < 1 || jstart < 1) fail(FGTPW_EINVAL, "ranges are 1-based");
        std::lock_guard<std::mutex> lk(g_mutex);
        select_device();
        Workspace& W = workspace();
        hipStream_t s = W.own;
        const size_t ng = dim, nh = dim == 2 ? 3 : 6;
        const double* src = source + dim * (size_t)(istart - 1);
        const double* tg = targ + dim * (size_t)(jstart - 1);
        double* d_src = W.get<double>("pd_src", dim * (size_t)ns);
        double* d_chg = ifcharge ? W.get<double>("pd_chg", (size_t)nd * ns) : nullptr;
        double* d_ws = nullptr;  // dipole weights d n, (ns, nd, 2)
        double* d_tg = W.get<double>("pd_tg", dim * (size_t)nt);
        double* d_p = W.get<double>("pd_pot", (size_t)nd * nt);
        double* d_g = W.get<double>("pd_grad", ng * nd * nt);
        double* d_h = W.get<double>("pd_hess", nh * nd * nt);
        HIPCHECK(hipMemcpyAsync(d_src, src, dim * (size_t)ns * 8, hipMemcpyHostToDevice, s));
        if (ifcharge)
            HIPCHECK(hipMemcpyAsync(d_chg, charge + (size_t)nd * (istart - 1), (size_t)nd * ns * 8,
                                    hipMemcpyHostToDevice, s));
        if (ifdipole) {
            double* d_rn = W.get<double>("pd_rn", dim * (size_t)ns);
            double* d_ds = W.get<double>("pd_ds", (size_t)nd * ns);
            unsigned* id = W.get<unsigned>("pd_iota", ns);
            d_ws = W.get<double>("pd_ws", dim * (size_t)nd * ns);
            HIPCHECK(hipMemcpyAsync(d_rn, rnormal + dim * (size_t)(istart - 1), dim * (size_t)ns * 8,
                                    hipMemcpyHostToDevice, s));
            HIPCHECK(hipMemcpyAsync(d_ds, dipstr + (size_t)nd * (istart - 1), (size_t)nd * ns * 8,
                                    hipMemcpyHostToDevice, s));
            launch_iota(id, ns, s);
            if (dim == 2)
                launch_gather_dip(d_rn, d_ds, nd, id, ns, d_ws, s);
            else
                launch_gather_dip3(d_rn, d_ds, nd, id, ns, d_ws, s);
        }
        HIPCHECK(hipMemcpyAsync(d_tg, tg, dim * (size_t)nt * 8, hipMemcpyHostToDevice, s));
        HIPCHECK(hipMemcpyAsync(d_p, pot + (size_t)nd * (jstart - 1), (size_t)nd * nt * 8,
                                hipMemcpyHostToDevice, s));
        if (ifpgh >= 2)
            HIPCHECK(hipMemcpyAsync(d_g, grad + ng * nd * (jstart - 1), ng * nd * nt * 8,
                                    hipMemcpyHostToDevice, s));
        if (ifpgh >= 3)
            HIPCHECK(hipMemcpyAsync(d_h, hess + nh * nd * (jstart - 1), nh * nd * nt * 8,
                                    hipMemcpyHostToDevice, s));
        if (dim == 2)
            launch_direct_all(d_src, d_chg, d_ws, nd, ns, d_tg, nt, delta, dmax, ifpgh, d_p, d_g,
                              d_h, s);
        else
            launch_direct_all3(d_src, d_chg, d_ws, nd, ns, d_tg, nt, delta, dmax, ifpgh, d_p, d_g,
                               d_h, s);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(pot + (size_t)nd * (jstart - 1), d_p, (size_t)nd * nt * 8,
                                hipMemcpyDeviceToHost, s));
        if (ifpgh >= 2)
            HIPCHECK(hipMemcpyAsync(grad + ng * nd * (jstart - 1), d_g, ng * nd * nt * 8,
                                    hipMemcpyDeviceToHost, s));
        if (ifpgh >= 3)
            HIPCHECK(hipMemcpyAsync(hess + nh * nd * (jstart - 1), d_h, nh * nd * nt * 8,
                                    hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
    });
}
}  // namespace
}  // namespace fgtpw

extern "C" {

int fgtpw_fgt2d(int nd, double delta, double eps, int ns, const double* sources, int ifcharge,
                const double* charge, int ifdipole, const double* rnormal, const double* dipstr,
                int iper, int ifpgh, double* pot, double* grad, double* hess, int nt,
                const double* targ, int ifpghtarg, double* pottarg, double* gradtarg,
                double* hesstarg)
{
    (void)iper;
    return run_host(2, nd, delta, eps, ns, sources, ifcharge, charge, ifdipole, rnormal, dipstr,
                    ifpgh, pot, grad, hess, nt, targ, ifpghtarg, pottarg, gradtarg, hesstarg);
}

int fgtpw_fgt3d(int nd, double delta, double eps, int ns, const double* sources, int ifcharge,
                const double* charge, int ifdipole, const double* rnormal, const double* dipstr,
                int iper, int ifpgh, double* pot, double* grad, double* hess, int nt,
                const double* targ, int ifpghtarg, double* pottarg, double* gradtarg,
                double* hesstarg)
{
    (void)iper;
    return run_host(3, nd, delta, eps, ns, sources, ifcharge, charge, ifdipole, rnormal, dipstr,
                    ifpgh, pot, grad, hess, nt, targ, ifpghtarg, pottarg, gradtarg, hesstarg);
}

int fgtpw_fgt3d_device(int nd, double delta, double eps, int ns, const double* sources,
                       int ifcharge, const double* charge, int ifdipole, const double* rnormal,
                       const double* dipstr, int iper, int ifpgh, double* pot, double* grad,
                       double* hess, int nt, const double* targ, int ifpghtarg, double* pottarg,
                       double* gradtarg, double* hesstarg, void* stream)
{
    (void)iper;
    return run_guarded([&] {
        validate(nd, delta, eps, ns, ifcharge, ifdipole, ifpgh, nt, ifpghtarg);
        check_strengths(ifcharge, charge, ifdipole, rnormal, dipstr);
        std::lock_guard<std::mutex> lk(g_mutex);
        select_device();
        Workspace& W = workspace();
        Call q{nd, delta, eps, ns, sources, ifcharge ? charge : nullptr, ifpgh, pot, grad, hess,
               nt, targ, ifpghtarg, pottarg, gradtarg, hesstarg};
        if (ifdipole) { q.rn = rnormal; q.ds = dipstr; }
        hipStream_t s = (hipStream_t)stream;
        run3d(W, q, s);
        HIPCHECK(hipStreamSynchronize(s));
    });
}

int fgtpw_fgt3dpart_direct(int nd, double delta, double dmax, int istart, int iend, int jstart,
                           int jend, const double* source, int ifcharge, const double* charge,
                           int ifdipole, const double* rnormal, const double* dipstr,
                           const double* targ, int ifpgh, double* pot, double* grad, double* hess)
{
    return run_part_direct(3, nd, delta, dmax, istart, iend, jstart, jend, source, ifcharge,
                           charge, ifdipole, rnormal, dipstr, targ, ifpgh, pot, grad, hess);
}

int fgtpw_fgt2dpart_direct(int nd, double delta, double eps, int istart, int iend, int jstart,
                           int jend, const double* source, int ifcharge, const double* charge,
                           int ifdipole, const double* rnormal, const double* dipstr,
                           const double* targ, int ifpgh, double* pot, double* grad, double* hess)
{
    if (!(eps > 0)) {
        t_err = "eps must be positive";
        return FGTPW_EINVAL;
    }
    return run_part_direct(2, nd, delta, std::log(1.0 / eps) * delta, istart, iend, jstart, jend,
                           source, ifcharge, charge, ifdipole, rnormal, dipstr, targ, ifpgh, pot,
                           grad, hess);
}

int fgtpw_last_tree2d(int* info, int cap, int* laddr, int* lev, int* par, int* nch, int* ch,
                      int* ncol, int* col, double* cen, int* isrc, int* isrcse, int* itarg,
                      int* itargse)
{
    return run_guarded([&] {
        std::lock_guard<std::mutex> lk(g_mutex);
        select_device();
        Workspace& W = workspace();
        if (!W.last_valid) fail(FGTPW_EINVAL, "no tree: the last call built none");
        if (W.last_dim != 2) fail(FGTPW_EINVAL, "the last call was a 3D transform");
        complete_colleagues(W.last);
        const Tree2D& T = W.last;
        const int nb = T.nboxes();
        info[0] = T.nlevels;
        info[1] = nb;
        info[2] = (int)T.leaf_of_rank.size();
        info[3] = T.nlmax;
        if (!laddr || cap < nb) return;
        std::memcpy(laddr, T.laddr.data(), sizeof(int) * T.laddr.size());
        std::memcpy(lev, T.lev.data(), sizeof(int) * nb);
        std::memcpy(par, T.par.data(), sizeof(int) * nb);
        std::memcpy(nch, T.nch.data(), sizeof(int) * nb);
        std::memcpy(ch, T.ch.data(), sizeof(int) * 4 * nb);
        std::memcpy(ncol, T.ncol.data(), sizeof(int) * nb);
        std::memcpy(col, T.col.data(), sizeof(int) * 9 * nb);
        std::memcpy(cen, T.cen.data(), sizeof(double) * 2 * nb);
        hipStream_t s = W.own;
        HIPCHECK(hipMemcpyAsync(isrc, W.get<unsigned>("isrc", W.last_ns), sizeof(int) * W.last_ns,
                                hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(isrcse, W.get<int2>("src_se", nb), sizeof(int2) * nb,
                                hipMemcpyDeviceToHost, s));
        if (W.last_nt > 0) {
            HIPCHECK(hipMemcpyAsync(itarg, W.get<unsigned>("itarg", W.last_nt),
                                    sizeof(int) * W.last_nt, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipMemcpyAsync(itargse, W.get<int2>("trg_se", nb), sizeof(int2) * nb,
                                    hipMemcpyDeviceToHost, s));
        }
        HIPCHECK(hipStreamSynchronize(s));
        // The transform keeps each leaf's points in Morton order; the
        // reference's stable partitions (sort_pts_to_children,
        // pts_tree2d4.f:594-689) keep them in input order.  Same sets per
        // leaf, so sorting each leaf's indices gives the reference order.
        for (int b = 0; b < nb; ++b) {
            if (T.nch[b] != 0) continue;
            std::sort(isrc + isrcse[2 * b], isrc + isrcse[2 * b + 1]);
            if (W.last_nt > 0) std::sort(itarg + itargse[2 * b], itarg + itargse[2 * b + 1]);
        }
    });
}

int fgtpw_last_tree3d(int* info, int cap, int* laddr, int* lev, int* par, int* nch, int* ch,
                      int* ncol, int* col, double* cen, int* isrc, int* isrcse, int* itarg,
                      int* itargse)
{
    return run_guarded([&] {
        std::lock_guard<std::mutex> lk(g_mutex);
        select_device();
        Workspace& W = workspace();
        if (!W.last_valid) fail(FGTPW_EINVAL, "no tree: the last call built none");
        if (W.last_dim != 3) fail(FGTPW_EINVAL, "the last call was a 2D transform");
        complete_colleagues(W.last3);
        const Tree3D& T = W.last3;
        const int nb = T.nboxes();
        info[0] = T.nlevels;
        info[1] = nb;
        info[2] = (int)T.leaf_of_rank.size();
        info[3] = T.nlmax;
        if (!laddr || cap < nb) return;
        std::memcpy(laddr, T.laddr.data(), sizeof(int) * T.laddr.size());
        std::memcpy(lev, T.lev.data(), sizeof(int) * nb);
        std::memcpy(par, T.par.data(), sizeof(int) * nb);
        std::memcpy(nch, T.nch.data(), sizeof(int) * nb);
        std::memcpy(ch, T.ch.data(), sizeof(int) * 8 * nb);
        std::memcpy(ncol, T.ncol.data(), sizeof(int) * nb);
        std::memcpy(col, T.col.data(), sizeof(int) * 27 * nb);
        std::memcpy(cen, T.cen.data(), sizeof(double) * 3 * nb);
        hipStream_t s = W.own;
        HIPCHECK(hipMemcpyAsync(isrc, W.get<unsigned>("isrc", W.last_ns), sizeof(int) * W.last_ns,
                                hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(isrcse, W.get<int2>("src_se", nb), sizeof(int2) * nb,
                                hipMemcpyDeviceToHost, s));
        if (W.last_nt > 0) {
            HIPCHECK(hipMemcpyAsync(itarg, W.get<unsigned>("itarg", W.last_nt),
                                    sizeof(int) * W.last_nt, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipMemcpyAsync(itargse, W.get<int2>("trg_se", nb), sizeof(int2) * nb,
                                    hipMemcpyDeviceToHost, s));
        }
        HIPCHECK(hipStreamSynchronize(s));
        // reference order inside each leaf (stable partitions keep input order)
        for (int b = 0; b < nb; ++b) {
            if (T.nch[b] != 0) continue;
            std::sort(isrc + isrcse[2 * b], isrc + isrcse[2 * b + 1]);
            if (W.last_nt > 0) std::sort(itarg + itargse[2 * b], itarg + itargse[2 * b + 1]);
        }
    });
}

int fgtpw_tree2d_from_leaves(int nlmax, int npwlevel, double bs0, const double* cen0, int nleaves,
                             const unsigned long long* keys, const int* levels, int* info,
                             int cap, int* laddr, int* lev, int* par, int* nch, int* ch,
                             int* ncol, int* col, double* cen)
{
    return run_guarded([&] {
        std::vector<LeafRun> leaves(nleaves);
        for (int i = 0; i < nleaves; ++i) leaves[i] = LeafRun{keys[i], levels[i], 0};
        Tree2D T;
        build_tree2d(T, leaves, nlmax, npwlevel, bs0, cen0);
        complete_colleagues(T);
        const int nb = T.nboxes();
        info[0] = T.nlevels;
        info[1] = nb;
        info[2] = (int)T.leaf_of_rank.size();
        info[3] = T.nlmax;
        if (!laddr || cap < nb) return;
        std::memcpy(laddr, T.laddr.data(), sizeof(int) * T.laddr.size());
        std::memcpy(lev, T.lev.data(), sizeof(int) * nb);
        std::memcpy(par, T.par.data(), sizeof(int) * nb);
        std::memcpy(nch, T.nch.data(), sizeof(int) * nb);
        std::memcpy(ch, T.ch.data(), sizeof(int) * 4 * nb);
        std::memcpy(ncol, T.ncol.data(), sizeof(int) * nb);
        std::memcpy(col, T.col.data(), sizeof(int) * 9 * nb);
        std::memcpy(cen, T.cen.data(), sizeof(double) * 2 * nb);
    });
}

int fgtpw_last_stats(int* out, int n)
{
    int k = std::min(n, (int)t_stats.size());
    for (int i = 0; i < k; ++i) out[i] = t_stats[i];
    return k;
}

int fgtpw_set_precision(int mode)
{
    if (mode != 0 && mode != 1) {
        t_err = "precision mode must be 0 (FP64) or 1 (mixed)";
        return FGTPW_EINVAL;
    }
    const int prev = g_precision;
    g_precision = mode;
    return prev;
}

int fgtpw_last_precision(void) { return t_mixed ? 1 : 0; }

const char* fgtpw_tuning_table(void)
{
    static std::string text;
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    const Tuning& t = tuning();
    text.clear();
    for (int i = 0; i < 4; ++i)
        text += "segf" + std::to_string(i + 1) + "=" + std::to_string(t.segf[i]) + " sege" +
                std::to_string(i + 1) + "=" + std::to_string(t.sege[i]) + " spw" +
                std::to_string(i + 1) + "=" + std::to_string(t.spw[i]) + " ";
    text += "sort=" + std::to_string(t.sort) + ";source=" + t.source;
    return text.c_str();
}

int fgtpw_set_timing(int on)
{
    g_timing = on != 0;
    g_serial = on == 2 || g_serial_env;
    return FGTPW_OK;
}

int fgtpw_stage_times(double* ms, int n)
{
    int k = std::min(n, (int)t_stage_ms.size());
    for (int i = 0; i < k; ++i) ms[i] = t_stage_ms[i];
    return k;
}

int fgtpw_kernel_launches(int* id, int* level, double* ms, long long* units, int cap)
{
    int n = (int)t_klaunch.size();
    for (int i = 0; i < std::min(n, cap); ++i) {
        id[i] = t_klaunch[i].id;
        level[i] = t_klaunch[i].level;
        ms[i] = t_klaunch[i].ms;
        units[i] = t_klaunch[i].units;
    }
    return n;
}

const char* fgtpw_kernel_name(int i)
{
    if (i < 0 || i >= KID_N) return "";
    return kKernels[i];
}

const char* fgtpw_stage_name(int i)
{
    if (i < 0 || i >= (int)(sizeof(kStages) / sizeof(kStages[0]))) return "";
    return kStages[i];
}

}  // extern "C"

// ---------------------------------------------------------------------------
// serving: a queue of transforms run by a pool of slots, each slot a worker
// thread with its own workspace (device buffers that stay allocated from one
// request to the next) and its own stream, so the host tree of one request
// overlaps the kernels of others and the kernels of several requests share
// the device.  Requests are admitted in arrival order when a slot is free
// and their estimated HBM footprint fits the budget beside those in flight.
// ---------------------------------------------------------------------------
namespace fgtpw {
namespace {
struct ServeRequest {
    long long id = 0;
    int dim = 2, nd = 1, ns = 0, nt = 0, ifpgh = 0, ifpghtarg = 0;
    double delta = 0, eps = 0;
    std::vector<double> src, chg, trg, rn, ds;  // copies made by submit
    double *pot = nullptr, *grad = nullptr, *hess = nullptr;
    double *pottarg = nullptr, *gradtarg = nullptr, *hesstarg = nullptr;
    size_t footprint = 0;
    int status = FGTPW_OK;
    std::string err;
    bool done = false;
    std::chrono::steady_clock::time_point t_submit, t_start, t_end;
};

struct Server {
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<ServeRequest*> queue;
    std::unordered_map<long long, ServeRequest*> tickets;
    std::vector<std::thread> workers;
    std::vector<Workspace*> slots;
    size_t budget = 0, in_flight = 0, peak_in_flight = 0;
    int running = 0, peak_running = 0, waiters = 0;
    long long next_id = 1;
    bool stopping = false;
    int device = 0;
};
Server* g_server = nullptr;
std::mutex g_server_mutex;  // start / stop

int quantities(int dim, int ifpgh)
{
    return ifpgh <= 0 ? 0 : ifpgh == 1 ? 1 : ifpgh == 2 ? 1 + dim : 1 + dim + (dim == 2 ? 3 : 6);
}

// HBM bytes a request is admitted with (an estimate): inputs and their
// sorted copies, outputs with their sorted-order accumulators and staging,
// keys/permutations/sort scratch (48 B per point), and two plane-wave
// expansions (npw^d/2 complex per density) per box at or below the cutoff
// level, whose side is about dc = sqrt(delta ln(1/eps)) (fgt2d.f:124-127).
// The tree reaches the cutoff level L only when boxes of level L-1 still
// hold more than ndiv0 = 10 points on average (fgt2d.f:158); then the boxes
// with expansions are the occupied cutoff cells plus those below, refined
// to ndiv points (80 in 2D, 100 in 3D; fgt2d.f:162).
size_t footprint(const ServeRequest& r, double extent)
{
    const double d = r.dim, n = (double)r.ns + r.nt, nch = std::pow(2.0, d);
    // per output quantity: the output, staging and a sorted accumulator
    // (2D: the near field's; 3D: the evaluations')
    const double per_out = 24.0;
    double bytes = 16.0 * ((d + r.nd) * r.ns + d * r.nt) +
                   per_out * r.nd * ((double)quantities(r.dim, r.ifpgh) * r.ns +
                                     (double)quantities(r.dim, r.ifpghtarg) * r.nt) +
                   48.0 * n;
    const double lne = std::log(1.0 / r.eps);
    const double dc = std::sqrt(r.delta * lne);
    const double lcut = std::max(0.0, std::ceil(std::log2(std::max(extent, 1e-300) / dc)));
    const double ltree = std::ceil(std::log(std::max(n / 10.0, 1.0)) / std::log(nch)) + 1.0;
    if (lcut <= ltree) {
        const double npw = 2.0 * std::floor(3.0 * lne / M_PI) + 2.0;
        const double nexp = std::pow(npw, d) / 2.0;
        const double cells = std::pow(std::min(std::pow(2.0, lcut), 1e6), d);
        const double ndiv = r.dim == 2 ? 80.0 : 100.0;
        const double nbox = std::min(cells, n) + nch / (nch - 1.0) * 4.0 * n / ndiv;
        bytes += 2.0 * 16.0 * nexp * r.nd * nbox;
    }
    return (size_t)std::min(bytes, 1e18);
}

void serve_worker(Server* S, int slot)
{
    (void)hipSetDevice(S->device);
    Workspace* W = S->slots[slot];
    for (;;) {
        ServeRequest* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->m);
            S->cv_work.wait(lk, [&] {
                if (S->queue.empty()) return S->stopping;
                const ServeRequest* f = S->queue.front();
                return S->running == 0 || S->in_flight + f->footprint <= S->budget;
            });
            if (S->queue.empty()) return;  // stopping and drained
            r = S->queue.front();
            S->queue.pop_front();
            S->in_flight += r->footprint;
            S->running += 1;
            S->peak_in_flight = std::max(S->peak_in_flight, S->in_flight);
            S->peak_running = std::max(S->peak_running, S->running);
        }
        r->t_start = std::chrono::steady_clock::now();
        r->status = run_guarded([&] {
            host_transform(*W, r->dim, r->nd, r->delta, r->eps, r->ns, r->src.data(),
                           r->chg.empty() ? nullptr : r->chg.data(),
                           r->rn.empty() ? nullptr : r->rn.data(), r->ds.data(), r->ifpgh,
                           r->pot, r->grad, r->hess, r->nt, r->nt > 0 ? r->trg.data() : nullptr,
                           r->ifpghtarg, r->pottarg, r->gradtarg, r->hesstarg);
        });
        if (r->status != FGTPW_OK) r->err = t_err;
        r->t_end = std::chrono::steady_clock::now();
        // release the request's inputs before it is reported done
        std::vector<double>().swap(r->src);
        std::vector<double>().swap(r->chg);
        std::vector<double>().swap(r->trg);
        std::vector<double>().swap(r->rn);
        std::vector<double>().swap(r->ds);
        {
            std::lock_guard<std::mutex> lk(S->m);
            S->in_flight -= r->footprint;
            S->running -= 1;
            r->done = true;
        }
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
}
}  // namespace
}  // namespace fgtpw

extern "C" {

int fgtpw_serve_start(int slots, double hbm_budget_gib)
{
    return run_guarded([&] {
        std::lock_guard<std::mutex> g(g_server_mutex);
        if (g_server) fail(FGTPW_EINVAL, "the server is already running");
        if (slots < 1 || slots > 16) fail(FGTPW_EINVAL, "slots must be in 1..16");
        if (!(hbm_budget_gib >= 0)) fail(FGTPW_EINVAL, "the HBM budget must be >= 0");
        select_device();
        auto* S = new Server();
        HIPCHECK(hipGetDevice(&S->device));
        size_t free_b = 0, total_b = 0;
        HIPCHECK(hipMemGetInfo(&free_b, &total_b));
        S->budget = hbm_budget_gib > 0 ? (size_t)(hbm_budget_gib * 1073741824.0)
                                       : (size_t)(0.6 * (double)free_b);
        try {
            for (int i = 0; i < slots; ++i) S->slots.push_back(new_workspace(false));
        } catch (...) {
            for (auto* w : S->slots) { w->release(); delete w; }
            delete S;
            throw;
        }
        for (int i = 0; i < slots; ++i) S->workers.emplace_back(serve_worker, S, i);
        g_server = S;
    });
}

long long fgtpw_submit(const fgtpw_call* c)
{
    long long ticket = 0;
    int rc = run_guarded([&] {
        if (!c) fail(FGTPW_EINVAL, "no call");
        if (c->dim != 2 && c->dim != 3) fail(FGTPW_EINVAL, "dim must be 2 or 3");
        validate(c->nd, c->delta, c->eps, c->ns, c->ifcharge, c->ifdipole, c->ifpgh, c->nt,
                 c->ifpghtarg);
        if (!c->sources || (c->nt > 0 && !c->targ)) fail(FGTPW_EINVAL, "missing input array");
        check_strengths(c->ifcharge, c->charge, c->ifdipole, c->rnormal, c->dipstr);

        auto r = std::make_unique<ServeRequest>();
        r->dim = c->dim;
        r->nd = c->nd;
        r->ns = c->ns;
        r->nt = c->nt;
        r->delta = c->delta;
        r->eps = c->eps;
        r->ifpgh = c->ifpgh;
        r->ifpghtarg = c->nt > 0 ? c->ifpghtarg : 0;
        r->src.assign(c->sources, c->sources + (size_t)c->dim * c->ns);
        if (c->ifcharge) r->chg.assign(c->charge, c->charge + (size_t)c->nd * c->ns);
        if (c->ifdipole) {
            r->rn.assign(c->rnormal, c->rnormal + (size_t)c->dim * c->ns);
            r->ds.assign(c->dipstr, c->dipstr + (size_t)c->nd * c->ns);
        }
        if (c->nt > 0) r->trg.assign(c->targ, c->targ + (size_t)c->dim * c->nt);
        r->pot = c->pot; r->grad = c->grad; r->hess = c->hess;
        r->pottarg = c->pottarg; r->gradtarg = c->gradtarg; r->hesstarg = c->hesstarg;
        double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
        for (const auto* v : {&r->src, &r->trg})
            for (size_t i = 0; i < v->size(); ++i) {
                const int k = (int)(i % c->dim);
                lo[k] = std::min(lo[k], (*v)[i]);
                hi[k] = std::max(hi[k], (*v)[i]);
            }
        double ext = 0;
        for (int k = 0; k < c->dim; ++k) ext = std::max(ext, hi[k] - lo[k]);
        if (!std::isfinite(ext)) fail(FGTPW_EINVAL, "non-finite coordinates");
        r->footprint = footprint(*r, ext);
        r->t_submit = std::chrono::steady_clock::now();
        std::lock_guard<std::mutex> g(g_server_mutex);
        Server* S = g_server;
        if (!S) fail(FGTPW_EINVAL, "the server is not running (fgtpw_serve_start)");
        {
            std::lock_guard<std::mutex> lk(S->m);
            r->id = S->next_id++;
            ticket = r->id;
            S->tickets[r->id] = r.get();
            S->queue.push_back(r.release());
        }
        S->cv_work.notify_all();
    });
    return rc == FGTPW_OK ? ticket : rc;
}

int fgtpw_wait(long long ticket, double* ms)
{
    int status = FGTPW_OK;
    std::string err;
    int rc = run_guarded([&] {
        Server* S;
        std::unique_lock<std::mutex> lk;
        {
            std::lock_guard<std::mutex> g(g_server_mutex);
            S = g_server;
            if (!S) fail(FGTPW_EINVAL, "the server is not running");
            lk = std::unique_lock<std::mutex>(S->m);
        }
        auto it = S->tickets.find(ticket);
        if (it == S->tickets.end()) fail(FGTPW_EINVAL, "unknown ticket");
        ServeRequest* r = it->second;
        S->tickets.erase(it);
        S->waiters += 1;
        S->cv_done.wait(lk, [&] { return r->done; });
        S->waiters -= 1;
        lk.unlock();
        S->cv_done.notify_all();
        if (ms) {
            auto d = [](std::chrono::steady_clock::time_point a,
                        std::chrono::steady_clock::time_point b) {
                return std::chrono::duration<double, std::milli>(b - a).count();
            };
            ms[0] = d(r->t_submit, r->t_start);
            ms[1] = d(r->t_start, r->t_end);
            ms[2] = d(r->t_submit, r->t_end);
        }
        status = r->status;
        err = r->err;
        delete r;
    });
    if (rc != FGTPW_OK) return rc;
    if (status != FGTPW_OK) t_err = err;
    return status;
}

int fgtpw_serve_stats(long long* out, int n)
{
    std::lock_guard<std::mutex> g(g_server_mutex);
    Server* S = g_server;
    if (!S) return 0;
    std::lock_guard<std::mutex> lk(S->m);
    const long long v[] = {(long long)S->slots.size(), (long long)S->budget,
                           (long long)S->queue.size(), (long long)S->running,
                           (long long)S->in_flight, (long long)S->peak_running,
                           (long long)S->peak_in_flight};
    const int k = std::min(n, (int)(sizeof(v) / sizeof(v[0])));
    for (int i = 0; i < k; ++i) out[i] = v[i];
    return k;
}

int fgtpw_serve_stop(void)
{
    return run_guarded([&] {
        std::lock_guard<std::mutex> g(g_server_mutex);
        Server* S = g_server;
        if (!S) return;
        {
            std::lock_guard<std::mutex> lk(S->m);
            S->stopping = true;
        }
        S->cv_work.notify_all();
        for (auto& t : S->workers) t.join();
        {
            // waits already blocked on a request return before the server goes
            std::unique_lock<std::mutex> lk(S->m);
            S->cv_done.wait(lk, [&] { return S->waiters == 0; });
        }
        // requests never waited for
        for (auto& kv : S->tickets) delete kv.second;
        for (auto* w : S->slots) { w->release(); delete w; }
        delete S;
        g_server = nullptr;
    });
}

}  // extern "C"
