// Batched multishift CG on the device: the host side of CGM::Invert
// (inverter_cgm.cpp:17-87) for many independent systems that share one gauge
// field, each with its own list of shifts (a multishift system) or a single
// shift (a refinement system started from an initial guess).
//
// One iteration is captured in a hipGraph (one per right-hand-side blocking and
// group count, with every grid sized for the systems that can be live) and
// replayed in chunks; all recurrence scalars live in device memory (CgmState),
// so a chunk needs no host round trip, and one batched hopping-term launch per
// half-iteration serves every live system, so each link is read from HBM once
// for all of them.
// Between chunks the host reads the states back, and in mixed precision
// performs reliable updates: the FP32 iterated solutions are folded into FP64
// accumulators and the base residual is recomputed in FP64, r = b - A_0 x,
// whenever |r| has dropped by `delta` since the last one. New systems may be
// added between chunks (e.g. when a multishift system drops a shift).
#pragma once

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <type_traits>
#include <vector>

#include "kq_cgm_kernels.h"
#include "kq_dslash.h"
#include "kq_layout.h"

namespace kq {

// Device-memory cache: released buffers are kept by size and handed out again to
// allocations of the same size, so repeated solves of one problem shape call
// hipMalloc/hipFree only the first time. Before a cached block is reused the
// library stream is synchronised (normally idle by then), so no work queued
// under the block's previous owner can still touch it. trim() frees the cache
// (gauge free, library end, and before retrying a failed hipMalloc).
class DevPool {
public:
    static DevPool &get()
    {
        // never destroyed: buffers owned by other static objects are released at exit
        static DevPool *pool = new DevPool;
        return *pool;
    }
    void set_stream(hipStream_t s) { stream_ = s; }
    void *take(size_t bytes)
    {
        auto it = free_.find(bytes);
        if (it != free_.end()) {
            void *p = it->second;
            free_.erase(it);
            if (stream_)
                (void)hipStreamSynchronize(stream_);
            else
                (void)hipDeviceSynchronize();
            return p;
        }
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            trim();
            if (hipMalloc(&p, bytes) != hipSuccess) {
                (void)hipGetLastError();
                return nullptr;
            }
        }
        return p;
    }
    void give(void *p, size_t bytes) { free_.emplace(bytes, p); }
    void trim()
    {
        if (free_.empty())
            return;
        if (stream_)
            (void)hipStreamSynchronize(stream_);
        for (auto &f : free_)
            (void)hipFree(f.second);
        free_.clear();
    }
    size_t cached_bytes() const
    {
        size_t n = 0;
        for (const auto &f : free_)
            n += f.first;
        return n;
    }

private:
    std::multimap<size_t, void *> free_;
    hipStream_t stream_ = nullptr;
};

// KQ_DIRECT_LAUNCH=1 makes the solvers launch their iteration kernels one by one
// instead of replaying hipGraphs (same kernels, grids and order; for tools that
// trace launches)
inline bool direct_launches()
{
    static const bool on = [] {
        const char *e = getenv("KQ_DIRECT_LAUNCH");
        return e != nullptr && e[0] == '1';
    }();
    return on;
}

// KQ_DEFERRED=0 makes the multishift systems update every shift's x and p each iteration
// instead of folding the updates per segment of recorded residuals
inline bool deferred_updates_on()
{
    static const bool on = [] {
        const char *e = getenv("KQ_DEFERRED");
        return !(e != nullptr && e[0] == '0');
    }();
    return on;
}

// KQ_PAIR_REL=0 keeps every reliable update to the systems that reached their own threshold
// (no partner joins an odd set; see CgmRunner::run)
inline bool pair_reliable_on()
{
    static const bool on = [] {
        const char *e = getenv("KQ_PAIR_REL");
        return !(e != nullptr && e[0] == '0');
    }();
    return on;
}

// History slots for the deferred updates of up to nsys multishift systems: each slot records
// kHistMax residuals of VR precision; the slots stay within 32 GiB (at 48^3x96 in FP64, four
// slots), and systems beyond them update every iteration
template <typename VR> inline int deferred_slots(int nsys, uint32_t Vh)
{
    if (!deferred_updates_on() || nsys <= 0)
        return 0;
    const size_t per = sizeof(c2_t<VR>) * 3 * (size_t)Vh * kHistMax;
    const size_t budget = (size_t)32 << 30;
    return (int)std::min<size_t>((size_t)nsys, std::max<size_t>(1, budget / per));
}

class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p_(o.p_), n_(o.n_)
    {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t bytes)
    {
        release();
        if (bytes) {
            p_ = DevPool::get().take(bytes);
            if (!p_)
                throw std::bad_alloc();
        }
        n_ = bytes;
    }
    void release()
    {
        if (p_)
            DevPool::get().give(p_, n_);
        p_ = nullptr;
        n_ = 0;
    }
    template <typename T> T *as() const { return reinterpret_cast<T *>(p_); }
    size_t bytes() const { return n_; }
    explicit operator bool() const { return p_ != nullptr; }

private:
    void *p_ = nullptr;
    size_t n_ = 0;
};

template <typename T> class Pinned {
public:
    Pinned() = default;
    explicit Pinned(size_t n) { alloc(n); }
    Pinned(const Pinned &) = delete;
    Pinned &operator=(const Pinned &) = delete;
    ~Pinned() { release(); }
    void alloc(size_t n)
    {
        release();
        if (n) {
            if (hipHostMalloc(reinterpret_cast<void **>(&p_), n * sizeof(T), hipHostMallocDefault) != hipSuccess) {
                p_ = nullptr;
                throw std::bad_alloc();
            }
            // pinned host memory is mapped into the device address space
            if (hipHostGetDevicePointer(reinterpret_cast<void **>(&d_), p_, 0) != hipSuccess) {
                (void)hipHostFree(p_);
                p_ = nullptr;
                throw std::bad_alloc();
            }
        }
        n_ = n;
    }
    void release()
    {
        if (p_)
            (void)hipHostFree(p_);
        p_ = nullptr;
        d_ = nullptr;
        n_ = 0;
    }
    T *get() const { return p_; }
    T *dev() const { return d_; }  // the same memory as seen by kernels
    T &operator[](size_t i) const { return p_[i]; }

private:
    T *p_ = nullptr;
    T *d_ = nullptr;
    size_t n_ = 0;
};


struct RunnerTuning {
    int nr = 1;      // right-hand sides per thread in the batched hopping term
    int nt = 256;    // threads per block of the hopping term
    int chunk = 8;   // iterations replayed between host checks
    int cap[5] = {0, 0, 0, 0, 0};  // occupancy cap of the hopping term for nr = 1, 2, 3, 4, 8
    int nt64 = 256;             // FP64 hopping term of the reliable updates: block size
    int cap64[2] = {0, 0};      // and occupancy cap for nr = 1, 2
    // nr by number of live systems (index clamped to the last entry); empty: always nr
    std::vector<int> nr_by_live;
    int cap_for(int r) const { return cap[r >= 8 ? 4 : r >= 4 ? 3 : r >= 3 ? 2 : r >= 2 ? 1 : 0]; }
    int pick_nr(int live) const
    {
        if (nr_by_live.empty())
            return nr;
        return nr_by_live[std::min<size_t>((size_t)std::max(live, 0), nr_by_live.size() - 1)];
    }
};

// Batched multishift CG over up to max_sys systems that share one gauge field.
// S: recurrence scalar type; VR: iteration-vector real type; LF/LL: link
// formats of the iteration. In reliable mode (VR sloppy) every system also
// carries an FP64 source and FP64 solutions, and its base residual is
// recomputed in FP64 whenever |r| dropped by delta since the last time.
template <typename S, typename VR, class LF, class LL> class CgmRunner {
public:
    using V2 = c2_t<VR>;
    using DropFn = std::function<void(int sys, int k)>;

    CgmRunner(const Geom &g, int parity, LF fat, LL lng, int max_sys, size_t pool_vecs, hipStream_t st,
              RunnerTuning tune)
        : g_(g), parity_(parity), fat_(fat), lng_(lng), max_sys_(max_sys), pool_vecs_(pool_vecs), st_(st), tune_(tune)
    {
        n3_ = (size_t)3 * g.Vh;
        r_.alloc(sizeof(V2) * max_sys * n3_);
        s_.alloc(sizeof(V2) * max_sys * n3_);
        tmp_.alloc(sizeof(V2) * max_sys * n3_);
        p_.alloc(sizeof(V2) * pool_vecs * n3_);
        x_.alloc(sizeof(V2) * pool_vecs * n3_);
        state_.alloc(sizeof(CgmState<S>) * max_sys);
        state_h_.alloc(max_sys);
        list_stage_.alloc((size_t)kStages * max_sys);
        std::memset(state_h_.get(), 0, sizeof(CgmState<S>) * max_sys);
        for (int i = 0; i < max_sys; ++i)
            state_h_[i].stop = 1;  // unused slots never run
        nblk_ = blas_nblk(g.Vh);
        nblk_dsl_ = dslash_nblk(g.Vh, tune.nt);
        // partials of both hopping terms the runner launches: the iteration's (block size
        // tune.nt, stride nblk_dsl_) and the FP64 one of the reliable updates (tune.nt64),
        // whose block sizes are tuned independently; rows for the smallest block size
        // launch_dslash accepts fit either
        part_a_.alloc(sizeof(double) * max_sys * dslash_nblk(g.Vh, kMinDslashNT));
        part_b_.alloc(sizeof(double) * max_sys * nblk_);
        ints_.alloc(sizeof(int) * (2 * max_sys + 4));
        sigma_.alloc(sizeof(double) * max_sys);
        off_.alloc(sizeof(size_t) * max_sys);
        off_h_.assign(max_sys, 0);
        sig_h_.assign(max_sys, 0.0);
        last_fold_.assign(max_sys, 0);
        prev_max_.assign(max_sys, 0);
        rel_rr_.assign(max_sys, {});
        finalized_.assign(max_sys, 1);
        KQ_HIP_CHECK(hipMemcpyAsync(state_.as<void>(), state_h_.get(), sizeof(CgmState<S>) * max_sys,
                                    hipMemcpyHostToDevice, st_));
        KQ_HIP_CHECK(hipStreamSynchronize(st_));
    }
    ~CgmRunner()
    {
        for (auto &g : graphs_)
            if (g.second)
                (void)hipGraphExecDestroy(g.second);
    }
    CgmRunner(const CgmRunner &) = delete;
    CgmRunner &operator=(const CgmRunner &) = delete;

    // Mixed precision: FP64 links for the true residual and FP64 accumulators.
    // A system pauses (waits for its reliable update) once |r|^2 falls below
    // pause_factor * delta^2 * (|r|^2 at the last update).
    // lng64: LinkD, or LinkD14 when the long links have that form
    template <class LL64> void enable_reliable(LinkD fat64, LL64 lng64, double delta, double pause_factor = 1e-2)
    {
        reliable_ = true;
        fat64_ = fat64;
        if constexpr (std::is_same<LL64, LinkD14>::value) {
            lng64r_ = lng64;
            l14_ = true;
        } else {
            lng64_ = lng64;
            l14_ = false;
        }
        delta_ = delta;
        pause2_ = pause_factor * delta * delta;
        b_hi_.alloc(sizeof(double2) * max_sys_ * n3_);
        x_hi_.alloc(sizeof(double2) * pool_vecs_ * n3_);
        tmp64_.alloc(sizeof(double2) * max_sys_ * n3_);
        ax64_.alloc(sizeof(double2) * max_sys_ * n3_);
    }

    // Deferred solution updates for up to `slots` multishift systems: each records its residuals
    // and recurrence scalars, keeps only p_0 current, and has its x_k and p_k brought up to date
    // by segment folds (kq_cgm_kernels.h). The folds form each segment's update in FP64 and add
    // it to x_hi in the reliable mode, else to the solutions themselves (VR = S: the FP64 solve
    // and the reference method in double or float).
    void enable_deferred(int slots)
    {
        if (!reliable_ && !std::is_same<S, VR>::value)
            throw std::logic_error("deferred solution updates: the solutions are x_hi or of the scalar type");
        if (tune_.chunk > kHistMax)
            throw std::invalid_argument("deferred solution updates: chunk longer than the history");
        ndslot_ = std::max(0, slots);
        if (ndslot_ == 0)
            return;
        shist_.alloc(sizeof(double) * hist_scalar_index(ndslot_, 0, 0, 0));
        rhist_.alloc(sizeof(V2) * (size_t)ndslot_ * kHistMax * n3_);
        p0snap_.alloc(sizeof(V2) * (size_t)ndslot_ * n3_);
        coef_.alloc(sizeof(double) * (size_t)ndslot_ * kMaxShift * kCoefStride);
    }

    // Register a system; it starts at the next run() step. shifts: reference
    // convention (absolute shifts[0], shifts[k>=1] relative to it); tolk: absolute
    // squared target per shift. b: device source, FP64 SoA in reliable mode, VR
    // SoA otherwise. x0 (reliable mode, single shift): FP64 initial guess or null.
    int add_system(int nshift, const S *shifts, const double *tolk, int max_iter, const void *b, const double *x0)
    {
        if (nsys_ >= max_sys_)
            throw std::runtime_error("solver: no free system slot");
        if (nshift < 1 || nshift > kMaxShift)
            throw std::invalid_argument("solver: bad shift count");
        if (pool_used_ + (size_t)nshift > pool_vecs_)
            throw std::runtime_error("solver: vector pool exhausted");
        if (x0 && (!reliable_ || nshift != 1))
            throw std::invalid_argument("solver: an initial guess needs the reliable single-shift mode");
        const int sys = nsys_++;
        off_h_[sys] = pool_used_ * n3_;
        pool_used_ += nshift;
        CgmState<S> &c = state_h_[sys];
        std::memset(&c, 0, sizeof(c));
        c.nshift = nshift;
        c.maxIter = max_iter;
        c.stop = 1;  // until started
        c.pause2 = reliable_ ? pause2_ : 0.0;
        c.dslot = (nshift > 1 && dslot_used_ < ndslot_) ? dslot_used_++ : -1;
        c.hist = 0;
        for (int k = 0; k < nshift; ++k) {
            c.shift[k] = shifts[k];
            c.tolk[k] = tolk[k];
        }
        sig_h_[sys] = (double)shifts[0];
        if (reliable_) {
            KQ_HIP_CHECK(hipMemcpyAsync(b_hi_.as<double2>() + sys * n3_, b, sizeof(double2) * n3_,
                                        hipMemcpyDeviceToDevice, st_));
            if (x0) {
                KQ_HIP_CHECK(hipMemcpyAsync(x_hi_.as<double2>() + off_h_[sys], x0, sizeof(double2) * n3_,
                                            hipMemcpyDeviceToDevice, st_));
                pending_x0_.push_back(sys);
            } else {
                KQ_HIP_CHECK(hipMemsetAsync(x_hi_.as<double2>() + off_h_[sys], 0, sizeof(double2) * nshift * n3_, st_));
                KQ_HIP_CHECK(hipMemsetAsync(ax64_.as<double2>() + sys * n3_, 0, sizeof(double2) * n3_, st_));
            }
        } else {
            KQ_HIP_CHECK(hipMemcpyAsync(r_.as<V2>() + sys * n3_, b, sizeof(V2) * n3_, hipMemcpyDeviceToDevice, st_));
        }
        KQ_HIP_CHECK(hipMemsetAsync(x_.as<V2>() + off_h_[sys], 0, sizeof(V2) * nshift * n3_, st_));
        pending_.push_back(sys);
        return sys;
    }

    // Iterate until every registered system has stopped. on_drop(sys, k) is
    // called (between chunks) for each shift k >= 1 that a system drops; by
    // then its solution is final in x_hi(sys, k) (reliable) or x_lo(sys, k)
    // once the work enqueued so far has run. after_launch() is called each time
    // a chunk of iterations has been enqueued, before the host waits for it: host
    // work placed there (copies on another stream) overlaps the chunk.
    void run(const DropFn &on_drop = nullptr, const std::function<void()> &after_launch = nullptr)
    {
        start_pending();
        static const bool dbg = getenv("KQ_DEBUG") != nullptr;
        using clk = std::chrono::steady_clock;
        double t_launch = 0, t_wait = 0, t_rel = 0;
        int nchunks = 0;
        const auto t_run0 = clk::now();
        for (;;) {
            // one graph per right-hand-side blocking, chosen from the live count
            const auto ta = clk::now();
            ++nchunks;
            int nlive = 0;
            for (int sys = 0; sys < nsys_; ++sys)
                nlive += (!state_h_[sys].stop && !state_h_[sys].pause) ? 1 : 0;
            if (direct_launches()) {
                const int nr = tune_.pick_nr(nlive);
                for (int c = 0; c < tune_.chunk; ++c)
                    iteration(nr, std::max(1, (nlive + nr - 1) / nr));
            } else {
                hipGraphExec_t gx = graph_for(nlive);
                for (int c = 0; c < tune_.chunk; ++c)
                    KQ_HIP_CHECK(hipGraphLaunch(gx, st_));
            }
            if (after_launch)
                after_launch();
            const auto tb = clk::now();
            download_states();
            const auto tc = clk::now();
            t_launch += std::chrono::duration<double>(tb - ta).count();
            t_wait += std::chrono::duration<double>(tc - tb).count();
            std::vector<int> fold, rel;
            std::vector<double> rr_iter(nsys_, 0.0);
            std::vector<std::pair<int, int>> drops;
            bool live = false;
            for (int sys = 0; sys < nsys_; ++sys) {
                CgmState<S> &c = state_h_[sys];
                if (finalized_[sys])
                    continue;
                if (dbg)
                    fprintf(stderr, "[kq] sys %d iter %d rr %.4e r2max %.4e maxTerm %d stop %d pause %d\n", sys, c.iter,
                            (double)c.rr, c.r2max, c.maxTerm, c.stop, c.pause);
                const bool dropped = c.maxTerm < prev_max_[sys];
                for (int k = std::max(1, c.maxTerm); k < prev_max_[sys]; ++k)
                    drops.emplace_back(sys, k);
                prev_max_[sys] = std::max(0, c.maxTerm);
                if (c.stop) {
                    finalized_[sys] = 1;
                    if (reliable_ || c.dslot >= 0) {
                        c.maxTermFold = last_fold_[sys];
                        fold.push_back(sys);
                    }
                    continue;
                }
                live = true;
                // a single-shift iteration whose residual grows far above the last true one has
                // lost its recurrence to the sloppy operator: update and restart it
                const bool diverging = reliable_ && c.maxTerm <= 1 && !((double)c.rr < 1e2 * c.r2max);
                const bool need_rel =
                    reliable_ && (c.pause || (double)c.rr < delta_ * delta_ * c.r2max || diverging);
                rr_iter[sys] = diverging ? 0.0 : (double)c.rr;
                c.discard = diverging ? 1 : 0;  // the updates since the last fold are not trusted
                // a recorded segment is folded before the next chunk could overrun the history
                const bool hist_full = c.dslot >= 0 && c.hist + tune_.chunk > kHistMax;
                if ((reliable_ || c.dslot >= 0) && (need_rel || dropped || hist_full)) {
                    c.maxTermFold = last_fold_[sys];
                    fold.push_back(sys);
                }
                if (need_rel) {
                    c.pause = 0;
                    rel.push_back(sys);
                }
            }
            // The FP64 operator reads its links once for a pair of systems (1.17x one system at
            // 32^4), and an odd set leaves one system of a pair slot unused: the live system
            // nearest its own threshold joins early if its residual has fallen at least halfway
            // there (|r|^2 < delta |r_last|^2, against delta^2 for its own update)
            if (reliable_ && rel.size() % 2 == 1 && pair_reliable_on()) {
                int best = -1;
                double bestq = delta_;
                for (int sys = 0; sys < nsys_; ++sys) {
                    const CgmState<S> &c = state_h_[sys];
                    if (finalized_[sys] || c.stop || c.discard || c.r2max <= 0.0 ||
                        std::find(rel.begin(), rel.end(), sys) != rel.end())
                        continue;
                    const double q = (double)c.rr / c.r2max;
                    if (q < bestq) {
                        bestq = q;
                        best = sys;
                    }
                }
                if (best >= 0) {
                    CgmState<S> &c = state_h_[best];
                    if (std::find(fold.begin(), fold.end(), best) == fold.end()) {
                        c.maxTermFold = last_fold_[best];
                        fold.push_back(best);
                    }
                    c.pause = 0;
                    rel.push_back(best);
                }
            }
            if (!fold.empty() || !rel.empty()) {
                const auto tr = clk::now();
                reliable_update(fold, rel);
                t_rel += std::chrono::duration<double>(clk::now() - tr).count();
            }
            // A single-shift system whose true residual has fallen by less than a decade over the
            // last kStallUpdates reliable updates has stalled on the sloppy operator (a healthy one
            // gains a decade or more per update; seen with 16-bit fat links on ill-conditioned
            // systems, where the base system crept to 1e-10 over ~39 000 iterations): it stops, and
            // the caller's FP64 verification refines it from a fresh start.
            std::vector<int> stalled;
            for (int sys : rel) {
                const CgmState<S> &c = state_h_[sys];
                std::vector<double> &h = rel_rr_[sys];
                h.push_back(c.r2max);
                if (h.size() > (size_t)kStallUpdates + 1)
                    h.erase(h.begin());
                if (h.size() == (size_t)kStallUpdates + 1 && !(h.back() < 1e-2 * h.front()) && !c.stop &&
                    c.maxTerm <= 1)
                    stalled.push_back(sys);
            }
            // restart single-shift systems whose iterated residual was far below the true one
            std::vector<int> restart;
            for (int sys : rel) {
                const CgmState<S> &c = state_h_[sys];
                if (!c.stop && c.maxTerm <= 1 && !((double)c.rr < 1e2 * rr_iter[sys]) &&
                    std::find(stalled.begin(), stalled.end(), sys) == stalled.end())
                    restart.push_back(sys);
            }
            if (!restart.empty()) {
                set_list(restart);
                copy_r_to_p<S, VR>(state_.as<CgmState<S>>(), vec(r_), p_.as<VR>(), off_.as<size_t>(), list(),
                                   list_n(), list_grid(), st_);
                cgm_restart_list<S>(state_.as<CgmState<S>>(), list(), list_n(), list_grid(), st_);
                if (ndslot_ > 0)
                    seg_reset<S, VR>(state_.as<CgmState<S>>(), p_.as<VR>(), p0snap_.as<VR>(), off_.as<size_t>(), g_.Vh,
                                     list(), list_n(), list_grid(), st_);
                restarts_ += (int)restart.size();
                download_states();
                if (dbg)
                    fprintf(stderr, "[kq] restarted %zu single-shift systems\n", restart.size());
            }
            if (!stalled.empty()) {
                for (int sys : stalled)
                    state_h_[sys].stop = 1;
                upload_states();
                stalls_ += (int)stalled.size();
                if (dbg)
                    fprintf(stderr, "[kq] stopped %zu stalled single-shift systems\n", stalled.size());
            }
            if (on_drop)
                for (const auto &d : drops)
                    on_drop(d.first, d.second);
            if (!pending_.empty()) {
                start_pending();
                live = true;
            }
            if (!live)
                break;
        }
        if (dbg)
            fprintf(stderr, "[kq] run: %.2f ms in %d chunks: launch %.2f ms, wait for chunk %.2f ms, reliable "
                            "updates %.2f ms\n",
                    1e3 * std::chrono::duration<double>(clk::now() - t_run0).count(), nchunks, 1e3 * t_launch,
                    1e3 * t_wait, 1e3 * t_rel);
    }

    double *x_hi(int sys, int k) const { return x_hi_.as<double>() + 2 * (off_h_[sys] + (size_t)k * n3_); }
    VR *x_lo(int sys, int k) const { return x_.as<VR>() + 2 * (off_h_[sys] + (size_t)k * n3_); }
    double *b_hi(int sys) const { return b_hi_.as<double>() + 2 * sys * n3_; }
    const CgmState<S> &state(int sys) const { return state_h_[sys]; }
    int nsys() const { return nsys_; }
    int reliable_updates() const { return reliable_updates_; }
    int restarts() const { return restarts_; }
    int stalls() const { return stalls_; }
    const Geom &geom() const { return g_; }

private:
    Vec<VR> vec(const DevBuf &b) const { return Vec<VR>{b.as<V2>(), g_.Vh, n3_}; }
    Vec<VR> vec_p0() const { return Vec<VR>{p_.as<V2>(), g_.Vh, 0, off_.as<size_t>()}; }
    int *rlist() const { return ints_.as<int>(); }
    int *nact() const { return ints_.as<int>() + max_sys_; }
    int *any() const { return ints_.as<int>() + max_sys_ + 1; }
    int *list() const { return ints_.as<int>() + max_sys_ + 2; }
    int *list_n() const { return ints_.as<int>() + 2 * max_sys_ + 2; }

    CgmLaunch launch(int nlaunch) const
    {
        CgmLaunch L;
        L.Vh = g_.Vh;
        L.nblk = nblk_;
        L.nblk_dsl = nblk_dsl_;
        L.nsys = max_sys_;
        L.nlaunch = std::min(std::max(nlaunch, 1), max_sys_);
        L.rlist = rlist();
        L.nact = nact();
        L.any = any();
        L.part_a = part_a_.as<double>();
        L.part_b = part_b_.as<double>();
        L.off = off_.as<size_t>();
        L.shist = ndslot_ > 0 ? shist_.as<double>() : nullptr;
        L.rhist = ndslot_ > 0 ? rhist_.as<void>() : nullptr;
        return L;
    }

    // the states move with a copy kernel through the mapped pinned buffer (a DMA
    // transfer would add a copy-engine round trip after every chunk)
    static_assert(sizeof(CgmState<S>) % 8 == 0, "CgmState is copied in 8-byte words");
    void download_states()
    {
        copy_words(state_.as<void>(), state_h_.dev(), sizeof(CgmState<S>) * max_sys_, st_);
        KQ_HIP_CHECK(hipStreamSynchronize(st_));
        stage_uses_ = 0;
    }
    void upload_states() { copy_words(state_h_.dev(), state_.as<void>(), sizeof(CgmState<S>) * max_sys_, st_); }
    // upload a list of systems for the list-driven kernels (blocking)
    // Upload a list of systems for the list-driven kernels without waiting: the host
    // writes it into one of kStages slots of mapped pinned memory and a kernel copies
    // it. A slot is reused only after the stream has been synchronised since its last
    // use (download_states synchronises; after kStages lists without one, wait here).
    static constexpr int kStages = 4;
    void set_list(const std::vector<int> &l)
    {
        const int n = (int)l.size();
        if (n > max_sys_)
            throw std::logic_error("set_list: more systems than the runner holds");
        if (stage_uses_ >= kStages) {
            KQ_HIP_CHECK(hipStreamSynchronize(st_));
            stage_uses_ = 0;
        }
        const size_t off = (size_t)stage_slot_ * max_sys_;
        std::copy(l.begin(), l.end(), list_stage_.get() + off);
        set_int_list(list_stage_.dev() + off, list(), list_n(), n, st_);
        stage_slot_ = (stage_slot_ + 1) % kStages;
        ++stage_uses_;
        list_size_ = std::max(n, 1);
    }
    int list_grid() const { return list_size_; }

    // ax64 = A_0 x_hi[.][0] for the n listed systems (FP64 links). The first hopping term reads
    // the links once for two systems; so does the second, with the fused sigma x - D tmp and
    // <x, Ax> epilogue, where the two-system term is the LDS-staged one (195 us against 2 x 173
    // us at 32^4); elsewhere it runs one system per thread, since the unstaged two-system form
    // costs more in FP64 registers than a second pass over the links (478 against 2 x 195 us)
    void apply_a64_x0(int n)
    {
        if (l14_)
            apply_a64_x0_t(n, lng64r_);
        else
            apply_a64_x0_t(n, lng64_);
    }
    template <class LL64> void apply_a64_x0_t(int n, LL64 lng64)
    {
        const int nr = n >= 2 ? 2 : 1;
        const int cap = tune_.cap64[nr - 1];
        DslashArgs<LinkD, LL64, double> a{};
        a.g = g_;
        a.fat = fat64_;
        a.lng = lng64;
        a.rlist = list();
        a.nact = list_n();
        a.nrhs = n;
        Vec<double> x0{x_hi_.as<double2>(), g_.Vh, 0, off_.as<size_t>()};
        Vec<double> t{tmp64_.as<double2>(), g_.Vh, n3_};
        Vec<double> ax{ax64_.as<double2>(), g_.Vh, n3_};
        a.dst_parity = 1 - parity_;
        a.src = x0;
        a.dst = t;
        launch_dslash(a, EPI_STORE, nr, tune_.nt64, st_, cap);
        a.dst_parity = parity_;
        a.src = t;
        a.dst = ax;
        a.aux = x0;
        a.sigma = sigma_.as<double>();
        a.partial = part_a_.as<double>();
        const int nr2 = (nr == 2 && dslash_staged(g_, 2, tune_.nt64)) ? 2 : 1;
        launch_dslash(a, EPI_AXMY_DOT, nr2, tune_.nt64, st_, tune_.cap64[nr2 - 1]);
    }

    void start_pending()
    {
        if (pending_.empty())
            return;
        for (int sys : pending_) {
            last_fold_[sys] = state_h_[sys].nshift;
            prev_max_[sys] = state_h_[sys].nshift;
            finalized_[sys] = 0;
            rel_rr_[sys].clear();
        }
        upload_states();
        KQ_HIP_CHECK(hipMemcpyAsync(off_.as<void>(), off_h_.data(), sizeof(size_t) * max_sys_, hipMemcpyHostToDevice,
                                    st_));
        KQ_HIP_CHECK(hipMemcpyAsync(sigma_.as<void>(), sig_h_.data(), sizeof(double) * max_sys_,
                                    hipMemcpyHostToDevice, st_));
        if (reliable_) {
            if (!pending_x0_.empty()) {
                set_list(pending_x0_);
                apply_a64_x0((int)pending_x0_.size());
            }
            set_list(pending_);
            rel_residual<VR>(b_hi_.as<double>(), ax64_.as<double>(), vec(r_), list(), list_n(), list_grid(),
                             part_b_.as<double>(), nblk_, st_);
        } else {
            set_list(pending_);
            vec_dot_partial<VR>(vec(r_), vec(r_), list_grid(), part_b_.as<double>(), nblk_, st_, list(), list_n());
        }
        copy_r_to_p<S, VR>(state_.as<CgmState<S>>(), vec(r_), p_.as<VR>(), off_.as<size_t>(), list(), list_n(),
                           list_grid(), st_);
        cgm_init_list<S>(state_.as<CgmState<S>>(), list(), list_n(), list_grid(), part_b_.as<double>(), nblk_, st_);
        if (ndslot_ > 0)
            seg_reset<S, VR>(state_.as<CgmState<S>>(), p_.as<VR>(), p0snap_.as<VR>(), off_.as<size_t>(), g_.Vh,
                             list(), list_n(), list_grid(), st_);
        pending_.clear();
        pending_x0_.clear();
        download_states();
    }

    void reliable_update(const std::vector<int> &fold, const std::vector<int> &rel)
    {
        upload_states();  // maxTermFold and cleared pause flags
        if (!fold.empty()) {
            std::vector<int> plain, seg;
            int kmax = 0;
            for (int sys : fold) {
                if (state_h_[sys].dslot >= 0) {
                    seg.push_back(sys);
                    kmax = std::max(kmax, state_h_[sys].maxTermFold);
                } else if (reliable_) {
                    plain.push_back(sys);
                }
            }
            if (!plain.empty()) {
                set_list(plain);
                rel_fold<S, VR>(state_.as<CgmState<S>>(), x_hi_.as<double>(), x_.as<VR>(), off_.as<size_t>(), g_.Vh,
                                list(), list_n(), list_grid(), st_);
            }
            if (!seg.empty()) {
                set_list(seg);
                CgmState<S> *stp = state_.as<CgmState<S>>();
                seg_coef<S>(stp, shist_.as<double>(), coef_.as<double>(), list(), list_n(), list_grid(), st_);
                // the folds accumulate into x_hi (FP64) in the reliable mode, else into the solutions
                // themselves (FP64 solve, reference method in either precision)
                if (reliable_) {
                    seg_fold<S, VR, double>(stp, coef_.as<double>(), x_hi_.as<double>(), p_.as<VR>(), p0snap_.as<VR>(),
                                            rhist_.as<VR>(), off_.as<size_t>(), g_.Vh, list(), list_n(), list_grid(),
                                            kmax, st_);
                } else if constexpr (std::is_same<S, VR>::value) {
                    seg_fold<S, VR, VR>(stp, coef_.as<double>(), x_.as<VR>(), p_.as<VR>(), p0snap_.as<VR>(),
                                        rhist_.as<VR>(), off_.as<size_t>(), g_.Vh, list(), list_n(), list_grid(), kmax,
                                        st_);
                }
                seg_reset<S, VR>(stp, p_.as<VR>(), p0snap_.as<VR>(), off_.as<size_t>(), g_.Vh, list(), list_n(),
                                 list_grid(), st_);
            }
            for (int sys : fold)
                last_fold_[sys] = state_h_[sys].stop ? 0 : std::max(0, state_h_[sys].maxTerm);
        }
        if (!rel.empty()) {
            set_list(rel);
            apply_a64_x0((int)rel.size());
            rel_residual<VR>(b_hi_.as<double>(), ax64_.as<double>(), vec(r_), list(), list_n(), list_grid(),
                             part_b_.as<double>(), nblk_, st_);
            rel_set_rr<S>(state_.as<CgmState<S>>(), list(), list_n(), list_grid(), part_b_.as<double>(), nblk_, st_);
            reliable_updates_ += 1;
        }
        download_states();  // keep the host copy current (rr, r2max set on the device)
    }

    // One iteration on the runner's stream (what the graphs hold), with the
    // hopping term blocked nr right-hand sides per thread and every per-system
    // grid sized for nr * groups live systems.
    void iteration(int nr, int groups)
    {
        const CgmLaunch L = launch(nr * groups);
        CgmState<S> *st = state_.as<CgmState<S>>();
        cgm_begin<S>(st, L, st_);
        DslashArgs<LF, LL, VR> d{};
        d.g = g_;
        d.fat = fat_;
        d.lng = lng_;
        d.run = L.any;
        d.rlist = L.rlist;
        d.nact = L.nact;
        d.nrhs = L.nlaunch;
        d.dst_parity = 1 - parity_;
        d.src = vec_p0();
        d.dst = vec(tmp_);
        launch_dslash(d, EPI_STORE, nr, tune_.nt, st_, tune_.cap_for(nr));
        d.dst_parity = parity_;
        d.src = vec(tmp_);
        d.dst = vec(s_);
        d.aux = vec_p0();
        d.sigma = sigma_.as<double>();
        d.partial = part_a_.as<double>();
        launch_dslash(d, EPI_AXMY_DOT, nr, tune_.nt, st_, tune_.cap_for(nr));
        cgm_scalars1<S>(st, L, st_);
        cgm_axpy_norm<S, VR>(st, vec(r_), vec(s_), L, st_);
        cgm_scalars2<S>(st, L, st_);
        cgm_xp_update<S, VR, VR>(st, x_.as<VR>(), p_.as<VR>(), vec(r_), L, st_);
    }

    // graph of one iteration for nlive live systems: right-hand-side blocking nr
    // from the tuning table and ceil(nlive / nr) groups (no grid slots for
    // systems that cannot be live during the chunk)
    hipGraphExec_t graph_for(int nlive)
    {
        const int nr = tune_.pick_nr(nlive);
        const int groups = std::max(1, (nlive + nr - 1) / nr);
        hipGraphExec_t &gx = graphs_[{nr, groups}];
        if (!gx) {
            hipGraph_t graph;
            KQ_HIP_CHECK(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
            iteration(nr, groups);
            KQ_HIP_CHECK(hipStreamEndCapture(st_, &graph));
            KQ_HIP_CHECK(hipGraphInstantiate(&gx, graph, nullptr, nullptr, 0));
            KQ_HIP_CHECK(hipGraphDestroy(graph));
        }
        return gx;
    }

    Geom g_;
    int parity_;
    LF fat_;
    LL lng_;
    int max_sys_;
    size_t pool_vecs_;
    hipStream_t st_;
    RunnerTuning tune_;
    size_t n3_ = 0;
    uint32_t nblk_ = 0, nblk_dsl_ = 0;
    DevBuf r_, s_, tmp_, p_, x_, state_, part_a_, part_b_, ints_, sigma_, off_;
    Pinned<CgmState<S>> state_h_;
    Pinned<int> list_stage_;  // kStages slots of max_sys_ system indices
    int stage_slot_ = 0, stage_uses_ = 0, list_size_ = 1;
    std::vector<size_t> off_h_;
    std::vector<double> sig_h_;
    int nsys_ = 0;
    size_t pool_used_ = 0;
    std::vector<int> pending_, pending_x0_;
    std::vector<int> last_fold_, prev_max_;
    std::vector<char> finalized_;
    bool reliable_ = false;
    LinkD fat64_{}, lng64_{};
    LinkD14 lng64r_{};
    bool l14_ = false;
    double delta_ = 0.1, pause2_ = 0.0;
    DevBuf b_hi_, x_hi_, tmp64_, ax64_;
    int ndslot_ = 0, dslot_used_ = 0;
    DevBuf shist_, rhist_, p0snap_, coef_;
    int reliable_updates_ = 0;
    int restarts_ = 0;
    int stalls_ = 0;
    static constexpr int kStallUpdates = 5;
    std::vector<std::vector<double>> rel_rr_;  // true |r|^2 at the last kStallUpdates + 1 reliable updates
    std::map<std::pair<int, int>, hipGraphExec_t> graphs_;
};

}  // namespace kq
