// Batched multishift CG on the device: the host side of CGM::Invert
// (inverter_cgm.cpp:17-87) for nsys independent systems that share one gauge
// field, each with its own list of shifts.
//
// The iteration is captured once in a hipGraph and replayed in chunks; all
// recurrence scalars live in device memory (CgmState), so a chunk needs no host
// round trip. Between chunks the host reads the states back, and in mixed
// precision performs reliable updates: the FP32 iterated solution is folded
// into FP64 accumulators and the residual of the base system is recomputed in
// FP64, r = b - A_0 x, whenever |r| has dropped by `delta` since the last one.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "kq_cgm_kernels.h"
#include "kq_dslash.h"
#include "kq_layout.h"

namespace kq {

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
            hipError_t e = hipMalloc(&p_, bytes);
            if (e != hipSuccess) {
                p_ = nullptr;
                throw std::bad_alloc();
            }
        }
        n_ = bytes;
    }
    void release()
    {
        if (p_)
            (void)hipFree(p_);
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
        }
        n_ = n;
    }
    void release()
    {
        if (p_)
            (void)hipHostFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T *get() const { return p_; }
    T &operator[](size_t i) const { return p_[i]; }

private:
    T *p_ = nullptr;
    size_t n_ = 0;
};

struct RunnerTuning {
    int nr = 1;       // right-hand sides per thread in the batched hopping term
    int nt = 256;     // threads per block of the hopping term
    int chunk = 16;   // iterations replayed between host checks
};

template <typename S, typename VR, class LF, class LL> class CgmRunner {
public:
    using V2 = c2_t<VR>;

    CgmRunner(const Geom &g, int parity, LF fat, LL lng, int nsys, int nshift, hipStream_t st, RunnerTuning tune)
        : g_(g), parity_(parity), fat_(fat), lng_(lng), nsys_(nsys), nshift_(nshift), st_(st), tune_(tune)
    {
        const size_t n3 = (size_t)3 * g.Vh;
        r_.alloc(sizeof(V2) * nsys * n3);
        s_.alloc(sizeof(V2) * nsys * n3);
        tmp_.alloc(sizeof(V2) * nsys * n3);
        p_.alloc(sizeof(V2) * nsys * nshift * n3);
        x_.alloc(sizeof(V2) * nsys * nshift * n3);
        state_.alloc(sizeof(CgmState<S>) * nsys);
        state_h_.alloc(nsys);
        nblk_ = blas_nblk(g.Vh);
        nblk_dsl_ = dslash_nblk(g.Vh, tune.nt);
        part_a_.alloc(sizeof(double) * nsys * nblk_dsl_);
        part_b_.alloc(sizeof(double) * nsys * nblk_);
        ints_.alloc(sizeof(int) * (2 * nsys + 4));
        sigma_.alloc(sizeof(double) * nsys);
        last_fold_.assign(nsys, nshift);
    }
    ~CgmRunner()
    {
        if (graph_exec_)
            (void)hipGraphExecDestroy(graph_exec_);
    }

    // Mixed precision: FP64 links for the true residual and FP64 accumulators.
    void enable_reliable(LinkD fat64, LinkD lng64, double delta)
    {
        reliable_ = true;
        fat64_ = fat64;
        lng64_ = lng64;
        delta_ = delta;
        const size_t n3 = (size_t)3 * g_.Vh;
        b_hi_.alloc(sizeof(double2) * nsys_ * n3);
        x_hi_.alloc(sizeof(double2) * nsys_ * nshift_ * n3);
        tmp64_.alloc(sizeof(double2) * nsys_ * n3);
        ax64_.alloc(sizeof(double2) * nsys_ * n3);
        nblk_dsl64_ = dslash_nblk(g_.Vh, 256);
        part_rel_.alloc(sizeof(double) * nsys_ * nblk_dsl64_);
    }

    // Buffers the caller fills before start():
    VR *r_buf() const { return r_.as<VR>(); }              // [nsys][3][Vh] source (plain mode)
    double *b_hi() const { return b_hi_.as<double>(); }    // [nsys][3][Vh] FP64 source (reliable mode)
    double *x_hi() const { return x_hi_.as<double>(); }    // [nsys][nshift][3][Vh] FP64 solution
    VR *x_buf() const { return x_.as<VR>(); }              // [nsys][nshift][3][Vh] iterated solution
    const Geom &geom() const { return g_; }
    int nsys() const { return nsys_; }
    int nshift() const { return nshift_; }

    // shifts: nsys*nshift in the reference convention (absolute shift[0],
    // shift[k>=1] relative to it); tol: absolute squared target per system.
    // initial_guess (reliable mode, nshift == 1 only): start from x_hi.
    void start(const std::vector<S> &shifts, const std::vector<double> &tol, int max_iter, bool initial_guess = false)
    {
        if (nshift_ > kMaxShift)
            throw std::invalid_argument("too many shifts");
        if (initial_guess && (!reliable_ || nshift_ != 1))
            throw std::invalid_argument("an initial guess needs the reliable single-shift mode");
        const size_t n3 = (size_t)3 * g_.Vh;
        std::vector<double> sig(nsys_);
        for (int r = 0; r < nsys_; ++r) {
            CgmState<S> &c = state_h_[r];
            std::memset(&c, 0, sizeof(c));
            c.nshift = nshift_;
            c.maxIter = max_iter;
            c.tol = tol[r];
            for (int k = 0; k < nshift_; ++k)
                c.shift[k] = shifts[(size_t)r * nshift_ + k];
            sig[r] = (double)shifts[(size_t)r * nshift_];
        }
        KQ_HIP_CHECK(hipMemcpyAsync(state_.as<void>(), state_h_.get(), sizeof(CgmState<S>) * nsys_,
                                    hipMemcpyHostToDevice, st_));
        KQ_HIP_CHECK(hipMemcpyAsync(sigma_.as<void>(), sig.data(), sizeof(double) * nsys_, hipMemcpyHostToDevice, st_));
        KQ_HIP_CHECK(hipMemsetAsync(x_.as<void>(), 0, x_.bytes(), st_));
        set_all_list();
        if (reliable_) {
            if (!initial_guess) {
                KQ_HIP_CHECK(hipMemsetAsync(x_hi_.as<void>(), 0, x_hi_.bytes(), st_));
                KQ_HIP_CHECK(hipMemsetAsync(ax64_.as<void>(), 0, ax64_.bytes(), st_));
            } else {
                apply_a64_x0();
            }
            rel_residual<VR>(b_hi_.as<double>(), ax64_.as<double>(), vec(r_), rel_list(), rel_n(), nsys_,
                             part_b_.as<double>(), nblk_, st_);
        } else {
            vec_dot_partial<VR>(vec(r_), vec(r_), nsys_, part_b_.as<double>(), nblk_, st_);
        }
        for (int r = 0; r < nsys_; ++r)
            for (int k = 0; k < nshift_; ++k)
                KQ_HIP_CHECK(hipMemcpyAsync(p_.as<V2>() + ((size_t)r * nshift_ + k) * n3, r_.as<V2>() + (size_t)r * n3,
                                            sizeof(V2) * n3, hipMemcpyDeviceToDevice, st_));
        cgm_init_state<S>(state_.as<CgmState<S>>(), launch(), st_);
        last_fold_.assign(nsys_, nshift_);
        reliable_updates_ = 0;
    }

    // Iterate every system to convergence (or max_iter).
    void run()
    {
        if (!graph_exec_)
            capture();
        int chunk = std::max(1, tune_.chunk);
        for (;;) {
            for (int c = 0; c < chunk; ++c)
                KQ_HIP_CHECK(hipGraphLaunch(graph_exec_, st_));
            KQ_HIP_CHECK(hipMemcpyAsync(state_h_.get(), state_.as<void>(), sizeof(CgmState<S>) * nsys_,
                                        hipMemcpyDeviceToHost, st_));
            KQ_HIP_CHECK(hipStreamSynchronize(st_));
            bool live = false;
            std::vector<int> rel;
            static const bool dbg = getenv("KQ_DEBUG") != nullptr;
            for (int r = 0; r < nsys_; ++r) {
                const CgmState<S> &c = state_h_[r];
                if (dbg)
                    fprintf(stderr, "[kq] sys %d iter %d rr %.4e r2max %.4e maxTerm %d stop %d B0 %.4e a0 %.4e\n", r,
                            c.iter, (double)c.rr, c.r2max, c.maxTerm, c.stop, (double)c.B[0], (double)c.a[0]);
                if (!c.stop)
                    live = true;
                if (reliable_) {
                    const bool drop = (double)c.rr < delta_ * delta_ * c.r2max;
                    if (c.stop || drop)
                        rel.push_back(r);
                }
            }
            if (reliable_ && !rel.empty())
                reliable_update(rel);
            if (!live)
                break;
        }
    }

    std::vector<CgmState<S>> states()
    {
        KQ_HIP_CHECK(hipMemcpyAsync(state_h_.get(), state_.as<void>(), sizeof(CgmState<S>) * nsys_,
                                    hipMemcpyDeviceToHost, st_));
        KQ_HIP_CHECK(hipStreamSynchronize(st_));
        return std::vector<CgmState<S>>(state_h_.get(), state_h_.get() + nsys_);
    }
    int reliable_updates() const { return reliable_updates_; }

    // One iteration, launched on the runner's stream (also what the graph holds).
    void iteration()
    {
        const CgmLaunch L = launch();
        CgmState<S> *st = state_.as<CgmState<S>>();
        cgm_begin<S>(st, L, st_);
        DslashArgs<LF, LL, VR> d1 = dargs(1 - parity_);
        d1.src = vec_p0();
        d1.dst = vec(tmp_);
        launch_dslash(d1, EPI_STORE, tune_.nr, tune_.nt, st_);
        DslashArgs<LF, LL, VR> d2 = dargs(parity_);
        d2.src = vec(tmp_);
        d2.dst = vec(s_);
        d2.aux = vec_p0();
        d2.sigma = sigma_.as<double>();
        d2.partial = part_a_.as<double>();
        launch_dslash(d2, EPI_AXMY_DOT, tune_.nr, tune_.nt, st_);
        cgm_scalars1<S>(st, L, st_);
        cgm_axpy_norm<S, VR>(st, vec(r_), vec(s_), L, st_);
        cgm_scalars2<S>(st, L, st_);
        cgm_xp_update<S, VR, VR>(st, x_.as<VR>(), p_.as<VR>(), vec(r_), nshift_, L, st_);
    }

private:
    Vec<VR> vec(const DevBuf &b) const { return Vec<VR>{b.as<V2>(), g_.Vh, (size_t)3 * g_.Vh}; }
    Vec<VR> vec_p0() const { return Vec<VR>{p_.as<V2>(), g_.Vh, (size_t)3 * g_.Vh * nshift_}; }
    int *rlist() const { return ints_.as<int>(); }
    int *nact() const { return ints_.as<int>() + nsys_; }
    int *any() const { return ints_.as<int>() + nsys_ + 1; }
    int *rel_list() const { return ints_.as<int>() + nsys_ + 2; }
    int *rel_n() const { return ints_.as<int>() + 2 * nsys_ + 2; }

    CgmLaunch launch() const
    {
        CgmLaunch L;
        L.Vh = g_.Vh;
        L.nblk = nblk_;
        L.nblk_dsl = nblk_dsl_;
        L.nrhs = nsys_;
        L.rlist = rlist();
        L.nact = nact();
        L.any = any();
        L.part_a = part_a_.as<double>();
        L.part_b = part_b_.as<double>();
        return L;
    }

    DslashArgs<LF, LL, VR> dargs(int dst_parity) const
    {
        DslashArgs<LF, LL, VR> a{};
        a.g = g_;
        a.dst_parity = dst_parity;
        a.fat = fat_;
        a.lng = lng_;
        a.run = any();
        a.rlist = rlist();
        a.nact = nact();
        a.nrhs = nsys_;
        return a;
    }

    // upload a list of systems for the reliable-update kernels (blocking)
    void set_list(const std::vector<int> &l)
    {
        const int n = (int)l.size();
        if (n > 0)
            KQ_HIP_CHECK(hipMemcpyAsync(rel_list(), l.data(), sizeof(int) * n, hipMemcpyHostToDevice, st_));
        KQ_HIP_CHECK(hipMemcpyAsync(rel_n(), &n, sizeof(int), hipMemcpyHostToDevice, st_));
        KQ_HIP_CHECK(hipStreamSynchronize(st_));
    }
    void set_all_list()
    {
        std::vector<int> l(nsys_);
        for (int r = 0; r < nsys_; ++r)
            l[r] = r;
        set_list(l);
    }

    // ax64 = A_0 x_hi[.][0] for the systems of rel_list (FP64 links).
    void apply_a64_x0()
    {
        DslashArgs<LinkD, LinkD, double> a{};
        a.g = g_;
        a.fat = fat64_;
        a.lng = lng64_;
        a.rlist = rel_list();
        a.nact = rel_n();
        a.nrhs = nsys_;
        Vec<double> x0{x_hi_.as<double2>(), g_.Vh, (size_t)3 * g_.Vh * nshift_};
        Vec<double> t{tmp64_.as<double2>(), g_.Vh, (size_t)3 * g_.Vh};
        Vec<double> ax{ax64_.as<double2>(), g_.Vh, (size_t)3 * g_.Vh};
        a.dst_parity = 1 - parity_;
        a.src = x0;
        a.dst = t;
        launch_dslash(a, EPI_STORE, 1, 256, st_);
        a.dst_parity = parity_;
        a.src = t;
        a.dst = ax;
        a.aux = x0;
        a.sigma = sigma_.as<double>();
        a.partial = part_rel_.as<double>();
        launch_dslash(a, EPI_AXMY_DOT, 1, 256, st_);
    }

    void reliable_update(const std::vector<int> &sys)
    {
        // fold the FP32 solutions into the FP64 accumulators
        for (int r : sys)
            state_h_[r].maxTermFold = last_fold_[r];
        KQ_HIP_CHECK(hipMemcpyAsync(state_.as<void>(), state_h_.get(), sizeof(CgmState<S>) * nsys_,
                                    hipMemcpyHostToDevice, st_));
        set_list(sys);
        rel_fold<S, VR>(state_.as<CgmState<S>>(), x_hi_.as<double>(), x_.as<VR>(), nshift_, g_.Vh, rel_list(), rel_n(),
                        nsys_, st_);
        std::vector<int> live;
        for (int r : sys) {
            last_fold_[r] = std::max(0, state_h_[r].maxTerm);
            if (!state_h_[r].stop)
                live.push_back(r);
        }
        if (!live.empty()) {
            set_list(live);
            apply_a64_x0();
            rel_residual<VR>(b_hi_.as<double>(), ax64_.as<double>(), vec(r_), rel_list(), rel_n(), nsys_,
                             part_b_.as<double>(), nblk_, st_);
            rel_set_rr<S>(state_.as<CgmState<S>>(), rel_list(), rel_n(), nsys_, part_b_.as<double>(), nblk_, st_);
            reliable_updates_ += 1;
        }
        KQ_HIP_CHECK(hipStreamSynchronize(st_));
    }

    void capture()
    {
        hipGraph_t graph;
        KQ_HIP_CHECK(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
        iteration();
        KQ_HIP_CHECK(hipStreamEndCapture(st_, &graph));
        KQ_HIP_CHECK(hipGraphInstantiate(&graph_exec_, graph, nullptr, nullptr, 0));
        KQ_HIP_CHECK(hipGraphDestroy(graph));
    }

    Geom g_;
    int parity_;
    LF fat_;
    LL lng_;
    int nsys_, nshift_;
    hipStream_t st_;
    RunnerTuning tune_;
    uint32_t nblk_ = 0, nblk_dsl_ = 0, nblk_dsl64_ = 0;
    DevBuf r_, s_, tmp_, p_, x_, state_, part_a_, part_b_, ints_, sigma_;
    Pinned<CgmState<S>> state_h_;
    bool reliable_ = false;
    LinkD fat64_{}, lng64_{};
    double delta_ = 0.1;
    DevBuf b_hi_, x_hi_, tmp64_, ax64_, part_rel_;
    std::vector<int> last_fold_;
    int reliable_updates_ = 0;
    hipGraphExec_t graph_exec_ = nullptr;
};

}  // namespace kq
