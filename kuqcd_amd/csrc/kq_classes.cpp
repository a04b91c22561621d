// The inner C++ API of kuqcd.h (fields, the HISQ operator, the multishift CG and
// BLAS-1; declared in include/kuqcd_classes.h) over the library's device formats:
// colour vectors in the SoA layout of the solver, links converted to the
// hopping-term formats, the solve run by the same device-resident CG runner as
// kuqcd_solve_multishift.
#include "kuqcd_classes.h"

#include <cmath>
#include <string>
#include <vector>

#include "kq_blas.h"
#include "kq_cgm_host.h"
#include "kq_context.h"
#include "kq_solve.h"

namespace kuqcd {

namespace {

hipStream_t lib_stream()
{
    kq::ensure_device();
    return kq::ctx().stream;
}

void *take(size_t bytes)
{
    void *p = kq::DevPool::get().take(bytes);
    if (!p)
        throw std::bad_alloc();
    KQ_HIP_CHECK(hipMemsetAsync(p, 0, bytes, lib_stream()));
    return p;
}

void give(void *p, size_t bytes)
{
    if (p)
        kq::DevPool::get().give(p, bytes);
}

uint32_t half_volume(const LattBase &lb) { return (uint32_t)(lb.lat.V / 2); }

kq::Geom geom_of(const LattBase &lb)
{
    const int L[4] = {lb.lat.nx, lb.lat.ny, lb.lat.nz, lb.lat.nt};
    return kq::make_geom(L);
}

// 0 for EVEN, 1 for ODD, -1 for both parities
int parity_of(KuQCDSiteSubset s) { return s == KUQCD_EVEN ? 0 : s == KUQCD_ODD ? 1 : -1; }

void check_lattice(const LattBase &lb, size_t real_bytes, unsigned nbatch, const char *what)
{
    if (lb.lat.V == 0 || lb.lat.nx % 2 != 0)
        throw std::invalid_argument(std::string(what) + ": the lattice is empty or nx is odd");
    if (lb.lat.V / 2 >= ((size_t)1 << 31))
        throw std::invalid_argument(std::string(what) + ": half volume must be below 2^31 sites");
    // the reference sizes the storage with latBase.prec (lattice_field.cpp:31)
    if ((size_t)lb.prec != real_bytes)
        throw std::invalid_argument(std::string(what) + ": LattBase precision differs from the field's real type");
    if (lb.Nbatch != nbatch)
        throw std::invalid_argument(std::string(what) + ": LattBase gauge order differs from the field's Nbatch");
}

// kernel link formats of T: type (18 reals) and rtype, the long links in fewer reals (LinkD14:
// a complex multiple of an SU(3) matrix; LinkF12: a real multiple), whose packer counts in *fail
// the links without that form
template <typename T> struct link_fmt;
template <> struct link_fmt<double> {
    using type = kq::LinkD;
    using rtype = kq::LinkD14;
    static void make(const double2 *canon, char *dst, uint32_t Vh, hipStream_t st)
    {
        kq::canon_to_linkD(canon, dst, Vh, st);
    }
    static void make_r(const double2 *canon, char *dst, uint32_t Vh, int *fail, hipStream_t st)
    {
        kq::canon_to_linkD14(canon, dst, Vh, fail, st);
    }
};
template <> struct link_fmt<float> {
    using type = kq::LinkF;
    using rtype = kq::LinkF12;
    static void make(const float2 *canon, char *dst, uint32_t Vh, hipStream_t st)
    {
        kq::canon_to_linkF<float>(canon, dst, Vh, st);
    }
    static void make_r(const float2 *canon, char *dst, uint32_t Vh, int *fail, hipStream_t st)
    {
        kq::canon_to_linkF12(canon, dst, Vh, fail, st);
    }
};

template <class V> struct nbatch_of;
template <typename T, int NB, int Ns> struct nbatch_of<SpinorField<T, NB, Ns>> {
    static constexpr int value = NB;
};

}  // namespace

// ---------------------------------------------------------------- SpinorField
template <typename T, int NB, int Ns> SpinorField<T, NB, Ns>::SpinorField(LattBase &lb) : latBase(lb)
{
    check_lattice(lb, sizeof(T), NB, "SpinorField");
    sites_ = lb.Sites();
    bytes_ = sizeof(kq::c2_t<T>) * 3 * sites_;
    dev_ = take(bytes_);
}

template <typename T, int NB, int Ns> SpinorField<T, NB, Ns>::SpinorField(const SpinorField &o) : latBase(o.latBase)
{
    sites_ = o.sites_;
    bytes_ = o.bytes_;
    dev_ = take(bytes_);
    KQ_HIP_CHECK(hipMemcpyAsync(dev_, o.dev_, bytes_, hipMemcpyDeviceToDevice, lib_stream()));
}

template <typename T, int NB, int Ns> SpinorField<T, NB, Ns>::~SpinorField()
{
    give(dev_, bytes_);
}

template <typename T, int NB, int Ns> SpinorField<T, NB, Ns> &SpinorField<T, NB, Ns>::operator=(const SpinorField &o)
{
    if (this == &o)
        return *this;
    if (o.sites_ != sites_)
        throw std::invalid_argument("SpinorField: assignment between fields of different lengths");
    KQ_HIP_CHECK(hipMemcpyAsync(dev_, o.dev_, bytes_, hipMemcpyDeviceToDevice, lib_stream()));
    return *this;
}

template <typename T, int NB, int Ns> void SpinorField<T, NB, Ns>::SetValues(const T *h)
{
    if (h == nullptr)
        throw std::invalid_argument("SpinorField::SetValues: null pointer");
    const uint32_t Vh = half_volume(latBase);
    const kq::Vec<T> v{static_cast<kq::c2_t<T> *>(dev_), Vh, (size_t)3 * Vh};
    kq::DevBuf stage;
    const int nhalf = latBase.site_subset == KUQCD_ALL ? 2 : 1;
    for (int q = 0; q < nhalf; ++q)
        kq::upload_vec<T, T>(h + (size_t)6 * Vh * q, v, q, stage, lib_stream());
}

template <typename T, int NB, int Ns> void SpinorField<T, NB, Ns>::GetValues(T *h) const
{
    if (h == nullptr)
        throw std::invalid_argument("SpinorField::GetValues: null pointer");
    const uint32_t Vh = half_volume(latBase);
    const kq::Vec<T> v{static_cast<kq::c2_t<T> *>(dev_), Vh, (size_t)3 * Vh};
    kq::DevBuf stage;
    const int nhalf = latBase.site_subset == KUQCD_ALL ? 2 : 1;
    for (int q = 0; q < nhalf; ++q)
        kq::download_vec<T, T>(v, q, h + (size_t)6 * Vh * q, stage, lib_stream());
}

template <typename T, int NB, int Ns> void SpinorField<T, NB, Ns>::Clear()
{
    KQ_HIP_CHECK(hipMemsetAsync(dev_, 0, bytes_, lib_stream()));
}

template <typename T, int NB, int Ns> static size_t ncomplex(const SpinorField<T, NB, Ns> &v)
{
    return (size_t)3 * v.Sites();
}

template <typename T, int NB, int Ns> static kq::c2_t<T> *cptr(const SpinorField<T, NB, Ns> &v)
{
    return static_cast<kq::c2_t<T> *>(v.DevicePtr());
}

template <typename T, int NB, int Ns>
static void lincomb(SpinorField<T, NB, Ns> &y, const SpinorField<T, NB, Ns> &x, T a, T b)
{
    if (y.Sites() != x.Sites())
        throw std::invalid_argument("BLAS-1: fields of different lengths");
    kq::flat_axpby<T>(cptr(x), cptr(y), ncomplex(y), a, b, lib_stream());
}

template <typename T, int NB, int Ns> SpinorField<T, NB, Ns> SpinorField<T, NB, Ns>::operator*(T scalar)
{
    SpinorField res(latBase);
    lincomb(res, *this, scalar, T(0));
    return res;
}

template <typename T, int NB, int Ns> SpinorField<T, NB, Ns> &SpinorField<T, NB, Ns>::operator-=(const SpinorField &v)
{
    lincomb(*this, v, T(-1), T(1));
    return *this;
}

template <typename T, int NB, int Ns> SpinorField<T, NB, Ns> SpinorField<T, NB, Ns>::operator+(const SpinorField &v)
{
    SpinorField res(*this);
    lincomb(res, v, T(1), T(1));
    return res;
}

template <typename T, int NB, int Ns>
SpinorField<T, NB, Ns> operator-(const SpinorField<T, NB, Ns> &a, const SpinorField<T, NB, Ns> &b)
{
    SpinorField<T, NB, Ns> res(a);
    lincomb(res, b, T(-1), T(1));
    return res;
}

template <typename T, int NB, int Ns> T RealDot(const SpinorField<T, NB, Ns> &bra, const SpinorField<T, NB, Ns> &ket)
{
    if (bra.Sites() != ket.Sites())
        throw std::invalid_argument("RealDot: fields of different lengths");
    const size_t n = ncomplex(bra);
    hipStream_t st = lib_stream();
    kq::DevBuf part(sizeof(double) * (kq::flat_dot_nblk(n) + 1));
    double *out = part.as<double>() + kq::flat_dot_nblk(n);
    kq::flat_dot<T>(cptr(bra), cptr(ket), n, part.as<double>(), out, st);
    double h = 0.0;
    KQ_HIP_CHECK(hipMemcpyAsync(&h, out, sizeof(double), hipMemcpyDeviceToHost, st));
    KQ_HIP_CHECK(hipStreamSynchronize(st));
    return (T)h;
}

template <typename T, int NB, int Ns> T Norm2(const SpinorField<T, NB, Ns> &v) { return RealDot(v, v); }

template <typename T, int NB, int Ns> T Norm(const SpinorField<T, NB, Ns> &v) { return std::sqrt(Norm2(v)); }

template <typename T, int NB, int Ns> void Axmy(SpinorField<T, NB, Ns> &y, const SpinorField<T, NB, Ns> &x, T a)
{
    lincomb(y, x, a, T(-1));
}

template <typename T, int NB, int Ns> void Axpy(SpinorField<T, NB, Ns> &y, const SpinorField<T, NB, Ns> &x, T a)
{
    lincomb(y, x, a, T(1));
}

template <typename T, int NB, int Ns> void Ymax(SpinorField<T, NB, Ns> &y, const SpinorField<T, NB, Ns> &x, T a)
{
    lincomb(y, x, T(-a), T(1));
}

template <typename T, int NB, int Ns>
void Axpby(SpinorField<T, NB, Ns> &y, const SpinorField<T, NB, Ns> &x, T a, T b)
{
    if (b == T(0)) {
        // y = a x: the kernel's b == 0 form does not read y (b y would be 0 unless y held inf/nan)
        lincomb(y, x, a, T(0));
        return;
    }
    lincomb(y, x, a, b);
}

// ---------------------------------------------------------------- GaugeField
template <typename T, int NB, KuQCDReconstructType R>
GaugeField<T, NB, R>::GaugeField(LattBase &lb, CommBase &c) : latBase(lb), comm(c)
{
    check_lattice(lb, sizeof(T), NB, "GaugeField");
    if (lb.site_subset != KUQCD_ALL)
        throw std::invalid_argument("GaugeField: links are stored for both parities (site_subset ALL)");
    bytes_ = sizeof(kq::c2_t<T>) * 36 * lb.lat.V;
    dev_ = take(bytes_);
}

template <typename T, int NB, KuQCDReconstructType R>
GaugeField<T, NB, R>::GaugeField(const GaugeField &o) : latBase(o.latBase), comm(o.comm)
{
    bytes_ = o.bytes_;
    dev_ = take(bytes_);
    KQ_HIP_CHECK(hipMemcpyAsync(dev_, o.dev_, bytes_, hipMemcpyDeviceToDevice, lib_stream()));
}

template <typename T, int NB, KuQCDReconstructType R> GaugeField<T, NB, R>::~GaugeField()
{
    give(dev_, bytes_);
}

template <typename T, int NB, KuQCDReconstructType R> void GaugeField<T, NB, R>::SetValues(const T *h)
{
    if (h == nullptr)
        throw std::invalid_argument("GaugeField::SetValues: null pointer");
    hipStream_t st = lib_stream();
    const size_t V = latBase.lat.V;
    const uint32_t Vh = (uint32_t)(V / 2);
    // stream the host links through a device staging buffer, 2^20 sites at a time
    const size_t chunk = std::min<size_t>(V, (size_t)1 << 20);
    kq::DevBuf stage(sizeof(T) * 72 * chunk);
    for (size_t i0 = 0; i0 < V; i0 += chunk) {
        const size_t n = std::min(chunk, V - i0);
        kq::copy_from_host(stage.as<void>(), h + 72 * i0, sizeof(T) * 72 * n, st);
        kq::gauge_host_to_soa<T, T>(stage.as<T>(), i0, n, static_cast<kq::c2_t<T> *>(dev_), Vh, 1.0, st);
        KQ_HIP_CHECK(hipStreamSynchronize(st));
    }
}

// ---------------------------------------------------------------- DiracHISQ
template <typename T, int NB, KuQCDReconstructType RS, KuQCDReconstructType RN>
DiracHISQ<T, NB, RS, RN>::DiracHISQ(KuQCDDslashType diracType, const LinkST &fat, const LinkNT &lng, double mass_,
                                    VecT *tmp_, T naikEps)
    : Dirac<VecT>(diracType), gaugeSmeared(fat), gaugeNaik(lng), tmp(tmp_), latBase(fat.latBase), mass(mass_),
      mass2((T)(mass_ * mass_)), c3000((T)((-1. / 48.0) * (1.0 + (double)naikEps)))
{
    // dirac_hisq.cpp:30-45 with the constants folded into the device links
    if (diracType != KUQCD_HISQ_DIRAC && diracType != KUQCD_HISQPC_DIRAC)
        throw std::invalid_argument("DiracHISQ: dslash type must be KUQCD_HISQ_DIRAC or KUQCD_HISQPC_DIRAC");
    const Lattice &a = fat.latBase.lat, &b = lng.latBase.lat;
    if (a.nx != b.nx || a.ny != b.ny || a.nz != b.nz || a.nt != b.nt)
        throw std::invalid_argument("DiracHISQ: fat and long links on different lattices");
    using LK = typename link_fmt<T>::type;
    hipStream_t st = lib_stream();
    const uint32_t Vh = half_volume(latBase);
    link_bytes_ = kq::link_field_bytes<LK>(Vh);
    const size_t ncomp = (size_t)36 * latBase.lat.V;
    kq::DevBuf scaled(sizeof(kq::c2_t<T>) * ncomp);
    fat_ = take(link_bytes_);
    lng_ = take(link_bytes_);
    // U' = U / 2 and N' = c3000 N (dirac_hisq.cpp:40, 473, 480)
    kq::flat_axpby<T>(static_cast<const kq::c2_t<T> *>(fat.DevicePtr()), scaled.as<kq::c2_t<T>>(), ncomp, T(0.5), T(0),
                      st);
    link_fmt<T>::make(scaled.as<kq::c2_t<T>>(), static_cast<char *>(fat_), Vh, st);
    kq::flat_axpby<T>(static_cast<const kq::c2_t<T> *>(lng.DevicePtr()), scaled.as<kq::c2_t<T>>(), ncomp, c3000, T(0),
                      st);
    link_fmt<T>::make(scaled.as<kq::c2_t<T>>(), static_cast<char *>(lng_), Vh, st);
    // the long links also in fewer reals when every one has that form (the operator and the
    // device CGM then read those)
    using RK = typename link_fmt<T>::rtype;
    lngr_bytes_ = kq::link_field_bytes<RK>(Vh);
    lngr_ = take(lngr_bytes_);
    kq::DevBuf fail(sizeof(int));
    KQ_HIP_CHECK(hipMemsetAsync(fail.as<void>(), 0, sizeof(int), st));
    link_fmt<T>::make_r(scaled.as<kq::c2_t<T>>(), static_cast<char *>(lngr_), Vh, fail.as<int>(), st);
    int nfail = 0;
    KQ_HIP_CHECK(hipMemcpyAsync(&nfail, fail.as<void>(), sizeof(int), hipMemcpyDeviceToHost, st));
    KQ_HIP_CHECK(hipStreamSynchronize(st));
    if (nfail != 0) {
        give(lngr_, lngr_bytes_);
        lngr_ = nullptr;
    }
}

template <typename T, int NB, KuQCDReconstructType RS, KuQCDReconstructType RN> DiracHISQ<T, NB, RS, RN>::~DiracHISQ()
{
    delete own_tmp_;
    delete own_tmp_base_;
    give(fat_, link_bytes_);
    give(lng_, link_bytes_);
    give(lngr_, lngr_bytes_);
}

// dst (parity q) = D src (parity 1 - q) for the half-lattice blocks at these pointers
template <typename T, class LF, class LL>
static void hop(const kq::Geom &g, LF fat, LL lng, const void *src, void *dst, int q, hipStream_t st)
{
    kq::DslashArgs<LF, LL, T> a{};
    a.g = g;
    a.dst_parity = q;
    a.fat = fat;
    a.lng = lng;
    a.src = kq::Vec<T>{static_cast<kq::c2_t<T> *>(const_cast<void *>(src)), g.Vh, (size_t)3 * g.Vh};
    a.dst = kq::Vec<T>{static_cast<kq::c2_t<T> *>(dst), g.Vh, (size_t)3 * g.Vh};
    a.nrhs = 1;
    kq::launch_dslash(a, kq::EPI_STORE, 1, kq::kReproNT, st, 0);
}

template <typename T, int NB, KuQCDReconstructType RS, KuQCDReconstructType RN>
void DiracHISQ<T, NB, RS, RN>::Dslash(const VecT &src, VecT &dst, bool update)
{
    (void)update;  // halo exchange only (dirac_hisq.cpp:412, 452); one rank has no halo
    using LK = typename link_fmt<T>::type;
    using RK = typename link_fmt<T>::rtype;
    const kq::Geom g = geom_of(latBase);
    if (src.latBase.lat.V != latBase.lat.V || dst.latBase.lat.V != latBase.lat.V)
        throw std::invalid_argument("Dslash: fields on another lattice than the links");
    const LK f{static_cast<char *>(fat_), kq::nb64_of(g.Vh)}, l{static_cast<char *>(lng_), kq::nb64_of(g.Vh)};
    const RK lr{static_cast<char *>(lngr_), kq::nb64_of(g.Vh)};
    hipStream_t st = lib_stream();
    // dst (parity q) = D src (parity 1 - q), reading the long links in fewer reals when held
    auto hop = [&](const void *s, void *d, int q) {
        if (lngr_ != nullptr)
            ::kuqcd::hop<T>(g, f, lr, s, d, q, st);
        else
            ::kuqcd::hop<T>(g, f, l, s, d, q, st);
    };
    const int qd = parity_of(dst.latBase.site_subset), qs = parity_of(src.latBase.site_subset);
    if (qd >= 0) {
        // dirac_hisq.cpp:415-419: the output parity is dst's; src holds the other one (or both)
        const char *s = static_cast<const char *>(src.DevicePtr());
        if (qs < 0)
            s += sizeof(kq::c2_t<T>) * 3 * (size_t)g.Vh * (1 - qd);
        else if (qs != 1 - qd)
            throw std::invalid_argument("Dslash: src must hold the parity opposite to dst's");
        hop(s, dst.DevicePtr(), qd);
        return;
    }
    if (this->diracType == KUQCD_HISQPC_DIRAC)
        throw std::invalid_argument("Dslash: the even-odd operator writes one parity (dst site_subset EVEN or ODD)");
    if (qs >= 0)
        throw std::invalid_argument("Dslash: the full operator maps a field of both parities (site_subset ALL)");
    // full lattice: each parity of dst from the other parity of src
    const size_t half = sizeof(kq::c2_t<T>) * 3 * (size_t)g.Vh;
    const char *s = static_cast<const char *>(src.DevicePtr());
    char *d = static_cast<char *>(dst.DevicePtr());
    hop(s + half, d, 0);
    hop(s, d + half, 1);
}

template <typename T, int NB, KuQCDReconstructType RS, KuQCDReconstructType RN>
void DiracHISQ<T, NB, RS, RN>::MdaggM(const VecT &src, VecT &dst, bool update)
{
    // dirac_hisq.cpp:457-462: D (D src) through tmp, which has the other parity
    VecT *t = tmp;
    const KuQCDSiteSubset want = src.latBase.site_subset == KUQCD_ALL  ? KUQCD_ALL
                                 : src.latBase.site_subset == KUQCD_EVEN ? KUQCD_ODD
                                                                         : KUQCD_EVEN;
    if (t == nullptr || t->latBase.site_subset != want || t->latBase.lat.V != src.latBase.lat.V) {
        if (own_tmp_ == nullptr || own_tmp_->latBase.site_subset != want ||
            own_tmp_->latBase.lat.V != src.latBase.lat.V) {
            delete own_tmp_;
            own_tmp_ = nullptr;
            delete own_tmp_base_;
            own_tmp_base_ = new LattBase(src.latBase);
            own_tmp_base_->site_subset = want;
            *own_tmp_base_ = LattBase(*own_tmp_base_);  // recompute lenFull / lenBulk
            own_tmp_ = new VecT(*own_tmp_base_);
        }
        t = own_tmp_;
    }
    Dslash(src, *t, update);
    Dslash(*t, dst, update);
}

// ---------------------------------------------------------------- CGM
namespace {

void print_iters(int iter, int max_iter)
{
    // inverter_cgm.cpp:80-84
    if (iter >= max_iter - 1)
        kq::log_info("CGM: Warning max iteration reached %d\n", iter);
    else
        kq::log_info("CGM: # iterations %d\n", iter);
}

// The recurrence of inverter_cgm.cpp:17-87 driven from the host over any Dirac operator
// (a caller's own Dirac subclass), with the inner API's fields and BLAS-1
template <typename T, typename VecT> struct FieldOps {
    Dirac<VecT> *dirac;
    std::vector<VecT> p;
    VecT ap, r;
    VecT **sol;
    T norm2_r() { return Norm2(r); }
    T apply(T sigma0)
    {
        dirac->MdaggM(p[0], ap, true);
        Axmy(ap, p[0], sigma0);  // ap = sigma_0 p0 - D D p0
        return RealDot(p[0], ap);
    }
    void update_r(T b0) { Axpy(r, ap, b0); }
    void update_x(int k, T bk) { Ymax(*sol[k], p[k], bk); }
    void update_p(int k, T zk, T ak) { Axpby(p[k], r, zk, ak); }
};

template <typename T, typename VecT>
int cgm_host(Dirac<VecT> *dirac, int ns, const T *shift, const VecT &src, VecT *sol[], double tol, int max_iter)
{
    FieldOps<T, VecT> ops{dirac, std::vector<VecT>(ns, src), VecT(src.latBase), VecT(src), sol};
    for (int k = 0; k < ns; ++k)
        sol[k]->Clear();
    return kq::cgm_recurrence<T>(ops, ns, shift, tol, max_iter);
}

}  // namespace

template <typename T, typename VecT>
int CGM<T, VecT>::Invert(int noShifts, T shifts[], const VecT &src, VecT *sol[], const double tol, int maxIters)
{
    if (noShifts <= 0 || shifts == nullptr || sol == nullptr)
        throw std::invalid_argument("CGM::Invert: no shifts");
    for (int k = 0; k < noShifts; ++k)
        if (sol[k] == nullptr || sol[k]->Sites() != src.Sites())
            throw std::invalid_argument("CGM::Invert: solution fields missing or of another length");
    if (this->dirac == nullptr)
        throw std::invalid_argument("CGM::Invert: no Dirac operator");
    constexpr int NB = nbatch_of<VecT>::value;
    using H = DiracHISQ<T, NB, KUQCD_GAUGE_RECON_NO, KUQCD_GAUGE_RECON_NO>;
    using LK = typename link_fmt<T>::type;
    H *h = dynamic_cast<H *>(this->dirac);
    const int parity = parity_of(src.latBase.site_subset);
    int iter;
    if (h != nullptr && h->diracType == KUQCD_HISQPC_DIRAC && parity >= 0 && noShifts <= kq::kMaxShift) {
        // the device runner of kuqcd_solve_multishift (one system, scalars in T)
        hipStream_t st = lib_stream();
        const kq::Geom g = geom_of(src.latBase);
        using RK = typename link_fmt<T>::rtype;
        const LK f{static_cast<char *>(const_cast<void *>(h->KernelLinks(0))), kq::nb64_of(g.Vh)};
        const LK l{static_cast<char *>(const_cast<void *>(h->KernelLinks(1))), kq::nb64_of(g.Vh)};
        const RK lr{static_cast<char *>(const_cast<void *>(h->KernelLinks(2))), kq::nb64_of(g.Vh)};
        kq::RunnerTuning tune;
        tune.nr = 1;
        tune.nt = kq::kReproNT;
        tune.chunk = 16;
        auto solve = [&](auto lng) {
            kq::CgmRunner<T, T, LK, decltype(lng)> run(g, parity, f, lng, 1, noShifts, st, tune);
            run.enable_deferred(kq::deferred_slots<T>(1, g.Vh));
            const std::vector<double> tolk(noShifts, tol);
            run.add_system(noShifts, shifts, tolk.data(), maxIters, src.DevicePtr(), nullptr);
            run.run();
            const size_t bytes = sizeof(kq::c2_t<T>) * 3 * (size_t)g.Vh;
            for (int k = 0; k < noShifts; ++k)
                KQ_HIP_CHECK(
                    hipMemcpyAsync(sol[k]->DevicePtr(), run.x_lo(0, k), bytes, hipMemcpyDeviceToDevice, st));
            KQ_HIP_CHECK(hipStreamSynchronize(st));
            return run.state(0).iter;
        };
        iter = h->KernelLinks(2) != nullptr ? solve(lr) : solve(l);
    } else {
        iter = cgm_host<T, VecT>(this->dirac, noShifts, shifts, src, sol, tol, maxIters);
    }
    print_iters(iter, maxIters);
    return iter;
}

// ---------------------------------------------------------------- instantiations
#define KQ_CLASSES_INST(T, NB)                                                                              \
    template class SpinorField<T, NB, 1>;                                                                   \
    template SpinorField<T, NB, 1> operator-(const SpinorField<T, NB, 1> &, const SpinorField<T, NB, 1> &); \
    template T Norm2(const SpinorField<T, NB, 1> &);                                                        \
    template T Norm(const SpinorField<T, NB, 1> &);                                                         \
    template T RealDot(const SpinorField<T, NB, 1> &, const SpinorField<T, NB, 1> &);                       \
    template void Axmy(SpinorField<T, NB, 1> &, const SpinorField<T, NB, 1> &, T);                          \
    template void Axpy(SpinorField<T, NB, 1> &, const SpinorField<T, NB, 1> &, T);                          \
    template void Ymax(SpinorField<T, NB, 1> &, const SpinorField<T, NB, 1> &, T);                          \
    template void Axpby(SpinorField<T, NB, 1> &, const SpinorField<T, NB, 1> &, T, T);                      \
    template class GaugeField<T, NB, KUQCD_GAUGE_RECON_NO>;                                                 \
    template class DiracHISQ<T, NB, KUQCD_GAUGE_RECON_NO, KUQCD_GAUGE_RECON_NO>;                            \
    template class CGM<T, SpinorField<T, NB, 1>>;
KQ_CLASSES_INST(float, 1)
KQ_CLASSES_INST(double, 1)
KQ_CLASSES_INST(float, 16)
KQ_CLASSES_INST(double, 8)

}  // namespace kuqcd
