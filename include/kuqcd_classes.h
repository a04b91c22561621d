/*
 * KuQCD inner C++ API, MI355X implementation: the classes the reference's
 * kuqcd.h brings into every caller (kuqcd.h:18-22), with the reference's names,
 * template parameters and semantics, over fields that live in GPU memory.
 *
 *   Lattice, LattBase               space_time.h:21-44, latt_base.h:20-55
 *   CommBase                        communication.h (single rank: nothing to exchange)
 *   SpinorField<T, Nbatch, 1>       spinor_field.h:20-40, lattice_field.h:22-37
 *   GaugeField<T, Nbatch, Recon>    gauge_field.h:21-45
 *   Dirac<VecT>, DiracHISQ<...>     dirac.h:20-29, dirac_hisq.h:25-68
 *   Inverter<T, VecT>, CGM<T, VecT> inverter.h:20-33, inverter_cgm.h:21-28
 *   Norm2, Norm, RealDot, Axmy,     spinor_field.h:42-64 (spinor_field.cpp:130-252)
 *   Axpy, Ymax, Axpby, operators
 *
 * Field data stay in GPU memory in the library's structure-of-arrays layout;
 * SetValues/GetValues copy from/to the reference's host layouts
 * (kuqcd.cpp:103-169). Operations are queued on the library's stream in call
 * order; the ones that return a number (Norm2, RealDot, Norm, CGM::Invert) and
 * GetValues wait for it. kuqcd_init() (or kq_init_grid) is called first, as
 * with the reference.
 *
 * Instantiated (as the reference, dirac_hisq.h:60-68, spinor_field.h:76-84):
 * T/Nbatch = float/1, double/1, float/16, double/8; Ns = 1; Recon = RECON_NO.
 * The Nbatch lanes of the reference's BATCH order are an internal storage
 * detail there (the host layouts are the NORMAL ones, kuqcd.cpp:118-121); here
 * every Nbatch uses the same device layout.
 */
#ifndef KUQCD_CLASSES_H
#define KUQCD_CLASSES_H

#include <array>
#include <cstddef>
#include <cstdio>
#include <stdexcept>

// Exported with default visibility, including the enumeration types that are
// template arguments here (an instantiation is only as visible as its arguments),
// also when the including code builds with -fvisibility=hidden.
#pragma GCC visibility push(default)
#include "kuqcd_enum.h"

#define kuqcd_class_export __attribute__((visibility("default")))

namespace kuqcd {

// space_time.h:21-44 -- extents of a 4-D lattice, x fastest
class Lattice {
public:
    unsigned short nx = 0;
    unsigned short ny = 0;
    unsigned short nz = 0;
    unsigned short nt = 0;
    size_t V = 0;

    Lattice() = default;
    Lattice(const Lattice &lat) = default;
    Lattice(unsigned short x, unsigned short y, unsigned short z, unsigned short t) : nx(x), ny(y), nz(z), nt(t)
    {
        V = (size_t)nx * ny * nz * nt;
    }
    Lattice &operator=(const Lattice &lat) = default;
    unsigned short &operator[](int dir) { return dir == 0 ? nx : dir == 1 ? ny : dir == 2 ? nz : nt; }
    const unsigned short &operator[](int dir) const { return dir == 0 ? nx : dir == 1 ? ny : dir == 2 ? nz : nt; }
    size_t ComputeVolume()
    {
        V = (size_t)nx * ny * nz * nt;
        return V;
    }
    void Print() const { std::printf("(%u, %u, %u, %u)\n", nx, ny, nz, nt); }
};

// latt_base.h:20-55 -- the lattice a field lives on: global extents, process grid
// (one rank here), precision, gauge order (Nbatch lanes along x in BATCH order)
// and site subset. For a field of one parity (EVEN/ODD) lenBulk = V/2/Nbatch,
// for ALL it is V/Nbatch (latt_base.cpp:116-133); with one rank there is no halo
// and lenFull = lenBulk.
class LattBase {
public:
    Lattice latSupBulkFL;  // whole lattice, x in units of Nbatch lanes
    Lattice latSupBulkCB;  // its checkerboard (x halved)
    Lattice latSubFullFL;  // this rank's part including halos (= bulk at one rank)
    Lattice latSubFullCB;
    Lattice latSubBulkFL;  // this rank's part
    Lattice latSubBulkCB;
    Lattice *latFull = nullptr;
    Lattice *latBulk = nullptr;
    size_t lenFull = 0;
    size_t lenBulk = 0;
    std::array<unsigned short, 4> gridDim{1, 1, 1, 1};
    unsigned short haloDepth = 0;
    unsigned int Nbatch = 1;
    int myRank = 0;
    int np = 1;
    KuQCDPrecision prec = KUQCD_DOUBLE;
    KuQCDGaugeOrder gauge_order = KUQCD_NORMAL_ORDER;
    KuQCDRankOrder rankOrder = KUQCD_TZYX_ORDER;
    KuQCDSiteSubset site_subset = KUQCD_ALL;
    Lattice lat;  // global extents in sites

    LattBase() = default;
    LattBase(const LattBase &o) { *this = o; }
    LattBase(const Lattice &lat_, const std::array<unsigned short, 4> &gridDim_, unsigned short depth,
             KuQCDGaugeOrder order_, KuQCDRankOrder rankOrder_, KuQCDPrecision prec_,
             KuQCDSiteSubset parity_ = KUQCD_ALL)
        : gridDim(gridDim_), haloDepth(depth), prec(prec_), gauge_order(order_), rankOrder(rankOrder_),
          site_subset(parity_), lat(lat_)
    {
        if (gridDim[0] * gridDim[1] * gridDim[2] * gridDim[3] != 1)
            throw std::invalid_argument("LattBase: one process per GPU (the process grid must be 1x1x1x1)");
        if (order_ == KUQCD_BATCH_ORDER)
            Nbatch = (prec_ == KUQCD_DOUBLE) ? 8 : 16;
        else
            Nbatch = 1;
        if (lat.nx % (2 * Nbatch) != 0)
            throw std::invalid_argument("LattBase: nx must be a multiple of 2 * Nbatch");
        latSupBulkFL = Lattice(lat.nx / Nbatch, lat.ny, lat.nz, lat.nt);
        latSupBulkCB = Lattice(lat.nx / Nbatch / 2, lat.ny, lat.nz, lat.nt);
        latSubFullFL = latSubBulkFL = latSupBulkFL;
        latSubFullCB = latSubBulkCB = latSupBulkCB;
        init_len();
    }
    LattBase &operator=(const LattBase &o)
    {
        latSupBulkFL = o.latSupBulkFL;
        latSupBulkCB = o.latSupBulkCB;
        latSubFullFL = o.latSubFullFL;
        latSubFullCB = o.latSubFullCB;
        latSubBulkFL = o.latSubBulkFL;
        latSubBulkCB = o.latSubBulkCB;
        gridDim = o.gridDim;
        haloDepth = o.haloDepth;
        Nbatch = o.Nbatch;
        myRank = o.myRank;
        np = o.np;
        prec = o.prec;
        gauge_order = o.gauge_order;
        rankOrder = o.rankOrder;
        site_subset = o.site_subset;
        lat = o.lat;
        init_len();
        return *this;
    }
    // sites of this field (all Nbatch lanes counted)
    size_t Sites() const { return site_subset == KUQCD_ALL ? lat.V : lat.V / 2; }

private:
    void init_len()
    {
        const bool pc = site_subset != KUQCD_ALL;
        latFull = pc ? &latSubFullCB : &latSubFullFL;
        latBulk = pc ? &latSubBulkCB : &latSubBulkFL;
        lenFull = latFull->V;
        lenBulk = latBulk->V;
    }
};

// communication.h -- the halo-exchange state of a field's lattice; one rank has
// no neighbours, so there is nothing to set up
class CommBase {
public:
    explicit CommBase(LattBase &latBase_) : latBase(latBase_) {}
    LattBase &latBase;
};

// spinor_field.h:20-40 -- a staggered colour vector on the sites of latBase
// (one parity for site_subset EVEN/ODD, both for ALL), in GPU memory.
template <typename T, int Nbatch, int Ns> class kuqcd_class_export SpinorField {
public:
    static_assert(Ns == 1, "staggered fields have one spin component");
    LattBase &latBase;

    SpinorField() = delete;
    explicit SpinorField(LattBase &latBase_);
    SpinorField(const SpinorField &o);  // device copy
    ~SpinorField();
    SpinorField &operator=(const SpinorField &o);  // device copy (same lattice)

    // host layout (kuqcd.cpp:129-169): h[2*(c + 3*i) + reim], i over the field's sites
    // (for ALL: i = parity * V/2 + checkerboard index)
    void SetValues(const T *h);
    void GetValues(T *h) const;
    void Clear();  // KMemory::Clear (kuqcd_memory.cpp:90-100): all zero
    int GetNbatch() const { return Nbatch; }

    // spinor_field.h:34-38: res = this * scalar, this -= v, res = this + v
    SpinorField operator*(T scalar);
    SpinorField &operator-=(const SpinorField &v);
    SpinorField operator+(const SpinorField &v);

    // the field's storage: 3 complex T per site, colour-major (c * sites + i)
    void *DevicePtr() const { return dev_; }
    size_t Sites() const { return sites_; }

private:
    void *dev_ = nullptr;
    size_t sites_ = 0;
    size_t bytes_ = 0;
};

template <typename T, int Nbatch, int Ns>
kuqcd_class_export SpinorField<T, Nbatch, Ns> operator-(const SpinorField<T, Nbatch, Ns> &a,
                                                        const SpinorField<T, Nbatch, Ns> &b);
// spinor_field.cpp:130-252
template <typename T, int Nbatch, int Ns> kuqcd_class_export T Norm2(const SpinorField<T, Nbatch, Ns> &v);
template <typename T, int Nbatch, int Ns> kuqcd_class_export T Norm(const SpinorField<T, Nbatch, Ns> &v);
template <typename T, int Nbatch, int Ns>
kuqcd_class_export T RealDot(const SpinorField<T, Nbatch, Ns> &bra, const SpinorField<T, Nbatch, Ns> &ket);
// y = a x - y
template <typename T, int Nbatch, int Ns>
kuqcd_class_export void Axmy(SpinorField<T, Nbatch, Ns> &y, const SpinorField<T, Nbatch, Ns> &x, T a);
// y = y + a x
template <typename T, int Nbatch, int Ns>
kuqcd_class_export void Axpy(SpinorField<T, Nbatch, Ns> &y, const SpinorField<T, Nbatch, Ns> &x, T a);
// y = y - a x
template <typename T, int Nbatch, int Ns>
kuqcd_class_export void Ymax(SpinorField<T, Nbatch, Ns> &y, const SpinorField<T, Nbatch, Ns> &x, T a);
// y = a x + b y
template <typename T, int Nbatch, int Ns>
kuqcd_class_export void Axpby(SpinorField<T, Nbatch, Ns> &y, const SpinorField<T, Nbatch, Ns> &x, T a, T b);

// gauge_field.h:21-45 -- one link field (fat or long) on all sites, in GPU memory
template <typename T, int Nbatch, KuQCDReconstructType Recon> class kuqcd_class_export GaugeField {
public:
    static_assert(Recon == KUQCD_GAUGE_RECON_NO, "links are loaded with all 18 reals (kuqcd.cpp:319-327)");
    LattBase &latBase;
    CommBase &comm;

    GaugeField() = delete;
    GaugeField(LattBase &latBase_, CommBase &comm_);
    GaugeField(const GaugeField &o);
    ~GaugeField();
    GaugeField &operator=(const GaugeField &) = delete;

    // host layout (kuqcd.cpp:103-127): h[2*(l + 9*(mu + 4*i)) + reim], i = parity * V/2 +
    // checkerboard index, l = 3 * row + col
    void SetValues(const T *h);
    // gauge_field.cpp:114-334: refresh the halos; one rank has none
    void Update(KuQCDHalo halo = KUQCD_COMMALL) { (void)halo; }

    void *DevicePtr() const { return dev_; }  // complex T [(q*4 + mu)*9 + l][V/2]

private:
    void *dev_ = nullptr;
    size_t bytes_ = 0;
};

// dirac.h:20-29
template <typename VecT> class Dirac {
public:
    KuQCDDslashType diracType = KUQCD_INVALID_DIRAC;
    explicit Dirac(KuQCDDslashType dt) : diracType(dt) {}
    virtual ~Dirac() {}
    virtual void Dslash(const VecT &src, VecT &dst, bool update) = 0;
    virtual void MdaggM(const VecT &src, VecT &dst, bool update) = 0;
};

// dirac_hisq.h:25-68 -- the HISQ hopping term of dirac_hisq.cpp:409-502 on the GPU:
//   D psi = sum_mu [ 1/2 (U psi(x+mu) - U^+ psi(x-mu)) + c3000 (N psi(x+3mu) - N^+ psi(x-3mu)) ]
// with c3000 = -1/48 (1 + naikEps); no mass term (the mass is stored, as in the reference, and
// enters a solve through the shifts). KUQCD_HISQPC_DIRAC: dst has one parity (EVEN/ODD), src the
// other. KUQCD_HISQ_DIRAC: both fields hold both parities (site_subset ALL), and D maps each
// parity of src to the other parity of dst. MdaggM(src, dst) = D(D src) through tmp (of the
// other parity; an internal field is used when tmp is null).
template <typename T, int Nbatch, KuQCDReconstructType ReconS, KuQCDReconstructType ReconN>
class kuqcd_class_export DiracHISQ : public Dirac<SpinorField<T, Nbatch, 1>> {
public:
    typedef SpinorField<T, Nbatch, 1> VecT;
    typedef GaugeField<T, Nbatch, ReconS> LinkST;
    typedef GaugeField<T, Nbatch, ReconN> LinkNT;
    const LinkST &gaugeSmeared;
    const LinkNT &gaugeNaik;
    VecT *tmp = nullptr;
    LattBase &latBase;
    double mass = 0;
    T mass2 = 0, c3000 = -1.0 / 48.0;
    int myRank = 0;
    int np = 1;

    DiracHISQ(KuQCDDslashType diracType, const LinkST &gaugeSmeared_, const LinkNT &gaugeNaik_, double mass_,
              VecT *tmp_ = nullptr, T naikEps = 0);
    ~DiracHISQ() override;
    DiracHISQ(const DiracHISQ &) = delete;
    DiracHISQ &operator=(const DiracHISQ &) = delete;
    void Dslash(const VecT &src, VecT &dst, bool update = false) override;
    void MdaggM(const VecT &src, VecT &dst, bool update = false) override;

    // the links in the kernel format, pre-multiplied by 1/2 (fat) and c3000 (long): which = 0 fat,
    // 1 long (18 reals), 2 long in fewer reals (14 in double, 12 in float; null when some long
    // link is not a multiple of an SU(3) matrix)
    const void *KernelLinks(int which) const { return which == 0 ? fat_ : which == 1 ? lng_ : lngr_; }

private:
    void *fat_ = nullptr, *lng_ = nullptr, *lngr_ = nullptr;
    size_t link_bytes_ = 0, lngr_bytes_ = 0;
    VecT *own_tmp_ = nullptr;
    LattBase *own_tmp_base_ = nullptr;
};

// inverter.h:20-33
template <typename T, typename VecT> class Inverter {
public:
    KuQCDInverterType inverterType;
    Dirac<VecT> *dirac = nullptr;
    Inverter(KuQCDInverterType it, Dirac<VecT> *dirac_) : inverterType(it), dirac(dirac_) {}
    virtual ~Inverter() {}
    virtual int Invert(int noShifts, T shifts[], const VecT &src, VecT *sol[], const double tol, int maxIters) = 0;
};

// inverter_cgm.h:21-28 -- multishift CG (inverter_cgm.cpp:17-87): solves
// (M + shifts[0]) x_0 = src and (M + shifts[0] + shifts[k]) x_k = src (k >= 1) with
// M = -MdaggM, stops shift k once Z_k^2 |r|^2 < tol (absolute), dropping converged shifts
// from the end; returns the iteration count. With this library's DiracHISQ the whole
// iteration runs on the GPU with scalars in T (the reference's recurrence); with any other
// Dirac it runs the same recurrence on the host over Dirac::MdaggM and the BLAS-1 functions.
template <typename T, typename VecT> class kuqcd_class_export CGM : public Inverter<T, VecT> {
public:
    CGM(KuQCDInverterType it, Dirac<VecT> *dirac_) : Inverter<T, VecT>(it, dirac_) {}
    ~CGM() override {}
    int Invert(int noShifts, T shifts[], const VecT &src, VecT *sol[], const double tol, int maxIters) override;
};

#define KUQCD_CLASSES_EXTERN(T, NB)                                                                          \
    extern template class SpinorField<T, NB, 1>;                                                             \
    extern template SpinorField<T, NB, 1> operator-(const SpinorField<T, NB, 1> &, const SpinorField<T, NB, 1> &); \
    extern template T Norm2(const SpinorField<T, NB, 1> &);                                                  \
    extern template T Norm(const SpinorField<T, NB, 1> &);                                                   \
    extern template T RealDot(const SpinorField<T, NB, 1> &, const SpinorField<T, NB, 1> &);                 \
    extern template void Axmy(SpinorField<T, NB, 1> &, const SpinorField<T, NB, 1> &, T);                    \
    extern template void Axpy(SpinorField<T, NB, 1> &, const SpinorField<T, NB, 1> &, T);                    \
    extern template void Ymax(SpinorField<T, NB, 1> &, const SpinorField<T, NB, 1> &, T);                    \
    extern template void Axpby(SpinorField<T, NB, 1> &, const SpinorField<T, NB, 1> &, T, T);                \
    extern template class GaugeField<T, NB, KUQCD_GAUGE_RECON_NO>;                                           \
    extern template class DiracHISQ<T, NB, KUQCD_GAUGE_RECON_NO, KUQCD_GAUGE_RECON_NO>;                      \
    extern template class CGM<T, SpinorField<T, NB, 1>>
KUQCD_CLASSES_EXTERN(float, 1);
KUQCD_CLASSES_EXTERN(double, 1);
KUQCD_CLASSES_EXTERN(float, 16);
KUQCD_CLASSES_EXTERN(double, 8);
#undef KUQCD_CLASSES_EXTERN

}  // namespace kuqcd

#pragma GCC visibility pop

#endif
