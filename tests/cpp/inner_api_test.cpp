// A C++ caller of the inner classes that include/kuqcd.h declares (Lattice,
// LattBase, GaugeField, SpinorField, DiracHISQ, CGM and the BLAS-1 functions),
// compiled with g++ against the header alone. It reads a problem written by
// tests/test_gpu_inner_api.py, applies the operator and the solver through the
// classes, and writes the results back for the test to check against the oracle.
//
// usage: inner_api_test double|float16 <input> <output>
//   double : T = double, Nbatch 1 (NORMAL order)
//   float16: T = float, Nbatch 16 (BATCH order; nx a multiple of 32)
//
// input : int32 lat[4], int32 ns, int32 max_iter, double tol, double shifts[ns],
//         T fat[V*72], T lng[V*72], T b_even[Vh*6], T src_odd[Vh*6], T full[V*6]
// output: T dslash_even[Vh*6], T mdaggm_even[Vh*6], double norm2_b, double dot_b_mdaggm,
//         T axpby[Vh*6] (y = 0.75 b - 1.25 mdaggm), int32 iter, T sol[ns][Vh*6],
//         int32 iter_generic, T sol_generic[ns][Vh*6], T full_dslash[V*6]
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "kuqcd.h"

using namespace kuqcd;

namespace {

struct Reader {
    FILE *f;
    template <typename X> void get(X *p, size_t n)
    {
        if (std::fread(p, sizeof(X), n, f) != n)
            throw std::runtime_error("short input file");
    }
};

struct Writer {
    FILE *f;
    template <typename X> void put(const X *p, size_t n)
    {
        if (std::fwrite(p, sizeof(X), n, f) != n)
            throw std::runtime_error("cannot write output");
    }
};

// a caller's own operator: forwards to another Dirac (so CGM runs its host-driven recurrence)
template <typename VecT> class Forward : public Dirac<VecT> {
public:
    explicit Forward(Dirac<VecT> *d) : Dirac<VecT>(d->diracType), inner(d) {}
    void Dslash(const VecT &src, VecT &dst, bool update) override { inner->Dslash(src, dst, update); }
    void MdaggM(const VecT &src, VecT &dst, bool update) override { inner->MdaggM(src, dst, update); }
    Dirac<VecT> *inner;
};

template <typename T, int NB> int run(const char *in_path, const char *out_path)
{
    FILE *fi = std::fopen(in_path, "rb");
    if (!fi)
        return 3;
    Reader rd{fi};
    int32_t lat[4], ns, max_iter;
    double tol;
    rd.get(lat, 4);
    rd.get(&ns, 1);
    rd.get(&max_iter, 1);
    rd.get(&tol, 1);
    std::vector<double> shifts_d(ns);
    rd.get(shifts_d.data(), ns);
    const size_t V = (size_t)lat[0] * lat[1] * lat[2] * lat[3], Vh = V / 2;
    std::vector<T> fat(V * 72), lng(V * 72), b(Vh * 6), src(Vh * 6), full(V * 6);
    rd.get(fat.data(), fat.size());
    rd.get(lng.data(), lng.size());
    rd.get(b.data(), b.size());
    rd.get(src.data(), src.size());
    rd.get(full.data(), full.size());
    std::fclose(fi);

    if (kuqcd_init() != KUQCD_SUCCESS)
        return 4;
    const KuQCDPrecision prec = sizeof(T) == 8 ? KUQCD_DOUBLE : KUQCD_SINGLE;
    const KuQCDGaugeOrder order = NB == 1 ? KUQCD_NORMAL_ORDER : KUQCD_BATCH_ORDER;
    const std::array<unsigned short, 4> grid{1, 1, 1, 1};
    // the declarations of the reference's test_solver.cpp:40-43
    Lattice latSup(lat[0], lat[1], lat[2], lat[3]);
    LattBase *latBaseOdd = nullptr;
    LattBase gaugeBase(latSup, grid, 3, order, KUQCD_TZYX_ORDER, prec, KUQCD_ALL);
    LattBase evenBase(latSup, grid, 0, order, KUQCD_TZYX_ORDER, prec, KUQCD_EVEN);
    latBaseOdd = new LattBase(latSup, grid, 0, order, KUQCD_TZYX_ORDER, prec, KUQCD_ODD);
    LattBase fullBase(latSup, grid, 0, order, KUQCD_TZYX_ORDER, prec, KUQCD_ALL);
    CommBase comm(gaugeBase);

    GaugeField<T, NB, KUQCD_GAUGE_RECON_NO> gFat(gaugeBase, comm), gLong(gaugeBase, comm);
    gFat.SetValues(fat.data());
    gLong.SetValues(lng.data());
    SpinorField<T, NB, 1> tmp(*latBaseOdd);
    DiracHISQ<T, NB, KUQCD_GAUGE_RECON_NO, KUQCD_GAUGE_RECON_NO> dirac(KUQCD_HISQPC_DIRAC, gFat, gLong, 0.0, &tmp);

    FILE *fo = std::fopen(out_path, "wb");
    if (!fo)
        return 5;
    Writer wr{fo};
    std::vector<T> h(Vh * 6);

    // hopping term: odd source -> even sites (dirac_hisq.cpp:409-455)
    SpinorField<T, NB, 1> so(*latBaseOdd), de(evenBase);
    so.SetValues(src.data());
    dirac.Dslash(so, de, false);
    de.GetValues(h.data());
    wr.put(h.data(), h.size());

    // D D on the even source (dirac_hisq.cpp:457-462)
    SpinorField<T, NB, 1> be(evenBase), me(evenBase);
    be.SetValues(b.data());
    dirac.MdaggM(be, me, false);
    me.GetValues(h.data());
    wr.put(h.data(), h.size());

    // BLAS-1 (spinor_field.cpp:130-252)
    const double n2 = (double)Norm2(be), dt = (double)RealDot(be, me);
    wr.put(&n2, 1);
    wr.put(&dt, 1);
    SpinorField<T, NB, 1> y(me);
    Axpby(y, be, (T)0.75, (T)-1.25);
    y.GetValues(h.data());
    wr.put(h.data(), h.size());

    // multishift CG on the GPU runner (inverter_cgm.cpp:17-87)
    std::vector<T> shifts(ns);
    for (int k = 0; k < ns; ++k)
        shifts[k] = (T)shifts_d[k];
    std::vector<std::unique_ptr<SpinorField<T, NB, 1>>> sol, solg;
    std::vector<SpinorField<T, NB, 1> *> sp, spg;
    for (int k = 0; k < ns; ++k) {
        sol.emplace_back(new SpinorField<T, NB, 1>(evenBase));
        solg.emplace_back(new SpinorField<T, NB, 1>(evenBase));
        sp.push_back(sol.back().get());
        spg.push_back(solg.back().get());
    }
    CGM<T, SpinorField<T, NB, 1>> cgm(KUQCD_CGM, &dirac);
    const int32_t iter = cgm.Invert(ns, shifts.data(), be, sp.data(), tol, max_iter);
    wr.put(&iter, 1);
    for (int k = 0; k < ns; ++k) {
        sp[k]->GetValues(h.data());
        wr.put(h.data(), h.size());
    }
    // the same recurrence driven from the host over a caller's operator
    Forward<SpinorField<T, NB, 1>> fwd(&dirac);
    CGM<T, SpinorField<T, NB, 1>> cgm2(KUQCD_CGM, &fwd);
    const int32_t iter2 = cgm2.Invert(ns, shifts.data(), be, spg.data(), tol, max_iter);
    wr.put(&iter2, 1);
    for (int k = 0; k < ns; ++k) {
        spg[k]->GetValues(h.data());
        wr.put(h.data(), h.size());
    }

    // the operator on both parities (KUQCD_HISQ_DIRAC with full-lattice fields)
    DiracHISQ<T, NB, KUQCD_GAUGE_RECON_NO, KUQCD_GAUGE_RECON_NO> full_op(KUQCD_HISQ_DIRAC, gFat, gLong, 0.0);
    SpinorField<T, NB, 1> fs(fullBase), fd(fullBase);
    fs.SetValues(full.data());
    full_op.Dslash(fs, fd, false);
    std::vector<T> hf(V * 6);
    fd.GetValues(hf.data());
    wr.put(hf.data(), hf.size());
    std::fclose(fo);

    sol.clear();
    solg.clear();
    delete latBaseOdd;
    std::printf("inner API: CGM %d iterations on the device runner, %d host-driven\n", iter, iter2);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s double|float16 <input> <output>\n", argv[0]);
        return 2;
    }
    try {
        const std::string mode = argv[1];
        int rc;
        if (mode == "double")
            rc = run<double, 1>(argv[2], argv[3]);
        else if (mode == "float16")
            rc = run<float, 16>(argv[2], argv[3]);
        else
            return 2;
        kuqcd_gauge_free();
        kuqcd_end();
        return rc;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "inner API test failed: %s\n", e.what());
        return 1;
    }
}
