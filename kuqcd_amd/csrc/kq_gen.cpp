// Seeded synthetic gauge fields and sources, written in the reference's host
// layouts (kuqcd.cpp:103-169). A counter-based generator keyed by
// (seed, site, direction) makes every link independent of thread count and
// evaluation order, so the same seed gives the same field on any machine.
//
//   fat link   U_mu(x) = fat_scale * eta_mu(x) * G(seed_fat; x, mu)
//   long link  N_mu(x) = c_naik * eta_mu(x) * W_mu(x) W_mu(x+mu) W_mu(x+2mu),  W = G(seed_long; ., mu)
//   G = Haar-random SU(3) (kind 0) or exp(i eps H), H = sum_a h_a lambda_a / 2, h_a ~ N(0,1) (kind 1)
//   eta_x = 1, eta_y = (-1)^x, eta_z = (-1)^(x+y), eta_t = (-1)^(x+y+z)   (staggered phases)
//   source: complex Gaussian with E|z|^2 = 1 on the sites of one parity
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <thread>
#include <vector>

#include "kq_gen.h"

namespace kq {

using cd = std::complex<double>;

static inline uint64_t smix(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct CRng {
    uint64_t s;
    CRng(uint64_t seed, uint64_t stream, uint64_t idx) : s(smix(smix(seed * 0x9e3779b97f4a7c15ull + 1) ^ smix(stream + 0x1234567ull) ^ smix(idx))) {}
    double uniform()  // (0, 1)
    {
        s = smix(s);
        return ((double)(s >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    }
    // standard normal pair (Box-Muller)
    void normal2(double &a, double &b)
    {
        const double u1 = uniform(), u2 = uniform();
        const double r = std::sqrt(-2.0 * std::log(u1));
        a = r * std::cos(2.0 * M_PI * u2);
        b = r * std::sin(2.0 * M_PI * u2);
    }
};

typedef cd Mat3[9];  // row-major m[row*3+col]

static void haar_su3(CRng &g, Mat3 m)
{
    cd z[9];
    for (int k = 0; k < 9; ++k) {
        double a, b;
        g.normal2(a, b);
        z[k] = cd(a, b) * std::sqrt(0.5);
    }
    // modified Gram-Schmidt on the columns: Q R with R_ii > 0 real, so Q is
    // Haar on U(3) (Mezzadri 2007)
    cd q[9];
    for (int c = 0; c < 3; ++c) {
        cd v[3] = {z[c], z[3 + c], z[6 + c]};
        for (int p = 0; p < c; ++p) {
            cd d = 0;
            for (int r = 0; r < 3; ++r)
                d += std::conj(q[3 * r + p]) * v[r];
            for (int r = 0; r < 3; ++r)
                v[r] -= d * q[3 * r + p];
        }
        double n = std::sqrt(std::norm(v[0]) + std::norm(v[1]) + std::norm(v[2]));
        for (int r = 0; r < 3; ++r)
            q[3 * r + c] = v[r] / n;
    }
    // divide by det^(1/3) -> SU(3)
    cd det = q[0] * (q[4] * q[8] - q[5] * q[7]) - q[1] * (q[3] * q[8] - q[5] * q[6]) + q[2] * (q[3] * q[7] - q[4] * q[6]);
    const double ph = std::arg(det) / 3.0;
    const cd f = std::polar(1.0, -ph);
    for (int k = 0; k < 9; ++k)
        m[k] = q[k] * f;
}

static void matmul(const Mat3 a, const Mat3 b, Mat3 c)
{
    cd t[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            cd s = 0;
            for (int k = 0; k < 3; ++k)
                s += a[3 * i + k] * b[3 * k + j];
            t[3 * i + j] = s;
        }
    for (int k = 0; k < 9; ++k)
        c[k] = t[k];
}

// exp(i eps H) by scaling and squaring of a Taylor series
static void weak_su3(CRng &g, double eps, Mat3 m)
{
    double h[8];
    for (int k = 0; k < 8; k += 2)
        g.normal2(h[k], h[k + 1]);
    const double s3 = 1.0 / std::sqrt(3.0);
    // H = sum_a h_a lambda_a / 2 (Gell-Mann basis)
    cd H[9];
    H[0] = 0.5 * (h[2] + h[7] * s3);
    H[1] = 0.5 * cd(h[0], -h[1]);
    H[2] = 0.5 * cd(h[3], -h[4]);
    H[3] = 0.5 * cd(h[0], h[1]);
    H[4] = 0.5 * (-h[2] + h[7] * s3);
    H[5] = 0.5 * cd(h[5], -h[6]);
    H[6] = 0.5 * cd(h[3], h[4]);
    H[7] = 0.5 * cd(h[5], h[6]);
    H[8] = 0.5 * (-2.0 * h[7] * s3);
    const int nsq = 6;
    const double sc = eps / (double)(1 << nsq);
    cd A[9];
    for (int k = 0; k < 9; ++k)
        A[k] = cd(0, sc) * H[k];
    cd E[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, T[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int n = 1; n <= 14; ++n) {
        matmul(T, A, T);
        for (int k = 0; k < 9; ++k)
            T[k] /= (double)n;
        for (int k = 0; k < 9; ++k)
            E[k] += T[k];
    }
    for (int i = 0; i < nsq; ++i)
        matmul(E, E, E);
    for (int k = 0; k < 9; ++k)
        m[k] = E[k];
}

static void base_link(int kind, double eps, uint64_t seed, uint64_t site, int mu, Mat3 m)
{
    CRng g(seed, 7 + (uint64_t)mu, site);
    if (kind == KQ_GEN_HAAR)
        haar_su3(g, m);
    else
        weak_su3(g, eps, m);
}

struct LatIdx {
    int L[4];
    uint64_t V, Vh;
    explicit LatIdx(const unsigned lat[4])
    {
        for (int i = 0; i < 4; ++i)
            L[i] = (int)lat[i];
        V = (uint64_t)L[0] * L[1] * L[2] * L[3];
        Vh = V / 2;
    }
    // full coordinates of storage index i = q*Vh + s
    void coords(uint64_t i, int c[4]) const
    {
        const uint64_t q = i / Vh, s = i - q * Vh;
        const uint64_t lxh = (uint64_t)L[0] / 2;
        uint64_t r = s / lxh;
        const int xcb = (int)(s - r * lxh);
        c[1] = (int)(r % L[1]);
        r /= L[1];
        c[2] = (int)(r % L[2]);
        c[3] = (int)(r / L[2]);
        c[0] = 2 * xcb + ((c[1] + c[2] + c[3] + (int)q) & 1);
    }
    // storage index of full coordinates
    uint64_t index(const int c[4]) const
    {
        const int q = (c[0] + c[1] + c[2] + c[3]) & 1;
        const uint64_t lxh = (uint64_t)L[0] / 2;
        return (uint64_t)q * Vh + (uint64_t)(c[0] >> 1) + lxh * ((uint64_t)c[1] + (uint64_t)L[1] * ((uint64_t)c[2] + (uint64_t)L[2] * c[3]));
    }
};

static int eta(int mu, const int c[4])
{
    int e = 0;
    for (int nu = 0; nu < mu; ++nu)
        e += c[nu];
    return (e & 1) ? -1 : 1;
}

template <typename F> static void parallel_for(uint64_t n, F f)
{
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096)
        nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t]() {
            const uint64_t b = n * t / nt, e = n * (t + 1) / nt;
            for (uint64_t i = b; i < e; ++i)
                f(i);
        });
    for (auto &x : th)
        x.join();
}

template <typename HR>
int gen_gauge(HR *fat, HR *lng, const unsigned lat[4], int kind, double eps, double fat_scale, double c_naik,
              uint64_t seed_fat, uint64_t seed_long, int phases)
{
    for (int i = 0; i < 4; ++i)
        if (lat[i] == 0)
            return -1;
    if (lat[0] % 2)
        return -1;
    const LatIdx li(lat);
    parallel_for(li.V, [&](uint64_t i) {
        int c[4];
        li.coords(i, c);
        for (int mu = 0; mu < 4; ++mu) {
            const double e = phases ? (double)eta(mu, c) : 1.0;
            Mat3 m;
            if (fat) {
                base_link(kind, eps, seed_fat, li.index(c), mu, m);
                for (int l = 0; l < 9; ++l) {
                    const cd v = m[l] * (fat_scale * e);
                    fat[2 * (l + 9 * (mu + 4 * i))] = (HR)v.real();
                    fat[2 * (l + 9 * (mu + 4 * i)) + 1] = (HR)v.imag();
                }
            }
            if (lng) {
                Mat3 w0, w1, w2;
                int cc[4] = {c[0], c[1], c[2], c[3]};
                base_link(kind, eps, seed_long, li.index(cc), mu, w0);
                cc[mu] = (c[mu] + 1) % li.L[mu];
                base_link(kind, eps, seed_long, li.index(cc), mu, w1);
                cc[mu] = (c[mu] + 2) % li.L[mu];
                base_link(kind, eps, seed_long, li.index(cc), mu, w2);
                matmul(w0, w1, m);
                matmul(m, w2, m);
                for (int l = 0; l < 9; ++l) {
                    const cd v = m[l] * (c_naik * e);
                    lng[2 * (l + 9 * (mu + 4 * i))] = (HR)v.real();
                    lng[2 * (l + 9 * (mu + 4 * i)) + 1] = (HR)v.imag();
                }
            }
        }
    });
    return 0;
}

template <typename HR> int gen_source(HR *b, const unsigned lat[4], int parity, uint64_t seed)
{
    for (int i = 0; i < 4; ++i)
        if (lat[i] == 0)
            return -1;
    const LatIdx li(lat);
    parallel_for(li.Vh, [&](uint64_t s) {
        CRng g(seed, 3 + (uint64_t)parity, s);
        for (int c = 0; c < 3; ++c) {
            double a, bb;
            g.normal2(a, bb);
            b[6 * s + 2 * c] = (HR)(a * std::sqrt(0.5));
            b[6 * s + 2 * c + 1] = (HR)(bb * std::sqrt(0.5));
        }
    });
    return 0;
}

template int gen_gauge<double>(double *, double *, const unsigned *, int, double, double, double, uint64_t, uint64_t,
                               int);
template int gen_gauge<float>(float *, float *, const unsigned *, int, double, double, double, uint64_t, uint64_t, int);
template int gen_source<double>(double *, const unsigned *, int, uint64_t);
template int gen_source<float>(float *, const unsigned *, int, uint64_t);

}  // namespace kq
