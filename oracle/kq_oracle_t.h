/*
 * Precision-generic body of the oracle (included twice by kq_oracle.c, once with
 * REAL=double / SFX=_d and once with REAL=float / SFX=_f).
 *
 * TEST INFRASTRUCTURE ONLY: this is the CPU checker for the HISQ even-odd
 * multishift CG path. It is never linked into the product library.
 *
 * All arrays use the reference's host layouts:
 *   gauge  : g[2*(l + 9*(mu + 4*i)) + ri], i = q*Vh + s (even block first),
 *            l = 3*row + col                     (kuqcd.cpp:103-127)
 *   spinor : v[2*(c + 3*s) + ri], s = checkerboard index of one parity
 *                                                (kuqcd.cpp:129-169)
 * Every REAL operation is done in REAL, as the reference does for its
 * template parameter T (float for KUQCD_SINGLE, double for KUQCD_DOUBLE).
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* res += U * (a*v)   -- UMultV applied to a pre-scaled spinor and accumulated
 * exactly as DiracHISQ::FwdDslash does (dirac_hisq.cpp:473, 480):
 * the spinor is scaled first (ColorSpinor * T), the 3x3 product is formed in a
 * fresh zero spinor summing col = 0..2 (kuqcd_math.cpp:20-35), then added. */
static void FN(umv_acc)(const REAL *u, const REAL *v, REAL a, REAL *res, int sign)
{
    REAL w[6];
    for (int k = 0; k < 6; ++k)
        w[k] = v[k] * a;
    for (int row = 0; row < 3; ++row) {
        REAL re = 0, im = 0;
        for (int col = 0; col < 3; ++col) {
            const REAL ur = u[2 * (col + 3 * row)], ui = u[2 * (col + 3 * row) + 1];
            re += ur * w[2 * col] - ui * w[2 * col + 1];
            im += ur * w[2 * col + 1] + ui * w[2 * col];
        }
        if (sign > 0) {
            res[2 * row] += re;
            res[2 * row + 1] += im;
        } else {
            res[2 * row] -= re;
            res[2 * row + 1] -= im;
        }
    }
}

/* res -= U^dagger * (a*v)  -- UDaggMultV (kuqcd_math.cpp:91-106) as used by
 * DiracHISQ::BwdDslash (dirac_hisq.cpp:493, 500). */
static void FN(udmv_sub)(const REAL *u, const REAL *v, REAL a, REAL *res)
{
    REAL w[6];
    for (int k = 0; k < 6; ++k)
        w[k] = v[k] * a;
    for (int row = 0; row < 3; ++row) {
        REAL re = 0, im = 0;
        for (int col = 0; col < 3; ++col) {
            const REAL ur = u[2 * (row + 3 * col)], ui = u[2 * (row + 3 * col) + 1];
            re += ur * w[2 * col] + ui * w[2 * col + 1];
            im += ur * w[2 * col + 1] - ui * w[2 * col];
        }
        res[2 * row] -= re;
        res[2 * row + 1] -= im;
    }
}

/* Even-odd HISQ hopping term, dirac_hisq.cpp:409-502:
 *   dst(x) = sum_mu [ 1/2 U_mu(x) src(x+mu) + c3000 N_mu(x) src(x+3mu) ]
 *          - sum_mu [ 1/2 U_mu(x-mu)^+ src(x-mu) + c3000 N_mu(x-3mu)^+ src(x-3mu) ]
 * with c3000 = -1/48 (dirac_hisq.cpp:40), dst on parity dst_parity, src on the
 * other parity. Accumulation order: forward mu=0..3 (fat then long), then
 * backward mu=0..3 (fat then long) (dirac_hisq.cpp:441-449). */
void FN(kqo_dslash)(const REAL *fat, const REAL *lng, const REAL *src, REAL *dst, const int L[4], int dst_parity)
{
    const long V = (long)L[0] * L[1] * L[2] * L[3];
    const long Vh = V / 2;
    const long gdst = dst_parity ? Vh : 0; /* link block of the output parity (forward hops) */
    const long gsrc = dst_parity ? 0 : Vh; /* link block of the input parity (backward hops) */
    const REAL half = (REAL)0.5;
    const REAL c3 = (REAL)(-1.0 / 48.0);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < Vh; ++s) {
        REAL res[6] = {0, 0, 0, 0, 0, 0};
        for (int mu = 0; mu < 4; ++mu) {
            long n1 = kqo_nb(L, dst_parity, s, mu, +1);
            FN(umv_acc)(fat + 2 * 9 * (mu + 4 * (gdst + s)), src + 6 * n1, half, res, +1);
            long n3 = kqo_nb(L, dst_parity, s, mu, +3);
            FN(umv_acc)(lng + 2 * 9 * (mu + 4 * (gdst + s)), src + 6 * n3, c3, res, +1);
        }
        for (int mu = 0; mu < 4; ++mu) {
            long n1 = kqo_nb(L, dst_parity, s, mu, -1);
            FN(udmv_sub)(fat + 2 * 9 * (mu + 4 * (gsrc + n1)), src + 6 * n1, half, res);
            long n3 = kqo_nb(L, dst_parity, s, mu, -3);
            FN(udmv_sub)(lng + 2 * 9 * (mu + 4 * (gsrc + n3)), src + 6 * n3, c3, res);
        }
        for (int k = 0; k < 6; ++k)
            dst[6 * s + k] = res[k];
    }
}

/* MdaggM (dirac_hisq.cpp:457-462): dst = D_{p,pbar} D_{pbar,p} src, via tmp on
 * the opposite parity. */
void FN(kqo_mdaggm)(const REAL *fat, const REAL *lng, const REAL *src, REAL *dst, REAL *tmp, const int L[4],
                    int parity)
{
    FN(kqo_dslash)(fat, lng, src, tmp, L, 1 - parity);
    FN(kqo_dslash)(fat, lng, tmp, dst, L, parity);
}

/* ---- BLAS-1, spinor_field.cpp:130-252 (n = number of REALs = 6*Vh) ---- */
void FN(kqo_axmy)(REAL *y, const REAL *x, REAL a, long n) /* y = a*x - y */
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i)
        y[i] = a * x[i] - y[i];
}
void FN(kqo_axpy)(REAL *y, const REAL *x, REAL a, long n) /* y += a*x */
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i)
        y[i] += a * x[i];
}
void FN(kqo_ymax)(REAL *y, const REAL *x, REAL a, long n) /* y -= a*x */
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i)
        y[i] -= a * x[i];
}
void FN(kqo_axpby)(REAL *y, const REAL *x, REAL a, REAL b, long n) /* y = a*x + b*y */
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i)
        y[i] = a * x[i] + b * y[i];
}
/* Sum of x[i]*y[i] accumulated in REAL. The reference uses an OpenMP
 * reduction in T (spinor_field.cpp:192-246), whose order depends on the
 * thread count; here blocks of KQO_BLK products are summed in parallel and
 * the block sums are added in order, so the result is the same for any
 * thread count. */
REAL FN(kqo_realdot)(const REAL *x, const REAL *y, long n)
{
    const long nb = (n + KQO_BLK - 1) / KQO_BLK;
    REAL *part = (REAL *)malloc(sizeof(REAL) * (size_t)(nb > 0 ? nb : 1));
#pragma omp parallel for schedule(static)
    for (long b = 0; b < nb; ++b) {
        REAL acc = 0;
        long e = (b + 1) * KQO_BLK < n ? (b + 1) * KQO_BLK : n;
        for (long i = b * KQO_BLK; i < e; ++i)
            acc += x[i] * y[i];
        part[b] = acc;
    }
    REAL res = 0;
    for (long b = 0; b < nb; ++b)
        res += part[b];
    free(part);
    return res;
}
REAL FN(kqo_norm2)(const REAL *x, long n)
{
    return FN(kqo_realdot)(x, x, n);
}

/* D D over fields of both parities (even block, then odd): the full operator of a
 * KUQCD_HISQ_DIRAC DiracHISQ (dirac_hisq.cpp:409-462 with offseti = offsetg = 0) maps the odd
 * block to the even one and back, so its square is MdaggM of each block on its own parity. */
static void FN(mdaggm_any)(const REAL *fat, const REAL *lng, const REAL *src, REAL *dst, REAL *tmp, const int L[4],
                           int parity)
{
    const long n6 = 6L * L[0] * L[1] * L[2] * L[3] / 2;
    if (parity >= 0) {
        FN(kqo_mdaggm)(fat, lng, src, dst, tmp, L, parity);
        return;
    }
    FN(kqo_mdaggm)(fat, lng, src, dst, tmp, L, 0);
    FN(kqo_mdaggm)(fat, lng, src + n6, dst + n6, tmp + n6, L, 1);
}

/* CGM::Invert, inverter_cgm.cpp:17-87: Jegerlehner multishift CG on
 *   A_0 = shifts[0] - D D,  A_k = shifts[0] + shifts[k] - D D (k >= 1),
 * over fields of one parity (parity 0 or 1) or of both (parity -1, mdaggm_any),
 * scalars in REAL, stop when zeta_k^2 |r|^2 < tol for the last active shift
 * (absolute squared residual), converged shifts dropped from the end.
 * The literal 1.0 in the zeta recurrence is a double, as in the reference
 * (inverter_cgm.cpp:52), so that denominator is formed in double for REAL=float.
 * x[k] are outputs (cleared here). Returns the iteration count. */
int FN(kqo_cgm)(const REAL *fat, const REAL *lng, const int L[4], int parity, int nshift, const REAL *shifts,
                const REAL *b, REAL **x, double tol, int max_iter)
{
    const long Vh = (long)L[0] * L[1] * L[2] * L[3] / 2;
    const long n = (parity < 0 ? 12 : 6) * Vh;
    const size_t bytes = sizeof(REAL) * (size_t)n;
    int maxTerm = nshift;
    REAL *a = (REAL *)calloc((size_t)nshift, sizeof(REAL));
    REAL *B = (REAL *)calloc((size_t)nshift, sizeof(REAL));
    REAL *Z = (REAL *)calloc((size_t)nshift, sizeof(REAL));
    REAL *Zm1 = (REAL *)calloc((size_t)nshift, sizeof(REAL));
    REAL **p = (REAL **)malloc(sizeof(REAL *) * (size_t)nshift);
    REAL *s = (REAL *)malloc(bytes), *r = (REAL *)malloc(bytes), *pi0 = (REAL *)malloc(bytes);
    REAL *tmp = (REAL *)malloc(bytes);
    for (int i = 0; i < nshift; ++i) {
        p[i] = (REAL *)malloc(bytes);
        memcpy(p[i], b, bytes);
        a[i] = 0;
        B[i] = 1;
        Z[i] = 1;
        Zm1[i] = 1;
        memset(x[i], 0, bytes);
    }
    memcpy(r, b, bytes);
    REAL rNorm2 = FN(kqo_norm2)(r, n);
    REAL pAp, lambda, lambda2, rr1, Bm1 = 1;
    int iter = 0;
    do {
        iter++;
        memcpy(pi0, p[0], bytes);
        FN(mdaggm_any)(fat, lng, pi0, s, tmp, L, parity);
        FN(kqo_axmy)(s, pi0, shifts[0], n);
        pAp = FN(kqo_realdot)(pi0, s, n);
        B[0] = -rNorm2 / pAp;
        FN(kqo_axpy)(r, s, B[0], n);
        for (int j = 1; j < maxTerm; ++j) {
            rr1 = Bm1 * Zm1[j] / (B[0] * a[0] * (Zm1[j] - Z[j]) + Zm1[j] * Bm1 * (1.0 - shifts[j] * B[0]));
            Zm1[j] = Z[j];
            Z[j] = Z[j] * rr1;
            B[j] = B[0] * rr1;
        }
        Bm1 = B[0];
        lambda2 = FN(kqo_norm2)(r, n);
        a[0] = lambda2 / rNorm2;
        rNorm2 = lambda2;
        for (int i = 0; i < maxTerm; ++i)
            FN(kqo_ymax)(x[i], p[i], B[i], n);
        for (int j = 1; j < maxTerm; ++j)
            a[j] = a[0] * Z[j] * B[j] / (Zm1[j] * B[0]);
        for (int i = 0; i < maxTerm; ++i)
            FN(kqo_axpby)(p[i], r, Z[i], a[i], n);
        do {
            lambda = Z[maxTerm - 1] * Z[maxTerm - 1] * lambda2;
            if (lambda < tol)
                maxTerm--;
        } while ((maxTerm > 0) && (lambda < tol));
    } while ((maxTerm > 0) && (iter < max_iter));
    for (int i = 0; i < nshift; ++i)
        free(p[i]);
    free(p);
    free(s);
    free(r);
    free(pi0);
    free(tmp);
    free(a);
    free(B);
    free(Z);
    free(Zm1);
    return iter;
}

#undef FN
#undef CAT
#undef CAT2
