// The multishift CG recurrence of CGM::Invert (inverter_cgm.cpp:17-87) driven from the host in
// the reference's scalar type T, over any operator and field storage. Ops supplies
//   T    norm2_r()                       |r|^2
//   T    apply(T sigma0)                 ap = sigma0 p_0 - D D p_0, returns <p_0, ap>
//   void update_r(T b0)                  r += b0 ap
//   void update_x(int k, T bk)           x_k -= bk p_k
//   void update_p(int k, T zk, T ak)     p_k = zk r + ak p_k
// and holds r = p_k = b, x_k = 0 on entry. Shifts are in the reference's convention (shift[0]
// absolute, shift[k >= 1] relative to it); tol is the absolute squared stop threshold. Returns
// the iteration count.
#pragma once

#include <vector>

namespace kq {

template <typename T, class Ops> int cgm_recurrence(Ops &ops, int ns, const T *shift, double tol, int max_iter)
{
    struct Shift {
        T alpha = 0, beta = 1, zeta = 1, zeta_prev = 1;
    };
    std::vector<Shift> sh(ns);
    T rr = ops.norm2_r(), beta_prev = 1.0;
    int live = ns, iter = 0;
    do {
        ++iter;
        const T pap = ops.apply(shift[0]);
        Shift &s0 = sh[0];
        s0.beta = -rr / pap;
        ops.update_r(s0.beta);
        for (int k = 1; k < live; ++k) {
            Shift &s = sh[k];
            // the double literal keeps the reference's promotion for T = float (inverter_cgm.cpp:52)
            const T ratio = beta_prev * s.zeta_prev /
                            (s0.beta * s0.alpha * (s.zeta_prev - s.zeta) + s.zeta_prev * beta_prev * (1.0 - shift[k] * s0.beta));
            s.zeta_prev = s.zeta;
            s.zeta = s.zeta * ratio;
            s.beta = s0.beta * ratio;
        }
        beta_prev = s0.beta;
        const T rr_new = ops.norm2_r();
        s0.alpha = rr_new / rr;
        rr = rr_new;
        for (int k = 0; k < live; ++k)
            ops.update_x(k, sh[k].beta);
        for (int k = 1; k < live; ++k)
            sh[k].alpha = s0.alpha * sh[k].zeta * sh[k].beta / (sh[k].zeta_prev * s0.beta);
        for (int k = 0; k < live; ++k)
            ops.update_p(k, sh[k].zeta, sh[k].alpha);
        // drop converged shifts from the end (inverter_cgm.cpp:71-76)
        while (live > 0 && sh[live - 1].zeta * sh[live - 1].zeta * rr_new < tol)
            --live;
    } while (live > 0 && iter < max_iter);
    return iter;
}

}  // namespace kq
