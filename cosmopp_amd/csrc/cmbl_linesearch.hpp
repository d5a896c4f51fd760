// The More-Thuente line search as a host-side state machine, so that each
// trial's function value and slope can come from the device.  Shared by the
// device-resident L-BFGS (cmbl_loops.hip) and the cmbl_mt_* entry points.
#pragma once

// every operation rounded on its own, as in the reference's line search
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>

namespace cmbl {

// The line search of Moré and Thuente (ACM TOMS 20, 1994; MINPACK-2 dcsrch /
// dcstep) with the parameters and conventions of the reference's
// moreThuenteSearch (include/line_search.hpp:160-272): run as a state machine
// so that each trial's function value and slope can come from the device.
struct MoreThuente {
    double ftol = 0.1, gtol = 0.5, xtol = 1e-15, stpmin = 1e-15, stpmax = 1e15;
    int maxfev = 100;

    double finit = 0, dginit = 0, dgtest = 0, width = 0, width1 = 0;
    double stx = 0, fx = 0, dgx = 0, sty = 0, fy = 0, dgy = 0, stmin = 0, stmax = 0, stp = 0;
    bool brackt = false, stage1 = true;
    int infoc = 1, nfev = 0;

    // Interval update and trial step (dcstep).  Returns 0 when the inputs are
    // inconsistent (the step is then left as it is), else the case 1-4.
    static int step(double& stx, double& fx, double& dx, double& sty, double& fy, double& dy, double& stp,
                    double fp, double dp, bool& brackt, double lo, double hi)
    {
        const double p66 = 0.66;
        if ((brackt && (stp <= std::min(stx, sty) || stp >= std::max(stx, sty))) || dx * (stp - stx) >= 0 || hi < lo)
            return 0;
        const double sgnd = dp * (dx / std::fabs(dx));
        int kase;
        bool bound;
        double stpf;
        double theta, s, gamma, p, q, r, stpc, stpq;
        if (fp > fx) {
            // higher function value: the minimum is bracketed; cubic, or the
            // midpoint of cubic and quadratic steps
            kase = 1;
            bound = true;
            theta = 3 * (fx - fp) / (stp - stx) + dx + dp;
            s = std::max(std::max(std::fabs(theta), std::fabs(dx)), std::fabs(dp));
            gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
            if (stp < stx) gamma = -gamma;
            p = (gamma - dx) + theta;
            q = ((gamma - dx) + gamma) + dp;
            r = p / q;
            stpc = stx + r * (stp - stx);
            stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2) * (stp - stx);
            if (std::fabs(stpc - stx) < std::fabs(stpq - stx))
                stpf = stpc;
            else
                stpf = stpc + (stpq - stpc) / 2;
            brackt = true;
        } else if (sgnd < 0) {
            // derivatives of opposite sign: bracketed; cubic or secant step
            kase = 2;
            bound = false;
            theta = 3 * (fx - fp) / (stp - stx) + dx + dp;
            s = std::max(std::max(std::fabs(theta), std::fabs(dx)), std::fabs(dp));
            gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
            if (stp > stx) gamma = -gamma;
            p = (gamma - dp) + theta;
            q = ((gamma - dp) + gamma) + dx;
            r = p / q;
            stpc = stp + r * (stx - stp);
            stpq = stp + (dp / (dp - dx)) * (stx - stp);
            if (std::fabs(stpc - stp) > std::fabs(stpq - stp))
                stpf = stpc;
            else
                stpf = stpq;
            brackt = true;
        } else if (std::fabs(dp) < std::fabs(dx)) {
            // same sign, decreasing magnitude: cubic step only if it tends
            // to infinity in the step direction, else the secant step
            kase = 3;
            bound = true;
            theta = 3 * (fx - fp) / (stp - stx) + dx + dp;
            s = std::max(std::max(std::fabs(theta), std::fabs(dx)), std::fabs(dp));
            gamma = s * std::sqrt(std::max(0.0, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
            if (stp > stx) gamma = -gamma;
            p = (gamma - dp) + theta;
            q = (gamma + (dx - dp)) + gamma;
            r = p / q;
            if (r < 0 && gamma != 0)
                stpc = stp + r * (stx - stp);
            else if (stp > stx)
                stpc = hi;
            else
                stpc = lo;
            stpq = stp + (dp / (dp - dx)) * (stx - stp);
            if (brackt)
                stpf = std::fabs(stp - stpc) < std::fabs(stp - stpq) ? stpc : stpq;
            else
                stpf = std::fabs(stp - stpc) > std::fabs(stp - stpq) ? stpc : stpq;
        } else {
            // same sign, no decrease in magnitude: cubic step to the other end
            // of the bracket, or to the step bound
            kase = 4;
            bound = false;
            if (brackt) {
                theta = 3 * (fp - fy) / (sty - stp) + dy + dp;
                s = std::max(std::max(std::fabs(theta), std::fabs(dy)), std::fabs(dp));
                gamma = s * std::sqrt((theta / s) * (theta / s) - (dy / s) * (dp / s));
                if (stp > sty) gamma = -gamma;
                p = (gamma - dp) + theta;
                q = ((gamma - dp) + gamma) + dy;
                r = p / q;
                stpc = stp + r * (sty - stp);
                stpf = stpc;
            } else if (stp > stx) {
                stpf = hi;
            } else {
                stpf = lo;
            }
        }
        // new interval of uncertainty
        if (fp > fx) {
            sty = stp;
            fy = fp;
            dy = dp;
        } else {
            if (sgnd < 0) {
                sty = stx;
                fy = fx;
                dy = dx;
            }
            stx = stp;
            fx = fp;
            dx = dp;
        }
        stpf = std::min(hi, stpf);
        stpf = std::max(lo, stpf);
        stp = stpf;
        if (brackt && bound) {
            if (sty > stx)
                stp = std::min(stx + p66 * (sty - stx), stp);
            else
                stp = std::max(stx + p66 * (sty - stx), stp);
        }
        return kase;
    }

    // The trial step for the next evaluation: the interval ends for the
    // current state and the safeguards (csrch's loop head).
    double next_trial()
    {
        const int xtrapf = 4;
        if (brackt) {
            stmin = std::min(stx, sty);
            stmax = std::max(stx, sty);
        } else {
            stmin = stx;
            stmax = stp + xtrapf * (stp - stx);
        }
        stp = std::max(stp, stpmin);
        stp = std::min(stp, stpmax);
        if ((brackt && (stp <= stmin || stp >= stmax)) || nfev >= maxfev - 1 || infoc == 0 ||
            (brackt && stmax - stmin <= xtol * stmax))
            stp = stx;
        return stp;
    }

    // Start a search from f0 with slope dg0 < 0 and initial step stp0;
    // returns the first trial step.
    double begin(double f0, double dg0, double stp0)
    {
        finit = f0;
        dginit = dg0;
        dgtest = ftol * dginit;
        width = stpmax - stpmin;
        width1 = 2 * width;
        stx = 0;
        fx = finit;
        dgx = dginit;
        sty = 0;
        fy = finit;
        dgy = dginit;
        brackt = false;
        stage1 = true;
        infoc = 1;
        nfev = 0;
        stp = stp0;
        return next_trial();
    }

    // Value f and slope dg at the last trial step.  Returns the search's
    // result code (1: sufficient decrease and curvature; 2-6: stopped by the
    // tolerances, bounds or evaluation count) or 0 with *trial set to the
    // next step to evaluate.
    int feed(double f, double dg, double* trial)
    {
        const double p66 = 0.66;
        ++nfev;
        const double ftest1 = finit + stp * dgtest;
        int info = 0;
        if ((brackt && (stp <= stmin || stp >= stmax)) || infoc == 0) info = 6;
        if (stp == stpmax && f <= ftest1 && dg <= dgtest) info = 5;
        if (stp == stpmin && (f > ftest1 || dg >= dgtest)) info = 4;
        if (nfev >= maxfev) info = 3;
        if (brackt && stmax - stmin <= xtol * stmax) info = 2;
        if (f <= ftest1 && std::fabs(dg) <= gtol * (-dginit)) info = 1;
        if (info != 0) return info;
        if (stage1 && f <= ftest1 && dg >= std::min(ftol, gtol) * dginit) stage1 = false;
        if (stage1 && f <= fx && f > ftest1) {
            // the modified function psi(stp) = f(stp) - stp * dgtest
            const double fm = f - stp * dgtest;
            double fxm = fx - stx * dgtest, fym = fy - sty * dgtest;
            const double dgm = dg - dgtest;
            double dgxm = dgx - dgtest, dgym = dgy - dgtest;
            infoc = step(stx, fxm, dgxm, sty, fym, dgym, stp, fm, dgm, brackt, stmin, stmax);
            fx = fxm + stx * dgtest;
            fy = fym + sty * dgtest;
            dgx = dgxm + dgtest;
            dgy = dgym + dgtest;
        } else {
            infoc = step(stx, fx, dgx, sty, fy, dgy, stp, f, dg, brackt, stmin, stmax);
        }
        if (brackt) {
            // force a sufficient shrink of the bracket, else bisect
            if (std::fabs(sty - stx) >= p66 * width1) stp = stx + (sty - stx) / 2;
            width1 = width;
            width = std::fabs(sty - stx);
        }
        *trial = next_trial();
        return 0;
    }
};

}  // namespace cmbl
