// adj_check.cpp -- CPU transpose-identity checks of the hand-written adjoint (adjoint.hpp):
// for a point function y = f(x), <J v, w> from dual-number forward mode equals <v, J^T w> from
// the adjoint, for random x, v, w.  Built by tests/native/Makefile; used by the CPU tests.
#include "../../oracle/dual.hpp"
#include "../../glmmhdturb_amd/csrc/adjoint.hpp"

#include <cmath>
#include <cstdint>
#include <random>
#include <algorithm>
#include <vector>

using namespace glm;

namespace {

template <typename T>
Phys<T> phys(double gamma, double ch, double llf_d, double llf_p)
{
    Phys<T> s;
    s.gamma = T(gamma);
    s.igm1 = T(1.0 / (gamma - 1.0));
    s.smallr = T(1e-6);
    s.pfl = T(1e-6 * 1e-8 / gamma);
    s.ch = T(ch);
    s.ch2 = T(ch * ch);
    s.half_ich = T(0.5 / ch);
    s.llf_dmin = T(llf_d);
    s.llf_pmin = T(llf_p);
    return s;
}

struct Rng {
    std::mt19937_64 g;
    explicit Rng(uint64_t seed) : g(seed) {}
    double u(double a, double b) { return std::uniform_real_distribution<double>(a, b)(g); }
    double n() { return std::normal_distribution<double>(0.0, 1.0)(g); }
};

// a physical primitive state: rho, v, p, B, psi
void prim(Rng& r, double q[9], double bscale)
{
    q[0] = r.u(0.3, 2.0);
    for (int i = 1; i <= 3; ++i) q[i] = r.u(-1.5, 1.5);
    q[4] = r.u(0.2, 2.0);
    for (int i = 5; i <= 7; ++i) q[i] = bscale * r.u(-1.0, 1.0);
    q[8] = 0.05 * r.u(-1.0, 1.0);
}

void prim2cons_d(const double q[9], double c[9], double gamma)
{
    Phys<double> s = phys<double>(gamma, 1.0, 0, 0);
    prim2cons(q, c, s);
}

template <int N>
void seedvec(Rng& r, double v[N])
{
    for (int i = 0; i < N; ++i) v[i] = r.n();
}

}  // namespace

extern "C" {

// which: 0 cons2prim, 1 prim2cons, 2..4 flux_dir<0..2>, 5 fast_speed, 6 half slope MonCen,
// 7 half slope minmod, 8 ppm_edges, 9 llf_x, 10 llf_x bound, 11 hll_x, 12 hlld_x,
// 13..15 riemann<HLLD, 0..2>, 16 riemann<HLL, 1>, 17 riemann<LLF, 2>, 18 riemann<HLLD> with the LLF switch,
// 19..23 face states (full/normal-only predictor, PPM/PLM, guards), 24..25 full Hancock predictor
// (MonCen/minmod), 26 update epilogue with the OU kick, 27 the same for a floored cell,
// 28 riemann<LLF bound, 1>.
// bscale sets |B| (0.05: weak field, 2: strong); out = {<Jv, w>, <v, J^T w>}.  Returns 0, or -1
// for an unknown function.
int adj_check(int which, uint64_t seed, double bscale, double* out)
{
    Rng r(seed);
    const double gamma = 5.0 / 3.0, ch = 2.5;
    Phys<double> s = phys<double>(gamma, ch, which == 18 ? 10.0 : 0.0, which == 18 ? 10.0 : 0.0);
    Phys<Dual> sd = phys<Dual>(gamma, ch, which == 18 ? 10.0 : 0.0, which == 18 ? 10.0 : 0.0);
    double jvw = 0.0, vjw = 0.0;
    if (which <= 4) {
        double q[9], x[9], v[9], w[9], xb[9] = {0};
        prim(r, q, bscale);
        if (which == 0) prim2cons_d(q, x, gamma);
        else for (int i = 0; i < 9; ++i) x[i] = q[i];
        seedvec<9>(r, v);
        seedvec<9>(r, w);
        Dual xd[9], yd[9];
        for (int i = 0; i < 9; ++i) xd[i] = Dual(x[i], v[i]);
        if (which == 0) { cons2prim(xd, yd, sd); cons2prim_adj(x, w, xb, s); }
        else if (which == 1) { prim2cons(xd, yd, sd); prim2cons_adj(x, w, xb, s); }
        else if (which == 2) { flux_dir<Dual, 0>(xd, yd, sd); flux_dir_adj<double, 0>(x, w, xb, s); }
        else if (which == 3) { flux_dir<Dual, 1>(xd, yd, sd); flux_dir_adj<double, 1>(x, w, xb, s); }
        else { flux_dir<Dual, 2>(xd, yd, sd); flux_dir_adj<double, 2>(x, w, xb, s); }
        for (int i = 0; i < 9; ++i) { jvw += yd[i].d * w[i]; vjw += v[i] * xb[i]; }
    } else if (which == 5) {
        double q[9];
        prim(r, q, bscale);
        double b2 = q[5] * q[5] + q[6] * q[6] + q[7] * q[7], v[4], w = r.n();
        seedvec<4>(r, v);
        Dual cf = fast_speed(Dual(q[0], v[0]), Dual(q[4], v[1]), Dual(b2, v[2]), Dual(q[5], v[3]), sd);
        double xb[4] = {0, 0, 0, 0};
        fast_speed_adj(q[0], q[4], b2, q[5], w, xb[0], xb[1], xb[2], xb[3], s);
        jvw = cf.d * w;
        for (int i = 0; i < 4; ++i) vjw += v[i] * xb[i];
    } else if (which <= 8) {
        double x[3] = {r.u(-1, 1), r.u(-1, 1), r.u(-1, 1)}, v[3], w[2], xb[3] = {0, 0, 0};
        seedvec<3>(r, v);
        seedvec<2>(r, w);
        Dual a(x[0], v[0]), b(x[1], v[1]), c(x[2], v[2]);
        if (which == 6) {
            Dual h = half_slope<Dual, SLOPE_MONCEN>(a, b, c);
            half_slope_adj<double, SLOPE_MONCEN>(x[0], x[1], x[2], w[0], xb[0], xb[1], xb[2]);
            jvw = h.d * w[0];
        } else if (which == 7) {
            Dual h = half_slope<Dual, SLOPE_MINMOD>(a, b, c);
            half_slope_adj<double, SLOPE_MINMOD>(x[0], x[1], x[2], w[0], xb[0], xb[1], xb[2]);
            jvw = h.d * w[0];
        } else {
            Dual lo, hi;
            ppm_edges(a, b, c, lo, hi);
            ppm_edges_adj(x[0], x[1], x[2], w[0], w[1], xb[0], xb[1], xb[2]);
            jvw = lo.d * w[0] + hi.d * w[1];
        }
        for (int i = 0; i < 3; ++i) vjw += v[i] * xb[i];
    } else if (which <= 18 || which == 28) {
        double L[9], R[9], vL[9], vR[9], w[9], Lb[9] = {0}, Rb[9] = {0};
        prim(r, L, bscale);
        prim(r, R, bscale);
        if (which <= 12) R[5] = L[5];   // solvers take a common normal field (the GLM bn*)
        seedvec<9>(r, vL);
        seedvec<9>(r, vR);
        seedvec<9>(r, w);
        if (which <= 12) { vR[5] = vL[5]; vL[8] = vR[8] = 0.0; w[5] = w[8] = 0.0; }
        Dual Ld[9], Rd[9], fd[9];
        for (int i = 0; i < 9; ++i) { Ld[i] = Dual(L[i], vL[i]); Rd[i] = Dual(R[i], vR[i]); }
        switch (which) {
        case 9: llf_x<Dual, false>(Ld, Rd, fd, sd); llf_x_adj<double, false>(L, R, w, Lb, Rb, s); break;
        case 10: llf_x<Dual, true>(Ld, Rd, fd, sd); llf_x_adj<double, true>(L, R, w, Lb, Rb, s); break;
        case 11: hll_x(Ld, Rd, fd, sd); hll_x_adj(L, R, w, Lb, Rb, s); break;
        case 12: hlld_x(Ld, Rd, fd, sd); hlld_x_adj(L, R, w, Lb, Rb, s); break;
        case 13: riemann<Dual, RIEM_HLLD, 0>(Ld, Rd, fd, sd); riemann_adj<double, RIEM_HLLD, 0>(L, R, w, Lb, Rb, s); break;
        case 14: riemann<Dual, RIEM_HLLD, 1>(Ld, Rd, fd, sd); riemann_adj<double, RIEM_HLLD, 1>(L, R, w, Lb, Rb, s); break;
        case 15: riemann<Dual, RIEM_HLLD, 2>(Ld, Rd, fd, sd); riemann_adj<double, RIEM_HLLD, 2>(L, R, w, Lb, Rb, s); break;
        case 16: riemann<Dual, RIEM_HLL, 1>(Ld, Rd, fd, sd); riemann_adj<double, RIEM_HLL, 1>(L, R, w, Lb, Rb, s); break;
        case 17: riemann<Dual, RIEM_LLF, 2>(Ld, Rd, fd, sd); riemann_adj<double, RIEM_LLF, 2>(L, R, w, Lb, Rb, s); break;
        case 28: riemann<Dual, RIEM_LLFB, 1>(Ld, Rd, fd, sd); riemann_adj<double, RIEM_LLFB, 1>(L, R, w, Lb, Rb, s); break;
        default: riemann<Dual, RIEM_HLLD, 0>(Ld, Rd, fd, sd); riemann_adj<double, RIEM_HLLD, 0>(L, R, w, Lb, Rb, s); break;
        }
        if (which <= 12) {
            // the solver's slot-5 input is the shared bn*: fold R's slot-5 adjoint into it
            Lb[5] += Rb[5];
            Rb[5] = 0.0;
            for (int i = 0; i < 9; ++i) jvw += fd[i].d * w[i];
            for (int i = 0; i < 9; ++i) vjw += vL[i] * Lb[i] + vR[i] * Rb[i];
        } else {
            for (int i = 0; i < 9; ++i) jvw += fd[i].d * w[i];
            for (int i = 0; i < 9; ++i) vjw += vL[i] * Lb[i] + vR[i] * Rb[i];
        }
    } else if (which <= 23) {
        // face states of one cell: inputs l, m0, r, mhf; outputs lo, hi
        double x[4][9], v[4][9], w[2][9], xb[4][9] = {{0}};
        for (int k = 0; k < 4; ++k) prim(r, x[k], bscale);
        for (int k = 0; k < 4; ++k) seedvec<9>(r, v[k]);
        seedvec<9>(r, w[0]);
        seedvec<9>(r, w[1]);
        Dual xd[4][9], lo[9], hi[9];
        for (int k = 0; k < 4; ++k) for (int i = 0; i < 9; ++i) xd[k][i] = Dual(x[k][i], v[k][i]);
        FaceOpts o{1, GUARD_IFLO};
        const double hd = 0.01;
        switch (which) {
        case 19:
            faces_fwd<Dual, PRED_FULL, RECON_PPM, SLOPE_MONCEN, 0>(xd[0], xd[1], xd[2], xd[3], true, Dual(hd), o, sd, lo, hi);
            faces_adj<double, PRED_FULL, RECON_PPM, SLOPE_MONCEN, 0>(x[0], x[1], x[2], x[3], true, hd, o, s, w[0], w[1], xb[0], xb[1], xb[2], xb[3]);
            break;
        case 20:
            faces_fwd<Dual, PRED_FULL, RECON_PPM, SLOPE_MONCEN, 1>(xd[0], xd[1], xd[2], xd[3], false, Dual(hd), o, sd, lo, hi);
            faces_adj<double, PRED_FULL, RECON_PPM, SLOPE_MONCEN, 1>(x[0], x[1], x[2], x[3], false, hd, o, s, w[0], w[1], xb[0], xb[1], xb[2], xb[3]);
            break;
        case 21:
            o.face_guard = GUARD_CLAMP;
            faces_fwd<Dual, PRED_FULL, RECON_PLM, SLOPE_MINMOD, 2>(xd[0], xd[1], xd[2], xd[3], true, Dual(hd), o, sd, lo, hi);
            faces_adj<double, PRED_FULL, RECON_PLM, SLOPE_MINMOD, 2>(x[0], x[1], x[2], x[3], true, hd, o, s, w[0], w[1], xb[0], xb[1], xb[2], xb[3]);
            break;
        case 22:
            faces_fwd<Dual, PRED_ONED, RECON_PPM, SLOPE_MONCEN, 1>(xd[0], xd[1], xd[2], xd[3], true, Dual(hd), o, sd, lo, hi);
            faces_adj<double, PRED_ONED, RECON_PPM, SLOPE_MONCEN, 1>(x[0], x[1], x[2], x[3], true, hd, o, s, w[0], w[1], xb[0], xb[1], xb[2], xb[3]);
            break;
        default:
            o.face_guard = GUARD_NONE;
            faces_fwd<Dual, PRED_ONED, RECON_PLM, SLOPE_MONCEN, 2>(xd[0], xd[1], xd[2], xd[3], true, Dual(hd), o, sd, lo, hi);
            faces_adj<double, PRED_ONED, RECON_PLM, SLOPE_MONCEN, 2>(x[0], x[1], x[2], x[3], true, hd, o, s, w[0], w[1], xb[0], xb[1], xb[2], xb[3]);
            break;
        }
        for (int i = 0; i < 9; ++i) jvw += lo[i].d * w[0][i] + hi[i].d * w[1][i];
        for (int k = 0; k < 4; ++k) for (int i = 0; i < 9; ++i) vjw += v[k][i] * xb[k][i];
    } else if (which <= 25) {
        // 24: full Hancock predictor from the cell and its six neighbours (MonCen; 25: minmod)
        double p0[9], nb[6][9], v0[9], vn[6][9], w[9], p0b[9] = {0}, nbb[6][9] = {{0}};
        prim(r, p0, bscale);
        for (int k = 0; k < 6; ++k) {
            for (int i = 0; i < 9; ++i) nb[k][i] = p0[i] * (1.0 + 0.2 * r.u(-1, 1)) + 0.05 * r.u(-1, 1);
            nb[k][0] = std::fabs(nb[k][0]) + 0.1;
            nb[k][4] = std::fabs(nb[k][4]) + 0.1;
        }
        seedvec<9>(r, v0);
        for (int k = 0; k < 6; ++k) seedvec<9>(r, vn[k]);
        seedvec<9>(r, w);
        Dual pd[9], nd[6][9], mh[9];
        for (int i = 0; i < 9; ++i) pd[i] = Dual(p0[i], v0[i]);
        for (int k = 0; k < 6; ++k) for (int i = 0; i < 9; ++i) nd[k][i] = Dual(nb[k][i], vn[k][i]);
        const double hd = 0.02;
        if (which == 24) {
            hancock_fwd<Dual, SLOPE_MONCEN>(pd, nd, mh, Dual(hd), 1, sd);
            hancock_adj<double, SLOPE_MONCEN>(p0, nb, w, p0b, nbb, hd, 1, s);
        } else {
            hancock_fwd<Dual, SLOPE_MINMOD>(pd, nd, mh, Dual(hd), 1, sd);
            hancock_adj<double, SLOPE_MINMOD>(p0, nb, w, p0b, nbb, hd, 1, s);
        }
        for (int i = 0; i < 9; ++i) jvw += mh[i].d * w[i];
        for (int i = 0; i < 9; ++i) vjw += v0[i] * p0b[i];
        for (int k = 0; k < 6; ++k) for (int i = 0; i < 9; ++i) vjw += vn[k][i] * nbb[k][i];
    } else if (which <= 27) {
        // 26: update epilogue with the OU kick; 27: a floored cell (rho below smallr)
        double q[9], c[9], a3[3], v[9], va[3], w[9], cb[9] = {0}, a3b[3] = {0};
        prim(r, q, bscale);
        prim2cons_d(q, c, gamma);
        if (which == 27) c[0] = 5e-7;
        for (int k = 0; k < 3; ++k) a3[k] = r.u(-3, 3);
        seedvec<9>(r, v);
        seedvec<3>(r, va);
        seedvec<9>(r, w);
        Dual cd[9], ad[3];
        for (int i = 0; i < 9; ++i) cd[i] = Dual(c[i], v[i]);
        for (int k = 0; k < 3; ++k) ad[k] = Dual(a3[k], va[k]);
        UpdOpts o{1, 1};
        const double fac = 0.97, smallc = 1e-4, tmr = 1e-5, dt = 0.01;
        epilogue_fwd<Dual>(cd, ad, Dual(fac), Dual(smallc), Dual(tmr), Dual(dt), o, sd);
        epilogue_adj<double>(c, a3, fac, smallc, tmr, dt, o, s, w, cb, a3b);
        for (int i = 0; i < 9; ++i) jvw += cd[i].d * w[i];
        for (int i = 0; i < 9; ++i) vjw += v[i] * cb[i];
        for (int k = 0; k < 3; ++k) vjw += va[k] * a3b[k];
    } else {
        return -1;
    }
    out[0] = jvw;
    out[1] = vjw;
    return 0;
}

// The division-free spj against its quotient form max / max(min, 1e-20) > 2 on n triples whose
// ratio sits within a few ulps of 2, plus random and tiny ones.  prec 8: double, 4: float.
// Returns the number of disagreements.
long spj_check(long n, uint64_t seed, int prec)
{
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> ue(-60.0, 60.0), uu(0.0, 1.0);
    long bad = 0;
    for (long i = 0; i < n; ++i) {
        const double lo = std::pow(10.0, ue(g) * 0.4);
        const int kind = (int)(g() % 4);
        auto run = [&](auto zero) {
            using T = decltype(zero);
            T l = (T)lo, h;
            if (kind == 0) {            // 2 lo moved by -8..8 ulps
                h = T(2) * l;
                const int k = (int)(g() % 17) - 8;
                for (int j = 0; j < (k < 0 ? -k : k); ++j) h = std::nextafter(h, k < 0 ? T(0) : T(1e30));
            } else if (kind == 1) {
                h = l * (T)(0.5 + 4.0 * uu(g));
            } else if (kind == 2) {
                h = (T)(lo * 1e-22);   // below the 1e-20 floor
            } else {
                h = T(2) * l * (T)(1.0 + (uu(g) - 0.5) * 1e-14);
            }
            T m = l + (h - l) * (T)uu(g);
            T x[3] = {l, m, h};
            std::shuffle(x, x + 3, g);
            T hi = std::max(x[0], std::max(x[1], x[2]));
            T lw = std::max(std::min(x[0], std::min(x[1], x[2])), T(1e-20));
            volatile T q = hi / lw;
            return (bool)(q > T(2)) != glm::spj(x[0], x[1], x[2]);
        };
        bad += prec == 4 ? run(0.0f) : run(0.0);
    }
    return bad;
}

}  // extern "C"

// Host restatement of the six reverse passes of adj_impl.hpp (no forcing), for CPU comparison
// with the dual-number oracle.  Scheme: pred 0/1, recon 0/1 (PLM/PPM), slope 0/1, riem 0..3,
// guard 0..2; mh_revert, update_floor on.
template <int PRED, int RECON, int SLOPE, int RIEM>
static void adj_step_host(const double* u0, const double* lam1, double* lam0, int nx, int ny, int nz, double gamma,
                          double dt, double dx, double ch, double glm_fac, int guard, double* dump = nullptr)
{
    const size_t n = (size_t)nx * ny * nz;
    Phys<double> s = phys<double>(gamma, ch, 0.0, 0.0);
    const double dtdx = dt / dx, hd = 0.5 * dtdx;
    FaceOpts o{1, guard};
    UpdOpts uo{1, 0};
    std::vector<double> W(9 * n), MH(9 * n), FS(54 * n), LPRE(9 * n), LFS(54 * n), CON(63 * n);
    auto idx = [&](int i, int j, int k) {
        i = (i % nx + nx) % nx; j = (j % ny + ny) % ny; k = (k % nz + nz) % nz;
        return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
    };
    auto ld = [&](const double* b, size_t c, double x[9]) { for (int v = 0; v < 9; ++v) x[v] = b[(size_t)v * n + c]; };
    auto st = [&](double* b, size_t c, const double x[9]) { for (int v = 0; v < 9; ++v) b[(size_t)v * n + c] = x[v]; };
    for (size_t c = 0; c < n; ++c) { double u[9], w[9]; ld(u0, c, u); cons2prim(u, w, s); st(W.data(), c, w); }
    auto stencil = [&](int i, int j, int k, double p0[9], double nb[6][9]) {
        ld(W.data(), idx(i, j, k), p0);
        ld(W.data(), idx(i - 1, j, k), nb[0]); ld(W.data(), idx(i + 1, j, k), nb[1]);
        ld(W.data(), idx(i, j - 1, k), nb[2]); ld(W.data(), idx(i, j + 1, k), nb[3]);
        ld(W.data(), idx(i, j, k - 1), nb[4]); ld(W.data(), idx(i, j, k + 1), nb[5]);
    };
    auto upflag = [&](const double p0[9], double nb[6][9]) {
        return RECON != RECON_PPM || !(spj(nb[0][4], p0[4], nb[1][4]) || spj(nb[2][4], p0[4], nb[3][4]) || spj(nb[4][4], p0[4], nb[5][4]));
    };
    for (int k = 0; k < nz; ++k) for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) {
        size_t c = idx(i, j, k);
        double p0[9], nb[6][9], mh[9], lo[9], hi[9];
        stencil(i, j, k, p0, nb);
        bool up = upflag(p0, nb);
        if (PRED == PRED_FULL) { hancock_fwd<double, SLOPE>(p0, nb, mh, hd, 1, s); st(MH.data(), c, mh); }
        faces_fwd<double, PRED, RECON, SLOPE, 0>(nb[0], p0, nb[1], mh, up, hd, o, s, lo, hi);
        st(FS.data() + 0 * 9 * n, c, lo); st(FS.data() + 1 * 9 * n, c, hi);
        faces_fwd<double, PRED, RECON, SLOPE, 1>(nb[2], p0, nb[3], mh, up, hd, o, s, lo, hi);
        st(FS.data() + 2 * 9 * n, c, lo); st(FS.data() + 3 * 9 * n, c, hi);
        faces_fwd<double, PRED, RECON, SLOPE, 2>(nb[4], p0, nb[5], mh, up, hd, o, s, lo, hi);
        st(FS.data() + 4 * 9 * n, c, lo); st(FS.data() + 5 * 9 * n, c, hi);
    }
    auto flux = [&](int d, size_t cl, size_t cr, double f[9]) {
        double L[9], R[9];
        ld(FS.data() + (size_t)(2 * d + 1) * 9 * n, cl, L);
        ld(FS.data() + (size_t)(2 * d) * 9 * n, cr, R);
        if (d == 0) riemann<double, RIEM, 0>(L, R, f, s);
        else if (d == 1) riemann<double, RIEM, 1>(L, R, f, s);
        else riemann<double, RIEM, 2>(L, R, f, s);
    };
    for (int k = 0; k < nz; ++k) for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) {
        size_t c = idx(i, j, k);
        double acc[9] = {0}, fl[9], fh[9], r[9], lam[9], rb[9] = {0}, a3[3] = {0, 0, 0}, a3b[3] = {0, 0, 0};
        flux(0, idx(i - 1, j, k), c, fl); flux(0, c, idx(i + 1, j, k), fh); for (int v = 0; v < 9; ++v) acc[v] += fl[v] - fh[v];
        flux(1, idx(i, j - 1, k), c, fl); flux(1, c, idx(i, j + 1, k), fh); for (int v = 0; v < 9; ++v) acc[v] += fl[v] - fh[v];
        flux(2, idx(i, j, k - 1), c, fl); flux(2, c, idx(i, j, k + 1), fh); for (int v = 0; v < 9; ++v) acc[v] += fl[v] - fh[v];
        double u[9]; ld(u0, c, u); ld(lam1, c, lam);
        for (int v = 0; v < 9; ++v) r[v] = u[v] + dtdx * acc[v];
        epilogue_adj<double>(r, a3, glm_fac, 1e-4, 1e-5, dt, uo, s, lam, rb, a3b);
        st(LPRE.data(), c, rb);
    }
    for (int k = 0; k < nz; ++k) for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) {
        size_t c = idx(i, j, k);
        size_t cls[3] = {idx(i - 1, j, k), idx(i, j - 1, k), idx(i, j, k - 1)};
        for (int d = 0; d < 3; ++d) {
            double L[9], R[9], lr[9], ll[9], ob[9], Lb[9] = {0}, Rb[9] = {0};
            ld(FS.data() + (size_t)(2 * d + 1) * 9 * n, cls[d], L);
            ld(FS.data() + (size_t)(2 * d) * 9 * n, c, R);
            ld(LPRE.data(), c, lr); ld(LPRE.data(), cls[d], ll);
            for (int v = 0; v < 9; ++v) ob[v] = dtdx * (lr[v] - ll[v]);
            if (d == 0) riemann_adj<double, RIEM, 0>(L, R, ob, Lb, Rb, s);
            else if (d == 1) riemann_adj<double, RIEM, 1>(L, R, ob, Lb, Rb, s);
            else riemann_adj<double, RIEM, 2>(L, R, ob, Lb, Rb, s);
            st(LFS.data() + (size_t)(2 * d + 1) * 9 * n, cls[d], Lb);
            st(LFS.data() + (size_t)(2 * d) * 9 * n, c, Rb);
        }
    }
    for (int k = 0; k < nz; ++k) for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) {
        size_t c = idx(i, j, k);
        double p0[9], nb[6][9], mh[9] = {0}, con[7][9] = {{0}}, mhb[9] = {0}, lob[9], hib[9];
        stencil(i, j, k, p0, nb);
        bool up = upflag(p0, nb);
        if (PRED == PRED_FULL) ld(MH.data(), c, mh);
        ld(LFS.data() + 0 * 9 * n, c, lob); ld(LFS.data() + 1 * 9 * n, c, hib);
        faces_adj<double, PRED, RECON, SLOPE, 0>(nb[0], p0, nb[1], mh, up, hd, o, s, lob, hib, con[1], con[0], con[2], mhb);
        ld(LFS.data() + 2 * 9 * n, c, lob); ld(LFS.data() + 3 * 9 * n, c, hib);
        faces_adj<double, PRED, RECON, SLOPE, 1>(nb[2], p0, nb[3], mh, up, hd, o, s, lob, hib, con[3], con[0], con[4], mhb);
        ld(LFS.data() + 4 * 9 * n, c, lob); ld(LFS.data() + 5 * 9 * n, c, hib);
        faces_adj<double, PRED, RECON, SLOPE, 2>(nb[4], p0, nb[5], mh, up, hd, o, s, lob, hib, con[5], con[0], con[6], mhb);
        if (PRED == PRED_FULL) hancock_adj<double, SLOPE>(p0, nb, mhb, con[0], &con[1], hd, 1, s);
        for (int q = 0; q < 7; ++q) st(CON.data() + (size_t)q * 9 * n, c, con[q]);
    }
    for (int k = 0; k < nz; ++k) for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) {
        size_t c = idx(i, j, k);
        double wb[9], t[9], u[9], lb[9];
        ld(CON.data(), c, wb);
        size_t lo3[3] = {idx(i - 1, j, k), idx(i, j - 1, k), idx(i, j, k - 1)};
        size_t hi3[3] = {idx(i + 1, j, k), idx(i, j + 1, k), idx(i, j, k + 1)};
        for (int d = 0; d < 3; ++d) {
            ld(CON.data() + (size_t)(2 + 2 * d) * 9 * n, lo3[d], t); for (int v = 0; v < 9; ++v) wb[v] += t[v];
            ld(CON.data() + (size_t)(1 + 2 * d) * 9 * n, hi3[d], t); for (int v = 0; v < 9; ++v) wb[v] += t[v];
        }
        ld(u0, c, u); ld(LPRE.data(), c, lb);
        cons2prim_adj(u, wb, lb, s);
        st(lam0, c, lb);
    }
    if (dump) {   // the scratch planes in the layout of adj_impl.hpp (W, MH, FS, LPRE, LFS, CON)
        double* o = dump;
        for (const std::vector<double>* v : {&W, &MH, &FS, &LPRE, &LFS, &CON}) {
            std::copy(v->begin(), v->end(), o);
            o += v->size();
        }
    }
}

extern "C" int adj_step_cpu(int scheme, const double* u0, const double* lam1, double* lam0, int nx, int ny, int nz,
                            double gamma, double dt, double dx, double ch, double glm_fac, double* dump)
{
    // 0: full PLM MonCen LLF-bound iflo; 1: full PLM MonCen HLL iflo; 2: full PPM MonCen HLLD iflo;
    // 3: full PLM minmod HLL none
    switch (scheme) {
    case 0: adj_step_host<PRED_FULL, RECON_PLM, SLOPE_MONCEN, RIEM_LLFB>(u0, lam1, lam0, nx, ny, nz, gamma, dt, dx, ch, glm_fac, GUARD_IFLO, dump); return 0;
    case 1: adj_step_host<PRED_FULL, RECON_PLM, SLOPE_MONCEN, RIEM_HLL>(u0, lam1, lam0, nx, ny, nz, gamma, dt, dx, ch, glm_fac, GUARD_IFLO, dump); return 0;
    case 2: adj_step_host<PRED_FULL, RECON_PPM, SLOPE_MONCEN, RIEM_HLLD>(u0, lam1, lam0, nx, ny, nz, gamma, dt, dx, ch, glm_fac, GUARD_IFLO, dump); return 0;
    case 3: adj_step_host<PRED_FULL, RECON_PLM, SLOPE_MINMOD, RIEM_HLL>(u0, lam1, lam0, nx, ny, nz, gamma, dt, dx, ch, glm_fac, GUARD_NONE, dump); return 0;
    default: return -1;
    }
}
