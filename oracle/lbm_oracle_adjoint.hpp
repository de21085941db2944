// lbm_oracle_adjoint.hpp — FP64 host adjoint of the oracle's time step.
//
// TEST INFRASTRUCTURE ONLY (see lbm_oracle.hpp). The reference has no adjoint; this is the
// "FP64 host-reference gradient" the GPU adjoint is checked against. It does not hand-derive
// any Jacobian: the oracle's own node update (macroscopics, Guo correction, equilibrium,
// BGK/MRT collision — the templates in lbm_oracle.hpp) is evaluated on forward-mode dual
// numbers seeded with the nine populations and the two force components, which gives the
// node's exact 9x11 Jacobian. The backward sweep then transposes the streaming permutation.
#pragma once

#include "lbm_oracle.hpp"

namespace oracle {

// value + derivatives with respect to K seeds
struct Dual {
    static constexpr int K = 11;
    double v;
    double d[K];
    Dual(double x = 0.0) : v(x) {
        for (int i = 0; i < K; i++) d[i] = 0.0;
    }
    explicit Dual(int x) : Dual(double(x)) {}
    static Dual seed(double x, int i) {
        Dual r(x);
        r.d[i] = 1.0;
        return r;
    }
};
inline Dual operator+(const Dual& a, const Dual& b) {
    Dual r(a.v + b.v);
    for (int i = 0; i < Dual::K; i++) r.d[i] = a.d[i] + b.d[i];
    return r;
}
inline Dual operator-(const Dual& a, const Dual& b) {
    Dual r(a.v - b.v);
    for (int i = 0; i < Dual::K; i++) r.d[i] = a.d[i] - b.d[i];
    return r;
}
inline Dual operator-(const Dual& a) {
    Dual r(-a.v);
    for (int i = 0; i < Dual::K; i++) r.d[i] = -a.d[i];
    return r;
}
inline Dual operator*(const Dual& a, const Dual& b) {
    Dual r(a.v * b.v);
    for (int i = 0; i < Dual::K; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
inline Dual operator/(const Dual& a, const Dual& b) {
    Dual r(a.v / b.v);
    for (int i = 0; i < Dual::K; i++) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v);
    return r;
}
inline Dual& operator+=(Dual& a, const Dual& b) { return a = a + b; }
inline Dual& operator-=(Dual& a, const Dual& b) { return a = a - b; }

// The node update exactly as Solver<R>::step composes it for a FLUID node
// (macroscopics.cu:5-38, 99-110; equilibrium.cu:5-39; BGK.cuh / MRT.cu).
template <class R>
void node_update(const lbm_config& c, const R* g, R fx, R fy, R* out) {
    R f[Q], feq[Q];
    R r = 0, jx = 0, jy = 0;
    for (int i = 0; i < Q; i++) {
        f[i] = g[i];
        r += g[i];
        jx += g[i] * R(CX[i]);
        jy += g[i] * R(CY[i]);
    }
    R ux = jx * (R(1) / r), uy = jy * (R(1) / r);
    ux += R(0.5) * fx / r;
    uy += R(0.5) * fy / r;
    equilibrium<R>(r, ux, uy, feq, false);
    if (c.collision == LBM_BGK) {
        collide_bgk<R>(f, feq, ux, uy, fx, fy, R(c.omega));
    } else {
        R S[Q];
        for (int k = 0; k < Q; k++) S[k] = R(c.S[k]);
        collide_mrt<R>(f, feq, ux, uy, fx, fy, S, false);
    }
    for (int q = 0; q < Q; q++) out[q] = f[q];
}

// J[q][k] = d out_q / d seed_k, seeds = g[0..8], fx, fy.
inline void node_jacobian(const lbm_config& c, const double* g, double fx, double fy, double* J) {
    Dual gd[Q], out[Q];
    for (int i = 0; i < Q; i++) gd[i] = Dual::seed(g[i], i);
    node_update<Dual>(c, gd, Dual::seed(fx, 9), Dual::seed(fy, 10), out);
    for (int q = 0; q < Q; q++)
        for (int k = 0; k < Dual::K; k++) J[q * Dual::K + k] = out[q].d[k];
}

// Adjoint of the periodic, FLUID-only step from the solver's current state (which must be
// the state written by init_from_macros) over T steps, for dL/du_T given. Outputs dL/dF
// (AoS [N][2]) and dL/du0 (AoS [N][2]).
inline void adjoint(Solver<double>& s, int T, const double* dLdu, double* gradF, double* gradU0) {
    const int nx = s.nx, ny = s.ny;
    const size_t N = s.N;
    auto wrap = [](int v, int n) { return (v % n + n) % n; };
    std::vector<double> rho0(s.rho), u0(s.u);  // initial macros (SoA u)
    std::vector<std::vector<double>> states;   // f*_0 .. f*_{T-1}
    for (int t = 0; t < T; t++) {
        states.push_back(s.f);
        s.step();
    }
    auto force = [&](size_t n, double& fx, double& fy) {
        if (s.force_field.empty()) {
            fx = s.cfg.force[0];
            fy = s.cfg.force[1];
        } else {
            fx = s.force_field[2 * n];
            fy = s.force_field[2 * n + 1];
        }
    };
    auto gather = [&](const std::vector<double>& st, int x, int y, double* g) {
        for (int q = 0; q < Q; q++) g[q] = st[q * N + size_t(wrap(y - CY[q], ny)) * nx + size_t(wrap(x - CX[q], nx))];
    };
    std::vector<double> mu(Q * N), mu_next(Q * N);
    for (size_t n = 0; n < 2 * N; n++) gradF[n] = 0.0;
    // seed: u_T = (sum g c + F/2) / rho on g_T
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            size_t n = size_t(y) * nx + x;
            double g[Q], fx, fy;
            gather(states[size_t(T - 1)], x, y, g);
            force(n, fx, fy);
            Dual gd[Q];
            Dual r = 0, jx = 0, jy = 0;
            for (int i = 0; i < Q; i++) {
                gd[i] = Dual::seed(g[i], i);
                r += gd[i];
                jx += gd[i] * Dual(CX[i]);
                jy += gd[i] * Dual(CY[i]);
            }
            Dual Fx = Dual::seed(fx, 9), Fy = Dual::seed(fy, 10);
            Dual ux = jx * (Dual(1.0) / r) + Dual(0.5) * Fx / r;
            Dual uy = jy * (Dual(1.0) / r) + Dual(0.5) * Fy / r;
            for (int k = 0; k < Q; k++) mu[k * N + n] = dLdu[2 * n] * ux.d[k] + dLdu[2 * n + 1] * uy.d[k];
            gradF[2 * n] += dLdu[2 * n] * ux.d[9] + dLdu[2 * n + 1] * uy.d[9];
            gradF[2 * n + 1] += dLdu[2 * n] * ux.d[10] + dLdu[2 * n + 1] * uy.d[10];
        }
    double J[Q * Dual::K];
    for (int t = T - 1; t >= 1; t--) {
        std::swap(mu, mu_next);
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                size_t n = size_t(y) * nx + x;
                double lam[Q], g[Q], fx, fy;
                for (int q = 0; q < Q; q++)
                    lam[q] = mu_next[q * N + size_t(wrap(y + CY[q], ny)) * nx + size_t(wrap(x + CX[q], nx))];
                gather(states[size_t(t - 1)], x, y, g);
                force(n, fx, fy);
                node_jacobian(s.cfg, g, fx, fy, J);
                for (int k = 0; k < Q; k++) {
                    double acc = 0.0;
                    for (int q = 0; q < Q; q++) acc += J[q * Dual::K + k] * lam[q];
                    mu[k * N + n] = acc;
                }
                for (int d = 0; d < 2; d++) {
                    double acc = 0.0;
                    for (int q = 0; q < Q; q++) acc += J[q * Dual::K + 9 + d] * lam[q];
                    gradF[2 * n + d] += acc;
                }
            }
    }
    // initial state f*_0 = feq(rho0, u0)
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            size_t n = size_t(y) * nx + x;
            double lam[Q];
            for (int q = 0; q < Q; q++)
                lam[q] = mu[q * N + size_t(wrap(y + CY[q], ny)) * nx + size_t(wrap(x + CX[q], nx))];
            Dual feq[Q];
            equilibrium<Dual>(Dual(rho0[n]), Dual::seed(u0[n], 0), Dual::seed(u0[N + n], 1), feq, false);
            double gx = 0, gy = 0;
            for (int q = 0; q < Q; q++) {
                gx += feq[q].d[0] * lam[q];
                gy += feq[q].d[1] * lam[q];
            }
            gradU0[2 * n] = gx;
            gradU0[2 * n + 1] = gy;
        }
}

}  // namespace oracle
