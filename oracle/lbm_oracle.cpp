// lbm_oracle.cpp — C entry points of the CPU oracle (liblbm_oracle.so), for ctypes.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg as the checker. Values cross this boundary as double; the solver runs
// in float (precision 32) or double (precision 64).
#include "lbm_oracle.hpp"
#include "lbm_oracle_adjoint.hpp"

#include <memory>
#include <omp.h>

using namespace oracle;

namespace {
struct Handle {
    int prec;
    std::unique_ptr<Solver<float>> s32;
    std::unique_ptr<Solver<double>> s64;
    std::string err;
};

template <class F>
int dispatch(Handle* h, F&& fn) {
    try {
        if (h->prec == 32) fn(*h->s32);
        else fn(*h->s64);
        return 0;
    } catch (const std::exception& e) {
        h->err = e.what();
        return -1;
    }
}
}  // namespace

extern "C" {

void* oracle_create(const lbm_config* cfg, int precision) {
    auto* h = new Handle();
    h->prec = precision == 64 ? 64 : 32;
    try {
        if (h->prec == 32) h->s32.reset(new Solver<float>(*cfg));
        else h->s64.reset(new Solver<double>(*cfg));
    } catch (...) {
        delete h;
        return nullptr;
    }
    return h;
}

void oracle_destroy(void* p) { delete static_cast<Handle*>(p); }

int oracle_set_eq_literal(void* p, int on) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) { s.eq_literal = on != 0; });
}

int oracle_set_flags(void* p, const uint8_t* flags) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) { s.flags.assign(flags, flags + s.N); });
}

int oracle_init_from_macros(void* p, const double* rho, const double* u) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) {
        using R = typename std::remove_reference_t<decltype(s.f)>::value_type;
        std::vector<R> r(rho, rho + s.N), uu(u, u + 2 * s.N);
        s.init_from_macros(r.data(), uu.data());
    });
}

int oracle_set_force_field(void* p, const double* F) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) {
        using R = typename std::remove_reference_t<decltype(s.f)>::value_type;
        if (!F) s.force_field.clear();
        else s.force_field.assign(F, F + 2 * s.N);
        (void)sizeof(R);
    });
}

int oracle_step(void* p, int n) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) {
        for (int i = 0; i < n; i++) s.step();
    });
}

int oracle_get_macros(void* p, double* rho, double* u) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) {
        for (size_t n = 0; n < s.N; n++) {
            if (rho) rho[n] = double(s.rho[n]);
            if (u) {
                u[2 * n] = double(s.u[n]);
                u[2 * n + 1] = double(s.u[s.N + n]);
            }
        }
    });
}

int oracle_get_pops(void* p, double* f) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) {
        for (size_t i = 0; i < Q * s.N; i++) f[i] = double(s.f[i]);
    });
}

int oracle_get_pops_back(void* p, double* f) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) {
        for (size_t i = 0; i < Q * s.N; i++) f[i] = double(s.f_back[i]);
    });
}

int oracle_set_pops(void* p, const double* f, const double* f_back) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) {
        using R = typename std::remove_reference_t<decltype(s.f)>::value_type;
        for (size_t i = 0; i < Q * s.N; i++) s.f[i] = R(f[i]);
        for (size_t i = 0; i < Q * s.N; i++) s.f_back[i] = R(f_back ? f_back[i] : f[i]);
        s.quantize(s.f);
        s.quantize(s.f_back);
    });
}

int oracle_get_avg(void* p, double* avg) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) {
        for (int i = 0; i < 3; i++) avg[i] = double(s.avg[i]);
    });
}

int oracle_ibm_set_markers(void* p, const double* xy, const double* ds, int n, int iters) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) {
        using R = typename std::remove_reference_t<decltype(s.f)>::value_type;
        std::vector<R> a(xy, xy + 2 * n), d;
        if (ds) d.assign(ds, ds + n);
        s.set_markers(a.data(), ds ? d.data() : nullptr, n, iters);
    });
}

// Lagrangian force summed over the IBM iterations of the last step, AoS [n][2], and the
// total force on the fluid sum_l F_l ds_l.
int oracle_ibm_get_forces(void* p, double* lag, double* total) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) {
        int np = int(s.lag_rho.size());
        double tx = 0, ty = 0;
        for (int l = 0; l < np; l++) {
            double fx = double(s.lag_force_sum[l]), fy = double(s.lag_force_sum[np + l]);
            if (lag) {
                lag[2 * l] = fx;
                lag[2 * l + 1] = fy;
            }
            double ds = (s.cfg.quirks & LBM_QUIRK_IBM_NO_DS) ? 1.0 : double(s.lag_ds[l]);
            tx += fx * ds;
            ty += fy * ds;
        }
        if (total) {
            total[0] = tx;
            total[1] = ty;
        }
    });
}

// The three IBM operators of one iteration on given fields (FP64), for pinning them one by
// one: interpolation (IBM_impl.cu:5-51) of rho and u (SoA [2][N]) to the markers -> u_lag
// (SoA [2][n]), rho_lag [n]; the Lagrangian force (IBM_impl.cuh:7-26) of those -> F_lag_out
// (SoA [2][n]); spreading (IBM_impl.cu:120-154) of F_lag_in (SoA [2][n]) -> F_eul (SoA [2][N]).
int oracle_ibm_operators(const lbm_config* cfg, const double* rho, const double* u, const double* xy,
                         const double* ds, int n, const double* F_lag_in, double* u_lag, double* rho_lag,
                         double* F_lag_out, double* F_eul) {
    try {
        Solver<double> s(*cfg);
        s.set_markers(xy, ds, n, 1);
        std::copy(rho, rho + s.N, s.rho.begin());
        std::copy(u, u + 2 * s.N, s.eul_u_prev.begin());
        s.ibm_interpolate();
        s.ibm_lagrangian();
        std::copy(s.lag_u.begin(), s.lag_u.end(), u_lag);
        std::copy(s.lag_rho.begin(), s.lag_rho.end(), rho_lag);
        std::copy(s.lag_force.begin(), s.lag_force.end(), F_lag_out);
        std::copy(F_lag_in, F_lag_in + 2 * n, s.lag_force.begin());
        std::fill(s.eul_force_iter.begin(), s.eul_force_iter.end(), 0.0);
        s.ibm_spread();
        std::copy(s.eul_force_iter.begin(), s.eul_force_iter.end(), F_eul);
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

int oracle_get_force(void* p, double* F) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) {
        for (size_t n = 0; n < s.N; n++) {
            F[2 * n] = double(s.force[n]);
            F[2 * n + 1] = double(s.force[s.N + n]);
        }
    });
}

// FP64 adjoint from the current (freshly initialised) state over T steps: dL/dF and dL/du0.
int oracle_adjoint(void* p, int T, const double* dLdu, double* gradF, double* gradU0) {
    auto* h = static_cast<Handle*>(p);
    if (h->prec != 64) {
        h->err = "the oracle adjoint runs in FP64";
        return -1;
    }
    try {
        adjoint(*h->s64, T, dLdu, gradF, gradU0);
    } catch (const std::exception& e) {
        h->err = e.what();
        return -1;
    }
    return 0;
}

// 9x11 Jacobian of the node update (seeds g[0..8], Fx, Fy), row-major.
int oracle_node_jacobian(const lbm_config* c, const double* g, const double* F, double* J) {
    node_jacobian(*c, g, F[0], F[1], J);
    return 0;
}

const char* oracle_last_error(void* p) { return static_cast<Handle*>(p)->err.c_str(); }

int oracle_num_threads(void) { return omp_get_max_threads(); }
void oracle_set_num_threads(int n) { omp_set_num_threads(n); }

// ---- node-local building blocks, for the unit tests -----------------------------------

void oracle_tables(double* w, int* cx, int* cy, int* opp, double* M, double* Minv) {
    for (int q = 0; q < Q; q++) {
        w[q] = weight<double>(q);
        cx[q] = CX[q];
        cy[q] = CY[q];
        opp[q] = OPP[q];
    }
    for (int i = 0; i < Q * Q; i++) {
        M[i] = MM[i];
        Minv[i] = double(MINV_NUM[i]) / double(MINV_DEN[i]);
    }
}

void oracle_equilibrium(int prec, int literal, double rho, double ux, double uy, double* feq) {
    if (prec == 32) {
        float o[Q];
        equilibrium<float>(float(rho), float(ux), float(uy), o, literal != 0);
        for (int q = 0; q < Q; q++) feq[q] = o[q];
    } else {
        equilibrium<double>(rho, ux, uy, feq, literal != 0);
    }
}

void oracle_cm_matrix_inverse(double ux, double uy, double* Mi) { cm_matrix_inverse<double>(Mi, ux, uy); }

double oracle_delta4(double r) { return delta4<double>(r); }

int oracle_adapter_rate(double rho, double j, double pi, const double* avg, double* out) {
    *out = optimal_adapter_rate<double>(rho, j, pi, avg);
    return 0;
}

// One node through macroscopics + Guo correction + collision, as the step does it.
// g: post-stream populations; F: force; avg: adapter averages (NULL: none).
int oracle_collide_node(int prec, const lbm_config* cfg, const double* g, const double* F,
                        const double* avg, double* out) {
    auto run = [&](auto tag) {
        using R = decltype(tag);
        R f[Q], feq[Q];
        for (int q = 0; q < Q; q++) f[q] = R(g[q]);
        R r = 0, jx = 0, jy = 0;
        for (int i = 0; i < Q; i++) {
            r += f[i];
            jx += f[i] * R(CX[i]);
            jy += f[i] * R(CY[i]);
        }
        R ux = jx * (R(1) / r), uy = jy * (R(1) / r);
        ux += R(0.5) * R(F[0]) / r;
        uy += R(0.5) * R(F[1]) / r;
        R Sv[Q];
        for (int k = 0; k < Q; k++) Sv[k] = R(cfg->S[k]);
        if (cfg->collision == LBM_BGK) {
            equilibrium<R>(r, ux, uy, feq, false);
            collide_bgk<R>(f, feq, ux, uy, R(F[0]), R(F[1]), R(cfg->omega));
        } else if (cfg->collision == LBM_MRT) {
            equilibrium<R>(r, ux, uy, feq, false);
            collide_mrt<R>(f, feq, ux, uy, R(F[0]), R(F[1]), Sv, (cfg->quirks & LBM_QUIRK_MRT_FORCE_ROWS) != 0);
        } else {
            R a[3];
            if (avg) for (int i = 0; i < 3; i++) a[i] = R(avg[i]);
            collide_cm<R>(f, ux, uy, R(F[0]), R(F[1]), Sv, avg ? a : nullptr);
        }
        for (int q = 0; q < Q; q++) out[q] = double(f[q]);
    };
    if (prec == 32) run(float(0));
    else run(double(0));
    return 0;
}

uint16_t oracle_f32_to_f16(float x) { return f32_to_f16(x); }
float oracle_f16_to_f32(uint16_t h) { return f16_to_f32(h); }
uint16_t oracle_f32_to_bf16(float x) { return f32_to_bf16(x); }
float oracle_bf16_to_f32(uint16_t h) { return bf16_to_f32(h); }

}  // extern "C"
