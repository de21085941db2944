// lbm_oracle3d.cpp — C entry points of the D3Q27 oracle (in liblbm_oracle.so), for ctypes.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg as the checker. Values cross this boundary as double; the solver runs in
// float (precision 32) or double (precision 64).
#include "lbm_oracle3d.hpp"

#include <memory>
#include <string>

using namespace oracle3d;

namespace {
struct Handle3 {
    int prec;
    std::unique_ptr<Solver3<float>> s32;
    std::unique_ptr<Solver3<double>> s64;
    std::string err;
};

template <class F>
int dispatch3(Handle3* h, F&& fn) {
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

void* oracle3d_create(const lbm3d_config* cfg, int precision) {
    auto* h = new Handle3();
    h->prec = precision == 64 ? 64 : 32;
    try {
        if (h->prec == 32) h->s32.reset(new Solver3<float>(*cfg));
        else h->s64.reset(new Solver3<double>(*cfg));
    } catch (...) {
        delete h;
        return nullptr;
    }
    return h;
}

void oracle3d_destroy(void* p) { delete static_cast<Handle3*>(p); }

int oracle3d_set_flags(void* p, const uint8_t* flags) {
    return dispatch3(static_cast<Handle3*>(p), [&](auto& s) { s.set_flags(flags); });
}

int oracle3d_init_from_macros(void* p, const double* rho, const double* u) {
    return dispatch3(static_cast<Handle3*>(p), [&](auto& s) {
        using R = typename std::remove_reference_t<decltype(s.f)>::value_type;
        std::vector<R> r(rho, rho + s.N), uu(u, u + 3 * s.N);
        s.init_from_macros(r.data(), uu.data());
    });
}

int oracle3d_set_force_field(void* p, const double* F) {
    return dispatch3(static_cast<Handle3*>(p), [&](auto& s) {
        if (!F) s.force_field.clear();
        else s.force_field.assign(F, F + 3 * s.N);
    });
}

int oracle3d_step(void* p, int n) {
    return dispatch3(static_cast<Handle3*>(p), [&](auto& s) {
        for (int i = 0; i < n; i++) s.step();
    });
}

// rho [N] and the corrected u, AoS [N][3], of the last step.
int oracle3d_get_macros(void* p, double* rho, double* u) {
    return dispatch3(static_cast<Handle3*>(p), [&](auto& s) {
        for (size_t n = 0; n < s.N; n++) {
            if (rho) rho[n] = double(s.rho[n]);
            if (u)
                for (int c = 0; c < 3; c++) u[3 * n + c] = double(s.u[c * s.N + n]);
        }
    });
}

int oracle3d_get_pops(void* p, double* f) {
    return dispatch3(static_cast<Handle3*>(p), [&](auto& s) {
        for (size_t i = 0; i < size_t(Q) * s.N; i++) f[i] = double(s.f[i]);
    });
}

int oracle3d_set_pops(void* p, const double* f) {
    return dispatch3(static_cast<Handle3*>(p), [&](auto& s) {
        using R = typename std::remove_reference_t<decltype(s.f)>::value_type;
        for (size_t i = 0; i < size_t(Q) * s.N; i++) s.f[i] = s.f_back[i] = R(f[i]);
        s.quantize(s.f);
        s.quantize(s.f_back);
    });
}

// Only the streaming stage: f -> f_back by the push rule, then the swap (pops read back).
// The second buffer alone (after oracle3d_set_pops): the values streaming leaves in the slots
// whose source lies outside a non-periodic domain.
int oracle3d_set_pops_back(void* p, const double* f) {
    return dispatch3(static_cast<Handle3*>(p), [&](auto& s) {
        using R = typename std::remove_reference_t<decltype(s.f)>::value_type;
        for (size_t i = 0; i < size_t(Q) * s.N; i++) s.f_back[i] = R(f[i]);
        s.quantize(s.f_back);
    });
}

int oracle3d_stream_only(void* p) {
    return dispatch3(static_cast<Handle3*>(p), [&](auto& s) {
        s.stream();
        std::swap(s.f, s.f_back);
    });
}

// Streaming then the boundary functors (main.cu:98-99), no collision: the post-boundary
// populations are read back with oracle3d_get_pops.
int oracle3d_stream_boundaries(void* p) {
    return dispatch3(static_cast<Handle3*>(p), [&](auto& s) {
        s.stream();
        std::swap(s.f, s.f_back);
        s.apply_boundaries();
    });
}

// Immersed-boundary markers (AoS [n][3], lattice units) with areas ds (or null: 1) and the
// number of multi-direct iterations (IBMManager<3>, IBMManager.cuh:54-110, 222-252).
int oracle3d_ibm_set_markers(void* p, const double* xyz, const double* ds, int n, int iters) {
    return dispatch3(static_cast<Handle3*>(p), [&](auto& s) {
        using R = typename std::remove_reference_t<decltype(s.f)>::value_type;
        std::vector<R> a(xyz, xyz + 3 * size_t(n)), d;
        if (ds) d.assign(ds, ds + n);
        s.ibm_set_markers(a.data(), ds ? d.data() : nullptr, n, iters);
    });
}

// Lagrangian forces summed over the iterations of the last step, AoS [n][3], and the total
// force on the fluid, sum_l F_l ds_l (ds = 1 under LBM_QUIRK_IBM_NO_DS).
int oracle3d_ibm_get_forces(void* p, double* lag, double* total) {
    return dispatch3(static_cast<Handle3*>(p), [&](auto& s) {
        const size_t n = s.lag_ds.size();
        double t[3] = {0, 0, 0};
        const bool no_ds = (s.cfg.quirks & LBM_QUIRK_IBM_NO_DS) != 0;
        for (size_t l = 0; l < n; l++)
            for (int d = 0; d < 3; d++) {
                const double v = double(s.lag_force_sum[d * n + l]);
                if (lag) lag[3 * l + d] = v;
                t[d] += v * (no_ds ? 1.0 : double(s.lag_ds[l]));
            }
        if (total)
            for (int d = 0; d < 3; d++) total[d] = t[d];
    });
}

const char* oracle3d_last_error(void* p) { return static_cast<Handle3*>(p)->err.c_str(); }

// ---- node-local building blocks, for the unit tests -------------------------------------

void oracle3d_tables(double* w, int* cx, int* cy, int* cz, int* opp) {
    for (int q = 0; q < Q; q++) {
        w[q] = weight<double>(q);
        cx[q] = CX[q];
        cy[q] = CY[q];
        cz[q] = CZ[q];
        opp[q] = OPP[q];
    }
}

void oracle3d_equilibrium(int prec, int literal, double rho, double ux, double uy, double uz, double* feq) {
    if (prec == 32) {
        float o[Q];
        equilibrium<float>(float(rho), float(ux), float(uy), float(uz), o, literal != 0);
        for (int q = 0; q < Q; q++) feq[q] = o[q];
    } else {
        equilibrium<double>(rho, ux, uy, uz, feq, literal != 0);
    }
}

// MRT<3> basis M and its inverse (row-major 27x27, FP64).
void oracle3d_mrt_basis(double* M, double* Minv) {
    const auto& b = mrt_basis();
    for (int i = 0; i < Q * Q; i++) {
        M[i] = b.M[i];
        Minv[i] = b.Minv[i];
    }
}

// One node through MRT<3> (FP64): g post-stream populations, u corrected velocity, F force,
// S[27] rates.
void oracle3d_collide_mrt(const double* g, const double* u, const double* F, const float* S, int literal, double* out) {
    double f[Q];
    for (int q = 0; q < Q; q++) f[q] = g[q];
    collide_mrt3<double>(f, u[0], u[1], u[2], F[0], F[1], F[2], S, literal != 0);
    for (int q = 0; q < Q; q++) out[q] = f[q];
}

// One node through CM<3, NoAdapter> (FP64): g post-stream populations, u corrected velocity.
void oracle3d_collide_cm(const double* g, const double* u, const double* F, double omega, double* out) {
    double f[Q];
    for (int q = 0; q < Q; q++) f[q] = g[q];
    collide_cm3<double>(f, u[0], u[1], u[2], F[0], F[1], F[2], omega);
    for (int q = 0; q < Q; q++) out[q] = f[q];
}

// Guo source of BGK<3> for all 27 directions (FP64).
void oracle3d_guo_source(double ux, double uy, double uz, double fx, double fy, double fz, double omega, double* out) {
    for (int q = 0; q < Q; q++) out[q] = guo_source<double>(q, ux, uy, uz, fx, fy, fz, omega);
}

}  // extern "C"
