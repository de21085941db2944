// lbm_constants.cuh — the names scenario headers use from src/core/lbm_constants.cuh:
// lattice tables, index helpers, viscosity helpers and the BC_flag enum (same names and
// order). Host-facing vector fields are AoS ([node][2]); populations are SoA.
#ifndef LBM_CONSTANTS_H
#define LBM_CONSTANTS_H

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "defines.hpp"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

constexpr uint8_t dimensions = 2;
constexpr uint8_t quadratures = 9;
const float cs = 0.57735026918962576f;
const float cs2 = 1.0f / 3.0f;

const float h_weights[9] = {4.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f,
                            1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f};
const int h_C[18] = {0, 0, 1, 0, 0, 1, -1, 0, 0, -1, 1, 1, -1, 1, -1, -1, 1, -1};
const int h_OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

// Device copies for functors that read them; set by LBM<2>::init.
static __constant__ float WEIGHTS[9];
static __constant__ int C[18];
static __constant__ int OPP[9];
static __constant__ float vis;
static __constant__ float tau;
static __constant__ float omega;

constexpr int num_nodes = NX * NY;

__host__ __device__ __forceinline__ int get_node_index(int node, int quadrature = 0) {
    return quadrature * num_nodes + node;
}
__host__ __device__ __forceinline__ int get_vec_index(int node, int component) { return node * dimensions + component; }
__host__ __device__ __forceinline__ int get_node_from_coords(int x, int y, int z = 0) { return y * NX + x; }
__host__ __device__ __forceinline__ void get_coords_from_node(int node, int& x, int& y, int& z) {
    x = node % NX;
    y = node / NX;
    z = 0;
}

constexpr inline float viscosity_to_tau(float v) { return 3 * v + 0.5f; }
constexpr inline float tau_to_viscosity(float t) { return (t - 0.5f) / 3.0f; }
constexpr float compute_reynolds(float u_max, float domain_size, float viscosity) { return (u_max * domain_size) / viscosity; }

enum BC_flag {
    FLUID,
    BOUNCE_BACK,
    ZOU_HE_TOP,
    ZOU_HE_LEFT,
    ZOU_HE_TOP_LEFT_TOP_INFLOW,
    ZOU_HE_TOP_RIGHT_TOP_INFLOW,
    CYLINDER,
    ZG_OUTFLOW,
    PRESSURE_OUTLET,
    REGULARIZED_INLET_TOP,
    REGULARIZED_INLET_LEFT,
    REGULARIZED_BOUNCE_BACK,
    REGULARIZED_BOUNCE_BACK_CORNER,
    EXTRAPOLATED_CORNER_EDGE,
    CORNER_EDGE_BOUNCE_BACK,
    PRESSURE_INLET_LEFT,
    GUO_VELOCITY_INLET,
    GUO_PRESSURE_OUTLET,
    REGULARIZED_OUTLET,
    MOVING_WALL_BOUNCE_BACK,
};

#endif  // LBM_CONSTANTS_H
