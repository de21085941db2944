// adapters.cuh — adapter type names of src/core/collision/adapters.cuh; each maps to an
// lbm_config.adapter value.
#ifndef ADAPTERS_H
#define ADAPTERS_H
#include "../../../../include/lbm_hip.h"

struct MomentInfo {  // lbm.cuh:25-31
    float rho_avg_norm;
    float j_avg_norm;
    float pi_avg_norm;
};

struct AdapterBase {
    static constexpr int i_star = 5;
};
struct NoAdapter : AdapterBase {
    static constexpr int kind = LBM_NO_ADAPTER;
};
struct OptimalAdapter : AdapterBase {
    static constexpr int kind = LBM_OPTIMAL_ADAPTER;
};
struct EmpiricalAdapter : AdapterBase {  // 3-D only in the reference
    static constexpr int kind = -1;
};
#endif
