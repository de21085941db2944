// collision.cuh — collision-operator type names of src/core/collision (BGK<2>, MRT<2>,
// CM<2, Adapter>). They select lbm_config.collision / adapter; the operators themselves are
// the fused kernels of liblbm_hip.so.
#ifndef LBM_COLLISION_H
#define LBM_COLLISION_H
#include "../lbm_constants.cuh"
#include "adapters.cuh"

template <int dim>
struct BGK {
    static_assert(dim == 2, "this build is D2Q9");
    static constexpr int kind = LBM_BGK;
    static constexpr int adapter = LBM_NO_ADAPTER;
};
template <int dim>
struct MRT {
    static_assert(dim == 2, "this build is D2Q9");
    static constexpr int kind = LBM_MRT;
    static constexpr int adapter = LBM_NO_ADAPTER;
};
template <int dim, typename AdapterType = NoAdapter>
struct CM {
    static_assert(dim == 2, "this build is D2Q9");
    static constexpr int kind = LBM_CM;
    static constexpr int adapter = AdapterType::kind;
};
#endif
