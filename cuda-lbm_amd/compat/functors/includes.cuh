// includes.cuh — functor umbrella header (src/functors/includes.cuh). Boundary conditions
// are selected per node by BC_flag and applied inside the fused kernel; the initial
// condition functor scenarios may default to is DefaultInit.
#ifndef FUNCTOR_INCLUDES_H
#define FUNCTOR_INCLUDES_H
#include "./initialConditions/defaultInit.cuh"
#endif
