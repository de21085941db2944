// IBMBody.cuh — the marker-set type of src/IBM/IBMBody.cuh:7-45 (host arrays, AoS [n][2]).
#ifndef IBM_BODY_H
#define IBM_BODY_H

#include <math.h>

struct IBMBody {
    int num_points;
    float* points;
    float* velocities;
};

inline void h_ibm_free(IBMBody body) {
    delete[] body.points;
    delete[] body.velocities;
}

#endif
