// IBM_generators.cuh — create_cylinder (src/IBM/IBM_generators.cu:5-25 behaviour): num_pts
// equispaced markers on a circle of radius r about (cx, cy), zero marker velocities.
#ifndef IBM_GENERATORS_H
#define IBM_GENERATORS_H

#include <math.h>

#include "IBMBody.cuh"

inline IBMBody create_cylinder(float cx, float cy, float r, int num_pts = 16) {
    IBMBody b{num_pts, new float[2 * num_pts], new float[2 * num_pts]};
    const float step = 2.0f * float(M_PI) / float(num_pts);
    for (int i = 0; i < num_pts; i++) {
        b.points[2 * i] = cx + r * cosf(float(i) * step);
        b.points[2 * i + 1] = cy + r * sinf(float(i) * step);
        b.velocities[2 * i] = b.velocities[2 * i + 1] = 0.0f;
    }
    return b;
}

#endif
