/*
 * state_vectors.h -- drop-in STATE_VECTOR / STATE_VECTORS
 * (state_vectors.h:57-129 of the reference): the platform state vectors
 * carried by the packets' sub-commutated ancillary words, and their linear
 * interpolation in time, through the C ABI (s1g_state_vectors,
 * s1g_interpolate_state).  The attitude is kept as the transmitted words
 * (quaternion q0..q3, angular rates); the reference's Euler-angle
 * conversion is not provided.
 */
#pragma once

#include "packet.h"

struct STATE_VECTOR {
    double time = 0.0;
    D_VEC_1D position;      /* ECEF, m   */
    D_VEC_1D velocity;      /* ECEF, m/s */
    D_VEC_1D quaternions;   /* q0..q3 as transmitted */
    D_VEC_1D angular_rate;  /* omega x, y, z as transmitted */
};

struct STATE_VECTORS {
    D_VEC_1D times;
    D_VEC_2D positions;
    D_VEC_2D velocities;
    D_VEC_2D quaternions;
    D_VEC_2D angular_rates;

    STATE_VECTORS() {}
    /* One vector per complete cycle of 64 ancillary words, kept while the
     * time stamps increase (state_vectors.cpp:19-91). */
    explicit STATE_VECTORS(PACKET_VEC_1D& packets);

    /* Linear interpolation between the bracketing vectors; the first / last
     * vector outside their span (state_vectors.cpp:141-200). */
    [[nodiscard]] STATE_VECTOR interpolate(const double& time) const;

    [[nodiscard]] size_t size() const { return times.size(); }
};
