// Point physics shared by the pointwise and the fused dense kernels
// (internal header).
#pragma once

#include "frhip_common.h"

namespace frhip {

// Inviscid Euler flux; pyfr/solvers/euler/kernels/flux.mako:3-26.
template <typename T>
__device__ __forceinline__ void inviscid_flux(const T s[5], T f[3][5], T &p,
                                              T v[3], T gm1)
{
    const T invrho = T(1)/s[0], E = s[4];
    const T rhov[3] = {s[1], s[2], s[3]};
    for (int i = 0; i < 3; ++i) v[i] = invrho*rhov[i];
    p = gm1*(E - T(0.5)*invrho*((rhov[0])*(rhov[0]) + (rhov[1])*(rhov[1]) +
                               (rhov[2])*(rhov[2])));
    for (int i = 0; i < 3; ++i) {
        f[i][0] = rhov[i];
        f[i][4] = (E + p)*v[i];
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            f[i][j + 1] = (i == j) ? rhov[i]*v[j] + p : rhov[i]*v[j];
}

}  // namespace frhip
