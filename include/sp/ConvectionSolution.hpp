// ConvectionSolution.hpp -- the initial condition / exact solution of the advection-diffusion
// test problem, with the signatures of src/Convection/ConvectionSolution.h:8-52:
//   exactQ(nu0, nufreq, cx, cy, cz, x, y, z, t)
//   fillQ(field, nu0, nufreq, cx, cy, cz, t, xstart, xend, ystart, yend, zstart, zend)
// q(x, t) = prod_d sin(2 pi (x_d - c_d t)) * exp(-12 pi^2 a(t)), a(t) = int_0^t nu, with
// nu(t) = nu0 (1 + sin(nufreq t) / 2); cell-centred points x_i = start + (i + 1/2) h.
// pi is the reference's 15-digit constant, so fields agree with it bit for bit.
#pragma once

#include <cmath>

namespace sp {

inline double exactQ(double nu0, double nufreq, double cx, double cy, double cz, double x, double y, double z,
                     double t)
{
    const double pi = 3.14159265358979;
    // integral of nu over [0, t]; constant nu when nufreq is 0
    const double a = nufreq > 0 ? nu0 * t - (nu0 / (2. * nufreq)) * (std::cos(nufreq * t) - 1.) : nu0 * t;
    const double wave = std::sin(2. * pi * (x - cx * t)) * std::sin(2. * pi * (y - cy * t)) *
                        std::sin(2. * pi * (z - cz * t));
    return wave * std::exp(-12. * pi * pi * a);
}

// Fills the interior of `q` (host element access) with exactQ at time t on the box
// [xstart, xend] x [ystart, yend] x [zstart, zend].
template <typename Field>
void fillQ(Field& q, double nu0, double nufreq, double cx, double cy, double cz, double t, double xstart,
           double xend, double ystart, double yend, double zstart, double zend)
{
    const auto d = q.calculationDomain();
    const int ni = d.iSize(), nj = d.jSize(), nk = d.kSize();
    const double hx = (xend - xstart) / ni / 2., hy = (yend - ystart) / nj / 2., hz = (zend - zstart) / nk / 2.;
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int k = 0; k < nk; ++k)
                q(i, j, k) = exactQ(nu0, nufreq, cx, cy, cz, xstart + (2 * i + 1) * hx, ystart + (2 * j + 1) * hy,
                                    zstart + (2 * k + 1) * hz, t);
}

}  // namespace sp
