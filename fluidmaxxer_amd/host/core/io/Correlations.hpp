#pragma once
// The physical correlations main.cpp uses to set up its scenario (main.cpp:43-99 of the reference),
// as free functions over DATA_TYPE:
//   computeEffectiveDiffusivity   Mackie-Meares (1955): D_eff = eps / (2 - eps)^2 * D_m
//   calculateInletVelocityFromPeclet   u_s = Pe D / d_p * eps (superficial velocity)
//   calculateInterstitialVelocity      u_i = u_s / eps
//   calculatekFilm   film mass-transfer coefficient: Sc = nu / D,
//                    Sh = 1.09 eps_p (eps_p Re)^0.33 Sc^0.33, k_film = Sh D / d_p
//   calculatekLDF    k_ldf = 3 (1 - eps_bed) / (d_p / 2) * k_film, in 1/step: * dt
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <type_traits>

#include "core/Types.hpp"

inline DATA_TYPE computeEffectiveDiffusivity(DATA_TYPE Dm, DATA_TYPE epsParticle) {
    const DATA_TYPE numerator = epsParticle;
    const DATA_TYPE denominator = std::pow(2.0 - epsParticle, 2.0);
    return (numerator / denominator) * Dm;
}

inline DATA_TYPE calculateInletVelocityFromPeclet(DATA_TYPE peclet, DATA_TYPE diffusivity, DATA_TYPE particleDiameter,
                                                  DATA_TYPE porosity) {
    if (peclet <= DATA_TYPE(0) || diffusivity <= DATA_TYPE(0) || particleDiameter <= DATA_TYPE(0))
        throw std::invalid_argument("Pe, D, and Lc must be positive.");
    if (porosity <= DATA_TYPE(0) || porosity > DATA_TYPE(1)) throw std::invalid_argument("Porosity eps must be in (0,1].");
    return (peclet * diffusivity / particleDiameter) * porosity;
}

inline DATA_TYPE calculateInterstitialVelocity(DATA_TYPE superficialVelocity, DATA_TYPE porosity) {
    if (superficialVelocity < DATA_TYPE(0)) throw std::invalid_argument("Superficial velocity must be non-negative.");
    if (porosity <= DATA_TYPE(0) || porosity > DATA_TYPE(1)) throw std::invalid_argument("Porosity eps must be in (0,1].");
    return superficialVelocity / porosity;
}

inline DATA_TYPE calculatekFilm(DATA_TYPE Re, DATA_TYPE kinematicViscosity, DATA_TYPE soluteDiffusivity,
                                DATA_TYPE particlePorosity, DATA_TYPE particleDiameter) {
    const DATA_TYPE Sc = kinematicViscosity / soluteDiffusivity;
    const DATA_TYPE Sh = 1.09 * particlePorosity * std::pow(particlePorosity * Re, 0.33) * std::pow(Sc, 0.33);
    std::printf("[kFilm] Re = %.2f Sc = %.2f Sh = %.2f\n", double(Re), double(Sc), double(Sh));
    return Sh * soluteDiffusivity / particleDiameter;
}

inline DATA_TYPE calculatekLDF(DATA_TYPE geometryPorosity, DATA_TYPE particleDiameter, DATA_TYPE k_film,
                               DATA_TYPE timestep) {
    const DATA_TYPE a_s = 3 * (1 - geometryPorosity) / (particleDiameter / 2);
    const DATA_TYPE k_ldf = a_s * k_film;
    const DATA_TYPE k_ldf_LU = k_ldf * timestep;
    std::printf("[config] Calculated k_ldf = %g -> %g [1/timestep]\n", double(k_ldf), double(k_ldf_LU));
    return k_ldf_LU;
}
