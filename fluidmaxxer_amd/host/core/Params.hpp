#pragma once
// LBMParams and BoundaryPhase, field for field as Params.hpp:6-44 of the reference.
#include "core/Types.hpp"

struct LBMParams {
    // domain size
    int Nx, Ny, Nz;
    // NS lattice: density and relaxation time
    DATA_TYPE rho0;
    DATA_TYPE tau;
    // AD lattice
    DATA_TYPE tau_ad;
    DATA_TYPE phi0;
    DATA_TYPE minConcentration = 1e-8;
    // periodic streaming
    bool periodicX = true;
    bool periodicY = true;
    bool periodicZ = true;
    // adsorption
    bool enableAdsorption = false;
    DATA_TYPE KHenry;  // Henry constant
    DATA_TYPE kA;      // LDF rate (per step), linear isotherm
    DATA_TYPE qMax;       // Langmuir capacity
    DATA_TYPE bLangmuir;  // Langmuir affinity
    DATA_TYPE kLDF;       // LDF rate (per step), Langmuir
    // porosity defaults
    DATA_TYPE porosityValue = 0.67;
    DATA_TYPE betaPorous = 1;
};

enum class BoundaryPhase { PreCollision, PostStreaming };
