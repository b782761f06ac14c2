#pragma once
// AD operator plugin API, as ADInterfaces.hpp:9-22 of the reference.
#include "core/Params.hpp"

template <typename Descriptor> class ADLattice;

template <typename Descriptor> class ADCollisionOperator {
public:
    virtual ~ADCollisionOperator() = default;
    virtual void initEq(ADLattice<Descriptor> &lat) = 0;
    virtual void collideStream(ADLattice<Descriptor> &lat) = 0;  // g -> g_new
};

template <typename Descriptor> class ADBoundaryCondition {
public:
    virtual ~ADBoundaryCondition() = default;
    virtual void apply(ADLattice<Descriptor> &lat, BoundaryPhase phase, int step) = 0;
};
