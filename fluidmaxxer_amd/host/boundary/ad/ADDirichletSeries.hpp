#pragma once
// Inlet concentration from a knot series (no reference counterpart; generalises
// ADDirichletTanhPulse to an arbitrary signal), PreCollision on g: g <- phi_in(step) * w on the
// mask, phi_in linear between knots placed every dt_knot steps (last value held afterwards),
// evaluated in double and rounded to DATA_TYPE once.
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "core/DeviceIndexList.hpp"
#include "core/interfaces/ADInterfaces.hpp"
#include "core/lattice/ADLattice.hpp"

template <typename Descriptor> class ADDirichletSeries : public ADBoundaryCondition<Descriptor> {
public:
    ADDirichletSeries(const LBMParams &p, std::vector<std::uint8_t> mask, std::vector<double> knots, double dt_knot)
        : params(p), mask_h(std::move(mask)), knots_(std::move(knots)), dt(dt_knot) {
        if (std::int64_t(mask_h.size()) != std::int64_t(p.Nx) * p.Ny * p.Nz)
            throw std::runtime_error("ADDirichletSeries mask mismatch");
        if (knots_.empty() || !(dt > 0.0)) throw std::runtime_error("ADDirichletSeries: empty knots or dt <= 0");
        d_cells.assign(fmx::maskCells(mask_h));
    }
    DATA_TYPE valueAt(int step) const {
        const double s = double(step) / dt;
        if (s <= 0.0) return DATA_TYPE(knots_[0]);
        const std::size_t k = std::size_t(std::floor(s));
        if (k + 1 >= knots_.size()) return DATA_TYPE(knots_.back());
        const double a = s - double(k);
        return static_cast<DATA_TYPE>(knots_[k] * (1.0 - a) + knots_[k + 1] * a);
    }
    void apply(ADLattice<Descriptor> &lat, BoundaryPhase phase, int step) override {
        if (phase != BoundaryPhase::PreCollision || d_cells.size() == 0) return;
        fmx::check(fmx_ad_bc_dirichlet(fmx::kPrec, lat.d_g_ptr(), d_cells.data(), d_cells.size(),
                                       double(valueAt(step)), lat.size64(), nullptr),
                   "fmx_ad_bc_dirichlet");
    }
    const std::vector<std::uint8_t> &mask() const { return mask_h; }

private:
    LBMParams params{};
    std::vector<std::uint8_t> mask_h;
    std::vector<double> knots_;
    double dt = 1.0;
    fmx::DeviceIndexList d_cells;
};
