#pragma once
// Smooth tanh inlet pulse (ADDirichletTanhPulse.hpp:12-38, ADDirichletTanhPulse.cu:11-100 of
// the reference), PreCollision on g: g <- phi_in(step) * w on the mask with
//   phi_in = base + (DATA_TYPE) (0.5 (tanh((s - t_on)/k) - tanh((s - t_off)/k))) * amp,
// the pulse evaluated in double on the host.
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "core/DeviceIndexList.hpp"
#include "core/interfaces/ADInterfaces.hpp"
#include "core/lattice/ADLattice.hpp"

template <typename Descriptor> class ADDirichletTanhPulse : public ADBoundaryCondition<Descriptor> {
public:
    ADDirichletTanhPulse(const LBMParams &p, std::vector<std::uint8_t> mask, DATA_TYPE phi_base_, DATA_TYPE phi_amp_,
                         int t_on_step, int t_off_step, DATA_TYPE k_step)
        : params(p), mask_h(std::move(mask)), phi_base(phi_base_), phi_amp(phi_amp_), t_on(t_on_step),
          t_off(t_off_step), k(k_step) {
        if (std::int64_t(mask_h.size()) != std::int64_t(p.Nx) * p.Ny * p.Nz)
            throw std::runtime_error("ADDirichletTanhPulse mask mismatch");
        d_cells.assign(fmx::maskCells(mask_h));
    }
    DATA_TYPE valueAt(int step) const {
        const double pulse = 0.5 * (std::tanh((double(step) - double(t_on)) / double(k)) -
                                    std::tanh((double(step) - double(t_off)) / double(k)));
        return phi_base + static_cast<DATA_TYPE>(pulse) * phi_amp;
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
    fmx::DeviceIndexList d_cells;
    DATA_TYPE phi_base = DATA_TYPE(0), phi_amp = DATA_TYPE(0);
    int t_on = 0, t_off = 0;
    DATA_TYPE k = DATA_TYPE(1);
};
