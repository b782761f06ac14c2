#pragma once
// Velocity inlet by non-equilibrium extrapolation (VelocityDirichletNEQ.hpp:14-121,
// VelocityDirichletNEQ.cu:13-117 of the reference), PreCollision on f:
//   f_B = feq(rho_N, u_BC) + f_N - feq(rho_N, u_N), N the cell one step inward.
// Constructor rules as the reference: the mask must be exactly one axis-aligned plane on the
// domain boundary; the neighbour is inferred from that plane.
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "core/DeviceIndexList.hpp"
#include "core/interfaces/NSInterfaces.hpp"
#include "core/lattice/NSLattice.hpp"

template <typename Descriptor> class VelocityDirichletNEQ : public NSBoundaryCondition<Descriptor> {
public:
    VelocityDirichletNEQ(const LBMParams &p, std::vector<std::uint8_t> mask, DATA_TYPE ux, DATA_TYPE uy, DATA_TYPE uz)
        : params(p), h_mask(std::move(mask)), u0x(ux), u0y(uy), u0z(uz) {
        const std::int64_t n = std::int64_t(params.Nx) * params.Ny * params.Nz;
        if (std::int64_t(h_mask.size()) != n) throw std::runtime_error("VelocityDirichletNEQ mask mismatch");
        std::vector<std::int64_t> bnd = fmx::maskCells(h_mask), nei;
        if (bnd.empty()) return;
        const int nx = params.Nx, ny = params.Ny, nz = params.Nz;
        int lo[3] = {nx - 1, ny - 1, nz - 1}, hi[3] = {0, 0, 0};
        for (std::int64_t c : bnd) {
            const int v[3] = {int(c % nx), int((c / nx) % ny), int(c / (std::int64_t(nx) * ny))};
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], v[a]);
                hi[a] = std::max(hi[a], v[a]);
            }
        }
        const int ext[3] = {nx, ny, nz};
        int axis = -1, planes = 0;
        for (int a = 0; a < 3; ++a)
            if (lo[a] == hi[a]) {
                axis = a;
                ++planes;
            }
        if (planes != 1)
            throw std::runtime_error(
                "VelocityDirichletNEQ: mask must describe exactly one axis-aligned plane (x=const OR y=const OR z=const).");
        const char *names[3] = {"x", "y", "z"};
        int step;
        if (lo[axis] == 0) step = 1;
        else if (lo[axis] == ext[axis] - 1) step = -1;
        else throw std::runtime_error(std::string("VelocityDirichletNEQ: ") + names[axis] + "-plane not at boundary");
        const std::int64_t stride = axis == 0 ? 1 : (axis == 1 ? std::int64_t(nx) : std::int64_t(nx) * ny);
        if (ext[axis] < 2) throw std::runtime_error("VelocityDirichletNEQ: inferred neighbor out of domain.");
        nei.reserve(bnd.size());
        for (std::int64_t c : bnd) nei.push_back(c + step * stride);
        plane_axis = axis;
        plane_side = step > 0 ? 0 : 1;
        d_bnd.assign(bnd);
        d_nei.assign(nei);
    }

    void apply(Lattice<Descriptor> &lat, BoundaryPhase phase, int) override {
        if (phase != BoundaryPhase::PreCollision || d_bnd.size() == 0) return;
        fmx::check(fmx_ns_bc_velocity_neq(fmx::kPrec, lat.d_f_ptr(), d_bnd.data(), d_nei.data(), d_bnd.size(),
                                          double(u0x), double(u0y), double(u0z), lat.size64(), nullptr),
                   "fmx_ns_bc_velocity_neq");
    }

    const std::vector<std::uint8_t> &mask() const { return h_mask; }
    DATA_TYPE ux() const { return u0x; }
    DATA_TYPE uy() const { return u0y; }
    DATA_TYPE uz() const { return u0z; }

private:
    LBMParams params{};
    std::vector<std::uint8_t> h_mask;
    fmx::DeviceIndexList d_bnd, d_nei;
    int plane_axis = -1, plane_side = 0;
    DATA_TYPE u0x = DATA_TYPE(0), u0y = DATA_TYPE(0), u0z = DATA_TYPE(0);
};
