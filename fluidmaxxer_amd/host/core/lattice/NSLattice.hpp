#pragma once
// Lattice<D3Q19>: the NS lattice container of NSLattice.hpp:14-264, over the C ABI.
//
// Same interface: owns f, f_new, obstacle, porosity (device), holds one collision
// operator and an ordered list of boundary conditions, step() runs
// PreCollision BCs -> collideStream -> PostStreaming BCs -> swap (NSLattice.hpp:188-209).
// Differences: populations are structure-of-arrays on the device (d_f_ptr()[q*N + c]);
// downloadPopulationsAoS() returns the reference layout; the redundant f -> f_new copy
// before each collide (NSLattice.hpp:197-198) is not made, because every collide writes
// all cells; indices are 64-bit.
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <type_traits>
#include <utility>
#include <vector>

#include "core/DeviceHelpers.hpp"
#include "core/Params.hpp"
#include "core/Types.hpp"
#include "core/interfaces/NSInterfaces.hpp"

template <typename Descriptor> class Lattice {
public:
    explicit Lattice(const LBMParams &p) : params(p) {
        if (params.Nx <= 0 || params.Ny <= 0 || params.Nz <= 0)
            throw std::runtime_error("NSLattice: invalid dimensions");
        nCells = std::int64_t(params.Nx) * params.Ny * params.Nz;
        bytesF = std::size_t(nCells) * Descriptor::Q * sizeof(DATA_TYPE);
        bytesCells = std::size_t(nCells) * sizeof(DATA_TYPE);
        d_f = static_cast<DATA_TYPE *>(fmx::mallocBytes(bytesF));
        d_f_new = static_cast<DATA_TYPE *>(fmx::mallocBytes(bytesF));
        d_obstacle = static_cast<std::uint8_t *>(fmx::mallocBytes(std::size_t(nCells)));
        fmx::memsetBytes(d_obstacle, 0, std::size_t(nCells));
        d_porosity = static_cast<DATA_TYPE *>(fmx::mallocBytes(bytesCells));
        fmx::memsetBytes(d_porosity, 0, bytesCells);
        h_rho.assign(std::size_t(nCells), DATA_TYPE(params.rho0));
        h_ux.assign(std::size_t(nCells), DATA_TYPE(0));
        h_uy.assign(std::size_t(nCells), DATA_TYPE(0));
        h_uz.assign(std::size_t(nCells), DATA_TYPE(0));
        h_obstacle.assign(std::size_t(nCells), 0);
    }

    ~Lattice() {
        for (void *p : {(void *)d_f, (void *)d_f_new, (void *)d_obstacle, (void *)d_porosity, (void *)d_rho,
                        (void *)d_ux, (void *)d_uy, (void *)d_uz})
            fmx::freeBytes(p);
    }

    Lattice(const Lattice &) = delete;
    Lattice &operator=(const Lattice &) = delete;

    std::int64_t cellIndex(int x, int y, int z) const {
        return x + std::int64_t(params.Nx) * (y + std::int64_t(params.Ny) * z);
    }

    // NSLattice.hpp:88-114
    void setPorosityFromMask(const std::vector<std::uint8_t> &mask, DATA_TYPE phiMasked, DATA_TYPE phiElse) {
        if (std::int64_t(mask.size()) != nCells)
            throw std::runtime_error("setPorosityFromMask: mask size != lattice size");
        std::vector<DATA_TYPE> h(mask.size(), phiElse);
        for (std::size_t i = 0; i < mask.size(); ++i)
            if (mask[i]) h[i] = phiMasked;
        setPorosityField(h);
    }
    // per-cell porosity (PorousBGKCollision) or solid fraction (PartialBounceBackCollision)
    void setPorosityField(const std::vector<DATA_TYPE> &field) {
        if (std::int64_t(field.size()) != nCells) throw std::runtime_error("setPorosityField: size != lattice size");
        fmx::memcpyHtoD(d_porosity, field.data(), bytesCells);
    }

    const LBMParams &P() const { return params; }
    int Size() const { return int(nCells); }
    std::int64_t size64() const { return nCells; }
    fmx_grid grid() const {
        return fmx::gridOf(params.Nx, params.Ny, params.Nz, params.periodicX, params.periodicY, params.periodicZ);
    }

    DATA_TYPE *d_f_ptr() { return d_f; }
    const DATA_TYPE *d_f_ptr() const { return d_f; }
    DATA_TYPE *d_f_new_ptr() { return d_f_new; }
    const DATA_TYPE *d_f_new_ptr() const { return d_f_new; }
    std::uint8_t *d_obstacle_ptr() { return d_obstacle; }
    const std::uint8_t *d_obstacle_ptr() const { return d_obstacle; }
    DATA_TYPE *d_rho_ptr() { return d_rho; }
    const DATA_TYPE *d_rho_ptr() const { return d_rho; }
    DATA_TYPE *d_ux_ptr() { return d_ux; }
    const DATA_TYPE *d_ux_ptr() const { return d_ux; }
    DATA_TYPE *d_uy_ptr() { return d_uy; }
    const DATA_TYPE *d_uy_ptr() const { return d_uy; }
    DATA_TYPE *d_uz_ptr() { return d_uz; }
    const DATA_TYPE *d_uz_ptr() const { return d_uz; }
    DATA_TYPE *d_porosity_ptr() { return d_porosity; }
    const DATA_TYPE *d_porosity_ptr() const { return d_porosity; }

    // NSLattice.hpp:151-156
    void setObstacleMask(const std::vector<std::uint8_t> &mask) {
        if (std::int64_t(mask.size()) != nCells) throw std::runtime_error("NSLattice::setObstacleMask: mask size mismatch");
        fmx::memcpyHtoD(d_obstacle, mask.data(), std::size_t(nCells));
    }
    const std::vector<std::uint8_t> &obstacleMaskHost() {
        downloadObstacle();
        return h_obstacle;
    }

    void setCollisionOperator(std::shared_ptr<NSCollisionOperator<Descriptor>> op) { collision = std::move(op); }
    const std::shared_ptr<NSCollisionOperator<Descriptor>> &collisionOperator() const { return collision; }

    template <typename TBoundary, typename... Args> TBoundary &addBoundary(Args &&...args) {
        static_assert(std::is_base_of_v<NSBoundaryCondition<Descriptor>, TBoundary>,
                      "addBoundary: TBoundary must derive from NSBoundaryCondition<Descriptor>");
        auto ptr = std::make_unique<TBoundary>(std::forward<Args>(args)...);
        TBoundary &ref = *ptr;
        boundaries.emplace_back(std::move(ptr));
        return ref;
    }
    const std::vector<std::unique_ptr<NSBoundaryCondition<Descriptor>>> &boundaryList() const { return boundaries; }

    // NSLattice.hpp:177-185
    void initEquilibrium() {
        if (!collision) throw std::runtime_error("NSLattice::initEquilibrium: collision operator not set");
        collision->initEq(*this);
        fmx::memcpyD2D(d_f_new, d_f, bytesF);
    }

    // NSLattice.hpp:188-209
    void step(int stepIndex) {
        if (!collision) throw std::runtime_error("NSLattice: collision operator not set, could not step()!");
        for (auto &bc : boundaries) bc->apply(*this, BoundaryPhase::PreCollision, stepIndex);
        collision->collideStream(*this);
        for (auto &bc : boundaries) bc->apply(*this, BoundaryPhase::PostStreaming, stepIndex);
        std::swap(d_f, d_f_new);
    }

    // Macros.cu:95-119
    void computeMacroscopicDevice() {
        static_assert(Descriptor::Q == 19, "moments are implemented for D3Q19");
        ensureMacroBuffers();
        fmx_grid g = grid();
        fmx::check(fmx_ns_moments(fmx::kPrec, d_f, d_rho, d_ux, d_uy, d_uz, &g, nullptr), "fmx_ns_moments");
    }
    // Macros.cu:63-93
    void downloadMacroscopic() {
        computeMacroscopicDevice();
        fmx::memcpyDtoH(h_rho.data(), d_rho, bytesCells);
        fmx::memcpyDtoH(h_ux.data(), d_ux, bytesCells);
        fmx::memcpyDtoH(h_uy.data(), d_uy, bytesCells);
        fmx::memcpyDtoH(h_uz.data(), d_uz, bytesCells);
    }
    void downloadObstacle() { fmx::memcpyDtoH(h_obstacle.data(), d_obstacle, std::size_t(nCells)); }
    // populations in the reference layout f[c*Q + q]
    std::vector<DATA_TYPE> downloadPopulationsAoS() {
        DATA_TYPE *tmp = static_cast<DATA_TYPE *>(fmx::mallocBytes(bytesF));
        fmx::check(fmx_pack_aos(fmx::kPrec, d_f, tmp, Descriptor::Q, nCells, nullptr), "fmx_pack_aos");
        std::vector<DATA_TYPE> h(std::size_t(nCells) * Descriptor::Q);
        fmx::memcpyDtoH(h.data(), tmp, bytesF);
        fmx::freeBytes(tmp);
        return h;
    }
    void uploadPopulationsAoS(const std::vector<DATA_TYPE> &h) {
        if (h.size() != std::size_t(nCells) * Descriptor::Q) throw std::runtime_error("uploadPopulationsAoS: size");
        DATA_TYPE *tmp = static_cast<DATA_TYPE *>(fmx::mallocBytes(bytesF));
        fmx::memcpyHtoD(tmp, h.data(), bytesF);
        fmx::check(fmx_unpack_aos(fmx::kPrec, tmp, d_f, Descriptor::Q, nCells, nullptr), "fmx_unpack_aos");
        fmx::checkThrow("uploadPopulationsAoS");
        fmx::freeBytes(tmp);
    }

    const std::vector<DATA_TYPE> &rhoHost() const { return h_rho; }
    const std::vector<DATA_TYPE> &uxHost() const { return h_ux; }
    const std::vector<DATA_TYPE> &uyHost() const { return h_uy; }
    const std::vector<DATA_TYPE> &uzHost() const { return h_uz; }
    const std::vector<std::uint8_t> &obstacleHost() const { return h_obstacle; }

    // swap the population buffers without stepping (fused steppers that write f_new)
    void swapPopulations() { std::swap(d_f, d_f_new); }

private:
    void ensureMacroBuffers() {
        if (!d_rho) d_rho = static_cast<DATA_TYPE *>(fmx::mallocBytes(bytesCells));
        if (!d_ux) d_ux = static_cast<DATA_TYPE *>(fmx::mallocBytes(bytesCells));
        if (!d_uy) d_uy = static_cast<DATA_TYPE *>(fmx::mallocBytes(bytesCells));
        if (!d_uz) d_uz = static_cast<DATA_TYPE *>(fmx::mallocBytes(bytesCells));
    }

    LBMParams params{};
    std::int64_t nCells = 0;
    std::size_t bytesF = 0, bytesCells = 0;
    DATA_TYPE *d_f = nullptr, *d_f_new = nullptr;
    std::uint8_t *d_obstacle = nullptr;
    DATA_TYPE *d_rho = nullptr, *d_ux = nullptr, *d_uy = nullptr, *d_uz = nullptr;
    DATA_TYPE *d_porosity = nullptr;
    std::vector<DATA_TYPE> h_rho, h_ux, h_uy, h_uz;
    std::vector<std::uint8_t> h_obstacle;
    std::shared_ptr<NSCollisionOperator<Descriptor>> collision;
    std::vector<std::unique_ptr<NSBoundaryCondition<Descriptor>>> boundaries;
};
