#pragma once
// ADLattice<D3Q7>: the advection-diffusion container of ADLattice.hpp:17-190, over the C ABI.
// Same interface and step order (ADLattice.hpp:144-162); populations are structure-of-arrays
// on the device; the relaxation-time field starts at LBMParams::tau_ad (the reference leaves
// it uninitialised until setTauField, ADLattice.hpp:46).
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <type_traits>
#include <utility>
#include <vector>

#include "core/DeviceHelpers.hpp"
#include "core/Params.hpp"
#include "core/Types.hpp"
#include "core/interfaces/ADInterfaces.hpp"

template <typename Descriptor> class Lattice;

template <typename Descriptor> class ADLattice {
public:
    explicit ADLattice(const LBMParams &p) : params(p) {
        if (params.Nx <= 0 || params.Ny <= 0 || params.Nz <= 0) throw std::runtime_error("ADLattice: invalid dimensions");
        nCells = std::int64_t(params.Nx) * params.Ny * params.Nz;
        bytesG = std::size_t(nCells) * Descriptor::Q * sizeof(DATA_TYPE);
        bytesCells = std::size_t(nCells) * sizeof(DATA_TYPE);
        d_g = static_cast<DATA_TYPE *>(fmx::mallocBytes(bytesG));
        d_g_new = static_cast<DATA_TYPE *>(fmx::mallocBytes(bytesG));
        for (DATA_TYPE **p : {&d_phi, &d_Sphi, &d_ux_adv, &d_uy_adv, &d_uz_adv, &d_tau}) {
            *p = static_cast<DATA_TYPE *>(fmx::mallocBytes(bytesCells));
            fmx::memsetBytes(*p, 0, bytesCells);
        }
        setTauField(std::vector<DATA_TYPE>(std::size_t(nCells), params.tau_ad));
        h_phi.assign(std::size_t(nCells), DATA_TYPE(0));
    }

    ~ADLattice() {
        for (void *p : {(void *)d_g, (void *)d_g_new, (void *)d_phi, (void *)d_Sphi, (void *)d_ux_adv, (void *)d_uy_adv,
                        (void *)d_uz_adv, (void *)d_tau})
            fmx::freeBytes(p);
    }

    ADLattice(const ADLattice &) = delete;
    ADLattice &operator=(const ADLattice &) = delete;

    const LBMParams &P() const { return params; }
    int Size() const { return int(nCells); }
    std::int64_t size64() const { return nCells; }
    fmx_grid grid() const {
        return fmx::gridOf(params.Nx, params.Ny, params.Nz, params.periodicX, params.periodicY, params.periodicZ);
    }

    DATA_TYPE *d_g_ptr() { return d_g; }
    DATA_TYPE *d_g_new_ptr() { return d_g_new; }
    DATA_TYPE *d_phi_ptr() { return d_phi; }
    DATA_TYPE *d_Sphi_ptr() { return d_Sphi; }
    DATA_TYPE *d_tau_ptr() { return d_tau; }
    DATA_TYPE *d_ux_adv_ptr() { return d_ux_adv; }
    DATA_TYPE *d_uy_adv_ptr() { return d_uy_adv; }
    DATA_TYPE *d_uz_adv_ptr() { return d_uz_adv; }

    // ADLattice.hpp:89-92
    void zeroSource() { fmx::memsetBytes(d_Sphi, 0, bytesCells); }

    // ADLattice.hpp:95-100
    void setTauField(const std::vector<DATA_TYPE> &tau) {
        if (std::int64_t(tau.size()) != nCells) throw std::runtime_error("ADLattice: tau field size mismatch");
        fmx::memcpyHtoD(d_tau, tau.data(), bytesCells);
        h_tau = tau;
    }
    const std::vector<DATA_TYPE> &tauFieldHost() const { return h_tau; }

    // ADLattice.hpp:104-111
    template <typename FluidDesc> void setAdvectionFromFluidDevice(const Lattice<FluidDesc> &fluid) {
        fmx::memcpyD2D(d_ux_adv, fluid.d_ux_ptr(), bytesCells);
        fmx::memcpyD2D(d_uy_adv, fluid.d_uy_ptr(), bytesCells);
        fmx::memcpyD2D(d_uz_adv, fluid.d_uz_ptr(), bytesCells);
    }

    void setCollisionOperator(std::shared_ptr<ADCollisionOperator<Descriptor>> op) { collision = std::move(op); }

    template <typename TBoundary, typename... Args> TBoundary &addBoundary(Args &&...args) {
        static_assert(std::is_base_of_v<ADBoundaryCondition<Descriptor>, TBoundary>,
                      "AD boundary must derive from ADBoundaryCondition");
        auto ptr = std::make_unique<TBoundary>(std::forward<Args>(args)...);
        TBoundary &ref = *ptr;
        boundaries.emplace_back(std::move(ptr));
        return ref;
    }
    const std::vector<std::unique_ptr<ADBoundaryCondition<Descriptor>>> &boundaryList() const { return boundaries; }

    // ADLattice.hpp:134-141 (device-to-device copy; the reference passes the HtoD kind here)
    void initEquilibrium() {
        if (!collision) throw std::runtime_error("ADLattice: collision operator not set, could not initEq()!");
        collision->initEq(*this);
        fmx::memcpyD2D(d_g_new, d_g, bytesG);
    }

    // ADLattice.hpp:144-162
    void step(int stepIndex) {
        if (!collision) throw std::runtime_error("ADLattice: collision operator not set, could not step()!");
        for (auto &bc : boundaries) bc->apply(*this, BoundaryPhase::PreCollision, stepIndex);
        collision->collideStream(*this);
        for (auto &bc : boundaries) bc->apply(*this, BoundaryPhase::PostStreaming, stepIndex);
        std::swap(d_g, d_g_new);
    }

    // ADLattice.hpp:166-169
    void downloadPhi() { fmx::memcpyDtoH(h_phi.data(), d_phi, bytesCells); }
    const std::vector<DATA_TYPE> &phiHost() const { return h_phi; }

    std::vector<DATA_TYPE> downloadPopulationsAoS() {
        DATA_TYPE *tmp = static_cast<DATA_TYPE *>(fmx::mallocBytes(bytesG));
        fmx::check(fmx_pack_aos(fmx::kPrec, d_g, tmp, Descriptor::Q, nCells, nullptr), "fmx_pack_aos");
        std::vector<DATA_TYPE> h(std::size_t(nCells) * Descriptor::Q);
        fmx::memcpyDtoH(h.data(), tmp, bytesG);
        fmx::freeBytes(tmp);
        return h;
    }

    void swapPopulations() { std::swap(d_g, d_g_new); }
    // replace the concentration buffer (fused steppers ping-pong it); returns the previous one
    DATA_TYPE *exchangePhi(DATA_TYPE *p) {
        std::swap(d_phi, p);
        return p;
    }

private:
    LBMParams params{};
    std::int64_t nCells = 0;
    std::size_t bytesG = 0, bytesCells = 0;
    DATA_TYPE *d_g = nullptr, *d_g_new = nullptr;
    DATA_TYPE *d_phi = nullptr, *d_Sphi = nullptr;
    DATA_TYPE *d_ux_adv = nullptr, *d_uy_adv = nullptr, *d_uz_adv = nullptr;
    DATA_TYPE *d_tau = nullptr;
    std::vector<DATA_TYPE> h_phi, h_tau;
    std::shared_ptr<ADCollisionOperator<Descriptor>> collision;
    std::vector<std::unique_ptr<ADBoundaryCondition<Descriptor>>> boundaries;
};
