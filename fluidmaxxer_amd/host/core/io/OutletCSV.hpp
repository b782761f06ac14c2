#pragma once
// Outlet breakthrough log (OutletCSV.hpp:12-67 of the reference): header
// "step,time_s,phi_out_avg,phi_out_min,phi_out_max", one row per log(step) with the FP64
// mean/min/max of the AD lattice's phi over the outlet mask.
#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <vector>

#include "core/DeviceIndexList.hpp"
#include "core/Types.hpp"

// The statistics are reduced on the device (fmx_reduce_outlet_mean: FP64, a fixed summation order)
// over the outlet cells, kept there as an index list, so a log call moves three doubles instead of
// the whole concentration field; breakthrough curves logged every step stay cheap.
template <typename ADLat, typename UnitConv> class OutletCSV {
public:
    OutletCSV(const char *path, const UnitConv &uc, const std::vector<std::uint8_t> &outletMask, ADLat &adLat)
        : uc_(uc), ad_(adLat), cells_(fmx::maskCells(outletMask)), out_(path) {
        if (!out_) throw std::runtime_error("Could not open outlet csv");
        out_ << "step,time_s,phi_out_avg,phi_out_min,phi_out_max\n";
        if (cells_.size()) stats_ = static_cast<double *>(fmx::mallocBytes(3 * sizeof(double)));
    }
    ~OutletCSV() { fmx::freeBytes(stats_); }
    OutletCSV(const OutletCSV &) = delete;
    OutletCSV &operator=(const OutletCSV &) = delete;

    void log(int step) {
        double r[3] = {0.0, 0.0, 0.0};  // mean, min, max (0 for an empty mask, as the reference)
        if (cells_.size()) {
            fmx::check(fmx_reduce_outlet_mean(fmx::kPrec, ad_.d_phi_ptr(), cells_.data(), cells_.size(), stats_,
                                              nullptr),
                       "fmx_reduce_outlet_mean");
            fmx::memcpyDtoH(r, stats_, sizeof r);
        }
        out_ << step << "," << uc_.timeLatticeToPhys(step) << "," << r[0] << "," << r[1] << "," << r[2] << "\n";
    }
    void flush() { out_.flush(); }

private:
    const UnitConv &uc_;
    ADLat &ad_;
    fmx::DeviceIndexList cells_;
    double *stats_ = nullptr;
    std::ofstream out_;
};
