#pragma once
// Outlet breakthrough log (OutletCSV.hpp:12-67 of the reference): header
// "step,time_s,phi_out_avg,phi_out_min,phi_out_max", one row per log(step) with the FP64
// mean/min/max of the AD lattice's phi over the outlet mask.
#include <algorithm>
#include <cstdint>
#include <fstream>
#include <limits>
#include <stdexcept>
#include <vector>

template <typename ADLat, typename UnitConv> class OutletCSV {
public:
    OutletCSV(const char *path, const UnitConv &uc, const std::vector<std::uint8_t> &outletMask, ADLat &adLat)
        : uc_(uc), outletMask_(outletMask), ad_(adLat), out_(path) {
        if (!out_) throw std::runtime_error("Could not open outlet csv");
        out_ << "step,time_s,phi_out_avg,phi_out_min,phi_out_max\n";
    }
    void log(int step) {
        ad_.downloadPhi();
        const auto &phi = ad_.phiHost();
        double sum = 0.0, mn = std::numeric_limits<double>::infinity(), mx = -std::numeric_limits<double>::infinity();
        std::size_t cnt = 0;
        for (std::size_t i = 0; i < outletMask_.size(); ++i) {
            if (!outletMask_[i]) continue;
            const double v = static_cast<double>(phi[i]);
            sum += v;
            mn = std::min(mn, v);
            mx = std::max(mx, v);
            ++cnt;
        }
        const double avg = cnt ? sum / double(cnt) : 0.0;
        if (!cnt) mn = mx = 0.0;
        out_ << step << "," << uc_.timeLatticeToPhys(step) << "," << avg << "," << mn << "," << mx << "\n";
    }
    void flush() { out_.flush(); }

private:
    const UnitConv &uc_;
    const std::vector<std::uint8_t> &outletMask_;
    ADLat &ad_;
    std::ofstream out_;
};
