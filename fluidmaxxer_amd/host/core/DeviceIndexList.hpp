#pragma once
// A device-resident list of 64-bit cell indices (compacted boundary masks).
#include <cstdint>
#include <vector>

#include "core/DeviceHelpers.hpp"

namespace fmx {

class DeviceIndexList {
public:
    DeviceIndexList() = default;
    explicit DeviceIndexList(const std::vector<std::int64_t> &h) { assign(h); }
    ~DeviceIndexList() { freeBytes(d_); }
    DeviceIndexList(const DeviceIndexList &) = delete;
    DeviceIndexList &operator=(const DeviceIndexList &) = delete;

    void assign(const std::vector<std::int64_t> &h) {
        freeBytes(d_);
        d_ = nullptr;
        n_ = std::int64_t(h.size());
        if (n_) {
            d_ = static_cast<std::int64_t *>(mallocBytes(h.size() * sizeof(std::int64_t)));
            memcpyHtoD(d_, h.data(), h.size() * sizeof(std::int64_t));
        }
    }
    const std::int64_t *data() const { return d_; }
    std::int64_t size() const { return n_; }

private:
    std::int64_t *d_ = nullptr;
    std::int64_t n_ = 0;
};

// cells of a mask, in cell order
inline std::vector<std::int64_t> maskCells(const std::vector<std::uint8_t> &mask) {
    std::vector<std::int64_t> out;
    for (std::size_t i = 0; i < mask.size(); ++i)
        if (mask[i]) out.push_back(std::int64_t(i));
    return out;
}

// true when the mask is exactly the plane x = xPlane
inline bool isFullXPlane(const std::vector<std::uint8_t> &mask, int Nx, int Ny, int Nz, int xPlane) {
    if (mask.size() != std::size_t(Nx) * Ny * Nz) return false;
    for (std::size_t i = 0; i < mask.size(); ++i)
        if ((mask[i] != 0) != (int(i % std::size_t(Nx)) == xPlane)) return false;
    return true;
}

}  // namespace fmx
