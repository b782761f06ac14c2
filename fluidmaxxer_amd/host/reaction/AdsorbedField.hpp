#pragma once
// Adsorbed load q on the device with a host mirror (AdsorbedField.hpp:7-60 of the reference).
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "core/DeviceHelpers.hpp"
#include "core/Params.hpp"

class AdsorbedField {
public:
    explicit AdsorbedField(const LBMParams &p) : p_(p) {
        n_ = std::int64_t(p_.Nx) * p_.Ny * p_.Nz;
        d_q_ = static_cast<DATA_TYPE *>(fmx::mallocBytes(sizeof(DATA_TYPE) * std::size_t(n_)));
        fmx::memsetBytes(d_q_, 0, sizeof(DATA_TYPE) * std::size_t(n_));
        h_q_.assign(std::size_t(n_), DATA_TYPE(0));
    }
    ~AdsorbedField() { fmx::freeBytes(d_q_); }
    AdsorbedField(const AdsorbedField &) = delete;
    AdsorbedField &operator=(const AdsorbedField &) = delete;

    DATA_TYPE *d_q() { return d_q_; }
    const DATA_TYPE *d_q() const { return d_q_; }
    void download() { fmx::memcpyDtoH(h_q_.data(), d_q_, sizeof(DATA_TYPE) * std::size_t(n_)); }
    void upload() { fmx::memcpyHtoD(d_q_, h_q_.data(), sizeof(DATA_TYPE) * std::size_t(n_)); }
    std::vector<DATA_TYPE> &host() { return h_q_; }
    const std::vector<DATA_TYPE> &host() const { return h_q_; }
    std::int64_t size64() const { return n_; }

private:
    LBMParams p_;
    std::int64_t n_ = 0;
    DATA_TYPE *d_q_ = nullptr;
    std::vector<DATA_TYPE> h_q_;
};
