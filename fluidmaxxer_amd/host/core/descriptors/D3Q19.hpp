#pragma once
// D3Q19 velocity set (D3Q19.hpp:5-50 of the reference): rest, 6 axis, 12 edge directions.
// The device kernels carry the same table as compile-time constants; there is no upload step.
#include <array>

#include "core/Types.hpp"

struct D3Q19 {
    static constexpr int Q = 19;
    static constexpr DATA_TYPE cs2 = DATA_TYPE(1) / DATA_TYPE(3);
    static constexpr DATA_TYPE inv_cs2 = DATA_TYPE(3);
    static constexpr std::array<int, Q> cx = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    static constexpr std::array<int, Q> cy = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
    static constexpr std::array<int, Q> cz = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
    static constexpr std::array<DATA_TYPE, Q> w = {
        DATA_TYPE(1) / DATA_TYPE(3),  DATA_TYPE(1) / DATA_TYPE(18), DATA_TYPE(1) / DATA_TYPE(18),
        DATA_TYPE(1) / DATA_TYPE(18), DATA_TYPE(1) / DATA_TYPE(18), DATA_TYPE(1) / DATA_TYPE(18),
        DATA_TYPE(1) / DATA_TYPE(18), DATA_TYPE(1) / DATA_TYPE(36), DATA_TYPE(1) / DATA_TYPE(36),
        DATA_TYPE(1) / DATA_TYPE(36), DATA_TYPE(1) / DATA_TYPE(36), DATA_TYPE(1) / DATA_TYPE(36),
        DATA_TYPE(1) / DATA_TYPE(36), DATA_TYPE(1) / DATA_TYPE(36), DATA_TYPE(1) / DATA_TYPE(36),
        DATA_TYPE(1) / DATA_TYPE(36), DATA_TYPE(1) / DATA_TYPE(36), DATA_TYPE(1) / DATA_TYPE(36),
        DATA_TYPE(1) / DATA_TYPE(36)};
    static constexpr std::array<int, Q> opp = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
};

// kept for source compatibility with the reference's start-up sequence; nothing to upload
inline void uploadD3Q19Constants() {}
