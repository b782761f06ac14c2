#pragma once
// D3Q7 velocity set (D3Q7.hpp:5-28 of the reference).
#include <array>

#include "core/Types.hpp"

struct D3Q7 {
    static constexpr int Q = 7;
    static constexpr DATA_TYPE cs2 = DATA_TYPE(1) / DATA_TYPE(4);
    static constexpr DATA_TYPE inv_cs2 = DATA_TYPE(4);
    static constexpr std::array<int, Q> cx = {0, 1, -1, 0, 0, 0, 0};
    static constexpr std::array<int, Q> cy = {0, 0, 0, 1, -1, 0, 0};
    static constexpr std::array<int, Q> cz = {0, 0, 0, 0, 0, 1, -1};
    static constexpr std::array<DATA_TYPE, Q> w = {DATA_TYPE(1) / DATA_TYPE(4), DATA_TYPE(1) / DATA_TYPE(8),
                                                   DATA_TYPE(1) / DATA_TYPE(8), DATA_TYPE(1) / DATA_TYPE(8),
                                                   DATA_TYPE(1) / DATA_TYPE(8), DATA_TYPE(1) / DATA_TYPE(8),
                                                   DATA_TYPE(1) / DATA_TYPE(8)};
    static constexpr std::array<int, Q> opp = {0, 2, 1, 4, 3, 6, 5};
};

inline void uploadD3Q7Constants() {}
