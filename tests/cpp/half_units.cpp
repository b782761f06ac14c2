// Host build of the FP16 decode used by the 16-bit population storage (fluidmaxxer_amd/csrc/fmx_half.hpp):
// writes half_bits_to_float(h) for h = 0 .. 65535 as raw float32 to argv[1].
#include <cstdio>
#include <vector>

#include "fmx_half.hpp"

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: half_units <out.f32>\n");
        return 2;
    }
    std::vector<float> out(65536);
    for (unsigned h = 0; h < 65536u; ++h) out[h] = fmx::half_bits_to_float(static_cast<std::uint16_t>(h));
    std::FILE *fp = std::fopen(argv[1], "wb");
    if (!fp || std::fwrite(out.data(), sizeof(float), out.size(), fp) != out.size()) return 1;
    return std::fclose(fp) == 0 ? 0 : 1;
}
