// geometry.cpp -- host-side geometry inputs: seeded sphere packs, x-plane masks and
// ASCII VTK ImageData (.vti) read/write.
//
// The VTI reader follows the reference's readVTI contract (VTIReader.hpp:35-160):
// dimensions from the first `WholeExtent="..."` attribute (else the first
// `Extent="..."`), values from the first <DataArray> element parsed as numbers until
// the first non-number, mask = value > threshold; an array of 3*N values is read as
// vectors and thresholded by magnitude; any other count fills min(count, N) cells
// and leaves the rest 0.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "fmx_internal.hpp"

namespace fmx {
namespace {

// SplitMix64 (Steele, Lea & Flood 2014) seeds xoshiro256** (Blackman & Vigna 2018).
struct Rng {
    std::uint64_t s[4];
    static std::uint64_t splitmix(std::uint64_t &x) {
        std::uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(std::uint64_t seed) {
        for (auto &w : s) w = splitmix(seed);
    }
    static std::uint64_t rotl(std::uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    std::uint64_t next() {
        const std::uint64_t result = rotl(s[1] * 5, 7) * 9;
        const std::uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    // uniform in [0, 1) from the top 53 bits
    double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Sphere {
    int x, y, z;
    double r;
};

}  // namespace
}  // namespace fmx

using namespace fmx;

namespace {

// RFC 4648 base64 (whitespace ignored, stops at padding)
std::string base64_decode(const std::string &in) {
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+') return 62;
        if (c == '/') return 63;
        return -1;
    };
    std::string out;
    unsigned acc = 0;
    int bits = 0;
    for (char c : in) {
        if (c == '=') break;
        const int v = val(c);
        if (v < 0) continue;
        acc = (acc << 6) | unsigned(v);
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back(char((acc >> bits) & 0xFFu));
        }
    }
    return out;
}

// little-endian array of a VTK scalar type -> doubles
bool decode_values(const std::string &type, const char *p, size_t nbytes, std::vector<double> &vals) {
    auto run = [&](auto tag) {
        using T = decltype(tag);
        const size_t n = nbytes / sizeof(T);
        vals.resize(n);
        for (size_t i = 0; i < n; ++i) {
            T v;
            std::memcpy(&v, p + i * sizeof(T), sizeof(T));
            vals[i] = double(v);
        }
        return true;
    };
    if (type == "UInt8") return run(std::uint8_t{});
    if (type == "Int8") return run(std::int8_t{});
    if (type == "UInt16") return run(std::uint16_t{});
    if (type == "Int16") return run(std::int16_t{});
    if (type == "UInt32") return run(std::uint32_t{});
    if (type == "Int32") return run(std::int32_t{});
    if (type == "UInt64") return run(std::uint64_t{});
    if (type == "Int64") return run(std::int64_t{});
    if (type == "Float32") return run(float{});
    if (type == "Float64") return run(double{});
    return false;
}

}  // namespace

extern "C" {

int fmx_geom_sphere_pack(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, double r_min, double r_max,
                         double target_porosity, int32_t overlap, int32_t clearance, int64_t max_tries,
                         uint8_t *solid, uint8_t *sphere, double *porosity) {
    FMX_TRY_BEGIN
    if (nx < 1 || ny < 3 || nz < 3 || !solid || !(r_min > 0.0) || r_max < r_min || clearance < 0 ||
        2 * clearance >= nx || max_tries < 1)
        return set_error(FMX_ERR_ARG, "fmx_geom_sphere_pack: bad arguments");
    const i64 N = i64(nx) * ny * nz;
    auto idx = [&](int x, int y, int z) { return i64(x) + i64(nx) * (i64(y) + i64(ny) * z); };
    std::memset(solid, 0, size_t(N));
    if (sphere) std::memset(sphere, 0, size_t(N));
    i64 n_solid = 0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) {
            if (!(y == 0 || y == ny - 1 || z == 0 || z == nz - 1)) continue;
            for (int x = 0; x < nx; ++x) {
                solid[idx(x, y, z)] = 1;
                ++n_solid;
            }
        }
    const double target_solid = (1.0 - target_porosity) * double(N);
    Rng rng(seed);
    std::vector<Sphere> placed;
    // uniform grid of placed centres for the overlap test
    const double cell_sz = std::max(1.0, 2.0 * r_max);
    const int gx = std::max(1, int(std::ceil(nx / cell_sz))), gy = std::max(1, int(std::ceil(ny / cell_sz))),
              gz = std::max(1, int(std::ceil(nz / cell_sz)));
    std::vector<std::vector<int>> bins(size_t(gx) * gy * gz);
    auto bin_of = [&](int x, int y, int z) {
        const int bx = std::min(gx - 1, int(x / cell_sz)), by = std::min(gy - 1, int(y / cell_sz)),
                  bz = std::min(gz - 1, int(z / cell_sz));
        return (size_t(bz) * gy + by) * gx + bx;
    };
    i64 fails = 0;
    const int span_x = nx - 2 * clearance;
    while (double(n_solid) < target_solid && fails < max_tries) {
        const double r = r_min + (r_max - r_min) * rng.uniform();
        const int cx = clearance + std::min(span_x - 1, int(rng.uniform() * span_x));
        const int cy = 1 + std::min(ny - 3, int(rng.uniform() * (ny - 2)));
        const int cz = 1 + std::min(nz - 3, int(rng.uniform() * (nz - 2)));
        if (!overlap) {
            bool hit = false;
            const int bx = std::min(gx - 1, int(cx / cell_sz)), by = std::min(gy - 1, int(cy / cell_sz)),
                      bz = std::min(gz - 1, int(cz / cell_sz));
            for (int dz = -1; dz <= 1 && !hit; ++dz)
                for (int dy = -1; dy <= 1 && !hit; ++dy)
                    for (int dx = -1; dx <= 1 && !hit; ++dx) {
                        const int ax = bx + dx, ay = by + dy, az = bz + dz;
                        if (ax < 0 || ay < 0 || az < 0 || ax >= gx || ay >= gy || az >= gz) continue;
                        for (int id : bins[(size_t(az) * gy + ay) * gx + ax]) {
                            const Sphere &o = placed[size_t(id)];
                            const double ddx = o.x - cx, ddy = o.y - cy, ddz = o.z - cz;
                            const double rr = o.r + r;
                            if (ddx * ddx + ddy * ddy + ddz * ddz < rr * rr) {
                                hit = true;
                                break;
                            }
                        }
                    }
            if (hit) {
                ++fails;
                continue;
            }
        }
        fails = 0;
        bins[bin_of(cx, cy, cz)].push_back(int(placed.size()));
        placed.push_back(Sphere{cx, cy, cz, r});
        const int R = int(std::ceil(r));
        const double r2 = r * r;
        for (int z = std::max(0, cz - R); z <= std::min(nz - 1, cz + R); ++z)
            for (int y = std::max(0, cy - R); y <= std::min(ny - 1, cy + R); ++y)
                for (int x = std::max(0, cx - R); x <= std::min(nx - 1, cx + R); ++x) {
                    const double d2 = double(x - cx) * (x - cx) + double(y - cy) * (y - cy) + double(z - cz) * (z - cz);
                    if (d2 > r2) continue;
                    const i64 c = idx(x, y, z);
                    if (sphere) sphere[c] = 1;
                    if (!solid[c]) {
                        solid[c] = 1;
                        ++n_solid;
                    }
                }
    }
    if (porosity) *porosity = 1.0 - double(n_solid) / double(N);
    return FMX_OK;
    FMX_TRY_END
}

int fmx_build_x_plane_mask(int32_t nx, int32_t ny, int32_t nz, int32_t x_plane, uint8_t *mask) {
    if (nx < 1 || ny < 1 || nz < 1 || x_plane < 0 || x_plane >= nx || !mask)
        return set_error(FMX_ERR_ARG, "fmx_build_x_plane_mask: bad arguments");
    const i64 N = i64(nx) * ny * nz;
    std::memset(mask, 0, size_t(N));
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) mask[i64(x_plane) + i64(nx) * (y + i64(ny) * z)] = 1;
    return FMX_OK;
}

int fmx_vti_write_mask(const char *path, int32_t nx, int32_t ny, int32_t nz, const uint8_t *mask, const char *name) {
    FMX_TRY_BEGIN
    if (!path || !mask || nx < 1 || ny < 1 || nz < 1) return set_error(FMX_ERR_ARG, "fmx_vti_write_mask: bad arguments");
    std::FILE *fp = std::fopen(path, "wb");
    if (!fp) return set_error(FMX_ERR_ARG, "fmx_vti_write_mask: cannot open %s", path);
    const char *nm = name ? name : "solid";
    std::fprintf(fp,
                 "<?xml version=\"1.0\"?>\n"
                 "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
                 "  <ImageData WholeExtent=\"0 %d 0 %d 0 %d\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n"
                 "    <Piece Extent=\"0 %d 0 %d 0 %d\">\n"
                 "      <PointData Scalars=\"%s\">\n"
                 "        <DataArray type=\"UInt8\" Name=\"%s\" format=\"ascii\">\n",
                 nx - 1, ny - 1, nz - 1, nx - 1, ny - 1, nz - 1, nm, nm);
    const i64 N = i64(nx) * ny * nz;
    std::vector<char> line;
    line.reserve(size_t(2 * nx + 2));
    for (i64 row = 0; row < N / nx; ++row) {
        line.clear();
        for (int x = 0; x < nx; ++x) {
            line.push_back(mask[row * nx + x] ? '1' : '0');
            line.push_back(x + 1 < nx ? ' ' : '\n');
        }
        std::fwrite(line.data(), 1, line.size(), fp);
    }
    std::fprintf(fp,
                 "        </DataArray>\n"
                 "      </PointData>\n"
                 "    </Piece>\n"
                 "  </ImageData>\n"
                 "</VTKFile>\n");
    const bool bad = std::ferror(fp) != 0;
    std::fclose(fp);
    if (bad) return set_error(FMX_ERR_ARG, "fmx_vti_write_mask: write error on %s", path);
    return FMX_OK;
    FMX_TRY_END
}

int fmx_vti_read_mask(const char *path, double threshold, int32_t *dims, uint8_t *mask, int64_t capacity) {
    FMX_TRY_BEGIN
    if (!path || !dims) return set_error(FMX_ERR_ARG, "fmx_vti_read_mask: bad arguments");
    std::ifstream in(path, std::ios::binary);
    if (!in) return set_error(FMX_ERR_ARG, "[VTIReader]: Fatal, could not open File: %s", path);
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string content = ss.str();
    auto attr = [&](const char *nm) -> std::string {
        const std::string pat = std::string(nm) + "=\"";
        size_t p = content.find(pat);
        if (p == std::string::npos) return {};
        p += pat.size();
        const size_t e = content.find('"', p);
        if (e == std::string::npos) return {};
        return content.substr(p, e - p);
    };
    std::string ext = attr("WholeExtent");
    if (ext.empty()) ext = attr("Extent");
    if (ext.empty()) return set_error(FMX_ERR_ARG, "[VTIReader]: Fatal, no Extent/WholeExtent found in VTI File!");
    int e6[6];
    {
        std::istringstream is(ext);
        for (int &v : e6)
            if (!(is >> v)) return set_error(FMX_ERR_ARG, "[VTIReader]: Fatal, could not parse Extent/WholeExtent! - %s", ext.c_str());
    }
    dims[0] = e6[1] - e6[0] + 1;
    dims[1] = e6[3] - e6[2] + 1;
    dims[2] = e6[5] - e6[4] + 1;
    if (!mask) return FMX_OK;
    const i64 N = i64(dims[0]) * dims[1] * dims[2];
    if (N <= 0 || capacity < N) return set_error(FMX_ERR_ARG, "fmx_vti_read_mask: capacity %lld < %lld cells",
                                                 (long long)capacity, (long long)N);
    const size_t da = content.find("<DataArray");
    if (da == std::string::npos) return set_error(FMX_ERR_ARG, "[VTIReader]: Fatal, no <DataArray> found in VTI File.");
    const size_t gt = content.find('>', da);
    if (gt == std::string::npos) return set_error(FMX_ERR_ARG, "[VTIReader]: Fatal, invalid DataArray-Tag.");
    const std::string tag = content.substr(da, gt - da + 1);
    auto tag_attr = [&](const char *nm) -> std::string {
        const std::string pat = std::string(" ") + nm + "=\"";
        size_t q = tag.find(pat);
        if (q == std::string::npos) return {};
        q += pat.size();
        const size_t e = tag.find('"', q);
        return e == std::string::npos ? std::string() : tag.substr(q, e - q);
    };
    const std::string format = tag_attr("format");
    std::vector<double> vals;
    if (format == "appended" || format == "binary") {
        // raw appended or inline base64 binary arrays (extension: the reference reads ASCII only)
        if (!attr("compressor").empty())
            return set_error(FMX_ERR_UNSUPPORTED, "[VTIReader]: compressed binary arrays are not supported");
        const size_t hsize = attr("header_type") == "UInt64" ? 8 : 4;
        std::string blob;
        if (format == "appended") {
            const std::string off = tag_attr("offset");
            const size_t ad = content.find("<AppendedData");
            const size_t us = ad == std::string::npos ? ad : content.find('_', content.find('>', ad));
            if (off.empty() || us == std::string::npos)
                return set_error(FMX_ERR_ARG, "[VTIReader]: Fatal, appended DataArray without offset or AppendedData.");
            const std::string enc = content.substr(ad, content.find('>', ad) - ad);
            if (enc.find("encoding=\"raw\"") == std::string::npos)
                return set_error(FMX_ERR_UNSUPPORTED, "[VTIReader]: only raw-encoded AppendedData is supported");
            const size_t at = us + 1 + size_t(std::stoull(off));
            if (at + hsize > content.size()) return set_error(FMX_ERR_ARG, "[VTIReader]: Fatal, appended offset out of range.");
            blob = content.substr(at);
        } else {
            const size_t end = content.find("</DataArray>", gt);
            if (end == std::string::npos)
                return set_error(FMX_ERR_ARG, "[VTIReader]: Fatal, could not found closing </DataArray>-Tag!");
            blob = base64_decode(content.substr(gt + 1, end - gt - 1));
        }
        if (blob.size() < hsize) return set_error(FMX_ERR_ARG, "[VTIReader]: Fatal, truncated binary array.");
        std::uint64_t nbytes = 0;
        std::memcpy(&nbytes, blob.data(), hsize);
        if (hsize + nbytes > blob.size()) return set_error(FMX_ERR_ARG, "[VTIReader]: Fatal, truncated binary array.");
        if (!decode_values(tag_attr("type"), blob.data() + hsize, size_t(nbytes), vals))
            return set_error(FMX_ERR_UNSUPPORTED, "[VTIReader]: unsupported DataArray type '%s'", tag_attr("type").c_str());
    } else {
        const size_t end = content.find("</DataArray>", gt);
        if (end == std::string::npos)
            return set_error(FMX_ERR_ARG, "[VTIReader]: Fatal, could not found closing </DataArray>-Tag!");
        // parse numbers in content[gt+1, end) until the first non-number
        std::string body = content.substr(gt + 1, end - gt - 1);
        vals.reserve(size_t(N));
        const char *p = body.c_str();
        for (;;) {
            char *q = nullptr;
            errno = 0;
            const double v = std::strtod(p, &q);
            if (q == p) break;
            vals.push_back(v);
            p = q;
        }
    }
    if (vals.empty()) return set_error(FMX_ERR_ARG, "[VTIReader]: Fatal, DataArray does not contain numerical values!");
    std::memset(mask, 0, size_t(N));
    const i64 nv = i64(vals.size());
    if (nv == 3 * N && nv != N) {
        for (i64 i = 0; i < N; ++i) {
            const double a = vals[3 * i], b = vals[3 * i + 1], c = vals[3 * i + 2];
            mask[i] = std::sqrt(a * a + b * b + c * c) > threshold ? 1 : 0;
        }
    } else {
        const i64 m = std::min(nv, N);
        for (i64 i = 0; i < m; ++i) mask[i] = vals[size_t(i)] > threshold ? 1 : 0;
        if (nv != N)
            std::fprintf(stderr,
                         "[VTIReader]: Warning: Number of values (%lld) does not fit to Nx*Ny*Nz (%lld) or 3*Nx*Ny*Nz.\n",
                         (long long)nv, (long long)N);
    }
    return FMX_OK;
    FMX_TRY_END
}

}  // extern "C"
