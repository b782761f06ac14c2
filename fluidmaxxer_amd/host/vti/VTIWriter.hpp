#pragma once
// VTIWriterPhys<FluidDescriptor> (VTIWriter.hpp:21-334 of the reference): VTK ImageData output of
// registered per-cell fields in physical units, one file per written step, optionally indexed in a
// ParaView collection (.pvd) with the physical time of each step.
//
// Same registration API and the same ASCII file layout (x-fastest point data, Spacing = dx,
// Float32/Float64 by DATA_TYPE, 7 / 15 fixed digits).  Addition: setBinary(true) writes the arrays as
// raw appended little-endian data (header_type UInt64) instead of ASCII, for grids where text output
// is too large; fmx_vti_read_mask reads both forms.
//
// writeImageData() is the lattice-independent writer underneath (usable without a device).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "core/Types.hpp"
#include "core/io/UnitConverter.hpp"
#include "core/lattice/ADLattice.hpp"
#include "core/lattice/NSLattice.hpp"

namespace fmx_vti {

struct Field {
    std::string name;
    int components = 1;
    std::function<DATA_TYPE(std::int64_t cell, int component)> value;
};

inline const char *dataTypeName() { return std::is_same_v<DATA_TYPE, float> ? "Float32" : "Float64"; }

// One ImageData file; cells are x-fastest, cell = x + Nx*(y + Ny*z).
inline void writeImageData(const std::string &path, int Nx, int Ny, int Nz, double spacing,
                           const std::vector<Field> &fields, bool binary) {
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("VTIWriter: could not open " + path);
    const double sz = Nz > 1 ? spacing : 1.0;
    const std::int64_t n = std::int64_t(Nx) * Ny * Nz;
    if constexpr (std::is_same_v<DATA_TYPE, float>) out << std::setprecision(7) << std::fixed;
    else out << std::setprecision(15) << std::fixed;
    out << "<?xml version=\"1.0\"?>\n<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\"";
    if (binary) out << " header_type=\"UInt64\"";
    out << ">\n  <ImageData WholeExtent=\"0 " << (Nx - 1) << " 0 " << (Ny - 1) << " 0 " << (Nz - 1)
        << "\" Origin=\"0 0 0\" Spacing=\"" << spacing << " " << spacing << " " << sz << "\">\n"
        << "    <Piece Extent=\"0 " << (Nx - 1) << " 0 " << (Ny - 1) << " 0 " << (Nz - 1) << "\">\n"
        << "      <PointData";
    if (!fields.empty()) out << " Scalars=\"" << fields.front().name << "\"";
    out << ">\n";
    std::uint64_t offset = 0;
    for (const Field &f : fields) {
        out << "        <DataArray type=\"" << dataTypeName() << "\" Name=\"" << f.name << "\"";
        if (f.components > 1) out << " NumberOfComponents=\"" << f.components << "\"";
        if (binary) {
            out << " format=\"appended\" offset=\"" << offset << "\"/>\n";
            offset += sizeof(std::uint64_t) + std::uint64_t(n) * f.components * sizeof(DATA_TYPE);
            continue;
        }
        out << " format=\"ascii\">\n          ";
        for (std::int64_t c = 0; c < n; ++c)
            for (int k = 0; k < f.components; ++k) out << f.value(c, k) << " ";
        out << "\n        </DataArray>\n";
    }
    out << "      </PointData>\n      <CellData>\n      </CellData>\n    </Piece>\n  </ImageData>\n";
    if (binary) {
        out << "  <AppendedData encoding=\"raw\">\n   _";
        std::vector<DATA_TYPE> buf;
        for (const Field &f : fields) {
            const std::uint64_t bytes = std::uint64_t(n) * f.components * sizeof(DATA_TYPE);
            out.write(reinterpret_cast<const char *>(&bytes), sizeof(bytes));
            buf.resize(std::size_t(n) * f.components);
            for (std::int64_t c = 0; c < n; ++c)
                for (int k = 0; k < f.components; ++k) buf[std::size_t(c) * f.components + k] = f.value(c, k);
            out.write(reinterpret_cast<const char *>(buf.data()), std::streamsize(bytes));
        }
        out << "\n  </AppendedData>\n";
    }
    out << "</VTKFile>\n";
    if (!out) throw std::runtime_error("VTIWriter: write failed for " + path);
}

// Adds a DataSet entry (file name relative to the .pvd) for `time`, creating the collection if needed.
inline void appendToPVD(const std::string &pvdFilename, const std::string &vtiFilename, double time) {
    const std::string file = std::filesystem::path(vtiFilename).filename().string();
    std::ostringstream entry;
    entry << "    <DataSet timestep=\"" << time << "\" group=\"\" part=\"0\" file=\"" << file << "\"/>\n";
    std::ifstream in(pvdFilename);
    std::string content;
    if (in) {
        std::stringstream buf;
        buf << in.rdbuf();
        content = buf.str();
    }
    const std::string tag = "</Collection>";
    const auto pos = content.rfind(tag);
    if (pos == std::string::npos) {
        content = "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
                  "  <Collection>\n" + entry.str() + "  </Collection>\n</VTKFile>\n";
    } else {
        content.insert(pos, entry.str());
    }
    std::ofstream out(pvdFilename);
    if (!out) throw std::runtime_error("VTIWriter: could not open " + pvdFilename);
    out << content;
}

}  // namespace fmx_vti

template <typename FluidDescriptor> class VTIWriterPhys {
public:
    explicit VTIWriterPhys(const UnitConverter &uc) : uc_(uc) {}

    void setOutputPattern(std::string prefix, int width = 6) {
        filePrefix_ = std::move(prefix);
        fileNumberWidth_ = width;
    }
    void setPVDFile(std::string pvdFilename) {
        pvdFile_ = std::move(pvdFilename);
        usePVD_ = true;
    }
    // raw appended binary arrays instead of ASCII
    void setBinary(bool on) { binary_ = on; }

    void addPreWriteHook(std::function<void()> hook) { preWriteHooks_.push_back(std::move(hook)); }

    void registerScalar(const std::string &name, std::function<DATA_TYPE(int)> getter) {
        fields_.push_back({name, 1, [getter](std::int64_t c, int) { return getter(int(c)); }});
    }
    void registerScalarField(const std::string &name, const std::vector<DATA_TYPE> &field) {
        if (field.empty()) throw std::runtime_error("registerScalarField: empty field");
        registerScalar(name, [&field](int c) -> DATA_TYPE { return field[std::size_t(c)]; });
    }
    void registerVector3(const std::string &name, std::function<void(int, DATA_TYPE &, DATA_TYPE &, DATA_TYPE &)> getter) {
        fields_.push_back({name, 3, [getter](std::int64_t c, int k) {
                               DATA_TYPE v[3] = {DATA_TYPE(0), DATA_TYPE(0), DATA_TYPE(0)};
                               getter(int(c), v[0], v[1], v[2]);
                               return v[k];
                           }});
    }

    void registerRhoPhys(const Lattice<FluidDescriptor> &lat, const std::string &name = "rho_phys") {
        addPreWriteHook([&lat]() { const_cast<Lattice<FluidDescriptor> &>(lat).downloadMacroscopic(); });
        registerScalar(name, [&lat, this](int c) -> DATA_TYPE {
            return static_cast<DATA_TYPE>(uc_.densityLatticeToPhys(static_cast<double>(lat.rhoHost()[c])));
        });
    }
    void registerVelocityPhys(const Lattice<FluidDescriptor> &lat, const std::string &name = "velocity_phys") {
        addPreWriteHook([&lat]() { const_cast<Lattice<FluidDescriptor> &>(lat).downloadMacroscopic(); });
        registerVector3(name, [&lat, this](int c, DATA_TYPE &x, DATA_TYPE &y, DATA_TYPE &z) {
            x = static_cast<DATA_TYPE>(uc_.velocityLatticeToPhys(static_cast<double>(lat.uxHost()[c])));
            y = static_cast<DATA_TYPE>(uc_.velocityLatticeToPhys(static_cast<double>(lat.uyHost()[c])));
            z = static_cast<DATA_TYPE>(uc_.velocityLatticeToPhys(static_cast<double>(lat.uzHost()[c])));
        });
    }
    void registerSpeedMagPhys(const Lattice<FluidDescriptor> &lat, const std::string &name = "u_mag_phys") {
        addPreWriteHook([&lat]() { const_cast<Lattice<FluidDescriptor> &>(lat).downloadMacroscopic(); });
        registerScalar(name, [&lat, this](int c) -> DATA_TYPE {
            const double ux = uc_.velocityLatticeToPhys(static_cast<double>(lat.uxHost()[c]));
            const double uy = uc_.velocityLatticeToPhys(static_cast<double>(lat.uyHost()[c]));
            const double uz = uc_.velocityLatticeToPhys(static_cast<double>(lat.uzHost()[c]));
            return static_cast<DATA_TYPE>(std::sqrt(ux * ux + uy * uy + uz * uz));
        });
    }
    void registerObstacle(const Lattice<FluidDescriptor> &lat, const std::string &name = "obstacle") {
        addPreWriteHook([&lat]() { const_cast<Lattice<FluidDescriptor> &>(lat).downloadObstacle(); });
        registerScalar(name,
                       [&lat](int c) -> DATA_TYPE { return static_cast<DATA_TYPE>(static_cast<int>(lat.obstacleHost()[c])); });
    }
    template <typename ADDescriptor>
    void registerConcentration(const ADLattice<ADDescriptor> &ad, const std::string &name = "concentration") {
        addPreWriteHook([&ad]() { const_cast<ADLattice<ADDescriptor> &>(ad).downloadPhi(); });
        registerScalar(name, [&ad](int c) -> DATA_TYPE { return ad.phiHost()[std::size_t(c)]; });
    }

    void writeStep(const Lattice<FluidDescriptor> &lat, double physTime, int stepIndex) {
        for (auto &hook : preWriteHooks_) hook();
        const std::string vti = makeStepFilename(stepIndex);
        std::error_code ec;
        const std::filesystem::path p(vti);
        if (p.has_parent_path()) std::filesystem::create_directories(p.parent_path(), ec);
        if (usePVD_ && std::filesystem::path(pvdFile_).has_parent_path())
            std::filesystem::create_directories(std::filesystem::path(pvdFile_).parent_path(), ec);
        const auto &prm = lat.P();
        fmx_vti::writeImageData(vti, prm.Nx, prm.Ny, prm.Nz, uc_.dx_phys, fields_, binary_);
        if (usePVD_) fmx_vti::appendToPVD(pvdFile_, vti, physTime);
        std::printf("[VTIWriter]: Wrote VTI: %s\n", vti.c_str());
    }

private:
    std::string makeStepFilename(int stepIndex) const {
        std::ostringstream ss;
        ss << filePrefix_ << std::setw(fileNumberWidth_) << std::setfill('0') << stepIndex << ".vti";
        return ss.str();
    }

    const UnitConverter &uc_;
    std::vector<fmx_vti::Field> fields_;
    std::vector<std::function<void()>> preWriteHooks_;
    bool usePVD_ = false, binary_ = false;
    std::string pvdFile_;
    std::string filePrefix_ = "out_";
    int fileNumberWidth_ = 6;
};
