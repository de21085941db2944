// lbm_io.hip — the reference's output files, written from host arrays (no device work).
//
//   lbm_write_vti       LBM::save_vtk (lbm.cuh:269-343): VTK XML ImageData, version 1.0,
//                       UInt64 headers, appended raw Float32 "Density" and 3-component
//                       "Velocity" (2-D velocities padded with u_z = 0), file
//                       <dir>/vtk/sim_data_%06d.vti.
//   lbm_write_macros    LBM::save_macroscopics (lbm.cuh:173-204): raw Float32
//                       <dir>/density/density_<t>.bin and <dir>/velocity/velocity_<t>.bin
//                       (velocity AoS, the "aos" layout of graphics/visualize.py).
//   lbm_write_midplane  LBM::save_midplane_slice (lbm.cuh:206-267): the z = NZ/2 plane of a
//                       3-D field, <dir>/midplane_density/density_<t>.bin and
//                       <dir>/midplane_velocity/velocity_<t>.bin ([NY][NX][3], the layout
//                       graphics/plot_midplane.py reads).
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

#include "../../include/lbm_hip.h"

namespace {

int make_dir(const std::string& d) {
    std::error_code ec;
    std::filesystem::create_directories(d, ec);
    return ec ? LBM_E_INVALID : LBM_OK;
}

struct File {
    FILE* f = nullptr;
    explicit File(const std::string& path) : f(std::fopen(path.c_str(), "wb")) {}
    ~File() {
        if (f) std::fclose(f);
    }
    bool put(const void* p, size_t n) { return n == 0 || std::fwrite(p, 1, n, f) == n; }
    bool str(const std::string& s) { return put(s.data(), s.size()); }
};

bool valid_shape(int nx, int ny, int nz) { return nx > 0 && ny > 0 && nz > 0 && (long long)nx * ny * nz < (1LL << 40); }

}  // namespace

extern "C" {

int lbm_write_vti(const char* dir, int64_t t, int32_t nx, int32_t ny, int32_t nz, int32_t dim, const float* rho,
                  const float* u, uint32_t flags) {
    if (!dir || !rho || !u || !valid_shape(nx, ny, nz) || t < 0 || (dim != 2 && dim != 3) || (dim == 2 && nz != 1))
        return LBM_E_INVALID;
    const std::string out_dir = std::string(dir) + "/vtk";
    if (make_dir(out_dir)) return LBM_E_INVALID;
    char name[64];
    std::snprintf(name, sizeof(name), "/sim_data_%06lld.vti", (long long)t);
    File file(out_dir + name);
    if (!file.f) return LBM_E_INVALID;
    const size_t n = size_t(nx) * size_t(ny) * size_t(nz);
    const uint64_t rho_bytes = n * sizeof(float), vel_bytes = 3 * n * sizeof(float);
    // The appended block holds [UInt64 length][Density][UInt64 length][Velocity]; the
    // Velocity array therefore starts 8 + rho_bytes after the '_'. The reference declares
    // rho_bytes (LBM_VTI_REFERENCE_OFFSET reproduces that).
    const uint64_t vel_offset = (flags & LBM_VTI_REFERENCE_OFFSET) ? rho_bytes : rho_bytes + sizeof(uint64_t);
    const std::string ext = "0 " + std::to_string(nx - 1) + " 0 " + std::to_string(ny - 1) + " 0 " + std::to_string(nz - 1);
    std::string head;
    head += "<?xml version=\"1.0\"?>\n";
    head += "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
    head += "  <ImageData WholeExtent=\"" + ext + "\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n";
    head += "    <Piece Extent=\"" + ext + "\">\n";
    head += "      <PointData Scalars=\"Density\" Vectors=\"Velocity\">\n";
    head += "        <DataArray type=\"Float32\" Name=\"Density\" format=\"appended\" offset=\"0\"/>\n";
    head += "        <DataArray type=\"Float32\" Name=\"Velocity\" NumberOfComponents=\"3\" format=\"appended\" offset=\"" +
            std::to_string(vel_offset) + "\"/>\n";
    head += "      </PointData>\n      <CellData>\n      </CellData>\n    </Piece>\n  </ImageData>\n";
    head += "  <AppendedData encoding=\"raw\">\n   _";
    bool ok = file.str(head) && file.put(&rho_bytes, 8) && file.put(rho, rho_bytes) && file.put(&vel_bytes, 8);
    // velocity as [node][3] in chunks (2-D: u_z = 0)
    std::vector<float> buf;
    const size_t chunk = 1 << 16;
    for (size_t i0 = 0; ok && i0 < n; i0 += chunk) {
        const size_t m = n - i0 < chunk ? n - i0 : chunk;
        buf.assign(3 * m, 0.0f);
        for (size_t i = 0; i < m; i++)
            for (int c = 0; c < dim; c++) buf[3 * i + size_t(c)] = u[size_t(dim) * (i0 + i) + size_t(c)];
        ok = file.put(buf.data(), buf.size() * sizeof(float));
    }
    ok = ok && file.str("\n  </AppendedData>\n</VTKFile>\n");
    return ok ? LBM_OK : LBM_E_INVALID;
}

int lbm_write_macros(const char* dir, int64_t t, int64_t n, int32_t dim, const float* rho, const float* u) {
    if (!dir || !rho || !u || n <= 0 || t < 0 || (dim != 2 && dim != 3)) return LBM_E_INVALID;
    const std::string d(dir);
    if (make_dir(d + "/density") || make_dir(d + "/velocity")) return LBM_E_INVALID;
    File fr(d + "/density/density_" + std::to_string(t) + ".bin");
    File fv(d + "/velocity/velocity_" + std::to_string(t) + ".bin");
    if (!fr.f || !fv.f) return LBM_E_INVALID;
    const bool ok = fr.put(rho, size_t(n) * sizeof(float)) && fv.put(u, size_t(n) * size_t(dim) * sizeof(float));
    return ok ? LBM_OK : LBM_E_INVALID;
}

int lbm_write_midplane(const char* dir, int64_t t, int32_t nx, int32_t ny, int32_t nz, const float* rho, const float* u) {
    if (!dir || !rho || !u || !valid_shape(nx, ny, nz) || t < 0) return LBM_E_INVALID;
    const std::string d(dir);
    if (make_dir(d + "/midplane_density") || make_dir(d + "/midplane_velocity")) return LBM_E_INVALID;
    File fr(d + "/midplane_density/density_" + std::to_string(t) + ".bin");
    File fv(d + "/midplane_velocity/velocity_" + std::to_string(t) + ".bin");
    if (!fr.f || !fv.f) return LBM_E_INVALID;
    const size_t plane = size_t(nx) * size_t(ny), z0 = size_t(nz / 2) * plane;
    // the plane is contiguous in both layouts ([z][y][x] and [z][y][x][3])
    const bool ok = fr.put(rho + z0, plane * sizeof(float)) && fv.put(u + 3 * z0, 3 * plane * sizeof(float));
    return ok ? LBM_OK : LBM_E_INVALID;
}

}  // extern "C"
