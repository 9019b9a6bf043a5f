// output.hip -- fields, monitors and files of the reference drivers' output.
//
//   velocity field     calcula per site (SimBoltz_Functions/LBM_functions.cpp:
//                      3517-3589) as rec_velocity writes it (:4218-4288)
//   vorticity field    vorticity (:3761-3872) as rec_vorticity writes it
//                      (:4302-4346)
//   enstrophy          calc_enstrophy (Other_functions.cpp:119-142)
//   VTK files          rec_velocity / rec_vorticity, the map writers of
//                      src/main_stability-map_taylor-green.cpp:181-200 and
//                      src/main_stability-map_obstacle.cpp:188-207, and the
//                      reader of tools/error_calculator.cpp:125-168
//   L2 errors          tools/error_calculator.cpp:35-114
//
// Fields are computed on the GPU from the populations in HBM (bitwise the
// reference's values: same moment sums, same difference quotients) and
// copied to the host once; the text formatting is the C++ stream default the
// reference uses (six significant digits, "%g").
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "collide.hpp"
#include "common.hpp"
#include "lattice_impl.hpp"

using namespace erlbm;
using namespace erlbm::lat;

namespace {

// u of every fluid site, written at its voxel (solid voxels keep 0).
template <int Q>
__global__ __launch_bounds__(BLOCK) void k_velocity(const double *__restrict__ aos, int n_fluid,
                                                    const int *__restrict__ vox_of_fluid, double *__restrict__ vel)
{
    const int p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= n_fluid) return;
    double f[Q];
    static_for<Q>([&](auto I) { f[(int)I] = aos[(long long)p * Q + (int)I]; });
    double rho, vx, vy, vz;
    moments_exact<Q>(f, rho, vx, vy, vz);
    const long long v = vox_of_fluid ? vox_of_fluid[p] : p;
    vel[3 * v + 0] = vx;
    vel[3 * v + 1] = vy;
    vel[3 * v + 2] = vz;
}

// vorticity() at every voxel: central differences of the neighbours'
// velocity with periodic wrap, a solid neighbour counting as u = 0.
// zero_solids: write 0 at solid voxels (rec_vorticity); otherwise the curl
// there too (calc_enstrophy visits every voxel).  enstrophy_term: write
// 0.5 |w|^2 to out[v] instead of the vector.
__global__ __launch_bounds__(BLOCK) void k_vorticity(const double *__restrict__ vel, const uint32_t *__restrict__ nbr,
                                                     int nx, int ny, int nz, int zero_solids, int enstrophy_term,
                                                     double *__restrict__ out)
{
    const long long v = (long long)blockIdx.x * BLOCK + threadIdx.x;
    const long long n = (long long)nx * ny * nz;
    if (v >= n) return;
    const int x = (int)(v % nx);
    const int y = (int)((v / nx) % ny);
    const int z = (int)(v / ((long long)nx * ny));
    const bool solid = !(nbr[v] & FLUID_BIT);
    double wx = 0.0, wy = 0.0, wz = 0.0;
    if (!(solid && zero_solids)) {
        auto at = [&](int xx, int yy, int zz) { return 3 * (xx + (long long)nx * (yy + (long long)ny * zz)); };
        const long long xp = at((x + 1) % nx, y, z), xm = at((x - 1 + nx) % nx, y, z);
        const long long yp = at(x, (y + 1) % ny, z), ym = at(x, (y - 1 + ny) % ny, z);
        const long long zp = at(x, y, (z + 1) % nz), zm = at(x, y, (z - 1 + nz) % nz);
        const double d_vz_y = (vel[yp + 2] - vel[ym + 2]) / 2.0;
        const double d_vy_z = (vel[zp + 1] - vel[zm + 1]) / 2.0;
        wx = d_vz_y - d_vy_z;
        const double d_vx_z = (vel[zp + 0] - vel[zm + 0]) / 2.0;
        const double d_vz_x = (vel[xp + 2] - vel[xm + 2]) / 2.0;
        wy = d_vx_z - d_vz_x;
        const double d_vy_x = (vel[xp + 1] - vel[xm + 1]) / 2.0;
        const double d_vx_y = (vel[yp + 0] - vel[ym + 0]) / 2.0;
        wz = d_vy_x - d_vx_y;
    }
    if (enstrophy_term) {
        out[v] = 0.5 * (wx * wx + wy * wy + wz * wz);
    } else {
        out[3 * v + 0] = wx;
        out[3 * v + 1] = wy;
        out[3 * v + 2] = wz;
    }
}

// calc_enstrophy's sum: one running FP64 sum from 0 with x outermost, then
// y, then z (the reference's loop order).
__global__ void k_sum_x_outer(const double *__restrict__ term, int nx, int ny, int nz, double *out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++)
            for (int z = 0; z < nz; z++) s = s + term[x + (long long)nx * (y + (long long)ny * z)];
    *out = s;
}

// u of every voxel on the device (3 nvox doubles, caller frees).
int velocity_device(erlbm_lattice *h, int which, double **vel)
{
    *vel = nullptr;
    double *aos = nullptr;
    void *owned = nullptr;
    int rc = to_aos_device(h, which, &aos, &owned);
    if (rc) return rc;
    double *d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(double) * 3 * (size_t)h->nvox);
    if (e == hipSuccess) e = hipMemsetAsync(d, 0, sizeof(double) * 3 * (size_t)h->nvox, h->stream);
    if (e == hipSuccess) {
        const int nb = mon_blocks(h);
        if (h->q == 9)
            hipLaunchKernelGGL((k_velocity<9>), dim3(nb), dim3(BLOCK), 0, h->stream, aos, h->n_fluid,
                               h->vox_of_fluid, d);
        else
            hipLaunchKernelGGL((k_velocity<19>), dim3(nb), dim3(BLOCK), 0, h->stream, aos, h->n_fluid,
                               h->vox_of_fluid, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(owned);
    if (e != hipSuccess) {
        (void)hipFree(d);
        set_last_error("velocity field", e);
        return ERLBM_ERR_HIP;
    }
    *vel = d;
    return ERLBM_OK;
}

int vorticity_device(erlbm_lattice *h, int which, bool enstrophy_terms, double **out)
{
    *out = nullptr;
    double *vel = nullptr;
    int rc = velocity_device(h, which, &vel);
    if (rc) return rc;
    double *d = nullptr;
    const size_t n = (enstrophy_terms ? 1 : 3) * (size_t)h->nvox;
    hipError_t e = hipMalloc(&d, sizeof(double) * n);
    if (e == hipSuccess) {
        const unsigned nb = (unsigned)((h->nvox + BLOCK - 1) / BLOCK);
        hipLaunchKernelGGL(k_vorticity, dim3(nb), dim3(BLOCK), 0, h->stream, vel, h->nbr, h->nx, h->ny, h->nz,
                           enstrophy_terms ? 0 : 1, enstrophy_terms ? 1 : 0, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(vel);
    if (e != hipSuccess) {
        (void)hipFree(d);
        set_last_error("vorticity field", e);
        return ERLBM_ERR_HIP;
    }
    *out = d;
    return ERLBM_OK;
}

int to_host(erlbm_lattice *h, double *d, double *host, size_t n)
{
    hipError_t e = hipMemcpy(host, d, sizeof(double) * n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) {
        set_last_error("field copy", e);
        return ERLBM_ERR_HIP;
    }
    (void)h;
    return ERLBM_OK;
}

// ---- text --------------------------------------------------------------------------
// The C++ stream default for doubles: "%g" with precision 6.
inline void put_g(std::string &s, double v)
{
    char b[40];
    const int n = std::snprintf(b, sizeof b, "%g", v);
    s.append(b, (size_t)n);
}

int write_vectors(const char *path, const char *title, const char *name, int nx, int ny, int nz, const double *v)
{
    if (!path || !v || nx < 1 || ny < 1 || nz < 1) return ERLBM_ERR_ARG;
    std::FILE *fp = std::fopen(path, "wb");
    if (!fp) {
        set_last_error_msg("cannot open the output file");
        return ERLBM_ERR_ARG;
    }
    std::string s;
    s.reserve(1 << 20);
    s += "# vtk DataFile Version 2.0\n";
    s += title;
    s += "\nASCII\nDATASET STRUCTURED_POINTS\nDIMENSIONS " + std::to_string(nx) + " " + std::to_string(ny) + " " +
         std::to_string(nz) + "\nASPECT_RATIO 1 1 1\nORIGIN 0 0 0\nPOINT_DATA " + std::to_string(nx * ny * nz) +
         "\nVECTORS " + name + " double\n";
    bool ok = true;
    for (int z = 0; z < nz && ok; z++)
        for (int y = 0; y < ny && ok; y++) {
            for (int x = 0; x < nx; x++) {
                const double *p = v + 3 * (x + (long long)nx * (y + (long long)ny * z));
                put_g(s, p[0]);
                s += ' ';
                put_g(s, p[1]);
                s += ' ';
                put_g(s, p[2]);
                s += ' ';
            }
            s += '\n';
            if (s.size() > (1u << 20)) {
                ok = std::fwrite(s.data(), 1, s.size(), fp) == s.size();
                s.clear();
            }
        }
    if (ok && !s.empty()) ok = std::fwrite(s.data(), 1, s.size(), fp) == s.size();
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) set_last_error_msg("write error");
    return ok ? ERLBM_OK : ERLBM_ERR_ARG;
}

int write_map(const char *path, const std::string &title, int n_i, int n_j, const int *map)
{
    if (!path || !map || n_i < 1 || n_j < 1) return ERLBM_ERR_ARG;
    std::FILE *fp = std::fopen(path, "wb");
    if (!fp) {
        set_last_error_msg("cannot open the output file");
        return ERLBM_ERR_ARG;
    }
    std::string s = "# vtk DataFile Version 2.0\n" + title + "\nASCII\nDATASET STRUCTURED_POINTS\nDIMENSIONS " +
                    std::to_string(n_i) + " " + std::to_string(n_j) + " 1\nASPECT_RATIO 1 1 1\nORIGIN 0 0 0\n" +
                    "POINT_DATA " + std::to_string(n_i * n_j) + "\nSCALARS Stability int\nLOOKUP_TABLE default\n";
    for (int j = 0; j < n_j; j++) {
        for (int i = 0; i < n_i; i++) {
            s += std::to_string(map[i + (size_t)j * n_i]);
            s += ' ';
        }
        s += '\n';
    }
    bool ok = std::fwrite(s.data(), 1, s.size(), fp) == s.size();
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) set_last_error_msg("write error");
    return ok ? ERLBM_OK : ERLBM_ERR_ARG;
}

std::string g_str(double v)
{
    std::string s;
    put_g(s, v);
    return s;
}

const char *collision_name(int col)
{
    switch (col) {
    case ERLBM_COL_BGK: return "BGK";
    case ERLBM_COL_MRT: return "MRT";
    case ERLBM_COL_REG: return "RLBM";
    case ERLBM_COL_EXT_REG: return "E-RLBM";
    default: return "?";
    }
}

}  // namespace

// ---- C ABI ---------------------------------------------------------------------------
extern "C" int erlbm_velocity_field(erlbm_lattice *h, int which, double *vel)
{
    if (!h || !vel) return ERLBM_ERR_ARG;
    double *d = nullptr;
    int rc = velocity_device(h, which, &d);
    if (rc) return rc;
    return to_host(h, d, vel, 3 * (size_t)h->nvox);
}

extern "C" int erlbm_vorticity_field(erlbm_lattice *h, int which, double *vort)
{
    if (!h || !vort) return ERLBM_ERR_ARG;
    double *d = nullptr;
    int rc = vorticity_device(h, which, false, &d);
    if (rc) return rc;
    return to_host(h, d, vort, 3 * (size_t)h->nvox);
}

extern "C" int erlbm_enstrophy(erlbm_lattice *h, int which, double *enstrophy)
{
    if (!h || !enstrophy) return ERLBM_ERR_ARG;
    double *terms = nullptr;
    int rc = vorticity_device(h, which, true, &terms);
    if (rc) return rc;
    double *d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(double));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_sum_x_outer, dim3(1), dim3(64), 0, h->stream, terms, h->nx, h->ny, h->nz, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(enstrophy, d, sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(terms);
    (void)hipFree(d);
    if (e != hipSuccess) {
        set_last_error("erlbm_enstrophy", e);
        return ERLBM_ERR_HIP;
    }
    return ERLBM_OK;
}

extern "C" int erlbm_write_velocity_vtk(const char *path, int nx, int ny, int nz, const double *vel)
{
    return write_vectors(path, "Velocidade", "velocidade", nx, ny, nz, vel);
}

extern "C" int erlbm_write_vorticity_vtk(const char *path, int nx, int ny, int nz, const double *vort)
{
    return write_vectors(path, "Velocidade", "vorticity", nx, ny, nz, vort);
}

extern "C" int erlbm_read_velocity_vtk(const char *path, double *vel, long long cap, int *nx, int *ny, int *nz)
{
    if (!path || !nx || !ny || !nz) return ERLBM_ERR_ARG;
    std::ifstream in(path);
    if (!in) {
        set_last_error_msg("erlbm_read_velocity_vtk: cannot open the file");
        return ERLBM_ERR_ARG;
    }
    std::string line, word;
    for (int i = 0; i < 4; i++) std::getline(in, line);
    std::getline(in, line);
    std::istringstream dims(line);
    int a = 0, b = 0, c = 0;
    dims >> word >> a >> b >> c;
    if (word != "DIMENSIONS" || a < 1 || b < 1 || c < 1) {
        set_last_error_msg("erlbm_read_velocity_vtk: no DIMENSIONS line where rec_velocity writes it");
        return ERLBM_ERR_ARG;
    }
    for (int i = 0; i < 4; i++) std::getline(in, line);
    if (line.find("VECTORS") == std::string::npos) {
        set_last_error_msg("erlbm_read_velocity_vtk: unexpected header (VECTORS, ASCII expected)");
        return ERLBM_ERR_ARG;
    }
    *nx = a;
    *ny = b;
    *nz = c;
    if (!vel) return ERLBM_OK;
    const long long n = 3LL * a * b * c;
    if (cap < n) {
        set_last_error_msg("erlbm_read_velocity_vtk: buffer too small");
        return ERLBM_ERR_ARG;
    }
    for (long long k = 0; k < n; k++) in >> vel[k];
    if (!in) {
        set_last_error_msg("erlbm_read_velocity_vtk: the file ends early");
        return ERLBM_ERR_ARG;
    }
    return ERLBM_OK;
}

extern "C" int erlbm_vtk_round(double *v, long long n)
{
    if (!v || n < 0) return ERLBM_ERR_ARG;
    char b[40];
    for (long long k = 0; k < n; k++) {
        std::snprintf(b, sizeof b, "%g", v[k]);
        v[k] = std::strtod(b, nullptr);
    }
    return ERLBM_OK;
}

extern "C" int erlbm_velocity_l2_error(const double *sim, int nx, int ny, const double *ref, int nx_ref, int ny_ref,
                                       double *e_vector, double *e_magnitude)
{
    if (!sim || !ref || nx < 1 || ny < 1 || nx_ref < 1 || ny_ref < 1) return ERLBM_ERR_ARG;
    if (nx_ref % nx != 0 || ny_ref % ny != 0 || nx_ref / nx != ny_ref / ny) {
        set_last_error_msg("erlbm_velocity_l2_error: the reference must be an integer refinement of the run");
        return ERLBM_ERR_ARG;
    }
    const int s = nx_ref / nx;
    double sum_dif2 = 0.0, sum_vec2 = 0.0, sum_ur2 = 0.0;
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            const size_t pc = 3 * ((size_t)x + (size_t)y * nx);
            const size_t pr = 3 * ((size_t)x * s + (size_t)y * s * nx_ref);
            const double rx = ref[pr], ry = ref[pr + 1];
            const double ur = std::sqrt(rx * rx + ry * ry);
            const double cx = sim[pc], cy = sim[pc + 1];
            const double us = std::sqrt(cx * cx + cy * cy);
            sum_dif2 = sum_dif2 + (us - ur) * (us - ur);
            sum_vec2 = sum_vec2 + (cx - rx) * (cx - rx) + (cy - ry) * (cy - ry);
            sum_ur2 = sum_ur2 + ur * ur;
        }
    if (e_magnitude) *e_magnitude = std::sqrt(sum_dif2 / sum_ur2);
    if (e_vector) *e_vector = std::sqrt(sum_vec2 / sum_ur2);
    return ERLBM_OK;
}

extern "C" int erlbm_write_tg_map_vtk(const char *path, int n_ma, int n_nu, int collision, double tau_nh,
                                      const int *map)
{
    const std::string title = std::string("Stability map Taylor-Green: ") + collision_name(collision) +
                              " tau_nh = " + g_str(tau_nh) + " ; x: U0 = Umin + (i+1)(cs-Umin)/" +
                              std::to_string(n_ma) + ", Umin = 0.05 cs ; y: log10 nu = -5 + 3 j/" +
                              std::to_string(n_nu);
    return write_map(path, title, n_ma, n_nu, map);
}

extern "C" int erlbm_write_obstacle_map_vtk(const char *path, int n_tau, int n_nu, double Ma, const int *map)
{
    const std::string title = "Stability map obstacle: E-RLBM, Ma = " + g_str(Ma) +
                              " ; x: tau_nh = 0.5 + 0.6 i/" + std::to_string(n_tau) +
                              " ; y: log10 nu = -5 + 4.3 j/" + std::to_string(n_nu);
    return write_map(path, title, n_tau, n_nu, map);
}
