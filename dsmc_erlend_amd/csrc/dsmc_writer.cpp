// dsmc_writer.cpp - Printing (src/Sampling.cpp:36-268) over the public C-ABI.
//
// Writes DSMC2D_Field.plt (two header lines, then one row per cell in k*NY+l order with the
// 8 columns X Y N Ux Uy Tx Ty T in %18.10e) and the DSMC2D_Field.json sidecar, in the formats
// the reference's post-processing reads (postprocessing/compare_runs.py, plot_field.py).
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "../../include/dsmc_c.h"

namespace {

const char* wall_name(int w)
{
    return w == DSMC_WALL_SPECULAR ? "specular" : w == DSMC_WALL_DIFFUSE ? "diffuse" : "cll";
}

}  // namespace

extern "C" int dsmc_write_field(dsmc_ctx* ctx, int case_index, const char* directory)
{
    dsmc_params p;
    if (int rc = dsmc_get_params(ctx, case_index, &p)) return rc;
    dsmc_constants k;
    if (int rc = dsmc_get_constants(ctx, case_index, &k)) return rc;
    const size_t m = (size_t)p.nx * (size_t)p.ny;
    std::vector<double> f(6 * m);
    if (int rc = dsmc_get_fields(ctx, case_index, f.data(), m)) return rc;
    if (!directory) return DSMC_E_ARGUMENT;

    std::error_code ec;
    const std::filesystem::path dir(directory);
    std::filesystem::create_directories(dir, ec);
    if (ec) return DSMC_E_IO;
    const std::string plt = (dir / "DSMC2D_Field.plt").string();
    FILE* out = std::fopen(plt.c_str(), "wt");
    if (!out) return DSMC_E_IO;
    std::fprintf(out, "VARIABLES = \"X\" \"Y\" \"N\" \"Ux\" \"Uy\" \"Tx\" \"Ty\" \"T\"\n");
    std::fprintf(out, "ZONE I=%d J=%d\n", p.nx, p.ny);
    for (int kx = 0; kx < p.nx; ++kx)
        for (int ly = 0; ly < p.ny; ++ly) {
            const double* r = f.data() + 6 * ((size_t)kx * p.ny + ly);
            const double x = p.x1 + (kx + 0.5) * k.dx;
            const double y = p.y1 + (ly + 0.5) * k.dy;
            std::fprintf(out, "%18.10e %18.10e %18.10e %18.10e %18.10e %18.10e %18.10e %18.10e\n",
                         x, y, r[0], r[1], r[2], r[3], r[4], r[5]);
        }
    const bool bad = std::ferror(out) != 0;
    if (std::fclose(out) != 0 || bad) return DSMC_E_IO;

    std::ofstream js(dir / "DSMC2D_Field.json");
    js << std::setprecision(17)
       << "{\n  \"obstacle\": {\"type\": \"box\", \"center\": [" << p.box_x << ", " << p.box_y
       << "], \"width\": " << p.box_width << ", \"height\": " << p.box_height << "},\n"
       << "  \"angle_of_attack_degrees\": " << p.attack_angle_deg << ",\n"
       << "  \"wall_model\": \"" << wall_name(p.wall_model) << "\",\n"
       << "  \"freestream\": {\"number_density\": " << p.n_free << ", \"temperature\": "
       << p.t_free << ", \"mach\": " << p.mach << ", \"molar_mass\": " << p.molar_mass << "},\n"
       << "  \"grid\": {\"nx\": " << p.nx << ", \"ny\": " << p.ny << ", \"bounds\": [" << p.x1
       << ", " << p.x2 << ", " << p.y1 << ", " << p.y2 << "], \"thickness\": " << p.dz << "},\n"
       << "  \"sampling\": {\"dt\": " << p.dt << ", \"steps\": " << p.n_step
       << ", \"first_step\": " << p.first_sampling_step << ", \"period\": " << p.sampling_period
       << ", \"particles_per_cell\": " << p.npc_free << "},\n"
       << "  \"collision_model\": {\"reference_diameter\": " << p.d_ref
       << ", \"viscosity_index\": " << p.viscosity_index << ", \"reference_temperature\": "
       << p.t_ref << "},\n"
       << "  \"wall_temperature\": " << p.t_wall;
    if (p.wall_model == DSMC_WALL_CLL)
        js << ",\n  \"cll_accommodation\": {\"normal\": " << p.acc_n << ", \"tangential\": "
           << p.acc_t << "}";
    js << "\n}\n";
    js.close();
    return js ? 0 : DSMC_E_IO;
}
