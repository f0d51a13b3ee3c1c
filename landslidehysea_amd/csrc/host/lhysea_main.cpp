// L-HySEA.exe <data.dat> [options]: the reference's command-line contract
// (src/GPU/main.cxx:62-182) on one MI355X through the C-ABI of liblhysea_hip.so.
//
// Output follows src/GPU/ShallowWater.cu:211-334 and 462-474: six NetCDF files
// <prefijo>_{eta1,q1x,q1y,eta2,q2x,q2y}.nc written every tiempo_guardar
// (-1 never, 0 every step) with max_height at the end, or, when a gauge file is
// given, <prefijo>_eta_puntos.txt.
//
// Options after the data file (the reference grammar ignores them):
//   --precision fp64|fp32c   arithmetic of the step (default fp64)
//   --steps N                run N steps instead of while (t < tiempo_tot)
//   --dry-run                load, write the initial output, no GPU
//   --dump                   print the parsed parameters as JSON and exit
//   --quiet                  no per-step line
//   --pouliquen              Pouliquen friction law: four friction angles in
//                            data.dat (the reference selects the law at build
//                            time, src/GPU/Constantes.hxx:17)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <chrono>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/lhysea.h"
#include "lhysea_netcdf.hpp"
#include "lhysea_problem.hpp"

using namespace lhysea;

namespace {

void usage(const char* prog)
{
    std::fprintf(stderr,
                 "Uso:\n%s ficheroDatos [--precision fp64|fp32c] [--steps N] [--dry-run] [--dump] [--quiet] [--pouliquen]\n\n"
                 "Formato de ficheroDatos:\n"
                 "\tNombre de la batimetria\n\tLeer condiciones iniciales de fichero (0|1)\n"
                 "\tSi 0: Xmin Xmax Ymin Ymax Volumenes_x Volumenes_y\n"
                 "\tSi 1: Fichero de topografia, Fichero de estado inicial\n"
                 "\tBorde superior, inferior, izquierdo, derecho (1: abierto, 0 o -1: pared)\n"
                 "\tTiempo de simulacion\n\tTiempo de guardado (-1: no guardar, 0: todo)\n"
                 "\tLeer fichero de puntos (0|1) [fichero]\n\tCFL\n\tRatio de densidades\n"
                 "\tAngulo de reposo (--pouliquen: cuatro angulos)\n\tFriccion entre capas, agua-fondo, sedimentos-fondo\n"
                 "\tVelocidad maxima agua, sedimentos\n\tNormalizar (0|1) [L H]\n"
                 "\tPrefijo de los ficheros de guardado\n",
                 prog);
}

// The six output files of the reference (src/GPU/netcdf.cu:29-335).
struct Output {
    std::unique_ptr<NcWriter> nc[6];
    int var[6] = {0}, time_var[6] = {0}, max_var = -1;
    int nx = 0, ny = 0;
    long rec = 0;
    std::vector<float> buf;

    bool open(const Problem& pb)
    {
        static const char* suffix[6] = {"eta1", "q1x", "q1y", "eta2", "q2x", "q2y"};
        static const char* lname[6] = {"Wave amplitude", "Mass flow of water along x", "Mass flow of water along y",
                                       "Sediments", "Mass flow of sediments along x",
                                       "Mass flow of sediments along y"};
        const lhysea_params& p = pb.p;
        nx = p.nx;
        ny = p.ny;
        const size_t N = (size_t)nx * ny;
        std::vector<float> xg(nx), yg(ny), bati(N);
        for (int i = 0; i < nx; i++) xg[i] = (float)(pb.xmin * pb.L + (i + 0.5) * p.dx * pb.L);
        for (int j = 0; j < ny; j++) yg[j] = (float)(pb.ymin * pb.L + (j + 0.5) * p.dy * pb.L);
        for (size_t c = 0; c < N; c++) bati[c] = (float)((pb.depth[c] + pb.Hmin) * pb.H);
        char when[32];
        std::time_t now = std::time(nullptr);
        std::strftime(when, sizeof when, "%Y-%m-%d %H:%M:%S", std::localtime(&now));
        const double fac = (pb.Q / pb.H) * std::sqrt(pb.L) / std::pow(pb.H, 7.0 / 6.0);
        for (int v = 0; v < 6; v++) {
            std::string path = pb.prefijo + "_" + suffix[v] + ".nc";
            nc[v].reset(new NcWriter(path));
            NcWriter& w = *nc[v];
            if (!w.ok()) return false;
            int dx = w.def_dim("lon", nx), dy = w.def_dim("lat", ny), gx = -1, gy = -1;
            if (v == 0) {
                gx = w.def_dim("grid_x", nx);
                gy = w.def_dim("grid_y", ny);
            }
            int dt = w.def_dim("time", 0);
            int xv = w.def_var("lon", NcWriter::NC_FLOAT, {dx});
            int yv = w.def_var("lat", NcWriter::NC_FLOAT, {dy});
            int gxv = -1, gyv = -1, bv = -1;
            if (v == 0) {
                gxv = w.def_var("grid_x", NcWriter::NC_FLOAT, {gx});
                gyv = w.def_var("grid_y", NcWriter::NC_FLOAT, {gy});
                bv = w.def_var("bathymetry", NcWriter::NC_FLOAT, {gy, gx});
            }
            time_var[v] = w.def_var("time", NcWriter::NC_FLOAT, {dt});
            if (v == 0) max_var = w.def_var("max_height", NcWriter::NC_FLOAT, {dy, dx});
            var[v] = w.def_var(suffix[v], NcWriter::NC_FLOAT, {dt, dy, dx});
            w.put_att_text(var[v], "units", (v == 0 || v == 3) ? "meters" : "meters/second");
            w.put_att_text(var[v], "long_name", lname[v]);
            const float fill = -1e30f;
            if (v == 0) {
                w.put_att_text(bv, "long_name", "Grid bathymetry");
                w.put_att_text(bv, "standard_name", "depth");
                w.put_att_text(bv, "units", "meters");
                w.put_att_float(bv, "missing_value", fill);
                w.put_att_float(bv, "_FillValue", fill);
                w.put_att_text(gxv, "long_name", "Grid x axis");
                w.put_att_text(gxv, "units", "meters");
                w.put_att_text(gyv, "long_name", "Grid y axis");
                w.put_att_text(gyv, "units", "meters");
                w.put_att_text(max_var, "long_name", "Maximum wave amplitude");
                w.put_att_text(max_var, "units", "meters");
                w.put_att_float(max_var, "missing_value", fill);
                w.put_att_float(max_var, "_FillValue", fill);
            }
            w.put_att_text(xv, "long_name", "x axis");
            w.put_att_text(xv, "units", "meters");
            w.put_att_text(yv, "long_name", "y axis");
            w.put_att_text(yv, "units", "meters");
            w.put_att_text(time_var[v], "long_name", "Time");
            w.put_att_text(time_var[v], "units", "seconds since 1970-01-01");
            w.put_att_float(var[v], "missing_value", fill);
            w.put_att_float(var[v], "_FillValue", fill);
            w.put_att_text(-1, "Conventions", "CF-1.0");
            w.put_att_text(-1, "title", "LandSlide-HySEA model output");
            w.put_att_text(-1, "history", when);
            w.put_att_text(-1, "grid_name", pb.nombre_bati);
            w.put_att_float(-1, "simulation_time", (float)(pb.tiempo_tot * pb.T));
            w.put_att_float(-1, "CFL", (float)p.cfl);
            w.put_att_float(-1, "ratio_of_densities", (float)p.r);
            // friction law and angles in degrees as src/GPU/netcdf.cu:236-244 writes them
            char a[128];
            const double d = 180.0 / M_PI;
            if (p.friction_law == 1) {
                w.put_att_text(-1, "friction_law", "Pouliquen");
                std::snprintf(a, sizeof a, "%.4f, %.4f, %.4f, %.4f", (float)(p.angle1 * d), (float)(p.angle2 * d),
                              (float)(p.angle3 * d), (float)(p.angle4 * d));
            } else {
                w.put_att_text(-1, "friction_law", "Coulomb");
                std::snprintf(a, sizeof a, "%.4f", (float)(p.angle1 * d));
            }
            w.put_att_text(-1, "angle_of_repose", a);
            w.put_att_float(-1, "friction_between_layers", (float)(p.mfc / pb.L));
            w.put_att_float(-1, "friction_water_bottom", (float)(p.mf0 / fac));
            w.put_att_float(-1, "friction_sediments_bottom", (float)(p.mfs / fac));
            w.put_att_float(-1, "max_speed_water", (float)(p.vmax1 * pb.Q / pb.H));
            w.put_att_float(-1, "max_speed_sediments", (float)(p.vmax2 * pb.Q / pb.H));
            if (!w.enddef()) return false;
            w.put_var(xv, xg.data());
            w.put_var(yv, yg.data());
            if (v == 0) {
                w.put_var(gxv, xg.data());
                w.put_var(gyv, yg.data());
                w.put_var(bv, bati.data());
            }
        }
        buf.resize(N);
        return true;
    }

    // one record of every field (src/GPU/ShallowWater.cu:276-316)
    bool write(const Problem& pb, const std::vector<double>& W, double t)
    {
        const size_t N = (size_t)nx * ny;
        float tt = (float)(t * pb.T);
        for (int v = 0; v < 6; v++) {
            for (size_t c = 0; c < N; c++) {
                double x;
                switch (v) {
                    case 0: x = (W[c] + W[3 * N + c] - pb.depth[c] - pb.Hmin) * pb.H; break;
                    case 3: x = W[3 * N + c] * pb.H; break;
                    case 1: x = W[N + c] * pb.Q; break;
                    case 2: x = W[2 * N + c] * pb.Q; break;
                    case 4: x = W[4 * N + c] * pb.Q; break;
                    default: x = W[5 * N + c] * pb.Q; break;
                }
                buf[c] = (float)x;
            }
            if (!nc[v]->put_record(time_var[v], rec, &tt) || !nc[v]->put_record(var[v], rec, buf.data()))
                return false;
        }
        rec++;
        return true;
    }

    bool close(const Problem& pb, const std::vector<double>& emax)
    {
        const size_t N = (size_t)nx * ny;
        for (size_t c = 0; c < N; c++) buf[c] = (float)((emax[c] - pb.Hmin) * pb.H);
        bool ok = nc[0]->put_var(max_var, buf.data());
        for (int v = 0; v < 6; v++) ok = nc[v]->close() && ok;
        return ok;
    }
};

void gauge_line(FILE* fp, const Problem& pb, const std::vector<double>& W, double t)
{
    std::fprintf(fp, "%e", t * pb.T);
    for (int g : pb.gauge_index) {
        if (g >= 0)
            std::fprintf(fp, " %.8e", (W[g] - pb.depth[g] - pb.Hmin) * pb.H);
        else
            std::fprintf(fp, " -999");
    }
    std::fprintf(fp, "\n");
}

int check(int rc)
{
    if (rc != LHYSEA_OK) std::fprintf(stderr, "Error: %s\n", lhysea_last_error());
    return rc;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) {
        usage(argv[0]);
        return 1;
    }
    std::string data = argv[1];
    int precision = LHYSEA_FP64;
    long fixed_steps = -1;
    bool dry = false, dump = false, quiet = false;
    int friction_law = 0;
    for (int i = 2; i < argc; i++) {
        std::string a = argv[i];
        if (a == "--precision" && i + 1 < argc) {
            std::string v = argv[++i];
            if (v == "fp64") precision = LHYSEA_FP64;
            else if (v == "fp32c") precision = LHYSEA_FP32C;
            else { usage(argv[0]); return 1; }
        } else if (a == "--steps" && i + 1 < argc) {
            fixed_steps = std::atol(argv[++i]);
        } else if (a == "--dry-run") {
            dry = true;
        } else if (a == "--dump") {
            dump = true;
        } else if (a == "--quiet") {
            quiet = true;
        } else if (a == "--pouliquen") {
            friction_law = 1;
        } else {
            usage(argv[0]);
            return 1;
        }
    }
    Problem pb;
    std::string msg;
    if (load_problem(data, pb, msg, friction_law) != 0) {
        std::fprintf(stderr, "%s\n", msg.c_str());
        return 1;
    }
    pb.p.precision = precision;
    const lhysea_params& p = pb.p;
    if (dump) {
        std::printf("{\"nombre_bati\": \"%s\", \"prefijo\": \"%s\", \"nx\": %d, \"ny\": %d, "
                    "\"xmin\": %.17g, \"xmax\": %.17g, \"ymin\": %.17g, \"ymax\": %.17g, \"dx\": %.17g, \"dy\": %.17g, "
                    "\"borde\": [%.17g, %.17g, %.17g, %.17g], \"tiempo_tot\": %.17g, \"tiempo_guardar\": %.17g, "
                    "\"leer_fichero_puntos\": %d, \"cfl\": %.17g, \"r\": %.17g, \"angle1_deg\": %.17g, "
                    "\"friction_law\": %d, \"angles_deg\": [%.17g, %.17g, %.17g, %.17g], "
                    "\"mfc\": %.17g, \"mf0\": %.17g, \"mfs\": %.17g, \"vmax1\": %.17g, \"vmax2\": %.17g, "
                    "\"normalizar\": %d, \"L\": %.17g, \"H\": %.17g, \"Q\": %.17g, \"T\": %.17g, \"g\": %.17g, "
                    "\"eps_h\": %.17g, \"Hmin\": %.17g, \"gauges\": %zu}\n",
                    pb.nombre_bati.c_str(), pb.prefijo.c_str(), p.nx, p.ny, pb.xmin, pb.xmax, pb.ymin, pb.ymax,
                    p.dx, p.dy, p.borde_sup, p.borde_inf, p.borde_izq, p.borde_der, pb.tiempo_tot,
                    pb.tiempo_guardar, pb.leer_fichero_puntos, p.cfl, p.r, p.angle1 * 180.0 / M_PI, p.friction_law,
                    p.angle1 * 180.0 / M_PI, p.angle2 * 180.0 / M_PI, p.angle3 * 180.0 / M_PI,
                    p.angle4 * 180.0 / M_PI, p.mfc, p.mf0,
                    p.mfs, p.vmax1, p.vmax2, pb.normalizar, pb.L, pb.H, pb.Q, pb.T, p.g, p.eps_h, pb.Hmin,
                    pb.gauge_index.size());
        return 0;
    }
    std::printf("%s", describe_problem(pb).c_str());
    std::printf("HMIN_GLOBAL = %g\n", pb.Hmin);
    const size_t N = (size_t)p.nx * p.ny;

    Output out;
    FILE* fp = nullptr;
    const bool use_nc = pb.leer_fichero_puntos == 0;
    if (use_nc) {
        if (!out.open(pb)) {
            std::fprintf(stderr, "Error: cannot create the NetCDF output files\n");
            return 1;
        }
    } else {
        fp = std::fopen((pb.prefijo + "_eta_puntos.txt").c_str(), "wt");
        if (!fp) {
            std::fprintf(stderr, "Error: cannot create the gauge output file\n");
            return 1;
        }
    }
    std::vector<double> W = pb.W, emax(N);
    for (size_t c = 0; c < N; c++) emax[c] = W[c] + W[3 * N + c] - pb.depth[c];
    if (dry) {
        if (use_nc) {
            out.write(pb, W, 0.0);
            out.close(pb, emax);
        } else {
            gauge_line(fp, pb, W, 0.0);
            std::fclose(fp);
        }
        return 0;
    }

    int probe = lhysea_device_probe();
    if (probe == 1) {
        std::fprintf(stderr, "Error: No hay tarjeta grafica\n");
        return 1;
    }
    if (probe == 2) {
        std::fprintf(stderr, "Error: No hay ninguna tarjeta grafica soportada\n");
        return 1;
    }
    lhysea_sim* sim = nullptr;
    int rc = check(lhysea_create(&p, &sim));
    if (rc) {
        std::fprintf(stderr, rc == LHYSEA_ERR_GPU_MEMORY ? "Error: No hay memoria GPU suficiente\n"
                                                        : "Error: cannot create the simulation\n");
        return 1;
    }
    rc = check(lhysea_upload(sim, W.data(), pb.depth.data()));
    if (!rc) rc = check(lhysea_track_eta_max(sim, 1));
    double dt = 0.0;
    if (!rc) rc = check(lhysea_initial_dt(sim, &dt));
    if (!rc) std::printf("deltaT inicial = %e seg\n", dt * pb.T);
    double t = 0.0, next_save = 0.0;
    int32_t steps = 0;
    auto t0 = std::chrono::steady_clock::now();
    while (!rc) {
        if (fixed_steps >= 0 ? steps >= fixed_steps : !(t < pb.tiempo_tot)) break;
        if (pb.tiempo_guardar >= 0.0 && t >= next_save) {
            rc = check(lhysea_download(sim, W.data()));
            if (rc) break;
            if (use_nc) {
                if (!out.write(pb, W, t)) { std::fprintf(stderr, "Error: NetCDF write failed\n"); rc = 1; break; }
            } else {
                gauge_line(fp, pb, W, t);
            }
            next_save += pb.tiempo_guardar;
        }
        if (quiet && pb.tiempo_guardar != 0.0) {
            // no per-step line: advance to the next save time (or the end) in
            // one call; the device stops at the first step that starts at or
            // after it, where the loop above saves.  A save time already
            // passed (dt above the save interval) takes one step, as the
            // reference's loop does.
            double target = (fixed_steps >= 0) ? INFINITY : pb.tiempo_tot;
            if (pb.tiempo_guardar > 0.0 && next_save < target) target = next_save;
            const long left = (fixed_steps >= 0) ? fixed_steps - steps : (long)INT32_MAX;
            int32_t done = 0;
            rc = (target > t) ? check(lhysea_run_until(sim, target, (int32_t)(left < INT32_MAX ? left : INT32_MAX),
                                                       &done))
                              : check(lhysea_step(sim, 1));
        } else {
            rc = check(lhysea_step(sim, 1));
        }
        if (!rc) rc = check(lhysea_get_time(sim, &t, &dt, &steps));
        if (!rc && !quiet) std::printf("Iteracion %3d, deltaT = %e seg, Tiempo = %g seg\n", steps, dt * pb.T, t * pb.T);
    }
    auto t1 = std::chrono::steady_clock::now();
    if (!rc) {
        std::vector<double> emt(N);
        rc = check(lhysea_eta_max(sim, emax.data(), emt.data()));
    }
    lhysea_destroy(sim);
    if (use_nc) {
        if (!out.close(pb, emax) && !rc) rc = 1;
    } else {
        std::fclose(fp);
    }
    if (rc) return 1;
    std::printf("\nTiempo: %e seg\n", std::chrono::duration<double>(t1 - t0).count());
    return 0;
}
