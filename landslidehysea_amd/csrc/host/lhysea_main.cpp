// L-HySEA.exe <data.dat> [options]: the reference's command-line contract
// (src/GPU/main.cxx:62-182) on one MI355X through the C-ABI of liblhysea_hip.so:
// the problem is loaded here and run by lhysea_shallow_water_output, which
// writes the output as src/GPU/ShallowWater.cu:211-334 and 462-474 do: six
// NetCDF files <prefijo>_{eta1,q1x,q1y,eta2,q2x,q2y}.nc written every
// tiempo_guardar (-1 never, 0 every step) with max_height at the end, or, when
// a gauge file is given, <prefijo>_eta_puntos.txt.
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
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../../include/lhysea.h"
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

    lhysea_output_spec spec{};
    spec.prefix = pb.prefijo.c_str();
    spec.grid_name = pb.nombre_bati.c_str();
    spec.save_interval = pb.tiempo_guardar;
    spec.xmin = pb.xmin;
    spec.ymin = pb.ymin;
    spec.Hmin = pb.Hmin;
    spec.Q = pb.Q;
    spec.T = pb.T;
    spec.sim_time = pb.tiempo_tot;
    if (pb.leer_fichero_puntos != 0) {
        spec.gauges = pb.gauge_index.data();
        spec.ngauges = (int32_t)pb.gauge_index.size();
    }
    std::vector<double> W = pb.W;
    if (dry) {
        // the initial state and its eta max, no GPU
        const size_t N = (size_t)p.nx * p.ny;
        std::vector<double> emax(N);
        for (size_t c = 0; c < N; c++) emax[c] = W[c] + W[3 * N + c] - pb.depth[c];
        lhysea_output* out = nullptr;
        int rc = check(lhysea_output_open(&spec, &p, pb.depth.data(), &out));
        if (!rc) rc = check(lhysea_output_write(out, W.data(), 0.0));
        const int rc2 = check(lhysea_output_close(out, pb.leer_fichero_puntos == 0 ? emax.data() : nullptr));
        return (rc || rc2) ? 1 : 0;
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
    double secs = 0.0;
    int32_t steps = 0;
    const double t_end = (fixed_steps >= 0) ? INFINITY : pb.tiempo_tot;
    const int32_t max_steps = (fixed_steps >= 0 && fixed_steps < INT32_MAX) ? (int32_t)fixed_steps : INT32_MAX;
    const int rc = check(lhysea_shallow_water_output(&p, W.data(), pb.depth.data(), t_end, max_steps, quiet ? 0 : 1,
                                                     &spec, &secs, &steps));
    if (rc == LHYSEA_ERR_GPU_MEMORY) std::fprintf(stderr, "Error: No hay memoria GPU suficiente\n");
    if (rc) return 1;
    std::printf("\nTiempo: %e seg\n", secs);
    return 0;
}
