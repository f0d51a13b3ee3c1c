// data.dat / .bat / .ini loading (src/GPU/Problema.cxx:139-549), one process.
#include "lhysea_problem.hpp"

#include <sys/stat.h>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <sstream>

namespace lhysea {

namespace {

bool exists(const std::string& f)
{
    struct stat st;
    return stat(f.c_str(), &st) == 0;
}

// nearest-index search of the reference (src/GPU/Problema.cxx:8-42)
int point_index(const std::vector<float>& X, const std::vector<float>& Y, float lon, float lat)
{
    int nx = (int)X.size(), ny = (int)Y.size(), i = 0, pos = -1;
    bool found = false;
    while (i < nx && !found) {
        if (X[i] >= lon) {
            found = true;
            int im = i > 0 ? i - 1 : 0;
            if (std::fabs(lon - X[im]) < std::fabs(lon - X[i])) i = im;
        } else {
            i++;
        }
    }
    if (found) {
        int j = 0;
        found = false;
        while (j < ny && !found) {
            if (Y[j] >= lat) {
                found = true;
                int jm = j > 0 ? j - 1 : 0;
                if (std::fabs(lat - Y[jm]) < std::fabs(lat - Y[j])) j = jm;
                pos = j * nx + i;
            } else {
                j++;
            }
        }
    }
    return pos;
}

}  // namespace

int load_problem(const std::string& fich_ent, Problem& pb, std::string& msg, int friction_law)
{
    // directory of data.dat: relative file names are resolved against it
    size_t slash = fich_ent.find_last_of("/\\");
    pb.directory = (slash == std::string::npos) ? "" : fich_ent.substr(0, slash + 1);
    std::ifstream f(fich_ent.c_str());
    if (!f) {
        msg = "Error: No se ha encontrado el fichero '" + fich_ent + "'";
        return 1;
    }
    lhysea_params& p = pb.p;
    p = lhysea_params{};
    std::string fich_topo, fich_est, fich_puntos;
    std::ifstream bat;
    int nx = 0, ny = 0;
    // positional stream >> grammar (src/GPU/Problema.cxx:189-287); an int read
    // stops at the first non-digit, which the 22-line Port Valdez file relies on
    f >> pb.nombre_bati >> pb.leer_de_fichero;
    if (pb.leer_de_fichero == 0) {
        f >> pb.xmin >> pb.xmax >> pb.ymin >> pb.ymax >> nx >> ny;
    } else {
        f >> fich_topo >> fich_est;
        fich_topo = pb.directory + fich_topo;
        fich_est = pb.directory + fich_est;
        if (!exists(fich_topo)) {
            msg = "Error: No se ha encontrado el fichero '" + fich_topo + "'";
            return 1;
        }
        if (!exists(fich_est)) {
            msg = "Error: No se ha encontrado el fichero '" + fich_est + "'";
            return 1;
        }
        bat.open(fich_topo.c_str());
        bat >> pb.xmin >> pb.xmax >> pb.ymin >> pb.ymax >> nx >> ny;
    }
    f >> p.borde_sup >> p.borde_inf >> p.borde_izq >> p.borde_der;
    f >> pb.tiempo_tot >> pb.tiempo_guardar;
    f >> pb.leer_fichero_puntos;
    if (pb.leer_fichero_puntos == 1) {
        f >> fich_puntos;
        fich_puntos = pb.directory + fich_puntos;
        if (!exists(fich_puntos)) {
            msg = "Error: No se ha encontrado el fichero '" + fich_puntos + "'";
            return 1;
        }
    }
    f >> p.cfl >> p.r;
    // Coulomb: one angle; Pouliquen: four (src/GPU/Problema.cxx:240-256, chosen
    // at build time there by COULOMB in src/GPU/Constantes.hxx:17)
    double a[4] = {0, 0, 0, 0};
    f >> a[0];
    if (friction_law == 1) f >> a[1] >> a[2] >> a[3];
    else a[1] = a[2] = a[3] = a[0];
    p.angle1 = a[0] * M_PI / 180.0;
    p.angle2 = a[1] * M_PI / 180.0;
    p.angle3 = a[2] * M_PI / 180.0;
    p.angle4 = a[3] * M_PI / 180.0;
    p.friction_law = friction_law;
    f >> p.mfc >> p.mf0 >> p.mfs >> p.vmax1 >> p.vmax2;
    f >> pb.normalizar;
    if (pb.normalizar == 0) {
        p.g = 9.81;
        pb.L = pb.H = pb.Q = pb.T = 1.0;
    } else {
        f >> pb.L >> pb.H;
        pb.Q = std::sqrt(9.81 * std::pow(pb.H, 3.0));
        pb.T = pb.L * pb.H / pb.Q;
        p.g = 1.0;
        pb.xmin /= pb.L;
        pb.xmax /= pb.L;
        pb.ymin /= pb.L;
        pb.ymax /= pb.L;
        pb.tiempo_tot /= pb.T;
        pb.tiempo_guardar /= pb.T;
        p.mfc *= pb.L;
        const double fac = (pb.Q / pb.H) * std::sqrt(pb.L) / std::pow(pb.H, 7.0 / 6.0);
        p.mf0 *= fac;
        p.mfs *= fac;
        p.vmax1 /= pb.Q / pb.H;
        p.vmax2 /= pb.Q / pb.H;
    }
    f >> pb.prefijo;
    if (!f && !f.eof()) {
        msg = "Error: malformed data file '" + fich_ent + "'";
        return 1;
    }
    p.eps_h = 5e-3 / pb.H;
    p.L = pb.L;
    p.H = pb.H;
    p.peso = p.beta = 1.0;
    if (nx < 1 || ny < 1) {
        msg = "Error: invalid grid size";
        return 1;
    }
    p.nx = nx;
    p.ny = ny;
    p.dx = (pb.xmax - pb.xmin) / nx;
    p.dy = (pb.ymax - pb.ymin) / ny;

    const size_t N = (size_t)nx * ny;
    // gauges (src/GPU/Problema.cxx:299-348), with the reference's fixed
    // lon/lat projection
    if (pb.leer_fichero_puntos == 1) {
        std::ifstream g(fich_puntos.c_str());
        int np = 0;
        g >> np;
        std::vector<float> X(nx), Y(ny);
        for (int i = 0; i < nx; i++) X[i] = (float)((pb.xmin + i * p.dx) * pb.L);
        for (int j = 0; j < ny; j++) Y[j] = (float)((pb.ymin + j * p.dy) * pb.L);
        std::vector<int> first(N, -1);
        pb.gauge_index.assign(np, -1);
        for (int k = 0; k < np; k++) {
            float lon = 0, lat = 0;
            g >> lon >> lat;
            lon = 63259.0f * (lon + 67.7f) / 0.6f;
            lat = 55604.0f * (lat - 18.3f) / 0.5f;
            int pos = point_index(X, Y, lon, lat);
            if (pos >= 0) {
                if (first[pos] == -1) first[pos] = k;
                pb.gauge_index[k] = pos;
            }
        }
    }

    pb.W.assign(6 * N, 0.0);
    pb.depth.assign(N, 0.0);
    if (pb.leer_de_fichero == 0) {
        // analytic initial condition of the reference build: flat 5 m bottom,
        // both layers empty (src/cond_ini.cxx:5-22, 36-48)
        for (size_t c = 0; c < N; c++) pb.depth[c] = 5.0 / pb.H;
        pb.Hmin = 0.0;
        return 0;
    }
    double hmin = 1e30;
    for (size_t c = 0; c < N; c++) {
        double v = 0;
        if (!(bat >> v)) {
            msg = "Error: short bathymetry file '" + fich_topo + "'";
            return 1;
        }
        v /= pb.H;
        pb.depth[c] = v;
        if (v < hmin) hmin = v;
    }
    // depth shift (src/GPU/Problema.cxx:459-481)
    pb.Hmin = (hmin >= 0.0) ? 0.0 : hmin;
    if (pb.Hmin < 0.0)
        for (size_t c = 0; c < N; c++) pb.depth[c] -= pb.Hmin;
    std::ifstream ini(fich_est.c_str());
    for (size_t c = 0; c < N; c++) {
        double w[6];
        for (int k = 0; k < 6; k++)
            if (!(ini >> w[k])) {
                msg = "Error: short initial-state file '" + fich_est + "'";
                return 1;
            }
        pb.W[c] = w[0] / pb.H;
        pb.W[N + c] = w[1] / pb.Q;
        pb.W[2 * N + c] = w[2] / pb.Q;
        pb.W[3 * N + c] = w[3] / pb.H;
        pb.W[4 * N + c] = w[4] / pb.Q;
        pb.W[5 * N + c] = w[5] / pb.Q;
    }
    return 0;
}

std::string describe_problem(const Problem& pb)
{
    const lhysea_params& p = pb.p;
    std::ostringstream o;
    const double fac = (pb.Q / pb.H) * std::sqrt(pb.L) / std::pow(pb.H, 7.0 / 6.0);
    o << "Info Problema\n";
    o << "Volumenes: " << p.nx << " x " << p.ny << ", Total: " << (long)p.nx * p.ny << "\n";
    o << "X: [" << pb.xmin * pb.L << ", " << pb.xmax * pb.L << "]\n";
    o << "Y: [" << pb.ymin * pb.L << ", " << pb.ymax * pb.L << "]\n";
    o << "CFL: " << p.cfl << "\n";
    o << "r: " << p.r << "\n";
    if (p.friction_law == 1)
        o << "Angulos de reposo: " << p.angle1 * 180.0 / M_PI << ", " << p.angle2 * 180.0 / M_PI << ", "
          << p.angle3 * 180.0 / M_PI << ", " << p.angle4 * 180.0 / M_PI << "\n";
    else
        o << "Angulo de reposo: " << p.angle1 * 180.0 / M_PI << "\n";
    o << "Friccion entre capas: " << p.mfc / pb.L << "\n";
    o << "Friccion agua-fondo: " << p.mf0 / fac << "\n";
    o << "Friccion sedimento-fondo: " << p.mfs / fac << "\n";
    o << "Velocidad maxima agua: " << p.vmax1 * (pb.Q / pb.H) << "\n";
    o << "Velocidad maxima sedimentos: " << p.vmax2 * (pb.Q / pb.H) << "\n";
    o << "Tiempo de simulacion: " << pb.tiempo_tot * pb.T << " seg.\n";
    return o.str();
}

}  // namespace lhysea
