// Problem loading for the L-HySEA.exe front end: the data.dat grammar and the
// .bat/.ini readers of the reference (src/GPU/Problema.cxx:139-549), single
// process (no row decomposition).
#pragma once
#include <string>
#include <vector>
#include "../../../include/lhysea.h"

namespace lhysea {

struct Problem {
    lhysea_params p{};           // internal (possibly normalised) units
    std::string nombre_bati, prefijo, directory;
    double xmin = 0, xmax = 0, ymin = 0, ymax = 0;
    double Hmin = 0;             // depth shift applied to H (<= 0)
    double tiempo_tot = 0, tiempo_guardar = 0;
    double L = 1, H = 1, Q = 1, T = 1;
    int leer_de_fichero = 1;
    int leer_fichero_puntos = 0;
    int normalizar = 0;
    std::vector<int> gauge_index;        // per requested gauge: flat cell or -1
    std::vector<double> W;               // 6 planes h1,q1x,q1y,h2,q2x,q2y (nx*ny each)
    std::vector<double> depth;           // H, shifted so that min >= 0
};

// Returns 0 on success, 1 if a file is missing or unreadable (the reference's
// error code); msg receives a description.  friction_law 0 (Coulomb) reads one
// friction angle, 1 (Pouliquen) four.
int load_problem(const std::string& data_dat, Problem& pb, std::string& msg, int friction_law = 0);

// Parameter echo in the reference's format (src/GPU/Problema.cxx:556-580).
std::string describe_problem(const Problem& pb);

}  // namespace lhysea
