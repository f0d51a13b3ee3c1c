// Output of a run as the reference's shallowWater writes it, and the one-shot
// driver that writes it (lhysea_output_* and lhysea_shallow_water_output of
// include/lhysea.h).
//
// src/GPU/ShallowWater.cu:211-334 and 462-474 with src/GPU/netcdf.cu:29-411:
// six NetCDF files <prefix>_{eta1,q1x,q1y,eta2,q2x,q2y}.nc (grid axes, the
// bathymetry in eta1, one record per save time, max_height in eta1 at the
// end), or, with gauges, the text file <prefix>_eta_puntos.txt (one line per
// save time: the time and eta1 at each gauge, -999 outside the grid).  Values
// are written in physical units: lengths times H, discharges times Q, times
// times T.  Host code only: no GPU is needed to write.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <ctime>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../lhysea_internal.hpp"
#include "../../../include/lhysea.h"
#include "lhysea_netcdf.hpp"

using namespace lhysea;

struct lhysea_output {
    lhysea_params p{};
    std::string prefix, grid_name;
    double save_interval = -1.0, xmin = 0, ymin = 0, Hmin = 0, Q = 1, T = 1, sim_time = 0;
    std::vector<double> depth;          // H of the run (internal units, Hmin-shifted)
    std::vector<int32_t> gauges;        // gauge text output when not empty
    bool use_gauges = false;
    // NetCDF
    std::unique_ptr<NcWriter> nc[6];
    int var[6] = {0}, time_var[6] = {0}, max_var = -1;
    long rec = 0;
    std::vector<float> buf;
    // gauge text
    FILE* fp = nullptr;

    bool open_nc();
    bool write_nc(const double* W, double t);
    bool close_nc(const double* emax);
    void gauge_line(const double* W, double t);
};

namespace {

int out_fail(int code, const std::string& msg)
{
    set_last_error(msg);
    return code;
}

}  // namespace

// The six files of src/GPU/netcdf.cu:29-335 with their dimensions, variables
// and attributes (an output format other tools read, kept as it is).
bool lhysea_output::open_nc()
{
    static const char* suffix[6] = {"eta1", "q1x", "q1y", "eta2", "q2x", "q2y"};
    static const char* lname[6] = {"Wave amplitude", "Mass flow of water along x", "Mass flow of water along y",
                                   "Sediments", "Mass flow of sediments along x", "Mass flow of sediments along y"};
    const int nx = p.nx, ny = p.ny;
    const size_t N = (size_t)nx * ny;
    const double L = p.L, H = p.H;
    std::vector<float> xg(nx), yg(ny), bati(N);
    for (int i = 0; i < nx; i++) xg[i] = (float)(xmin * L + (i + 0.5) * p.dx * L);
    for (int j = 0; j < ny; j++) yg[j] = (float)(ymin * L + (j + 0.5) * p.dy * L);
    for (size_t c = 0; c < N; c++) bati[c] = (float)((depth[c] + Hmin) * H);
    char when[32];
    std::time_t now = std::time(nullptr);
    std::strftime(when, sizeof when, "%Y-%m-%d %H:%M:%S", std::localtime(&now));
    const double fac = (Q / H) * std::sqrt(L) / std::pow(H, 7.0 / 6.0);
    for (int v = 0; v < 6; v++) {
        nc[v].reset(new (std::nothrow) NcWriter(prefix + "_" + suffix[v] + ".nc"));
        if (!nc[v] || !nc[v]->ok()) return false;
        NcWriter& w = *nc[v];
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
        w.put_att_text(-1, "grid_name", grid_name);
        w.put_att_float(-1, "simulation_time", (float)(sim_time * T));
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
        w.put_att_float(-1, "friction_between_layers", (float)(p.mfc / L));
        w.put_att_float(-1, "friction_water_bottom", (float)(p.mf0 / fac));
        w.put_att_float(-1, "friction_sediments_bottom", (float)(p.mfs / fac));
        w.put_att_float(-1, "max_speed_water", (float)(p.vmax1 * Q / H));
        w.put_att_float(-1, "max_speed_sediments", (float)(p.vmax2 * Q / H));
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
bool lhysea_output::write_nc(const double* W, double t)
{
    const size_t N = (size_t)p.nx * p.ny;
    const float tt = (float)(t * T);
    for (int v = 0; v < 6; v++) {
        for (size_t c = 0; c < N; c++) {
            double x;
            switch (v) {
                case 0: x = (W[c] + W[3 * N + c] - depth[c] - Hmin) * p.H; break;
                case 3: x = W[3 * N + c] * p.H; break;
                case 1: x = W[N + c] * Q; break;
                case 2: x = W[2 * N + c] * Q; break;
                case 4: x = W[4 * N + c] * Q; break;
                default: x = W[5 * N + c] * Q; break;
            }
            buf[c] = (float)x;
        }
        if (!nc[v]->put_record(time_var[v], rec, &tt) || !nc[v]->put_record(var[v], rec, buf.data())) return false;
    }
    rec++;
    return true;
}

// max_height (src/GPU/ShallowWater.cu:462-474) and the files closed
bool lhysea_output::close_nc(const double* emax)
{
    const size_t N = (size_t)p.nx * p.ny;
    bool ok = true;
    if (emax) {
        for (size_t c = 0; c < N; c++) buf[c] = (float)((emax[c] - Hmin) * p.H);
        ok = nc[0]->put_var(max_var, buf.data());
    }
    for (int v = 0; v < 6; v++) ok = nc[v]->close() && ok;
    return ok;
}

// src/GPU/ShallowWater.cu:318-331
void lhysea_output::gauge_line(const double* W, double t)
{
    std::fprintf(fp, "%e", t * T);
    for (int32_t g : gauges) {
        if (g >= 0)
            std::fprintf(fp, " %.8e", (W[g] - depth[g] - Hmin) * p.H);
        else
            std::fprintf(fp, " -999");
    }
    std::fprintf(fp, "\n");
}

extern "C" {

int lhysea_output_open(const lhysea_output_spec* spec, const lhysea_params* p, const double* H, lhysea_output** out)
{
    if (!spec || !p || !H || !out || !spec->prefix) return out_fail(LHYSEA_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    if (p->nx <= 0 || p->ny <= 0 || spec->ngauges < 0 || (spec->ngauges > 0 && !spec->gauges))
        return out_fail(LHYSEA_ERR_ARGUMENT, "bad grid or gauge list");
    const size_t N = (size_t)p->nx * p->ny;
    for (int32_t g = 0; g < spec->ngauges; g++)
        if (spec->gauges[g] < -1 || spec->gauges[g] >= (int64_t)N)
            return out_fail(LHYSEA_ERR_ARGUMENT, "gauge cell outside the grid (use -1)");
    lhysea_output* o = new (std::nothrow) lhysea_output;
    if (!o) return out_fail(LHYSEA_ERR_HOST_MEMORY, "host allocation failed");
    try {
        o->p = *p;
        o->prefix = spec->prefix;
        o->grid_name = spec->grid_name ? spec->grid_name : "";
        o->save_interval = spec->save_interval;
        o->xmin = spec->xmin;
        o->ymin = spec->ymin;
        o->Hmin = spec->Hmin;
        o->Q = spec->Q;
        o->T = spec->T;
        o->sim_time = spec->sim_time;
        o->depth.assign(H, H + N);
        o->use_gauges = spec->gauges != nullptr;
        if (o->use_gauges) o->gauges.assign(spec->gauges, spec->gauges + spec->ngauges);
    } catch (const std::bad_alloc&) {
        delete o;
        return out_fail(LHYSEA_ERR_HOST_MEMORY, "host allocation failed");
    }
    if (o->use_gauges) {
        o->fp = std::fopen((o->prefix + "_eta_puntos.txt").c_str(), "wt");
        if (!o->fp) {
            delete o;
            return out_fail(LHYSEA_ERR_IO, "cannot create the gauge output file");
        }
    } else if (!o->open_nc()) {
        delete o;
        return out_fail(LHYSEA_ERR_IO, "cannot create the NetCDF output files");
    }
    *out = o;
    return LHYSEA_OK;
}

int lhysea_output_write(lhysea_output* o, const double* W, double t)
{
    if (!o || !W) return out_fail(LHYSEA_ERR_ARGUMENT, "null argument");
    if (o->use_gauges) {
        o->gauge_line(W, t);
        return std::ferror(o->fp) ? out_fail(LHYSEA_ERR_IO, "gauge file write failed") : LHYSEA_OK;
    }
    return o->write_nc(W, t) ? LHYSEA_OK : out_fail(LHYSEA_ERR_IO, "NetCDF write failed");
}

int lhysea_output_close(lhysea_output* o, const double* eta_max)
{
    if (!o) return LHYSEA_OK;
    bool ok;
    if (o->use_gauges) {
        ok = std::fclose(o->fp) == 0;
    } else {
        ok = o->close_nc(eta_max);
    }
    delete o;
    return ok ? LHYSEA_OK : out_fail(LHYSEA_ERR_IO, "closing the output files failed");
}

int lhysea_shallow_water_output(const lhysea_params* p, double* W, const double* H, double t_end, int32_t max_steps,
                                int32_t verbose, const lhysea_output_spec* spec, double* elapsed_s,
                                int32_t* steps_done)
{
    if (!p || !W || !H || !spec || max_steps < 0) return out_fail(LHYSEA_ERR_ARGUMENT, "bad arguments");
    const size_t N = (size_t)p->nx * p->ny;
    if (steps_done) *steps_done = 0;
    if (elapsed_s) *elapsed_s = 0.0;
    lhysea_output* out = nullptr;
    int rc = lhysea_output_open(spec, p, H, &out);
    if (rc) return rc;
    lhysea_sim* s = nullptr;
    rc = lhysea_create(p, &s);
    if (!rc) rc = lhysea_upload(s, W, H);
    if (!rc) rc = lhysea_track_eta_max(s, 1);
    double dt = 0.0, t = 0.0, next_save = 0.0;
    if (!rc) rc = lhysea_initial_dt(s, &dt);
    if (!rc && verbose) std::printf("deltaT inicial = %e seg\n", dt * spec->T);
    int32_t steps = 0;
    const double save = spec->save_interval;
    const auto t0 = std::chrono::steady_clock::now();
    // while (tiempo_act < tiempo_tot) with the save at its top
    // (src/GPU/ShallowWater.cu:269-334), at most max_steps steps
    while (!rc && t < t_end && steps < max_steps) {
        if (save >= 0.0 && t >= next_save) {
            rc = lhysea_download(s, W);
            if (!rc) rc = lhysea_output_write(out, W, t);
            if (rc) break;
            next_save += save;
        }
        if (verbose || save == 0.0) {
            rc = lhysea_step(s, 1);
        } else {
            // to the next save time (or the end) in one call: the device stops
            // at the first step that starts at or after it, where the loop
            // saves; a save time already passed takes one step
            double target = t_end;
            if (save > 0.0 && next_save < target) target = next_save;
            int32_t done = 0;
            rc = (target > t) ? lhysea_run_until(s, target, max_steps - steps, &done) : lhysea_step(s, 1);
        }
        if (!rc) rc = lhysea_get_time(s, &t, &dt, &steps);
        if (!rc && verbose)
            std::printf("Iteracion %3d, deltaT = %e seg, Tiempo = %g seg\n", steps, dt * spec->T, t * spec->T);
    }
    const auto t1 = std::chrono::steady_clock::now();
    if (elapsed_s) *elapsed_s = std::chrono::duration<double>(t1 - t0).count();
    if (steps_done) *steps_done = steps;
    std::vector<double> emax, emt;
    if (!rc) {
        try {
            emax.resize(N);
            emt.resize(N);
        } catch (const std::bad_alloc&) {
            rc = out_fail(LHYSEA_ERR_HOST_MEMORY, "host allocation failed");
        }
    }
    if (!rc) rc = lhysea_eta_max(s, emax.data(), emt.data());
    if (!rc) rc = lhysea_download(s, W);
    if (s) lhysea_destroy(s);
    const std::string err = rc ? std::string(lhysea_last_error()) : std::string();
    const int rc_close = lhysea_output_close(out, rc ? nullptr : emax.data());
    if (rc) {
        set_last_error(err);
        return rc;
    }
    return rc_close;
}

}  // extern "C"
