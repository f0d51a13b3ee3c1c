// C-ABI of the ensemble batch: many independent FP32-compensated scenarios
// stepped together, one batched launch (plus one finalize launch) per step,
// replayed from a captured hipGraph.
//
// The reference runs one scenario per process (src/GPU/main.cxx:157-164 calls
// shallowWater once); a batch is a set of such runs sharing launches.  Each
// scenario keeps its own grid, parameters, dt, time and stop condition (the
// reference's `while (tiempo_act < tiempo_tot)`, src/GPU/ShallowWater.cu:269),
// and its results are those of the single-simulation FP32C path: the same
// tile function computes both, and its result does not depend on the tile
// shape, so the autotuner may pick any shape.
#include <climits>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#include "lhysea_internal.hpp"
#include "../../include/lhysea.h"

using namespace lhysea;

namespace {

int bfail(int code, const std::string& msg)
{
    set_last_error(msg);
    return code;
}

int bhip_fail(hipError_t e, const char* where)
{
    return bfail((e == hipErrorOutOfMemory) ? LHYSEA_ERR_GPU_MEMORY : LHYSEA_ERR_HIP,
                 std::string(where) + ": " + hipGetErrorString(e));
}

#define BHIPCHK(expr)                                     \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return bhip_fail(e_, #expr); \
    } while (0)

// steps per captured graph; longer runs replay it, the remainder is launched directly
constexpr int kGraphSteps = 16;
constexpr int kMaxScen = kMaxBatchScen; // scenario index field of the block table (12 bits)
constexpr int kMaxTilesPerScen = 1 << 20;

struct Scen {
    lhysea_params p{};
    size_t N = 0;
    float* W[2] = {nullptr, nullptr};
    double* H = nullptr;
    double* emax = nullptr;
    int* gauges = nullptr;
    int ng = 0;
    double* gseries = nullptr;
    int gcap = 0, gstart = 0;
    double* dt_hist = nullptr;
    int hcap = 0;
    bool track_emax = false;
    double t_end = INFINITY;
    int max_steps = INT_MAX;
    // class tiles of the split shape (62 x 14): table and list per parity
    int* cls = nullptr;
    int* list = nullptr;
    int ctx = 0, cty = 0;
    bool cls_dirty = true;           // the class table does not describe the current state
};

}  // namespace

struct lhysea_batch {
    std::vector<Scen> sc;
    Ctl* ctl = nullptr;              // one per scenario
    BatchScen* d_sd = nullptr;       // device descriptors
    unsigned* d_tab = nullptr;       // block -> (scenario, tile)
    int nblocks = 0;
    int tile = 0;
    hipStream_t stream = nullptr;
    hipGraphExec_t exec = nullptr;
    bool sd_dirty = true;            // device descriptors differ from the host settings
    bool sd_outputs = true;          // device descriptors carry the gauge / eta max / dt outputs
    bool tab_dirty = true;
    double* stage = nullptr;         // FP64 values of one scenario staged for upload / download
    size_t stage_n = 0;              // cells the staging buffer holds
};

namespace {

// staging buffer of at least N cells (6 FP64 planes)
bool ensure_stage(lhysea_batch* b, size_t N)
{
    if (b->stage_n >= N) return true;
    hipFree(b->stage);
    b->stage_n = 0;
    if (hipMalloc(&b->stage, 6 * N * sizeof(double)) != hipSuccess) {
        b->stage = nullptr;
        return false;
    }
    b->stage_n = N;
    return true;
}

void drop_graph(lhysea_batch* b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
}

BatchScen descriptor(const lhysea_batch* b, int i, bool with_outputs)
{
    const Scen& s = b->sc[i];
    BatchScen d{};
    d.k = make_consts(s.p);
    d.H = s.H;
    d.W[0] = s.W[0];
    d.W[1] = s.W[1];
    d.ctl = b->ctl + i;
    d.t_end = s.t_end;
    d.max_steps = s.max_steps;
    const BatchTile& t = kBatchTiles[b->tile];
    d.tiles_x = (s.p.nx + t.tx - 1) / t.tx;
    d.tiles_y = (s.p.ny + t.ty - 1) / t.ty;
    d.cls = s.cls;
    d.list = s.list;
    d.ctiles_x = s.ctx;
    d.ctiles_y = s.cty;
    if (with_outputs) {
        d.emax = s.track_emax ? s.emax : nullptr;
        d.gauges = s.gauges;
        d.ng = s.ng;
        // gauge rows are indexed by the step count; offset the base so row 0 is
        // the first step after the gauges were set
        d.gseries = s.gseries ? s.gseries - (size_t)s.gstart * s.ng : nullptr;
        d.gcap = s.gstart + s.gcap;
        d.dt_hist = s.dt_hist;
        d.hcap = s.hcap;
    }
    return d;
}

int upload_descriptors(lhysea_batch* b, bool with_outputs)
{
    std::vector<BatchScen> hs(b->sc.size());
    for (size_t i = 0; i < hs.size(); i++) hs[i] = descriptor(b, (int)i, with_outputs);
    BHIPCHK(hipStreamSynchronize(b->stream));
    BHIPCHK(hipMemcpy(b->d_sd, hs.data(), hs.size() * sizeof(BatchScen), hipMemcpyHostToDevice));
    b->sd_dirty = false;
    b->sd_outputs = with_outputs;
    return LHYSEA_OK;
}

// Concatenated tile list of all scenarios, dealt to blocks so that XCD x
// (blocks x, x+8, ...) runs one contiguous run of it.
int build_table(lhysea_batch* b)
{
    const BatchTile& t = kBatchTiles[b->tile];
    std::vector<unsigned> tiles;
    for (size_t i = 0; i < b->sc.size(); i++) {
        const long nt = (long)((b->sc[i].p.nx + t.tx - 1) / t.tx) * ((b->sc[i].p.ny + t.ty - 1) / t.ty);
        if (nt >= kMaxTilesPerScen) return bfail(LHYSEA_ERR_ARGUMENT, "scenario too large for the block table");
        for (long k = 0; k < nt; k++) tiles.push_back(((unsigned)i << 20) | (unsigned)k);
    }
    const int nb = (int)tiles.size();
    std::vector<unsigned> tab(nb);
    const int q = nb >> 3, rem = nb & 7;
    for (int blk = 0; blk < nb; blk++) {
        const int xcd = blk & 7, slot = blk >> 3;
        tab[blk] = tiles[xcd * q + (xcd < rem ? xcd : rem) + slot];
    }
    BHIPCHK(hipStreamSynchronize(b->stream));
    hipFree(b->d_tab);
    b->d_tab = nullptr;
    BHIPCHK(hipMalloc(&b->d_tab, (size_t)nb * sizeof(unsigned)));
    BHIPCHK(hipMemcpy(b->d_tab, tab.data(), (size_t)nb * sizeof(unsigned), hipMemcpyHostToDevice));
    b->nblocks = nb;
    b->tab_dirty = false;
    drop_graph(b);
    return LHYSEA_OK;
}

int prepare(lhysea_batch* b, bool with_outputs)
{
    int rc;
    if (b->tab_dirty && (rc = build_table(b)) != LHYSEA_OK) return rc;
    if ((b->sd_dirty || b->sd_outputs != with_outputs) && (rc = upload_descriptors(b, with_outputs)) != LHYSEA_OK)
        return rc;
    if (b->tile == kBatchSplit) {
        // class tables of scenarios whose state changed other than by a split step
        for (size_t i = 0; i < b->sc.size(); i++) {
            if (!b->sc[i].cls_dirty) continue;
            BHIPCHK(f32c_batch_classify(descriptor(b, (int)i, with_outputs), b->d_sd, (int)i, b->stream));
            b->sc[i].cls_dirty = false;
        }
    } else {
        // the other shapes do not keep the class tables
        for (Scen& sc : b->sc) sc.cls_dirty = true;
    }
    return LHYSEA_OK;
}

int launch_direct(lhysea_batch* b, int nsteps)
{
    bool buf = true;
    for (const Scen& sc : b->sc) buf = buf && state_buf_fits(sc.N);
    for (int n = 0; n < nsteps; n++) {
        hipError_t e = f32c_batch_step(b->tile, b->d_sd, (int)b->sc.size(), b->d_tab, b->nblocks, buf, b->stream);
        if (e != hipSuccess) return bhip_fail(e, "batch step launch");
    }
    return LHYSEA_OK;
}

// Every argument of the step launches is fixed (state, parity, dt and time
// live in device memory), so one captured graph of kGraphSteps steps serves
// every replay; it holds kernel nodes only.
int ensure_graph(lhysea_batch* b)
{
    if (b->exec) return LHYSEA_OK;
    hipGraph_t g = nullptr;
    BHIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_direct(b, kGraphSteps);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc != LHYSEA_OK) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return bhip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->exec = nullptr;
        return bhip_fail(e, "hipGraphInstantiate");
    }
    return LHYSEA_OK;
}

int run_steps(lhysea_batch* b, int nsteps, bool use_graph, bool with_outputs = true)
{
    int rc = prepare(b, with_outputs);
    if (rc != LHYSEA_OK) return rc;
    int left = nsteps;
    if (use_graph && left >= kGraphSteps) {
        if ((rc = ensure_graph(b)) != LHYSEA_OK) return rc;
        for (; left >= kGraphSteps; left -= kGraphSteps) BHIPCHK(hipGraphLaunch(b->exec, b->stream));
    }
    return launch_direct(b, left);
}

int read_ctl(lhysea_batch* b, std::vector<Ctl>& c)
{
    c.resize(b->sc.size());
    BHIPCHK(hipStreamSynchronize(b->stream));
    BHIPCHK(hipMemcpy(c.data(), b->ctl, c.size() * sizeof(Ctl), hipMemcpyDeviceToHost));
    return LHYSEA_OK;
}

// Wave-speed pass for the initial dt (src/GPU/ShallowWater.cu:236-257), for
// every scenario or only those that have none yet (dt <= 0 after an upload).
int initial_dt(lhysea_batch* b, bool missing_only)
{
    std::vector<Ctl> c;
    int rc = read_ctl(b, c);
    if (rc != LHYSEA_OK) return rc;
    for (size_t i = 0; i < b->sc.size(); i++) {
        if (missing_only && c[i].dt_cur > 0) continue;
        const Consts k = make_consts(b->sc[i].p);
        hipError_t e = f32c_initial_dt(k, b->sc[i].W[c[i].step & 1], b->ctl + i, b->stream);
        if (e != hipSuccess) return bhip_fail(e, "initial dt");
    }
    BHIPCHK(hipStreamSynchronize(b->stream));
    return LHYSEA_OK;
}

bool bad_sid(const lhysea_batch* b, int sid) { return !b || sid < 0 || sid >= (int)b->sc.size(); }

// HIP-event time of nsteps steps (graph replays plus the remainder)
int time_run(lhysea_batch* b, int nsteps, bool with_outputs, double* total_ms)
{
    int rc = prepare(b, with_outputs);
    if (rc != LHYSEA_OK) return rc;
    if (nsteps >= kGraphSteps && (rc = ensure_graph(b)) != LHYSEA_OK) return rc;
    hipEvent_t e0, e1;
    BHIPCHK(hipEventCreate(&e0));
    hipError_t e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        hipEventDestroy(e0);
        return bhip_fail(e, "hipEventCreate");
    }
    e = hipStreamSynchronize(b->stream);
    if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
    if (e == hipSuccess) rc = run_steps(b, nsteps, true, with_outputs);
    if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (e != hipSuccess) return bhip_fail(e, "batch timing");
    if (rc != LHYSEA_OK) return rc;
    *total_ms = ms;
    return LHYSEA_OK;
}

}  // namespace

extern "C" {

int lhysea_batch_create(const lhysea_params* p, int32_t n, lhysea_batch** out)
{
    if (!p || !out) return bfail(LHYSEA_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    if (n < 1 || n > kMaxScen) return bfail(LHYSEA_ERR_ARGUMENT, "batch size must be in [1, 4096]");
    for (int i = 0; i < n; i++) {
        const lhysea_params& q = p[i];
        if (q.nx < 1 || q.ny < 1) return bfail(LHYSEA_ERR_ARGUMENT, "nx and ny must be positive");
        if (q.precision != LHYSEA_FP32C) return bfail(LHYSEA_ERR_ARGUMENT, "batches run the FP32C path");
        if (!(q.dx > 0) || !(q.dy > 0) || !(q.eps_h > 0) || !(q.H > 0))
            return bfail(LHYSEA_ERR_ARGUMENT, "dx, dy, eps_h and H must be positive");
        if (q.friction_law < 0 || q.friction_law > 1 || q.peso != 1.0 || q.beta != 1.0)
            return bfail(LHYSEA_ERR_ARGUMENT,
                         "the FP32C path implements the Coulomb and Pouliquen laws with peso = beta = 1");
    }
    if (lhysea_device_probe() != 0) return bfail(LHYSEA_ERR_NO_DEVICE, "no usable HIP device");
    lhysea_batch* b = new (std::nothrow) lhysea_batch;
    if (!b) return bfail(LHYSEA_ERR_HOST_MEMORY, "host allocation failed");
    auto cleanup = [&](int code, const char* m) {
        lhysea_batch_destroy(b);
        return bfail(code, m);
    };
    b->sc.resize(n);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(LHYSEA_ERR_HIP, "stream creation failed");
    if (hipMalloc(&b->ctl, (size_t)n * sizeof(Ctl)) != hipSuccess ||
        hipMalloc(&b->d_sd, (size_t)n * sizeof(BatchScen)) != hipSuccess)
        return cleanup(LHYSEA_ERR_GPU_MEMORY, "device allocation failed");
    std::vector<Ctl> c0(n);
    for (auto& c : c0) c.dt_cur = -1.0;
    if (hipMemcpy(b->ctl, c0.data(), (size_t)n * sizeof(Ctl), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(LHYSEA_ERR_HIP, "control block upload failed");
    for (int i = 0; i < n; i++) {
        Scen& s = b->sc[i];
        s.p = p[i];
        s.N = (size_t)p[i].nx * p[i].ny;
        s.hcap = 1 << 16;
        for (int k = 0; k < 2; k++)
            if (hipMalloc(&s.W[k], 12 * s.N * sizeof(float)) != hipSuccess)
                return cleanup(LHYSEA_ERR_GPU_MEMORY, "device state allocation failed");
        s.ctx = (p[i].nx + kBatchTiles[kBatchSplit].tx - 1) / kBatchTiles[kBatchSplit].tx;
        s.cty = (p[i].ny + kBatchTiles[kBatchSplit].ty - 1) / kBatchTiles[kBatchSplit].ty;
        const size_t nct = (size_t)s.ctx * s.cty;
        if (hipMalloc(&s.H, s.N * sizeof(double)) != hipSuccess ||
            hipMalloc(&s.emax, 2 * s.N * sizeof(double)) != hipSuccess ||
            hipMalloc(&s.dt_hist, (size_t)s.hcap * sizeof(double)) != hipSuccess ||
            hipMalloc(&s.cls, 2 * nct * sizeof(int)) != hipSuccess ||
            hipMalloc(&s.list, 2 * nct * sizeof(int)) != hipSuccess)
            return cleanup(LHYSEA_ERR_GPU_MEMORY, "device allocation failed");
    }
    *out = b;
    return LHYSEA_OK;
}

int lhysea_batch_destroy(lhysea_batch* b)
{
    if (!b) return LHYSEA_OK;
    if (b->stream) hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (Scen& s : b->sc) {
        hipFree(s.W[0]);
        hipFree(s.W[1]);
        hipFree(s.H);
        hipFree(s.emax);
        hipFree(s.gauges);
        hipFree(s.gseries);
        hipFree(s.dt_hist);
        hipFree(s.cls);
        hipFree(s.list);
    }
    hipFree(b->ctl);
    hipFree(b->stage);
    hipFree(b->d_sd);
    hipFree(b->d_tab);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return LHYSEA_OK;
}

int lhysea_batch_size(const lhysea_batch* b, int32_t* n)
{
    if (!b || !n) return bfail(LHYSEA_ERR_ARGUMENT, "null argument");
    *n = (int32_t)b->sc.size();
    return LHYSEA_OK;
}

int lhysea_batch_upload(lhysea_batch* b, int32_t sid, const double* W, const double* H)
{
    if (bad_sid(b, sid) || !W || !H) return bfail(LHYSEA_ERR_ARGUMENT, "bad scenario index or null argument");
    Scen& s = b->sc[sid];
    const size_t N = s.N;
    BHIPCHK(hipStreamSynchronize(b->stream));
    // the FP64 state staged on the device, split there into FP32 values and
    // compensations; eta max = (h1 + h2 - H, time 0) from it
    if (!ensure_stage(b, N)) return bfail(LHYSEA_ERR_GPU_MEMORY, "staging buffer allocation failed");
    BHIPCHK(hipMemcpyAsync(b->stage, W, 6 * N * sizeof(double), hipMemcpyHostToDevice, b->stream));
    BHIPCHK(hipMemcpyAsync(s.H, H, N * sizeof(double), hipMemcpyHostToDevice, b->stream));
    BHIPCHK(f32c_split_state(b->stage, s.W[0], N, b->stream));
    BHIPCHK(init_eta_max(b->stage, s.H, s.emax, N, b->stream));
    BHIPCHK(hipStreamSynchronize(b->stream));   // the caller's W and H are read
    Ctl c{};
    c.dt_cur = -1.0;
    BHIPCHK(hipMemcpy(b->ctl + sid, &c, sizeof(Ctl), hipMemcpyHostToDevice));
    if (s.gstart != 0) b->sd_dirty = true;
    s.gstart = 0;
    s.cls_dirty = true;
    return LHYSEA_OK;
}

int lhysea_batch_download(lhysea_batch* b, int32_t sid, double* W)
{
    if (bad_sid(b, sid) || !W) return bfail(LHYSEA_ERR_ARGUMENT, "bad scenario index or null argument");
    Ctl c;
    BHIPCHK(hipStreamSynchronize(b->stream));
    BHIPCHK(hipMemcpy(&c, b->ctl + sid, sizeof(Ctl), hipMemcpyDeviceToHost));
    const Scen& s = b->sc[sid];
    const size_t N = s.N;
    if (!ensure_stage(b, N)) return bfail(LHYSEA_ERR_GPU_MEMORY, "staging buffer allocation failed");
    BHIPCHK(f32c_merge_state(s.W[c.step & 1], b->stage, N, b->stream));
    BHIPCHK(hipStreamSynchronize(b->stream));
    BHIPCHK(hipMemcpy(W, b->stage, 6 * N * sizeof(double), hipMemcpyDeviceToHost));
    return LHYSEA_OK;
}

int lhysea_batch_initial_dt(lhysea_batch* b)
{
    if (!b) return bfail(LHYSEA_ERR_ARGUMENT, "null handle");
    return initial_dt(b, false);
}

int lhysea_batch_set_stop(lhysea_batch* b, int32_t sid, double t_end, int32_t max_steps)
{
    if (bad_sid(b, sid)) return bfail(LHYSEA_ERR_ARGUMENT, "bad scenario index");
    b->sc[sid].t_end = t_end;
    b->sc[sid].max_steps = (max_steps < 0) ? INT_MAX : max_steps;
    b->sd_dirty = true;
    return LHYSEA_OK;
}

int lhysea_batch_step(lhysea_batch* b, int32_t nsteps)
{
    if (!b || nsteps < 0) return bfail(LHYSEA_ERR_ARGUMENT, "bad handle or step count");
    int rc = initial_dt(b, true);
    return (rc != LHYSEA_OK) ? rc : run_steps(b, nsteps, true);
}

int lhysea_batch_step_eager(lhysea_batch* b, int32_t nsteps)
{
    if (!b || nsteps < 0) return bfail(LHYSEA_ERR_ARGUMENT, "bad handle or step count");
    int rc = initial_dt(b, true);
    return (rc != LHYSEA_OK) ? rc : run_steps(b, nsteps, false);
}

int lhysea_batch_sync(lhysea_batch* b)
{
    if (!b) return bfail(LHYSEA_ERR_ARGUMENT, "null handle");
    BHIPCHK(hipStreamSynchronize(b->stream));
    return LHYSEA_OK;
}

int lhysea_batch_get_time(lhysea_batch* b, int32_t sid, double* t, double* dt_next, int32_t* steps)
{
    if (bad_sid(b, sid)) return bfail(LHYSEA_ERR_ARGUMENT, "bad scenario index");
    Ctl c;
    BHIPCHK(hipStreamSynchronize(b->stream));
    BHIPCHK(hipMemcpy(&c, b->ctl + sid, sizeof(Ctl), hipMemcpyDeviceToHost));
    if (t) *t = c.t;
    if (dt_next) *dt_next = c.dt_cur;
    if (steps) *steps = c.step;
    return LHYSEA_OK;
}

int lhysea_batch_track_eta_max(lhysea_batch* b, int32_t sid, int enable)
{
    if (bad_sid(b, sid)) return bfail(LHYSEA_ERR_ARGUMENT, "bad scenario index");
    b->sc[sid].track_emax = enable != 0;
    b->sd_dirty = true;
    return LHYSEA_OK;
}

int lhysea_batch_eta_max(lhysea_batch* b, int32_t sid, double* val, double* time)
{
    if (bad_sid(b, sid)) return bfail(LHYSEA_ERR_ARGUMENT, "bad scenario index");
    const Scen& s = b->sc[sid];
    std::vector<double> em(2 * s.N);
    BHIPCHK(hipStreamSynchronize(b->stream));
    BHIPCHK(hipMemcpy(em.data(), s.emax, 2 * s.N * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < s.N; i++) {
        if (val) val[i] = em[2 * i];
        if (time) time[i] = em[2 * i + 1];
    }
    return LHYSEA_OK;
}

int lhysea_batch_set_gauges(lhysea_batch* b, int32_t sid, const int32_t* cells, int32_t ng, int32_t max_steps)
{
    if (bad_sid(b, sid) || ng < 0 || max_steps < 0 || (ng > 0 && !cells))
        return bfail(LHYSEA_ERR_ARGUMENT, "bad gauge arguments");
    Scen& s = b->sc[sid];
    for (int g = 0; g < ng; g++)
        if (cells[g] < 0 || (size_t)cells[g] >= s.N) return bfail(LHYSEA_ERR_ARGUMENT, "gauge cell out of range");
    BHIPCHK(hipStreamSynchronize(b->stream));
    hipFree(s.gauges);
    hipFree(s.gseries);
    s.gauges = nullptr;
    s.gseries = nullptr;
    s.ng = 0;
    s.gcap = 0;
    b->sd_dirty = true;
    if (ng == 0 || max_steps == 0) return LHYSEA_OK;
    BHIPCHK(hipMalloc(&s.gauges, (size_t)ng * sizeof(int)));
    BHIPCHK(hipMemcpy(s.gauges, cells, (size_t)ng * sizeof(int), hipMemcpyHostToDevice));
    BHIPCHK(hipMalloc(&s.gseries, (size_t)ng * max_steps * sizeof(double)));
    Ctl c;
    BHIPCHK(hipMemcpy(&c, b->ctl + sid, sizeof(Ctl), hipMemcpyDeviceToHost));
    s.ng = ng;
    s.gcap = max_steps;
    s.gstart = c.step;
    return LHYSEA_OK;
}

int lhysea_batch_get_gauges(lhysea_batch* b, int32_t sid, double* out, int32_t* nrec)
{
    if (bad_sid(b, sid)) return bfail(LHYSEA_ERR_ARGUMENT, "bad scenario index");
    const Scen& s = b->sc[sid];
    Ctl c;
    BHIPCHK(hipStreamSynchronize(b->stream));
    BHIPCHK(hipMemcpy(&c, b->ctl + sid, sizeof(Ctl), hipMemcpyDeviceToHost));
    int n = c.step - s.gstart;
    if (n > s.gcap) n = s.gcap;
    if (n < 0 || !s.gseries) n = 0;
    if (nrec) *nrec = n;
    if (out && n > 0) BHIPCHK(hipMemcpy(out, s.gseries, (size_t)n * s.ng * sizeof(double), hipMemcpyDeviceToHost));
    return LHYSEA_OK;
}

int lhysea_batch_health(lhysea_batch* b, int32_t sid, uint32_t* flags, int32_t* first_step)
{
    if (bad_sid(b, sid) || !flags) return bfail(LHYSEA_ERR_ARGUMENT, "bad arguments");
    Ctl c;
    BHIPCHK(hipStreamSynchronize(b->stream));
    BHIPCHK(hipMemcpy(&c, b->ctl + sid, sizeof(Ctl), hipMemcpyDeviceToHost));
    *flags = c.health;
    if (first_step) *first_step = c.health ? c.health_step : -1;
    return LHYSEA_OK;
}

int lhysea_batch_dt_history(lhysea_batch* b, int32_t sid, double* out, int32_t max_steps, int32_t* nrec)
{
    if (bad_sid(b, sid) || max_steps < 0) return bfail(LHYSEA_ERR_ARGUMENT, "bad arguments");
    const Scen& s = b->sc[sid];
    Ctl c;
    BHIPCHK(hipStreamSynchronize(b->stream));
    BHIPCHK(hipMemcpy(&c, b->ctl + sid, sizeof(Ctl), hipMemcpyDeviceToHost));
    int n = c.step < s.hcap ? c.step : s.hcap;
    if (n > max_steps) n = max_steps;
    if (nrec) *nrec = n;
    if (out && n > 0) BHIPCHK(hipMemcpy(out, s.dt_hist, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    return LHYSEA_OK;
}

int lhysea_batch_tile_shapes(int32_t* shapes, int32_t* count)
{
    if (count) *count = kNumBatchTiles;
    if (shapes)
        for (int i = 0; i < kNumBatchTiles; i++) {
            shapes[3 * i] = kBatchTiles[i].tx;
            shapes[3 * i + 1] = kBatchTiles[i].ty;
            shapes[3 * i + 2] = kBatchTiles[i].nt;
        }
    return LHYSEA_OK;
}

int lhysea_batch_set_tile(lhysea_batch* b, int32_t tile)
{
    if (!b || tile < 0 || tile >= kNumBatchTiles) return bfail(LHYSEA_ERR_ARGUMENT, "bad handle or tile index");
    if (tile != b->tile) {
        BHIPCHK(hipStreamSynchronize(b->stream));
        b->tile = tile;
        b->tab_dirty = true;
        b->sd_dirty = true;
        drop_graph(b);
    }
    return LHYSEA_OK;
}

int lhysea_batch_get_tile(const lhysea_batch* b, int32_t* tile)
{
    if (!b || !tile) return bfail(LHYSEA_ERR_ARGUMENT, "null argument");
    *tile = b->tile;
    return LHYSEA_OK;
}

int lhysea_batch_time_steps(lhysea_batch* b, int32_t nsteps, double* total_ms)
{
    if (!b || nsteps < 1 || !total_ms) return bfail(LHYSEA_ERR_ARGUMENT, "bad arguments");
    int rc = initial_dt(b, true);
    return (rc != LHYSEA_OK) ? rc : time_run(b, nsteps, true, total_ms);
}

// Time every tile shape over nsteps steps from the current state and keep the
// fastest.  The state and times are restored afterwards: each scenario's
// current buffer and the control blocks are saved first and written back
// after every candidate; gauges, eta max and dt history are not written
// while timing.
int lhysea_batch_autotune(lhysea_batch* b, int32_t nsteps, double* ms_per_tile, int32_t* best)
{
    if (!b || nsteps < 1) return bfail(LHYSEA_ERR_ARGUMENT, "bad handle or step count");
    const int n = (int)b->sc.size();
    std::vector<Ctl> c;
    int rc = initial_dt(b, true);
    if (rc == LHYSEA_OK) rc = read_ctl(b, c);
    if (rc != LHYSEA_OK) return rc;
    std::vector<float*> save(n, nullptr);
    hipError_t e = hipSuccess;
    for (int i = 0; i < n && e == hipSuccess; i++) {
        const Scen& s = b->sc[i];
        e = hipMalloc(&save[i], 12 * s.N * sizeof(float));
        if (e == hipSuccess)
            e = hipMemcpy(save[i], s.W[c[i].step & 1], 12 * s.N * sizeof(float), hipMemcpyDeviceToDevice);
    }
    auto restore = [&]() -> int {
        BHIPCHK(hipStreamSynchronize(b->stream));
        for (int i = 0; i < n; i++) {
            const Scen& s = b->sc[i];
            BHIPCHK(hipMemcpy(s.W[c[i].step & 1], save[i], 12 * s.N * sizeof(float), hipMemcpyDeviceToDevice));
        }
        BHIPCHK(hipMemcpy(b->ctl, c.data(), (size_t)n * sizeof(Ctl), hipMemcpyHostToDevice));
        for (Scen& sc : b->sc) sc.cls_dirty = true;
        return LHYSEA_OK;
    };
    const int keep = b->tile;
    int pick = keep;
    double best_ms = INFINITY;
    rc = (e == hipSuccess) ? LHYSEA_OK : bhip_fail(e, "autotune snapshot");
    for (int t = 0; t < kNumBatchTiles && rc == LHYSEA_OK; t++) {
        rc = lhysea_batch_set_tile(b, t);
        if (rc == LHYSEA_OK) rc = prepare(b, false);
        if (rc == LHYSEA_OK) rc = launch_direct(b, 1);           // first-launch cost out of the timing
        if (rc == LHYSEA_OK) rc = restore();
        double ms = 0.0;
        if (rc == LHYSEA_OK) rc = time_run(b, nsteps, false, &ms);
        if (rc == LHYSEA_OK) rc = restore();
        if (rc != LHYSEA_OK) break;
        if (ms_per_tile) ms_per_tile[t] = ms;
        if (ms < best_ms) {
            best_ms = ms;
            pick = t;
        }
    }
    hipStreamSynchronize(b->stream);
    for (float* p : save) hipFree(p);
    const int rc2 = lhysea_batch_set_tile(b, rc == LHYSEA_OK ? pick : keep);
    b->sd_dirty = true;
    if (rc != LHYSEA_OK) return rc;
    if (rc2 != LHYSEA_OK) return rc2;
    if (best) *best = pick;
    return LHYSEA_OK;
}

}  // extern "C"
