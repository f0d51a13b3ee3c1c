// C-ABI of the LandSlide-HySEA MI355X library: handle management, uploads,
// the step loop and the one-shot shallowWater-style driver.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <new>
#include <string>
#include <vector>
#include "lhysea_internal.hpp"
#include "../../include/lhysea.h"

using namespace lhysea;

namespace {

thread_local std::string g_err;

// Every failure return records its message and clears HIP's per-thread last
// error: a failed runtime call (an allocation beyond the device's memory, say)
// leaves that error set, and the next kernel-launch check would report it.
int fail(int code, const std::string& msg)
{
    g_err = msg;
    (void)hipGetLastError();
    return code;
}

int hip_fail(hipError_t e, const char* where)
{
    g_err = std::string(where) + ": " + hipGetErrorString(e);
    (void)hipGetLastError();
    return (e == hipErrorOutOfMemory) ? LHYSEA_ERR_GPU_MEMORY : LHYSEA_ERR_HIP;
}

#define HIPCHK(expr)                                   \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

}  // namespace

void lhysea::set_last_error(const std::string& msg)
{
    g_err = msg;
    (void)hipGetLastError();
}

int lhysea::numeric_fail(const Ctl& c)
{
    std::string m = "bad state from step " + std::to_string(c.health_step) + ":";
    if (c.health & HEALTH_NONFINITE) m += " non-finite value";
    if (c.health & HEALTH_NEGATIVE_DEPTH) m += " depth below -eps_h";
    g_err = m;
    return LHYSEA_ERR_NUMERIC;
}

int lhysea::check_params(const lhysea_params& p)
{
    if (p.nx < 1 || p.ny < 1) return fail(LHYSEA_ERR_ARGUMENT, "nx and ny must be positive");
    if (p.precision != LHYSEA_FP64 && p.precision != LHYSEA_FP32C)
        return fail(LHYSEA_ERR_ARGUMENT, "unknown precision");
    if (!(p.dx > 0) || !(p.dy > 0) || !(p.eps_h > 0) || !(p.H > 0) || !(p.peso > 0))
        return fail(LHYSEA_ERR_ARGUMENT, "dx, dy, eps_h, H and peso must be positive");
    if (p.friction_law != 0 && p.friction_law != 1)
        return fail(LHYSEA_ERR_ARGUMENT, "friction_law must be 0 (Coulomb) or 1 (Pouliquen)");
    if (p.precision == LHYSEA_FP32C && (p.peso != 1.0 || p.beta != 1.0))
        return fail(LHYSEA_ERR_ARGUMENT, "the FP32C path implements peso = beta = 1");
    return LHYSEA_OK;
}

struct lhysea_adj_work;

struct lhysea_sim {
    lhysea_params p{};
    Consts k{};
    size_t N = 0;
    hipStream_t stream = nullptr;
    int cur = 0;                 // buffer index holding the current state
    int mode = MODE_FUSED;
    int adj_mode = MODE_FUSED;   // FP32C reverse step: fused tile kernel or four passes
    Buf64 b64{};
    Buf32 b32{};
    Ctl* ctl = nullptr;
    double* emax = nullptr;      // 2N interleaved (value, time)
    double* stage = nullptr;     // FP32C: 6N FP64 values staged for upload / download (allocated on first use)
    bool track_emax = false;
    int* gauges = nullptr;
    int ng = 0;
    double* gseries = nullptr;
    int gcap = 0;
    int gstart = 0;              // step count when the gauges were set
    double* dt_hist = nullptr;
    int hcap = 0;
    double* dt_seq = nullptr;    // frozen dt sequence (adjoint replay)
    lhysea_adj_work* adj = nullptr;  // adjoint forward pass awaiting its reverse sweep
    // FP32C: tile classes of the step and the launch split they select
    int* tcls = nullptr;
    int* tlist = nullptr;
    int ntiles = 0;
    int split = 0;
    int split_mode = -1;         // -1 chosen per launch piece, 0 never split, 1 always
    int step_host = 0;           // ctl->step when the next launched step starts
};


// Calls that step or reconfigure the handle are refused between an adjoint
// forward pass and its reverse sweep, which replays the recorded steps.
static int adj_busy()
{
    return fail(LHYSEA_ERR_ARGUMENT, "an adjoint forward pass awaits lhysea_adjoint_backward");
}

Consts lhysea::make_consts(const lhysea_params& p)
{
    Consts k{};
    k.nx = p.nx;
    k.ny = p.ny;
    k.friction_law = p.friction_law;
    k.dx = p.dx;
    k.dy = p.dy;
    k.area = p.dx * p.dy;
    k.borde_sup = p.borde_sup;
    k.borde_inf = p.borde_inf;
    k.borde_izq = p.borde_izq;
    k.borde_der = p.borde_der;
    k.cfl = p.cfl;
    k.r = p.r;
    k.g = p.g;
    k.eps_h = p.eps_h;
    // per-run constants of defTerminoFriccion (src/GPU/Arista_kernel.cu:11-58)
    k.mu = std::fabs(std::tan(p.angle1)) * (p.L / p.H);
    k.tan1 = std::tan(p.angle1);
    k.tan2 = std::tan(p.angle2);
    k.tan3 = std::tan(p.angle3);
    k.tan4 = std::tan(p.angle4);
    k.LoverH = p.L / p.H;
    k.mfc = p.mfc;
    k.mf0 = p.mf0;
    k.mfs = p.mfs;
    k.vmax1 = p.vmax1;
    k.vmax2 = p.vmax2;
    k.peso = p.peso;
    k.beta = p.beta;
    ConstsF32& f = k.f;
    f.eps = (float)k.eps_h;
    f.eps4 = f.eps * f.eps * f.eps * f.eps;
    f.g = (float)k.g;
    f.gp = (float)(k.g * (1.0 - k.r));
    f.r = (float)k.r;
    f.mu = (float)k.mu;
    f.peso = (float)k.peso;
    f.area = (float)k.area;
    f.len_x = (float)k.dy;
    f.len_y = (float)k.dx;
    f.igp = (float)(1.0 / (k.g * (1.0 - k.r)));
    f.ten_eps = (float)(10.0 / k.eps_h);
    f.ieps = (float)(1.0 / k.eps_h);
    f.mfc = (float)k.mfc;
    f.mf0 = (float)k.mf0;
    f.mfs = (float)k.mfs;
    f.vmax1 = (float)k.vmax1;
    f.vmax2 = (float)k.vmax2;
    f.tan1 = (float)k.tan1;
    f.tan2 = (float)k.tan2;
    f.tan3 = (float)k.tan3;
    f.tan4 = (float)k.tan4;
    f.LoverH = (float)k.LoverH;
    return k;
}

static StepArgs step_args(lhysea_sim* s)
{
    StepArgs a{};
    a.ctl = s->ctl;
    a.dt_seq = s->dt_seq;
    a.emax = s->track_emax ? s->emax : nullptr;
    a.gauges = s->gauges;
    a.ng = s->ng;
    a.gseries = s->gseries;
    a.gcap = s->gcap;
    a.dt_hist = s->dt_hist;
    a.hcap = s->hcap;
    a.t_end = INFINITY;
    a.tcls = s->tcls;
    a.tlist = s->tlist;
    a.split = s->split;
    a.wn_prev = 1;
    return a;
}

// Launch split of the next steps from the control block just read: the
// sediment-free launch pays when enough tiles are sediment-free.
static void choose_split(lhysea_sim* s, const Ctl& c)
{
    s->step_host = c.step;
    s->split = (s->split_mode >= 0) ? s->split_mode
                                    : (s->ntiles - (int)c.gcount[c.step & 1]) >= f32c_split_min_bare();
}

// Tile classes of the current FP32C state at the current step; needed after
// every change of the state or step count other than a step.
// With choose, the launch split is chosen from the new classes (a host sync).
// Tile classes of the current state (or of the state at src, in the current
// state's layout).
static hipError_t classify(lhysea_sim* s, bool choose, const void* src = nullptr)
{
    hipError_t e = (s->p.precision == LHYSEA_FP32C)
                       ? f32c_classify(s->k, src ? (const float*)src : s->b32.W[s->cur], step_args(s), s->stream)
                       : f64_classify(s->k, src ? (const double*)src : s->b64.W[s->cur], step_args(s), s->stream);
    if (!choose) return e;
    Ctl c;
    if (e == hipSuccess) e = hipMemcpyAsync(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) choose_split(s, c);
    return e;
}

static StepArgs launch_args(lhysea_sim* s, hipEvent_t ev0, hipEvent_t ev1)
{
    StepArgs a = step_args(s);
    a.ev0 = ev0;
    a.ev1 = ev1;
    if (s->gseries) {
        // gauge rows are indexed by ctl->step; offset the base so row 0 is the
        // first step after lhysea_set_gauges
        a.gseries = s->gseries - (size_t)s->gstart * s->ng;
        a.gcap = s->gstart + s->gcap;
    }
    return a;
}

// One step from state buffer W to Wn (either precision's layout).
static hipError_t launch_one(lhysea_sim* s, const void* W, void* Wn, const StepArgs& a0)
{
    StepArgs a = a0;
    a.step = s->step_host++;
    if (s->p.precision == LHYSEA_FP64) {
        if (s->mode == MODE_FUSED) return f64_step_fused(s->k, (const double*)W, (double*)Wn, s->b64.H, a, s->stream);
        return f64_step_fourpass(s->k, (const double*)W, (double*)Wn, s->b64.H, s->b64.acc, a, s->stream);
    }
    return f32c_step_fused(s->k, (const float*)W, (float*)Wn, s->b32.H, a, s->stream);
}

static void* state_buf(lhysea_sim* s, int b)
{
    return s->p.precision == LHYSEA_FP64 ? (void*)s->b64.W[b] : (void*)s->b32.W[b];
}

static void set_state_buf(lhysea_sim* s, int b, void* p)
{
    if (s->p.precision == LHYSEA_FP64)
        s->b64.W[b] = (double*)p;
    else
        s->b32.W[b] = (float*)p;
}

static int launch_steps(lhysea_sim* s, int nsteps, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr,
                        double t_end = INFINITY)
{
    StepArgs a = launch_args(s, ev0, ev1);
    a.t_end = t_end;
    for (int n = 0; n < nsteps; n++) {
        const int nxt = 1 - s->cur;
        hipError_t e = launch_one(s, state_buf(s, s->cur), state_buf(s, nxt), a);
        if (e != hipSuccess) return hip_fail(e, "step launch");
        s->cur = nxt;
    }
    return LHYSEA_OK;
}

extern "C" {

static void adj_release(lhysea_sim* s);

int lhysea_device_probe(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return 1;
    for (int d = 0; d < n; d++) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) == hipSuccess && prop.major >= 9) return 0;
    }
    return 2;
}

const char* lhysea_last_error(void) { return g_err.c_str(); }

int lhysea_create(const lhysea_params* p, lhysea_sim** out)
{
    if (!p || !out) return fail(LHYSEA_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    if (const int rc = check_params(*p)) return rc;
    if (lhysea_device_probe() != 0) return fail(LHYSEA_ERR_NO_DEVICE, "no usable HIP device");
    lhysea_sim* s = new (std::nothrow) lhysea_sim;
    if (!s) return fail(LHYSEA_ERR_HOST_MEMORY, "host allocation failed");
    s->p = *p;
    s->k = make_consts(*p);
    s->N = (size_t)p->nx * p->ny;
    auto cleanup = [&](int code, const std::string& m) {
        lhysea_destroy(s);
        return fail(code, m);
    };
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return cleanup(LHYSEA_ERR_HIP, "stream creation failed");
    const size_t N = s->N;
    if (p->precision == LHYSEA_FP64) {
        for (int b = 0; b < 2; b++)
            if (hipMalloc(&s->b64.W[b], 6 * N * sizeof(double)) != hipSuccess)
                return cleanup(LHYSEA_ERR_GPU_MEMORY, "device state allocation failed");
        if (hipMalloc(&s->b64.H, N * sizeof(double)) != hipSuccess ||
            hipMalloc(&s->b64.acc, 7 * N * sizeof(double)) != hipSuccess)
            return cleanup(LHYSEA_ERR_GPU_MEMORY, "device allocation failed");
        if (hipMemsetAsync(s->b64.acc, 0, 7 * N * sizeof(double), s->stream) != hipSuccess)
            return cleanup(LHYSEA_ERR_HIP, "accumulator clear failed");
        s->ntiles = f64_tiles(s->k);
    } else {
        for (int b = 0; b < 2; b++)
            if (hipMalloc(&s->b32.W[b], 12 * N * sizeof(float)) != hipSuccess)
                return cleanup(LHYSEA_ERR_GPU_MEMORY, "device state allocation failed");
        if (hipMalloc(&s->b32.H, N * sizeof(double)) != hipSuccess)
            return cleanup(LHYSEA_ERR_GPU_MEMORY, "device allocation failed");
        s->ntiles = f32c_tiles(s->k);
    }
    // tile classes of the step: two parities of stamps and lists
    if (hipMalloc(&s->tcls, 2 * (size_t)s->ntiles * sizeof(int)) != hipSuccess ||
        hipMalloc(&s->tlist, 2 * (size_t)s->ntiles * sizeof(int)) != hipSuccess)
        return cleanup(LHYSEA_ERR_GPU_MEMORY, "device allocation failed");
    if (hipMalloc(&s->ctl, sizeof(Ctl)) != hipSuccess || hipMalloc(&s->emax, 2 * N * sizeof(double)) != hipSuccess)
        return cleanup(LHYSEA_ERR_GPU_MEMORY, "device allocation failed");
    {
        const Ctl c0{};   // no class table yet (cls_from past every step)
        if (hipMemcpy(s->ctl, &c0, sizeof(Ctl), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(LHYSEA_ERR_HIP, "control block upload failed");
    }
    s->hcap = 1 << 20;
    if (hipMalloc(&s->dt_hist, (size_t)s->hcap * sizeof(double)) != hipSuccess)
        return cleanup(LHYSEA_ERR_GPU_MEMORY, "device allocation failed");
    if (hipStreamSynchronize(s->stream) != hipSuccess) return cleanup(LHYSEA_ERR_HIP, "stream sync failed");
    *out = s;
    return LHYSEA_OK;
}

int lhysea_destroy(lhysea_sim* s)
{
    if (!s) return LHYSEA_OK;
    adj_release(s);
    if (s->stream) hipStreamSynchronize(s->stream);
    for (int b = 0; b < 2; b++) {
        hipFree(s->b64.W[b]);
        hipFree(s->b32.W[b]);
    }
    hipFree(s->b64.H);
    hipFree(s->b64.acc);
    hipFree(s->b32.H);
    hipFree(s->b32.acc);
    hipFree(s->ctl);
    hipFree(s->emax);
    hipFree(s->stage);
    hipFree(s->gauges);
    hipFree(s->gseries);
    hipFree(s->dt_hist);
    hipFree(s->dt_seq);
    hipFree(s->tcls);
    hipFree(s->tlist);
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
    return LHYSEA_OK;
}

int lhysea_upload(lhysea_sim* s, const double* W, const double* H)
{
    if (!s || !W || !H) return fail(LHYSEA_ERR_ARGUMENT, "null argument");
    if (s->adj) return adj_busy();
    const size_t N = s->N;
    HIPCHK(hipStreamSynchronize(s->stream));
    s->cur = 0;
    // the FP64 state on the device (FP32C: staged, then split into values and
    // compensations there), and eta max = (h1 + h2 - H, time 0) from it; the
    // copies go on the handle's stream, ahead of the kernels that read them
    const double* Wd;
    const double* Hd;
    if (s->p.precision == LHYSEA_FP64) {
        HIPCHK(hipMemcpyAsync(s->b64.W[0], W, 6 * N * sizeof(double), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(s->b64.H, H, N * sizeof(double), hipMemcpyHostToDevice, s->stream));
        Wd = s->b64.W[0];
        Hd = s->b64.H;
    } else {
        if (!s->stage && hipMalloc(&s->stage, 6 * N * sizeof(double)) != hipSuccess) {
            s->stage = nullptr;
            return fail(LHYSEA_ERR_GPU_MEMORY, "staging buffer allocation failed");
        }
        HIPCHK(hipMemcpyAsync(s->stage, W, 6 * N * sizeof(double), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(s->b32.H, H, N * sizeof(double), hipMemcpyHostToDevice, s->stream));
        HIPCHK(f32c_split_state(s->stage, s->b32.W[0], N, s->stream));
        Wd = s->stage;
        Hd = s->b32.H;
    }
    HIPCHK(init_eta_max(Wd, Hd, s->emax, N, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));   // the caller's W and H are read
    Ctl c{};
    c.dt_cur = -1.0;
    HIPCHK(hipMemcpy(s->ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice));
    s->gstart = 0;
    HIPCHK(classify(s, true));
    return LHYSEA_OK;
}

int lhysea_download(lhysea_sim* s, double* W)
{
    if (!s || !W) return fail(LHYSEA_ERR_ARGUMENT, "null argument");
    const size_t N = s->N;
    HIPCHK(hipStreamSynchronize(s->stream));
    if (s->p.precision == LHYSEA_FP64) {
        HIPCHK(hipMemcpy(W, s->b64.W[s->cur], 6 * N * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        if (!s->stage && hipMalloc(&s->stage, 6 * N * sizeof(double)) != hipSuccess) {
            s->stage = nullptr;
            return fail(LHYSEA_ERR_GPU_MEMORY, "staging buffer allocation failed");
        }
        HIPCHK(f32c_merge_state(s->b32.W[s->cur], s->stage, N, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        HIPCHK(hipMemcpy(W, s->stage, 6 * N * sizeof(double), hipMemcpyDeviceToHost));
    }
    return LHYSEA_OK;
}

int lhysea_initial_dt(lhysea_sim* s, double* dt)
{
    if (!s) return fail(LHYSEA_ERR_ARGUMENT, "null handle");
    if (s->adj) return adj_busy();
    hipError_t e = (s->p.precision == LHYSEA_FP64) ? f64_initial_dt(s->k, s->b64.W[s->cur], s->ctl, s->stream)
                                                   : f32c_initial_dt(s->k, s->b32.W[s->cur], s->ctl, s->stream);
    if (e != hipSuccess) return hip_fail(e, "initial dt");
    if (dt) {
        Ctl c;
        HIPCHK(hipMemcpyAsync(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        *dt = c.dt_cur;
    }
    return LHYSEA_OK;
}

int lhysea_set_dt(lhysea_sim* s, double dt)
{
    if (!s || !(dt > 0)) return fail(LHYSEA_ERR_ARGUMENT, "bad handle or dt");
    if (s->adj) return adj_busy();
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy((char*)s->ctl + offsetof(Ctl, dt_cur), &dt, sizeof(double), hipMemcpyHostToDevice));
    return LHYSEA_OK;
}

int lhysea_step(lhysea_sim* s, int nsteps)
{
    if (!s || nsteps < 0) return fail(LHYSEA_ERR_ARGUMENT, "bad handle or step count");
    if (s->adj) return adj_busy();
    Ctl c;
    HIPCHK(hipMemcpyAsync(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (c.health) return numeric_fail(c);
    if (!(c.dt_cur > 0) && !s->dt_seq) {
        int rc = lhysea_initial_dt(s, nullptr);
        if (rc) return rc;
    }
    // pieces of at most 1024 steps, the launch split chosen again before each
    for (int done = 0; done < nsteps;) {
        if (done > 0) {
            HIPCHK(hipMemcpyAsync(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s->stream));
            HIPCHK(hipStreamSynchronize(s->stream));
            if (c.health) return numeric_fail(c);
        }
        choose_split(s, c);
        const int n = (nsteps - done < 1024) ? nsteps - done : 1024;
        const int rc = launch_steps(s, n);
        if (rc) return rc;
        done += n;
    }
    return LHYSEA_OK;
}

int lhysea_run_until(lhysea_sim* s, double t_end, int32_t max_steps, int32_t* steps_done)
{
    if (!s || max_steps < 0) return fail(LHYSEA_ERR_ARGUMENT, "bad handle or step budget");
    if (s->adj) return adj_busy();
    Ctl c;
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    if (!(c.dt_cur > 0) && !s->dt_seq) {
        int rc = lhysea_initial_dt(s, nullptr);
        if (rc) return rc;
        HIPCHK(hipMemcpy(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    }
    const int step0 = c.step;
    if (steps_done) *steps_done = 0;
    if (c.health) return numeric_fail(c);
    const bool device_stop = (s->mode == MODE_FUSED);
    while (c.t < t_end && c.step - step0 < max_steps) {
        // chunks sized from the current dt; steps launched after t reaches
        // t_end do nothing (fused kernels), so the host syncs once per chunk
        int chunk = 1;
        if (device_stop) {
            const double est = (c.dt_cur > 0) ? (t_end - c.t) / c.dt_cur : 1.0;
            chunk = (est > 511.0) ? 512 : (est >= 0.0) ? (int)est + 1 : 1;
        }
        if (chunk > max_steps - (c.step - step0)) chunk = max_steps - (c.step - step0);
        choose_split(s, c);
        const int cur0 = s->cur, stepc = c.step;
        int rc = launch_steps(s, chunk, nullptr, nullptr, device_stop ? t_end : INFINITY);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(s->stream));
        HIPCHK(hipMemcpy(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
        // the state is in the buffer the last executed step wrote
        s->cur = cur0 ^ ((c.step - stepc) & 1);
        if (c.health) {
            if (steps_done) *steps_done = c.step - step0;
            return numeric_fail(c);
        }
    }
    if (steps_done) *steps_done = c.step - step0;
    return LHYSEA_OK;
}

int lhysea_health(lhysea_sim* s, uint32_t* flags, int32_t* first_step)
{
    if (!s || !flags) return fail(LHYSEA_ERR_ARGUMENT, "null argument");
    Ctl c;
    HIPCHK(hipMemcpyAsync(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    *flags = c.health;
    if (first_step) *first_step = c.health ? c.health_step : -1;
    return LHYSEA_OK;
}

int lhysea_sync(lhysea_sim* s)
{
    if (!s) return fail(LHYSEA_ERR_ARGUMENT, "null handle");
    HIPCHK(hipStreamSynchronize(s->stream));
    return LHYSEA_OK;
}

int lhysea_get_time(lhysea_sim* s, double* t, double* dt_next, int32_t* steps)
{
    if (!s) return fail(LHYSEA_ERR_ARGUMENT, "null handle");
    Ctl c;
    HIPCHK(hipMemcpyAsync(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (t) *t = c.t;
    if (dt_next) *dt_next = c.dt_cur;
    if (steps) *steps = c.step;
    return LHYSEA_OK;
}

int lhysea_track_eta_max(lhysea_sim* s, int enable)
{
    if (!s) return fail(LHYSEA_ERR_ARGUMENT, "null handle");
    if (s->adj) return adj_busy();
    s->track_emax = enable != 0;
    return LHYSEA_OK;
}

int lhysea_eta_max(lhysea_sim* s, double* val, double* time)
{
    if (!s) return fail(LHYSEA_ERR_ARGUMENT, "null handle");
    std::vector<double> em(2 * s->N);
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(em.data(), s->emax, 2 * s->N * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < s->N; i++) {
        if (val) val[i] = em[2 * i];
        if (time) time[i] = em[2 * i + 1];
    }
    return LHYSEA_OK;
}

int lhysea_set_gauges(lhysea_sim* s, const int32_t* cells, int32_t ng, int32_t max_steps)
{
    if (!s || ng < 0 || max_steps < 0 || (ng > 0 && !cells)) return fail(LHYSEA_ERR_ARGUMENT, "bad gauge arguments");
    if (s->adj) return adj_busy();
    for (int g = 0; g < ng; g++)
        if (cells[g] < 0 || (size_t)cells[g] >= s->N) return fail(LHYSEA_ERR_ARGUMENT, "gauge cell out of range");
    HIPCHK(hipStreamSynchronize(s->stream));
    hipFree(s->gauges);
    hipFree(s->gseries);
    s->gauges = nullptr;
    s->gseries = nullptr;
    s->ng = 0;
    s->gcap = 0;
    if (ng == 0 || max_steps == 0) return LHYSEA_OK;
    HIPCHK(hipMalloc(&s->gauges, ng * sizeof(int)));
    HIPCHK(hipMalloc(&s->gseries, (size_t)ng * max_steps * sizeof(double)));
    HIPCHK(hipMemcpy(s->gauges, cells, ng * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(s->gseries, 0, (size_t)ng * max_steps * sizeof(double)));
    Ctl c;
    HIPCHK(hipMemcpy(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    s->gstart = c.step;
    s->ng = ng;
    s->gcap = max_steps;
    return LHYSEA_OK;
}

int lhysea_get_gauges(lhysea_sim* s, double* out, int32_t* nrec)
{
    if (!s) return fail(LHYSEA_ERR_ARGUMENT, "null handle");
    Ctl c;
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    int n = c.step - s->gstart;
    if (n > s->gcap) n = s->gcap;
    if (n < 0) n = 0;
    if (nrec) *nrec = n;
    if (out && n > 0 && s->ng > 0)
        HIPCHK(hipMemcpy(out, s->gseries, (size_t)n * s->ng * sizeof(double), hipMemcpyDeviceToHost));
    return LHYSEA_OK;
}

int lhysea_dt_history(lhysea_sim* s, double* out, int32_t max_steps, int32_t* nrec)
{
    if (!s || max_steps < 0) return fail(LHYSEA_ERR_ARGUMENT, "bad arguments");
    Ctl c;
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    int n = c.step < s->hcap ? c.step : s->hcap;
    if (n > max_steps) n = max_steps;
    if (nrec) *nrec = n;
    if (out && n > 0) HIPCHK(hipMemcpy(out, s->dt_hist, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    return LHYSEA_OK;
}

int lhysea_set_kernel(lhysea_sim* s, int32_t mode)
{
    if (!s || (mode != MODE_FUSED && mode != MODE_FOURPASS)) return fail(LHYSEA_ERR_ARGUMENT, "bad kernel mode");
    if (s->adj) return adj_busy();
    if (mode == MODE_FOURPASS && s->p.precision != LHYSEA_FP64)
        return fail(LHYSEA_ERR_ARGUMENT, "the four-pass kernels exist in FP64 only");
    // the four-pass path keeps no tile classes: rebuild them on the way back
    if (mode == MODE_FUSED && s->mode != MODE_FUSED) HIPCHK(classify(s, true));
    s->mode = mode;
    return LHYSEA_OK;
}

int lhysea_set_adjoint_kernel(lhysea_sim* s, int32_t mode)
{
    if (!s || (mode != MODE_FUSED && mode != MODE_FOURPASS)) return fail(LHYSEA_ERR_ARGUMENT, "bad kernel mode");
    if (s->adj) return adj_busy();
    s->adj_mode = mode;
    return LHYSEA_OK;
}

int lhysea_set_tile_split(lhysea_sim* s, int32_t mode)
{
    if (!s || mode < -1 || mode > 1) return fail(LHYSEA_ERR_ARGUMENT, "bad handle or split mode");
    if (s->adj) return adj_busy();
    s->split_mode = mode;
    return LHYSEA_OK;
}

int lhysea_tile_classes(lhysea_sim* s, int32_t* tiles, int32_t* general, int32_t* split)
{
    if (!s) return fail(LHYSEA_ERR_ARGUMENT, "null handle");
    Ctl c;
    HIPCHK(hipMemcpyAsync(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (tiles) *tiles = s->ntiles;
    if (general) *general = (int32_t)c.gcount[c.step & 1];
    if (split) *split = s->split;
    return LHYSEA_OK;
}

int lhysea_time_steps(lhysea_sim* s, int32_t nsteps, double* total_ms, double* kernel_ms)
{
    if (!s || nsteps < 1) return fail(LHYSEA_ERR_ARGUMENT, "bad arguments");
    if (s->adj) return adj_busy();
    Ctl c;
    HIPCHK(hipMemcpyAsync(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (!(c.dt_cur > 0)) {
        int rc = lhysea_initial_dt(s, nullptr);
        if (rc) return rc;
    }
    choose_split(s, c);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, s->stream));
    int rc = launch_steps(s, nsteps);
    if (rc) return rc;
    HIPCHK(hipEventRecord(e1, s->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    if (total_ms) *total_ms = ms;
    if (kernel_ms) {
        // events recorded by the launcher immediately around the main step
        // kernel (fused kernel, or the cell-update kernel of the four-pass path
        // is not bracketed: there the whole step is bracketed instead)
        double sum = 0.0;
        for (int n = 0; n < nsteps; n++) {
            if (s->mode == MODE_FUSED) {
                rc = launch_steps(s, 1, e0, e1);
                if (rc) return rc;
            } else {
                HIPCHK(hipEventRecord(e0, s->stream));
                rc = launch_steps(s, 1);
                if (rc) return rc;
                HIPCHK(hipEventRecord(e1, s->stream));
            }
            HIPCHK(hipEventSynchronize(e1));
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            sum += ms;
        }
        *kernel_ms = sum / nsteps;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return LHYSEA_OK;
}

// Workspace of one adjoint run, held by the handle from the forward pass to
// the reverse sweep: the snapshot store, the reverse-step buffers and the
// loss seeds (all in HBM).
struct lhysea_adj_work {
    lhysea_sim* s = nullptr;
    char* snap = nullptr;  // snapshots of the state (or every state, store_all)
    std::vector<char*> slot; // the buffer each snapshot slot holds (stores of the reverse sweep swap buffers)
    void* own[2] = {nullptr, nullptr};  // the handle's state buffers, given back when the workspace goes
    bool swap_store = false; // reverse sweep: a store hands the working buffer to the slot
    size_t state_bytes = 0;
    void* P = nullptr;     // partial sums [3][2][N]
    void* ACC = nullptr;   // [6][N]
    void* AB = nullptr;    // [6][N]
    double* LAM = nullptr;   // [6][N] state adjoint lambda of the current step
    double* LAM2 = nullptr;  // [6][N] second lambda buffer of the fused FP32 reverse step
    double* gmu = nullptr;   // friction adjoint slots: ngmu per component (Coulomb 1, Pouliquen 4)
    size_t ngmu = 0;
    int ngcomp = 1;
    const char* W = nullptr; // stored state of the reverse step (all states kept), else null
    const char* src = nullptr; // restored snapshot that the next advance or reverse step reads in place
    int* gauges = nullptr;
    double* gser = nullptr;  // gauge series of the forward pass
    double* G = nullptr;     // [nsteps][ng] seeds dL/deta
    int ng = 0;
    long nsteps = 0;
    int nsnap = 0;
    bool store_all = false;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    double fwd_ms = 0;       // forward pass time (with the first-descent snapshots)
    const double* saved_seq = nullptr;
    bool saved_emax = false;
};

namespace {

// Frees a workspace and gives the handle back its stepping configuration.
// The gauge buffers belong to the adjoint run; the handle's state must be
// uploaded again before further steps.
void adj_free(lhysea_sim* s, lhysea_adj_work* a)
{
    if (!a) return;
    hipStreamSynchronize(s->stream);
    if (a->own[0]) {
        set_state_buf(s, 0, a->own[0]);
        set_state_buf(s, 1, a->own[1]);
    }
    hipFree(a->snap);
    hipFree(a->P);
    hipFree(a->ACC);
    hipFree(a->AB);
    hipFree(a->LAM);
    hipFree(a->LAM2);
    hipFree(a->gmu);
    hipFree(a->G);
    if (a->e0) hipEventDestroy(a->e0);
    if (a->e1) hipEventDestroy(a->e1);
    if (a->e2) hipEventDestroy(a->e2);
    s->dt_seq = const_cast<double*>(a->saved_seq);
    s->track_emax = a->saved_emax;
    hipFree(a->gser);
    if (a->gauges) hipFree(a->gauges);
    if (s->gauges == a->gauges) s->gauges = nullptr;
    if (s->gseries == a->gser) s->gseries = nullptr;
    s->ng = 0;
    s->gcap = 0;
    delete a;
}

void* state_ptr(lhysea_sim* s) { return s->p.precision == LHYSEA_FP64 ? (void*)s->b64.W[s->cur] : (void*)s->b32.W[s->cur]; }

hipError_t hook_advance(void* c, int k0, int n)
{
    lhysea_adj_work* a = (lhysea_adj_work*)c;
    lhysea_sim* s = a->s;
    hipError_t e = adj_set_step(s->ctl, k0, s->stream);
    s->step_host = k0;
    // the restored state's tile classes (the launch split of the forward pass is kept)
    if (e == hipSuccess) e = classify(s, false, a->src);
    if (e != hipSuccess) return e;
    if (a->src && n > 0) {
        // the first step reads the restored snapshot in place; the snapshot
        // stays intact (the step writes the other state buffer)
        const int nxt = 1 - s->cur;
        e = launch_one(s, a->src, state_buf(s, nxt), launch_args(s, nullptr, nullptr));
        if (e != hipSuccess) return e;
        s->cur = nxt;
        a->src = nullptr;
        n--;
    }
    return launch_steps(s, n) == LHYSEA_OK ? hipSuccess : hipErrorUnknown;
}

// In the reverse sweep a store of the working state copies nothing either:
// the working buffer becomes the slot's and the slot's old buffer the
// working one (the schedule restores the slot next, and the advance after a
// restore rebuilds the tile classes and writes every plane of its first two
// steps, so the old contents are never read).  The forward pass copies: it
// steps on from the stored state.
hipError_t hook_store(void* c, int slot)
{
    lhysea_adj_work* a = (lhysea_adj_work*)c;
    lhysea_sim* s = a->s;
    char* dst = a->slot[slot];
    const void* from = a->src ? (const void*)a->src : state_ptr(s);
    if (from == dst) return hipSuccess;
    if (a->swap_store && !a->src) {
        a->slot[slot] = (char*)state_ptr(s);
        set_state_buf(s, s->cur, dst);
        return hipSuccess;
    }
    return hipMemcpyAsync(dst, from, a->state_bytes, hipMemcpyDeviceToDevice, s->stream);
}

// A restore copies nothing: the next advance or reverse step reads the
// snapshot slot in place.
hipError_t hook_restore(void* c, int slot)
{
    lhysea_adj_work* a = (lhysea_adj_work*)c;
    a->src = a->slot[slot];
    return hipSuccess;
}

hipError_t hook_reverse(void* c, int k)
{
    lhysea_adj_work* a = (lhysea_adj_work*)c;
    lhysea_sim* s = a->s;
    hipError_t e;
    // the state W^k: its stored slot (every state kept, or a restored
    // snapshot), else the working buffer
    const void* W = a->W ? (const void*)a->W : a->src ? (const void*)a->src : state_ptr(s);
    // the seeds of the gauge series at step k (rows of G: after step k) join
    // lambda^k as it is formed
    const double* Gk = (k >= 1 && a->ng > 0) ? a->G + (size_t)(k - 1) * a->ng : nullptr;
    if (s->p.precision == LHYSEA_FP64) {
        e = adj_reverse_step_f64(s->k, (const double*)W, s->b64.H, (double*)a->P, (double*)a->ACC, (double*)a->AB,
                                 a->LAM, a->gmu, a->ngmu, s->dt_hist, k, s->stream);
        if (e == hipSuccess && Gk) e = adj_inject(a->LAM, s->N, a->gauges, a->ng, Gk, s->stream);
    } else if (s->adj_mode == MODE_FOURPASS) {
        e = adj_reverse_step_f32(s->k, (const float*)W, s->b32.H, (float*)a->P, (float*)a->ACC, (float*)a->AB, a->LAM,
                                 a->gmu, a->ngmu, s->dt_hist, k, s->stream);
        if (e == hipSuccess && Gk) e = adj_inject(a->LAM, s->N, a->gauges, a->ng, Gk, s->stream);
    } else {
        // lambda^{n+1} in LAM, lambda^n (with the seeds) into LAM2, then swap
        e = adj_reverse_step_f32_fused(s->k, (const float*)W, s->b32.H, (float*)a->P, (float*)a->ACC, a->LAM,
                                       a->LAM2, a->gmu, a->ngmu, s->dt_hist, k, a->gauges, Gk ? a->ng : 0, Gk,
                                       s->stream);
        std::swap(a->LAM, a->LAM2);
    }
    return e;
}

}  // namespace

static void adj_release(lhysea_sim* s)
{
    lhysea_adj_work* a = s->adj;
    s->adj = nullptr;
    adj_free(s, a);
}

#define ADJCHK(expr)                                          \
    do {                                                      \
        hipError_t e_ = (expr);                               \
        if (e_ != hipSuccess) return done(hip_fail(e_, #expr)); \
    } while (0)

int lhysea_adjoint_forward(lhysea_sim* s, int32_t nsteps, int32_t nsnap, const int32_t* gauges, int32_t ng,
                           double* eta)
{
    if (!s || nsteps < 1 || nsnap < 1 || ng < 0 || (ng > 0 && !gauges))
        return fail(LHYSEA_ERR_ARGUMENT, "bad adjoint arguments");
    if (nsteps > s->hcap) return fail(LHYSEA_ERR_ARGUMENT, "too many steps for the dt history");
    for (int g = 0; g < ng; g++)
        if (gauges[g] < 0 || (size_t)gauges[g] >= s->N) return fail(LHYSEA_ERR_ARGUMENT, "gauge cell out of range");
    // a forward pass whose reverse sweep never came is discarded
    adj_release(s);
    const size_t N = s->N;
    const bool f64 = s->p.precision == LHYSEA_FP64;
    const size_t fsz = f64 ? sizeof(double) : sizeof(float);
    lhysea_adj_work* a = new (std::nothrow) lhysea_adj_work{};
    if (!a) return fail(LHYSEA_ERR_HOST_MEMORY, "adjoint workspace");
    a->s = s;
    a->state_bytes = 48 * N;       // 6 doubles or 12 floats per cell
    a->ng = ng;
    a->nsteps = nsteps;
    a->nsnap = nsnap;
    a->saved_seq = s->dt_seq;
    a->saved_emax = s->track_emax;
    auto done = [&](int code) {
        adj_free(s, a);
        return code;
    };
    // with a snapshot per state (nsnap > nsteps) every state stays in HBM:
    // the forward writes step k's result straight into slot k + 1 and the
    // reverse sweep reads slot k in place (no re-advance, no copy)
    a->store_all = nsnap > nsteps;
    const size_t nslots = a->store_all ? (size_t)nsteps + 1 : (size_t)nsnap;
    a->ngmu = adj_gmu_slots(s->k);
    a->ngcomp = s->p.friction_law == 1 ? 4 : 1;   // Pouliquen: the four angles' tangents
    if (hipMalloc(&a->snap, nslots * a->state_bytes) != hipSuccess || hipMalloc(&a->P, 6 * N * fsz) != hipSuccess ||
        hipMalloc(&a->ACC, 6 * N * fsz) != hipSuccess || hipMalloc(&a->AB, 6 * N * fsz) != hipSuccess ||
        hipMalloc(&a->LAM, 6 * N * sizeof(double)) != hipSuccess ||
        (!f64 && hipMalloc(&a->LAM2, 6 * N * sizeof(double)) != hipSuccess) ||
        hipMalloc(&a->gmu, a->ngcomp * a->ngmu * sizeof(double)) != hipSuccess ||
        hipMalloc(&a->G, (size_t)nsteps * (ng > 0 ? ng : 1) * sizeof(double)) != hipSuccess)
        return done(fail(LHYSEA_ERR_GPU_MEMORY, "adjoint workspace allocation failed"));
    a->slot.resize(nslots);
    for (size_t i = 0; i < nslots; i++) a->slot[i] = a->snap + i * a->state_bytes;
    a->own[0] = state_buf(s, 0);
    a->own[1] = state_buf(s, 1);
    ADJCHK(hipEventCreate(&a->e0));
    ADJCHK(hipEventCreate(&a->e1));
    ADJCHK(hipEventCreate(&a->e2));

    // forward pass from the uploaded state: dt history, gauge series and the
    // snapshots of the first descent of the checkpoint tree
    int rc = lhysea_set_gauges(s, gauges, ng, nsteps);
    if (rc) return done(rc);
    a->gauges = s->gauges;
    a->gser = s->gseries;
    s->track_emax = false;
    s->dt_seq = nullptr;
    Ctl c0;
    ADJCHK(hipMemcpy(&c0, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    if (c0.step != 0) return done(fail(LHYSEA_ERR_ARGUMENT, "the adjoint starts from a freshly uploaded state"));
    if (!(c0.dt_cur > 0)) {
        rc = lhysea_initial_dt(s, nullptr);
        if (rc) return done(rc);
    }
    ADJCHK(hipEventRecord(a->e0, s->stream));
    ADJCHK(hook_store(a, 0));
    if (a->store_all) {
        StepArgs sa = launch_args(s, nullptr, nullptr);
        sa.wn_prev = 0;    // each step writes a fresh slot
        for (long k = 0; k < nsteps; k++)
            ADJCHK(launch_one(s, a->snap + (size_t)k * a->state_bytes, a->snap + (size_t)(k + 1) * a->state_bytes, sa));
    } else {
        std::vector<std::pair<long, int>> chain;
        revolve_chain(nsteps, nsnap, chain);
        long pos = 0;
        for (auto& ch : chain) {
            rc = launch_steps(s, (int)(ch.first - pos));
            if (rc) return done(rc);
            ADJCHK(hook_store(a, ch.second));
            pos = ch.first;
        }
        rc = launch_steps(s, (int)(nsteps - pos));
        if (rc) return done(rc);
    }
    ADJCHK(hipEventRecord(a->e1, s->stream));
    ADJCHK(hipStreamSynchronize(s->stream));
    ADJCHK(hipGetLastError());
    float f_ms = 0;
    (void)hipEventElapsedTime(&f_ms, a->e0, a->e1);   // a timing statistic: 0 if unavailable
    a->fwd_ms = f_ms;
    if (eta && ng > 0) {
        int32_t nrec = 0;
        rc = lhysea_get_gauges(s, eta, &nrec);
        if (rc) return done(rc);
    }
    // gauges off for the re-advances; the dt history is frozen from here on
    s->gseries = nullptr;
    s->ng = 0;
    s->dt_seq = s->dt_hist;
    s->adj = a;
    return LHYSEA_OK;
}

}  // extern "C"

// The reverse sweep of the pending adjoint workspace, which it consumes
// whatever happens: dL/dW0 into grad_W0 (may be null) and the adjoint of the
// friction coefficient, fr[0] = dL/dmu (Coulomb) or fr[i] = dL/dtan(angle
// i+1), i = 0..3 (Pouliquen).
static int adjoint_sweep(lhysea_sim* s, const double* seeds, double* grad_W0, double fr[4], double* stats)
{
    if (!s) return fail(LHYSEA_ERR_ARGUMENT, "null handle");
    if (!s->adj) return fail(LHYSEA_ERR_ARGUMENT, "no adjoint forward pass to reverse");
    if (s->adj->ng > 0 && !seeds) return fail(LHYSEA_ERR_ARGUMENT, "missing seeds");
    // the workspace goes whatever happens below
    lhysea_adj_work* a = s->adj;
    s->adj = nullptr;
    auto done = [&](int code) {
        adj_free(s, a);
        return code;
    };
    const size_t N = s->N;
    const long nsteps = a->nsteps;
    const int ng = a->ng;
    if (ng > 0) ADJCHK(hipMemcpy(a->G, seeds, (size_t)nsteps * ng * sizeof(double), hipMemcpyHostToDevice));
    ADJCHK(hipEventRecord(a->e1, s->stream));
    ADJCHK(hipMemsetAsync(a->LAM, 0, 6 * N * sizeof(double), s->stream));
    ADJCHK(hipMemsetAsync(a->gmu, 0, a->ngcomp * a->ngmu * sizeof(double), s->stream));
    ADJCHK(hook_restore(a, 0));
    ADJCHK(adj_inject(a->LAM, N, a->gauges, ng, a->G + (size_t)(nsteps - 1) * ng, s->stream));

    // reverse sweep with binomial checkpointing (or over the stored states)
    AdjointHooks h{hook_advance, hook_store, hook_restore, hook_reverse, a, 0, 0};
    hipError_t he = hipSuccess;
    if (a->store_all) {
        for (long k = nsteps - 1; k >= 0 && he == hipSuccess; k--) {
            a->W = a->snap + (size_t)k * a->state_bytes;
            he = hook_reverse(a, (int)k);
            h.reverses++;
        }
        a->W = nullptr;
    } else {
        a->swap_store = true;
        he = run_treeverse(h, a->nsnap, nsteps);
    }
    if (he != hipSuccess) return done(hip_fail(he, "adjoint sweep"));
    ADJCHK(hipEventRecord(a->e2, s->stream));
    ADJCHK(hipStreamSynchronize(s->stream));
    ADJCHK(hipGetLastError());

    if (grad_W0) ADJCHK(hipMemcpy(grad_W0, a->LAM, 6 * N * sizeof(double), hipMemcpyDeviceToHost));
    // slot sums in slot order, per component
    std::vector<double> slots((size_t)a->ngcomp * a->ngmu);
    ADJCHK(hipMemcpy(slots.data(), a->gmu, slots.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < 4; i++) {
        double v = 0.0;
        if (i < a->ngcomp)
            for (size_t k = 0; k < a->ngmu; k++) v += slots[(size_t)i * a->ngmu + k];
        fr[i] = v;
    }
    if (stats) {
        float r_ms = 0;
        (void)hipEventElapsedTime(&r_ms, a->e1, a->e2);   // a timing statistic: 0 if unavailable
        stats[0] = a->fwd_ms;                  // forward pass (with first-descent snapshots)
        stats[1] = r_ms;                       // reverse sweep (re-advances + reverse steps)
        stats[2] = (double)h.advances;         // re-advanced forward steps
        stats[3] = (double)h.reverses;         // reverse steps
        stats[4] = (double)a->nsnap;
        stats[5] = (double)revolve_advances(nsteps, a->nsnap);
    }
    return done(LHYSEA_OK);
}

// dL/d(angle i+1) from the friction-coefficient adjoint of adjoint_sweep()
static double angle_gradient(const lhysea_params& p, int i, const double fr[4])
{
    const double ang[4] = {p.angle1, p.angle2, p.angle3, p.angle4};
    const double t = std::tan(ang[i]);
    if (p.friction_law == 1) return fr[i] * (1.0 + t * t);   // d tan(a) / da = 1 + tan^2 a
    if (i > 0) return 0.0;
    // mu = |tan a| L/H  ->  dmu/da = sign(tan a) (1 + tan^2 a) L/H
    return fr[0] * (t >= 0 ? 1.0 : -1.0) * (1.0 + t * t) * (p.L / p.H);
}

extern "C" {

int lhysea_adjoint_backward(lhysea_sim* s, const double* seeds, double* grad_W0, double* grad_mu,
                            double* grad_angle1, double* stats)
{
    if (s && s->p.friction_law == 1 && grad_mu)
        return fail(LHYSEA_ERR_ARGUMENT, "dL/dmu belongs to the Coulomb law; under the Pouliquen law "
                                         "lhysea_adjoint_backward_angles gives the four angles' gradients");
    double fr[4];
    const int rc = adjoint_sweep(s, seeds, grad_W0, fr, stats);
    if (rc) return rc;
    if (grad_mu) *grad_mu = fr[0];
    if (grad_angle1) *grad_angle1 = angle_gradient(s->p, 0, fr);
    return LHYSEA_OK;
}

int lhysea_adjoint_backward_angles(lhysea_sim* s, const double* seeds, double* grad_W0, double* grad_angles,
                                   double* stats)
{
    double fr[4];
    const int rc = adjoint_sweep(s, seeds, grad_W0, fr, stats);
    if (rc) return rc;
    if (grad_angles)
        for (int i = 0; i < 4; i++) grad_angles[i] = angle_gradient(s->p, i, fr);
    return LHYSEA_OK;
}

int lhysea_adjoint(lhysea_sim* s, int32_t nsteps, int32_t nsnap, const int32_t* gauges, int32_t ng,
                   const double* obs, double* loss, double* grad_W0, double* grad_mu, double* grad_angle1,
                   double* stats)
{
    if (!s || nsteps < 1 || ng < 0 || (ng > 0 && (!gauges || !obs)))
        return fail(LHYSEA_ERR_ARGUMENT, "bad adjoint arguments");
    // gauge misfit L = sum (eta - obs)^2 and its seeds dL/deta = 2 (eta - obs)
    std::vector<double> G((size_t)nsteps * ng);
    int rc = lhysea_adjoint_forward(s, nsteps, nsnap, gauges, ng, G.data());
    if (rc) return rc;
    double L = 0.0;
    for (size_t i = 0; i < G.size(); i++) {
        const double d = G[i] - obs[i];
        L += d * d;
        G[i] = 2.0 * d;
    }
    if (loss) *loss = L;
    return lhysea_adjoint_backward(s, G.data(), grad_W0, grad_mu, grad_angle1, stats);
}

int lhysea_shallow_water(const lhysea_params* p, double* W, const double* H, double t_end, double* eta_max,
                         double* eta_max_time, double* elapsed_s, int32_t* steps_done)
{
    lhysea_sim* s = nullptr;
    int rc = lhysea_create(p, &s);
    if (rc) return rc;
    rc = lhysea_upload(s, W, H);
    if (!rc) rc = lhysea_track_eta_max(s, 1);
    if (!rc) rc = lhysea_initial_dt(s, nullptr);
    int32_t n = 0;
    auto t0 = std::chrono::steady_clock::now();
    // while (t < t_end) (src/GPU/ShallowWater.cu:269), stepped on the device
    if (!rc) rc = lhysea_run_until(s, t_end, INT32_MAX, &n);
    auto t1 = std::chrono::steady_clock::now();
    if (elapsed_s) *elapsed_s = std::chrono::duration<double>(t1 - t0).count();
    if (steps_done) *steps_done = n;
    if (!rc) rc = lhysea_download(s, W);
    if (!rc && (eta_max || eta_max_time)) rc = lhysea_eta_max(s, eta_max, eta_max_time);
    lhysea_destroy(s);
    return rc;
}

}  // extern "C"
