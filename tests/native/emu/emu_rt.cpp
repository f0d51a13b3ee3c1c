#include <ucontext.h>
#include <algorithm>
#include <hip/hip_runtime.h>
EmuIdx threadIdx, blockIdx;
dim3 blockDim, gridDim;
std::atomic<int> emu_or_flag{0};
std::vector<double> emu_shfl_d;
std::mutex emu_atomic_mu;
struct EmuNode { dim3 g, b; std::function<void()> body; };
typedef std::vector<EmuNode> EmuGraph;
static EmuGraph* emu_capture = nullptr;
long emu_graph_launches = 0;
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { emu_capture = new EmuGraph; return 0; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) { *g = emu_capture; emu_capture = nullptr; return 0; }
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, char*, size_t) { *e = new EmuGraph(*(EmuGraph*)g); return 0; }
hipError_t hipGraphDestroy(hipGraph_t g) { delete (EmuGraph*)g; return 0; }
hipError_t hipGraphExecDestroy(hipGraphExec_t g) { delete (EmuGraph*)g; return 0; }
static void emu_run(dim3 grid, dim3 block, const std::function<void()>& body);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    emu_graph_launches++;
    for (auto& n : *(EmuGraph*)e) emu_run(n.g, n.b, n.body);
    return 0;
}
void emu_launch(dim3 grid, dim3 block, const std::function<void()>& body) {
    if (emu_capture) { emu_capture->push_back({grid, block, body}); return; }
    emu_run(grid, block, body);
}
// Lanes as fibers: a lane runs until it waits at a barrier or returns; the
// scheduler passes over the lanes of the workgroup until every one returned.
struct EmuBar { unsigned expected, arrived, gen; };
static EmuBar emu_blk;
static std::vector<EmuBar> emu_wav;
static std::vector<ucontext_t> emu_fctx;
static std::vector<char> emu_done;
static ucontext_t emu_sched;
static char* emu_stacks = nullptr;
static size_t emu_nstacks = 0;
static constexpr size_t kEmuStack = 256 * 1024;
static unsigned emu_cur;
static const std::function<void()>* emu_body;

static void emu_release(EmuBar& b) {
    if (b.arrived > 0 && b.arrived >= b.expected) { b.arrived = 0; b.gen++; }
}
static void emu_wait(EmuBar& b) {
    const unsigned g = b.gen;
    b.arrived++;
    emu_release(b);
    while (b.gen == g) swapcontext(&emu_fctx[emu_cur], &emu_sched);
}
void emu_block_sync() { emu_wait(emu_blk); }
void emu_wave_sync() { emu_wait(emu_wav[emu_cur / 64]); }
static void emu_fiber() {
    (*emu_body)();
    emu_done[emu_cur] = 1;
    // a lane that returned holds no barrier up
    emu_blk.expected--; emu_release(emu_blk);
    EmuBar& w = emu_wav[emu_cur / 64];
    w.expected--; emu_release(w);
}
void* emu_device_alloc(size_t n) {
    static const size_t cap = [] {
        const char* c = std::getenv("EMU_MEM_CAP");
        return c ? (size_t)std::strtoull(c, nullptr, 10) : (size_t)-1;
    }();
    if (n > cap) return nullptr;
    return calloc(n ? n : 1, 1);
}
static void emu_run(dim3 grid, dim3 block, const std::function<void()>& body) {
    gridDim = grid; blockDim = block;
    const unsigned nt = block.x * block.y * block.z;
    emu_shfl_d.assign(nt, 0.0);
    emu_fctx.resize(nt);
    emu_done.resize(nt);
    emu_wav.resize((nt + 63) / 64);
    if (emu_nstacks < nt) {
        free(emu_stacks);
        emu_stacks = (char*)malloc(nt * kEmuStack);
        emu_nstacks = nt;
    }
    emu_body = &body;
    // workgroups in index order, or from the last one down when EMU_ORDER=reverse
    static const bool reverse = [] {
        const char* o = std::getenv("EMU_ORDER");
        return o && std::strcmp(o, "reverse") == 0;
    }();
    const unsigned nwg = grid.x * grid.y;
    for (unsigned q = 0; q < nwg; q++) {
        const unsigned wg = reverse ? nwg - 1 - q : q;
        const unsigned bx = wg % grid.x, by = wg / grid.x;
        blockIdx = {bx, by, 0};
        emu_blk = {nt, 0, 0};
        for (unsigned w = 0; w < emu_wav.size(); w++) emu_wav[w] = {std::min(64u, nt - 64 * w), 0, 0};
        for (unsigned t = 0; t < nt; t++) {
            emu_done[t] = 0;
            getcontext(&emu_fctx[t]);
            emu_fctx[t].uc_stack.ss_sp = emu_stacks + t * kEmuStack;
            emu_fctx[t].uc_stack.ss_size = kEmuStack;
            emu_fctx[t].uc_link = &emu_sched;
            makecontext(&emu_fctx[t], emu_fiber, 0);
        }
        unsigned left = nt;
        while (left) {
            for (unsigned t = 0; t < nt; t++) {
                if (emu_done[t]) continue;
                emu_cur = t;
                threadIdx = {t % block.x, (t / block.x) % block.y, 0};
                swapcontext(&emu_sched, &emu_fctx[t]);
                if (emu_done[t]) left--;
            }
        }
    }
}
