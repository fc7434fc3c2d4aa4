// The extern "C" boundary (include/hyperflow.h): argument checks, host/device
// staging, all-or-nothing commit, lowest-failing-index reporting and the
// dispatch to the kernels compiled for the mechanism's topology.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "host_thermo.hpp"
#include "hyperflow.h"
#include "mechanism.hpp"
#include "topo_ops.hpp"

using namespace hf;

// Scratch of hf_flux_stage_noalloc: the failure word and counters, the box
// table, the face-flux descriptors and the tile table of one call, carved at
// fixed offsets. Every GPU the library runs on holds one copy (1.4 MB) until
// the process ends; calls that overlap on one GPU share it and race.
constexpr int kNoAllocMaxBoxes = 4096;
constexpr int kNoAllocMaxTiles = 65536;
constexpr size_t kNoAllocBoxOff = 256;
constexpr size_t kNoAllocFluxOff = kNoAllocBoxOff + sizeof(hf_box) * kNoAllocMaxBoxes;
constexpr size_t kNoAllocTileOff = kNoAllocFluxOff + sizeof(hf_face_fluxes) * kNoAllocMaxBoxes;
constexpr size_t kNoAllocBytes = kNoAllocTileOff + sizeof(int32_t) * 3 * kNoAllocMaxTiles;
static_assert(kNoAllocFluxOff % 256 == 0 && kNoAllocTileOff % 256 == 0, "scratch regions stay 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

thread_local std::string g_create_error;

// Device address of `scratch` per device, taken once with that device
// current (each device holds its own copy at its own address). Host code
// uses these addresses, never the array's name.
constexpr int kMaxDevices = 64;
std::atomic<void*> g_scratch_ptr[kMaxDevices];

int scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return HF_EDEVICE;
    if (g_scratch_ptr[device].load()) return HF_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        (void)hipGetLastError();
        return HF_EDEVICE;
    }
    g_scratch_ptr[device].store(p);
    return HF_OK;
}

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    ~DevBuf() { release(); }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
    }
    bool ensure(size_t bytes) {
        if (bytes <= n) return true;
        release();
        if (hipMalloc(&p, bytes) != hipSuccess) {
            p = nullptr;
            return false;
        }
        n = bytes;
        return true;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// True when doubles [lo, hi] relative to p (element indices, inclusive) lie
// inside the device allocation that holds p (hipMemGetAddressRange). A
// pointer the runtime cannot place passes: host memory is staged and
// checked by the copies themselves.
bool within_allocation(const double* p, int64_t lo, int64_t hi) {
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<double*>(p)) != hipSuccess || !base) {
        (void)hipGetLastError();
        return true;
    }
    const intptr_t b = (intptr_t)base, a0 = (intptr_t)p + (intptr_t)lo * 8, a1 = (intptr_t)p + (intptr_t)hi * 8 + 8;
    return a0 >= b && a1 <= b + (intptr_t)size;
}

bool is_device_ptr(const void* ptr) {
    if (!ptr) return false;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, ptr) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

}  // namespace

// Online tuning of the batched form (every hf_options tuning value left at
// 0): coordinate descent over the refill group, the resident waves per CU,
// the stiffness bins and the Jacobian / setup grouping waits. The first call
// of a size class (cell counts within a factor of two) runs the defaults;
// each following call tries one candidate of one parameter, and a candidate
// faster per cell than the best so far is kept. After every candidate has
// run, the best configuration serves the size class. Each cell's result is
// independent of every one of these values, so tuning never changes a
// result.
struct ChemTuner {
    enum { CHUNK, WAVES, BINS, JWAIT, SWAIT, NPARAM };
    static constexpr int kCand = 3;
    // candidate values per parameter, the default first
    static constexpr int kLane[NPARAM][kCand] = {{8, 4, 16}, {4, 2, 8}, {64, 32, 128}, {16, 8, 32}, {2, 1, 4}};
    static constexpr int kWave[NPARAM][kCand] = {{8, 8, 8}, {8, 4, 16}, {32, 16, 64}, {16, 16, 16}, {2, 2, 2}};
    int64_t size = 0;
    bool wave = false;               // one cell per wave (large mechanisms): only waves and bins matter
    int best[NPARAM] = {};
    double best_t = 0;
    int param = 0, cand = 0;         // the trial this call runs (cand 0: the defaults, first call)
    bool done = false;
    const int (*table())[kCand] { return wave ? kWave : kLane; }
    bool tunable(int p) const { return !wave || p == WAVES || p == BINS; }
    // the configuration of this call
    void pick(int64_t ncell, bool wave_form, int* cfg) {
        if (size == 0 || ncell > 2 * size || 2 * ncell < size || wave_form != wave) {
            size = ncell;
            wave = wave_form;
            for (int p = 0; p < NPARAM; ++p) best[p] = table()[p][0];
            best_t = 0;
            param = 0;
            cand = 0;
            done = false;
        }
        for (int p = 0; p < NPARAM; ++p) cfg[p] = best[p];
        if (!done && cand > 0) cfg[param] = table()[param][cand];
    }
    void advance() {
        for (;;) {
            if (++cand >= kCand) {
                cand = 1;
                if (++param >= NPARAM) {
                    done = true;
                    return;
                }
            }
            if (tunable(param)) return;
        }
    }
    void record(double ms, int64_t ncell) {
        if (done || ncell <= 0) return;
        const double t = ms / (double)ncell;
        if (cand == 0) {
            best_t = t;
            param = 0;
        } else if (t < best_t) {
            best_t = t;
            best[param] = table()[param][cand];
        }
        advance();
    }
};

// Online choice of the 2-D flux tile shape: the first calls of a size class
// (cell counts within a factor of two) run each shape kRounds times in turn,
// and the shape with the fastest of its runs per cell serves the following
// calls (two runs each keep a single slow launch from deciding).
struct FluxTuner {
    static constexpr int kRounds = 2;
    int64_t size = 0;
    int calls = 0, best = 0, n = 1;
    double per_cell[3] = {};
    int pick(int64_t cells, int nshape) {
        if (size == 0 || cells > 2 * size || 2 * cells < size || nshape != n) {
            size = cells;
            n = nshape;
            calls = 0;
            best = 0;
        }
        return calls < kRounds * n ? calls % n : best;
    }
    void record(double ms, int64_t cells) {
        if (calls >= kRounds * n || cells <= 0 || !(ms > 0)) return;
        const int k = calls % n;
        const double t = ms / (double)cells;
        per_cell[k] = calls < n ? t : std::min(per_cell[k], t);
        if (++calls == kRounds * n) {
            best = 0;
            for (int j = 1; j < n; ++j)
                if (per_cell[j] < per_cell[best]) best = j;
        }
    }
};

// One captured chemistry call (a HIP graph of its kernels), replayed while a
// device-resident call repeats the same arguments and buffers.
struct ChemGraph {
    std::vector<uint64_t> key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        key.clear();
    }
    ~ChemGraph() {
        reset();
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
    }
};

struct hf_ctx {
    Mechanism mech;
    ChemTuner tuner;
    FluxTuner ftuner;
    ChemGraph graph;
    const TopoOps* ops = nullptr;
    hf_options opt{};
    int num_cus = 256;
    hipStream_t stream = nullptr;
    void* d_mech = nullptr;
    std::string err;
    std::atomic<bool> busy{false};
    DevBuf U, out, dt, status, scal, work, boxes, tiles, blocks, aux[6];
    // hf_box_strang_step: its captured graph and the tables and scalars it reads
    ChemGraph step_graph;
    DevBuf sbox, stiles, scells, sscal;
    // hf_amr_strang_step: its captured graph and the device tables it reads
    ChemGraph amr_graph;
    DevBuf atab;
};

namespace {

struct BusyGuard {
    hf_ctx* c;
    bool ok;
    explicit BusyGuard(hf_ctx* ctx) : c(ctx) {
        bool expected = false;
        ok = c->busy.compare_exchange_strong(expected, true);
    }
    ~BusyGuard() {
        if (ok) c->busy.store(false);
    }
};

int fail(hf_ctx* c, int code, const std::string& msg) {
    c->err = msg;
    return code;
}

// Bind the context's device: stream, CU count and the mechanism tables.
int ensure_device(hf_ctx* c) {
    if (c->d_mech) return hipSetDevice(c->opt.device) == hipSuccess ? HF_OK : fail(c, HF_EDEVICE, "hipSetDevice failed");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= c->opt.device) {
        (void)hipGetLastError();
        return fail(c, HF_EDEVICE, "no HIP device " + std::to_string(c->opt.device));
    }
    if (hipSetDevice(c->opt.device) != hipSuccess) return fail(c, HF_EDEVICE, "hipSetDevice failed");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->opt.device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    if (!c->stream && hipStreamCreate(&c->stream) != hipSuccess) return fail(c, HF_EDEVICE, "hipStreamCreate failed");
    std::vector<char> host(c->ops->mech_bytes);
    c->ops->fill(c->mech, host.data());
    void* d = nullptr;
    if (hipMalloc(&d, host.size()) != hipSuccess) return fail(c, HF_ENOMEM, "cannot allocate mechanism tables");
    if (hipMemcpy(d, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return fail(c, HF_EDEVICE, "cannot upload mechanism tables");
    }
    c->d_mech = d;
    // the static scratch of the no-allocation entry point, once per device
    if (scratch_setup(c->opt.device) != HF_OK) return fail(c, HF_EDEVICE, "hipGetSymbolAddress(scratch) failed");
    return HF_OK;
}

#define HF_DEVICE(c)                       \
    do {                                   \
        const int r_ = ensure_device(c);   \
        if (r_ != HF_OK) return r_;        \
    } while (0)

hf_options defaults() {
    hf_options o{};
    o.rtol = 1e-8;
    o.atol = 1e-15;
    o.max_steps = 5000;
    o.precision = HF_PREC_FP64;
    o.device = 0;
    return o;
}

int hip_fail(hf_ctx* c, hipError_t e, const char* what) {
    return fail(c, HF_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HF_HIP(c, call)                                   \
    do {                                                  \
        hipError_t e_ = (call);                           \
        if (e_ != hipSuccess) return hip_fail(c, e_, #call); \
    } while (0)

}  // namespace

extern "C" {

const char* hf_create_error(void) { return g_create_error.c_str(); }

int hf_create(const char* mechanism_yaml, const hf_options* opt, hf_ctx** out) {
    if (!out || !mechanism_yaml) {
        g_create_error = "null argument";
        return HF_EINVAL;
    }
    *out = nullptr;
    auto c = std::make_unique<hf_ctx>();
    c->opt = opt ? *opt : defaults();
    if (!(c->opt.rtol > 0) || !(c->opt.atol > 0) || c->opt.max_steps < 1) {
        g_create_error = "invalid tolerances or step limit";
        return HF_EINVAL;
    }
    try {
        c->mech = load_mechanism(mechanism_yaml);
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return HF_EINVAL;
    }
    for (auto& r : c->mech.reactions)
        if (!(r.A > 0) || (r.kind == Reaction::FALLOFF && !(r.A0 > 0))) {
            g_create_error = "reaction '" + r.equation + "' has a non-positive pre-exponential factor";
            return HF_EINVAL;
        }
    const std::string topo = c->mech.topology();
    for (const TopoOps* const* t = compiled_topologies(); *t; ++t)
        if (topo == (*t)->topology) c->ops = *t;
    if (!c->ops) {
        g_create_error = "mechanism '" + std::string(mechanism_yaml) +
                         "' has no compiled kernels (its species/reaction topology matches no built mechanism)";
        return HF_EINVAL;
    }
    if (c->opt.device < 0) {
        g_create_error = "invalid device ordinal";
        return HF_EINVAL;
    }
    // The device is bound here when it exists (stream, mechanism tables, the
    // no-allocation scratch address), else on the first call that needs it,
    // so host-only set-up helpers work on a machine without a GPU.
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > c->opt.device) {
        if (ensure_device(c.get()) != HF_OK) {
            g_create_error = c->err;
            hf_destroy(c.release());
            return HF_EDEVICE;
        }
    } else {
        (void)hipGetLastError();
    }
    *out = c.release();
    return HF_OK;
}

void hf_destroy(hf_ctx* c) {
    if (!c) return;
    if (c->d_mech || c->stream) {
        (void)hipSetDevice(c->opt.device);
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        c->graph.reset();
        c->step_graph.reset();
        if (c->d_mech) (void)hipFree(c->d_mech);
        if (c->stream) (void)hipStreamDestroy(c->stream);
    }
    delete c;
}

const char* hf_last_error(const hf_ctx* c) { return c ? c->err.c_str() : "null context"; }
int hf_num_species(const hf_ctx* c) { return c ? c->mech.ns() : 0; }
int hf_num_reactions(const hf_ctx* c) { return c ? c->mech.nr() : 0; }
const char* hf_species_name(const hf_ctx* c, int k) {
    return (c && k >= 0 && k < c->mech.ns()) ? c->mech.species[k].c_str() : nullptr;
}
int hf_molecular_weights(const hf_ctx* c, double* W) {
    if (!c || !W) return HF_EINVAL;
    for (int k = 0; k < c->mech.ns(); ++k) W[k] = c->mech.W[k];
    return HF_OK;
}
int hf_set_options(hf_ctx* c, const hf_options* o) {
    if (!c || !o) return HF_EINVAL;
    if (!(o->rtol > 0) || !(o->atol > 0) || o->max_steps < 1) return fail(c, HF_EINVAL, "invalid tolerances or step limit");
    if (o->device != c->opt.device) return fail(c, HF_EINVAL, "the device of a context cannot change");
    c->opt = *o;
    return HF_OK;
}
int hf_get_options(const hf_ctx* c, hf_options* o) {
    if (!c || !o) return HF_EINVAL;
    *o = c->opt;
    return HF_OK;
}

namespace {

// Replays the cached graph when the launch repeats the cached arguments
// (pointers, sizes, tolerances and tuning values), otherwise captures the
// kernels of this launch into a new graph first. The graph holds kernel
// nodes only (no memset or memcpy nodes); *ms is the device time of the
// whole graph, measured with events around its launch.
int chem_graph_launch(hf_ctx* c, ChemLaunch L, float* ms, int* how) {
    auto bits = [](double x) {
        uint64_t u;
        std::memcpy(&u, &x, sizeof(u));
        return u;
    };
    const std::vector<uint64_t> key = {
        (uint64_t)(uintptr_t)L.mech, (uint64_t)(uintptr_t)L.U, (uint64_t)L.ncell, (uint64_t)L.ld,
        (uint64_t)(uintptr_t)L.dt, (uint64_t)L.dt_per_cell, (uint64_t)(uintptr_t)L.out, (uint64_t)L.ldo,
        (uint64_t)(uintptr_t)L.status, (uint64_t)(uintptr_t)L.first_failed, (uint64_t)(uintptr_t)L.counters,
        bits(L.rtol), bits(L.atol), (uint64_t)L.maxsteps, (uint64_t)L.mode, (uint64_t)L.precision,
        (uint64_t)L.num_cus, (uint64_t)L.waves_per_cu, (uint64_t)L.chunk, (uint64_t)L.nbins, (uint64_t)L.jac_wait,
        (uint64_t)L.setup_wait,
        (uint64_t)(uintptr_t)L.workspace, (uint64_t)L.workspace_bytes, (uint64_t)(uintptr_t)L.commit,
        (uint64_t)L.commit_ld, (uint64_t)(uintptr_t)L.blocks, (uint64_t)L.nblock};
    ChemGraph& g = c->graph;
    if (!g.t0 && (hipEventCreate(&g.t0) != hipSuccess || hipEventCreate(&g.t1) != hipSuccess)) return HF_EDEVICE;
    *how = 1;
    if (!g.exec || g.key != key) {
        *how = 2;
        g.reset();
        L.kernel_ms = nullptr;   // events are recorded around the graph launch instead
        if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return HF_EDEVICE;
        const int r = c->ops->chem(L);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (r != HF_OK || e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return r != HF_OK ? r : HF_EDEVICE;
        }
        g.graph = graph;
        if (hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0) != hipSuccess) {
            g.reset();
            return HF_EDEVICE;
        }
        g.key = key;
    }
    if (hipEventRecord(g.t0, c->stream) != hipSuccess || hipGraphLaunch(g.exec, c->stream) != hipSuccess ||
        hipEventRecord(g.t1, c->stream) != hipSuccess)
        return HF_EDEVICE;
    if (hipEventSynchronize(g.t1) != hipSuccess || hipEventElapsedTime(ms, g.t0, g.t1) != hipSuccess) return HF_EDEVICE;
    return HF_OK;
}

}  // namespace

namespace {
// hf_chem_advance over flat planes (blocks null) or over the leaf cells of a
// block table (hf_chem_advance_cells: ncell = the blocks' cell count, U and
// ld unused, the cells addressed through the uploaded table).
int chem_impl(hf_ctx* c, double* U, int64_t ncell, int64_t ld, const double* dt, int dt_per_cell, int mode,
              hf_chem_stats* stats, int64_t* first_failed, const hf_cells* blocks, int nblock) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "nested or concurrent chemistry batch is unsupported");
    if (first_failed) *first_failed = -1;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    std::vector<int64_t> bstart;
    if (blocks) {
        // every block in device memory, one interval for the whole batch (BlockAMR::react)
        bstart.assign(1, 0);
        for (int b = 0; b < nblock; ++b) {
            const hf_cells& B = blocks[b];
            if (B.nx < 1 || B.ny < 1 || !B.U || B.pitch < B.nx || B.plane < B.pitch * B.ny || B.offset < 0 ||
                (B.mask && B.mask_pitch < B.nx))
                return fail(c, HF_EINVAL, "block " + std::to_string(b) + " has an invalid extent or stride");
            bstart.push_back(bstart.back() + (int64_t)B.nx * B.ny);
        }
        ncell = ld = bstart.back();
        U = nullptr;
        if (dt_per_cell) return fail(c, HF_EINVAL, "a block batch takes one interval");
    }
    if (ncell < 0 || ld < ncell || !dt || (ncell > 0 && !U && !blocks) ||
        (mode != HF_CHEM_PLAIN && mode != HF_CHEM_BATCHED))
        return fail(c, HF_EINVAL, "invalid chemistry batch extent or interval");
    const bool dev_dt = is_device_ptr(dt);
    if (!dt_per_cell) {
        double h = 0;
        if (dev_dt) HF_HIP(c, hipMemcpy(&h, dt, sizeof(double), hipMemcpyDeviceToHost));
        else h = dt[0];
        if (!(h > 0) || !std::isfinite(h)) return fail(c, HF_EINVAL, "invalid chemistry batch extent or interval");
    }
    if (ncell == 0) return HF_OK;
    HF_DEVICE(c);
    const int ns = c->mech.ns(), nv = ns + 3;
    if (blocks) {
        for (int b = 0; b < nblock; ++b)
            if (!is_device_ptr(blocks[b].U) || (blocks[b].mask && !is_device_ptr(blocks[b].mask)))
                return fail(c, HF_EINVAL, "block " + std::to_string(b) + " is not in device memory");
        const size_t tb = sizeof(hf_cells) * nblock, sb = sizeof(int64_t) * (nblock + 1);
        if (!c->blocks.ensure(tb + sb)) return fail(c, HF_ENOMEM, "device scratch");
        HF_HIP(c, hipMemcpyAsync(c->blocks.p, blocks, tb, hipMemcpyHostToDevice, c->stream));
        HF_HIP(c, hipMemcpyAsync(static_cast<char*>(c->blocks.p) + tb, bstart.data(), sb, hipMemcpyHostToDevice,
                                 c->stream));
    }
    const bool dev_U = blocks || is_device_ptr(U);
    const double* dU = U;
    int64_t dld = ld;
    if (!dev_U) {
        if (!c->U.ensure(sizeof(double) * nv * ncell)) return fail(c, HF_ENOMEM, "device staging");
        HF_HIP(c, hipMemcpy2DAsync(c->U.p, ncell * sizeof(double), U, ld * sizeof(double), ncell * sizeof(double), nv,
                                   hipMemcpyHostToDevice, c->stream));
        dU = c->U.as<double>();
        dld = ncell;
    }
    const double* ddt = dt;
    const int64_t ndt = dt_per_cell ? ncell : 1;
    if (!dev_dt) {
        if (!c->dt.ensure(sizeof(double) * ndt)) return fail(c, HF_ENOMEM, "device staging");
        HF_HIP(c, hipMemcpyAsync(c->dt.p, dt, sizeof(double) * ndt, hipMemcpyHostToDevice, c->stream));
        ddt = c->dt.as<double>();
    }
    if (!c->out.ensure(sizeof(double) * ns * ncell) || !c->status.ensure(ncell) ||
        !c->scal.ensure(sizeof(unsigned long long) * 16))
        return fail(c, HF_ENOMEM, "device scratch");
    unsigned long long* scal = c->scal.as<unsigned long long>();
    // scal[0] = first failed (init ~0), scal[1..8] = counters
    std::vector<unsigned long long> init(16, 0);
    init[0] = ~0ull;
    HF_HIP(c, hipMemcpyAsync(scal, init.data(), sizeof(unsigned long long) * 16, hipMemcpyHostToDevice, c->stream));
    ChemLaunch L{};
    L.mech = c->d_mech;
    L.U = dU;
    L.ncell = ncell;
    L.ld = dld;
    if (blocks) {
        L.blocks = c->blocks.as<hf_cells>();
        L.block_start = reinterpret_cast<const int64_t*>(c->blocks.as<char>() + sizeof(hf_cells) * nblock);
        L.nblock = nblock;
    }
    L.dt = ddt;
    L.dt_per_cell = dt_per_cell;
    L.out = c->out.as<double>();
    L.ldo = ncell;
    L.status = c->status.as<int8_t>();
    L.first_failed = scal;
    L.counters = scal + 1;
    L.rtol = c->opt.rtol;
    L.atol = c->opt.atol;
    L.maxsteps = c->opt.max_steps;
    L.mode = mode;
    L.precision = c->opt.precision;
    L.num_cus = c->num_cus;
    L.waves_per_cu = c->opt.waves_per_cu;
    L.chunk = c->opt.chunk;
    L.nbins = c->opt.nbins;
    L.jac_wait = c->opt.jac_wait;
    L.setup_wait = c->opt.setup_wait;
    // online tuning when every tuning value is left to the library
    const bool tune = mode == HF_CHEM_BATCHED && !c->opt.chunk && !c->opt.waves_per_cu && !c->opt.nbins &&
                      !c->opt.jac_wait && !c->opt.setup_wait;
    if (tune) {
        int cfg[ChemTuner::NPARAM];
        c->tuner.pick(ncell, c->ops->ns > 16, cfg);
        L.chunk = cfg[ChemTuner::CHUNK];
        L.waves_per_cu = cfg[ChemTuner::WAVES];
        L.nbins = cfg[ChemTuner::BINS];
        L.jac_wait = cfg[ChemTuner::JWAIT];
        L.setup_wait = cfg[ChemTuner::SWAIT];
    }
    const size_t wbytes = c->ops->chem_workspace(ncell, L.nbins, c->num_cus, L.waves_per_cu);
    if (wbytes) {
        if (!c->work.ensure(wbytes)) return fail(c, HF_ENOMEM, "device workspace");
        L.workspace = c->work.p;
        L.workspace_bytes = wbytes;
    }
    L.stream = c->stream;
    float ms = 0;
    L.kernel_ms = &ms;
    if (dev_U) {
        // device-resident field: the commit is a kernel that runs only when
        // no cell failed, so the whole call is one capturable sequence
        // (block form: non-null to enable it; the kernel writes through the block table)
        L.commit = blocks ? blocks[0].U : U;
        L.commit_ld = ld;
    }
    int r, how = 0;
    if (dev_U && dev_dt && !c->opt.eager) {
        r = chem_graph_launch(c, L, &ms, &how);
    } else {
        r = c->ops->chem(L);
    }
    if (r != HF_OK) return fail(c, r, r == HF_EDEVICE ? "chemistry kernel launch failed" : "invalid chemistry launch");
    unsigned long long host_scal[9];
    HF_HIP(c, hipMemcpyAsync(host_scal, scal, sizeof(host_scal), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipStreamSynchronize(c->stream));
    if (stats) {
        stats->cells = ncell;
        stats->steps = (int64_t)host_scal[1];
        stats->rhs = (int64_t)host_scal[2];
        stats->jac = (int64_t)host_scal[3];
        stats->lu = (int64_t)host_scal[4];
        stats->newton = (int64_t)host_scal[5];
        stats->conv_failures = (int64_t)host_scal[6];
        stats->err_failures = (int64_t)host_scal[7];
        stats->max_cell_steps = (int64_t)host_scal[8];
        stats->team = (int)std::min<int64_t>(ncell, INT32_MAX);
        stats->kernel_ms = ms;
        stats->chunk = mode == HF_CHEM_BATCHED ? L.chunk : 0;
        stats->graph = how;
        stats->waves_per_cu = mode == HF_CHEM_BATCHED ? L.waves_per_cu : 0;
        stats->nbins = mode == HF_CHEM_BATCHED ? L.nbins : 0;
        stats->jac_wait = mode == HF_CHEM_BATCHED ? L.jac_wait : 0;
        stats->setup_wait = mode == HF_CHEM_BATCHED ? L.setup_wait : 0;
        stats->tuned = tune ? (c->tuner.done ? 2 : 1) : 0;
    }
    if (tune && host_scal[0] == ~0ull) c->tuner.record(ms, ncell);
    if (host_scal[0] != ~0ull) {
        const int64_t k = (int64_t)host_scal[0];
        int8_t code = 0;
        HF_HIP(c, hipMemcpy(&code, c->status.as<int8_t>() + k, 1, hipMemcpyDeviceToHost));
        const char* what = code == HF_EINVAL ? "invalid chemistry interval"
                           : code == HF_EDOMAIN ? "invalid accepted species or inadmissible state"
                                                : "integrator failed (CVode-equivalent error)";
        if (blocks) {
            // (block << 32) | (j*nx + i), the flux stage's convention
            const int b = (int)(std::upper_bound(bstart.begin(), bstart.end(), k) - bstart.begin()) - 1;
            const int64_t local = k - bstart[b];
            if (first_failed) *first_failed = ((int64_t)b << 32) | local;
            return fail(c, code ? code : HF_ECONV,
                        "chemistry block " + std::to_string(b) + " cell (" + std::to_string(local % blocks[b].nx) +
                            ", " + std::to_string(local / blocks[b].nx) + "): " + what);
        }
        if (first_failed) *first_failed = k;
        return fail(c, code ? code : HF_ECONV, "chemistry cell " + std::to_string(k) + ": " + what);
    }
    // commit the species planes (device-resident fields were committed by the kernels)
    if (!dev_U) {
        HF_HIP(c, hipMemcpy2DAsync(U, ld * sizeof(double), c->out.p, ncell * sizeof(double), ncell * sizeof(double), ns,
                                   hipMemcpyDeviceToHost, c->stream));
        HF_HIP(c, hipStreamSynchronize(c->stream));
    }
    return HF_OK;
}
}  // namespace

int hf_chem_advance(hf_ctx* c, double* U, int64_t ncell, int64_t ld, const double* dt, int dt_per_cell, int mode,
                    hf_chem_stats* stats, int64_t* first_failed) {
    return chem_impl(c, U, ncell, ld, dt, dt_per_cell, mode, stats, first_failed, nullptr, 0);
}

int hf_chem_advance_cells(hf_ctx* c, const hf_cells* blocks, int nblock, const double* dt, int mode,
                          hf_chem_stats* stats, int64_t* first_failed) {
    if (c && (nblock < 0 || (nblock > 0 && !blocks))) return fail(c, HF_EINVAL, "invalid block table");
    if (c && nblock == 0) {
        if (first_failed) *first_failed = -1;
        if (stats) std::memset(stats, 0, sizeof(*stats));
        return HF_OK;
    }
    return chem_impl(c, nullptr, 0, 0, dt, 0, mode, stats, first_failed, blocks, nblock);
}

namespace {
// noalloc: the tables and counters live in the device's static scratch (no
// device allocation, no timing events; device-resident arrays only).
// in_ptrs/out_ptrs: one device base pointer per box (then Uin/Uout are unused).
int flux_stage_impl(hf_ctx* c, const hf_box* boxes, int nbox, const double* Uin, double* Uout, int first_order,
                    const hf_face_fluxes* fluxes, hf_flux_stats* stats, int64_t* first_failed, bool noalloc = false,
                    const double* const* in_ptrs = nullptr, double* const* out_ptrs = nullptr) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent flux stage is unsupported");
    if (first_failed) *first_failed = -1;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    const bool per_box = in_ptrs || out_ptrs;
    if (nbox < 0 || (nbox > 0 && (!boxes || (per_box ? !(in_ptrs && out_ptrs) : (!Uin || !Uout)))))
        return fail(c, HF_EINVAL, "invalid flux-stage arguments");
    if (nbox == 0) return HF_OK;
    char* fixed = nullptr;
    if (noalloc) {
        // the device was bound when the context was created: binding it now would allocate
        if (!c->d_mech) return fail(c, HF_EDEVICE, "no device was bound when the context was created");
        if (hipSetDevice(c->opt.device) != hipSuccess) return fail(c, HF_EDEVICE, "hipSetDevice failed");
        fixed = static_cast<char*>(g_scratch_ptr[c->opt.device].load());
        if (!fixed) return fail(c, HF_EDEVICE, "the static scratch has no device address");
        if (nbox > kNoAllocMaxBoxes)
            return fail(c, HF_EINVAL, "more than " + std::to_string(kNoAllocMaxBoxes) + " boxes for the no-allocation stage");
        if (!is_device_ptr(Uin) || !is_device_ptr(Uout))
            return fail(c, HF_EINVAL, "the no-allocation stage needs device arrays");
    }
    HF_DEVICE(c);
    const int nv = c->mech.ns() + 3;
    // all boxes 1-D (x only) or all 2-D
    bool one_d = true;
    int64_t in_extent = 0, out_extent = 0, cells = 0;
    for (int b = 0; b < nbox; ++b) {
        const hf_box& B = boxes[b];
        if (B.nx < 1 || B.ny < 1 || B.in_pitch < B.nx + 4 || B.out_pitch < B.nx || B.in_plane < B.in_pitch * (B.ny + 4) ||
            B.out_plane < B.out_pitch * B.ny || B.in_offset < 0 || B.out_offset < 0 || !std::isfinite(B.dtdx) ||
            !std::isfinite(B.dtdy))
            return fail(c, HF_EINVAL, "box " + std::to_string(b) + " has an invalid extent or stride");
        if (!(B.ny == 1 && B.dtdy == 0)) one_d = false;
        in_extent = std::max<int64_t>(in_extent, B.in_offset + (nv - 1) * B.in_plane + (B.ny + 3) * B.in_pitch + B.nx + 4);
        out_extent = std::max<int64_t>(out_extent, B.out_offset + (nv - 1) * B.out_plane + (B.ny - 1) * B.out_pitch + B.nx);
        cells += (int64_t)B.nx * B.ny;
    }
    // 2-D tile shape: tuned online per size class (results do not depend on it)
    const bool tune_shape = !one_d && !noalloc && c->ops->flux_nshape > 1;
    const int shape = tune_shape ? c->ftuner.pick(cells, c->ops->flux_nshape) : 0;
    const int TX = one_d ? c->ops->flux_tx1 : c->ops->flux_shape_tx[shape];
    const int TY = one_d ? 1 : c->ops->flux_shape_ty[shape];
    std::vector<int32_t> tiles;
    for (int b = 0; b < nbox; ++b)
        for (int tj = 0; tj < (boxes[b].ny + TY - 1) / TY; ++tj)
            for (int ti = 0; ti < (boxes[b].nx + TX - 1) / TX; ++ti) {
                tiles.push_back(b);
                tiles.push_back(ti);
                tiles.push_back(tj);
            }
    const int ntiles = (int)(tiles.size() / 3);
    if (noalloc && ntiles > kNoAllocMaxTiles)
        return fail(c, HF_EINVAL, std::to_string(ntiles) + " tiles exceed the no-allocation limit of " +
                                      std::to_string(kNoAllocMaxTiles));
    if (fluxes)
        for (int b = 0; b < nbox; ++b) {
            const hf_face_fluxes& F = fluxes[b];
            const hf_box& B = boxes[b];
            if ((F.fx && !is_device_ptr(F.fx)) || (F.fy && !is_device_ptr(F.fy)))
                return fail(c, HF_EINVAL, "face fluxes need device memory");
            // the faces the stage writes: x (0..nx, 0..ny-1), y (0..nx-1, 0..ny)
            if ((F.fx && !within_allocation(F.fx, F.fx_offset,
                                            F.fx_offset + (nv - 1) * F.fx_plane + (B.ny - 1) * F.fx_pitch + B.nx)) ||
                (F.fy && !one_d &&
                 !within_allocation(F.fy, F.fy_offset, F.fy_offset + (nv - 1) * F.fy_plane + B.ny * F.fy_pitch + B.nx - 1)))
                return fail(c, HF_EINVAL, "face fluxes of box " + std::to_string(b) + " lie outside their allocation");
        }
    if (per_box) {
        // every box inside its own input and output allocations
        for (int b = 0; b < nbox; ++b) {
            const hf_box& B = boxes[b];
            if (!is_device_ptr(in_ptrs[b]) || !is_device_ptr(out_ptrs[b]))
                return fail(c, HF_EINVAL, "box " + std::to_string(b) + ": per-box arrays need device memory");
            if (!within_allocation(in_ptrs[b], B.in_offset,
                                   B.in_offset + (nv - 1) * B.in_plane + (B.ny + 3) * B.in_pitch + B.nx + 3))
                return fail(c, HF_EINVAL, "box " + std::to_string(b) + " reads outside its input allocation");
            if (!within_allocation(out_ptrs[b], B.out_offset,
                                   B.out_offset + (nv - 1) * B.out_plane + (B.ny - 1) * B.out_pitch + B.nx - 1))
                return fail(c, HF_EINVAL, "box " + std::to_string(b) + " writes outside its output allocation");
        }
    } else if (is_device_ptr(Uin) && !within_allocation(Uin, 0, in_extent - 1))
        return fail(c, HF_EINVAL, "a box reads outside the input allocation");
    if (!per_box && is_device_ptr(Uout)) {
        for (int b = 0; b < nbox; ++b)
            if (!within_allocation(Uout, boxes[b].out_offset, boxes[b].out_offset + (nv - 1) * boxes[b].out_plane +
                                                                   (boxes[b].ny - 1) * boxes[b].out_pitch +
                                                                   boxes[b].nx - 1))
                return fail(c, HF_EINVAL, "box " + std::to_string(b) + " writes outside the output allocation");
    }
    const size_t flux_table = fluxes ? sizeof(hf_face_fluxes) * nbox : 0;
    const size_t ptr_table = per_box ? 2 * sizeof(void*) * nbox : 0;
    char* box_dev;
    int32_t* tile_dev;
    unsigned long long* scal;
    if (noalloc) {
        scal = reinterpret_cast<unsigned long long*>(fixed);
        box_dev = fixed + kNoAllocBoxOff;
        tile_dev = reinterpret_cast<int32_t*>(fixed + kNoAllocTileOff);
    } else {
        if (!c->boxes.ensure(sizeof(hf_box) * nbox + flux_table + ptr_table) ||
            !c->tiles.ensure(sizeof(int32_t) * tiles.size()) ||
            !c->scal.ensure(sizeof(unsigned long long) * 16))
            return fail(c, HF_ENOMEM, "device scratch");
        scal = c->scal.as<unsigned long long>();
        box_dev = static_cast<char*>(c->boxes.p);
        tile_dev = c->tiles.as<int32_t>();
    }
    char* flux_dev = noalloc ? fixed + kNoAllocFluxOff : box_dev + sizeof(hf_box) * nbox;
    HF_HIP(c, hipMemcpyAsync(box_dev, boxes, sizeof(hf_box) * nbox, hipMemcpyHostToDevice, c->stream));
    if (fluxes) HF_HIP(c, hipMemcpyAsync(flux_dev, fluxes, flux_table, hipMemcpyHostToDevice, c->stream));
    char* ptr_dev = box_dev + sizeof(hf_box) * nbox + flux_table;
    if (per_box) {
        // the call synchronises before it returns, so the caller's host tables outlive these copies
        HF_HIP(c, hipMemcpyAsync(ptr_dev, in_ptrs, sizeof(void*) * nbox, hipMemcpyHostToDevice, c->stream));
        HF_HIP(c, hipMemcpyAsync(ptr_dev + sizeof(void*) * nbox, out_ptrs, sizeof(void*) * nbox, hipMemcpyHostToDevice,
                                 c->stream));
    }
    HF_HIP(c, hipMemcpyAsync(tile_dev, tiles.data(), sizeof(int32_t) * tiles.size(), hipMemcpyHostToDevice, c->stream));
    const bool dev_in = per_box || is_device_ptr(Uin), dev_out = per_box || is_device_ptr(Uout);
    const double* din = Uin;
    double* dout = Uout;
    if (!dev_in) {
        if (!c->U.ensure(sizeof(double) * in_extent)) return fail(c, HF_ENOMEM, "device staging");
        HF_HIP(c, hipMemcpyAsync(c->U.p, Uin, sizeof(double) * in_extent, hipMemcpyHostToDevice, c->stream));
        din = c->U.as<double>();
    }
    if (!dev_out) {
        if (!c->out.ensure(sizeof(double) * out_extent)) return fail(c, HF_ENOMEM, "device staging");
        HF_HIP(c, hipMemcpyAsync(c->out.p, Uout, sizeof(double) * out_extent, hipMemcpyHostToDevice, c->stream));
        dout = c->out.as<double>();
    }
    // the failure word and counters start fresh on every call (the static
    // scratch keeps the previous call's values)
    unsigned long long init[3] = {~0ull, 0, 0};
    HF_HIP(c, hipMemcpyAsync(scal, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
    FluxLaunch L{};
    L.mech = c->d_mech;
    L.boxes = reinterpret_cast<const hf_box*>(box_dev);
    L.nbox = nbox;
    L.tiles = tile_dev;
    L.ntiles = ntiles;
    L.Uin = din;
    L.Uout = dout;
    L.first_order = first_order ? 1 : 0;
    L.one_d = one_d ? 1 : 0;
    L.shape = shape;
    L.first_failed = scal;
    L.counters = scal + 1;
    L.fluxes = fluxes ? reinterpret_cast<const hf_face_fluxes*>(flux_dev) : nullptr;
    L.in_ptrs = per_box ? reinterpret_cast<const double* const*>(ptr_dev) : nullptr;
    L.out_ptrs = per_box ? reinterpret_cast<double* const*>(ptr_dev + sizeof(void*) * nbox) : nullptr;
    L.stream = c->stream;
    float ms = 0;
    L.kernel_ms = noalloc ? nullptr : &ms;
    const int r = c->ops->flux(L);
    if (r != HF_OK) return fail(c, r, "flux kernel launch failed");
    unsigned long long hs[3];
    HF_HIP(c, hipMemcpyAsync(hs, scal, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
    if (!dev_out)
        HF_HIP(c, hipMemcpyAsync(Uout, dout, sizeof(double) * out_extent, hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipStreamSynchronize(c->stream));
    if (stats) {
        stats->cells = cells;
        stats->faces = (int64_t)hs[1];
        stats->first_order_sides = (int64_t)hs[2];
        stats->kernel_ms = ms;
        stats->tile_x = TX;
        stats->tile_y = TY;
    }
    if (tune_shape) c->ftuner.record(ms, cells);
    if (hs[0] != ~0ull) {
        if (first_failed) *first_failed = (int64_t)hs[0];
        return fail(c, HF_EDOMAIN, "inadmissible state after flux stage at box " + std::to_string(hs[0] >> 32) +
                                       " cell " + std::to_string(hs[0] & 0xffffffffull));
    }
    return HF_OK;
}
}  // namespace

int hf_flux_stage(hf_ctx* c, const hf_box* boxes, int nbox, const double* Uin, double* Uout, int first_order,
                  hf_flux_stats* stats, int64_t* first_failed) {
    return flux_stage_impl(c, boxes, nbox, Uin, Uout, first_order, nullptr, stats, first_failed);
}

int hf_flux_stage_fluxes(hf_ctx* c, const hf_box* boxes, int nbox, const double* Uin, double* Uout, int first_order,
                         const hf_face_fluxes* fluxes, hf_flux_stats* stats, int64_t* first_failed) {
    if (c && nbox > 0 && !fluxes) return fail(c, HF_EINVAL, "missing face-flux descriptors");
    return flux_stage_impl(c, boxes, nbox, Uin, Uout, first_order, fluxes, stats, first_failed);
}

int hf_flux_stage_boxes(hf_ctx* c, const hf_box* boxes, int nbox, const double* const* in, double* const* out,
                        int first_order, const hf_face_fluxes* fluxes, hf_flux_stats* stats, int64_t* first_failed) {
    return flux_stage_impl(c, boxes, nbox, nullptr, nullptr, first_order, fluxes, stats, first_failed, false, in, out);
}

int hf_flux_stage_noalloc(hf_ctx* c, const hf_box* boxes, int nbox, const double* Uin, double* Uout, int first_order,
                          const hf_face_fluxes* fluxes, hf_flux_stats* stats, int64_t* first_failed) {
    return flux_stage_impl(c, boxes, nbox, Uin, Uout, first_order, fluxes, stats, first_failed, true);
}

static int props_common(hf_ctx* c, const double* U, int64_t ncell, int64_t ld, double* p, double* a, double* T,
                        int64_t* first_failed, double dx, double dy, double* rate) {
    if (ncell < 0 || ld < ncell || (ncell > 0 && !U)) return fail(c, HF_EINVAL, "invalid extent");
    if (first_failed) *first_failed = -1;
    if (ncell == 0) return HF_OK;
    HF_DEVICE(c);
    const int nv = c->mech.ns() + 3;
    const double* dU = U;
    int64_t dld = ld;
    if (!is_device_ptr(U)) {
        if (!c->U.ensure(sizeof(double) * nv * ncell)) return fail(c, HF_ENOMEM, "device staging");
        HF_HIP(c, hipMemcpy2DAsync(c->U.p, ncell * sizeof(double), U, ld * sizeof(double), ncell * sizeof(double), nv,
                                   hipMemcpyHostToDevice, c->stream));
        dU = c->U.as<double>();
        dld = ncell;
    }
    double* outs[3] = {p, a, T};
    double* douts[3] = {nullptr, nullptr, nullptr};
    bool host_out[3] = {false, false, false};
    for (int i = 0; i < 3; ++i) {
        if (!outs[i]) continue;
        if (is_device_ptr(outs[i])) {
            douts[i] = outs[i];
        } else {
            if (!c->aux[i].ensure(sizeof(double) * ncell)) return fail(c, HF_ENOMEM, "device staging");
            douts[i] = c->aux[i].as<double>();
            host_out[i] = true;
        }
    }
    if (!c->scal.ensure(sizeof(unsigned long long) * 16)) return fail(c, HF_ENOMEM, "device scratch");
    unsigned long long* scal = c->scal.as<unsigned long long>();
    unsigned long long init[2] = {~0ull, 0};
    HF_HIP(c, hipMemcpyAsync(scal, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
    PropsLaunch L{};
    L.mech = c->d_mech;
    L.U = dU;
    L.ncell = ncell;
    L.ld = dld;
    L.p = douts[0];
    L.a = douts[1];
    L.T = douts[2];
    L.first_failed = scal;
    L.dx = dx;
    L.dy = dy;
    L.rate_max = rate ? reinterpret_cast<double*>(scal + 1) : nullptr;
    L.stream = c->stream;
    const int r = c->ops->props(L);
    if (r != HF_OK) return fail(c, r, "properties kernel launch failed");
    unsigned long long hs[2];
    HF_HIP(c, hipMemcpyAsync(hs, scal, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
    for (int i = 0; i < 3; ++i)
        if (host_out[i])
            HF_HIP(c, hipMemcpyAsync(outs[i], douts[i], sizeof(double) * ncell, hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipStreamSynchronize(c->stream));
    if (rate) std::memcpy(rate, &hs[1], sizeof(double));
    if (hs[0] != ~0ull) {
        if (first_failed) *first_failed = (int64_t)hs[0];
        return fail(c, HF_EDOMAIN, "inadmissible conserved state at cell " + std::to_string(hs[0]));
    }
    return HF_OK;
}

int hf_properties(hf_ctx* c, const double* U, int64_t ncell, int64_t ld, double* p, double* a, double* T,
                  int64_t* first_failed) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    return props_common(c, U, ncell, ld, p, a, T, first_failed, 0, 0, nullptr);
}

int hf_stable_dt(hf_ctx* c, const double* U, int64_t ncell, int64_t ld, double dx, double dy, double cfl, double* dt) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    if (!dt || !(dx > 0) || !(dy >= 0) || !(cfl > 0)) return fail(c, HF_EINVAL, "invalid CFL arguments");
    double rate = 0;
    int64_t ff;
    const int r = props_common(c, U, ncell, ld, nullptr, nullptr, nullptr, &ff, dx, dy, &rate);
    if (r != HF_OK) return r;
    *dt = rate > 0 ? cfl / rate : INFINITY;
    return HF_OK;
}

namespace {

// Host or device arrays of the sensitivity entry points: device pointers are
// used in place, host arrays are staged through one device buffer (inputs
// copied in, outputs copied back after the kernels).
struct Staging {
    struct Item {
        const void* host;
        void* hout;
        size_t off, bytes;
    };
    std::vector<Item> in, out;
    size_t total = 0;
    static size_t pad(size_t b) { return (b + 255) & ~size_t(255); }
    // returns the offset (bytes) of a staged array, or SIZE_MAX when used in place
    size_t add_in(const double* p, size_t n) {
        if (!p || is_device_ptr(p)) return SIZE_MAX;
        in.push_back({p, nullptr, total, n * sizeof(double)});
        total += pad(n * sizeof(double));
        return in.back().off;
    }
    size_t add_out(double* p, size_t n) {
        if (!p || is_device_ptr(p)) return SIZE_MAX;
        out.push_back({nullptr, p, total, n * sizeof(double)});
        total += pad(n * sizeof(double));
        return out.back().off;
    }
};

int sens_common(hf_ctx* c, bool tangent, int64_t ncell, const double* rho, const double* T0, const double* Y0,
                double h, int nsteps, int ckpt_every, const double* lnA_shift, const double* v_lnA,
                const double* v_T0, const double* v_Y0, double* J, double* grad_lnA, double* grad_T0,
                double* grad_Y0, double* dJ) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    if (ncell < 0) return fail(c, HF_EINVAL, "invalid extent");
    if (!(h > 0) || !std::isfinite(h) || nsteps < 1) return fail(c, HF_EINVAL, "invalid step size or step count");
    if (!tangent && ckpt_every < 1) return fail(c, HF_EINVAL, "invalid checkpoint interval");
    if (ncell == 0) return HF_OK;
    if (!rho || !T0 || !Y0 || !J) return fail(c, HF_EINVAL, "missing state or objective array");
    if (tangent ? (!v_lnA || !v_T0 || !v_Y0 || !dJ) : (!grad_lnA || !grad_T0 || !grad_Y0))
        return fail(c, HF_EINVAL, "missing direction or gradient array");
    HF_DEVICE(c);
    const size_t n = (size_t)ncell, ns = (size_t)c->ops->ns, nr = (size_t)c->ops->nr;
    Staging st;
    const size_t o_rho = st.add_in(rho, n), o_T0 = st.add_in(T0, n), o_Y0 = st.add_in(Y0, n * ns);
    const size_t o_sh = st.add_in(lnA_shift, nr);
    size_t o_vl = SIZE_MAX, o_vt = SIZE_MAX, o_vy = SIZE_MAX, o_gl = SIZE_MAX, o_gt = SIZE_MAX, o_gy = SIZE_MAX,
           o_dj = SIZE_MAX;
    const size_t o_J = st.add_out(J, n);
    if (tangent) {
        o_vl = st.add_in(v_lnA, n * nr);
        o_vt = st.add_in(v_T0, n);
        o_vy = st.add_in(v_Y0, n * ns);
        o_dj = st.add_out(dJ, n);
    } else {
        o_gl = st.add_out(grad_lnA, n * nr);
        o_gt = st.add_out(grad_T0, n);
        o_gy = st.add_out(grad_Y0, n * ns);
    }
    const size_t wbytes = tangent ? 0 : c->ops->sens_workspace(ncell, nsteps, ckpt_every);
    if (!c->aux[0].ensure(st.total + 256)) return fail(c, HF_ENOMEM, "device staging buffer");
    if (wbytes && !c->work.ensure(wbytes)) return fail(c, HF_ENOMEM, "device checkpoint workspace");
    if (!c->scal.ensure(256)) return fail(c, HF_ENOMEM, "device scalars");
    char* base = c->aux[0].as<char>();
    auto dev = [&](const double* p, size_t off) { return off == SIZE_MAX ? p : reinterpret_cast<const double*>(base + off); };
    auto devo = [&](double* p, size_t off) { return off == SIZE_MAX ? p : reinterpret_cast<double*>(base + off); };
    for (const auto& it : st.in)
        HF_HIP(c, hipMemcpyAsync(base + it.off, it.host, it.bytes, hipMemcpyHostToDevice, c->stream));
    AdjointLaunch L{};
    L.mech = c->d_mech;
    L.ncell = ncell;
    L.rho = dev(rho, o_rho);
    L.T0 = dev(T0, o_T0);
    L.Y0 = dev(Y0, o_Y0);
    L.h = h;
    L.nsteps = nsteps;
    L.ckpt_every = ckpt_every;
    L.lnA_shift = dev(lnA_shift, o_sh);
    L.J = devo(J, o_J);
    if (tangent) {
        L.v_lnA = dev(v_lnA, o_vl);
        L.v_T0 = dev(v_T0, o_vt);
        L.v_Y0 = dev(v_Y0, o_vy);
        L.dJ = devo(dJ, o_dj);
    } else {
        L.grad_lnA = devo(grad_lnA, o_gl);
        L.grad_T0 = devo(grad_T0, o_gt);
        L.grad_Y0 = devo(grad_Y0, o_gy);
        L.workspace = c->work.as<double>();
        L.workspace_bytes = wbytes;
    }
    L.first_failed = c->scal.as<unsigned long long>();
    L.stream = c->stream;
    const int r = tangent ? c->ops->tangent(L) : c->ops->adjoint(L);
    if (r != HF_OK) {
        (void)hipStreamSynchronize(c->stream);
        return fail(c, r, r == HF_EDEVICE ? "sensitivity kernel launch failed" : "invalid sensitivity arguments");
    }
    for (const auto& it : st.out)
        HF_HIP(c, hipMemcpyAsync(it.hout, base + it.off, it.bytes, hipMemcpyDeviceToHost, c->stream));
    unsigned long long ff = ~0ull;
    HF_HIP(c, hipMemcpyAsync(&ff, L.first_failed, sizeof(ff), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipStreamSynchronize(c->stream));
    if (ff != ~0ull)
        return fail(c, HF_ECONV, "sensitivity cell " + std::to_string(ff) +
                                     ": backward-Euler Newton iteration failed (results of that cell are NaN)");
    return HF_OK;
}

}  // namespace

int hf_chem_adjoint(hf_ctx* c, int64_t ncell, const double* rho, const double* T0, const double* Y0, double h,
                    int nsteps, int ckpt_every, const double* lnA_shift, double* J, double* grad_lnA,
                    double* grad_T0, double* grad_Y0) {
    return sens_common(c, false, ncell, rho, T0, Y0, h, nsteps, ckpt_every, lnA_shift, nullptr, nullptr, nullptr, J,
                       grad_lnA, grad_T0, grad_Y0, nullptr);
}

int hf_chem_tangent(hf_ctx* c, int64_t ncell, const double* rho, const double* T0, const double* Y0, double h,
                    int nsteps, const double* lnA_shift, const double* v_lnA, const double* v_T0,
                    const double* v_Y0, double* J, double* dJ) {
    return sens_common(c, true, ncell, rho, T0, Y0, h, nsteps, 1, lnA_shift, v_lnA, v_T0, v_Y0, J, nullptr, nullptr,
                       nullptr, dJ);
}

// ------------------------------------------------------------------ single-box driver operations
namespace {

__global__ void __launch_bounds__(256) lincomb_kernel(double* __restrict__ y, const double* __restrict__ x, double a,
                                                      double b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = fma(a, y[i], b * x[i]);
}

int box_common(hf_ctx* c, BoxLaunch& L, const hf_box* box, const void* U) {
    if (!box || !U) return fail(c, HF_EINVAL, "missing box or array");
    if (box->nx <= 0 || box->ny <= 0 || box->in_pitch < box->nx + 4 || box->in_plane < box->in_pitch * (box->ny + 4))
        return fail(c, HF_EINVAL, "invalid box extent");
    HF_DEVICE(c);
    if (!is_device_ptr(U)) return fail(c, HF_EINVAL, "box operations need device memory");
    if (!c->scal.ensure(sizeof(unsigned long long) * 256)) return fail(c, HF_ENOMEM, "device scalars");
    L.mech = c->d_mech;
    L.box = *box;
    L.scal = c->scal.as<unsigned long long>() + 8;
    L.first_failed = c->scal.as<unsigned long long>();
    L.stream = c->stream;
    return HF_OK;
}

}  // namespace

int hf_fill_ghosts(hf_ctx* c, double* U, const hf_box* box, const int32_t bc[4]) {
    if (!c) return HF_EINVAL;
    if (bc)
        for (int k = 0; k < 4; ++k)
            if (bc[k] == HF_BC_INFLOW || bc[k] == HF_BC_WEDGE || bc[k] == HF_BC_KEEP)
                return fail(c, HF_EINVAL, "inflow, wedge and kept sides need hf_fill_ghosts_flow");
    return hf_fill_ghosts_flow(c, U, box, bc, nullptr);
}

namespace {
// Boundary codes of a ghost fill (hf_fill_ghosts_flow's rules).
int check_bc(hf_ctx* c, const hf_box* box, const int32_t bc[4], const double* inlet) {
    if (!bc) return fail(c, HF_EINVAL, "missing boundary codes");
    const bool two_d = bc[2] >= 0 && bc[3] >= 0;
    for (int k = 0; k < 4; ++k) {
        if (bc[k] > HF_BC_KEEP || (k < 2 && bc[k] < 0)) return fail(c, HF_EINVAL, "unknown boundary code");
        if (bc[k] == HF_BC_WEDGE && (k != 2 || !two_d))
            return fail(c, HF_EINVAL, "the wedge boundary exists only on the y-lo side of a 2-D box");
        if (bc[k] == HF_BC_INFLOW && !inlet) return fail(c, HF_EINVAL, "an inflow boundary needs the inlet state");
    }
    if (bc[2] == HF_BC_WEDGE && box && box->ny < 2)
        return fail(c, HF_EINVAL, "a wedge bottom needs at least two rows");
    return HF_OK;
}

void set_fill(BoxLaunch& L, double* U, const int32_t bc[4], const double* inlet) {
    L.U = U;
    L.op = BoxLaunch::FILL;
    for (int k = 0; k < 4; ++k) L.bc[k] = bc[k];
    L.two_d = bc[2] >= 0 && bc[3] >= 0;
    L.inlet = inlet;
}
}  // namespace

int hf_fill_ghosts_flow(hf_ctx* c, double* U, const hf_box* box, const int32_t bc[4], const double* inlet) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    const int rb = check_bc(c, box, bc, inlet);
    if (rb != HF_OK) return rb;
    BoxLaunch L{};
    const int r = box_common(c, L, box, U);
    if (r != HF_OK) return r;
    set_fill(L, U, bc, inlet);
    if (c->ops->box_ops(L) != HF_OK) return fail(c, HF_EDEVICE, "ghost fill launch failed");
    HF_HIP(c, hipStreamSynchronize(c->stream));
    return HF_OK;
}

int hf_euler_stage(hf_ctx* c, double* Uin, double* Uout, const hf_box* box, const int32_t bc[4], const double* inlet,
                   double dx, double dy, int64_t* repairs, int32_t* first_order, int64_t* first_failed) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    if (repairs) *repairs = 0;
    if (first_order) *first_order = 0;
    if (first_failed) *first_failed = -1;
    const int rb = check_bc(c, box, bc, inlet);
    if (rb != HF_OK) return rb;
    if (!(dx > 0) || !(dy >= 0)) return fail(c, HF_EINVAL, "invalid cell size");
    BoxLaunch L{};
    const int r = box_common(c, L, box, Uin);
    if (r != HF_OK) return r;
    const hf_box& B = *box;
    if (!Uout || !is_device_ptr(Uout)) return fail(c, HF_EINVAL, "the stage needs device memory");
    if (B.out_pitch < B.nx || B.out_plane < B.out_pitch * B.ny || B.out_offset < 0 || !std::isfinite(B.dtdx) ||
        !std::isfinite(B.dtdy))
        return fail(c, HF_EINVAL, "invalid output extent or step");
    const int ns = c->ops->ns;
    // ghost fill of Uin
    set_fill(L, Uin, bc, inlet);
    if (c->ops->box_ops(L) != HF_OK) return fail(c, HF_EDEVICE, "ghost fill launch failed");
    // flux tiles of the box (device copies shared by both attempts)
    const bool one_d = B.ny == 1 && B.dtdy == 0;
    const int TX = one_d ? c->ops->flux_tx1 : c->ops->flux_tx2, TY = one_d ? 1 : c->ops->flux_ty2;
    std::vector<int32_t> tiles;
    for (int tj = 0; tj < (B.ny + TY - 1) / TY; ++tj)
        for (int ti = 0; ti < (B.nx + TX - 1) / TX; ++ti) {
            tiles.push_back(0);
            tiles.push_back(ti);
            tiles.push_back(tj);
        }
    if (!c->boxes.ensure(sizeof(hf_box)) || !c->tiles.ensure(sizeof(int32_t) * tiles.size()))
        return fail(c, HF_ENOMEM, "device scratch");
    HF_HIP(c, hipMemcpyAsync(c->boxes.p, &B, sizeof(hf_box), hipMemcpyHostToDevice, c->stream));
    HF_HIP(c, hipMemcpyAsync(c->tiles.p, tiles.data(), sizeof(int32_t) * tiles.size(), hipMemcpyHostToDevice,
                             c->stream));
    unsigned long long* scal = c->scal.as<unsigned long long>();
    // status block read back once per attempt: [0] the failure index shared by
    // the repair's material-negative test and the validation, [8 + 3 ns] the
    // repair count, [8 + 3 ns + 1] the repair's donor failure
    std::vector<unsigned long long> hv(8 + 3 * ns + 2);
    for (int fo = 0; fo < 2; ++fo) {
        // flux stage: its status words are not read here (the repair and the
        // validation judge the raw update), so they need no reset
        FluxLaunch F{};
        F.mech = c->d_mech;
        F.boxes = c->boxes.as<hf_box>();
        F.nbox = 1;
        F.tiles = c->tiles.as<int32_t>();
        F.ntiles = (int)(tiles.size() / 3);
        F.Uin = Uin;
        F.Uout = Uout;
        F.first_order = fo;
        F.one_d = one_d ? 1 : 0;
        F.first_failed = scal;
        F.counters = scal + 1;
        F.skip_validation = 1;             // the repair and validation kernels below judge the update
        F.fluxes = nullptr;
        F.stream = c->stream;
        F.kernel_ms = nullptr;
        const int rf = c->ops->flux(F);
        if (rf != HF_OK) return fail(c, rf, "flux kernel launch failed");
        // repair (its init kernel resets the failure index), then the validation
        // on the same index; the validation's CFL rate lands in a slot the
        // repair no longer needs
        BoxLaunch R = L;
        R.U = Uout;
        R.op = BoxLaunch::REPAIR;
        if (c->ops->box_ops(R) != HF_OK) return fail(c, HF_EDEVICE, "roundoff repair launch failed");
        BoxLaunch V = L;
        V.U = Uout;
        V.op = BoxLaunch::PROPS;
        V.dx = dx;
        V.dy = dy;
        if (c->ops->box_ops(V) != HF_OK) return fail(c, HF_EDEVICE, "validation launch failed");
        HF_HIP(c, hipMemcpyAsync(hv.data(), scal, sizeof(unsigned long long) * hv.size(), hipMemcpyDeviceToHost,
                                 c->stream));
        HF_HIP(c, hipStreamSynchronize(c->stream));
        const unsigned long long bad = hv[0];
        const bool donor_failed = hv[8 + 3 * ns + 1] != 0;
        if (bad == ~0ull && !donor_failed) {
            if (repairs) *repairs = (int64_t)hv[8 + 3 * ns];
            return HF_OK;
        }
        if (first_order) *first_order = 1;
        if (fo == 1) {
            const unsigned long long at = bad == ~0ull ? 0ull : bad;
            if (first_failed) *first_failed = (int64_t)at;
            return fail(c, HF_EDOMAIN, donor_failed ? std::string("insufficient positive donor for roundoff "
                                                                  "redistribution after the first-order stage")
                                                    : "inadmissible state after the first-order stage at cell " +
                                                          std::to_string(at));
        }
    }
    return HF_OK;
}

namespace {
// Cells (i0..i1, j0..j1) of every component of a patch lie inside the
// device allocation of p.
bool patch_in_allocation(const double* p, const hf_patch& P, int nv, int64_t i0, int64_t j0, int64_t i1, int64_t j1) {
    return within_allocation(p, P.offset + j0 * P.pitch + i0, P.offset + (nv - 1) * P.plane + j1 * P.pitch + i1);
}

// The cells an operation reads and writes, against their allocations:
// INTERP (coarse: grow(coarsen(region), 1), fine: region), AVGDOWN (fine:
// refine(region), coarse: region), REFLUX (coarse: region grown by 1; the
// face fluxes of its faces and of their fine faces).
int check_extents(hf_ctx* c, int op, const double* src, const hf_patch& sp, const double* dst, const hf_patch& dp,
                  const int32_t* g, const hf_face_fluxes* cf = nullptr, const hf_face_fluxes* ff = nullptr) {
    const int nv = c->mech.ns() + 3;
    bool ok = true;
    if (op == BoxLaunch::INTERP) {
        ok = patch_in_allocation(src, sp, nv, (g[0] >> 1) - 1, (g[1] >> 1) - 1, (g[2] >> 1) + 1, (g[3] >> 1) + 1) &&
             patch_in_allocation(dst, dp, nv, g[0], g[1], g[2], g[3]);
    } else if (op == BoxLaunch::AVGDOWN) {
        ok = patch_in_allocation(src, sp, nv, 2 * g[0], 2 * g[1], 2 * g[2] + 1, 2 * g[3] + 1) &&
             patch_in_allocation(dst, dp, nv, g[0], g[1], g[2], g[3]);
    } else if (op == BoxLaunch::REFLUX_CELLS) {
        auto faces = [&](const hf_face_fluxes& F, int64_t i0, int64_t j0, int64_t i1, int64_t j1) {
            return within_allocation(F.fx, F.fx_offset + j0 * F.fx_pitch + i0,
                                     F.fx_offset + (nv - 1) * F.fx_plane + j1 * F.fx_pitch + i1 + 1) &&
                   within_allocation(F.fy, F.fy_offset + j0 * F.fy_pitch + i0,
                                     F.fy_offset + (nv - 1) * F.fy_plane + (j1 + 1) * F.fy_pitch + i1);
        };
        ok = patch_in_allocation(dst, dp, nv, g[0], g[1], g[2], g[3]) && faces(*cf, g[0], g[1], g[2], g[3]) &&
             faces(*ff, 2 * g[0], 2 * g[1], 2 * g[2] + 1, 2 * g[3] + 1);
    }
    return ok ? HF_OK : fail(c, HF_EINVAL, "a patch, region or face-flux array lies outside its allocation");
}

int patch_common(hf_ctx* c, BoxLaunch& L, const double* src, const hf_patch* sp, double* dst, const hf_patch* dp,
                 const int32_t region[4]) {
    if (!src || !sp || !dst || !dp || !region) return fail(c, HF_EINVAL, "missing patch, array or region");
    if (region[2] < region[0] || region[3] < region[1]) return fail(c, HF_EINVAL, "empty region");
    if (sp->pitch <= 0 || sp->plane <= 0 || dp->pitch <= 0 || dp->plane <= 0)
        return fail(c, HF_EINVAL, "invalid patch stride");
    HF_DEVICE(c);
    if (!is_device_ptr(src) || !is_device_ptr(dst)) return fail(c, HF_EINVAL, "patch operations need device memory");
    if (!c->scal.ensure(sizeof(unsigned long long) * 256)) return fail(c, HF_ENOMEM, "device scalars");
    L.mech = c->d_mech;
    L.src = src;
    L.dst = dst;
    L.sp = *sp;
    L.dp = *dp;
    for (int k = 0; k < 4; ++k) L.region[k] = region[k];
    L.scal = c->scal.as<unsigned long long>() + 8;
    L.first_failed = c->scal.as<unsigned long long>();
    L.stream = c->stream;
    return HF_OK;
}
}  // namespace

int hf_interp_coarse_to_fine(hf_ctx* c, const double* coarse, const hf_patch* cp, double* fine, const hf_patch* fp,
                             const int32_t region[4], int64_t* limited, int64_t* first_failed) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    if (limited) *limited = 0;
    if (first_failed) *first_failed = -1;
    BoxLaunch L{};
    const int r = patch_common(c, L, coarse, cp, fine, fp, region);
    if (r != HF_OK) return r;
    if (check_extents(c, BoxLaunch::INTERP, coarse, *cp, fine, *fp, region) != HF_OK) return HF_EINVAL;
    L.op = BoxLaunch::INTERP;
    if (c->ops->box_ops(L) != HF_OK) return fail(c, HF_EDEVICE, "interpolation launch failed");
    unsigned long long h[2] = {0, ~0ull};
    HF_HIP(c, hipMemcpyAsync(&h[0], L.scal, sizeof(h[0]), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipMemcpyAsync(&h[1], L.first_failed, sizeof(h[1]), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipStreamSynchronize(c->stream));
    if (limited) *limited = (int64_t)h[0];
    if (h[1] != ~0ull) {
        if (first_failed) *first_failed = (int64_t)h[1];
        return fail(c, HF_EDOMAIN, "inadmissible coarse parent " + std::to_string(h[1]) + " in interpolation");
    }
    return HF_OK;
}

int hf_reflux(hf_ctx* c, double* Uc, const hf_patch* cp, const hf_face_fluxes* coarse, const hf_face_fluxes* fine,
              const int32_t fine_box[4], int32_t sides, double h, double dxc, double dyc) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    if (!coarse || !fine || !fine_box) return fail(c, HF_EINVAL, "missing fluxes or fine box");
    if (sides < 0 || sides > 15) return fail(c, HF_EINVAL, "invalid side mask");
    if ((fine_box[0] & 1) || (fine_box[1] & 1) || !(fine_box[2] & 1) || !(fine_box[3] & 1))
        return fail(c, HF_EINVAL, "the fine box must be aligned to the refinement ratio 2");
    if (!(h > 0) || !(dxc > 0) || !(dyc > 0) || !std::isfinite(h) || !std::isfinite(dxc) || !std::isfinite(dyc))
        return fail(c, HF_EINVAL, "invalid step or cell size");
    for (const hf_face_fluxes* f : {coarse, fine})
        if (!f->fx || !f->fy || !is_device_ptr(f->fx) || !is_device_ptr(f->fy))
            return fail(c, HF_EINVAL, "face fluxes need device memory");
    if (sides == 0) return HF_OK;
    BoxLaunch L{};
    const int r = patch_common(c, L, Uc, cp, Uc, cp, fine_box);
    if (r != HF_OK) return r;
    L.op = BoxLaunch::REFLUX;
    L.cflux = *coarse;
    L.fflux = *fine;
    L.sides = sides;
    L.h = h;
    L.dx = dxc;
    L.dy = dyc;
    if (c->ops->box_ops(L) != HF_OK) return fail(c, HF_EDEVICE, "reflux launch failed");
    HF_HIP(c, hipStreamSynchronize(c->stream));
    return HF_OK;
}

int hf_average_down(hf_ctx* c, const double* fine, const hf_patch* fp, double* coarse, const hf_patch* cp,
                    const int32_t cregion[4]) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    BoxLaunch L{};
    const int r = patch_common(c, L, fine, fp, coarse, cp, cregion);
    if (r != HF_OK) return r;
    if (check_extents(c, BoxLaunch::AVGDOWN, fine, *fp, coarse, *cp, cregion) != HF_OK) return HF_EINVAL;
    L.op = BoxLaunch::AVGDOWN;
    if (c->ops->box_ops(L) != HF_OK) return fail(c, HF_EDEVICE, "average-down launch failed");
    HF_HIP(c, hipStreamSynchronize(c->stream));
    return HF_OK;
}

namespace {
// A device copy of a block table and its cell-id prefix sums in L (the
// context's block buffer), after checking every block.
int cells_common(hf_ctx* c, BoxLaunch& L, const hf_cells* blocks, int nblock, std::vector<int64_t>& start) {
    if (nblock < 1 || !blocks) return fail(c, HF_EINVAL, "invalid block table");
    HF_DEVICE(c);
    start.assign(1, 0);
    for (int b = 0; b < nblock; ++b) {
        const hf_cells& B = blocks[b];
        if (B.nx < 1 || B.ny < 1 || !B.U || B.pitch < B.nx || B.plane < B.pitch * B.ny || B.offset < 0 ||
            (B.mask && B.mask_pitch < B.nx) || !(B.dx > 0) || !(B.dy >= 0))
            return fail(c, HF_EINVAL, "block " + std::to_string(b) + " has an invalid extent, stride or cell size");
        if (!is_device_ptr(B.U) || (B.mask && !is_device_ptr(B.mask)))
            return fail(c, HF_EINVAL, "block " + std::to_string(b) + " is not in device memory");
        start.push_back(start.back() + (int64_t)B.nx * B.ny);
    }
    const size_t tb = sizeof(hf_cells) * nblock, sb = sizeof(int64_t) * (nblock + 1);
    if (!c->blocks.ensure(tb + sb) || !c->scal.ensure(sizeof(unsigned long long) * 256))
        return fail(c, HF_ENOMEM, "device scratch");
    HF_HIP(c, hipMemcpyAsync(c->blocks.p, blocks, tb, hipMemcpyHostToDevice, c->stream));
    HF_HIP(c, hipMemcpyAsync(c->blocks.as<char>() + tb, start.data(), sb, hipMemcpyHostToDevice, c->stream));
    L.mech = c->d_mech;
    L.blocks = c->blocks.as<hf_cells>();
    L.block_start = reinterpret_cast<const int64_t*>(c->blocks.as<char>() + tb);
    L.nblock = nblock;
    L.total = start.back();
    L.scal = c->scal.as<unsigned long long>() + 8;
    L.first_failed = c->scal.as<unsigned long long>();
    L.stream = c->stream;
    return HF_OK;
}

std::string cell_name(const std::vector<int64_t>& start, const hf_cells* blocks, unsigned long long code) {
    const int b = (int)(code >> 32);
    const int64_t k = (int64_t)(code & 0xffffffffull);
    (void)start;
    return "block " + std::to_string(b) + " cell (" + std::to_string(k % blocks[b].nx) + ", " +
           std::to_string(k / blocks[b].nx) + ")";
}
}  // namespace

int hf_validate_cells(hf_ctx* c, const hf_cells* blocks, int nblock, double* rate, int64_t* first_failed) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    if (first_failed) *first_failed = -1;
    if (rate) *rate = 0;
    BoxLaunch L{};
    std::vector<int64_t> start;
    const int r = cells_common(c, L, blocks, nblock, start);
    if (r != HF_OK) return r;
    L.op = BoxLaunch::CELLS_PROPS;
    unsigned long long init[2] = {~0ull, 0};
    HF_HIP(c, hipMemcpyAsync(L.first_failed, &init[0], sizeof(init[0]), hipMemcpyHostToDevice, c->stream));
    HF_HIP(c, hipMemcpyAsync(L.scal, &init[1], sizeof(init[1]), hipMemcpyHostToDevice, c->stream));
    if (c->ops->box_ops(L) != HF_OK) return fail(c, HF_EDEVICE, "validation launch failed");
    unsigned long long out[2];
    HF_HIP(c, hipMemcpyAsync(&out[0], L.first_failed, sizeof(out[0]), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipMemcpyAsync(&out[1], L.scal, sizeof(out[1]), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipStreamSynchronize(c->stream));
    if (out[0] != ~0ull) {
        if (first_failed) *first_failed = (int64_t)out[0];
        return fail(c, HF_EDOMAIN, "inadmissible state at " + cell_name(start, blocks, out[0]));
    }
    if (rate) std::memcpy(rate, &out[1], sizeof(double));
    return HF_OK;
}

int hf_repair_cells(hf_ctx* c, const hf_cells* blocks, int nblock, int64_t* repairs, int64_t* first_failed) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    if (first_failed) *first_failed = -1;
    if (repairs) *repairs = 0;
    BoxLaunch L{};
    std::vector<int64_t> start;
    const int r = cells_common(c, L, blocks, nblock, start);
    if (r != HF_OK) return r;
    const int ns = c->ops->ns;
    const int64_t groups = (L.total + 255) / 256;
    if (!c->aux[5].ensure(sizeof(double) * ns * groups)) return fail(c, HF_ENOMEM, "device scratch");
    L.partials = c->aux[5].as<double>();
    L.op = BoxLaunch::CELLS_REPAIR;
    if (c->ops->box_ops(L) != HF_OK) return fail(c, HF_EDEVICE, "roundoff repair launch failed");
    std::vector<unsigned long long> h(3 * ns + 2);
    unsigned long long ff = ~0ull;
    HF_HIP(c, hipMemcpyAsync(h.data(), L.scal, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipMemcpyAsync(&ff, L.first_failed, sizeof(ff), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipStreamSynchronize(c->stream));
    if (repairs) *repairs = (int64_t)h[3 * ns];
    if (ff != ~0ull) {
        if (first_failed) *first_failed = (int64_t)ff;
        return fail(c, HF_EDOMAIN, "materially negative species after composite flow stage at " +
                                       cell_name(start, blocks, ff));
    }
    if (h[3 * ns + 1]) return fail(c, HF_EDOMAIN, "insufficient positive donor for roundoff redistribution");
    return HF_OK;
}

namespace {
// Integrals over the leaf cells of a block table: per workgroup, a fixed-order
// sum per component (butterfly within each wave, then the four waves in
// order) into partials[group * nv + n]
__global__ void __launch_bounds__(256) totals_partial_kernel(CellMap S, int64_t total, int nv, int absolute,
                                                             double* __restrict__ partials) {
    __shared__ double part[4][64];
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double* q = nullptr;
    int64_t plane = 0;
    double vol = 0;
    if (c < total) {
        int b, i, j;
        const hf_cells& B = S.find(c, b, i, j);
        if (!B.mask || B.mask[(int64_t)j * B.mask_pitch + i] == 1) {
            q = B.U + B.offset + (int64_t)j * B.pitch + i;
            plane = B.plane;
            vol = B.dy > 0 ? B.dx * B.dy : B.dx;
        }
    }
    for (int n = 0; n < nv; ++n) {
        double x = 0.0;
        if (q) {
            const double v = q[(int64_t)n * plane];
            x = (absolute ? fabs(v) : v) * vol;
        }
        #pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6][n] = x;
    }
    __syncthreads();
    if (threadIdx.x < nv) {
        const int n = threadIdx.x;
        partials[(int64_t)blockIdx.x * nv + n] = ((part[0][n] + part[1][n]) + part[2][n]) + part[3][n];
    }
}

// one workgroup per component: the workgroup partials summed in a fixed order
__global__ void __launch_bounds__(256) totals_final_kernel(const double* __restrict__ partials, int ngroups, int nv,
                                                           double* __restrict__ out) {
    __shared__ double wsum[4];
    const int n = blockIdx.x;
    double x = 0.0;
    for (int g = threadIdx.x; g < ngroups; g += 256) x += partials[(int64_t)g * nv + n];
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) out[n] = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
}
}  // namespace

int hf_totals_cells(hf_ctx* c, const hf_cells* blocks, int nblock, int absolute, double* totals) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    if (!totals) return fail(c, HF_EINVAL, "missing output");
    const int nv = c->mech.ns() + 3;
    for (int n = 0; n < nv; ++n) totals[n] = 0.0;
    if (nv > 64) return fail(c, HF_EINVAL, "mechanism too large for the totals reduction");
    BoxLaunch L{};
    std::vector<int64_t> start;
    const int r = cells_common(c, L, blocks, nblock, start);
    if (r != HF_OK) return r;
    const int64_t groups = (L.total + 255) / 256;
    if (groups > INT32_MAX || !c->aux[4].ensure(sizeof(double) * (nv * groups + 64)))
        return fail(c, HF_ENOMEM, "device scratch");
    double* part = c->aux[4].as<double>();
    double* dout = part + nv * groups;
    const CellMap S{nullptr, 0, L.blocks, L.block_start, L.nblock};
    hipLaunchKernelGGL(totals_partial_kernel, dim3((unsigned)groups), dim3(256), 0, c->stream, S, L.total, nv,
                       absolute ? 1 : 0, part);
    hipLaunchKernelGGL(totals_final_kernel, dim3((unsigned)nv), dim3(256), 0, c->stream, part, (int)groups, nv, dout);
    if (hipGetLastError() != hipSuccess) return fail(c, HF_EDEVICE, "totals launch failed");
    HF_HIP(c, hipMemcpyAsync(totals, dout, sizeof(double) * nv, hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipStreamSynchronize(c->stream));
    return HF_OK;
}

int hf_interp_regions(hf_ctx* c, const double* coarse, const hf_patch* cp, double* fine, const hf_patch* fp,
                      const int32_t* regions, int nregion, const uint8_t* fine_mask, int64_t fine_mask_pitch,
                      const int32_t fine_extent[2], int64_t* limited, int64_t* first_failed) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    if (limited) *limited = 0;
    if (first_failed) *first_failed = -1;
    if (nregion < 0 || (nregion > 0 && !regions)) return fail(c, HF_EINVAL, "invalid region list");
    if (nregion == 0) return HF_OK;
    if (fine_mask && (!fine_extent || fine_extent[0] < 1 || fine_extent[1] < 1 || fine_mask_pitch < fine_extent[0]))
        return fail(c, HF_EINVAL, "invalid fine mask extent");
    int64_t most = 0;
    for (int k = 0; k < nregion; ++k) {
        const int32_t* g = regions + 4 * k;
        if (g[2] < g[0] || g[3] < g[1]) return fail(c, HF_EINVAL, "empty region " + std::to_string(k));
        most = std::max<int64_t>(most, (int64_t)((g[2] >> 1) - (g[0] >> 1) + 1) * ((g[3] >> 1) - (g[1] >> 1) + 1));
    }
    BoxLaunch L{};
    const int r = patch_common(c, L, coarse, cp, fine, fp, regions);
    if (r != HF_OK) return r;
    for (int k = 0; k < nregion; ++k)
        if (check_extents(c, BoxLaunch::INTERP, coarse, *cp, fine, *fp, regions + 4 * k) != HF_OK) return HF_EINVAL;
    if (fine_mask && !is_device_ptr(fine_mask)) return fail(c, HF_EINVAL, "the fine mask needs device memory");
    if (!c->tiles.ensure(sizeof(int32_t) * 4 * nregion)) return fail(c, HF_ENOMEM, "device scratch");
    HF_HIP(c, hipMemcpyAsync(c->tiles.p, regions, sizeof(int32_t) * 4 * nregion, hipMemcpyHostToDevice, c->stream));
    L.op = BoxLaunch::INTERP;
    L.regions = c->tiles.as<int32_t>();
    L.nregion = nregion;
    L.total = most;
    L.mask = fine_mask;
    L.mask_pitch = fine_mask_pitch;
    L.mask_nx = fine_mask ? fine_extent[0] : 0;
    L.mask_ny = fine_mask ? fine_extent[1] : 0;
    if (c->ops->box_ops(L) != HF_OK) return fail(c, HF_EDEVICE, "interpolation launch failed");
    unsigned long long h[2] = {0, ~0ull};
    HF_HIP(c, hipMemcpyAsync(&h[0], L.scal, sizeof(h[0]), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipMemcpyAsync(&h[1], L.first_failed, sizeof(h[1]), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipStreamSynchronize(c->stream));
    if (limited) *limited = (int64_t)h[0];
    if (h[1] != ~0ull) {
        if (first_failed) *first_failed = (int64_t)h[1];
        return fail(c, HF_EDOMAIN, "inadmissible coarse parent " + std::to_string(h[1] & 0xffffffffull) +
                                       " of region " + std::to_string(h[1] >> 32) + " in interpolation");
    }
    return HF_OK;
}

int hf_average_down_masked(hf_ctx* c, const double* fine, const hf_patch* fp, double* coarse, const hf_patch* cp,
                           const uint8_t* coarse_mask, int64_t coarse_mask_pitch, const int32_t cregion[4]) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    if (!coarse_mask || !cregion || cregion[0] < 0 || cregion[1] < 0 || coarse_mask_pitch <= cregion[2])
        return fail(c, HF_EINVAL, "invalid coarse mask or region");
    BoxLaunch L{};
    const int r = patch_common(c, L, fine, fp, coarse, cp, cregion);
    if (r != HF_OK) return r;
    if (check_extents(c, BoxLaunch::AVGDOWN, fine, *fp, coarse, *cp, cregion) != HF_OK) return HF_EINVAL;
    if (!is_device_ptr(coarse_mask)) return fail(c, HF_EINVAL, "the coarse mask needs device memory");
    L.op = BoxLaunch::AVGDOWN;
    L.mask = coarse_mask;
    L.mask_pitch = coarse_mask_pitch;
    if (c->ops->box_ops(L) != HF_OK) return fail(c, HF_EDEVICE, "average-down launch failed");
    HF_HIP(c, hipStreamSynchronize(c->stream));
    return HF_OK;
}

int hf_reflux_masked(hf_ctx* c, double* Uc, const hf_patch* cp, const hf_face_fluxes* coarse,
                     const hf_face_fluxes* fine, const uint8_t* coarse_mask, int64_t coarse_mask_pitch,
                     const int32_t cregion[4], double h, double dxc, double dyc) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    if (!coarse || !fine || !coarse_mask || !cregion) return fail(c, HF_EINVAL, "missing fluxes, mask or region");
    if (cregion[0] < 0 || cregion[1] < 0 || coarse_mask_pitch <= cregion[2])
        return fail(c, HF_EINVAL, "invalid coarse mask or region");
    if (!(h > 0) || !(dxc > 0) || !(dyc > 0) || !std::isfinite(h) || !std::isfinite(dxc) || !std::isfinite(dyc))
        return fail(c, HF_EINVAL, "invalid step or cell size");
    for (const hf_face_fluxes* f : {coarse, fine})
        if (!f->fx || !f->fy || !is_device_ptr(f->fx) || !is_device_ptr(f->fy))
            return fail(c, HF_EINVAL, "face fluxes need device memory");
    BoxLaunch L{};
    const int r = patch_common(c, L, Uc, cp, Uc, cp, cregion);
    if (r != HF_OK) return r;
    if (check_extents(c, BoxLaunch::REFLUX_CELLS, Uc, *cp, Uc, *cp, cregion, coarse, fine) != HF_OK) return HF_EINVAL;
    if (!is_device_ptr(coarse_mask)) return fail(c, HF_EINVAL, "the coarse mask needs device memory");
    L.op = BoxLaunch::REFLUX_CELLS;
    L.cflux = *coarse;
    L.fflux = *fine;
    L.mask = coarse_mask;
    L.mask_pitch = coarse_mask_pitch;
    L.h = h;
    L.dx = dxc;
    L.dy = dyc;
    if (c->ops->box_ops(L) != HF_OK) return fail(c, HF_EDEVICE, "reflux launch failed");
    HF_HIP(c, hipStreamSynchronize(c->stream));
    return HF_OK;
}

int hf_tag_cells(hf_ctx* c, const double* U, const hf_box* box, const hf_tag_params* prm, uint8_t* tags,
                 int64_t* ntagged, int64_t* first_failed) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    if (ntagged) *ntagged = 0;
    if (first_failed) *first_failed = -1;
    if (!prm || !tags) return fail(c, HF_EINVAL, "missing tagging parameters or output");
    if (prm->fuel < 0 || prm->fuel >= c->mech.ns()) return fail(c, HF_EINVAL, "fuel species index out of range");
    if (!(prm->dx > 0) || !(prm->dy > 0)) return fail(c, HF_EINVAL, "invalid cell size");
    BoxLaunch L{};
    const int r = box_common(c, L, box, U);
    if (r != HF_OK) return r;
    if (!is_device_ptr(tags)) return fail(c, HF_EINVAL, "tags need device memory");
    L.U = const_cast<double*>(U);
    L.op = BoxLaunch::TAG;
    L.tag = *prm;
    L.tags = tags;
    if (c->ops->box_ops(L) != HF_OK) return fail(c, HF_EDEVICE, "tagging launch failed");
    unsigned long long h[2] = {0, ~0ull};
    HF_HIP(c, hipMemcpyAsync(&h[0], L.scal, sizeof(h[0]), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipMemcpyAsync(&h[1], L.first_failed, sizeof(h[1]), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipStreamSynchronize(c->stream));
    if (ntagged) *ntagged = (int64_t)h[0];
    if (h[1] != ~0ull) {
        if (first_failed) *first_failed = (int64_t)h[1];
        return fail(c, HF_EDOMAIN, "inadmissible state while tagging at cell " + std::to_string(h[1]));
    }
    return HF_OK;
}

// ------------------------------------------------------------------ a whole Strang step as one graph
namespace {

// Step scalars (device, 64-bit words): [0] CFL rate bits, [1] dt, [2] h =
// dt/2, [3] time left to the target (<= 0: none), [4] status (bit 0 a flow
// stage or validation rejected, bit 1 the chemistry failed), [5] repairs,
// [8] chemistry failure index, [9..16] chemistry counters.
enum { SS_RATE = 0, SS_DT = 1, SS_H = 2, SS_LEFT = 3, SS_STATUS = 4, SS_REPAIRS = 5, SS_CHEM = 8, SS_WORDS = 32 };

// validation reset: the shared failure index and the CFL-rate slot
__global__ void step_props_reset(unsigned long long* ff, unsigned long long* rate) {
    if (threadIdx.x == 0) {
        *ff = ~0ull;
        *rate = 0;
    }
}

// dt = cfl / rate (min with the time left), h = dt / 2 (BlockAMR::run and
// stableDt, BlockAMR.cpp:335-346, 409); the status starts from the initial
// validation
__global__ void step_dt(unsigned long long* s, const unsigned long long* ff, const unsigned long long* rate,
                        double cfl) {
    if (threadIdx.x != 0) return;
    double* d = reinterpret_cast<double*>(s);
    double dt = cfl / __longlong_as_double((long long)*rate);
    const double left = d[SS_LEFT];
    if (left > 0) dt = fmin(left, dt);
    d[SS_DT] = dt;
    d[SS_H] = 0.5 * dt;
    s[SS_STATUS] = *ff != ~0ull ? 1ull : 0ull;
    s[SS_REPAIRS] = 0;
}

// the flux stage's h/dx and h/dy from the device h, in the box table
__global__ void step_set_h(hf_box* box, const unsigned long long* s, double dx, double dy, int one_d) {
    if (threadIdx.x != 0) return;
    const double h = reinterpret_cast<const double*>(s)[SS_H];
    box->dtdx = h / dx;
    box->dtdy = one_d ? 0.0 : h / dy;
}

// fold a repair + validation (or the chemistry) into the step status
__global__ void step_fold(unsigned long long* s, const unsigned long long* ff, const unsigned long long* donor_fail,
                          const unsigned long long* repairs, unsigned long long bit) {
    if (threadIdx.x != 0) return;
    if (*ff != ~0ull || (donor_fail && *donor_fail)) s[SS_STATUS] |= bit;
    if (repairs) s[SS_REPAIRS] += *repairs;
}

__global__ void __launch_bounds__(256) step_copy(double* __restrict__ dst, const double* __restrict__ src, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// out = a y + b x, the arithmetic of hf_lincomb
__global__ void __launch_bounds__(256) step_lincomb(double* __restrict__ out, const double* __restrict__ y,
                                                    const double* __restrict__ x, double a, double b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = fma(a, y[i], b * x[i]);
}

struct StepPlan {
    double *U, *U1, *U2, *old, *backup;
    hf_box box;
    int32_t bc[4];
    double inlet[64];
    int has_inlet;
    double dx, dy, cfl;
    int chem_mode, ntiles;
    int64_t n;                      // doubles per array
};

// Enqueue one Strang step on the context's stream: kernels only (no copies
// from the host, no synchronisation), so the sequence can be captured.
int step_enqueue(hf_ctx* c, const StepPlan& P) {
    const int ns = c->ops->ns;
    const bool one_d = P.box.ny == 1 && P.bc[2] < 0;
    auto* s = c->sscal.as<unsigned long long>();
    unsigned long long* scal = c->scal.as<unsigned long long>();     // box operations' words
    unsigned long long *ff = scal, *rate = scal + 8;
    const unsigned gcopy = (unsigned)std::min<int64_t>((P.n + 255) / 256, 8192);
    BoxLaunch L{};
    L.mech = c->d_mech;
    L.box = P.box;
    L.scal = scal + 8;
    L.first_failed = scal;
    L.stream = c->stream;
    L.dx = P.dx;
    L.dy = one_d ? 0.0 : P.dy;
    auto validate = [&](double* A) {
        hipLaunchKernelGGL(step_props_reset, dim3(1), dim3(64), 0, c->stream, ff, rate);
        BoxLaunch V = L;
        V.U = A;
        V.op = BoxLaunch::PROPS;
        return c->ops->box_ops(V);
    };
    auto stage = [&](double* src, double* dst) {
        BoxLaunch Fl = L;
        set_fill(Fl, src, P.bc, P.has_inlet ? P.inlet : nullptr);
        if (c->ops->box_ops(Fl) != HF_OK) return HF_EDEVICE;
        FluxLaunch F{};
        F.mech = c->d_mech;
        F.boxes = c->sbox.as<hf_box>();
        F.nbox = 1;
        F.tiles = c->stiles.as<int32_t>();
        F.ntiles = P.ntiles;
        F.Uin = src;
        F.Uout = dst;
        F.first_order = 0;
        F.one_d = one_d ? 1 : 0;
        F.first_failed = scal + 2;           // the raw update is judged by the repair and validation
        F.skip_validation = 1;
        F.counters = scal + 3;
        F.stream = c->stream;
        if (c->ops->flux(F) != HF_OK) return HF_EDEVICE;
        BoxLaunch R = L;
        R.U = dst;
        R.op = BoxLaunch::REPAIR;
        if (c->ops->box_ops(R) != HF_OK) return HF_EDEVICE;
        BoxLaunch V = L;
        V.U = dst;
        V.op = BoxLaunch::PROPS;
        if (c->ops->box_ops(V) != HF_OK) return HF_EDEVICE;
        hipLaunchKernelGGL(step_fold, dim3(1), dim3(64), 0, c->stream, s, ff, scal + 8 + 3 * ns + 1, scal + 8 + 3 * ns,
                           1ull);
        return HF_OK;
    };
    // dt from the CFL rate of the state
    if (validate(P.U) != HF_OK) return HF_EDEVICE;
    hipLaunchKernelGGL(step_dt, dim3(1), dim3(64), 0, c->stream, s, ff, rate, P.cfl);
    hipLaunchKernelGGL(step_copy, dim3(gcopy), dim3(256), 0, c->stream, P.backup, P.U, P.n);
    for (int half = 0; half < 2; ++half) {
        // flow(dt/2): Heun, U <- U/2 + E(E(U))/2, validated
        hipLaunchKernelGGL(step_copy, dim3(gcopy), dim3(256), 0, c->stream, P.old, P.U, P.n);
        hipLaunchKernelGGL(step_set_h, dim3(1), dim3(64), 0, c->stream, c->sbox.as<hf_box>(), s, P.dx, P.dy,
                           one_d ? 1 : 0);
        if (stage(P.U, P.U1) != HF_OK || stage(P.U1, P.U2) != HF_OK) return HF_EDEVICE;
        hipLaunchKernelGGL(step_lincomb, dim3(gcopy), dim3(256), 0, c->stream, P.U, P.U2, P.old, 0.5, 0.5, P.n);
        if (validate(P.U) != HF_OK) return HF_EDEVICE;
        hipLaunchKernelGGL(step_fold, dim3(1), dim3(64), 0, c->stream, s, ff, nullptr, nullptr, 1ull);
        if (half == 1) break;
        // react(dt): one batch over the box's cells, in place
        ChemLaunch C{};
        const int64_t ncell = (int64_t)P.box.nx * P.box.ny;
        C.mech = c->d_mech;
        C.ncell = ncell;
        C.ld = ncell;
        C.blocks = c->scells.as<hf_cells>();
        C.block_start = reinterpret_cast<const int64_t*>(c->scells.as<char>() + sizeof(hf_cells));
        C.nblock = 1;
        C.dt = reinterpret_cast<const double*>(s + SS_DT);
        C.dt_per_cell = 0;
        C.out = c->out.as<double>();
        C.ldo = ncell;
        C.status = c->status.as<int8_t>();
        C.first_failed = s + SS_CHEM;
        C.counters = s + SS_CHEM + 1;
        C.rtol = c->opt.rtol;
        C.atol = c->opt.atol;
        C.maxsteps = c->opt.max_steps;
        C.mode = P.chem_mode;
        C.precision = c->opt.precision;
        C.num_cus = c->num_cus;
        C.waves_per_cu = c->opt.waves_per_cu;
        C.chunk = c->opt.chunk > 0 ? c->opt.chunk : 8;
        C.nbins = c->opt.nbins;
        C.jac_wait = c->opt.jac_wait;
        C.setup_wait = c->opt.setup_wait;
        C.workspace = c->work.p;
        C.workspace_bytes = c->work.n;
        C.stream = c->stream;
        C.commit = P.U;
        C.commit_ld = ncell;
        if (c->ops->chem(C) != HF_OK) return HF_EDEVICE;
        hipLaunchKernelGGL(step_fold, dim3(1), dim3(64), 0, c->stream, s, s + SS_CHEM, nullptr, nullptr, 2ull);
    }
    return hipGetLastError() == hipSuccess ? HF_OK : HF_EDEVICE;
}

}  // namespace

int hf_box_strang_step(hf_ctx* c, double* U, double* U1, double* U2, double* old, double* backup, const hf_box* box,
                       const int32_t bc[4], const double* inlet, double dx, double dy, double cfl, int chem_mode,
                       double time_left, double* dt_taken, int32_t* taken, int64_t* repairs) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    if (taken) *taken = 0;
    if (dt_taken) *dt_taken = 0;
    if (repairs) *repairs = 0;
    const int rb = check_bc(c, box, bc, inlet);
    if (rb != HF_OK) return rb;
    if (!(dx > 0) || !(dy >= 0) || !(cfl > 0) || (chem_mode != HF_CHEM_PLAIN && chem_mode != HF_CHEM_BATCHED))
        return fail(c, HF_EINVAL, "invalid cell size, CFL number or chemistry form");
    BoxLaunch L0{};
    const int r = box_common(c, L0, box, U);
    if (r != HF_OK) return r;
    for (double* p : {U1, U2, old, backup})
        if (!p || !is_device_ptr(p)) return fail(c, HF_EINVAL, "the step's arrays need device memory");
    const int nv = c->mech.ns() + 3;
    if (nv > 64) return fail(c, HF_EINVAL, "mechanism too large for the step graph's inlet state");
    StepPlan P{};
    P.U = U;
    P.U1 = U1;
    P.U2 = U2;
    P.old = old;
    P.backup = backup;
    P.box = *box;
    P.box.out_offset = box->in_offset + 2 * box->in_pitch + 2;     // the stage writes the ghosted output's cells
    P.box.out_pitch = box->in_pitch;
    P.box.out_plane = box->in_plane;
    for (int k = 0; k < 4; ++k) P.bc[k] = bc[k];
    if (inlet) {
        P.has_inlet = 1;
        for (int n = 0; n < nv; ++n) P.inlet[n] = inlet[n];
    }
    P.dx = dx;
    P.dy = dy;
    P.cfl = cfl;
    P.chem_mode = chem_mode;
    P.n = box->in_offset + (int64_t)nv * box->in_plane;
    for (double* p : {U, U1, U2, old, backup})
        if (!within_allocation(p, 0, P.n - 1)) return fail(c, HF_EINVAL, "a step array is smaller than the box");
    const bool one_d = box->ny == 1 && bc[2] < 0;
    const int TX = one_d ? c->ops->flux_tx1 : c->ops->flux_tx2, TY = one_d ? 1 : c->ops->flux_ty2;
    std::vector<int32_t> tiles;
    for (int tj = 0; tj < (box->ny + TY - 1) / TY; ++tj)
        for (int ti = 0; ti < (box->nx + TX - 1) / TX; ++ti) {
            tiles.push_back(0);
            tiles.push_back(ti);
            tiles.push_back(tj);
        }
    P.ntiles = (int)(tiles.size() / 3);
    const int64_t ncell = (int64_t)box->nx * box->ny;
    const int ns = c->mech.ns();
    const size_t wbytes = c->ops->chem_workspace(ncell, c->opt.nbins, c->num_cus, c->opt.waves_per_cu);
    hf_cells cells{U, box->in_offset + 2 * box->in_pitch + 2, box->in_pitch, box->in_plane, box->nx, box->ny,
                   nullptr, 0, dx, dy};
    const int64_t start[2] = {0, ncell};
    if (!c->sbox.ensure(sizeof(hf_box)) || !c->stiles.ensure(sizeof(int32_t) * tiles.size()) ||
        !c->scells.ensure(sizeof(hf_cells) + sizeof(start)) || !c->sscal.ensure(8 * SS_WORDS) ||
        !c->scal.ensure(sizeof(unsigned long long) * 256) || !c->out.ensure(sizeof(double) * ns * ncell) ||
        !c->status.ensure(ncell) || (wbytes && !c->work.ensure(wbytes)))
        return fail(c, HF_ENOMEM, "device scratch");
    // a new key (arrays, box, conditions, options, and the context buffers the
    // kernels address, which other calls may have grown) captures a new graph
    std::vector<uint64_t> key(sizeof(StepPlan) / 8 + 1, 0);
    std::memcpy(key.data(), &P, sizeof(StepPlan));
    auto bits = [](double x) {
        uint64_t u;
        std::memcpy(&u, &x, sizeof(u));
        return u;
    };
    for (uint64_t v : {bits(c->opt.rtol), bits(c->opt.atol), (uint64_t)c->opt.max_steps, (uint64_t)c->opt.precision,
                       (uint64_t)c->opt.waves_per_cu, (uint64_t)c->opt.chunk, (uint64_t)c->opt.nbins,
                       (uint64_t)c->opt.jac_wait, (uint64_t)c->opt.setup_wait, (uint64_t)c->opt.eager})
        key.push_back(v);
    for (const DevBuf* b : {&c->sbox, &c->stiles, &c->scells, &c->sscal, &c->scal, &c->out, &c->status, &c->work}) {
        key.push_back((uint64_t)(uintptr_t)b->p);
        key.push_back((uint64_t)b->n);
    }
    ChemGraph& g = c->step_graph;
    const bool fresh = !g.exec || g.key != key;
    if (fresh) {
        g.reset();
        HF_HIP(c, hipMemcpyAsync(c->sbox.p, &P.box, sizeof(hf_box), hipMemcpyHostToDevice, c->stream));
        HF_HIP(c, hipMemcpyAsync(c->stiles.p, tiles.data(), sizeof(int32_t) * tiles.size(), hipMemcpyHostToDevice,
                                 c->stream));
        HF_HIP(c, hipMemcpyAsync(c->scells.p, &cells, sizeof(cells), hipMemcpyHostToDevice, c->stream));
        HF_HIP(c, hipMemcpyAsync(c->scells.as<char>() + sizeof(hf_cells), start, sizeof(start), hipMemcpyHostToDevice,
                                 c->stream));
    }
    double left = time_left;
    HF_HIP(c, hipMemcpyAsync(c->sscal.as<char>() + 8 * SS_LEFT, &left, 8, hipMemcpyHostToDevice, c->stream));
    int how = 0;
    if (c->opt.eager) {
        if (step_enqueue(c, P) != HF_OK) return fail(c, HF_EDEVICE, "step launch failed");
    } else {
        if (fresh) {
            if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return fail(c, HF_EDEVICE, "graph capture failed");
            const int rs = step_enqueue(c, P);
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(c->stream, &graph);
            if (rs != HF_OK || e != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return fail(c, HF_EDEVICE, "step capture failed");
            }
            g.graph = graph;
            if (hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0) != hipSuccess) {
                g.reset();
                return fail(c, HF_EDEVICE, "step graph instantiation failed");
            }
            g.key = key;
            how = 2;
        } else {
            how = 1;
        }
        HF_HIP(c, hipGraphLaunch(g.exec, c->stream));
    }
    unsigned long long hs[SS_REPAIRS + 1];
    HF_HIP(c, hipMemcpyAsync(hs, c->sscal.p, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipStreamSynchronize(c->stream));
    if (hs[SS_STATUS] != 0) {
        // a stage needed its first-order redo, a validation or the chemistry
        // failed: the state goes back to the step's start for the caller's
        // step with host decisions
        const unsigned gcopy = (unsigned)std::min<int64_t>((P.n + 255) / 256, 8192);
        hipLaunchKernelGGL(step_copy, dim3(gcopy), dim3(256), 0, c->stream, U, backup, P.n);
        HF_HIP(c, hipStreamSynchronize(c->stream));
        return HF_OK;
    }
    if (taken) *taken = how == 0 ? 3 : how;
    if (dt_taken) std::memcpy(dt_taken, &hs[SS_DT], sizeof(double));
    if (repairs) *repairs = (int64_t)hs[SS_REPAIRS];
    return HF_OK;
}

// ------------------------------------------------------------------ a hierarchy's Strang step as one graph
namespace {

// h/dx and h/dy of every flux box from the device h (hd: each box's level cell size)
__global__ void amr_set_h(hf_box* boxes, const double2* hd, int nbox, const unsigned long long* s) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    const double h = reinterpret_cast<const double*>(s)[SS_H];
    boxes[b].dtdx = h / hd[b].x;
    boxes[b].dtdy = h / hd[b].y;
}

enum { AK_U = 0, AK_U1 = 1, AK_U2 = 2, AK_OLD = 3, AK_BACKUP = 4 };

// dst = fma(a, y, b x) (x null: dst = y) on every component of every cell
// of a block table: the valid cells only, not the level arrays' empty space
// (the tables share one geometry and differ in their arrays)
__global__ void __launch_bounds__(256) step_cells_combine(const hf_cells* __restrict__ dst,
                                                          const hf_cells* __restrict__ y,
                                                          const hf_cells* __restrict__ x,
                                                          const int64_t* __restrict__ start, int nblock,
                                                          int64_t total, int nv, double a, double b) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= total) return;
    const CellMap S{nullptr, 0, dst, start, nblock};
    int bi, i, j;
    const hf_cells& B = S.find(c, bi, i, j);
    const int64_t at = B.offset + (int64_t)j * B.pitch + i;
    double* __restrict__ out = dst[bi].U + at;
    const double* __restrict__ py = y[bi].U + at;
    const double* __restrict__ px = x ? x[bi].U + at : nullptr;
    for (int n = 0; n < nv; ++n) {
        const int64_t o = (int64_t)n * B.plane;
        out[o] = px ? fma(a, py[o], b * px[o]) : py[o];
    }
}

struct AmrLevel {
    double* A[5];                   // U, U1, U2, old, backup
    double *fx, *fy;
    const uint8_t* mask;
    int nx, ny;
    double dx, dy;
    int wedge_i;
    const int32_t* ghosts;          // host strips
    int nghost;
    const int32_t* ghosts_dev;
    int64_t most;                   // the largest strip's parent count
    int64_t n;                      // doubles per array
    hf_patch patch() const { return hf_patch{2 * (int64_t)(nx + 4) + 2, nx + 4, (int64_t)(nx + 4) * (ny + 4)}; }
    hf_face_fluxes faces() const {
        return hf_face_fluxes{fx, 0, nx + 1, (int64_t)ny * (nx + 1), fy, 0, nx, (int64_t)(ny + 1) * nx};
    }
};

struct AmrPlan {
    std::vector<AmrLevel> lev;
    int32_t bc[4];
    double inlet[64];
    int has_inlet;
    double cfl;
    int chemistry, chem_mode;
    int nbox, ntiles, nblock;
    int64_t total;                  // cells of every box of every level (the block tables' cell count)
    int cfg[ChemTuner::NPARAM];     // chemistry refill group, waves per CU, bins, waits
    // device tables
    hf_box* boxes;
    const hf_face_fluxes* fluxes;
    const double2* hd;
    const double* const* in_ptrs[2];  // the two stages of a flow: U -> U1, U1 -> U2
    double* const* out_ptrs[2];
    const int32_t* tiles;
    const hf_cells* cells[5];       // the boxes' blocks in U, U1, U2, old, backup
    const int64_t* start;
};

// Enqueue one hierarchy Strang step on the context's stream: kernels only,
// so the sequence can be captured. Every failure folds into the step status.
int amr_enqueue(hf_ctx* c, const AmrPlan& P) {
    const int ns = c->ops->ns;
    const int nl = (int)P.lev.size();
    auto* s = c->sscal.as<unsigned long long>();
    unsigned long long* scal = c->scal.as<unsigned long long>();
    unsigned long long *ff = scal, *rate = scal + 8;
    BoxLaunch B0{};
    B0.mech = c->d_mech;
    B0.scal = scal + 8;
    B0.first_failed = ff;
    B0.stream = c->stream;
    auto fold = [&](const unsigned long long* f, const unsigned long long* donor, const unsigned long long* rep,
                    unsigned long long bit) {
        hipLaunchKernelGGL(step_fold, dim3(1), dim3(64), 0, c->stream, s, f, donor, rep, bit);
    };
    const unsigned gcells = (unsigned)((P.total + 255) / 256);
    auto combine = [&](int dst, int y, int x, double a, double b) {
        hipLaunchKernelGGL(step_cells_combine, dim3(gcells), dim3(256), 0, c->stream, P.cells[dst], P.cells[y],
                           x < 0 ? nullptr : P.cells[x], P.start, P.nblock, P.total, ns + 3, a, b);
    };
    auto with_cells = [&](BoxLaunch& V, int key) {
        V.blocks = P.cells[key];
        V.block_start = P.start;
        V.nblock = P.nblock;
        V.total = P.total;
    };
    // validation of every valid cell, its CFL rate over the leaves
    auto validate = [&](int key) {
        hipLaunchKernelGGL(step_props_reset, dim3(1), dim3(64), 0, c->stream, ff, rate);
        BoxLaunch V = B0;
        with_cells(V, key);
        V.op = BoxLaunch::CELLS_PROPS;
        return c->ops->box_ops(V);
    };
    auto average = [&](int key) {
        for (int l = nl - 1; l >= 1; --l) {
            const AmrLevel &F = P.lev[l], &C = P.lev[l - 1];
            BoxLaunch A = B0;
            A.op = BoxLaunch::AVGDOWN;
            A.src = F.A[key];
            A.sp = F.patch();
            A.dst = C.A[key];
            A.dp = C.patch();
            A.region[0] = A.region[1] = 0;
            A.region[2] = C.nx - 1;
            A.region[3] = C.ny - 1;
            A.mask = C.mask;
            A.mask_pitch = C.nx;
            if (c->ops->box_ops(A) != HF_OK) return HF_EDEVICE;
        }
        return HF_OK;
    };
    // one composite Euler stage src -> dst (BlockAMR::euler over every level)
    auto stage = [&](int src, int dst, int pass) {
        for (int l = 0; l < nl; ++l) {
            const AmrLevel& L = P.lev[l];
            if (l > 0 && L.nghost > 0) {
                const AmrLevel& C = P.lev[l - 1];
                BoxLaunch I = B0;
                I.op = BoxLaunch::INTERP;
                I.src = C.A[src];
                I.sp = C.patch();
                I.dst = L.A[src];
                I.dp = L.patch();
                for (int k = 0; k < 4; ++k) I.region[k] = L.ghosts[k];
                I.regions = L.ghosts_dev;
                I.nregion = L.nghost;
                I.total = L.most;
                I.mask = L.mask;
                I.mask_pitch = L.nx;
                I.mask_nx = L.nx;
                I.mask_ny = L.ny;
                if (c->ops->box_ops(I) != HF_OK) return HF_EDEVICE;
                fold(ff, nullptr, nullptr, 1ull);
            }
            BoxLaunch Fl = B0;
            Fl.box = hf_box{L.nx, L.ny, 0, L.wedge_i, 0, L.nx + 4, (int64_t)(L.nx + 4) * (L.ny + 4), 0, L.nx,
                            (int64_t)L.nx * L.ny, 0.0, 0.0};
            set_fill(Fl, L.A[src], P.bc, P.has_inlet ? P.inlet : nullptr);
            if (c->ops->box_ops(Fl) != HF_OK) return HF_EDEVICE;
        }
        FluxLaunch F{};
        F.mech = c->d_mech;
        F.boxes = P.boxes;
        F.nbox = P.nbox;
        F.tiles = P.tiles;
        F.ntiles = P.ntiles;
        F.first_order = 0;
        F.one_d = 0;
        F.skip_validation = 1;              // the repair and the validation below judge the update
        F.first_failed = scal + 2;
        F.counters = scal + 3;
        F.fluxes = P.fluxes;
        F.in_ptrs = P.in_ptrs[pass];
        F.out_ptrs = P.out_ptrs[pass];
        F.stream = c->stream;
        if (c->ops->flux(F) != HF_OK) return HF_EDEVICE;
        for (int l = nl - 1; l >= 1; --l) {
            const AmrLevel &Fi = P.lev[l], &C = P.lev[l - 1];
            BoxLaunch R = B0;
            R.op = BoxLaunch::REFLUX_CELLS;
            R.dst = C.A[dst];
            R.dp = C.patch();
            R.cflux = C.faces();
            R.fflux = Fi.faces();
            R.region[0] = R.region[1] = 0;
            R.region[2] = C.nx - 1;
            R.region[3] = C.ny - 1;
            R.mask = C.mask;
            R.mask_pitch = C.nx;
            R.hp = reinterpret_cast<const double*>(s + SS_H);
            R.dx = C.dx;
            R.dy = C.dy;
            if (c->ops->box_ops(R) != HF_OK) return HF_EDEVICE;
        }
        if (average(dst) != HF_OK) return HF_EDEVICE;
        BoxLaunch Rp = B0;
        with_cells(Rp, dst);
        Rp.op = BoxLaunch::CELLS_REPAIR;
        Rp.partials = c->aux[5].as<double>();
        if (c->ops->box_ops(Rp) != HF_OK) return HF_EDEVICE;
        fold(ff, scal + 8 + 3 * ns + 1, scal + 8 + 3 * ns, 1ull);
        // the average-down again: a repaired leaf changes its coarse parent
        // (with nothing repaired it rewrites the same values)
        if (average(dst) != HF_OK || validate(dst) != HF_OK) return HF_EDEVICE;
        fold(ff, nullptr, nullptr, 1ull);
        return HF_OK;
    };
    // dt from the CFL rate of the leaves
    if (validate(AK_U) != HF_OK) return HF_EDEVICE;
    hipLaunchKernelGGL(step_dt, dim3(1), dim3(64), 0, c->stream, s, ff, rate, P.cfl);
    hipLaunchKernelGGL(amr_set_h, dim3((unsigned)((P.nbox + 255) / 256)), dim3(256), 0, c->stream, P.boxes, P.hd,
                       P.nbox, s);
    // the step's start and the Heun average touch the valid cells only: the
    // ghost cells are refilled before every read
    combine(AK_BACKUP, AK_U, -1, 1.0, 0.0);
    for (int half = 0; half < 2; ++half) {
        // flow(dt/2): U <- U/2 + E(E(U))/2 on every level, average-down, validation
        combine(AK_OLD, AK_U, -1, 1.0, 0.0);
        if (stage(AK_U, AK_U1, 0) != HF_OK || stage(AK_U1, AK_U2, 1) != HF_OK) return HF_EDEVICE;
        combine(AK_U, AK_U2, AK_OLD, 0.5, 0.5);
        if (average(AK_U) != HF_OK || validate(AK_U) != HF_OK) return HF_EDEVICE;
        fold(ff, nullptr, nullptr, 1ull);
        if (half == 1 || !P.chemistry) continue;
        // react(dt): one batch over every leaf cell, in place (BlockAMR::react)
        ChemLaunch C{};
        C.mech = c->d_mech;
        C.ncell = P.total;
        C.ld = P.total;
        C.blocks = P.cells[AK_U];
        C.block_start = P.start;
        C.nblock = P.nblock;
        C.dt = reinterpret_cast<const double*>(s + SS_DT);
        C.dt_per_cell = 0;
        C.out = c->out.as<double>();
        C.ldo = P.total;
        C.status = c->status.as<int8_t>();
        C.first_failed = s + SS_CHEM;
        C.counters = s + SS_CHEM + 1;
        C.rtol = c->opt.rtol;
        C.atol = c->opt.atol;
        C.maxsteps = c->opt.max_steps;
        C.mode = P.chem_mode;
        C.precision = c->opt.precision;
        C.num_cus = c->num_cus;
        C.chunk = P.cfg[ChemTuner::CHUNK];
        C.waves_per_cu = P.cfg[ChemTuner::WAVES];
        C.nbins = P.cfg[ChemTuner::BINS];
        C.jac_wait = P.cfg[ChemTuner::JWAIT];
        C.setup_wait = P.cfg[ChemTuner::SWAIT];
        C.workspace = c->work.p;
        C.workspace_bytes = c->work.n;
        C.stream = c->stream;
        C.commit = P.lev[0].A[AK_U];        // non-null: commit through the block table when no cell failed
        C.commit_ld = P.total;
        if (c->ops->chem(C) != HF_OK) return HF_EDEVICE;
        fold(s + SS_CHEM, nullptr, nullptr, 2ull);
        if (average(AK_U) != HF_OK || validate(AK_U) != HF_OK) return HF_EDEVICE;
        fold(ff, nullptr, nullptr, 1ull);
    }
    return hipGetLastError() == hipSuccess ? HF_OK : HF_EDEVICE;
}

// byte offset of the next table section, 256-byte aligned
size_t amr_align(size_t n) { return (n + 255) & ~size_t(255); }

}  // namespace

int hf_amr_strang_step(hf_ctx* c, const hf_level* levels, int nlevel, const int32_t bc[4], const double* inlet,
                       double cfl, int chemistry, int chem_mode, double time_left, double* dt_taken, int32_t* taken,
                       int64_t* repairs) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    if (taken) *taken = 0;
    if (dt_taken) *dt_taken = 0;
    if (repairs) *repairs = 0;
    if (!levels || nlevel < 1 || nlevel > 16) return fail(c, HF_EINVAL, "invalid level list");
    if (!(cfl > 0) || !std::isfinite(cfl) || (chem_mode != HF_CHEM_PLAIN && chem_mode != HF_CHEM_BATCHED))
        return fail(c, HF_EINVAL, "invalid CFL number or chemistry form");
    const int rb = check_bc(c, nullptr, bc, inlet);
    if (rb != HF_OK) return rb;
    if (bc[2] < 0 || bc[3] < 0) return fail(c, HF_EINVAL, "a hierarchy step is 2-D: y boundary codes are needed");
    for (int k = 0; k < 4; ++k)
        if (bc[k] == HF_BC_KEEP)
            return fail(c, HF_EINVAL, "a hierarchy step refills every ghost cell: HF_BC_KEEP is not a domain side");
    HF_DEVICE(c);
    const int ns = c->mech.ns(), nv = ns + 3;
    if (nv > 64) return fail(c, HF_EINVAL, "mechanism too large for the step graph's inlet state");
    AmrPlan P{};
    for (int k = 0; k < 4; ++k) P.bc[k] = bc[k];
    if (inlet) {
        P.has_inlet = 1;
        for (int n = 0; n < nv; ++n) P.inlet[n] = inlet[n];
    }
    P.cfl = cfl;
    P.chemistry = chemistry ? 1 : 0;
    P.chem_mode = chem_mode;
    // levels: extents, ratio 2, device arrays of the right size, boxes and strips inside
    std::vector<hf_box> boxes;
    std::vector<hf_face_fluxes> fluxes;
    std::vector<double2> hd;
    std::vector<const double*> ptrs[4];
    std::vector<hf_cells> cells[5];
    std::vector<int64_t> start(1, 0);
    std::vector<int32_t> ghosts;
    std::vector<size_t> ghost_at;
    for (int l = 0; l < nlevel; ++l) {
        const hf_level& H = levels[l];
        const std::string name = "level " + std::to_string(l);
        if (H.nx < 1 || H.ny < 1 || !(H.dx > 0) || !(H.dy > 0) || !std::isfinite(H.dx) || !std::isfinite(H.dy))
            return fail(c, HF_EINVAL, name + " has an invalid extent or cell size");
        if (l > 0 && (H.nx != 2 * levels[l - 1].nx || H.ny != 2 * levels[l - 1].ny))
            return fail(c, HF_EINVAL, name + " is not a refinement by 2 of the level below");
        if (H.nbox < 1 || !H.boxes || H.nghost < 0 || (H.nghost > 0 && !H.ghosts) || (l == 0 && H.nghost > 0))
            return fail(c, HF_EINVAL, name + " has an invalid box or ghost-strip list");
        AmrLevel L{};
        double* arr[5] = {H.U, H.U1, H.U2, H.old, H.backup};
        L.n = (int64_t)nv * (H.nx + 4) * (H.ny + 4);
        for (int k = 0; k < 5; ++k) {
            if (!arr[k] || !is_device_ptr(arr[k]) || !within_allocation(arr[k], 0, L.n - 1))
                return fail(c, HF_EINVAL, name + ": every level array must be device memory of nv (nx+4)(ny+4) values");
            L.A[k] = arr[k];
        }
        if (!H.fx || !H.fy || !is_device_ptr(H.fx) || !is_device_ptr(H.fy) || !H.mask || !is_device_ptr(H.mask) ||
            !within_allocation(H.fx, 0, (int64_t)nv * H.ny * (H.nx + 1) - 1) ||
            !within_allocation(H.fy, 0, (int64_t)nv * (H.ny + 1) * H.nx - 1))
            return fail(c, HF_EINVAL, name + ": face fluxes and mask must be device memory of the level's size");
        L.fx = H.fx;
        L.fy = H.fy;
        L.mask = H.mask;
        L.nx = H.nx;
        L.ny = H.ny;
        L.dx = H.dx;
        L.dy = H.dy;
        L.wedge_i = H.wedge_i;
        L.ghosts = H.ghosts;
        L.nghost = H.nghost;
        if (l == 0 && (H.nbox != 1 || H.boxes[0] != 0 || H.boxes[1] != 0 || H.boxes[2] != H.nx - 1 ||
                       H.boxes[3] != H.ny - 1))
            return fail(c, HF_EINVAL, "level 0 must be one box over the whole level");
        const int64_t gx = H.nx + 4, plane = gx * (H.ny + 4);
        for (int b = 0; b < H.nbox; ++b) {
            const int32_t* g = H.boxes + 4 * b;
            if (g[0] < 0 || g[1] < 0 || g[2] >= H.nx || g[3] >= H.ny || g[2] < g[0] || g[3] < g[1])
                return fail(c, HF_EINVAL, name + " box " + std::to_string(b) + " lies outside the level");
            int walls = 0;
            const bool touch[4] = {g[0] == 0, g[2] == H.nx - 1, g[1] == 0, g[3] == H.ny - 1};
            for (int k = 0; k < 4; ++k)
                if (touch[k] && (bc[k] == HF_BC_REFLECTING || (k == 2 && bc[k] == HF_BC_WEDGE))) walls |= 1 << k;
            const int32_t bnx = g[2] - g[0] + 1, bny = g[3] - g[1] + 1;
            boxes.push_back(hf_box{bnx, bny, walls, H.wedge_i - g[0], g[1] * gx + g[0], gx, plane,
                                   (g[1] + 2) * gx + g[0] + 2, gx, plane, 0.0, 0.0});
            // box-local face (a, b) is the level's face (i0 + a, j0 + b)
            fluxes.push_back(hf_face_fluxes{H.fx, (int64_t)g[1] * (H.nx + 1) + g[0], H.nx + 1,
                                            (int64_t)H.ny * (H.nx + 1), H.fy, (int64_t)g[1] * H.nx + g[0], H.nx,
                                            (int64_t)(H.ny + 1) * H.nx});
            hd.push_back(make_double2(H.dx, H.dy));
            ptrs[0].push_back(H.U);
            ptrs[1].push_back(H.U1);
            ptrs[2].push_back(H.U1);
            ptrs[3].push_back(H.U2);
            double* keyed[5] = {H.U, H.U1, H.U2, H.old, H.backup};
            for (int k = 0; k < 5; ++k)
                cells[k].push_back(hf_cells{keyed[k], (g[1] + 2) * gx + g[0] + 2, gx, plane, bnx, bny,
                                            H.mask + (int64_t)g[1] * H.nx + g[0], H.nx, H.dx, H.dy});
            start.push_back(start.back() + (int64_t)bnx * bny);
        }
        ghost_at.push_back(ghosts.size());
        if (l > 0) {
            const AmrLevel& C = P.lev[l - 1];
            const hf_patch cp = C.patch(), fp = L.patch();
            for (int k = 0; k < H.nghost; ++k) {
                const int32_t* g = H.ghosts + 4 * k;
                if (g[2] < g[0] || g[3] < g[1]) return fail(c, HF_EINVAL, name + " has an empty ghost strip");
                if (check_extents(c, BoxLaunch::INTERP, C.A[0], cp, L.A[0], fp, g) != HF_OK) return HF_EINVAL;
                L.most = std::max<int64_t>(L.most, (int64_t)((g[2] >> 1) - (g[0] >> 1) + 1) *
                                                       ((g[3] >> 1) - (g[1] >> 1) + 1));
                for (int q = 0; q < 4; ++q) ghosts.push_back(g[q]);
            }
        }
        P.lev.push_back(L);
    }
    P.nbox = (int)boxes.size();
    P.nblock = P.nbox;
    P.total = start.back();
    const int TX = c->ops->flux_tx2, TY = c->ops->flux_ty2;
    std::vector<int32_t> tiles;
    for (int b = 0; b < P.nbox; ++b)
        for (int tj = 0; tj < (boxes[b].ny + TY - 1) / TY; ++tj)
            for (int ti = 0; ti < (boxes[b].nx + TX - 1) / TX; ++ti) {
                tiles.push_back(b);
                tiles.push_back(ti);
                tiles.push_back(tj);
            }
    P.ntiles = (int)(tiles.size() / 3);
    // the chemistry's tuning values: the options', else the tuner's result
    // for this size class once it has one, else the defaults (no result
    // depends on them)
    const bool wave = c->ops->ns > 16;
    const bool tune = !c->opt.chunk && !c->opt.waves_per_cu && !c->opt.nbins && !c->opt.jac_wait && !c->opt.setup_wait;
    if (tune) {
        const ChemTuner& T = c->tuner;
        const bool same = T.done && T.wave == wave && P.total <= 2 * T.size && 2 * P.total >= T.size;
        for (int p = 0; p < ChemTuner::NPARAM; ++p)
            P.cfg[p] = same ? T.best[p] : (wave ? ChemTuner::kWave : ChemTuner::kLane)[p][0];
    } else {
        P.cfg[ChemTuner::CHUNK] = c->opt.chunk;
        P.cfg[ChemTuner::WAVES] = c->opt.waves_per_cu;
        P.cfg[ChemTuner::BINS] = c->opt.nbins;
        P.cfg[ChemTuner::JWAIT] = c->opt.jac_wait;
        P.cfg[ChemTuner::SWAIT] = c->opt.setup_wait;
    }
    // device tables, one buffer
    const size_t nb = (size_t)P.nbox;
    size_t at = 0;
    const size_t o_box = at;
    at = amr_align(at + sizeof(hf_box) * nb);
    const size_t o_flux = at;
    at = amr_align(at + sizeof(hf_face_fluxes) * nb);
    const size_t o_hd = at;
    at = amr_align(at + sizeof(double2) * nb);
    const size_t o_ptr = at;
    at = amr_align(at + 4 * sizeof(void*) * nb);
    const size_t o_tile = at;
    at = amr_align(at + sizeof(int32_t) * tiles.size());
    const size_t o_ghost = at;
    at = amr_align(at + sizeof(int32_t) * std::max<size_t>(ghosts.size(), 1));
    const size_t o_cells = at;
    at = amr_align(at + 5 * sizeof(hf_cells) * nb);
    const size_t o_start = at;
    at = amr_align(at + sizeof(int64_t) * start.size());
    const size_t groups = (size_t)((P.total + 255) / 256);
    const size_t wbytes =
        P.chemistry ? c->ops->chem_workspace(P.total, P.cfg[ChemTuner::BINS], c->num_cus, P.cfg[ChemTuner::WAVES]) : 0;
    if (!c->atab.ensure(at) || !c->sscal.ensure(8 * SS_WORDS) || !c->scal.ensure(sizeof(unsigned long long) * 256) ||
        !c->aux[5].ensure(sizeof(double) * ns * groups) || (P.chemistry && (!c->out.ensure(sizeof(double) * ns * P.total) ||
                                                                             !c->status.ensure(P.total))) ||
        (wbytes && !c->work.ensure(wbytes)))
        return fail(c, HF_ENOMEM, "device scratch");
    char* T = c->atab.as<char>();
    P.boxes = reinterpret_cast<hf_box*>(T + o_box);
    P.fluxes = reinterpret_cast<const hf_face_fluxes*>(T + o_flux);
    P.hd = reinterpret_cast<const double2*>(T + o_hd);
    auto* pt = reinterpret_cast<double**>(T + o_ptr);
    P.in_ptrs[0] = pt;
    P.out_ptrs[0] = pt + nb;
    P.in_ptrs[1] = pt + 2 * nb;
    P.out_ptrs[1] = pt + 3 * nb;
    P.tiles = reinterpret_cast<const int32_t*>(T + o_tile);
    for (int k = 0; k < 5; ++k) P.cells[k] = reinterpret_cast<const hf_cells*>(T + o_cells) + k * nb;
    P.start = reinterpret_cast<const int64_t*>(T + o_start);
    for (int l = 1; l < nlevel; ++l)
        P.lev[l].ghosts_dev = reinterpret_cast<const int32_t*>(T + o_ghost) + ghost_at[l];
    // the key: every table, level, condition, option and context buffer the
    // graph's kernels address (other calls may have grown the buffers)
    std::vector<uint64_t> key;
    auto bits = [](double x) {
        uint64_t u;
        std::memcpy(&u, &x, sizeof(u));
        return u;
    };
    auto push_bytes = [&](const void* p, size_t n) {
        const size_t w = key.size();
        key.resize(w + (n + 7) / 8, 0);
        std::memcpy(key.data() + w, p, n);
    };
    push_bytes(boxes.data(), sizeof(hf_box) * nb);
    push_bytes(fluxes.data(), sizeof(hf_face_fluxes) * nb);
    push_bytes(ghosts.data(), sizeof(int32_t) * ghosts.size());
    for (const hf_cells& b : cells[0]) key.push_back((uint64_t)(uintptr_t)b.mask);
    for (const AmrLevel& L : P.lev) {
        for (double* a : L.A) key.push_back((uint64_t)(uintptr_t)a);
        for (uint64_t v : {(uint64_t)L.nx, (uint64_t)L.ny, bits(L.dx), bits(L.dy), (uint64_t)(int64_t)L.nghost})
            key.push_back(v);
    }
    for (int k = 0; k < 4; ++k) key.push_back((uint64_t)(int64_t)bc[k]);
    key.push_back((uint64_t)P.has_inlet);
    for (int n = 0; n < nv; ++n) key.push_back(bits(P.inlet[n]));
    for (uint64_t v : {bits(cfl), (uint64_t)P.chemistry, (uint64_t)chem_mode, (uint64_t)nlevel, (uint64_t)TX,
                       (uint64_t)TY, bits(c->opt.rtol), bits(c->opt.atol), (uint64_t)c->opt.max_steps,
                       (uint64_t)c->opt.precision, (uint64_t)c->opt.eager})
        key.push_back(v);
    for (int p = 0; p < ChemTuner::NPARAM; ++p) key.push_back((uint64_t)(int64_t)P.cfg[p]);
    for (const DevBuf* b : {&c->atab, &c->sscal, &c->scal, &c->aux[5], &c->out, &c->status, &c->work}) {
        key.push_back((uint64_t)(uintptr_t)b->p);
        key.push_back((uint64_t)b->n);
    }
    ChemGraph& g = c->amr_graph;
    const bool fresh = !g.exec || g.key != key;
    if (fresh) {
        g.reset();
        std::vector<const double*> pflat;
        for (int k = 0; k < 4; ++k) pflat.insert(pflat.end(), ptrs[k].begin(), ptrs[k].end());
        std::vector<hf_cells> cflat;
        for (int k = 0; k < 5; ++k) cflat.insert(cflat.end(), cells[k].begin(), cells[k].end());
        HF_HIP(c, hipMemcpyAsync(T + o_box, boxes.data(), sizeof(hf_box) * nb, hipMemcpyHostToDevice, c->stream));
        HF_HIP(c, hipMemcpyAsync(T + o_flux, fluxes.data(), sizeof(hf_face_fluxes) * nb, hipMemcpyHostToDevice,
                                 c->stream));
        HF_HIP(c, hipMemcpyAsync(T + o_hd, hd.data(), sizeof(double2) * nb, hipMemcpyHostToDevice, c->stream));
        HF_HIP(c, hipMemcpyAsync(T + o_ptr, pflat.data(), sizeof(void*) * pflat.size(), hipMemcpyHostToDevice,
                                 c->stream));
        HF_HIP(c, hipMemcpyAsync(T + o_tile, tiles.data(), sizeof(int32_t) * tiles.size(), hipMemcpyHostToDevice,
                                 c->stream));
        if (!ghosts.empty())
            HF_HIP(c, hipMemcpyAsync(T + o_ghost, ghosts.data(), sizeof(int32_t) * ghosts.size(),
                                     hipMemcpyHostToDevice, c->stream));
        HF_HIP(c, hipMemcpyAsync(T + o_cells, cflat.data(), sizeof(hf_cells) * cflat.size(), hipMemcpyHostToDevice,
                                 c->stream));
        HF_HIP(c, hipMemcpyAsync(T + o_start, start.data(), sizeof(int64_t) * start.size(), hipMemcpyHostToDevice,
                                 c->stream));
        // the host tables above are locals: the copies finish before they go
        HF_HIP(c, hipStreamSynchronize(c->stream));
    }
    double left = time_left;
    HF_HIP(c, hipMemcpyAsync(c->sscal.as<char>() + 8 * SS_LEFT, &left, 8, hipMemcpyHostToDevice, c->stream));
    int how = 0;
    if (c->opt.eager) {
        if (amr_enqueue(c, P) != HF_OK) return fail(c, HF_EDEVICE, "step launch failed");
    } else {
        if (fresh) {
            if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return fail(c, HF_EDEVICE, "graph capture failed");
            const int rs = amr_enqueue(c, P);
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(c->stream, &graph);
            if (rs != HF_OK || e != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return fail(c, HF_EDEVICE, "step capture failed");
            }
            g.graph = graph;
            if (hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0) != hipSuccess) {
                g.reset();
                return fail(c, HF_EDEVICE, "step graph instantiation failed");
            }
            g.key = key;
            how = 2;
        } else {
            how = 1;
        }
        HF_HIP(c, hipGraphLaunch(g.exec, c->stream));
    }
    unsigned long long hs[SS_REPAIRS + 1];
    HF_HIP(c, hipMemcpyAsync(hs, c->sscal.p, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipStreamSynchronize(c->stream));
    if (hs[SS_STATUS] != 0) {
        // a stage, a validation or the chemistry failed: every level goes
        // back to the step's start for the caller's step with host decisions
        hipLaunchKernelGGL(step_cells_combine, dim3((unsigned)((P.total + 255) / 256)), dim3(256), 0, c->stream,
                           P.cells[AK_U], P.cells[AK_BACKUP], nullptr, P.start, P.nblock, P.total, nv, 1.0, 0.0);
        HF_HIP(c, hipStreamSynchronize(c->stream));
        return HF_OK;
    }
    if (taken) *taken = how == 0 ? 3 : how;
    if (dt_taken) std::memcpy(dt_taken, &hs[SS_DT], sizeof(double));
    if (repairs) *repairs = (int64_t)hs[SS_REPAIRS];
    return HF_OK;
}

int hf_repair_roundoff(hf_ctx* c, double* U, const hf_box* box, int64_t* repairs, int64_t* first_failed) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    if (first_failed) *first_failed = -1;
    BoxLaunch L{};
    const int r = box_common(c, L, box, U);
    if (r != HF_OK) return r;
    L.U = U;
    L.op = BoxLaunch::REPAIR;
    if (c->ops->box_ops(L) != HF_OK) return fail(c, HF_EDEVICE, "roundoff repair launch failed");
    const int ns = c->ops->ns;
    std::vector<unsigned long long> h(3 * ns + 2);
    unsigned long long ff = ~0ull;
    HF_HIP(c, hipMemcpyAsync(h.data(), L.scal, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipMemcpyAsync(&ff, L.first_failed, sizeof(ff), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipStreamSynchronize(c->stream));
    if (repairs) *repairs = (int64_t)h[3 * ns];
    if (ff != ~0ull || h[3 * ns + 1]) {
        if (first_failed) *first_failed = ff == ~0ull ? 0 : (int64_t)ff;
        return fail(c, HF_EDOMAIN, h[3 * ns + 1] ? "insufficient positive donor for roundoff redistribution"
                                                 : "materially negative species after flow stage at cell " +
                                                       std::to_string(ff));
    }
    return HF_OK;
}

int hf_validate_box(hf_ctx* c, const double* U, const hf_box* box, double dx, double dy, double* rate,
                    int64_t* first_failed) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    if (first_failed) *first_failed = -1;
    if (!(dx > 0) || !(dy >= 0)) return fail(c, HF_EINVAL, "invalid cell size");
    BoxLaunch L{};
    const int r = box_common(c, L, box, U);
    if (r != HF_OK) return r;
    L.U = const_cast<double*>(U);
    L.op = BoxLaunch::PROPS;
    L.dx = dx;
    L.dy = dy;
    unsigned long long init[2] = {~0ull, 0};
    HF_HIP(c, hipMemcpyAsync(L.first_failed, &init[0], sizeof(init[0]), hipMemcpyHostToDevice, c->stream));
    HF_HIP(c, hipMemcpyAsync(L.scal, &init[1], sizeof(init[1]), hipMemcpyHostToDevice, c->stream));
    if (c->ops->box_ops(L) != HF_OK) return fail(c, HF_EDEVICE, "validation launch failed");
    unsigned long long out[2];
    HF_HIP(c, hipMemcpyAsync(&out[0], L.first_failed, sizeof(out[0]), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipMemcpyAsync(&out[1], L.scal, sizeof(out[1]), hipMemcpyDeviceToHost, c->stream));
    HF_HIP(c, hipStreamSynchronize(c->stream));
    if (out[0] != ~0ull) {
        if (first_failed) *first_failed = (int64_t)out[0];
        return fail(c, HF_EDOMAIN, "inadmissible state at cell " + std::to_string(out[0]));
    }
    if (rate) {
        double v;
        std::memcpy(&v, &out[1], sizeof(v));
        *rate = v;
    }
    return HF_OK;
}

int hf_lincomb(hf_ctx* c, double* y, const double* x, double a, double b, int64_t n) {
    if (!c) return HF_EINVAL;
    BusyGuard guard(c);
    if (!guard.ok) return fail(c, HF_EBUSY, "concurrent call is unsupported");
    if (n < 0 || (n > 0 && (!y || !x))) return fail(c, HF_EINVAL, "invalid extent");
    if (n == 0) return HF_OK;
    HF_DEVICE(c);
    if (!is_device_ptr(y) || !is_device_ptr(x)) return fail(c, HF_EINVAL, "lincomb needs device memory");
    const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(lincomb_kernel, dim3(grid), dim3(256), 0, c->stream, y, x, a, b, n);
    if (hipGetLastError() != hipSuccess) return fail(c, HF_EDEVICE, "lincomb launch failed");
    HF_HIP(c, hipStreamSynchronize(c->stream));
    return HF_OK;
}

int hf_synth_cells(hf_ctx* c, int which, uint64_t seed, int64_t ncell, double* U, double* dt) {
    if (!c) return HF_EINVAL;
    if (ncell < 0 || (ncell > 0 && !U)) return fail(c, HF_EINVAL, "invalid extent");
    try {
        synth_cells(c->mech, which, seed, ncell, U, dt);
    } catch (const std::exception& e) {
        return fail(c, HF_EINVAL, e.what());
    }
    return HF_OK;
}

int hf_equilibrate_tp(hf_ctx* c, double T, double p, const double* X, double* Y) {
    if (!c) return HF_EINVAL;
    if (!X || !Y || !(T > 0) || !(p > 0)) return fail(c, HF_EINVAL, "invalid equilibrium arguments");
    try {
        equilibrate_tp(c->mech, T, p, X, Y);
    } catch (const std::exception& e) {
        return fail(c, HF_ECONV, e.what());
    }
    return HF_OK;
}

int hf_state_from_TPY(hf_ctx* c, double T, double p, const double* Y, double u, double v, double* q) {
    if (!c) return HF_EINVAL;
    if (!Y || !q || !(T > 0) || !(p > 0)) return fail(c, HF_EINVAL, "invalid state arguments");
    state_from_TPY(c->mech, T, p, Y, u, v, q);
    return HF_OK;
}

}  // extern "C"
