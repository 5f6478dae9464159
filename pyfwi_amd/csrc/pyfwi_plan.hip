// Plan objects and the extern "C" entry points declared in include/pyfwi_hip.h.
//
// A plan owns, on one GPU:
//   model   lam, mu, buoyancy in the padded layout; damping profiles (1-D)
//   fields  [nshots][5][fstride]  forward state of every shot of the batch
//   adj     [nshots][5][fstride]  adjoint state
//   av      2 x [nshots][2][fstride]  exact mode velocity adjoint ping-pong
//   gacc    [nsets][6][fstride]  gradient accumulators: one set per block's group
//           of shots
//   snaps   [nshots][nchp][5][fstride]  reference checkpoints (pyfwi mode)
//   ckpts   [nshots][ncp][5][fstride]   exact-mode checkpoints every K steps
//   seg     [nshots][K][5][fstride]     exact-mode recomputed segment
// Batched shots are the z dimension of every stencil launch, so one launch
// updates all shots of a step (the reference runs shots one after another,
// src/PyFWI/wave_propagation.py:639-645, 740-749).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <initializer_list>
#include <mutex>
#include <vector>

#include "../../include/pyfwi_hip.h"
#include "pyfwi_kernels.h"

using namespace pyfwi;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                        \
            return fail(PYFWI_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------ allocation-free forward
// pyfwi_forward_noalloc keeps the forward states and the shot table of its
// call in this array instead of allocating.  Each GPU the library runs on
// holds its own copy until the process ends; host code uses the per-device
// address from hipGetSymbolAddress, never the array's name.
static __device__ __align__(256) float scratch[PYFWI_SCRATCH_FLOATS];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[MAX_DEVICES] = {};
bool g_scratch_done[MAX_DEVICES] = {};

// Address of `scratch` on device `dev`, looked up once per device with that
// device current (plan creation); null when the lookup failed.
float *scratch_on(int dev)
{
    if (dev < 0 || dev >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_done[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            ptr = nullptr;
        }
        g_scratch_ptr[dev] = static_cast<float *>(ptr);
        g_scratch_done[dev] = true;
    }
    return g_scratch_ptr[dev];
}

struct DevBuf {
    float *p = nullptr;
    size_t n = 0;  // floats
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// Grow-only device buffer; zeroed on (re)allocation.
int ensure(DevBuf &b, size_t n, hipStream_t st, const char *what)
{
    if (b.n >= n && b.p) return 0;
    b.release();
    if (n == 0) return 0;
    hipError_t e = hipMalloc(&b.p, n * sizeof(float));
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(PYFWI_ENOMEM, std::string("hipMalloc failed for ") + what + " (" +
                                      std::to_string(n * sizeof(float)) + " bytes)");
    }
    b.n = n;
    launch_zero(b.p, (int64_t)n, st);  // halos and padding must read as zero
    return 0;
}

} // namespace

struct pyfwi_plan {
    pyfwi_grid grid{};
    DevGrid dg{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    bool model_set = false;
    DevBuf lam, mu, bu, dxp_buf, dzp_buf, scratch_in;
    // damping profiles with HALO zero entries on both sides: dxp[-HALO..nx+HALO)
    float *dxp = nullptr, *dzp = nullptr;
    DevBuf fields, fields2, adj, adj2, av0, av1, gacc, snaps, ckpts, seg;
    float *cur = nullptr;  // buffer (fields or fields2) holding the current forward state
    // tuning: shots per pass (0: all), fused steps, strip height (0: automatic)
    int tune_group = 0, tune_fused = 0, tune_rows = 0;
    int adj_fused = 0;  // reference adjoint in one launch a step (PYFWI_TUNE_ADJ_FUSED)
    int fwd_two = 0;    // kept-state forward two steps a launch (PYFWI_TUNE_FWD_TWO)
    bool ckpts_all = false;  // ckpts holds every shot of the last forward
    DevBuf src_stage, seis_stage, res_stage, grad_stage, shot_tab_f, resid;
    // record filter of the misfit (pyfwi_set_record_filter): taps k[0..L]
    DevBuf fir_taps, filt_a, filt_b;
    int fir_L = -1, fir_pad = 0;
    double *d_misfit = nullptr;
    int *d_check = nullptr;  // two ints for the model check
    // hipGraph replay of the time loops (pyfwi_plan_set_graphs)
    int graphs = 0;
    hipStream_t cap_stream = nullptr;  // capture only: nothing runs on it
    std::vector<std::pair<uint64_t, hipGraphExec_t>> graph_cache;
    int64_t graph_captures = 0, graph_replays = 0;
    std::vector<int32_t> last_passes;  // shots per pass of the last call
    std::vector<int32_t> last_order;   // the caller's shot indices in the order they ran
    int last_ckpt_every = 0;           // exact mode: checkpoint interval of the last call
    pyfwi_shot *d_shots = nullptr;
    int shots_cap = 0;
    // receiver table (pyfwi_set_receivers): host copy for the shot checks
    std::vector<int32_t> h_rtab;
    int32_t *d_rtab = nullptr;
    int64_t rtab_cap = 0;
    // state of the last forward
    std::vector<pyfwi_shot> last_shots;
    std::vector<pyfwi_shot> h_shots;  // host copy of the table on the device
    std::vector<int32_t> chp;
    int ckpt_every = 0;
    int ncp = 0;
    bool have_forward = false;
    // every source of the uploaded shots lies on an active cell (rows and
    // columns sdo..N-sdo): the forward stencil launches then apply the source
    // term to the values they load.  Otherwise it is added in memory, where on
    // a never-updated cell it accumulates from step to step, as the
    // reference's in-place additions do.
    bool src_inside = true;
    // every shot records a receiver line on active cells: the stress update
    // writes the record samples and the adjoint kernels add the residual.
    // Otherwise a gather launch records and the residual is added in memory.
    bool rec_inline = true;
    // timing
    int timing = 0;  // 0 off, 1 every stencil launch, 2 the time loops
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_kind;
    std::vector<int> ev_count;  // steps a bracket covers (1 for a launch)
    size_t ev_used = 0;
    double kt_ms[PYFWI_K_COUNT] = {0};
    int64_t kt_n[PYFWI_K_COUNT] = {0};
};

namespace {

int64_t fields_per_shot(const pyfwi_plan *p) { return 5 * p->dg.fstride; }

int max_nt(const std::vector<pyfwi_shot> &s)
{
    int m = 0;
    for (auto &x : s) m = std::max(m, x.nt);
    return m;
}

// Bracket a stencil launch (timing mode 1) or a time loop of `count` steps
// (mode 2) with an event pair.
struct Timed {
    pyfwi_plan *p;
    int kind;
    size_t slot = 0;
    bool on;
    Timed(pyfwi_plan *p_, int k, int mode = 1, int count = 1)
        : p(p_), kind(k), on(p_->timing == mode)
    {
        if (!on) return;
        if (p->ev_used + 2 > p->ev_pool.size()) {
            for (int n = 0; n < 256; ++n) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) { on = false; return; }
                p->ev_pool.push_back(e);
                p->ev_kind.push_back(-1);
                p->ev_count.push_back(1);
            }
        }
        slot = p->ev_used;
        p->ev_used += 2;
        p->ev_kind[slot] = kind;
        p->ev_count[slot] = count;
        (void)hipEventRecord(p->ev_pool[slot], p->stream);
    }
    ~Timed()
    {
        if (on) (void)hipEventRecord(p->ev_pool[slot + 1], p->stream);
    }
};

// Fold the recorded event pairs into the per-kind totals.
int harvest_timing(pyfwi_plan *p)
{
    if (!p->timing || p->ev_used == 0) return 0;
    HIPCHK(hipStreamSynchronize(p->stream));
    for (size_t k = 0; k + 1 < p->ev_used; k += 2) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, p->ev_pool[k], p->ev_pool[k + 1]));
        int kind = p->ev_kind[k];
        if (kind >= 0 && kind < PYFWI_K_COUNT) {
            p->kt_ms[kind] += ms;
            p->kt_n[kind] += p->ev_count[k];
        }
    }
    p->ev_used = 0;
    return 0;
}

// Flush the event pool when it grows large (long time loops).
int maybe_harvest(pyfwi_plan *p)
{
    if (p->timing && p->ev_used >= 8192) return harvest_timing(p);
    return 0;
}

int check_shots(const pyfwi_plan *p, const pyfwi_shot *shots, int nshots)
{
    const pyfwi_grid &g = p->grid;
    if (!shots || nshots <= 0) return fail(PYFWI_EINVAL, "no shots");
    // launches put the shots of a pass in grid.y, at most 65535
    if (nshots > 65535) return fail(PYFWI_EINVAL, "at most 65535 shots a call");
    for (int s = 0; s < nshots; ++s) {
        const pyfwi_shot &h = shots[s];
        std::string w = "shot " + std::to_string(s) + ": ";
        if (h.nt < 1) return fail(PYFWI_EINVAL, w + "nt must be >= 1");
        if (h.dt_scale < 1) return fail(PYFWI_EINVAL, w + "dt_scale must be >= 1");
        if (h.n_samples < 1) return fail(PYFWI_EINVAL, w + "n_samples must be >= 1");
        if (h.src_x < 0 || h.src_x >= g.tnx || h.src_z < 0 || h.src_z >= g.tnz)
            return fail(PYFWI_EINVAL, w + "source cell outside the grid");
        if (h.n_rec < 0 || h.rec_stride < h.n_rec)
            return fail(PYFWI_EINVAL, w + "need 0 <= n_rec <= rec_stride");
        if (h.rec_flags & ~PYFWI_REC_VZ_FROM_VX) return fail(PYFWI_EINVAL, w + "unknown receiver flag");
        if (h.rec_layout == PYFWI_REC_TABLE) {
            if (h.rec_table < 0 ||
                h.rec_table + 2 * (int64_t)h.n_rec > (int64_t)p->h_rtab.size())
                return fail(PYFWI_EINVAL, w + "receiver entries outside the plan's receiver table");
        } else if (h.rec_layout != PYFWI_REC_LINE) {
            return fail(PYFWI_EINVAL, w + "unknown receiver layout");
        } else if (h.n_rec > 0) {
            if (h.rec_dx < 1) return fail(PYFWI_EINVAL, w + "receiver spacing must be >= 1 cell");
            int64_t last = (int64_t)h.rec_x0 + (int64_t)(h.n_rec - 1) * h.rec_dx;
            if (h.rec_z < 0 || h.rec_z >= g.tnz || h.rec_x0 < 0 || last >= g.tnx)
                return fail(PYFWI_EINVAL, w + "receiver line leaves the grid");
        }
        if (h.src_offset < 0 || h.rec_offset < 0) return fail(PYFWI_EINVAL, w + "negative offset");
    }
    return 0;
}

size_t src_extent(const pyfwi_shot *shots, int nshots)
{
    size_t m = 0;
    for (int s = 0; s < nshots; ++s) {
        size_t ninj = (size_t)((shots[s].nt + shots[s].dt_scale - 1) / shots[s].dt_scale);
        m = std::max(m, (size_t)shots[s].src_offset + ninj * 5);
    }
    return m;
}

size_t rec_extent(const pyfwi_shot *shots, int nshots)
{
    size_t m = 0;
    for (int s = 0; s < nshots; ++s)
        m = std::max(m, (size_t)shots[s].rec_offset +
                            (size_t)5 * shots[s].n_samples * shots[s].rec_stride);
    return m;
}

// Where the shots' sparse terms go: the stencil kernels apply sources on
// active cells and record / inject receiver lines on active cells; other
// shots take the in-memory gather and injection launches.
void shot_flags(pyfwi_plan *p, const pyfwi_shot *shots, int nshots)
{
    const int sdo = p->grid.sdo, nz = p->grid.tnz, nx = p->grid.tnx;
    auto inside = [&](int z, int x) { return z >= sdo && z <= nz - sdo && x >= sdo && x <= nx - sdo; };
    p->src_inside = p->rec_inline = true;
    for (int k = 0; k < nshots; ++k) {
        const pyfwi_shot &sh = shots[k];
        if (!inside(sh.src_z, sh.src_x)) p->src_inside = false;
        if (sh.rec_layout != PYFWI_REC_LINE || sh.rec_flags != 0) {
            p->rec_inline = false;
            continue;
        }
        const int last = sh.rec_x0 + (sh.n_rec > 0 ? sh.n_rec - 1 : 0) * sh.rec_dx;
        if (sh.n_rec > 0 && (!inside(sh.rec_z, sh.rec_x0) || !inside(sh.rec_z, last)))
            p->rec_inline = false;
    }
}

int upload_shots(pyfwi_plan *p, const pyfwi_shot *shots, int nshots)
{
    p->h_shots.assign(shots, shots + nshots);
    shot_flags(p, shots, nshots);
    if (nshots > p->shots_cap) {
        if (p->d_shots) (void)hipFree(p->d_shots);
        p->d_shots = nullptr;
        HIPCHK(hipMalloc(&p->d_shots, sizeof(pyfwi_shot) * nshots));
        p->shots_cap = nshots;
    }
    HIPCHK(hipMemcpyAsync(p->d_shots, shots, sizeof(pyfwi_shot) * nshots,
                          hipMemcpyHostToDevice, p->stream));
    // the caller's table may be freed as soon as the entry point returns
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

// Device pointer for a caller array: the pointer itself for device memory,
// otherwise a staged copy (direction given by `in`).
// With PYFWI_MEM_DEVICE every caller pointer must be memory HIP knows
// (device, managed or pinned host): the kernels dereference it as it is, and
// a plain host allocation read by a kernel faults the GPU.  Null passes.
int check_device_ptrs(int mem, std::initializer_list<std::pair<const void *, const char *>> ptrs)
{
    if (mem != PYFWI_MEM_DEVICE) return 0;
    for (const auto &pw : ptrs) {
        if (!pw.first) continue;
        hipPointerAttribute_t a{};
        const hipError_t e = hipPointerGetAttributes(&a, pw.first);
        if (e != hipSuccess) (void)hipGetLastError();
        if (e != hipSuccess || a.type == hipMemoryTypeUnregistered)
            return fail(PYFWI_EINVAL, std::string(pw.second) +
                                          " is not memory the GPU can address (PYFWI_MEM_DEVICE "
                                          "needs a device, managed or pinned pointer)");
    }
    return 0;
}

int stage_in(pyfwi_plan *p, DevBuf &stage, const float *ptr, size_t n, int mem,
             const float **out, const char *what)
{
    if (mem == PYFWI_MEM_DEVICE) {
        *out = ptr;
        return 0;
    }
    if (int e = ensure(stage, n, p->stream, what)) return e;
    HIPCHK(hipMemcpyAsync(stage.p, ptr, n * sizeof(float), hipMemcpyHostToDevice, p->stream));
    *out = stage.p;
    return 0;
}

// ---------------------------------------------------------- graph replay
// With graphs on, a time loop is captured into a hipGraph the first time it
// runs with a given configuration and replayed by one hipGraphLaunch
// afterwards: the loop issues thousands of small launches, which the host
// otherwise submits one by one.  The key hashes everything the loop's
// launches take as arguments (buffers, shot table, model, steps, flags), so
// a changed call captures a new graph.  The loops hold no allocation, memset
// or synchronisation; capture happens on a private stream and the graph runs
// on the plan's stream.

struct KeyHash {
    uint64_t h = 1469598103934665603ull;  // FNV-1a
    KeyHash &add(const void *data, size_t n)
    {
        const unsigned char *c = static_cast<const unsigned char *>(data);
        for (size_t k = 0; k < n; ++k) h = (h ^ c[k]) * 1099511628211ull;
        return *this;
    }
    template <class T> KeyHash &val(const T &v) { return add(&v, sizeof(T)); }
};

// Hash of the plan state every captured loop depends on.
KeyHash loop_key(const pyfwi_plan *p, int kind, int g0, int ng)
{
    KeyHash k;
    k.val(kind).val(g0).val(ng).val(p->dg).val(p->src_inside).val(p->rec_inline)
        .val(p->d_shots).val(p->d_rtab).val(p->lam.p).val(p->mu.p).val(p->bu.p)
        .val(p->dxp).val(p->dzp).val(p->fields.p).val(p->adj.p).val(p->adj2.p)
        .val(p->adj_fused).val(p->gacc.p);
    k.add(p->h_shots.data() + g0, sizeof(pyfwi_shot) * (size_t)ng);
    return k;
}

constexpr size_t GRAPH_CACHE_MAX = 256;

// Run `body` (a loop of launches on p->stream) directly, or as a cached
// graph when graphs are on.
template <class F>
int run_loop(pyfwi_plan *p, uint64_t key, F body)
{
    if (!p->graphs || p->timing || p->tune_fused) return body();
    for (auto &ge : p->graph_cache)
        if (ge.first == key) {
            HIPCHK(hipGraphLaunch(ge.second, p->stream));
            p->graph_replays++;
            return 0;
        }
    if (!p->cap_stream) HIPCHK(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
    hipStream_t user = p->stream;
    p->stream = p->cap_stream;
    hipError_t be = hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal);
    if (be != hipSuccess) {
        p->stream = user;
        return fail(PYFWI_EHIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(be));
    }
    const int e = body();
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(p->cap_stream, &graph);
    p->stream = user;
    if (e) {
        if (ce == hipSuccess && graph) (void)hipGraphDestroy(graph);
        return e;
    }
    if (ce != hipSuccess)
        return fail(PYFWI_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess)
        return fail(PYFWI_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    if (p->graph_cache.size() >= GRAPH_CACHE_MAX) {
        HIPCHK(hipStreamSynchronize(p->stream));
        for (auto &ge : p->graph_cache) (void)hipGraphExecDestroy(ge.second);
        p->graph_cache.clear();
    }
    p->graph_cache.emplace_back(key, exec);
    p->graph_captures++;
    HIPCHK(hipGraphLaunch(exec, p->stream));
    return 0;
}

} // namespace

// =================================================================== C ABI

extern "C" {

int32_t pyfwi_version(void) { return 100; }

const char *pyfwi_last_error(void) { return g_err.c_str(); }

int32_t pyfwi_plan_create(const pyfwi_grid *grid, void *stream, pyfwi_plan **out)
{
    if (!grid || !out) return fail(PYFWI_EINVAL, "null argument");
    *out = nullptr;
    const pyfwi_grid &g = *grid;
    if (g.sdo != 2 && g.sdo != 4) return fail(PYFWI_EINVAL, "sdo must be 2 (order 4) or 4 (order 8)");
    if (g.tnz < 2 * g.sdo + 1 || g.tnx < 2 * g.sdo + 1)
        return fail(PYFWI_EINVAL, "grid smaller than the stencil");
    if (g.npml < 0) return fail(PYFWI_EINVAL, "npml must be >= 0");
    if (!(g.dt > 0) || !(g.dx > 0) || !(g.dz > 0) || !std::isfinite(g.dt))
        return fail(PYFWI_EINVAL, "dt, dx, dz must be positive");
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    pyfwi_plan *p = new pyfwi_plan();
    p->grid = g;
    p->device = dev;
    DevGrid &d = p->dg;
    d.nz = g.tnz;
    d.nx = g.tnx;
    d.pitch = ((g.tnx + HALO + 63) / 64) * 64;  // >= tnx + 8 zero columns
    d.sdo = g.sdo;
    d.npml = g.npml;
    // column sdo (the first active one, where every column block starts)
    // falls on a 256-byte boundary, so each wavefront's row segment covers
    // whole cache lines
    d.origin = (int64_t)HALO * d.pitch + (64 - g.sdo);
    d.fstride = (int64_t)(g.tnz + 2 * HALO) * d.pitch + 64;  // tail for the shifted last row
    d.fstride = ((d.fstride + 63) / 64) * 64;
    for (int k = 0; k < 4; ++k) {
        d.cx[k] = (float)(g.c[k] / g.dx);
        d.cz[k] = (float)(g.c[k] / g.dz);
    }
    d.dt = (float)g.dt;
    if (ncu > 0) d.ncu = ncu;
    (void)scratch_on(dev);  // set-up of pyfwi_forward_noalloc on this device
    if (stream) {
        p->stream = (hipStream_t)stream;
    } else {
        hipError_t e = hipStreamCreate(&p->own_stream);
        if (e != hipSuccess) {
            delete p;
            return fail(PYFWI_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        }
        p->stream = p->own_stream;
    }
    *out = p;
    return 0;
}

int32_t pyfwi_plan_destroy(pyfwi_plan *p)
{
    if (!p) return 0;
    (void)hipStreamSynchronize(p->stream);
    for (DevBuf *b : {&p->fir_taps, &p->filt_a, &p->filt_b})
        b->release();
    for (DevBuf *b : {&p->lam, &p->mu, &p->bu, &p->dxp_buf, &p->dzp_buf, &p->scratch_in, &p->fields, &p->fields2,
                      &p->adj, &p->adj2, &p->av0, &p->av1, &p->gacc, &p->snaps, &p->ckpts, &p->seg,
                      &p->src_stage, &p->seis_stage, &p->res_stage, &p->grad_stage,
                      &p->shot_tab_f, &p->resid})
        b->release();
    if (p->d_misfit) (void)hipFree(p->d_misfit);
    if (p->d_check) (void)hipFree(p->d_check);
    if (p->d_shots) (void)hipFree(p->d_shots);
    if (p->d_rtab) (void)hipFree(p->d_rtab);
    for (auto e : p->ev_pool) (void)hipEventDestroy(e);
    for (auto &ge : p->graph_cache) (void)hipGraphExecDestroy(ge.second);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    if (p->own_stream) (void)hipStreamDestroy(p->own_stream);
    delete p;
    return 0;
}

int32_t pyfwi_plan_set_stream(pyfwi_plan *p, void *stream)
{
    if (!p) return fail(PYFWI_EINVAL, "null plan");
    HIPCHK(hipStreamSynchronize(p->stream));
    p->stream = stream ? (hipStream_t)stream : p->own_stream;
    if (!p->stream) {
        HIPCHK(hipStreamCreate(&p->own_stream));
        p->stream = p->own_stream;
    }
    return 0;
}

int32_t pyfwi_set_receivers(pyfwi_plan *p, const int32_t *table, int64_t count)
{
    if (!p) return fail(PYFWI_EINVAL, "null plan");
    if (count < 0 || (count > 0 && !table)) return fail(PYFWI_EINVAL, "bad receiver table");
    const int64_t ncell = (int64_t)p->grid.tnz * p->grid.tnx;
    for (int64_t k = 0; k < count; ++k)
        if (table[k] < -1 || table[k] >= ncell)
            return fail(PYFWI_EINVAL, "receiver table entry " + std::to_string(k) +
                                          " is not a cell of the grid (or -1)");
    // kernels of earlier calls may still read the old table
    HIPCHK(hipStreamSynchronize(p->stream));
    if (count > p->rtab_cap) {
        if (p->d_rtab) HIPCHK(hipFree(p->d_rtab));
        p->d_rtab = nullptr;
        p->rtab_cap = 0;
        if (hipMalloc(&p->d_rtab, (size_t)count * sizeof(int32_t)) != hipSuccess) {
            p->d_rtab = nullptr;
            p->h_rtab.clear();
            return fail(PYFWI_ENOMEM, "hipMalloc failed for the receiver table");
        }
        p->rtab_cap = count;
    }
    p->h_rtab.assign(table, table + count);
    if (count > 0)
        HIPCHK(hipMemcpy(p->d_rtab, table, (size_t)count * sizeof(int32_t), hipMemcpyHostToDevice));
    return 0;
}

int32_t pyfwi_set_model(pyfwi_plan *p, const float *lam, const float *mu, const float *buoy,
                        const float *dx_prof, const float *dz_prof, int32_t mem)
{
    if (!p || !lam || !mu || !buoy || !dx_prof || !dz_prof) return fail(PYFWI_EINVAL, "null argument");
    if (int e = check_device_ptrs(mem, {{lam, "lam"}, {mu, "mu"}, {buoy, "buoyancy"},
                                        {dx_prof, "dx profile"}, {dz_prof, "dz profile"}}))
        return e;
    const DevGrid &g = p->dg;
    const size_t n = (size_t)g.nz * g.nx;
    if (mem == PYFWI_MEM_HOST) {
        for (const float *a : {lam, mu, buoy})
            for (size_t k = 0; k < n; ++k)
                if (!std::isfinite(a[k])) return fail(PYFWI_ENONFINITE, "model holds NaN or Inf");
        for (int k = 0; k < g.nx; ++k)
            if (!std::isfinite(dx_prof[k])) return fail(PYFWI_ENONFINITE, "damping holds NaN or Inf");
        for (int k = 0; k < g.nz; ++k)
            if (!std::isfinite(dz_prof[k])) return fail(PYFWI_ENONFINITE, "damping holds NaN or Inf");
    }
    for (DevBuf *b : {&p->lam, &p->mu, &p->bu})
        if (int e = ensure(*b, g.fstride, p->stream, "model")) return e;
    if (int e = ensure(p->dxp_buf, g.nx + 2 * HALO, p->stream, "damping")) return e;
    if (int e = ensure(p->dzp_buf, g.nz + 2 * HALO, p->stream, "damping")) return e;
    p->dxp = p->dxp_buf.p + HALO;
    p->dzp = p->dzp_buf.p + HALO;
    const hipMemcpyKind kind = mem == PYFWI_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (mem == PYFWI_MEM_DEVICE) {
        launch_pad(g, lam, p->lam.p, p->stream);
        launch_pad(g, mu, p->mu.p, p->stream);
        launch_pad(g, buoy, p->bu.p, p->stream);
    } else {
        if (int e = ensure(p->scratch_in, n, p->stream, "staging")) return e;
        const float *srcs[3] = {lam, mu, buoy};
        DevBuf *dsts[3] = {&p->lam, &p->mu, &p->bu};
        for (int k = 0; k < 3; ++k) {
            HIPCHK(hipMemcpyAsync(p->scratch_in.p, srcs[k], n * sizeof(float), kind, p->stream));
            launch_pad(g, p->scratch_in.p, dsts[k]->p, p->stream);
        }
    }
    HIPCHK(hipMemcpyAsync(p->dxp, dx_prof, g.nx * sizeof(float), kind, p->stream));
    HIPCHK(hipMemcpyAsync(p->dzp, dz_prof, g.nz * sizeof(float), kind, p->stream));
    p->model_set = false;
    p->have_forward = false;
    // largest P velocity (and NaN / Inf on device arrays) for the stability check
    if (!p->d_check) HIPCHK(hipMalloc(&p->d_check, 2 * sizeof(int)));
    launch_zero(reinterpret_cast<float *>(p->d_check), 2, p->stream);
    launch_model_check(g, p->lam.p, p->mu.p, p->bu.p, p->d_check, p->stream);
    HIPCHK(hipGetLastError());
    int chk[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(chk, p->d_check, sizeof(chk), hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));  // also: host arrays may go away
    if (chk[1]) return fail(PYFWI_ENONFINITE, "model holds NaN or Inf");
    float vp2 = 0.0f;
    std::memcpy(&vp2, &chk[0], sizeof(float));
    const pyfwi_grid &gr = p->grid;
    const double csum = std::fabs(gr.c[0]) + std::fabs(gr.c[1]) + std::fabs(gr.c[2]) +
                        std::fabs(gr.c[3]);
    const double a = csum / (gr.dx * gr.dx) + csum / (gr.dz * gr.dz);
    const double vmax = std::sqrt((double)vp2);
    if (vmax > 0) {
        const double dt_max = 2.0 / vmax / std::sqrt(a * 5.0);
        if (gr.dt > dt_max * (1.0 + 1e-6))
            return fail(PYFWI_ECFL, "dt = " + std::to_string(gr.dt) +
                                        " exceeds the stability limit " + std::to_string(dt_max) +
                                        " of this model (largest P velocity " +
                                        std::to_string(vmax) + ")");
    }
    p->model_set = true;
    return 0;
}

namespace {

// One forward step of the shots of a pass, the reference's loop body at
// wave_propagation.py:704-708: gather + inject on the current state, then the
// velocity and stress half steps.  The two-launch form applies the source
// term to the values it loads and writes the next step's record sample from
// the stress update (receiver tables: a gather launch first); the fused form,
// and sources off the active cells, record and inject in memory first.
int forward_step(pyfwi_plan *p, const pyfwi_shot *dsh, int ng, int t, const float *src,
                 float *seis)
{
    const DevGrid &g = p->dg;
    const int64_t fps = fields_per_shot(p);
    if (!p->tune_fused && !p->src_inside) {
        {
            Timed ts(p, PYFWI_K_FWD_SPARSE);
            launch_inject_record(g, ng, dsh, t, p->cur, fps, src, seis, p->d_rtab, p->stream);
        }
        {
            Timed tm(p, PYFWI_K_FWD_VEL);
            launch_vel(g, +1, ng, dsh, t, p->cur, fps, p->cur, fps, p->bu.p, p->dxp, p->dzp,
                       p->stream);
        }
        Timed tm(p, PYFWI_K_FWD_STR);
        launch_str(g, +1, ng, dsh, t, p->cur, fps, p->cur, fps, p->lam.p, p->mu.p, p->dxp, p->dzp,
                   p->stream);
    } else if (p->tune_fused) {
        // sources on the active cells: the kernel adds the source term to
        // the values it loads, and writes the record sample itself unless
        // the receivers need a gather launch
        const bool inl = p->src_inside, rec = inl && p->rec_inline;
        if (!rec) {
            Timed ts(p, PYFWI_K_FWD_SPARSE);
            launch_inject_record(g, ng, dsh, t, p->cur, fps, inl ? nullptr : src, seis,
                                 p->d_rtab, p->stream);
        }
        float *nxt = p->cur == p->fields.p ? p->fields2.p : p->fields.p;
        Timed tm(p, PYFWI_K_FWD_STEP);
        launch_fused_step(g, +1, ng, dsh, t, p->cur, fps, nxt, fps, p->bu.p, p->lam.p, p->mu.p,
                          p->dxp, p->dzp, p->stream, inl ? src : nullptr, rec ? seis : nullptr);
        p->cur = nxt;
    } else {
        if (!p->rec_inline) {
            Timed ts(p, PYFWI_K_FWD_SPARSE);
            launch_inject_record(g, ng, dsh, t, p->cur, fps, nullptr, seis, p->d_rtab, p->stream);
        }
        {
            Timed tm(p, PYFWI_K_FWD_VEL);
            launch_vel(g, +1, ng, dsh, t, p->cur, fps, p->cur, fps, p->bu.p, p->dxp, p->dzp,
                       p->stream, src);
        }
        Timed tm(p, PYFWI_K_FWD_STR);
        launch_str(g, +1, ng, dsh, t, p->cur, fps, p->cur, fps, p->lam.p, p->mu.p, p->dxp, p->dzp,
                   p->stream, src, p->rec_inline ? seis : nullptr);
    }
    return 0;
}

// Forward step of a kept-state run: the state of step t-1 is read from `in`
// (per-shot stride in_stride) and step t written to `out`; the source term
// never enters memory, so every slot holds the reference's snapshot.  The
// stress update writes the record sample when the receivers lie on one row
// of the active cells; otherwise the caller gathers the records from the
// slots after the loop (seis == nullptr here).
int forward_step_kept(pyfwi_plan *p, const pyfwi_shot *dsh, int ng, int t, const float *src,
                      float *seis, const float *in, int64_t in_stride, float *out,
                      int64_t out_stride)
{
    const DevGrid &g = p->dg;
    if (p->tune_fused) {
        Timed tm(p, PYFWI_K_FWD_STEP);
        launch_fused_step(g, +1, ng, dsh, t, in, in_stride, out, out_stride, p->bu.p, p->lam.p,
                          p->mu.p, p->dxp, p->dzp, p->stream, src, seis);
        return 0;
    }
    {
        Timed tm(p, PYFWI_K_FWD_VEL);
        launch_vel(g, +1, ng, dsh, t, in, in_stride, out, out_stride, p->bu.p, p->dxp, p->dzp,
                   p->stream, src);
    }
    Timed tm(p, PYFWI_K_FWD_STR);
    launch_str(g, +1, ng, dsh, t, in, in_stride, out, out_stride, p->lam.p, p->mu.p, p->dxp,
               p->dzp, p->stream, src, seis);
    return 0;
}

// Time-reversed step (wave_propagation.py:780-812).
int reverse_step(pyfwi_plan *p, const pyfwi_shot *dsh, int ng, int t)
{
    const DevGrid &g = p->dg;
    if (p->tune_fused) {
        float *nxt = p->cur == p->fields.p ? p->fields2.p : p->fields.p;
        Timed tm(p, PYFWI_K_REV_STEP);
        const int64_t fps = fields_per_shot(p);
        launch_fused_step(g, -1, ng, dsh, t, p->cur, fps, nxt, fps, p->bu.p, p->lam.p, p->mu.p,
                          p->dxp, p->dzp, p->stream);
        p->cur = nxt;
    } else {
        const int64_t fps = fields_per_shot(p);
        {
            Timed tm(p, PYFWI_K_REV_STR);
            launch_str(g, -1, ng, dsh, t, p->cur, fps, p->cur, fps, p->lam.p, p->mu.p, p->dxp,
                       p->dzp, p->stream);
        }
        Timed tm(p, PYFWI_K_REV_VEL);
        launch_vel(g, -1, ng, dsh, t, p->cur, fps, p->cur, fps, p->bu.p, p->dxp, p->dzp,
                   p->stream);
    }
    return 0;
}

// True when the checkpoint list is every step 0..n-1 with n >= nt (slot t =
// state after step t), so it keeps every step of shots of up to nt steps.
bool keeps_every_step(const std::vector<int32_t> &cl, int nt)
{
    if ((int)cl.size() < nt || cl.empty()) return false;
    for (int k = 0; k < (int)cl.size(); ++k)
        if (cl[k] != k) return false;
    return true;
}

void every_step(int nt, std::vector<int32_t> &cl)
{
    cl.resize(nt);
    for (int k = 0; k < nt; ++k) cl[k] = k;
}

// A checkpoint list other than every step is a list of the reference's
// time indices, which only has a meaning when all shots share nt; "every
// step" keeps each shot's own steps 0..nt_s-1.
int check_chp_nt(const std::vector<pyfwi_shot> &sh, const std::vector<int32_t> &cl, int nt)
{
    if (cl.empty() || keeps_every_step(cl, nt)) return 0;
    for (auto &x : sh)
        if (x.nt != nt)
            return fail(PYFWI_EINVAL, "a checkpoint list other than every step 0..nt-1 needs "
                                      "equal nt for all shots");
    return 0;
}

// The shots of a call run in the order given.
void identity_order(pyfwi_plan *p, int nshots)
{
    p->last_order.resize((size_t)nshots);
    for (int k = 0; k < nshots; ++k) p->last_order[(size_t)k] = k;
}

int pass_size(const pyfwi_plan *p, int nshots)
{
    return p->tune_group > 0 ? std::min(p->tune_group, nshots) : nshots;
}

// Zeroed state buffers for a pass of ng shots.
int prepare_fields(pyfwi_plan *p, int ng)
{
    const int64_t fps = fields_per_shot(p);
    if (int e = ensure(p->fields, (size_t)ng * fps, p->stream, "fields")) return e;
    if (p->tune_fused)
        if (int e = ensure(p->fields2, (size_t)ng * fps, p->stream, "fields")) return e;
    launch_zero(p->fields.p, (int64_t)ng * fps, p->stream);
    if (p->tune_fused)
        launch_zero(p->fields2.p, (int64_t)ng * fps, p->stream);
    p->cur = p->fields.p;
    return 0;
}

// Time-major layout of the kept states of a pass that keeps every step:
// state t of the shots live at step t (t < nt_s) side by side, the pass's
// shots ordered by decreasing nt so that they are a prefix.  State t of shot
// s lies off[t] + s states from the buffer start, so a pass holds the sum of
// its shots' nt states, where one slot per step of the longest shot for
// every shot would hold ng * max nt.  The launches of step t then advance
// only its live[t] shots (the forward's step t and the backward's step t
// both take the shots with t < nt_s), so a step where one shot is left runs
// the one-shot kernels.
struct TimeMajor {
    std::vector<int64_t> off;
    std::vector<int> live;
};

int time_major(const pyfwi_shot *sh, int ng, TimeMajor &tm)
{
    const int nt = sh[0].nt;
    for (int k = 1; k < ng; ++k)
        if (sh[k].nt > sh[k - 1].nt)
            return fail(PYFWI_EINVAL, "time-major kept states need shots by decreasing nt");
    tm.off.assign((size_t)nt, 0);
    tm.live.assign((size_t)nt, 0);
    int64_t cum = 0;
    int live = ng;
    for (int t = 0; t < nt; ++t) {
        while (live > 0 && sh[live - 1].nt <= t) --live;
        tm.off[(size_t)t] = cum;
        tm.live[(size_t)t] = live;
        cum += live;
    }
    return 0;
}

// Forward time loop of one pass of ng shots starting at shot g0.  `snaps`
// and `ckpts` (or null) have per-shot strides nchp*fps and ncp*fps, or, with
// `tm` (every step kept), the kept states are laid out time-major.
int forward_pass(pyfwi_plan *p, int g0, int ng, const float *d_src, float *d_seis,
                 const std::vector<int32_t> &cl, float *snaps, float *ckpts, int ncp, int K,
                 const TimeMajor *tm = nullptr)
{
    const int64_t fps = fields_per_shot(p);
    const pyfwi_shot *dsh = p->d_shots + g0;
    std::vector<pyfwi_shot> sh(p->h_shots.begin() + g0, p->h_shots.begin() + g0 + ng);
    const int nt = max_nt(sh);
    const int nchp = (int)cl.size();
    if (int e = prepare_fields(p, ng)) return e;
    Timed loop_tm(p, PYFWI_K_FWD_LOOP, 2, nt);
    KeyHash key = loop_key(p, 1, g0, ng);
    key.val(d_src).val(d_seis).val(snaps).val(ckpts).val(ncp).val(K).val(tm != nullptr);
    key.add(cl.data(), cl.size() * sizeof(int32_t));
    // every step kept (chpr=100): the forward kernels step from one slot
    // into the next, so the kept states cost no extra traffic
    if (snaps && keeps_every_step(cl, nt) && p->src_inside) {
        return run_loop(p, key.h, [&]() -> int {
            const int64_t ss = tm ? fps : (int64_t)nchp * fps;  // between shots
            const float *in = p->fields.p;
            int64_t in_stride = fps;
            for (int t = 0; t < nt;) {
                if (ckpts && t % K == 0)
                    launch_copy_shots(ckpts + (int64_t)(t / K) * fps, (int64_t)ncp * fps, in,
                                      in_stride, fps, ng, p->stream);
                const int live = tm ? tm->live[(size_t)t] : ng;
                float *out = snaps + (tm ? tm->off[(size_t)t] : (int64_t)t) * fps;
                // two steps a launch when the same shots are live at both
                if (p->fwd_two && p->dg.sdo == 2 && t + 1 < nt &&
                    (tm ? tm->live[(size_t)t + 1] : ng) == live) {
                    float *out1 = snaps + (tm ? tm->off[(size_t)t + 1] : (int64_t)t + 1) * fps;
                    {
                        Timed tk(p, PYFWI_K_FWD_TWO);
                        launch_fwd_two(p->dg, live, dsh, t, in, in_stride, out, out1, ss,
                                       p->bu.p, p->lam.p, p->mu.p, p->dxp, p->dzp, p->stream,
                                       d_src, p->rec_inline ? d_seis : nullptr);
                    }
                    if (ckpts && (t + 1) % K == 0)
                        launch_copy_shots(ckpts + (int64_t)((t + 1) / K) * fps,
                                          (int64_t)ncp * fps, out, ss, fps, ng, p->stream);
                    in = out1;
                    t += 2;
                } else {
                    forward_step_kept(p, dsh, live, t, d_src, p->rec_inline ? d_seis : nullptr,
                                      in, in_stride, out, ss);
                    in = out;
                    t += 1;
                }
                in_stride = ss;
                if (int e = maybe_harvest(p)) return e;
            }
            // receiver tables: one gather launch over all steps instead of one
            // per step (time-major: one per step, the step's states side by side)
            if (!p->rec_inline) {
                Timed ts(p, PYFWI_K_FWD_SPARSE);
                if (!tm) {
                    launch_record_kept(p->dg, ng, dsh, 0, nt, snaps, ss, fps, d_seis, p->d_rtab,
                                       p->stream);
                } else {
                    // the gather of step t reads state t-1 (none for t = 0)
                    for (int t = 0; t < nt; ++t)
                        launch_record_kept(p->dg, ng, dsh, t, t + 1,
                                           snaps + (t > 0 ? tm->off[(size_t)t - 1] - (t - 1) : 0) * fps,
                                           fps, fps, d_seis, p->d_rtab, p->stream);
                }
            }
            return 0;
        });
    }
    return run_loop(p, key.h, [&]() -> int {
        int chpc = 0;
        for (int t = 0; t < nt; ++t) {
            if (ckpts && t % K == 0)
                launch_copy_shots(ckpts + (int64_t)(t / K) * fps, (int64_t)ncp * fps, p->cur, fps,
                                  fps, ng, p->stream);
            forward_step(p, dsh, ng, t, d_src, d_seis);
            if (snaps && tm) {
                // every step kept, time-major: the live shots' states of step t
                launch_copy_shots(snaps + tm->off[(size_t)t] * fps, fps, p->cur, fps, fps,
                                  tm->live[(size_t)t], p->stream);
            } else if (snaps && std::find(cl.begin(), cl.end(), t) != cl.end()) {
                if (chpc < nchp)
                    launch_copy_shots(snaps + (int64_t)chpc * fps, (int64_t)nchp * fps, p->cur,
                                      fps, fps, ng, p->stream);
                chpc++;
            }
            if (int e = maybe_harvest(p)) return e;
        }
        return 0;
    });
}

} // namespace

namespace {

// Bytes of device memory the plan may use for a call: 90 % of what is free
// plus the large buffers it holds and may give back (kept states and
// checkpoints).
double plan_budget(const pyfwi_plan *p)
{
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        (void)hipGetLastError();
        return 0.0;
    }
    const double held = (double)(p->snaps.n + p->ckpts.n + p->seg.n) * sizeof(float);
    return 0.9 * ((double)free_b + held);
}

// Checkpoint interval of the exact adjoint for shots of up to nt steps: the
// two-level schedule keeps ceil(nt/K) checkpoints and recomputes one K-step
// segment at a time into K slots, so K = ceil(sqrt(nt)) minimises the states
// held per shot (ceil(nt/K) + K), at the cost of one extra forward sweep.
int auto_interval(int nt) { return std::max(1, (int)std::ceil(std::sqrt((double)nt))); }

// States (5 fields each) the exact adjoint holds per shot of a pass: the
// checkpoints, the segment, the forward and adjoint states and the two
// adjoint-velocity buffers (2 x 2 of 5 fields), rounded up.
double exact_states_per_shot(int nt, int K, bool with_ckpts)
{
    const int ncp = (nt + K - 1) / K;
    return (with_ckpts ? ncp : 0) + K + 2 + 0.8 + 1.2 / 4;  // gacc: 6 fields per group of >= 4
}

// Shots per pass for the exact adjoint of `nshots` shots of up to nt steps
// within `budget` bytes: as many as fit, then balanced over the passes.
int exact_pass_size(const pyfwi_plan *p, int nshots, int nt, int K, bool with_ckpts,
                    double budget)
{
    const double per = exact_states_per_shot(nt, K, with_ckpts) * (double)fields_per_shot(p) *
                       sizeof(float);
    int g = (int)std::floor(budget / per);
    if (g < 1) return 0;
    g = std::min(g, nshots);
    const int npass = (nshots + g - 1) / g;
    return (nshots + npass - 1) / npass;
}

// Give back the kept-state buffer when a call needs the memory.
void release_snaps(pyfwi_plan *p)
{
    if (!p->snaps.p) return;
    (void)hipStreamSynchronize(p->stream);
    p->snaps.release();
    p->chp.clear();  // a reference-mode gradient now needs a new forward
}

} // namespace

int32_t pyfwi_forward(pyfwi_plan *p, const pyfwi_shot *shots, int32_t nshots, const float *src,
                      float *seis, const int32_t *chp, int32_t nchp, int32_t ckpt_every,
                      int32_t mem)
{
    if (!p || !src || !seis) return fail(PYFWI_EINVAL, "null argument");
    if (int e = check_device_ptrs(mem, {{src, "source table"}, {seis, "records"}})) return e;
    if (!p->model_set) return fail(PYFWI_ESTATE, "pyfwi_set_model must come first");
    if (int e = check_shots(p, shots, nshots)) return e;
    if (nchp < 0 || (nchp > 0 && !chp)) return fail(PYFWI_EINVAL, "bad checkpoint list");
    if (ckpt_every < PYFWI_CKPT_AUTO)
        return fail(PYFWI_EINVAL, "ckpt_every must be >= 0 or PYFWI_CKPT_AUTO");
    std::vector<pyfwi_shot> sh(shots, shots + nshots);
    const int nt = max_nt(sh);
    std::vector<int32_t> cl(chp, chp + std::max(nchp, 0));
    if (int e = check_chp_nt(sh, cl, nt)) return e;
    const int64_t fps = fields_per_shot(p);
    p->have_forward = false;
    p->ckpts_all = false;
    if (nchp > 0)
        if (int e = ensure(p->snaps, (size_t)nshots * nchp * fps, p->stream, "checkpoints")) return e;
    int ncp = 0;
    if (ckpt_every == PYFWI_CKPT_AUTO) {
        // keep every shot's checkpoints only when they and one shot's exact
        // adjoint working set fit; otherwise the gradient recomputes them
        // pass by pass
        ckpt_every = auto_interval(nt);
        const int n_cp = (nt + ckpt_every - 1) / ckpt_every;
        const double state = (double)fps * sizeof(float);
        const double need = ((double)nshots * n_cp +
                             exact_states_per_shot(nt, ckpt_every, false)) * state;
        if (need > plan_budget(p) - (double)p->snaps.n * sizeof(float)) ckpt_every = 0;
    }
    if (ckpt_every > 0) {
        ncp = (nt + ckpt_every - 1) / ckpt_every;
        if (int e = ensure(p->ckpts, (size_t)nshots * ncp * fps, p->stream, "checkpoints")) return e;
    }
    if (int e = upload_shots(p, shots, nshots)) return e;
    const size_t nsrc = src_extent(shots, nshots), nrec = rec_extent(shots, nshots);
    const float *d_src = nullptr;
    if (int e = stage_in(p, p->src_stage, src, nsrc, mem, &d_src, "source table")) return e;
    float *d_seis = seis;
    if (mem != PYFWI_MEM_DEVICE) {
        if (int e = ensure(p->seis_stage, nrec, p->stream, "records")) return e;
        d_seis = p->seis_stage.p;
    }
    launch_zero(d_seis, (int64_t)nrec, p->stream);
    const int G = pass_size(p, nshots);
    p->last_passes.clear();
    identity_order(p, nshots);
    p->last_ckpt_every = ckpt_every;
    for (int g0 = 0; g0 < nshots; g0 += G) {
        const int ng = std::min(G, nshots - g0);
        p->last_passes.push_back(ng);
        float *sn = nchp > 0 ? p->snaps.p + (int64_t)g0 * nchp * fps : nullptr;
        float *ck = ncp > 0 ? p->ckpts.p + (int64_t)g0 * ncp * fps : nullptr;
        if (int e = forward_pass(p, g0, ng, d_src, d_seis, cl, sn, ck, ncp, ckpt_every)) return e;
    }
    HIPCHK(hipGetLastError());
    if (mem != PYFWI_MEM_DEVICE) {
        HIPCHK(hipMemcpyAsync(seis, d_seis, nrec * sizeof(float), hipMemcpyDeviceToHost, p->stream));
        HIPCHK(hipStreamSynchronize(p->stream));
    }
    if (int e = harvest_timing(p)) return e;
    p->last_shots = sh;
    p->chp = cl;
    p->ckpt_every = ckpt_every;
    p->ncp = ncp;
    p->ckpts_all = ncp > 0;
    p->have_forward = true;
    return 0;
}

namespace {

// Reference adjoint of one pass (shots g0 .. g0+ng-1) into gacc[adj_acc_sets(g.sdo, ng)].
// `snaps` holds the pass's kept states, per-shot stride cl.size() states, at
// the steps of `cl`, or laid out time-major (`tm`, every step kept).
int gradient_pyfwi(pyfwi_plan *p, int g0, int ng, const float *d_res, bool precond,
                   const float *snaps, const std::vector<int32_t> &cl,
                   const TimeMajor *tm = nullptr)
{
    const DevGrid &g = p->dg;
    const int64_t fps = fields_per_shot(p);
    const pyfwi_shot *dsh = p->d_shots + g0;
    const int nt = max_nt(std::vector<pyfwi_shot>(p->h_shots.begin() + g0,
                                                  p->h_shots.begin() + g0 + ng));
    const int nchp = (int)cl.size();
    if (int e = ensure(p->adj, (size_t)ng * fps, p->stream, "adjoint fields")) return e;
    launch_zero(p->adj.p, (int64_t)ng * fps, p->stream);
    // one launch a step (order 4, receivers on the active cells): the
    // adjoint state moves between two buffers, read from one and written to
    // the other, so both start at zero and stay zero off the active cells
    const bool one_launch = p->adj_fused && p->rec_inline && g.sdo == 2;
    if (one_launch) {
        if (int e = ensure(p->adj2, (size_t)ng * fps, p->stream, "adjoint fields")) return e;
        launch_zero(p->adj2.p, (int64_t)ng * fps, p->stream);
    }
    if (int e = prepare_fields(p, ng)) return e;
    // steps at which a shot of the pass starts its backward sweep
    std::vector<char> first_step(nt + 1, 0);
    for (int k = 0; k < ng; ++k) {
        const int ft = p->h_shots[g0 + k].nt - 1;
        if (ft >= 1 && ft <= nt) first_step[ft] = 1;
    }
    // every step kept: read the forward state of step t from its slot
    const bool direct = keeps_every_step(cl, nt);
    Timed loop_tm(p, PYFWI_K_BWD_LOOP, 2, nt - 1);
    KeyHash key = loop_key(p, 2, g0, ng);
    key.val(d_res).val(precond).val(snaps).val(tm != nullptr);
    key.add(cl.data(), cl.size() * sizeof(int32_t));
    return run_loop(p, key.h, [&]() -> int {
    int chpc = nchp - 1;
    float *ain = p->adj.p, *aout = one_launch ? p->adj2.p : p->adj.p;
    for (int t = nt - 1; t > 0; --t) {
        const float *fwd = p->cur;
        int64_t fwd_stride = fps;
        if (direct) {
            fwd = snaps + (tm ? tm->off[(size_t)t] : (int64_t)t) * fps;
            fwd_stride = tm ? fps : (int64_t)nchp * fps;
        } else if (chpc >= 0 && t == cl[chpc]) {
            launch_copy_shots(p->cur, fps, snaps + (int64_t)chpc * fps, (int64_t)nchp * fps, fps,
                              ng, p->stream);
            chpc--;
            fwd = p->cur;
        } else {
            reverse_step(p, dsh, ng, t);
            fwd = p->cur;
        }
        // the residual of step t: the adjoint kernels add it (the velocity
        // update of step t+1 already wrote the adjoint velocities with it,
        // except for shots starting here); with receivers off the active
        // cells or from a table, in memory every step
        const float *r_on = p->rec_inline ? d_res : nullptr;
        if (!p->rec_inline) {
            Timed ts(p, PYFWI_K_BWD_SPARSE);
            launch_adj_inject(g, ng, dsh, t, ain, d_res, p->d_rtab, p->stream);
        } else if (first_step[t]) {
            Timed ts(p, PYFWI_K_BWD_SPARSE);
            launch_adj_inject(g, ng, dsh, t, ain, d_res, p->d_rtab, p->stream, 0, 2, true);
        }
        // time-major: the live shots of step t only (gacc holds the sets of
        // every launch size up to ng, adj_acc_sets_upto)
        const int nl = tm && direct ? tm->live[(size_t)t] : ng;
        if (one_launch) {
            {
                Timed tmr(p, PYFWI_K_ADJ_STEP);
                if (!launch_adj_fused(g, nl, dsh, t, fwd, fwd_stride, ain, aout, p->lam.p,
                                      p->mu.p, p->bu.p, p->dxp, p->dzp, p->gacc.p, precond,
                                      p->stream, r_on))
                    return fail(PYFWI_EINVAL, "the one-launch adjoint step needs order 4");
            }
            std::swap(ain, aout);
            if (int e = maybe_harvest(p)) return e;
            continue;
        }
        {
            Timed tmr(p, PYFWI_K_ADJ_STR);
            launch_adj_str(g, nl, dsh, t, fwd, fwd_stride, p->adj.p, p->bu.p, p->dxp, p->dzp,
                           p->gacc.p, precond, p->stream, r_on);
        }
        {
            Timed tmr(p, PYFWI_K_ADJ_VEL);
            launch_adj_vel_grad(g, nl, dsh, t, fwd, fwd_stride, p->adj.p, p->lam.p, p->mu.p,
                                p->bu.p, p->dxp, p->dzp, p->gacc.p, precond, p->stream, r_on);
        }
        if (int e = maybe_harvest(p)) return e;
    }
    return 0;
    });
}

// Exact adjoint of one pass, forward states recomputed from checkpoints
// every K steps (ckpts: per-shot stride ncp*fps, first shot of the pass).
int gradient_exact(pyfwi_plan *p, int g0, int ng, const float *d_src, const float *d_res,
                   const float *ckpts, int ncp, int K)
{
    const DevGrid &g = p->dg;
    const int64_t fps = fields_per_shot(p);
    const pyfwi_shot *dsh = p->d_shots + g0;
    std::vector<pyfwi_shot> sh(p->h_shots.begin() + g0, p->h_shots.begin() + g0 + ng);
    const int nt = max_nt(sh);
    if (int e = ensure(p->adj, (size_t)ng * fps, p->stream, "adjoint fields")) return e;
    if (int e = ensure(p->av0, (size_t)ng * 2 * g.fstride, p->stream, "adjoint velocity")) return e;
    if (int e = ensure(p->av1, (size_t)ng * 2 * g.fstride, p->stream, "adjoint velocity")) return e;
    if (int e = ensure(p->seg, (size_t)ng * K * fps, p->stream, "segment")) return e;
    if (int e = ensure(p->fields, (size_t)ng * fps, p->stream, "fields")) return e;
    launch_zero(p->adj.p, (int64_t)ng * fps, p->stream);
    launch_zero(p->av0.p, (int64_t)ng * 2 * g.fstride, p->stream);
    launch_zero(p->av1.p, (int64_t)ng * 2 * g.fstride, p->stream);
    float *av_cur = p->av0.p, *av_nxt = p->av1.p;
    const int64_t seg_stride = (int64_t)K * fps;  // per shot
    for (int c = ncp - 1; c >= 0; --c) {
        const int t0 = c * K, t1 = std::min(nt, t0 + K);
        // recompute the segment: after the velocity update of step t the field
        // buffer holds V_t (velocities) and the stresses that step t updates
        launch_copy_shots(p->fields.p, fps, ckpts + (int64_t)c * fps, (int64_t)ncp * fps, fps, ng,
                          p->stream);
        for (int t = t0; t < t1; ++t) {
            launch_inject_record(g, ng, dsh, t, p->fields.p, fps, d_src, p->seis_stage.p,
                                 p->d_rtab, p->stream);
            {
                Timed tm(p, PYFWI_K_FWD_VEL);
                launch_vel(g, +1, ng, dsh, t, p->fields.p, fps, p->fields.p, fps, p->bu.p, p->dxp,
                           p->dzp, p->stream);
            }
            launch_copy_shots(p->seg.p + (int64_t)(t - t0) * fps, seg_stride, p->fields.p, fps,
                              fps, ng, p->stream);
            if (t + 1 < t1) {
                Timed tm(p, PYFWI_K_FWD_STR);
                launch_str(g, +1, ng, dsh, t, p->fields.p, fps, p->fields.p, fps, p->lam.p,
                           p->mu.p, p->dxp, p->dzp, p->stream);
            }
        }
        for (int t = t1 - 1; t >= t0; --t) {
            const float *slot = p->seg.p + (int64_t)(t - t0) * fps;
            {
                Timed tm(p, PYFWI_K_EXA_STR);
                launch_exact_str(g, ng, dsh, t, slot, seg_stride, p->adj.p, av_cur, p->lam.p,
                                 p->mu.p, p->dxp, p->dzp, p->gacc.p, p->stream);
            }
            {
                Timed tm(p, PYFWI_K_EXA_VEL);
                launch_exact_vel(g, ng, dsh, t, slot + 2 * g.fstride, seg_stride, p->adj.p, av_cur,
                                 av_nxt, p->bu.p, p->dxp, p->dzp, p->gacc.p, p->stream);
            }
            std::swap(av_cur, av_nxt);
            launch_exact_inject(g, ng, dsh, t, p->adj.p, av_cur, d_res, p->d_rtab, p->stream);
            if (int e = maybe_harvest(p)) return e;
        }
    }
    return 0;
}

// Exact adjoint of one pass from the states the forward kept at every step
// (`snaps`: the pass's first shot, per-shot stride nchp*fps, slot t = state
// after step t): step t's velocity half step produced the velocities of slot
// t from the stresses of slot t-1 (zero before step 0) with step t's source
// term, which k_exact_vel adds as it loads them.  No segment is recomputed.
int gradient_exact_kept(pyfwi_plan *p, int g0, int ng, const float *d_src, const float *d_res,
                        const float *snaps, int nchp)
{
    const DevGrid &g = p->dg;
    const int64_t fps = fields_per_shot(p);
    const pyfwi_shot *dsh = p->d_shots + g0;
    const int nt = max_nt(std::vector<pyfwi_shot>(p->h_shots.begin() + g0,
                                                  p->h_shots.begin() + g0 + ng));
    if (nt > nchp) return fail(PYFWI_EINVAL, "kept states shorter than the shots");
    if (int e = ensure(p->adj, (size_t)ng * fps, p->stream, "adjoint fields")) return e;
    if (int e = ensure(p->av0, (size_t)ng * 2 * g.fstride, p->stream, "adjoint velocity")) return e;
    if (int e = ensure(p->av1, (size_t)ng * 2 * g.fstride, p->stream, "adjoint velocity")) return e;
    if (int e = ensure(p->fields, (size_t)ng * fps, p->stream, "fields")) return e;
    launch_zero(p->adj.p, (int64_t)ng * fps, p->stream);
    launch_zero(p->av0.p, (int64_t)ng * 2 * g.fstride, p->stream);
    launch_zero(p->av1.p, (int64_t)ng * 2 * g.fstride, p->stream);
    launch_zero(p->fields.p, (int64_t)ng * fps, p->stream);  // the state before step 0
    float *av_cur = p->av0.p, *av_nxt = p->av1.p;
    const int64_t ss = (int64_t)nchp * fps;  // between shots
    for (int t = nt - 1; t >= 0; --t) {
        const float *vel = snaps + (int64_t)t * fps;
        const float *tau = (t > 0 ? snaps + (int64_t)(t - 1) * fps : p->fields.p) + 2 * g.fstride;
        const int64_t tau_stride = t > 0 ? ss : fps;
        {
            Timed tm(p, PYFWI_K_EXA_STR);
            launch_exact_str(g, ng, dsh, t, vel, ss, p->adj.p, av_cur, p->lam.p, p->mu.p, p->dxp,
                             p->dzp, p->gacc.p, p->stream);
        }
        {
            Timed tm(p, PYFWI_K_EXA_VEL);
            launch_exact_vel(g, ng, dsh, t, tau, tau_stride, p->adj.p, av_cur, av_nxt, p->bu.p,
                             p->dxp, p->dzp, p->gacc.p, p->stream, d_src);
        }
        std::swap(av_cur, av_nxt);
        launch_exact_inject(g, ng, dsh, t, p->adj.p, av_cur, d_res, p->d_rtab, p->stream);
        if (int e = maybe_harvest(p)) return e;
    }
    return 0;
}

} // namespace

int32_t pyfwi_gradient(pyfwi_plan *p, const pyfwi_shot *shots, int32_t nshots, const float *src,
                       const float *res, float *grad, int32_t mode, int32_t ckpt_every,
                       int32_t flags, int32_t mem)
{
    if (!p || !res || !grad) return fail(PYFWI_EINVAL, "null argument");
    if (int e = check_device_ptrs(mem, {{src, "source table"}, {res, "residuals"},
                                        {grad, "gradient"}}))
        return e;
    if (!p->model_set) return fail(PYFWI_ESTATE, "pyfwi_set_model must come first");
    if (mode != PYFWI_MODE_PYFWI && mode != PYFWI_MODE_EXACT) return fail(PYFWI_EINVAL, "unknown mode");
    if (int e = check_shots(p, shots, nshots)) return e;
    if (ckpt_every < 0) return fail(PYFWI_EINVAL, "ckpt_every must be >= 0 (0: automatic)");
    const DevGrid &g = p->dg;
    const int64_t fps = fields_per_shot(p);
    const size_t nrec = rec_extent(shots, nshots);
    bool same = p->have_forward && (int)p->last_shots.size() == nshots &&
                std::memcmp(p->last_shots.data(), shots, sizeof(pyfwi_shot) * nshots) == 0;
    if (mode == PYFWI_MODE_PYFWI && (!same || p->chp.empty()))
        return fail(PYFWI_ESTATE, "pyfwi-mode gradient must follow a forward of the same shots "
                                  "with a checkpoint list");
    const float *d_res = nullptr;
    if (int e = stage_in(p, p->res_stage, res, nrec, mem, &d_res, "residuals")) return e;
    const float *d_src = nullptr;
    if (mode == PYFWI_MODE_EXACT) {
        if (!src) return fail(PYFWI_EINVAL, "exact mode needs the source table");
        if (int e = stage_in(p, p->src_stage, src, src_extent(shots, nshots), mem, &d_src,
                             "source table"))
            return e;
        if (int e = ensure(p->seis_stage, nrec, p->stream, "records")) return e;
    }
    if (int e = upload_shots(p, shots, nshots)) return e;
    float *d_grad = grad;
    const size_t n = (size_t)g.nz * g.nx;
    if (mem != PYFWI_MEM_DEVICE) {
        if (int e = ensure(p->grad_stage, 6 * n, p->stream, "gradient output")) return e;
        d_grad = p->grad_stage.p;
    }
    int G = pass_size(p, nshots);
    const int nt_all = max_nt(std::vector<pyfwi_shot>(shots, shots + nshots));
    // exact mode after a forward that kept every step of these shots (the
    // reference's chpr = 100) with its sources on active cells: the backward
    // reads the kept states instead of recomputing segments from checkpoints
    const bool kept = mode == PYFWI_MODE_EXACT && same && p->snaps.p && p->src_inside &&
                      keeps_every_step(p->chp, nt_all) &&
                      p->snaps.n >= (size_t)nshots * p->chp.size() * fps;
    const bool reuse = mode == PYFWI_MODE_EXACT && !kept && same && p->ckpts_all &&
                       (ckpt_every == 0 || ckpt_every == p->ckpt_every);
    const int K = kept ? 1 : reuse ? p->ckpt_every : (ckpt_every > 0 ? ckpt_every : auto_interval(nt_all));
    p->last_ckpt_every = mode == PYFWI_MODE_EXACT ? (kept ? 0 : K) : 0;
    if (mode == PYFWI_MODE_EXACT && !kept && p->tune_group == 0) {
        // checkpoint schedule sized to the device memory: shots per pass from
        // what the pass's checkpoints (unless kept by the forward), segment
        // and working states take
        double budget = plan_budget(p) - (reuse ? (double)p->ckpts.n * sizeof(float) : 0.0);
        G = exact_pass_size(p, nshots, nt_all, K, !reuse, budget - (double)p->snaps.n * 4.0);
        if (G < 1) {
            release_snaps(p);
            G = exact_pass_size(p, nshots, nt_all, K, !reuse, budget);
        }
        if (G < 1)
            return fail(PYFWI_ENOMEM, "the exact adjoint of one shot (checkpoints every " +
                                          std::to_string(K) + " steps) does not fit in device "
                                          "memory");
        // buffers are only grown: give back those larger than this schedule's
        // so that the others can grow
        const size_t ncp_all = (size_t)((nt_all + K - 1) / K);
        HIPCHK(hipStreamSynchronize(p->stream));
        if (!reuse && p->ckpts.n > (size_t)G * ncp_all * fps) {
            p->ckpts.release();
            p->ckpts_all = false;
        }
        if (p->seg.n > (size_t)G * K * fps) p->seg.release();
    }
    p->last_passes.clear();
    identity_order(p, nshots);
    // exact mode sized to memory: the passes' shots spread evenly (at most G)
    const bool even = mode == PYFWI_MODE_EXACT && !kept && p->tune_group == 0;
    const int npass = (nshots + G - 1) / G;
    for (int g0 = 0, k = 0; g0 < nshots; g0 += G, ++k) {
        if (even) G = nshots / npass + (k < nshots % npass ? 1 : 0);
        const int ng = std::min(G, nshots - g0);
        p->last_passes.push_back(ng);
        const int nsets = adj_acc_sets(g.sdo, ng);
        if (int e = ensure(p->gacc, (size_t)nsets * 6 * g.fstride, p->stream, "gradient"))
            return e;
        launch_zero(p->gacc.p, (int64_t)nsets * 6 * g.fstride, p->stream);
        if (mode == PYFWI_MODE_PYFWI) {
            const float *sn = p->snaps.p + (int64_t)g0 * p->chp.size() * fps;
            if (int e = gradient_pyfwi(p, g0, ng, d_res, (flags & PYFWI_F_PRECOND) != 0, sn,
                                       p->chp))
                return e;
        } else if (kept) {
            const float *sn = p->snaps.p + (int64_t)g0 * p->chp.size() * fps;
            if (int e = gradient_exact_kept(p, g0, ng, d_src, d_res, sn, (int)p->chp.size()))
                return e;
        } else {
            std::vector<pyfwi_shot> shp(shots + g0, shots + g0 + ng);
            const int ncp = (max_nt(shp) + K - 1) / K;
            const float *ck;
            if (reuse) {
                ck = p->ckpts.p + (int64_t)g0 * p->ncp * fps;
                if (int e = gradient_exact(p, g0, ng, d_src, d_res, ck, p->ncp, K)) return e;
            } else {
                // forward of this pass again, keeping its checkpoints
                if (int e = ensure(p->ckpts, (size_t)ng * ncp * fps, p->stream, "checkpoints"))
                    return e;
                p->ckpts_all = false;
                std::vector<int32_t> none;
                if (int e = forward_pass(p, g0, ng, d_src, p->seis_stage.p, none, nullptr,
                                         p->ckpts.p, ncp, K))
                    return e;
                if (int e = gradient_exact(p, g0, ng, d_src, d_res, p->ckpts.p, ncp, K)) return e;
            }
        }
        launch_grad_sum(g, nsets, 6, p->gacc.p, d_grad, p->bu.p, mode == PYFWI_MODE_EXACT,
                        g0 > 0, p->stream);
    }
    HIPCHK(hipGetLastError());
    if (mem != PYFWI_MEM_DEVICE) {
        HIPCHK(hipMemcpyAsync(grad, d_grad, 6 * n * sizeof(float), hipMemcpyDeviceToHost, p->stream));
        HIPCHK(hipStreamSynchronize(p->stream));
    }
    if (int e = harvest_timing(p)) return e;
    return 0;
}

int32_t pyfwi_misfit_gradient(pyfwi_plan *p, const pyfwi_shot *shots, int32_t nshots,
                              const float *src, const float *obs, float *seis, float *grad,
                              double *misfit, const int32_t *chp, int32_t nchp,
                              int32_t comp_mask, int32_t flags, int32_t mem)
{
    if (!p || !src || !obs || !grad || !misfit || !chp || nchp <= 0)
        return fail(PYFWI_EINVAL, "null argument or empty checkpoint list");
    if (int e = check_device_ptrs(mem, {{src, "source table"}, {obs, "observed records"},
                                        {seis, "records"}, {grad, "gradient"}}))
        return e;
    if (!p->model_set) return fail(PYFWI_ESTATE, "pyfwi_set_model must come first");
    if (int e = check_shots(p, shots, nshots)) return e;
    if (comp_mask < 0 || comp_mask > 31) return fail(PYFWI_EINVAL, "comp_mask must be 0..31");
    std::vector<pyfwi_shot> sh(shots, shots + nshots);
    const int nt = max_nt(sh);
    std::vector<int32_t> cl(chp, chp + nchp);
    if (int e = check_chp_nt(sh, cl, nt)) return e;
    const bool every = keeps_every_step(cl, nt);
    const DevGrid &g = p->dg;
    const int64_t fps = fields_per_shot(p);
    const size_t nrec = rec_extent(shots, nshots), nsrc = src_extent(shots, nshots);
    const size_t n = (size_t)g.nz * g.nx;
    const float *d_src = nullptr, *d_obs = nullptr;
    if (int e = stage_in(p, p->src_stage, src, nsrc, mem, &d_src, "source table")) return e;
    if (int e = stage_in(p, p->res_stage, obs, nrec, mem, &d_obs, "observed records")) return e;
    float *d_seis = seis;
    if (mem != PYFWI_MEM_DEVICE || !seis) {
        if (int e = ensure(p->seis_stage, nrec, p->stream, "records")) return e;
        d_seis = p->seis_stage.p;
    }
    float *d_grad = grad;
    if (mem != PYFWI_MEM_DEVICE) {
        if (int e = ensure(p->grad_stage, 6 * n, p->stream, "gradient output")) return e;
        d_grad = p->grad_stage.p;
    }
    if (int e = ensure(p->resid, nrec, p->stream, "residuals")) return e;
    const bool filtered = p->fir_L >= 0;
    if (filtered) {
        for (DevBuf *b : {&p->filt_a, &p->filt_b})
            if (int e = ensure(*b, nrec, p->stream, "filtered records")) return e;
    }
    if (!p->d_misfit) HIPCHK(hipMalloc(&p->d_misfit, sizeof(double)));
    HIPCHK(hipMemsetAsync(p->d_misfit, 0, sizeof(double), p->stream));
    launch_zero(d_seis, (int64_t)nrec, p->stream);
    p->chp = cl;
    p->last_shots = sh;
    p->have_forward = false;  // kept states belong to the last pass only
    // the passes: runs of `order` (the caller's shot indices) of the sizes in
    // `sizes`, each run by decreasing nt
    std::vector<int32_t> order(nshots), sizes;
    for (int k = 0; k < nshots; ++k) order[k] = k;
    // every step kept: states are laid out time-major in each pass (a pass
    // holds the sum of its shots' nt), the shots longest first; else each
    // shot of a pass holds the nchp states of the checkpoint list
    if (every)
        std::stable_sort(order.begin(), order.end(),
                         [&](int32_t a, int32_t b) { return sh[a].nt > sh[b].nt; });
    auto states_of = [&](int32_t k) -> size_t { return every ? (size_t)sh[k].nt : (size_t)nchp; };
    size_t need = 0;
    for (int k = 0; k < nshots; ++k) need += states_of(k);
    const size_t first = states_of(order[0]);
    // kept-state buffer, in states: with a tuned pass size, the largest of
    // those passes; else the buffer held from an earlier call when it takes
    // the longest shot (a reallocation of this size costs seconds), or as
    // many states as 90 % of the free memory holds (the buffer held counts as
    // free), never more than all shots need
    size_t slots = 0;
    if (p->tune_group > 0) {
        for (int g0 = 0; g0 < nshots; g0 += p->tune_group) {
            size_t sum = 0;
            for (int k = g0; k < std::min(nshots, g0 + p->tune_group); ++k)
                sum += states_of(order[k]);
            slots = std::max(slots, sum);
            sizes.push_back(std::min(p->tune_group, nshots - g0));
        }
    } else {
        if (p->snaps.n / (size_t)fps >= first) {
            slots = std::min(need, p->snaps.n / (size_t)fps);
        } else {
            size_t free_b = 0, total_b = 0;
            HIPCHK(hipMemGetInfo(&free_b, &total_b));
            const double avail = 0.9 * ((double)free_b + (double)p->snaps.n * sizeof(float));
            slots = std::min(need, (size_t)(avail / ((double)fps * sizeof(float))));
            if (slots < first)
                return fail(PYFWI_ENOMEM, "the kept states of one shot (" +
                                              std::to_string((long long)(first * fps * sizeof(float))) +
                                              " bytes) do not fit in device memory");
        }
        // first fit by decreasing nt: each shot joins the first pass with
        // room for its states.  A launch costs a fixed part plus a part per
        // live shot, so the step count summed over the passes (one launch
        // pair per step of a pass's longest shot) is what the packing saves:
        // long shots share their passes with short ones instead of running
        // alone.  Each pass keeps its shots by decreasing nt.
        std::vector<std::vector<int32_t>> bins;
        std::vector<size_t> used;
        for (int32_t k : order) {
            size_t b = 0;
            while (b < bins.size() && used[b] + states_of(k) > slots) ++b;
            if (b == bins.size()) {
                bins.emplace_back();
                used.push_back(0);
            }
            bins[b].push_back(k);
            used[b] += states_of(k);
        }
        order.clear();
        for (auto &b : bins) {
            order.insert(order.end(), b.begin(), b.end());
            sizes.push_back((int32_t)b.size());
        }
    }
    std::vector<pyfwi_shot> shs(nshots);
    for (int k = 0; k < nshots; ++k) shs[k] = sh[order[k]];
    if (int e = upload_shots(p, shs.data(), nshots)) return e;
    if (int e = ensure(p->snaps, slots * fps, p->stream, "checkpoints")) return e;
    p->last_passes = sizes;
    p->last_order = order;
    p->last_ckpt_every = 0;
    for (int g0 = 0, b = 0; g0 < nshots; ++b) {
        const int ng = sizes[b];
        const int nt_pass = every ? shs[g0].nt : nt;
        std::vector<int32_t> cl_pass = cl;
        TimeMajor tm;
        if (every) {
            every_step(nt_pass, cl_pass);
            if (int e = time_major(shs.data() + g0, ng, tm)) return e;
        }
        const TimeMajor *tmp = every ? &tm : nullptr;
        if (int e = forward_pass(p, g0, ng, d_src, d_seis, cl_pass, p->snaps.p, nullptr, 0, 1,
                                 tmp))
            return e;
        if (filtered) {
            // the misfit of the lowpassed records, K d - K d_obs = K (d - d_obs)
            // (the filter is linear): the residual is filtered once, so the
            // float32 sums see the residual, not the records it is the small
            // difference of.  Its adjoint source is K K r (K is symmetric).
            const pyfwi_shot *dsp = p->d_shots + g0, *hsp = p->h_shots.data() + g0;
            launch_l2_residual(ng, dsp, d_seis, d_obs, p->filt_a.p, comp_mask, nullptr,
                               p->stream);
            launch_fir(ng, dsp, hsp, p->filt_a.p, p->filt_b.p, comp_mask, p->fir_taps.p,
                       p->fir_L, p->fir_pad, p->stream);
            launch_l2_residual(ng, dsp, p->filt_b.p, nullptr, p->filt_a.p, comp_mask,
                               p->d_misfit, p->stream);
            launch_zero(p->resid.p, (int64_t)nrec, p->stream);
            launch_fir(ng, dsp, hsp, p->filt_a.p, p->resid.p, comp_mask, p->fir_taps.p,
                       p->fir_L, p->fir_pad, p->stream);
        } else {
            launch_l2_residual(ng, p->d_shots + g0, d_seis, d_obs, p->resid.p, comp_mask,
                               p->d_misfit, p->stream);
        }
        const int nsets = every ? adj_acc_sets_upto(g.sdo, ng) : adj_acc_sets(g.sdo, ng);
        if (int e = ensure(p->gacc, (size_t)nsets * 6 * g.fstride, p->stream, "gradient"))
            return e;
        launch_zero(p->gacc.p, (int64_t)nsets * 6 * g.fstride, p->stream);
        if (int e = gradient_pyfwi(p, g0, ng, p->resid.p, (flags & PYFWI_F_PRECOND) != 0,
                                   p->snaps.p, cl_pass, tmp))
            return e;
        launch_grad_sum(g, nsets, 6, p->gacc.p, d_grad, p->bu.p, 0, g0 > 0, p->stream);
        g0 += ng;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(misfit, p->d_misfit, sizeof(double), hipMemcpyDeviceToHost, p->stream));
    if (mem != PYFWI_MEM_DEVICE) {
        HIPCHK(hipMemcpyAsync(grad, d_grad, 6 * n * sizeof(float), hipMemcpyDeviceToHost, p->stream));
        if (seis)
            HIPCHK(hipMemcpyAsync(seis, d_seis, nrec * sizeof(float), hipMemcpyDeviceToHost,
                                  p->stream));
    }
    HIPCHK(hipStreamSynchronize(p->stream));
    if (int e = harvest_timing(p)) return e;
    return 0;
}

int32_t pyfwi_forward_noalloc(pyfwi_plan *p, const pyfwi_shot *shots, int32_t nshots,
                              const float *src, float *seis)
{
    if (!p || !src || !seis) return fail(PYFWI_EINVAL, "null argument");
    if (int e = check_shots(p, shots, nshots)) return e;
    const int64_t fps = fields_per_shot(p);
    // scratch layout: the shot table (256-byte aligned), then the states
    const int64_t tab = ((int64_t)nshots * (int64_t)sizeof(pyfwi_shot) + 255) / 256 * 64;
    const int64_t need = tab + (int64_t)nshots * fps;
    if (need > (int64_t)PYFWI_SCRATCH_FLOATS)
        return fail(PYFWI_EINVAL, "pyfwi_forward_noalloc: " + std::to_string(nshots) +
                                      " shots of this grid need " + std::to_string(need * 4) +
                                      " bytes of scratch; the maximum is " +
                                      std::to_string((int64_t)PYFWI_SCRATCH_FLOATS * 4));
    if (!p->model_set) return fail(PYFWI_ESTATE, "pyfwi_set_model must come first");
    if (int e = check_device_ptrs(PYFWI_MEM_DEVICE, {{src, "source table"}, {seis, "records"}}))
        return e;
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev != p->device) return fail(PYFWI_EINVAL, "the plan's device is not current");
    float *base = scratch_on(dev);
    if (!base) return fail(PYFWI_EHIP, "hipGetSymbolAddress failed for the scratch array");
    pyfwi_shot *dsh = reinterpret_cast<pyfwi_shot *>(base);
    float *F = base + tab;
    shot_flags(p, shots, nshots);
    p->have_forward = false;  // no states are kept for a gradient
    HIPCHK(hipMemcpyAsync(dsh, shots, sizeof(pyfwi_shot) * nshots, hipMemcpyHostToDevice,
                          p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));  // the caller's table may go away
    const DevGrid &g = p->dg;
    launch_zero(seis, (int64_t)rec_extent(shots, nshots), p->stream);
    launch_zero(F, (int64_t)nshots * fps, p->stream);
    const int nt = max_nt(std::vector<pyfwi_shot>(shots, shots + nshots));
    for (int t = 0; t < nt; ++t) {
        // the two-launch step of pyfwi_forward (forward_step), untimed
        if (!p->src_inside || !p->rec_inline)
            launch_inject_record(g, nshots, dsh, t, F, fps, p->src_inside ? nullptr : src, seis,
                                 p->d_rtab, p->stream);
        const float *s_in = p->src_inside ? src : nullptr;
        launch_vel(g, +1, nshots, dsh, t, F, fps, F, fps, p->bu.p, p->dxp, p->dzp, p->stream,
                   s_in);
        launch_str(g, +1, nshots, dsh, t, F, fps, F, fps, p->lam.p, p->mu.p, p->dxp, p->dzp,
                   p->stream, s_in, p->src_inside && p->rec_inline ? seis : nullptr);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

int32_t pyfwi_set_record_filter(pyfwi_plan *p, const double *taps, int32_t half_len,
                                int32_t pad)
{
    if (!p) return fail(PYFWI_EINVAL, "null plan");
    if (half_len < 0) {
        p->fir_L = -1;
        return 0;
    }
    if (!taps || pad < 0 || half_len > PYFWI_FIR_MAX_HALF)
        return fail(PYFWI_EINVAL, "record filter: taps needed, pad >= 0, half length <= " +
                                      std::to_string(PYFWI_FIR_MAX_HALF));
    std::vector<float> t(taps, taps + half_len + 1);
    for (float v : t)
        if (!std::isfinite(v)) return fail(PYFWI_ENONFINITE, "record filter holds NaN or Inf");
    HIPCHK(hipStreamSynchronize(p->stream));  // earlier calls may still read the old taps
    if (int e = ensure(p->fir_taps, t.size(), p->stream, "record filter")) return e;
    HIPCHK(hipMemcpy(p->fir_taps.p, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
    p->fir_L = half_len;
    p->fir_pad = pad;
    return 0;
}

int32_t pyfwi_plan_set_graphs(pyfwi_plan *p, int32_t enable)
{
    if (!p) return fail(PYFWI_EINVAL, "null plan");
    if (enable != 0 && enable != 1) return fail(PYFWI_EINVAL, "enable must be 0 or 1");
    if (!enable && !p->graph_cache.empty()) {
        HIPCHK(hipStreamSynchronize(p->stream));
        for (auto &ge : p->graph_cache) (void)hipGraphExecDestroy(ge.second);
        p->graph_cache.clear();
    }
    p->graphs = enable;
    return 0;
}

int32_t pyfwi_graph_stats(pyfwi_plan *p, int64_t *captures, int64_t *replays)
{
    if (!p || !captures || !replays) return fail(PYFWI_EINVAL, "null argument");
    *captures = p->graph_captures;
    *replays = p->graph_replays;
    return 0;
}

int32_t pyfwi_plan_set_tuning(pyfwi_plan *p, int32_t shots_per_pass, int32_t fused,
                              int32_t strip_rows)
{
    if (!p) return fail(PYFWI_EINVAL, "null plan");
    if (shots_per_pass < 0 || strip_rows < 0 || (strip_rows > 0 && strip_rows < 2 * p->grid.sdo))
        return fail(PYFWI_EINVAL, "shots_per_pass and strip_rows must be >= 0 "
                                  "(strip_rows 0 or >= 2*sdo)");
    p->tune_group = shots_per_pass;
    p->tune_fused = (fused & PYFWI_TUNE_FUSED_STEPS) ? 1 : 0;
    p->adj_fused = (fused & PYFWI_TUNE_ADJ_FUSED) ? 1 : 0;
    p->fwd_two = (fused & PYFWI_TUNE_FWD_TWO) ? 1 : 0;
    p->tune_rows = strip_rows;
    p->dg.rows_override = strip_rows;
    p->have_forward = false;
    return 0;
}

int32_t pyfwi_get_snapshot(pyfwi_plan *p, int32_t shot, int32_t slot, float *out, int32_t mem)
{
    if (!p || !out) return fail(PYFWI_EINVAL, "null argument");
    if (int e = check_device_ptrs(mem, {{out, "snapshot output"}})) return e;
    if (!p->have_forward) return fail(PYFWI_ESTATE, "no forward run yet");
    const int nchp = (int)p->chp.size();
    if (shot < 0 || shot >= (int)p->last_shots.size() || slot < 0 || slot >= nchp)
        return fail(PYFWI_EINVAL, "snapshot index out of range");
    const DevGrid &g = p->dg;
    const int64_t fps = fields_per_shot(p);
    const size_t n = (size_t)g.nz * g.nx;
    float *d_out = out;
    if (mem != PYFWI_MEM_DEVICE) {
        if (int e = ensure(p->grad_stage, 6 * n, p->stream, "snapshot staging")) return e;
        d_out = p->grad_stage.p;
    }
    launch_unpad(g, p->snaps.p + ((int64_t)shot * nchp + slot) * fps, d_out, 5, p->stream);
    HIPCHK(hipGetLastError());
    if (mem != PYFWI_MEM_DEVICE) {
        HIPCHK(hipMemcpyAsync(out, d_out, 5 * n * sizeof(float), hipMemcpyDeviceToHost, p->stream));
        HIPCHK(hipStreamSynchronize(p->stream));
    }
    return 0;
}

int32_t pyfwi_last_passes(pyfwi_plan *p, int32_t *sizes, int32_t cap, int32_t *count,
                          int32_t *ckpt_every)
{
    if (!p || !count || cap < 0 || (cap > 0 && !sizes)) return fail(PYFWI_EINVAL, "null argument");
    if (ckpt_every) *ckpt_every = p->last_ckpt_every;
    *count = (int32_t)p->last_passes.size();
    for (int k = 0; k < cap && k < *count; ++k) sizes[k] = p->last_passes[k];
    return 0;
}

int32_t pyfwi_last_order(pyfwi_plan *p, int32_t *order, int32_t cap, int32_t *count)
{
    if (!p || !count || cap < 0 || (cap > 0 && !order)) return fail(PYFWI_EINVAL, "null argument");
    *count = (int32_t)p->last_order.size();
    for (int k = 0; k < cap && k < *count; ++k) order[k] = p->last_order[(size_t)k];
    return 0;
}

int32_t pyfwi_synchronize(pyfwi_plan *p)
{
    if (!p) return fail(PYFWI_EINVAL, "null plan");
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

int32_t pyfwi_plan_set_timing(pyfwi_plan *p, int32_t enable)
{
    if (!p) return fail(PYFWI_EINVAL, "null plan");
    if (int e = harvest_timing(p)) return e;
    if (enable < 0 || enable > 2) return fail(PYFWI_EINVAL, "timing mode must be 0, 1 or 2");
    p->timing = enable;
    for (int k = 0; k < PYFWI_K_COUNT; ++k) {
        p->kt_ms[k] = 0;
        p->kt_n[k] = 0;
    }
    return 0;
}

int32_t pyfwi_kernel_timing(pyfwi_plan *p, int32_t kind, double *total_ms, int64_t *launches)
{
    if (!p || !total_ms || !launches) return fail(PYFWI_EINVAL, "null argument");
    if (kind < 0 || kind >= PYFWI_K_COUNT) return fail(PYFWI_EINVAL, "unknown kernel kind");
    if (!p->timing) return fail(PYFWI_ESTATE, "timing is not enabled");
    if (int e = harvest_timing(p)) return e;
    *total_ms = p->kt_ms[kind];
    *launches = p->kt_n[kind];
    return 0;
}

} // extern "C"
