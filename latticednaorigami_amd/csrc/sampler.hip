// sampler.hip - kernels and C ABI of libldo_hip.so (gfx950).
//
// Kernels
//   ldo_init_kernel      build / replace a replica's configuration
//                        (OrigamiSystem constructor, set_config)
//   ldo_sweep_kernel     persistent MC sweep: each replica runs `steps`
//                        steps of GCMCSimulation::simulate; in PTMC mode it
//                        first switches to its ladder slot (update_control_qs)
//                        and ends with its dependent quantities
//   ldo_sweep_lds_kernel the same sweep with the replica held in LDS: one
//                        workgroup per replica copies its block, the launch
//                        parameters and the engine object into LDS, runs the
//                        chain there and writes the block back
//   ldo_exchange_kernel  one PTMC exchange round over even/odd ladder pairs
//                        (OneDPTGCMCSimulation::attempt_exchange)
// A replica is driven by one lane.  With lanes_per_replica = 64 every
// replica has a wavefront of its own (lane 0 works), which keeps replicas on
// separate CUs when there are few of them, and its block is held in that
// CU's LDS when it fits; with 1, 64 replicas share a wavefront.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ldo.h"
#include "layout.h"
#include "plan.h"
#include "replica_ops.h"

// ------------------------------------------------------------------ kernels
__device__ __forceinline__ int replica_of_thread(const Params& P, bool& active) {
    int r;
    if (P.lanes_per_replica == 1) {
        r = blockIdx.x * blockDim.x + threadIdx.x;
        active = r < P.n_replicas;
    }
    else {
        r = blockIdx.x;
        active = threadIdx.x == 0 && r < P.n_replicas;
    }
    return r;
}

extern "C" __global__ void ldo_init_kernel(Params P, const int32_t* c_idx, const int32_t* c_ident,
                                           const int32_t* c_pos, const int32_t* c_ore,
                                           int32_t n_chains, int32_t only_rep, int32_t mode,
                                           const int32_t* rep_slot) {
    bool active;
    int r = replica_of_thread(P, active);
    if (!active || (only_rep >= 0 && r != only_rep)) {
        return;
    }
    init_replica(P, r, c_idx, c_ident, c_pos, c_ore, n_chains, mode, rep_slot);
}

extern "C" __global__ void ldo_sweep_kernel(Params P, int64_t steps, const int32_t* rep_slot) {
    bool active;
    int r = replica_of_thread(P, active);
    if (active) {
        sweep_replica(P, r, steps, rep_slot);
    }
}

// One workgroup of 64 lanes per replica.  Every access of the Markov chain
// to its replica state (the block, the engine's section pointers and the
// launch parameters) goes to LDS instead of device memory or scratch.  The
// dynamic LDS holds the block (P.L.bytes, a multiple of 256); the host picks
// this kernel only when that fits.
extern "C" __global__ void __launch_bounds__(64)
    ldo_sweep_lds_kernel(Params P, int64_t steps, const int32_t* rep_slot) {
    extern __shared__ uint4 lds_block[];
    __shared__ Params lds_params;
    __shared__ Eng lds_engine;
    int r = blockIdx.x;
    if (r >= P.n_replicas) {
        return;
    }
    size_t n16 = (size_t)P.L.bytes / 16;
    uint4* g = (uint4*)(P.blocks + (size_t)r * P.L.bytes);
    for (size_t i = threadIdx.x; i < n16; i += blockDim.x) {
        lds_block[i] = g[i];
    }
    if (threadIdx.x == 0) {
        lds_params = P;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        sweep_replica_on(lds_engine, lds_params, r, (char*)lds_block, steps, rep_slot);
    }
    __syncthreads();
    for (size_t i = threadIdx.x; i < n16; i += blockDim.x) {
        g[i] = lds_block[i];
    }
}

extern "C" __global__ void ldo_dependent_kernel(Params P) {
    bool active;
    int r = replica_of_thread(P, active);
    if (active) {
        dependent_replica(P, r);
    }
}

extern "C" __global__ void ldo_exchange_kernel(Params P, XState X, int64_t swap_i) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (P.exchange_mode == LDO_EXCHANGE_WINDOWS) {
            exchange_windows_round(P, X, swap_i);
        }
        else {
            exchange_round(P, X, swap_i);
        }
    }
}

extern "C" __global__ void ldo_status_kernel(Params P, int32_t* out) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < P.n_replicas) {
        out[r] = replica_status(P, r);
    }
}

// ------------------------------------------------------------------- host
namespace {

using ldo::AbiError;

std::string g_create_error;

struct Ctx {
    ldo::Plan plan; // host-side description of the run
    Params P {};    // plan.P with the device pointers filled in
    int R = 0;
    std::vector<std::vector<uint64_t>> tapes;
    hipStream_t stream = nullptr;
    int64_t swap_counter = 0;
    std::string err;
    // kernel timing (ldo_set_profiling): event pairs awaiting readout and a
    // pool of spent events for reuse
    struct EvPair {
        hipEvent_t start, stop;
        int kind; // 0 sweep, 1 exchange
    };
    bool profiling = false;
    std::vector<EvPair> pending;
    std::vector<hipEvent_t> ev_pool;
    double prof_ms[2] = {0, 0};
    int64_t prof_n[2] = {0, 0};
    // device buffers
    char* d_blocks = nullptr;
    int32_t* d_type_len = nullptr;
    int16_t* d_type_ident = nullptr;
    double* d_tab = nullptr;        // hyb_e | hyb_h | hyb_s | stack
    double* d_init = nullptr;       // init_e | init_h | init_s
    uint64_t* d_tapes = nullptr;
    int64_t* d_tape_meta = nullptr; // off | len
    int32_t* d_xint = nullptr;      // q_to_rep | rep_slot | status
    double* d_xdbl = nullptr;       // ladder_t | ladder_u | ladder_s | reduced_u
    int64_t* d_xcnt = nullptr;      // exchange attempts, accepts
    int32_t* d_cfg = nullptr;
    size_t cfg_cap = 0;
    ldo_window* d_windows = nullptr;
    double* d_grid = nullptr;
    double* d_win_stored = nullptr;
    std::vector<ldo_window> windows; // host copy
    bool lds_sweep = false;           // sweep with the replica blocks in LDS
};

struct HipError {
    hipError_t e;
    const char* what;
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        throw HipError {e, what};
    }
}

// Every transfer and fill runs on the context's stream, in order with its
// kernels: work on the legacy null stream is not ordered with a stream that
// was created non-blocking, and a host-side return from hipMemset/hipMemcpy
// does not promise that the device has finished.  Uploads synchronize before
// returning so that the host staging buffer may be released.
void upload(Ctx& c, void* dst, const void* src, size_t bytes, const char* what) {
    if (bytes == 0) {
        return;
    }
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c.stream), what);
    hip_check(hipStreamSynchronize(c.stream), what);
}

void download(Ctx& c, void* dst, const void* src, size_t bytes, const char* what) {
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream), what);
    hip_check(hipStreamSynchronize(c.stream), what);
}

void zero(Ctx& c, void* dst, size_t bytes, const char* what) {
    hip_check(hipMemsetAsync(dst, 0, bytes, c.stream), what);
}

// Reads out every recorded event pair into the running totals.
void drain_profile(Ctx& c) {
    if (c.pending.empty()) {
        return;
    }
    hip_check(hipStreamSynchronize(c.stream), "sync profile");
    for (const Ctx::EvPair& p: c.pending) {
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, p.start, p.stop), "event time");
        c.prof_ms[p.kind] += ms;
        c.prof_n[p.kind] += 1;
        c.ev_pool.push_back(p.start);
        c.ev_pool.push_back(p.stop);
    }
    c.pending.clear();
}

hipEvent_t take_event(Ctx& c) {
    if (!c.ev_pool.empty()) {
        hipEvent_t e = c.ev_pool.back();
        c.ev_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "event create");
    return e;
}

// Launches through `launch`, bracketed by events on the sampler's stream
// when profiling is on.
template <typename F>
void launch_timed(Ctx& c, int kind, F launch) {
    if (!c.profiling) {
        launch();
        return;
    }
    hipEvent_t a = take_event(c), b = take_event(c);
    hip_check(hipEventRecord(a, c.stream), "event record");
    launch();
    hip_check(hipEventRecord(b, c.stream), "event record");
    c.pending.push_back({a, b, kind});
    if (c.pending.size() >= 4096) {
        drain_profile(c);
    }
}

template <typename T>
void release(Ctx& c, T*& p, const char* what) {
    if (p) {
        hip_check(hipStreamSynchronize(c.stream), what);
        hip_check(hipFree(p), what);
        p = nullptr;
    }
}

void upload_tables(Ctx& c) {
    int n = c.R;
    size_t n2 = (size_t)c.plan.spec.n_ident * c.plan.spec.n_ident;
    std::vector<double> tab, init;
    ldo::slot_tables(c.plan, tab, init);
    if (c.d_tab == nullptr) {
        hip_check(hipMalloc(&c.d_tab, tab.size() * sizeof(double)), "hipMalloc tables");
        hip_check(hipMalloc(&c.d_init, init.size() * sizeof(double)), "hipMalloc init");
    }
    upload(c, c.d_tab, tab.data(), tab.size() * sizeof(double), "copy tables");
    upload(c, c.d_init, init.data(), init.size() * sizeof(double), "copy init");
    c.P.hyb_e = c.d_tab;
    c.P.hyb_h = c.d_tab + (size_t)n * n2;
    c.P.hyb_s = c.d_tab + 2 * (size_t)n * n2;
    c.P.stack = c.d_tab + 3 * (size_t)n * n2;
    c.P.init_e = c.d_init;
    c.P.init_h = c.d_init + n;
    c.P.init_s = c.d_init + 2 * n;
    std::vector<double> xd = ldo::ladder_values(c.plan);
    if (c.d_xdbl == nullptr) {
        hip_check(hipMalloc(&c.d_xdbl, xd.size() * sizeof(double)), "hipMalloc ladder");
    }
    upload(c, c.d_xdbl, xd.data(), xd.size() * sizeof(double), "copy ladder");
}

dim3 grid_of(const Ctx& c) {
    if (c.P.lanes_per_replica == 1) {
        return dim3((c.R + 63) / 64);
    }
    return dim3(c.R);
}

// Whether each replica's block can live in its workgroup's LDS during a
// sweep: one wavefront per replica, and the block plus the kernel's static
// LDS within what the device grants a workgroup.
void choose_sweep_kernel(Ctx& c) {
    c.lds_sweep = false;
    const char* force = std::getenv("LDO_SWEEP_KERNEL");
    if (force != nullptr && std::strcmp(force, "global") == 0) {
        return;
    }
    if (c.P.lanes_per_replica != 64) {
        return;
    }
    int dev = 0, optin = 0;
    hip_check(hipGetDevice(&dev), "get device");
    hip_check(hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev),
              "shared memory opt-in limit");
    hipFuncAttributes fa;
    hip_check(hipFuncGetAttributes(&fa, (const void*)ldo_sweep_lds_kernel), "kernel attributes");
    size_t need = (size_t)c.P.L.bytes + fa.sharedSizeBytes;
    if (optin <= 0 || need > (size_t)optin) {
        return;
    }
    hip_check(hipFuncSetAttribute((const void*)ldo_sweep_lds_kernel,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)c.P.L.bytes),
              "dynamic LDS size");
    c.lds_sweep = true;
}

void launch_sweep(Ctx& c, int64_t steps, const int32_t* rep_slot) {
    launch_timed(c, 0, [&]() {
        if (c.lds_sweep) {
            hipLaunchKernelGGL(ldo_sweep_lds_kernel, dim3(c.R), dim3(64), (size_t)c.P.L.bytes,
                               c.stream, c.P, steps, rep_slot);
        }
        else {
            hipLaunchKernelGGL(ldo_sweep_kernel, grid_of(c), dim3(64), 0, c.stream, c.P, steps,
                               rep_slot);
        }
    });
    hip_check(hipGetLastError(), "launch sweep");
}

void upload_config(Ctx& c, const ldo_config* cfg, int& n_domains) {
    std::vector<int32_t> h;
    ldo::pack_config(c.plan, cfg, h, n_domains);
    if (h.size() > c.cfg_cap) {
        release(c, c.d_cfg, "hipFree cfg");
        hip_check(hipMalloc(&c.d_cfg, h.size() * sizeof(int32_t)), "hipMalloc cfg");
        c.cfg_cap = h.size();
    }
    upload(c, c.d_cfg, h.data(), h.size() * sizeof(int32_t), "copy cfg");
}

void check_statuses(Ctx& c) {
    hipLaunchKernelGGL(ldo_status_kernel, dim3((c.R + 255) / 256), dim3(256), 0, c.stream, c.P,
                       c.d_xint + 2 * c.R);
    hip_check(hipGetLastError(), "launch status");
    std::vector<int32_t> st(c.R);
    download(c, st.data(), c.d_xint + 2 * c.R, c.R * sizeof(int32_t), "copy status");
    for (int r = 0; r < c.R; r++) {
        ldo::check_status(st[r], r);
    }
}

void upload_tapes(Ctx& c) {
    std::vector<int64_t> meta(2 * c.R, 0);
    std::vector<uint64_t> all;
    for (int r = 0; r < c.R; r++) {
        meta[r] = (int64_t)all.size();
        meta[c.R + r] = (int64_t)c.tapes[r].size();
        all.insert(all.end(), c.tapes[r].begin(), c.tapes[r].end());
    }
    release(c, c.d_tapes, "hipFree tapes");
    hip_check(hipMalloc(&c.d_tapes, std::max<size_t>(1, all.size()) * sizeof(uint64_t)),
              "hipMalloc tapes");
    upload(c, c.d_tapes, all.data(), all.size() * sizeof(uint64_t), "copy tapes");
    upload(c, c.d_tape_meta, meta.data(), meta.size() * sizeof(int64_t), "copy tape meta");
    c.P.tapes = c.d_tapes;
    c.P.tape_off = c.d_tape_meta;
    c.P.tape_len = c.d_tape_meta + c.R;
}

template <typename F>
int guarded(Ctx* c, F f) {
    std::string& err = c ? c->err : g_create_error;
    try {
        f();
    } catch (AbiError& e) {
        err = e.msg;
        return e.code;
    } catch (HipError& e) {
        err = std::string(e.what) + ": " + hipGetErrorString(e.e);
        return LDO_E_HIP;
    } catch (std::exception& e) {
        err = e.what();
        return LDO_E_CONSTRAINT;
    }
    return LDO_OK;
}

void free_ctx(Ctx* c) {
    if (c->stream) {
        (void)hipStreamSynchronize(c->stream);
    }
    for (const Ctx::EvPair& p: c->pending) {
        c->ev_pool.push_back(p.start);
        c->ev_pool.push_back(p.stop);
    }
    for (hipEvent_t e: c->ev_pool) {
        (void)hipEventDestroy(e);
    }
    void* bufs[] = {c->d_blocks, c->d_type_len, c->d_type_ident, c->d_tab, c->d_init,
                    c->d_tapes, c->d_tape_meta, c->d_xint, c->d_xdbl, c->d_xcnt, c->d_cfg,
                    c->d_windows, c->d_grid, c->d_win_stored};
    for (void* b: bufs) {
        if (b) {
            (void)hipFree(b);
        }
    }
    if (c->stream) {
        (void)hipStreamDestroy(c->stream);
    }
    delete c;
}

} // namespace

extern "C" {

int ldo_create(const ldo_system_desc* sys, const ldo_run_params* run, const ldo_config* init,
               int32_t n_replicas, const double* temps, void** out_ctx) {
    Ctx* c = new Ctx;
    int rc = guarded(nullptr, [&]() {
        if (!init || !out_ctx) {
            throw AbiError {LDO_E_ARG, "null argument or no replicas"};
        }
        ldo::make_plan(sys, run, n_replicas, temps, c->plan);
        c->P = c->plan.P;
        c->R = n_replicas;
        Params& P = c->P;
        hip_check(hipStreamCreate(&c->stream), "stream");
        hip_check(hipMalloc(&c->d_type_len, c->plan.type_len.size() * sizeof(int32_t)),
                  "hipMalloc type_len");
        upload(*c, c->d_type_len, c->plan.type_len.data(),
               c->plan.type_len.size() * sizeof(int32_t), "copy type_len");
        P.type_len = c->d_type_len;
        hip_check(hipMalloc(&c->d_type_ident, c->plan.type_ident.size() * sizeof(int16_t)),
                  "hipMalloc type_ident");
        upload(*c, c->d_type_ident, c->plan.type_ident.data(),
               c->plan.type_ident.size() * sizeof(int16_t), "copy type_ident");
        P.type_ident = c->d_type_ident;
        upload_tables(*c);
        hip_check(hipMalloc(&c->d_xint, 3 * (size_t)n_replicas * sizeof(int32_t)), "hipMalloc x");
        std::vector<int32_t> xi(3 * n_replicas, 0);
        for (int r = 0; r < n_replicas; r++) {
            xi[r] = r;
            xi[n_replicas + r] = r;
        }
        upload(*c, c->d_xint, xi.data(), xi.size() * sizeof(int32_t), "copy x");
        hip_check(hipMalloc(&c->d_xcnt, 2 * sizeof(int64_t)), "hipMalloc xcnt");
        zero(*c, c->d_xcnt, 2 * sizeof(int64_t), "clear xcnt");
        hip_check(hipMalloc(&c->d_tape_meta, 2 * (size_t)n_replicas * sizeof(int64_t)),
                  "hipMalloc tape meta");
        c->tapes.assign(n_replicas, {});
        upload_tapes(*c);
        // replica blocks
        size_t total = (size_t)P.L.bytes * n_replicas;
        hip_check(hipMalloc(&c->d_blocks, total), "hipMalloc blocks");
        zero(*c, c->d_blocks, total, "clear blocks");
        P.blocks = c->d_blocks;
        int nd = 0;
        upload_config(*c, init, nd);
        int32_t* cf = c->d_cfg;
        hipLaunchKernelGGL(ldo_init_kernel, grid_of(*c), dim3(64), 0, c->stream, c->P, cf,
                           cf + init->n_chains, cf + 2 * init->n_chains,
                           cf + 2 * init->n_chains + 3 * nd, init->n_chains, -1, 0,
                           (const int32_t*)nullptr);
        hip_check(hipGetLastError(), "launch init");
        check_statuses(*c);
        choose_sweep_kernel(*c);
    });
    if (rc != LDO_OK) {
        free_ctx(c);
        return rc;
    }
    *out_ctx = c;
    return LDO_OK;
}

void ldo_destroy(void* ctx) {
    if (ctx) {
        free_ctx(static_cast<Ctx*>(ctx));
    }
}

const char* ldo_last_error(void* ctx) {
    if (!ctx) {
        return g_create_error.c_str();
    }
    return static_cast<Ctx*>(ctx)->err.c_str();
}

int ldo_set_config(void* ctx, int32_t replica, const ldo_config* cfg) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (replica < 0 || replica >= c->R) {
            throw AbiError {LDO_E_ARG, "replica out of range"};
        }
        int nd = 0;
        upload_config(*c, cfg, nd);
        int32_t* cf = c->d_cfg;
        hipLaunchKernelGGL(ldo_init_kernel, grid_of(*c), dim3(64), 0, c->stream, c->P, cf,
                           cf + cfg->n_chains, cf + 2 * cfg->n_chains,
                           cf + 2 * cfg->n_chains + 3 * nd, cfg->n_chains, replica, 1,
                           (const int32_t*)nullptr);
        hip_check(hipGetLastError(), "launch set_config");
        check_statuses(*c);
    });
}

int ldo_get_config(void* ctx, int32_t replica, ldo_config* cfg, int32_t cap_chains,
                   int32_t cap_domains) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (replica < 0 || replica >= c->R) {
            throw AbiError {LDO_E_ARG, "replica out of range"};
        }
        std::vector<char> blk(c->P.L.bytes);
        download(*c, blk.data(), c->d_blocks + (size_t)replica * c->P.L.bytes, c->P.L.bytes,
                 "copy block");
        ldo::read_config(c->plan, blk.data(), cfg, cap_chains, cap_domains);
    });
}

int ldo_set_ladder(void* ctx, int32_t n, const double* temps, const double* u, const double* b,
                   const double* s) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (n != c->R) {
            throw AbiError {LDO_E_ARG, "ladder size must equal the replica count"};
        }
        c->plan.ladder_t.assign(temps, temps + n);
        c->plan.ladder_u.assign(u, u + n);
        c->plan.ladder_b.assign(b, b + n);
        c->plan.ladder_s.assign(s, s + n);
        hip_check(hipStreamSynchronize(c->stream), "sync");
        upload_tables(*c);
    });
}

int ldo_set_windows(void* ctx, int32_t n, const ldo_window* windows, const double* grid,
                    int32_t grid_len) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        ldo::check_windows(c->plan, n, windows, grid_len);
        std::vector<int32_t> q2r(c->R);
        download(*c, q2r.data(), c->d_xint, c->R * sizeof(int32_t), "copy permutation");
        std::vector<RepScalars> scalars(c->R);
        for (int r = 0; r < c->R; r++) {
            download(*c, &scalars[r], c->d_blocks + (size_t)r * c->P.L.bytes, sizeof(RepScalars),
                     "copy scalars");
        }
        std::vector<double> stored = ldo::initial_window_state(windows, n, q2r, scalars);
        release(*c, c->d_windows, "hipFree windows");
        release(*c, c->d_grid, "hipFree grid");
        release(*c, c->d_win_stored, "hipFree window state");
        size_t ng = std::max<size_t>(1, (size_t)n * grid_len);
        hip_check(hipMalloc(&c->d_windows, n * sizeof(ldo_window)), "hipMalloc windows");
        hip_check(hipMalloc(&c->d_grid, ng * sizeof(double)), "hipMalloc grid");
        hip_check(hipMalloc(&c->d_win_stored, stored.size() * sizeof(double)),
                  "hipMalloc window state");
        upload(*c, c->d_windows, windows, n * sizeof(ldo_window), "copy windows");
        upload(*c, c->d_grid, grid, (size_t)n * grid_len * sizeof(double), "copy grid");
        upload(*c, c->d_win_stored, stored.data(), stored.size() * sizeof(double),
               "copy window state");
        c->windows.assign(windows, windows + n);
        c->P.windows = c->d_windows;
        c->P.win_grid = c->d_grid;
        c->P.grid_len = grid_len;
        c->P.win_stored = c->d_win_stored;
    });
}

int ldo_set_tape(void* ctx, int32_t replica, const uint64_t* tape, int64_t n) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (replica < 0 || replica >= c->R || n < 0) {
            throw AbiError {LDO_E_ARG, "replica out of range"};
        }
        c->tapes[replica].assign(tape, tape + n);
        hip_check(hipStreamSynchronize(c->stream), "sync");
        upload_tapes(*c);
        // restart the replica's tape cursor
        int64_t cursor = 0;
        upload(*c, c->d_blocks + (size_t)replica * c->P.L.bytes + offsetof(RepScalars, tape_pos),
               &cursor, sizeof cursor, "reset tape cursor");
    });
}

int ldo_sweep(void* ctx, int64_t steps) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (steps < 0) {
            throw AbiError {LDO_E_ARG, "negative step count"};
        }
        launch_sweep(*c, steps, nullptr);
        check_statuses(*c);
    });
}

int ldo_run(void* ctx, int64_t n_swaps, int64_t interval) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (n_swaps < 0 || interval < 0) {
            throw AbiError {LDO_E_ARG, "negative swap count or interval"};
        }
        XState X;
        X.q_to_rep = c->d_xint;
        X.rep_slot = c->d_xint + c->R;
        X.ladder_t = c->d_xdbl;
        X.ladder_u = c->d_xdbl + c->R;
        X.ladder_s = c->d_xdbl + 2 * c->R;
        X.reduced_u = c->d_xdbl + 3 * c->R;
        X.counts = c->d_xcnt;
        for (int64_t s = 0; s < n_swaps; s++) {
            int64_t swap_i = ++c->swap_counter;
            launch_sweep(*c, interval, X.rep_slot);
            launch_timed(*c, 1, [&]() {
                hipLaunchKernelGGL(ldo_exchange_kernel, dim3(1), dim3(64), 0, c->stream, c->P, X,
                                   swap_i);
            });
            hip_check(hipGetLastError(), "launch exchange");
        }
        check_statuses(*c);
    });
}

int ldo_get_observables(void* ctx, int32_t replica, ldo_obs* out) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (replica < 0 || replica >= c->R) {
            throw AbiError {LDO_E_ARG, "replica out of range"};
        }
        hipLaunchKernelGGL(ldo_dependent_kernel, grid_of(*c), dim3(64), 0, c->stream, c->P);
        hip_check(hipGetLastError(), "launch dependent");
        RepScalars S;
        download(*c, &S, c->d_blocks + (size_t)replica * c->P.L.bytes, sizeof S, "copy scalars");
        ldo::fill_observables(c->plan, S, out);
        out->bias = 0;
        if (c->P.windows != nullptr) {
            // the window the replica holds now: rep_slot, which a swap
            // updates at once (temp_slot follows at the next interval)
            int32_t q = 0;
            download(*c, &q, c->d_xint + c->R + replica, sizeof q, "copy slot");
            double st[2];
            download(*c, st, c->d_win_stored + 2 * q, sizeof st, "copy window state");
            out->bias = (st[0] + st[1]) * c->windows[q].bias_mult;
        }
    });
}

int ldo_get_counters(void* ctx, ldo_counters* out) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        std::memset(out, 0, sizeof *out);
        for (int r = 0; r < c->R; r++) {
            RepScalars S;
            download(*c, &S, c->d_blocks + (size_t)r * c->P.L.bytes, sizeof S, "copy scalars");
            ldo::add_counters(S, out);
        }
        int64_t x[2];
        download(*c, x, c->d_xcnt, sizeof x, "copy xcnt");
        out->exchange_attempts = x[0];
        out->exchange_accepts = x[1];
    });
}

int ldo_get_permutation(void* ctx, int32_t* q_to_rep) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        download(*c, q_to_rep, c->d_xint, c->R * sizeof(int32_t), "copy permutation");
    });
}

void* ldo_get_stream(void* ctx) { return ctx ? (void*)static_cast<Ctx*>(ctx)->stream : nullptr; }

const char* ldo_kernel_name(void* ctx) {
    const Ctx* c = static_cast<const Ctx*>(ctx);
    return (c != nullptr && c->lds_sweep) ? "ldo_sweep_lds_kernel" : "ldo_sweep_kernel";
}

int ldo_set_profiling(void* ctx, int32_t on) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        drain_profile(*c);
        c->profiling = on != 0;
        if (c->profiling) {
            c->prof_ms[0] = c->prof_ms[1] = 0;
            c->prof_n[0] = c->prof_n[1] = 0;
        }
    });
}

int ldo_get_profile(void* ctx, double* sweep_ms, int64_t* sweep_launches, double* exchange_ms,
                    int64_t* exchange_launches) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        drain_profile(*c);
        *sweep_ms = c->prof_ms[0];
        *sweep_launches = c->prof_n[0];
        *exchange_ms = c->prof_ms[1];
        *exchange_launches = c->prof_n[1];
    });
}

} // extern "C"
