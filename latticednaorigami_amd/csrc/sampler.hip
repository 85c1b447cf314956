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
// With lanes_per_replica = 64 every replica has a wavefront of its own, which
// keeps replicas on separate CUs when there are few of them; its 64 lanes run
// the chain together on identical values (any lane's store is every lane's
// store) and split the independent reads of the lane-parallel sections
// (endpoint scans, trial-site lookups, the random domain draw) between them.
// Its block is held in that CU's LDS when it fits.  With 1, 64 replicas
// share a wavefront, one lane each.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ldo.h"
#include "../../include/ldo_mlp.h"
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

// The replica a workgroup of the one-replica-per-workgroup sweeps runs.
// Workgroups are dealt to the 8 XCDs round-robin, so workgroup b runs on XCD
// b % 8; when the ladder divides into 8 equal runs of slots, XCD x takes the
// replicas of slots [x R/8, (x+1) R/8), so that each XCD's L2 holds the
// energy tables of few temperatures (slots are temperature-major).  q_to_rep
// (slot -> replica) is the exchange's permutation, or null for the identity.
__device__ inline int replica_of_block(const Params& P, const int32_t* q_to_rep) {
    int b = blockIdx.x, R = P.n_replicas;
    int q = (R % 8 == 0) ? (b % 8) * (R / 8) + b / 8 : b;
    return q_to_rep != nullptr ? q_to_rep[q] : q;
}

// With a wavefront per replica, its 64 lanes run the replica's chain
// together (lane-parallel sections spread independent reads over them); the
// engine object sits in LDS, so its section pointers are not re-read from
// the lanes' stacks in device memory.
extern "C" __global__ void ldo_sweep_kernel(Params P, int64_t steps, const int32_t* rep_slot,
                                            const int32_t* q_to_rep) {
    __shared__ Eng lds_engine;
    if (P.lanes_per_replica == 1) {
        bool active;
        int r = replica_of_thread(P, active);
        if (active) {
            sweep_replica(P, r, steps, rep_slot);
        }
        return;
    }
    if ((int)blockIdx.x < P.n_replicas) {
        int r = replica_of_block(P, q_to_rep);
        sweep_replica_on(lds_engine, P, r, P.blocks + (size_t)r * P.L.bytes, steps, rep_slot, 1);
    }
}

// One workgroup of 64 lanes per replica.  Every access of the Markov chain
// to its replica state (the block, the engine's section pointers and the
// launch parameters) goes to LDS instead of device memory or scratch.  The
// dynamic LDS holds the block (P.L.bytes, a multiple of 256); the host picks
// this kernel only when that fits.
extern "C" __global__ void __launch_bounds__(64)
    ldo_sweep_lds_kernel(Params P, int64_t steps, const int32_t* rep_slot,
                         const int32_t* q_to_rep) {
    extern __shared__ uint4 lds_block[];
    __shared__ Params lds_params;
    __shared__ Eng lds_engine;
    if ((int)blockIdx.x >= P.n_replicas) {
        return;
    }
    int r = replica_of_block(P, q_to_rep);
    size_t n16 = (size_t)P.L.bytes / 16;
    uint4* g = (uint4*)(P.blocks + (size_t)r * P.L.bytes);
    for (size_t i = threadIdx.x; i < n16; i += blockDim.x) {
        lds_block[i] = g[i];
    }
    if (threadIdx.x == 0) {
        lds_params = P;
    }
    __syncthreads();
    sweep_replica_on(lds_engine, lds_params, r, (char*)lds_block, steps, rep_slot, 1);
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

extern "C" __global__ void ldo_exchange_kernel(Params P, XState X, int64_t swap_i,
                                               int64_t rng_ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (P.exchange_mode == LDO_EXCHANGE_WINDOWS) {
            exchange_windows_round(P, X, swap_i, rng_ctr);
        }
        else {
            exchange_round(P, X, swap_i, rng_ctr);
        }
    }
}

// one site query on a copy of its replica's block (ldo_site_eval)
extern "C" __global__ void ldo_site_eval_kernel(Params P, char* blk, ldo_site_query q,
                                                ldo_site_result* out, int32_t* found) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ldo_site_result r;
        *found = site_eval_on(P, blk, q, r) ? 1 : 0;
        *out = r;
    }
}

extern "C" __global__ void ldo_status_kernel(Params P, int32_t* out) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < P.n_replicas) {
        out[r] = replica_status(P, r);
    }
}

// `attempts` learned domain-regrowth attempts per replica, one workgroup of
// 64 lanes per replica: lane 0 runs the Markov chain, all lanes build the
// input box and evaluate the network between barriers.  Each replica acts
// with the control variables of its last sweep.
extern "C" __global__ void __launch_bounds__(64)
    ldo_learned_kernel(Params P, int64_t attempts) {
    __shared__ Params sp;
    __shared__ Eng e;
    __shared__ LearnedMove m;
    __shared__ int go;
    int r = blockIdx.x, lane = threadIdx.x;
    if (r >= P.n_replicas) {
        return;
    }
    if (lane == 0) {
        sp = P;
        e.attach(&sp, sp.blocks + (size_t)r * sp.L.bytes, r);
        e.select_tables();
    }
    __syncthreads();
    for (int64_t it = 0; it < attempts; it++) {
        if (lane == 0) {
            go = e.S->status == ST_OK ? e.learned_begin(m) : -1;
        }
        __syncthreads();
        int g = go;
        if (g < 0) {
            break;
        }
        // an attempt asks for up to three network evaluations (a domain
        // regrowth one, an exchange one per growth path and placement)
        while (g == 1) {
            for (int ph = 0; ph < LDO_LEARNED_PHASES; ph++) {
                learned_phase_lane(e, sp, m, ph, lane, 64);
                __syncthreads();
            }
            if (lane == 0) {
                go = e.learned_continue(m);
            }
            __syncthreads();
            g = go;
        }
        __syncthreads();
    }
}

// Training samples of the learned proposal from the current configurations:
// for every staple and both of its ends, the input box around the end's
// chain neighbour (the end left out), the legal outputs and the end's
// actual site and orientation.  Replica r writes from sample r * cap on and
// its count to counts[r].
extern "C" __global__ void __launch_bounds__(64)
    ldo_learned_samples_kernel(Params P, int32_t cap, uint8_t* ch_out, int32_t* id_out,
                               uint8_t* legal_out, int32_t* tgt_out, int32_t* counts) {
    __shared__ Params sp;
    __shared__ Eng e;
    __shared__ LearnedMove m;
    __shared__ int n_st;
    int r = blockIdx.x, lane = threadIdx.x;
    if (r >= P.n_replicas) {
        return;
    }
    if (lane == 0) {
        sp = P;
        e.attach(&sp, sp.blocks + (size_t)r * sp.L.bytes, r);
        n_st = e.S->status == ST_OK ? e.S->num_staples : 0;
    }
    __syncthreads();
    int n = 0;
    for (int w = 1; w <= n_st; w++) {
        for (int end = 0; end < 2 && n < cap; end++) {
            if (lane == 0) {
                int cs = e.order[w];
                int len = e.ch[cs].len;
                int d = end == 0 ? e.slot(cs, 0) : e.slot(cs, len - 1);
                int a = end == 0 ? e.slot(cs, 1) : e.slot(cs, len - 2);
                int idv = e.ident_of(d);
                m.d = d;
                m.a = a;
                m.id = (idv < 0 ? -idv : idv) - 1;
                I3 ap = e.pos(a);
                m.ax = ap.x;
                m.ay = ap.y;
                m.az = ap.z;
                m.k_old = 6 * unit_index(sub3(e.pos(d), ap)) + e.dom[d].ore;
            }
            __syncthreads();
            learned_context_lane(e, m, lane, 64);
            __syncthreads();
            size_t o = (size_t)r * cap + n;
            for (int s = lane; s < 343; s += 64) {
                ch_out[o * 343 + s] = m.ch[s];
            }
            if (lane < 36) {
                legal_out[o * 36 + lane] = m.legal[lane / 6];
            }
            if (lane == 0) {
                id_out[o] = m.id;
                tgt_out[o] = m.k_old;
            }
            n++;
            __syncthreads();
        }
    }
    if (lane == 0) {
        counts[r] = n;
    }
}


// ---------------------------------------------------------------- batches
// Chain k of a batch is replica 0 of its own system: plist[k] carries that
// system's layout, tables and block pointer (P.blocks = the chain's block).
struct BatchCfg {
    int64_t off;       // first int of the chain's packed configuration
    int32_t n_chains;
    int32_t n_domains;
};

extern "C" __global__ void ldo_batch_init_kernel(const Params* plist, int32_t n,
                                                 const int32_t* cfg, const BatchCfg* meta) {
    int k = blockIdx.x;
    if (k >= n || threadIdx.x != 0) {
        return;
    }
    const BatchCfg& m = meta[k];
    const int32_t* h = cfg + m.off;
    int nc = m.n_chains, nd = m.n_domains;
    init_replica(plist[k], 0, h, h + nc, h + 2 * nc, h + 2 * nc + 3 * nd, nc, 0, nullptr);
}

// One workgroup per chain with the chain's block, its parameters and the
// engine object in LDS (dynamic LDS = the largest block of the batch).
extern "C" __global__ void __launch_bounds__(64)
    ldo_batch_sweep_lds_kernel(const Params* plist, int32_t n, int64_t steps) {
    extern __shared__ uint4 lds_block[];
    __shared__ Params lds_params;
    __shared__ Eng lds_engine;
    int k = blockIdx.x;
    if (k >= n) {
        return;
    }
    if (threadIdx.x == 0) {
        lds_params = plist[k];
    }
    __syncthreads();
    size_t n16 = (size_t)lds_params.L.bytes / 16;
    uint4* g = (uint4*)lds_params.blocks;
    for (size_t i = threadIdx.x; i < n16; i += blockDim.x) {
        lds_block[i] = g[i];
    }
    __syncthreads();
    sweep_replica_on(lds_engine, lds_params, 0, (char*)lds_block, steps, nullptr, 1);
    __syncthreads();
    for (size_t i = threadIdx.x; i < n16; i += blockDim.x) {
        g[i] = lds_block[i];
    }
}

// Blocks in device memory: mode 0, one chain per wavefront run by all its
// lanes; 1, one chain per lane; 2, one chain per wavefront run by lane 0.
// With a chain per wavefront its parameters and engine object sit in LDS, so
// their fields are not re-read from the lanes' stacks in device memory.
extern "C" __global__ void ldo_batch_sweep_kernel(const Params* plist, int32_t n, int64_t steps,
                                                  int32_t mode) {
    __shared__ Params lds_params;
    __shared__ Eng lds_engine;
    int k;
    if (mode == 1) {
        k = blockIdx.x * blockDim.x + threadIdx.x;
        if (k < n) {
            sweep_replica(plist[k], 0, steps, nullptr);
        }
        return;
    }
    k = blockIdx.x;
    if (k >= n) {
        return;
    }
    if (threadIdx.x == 0) {
        lds_params = plist[k];
    }
    __syncthreads();
    if (mode == 0) {
        sweep_replica_on(lds_engine, lds_params, 0, lds_params.blocks, steps, nullptr, 1);
    }
    else if (threadIdx.x == 0) {
        sweep_replica_on(lds_engine, lds_params, 0, lds_params.blocks, steps, nullptr, 0);
    }
}

extern "C" __global__ void ldo_batch_scalars_kernel(const Params* plist, int32_t n,
                                                    int32_t dependent, RepScalars* out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) {
        return;
    }
    if (dependent) {
        dependent_replica(plist[k], 0);
    }
    out[k] = *(const RepScalars*)plist[k].blocks;
}

// Staging of ldo_set_config_noalloc: the packed configuration (chain
// indices | identities | positions | orientations) of the latest call.  One
// array per GPU, held until the process ends; host code reaches it through
// the address hipGetSymbolAddress returns on that device, never by name.
static __device__ __align__(256) int32_t scratch[LDO_NOALLOC_CONFIG_INTS];

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
    int64_t swap_counter = 0; // exchange rounds so far (the exchange stream's counter)
    int64_t swap_base = 0;    // swap_counter when the pairing index last restarted
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
    int32_t* d_xint = nullptr;      // q_to_rep | rep_slot | status | slot_key
    double* d_xdbl = nullptr;       // ladder_t | ladder_u | ladder_s | reduced_u | rep_u
    int64_t* d_xcnt = nullptr;      // exchange attempts, accepts, per-pair counts [4R]
    int n_classes = 1;              // table key classes of the ladder
    char* d_site_blk = nullptr;     // block copy of a site evaluation
    ldo_site_result* d_site_res = nullptr;
    int32_t* d_cfg = nullptr;
    size_t cfg_cap = 0;
    int32_t* scratch = nullptr;       // this device's `scratch` (ldo_set_config_noalloc)
    ldo_window* d_windows = nullptr;
    double* d_grid = nullptr;
    double* d_win_stored = nullptr;
    int64_t* d_hist = nullptr;        // window order-parameter visits
    std::vector<ldo_window> windows; // host copy
    bool lds_sweep = false;           // sweep with the replica blocks in LDS
    float* d_net = nullptr;           // learned proposal parameters
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
    std::vector<int32_t> key;
    c.n_classes = ldo::slot_classes(c.plan, key);
    upload(c, c.d_xint + 3 * (size_t)n, key.data(), key.size() * sizeof(int32_t), "copy keys");
    c.P.slot_key = c.d_xint + 3 * (size_t)n;
    c.P.pt_dim2 = c.plan.dim2;
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

// The initiation-term bookkeeping of a replica keeps one bit per table key
// class of the ladder (Eng::begin_interval).
void check_classes(const Ctx& c) {
    if (c.n_classes > 64) {
        throw AbiError {LDO_E_ARG, "a ladder holds at most 64 (temperature, stacking) pairs"};
    }
}

void launch_sweep(Ctx& c, int64_t steps, const int32_t* rep_slot) {
    // with a ladder, workgroups follow the slot order (XCD-aware placement)
    const int32_t* q_to_rep = rep_slot != nullptr ? c.d_xint : nullptr;
    launch_timed(c, 0, [&]() {
        if (c.lds_sweep) {
            hipLaunchKernelGGL(ldo_sweep_lds_kernel, dim3(c.R), dim3(64), (size_t)c.P.L.bytes,
                               c.stream, c.P, steps, rep_slot, q_to_rep);
        }
        else {
            hipLaunchKernelGGL(ldo_sweep_kernel, grid_of(c), dim3(64), 0, c.stream, c.P, steps,
                               rep_slot, q_to_rep);
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

// Device address of `scratch` on the current device, looked up once per
// device (at ldo_create, with that device current).
std::mutex g_scratch_mu;
int32_t* g_scratch_ptr[64] = {};

int32_t* scratch_address(int dev) {
    if (dev < 0 || dev >= 64) {
        throw AbiError {LDO_E_HIP, "device index out of range for the no-alloc staging"};
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress scratch");
        g_scratch_ptr[dev] = static_cast<int32_t*>(p);
    }
    return g_scratch_ptr[dev];
}

// Status word of one replica (no host allocation).
void check_replica_status(Ctx& c, int replica) {
    hipLaunchKernelGGL(ldo_status_kernel, dim3((c.R + 255) / 256), dim3(256), 0, c.stream, c.P,
                       c.d_xint + 2 * c.R);
    hip_check(hipGetLastError(), "launch status");
    int32_t st = 0;
    download(c, &st, c.d_xint + 2 * c.R + replica, sizeof st, "copy status");
    ldo::check_status(st, replica);
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
                    c->d_windows, c->d_grid, c->d_win_stored, c->d_net, c->d_hist,
                    c->d_site_blk, c->d_site_res};
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

size_t hist_bins(const Ctx& c) {
    return (size_t)c.R * (c.P.grid_len + 1);
}

// (Re)allocates the window histogram for the current grid, zeroed.
void reset_histogram(Ctx& c, bool realloc) {
    if (realloc) {
        release(c, c.d_hist, "hipFree histogram");
        hip_check(hipMalloc(&c.d_hist, hist_bins(c) * sizeof(int64_t)), "hipMalloc histogram");
    }
    std::vector<int64_t> zero(hist_bins(c), 0);
    upload(c, c.d_hist, zero.data(), zero.size() * sizeof(int64_t), "clear histogram");
    c.P.hist = c.d_hist;
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
        int dev = 0;
        hip_check(hipGetDevice(&dev), "get device");
        c->scratch = scratch_address(dev);
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
        hip_check(hipMalloc(&c->d_xint, 4 * (size_t)n_replicas * sizeof(int32_t)), "hipMalloc x");
        std::vector<int32_t> xi(4 * n_replicas, 0);
        for (int r = 0; r < n_replicas; r++) {
            xi[r] = r;
            xi[n_replicas + r] = r;
        }
        upload(*c, c->d_xint, xi.data(), xi.size() * sizeof(int32_t), "copy x");
        upload_tables(*c);
        size_t xcnt_bytes = (2 + 4 * (size_t)n_replicas) * sizeof(int64_t);
        hip_check(hipMalloc(&c->d_xcnt, xcnt_bytes), "hipMalloc xcnt");
        zero(*c, c->d_xcnt, xcnt_bytes, "clear xcnt");
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

// The configuration goes into this device's `scratch` array instead of a
// buffer of the context; no device memory is allocated or freed.
int ldo_set_config_noalloc(void* ctx, int32_t replica, const ldo_config* cfg) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (replica < 0 || replica >= c->R) {
            throw AbiError {LDO_E_ARG, "replica out of range"};
        }
        int nd = ldo::config_domains(c->plan, cfg);
        int64_t n = cfg->n_chains;
        if (2 * n + 6 * (int64_t)nd > LDO_NOALLOC_CONFIG_INTS) {
            throw AbiError {LDO_E_ARG, "configuration larger than the no-alloc staging "
                                       "(2 per chain + 6 per domain > LDO_NOALLOC_CONFIG_INTS)"};
        }
        int32_t* cf = c->scratch;
        upload(*c, cf, cfg->chain_index, n * sizeof(int32_t), "stage chain indices");
        upload(*c, cf + n, cfg->chain_identity, n * sizeof(int32_t), "stage identities");
        upload(*c, cf + 2 * n, cfg->positions, 3 * (size_t)nd * sizeof(int32_t),
               "stage positions");
        upload(*c, cf + 2 * n + 3 * nd, cfg->orientations, 3 * (size_t)nd * sizeof(int32_t),
               "stage orientations");
        hipLaunchKernelGGL(ldo_init_kernel, grid_of(*c), dim3(64), 0, c->stream, c->P, cf, cf + n,
                           cf + 2 * n, cf + 2 * n + 3 * nd, (int32_t)n, replica, 1,
                           (const int32_t*)nullptr);
        hip_check(hipGetLastError(), "launch set_config");
        check_replica_status(*c, replica);
    });
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
        if (!temps || !u || !b || !s) {
            throw AbiError {LDO_E_ARG, "null ladder array"};
        }
        c->plan.ladder_t.assign(temps, temps + n);
        c->plan.ladder_u.assign(u, u + n);
        c->plan.ladder_b.assign(b, b + n);
        c->plan.ladder_s.assign(s, s + n);
        c->plan.dim2 = 1;
        hip_check(hipStreamSynchronize(c->stream), "sync");
        upload_tables(*c);
    });
}

int ldo_set_ladder_2d(void* ctx, int32_t n_temps, const double* temps, int32_t n_stacking,
                      const double* stacking_mults, double staple_u_mult) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        ldo::set_ladder_2d(c->plan, n_temps, temps, n_stacking, stacking_mults, staple_u_mult);
        hip_check(hipStreamSynchronize(c->stream), "sync");
        upload_tables(*c);
    });
}

int ldo_set_step(void* ctx, int64_t step) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (step < 0) {
            throw AbiError {LDO_E_ARG, "negative step"};
        }
        for (int r = 0; r < c->R; r++) {
            upload(*c, c->d_blocks + (size_t)r * c->P.L.bytes + offsetof(RepScalars, step), &step,
                   sizeof step, "copy step");
        }
    });
}

int ldo_get_exchange_counts(void* ctx, int32_t n, int64_t* attempts, int64_t* accepts) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (n != 2 * c->R || !attempts || !accepts) {
            throw AbiError {LDO_E_ARG, "exchange counts hold 2 values per replica"};
        }
        download(*c, attempts, c->d_xcnt + 2, (size_t)n * sizeof(int64_t), "copy pair counts");
        download(*c, accepts, c->d_xcnt + 2 + n, (size_t)n * sizeof(int64_t), "copy pair counts");
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
        if (c->d_hist) {
            reset_histogram(*c, true);
        }
    });
}

int ldo_set_window_grid(void* ctx, const double* grid) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (!c->P.windows) {
            throw AbiError {LDO_E_ARG, "no umbrella windows set"};
        }
        if (!grid && c->P.grid_len > 0) {
            throw AbiError {LDO_E_ARG, "grid values required"};
        }
        upload(*c, c->d_grid, grid, (size_t)c->R * c->P.grid_len * sizeof(double), "copy grid");
    });
}

int ldo_set_histogram(void* ctx, int32_t on) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (!on) {
            c->P.hist = nullptr;
            release(*c, c->d_hist, "hipFree histogram");
            return;
        }
        if (!c->P.windows) {
            throw AbiError {LDO_E_ARG, "the histogram counts a window's order parameter: set windows first"};
        }
        reset_histogram(*c, true);
    });
}

int ldo_get_histogram(void* ctx, int64_t* counts, int32_t reset) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (!c->d_hist) {
            throw AbiError {LDO_E_ARG, "histogram not enabled"};
        }
        if (!counts) {
            throw AbiError {LDO_E_ARG, "output required"};
        }
        download(*c, counts, c->d_hist, hist_bins(*c) * sizeof(int64_t), "copy histogram");
        if (reset) {
            reset_histogram(*c, false);
        }
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

int ldo_learned_set_network(void* ctx, int32_t n_ident, const float* params, double fallback) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (!params || n_ident < 1) {
            throw AbiError {LDO_E_ARG, "parameters and n_ident >= 1 required"};
        }
        ldo::check_learned_types(c->plan);
        ldo_mlp_layout L;
        ldo::mlp_layout(n_ident, L);
        if (!(fallback >= 0.0 && fallback <= 1.0)) {
            throw AbiError {LDO_E_ARG, "fallback probability outside [0, 1]"};
        }
        if (c->P.n_moves > LDO_MAX_MOVES - 3) {
            throw AbiError {LDO_E_ARG, "the learned moves need three free counter slots"};
        }
        if (c->P.rng_mode != 0) {
            throw AbiError {LDO_E_ARG, "the learned move draws from the Philox streams"};
        }
        if (n_ident < c->P.L.ns) {
            throw AbiError {LDO_E_ARG, "n_ident must cover every scaffold domain identity"};
        }
        release(*c, c->d_net, "hipFree network");
        hip_check(hipMalloc(&c->d_net, L.total * sizeof(float)), "hipMalloc network");
        upload(*c, c->d_net, params, L.total * sizeof(float), "copy network");
        c->P.net = c->d_net;
        int64_t offs[10] = {L.w1, L.emb, L.w1e, L.b1, L.w2, L.b2, L.w3, L.b3, L.w4, L.b4};
        for (int k = 0; k < 10; k++) {
            c->P.net_off[k] = offs[k];
        }
        c->P.net_n_ident = n_ident;
        c->P.learned_slot = LDO_MAX_MOVES - 1;
        c->P.learned_fallback = fallback;
    });
}

int ldo_learned_sweep(void* ctx, int64_t attempts) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (attempts < 0) {
            throw AbiError {LDO_E_ARG, "negative attempt count"};
        }
        if (!c->P.net) {
            throw AbiError {LDO_E_MISUSE, "no network (ldo_learned_set_network)"};
        }
        hipLaunchKernelGGL(ldo_learned_kernel, dim3(c->R), dim3(64), 0, c->stream, c->P,
                           attempts);
        hip_check(hipGetLastError(), "launch learned");
        check_statuses(*c);
    });
}

int ldo_learned_samples(void* ctx, int32_t cap, uint8_t* channels, int32_t* ident,
                        uint8_t* legal, int32_t* target, int32_t* n_out) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (cap < 0 || !channels || !ident || !legal || !target || !n_out) {
            throw AbiError {LDO_E_ARG, "buffers and cap >= 0 required"};
        }
        ldo::check_learned_types(c->plan);
        int per = 2 * c->P.max_total_staples;
        size_t n = (size_t)c->R * per;
        uint8_t *d_ch = nullptr, *d_lg = nullptr;
        int32_t *d_id = nullptr, *d_tg = nullptr, *d_cnt = nullptr;
        auto cleanup = [&]() {
            for (void* b: {(void*)d_ch, (void*)d_lg, (void*)d_id, (void*)d_tg, (void*)d_cnt}) {
                if (b) {
                    (void)hipFree(b);
                }
            }
        };
        try {
            hip_check(hipMalloc(&d_ch, std::max<size_t>(1, n * 343)), "hipMalloc samples");
            hip_check(hipMalloc(&d_lg, std::max<size_t>(1, n * 36)), "hipMalloc samples");
            hip_check(hipMalloc(&d_id, std::max<size_t>(1, n) * 4), "hipMalloc samples");
            hip_check(hipMalloc(&d_tg, std::max<size_t>(1, n) * 4), "hipMalloc samples");
            hip_check(hipMalloc(&d_cnt, (size_t)c->R * 4), "hipMalloc samples");
            hipLaunchKernelGGL(ldo_learned_samples_kernel, dim3(c->R), dim3(64), 0, c->stream,
                               c->P, per, d_ch, d_id, d_lg, d_tg, d_cnt);
            hip_check(hipGetLastError(), "launch samples");
            std::vector<int32_t> cnt(c->R), id(n), tg(n);
            std::vector<uint8_t> ch(n * 343), lg(n * 36);
            download(*c, cnt.data(), d_cnt, cnt.size() * 4, "copy counts");
            download(*c, ch.data(), d_ch, ch.size(), "copy samples");
            download(*c, lg.data(), d_lg, lg.size(), "copy samples");
            download(*c, id.data(), d_id, id.size() * 4, "copy samples");
            download(*c, tg.data(), d_tg, tg.size() * 4, "copy samples");
            int32_t k = 0;
            for (int r = 0; r < c->R; r++) {
                for (int i = 0; i < cnt[r] && k < cap; i++, k++) {
                    size_t o = (size_t)r * per + i;
                    std::memcpy(channels + (size_t)k * 343, &ch[o * 343], 343);
                    std::memcpy(legal + (size_t)k * 36, &lg[o * 36], 36);
                    ident[k] = id[o];
                    target[k] = tg[o];
                }
            }
            *n_out = k;
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
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
        check_classes(*c);
        XState X;
        X.q_to_rep = c->d_xint;
        X.rep_slot = c->d_xint + c->R;
        X.ladder_t = c->d_xdbl;
        X.ladder_u = c->d_xdbl + c->R;
        X.ladder_s = c->d_xdbl + 2 * c->R;
        X.reduced_u = c->d_xdbl + 3 * c->R;
        X.rep_u = c->d_xdbl + 3 * c->R + c->plan.reduced_u.size();
        X.counts = c->d_xcnt;
        for (int64_t s = 0; s < n_swaps; s++) {
            int64_t rng_ctr = ++c->swap_counter;
            int64_t swap_i = rng_ctr - c->swap_base;
            launch_sweep(*c, interval, X.rep_slot);
            launch_timed(*c, 1, [&]() {
                hipLaunchKernelGGL(ldo_exchange_kernel, dim3(1), dim3(64), 0, c->stream, c->P, X,
                                   swap_i, rng_ctr);
            });
            hip_check(hipGetLastError(), "launch exchange");
        }
        check_statuses(*c);
    });
}

int ldo_exchange(void* ctx, int64_t swap_index) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (swap_index < 0) {
            throw AbiError {LDO_E_ARG, "negative swap index"};
        }
        check_classes(*c);
        XState X;
        X.q_to_rep = c->d_xint;
        X.rep_slot = c->d_xint + c->R;
        X.ladder_t = c->d_xdbl;
        X.ladder_u = c->d_xdbl + c->R;
        X.ladder_s = c->d_xdbl + 2 * c->R;
        X.reduced_u = c->d_xdbl + 3 * c->R;
        X.rep_u = c->d_xdbl + 3 * c->R + c->plan.reduced_u.size();
        X.counts = c->d_xcnt;
        int64_t rng_ctr = ++c->swap_counter;
        // update_control_qs and update_dependent_qs: an interval of 0 steps
        launch_sweep(*c, 0, X.rep_slot);
        launch_timed(*c, 1, [&]() {
            hipLaunchKernelGGL(ldo_exchange_kernel, dim3(1), dim3(64), 0, c->stream, c->P, X,
                               swap_index, rng_ctr);
        });
        hip_check(hipGetLastError(), "launch exchange");
        check_statuses(*c);
    });
}

int ldo_site_eval(void* ctx, const ldo_site_query* q, int32_t n, ldo_site_result* res) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (n < 0 || (n > 0 && (!q || !res))) {
            throw AbiError {LDO_E_ARG, "null queries or results"};
        }
        for (int32_t k = 0; k < n; k++) {
            int a = std::abs(q[k].ore[0]) + std::abs(q[k].ore[1]) + std::abs(q[k].ore[2]);
            if (q[k].replica < 0 || q[k].replica >= c->R || a != 1) {
                throw AbiError {LDO_E_ARG, "query " + std::to_string(k) +
                                               ": replica out of range or ore not a unit vector"};
            }
        }
        if (n == 0) {
            return;
        }
        if (c->d_site_blk == nullptr) {
            hip_check(hipMalloc(&c->d_site_blk, c->P.L.bytes), "hipMalloc site block");
            hip_check(hipMalloc(&c->d_site_res, sizeof(ldo_site_result) + sizeof(int32_t)),
                      "hipMalloc site result");
        }
        int32_t* d_found = (int32_t*)(c->d_site_res + 1);
        for (int32_t k = 0; k < n; k++) {
            hip_check(hipMemcpyAsync(c->d_site_blk,
                                     c->d_blocks + (size_t)q[k].replica * c->P.L.bytes,
                                     c->P.L.bytes, hipMemcpyDeviceToDevice, c->stream),
                      "copy block");
            hipLaunchKernelGGL(ldo_site_eval_kernel, dim3(1), dim3(64), 0, c->stream, c->P,
                               c->d_site_blk, q[k], c->d_site_res, d_found);
            hip_check(hipGetLastError(), "launch site eval");
            int32_t found = 0;
            download(*c, &res[k], c->d_site_res, sizeof(ldo_site_result), "copy site result");
            download(*c, &found, d_found, sizeof found, "copy site flag");
            if (!found) {
                throw AbiError {LDO_E_ARG, "query " + std::to_string(k) +
                                               ": no such chain or domain in the replica"};
            }
        }
    });
}

int ldo_restart_swaps(void* ctx) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() { c->swap_base = c->swap_counter; });
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

int ldo_get_move_tracking(void* ctx, int32_t replica, int64_t* out, int64_t cap,
                          int64_t* n_out) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (replica < 0 || replica >= c->R || !out || !n_out) {
            throw AbiError {LDO_E_ARG, "replica out of range or null buffer"};
        }
        int64_t n = c->P.L.track_len;
        if (cap < n) {
            throw AbiError {LDO_E_ARG, "tracking capacity too small"};
        }
        download(*c, out, c->d_blocks + (size_t)replica * c->P.L.bytes + c->P.L.off_track,
                 (size_t)n * sizeof(int64_t), "copy tracking");
        *n_out = n;
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

int ldo_set_permutation(void* ctx, const int32_t* q_to_rep) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (!q_to_rep) {
            throw AbiError {LDO_E_ARG, "permutation required"};
        }
        std::vector<int32_t> both(2 * c->R, -1);
        for (int q = 0; q < c->R; q++) {
            int r = q_to_rep[q];
            if (r < 0 || r >= c->R || both[c->R + r] >= 0) {
                throw AbiError {LDO_E_ARG, "not a permutation of the replicas"};
            }
            both[q] = r;
            both[c->R + r] = q;
        }
        upload(*c, c->d_xint, both.data(), both.size() * sizeof(int32_t), "copy permutation");
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

// ------------------------------------------------------------ batch host
namespace {

struct BatchCtx {
    int n = 0;
    std::vector<ldo::Plan> plans;
    std::vector<Params> params;
    hipStream_t stream = nullptr;
    std::string err;
    char* d_blocks = nullptr;
    double* d_tab = nullptr;
    double* d_init = nullptr;
    int32_t* d_tlen = nullptr;
    int16_t* d_tident = nullptr;
    uint64_t* d_tapes = nullptr;
    int64_t* d_tmeta = nullptr;
    Params* d_params = nullptr;
    int32_t* d_cfg = nullptr;
    BatchCfg* d_meta = nullptr;
    RepScalars* d_scalars = nullptr;
    size_t max_block = 0;
    bool lds_ok = false;
    int kernel = LDO_BATCH_WAVE;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

std::string g_batch_create_error;

template <typename F>
int bguarded(BatchCtx* b, F f) {
    std::string& err = b ? b->err : g_batch_create_error;
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

template <typename T>
void bupload(BatchCtx& b, T*& dst, const std::vector<T>& src, const char* what) {
    size_t bytes = std::max<size_t>(1, src.size()) * sizeof(T);
    hip_check(hipMalloc(&dst, bytes), what);
    if (!src.empty()) {
        hip_check(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice,
                                 b.stream),
                  what);
        hip_check(hipStreamSynchronize(b.stream), what);
    }
}

void bdownload(BatchCtx& b, void* dst, const void* src, size_t bytes, const char* what) {
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, b.stream), what);
    hip_check(hipStreamSynchronize(b.stream), what);
}

template <typename T>
void bfree(T*& p) {
    if (p) {
        (void)hipFree(p);
        p = nullptr;
    }
}

void free_batch(BatchCtx* b) {
    if (b->stream) {
        (void)hipStreamSynchronize(b->stream);
    }
    bfree(b->d_blocks);
    bfree(b->d_tab);
    bfree(b->d_init);
    bfree(b->d_tlen);
    bfree(b->d_tident);
    bfree(b->d_tapes);
    bfree(b->d_tmeta);
    bfree(b->d_params);
    bfree(b->d_cfg);
    bfree(b->d_meta);
    bfree(b->d_scalars);
    if (b->ev0) {
        (void)hipEventDestroy(b->ev0);
    }
    if (b->ev1) {
        (void)hipEventDestroy(b->ev1);
    }
    if (b->stream) {
        (void)hipStreamDestroy(b->stream);
    }
    delete b;
}

void batch_scalars(BatchCtx& b, bool dependent, std::vector<RepScalars>& out) {
    hipLaunchKernelGGL(ldo_batch_scalars_kernel, dim3((b.n + 63) / 64), dim3(64), 0, b.stream,
                       b.d_params, b.n, (int32_t)dependent, b.d_scalars);
    hip_check(hipGetLastError(), "launch scalars");
    out.resize(b.n);
    bdownload(b, out.data(), b.d_scalars, (size_t)b.n * sizeof(RepScalars), "copy scalars");
}

void batch_check(BatchCtx& b) {
    std::vector<RepScalars> s;
    batch_scalars(b, false, s);
    for (int k = 0; k < b.n; k++) {
        ldo::check_status(s[k].status, k);
    }
}

// Launches `steps` steps on every chain with the batch's sweep kernel and
// returns the kernel time.
double batch_launch(BatchCtx& b, int kernel, int64_t steps) {
    hip_check(hipEventRecord(b.ev0, b.stream), "event record");
    if (kernel == LDO_BATCH_LDS) {
        hipLaunchKernelGGL(ldo_batch_sweep_lds_kernel, dim3(b.n), dim3(64), b.max_block, b.stream,
                           b.d_params, b.n, steps);
    }
    else if (kernel == LDO_BATCH_WAVE || kernel == LDO_BATCH_WAVE_SOLO) {
        hipLaunchKernelGGL(ldo_batch_sweep_kernel, dim3(b.n), dim3(64), 0, b.stream, b.d_params,
                           b.n, steps, kernel == LDO_BATCH_WAVE ? 0 : 2);
    }
    else {
        hipLaunchKernelGGL(ldo_batch_sweep_kernel, dim3((b.n + 63) / 64), dim3(64), 0, b.stream,
                           b.d_params, b.n, steps, 1);
    }
    hip_check(hipGetLastError(), "launch batch sweep");
    hip_check(hipEventRecord(b.ev1, b.stream), "event record");
    hip_check(hipEventSynchronize(b.ev1), "sync batch sweep");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, b.ev0, b.ev1), "event time");
    batch_check(b);
    return ms;
}

} // namespace

extern "C" {

int ldo_batch_create(int32_t n_chains, const ldo_system_desc* systems, const ldo_run_params* runs,
                     const ldo_config* init, const double* temps, void** out_batch) {
    BatchCtx* b = new BatchCtx;
    int rc = bguarded(nullptr, [&]() {
        if (n_chains < 1 || !systems || !runs || !init || !temps || !out_batch) {
            throw AbiError {LDO_E_ARG, "null argument or no chains"};
        }
        b->n = n_chains;
        b->plans.resize(n_chains);
        std::vector<double> tab, init_e;
        std::vector<int32_t> tlen, cfg;
        std::vector<int16_t> tident;
        std::vector<size_t> tab_off(n_chains), tlen_off(n_chains), tident_off(n_chains);
        std::vector<size_t> block_off(n_chains);
        std::vector<BatchCfg> meta(n_chains);
        size_t blocks = 0;
        for (int k = 0; k < n_chains; k++) {
            if (runs[k].rng_mode != LDO_RNG_PHILOX || runs[k].exchange_mode != LDO_EXCHANGE_NONE) {
                throw AbiError {LDO_E_ARG, "batch chains use Philox streams and no exchange"};
            }
            ldo::Plan& pl = b->plans[k];
            try {
                ldo::make_plan(&systems[k], &runs[k], 1, &temps[k], pl);
            } catch (AbiError& e) {
                throw AbiError {e.code, "chain " + std::to_string(k) + ": " + e.msg};
            }
            std::vector<double> t1, i1;
            ldo::slot_tables(pl, t1, i1);
            tab_off[k] = tab.size();
            tab.insert(tab.end(), t1.begin(), t1.end());
            init_e.insert(init_e.end(), i1.begin(), i1.end());
            tlen_off[k] = tlen.size();
            tlen.insert(tlen.end(), pl.type_len.begin(), pl.type_len.end());
            tident_off[k] = tident.size();
            tident.insert(tident.end(), pl.type_ident.begin(), pl.type_ident.end());
            block_off[k] = blocks;
            blocks += (size_t)pl.P.L.bytes;
            b->max_block = std::max(b->max_block, (size_t)pl.P.L.bytes);
            std::vector<int32_t> h;
            int nd = 0;
            try {
                ldo::pack_config(pl, &init[k], h, nd);
            } catch (AbiError& e) {
                throw AbiError {e.code, "chain " + std::to_string(k) + ": " + e.msg};
            }
            meta[k] = BatchCfg {(int64_t)cfg.size(), init[k].n_chains, nd};
            cfg.insert(cfg.end(), h.begin(), h.end());
        }
        hip_check(hipStreamCreate(&b->stream), "stream");
        hip_check(hipEventCreate(&b->ev0), "event create");
        hip_check(hipEventCreate(&b->ev1), "event create");
        bupload(*b, b->d_tab, tab, "copy tables");
        bupload(*b, b->d_init, init_e, "copy init");
        bupload(*b, b->d_tlen, tlen, "copy type_len");
        bupload(*b, b->d_tident, tident, "copy type_ident");
        bupload(*b, b->d_tapes, std::vector<uint64_t>(1, 0), "copy tapes");
        bupload(*b, b->d_tmeta, std::vector<int64_t>(2, 0), "copy tape meta");
        bupload(*b, b->d_cfg, cfg, "copy configurations");
        bupload(*b, b->d_meta, meta, "copy configuration index");
        hip_check(hipMalloc(&b->d_blocks, blocks), "hipMalloc blocks");
        hip_check(hipMemsetAsync(b->d_blocks, 0, blocks, b->stream), "clear blocks");
        hip_check(hipMalloc(&b->d_scalars, (size_t)n_chains * sizeof(RepScalars)),
                  "hipMalloc scalars");
        b->params.resize(n_chains);
        for (int k = 0; k < n_chains; k++) {
            const ldo::Plan& pl = b->plans[k];
            Params P = pl.P;
            size_t n2 = (size_t)pl.spec.n_ident * pl.spec.n_ident;
            double* t0 = b->d_tab + tab_off[k];
            P.hyb_e = t0;
            P.hyb_h = t0 + n2;
            P.hyb_s = t0 + 2 * n2;
            P.stack = t0 + 3 * n2;
            P.init_e = b->d_init + 3 * (size_t)k;
            P.init_h = b->d_init + 3 * (size_t)k + 1;
            P.init_s = b->d_init + 3 * (size_t)k + 2;
            P.type_len = b->d_tlen + tlen_off[k];
            P.type_ident = b->d_tident + tident_off[k];
            P.tapes = b->d_tapes;
            P.tape_off = b->d_tmeta;
            P.tape_len = b->d_tmeta + 1;
            P.blocks = b->d_blocks + block_off[k];
            P.windows = nullptr;
            P.win_grid = nullptr;
            P.win_stored = nullptr;
            b->params[k] = P;
        }
        bupload(*b, b->d_params, b->params, "copy parameters");
        hipLaunchKernelGGL(ldo_batch_init_kernel, dim3(n_chains), dim3(64), 0, b->stream,
                           b->d_params, n_chains, b->d_cfg, b->d_meta);
        hip_check(hipGetLastError(), "launch batch init");
        batch_check(*b);
        // LDS residency when the largest block fits a workgroup's LDS
        int dev = 0, optin = 0;
        hip_check(hipGetDevice(&dev), "get device");
        hip_check(hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev),
                  "shared memory opt-in limit");
        hipFuncAttributes fa;
        hip_check(hipFuncGetAttributes(&fa, (const void*)ldo_batch_sweep_lds_kernel),
                  "kernel attributes");
        if (optin > 0 && b->max_block + fa.sharedSizeBytes <= (size_t)optin) {
            hip_check(hipFuncSetAttribute((const void*)ldo_batch_sweep_lds_kernel,
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)b->max_block),
                      "dynamic LDS size");
            b->lds_ok = true;
            b->kernel = LDO_BATCH_LDS;
        }
    });
    if (rc != LDO_OK) {
        free_batch(b);
        return rc;
    }
    *out_batch = b;
    return LDO_OK;
}

void ldo_batch_destroy(void* batch) {
    if (batch) {
        free_batch(static_cast<BatchCtx*>(batch));
    }
}

const char* ldo_batch_last_error(void* batch) {
    return batch ? static_cast<BatchCtx*>(batch)->err.c_str() : g_batch_create_error.c_str();
}

int ldo_batch_set_kernel(void* batch, int32_t kernel) {
    BatchCtx* b = static_cast<BatchCtx*>(batch);
    return bguarded(b, [&]() {
        if (kernel < LDO_BATCH_LDS || kernel > LDO_BATCH_WAVE_SOLO) {
            throw AbiError {LDO_E_ARG, "unknown batch kernel"};
        }
        if (kernel == LDO_BATCH_LDS && !b->lds_ok) {
            throw AbiError {LDO_E_ARG, "the largest chain block does not fit in LDS"};
        }
        b->kernel = kernel;
    });
}

int ldo_batch_tune(void* batch, int64_t steps_per_trial, int32_t* chosen, double* ms) {
    BatchCtx* b = static_cast<BatchCtx*>(batch);
    return bguarded(b, [&]() {
        if (steps_per_trial < 1) {
            throw AbiError {LDO_E_ARG, "steps_per_trial must be positive"};
        }
        double best = -1;
        for (int k = LDO_BATCH_LDS; k <= LDO_BATCH_WAVE_SOLO; k++) {
            double t = -1;
            if (k != LDO_BATCH_LDS || b->lds_ok) {
                t = batch_launch(*b, k, steps_per_trial);
                if (best < 0 || t < best) {
                    best = t;
                    b->kernel = k;
                }
            }
            if (ms) {
                ms[k] = t;
            }
        }
        if (chosen) {
            *chosen = b->kernel;
        }
    });
}

int ldo_batch_sweep(void* batch, int64_t steps, double* kernel_ms) {
    BatchCtx* b = static_cast<BatchCtx*>(batch);
    return bguarded(b, [&]() {
        if (steps < 0) {
            throw AbiError {LDO_E_ARG, "negative step count"};
        }
        double t = batch_launch(*b, b->kernel, steps);
        if (kernel_ms) {
            *kernel_ms = t;
        }
    });
}

int ldo_batch_get_observables(void* batch, int32_t first, int32_t n, ldo_obs* out) {
    BatchCtx* b = static_cast<BatchCtx*>(batch);
    return bguarded(b, [&]() {
        if (first < 0 || n < 0 || first + n > b->n || (n > 0 && !out)) {
            throw AbiError {LDO_E_ARG, "chain range out of bounds"};
        }
        std::vector<RepScalars> s;
        batch_scalars(*b, true, s);
        for (int i = 0; i < n; i++) {
            ldo::fill_observables(b->plans[first + i], s[first + i], &out[i]);
            out[i].bias = 0;
        }
    });
}

int ldo_batch_get_counters(void* batch, ldo_counters* out) {
    BatchCtx* b = static_cast<BatchCtx*>(batch);
    return bguarded(b, [&]() {
        std::memset(out, 0, sizeof *out);
        std::vector<RepScalars> s;
        batch_scalars(*b, false, s);
        for (int k = 0; k < b->n; k++) {
            ldo::add_counters(s[k], out);
        }
    });
}

int ldo_batch_get_config(void* batch, int32_t chain, ldo_config* cfg, int32_t cap_chains,
                         int32_t cap_domains) {
    BatchCtx* b = static_cast<BatchCtx*>(batch);
    return bguarded(b, [&]() {
        if (chain < 0 || chain >= b->n) {
            throw AbiError {LDO_E_ARG, "chain out of range"};
        }
        const Params& P = b->params[chain];
        std::vector<char> blk(P.L.bytes);
        bdownload(*b, blk.data(), P.blocks, P.L.bytes, "copy block");
        ldo::read_config(b->plans[chain], blk.data(), cfg, cap_chains, cap_domains);
    });
}

} // extern "C"
