// sampler.hip - kernels and C ABI of libldo_hip.so (gfx950).
//
// Kernels
//   ldo_init_kernel      build / replace a replica's configuration
//                        (OrigamiSystem constructor, set_config)
//   ldo_sweep_kernel     persistent MC sweep: each replica runs `steps`
//                        steps of GCMCSimulation::simulate; in PTMC mode it
//                        first switches to its ladder slot (update_control_qs)
//                        and ends with its dependent quantities
//   ldo_exchange_kernel  one PTMC exchange round over even/odd ladder pairs
//                        (OneDPTGCMCSimulation::attempt_exchange)
// A replica is driven by one lane.  With lanes_per_replica = 64 every
// replica has a wavefront of its own (lane 0 works), which keeps replicas on
// separate SIMDs when there are few of them; with 1, 64 replicas share a
// wavefront.
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
#include "engine.h"
#include "engine_fourbody.h"
#include "engine_moves.h"
#include "engine_system.h"
#include "layout.h"
#include "tables.h"

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
    Eng e;
    char* blk = P.blocks + (size_t)r * P.L.bytes;
    RepScalars* S = (RepScalars*)blk;
    if (mode == 0) {
        S->temp_slot = rep_slot ? rep_slot[r] : r;
        S->init_slot = S->temp_slot;
        S->visited = 0;
        S->status = ST_OK;
        e.attach(&P, blk, r);
        for (int i = 0; i < P.L.hcap; i++) {
            e.hkey[i] = LDO_HASH_EMPTY;
        }
        for (int c = 0; c < P.L.nchain; c++) {
            e.ch[c].ident = -1;
            e.ch[c].len = 0;
        }
        for (int t = 0; t < P.L.ntypes; t++) {
            e.tcount[t] = 0;
        }
        S->n_free = P.L.smax;
        for (int i = 0; i < P.L.smax; i++) {
            e.freel[i] = (int16_t)(P.L.smax - i);
        }
        // initialize_scaffold (origami_system.cpp:680-693)
        e.add_chain_with(0, c_idx[0]);
        S->num_staples--;
    }
    else {
        e.attach(&P, blk, r);
        // set_config (origami_system.cpp:327-341)
        for (int w = 0; w < S->n_chains; w++) {
            int cs = e.order[w];
            for (int d = 0; d < e.ch[cs].len; d++) {
                e.unassign_domain(e.slot(cs, d));
            }
        }
        for (int w = S->n_chains - 1; w != 0; w--) {
            e.delete_chain(e.order[w]);
        }
    }
    // initialize_staples (origami_system.cpp:659-678)
    int maxidx = c_idx[0];
    for (int c = 1; c < n_chains; c++) {
        e.add_chain_with(c_ident[c], c_idx[c]);
        maxidx = c_idx[c] > maxidx ? c_idx[c] : maxidx;
    }
    S->energy += (double)(S->n_chains - 1) * e.init_e;
    S->current_c_i = maxidx;
    // set_all_domains(config) (origami_system.cpp:588-616)
    int k = 0;
    for (int w = 0; w < S->n_chains; w++) {
        int cs = e.order[w];
        for (int d = 0; d < e.ch[cs].len; d++, k++) {
            int s = e.slot(cs, d);
            e.dom[s].x = c_pos[3 * k];
            e.dom[s].y = c_pos[3 * k + 1];
            e.dom[s].z = c_pos[3 * k + 2];
            I3 o = i3(c_ore[3 * k], c_ore[3 * k + 1], c_ore[3 * k + 2]);
            e.dom[s].ore = (int8_t)unit_index(o);
        }
    }
    e.set_all_domains();
}

extern "C" __global__ void ldo_sweep_kernel(Params P, int64_t steps, const int32_t* rep_slot) {
    bool active;
    int r = replica_of_thread(P, active);
    if (!active) {
        return;
    }
    Eng e;
    e.attach(&P, P.blocks + (size_t)r * P.L.bytes, r);
    if (e.S->status != ST_OK) {
        return;
    }
    if (rep_slot != nullptr) {
        e.begin_interval(rep_slot[r]);
    }
    e.simulate(steps);
    if (rep_slot != nullptr) {
        e.update_enthalpy_and_entropy();
    }
}

extern "C" __global__ void ldo_dependent_kernel(Params P) {
    bool active;
    int r = replica_of_thread(P, active);
    if (!active) {
        return;
    }
    Eng e;
    e.attach(&P, P.blocks + (size_t)r * P.L.bytes, r);
    e.update_enthalpy_and_entropy();
}

// Exchange state on the device
struct XState {
    int32_t* q_to_rep;
    int32_t* rep_slot;
    const double* ladder_t;
    const double* ladder_u;
    const double* ladder_s;
    const double* reduced_u; // [n_replicas * (ntypes-1)]
    int64_t* counts;         // attempts, accepts
};

__device__ double exchange_uniform(const Params& P, int64_t swap_i, uint32_t& idx) {
    if (P.rng_mode == 1) {
        RepScalars* S0 = (RepScalars*)P.blocks;
        int64_t pos = S0->tape_pos;
        if (pos >= P.tape_len[0]) {
            S0->status = ST_TAPE;
            return 0.0;
        }
        S0->tape_pos = pos + 1;
        return __longlong_as_double((long long)P.tapes[P.tape_off[0] + pos]);
    }
    uint32_t c0 = idx++, c1 = 0u, c2 = (uint32_t)(uint64_t)swap_i,
             c3 = (uint32_t)((uint64_t)swap_i >> 32);
    uint32_t k0 = (uint32_t)P.seed, k1 = 0xFFFFFFFFu;
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    uint64_t x = (uint64_t)c0 | ((uint64_t)c1 << 32);
    return (double)(x >> 11) * 0x1.0p-53;
}

// OneDPTGCMCSimulation::attempt_exchange + calc_acceptance_p
// (ptmc_simulation.cpp:255-313, 360-412).  The per-staple chemical
// potential term runs over the staple types.
extern "C" __global__ void ldo_exchange_kernel(Params P, XState X, int64_t swap_i) {
    if (threadIdx.x != 0 || blockIdx.x != 0) {
        return;
    }
    int R = P.n_replicas;
    int nst = P.L.ntypes - 1;
    uint32_t idx = 0;
    for (int i = (int)(swap_i % 2); i < R - 1; i += 2) {
        X.counts[0]++;
        int r1 = X.q_to_rep[i], r2 = X.q_to_rep[i + 1];
        RepScalars* S1 = (RepScalars*)(P.blocks + (size_t)r1 * P.L.bytes);
        RepScalars* S2 = (RepScalars*)(P.blocks + (size_t)r2 * P.L.bytes);
        const int16_t* tc1 = (const int16_t*)((char*)S1 + P.L.off_tcount);
        const int16_t* tc2 = (const int16_t*)((char*)S2 + P.L.off_tcount);
        double t1 = X.ladder_t[i], t2 = X.ladder_t[i + 1];
        double sm1 = X.ladder_s[i], sm2 = X.ladder_s[i + 1];
        double dbu_dn = 0;
        for (int k = 0; k < nst; k++) {
            double u1 = X.reduced_u[(size_t)r1 * nst + k] * t1 * X.ladder_u[i];
            double u2 = X.reduced_u[(size_t)r2 * nst + k] * t2 * X.ladder_u[i + 1];
            dbu_dn += (u2 / t2 - u1 / t1) * ((double)tc2[k + 1] - (double)tc1[k + 1]);
        }
        double DB = 1 / t2 - 1 / t1;
        double DH = S2->hyb_h * t2 - S1->hyb_h * t1;
        double Dst = S2->stack_e * t2 - S1->stack_e * t1;
        double DBM = sm2 / t2 - sm1 / t1;
        double DBias = 0.0 * t2 - 0.0 * t1;
        double p = fmin(1.0, exp(DB * (DH + DBias) + DBM * Dst - dbu_dn));
        bool accept = (p == 1) ? true : (p > exchange_uniform(P, swap_i, idx));
        if (accept) {
            X.counts[1]++;
            X.q_to_rep[i] = r2;
            X.q_to_rep[i + 1] = r1;
            X.rep_slot[r2] = i;
            X.rep_slot[r1] = i + 1;
        }
    }
}

extern "C" __global__ void ldo_status_kernel(Params P, int32_t* out) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < P.n_replicas) {
        out[r] = ((RepScalars*)(P.blocks + (size_t)r * P.L.bytes))->status;
    }
}

// ------------------------------------------------------------------- host
namespace {

std::string g_create_error;

struct Ctx {
    Params P {};
    ldo::SystemSpec spec;
    int R = 0;
    int ntypes = 0;
    std::vector<double> ladder_t, ladder_u, ladder_b, ladder_s;
    std::vector<double> reduced_u;
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
    double* d_tab = nullptr; // hyb_e | hyb_h | hyb_s | stack
    double* d_init = nullptr; // init_e | init_h | init_s
    uint64_t* d_tapes = nullptr;
    int64_t* d_tape_meta = nullptr; // off | len
    int32_t* d_xint = nullptr;     // q_to_rep | rep_slot | status
    double* d_xdbl = nullptr;      // ladder_t | ladder_u | ladder_s | reduced_u
    int64_t* d_xcnt = nullptr;     // exchange attempts, accepts
    int32_t* d_cfg = nullptr;
    size_t cfg_cap = 0;
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

struct AbiError {
    int code;
    std::string msg;
};

int64_t round_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

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

Layout make_layout(int ns, int lmax, int smax, int ntypes, int tcap) {
    Layout L;
    std::memset(&L, 0, sizeof L);
    L.ns = ns;
    L.lmax = lmax;
    L.smax = smax;
    L.nslots = ns + smax * lmax;
    L.nchain = 1 + smax;
    L.ntypes = ntypes;
    L.tcap = tcap;
    int hcap = 64, hbits = 6;
    while (hcap < 4 * L.nslots) {
        hcap *= 2;
        hbits++;
    }
    L.hcap = hcap;
    L.hbits = hbits;
    L.ecap = 4 * L.nslots + 16;
    int64_t off = round_up(sizeof(RepScalars), 16);
    auto take = [&](int64_t& field, int64_t bytes) {
        field = off;
        off = round_up(off + bytes, 16);
    };
    int64_t S = L.nslots, C = L.nchain;
    take(L.off_dom, S * (int64_t)sizeof(DomRec));
    take(L.off_dident, S * 2);
    take(L.off_chain, C * (int64_t)sizeof(ChainRec));
    take(L.off_order, C * 2);
    take(L.off_free, C * 2);
    take(L.off_tlist, (int64_t)ntypes * tcap * 2);
    take(L.off_tcount, (int64_t)ntypes * 2);
    take(L.off_hkey, (int64_t)hcap * 8);
    take(L.off_hval, (int64_t)hcap * 4);
    take(L.off_modified, 2 * S * 2);
    take(L.off_assigned, 2 * S * 2);
    take(L.off_prev, S * (int64_t)sizeof(PosOre));
    take(L.off_old, S * (int64_t)sizeof(PosOre));
    take(L.off_seg, S);
    take(L.off_gp, S * 2);
    take(L.off_stem, S * 2);
    take(L.off_inact, S * 2);
    take(L.off_netseg, S);
    take(L.off_netds_flag, S);
    take(L.off_checked, C);
    take(L.off_netcs, C);
    take(L.off_regrow, C);
    take(L.off_dir0, C);
    take(L.off_dir1, C);
    take(L.off_active, (int64_t)L.ecap * (int64_t)sizeof(Endpoint));
    take(L.off_active0, (int64_t)L.ecap * (int64_t)sizeof(Endpoint));
    take(L.off_gps, 2 * S * 2);
    take(L.off_iaes, 2 * S * 2);
    take(L.off_netds, S * 2);
    take(L.off_frames, C * (int64_t)sizeof(NetFrame));
    take(L.off_part, C);
    take(L.off_sframes, C * (int64_t)sizeof(NetFrame));
    L.bytes = round_up(off, 256);
    return L;
}

void upload_tables(Ctx& c) {
    int n = c.R;
    size_t n2 = (size_t)c.spec.n_ident * c.spec.n_ident;
    std::vector<double> tab(4 * n2 * n);
    std::vector<double> init(3 * n);
    for (int q = 0; q < n; q++) {
        ldo::TableSet t = ldo::build_tables(c.spec, c.ladder_t[q], c.ladder_s[q]);
        std::copy(t.hyb_e.begin(), t.hyb_e.end(), tab.begin() + (0 * n + q) * n2);
        std::copy(t.hyb_h.begin(), t.hyb_h.end(), tab.begin() + (1 * (size_t)n + q) * n2);
        std::copy(t.hyb_s.begin(), t.hyb_s.end(), tab.begin() + (2 * (size_t)n + q) * n2);
        std::copy(t.stack.begin(), t.stack.end(), tab.begin() + (3 * (size_t)n + q) * n2);
        init[q] = t.init_e;
        init[n + q] = t.init_h;
        init[2 * n + q] = t.init_s;
    }
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
    c.P.n_slots = n;
    std::vector<double> xd(3 * n + c.reduced_u.size());
    for (int q = 0; q < n; q++) {
        xd[q] = c.ladder_t[q];
        xd[n + q] = c.ladder_u[q];
        xd[2 * n + q] = c.ladder_s[q];
    }
    std::copy(c.reduced_u.begin(), c.reduced_u.end(), xd.begin() + 3 * n);
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

void upload_config(Ctx& c, const ldo_config* cfg, int& n_domains) {
    if (cfg->n_chains < 1 || cfg->chain_identity[0] != 0) {
        throw AbiError {LDO_E_ARG, "configuration must start with the scaffold chain"};
    }
    if (cfg->n_chains - 1 > c.P.max_total_staples) {
        throw AbiError {LDO_E_ARG, "configuration holds more staples than max_total_staples"};
    }
    n_domains = 0;
    for (int i = 0; i < cfg->n_chains; i++) {
        int id = cfg->chain_identity[i];
        if (id < 0 || id >= c.ntypes) {
            throw AbiError {LDO_E_ARG, "chain identity out of range"};
        }
        n_domains += (int)c.spec.idents[id].size();
    }
    for (int k = 0; k < 3 * n_domains; k++) {
        int v = cfg->orientations[k];
        if (v < -1 || v > 1) {
            throw AbiError {LDO_E_ARG, "orientation is not a unit vector"};
        }
    }
    size_t need = (size_t)2 * cfg->n_chains + (size_t)6 * n_domains;
    if (need > c.cfg_cap) {
        release(c, c.d_cfg, "hipFree cfg");
        hip_check(hipMalloc(&c.d_cfg, need * sizeof(int32_t)), "hipMalloc cfg");
        c.cfg_cap = need;
    }
    std::vector<int32_t> h(need);
    std::copy(cfg->chain_index, cfg->chain_index + cfg->n_chains, h.begin());
    std::copy(cfg->chain_identity, cfg->chain_identity + cfg->n_chains, h.begin() + cfg->n_chains);
    std::copy(cfg->positions, cfg->positions + 3 * n_domains, h.begin() + 2 * cfg->n_chains);
    std::copy(cfg->orientations, cfg->orientations + 3 * n_domains,
              h.begin() + 2 * cfg->n_chains + 3 * n_domains);
    upload(c, c.d_cfg, h.data(), need * sizeof(int32_t), "copy cfg");
}

void check_statuses(Ctx& c) {
    hipLaunchKernelGGL(ldo_status_kernel, dim3((c.R + 255) / 256), dim3(256), 0, c.stream, c.P,
                       c.d_xint + 2 * c.R);
    hip_check(hipGetLastError(), "launch status");
    std::vector<int32_t> st(c.R);
    hip_check(hipMemcpyAsync(st.data(), c.d_xint + 2 * c.R, c.R * sizeof(int32_t),
                             hipMemcpyDeviceToHost, c.stream),
              "copy status");
    hip_check(hipStreamSynchronize(c.stream), "sync");
    for (int r = 0; r < c.R; r++) {
        if (st[r] != ST_OK) {
            const char* why = "unknown";
            int code = LDO_E_MISUSE;
            switch (st[r]) {
            case ST_CONSTRAINT:
                why = "constraint violation (OrigamiMisuse)";
                code = LDO_E_CONSTRAINT;
                break;
            case ST_CAPACITY:
                why = "replica capacity exceeded";
                break;
            case ST_TAPE:
                why = "recorded draw sequence exhausted";
                break;
            case ST_COORD:
                why = "lattice coordinate outside the packed-key range";
                break;
            case ST_MISUSE:
                why = "sampler misuse";
                break;
            }
            throw AbiError {code, "replica " + std::to_string(r) + ": " + why};
        }
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
    try {
        f();
    } catch (AbiError& e) {
        if (c) {
            c->err = e.msg;
        }
        else {
            g_create_error = e.msg;
        }
        return e.code;
    } catch (HipError& e) {
        std::string m = std::string(e.what) + ": " + hipGetErrorString(e.e);
        if (c) {
            c->err = m;
        }
        else {
            g_create_error = m;
        }
        return LDO_E_HIP;
    } catch (std::exception& e) {
        if (c) {
            c->err = e.what();
        }
        else {
            g_create_error = e.what();
        }
        return LDO_E_CONSTRAINT;
    }
    return LDO_OK;
}

// The engine kernels (sweep, exchange, dependent quantities) took an
// illegal-address fault in their first run on an MI355X.  The cause has not
// been identified: the same code is clean under AddressSanitizer and
// UndefinedBehaviorSanitizer in a CPU build over hundreds of seeds.  Until a
// cause is shown the engine stays switched off by default.
bool engine_enabled() {
    const char* v = std::getenv("LDO_ENABLE_UNVERIFIED_ENGINE");
    return v != nullptr && std::strcmp(v, "1") == 0;
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
                    c->d_tapes, c->d_tape_meta, c->d_xint, c->d_xdbl, c->d_xcnt, c->d_cfg};
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
        if (!engine_enabled()) {
            throw AbiError {LDO_E_NOTIMPL,
                            "the HIP engine is switched off: its sweep kernel took an "
                            "unexplained illegal-address fault on gfx950; set "
                            "LDO_ENABLE_UNVERIFIED_ENGINE=1 to run it anyway"};
        }
        if (!sys || !run || !init || !temps || !out_ctx || n_replicas < 1) {
            throw AbiError {LDO_E_ARG, "null argument or no replicas"};
        }
        if (sys->apply_mean_field_cor) {
            throw AbiError {LDO_E_NOTIMPL, "apply_mean_field_cor is not supported"};
        }
        if (sys->n_chain_types < 2) {
            throw AbiError {LDO_E_ARG, "a system needs a scaffold and at least one staple type"};
        }
        if (run->n_movetypes < 1 || run->n_movetypes > LDO_MAX_MOVES) {
            throw AbiError {LDO_E_ARG, "moveset must hold 1..8 movetypes"};
        }
        if (run->exchange_mode == LDO_EXCHANGE_TEMPERATURE && n_replicas > 64) {
            throw AbiError {LDO_E_ARG, "a temperature ladder holds at most 64 slots"};
        }
        c->R = n_replicas;
        c->ntypes = sys->n_chain_types;
        ldo::SystemSpec& sp = c->spec;
        int k = 0, lo = 0, hi = 0, lmax = 0;
        bool first = true;
        for (int t = 0; t < sys->n_chain_types; t++) {
            std::vector<int> ids;
            std::vector<std::string> sq;
            for (int d = 0; d < sys->chain_lengths[t]; d++, k++) {
                int v = sys->identities[k];
                ids.push_back(v);
                if (first) {
                    lo = hi = v;
                    first = false;
                }
                lo = std::min(lo, v);
                hi = std::max(hi, v);
                if (sys->hybridization_pot == LDO_HYB_NEARESTNEIGHBOUR ||
                    sys->stacking_pot == LDO_STACK_SEQUENCESPECIFIC) {
                    if (!sys->sequences) {
                        throw AbiError {LDO_E_ARG, "sequences are required by this potential"};
                    }
                    sq.push_back(sys->sequences[k]);
                }
                else {
                    sq.push_back("");
                }
            }
            if (t > 0) {
                lmax = std::max(lmax, (int)ids.size());
            }
            sp.idents.push_back(ids);
            sp.seqs.push_back(sq);
        }
        if (lmax > 8 || lmax > run->max_staple_size) {
            throw AbiError {LDO_E_ARG, "staples longer than max_staple_size (at most 8)"};
        }
        sp.ident_lo = lo;
        sp.n_ident = hi - lo + 1;
        sp.hybridization_pot = sys->hybridization_pot;
        sp.stacking_pot = sys->stacking_pot;
        sp.binding_h = sys->binding_h;
        sp.binding_s = sys->binding_s;
        sp.misbinding_h = sys->misbinding_h;
        sp.misbinding_s = sys->misbinding_s;
        sp.stacking_ene = sys->stacking_ene;
        sp.cation_M = sys->cation_M;
        for (int i = 0; i < sys->n_specified; i++) {
            sp.spec_h.push_back(sys->enthalpies[i]);
            sp.spec_s.push_back(sys->entropies[i]);
        }
        int ns = (int)sp.idents[0].size();
        int smax = run->max_total_staples;
        if (smax < 1 || smax > 4000) {
            throw AbiError {LDO_E_ARG, "max_total_staples must be in 1..4000"};
        }
        if (ns + (int64_t)smax * lmax > 30000) {
            throw AbiError {LDO_E_ARG, "too many domain slots for 16-bit slot indices"};
        }
        int tcap = std::min(std::max(run->max_type_staples, 1), smax);
        Params& P = c->P;
        P.L = make_layout(ns, lmax, smax, sys->n_chain_types, tcap);
        P.n_replicas = n_replicas;
        P.domain_type = sys->domain_type;
        P.misbinding_pot = sys->misbinding_pot;
        P.cyclic = sys->cyclic;
        P.n_moves = run->n_movetypes;
        double cum = 0;
        for (int m = 0; m < run->n_movetypes; m++) {
            int kind = run->movetype_kind[m];
            if (kind < 0 || kind > 3) {
                throw AbiError {LDO_E_NOTIMPL, "movetype not supported"};
            }
            P.move_kind[m] = kind;
            cum += run->movetype_freq[m];
            P.move_cum[m] = cum;
            P.move_maxreg[m] = run->movetype_max_regrowth ? run->movetype_max_regrowth[m] : 0;
            if (kind == 3 && P.move_maxreg[m] < 2) {
                throw AbiError {LDO_E_ARG, "CTCB max_regrowth must be at least 2"};
            }
        }
        P.max_total_staples = smax;
        P.max_type_staples = run->max_type_staples;
        P.centering_freq = run->centering_freq;
        P.centering_domain = run->centering_domain;
        if (P.centering_domain < 0 || P.centering_domain >= ns) {
            throw AbiError {LDO_E_ARG, "centering_domain out of range"};
        }
        P.check_freq = run->constraint_check_freq;
        P.rng_mode = run->rng_mode;
        P.seed = run->seed;
        P.ident_lo = lo;
        P.n_ident = sp.n_ident;
        P.fugacity = sys->staple_M;
        int lpr = run->lanes_per_replica;
        if (lpr == 0) {
            lpr = n_replicas >= 2048 ? 1 : 64;
        }
        if (lpr != 1 && lpr != 64) {
            throw AbiError {LDO_E_ARG, "lanes_per_replica must be 0, 1 or 64"};
        }
        P.lanes_per_replica = lpr;

        hip_check(hipStreamCreate(&c->stream), "stream");
        // chain lengths by identity
        std::vector<int32_t> tl(sys->n_chain_types);
        for (int t = 0; t < sys->n_chain_types; t++) {
            tl[t] = (int)sp.idents[t].size();
        }
        hip_check(hipMalloc(&c->d_type_len, tl.size() * sizeof(int32_t)), "hipMalloc type_len");
        upload(*c, c->d_type_len, tl.data(), tl.size() * sizeof(int32_t), "copy type_len");
        P.type_len = c->d_type_len;
        // ladder / per-replica chemical potentials (molarity_to_chempots,
        // chempots_to_reduced_chempots: origami_system.cpp:965-989)
        int nst = sys->n_chain_types - 1;
        c->reduced_u.assign((size_t)n_replicas * nst, 0.0);
        for (int r = 0; r < n_replicas; r++) {
            c->ladder_t.push_back(temps[r]);
            c->ladder_u.push_back(1.0);
            c->ladder_b.push_back(1.0);
            c->ladder_s.push_back(1.0);
            for (int t = 0; t < nst; t++) {
                double extra = (2 * sp.idents[t + 1].size() - 1) * std::log(6);
                double u = temps[r] * (std::log(sys->staple_M) - extra);
                c->reduced_u[(size_t)r * nst + t] = u / temps[r];
            }
        }
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
        // per-slot identity index of scaffold domains and staple types
        {
            int16_t* d_ti = nullptr;
            std::vector<int16_t> h((size_t)sys->n_chain_types * std::max(lmax, ns), 0);
            int st = std::max(lmax, ns);
            for (int t = 0; t < sys->n_chain_types; t++) {
                for (size_t d = 0; d < sp.idents[t].size(); d++) {
                    h[(size_t)t * st + d] = (int16_t)(sp.idents[t][d] - lo);
                }
            }
            hip_check(hipMalloc(&d_ti, h.size() * sizeof(int16_t)), "hipMalloc type_ident");
            c->d_type_ident = d_ti;
            upload(*c, d_ti, h.data(), h.size() * sizeof(int16_t), "copy type_ident");
            P.type_ident = d_ti;
            P.type_ident_stride = st;
        }
        int32_t* cf = c->d_cfg;
        hipLaunchKernelGGL(ldo_init_kernel, grid_of(*c), dim3(64), 0, c->stream, c->P, cf,
                           cf + init->n_chains, cf + 2 * init->n_chains,
                           cf + 2 * init->n_chains + 3 * nd, init->n_chains, -1, 0,
                           (const int32_t*)nullptr);
        hip_check(hipGetLastError(), "launch init");
        check_statuses(*c);
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
        Ctx* c = static_cast<Ctx*>(ctx);
        (void)hipStreamSynchronize(c->stream);
        free_ctx(c);
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
        const Layout& L = c->P.L;
        std::vector<char> blk(L.bytes);
        hip_check(hipMemcpyAsync(blk.data(), c->d_blocks + (size_t)replica * L.bytes, L.bytes,
                                 hipMemcpyDeviceToHost, c->stream),
                  "copy block");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        RepScalars* S = (RepScalars*)blk.data();
        DomRec* dom = (DomRec*)(blk.data() + L.off_dom);
        ChainRec* ch = (ChainRec*)(blk.data() + L.off_chain);
        int16_t* order = (int16_t*)(blk.data() + L.off_order);
        if (S->n_chains > cap_chains) {
            throw AbiError {LDO_E_ARG, "chain capacity too small"};
        }
        static const int U[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0},
                                    {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        int k = 0;
        for (int w = 0; w < S->n_chains; w++) {
            int cs = order[w];
            cfg->chain_index[w] = ch[cs].uid;
            cfg->chain_identity[w] = ch[cs].ident;
            cfg->chain_length[w] = ch[cs].len;
            for (int d = 0; d < ch[cs].len; d++, k++) {
                if (k >= cap_domains) {
                    throw AbiError {LDO_E_ARG, "domain capacity too small"};
                }
                int s = cs == 0 ? d : L.ns + (cs - 1) * L.lmax + d;
                cfg->positions[3 * k] = dom[s].x;
                cfg->positions[3 * k + 1] = dom[s].y;
                cfg->positions[3 * k + 2] = dom[s].z;
                for (int q = 0; q < 3; q++) {
                    cfg->orientations[3 * k + q] = U[dom[s].ore][q];
                }
            }
        }
        cfg->n_chains = S->n_chains;
    });
}

int ldo_set_ladder(void* ctx, int32_t n, const double* temps, const double* u, const double* b,
                   const double* s) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        if (n != c->R) {
            throw AbiError {LDO_E_ARG, "ladder size must equal the replica count"};
        }
        c->ladder_t.assign(temps, temps + n);
        c->ladder_u.assign(u, u + n);
        c->ladder_b.assign(b, b + n);
        c->ladder_s.assign(s, s + n);
        hip_check(hipStreamSynchronize(c->stream), "sync");
        upload_tables(*c);
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
        launch_timed(*c, 0, [&]() {
            hipLaunchKernelGGL(ldo_sweep_kernel, grid_of(*c), dim3(64), 0, c->stream, c->P,
                               steps, (const int32_t*)nullptr);
        });
        hip_check(hipGetLastError(), "launch sweep");
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
            launch_timed(*c, 0, [&]() {
                hipLaunchKernelGGL(ldo_sweep_kernel, grid_of(*c), dim3(64), 0, c->stream, c->P,
                                   interval, (const int32_t*)X.rep_slot);
            });
            hip_check(hipGetLastError(), "launch sweep");
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
        hip_check(hipMemcpyAsync(&S, c->d_blocks + (size_t)replica * c->P.L.bytes, sizeof S,
                                 hipMemcpyDeviceToHost, c->stream),
                  "copy scalars");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        out->energy = S.energy;
        out->enthalpy = S.hyb_h;
        out->entropy = S.hyb_s;
        out->stacking_energy = S.stack_e;
        out->temperature = c->ladder_t[S.temp_slot];
        out->step = S.step;
        out->num_staples = S.num_staples;
        out->num_domains = S.num_domains;
        out->num_bound_domain_pairs = S.num_bound;
        out->num_fully_bound_domain_pairs = S.num_fully;
        out->num_misbound_domain_pairs = S.num_bound - S.num_fully;
        out->num_stacked_domain_pairs = S.num_stacked;
        out->num_unassigned_domains = S.num_unassigned;
        out->status = S.status;
    });
}

int ldo_get_counters(void* ctx, ldo_counters* out) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        std::memset(out, 0, sizeof *out);
        std::vector<RepScalars> all(c->R);
        for (int r = 0; r < c->R; r++) {
            hip_check(hipMemcpyAsync(&all[r], c->d_blocks + (size_t)r * c->P.L.bytes,
                                     sizeof(RepScalars), hipMemcpyDeviceToHost, c->stream),
                      "copy scalars");
        }
        int64_t x[2];
        hip_check(hipMemcpyAsync(x, c->d_xcnt, sizeof x, hipMemcpyDeviceToHost, c->stream),
                  "copy xcnt");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        for (auto& S: all) {
            for (int m = 0; m < LDO_MAX_MOVES; m++) {
                out->attempts[m] += S.attempts[m];
                out->accepts[m] += S.accepts[m];
            }
            out->site_empty += S.site_empty;
            out->site_bind += S.site_bind;
            out->commits += S.commits;
        }
        out->exchange_attempts = x[0];
        out->exchange_accepts = x[1];
    });
}

int ldo_get_permutation(void* ctx, int32_t* q_to_rep) {
    Ctx* c = static_cast<Ctx*>(ctx);
    return guarded(c, [&]() {
        hip_check(hipMemcpyAsync(q_to_rep, c->d_xint, c->R * sizeof(int32_t),
                                 hipMemcpyDeviceToHost, c->stream),
                  "copy permutation");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

void* ldo_get_stream(void* ctx) { return ctx ? (void*)static_cast<Ctx*>(ctx)->stream : nullptr; }

const char* ldo_kernel_name(void) { return "ldo_sweep_kernel"; }

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
