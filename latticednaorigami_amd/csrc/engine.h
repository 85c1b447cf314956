// engine.h - device-side replica engine of the sampler (gfx950).
//
// One Eng instance drives one replica: it owns pointers into the replica's
// block (layout.h) and implements the reference's lattice state updates,
// FourBody energy and move types on slot-indexed arrays and an
// open-addressed site hash.  Function names follow the reference's
// operations; each cites the reference function whose semantics it
// reproduces (the random draws are made in the same order, so a replica's
// trajectory is a function of its random stream alone).
#ifndef LDO_ENGINE_H
#define LDO_ENGINE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "layout.h"

// Engine functions compile for both the device (the kernels) and the host,
// where a test build runs the same code against the CPU oracle.
#define DEV __host__ __device__ __forceinline__
// Call boundaries: DEVNI functions are compiled once and called; DEVIN
// functions (the FourBody walk's parts, the occupancy updates) are inlined
// into the DEVNI function that calls them, so a binding evaluation runs as
// a few large functions with its operands in registers.
#define DEVNI __host__ __device__ __noinline__
#define DEVIN __host__ __device__ __forceinline__

struct I3 {
    int x, y, z;
};

// the double whose IEEE-754 bit pattern is b (a recorded uniform_real)
DEV double bits_to_double(uint64_t b) {
    double d;
    __builtin_memcpy(&d, &b, sizeof d);
    return d;
}

DEV I3 i3(int x, int y, int z) {
    I3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
DEV I3 add3(I3 a, I3 b) { return i3(a.x + b.x, a.y + b.y, a.z + b.z); }
DEV I3 sub3(I3 a, I3 b) { return i3(a.x - b.x, a.y - b.y, a.z - b.z); }
DEV I3 neg3(I3 a) { return i3(-a.x, -a.y, -a.z); }
DEV bool eq3(I3 a, I3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// unit vector of orientation index o (+x,-x,+y,-y,+z,-z)
DEV I3 unit3(int o) {
    int a = o >> 1;
    int s = (o & 1) ? -1 : 1;
    return i3(a == 0 ? s : 0, a == 1 ? s : 0, a == 2 ? s : 0);
}

// index of a unit vector (never called on other vectors)
DEV int unit_index(I3 v) {
    if (v.x != 0) {
        return v.x > 0 ? 0 : 1;
    }
    if (v.y != 0) {
        return v.y > 0 ? 2 : 3;
    }
    return v.z > 0 ? 4 : 5;
}

// VectorThree::rotate_half (utility.cpp:68-85)
DEV I3 rotate_half3(I3 v, I3 axis) {
    int ax = axis.x < 0 ? -axis.x : axis.x;
    int ay = axis.y < 0 ? -axis.y : axis.y;
    int az = axis.z < 0 ? -axis.z : axis.z;
    if (ax == 1 && ay == 0 && az == 0) {
        return i3(v.x, -v.y, -v.z);
    }
    if (ax == 0 && ay == 1 && az == 0) {
        return i3(-v.x, v.y, -v.z);
    }
    if (ax == 0 && ay == 0 && az == 1) {
        return i3(-v.x, -v.y, v.z);
    }
    return v;
}

// VectorThree::rotate(axis, -1) (utility.cpp:102-142): the quarter turn the
// ThreeQuarterTurn twist rule uses.
DEV I3 rotate_minus_quarter3(I3 v, I3 axis) {
    int dir = -1;
    int ax = axis.x < 0 ? -axis.x : axis.x;
    int ay = axis.y < 0 ? -axis.y : axis.y;
    int az = axis.z < 0 ? -axis.z : axis.z;
    if (axis.x != ax || axis.y != ay || axis.z != az) {
        dir = -dir;
    }
    if (ax == 1 && ay == 0 && az == 0) {
        return i3(v.x, -dir * v.z, dir * v.y);
    }
    if (ax == 0 && ay == 1 && az == 0) {
        return i3(-dir * v.z, v.y, dir * v.x);
    }
    if (ax == 0 && ay == 0 && az == 1) {
        return i3(-dir * v.y, dir * v.x, v.z);
    }
    return v;
}

struct Delta {
    double e;
    int st;
};

// The CTRG level being grown (CTRGRegrowthMCMovetype's m_di, m_d, m_dir,
// m_ref_d, m_c_attempts, m_d_max_c_attempts, m_stemd and m_avail_cis, the
// configurations left as bits of the 36 (site, orientation) pairs)
struct RgCur {
    int di, d, dir, ref, att, maxatt;
    bool stemd;
    uint64_t avail;
};
#define LDO_RG_ALL ((1ull << 36) - 1)
// how a CTRG assignment is recorded (Eng::rg_set_config)
enum { RG_NONE = 0, RG_LIST = 1, RG_DUP = 2 };

struct Eng {
    const Params* P;
    int rep;
    RepScalars* S;
    DomRec* dom;
    int16_t* dident;
    ChainRec* ch;
    int16_t* order;
    int16_t* freel;
    int16_t* tlist;
    int16_t* tcount;
    uint64_t* hkey;
    uint32_t* hval;
    int16_t* modified;
    int16_t* assigned;
    PosOre* prev;
    PosOre* old;
    int8_t* seg;
    int16_t* gp;
    int16_t* stem;
    int16_t* inact;
    int8_t* netseg;
    int8_t* netds_flag;
    int8_t* checked;
    int8_t* netcs;
    int8_t* regrow;
    int8_t* dir0;
    int8_t* dir1;
    Endpoint* active;
    Endpoint* active0;
    int16_t* gps;
    int16_t* iaes;
    int16_t* netds;
    NetFrame* frames;
    int8_t* part;
    NetFrame* sframes;
    GrowFrame* gstack;
    int16_t* gpl;
    int16_t* rgs;
    int8_t* rgatt;
    int8_t* rgwatt;
    uint64_t* rgavail;
    uint64_t* rgwavail;
    double* rgopen;
    Endpoint* erstk;
    int16_t* ercnt;
    int8_t* jdom;
    int16_t* jps;
    JumpSeg* jseg;
    int16_t* jstem;
    int8_t* stemseg;
    int8_t* sdir;
    int64_t* track;
    const double* t_hyb;
    const double* t_stack;
    double init_e;
    int ns, lshift, lmask, nslots, nchain, hmask, hbits, n_ident;
    // launch parameters the hot paths read, copied out of Params so that a
    // read is one access to the engine object rather than two dependent ones
    int cyc, dtype, mispot, rngmode;
    uint64_t seed;
    // FourBody accumulators (BindingPotential::m_delta_config)
    double dc_e;
    int dc_st;
    bool fb_viol;
    bool pot_viol;
    int n_gps, n_iaes, n_netds;
    bool net_external;
    // CTRG: the move's recoil and attempt limits, and the assignments the
    // reference records beyond the recoil growth's (see rg_set_config)
    int rg_max_recoils, rg_max_att, rg_max_segreg;
    int64_t rg_dup;
    // par = 1: the 64 lanes of a wavefront run this engine together on
    // identical values (every lane computes and stores the same thing), and
    // the lane-parallel sections split independent reads over the lanes and
    // combine them with wave votes and shuffles.  par = 0: one lane runs the
    // engine and the sections run serially.
    int par;

    // ------------------------------------------------------------ lanes
    DEV int plane() const {
#ifdef __HIP_DEVICE_COMPILE__
        return par ? (int)(threadIdx.x & 63u) : 0;
#else
        return 0;
#endif
    }
    DEV int pwidth() const { return par ? 64 : 1; }
    // whether b holds on any lane of the section
    DEV bool pany(bool b) const {
#ifdef __HIP_DEVICE_COMPILE__
        return par ? __any((int)b) != 0 : b;
#else
        return b;
#endif
    }
    // lanes where b holds (bit l: lane l)
    DEV uint64_t pballot(bool b) const {
#ifdef __HIP_DEVICE_COMPILE__
        return par ? (uint64_t)__ballot((int)b) : (b ? 1ull : 0ull);
#else
        return b ? 1ull : 0ull;
#endif
    }
    // value v of lane src
    DEV int pshfl(int v, int src) const {
#ifdef __HIP_DEVICE_COMPILE__
        return par ? __shfl(v, src) : v;
#else
        (void)src;
        return v;
#endif
    }
    // inclusive prefix sum of v over lanes 0..plane()
    DEV int pscan(int v) const {
#ifdef __HIP_DEVICE_COMPILE__
        if (par) {
            int lane = plane();
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                int t = __shfl_up(v, off);
                if (lane >= off) {
                    v += t;
                }
            }
        }
#endif
        return v;
    }

    // ------------------------------------------------------------ wiring
    // Fills a block section of `bytes` bytes with a 32-bit pattern in 16-byte
    // stores.  Sections start 16-byte aligned and own the padding up to the
    // next 16-byte boundary (the replica-block layout rounds every section).
    DEV void fill_section(void* base, int64_t bytes, uint32_t pattern) {
        struct alignas(16) Quad {
            uint32_t a, b, c, d;
        };
        Quad q {pattern, pattern, pattern, pattern};
        Quad* dst = (Quad*)base;
        for (int64_t i = plane(), n = (bytes + 15) / 16; i < n; i += pwidth()) {
            dst[i] = q;
        }
    }

    DEV void attach(const Params* p, char* blk, int r, int par_ = 0) {
        P = p;
        rep = r;
        par = par_;
        const Layout& L = p->L;
        S = (RepScalars*)blk;
        dom = (DomRec*)(blk + L.off_dom);
        dident = (int16_t*)(blk + L.off_dident);
        ch = (ChainRec*)(blk + L.off_chain);
        order = (int16_t*)(blk + L.off_order);
        freel = (int16_t*)(blk + L.off_free);
        tlist = (int16_t*)(blk + L.off_tlist);
        tcount = (int16_t*)(blk + L.off_tcount);
        hkey = (uint64_t*)(blk + L.off_hkey);
        hval = (uint32_t*)(blk + L.off_hval);
        modified = (int16_t*)(blk + L.off_modified);
        assigned = (int16_t*)(blk + L.off_assigned);
        prev = (PosOre*)(blk + L.off_prev);
        old = (PosOre*)(blk + L.off_old);
        seg = (int8_t*)(blk + L.off_seg);
        gp = (int16_t*)(blk + L.off_gp);
        stem = (int16_t*)(blk + L.off_stem);
        inact = (int16_t*)(blk + L.off_inact);
        netseg = (int8_t*)(blk + L.off_netseg);
        netds_flag = (int8_t*)(blk + L.off_netds_flag);
        checked = (int8_t*)(blk + L.off_checked);
        netcs = (int8_t*)(blk + L.off_netcs);
        regrow = (int8_t*)(blk + L.off_regrow);
        dir0 = (int8_t*)(blk + L.off_dir0);
        dir1 = (int8_t*)(blk + L.off_dir1);
        active = (Endpoint*)(blk + L.off_active);
        active0 = (Endpoint*)(blk + L.off_active0);
        gps = (int16_t*)(blk + L.off_gps);
        iaes = (int16_t*)(blk + L.off_iaes);
        netds = (int16_t*)(blk + L.off_netds);
        frames = (NetFrame*)(blk + L.off_frames);
        part = (int8_t*)(blk + L.off_part);
        sframes = (NetFrame*)(blk + L.off_sframes);
        gstack = (GrowFrame*)(blk + L.off_gstack);
        gpl = (int16_t*)(blk + L.off_gpl);
        rgs = (int16_t*)(blk + L.off_rgs);
        rgatt = (int8_t*)(blk + L.off_rgatt);
        rgwatt = (int8_t*)(blk + L.off_rgwatt);
        rgavail = (uint64_t*)(blk + L.off_rgavail);
        rgwavail = (uint64_t*)(blk + L.off_rgwavail);
        rgopen = (double*)(blk + L.off_rgopen);
        erstk = (Endpoint*)(blk + L.off_erstk);
        ercnt = (int16_t*)(blk + L.off_ercnt);
        jdom = (int8_t*)(blk + L.off_jdom);
        jps = (int16_t*)(blk + L.off_jps);
        jseg = (JumpSeg*)(blk + L.off_jseg);
        jstem = (int16_t*)(blk + L.off_jstem);
        stemseg = (int8_t*)(blk + L.off_stemseg);
        sdir = (int8_t*)(blk + L.off_sdir);
        track = (int64_t*)(blk + L.off_track);
        ns = L.ns;
        lshift = L.lshift;
        lmask = (1 << L.lshift) - 1;
        nslots = L.nslots;
        nchain = L.nchain;
        hbits = L.hbits;
        hmask = L.hcap - 1;
        n_ident = p->n_ident;
        cyc = p->cyclic;
        dtype = p->domain_type;
        mispot = p->misbinding_pot;
        rngmode = p->rng_mode;
        seed = p->seed;
        select_tables();
    }

    DEV void select_tables() {
        size_t n2 = (size_t)P->n_ident * P->n_ident;
        t_hyb = P->hyb_e + (size_t)S->temp_slot * n2;
        t_stack = P->stack + (size_t)S->temp_slot * n2;
        init_e = P->init_e[S->init_slot];
    }

    // bounded appends to the move's domain lists
    DEV void push_modified(int s) {
        if (S->n_modified >= 2 * nslots) {
            fail(ST_CAPACITY);
            return;
        }
        modified[S->n_modified++] = (int16_t)s;
    }
    DEV void push_assigned(int s) {
        if (S->n_assigned >= 2 * nslots) {
            fail(ST_CAPACITY);
            return;
        }
        assigned[S->n_assigned++] = (int16_t)s;
    }

    DEV void fail(int code) {
        if (S->status == ST_OK) {
            S->status = code;
        }
    }

    // ------------------------------------------------------------ rng
    // Philox4x32-10 (Salmon et al., SC'11); counter (draw, stream, step lo,
    // hi) with stream 0 for MC steps and 1 for learned-move attempts (whose
    // own count replaces the step), key (seed lo, replica)
    DEV uint64_t philox64() {
        uint64_t ctr = S->rng_sub ? (uint64_t)S->learned_step : (uint64_t)S->step;
        uint32_t c0 = (uint32_t)S->rng_idx, c1 = (uint32_t)S->rng_sub;
        uint32_t c2 = (uint32_t)ctr, c3 = (uint32_t)(ctr >> 32);
        uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)rep;
#pragma unroll
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
        S->rng_idx++;
        return (uint64_t)c0 | ((uint64_t)c1 << 32);
    }

    DEV uint64_t tape_next() {
        int64_t pos = S->tape_pos;
        if (pos >= P->tape_len[rep]) {
            fail(ST_TAPE);
            return 0;
        }
        S->tape_pos = pos + 1;
        return P->tapes[P->tape_off[rep] + pos];
    }

    // RandomGens::uniform_real (random_gens.cpp:22-24)
    DEV double uniform_real() {
        if (rngmode == 1) {
            return bits_to_double(tape_next());
        }
        return (double)(philox64() >> 11) * 0x1.0p-53;
    }

    // RandomGens::uniform_int (random_gens.cpp:26-43)
    DEV int uniform_int(int lo, int hi) {
        if (rngmode == 1) {
            return (int)(int64_t)tape_next();
        }
        uint64_t x = philox64();
        uint64_t n = (uint64_t)((int64_t)hi - (int64_t)lo + 1);
        return lo + (int)(((x >> 32) * n) >> 32);
    }

    // ------------------------------------------------------------ slots
    DEV int cslot_of(int s) const { return s < ns ? 0 : 1 + ((s - ns) >> lshift); }
    DEV int d_of(int s) const { return s < ns ? s : ((s - ns) & lmask); }
    DEV int slot(int cs, int d) const { return cs == 0 ? d : ns + ((cs - 1) << lshift) + d; }
    DEV int len_of(int cs) const { return ch[cs].len; }

    // Domain::m_forward_domain / m_backward_domain (origami_system.cpp:427-433, 686-692)
    // the scaffold occupies slots [0, ns) and is ns domains long
    DEV int fwd(int s) const {
        if (s < 0) {
            return LDO_NONE;
        }
        if (s < ns) {
            return s + 1 < ns ? s + 1 : (cyc ? 0 : LDO_NONE);
        }
        return d_of(s) + 1 < ch[cslot_of(s)].len ? s + 1 : LDO_NONE;
    }
    DEV int bwd(int s) const {
        if (s < 0) {
            return LDO_NONE;
        }
        if (s < ns) {
            return s > 0 ? s - 1 : (cyc ? ns - 1 : LDO_NONE);
        }
        return d_of(s) > 0 ? s - 1 : LDO_NONE;
    }
    // Domain::operator+ (domain.cpp:9-31)
    DEV int walk(int s, int incr) const {
        while (s >= 0 && incr > 0) {
            s = fwd(s);
            incr--;
        }
        while (s >= 0 && incr < 0) {
            s = bwd(s);
            incr++;
        }
        return s;
    }

    DEV I3 pos(int s) const { return i3(dom[s].x, dom[s].y, dom[s].z); }
    DEV I3 ore(int s) const { return unit3(dom[s].ore); }
    DEV int st(int s) const { return dom[s].state; }
    DEV int bd(int s) const { return dom[s].bound; }
    DEV bool eb(int s) const { return s >= 0 && dom[s].state == OCC_BOUND; }
    DEV int ident_of(int s) const { return dident[s] + P->ident_lo; }
    DEV double hyb(int a, int b) const { return t_hyb[(size_t)dident[a] * n_ident + dident[b]]; }
    DEV double stk(int a, int b) const { return t_stack[(size_t)dident[a] * n_ident + dident[b]]; }

    // ------------------------------------------------------------ site hash
    DEV uint64_t pkey(I3 p) const {
        return ((uint64_t)(uint32_t)(p.x + LDO_COORD_OFFSET) << 42) |
               ((uint64_t)(uint32_t)(p.y + LDO_COORD_OFFSET) << 21) |
               (uint64_t)(uint32_t)(p.z + LDO_COORD_OFFSET);
    }
    DEV int home(uint64_t k) const {
        return (int)((k * 0x9E3779B97F4A7C15ull) >> (64 - hbits));
    }
    DEV int hfind(uint64_t k) const {
        int i = home(k);
        for (int n = 0; n <= hmask; n++) {
            uint64_t kk = hkey[i];
            if (kk == k) {
                return i;
            }
            if (kk == LDO_HASH_EMPTY) {
                return -1;
            }
            i = (i + 1) & hmask;
        }
        return -1;
    }
    // OrigamiSystem::position_occupancy (origami_system.cpp:193-202)
    DEV int occupancy(I3 p) {
        int i = hfind(pkey(p));
        return i < 0 ? OCC_UNASSIGNED : (int)((hval[i] >> 16) & 3u);
    }
    DEV int unbound_at(I3 p) {
        int i = hfind(pkey(p));
        int s = i < 0 ? -1 : (int)(int16_t)(hval[i] & 0xFFFFu);
        if (s < 0 || s >= nslots) {
            fail(ST_CONSTRAINT);
            return 0;
        }
        return s;
    }
    DEV void hput(I3 p, int occ, int s) {
        if (p.x <= -LDO_COORD_OFFSET + 2 || p.x >= LDO_COORD_OFFSET - 2 ||
            p.y <= -LDO_COORD_OFFSET + 2 || p.y >= LDO_COORD_OFFSET - 2 ||
            p.z <= -LDO_COORD_OFFSET + 2 || p.z >= LDO_COORD_OFFSET - 2) {
            fail(ST_COORD);
        }
        uint64_t k = pkey(p);
        int i = home(k);
        int n = 0;
        while (hkey[i] != k && hkey[i] != LDO_HASH_EMPTY) {
            i = (i + 1) & hmask;
            if (++n > hmask) {
                fail(ST_CAPACITY);
                return;
            }
        }
        hkey[i] = k;
        hval[i] = ((uint32_t)occ << 16) | (uint32_t)(uint16_t)(int16_t)s;
    }
    // linear-probing deletion with backward shift (Knuth, TAOCP 6.4 Alg. R)
    DEV void herase(I3 p) {
        int i = hfind(pkey(p));
        if (i < 0) {
            return;
        }
        int j = i;
        for (int n = 0; n <= hmask; n++) {
            j = (j + 1) & hmask;
            uint64_t kj = hkey[j];
            if (kj == LDO_HASH_EMPTY) {
                break;
            }
            int k = home(kj);
            bool stay = (i <= j) ? (i < k && k <= j) : (i < k || k <= j);
            if (stay) {
                continue;
            }
            hkey[i] = kj;
            hval[i] = hval[j];
            i = j;
        }
        hkey[i] = LDO_HASH_EMPTY;
    }

    // ------------------------------------------------------------ rules
    // HalfTurnDomain / ThreeQuarterTurnDomain::check_twist_constraint
    // (domain.cpp:33-40, 69-78)
    DEV bool twist_ok(int a, I3 ndr, int b) const {
        I3 r = (dtype == 0) ? rotate_half3(ore(a), ndr)
                                     : rotate_minus_quarter3(ore(a), ndr);
        return eq3(r, ore(b));
    }
    // check_kink_constraint (domain.cpp:42-57, 80-97)
    DEV bool kink_ok(int a, I3 ndr, int b) const {
        I3 oa = ore(a), ob = ore(b);
        if (eq3(ndr, neg3(oa))) {
            return false;
        }
        if (eq3(ndr, oa)) {
            if (dtype == 0) {
                return !eq3(ob, neg3(oa));
            }
            return !(eq3(oa, ob) || eq3(oa, neg3(ob)));
        }
        if (eq3(ndr, ob) || eq3(ndr, neg3(ob))) {
            return false;
        }
        return true;
    }
    // check_junction_constraint (domain.cpp:59-67, 99-118)
    DEV bool junction_ok(int j1, int j2, int k1, int k2) const {
        if (dtype == 0) {
            return true;
        }
        I3 nk = sub3(pos(k2), pos(k1));
        if (eq3(nk, ore(k1))) {
            I3 n1 = sub3(pos(j2), pos(j1));
            if (d_of(j1) > d_of(j2)) {
                n1 = neg3(n1);
            }
            if (!twist_ok(k1, n1, k2)) {
                return false;
            }
        }
        return true;
    }
    // doubly_contiguous (origami_potential.cpp:51-68)
    DEV bool doubly_contig(int a, int b) const {
        if (cslot_of(a) != cslot_of(b) || d_of(b) != d_of(a) + 1) {
            return false;
        }
        int ba = bd(a), bb = bd(b);
        if (cslot_of(ba) != cslot_of(bb)) {
            return false;
        }
        int dd = d_of(bb) - d_of(ba);
        return dd == 1 || dd == -1;
    }
    // check_pair_stacked (origami_potential.cpp:70-83)
    DEV bool pair_stacked(int a, int b) const {
        if (d_of(a) > d_of(b)) {
            int t = a;
            a = b;
            b = t;
        }
        I3 ndr = sub3(pos(b), pos(a));
        I3 oa = ore(a);
        if (!eq3(ndr, oa) && !eq3(ndr, neg3(oa))) {
            return twist_ok(a, ndr, b);
        }
        return false;
    }

    // ------------------------------------------------------------ FourBody
    DEVIN void fb_single(int h1, int h2, int h3);
    DEVIN void fb_double(int h1, int h2, int h3);
    DEVIN void fb_triply(int h1, int h2, int h3);
    DEVNI void fb_junction(int j1, int j2, int j3, int j4, int k1, int k2);
    DEVIN void fb_junction_pair(int a, int b, int j);
    DEVIN void fb_backward_triplets(int a, int b);
    DEVIN void fb_forward_triplets(int a, int b);
    DEVIN void fb_central_triplets(int di, int dj);
    DEVIN void fb_backward_single_junction(int a, int b);
    DEVIN void fb_backward_outer(int j3, int j4, int k2, int k1);
    DEVIN void fb_forward_outer(int j1, int j2, int k1, int k2);
    DEVIN void fb_central_outer(int j1, int j2, int k1, int k2, int k2b);
    DEVIN void fb_forward_single_junction(int a, int b);
    DEVIN void fb_central_single_junction(int a, int b);
    DEVIN void fb_regular_pair(int a, int b, int i);
    DEVIN void fb_helix_pair(int a, int b, int i, int j);
    DEVIN void fb_around(int cd, int j);
    DEVNI void fb_terms(int di, int dj);
    DEVIN Delta bind_domain(int di);

    // ------------------------------------------------------------ system
    DEVIN void update_occupancies(int s, I3 p);
    DEVIN Delta internal_check(int s, I3 p, int o);
    DEVIN Delta internal_unassign(int s, bool need_delta);
    DEVNI double check_domain_constraints(int s, I3 p, int o);
    DEVNI double unassign_domain(int s);
    DEVNI double set_checked_domain_config(int s, I3 p, int o);
    DEVNI double set_domain_config(int s, I3 p, int o);
    DEV void set_domain_orientation(int s, int o) {
        if (dom[s].state == OCC_BOUND) {
            S->violated = 1;
        }
        else {
            dom[s].ore = (int8_t)o;
        }
    }
    DEVNI int add_chain_with(int ident, int uid);
    DEVNI int add_chain(int ident);
    DEVNI void delete_chain(int cs);
    DEVNI void center();
    DEVNI void set_all_domains();
    DEVNI void check_all_constraints();
    DEVNI void update_energy();
    DEVNI void update_enthalpy_and_entropy();

    // ------------------------------------------------------------ moves
    DEVIN int select_random_domain();
    DEVIN bool test_acceptance(double p);
    DEVNI bool staple_is_connector(int cs);
    DEVNI bool scan_for_scaffold(int dom0, int excl_cs);
    DEVIN int num_bound_staple_domains(int cs);
    DEVNI void reset_origami();
    DEVIN double set_growth_point(int gnew, int gold);
    DEVNI void grow_staple(int cs, int d_i);
    DEVNI void grow_chain_seq(int cs, int from, int to);
    DEVNI bool orientation_move();
    DEVNI bool insert_staple();
    DEVNI bool delete_staple();
    DEVIN void cb_unassign_chain(int cs, int from, int to);
    DEVNI void cb_select_and_set(int d, int prev_d);
    DEVIN void cb_setup_regrow_old();
    DEVNI bool cb_test_acceptance();
    DEVIN void cb_set_old_growth_point(int gnew, int gold);
    DEVNI bool cb_staple_move();
    // constraint points
    DEVNI void cp_reset();
    DEVIN void cp_add_active(int s, I3 p, int sg);
    DEVIN void cp_remove_active(int s, bool rec = false);
    DEVIN void cp_remove_activated(int s);
    DEVIN void cp_update_endpoints(int s);
    DEVIN bool cp_endpoint_reached(int s, I3 p);
    DEVIN bool cp_walks_remain(int s, I3 p);
    DEVIN bool cp_walks_seg(int s, I3 p, int sg);
    DEVIN int cp_seg_dir(int cs, int sg);
    DEVIN bool cp_selected(int b);
    DEVNI void cp_save_initial();
    DEVNI void cp_scan_network(int d);
    DEVNI void cp_find(int first, int count, int dir, int sg);
    DEVIN bool ct_select_indices(int max_regrowth, int& first, int& count, int& dir);
    DEVNI void ctcb_grow_chain_scaffold(int first, int count, int dir, bool first_gp = true);
    DEVNI void ctcb_grow_staple_from(int gold);
    DEVNI bool ctcb_scaffold_move(int max_regrowth);
    DEVIN int rg_get_dir(int s);
    DEVNI double rg_unassign_and_save(bool save);
    DEVNI double rg_recoil_regrow();
    DEVNI double rg_set_config(int d, I3 p, int o, int rec);
    DEVIN void rg_prepare_for_growth(RgCur& c);
    DEVNI double rg_prepare_for_regrowth(RgCur& c, int rec);
    DEVNI void rg_restore_endpoints(int d);
    DEVIN void rg_select(RgCur& c, I3& p, int& o);
    DEVNI double rg_p_open(const RgCur& c, I3 p, int o);
    DEVIN bool rg_test_open(double p);
    DEVNI double rg_calc_weights(const PosOre* actual);
    DEVNI bool rg_test_avail(RgCur& c);
    DEVNI void rg_calc_old_c_opens();
    DEVIN void rg_reset_active(bool remove_first);
    DEVNI bool rg_regrow_and_accept(bool rm_new, bool rm_opens, bool rm_old);
    DEVNI bool ctrg_scaffold_move(int max_regrowth);
    DEVIN void jump_check_stems(int cur, int& tail);
    DEVNI bool jump_fill(int start, int have, int max_length, int seg_max, int dir, int& tail,
                         int& last);
    DEVNI bool jump_select(int max_regrowth);
    DEVNI bool ctrg_jump_move(int max_regrowth);
    DEVIN void ctcb_set_first_seg_domain(int stemd);
    DEVNI void ctcb_jump_grow();
    DEVNI bool ctcb_jump_move(int max_regrowth);
    DEVIN bool attempt(int m);
    DEVNI bool attempt_kind(int kind, int max_regrowth, int m = -1);
    DEVIN void track_move(int m, bool acc);
    // learned domain regrowth (lane 0 parts)
    DEVNI bool chain_connected(int cs);
    DEVNI int learned_begin(LearnedMove& m);
    DEVNI void learned_finish(LearnedMove& m);
    DEVNI int learned_continue(LearnedMove& m);
    DEVNI int learned_exchange_begin(LearnedMove& m);
    DEVNI int learned_exchange_continue(LearnedMove& m);
    DEVNI int learned_path(LearnedMove& m, int i);
    DEVNI int complement_of(int s);
    DEVNI int learned_end(LearnedMove& m, bool changed, bool accepted);
    DEVNI void learned_softmax(const LearnedMove& m, double* w, double& z);
    // channel of the learned proposal's input at site p, leaving domain
    // `excl` out: 0 empty; 1 + o for a scaffold domain, 7 + o for a staple
    // domain only; with two domains of one kind the smaller orientation index
    // (read-only: any lane may call it)
    DEV int site_channel(I3 p, int excl) const {
        int i = hfind(pkey(p));
        if (i < 0) {
            return 0;
        }
        int s = (int)(int16_t)(hval[i] & 0xFFFFu);
        int occ = (int)((hval[i] >> 16) & 3u);
        int c[2] = {s, (occ != OCC_UNBOUND && s >= 0) ? (int)dom[s].bound : (int)LDO_NONE};
        int scaf = 6, stap = 6;
        for (int k = 0; k < 2; k++) {
            int x = c[k];
            if (x < 0 || x == excl) {
                continue;
            }
            int o = dom[x].ore;
            if (cslot_of(x) == 0) {
                scaf = o < scaf ? o : scaf;
            }
            else {
                stap = o < stap ? o : stap;
            }
        }
        if (scaf < 6) {
            return 1 + scaf;
        }
        return stap < 6 ? 7 + stap : 0;
    }
    // domains at site p other than `excl` (0, 1 or 2)
    DEV int site_count(I3 p, int excl) const {
        int i = hfind(pkey(p));
        if (i < 0) {
            return 0;
        }
        int s = (int)(int16_t)(hval[i] & 0xFFFFu);
        int occ = (int)((hval[i] >> 16) & 3u);
        int n = (s >= 0 && s != excl) ? 1 : 0;
        if (occ != OCC_UNBOUND && s >= 0) {
            int b = dom[s].bound;
            n += (b >= 0 && b != excl) ? 1 : 0;
        }
        return n;
    }
    DEVNI void simulate(int64_t steps);
    DEVNI void begin_interval(int slot, bool recompute);
    DEVNI int window_op(int staple_offset);
    DEVNI void count_window_op();
    DEVNI double calc_move_bias(int staple_offset);

    // move-local scratch for the CTCB segment (kept in registers)
    int seg_first, seg_count, seg_dir;
};

#endif
