// layout.h - device data layout of the sampler (host + device).
//
// Every replica owns one fixed-size byte block ("replica block") in HBM that
// holds its whole lattice state and the scratch of the move in flight.  The
// block is built from structure-of-arrays sections at offsets computed once
// on the host (Layout) so the same code addresses it in HBM (one lane per
// replica) or after staging it into LDS (one wavefront per replica).
//
// Domains live in fixed slots: slots [0, ns) are the scaffold, staple chain
// slot k (1 <= k <= max_total_staples) owns slots ns + (k-1)*lmax .. + lmax.
// A lattice site is one entry of an open-addressed, linear-probing hash
// keyed by the packed position; its value holds the occupancy state and, for
// an unbound site, the slot of the domain on it.
#ifndef LDO_LAYOUT_H
#define LDO_LAYOUT_H

#include <stdint.h>

#include "../../include/ldo.h"

#define LDO_NONE ((int16_t)-1)
#define LDO_HASH_EMPTY 0xFFFFFFFFFFFFFFFFull
#define LDO_COORD_BITS 21
#define LDO_COORD_OFFSET (1 << 20)

enum { OCC_UNASSIGNED = 0, OCC_UNBOUND = 1, OCC_BOUND = 2, OCC_MISBOUND = 3 };

// Domain record, 16 bytes.
struct DomRec {
    int32_t x, y, z;
    int8_t ore;   // index into the six unit vectors, (+x,-x,+y,-y,+z,-z)
    int8_t state; // OCC_*
    int16_t bound; // slot of the bound partner or LDO_NONE
};

// Saved position/orientation (prev/old configuration), 16 bytes.
struct PosOre {
    int32_t x, y, z;
    int32_t ore;
};

struct ChainRec {
    int32_t uid;   // unique chain index (the reference's m_c)
    int16_t ident; // chain identity, -1 when the slot is free
    int16_t len;
};

// Active endpoint of the constraint-point machinery, 16 bytes.
struct Endpoint {
    int16_t cslot; // chain slot of the endpoint domain's chain
    int16_t d;     // domain index within the chain (scaffolds up to 30000 domains)
    int32_t seg;
    int32_t x, y, z;
};

struct NetFrame {
    int16_t growth; // growth domain slot of this recursion level
    int16_t idx;    // position in that staple's domain stack
};

// One level of the CTCB growth recursion (scaffold segment or staple grown
// from a growthpoint), kept on an explicit stack in the replica block.
struct GrowFrame {
    int16_t first, count, i, prev;
    int16_t gold, gnew;
    int8_t dir, kind, phase, pad;
};
#define LDO_GROW_DEPTH 32
// growth-point pairs of the staple a CB staple regrowth moves
#define LDO_MAX_STAPLE_LEN 8

#define LDO_MAX_MOVES 8

// non-contiguous scaffold selection: at most LDO_JUMP_MAX_SEGS segments
// (1 + 2 per stem), LDO_JUMP_MAX_SEGS / 2 stems; max_regrowth is capped so
// that these and the possible-stem queue (LDO_JUMP_QUEUE) always suffice
#define LDO_JUMP_MAX_SEGS 128
#define LDO_JUMP_QUEUE 512
#define LDO_JUMP_MAX_REGROWTH 60

// one segment of a non-contiguous selection: count domains from first in
// direction dir; last is the domain its end endpoint lies past (the stem for
// an empty fill)
struct JumpSeg {
    int16_t first, count, last;
    int8_t dir, pad;
};

// Per-replica scalars at the head of a replica block (kept 8-byte aligned).
struct RepScalars {
    int64_t step;
    double energy;
    double hyb_h;
    double hyb_s;
    double stack_e;
    double modifier;
    double delta_e;
    double bias;
    double new_bias;
    double new_modifier;
    int64_t tape_pos;
    uint64_t visited; // ladder slots whose tables this replica has built
    int64_t attempts[LDO_MAX_MOVES];
    int64_t accepts[LDO_MAX_MOVES];
    int64_t site_empty;
    int64_t site_bind;
    int64_t commits;
    int64_t learned_step; // learned-move attempts so far (their Philox counter)
    int32_t num_domains;
    int32_t num_staples;
    int32_t num_bound;
    int32_t num_fully;
    int32_t num_self;
    int32_t num_unassigned;
    int32_t num_stacked;
    int32_t current_c_i;
    int32_t n_chains; // working order length
    int32_t status;
    int32_t violated;
    int32_t temp_slot; // ladder slot whose tables are in use
    int32_t init_slot; // ladder slot the initiation terms come from
    int32_t n_free;
    int32_t n_modified;
    int32_t n_assigned;
    int32_t n_added;
    int32_t rejected;
    int32_t regrow_old;
    int32_t insertion_sites;
    int32_t kind;
    int32_t n_active;
    int32_t n_active0;
    int32_t rng_idx;
    int32_t added[4];
    int32_t rng_sub; // 0: MC step stream; 1: learned-move stream
    // the move in flight's tracking key (the movetypes' m_tracker): staple
    // type, insertion or deletion, no staples to regrow, CTCB segment length
    // and staples regrown
    int32_t trk_type;
    int32_t trk_ins;
    int32_t trk_nost;
    int32_t trk_len;
    int32_t trk_nst;
    int32_t last_move; // move slot of the last step and whether it was accepted
    int32_t last_acc;
    // CTRG: regrowth stack length, erased-endpoint stack height (positions)
    // and depth (levels)
    int32_t n_rgs;
    int32_t er_top;
    int32_t er_depth;
    // non-contiguous selection: segments (0 for a contiguous move), stems,
    // selected scaffold domains
    int32_t n_jsegs;
    int32_t n_jstems;
    int32_t n_jdom;
    int32_t pad_rg;
};

// Byte offsets of the sections of a replica block, and sizes.
struct Layout {
    int32_t ns;        // scaffold domains
    int32_t lmax;      // longest staple
    int32_t lshift;    // log2 of the slot stride of a staple chain
    int32_t smax;      // max_total_staples
    int32_t nslots;    // ns + smax << lshift
    int32_t nchain;    // 1 + smax chain slots
    int32_t ntypes;    // chain identities (scaffold + staple types)
    int32_t tcap;      // per-type list capacity
    int32_t hcap;      // hash capacity (power of two)
    int32_t hbits;
    int32_t ecap;      // endpoint list capacity
    int32_t track_len; // per-move tracking counters (see off_track)
    int64_t off_dom, off_dident, off_chain, off_order, off_free, off_tlist, off_tcount;
    int64_t off_hkey, off_hval, off_modified, off_assigned, off_prev, off_old;
    int64_t off_seg, off_gp, off_stem, off_inact, off_netseg, off_netds_flag;
    int64_t off_checked, off_netcs, off_regrow, off_dir0, off_dir1;
    int64_t off_active, off_active0, off_gps, off_iaes, off_netds, off_frames;
    int64_t off_part, off_sframes, off_gstack, off_gpl;
    // CTRG recoil growth (empty without a CTRG move): regrowth stack, per
    // level attempts (growth, weights), configurations left as a 36-bit mask
    // (growth, weights) and open probability, erased-endpoint stack and counts
    int64_t off_rgs, off_rgatt, off_rgwatt, off_rgavail, off_rgwavail, off_rgopen;
    int64_t off_erstk, off_ercnt;
    // non-contiguous selection (empty without a jump move): selected
    // scaffold domains, possible-stem queue, segments, stems, the first
    // segment of each scaffold stem's pair, and segment directions
    int64_t off_jdom, off_jps, off_jseg, off_jstem, off_stemseg, off_sdir;
    // int64 tracking counters by staple type T = ntypes (index = identity)
    // and scaffold length / staple count: Met insertion attempts[T],
    // accepts[T], deletion attempts[T], accepts[T]; CB attempts[T],
    // accepts[T]; CTCB by segment length attempts[ns+1], accepts[ns+1]; by
    // staples regrown attempts[smax+1], accepts[smax+1]; CTRG by segment
    // length attempts[ns+1], accepts[ns+1]; CTRG jump attempts, accepts;
    // CTCB jump by staples regrown attempts[smax+1], accepts[smax+1]
    int64_t off_track;
    int64_t bytes; // block size (multiple of 256)
};

// Kernel-wide constants (one copy, passed by value to every launch).
struct Params {
    Layout L;
    int32_t n_replicas;
    int32_t domain_type;
    int32_t misbinding_pot;
    int32_t cyclic;
    int32_t n_moves;
    int32_t move_kind[LDO_MAX_MOVES];
    double move_cum[LDO_MAX_MOVES];
    int32_t move_maxreg[LDO_MAX_MOVES];
    int32_t move_recoils[LDO_MAX_MOVES];   // CTRG max_num_recoils
    int32_t move_cattempts[LDO_MAX_MOVES]; // CTRG max_c_attempts
    int32_t move_segreg[LDO_MAX_MOVES];    // CTRG jump max_seg_regrowth
    int32_t max_total_staples;
    int32_t max_type_staples;
    int32_t centering_freq;
    int32_t centering_domain;
    int32_t check_freq;
    int32_t rng_mode;
    uint64_t seed;
    int32_t ident_lo;
    int32_t n_ident;
    int32_t n_slots; // ladder slots with tables
    int32_t lanes_per_replica;
    int32_t flags;   // ldo_run_flags
    double fugacity; // reduced fugacity = staple_M (never updated)
    const int32_t* type_len;   // [ntypes]
    const int16_t* type_ident; // [ntypes * type_ident_stride] table index of each domain
    int32_t type_ident_stride;
    const double* hyb_e;       // [n_slots * n_ident^2]
    const double* hyb_h;
    const double* hyb_s;
    const double* stack;
    const double* init_e; // [n_slots]
    const double* init_h;
    const double* init_s;
    char* blocks;
    const uint64_t* tapes;    // concatenated tapes
    const int64_t* tape_off;  // [n_replicas]
    const int64_t* tape_len;  // [n_replicas]
    int32_t exchange_mode;    // ldo_exchange_mode
    int32_t pt_dim2;          // stacking multipliers of a 2D ladder grid (else 1)
    const int32_t* slot_key;  // [n_slots] table key class of each ladder slot
    int32_t grid_len;         // grid values per window
    const ldo_window* windows; // [n_slots] umbrella windows, or null
    const double* win_grid;    // [n_slots * grid_len]
    double* win_stored;        // [n_slots * 2] stored restraint and grid bias
    // visits of each window's order parameter after every step, or null:
    // [n_slots * (grid_len + 1)], the last bin of a window counting values
    // off its grid
    int64_t* hist;
    // learned proposal (ldo_mlp.h layout), or null
    const float* net;
    int64_t net_off[10]; // w1, emb, w1e, b1, w2, b2, w3, b3, w4, b4
    int32_t net_n_ident;
    int32_t learned_slot; // counters: [slot] learned moves, [slot - 1] CB fallbacks
    double learned_fallback;
};

// One learned-move attempt, shared by the lanes of a replica's wavefront:
// lane 0 fills d..k_old, the lanes fill ch/legal and the network's logits.
struct LearnedMove {
    int32_t mode; // 1: the network is needed
    int32_t d;    // moving staple end domain (the domain the network places)
    int32_t a;    // its chain neighbour (the anchor)
    int32_t id;   // identity index of d (embedding row)
    int32_t k_old; // output d occupies (scored), or -1 (sampled)
    int32_t ax, ay, az;
    int32_t kind;  // 1 domain regrowth, 2 staple insertion, 3 staple deletion
    int32_t stage; // growth path being scored (exchange)
    int32_t cs;    // staple chain slot (exchange)
    int32_t valid[2];
    double qsum;   // summed proposal probability of the staple's growth paths
    uint8_t legal[6];
    uint8_t ch[343];
    float h[2][256];
    float logits[36];
};

// move kinds held in RepScalars::kind (the grow_chain dispatch)
#define LDO_MOVE_KIND_ORIENT 0
#define LDO_MOVE_KIND_MET 1
#define LDO_MOVE_KIND_CB 2
#define LDO_MOVE_KIND_CTCB 3
#define LDO_MOVE_KIND_CTRG 4
#define LDO_MOVE_KIND_CTRG_JUMP 5

// status codes written into RepScalars::status
#define ST_OK 0
#define ST_CONSTRAINT 1
#define ST_MISUSE 2
#define ST_CAPACITY 7
#define ST_TAPE 8
#define ST_COORD 9

#endif
