"""System descriptions: the reference's input files and the synthetic
BASELINE workloads, lowered to the C ABI structs of include/ldo.h.

A `System` holds what setup_origami() reads from the system JSON and the
.inp file (origami_system.cpp:991-1030, files.cpp:36-104, parser.cpp:27-449)
plus the moveset (files.cpp:264-280).  The generators build the synthetic
systems the benchmark configurations are defined on; their seeds are part of
the workload definition.
"""
import ctypes as C
import json
import math
from dataclasses import dataclass, field, replace
from fractions import Fraction

import numpy as np

from . import abi

# kcal/mol -> K (for H) and kcal/mol/K -> k_B (for S): 4184 / R
KCAL_PER_MOL_TO_K = 4184.0 / 8.3144598

UNIT_VECTORS = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]


@dataclass
class Move:
    kind: int
    freq: float
    max_regrowth: int = 0
    # CTRG recoil growth: recoils allowed and trial configurations per domain
    max_num_recoils: int = 0
    max_c_attempts: int = 0
    max_seg_regrowth: int = 0  # CTRG jump
    label: str = None          # the moveset file's label (None: the type name)


@dataclass
class Chain:
    index: int
    identity: int
    positions: list
    orientations: list


@dataclass
class System:
    identities: list
    sequences: list = None
    cyclic: bool = False
    domain_type: int = abi.HALFTURN
    hybridization_pot: int = abi.HYB_NEARESTNEIGHBOUR
    misbinding_pot: int = abi.MISBIND_OPPOSING
    stacking_pot: int = abi.STACK_CONSTANT
    binding_h: float = 0.0
    binding_s: float = 0.0
    misbinding_h: float = 0.0
    misbinding_s: float = 0.0
    stacking_ene: float = 1.0
    cation_M: float = 1.0
    staple_M: float = 1.0
    enthalpies: list = None
    entropies: list = None
    chains: list = field(default_factory=list)
    moves: list = field(default_factory=list)
    max_total_staples: int = 999
    max_type_staples: int = 999
    max_staple_size: int = 2
    centering_freq: int = 0
    centering_domain: int = 0
    constraint_check_freq: int = 0
    seed: int = 0
    name: str = ""
    temps: list = None
    flags: int = 0  # ldo_run_flags (abi.FLAG_*); 0 reproduces the reference

    # ------------------------------------------------------------- sizes
    @property
    def n_scaffold(self):
        return len(self.identities[0])

    def max_domains(self):
        """Largest number of domains a replica can hold."""
        longest = max((len(c) for c in self.identities[1:]), default=0)
        return self.n_scaffold + self.max_total_staples * longest

    # ------------------------------------------------------------- lowering
    def to_abi(self, rng_mode=abi.RNG_PHILOX, exchange_mode=abi.EXCHANGE_NONE,
               lanes_per_replica=0):
        """Return (SystemDesc, RunParams, Config, keepalive)."""
        keep = []

        def arr(ctype, values):
            a = (ctype * max(1, len(values)))(*values)
            keep.append(a)
            return a

        lengths = [len(c) for c in self.identities]
        flat = [v for c in self.identities for v in c]
        sd = abi.SystemDesc()
        sd.n_chain_types = len(self.identities)
        sd.chain_lengths = arr(C.c_int32, lengths)
        sd.identities = arr(C.c_int32, flat)
        if self.sequences is not None:
            seqs = [s.encode() for c in self.sequences for s in c]
            sd.sequences = arr(C.c_char_p, seqs)
        if self.enthalpies is not None:
            sd.enthalpies = arr(C.c_double, self.enthalpies)
            sd.entropies = arr(C.c_double, self.entropies)
            sd.n_specified = len(self.enthalpies)
        sd.cyclic = int(self.cyclic)
        sd.domain_type = self.domain_type
        sd.hybridization_pot = self.hybridization_pot
        sd.misbinding_pot = self.misbinding_pot
        sd.stacking_pot = self.stacking_pot
        sd.binding_h = self.binding_h
        sd.binding_s = self.binding_s
        sd.misbinding_h = self.misbinding_h
        sd.misbinding_s = self.misbinding_s
        sd.stacking_ene = self.stacking_ene
        sd.cation_M = self.cation_M
        sd.staple_M = self.staple_M

        rp = abi.RunParams()
        rp.n_movetypes = len(self.moves)
        rp.movetype_kind = arr(C.c_int32, [m.kind for m in self.moves])
        rp.movetype_freq = arr(C.c_double, [m.freq for m in self.moves])
        rp.movetype_max_regrowth = arr(C.c_int32, [m.max_regrowth for m in self.moves])
        rp.movetype_max_num_recoils = arr(C.c_int32, [m.max_num_recoils for m in self.moves])
        rp.movetype_max_c_attempts = arr(C.c_int32, [m.max_c_attempts for m in self.moves])
        rp.movetype_max_seg_regrowth = arr(C.c_int32, [m.max_seg_regrowth for m in self.moves])
        rp.max_total_staples = self.max_total_staples
        rp.max_type_staples = self.max_type_staples
        rp.max_staple_size = self.max_staple_size
        rp.centering_freq = self.centering_freq
        rp.centering_domain = self.centering_domain
        rp.constraint_check_freq = self.constraint_check_freq
        rp.seed = self.seed
        rp.rng_mode = rng_mode
        rp.exchange_mode = exchange_mode
        rp.lanes_per_replica = lanes_per_replica
        rp.flags = self.flags

        cfg = config_to_abi(self.chains, keep)
        return sd, rp, cfg, keep


def config_to_abi(chains, keep):
    def arr(ctype, values):
        a = (ctype * max(1, len(values)))(*values)
        keep.append(a)
        return a

    cfg = abi.Config()
    cfg.n_chains = len(chains)
    cfg.chain_index = arr(C.c_int32, [c.index for c in chains])
    cfg.chain_identity = arr(C.c_int32, [c.identity for c in chains])
    cfg.chain_length = arr(C.c_int32, [len(c.positions) for c in chains])
    cfg.positions = arr(C.c_int32, [v for c in chains for p in c.positions for v in p])
    cfg.orientations = arr(C.c_int32, [v for c in chains for p in c.orientations for v in p])
    return cfg


def empty_config(cap_chains, cap_domains):
    keep = []
    cfg = abi.Config()
    for name, n in (("chain_index", cap_chains), ("chain_identity", cap_chains),
                    ("chain_length", cap_chains), ("positions", 3 * cap_domains),
                    ("orientations", 3 * cap_domains)):
        a = (C.c_int32 * n)()
        keep.append(a)
        setattr(cfg, name, a)
    return cfg, keep


def config_from_abi(cfg):
    chains = []
    k = 0
    for c in range(cfg.n_chains):
        n = cfg.chain_length[c]
        pos = [tuple(cfg.positions[3 * (k + d): 3 * (k + d) + 3]) for d in range(n)]
        ore = [tuple(cfg.orientations[3 * (k + d): 3 * (k + d) + 3]) for d in range(n)]
        chains.append(Chain(cfg.chain_index[c], cfg.chain_identity[c], pos, ore))
        k += n
    return chains


# ----------------------------------------------------------- reference files
def parse_fraction(text):
    """utility::Fraction (utility.cpp:266-282): "a/b" or a real number."""
    text = str(text)
    if "/" in text:
        num, den = text.split("/", 1)
        return float(num) / float(den)
    return float(text)


def read_system_json(path_or_dict):
    """OrigamiInputFile::read_file (files.cpp:36-104)."""
    if isinstance(path_or_dict, dict):
        root = path_or_dict
    else:
        with open(path_or_dict) as f:
            root = json.load(f)
    o = root["origami"]
    chains = []
    for ch in o["configurations"][0]["chains"]:
        chains.append(Chain(int(ch["index"]), int(ch["identity"]),
                            [tuple(p) for p in ch["positions"]],
                            [tuple(p) for p in ch["orientations"]]))
    return dict(identities=o["identities"], sequences=o.get("sequences"),
                enthalpies=o.get("enthalpies"), entropies=o.get("entropies"),
                cyclic=bool(o.get("cyclic", False)), chains=chains)


def read_moveset_json(path_or_dict):
    """OrigamiMovetypeFile (files.cpp:264-314)."""
    if isinstance(path_or_dict, dict):
        root = path_or_dict
    else:
        with open(path_or_dict) as f:
            root = json.load(f)
    moves = []
    for m in root["origami"]["movetypes"]:
        if m["type"] not in abi.MOVETYPE_NAMES:
            raise NotImplementedError(f"{m['type']}: movetype not supported")
        moves.append(Move(abi.MOVETYPE_NAMES[m["type"]], parse_fraction(m["freq"]),
                          int(m.get("max_regrowth", 0)), int(m.get("max_num_recoils", 0)),
                          int(m.get("max_c_attempts", 0)), int(m.get("max_seg_regrowth", 0)),
                          m.get("label")))
    return moves


# ------------------------------------------------------- synthetic workloads
def standard_moveset(max_regrowth):
    """Orientation 1/10, staple exchange 4/10, CB staple 3/10, CTCB 2/10."""
    return [Move(abi.MOVE_ORIENTATION_ROTATION, 1 / 10),
            Move(abi.MOVE_MET_STAPLE_EXCHANGE, 4 / 10),
            Move(abi.MOVE_CB_STAPLE_REGROWTH, 3 / 10),
            Move(abi.MOVE_CTCB_SCAFFOLD_REGROWTH, 2 / 10, max_regrowth)]


def ctrg_moveset(max_regrowth, max_num_recoils=1, max_c_attempts=36):
    """The reference's examples/moveset_standard.json with its contiguous
    CTRG scaffold regrowth taking the non-contiguous one's share: orientation
    2/6, staple exchange 1/6, CB staple 1/6, CTRG 2/6."""
    return [Move(abi.MOVE_ORIENTATION_ROTATION, 2 / 6),
            Move(abi.MOVE_MET_STAPLE_EXCHANGE, 1 / 6),
            Move(abi.MOVE_CB_STAPLE_REGROWTH, 1 / 6),
            Move(abi.MOVE_CTRG_SCAFFOLD_REGROWTH, 2 / 6, max_regrowth, max_num_recoils,
                 max_c_attempts)]


def jump_moveset(max_regrowth, max_num_recoils=1, max_c_attempts=36, max_seg_regrowth=2):
    """The reference's examples/moveset_standard.json: orientation 2/6,
    staple exchange 1/6, CB staple 1/6, contiguous and non-contiguous CTRG
    1/6 each."""
    return [Move(abi.MOVE_ORIENTATION_ROTATION, 2 / 6),
            Move(abi.MOVE_MET_STAPLE_EXCHANGE, 1 / 6),
            Move(abi.MOVE_CB_STAPLE_REGROWTH, 1 / 6),
            Move(abi.MOVE_CTRG_SCAFFOLD_REGROWTH, 1 / 6, max_regrowth, max_num_recoils,
                 max_c_attempts),
            Move(abi.MOVE_CTRG_JUMP_SCAFFOLD_REGROWTH, 1 / 6, max_regrowth, max_num_recoils,
                 max_c_attempts, max_seg_regrowth)]


def four_domain_system(seed=20261202, max_total_staples=4):
    """The enumerable plain-method workload: a 4-domain HalfTurn scaffold,
    staple types [1,2] and [3,4], Uniform hybridization with
    dH = -120 kcal/mol and dS = -0.33 kcal/mol/K, Opposing misbinding,
    constant stacking, 1 uM staples at 330 K."""
    return System(
        identities=[[-1, -2, -3, -4], [1, 2], [3, 4]],
        cyclic=False,
        domain_type=abi.HALFTURN,
        hybridization_pot=abi.HYB_UNIFORM,
        misbinding_pot=abi.MISBIND_OPPOSING,
        binding_h=-120.0 * KCAL_PER_MOL_TO_K,
        binding_s=-0.33 * KCAL_PER_MOL_TO_K,
        stacking_ene=-1000.0,
        cation_M=1.0,
        staple_M=1e-6,
        chains=[Chain(0, 0, [(i, 0, 0) for i in range(4)], [(0, 0, 1)] * 4)],
        moves=standard_moveset(4),
        max_total_staples=max_total_staples,
        max_type_staples=4,
        max_staple_size=2,
        centering_freq=10000,
        constraint_check_freq=10000,
        seed=seed,
        name="four-domain HalfTurn scaffold, 2 two-domain staple types, Uniform, 330 K",
        temps=[330.0],
    )


def hairpin_system(n_domains=4, seed=20261202, max_total_staples=None):
    """A scaffold of n_domains (even) on a line with crossover staples
    [k+1, n-k] that fold it into a hairpin: each staple binds two scaffold
    domains far apart along the chain, so the non-contiguous scaffold
    selection finds stem domains.  Energies and concentrations as the
    four-domain workload; at most one staple of each type."""
    n = n_domains
    staples = [[k + 1, n - k] for k in range(n // 2)]
    s = four_domain_system(seed=seed, max_total_staples=max_total_staples or n // 2)
    s.identities = [[-(i + 1) for i in range(n)]] + staples
    s.chains = [Chain(0, 0, [(i, 0, 0) for i in range(n)], [(0, 0, 1)] * n)]
    s.max_type_staples = 1
    s.name = f"{n}-domain hairpin scaffold, {len(staples)} crossover staple types, Uniform"
    return s


def three_domain_system(domain_type=abi.HALFTURN, misbinding_pot=abi.MISBIND_OPPOSING,
                        max_total_staples=1, seed=20261202, flags=0):
    """An enumerable system with one three-domain staple type: a 3-domain
    scaffold [-1,-2,-3] and staple [1,2,3], energies and moveset as the
    four-domain workload (CTCB max_regrowth 3)."""
    return System(
        identities=[[-1, -2, -3], [1, 2, 3]],
        cyclic=False,
        domain_type=domain_type,
        hybridization_pot=abi.HYB_UNIFORM,
        misbinding_pot=misbinding_pot,
        binding_h=-120.0 * KCAL_PER_MOL_TO_K,
        binding_s=-0.33 * KCAL_PER_MOL_TO_K,
        stacking_ene=-1000.0,
        cation_M=1.0,
        staple_M=1e-6,
        chains=[Chain(0, 0, [(i, 0, 0) for i in range(3)], [(0, 0, 1)] * 3)],
        moves=standard_moveset(3),
        max_total_staples=max_total_staples,
        max_type_staples=max_total_staples,
        max_staple_size=3,
        centering_freq=10000,
        constraint_check_freq=10000,
        seed=seed,
        name="three-domain scaffold, one three-domain staple type, Uniform",
        temps=[349.0],
        flags=flags,
    )


def raster_position(i, width):
    row, col = divmod(i, width)
    if row % 2 == 1:
        col = width - 1 - col
    return (col, row, 0)


def domino_tiling(rows, cols, flips, seed):
    """Seeded random domino tiling of a rows x cols grid: start from vertical
    dominoes and apply `flips` random 2x2 rotations."""
    partner = {}
    for r in range(0, rows - 1, 2):
        for c in range(cols):
            partner[(r, c)] = (r + 1, c)
            partner[(r + 1, c)] = (r, c)
    rng = np.random.default_rng(seed)
    picks = rng.integers(0, [rows - 1, cols - 1], size=(flips, 2))
    for r, c in picks:
        a, b, d, e = (r, c), (r, c + 1), (r + 1, c), (r + 1, c + 1)
        if partner.get(a) == d and partner.get(b) == e:
            partner[a], partner[b] = b, a
            partner[d], partner[e] = e, d
        elif partner.get(a) == b and partner.get(d) == e:
            partner[a], partner[d] = d, a
            partner[b], partner[e] = e, b
    pairs = sorted({tuple(sorted((k, v))) for k, v in partner.items()})
    return pairs


def _revcomp(s):
    return s[::-1].translate(str.maketrans("ACGT", "TGCA"))


def random_sequences(n, length, seed):
    rng = np.random.default_rng(seed)
    while True:
        seqs = ["".join(rng.choice(list("ACGT"), size=length)) for _ in range(n)]
        rcs = {_revcomp(s) for s in seqs}
        if len(set(seqs)) == n and not (rcs & set(seqs)):
            return seqs


def raster_system(n_domains=64, width=8, replicas=32, t_lo=300.0, t_hi=360.0,
                  seed=20261202, tiling_seed=1, sequence_seed=2):
    """The PTMC workload: a 64-domain HalfTurn scaffold on an 8x8
    boustrophedon raster in z=0, 32 two-domain staple types from a seeded
    domino tiling, random 16-nt sequences, NearestNeighbour energies, 0.5 M
    cation, 1 uM staples, 32 replicas at 300-360 K."""
    rows = n_domains // width
    index_of = {}
    for i in range(n_domains):
        col, row, _ = raster_position(i, width)
        index_of[(row, col)] = i
    pairs = domino_tiling(rows, width, 10000, tiling_seed)
    staples = []
    for a, b in pairs:
        i, j = sorted((index_of[a], index_of[b]))
        staples.append([i + 1, j + 1])
    staples.sort()
    seqs = random_sequences(n_domains, 16, sequence_seed)
    seq_of = {-(i + 1): s for i, s in enumerate(seqs)}
    identities = [[-(i + 1) for i in range(n_domains)]] + staples
    sequences = [seqs] + [[_revcomp(seq_of[-v]) for v in st] for st in staples]
    temps = [t_lo + (t_hi - t_lo) * k / max(1, replicas - 1) for k in range(replicas)]
    sysd = System(
        identities=identities,
        sequences=sequences,
        cyclic=False,
        domain_type=abi.HALFTURN,
        hybridization_pot=abi.HYB_NEARESTNEIGHBOUR,
        misbinding_pot=abi.MISBIND_OPPOSING,
        stacking_ene=-1000.0,
        cation_M=0.5,
        staple_M=1e-6,
        chains=[Chain(0, 0, [raster_position(i, width) for i in range(n_domains)],
                      [(0, 0, 1)] * n_domains)],
        moves=standard_moveset(12),
        max_total_staples=n_domains,
        max_type_staples=2,
        max_staple_size=2,
        centering_freq=0,
        constraint_check_freq=0,
        seed=seed,
        name=f"{n_domains}-domain HalfTurn raster scaffold, {len(staples)} staple types, NN",
        temps=temps,
    )
    return sysd


def layered_position(i, width, layers_cells):
    """Position of scaffold domain i on a stack of width x width boustrophedon
    layers: even layers run the raster forwards, odd layers backwards, so the
    path steps +z between layers."""
    layer, k = divmod(i, layers_cells)
    if layer % 2 == 1:
        k = layers_cells - 1 - k
    col, row, _ = raster_position(k, width)
    return (col, row, layer)


def _is_path3(cells):
    """Three lattice cells forming a path a-b-c: returns (a, b, c) or None."""
    def adj(u, v):
        return sum(abs(x - y) for x, y in zip(u, v)) == 1
    for mid in cells:
        ends = [c for c in cells if c != mid]
        if adj(mid, ends[0]) and adj(mid, ends[1]):
            return ends[0], mid, ends[1]
    return None


def tromino_tiling(path, flips, seed):
    """Seeded random tiling of the cells of `path` (a lattice path) into
    paths of three cells, plus one pair when len(path) % 3 == 2: start from
    consecutive triples along the path and apply `flips` random retilings of
    two adjacent triples (a uniform choice among the ways to split their six
    cells into two paths of three).  Returns a list of cell tuples."""
    import itertools
    rng = np.random.default_rng(seed)
    n = len(path)
    pieces = [tuple(path[i:i + 3]) for i in range(0, n - n % 3, 3)]
    rest = tuple(path[n - n % 3:])
    owner = {}
    for k, pc in enumerate(pieces):
        for c in pc:
            owner[c] = k
    steps = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    for _ in range(flips):
        a = int(rng.integers(len(pieces)))
        c = pieces[a][int(rng.integers(3))]
        d = steps[int(rng.integers(6))]
        nb = (c[0] + d[0], c[1] + d[1], c[2] + d[2])
        b = owner.get(nb)
        if b is None or b == a:
            continue
        cells = sorted(pieces[a] + pieces[b])
        splits = []
        for trio in itertools.combinations(cells, 3):
            if cells[0] not in trio:
                continue
            other = tuple(x for x in cells if x not in trio)
            p1, p2 = _is_path3(list(trio)), _is_path3(list(other))
            if p1 and p2:
                splits.append((p1, p2))
        p1, p2 = splits[int(rng.integers(len(splits)))]
        pieces[a], pieces[b] = p1, p2
        for x in p1:
            owner[x] = a
        for x in p2:
            owner[x] = b
    return pieces + ([rest] if rest else [])


def mwus3d_system(width=5, layers=8, windows=64, temps=(310.0, 315.0, 320.0, 325.0),
                  seed=20261202, tiling_seed=1, sequence_seed=2,
                  flags=abi.FLAG_CTCB_SCAFFOLD_ENDPOINTS):
    """The large 3D umbrella workload: a 200-domain ThreeQuarterTurn scaffold
    on a 5x5 raster x 8 layers, 66 three-domain staple types and one
    two-domain type from a seeded tromino tiling of the scaffold's cells
    (every staple binds lattice-adjacent scaffold domains, its middle domain
    adjacent to both ends), random 8-nt sequences, NearestNeighbour energies,
    0.5 M cation, 1 uM staples, at most 2 staples per type.  Replicas:
    `windows` numfulldomains windows at each of the temperatures, ordered
    temperature-major, so each temperature's windows form one ladder.
    CTCB constrains the in-segment scaffold domains of externally bound
    staple networks (abi.FLAG_CTCB_SCAFFOLD_ENDPOINTS) by default: under the
    reference's rule a three-domain staple bound across a segment boundary
    is left dangling, its site then blocks the old-configuration regrowth,
    and the rejected move leaves scaffold domains unassigned (the run stops
    with a constraint error within a few thousand steps).
    Returns (system, window list, grid values, grid length)."""
    cells = width * width
    n = cells * layers
    path = [layered_position(i, width, cells) for i in range(n)]
    index_of = {p: i for i, p in enumerate(path)}
    pieces = tromino_tiling(path, 20 * n, tiling_seed)
    staples = []
    for pc in pieces:
        ids = [index_of[c] + 1 for c in pc]
        if ids[-1] < ids[0]:
            ids.reverse()
        staples.append(ids)
    staples.sort()
    seqs = random_sequences(n, 8, sequence_seed)
    identities = [[-(i + 1) for i in range(n)]] + staples
    sequences = [seqs] + [[_revcomp(seqs[v - 1]) for v in st] for st in staples]
    all_temps = [t for t in temps for _ in range(windows)]
    sysd = System(
        identities=identities,
        sequences=sequences,
        cyclic=False,
        domain_type=abi.THREEQUARTERTURN,
        hybridization_pot=abi.HYB_NEARESTNEIGHBOUR,
        misbinding_pot=abi.MISBIND_OPPOSING,
        stacking_ene=-1000.0,
        cation_M=0.5,
        staple_M=1e-6,
        chains=[Chain(0, 0, path, [(0, 0, 1)] * n)],
        moves=standard_moveset(12),
        max_total_staples=2 * len(staples),
        max_type_staples=2,
        max_staple_size=3,
        centering_freq=0,
        constraint_check_freq=0,
        seed=seed,
        name=(f"{n}-domain ThreeQuarterTurn {width}x{width}x{layers} scaffold, "
              f"{len(staples)} staple types, NN"),
        temps=all_temps,
        flags=flags,
    )
    # LinearStepWell windows tiling numfulldomains 0..n with one point of
    # overlap on each side, 1 kT per point outside the well, no grid bias
    wins = []
    for w in range(windows):
        lo = max(0, (w * n) // windows - 1)
        hi = min(n, ((w + 1) * n) // windows + 1)
        wins.append(abi.Window(abi.OP_NUMFULLDOMAINS, lo, hi, 0, 0.0, 0.0, 1.0, 0.0, 1.0))
    wins = wins * len(temps)
    return sysd, wins, [], 0


def batch_chain_system(n_domains, width=8, seed=20261202, tiling_seed=1):
    """One chain of the batching workload: an `n_domains` HalfTurn scaffold on
    a width-8 boustrophedon raster in z=0 and its two-domain staple types.
    The staples come from the domino rule of the raster workload on the
    enclosing full rows (dominoes with both cells on the scaffold), then
    consecutive scaffold domains left unpaired are paired along the path, so
    every type binds two lattice-adjacent scaffold domains.  Uniform energies
    as the four-domain workload, 1 uM staples."""
    rows = -(-n_domains // width)
    index_of = {}
    for i in range(rows * width):
        col, row, _ = raster_position(i, width)
        index_of[(row, col)] = i
    pairs = domino_tiling(rows + (rows % 2), width, 10 * rows * width, tiling_seed)
    used = set()
    staples = []
    for a, b in pairs:
        i, j = index_of.get(a, n_domains), index_of.get(b, n_domains)
        if i < n_domains and j < n_domains:
            i, j = sorted((i, j))
            staples.append([i + 1, j + 1])
            used.update((i, j))
    i = 0
    while i + 1 < n_domains:
        if i not in used and i + 1 not in used:
            staples.append([i + 1, i + 2])
            used.update((i, i + 1))
            i += 2
        else:
            i += 1
    staples.sort()
    n_types = len(staples)
    return System(
        identities=[[-(i + 1) for i in range(n_domains)]] + staples,
        cyclic=False,
        domain_type=abi.HALFTURN,
        hybridization_pot=abi.HYB_UNIFORM,
        misbinding_pot=abi.MISBIND_OPPOSING,
        binding_h=-120.0 * KCAL_PER_MOL_TO_K,
        binding_s=-0.33 * KCAL_PER_MOL_TO_K,
        stacking_ene=-1000.0,
        cation_M=1.0,
        staple_M=1e-6,
        chains=[Chain(0, 0, [raster_position(i, width) for i in range(n_domains)],
                      [(0, 0, 1)] * n_domains)],
        moves=standard_moveset(min(12, n_domains)),
        max_total_staples=2 * n_types,
        max_type_staples=2,
        max_staple_size=2,
        centering_freq=0,
        constraint_check_freq=0,
        seed=seed,
        name=f"{n_domains}-domain HalfTurn raster chain, {n_types} staple types, Uniform",
        temps=[330.0],
    )


def batch_systems(n_chains=4096, len_lo=8, len_hi=128, t_lo=310.0, t_hi=350.0, length_seed=3,
                  temp_seed=4, seed=20261202):
    """The batching workload: `n_chains` independent constant-T chains with
    scaffold lengths L ~ U{8..128} and temperatures T ~ U(310, 350) K; chain
    k has seed `seed + k`.  Returns (systems, temperatures)."""
    lengths = np.random.default_rng(length_seed).integers(len_lo, len_hi + 1, size=n_chains)
    temps = np.random.default_rng(temp_seed).uniform(t_lo, t_hi, size=n_chains)
    cache = {}
    systems = []
    for k, L in enumerate(lengths):
        L = int(L)
        if L not in cache:
            cache[L] = batch_chain_system(L)
        s = replace(cache[L], seed=seed + k, temps=[float(temps[k])])
        systems.append(s)
    return systems, [float(t) for t in temps]
