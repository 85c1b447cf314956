"""ldo_run: the reference's input files in, its output files out, every
replica on the GPU.  The expected lines are built from the CPU oracle run on
the same system and seed, in the reference's formats (files.cpp:493-791,
ptmc_simulation.cpp:91-104, 315-322)."""
import json
import sys
import os
import shutil
import subprocess

import pytest

from latticednaorigami_amd import abi, systems
from latticednaorigami_amd.build import DRIVER

MOVE_TYPES = {abi.MOVE_ORIENTATION_ROTATION: "OrientationRotation",
              abi.MOVE_MET_STAPLE_EXCHANGE: "MetStapleExchange",
              abi.MOVE_CB_STAPLE_REGROWTH: "CBStapleRegrowth",
              abi.MOVE_CTCB_SCAFFOLD_REGROWTH: "CTCBScaffoldRegrowth",
              abi.MOVE_CTRG_SCAFFOLD_REGROWTH: "CTRGScaffoldRegrowth",
              abi.MOVE_CTRG_JUMP_SCAFFOLD_REGROWTH: "CTRGJumpScaffoldRegrowth",
              abi.MOVE_CTCB_JUMP_SCAFFOLD_REGROWTH: "CTCBJumpScaffoldRegrowth"}
HYB = {abi.HYB_NEARESTNEIGHBOUR: "NearestNeighbour", abi.HYB_UNIFORM: "Uniform"}
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# the reference's examples/moveset_standard.json
MOVESET_STANDARD = os.path.join(GOLDEN, "moveset_standard.json")


def label_of(m):
    return m.label or MOVE_TYPES[m.kind]


def write_inputs(d, s, extra):
    """The reference's three input files for system `s`; returns the .inp path."""
    sysfile = {"origami": {
        "cyclic": s.cyclic,
        "identities": s.identities,
        "configurations": [{"step": 0, "chains": [
            {"index": c.index, "identity": c.identity, "positions": [list(p) for p in c.positions],
             "orientations": [list(o) for o in c.orientations]} for c in s.chains]}]}}
    if s.sequences is not None:
        sysfile["origami"]["sequences"] = s.sequences
    with open(os.path.join(d, "system.json"), "w") as f:
        json.dump(sysfile, f)
    def entry(m):
        e = {"label": label_of(m), "type": MOVE_TYPES[m.kind], "freq": m.freq,
             "max_regrowth": m.max_regrowth}
        if m.kind in (abi.MOVE_CTRG_SCAFFOLD_REGROWTH, abi.MOVE_CTRG_JUMP_SCAFFOLD_REGROWTH):
            e["max_num_recoils"] = m.max_num_recoils
            e["max_c_attempts"] = m.max_c_attempts
        if m.kind in (abi.MOVE_CTRG_JUMP_SCAFFOLD_REGROWTH, abi.MOVE_CTCB_JUMP_SCAFFOLD_REGROWTH):
            e["max_seg_regrowth"] = m.max_seg_regrowth
        return e
    moves = {"origami": {"movetypes": [entry(m) for m in s.moves]}}
    with open(os.path.join(d, "moves.json"), "w") as f:
        json.dump(moves, f)
    lines = {
        "origami_input_filename": os.path.join(d, "system.json"),
        "movetype_file": os.path.join(d, "moves.json"),
        "domain_type": "HalfTurn" if s.domain_type == abi.HALFTURN else "ThreeQuarterTurn",
        "binding_pot": "FourBody", "misbinding_pot": "Opposing", "stacking_pot": "Constant",
        "hybridization_pot": HYB[s.hybridization_pot],
        "binding_h": repr(s.binding_h), "binding_s": repr(s.binding_s),
        "stacking_ene": repr(s.stacking_ene), "cation_M": repr(s.cation_M),
        "staple_M": repr(s.staple_M), "max_total_staples": s.max_total_staples,
        "max_type_staples": s.max_type_staples, "max_staple_size": s.max_staple_size,
        "centering_freq": s.centering_freq, "constraint_check_freq": s.constraint_check_freq,
        "random_seed": s.seed, "output_filebase": os.path.join(d, "out"),
        "ops_to_output": "numfulldomains nummisdomains numstackedpairs numstaples",
    }
    lines.update(extra)
    path = os.path.join(d, "run.inp")
    with open(path, "w") as f:
        f.write("# written by the test\n")
        for k, v in lines.items():
            f.write(f"{k}={v}\n")
    return path


def g10(x):
    """A double as std::ostream writes it with precision(10)."""
    return "%.10g" % x


def expect_lines(ref, r, s, step):
    """The .ene, .ops, .counts and .trj records of replica r at `step`."""
    ob = ref.observables(r)
    cfg = ref.config(r)
    ene = (f"{step} {g10(ob['energy'])} {g10(ob['enthalpy'])} {g10(ob['entropy'])} "
           f"{g10(ob['stacking_energy'])} {g10(ob['bias'])} ")
    ops = (f"{step} {ob['num_fully_bound_domain_pairs']} {ob['num_misbound_domain_pairs']} "
           f"{ob['num_stacked_domain_pairs']} {ob['num_staples']}")
    unique = len({c.identity for c in cfg[1:]})
    counts = (f"{step} {ob['num_staples']} {unique} {ob['num_bound_domain_pairs']} "
              f"{ob['num_fully_bound_domain_pairs']} {ob['num_misbound_domain_pairs']} ")
    trj = [str(step)]
    for c in cfg:
        trj.append(f"{c.index} {c.identity}")
        trj.append("".join(f"{v} " for p in c.positions for v in p))
        trj.append("".join(f"{v} " for o in c.orientations for v in o))
    trj.append("")
    return ene, ops, counts, trj


def expect_config_records(s, cfg, step):
    """The .vcf frame, .states and .ores lines and the .staples and
    .staplestates lines of a configuration (files.cpp:529-691).  Every domain
    of a saved configuration is assigned; two domains on one site are a bound
    pair (complementary identities) or a misbound one."""
    max_dom = s.max_staple_size * s.max_total_staples + len(s.identities[0])
    site = {}
    for ci, c in enumerate(cfg):
        for d, p in enumerate(c.positions):
            site.setdefault(tuple(p), []).append((ci, d, s.identities[c.identity][d]))
    state = {}
    for occ in site.values():
        if len(occ) == 1:
            state[occ[0][:2]] = 1
        else:
            a, b = occ
            state[a[:2]] = state[b[:2]] = 2 if a[2] == -b[2] else 3
    vcf, states, ores, n = ["timestep"], "", "", 0
    for ci, c in enumerate(cfg):
        k = 0
        for d, p in enumerate(c.positions):
            vcf.append("".join(f"{v} " for v in p))
            states += f"{state[(ci, d)]} "
            ores += "".join(f"{v} " for v in c.orientations[d])
            k += 1
        for _ in range(k, s.max_staple_size):
            vcf.append("0 0 0 ")
            states += "-1 "
            ores += "0 0 0 "
        n += max(k, s.max_staple_size)
    for _ in range(n, max_dom):
        vcf.append("0 0 0 ")
        states += "-1 "
        ores += "0 0 0 "
    counts = [0] * (len(s.identities) - 1)
    full = [0] * (len(s.identities) - 1)
    for ci, c in enumerate(cfg[1:], start=1):
        counts[c.identity - 1] += 1
        if all(state[(ci, d)] == 2 for d in range(len(c.positions))):
            full[c.identity - 1] = 1
    staples = f"{step} " + "".join(f"{x} " for x in counts)
    staplestates = f"{step} " + "".join(f"{x} " for x in full)
    return vcf + [""], states, ores, staples, staplestates


def gdef(x):
    """A double as std::ostream writes it by default (6 significant digits)."""
    return "%g" % x


def ratio(a, b, single=False):
    """a / b as the movetype summaries print it: a float or double quotient,
    and 0 / 0 as glibc prints the x86 default NaN."""
    import numpy as np
    if b == 0:
        return "-nan"
    return gdef(float(np.float32(a) / np.float32(b)) if single else a / b)


def expect_log_entry(ref, r, s, step):
    """GCMCSimulation::write_log_entry (simulation.cpp:667-704)."""
    ob = ref.observables(r)
    cfg = ref.config(r)
    counts = [0] * (len(s.identities) - 1)
    for c in cfg[1:]:
        counts[c.identity - 1] += 1
    label = label_of(s.moves[ob["last_move"]])
    return (f"Step: {step}\nTemperature: {gdef(ob['temperature'])}\n"
            f"Bound staples: {ob['num_staples']}\n"
            f"Unique bound staples: {sum(1 for n in counts if n)}\n"
            f"Fully bound domain pairs: {ob['num_fully_bound_domain_pairs']}\n"
            f"Misbound domain pairs: {ob['num_misbound_domain_pairs']}\n"
            f"Stacked domain pairs: {ob['num_stacked_domain_pairs']}\n"
            f"Linear helix triplets: 0\nStacked junction quadruplets: 0\n"
            f"Staple counts: {''.join(f'{n} ' for n in counts)}\n"
            f"System energy: {gdef(ob['energy'])}\nTotal external bias: {gdef(ob['bias'])}\n"
            f"Domain update external bias: 0\nMove update external bias: {gdef(ob['bias'])}\n"
            f"Movetype: {label}\nAccepted: {'true' if ob['last_accepted'] else 'false'}\n\n")


def expect_summary(ref, s):
    """The run summary written to the log and the .moves file
    (simulation.cpp:706-718; movetypes.cpp:88-96; met_movetypes.cpp:145-190;
    cb_movetypes.cpp:264-290, 626-678; rg_movetypes.cpp:583-627)."""
    c = ref.counters()
    tr = ref.move_tracking(0)
    log, moves = ["Run summary\n\n"], []
    for m, mv in enumerate(s.moves):
        label = label_of(mv)
        head = (f"Movetype: {label}\n    Attempts: {c.attempts[m]}\n    Accepts: {c.accepts[m]}\n"
                f"    Frequency: {ratio(c.accepts[m], c.attempts[m])}\n")
        log.append(head + "\n")
        if mv.kind == abi.MOVE_MET_STAPLE_EXCHANGE:
            types = [t for t in range(1, len(s.identities))
                     if tr["met_insert_attempts"][t] + tr["met_delete_attempts"][t] > 0]
            moves.append(head + "       Exchange multipliers\n        " + "1, " * len(types) + "\n")
            for t in types:
                ia, ic = tr["met_insert_attempts"][t], tr["met_insert_accepts"][t]
                da, dc = tr["met_delete_attempts"][t], tr["met_delete_accepts"][t]
                moves.append(f"    Staple type: {t}\n        Insertion attempts: {ia}\n"
                             f"        Insertion accepts: {ic}\n"
                             f"        Insertion frequency: {ratio(ic, ia, True)}\n"
                             f"        Deletion attempts: {da}\n        Deletion accepts: {dc}\n"
                             f"        Deletion frequency: {ratio(dc, da, True)}\n")
        elif mv.kind == abi.MOVE_CB_STAPLE_REGROWTH:
            moves.append(head + "\n")
            for t in range(1, len(s.identities)):
                a, k = tr["cb_attempts"][t], tr["cb_accepts"][t]
                if a:
                    moves.append(f"    Staple type: {t}\n        Attempts: {a}\n        Accepts: {k}\n"
                                 f"        Frequency: {ratio(k, a)}\n")
        elif mv.kind == abi.MOVE_CTCB_SCAFFOLD_REGROWTH:
            moves.append(head)
            for name, sep in (("length", "Number of scaffold domains"), ("staples", "Number of staples")):
                for n, (a, k) in enumerate(zip(tr[f"ctcb_{name}_attempts"], tr[f"ctcb_{name}_accepts"])):
                    if a:
                        moves.append(f"    {sep}: {n}\n        Attempts: {a}\n        Accepts: {k}\n"
                                     f"        Frequency: {ratio(k, a)}\n")
                if name == "length":
                    moves.append("\n")
        elif mv.kind == abi.MOVE_CTRG_SCAFFOLD_REGROWTH:
            # rg_movetypes.cpp:583-627
            moves.append(head)
            for n, (a, k) in enumerate(zip(tr["ctrg_length_attempts"], tr["ctrg_length_accepts"])):
                if a:
                    moves.append(f"    Number of scaffold domains: {n}\n        Attempts: {a}\n"
                                 f"        Accepts: {k}\n        Frequency: {ratio(k, a)}\n")
            moves.append("\n")
        elif mv.kind == abi.MOVE_CTCB_JUMP_SCAFFOLD_REGROWTH:
            # cb_movetypes.cpp:816-870, every move under 0 scaffold domains
            moves.append(head)
            att, acc = tr["ctcb_jump_staples_attempts"], tr["ctcb_jump_staples_accepts"]
            if sum(att):
                moves.append(f"    Number of scaffold domains: 0\n        Attempts: {sum(att)}\n"
                             f"        Accepts: {sum(acc)}\n"
                             f"        Frequency: {ratio(sum(acc), sum(att))}\n")
            moves.append("\n")
            for n, (a, k) in enumerate(zip(att, acc)):
                if a:
                    moves.append(f"    Number of staples: {n}\n        Attempts: {a}\n"
                                 f"        Accepts: {k}\n        Frequency: {ratio(k, a)}\n")
        elif mv.kind == abi.MOVE_CTRG_JUMP_SCAFFOLD_REGROWTH:
            # rg_movetypes.cpp:745-789, every move under 0 scaffold domains
            moves.append(head)
            a, k = tr["ctrg_jump_attempts"][0], tr["ctrg_jump_accepts"][0]
            if a:
                moves.append(f"    Number of scaffold domains: 0\n        Attempts: {a}\n"
                             f"        Accepts: {k}\n        Frequency: {ratio(k, a)}\n")
            moves.append("\n")
    return "".join(log), "".join(moves)


def expect_acceptance_freqs(ref, v1s, v2s=None):
    """OneDPTGCMCSimulation / TwoDPTGCMCSimulation::write_acceptance_freqs
    (ptmc_simulation.cpp:414-426, 562-592) from the oracle's per-pair
    counts."""
    att, acc = ref.exchange_counts()
    out = ""
    if v2s is None:
        for i in range(len(v1s) - 1):
            out += (f"{gdef(v1s[i])} {gdef(v1s[i + 1])} {acc[0][i]} {att[0][i]} "
                    f"{ratio(acc[0][i], att[0][i])} \n")
        return out + "\n"
    n2 = len(v2s)
    for i in range(len(v1s) - 1):
        for j in range(n2):
            q = i * n2 + j
            out += (f"{gdef(v1s[i])} {gdef(v1s[i + 1])} {gdef(v2s[j])} {acc[0][q]} {att[0][q]} "
                    f"{ratio(acc[0][q], att[0][q])} \n")
    out += "\n"
    for j in range(n2 - 1):
        for i in range(len(v1s)):
            q = i * n2 + j
            out += (f"{gdef(v2s[j])} {gdef(v2s[j + 1])} {gdef(v1s[i])} {acc[1][q]} {att[1][q]} "
                    f"{ratio(acc[1][q], att[1][q])} \n")
    return out + "\n"


def read(d, name):
    with open(os.path.join(d, name)) as f:
        return f.read().split("\n")


def run_driver(inp):
    print("case: ldo_run", inp, flush=True)
    p = subprocess.run([DRIVER, inp], capture_output=True, text=True, timeout=180)
    assert p.returncode == 0, p.stderr
    return p
    return p


def test_driver_rejects_unsupported_inputs(tmp_path):
    """Parsing errors stop the run before the GPU is touched: the reference's
    own default domain_type 'Halfturn' (rejected by its add_chain too), and
    an unsupported simulation type or movetype."""
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    s = systems.four_domain_system()
    for extra, msg in (({"domain_type": "Halfturn"}, "domain_type Halfturn"),
                       ({"simulation_type": "enumerate"}, "simulation_type enumerate"),
                       ({"simulation_type": "annealing", "max_temp": "340", "min_temp": "330",
                         "temp_interval": "3"}, "Bad temperature interval"),
                       ({"simulation_type": "2d_parallel_tempering", "num_reps": "5",
                         "temps": "330 340", "stacking_mults": "1 1.1"},
                        "num_reps must equal"),
                       ({"simulation_type": "st_parallel_tempering", "num_reps": "2",
                         "temps": "330 340", "stacking_mults": "1"},
                        "stacking_mults must list"),
                       # biases outside the umbrella-sampling drivers: refused,
                       # not silently dropped
                       ({"bias_functions_file": "biases.json"}, "bias_functions_file"),
                       ({"simulation_type": "ut_parallel_tempering", "num_reps": "2",
                         "temps": "330 340", "bias_functions_file": "biases.json"},
                        "bias_functions_file")):
        inp = write_inputs(str(tmp_path), s, {"simulation_type": "constant_temp", **extra})
        p = subprocess.run([DRIVER, inp], capture_output=True, text=True, timeout=60)
        assert p.returncode == 1 and msg in p.stderr, p.stderr
    s.moves = [systems.Move(abi.MOVE_ORIENTATION_ROTATION, 1.0)]
    inp = write_inputs(str(tmp_path), s, {"simulation_type": "constant_temp"})
    with open(os.path.join(str(tmp_path), "moves.json")) as f:
        text = f.read().replace("OrientationRotation", "CTRGLinkerRegrowth")
    with open(os.path.join(str(tmp_path), "moves.json"), "w") as f:
        f.write(text)
    p = subprocess.run([DRIVER, inp], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1 and "movetype CTRGLinkerRegrowth" in p.stderr, p.stderr
    # a CTRG entry without its recoil-growth limits (the reference's
    # get_int_option throws for them too)
    with open(os.path.join(str(tmp_path), "moves.json"), "w") as f:
        f.write(text.replace("CTRGLinkerRegrowth", "CTRGScaffoldRegrowth")
                .replace('"max_regrowth": 0', '"max_regrowth": 4'))
    p = subprocess.run([DRIVER, inp], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1 and "max_c_attempts" in p.stderr, p.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("moveset", ["ctcb", "standard_json"])
def test_driver_constant_temp_matches_oracle(tmp_path, oracle_mod, moveset):
    s = systems.four_domain_system()
    if moveset == "standard_json":
        # the reference's examples/moveset_standard.json as it is, labels and
        # "a/b" frequencies included (max_regrowth 12 > 4 domains: a
        # contiguous selection takes at most the whole scaffold)
        s.moves = systems.read_moveset_json(MOVESET_STANDARD)
    d = str(tmp_path)
    inp = write_inputs(d, s, {"simulation_type": "constant_temp", "temp": "330",
                              "ct_steps": 20000, "configs_output_freq": 4000,
                              "vtf_output_freq": 4000, "counts_output_freq": 2000,
                              "times_output_freq": 5000, "energies_output_freq": 2000,
                              "order_params_output_freq": 2000, "logging_freq": 4000})
    if moveset == "standard_json":
        shutil.copyfile(MOVESET_STANDARD, os.path.join(d, "moves.json"))
    out = run_driver(inp).stdout
    ref = oracle_mod.Oracle(s, temps=[330.0])
    ene, ops, counts, trj = read(d, "out.ene"), read(d, "out.ops"), read(d, "out.counts"), \
        read(d, "out.trj")
    assert ene[0] == "step tenergy henthalpy hentropy stacking bias"
    assert ops[0] == "numfulldomains, nummisdomains, numstackedpairs, numstaples, "
    want_trj, want_vcf, want_states, want_ores = [], [], [], []
    want_log = ""
    staples, staplestates = read(d, "out.staples"), read(d, "out.staplestates")
    for k in range(1, 11):
        ref.sweep(2000)
        e, o, c, t = expect_lines(ref, 0, s, 2000 * k)
        assert ene[k] == e and ops[k] == o and counts[k - 1] == c, k
        vcf, st, orr, stp, sts = expect_config_records(s, ref.config(0), 2000 * k)
        assert staples[k - 1] == stp and staplestates[k - 1] == sts, k
        if k % 2 == 0:
            want_trj += t
            want_vcf += vcf
            want_states.append(st)
            want_ores.append(orr)
        if (2000 * k) % 4000 == 0:
            want_log += expect_log_entry(ref, 0, s, 2000 * k)
    assert trj[:len(want_trj)] == want_trj
    assert read(d, "out.vcf")[:len(want_vcf)] == want_vcf
    assert read(d, "out.states")[:5] == want_states
    assert read(d, "out.ores")[:5] == want_ores
    times = read(d, "out.times")
    assert times[0] == "step time" and [x.split()[0] for x in times[1:5]] == \
        ["5000", "10000", "15000", "20000"]
    # the per-step log and the run summary on standard output, the movetype
    # summaries in the .moves file
    log, moves = expect_summary(ref, s)
    assert out == want_log + log
    with open(os.path.join(d, "out.moves")) as f:
        assert f.read() == moves
    assert read(d, "out.vsf") == ["atom 0:4 radius 0.25 type scaffold",
                                  "atom 4:11 radius 0.25 type staple"]


@pytest.mark.gpu
def test_driver_ptmc_matches_oracle(tmp_path, oracle_mod):
    s = systems.raster_system(replicas=4)
    d = str(tmp_path)
    temps = s.temps
    inp = write_inputs(d, s, {"simulation_type": "ut_parallel_tempering", "num_reps": 4,
                              "temps": " ".join(repr(t) for t in temps),
                              "chem_pot_mults": "1 1 1 1", "bias_mults": "1 1 1 1",
                              "stacking_mults": "1 1 1 1", "swaps": 8,
                              "exchange_interval": 500, "configs_output_freq": 1000,
                              "energies_output_freq": 1000, "order_params_output_freq": 1000,
                              "logging_freq": 1000})
    out = run_driver(inp).stdout
    ref = oracle_mod.Oracle(s, temps=temps, exchange_mode=abi.EXCHANGE_TEMPERATURE)
    swp = read(d, "out.swp")
    assert swp[0] == "".join(f"{g10(t)}/1/ " for t in temps)
    logs = [""] * 4
    for k in range(1, 5):
        # each entry is written before the round's exchange
        # (ptmc_simulation.cpp:140-147)
        ref.run(1, 500)
        before = ref.permutation()
        ref.run(1, 500)
        assert swp[k] == "".join(f"{q} " for q in before)
        for r in range(4):
            e, o, _, _ = expect_lines(ref, r, s, 1000 * k)
            assert read(d, f"out-{r}.ene")[k] == e
            assert read(d, f"out-{r}.ops")[k] == o
            logs[r] += expect_log_entry(ref, r, s, 1000 * k)
    # the closing entry after the last exchange
    assert swp[5:] == ["".join(f"{q} " for q in ref.permutation()), ""]
    # each rank's log (ptmc_simulation.cpp:48-56), no run summary
    for r in range(4):
        with open(os.path.join(d, f"out-{r}.out")) as f:
            assert f.read() == logs[r], r
    assert out.endswith(expect_acceptance_freqs(ref, temps))


PT_VARIANTS = {
    # simulation_type: (exchange mode, exchanged control quantities)
    "t_parallel_tempering": (abi.EXCHANGE_T, "t"),
    "hut_parallel_tempering": (abi.EXCHANGE_HUT, "tub"),
    "st_parallel_tempering": (abi.EXCHANGE_ST, "ts"),
    "2d_parallel_tempering": (abi.EXCHANGE_2D, "tus"),
}


@pytest.mark.gpu
@pytest.mark.parametrize("sim", sorted(PT_VARIANTS))
def test_driver_pt_variants_match_oracle(tmp_path, oracle_mod, sim):
    """t_, hut_, st_ and 2d_parallel_tempering (ptmc_simulation.cpp:428-680)
    on the four-domain system: the swap file's header of exchanged control
    quantities (91-104), its entries, every replica's energies and log, and
    the acceptance frequencies on standard output, against the oracle."""
    s = systems.four_domain_system()
    d = str(tmp_path)
    mode, qs = PT_VARIANTS[sim]
    common = {"simulation_type": sim, "swaps": 8, "exchange_interval": 250,
              "configs_output_freq": 500, "energies_output_freq": 500,
              "logging_freq": 500, "staple_u_mult": "1"}
    if mode == abi.EXCHANGE_2D:
        v1s, v2s = [326.0, 334.0], [1.0, 1.15, 0.85]
        temps = [t for t in v1s for _ in v2s]
        ctrl = {"t": temps, "u": [1.0] * 6, "s": v2s * 2}
        extra = {"num_reps": 6, "temps": "326 334", "stacking_mults": "1 1.15 0.85"}
    else:
        temps = [320.0, 326.0, 326.0, 332.0, 338.0, 344.0]
        ctrl = {"t": temps, "u": [1.0, 1.1, 0.9, 1.05, 1.0, 0.95],
                "b": [1.0, 0.9, 1.1, 1.0, 1.2, 1.0], "s": [1.0, 1.1, 0.9, 1.2, 1.0, 0.8]}
        extra = {"num_reps": 6, "temps": " ".join(repr(t) for t in temps),
                 "chem_pot_mults": " ".join(repr(x) for x in ctrl["u"]),
                 "bias_mults": " ".join(repr(x) for x in ctrl["b"]),
                 "stacking_mults": " ".join(repr(x) for x in ctrl["s"])}
    out = run_driver(write_inputs(d, s, {**common, **extra})).stdout
    ref = oracle_mod.Oracle(s, temps=temps, exchange_mode=mode)
    if mode == abi.EXCHANGE_2D:
        ref.set_ladder_2d(v1s, v2s, 1.0)
    else:
        ref.set_ladder(temps, ctrl["u"], ctrl["b"], ctrl["s"])
    swp = read(d, "out.swp")
    assert swp[0] == "".join("".join(f"{gdef(ctrl[k][q])}/" for k in qs) + " "
                             for q in range(6))
    logs = [""] * 6
    for k in range(1, 5):
        ref.run(1, 250)
        before = ref.permutation()
        ref.run(1, 250)
        assert swp[k] == "".join(f"{q} " for q in before), k
        for r in range(6):
            e, o, _, _ = expect_lines(ref, r, s, 500 * k)
            assert read(d, f"out-{r}.ene")[k] == e, (k, r)
            logs[r] += expect_log_entry(ref, r, s, 500 * k)
    assert swp[5:] == ["".join(f"{q} " for q in ref.permutation()), ""]
    for r in range(6):
        with open(os.path.join(d, f"out-{r}.out")) as f:
            assert f.read() == logs[r], r
    att, _ = ref.exchange_counts()
    assert att.sum() > 0
    if mode == abi.EXCHANGE_2D:
        assert out.endswith(expect_acceptance_freqs(ref, v1s, v2s))
    else:
        assert out.endswith(expect_acceptance_freqs(ref, temps))


@pytest.mark.gpu
def test_driver_annealing_matches_oracle(tmp_path, oracle_mod):
    """simulation_type annealing (AnnealingGCMCSimulation::run,
    annealing_simulation.cpp:38-50): 340, 335 and 330 K for 1000 steps each,
    each temperature starting at the step the previous simulate call
    returned plus the steps before it (1..1000, 1002..2001, 3004..4003), the
    log on standard output with a run summary after every temperature, and
    the .moves file of the whole run."""
    s = systems.four_domain_system()
    d = str(tmp_path)
    inp = write_inputs(d, s, {"simulation_type": "annealing", "max_temp": "340",
                              "min_temp": "330", "temp_interval": "5", "steps_per_temp": 1000,
                              "energies_output_freq": 500, "order_params_output_freq": 500,
                              "configs_output_freq": 500, "logging_freq": 1000})
    out = run_driver(inp).stdout
    ref = oracle_mod.Oracle(s, temps=[340.0])
    want_ene, want_trj, want_log = [], [], ""
    step = 0
    for temp in (340.0, 335.0, 330.0):
        ref.set_ladder([temp])
        ref.run(1, 0)
        ref.set_step(step)
        cur, end = step, step + 1000
        for at in sorted(set(range((step // 500 + 1) * 500, end + 1, 500)) | {end}):
            ref.sweep(at - cur)
            cur = at
            if at % 500 == 0:
                e, _, _, trj = expect_lines(ref, 0, s, at)
                want_ene.append(e)
                want_trj += trj
            if at % 1000 == 0:
                want_log += expect_log_entry(ref, 0, s, at)
        want_log += expect_summary(ref, s)[0]
        step += step + 1000 + 1
    assert read(d, "out.ene")[1:1 + len(want_ene)] == want_ene
    assert read(d, "out.trj")[:len(want_trj)] == want_trj
    assert out == want_log
    with open(os.path.join(d, "out.moves")) as f:
        assert f.read() == expect_summary(ref, s)[1]


@pytest.mark.gpu
def test_driver_ptmc_restart_matches_oracle(tmp_path, oracle_mod):
    """A PTMC run restarted from frame 2 of each replica's trajectory and the
    last line of the swap file (restart_from_config / restart_from_swap)
    continues as the oracle does from the same configurations and slots."""
    s = systems.raster_system(replicas=4)
    d = str(tmp_path)
    temps = s.temps
    common = {"simulation_type": "ut_parallel_tempering", "num_reps": 4,
              "temps": " ".join(repr(t) for t in temps), "chem_pot_mults": "1 1 1 1",
              "bias_mults": "1 1 1 1", "stacking_mults": "1 1 1 1", "exchange_interval": 500,
              "configs_output_freq": 1000, "energies_output_freq": 1000}
    run_driver(write_inputs(d, s, {**common, "swaps": 6}))
    first = os.path.join(d, "first")
    os.mkdir(first)
    for name in os.listdir(d):
        if name.startswith("out"):
            os.rename(os.path.join(d, name), os.path.join(first, name))
    run_driver(write_inputs(d, s, {**common, "swaps": 4, "restart_from_config": "true",
                                   "restart_traj_filebase": os.path.join(first, "out"),
                                   "restart_step": 2, "restart_from_swap": "true",
                                   "restart_swap_file": os.path.join(first, "out.swp")}))
    ref = oracle_mod.Oracle(s, temps=temps, exchange_mode=abi.EXCHANGE_TEMPERATURE)
    ref0 = oracle_mod.Oracle(s, temps=temps, exchange_mode=abi.EXCHANGE_TEMPERATURE)
    ref0.run(6, 500)  # frame 2 is step 3000, the run's end; the swap file's last line too
    for r in range(4):
        ref.set_config(r, ref0.config(r))
    ref.set_permutation(ref0.permutation())
    for k in range(1, 3):
        ref.run(2, 500)
        for r in range(4):
            e, _, _, _ = expect_lines(ref, r, s, 1000 * k)
            assert read(d, f"out-{r}.ene")[k] == e, (k, r)


def write_mwus_inputs(d, s, windows, extra):
    """The reference's order-parameter, bias-function and windows files for
    numstaples windows with a LinearStepWell restraint and a Grid bias."""
    with open(os.path.join(d, "ops.json"), "w") as f:
        json.dump({"origami": {"order_params": [
            {"label": "Number of bound staples", "tag": "numstaples", "type": "NumStaples",
             "level": 0},
            {"label": "Number of fully bound domain pairs", "tag": "numfulldomains",
             "type": "NumBoundDomainPairs", "level": 0}]}}, f)
    with open(os.path.join(d, "biases.json"), "w") as f:
        json.dump({"origami": {"bias_functions": [
            {"label": "well", "tag": "lswnumstaples", "type": "LinearStepWell", "min_op": 0,
             "max_op": 24, "well_bias": 0, "min_bias": 2, "slope": 1.5, "ops": ["numstaples"],
             "update_per_domain": False},
            {"label": "grid", "tag": "grid", "type": "Grid", "ops": ["numstaples"],
             "update_per_domain": False}]}}, f)
    with open(os.path.join(d, "run.windows"), "w") as f:
        f.write("lswnumstaples\n" + "".join(f"{lo}, {hi}\n" for lo, hi in windows))
    return write_inputs(d, s, {
        "simulation_type": "mw_umbrella_sampling", "temp": "330",
        "order_parameter_file": os.path.join(d, "ops.json"),
        "bias_functions_file": os.path.join(d, "biases.json"),
        "windows_file": os.path.join(d, "run.windows"), "us_grid_bias_tag": "grid",
        "multi_window": "true", "max_equil_dur": 1000000000, "max_iter_dur": 1000000000,
        **extra})


def test_driver_rejects_unsupported_umbrella_inputs(tmp_path):
    """Umbrella inputs the driver cannot honour stop the run before the GPU
    is touched: a bias function besides the window restraint and the grid,
    and windows over two order parameters."""
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    d = str(tmp_path)
    s = systems.raster_system(replicas=2)
    inp = write_mwus_inputs(d, s, [(0, 2), (1, 3)], {})
    with open(os.path.join(d, "biases.json")) as f:
        b = json.load(f)
    b["origami"]["bias_functions"].append({"label": "x", "tag": "extra", "type": "SquareWell",
                                           "ops": ["numfulldomains"], "min_op": 0, "max_op": 1,
                                           "well_bias": 0, "outside_bias": 1})
    with open(os.path.join(d, "biases.json"), "w") as f:
        json.dump(b, f)
    p = subprocess.run([DRIVER, inp], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1 and "bias function extra" in p.stderr, p.stderr
    inp = write_mwus_inputs(d, s, [(0, 2), (1, 3)], {})
    with open(os.path.join(d, "run.windows"), "w") as f:
        f.write("lswnumstaples\n0 1, 2 3\n")
    p = subprocess.run([DRIVER, inp], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1 and "one order parameter" in p.stderr, p.stderr


@pytest.mark.gpu
def test_driver_mwus_iterations_match_oracle(tmp_path, oracle_mod):
    """Three numstaples windows, an equilibration run and three adaptive
    iterations (max_D_bias 1): every window's per-iteration energies, its
    -inp.biases and .biases weights and its .out summary equal the oracle's
    run with the same histograms fed through the restated update
    (oracle/us_ref.py)."""
    import numpy as np
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "oracle"))
    import us_ref
    s = systems.raster_system(replicas=3)
    d = str(tmp_path)
    windows = [(0, 2), (1, 3), (2, 4)]
    run_driver(write_mwus_inputs(d, s, windows, {
        "equil_steps": 1000, "iter_steps": 2000, "max_num_iters": 3, "max_D_bias": 1.0,
        "energies_output_freq": 500}))
    W = len(windows)
    glen = 64 + s.max_total_staples * s.max_staple_size + 1
    wins = [abi.Window(abi.OP_NUMSTAPLES, lo, hi, 0, 0.0, 2.0, 1.5, 0.0, 1.0)
            for lo, hi in windows]
    ref = oracle_mod.Oracle(s, temps=[330.0] * W)
    ref.set_windows(wins, [0.0] * (W * glen), glen)
    ref.set_histogram(True)
    post = [f"out_win-{lo}--{hi}" for lo, hi in windows]
    for k in (1, 2):
        ref.sweep(500)
        for w in range(W):
            e, _, _, _ = expect_lines(ref, w, s, 500 * k)
            assert read(d, post[w] + "_iter-equil.ene")[k] == e, (k, w)
    ref.histogram(reset=True)
    us = [us_ref.WindowState() for _ in range(W)]
    summaries = [""] * W
    visited = 0
    for n in range(3):
        for w in range(W):
            with open(os.path.join(d, f"{post[w]}_iter-{n}-inp.biases")) as f:
                assert f.read() == us_ref.weights_json(us[w]), (n, w)
        for k in range(1, 5):
            ref.sweep(500)
            for w in range(W):
                e, _, _, _ = expect_lines(ref, w, s, 500 * k)
                assert read(d, f"{post[w]}_iter-{n}.ene")[k] == e, (n, k, w)
        h = ref.histogram(reset=True)
        assert (h[:, glen] == 0).all() and (h.sum(axis=1) == 2000).all()
        for w in range(W):
            us_ref.process_iteration(us[w], [int(x) for x in h[w, :glen]], us[w].grid(glen),
                                     2001, 1.0)
            summaries[w] += us_ref.summary(us[w], n)
            with open(os.path.join(d, f"{post[w]}_iter-{n}.biases")) as f:
                assert f.read() == us_ref.weights_json(us[w]), (n, w)
            visited += len(us[w].S_n)
        ref.set_window_grid(np.concatenate([us[w].grid(glen) for w in range(W)]).tolist())
    for w in range(W):
        with open(os.path.join(d, post[w] + ".out")) as f:
            assert f.read() == summaries[w], w
    assert visited > 3 * W  # the windows sampled more than one point each


@pytest.mark.gpu
def test_driver_mwus_restart_from_named_files(tmp_path, oracle_mod):
    """restart_from_config with restart_traj_files (one file per window,
    us_simulation.cpp:542-546) and restart_steps: each window starts from the
    named frame of another window's equilibration trajectory and continues
    as the oracle does from the same configurations."""
    s = systems.raster_system(replicas=3)
    d = str(tmp_path)
    windows = [(0, 2), (1, 3), (2, 4)]
    W = len(windows)
    common = {"max_num_iters": 0, "configs_output_freq": 500, "energies_output_freq": 500}
    run_driver(write_mwus_inputs(d, s, windows, {**common, "equil_steps": 1000}))
    first = os.path.join(d, "first")
    os.mkdir(first)
    for name in os.listdir(d):
        if name.startswith("out"):
            os.rename(os.path.join(d, name), os.path.join(first, name))
    post = [f"out_win-{lo}--{hi}" for lo, hi in windows]
    src = [(2, 1), (0, 0), (1, 1)]  # window w starts from (window, frame)
    files = " ".join(os.path.join(first, post[v] + "_iter-equil.trj") for v, _ in src)
    run_driver(write_mwus_inputs(d, s, windows, {
        **common, "equil_steps": 1000, "restart_from_config": "true",
        "restart_traj_files": files, "restart_steps": " ".join(str(f) for _, f in src)}))
    glen = 64 + s.max_total_staples * s.max_staple_size + 1
    wins = [abi.Window(abi.OP_NUMSTAPLES, lo, hi, 0, 0.0, 2.0, 1.5, 0.0, 1.0)
            for lo, hi in windows]
    ref0 = oracle_mod.Oracle(s, temps=[330.0] * W)
    ref0.set_windows(wins, [0.0] * (W * glen), glen)
    frames = []
    for _ in range(2):
        ref0.sweep(500)
        frames.append([ref0.config(w) for w in range(W)])
    assert len({repr(frames[f][v]) for v, f in src}) == W
    ref = oracle_mod.Oracle(s, temps=[330.0] * W)
    ref.set_windows(wins, [0.0] * (W * glen), glen)
    for w, (v, f) in enumerate(src):
        ref.set_config(w, frames[f][v])
    for k in (1, 2):
        ref.sweep(500)
        for w in range(W):
            e, _, _, _ = expect_lines(ref, w, s, 500 * k)
            assert read(d, post[w] + "_iter-equil.ene")[k] == e, (k, w)


@pytest.mark.gpu
def test_driver_single_window_us_matches_oracle(tmp_path, oracle_mod):
    """simulation_type umbrella_sampling (SimpleUSGCMCSimulation): one window
    whose restraint is the bias file's own LinearStepWell (min_op 1, max_op
    3), an equilibration run and two adaptive iterations. The iteration
    energies and bias files equal the oracle's one-window run, and the
    constructor's messages and the iteration summaries go to standard
    output, as the reference's stream defaults to it for a single window
    (us_simulation.hpp:78)."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "oracle"))
    import us_ref
    s = systems.raster_system(replicas=1)
    d = str(tmp_path)
    inp = write_mwus_inputs(d, s, [], {
        "simulation_type": "umbrella_sampling", "equil_steps": 1000, "iter_steps": 2000,
        "max_num_iters": 2, "max_D_bias": 1.0, "energies_output_freq": 500})
    with open(os.path.join(d, "biases.json")) as f:
        b = json.load(f)
    b["origami"]["bias_functions"][0].update({"min_op": 1, "max_op": 3})
    with open(os.path.join(d, "biases.json"), "w") as f:
        json.dump(b, f)
    p = run_driver(inp)
    glen = 64 + s.max_total_staples * s.max_staple_size + 1
    ref = oracle_mod.Oracle(s, temps=[330.0])
    ref.set_windows([abi.Window(abi.OP_NUMSTAPLES, 1, 3, 0, 0.0, 2.0, 1.5, 0.0, 1.0)],
                    [0.0] * glen, glen)
    ref.set_histogram(True)
    for k in (1, 2):
        ref.sweep(500)
        e, _, _, _ = expect_lines(ref, 0, s, 500 * k)
        assert read(d, "out_iter-equil.ene")[k] == e, k
    ref.histogram(reset=True)
    us = us_ref.WindowState()
    summary = ""
    for n in range(2):
        with open(os.path.join(d, f"out_iter-{n}-inp.biases")) as f:
            assert f.read() == us_ref.weights_json(us), n
        for k in range(1, 5):
            ref.sweep(500)
            e, _, _, _ = expect_lines(ref, 0, s, 500 * k)
            assert read(d, f"out_iter-{n}.ene")[k] == e, (n, k)
        h = ref.histogram(reset=True)
        us_ref.process_iteration(us, [int(x) for x in h[0, :glen]], us.grid(glen), 2001, 1.0)
        summary += us_ref.summary(us, n)
        with open(os.path.join(d, f"out_iter-{n}.biases")) as f:
            assert f.read() == us_ref.weights_json(us), n
        ref.set_window_grid(list(us.grid(glen)))
    assert p.stdout == ("No biases read in\nStarting from configuration in system file\n"
                        "Starting new iteration\n" + summary), p.stdout
    assert not os.path.exists(os.path.join(d, "out.out"))


@pytest.mark.gpu
def test_driver_ptmwus_iterations_match_oracle(tmp_path, oracle_mod):
    """Three numstaples windows exchanging configurations every 500 steps
    (ptmw_umbrella_sampling): equilibration, then two iterations of 3 swaps.
    Each window's energies come from the configuration it held before the
    round's swap, its bias column from its own stored value; the swap file
    lists each window's configuration; the weights divide the visits by the
    1500 steps of the iteration (run_swaps counts no extra step).  The odd
    swap count checks that every iteration pairs from swap_i = 1 again
    (us_simulation.cpp:670) while the exchange's random stream continues."""
    import numpy as np
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "oracle"))
    import us_ref
    s = systems.raster_system(replicas=3)
    d = str(tmp_path)
    windows = [(0, 2), (0, 3), (1, 4)]
    p = run_driver(write_mwus_inputs(d, s, windows, {
        "simulation_type": "ptmw_umbrella_sampling", "equil_steps": 1000, "iter_swaps": 3,
        "exchange_interval": 500, "max_num_iters": 2, "max_D_bias": 1.0,
        "energies_output_freq": 500, "configs_output_freq": 1000}))
    W = len(windows)
    glen = 64 + s.max_total_staples * s.max_staple_size + 1
    wins = [abi.Window(abi.OP_NUMSTAPLES, lo, hi, 0, 0.0, 2.0, 1.5, 0.0, 1.0)
            for lo, hi in windows]
    ref = oracle_mod.Oracle(s, temps=[330.0] * W, exchange_mode=abi.EXCHANGE_WINDOWS)
    ref.set_windows(wins, [0.0] * (W * glen), glen)
    ref.set_histogram(True)
    post = [f"out_win-{lo}--{hi}" for lo, hi in windows]
    ref.sweep(1000)
    ref.histogram(reset=True)
    us = [us_ref.WindowState() for _ in range(W)]
    summaries = [""] * W
    swapped = False
    K = 3
    freqs = ""
    for n in range(2):
        ref.restart_swaps()
        att0, acc0 = ref.exchange_counts()
        swp = [" ".join(str(q) for q in ref.permutation()) + " "]
        for k in range(1, K + 1):
            before = ref.permutation()
            ref.run(1, 500)
            after = ref.permutation()
            swapped = swapped or before != after
            for w in range(W):
                e, _, _, _ = expect_lines(ref, before[w], s, 500 * k)
                e = e.rsplit(" ", 2)[0] + " " + g10(ref.observables(after[w])["bias"]) + " "
                assert read(d, f"{post[w]}_iter-{n}.ene")[k] == e, (n, k, w)
            if (500 * k) % 1000 == 0:
                swp.append(" ".join(str(q) for q in before) + " ")
        if (500 * K) % 1000 == 0:
            swp.append(" ".join(str(q) for q in ref.permutation()) + " ")
        assert read(d, f"out_iter-{n}.swp")[:-1] == swp, n
        # write_acceptance_freqs (us_simulation.cpp:884-898) per iteration
        att1, acc1 = ref.exchange_counts()
        freqs += f"Iter {n}\nWindow 1, Window 2, Swaps, Attempts, Frequency\n"
        for i in range(W - 1):
            a, c = int(att1[0, i] - att0[0, i]), int(acc1[0, i] - acc0[0, i])
            assert a > 0
            freqs += f"{i} {i + 1} {c} {a} {c / a:g} \n"
        freqs += "\n"
        h = ref.histogram(reset=True)
        for w in range(W):
            us_ref.process_iteration(us[w], [int(x) for x in h[w, :glen]], us[w].grid(glen),
                                     500 * K, 1.0)
            summaries[w] += us_ref.summary(us[w], n)
            with open(os.path.join(d, f"{post[w]}_iter-{n}.biases")) as f:
                assert f.read() == us_ref.weights_json(us[w]), (n, w)
        ref.set_window_grid(np.concatenate([us[w].grid(glen) for w in range(W)]).tolist())
    for w in range(W):
        with open(os.path.join(d, post[w] + ".out")) as f:
            assert f.read() == summaries[w], w
    assert p.stdout == freqs, p.stdout
    assert swapped
