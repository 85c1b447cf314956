"""The C ABI library loads and exports every entry point include/ldo.h
declares (no GPU calls)."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    from latticednaorigami_amd import abi
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            header = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"\b(ldo_[a-z_]+)\s*\(", header))
    assert "ldo_create" in names and "ldo_sweep" in names and "ldo_mlp_loss_grad" in names
    lib = C.CDLL(abi.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_sizes_match_header():
    """ctypes mirrors have the C layout (checked against offsetof values
    compiled from include/ldo.h)."""
    import subprocess
    import tempfile
    from latticednaorigami_amd import abi
    src = r'''
#include <stddef.h>
#include <stdio.h>
#include "ldo.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(ldo_system_desc), sizeof(ldo_run_params),
         sizeof(ldo_config), sizeof(ldo_obs), sizeof(ldo_counters),
         offsetof(ldo_system_desc, binding_h), offsetof(ldo_run_params, seed));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    got = [int(v) for v in out]
    want = [C.sizeof(abi.SystemDesc), C.sizeof(abi.RunParams), C.sizeof(abi.Config),
            C.sizeof(abi.Obs), C.sizeof(abi.Counters), abi.SystemDesc.binding_h.offset,
            abi.RunParams.seed.offset]
    assert got == want


def test_reference_example_moveset_parses():
    """The reference's examples/moveset_standard.json (a fixture copy): every
    movetype it names is supported, with its labels, fractions and limits."""
    import os
    from latticednaorigami_amd import abi, systems
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "moveset_standard.json")
    moves = systems.read_moveset_json(path)
    assert [m.kind for m in moves] == [abi.MOVE_ORIENTATION_ROTATION, abi.MOVE_MET_STAPLE_EXCHANGE,
                                       abi.MOVE_CB_STAPLE_REGROWTH,
                                       abi.MOVE_CTRG_SCAFFOLD_REGROWTH,
                                       abi.MOVE_CTRG_JUMP_SCAFFOLD_REGROWTH]
    assert [m.freq for m in moves] == [2 / 6, 1 / 6, 1 / 6, 1 / 6, 1 / 6]
    assert (moves[3].max_num_recoils, moves[3].max_c_attempts, moves[3].max_regrowth) == (1, 36, 12)
    assert (moves[4].max_num_recoils, moves[4].max_c_attempts, moves[4].max_regrowth,
            moves[4].max_seg_regrowth) == (1, 36, 12, 2)
    assert moves[3].label == "Contiguous CTRG scaffold regrowth"
