"""ldo_set_config_noalloc against ldo_set_config on an MI355X: two samplers
of the raster PTMC system receive the same configuration, one through each
entry point, and must then hold the same state (configuration, observables)
and run the same trajectory bit for bit.  Two calls in a row, the second
with a smaller configuration than the first."""
import pytest

from latticednaorigami_amd import abi, systems

pytestmark = [pytest.mark.gpu]

STEPS = 2000


def _same(a, b):
    oa, ob = a.observables(0), b.observables(0)
    assert oa == ob
    assert a.config(0) == b.config(0)


def test_set_config_noalloc_matches_set_config(oracle_mod):
    from latticednaorigami_amd.sampler import Sampler
    s = systems.raster_system()
    temps = s.temps[:1]
    # two configurations from the oracle: many staples, then fewer
    ref = oracle_mod.Oracle(s, temps=[300.0])
    ref.sweep(40000)
    big = ref.config(0)
    small = [c for c in big if c.identity == 0] + [c for c in big if c.identity != 0][:2]
    assert len(big) > len(small) > 1
    a = Sampler(s, temps=temps)
    b = Sampler(s, temps=temps)
    for cfg in (big, small):
        print(f"configuration with {len(cfg)} chains", flush=True)
        a.set_config(0, cfg)
        b.set_config_noalloc(0, cfg)
        assert b.config(0) == cfg
        _same(a, b)
        a.sweep(STEPS)
        b.sweep(STEPS)
        _same(a, b)
    # argument checks match the allocating entry point
    with pytest.raises(abi.LdoError) as e:
        b.set_config_noalloc(5, small)
    assert e.value.code == abi.LDO_E_ARG
    bad = [systems.Chain(1, 1, small[1].positions, small[1].orientations)]
    with pytest.raises(abi.LdoError) as e:
        b.set_config_noalloc(0, bad)
    assert e.value.code == abi.LDO_E_ARG
