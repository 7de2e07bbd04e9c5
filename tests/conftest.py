import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def systems(tmp_path_factory):
    """Seeded synthetic windows (topology/fep/inp) shared by the tests."""
    from q_amd import synth
    base = tmp_path_factory.mktemp("systems")
    cache = {}

    systems_cache = {}

    def get(config, seed=0, steps=20, **kw):
        key = (config, seed, steps, tuple(sorted(kw.items())))
        if key not in cache:
            d = base / f"{config}_{seed}_{steps}_{len(cache)}"
            builder, bkw, _ = synth.CONFIGS[config]
            # configurations that differ only in their .inp share the system
            skey = (builder.__name__, tuple(sorted(bkw.items())), seed)
            if skey not in systems_cache:
                systems_cache[skey] = builder(seed=seed, **bkw)
            cache[key] = synth.make_window(config, str(d), seed=seed, steps=steps,
                                           system=systems_cache[skey], **kw)
        return cache[key]

    return get


@pytest.fixture(scope="session")
def golden_inp(systems, tmp_path_factory):
    """Input file of the window a golden vector set of tests/golden/ was made
    from (same generator arguments as tests/golden/make_golden.py)."""
    from tests.golden import make_golden as M
    cache = {}

    def get(cfg, seed):
        if (cfg, seed) not in cache:
            if cfg == "kat":
                cache[(cfg, seed)] = M.kat_window(str(tmp_path_factory.mktemp("kat")))
            else:
                cache[(cfg, seed)] = systems(cfg, seed=seed, steps=M.STEPS, energy=M.ENERGY, output=M.OUTPUT)
        return cache[(cfg, seed)]

    return get
