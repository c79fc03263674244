"""Each oracle pin rejects the mistake it is there to catch (CPU only).

The mutants (tests/oracle_mutants.py) are copies of oracle/fsmc_oracle.c with
one planted mistake each, compiled with the oracle's flags; every pin must
pass on the real oracle (test_oracle_pins.py) and fail on its mutants here.
"""
import concurrent.futures as cf

import pytest

import oracle as orc
import oracle_mutants as om
import test_oracle_pins as pins

CASES = [
    ("lsmc_no_terminal_scale", pins.pin_lsmc_ref_two_dates),
    ("lsmc_terminal_at_n_steps", pins.pin_lsmc_ref_two_dates),
    ("lsmc_no_terminal_scale", pins.pin_lsmc_ref_lag_and_ridge),
    ("lsmc_terminal_at_n_steps", pins.pin_lsmc_ref_lag_and_ridge),
    ("lsmc_no_lag", pins.pin_lsmc_ref_lag_and_ridge),
    ("lsmc_no_ridge", pins.pin_lsmc_ref_lag_and_ridge),
    ("libor_no_drift", pins.pin_libor_spot_measure_drift),
    ("libor_flipped_ito", pins.pin_libor_spot_measure_drift),
    ("libor_negative_drift", pins.pin_libor_spot_measure_drift),
    ("heston_val2_chunk_mean", pins.pin_heston_running_val2),
    ("heston_val2_ignored", pins.pin_heston_running_val2),
    ("geometric_continuous_variance", pins.pin_geometric_asian_discrete),
    ("geometric_continuous_mean", pins.pin_geometric_asian_discrete),
]


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("oracle_mutants"))
    names = sorted({name for name, _ in CASES})
    with cf.ThreadPoolExecutor(4) as ex:
        paths = list(ex.map(lambda n: om.build(n, out), names))
    return dict(zip(names, paths))


def test_every_mutant_is_listed():
    assert {name for name, _ in CASES} == set(om.MUTANTS)


@pytest.mark.parametrize("name,pin", CASES, ids=[f"{n}-{p.__name__}" for n, p in CASES])
def test_pin_rejects_mutant(mutant_libs, name, pin):
    pin()  # the real oracle passes
    with orc.use(mutant_libs[name]):
        with pytest.raises(AssertionError):
            pin()
