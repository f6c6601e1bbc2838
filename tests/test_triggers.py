"""The snapshot trigger chain of the time loop (host logic, CPU), against
write sequences worked out by hand from src/CosmoBEC_MPI_stan.cpp:238-284
and src/CosmoBEC_MPI.cpp:238-284 for a fixed scale-factor trajectory."""
import pytest


def _run(variant, a_triggers=True, twrite=100, t_stop=1000, step=0.03, a0=0.021):
    import ifdm_code_amd as F
    trig = F.OutputTriggers(variant, a_triggers)
    out = []
    for tn in range(t_stop):
        a = a0 + step * (tn + 1)          # par.ai after Timestep_criteria of step tn
        t, stop = trig(a, tn, twrite, t_stop)
        if t is not None:
            out.append((tn, t))
        if stop:
            return out, tn
    return out, None


def test_stan_chain_exact():
    out, stop_tn = _run(0)
    a = [0.021 + 0.03 * (tn + 1) for tn in range(20)]
    first_105 = next(tn for tn in range(20) if a[tn] > 0.105)
    first_25 = next(tn for tn in range(20) if a[tn] > 0.25)
    first_5 = next(tn for tn in range(20) if a[tn] > 0.5)
    expect = [(0, 1), (first_105, 999994), (first_25, 999995)]
    expect += [(tn, 2 + k) for k, tn in enumerate(range(first_25 + 1, first_5))]
    expect += [(first_5, 999997)]
    assert out == expect
    assert stop_tn == first_5


def test_mpi_variant_chain():
    out, stop_tn = _run(1)
    a = [0.021 + 0.03 * (tn + 1) for tn in range(20)]
    first_105 = next(tn for tn in range(20) if a[tn] > 0.105)
    first_25 = next(tn for tn in range(20) if a[tn] > 0.25)
    first_333 = next(tn for tn in range(20) if a[tn] > 0.333)
    first_5 = next(tn for tn in range(20) if a[tn] > 0.5)
    expect = [(0, 1), (first_105, 999994), (first_25, 999995)]
    expect += [(tn, 2 + k) for k, tn in enumerate(range(first_25 + 1, first_333))]
    expect += [(first_333, 999996), (first_5, 999997)]
    assert out == expect
    assert stop_tn == first_5


def test_no_a_triggers_and_last_step():
    out, stop_tn = _run(0, a_triggers=False, twrite=4, t_stop=10)
    assert out == [(0, 1), (4, 5), (8, 9), (9, 10)]
    assert stop_tn is None


def test_a_beyond_one_stops():
    import ifdm_code_amd as F
    trig = F.OutputTriggers(0, True)
    assert trig(1.2, 5, 100, 1000) == (999999, True)


def test_oracle_trigger_chain_matches():
    """The oracle's restatement of the trigger chain (oracle.Triggers, used
    for the 64^3 driver fixture) gives the same hand-worked sequence."""
    from oracle import oracle as O
    t = O.Triggers()
    out, stop_tn = [], None
    for tn in range(1000):
        a = 0.021 + 0.03 * (tn + 1)
        w, stop = t(a, tn, 100, 1000)
        if w is not None:
            out.append((tn, w))
        if stop:
            stop_tn = tn
            break
    expect, lib_stop = _run(0)
    assert out == expect and stop_tn == lib_stop
