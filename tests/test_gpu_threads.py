"""Engines used from several host threads at once (one engine per thread, as the reference's
engine contract allows): each engine has its own HIP stream and buffers, so results are bitwise
those of the same calls made one after another."""
import threading

import numpy as np
import pytest

from hawkes_amd import Engine, HphError, synth

pytestmark = pytest.mark.gpu


def run_engine(locs, times, params):
    e = Engine(2, len(times))
    e.update_locations(locs)
    e.set_times(times)
    out = []
    for p in params:
        e.set_parameters(p)
        ll, g = e.log_likelihood_and_gradient()
        out.append((ll, g, e.probs_self_excite()))
    e.close()
    return out


def test_engines_in_parallel_threads():
    data = [synth.cluster_process(2500 + 700 * s, seed=60 + s) for s in range(4)]
    params = [synth.jitter_params(5, seed=70 + s) for s in range(4)]
    sequential = [run_engine(l, t, p) for (l, t), p in zip(data, params)]
    results = [None] * 4
    errors = []

    def worker(s):
        try:
            results[s] = run_engine(data[s][0], data[s][1], params[s])
        except Exception as exc:  # reported below
            errors.append(exc)

    threads = [threading.Thread(target=worker, args=(s,)) for s in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not errors, errors
    for s in range(4):
        for (a_ll, a_g, a_p), (b_ll, b_g, b_p) in zip(results[s], sequential[s]):
            assert a_ll == b_ll and np.array_equal(a_g, b_g) and np.array_equal(a_p, b_p)


def test_last_error_is_per_thread():
    """An error raised in one thread does not overwrite the message seen in another."""
    messages = {}

    def bad(name, params):
        e = Engine(2, 10)
        for _ in range(200):
            try:
                e.set_parameters(params)
            except HphError as exc:
                messages.setdefault(name, set()).add(str(exc))
        e.close()

    t1 = threading.Thread(target=bad, args=("length", [1.0, 2.0]))
    t2 = threading.Thread(target=bad, args=("finite", [1.0, 1.0, float("nan"), 1.0, 1.0, 1.0]))
    t1.start()
    t2.start()
    t1.join(timeout=60)
    t2.join(timeout=60)
    assert all("length" in m for m in messages["length"])
    assert all("finite" in m for m in messages["finite"])
