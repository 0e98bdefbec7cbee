"""What the learned warm start gains over the method without it (the freestream fill), on the
GPU: the network is trained by the library's own kernels on steady fields of the batched GPU
DSMC, then measured on held-out cases against their converged freestream-initialised fields."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_warm_start_beats_freestream_on_held_out_cases():
    from dsmc_erlend_amd.surrogate import Surrogate
    from dsmc_erlend_amd.sweep import (draw_cases, field_error, freestream_fields,
                                       reference_fields, run_to_steady, steady_fields)
    R = 32
    train = draw_cases(24, seed=1, grid=R, npc=30, seed_base=100)
    ftrain, rate = steady_fields(train, 3000, 1500)
    assert rate > 0 and all(np.isfinite(f).all() for f in ftrain)
    with Surrogate(height=R, width=R, max_batch=8, seed=2) as net:
        net.set_dataset(train, np.stack(ftrain))
        hist = net.train(epochs=60, batch=8, lr=2e-3, seed=3)
        test = draw_cases(4, seed=2, grid=R, npc=30, seed_base=200)
        pred = net.predict(test)
    assert hist[-6:].mean() < 0.5 * hist[:3].mean()
    refs, refs_b = reference_fields(test, 6000, 2000, bins=16, seed=77)
    e_net = np.array([max(field_error(pred[i], refs[i])) for i in range(len(test))])
    e_free = np.array([max(field_error(freestream_fields(p), refs[i])) for i, p in enumerate(test)])
    print("held-out field error, network:", e_net, "freestream:", e_free)
    assert e_net.mean() < 0.6 * e_free.mean(), (e_net, e_free)
    # seeded from the prediction, the average over the first 300 steps is closer to the
    # converged field (on 16 x 16 blocks) than the same average from the freestream fill
    _, _, warm = run_to_steady(test, refs_b, 0.0, 300, 300, 300, 16, init_fields=pred)
    _, _, free = run_to_steady(test, refs_b, 0.0, 300, 300, 300, 16)
    e_w = np.array([max(h[-1][1]) for h in warm])
    e_f = np.array([max(h[-1][1]) for h in free])
    print("running-average error after 300 steps, warm:", e_w, "freestream:", e_f)
    assert e_w.mean() < e_f.mean(), (e_w, e_f)
