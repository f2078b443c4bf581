"""The learned warm start end to end on the GPU, at a size that trains in seconds.

Pairs come from the GPU solver (64 x 64 box, 12 seeded cyclone tracks cycling through
plain, heterogeneous and leads ice; targets are the velocity mEVP reaches after up to
6000 subcycles); the network is trained by the library's own kernels; the last three
tracks are held out.

What the network gains over the method without it (mEVP starting from u0, the previous
timestep's velocity, whose predicted change is zero): its held-out prediction error of
the converged velocity change is a fraction of the cold start's.

What the residual gate does with it: the momentum residual passes the velocity through
the VP law, where any strain-rate noise well above DeltaMin = 2e-9 1/s turns into O(P)
stress, so a prediction that is close in L2 can still carry a larger residual than u0.
The gate then restores the cold start exactly, and the residual-targeted solve from it
is the plain one.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def trained():
    from nextsimdg_amd import learned as L

    nx = 64
    seeds = list(range(12))
    X, T, W, stats = L.generate_pairs(nx, seeds, 3, steps_between=2, spinup=2, converge=6000, r_stop=0.05,
                                      check_every=1000)
    r0 = np.median([s[3] for s in stats])
    rT = np.median([s[5] for s in stats])
    _say(f"  pairs {X.shape[0]}: median residual of the start {r0:.3e}, of the targets {rT:.3e}")
    assert rT < 0.2 * r0
    net = L.WarmStartNet(3)
    ntest = 9
    train, test = list(range(X.shape[0] - ntest)), list(range(X.shape[0] - ntest, X.shape[0]))
    hist = L.train(net, X, T, W, epochs=60, batch=3, lr=2e-3, seed=1, idx=train)
    _say(f"  training loss {hist[0]:.4e} -> {hist[-1]:.4e}")
    return net, X, T, W, test


def test_network_beats_the_cold_start_on_held_out_tracks(trained):
    from nextsimdg_amd import learned as L

    net, X, T, W, test = trained
    ln, lz = L.evaluate(net, X, T, W, test)
    _say(f"  held-out loss: network {ln:.4e}, cold start {lz:.4e}, ratio {ln / lz:.3f}")
    assert ln < 0.3 * lz


def test_gate_keeps_the_better_start_and_restores_the_cold_one_exactly(trained):
    from nextsimdg_amd import learned as L

    net = trained[0]
    k, off = L.new_track(64, 77, "heterogeneous")
    for s in range(2):
        L.plain_step(k, off, s)
    L.begin_step(k, off, 2)
    u0, v0 = k.get_cg("u"), k.get_cg("v")
    s0 = k.getDGData("s11", 8)
    blob = k.save_state()
    rp, rc, used = L.warm_start(k, net)
    _say(f"  r(predicted) {rp:.4e}  r(cold) {rc:.4e}  kept: {used}")
    assert rc == pytest.approx(k.momentum_residual()[0] if not used else rc)
    if used:
        assert rp <= rc
        assert not np.array_equal(k.get_cg("u"), u0)
    else:
        assert rp > rc
        assert np.array_equal(k.get_cg("u"), u0) and np.array_equal(k.get_cg("v"), v0)
        assert np.array_equal(k.getDGData("s11", 8), s0)
    # the residual-targeted solve from the gated start against plain subcycles from the cold one
    k.load_state(blob)
    k.subcycles(100)
    r_plain = k.momentum_residual()[0]
    k.load_state(blob)
    L.warm_start(k, net)
    done, r = L.subcycles_to_residual(k, r_plain, 400, 10)
    _say(f"  plain r after 100: {r_plain:.4e}; gated start reaches {r:.4e} after {done}")
    assert r <= r_plain and done <= 110
    # the same solve started from the residual the gate computed: the same subcycles, bits
    k.load_state(blob)
    rp2, rc2, used2 = L.warm_start(k, net)
    assert (rp2, rc2, used2) == (rp, rc, used)
    assert L.subcycles_to_residual(k, r_plain, 400, 10, rp2 if used2 else rc2) == (done, r)
