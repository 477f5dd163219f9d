"""The batched engine's persistent step (one launch per step: every run on one XCD, whose
workgroups take the spawn, sort, merge, copy and final items of its runs from a queue, each
item waiting for its run's previous phase) against its five-launch step: bit-identical
histories, statuses and vectors for runs of 1 to 49 spawn slices, a run whose population
dies and one that outgrows max_length, at every block size, with and without graph
capture, and over the 256-run sweep; and the bounded wait giving up when the host withholds
a run's spawn reports."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("shift", "norm", "len", "vproj", "hproj", "spawn_attempts")


def _mixed(rimu):
    out = []
    for i, (M, N, u, target) in enumerate([(6, 6, 4.0, 500.0), (9, 5, 2.5, 3000.0), (10, 8, 4.0, 4000.0),
                                           (12, 10, 6.0, 16000.0), (14, 12, 3.0, 1e5), (7, 4, 1.0, 2000.0)]):
        h = rimu.HubbardReal1D(rimu.near_uniform(N, M), u=u)
        out.append(rimu.RunSpec(h, target_walkers=target, dtau=1e-3, seed=300 + i))
    # the reference's abort set-ups (test/projector_monte_carlo_problem.jl:252-312)
    out.append(rimu.RunSpec(rimu.HubbardReal1D(rimu.BoseFS(2, 3), u=20.0), target_walkers=1000,
                            dtau=0.01, seed=7, shift=0.0, zeta=0.0, xi=0.0))
    out.append(rimu.RunSpec(rimu.HubbardReal1D(rimu.BoseFS(1, 3, 5, 2, 1)), target_walkers=100,
                            dtau=0.1, seed=7))
    return out


def _run(rimu, sp, nsteps, persistent, block, radix, spg):
    b = rimu.BatchedRuns(sp, history_steps=nsteps)
    b.set_config(block, radix, persistent=persistent)
    b.step(nsteps, steps_per_graph=spg)
    h = b.history()
    st, ln = b.status()
    vecs = [b.vector(r) for r in range(len(sp))]
    b.close()
    return h, st, ln, vecs


@pytest.mark.parametrize("block,radix,spg", [(512, 6, 50), (1024, 8, 1), (256, 4, 25)])
def test_persistent_step_equals_five_launches(gpu, block, radix, spg):
    sp = _mixed(gpu)
    nsteps = 300
    print(f"persistent step: block {block}, radix {radix}, {spg} steps per graph", flush=True)
    # one five-launch configuration for all: results do not depend on the launch shape
    ref = _run(gpu, sp, nsteps, False, 512, 6, 50)
    got = _run(gpu, sp, nsteps, True, block, radix, spg)
    for f in FIELDS:
        np.testing.assert_array_equal(got[0][f], ref[0][f], err_msg=f)
    assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2])
    for r, ((ka, va), (kb, vb)) in enumerate(zip(got[3], ref[3])):
        assert np.array_equal(ka, kb) and np.array_equal(va, vb), f"run {r}"
    # the abort rules fired and the 1e5-walker run spread over 49 slices grew
    assert ref[1][6] == 1 and ref[1][7] == 2 and (ref[1][:6] == 0).all()
    assert ref[0]["norm"][-1, 4] > 1000


def test_persistent_wait_gives_up(gpu):
    """The first run of the first queue never reports its spawn slices (forced from the host):
    its sort's bounded wait gives up, every later item is skipped, every workgroup exits and
    the step reports the error; the runs keep a consistent state (a step is either complete
    or not begun) and step on from it."""
    sp = _mixed(gpu)[:6]
    b = gpu.BatchedRuns(sp, history_steps=10)
    b.set_persistent(True)
    b.step(2, steps_per_graph=1)
    b.set_wait(2000, withhold_first=True)
    with pytest.raises(gpu._native.RimuError, match="gave up"):
        b.step(1, steps_per_graph=1)
    st, ln = b.status()
    assert (st == 0).all() and (ln > 0).all()
    # the batch goes on from that state: every run's vector, length and norm agree with its
    # last history row (a run whose step was cut off kept its vector and slice bounds)
    b.set_wait(1 << 25, withhold_first=False)
    b.step(5, steps_per_graph=1)
    st, ln = b.status()
    h = b.history()
    assert (st == 0).all()
    for r in range(len(sp)):
        keys, vals = b.vector(r)
        assert len(keys) == ln[r] and (keys[1:] > keys[:-1]).all()
        last = h["norm"][:, r].nonzero()[0][-1]
        assert h["len"][last, r] == ln[r] and h["norm"][last, r] == np.abs(vals).sum()
    b.close()


@pytest.mark.parametrize("block", [256, 512, 1024])
def test_persistent_step_equals_five_launches_on_the_sweep(gpu, block):
    """The 256-run U-sweep for 1500 steps: every history entry of the persistent step equals
    the five launches' (hand-overs stay inside one XCD)."""
    specs = gpu.u_sweep_specs(256)
    print(f"sweep, persistent against five launches, block {block}", flush=True)
    out = []
    for p in (False, True):
        b = gpu.BatchedRuns(specs, history_steps=1500)
        b.set_config(block, 6, persistent=p)
        b.step(1500, steps_per_graph=100)
        out.append((b.history(), b.status()[0]))
        b.close()
    (h0, s0), (h1, s1) = out
    assert np.array_equal(s0, s1)
    for f in FIELDS:
        np.testing.assert_array_equal(h1[f], h0[f], err_msg=f)
