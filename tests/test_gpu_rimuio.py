"""Checkpoints of device runs through Rimu's state files: a run saved with save_state and
resumed from load_state continues bit for bit as if it had never stopped (every random
draw is addressed by (seed, address, attempt, step))."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(rimu, H, **kw):
    return rimu.ProjectorMonteCarloProblem(
        H, style=rimu.IsStochasticInteger(), time_step=1e-3, random_seed=99,
        shift_strategy=rimu.DoubleLogUpdate(target_walkers=5000), **kw)


def test_resume_from_state_file_is_bit_exact(gpu, tmp_path):
    H = gpu.HubbardReal1D(gpu.near_uniform(10, 10), u=6.0)
    full = gpu.solve(_problem(gpu, H, last_step=400))
    fk, fv = full.engine.download()
    fdf = full.df
    full.close()
    part = gpu.solve(_problem(gpu, H, last_step=200))
    f = tmp_path / "state.arrow"
    gpu.save_state(f, part, step=200, shift=part.shift_parameters.shift)
    part.close()
    dv, meta = gpu.load_state(f)
    assert meta["step"] == 200 and isinstance(dv.style, gpu.IsStochasticInteger)
    resumed = gpu.solve(_problem(gpu, H, start_at=dv, shift=meta["shift"], starting_step=200,
                                 last_step=400))
    rk, rv = resumed.engine.download()
    rdf = resumed.df
    resumed.close()
    assert len(rk) > 1000
    assert np.array_equal(fk, rk) and np.array_equal(fv, rv)
    assert rdf["step"].tolist() == list(range(201, 401))
    assert np.array_equal(fdf["shift"].to_numpy()[200:], rdf["shift"].to_numpy())
    assert np.array_equal(fdf["len"].to_numpy()[200:], rdf["len"].to_numpy())


def test_engine_state_round_trip_float(gpu, tmp_path):
    """IsDynamicSemistochastic FP64 state: save, load into a fresh engine, step on."""
    M = N = 12
    H = gpu.HubbardReal1D(gpu.near_uniform(N, M), u=6.0)
    sty = gpu.IsDynamicSemistochastic()

    def engine():
        return gpu._native.Engine(H.descriptor(), sty.descriptor(), 5, 2_000_000, 1 << 20)

    a = engine()
    a.upload(np.array([H.starting_address().bits], dtype=np.uint64), np.array([5000.0]))
    for s in range(1, 41):
        a.step(s, 2.0, 1e-2)
    k, v = a.download()
    f = tmp_path / "s.arrow"
    from rimu_amd import rimuio
    t = rimuio.state_table(M, N, k, v, step=40)
    import pyarrow.ipc as ipc
    with ipc.new_file(str(f), t.schema, options=ipc.IpcWriteOptions(compression="zstd")) as w:
        w.write_table(t)
    _, _, k2, v2, meta = rimuio.read_state_arrays(f)
    assert np.array_equal(k, k2) and np.array_equal(v.view(np.uint64), v2.view(np.uint64))
    b = engine()
    b.upload(k2, v2)
    for s in range(41, 61):
        a.step(s, 2.0, 1e-2)
        b.step(s, 2.0, 1e-2)
    ka, va = a.download()
    kb, vb = b.download()
    a.close()
    b.close()
    assert len(ka) > 1000
    assert np.array_equal(ka, kb) and np.array_equal(va.view(np.uint64), vb.view(np.uint64))
