"""Batched replicas (ch_batch_*): with plain rounding every replica of a
mixed-size batch evolves exactly as the same system run on its own
(bit-identical fields: the batch kernel shares the per-cell arithmetic), FP64
replicas match the oracle to 1e-10, FP32 replicas stay within 1e-4 (relative
L-inf, 200 steps) of FP64, and the per-replica FP64 observables match the
oracle.  The compensated FP32 modes give the same bits for every region split,
and the default one keeps the 256-replica workload within 1e-5 of FP64 over
its 5000 steps."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_linf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(ch):
    c = ch.Context(0)
    yield c
    c.close()


def make_replicas(ch, precision, Ns, seed=0):
    rng = np.random.default_rng(seed)
    descs, fields = [], []
    for N in Ns:
        eps = rng.uniform(0.5, 1.0)
        c = rng.uniform(-0.3, 0.3)
        descs.append(ch.make_desc(dims=2, N=N, integrator="euler", precision=precision, dt=0.01, k=1.0,
                                  epsilon=float(eps)))
        fields.append(c + rng.uniform(-0.01, 0.01, N * N))
    return descs, fields


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_equals_single_systems(ch, ctx, monkeypatch, precision):
    """Plain rounding (CH_COMP_NONE): every region split (CH_BATCH_SHAPE) and
    launch depth (CH_BATCH_TB) gives each replica the bits of the same system
    run alone."""
    Ns = [64, 128, 64, 256, 128, 512, 64]
    descs, fields = make_replicas(ch, precision, Ns)
    for shape, tb in (("0", "1"), ("0", "2"), ("0", "4"), ("1", "4"), ("2", "4"), ("1", "8"), ("2", "8"),
                      ("3", "1"), ("3", "4"), ("3", "8"), ("4", "2"), ("4", "8"), ("5", "2"), ("5", "4"),
                      ("5", "8"), ("6", "1"), ("6", "4"), ("7", "2"), ("7", "4"), ("7", "8"),
                      ("8", "1"), ("8", "2"), ("8", "4"), ("8", "8")):
        monkeypatch.setenv("CH_BATCH_TB", tb)
        monkeypatch.setenv("CH_BATCH_SHAPE", shape)
        b = ch.Batch(ctx, descs, compensation="none")
        assert b.steps_per_launch == int(tb)
        for r, f in enumerate(fields):
            b.upload(r, f)
        b.evolve(77)
        outs = [b.download(r) for r in range(len(Ns))]
        b.close()
        for r, (d, f) in enumerate(zip(descs, fields)):
            s = ch.System(ctx, d)
            s.upload(f)
            s.evolve(77)
            ref = s.download()
            s.close()
            assert np.array_equal(outs[r], ref), (shape, tb, r, Ns[r])


def test_batch_against_oracle(ch, ctx):
    Ns = [64, 128, 256]
    descs64, fields = make_replicas(ch, "fp64", Ns, seed=1)
    descs32, _ = make_replicas(ch, "fp32", Ns, seed=1)
    res = {}
    for name, descs in (("fp64", descs64), ("fp32", descs32)):
        b = ch.Batch(ctx, descs)
        for r, f in enumerate(fields):
            b.upload(r, f)
        b.evolve(200)
        res[name] = ([b.download(r) for r in range(3)], b.observables())
        b.close()
    for r, (N, d) in enumerate(zip(Ns, descs64)):
        p = O.Params(dims=2, N=N, dt=0.01, k=1.0)
        m = O.Model(epsilon=d.epsilon)
        ref = O.run(O.EULER, p, m, fields[r], 200)
        assert rel_linf(res["fp64"][0][r], ref) < 1e-10
        err32 = rel_linf(res["fp32"][0][r], ref)
        print(f"replica N={N} fp32 rel L-inf vs fp64 oracle = {err32:.3e}")
        assert err32 < 1e-4
        mass, fe = res["fp64"][1]
        assert mass[r] == pytest.approx(O.average_mass(p, ref), rel=1e-10, abs=1e-14)
        assert fe[r] == pytest.approx(O.average_free_energy(p, m, ref), rel=1e-10)


def test_batch_rejects_unsupported(ch, ctx):
    with pytest.raises(ch.CHError):
        ch.Batch(ctx, [ch.make_desc(dims=2, N=32)])
    with pytest.raises(ch.CHError):
        ch.Batch(ctx, [ch.make_desc(dims=2, N=64), ch.make_desc(dims=2, N=64, precision="fp32")])
    with pytest.raises(ch.CHError):
        ch.Batch(ctx, [ch.make_desc(dims=3, N=64)])


def test_autotune_keeps_state_and_results(ch, ctx, tmp_path, monkeypatch):
    """The autotuner's trial launches leave the fields unchanged, it picks a
    valid configuration, the tuned batch gives the single-system bits, and the
    result is recorded (CH_TUNE_CACHE) for the next batch of the composition."""
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("CH_TUNE_CACHE", str(cache))
    Ns = [64, 256, 128]
    descs, fields = make_replicas(ch, "fp32", Ns, seed=4)
    b = ch.Batch(ctx, descs, compensation="none")
    for r, f in enumerate(fields):
        b.upload(r, f)
    shape, k = b.autotune()
    print(f"autotuned: shape {shape}, {k} steps per launch")
    assert shape in range(9) and k in (2, 4, 8) and b.steps_per_launch == k
    for r, f in enumerate(fields):
        assert np.array_equal(b.download(r), f.astype(np.float32).astype(np.float64))
    b.evolve(45)
    outs = [b.download(r) for r in range(len(Ns))]
    b.close()
    for r, (d, f) in enumerate(zip(descs, fields)):
        s = ch.System(ctx, d)
        s.upload(f)
        s.evolve(45)
        assert np.array_equal(outs[r], s.download())
        s.close()
    lines = cache.read_text().split("\n")
    assert any(ln.endswith(f" {shape} {k}") and "/c0/" in ln and "/64x1/128x1/256x1" in ln for ln in lines), lines
    b2 = ch.Batch(ctx, descs, compensation="none")
    assert b2.steps_per_launch == k
    b2.close()


def test_compensation_modes(ch, ctx):
    """FP32 batches default to per-launch compensation, FP64 batches take none;
    invalid requests are refused."""
    d32 = [ch.make_desc(dims=2, N=64, precision="fp32"), ch.make_desc(dims=2, N=128, precision="fp32")]
    d64 = [ch.make_desc(dims=2, N=64), ch.make_desc(dims=2, N=128)]
    b = ch.Batch(ctx, d32)
    assert b.compensation == "launch" and b.steps_per_launch >= 4
    b.close()
    b = ch.Batch(ctx, d64)
    assert b.compensation == "none"
    b.close()
    for mode in ("launch", "stored", "low8"):
        with pytest.raises(ch.CHError):
            ch.Batch(ctx, d64, compensation=mode)
    import ctypes
    from cahn_hilliard_amd import _lib as L
    arr = (L.ChDesc * 2)(*d32)
    h = ctypes.c_void_p()
    assert L.lib().ch_batch_create_ex(ctx.handle, 2, arr, 7, ctypes.byref(h)) == 1  # CH_EINVAL


@pytest.mark.parametrize("mode", ["launch", "low8", "stored"])
def test_compensated_split_invariance(ch, ctx, monkeypatch, mode):
    """At a fixed launch depth every region split gives the same bits (the
    compensation is a per-cell computation), and after 141 steps the FP32
    fields are within 1e-6 of FP64.  The stored mode does not depend on the
    depth either."""
    rng = np.random.default_rng(5)
    Ns = [64, 128, 64, 256, 128, 512, 64]
    eps = rng.uniform(0.5, 1.0, len(Ns))
    fields = [rng.uniform(-0.3, 0.3) + rng.uniform(-0.01, 0.01, N * N) for N in Ns]
    d64 = [ch.make_desc(dims=2, N=N, integrator="euler", dt=0.01, epsilon=float(e)) for N, e in zip(Ns, eps)]
    d32 = [ch.make_desc(dims=2, N=N, integrator="euler", precision="fp32", dt=0.01, epsilon=float(e))
           for N, e in zip(Ns, eps)]
    b = ch.Batch(ctx, d64)
    for r, f in enumerate(fields):
        b.upload(r, f)
    b.evolve(141)
    ref = [b.download(r) for r in range(len(Ns))]
    b.close()
    outs = {}
    for tb in ("4", "8"):
        for shape in ("0", "1", "2", "3", "4", "5", "6", "7", "8"):
            monkeypatch.setenv("CH_BATCH_SHAPE", shape)
            monkeypatch.setenv("CH_BATCH_TB", tb)
            b = ch.Batch(ctx, d32, compensation=mode)
            assert b.compensation == mode and b.steps_per_launch == int(tb)
            for r, f in enumerate(fields):
                b.upload(r, f)
            b.evolve(77)
            b.evolve(64)
            outs[shape, tb] = [b.download(r) for r in range(len(Ns))]
            b.close()
    for (shape, tb), v in outs.items():
        base = outs["0", "4" if mode == "stored" else tb]
        assert all(np.array_equal(a, c) for a, c in zip(v, base)), (mode, shape, tb)
        err = max(rel_linf(a, c) for a, c in zip(v, ref))
        assert err < 1e-6, (mode, shape, tb, err)


def c3_errors(ch, ctx, compensation=None, marks=(1000, 5000)):
    """The 256-replica Landau workload (64 each at N = 64, 128, 256, 512; per
    replica drand48 seeded 123456 + r: eps ~ U[0.5, 1], psi0 ~ c + U[-0.01, 0.01]
    with c ~ U[-0.3, 0.3]): the worst replica's FP32 relative L-inf error
    against the FP64 batch after each of `marks` steps."""
    from cahn_hilliard_amd.init import replica_batch
    Ns, eps, fields = replica_batch()
    outs = {}
    for prec, comp in (("fp64", None), ("fp32", compensation)):
        descs = [ch.make_desc(dims=2, N=N, integrator="euler", precision=prec, dt=0.01, dx=1.0, k=1.0, epsilon=e)
                 for N, e in zip(Ns, eps)]
        b = ch.Batch(ctx, descs, compensation=comp)
        for r, f in enumerate(fields):
            b.upload(r, f)
        done = 0
        for t in marks:
            b.evolve(t - done)
            done = t
            outs[prec, t] = [b.download(r) for r in range(len(Ns))]
        b.close()
    return {t: max(rel_linf(a, c) for a, c in zip(outs["fp32", t], outs["fp64", t])) for t in marks}


@pytest.mark.parametrize("mode, bound", [(None, 1e-5), ("low8", 3e-6), ("stored", 3e-6)])
def test_c3_full_length_fp32_error(ch, ctx, mode, bound):
    """The 256-replica workload over its full 5000 steps: the default FP32 batch
    (per-launch compensation) within 1e-5 of FP64 (measured 9.9e-6), the 8-bit
    low-order and the stored-compensation modes within 3e-6 (1.4e-6)."""
    errs = c3_errors(ch, ctx, mode)
    print(f"C3 {mode or 'default'}: fp32 rel L-inf vs fp64 = {errs[1000]:.3e} after 1000 steps, "
          f"{errs[5000]:.3e} after 5000")
    assert errs[1000] < 1e-6
    assert errs[5000] < bound


@pytest.mark.xfail(strict=True, reason="plain FP32 rounding of every update accumulates to 5.6e-5 over the "
                                       "5000 steps (the compensated modes exist for this)")
def test_c3_plain_fp32_error(ch, ctx):
    errs = c3_errors(ch, ctx, "none")
    print(f"C3 plain: fp32 rel L-inf vs fp64 = {errs[1000]:.3e} after 1000 steps, {errs[5000]:.3e} after 5000")
    assert errs[5000] < 1e-5
