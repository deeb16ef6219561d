"""ch_observables_device_noalloc: the observables of a caller-held device field
with no allocation (the partial sums in the library's fixed per-device scratch
array).  Checked through the C ABI against ch_sys_observables on the same
field -- same kernels, so the same bits -- on consecutive calls, the second
with a smaller field, in FP64 and FP32, for both single-species models."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(ch):
    c = ch.Context(0)
    yield c
    c.close()


def system_observables(ch, ctx, desc, rho):
    s = ch.System(ctx, desc)
    s.upload(rho)
    out = s.observables()
    s.close()
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("model", ["landau", "simple_wertheim"])
def test_noalloc_matches_system_observables(ch, ctx, precision, model):
    import torch
    rng = np.random.default_rng(11)
    kw = dict(epsilon=0.9) if model == "landau" else dict(B2=1.0, delta=1e4, valence=4)
    # two calls in a row: 3D N=64 (2.6e5 cells, all 1024 partials used), then 2D N=32
    for dims, N in ((3, 64), (2, 32)):
        desc = ch.make_desc(dims=dims, N=N, integrator="euler", free_energy=model, precision=precision, k=1.0,
                            dx=0.8, **kw)
        base = 0.0 if model == "landau" else 0.05
        rho = base + 0.04 * (2 * rng.random(N ** dims) - 1)
        if precision == "fp32":
            rho = rho.astype(np.float32).astype(np.float64)
        dev = torch.tensor(rho, dtype=torch.float64 if precision == "fp64" else torch.float32, device="cuda")
        torch.cuda.synchronize()
        got = ch.observables_device_noalloc(ctx, desc, dev.data_ptr())
        want = system_observables(ch, ctx, desc, rho)
        print(f"{model} {precision} {dims}D N={N}: noalloc {got}, system {want}")
        assert got == want


def test_noalloc_rejects(ch, ctx):
    import torch
    dev = torch.zeros(16, dtype=torch.float64, device="cuda")
    with pytest.raises(ch.CHError):  # 2^33 cells: refused before any device work
        ch.observables_device_noalloc(ctx, ch.make_desc(dims=3, N=2048), dev.data_ptr())
    with pytest.raises(ch.CHError):
        ch.observables_device_noalloc(ctx, ch.make_desc(dims=2, N=4), 0)
