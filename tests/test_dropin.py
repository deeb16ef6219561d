"""The drop-in drivers ch_{1,2,3}D (cahn-hilliard_amd/host) run as the
reference's own binaries do in its test-suite (tests/test_landau.py:26-65):
the same TOML input, the same output files, the same comparisons.

CPU tests cover configuration errors that stop before any device work; the
GPU tests rerun the reference's regression and mass tests on our binaries.
"""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import REPO

BIN = os.path.join(REPO, "cahn-hilliard_amd", "bin")
GOLD = os.path.join(REPO, "tests", "golden")


def run_bin(dims, cfg_text, tmp_path, check=True):
    cfg = tmp_path / "input.toml"
    cfg.write_text(cfg_text)
    exe = os.path.join(BIN, f"ch_{dims}D")
    return subprocess.run([exe, str(cfg)], cwd=tmp_path, capture_output=True, text=True, check=check, timeout=300)


def simple_landau(integrator):
    with open(os.path.join(GOLD, "simple_landau.toml")) as f:
        return f'\nintegrator = "{integrator}"\n' + f.read()


@pytest.mark.parametrize("text,msg", [
    ("steps = 10\nN = 48\ndt = 0.1\nk = 1.0\nfree_energy = \"landau\"\ninitial_density = 0.0\n[landau]\nepsilon = 0.9\n",
     "N should be a power of 2"),
    ("steps = 10\nN = 16\ndt = 0.1\nk = 1.0\nfree_energy = \"landau\"\n[landau]\nepsilon = 0.9\n",
     "Either 'initial_density' or 'load_from' should be specified"),
    ("steps = 10\nN = 16\ndt = 0.1\nfree_energy = \"landau\"\ninitial_density = 0.0\n[landau]\nepsilon = 0.9\n",
     "Mandatory option 'k' not found"),
    ("steps = 10\nN = 16\ndt = 0.1\nk = 1.0\nfree_energy = \"landau\"\ninitial_density = 0.0\n[landau]\n",
     "requires either 'epsilon' or 'T'"),
    ("steps = 10\nN = 16\ndt = 0.1\nk = 1.0\nfree_energy = \"foo\"\ninitial_density = 0.0\n",
     "Unsupported free energy model 'foo'"),
    ("steps = 10\nN = 16 16\n", "Parsing failed"),
])
def test_config_errors_before_device_work(tmp_path, text, msg):
    r = run_bin(2, text, tmp_path, check=False)
    assert r.returncode == 1
    assert msg in (r.stderr + r.stdout)


def test_toml_subset(tmp_path):
    """Dotted keys, inline tables, arrays, 1e7-style integers and comments parse
    (the run then stops at the power-of-two check, before any device work)."""
    text = ("steps = 1e7  # float literal for an integer option\nN = 24\ndt = 1e-3\nk = [1.0]\n"
            "free_energy = \"simple_wertheim\"\ninitial_density = 5e-5\nwertheim.valence = 4\n"
            "[wertheim]\nB2 = 2190\ndelta = { T = 300, deltaH = -42200, deltaS = -119.1, salt = 0.5, sticky_size = 6 }\n")
    r = run_bin(3, text, tmp_path, check=False)
    assert r.returncode == 1 and "N should be a power of 2" in r.stderr
    assert "valence = 4" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", ["euler", "pseudospectral"])
@pytest.mark.parametrize("dims", [1, 2, 3])
def test_reference_regression_on_dropin_binary(tmp_path, integrator, dims):
    """tests/test_landau.py:41-65 of the reference, pointed at ch_{dims}D."""
    run_bin(dims, simple_landau(integrator), tmp_path)
    phi0 = np.loadtxt(tmp_path / "init_0.dat")
    phi1 = np.loadtxt(tmp_path / "last_0.dat")
    assert abs(phi0.sum() - phi1.sum()) < 1e-4
    ref = np.load(os.path.join(GOLD, "reference_landau_fixtures.npz"))[f"{integrator}_{dims}D"]
    np.testing.assert_allclose(phi1.ravel(), ref, rtol=1e-12, atol=1e-14)
    # the header line the reference writes (src/utils/Printer.cpp:183), as in its
    # own fixture tests/data/reference/landau_euler_2D.dat: "# step = 100, t = 0.10000, size = 64x64x1"
    head = (tmp_path / "last_0.dat").read_text().split("\n", 1)[0]
    assert head.startswith("# step = 100, t = 0.10000, size = 64x") and head.endswith("| density")


@pytest.mark.gpu
def test_energy_file_and_restart(tmp_path):
    """energy.dat cadence (a line *before* every print_every-th step), FP64
    observables, and a load_from restart continuing the step counter."""
    text = ("integrator = \"euler\"\nsteps = 300\nprint_every = 100\nprint_last_every = 150\nN = 32\ndt = 0.01\n"
            "k = 1.0\nfree_energy = \"landau\"\ninitial_density = 0.0\ninitial_A = 0.3\nseed = 123456\n"
            "[landau]\nepsilon = 0.9\n")
    r = run_bin(2, text, tmp_path)
    lines = (tmp_path / "energy.dat").read_text().split("\n")
    lines = [l for l in lines if l]
    assert [l.split()[3] for l in lines] == ["0", "100", "200"]
    # "{:.5}" of fmt 7.1.3 (src/main.cpp:121): trailing zeros kept (tests/golden/fmt_vectors.json)
    assert lines[0].split()[0] == "0.0000" and lines[1].split()[0] == "1.0000"
    assert len(lines[0].split()[1].replace("-", "").split("e")[0].replace(".", "").lstrip("0")) == 8
    assert r.stdout.split("\n")[0] == lines[0]
    import sys
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    p = O.Params(dims=2, N=32, dt=0.01, k=1.0)
    rho0 = O.init_density(2, 32, 0.0, 0.3, 0, 123456)
    F0, m0 = float(lines[0].split()[1]), float(lines[0].split()[2])
    assert F0 == pytest.approx(O.average_free_energy(p, O.Model(epsilon=0.9), rho0), rel=1e-7)
    assert m0 == pytest.approx(O.average_mass(p, rho0), rel=1e-4, abs=1e-9)
    # last_0.dat written at t = 150 (print_last_every) then overwritten at the end
    head = (tmp_path / "last_0.dat").read_text().split("\n", 1)[0]
    assert head.startswith("# step = 300,")
    # restart from it: step counter continues at 300
    rs = tmp_path / "restart"
    rs.mkdir()
    shutil.copy(tmp_path / "last_0.dat", rs / "start.dat")
    text2 = text.replace("initial_density = 0.0\n", "load_from = \"start.dat\"\n").replace("steps = 300", "steps = 100")
    run_bin(2, text2, rs)
    e2 = [l for l in (rs / "energy.dat").read_text().split("\n") if l]
    assert e2[0].split()[3] == "300"


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", ["euler", "pseudospectral"])
def test_lossless_restart_is_exact(tmp_path, integrator):
    """output.lossless = true: 200 steps, restart from last_0.dat, 100 more steps
    against 300 steps in one run (the 6-digit default cannot match).  Euler's
    state is the field itself: same bits.  The pseudospectral state is rho_hat,
    which the restart rebuilds as r2c(rho): equal to round-off."""
    base = (f"integrator = \"{integrator}\"\nN = 64\ndt = 0.01\nk = 1.0\nfree_energy = \"landau\"\n"
            "initial_A = 0.3\nseed = 123456\nprint_every = 0\n[output]\nlossless = true\n[landau]\nepsilon = 0.9\n")
    one = tmp_path / "one"
    one.mkdir()
    run_bin(2, "steps = 300\ninitial_density = 0.0\n" + base, one)
    a = tmp_path / "a"
    a.mkdir()
    run_bin(2, "steps = 200\ninitial_density = 0.0\n" + base, a)
    b = tmp_path / "b"
    b.mkdir()
    shutil.copy(a / "last_0.dat", b / "start.dat")
    run_bin(2, "steps = 100\nload_from = \"start.dat\"\n" + base, b)
    full = np.loadtxt(one / "last_0.dat")
    resumed = np.loadtxt(b / "last_0.dat")
    if integrator == "euler":
        assert np.array_equal(full, resumed)
    else:
        assert np.abs(full - resumed).max() <= 1e-12 * np.abs(full).max()
    assert len((one / "last_0.dat").read_text().split("\n")[1].split()[0]) > 12


GENERIC_TOML = """steps = {steps}
print_every = {every}
N = {N}
dt = 1e-4
dx = 1.0
initial_density = [0.02, 0.03, 0.01]
initial_A = 0.05
free_energy = "generic_wertheim"
k = [1.0, 1.5, 2.0]
seed = 123456
output.lossless = true

[generic_wertheim]
allow_unspecified_B2s = false

[[generic_wertheim.species]]
patches = [0, 0, 0, 0]

[[generic_wertheim.species]]
patches = [1, 1, 1, 1]

[[generic_wertheim.species]]
patches = [0, 1]

[[generic_wertheim.B2s]]
interaction = "0-0"
value = 0.7
[[generic_wertheim.B2s]]
interaction = "1-1"
value = 0.7
[[generic_wertheim.B2s]]
interaction = "2-2"
value = 0.7
[[generic_wertheim.B2s]]
interaction = "0-1"
value = 0.7
[[generic_wertheim.B2s]]
interaction = "0-2"
value = 0.7
[[generic_wertheim.B2s]]
interaction = "1-2"
value = 0.7

[[generic_wertheim.deltas]]
interaction = "0-0"
value = 1000.0

[[generic_wertheim.deltas]]
interaction = "1-1"
T = 293.15
deltaH = -42200
deltaS = -119.1
salt = 0.5
sticky_size = 6
"""


@pytest.mark.parametrize("model,extra", [
    ("generic_wertheim", ""),
    ("saleh", '[saleh]\nvalence = [4, 4, 2]\nB2 = 0.7\ndelta_AA = 1e3\ndelta_BB = { value = 2e3 }\n'),
    ("ricci", '[ricci]\nB2 = 0.7\ndelta_00 = 1e3\ndelta_12 = 5e2\n'),
])
def test_multi_species_configs_parse(tmp_path, model, extra):
    """[[table]] arrays, per-species arrays and the three models' keys are read
    (the run then stops at the power-of-two check, before any device work)."""
    if model == "generic_wertheim":
        text = GENERIC_TOML.format(steps=10, every=0, N=24)
    else:
        text = (f'steps = 10\nN = 24\ndt = 1e-4\ninitial_density = 0.02\nfree_energy = "{model}"\nk = 1.0\n' + extra)
    r = run_bin(2, text, tmp_path, check=False)
    assert r.returncode == 1 and "N should be a power of 2" in r.stderr, r.stderr[-2000:]
    if model == "generic_wertheim":
        assert "Adding 0-0 interaction with delta = 1000" in r.stderr
        assert "Adding 1-2 virial coefficient" in r.stderr


@pytest.mark.gpu
def test_generic_wertheim_on_dropin_binary(tmp_path):
    """A three-species GenericWertheim run through ch_2D: one output file per
    species, every species' mass conserved, finite energies."""
    run_bin(2, GENERIC_TOML.format(steps=200, every=50, N=32), tmp_path)
    for s in range(3):
        a = np.loadtxt(tmp_path / f"init_{s}.dat")
        b = np.loadtxt(tmp_path / f"last_{s}.dat")
        assert a.shape == (32, 32) and np.isfinite(b).all()
        assert abs(a.sum() - b.sum()) < 1e-12 * a.sum()  # 17-digit output (output.lossless)
        assert np.abs(a - b).max() > 0
    e = np.loadtxt(tmp_path / "energy.dat")
    assert e.shape[0] == 4 and np.isfinite(e).all()  # t = 0, 50, 100, 150 (main.cpp:104-127)


LANDAU_BASE = ("integrator = \"euler\"\nN = 32\ndt = 0.01\nk = 1.0\nfree_energy = \"landau\"\n"
               "initial_density = 0.0\ninitial_A = 0.3\nseed = 7\n[landau]\nepsilon = 0.9\n")


def native_frames(path):
    """(step, header, values) of each frame of a native trajectory file."""
    frames, cur = [], None
    for line in path.read_text().split("\n"):
        if line.startswith("# step = "):
            cur = [int(line.split()[3].rstrip(",")), line, []]
            frames.append(cur)
        elif line.strip():
            cur[2].extend(float(v) for v in line.split())
    return [(s, h, np.array(v)) for s, h, v in frames]


@pytest.mark.gpu
def test_vtk_output(tmp_path):
    """output.print_vtk (src/utils/Printer.cpp:209-262): init_/last_ VTK files and
    one trajectory file per frame, STRUCTURED_POINTS ASCII, values in
    std::scientific equal to the native files' values."""
    traj = tmp_path / "traj"
    traj.mkdir()
    text = ("steps = 40\nprint_trajectory_every = 20\n" + LANDAU_BASE.replace("N = 32", "N = 16") +
            f"[output]\nprint_vtk = true\ntrajectory_path = \"{traj}\"\n")
    run_bin(2, text, tmp_path)
    for name in ("init_0.vtk", "init_density.vtk", "last_0.vtk", "last_density.vtk"):
        assert (tmp_path / name).exists(), name
    assert sorted(p.name for p in traj.iterdir()) == ["traj_0_0.vtk", "traj_0_20.vtk"]
    lines = (tmp_path / "last_0.vtk").read_text().split("\n")
    assert lines[0] == "# vtk DataFile Version 3.0"
    assert lines[1] == "step = 40, t = 0.40000, size = 16x16x1 | density"
    assert lines[2:9] == ["ASCII", "DATASET STRUCTURED_POINTS", "DIMENSIONS 16 16 1", "ORIGIN 0 0 0",
                          "SPACING 1 1 1", "POINT_DATA 256", "SCALARS species_0 double 1"]
    assert lines[9] == "LOOKUP_TABLE default"
    vals = [l for l in lines[10:] if l]
    assert len(vals) == 256 and all("e" in v for v in vals)
    native = np.loadtxt(tmp_path / "last_0.dat").ravel()
    np.testing.assert_allclose(np.array([float(v) for v in vals]), native, rtol=1e-5)
    # the 2D trajectory frame at t = 20 names its step in the header
    assert (traj / "traj_0_20.vtk").read_text().split("\n")[1].startswith("step = 20, t = 0.20000,")


@pytest.mark.gpu
def test_log_trajectory_cadence(tmp_path):
    """print_trajectory_strategy = "log" (Printer.cpp:145-155): frames at
    t = round(log_n0 log_fact^n) for n = 0, 1, 2, ... (the counter advances
    per frame; the reference never advances it and writes only t = log_n0)."""
    text = ("steps = 50\nprint_trajectory_strategy = \"log\"\nprint_last_every = 25\n" + LANDAU_BASE +
            "[output]\nlog_n0 = 2\nlog_fact = 1.5\n")
    run_bin(2, text, tmp_path)
    steps = [s for s, _, _ in native_frames(tmp_path / "traj_0.dat")]
    want = []
    n = 0
    # C's round(): halves away from zero (round(4.5) = 5), unlike Python's round
    while math.floor(2 * 1.5 ** n + 0.5) < 50:
        want.append(math.floor(2 * 1.5 ** n + 0.5))
        n += 1
    assert steps == want


def landau_pressure(rho, N, k=1.0, eps=0.9, dx=1.0):
    """p = psi f'(psi) - f(psi) - k/2 sum_d (forward difference)^2 (CahnHilliard.cpp:322-361)."""
    f = -0.5 * eps * rho ** 2 + 0.25 * rho ** 4
    fp = -eps * rho + rho ** 3
    r = rho.reshape(N, N)
    g2 = ((np.roll(r, -1, 1) - r) / dx) ** 2 + ((np.roll(r, -1, 0) - r) / dx) ** 2
    return (rho * fp - f) - 0.5 * k * g2.ravel()


@pytest.mark.gpu
@pytest.mark.parametrize("strategy", ["linear", "log"])
def test_pressure_trajectory_cadence(tmp_path, strategy):
    """output.print_pressure (Printer.cpp:33-46, 157-170; main.cpp:114-128):
    pressure frames every print_pressure_every steps (linear) or on the log
    schedule, one pressure_<species>.dat per species holding the pressure
    field (the reference's add_to_trajectory writes the density there
    instead); the energy lines gain the average pressure column."""
    out = ("[output]\nprint_pressure = true\nprint_pressure_every = 20\n" if strategy == "linear" else
           "[output]\nprint_pressure = true\nprint_pressure_strategy = \"log\"\nlog_n0 = 1\nlog_fact = 3\n")
    text = "steps = 60\nprint_every = 20\n" + LANDAU_BASE + out
    run_bin(2, text, tmp_path)
    frames = native_frames(tmp_path / "pressure_0.dat")
    steps = [s for s, _, _ in frames]
    assert steps == ([0, 20, 40] if strategy == "linear" else [1, 3, 9, 27])
    assert frames[0][1].endswith("| pressure")
    if strategy == "linear":
        rho0 = np.loadtxt(tmp_path / "init_0.dat").ravel()
        np.testing.assert_allclose(frames[0][2], landau_pressure(rho0, 32), rtol=2e-5, atol=1e-9)
    lines = [l for l in (tmp_path / "energy.dat").read_text().split("\n") if l]
    assert [len(l.split()) for l in lines] == [5, 5, 5]
    if strategy == "linear":
        assert float(lines[0].split()[4]) == pytest.approx(frames[0][2].mean(), rel=1e-4, abs=1e-9)


@pytest.mark.gpu
def test_pressure_strategy_rejected(tmp_path):
    """The Printer reads its keys after the system (and its device context) is
    built, as the reference's Manager does (src/main.cpp:20-89)."""
    text = "steps = 1\n" + LANDAU_BASE + "[output]\nprint_pressure = true\nprint_pressure_strategy = \"cubic\"\n"
    r = run_bin(2, text, tmp_path, check=False)
    assert r.returncode != 0 and "Unsupported pressure printing strategy 'cubic'" in (r.stdout + r.stderr)


def _load_from_rows(path, N):
    """The rows a 2D load_from reads (src/CahnHilliard.cpp:72-132): lines not
    starting with '#', N whitespace-separated values each, N of them."""
    rows = [ln.split() for ln in open(path).read().split("\n") if ln.strip() and not ln.lstrip().startswith("#")]
    assert len(rows) == N and all(len(r) == N for r in rows)
    return np.array([[float(v) for v in r] for r in rows])


def test_replica_files_round_trip(ch, tmp_path):
    """write_replicas: each replica's field as a native file that load_from
    reads back exactly (17 digits), and an input naming it with the replica's
    epsilon."""
    import tomli
    from cahn_hilliard_amd.init import replica_batch, write_replicas
    Ns, eps, fields = replica_batch(sizes=(64, 128), per_size=2)
    tomls = write_replicas(str(tmp_path), Ns, eps, fields, steps=10, print_every=0)
    assert len(tomls) == 4
    for r, t in enumerate(tomls):
        cfg = tomli.loads(open(t).read())
        assert cfg["N"] == Ns[r] and cfg["landau"]["epsilon"] == eps[r] and cfg["integrator"] == "euler"
        got = _load_from_rows(cfg["load_from"], Ns[r])
        assert np.array_equal(got.ravel(), fields[r])


@pytest.mark.gpu
def test_replica_file_runs_on_dropin_binary(ch, tmp_path):
    """A C3 replica written by write_replicas, run by the drop-in driver for 50
    steps (lossless output), equals the same replica advanced in an FP64 batch,
    bit for bit."""
    from cahn_hilliard_amd.init import replica_batch, write_replicas
    Ns, eps, fields = replica_batch(sizes=(64, 128), per_size=1)
    tomls = write_replicas(str(tmp_path / "in"), Ns, eps, fields, steps=50, print_every=0)
    ctx = ch.Context(0)
    descs = [ch.make_desc(dims=2, N=N, integrator="euler", epsilon=e, k=1.0, dt=0.01) for N, e in zip(Ns, eps)]
    b = ch.Batch(ctx, descs)
    for r, f in enumerate(fields):
        b.upload(r, f)
    b.evolve(50)
    want = [b.download(r) for r in range(len(Ns))]
    b.close()
    ctx.close()
    for r, t in enumerate(tomls):
        run = tmp_path / f"run{r}"
        run.mkdir()
        run_bin(2, open(t).read() + "\n[output]\nlossless = true\n", run)
        got = np.loadtxt(run / "last_0.dat")
        assert np.array_equal(got.ravel(), want[r]), r

