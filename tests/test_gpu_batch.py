"""GPU batches of realisations (ensembles / coupling scans): B members of one
lattice size advanced by batched launches must reproduce, member by member,
single-context runs with the same initial conditions and couplings, and the
oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

PHI0, DPHI0 = 1.0093430384226378929425913902459, -0.7137133070120812430962278466136


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def member_ic(n, L, seed, mpl=2e5):
    fld = np.empty((2, n, n, n))
    fldp = np.empty_like(fld)
    fld[0] = PHI0 + O.sample(n, L, -0.25, 1.0, mpl, seed, 0)
    fldp[0] = DPHI0 + O.sample(n, L, 0.25, 1.0, mpl, seed, 1)
    fld[1] = O.sample(n, L, -0.25, 1.0, mpl, seed, 2)
    fldp[1] = O.sample(n, L, 0.25, 1.0, mpl, seed, 3)
    return fld, fldp


@pytest.mark.parametrize("lap", ["spectral", "fd7"])
@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_batch_members_match_single_runs(gpu, lap, precision):
    print("case batch", lap, precision, flush=True)
    n, L, dt, nsteps = 16, 10.0, 1e-3, 4
    g2s = [1e4, 1e5, 2.5e5]
    kw = dict(nfld=2, length=L, time_var=gpu.COSMIC, model=gpu.MODEL_TWOFIELD,
              laplacian=gpu.LAP_SPECTRAL if lap == "spectral" else gpu.LAP_STENCIL,
              precision=gpu.F64 if precision == "f64" else gpu.MIXED)
    ics = [member_ic(n, L, 100 + m) for m in range(len(g2s))]
    singles = []
    for m, g2 in enumerate(g2s):
        with gpu.Lattice(n, model_params=(1.0, g2), **kw) as lat:
            lat.set_state(*ics[m], 1.0, 0.0)
            lat.calc_metric()
            lat.symp4(dt, nsteps)
            singles.append(lat.get_state() + (lat.energies(), lat.output_spectra(1)))
    with gpu.Lattice(n, members=[(1.0, g2) for g2 in g2s], **kw) as batch:
        assert batch.members == len(g2s)
        views = [batch.member(m) for m in range(len(g2s))]
        for m, v in enumerate(views):
            v.set_state(*ics[m], 1.0, 0.0)
            v.calc_metric()
        with pytest.raises(gpu.LatticeError):
            batch.energies()
        batch.symp4(dt, nsteps)
        got = [v.get_state() + (v.energies(), v.output_spectra(1)) for v in views]
        for v in views:
            v.close()
    for m in range(len(g2s)):
        gf, gp, gy, gpy, ge, gs = got[m]
        sf, sp, sy, spy, se, ss = singles[m]
        # same kernels, same per-member grids and reduction order: identical
        assert np.array_equal(gf, sf) and np.array_equal(gp, sp), m
        assert gy == sy and gpy == spy, m
        assert ge["rho"] == se["rho"], m
        assert np.array_equal(gs, ss), m
    # and the members differ from each other (distinct couplings and seeds)
    assert rel(got[0][0], got[2][0]) > 1e-9


def test_batch_member_matches_oracle(gpu):
    print("case batch oracle", flush=True)
    n, L, dt, nsteps = 16, 10.0, 1e-3, 3
    g2s = [5e4, 2e5]
    with gpu.Lattice(n, nfld=2, length=L, members=[(1.0, g2) for g2 in g2s]) as batch:
        views = [batch.member(m) for m in range(2)]
        ics = [member_ic(n, L, 7 + m) for m in range(2)]
        for m, v in enumerate(views):
            v.set_state(*ics[m], 1.0, 0.0)
            v.calc_metric()
        batch.symp(6, dt, nsteps)
        got = [v.get_state() for v in views]
    for m, g2 in enumerate(g2s):
        sim = O.Simulation(n, 2, L, O.COSMIC, model=O.MODEL_TWOFIELD, mp=(1.0, g2))
        sim.set_state(*ics[m], 1.0, 0.0)
        sim.calc_metric()
        sim.symp(6, dt, nsteps)
        assert rel(got[m][0], sim.fld) < 1e-12
        assert rel(got[m][1], sim.fldp) < 1e-12
        assert abs(got[m][2] - sim.yscl) < 1e-13


def test_batch_rejects_per_member_calls_and_gw(gpu):
    with gpu.Lattice(8, nfld=2, length=5.0, members=[(1.0, 1.0), (1.0, 2.0)]) as batch:
        with pytest.raises(gpu.LatticeError):
            batch.set_state(np.zeros((2, 8, 8, 8)), np.zeros((2, 8, 8, 8)))
        with pytest.raises(gpu.LatticeError):
            batch.member(2)
    with pytest.raises(gpu.LatticeError):
        gpu.Lattice(8, nfld=2, length=5.0, gw=True, members=[(1.0, 1.0), (1.0, 2.0)])


def _ens_run(gpu, classes, params, ics, dts, calls, graph):
    with gpu.Ensemble(classes, members=params) as ens:
        ens.use_graph(graph)
        views = [ens.member(i) for i in range(ens.size)]
        for v, ic in zip(views, ics):
            v.set_state(*ic, 1.0, 0.0)
            v.calc_metric()
        for nsteps in calls:
            ens.symp(4, dts, nsteps)
        ens.synchronize()
        out = [v.get_state() for v in views]
        for v in views:
            v.close()
    return out


def test_ensemble_graph_replays_match_eager_and_single_runs(gpu):
    """Two size classes, per-member couplings and seeds; hipGraph capture then
    replay of the same call must equal the eager path bitwise, and each member
    must match a single-context run."""
    print("case ensemble", flush=True)
    classes = [(16, 3, 10.0), (32, 2, 10.0)]
    g2s = [1e4, 5e4, 2.5e5, 1e5, 2e5]
    params = [(1.0, g) for g in g2s]
    sizes = [16, 16, 16, 32, 32]
    ics = [member_ic(n, 10.0, 300 + i) for i, n in enumerate(sizes)]
    dts = [1e-3, 5e-4]
    calls = [2, 2, 3]  # the second call replays the graph captured by the first
    eager = _ens_run(gpu, classes, params, ics, dts, calls, graph=False)
    graph = _ens_run(gpu, classes, params, ics, dts, calls, graph=True)
    for i in range(len(g2s)):
        for a, b in zip(eager[i], graph[i]):
            assert np.array_equal(np.asarray(a), np.asarray(b)), i
    for i, (n, g2) in enumerate(zip(sizes, g2s)):
        with gpu.Lattice(n, nfld=2, length=10.0, model_params=(1.0, g2)) as lat:
            lat.set_state(*ics[i], 1.0, 0.0)
            lat.calc_metric()
            dt = dts[0] if n == 16 else dts[1]
            for nsteps in calls:
                lat.symp4(dt, nsteps)
            ref = lat.get_state()
        assert rel(graph[i][0], ref[0]) < 1e-13, i
        assert rel(graph[i][1], ref[1]) < 1e-12, i
        assert abs(graph[i][2] - ref[2]) < 1e-14, i


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_ensemble_single_launch_matches_per_class_launches(gpu, precision):
    """One launch per sweep for three size classes (a radix-3 size among them)
    against one launch per sweep per class, with the same launch shapes: each
    member keeps its tile loop, radix schedule and reduction slots, so states
    and background scalars agree bitwise, eagerly and through the graph."""
    print("case ensemble single launch", precision, flush=True)
    classes = [(16, 2, 10.0), (48, 1, 10.0), (32, 2, 10.0)]
    g2s = [1e4, 2e5, 5e4, 1e5, 2.5e5]
    params = [(1.0, g) for g in g2s]
    sizes = [16, 16, 48, 32, 32]
    ics = [member_ic(n, 10.0, 400 + i) for i, n in enumerate(sizes)]
    dts = [1e-3, 3e-4, 5e-4]
    shapes = [(1, 1, 1, 1), (0, 0, 0, 0), (0, 1, 0, 1)]
    prec = gpu.F64 if precision == "f64" else gpu.MIXED
    out = {}
    for single, graph in ((True, True), (True, False), (False, False)):
        with gpu.Ensemble(classes, precision=prec, members=params) as ens:
            assert ens.single_launch(None) in (True, False)  # the autotuner's pick
            assert ens.single_launch(single) is single  # these classes allow both
            for k, sh in enumerate(shapes):
                ens.set_tuning(k, sh)
                assert ens.tuning(k) == sh
            ens.use_graph(graph)
            views = [ens.member(i) for i in range(ens.size)]
            for v, ic in zip(views, ics):
                v.set_state(*ic, 1.0, 0.0)
                v.calc_metric()
            for nsteps in (2, 2, 1):  # the second call replays the first's graph
                ens.symp(4, dts, nsteps)
            ens.synchronize()
            out[(single, graph)] = [v.get_state() for v in views]
            for v in views:
                v.close()
    ref = out[(False, False)]
    for key in ((True, True), (True, False)):
        for i in range(len(sizes)):
            for a, b in zip(out[key][i], ref[i]):
                assert np.array_equal(np.asarray(a), np.asarray(b)), (key, i)
    # classes the single launch does not cover run one launch per class
    with gpu.Ensemble([(8, 1, 5.0), (16, 1, 5.0)]) as ens:
        assert ens.single_launch(True) is False


@pytest.mark.parametrize("n", [32, 64, 128, 192])
@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_tuned_launch_shapes_match_oracle(gpu, n, precision):
    """Both launch shapes of every sweep (default, and fewer rows per x
    workgroup / fewer columns per y, z tile) against the oracle: FP64 to
    1e-12; the mixed form to its FP32 level on the fluctuations."""
    print("case tuning", n, precision, flush=True)
    L, dt, nsteps = 10.0, 2e-4, 2
    fld, fldp = member_ic(n, L, 55)
    prec = gpu.F64 if precision == "f64" else gpu.MIXED
    sim = O.Simulation(n, 2, L, O.COSMIC, model=O.MODEL_TWOFIELD, mp=(1.0, 2.5e5))
    sim.set_state(fld, fldp, 1.0, 0.0)
    sim.calc_metric()
    sim.symp(4, dt, nsteps)
    out = []
    for shape in ((0, 0, 0, 0), (1, 1, 1, 1)):
        with gpu.Lattice(n, nfld=2, length=L, precision=prec) as lat:
            lat.set_tuning(*shape)
            assert tuple(lat.tuning().values()) == shape
            lat.set_state(fld, fldp, 1.0, 0.0)
            lat.calc_metric()
            lat.symp4(dt, nsteps)
            out.append(lat.get_state())
    for st in out:
        if precision == "f64":
            assert rel(st[0], sim.fld) < 1e-12 and rel(st[1], sim.fldp) < 1e-12
        else:
            for l in range(2):
                for a, b in ((st[0][l], sim.fld[l]), (st[1][l], sim.fldp[l])):
                    assert rel(a - a.mean(), b - b.mean()) < 1e-5, l
        assert abs(st[2] - sim.yscl) < 1e-13
    with gpu.Lattice(n, nfld=2, length=L, precision=prec) as lat:
        t = lat.autotune()
        assert set(t.values()) <= {0, 1}


def test_ensemble_96_members_match_oracle(gpu):
    """The 96-realisation ensemble (32 x 64^3 + 48 x 128^3 + 16 x 192^3, fixed
    L = 10, g^2/m^2 log-uniform in [1e4, 1e6]), autotuned launch mode and
    shapes, graph replay: one member of each class against the oracle from
    the member's own initial state."""
    print("case ensemble 96", flush=True)
    classes = [(64, 32, 10.0), (128, 48, 10.0), (192, 16, 10.0)]
    rng = np.random.default_rng(11)
    g2s = np.exp(rng.uniform(np.log(1e4), np.log(1e6), 96))
    params = [(1.0, float(g)) for g in g2s]
    dts = [1e-3, 1e-3, 1e-3]  # omega dt <= 1 for g^2 up to 1e6
    picks = [5, 32 + 17, 80 + 9]
    with gpu.Ensemble(classes, members=params) as ens:
        views = [ens.member(i) for i in range(ens.size)]
        ics = {}
        for i, v in enumerate(views):
            g2 = float(g2s[i])
            v.sample_field(0, 0, PHI0, -0.25, 1.0, 2e5, 3000 + i, 0)
            v.sample_field(0, 1, DPHI0, 0.25, 1.0, 2e5, 3000 + i, 1)
            v.sample_field(1, 0, 0.0, -0.25, g2 * PHI0 ** 2, 2e5, 3000 + i, 2)
            v.sample_field(1, 1, 0.0, 0.25, g2 * PHI0 ** 2, 2e5, 3000 + i, 3)
            v.set_scalars(1.0, 0.0)
            v.calc_metric()
            if i in picks:
                ics[i] = v.get_state()
        for _ in range(2):  # capture, then replay
            ens.symp(4, dts, 5)
        ens.synchronize()
        got = {i: views[i].get_state() for i in picks}
        for v in views:
            v.close()
    for i in picks:
        n = 64 if i < 32 else (128 if i < 80 else 192)
        sim = O.Simulation(n, 2, 10.0, O.COSMIC, model=O.MODEL_TWOFIELD, mp=params[i])
        f, p, y, py = ics[i]
        sim.set_state(f, p, y, py)
        for _ in range(2):
            sim.symp(4, 1e-3, 5)
        print("ensemble member", i, n, rel(got[i][0], sim.fld), rel(got[i][1], sim.fldp), flush=True)
        assert rel(got[i][0], sim.fld) < 1e-12, i
        assert rel(got[i][1], sim.fldp) < 1e-12, i
        assert abs(got[i][2] - sim.yscl) < 1e-13 * abs(sim.yscl), i
        assert abs(got[i][3] - sim.ysclp) < 1e-11 * abs(sim.ysclp), i


def test_member_views_free_their_scratch(gpu):
    """A mixed-form 7-point batch holds no k-space arrays; a member view's
    output spectra allocate them on demand, and destroying the view frees
    them: device memory returns to its level before the views."""
    import torch

    print("case view scratch", flush=True)
    n = 128
    with gpu.Lattice(n, nfld=2, length=10.0, precision=gpu.MIXED, laplacian=gpu.LAP_STENCIL,
                     members=[(1.0, 2.5e5)] * 2) as batch:
        v = batch.member(0)
        v.output_spectra(0)  # first use: the batch's shared spectrum buffers
        v.close()
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        for _ in range(6):
            v = batch.member(1)
            v.output_spectra(1)
            v.close()
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
    nk_bytes = 2 * 8 * (n // 2 + 8) * n * n  # one view's two complex FP32 arrays
    assert free0 - free1 < nk_bytes, (free0, free1, nk_bytes)
