"""What the trained corrector gains: trained on the GPU on pairs the plain
integrator generates (coarse leapfrog step of 8 dt against 8 Yoshida-4
steps of dt), it must cut the one-step error of the coarse step against
the plain integrator on held-out states (couplings and seeds the training
never saw) to at most GAIN_BAR of the leapfrog step's alone, for the fields
and for the momenta; and the judged stepping must keep the trajectory
error below the leapfrog's while using fewer force evaluations than the
plain integrator.  The held-out lattices are destroyed before the
corrector (a corrector must not keep a destroyed context's stream)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PHI0, DPHI0 = 1.0093430384226378929425913902459, -0.7137133070120812430962278466136
N, L, CF = 64, 2.5, 8           # dx of the 256^3, L = 10 configuration
MPL = 2.0e5                     # vacuum amplitude of the configurations' initial conditions
DT = (L / N) / 40.0
GAIN_BAR = 0.7


def init(lat, seed, g2):
    m2 = [1.0, g2 * PHI0 ** 2]
    lat.sample_field(0, 0, PHI0, -0.25, m2[0], MPL, seed, 0)
    lat.sample_field(0, 1, DPHI0, 0.25, m2[0], MPL, seed, 1)
    lat.sample_field(1, 0, 0.0, -0.25, m2[1], MPL, seed, 2)
    lat.sample_field(1, 1, 0.0, 0.25, m2[1], MPL, seed, 3)
    lat.set_scalars(1.0, 0.0)
    lat.calc_metric()


def fluct_err(st, ref):
    out = []
    for k in (0, 1):
        d = np.linalg.norm(st[k] - ref[k])
        r = np.linalg.norm(ref[k] - ref[k].mean(axis=(1, 2, 3), keepdims=True))
        out.append(d / r)
    return out


def test_trained_corrector_beats_leapfrog_on_held_out_states(gpu):
    print("case corrector gain", N, CF, flush=True)
    rng = np.random.default_rng(3)
    g2max = (4.0 - 3.0 * math.pi ** 2 * (CF / 40.0) ** 2) / ((CF * DT) ** 2 * PHI0 ** 2)
    with gpu.Corrector(seed=11) as corr:
        # four trajectories trained on in turn, each replaced by a fresh one
        # (new coupling and seed) after 50 pairs, so that the training
        # states stay in the early preheating phase the held-out states are in
        def fresh(k):
            g2 = math.sqrt(1e4 * g2max) if k == 0 else float(np.exp(rng.uniform(np.log(1e4), np.log(g2max))))
            lat = gpu.Lattice(N, nfld=2, length=L, model_params=(1.0, g2))
            init(lat, 700 + k, g2)
            warm = int(rng.integers(0, 200))
            if warm:
                lat.symp4(DT, warm)
            return lat

        streams = [fresh(k) for k in range(4)]
        nq, npairs = 4, 3000
        for q in range(npairs):
            lat = streams[q % 4]
            loss = corr.train(lat, DT, CF, npairs=1, lr=5e-4 * 0.1 ** (q / npairs))[0]
            assert np.isfinite(loss)
            if (q // 4) % 50 == 49 and q % 4 == 3:
                for lat in streams:
                    lat.close()
                streams = [fresh(nq + k) for k in range(4)]
                nq += 4
        for lat in streams:
            lat.close()
        ratios = []
        for g2, seed, warm in ((1.3e4, 901, 40), (2.5e4, 902, 80), (3.5e4, 903, 20)):
            with gpu.Lattice(N, nfld=2, length=L, model_params=(1.0, g2)) as lat:
                init(lat, seed, g2)
                lat.symp4(DT, warm)
                start = lat.get_state()
                lat.symp4(DT, CF)
                fine = lat.get_state()
                lat.set_state(*start)
                lat.symp(2, CF * DT, 1)
                leap = lat.get_state()
                lat.set_state(*start)
                corr.step(lat, DT, CF, nsteps=1, threshold=1e300)
                cor = lat.get_state()
            el, ec = fluct_err(leap, fine), fluct_err(cor, fine)
            ratios.append([ec[0] / el[0], ec[1] / el[1]])
            print("held out g2", g2, "leapfrog", el, "corrected", ec, flush=True)
        ratios = np.array(ratios)
        print("ratios", ratios.tolist(), flush=True)
        assert np.median(ratios[:, 0]) <= GAIN_BAR and np.median(ratios[:, 1]) <= GAIN_BAR
        # judged stepping over 10 coarse steps: cheaper than the plain run,
        # closer to it than the leapfrog alone
        g2 = 2e4
        with gpu.Lattice(N, nfld=2, length=L, model_params=(1.0, g2)) as lat:
            init(lat, 905, g2)
            start = lat.get_state()
            lat.symp4(DT, 10 * CF)
            ref = lat.get_state()
            lat.set_state(*start)
            lat.symp(2, CF * DT, 10)
            leap = lat.get_state()
            lat.set_state(*start)
            st = corr.step_judged(lat, DT, CF, nsteps=10, threshold=1e-6, judge_every=5, err_tol=0.5)
            judged = lat.get_state()
        print("judged", st, fluct_err(judged, ref), fluct_err(leap, ref), flush=True)
        assert st["force_evals"] < 3 * CF * 10 / 2
        ej, el = fluct_err(judged, ref), fluct_err(leap, ref)
        assert ej[0] < el[0] and ej[1] < el[1]
