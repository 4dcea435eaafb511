"""GPU checks of the captured RK2 step of a sweep (eb_sweep_field_rk2_graph): the device-
controlled batched DGMRES inside it gives the host-controlled solves' results and iteration
counts, every replay of the graph matches the eager step (with eager work on the same stream
between replays), and a budget too small for some system falls back to the eager step."""
import numpy as np
import pytest

from electricbubbles_amd.workloads import perturbed_circle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    if not _t.cuda.is_available():
        pytest.skip("no GPU")
    return _t


def _case(nsim=12, seed=21):
    rng = np.random.default_rng(seed)
    nd = (np.round(np.exp(rng.uniform(np.log(128), np.log(512), nsim)) / 2) * 2).astype(np.int32)
    rku = np.concatenate([perturbed_circle(int(n), [(2, 0.05)])[0] for n in nd]).view(np.float64).copy()
    e0 = rng.uniform(0.1, 2.0, nsim) ** 2
    cap = rng.uniform(0.0, 0.2, nsim)
    return nd, rku, e0, cap


def _field(torch, nd, e0, cap, stream):
    from electricbubbles_amd import field, spectral
    B = spectral.Bodies(nd, stream=stream)
    return B, field.SweepField(B, np.arange(nd.size + 1), cap, e0, stream=stream)


def _eager_step(SF, spectral, ybar, dt, stream):
    phi0, y = spectral.rk2_stage1(dt, ybar, SF.rhs(ybar), stream=stream)
    spectral.rk2_stage2(dt, phi0, SF.rhs(y), ybar, stream=stream)


def test_graph_step_matches_eager(torch):
    from electricbubbles_amd import spectral
    nd, rku, e0, cap = _case()
    print("graph step: sweep of", nd.size, "simulations", flush=True)
    st = torch.cuda.Stream()
    B, SF = _field(torch, nd, e0, cap, st)
    B2, SF2 = _field(torch, nd, e0, cap, st)
    y0 = torch.tensor(rku, device="cuda")
    torch.cuda.synchronize()
    yg, ye = y0.clone(), y0.clone()
    dt = 1e-3
    with torch.cuda.stream(st):
        for step in range(4):
            redone = SF.rk2_graph(yg, dt, budget=12)
            it_g = SF.iterations()
            _eager_step(SF2, spectral, ye, dt, st)
            it_e = SF2.iterations()
            SF.rhs(ye)  # eager work of the graph's own field between replays
            st.synchronize()
            assert not redone
            for a, b in zip(it_g, it_e):
                assert np.array_equal(a, b), (step, a, b)
            d = float((yg - ye).norm() / ye.norm())
            print("step", step, "rel diff", d, flush=True)
            assert d < 1e-13
    torch.cuda.synchronize()


def test_graph_budget_fallback(torch):
    from electricbubbles_amd import spectral
    nd, rku, e0, cap = _case(nsim=6, seed=5)
    print("graph budget fallback:", nd.size, "simulations", flush=True)
    st = torch.cuda.Stream()
    B, SF = _field(torch, nd, e0, cap, st)
    B2, SF2 = _field(torch, nd, e0, cap, st)
    y0 = torch.tensor(rku, device="cuda")
    torch.cuda.synchronize()
    yg, ye = y0.clone(), y0.clone()
    with torch.cuda.stream(st):
        redone = SF.rk2_graph(yg, 1e-3, budget=2)
        _eager_step(SF2, spectral, ye, 1e-3, st)
        st.synchronize()
    assert redone
    assert float((yg - ye).norm() / ye.norm()) < 1e-13


def _plain_case(seed=3):
    """One simulation of three bodies (a plain field): 384, 256 and 512 points."""
    rng = np.random.default_rng(seed)
    nd = np.array([384, 256, 512], np.int32)
    shifts = [0.0, 2.6 + 0.4j, -0.3 + 2.7j]
    parts = []
    for n, c in zip(nd, shifts):
        parts.append(perturbed_circle(int(n), [(2, 0.05), (3, 0.02 * rng.uniform())], center=c)[0])
    return nd, np.concatenate(parts).view(np.float64).copy()


def _plain_eager_step(F, spectral, ybar, dt, stream):
    phi0, y = spectral.rk2_stage1(dt, ybar, F.rhs(ybar), stream=stream)
    spectral.rk2_stage2(dt, phi0, F.rhs(y), ybar, stream=stream)


@pytest.mark.parametrize("mode", ["auto", "fmm"])
def test_plain_graph_step_matches_eager(torch, mode):
    """eb_field_rk2_graph: the device-controlled solves of one system give the host-controlled
    solves' iteration counts and, up to the reductions' rounding, their state; every replay
    matches the eager step, with eager work of the same field between replays.  In FMM mode the
    captured step holds the near field's second stream."""
    import electricbubbles_amd as eb
    from electricbubbles_amd import field, spectral
    nd, rku = _plain_case()
    print("plain graph step:", mode, nd.tolist(), flush=True)
    st = torch.cuda.Stream()
    Fg = field.Field(spectral.Bodies(nd, stream=st), cap=0.1, e0=1.5, stream=st)
    Fe = field.Field(spectral.Bodies(nd, stream=st), cap=0.1, e0=1.5, stream=st)
    if mode == "fmm":
        Fg.set_mode(eb.MODE_FMM)
        Fe.set_mode(eb.MODE_FMM)
    y0 = torch.tensor(rku, device="cuda")
    torch.cuda.synchronize()
    yg, ye = y0.clone(), y0.clone()
    dt = 1e-3
    with torch.cuda.stream(st):
        for step in range(4):
            redone = Fg.rk2_graph(yg, dt, budget=40)
            it_g = Fg.iterations()
            _plain_eager_step(Fe, spectral, ye, dt, st)
            it_e = Fe.iterations()
            Fg.rhs(ye)  # eager work of the graph's own field between replays
            st.synchronize()
            print("step", step, "iterations graph", it_g, "eager", it_e, "redone", redone, flush=True)
            assert not redone
            assert it_g == it_e, (step, it_g, it_e)
            d = float((yg - ye).norm() / ye.norm())
            print("step", step, "iterations", it_g, "rel diff", d, flush=True)
            assert d < 1e-12
    torch.cuda.synchronize()


def test_plain_graph_budget_fallback(torch):
    from electricbubbles_amd import field, spectral
    nd, rku = _plain_case(seed=4)
    print("plain graph budget fallback:", nd.tolist(), flush=True)
    st = torch.cuda.Stream()
    Fg = field.Field(spectral.Bodies(nd, stream=st), cap=0.1, e0=1.5, stream=st)
    Fe = field.Field(spectral.Bodies(nd, stream=st), cap=0.1, e0=1.5, stream=st)
    y0 = torch.tensor(rku, device="cuda")
    torch.cuda.synchronize()
    yg, ye = y0.clone(), y0.clone()
    with torch.cuda.stream(st):
        redone = Fg.rk2_graph(yg, 1e-3, budget=2)
        _plain_eager_step(Fe, spectral, ye, 1e-3, st)
        st.synchronize()
    assert redone
    assert float((yg - ye).norm() / ye.norm()) < 1e-13


def _rk4_eager_step(F, spectral, y, dt, stream, torch):
    acc, ys = torch.empty_like(y), torch.empty_like(y)
    for q in range(4):
        spectral.rk4_stage(q, dt, y, F.rhs(y if q == 0 else ys), acc, ys, stream=stream)


def test_rk4_graph_steps_match_eager(torch):
    """eb_sweep_field_rk4_graph and eb_field_rk4_graph against the eager RK4 step (four F, the
    stage updates of eb_rk4_stage) over three replays; switching a field from RK2 to RK4 rebuilds
    its graph; a one-step budget falls back to the host-controlled solves."""
    from electricbubbles_amd import field, spectral
    nd, rku, e0, cap = _case(nsim=8, seed=9)
    st = torch.cuda.Stream()
    B, SF = _field(torch, nd, e0, cap, st)
    B2, SF2 = _field(torch, nd, e0, cap, st)
    pnd, prku = _plain_case(seed=6)
    Fg = field.Field(spectral.Bodies(pnd, stream=st), cap=0.1, e0=1.5, stream=st)
    Fe = field.Field(spectral.Bodies(pnd, stream=st), cap=0.1, e0=1.5, stream=st)
    dt = 1e-3
    with torch.cuda.stream(st):
        for name, G, E, y0 in (("sweep", SF, SF2, rku), ("plain", Fg, Fe, prku)):
            yg = torch.tensor(y0, device="cuda")
            ye = yg.clone()
            assert not G.rk2_graph(yg, dt, budget=40)  # an RK2 graph first: the RK4 call replaces it
            _plain_eager_step(E, spectral, ye, dt, st) if name == "plain" else _eager_step(E, spectral, ye, dt, st)
            for step in range(3):
                assert not G.rk4_graph(yg, dt, budget=40)
                _rk4_eager_step(E, spectral, ye, dt, st, torch)
                st.synchronize()
                d = float((yg - ye).norm() / ye.norm())
                print(name, "rk4 step", step, "rel diff", d, flush=True)
                assert d < 1e-12
            yb, yc = yg.clone(), yg.clone()
            assert G.rk4_graph(yb, dt, budget=1)  # redone with host-controlled solves
            _rk4_eager_step(E, spectral, yc, dt, st, torch)
            st.synchronize()
            assert float((yb - yc).norm() / yc.norm()) < 1e-12
    torch.cuda.synchronize()
