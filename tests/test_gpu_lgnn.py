"""GPU parity of the learned preconditioner (qcu_mrhs_amd/csrc/lgnn.hip)
against the CPU oracle (oracle/lgnn.py), through the C entry points, and what
the trained network gains over the solver without it.

The GPU evaluates the network with FP32 transports and an FP32 collapsed
stencil (FP64 partial sums for the gradient reductions), so forward results
agree with the FP64 oracle to a relative 1e-5 and gradients to 1e-4."""
import numpy as np
import pytest

from oracle import lgnn as olgnn
from oracle import oracle
from qcu_mrhs_amd import api, fields, lgnn

pytestmark = pytest.mark.gpu

L, Nc, R, KAPPA = (4, 4, 4, 8), 3, 4, 0.13
RTOL_FWD = 1e-5
RTOL_GRAD = 1e-4


def _case(*a):
    print("case", *a, flush=True)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _setup(L_, Nc_, R_, gauge, kappa, inner=2):
    api.initGridSize(L_, Nc_, R_, 2, inner)
    d_g = api.DeviceArray.from_complex(gauge.reshape(-1))
    api.loadQcuGauge(d_g, 2)
    api.getDslash(0, 1.0 / (2 * kappa) - 4.0)
    return d_g


def _odd_field(seed, shape):
    g = np.random.default_rng(seed)
    return g.standard_normal(shape) + 1j * g.standard_normal(shape)


def _abi_weights(net):
    """oracle W1 [C][1][9][4][4], W2 [1][C][9][4][4] -> C ABI [C][9][4][4] each."""
    return net.W1[:, 0].copy(), net.W2[0].copy()


# FP16 / BF16 storage runs the matrix-core path: inputs, M and the mixed hop
# fields are rounded to 16 bits (several roundings of 2^-11 / 2^-8 relative)
@pytest.mark.parametrize("prec,tol", [(2, RTOL_FWD), (1, RTOL_FWD), (0, 4e-3), (4, 2.5e-2)])
def test_forward_matches_oracle(prec, tol):
    _case("lgnn forward", prec)
    gauge = fields.gauge_field(L, Nc, 71, kind="haar")
    d_g = _setup(L, Nc, R, gauge, KAPPA)
    Vh = fields.half_volume(L)
    psi = _odd_field(72, (Vh, 4, Nc, R))
    net = olgnn.Net.random(8, 73, scale=0.3)
    lgnn.set_weights(*_abi_weights(net))
    d_in = api.DeviceArray.from_complex(psi.reshape(-1), prec)
    d_out = api.DeviceArray(api.storage_bytes(psi.size, prec))
    lgnn.apply(d_out, d_in, prec)
    got = d_out.complex(psi.size, prec)
    # the oracle sees the input as stored (rounded to the storage precision)
    psi_s = api.from_storage(api.to_storage(psi.reshape(-1), prec), prec).reshape(psi.shape)
    ref, _ = net.forward(olgnn.Transport(gauge, L, Nc), psi_s, L, Nc, R)
    err = _rel(got, ref.reshape(-1))
    print(f"precision {prec}: forward relative error {err:.2e}")
    assert err < tol
    api.finalizeQcu()
    del d_g


@pytest.mark.parametrize("Nc_,R_,prec", [(3, 17, 0), (2, 16, 0), (4, 5, 4), (3, 32, 4)])
def test_forward_matrix_core_path_shapes(Nc_, R_, prec):
    """FP16/BF16 application runs on the matrix-core kernels (transport + MFMA
    mixing, MFMA gather); RHS chunks beyond 16, partial chunks and Nc 2..4
    against the FP64 oracle on inputs rounded to the storage precision."""
    _case("lgnn forward mfma", Nc_, R_, prec)
    gauge = fields.gauge_field(L, Nc_, 171 + Nc_, kind="haar")
    d_g = _setup(L, Nc_, R_, gauge, KAPPA)
    Vh = fields.half_volume(L)
    psi = _odd_field(172 + R_, (Vh, 4, Nc_, R_))
    net = olgnn.Net.random(8, 173, scale=0.3)
    lgnn.set_weights(*_abi_weights(net))
    d_in = api.DeviceArray.from_complex(psi.reshape(-1), prec)
    d_out = api.DeviceArray(api.storage_bytes(psi.size, prec))
    lgnn.apply(d_out, d_in, prec)
    got = d_out.complex(psi.size, prec)
    psi_s = api.from_storage(api.to_storage(psi.reshape(-1), prec), prec).reshape(psi.shape)
    ref, _ = net.forward(olgnn.Transport(gauge, L, Nc_), psi_s, L, Nc_, R_)
    err = _rel(got, ref.reshape(-1))
    print(f"Nc {Nc_} R {R_} precision {prec}: forward relative error {err:.2e}")
    assert err < (4e-3 if prec == 0 else 2.5e-2)
    api.finalizeQcu()
    del d_g


@pytest.mark.parametrize("Nc_,R_,prec", [(3, 4, 0), (3, 17, 0), (2, 16, 4), (4, 5, 0), (3, 32, 4)])
def test_projected_network_matrix_core_path(Nc_, R_, prec):
    """Projected network (second-layer weights in the image of each hop's
    Wilson projector): the library projects the weights it is given as the
    package's project_w2 does, and its FP16 / BF16 application -- hidden hop
    fields of two spins, the other two rebuilt in the gather -- matches the
    FP64 oracle network with the projected weights, on inputs rounded to the
    storage precision; the same weights unprojected give a different network
    (the projection is not a no-op here)."""
    _case("lgnn projected", Nc_, R_, prec)
    gauge = fields.gauge_field(L, Nc_, 271 + Nc_, kind="haar")
    d_g = _setup(L, Nc_, R_, gauge, KAPPA)
    Vh = fields.half_volume(L)
    psi = _odd_field(272 + R_, (Vh, 4, Nc_, R_))
    net = olgnn.Net.random(8, 273, scale=0.3)
    w1, w2 = _abi_weights(net)
    tol = 4e-3 if prec == 0 else 2.5e-2
    lgnn.set_projected(True)
    try:
        lgnn.set_weights(w1, w2)
        gw1, gw2 = lgnn.get_weights(8)
        w2p = lgnn.project_w2(w2)
        assert np.array_equal(gw1, w1) and np.abs(gw2 - w2p).max() < 1e-14
        d_in = api.DeviceArray.from_complex(psi.reshape(-1), prec)
        d_out = api.DeviceArray(api.storage_bytes(psi.size, prec))
        lgnn.apply(d_out, d_in, prec)
        got = d_out.complex(psi.size, prec)
    finally:
        lgnn.set_projected(False)
    psi_s = api.from_storage(api.to_storage(psi.reshape(-1), prec), prec).reshape(psi.shape)
    tr = olgnn.Transport(gauge, L, Nc_)
    ref, _ = olgnn.Net(net.W1, w2p[None]).forward(tr, psi_s, L, Nc_, R_)
    ref_full, _ = net.forward(tr, psi_s, L, Nc_, R_)
    err = _rel(got, ref.reshape(-1))
    print(f"Nc {Nc_} R {R_} precision {prec}: projected forward relative error {err:.2e} "
          f"(unprojected network differs by {_rel(ref_full, ref):.2e})")
    assert err < tol and _rel(ref_full, ref) > 10 * tol
    api.finalizeQcu()
    del d_g


@pytest.mark.parametrize("prec", [0, 4])
def test_projected_application_at_production_shape(prec):
    """24^3 x 48, SU(3), 16 RHS, smooth links: with weights in the projected
    family the half-traffic kernels (hidden fields of two spins, 2 GB) agree
    with the general kernels (four spins, pinned against the oracle at small
    volume) on the same input, and the output is finite -- the full-size
    buffer offsets of both forms.  With such weights the general kernel's
    spins 2, 3 of each hidden field are i^kr times spins 0, 1 up to the FP32
    accumulation order, and so equal the rebuilt ones after 16-bit rounding
    but for rare rounding flips (measured 1e-6 relative)."""
    tol = 1e-4
    Lp, Rp = (24, 24, 24, 48), 16
    _case("lgnn projected production", prec)
    gauge = fields.gauge_field(Lp, Nc, 3001, kind="smooth", eps=0.5)
    d_g = _setup(Lp, Nc, Rp, gauge, 0.1355)
    n = fields.half_volume(Lp) * 4 * Nc * Rp
    psi = _odd_field(3002, (n,))
    w1, w2 = lgnn.neumann_weights(0.1355, channels=8, noise=1e-2, seed=3)
    w2 = lgnn.project_w2(w2)
    d_in = api.DeviceArray.from_complex(psi, prec)
    d_out = api.DeviceArray(api.storage_bytes(n, prec))
    out = {}
    try:
        for proj in (False, True):
            lgnn.set_projected(proj)
            lgnn.set_weights(w1, w2)
            lgnn.apply(d_out, d_in, prec)
            out[proj] = d_out.complex(n, prec)
    finally:
        lgnn.set_projected(False)
    err = _rel(out[True], out[False])
    print(f"precision {prec}: projected against general kernels {err:.2e}")
    assert np.all(np.isfinite(out[True])) and err < tol
    api.finalizeQcu()
    del d_g


def test_application_follows_a_reloaded_gauge_field():
    """The network keeps its single- and 16-bit copies of the gauge field
    between calls while the loaded field is unchanged, and a loadQcuGauge of
    another field into the same context replaces them: FP32 and FP16
    applications before and after the reload against the oracle on each
    field (and the two fields give different results)."""
    _case("lgnn gauge reload")
    gA = fields.gauge_field(L, Nc, 301, kind="haar")
    gB = fields.gauge_field(L, Nc, 302, kind="haar")
    d_g = _setup(L, Nc, R, gA, KAPPA)
    Vh = fields.half_volume(L)
    psi = _odd_field(303, (Vh, 4, Nc, R))
    net = olgnn.Net.random(8, 304, scale=0.3)
    lgnn.set_weights(*_abi_weights(net))
    for gauge in (gA, gA, gB, gB, gA):
        d_g.upload(gauge.reshape(-1))
        api.loadQcuGauge(d_g, 2)
        for prec, tol in ((1, RTOL_FWD), (0, 4e-3)):
            d_in = api.DeviceArray.from_complex(psi.reshape(-1), prec)
            d_out = api.DeviceArray(api.storage_bytes(psi.size, prec))
            for _ in range(2):  # the second call reuses the copies
                lgnn.apply(d_out, d_in, prec)
                got = d_out.complex(psi.size, prec)
                psi_s = api.from_storage(api.to_storage(psi.reshape(-1), prec), prec).reshape(psi.shape)
                ref, _ = net.forward(olgnn.Transport(gauge, L, Nc), psi_s, L, Nc, R)
                other, _ = net.forward(olgnn.Transport(gB if gauge is gA else gA, L, Nc), psi_s, L, Nc, R)
                assert _rel(got, ref.reshape(-1)) < tol and _rel(other, ref) > 100 * tol
    api.finalizeQcu()
    del d_g


def test_neumann_weights_reproduce_schur_neumann_term():
    """Weights (1 -+ gamma_mu) built by the package (not the oracle): the GPU
    network equals psi + kappa^2 D_oe D_eo psi = 2 psi - A psi, with A the odd
    Schur operator of the FP64 dslash oracle."""
    _case("lgnn neumann")
    gauge = fields.gauge_field(L, Nc, 81, kind="haar")
    d_g = _setup(L, Nc, R, gauge, KAPPA)
    Vh = fields.half_volume(L)
    psi = _odd_field(82, (Vh, 4, Nc, R)).reshape(-1)
    lgnn.set_weights(*lgnn.neumann_weights(KAPPA, channels=8))
    d_in = api.DeviceArray.from_complex(psi)
    d_out = api.DeviceArray(psi.nbytes)
    lgnn.apply(d_out, d_in, 2)
    got = d_out.complex(psi.size)
    ref = 2 * psi - oracle.schur(psi, gauge, L, Nc, R, KAPPA)
    assert _rel(got, ref) < RTOL_FWD
    api.finalizeQcu()
    del d_g


def test_loss_and_gradients_match_oracle():
    _case("lgnn gradient")
    gauge = fields.gauge_field(L, Nc, 91, kind="haar")
    d_g = _setup(L, Nc, R, gauge, KAPPA)
    Vh = fields.half_volume(L)
    shape = (Vh, 4, Nc, R)
    r = _odd_field(92, shape).astype(np.complex64).astype(np.complex128)
    e = _odd_field(93, shape).astype(np.complex64).astype(np.complex128)
    net = olgnn.Net.random(8, 94, scale=0.3)
    lgnn.set_weights(*_abi_weights(net))
    d_r = api.DeviceArray.from_complex(r.reshape(-1), 1)
    d_e = api.DeviceArray.from_complex(e.reshape(-1), 1)
    loss, g1, g2 = lgnn.loss_grad(d_r, d_e, 8)
    ref_loss, ref_g1, ref_g2 = net.loss_and_grad(olgnn.Transport(gauge, L, Nc), r, e, L, Nc, R)
    print(f"loss {loss:.8e} (oracle {ref_loss:.8e}); gradient errors "
          f"{_rel(g1, ref_g1[:, 0]):.2e}, {_rel(g2, ref_g2[0]):.2e}")
    assert abs(loss - ref_loss) <= RTOL_FWD * abs(ref_loss)
    assert _rel(g1, ref_g1[:, 0]) < RTOL_GRAD
    assert _rel(g2, ref_g2[0]) < RTOL_GRAD
    # a second call on the same batch gives the same numbers (gradients are not accumulated)
    loss2, g1b, _ = lgnn.loss_grad(d_r, d_e, 8)
    assert loss2 == loss and np.array_equal(g1b, g1)
    api.finalizeQcu()
    del d_g


@pytest.mark.parametrize("K", [1, 3])
def test_power_loss_and_gradients_match_oracle(K):
    """Residual-power loss ||(1 - A N)^K v / ||v||||^2 (forward and back-propagation
    through A, A^+, N, N^+ on the GPU in FP32) against the oracle's layer-form
    back-propagation with the FP64 Schur operator."""
    _case("lgnn power gradient", K)
    gauge = fields.gauge_field(L, Nc, 95, kind="smooth", eps=0.4)
    d_g = _setup(L, Nc, R, gauge, KAPPA)
    Vh = fields.half_volume(L)
    v = _odd_field(96, (Vh, 4, Nc, R)).astype(np.complex64).astype(np.complex128)
    net = olgnn.Net.random(8, 97, scale=0.15)
    lgnn.set_weights(*_abi_weights(net))
    d_v = api.DeviceArray.from_complex(v.reshape(-1), 1)
    loss, g1, g2 = lgnn.power_loss_grad(d_v, K, 8)

    def schur(x, dag):
        return oracle.schur(x.reshape(-1), gauge, L, Nc, R, KAPPA, dag)

    ref_loss, ref_g1, ref_g2 = net.power_loss_and_grad(olgnn.Transport(gauge, L, Nc), schur, v, K, L, Nc, R)
    print(f"K {K}: loss {loss:.8e} (oracle {ref_loss:.8e}); gradient errors "
          f"{_rel(g1, ref_g1[:, 0]):.2e}, {_rel(g2, ref_g2[0]):.2e}")
    assert abs(loss - ref_loss) <= 1e-5 * abs(ref_loss)
    assert _rel(g1, ref_g1[:, 0]) < RTOL_GRAD
    assert _rel(g2, ref_g2[0]) < RTOL_GRAD
    api.finalizeQcu()
    del d_g


def _solve(b, use_pc, inner=2, tol=1e-10, delta=0.0):
    R_ = b.shape[0]
    api.set_solver_options(inner, delta, use_pc)
    bs = [api.DeviceArray.from_complex(b[k]) for k in range(R_)]
    xs = [api.DeviceArray(b[k].nbytes) for k in range(R_)]
    for k in range(R_):
        api.pushBackFermions(xs[k], bs[k])
    info = api.qcuInvertEx(5000, tol)
    x = np.stack([xs[k].complex(b.shape[1]) for k in range(R_)])
    return x, info


def _train(Lt, Rt, eps, kappa, n_cfg, epochs=12, lr=2e-3, power=3, projected=False):
    """Batches from plain FP64 solves of n_cfg seeded configurations (their
    gauge fields; 16 iterate snapshots + 16 Gaussian errors per RHS for the
    pair loss), training from the first Neumann term with the residual-power
    loss on the GPU (projected: in the projected family, which the library
    keeps for the rest of the process until switched off).  Returns (held-out
    gauge, held-out sources, device gauge, Neumann start weights, trained
    weights, losses)."""
    gauges = [fields.gauge_field(Lt, Nc, 7000 + k, kind="smooth", eps=eps) for k in range(n_cfg)]
    d_g = _setup(Lt, Nc, Rt, gauges[0], kappa)
    lgnn.set_projected(projected)
    w1, w2 = lgnn.neumann_weights(kappa, channels=8, noise=1e-3, seed=11)
    lgnn.set_weights(w1, w2)
    lgnn.clear_data()
    for k in range(n_cfg):
        d_g.upload(gauges[k].reshape(-1))
        api.loadQcuGauge(d_g, 2)
        b = fields.gaussian_sources(Lt, Nc, Rt, 7100 + k)
        bs = [api.DeviceArray.from_complex(b[j]) for j in range(Rt)]
        xs = [api.DeviceArray(b[j].nbytes) for j in range(Rt)]
        for j in range(Rt):
            api.pushBackFermions(xs[j], bs[j])
        lgnn.collect(5000, 1e-10, 16, n_random=16, seed=k)
    lgnn.set_objective(power)
    losses = lgnn.train(epochs, lr, seed=5)
    lgnn.set_objective(0)
    trained = lgnn.get_weights(8)
    lgnn.clear_data()
    held = fields.gauge_field(Lt, Nc, 7900, kind="smooth", eps=eps)
    d_g.upload(held.reshape(-1))
    api.loadQcuGauge(d_g, 2)
    b = fields.gaussian_sources(Lt, Nc, Rt, 7901)
    return held, b, d_g, lgnn.neumann_weights(kappa, channels=8), trained, losses


def _true_rel(x, b, gauge, Lt, Rt, kappa):
    bm = fields.to_mrhs(b, Nc)
    r = bm - oracle.mat(fields.to_mrhs(x, Nc), gauge, Lt, Nc, Rt, kappa)
    return (oracle.norm(r, Rt) / oracle.norm(bm, Rt)).max()


def _min_time(b, use_pc, inner, reps=3):
    """Fastest of `reps` identical solves (wall time of qcuInvertEx) and its info."""
    best = None
    for _ in range(reps):
        x, info = _solve(b, use_pc, inner)
        if best is None or info["seconds"] < best[1]["seconds"]:
            best = (x, info)
    return best


def test_training_beats_the_untrained_start_on_held_out_config():
    """8^4 SU(3), exp(i 0.3 H) links, kappa 0.135, R = 4, 4 training
    configurations, residual-power loss (K = 3): on a held-out configuration,
    in the plain FP64 solve (deterministic recurrences), the trained network as
    right preconditioner takes fewer iterations and less time (fastest of
    three solves each) than its untrained start (the first Neumann term), which
    takes fewer than the solver without it.  With FP16 inner precision both
    preconditioned solves beat the unpreconditioned one; their order is not
    asserted there: the reliable updates restart the inner solve from
    FP16-rounded residuals, so a change at the level of FP64 rounding moves
    those counts by a few iterations either way.  Every solve reaches a FP64
    true residual of 1e-10 through the oracle operator."""
    Lt, Rt, kappa = (8, 8, 8, 8), 4, 0.135
    _case("lgnn train", Lt)
    held, b, d_g, neumann, trained, losses = _train(Lt, Rt, 0.3, kappa, 4)
    print("training loss per epoch:", " ".join(f"{v:.3e}" for v in losses))
    assert np.all(np.isfinite(losses)) and losses[-1] < losses[0]
    res = {}
    for inner in (2, 0):
        _, plain = _min_time(b, False, inner)
        lgnn.set_weights(*neumann)
        _, init = _min_time(b, True, inner)
        lgnn.set_weights(*trained)
        x, pc = _min_time(b, True, inner)
        tr = _true_rel(x, b, held, Lt, Rt, kappa)
        res[inner] = (plain, init, pc)
        print(f"inner {inner}: iterations plain {plain['iterations']}, Neumann start {init['iterations']}, "
              f"trained {pc['iterations']} ({pc['reliable_updates']} reliable updates); times "
              f"{plain['seconds'] * 1e3:.2f} / {init['seconds'] * 1e3:.2f} / {pc['seconds'] * 1e3:.2f} ms; "
              f"trained true residual {tr:.2e}")
        assert plain["converged"] and init["converged"] and pc["converged"]
        assert init["preconditioned"] == 1 and pc["preconditioned"] == 1 and pc["fallback"] == 0
        assert tr < 1.5e-10 and pc["true_rel_residual"] < 1.5e-10
        if inner == 2:
            assert pc["iterations"] < init["iterations"] < plain["iterations"]
            assert pc["seconds"] < init["seconds"]
        else:
            assert max(pc["iterations"], init["iterations"]) < plain["iterations"]
    api.finalizeQcu()
    del d_g


def test_projected_training_on_held_out_config():
    """The 8^4 training of the test above in the projected family: every
    training step leaves the second-layer weights in the image of the hops'
    Wilson projectors; on the held-out configuration the trained network needs
    no more FP64 iterations than its start (the first Neumann term, which is
    in the family) and fewer than the solver without it, and the FP16 solve
    with the half-traffic matrix-core application converges to an FP64 true
    residual of 1e-10 through the oracle operator in fewer iterations than
    without the network."""
    Lt, Rt, kappa = (8, 8, 8, 8), 4, 0.135
    _case("lgnn train projected", Lt)
    try:
        held, b, d_g, neumann, trained, losses = _train(Lt, Rt, 0.3, kappa, 4, projected=True)
        print("training loss per epoch:", " ".join(f"{v:.3e}" for v in losses))
        assert np.all(np.isfinite(losses)) and losses[-1] < losses[0]
        w2 = trained[1]
        assert np.abs(lgnn.project_w2(w2) - w2).max() < 1e-12 * np.abs(w2).max()
        it = {}
        for inner in (2, 0):
            _, plain = _solve(b, False, inner)
            lgnn.set_weights(*neumann)
            _, init = _solve(b, True, inner)
            lgnn.set_weights(*trained)
            x, pc = _solve(b, True, inner)
            tr = _true_rel(x, b, held, Lt, Rt, kappa)
            print(f"inner {inner}: iterations plain {plain['iterations']}, Neumann start {init['iterations']}, "
                  f"trained (projected) {pc['iterations']}; trained true residual {tr:.2e}")
            assert pc["converged"] and pc["preconditioned"] == 1 and pc["fallback"] == 0 and tr < 1.5e-10
            it[inner] = (plain["iterations"], init["iterations"], pc["iterations"])
        assert it[2][2] <= it[2][1] < it[2][0]
        assert it[0][2] < it[0][0]
    finally:
        lgnn.set_projected(False)
    api.finalizeQcu()
    del d_g


@pytest.mark.parametrize("kappa", [0.130, 0.1345, 0.1355])
def test_training_at_production_volume_and_kappa(kappa):
    """16^3 x 32 SU(3), exp(i 0.5 H) links, 16 RHS, kappa 0.130 / 0.1345 /
    0.1355, 4 training configurations, residual-power loss (K = 3): on a
    held-out configuration, in the plain FP64 solve, the trained network needs
    no more iterations than its untrained start and fewer than the solver
    without it; with FP16 inner precision (FP64 reliable updates) the trained
    network beats the unpreconditioned solve (its order against the start is
    printed: FP16 counts move by an iteration or two with rounding).  FP64 true
    residual 1e-10 through the oracle for the trained network's solves."""
    Lt, Rt = (16, 16, 16, 32), 16
    _case("lgnn train production", kappa)
    held, b, d_g, neumann, trained, losses = _train(Lt, Rt, 0.5, kappa, 4)
    assert np.all(np.isfinite(losses)) and losses[-1] <= losses[0]
    it = {}
    for inner in (2, 0):
        _, plain = _solve(b, False, inner)
        lgnn.set_weights(*neumann)
        _, init = _solve(b, True, inner)
        lgnn.set_weights(*trained)
        x, pc = _solve(b, True, inner)
        tr = _true_rel(x, b, held, Lt, Rt, kappa)
        name = {2: "FP64", 0: "FP16 inner"}[inner]
        print(f"kappa {kappa}: {name} iterations plain {plain['iterations']}, Neumann start "
              f"{init['iterations']}, trained {pc['iterations']}; times {plain['seconds'] * 1e3:.1f} / "
              f"{init['seconds'] * 1e3:.1f} / {pc['seconds'] * 1e3:.1f} ms; trained true residual {tr:.2e}")
        assert pc["converged"] and pc["fallback"] == 0 and tr < 1.5e-10
        it[inner] = (plain["iterations"], init["iterations"], pc["iterations"])
    assert it[2][2] <= it[2][1] < it[2][0]
    assert it[0][2] < it[0][0]
    api.finalizeQcu()
    del d_g
