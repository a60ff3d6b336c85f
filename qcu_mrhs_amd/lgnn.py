"""Learned gauge-covariant right preconditioner over libqcu.so.

Model: two linear parallel-transport convolution layers 1 -> C -> 1 on the
odd-parity Schur system (nearest-neighbour hops, complex 4 x 4 spin-matrix
weights); see include/qcu.h for the C entry points.  Weights are numpy arrays
``W1[C][9][4][4]`` (layer 1, output channel o, hop h) and ``W2[C][9][4][4]``
(layer 2, input channel o, hop h), hop 0 the identity, hop 1 + 2 mu (+1) the
forward (backward) transport along mu.

Typical use, after ``initGridSize`` / ``loadQcuGauge`` / ``getDslash``::

    w1, w2 = neumann_weights(kappa, channels=8, noise=1e-3, seed=1)
    set_weights(w1, w2)
    for cfg in configurations:                 # load each gauge field, queue sources
        collect(max_iteration, 1e-10, n_snapshots)
    losses = train(epochs, learning_rate)
    api.set_solver_options(inner, 0.1, use_preconditioner=True)
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import api

NHOP = 9

# Euclidean gamma matrices, chiral basis (gamma_x, gamma_y, gamma_z, gamma_t):
# the projectors (1 -+ gamma_mu) of the Wilson hopping term in the dslash
# kernels (spin projection onto two components plus reconstruction).
_I = 1j
GAMMA = np.array([
    [[0, 0, 0, _I], [0, 0, _I, 0], [0, -_I, 0, 0], [-_I, 0, 0, 0]],
    [[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]],
    [[0, 0, _I, 0], [0, 0, 0, -_I], [-_I, 0, 0, 0], [0, _I, 0, 0]],
    [[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]],
], dtype=np.complex128)


def neumann_weights(kappa, channels=8, noise=0.0, seed=0):
    """Weights for which the network equals 1 + kappa^2 D_oe D_eo (the first
    Neumann term of the inverse of the odd Schur operator 1 - kappa^2 D_oe D_eo):
    channel 0 carries the identity, channel 1 the hopping term (1 -+ gamma_mu)
    in both layers; every entry gets complex Gaussian noise of size ``noise``
    so the remaining channels receive gradients."""
    if channels < 2:
        raise ValueError("the Neumann initialisation needs at least two channels")
    W1 = np.zeros((channels, NHOP, 4, 4), dtype=np.complex128)
    W2 = np.zeros((channels, NHOP, 4, 4), dtype=np.complex128)
    I4 = np.eye(4)
    W1[0, 0] = I4
    W2[0, 0] = I4
    for mu in range(4):
        W1[1, 1 + 2 * mu] = kappa * (I4 - GAMMA[mu])
        W1[1, 2 + 2 * mu] = kappa * (I4 + GAMMA[mu])
        W2[1, 1 + 2 * mu] = kappa * (I4 - GAMMA[mu])
        W2[1, 2 + 2 * mu] = kappa * (I4 + GAMMA[mu])
    if noise:
        g = np.random.default_rng(seed)
        for W in (W1, W2):
            W += noise * (g.standard_normal(W.shape) + 1j * g.standard_normal(W.shape))
    return W1, W2


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def set_weights(w1, w2):
    w1 = np.ascontiguousarray(w1, dtype=np.complex128)
    w2 = np.ascontiguousarray(w2, dtype=np.complex128)
    if w1.shape != w2.shape or w1.shape[1:] != (NHOP, 4, 4):
        raise ValueError(f"weights must be [C][9][4][4], got {w1.shape} and {w2.shape}")
    api.check(api.lib().qcuLgnnSetWeights(w1.shape[0], _dptr(w1), _dptr(w2)))


def get_weights(channels):
    w1 = np.zeros((channels, NHOP, 4, 4), dtype=np.complex128)
    w2 = np.zeros_like(w1)
    api.check(api.lib().qcuLgnnGetWeights(_dptr(w1), _dptr(w2)))
    return w1, w2


def apply(out, inp, precision=2):
    """out = N inp on odd-parity multi-RHS half vectors (device buffers)."""
    api.check(api.lib().qcuLgnnApply(api._p(out), api._p(inp), int(precision)))


def loss_grad(inp, target, channels):
    """(loss, dL/dRe W1 + i dL/dIm W1, same for W2) for one batch of FP32
    odd-parity multi-RHS pairs (device buffers)."""
    g1 = np.zeros((channels, NHOP, 4, 4), dtype=np.complex128)
    g2 = np.zeros_like(g1)
    loss = ctypes.c_double(0.0)
    api.check(api.lib().qcuLgnnLossGrad(api._p(inp), api._p(target), ctypes.byref(loss), _dptr(g1), _dptr(g2)))
    return loss.value, g1, g2


def collect(max_iteration, tol, n_snapshots, n_random=0, seed=0):
    """Solve the queued systems (plain FP64) and keep n_snapshots
    (error, residual) pairs per RHS from the solver's iterates plus n_random
    pairs (Gaussian error, its residual); returns the number of stored batches."""
    nb = ctypes.c_int(0)
    api.check(api.lib().qcuLgnnCollect(int(max_iteration), float(tol), int(n_snapshots), int(n_random),
                                       int(seed) & 0xFFFFFFFF, ctypes.byref(nb)))
    return nb.value


def train(epochs, learning_rate, seed=0):
    """Adam over the stored batches (cosine learning-rate decay from
    ``learning_rate`` to 0 over the epochs); returns the mean loss per pair
    per epoch."""
    losses = np.zeros(int(epochs), dtype=np.float64)
    api.check(api.lib().qcuLgnnTrain(int(epochs), float(learning_rate), int(seed) & 0xFFFFFFFF, _dptr(losses)))
    return losses


def clear_data():
    api.check(api.lib().qcuLgnnClearData())


def set_objective(power):
    """Training objective of train(): 0 = pair loss sum ||N r - e||^2 / ||e||^2
    on the collected (error, residual) pairs; K = 1..8 = residual-power loss
    sum ||(1 - A N)^K v||^2 on a fixed unit Gaussian vector per stored batch,
    which weights the modes that set the preconditioned iteration count."""
    api.check(api.lib().qcuLgnnSetObjective(int(power)))


def projector(h):
    """The Wilson projector of hop h = 1..8 (h = 1 + 2 mu forward, 2 + 2 mu
    backward): (1 - gamma_mu) / 2 forward, (1 + gamma_mu) / 2 backward."""
    mu, back = (h - 1) >> 1, (h - 1) & 1
    return 0.5 * (np.eye(4) + (1 if back else -1) * GAMMA[mu])


def project_w2(w2):
    """Second-layer weights moved into the projected family: W2[o][h] =
    P_h W2[o][h] for the hops h = 1..8 (hop 0, the identity, unchanged)."""
    w2 = np.array(w2, dtype=np.complex128, copy=True)
    for h in range(1, NHOP):
        w2[:, h] = np.einsum("st,otu->osu", projector(h), w2[:, h])
    return w2


def set_projected(on):
    """Projected network mode of the library (second-layer weights kept in
    the image of each hop's Wilson projector; the FP16 / BF16 application then
    stores half the hidden hop fields).  Persists across set_weights."""
    api.check(api.lib().qcuLgnnSetProjected(1 if on else 0))


def power_loss_grad(v, power, channels):
    """(loss, dL/dRe W1 + i dL/dIm W1, same for W2) of the residual-power loss
    sum_r ||(1 - A N)^power v_r / ||v_r||||^2 for one batch of FP32 odd-parity
    multi-RHS vectors (device buffer) on the loaded gauge field."""
    g1 = np.zeros((channels, NHOP, 4, 4), dtype=np.complex128)
    g2 = np.zeros_like(g1)
    loss = ctypes.c_double(0.0)
    api.check(api.lib().qcuLgnnPowerLossGrad(api._p(v), int(power), ctypes.byref(loss), _dptr(g1), _dptr(g2)))
    return loss.value, g1, g2
