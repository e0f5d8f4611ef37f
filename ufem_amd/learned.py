"""Training set and training loop of the learned preconditioner (ufem_net_*).

Pairs come from plain Jacobi-PCG solves of the same cantilever under fresh
random material fields (E = exp of a Gaussian random field per brick): each
solve snapshots (x_k, r_k) every `every` iterations, runs to a tight
tolerance so that e_k = u - x_k is accurate, and keeps `pairs_per_field`
evenly spread pairs (r_k, e_k) together with that field's matrix (the
energy loss needs K e for every pair).
"""

from __future__ import annotations

import time

import numpy as np

from . import gpu as G
from .mesh import grf_E


def field_E(mesh, seed):
    """a fresh E field for the mesh's box (squared-exponential GRF, 0.2 m, log-std 0.5)"""
    L, B, H = (mesh.coords[:, i].max() - mesh.coords[:, i].min() for i in range(3))
    return grf_E(seed, mesh.nx, mesh.ny, mesh.nz, (L / mesh.nx, B / mesh.ny, H / mesh.nz))


def collect(system, net, mesh, seeds, pairs_per_field=32, every=8, eps=1e-11, npts=2, log=None, field=None):
    """Assemble and solve `system` for each E field of `seeds` (field(mesh,
    seed), by default `field_E`), adding pairs to `net`; leaves the system
    holding the last field.  Returns (pairs, seconds)."""
    field = field or field_E
    t0 = time.perf_counter()
    added = 0
    for k, sd in enumerate(seeds):
        E = field(mesh, sd)
        system.set_E(E)
        system.fill_K(npts)
        system.add_disps()
        net.bind(system, E)
        n, info = net.collect(pairs_per_field, eps=eps, every=every)
        added += n
        if log and (k % 8 == 0 or k == len(seeds) - 1):
            log(f"field {k + 1}/{len(seeds)}: {info['iters']} PCG iterations, {added} pairs")
    return added, time.perf_counter() - t0


def rebind(system, net, mesh, E, npts=2):
    """put `system` back on the material field E and bind the network to it"""
    system.set_E(E)
    system.fill_K(npts)
    system.add_disps()
    net.bind(system, E)


def train(net, steps, batch=4, lr=3e-3, seed=0):
    """Adam on the energy loss; returns (loss per step, device ms)"""
    loss, ms = net.train(steps, batch=batch, lr=lr, seed=seed)
    return np.asarray(loss), ms
