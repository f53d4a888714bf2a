"""Helpers for the GPU parity tests: run the HIP path and the CPU oracle on the
same inputs and return both results."""

import numpy as np

import oracle


def oracle_rng(runtime):
    """The oracle's RNG streams for a Runtime (same seed, same stream count)."""
    return oracle.rng_init(runtime.R * runtime.L * runtime.L, runtime.seed)


def device_rng_as_pairs(dev):
    host = dev.copy_to_host()
    return np.stack([host["s0"], host["s1"]], axis=-1).astype(np.uint64)


def upload_pairs(dev, pairs):
    """Write (n, 2) uint64 states into a structured device RNG array."""
    dev.copy_to_device(np.ascontiguousarray(pairs, dtype=np.uint64).view(dev.dtype).reshape(dev.shape))


def oracle_sweep(model_kind, state, rng, betas, model, nsweeps):
    """Run nsweeps of the oracle's colour updates in order 0..3."""
    for _ in range(nsweeps):
        for color in range(4):
            oracle_color(model_kind, state, rng, betas, model, color)


def oracle_color(model_kind, state, rng, betas, model, color):
    if model_kind == "ising":
        oracle.ising_color(state, rng, betas, color, model.J)
    elif model_kind == "xy":
        oracle.xy_color(state, rng, betas, color, model.J, model.delta)
    elif model_kind == "kawamura":
        oracle.kawamura_color(state, rng, betas, color, model.J, model.delta)
    elif model_kind == "he3":
        oracle.he3_color(state, rng, betas, color, model.betas, model.K, model.delta, model.g_d)
    else:
        raise ValueError(model_kind)
