"""Uniform lists for the parity mode of dvm_set_uniform_stream / the oracle's
orc_set_uniform_stream (test helpers)."""
import ctypes as C

import numpy as np

from tests._oracle import load_oracle


def sample_order(solver, ntime, ncell):
    """(t, s) of every sample of every cell, in the order a collision at step
    ntime draws them, from the solver's parity trace (counter-based mode)."""
    out = []
    for cell in range(ncell):
        _, _, n = solver.collide_trace(ntime, cell, 0)
        irec, _, _ = solver.collide_trace(ntime, cell, n)
        out.append([(int(r[0]) - 1, int(r[1])) for r in irec])
    return out


def philox_stream(seed, ntime, order, num_repl):
    """The list that reproduces the counter-based generator: per cell (in
    order) and sample (t, s) the words of the Philox4x32-10 blocks
    (s, t + 1 | r << 2, cell, ntime) for replenish locations r = 1..NRepl,
    all four words of the first block and the first two of the others,
    each * 2^-32."""
    lib = load_oracle()
    key = (C.c_uint32 * 2)(seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    ctr, o = (C.c_uint32 * 4)(), (C.c_uint32 * 4)()
    words = []
    for cell, samples in enumerate(order):
        for t, s in samples:
            ctr[0], ctr[1], ctr[2], ctr[3] = s, t + 1, cell, ntime
            lib.orc_philox4x32_10(ctr, key, o)
            words += [o[0], o[1], o[2], o[3]]
            for r in range(2, num_repl + 1):
                ctr[1] = (t + 1) | (r << 2)
                lib.orc_philox4x32_10(ctr, key, o)
                words += [o[0], o[1]]
    return np.asarray(words, dtype=np.float64) * 2.0 ** -32


def per_sample(cfg):
    return 2 + 2 * cfg.num_repl
