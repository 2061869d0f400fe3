"""Shared helpers of the GPU parity tests."""
import numpy as np

import oracle
from cuco_amd import engine as E
from cuco_amd import ics


def oracle_params(box, nthreads=0):
    return oracle.params(box.M, box.L, ics.G_INTERNAL, box.asmth, box.rcut,
                         box.soft_h, nthreads=nthreads)


def make_engine(box, capacity=None):
    return E.Engine(max_particles=capacity or box.N, **box.engine_kwargs())


def rel_err(a, ref, denom):
    """Per-particle |a - ref| / |denom| (denominator: |a_ref,total|)."""
    return np.linalg.norm(a - ref, axis=1) / np.linalg.norm(denom, axis=1)


def percentiles(err):
    return dict(median=float(np.median(err)), p99=float(np.percentile(err, 99)),
                max=float(err.max()))


def halo_box(n_side=32, M=64, L_mpc=100.0, n_halos=20, seed=77):
    return ics.small_box(f"halo{n_side}", n_side, M, L_mpc, seed, "halos",
                         n_halos=n_halos)
