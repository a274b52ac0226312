"""Shared test helpers: small workloads and error metrics."""
import functools

import numpy as np

from thrid_amd import abi, systems


def rms_rel(a, b):
    """rms relative force error: sqrt(<|a-b|^2> / <|b|^2>)."""
    return float(np.sqrt(np.mean(np.sum((a - b) ** 2, 1)) / np.mean(np.sum(b ** 2, 1))))


def ene(r, name):
    return float(r["ene"][abi.ENE_NAMES.index(name)])


@functools.lru_cache(maxsize=None)
def water648():
    return systems.water_box()


@functools.lru_cache(maxsize=None)
def small_slab(form="rb"):
    """1.4k-atom water/1-propanol slab with bonded terms and 1-4 pairs."""
    s = systems.propanol_water_slab(nwater=300, nprop=40, box=(24.0, 24.0, 60.0),
                                    nfft=(24, 24, 60), rcut=10.0, rcut_book=11.0, alpha=0.3,
                                    torsion_form=form)
    # displace atoms off the template geometry so every bonded term has a force
    rng = np.random.default_rng(7)
    s["pos"] = s["pos"] + rng.normal(scale=0.05, size=s["pos"].shape)
    return s


@functools.lru_cache(maxsize=None)
def butane_box():
    s = systems.mixture_box(2000, "butane", seed=11, rcut=9.0, rcut_book=10.0)
    rng = np.random.default_rng(8)
    s["pos"] = s["pos"] + rng.normal(scale=0.05, size=s["pos"].shape)
    return s


def pairs_from_csr(idx, lst, n):
    """Reference CSR (1-based, 0 sentinels) -> sorted (i, j) pairs."""
    out = []
    for i in range(1, n):
        s, e = idx[i - 1], idx[i] if i < n else len(lst) + 1
        for k in range(s - 1, e - 1):
            j = lst[k]
            if j > 0:
                out.append((i, j))
    out.sort()
    return np.array(out, np.int32).reshape(-1, 2)
