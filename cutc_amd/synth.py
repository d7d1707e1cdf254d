"""Seeded synthetic systems for the transcorrelated integral path.

Arrays are numpy, C-ordered with the reversed shape of the Fortran arrays that
``cutc_int`` takes (reference ``src/cutc_module.f90:23-35``), so their memory is
exactly the Fortran layout:

=============  =====================  ==========================
field          numpy shape            Fortran array
=============  =====================  ==========================
``rn``         (n_nuc, 3)             rn(3, n_nuc)
``r1, wr1``    (n1, 3), (n1,)         r1(3, n_grid1), wr1(n_grid1)
``r2, wr2``    (n2, 3), (n2,)         r2(3, n_grid2), wr2(n_grid2)
``aos1``       (4, n_ao, n1)          aos_data1(n_grid1, n_ao, 4)
``aos2``       (4, n_ao, n2)          aos_data2(n_grid2, n_ao, 4)
``c_bh`` ...   (n_nuc, size_bh)       c_bh(size_bh, n_nuc)
=============  =====================  ==========================

The numbers come from ``libcutcsynth.so`` (``cutc_amd/csrc/synth.c``).
"""
from __future__ import annotations

import ctypes
import dataclasses
import os

import numpy as np

from . import _lib

BASE_SEED = 20261203

# nuclei of the 3-nucleus system (bohr)
TRIATOMIC = np.array([[0.0, 0.0, 0.0], [0.0, 1.43, 1.11], [0.0, -1.43, 1.11]])

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


_SYNTH = None


def _synth():
    global _SYNTH
    if _SYNTH is None:
        lib = ctypes.CDLL(_lib.lib_path("libcutcsynth.so"))
        lib.synth_nuclei_cube.restype = ctypes.c_int
        lib.synth_nuclei_cube.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]
        lib.synth_basis.restype = ctypes.c_int
        lib.synth_basis.argtypes = [ctypes.c_uint64, ctypes.c_int, _ip, _ip, ctypes.c_int, _ip, _ip, _dp, _dp]
        lib.synth_grid.restype = None
        lib.synth_grid.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
        lib.synth_aos.restype = None
        lib.synth_aos.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _ip, _ip, _dp, _dp, _dp]
        lib.synth_jastrow.restype = None
        lib.synth_jastrow.argtypes = [ctypes.c_uint64, ctypes.c_int, _dp, _ip, _ip, _ip]
        _SYNTH = lib
    return _SYNTH


@dataclasses.dataclass
class System:
    """One molecule-like system: nuclei, AO basis, two grids, Jastrow table."""

    rn: np.ndarray
    r1: np.ndarray
    wr1: np.ndarray
    r2: np.ndarray
    wr2: np.ndarray
    aos1: np.ndarray
    aos2: np.ndarray
    c_bh: np.ndarray
    m_bh: np.ndarray
    n_bh: np.ndarray
    o_bh: np.ndarray
    seed: int = 0
    name: str = ""
    # the AO basis behind aos1 / aos2 in the arrays cutc_aos takes (numpy C
    # order of the Fortran layouts): ao_nucl (n_ao,) 1-based, ao_power
    # (3, n_ao), ao_prim_num (n_ao,), ao_expo / ao_coef (n_ao, prim_max)
    basis: dict | None = None

    @property
    def n_nuc(self):
        return self.rn.shape[0]

    @property
    def n_ao(self):
        return self.aos1.shape[1]

    @property
    def n1(self):
        return self.r1.shape[0]

    @property
    def n2(self):
        return self.r2.shape[0]

    @property
    def size_bh(self):
        return self.c_bh.shape[1]

    def subset_r1(self, n: int) -> "System":
        """The same system with only the first ``n`` grid-1 points."""
        return dataclasses.replace(
            self, r1=np.ascontiguousarray(self.r1[:n]), wr1=np.ascontiguousarray(self.wr1[:n]),
            aos1=np.ascontiguousarray(self.aos1[:, :, :n]), name=f"{self.name}[r1<{n}]")

    def subset_r2(self, n: int) -> "System":
        return dataclasses.replace(
            self, r2=np.ascontiguousarray(self.r2[:n]), wr2=np.ascontiguousarray(self.wr2[:n]),
            aos2=np.ascontiguousarray(self.aos2[:, :, :n]), name=f"{self.name}[r2<{n}]")

    def shape_str(self):
        return f"n_nuc={self.n_nuc} n_ao={self.n_ao} n1={self.n1} n2={self.n2} size_bh={self.size_bh}"


def round_robin_shells(n_nuc: int, n_ao: int):
    """Shells s, p, d handed out round-robin over the nuclei (all nuclei get an
    s shell, then a p shell, then a d shell, then again), until n_ao AOs."""
    ncart = (1, 3, 6)
    shells, tot, t = [], 0, 0
    while tot < n_ao:
        L = (t // n_nuc) % 3
        shells.append((t % n_nuc, L))
        tot += ncart[L]
        t += 1
    return shells


def make_system(n_nuc: int, n_ao: int, n1: int, n2: int, seed: int, nuclei=None, shells=None,
                name: str = "") -> System:
    """Draw a system.  ``nuclei``: (n_nuc, 3) array, or None for the
    12-bohr cube with 1.8-bohr minimum spacing.  ``shells``: list of
    (nucleus, L), or None for :func:`round_robin_shells`."""
    lib = _synth()
    if nuclei is None:
        rn = np.zeros((n_nuc, 3))
        if lib.synth_nuclei_cube(seed, n_nuc, 12.0, 1.8, _ptr(rn)) != 0:
            raise RuntimeError("could not place the nuclei")
    else:
        rn = np.ascontiguousarray(np.asarray(nuclei, dtype=np.float64).reshape(n_nuc, 3))
    if shells is None:
        shells = round_robin_shells(n_nuc, n_ao)
    sh_nuc = np.array([s[0] for s in shells], dtype=np.int32)
    sh_L = np.array([s[1] for s in shells], dtype=np.int32)
    ao_nuc = np.zeros(n_ao, np.int32)
    ao_pow = np.zeros(3 * n_ao, np.int32)
    ao_alpha = np.zeros(n_ao)
    ao_norm = np.zeros(n_ao)
    got = lib.synth_basis(seed, len(shells), _ptr(sh_nuc, _ip), _ptr(sh_L, _ip), n_ao, _ptr(ao_nuc, _ip),
                          _ptr(ao_pow, _ip), _ptr(ao_alpha), _ptr(ao_norm))
    if got != n_ao:
        raise RuntimeError(f"shell list gives {got} AOs, wanted {n_ao}")

    def grid(stream, n):
        r = np.zeros((n, 3))
        w = np.zeros(n)
        lib.synth_grid(seed, stream, n_nuc, _ptr(rn), n, _ptr(r), _ptr(w))
        aos = np.zeros((4, n_ao, n))
        lib.synth_aos(n, _ptr(r), n_ao, _ptr(rn), _ptr(ao_nuc, _ip), _ptr(ao_pow, _ip), _ptr(ao_alpha),
                      _ptr(ao_norm), _ptr(aos))
        return r, w, aos

    r1, wr1, aos1 = grid(1, n1)
    r2, wr2, aos2 = grid(2, n2)
    c = np.zeros((n_nuc, 6))
    m = np.zeros((n_nuc, 6), np.int32)
    n = np.zeros((n_nuc, 6), np.int32)
    o = np.zeros((n_nuc, 6), np.int32)
    lib.synth_jastrow(seed, n_nuc, _ptr(c), _ptr(m, _ip), _ptr(n, _ip), _ptr(o, _ip))
    basis = {"ao_nucl": ao_nuc + 1, "ao_power": np.ascontiguousarray(ao_pow.reshape(n_ao, 3).T),
             "ao_prim_num": np.ones(n_ao, np.int32), "ao_expo": ao_alpha.reshape(n_ao, 1).copy(),
             "ao_coef": ao_norm.reshape(n_ao, 1).copy()}
    return System(rn, r1, wr1, r2, wr2, aos1, aos2, c, m, n, o, seed=seed, name=name, basis=basis)


def triatomic(n1=4000, n2=16000, seed=BASE_SEED + 0):
    """3 nuclei at (0,0,0) (0,1.43,1.11) (0,-1.43,1.11); 24 AOs: nucleus 0
    gets s,s,p,p,d and nuclei 1, 2 get s,s,p."""
    shells = [(0, 0), (0, 0), (0, 1), (0, 1), (0, 2), (1, 0), (1, 0), (1, 1), (2, 0), (2, 0), (2, 1)]
    return make_system(3, 24, n1, n2, seed, nuclei=TRIATOMIC, shells=shells, name=f"triatomic-{n1}x{n2}")


def cube20(n1=60000, n2=200000, seed=BASE_SEED + 1):
    """20 nuclei in a 12-bohr cube, n_ao = 200 (s, p, d on every nucleus)."""
    return make_system(20, 200, n1, n2, seed, name=f"cube20-{n1}x{n2}")


def cube40(n1=80000, n2=300000, seed=BASE_SEED + 4):
    """40 nuclei in a 12-bohr cube, n_ao = 300."""
    return make_system(40, 300, n1, n2, seed, name=f"cube40-{n1}x{n2}")


def batch_systems(count=64, seed=BASE_SEED + 3, scale=1.0):
    """Independent systems with n_nuc U{3..12}, n_ao U{30..150},
    n1 U{5k..30k}, n2 U{20k..100k} (grid sizes times ``scale``)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nn = int(rng.integers(3, 13))
        nao = int(rng.integers(30, 151))
        n1 = max(1, int(int(rng.integers(5000, 30001)) * scale))
        n2 = max(1, int(int(rng.integers(20000, 100001)) * scale))
        out.append(make_system(nn, nao, n1, n2, seed * 1000 + i, name=f"batch{i}"))
    return out
