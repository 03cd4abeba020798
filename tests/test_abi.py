"""The C-ABI library loads and exports every entry point include/cbm_hip.h
declares; configuration errors are reported before any device work.  CPU only."""
import ctypes as C
import os
import re

import pytest

import cbm._abi as A
from helpers import model, sim

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "cbm_hip.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(cbm_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert "cbm_create" in names and "cbm_step" in names and "cbm_adjoint" in names
    lib = A.lib()
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(A.SIGNATURES), set(names) ^ set(A.SIGNATURES)
    assert b"gfx950" in lib.cbm_version()


@pytest.mark.parametrize("bad", [
    dict(dims=4),
    dict(dims=2, integrator=9),
    dict(dims=2, neighbors=A.CELL_LINKED, cell_edge=0.0),
    dict(dims=3, neighbors=A.VERLET_DISPLACEMENT, skin=0.0, max_neighbors=8),
    dict(dims=3, neighbors=A.VERLET_DISPLACEMENT, skin=2.0, max_neighbors=0),
    dict(dims=2, dt=0.0),
])
def test_bad_configuration_is_rejected(bad):
    dims = bad.pop("dims")
    m = model(dims if dims <= 3 else 3, **bad)
    if dims > 3:
        m.dims = dims
    s = sim(10, [[0, 10]] * 3)
    h = C.c_void_p()
    rc = A.lib().cbm_create(C.byref(h), C.byref(m), 1, C.byref(s))
    assert rc == A.ERR_CONFIG
    assert A.lib().cbm_last_error(h)
    A.lib().cbm_destroy(h)


def test_capacity_below_n_is_rejected():
    m = model(2, neighbors=A.FULL)
    s = sim(10, [[0, 10]] * 2, n_max=5)
    h = C.c_void_p()
    assert A.lib().cbm_create(C.byref(h), C.byref(m), 1, C.byref(s)) == A.ERR_CONFIG
    A.lib().cbm_destroy(h)


def _create(m, s):
    h = C.c_void_p()
    rc = A.lib().cbm_create(C.byref(h), C.byref(m), 1, C.byref(s))
    A.lib().cbm_destroy(h)
    return rc


@pytest.mark.parametrize("n,alg", [((3, 8), A.MEDIUM_DGADI), ((8, 8), 7)])
def test_bad_medium_is_rejected(n, alg):
    m = model(2, neighbors=A.FULL)
    m.medium_n[0], m.medium_n[1] = n
    m.medium_alg = alg
    assert _create(m, sim(4, [[0, 8]] * 2)) == A.ERR_CONFIG


def test_unknown_rule_is_rejected():
    m = model(2, neighbors=A.FULL, rule=9)
    assert _create(m, sim(4, [[0, 8]] * 2)) == A.ERR_CONFIG


@pytest.mark.parametrize("law,k", [(A.LAW_CUSTOM, 5), (A.LAW_LINEAR, 1), (A.LAW_CUSTOM, -1)])
def test_bad_user_de_is_rejected(law, k):
    m = model(2, neighbors=A.FULL, law=law)
    m.user_de = k
    assert _create(m, sim(4, [[0, 8]] * 2)) == A.ERR_CONFIG


@pytest.mark.parametrize("walk", [2, 300, 4096])
def test_walk_option_outside_its_range_is_rejected(walk):
    m = model(3, neighbors=A.VERLET_DISPLACEMENT, skin=2.0, max_neighbors=8)
    m.walk = walk
    s = sim(10, [[0, 10]] * 3)
    h = C.c_void_p()
    rc = A.lib().cbm_create(C.byref(h), C.byref(m), 1, C.byref(s))
    assert rc == A.ERR_CONFIG
    assert b"walk" in A.lib().cbm_last_error(h)
