"""The registration server (nrh_server_*): Aladin then F3D per request, requests admitted by a
device-memory budget and run concurrently.

Each request's result must be the same alone (one request at a time) and in a mix of concurrent
NMI and LNCC requests of different sizes, bit for bit; a request with a flat image fails alone
with NRH_ERR_NUMERIC (the reference stops the process there: reg-lib/cpu/_reg_blockMatching.cpp:
265-266) while the others finish; admission keeps the running requests' estimated memory within
the budget.
"""
import numpy as np
import pytest

from niftyreg_amd import _lib, phantoms
from niftyreg_amd.geometry import Geom
from niftyreg_amd.serve import Server, request_config

pytestmark = pytest.mark.gpu


def make_request(dims, seed):
    g = Geom.identity(dims)
    ref = phantoms.shell_phantom(dims, seed)
    A = phantoms.random_affine(seed + 500, rot_deg=4.0, trans_mm=2.0, scale=(0.97, 1.03), shear=0.01)
    flo = phantoms.bspline_warp(phantoms.affine_warp(ref, A), 5, 1.5, seed + 900)
    return g, ref.ravel(), g, flo.ravel()


def config(measure):
    return request_config(levels=2, maxit=15, aladin_levels=2, aladin_maxit=3, measure=measure)


REQUESTS = [((56, 48, 40), 1, _lib.MEASURE_NMI), ((48, 60, 44), 2, _lib.MEASURE_LNCC), ((64, 56, 48), 3, _lib.MEASURE_NMI),
            ((52, 52, 52), 4, _lib.MEASURE_LNCC), ((72, 48, 40), 5, _lib.MEASURE_NMI)]


def _same(a, b):
    assert a["status"] == b["status"]
    if a["status"] != _lib.NRH_OK:
        return
    assert np.array_equal(a["affine"], b["affine"])
    assert a["iterations"] == b["iterations"]
    assert a["similarity"] == b["similarity"]
    assert np.array_equal(a["cpp"], b["cpp"])


def test_results_same_alone_and_in_mix(gpu):
    reqs = [(make_request(d, s), config(m)) for d, s, m in REQUESTS]
    g = Geom.identity((48, 48, 48))
    flat = np.full(g.nvox, 0.5, np.float32)
    reqs.insert(2, ((g, flat, g, flat), config(_lib.MEASURE_NMI)))
    alone = []
    srv = Server(budget_bytes=64 << 30, max_active=1)
    try:
        for (pair, cfg) in reqs:
            alone.append(srv.wait(srv.submit(*pair, cfg)))
    finally:
        srv.close()
    srv = Server(budget_bytes=64 << 30, max_active=4)
    try:
        ids = [srv.submit(*pair, cfg) for (pair, cfg) in reqs]
        mixed = [srv.wait(i) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    print(f"mix: {st}")
    assert st["peak_active"] > 1
    assert alone[2]["status"] == _lib.NRH_ERR_NUMERIC and "active blocks" in alone[2]["error"]
    for q in range(len(reqs)):
        if q != 2:
            assert alone[q]["status"] == _lib.NRH_OK, alone[q]["error"]
            assert alone[q]["similarity"] > 0
        _same(alone[q], mixed[q])
    assert st["completed"] == len(reqs) - 1 and st["failed"] == 1


def test_admission_keeps_the_budget(gpu):
    reqs = [(make_request(d, s), config(m)) for d, s, m in REQUESTS[:3]]
    # room for one request at a time: requests wait for each other
    srv = Server(budget_bytes=100 << 20, max_active=4)
    try:
        res = [srv.wait(i) for i in [srv.submit(*p, c) for p, c in reqs]]
        st = srv.stats()
    finally:
        srv.close()
    assert all(r["status"] == _lib.NRH_OK for r in res)
    assert st["peak_active"] == 1 and st["peak_bytes"] <= st["budget_bytes"]
    for a, b in zip(res, res[1:]):
        assert b["started"] >= a["finished"] - 1e-6 or b["started"] >= a["started"]
    with pytest.raises(_lib.NiftyRegError):
        small = Server(budget_bytes=1 << 20, max_active=1)
        try:
            small.submit(*reqs[0][0], reqs[0][1])
        finally:
            small.close()
