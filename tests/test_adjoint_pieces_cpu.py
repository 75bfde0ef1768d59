"""The hand-written reverse-mode pieces (csrc/swe_adjoint.hpp), built for the
host in double precision, against central finite differences of the same
forward, and that forward against the FP64 oracle's face fluxes."""

from __future__ import annotations

import ctypes
import os
import zlib

import numpy as np
import pytest

from conftest import PKG_DIR

G, HG, DRY = 9.81, 0.5 * 9.81, 1.0e-4
P73 = float(np.float32(7.0 / 3.0))


@pytest.fixture(scope="module")
def hc():
    path = os.path.join(PKG_DIR, "hybrid_flood_amd", "libswe_hostcheck.so")
    if not os.path.exists(path):
        pytest.skip("libswe_hostcheck.so not built")
    lib = ctypes.CDLL(path)
    D = ctypes.POINTER(ctypes.c_double)
    lib.swe_hc_face_fwd.argtypes = [ctypes.c_int32, D, D]
    lib.swe_hc_face_vjp.argtypes = [ctypes.c_int32, D, D, D]
    lib.swe_hc_point_fwd.argtypes = [D, D]
    lib.swe_hc_point_vjp.argtypes = [D, D]
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def face_fwd(hc, xdir, x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros(5)
    hc.swe_hc_face_fwd(xdir, _p(x), _p(out))
    return out


def face_vjp(hc, xdir, x, gout):
    x = np.ascontiguousarray(x, np.float64)
    gout = np.ascontiguousarray(gout, np.float64)
    gin = np.zeros(6)
    hc.swe_hc_face_vjp(xdir, _p(x), _p(gout), _p(gin))
    return gin


STATE_IDX = [3, 4, 5, 8, 9, 10]  # hA, qnA, qtA, hB, qnB, qtB in the input vector


def _face_inputs(rng, regime):
    hA, hB = rng.uniform(0.05, 0.6, 2)
    bA, bB = rng.uniform(-0.2, 0.2, 2)
    uA, uB, vA, vB = rng.normal(0, 0.6, 4)
    aA = aB = 1
    if regime == "supersonic_right":
        uA, uB = 4.0, 3.5
    elif regime == "supersonic_left":
        uA, uB = -4.0, -3.6
    elif regime == "wallA":
        aB = 0
    elif regime == "wallB":
        aA = 0
    elif regime == "step":  # B's bed above A's surface: B side dries in reconstruction
        bB = hA + bA + 0.3
    return np.array([G, HG, DRY, hA, hA * uA, hA * vA, bA, aA, hB, hB * uB, hB * vB, bB, aB], np.float64)


@pytest.mark.parametrize("xdir", [1, 0])
@pytest.mark.parametrize("regime", ["sub", "supersonic_right", "supersonic_left", "wallA", "wallB", "step"])
def test_face_vjp_matches_finite_differences(hc, xdir, regime):
    rng = np.random.default_rng(zlib.crc32(f"{xdir}{regime}".encode()))
    for trial in range(20):
        x = _face_inputs(rng, regime)
        gout = rng.normal(size=5)
        gin = face_vjp(hc, xdir, x, gout)
        fd = np.zeros(6)
        for k, idx in enumerate(STATE_IDX):
            eps = 1e-6 * max(1.0, abs(x[idx]))
            xp, xm = x.copy(), x.copy()
            xp[idx] += eps
            xm[idx] -= eps
            fd[k] = (face_fwd(hc, xdir, xp) - face_fwd(hc, xdir, xm)) @ gout / (2 * eps)
        np.testing.assert_allclose(gin, fd, rtol=1e-5, atol=1e-7, err_msg=f"{regime} trial {trial}")


def test_irrelevant_face_has_no_gradient(hc):
    x = _face_inputs(np.random.default_rng(0), "sub")
    x[7] = x[12] = 0
    assert np.all(face_vjp(hc, 1, x, np.ones(5)) == 0)


@pytest.mark.parametrize("xdir", [1, 0])
def test_face_fwd_matches_fp64_oracle(hc, xdir):
    from oracle import swe_oracle as o

    rng = np.random.default_rng(5 + xdir)
    for regime in ("sub", "supersonic_right", "wallA", "wallB", "step"):
        x = _face_inputs(rng, regime)
        a = lambda v: np.array([v], np.float64)  # noqa: E731
        want = o.face_fluxes(a(x[3]), a(x[4]), a(x[5]), a(x[6]), np.array([bool(x[7])]),
                             a(x[8]), a(x[9]), a(x[10]), a(x[11]), np.array([bool(x[12])]),
                             np.float64(G), np.float64(HG), np.float64(DRY), "x" if xdir else "y")
        got = face_fwd(hc, xdir, x)
        np.testing.assert_allclose(got, [w[0] for w in want], rtol=1e-13, atol=1e-15)


def point_fwd(hc, x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros(3)
    hc.swe_hc_point_fwd(_p(x), _p(out))
    return out


@pytest.mark.parametrize("regime", ["wet", "clip", "shallow", "still"])
def test_point_vjp_matches_finite_differences(hc, regime):
    rng = np.random.default_rng(zlib.crc32(regime.encode()))
    for _ in range(30):
        hs = rng.uniform(0.01, 0.5)
        hus, hvs = rng.normal(0, 0.2, 2)
        if regime == "clip":
            hs = -rng.uniform(1e-4, 1e-2)
        elif regime == "shallow":
            hs = rng.uniform(2 * DRY, 5e-3)
        elif regime == "still":
            hus = hvs = 0.0
        n, dtrate, dtg = rng.uniform(0.01, 0.06), rng.uniform(0, 1e-5), rng.uniform(0.1, 5.0)
        lam = rng.normal(size=3)
        x = np.array([DRY, P73, hs, hus, hvs, n, 1, dtrate, dtg, *lam], np.float64)
        out = np.zeros(5)
        hc.swe_hc_point_vjp(_p(x), _p(out))
        # finite differences of point_fwd in (hs, hus, hvs, n); rain enters hs additively
        fwd_in = lambda hs_, hus_, hvs_, n_: np.array([DRY, P73, hs_, hus_, hvs_, n_, 1, dtg])  # noqa: E731
        def f(hs_, hus_, hvs_, n_):
            return point_fwd(hc, fwd_in(hs_, hus_, hvs_, n_)) @ lam
        fd = []
        base = [hs, hus, hvs, n]
        for k in range(4):
            eps = 1e-7 * max(1.0, abs(base[k]))
            bp, bm = list(base), list(base)
            bp[k] += eps
            bm[k] -= eps
            fd.append((f(*bp) - f(*bm)) / (2 * eps))
        # still water sits on the |q| kink: the central difference there is off by
        # O(K/q * eps); compare it with a looser bar
        rtol = 1e-3 if regime == "still" else 1e-5
        np.testing.assert_allclose(out[:3], fd[:3], rtol=rtol, atol=1e-8)
        np.testing.assert_allclose(out[3], fd[3], rtol=1e-5, atol=1e-8)
        np.testing.assert_allclose(out[4], out[0] * dtrate, rtol=1e-15)


def test_point_vjp_inactive_cell_is_zero(hc):
    x = np.array([DRY, P73, 0.3, 0.1, 0.2, 0.03, 0, 1e-6, 1.0, 1.0, 1.0, 1.0])
    out = np.ones(5)
    hc.swe_hc_point_vjp(_p(x), _p(out))
    assert np.all(out == 0)
