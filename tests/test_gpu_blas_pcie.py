"""The BLAS-1 primitives (BlasHandle::{dotc_c64, axpy_c64, scal},
src/device/blas.rs:24-241) against known answers, and the PcieAccount
invariant of the diagonalize hot path (src/scf.rs:240-251,
src/device/pcie.rs:21-70) through the context's transfer counters.

Every host<->device copy the library makes is counted. The reference
expects the filter + Rayleigh-Ritz chain to move nothing but psi in and psi
plus the eigenvalues out (its assertion miscounts the eigenvalues,
discrepancy D4). Here the filter moves exactly its window {c, 1/e} in
(16 B) and its divergence flags out once per call (32 B: iteration, norms,
growth), instead of the reference's two norm read-backs per step; Rayleigh-
Ritz moves the Cholesky status (4 B) and the n_b eigenvalues (8 n_b B)
out and nothing in."""
import ctypes as C

import numpy as np
import pytest

import chs
import pyoracle as O

pytestmark = pytest.mark.gpu

dp = C.POINTER(C.c_double)


def _c(a):
    return np.ascontiguousarray(a, np.complex128)


def test_zdotc_kat(ctx):
    # <a|b> = conj(1+i) 2 + conj(2i) (1-i) = -4i (cg_helpers.rs:651-666)
    a = chs.Gpu.from_host(ctx, _c([1 + 1j, 2j]))
    b = chs.Gpu.from_host(ctx, _c([2, 1 - 1j]))
    out = np.zeros(2)
    ctx.check(chs.lib().chs_zdotc(ctx.h, 2, a.ptr, b.ptr,
                                  out.ctypes.data_as(dp)))
    assert out[0] == 0.0 and out[1] == -4.0
    # long vector against numpy (grid-stride partial sums, fixed order)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(100_001) + 1j * rng.standard_normal(100_001)
    y = rng.standard_normal(100_001) + 1j * rng.standard_normal(100_001)
    dx, dy = chs.Gpu.from_host(ctx, x), chs.Gpu.from_host(ctx, y)
    ctx.check(chs.lib().chs_zdotc(ctx.h, len(x), dx.ptr, dy.ptr, out.ctypes.data_as(dp)))
    ref = np.vdot(x, y)
    assert abs(complex(*out) - ref) < 1e-11 * abs(ref)


def test_zaxpy_zscal_kat(ctx):
    # y = 2 x on x = [1, 2, 3, 4] (the reference's axpy KAT, blas.rs:352-368)
    x = chs.Gpu.from_host(ctx, _c([1, 2, 3, 4]))
    y = chs.Gpu.from_host(ctx, np.zeros(4, complex))
    alpha = np.array([2.0, 0.0])
    ctx.check(chs.lib().chs_zaxpy(ctx.h, 4, alpha.ctypes.data_as(dp), x.ptr,
                                  y.ptr))
    assert np.array_equal(y.sync_to_host(), np.array([2, 4, 6, 8], complex))
    # complex alpha: y += i x
    alpha = np.array([0.0, 1.0])
    ctx.check(chs.lib().chs_zaxpy(ctx.h, 4, alpha.ctypes.data_as(dp), x.ptr,
                                  y.ptr))
    assert np.array_equal(y.sync_to_host(), np.array([2 + 1j, 4 + 2j, 6 + 3j, 8 + 4j]))
    # x <- (0 + i)(1 + 2i) = -2 + i
    z = chs.Gpu.from_host(ctx, _c([1 + 2j, -3j]))
    ctx.check(chs.lib().chs_zscal(ctx.h, 2, alpha.ctypes.data_as(dp), z.ptr))
    assert np.array_equal(z.sync_to_host(), np.array([-2 + 1j, 3 + 0j]))


@pytest.mark.parametrize("npa", [0, 4])
def test_pcie_account_filter_and_rayleigh_ritz(npa):
    c = chs.Context(0)
    p = O.default_params(n_rep=1, n_grid=36, n_bands=16, n_proj_per_atom=npa,
                         q0=0.25 if npa else 0.0)
    m = O.Model(p, O.positions(1))
    mil, idx, kin = m.pw()
    H = chs.Hamiltonian(c, 36, mil, kin, m.vion())
    if npa:
        H.bind_projectors(chs.Gpu.from_host(c, m.beta()), chs.Gpu.from_host(c, m.d()),
                          chs.Gpu.from_host(c, np.full(m.n_proj, 0.25)))
    psi = chs.Gpu.from_host(c, m.initial_psi(16))
    b4 = O.spectral_bounds(None, m.grid_lambda_max())
    # warm-up: first-use tables (twiddles, GEMM choices) move set-up bytes
    y = chs.chebyshev_filter(c, psi, 8, b4)
    chs.rayleigh_ritz(c, y)
    c.synchronize()
    c.transfer_reset()
    y = chs.chebyshev_filter(c, psi, 8, b4)
    c.synchronize()
    assert c.transfer_bytes() == (16, 32)
    c.transfer_reset()
    eps, _ = chs.rayleigh_ritz(c, y)
    c.synchronize()
    assert c.transfer_bytes() == (0, 4 + 16 * 8)
    assert np.all(np.diff(eps) >= 0)
    del H, psi, y
    c.close()
