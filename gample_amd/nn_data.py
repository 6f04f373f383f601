"""Training data for the learned Poisson warm start / preconditioner.

The network is trained on (b, x) pairs of the solver's SPD system S x = b.
Pairs come from the simulator itself (Context.nn_add_pair after a field
solve) and, to cover the spectra the flexible CG shows the network (its
residuals get rougher as it converges), from random right-hand sides with a
power-law spectrum, solved by MG-CG (Context.solve_rhs) to 1e-8.
"""
import numpy as np


def random_rhs(mx, my, periodic_x, periodic_y, rng, slope=None):
    """A random right-hand side on the unknown grid (my x mx, x fastest):
    Fourier (periodic axis) or sine (Dirichlet axis) modes with amplitudes
    ~ k^-slope, slope drawn from [0, 3] unless given; zero mean if both axes
    are periodic."""
    if slope is None:
        slope = rng.uniform(0.0, 3.0)

    def axis_basis(m, per):
        j = np.arange(m)
        if per:
            ks = np.arange(1, m // 2 + 1)
            B = np.concatenate([np.cos(2 * np.pi * np.outer(ks, j) / m), np.sin(2 * np.pi * np.outer(ks, j) / m)])
            kk = np.concatenate([ks, ks]).astype(float)
        else:
            ks = np.arange(1, m + 1)
            B = np.sin(np.pi * np.outer(ks, j + 1) / (m + 1))
            kk = ks.astype(float) / 2.0
        return B, kk

    Bx, kx = axis_basis(mx, periodic_x)
    By, ky = axis_basis(my, periodic_y)
    kk = np.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2)
    c = rng.standard_normal(kk.shape) * kk ** (-slope)
    b = By.T @ c @ Bx
    if periodic_x and periodic_y:
        b -= b.mean()
    return (b / np.sqrt(np.mean(b * b))).ravel()


def add_random_pairs(ctx, n, rng, periodic_x, periodic_y, tol=1e-8):
    """Solve n random right-hand sides with MG-CG and append the pairs to the
    context's training set; returns the arrays (b, x)."""
    from . import F_SOL
    my, mx = ctx.field_shape(F_SOL)[0]
    bs, xs = [], []
    for _ in range(n):
        b = random_rhs(mx, my, periodic_x, periodic_y, rng)
        x, st, rc = ctx.solve_rhs(b, tol=tol, maxiter=400, solver=0)
        bs.append(b)
        xs.append(x)
    bs, xs = np.array(bs), np.array(xs)
    ctx.nn_train(epochs=0, rho=bs, phi=xs)
    return bs, xs
