"""Benchmark problem families and their seeded ensemble inputs.

Right-hand sides are ordinary Python functions f(du, u, p, t) (traced to HIP by
odemi.tracer).  Input generators use numpy's PCG64 with a fixed seed per
workload, so the GPU engine and the CPU oracle see identical arrays.
"""
import numpy as np

from .api import ContinuousCallback, ODEFunction


def lorenz(du, u, p, t):
    x, y, z = u[0], u[1], u[2]
    du[0] = p[0] * (y - x)
    du[1] = x * (p[1] - z) - y
    du[2] = x * y - p[2] * z


def lorenz_jac(J, u, p, t):
    J[0, 0] = -p[0]; J[0, 1] = p[0]; J[0, 2] = 0.0
    J[1, 0] = p[1] - u[2]; J[1, 1] = -1.0; J[1, 2] = -u[0]
    J[2, 0] = u[1]; J[2, 1] = u[0]; J[2, 2] = -p[2]


def robertson(du, u, p, t):
    y1, y2, y3 = u[0], u[1], u[2]
    du[0] = -p[0] * y1 + p[2] * y2 * y3
    du[1] = p[0] * y1 - p[1] * (y2 * y2) - p[2] * y2 * y3
    du[2] = p[1] * (y2 * y2)


def robertson_jac(J, u, p, t):
    y2, y3 = u[1], u[2]
    J[0, 0] = -p[0]; J[0, 1] = p[2] * y3; J[0, 2] = p[2] * y2
    J[1, 0] = p[0]; J[1, 1] = -2.0 * p[1] * y2 - p[2] * y3; J[1, 2] = -p[2] * y2
    J[2, 0] = 0.0; J[2, 1] = 2.0 * p[1] * y2; J[2, 2] = 0.0


def chain(du, u, p, t):
    """Damped nonlinear oscillator chain, n = 2m, fixed ends x_0 = x_{m+1} = 0:
    x_i' = v_i,  v_i' = -k (2 x_i - x_{i-1} - x_{i+1}) - c v_i - alpha x_i^3."""
    m = len(u) // 2
    k, c, al = p[0], p[1], p[2]
    for i in range(m):
        du[i] = u[m + i]
    for i in range(m):
        x = u[i]
        s = 2.0 * x
        if i > 0:
            s = s - u[i - 1]
        if i < m - 1:
            s = s - u[i + 1]
        du[m + i] = -k * s - c * u[m + i] - al * (x * x * x)


def chain_jac(J, u, p, t):
    m = len(u) // 2
    k, c, al = p[0], p[1], p[2]
    for i in range(m):
        J[i, m + i] = 1.0
    for i in range(m):
        x = u[i]
        r = m + i
        J[r, i] = -k * 2.0 - al * (3.0 * (x * x))
        if i > 0:
            J[r, i - 1] = k
        if i < m - 1:
            J[r, i + 1] = k
        J[r, m + i] = -c


def ball(du, u, p, t):
    du[0] = u[1]
    du[1] = -9.81


def ball_jac(J, u, p, t):
    J[0, 0] = 0.0
    J[0, 1] = 1.0
    J[1, 0] = 0.0
    J[1, 1] = 0.0


def _bounce(integrator):
    integrator.u[1] = -integrator.p[0] * integrator.u[1]


LORENZ = ODEFunction(lorenz, jac=lorenz_jac)
BALL = ODEFunction(ball, jac=ball_jac)
# the floor: a downcrossing of the height flips the velocity with restitution p[0]
BALL_BOUNCE = ContinuousCallback(lambda u, t, integrator: u[0], None, _bounce, save_positions=(False, False))
ROBERTSON = ODEFunction(robertson, jac=robertson_jac)
CHAIN = ODEFunction(chain, jac=chain_jac)


def lorenz_small(N=1024, seed=1):
    """1,024 Lorenz trajectories: sigma=10, beta=8/3, rho~U[0,21], u0=(1,0,0), tspan (0,1)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rho = rng.uniform(0.0, 21.0, N)
    p = np.stack([np.full(N, 10.0), rho, np.full(N, 8.0 / 3.0)])
    u0 = np.stack([np.ones(N), np.zeros(N), np.zeros(N)])
    return dict(u0=u0, p=p, t0=np.zeros(N), tf=np.ones(N), reltol=1e-8, abstol=1e-8)


def lorenz_large(N=1048576, seed=2):
    """Lorenz: rho~U[0,28], sigma~U[8,12], beta=8/3, u0~U[-1,1]^3, t_end~U[1,10]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rho = rng.uniform(0.0, 28.0, N)
    sigma = rng.uniform(8.0, 12.0, N)
    u0 = rng.uniform(-1.0, 1.0, (3, N))
    tf = rng.uniform(1.0, 10.0, N)
    p = np.stack([sigma, rho, np.full(N, 8.0 / 3.0)])
    return dict(u0=u0, p=p, t0=np.zeros(N), tf=tf, reltol=1e-6, abstol=1e-6)


def robertson_ensemble(N=262144, seed=3):
    """Robertson: k=(0.04,3e7,1e4) each scaled by LogUniform[0.5,2], u0=(1,0,0), tspan (0,1e5)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    s = np.exp(rng.uniform(np.log(0.5), np.log(2.0), (3, N)))
    p = np.array([0.04, 3e7, 1e4])[:, None] * s
    u0 = np.stack([np.ones(N), np.zeros(N), np.zeros(N)])
    return dict(u0=u0, p=p, t0=np.zeros(N), tf=np.full(N, 1e5), reltol=1e-6, abstol=1e-10)


def chain_ensemble(N=16384, seed=4, m_lo=4, m_hi=64):
    """Oscillator chains: m~U{m_lo..m_hi}, n=2m, k~U[1,1e4], c~U[0.01,1], alpha~U[0,1],
    x0~N(0,0.1^2), v0=0, tspan (0,10).  u0 is packed to nmax = 2*m_hi rows (zeros beyond n)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = rng.integers(m_lo, m_hi + 1, N)
    k = rng.uniform(1.0, 1e4, N)
    c = rng.uniform(0.01, 1.0, N)
    al = rng.uniform(0.0, 1.0, N)
    nmax = 2 * m_hi
    u0 = np.zeros((nmax, N))
    x0 = rng.normal(0.0, 0.1, (m_hi, N))
    for j in range(N):
        u0[: m[j], j] = x0[: m[j], j]
    p = np.stack([k, c, al])
    return dict(u0=u0, p=p, t0=np.zeros(N), tf=np.full(N, 10.0), nvec=(2 * m).astype(np.int32),
                reltol=1e-6, abstol=1e-6)


def ball_ensemble(N=262144, seed=5):
    """Bouncing balls: h0~U[10,50], v0~U[-5,5], restitution~U[0.8,0.95], tspan (0,10);
    every trajectory bounces 1-5 times (the accumulation of bounces lies beyond t = 10)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    u0 = np.stack([rng.uniform(10.0, 50.0, N), rng.uniform(-5.0, 5.0, N)])
    p = rng.uniform(0.8, 0.95, (1, N))
    return dict(u0=u0, p=p, t0=np.zeros(N), tf=np.full(N, 10.0), reltol=1e-6, abstol=1e-6)
