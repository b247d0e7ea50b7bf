"""Initial-condition generators for the benchmark / parity workloads.

All generators are deterministic and return global fields in the layout of
``kratos_amd.mesh`` (u: (5, Nz, Ny, Nx); face fields bx, by, bz).  The
random numbers come from SplitMix64 (Steele, Lea & Flood 2014) so that the
same seed gives the same state on every machine.

* ``sod``        -- Sod shock tube (config 1: 512x2x1, gamma 1.4).
* ``turbulence`` -- seeded Fourier-mode velocity plus a face-centred curl of
  a seeded edge vector potential (div B = 0 to round-off), rescaled to the
  requested rms Mach number and plasma beta (configs 2, 3, 5).
* ``blast``      -- MHD blast wave in a uniform oblique field (the refined-mesh blast).
* ``orszag_tang``, ``brio_wu`` -- the 2-D and 1-D MHD classics.
"""
import math

import numpy as np

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = int(seed) & _MASK

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & _MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self):
        """Uniform double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)


def _modes(rng, n_modes, k_max):
    """n_modes (k, amplitude vector, phase) with k in Z^3, 0 < |k| <= k_max."""
    out = []
    span = 2 * k_max + 1
    while len(out) < n_modes:
        k = [int(rng.uniform() * span) - k_max for _ in range(3)]
        k2 = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
        if k2 == 0 or k2 > k_max * k_max:
            continue
        amp = [2.0 * rng.uniform() - 1.0 for _ in range(3)]
        phase = 2.0 * math.pi * rng.uniform()
        out.append((k, amp, phase))
    return out


def _fields(modes, comps, xs, ys, zs):
    """sum_m amp_m[c] * sin(2 pi k_m . x + phase_m) on the grid xs x ys x zs for
    each c in comps, evaluated separably: sin(t_yz + t_x) = S_yz cos t_x + C_yz sin t_x.
    Uses torch (multi-threaded, CPU) when importable, numpy otherwise."""
    try:
        import torch
        xp, T = torch, True
        to = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    except Exception:  # pragma: no cover - numpy fallback for IC generation only
        xp, T = np, False
        to = lambda a: a
    xs, ys, zs = to(np.asarray(xs, dtype=np.float64)), to(np.asarray(ys, dtype=np.float64)), \
        to(np.asarray(zs, dtype=np.float64))
    shape = (len(zs), len(ys), len(xs))
    out = [xp.zeros(shape, dtype=xp.float64) for _ in comps]
    two_pi = 2.0 * math.pi
    for k, amp, ph in modes:
        tx = two_pi * k[0] * xs
        tyz = two_pi * (k[1] * ys[None, :] + k[2] * zs[:, None]) + ph
        cx, sx = xp.cos(tx), xp.sin(tx)
        cyz, syz = xp.cos(tyz), xp.sin(tyz)
        s3 = syz[:, :, None] * cx[None, None, :]
        if T:
            s3.addcmul_(cyz[:, :, None], sx[None, None, :])
        else:
            s3 += cyz[:, :, None] * sx[None, None, :]
        for o, c in zip(out, comps):
            if amp[c] == 0.0:
                continue
            if T:
                o.add_(s3, alpha=amp[c])
            else:
                o += amp[c] * s3
    return [o.numpy() if T else o for o in out]


def cell_centred_b(bx, by, bz):
    return (0.5 * (bx[:, :, :-1] + bx[:, :, 1:]), 0.5 * (by[:, :-1, :] + by[:, 1:, :]),
            0.5 * (bz[:-1] + bz[1:]))


def total_energy(rho, pre, vel, bcc, gamma):
    """E = p/(gamma-1) + rho v^2/2 + |Bcc|^2/2 with the c2p's Bcc convention."""
    e = pre / (gamma - 1.0) + 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
    if bcc is not None:
        e = e + 0.5 * (bcc[0] ** 2 + bcc[1] ** 2 + bcc[2] ** 2)
    return e


def turbulence(N, seed=0, gamma=5.0 / 3.0, mach=0.5, beta=1.0, n_modes=32, k_max=4,
               beta_range=None):
    """Seeded MHD turbulence in the periodic unit box with N = (Nx, Ny, Nz) cells.

    rho = 1, p = 1/gamma (c_s = 1).  Velocity: n_modes Fourier modes with
    0 < |k| <= k_max, rescaled to rms Mach ``mach``.  B: face-centred discrete
    curl of an edge vector potential built the same way (each component its
    own n_modes), rescaled to plasma beta = 2 p / <|Bcc|^2>.  With
    ``beta_range=(lo, hi)`` beta is drawn uniformly from the same stream
    after the modes.  Returns (u, (bx, by, bz), info).
    """
    Nx, Ny, Nz = N
    rng = SplitMix64(seed)
    vmodes = _modes(rng, n_modes, k_max)
    amodes = [_modes(rng, n_modes, k_max) for _ in range(3)]
    if beta_range is not None:
        beta = beta_range[0] + (beta_range[1] - beta_range[0]) * rng.uniform()
    xc = (np.arange(Nx) + 0.5) / Nx
    yc = (np.arange(Ny) + 0.5) / Ny
    zc = (np.arange(Nz) + 0.5) / Nz
    xf = np.arange(Nx + 1) / Nx
    yf = np.arange(Ny + 1) / Ny
    zf = np.arange(Nz + 1) / Nz
    rho = np.ones((Nz, Ny, Nx))
    pre = np.full((Nz, Ny, Nx), 1.0 / gamma)
    vel = _fields(vmodes, (0, 1, 2), xc, yc, zc)
    vrms = math.sqrt(float(np.mean(vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)))
    cs = math.sqrt(gamma * (1.0 / gamma) / 1.0)
    s = mach * cs / vrms
    vel = [v * s for v in vel]
    # edge potential: A_x at (xc, yf, zf), A_y at (xf, yc, zf), A_z at (xf, yf, zc)
    ax = _fields(amodes[0], (0,), xc, yf, zf)[0]
    ay = _fields(amodes[1], (1,), xf, yc, zf)[0]
    az = _fields(amodes[2], (2,), xf, yf, zc)[0]
    dx, dy, dz = 1.0 / Nx, 1.0 / Ny, 1.0 / Nz
    bx = (az[:, 1:, :] - az[:, :-1, :]) / dy - (ay[1:, :, :] - ay[:-1, :, :]) / dz
    by = (ax[1:, :, :] - ax[:-1, :, :]) / dz - (az[:, :, 1:] - az[:, :, :-1]) / dx
    bz = (ay[:, :, 1:] - ay[:, :, :-1]) / dx - (ax[:, 1:, :] - ax[:, :-1, :]) / dy
    bcc = cell_centred_b(bx, by, bz)
    b2 = float(np.mean(bcc[0] ** 2 + bcc[1] ** 2 + bcc[2] ** 2))
    sb = math.sqrt(2.0 * (1.0 / gamma) / beta / b2)
    bx, by, bz = bx * sb, by * sb, bz * sb
    bcc = cell_centred_b(bx, by, bz)
    u = np.empty((5, Nz, Ny, Nx))
    u[0] = rho
    u[1] = total_energy(rho, pre, vel, bcc, gamma)
    for c in range(3):
        u[2 + c] = rho * vel[c]
    return u, (bx, by, bz), {"beta": beta, "mach": mach, "seed": seed}


def sod(N=(512, 2, 1), gamma=1.4, x_split=0.5, left=(1.0, 1.0), right=(0.125, 0.1),
        x_min=0.0, x_max=1.0):
    """Sod tube along x: (rho, p) = left for x < x_split, right otherwise."""
    Nx, Ny, Nz = N
    dx = (x_max - x_min) / Nx
    xc = x_min + (np.arange(Nx) + 0.5) * dx
    rho = np.where(xc < x_split, left[0], right[0])
    pre = np.where(xc < x_split, left[1], right[1])
    u = np.zeros((5, Nz, Ny, Nx))
    u[0] = rho[None, None, :]
    u[1] = (pre / (gamma - 1.0))[None, None, :]
    return u


def blast(N, gamma=5.0 / 3.0, rho0=1.0, p0=0.1, p_in=10.0, r_in=0.1,
          b0=(1.0 / math.sqrt(2.0), 1.0 / math.sqrt(2.0), 0.0),
          x_min=(-0.5, -0.5, -0.5), x_max=(0.5, 0.5, 0.5)):
    """MHD blast: uniform rho, p, oblique B; p = p_in inside radius r_in."""
    Nx, Ny, Nz = N
    xs = [x_min[a] + (np.arange(N[a]) + 0.5) * (x_max[a] - x_min[a]) / N[a] for a in range(3)]
    r2 = (xs[2][:, None, None] ** 2 + xs[1][None, :, None] ** 2 + xs[0][None, None, :] ** 2)
    pre = np.where(r2 < r_in * r_in, p_in, p0)
    rho = np.full((Nz, Ny, Nx), rho0)
    bx = np.full((Nz, Ny, Nx + 1), b0[0])
    by = np.full((Nz, Ny + 1, Nx), b0[1])
    bz = np.full((Nz + 1, Ny, Nx), b0[2])
    u = np.zeros((5, Nz, Ny, Nx))
    u[0] = rho
    u[1] = total_energy(rho, pre, [0 * rho] * 3, cell_centred_b(bx, by, bz), gamma)
    return u, (bx, by, bz)


def orszag_tang(N=(256, 256, 1), gamma=5.0 / 3.0):
    """Orszag-Tang vortex (Orszag & Tang 1979, in the normalisation of
    Stone et al. 2008, ApJS 178, sec. 5.5) on the periodic unit square:
    rho = 25/(36 pi), p = 5/(12 pi), v = (-sin 2 pi y, sin 2 pi x, 0) and
    B = B0 (-sin 2 pi y, sin 4 pi x, 0) with B0 = 1/sqrt(4 pi).  bx is
    sampled at the x-faces (a function of y only) and by at the y-faces (of
    x only), so the discrete div B is exactly zero.  N[2] = 1 gives the 2-D
    problem (bz has its two coincident z-faces)."""
    Nx, Ny, Nz = N
    xc = (np.arange(Nx) + 0.5) / Nx
    yc = (np.arange(Ny) + 0.5) / Ny
    b0 = 1.0 / math.sqrt(4.0 * math.pi)
    rho = np.full((Nz, Ny, Nx), 25.0 / (36.0 * math.pi))
    pre = np.full((Nz, Ny, Nx), 5.0 / (12.0 * math.pi))
    vx = np.broadcast_to(-np.sin(2 * math.pi * yc)[None, :, None], (Nz, Ny, Nx))
    vy = np.broadcast_to(np.sin(2 * math.pi * xc)[None, None, :], (Nz, Ny, Nx))
    vel = [vx, vy, 0 * rho]
    bx = np.broadcast_to((-b0 * np.sin(2 * math.pi * yc))[None, :, None], (Nz, Ny, Nx + 1)).copy()
    by = np.broadcast_to((b0 * np.sin(4 * math.pi * xc))[None, None, :], (Nz, Ny + 1, Nx)).copy()
    bz = np.zeros((Nz + 1, Ny, Nx))
    u = np.zeros((5, Nz, Ny, Nx))
    u[0] = rho
    for a in range(3):
        u[2 + a] = rho * vel[a]
    u[1] = total_energy(rho, pre, vel, cell_centred_b(bx, by, bz), gamma)
    return u, (bx, by, bz)


def field_loop(N=(128, 64, 1), gamma=5.0 / 3.0, vel=(2.0, 1.0, 0.0), a0=1e-3, r0=0.3,
               rho0=1.0, p0=1.0, x_min=(-1.0, -0.5, -0.5), x_max=(1.0, 0.5, 0.5)):
    """Field-loop advection (Gardiner & Stone 2005, JCP 205, 509, sec. 5.1):
    a weak loop of field lines, B = curl(A_z e_z) with A_z = a0 (r0 - r) for
    r < r0 (r in the x-y plane), carried by the uniform flow `vel` through
    the periodic box at rho0, p0.  A_z sits on the z edges, so the face
    field has div B = 0 exactly and Bz = 0; the loop is uniform along z
    (N[2] = 1 gives the 2-D problem).  With vel = (2, 1, .) on the default
    2 x 1 box the loop returns to its start at t = 1."""
    Nx, Ny, Nz = N
    dx = (x_max[0] - x_min[0]) / Nx
    dy = (x_max[1] - x_min[1]) / Ny
    xf = x_min[0] + np.arange(Nx + 1) * dx
    yf = x_min[1] + np.arange(Ny + 1) * dy
    r = np.sqrt(xf[None, :] ** 2 + yf[:, None] ** 2)
    az = np.where(r < r0, a0 * (r0 - r), 0.0)                # (Ny+1, Nx+1)
    bx2 = (az[1:, :] - az[:-1, :]) / dy                       # (Ny, Nx+1)
    by2 = -(az[:, 1:] - az[:, :-1]) / dx                      # (Ny+1, Nx)
    bx = np.broadcast_to(bx2[None], (Nz, Ny, Nx + 1)).copy()
    by = np.broadcast_to(by2[None], (Nz, Ny + 1, Nx)).copy()
    bz = np.zeros((Nz + 1, Ny, Nx))
    rho = np.full((Nz, Ny, Nx), rho0)
    pre = np.full((Nz, Ny, Nx), p0)
    v = [np.full((Nz, Ny, Nx), vel[a]) for a in range(3)]
    u = np.zeros((5, Nz, Ny, Nx))
    u[0] = rho
    for a in range(3):
        u[2 + a] = rho * v[a]
    u[1] = total_energy(rho, pre, v, cell_centred_b(bx, by, bz), gamma)
    return u, (bx, by, bz)


def magnetic_energy(bfs):
    """Sum over cells of |Bcc|^2 / 2 (per unit cell volume)."""
    bcc = cell_centred_b(*bfs)
    return 0.5 * float(sum(np.sum(c * c) for c in bcc))


def brio_wu(N=(512, 1, 1), gamma=2.0, x_split=0.5, bx0=0.75):
    """Brio & Wu (1988) MHD shock tube along x on [0, 1]: (rho, p, By) =
    (1, 1, 1) for x < x_split and (0.125, 0.1, -1) beyond, Bx = 0.75,
    gamma = 2.  N = (n, 1, 1) is the 1-D problem."""
    Nx, Ny, Nz = N
    xc = (np.arange(Nx) + 0.5) / Nx
    left = xc < x_split
    rho = np.broadcast_to(np.where(left, 1.0, 0.125)[None, None, :], (Nz, Ny, Nx)).copy()
    pre = np.broadcast_to(np.where(left, 1.0, 0.1)[None, None, :], (Nz, Ny, Nx))
    bx = np.full((Nz, Ny, Nx + 1), bx0)
    by = np.broadcast_to(np.where(left, 1.0, -1.0)[None, None, :], (Nz, Ny + 1, Nx)).copy()
    bz = np.zeros((Nz + 1, Ny, Nx))
    u = np.zeros((5, Nz, Ny, Nx))
    u[0] = rho
    u[1] = total_energy(rho, pre, [0 * rho] * 3, cell_centred_b(bx, by, bz), gamma)
    return u, (bx, by, bz)


# ----------------------------------------------------------------------
# Per-block initial states (refined meshes have no global array): each
# generator evaluates its fields at the block's own cell and face centres
# and returns the reference's host AoS arrays of every block.
def _pack_block(n, rho, pre, vel, bfs, gamma):
    """(u_aos, bf_aos) of one block from cell fields (nz, ny, nx) and, for
    MHD, face fields bx (nz, ny, nx+1), by (nz, ny+1, nx), bz (nz+1, ny, nx)."""
    mhd = bfs is not None
    nvar = 8 if mhd else 5
    ub = np.zeros((nvar, n[2], n[1], n[0]))
    ub[0] = rho
    for a in range(3):
        ub[2 + a] = rho * vel[a]
    bcc = cell_centred_b(*bfs) if mhd else [0 * rho] * 3
    ub[1] = total_energy(rho, pre, vel, bcc, gamma)
    bf_aos = None
    if mhd:
        for a in range(3):
            ub[5 + a] = bcc[a]
        f = np.zeros((n[2] + 5, n[1] + 5, n[0] + 5, 3))
        f[2:n[2] + 2, 2:n[1] + 2, 2:n[0] + 3, 0] = bfs[0]
        f[2:n[2] + 2, 2:n[1] + 3, 2:n[0] + 2, 1] = bfs[1]
        f[2:n[2] + 3, 2:n[1] + 2, 2:n[0] + 2, 2] = bfs[2]
        bf_aos = f.reshape(-1)
    return np.ascontiguousarray(np.moveaxis(ub, 0, -1)).reshape(-1), bf_aos


def _block_axes(spec, b):
    """Cell-centre and face coordinates of block b along each axis."""
    xf0, dx0 = spec.geometry(b)
    n = spec.n_cell_block
    cc = [xf0[a] + (np.arange(n[a]) + 0.5) * dx0[a] for a in range(3)]
    fc = [xf0[a] + np.arange(n[a] + 1) * dx0[a] for a in range(3)]
    return cc, fc, dx0


def blast_blocks(spec, gamma=5.0 / 3.0, rho0=1.0, p0=0.1, p_in=10.0, r_in=0.1,
                 b0=(1.0 / math.sqrt(2.0), 1.0 / math.sqrt(2.0), 0.0), mhd=True):
    """MHD blast on every block of `spec`: uniform rho and oblique B, p = p_in
    inside radius r_in around the origin, p0 outside."""
    n = spec.n_cell_block
    out = []
    for b in range(spec.n_blocks):
        cc, _, _ = _block_axes(spec, b)
        r2 = cc[2][:, None, None] ** 2 + cc[1][None, :, None] ** 2 + cc[0][None, None, :] ** 2
        pre = np.where(r2 < r_in * r_in, p_in, p0)
        rho = np.full((n[2], n[1], n[0]), rho0)
        vel = [0 * rho] * 3
        bfs = None
        if mhd:
            bfs = (np.full((n[2], n[1], n[0] + 1), b0[0]), np.full((n[2], n[1] + 1, n[0]), b0[1]),
                   np.full((n[2] + 1, n[1], n[0]), b0[2]))
        out.append(_pack_block(n, rho, pre, vel, bfs, gamma))
    return out


def smooth_blocks(spec, seed=0, gamma=5.0 / 3.0, mhd=True, amp=0.1, mach=0.3, b_amp=0.04,
                  b0=(0.4, -0.3, 0.2)):
    """Smooth periodic state on every block of `spec`: a few seeded Fourier
    modes (integer wavenumbers on the box) in rho, p and v, and for MHD a
    uniform field plus the discrete curl of a seeded edge vector potential,
    so div B = 0 in every block at its own resolution."""
    rng = SplitMix64(seed)
    L = [spec.x_max[a] - spec.x_min[a] for a in range(3)]

    def modes(k):
        return [(tuple(int(rng.uniform() * 3) + 1 for _ in range(3)),
                 2 * math.pi * rng.uniform(), rng.uniform() - 0.5) for _ in range(k)]

    def field(ms, x, y, z):
        s = 0.0
        for kv, ph, am in ms:
            s = s + am * np.sin(2 * math.pi * (kv[0] * (x - spec.x_min[0]) / L[0]
                                                + kv[1] * (y - spec.x_min[1]) / L[1]
                                                + kv[2] * (z - spec.x_min[2]) / L[2]) + ph)
        return s

    m_rho, m_pre = modes(3), modes(3)
    m_v = [modes(3) for _ in range(3)]
    m_a = [modes(2) for _ in range(3)]
    n = spec.n_cell_block
    out = []
    for b in range(spec.n_blocks):
        cc, fc, dx = _block_axes(spec, b)
        Z, Y, X = np.meshgrid(cc[2], cc[1], cc[0], indexing="ij")
        rho = 1.0 + amp * field(m_rho, X, Y, Z)
        pre = 1.0 / gamma + amp / gamma * field(m_pre, X, Y, Z)
        vel = [mach * field(m_v[a], X, Y, Z) for a in range(3)]
        bfs = None
        if mhd:
            # A_a on the edges along a: cell centre along a, faces across
            Ax = b_amp * field(m_a[0], *_grid(cc[0], fc[1], fc[2]))
            Ay = b_amp * field(m_a[1], *_grid(fc[0], cc[1], fc[2]))
            Az = b_amp * field(m_a[2], *_grid(fc[0], fc[1], cc[2]))
            # arrays are (z, y, x); B = curl A on the faces
            bx = (Az[:, 1:, :] - Az[:, :-1, :]) / dx[1] - (Ay[1:, :, :] - Ay[:-1, :, :]) / dx[2]
            by = (Ax[1:, :, :] - Ax[:-1, :, :]) / dx[2] - (Az[:, :, 1:] - Az[:, :, :-1]) / dx[0]
            bz = (Ay[:, :, 1:] - Ay[:, :, :-1]) / dx[0] - (Ax[:, 1:, :] - Ax[:, :-1, :]) / dx[1]
            bfs = (bx + b0[0], by + b0[1], bz + b0[2])
        out.append(_pack_block(n, rho, pre, vel, bfs, gamma))
    return out


def _grid(xs, ys, zs):
    """(x, y, z) coordinate arrays of shape (len(zs), len(ys), len(xs))."""
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    return X, Y, Z


def wave_blocks(spec, amp=0.1, vel=(1.0, 0.0, 0.0), pre=1.0, gamma=5.0 / 3.0, mhd=False,
                b0=(0.0, 0.0, 0.0)):
    """Density wave rho = 1 + amp sin(2 pi (x - x_min) / L_x) in a uniform flow
    `vel` at uniform pressure (and uniform B for MHD): the exact solution is
    the translated profile.  Cells hold the exact cell average of rho."""
    n = spec.n_cell_block
    L = spec.x_max[0] - spec.x_min[0]
    out = []
    for b in range(spec.n_blocks):
        _, fc, dx = _block_axes(spec, b)
        xl = (fc[0][:-1] - spec.x_min[0]) / L
        xr = (fc[0][1:] - spec.x_min[0]) / L
        avg = 1.0 + amp * (np.cos(2 * math.pi * xl) - np.cos(2 * math.pi * xr)) / (
            2 * math.pi * (xr - xl))
        rho = np.broadcast_to(avg[None, None, :], (n[2], n[1], n[0])).copy()
        p = np.full_like(rho, pre)
        v = [np.full_like(rho, vel[a]) for a in range(3)]
        bfs = None
        if mhd:
            bfs = (np.full((n[2], n[1], n[0] + 1), b0[0]), np.full((n[2], n[1] + 1, n[0]), b0[1]),
                   np.full((n[2] + 1, n[1], n[0]), b0[2]))
        out.append(_pack_block(n, rho, p, v, bfs, gamma))
    return out


def wave_exact(spec, b, t, amp=0.1, vel=1.0):
    """Exact cell averages of the translated density wave on block b at time t."""
    n = spec.n_cell_block
    L = spec.x_max[0] - spec.x_min[0]
    _, fc, _ = _block_axes(spec, b)
    xl = (fc[0][:-1] - spec.x_min[0] - vel * t) / L
    xr = (fc[0][1:] - spec.x_min[0] - vel * t) / L
    avg = 1.0 + amp * (np.cos(2 * math.pi * xl) - np.cos(2 * math.pi * xr)) / (
        2 * math.pi * (xr - xl))
    return np.broadcast_to(avg[None, None, :], (n[2], n[1], n[0]))
