"""Seeded synthetic inputs of the benchmark and parity workloads.

Every generator returns host numpy arrays in the FAB layout of the C ABI
(``cell_shape`` / ``face_shape``) plus the parameter / BC descriptions, so
the CPU oracle and the GPU library receive bit-identical inputs.

Workloads (named by what they are):
  * ``diffusive_64``  — 64x64 diffusive solidification, unit square,
    periodic in x, cold bottom theta = -1, warm top theta = 1.2, St = 5,
    CR = 2, Le = 200, H = St + 1.2, Theta = -1 + 1e-3 U(-1, 1), dt = 1e-3,
    10 backward-Euler steps.
  * ``mushy_256`` — 256x256 mushy layer with porosity from a smooth seeded
    Gaussian random field (correlation length 0.1) in [0.05, 1], St = 5,
    CR = 2, Kozeny-Carman permeability, Da = 1e-3, Ra_C = 200.
  * ``heat_flux`` — 128x128 diffusive solidification driven by a bottom
    heat-flux profile (the adjoint workload): dt = 5e-4, 50 steps.
  * ``batch_instances`` — independent mushy layers for the batched form:
    grid sizes drawn by seed in ratio 8:4:2:1 from 64^2, 128^2, 256^2,
    512x256; St in [1, 10], CR in [1, 20], Ra_C in [0, 400], Da in
    [1e-4, 1e-2], uniformly.
"""
import numpy as np

DIRICHLET, NEUMANN = 0, 1


def cell_shape(n, ncomp):
    if len(n) == 3:
        return (ncomp, n[2] + 2, n[1] + 2, n[0] + 2)
    return (ncomp, 1, n[1] + 2, n[0] + 2)


def face_shape(n, d, ncomp):
    m = list(n) + [1] * (3 - len(n))
    m[d] += 1
    return (ncomp, m[2], m[1], m[0])


def valid(n, x):
    if len(n) == 3:
        return x[:, 1:-1, 1:-1, 1:-1]
    return x[:, :, 1:-1, 1:-1]


def closure_point(H, C, St, CR, pc=1e-5, c=1.0, te=0.0, Te=0.0):
    """Scalar phase-diagram closure used to derive boundary values
    (computeEnthalpyVars, src/phaseDiagram.cpp:180-223)."""
    hs = 0.0 if C < -CR else c * ((-C - CR) / pc)
    hs = max(hs, 0.0) + te * c
    he = (CR + C) / (CR + Te) * (St + te * (1 - c)) + c * te
    hl = St - C + te + Te
    if H <= hs:
        return H / c, 0.0
    if H <= he:
        return te, Te
    if H < hl:
        vc = max(C, -CR)
        a = CR * (c - 1) + St * (pc - 1)
        b = CR * (1 - 2 * c) + H * (1 - pc) - vc * (c - 1) - pc * St
        q = (CR + vc) * c + pc * H
        chi = (-b - np.sqrt(b * b - 4 * a * q)) / (2 * a)
        cl = (C + CR * (1 - chi)) / (chi + pc * (1 - chi))
        return -cl, cl
    return H - St, C


def diffusive(n=(64, 64), seed=1, St=5.0, CR=2.0, lewis=200.0, theta_lo=-1.0, theta_top=1.2,
              dt=1e-3, steps=10, noise=1e-3):
    """Diffusive solidification (no flow): HC state, derived-variable BCs."""
    n = list(n)
    dim = len(n)
    H_top = St + theta_top
    C0 = -1.0
    # boundary (theta, Theta_l) from the boundary (H, Theta) by the closure
    # (MushyLayerParams::computeDerivedBCs, src/MushyLayerParams.cpp:130-170)
    t_lo, cl_lo = closure_point(theta_lo, C0, St, CR)     # H_lo = theta_lo (solid)
    t_hi, cl_hi = closure_point(H_top, C0, St, CR)
    vd = dim - 1
    bc = {(vd, 0, 0): (DIRICHLET, t_lo), (vd, 1, 0): (DIRICHLET, t_hi),
          (vd, 0, 1): (DIRICHLET, cl_lo), (vd, 1, 1): (DIRICHLET, cl_hi)}
    periodic = [1] * (dim - 1) + [0]
    rng = np.random.default_rng(seed)
    phi = np.zeros(cell_shape(n, 2))
    v = valid(n, phi)
    v[0] = H_top
    v[1] = C0 + noise * rng.uniform(-1.0, 1.0, size=v[1].shape)
    return dict(n=n, dim=dim, dx=1.0 / n[0], periodic=periodic, bc=bc, phi=phi, dt=dt, steps=steps,
                params=dict(stefan=St, compositionRatio=CR, lewis=lewis))


def gaussian_field(n, corr, seed, lo=0.05, hi=1.0):
    """Periodic smooth Gaussian random field affinely mapped to [lo, hi]:
    white noise convolved with a separable periodic Gaussian of standard
    deviation corr/dx cells, truncated at 4 sigma."""
    rng = np.random.default_rng(seed)
    shape = tuple(reversed(n))
    f = rng.standard_normal(shape)
    sig = corr * n[0]
    r = int(np.ceil(4 * sig))
    k = np.exp(-0.5 * (np.arange(-r, r + 1) / sig) ** 2)
    k /= k.sum()
    for ax in range(len(shape)):
        acc = np.zeros_like(f)
        for o, w in zip(range(-r, r + 1), k):
            acc += w * np.roll(f, o, axis=ax)
        f = acc
    f = (f - f.min()) / (f.max() - f.min())
    return lo + (hi - lo) * f


def mushy(n=(256, 256), seed=2, St=5.0, CR=2.0, lewis=200.0, corr=0.1, theta_bottom=0.2, theta_top=1.0,
          theta_wall=0.0, dt=1e-3, steps=20):
    """Mushy layer whose porosity is a seeded smooth random field.

    The temperature starts linear in z from theta_bottom to theta_top with
    the liquid on the liquidus (Theta_l = -theta); each cell's (H, Theta)
    follows from its porosity by inverting the mush relations
    Theta = Theta_l (chi + pc(1-chi)) - CR(1-chi), H = chi St + theta (c = 1).
    The bottom wall is held at theta_wall (the eutectic by default, colder
    than the initial profile, so a solidification front starts there); the
    top wall keeps theta_top.  Periodic sideways."""
    n = list(n)
    dim = len(n)
    pc = 1e-5
    chi = gaussian_field(n, corr, seed)
    z = (np.arange(n[-1]) + 0.5) / n[-1]
    shape = [n[-1]] + [1] * (dim - 1)
    theta = (theta_bottom + (theta_top - theta_bottom) * z).reshape(shape)
    theta_l = -theta
    Theta = theta_l * (chi + pc * (1 - chi)) - CR * (1 - chi)
    H = chi * St + theta
    phi = np.zeros(cell_shape(n, 2))
    v = valid(n, phi)
    v[0] = H.reshape(v[0].shape)
    v[1] = Theta.reshape(v[1].shape)
    vd = dim - 1
    bc = {(vd, 0, 0): (DIRICHLET, theta_wall), (vd, 1, 0): (DIRICHLET, theta_top),
          (vd, 0, 1): (DIRICHLET, -theta_wall), (vd, 1, 1): (DIRICHLET, -theta_top)}
    periodic = [1] * (dim - 1) + [0]
    return dict(n=n, dim=dim, dx=1.0 / n[0], periodic=periodic, bc=bc, phi=phi, dt=dt, steps=steps,
                params=dict(stefan=St, compositionRatio=CR, lewis=lewis), chi=chi)


def face_shape(n, d, ncomp):
    m = list(n) + [1] * (3 - len(n))
    m[d] += 1
    return (ncomp, m[2], m[1], m[0])


def geometric_faces(n, periodic, chi):
    """Face porosity sqrt(|chi_lo chi_hi|) (src/CellToEdge2F.ChF:20-46) from a
    valid-cell field (shape reversed(n)); wall faces use the interior cell."""
    dim = len(n)
    out = []
    for d in range(dim):
        ax = dim - 1 - d
        if periodic[d]:
            lo = np.concatenate([np.take(chi, [-1], axis=ax), chi], axis=ax)
            hi = np.concatenate([chi, np.take(chi, [0], axis=ax)], axis=ax)
        else:
            lo = np.concatenate([np.take(chi, [0], axis=ax), chi], axis=ax)
            hi = np.concatenate([chi, np.take(chi, [-1], axis=ax)], axis=ax)
        f = np.sqrt(np.abs(lo * hi))
        out.append(np.ascontiguousarray(f.reshape(face_shape(n, d, 1))))
    return out


def darcy_brinkman(n=(256, 256), seed=2, Pr=10.0, Da=1e-3, corr=0.1, dt=1e-3, steps=20, Ra_C=200.0):
    """Coefficients of the Darcy-Brinkman U* solve and of the projection on a
    porosity field from the same seeded random field as `mushy`:
      U*:         a = 1, b = -Pr, c = (Pr/Da) chi/Pi, Pi = chi^3/(1-chi)^2
                  (Kozeny-Carman, src/MushyLayerParams.cpp:1046-1049;
                  srcSubcycle/AMRLevelMushyLayerInit.cpp:665-689), no-slip
                  (homogeneous Dirichlet) walls in the last direction;
      projection: b = -chi_face, homogeneous Neumann walls.
    Periodic sideways.  `u0` is a smooth seeded initial velocity component
    and `src` a buoyancy-like source Ra_C*(chi - mean chi)."""
    n = list(n)
    dim = len(n)
    chi = gaussian_field(n, corr, seed)
    with np.errstate(divide="ignore"):
        perm = chi ** 3 / (1.0 - chi) ** 2
        cval = np.where(np.isinf(perm), 0.0, chi / np.where(perm == 0, 1.0, perm)) * (Pr / Da)
    c = np.zeros(cell_shape(n, 1))
    valid(n, c)[0] = cval.reshape(valid(n, c)[0].shape)
    b_u = [np.full(face_shape(n, d, 1), -Pr) for d in range(dim)]
    periodic = [1] * (dim - 1) + [0]
    b_p = [-f for f in geometric_faces(n, periodic, chi)]
    u0 = np.zeros(cell_shape(n, 1))
    valid(n, u0)[0] = (0.1 * gaussian_field(n, corr, seed + 2, -1.0, 1.0)).reshape(valid(n, u0)[0].shape)
    src = np.zeros(cell_shape(n, 1))
    valid(n, src)[0] = (Ra_C * (chi - chi.mean())).reshape(valid(n, src)[0].shape)
    vd = dim - 1
    bc_u = {(vd, 0, 0): (DIRICHLET, 0.0), (vd, 1, 0): (DIRICHLET, 0.0)}
    bc_p = {(vd, 0, 0): (NEUMANN, 0.0), (vd, 1, 0): (NEUMANN, 0.0)}
    return dict(n=n, dim=dim, dx=1.0 / n[0], periodic=periodic, chi=chi, c=c, b_u=b_u, b_p=b_p, u0=u0, src=src,
                bc_u=bc_u, bc_p=bc_p, dt=dt, steps=steps, Pr=Pr, Da=Da)


BATCH_SIZES = ((64, 64), (128, 128), (256, 256), (512, 256))


def batch_instances(count=512, seed=5, sizes=BATCH_SIZES, weights=(8, 4, 2, 1), corr=0.1, dt=1e-3):
    """Independent mushy-layer instances: each a `mushy` case of its own seed,
    grid size (drawn with probabilities proportional to `weights`) and
    parameters St ~ U[1, 10], CR ~ U[1, 20]; Ra_C ~ U[0, 400] and
    Da ~ U[1e-4, 1e-2] are drawn too and stored with the case (they enter
    only the velocity solves)."""
    rng = np.random.default_rng(seed)
    w = np.asarray(weights, dtype=float)
    pick = rng.choice(len(sizes), size=count, p=w / w.sum())
    out = []
    for i, k in enumerate(pick):
        St = float(rng.uniform(1.0, 10.0))
        CR = float(rng.uniform(1.0, 20.0))
        Ra_C = float(rng.uniform(0.0, 400.0))
        Da = float(rng.uniform(1e-4, 1e-2))
        c = mushy(n=sizes[k], seed=seed * 100003 + i, St=St, CR=CR, corr=corr, dt=dt)
        c["Ra_C"], c["Da"] = Ra_C, Da
        out.append(c)
    return out


def heat_flux(n=(128, 128), seed=4, St=5.0, CR=2.0, lewis=200.0, theta_top=1.2, dt=5e-4, steps=50, noise=1e-3):
    """Diffusive solidification cooled through the bottom by a heat-flux
    profile: the `diffusive` liquid start and top BC, with the bottom theta
    BC Neumann (its value per face is the flux profile, applied as the
    equivalent source of the step) and the bottom Theta_l BC Neumann 0.
    Periodic in x."""
    c = diffusive(n=n, seed=seed, St=St, CR=CR, lewis=lewis, theta_top=theta_top, dt=dt, steps=steps, noise=noise)
    vd = len(n) - 1
    c["bc"][(vd, 0, 0)] = (NEUMANN, 0.0)
    c["bc"][(vd, 0, 1)] = (NEUMANN, 0.0)
    return c


def flux_profile(nx, F0=10.0, modes=((1, 0.3, 0.0),)):
    """F(x) = F0 (1 + sum_k a_k sin(2 pi k x + p_k)) at the bottom face
    centres x = (i + 1/2)/nx."""
    x = (np.arange(nx) + 0.5) / nx
    f = np.ones(nx)
    for k, a, ph in modes:
        f = f + a * np.sin(2 * np.pi * k * x + ph)
    return F0 * f


def flow(dim=2, Pr=10.0, Da=1e-3, Ra_C=200.0, Ra_T=0.0, permeability="kozeny_carman"):
    """Flow parameters of the Darcy-Brinkman step on a `mushy` case: Prandtl
    number (the reference's Darcy-Brinkman example uses Pr = 10,
    examples/mushylayer_darcybrinkman/inputs), Darcy number, solutal and
    thermal Rayleigh numbers, the permeability law; no-slip walls in the last
    direction for the velocity, homogeneous Neumann for the pressure, periodic
    sideways (as the case).  The velocity and pressure start at zero."""
    vd = dim - 1
    return dict(Pr=Pr, Da=Da, Ra_C=Ra_C, Ra_T=Ra_T, permeability=permeability,
                bc_u={(vd, 0, 0): (DIRICHLET, 0.0), (vd, 1, 0): (DIRICHLET, 0.0)},
                bc_p={(vd, 0, 0): (NEUMANN, 0.0), (vd, 1, 0): (NEUMANN, 0.0)})
