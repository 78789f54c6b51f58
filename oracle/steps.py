"""Whole implicit steps restated with the C oracle.

TEST INFRASTRUCTURE ONLY: used by tests/ and by the CPU-baseline leg of
bench.py as the checker of the GPU steppers (mushy-layer_amd/steps.py).

DarcyBrinkmanStep is one Darcy-Brinkman mushy-layer step on a single-box
level, in the reference's order (AMRLevelMushyLayer::advance,
srcSubcycle/AMRLevelMushyLayer.cpp:451-787):
  1. enthalpy variables at t^n; HC face coefficients from the porosity at
     t^n (srcSubcycle/AMRLevelMushyLayerInit.cpp:844-953);
  2. the advective source of (H, Theta) (computeScalarAdvectiveSrcHC,
     :2074-2104): (theta, Theta_l)^n traced to the faces with the diffusive
     source of (H, Theta)^n, advected by the face average of the velocity
     u^n (the reference's MAC-projected advection velocity is not built);
     then the backward-Euler FAS solve of (H, Theta) with that source
     (multiCompAdvectDiffuse, :1645-1831);
  3. enthalpy variables at t^{n+1} (updateEnthalpyVariables, :729);
  4. cell-centred velocity (computeCCvelocity,
     srcSubcycle/AMRLevelMushyLayerVel.cpp:673-917):
       U* source (buoyancy - porosity * grad P^{n-1/2}) at the half time;
       U* coefficients a = 1, b = -Pr, c = (Pr/Da) chi^n / Pi(chi^n)
       (defineUstarMultigrid, srcSubcycle/AMRLevelMushyLayerInit.cpp
       :579-747, coefficients at the old time);
       one backward-Euler solve per velocity component (:1326-1403) from
       the previous U*;
       u = U* + dt chi grad P^{n-1/2}; u = 0 where chi^{n+1} <= 0.05;
       projection: div(chi_f grad pi) = div(u)/dt from pi = 0
       (Projector::levelProject, srcSubcycle/Projector.cpp:1340-1486;
       solveMGlevel zeroes pi, :3164-3173); u -= dt chi grad pi;
       u = 0 where chi^{n+1} <= 0.05; the pressure kept for the next step.
"""
import numpy as np

from oracle import mlref as O

# multigrid options of the velocity solves: U* (VelocityMultigrid,
# srcSubcycle/MushyLayerSubcycleUtils.cpp:654-659: smoothing 2 down, up and
# before the bottom, 10 iterations, tolerance 1e-10; BiCGStab bottom with
# Chombo's default 80 iterations, srcSubcycle/AMRLevelMushyLayerInit.cpp:739)
# and the projection (srcSubcycle/Projector.cpp:53-63: down 2, up 4, 4 before
# the bottom, tolerance and hang 1e-15, 20 iterations; BiCGStab bottom of 20
# iterations, :3004-3023)
U_OPTS = dict(nu1=2, nu2=2, bottom=2, max_iter=10, eps=1e-10, hang=1e-10, norm_thresh=1e-10,
              bottom_solver=O.BOTTOM_BICGSTAB, bottom_max_iter=80)
P_OPTS = dict(nu1=2, nu2=4, bottom=4, max_iter=20, eps=1e-15, hang=1e-15, norm_thresh=1e-15,
              bottom_solver=O.BOTTOM_BICGSTAB, bottom_max_iter=20)
# solidPorosity = ccVelPorosityLimit (srcSubcycle/MushyLayerSubcycleUtils.cpp:223-233)
POROSITY_LIMIT = 0.05
# hang rule off: a solve stops on eps, norm_thresh or max_iter only
# (|r| > (1 - hang)*|r_prev| never holds for hang = -inf)
NO_HANG = float("-inf")


def flow_coeffs(Pr, Da, Ra_C, Ra_T=0.0):
    """Diffusive time scale, advective velocity scale
    (src/MushyLayerParams.cpp:736-753): viscosity Pr, buoyancy Pr Ra_T and
    Pr Ra_C, Darcy coefficient Pr/Da."""
    return dict(viscosity=Pr, buoyT=Pr * Ra_T, buoyS=Pr * Ra_C, darcy=Pr / Da)


def darcy_ccoef(chi, coeff, law="kozeny_carman"):
    """c = (chi / Pi) * coeff on a valid-cell porosity array (any shape), by
    the oracle's mlr_darcy_ccoef (pow(x, 3) and pow(x, 2) as once-rounded
    cube and square)."""
    chi = np.asarray(chi, dtype=np.float64)
    flat = chi.reshape(-1)
    g = O.make_grid([flat.size, 1], 1.0, [0, 0, 0])
    cell = np.zeros(O.cell_shape(g, 1))
    O.valid(g, cell)[0, 0, 0, :] = flat
    return O.valid(g, O.darcy_ccoef(g, cell, coeff, law))[0, 0, 0, :].reshape(chi.shape)


class HCStep:
    """The (H, Theta) backward-Euler step alone."""

    def __init__(self, case, max_iter=100, hang=NO_HANG):
        self.n = list(case["n"])
        per = list(case["periodic"]) + [0] * (3 - len(case["periodic"]))
        self.g = O.make_grid(self.n, case["dx"], per)
        self.p = O.default_params(**case["params"])
        self.bc = O.make_bc(case["bc"], apply=case.get("apply_bc", 1))
        self.dt = case["dt"]
        self.phi = case["phi"].copy()
        self.opts = O.default_opts(max_iter=max_iter, hang=hang)
        self.steps = 0
        self.hc_stats = []

    def hc_src(self, vold, b):
        """Source of the (H, Theta) step (none without flow)."""
        return None

    def hc_step(self):
        g, p = self.g, self.p
        vold = O.enthalpy_vars(g, p, self.phi)
        chi = np.ascontiguousarray(vold[3:4])
        b = O.hc_bcoef(g, p, chi)
        src = self.hc_src(vold, b)
        old = self.phi.copy()
        r = O.hc_be_step(g, p, self.bc, None, b, self.dt, self.phi, old, src=src, opts=self.opts)
        self.hc_stats.append((r.exit_status, r.iterations))
        return vold

    def step(self):
        self.hc_step()
        self.steps += 1


class DarcyBrinkmanStep(HCStep):
    """The Darcy-Brinkman step (module docstring) on the oracle."""

    def __init__(self, case, flow, max_iter=100, hang=NO_HANG, advect=True):
        super().__init__(case, max_iter, hang)
        self.flow = flow
        self.advect = advect
        self.abc = O.make_advect_bc(hc_bc=self.bc)
        dim = len(self.n)
        g = self.g
        self.bcu = O.make_bc(flow["bc_u"])
        self.bcp = O.make_bc(flow["bc_p"])
        self.coef = flow_coeffs(flow["Pr"], flow["Da"], flow["Ra_C"], flow.get("Ra_T", 0.0))
        self.law = flow.get("permeability", "kozeny_carman")
        shp = O.cell_shape(g, 1)
        self.u = [np.zeros(shp) for _ in range(dim)]
        self.ustar = [np.zeros(shp) for _ in range(dim)]
        self.pres = np.zeros(shp)
        self.b_u = [np.full(O.face_shape(g, d, 1), -self.coef["viscosity"]) for d in range(dim)]
        self.uopts = O.default_opts(**U_OPTS)
        self.popts = O.default_opts(**P_OPTS)
        self.vstats = []

    def hc_src(self, vold, b):
        """The advective source (step 2): diffusive source of (H, Theta)^n,
        face velocities of u^n, the Godunov fluxes' divergence."""
        if not self.advect:
            return None
        g = self.g
        S = O.hc_apply_op(g, self.p, self.bc, 0.0, None, 1.0, b, self.phi.copy())
        uf = O.face_velocity(g, self.u)
        src, _, _ = O.advect_src(g, self.abc, self.dt, 1.0, vold, uf, S=S)
        return src

    def step(self):
        g, dt = self.g, self.dt
        vold = self.hc_step()
        vnew = O.enthalpy_vars(g, self.p, self.phi)
        co = self.coef
        # U* coefficients from the porosity at t^n
        c = O.darcy_ccoef(g, np.ascontiguousarray(vold[3:4]), co["darcy"], self.law)
        src = O.ustar_src(g, co["buoyT"], co["buoyS"], 1, vold, vnew, self.pres)
        vs = []
        for d in range(g.dim):
            rhs = np.zeros_like(self.u[d])
            O.valid(g, rhs)[...] = O.valid(g, self.u[d]) * 1.0 + dt * O.valid(g, src[d])
            st = O.lin_mg_solve(g, 1, self.bcu, 1.0, None, -dt, c, self.b_u, self.ustar[d], rhs, opts=self.uopts)
            vs.append((st.exit_status, st.iterations))
        u, prhs = O.cc_project_rhs(g, dt, POROSITY_LIMIT, 1, vold, vnew, self.pres, self.ustar)
        bp = O.pressure_faces(g, vold, vnew)
        self.pres = np.zeros(O.cell_shape(g, 1))
        st = O.lin_mg_solve(g, 1, self.bcp, 0.0, None, 1.0, None, bp, self.pres, prhs, opts=self.popts)
        vs.append((st.exit_status, st.iterations))
        O.cc_correct(g, dt, POROSITY_LIMIT, 1, vold, vnew, self.pres, u)
        self.u = u
        self.vstats.append(vs)
        self.steps += 1
