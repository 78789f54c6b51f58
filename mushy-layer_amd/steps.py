"""Whole implicit steps on the GPU, composed from the C-ABI entry points.

HCStepper: one backward-Euler step of (H, Theta) — enthalpy variables at
t^n -> HC face coefficients from the porosity at t^n -> FAS solve
(multiCompAdvectDiffuse, srcSubcycle/AMRLevelMushyLayer.cpp:1645-1831).

DarcyBrinkmanStepper: one Darcy-Brinkman mushy-layer step in the order of
AMRLevelMushyLayer::advance (srcSubcycle/AMRLevelMushyLayer.cpp:451-787):
the (H, Theta) step, the enthalpy variables at t^{n+1}
(updateEnthalpyVariables), then the cell-centred velocity
(computeCCvelocity, srcSubcycle/AMRLevelMushyLayerVel.cpp:673-917):
  U* source = buoyancy - porosity * grad P^{n-1/2} at the half time
    (computeUstarSrc, :1724-1853);
  U* coefficients a = 1, b = -Pr, c = (Pr/Da) chi^n / Pi(chi^n)
    (defineUstarMultigrid, srcSubcycle/AMRLevelMushyLayerInit.cpp:579-747);
  one backward-Euler solve per velocity component from the previous U*
    (:1326-1403), run concurrently;
  u = U* + dt chi grad P^{n-1/2}, zero where chi^{n+1} <= 0.05, projection
    rhs div(u)/dt, faces -(porosity at the half time), solve from pi = 0
    (Projector::levelProject, srcSubcycle/Projector.cpp:1340-1486,
    solveMGlevel :3164-3173), u -= dt chi grad pi, zero where solid.
Before the (H, Theta) solve the stepper forms its advective source
(computeScalarAdvectiveSrcHC, srcSubcycle/AMRLevelMushyLayer.cpp:2074-2104):
the diffusive source of (H, Theta)^n (computeScalDiffusion), the face
average of the velocity u^n as the advection velocity (the reference
advects with its MAC-projected predictor, which is not built here), and
the Godunov face values of (theta, Theta_l)^n.  Everything stays on the
device.
"""
import numpy as np

from . import (DeviceArray, HCOp, VCOp, default_opts, face_shape, make_advect_params, make_bc, make_flow_params,
               make_params, valid, hc_db_advect_src_many, hc_db_coefs_many, hc_set_coefs_many, vc_db_correct_many,
               vc_db_project_many, vc_db_ustar_many, vc_set_coefs_many)

# multigrid options of the velocity solves: U* (VelocityMultigrid,
# srcSubcycle/MushyLayerSubcycleUtils.cpp:654-659: smoothing 2 down, up and
# before the bottom, 10 iterations, tolerance 1e-10; BiCGStab bottom
# s_botSolverUStar with Chombo's default 80 iterations,
# srcSubcycle/AMRLevelMushyLayerInit.cpp:739) and the projection
# (srcSubcycle/Projector.cpp:53-63: down 2, up 4, 4 before the bottom,
# tolerance and hang 1e-15, 20 iterations; BiCGStab bottom of 20 iterations,
# :3004-3023)
U_OPTS = dict(nu1=2, nu2=2, bottom=2, max_iter=10, eps=1e-10, hang=1e-10, norm_thresh=1e-10,
              bottom_solver=1, bottom_max_iter=80)
P_OPTS = dict(nu1=2, nu2=4, bottom=4, max_iter=20, eps=1e-15, hang=1e-15, norm_thresh=1e-15,
              bottom_solver=1, bottom_max_iter=20)
POROSITY_LIMIT = 0.05   # solidPorosity (srcSubcycle/MushyLayerSubcycleUtils.cpp:223-233)
# hang rule off: a solve stops on eps, norm_thresh or max_iter only
# (|r| > (1 - hang)*|r_prev| never holds for hang = -inf)
NO_HANG = float("-inf")


def flow_coeffs(Pr, Da, Ra_C, Ra_T=0.0):
    """Diffusive time scale, advective velocity scale
    (src/MushyLayerParams.cpp:736-753)."""
    return dict(viscosity=Pr, buoyT=Pr * Ra_T, buoyS=Pr * Ra_C, darcy=Pr / Da)


class HCStepper:
    """Device-resident (H, Theta) state and its implicit step."""

    def __init__(self, ml, ctx, case, opts):
        self.ml, self.ctx, self.case, self.opts = ml, ctx, case, opts
        n = case["n"]
        self.n = list(n)
        self.dt = case["dt"]
        self.op = HCOp(ctx, n, case["dx"], case["periodic"], make_params(**case["params"]),
                       make_bc(case["bc"], apply=case.get("apply_bc", 1)))
        shp = case["phi"].shape
        self.init = DeviceArray(ctx, shp, host=case["phi"])
        self.phi = DeviceArray(ctx, shp, host=case["phi"])
        self.old = DeviceArray(ctx, shp)
        self.vold = DeviceArray(ctx, (4,) + shp[1:])
        self.vnew = DeviceArray(ctx, (4,) + shp[1:])
        self.b = [DeviceArray(ctx, face_shape(n, d, 2)) for d in range(len(n))]
        self.stats = []

    def copy(self, dst, src):
        self.ctx.check(self.ctx.lib.ml_buf_copy(self.ctx.h, dst.ptr, src.ptr, src.nbytes), "ml_buf_copy")

    def reset(self):
        self.copy(self.phi, self.init)
        self.stats = []

    def set_flags(self, flags):
        self.opts.flags = flags

    def hc_src(self):
        """Source of the (H, Theta) step (none without flow)."""
        return None

    def hc_step(self):
        op = self.op
        op.enthalpy_vars(self.phi, self.vold)
        op.assemble_bcoef(self.vold.comp_ptr(3), self.b)
        op.set_coefs(None, self.b, 1.0, -self.dt)
        src = self.hc_src()
        self.copy(self.old, self.phi)
        st = op.solve_be(self.phi, self.old, self.dt, src=src, opts=self.opts)
        self.stats.append((st.exit_status, st.iterations, st.final_residual))

    def step(self):
        self.hc_step()

    # ---- checkpoint / restart (the reference's Chombo HDF5 layout,
    # srcSubcycle/AMRLevelMushyLayerIO.cpp:131-420)
    def checkpoint_fields(self):
        """Scalar and vector fields of a checkpoint, host copies in the
        package's layout: Enthalpy and Bulk concentration."""
        phi = self.phi.get()
        return {"Enthalpy": phi[0:1], "Bulk concentration": phi[1:2]}, {}, {}

    def save_checkpoint(self, path, iteration, time):
        """A single-level checkpoint of the current state (writeCheckpointHeader
        / writeCheckpointLevel's layout; needs libmlio.so)."""
        from . import chombo_io
        self.ctx.sync()
        scalars, vectors, faces = self.checkpoint_fields()
        c = self.case
        chombo_io.write_checkpoint(path, self.n, c["dx"], self.dt, time, iteration, c["periodic"], scalars, vectors,
                                   faces)

    def load_checkpoint(self, path):
        """Restore the state from a checkpoint of the same grid
        (readCheckpointLevel); returns the file's (iteration, time)."""
        from . import chombo_io
        lev = chombo_io.read_level(path, 0)
        h = lev["header"]
        if [h["hi"][d] - h["lo"][d] + 1 for d in range(len(self.n))] != list(self.n):
            raise ValueError("checkpoint grid %s differs from %s" % (h, self.n))
        phi = np.concatenate([chombo_io.domain_array(lev, "Enthalpy"),
                              chombo_io.domain_array(lev, "Bulk concentration")], axis=0)
        self.phi.set(np.ascontiguousarray(phi))
        self.restore_extra(lev)
        self.ctx.sync()
        return lev["root"]["iteration"], lev["root"]["time"]

    def restore_extra(self, lev):
        pass

    def porosity(self):
        """Host copy of the porosity of the current state (valid cells)."""
        self.op.enthalpy_vars(self.phi, self.vnew)
        return valid(self.n, self.vnew.get())[3]

    def close(self):
        self.op.close()


class DarcyBrinkmanStepper(HCStepper):
    """The Darcy-Brinkman step (module docstring).  `flow`: Pr, Da, Ra_C
    (Ra_T), the velocity and pressure BC specs, the permeability law."""

    def __init__(self, ml, ctx, case, opts, flow, advect=True):
        super().__init__(ml, ctx, case, opts)
        n = self.n
        dim = len(n)
        self.flow = flow
        self.advect = advect
        self.coef = flow_coeffs(flow["Pr"], flow["Da"], flow["Ra_C"], flow.get("Ra_T", 0.0))
        self.law = flow.get("permeability", "kozeny_carman")
        self.fp = make_flow_params(self.coef["buoyT"], self.coef["buoyS"], POROSITY_LIMIT, 1)
        self.uop = VCOp(ctx, n, case["dx"], case["periodic"], make_bc(flow["bc_u"]))
        self.pop = VCOp(ctx, n, case["dx"], case["periodic"], make_bc(flow["bc_p"]))
        shp1 = (1,) + case["phi"].shape[1:]
        self.u = [DeviceArray(ctx, shp1) for _ in range(dim)]
        self.ustar = [DeviceArray(ctx, shp1) for _ in range(dim)]
        self.uold = [DeviceArray(ctx, shp1) for _ in range(dim)]
        self.src = [DeviceArray(ctx, shp1) for _ in range(dim)]
        self.c = DeviceArray(ctx, shp1)
        self.bu = [DeviceArray(ctx, face_shape(n, d, 1), host=np.full(face_shape(n, d, 1), -self.coef["viscosity"]))
                   for d in range(dim)]
        self.bp = [DeviceArray(ctx, face_shape(n, d, 1)) for d in range(dim)]
        self.prhs = DeviceArray(ctx, shp1)
        self.pi = DeviceArray(ctx, shp1)
        shp2 = (2,) + case["phi"].shape[1:]
        self.dsrc = DeviceArray(ctx, shp2)
        self.hsrc = DeviceArray(ctx, shp2)
        self.uf = [DeviceArray(ctx, face_shape(n, d, 1)) for d in range(dim)]
        self.ap = make_advect_params(self.op.bc, 1.0)
        self.uopts = default_opts(**U_OPTS)
        self.popts = default_opts(**P_OPTS)
        self.vstats = []
        # the U* components share every launch of one batched solve where the
        # batched engine covers the grid (2D, levels up to 512^2 cells);
        # otherwise they run concurrently on two streams (same results)
        self.ub = None
        try:
            from . import VCBatch
            self.ub = VCBatch(ctx, [self.uop], per_op=dim)
        except RuntimeError:
            self.ub = None

    def _zero(self, x):
        self.ctx.check(self.ctx.lib.ml_vec_set(self.ctx.h, x.nbytes // 8, x.ptr, 0.0), "ml_vec_set")

    def reset(self):
        super().reset()
        for x in self.u + self.ustar + [self.pi]:
            self._zero(x)
        self.vstats = []

    def set_flags(self, flags):
        super().set_flags(flags)
        self.uopts.flags = flags
        self.popts.flags = flags

    def hc_src(self):
        """The advective source: diffusive source of (H, Theta)^n, face
        velocities of u^n, src = -div(Godunov fluxes of (theta, Theta_l)^n)."""
        if not self.advect:
            return None
        op = self.op
        op.diffusive_src(self.phi, self.dsrc)
        self.uop.face_velocity(self.u, self.uf)
        self._zero(self.hsrc)
        op.advect_src(self.ap, self.dt, self.vold, self.uf, self.dsrc, self.hsrc)
        return self.hsrc

    def step(self):
        self.hc_step()
        dt = self.dt
        self.op.enthalpy_vars(self.phi, self.vnew)
        uop, pop = self.uop, self.pop
        uop.darcy_ccoef_law(self.vold.comp_ptr(3), self.coef["darcy"], self.law, self.c)
        uop.set_coefs(None, self.bu, self.c, 1.0, -dt)
        uop.ustar_src(self.fp, self.vold, self.vnew, self.pi, self.src)
        for u, uo in zip(self.u, self.uold):
            self.copy(uo, u)
        if self.ub is not None and self.uopts.flags == 0:
            sts = self.ub.solve_be(self.ustar, self.uold, dt, srcs=self.src, opts=self.uopts)
        else:
            sts = uop.solve_be_many(self.ustar, self.uold, dt, src=self.src, opts=self.uopts)
        vs = [(st.exit_status, st.iterations) for st in sts]
        uop.cc_project_rhs(self.fp, dt, self.vold, self.vnew, self.pi, self.ustar, self.u, self.prhs)
        pop.pressure_faces(self.vold, self.vnew, self.bp)
        pop.set_coefs(None, self.bp, None, 0.0, 1.0)
        self._zero(self.pi)
        st = pop.solve(self.pi, self.prhs, opts=self.popts)
        vs.append((st.exit_status, st.iterations))
        uop.cc_correct(self.fp, dt, self.vold, self.vnew, self.pi, self.u)
        self.vstats.append(vs)

    def checkpoint_fields(self):
        """The reference's checkpoint variables of a Darcy-Brinkman state
        (Enthalpy, Bulk concentration, Pressure, Darcy velocity, advVel), and
        U*, the initial guess of the next U* solves (an extra variable, so a
        restart continues bit for bit)."""
        scalars, _, _ = super().checkpoint_fields()
        scalars["Pressure"] = self.pi.get()
        vectors = {"Darcy velocity": np.concatenate([u.get() for u in self.u], axis=0),
                   "U*": np.concatenate([u.get() for u in self.ustar], axis=0)}
        self.uop.face_velocity(self.u, self.uf)
        faces = {"advVel": [f.get() for f in self.uf]}
        return scalars, vectors, faces

    def restore_extra(self, lev):
        from . import chombo_io
        self.pi.set(np.ascontiguousarray(chombo_io.domain_array(lev, "Pressure")))
        vel = chombo_io.domain_array(lev, "Darcy velocity")
        ust = chombo_io.domain_array(lev, "U*") if "U*" in lev["fabs"] else np.zeros_like(vel)
        for d in range(len(self.n)):
            self.u[d].set(np.ascontiguousarray(vel[d:d + 1]))
            self.ustar[d].set(np.ascontiguousarray(ust[d:d + 1]))

    def flow_fields(self):
        """Host copies of the velocity components and the pressure (valid cells)."""
        out = {"u%d" % d: valid(self.n, u.get())[0] for d, u in enumerate(self.u)}
        out["p"] = valid(self.n, self.pi.get())[0]
        return out

    def close(self):
        if self.ub is not None:
            self.ub.close()
        super().close()
        self.uop.close()
        self.pop.close()


class BatchDarcyBrinkmanStepper:
    """Many independent Darcy-Brinkman instances (own grid size, St, CR,
    Ra_C, Da) stepping together: each instance's step is exactly the
    DarcyBrinkmanStepper step with the FP32-refined (H, Theta) solve, but
    the solves of all instances share launches -- the (H, Theta) solves in
    one HCBatch step with every instance's advective source, the U* solves
    (both components of every instance) in one VCBatch, the projections in
    another -- and so do the coefficient hierarchies of each operator family
    (hc_set_coefs_many, vc_set_coefs_many), the advective sources
    (hc_db_advect_src_many) and the closure, source, projection-rhs and
    correction kernels (hc_db_coefs_many, vc_db_ustar_many,
    vc_db_project_many, vc_db_correct_many): a step is a fixed number of
    launches whatever the number of instances.  `cases`: `workloads.batch_instances` cases (each
    carrying Ra_C and Da); `flow_of(case)`: the flow dict of a case."""

    def __init__(self, ml, ctx, cases, flow_of, hc_opts):
        from . import HCBatch, VCBatch
        self.ml, self.ctx = ml, ctx
        self.dt = cases[0]["dt"]
        if any(c["dt"] != self.dt for c in cases):
            raise ValueError("the instances of a batch share dt")
        self.st = [DarcyBrinkmanStepper(ml, ctx, c, hc_opts, flow_of(c)) for c in cases]
        self.hc_opts = hc_opts
        self.hcb = HCBatch(ctx, [s.op for s in self.st])
        self.ub = VCBatch(ctx, [s.uop for s in self.st], per_op=len(self.st[0].n))
        self.pb = VCBatch(ctx, [s.pop for s in self.st], per_op=1)
        self.uopts = default_opts(**U_OPTS)
        self.popts = default_opts(**P_OPTS)
        self.stats = []

    def reset(self):
        for s in self.st:
            s.reset()
        self.stats = []

    def step(self):
        dt = self.dt
        sts = self.st
        ctx = self.ctx
        # (H, Theta): enthalpy variables and coefficients at t^n, then every
        # instance's advective source, then one batched solve (which
        # assembles the same coefficients again from the state)
        hc_db_coefs_many(ctx, [s.op for s in sts], [s.phi for s in sts], [s.vold for s in sts], [s.b for s in sts])
        hc_set_coefs_many(ctx, [s.op for s in sts], None, [s.b for s in sts], 1.0, -dt)
        if all(s.advect for s in sts):
            hc_db_advect_src_many(ctx, [s.op for s in sts], [s.ap for s in sts], dt, [s.phi for s in sts],
                                  [s.vold for s in sts], [s.u for s in sts], [s.dsrc for s in sts],
                                  [s.uf for s in sts], [s.hsrc for s in sts])
            srcs = [s.hsrc for s in sts]
        else:
            srcs = [s.hc_src() for s in sts]
        hst = self.hcb.step([s.phi for s in sts], dt, self.hc_opts, srcs=srcs)
        # enthalpy variables at t^{n+1}, U* coefficients and sources
        vc_db_ustar_many(ctx, [s.op for s in sts], [s.uop for s in sts], [s.fp for s in sts],
                         [s.coef["darcy"] for s in sts], [s.law for s in sts], [0.0] * len(sts),
                         [s.phi for s in sts], [s.vold for s in sts], [s.vnew for s in sts], [s.pi for s in sts],
                         [s.c for s in sts], [s.src for s in sts], [s.u for s in sts], [s.uold for s in sts])
        vc_set_coefs_many(ctx, [s.uop for s in sts], None, [s.bu for s in sts], [s.c for s in sts], 1.0, -dt)
        ust = self.ub.solve_be([x for s in sts for x in s.ustar], [x for s in sts for x in s.uold], dt,
                               srcs=[x for s in sts for x in s.src], opts=self.uopts)
        # projection
        vc_db_project_many(ctx, [s.uop for s in sts], [s.fp for s in sts], dt, [s.vold for s in sts],
                           [s.vnew for s in sts], [s.pi for s in sts], [s.ustar for s in sts], [s.u for s in sts],
                           [s.prhs for s in sts], [s.bp for s in sts])
        vc_set_coefs_many(ctx, [s.pop for s in sts], None, [s.bp for s in sts], None, 0.0, 1.0)
        pst = self.pb.solve([s.pi for s in sts], [s.prhs for s in sts], opts=self.popts)
        vc_db_correct_many(ctx, [s.uop for s in sts], [s.fp for s in sts], dt, [s.vold for s in sts],
                           [s.vnew for s in sts], [s.pi for s in sts], [s.u for s in sts])
        d = len(sts[0].n)
        out = []
        for i, s in enumerate(sts):
            vs = [(x.exit_status, x.iterations) for x in ust[i * d:(i + 1) * d]] + \
                 [(pst[i].exit_status, pst[i].iterations)]
            s.stats.append((hst[i].exit_status, hst[i].iterations, hst[i].final_residual))
            s.vstats.append(vs)
            out.append((hst[i], vs))
        self.stats.append(out)
        return out

    def close(self):
        self.hcb.close()
        self.ub.close()
        self.pb.close()
        for s in self.st:
            s.close()
