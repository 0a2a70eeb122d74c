"""Stage-by-stage FP64 oracle model (test infrastructure only).

`OracleModel` strings the restated CPU twins of pd_oracle.c together in the
order the reference's solvers call them (Solve_Plate_Tensile.cu:48-89,
Solve_Plate_Crack_Grow.cu:48-97, Solve_Kalthoff_Winkler.cu:303-369), holding
the state as numpy arrays.  Tests drive it side by side with the GPU engine.
"""
import numpy as np

from . import oracle as O


class OracleModel:
    def __init__(self, sc, x, y, z=None, n_int=None):
        """sc: dict of model scalars (dim, model, dx, delta, vol, bc, mass, K, G,
        emod, pratio, dens, s0, dt, break_y_min, strip_velocity, n_top, n_bot,
        theta_skip_broken, reduce_lanes)."""
        self.sc = dict(sc)
        self.dim = sc["dim"]
        self.X = [np.ascontiguousarray(x, dtype=np.float64), np.ascontiguousarray(y, dtype=np.float64)]
        if self.dim == 3:
            self.X.append(np.ascontiguousarray(z, dtype=np.float64))
        self.N = len(x)
        self.N_int = self.N if n_int is None else n_int
        N, D = self.N, self.dim
        z0 = lambda: [np.zeros(N) for _ in range(D)]  # noqa: E731
        self.U, self.V, self.A, self.B, self.F = z0(), z0(), z0(), z0(), z0()
        self.VH, self.VHO, self.FO = z0(), z0(), z0()
        self.dmg = np.zeros(N)
        self.cyl = None

    # ---------------------------------------------------------------- setup
    def build_neighbors(self, brute=False, MN=None):
        x, y = self.X[0], self.X[1]
        z = self.X[2] if self.dim == 3 else None
        if brute:
            MN = MN or (32 if self.dim == 2 else 128)
            NN, NL = O.neighbors_brute(self.dim, x, y, z, self.sc["delta"], MN)
            self.row_ptr, self.col = O.dense_to_csr(NN, NL, MN)
        else:
            self.row_ptr, self.col, _ = O.neighbors_cell(self.dim, x, y, z, self.sc["delta"])
        nb = len(self.col)
        self.fail = np.ones(nb, dtype=np.uint8)
        self.scr = np.ones(nb)
        self.vol_corr()

    def vol_corr(self):
        z = self.X[2] if self.dim == 3 else None
        self.idist, self.fac = O.vol_corr(self.dim, self.row_ptr, self.col, self.X[0], self.X[1], z,
                                          self.sc["delta"], self.sc["dx"])

    def surface_correct(self, scr_mode=0):
        sc = self.sc
        sed = (0.5 * sc["emod"] / (1.0 - sc["pratio"] * sc["pratio"]) * 1.0e-6 if self.dim == 2
               else 0.6 * sc["emod"] * 1.0e-6)
        z = self.X[2] if self.dim == 3 else None
        self.scr, self.fncst = O.surface_correct(self.dim, self.row_ptr, self.col, self.X[0], self.X[1], z,
                                                 self.idist, self.fac, sc["bc"], sc["vol"], sed, scr_mode)

    def init_state(self):
        self.fail[:] = 1
        for arrs in (self.U, self.V, self.A, self.B, self.F, self.VH, self.VHO, self.FO):
            for a in arrs:
                a[:] = 0.0
        self.dmg[:] = 0.0

    def set_crack_2d(self, clength, lx, ly, theta):
        O.set_crack_2d(self.N_int, self.row_ptr, self.col, self.X[0], self.X[1], clength, lx, ly, theta, self.fail)

    def set_notches_kw(self, rad):
        O.set_crack_kw(self.N_int, self.row_ptr, self.col, self.X[0], self.X[1], rad, self.fail)

    def lps_weights(self):
        self.w, self.m = O.lps_weights(self.row_ptr, self.idist, self.fac, self.sc["delta"], self.sc["vol"])
        self.e = np.zeros(len(self.col))
        self.theta = np.zeros(self.N)

    # ----------------------------------------------------------------- step
    def ymin(self):
        return self.sc["break_y_min"] if self.dim == 3 else -np.inf

    def force(self):
        sc = self.sc
        if sc["model"] == 1:
            O.pmb_force(self.dim, self.N, self.row_ptr, self.col, self.X, self.U, self.idist, self.fac, self.scr,
                        sc["vol"], sc["bc"], sc["s0"], self.ymin(), self.fail, self.F)
        else:
            O.lps_theta(self.dim, self.N, self.row_ptr, self.col, self.X, self.U, self.idist, self.fac, self.w,
                        self.m, self.fail, sc["vol"], sc["pratio"], sc.get("theta_skip_broken", 1), self.e,
                        self.theta)
            O.lps_force(self.dim, self.N_int, self.row_ptr, self.col, self.X, self.U, self.idist, self.fac, self.w,
                        self.m, self.theta, self.e, sc["vol"], sc["pratio"], sc["K"], sc["G"], sc["s0"],
                        self.ymin(), self.fail, self.F)

    def verlet_half1(self):
        sc = self.sc
        O.quick_load(self.N_int, sc.get("n_top", 0), sc.get("n_bot", 0), sc.get("strip_velocity", 0.0), sc["dt"],
                     self.U[1])
        O.verlet_half1(self.dim, self.N_int, sc["dt"], self.V, self.A, self.U)
        if self.cyl is not None:
            O.cylinder_half1(sc["dt"], self.cst)

    def verlet_half2(self):
        O.verlet_half2(self.dim, self.N_int, self.sc["dt"], self.sc["dens"], self.V, self.A, self.F, self.B)

    def cylinder_init(self, cyl):
        self.cyl = np.asarray(cyl, dtype=np.float64)
        self.cst = O.cylinder_init(self.cyl)

    def contact_step(self):
        sc = self.sc
        T = sc.get("reduce_lanes", 4096)
        R = O.contact(self.N_int, self.dmg, self.X, self.U, self.V, self.cyl, self.cst["disp"], sc["dt"],
                      sc["dens"], sc["vol"])
        for d in range(3):
            self.cst["force"][d] = O.sum_tree(R[d], T)
        O.cylinder_half2(sc["dt"], self.cyl[1], self.cst)

    def step(self):
        """One explicit step, in the order of Solve_Kalthoff_Winkler.cu:332-357."""
        self.verlet_half1()
        self.force()
        if self.cyl is not None:
            self.contact_step()
        self.verlet_half2()

    def adr_iterate(self, ct):
        """Force, canonical-tree damping/residual reduction, ADR update."""
        sc = self.sc
        T = sc.get("reduce_lanes", 4096)
        self.force()
        cn, disp = O.adr_terms(self.dim, self.N_int, sc["mass"], sc["dt"], self.U, self.F, self.FO, self.VHO)
        rn, rd = O.resid_terms(self.dim, self.N_int, self.N, self.F, self.B)
        cn1, cn2 = O.sum_tree(cn, T), O.sum_tree(disp, T)
        rnum, rden = O.sum_tree(rn, T), O.sum_tree(rd, T)
        c = O.adr_c(cn1, cn2)
        r = np.sqrt(rnum) / np.sqrt(rden) if rden > 0 else (np.inf if rnum > 0 else 0.0)
        O.adr_integrate(self.dim, self.N_int, ct, c, sc["mass"], sc["dt"], self.F, self.B, self.VH, self.V,
                        self.VHO, self.U, self.FO)
        return c, r

    def damage(self):
        self.dmg[: self.N_int] = O.damage(self.N_int, self.row_ptr, self.fail, self.fac, self.sc["vol"])

    def stack(self, name):
        return np.stack(getattr(self, name))


# ------------------------------------------------------- batched-plate ADR
def plate_specimen(sc, n, clen, ang):
    """One plate of the batched family (Coord.cu:47-66 layout with 3 strip
    rows above and below, centred crack of length clen at angle ang) set up
    like pd_batch_create does: CSR, volume correction, intact bonds, crack.
    sc: dict of model scalars (see OracleModel)."""
    x, y = O.coord_plate_strips(n, n, sc["dx"])
    sc = dict(sc)
    sc["n_top"] = sc["n_bot"] = 3 * n
    m = OracleModel(sc, x, y, None, n * n)
    m.build_neighbors()
    m.init_state()
    if clen > 0:
        m.set_crack_2d(clen, 0.0, 0.0, ang)
    return m


def adr_increment(m, du, tol, max_iters, T):
    """One load increment of the batched ADR: strips move by +du / -du, then
    ADR (Integrate.cu:7-129 with the canonical reduction tree over T lanes)
    until the relative residual ||F+b|| free / ||F|| strips <= tol or
    max_iters.  Returns (iterations, residual)."""
    Ni, nt = m.N_int, m.sc["n_top"]
    m.U[1][Ni:Ni + nt] += du
    m.U[1][Ni + nt:] += -du
    sc = m.sc
    it = 0
    while True:
        m.force()
        cn, disp = O.adr_terms(2, Ni, sc["mass"], sc["dt"], m.U, m.F, m.FO, m.VHO)
        rn, rd = O.resid_terms(2, Ni, m.N, m.F, m.B)
        c = O.adr_c(O.sum_tree(cn, T), O.sum_tree(disp, T))
        rnum, rden = O.sum_tree(rn, T), O.sum_tree(rd, T)
        r = np.sqrt(rnum) / np.sqrt(rden) if rden > 0 else (np.inf if rnum > 0 else 0.0)
        if r <= tol or it >= max_iters or r != r:
            return it, r
        O.adr_integrate(2, Ni, it, c, sc["mass"], sc["dt"], m.F, m.B, m.VH, m.V, m.VHO, m.U, m.FO)
        it += 1
