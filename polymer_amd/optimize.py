"""Geometry optimisation of a molecule (JOBTYPE 2: OPTIMIZATION,
src/poptimization.f90:1-489) on the device SCF and nuclear gradient.

The reference's optimiser, restated: each cycle runs a single point (SCF and
the RHF gradient dE/dR), stores the geometry and gradient, stops when every
gradient component is at most OPTCONV (DOPTN(41), GDONE :247-299), and
otherwise moves the atoms by

* GRELAX (:303-389), steepest descent R -= RELAX1 g, with RELAX1 = 1 at the
  start and, on cycles that follow a relaxation step, rescaled by the secant
  ratio g_prev / (g_prev - g_cur) of the gradient component that was largest
  on the previous cycle; RELAX1 <= 0 becomes 10 and values above 100 are cut
  to 100;
* GDIIS (:393-489) every GDIISTERM-th cycle (IOPTN(42), default 4; 0 = never,
  1 is rejected): the geometry sum_j c_j R_j over the last GDIISTERM cycles
  with sum_j c_j = 1 minimising |sum_j c_j g_j|^2 (the bordered Gram matrix
  of the stored gradients).

CHECK_GEOMETRY's molecular test (src/pread.f90:1791-1900): two atoms at the
same position abort with "ZERO INTERNUCLEAR SEPARATION".  Periodic cells and
the helix angle (RELAX2, RELAX3) are not optimised here.
"""
import numpy as np

MAXOPTCYCLES = 100     # src/pmodules.f90:232


class OptimizationError(RuntimeError):
    """An abort of the optimiser; `history` holds the cycles run, when any."""
    history = None


def _check_geometry(x):
    n = len(x)
    for i in range(n):
        for j in range(i + 1, n):
            if np.linalg.norm(x[i] - x[j]) == 0.0:
                raise OptimizationError("ZERO INTERNUCLEAR SEPARATION")


def _gdiis(xs, gs):
    """Geometry with the coefficients that minimise the extrapolated
    gradient under sum c = 1 (the reference's bordered Gram system)."""
    m = len(xs)
    A = np.empty((m + 1, m + 1))
    for a in range(m):
        for b in range(m):
            A[a, b] = float(np.sum(gs[a] * gs[b]))
    A[:m, m] = A[m, :m] = -1.0
    A[m, m] = 0.0
    rhs = np.zeros(m + 1)
    rhs[m] = -1.0
    c = np.linalg.solve(A, rhs)[:m]
    return sum(ci * xi for ci, xi in zip(c, xs)), c


def optimize(evaluate, x0, optconv=4e-4, gdiisterm=4, maxcycles=MAXOPTCYCLES):
    """Run the reference's optimiser from coordinates x0 (natom x 3, bohr).
    evaluate(x) -> (energy, gradient (natom x 3, Eh/bohr)).  Returns a dict
    with converged, cycles, x (final geometry), energy, gradient and the
    per-cycle history (x, energy, gradient, step kind)."""
    if gdiisterm < 0 or gdiisterm == 1:
        raise OptimizationError("INVALID GDIISTERM")
    x = np.array(x0, dtype=np.float64)
    _check_geometry(x)
    relax1 = 1.0
    imax = None
    hist = []
    for cycle in range(1, maxcycles + 1):
        e, g = evaluate(x)
        g = np.asarray(g, dtype=np.float64)
        hist.append({"x": x.copy(), "energy": float(e), "gradient": g.copy()})
        if np.max(np.abs(g)) <= optconv:
            return {"converged": True, "cycles": cycle, "x": x, "energy": float(e), "gradient": g,
                    "history": hist}
        if gdiisterm != 0 and cycle != 1 and cycle % gdiisterm == 0:
            x, c = _gdiis([h["x"] for h in hist[-gdiisterm:]], [h["gradient"] for h in hist[-gdiisterm:]])
            hist[-1]["step"] = ("gdiis", c)
        else:
            if (gdiisterm != 0 and cycle % gdiisterm != 1) or (gdiisterm == 0 and cycle != 1):
                gp, gc = hist[-2]["gradient"][imax], g[imax]
                if gp - gc != 0.0:
                    relax1 *= gp / (gp - gc)
            if relax1 <= 0.0:
                relax1 = 10.0
            if relax1 > 100.0:
                relax1 = 100.0
            x = x - relax1 * g
            imax = np.unravel_index(int(np.argmax(np.abs(g))), g.shape)   # first largest, atom-major
            hist[-1]["step"] = ("relax", relax1)
        _check_geometry(x)
    err = OptimizationError("OPTMSG: TOO MANY ITERATIONS IN GEOMETRY OPTIMIZATION")
    err.history = hist
    raise err


def optimize_molecule(basis_name, atoms, optconv=4e-4, gdiisterm=4, maxcycles=MAXOPTCYCLES, schwarz_tol=1e-12,
                      scf=None, device=0, restart=True):
    """The optimiser on the device path: each cycle builds a context for the
    current geometry, runs its SCF (options scf, as for
    Context.self_consistent_field) and takes Context.scf_gradient.  With
    restart, every cycle after the first starts its SCF from the previous
    cycle's converged density, as the reference's single point restores the
    stored density before its SCF (RESTORE_DENSITYMATRIX,
    src/psinglepoint.f90:192-194).  atoms: (Z, x, y, z) in bohr.  Returns
    optimize()'s dict with atoms at the final geometry and the SCF
    iterations of each cycle."""
    import polymer_amd as pa
    z = [int(a[0]) for a in atoms]
    opts = dict(scfconv=1e-10, diis=2, scfcycles=100)
    if scf:
        opts.update(scf)
    state = {"P": None, "iterations": []}

    def evaluate(x):
        geom = [(zi, float(p[0]), float(p[1]), float(p[2])) for zi, p in zip(z, x)]
        ctx = pa.Context(pa.read_basis(basis_name, geom), schwarz_tol=schwarz_tol, device=device)
        try:
            r = ctx.self_consistent_field(guess=state["P"] if restart else None, **opts)
            if not r["converged"]:
                raise OptimizationError("SCF did not converge during the geometry optimisation")
            state["P"] = r["P"]
            state["iterations"].append(r["iterations"])
            return r["e_total"], ctx.scf_gradient(r)
        finally:
            ctx.close()

    out = optimize(evaluate, np.array([[a[1], a[2], a[3]] for a in atoms]), optconv, gdiisterm, maxcycles)
    out["atoms"] = [(zi, float(p[0]), float(p[1]), float(p[2])) for zi, p in zip(z, out["x"])]
    out["scf_iterations"] = state["iterations"]
    return out
