"""Pins of the oracle's NonlinearForm Newton assembly and fixpoint loop
(src/assemblypatterns/nonlinearform.jl:44-246, src/solvers.jl:812-1017) on
the nonlinear Poisson problem -div((1 + u^2) grad u) = f of
examples/Example205_NonlinearPoisson2D.jl (q = 1):

* the Newton matrix is the Jacobian of the nonlinear residual
  R(u) = A(u) u - b(u) = int alpha(u) grad u . grad v - f v: central finite
  differences of R agree to 1e-7 (a dropped alpha' term or a transposed
  matrix fails);
* with P2 and u = x y (in the space; every integrand exact at the chosen
  quadrature orders) the iterates converge to u at every dof;
* the residuals fall quadratically."""
import numpy as np

C = (1.0, 0.0, 1.0)                       # alpha(u) = 1 + u^2
F = [0.0] * 15
F[11], F[13] = -2.0, -2.0                 # f = -div((1 + x^2 y^2)(y, x)) = -2 x^3 y - 2 x y^3


def _mesh(O, N, fe):
    from meshes import shuffled_square
    coords, cells = shuffled_square(O, N, 4)
    cf, fn, fc = O.tri_faces(cells, coords.shape[0])
    cd, nd = O.celldofs(fe, 1, cells, coords.shape[0], cf, fn.shape[0])
    fixed = O.boundary_dofs(fe, 1, coords.shape[0], fn, fc, fn.shape[0])
    pts = coords if fe == O.FE_P1 else np.vstack([coords, 0.5 * (coords[fn[:, 0]] + coords[fn[:, 1]])])
    return coords, cells, cd, nd, fixed, pts


def _residual(O, fe, mesh, u, q):
    coords, cells, cd, nd, fixed, pts = mesh
    A, bn = O.newton_assemble(fe, coords, cells, cd, nd, u, C, q)
    f = O.assemble_lf_poly(fe, F, fe + 4, coords, cells, cd, nd)
    return A @ u - (f + bn), A


def test_newton_matrix_is_the_jacobian(O):
    for fe in (O.FE_P1, O.FE_P2):
        mesh = _mesh(O, 5, fe)
        nd = mesh[3]
        q = fe + 2 * (fe - 1) + 2
        rng = np.random.default_rng(fe)
        u = rng.standard_normal(nd)
        d = rng.standard_normal(nd)
        R0, A = _residual(O, fe, mesh, u, q)
        eps = 1e-6
        Rp, _ = _residual(O, fe, mesh, u + eps * d, q)
        Rm, _ = _residual(O, fe, mesh, u - eps * d, q)
        fd = (Rp - Rm) / (2 * eps)
        Jd = A @ d
        assert np.max(np.abs(fd - Jd)) <= 1e-7 * np.max(np.abs(Jd)), fe


def test_p2_reproduces_xy_with_quadratic_convergence(O):
    mesh = _mesh(O, 6, O.FE_P2)
    coords, cells, cd, nd, fixed, pts = mesh
    u_ex = pts[:, 0] * pts[:, 1]
    x, res = O.newton_solve(O.FE_P2, coords, cells, cd, nd, fixed, u_ex, C, F, quadorder=2 + 1 + 1 + 2,
                            lf_quadorder=2 + 4, maxit=12, target_residual=1e-13)
    assert res[-1] < 1e-13 and len(res) <= 8, res
    assert np.max(np.abs(x - u_ex)) < 1e-12
    # quadratic: once below 1e-2 each residual is below a constant times the square of the last
    for r0, r1 in zip(res, res[1:]):
        if 1e-12 < r0 < 1e-2:
            assert r1 <= 10.0 * r0 * r0, res
